"""Infinite-size DMRG without a GPU: the oracle's restatement pinned by
mathematics (the cell tensors of a two-cell ring against the dense product
of the shift gates, the dominant eigenvector against a dense symmetric
eigensolve, the transverse-field Ising chain against its exact energy), and
the device driver's descriptors run by a numpy emulation of the C-ABI calls
against the oracle."""
import os
import sys

import numpy as np
import pytest
from scipy.integrate import quad

sys.path.insert(0, os.path.dirname(__file__))

from oracle import idmrg_ref as I  # noqa: E402
from tnalg_amd import parameters as P  # noqa: E402


def _para(dtype, chi, sym, sweeps=20, dt_ob=1, **kw):
    p = dict(dmrg_type=dtype, spin='half', jxy=1.0, jz=1.0, hx=0.0, hz=0.0, n_site=2, chi=chi,
             sweep_time=sweeps, tau=1e-4, eigs_tol=1e-15, break_tol=2e-10, is_symme_env=sym,
             is_real=True, form='center_ort', dt_ob=dt_ob)
    p.update(kw)
    return P.make_para_consistent_idmrg(p)


def _gate_on(g, i, j, nsp=4):
    """The 2-spin operator g (4 x 4, on (s_i, s_j)) on nsp spins, dense."""
    g4 = g.reshape(2, 2, 2, 2)
    full = np.zeros((2,) * (2 * nsp))
    it = np.ndindex(*(2,) * (2 * nsp))
    for idx in it:
        out, inn = idx[:nsp], idx[nsp:]
        if any(out[k] != inn[k] for k in range(nsp) if k not in (i, j)):
            continue
        full[idx] = g4[out[i], out[j], inn[i], inn[j]]
    return full.reshape(2 ** nsp, 2 ** nsp)


@pytest.mark.parametrize('tau,way', [(1e-4, 'shift'), (0.3, 'shift'), (0.2, 'expm')])
def test_cell_tensor_ring_is_the_gate_product(tau, way):
    """Two cells on a ring (MPO bonds traced) are the intra-cell gates on
    spins (0,1), (2,3) followed by the inter-cell gates on (1,2), (3,0);
    the product's and the oracle's cell tensors agree."""
    from scipy.linalg import expm
    h = I.bond_hamiltonian(0.7, 1.2, 0.3, 0.2)
    t = I.cell_tensor(h, tau, way)
    assert np.array_equal(t, P.hamiltonian2cell_tensor(h, tau, way))
    g = np.eye(4) - tau * h if way == 'shift' else expm(-tau * h)
    ring = np.einsum('kXCq,qYEk->XYCE', t, t).reshape(4, 4, 4, 4)   # (x y)(x' y') / (c e)(c' e')
    ring = ring.reshape(16, 16)
    intra = _gate_on(g, 0, 1) @ _gate_on(g, 2, 3)
    inter = _gate_on(g, 1, 2) @ _gate_on(g, 3, 0)
    assert np.allclose(ring, inter @ intra, rtol=0, atol=1e-13)
    # a plausible slip (gates in the other order, off by O(tau^2)) is told apart
    assert np.max(np.abs(ring - intra @ inter)) > 1e-10


def test_dominant_eigenvector_is_the_top_of_the_spectrum():
    """The oracle's 'mpo' centre (eigsh, which='LM') is the eigenvector of
    the largest eigenvalue of the dense symmetric h (numpy.linalg.eigh) at
    every one of the first iterations: the device path solves for the
    lowest eigenvector of -h."""
    para = _para('mpo', 3, False, jxy=0.8, hx=0.3, hz=0.1)
    rng = np.random.RandomState(5)
    A = I.RefInfinite(4, 3, 4, 'mpo', np.asarray(para['hamilt_index']), False, rng)
    t = I.cell_tensor(I.bond_hamiltonian(0.8, 1.0, 0.15, 0.05), para['tau'])
    A.update_ort_tensor_mps()
    A.update_left_env(t)
    A.update_right_env(t)
    for _ in range(4):
        h = A.effective_hamiltonian(t)
        w, v = np.linalg.eigh(h)
        assert w[-1] > abs(w[0])
        A.update_central_tensor(t, para['tau'])
        assert abs(abs(v[:, -1] @ A.mps[1].reshape(-1)) - 1) < 1e-10
        A.update_ort_tensor_mps()
        A.update_left_env(t)
        A.update_right_env(t)


def test_white_fixed_point_is_the_exact_ising_energy():
    """'white' iDMRG of the transverse-field Ising chain
    H = sum Sz Sz - hx sum Sx (jxy = 0, gapped at hx = 1) converges to the
    exact ground energy per bond -(1/pi) int_0^pi sqrt(J^2 + g^2 + 2 J g
    cos k) dk with J = 1/4, g = hx / 2."""
    para = _para('white', 8, True, sweeps=300, dt_ob=10, jxy=0.0, hx=1.0)
    J, g = 0.25, 0.5
    exact = -quad(lambda k: np.sqrt(J * J + g * g + 2 * J * g * np.cos(k)), 0, np.pi)[0] / np.pi
    rec = []
    _, ob = I.dmrg_infinite_size(para, seed=1, record=rec)
    assert rec[-1][0] < 290                 # stopped by the energy test
    assert abs(float(ob['eb']) - exact) < 1e-11


@pytest.mark.parametrize('dtype,sym,chi,kw', [
    ('mpo', False, 5, dict(jxy=0.8, hx=0.3, hz=0.1)),
    ('mpo', True, 4, dict(jxy=0.8, hx=0.3, hz=0.1)),
    ('white', True, 6, dict(jxy=0.0, hx=1.0)),
    ('white', True, 6, dict(jxy=0.4, hx=0.2, hz=0.3)),
])
def test_driver_descriptors_match_the_oracle(dtype, sym, chi, kw):
    """The device driver's descriptors, run by the numpy emulation of the
    C-ABI calls, give the oracle's energies at every iteration, its Schmidt
    values and (gauge-free) spectra of its environments / bath."""
    import torch
    from idmrg_emul import EmulatedLib
    from tnalg_amd.idmrg import InfiniteRun, MpsInfinite
    para = _para(dtype, chi, sym, **kw)
    rec = []
    Aref, _ = I.dmrg_infinite_size(para, seed=3, record=rec)
    h = P.hamiltonian_heisenberg('half', para['jxy'], para['jxy'], para['jz'], para['hx'] / 2,
                                 para['hz'] / 2)
    A = MpsInfinite('center_ort', 4 if dtype == 'mpo' else 2, chi, 4, n_site=2,
                    is_symme_env=para['is_symme_env'], dmrg_type=dtype,
                    hamilt_index=para['hamilt_index'], rng=np.random.RandomState(3),
                    device='cpu')
    (e1, info), = InfiniteRun([A], [para], [h], lib=EmulatedLib(), stream=lambda: None).run()
    obs = info['observations']
    assert [t for t, _ in obs] == [t for t, _ in rec]
    for (_, a), (_, b) in zip(obs, rec):
        assert np.allclose(a, b, rtol=0, atol=1e-10)
    assert np.allclose(A.lm[0], Aref.lm[0], rtol=0, atol=1e-10)
    if dtype == 'mpo':
        for k in range(2):
            ours, ref = A.env[k], Aref.env[k]
            for m in range(ours.shape[1]):
                assert np.allclose(np.linalg.eigvalsh(ours[:, m]), np.linalg.eigvalsh(ref[:, m]),
                                   rtol=0, atol=1e-10)
    else:
        assert np.allclose(np.linalg.eigvalsh(A.bath_op_onsite),
                           np.linalg.eigvalsh(Aref.bath_op_onsite), rtol=0, atol=1e-9)
    assert torch.isfinite(A.psi).all()


def test_unsupported_configurations_raise():
    from tnalg_amd.idmrg import MpsInfinite
    with pytest.raises(ValueError):
        MpsInfinite('center_ort', 2, 4, 4, n_site=1, dmrg_type='white', device='cpu')
    with pytest.raises(ValueError):
        MpsInfinite('translation_invariant', 2, 4, 4, n_site=2, dmrg_type='white', device='cpu')
    with pytest.raises(ValueError):
        MpsInfinite('center_ort', 4, 4, 2, n_site=2, dmrg_type='mpo', device='cpu')
