"""Imaginary-time TEBD on the device kernels (tebd_standard, the
TEBDalgo.py:11-87 drop-in) against the oracle's restatement and the dense
product of the gates."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..'))

pytestmark = pytest.mark.gpu


def _para(l, chi, **kw):
    from tnalg_amd.parameters import generate_parameters_standard_tebd, make_para_consistent_tebd
    para = generate_parameters_standard_tebd()
    para.update({'l': l, 'chi': chi, 'jxy': 1.0, 'jz': 0.7, 'hx': 0.3, 'hz': 0.2})
    para.update(kw)
    return make_para_consistent_tebd(para)


def test_tebd_sweeps_equal_dense_gates(gpu):
    """Without truncation (chi = 2^(l/2)) three sweeps of gates on the device
    give the dense product of the gates on the seeded initial state."""
    from scipy.linalg import expm
    from oracle import tebd_ref
    from oracle.model import operator_list
    from tnalg_amd.tebd import tebd_standard
    l, chi, seed = 6, 8, 3
    para = _para(l, chi, taut=1, iterate_time=3, if_break=False)
    print('tebd dense check l', l, 'chi', chi, flush=True)
    A, ob, _ = tebd_standard(para, seed=seed)
    ref = tebd_ref.RefMps(l, 2, chi, operator_list(), np.random.RandomState(seed))
    psi = ref.full_state()
    psi /= np.linalg.norm(psi)
    gate = expm(-0.1 * tebd_ref.bond_hamiltonian(1.0, 0.7, 0.15, 0.1))
    for _ in range(3):
        psi = tebd_ref.dense_gate_sweep(psi, l, list(range(l - 1)), gate)
    got = A.full_coefficients()
    assert abs(np.linalg.norm(got) - 1) < 1e-12
    assert abs(abs(got @ psi) - 1) < 1e-12
    assert ob['e_site'] == []     # neither 'all' records nor the energy test ran


def test_tebd_matches_oracle_with_truncation(gpu):
    """Truncation to chi = 6 on a 10-site chain with fields (no degenerate
    Schmidt values at the cut: the kept/dropped ratio stays above 2), two
    values of tau, outputs recorded every 10 sweeps: every record equals the
    oracle's."""
    from oracle import tebd_ref
    from tnalg_amd.tebd import tebd_standard
    para = _para(10, 6, taut=2, iterate_time=40, dt_ob=10, save_mode='all')
    print('tebd vs oracle l 10 chi 6', flush=True)
    A, ob, _ = tebd_standard(para, seed=4)
    _, ref = tebd_ref.tebd_standard(para, seed=4)
    assert len(ob['e_site']) == len(ref['e_site']) == 8
    assert np.allclose(ob['e_site'], ref['e_site'], rtol=0, atol=1e-10)
    for k in range(8):
        for key in ('mx', 'mz', 'eb'):
            assert np.allclose(ob[key][k], ref[key][k], rtol=0, atol=1e-9), (key, k)
        for a, b in zip(ob['lm'][k], ref['lm'][k]):
            assert np.allclose(np.sort(a)[::-1], np.sort(b)[::-1], rtol=0, atol=1e-9), k
        assert np.allclose(ob['ent'][k], ref['ent'][k], rtol=0, atol=1e-9), k
    assert A.chi == 6 and max(A.virtual_dim) == 6


def test_tebd_converges_to_the_oracle_fixed_point(gpu):
    """The energy test (if_break) stops each tau at the same sweep as the
    oracle and the fixed point agrees with it and, within the Trotter error,
    with exact diagonalisation."""
    from oracle import ed, tebd_ref
    from tnalg_amd.tebd import tebd_standard
    sys.path.insert(0, os.path.dirname(__file__))
    from test_tebd_oracle import tebd_hamiltonian
    l = 8
    para = _para(l, 16, tau0=0.2, dtau=0.25, taut=2, iterate_time=800, dt_ob=10, break_tol=1e-8)
    print('tebd fixed point l', l, flush=True)
    A, ob, _ = tebd_standard(para, seed=2)
    _, ref = tebd_ref.tebd_standard(para, seed=2)
    assert len(ob['e_site']) == len(ref['e_site']) == 2
    assert np.allclose(ob['e_site'], ref['e_site'], rtol=0, atol=1e-10)
    e0, _ = ed.ground_state(l, *tebd_hamiltonian(l, 1.0, 0.7, 0.3, 0.2))
    w = np.array([0.5] + [1.0] * (l - 2) + [0.5])
    e = float(np.sum(ob['eb'][-1])) - 0.3 * float(w @ ob['mx'][-1].reshape(-1)) \
        - 0.2 * float(w @ ob['mz'][-1].reshape(-1))
    assert abs(e - e0) < 1e-4 * abs(e0), (e, e0)
