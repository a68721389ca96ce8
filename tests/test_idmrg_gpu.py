"""Infinite-size DMRG on the device (dmrg_infinite_size, the
DMRG_anyH.py:106-188 drop-in): the two new kernels against numpy, the
driver against the oracle's restatement at every iteration, batching, and
the transverse-field Ising chain against its exact energy."""
import ctypes
import os
import sys

import numpy as np
import pytest
from scipy.integrate import quad

sys.path.insert(0, os.path.dirname(__file__))

pytestmark = pytest.mark.gpu


def _vp(x):
    return ctypes.c_void_p(int(x))


def _stream():
    import torch
    return _vp(torch.cuda.current_stream().cuda_stream)


def _upload(arrs):
    import torch
    return [torch.from_numpy(np.frombuffer(np.ascontiguousarray(a).tobytes(), np.uint8).copy()
                             ).cuda() for a in arrs]


def _para(dtype, chi, sym, sweeps=20, dt_ob=1, **kw):
    from tnalg_amd import parameters as P
    p = dict(dmrg_type=dtype, spin='half', jxy=1.0, jz=1.0, hx=0.0, hz=0.0, n_site=2, chi=chi,
             sweep_time=sweeps, tau=1e-4, eigs_tol=1e-15, break_tol=2e-10, is_symme_env=sym,
             is_real=True, form='center_ort', dt_ob=dt_ob)
    p.update(kw)
    return P.make_para_consistent_idmrg(p)


def test_kron_and_symmetrize_match_numpy(gpu):
    """tnx_kron_f64: sums of Kronecker products with strided factors and
    identity factors; tnx_symmetrize_f64: (X + X^T) / 2 of strided blocks."""
    import torch
    from tnalg_amd import _lib as L
    lib = L.load()
    rng = np.random.RandomState(7)
    keep, jobs, terms, refs = [], [], [], []
    blocks = 0
    for p, q, nt in [(5, 4, 3), (16, 2, 5), (2, 33, 2), (64, 4, 4), (1, 1, 1)]:
        ref = np.zeros((p * q, p * q))
        t0 = len(terms)
        for t in range(nt):
            lda, ldb = p + 3, q + 1
            a = rng.randn(p, lda)
            b = rng.randn(q, ldb)
            c = rng.randn()
            ida, idb = (t == 1), (t == 2)
            da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
            keep += [da, db]
            terms.append((0 if ida else da.data_ptr(), 0 if idb else db.data_ptr(), lda, ldb, c))
            A = np.eye(p) if ida else a[:, :p]
            B = np.eye(q) if idb else b[:, :q]
            ref += c * np.kron(A, B)
        ldd = p * q + 2
        dst = torch.full((p * q * ldd,), 7.0, dtype=torch.float64, device='cuda')
        keep.append(dst)
        jobs.append((dst.data_ptr(), p, q, ldd, t0, nt, blocks))
        blocks += -(-(p * q) ** 2 // L.KRON_BLOCK)
        refs.append((dst, ref, ldd))
    dj, dt = _upload([np.array(jobs, dtype=L.KRON_JOB), np.array(terms, dtype=L.KRON_TERM)])
    print('kron jobs', len(jobs), 'blocks', blocks, flush=True)
    assert lib.tnx_kron_f64(_vp(dj.data_ptr()), len(jobs), _vp(dt.data_ptr()), blocks,
                            _stream()) == 0
    torch.cuda.synchronize()
    for dst, ref, ldd in refs:
        n = ref.shape[0]
        got = dst.cpu().numpy().reshape(n, ldd)
        assert np.allclose(got[:, :n], ref, rtol=0, atol=1e-12)
        assert np.all(got[:, n:] == 7.0)          # outside the matrix: untouched
    # symmetrisation of count blocks with leading dimension and stride
    sjobs, srefs, blocks = [], [], 0
    for n, cnt in [(3, 4), (17, 2), (64, 1)]:
        ld, stride = cnt * n, n
        x = rng.randn(n * ld)
        dx = torch.from_numpy(x.copy()).cuda()
        keep.append(dx)
        sjobs.append((dx.data_ptr(), n, ld, cnt, stride, blocks, 0))
        blocks += -(-(cnt * n * n) // L.KRON_BLOCK)
        X = x.reshape(n, cnt, n)
        srefs.append((dx, (X + X.transpose(2, 1, 0)) / 2))
    ds, = _upload([np.array(sjobs, dtype=L.SYM_JOB)])
    print('symmetrize jobs', len(sjobs), flush=True)
    assert lib.tnx_symmetrize_f64(_vp(ds.data_ptr()), len(sjobs), blocks, _stream()) == 0
    torch.cuda.synchronize()
    for dx, ref in srefs:
        assert np.array_equal(dx.cpu().numpy().reshape(ref.shape), ref)


@pytest.mark.parametrize('dtype,sym,chi,kw', [
    ('mpo', False, 5, dict(jxy=0.8, hx=0.3, hz=0.1)),
    ('mpo', True, 4, dict(jxy=0.8, hx=0.3, hz=0.1)),
    ('white', True, 6, dict(jxy=0.0, hx=1.0)),
    ('white', True, 6, dict(jxy=0.4, hx=0.2, hz=0.3)),
])
def test_matches_oracle_every_iteration(gpu, dtype, sym, chi, kw):
    """The device run gives the oracle's energies at every iteration
    (1e-9), its Schmidt values and the spectra of its environments / bath
    (gauge-free: degenerate Schmidt pairs leave a rotation free)."""
    from oracle import idmrg_ref as I
    from tnalg_amd.idmrg import dmrg_infinite_size
    para = _para(dtype, chi, sym, **kw)
    rec = []
    Aref, _ = I.dmrg_infinite_size(para, seed=3, record=rec)
    print('idmrg', dtype, sym, chi, kw, flush=True)
    from tnalg_amd.idmrg import InfiniteRun, MpsInfinite
    from tnalg_amd.parameters import hamiltonian_heisenberg
    h = hamiltonian_heisenberg('half', para['jxy'], para['jxy'], para['jz'], para['hx'] / 2,
                               para['hz'] / 2)
    A = MpsInfinite('center_ort', 4 if dtype == 'mpo' else 2, chi, 4, n_site=2,
                    is_symme_env=para['is_symme_env'], dmrg_type=dtype,
                    hamilt_index=para['hamilt_index'], rng=np.random.RandomState(3))
    (e1, info), = InfiniteRun([A], [para], [h]).run()
    assert info['iterations'] == len(rec) and info['unconverged_solves'] == 0
    for (ta, a), (tb, b) in zip(info['observations'], rec):
        assert ta == tb and np.allclose(a, b, rtol=0, atol=1e-9), (ta, a, b)
    # the drop-in entry point runs the same computation
    A2, ob, info2 = dmrg_infinite_size(para, rng=np.random.RandomState(3), is_print=False)
    assert np.array_equal(np.asarray(ob['eb']), np.asarray(e1))
    assert np.allclose(A.lm[0], Aref.lm[0], rtol=0, atol=1e-9)
    if dtype == 'mpo':
        for k in range(2):
            for m in range(4):
                assert np.allclose(np.linalg.eigvalsh(A.env[k][:, m]),
                                   np.linalg.eigvalsh(Aref.env[k][:, m]), rtol=0, atol=1e-9)
        assert len(A.rho) == 3
    else:
        assert np.allclose(np.linalg.eigvalsh(A.bath_op_onsite),
                           np.linalg.eigvalsh(Aref.bath_op_onsite), rtol=0, atol=1e-8)
    # the returned state: centre normalised, left part an isometry
    c = A.mps[1].reshape(-1)
    assert abs(np.linalg.norm(c) - 1) < 1e-12
    a0 = A.mps[0].reshape(-1, chi)
    assert np.allclose(a0.T @ a0, np.eye(chi), rtol=0, atol=1e-12)


def test_every_observation_matches_oracle(gpu):
    """Per-iteration energies of an 'mpo' run (record of every observation)."""
    from oracle import idmrg_ref as I
    from tnalg_amd.idmrg import InfiniteRun, MpsInfinite
    from tnalg_amd.parameters import hamiltonian_heisenberg
    para = _para('mpo', 6, False, sweeps=12, jxy=0.5, jz=1.0, hx=0.4)
    rec = []
    I.dmrg_infinite_size(para, seed=9, record=rec)
    h = hamiltonian_heisenberg('half', 0.5, 0.5, 1.0, 0.2, 0.0)
    A = MpsInfinite('center_ort', 4, 6, 4, n_site=2, is_symme_env=False, dmrg_type='mpo',
                    hamilt_index=para['hamilt_index'], rng=np.random.RandomState(9))
    print('idmrg observations mpo chi 6', flush=True)
    (e1, info), = InfiniteRun([A], [para], [h]).run()
    assert len(info['observations']) == len(rec) == 12
    for (ta, a), (tb, b) in zip(info['observations'], rec):
        assert ta == tb and np.allclose(a, b, rtol=0, atol=1e-9), (ta, a, b)


def test_batch_equals_single_runs(gpu):
    """Three chains of one structure in a batch give bit-identical results to
    their runs alone (a chain's kernels read only its own data), including a
    chain that converges earlier and leaves the batch."""
    from tnalg_amd.idmrg import dmrg_infinite_size, dmrg_infinite_size_batch
    paras = [_para('white', 8, True, sweeps=60, dt_ob=5, jxy=0.0, hx=hx) for hx in (1.0, 1.6)]
    paras.append(_para('white', 8, True, sweeps=25, dt_ob=5, jxy=0.3, hx=0.2, hz=0.4))
    print('idmrg batch of 3 white chi 8', flush=True)
    batch = dmrg_infinite_size_batch(paras, seeds=[1, 2, 3])
    for k, para in enumerate(paras):
        A, ob, info = dmrg_infinite_size(para, rng=np.random.RandomState(k + 1), is_print=False)
        Ab, obb, infob = batch[k]
        assert np.array_equal(np.asarray(ob['eb']), np.asarray(obb['eb']))
        assert info['iterations'] == infob['iterations']
        assert np.array_equal(A.lm[0], Ab.lm[0])
        assert np.array_equal(A.mps[1], Ab.mps[1])


def test_white_reaches_exact_ising_energy(gpu):
    """The transverse-field Ising chain (jxy = 0, hx = 1: J = 1/4, g = 1/2)
    at chi = 16 converges to the exact energy per bond."""
    from tnalg_amd.idmrg import dmrg_infinite_size
    J, g = 0.25, 0.5
    exact = -quad(lambda k: np.sqrt(J * J + g * g + 2 * J * g * np.cos(k)), 0, np.pi)[0] / np.pi
    para = _para('white', 16, True, sweeps=400, dt_ob=10, jxy=0.0, hx=1.0)
    print('idmrg white ising chi 16', flush=True)
    A, ob, info = dmrg_infinite_size(para, rng=np.random.RandomState(4), is_print=False)
    assert info['converged']
    assert abs(float(ob['eb']) - exact) < 1e-11


def test_default_parameters_reach_the_critical_ising_energy(gpu):
    """generate_parameters_infinite_dmrg() (the reference's defaults: 'mpo',
    chi 16, H = sum Sz Sz - 0.5 sum Sx, the critical Ising chain J = g = 1/4)
    approaches the exact energy per bond -4 J / pi = -1 / pi on all three
    bonds of the centre within 200 iterations."""
    from tnalg_amd.idmrg import dmrg_infinite_size
    from tnalg_amd.parameters import generate_parameters_infinite_dmrg
    para = generate_parameters_infinite_dmrg()
    para['sweep_time'] = 200
    print('idmrg defaults', flush=True)
    A, ob, info = dmrg_infinite_size(para, rng=np.random.RandomState(0), is_print=True)
    print('unconverged eigensolves at', info['unconverged_at'], 'largest residual / theta',
          info['max_resid_rel'], flush=True)
    # at criticality an early solve (the second, from a random-state block)
    # may end at the restart limit a little above the eps |theta| test
    assert info['unconverged_solves'] <= 2 and info['max_resid_rel'] < 1e-11
    eb = np.asarray(ob['eb'])
    assert eb.shape == (3,)
    assert np.max(np.abs(eb + 1 / np.pi)) < 1e-6, eb
