"""tnx_lanczos_run_multi_static (the entry point that allocates nothing: its
Lanczos state lives in the library's own device array) against
tnx_lanczos_run_multi with a caller workspace, through the C ABI: the same
batch of dense symmetric eigenproblems at two matvec precisions gives
bit-identical results (Ritz values, residuals, matvec counts, convergence
flags and eigenvectors), on two calls in a row, the second with a smaller
batch; a batch whose state exceeds the array is rejected before any launch.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NCV = 20


def _vp(x):
    return ctypes.c_void_p(int(x))


def _batch(sizes, dts, seed):
    import torch
    from tnalg_amd import _lib as L
    rng = np.random.RandomState(seed)
    tdt = {L.F64: torch.float64, L.F32: torch.float32}
    bufs = []
    for n, dt in zip(sizes, dts):
        A = rng.randn(n, n)
        diag = np.linspace(0.0, 10.0, n)
        diag[0] = -2.0
        H = 0.002 * (A + A.T) / 2 + np.diag(diag)
        bufs.append(dict(n=n, dt=dt,
                         H=torch.from_numpy(H).to(tdt[dt]).cuda(),
                         V=torch.zeros((NCV + 1) * n, dtype=torch.float64, device='cuda'),
                         w=torch.zeros(n, dtype=tdt[dt], device='cuda'),
                         x=torch.zeros(n, dtype=tdt[dt], device='cuda'),
                         init=torch.from_numpy(rng.randn(n)).cuda(),
                         out=torch.zeros(n, dtype=torch.float64, device='cuda')))
    return bufs


def _descriptors(bufs):
    """Matvec plans (one per dtype) and the solve records of a batch."""
    import torch
    from tnalg_amd import _lib as L
    from tnalg_amd import records as R

    def up(a):
        return torch.from_numpy(np.frombuffer(np.ascontiguousarray(a).tobytes(),
                                              np.uint8).copy()).cuda()
    plans, keep = [], []
    for dt, mode in ((L.F64, L.GEMM_F64), (L.F32, L.GEMM_F32)):
        g = R.GemmList()
        for pos, b in enumerate(bufs):
            if b['dt'] == dt:
                g.add(b['w'].data_ptr(), b['n'], 1, 1,
                      [(b['H'].data_ptr(), b['n'], 0, b['x'].data_ptr(), 1, 0, b['n'], 1.0)],
                      active=pos)
        p, s = g.arrays()
        if len(p):
            dp, ds = up(p), up(s)
            keep += [dp, ds]
            plans.append(L.MatvecPlan(0, 0, 0, 0, 0, 0, 0, 0, dp.data_ptr(), ds.data_ptr(),
                                      len(p), R.gemm_tiles(p), 0, 0, mode))
    recs = [(b['V'].data_ptr(), b['w'].data_ptr(), b['x'].data_ptr(), b['init'].data_ptr(),
             b['out'].data_ptr(), b['n'], b['n'], min(NCV, b['n']), 0, 1e-12 / 1e-4, 1e-4,
             2e-6 if b['dt'] == L.F32 else 0.0, b['dt'], b['dt']) for b in bufs]
    lz = L.lanczos_records(recs)
    ch = (lz['n'] + L.LZ_CHUNK - 1) // L.LZ_CHUNK
    lz['chunk_begin'] = np.concatenate([[0], np.cumsum(ch)[:-1]])
    dlz = up(lz)
    keep.append(dlz)
    return (L.MatvecPlan * len(plans))(*plans), len(plans), dlz, int(ch.sum()), keep


def _run(bufs, static):
    import torch
    from tnalg_amd import _lib as L
    lib = L.load()
    plans, nplan, dlz, nch, keep = _descriptors(bufs)
    res = np.zeros(len(bufs), dtype=L.LANCZOS_RESULT)
    stream = _vp(torch.cuda.current_stream().cuda_stream)
    for b in bufs:
        b['out'].zero_()
    if static:
        rc = lib.tnx_lanczos_run_multi_static(plans, nplan, _vp(dlz.data_ptr()), len(bufs), nch,
                                              NCV, 50, 1, _vp(res.ctypes.data), stream)
    else:
        state = torch.zeros(int(lib.tnx_lanczos_state_bytes(len(bufs), nch, NCV)),
                            dtype=torch.uint8, device='cuda')
        keep.append(state)
        rc = lib.tnx_lanczos_run_multi(plans, nplan, _vp(dlz.data_ptr()), len(bufs), nch, NCV,
                                       50, 1, _vp(state.data_ptr()), _vp(res.ctypes.data),
                                       stream)
    assert rc == 0, (rc, lib.tnx_last_error())
    torch.cuda.synchronize()
    return res, [b['out'].cpu().numpy() for b in bufs]


def test_static_lanczos_equals_workspace_lanczos(gpu):
    from tnalg_amd import _lib as L
    cases = [([16, 300, 2000, 4100, 64, 1, 3], [L.F64, L.F32, L.F64, L.F32, L.F32, L.F64, L.F64]),
             ([40, 1500], [L.F32, L.F64])]    # the second call: a smaller batch
    for k, (sizes, dts) in enumerate(cases):
        print('case: static vs workspace Lanczos, call %d, sizes %s' % (k, sizes), flush=True)
        bufs = _batch(sizes, dts, seed=30 + k)
        ref, ref_out = _run(bufs, static=False)
        got, got_out = _run(bufs, static=True)
        assert np.array_equal(ref, got), (ref, got)
        for a, b in zip(ref_out, got_out):
            assert np.array_equal(a, b)
        assert ref['converged'].all() and (ref['iters'] > 0).all()


def test_static_lanczos_rejects_oversized_batch(gpu):
    """The size check runs on the host before any launch."""
    from tnalg_amd import _lib as L
    lib = L.load()
    nsolve, nch = 64, 20000   # about 10 MB of partial sums
    assert lib.tnx_lanczos_state_bytes(nsolve, nch, NCV) > 4 << 20
    plan = (L.MatvecPlan * 1)(L.MatvecPlan())
    res = np.zeros(nsolve, dtype=L.LANCZOS_RESULT)
    rc = lib.tnx_lanczos_run_multi_static(plan, 1, _vp(8), nsolve, nch, NCV, 5, 1,
                                          _vp(res.ctypes.data), None)
    assert rc == L.EINVAL
    assert b'scratch' in lib.tnx_last_error()
