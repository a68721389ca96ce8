"""Kernel-level parity of libtnalg_hip.so against numpy (f64 references).

Each test prints the case it runs before launching, so a device fault names it.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _vp(x):
    return ctypes.c_void_p(int(x))


def _stream():
    import torch
    return _vp(torch.cuda.current_stream().cuda_stream)


def _dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def _upload(arrs):
    import torch
    out = []
    for a in arrs:
        b = torch.from_numpy(np.frombuffer(np.ascontiguousarray(a).tobytes(), np.uint8).copy()).cuda()
        out.append(b)
    return out


MODES = {
    'f64': (0, 'float64', 'float64', 1e-12),
    'f32': (1, 'float32', 'float32', 2e-5),
    'bf16_f32': (2, 'bfloat16', 'float32', 2e-2),
    'bf16_bf16': (3, 'bfloat16', 'bfloat16', 3e-2),
}


@pytest.mark.parametrize('mode', list(MODES))
def test_gemm_grouped_matches_numpy(gpu, mode):
    import torch
    from tnalg_amd import _lib as L
    from tnalg_amd import records as R
    lib = L.load()
    code, tin, tout, tol = MODES[mode]
    tin_t, tout_t = getattr(torch, tin), getattr(torch, tout)
    rng = np.random.RandomState(7)
    shapes = [(1, 1, 1), (2, 8, 4), (64, 64, 64), (65, 130, 33), (128, 256, 128), (37, 3, 200),
              (256, 128, 512)]
    keep = []
    g = R.GemmList()
    expect = []
    outs = []
    for pi, (M, N, K) in enumerate(shapes):
        nseg = 1 + pi % 3
        ref = np.zeros((M, N))
        segs = []
        for s in range(nseg):
            ta, tb = (pi + s) % 2, (pi // 2 + s) % 2
            A = rng.randn(K, M) if ta else rng.randn(M, K)
            B = rng.randn(N, K) if tb else rng.randn(K, N)
            alpha = 0.5 + s
            Ad, Bd = _dev(A).to(tin_t), _dev(B).to(tin_t)
            keep += [Ad, Bd]
            Aq = Ad.to(torch.float64).cpu().numpy()
            Bq = Bd.to(torch.float64).cpu().numpy()
            ref += alpha * ((Aq.T if ta else Aq) @ (Bq.T if tb else Bq))
            segs.append((Ad.data_ptr(), A.shape[1], ta, Bd.data_ptr(), B.shape[1], tb, K, alpha))
        C0 = rng.randn(M, N)
        C0d = _dev(C0).to(tout_t)
        keep.append(C0d)
        ref += -0.75 * C0d.to(torch.float64).cpu().numpy()
        C = torch.zeros((M, N), dtype=tout_t, device='cuda')
        outs.append(C)
        g.add(C.data_ptr(), M, N, N, segs, c0=C0d.data_ptr(), ldc0=N, beta=-0.75)
        expect.append(ref)
    p, s = g.arrays()
    nt = R.gemm_tiles(p)
    dp, ds = _upload([p, s])
    print('gemm mode', mode, 'problems', len(p), 'tiles', nt, flush=True)
    L.check(lib.tnx_gemm_grouped(code, _vp(dp.data_ptr()), len(p), _vp(ds.data_ptr()), nt, None,
                                 _stream()), 'gemm')
    torch.cuda.synchronize()
    for C, ref, sh in zip(outs, expect, shapes):
        got = C.to(torch.float64).cpu().numpy()
        scale = np.abs(ref).max() + 1e-300
        err = np.abs(got - ref).max() / scale
        assert err < tol, (mode, sh, err)


@pytest.mark.parametrize('mode_name', ['f64', 'f32', 'bf16_f32', 'bf16_bf16'])
def test_gemm_variants_agree(gpu, mode_name):
    """Every kernel variant of a GEMM mode gives the same bits (they issue
    the same MFMAs in the same k order for every output block); an unknown
    variant is refused."""
    import torch
    from tnalg_amd import _lib as L
    from tnalg_amd import records as R
    lib = L.load()
    mode, tin, tout = {'f64': (L.GEMM_F64, torch.float64, torch.float64),
                       'f32': (L.GEMM_F32, torch.float32, torch.float32),
                       'bf16_f32': (L.GEMM_BF16_F32, torch.bfloat16, torch.float32),
                       'bf16_bf16': (L.GEMM_BF16_BF16, torch.bfloat16, torch.bfloat16)}[mode_name]
    nvar = lib.tnx_gemm_variants(mode)
    assert nvar >= (1 if mode == L.GEMM_F32 else 2)
    rng = np.random.RandomState(12)
    g = R.GemmList()
    keep, outs = [], []
    for (M, N, K, ta, tb) in [(64, 64, 128, 0, 0), (130, 70, 512, 1, 0), (33, 200, 37, 0, 1),
                              (256, 128, 300, 1, 1)]:
        A = _dev(rng.randn(K, M) if ta else rng.randn(M, K), tin)
        B = _dev(rng.randn(N, K) if tb else rng.randn(K, N), tin)
        C0 = _dev(rng.randn(M, N), tout)
        keep += [A, B, C0]
        C = torch.zeros((M, N), dtype=tout, device='cuda')
        outs.append(C)
        g.add(C.data_ptr(), M, N, N, [(A.data_ptr(), A.shape[1], ta, B.data_ptr(), B.shape[1], tb,
                                       K, 0.75)], c0=C0.data_ptr(), ldc0=N, beta=-0.5)
    p, s = g.arrays()
    dp, ds = _upload([p, s])
    got = []
    for v in range(nvar):
        print('gemm', mode_name, 'variant', v, flush=True)
        for C in outs:
            C.zero_()
        L.check(lib.tnx_gemm_grouped(mode | (v << 8), _vp(dp.data_ptr()), len(p),
                                     _vp(ds.data_ptr()), R.gemm_tiles(p), None, _stream()), 'gemm')
        torch.cuda.synchronize()
        got.append([C.float().cpu().numpy().copy() for C in outs])
    for v in range(1, nvar):
        for a, b in zip(got[0], got[v]):
            assert np.array_equal(a, b), v
    rc = lib.tnx_gemm_grouped(mode | (nvar << 8), _vp(dp.data_ptr()), len(p),
                              _vp(ds.data_ptr()), R.gemm_tiles(p), None, _stream())
    assert rc == L.EINVAL


def test_gemm_active_mask_skips_problems(gpu):
    import torch
    from tnalg_amd import _lib as L
    from tnalg_amd import records as R
    lib = L.load()
    A = _dev(np.ones((8, 8)))
    C1 = torch.full((8, 8), 7.0, dtype=torch.float64, device='cuda')
    C2 = torch.full((8, 8), 7.0, dtype=torch.float64, device='cuda')
    g = R.GemmList()
    g.add(C1.data_ptr(), 8, 8, 8, [(A.data_ptr(), 8, 0, A.data_ptr(), 8, 0, 8, 1.0)], active=0)
    g.add(C2.data_ptr(), 8, 8, 8, [(A.data_ptr(), 8, 0, A.data_ptr(), 8, 0, 8, 1.0)], active=1)
    p, s = g.arrays()
    act = torch.tensor([1, 0], dtype=torch.int32, device='cuda')
    dp, ds = _upload([p, s])
    L.check(lib.tnx_gemm_grouped(0, _vp(dp.data_ptr()), 2, _vp(ds.data_ptr()), R.gemm_tiles(p),
                                 _vp(act.data_ptr()), _stream()), 'gemm')
    torch.cuda.synchronize()
    assert torch.all(C1 == 8.0) and torch.all(C2 == 7.0)


def test_mix_matches_numpy(gpu):
    import torch
    from tnalg_amd import _lib as L
    from tnalg_amd import records as R
    lib = L.load()
    rng = np.random.RandomState(3)
    flat = rng.randn(200)   # operator k of a job of dimension d: flat[k d^2 : (k + 1) d^2]
    opsd = _dev(flat)
    mx = R.MixList()
    keep, expect, outs = [], [], []
    # jobs of every physical dimension in one launch (d <= 2 and d > 2 jobs
    # cover different numbers of (a, b) pairs per workgroup)
    for (Dl, d, Dr, acc) in [(1, 2, 1, 0), (3, 2, 5, 1), (40, 3, 33, 0), (128, 2, 64, 1),
                             (70, 4, 50, 1), (9, 1, 300, 0), (64, 3, 64, 1)]:
        op = lambda k: flat[k * d * d:(k + 1) * d * d].reshape(d, d)
        srcs = [rng.randn(Dl, d, Dr) for _ in range(3)]
        sd = [_dev(s) for s in srcs]
        keep += sd
        dst0 = rng.randn(Dl, d, Dr)
        dst = _dev(dst0)
        outs.append(dst)
        terms = [(sd[0].data_ptr(), L.F64, 0.5, -1), (sd[1].data_ptr(), L.F64, -1.5, 2),
                 (sd[2].data_ptr(), L.F64, 2.0, 4)]
        ref = 0.5 * srcs[0] - 1.5 * np.einsum('st,atb->asb', op(2), srcs[1]) + \
            2.0 * np.einsum('st,atb->asb', op(4), srcs[2])
        if acc:
            ref = ref + dst0
        expect.append(ref)
        mx.add(dst.data_ptr(), L.F64, Dl, d, Dr, terms, accumulate=acc)
    j, t = mx.arrays()
    dj, dt = _upload([j, t])
    print('mix jobs', len(j), flush=True)
    L.check(lib.tnx_mix(_vp(dj.data_ptr()), len(j), _vp(dt.data_ptr()), R.mix_blocks(j),
                        _vp(opsd.data_ptr()), None, _stream()), 'mix')
    torch.cuda.synchronize()
    for o, r in zip(outs, expect):
        assert np.allclose(o.cpu().numpy(), r, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize('sdt,ddt', [('float32', 'float32'), ('bfloat16', 'float32'),
                                     ('bfloat16', 'bfloat16'), ('float32', 'float64')])
def test_mix_dtypes_and_alignment(gpu, sdt, ddt):
    """Low-precision sources and destinations, through the vector path
    (rows of 4k pairs) and the scalar one (a source or destination that is
    not 4-element aligned, rows of other lengths): numpy on the rounded
    inputs, to the destination's rounding."""
    import torch
    from tnalg_amd import _lib as L
    from tnalg_amd import records as R
    lib = L.load()
    code = {'float64': L.F64, 'float32': L.F32, 'bfloat16': L.BF16}
    st, dtt = getattr(torch, sdt), getattr(torch, ddt)
    rng = np.random.RandomState(9)
    ops = rng.randn(3, 2, 2)
    opsd = _dev(ops.reshape(-1))
    mx = R.MixList()
    keep, expect, outs = [], [], []
    for (Dl, Dr, acc, shift) in [(16, 64, 0, 0), (33, 128, 1, 0), (8, 64, 0, 1), (5, 7, 1, 0),
                                 (128, 128, 1, 0)]:
        srcs = []
        for _ in range(2):
            buf = torch.from_numpy(rng.randn(Dl * 2 * Dr + shift)).to(st).cuda()
            keep.append(buf)
            srcs.append(buf[shift:])
        ref_src = [x.to(torch.float64).cpu().numpy().reshape(Dl, 2, Dr) for x in srcs]
        dbuf = torch.from_numpy(rng.randn(Dl * 2 * Dr + shift)).to(dtt).cuda()
        keep.append(dbuf)
        dst = dbuf[shift:]
        d0 = dst.to(torch.float64).cpu().numpy().reshape(Dl, 2, Dr)
        ref = 0.75 * ref_src[0] + np.einsum('st,atb->asb', ops[1], ref_src[1]) * (-2.0)
        if acc:
            ref = ref + d0
        expect.append(ref)
        outs.append(dst)
        mx.add(dst.data_ptr(), code[ddt], Dl, 2, Dr,
               [(srcs[0].data_ptr(), code[sdt], 0.75, -1), (srcs[1].data_ptr(), code[sdt], -2.0, 1)],
               accumulate=acc)
    j, t = mx.arrays()
    dj, dtd = _upload([j, t])
    print('mix dtypes', sdt, ddt, flush=True)
    L.check(lib.tnx_mix(_vp(dj.data_ptr()), len(j), _vp(dtd.data_ptr()), R.mix_blocks(j),
                        _vp(opsd.data_ptr()), None, _stream()), 'mix')
    torch.cuda.synchronize()
    tol = {'float64': 1e-13, 'float32': 1e-6, 'bfloat16': 1e-2}[ddt]
    for o, r in zip(outs, expect):
        got = o.to(torch.float64).cpu().numpy().reshape(r.shape)
        assert np.allclose(got, r, rtol=tol, atol=tol * np.abs(r).max()), (sdt, ddt, r.shape)


@pytest.mark.parametrize('m,n,trans', [(4, 2, 0), (2, 16, 0), (256, 128, 0), (256, 128, 1),
                                       (100, 37, 1), (64, 64, 0), (8, 8, 1)])
def test_qr_f64(gpu, m, n, trans):
    import torch
    from tnalg_amd import _lib as L
    lib = L.load()
    rng = np.random.RandomState(m * 31 + n)
    M = rng.randn(m, n)
    M[:, -1] = M[:, 0]  # rank deficient column
    src = M.T.copy() if trans else M.copy()
    k = min(m, n)
    srcd = _dev(src)
    q = torch.zeros(m * k, dtype=torch.float64, device='cuda')
    r = torch.zeros(k * n, dtype=torch.float64, device='cuda')
    work = torch.zeros(m * n + m * k, dtype=torch.float64, device='cuda')
    ld = m if trans else n
    job = np.array([(srcd.data_ptr(), q.data_ptr(), r.data_ptr(), work.data_ptr(), m, n, ld, trans,
                     trans, 0)], dtype=L.QR_JOB)
    (dj,) = _upload([job])
    print('qr', m, n, trans, flush=True)
    L.check(lib.tnx_qr_f64(_vp(dj.data_ptr()), 1, _stream()), 'qr')
    torch.cuda.synchronize()
    Q = q.cpu().numpy().reshape(k, m).T if trans else q.cpu().numpy().reshape(m, k)
    Rm = r.cpu().numpy().reshape(k, n)
    assert np.allclose(Q.T @ Q, np.eye(k), atol=1e-13)
    assert np.allclose(Q @ Rm, M, atol=1e-12 * np.abs(M).max())
    assert np.allclose(np.tril(Rm, -1), 0)


@pytest.mark.parametrize('m,n,trans', [(4, 2, 0), (256, 128, 0), (256, 128, 1), (33, 33, 0),
                                       (96, 17, 1)])
def test_svd_jacobi_f64(gpu, m, n, trans):
    import torch
    from tnalg_amd import _lib as L
    lib = L.load()
    rng = np.random.RandomState(m + 7 * n)
    M = rng.randn(m, n) @ np.diag(np.logspace(0, -9, n))
    src = M.T.copy() if trans else M.copy()
    srcd = _dev(src)
    u = torch.zeros(m * n, dtype=torch.float64, device='cuda')
    s = torch.zeros(n, dtype=torch.float64, device='cuda')
    svt = torch.zeros(n * n, dtype=torch.float64, device='cuda')
    work = torch.zeros(m * n + n * n, dtype=torch.float64, device='cuda')
    job = np.array([(srcd.data_ptr(), u.data_ptr(), s.data_ptr(), svt.data_ptr(), work.data_ptr(),
                     m, n, m if trans else n, trans, trans, 0)], dtype=L.SVD_JOB)
    (dj,) = _upload([job])
    print('svd', m, n, trans, flush=True)
    L.check(lib.tnx_svd_jacobi_f64(_vp(dj.data_ptr()), 1, 60, _stream()), 'svd')
    torch.cuda.synchronize()
    U = u.cpu().numpy().reshape(n, m).T if trans else u.cpu().numpy().reshape(m, n)
    sv = s.cpu().numpy()
    SVT = svt.cpu().numpy().reshape(n, n)
    ref = np.linalg.svd(M, compute_uv=False)
    assert np.allclose(sv, ref, rtol=1e-10, atol=1e-14 * ref[0])
    assert np.all(np.diff(sv) <= 0)
    assert np.allclose(U.T @ U, np.eye(n), atol=1e-10)
    assert np.allclose(U @ SVT, M, atol=1e-12 * np.abs(M).max())


def test_chol_inv_f64(gpu):
    """One launch over Gram matrices of every block shape (n = 1 .. 128,
    partial 16-blocks, leading dimension above n, one matrix with
    condition 1e12): R upper with a positive diagonal, R^T R = S + shift I
    to rounding, R R^{-1} = I to the conditioning of R, R as numpy's
    Cholesky factor (well-conditioned cases), status 0, nothing written
    beyond column n."""
    import torch
    from tnalg_amd import _lib as L
    lib = L.load()
    shift = 1e-12
    cases = [(1, 1, 0), (8, 8, 0), (16, 19, 0), (17, 17, 0), (37, 40, 0), (64, 64, 0),
             (100, 100, 0), (127, 130, 0), (128, 128, 0), (128, 128, 1)]
    bufs, jobs, mats = [], [], []
    for idx, (n, ld, ill) in enumerate(cases):
        rng = np.random.RandomState(100 + idx)
        A = rng.randn(2 * n + 3, n)
        if ill:
            A = A @ np.diag(np.logspace(0, -6, n))
        G = np.zeros((n, ld))
        G[:, :n] = A.T @ A
        Gd = _dev(G.ravel())
        r = torch.full((n * ld,), 7.0, dtype=torch.float64, device='cuda')
        ri = torch.full((n * ld,), 7.0, dtype=torch.float64, device='cuda')
        st = torch.full((1,), 5, dtype=torch.int32, device='cuda')
        bufs.append((Gd, r, ri, st))
        mats.append(G[:, :n])
        jobs.append((Gd.data_ptr(), r.data_ptr(), ri.data_ptr(), st.data_ptr(), 0, n, ld, shift))
    (dj,) = _upload([np.array(jobs, dtype=L.CHOL_JOB)])
    print('chol_inv batch', [c[0] for c in cases], flush=True)
    L.check(lib.tnx_chol_inv_f64(_vp(dj.data_ptr()), len(jobs), _stream()), 'chol')
    torch.cuda.synchronize()
    for (n, ld, ill), G, (Gd, r, ri, st) in zip(cases, mats, bufs):
        Rm = r.cpu().numpy().reshape(n, ld)
        Ri = ri.cpu().numpy().reshape(n, ld)
        assert np.all(Rm[:, n:] == 7.0) and np.all(Ri[:, n:] == 7.0), n   # padding untouched
        Rm, Ri = Rm[:, :n], Ri[:, :n]
        S = G + shift * np.trace(G) * np.eye(n)
        assert np.array_equal(np.tril(Rm, -1), np.zeros((n, n))) and np.all(np.diag(Rm) > 0), n
        assert np.array_equal(np.tril(Ri, -1), np.zeros((n, n))), n
        assert np.allclose(Rm.T @ Rm, S, rtol=0, atol=1e-13 * np.abs(S).max()), n
        cond = np.linalg.cond(Rm)
        assert np.abs(Rm @ Ri - np.eye(n)).max() < 1e-14 * max(cond, 10.0), (n, cond)
        if not ill:   # forward error of the factor: a few n u cond(S) apart from numpy's
            Rn = np.linalg.cholesky(S).T
            assert np.abs(Rm - Rn).max() < 1e-14 * cond ** 2 * np.abs(Rn).max(), n
        assert int(st.cpu()[0]) == 0, n


def test_chol_inv_indefinite_status(gpu):
    """A matrix that is not positive definite sets the job's status."""
    import torch
    from tnalg_amd import _lib as L
    lib = L.load()
    n = 40
    S = np.eye(n)
    S[23, 23] = -1.0
    Gd = _dev(S)
    r = torch.zeros(n * n, dtype=torch.float64, device='cuda')
    ri = torch.zeros(n * n, dtype=torch.float64, device='cuda')
    st = torch.zeros(1, dtype=torch.int32, device='cuda')
    jobs = np.array([(Gd.data_ptr(), r.data_ptr(), ri.data_ptr(), st.data_ptr(), 0, n, n, 0.0)],
                    dtype=L.CHOL_JOB)
    (dj,) = _upload([jobs])
    print('chol_inv indefinite', flush=True)
    L.check(lib.tnx_chol_inv_f64(_vp(dj.data_ptr()), 1, _stream()), 'chol')
    torch.cuda.synchronize()
    assert int(st.cpu()[0]) == 1


@pytest.mark.parametrize('eps', [1e-15, 3e-11, 1e-8])
def test_chol_inv_near_identity(gpu, eps):
    """A Gram matrix I + E with a tiny symmetric E (the last CholeskyQR3
    pass) takes the first-order path (|E| <= 1e-10) or the factorisation;
    both give R^T R = S and R R^{-1} = I to rounding and numpy's factor."""
    import torch
    from tnalg_amd import _lib as L
    lib = L.load()
    n = 100
    rng = np.random.RandomState(17)
    E = rng.randn(n, n)
    E = eps * (E + E.T) / 2
    S = np.eye(n) + E
    Gd = _dev(S)
    r = torch.zeros(n * n, dtype=torch.float64, device='cuda')
    ri = torch.zeros(n * n, dtype=torch.float64, device='cuda')
    jobs = np.array([(Gd.data_ptr(), r.data_ptr(), ri.data_ptr(), 0, 0, n, n, 0.0)],
                    dtype=L.CHOL_JOB)
    (dj,) = _upload([jobs])
    print('chol_inv near identity eps', eps, flush=True)
    L.check(lib.tnx_chol_inv_f64(_vp(dj.data_ptr()), 1, _stream()), 'chol')
    torch.cuda.synchronize()
    Rm = r.cpu().numpy().reshape(n, n)
    Ri = ri.cpu().numpy().reshape(n, n)
    tol = 2e-15 if eps <= 1e-10 else 1e-14
    assert np.array_equal(np.tril(Rm, -1), np.zeros((n, n)))
    assert np.allclose(Rm.T @ Rm, S, rtol=0, atol=tol)
    assert np.allclose(Rm @ Ri, np.eye(n), rtol=0, atol=tol)
    assert np.allclose(Rm, np.linalg.cholesky(S).T, rtol=0, atol=tol)


def test_svd_block_jacobi(gpu):
    """Batched block-Jacobi SVD of triangular factors of mixed orders (one
    launch per round over all jobs): R = u^T v, u orthogonal, s the singular
    values of R in descending order (numpy's SVD), a rank-deficient R kept
    exact."""
    import torch
    from tnalg_amd import _lib as L
    lib = L.load()
    rng = np.random.RandomState(5)
    Rs = []
    for n in (8, 37, 128, 300, 512):
        A = rng.randn(2 * n, n) @ np.diag(np.logspace(0, -10, n))
        Rs.append(np.linalg.qr(A)[1])
    Rd = np.triu(rng.randn(24, 24))
    Rd[9:] = 0.0           # rank 9
    Rs.append(Rd)
    # singular values down to 1e-22 of the largest: the tail is rounding
    # residue, which must not keep the sweeps going (a non-zero return fails)
    A = rng.randn(256, 256) @ np.diag(np.logspace(0, -22, 256)) @ np.linalg.qr(rng.randn(256, 256))[0]
    Rs.append(np.linalg.qr(A)[1])
    bufs, recs = [], []
    for R in Rs:
        n = R.shape[0]
        bw, nblk = L.svdj_blocking(n)
        r = _dev(R)
        work = torch.zeros(2 * n * n + n + 1, dtype=torch.float64, device='cuda')
        u = torch.zeros(n * n, dtype=torch.float64, device='cuda')
        v = torch.zeros(n * n, dtype=torch.float64, device='cuda')
        s = torch.zeros(n, dtype=torch.float64, device='cuda')
        bufs.append((r, work, u, v, s))
        recs.append((r.data_ptr(), work.data_ptr(), u.data_ptr(), v.data_ptr(), s.data_ptr(),
                     n, n, bw, nblk, 0, 0))
    jobs = np.array(recs, dtype=L.SVDJ_JOB)
    jobs['pair_begin'] = np.concatenate([[0], np.cumsum(jobs['nblk'] // 2)[:-1]])
    (dj,) = _upload([jobs])
    flag = torch.zeros(4, dtype=torch.int32, device='cuda')
    print('svd_block_jacobi orders', [R.shape[0] for R in Rs], flush=True)
    L.check(lib.tnx_svd_block_jacobi(_vp(dj.data_ptr()), len(jobs), int((jobs['nblk'] // 2).sum()),
                                     int(jobs['nblk'].max()), int(jobs['n'].max()),
                                     _vp(flag.data_ptr()), 60, _stream()), 'svd_block_jacobi')
    torch.cuda.synchronize()
    for R, (_, _, u, v, s) in zip(Rs, bufs):
        n = R.shape[0]
        U = u.cpu().numpy().reshape(n, n)
        Vs = v.cpu().numpy().reshape(n, n)
        sv = s.cpu().numpy()
        ref = np.linalg.svd(R, compute_uv=False)
        assert np.abs(U.T @ Vs - R).max() < 1e-12 * np.abs(R).max() * n, n
        assert np.abs(U @ U.T - np.eye(n)).max() < 1e-12 * n, n
        assert np.all(np.diff(sv) <= 0), n
        assert np.abs(sv - ref).max() < 1e-13 * ref[0] * n, n
        assert np.allclose(np.linalg.norm(Vs, axis=1), sv, rtol=1e-12, atol=1e-14), n


def test_trace_dot(gpu):
    import torch
    from tnalg_amd import _lib as L
    lib = L.load()
    rng = np.random.RandomState(1)
    X, Y = rng.randn(37, 37), rng.randn(37, 37)
    Xd, Yd = _dev(X), _dev(Y)
    jobs = np.array([(Xd.data_ptr(), Yd.data_ptr(), 37, 37, 37, 37), (Xd.data_ptr(), 0, 37, 37, 37, 0)],
                    dtype=L.DOT_JOB)
    (dj,) = _upload([jobs])
    out = torch.zeros(2, dtype=torch.float64, device='cuda')
    L.check(lib.tnx_trace_dot(_vp(dj.data_ptr()), 2, _vp(out.data_ptr()), _stream()), 'dot')
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert abs(o[0] - np.sum(X * Y)) < 1e-11 and abs(o[1] - np.trace(X)) < 1e-12


def test_lanczos_dense_symmetric(gpu):
    """Lowest eigenpairs of a batch of dense symmetric matrices (matvec = one
    GEMM per solve) against numpy.linalg.eigh."""
    import torch
    from tnalg_amd import _lib as L
    from tnalg_amd import records as R
    lib = L.load()
    rng = np.random.RandomState(11)
    sizes = [1, 3, 16, 200, 2500, 5000]
    ncv = 20
    keep, solves, refs = [], [], []
    g2 = R.GemmList()
    for k, n in enumerate(sizes):
        A = rng.randn(n, n)
        diag = np.linspace(0.0, 10.0, n)
        diag[0] = -2.0
        H = 0.002 * (A + A.T) / 2 + np.diag(diag)
        w_ref, v_ref = np.linalg.eigh(H)
        refs.append((w_ref[0], v_ref[:, 0], H))
        Hd = _dev(H)
        V = torch.zeros((ncv + 1) * n, dtype=torch.float64, device='cuda')
        w = torch.zeros(n, dtype=torch.float64, device='cuda')
        x = torch.zeros(n, dtype=torch.float64, device='cuda')
        init = _dev(rng.randn(n))
        out = torch.zeros(n, dtype=torch.float64, device='cuda')
        keep += [Hd, V, w, x, init, out]
        g2.add(w.data_ptr(), n, 1, 1, [(Hd.data_ptr(), n, 0, x.data_ptr(), 1, 0, n, 1.0)], active=k)
        solves.append((V.data_ptr(), w.data_ptr(), x.data_ptr(), init.data_ptr(), out.data_ptr(), n,
                       n, min(ncv, n), 0, 1e-12 / 1e-4, 1e-4, 0.0, 0, 0))
    lz = L.lanczos_records(solves)
    ch = (lz['n'] + L.LZ_CHUNK - 1) // L.LZ_CHUNK
    lz['chunk_begin'] = np.concatenate([[0], np.cumsum(ch)[:-1]])
    nch = int(ch.sum())
    p, s = g2.arrays()
    dp, ds, dlz = _upload([p, s, lz])
    plan = L.MatvecPlan(0, 0, 0, 0, 0, 0, 0, 0, dp.data_ptr(), ds.data_ptr(), len(p),
                        R.gemm_tiles(p), 0, 0, 0)
    state = torch.zeros(int(lib.tnx_lanczos_state_bytes(len(sizes), nch, ncv)), dtype=torch.uint8,
                        device='cuda')
    res = np.zeros(len(sizes), dtype=L.LANCZOS_RESULT)
    print('lanczos sizes', sizes, flush=True)
    rc = lib.tnx_lanczos_run_ex(plan, _vp(dlz.data_ptr()), len(sizes), nch, ncv, 50, 1, 0, 0,
                                _vp(state.data_ptr()), _vp(res.ctypes.data), _stream())
    assert rc == 0, (rc, lib.tnx_last_error())
    for k, n in enumerate(sizes):
        e0, v0, H = refs[k]
        out = keep[6 * k + 5].cpu().numpy()
        assert abs(res[k]['theta'] - e0) < 1e-9 * max(1, abs(e0)), (n, res[k], e0)
        assert abs(abs(out @ v0) - 1) < 1e-8, n
        assert np.linalg.norm(H @ out - e0 * out) < 1e-6 * max(1, abs(e0))


def test_lanczos_basis_stays_orthonormal(gpu):
    """A full cycle past convergence (every step forced, no restart) of
    matrices whose lowest eigenvalue is well separated, the case in which a
    Lanczos basis loses orthogonality first: the ncv + 1 basis vectors stay
    orthonormal to 1e-10 (the second projection is applied whenever it
    exceeds 1e-12 of the new coupling, and left out otherwise), and the
    tridiagonal's lowest eigenvalue matches numpy.linalg.eigh."""
    import torch
    from tnalg_amd import _lib as L
    from tnalg_amd import records as R
    lib = L.load()
    rng = np.random.RandomState(5)
    n, ncv, ns = 3000, 20, 3
    keep, solves, refs = [], [], []
    g2 = R.GemmList()
    for k in range(ns):
        diag = np.linspace(0.0, 10.0, n)
        diag[0] = -5.0 - k
        A = rng.randn(n, n)
        H = 0.01 * (A + A.T) / 2 + np.diag(diag)
        refs.append(np.linalg.eigvalsh(H)[0])
        Hd = _dev(H)
        V = torch.zeros(ncv + 1, n, dtype=torch.float64, device='cuda')
        w, x, out = (torch.zeros(n, dtype=torch.float64, device='cuda') for _ in range(3))
        init = _dev(rng.randn(n))
        keep += [Hd, V, w, x, init, out]
        g2.add(w.data_ptr(), n, 1, 1, [(Hd.data_ptr(), n, 0, x.data_ptr(), 1, 0, n, 1.0)], active=k)
        solves.append((V.data_ptr(), w.data_ptr(), x.data_ptr(), init.data_ptr(), out.data_ptr(), n,
                       n, 0, 0, 1e-30, 1e-4, 0.0, 0, 0))
    lz = L.lanczos_records(solves)
    ch = (lz['n'] + L.LZ_CHUNK - 1) // L.LZ_CHUNK
    lz['chunk_begin'] = np.concatenate([[0], np.cumsum(ch)[:-1]])
    nch = int(ch.sum())
    p, s = g2.arrays()
    dp, ds, dlz = _upload([p, s, lz])
    plan = L.MatvecPlan(0, 0, 0, 0, 0, 0, 0, 0, dp.data_ptr(), ds.data_ptr(), len(p),
                        R.gemm_tiles(p), 0, 0, 0)
    state = torch.zeros(int(lib.tnx_lanczos_state_bytes(ns, nch, ncv)), dtype=torch.uint8,
                        device='cuda')
    res = np.zeros(ns, dtype=L.LANCZOS_RESULT)
    print('lanczos basis orthogonality n', n, 'ncv', ncv, flush=True)
    rc = lib.tnx_lanczos_run_ex(plan, _vp(dlz.data_ptr()), ns, nch, ncv, 0, ncv, 0, 0,
                                _vp(state.data_ptr()), _vp(res.ctypes.data), _stream())
    assert rc == L.ENOCONV, (rc, lib.tnx_last_error())
    for k in range(ns):
        assert res[k]['iters'] == ncv, res[k]
        Vh = keep[6 * k + 1].cpu().numpy()
        G = Vh @ Vh.T
        assert np.abs(G - np.eye(ncv + 1)).max() < 1e-10, np.abs(G - np.eye(ncv + 1)).max()
        assert abs(res[k]['theta'] - refs[k]) < 1e-10 * abs(refs[k]), (res[k], refs[k])


def test_lanczos_f32_basis(gpu):
    """Solves whose basis is stored in f32 (v_dtype): a full forced cycle
    keeps the stored rows orthonormal to f32 level, and a converging solve
    gives the lowest eigenpair to f32 accuracy, beside the same problem with
    an f64 basis in the same call."""
    import torch
    from tnalg_amd import _lib as L
    from tnalg_amd import records as R
    lib = L.load()
    rng = np.random.RandomState(8)
    n, ncv = 3000, 20
    diag = np.linspace(0.0, 10.0, n)
    diag[0] = -4.0
    A = rng.randn(n, n)
    H = 0.01 * (A + A.T) / 2 + np.diag(diag)
    wref, vref = np.linalg.eigh(H)
    Hd = _dev(H)
    init0 = rng.randn(n)
    keep, solves = [], []
    g2 = R.GemmList()
    # (f32 basis, forced full cycle), (f32 basis, converging), (f64 basis, converging)
    for k, (vdt, tol) in enumerate([(L.F32, 1e-30), (L.F32, 1e-9), (L.F64, 1e-9)]):
        V = torch.zeros((ncv + 1) * n, dtype=torch.float64, device='cuda')
        w, x, out = (torch.zeros(n, dtype=torch.float64, device='cuda') for _ in range(3))
        init = _dev(init0)
        keep.append((V, out))
        keep += [w, x, init]
        g2.add(w.data_ptr(), n, 1, 1, [(Hd.data_ptr(), n, 0, x.data_ptr(), 1, 0, n, 1.0)], active=k)
        solves.append((V.data_ptr(), w.data_ptr(), x.data_ptr(), init.data_ptr(), out.data_ptr(), n,
                       n, 0, 0, tol / 1e-4, 1e-4, 0.0, 0, 0, vdt, 0))
    lz = L.lanczos_records(solves)
    ch = (lz['n'] + L.LZ_CHUNK - 1) // L.LZ_CHUNK
    lz['chunk_begin'] = np.concatenate([[0], np.cumsum(ch)[:-1]])
    nch = int(ch.sum())
    p, s_ = g2.arrays()
    dp, ds, dlz = _upload([p, s_, lz])
    plan = L.MatvecPlan(0, 0, 0, 0, 0, 0, 0, 0, dp.data_ptr(), ds.data_ptr(), len(p),
                        R.gemm_tiles(p), 0, 0, 0)
    state = torch.zeros(int(lib.tnx_lanczos_state_bytes(3, nch, ncv)), dtype=torch.uint8,
                        device='cuda')
    res = np.zeros(3, dtype=L.LANCZOS_RESULT)
    print('lanczos f32 basis n', n, flush=True)
    rc = lib.tnx_lanczos_run_ex(plan, _vp(dlz.data_ptr()), 3, nch, ncv, 30, 1, 0, 0,
                                _vp(state.data_ptr()), _vp(res.ctypes.data), _stream())
    assert rc in (L.OK, L.ENOCONV), (rc, lib.tnx_last_error())
    torch.cuda.synchronize()
    # the forced solve: ncv + 1 stored f32 rows, orthonormal to f32 level
    Vf = keep[0][0].cpu().numpy().view(np.float32)[:(ncv + 1) * n].reshape(ncv + 1, n)
    Vf = Vf.astype(np.float64)
    if res[0]['iters'] == ncv:   # no restart happened: every row is a Lanczos vector
        G = Vf @ Vf.T
        assert np.abs(G - np.eye(ncv + 1)).max() < 2e-6, np.abs(G - np.eye(ncv + 1)).max()
    # both converging solves stop at |r| <= 1e-5 (the same test); the f32
    # basis leaves the eigenvector's direction as accurate as that test asks
    # and its norm within the f32 rounding of the basis rows
    for k in (1, 2):
        assert res[k]['converged'] == 1, res[k]
        v = keep[4 * k][1].cpu().numpy()
        nv = np.linalg.norm(v)
        assert abs(nv - 1) < (1e-6 if k == 1 else 1e-12), (k, nv)
        v = v / nv
        th = res[k]['theta']
        assert abs(th - wref[0]) < 1e-7 * np.abs(wref).max(), (k, th, wref[0])
        assert abs(abs(v @ vref[:, 0]) - 1) < 1e-10, (k, abs(v @ vref[:, 0]) - 1)
        assert np.linalg.norm(H @ v - th * v) < 2e-5, (k, np.linalg.norm(H @ v - th * v))


def test_profile_times_the_first_krylov_step(gpu):
    """Sampled kernel timing around one Lanczos call (every call sampled):
    each matvec GEMM, k_dots, k_update_dots and k_update_norm launch of the
    first Krylov step is timed once by its blocks' wall-clock stamps, with
    its algorithmic work; launches are counted
    for every step issued (the host polls one step behind, so one step may
    be issued after the last one the solve needed)."""
    import torch
    from tnalg_amd import _lib as L
    from tnalg_amd import records as R
    lib = L.load()
    rng = np.random.RandomState(4)
    n, ncv = 3000, 20
    A = rng.randn(n, n)
    H = (A + A.T) / 2 + np.diag(np.linspace(0, 50, n))
    Hd = _dev(H)
    V = torch.zeros((ncv + 1) * n, dtype=torch.float64, device='cuda')
    w, x = (torch.zeros(n, dtype=torch.float64, device='cuda') for _ in range(2))
    init, out = _dev(rng.randn(n)), torch.zeros(n, dtype=torch.float64, device='cuda')
    g2 = R.GemmList()
    g2.add(w.data_ptr(), n, 1, 1, [(Hd.data_ptr(), n, 0, x.data_ptr(), 1, 0, n, 1.0)], active=0)
    lz = L.lanczos_records([(V.data_ptr(), w.data_ptr(), x.data_ptr(), init.data_ptr(),
                            out.data_ptr(), n, n, 0, 0, 1e-10 / 1e-4, 1e-4, 0.0, 0, 0)])
    nch = -(-n // L.LZ_CHUNK)
    p, s = g2.arrays()
    dp, ds, dlz = _upload([p, s, lz])
    flops = 2.0 * n * n
    plan = L.MatvecPlan(0, 0, 0, 0, 0, 0, 0, 0, dp.data_ptr(), ds.data_ptr(), len(p),
                        R.gemm_tiles(p), 0, 0, 0, (ctypes.c_double * 3)(0.0, 0.0, flops))
    state = torch.zeros(int(lib.tnx_lanczos_state_bytes(1, nch, ncv)), dtype=torch.uint8,
                        device='cuda')
    res = np.zeros(1, dtype=L.LANCZOS_RESULT)
    print('profiled lanczos n', n, flush=True)
    assert lib.tnx_profile_begin(1, 64) == 0
    rc = lib.tnx_lanczos_run_ex(plan, _vp(dlz.data_ptr()), 1, nch, ncv, 10, 1, 0, 0,
                                _vp(state.data_ptr()), _vp(res.ctypes.data), _stream())
    stats = (L.ProfStat * L.PROF_NCLASS)()
    assert lib.tnx_profile_end(stats, L.PROF_NCLASS) == L.PROF_NCLASS
    assert rc == 0
    by = {st.name.decode(): st for st in stats}
    steps = int(res[0]['iters'])
    g = by['gemm f64 matvec w = U R^T']
    assert g.samples == 1 and steps <= g.launches <= steps + 1
    assert g.work == flops and g.seconds > 0
    assert g.work_all == flops * g.launches      # every launch counted with its work
    # GEMM 1 has no problems: nothing is launched, so nothing is timed
    assert by['gemm f64 matvec T = F x'].samples == 0
    for k, words in (('k_dots', 8 + 8), ('k_update_dots', 8 + 8 + 8),
                     ('k_update_norm', 16 + 8)):
        st = by[k]
        assert st.samples == 1 and st.launches == g.launches and st.unit == 1, k
        assert st.work == nch * L.LZ_CHUNK * words, (k, st.work)
    assert by['mixer matvec'].samples == 0


def test_lanczos_multi_precision(gpu):
    """One tnx_lanczos_run_multi call over solves whose matvecs run at two
    precisions (an f64 and an f32 matvec plan, per-solve x/w dtypes): every
    solve's result is bit-identical to a separate single-precision call on
    its own group, and matches numpy.linalg.eigh (1e-9 for f64, the f32
    matvec floor for f32)."""
    import torch
    from tnalg_amd import _lib as L
    from tnalg_amd import records as R
    lib = L.load()
    rng = np.random.RandomState(5)
    sizes = [16, 300, 2000, 4100, 64]
    dts = [L.F64, L.F32, L.F64, L.F32, L.F32]
    ncv = 20
    tdt = {L.F64: torch.float64, L.F32: torch.float32}
    bufs, refs = [], []
    for k, (n, dt) in enumerate(zip(sizes, dts)):
        A = rng.randn(n, n)
        diag = np.linspace(0.0, 10.0, n)
        diag[0] = -2.0
        H = 0.002 * (A + A.T) / 2 + np.diag(diag)
        w_ref, v_ref = np.linalg.eigh(H)
        refs.append((w_ref[0], v_ref[:, 0]))
        bufs.append(dict(H=_dev(H, tdt[dt]), V=torch.zeros((ncv + 1) * n, dtype=torch.float64,
                                                              device='cuda'),
                         w=torch.zeros(n, dtype=tdt[dt], device='cuda'),
                         x=torch.zeros(n, dtype=tdt[dt], device='cuda'), init=_dev(rng.randn(n)),
                         out=torch.zeros(n, dtype=torch.float64, device='cuda')))

    def run(idx, multi):
        """idx: solve indices; multi: one call with a plan per dtype, else
        the call's own dtypes (all idx share one dtype)."""
        plans, keep = [], []
        for dt, mode in ((L.F64, L.GEMM_F64), (L.F32, L.GEMM_F32)):
            g = R.GemmList()
            for pos, k in enumerate(idx):
                if dts[k] == dt:
                    b, n = bufs[k], sizes[k]
                    g.add(b['w'].data_ptr(), n, 1, 1,
                          [(b['H'].data_ptr(), n, 0, b['x'].data_ptr(), 1, 0, n, 1.0)],
                          active=pos)
            p, s = g.arrays()
            if len(p):
                dp, ds = _upload([p, s])
                keep += [dp, ds]
                plans.append(L.MatvecPlan(0, 0, 0, 0, 0, 0, 0, 0, dp.data_ptr(), ds.data_ptr(),
                                          len(p), R.gemm_tiles(p), 0, 0, mode))
        recs = []
        for k in idx:
            b, n = bufs[k], sizes[k]
            floor = 2e-6 if dts[k] == L.F32 else 0.0
            recs.append((b['V'].data_ptr(), b['w'].data_ptr(), b['x'].data_ptr(),
                         b['init'].data_ptr(), b['out'].data_ptr(), n, n, min(ncv, n), 0,
                         1e-12 / 1e-4, 1e-4, floor, dts[k], dts[k]))
        lz = L.lanczos_records(recs)
        ch = (lz['n'] + L.LZ_CHUNK - 1) // L.LZ_CHUNK
        lz['chunk_begin'] = np.concatenate([[0], np.cumsum(ch)[:-1]])
        nch = int(ch.sum())
        (dlz,) = _upload([lz])
        state = torch.zeros(int(lib.tnx_lanczos_state_bytes(len(idx), nch, ncv)),
                            dtype=torch.uint8, device='cuda')
        res = np.zeros(len(idx), dtype=L.LANCZOS_RESULT)
        if multi:
            arr = (L.MatvecPlan * len(plans))(*plans)
            rc = lib.tnx_lanczos_run_multi(arr, len(plans), _vp(dlz.data_ptr()), len(idx), nch,
                                           ncv, 50, 1, _vp(state.data_ptr()),
                                           _vp(res.ctypes.data), _stream())
        else:
            assert len(plans) == 1
            dt = dts[idx[0]]
            rc = lib.tnx_lanczos_run_ex(plans[0], _vp(dlz.data_ptr()), len(idx), nch, ncv, 50, 1,
                                        dt, dt, _vp(state.data_ptr()), _vp(res.ctypes.data),
                                        _stream())
        assert rc == 0, (rc, lib.tnx_last_error())
        torch.cuda.synchronize()
        return res, [bufs[k]['out'].cpu().numpy() for k in idx]

    print('lanczos multi sizes', sizes, flush=True)
    allidx = list(range(len(sizes)))
    res_m, out_m = run(allidx, True)
    for dt in (L.F64, L.F32):
        idx = [k for k in allidx if dts[k] == dt]
        res_s, out_s = run(idx, False)
        for pos, k in enumerate(idx):
            assert res_s[pos]['theta'] == res_m[k]['theta'], (k, res_s[pos], res_m[k])
            assert res_s[pos]['iters'] == res_m[k]['iters'], k
            assert np.array_equal(out_s[pos], out_m[k]), k
    for k in allidx:
        e0, v0 = refs[k]
        tol = 1e-9 if dts[k] == L.F64 else 2e-5
        assert abs(res_m[k]['theta'] - e0) < tol * max(1, abs(e0)), (k, res_m[k], e0)
        assert abs(abs(out_m[k] @ v0) - 1) < (1e-8 if dts[k] == L.F64 else 1e-3), k


def test_lanczos_implicit_restart_clustered(gpu):
    """Lowest eigenpairs whose next eigenvalue lies 1e-4 .. 1e-6 above: a
    restart from the Ritz vector alone stalls on these (thousands of matvecs
    on a CPU model of it); the implicit restart keeps the lowest Ritz
    vectors and converges in a few hundred.  Checked against
    numpy.linalg.eigh: Ritz value, eigenvector and residual."""
    import torch
    from tnalg_amd import _lib as L
    from tnalg_amd import records as R
    lib = L.load()
    rng = np.random.RandomState(21)
    cases = [(600, 1e-4), (1000, 1e-5), (1500, 1e-6), (300, 0.5)]
    ncv = 20
    keep, solves, refs = [], [], []
    g2 = R.GemmList()
    for k, (n, gap) in enumerate(cases):
        d = np.sort(rng.rand(n)) * 10.0
        d[1] = d[0] + gap
        Q, _ = np.linalg.qr(rng.randn(n, n))
        H = (Q * d) @ Q.T
        H = (H + H.T) / 2
        refs.append((d[0], Q[:, :2], H))
        Hd = _dev(H)
        V = torch.zeros((ncv + 1) * n, dtype=torch.float64, device='cuda')
        w = torch.zeros(n, dtype=torch.float64, device='cuda')
        x = torch.zeros(n, dtype=torch.float64, device='cuda')
        init = _dev(rng.randn(n))
        out = torch.zeros(n, dtype=torch.float64, device='cuda')
        keep += [Hd, V, w, x, init, out]
        g2.add(w.data_ptr(), n, 1, 1, [(Hd.data_ptr(), n, 0, x.data_ptr(), 1, 0, n, 1.0)], active=k)
        solves.append((V.data_ptr(), w.data_ptr(), x.data_ptr(), init.data_ptr(), out.data_ptr(), n,
                       n, 0, 0, 1e-10 / 1e-4, 1e-4, 0.0, 0, 0))
    lz = L.lanczos_records(solves)
    ch = (lz['n'] + L.LZ_CHUNK - 1) // L.LZ_CHUNK
    lz['chunk_begin'] = np.concatenate([[0], np.cumsum(ch)[:-1]])
    nch = int(ch.sum())
    p, s = g2.arrays()
    dp, ds, dlz = _upload([p, s, lz])
    plan = L.MatvecPlan(0, 0, 0, 0, 0, 0, 0, 0, dp.data_ptr(), ds.data_ptr(), len(p),
                        R.gemm_tiles(p), 0, 0, 0)
    state = torch.zeros(int(lib.tnx_lanczos_state_bytes(len(cases), nch, ncv)), dtype=torch.uint8,
                        device='cuda')
    res = np.zeros(len(cases), dtype=L.LANCZOS_RESULT)
    print('lanczos clustered', cases, flush=True)
    rc = lib.tnx_lanczos_run_ex(plan, _vp(dlz.data_ptr()), len(cases), nch, ncv, 100, 1, 0, 0,
                                _vp(state.data_ptr()), _vp(res.ctypes.data), _stream())
    assert rc == 0, (rc, lib.tnx_last_error(), res)
    for k, (n, gap) in enumerate(cases):
        e0, v01, H = refs[k]
        v = keep[6 * k + 5].cpu().numpy()
        assert res[k]['converged'] == 1
        assert abs(res[k]['theta'] - e0) < 1e-9 * max(1.0, abs(e0)), (n, gap, res[k], e0)
        assert abs(np.linalg.norm(v) - 1) < 1e-12
        assert np.linalg.norm(H @ v - res[k]['theta'] * v) < 2e-6, (n, gap)
        # the eigenvector: within residual/gap of the lowest one, i.e. in
        # the span of the two lowest for the clustered cases
        c = v @ v01
        if gap >= 0.1:
            assert abs(abs(c[0]) - 1) < 1e-9, (n, gap, c)
        else:
            assert 1 - (c @ c) < 1e-6, (n, gap, c)
        assert res[k]['iters'] < 1200, (n, gap, res[k])


def test_fill_normal_stream(gpu):
    """The counter-based normal fill: the oracle's restatement of the
    stream (Philox4x32-10, pinned by its known-answer vectors, and
    Box-Muller) at the start and past 2^32 pairs, standard normal moments,
    the same values however the stream is cut into jobs (pair boundaries
    included), a different stream for another seed."""
    import torch
    from tnalg_amd import _lib as L
    lib = L.load()
    n = 1 << 20
    a = torch.zeros(n, dtype=torch.float64, device='cuda')
    b = torch.zeros(n, dtype=torch.float64, device='cuda')
    c = torch.zeros(n, dtype=torch.float64, device='cuda')
    cuts = [0, 1, 1000, 1001, 77777, n]
    jobs = [np.array([(a.data_ptr(), n, 12345, 0)], dtype=L.FILL_JOB),
            np.array([(b.data_ptr() + 8 * lo, hi - lo, 12345, lo)
                      for lo, hi in zip(cuts[:-1], cuts[1:])], dtype=L.FILL_JOB),
            np.array([(c.data_ptr(), n, 12346, 0)], dtype=L.FILL_JOB)]
    print('fill normal n', n, flush=True)
    for j in jobs:
        (dj,) = _upload([j])
        L.check(lib.tnx_fill_normal_f64(_vp(dj.data_ptr()), len(j), _stream()), 'fill')
        torch.cuda.synchronize()
    x, y, z = a.cpu().numpy(), b.cpu().numpy(), c.cpu().numpy()
    assert np.array_equal(x, y)
    assert not np.array_equal(x, z)
    assert abs(x.mean()) < 5e-3 and abs(x.std() - 1) < 5e-3
    assert abs(np.mean(x ** 4) - 3) < 0.05             # Gaussian kurtosis
    assert abs(np.corrcoef(x[0::2], x[1::2])[0, 1]) < 5e-3   # the two halves of a pair
    from oracle.rng_ref import normal
    far = (1 << 33) + 5
    d = torch.zeros(64, dtype=torch.float64, device='cuda')
    (dj,) = _upload([np.array([(d.data_ptr(), 64, (1 << 40) + 99, far)], dtype=L.FILL_JOB)])
    L.check(lib.tnx_fill_normal_f64(_vp(dj.data_ptr()), 1, _stream()), 'fill')
    torch.cuda.synchronize()
    ref = np.array([normal(12345, g) for g in range(256)])
    assert np.allclose(x[:256], ref, rtol=1e-14, atol=1e-14)
    ref = np.array([normal((1 << 40) + 99, far + g) for g in range(64)])
    assert np.allclose(d.cpu().numpy(), ref, rtol=1e-14, atol=1e-14)
