"""A numpy stand-in for the C-ABI calls of the infinite-size driver, over
host memory (CPU torch tensors): it reads the same descriptor records at
the same addresses, so the driver's descriptors are checked on the CPU.
The Lanczos call is emulated with scipy's eigsh on the plan's matvec."""
import ctypes

import numpy as np
from scipy.sparse.linalg import LinearOperator, eigsh

from tnalg_amd import _lib as L


def _records(addr, dtype, n):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    buf = (ctypes.c_char * (dtype.itemsize * n)).from_address(int(addr))
    return np.frombuffer(buf, dtype=dtype, count=n)


def _mat(addr, rows, cols, ld):
    """rows x cols view with row stride ld (doubles) at addr."""
    size = (rows - 1) * ld + cols
    base = np.ctypeslib.as_array((ctypes.c_double * size).from_address(int(addr)))
    return np.lib.stride_tricks.as_strided(base, shape=(rows, cols), strides=(8 * ld, 8))


def _vec(addr, n):
    return np.ctypeslib.as_array((ctypes.c_double * n).from_address(int(addr)))


def _addr(p):
    return p.value if isinstance(p, ctypes.c_void_p) else int(p)


class EmulatedLib:
    def tnx_last_error(self):
        return b''

    def tnx_lanczos_state_bytes(self, nsolve, nchunks, ncv):
        return 256

    def _gemm(self, prob, nprob, seg):
        probs = _records(prob, L.GEMM_PROBLEM, nprob)
        segs = _records(seg, L.GEMM_SEG, int((probs['seg_begin'] + probs['seg_count']).max())
                        if nprob else 0)
        for p in probs:
            M, N = int(p['M']), int(p['N'])
            acc = np.zeros((M, N))
            for sg in segs[p['seg_begin']:p['seg_begin'] + p['seg_count']]:
                K = int(sg['K'])
                A = _mat(sg['a'], K, M, sg['lda']).T if sg['transA'] else _mat(sg['a'], M, K, sg['lda'])
                B = _mat(sg['b'], N, K, sg['ldb']).T if sg['transB'] else _mat(sg['b'], K, N, sg['ldb'])
                acc += sg['alpha'] * (A @ B)
            if p['c0'] and p['beta'] != 0.0:
                acc += p['beta'] * _mat(p['c0'], M, N, p['ldc0'])
            _mat(p['c'], M, N, p['ldc'])[:, :] = acc
        return 0

    def tnx_gemm_grouped(self, mode, prob, nprob, seg, ntiles, active, stream):
        return self._gemm(_addr(prob), nprob, _addr(seg))

    def tnx_kron_f64(self, jobs, njob, terms, nblocks, stream):
        js = _records(_addr(jobs), L.KRON_JOB, njob)
        ts = _records(_addr(terms), L.KRON_TERM, int((js['term_begin'] + js['term_count']).max()))
        for j in js:
            p, q = int(j['p']), int(j['q'])
            acc = np.zeros((p * q, p * q))
            for t in ts[j['term_begin']:j['term_begin'] + j['term_count']]:
                a = _mat(t['a'], p, p, t['lda']) if t['a'] else np.eye(p)
                b = _mat(t['b'], q, q, t['ldb']) if t['b'] else np.eye(q)
                acc += t['coef'] * np.kron(a, b)
            _mat(j['dst'], p * q, p * q, j['ldd'])[:, :] = acc
        return 0

    def tnx_symmetrize_f64(self, jobs, njob, nblocks, stream):
        for j in _records(_addr(jobs), L.SYM_JOB, njob):
            for m in range(int(j['count'])):
                X = _mat(int(j['x']) + 8 * m * int(j['stride']), j['n'], j['n'], j['ld'])
                X[:, :] = (X + X.T.copy()) / 2
        return 0

    def tnx_normalize_f64(self, jobs, njob, stream):
        for j in _records(_addr(jobs), L.NORM_JOB, njob):
            x = _vec(j['x'], int(j['n']))
            nrm = np.linalg.norm(x)
            if nrm > 0:
                x /= nrm
        return 0

    def tnx_svd_jacobi_f64(self, jobs, njob, sweeps, stream):
        for j in _records(_addr(jobs), L.SVD_JOB, njob):
            m, n = int(j['m']), int(j['n'])
            A = _mat(j['src'], n, m, j['ld']).T if j['trans'] else _mat(j['src'], m, n, j['ld'])
            u, s, vt = np.linalg.svd(A, full_matrices=False)
            _mat(j['u'], m, n, n)[:, :] = u
            _vec(j['s'], n)[:] = s
            _mat(j['svt'], n, n, n)[:, :] = s[:, None] * vt
        return 0

    def tnx_mix(self, jobs, njob, terms, nblocks, ops, active, stream):
        """Linear combinations only (every term's operator is the identity)."""
        js = _records(_addr(jobs), L.MIX_JOB, njob)
        ts = _records(_addr(terms), L.MIX_TERM, int((js['term_begin'] + js['term_count']).max()))
        for j in js:
            n = int(j['Dl']) * int(j['d']) * int(j['Dr'])
            acc = np.zeros(n)
            for t in ts[j['term_begin']:j['term_begin'] + j['term_count']]:
                assert t['op'] < 0
                acc += t['coef'] * _vec(t['src'], n)
            _vec(j['dst'], n)[:] = acc
        return 0

    def tnx_lanczos_run_ex(self, plan, solves, nsolve, nchunks, ncv, max_restarts, min_iters,
                           xd, wd, state, results, stream):
        pl = plan
        lz = _records(_addr(solves), L.LANCZOS_SOLVE, nsolve)
        res = _records(_addr(results), L.LANCZOS_RESULT, nsolve)
        res = np.frombuffer((ctypes.c_char * res.nbytes).from_address(_addr(results)),
                            dtype=L.LANCZOS_RESULT, count=nsolve)
        for k, D in enumerate(lz):
            n = int(D['n'])
            x, w = _vec(D['x'], n), _vec(D['w'], n)
            count = [0]

            def mv(v):
                x[:] = np.ravel(v)
                self._gemm(pl.g1_prob, pl.g1_nprob, pl.g1_seg)
                self._gemm(pl.g2_prob, pl.g2_nprob, pl.g2_seg)
                count[0] += 1
                return w.copy()
            op = LinearOperator((n, n), matvec=mv, dtype=np.float64)
            th, v = eigsh(op, 1, which='SA', v0=_vec(D['init'], n).copy(), tol=0)
            _vec(D['out'], n)[:] = v[:, 0]
            res[k] = (th[0], 0.0, count[0], 1)
        return 0
