"""ctypes binding of ``libtnalg_hip.so`` (the C ABI of include/tnalg_hip.h).

The library is required: importing the engine on a machine where it has not
been built raises ``RuntimeError`` (there is no CPU fallback).  The numpy
structured dtypes below mirror the C descriptor structs field by field; their
sizes are checked against ``tnx_abi_sizes()`` at load time.
"""
import ctypes
import os

import numpy as np

from .build import LIB_PATH

F64, F32, BF16 = 0, 1, 2
GEMM_F64, GEMM_F32, GEMM_BF16_F32, GEMM_BF16_BF16 = 0, 1, 2, 3

GEMM_SEG = np.dtype([('a', '<u8'), ('b', '<u8'), ('lda', '<i4'), ('ldb', '<i4'),
                     ('transA', '<i4'), ('transB', '<i4'), ('K', '<i4'), ('reserved', '<i4'),
                     ('alpha', '<f8')])
GEMM_PROBLEM = np.dtype([('c', '<u8'), ('c0', '<u8'), ('M', '<i4'), ('N', '<i4'),
                         ('ldc', '<i4'), ('ldc0', '<i4'), ('seg_begin', '<i4'),
                         ('seg_count', '<i4'), ('tile_begin', '<i4'), ('active', '<i4'),
                         ('beta', '<f8')])
MIX_JOB = np.dtype([('dst', '<u8'), ('Dl', '<i4'), ('d', '<i4'), ('Dr', '<i4'),
                    ('dst_dtype', '<i4'), ('term_begin', '<i4'), ('term_count', '<i4'),
                    ('accumulate', '<i4'), ('active', '<i4'), ('block_begin', '<i4'),
                    ('reserved', '<i4')])
MIX_TERM = np.dtype([('src', '<u8'), ('coef', '<f8'), ('op', '<i4'), ('src_dtype', '<i4')])
LANCZOS_SOLVE = np.dtype([('V', '<u8'), ('w', '<u8'), ('x', '<u8'), ('init', '<u8'),
                          ('out', '<u8'), ('n', '<i4'), ('ldv', '<i4'), ('m', '<i4'),
                          ('chunk_begin', '<i4'), ('tol_scale', '<f8'), ('tau', '<f8'),
                          ('floor_rel', '<f8'), ('x_dtype', '<i4'), ('w_dtype', '<i4'),
                          ('v_dtype', '<i4'), ('reserved', '<i4')])


def lanczos_records(rows):
    """LANCZOS_SOLVE records from tuples of the first 14 fields (V ... w_dtype)
    or of all of them; the basis is f64 unless a row gives v_dtype."""
    full = len(LANCZOS_SOLVE.names)
    return np.array([tuple(r) + (0,) * (full - len(r)) for r in rows], dtype=LANCZOS_SOLVE)
LANCZOS_RESULT = np.dtype([('theta', '<f8'), ('resid', '<f8'), ('iters', '<i4'),
                           ('converged', '<i4')])
QR_JOB = np.dtype([('src', '<u8'), ('q', '<u8'), ('r', '<u8'), ('work', '<u8'),
                   ('m', '<i4'), ('n', '<i4'), ('ld', '<i4'), ('trans', '<i4'),
                   ('q_trans', '<i4'), ('reserved', '<i4')])
SVD_JOB = np.dtype([('src', '<u8'), ('u', '<u8'), ('s', '<u8'), ('svt', '<u8'), ('work', '<u8'),
                    ('m', '<i4'), ('n', '<i4'), ('ld', '<i4'), ('trans', '<i4'),
                    ('u_trans', '<i4'), ('reserved', '<i4')])
NORM_JOB = np.dtype([('x', '<u8'), ('n', '<i8')])
FILL_JOB = np.dtype([('dst', '<u8'), ('n', '<i8'), ('seed', '<u8'), ('offset', '<i8')])
KRON_TERM = np.dtype([('a', '<u8'), ('b', '<u8'), ('lda', '<i4'), ('ldb', '<i4'),
                      ('coef', '<f8')])
KRON_JOB = np.dtype([('dst', '<u8'), ('p', '<i4'), ('q', '<i4'), ('ldd', '<i4'),
                     ('term_begin', '<i4'), ('term_count', '<i4'), ('block_begin', '<i4')])
SYM_JOB = np.dtype([('x', '<u8'), ('n', '<i4'), ('ld', '<i4'), ('count', '<i4'),
                    ('stride', '<i4'), ('block_begin', '<i4'), ('reserved', '<i4')])
KRON_BLOCK = 256   # dst elements per workgroup of tnx_kron_f64 / tnx_symmetrize_f64
OP_REC = np.dtype([('kind', '<i4'), ('mode', '<i4'), ('n0', '<i4'), ('n1', '<i4'), ('n2', '<i4'),
                   ('n3', '<i4'), ('a0', '<u8'), ('a1', '<u8'), ('a2', '<u8')])
OP_GEMM, OP_MIX, OP_QR, OP_SVD, OP_CHOL, OP_SVDJ, OP_NORM = 1, 2, 3, 4, 5, 6, 7
DOT_JOB = np.dtype([('x', '<u8'), ('y', '<u8'), ('rows', '<i4'), ('cols', '<i4'),
                    ('ldx', '<i4'), ('ldy', '<i4')])
CHOL_JOB = np.dtype([('g', '<u8'), ('r', '<u8'), ('rinv', '<u8'), ('status', '<u8'),
                     ('shift_ref', '<u8'), ('n', '<i4'), ('ld', '<i4'), ('shift_scale', '<f8')])
SVDJ_JOB = np.dtype([('r', '<u8'), ('work', '<u8'), ('u', '<u8'), ('v', '<u8'), ('s', '<u8'),
                     ('n', '<i4'), ('ld_r', '<i4'), ('bw', '<i4'), ('nblk', '<i4'),
                     ('pair_begin', '<i4'), ('reserved', '<i4')])
SVDJ_NMAX = 1024
OK, EINVAL, ENOCONV = 0, -1, -3   # tnx_status values


def svdj_blocking(n, values_only=False):
    """(bw, nblk) of tnx_svd_block_jacobi for order n: bw the largest power of
    two <= min(32, 4096 / n) (at least 4), nblk = ceil(n / bw) rounded up to
    even.  Values-only jobs (no W in LDS): bw <= min(64, 8192 / n) and at most
    n / 2, so an order up to 128 is one block pair."""
    if values_only:
        bw = 64
        while bw > 4 and (bw * n > 8192 or 2 * bw > max(n, 8)):
            bw //= 2
    else:
        bw = 32
        while bw > 4 and bw * n > 4096:
            bw //= 2
    nblk = -(-n // bw)
    return bw, nblk + (nblk % 2)
CHOL_NMAX = 128

GEMM_TILE = 64
MIX_PAIRS = 512    # (a, b) pairs per mixer workgroup for d > 2 (1024 for d <= 2)


def mix_pairs(d):
    """(a, b) pairs per mixer workgroup of a job of physical dimension ``d``
    (scalar or array): 4 pairs per lane for d <= 2, 2 otherwise."""
    import numpy as _np
    return _np.where(_np.asarray(d) <= 2, 2 * MIX_PAIRS, MIX_PAIRS)


MIX_DMAX = 4       # largest physical dimension of the mixer
LZ_CHUNK = 1024
LZ_NCV_MAX = 32
QR_KMAX = 4096


class MatvecPlan(ctypes.Structure):
    _fields_ = [('g1_prob', ctypes.c_void_p), ('g1_seg', ctypes.c_void_p),
                ('g1_nprob', ctypes.c_int32), ('g1_ntiles', ctypes.c_int32),
                ('mix_job', ctypes.c_void_p), ('mix_term', ctypes.c_void_p),
                ('mix_njob', ctypes.c_int32), ('mix_nblocks', ctypes.c_int32),
                ('g2_prob', ctypes.c_void_p), ('g2_seg', ctypes.c_void_p),
                ('g2_nprob', ctypes.c_int32), ('g2_ntiles', ctypes.c_int32),
                ('ops', ctypes.c_void_p), ('mode', ctypes.c_int32), ('mode2', ctypes.c_int32),
                ('work', ctypes.c_double * 3)]


class ProfStat(ctypes.Structure):
    """One kernel class of the sampled kernel timing (tnx_prof_stat)."""
    _fields_ = [('name', ctypes.c_char * 48), ('launches', ctypes.c_int64),
                ('samples', ctypes.c_int64),
                ('seconds', ctypes.c_double), ('work', ctypes.c_double),
                ('work_all', ctypes.c_double),
                ('unit', ctypes.c_int32), ('reserved', ctypes.c_int32)]


PROF_NCLASS = 10


# exported symbol -> (restype, argtypes)
_vp, _i, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
SIGNATURES = {
    'tnx_version': (_i, []),
    'tnx_last_error': (ctypes.c_char_p, []),
    'tnx_abi_sizes': (_i, [_vp, _i]),
    'tnx_gemm_grouped': (_i, [_i, _vp, _i, _vp, _i, _vp, _vp]),
    'tnx_gemm_variants': (_i, [_i]),
    'tnx_mix': (_i, [_vp, _i, _vp, _i, _vp, _vp, _vp]),
    'tnx_heff_matvec': (_i, [ctypes.POINTER(MatvecPlan), _vp, _vp]),
    'tnx_lanczos_state_bytes': (_sz, [_i, _i, _i]),
    'tnx_lanczos_run': (_i, [ctypes.POINTER(MatvecPlan), _vp, _i, _i, _i, _i, _i, _i, _vp, _vp,
                             _vp]),
    'tnx_lanczos_run_ex': (_i, [ctypes.POINTER(MatvecPlan), _vp, _i, _i, _i, _i, _i, _i, _i, _vp,
                                _vp, _vp]),
    'tnx_lanczos_run_multi': (_i, [ctypes.POINTER(MatvecPlan), _i, _vp, _i, _i, _i, _i, _i, _vp,
                                   _vp, _vp]),
    'tnx_lanczos_run_multi_then': (_i, [ctypes.POINTER(MatvecPlan), _i, _vp, _i, _i, _i, _i, _i,
                                        _vp, _vp, _vp, _i, _vp, _vp]),
    'tnx_lanczos_run_multi_static': (_i, [ctypes.POINTER(MatvecPlan), _i, _vp, _i, _i, _i, _i,
                                          _i, _vp, _vp]),
    'tnx_qr_f64': (_i, [_vp, _i, _vp]),
    'tnx_svd_jacobi_f64': (_i, [_vp, _i, _i, _vp]),
    'tnx_chol_inv_f64': (_i, [_vp, _i, _vp]),
    'tnx_svd_block_jacobi': (_i, [_vp, _i, _i, _i, _i, _vp, _i, _vp]),
    'tnx_trace_dot': (_i, [_vp, _i, _vp, _vp]),
    'tnx_normalize_f64': (_i, [_vp, _i, _vp]),
    'tnx_fill_normal_f64': (_i, [_vp, _i, _vp]),
    'tnx_kron_f64': (_i, [_vp, _i, _vp, _i, _vp]),
    'tnx_symmetrize_f64': (_i, [_vp, _i, _i, _vp]),
    'tnx_run_ops': (_i, [_vp, _i, _vp]),
    'tnx_profile_begin': (_i, [_i, _i]),
    'tnx_profile_end': (_i, [ctypes.POINTER(ProfStat), _i]),
}

_LIB = None


class HipError(RuntimeError):
    pass


def load(path=None):
    """Load (once) and return the ctypes library; raise if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError('libtnalg_hip.so is not built (%s); run __graft_entry__.build() or '
                           'python -m tnalg_amd.build' % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    sizes = (ctypes.c_int32 * 32)()
    n = lib.tnx_abi_sizes(sizes, 32)
    want = [GEMM_SEG.itemsize, GEMM_PROBLEM.itemsize, MIX_JOB.itemsize, MIX_TERM.itemsize,
            ctypes.sizeof(MatvecPlan), LANCZOS_SOLVE.itemsize, LANCZOS_RESULT.itemsize,
            QR_JOB.itemsize, SVD_JOB.itemsize, DOT_JOB.itemsize, CHOL_JOB.itemsize,
            SVDJ_JOB.itemsize, ctypes.sizeof(ProfStat), NORM_JOB.itemsize, FILL_JOB.itemsize,
            KRON_TERM.itemsize, KRON_JOB.itemsize, SYM_JOB.itemsize, OP_REC.itemsize]
    got = list(sizes[:n])
    if got != want:
        raise RuntimeError('ABI size mismatch: library %s, binding %s' % (got, want))
    _LIB = lib
    return lib


def check(rc, what=''):
    if rc != 0:
        msg = load().tnx_last_error().decode(errors='replace')
        raise HipError('%s failed (rc=%d): %s' % (what, rc, msg))
    return rc
