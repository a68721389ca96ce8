"""Descriptor records for the C ABI and their concatenation across solves.

A solve contributes, per phase of a sweep step, small numpy record arrays
(GEMM problems + K-segments, mixer jobs + terms, QR/SVD jobs, Lanczos
solves) whose prefix-sum fields (``tile_begin``, ``seg_begin``,
``term_begin``, ``block_begin``, ``chunk_begin``) are local.  ``concat_*``
merges the arrays of all solves of a batch into one launch, rebasing those
fields.
"""
import numpy as np

from . import _lib as L

GEMM_SEG, GEMM_PROBLEM = L.GEMM_SEG, L.GEMM_PROBLEM
MIX_JOB, MIX_TERM = L.MIX_JOB, L.MIX_TERM

EMPTY_PROB = np.zeros(0, GEMM_PROBLEM)
EMPTY_SEG = np.zeros(0, GEMM_SEG)
EMPTY_JOB = np.zeros(0, MIX_JOB)
EMPTY_TERM = np.zeros(0, MIX_TERM)


class GemmList:
    """Accumulates GEMM problems: ``C = sum alpha op(A) op(B) + beta C0``."""

    def __init__(self):
        self.probs = []
        self.segs = []

    def add(self, c, M, N, ldc, segs, c0=0, ldc0=0, beta=0.0, active=-1):
        """``segs``: iterable of ``(a, lda, transA, b, ldb, transB, K, alpha)``."""
        if M <= 0 or N <= 0:
            return
        segs = [s for s in segs if s[6] > 0]
        start = len(self.segs)
        for (a, lda, ta, b, ldb, tb, K, alpha) in segs:
            self.segs.append((a, b, lda, ldb, ta, tb, K, 0, alpha))
        self.probs.append((c, c0, M, N, ldc, ldc0, start, len(segs), 0, active, beta))

    def arrays(self):
        p = np.array(self.probs, dtype=GEMM_PROBLEM) if self.probs else EMPTY_PROB.copy()
        s = np.array(self.segs, dtype=GEMM_SEG) if self.segs else EMPTY_SEG.copy()
        if len(p):
            tiles = ((p['M'] + L.GEMM_TILE - 1) // L.GEMM_TILE) * \
                ((p['N'] + L.GEMM_TILE - 1) // L.GEMM_TILE)
            p['tile_begin'] = np.concatenate([[0], np.cumsum(tiles)[:-1]])
        return p, s


class MixList:
    """Accumulates mixer jobs ``dst[a,s,b] (=|+=) sum coef O[s,t] src[a,t,b]``."""

    def __init__(self):
        self.jobs = []
        self.terms = []

    def add(self, dst, dst_dtype, Dl, d, Dr, terms, accumulate=0, active=-1):
        """``terms``: iterable of ``(src, src_dtype, coef, op)`` (op = -1: identity)."""
        if Dl * d * Dr <= 0:
            return
        start = len(self.terms)
        for (src, sdt, coef, op) in terms:
            self.terms.append((src, coef, op, sdt))
        self.jobs.append((dst, Dl, d, Dr, dst_dtype, start, len(self.terms) - start, accumulate,
                          active, 0, 0))

    def arrays(self):
        j = np.array(self.jobs, dtype=MIX_JOB) if self.jobs else EMPTY_JOB.copy()
        t = np.array(self.terms, dtype=MIX_TERM) if self.terms else EMPTY_TERM.copy()
        if len(j):
            blocks = (j['Dl'].astype(np.int64) * j['d'] * j['Dr'] + L.MIX_BLOCK - 1) // L.MIX_BLOCK
            j['block_begin'] = np.concatenate([[0], np.cumsum(blocks)[:-1]])
        return j, t


def gemm_tiles(p):
    if len(p) == 0:
        return 0
    last = p[-1]
    return int(last['tile_begin'] + ((last['M'] + 63) // 64) * ((last['N'] + 63) // 64))


def mix_blocks(j):
    if len(j) == 0:
        return 0
    last = j[-1]
    return int(last['block_begin'] + (int(last['Dl']) * int(last['d']) * int(last['Dr'])
                                      + L.MIX_BLOCK - 1) // L.MIX_BLOCK)


def concat_gemm(parts, actives=None):
    """``parts``: list of ``(probs, segs)`` -> merged ``(probs, segs, ntiles)``.
    ``actives[k]`` (optional) overrides the ``active`` field of part k."""
    ps, ss = [], []
    tile_off = seg_off = 0
    for k, (p, s) in enumerate(parts):
        if len(p) == 0:
            continue
        p = p.copy()
        p['seg_begin'] += seg_off
        p['tile_begin'] += tile_off
        if actives is not None:
            p['active'] = actives[k]
        tile_off += gemm_tiles(parts[k][0])
        seg_off += len(s)
        ps.append(p)
        ss.append(s)
    if not ps:
        return EMPTY_PROB.copy(), EMPTY_SEG.copy(), 0
    return np.concatenate(ps), np.concatenate(ss), tile_off


def concat_mix(parts, actives=None):
    js, ts = [], []
    blk_off = term_off = 0
    for k, (j, t) in enumerate(parts):
        if len(j) == 0:
            continue
        j = j.copy()
        j['term_begin'] += term_off
        j['block_begin'] += blk_off
        if actives is not None:
            j['active'] = actives[k]
        blk_off += mix_blocks(parts[k][0])
        term_off += len(t)
        js.append(j)
        ts.append(t)
    if not js:
        return EMPTY_JOB.copy(), EMPTY_TERM.copy(), 0
    return np.concatenate(js), np.concatenate(ts), blk_off


def concat_plain(parts, dtype):
    parts = [p for p in parts if len(p)]
    if not parts:
        return np.zeros(0, dtype)
    return np.concatenate(parts)
