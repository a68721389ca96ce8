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
        if d > L.MIX_DMAX:
            raise ValueError('physical dimension %d exceeds the mixer limit %d' % (d, L.MIX_DMAX))
        start = len(self.terms)
        for (src, sdt, coef, op) in terms:
            self.terms.append((src, coef_field(coef), op, sdt))
        self.jobs.append((dst, Dl, d, Dr, dst_dtype, start, len(self.terms) - start, accumulate,
                          active, 0, 0))

    def arrays(self):
        j = np.array(self.jobs, dtype=MIX_JOB) if self.jobs else EMPTY_JOB.copy()
        t = np.array(self.terms, dtype=MIX_TERM) if self.terms else EMPTY_TERM.copy()
        if len(j):
            blocks = (j['Dl'].astype(np.int64) * j['Dr'] + L.MIX_PAIRS - 1) // L.MIX_PAIRS
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
    return int(last['block_begin'] + (int(last['Dl']) * int(last['Dr'])
                                      + L.MIX_PAIRS - 1) // L.MIX_PAIRS)


# ------------------------------------------------------------ relocation
# A plan is built once per solve signature with the addresses of one solve;
# ``make_relative`` rewrites it against that solve's arena and operator slot,
# ``relocate`` moves a copy to any other solve.  Every 64-bit unsigned field
# of a record is a device address (0 = null, left alone); the ``op`` field of
# a mixer term is an operator-table index (-1 = identity, left alone); the
# ``coef`` of a mixer term that comes from the coupling tables is written
# as a quiet NaN whose payload is the coefficient's index + 1, and takes the
# solve's value on relocation.
_PTR_FIELDS = {}
_U64 = np.dtype('<u8')
_MASK64 = (1 << 64) - 1
_QNAN = 0x7FF8000000000000
_PAYLOAD = (1 << 51) - 1
_LITERAL_COEFS = (1.0, -1.0, 0.0)


def coef_field(c):
    """The ``coef`` field of a mixer term: the index of a tagged coupling
    coefficient (``terms.Coef``), NaN-boxed, or the literal value."""
    k = getattr(c, 'k', None)
    if k is None:
        return float(c)
    return float(np.array(_QNAN | (k + 1), dtype=np.uint64).view(np.float64))


def _unbox(coefs, values):
    """In place: NaN-boxed coefficient indices -> ``values[k]``."""
    bits = coefs.view(np.uint64)
    boxed = (bits & np.uint64(0xFFF8000000000000)) == np.uint64(_QNAN)
    boxed &= (bits & np.uint64(_PAYLOAD)) != 0
    if boxed.any():
        coefs[boxed] = values[(bits[boxed] & np.uint64(_PAYLOAD)).astype(np.int64) - 1]


def ptr_fields(dt):
    f = _PTR_FIELDS.get(dt)
    if f is None:
        f = tuple(n for n in dt.names if dt.fields[n][0] == _U64)
        _PTR_FIELDS[dt] = f
    return f


def _shift(arr, delta, slot_delta):
    if len(arr) == 0:
        return
    d = np.uint64(delta & _MASK64)
    for f in ptr_fields(arr.dtype):
        v = arr[f]
        nz = v != 0
        v[nz] += d
    if slot_delta and 'op' in arr.dtype.names:
        o = arr['op']
        o[o >= 0] += slot_delta


def relocate(arr, reloc):
    """Copy of a relative record array moved to ``reloc = (base, slot[,
    coefficient values])``."""
    out = arr.copy()
    if reloc is not None:
        _shift(out, reloc[0], reloc[1])
        if len(reloc) > 2 and len(out) and 'coef' in out.dtype.names:
            _unbox(out['coef'], reloc[2])
    return out


def _rel_one(arr, lo, hi, slot):
    for f in ptr_fields(arr.dtype):
        v = arr[f]
        nz = v[v != 0]
        if len(nz) and (int(nz.min()) < lo or int(nz.max()) >= hi):
            raise ValueError('plan record field %r points outside the solve arena' % f)
    if len(arr) and 'coef' in arr.dtype.names:
        c = arr['coef']
        lit = ~np.isnan(c)
        if not np.isin(c[lit], _LITERAL_COEFS).all():
            # a coupling value that lost its tag would be shared by solves
            # with other couplings
            raise ValueError('plan mixer term has an untagged coefficient')
    _shift(arr, -lo, -slot)


def make_relative(ph, lo, hi, slot):
    """In place: every record array of plan ``ph`` made relative to the
    arena ``[lo, hi)`` and operator slot ``slot``."""
    def walk(obj):
        if isinstance(obj, np.ndarray) and obj.dtype.names is not None:
            _rel_one(obj, lo, hi, slot)
        elif isinstance(obj, (tuple, list)):
            for o in obj:
                walk(o)
    for k, v in ph.items():
        if k not in ('dims_after', 'uid', 'prec', 'mv_flops', 'svd_n', 'lm_bond'):
            walk(v)


def _cat(arrs, dtype):
    """Concatenation of same-dtype record arrays as raw bytes (np.concatenate
    on structured dtypes re-derives the field promotion per call)."""
    if not arrs:
        return np.zeros(0, dtype)
    if len(arrs) == 1:
        return arrs[0]
    return np.concatenate([a.view(np.uint8) for a in arrs]).view(dtype)


def concat_gemm(parts, actives=None, relocs=None):
    """``parts``: list of ``(probs, segs)`` -> merged ``(probs, segs, ntiles)``.
    ``actives[k]`` (optional) overrides the ``active`` field of part k;
    ``relocs[k]`` (optional) moves part k from relative to device addresses."""
    ps, ss = [], []
    tile_off = seg_off = 0
    for k, (p, s) in enumerate(parts):
        if len(p) == 0:
            continue
        rl = relocs[k] if relocs is not None else None
        p = relocate(p, rl)
        p['seg_begin'] += seg_off
        p['tile_begin'] += tile_off
        if actives is not None:
            p['active'] = actives[k]
        tile_off += gemm_tiles(parts[k][0])
        seg_off += len(s)
        ps.append(p)
        ss.append(relocate(s, rl) if rl is not None else s)
    if not ps:
        return EMPTY_PROB.copy(), EMPTY_SEG.copy(), 0
    segs = _cat(ss, GEMM_SEG)
    return _cat(ps, GEMM_PROBLEM), (segs.copy() if len(ss) == 1 and relocs is None else segs), \
        tile_off


def concat_mix(parts, actives=None, relocs=None):
    js, ts = [], []
    blk_off = term_off = 0
    for k, (j, t) in enumerate(parts):
        if len(j) == 0:
            continue
        rl = relocs[k] if relocs is not None else None
        j = relocate(j, rl)
        j['term_begin'] += term_off
        j['block_begin'] += blk_off
        if actives is not None:
            j['active'] = actives[k]
        blk_off += mix_blocks(parts[k][0])
        term_off += len(t)
        js.append(j)
        ts.append(relocate(t, rl) if rl is not None else t)
    if not js:
        return EMPTY_JOB.copy(), EMPTY_TERM.copy(), 0
    terms = _cat(ts, MIX_TERM)
    return _cat(js, MIX_JOB), (terms.copy() if len(ts) == 1 and relocs is None else terms), \
        blk_off


def concat_plain(parts, dtype, relocs=None):
    if relocs is not None:
        parts = [relocate(p, rl) for p, rl in zip(parts, relocs) if len(p)]
        return _cat(parts, dtype)
    parts = [p for p in parts if len(p)]
    if not parts:
        return np.zeros(0, dtype)
    return _cat(parts, dtype).copy() if len(parts) == 1 else _cat(parts, dtype)
