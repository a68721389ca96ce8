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
        self.tiles = 0   # output tiles of the problems so far (tile_begin of the next)

    def add(self, c, M, N, ldc, segs, c0=0, ldc0=0, beta=0.0, active=-1):
        """``segs``: iterable of ``(a, lda, transA, b, ldb, transB, K, alpha)``."""
        if M <= 0 or N <= 0:
            return
        segs = [s for s in segs if s[6] > 0]
        start = len(self.segs)
        for (a, lda, ta, b, ldb, tb, K, alpha) in segs:
            self.segs.append((a, b, lda, ldb, ta, tb, K, 0, alpha))
        self.probs.append((c, c0, M, N, ldc, ldc0, start, len(segs), self.tiles, active, beta))
        self.tiles += (-(-M // L.GEMM_TILE)) * (-(-N // L.GEMM_TILE))

    def arrays(self):
        p = np.array(self.probs, dtype=GEMM_PROBLEM) if self.probs else EMPTY_PROB.copy()
        s = np.array(self.segs, dtype=GEMM_SEG) if self.segs else EMPTY_SEG.copy()
        return p, s


class MixList:
    """Accumulates mixer jobs ``dst[a,s,b] (=|+=) sum coef O[s,t] src[a,t,b]``."""

    def __init__(self):
        self.jobs = []
        self.terms = []
        self.blocks = 0   # workgroups of the jobs so far (block_begin of the next)

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
                          active, self.blocks, 0))
        self.blocks += -(-(Dl * Dr) // int(L.mix_pairs(d)))

    def arrays(self):
        j = np.array(self.jobs, dtype=MIX_JOB) if self.jobs else EMPTY_JOB.copy()
        t = np.array(self.terms, dtype=MIX_TERM) if self.terms else EMPTY_TERM.copy()
        return j, t


def gemm_tiles(p):
    if len(p) == 0:
        return 0
    last = p[-1]
    return int(last['tile_begin'] + ((last['M'] + 63) // 64) * ((last['N'] + 63) // 64))


def gemm_flops(p, s):
    """Algorithmic flops of a grouped-GEMM launch: 2 M N K summed over the
    problems and their K segments."""
    if len(p) == 0:
        return 0.0
    if len(s) == 0:
        return 0.0
    # (a problem without segments may start at len(s); it is masked below)
    ksum = np.add.reduceat(s['K'].astype(np.float64), np.minimum(p['seg_begin'], len(s) - 1))
    ksum = np.where(p['seg_count'] > 0, ksum, 0.0)
    return float((2.0 * p['M'].astype(np.float64) * p['N'] * ksum).sum())


_ITEM = {0: 8, 1: 4, 2: 2}   # dtype -> bytes


def mix_bytes(j, t):
    """Algorithmic bytes of a mixer launch: every term's source and the
    destination (twice when accumulated) once per element."""
    if len(j) == 0:
        return 0.0
    el = j['Dl'].astype(np.float64) * j['d'] * j['Dr']
    isz = np.vectorize(_ITEM.get)
    dst = isz(j['dst_dtype']) * (1 + j['accumulate'])
    src = np.zeros(len(j))
    if len(t):
        ts = isz(t['src_dtype']).astype(np.float64)
        src = np.add.reduceat(ts, np.minimum(j['term_begin'], len(ts) - 1))
        src = np.where(j['term_count'] > 0, src, 0.0)
    return float((el * (dst + src)).sum())


def mix_blocks(j):
    if len(j) == 0:
        return 0
    last = j[-1]
    pairs = int(L.mix_pairs(int(last['d'])))
    return int(last['block_begin'] + (int(last['Dl']) * int(last['Dr']) + pairs - 1) // pairs)


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


_PTR_WORDS = {}


def ptr_fields(dt):
    f = _PTR_FIELDS.get(dt)
    if f is None:
        f = tuple(n for n in dt.names if dt.fields[n][0] == _U64)
        _PTR_FIELDS[dt] = f
    return f


def _ptr_block(arr):
    """The device-address fields of ``arr`` as an ``(n, k)`` uint64 view.
    Every record type of the ABI starts with its address fields, packed."""
    dt = arr.dtype
    k = _PTR_WORDS.get(dt)
    if k is None:
        offs = [dt.fields[n][1] for n in ptr_fields(dt)]
        if offs != [8 * i for i in range(len(offs))] or dt.itemsize % 8:
            raise TypeError('record type without a leading block of address fields: %r' % dt)
        k = len(offs)
        _PTR_WORDS[dt] = k
    if k == 0 or len(arr) == 0:
        return None
    return arr.view(np.uint64).reshape(len(arr), dt.itemsize // 8)[:, :k]


def _shift(arr, delta, slot_delta):
    w = _ptr_block(arr)
    if w is not None:
        np.add(w, np.uint64(delta & _MASK64), out=w, where=w != 0)
    if slot_delta and len(arr) and 'op' in arr.dtype.names:
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


def relocate_tree(obj, reloc):
    """``relocate`` applied to every record array of a nested tuple/list
    (other leaves, e.g. stage kinds, are kept)."""
    if isinstance(obj, np.ndarray) and obj.dtype.names is not None:
        return relocate(obj, reloc)
    if isinstance(obj, tuple):
        return tuple(relocate_tree(o, reloc) for o in obj)
    if isinstance(obj, list):
        return [relocate_tree(o, reloc) for o in obj]
    return obj


def _rel_one(arr, lo, hi, slot):
    w = _ptr_block(arr)
    if w is not None:
        nz = w != 0
        if nz.any():
            v = w[nz]
            if int(v.min()) < lo or int(v.max()) >= hi:
                raise ValueError('plan record of type %r points outside the solve arena'
                                 % (arr.dtype.names,))
            np.subtract(w, np.uint64(lo), out=w, where=nz)
    if len(arr) and 'coef' in arr.dtype.names:
        c = arr['coef']
        ok = np.isnan(c)
        for lit in _LITERAL_COEFS:
            ok |= c == lit
        if not ok.all():
            # a coupling value that lost its tag would be shared by solves
            # with other couplings
            raise ValueError('plan mixer term has an untagged coefficient')
    if slot and len(arr) and 'op' in arr.dtype.names:
        o = arr['op']
        o[o >= 0] -= slot


PLAN_META = ('dims_after', 'dims_after_t', 'uid', 'prec', 'mv_flops', 'svd_n', 'lm_bond',
             'lm_bonds', 'packed')


def make_relative(ph, lo, hi, slot):
    """In place: every record array of plan ``ph`` made relative to the
    arena ``[lo, hi)`` and operator slot ``slot`` (array by array)."""
    def walk(obj):
        if isinstance(obj, np.ndarray) and obj.dtype.names is not None:
            _rel_one(obj, lo, hi, slot)
        elif isinstance(obj, (tuple, list)):
            for o in obj:
                walk(o)
    for k, v in ph.items():
        if k not in PLAN_META:
            walk(v)


# ------------------------------------------------------------ packed plans
# A plan holds a few dozen small record arrays.  Packed, they sit in one byte
# buffer with the word indices of every non-null device address, every
# NaN-boxed coefficient (with its index) and every operator index >= 0, so
# that making the plan relative and moving it to a solve are a fixed number of
# array operations whatever the number of records.
class _Leaf:
    __slots__ = ('i',)

    def __init__(self, i):
        self.i = i


_FIELD_WORDS = {}


def _field_words(dt):
    """(stride in 8-byte words, address words, coef word or -1, op int32 or -1)."""
    got = _FIELD_WORDS.get(dt)
    if got is None:
        _ptr_block(np.zeros(1, dt))      # checks the leading address block
        cw = dt.fields['coef'][1] // 8 if 'coef' in dt.names else -1
        ow = dt.fields['op'][1] // 4 if 'op' in dt.names else -1
        got = (dt.itemsize // 8, _PTR_WORDS[dt], cw, ow)
        _FIELD_WORDS[dt] = got
    return got


def _flatten(obj, arrs):
    """``obj`` with each record array replaced by a leaf holding its index
    in ``arrs`` (appended).  Record arrays inside a tuple or list are taken
    in place, without a call each."""
    to = type(obj)
    if to is np.ndarray and obj.dtype.names is not None:
        arrs.append(obj)
        return _Leaf(len(arrs) - 1)
    if to is tuple or to is list:
        out = []
        for o in obj:
            if type(o) is np.ndarray and o.dtype.names is not None:
                arrs.append(o)
                out.append(_Leaf(len(arrs) - 1))
            else:
                out.append(_flatten(o, arrs))
        return tuple(out) if to is tuple else out
    return obj


def _inflate(tree, buf, specs):
    """``tree`` with each leaf replaced by its records viewed in ``buf``
    (``frombuffer``: writable, and cheaper than a structured ``view``).
    Leaves inside a tuple or list are viewed in place, without a call each."""
    tt = type(tree)
    if tt is _Leaf:
        off, dt, n = specs[tree.i]
        return np.frombuffer(buf, dt, n, off)
    if tt is tuple or tt is list:
        out = []
        for t in tree:
            if type(t) is _Leaf:
                off, dt, n = specs[t.i]
                out.append(np.frombuffer(buf, dt, n, off))
            else:
                out.append(_inflate(t, buf, specs))
        return tuple(out) if tt is tuple else out
    return tree


_LAZY = object()


class Plan(dict):
    """A step plan.  Once packed, its record phases are views of the packed
    relative buffer made on first read (the steps read phases through
    ``Solve.part``, at the solve's addresses, so most are never read here)."""
    __slots__ = ()

    def __getitem__(self, k):
        v = dict.__getitem__(self, k)
        if v is _LAZY:
            pk = dict.__getitem__(self, 'packed')
            v = pk.views(pk.buf, k)
            dict.__setitem__(self, k, v)
        return v

    def get(self, k, default=None):
        return self[k] if k in self else default


class Packed:
    """Relative plan records in one buffer.  ``ptr``: uint64 word indices of
    the non-null addresses; ``cidx``/``ck``: float64 word indices of boxed
    coefficients and their coefficient indices; ``op``: int32 indices of
    operator-table references."""
    __slots__ = ('buf', 'trees', 'specs', 'ptr', 'cidx', 'ck', 'op')

    def views(self, buf, key):
        """The records of phase ``key`` as views into ``buf``."""
        return _inflate(self.trees[key], buf, self.specs)

    def relocate(self, reloc):
        """A copy of the buffer moved to ``reloc = (base, slot, coefficient
        values)``."""
        buf = self.buf.copy()
        if len(self.ptr):
            w = buf.view(np.uint64)
            w[self.ptr] += np.uint64(reloc[0] & _MASK64)
        if len(self.op) and reloc[1]:
            buf.view(np.int32)[self.op] += np.int32(reloc[1])
        if len(self.cidx):
            buf.view(np.float64)[self.cidx] = np.asarray(reloc[2], dtype=np.float64)[self.ck]
        return buf


def _shift_tree(tree, d):
    tt = type(tree)
    if tt is _Leaf:
        return _Leaf(tree.i + d)
    if tt is tuple or tt is list:
        out = [_Leaf(t.i + d) if type(t) is _Leaf else _shift_tree(t, d) for t in tree]
        return tuple(out) if tt is tuple else out
    return tree


def merge_packed(subs, meta):
    """One plan from packed relative sub-plans with disjoint phases (e.g. the
    precision-independent gauge/environment records shared by every
    precision, and one precision's local records): buffers concatenated,
    indices and leaves shifted.  ``meta``: the plan's non-record entries."""
    pk = Packed()
    pks = [s['packed'] for s in subs]
    offs, off = [], 0
    for p in pks:
        offs.append(off)
        off += len(p.buf)
    pk.buf = np.concatenate([p.buf for p in pks]) if pks else np.zeros(0, np.uint8)
    pk.ptr = np.concatenate([p.ptr + o // 8 for p, o in zip(pks, offs)])
    pk.cidx = np.concatenate([p.cidx + o // 8 for p, o in zip(pks, offs)])
    pk.ck = np.concatenate([p.ck for p in pks])
    pk.op = np.concatenate([p.op + o // 4 for p, o in zip(pks, offs)])
    pk.specs, pk.trees = [], {}
    for p, o in zip(pks, offs):
        d = len(pk.specs)
        pk.specs += [(so + o, dt, n) for so, dt, n in p.specs]
        for k, t in p.trees.items():
            if k in pk.trees:
                raise ValueError('sub-plans share the phase %r' % k)
            pk.trees[k] = _shift_tree(t, d)
    ph = Plan(meta)
    for k in pk.trees:
        dict.__setitem__(ph, k, _LAZY)
    ph['packed'] = pk
    return ph


def pack_relative(ph, lo, hi, slot):
    """Plan ``ph`` packed and made relative to the arena ``[lo, hi)`` and
    operator slot ``slot``; its phases are replaced by views of the packed
    buffer and the packing is stored under ``ph['packed']``."""
    arrs = []
    keys = [k for k in ph if k not in PLAN_META]
    trees = {k: _flatten(ph[k], arrs) for k in keys}
    pk = Packed()
    pk.trees = trees
    if arrs:
        sizes = np.fromiter((a.nbytes for a in arrs), np.int64, len(arrs))
        offs = np.cumsum(sizes) - sizes
        buf = np.frombuffer(bytearray(b''.join([a.tobytes() for a in arrs])), np.uint8)
        fw = np.array([_field_words(a.dtype) for a in arrs], dtype=np.int64).reshape(-1, 4)
        n = np.fromiter((len(a) for a in arrs), np.int64, len(arrs))
    else:
        offs = sizes = n = np.zeros(0, np.int64)
        buf = np.zeros(0, np.uint8)
        fw = np.zeros((0, 4), np.int64)
    pk.specs = list(zip(offs.tolist(), [a.dtype for a in arrs], n.tolist()))
    # per record: array id and first word
    aid = np.repeat(np.arange(len(arrs)), n)
    first = np.cumsum(n) - n
    rec = offs[aid] // 8 + (np.arange(len(aid)) - first[aid]) * fw[aid, 0]
    kp = fw[aid, 1]
    ptr = np.repeat(rec, kp) + (np.arange(int(kp.sum())) - np.repeat(np.cumsum(kp) - kp, kp))
    w = buf.view(np.uint64)
    v = w[ptr]
    nz = v != 0
    ptr, v = ptr[nz], v[nz]
    if len(v) and (int(v.min()) < lo or int(v.max()) >= hi):
        raise ValueError('plan record points outside the solve arena')
    w[ptr] = v - np.uint64(lo)
    pk.ptr = ptr
    has_c = fw[aid, 2] >= 0
    cidx = rec[has_c] + fw[aid[has_c], 2]
    bits = w[cidx]
    boxed = ((bits & np.uint64(0xFFF8000000000000)) == np.uint64(_QNAN)) & \
        ((bits & np.uint64(_PAYLOAD)) != 0)
    c = buf.view(np.float64)[cidx[~boxed]]
    ok = np.zeros(len(c), bool)
    for lit in _LITERAL_COEFS:
        ok |= c == lit
    if not ok.all():
        # a coupling value that lost its tag would be shared by solves with
        # other couplings
        raise ValueError('plan mixer term has an untagged coefficient')
    pk.cidx = cidx[boxed]
    pk.ck = (bits[boxed] & np.uint64(_PAYLOAD)).astype(np.int64) - 1
    has_o = fw[aid, 3] >= 0
    oidx = 2 * rec[has_o] + fw[aid[has_o], 3]
    iv = buf.view(np.int32)
    o = iv[oidx]
    pk.op = oidx[o >= 0]
    iv[pk.op] -= np.int32(slot)
    pk.buf = buf
    ph['packed'] = pk
    for k in keys:
        if type(ph) is Plan:
            dict.__setitem__(ph, k, _LAZY)
        else:
            ph[k] = pk.views(buf, k)
    return pk


def _cat(arrs, dtype):
    """Concatenation of same-dtype record arrays as raw bytes, into a new
    writable array (np.concatenate on structured dtypes re-derives the field
    promotion per call, and a byte ``view`` of a structured array runs a
    Python-level safety check per part; joining ``tobytes`` is several times
    faster for the few-record parts of a step)."""
    if not arrs:
        return np.zeros(0, dtype)
    if len(arrs) == 1:
        return arrs[0]
    return np.frombuffer(bytearray(b''.join([a.tobytes() for a in arrs])), dtype)


def _cat_copy(arrs, dtype):
    """Concatenation that is always a new array (the caller edits it)."""
    if not arrs:
        return np.zeros(0, dtype)
    if len(arrs) == 1:
        return arrs[0].copy()
    return _cat(arrs, dtype)


def _rebase(arr, lens, counts, field):
    """``arr[field] +=`` the exclusive prefix sum of ``counts`` (one entry per
    part), each entry repeated over that part's ``lens`` records."""
    counts = np.asarray(counts, dtype=np.int64)
    arr[field] += np.repeat(np.cumsum(counts) - counts, lens).astype(arr.dtype[field])


def _merge(parts, actives, relocs, rec_dt, sub_dt, sub_field, unit_field, units_of_last,
           empty):
    """Shared body of ``concat_gemm`` / ``concat_mix``: one byte-level
    concatenation per record kind, then vectorised fix-ups of the prefix
    fields (sub-record index ``sub_field`` and work-unit index ``unit_field``)
    and of ``active``."""
    if relocs is not None:
        parts = [(relocate(a, r), relocate(b, r)) for (a, b), r in zip(parts, relocs)]
    keep = [k for k, (a, _) in enumerate(parts) if len(a)]
    if not keep:
        return empty[0].copy(), empty[1].copy(), 0
    rs = [parts[k][0] for k in keep]
    ss = [parts[k][1] for k in keep]
    lens = np.fromiter((len(a) for a in rs), np.int64, len(rs))
    out = _cat_copy(rs, rec_dt)
    last = np.cumsum(lens) - 1
    units = units_of_last(out[last])     # work units of each part: its last record's end
    _rebase(out, lens, [len(b) for b in ss], sub_field)
    _rebase(out, lens, units, unit_field)
    if actives is not None:
        out['active'] = np.repeat(np.asarray([actives[k] for k in keep], dtype=np.int32), lens)
    return out, _cat_copy(ss, sub_dt), int(units.sum())


def _gemm_end(last):
    return last['tile_begin'].astype(np.int64) + \
        ((last['M'].astype(np.int64) + 63) // 64) * ((last['N'].astype(np.int64) + 63) // 64)


def _mix_end(last):
    pairs = L.mix_pairs(last['d']).astype(np.int64)
    return last['block_begin'].astype(np.int64) + \
        (last['Dl'].astype(np.int64) * last['Dr'] + pairs - 1) // pairs


def concat_gemm(parts, actives=None, relocs=None):
    """``parts``: list of ``(probs, segs)`` -> merged ``(probs, segs, ntiles)``.
    ``actives[k]`` (optional) overrides the ``active`` field of part k;
    ``relocs[k]`` (optional) moves part k from relative to device addresses
    (parts given already relocated skip that)."""
    return _merge(parts, actives, relocs, GEMM_PROBLEM, GEMM_SEG, 'seg_begin', 'tile_begin',
                  _gemm_end, (EMPTY_PROB, EMPTY_SEG))


def concat_mix(parts, actives=None, relocs=None):
    """``parts``: list of ``(jobs, terms)`` -> merged ``(jobs, terms, nblocks)``
    (as ``concat_gemm``)."""
    return _merge(parts, actives, relocs, MIX_JOB, MIX_TERM, 'term_begin', 'block_begin',
                  _mix_end, (EMPTY_JOB, EMPTY_TERM))


def concat_plain(parts, dtype, relocs=None):
    if relocs is not None:
        parts = [relocate(p, rl) for p, rl in zip(parts, relocs)]
    return _cat_copy([p for p in parts if len(p)], dtype)
