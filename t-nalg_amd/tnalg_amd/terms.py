"""Compile ``para`` coupling lists into block-environment term tables.

The reference keeps one effective operator per coupling row and bond
(``library/MPSClass.py:284-611``) and sums them into ``opt_env`` classes at
every step (``MPSClass.py:715-815``).  The engine instead carries, at every
bond ``b``, a left block (sites ``< b``) made of ``H_L`` plus one matrix per
*open* operator ``(i, a)`` (site ``i < b`` whose operator ``a`` still has a
partner ``j >= b``), and the mirror-image right block.  The effective
Hamiltonian at site ``p`` is then (``x`` = the ``(Dl, d, Dr)`` site tensor)

    Heff x = H_L x + x H_R^T + h_p x
             + sum_b O_b (Lt_b x)              Lt_b = sum_{rows (i,p,a,b)} c L[(i,a)]
             + sum_a O_a (x Rt_a^T)            Rt_a = sum_{rows (p,j,a,b)} c R[(j,b)]
             + sum_{rows i<p<j} c L[(i,a)] x R[(j,b)]^T

which is the reference's ``Heff`` term for term (same rows, same ``|c| > tol``
cut of ``MPSClass.py:725-734``).  The straddling double sum is regrouped
over the smaller side: ``sum_l L_l x (sum_r C[l,r] R_r)^T`` (or the mirror),
i.e. an exact refactorisation with ``min(#left, #right)`` GEMM pairs.
"""
import collections
import hashlib
import threading
from dataclasses import dataclass, field

import numpy as np


class Coef(float):
    """A coefficient of the term tables: its value, and its index ``k`` in
    ``Terms.coef_values``.  Step plans refer to coefficients by index, so
    solves whose couplings differ only in value share one plan."""

    def __new__(cls, value, k):
        obj = float.__new__(cls, value)
        obj.k = k
        return obj


@dataclass
class SiteTerms:
    field: list            # [(op, coef)] on-site operators at p
    lm: list               # [(op_b, [(left_open_idx, coef)])]  groups paired with O_b at p
    mr: list               # [(op_a, [(right_open_idx, coef)])] groups paired with O_a at p
    strad_left: bool       # True: factors are raw left opens, right side combined
    strad: list            # [(raw_idx, [(other_idx, coef)])]


@dataclass
class Terms:
    n: int
    d: int
    ops: np.ndarray                       # (n_ops, d, d) float64
    left_opens: list = field(default_factory=list)    # per bond b: [(i, a)]
    right_opens: list = field(default_factory=list)   # per bond b: [(j, b)]
    sites: list = field(default_factory=list)         # per site: SiteTerms
    n_two: int = 0
    n_one: int = 0
    layouts: dict = field(default_factory=dict, repr=False)  # device layouts built from these tables
    coef_values: np.ndarray = field(default=None, repr=False)  # value of coefficient k
    coef_digest: str = field(default=None, repr=False)         # hash of coef_values
    # two-site tables (compiled on request): per pair (p, p+1) the terms of
    # the merged site of physical dimension d^2 between bonds p and p+2, over
    # the merged operators ops2 (kron products, keys ops2_keys)
    pairs: list = field(default_factory=list)
    ops2: np.ndarray = field(default=None, repr=False)
    ops2_keys: list = field(default_factory=list)

    def structure(self):
        """Everything about the tables except the coefficient values (each
        coefficient appears as its index)."""
        def sub(v):
            return ('c', v.k) if isinstance(v, Coef) else v

        def site_key(st):
            return (tuple((op, sub(c)) for op, c in st.field),
                    tuple((b, tuple((i, sub(c)) for i, c in g)) for b, g in st.lm),
                    tuple((a, tuple((j, sub(c)) for j, c in g)) for a, g in st.mr),
                    st.strad_left,
                    tuple((r, tuple((o, sub(c)) for o, c in g)) for r, g in st.strad))
        return (self.n, self.d, self.ops.shape[0], tuple(map(tuple, self.left_opens)),
                tuple(map(tuple, self.right_opens)), tuple(site_key(s) for s in self.sites),
                tuple(site_key(s) for s in self.pairs), tuple(self.ops2_keys))

    def left_index(self, b):
        return {k: t for t, k in enumerate(self.left_opens[b])}

    def right_index(self, b):
        return {k: t for t, k in enumerate(self.right_opens[b])}


def _real_ops(ops, used):
    arr = []
    for k, o in enumerate(ops):
        o = np.asarray(o)
        if k in used and np.iscomplexobj(o) and np.abs(o.imag).max() > 0:
            raise ValueError('operator %d has an imaginary part; the real-arithmetic engine '
                             'supports real Hamiltonians (is_real=True)' % k)
        arr.append(np.real(o).astype(np.float64))
    return np.stack(arr)


def _regroup(pairs):
    """Straddling terms ``{(l, r): c}`` grouped over the smaller side:
    ``(strad_left, [(raw, [(other, c)])])``."""
    lset = sorted({l for (l, _) in pairs})
    rset = sorted({r for (_, r) in pairs})
    strad_left = len(lset) <= len(rset)
    strad = []
    if strad_left:
        for l in lset:
            strad.append((l, [(r, c) for (ll, r), c in sorted(pairs.items()) if ll == l]))
    else:
        for r in rset:
            strad.append((r, [(l, c) for (l, rr), c in sorted(pairs.items()) if rr == r]))
    return strad_left, strad


def _pair_tables(T, two, one, tag):
    """Two-site tables: for each pair (p, p+1) the sites merge into one of
    physical index s = s_p * d + s_{p+1}; its left block is bond p and its
    right block bond p+2.  Rows are sorted exactly as for one site:
    - a row inside the pair (i = p, j = p+1) is an on-site term O_a x O_b;
    - a one-site row at p or p+1 is O_k x 1 or 1 x O_k;
    - a row from a left open operator to p or p+1 pairs it with O_b x 1 or 1 x O_b;
    - a row from p or p+1 to a right open operator likewise on the right;
    - a row from i < p to j > p+1 straddles the merged site.
    The merged operators are numbered in the order the tables first use them."""
    n, d = T.n, T.d
    eye = np.eye(d)
    keys = {}

    def mop(key):
        if key not in keys:
            keys[key] = len(keys)
        return keys[key]

    def kron_of(key):
        if key[0] == 'L':
            return np.kron(T.ops[key[1]], eye)
        if key[0] == 'R':
            return np.kron(eye, T.ops[key[1]])
        return np.kron(T.ops[key[1]], T.ops[key[2]])

    for p in range(n - 1):
        lidx = T.left_index(p)
        ridx = T.right_index(p + 2)
        fld, lm, mr, pairs = {}, {}, {}, {}
        for (q, k, c) in one:
            if q == p or q == p + 1:
                key = ('L', k) if q == p else ('R', k)
                fld[key] = fld.get(key, 0.0) + c
        for (i, j, a, b, c) in two:
            if i == p and j == p + 1:
                key = ('LR', a, b)
                fld[key] = fld.get(key, 0.0) + c
            elif i < p and j in (p, p + 1):
                lm.setdefault(('L', b) if j == p else ('R', b), []).append((lidx[(i, a)], c))
            elif i in (p, p + 1) and j > p + 1:
                mr.setdefault(('L', a) if i == p else ('R', a), []).append((ridx[(j, b)], c))
            elif i < p and j > p + 1:
                kk = (lidx[(i, a)], ridx[(j, b)])
                pairs[kk] = pairs.get(kk, 0.0) + c
        strad_left, strad = _regroup(pairs)
        field_t = [(mop(k), tag(c)) for k, c in sorted(fld.items())]
        lm_t = [(mop(k), [(i, tag(c)) for i, c in v]) for k, v in sorted(lm.items())]
        mr_t = [(mop(k), [(j, tag(c)) for j, c in v]) for k, v in sorted(mr.items())]
        strad_t = [(r, [(o, tag(c)) for o, c in g]) for r, g in strad]
        T.pairs.append(SiteTerms(field=field_t, lm=lm_t, mr=mr_t, strad_left=strad_left,
                                 strad=strad_t))
    T.ops2_keys = sorted(keys, key=keys.get)
    T.ops2 = (np.stack([kron_of(k) for k in T.ops2_keys]) if keys
              else np.zeros((1, d * d, d * d)))


def compile_terms(para, pairs=False):
    """Build the :class:`Terms` of a ``para`` dict (keys ``l, op, index1, coeff1,
    index2, coeff2, eigs_tol``); ``pairs``: also the two-site tables."""
    n = int(para['l'])
    tol = float(para.get('eigs_tol', 0.0))
    index1 = np.asarray(para['index1'], dtype=int).reshape(-1, 2)
    index2 = np.asarray(para['index2'], dtype=int).reshape(-1, 4)
    coeff1 = np.asarray(para['coeff1'], dtype=float).reshape(-1)
    coeff2 = np.asarray(para['coeff2'], dtype=float).reshape(-1)
    ops_in = para['op']
    two, one = [], []
    used = set()
    for r in range(index2.shape[0]):
        c = coeff2[r]
        if not abs(c) > tol:
            continue
        i, j, a, b = (int(x) for x in index2[r])
        if i == j:
            raise ValueError('index2 row %d couples site %d to itself' % (r, i))
        if i > j:
            i, j, a, b = j, i, b, a
        if not (0 <= i and j < n):
            raise ValueError('index2 row %d has a site outside 0..%d' % (r, n - 1))
        two.append((i, j, a, b, float(c)))
        used.update((a, b))
    for r in range(index1.shape[0]):
        c = coeff1[r]
        q, k = (int(x) for x in index1[r])
        if abs(c) > tol and np.linalg.norm(np.asarray(ops_in[k]).reshape(-1)) > tol:
            if not 0 <= q < n:
                raise ValueError('index1 row %d has a site outside 0..%d' % (r, n - 1))
            one.append((q, k, float(c)))
            used.add(k)
    ops = _real_ops(ops_in, used)
    d = ops.shape[1]
    if 'd' in para and int(para['d']) != d:
        raise ValueError('para["d"]=%s but the operators are %dx%d' % (para['d'], d, d))
    T = Terms(n=n, d=d, ops=ops, n_two=len(two), n_one=len(one))
    coef_values = []

    def tag(c):
        coef_values.append(float(c))
        return Coef(c, len(coef_values) - 1)
    for b in range(n + 1):
        T.left_opens.append(sorted({(i, a) for (i, j, a, _, _) in two if i < b <= j}))
        T.right_opens.append(sorted({(j, bb) for (i, j, _, bb, _) in two if i < b <= j}))
    for p in range(n):
        lidx = T.left_index(p)
        ridx = T.right_index(p + 1)
        fld = {}
        for (q, k, c) in one:
            if q == p:
                fld[k] = fld.get(k, 0.0) + c
        lm, mr = {}, {}
        strad_c = {}
        for (i, j, a, b, c) in two:
            if j == p:
                lm.setdefault(b, []).append((lidx[(i, a)], c))
            elif i == p:
                mr.setdefault(a, []).append((ridx[(j, b)], c))
            elif i < p < j:
                key = (lidx[(i, a)], ridx[(j, b)])
                strad_c[key] = strad_c.get(key, 0.0) + c
        strad_left, strad = _regroup(strad_c)
        # tag every coefficient with its index, in an order fixed by the
        # structure alone (sorted keys, then list order)
        field_t = [(k, tag(c)) for k, c in sorted(fld.items(), key=lambda kv: kv[0])]
        lm_t = [(b, [(i, tag(c)) for i, c in v]) for b, v in sorted(lm.items(), key=lambda kv: kv[0])]
        mr_t = [(a, [(j, tag(c)) for j, c in v]) for a, v in sorted(mr.items(), key=lambda kv: kv[0])]
        strad_t = [(r, [(o, tag(c)) for o, c in g]) for r, g in strad]
        T.sites.append(SiteTerms(field=field_t, lm=lm_t, mr=mr_t, strad_left=strad_left,
                                 strad=strad_t))
    if pairs:
        _pair_tables(T, two, one, tag)
    T.coef_values = np.array(coef_values, dtype=np.float64)
    return T


_COMPILED = collections.OrderedDict()
_COMPILED_LOCK = threading.Lock()   # callers include request-preparation threads


def compile_terms_cached(para, maxsize=128, pairs=False):
    """``compile_terms`` memoised on the content it reads (``l``, ``eigs_tol``,
    ``index1/2``, ``coeff1/2``, ``op``, ``d``): solves of the same
    Hamiltonian share one read-only ``Terms``."""
    h = hashlib.sha1()
    h.update(repr((int(para['l']), float(para.get('eigs_tol', 0.0)), para.get('d'),
                   bool(pairs))).encode())
    for k, dt in (('index1', np.int64), ('index2', np.int64), ('coeff1', np.float64),
                  ('coeff2', np.float64)):
        a = np.ascontiguousarray(np.asarray(para[k], dtype=dt))
        h.update(repr(a.shape).encode())
        h.update(a.tobytes())
    for o in para['op']:
        a = np.ascontiguousarray(np.asarray(o))
        h.update(repr((a.dtype.str, a.shape)).encode())
        h.update(a.tobytes())
    key = h.hexdigest()
    with _COMPILED_LOCK:
        T = _COMPILED.get(key)
        if T is not None:
            _COMPILED.move_to_end(key)
            return T
    T = compile_terms(para, pairs=pairs)   # outside the lock; a race compiles twice
    with _COMPILED_LOCK:
        T = _COMPILED.setdefault(key, T)
        _COMPILED.move_to_end(key)
        while len(_COMPILED) > maxsize:
            _COMPILED.popitem(last=False)
    return T


def dense_heff(T, p, left_env, right_env, dl, dr, pair=False):
    """Dense ``Heff`` at site ``p`` (``pair``: at the merged sites p, p+1)
    from explicit block environments (host numpy; used by the tests to check
    the term tables against the oracle).  ``left_env = (HL, [L_open...])``
    at bond p, ``right_env = (HR, [R_open...])`` at bond p+1 (p+2)."""
    d, ops, st = (T.d * T.d, T.ops2, T.pairs[p]) if pair else (T.d, T.ops, T.sites[p])
    last = p + 1 if pair else p
    dim = dl * d * dr
    H = np.zeros((dim, dim))
    el, ed, er = np.eye(dl), np.eye(d), np.eye(dr)

    def k3(a, b, c):
        return np.kron(np.kron(a, b), c)

    HL, Lo = left_env
    HR, Ro = right_env
    if p > 0:
        H += k3(HL, ed, er)
    if last < T.n - 1:
        H += k3(el, ed, HR)
    for k, c in st.field:
        H += c * k3(el, ops[k], er)
    for b, grp in st.lm:
        Lt = sum(c * Lo[l] for l, c in grp)
        H += k3(Lt, ops[b], er)
    for a, grp in st.mr:
        Rt = sum(c * Ro[r] for r, c in grp)
        H += k3(el, ops[a], Rt)
    for raw, grp in st.strad:
        if st.strad_left:
            H += k3(Lo[raw], ed, sum(c * Ro[r] for r, c in grp))
        else:
            H += k3(sum(c * Lo[l] for l, c in grp), ed, Ro[raw])
    return H
