"""Device expectation values of MPS states (one- and two-point functions).

Restates the observation routines of ``library/MPSClass.py:1105-1212``
(``observation_s1``, ``observation_s1_s2``, ``observe_magnetization``,
``observe_bond_energy``, ``observe_bond_energy_from_jxyz``,
``observe_correlators_from_middle``) as one gauge-independent pass:

* a right-to-left pass builds the right identity environments ``F_b``
  (``F_N = 1``, ``F_b = sum_s A_s F_{b+1} A_s^T``);
* a left-to-right pass carries the left identity environment ``E_b`` and one
  "string" matrix per open left operator ``(i, a)`` needed by the requested
  pairs; at site ``s`` every carried matrix ``X`` is transferred through
  ``A_s`` once (``Z = X A_s``), mixed with the operators that close there,
  contracted back (``A_s^T (O Z)``) and closed against ``F_{s+1}`` with a
  Frobenius product:  ``<O_a(i) O_b(j)> = < T(X_(i,a), A_j, O_b), F_{j+1} >``.

All sites' transfers of all solves are batched into one grouped-GEMM /
mixer / GEMM / dot launch per site.  Values are divided by the norm
``<psi|psi> = Tr E_N``.
"""
import collections
import ctypes

import numpy as np
import torch

from . import _lib as L
from . import records as R
from .device import ctypes_void


def _p(t, off=0):
    return t.data_ptr() + 8 * off


class _MeasurePlan:
    """Buffers and bookkeeping of one solve's measurement.

    A string ``(i, a)`` lives from site ``i`` (where it starts) to the last
    site that closes it; strings whose lives do not overlap share a slot of
    the string buffers, so the buffers hold the strings alive at one site
    (three for the bond rows of a chain) rather than all of them.  The left
    identity environment needs the current and the next bond only."""

    def __init__(self, solve, one, two):
        self.s = solve
        n = solve.n
        self.one = [(int(i), int(k)) for i, k in one]
        rows = []
        for (i, j, a, b) in two:
            i, j, a, b = int(i), int(j), int(a), int(b)
            if i == j:
                raise ValueError('two-point rows need two different sites (got %d, %d)' % (i, j))
            if i > j:
                i, j, a, b = j, i, b, a
            rows.append((i, j, a, b))
        self.two = rows
        # chains: (i, a) -> last site needed
        self.chain_end = {}
        for (i, j, a, b) in rows:
            self.chain_end[(i, a)] = max(self.chain_end.get((i, a), -1), j)
        self.chains = sorted(self.chain_end)
        # slot of each string: a slot is free again for strings that start at
        # or after the site that last reads it
        self.cidx, busy, nslot = {}, [], 0
        for c in self.chains:
            busy.sort()
            if busy and busy[0][0] <= c[0]:
                _, slot = busy.pop(0)
            else:
                slot, nslot = nslot, nslot + 1
            self.cidx[c] = slot
            busy.append((self.chain_end[c], slot))
        dm = solve.dmax
        self.dsq = max(x * x for x in dm)
        self.psi = solve.psi_max
        nc = max(1, nslot)
        dev = solve.dev
        f64 = torch.float64
        # close counts per site
        self.close_at = [[] for _ in range(n)]   # (kind, ref, op, out_index)
        self.values = np.zeros(len(self.one) + len(rows))
        for t, (i, k) in enumerate(self.one):
            self.close_at[i].append(('E', None, k, t))
        for t, (i, j, a, b) in enumerate(rows):
            self.close_at[j].append(('X', self.cidx[(i, a)], b, len(self.one) + t))
        starts = [0] * n
        for (i, a) in self.chains:
            starts[i] += 1
        nm = max(1, max(len(c) + st for c, st in zip(self.close_at, starts)))
        self.E = torch.zeros(2 * self.dsq, dtype=f64, device=dev)
        self.F = torch.zeros((n + 1) * self.dsq, dtype=f64, device=dev)
        self.X = [torch.zeros(nc * self.dsq, dtype=f64, device=dev) for _ in range(2)]
        self.Z = torch.zeros((nc + 1) * self.psi, dtype=f64, device=dev)
        self.M = torch.zeros(nm * self.psi, dtype=f64, device=dev)
        self.Y = torch.zeros(nm * self.dsq, dtype=f64, device=dev)
        self.out = torch.zeros(1, dtype=f64, device=dev)
        self.E[0] = 1.0
        self.F[n * self.dsq] = 1.0

    def F_at(self, b):
        return _p(self.F, b * self.dsq)

    def E_at(self, b):
        return _p(self.E, (b % 2) * self.dsq)


def real_rows(ops, a, b, tol=1e-12):
    """Rows ``[(a', b', c)]`` of real operators whose sum ``sum c <a'(i) b'(j)>``
    is ``Re <op_a(i) op_b(j)>`` for a real state.

    The device table holds the real part ``R`` of every operator.  With
    ``op = R + iI`` and a real state, ``<I_i R_j>`` is real, so only the
    product of two imaginary parts reaches the real part:
    ``Re <op_a op_b> = <R_a R_b> - <I_a I_b>``.  ``I_a`` and ``I_b`` are
    written as real combinations of the operator list (``Sy = i (S+ - S-)/2``
    is ``(op4 - op5) i/2``); an imaginary part outside that span raises.
    The reference takes the same real part (``observe_bond_energy_from_jxyz``,
    ``MPSClass.py:1184-1194``)."""
    oa, ob = np.asarray(ops[a]), np.asarray(ops[b])
    ia, ib = np.imag(oa), np.imag(ob)
    if not (np.abs(ia).max() > 0 and np.abs(ib).max() > 0):
        return [(a, b, 1.0)]
    basis = [k for k, o in enumerate(ops) if np.abs(np.real(np.asarray(o))).max() > 0]
    M = np.stack([np.real(np.asarray(ops[k])).reshape(-1) for k in basis], axis=1)

    def decompose(im, k):
        c, *_ = np.linalg.lstsq(M, im.reshape(-1), rcond=None)
        if np.abs(M @ c - im.reshape(-1)).max() > tol * max(1.0, np.abs(im).max()):
            raise ValueError('operator %d: its imaginary part is not a real combination of the '
                             'operator list, so <op op> cannot be measured in real arithmetic'
                             % k)
        return [(basis[t], float(x)) for t, x in enumerate(c) if abs(x) > tol]
    rows = [] if not (np.abs(np.real(oa)).max() > 0 and np.abs(np.real(ob)).max() > 0) \
        else [(a, b, 1.0)]
    for ka, ca in decompose(ia, a):
        for kb, cb in decompose(ib, b):
            rows.append((ka, kb, -ca * cb))
    return rows


def measure(engine, solves, requests, tags=None):
    """``requests[k] = (one, two)`` for ``solves[k]``: ``one`` = [(site, op)],
    ``two`` = [(i, j, op_i, op_j)].  Returns, per solve, ``(one_vals, two_vals,
    norm)`` as numpy arrays: the real parts of the expectation values
    (``one``: ``Re <op>``, which for a real state is that of the operator's
    real part; ``two``: see :func:`real_rows`)."""
    return measure_end(measure_begin(engine, solves, requests, tags))


def measure_begin(engine, solves, requests, tags=None):
    """The launches of :func:`measure` (on the engine's stream, reading the
    solves' current tensors); :func:`measure_end` waits for and returns the
    values.  Later work on the stream may change the tensors meanwhile.
    ``tags[k]`` (optional, hashable) names request k: equal tags promise
    equal requests on solves with equal operator lists, so their expansion
    and program lookup are keyed by the tag instead of the rows."""
    expanded, maps, keys = [], [], []
    memo = engine.__dict__.setdefault('_measure_req_memo', {})
    for k, (s, (one, two)) in enumerate(zip(solves, requests)):
        tag = None if tags is None else tags[k]
        got = memo.get(tag) if tag is not None else None
        if got is None:
            ops = s.para['op']
            rows, where = [], []
            cache = {}
            for (i, j, a, b) in two:
                key = (int(a), int(b))
                if key not in cache:
                    cache[key] = real_rows(ops, *key)
                lst = []
                for (ka, kb, c) in cache[key]:
                    lst.append((len(rows), c))
                    rows.append((i, j, ka, kb))
                where.append(lst)
            simple = all(len(lst) == 1 and lst[0][1] == 1.0 for lst in where)
            got = (one, rows, where, simple)
            if tag is not None:
                if len(memo) > 4096:
                    memo.clear()
                memo[tag] = got
        one_, rows, where, simple = got
        expanded.append((one_, rows))
        maps.append((where, simple))
        keys.append(None if tag is None else (_solve_key(s), tag))
    return _measure_real_begin(engine, solves, expanded, keys), maps


def measure_end(handle):
    real, maps = handle
    out = []
    for (vone, vtwo, norm), (where, simple) in zip(_measure_real_end(real), maps):
        if simple:   # one row per request row, coefficient 1: the values themselves
            two = np.asarray(vtwo, dtype=np.float64).copy()
        else:
            two = np.array([sum(c * vtwo[t] for t, c in lst) for lst in where],
                           dtype=np.float64)
        out.append((vone, two, norm))
    return out


def _solve_key(s):
    """Everything the records of a measurement take from a solve: the device
    addresses (arena base and layout) and the operator slot."""
    return (s.base, s.n, s.d, tuple(s.dims), tuple(s.dmax), s.psi_max, s.ops_slot)


class _MeasureProgram:
    """Every launch of one measurement of a batch of solves, built once: the
    plans' device buffers, all descriptor records in one device buffer and
    the launch list.  The records depend only on the solves' addresses, bond
    dimensions and operator slots (``_solve_key``), so a batch of the same
    shape on the same arenas (the next batch of a solver service, or the next
    observation of the same solve) replays it with no host work beyond the
    launches."""

    def __init__(self, engine, solves, requests):
        self.plans = plans = [_MeasurePlan(s, one, two) for s, (one, two) in zip(solves, requests)]
        self.nbytes = sum(t.numel() * 8 for P in plans
                          for t in [P.E, P.F, P.Z, P.M, P.Y, P.out] + P.X)
        arrays, launches = [], []
        nmax = max(s.n for s in solves)

        def add_gemm(g):
            pr, sg = g.arrays()
            nt = R.gemm_tiles(pr)
            if nt:
                launches.append(('g', len(arrays), len(pr), nt))
                arrays.extend([pr, sg])

        def add_mix(mx):
            j, t = mx.arrays()
            nb = R.mix_blocks(j)
            if nb:
                launches.append(('m', len(arrays), len(j), nb))
                arrays.extend([j, t])

        # right identity environments F_b, b = N-1 .. 0
        for b in range(nmax - 1, -1, -1):
            g1, g2 = R.GemmList(), R.GemmList()
            for P in plans:
                s = P.s
                if b >= s.n:
                    continue
                d, Dl, Dr = s.d, s.dims[b], s.dims[b + 1]
                A = s.mps.ptr(b)
                g1.add(_p(P.Z), Dl * d, Dr, Dr, [(A, Dr, 0, P.F_at(b + 1), Dr, 1, Dr, 1.0)])
                g2.add(P.F_at(b), Dl, Dl, Dl, [(A, d * Dr, 0, _p(P.Z), d * Dr, 1, d * Dr, 1.0)])
            add_gemm(g1)
            add_gemm(g2)

        # left pass with strings; every closing is a Frobenius product into
        # dot slot k of the program
        self.dot_map = []   # (slot, plan, output index)
        ndot = 0
        for site in range(nmax):
            g1, mx, g2 = R.GemmList(), R.MixList(), R.GemmList()
            dots = []
            for P in plans:
                s = P.s
                if site >= s.n:
                    continue
                d, Dl, Dr = s.d, s.dims[site], s.dims[site + 1]
                A = s.mps.ptr(site)
                cur, nxt = P.X[site % 2], P.X[(site + 1) % 2]
                psi = P.psi
                Zp = lambda t: _p(P.Z, t * psi)
                Mp = lambda t: _p(P.M, t * psi)
                # Z_E and Z_X for chains alive at this site (started before it)
                g1.add(Zp(0), Dl, d * Dr, d * Dr, [(P.E_at(site), Dl, 0, A, d * Dr, 0, Dl, 1.0)])
                g2.add(P.E_at(site + 1), Dr, Dr, Dr, [(A, Dr, 1, Zp(0), Dr, 0, Dl * d, 1.0)])
                alive = [c for c in P.chains if c[0] < site <= P.chain_end[c]]
                for (i, a) in alive:
                    t = P.cidx[(i, a)]
                    g1.add(Zp(1 + t), Dl, d * Dr, d * Dr,
                           [(_p(cur, t * P.dsq), Dl, 0, A, d * Dr, 0, Dl, 1.0)])
                    if P.chain_end[(i, a)] > site:
                        g2.add(_p(nxt, t * P.dsq), Dr, Dr, Dr,
                               [(A, Dr, 1, Zp(1 + t), Dr, 0, Dl * d, 1.0)])
                nm = 0
                for (kind, ref, op, out) in P.close_at[site]:
                    src = Zp(0) if kind == 'E' else Zp(1 + ref)
                    mx.add(Mp(nm), L.F64, Dl, d, Dr, [(src, L.F64, 1.0, s.op_index(op))])
                    g2.add(_p(P.Y, nm * P.dsq), Dr, Dr, Dr, [(A, Dr, 1, Mp(nm), Dr, 0, Dl * d, 1.0)])
                    dots.append((_p(P.Y, nm * P.dsq), P.F_at(site + 1), Dr, Dr, Dr, Dr))
                    self.dot_map.append((ndot + len(dots) - 1, P, out))
                    nm += 1
                # strings that start here: X_(site, a) = A^T O_a (E A)
                for (i, a) in P.chains:
                    if i != site:
                        continue
                    t = P.cidx[(i, a)]
                    mx.add(Mp(nm), L.F64, Dl, d, Dr, [(Zp(0), L.F64, 1.0, s.op_index(a))])
                    g2.add(_p(nxt, t * P.dsq), Dr, Dr, Dr,
                           [(A, Dr, 1, Mp(nm), Dr, 0, Dl * d, 1.0)])
                    nm += 1
            add_gemm(g1)
            add_mix(mx)
            add_gemm(g2)
            if dots:
                launches.append(('d', len(arrays), len(dots), ndot))
                arrays.append(np.array(dots, dtype=L.DOT_JOB))
                ndot += len(dots)
        # the norms <psi|psi> = Tr E_N, into the slots after the values
        self.norm_slot = ndot
        launches.append(('d', len(arrays), len(plans), ndot))
        arrays.append(np.array([(P.E_at(P.s.n), 0, 1, 1, 1, 1) for P in plans], dtype=L.DOT_JOB))
        ndot += len(plans)
        self.dot_out = torch.zeros(max(ndot, 1), dtype=torch.float64, device=solves[0].dev)
        # all records in one device buffer (256-byte aligned arrays)
        offs, total = [], 0
        for a in arrays:
            offs.append(total)
            total += (a.nbytes + 255) // 256 * 256
        host = np.zeros(max(total, 256), dtype=np.uint8)
        for a, o in zip(arrays, offs):
            host[o:o + a.nbytes] = np.frombuffer(np.ascontiguousarray(a).tobytes(), np.uint8)
        self.records = torch.from_numpy(host).to(solves[0].dev)
        base = self.records.data_ptr()
        self.ptrs = [base + o for o in offs]
        self.launches = launches
        self.nbytes += self.records.numel() + self.dot_out.numel() * 8
        for P in plans:
            P.s = None   # the records hold what the replay needs

    def run(self, engine):
        """Launch the measurement on the engine's stream; returns per plan
        ``(one_vals, two_vals, norm)``."""
        self.launch(engine)
        return self.collect()

    def launch(self, engine):
        """Queue the measurement's kernels on the engine's stream (they read
        the solves' tensors as they are at this point of the stream)."""
        lib, stream, ptr = engine.lib, engine.stream(), self.ptrs
        opsp = ctypes_void(engine.ops.data_ptr())
        out = self.dot_out.data_ptr()
        for kind, i, n, x in self.launches:
            if kind == 'g':
                L.check(lib.tnx_gemm_grouped(L.GEMM_F64, ctypes_void(ptr[i]), n,
                                             ctypes_void(ptr[i + 1]), x, None, stream), 'gemm')
            elif kind == 'm':
                L.check(lib.tnx_mix(ctypes_void(ptr[i]), n, ctypes_void(ptr[i + 1]), x, opsp, None,
                                    stream), 'mix')
            else:
                L.check(lib.tnx_trace_dot(ctypes_void(ptr[i]), n, ctypes_void(out + 8 * x),
                                          stream), 'trace_dot')

    def collect(self):
        """The values of the last launch (waits for the stream up to here)."""
        vals = self.dot_out.cpu().numpy()
        for P in self.plans:
            P.values[:] = 0.0
        for (k, P, o) in self.dot_map:
            P.values[o] = vals[k]
        results = []
        for t, P in enumerate(self.plans):
            norm = float(vals[self.norm_slot + t])
            v = P.values / norm
            results.append((v[:len(P.one)], v[len(P.one):], norm))
        return results


MEASURE_CACHE = 8            # programs kept per engine (least recently used dropped)
MEASURE_CACHE_BYTES = 8 << 30


def _measure_real(engine, solves, requests):
    """:func:`measure` for operators whose real parts are meant."""
    return _measure_real_end(_measure_real_begin(engine, solves, requests))


def _measure_real_begin(engine, solves, requests, keys=None):
    engine.settle()   # solve state is final once the engine's host work is done
    # a program serves any batch of the same (solve, request) pairs, in any
    # order: its results are permuted back to the caller's order
    keys = [k if keys is not None and keys[t] is not None else
            (_solve_key(s), tuple(map(tuple, one)), tuple(map(tuple, two)))
            for t, (s, (one, two), k) in enumerate(zip(solves, requests, keys or [None] * len(solves)))]
    order = sorted(range(len(keys)), key=lambda k: keys[k])
    key = tuple(keys[k] for k in order)
    cache = getattr(engine, '_measure_cache', None)
    if cache is None:
        cache = engine._measure_cache = collections.OrderedDict()
    prog = cache.get(key)
    if prog is None:
        prog = _MeasureProgram(engine, [solves[k] for k in order], [requests[k] for k in order])
        cache[key] = prog
        while len(cache) > MEASURE_CACHE or (
                len(cache) > 1 and sum(p.nbytes for p in cache.values()) > MEASURE_CACHE_BYTES):
            cache.popitem(last=False)
    else:
        cache.move_to_end(key)
    prog.launch(engine)
    return prog, order


def _measure_real_end(handle):
    prog, order = handle
    res = prog.collect()
    out = [None] * len(order)
    for t, k in enumerate(order):
        out[k] = res[t]
    return out
