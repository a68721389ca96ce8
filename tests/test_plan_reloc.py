"""Shared step plans (no GPU): a plan built by one solve and relocated to
another solve's arena and operator slot equals the plan that solve builds for
itself from its own addresses."""
import numpy as np
import pytest
import torch

from tnalg_amd import _lib as L
from tnalg_amd import parameters as P
from tnalg_amd import records as R
from tnalg_amd.engine import Solve, Step
from tnalg_amd.terms import Coef


class CpuEngine:
    """The parts of Engine that Solve uses, on host memory."""
    device = torch.device('cpu')
    ncv = 20
    init_rng = 'numpy'

    def __init__(self, lowp):
        self.uses_lowp = lowp
        self.plan_cache = {}
        self._next_slot = 0
        self._moved_pool = {}
        self.moved_pool_limit = 0

    def arena_get(self, nbytes):
        return torch.empty(nbytes, dtype=torch.uint8)

    def arena_fill(self, arena, stage):
        arena[:stage.numel()].copy_(stage)
        arena[stage.numel():].zero_()

    def arena_release(self, arena):
        pass

    ops_off2 = 1024

    def alloc_ops(self, ops, ops2=None):
        s = self._next_slot
        n = ops.shape[0] if ops2 is None else max(ops.shape[0], ops2.shape[0])
        self._next_slot += n + 3   # distinct slots per solve
        return s

    def free_ops(self, slot):
        pass


def _para(lattice, **kw):
    para = P.generate_parameters_dmrg(lattice)
    para.update(kw)
    return P.make_consistent_parameter_dmrg(para)


def _records(ph):
    """(path, record array) of every record array of a plan, in a fixed order."""
    out = []

    def walk(path, obj):
        if isinstance(obj, np.ndarray) and obj.dtype.names is not None:
            out.append((path, obj))
        elif isinstance(obj, (tuple, list)):
            for k, o in enumerate(obj):
                walk(path + (k,), o)
    for key in sorted(k for k in ph if k not in R.PLAN_META):
        walk((key,), ph[key])
    return out


@pytest.mark.parametrize('lattice,kw,kw_b,prec', [
    ('chain', dict(l=10, chi=16), {}, 'bf16'),
    ('chain', dict(l=12, chi=300), {}, 'fp64'),          # blocked CholeskyQR3 in the middle
    ('snakeJ1J2', dict(square_width=3, square_height=4, chi=12), {}, 'fp32'),
    # same structure, other couplings and field: the plan is still shared
    ('chain', dict(l=10, chi=16, jz=1.0, hz=0.2), dict(jz=0.7, hz=0.45), 'bf16'),
    ('snakeJ1J2', dict(square_width=3, square_height=4, chi=12), dict(j2=0.3), 'fp64'),
    # two-site sweeps: truncating SVD splits, small QR splits, growing bonds
    ('chain', dict(l=10, chi=16, dmrg_sites=2), {}, 'bf16'),
    ('chain', dict(l=7, chi=8, chi_init=2, dmrg_sites=2, ob_position=3), {}, 'fp64'),
    ('chain', dict(l=10, chi=16, dmrg_sites=2, jz=1.0, hz=0.2), dict(jz=0.7, hz=0.45), 'fp32'),
    ('snakeJ1J2', dict(square_width=3, square_height=4, chi=12, dmrg_sites=2), dict(j2=0.3),
     'fp64'),
])
def test_relocated_plan_equals_own_plan(lattice, kw, kw_b, prec):
    para = _para(lattice, **kw)
    para_b = _para(lattice, **dict(kw, **kw_b))
    eng = CpuEngine(lowp=prec != 'fp64')
    a = Solve(eng, para, seed=1)
    b = Solve(eng, para_b, seed=2)
    assert a.sig == b.sig and a.reloc_key != b.reloc_key
    if kw_b:
        assert not np.array_equal(a.T.coef_values, b.T.coef_values)
    a.schedule_canonicalise()
    a.schedule_sweep(0)
    checked = 0
    for st in list(a.queue):
        dims0 = list(b.dims)
        ph_rel = a.plan(st, prec)                      # shared, relative
        own = b._build_plan(st, prec)                  # b's addresses and coefficients
        b.dims = dims0
        assert b.plan(st, prec) is ph_rel              # cache hit for b
        assert list(b.dims) == own['dims_after']
        ra, rb = _records(ph_rel), _records(own)
        assert [p for p, _ in ra] == [p for p, _ in rb]
        for (path, rel), (_, absr) in zip(ra, rb):
            moved = R.relocate(rel, b.reloc)
            want = R.relocate(absr, (0, 0, b.T.coef_values))   # b's coefficient values
            assert moved.tobytes() == want.tobytes(), (st, path)
            # the packed relocation that steps use gives the same records
            got = b.part(ph_rel, path[0])
            for k in path[1:]:
                got = got[k]
            assert got.tobytes() == want.tobytes(), (st, path)
            checked += 1
    assert checked > 0


def test_relocation_keeps_null_and_identity():
    g = R.GemmList()
    g.add(4096, 2, 2, 2, [(8192, 2, 0, 8200, 2, 0, 2, 1.0)])          # c0 = 0 (null)
    p, s = g.arrays()
    m = R.MixList()
    m.add(4096, L.F64, 1, 2, 1, [(8192, L.F64, 1.0, -1), (8192, L.F64, Coef(2.0, 1), 3)])
    j, t = m.arrays()
    ph = {'g': (p, s), 'm': (j, t)}
    R.make_relative(ph, 4000, 9000, 2)
    assert p['c'][0] == 96 and p['c0'][0] == 0 and s['b'][0] == 4200
    assert t['op'].tolist() == [-1, 1]
    q = R.relocate(t, (100000, 10, np.array([5.0, -0.25])))
    assert q['src'].tolist() == [104192, 104192] and q['op'].tolist() == [-1, 11]
    assert q['coef'].tolist() == [1.0, -0.25]   # the tagged coefficient takes the solve's value
    with pytest.raises(ValueError):   # an address outside the arena is a planning bug
        R.make_relative({'s': s.copy()}, 9000, 9100, 0)
    m = R.MixList()
    m.add(4096, L.F64, 1, 2, 1, [(8192, L.F64, 2.0, 3)])
    with pytest.raises(ValueError):   # an untagged coupling value cannot be shared
        R.make_relative({'m': m.arrays()}, 4000, 9000, 0)
    with pytest.raises(ValueError):
        R.pack_relative({'m': m.arrays()}, 4000, 9000, 0)


def test_packed_plan_relocation():
    """Packing: nulls, identity operators and literal coefficients stay, every
    other address, operator index and tagged coefficient moves; empty phases
    and nested stage lists keep their shape."""
    g = R.GemmList()
    g.add(4096, 2, 2, 2, [(8192, 2, 0, 8200, 2, 0, 2, 1.0)])          # c0 = 0 (null)
    m = R.MixList()
    m.add(4104, L.F64, 1, 2, 1, [(8192, L.F64, 1.0, -1), (8192, L.F64, Coef(2.0, 1), 3),
                                 (8800, L.F64, Coef(7.0, 0), 2)])
    e = R.MixList()
    ph = {'g': g.arrays(), 'm': m.arrays(), 'e': e.arrays(),
          'cq': [('g', g.arrays()), ('c', np.zeros(0, L.CHOL_JOB))], 'uid': 5, 'prec': 'bf16'}
    pk = R.pack_relative(ph, 4000, 9000, 2)
    assert ph['packed'] is pk and ph['uid'] == 5
    p, s = ph['g']
    assert p['c'].tolist() == [96] and p['c0'].tolist() == [0] and s['b'].tolist() == [4200]
    j, t = ph['m']
    assert j['dst'].tolist() == [104] and t['op'].tolist() == [-1, 1, 0]
    assert len(ph['e'][0]) == 0 and ph['cq'][0][0] == 'g' and len(ph['cq'][1][1]) == 0
    buf = pk.relocate((100000, 10, np.array([5.0, -0.25])))
    j2, t2 = pk.views(buf, 'm')
    assert j2['dst'].tolist() == [100104] and t2['src'].tolist() == [104192, 104192, 104800]
    assert t2['op'].tolist() == [-1, 11, 10]
    assert t2['coef'].tolist() == [1.0, -0.25, 5.0]
    kind, (p2, s2) = pk.views(buf, 'cq')[0]
    assert p2['c0'].tolist() == [0] and s2['a'].tolist() == [104192]
    assert t['src'].tolist() == [4192, 4192, 4800]    # the relative plan is unchanged
    with pytest.raises(ValueError):   # an address outside the arena is a planning bug
        R.pack_relative({'s': g.arrays()}, 9000, 9100, 0)


def test_concat_relocates_each_part():
    g = R.GemmList()
    g.add(4096, 70, 2, 2, [(8192, 2, 0, 8200, 2, 0, 2, 1.0)])
    part = g.arrays()
    R.make_relative({'g': part}, 4000, 9000, 0)
    p, s, nt = R.concat_gemm([part, part], relocs=[(1 << 20, 0), (1 << 30, 0)])
    assert nt == 4 and p['tile_begin'].tolist() == [0, 2] and p['seg_begin'].tolist() == [0, 1]
    assert p['c'].tolist() == [(1 << 20) + 96, (1 << 30) + 96]
    assert s['a'].tolist() == [(1 << 20) + 4192, (1 << 30) + 4192]
    assert part[0]['c'][0] == 96      # the shared relative plan is left unchanged


def _naive_concat(parts, rec_dt, sub_dt, sub_field, unit_field, units, actives=None):
    """Part-by-part concatenation (the definition the batched form must match)."""
    recs, subs = [], []
    sub_off = unit_off = 0
    for k, (a, b) in enumerate(parts):
        if len(a) == 0:
            continue
        a = a.copy()
        a[sub_field] += sub_off
        a[unit_field] += unit_off
        if actives is not None:
            a['active'] = actives[k]
        unit_off += units(parts[k][0])
        sub_off += len(b)
        recs.append(a)
        subs.append(b)
    if not recs:
        return np.zeros(0, rec_dt), np.zeros(0, sub_dt), 0
    return np.concatenate(recs), np.concatenate(subs), unit_off


def test_batched_concat_matches_part_by_part():
    rng = np.random.RandomState(3)
    gparts, mparts = [], []
    for k in range(9):
        g = R.GemmList()
        m = R.MixList()
        for _ in range(rng.randint(0, 4)):   # some parts empty
            M, N = int(rng.randint(1, 200)), int(rng.randint(1, 200))
            g.add(4096 * k, M, N, M, [(8192, M, 0, 8200, 2, 0, int(rng.randint(1, 9)), 1.0)] *
                  int(rng.randint(1, 4)))
            Dl, Dr = int(rng.randint(1, 90)), int(rng.randint(1, 90))
            m.add(4096 * k, L.F64, Dl, 2, Dr, [(8192, L.F64, 1.0, -1)] * int(rng.randint(1, 3)))
        gparts.append(g.arrays())
        mparts.append(m.arrays())
    acts = list(range(100, 109))
    for actives in (None, acts):
        p, s, nt = R.concat_gemm(gparts, actives)
        q, t, nq = _naive_concat(gparts, L.GEMM_PROBLEM, L.GEMM_SEG, 'seg_begin', 'tile_begin',
                                 R.gemm_tiles, actives)
        assert nt == nq and p.tobytes() == q.tobytes() and s.tobytes() == t.tobytes()
        j, u, nb = R.concat_mix(mparts, actives)
        q, t, nq = _naive_concat(mparts, L.MIX_JOB, L.MIX_TERM, 'term_begin', 'block_begin',
                                 R.mix_blocks, actives)
        assert nb == nq and j.tobytes() == q.tobytes() and u.tobytes() == t.tobytes()
    assert R.concat_gemm([gparts[0]][:0])[2] == 0


def test_assembled_step_matches_relocate_then_concat():
    """A batch of solves of different structures at different steps: the
    assembled descriptors equal relocating every part and concatenating part
    by part, and the cached relocated parts are left unchanged."""
    eng = CpuEngine(lowp=True)
    solves = [Solve(eng, _para('chain', l=10, chi=16), seed=1),
              Solve(eng, _para('chain', l=12, chi=300), seed=2),
              Solve(eng, _para('chain', l=10, chi=16, jz=0.6), seed=3),
              Solve(eng, _para('snakeJ1J2', square_width=3, square_height=4, chi=12), seed=4)]
    for s in solves:
        s.schedule_canonicalise()
        s.schedule_sweep(0)
    from tnalg_amd.engine import assemble_pre, assemble_lanczos
    steps = 0
    while all(s.queue for s in solves):
        pairs = [(s, s.plan(s.queue.pop(0), 'bf16')) for s in solves]
        arrs, meta = assemble_pre(pairs)
        # the same phases, relocated part by part
        for i, key in enumerate(('absorb', 'env_g1', 'env_gT')):
            parts = [R.relocate_tree(ph[key], s.reloc) for s, ph in pairs if key in ph]
            p, sg, nt = _naive_concat(parts, L.GEMM_PROBLEM, L.GEMM_SEG, 'seg_begin',
                                      'tile_begin', R.gemm_tiles)
            assert arrs[2 + 2 * i].tobytes() == p.tobytes() and meta['gemm'][i][1] == nt
            assert arrs[3 + 2 * i].tobytes() == sg.tobytes()
        for i, key in enumerate(('copy', 'env_lc', 'env_mix', 'cast')):
            parts = [R.relocate_tree(ph[key], s.reloc) for s, ph in pairs if key in ph]
            j, t, nb = _naive_concat(parts, L.MIX_JOB, L.MIX_TERM, 'term_begin', 'block_begin',
                                     R.mix_blocks)
            assert arrs[8 + 2 * i].tobytes() == j.tobytes() and meta['mix'][i][1] == nb
            assert arrs[9 + 2 * i].tobytes() == t.tobytes()
        group = [(s, ph) for s, ph in pairs if 'lz' in ph]
        if group:
            # two groups (as for two matvec precisions): the second group's
            # actives continue the first's, the solve records follow group order
            groups = [group[:1], group[1:]] if len(group) > 1 else [group]
            la, lm = assemble_lanczos(groups)
            first = 0
            for k, g in enumerate(groups):
                parts = [R.relocate_tree(ph['mv_g1'], s.reloc) for s, ph in g]
                p, sg, nt = _naive_concat(parts, L.GEMM_PROBLEM, L.GEMM_SEG, 'seg_begin',
                                          'tile_begin', R.gemm_tiles,
                                          list(range(first, first + len(g))))
                assert la[3 + 6 * k].tobytes() == p.tobytes() and lm['groups'][k][2][1] == nt
                first += len(g)
            lz = np.concatenate([R.relocate(ph['lz'], s.reloc) for s, ph in group])
            assert (la[2]['V'] == lz['V']).all() and la[2]['chunk_begin'][0] == 0
            assert (la[2]['x_dtype'] == L.BF16).all() and (la[2]['w_dtype'] == L.F32).all()
        # a cached relocated part is returned as is the next time
        s, ph = pairs[0]
        key = next(k for k in ('env_g1', 'absorb', 'copy') if k in ph)
        before = R.relocate_tree(ph[key], s.reloc)
        got = s.part(ph, key)
        assert got is s.part(ph, key)
        assert all(a.tobytes() == b.tobytes() for a, b in zip(got, before))
        steps += 1
    assert steps > 5


def test_plan_ahead_matches_inline_planning():
    """The worker's next-step host work: each solve's planned-ahead step is
    the plan inline planning returns, the solve's dims are left as they were,
    and the speculative descriptor sets equal the ones assembled inline."""
    from tnalg_amd.engine import (Engine, assemble_lanczos, assemble_pre, _lanczos_groups,
                                  _lanczos_key, _split_groups)

    class AheadEngine(CpuEngine):
        _plan_ahead = Engine._plan_ahead
        _prec_of = Engine._prec_of
        sweep_precision = Engine.sweep_precision

        def __init__(self):
            super().__init__(lowp=True)
            self._resident = {}
            self.precision_schedule = lambda sweep, nsweeps: ('bf16', 'fp32', 'fp64')[sweep]
            self.lanczos_streams, self.split_min = 2, 2   # the Lanczos batch in two calls

    eng = AheadEngine()
    solves = [Solve(eng, _para('chain', l=10, chi=16, sweep_time=3), seed=1),
              Solve(eng, _para('chain', l=8, chi=8, sweep_time=3), seed=2),
              Solve(eng, _para('snakeJ1J2', square_width=3, square_height=4, chi=12,
                               sweep_time=3), seed=3)]
    for s in solves:
        s.schedule_canonicalise()
        for t in range(3):
            s.schedule_sweep(t)
    steps = 0
    while all(s.queue for s in solves):
        dims0 = [list(s.dims) for s in solves]
        spec = eng._plan_ahead(solves)
        assert [list(s.dims) for s in solves] == dims0
        pairs = []
        for s in solves:
            st, prec, ph = s._next
            assert st is s.queue[0] and prec == eng._prec_of(s, st)
            assert s.plan(st, prec) is ph          # inline planning: the same shared plan
            assert list(s.dims) == list(ph['dims_after'])
            pairs.append((s, ph))
            s.queue.pop(0)
            s._next = None
        arrs, meta = assemble_pre(pairs)
        assert spec['pre'][1][1] == meta
        assert all(a.tobytes() == b.tobytes() for a, b in zip(spec['pre'][1][0], arrs))
        groups = _lanczos_groups(pairs)
        if groups:
            la, lm = assemble_lanczos(groups)
            parts = _split_groups(groups, 2, 2)
            nsolve = sum(1 for _, ph in pairs if 'lz' in ph)
            assert len(parts) == (2 if nsolve >= 2 else 1)
            assert sum(len(g) for p in parts for g in p) == nsolve
            assert [k for k, _ in spec['lz']] == [_lanczos_key(p) for p in parts]
            for part, (_, (arrs_a, meta_a)) in zip(parts, spec['lz']):
                la, lm = assemble_lanczos(part)
                assert meta_a == lm
                assert all(a.tobytes() == b.tobytes() for a, b in zip(arrs_a, la))
        else:
            assert not spec.get('lz')
        steps += 1
    assert steps > 20


def test_gemm_kmax_and_tuner_encoding():
    """Largest total K of a launch's problems; the tuner's shape classes and
    the variant carried in the mode's second byte (untuned classes: 0)."""
    from tnalg_amd.engine import GemmTuner, gemm_kmax
    g = R.GemmList()
    g.add(4096, 8, 8, 8, [(8192, 8, 0, 8200, 8, 0, 100, 1.0), (8192, 8, 0, 8200, 8, 0, 28, 1.0)])
    g.add(4096, 8, 8, 8, [(8192, 8, 0, 8200, 8, 0, 300, 1.0)])
    g.add(4096, 8, 8, 8, [(8192, 8, 0, 8200, 8, 0, 0, 1.0)])   # no segment with K > 0
    p, s = g.arrays()
    assert gemm_kmax(p, s) == 300 and gemm_kmax(p[:1], s) == 128
    assert gemm_kmax(p[:0], s) == 0

    class FakeLib:
        @staticmethod
        def tnx_gemm_variants(m):
            return 2 if m == L.GEMM_F64 else 1

    t = GemmTuner(FakeLib())
    assert t.nvar[L.GEMM_F64] == 2 and t.nvar[L.GEMM_BF16_F32] == 1
    assert t.mode_for(L.GEMM_F64, 2048, 128) == L.GEMM_F64
    t.table[t.shape_class(L.GEMM_F64, 2048, 128)] = 1
    assert t.mode_for(L.GEMM_F64, 2048, 128) == L.GEMM_F64 | (1 << 8)
    assert t.mode_for(L.GEMM_F64, 3000, 200) == L.GEMM_F64 | (1 << 8)   # same class
    assert t.mode_for(L.GEMM_F64, 1500, 100) == L.GEMM_F64              # another class
    assert t.mode_for(L.GEMM_F64, 512, 512) == L.GEMM_F64
    t.tune(L.GEMM_F32, 512, 512, None)      # a mode with one variant is not timed
    assert t.shape_class(L.GEMM_F32, 512, 512) not in t.table
