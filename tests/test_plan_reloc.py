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

    def __init__(self, lowp):
        self.uses_lowp = lowp
        self.plan_cache = {}
        self._next_slot = 0

    def arena_get(self, nbytes):
        return torch.empty(nbytes, dtype=torch.uint8)

    def arena_fill(self, arena, stage):
        arena[:stage.numel()].copy_(stage)
        arena[stage.numel():].zero_()

    def arena_release(self, arena):
        pass

    def alloc_ops(self, ops):
        s = self._next_slot
        self._next_slot += ops.shape[0] + 3   # distinct slots per solve
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
    for key in sorted(k for k in ph if k not in ('dims_after', 'uid', 'prec', 'mv_flops',
                                                  'svd_n', 'lm_bond')):
        walk((key,), ph[key])
    return out


@pytest.mark.parametrize('lattice,kw,kw_b,prec', [
    ('chain', dict(l=10, chi=16), {}, 'bf16'),
    ('chain', dict(l=12, chi=300), {}, 'fp64'),          # blocked CholeskyQR3 in the middle
    ('snakeJ1J2', dict(square_width=3, square_height=4, chi=12), {}, 'fp32'),
    # same structure, other couplings and field: the plan is still shared
    ('chain', dict(l=10, chi=16, jz=1.0, hz=0.2), dict(jz=0.7, hz=0.45), 'bf16'),
    ('snakeJ1J2', dict(square_width=3, square_height=4, chi=12), dict(j2=0.3), 'fp64'),
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
        assert b.dims == own['dims_after']
        ra, rb = _records(ph_rel), _records(own)
        assert [p for p, _ in ra] == [p for p, _ in rb]
        for (path, rel), (_, absr) in zip(ra, rb):
            moved = R.relocate(rel, b.reloc)
            want = R.relocate(absr, (0, 0, b.T.coef_values))   # b's coefficient values
            assert moved.tobytes() == want.tobytes(), (st, path)
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
