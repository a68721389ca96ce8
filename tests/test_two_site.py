"""Two-site DMRG (no GPU): the merged-pair term tables, the sweep schedule and
the step plans."""
import numpy as np
import pytest

from oracle import block_dmrg, dmrg_ref, ed, model
from tnalg_amd import parameters as P
from tnalg_amd.terms import compile_terms, dense_heff


def _long_range_para():
    bonds = [(i, i + 1) for i in range(7)] + [(0, 3), (2, 6), (1, 7), (5, 7), (3, 4)]
    jj = [1.0] * 7 + [0.3, -0.4, 0.2, 0.7, 0.25]
    c = model.couplings_from_bonds(8, bonds, jj, [0.5 * x for x in jj], hz=0.2, hx=0.1)
    return c, dmrg_ref.make_para(c, 6, 1, eigs_tol=1e-12)


def _basis(mps, lo, hi):
    """Row basis of sites lo..hi-1 (states x bond): left block for lo = 0,
    right block (transposed) for hi = N."""
    if lo == 0:
        v = np.ones((1, 1))
        for k in range(hi):
            a = mps[k]
            v = (v @ a.reshape(a.shape[0], -1)).reshape(-1, a.shape[2])
        return v
    v = np.ones((1, 1))
    for k in range(hi - 1, lo - 1, -1):
        a = mps[k]
        v = (a.reshape(-1, a.shape[2]) @ v).reshape(a.shape[0], -1)
    return v.T


@pytest.mark.parametrize('seed', [0, 1])
def test_pair_tables_reproduce_projected_hamiltonian(seed):
    """Dense Heff of every merged pair (p, p+1) from the two-site tables and
    the block environments at bonds p and p+2 equals P^T H P, P the isometry
    of the canonical MPS around the pair and H the exact Hamiltonian."""
    c, para = _long_range_para()
    n = 8
    eng = block_dmrg.BlockDmrg(para, np.random.RandomState(seed))
    eng.A.correct_orthogonal_center(0)
    eng.initial_environments()
    T = compile_terms(para, pairs=True)
    assert len(T.pairs) == n - 1
    H = ed.hamiltonian(n, c['op'], c['index1'], c['coeff1'], c['index2'], c['coeff2']).toarray()
    for p in range(n - 1):
        eng.move_center(p)
        le, re = eng.left[p], eng.right[p + 2]
        left = (le['H'], [le[k] for k in T.left_opens[p]])
        right = (re['H'], [re[k] for k in T.right_opens[p + 2]])
        mps = eng.A.mps
        dl, dr = mps[p].shape[0], mps[p + 1].shape[2]
        h = dense_heff(T, p, left, right, dl, dr, pair=True)
        PL, PR = _basis(mps, 0, p), _basis(mps, p + 2, n)
        Pm = np.kron(np.kron(PL, np.eye(4)), PR)
        assert np.allclose(Pm.T @ H @ Pm, h, atol=1e-12), p


def test_pair_tables_share_coefficients_with_one_site():
    """The two-site tables tag their coefficients into the same value list,
    after the one-site ones; the one-site tables are unchanged."""
    _, para = _long_range_para()
    a, b = compile_terms(para), compile_terms(para, pairs=True)
    assert np.array_equal(b.coef_values[:len(a.coef_values)], a.coef_values)
    assert len(b.coef_values) > len(a.coef_values)
    assert [s.lm for s in a.sites] == [s.lm for s in b.sites]
    assert a.structure() != b.structure()
    # merged operators are kron products of the site operators
    for key, o in zip(b.ops2_keys, b.ops2):
        if key[0] == 'LR':
            assert np.allclose(o, np.kron(b.ops[key[1]], b.ops[key[2]]))


@pytest.mark.parametrize('n,ob', [(2, 0), (2, 1), (3, 0), (8, 0), (8, 3), (8, 7)])
def test_two_site_sweep_route(n, ob):
    """Every pair step finds the centre on one of its sites and the split
    direction leads to the next pair; the sweep ends with the centre at ob."""
    from test_plan_reloc import CpuEngine
    from tnalg_amd.engine import Solve
    para = P.make_consistent_parameter_dmrg(dict(P.generate_parameters_dmrg('chain'), l=n,
                                                 chi=4, dmrg_sites=2, ob_position=ob))
    s = Solve(CpuEngine(lowp=False), para, seed=0)
    s.schedule_canonicalise()
    s.queue.clear()
    s.schedule_sweep(0)
    steps = list(s.queue)
    assert steps[-1].kind == 'split2' and all(st.kind == 'solve2' for st in steps[:-1])
    centre = ob
    for t, st in enumerate(steps):
        if st.kind == 'solve2':
            if t > 0:
                prev = st.site - st.dir
                assert prev == steps[t - 1].site and st.dir in (1, -1)
                centre = prev + (st.dir > 0)
            assert centre in (st.site, st.site + 1), (t, st, centre)
        else:
            centre = st.site + (st.dir > 0)
    assert centre == ob
    pairs = [st.site for st in steps[:-1]]
    assert sorted(set(pairs)) == list(range(n - 1))   # every pair solved
    # from an end: 2N - 3 pair solves (each turning pair once); from inside: 2N - 4
    want = 1 if n == 2 else (2 * n - 3 if ob in (0, n - 1) else 2 * n - 4)
    assert len(pairs) == want


def test_two_site_needs_d_le_2():
    para = P.generate_parameters_dmrg('chain')
    para.update(spin='one', dmrg_sites=2)
    para = P.make_consistent_parameter_dmrg(para)
    from tnalg_amd.engine import solve_geometry
    if para['d'] == 2:
        pytest.skip('spin-1 lattice not available')
    with pytest.raises(ValueError):
        solve_geometry(para, 8, False)


def test_two_site_then_one_site_sweeps():
    """two_site_sweeps = 1: the first sweep solves pairs, the next ones sites."""
    from test_plan_reloc import CpuEngine
    from tnalg_amd.engine import Solve
    para = P.make_consistent_parameter_dmrg(dict(P.generate_parameters_dmrg('chain'), l=8,
                                                 chi=8, dmrg_sites=2, two_site_sweeps=1))
    s = Solve(CpuEngine(lowp=False), para, seed=0)
    s.schedule_canonicalise()
    s.queue.clear()
    s.schedule_sweep(0)
    kinds0 = {st.kind for st in s.queue}
    s.queue.clear()
    s.schedule_sweep(1)
    kinds1 = {st.kind for st in s.queue}
    assert kinds0 == {'solve2', 'split2'} and kinds1 <= {'solve', 'move'} and 'solve' in kinds1


@pytest.mark.parametrize('kw', [dict(chi=2), dict(chi=600), dict(chi=5, chi_init=3),
                                dict(chi=8, chi_init=9), dict(dmrg_sites=3)])
def test_two_site_limits_raise_at_admission(kw):
    """Limits of the two-site split are checked when the solve is created
    (a server rejects such a request alone), not when a step is planned."""
    from tnalg_amd.engine import solve_geometry
    para = P.make_consistent_parameter_dmrg(dict(P.generate_parameters_dmrg('chain'), l=10,
                                                 **dict(dict(dmrg_sites=2), **kw)))
    with pytest.raises(ValueError):
        solve_geometry(para, 8, False)


def test_compiled_terms_cache_is_thread_safe():
    """Request preparation compiles terms on worker threads: concurrent
    lookups, inserts and evictions of a small cache neither raise nor lose
    entries' structure."""
    import concurrent.futures
    from tnalg_amd.terms import compile_terms_cached
    paras = [P.make_consistent_parameter_dmrg(dict(P.generate_parameters_dmrg('chain'), l=6 + k % 5,
                                                   jz=0.5 + 0.01 * k)) for k in range(40)]

    def work(k):
        T = compile_terms_cached(paras[k % 40], maxsize=4)
        return T.n == paras[k % 40]['l']
    with concurrent.futures.ThreadPoolExecutor(8) as ex:
        assert all(ex.map(work, range(400)))
