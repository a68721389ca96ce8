"""Host logic (no GPU): coupling-list generators and the term compiler."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import block_dmrg, dmrg_ref, model
from tnalg_amd import parameters as P
from tnalg_amd.terms import compile_terms, dense_heff


def test_chain_rows_match_reference_convention():
    para = P.generate_parameters_dmrg('chain')
    c = model.chain_couplings(para['l'], para['jxy'], para['jz'])
    assert np.array_equal(para['index2'], c['index2'])
    assert np.allclose(para['coeff2'], c['coeff2'])
    assert np.array_equal(para['index1'], c['index1'])
    assert para['nh'] == 3 * (para['l'] - 1)
    assert para['d'] == 2 and len(para['op']) == 7
    assert np.allclose(para['op'][6], -para['hx'] * para['op'][1] - para['hz'] * para['op'][3])


def test_override_and_remake_pattern():
    """Callers override para entries and re-call make_consistent_parameter_dmrg
    (ScriptRun/DMRG/runDMRGQIsingChain.py:12-33)."""
    para = P.generate_parameters_dmrg('chain')
    para['l'], para['jz'], para['hz'], para['chi'] = 20, 0.5, 0.25, 64
    para = P.make_consistent_parameter_dmrg(para)
    assert para['index2'].shape == (57, 4)
    assert np.allclose(para['coeff2'][2::3], 0.5)
    assert np.allclose(para['op'][6], np.diag([-0.125, 0.125]))
    assert para['data_exp'] == 'chainN20_j(1,0.5)_h(0,0.25)_chi64open'


@pytest.mark.parametrize('lat', ['square', 'jigsaw', 'full', 'longRange', 'husimi', 'arbitrary'])
def test_other_lattices_build(lat):
    para = P.generate_parameters_dmrg(lat)
    assert para['index2'].shape[1] == 4
    assert para['positions_h2'].max() == para['l'] - 1
    assert len(np.ravel(para['coeff2'])) == para['index2'].shape[0]


def test_square_row_major_bonds():
    pos = P.positions_nearest_neighbor_square(3, 2)
    assert pos.tolist() == [[0, 1], [1, 2], [3, 4], [4, 5], [0, 3], [1, 4], [2, 5]]


def test_snake_j1j2_counts_and_ranges():
    p1, p2 = P.positions_snake_j1j2(6, 6)
    assert len(p1) == 60 and len(p2) == 50
    assert np.all(p1[:, 0] < p1[:, 1]) and np.all(p2[:, 0] < p2[:, 1])
    assert int(np.max(p1[:, 1] - p1[:, 0])) == 11
    p1, p2 = P.positions_snake_j1j2(8, 8)
    assert len(p1) + len(p2) == 112 + 98
    # every lattice neighbour pair is at Manhattan distance 1 (J1) or diagonal (J2)
    def xy(k, w=8):
        y = k // w
        x = k % w if y % 2 == 0 else w - 1 - k % w
        return x, y
    for a, b in p1:
        (xa, ya), (xb, yb) = xy(a), xy(b)
        assert abs(xa - xb) + abs(ya - yb) == 1
    for a, b in p2:
        (xa, ya), (xb, yb) = xy(a), xy(b)
        assert abs(xa - xb) == 1 and abs(ya - yb) == 1


def test_snake_para_c3():
    para = P.generate_parameters_dmrg('snakeJ1J2')
    assert para['l'] == 36 and para['nh'] == 3 * 110
    j = np.ravel(para['coeff2'])[2::3]
    assert sorted(set(j.tolist())) == [0.5, 1.0]


def test_from_index2_positions():
    idx = np.array([[2, 3, 3, 3], [0, 1, 4, 5], [0, 1, 5, 4], [1, 2, 3, 3]])
    assert P.from_index2_to_positions_h2(idx).tolist() == [[0, 1], [1, 2], [2, 3]]


def _block_envs(eng, p):
    T = eng.T_compiled
    le, re = eng.left[p], eng.right[p + 1]
    left = (le['H'], [le[k] for k in T.left_opens[p]])
    right = (re['H'], [re[k] for k in T.right_opens[p + 1]])
    return left, right


@pytest.mark.parametrize('seed', [0, 1])
def test_term_tables_reproduce_reference_heff(seed):
    """Dense Heff from the compiled term tables + block environments equals the
    reference's per-term sum_c kron(vL, vM, vR) (MPSClass.py:613-662) at every site."""
    bonds = [(i, i + 1) for i in range(7)] + [(0, 3), (2, 6), (1, 7), (5, 7)]
    jj = [1.0] * 7 + [0.3, -0.4, 0.2, 0.7]
    c = model.couplings_from_bonds(8, bonds, jj, [0.5 * x for x in jj], hz=0.2, hx=0.1)
    para = dmrg_ref.make_para(c, 6, 1, eigs_tol=1e-12)
    eng = block_dmrg.BlockDmrg(para, np.random.RandomState(seed))
    eng.A.correct_orthogonal_center(0)
    eng.initial_environments()
    eng.T_compiled = compile_terms(para)
    for p in range(8):
        eng.move_center(p)
        left, right = _block_envs(eng, p)
        s = eng.A.mps[p].shape
        h_terms = dense_heff(eng.T_compiled, p, left, right, s[0], s[2])
        eng.A.update_all_effective_id()
        h_ref, _ = eng.A.effective_hamiltonian_dense(p, para['index1'], para['index2'],
                                                     para['coeff1'], para['coeff2'], tol=1e-12)
        assert np.allclose(h_terms, h_ref, atol=1e-12), p


def test_terms_reject_bad_rows():
    para = P.generate_parameters_dmrg('arbitrary')
    para['index2'] = np.array([[1, 1, 3, 3]])
    para['coeff2'] = np.array([[1.0]])
    with pytest.raises(ValueError):
        compile_terms(para)


def test_terms_swap_reversed_rows():
    base = model.chain_couplings(4, 1.0, 1.0)
    para = dict(base, eigs_tol=0.0)
    rev = dict(para)
    rev['index2'] = para['index2'][:, [1, 0, 3, 2]]
    a, b = compile_terms(para), compile_terms(rev)
    assert a.left_opens == b.left_opens and a.right_opens == b.right_opens
    assert [s.lm for s in a.sites] == [s.lm for s in b.sites]


# ---- the C-ABI library (loads on a CPU-only host; no compute calls)
def test_library_exports_header_symbols():
    from tnalg_amd import _lib
    lib = _lib.load()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, 'include', 'tnalg_hip.h')).read()
    names = set(re.findall(r'\b(tnx_[a-z0-9_]+)\s*\(', hdr))
    assert names, 'no declarations found'
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert lib.tnx_version() == 10000
    assert set(_lib.SIGNATURES) == names


def test_library_abi_struct_sizes():
    from tnalg_amd import _lib
    lib = _lib.load()
    sizes = (ctypes.c_int32 * 32)()
    n = lib.tnx_abi_sizes(sizes, 32)
    assert list(sizes[:n]) == [48, 56, 48, 24, 112, 96, 24, 56, 64, 32, 56, 64, 96, 16, 32,
                               32, 32, 32, 48]


def test_easystart_coefficient_convention():
    """coeff_convention='easystart' gives the (jxy, jxy, jz) rows of
    EasyStart_DMRG.py:142-150 for chains and square lattices; the default
    stays Parameters.py's (jxy/2, jxy/2, jz)."""
    para = P.generate_parameters_dmrg('chain')
    para.update(l=12, jxy=1.0, jz=1.0, coeff_convention='easystart')
    para = P.make_consistent_parameter_dmrg(para)
    c = model.chain_couplings(12, 1.0, 1.0, convention='easystart')
    assert np.array_equal(para['index2'], c['index2'])
    assert np.array_equal(para['coeff2'], c['coeff2'])
    assert np.all(para['coeff2'] == 1.0) and para['coeff2'].shape == (33, 1)
    # the reference's saved runs are this Hamiltonian: its exact ground energy
    from oracle import ed
    e0, _ = ed.ground_state(12, para['op'], para['index1'], para['coeff1'], para['index2'],
                            para['coeff2'])
    assert abs(e0 - (-8.676296661854156)) < 1e-10
    sq = P.generate_parameters_dmrg('square')
    sq.update(jxy=0.6, jz=1.3, coeff_convention='easystart')
    sq = P.make_consistent_parameter_dmrg(sq)
    assert np.allclose(sq['coeff2'].reshape(-1, 3), [0.6, 0.6, 1.3])
    sq['coeff_convention'] = 'parameters'
    sq = P.make_consistent_parameter_dmrg(sq)
    assert np.allclose(sq['coeff2'].reshape(-1, 3), [0.3, 0.3, 1.3])
    bad = P.generate_parameters_dmrg('chain')
    bad['coeff_convention'] = 'halved'
    with pytest.raises(ValueError):
        P.make_consistent_parameter_dmrg(bad)


def test_initial_mps_is_the_reference_draw():
    """The engine's seeded initial MPS is the reference's draw order
    (TensorBasicModule.py:179-211): first tensor, last tensor, then the bulk,
    standard normals from one RandomState."""
    from oracle import tensor_ops as tb
    from tnalg_amd.engine import random_mps_host
    for n, chi, seed in ((12, 16, 0), (7, 5, 3), (2, 4, 9)):
        got = random_mps_host(n, 2, chi, seed)
        want, dims = tb.random_open_mps(n, 2, chi, np.random.RandomState(seed))
        assert len(got) == n
        for a, b in zip(got, want):
            assert a.shape == b.shape and np.array_equal(a, b)
        assert [a.shape[0] for a in got] + [1] == dims
    # a RandomState passed in continues its stream (dmrg_finite_size(seed=None)
    # draws from numpy.random's global state)
    r1, r2 = np.random.RandomState(5), np.random.RandomState(5)
    r1.randn(3)
    r2.randn(3)
    a = random_mps_host(6, 2, 4, r1)
    b, _ = tb.random_open_mps(6, 2, 4, r2)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_library_profile_argument_checks():
    """The sampled-timing entry points check their arguments before any
    device call (so this runs without a GPU)."""
    from tnalg_amd import _lib
    lib = _lib.load()
    assert lib.tnx_profile_begin(0, 16) == _lib.EINVAL
    assert lib.tnx_profile_begin(4, 0) == _lib.EINVAL
    assert lib.tnx_profile_end(None, 0) == _lib.EINVAL   # not armed
    assert b'not armed' in lib.tnx_last_error()
