"""Pins of the CPU oracle (no GPU): the reference's docstring known-answer
examples, its saved runs, exact diagonalisation and closed forms."""
import json
import os

import numpy as np
import pytest

from oracle import block_dmrg, dmrg_ref, ed, model
from oracle import tensor_ops as tb

GOLD = os.path.join(os.path.dirname(__file__), 'golden')


def _gold(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# ---- TensorBasicModule docstring KATs (TBM:497-504,536-543,579-586,644-652,690-698,760-770)
T3 = np.array([[[1, 2, 1], [2, 1, 2]], [[2, 0, 2], [1, 3, 1]], [[3, 1, 0], [2, 2, 1]]], float)
V3 = np.array([[1, 1, 1], [1, 2, 1], [2, 2, 1]], float)


def test_kat_bound_vec_left2right():
    assert np.array_equal(tb.bound_vec_operator_left2right(T3),
                          [[23, 14, 12], [14, 19, 9], [12, 9, 11]])
    assert np.array_equal(tb.bound_vec_operator_left2right(T3, v=V3),
                          [[81, 65, 58], [60, 64, 45], [46, 40, 33]])


def test_kat_bound_vec_right2left():
    assert np.array_equal(tb.bound_vec_operator_right2left(T3),
                          [[15, 11, 13], [11, 19, 15], [13, 15, 19]])
    assert np.array_equal(tb.bound_vec_operator_right2left(T3, v=V3),
                          [[55, 54, 57], [53, 58, 59], [48, 51, 50]])


def test_kat_absorb():
    T = np.array([[[1, 2], [2, 3]], [[3, 4], [4, 5]]], float)
    M = np.array([[1, 3], [2, 4]], float)
    assert np.array_equal(tb.absorb_matrix2tensor(T, M, 2),
                          [[[5, 11], [8, 18]], [[11, 25], [14, 32]]])
    ones = np.ones((2, 2, 2))
    Ms = [np.array([[1, 2], [2, 3]], float), np.array([[2, 3], [3, 4]], float),
          np.array([[3, 4], [4, 5]], float)]
    assert np.array_equal(tb.absorb_matrices2tensor(ones, Ms),
                          [[[105, 135], [147, 189]], [[175, 225], [245, 315]]])
    assert np.array_equal(tb.absorb_matrices2tensor(ones, Ms[:2]),
                          [[[15, 15], [21, 21]], [[25, 25], [35, 35]]])


def test_kat_transfer_matrix():
    want = [[5, 4, 5, 4, 5, 4, 5, 4, 5], [4, 6, 4, 5, 3, 5, 4, 6, 4], [7, 5, 2, 8, 4, 1, 7, 5, 2],
            [4, 5, 4, 6, 3, 6, 4, 5, 4], [5, 3, 5, 3, 9, 3, 5, 3, 5], [8, 4, 1, 6, 6, 3, 8, 4, 1],
            [7, 8, 7, 5, 4, 5, 2, 1, 2], [8, 6, 8, 4, 6, 4, 1, 3, 1], [13, 7, 2, 7, 5, 2, 2, 2, 1]]
    assert np.array_equal(tb.transfer_matrix_mps(T3), want)


def test_operator_transfer_against_brute_force():
    """bound_vec_* with an operator against an explicit contraction."""
    rng = np.random.RandomState(0)
    A = rng.randn(3, 2, 4)
    O = rng.randn(2, 2)
    v = rng.randn(3, 3)
    ref = np.zeros((4, 4))
    for x in range(3):
        for a in range(3):
            for s in range(2):
                for t in range(2):
                    ref += A[x, s, :][:, None] * v[x, a] * O[s, t] * A[a, t, :][None, :]
    assert np.allclose(tb.bound_vec_operator_left2right(A, O, v), ref)


def test_entanglement_entropy_formula():
    # maximally entangled pair: two equal Schmidt values -> ln 2
    assert abs(tb.entanglement_entropy(np.array([1.0, 1.0])) - np.log(2)) < 1e-15
    assert tb.entanglement_entropy(np.array([1.0, 0.0])) == 0.0


# ---- exact diagonalisation and closed forms
def test_ed_values_golden():
    g = _gold('ed_values.json')
    c = model.chain_couplings(12, 1.0, 1.0)
    e, _ = ed.ground_state(12, c['op'], c['index1'], c['coeff1'], c['index2'], c['coeff2'])
    assert abs(e - g['heisenberg_N12_obc']) < 1e-11
    # the survey's independent ED value of the same Hamiltonian
    assert abs(e - (-5.142090632840530)) < 1e-11


def test_ed_xx_free_fermions():
    c = model.chain_couplings(10, 1.0, 0.0)
    e, _ = ed.ground_state(10, c['op'], c['index1'], c['coeff1'], c['index2'], c['coeff2'])
    assert abs(e - ed.xx_chain_energy(10)) < 1e-12


def test_ed_majumdar_ghosh():
    g = _gold('ed_values.json')
    assert abs(g['j1j2_N10_mg'] - (-3.75)) < 1e-12


# ---- the reference restatement against the reference's saved runs
@pytest.mark.parametrize('chi,sweeps', [(16, 12)])
def test_oracle_reproduces_saved_run(chi, sweeps):
    g = _gold('reference_runs.json')['chi%d' % chi]
    c = model.chain_couplings(12, 1.0, 1.0, convention='easystart')
    p = dmrg_ref.make_para(c, chi, sweeps, tau=1e-3, eigs_tol=1e-10, break_tol=1e-13, dt_ob=2)
    ob, A, info, _ = dmrg_ref.dmrg_finite_size(p, seed=1)
    assert abs(float(ob['e_per_site'][0]) * 12 - g['E_total']) < 1e-8
    assert np.allclose(ob['eb_full'][:3, 0], g['eb_full_0_3'], atol=1e-7)
    assert np.allclose(ob['eb'][:2, 0], g['eb_0_2'], atol=1e-7)
    assert abs(A.ent[5, 0] - g['ent_5']) < 1e-7
    assert np.allclose(A.lm[5][:4], g['lm_5_0_4'], atol=1e-7)
    assert A.virtual_dim == g['bond_dims']


def test_oracle_chi_exact_matches_ed():
    c = model.chain_couplings(8, 1.0, 0.7, hz=0.2)
    p = dmrg_ref.make_para(c, 16, 6, eigs_tol=1e-12, break_tol=1e-14, dt_ob=1)
    ob, A, info, _ = dmrg_ref.dmrg_finite_size(p, seed=0)
    e0, psi = ed.ground_state(8, c['op'], c['index1'], c['coeff1'], c['index2'], c['coeff2'])
    assert abs(float(ob['e_per_site'][0]) * 8 - e0) < 1e-10
    v = A.full_state()
    assert abs(abs(v @ psi) / np.linalg.norm(v) - 1) < 1e-8


def test_block_oracle_agrees_with_reference_restatement():
    bonds = [(i, i + 1) for i in range(7)] + [(0, 3), (2, 6), (1, 7)]
    jj = [1.0] * 7 + [0.3, -0.4, 0.2]
    c = model.couplings_from_bonds(8, bonds, jj, jj, hz=0.1)
    pa = dmrg_ref.make_para(c, 8, 5, eigs_tol=1e-12, break_tol=1e-14, dt_ob=1)
    ob1, _, _, _ = dmrg_ref.dmrg_finite_size(pa, seed=2)
    ob2, _, _, _ = block_dmrg.dmrg_block(dict(pa), seed=2)
    assert abs(float(ob1['e_per_site'][0]) - float(ob2['e_per_site'][0])) < 1e-11
    assert np.allclose(ob1['eb_full'], ob2['eb_full'], atol=1e-8)


def test_dense_heff_brute_force():
    """P^T H P for the isometry of a canonical MPS equals the oracle's Heff."""
    c = model.chain_couplings(6, 1.0, 0.8, hz=0.3)
    p = dmrg_ref.make_para(c, 4, 1, eigs_tol=0.0)
    from oracle.mps_ref import RefMps
    A = RefMps(6, 2, 4, c['op'], np.random.RandomState(1), eig_way=0)
    A.correct_orthogonal_center(0)
    A.correct_orthogonal_center(3)
    A.update_all_effective_id()
    h, s = A.effective_hamiltonian_dense(3, p['index1'], p['index2'], p['coeff1'], p['coeff2'])
    # isometry: left block (sites 0..2), site 3 identity, right block (sites 4,5)
    def left_basis():
        v = A.mps[0].reshape(-1, A.mps[0].shape[2])
        for k in (1, 2):
            a = A.mps[k]
            v = (v @ a.reshape(a.shape[0], -1)).reshape(-1, a.shape[2])
        return v  # (2^3, Dl)

    def right_basis():
        v = A.mps[5].reshape(A.mps[5].shape[0], -1)
        a = A.mps[4]
        v = (a.reshape(-1, a.shape[2]) @ v).reshape(a.shape[0], -1)
        return v.T  # (2^2, Dr)

    PL, PR = left_basis(), right_basis()
    P = np.kron(np.kron(PL, np.eye(2)), PR)
    H = ed.hamiltonian(6, c['op'], c['index1'], c['coeff1'], c['index2'], c['coeff2']).toarray()
    assert np.allclose(P.T @ H @ P, h, atol=1e-12)
