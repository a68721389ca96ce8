"""The TEBD restatement in oracle/ pinned by what the mathematics fixes: one
sweep without truncation equals the dense product of the gates on the full
state, and the imaginary-time fixed point approaches the ground energy of
exact diagonalisation as tau shrinks (the sequential product's energy error
is second order in tau)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..'))

from oracle import ed, tebd_ref  # noqa: E402
from oracle.model import operator_list  # noqa: E402


def tebd_para(l, chi, jxy=1.0, jz=0.7, hx=0.3, hz=0.2, **kw):
    para = {'l': l, 'chi': chi, 'jxy': jxy, 'jz': jz, 'hx': hx, 'hz': hz, 'spin': 'half',
            'tau0': 0.1, 'dtau': 0.1, 'taut': 3, 'dt_ob': 10, 'iterate_time': 3000,
            'save_mode': 'final', 'if_break': True, 'break_tol': 1e-10,
            'positions_h2': np.array([(p, p + 1) for p in range(l - 1)])}
    para.update(kw)
    return para


def tebd_hamiltonian(l, jxy, jz, hx, hz):
    """The sum of the TEBD bond Hamiltonians: every bond carries hx/2, hz/2
    on both its sites, so the chain ends see half the bulk field."""
    ops = operator_list(hx, hz)
    index2, coeff2, index1, coeff1 = [], [], [], []
    for p in range(l - 1):
        for a, b, c in ((4, 5, jxy / 2), (5, 4, jxy / 2), (3, 3, jz)):
            index2.append((p, p + 1, a, b))
            coeff2.append(c)
    for i in range(l):
        index1.append((i, 6))
        coeff1.append(1.0 if 0 < i < l - 1 else 0.5)
    return ops, np.array(index1), np.array(coeff1), np.array(index2), np.array(coeff2)


def test_sweep_without_truncation_equals_dense_gates():
    l, chi = 6, 8          # chi = 2^(l/2): nothing is truncated
    para = tebd_para(l, chi)
    rng = np.random.RandomState(3)
    mps = tebd_ref.RefMps(l, 2, chi, operator_list(0.3, 0.2), rng)
    tebd_ref.move_center(mps, 0)
    psi0 = mps.full_state()
    psi0 = psi0 / np.linalg.norm(psi0)
    h = tebd_ref.bond_hamiltonian(para['jxy'], para['jz'], para['hx'] / 2, para['hz'] / 2)
    from scipy.linalg import expm
    gate = expm(-0.1 * h)
    bonds = list(range(l - 1))
    for sweep in range(3):
        tebd_ref.gate_sweep(mps, bonds, gate, chi)
        psi0 = tebd_ref.dense_gate_sweep(psi0, l, bonds, gate)
        psi = mps.full_state()
        assert abs(np.linalg.norm(psi) - 1) < 1e-12
        assert abs(abs(psi @ psi0) - 1) < 1e-12, sweep


def test_truncated_sweep_keeps_the_largest_schmidt_values():
    """With truncation, one gate step keeps the chi largest singular values
    of G theta: the kept state's overlap with the untruncated one is the
    kept weight."""
    l, chi = 8, 4
    rng = np.random.RandomState(5)
    mps = tebd_ref.RefMps(l, 2, 16, operator_list(), rng)
    tebd_ref.move_center(mps, 3)
    from scipy.linalg import expm
    gate = expm(-0.2 * tebd_ref.bond_hamiltonian(1.0, 0.7, 0.15, 0.1))
    a, b = mps.mps[3], mps.mps[4]
    theta = np.einsum('st,atb->asb', gate,
                      np.tensordot(a, b, ([2], [0])).reshape(a.shape[0], 4, b.shape[2]))
    s = np.linalg.svd(theta.reshape(a.shape[0] * 2, -1), compute_uv=False)
    before = mps.full_state()
    full = tebd_ref.dense_gate_sweep(before, l, [3], gate)
    tebd_ref.gate_pair(mps, 3, gate, chi)
    after = mps.full_state()
    kept = np.sqrt(np.sum(s[:chi] ** 2) / np.sum(s ** 2))
    assert abs(abs(after @ full) - kept) < 1e-12


def test_fixed_point_energy_approaches_ed():
    l, chi = 8, 16
    para = tebd_para(l, chi)
    mps, ob = tebd_ref.tebd_standard(para, seed=2)
    e0, _ = ed.ground_state(l, *tebd_hamiltonian(l, 1.0, 0.7, 0.3, 0.2))
    # the TEBD energy: bond terms plus the fields with the edge weights
    w = np.array([0.5] + [1.0] * (l - 2) + [0.5])
    mx = mps.observe_magnetization(1).reshape(-1)
    mz = mps.observe_magnetization(3).reshape(-1)
    pos = np.array([(p, p + 1) for p in range(l - 1)])
    e = float(np.sum(mps.observe_bond_energy_from_jxyz(pos, 1.0, 1.0, 0.7))) \
        - 0.3 * float(w @ mx) - 0.2 * float(w @ mz)
    assert abs(e - e0) < 2e-5 * abs(e0), (e, e0)
    assert len(ob['e_site']) >= 1 and len(ob['lm'][-1]) == l - 1
