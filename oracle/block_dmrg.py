"""Block-environment one-site DMRG (oracle; test use only).

Same effective Hamiltonian, gauge moves, ARPACK local solve and sweep order as
``dmrg_ref`` (``algorithms/DMRG_anyH.py:18-101``), but the per-term effective
operators of ``library/MPSClass.py:284-611`` are replaced by block
environments carried bond by bond: at bond ``b`` the left block keeps
``H_L`` (all terms inside sites ``< b``) plus one matrix per open operator
``(site i < b, op a)`` that still has a partner at ``>= b``; the right block
likewise.  Each local step then costs O(1) transfers.  It is the fast CPU
checker at sizes where the reference's O(N^3 chi^3)-per-sweep recompute is
too slow, and the second ("block") CPU baseline.  Pinned against ``dmrg_ref``
(same seed) and exact diagonalisation in the tests.
"""
import numpy as np
from scipy.sparse.linalg import LinearOperator, eigsh

from . import tensor_ops as tb
from .dmrg_ref import get_bond_energies, local_step_order
from .mps_ref import RefMps


def normalised_rows(para):
    """Two-body rows with ``i < j`` (operators swapped with the sites) and
    one-body rows, after the reference's coefficient cut ``|c| > eigs_tol``
    (``MPSClass.py:725-734``)."""
    tol = para['eigs_tol']
    ops = para['op']
    two, one = [], []
    for r in range(para['index2'].shape[0]):
        c = float(np.ravel(para['coeff2'])[r])
        if abs(c) <= tol:
            continue
        i, j, a, b = (int(x) for x in para['index2'][r])
        if i > j:
            i, j, a, b = j, i, b, a
        two.append((i, j, a, b, c))
    for r in range(para['index1'].shape[0]):
        c = float(np.ravel(para['coeff1'])[r])
        i, k = (int(x) for x in para['index1'][r])
        if abs(c) > tol and np.linalg.norm(ops[k].reshape(-1)) > tol:
            one.append((i, k, c))
    return two, one


class BlockDmrg:
    def __init__(self, para, rng):
        self.para = para
        self.n = para['l']
        self.ops = [np.real(o) for o in para['op']]
        self.two, self.one = normalised_rows(para)
        self.A = RefMps(self.n, para['d'], para['chi'], para['op'], rng, way='qr',
                        is_save_op=False)
        self.left = [None] * (self.n + 1)
        self.right = [None] * (self.n + 1)
        self.n_matvec = 0

    # open operators at bond b: (i, a) with i < b <= j, and (j, b') likewise
    def left_opens(self, b):
        return sorted({(i, a) for (i, j, a, _, _) in self.two if i < b <= j})

    def right_opens(self, b):
        return sorted({(j, bb) for (i, j, _, bb, _) in self.two if i < b <= j})

    def site_field(self, p):
        h = np.zeros((self.para['d'],) * 2)
        for (i, k, c) in self.one:
            if i == p:
                h = h + c * self.ops[k]
        return h

    def grow_left(self, p):
        """Left block at bond p+1 from bond p through the left-canonical A[p]."""
        a = self.A.mps[p]
        env = self.left[p]
        new = {}
        h = tb.bound_vec_operator_left2right(a, self.site_field(p), None)
        if p > 0:
            h = h + tb.bound_vec_operator_left2right(a, None, env['H'])
        for (i, j, oa, ob, c) in self.two:
            if j == p:
                h = h + c * tb.bound_vec_operator_left2right(a, self.ops[ob], env[(i, oa)])
        new['H'] = h
        for (i, oa) in self.left_opens(p + 1):
            if i == p:
                new[(i, oa)] = tb.bound_vec_operator_left2right(a, self.ops[oa], None)
            else:
                new[(i, oa)] = tb.bound_vec_operator_left2right(a, None, env[(i, oa)])
        self.left[p + 1] = new

    def grow_right(self, p):
        """Right block at bond p from bond p+1 through the right-canonical A[p]."""
        a = self.A.mps[p]
        env = self.right[p + 1]
        new = {}
        h = tb.bound_vec_operator_right2left(a, self.site_field(p), None)
        if p < self.n - 1:
            h = h + tb.bound_vec_operator_right2left(a, None, env['H'])
        for (i, j, oa, ob, c) in self.two:
            if i == p:
                h = h + c * tb.bound_vec_operator_right2left(a, self.ops[oa], env[(j, ob)])
        new['H'] = h
        for (j, ob) in self.right_opens(p):
            if j == p:
                new[(j, ob)] = tb.bound_vec_operator_right2left(a, self.ops[ob], None)
            else:
                new[(j, ob)] = tb.bound_vec_operator_right2left(a, None, env[(j, ob)])
        self.right[p] = new

    def heff(self, p, x):
        s = self.A.mps[p].shape
        t = x.reshape(s)
        le, re = self.left[p], self.right[p + 1]
        out = np.einsum('st,atb->asb', self.site_field(p), t)
        if p > 0:
            out += np.tensordot(le['H'], t, axes=([1], [0]))
        if p < self.n - 1:
            out += np.tensordot(t, re['H'], axes=([2], [1]))
        for (i, j, oa, ob, c) in self.two:
            if j == p:
                y = np.tensordot(le[(i, oa)], t, axes=([1], [0]))
                out += c * np.einsum('st,atb->asb', self.ops[ob], y)
            elif i == p:
                y = np.tensordot(t, re[(j, ob)], axes=([2], [1]))
                out += c * np.einsum('st,atb->asb', self.ops[oa], y)
            elif i < p < j:
                y = np.tensordot(le[(i, oa)], t, axes=([1], [0]))
                out += c * np.tensordot(y, re[(j, ob)], axes=([2], [1]))
        return out.reshape(-1)

    def move_center(self, p):
        c = self.A.center
        while c < p:
            self.A.correct_orthogonal_center(c + 1)
            self.grow_left(c)
            c += 1
        while c > p:
            self.A.correct_orthogonal_center(c - 1)
            self.grow_right(c)
            c -= 1

    def solve_site(self, p):
        para = self.para
        self.move_center(p)
        s = self.A.mps[p].shape
        dim = int(np.prod(s))
        tau = para['tau']

        def mv(x):
            self.n_matvec += 1
            x = np.asarray(x).reshape(-1)
            return x - tau * self.heff(p, x)

        if dim == 1:
            return
        op = LinearOperator((dim, dim), matvec=mv, dtype=float)
        vec = eigsh(op, k=1, which='LM', v0=self.A.mps[p].reshape(-1), tol=para['eigs_tol'])[1]
        self.A.mps[p] = np.real(vec).reshape(s)

    def initial_environments(self):
        """After ``correct_orthogonal_center(ob_position)``: grow all blocks."""
        c = self.A.center
        self.left[0] = {'H': np.zeros((1, 1))}
        self.right[self.n] = {'H': np.zeros((1, 1))}
        for p in range(0, c):
            self.grow_left(p)
        for p in range(self.n - 1, c, -1):
            self.grow_right(p)


def dmrg_block(para, seed=0, rng=None, max_steps=None):
    """``dmrg_finite_size`` on block environments; outputs as ``dmrg_ref``."""
    if rng is None:
        rng = np.random.RandomState(seed)
    eng = BlockDmrg(para, rng)
    A = eng.A
    A.correct_orthogonal_center(para['ob_position'])
    eng.initial_environments()
    e0 = 0.0
    info = {'convergence': 1, 'sweeps': 0, 'energies': []}
    ob = {}
    steps = 0
    order = local_step_order(para['l'], para['ob_position'])
    for t in range(para['sweep_time']):
        if_ob = ((t + 1) % para['dt_ob'] == 0) or t == para['sweep_time'] - 1
        for n in order:
            eng.solve_site(n)
            steps += 1
            if max_steps is not None and steps >= max_steps:
                info['steps'] = steps
                return None, A, info, para
        info['sweeps'] = t + 1
        if if_ob:
            ob['eb_full'] = A.observe_bond_energy(para['index2'], para['coeff2'])
            ob['mx'] = A.observe_magnetization(1)
            ob['mz'] = A.observe_magnetization(3)
            ob['e_per_site'] = (sum(ob['eb_full']) - para['hx'] * sum(ob['mx'])
                                - para['hz'] * sum(ob['mz'])) / A.length
            info['energies'].append(float(np.ravel(ob['e_per_site'])[0]) * A.length)
            info['convergence'] = abs(ob['e_per_site'] - e0)
            if info['convergence'] < para['break_tol']:
                break
            e0 = ob['e_per_site']
    ob['eb'] = get_bond_energies(ob['eb_full'], para['positions_h2'], para['index2'])
    A.calculate_entanglement_spectrum()
    A.calculate_entanglement_entropy()
    ob['corr_x'] = A.observe_correlators_from_middle(1, 1)
    ob['corr_z'] = A.observe_correlators_from_middle(3, 3)
    info['steps'] = steps
    info['n_matvec'] = eng.n_matvec
    return ob, A, info, para
