"""Infinite-size DMRG of the reference, restated in numpy (oracle; test use only).

Restates ``algorithms/DMRG_anyH.py:106-188`` (``dmrg_infinite_size``) with
the infinite MPS of ``library/MPSClass.py:1550-1891`` (``MpsInfinite``,
form 'center_ort', two-site centre) and the cell tensor of
``library/HamiltonianModule.py:177-200`` (``hamiltonian2cell_tensor``), for
both ``dmrg_type`` values the driver runs:

* 'mpo': the centre (chi, D, D, chi) of two cells of D = d^2 spins is the
  dominant eigenvector (``eigsh`` of ``MPSClass.py:1861-1870`` is scipy's
  ``eigsh``, which='LM') of the symmetrised
  h = env_L . T . T . env_R (``MPSClass.py:1816-1845``); the environments
  are transfer-matrix fixed-point iterates through the cell tensor,
  normalised and symmetrised (``MPSClass.py:1652-1664``);
* 'white': the centre (chi, d, d, chi) is the dominant eigenvector of
  I - tau H with H built from the on-site bath and the effective operators
  (``MPSClass.py:1709-1727``), updated by ``update_bath_onsite`` and
  ``update_effective_ops`` (``MPSClass.py:1737-1769``).

The split of the centre is the SVD of ``update_ort_tensor_mps`` for
n_site = 2 (``MPSClass.py:1680-1690``).  Pinned by tests: the cell tensors
of a two-cell ring against the dense product of the shift gates, the
dominant eigenvector against a dense symmetric eigensolve, and the fixed
points against the Bethe-ansatz energy of the Heisenberg chain and the
exact energy of the critical transverse-field Ising chain.
"""
import numpy as np
from scipy.linalg import expm
from scipy.sparse.linalg import eigsh

from .model import spin_half_operators


def bond_hamiltonian(jxy, jz, hx, hz):
    """``hamiltonian_heisenberg('half', jxy, jxy, jz, hx, hz)``
    (``library/HamiltonianModule.py:141-147``), 4 x 4."""
    o = spin_half_operators()
    h = jxy * np.kron(o['sx'], o['sx']) + jxy * np.kron(o['sy'], o['sy']).real + \
        jz * np.kron(o['sz'], o['sz'])
    h = h - hx * (np.kron(o['id'], o['sx']) + np.kron(o['sx'], o['id']))
    h = h - hz * (np.kron(o['id'], o['sz']) + np.kron(o['sz'], o['id']))
    return np.real(h)


def cell_tensor(h, tau, way='shift'):
    """``hamiltonian2cell_tensor`` (``HamiltonianModule.py:177-200``): the
    gate g = I - tau h ('shift') or expm(-tau h), split by an SVD over
    (s1 s1') | (s2 s2') into vl (s1, s1', k) and vr (k, s2, s2') with the
    square roots of the singular values on both, and
    T[k, (x y), (c e), q] = sum_ab vr[k, x, a] g[a, b, c, e] vl[y, b, q]."""
    dd = h.shape[0]
    d = int(round(dd ** 0.5))
    g = np.eye(dd) - tau * h if way == 'shift' else expm(-tau * h)
    g4 = g.reshape(d, d, d, d)
    u, s, vt = np.linalg.svd(g4.transpose(0, 2, 1, 3).reshape(dd, dd))
    sq = np.sqrt(s)
    vl = (u * sq[None, :]).reshape(d, d, dd)
    vr = (sq[:, None] * vt).reshape(dd, d, d)
    t = np.einsum('abce,kxa,ybq->kxyceq', g4, vr, vl)
    return t.reshape(dd, dd, dd, dd)


def operators_list():
    """``[id, sx, sy, sz, su, sd]`` of ``MpsInfinite.__init__``
    (``MPSClass.py:1597-1599``); sy is complex."""
    o = spin_half_operators()
    return [o['id'], o['sx'], o['sy'], o['sz'], o['su'], o['sd']]


def hamiltonian_index(jxy, jz, hx, hz):
    """``hamiltonian_indexes('heisenberg', (jxy, jz, -hx/2, -hz/2))``
    (``HamiltonianModule.py:97-124``, called by ``Parameters.py:455-467``):
    rows (op1, op2, coefficient)."""
    return np.array([[4, 5, jxy], [5, 4, jxy], [3, 3, jz], [0, 1, -hx / 2], [1, 0, -hx / 2],
                     [0, 3, -hz / 2], [3, 0, -hz / 2]], dtype=float)


def bvol(tensor, op=None, v=None):
    """``bound_vec_operator_left2right`` (``TensorBasicModule.py:632-675``):
    v'[b, c] = sum A*[x, s, b] v[x, a] op[s, t] A[a, t, c]."""
    ket = tensor if op is None else np.einsum('st,atc->asc', op, tensor)
    if v is not None:
        ket = np.tensordot(v, ket, axes=([1], [0]))
    return np.tensordot(tensor.conj(), ket, axes=([0, 1], [0, 1]))


class RefInfinite:
    """``MpsInfinite`` with form 'center_ort' and n_site = 2
    (``MPSClass.py:1550-1891``).  ``d`` is the dimension of one index of the
    centre (D = d_spin^2 for 'mpo', d_spin for 'white'); ``D`` the MPO bond
    of the environments."""

    def __init__(self, d, chi, D, dmrg_type, hamilt_index, is_symme_env, rng):
        self.d, self.chi, self.D = d, chi, D
        self.n_site = 2
        self.dmrg_type = dmrg_type
        self.is_symme_env = is_symme_env
        self.hamilt_index = hamilt_index
        self.operators = operators_list()
        self.mps = [None, None, None]
        self.lm = [np.zeros(0)] * 3
        # initialize_imps (MPSClass.py:1614-1627)
        c = rng.randn(chi, d, d, chi)
        self.mps[1] = c / np.linalg.norm(c.reshape(-1))
        self.update_ort_tensor_mps()
        self.env = [np.ones((chi, D, chi)), np.ones((chi, D, chi))]
        if is_symme_env:
            self.env[1] = self.env[0]
        self.bath_op_onsite = np.eye(chi)
        self.effective_ops = [np.eye(chi) for _ in range(len(self.operators))]
        self.op_index = set(range(len(self.operators)))
        self.rho = None

    def update_ort_tensor_mps(self):
        """n_site = 2 (``MPSClass.py:1680-1690``): full SVD of the centre,
        chi kept; mps[2] is stored (right bond, s, new bond)."""
        s = self.mps[1].shape
        u, lm, v = np.linalg.svd(self.mps[1].reshape(s[0] * s[1], s[2] * s[3]))
        dc = min(self.chi, s[0] * s[1])
        self.mps[0] = u[:, :dc].reshape(s[0], s[1], dc)
        self.mps[2] = v[:dc, :].reshape(dc, s[2], s[3]).transpose(2, 1, 0)
        self.lm[0] = lm[:dc]

    def update_left_env(self, t):
        """``MPSClass.py:1652-1657``."""
        e = np.einsum('pxa,mxcq,rcb,pmr->aqb', self.mps[0].conj(), t, self.mps[0], self.env[0])
        e = e / np.linalg.norm(e.reshape(-1))
        self.env[0] = (e + e.transpose(2, 1, 0)) / 2

    def update_right_env(self, t):
        """``MPSClass.py:1659-1664``."""
        e = np.einsum('pxa,nxck,rcb,pkr->anb', self.mps[2].conj(), t, self.mps[2], self.env[1])
        e = e / np.linalg.norm(e.reshape(-1))
        self.env[1] = (e + e.transpose(2, 1, 0)) / 2

    def effective_hamiltonian(self, t):
        """n_site = 2 (``MPSClass.py:1831-1845``), symmetrised."""
        er = self.env[0] if self.is_symme_env else self.env[1]
        h = np.einsum('amA,mxcn,nyek,bkB->axybAceB', self.env[0], t, t, er)
        s = h.shape
        h = h.reshape(s[0] * s[1] * s[2] * s[3], -1)
        return (h + h.T) / 2

    def effective_hamilt_from_op(self):
        """``MPSClass.py:1709-1727`` (dense H of the 'white' centre)."""
        chi, d = self.chi, self.d
        id0, id1 = np.eye(d), np.eye(chi)
        n = chi * chi * d * d
        h = np.zeros((n, n))
        for i1, i2, j in self.hamilt_index:
            op1, op2 = self.operators[int(i1)], self.operators[int(i2)]
            h = h + j * np.kron(id1, np.kron(np.kron(op1, op2), id1))
        h = h + np.kron(np.kron(np.kron(self.bath_op_onsite, id0), id0), id1)
        h = h + np.kron(np.kron(np.kron(id1, id0), id0), self.bath_op_onsite)
        for i1, i2, j in self.hamilt_index:
            op1, op2 = self.effective_ops[int(i1)], self.operators[int(i2)]
            h = h + j * np.kron(np.kron(np.kron(op1, op2), id0), id1)
            h = h + j * np.kron(np.kron(np.kron(id1, id0), op2), op1)
        return np.real(h)

    def update_central_tensor(self, t, tau):
        """``MPSClass.py:1861-1870``: eigsh (which='LM', tol=0) from the
        current centre."""
        s = self.mps[1].shape
        if self.dmrg_type == 'white':
            h = np.eye(self.mps[1].size) - tau * self.effective_hamilt_from_op()
        else:
            h = self.effective_hamiltonian(t)
        v = eigsh(h, 1, v0=self.mps[1].reshape(-1, 1))[1]
        self.mps[1] = np.real(v).reshape(s)

    def update_effective_ops(self):
        """``MPSClass.py:1737-1752`` (every operator)."""
        for n in sorted(self.op_index):
            self.effective_ops[n] = bvol(self.mps[0], self.operators[n])

    def update_bath_onsite(self):
        """``MPSClass.py:1754-1763``."""
        op = bvol(self.mps[0], np.eye(self.d), self.bath_op_onsite)
        for i1, i2, j in self.hamilt_index:
            op = op + j * bvol(self.mps[0], self.operators[int(i2)], self.effective_ops[int(i1)])
        self.bath_op_onsite = (op + op.conj().T) / 2

    def rho_from_central_tensor(self):
        """n_site = 2 (``MPSClass.py:1872-1886``)."""
        if self.dmrg_type == 'mpo':
            d0 = 2
            tmp = self.mps[1].reshape(self.chi, d0, d0, d0, d0, self.chi)
            self.rho = []
            for n in range(1, 4):
                con = list(range(0, n)) + list(range(n + 2, 6))
                self.rho.append(np.tensordot(tmp, tmp, (con, con)).reshape(self.d, self.d))
        else:
            self.rho = np.tensordot(self.mps[1].conj(), self.mps[1], ([0, 3], [0, 3])).reshape(
                self.d ** 2, self.d ** 2)

    def observe_energy(self, h):
        """``MPSClass.py:1888-1899``."""
        if isinstance(self.rho, list):
            return np.array([np.trace(r.dot(h)) for r in self.rho])
        return np.trace(self.rho.dot(h))


def dmrg_infinite_size(para, seed=0, rng=None, hamilt=None, record=None):
    """``DMRG_anyH.py:106-188`` for n_site = 2.  ``record``: optional list
    that receives (t, e1) at every observation.  Returns (A, ob)."""
    rng = np.random.RandomState(seed) if rng is None else rng
    if hamilt is None:
        hamilt = bond_hamiltonian(para['jxy'], para['jz'], para['hx'] / 2, para['hz'] / 2)
    if para['n_site'] != 2:
        raise ValueError('the reference runs n_site = 2 only')
    dspin = para['d']
    d = dspin ** 2 if para['dmrg_type'] == 'mpo' else dspin
    A = RefInfinite(d, para['chi'], dspin ** 2, para['dmrg_type'],
                    np.asarray(para['hamilt_index'], dtype=float), para['is_symme_env'], rng)
    t_cell = cell_tensor(hamilt, para['tau']) if A.dmrg_type == 'mpo' else None
    e0, e1 = np.zeros((1, 3)), np.ones((1, 3))

    def env_step():
        if A.is_symme_env:
            A.update_ort_tensor_mps()
            if A.dmrg_type == 'white':
                A.update_bath_onsite()
                A.update_effective_ops()
            else:
                A.update_left_env(t_cell)
        else:
            A.update_ort_tensor_mps()
            A.update_left_env(t_cell)
            A.update_right_env(t_cell)

    env_step()
    for t in range(para['sweep_time']):
        A.update_central_tensor(t_cell, para['tau'])
        if t % para['dt_ob'] == 0:
            A.rho_from_central_tensor()
            e1 = A.observe_energy(hamilt)
            if record is not None:
                record.append((t, np.array(e1, copy=True)))
            de = np.sum(abs(e0 - e1)) / A.n_site
            if de > para['break_tol']:
                e0 = e1
            else:
                break
        env_step()
    return A, {'eb': e1}
