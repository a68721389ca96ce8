"""The reference's open-boundary MPS and one-site local solver (oracle; test use only).

Restates ``library/MPSClass.py::MpsOpenBoundaryClass`` for the path driven by
``dmrg_finite_size``: QR gauge moves, the per-coupling-term effective-operator
cache (``is_save_op=True``, the default of ``library/Parameters.py:25``),
the ``opt_env`` aggregation of ``eig_way=1`` (``library/Parameters.py:22``),
the matrix-free ``psi -> psi - tau*Heff psi`` handed to ARPACK
``eigsh(k=1, which='LM')``, and the observables/entanglement used at the end.
The dense ``eig_way=0`` branch (``MPSClass.py:613-662``) is also provided; the
reference's saved runs used it.

The cache reproduces the reference's cost profile on purpose (one-body
operators extend the nearest cached bond; two-body operators are
re-propagated from their second site whenever their bond entry is missing,
``MPSClass.py:585-611``), because this class is also the timed CPU baseline.
"""
import numpy as np
from scipy.sparse.linalg import LinearOperator, eigsh

from . import tensor_ops as tb


class RefMps:
    """``MpsOpenBoundaryClass`` (``library/MPSClass.py:55-108``), real arithmetic."""

    def __init__(self, length, d, chi, operators, rng, way='qr', is_save_op=True, eig_way=1,
                 mps=None):
        self.length = length
        self.phys_dim = d
        self.decomp_way = way
        self.operators = [np.asarray(o) for o in operators]
        self.orthogonality = np.zeros((length, 1))
        self.center = -1
        self.lm = [np.zeros(0) for _ in range(length - 1)]
        self.ent = np.zeros((length - 1, 1))
        if mps is None:
            self.mps, self.virtual_dim = tb.random_open_mps(length, d, chi, rng)
        else:
            self.mps = [np.array(m, dtype=float) for m in mps]
            self.virtual_dim = [m.shape[0] for m in self.mps] + [1]
        self.is_save_op = is_save_op
        self.eig_way = eig_way
        self.effect_s = {}     # (sn, p0, p1) -> one-body effective operator at bond p1
        self.effect_ss = {}    # (sn, snn, p0, q0, p1) -> two-body effective operator at bond p1
        self.effective_id = {}
        self.opt_env = {}
        self.n_matvec = 0
        self.n_transfer = 0

    # ------------------------------------------------------------------ gauge
    def orthogonalize_mps(self, l0, l1):
        """``MPSClass.py:186-247`` (no truncation, no normalisation)."""
        if l0 < l1:
            for n in range(l0, l1):
                self.mps[n], mat, self.virtual_dim[n + 1], lm = \
                    tb.left2right_decompose_tensor(self.mps[n], self.decomp_way)
                self.mps[n + 1] = tb.absorb_matrix2tensor(self.mps[n + 1], mat, 0)
                if lm.size > 0 and self.center > -1:
                    self.lm[n] = lm.copy()
            self.orthogonality[l0:l1] = -1
            self.orthogonality[l1] = 0
        elif l0 > l1:
            for n in range(l0, l1, -1):
                self.mps[n], mat, self.virtual_dim[n], lm = \
                    tb.right2left_decompose_tensor(self.mps[n], self.decomp_way)
                self.mps[n - 1] = tb.absorb_matrix2tensor(self.mps[n - 1], mat, 2)
                if lm.size > 0 and self.center > -1:
                    self.lm[n - 1] = lm.copy()
            self.orthogonality[l0:l1:-1] = 1
            self.orthogonality[l1] = 0

    def correct_orthogonal_center(self, p):
        """``MPSClass.py:250-279``."""
        if self.center < -0.5:
            self.orthogonalize_mps(0, p)
            self.orthogonalize_mps(self.length - 1, p)
        elif self.center != p:
            self.orthogonalize_mps(self.center, p)
        self.center = p

    # ------------------------------------------------------------ transfers
    def _l2r(self, n, op=None, v=None):
        self.n_transfer += 1
        return tb.bound_vec_operator_left2right(self.mps[n], op, v)

    def _r2l(self, n, op=None, v=None):
        self.n_transfer += 1
        return tb.bound_vec_operator_right2left(self.mps[n], op, v)

    def contract_v_l0_to_l1(self, l0, l1, v=None):
        """``MPSClass.py:532-539``."""
        if l0 < l1:
            for n in range(l0, l1):
                v = self._l2r(n, None, v)
        elif l0 > l1:
            for n in range(l0, l1, -1):
                v = self._r2l(n, None, v)
        return v

    # -------------------------------------------------------- operator cache
    def _extend_one_body(self, l0, l1, v, sn, p0, update):
        """``update_effect_op_l0_to_l1`` (``MPSClass.py:555-570``)."""
        if l0 < l1:
            for n in range(l0, l1):
                v = self._l2r(n, None, v)
                if update:
                    self.effect_s[(sn, p0, n + 1)] = v
        elif l0 > l1:
            for n in range(l0, l1, -1):
                v = self._r2l(n, None, v)
                if update:
                    self.effect_s[(sn, p0, n)] = v
        return v

    def update_all_effective_id(self):
        """``MPSClass.py:572-583``."""
        self.effective_id = {}
        c = self.center
        v = None
        for n in range(c, self.length):
            v = self._l2r(n, None, v)
            self.effective_id[(c, n + 1)] = v
        v = None
        for n in range(c, 0, -1):
            v = self._r2l(n, None, v)
            self.effective_id[(c, n)] = v

    def get_one_body(self, sn, p0, p1, update=True):
        """``get_effective_operators_one_body`` + ``find_nearest_key_one_body``
        (``MPSClass.py:308-369``): operator ``sn`` at site ``p0`` carried to bond
        ``p1`` (left-type when ``p0 < p1``, right-type otherwise)."""
        key = (sn, p0, p1)
        if key in self.effect_s:
            return self.effect_s[key]
        op = self.operators[sn]
        cands = [k[2] for k in self.effect_s if k[0] == sn and k[1] == p0]
        if p0 < p1:
            cands = [b for b in cands if p0 < b < p1]
            if not cands:
                v = self.effective_id[(self.center, p0)] if self.center < p0 else None
                v = self._l2r(p0, op, v)
                if update:
                    self.effect_s[(sn, p0, p0 + 1)] = v
                return self._extend_one_body(p0 + 1, p1, v, sn, p0, update)
            b = max(cands)
            return self._extend_one_body(b, p1, self.effect_s[(sn, p0, b)], sn, p0, update)
        cands = [b for b in cands if p1 < b <= p0]
        if not cands:
            v = self.effective_id[(self.center, p0 + 1)] if self.center > p0 else None
            v = self._r2l(p0, op, v)
            if update:
                self.effect_s[(sn, p0, p0)] = v
            return self._extend_one_body(p0 - 1, p1 - 1, v, sn, p0, update)
        b = min(cands)
        return self._extend_one_body(b - 1, p1 - 1, self.effect_s[(sn, p0, b)], sn, p0, update)

    def get_two_body(self, sn, snn, p0, q0, p1, update=True):
        """``get_effective_operator_two_body`` + ``update_effect_from_op1_to_op2``
        (``MPSClass.py:371-385,585-611``); requires ``p0 < q0 < p1`` or
        ``p1 <= p0 < q0``.  A missing bond entry is re-propagated from the
        operator sites, as the reference does."""
        if p0 > q0:
            p0, q0, sn, snn = q0, p0, snn, sn
        key = (sn, snn, p0, q0, p1)
        if key in self.effect_ss:
            return self.effect_ss[key]
        if q0 < p1:
            v = self.get_one_body(sn, p0, q0, update)
            v = self._l2r(q0, self.operators[snn], v)
            if update:
                self.effect_ss[(sn, snn, p0, q0, q0 + 1)] = v
            for n in range(q0 + 1, p1):
                v = self._l2r(n, None, v)
                if update:
                    self.effect_ss[(sn, snn, p0, q0, n + 1)] = v
            return v
        if p1 <= p0:
            v = self.get_one_body(snn, q0, p0 + 1, update)
            v = self._r2l(p0, self.operators[sn], v)
            if update:
                self.effect_ss[(sn, snn, p0, q0, p0)] = v
            for n in range(p0 - 1, p1 - 1, -1):
                v = self._r2l(n, None, v)
                if update:
                    self.effect_ss[(sn, snn, p0, q0, n)] = v
            return v
        raise RuntimeError('two-body operator requested across its own sites')

    def del_bad_effective_operators(self, p):
        """``MPSClass.py:387-406``: drop entries that depend on tensor ``p``."""
        self.effect_s = {k: v for k, v in self.effect_s.items()
                         if (k[1] < p and k[2] <= p) or (k[1] > p and k[2] > p)}
        self.effect_ss = {k: v for k, v in self.effect_ss.items()
                          if (k[3] < p and k[4] <= p) or (k[2] > p and k[4] > p)}

    # ------------------------------------------------------- environments
    def environment_s1(self, p, sn, pos):
        """``MPSClass.py:489-522`` -> ``(vL, vM, vR)``; ``None`` is identity."""
        if p < pos:
            return None, None, self._one_body_env(sn, pos, p + 1)
        if p > pos:
            return self._one_body_env(sn, pos, p), None, None
        return None, self.operators[sn], None

    def _one_body_env(self, sn, pos, bond):
        if self.is_save_op:
            return self.get_one_body(sn, pos, bond)
        if pos < bond:
            v = self._l2r(pos, self.operators[sn])
            return self.contract_v_l0_to_l1(pos + 1, bond, v)
        v = self._r2l(pos, self.operators[sn])
        return self.contract_v_l0_to_l1(pos - 1, bond - 1, v)

    def environment_s1_s2(self, p, sns, positions):
        """``MPSClass.py:417-479`` -> ``(vL, vM, vR)``; ``None`` is identity.

        Like the reference, the operator indices are NOT re-ordered when the
        positions are (``MPSClass.py:426-428``); all generators emit ``i < j``.
        """
        ops = [self.operators[sns[0]], self.operators[sns[1]]]
        positions = list(positions)
        if positions[0] > positions[1]:
            positions = positions[::-1]
            ops = ops[::-1]
        p0, q0 = positions
        if p < p0:
            if self.is_save_op:
                vr = self.get_two_body(sns[0], sns[1], p0, q0, p + 1)
            else:
                vr = self._r2l(q0, ops[1])
                vr = self.contract_v_l0_to_l1(q0 - 1, p0, vr)
                vr = self._r2l(p0, ops[0], vr)
                vr = self.contract_v_l0_to_l1(p0 - 1, p, vr)
            return None, None, vr
        if p > q0:
            if self.is_save_op:
                vl = self.get_two_body(sns[0], sns[1], p0, q0, p)
            else:
                vl = self._l2r(p0, ops[0])
                vl = self.contract_v_l0_to_l1(p0 + 1, q0, vl)
                vl = self._l2r(q0, ops[1], vl)
                vl = self.contract_v_l0_to_l1(q0 + 1, p, vl)
            return vl, None, None
        if p == p0:
            return None, ops[0], self._one_body_env(sns[1], q0, p + 1)
        if p == q0:
            return self._one_body_env(sns[0], p0, p), ops[1], None
        return self._one_body_env(sns[0], p0, p), None, self._one_body_env(sns[1], q0, p + 1)

    def all_environments_optimized(self, p, index1, index2, coeff1, coeff2, tol):
        """``MPSClass.py:715-734`` + ``classify_and_update_env`` (766-815)."""
        self.opt_env = {}
        env = self.opt_env

        def acc(key, mat):
            env[key] = env[key] + mat if key in env else mat.copy()

        for n in range(index1.shape[0]):
            c = float(np.ravel(coeff1)[n])
            sn = int(index1[n, 1])
            if abs(c) > tol and np.linalg.norm(self.operators[sn].reshape(-1)) > tol:
                pos = int(index1[n, 0])
                vl, vm, vr = self.environment_s1(p, sn, pos)
                if pos < p:
                    acc(('L',), vl * c)
                elif pos == p:
                    acc(('M', sn), vm * c)
                else:
                    acc(('R',), vr * c)
        for n in range(index2.shape[0]):
            c = float(np.ravel(coeff2)[n])
            if abs(c) > tol:
                pos = (int(index2[n, 0]), int(index2[n, 1]))
                sns = (int(index2[n, 2]), int(index2[n, 3]))
                vl, vm, vr = self.environment_s1_s2(p, sns, pos)
                if max(pos) < p:
                    acc(('L',), vl * c)
                elif min(pos) > p:
                    acc(('R',), vr * c)
                elif max(pos) == p:
                    acc(('LM', sns[1]), vl * c)
                elif min(pos) == p:
                    acc(('MR', sns[0]), vr * c)
                else:
                    env.setdefault(('LR',), []).append((c, vl, vr))

    def heff_apply(self, x, s):
        """``Heff x`` from ``opt_env`` (the ``tau`` part of ``MPSClass.py:837-858``)."""
        t = x.reshape(s)
        out = np.zeros_like(t)
        for key, e in self.opt_env.items():
            if key[0] == 'L':
                out += np.tensordot(e, t, axes=([1], [0]))
            elif key[0] == 'R':
                out += np.tensordot(t, e, axes=([2], [1]))
            elif key[0] == 'M':
                out += np.einsum('st,atb->asb', e, t)
            elif key[0] == 'LM':
                y = np.tensordot(e, t, axes=([1], [0]))
                out += np.einsum('st,atb->asb', self.operators[key[1]], y)
            elif key[0] == 'MR':
                y = np.tensordot(t, e, axes=([2], [1]))
                out += np.einsum('st,atb->asb', self.operators[key[1]], y)
            else:
                for c, vl, vr in e:
                    y = np.tensordot(vl, t, axes=([1], [0]))
                    out += c * np.tensordot(y, vr, axes=([2], [1]))
        return out.reshape(-1)

    def effective_hamiltonian_dense(self, p, index1, index2, coeff1, coeff2, tol=1e-12):
        """``MPSClass.py:613-662`` (``eig_way=0``): ``sum c kron(vL, vM, vR)``."""
        s = (self.virtual_dim[p], self.phys_dim, self.virtual_dim[p + 1])
        eye = [np.eye(s[0]), np.eye(s[1]), np.eye(s[2])]
        h = np.zeros((int(np.prod(s)),) * 2)

        def kron3(v):
            v = [eye[i] if v[i] is None else v[i] for i in range(3)]
            return np.kron(np.kron(v[0], v[1]), v[2])

        for n in range(index1.shape[0]):
            c = float(np.ravel(coeff1)[n])
            sn = int(index1[n, 1])
            if abs(c) > tol and np.linalg.norm(self.operators[sn].reshape(-1)) > tol:
                h += c * kron3(self.environment_s1(p, sn, int(index1[n, 0])))
        for n in range(index2.shape[0]):
            c = float(np.ravel(coeff2)[n])
            if abs(c) > tol:
                h += c * kron3(self.environment_s1_s2(
                    p, (int(index2[n, 2]), int(index2[n, 3])),
                    (int(index2[n, 0]), int(index2[n, 1]))))
        return h, s

    def update_tensor_eigs(self, p, index1, index2, coeff1, coeff2, tau, is_real, tol):
        """``MPSClass.py:860-891``: gauge move, environments, ARPACK on ``I - tau*Heff``."""
        c0 = self.center
        self.correct_orthogonal_center(p)
        if self.is_save_op:
            lo, hi = (c0, p) if c0 < p else (p, c0)
            for n in range(lo, hi + 1):
                self.del_bad_effective_operators(n)
            self.update_all_effective_id()
        if self.eig_way == 0:
            h, s = self.effective_hamiltonian_dense(p, index1, index2, coeff1, coeff2)
            h = np.eye(h.shape[0]) - tau * h
            op = h
        else:
            self.all_environments_optimized(p, index1, index2, coeff1, coeff2, tol)
            s = (self.virtual_dim[p], self.phys_dim, self.virtual_dim[p + 1])
            dim = int(np.prod(s))

            def mv(x):
                self.n_matvec += 1
                x = np.asarray(x).reshape(-1)
                return x - tau * self.heff_apply(x, s)

            op = LinearOperator((dim, dim), matvec=mv, dtype=float)
        v0 = self.mps[p].reshape(-1)
        if v0.size == 1:
            vec = np.ones(1)
        else:
            vec = eigsh(op, k=1, which='LM', v0=v0, tol=tol)[1]
        self.mps[p] = np.asarray(vec).reshape(s)
        if is_real:
            self.mps[p] = self.mps[p].real
        self.opt_env = {}

    # -------------------------------------------------------- observables
    def observation_s1(self, sn, position):
        """``MPSClass.py:1105-1120`` (direct transfer chain, same value as the cache)."""
        op = self.operators[sn]
        if position > self.center:
            v = self._r2l(position, op)
            v = self.contract_v_l0_to_l1(position - 1, self.center - 1, v)
        else:
            v = self._l2r(position, op)
            v = self.contract_v_l0_to_l1(position + 1, self.center + 1, v)
        return np.trace(v)

    def observation_s1_s2(self, sns, positions):
        """``MPSClass.py:1122-1157`` (non-cache branch 1147-1157)."""
        sns = list(sns)
        positions = list(positions)
        if positions[0] > positions[1]:
            sns = sns[::-1]
            positions = positions[::-1]
        p0, q0 = positions
        v = self.contract_v_l0_to_l1(self.center, p0) if self.center < p0 else None
        v = tb.bound_vec_operator_left2right(self.mps[p0], self.operators[sns[0]], v)
        v = self.contract_v_l0_to_l1(p0 + 1, q0, v)
        v = tb.bound_vec_operator_left2right(self.mps[q0], self.operators[sns[1]], v)
        if q0 < self.center:
            v = self.contract_v_l0_to_l1(q0 + 1, self.center + 1, v)
        return np.trace(v)

    def observe_magnetization(self, sn):
        """``MPSClass.py:1159-1169``."""
        return np.array([[self.observation_s1(sn, i).real] for i in range(self.length)])

    def observe_bond_energy(self, index2, coeff2):
        """``MPSClass.py:1171-1182``."""
        nh = index2.shape[0]
        eb = np.zeros((nh, 1))
        for n in range(nh):
            eb[n] = np.ravel(coeff2)[n] * self.observation_s1_s2(index2[n, 2:4], index2[n, :2]).real
        return eb

    def observe_bond_energy_from_jxyz(self, pos2, jx, jy, jz, tol=1e-20):
        """``MPSClass.py:1184-1194``: ``jx<SxSx> + jy Re<SySy> + jz<SzSz>`` per pair."""
        nh = pos2.shape[0]
        eb = np.zeros((nh, 1))
        for n in range(nh):
            if abs(jx) > tol:
                eb[n] += jx * self.observation_s1_s2((1, 1), pos2[n, :2]).real
            if abs(jy) > tol:
                eb[n] += jy * np.real(self.observation_s1_s2((2, 2), pos2[n, :2]))
            if abs(jz) > tol:
                eb[n] += jz * self.observation_s1_s2((3, 3), pos2[n, :2]).real
        return eb

    def observe_correlators_from_middle(self, op1, op2, ob_len=None):
        """``MPSClass.py:1196-1212``."""
        if ob_len is None:
            ob_len = self.length
        corr = []
        pos1 = int(round(self.length / 2))
        pos2 = pos1 + 1
        k = 0
        while pos1 > -0.1 and pos2 < self.length and ob_len > -0.1:
            corr.append(self.observation_s1_s2((op1, op2), (pos1, pos2)).real)
            if k % 2 == 0:
                pos1 -= 1
            else:
                pos2 += 1
            k += 1
            ob_len -= 1
        return np.array(corr)

    def calculate_entanglement_spectrum(self):
        """``MPSClass.py:965-985`` (``if_fast=True``): SVD centre moves fill ``lm``."""
        way, c = self.decomp_way, self.center
        self.decomp_way = 'svd'
        p0, p1 = self.length - 1, 0
        for n in range(self.length - 1):
            if self.lm[n].size == 0:
                p0 = min(p0, n)
                p1 = max(p1, n)
        self.correct_orthogonal_center(p0)
        self.correct_orthogonal_center(p1 + 1)
        self.correct_orthogonal_center(c)
        self.decomp_way = way

    def calculate_entanglement_entropy(self):
        """``MPSClass.py:987-992``."""
        for i in range(self.length - 1):
            self.ent[i] = -1 if self.lm[i].size == 0 else tb.entanglement_entropy(self.lm[i])

    def clean_to_save(self):
        """``MPSClass.py:1396-1405``."""
        self.effect_s, self.effect_ss, self.effective_id, self.opt_env = {}, {}, {}, {}

    def full_state(self):
        """Contract the MPS to a dense vector (small N only)."""
        v = self.mps[0].reshape(-1, self.mps[0].shape[2])
        for n in range(1, self.length):
            a = self.mps[n]
            v = (v @ a.reshape(a.shape[0], -1)).reshape(-1, a.shape[2])
        return v.reshape(-1)
