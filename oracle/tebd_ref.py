"""Imaginary-time TEBD of the reference, restated in numpy (oracle; test use only).

Restates ``algorithms/TEBDalgo.py:11-87`` (``tebd_standard``) with the MPS
of ``library/MPSClass.py:1408-1447`` (``MpsStandardTEBD``) for
nearest-neighbour pairs: the gate ``expm(-tau h)`` of the bond Hamiltonian
``hamiltonian_heisenberg(spin, jxy, jxy, jz, hx/2, hz/2)``
(``library/HamiltonianModule.py:141-147``) acts on the pair tensor
theta = A_p A_{p+1}, and the truncated SVD of G theta (at most chi values
kept, the centre moving to p+1, the centre tensor normalised) replaces the
reference's two gate halves (``evolve_gate_tebd``) followed by its QR and
truncating SVD moves (``truncate_mps_tebd``), which give the same singular
values and kept vectors for a nearest-neighbour pair.  Pinned by tests
against the dense product of the gates on the full state (no truncation)
and against exact diagonalisation (small tau).
"""
import numpy as np
from scipy.linalg import expm

from . import tensor_ops as tb
from .model import operator_list, spin_half_operators
from .mps_ref import RefMps


def bond_hamiltonian(jxy, jz, hx, hz):
    """``hamiltonian_heisenberg('half', jxy, jxy, jz, hx, hz)``
    (``library/HamiltonianModule.py:141-147``), 4 x 4 on ``s1 d + s2``."""
    o = spin_half_operators()
    h = jxy * np.kron(o['sx'], o['sx']) + jxy * np.kron(o['sy'], o['sy']).real + \
        jz * np.kron(o['sz'], o['sz'])
    h = h - hx * (np.kron(o['id'], o['sx']) + np.kron(o['sx'], o['id']))
    h = h - hz * (np.kron(o['id'], o['sz']) + np.kron(o['sz'], o['id']))
    return np.real(h)


def move_center(mps, p):
    """``orthogonalize_mps(center, p, normalize=True)`` by QR (MPSClass.py:186-247)."""
    mps.correct_orthogonal_center(p)
    mps.mps[p] = mps.mps[p] / np.linalg.norm(mps.mps[p])


def gate_pair(mps, p, gate, chi):
    """Gate on the pair (p, p+1) with the centre on it, truncated SVD,
    centre to p+1, normalised (``MPSClass.py:1416-1447``)."""
    a, b = mps.mps[p], mps.mps[p + 1]
    dl, d, _ = a.shape
    dr = b.shape[2]
    theta = np.tensordot(a, b, ([2], [0]))                     # (dl, d, d, dr)
    theta = np.einsum('st,atb->asb', gate, theta.reshape(dl, d * d, dr))
    u, s, vt = np.linalg.svd(theta.reshape(dl * d, d * dr), full_matrices=False)
    k = min(chi, len(s))
    mps.mps[p] = u[:, :k].reshape(dl, d, k)
    c = (s[:k, None] * vt[:k]).reshape(k, d, dr)
    mps.mps[p + 1] = c / np.linalg.norm(c)
    mps.lm[p] = s[:k].copy()
    mps.virtual_dim[p + 1] = k
    mps.center = p + 1


def gate_sweep(mps, bonds, gate, chi):
    """One sweep of gates over ``bonds`` in order, then ``norm_mps(True)``
    (``TEBDalgo.py:50-53``)."""
    for p in bonds:
        if mps.center not in (p, p + 1):
            move_center(mps, p if mps.center < p else p + 1)
        if mps.center == p + 1:
            move_center(mps, p)
        gate_pair(mps, p, gate, chi)
    mps.mps[mps.center] = mps.mps[mps.center] / np.linalg.norm(mps.mps[mps.center])


def bond_energy_sum(mps, bonds, jxy, jz):
    pos = np.array([(p, p + 1) for p in bonds]).reshape(-1, 2)
    return float(np.sum(mps.observe_bond_energy_from_jxyz(pos, jxy, jxy, jz)))


def record(ob, mps, para, bonds):
    """``TEBDalgo.py:66-77`` for the current state (spectra of every bond by
    SVD moves, as the device driver records them)."""
    mps.lm = [np.zeros(0) for _ in range(mps.length - 1)]
    mps.calculate_entanglement_spectrum()
    mps.calculate_entanglement_entropy()
    ob['lm'].append([x.copy() for x in mps.lm])
    ob['ent'].append(mps.ent.copy())
    mx, mz = mps.observe_magnetization(1), mps.observe_magnetization(3)
    pos = np.array([(p, p + 1) for p in bonds]).reshape(-1, 2)
    eb = mps.observe_bond_energy_from_jxyz(pos, para['jxy'], para['jxy'], para['jz'])
    ob['mx'].append(mx)
    ob['mz'].append(mz)
    ob['eb'].append(eb)
    l = para['l']
    ob['e_site'].append(float(np.sum(eb)) / l + float(np.sum(mx)) * para['hx'] / l +
                        float(np.sum(mz)) * para['hz'] / l)


def tebd_standard(para, seed=0, rng=None):
    """``TEBDalgo.py:11-87`` (``save_mode`` 'final' or 'all'); the random
    initial MPS is ``random_open_mps(l, d, chi, RandomState(seed))``.
    Returns ``(mps, ob)``."""
    rng = np.random.RandomState(seed) if rng is None else rng
    l, chi = int(para['l']), int(para['chi'])
    bonds = [min(a, b) for a, b in np.asarray(para['positions_h2']).reshape(-1, 2).tolist()]
    mps = RefMps(l, 2, chi, operator_list(para['hx'], para['hz']), rng, way='qr')
    move_center(mps, 0)
    h = bond_hamiltonian(para['jxy'], para['jz'], para['hx'] / 2, para['hz'] / 2)
    ob = {'lm': [], 'ent': [], 'mx': [], 'mz': [], 'eb': [], 'e_site': []}
    e0 = 100.0
    for nt in range(int(para['taut'])):
        gate = expm(-para['tau0'] * para['dtau'] ** nt * h)
        for t in range(int(para['iterate_time'])):
            gate_sweep(mps, bonds, gate, chi)
            if para.get('save_mode') == 'all':
                if t % para['dt_ob'] == 0:
                    record(ob, mps, para, bonds)
            elif para.get('if_break', True) and t % para['dt_ob'] == 0:
                e1 = bond_energy_sum(mps, bonds, para['jxy'], para['jz']) / l
                if abs(e0 - e1) < para['break_tol']:
                    record(ob, mps, para, bonds)
                    break
                e0 = e1
    return mps, ob


def dense_gate_sweep(psi, l, bonds, gate):
    """The sweep on a dense state vector: ``G_b`` on sites (p, p+1) for every
    bond in order, then normalised (the exact result without truncation)."""
    for p in bonds:
        t = psi.reshape(2 ** p, 4, 2 ** (l - p - 2))
        psi = np.einsum('st,atb->asb', gate, t).reshape(-1)
    return psi / np.linalg.norm(psi)
