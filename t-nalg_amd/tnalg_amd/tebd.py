"""Imaginary-time TEBD on the device kernels: the drop-in for
``tebd_standard(para)`` (``algorithms/TEBDalgo.py:11-87``) with the MPS of
``MpsStandardTEBD`` (``library/MPSClass.py:1408-1447``).

Every gate step reuses the kernels of the DMRG path: theta = A_p A_{p+1}
(grouped GEMM), the gate on the merged physical index (the mixer, with the
gate as a d^2 x d^2 operator of the solve's slot of the operator table),
the truncated split (block Jacobi SVD, or Householder QR for small pairs)
and the centre moves (CholeskyQR3 / Householder QR); the normalisations of
``orthogonalize_mps(normalize=True)`` and ``norm_mps(True)`` run as a
batched device normalisation.  No environment is built.

For a nearest-neighbour pair, applying the two gate halves of
``evolve_gate_tebd`` (the SVD split of ``expm(-tau h)``) and truncating
with ``truncate_mps_tebd`` is the two-site update theta -> G theta followed
by the truncated SVD of theta with the centre moving to p2: the gate
halves contracted over their bond are G itself, and the reference's QR of
A_{p2} before its SVD of A_{p1} leaves the singular values and kept vectors
of G theta.  Only nearest-neighbour pairs (p, p + 1) are supported (the
reference's identity strings for longer pairs are not built).

Outputs follow the reference's ``ob`` keys.  ``lm`` and ``ent`` are the
entanglement spectra and entropies of the recorded state (SVD centre moves
at record time); the reference records the raw singular values each
truncation produced during the last sweep of gates, before the later gates
of that sweep acted."""
import numpy as np

from .dmrg import MpsOpenBoundaryClass
from .parameters import (common_parameters_dmrg, generate_parameters_standard_tebd,
                         hamiltonian_heisenberg, make_consistent_parameter_dmrg,
                         parameter_dmrg_chain)


class MpsStandardTEBD(MpsOpenBoundaryClass):
    """``MPSClass.py:1408-1414``: the open-boundary MPS of a TEBD run (host
    copies of the tensors plus the device solve)."""

    def __init__(self, solve, engine, chi, spin='half'):
        MpsOpenBoundaryClass.__init__(self, solve, engine, spin=spin)
        self.chi = chi


def geometry_para(para):
    """The solve parameters of a TEBD run: the XXZ chain of ``para`` as a
    two-site DMRG problem (its tables size the buffers and give the
    observables); the initial MPS is the random one of bond dimension chi."""
    pd = dict(parameter_dmrg_chain(), **common_parameters_dmrg())
    for k in ('l', 'chi', 'jxy', 'jz', 'hx', 'hz', 'spin'):
        pd[k] = para[k]
    pd['bound_cond'] = 'open'
    pd = make_consistent_parameter_dmrg(pd)
    pd['dmrg_sites'] = 2
    pd['chi_init'] = int(para['chi'])
    pd['ob_position'] = 0
    return pd


def _bonds(para):
    pos = np.asarray(para['positions_h2'], dtype=int).reshape(-1, 2)
    out = []
    for a, b in pos.tolist():
        p1, p2 = min(a, b), max(a, b)
        if p2 != p1 + 1:
            raise ValueError('TEBD gates act on nearest-neighbour pairs only (got %d, %d)' % (a, b))
        out.append(p1)
    return out


def schedule_gate_sweep(solve, bonds, gate=0):
    """Queue one sweep of gates over the pairs ``(p, p+1)`` of ``bonds`` in
    order (``TEBDalgo.py:50-53``), then the normalisation of the centre.
    Consecutive pairs share a step: 'gate2' at p splits the pair p - 1 (the
    centre moves onto p) and applies the gate to (p, p+1)."""
    from .engine import Step
    c = solve._sched_center
    q = solve.queue
    pending = None   # left site of the pair whose gated theta awaits its split
    for p in bonds:
        if pending is not None and p == pending + 1:
            q.append(Step('gate2', p, +1, gate=gate))
            pending = p
            continue
        if pending is not None:
            q.append(Step('tsplit', pending, +1))
            c = pending + 1
            pending = None
        tgt = p if c <= p else p + 1   # the centre on the pair
        while c != tgt:
            dirn = 1 if tgt > c else -1
            q.append(Step('tmove', c, dirn))
            c += dirn
        q.append(Step('gate2', p, 0, gate=gate))
        pending = p
    if pending is not None:
        q.append(Step('tsplit', pending, +1))
        c = pending + 1
    q.append(Step('tnorm', c, 0))
    solve._sched_center = c


def _bond_energy(engine, solve, bonds, jxy, jz):
    """``observe_bond_energy_from_jxyz(positions_h2, jxy, jxy, jz)`` summed
    (``TEBDalgo.py:63-64``), one batched measurement."""
    from .observe import measure
    rows, coefs = [], []
    for p in bonds:
        for op, c in ((1, jxy), (2, jxy), (3, jz)):
            if abs(c) > 1e-20:
                rows.append((p, p + 1, op, op))
                coefs.append(c)
    if not rows:
        return 0.0
    _, v, _ = measure(engine, [solve], [([], rows)])[0]
    return float(np.dot(coefs, v))


def _record(ob, A, para, bonds):
    """The outputs of ``TEBDalgo.py:66-77`` for the current state."""
    A.refresh_from_device()
    A.calculate_entanglement_spectrum(if_fast=False)
    A.calculate_entanglement_entropy()
    ob['lm'].append([np.asarray(x).copy() for x in A.lm])
    ob['ent'].append(A.ent.copy())
    mx = A.observe_magnetization(1)
    mz = A.observe_magnetization(3)
    eb = A.observe_bond_energy_from_jxyz(np.array([(p, p + 1) for p in bonds]).reshape(-1, 2),
                                         para['jxy'], para['jxy'], para['jz'])
    ob['mx'].append(mx)
    ob['mz'].append(mz)
    ob['eb'].append(eb)
    l = para['l']
    ob['e_site'].append(float(np.sum(eb)) / l + float(np.sum(mx)) * para['hx'] / l +
                        float(np.sum(mz)) * para['hz'] / l)


def tebd_standard(para=None, seed=None, engine=None, verbose=False):
    """``tebd_standard(para)`` (``TEBDalgo.py:11-87``) on the device: random
    MPS (seeded), centre to site 0 and normalised; for each ``tau = tau0
    dtau^nt`` the gate ``expm(-tau h)`` of the bond Hamiltonian
    ``hamiltonian_heisenberg(spin, jxy, jxy, jz, hx/2, hz/2)``, sweeps of
    gates over ``positions_h2`` with truncation to chi and normalisation;
    with ``if_break``, every ``dt_ob`` sweeps the bond energy per site is
    compared with the previous test's and the run for this tau stops (after
    recording the outputs) once they differ by less than ``break_tol``;
    ``save_mode='all'`` records the outputs every ``dt_ob`` sweeps instead.
    Returns ``(mps, ob, para)``; the device state of ``mps`` is released
    (``clean_to_save``)."""
    from scipy.linalg import expm
    from .engine import Engine, Step
    if para is None:
        para = generate_parameters_standard_tebd()
    bonds = _bonds(para)
    pd = geometry_para(para)
    eng = engine if engine is not None else Engine(lowp=False, plan_ahead=False)
    s = eng.add_solve(pd, seed=0 if seed is None else seed, gates=1)
    s.schedule_canonicalise()
    s.queue.append(Step('tnorm', 0, 0))
    eng.run_queues([s])
    A = MpsStandardTEBD(s, eng, int(para['chi']), spin=para['spin'])
    hamilt = hamiltonian_heisenberg(para['spin'], para['jxy'], para['jxy'], para['jz'],
                                    para['hx'] / 2, para['hz'] / 2)
    ob = {'lm': [], 'ent': [], 'mx': [], 'mz': [], 'eb': [], 'e_site': []}
    time_now = 0.0
    e0 = 100.0
    l = int(para['l'])
    for nt in range(int(para['taut'])):
        tau = para['tau0'] * para['dtau'] ** nt
        if verbose:
            print('Now tau = ' + str(tau))
        s.set_gate(0, expm(-tau * hamilt))
        for t in range(int(para['iterate_time'])):
            schedule_gate_sweep(s, bonds)
            eng.run_queues([s])
            time_now += tau
            if para.get('save_mode') == 'all':
                if t % para['dt_ob'] == 0:
                    _record(ob, A, para, bonds)
            elif para.get('if_break', True) and t % para['dt_ob'] == 0:
                e1 = _bond_energy(eng, s, bonds, para['jxy'], para['jz']) / l
                if abs(e0 - e1) < para['break_tol']:
                    _record(ob, A, para, bonds)
                    if verbose:
                        print('effective beta = %g; converged with E = %.12g'
                              % (time_now, ob['e_site'][-1]))
                    break
                e0 = e1
    A.refresh_from_device()
    A.time = time_now
    A.clean_to_save()
    return A, ob, para
