"""The reference's finite-size DMRG sweep driver (oracle; test use only).

Restates ``algorithms/DMRG_anyH.py:18-101`` (``dmrg_finite_size``) and
``get_bond_energies`` (``DMRG_anyH.py:338-346``) on top of ``mps_ref.RefMps``.
Printing, thread pools and the wall clock are left out; the sweep schedule,
the convergence test and the outputs are the reference's.
"""
import numpy as np

from .mps_ref import RefMps


def get_bond_energies(eb_full, positions, index2):
    """``DMRG_anyH.py:338-346``: sum the rows of ``eb_full`` per bond."""
    eb = np.zeros((positions.shape[0], 1))
    for i in range(eb_full.size):
        hit = (index2[i, 0] == positions[:, 0]) & (index2[i, 1] == positions[:, 1])
        eb[np.nonzero(hit)] += np.ravel(eb_full)[i]
    return eb


def local_step_order(length, ob_position):
    """Sites updated in one sweep (``DMRG_anyH.py:47-64``)."""
    return (list(range(ob_position + 1, length)) + list(range(length - 2, -1, -1))
            + list(range(1, ob_position)))


def dmrg_finite_size(para, seed=0, rng=None, max_steps=None):
    """``DMRG_anyH.py:18-101`` with a seeded initial MPS.

    ``max_steps`` (optional) stops after that many local updates; it is used
    only to time a bounded sample and returns ``None`` outputs in that case.
    Returns ``(ob, A, info, para)``.
    """
    if rng is None:
        rng = np.random.RandomState(seed)
    A = RefMps(para['l'], para['d'], para['chi'], para['op'], rng, way='qr',
               is_save_op=para.get('is_save_op', True), eig_way=para.get('eigWay', 1))
    A.correct_orthogonal_center(para['ob_position'])
    e0 = 0.0
    info = {'convergence': 1, 'sweeps': 0, 'energies': []}
    ob = {}
    steps = 0
    order = local_step_order(para['l'], para['ob_position'])
    for t in range(para['sweep_time']):
        if_ob = ((t + 1) % para['dt_ob'] == 0) or t == para['sweep_time'] - 1
        for n in order:
            A.update_tensor_eigs(n, para['index1'], para['index2'], para['coeff1'],
                                 para['coeff2'], para['tau'], para['is_real'], para['eigs_tol'])
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
    A.clean_to_save()
    return ob, A, info, para


def make_para(couplings, chi, sweep_time, tau=1e-4, eigs_tol=1e-5, break_tol=1e-8, dt_ob=4,
              ob_position=0, eig_way=1, is_save_op=True):
    """A ``para`` dict with the keys ``dmrg_finite_size`` reads (``PAR:8-26``)."""
    para = dict(couplings)
    para.update(chi=chi, sweep_time=sweep_time, tau=tau, eigs_tol=eigs_tol, break_tol=break_tol,
                is_real=True, dt_ob=dt_ob, ob_position=ob_position, eigWay=eig_way,
                isParallel=False, isParallelEnvLMR=False, is_save_op=is_save_op,
                if_print_detail=False, d=2)
    para['nh'] = para['index2'].shape[0]
    return para
