"""End-to-end parity of the GPU DMRG engine (``dmrg_finite_size``) with the
oracle, exact diagonalisation and the reference's saved runs.

Golden numbers of the reference runs (``data_dmrg/chainN12_j(1,1)_h(0,0)_chi16open.pr``
and ``..._chi24open.pr``, N=12 open chain, every coeff2 = 1): converged
energies -8.676296538186877 (chi=16) and -8.676296659449305 (chi=24); chi=16
bond energies eb_full[0:3] = (-0.44060803847908186, -0.44060803847908186,
-0.20531997892240417), eb[0:2] = (-1.086536055880568, -0.5156163135305695),
ent[5] = 0.567208838192397, lm[5][:4] = (0.92381354272716787,
0.24934275621291499, 0.24934275619104823, 0.14810633922157662).
"""
import numpy as np
import pytest

from oracle import ed, model

pytestmark = pytest.mark.gpu


def _para(couplings, chi, sweeps, tau=1e-4, eigs_tol=1e-10, break_tol=1e-14, dt_ob=1):
    from tnalg_amd import parameters as P
    para = P.common_parameters_dmrg()
    para.update(couplings)
    para.update(chi=chi, sweep_time=sweeps, tau=tau, eigs_tol=eigs_tol, break_tol=break_tol,
                dt_ob=dt_ob, spin='half', d=2, lattice='arbitrary')
    para['nh'] = para['index2'].shape[0]
    return para


def test_heisenberg_n8_exact(gpu):
    from tnalg_amd import dmrg_finite_size
    c = model.chain_couplings(8, 1.0, 1.0, hz=0.2)
    para = _para(c, 16, 6)
    print('case: N=8 chi=16 heisenberg hz=0.2', flush=True)
    ob, A, info, _ = dmrg_finite_size(para, seed=1)
    e0, _ = ed.ground_state(8, c['op'], c['index1'], c['coeff1'], c['index2'], c['coeff2'])
    e = float(ob['e_per_site'][0]) * 8
    assert abs(e - e0) < 1e-10 * abs(e0), (e, e0)
    assert abs(A.norm_mps() - 1) < 1e-12


def test_fixture_chain_chi16(gpu):
    from tnalg_amd import dmrg_finite_size
    c = model.chain_couplings(12, 1.0, 1.0, convention='easystart')
    para = _para(c, 16, 30, tau=1e-3, eigs_tol=1e-12)
    print('case: fixture N=12 chi=16', flush=True)
    ob, A, info, _ = dmrg_finite_size(para, seed=3)
    e = float(ob['e_per_site'][0]) * 12
    assert abs(e - (-8.676296538186877)) < 2e-8 * 8.7, e
    assert np.allclose(ob['eb_full'][:3, 0], [-0.44060803847908186, -0.44060803847908186,
                                              -0.20531997892240417], atol=2e-7)
    assert np.allclose(ob['eb'][:2, 0], [-1.086536055880568, -0.5156163135305695], atol=2e-7)
    assert np.allclose(A.lm[5][:4], [0.92381354272716787, 0.24934275621291499,
                                     0.24934275619104823, 0.14810633922157662], atol=1e-6)
    assert abs(A.ent[5, 0] - 0.567208838192397) < 1e-6


def test_matches_oracle_c1(gpu):
    """Config-1 Hamiltonian (Parameters convention) with a field: GPU vs the
    reference restatement from the same seeded MPS, converged."""
    from oracle import dmrg_ref
    from tnalg_amd import dmrg_finite_size
    c = model.chain_couplings(12, 1.0, 1.0, hz=0.15)
    para = _para(c, 16, 12)
    print('case: C1-like N=12 chi=16 hz=0.15', flush=True)
    ob, A, info, _ = dmrg_finite_size(para, seed=5)
    ref = dmrg_ref.make_para(c, 16, 12, eigs_tol=1e-10, break_tol=1e-14, dt_ob=1)
    rob, rA, rinfo, _ = dmrg_ref.dmrg_finite_size(ref, seed=5)
    e, er = float(ob['e_per_site'][0]), float(rob['e_per_site'][0])
    assert abs(e - er) < 1e-9 * abs(er), (e, er)
    assert np.allclose(ob['mz'], rob['mz'], atol=1e-6)
    assert np.allclose(ob['eb_full'], rob['eb_full'], atol=1e-6)
    assert np.allclose(ob['corr_z'], rob['corr_z'], atol=1e-6)
    for b in range(11):
        k = min(len(A.lm[b]), len(rA.lm[b]))
        assert np.allclose(A.lm[b][:k], rA.lm[b][:k], atol=1e-5)


@pytest.mark.parametrize('n,chi', [(16, 64), (20, 256)])
def test_spectrum_is_schmidt_values(gpu, n, chi):
    """The spectra of the SVD moves (block-Jacobi SVD of the CholeskyQR3
    factor; chi=256 takes 15 rounds per sweep) are the Schmidt coefficients
    of the final MPS state at every bond."""
    from tnalg_amd import dmrg_finite_size
    c = model.chain_couplings(n, 1.0, 0.8, hz=0.1)
    para = _para(c, chi, 2)
    print('case: spectrum N=%d chi=%d' % (n, chi), flush=True)
    ob, A, info, _ = dmrg_finite_size(para, seed=2)
    psi = A.mps[0].reshape(-1, A.mps[0].shape[2])
    for a in A.mps[1:]:
        psi = (psi @ a.reshape(a.shape[0], -1)).reshape(-1, a.shape[2])
    psi = psi.reshape(-1)
    assert abs(np.linalg.norm(psi) - 1.0) < 1e-10
    for b in range(n - 1):
        sv = np.linalg.svd(psi.reshape(2 ** (b + 1), -1), compute_uv=False)
        lm = np.asarray(A.lm[b])
        assert len(lm) == min(chi, 2 ** (b + 1), 2 ** (n - b - 1)), b
        assert np.all(np.diff(lm) <= 0), b
        assert np.allclose(lm, sv[:len(lm)], atol=1e-10), (b, np.abs(lm - sv[:len(lm)]).max())


def test_long_range_j1j2_exact(gpu):
    """Straddling (long-range) terms: J1-J2 chain, chi large enough to be exact."""
    from tnalg_amd import dmrg_finite_size
    bonds = [(i, i + 1) for i in range(9)] + [(i, i + 2) for i in range(8)] + [(0, 9), (2, 7)]
    jj = [1.0] * 9 + [0.4] * 8 + [0.3, -0.2]
    c = model.couplings_from_bonds(10, bonds, jj, jj, hz=0.1)
    para = _para(c, 32, 8)
    print('case: J1-J2 N=10 chi=32 with long bonds', flush=True)
    ob, A, info, _ = dmrg_finite_size(para, seed=2)
    e0, _ = ed.ground_state(10, c['op'], c['index1'], c['coeff1'], c['index2'], c['coeff2'])
    e = float(ob['e_per_site'][0]) * 10
    assert abs(e - e0) < 1e-9 * abs(e0), (e, e0)


def test_snake_j1j2_4x4_exact(gpu):
    """The 2D path of the dense-coupling configuration at a size ED reaches:
    J1-J2 (J2 = 0.5) on a 4x4 open square lattice mapped along the snake
    (couplings up to 7 sites apart), chi = 256 (exact), mixed precision,
    against exact diagonalisation of the same coupling list."""
    from tnalg_amd import dmrg_finite_size
    from tnalg_amd import parameters as P
    para = P.generate_parameters_dmrg('snakeJ1J2')
    para.update(square_width=4, square_height=4, chi=256, sweep_time=8, dt_ob=8,
                eigs_tol=1e-10, tau=1e-4, break_tol=0.0)
    para = P.make_consistent_parameter_dmrg(para)
    print('case: snake J1-J2 4x4 chi=256 mixed', flush=True)
    ob, A, info, _ = dmrg_finite_size(para, seed=4, precision='mixed')
    e0, _ = ed.ground_state(16, para['op'], para['index1'], para['coeff1'], para['index2'],
                            para['coeff2'])
    e = float(ob['e_per_site'][0]) * 16
    assert abs(e - e0) < 1e-9 * abs(e0), (e, e0)


@pytest.mark.parametrize('precision', ['fp32', 'bf16'])
def test_mixed_precision_reaches_fp64(gpu, precision):
    from tnalg_amd import dmrg_finite_size
    c = model.chain_couplings(16, 1.0, 1.0, hz=0.1)
    print('case: mixed', precision, flush=True)
    ob64, _, _, _ = dmrg_finite_size(_para(c, 32, 8), seed=4, precision='fp64')
    obm, _, _, _ = dmrg_finite_size(_para(c, 32, 8), seed=4, precision=precision)
    e64, em = float(ob64['e_per_site'][0]), float(obm['e_per_site'][0])
    assert abs(e64 - em) < 1e-9 * abs(e64), (e64, em)


def test_saved_result_reloads(gpu, tmp_path):
    """``save_pr`` / ``load_pr`` of ``(ob, A, info, para)`` (the reference's
    result cache, ``library/BasicFunctions.py:29-95``): the loaded ``A``
    rebuilds its device state from the host tensors and gives the same
    observables; ``calculate_entanglement_spectrum(if_fast=False)`` on it
    recomputes every spectrum with SVD moves and gives the same spectra and
    entropies; an ``A`` released by ``clean_to_save`` still observes."""
    from tnalg_amd import dmrg_finite_size, load_pr, save_pr
    c = model.chain_couplings(10, 1.0, 0.9, hz=0.15)
    para = _para(c, 16, 3)
    print('case: save/load N=10 chi=16', flush=True)
    ob, A, info, para = dmrg_finite_size(para, seed=4)
    mz = A.observe_magnetization(3)
    eb = A.observe_bond_energy(para['index2'], para['coeff2'])
    save_pr(str(tmp_path), 'run.pr', (ob, A, info, para), ('ob', 'A', 'info', 'para'))
    B, ob2 = load_pr(str(tmp_path / 'run.pr'), ('A', 'ob'))
    assert B._solve is None and np.array_equal(ob2['eb_full'], ob['eb_full'])
    assert np.allclose(B.observe_magnetization(3), mz, atol=1e-12)
    assert np.allclose(B.observe_bond_energy(para['index2'], para['coeff2']), eb, atol=1e-12)
    B.calculate_entanglement_spectrum(if_fast=False)
    B.calculate_entanglement_entropy()
    for b in range(9):
        assert np.allclose(B.lm[b], A.lm[b], atol=1e-10), b
    assert np.allclose(B.ent, A.ent, atol=1e-9)
    assert abs(B.norm_mps() - 1) < 1e-12
    A.clean_to_save()
    assert A._solve is None
    assert np.allclose(A.observe_magnetization(3), mz, atol=1e-12)


def test_device_initial_draw(gpu):
    """init_rng='device': the initial MPS of a seed is drawn on the GPU with
    the shapes of the host draw, the same bits alone and in a batch, standard
    normal entries; the solve then runs to the same ground state as one
    started from the host draw."""
    from tnalg_amd.dmrg import _schedule_from
    from tnalg_amd.engine import Engine
    para = _para(model.chain_couplings(12, 1.0, 1.0, hz=0.1), 16, 4)
    print('device initial draw', flush=True)
    e1 = Engine(precision_schedule=_schedule_from('fp64'), lowp=False, init_rng='device')
    a = e1.add_solves([(para, 5), (para, 6)])
    e2 = Engine(precision_schedule=_schedule_from('fp64'), lowp=False, init_rng='device')
    b = e2.add_solves([(para, 5)])
    ha, hb = a[0].host_mps(), b[0].host_mps()
    hn = Engine(precision_schedule=_schedule_from('fp64'), lowp=False).add_solves(
        [(para, 5)])[0].host_mps()
    assert [x.shape for x in ha] == [x.shape for x in hn]
    assert all(np.array_equal(x, y) for x, y in zip(ha, hb))
    assert not np.array_equal(ha[1], a[1].host_mps()[1])
    v = np.concatenate([x.ravel() for x in ha])
    assert abs(v.mean()) < 0.05 and abs(v.std() - 1) < 0.05
    energies = []
    for eng, s in ((e1, a[0]), (e2, b[0])):
        s.schedule_canonicalise()
        for t in range(4):
            s.schedule_sweep(t)
        eng.run_queues([s])
        energies.append(s.last_theta)
    assert energies[0] == energies[1]


def test_chained_next_step_same_bits(gpu):
    """A step's Lanczos call that queues the next step's recorded launches
    behind its own kernels (tnx_lanczos_run_multi_then) gives the same bits
    as queuing them after the call returns: energies, Ritz values of every
    local solve and the final MPS tensors, mixed precision schedule."""
    import torch
    from tnalg_amd import parameters as P
    from tnalg_amd.dmrg import _schedule_from
    from tnalg_amd.engine import Engine
    para = P.generate_parameters_dmrg('chain')
    para.update(l=24, chi=32, jxy=1.0, jz=0.8, hz=0.1, sweep_time=4, eigs_tol=1e-10, tau=1e-4,
                dt_ob=4, break_tol=0.0)
    para = P.make_consistent_parameter_dmrg(para)
    out = []
    for chain in (True, False):
        eng = Engine(precision_schedule=_schedule_from('mixed'), lowp=True)
        eng.chain_next = chain
        solves = eng.add_solves([(para, s) for s in (3, 4, 5)])
        for s in solves:
            s.schedule_canonicalise()
        eng.run_queues(solves)
        for t in range(para['sweep_time']):
            for s in solves:
                s.schedule_sweep(t)
            eng.run_queues(solves)
        print('chain_next', chain, 'steps', eng.counters['steps'], flush=True)
        e = eng.finalize(solves)
        torch.cuda.synchronize()
        out.append((np.array(e), [list(s.thetas) for s in solves],
                    [[s.mps[i][:s.dims[i] * s.d * s.dims[i + 1]].cpu().numpy().copy()
                      for i in range(para['l'])] for s in solves]))
    (e1, th1, m1), (e2, th2, m2) = out
    assert np.array_equal(e1, e2)
    assert th1 == th2
    for a, b in zip(m1, m2):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
