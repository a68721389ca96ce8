"""Two-site sweeps on the GPU (``para['dmrg_sites'] = 2``): exact ground
states where chi holds the whole state, bond growth from a small chi_init,
canonical form after truncation, batch/alone bit-identity, and agreement with
the one-site engine.  The reference's finite-size driver is one-site only, so
these results are pinned by exact diagonalisation (``oracle/ed.py``) and the
one-site results, not by a reference run."""
import numpy as np
import pytest

from oracle import ed, model

pytestmark = pytest.mark.gpu


def _para(couplings, chi, sweeps, **kw):
    from tnalg_amd import parameters as P
    para = P.common_parameters_dmrg()
    para.update(couplings)
    para.update(chi=chi, sweep_time=sweeps, tau=1e-4, eigs_tol=1e-10, break_tol=1e-14,
                dt_ob=1, spin='half', d=2, lattice='arbitrary', dmrg_sites=2)
    para.update(kw)
    para['nh'] = para['index2'].shape[0]
    return para


def _exact(c, n):
    e0, _ = ed.ground_state(n, c['op'], c['index1'], c['coeff1'], c['index2'], c['coeff2'])
    return e0


def _canonical_errors(A):
    """Largest deviation from the identity of the left (sites < centre) and
    right (sites > centre) orthonormality conditions."""
    err = 0.0
    for p, a in enumerate(A.mps):
        Dl, d, Dr = a.shape
        if p < A.center:
            m = a.reshape(Dl * d, Dr)
            err = max(err, np.abs(m.T @ m - np.eye(Dr)).max())
        elif p > A.center:
            m = a.reshape(Dl, d * Dr)
            err = max(err, np.abs(m @ m.T - np.eye(Dl)).max())
    return err


def test_two_site_exact_chain(gpu):
    """N=10 XXZ chain with a field, chi=32 (holds the whole state): the
    two-site energy is the exact ground-state energy."""
    from tnalg_amd import dmrg_finite_size
    c = model.chain_couplings(10, 1.0, 0.8, hz=0.15)
    para = _para(c, 32, 4)
    print('case: two-site N=10 chi=32 exact', flush=True)
    ob, A, info, _ = dmrg_finite_size(para, seed=1)
    e0 = _exact(c, 10)
    e = float(ob['e_per_site'][0]) * 10
    assert abs(e - e0) < 1e-10 * abs(e0), (e, e0)
    assert abs(A.norm_mps() - 1) < 1e-10
    assert _canonical_errors(A) < 1e-10


def test_two_site_grows_bonds_from_chi_init(gpu):
    """Starting from a chi_init = 2 product-like MPS, the SVD splits grow the
    bonds to min(2^b, 2^(N-b), chi) and reach the exact energy (N=12,
    chi=64, a long-range J1-J2 snake)."""
    from tnalg_amd import dmrg_finite_size, parameters as P
    w, h = 3, 4
    p1, p2 = P.positions_snake_j1j2(w, h)
    bonds = [tuple(x) for x in p1] + [tuple(x) for x in p2]
    jj = [1.0] * len(p1) + [0.5] * len(p2)
    c = model.couplings_from_bonds(12, bonds, jj, jj)
    para = _para(c, 64, 5, chi_init=2)
    print('case: two-site N=12 J1-J2 chi_init=2 -> 64', flush=True)
    ob, A, info, _ = dmrg_finite_size(para, seed=2)
    e0 = _exact(c, 12)
    e = float(ob['e_per_site'][0]) * 12
    assert abs(e - e0) < 1e-9 * abs(e0), (e, e0)
    dims = A.virtual_dim
    assert dims == [min(2 ** b, 2 ** (12 - b), 64) for b in range(13)], dims
    assert _canonical_errors(A) < 1e-10


def test_two_site_truncation_keeps_canonical_form(gpu):
    """N=24 Heisenberg chain at chi=12: the middle bonds are truncated to
    chi, the MPS stays canonical and normalised, and the energy agrees with
    the one-site result at the same chi within the truncation error."""
    from tnalg_amd import dmrg_finite_size
    c = model.chain_couplings(24, 1.0, 1.0)
    para = _para(c, 12, 6)
    print('case: two-site N=24 chi=12 truncated', flush=True)
    ob, A, info, _ = dmrg_finite_size(para, seed=3)
    e2 = float(ob['e_per_site'][0]) * 24
    dims = A.virtual_dim
    assert max(dims) == 12 and dims[12] == 12
    assert _canonical_errors(A) < 1e-10
    assert abs(A.norm_mps() - 1) < 1e-10
    para1 = dict(para, dmrg_sites=1)
    print('case: one-site N=24 chi=12', flush=True)
    ob1, _, _, _ = dmrg_finite_size(para1, seed=3)
    e1 = float(ob1['e_per_site'][0]) * 24
    # both are variational at chi=12; they differ by the truncation error
    assert abs(e2 - e1) < 1e-4 * abs(e1), (e2, e1)


def test_two_site_batch_equals_alone(gpu):
    """Two-site and one-site solves of different lengths in one batch give
    the same bits as each solve alone."""
    from tnalg_amd.dmrg import _schedule_from
    from tnalg_amd.engine import Engine

    def run(paras, seeds):
        eng = Engine(precision_schedule=_schedule_from('mixed'), lowp=True, plan_ahead=False)
        ss = [eng.add_solve(p, seed=s) for p, s in zip(paras, seeds)]
        for s, p in zip(ss, paras):
            s.schedule_canonicalise()
            for t in range(p['sweep_time']):
                s.schedule_sweep(t)
        eng.run_queues(ss)
        return eng.finalize(ss)

    paras = [_para(model.chain_couplings(12, 1.0, 1.0), 16, 3),
             _para(model.chain_couplings(9, 1.0, 0.6, hz=0.2), 8, 3, chi_init=2),
             _para(model.chain_couplings(10, 1.0, 1.0), 16, 3, dmrg_sites=1)]
    print('case: two-site batch of 3 vs alone', flush=True)
    together = run(paras, [5, 6, 7])
    for k in range(3):
        alone = run([paras[k]], [5 + k])
        assert alone[0] == together[k], (k, alone[0], together[k])


def test_two_site_mixed_precision_matches_fp64(gpu):
    """The mixed-precision sweep schedule ends within 1e-9 of the f64 two-site
    energy (N=32 XXZ, chi=32)."""
    from tnalg_amd import dmrg_finite_size
    c = model.chain_couplings(32, 1.0, 0.7, hz=0.1)
    para = _para(c, 32, 8)
    print('case: two-site N=32 chi=32 fp64 vs mixed', flush=True)
    ob64, _, _, _ = dmrg_finite_size(para, seed=4, precision='fp64')
    obm, _, _, _ = dmrg_finite_size(dict(para), seed=4, precision='mixed')
    e64, em = float(ob64['e_per_site'][0]), float(obm['e_per_site'][0])
    assert abs(em - e64) < 1e-9 * abs(e64), (em, e64)


def test_two_site_growth_then_one_site(gpu):
    """Three two-site sweeps grow the bonds from chi_init = 2 to their
    maxima (each pass about doubles them), the one-site sweeps after them
    finish at the exact energy (N=12 XXZ chain, chi=64)."""
    from tnalg_amd import dmrg_finite_size
    c = model.chain_couplings(12, 1.0, 0.9, hz=0.05)
    para = _para(c, 64, 6, chi_init=2, two_site_sweeps=3)
    print('case: two-site x3 then one-site, N=12 chi_init=2 -> 64', flush=True)
    ob, A, info, _ = dmrg_finite_size(para, seed=6)
    e0 = _exact(c, 12)
    e = float(ob['e_per_site'][0]) * 12
    assert abs(e - e0) < 1e-10 * abs(e0), (e, e0)
    assert A.virtual_dim == [min(2 ** b, 2 ** (12 - b), 64) for b in range(13)]
