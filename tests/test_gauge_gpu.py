"""Gauge moves on the GPU (Householder for wide panels, CholeskyQR3 for tall
ones, blocked CholeskyQR3 above 128 columns) leave the physical state
unchanged and make the tensors isometries."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _chain(n, chi):
    from tnalg_amd import parameters as P
    para = P.parameter_dmrg_chain()
    para.update(P.common_parameters_dmrg())
    para.update(l=n, chi=chi, jxy=1.0, jz=0.8, hz=0.1, sweep_time=2, eigs_tol=1e-10)
    return P.make_consistent_parameter_dmrg(para)


@pytest.mark.parametrize('n,chi', [(10, 16), (16, 64), (20, 256)])
def test_canonicalisation_preserves_state(gpu, n, chi):
    from tnalg_amd.engine import Engine
    from tnalg_amd.observe import measure
    para = _chain(n, chi)
    eng = Engine(lowp=False)
    s = eng.add_solve(para, seed=7)
    rows = [tuple(r) for r in para['index2']]
    print('canonicalise N', n, 'chi', chi, flush=True)
    (_, v0, norm0), = measure(eng, [s], [([], rows)])
    s.schedule_canonicalise()
    eng.run_queues([s])
    (_, v1, norm1), = measure(eng, [s], [([], rows)])
    assert np.allclose(v0, v1, rtol=1e-9, atol=1e-12), np.abs(v0 - v1).max()
    A = s.host_mps()
    for k in range(1, n):   # right-canonical: sum_{s,b} A[a,s,b] A[c,s,b] = delta
        m = A[k].reshape(A[k].shape[0], -1)
        assert np.allclose(m @ m.T, np.eye(m.shape[0]), atol=1e-10), k
    assert abs(norm1 - np.linalg.norm(A[0]) ** 2) < 1e-8 * norm1


def test_large_bond_sweep_is_variational(gpu):
    """chi=256 (blocked CholeskyQR3 in the middle of the chain): two f64
    sweeps give an energy at or below the chi=64 result."""
    from tnalg_amd import dmrg_finite_size
    print('chi 256 sweep', flush=True)
    p64 = _chain(20, 64)
    p256 = _chain(20, 256)
    ob64, _, _, _ = dmrg_finite_size(p64, seed=1)
    ob256, A, _, _ = dmrg_finite_size(p256, seed=1)
    e64, e256 = float(ob64['e_per_site'][0]), float(ob256['e_per_site'][0])
    assert e256 <= e64 + 1e-10, (e256, e64)
    assert max(A.virtual_dim) == 256
