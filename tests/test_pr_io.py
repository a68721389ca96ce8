"""``.pr`` result files (no GPU): ``save_pr``/``load_pr`` follow
``library/BasicFunctions.py:29-95``, the result MPS pickles as host data, and
the loader refuses any type outside its allow-list."""
import os
import pickle

import numpy as np
import pytest

from tnalg_amd import load_pr, save_pr
from tnalg_amd.dmrg import MpsOpenBoundaryClass


def _host_mps():
    """An ``A`` as a loaded file holds it: host data, no device solve."""
    rng = np.random.RandomState(0)
    A = MpsOpenBoundaryClass.__new__(MpsOpenBoundaryClass)
    dims = [1, 2, 4, 2, 1]
    A.__dict__.update(
        _solve=None, _engine=None, _para={'l': 4, 'chi': 4, 'index2': np.array([[0, 1, 3, 3]]),
                                          'op': [np.eye(2), np.array([[0, .5j], [-.5j, 0]])]},
        spin='half', decomp_way='qr', length=4, phys_dim=2,
        mps=[rng.randn(dims[i], 2, dims[i + 1]) for i in range(4)], virtual_dim=dims, center=1,
        orthogonality=np.array([[-1.], [0.], [1.], [1.]]), operators=[np.eye(2)],
        lm=[np.array([1.0]), np.array([0.8, 0.6]), np.zeros(0)], ent=np.array([[0.], [.3], [-1.]]))
    return A


@pytest.mark.filterwarnings("ignore::PendingDeprecationWarning")
def test_save_load_round_trip(tmp_path):
    ob = {'eb_full': np.arange(6.0).reshape(-1, 1), 'e_per_site': np.array([-0.44]),
          'corr_x': np.zeros(0), 'mx': np.matrix([[1.0], [2.0]])}
    info = {'convergence': np.float64(1e-12), 't_cost': 1.5, 'stats': {'steps': 7},
            'energies': [-1.0, -2.0]}
    para = {'l': np.int64(4), 'chi': 4, 'lattice': 'chain', 'index2': np.array([[0, 1, 4, 5]]),
            'op': [np.eye(2), np.array([[0, .5j], [-.5j, 0]])], 'tags': {1, 2}, 'c': 1 + 2j}
    A = _host_mps()
    save_pr(str(tmp_path / 'sub'), 'run.pr', (ob, A, info, para), ('ob', 'A', 'info', 'para'))
    got = load_pr(str(tmp_path / 'sub' / 'run.pr'))
    assert sorted(got) == ['A', 'info', 'ob', 'para']
    assert np.array_equal(got['ob']['eb_full'], ob['eb_full'])
    assert isinstance(got['ob']['mx'], np.matrix)
    assert got['info'] == info and got['para']['lattice'] == 'chain' and got['para']['tags'] == {1, 2}
    assert got['para']['c'] == 1 + 2j and np.array_equal(got['para']['op'][1], para['op'][1])
    B = got['A']
    assert isinstance(B, MpsOpenBoundaryClass) and B._solve is None and B.center == 1
    assert all(np.array_equal(a, b) for a, b in zip(A.mps, B.mps))
    assert [x.tolist() for x in B.lm] == [x.tolist() for x in A.lm]
    assert np.array_equal(B._para['index2'], A._para['index2'])
    # one name / several names / no file (BasicFunctions.py:80-92)
    p = str(tmp_path / 'sub' / 'run.pr')
    assert load_pr(p, 'info') == info
    a2, ob2 = load_pr(p, ['A', 'ob'])
    assert isinstance(a2, MpsOpenBoundaryClass) and np.array_equal(ob2['eb_full'], ob['eb_full'])
    assert load_pr(str(tmp_path / 'missing.pr')) is False
    B.calculate_entanglement_entropy()          # host only: -1 where no spectrum is known
    assert B.ent[2, 0] == -1 and abs(B.ent[0, 0]) < 1e-15
    lm = np.array([0.8, 0.6])
    assert abs(B.ent[1, 0] + 2 * (lm ** 2 @ np.log(lm))) < 1e-15


class _Runs:
    def __reduce__(self):
        return (os.system, ('echo should-not-run',))


def test_loader_refuses_other_types(tmp_path):
    p = tmp_path / 'bad.pr'
    p.write_bytes(pickle.dumps({'x': _Runs()}))     # dumping runs nothing
    with pytest.raises(pickle.UnpicklingError):
        load_pr(str(p))
    with pytest.raises(ValueError):
        save_pr(str(tmp_path), 'x.pr', (1, 2), ('a',))
