"""Host logic of the device measurement (no GPU): the string slots of a
measurement plan are shared only by strings whose lives do not overlap."""
import numpy as np

from tnalg_amd.observe import _MeasurePlan


class _FakeSolve:
    def __init__(self, n, chi):
        self.n = n
        self.dmax = [min(2 ** b, 2 ** (n - b), chi) for b in range(n + 1)]
        self.psi_max = 2 * chi * chi
        self.dev = 'cpu'


def _check_slots(P):
    by_slot = {}
    for c, t in P.cidx.items():
        by_slot.setdefault(t, []).append((c[0], P.chain_end[c]))
    for t, lives in by_slot.items():
        lives.sort()
        for (s0, e0), (s1, e1) in zip(lives, lives[1:]):
            # a string may start where the previous one is last read
            assert s1 >= e0, (t, lives)
    return len(by_slot)


def test_bond_rows_use_three_slots():
    n = 40
    rows = [(i, i + 1, a, b) for i in range(n - 1) for a, b in ((4, 5), (5, 4), (3, 3))]
    P = _MeasurePlan(_FakeSolve(n, 16), [(i, 3) for i in range(n)], rows)
    assert len(P.chains) == 3 * (n - 1)
    assert _check_slots(P) == 3
    assert P.X[0].numel() == 3 * P.dsq


def test_long_strings_keep_their_slots():
    n = 24
    rows = [(i, j, 1, 1) for i in range(n) for j in range(i + 1, n)]   # all pairs
    P = _MeasurePlan(_FakeSolve(n, 8), [], rows)
    nslot = _check_slots(P)
    # string (i, 1) lives from i to n-1: every one overlaps every other
    assert nslot == n - 1
    rng = np.random.RandomState(0)
    mid = [(int(i), int(j), 3, 3) for i, j in rng.randint(0, n, size=(30, 2)) if i != j]
    _check_slots(_MeasurePlan(_FakeSolve(n, 8), [], mid))
