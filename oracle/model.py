"""Spin operators and coupling lists, restated for the oracle (test use only).

Conventions follow ``library/HamiltonianModule.py`` and
``library/Parameters.py``: basis ``|0>=up, |1>=down``; operator list
``op = [id, sx, sy, sz, su, sd, -hx*sx - hz*sz]``; an ``index2`` row
``(i, j, a, b)`` with coefficient ``c`` is the term ``c * op[a]_i op[b]_j``; an
``index1`` row ``(i, k)`` with coefficient ``c`` is ``c * op[k]_i``.
"""
import numpy as np


def spin_half_operators():
    """``library/HamiltonianModule.py:8-24`` (spin 1/2)."""
    op = {
        'id': np.eye(2),
        'sx': np.array([[0.0, 0.5], [0.5, 0.0]]),
        'sy': np.array([[0.0, 0.5j], [-0.5j, 0.0]]),
        'sz': np.array([[0.5, 0.0], [0.0, -0.5]]),
        'su': np.array([[0.0, 1.0], [0.0, 0.0]]),
        'sd': np.array([[0.0, 0.0], [1.0, 0.0]]),
    }
    return op


def operator_list(hx=0.0, hz=0.0):
    """``library/Parameters.py:324-328``: ``[id, sx, sy, sz, su, sd, field]``."""
    o = spin_half_operators()
    return [o['id'], o['sx'], o['sy'], o['sz'], o['su'], o['sd'],
            -hx * o['sx'] - hz * o['sz']]


def heisenberg_rows(positions):
    """Bond list -> ``index2`` rows ``(i,j,4,5),(i,j,5,4),(i,j,3,3)``.

    ``library/HamiltonianModule.py:351-370``.
    """
    positions = np.asarray(positions, dtype=int).reshape(-1, 2)
    rows = []
    for i, j in positions:
        rows += [(i, j, 4, 5), (i, j, 5, 4), (i, j, 3, 3)]
    return np.array(rows, dtype=int).reshape(-1, 4)


def chain_couplings(length, jxy, jz, hx=0.0, hz=0.0, convention='parameters'):
    """Open-chain coupling lists.

    ``convention='parameters'``: ``coeff2 = (jxy/2, jxy/2, jz)`` per bond
    (``library/Parameters.py:199-211``), i.e. ``H = sum jxy(SxSx+SySy) + jz SzSz``.
    ``convention='easystart'``: ``(jxy, jxy, jz)`` (``EasyStart_DMRG.py:142-150``),
    the convention of the reference's saved runs.
    Returns ``dict(index1, coeff1, index2, coeff2, positions_h2, op)``.
    """
    pos = np.array([(n, n + 1) for n in range(length - 1)], dtype=int)
    index2 = heisenberg_rows(pos)
    f = 0.5 if convention == 'parameters' else 1.0
    coeff2 = np.tile(np.array([f * jxy, f * jxy, jz], dtype=float), len(pos)).reshape(-1, 1)
    index1 = np.stack([np.arange(length), np.full(length, 6)], axis=1).astype(int)
    coeff1 = np.ones((length, 1))
    return dict(index1=index1, coeff1=coeff1, index2=index2, coeff2=coeff2,
                positions_h2=pos, op=operator_list(hx, hz), l=length, hx=hx, hz=hz)


def couplings_from_bonds(length, bonds, jxy_per_bond, jz_per_bond, hx=0.0, hz=0.0):
    """Arbitrary bond list with per-bond couplings (Parameters convention)."""
    pos = np.asarray(bonds, dtype=int).reshape(-1, 2)
    index2 = heisenberg_rows(pos)
    jxy = np.broadcast_to(np.asarray(jxy_per_bond, dtype=float), (len(pos),))
    jz = np.broadcast_to(np.asarray(jz_per_bond, dtype=float), (len(pos),))
    coeff2 = np.stack([jxy / 2, jxy / 2, jz], axis=1).reshape(-1, 1)
    index1 = np.stack([np.arange(length), np.full(length, 6)], axis=1).astype(int)
    coeff1 = np.ones((length, 1))
    return dict(index1=index1, coeff1=coeff1, index2=index2, coeff2=coeff2,
                positions_h2=pos, op=operator_list(hx, hz), l=length, hx=hx, hz=hz)
