"""Exact diagonalisation of coupling-list Hamiltonians (oracle; test use only).

Builds ``H = sum_rows c op[a]_i op[b]_j + sum_rows c op[k]_i`` in the
``index1/index2`` convention of ``library/Parameters.py`` as a sparse matrix
on ``2^N`` states (site 0 is the most significant factor, matching the MPS
contraction order) and returns its ground state.
"""
import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import eigsh


def _site_op(op, i, n):
    left = sp.identity(2 ** i, format='csr')
    right = sp.identity(2 ** (n - i - 1), format='csr')
    return sp.kron(sp.kron(left, sp.csr_matrix(op)), right, format='csr')


def hamiltonian(n, ops, index1, coeff1, index2, coeff2):
    """Sparse ``2^n x 2^n`` Hamiltonian (real part; the path uses real operators)."""
    h = sp.csr_matrix((2 ** n, 2 ** n))
    for r in range(index2.shape[0]):
        c = float(np.ravel(coeff2)[r])
        if c == 0:
            continue
        i, j, a, b = (int(x) for x in index2[r])
        h = h + c * (_site_op(np.real(ops[a]), i, n) @ _site_op(np.real(ops[b]), j, n))
    for r in range(index1.shape[0]):
        c = float(np.ravel(coeff1)[r])
        i, k = (int(x) for x in index1[r])
        if c == 0 or np.linalg.norm(ops[k]) == 0:
            continue
        h = h + c * _site_op(np.real(ops[k]), i, n)
    return h


def ground_state(n, ops, index1, coeff1, index2, coeff2):
    """``(E0, psi0)``; dense ``eigh`` for ``n <= 8``, ``eigsh`` above."""
    h = hamiltonian(n, ops, index1, coeff1, index2, coeff2)
    if n <= 8:
        w, v = np.linalg.eigh(h.toarray())
        return float(w[0]), v[:, 0]
    w, v = eigsh(h, k=1, which='SA', tol=1e-14, ncv=40)
    return float(w[0]), v[:, 0]


def xx_chain_energy(n, j=1.0):
    """Closed form: open XX chain ``H = j sum (SxSx + SySy)`` maps to free
    fermions with hopping ``j/2``; E0 = sum of negative single-particle
    energies ``j cos(k pi/(n+1))``, k = 1..n."""
    eps = j * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    return float(np.sum(eps[eps < 0]))
