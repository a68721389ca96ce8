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


# ---------------------------------------------------------------------------
# Brute-force expectation values on a dense state vector (site 0 is the most
# significant factor, as above).  These fix, by direct summation over the
# 2^N amplitudes, what the reference's transfer-chain observables compute
# (library/MPSClass.py:1105-1212): <O_i>, <O_i O'_j> and their combinations.

def apply_site_op(op, i, psi, n, d=2):
    """``(1 x .. x op_i x .. x 1) psi`` by reshaping, for any (complex) op."""
    v = np.asarray(psi).reshape(d ** i, d, d ** (n - i - 1))
    return np.einsum('st,atb->asb', np.asarray(op), v).reshape(-1)


def expect_one(psi, op, i, n):
    """``<psi|op_i|psi> / <psi|psi>`` (real part)."""
    psi = np.asarray(psi)
    nrm = np.vdot(psi, psi)
    return float(np.real(np.vdot(psi, apply_site_op(op, i, psi, n)) / nrm))


def expect_two(psi, op_a, i, op_b, j, n):
    """``<psi|op_a(i) op_b(j)|psi> / <psi|psi>`` (real part), i != j."""
    psi = np.asarray(psi)
    nrm = np.vdot(psi, psi)
    w = apply_site_op(op_b, j, psi, n)
    w = apply_site_op(op_a, i, w, n)
    return float(np.real(np.vdot(psi, w) / nrm))


def spin_dot(psi, i, j, n, ops):
    """``<S_i . S_j>`` with the complex ``sy`` (``ops`` = the operator list)."""
    return sum(expect_two(psi, ops[k], i, ops[k], j, n) for k in (1, 2, 3))
