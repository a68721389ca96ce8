"""Tensor primitives of the DMRG path, restated in numpy (oracle; test use only).

MPS tensors are ``(Dl, d, Dr)`` arrays; boundary ("environment") matrices are
indexed ``[bra, ket]``.  Every function cites the reference routine whose
result it reproduces.
"""
import numpy as np


def random_open_mps(length, d, chi, rng):
    """Random open-boundary MPS in the reference's draw order.

    ``library/TensorBasicModule.py:179-211`` (``is_eco=False``): the first
    tensor ``(1, d, chi)`` is drawn first, then the last ``(chi, d, 1)``, then
    the bulk ``(chi, d, chi)`` in site order.  ``rng`` is a
    ``numpy.random.RandomState`` (the reference uses the unseeded global one).
    Returns ``(mps, virtual_dim)``; ``virtual_dim`` has ``length + 1`` entries.
    """
    mps = [None] * length
    mps[0] = rng.randn(1, d, chi)
    mps[length - 1] = rng.randn(chi, d, 1)
    for n in range(1, length - 1):
        mps[n] = rng.randn(chi, d, chi)
    dims = [1] + [chi] * (length - 1) + [1]
    return mps, dims


def left2right_decompose_tensor(tensor, way='qr'):
    """Left-orthonormalise ``tensor`` on its third bond.

    ``library/TensorBasicModule.py:416-450``: reshape to ``(Dl*d, Dr)``, reduced
    QR (or SVD), keep ``dim = min(Dl*d, Dr)`` columns.  Returns
    ``(isometry (Dl, d, dim), mat (Dr, dim), dim, lm)`` where the caller absorbs
    ``mat`` into the next tensor's first bond (``tensor_next <- mat^T @ ...``).
    """
    dl, d, dr = tensor.shape
    dim = min(dl * d, dr)
    m = tensor.reshape(dl * d, dr)
    if way == 'svd':
        u, lm, vt = np.linalg.svd(m, full_matrices=False)
        q = u
        r = np.diag(lm[:dim]) @ vt[:dim, :]
    else:
        q, r = np.linalg.qr(m)
        lm = np.zeros(0)
    return q[:, :dim].reshape(dl, d, dim), r.T, dim, lm


def right2left_decompose_tensor(tensor, way='qr'):
    """Right-orthonormalise ``tensor`` on its first bond.

    ``library/TensorBasicModule.py:453-486``: decompose ``tensor^T`` reshaped to
    ``(d*Dr, Dl)``; keep ``dim = min(Dl, d*Dr)``.  Returns
    ``(isometry (dim, d, Dr), mat (Dl, dim), dim, lm)``; the caller absorbs
    ``mat`` into the previous tensor's last bond.
    """
    dl, d, dr = tensor.shape
    dim = min(dl, d * dr)
    m = tensor.reshape(dl, d * dr).T
    if way == 'svd':
        u, lm, vt = np.linalg.svd(m, full_matrices=False)
        q = u
        r = np.diag(lm[:dim]) @ vt[:dim, :]
    else:
        q, r = np.linalg.qr(m)
        lm = np.zeros(0)
    return q[:, :dim].T.reshape(dim, d, dr), r.T, dim, lm


def absorb_matrix2tensor(tensor, mat, bond):
    """Contract ``mat``'s FIRST index with ``tensor``'s index ``bond``.

    ``library/TensorBasicModule.py:489-526``; the new index takes the place of
    the old one: ``out[..., j, ...] = sum_i tensor[..., i, ...] mat[i, j]``.
    """
    out = np.tensordot(tensor, mat, axes=([bond], [0]))
    return np.moveaxis(out, -1, bond)


def absorb_matrices2tensor(tensor, mats, bonds=None):
    """Sequentially absorb ``mats[i]`` on ``bonds[i]`` (first matrix index).

    ``library/TensorBasicModule.py:529-613`` (both the ``full_fast`` path used
    when every bond gets a matrix and the sequential path give the same
    result).
    """
    if bonds is None:
        bonds = range(len(mats))
    for m, b in zip(mats, bonds):
        tensor = absorb_matrix2tensor(tensor, m, b)
    return tensor


def bound_vec_operator_left2right(tensor, op=None, v=None):
    """Left boundary transfer ``v' = sum A*[x,s,b] v[x,a] O[s,t] A[a,t,c]``.

    ``library/TensorBasicModule.py:632-675``.  ``v=None`` means identity,
    ``op=None`` means identity.  Output ``(Dr, Dr)`` indexed ``[bra, ket]``.
    """
    ket = tensor if op is None else np.einsum('st,atc->asc', op, tensor)
    if v is not None:
        ket = np.tensordot(v, ket, axes=([1], [0]))
    return np.tensordot(tensor.conj(), ket, axes=([0, 1], [0, 1]))


def bound_vec_operator_right2left(tensor, op=None, v=None):
    """Right boundary transfer ``v' = sum A*[a,s,b] v[b,f] O[s,t] A[c,t,f]``.

    ``library/TensorBasicModule.py:678-721``.  Output ``(Dl, Dl)`` indexed
    ``[bra, ket]``.
    """
    ket = tensor if op is None else np.einsum('st,atc->asc', op, tensor)
    bra = tensor.conj()
    if v is not None:
        bra = np.tensordot(bra, v, axes=([2], [0]))
    return np.tensordot(bra, ket, axes=([1, 2], [1, 2]))


def transfer_matrix_mps(tensor):
    """``TM[(a,a'),(b,b')] = sum_s A*[a,s,b] A[a',s,b']`` (TBM:753-775)."""
    dl, d, dr = tensor.shape
    tm = np.einsum('asb,csd->acbd', tensor.conj(), tensor)
    return tm.reshape(dl * dl, dr * dr)


def entanglement_entropy(lm, tol=1e-20):
    """``-2 sum lam^2 ln lam`` over normalised ``lam > tol`` (TBM:888-904)."""
    lm = np.sort(np.asarray(lm, dtype=float).reshape(-1))[::-1]
    lm = lm[lm > tol]
    lm = lm / np.linalg.norm(lm)
    return float(-2.0 * np.dot(lm ** 2, np.log(lm)))
