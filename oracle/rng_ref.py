"""Counter-based normal stream of the device initial-MPS draw, restated in
Python (oracle; test use only).

Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
as 1, 2, 3", SC 2011): ten rounds of two 32 x 32 -> 64-bit multiplies
(0xD2511F53, 0xCD9E8D57) on a 128-bit counter under a 64-bit key bumped by
the Weyl constants (0x9E3779B9, 0xBB67AE85).  Pinned by the published
known-answer vectors of the algorithm.  Number g of the stream of ``seed``:
pair q = g // 2 is Philox of counter (q lo, q hi, 0, 0) under the key (seed
lo, seed hi); words 0-1 and 2-3 give two 53-bit uniforms in (0, 1); the
Box-Muller transform gives (r cos, r sin) for even and odd g.
"""
import numpy as np

_M0, _M1, _W0, _W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
_MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    c = [int(x) & _MASK for x in ctr]
    k0, k1 = int(key[0]) & _MASK, int(key[1]) & _MASK
    for _ in range(10):
        p0, p1 = _M0 * c[0], _M1 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k0) & _MASK, p1 & _MASK,
             ((p0 >> 32) ^ c[3] ^ k1) & _MASK, p0 & _MASK]
        k0, k1 = (k0 + _W0) & _MASK, (k1 + _W1) & _MASK
    return c


def normal(seed, g):
    """Number ``g`` of the normal stream of ``seed``."""
    q = g >> 1
    c = philox4x32_10((q & _MASK, q >> 32, 0, 0), (seed & _MASK, (seed >> 32) & _MASK))
    a = (c[0] << 32) | c[1]
    b = (c[2] << 32) | c[3]
    u1 = ((a >> 11) + 0.5) * 2.0 ** -53
    u2 = ((b >> 11) + 0.5) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    return float(r * (np.sin(2 * np.pi * u2) if g & 1 else np.cos(2 * np.pi * u2)))
