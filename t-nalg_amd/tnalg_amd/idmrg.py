"""Infinite-size DMRG on the device kernels: the drop-in for
``dmrg_infinite_size(para, A, hamilt)`` (``algorithms/DMRG_anyH.py:106-188``)
with the infinite MPS of ``MpsInfinite`` (``library/MPSClass.py:1550-1891``),
form 'center_ort', two-site centre, ``dmrg_type`` 'mpo' or 'white'.

The centre psi (chi, D, D, chi) is held as the square matrix
Psi[(a x1), (x2 b)] of order n = chi D.  Every effective Hamiltonian of the
reference is a sum of Kronecker products, Psi -> sum_r c_r L_r Psi R_r^T,
with blocks of order n built by ``tnx_kron_f64``:

* 'mpo' (``MPSClass.py:1831-1845``): h = env_L . T . T . env_R with the
  cell tensor T of ``hamiltonian2cell_tensor``; with
  L_n = sum_m envL_m (x) T_mn and R_n = sum_k T_nk (x) envR_k (n over the
  MPO bond), h = sum_n L_n (x) R_n.  The reference symmetrises h and takes
  its largest-magnitude eigenvector (``eigsh``, which='LM'); the environment
  blocks are symmetric (every update symmetrises them), so the symmetrised
  h is (1/2) sum_n (L_n (x) R_n + L_n^T (x) R_n^T), and its dominant
  eigenvector is the lowest one of -h, which the Lanczos kernels find.
* 'white' (``MPSClass.py:1709-1727``): H = L_H (x) 1 + 1 (x) R_H +
  sum_pairs j (1 (x) op1) (x) (op2 (x) 1) with L_H = bath (x) 1 +
  sum j eff_i1 (x) op_i2 (+ the pair terms whose second operator is the
  identity) and R_H its mirror; the reference's largest-magnitude
  eigenvector of I - tau H is the lowest one of H, with ARPACK's
  convergence test on I - tau H.

So the matvec is the finite path's GEMM 1 (L_r Psi) and K-concatenated
GEMM 2 (sum T_r R_r^T), no mixer.  The same blocks update the environments
after the centre is split by the one-sided Jacobi SVD (``MPSClass.py:1680-1690``):
envL'_k = A^T L_k A and envR'_n = B R_n B^T with A the kept left singular
vectors and B the kept right ones (``update_left_env`` /
``update_right_env``, ``MPSClass.py:1652-1664``), then normalised
(``tnx_normalize_f64``) and symmetrised (``tnx_symmetrize_f64``); 'white'
updates bath' = sym(A^T K A), K = bath (x) 1 + sum j eff_i1 (x) op_i2, and
eff'_n = A^T (1 (x) op_n) A (``MPSClass.py:1737-1763``).  The reduced
density matrices of the observations are GEMMs over the centre
(``MPSClass.py:1872-1886``); the host closes tr(rho h).

All arithmetic is f64.  Chains with the same structure (type, chi,
symmetric environments) run as one batch: every launch covers all of them,
and a chain that passes its energy test leaves the batch.

Start vectors: the reference restarts each eigensolve from the previous
centre.  With symmetric environments the ground state of the new problem
often lies in the other reflection sector, so that start vector overlaps it
only at the rounding level (1e-12) and ARPACK reaches it by amplifying that
rounding; here the start vector is the previous centre plus 1e-2 times a
fixed random unit vector, and the Krylov test begins after ncv steps (as
ARPACK completes its first factorisation), so the solve reaches the same
(unique) ground state without depending on rounding.

Differences from the reference: only n_site = 2 with spin-1/2 operators is
built (the reference's n_site = 1 'mpo' path contracts an environment of
bond d with a cell tensor of bond d^2, and its n_site = 1 'white' path
reshapes the centre to two sites; neither runs); a centre with a zero
singular value among the kept ones leaves that row of mps[2] zero (numpy's
SVD completes it with an orthonormal vector).
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib as L
from . import records as R
from .device import ctypes_void, require_gpu
from .parameters import (generate_parameters_infinite_dmrg, hamiltonian2cell_tensor,
                         hamiltonian_heisenberg, spin_operators)

NCV = 20
MAX_RESTARTS = 60
# start vector of each eigensolve: the previous centre plus KICK times a
# fixed unit random vector (see InfiniteRun._build)
KICK = 1e-2
KICK_SEED = 20180913
# 'mpo' eigensolves: residual estimate <= MPO_FLOOR |theta| (ARPACK's tol = 0
# test is eps |theta|; the estimate |beta_j y_j| of the device tridiagonal
# solve bottoms out a few eps above that)
MPO_FLOOR = 1e-15
EPS = 2.220446049250313e-16
SVD_SWEEPS = 60
BLK = L.KRON_BLOCK


def _real_operators():
    """[id, sx, sy, sz, su, sd] (``MPSClass.py:1597-1599``) as real 2 x 2
    matrices; sy is stored divided by i (it is imaginary)."""
    o = spin_operators('half')
    return [np.real(o['id']), np.real(o['sx']), np.imag(o['sy']), np.real(o['sz']),
            np.real(o['su']), np.real(o['sd'])]


class MpsInfinite:
    """``MpsInfinite`` (``MPSClass.py:1550-1612``) with form 'center_ort' and
    n_site = 2, held on the device.  ``d``: dimension of one centre index
    (d_spin^2 for 'mpo', d_spin for 'white'); ``D``: MPO bond of the
    environments.  The initial centre is drawn from ``rng`` (numpy's global
    generator by default, as the reference) on the host and uploaded.  The
    reference's attributes (``mps``, ``lm``, ``env``, ``rho``,
    ``bath_op_onsite``, ``effective_ops``) are host copies read from the
    device on access."""

    def __init__(self, form, d, chi, D, n_tensor=3, n_site=1, spin='half', dmrg_type='mpo',
                 way='qr', operators=None, hamilt_index=None, is_symme_env=True, is_real=True,
                 debug=False, rng=None, device='cuda'):
        if form != 'center_ort':
            raise ValueError("only form='center_ort' runs the iDMRG driver")
        if n_site != 2:
            raise ValueError('only n_site = 2 is built (the reference runs no other)')
        if spin != 'half' or operators is not None:
            raise ValueError('only the spin-1/2 operators of MpsInfinite are built')
        if dmrg_type not in ('mpo', 'white'):
            raise ValueError('dmrg_type must be mpo or white')
        self.form, self.d, self.chi, self.D = form, int(d), int(chi), int(D)
        self.n_tensor, self.n_site, self.spin = 3, 2, spin
        self.dmrg_type, self.decomp_way = dmrg_type, way
        self.is_symme_env, self.is_real, self._debug = bool(is_symme_env), is_real, debug
        self.hamilt_index = None if hamilt_index is None else np.asarray(hamilt_index, float)
        op = spin_operators('half')
        self.operators = [op['id'], op['sx'], op['sy'], op['sz'], op['su'], op['sd']]
        self.op_index = set(range(len(self.operators)))
        self.orthogonality = np.array([-1, 0, 1])
        self.is_center_ort, self.is_canonical = True, False
        if dmrg_type == 'mpo' and self.D != self.d:
            raise ValueError("'mpo' needs D = d (the cell tensor's MPO bond is d_spin^2)")
        self.device = torch.device(device)
        chi, dc = self.chi, self.d
        self.n = n = chi * dc
        f64 = dict(dtype=torch.float64, device=self.device)
        rng = np.random if rng is None else rng
        c = rng.randn(chi, dc, dc, chi)
        c = c / np.linalg.norm(c.reshape(-1))
        self.psi = torch.from_numpy(np.ascontiguousarray(c).reshape(-1)).to(**f64)
        N = n * n
        self.init = torch.empty(N, **f64)
        kick = np.random.RandomState(KICK_SEED).randn(N)
        self.kick = torch.from_numpy(kick / np.linalg.norm(kick)).to(**f64)
        self.V = torch.empty((NCV + 1) * N, **f64)
        self.w = torch.empty(N, **f64)
        self.x = torch.empty(N, **f64)
        self.u = torch.zeros(N, **f64)
        self.s = torch.zeros(n, **f64)
        self.svt = torch.zeros(N, **f64)
        self.svd_work = torch.empty(2 * N, **f64)
        self.rho_buf = torch.zeros(3 * 16, **f64)
        Dm = self.D
        if dmrg_type == 'mpo':
            self.envL = torch.ones(chi * Dm * chi, **f64)
            self.envR = self.envL if self.is_symme_env else torch.ones(chi * Dm * chi, **f64)
            self.Lb = torch.empty(Dm * N, **f64)
            self.Rb = torch.empty(Dm * N, **f64)
            self.T1 = torch.empty(2 * Dm * N, **f64)
            self.Y = torch.empty(2 * Dm * n * chi, **f64)
        else:
            nops = len(self.operators)
            self.bath = torch.eye(chi, **f64).reshape(-1)
            self.eff = torch.eye(chi, **f64).reshape(1, -1).repeat(nops, 1).reshape(-1)
            self.opt = torch.from_numpy(np.stack(_real_operators()).reshape(-1)).to(**f64)
            self.LH = torch.empty(N, **f64)
            self.RH = torch.empty(N, **f64)
            self.K = torch.empty(N, **f64)
            self.Pop = torch.empty(nops * N, **f64)   # 1 (x) op_n, every operator
            self.P2 = torch.empty(nops * N, **f64)    # op_n (x) 1
            self.T1 = torch.empty((1 + 9) * N, **f64)
            self.Y = torch.empty((1 + nops) * n * chi, **f64)
        self.rho = None

    # ------------------------------------------------------- host views
    def _get(self, t):
        return t.detach().cpu().numpy().copy()

    @property
    def mps(self):
        chi, dc = self.chi, self.d
        u = self._get(self.u).reshape(self.n, self.n)[:, :chi]
        vt = self._get(self.svt).reshape(self.n, self.n)[:chi]
        return [u.reshape(chi, dc, chi), self._get(self.psi).reshape(chi, dc, dc, chi),
                vt.reshape(chi, dc, chi).transpose(2, 1, 0)]

    @property
    def lm(self):
        return [self._get(self.s)[:self.chi], np.zeros(0), np.zeros(0)]

    @property
    def env(self):
        chi, Dm = self.chi, self.D
        if self.dmrg_type == 'mpo':
            e0 = self._get(self.envL).reshape(chi, Dm, chi)
            # the reference's update_left_env rebinds env[0], so with
            # symmetric environments env[1] keeps its initial value
            e1 = np.ones((chi, Dm, chi)) if self.is_symme_env else \
                self._get(self.envR).reshape(chi, Dm, chi)
            return [e0, e1]
        e = np.ones((chi, Dm, chi))
        return [e, e]

    @property
    def bath_op_onsite(self):
        if self.dmrg_type == 'mpo':
            return np.eye(self.chi)
        return self._get(self.bath).reshape(self.chi, self.chi)

    @property
    def effective_ops(self):
        chi = self.chi
        if self.dmrg_type == 'mpo':
            return [np.eye(chi) for _ in self.operators]
        e = self._get(self.eff).reshape(-1, chi, chi)
        return [e[k] if k != 2 else 1j * e[k] for k in range(len(self.operators))]


def _vp(t, off=0):
    return int(t.data_ptr()) + 8 * int(off)


class InfiniteRun:
    """Lockstep iDMRG iterations of chains with the same structure.

    ``chains``: MpsInfinite objects (same dmrg_type, chi, d, D and
    is_symme_env); ``paras``: their parameter dicts; ``hamilts``: their
    4 x 4 bond Hamiltonians.  ``lib`` / ``stream``: the C library and the
    stream handle getter (tests pass an emulation)."""

    def __init__(self, chains, paras, hamilts, lib=None, stream=None):
        a0 = chains[0]
        for a in chains:
            if (a.dmrg_type, a.chi, a.d, a.D, a.is_symme_env) != \
                    (a0.dmrg_type, a0.chi, a0.d, a0.D, a0.is_symme_env):
                raise ValueError('chains of one run need the same structure')
        self.chains, self.paras = list(chains), list(paras)
        self.hamilts = [np.asarray(h, dtype=float) for h in hamilts]
        self.lib = lib if lib is not None else L.load()
        self._stream = stream
        self.type = a0.dmrg_type
        self.device = a0.device
        self.keep = []
        self.state = None
        for a, p, h in zip(self.chains, self.paras, self.hamilts):
            if self.type == 'mpo':
                t = hamiltonian2cell_tensor(h, p['tau'])
                # T_mn = T[m, :, :, n] as contiguous d x d blocks [m][n]
                a.Tmn = torch.from_numpy(np.ascontiguousarray(t.transpose(0, 3, 1, 2)).reshape(
                    -1)).to(dtype=torch.float64, device=self.device)
            else:
                if a.hamilt_index is None:
                    raise ValueError("'white' needs hamilt_index")
                if np.any(a.hamilt_index[:, :2] == 2):
                    raise ValueError('rows with sy (imaginary) are not built')
        if self.type == 'white':
            self._build_constants()

    def stream(self):
        if self._stream is not None:
            return self._stream()
        return ctypes_void(torch.cuda.current_stream().cuda_stream)

    def _upload(self, arrays):
        out = []
        for a in arrays:
            buf = torch.from_numpy(np.frombuffer(a.tobytes(), dtype=np.uint8).copy()).to(
                self.device) if a.nbytes else torch.zeros(256, dtype=torch.uint8,
                                                          device=self.device)
            self.keep.append(buf)
            out.append(buf)
        return out

    # --------------------------------------------------------- descriptors
    def _kron_arrays(self, jobs):
        """``jobs``: list of (dst, ldd, p, q, terms) with terms
        (a, lda, b, ldb, coef) (a / b = 0: identity)."""
        jrec, trec, blocks = [], [], 0
        for dst, ldd, p, q, terms in jobs:
            jrec.append((dst, p, q, ldd, len(trec), len(terms), blocks))
            for (a, lda, b, ldb, c) in terms:
                trec.append((a, b, lda, ldb, c))
            blocks += -(-(p * q) ** 2 // BLK)
        return (np.array(jrec, dtype=L.KRON_JOB), np.array(trec, dtype=L.KRON_TERM), blocks)

    def _kron(self, arrs):
        jobs, terms, nb = arrs
        dj, dt = self._upload([jobs, terms])
        return (dj, len(jobs), dt, nb)

    def _run_kron(self, k):
        dj, nj, dt, nb = k
        L.check(self.lib.tnx_kron_f64(ctypes_void(dj.data_ptr()), nj, ctypes_void(dt.data_ptr()),
                                      nb, self.stream()), 'tnx_kron_f64')

    def _gemm(self, glist):
        p, s = glist.arrays()
        dp, ds = self._upload([p, s])
        return (dp, len(glist.probs), ds, R.gemm_tiles(p))

    def _run_gemm(self, g):
        dp, npb, ds, nt = g
        if npb:
            L.check(self.lib.tnx_gemm_grouped(L.GEMM_F64, ctypes_void(dp.data_ptr()), npb,
                                              ctypes_void(ds.data_ptr()), nt, None,
                                              self.stream()), 'tnx_gemm_grouped')

    def _build_constants(self):
        """'white': the fixed blocks 1 (x) op_n and op_n (x) 1 of every chain."""
        jobs = []
        for a in self.chains:
            chi, d, n = a.chi, a.d, a.n
            N = n * n
            for k in range(len(a.operators)):
                op = _vp(a.opt, 4 * k)
                jobs.append((_vp(a.Pop, k * N), n, chi, d, [(0, 0, op, d, 1.0)]))
                jobs.append((_vp(a.P2, k * N), n, d, chi, [(op, d, 0, 0, 1.0)]))
        self._run_kron(self._kron(self._kron_arrays(jobs)))

    def _build(self, active):
        """Descriptor arrays of the iteration for the chains ``active``."""
        st = {'active': list(active)}
        kj, kj_env = [], []
        g1, g2 = R.GemmList(), R.GemmList()
        e1, e2 = R.GemmList(), R.GemmList()
        rho = R.GemmList()
        lz, svd, nrm_rows, nrm_env, sym = [], [], [], [], []
        kick = R.MixList()
        for q, ci in enumerate(active):
            a = self.chains[ci]
            chi, d, n, Dm = a.chi, a.d, a.n, a.D
            N = n * n
            x, w = _vp(a.x), _vp(a.w)
            kick.add(_vp(a.init), L.F64, 1, 1, N, [(_vp(a.psi), L.F64, 1.0, -1),
                                                  (_vp(a.kick), L.F64, KICK, -1)])
            if self.type == 'mpo':
                envR = a.envL if a.is_symme_env else a.envR
                for nn in range(Dm):
                    kj.append((_vp(a.Lb, nn * N), n, chi, d,
                               [(_vp(a.envL, m * chi), Dm * chi,
                                 _vp(a.Tmn, (m * Dm + nn) * d * d), d, 1.0) for m in range(Dm)]))
                    kj.append((_vp(a.Rb, nn * N), n, d, chi,
                               [(_vp(a.Tmn, (nn * Dm + k) * d * d), d,
                                 _vp(envR, k * chi), Dm * chi, 1.0) for k in range(Dm)]))
                segs = []
                for nn in range(Dm):
                    Lp, Rp = _vp(a.Lb, nn * N), _vp(a.Rb, nn * N)
                    t0, t1 = _vp(a.T1, nn * N), _vp(a.T1, (Dm + nn) * N)
                    g1.add(t0, n, n, n, [(Lp, n, 0, x, n, 0, n, 1.0)], active=q)
                    g1.add(t1, n, n, n, [(Lp, n, 1, x, n, 0, n, 1.0)], active=q)
                    segs.append((t0, n, 0, Rp, n, 1, n, -0.5))
                    segs.append((t1, n, 0, Rp, n, 0, n, -0.5))
                g2.add(w, n, n, n, segs, active=q)
                lz.append((_vp(a.V), w, x, _vp(a.init), _vp(a.psi), N, N, min(NCV, N), 0,
                           0.0, 0.0, MPO_FLOOR, L.F64, L.F64))
                # environment updates: Y_k = L_k A, Z_n = R_n B^T; env = A^T Y, B Z
                for k in range(Dm):
                    yk = _vp(a.Y, k * n * chi)
                    e1.add(yk, n, chi, chi, [(_vp(a.Lb, k * N), n, 0, _vp(a.u), n, 0, n, 1.0)])
                    e2.add(_vp(a.envL, k * chi), chi, chi, Dm * chi,
                           [(_vp(a.u), n, 1, yk, chi, 0, n, 1.0)])
                    if not a.is_symme_env:
                        zk = _vp(a.Y, (Dm + k) * n * chi)
                        e1.add(zk, n, chi, chi, [(_vp(a.Rb, k * N), n, 0, _vp(a.svt), n, 1, n,
                                                  1.0)])
                        e2.add(_vp(a.envR, k * chi), chi, chi, Dm * chi,
                               [(_vp(a.svt), n, 0, zk, chi, 0, n, 1.0)])
                envs = [a.envL] if a.is_symme_env else [a.envL, a.envR]
                for e in envs:
                    nrm_env.append((_vp(e), chi * Dm * chi))
                    sym.append((_vp(e), chi, Dm * chi, Dm, chi))
                # reduced density matrices of the spin pairs (0,1), (1,2), (2,3)
                for j in range(3):
                    P, Q = chi * 2 ** j, (4 * chi) // 2 ** j
                    rho.add(_vp(a.rho_buf, 16 * j), 4, 4, 4,
                            [(_vp(a.psi, p * 4 * Q), Q, 0, _vp(a.psi, p * 4 * Q), Q, 1, Q, 1.0)
                             for p in range(P)])
            else:
                hi = a.hamilt_index
                op = lambda k: _vp(a.opt, 4 * int(k))
                eff = lambda k: _vp(a.eff, int(k) * chi * chi)
                bath = _vp(a.bath)
                lh = [(bath, chi, 0, 0, 1.0)]
                rh = [(0, 0, bath, chi, 1.0)]
                kk = [(bath, chi, 0, 0, 1.0)]
                pairs = []
                for i1, i2, j in hi:
                    lh.append((eff(i1), chi, op(i2), d, j))
                    kk.append((eff(i1), chi, op(i2), d, j))
                    rh.append((op(i2), d, eff(i1), chi, j))
                    if int(i2) == 0:
                        lh.append((0, 0, op(i1), d, j))
                    elif int(i1) == 0:
                        rh.append((op(i2), d, 0, 0, j))
                    else:
                        pairs.append((int(i1), int(i2), j))
                if len(pairs) > 9:
                    raise ValueError('at most 9 two-operator rows')
                kj.append((_vp(a.LH), n, chi, d, lh))
                kj.append((_vp(a.RH), n, d, chi, rh))
                kj.append((_vp(a.K), n, chi, d, kk))
                th = _vp(a.T1)
                g1.add(th, n, n, n, [(_vp(a.LH), n, 0, x, n, 0, n, 1.0)], active=q)
                segs = [(x, n, 0, _vp(a.RH), n, 1, n, 1.0)]
                for k, (i1, i2, j) in enumerate(pairs):
                    tk = _vp(a.T1, (1 + k) * N)
                    g1.add(tk, n, n, n, [(_vp(a.Pop, i1 * N), n, 0, x, n, 0, n, 1.0)], active=q)
                    segs.append((tk, n, 0, _vp(a.P2, i2 * N), n, 1, n, j))
                g2.add(w, n, n, n, segs, c0=th, ldc0=n, beta=1.0, active=q)
                tau = float(self.paras[ci]['tau'])
                lz.append((_vp(a.V), w, x, _vp(a.init), _vp(a.psi), N, N, min(NCV, N), 0,
                           EPS / tau, tau, 0.0, L.F64, L.F64))
                # bath' = A^T K A, eff'_n = A^T (1 (x) op_n) A
                yb = _vp(a.Y)
                e1.add(yb, n, chi, chi, [(_vp(a.K), n, 0, _vp(a.u), n, 0, n, 1.0)])
                e2.add(bath, chi, chi, chi, [(_vp(a.u), n, 1, yb, chi, 0, n, 1.0)])
                for k in range(len(a.operators)):
                    yk = _vp(a.Y, (1 + k) * n * chi)
                    e1.add(yk, n, chi, chi, [(_vp(a.Pop, k * N), n, 0, _vp(a.u), n, 0, n, 1.0)])
                    e2.add(eff(k), chi, chi, chi, [(_vp(a.u), n, 1, yk, chi, 0, n, 1.0)])
                sym.append((bath, chi, chi, 1, 0))
                rho.add(_vp(a.rho_buf), 4, 4, 4,
                        [(_vp(a.psi, p * 4 * chi), chi, 0, _vp(a.psi, p * 4 * chi), chi, 1, chi,
                          1.0) for p in range(chi)])
            svd.append((_vp(a.psi), _vp(a.u), _vp(a.s), _vp(a.svt), _vp(a.svd_work),
                        n, n, n, 0, 0, 0))
            for r in range(chi):
                nrm_rows.append((_vp(a.svt, r * n), n))
        st['kron'] = self._kron(self._kron_arrays(kj))
        kj_, kt_ = kick.arrays()
        dkj, dkt = self._upload([kj_, kt_])
        st['kick'] = (dkj, len(kj_), dkt, kick.blocks)
        lz = L.lanczos_records(lz)
        ch = (lz['n'] + L.LZ_CHUNK - 1) // L.LZ_CHUNK
        lz['chunk_begin'] = np.concatenate([[0], np.cumsum(ch)[:-1]])
        st['nchunks'] = int(ch.sum())
        p1, s1 = g1.arrays()
        p2, s2 = g2.arrays()
        dp1, ds1, dp2, ds2, dlz = self._upload([p1, s1, p2, s2, lz])
        st['plan'] = L.MatvecPlan(dp1.data_ptr(), ds1.data_ptr(), len(g1.probs), R.gemm_tiles(p1),
                                  0, 0, 0, 0, dp2.data_ptr(), ds2.data_ptr(), len(g2.probs),
                                  R.gemm_tiles(p2), 0, L.GEMM_F64, L.GEMM_F64,
                                  (ctypes.c_double * 3)(R.gemm_flops(p1, s1), 0.0,
                                                        R.gemm_flops(p2, s2)))
        st['lz'] = (dlz, len(lz))
        st['res'] = np.zeros(len(lz), dtype=L.LANCZOS_RESULT)
        need = int(self.lib.tnx_lanczos_state_bytes(len(lz), st['nchunks'], NCV))
        st['lz_state'] = torch.empty(max(need, 256), dtype=torch.uint8, device=self.device)
        st['env1'], st['env2'] = self._gemm(e1), self._gemm(e2)
        st['rho'] = self._gemm(rho)
        st['svd'] = (self._upload([np.array(svd, dtype=L.SVD_JOB)])[0], len(svd))
        st['nrm_rows'] = (self._upload([np.array(nrm_rows, dtype=L.NORM_JOB)])[0], len(nrm_rows))
        st['nrm_env'] = (self._upload([np.array(nrm_env, dtype=L.NORM_JOB)])[0], len(nrm_env)) \
            if nrm_env else None
        if sym:
            srec, blocks = [], 0
            for (xp, nn, ld, cnt, stride) in sym:
                srec.append((xp, nn, ld, cnt, stride, blocks, 0))
                blocks += -(-(cnt * nn * nn) // BLK)
            st['sym'] = (self._upload([np.array(srec, dtype=L.SYM_JOB)])[0], len(srec), blocks)
        else:
            st['sym'] = None
        self.state = st

    # ----------------------------------------------------------- phases
    def _split_and_update(self):
        """SVD of the centres, kept right vectors normalised, environment
        (or bath and effective operator) updates."""
        st, lib, s = self.state, self.lib, self.stream()
        dj, nj = st['svd']
        L.check(lib.tnx_svd_jacobi_f64(ctypes_void(dj.data_ptr()), nj, SVD_SWEEPS, s), 'svd')
        dn, nn = st['nrm_rows']
        L.check(lib.tnx_normalize_f64(ctypes_void(dn.data_ptr()), nn, s), 'normalize rows')
        self._run_gemm(st['env1'])
        self._run_gemm(st['env2'])
        if st['nrm_env'] is not None:
            dn, nn = st['nrm_env']
            L.check(lib.tnx_normalize_f64(ctypes_void(dn.data_ptr()), nn, s), 'normalize env')
        if st['sym'] is not None:
            dsy, ns, nb = st['sym']
            L.check(lib.tnx_symmetrize_f64(ctypes_void(dsy.data_ptr()), ns, nb, s), 'symmetrize')

    def _solve(self):
        st = self.state
        dkj, nkj, dkt, nkb = st['kick']
        L.check(self.lib.tnx_mix(ctypes_void(dkj.data_ptr()), nkj, ctypes_void(dkt.data_ptr()), nkb,
                                 None, None, self.stream()), 'tnx_mix')
        dlz, nlz = st['lz']
        rc = self.lib.tnx_lanczos_run_ex(st['plan'], ctypes_void(dlz.data_ptr()), nlz,
                                         st['nchunks'], NCV, MAX_RESTARTS, NCV, L.F64, L.F64,
                                         ctypes_void(st['lz_state'].data_ptr()),
                                         ctypes_void(st['res'].ctypes.data), self.stream())
        if rc not in (L.OK, L.ENOCONV):
            raise L.HipError('tnx_lanczos_run_ex failed (rc=%d): %s'
                             % (rc, self.lib.tnx_last_error().decode(errors='replace')))
        return st['res'].copy()

    def _observe(self, obs):
        """tr(rho h) of the chains ``obs`` (the rho GEMMs run for every
        active chain; the host closes the trace of the observed ones)."""
        self._run_gemm(self.state['rho'])
        out = {}
        for ci in obs:
            a = self.chains[ci]
            r = a._get(a.rho_buf)
            h = self.hamilts[ci]
            if self.type == 'mpo':
                a.rho = [r[16 * j:16 * (j + 1)].reshape(4, 4) for j in range(3)]
                out[ci] = np.array([np.trace(x.dot(h)) for x in a.rho])
            else:
                a.rho = r[:16].reshape(4, 4)
                out[ci] = np.trace(a.rho.dot(h))
        return out

    def run(self, is_print=False):
        """The driver loop of ``DMRG_anyH.py:146-182`` for every chain;
        returns per chain (e1, info) with info the iteration count, the
        convergence flag and the Lanczos statistics."""
        nch = len(self.chains)
        e0 = [np.zeros((1, 3)) for _ in range(nch)]
        e1 = [np.ones((1, 3)) for _ in range(nch)]
        info = [{'iterations': 0, 'converged': False, 'matvecs': 0, 'unconverged_solves': 0,
                 'unconverged_at': [], 'max_resid_rel': 0.0, 'observations': []}
                for _ in range(nch)]
        active = list(range(nch))
        self._build(active)
        self._run_kron(self.state['kron'])
        self._split_and_update()
        t = 0
        while active and t < max(p['sweep_time'] for p in self.paras):
            active = [ci for ci in active if t < self.paras[ci]['sweep_time']]
            if active != self.state['active']:
                if not active:
                    break
                self._build(active)
            self._run_kron(self.state['kron'])
            res = self._solve()
            for q, ci in enumerate(active):
                info[ci]['iterations'] += 1
                info[ci]['matvecs'] += int(res['iters'][q])
                if res['converged'][q] == 0:
                    info[ci]['unconverged_solves'] += 1
                    info[ci]['unconverged_at'].append(t)
                rel = float(res['resid'][q]) / max(abs(float(res['theta'][q])), 1e-300)
                info[ci]['max_resid_rel'] = max(info[ci]['max_resid_rel'], rel)
            obs = [ci for ci in active if t % self.paras[ci]['dt_ob'] == 0]
            done = []
            if obs:
                vals = self._observe(obs)
                for ci in obs:
                    e1[ci] = vals[ci]
                    info[ci]['observations'].append((t, np.array(vals[ci], copy=True)))
                    if is_print:
                        print('At the %g-th sweep: Eb = ' % t + str(e1[ci]))
                    de = np.sum(abs(e0[ci] - e1[ci])) / self.chains[ci].n_site
                    if de > self.paras[ci]['break_tol']:
                        e0[ci] = e1[ci]
                    else:
                        if is_print:
                            print('Converged with de = %g' % de)
                        info[ci]['converged'] = True
                        done.append(ci)
            if done:
                # converged chains stop before their environment update
                active = [ci for ci in active if ci not in done]
                if not active:
                    break
                self._build(active)
            self._split_and_update()
            t += 1
        if self.device.type == 'cuda':
            torch.cuda.current_stream().synchronize()
        return list(zip(e1, info))


def dmrg_infinite_size(para=None, A=None, hamilt=None, is_print=True, rng=None):
    """``DMRG_anyH.py:106-188``: returns ``(A, ob, info)`` with
    ``ob = {'eb': e1}`` (the three bond energies inside the 'mpo' centre, or
    the centre bond energy of 'white') and ``info['t_cost']``.  ``rng``:
    generator of the initial centre (numpy's global one by default)."""
    t_start = time.time()
    if para is None:
        raise TypeError('para is required (the reference reads para before its default)')
    if is_print:
        print('Start ' + str(para['n_site']) + '-site iDMRG calculation')
    require_gpu()
    if hamilt is None:
        hamilt = hamiltonian_heisenberg(para['spin'], para['jxy'], para['jxy'], para['jz'],
                                        para['hx'] / 2, para['hz'] / 2)
    if A is None:
        d = para['d'] ** para['n_site'] if para['dmrg_type'] == 'mpo' else para['d']
        A = MpsInfinite(para['form'], d, para['chi'], para['d'] ** para['n_site'],
                        n_site=para['n_site'], is_symme_env=para['is_symme_env'],
                        dmrg_type=para['dmrg_type'], hamilt_index=para['hamilt_index'], rng=rng)
    run = InfiniteRun([A], [para], [hamilt])
    (e1, inf), = run.run(is_print)
    info = {'t_cost': time.time() - t_start}
    info.update((k, inf[k]) for k in ('iterations', 'converged', 'matvecs', 'unconverged_solves',
                                      'unconverged_at', 'max_resid_rel'))
    if is_print:
        print('Total time cost: %g' % info['t_cost'])
    return A, {'eb': e1}, info


def dmrg_infinite_size_batch(paras, seeds=None, is_print=False):
    """Several infinite-size runs of the same structure (dmrg_type, chi,
    symmetric environments) in one batch: every launch covers all chains.
    Returns a list of ``(A, ob, info)``; chain k's initial centre is drawn
    from ``RandomState(seeds[k])`` (numpy's global generator if None)."""
    require_gpu()
    t_start = time.time()
    chains, hams = [], []
    for k, para in enumerate(paras):
        rng = None if seeds is None else np.random.RandomState(seeds[k])
        h = hamiltonian_heisenberg(para['spin'], para['jxy'], para['jxy'], para['jz'],
                                   para['hx'] / 2, para['hz'] / 2)
        d = para['d'] ** para['n_site'] if para['dmrg_type'] == 'mpo' else para['d']
        chains.append(MpsInfinite(para['form'], d, para['chi'], para['d'] ** para['n_site'],
                                  n_site=para['n_site'], is_symme_env=para['is_symme_env'],
                                  dmrg_type=para['dmrg_type'],
                                  hamilt_index=para['hamilt_index'], rng=rng))
        hams.append(h)
    out = InfiniteRun(chains, paras, hams).run(is_print)
    t = time.time() - t_start
    return [(A, {'eb': e1}, dict(inf, t_cost=t)) for A, (e1, inf) in zip(chains, out)]


__all__ = ['MpsInfinite', 'InfiniteRun', 'dmrg_infinite_size', 'dmrg_infinite_size_batch',
           'generate_parameters_infinite_dmrg']
