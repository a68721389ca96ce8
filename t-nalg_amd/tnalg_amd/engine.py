"""Batched one-site DMRG engine on the gfx950 HIP library.

Every solve keeps its MPS, its block environments (``H`` plus open
operators per bond, see ``terms.py``) and its step scratch resident in HBM
(torch-ROCm allocations).  A *global step* advances every active solve of
the batch by one entry of its own schedule; each phase of the step
(QR gauge move, absorption, environment update, linear combinations,
Lanczos with the grouped-GEMM matvec) is ONE launch over all solves that
take part, built by concatenating the solves' descriptor records.

Reference correspondence (``library/MPSClass.py``):
  gauge move      correct_orthogonal_center / orthogonalize_mps   :186-279
  env update      update_effect_op_l0_to_l1 / bound_vec_operator_* :555-611, TBM:632-721
  Heff / matvec   all_environments_optimized + f_handle_optimized :715-858
  local solve     eigsh(I - tau Heff, k=1, 'LM', v0=A[p])          :883-889
Precision: ``'fp64'`` (plain) runs every GEMM in f64 MFMA; ``'fp32'`` and
``'bf16'`` run the matvec GEMMs in f32 / bf16 MFMA (f32 accumulation) on
low-precision copies of the environments, while gauge moves, environment
updates, the Krylov basis, reorthogonalisation and the tridiagonal solve stay
f64.
"""
import collections
import concurrent.futures
import ctypes
import hashlib
import itertools
import time

import numpy as np
import torch

from . import _lib as L
from . import records as R
from .device import ITEMSIZE, TORCH_DTYPE, Uploader, ctypes_void, require_gpu
from .terms import compile_terms_cached

PREC = {
    # name: (TI dtype, TC dtype, GEMM1 mode, GEMM2 mode, floor_rel)
    'fp64': (L.F64, L.F64, L.GEMM_F64, L.GEMM_F64, 0.0),
    'fp32': (L.F32, L.F32, L.GEMM_F32, L.GEMM_F32, 2e-6),
    'bf16': (L.BF16, L.F32, L.GEMM_BF16_BF16, L.GEMM_BF16_F32, 3e-3),
}


_PLAN_UID = itertools.count(1)
ARENA_ALIGN = 256
ARENA_HEADER = 256


def _p(t, off_elems=0, itemsize=8):
    return t.data_ptr() + off_elems * itemsize


class _ArenaViews:
    """Per-index buffers of one kind (site tensors, environments) inside a
    solve's arena.  Descriptors need only addresses (``ptr``); a tensor view
    is made on first indexing and kept (a solve has hundreds of buffers and
    its host code touches few of them as tensors)."""
    __slots__ = ('arena', 'base', 'ent', 'cache')

    def __init__(self, arena, n):
        self.arena = arena
        self.base = arena.data_ptr()
        self.ent = [None] * n      # (offset, bytes, f64 view?)
        self.cache = [None] * n

    def __len__(self):
        return len(self.ent)

    def __getitem__(self, i):
        v = self.cache[i]
        if v is None:
            off, nb, f64 = self.ent[i]
            v = self.arena[off:off + nb]
            if f64:
                v = v.view(torch.float64)
            self.cache[i] = v
        return v

    def ptr(self, i, off_elems=0, itemsize=8):
        return self.base + self.ent[i][0] + off_elems * itemsize


class _Stager:
    """Pinned host staging for uploads into persistent device buffers: a
    ring of pinned buffers that only grow, each reused once the copy out of
    it has completed (its event)."""

    def __init__(self, ring=4):
        self.host = [None] * ring
        self.events = [None] * ring
        self.slot = 0

    def upload(self, dst, arrays, offs):
        """dst[offs[i]:] <- arrays[i] (stream-ordered, no host sync unless the
        ring has wrapped onto a copy still in flight)."""
        total = dst.numel()
        k = self.slot
        self.slot = (self.slot + 1) % len(self.host)
        if self.events[k] is not None:
            self.events[k].synchronize()
        if self.host[k] is None or self.host[k].numel() < total:
            cap = 1 << max(16, (total - 1).bit_length())
            self.host[k] = torch.empty(cap, dtype=torch.uint8, pin_memory=True)
        hv = self.host[k].numpy()
        for a, o in zip(arrays, offs):
            if a.nbytes:
                hv[o:o + a.nbytes] = np.ascontiguousarray(a).reshape(-1).view(np.uint8)
        dst.copy_(self.host[k][:total], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self.events[k] = ev


def gemm_kmax(p, s):
    """Largest total K (sum over its segments) of the problems of a grouped
    GEMM launch."""
    if len(p) == 0:
        return 0
    k = np.concatenate([[0], np.cumsum(s['K'].astype(np.int64))])
    return int((k[p['seg_begin'] + p['seg_count']] - k[p['seg_begin']]).max())


class GemmTuner:
    """Kernel variant of each grouped-GEMM launch, by shape class (mode, bit
    length of the tile count, bit length of the largest K).  A class is
    tuned on its first matvec launches: every variant of the mode is timed
    with HIP events on the launch's stream and the fastest is kept.  The
    variants give the same bits, so tuning never changes results."""

    def __init__(self, lib, reps=3):
        self.lib = lib
        self.reps = reps
        self.nvar = {m: int(lib.tnx_gemm_variants(m)) for m in (L.GEMM_F64, L.GEMM_F32,
                                                                  L.GEMM_BF16_F32,
                                                                  L.GEMM_BF16_BF16)}
        self.table = {}
        self.version = 0   # bumped whenever the table changes

    @staticmethod
    def shape_class(mode, ntiles, kmax):
        return (mode, int(ntiles).bit_length(), int(kmax).bit_length())

    def mode_for(self, mode, ntiles, kmax):
        """The encoded mode (variant in the second byte) of a launch."""
        return mode | (self.table.get(self.shape_class(mode, ntiles, kmax), 0) << 8)

    def tune(self, mode, ntiles, kmax, launch):
        """Time every variant of ``mode`` with ``launch(encoded_mode)`` (a
        launch that may run repeatedly: it must not read its own output)."""
        key = self.shape_class(mode, ntiles, kmax)
        if key in self.table or self.nvar.get(mode, 1) < 2 or ntiles == 0:
            return
        best, best_t = 0, None
        for v in range(self.nvar[mode]):
            m = mode | (v << 8)
            launch(m)   # warm-up
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(self.reps):
                launch(m)
            e1.record()
            e1.synchronize()
            t = e0.elapsed_time(e1)
            if best_t is None or t < best_t:
                best, best_t = v, t
        self.table[key] = best
        self.version += 1


# the current raw stream of a device (torch's own accessor of it)
_RAW_STREAM = getattr(torch._C, '_cuda_getCurrentRawStream', None)


class _Resident:
    """A set of descriptor arrays uploaded once and kept in device memory
    (steps that repeat every sweep reuse them without host work).  ``host``:
    host objects built for it once (a Lanczos call's plan array and result
    array; a step's recorded pre-Lanczos launches)."""
    __slots__ = ('buf', 'ptrs', 'meta', 'host')

    def __init__(self, device, arrays, meta, stager):
        offs, total = [], 0
        for a in arrays:
            offs.append(total)
            total += (a.nbytes + 255) // 256 * 256
        self.buf = torch.empty(max(total, 256), dtype=torch.uint8, device=device)
        stager.upload(self.buf, arrays, offs)
        base = self.buf.data_ptr()
        self.ptrs = [base + o for o in offs]
        self.meta = meta
        self.host = None


def max_dims(dims, d, ob):
    """Largest dimension each bond takes: the initial (random MPS) dims and
    the dims after the initial canonicalisation around ``ob`` (bonds only
    shrink afterwards, via the reduced QR of the first sweep)."""
    n = len(dims) - 1
    cd = list(dims)
    for s in range(0, ob):
        cd[s + 1] = min(cd[s] * d, cd[s + 1])
    for s in range(n - 1, ob, -1):
        cd[s] = min(cd[s], d * cd[s + 1])
    return [max(a, b) for a, b in zip(dims, cd)]


def rsv_stride(dsq, dmx):
    """Doubles per bond slot of the saved R factors (a values-only SVD work
    area: n^2 + n + 1 doubles), rounded to 256 bytes."""
    return (dsq + dmx + 1 + 31) // 32 * 32


def buffer_layout(T, dm, d, ncv, lowp):
    """Sizes of every device buffer of a solve (elements for f64 buffers,
    bytes for the ``uint8`` ones); also the HBM footprint used for admission.
    With two-site tables the Krylov and matvec buffers hold a merged pair
    (Dl d^2 Dr), the pair tensor theta has its own buffer, and the split's
    square matrices (order up to max(Dl d, d Dr)) fit the CholeskyQR and
    scratch buffers."""
    n = T.n
    st = T.sites + T.pairs
    psi = max(dm[s] * d * dm[s + 1] for s in range(n))
    dsq = max(x * x for x in dm)
    dmx = max(dm)
    if T.pairs:
        psi = max(psi, max(dm[p] * d * d * dm[p + 2] for p in range(n - 1)))
        dmx = max(dmx, max(max(dm[p], dm[p + 2]) * d for p in range(n - 1)))
        dsq = max(dsq, dmx * dmx)
    nL = [len(T.left_opens[b]) for b in range(n + 1)]
    nR = [len(T.right_opens[b]) for b in range(n + 1)]
    n_lm = max(len(s.lm) for s in st)
    n_mr = max(len(s.mr) for s in st)
    n_sd = max(len(s.strad) for s in st)
    n_lc = max(len(s.lm) + len(s.mr) + len(s.strad) for s in st)
    lay = {
        'psi': psi, 'dsq': dsq,
        'mps': [dm[s] * d * dm[s + 1] for s in range(n)],
        'Lenv': [(1 + nL[b]) * dm[b] * dm[b] for b in range(n + 1)],
        'Renv': [(1 + nR[b]) * dm[b] * dm[b] for b in range(n + 1)],
        'lc64': max(1, max(n_lm, n_mr)) * dsq,
        'lc_lp': max(1, n_lc) * dsq * 8,
        'Tb': (1 + n_lm + n_sd) * psi * 8,
        'Ub': (1 + n_mr) * psi * 8,
        'Yb': (1 + max(n_lm, n_mr) + max(max(nL), max(nR))) * psi,
        'Wb': (1 + max(max(nL), max(nR))) * psi,
        'V': (ncv + 1) * psi,
        'Rq': dsq + dmx + 64,
        'qrw': 2 * psi + dsq,
        'chol': 6 * dsq + dmx + 64,
        'lmv': max(1, n - 1) * dmx,   # entanglement spectra, one row of dmx per bond
        # R factors of the QR moves of a spectrum pass, one slot per bond,
        # each the work area of a values-only SVD (n^2 + n + 1 doubles)
        'rsv': max(1, n - 1) * rsv_stride(dsq, dmx),
    }
    # arena: (name, index, bytes, f64 view?) in address order.  The MPS and
    # the CholeskyQR constants come first (staged from the host in one
    # copy); everything after them starts zeroed.  The first ARENA_HEADER
    # bytes stay unused, so no buffer sits at relative address 0 (0 marks a
    # null pointer in the descriptor records).
    spec = [('mps', s, 8 * x, True) for s, x in enumerate(lay['mps'])]
    spec.append(('cholb', None, 8 * lay['chol'], True))
    stage_end = len(spec)
    spec += [('Lenv', b, 8 * x, True) for b, x in enumerate(lay['Lenv'])]
    spec += [('Renv', b, 8 * x, True) for b, x in enumerate(lay['Renv'])]
    if lowp:
        spec += [('Lenv_lp', b, 4 * x, False) for b, x in enumerate(lay['Lenv'])]
        spec += [('Renv_lp', b, 4 * x, False) for b, x in enumerate(lay['Renv'])]
    spec += [('lc64', None, 8 * lay['lc64'], True), ('lc_lp', None, lay['lc_lp'], False),
             ('Tb', None, lay['Tb'], False), ('Ub', None, lay['Ub'], False),
             ('Yb', None, 8 * lay['Yb'], True), ('Wb', None, 8 * lay['Wb'], True),
             ('V', None, 8 * lay['V'], True), ('w', None, 8 * psi, True),
             ('x', None, 8 * psi, False), ('init', None, 8 * psi, True),
             ('Rq', None, 8 * lay['Rq'], True), ('qrw', None, 8 * lay['qrw'], True),
             ('lmv', None, 8 * lay['lmv'], True), ('rsv', None, 8 * lay['rsv'], True)]
    if T.pairs:
        spec.append(('th', None, 8 * psi, True))
    offs, o = [], ARENA_HEADER
    for k, (_, _, nb, _) in enumerate(spec):
        if k == stage_end:
            lay['stage_bytes'] = o
        offs.append(o)
        o += (nb + ARENA_ALIGN - 1) // ARENA_ALIGN * ARENA_ALIGN
    lay['arena'] = [(name, idx, off, nb, f64) for (name, idx, nb, f64), off in zip(spec, offs)]
    lay['offsets'] = {(name, idx): off for name, idx, off, _, _ in lay['arena']}
    lay['dmax_max'] = dmx
    lay['bytes'] = o
    return lay


def dmrg_sites(para):
    """1 (one-site sweeps, the reference's scheme) or 2 (two-site sweeps with
    SVD truncation to chi), from ``para['dmrg_sites']`` (default 1)."""
    k = int(para.get('dmrg_sites', 1))
    if k not in (1, 2):
        raise ValueError('dmrg_sites must be 1 or 2 (got %r)' % para.get('dmrg_sites'))
    return k


def initial_chi(para):
    """Bond dimension of the random initial MPS: ``chi``, or for two-site
    sweeps ``para['chi_init']`` when given (the truncation then grows the
    bonds up to ``chi``)."""
    chi = int(para['chi'])
    if dmrg_sites(para) == 2:
        chi0 = int(para.get('chi_init', chi))
        if not 1 <= chi0 <= chi:
            raise ValueError('chi_init must be in [1, chi]')
        # a bond of 3 with chi < 6 would need a truncating split of order 6,
        # below the block SVD's range (bonds 1 and 2 are never truncated,
        # bonds >= 4 give orders >= 8)
        if chi0 == 3 and chi < 6:
            raise ValueError('two-site sweeps: chi_init = 3 needs chi >= 6')
        return chi0
    return chi


def solve_geometry(para, ncv, lowp, dims=None):
    """(term tables, initial bond dims, per-bond maximum dims, arena layout)
    of a solve of ``para``.  ``dims``: bond dims of a given initial MPS;
    default those of the random MPS, [1, chi, ..., chi, 1] (chi_init for
    two-site sweeps).  Buffers are sized for bonds up to chi.  Layouts are
    kept on the (shared, read-only) term tables."""
    two = dmrg_sites(para) == 2
    T = compile_terms_cached(para, pairs=two)
    chi = int(para['chi'])
    if dims is None:
        chi0 = initial_chi(para)
        dims = [1] + [chi0] * (T.n - 1) + [1]
    ob = int(para.get('ob_position', 0))
    if two:
        if T.d * T.d > L.MIX_DMAX:
            raise ValueError('two-site sweeps need d^2 <= %d (d = %d)' % (L.MIX_DMAX, T.d))
        if chi < 4:
            raise ValueError('two-site sweeps need chi >= 4')
        full = [1] + [chi] * (T.n - 1) + [1]
        dm = max_dims([max(a, b) for a, b in zip(dims, full)], T.d, ob)
        if T.d * max(dm) > L.SVDJ_NMAX:
            raise ValueError('two-site split of order %d exceeds the SVD limit %d (chi <= %d)'
                             % (T.d * max(dm), L.SVDJ_NMAX, L.SVDJ_NMAX // T.d))
    else:
        dm = max_dims(dims, T.d, ob)
    key = (tuple(dims), ob, ncv, bool(lowp), chi if two else 0)
    lay = T.layouts.get(key)
    if lay is None:
        lay = buffer_layout(T, dm, T.d, ncv, lowp)
        T.layouts[key] = lay
    return T, list(dims), dm, lay


def make_stage(lay, host, pin=True):
    """Host image (pinned for a GPU engine) of the arena's staged prefix:
    every site of the initial MPS ``host`` at its offset (zero padding up to
    the largest shape the site takes) and the blocked-CholeskyQR identity
    (stride dmax) and ones."""
    st = torch.zeros(lay['stage_bytes'], dtype=torch.uint8, pin_memory=pin)
    v = st.numpy()
    offs = lay['offsets']
    for s, a in enumerate(host):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        o = offs[('mps', s)]
        v[o:o + a.nbytes] = a.view(np.uint8)
    cb = v[offs[('cholb', None)]:].view(np.float64)
    dsq, dmx = lay['dsq'], lay['dmax_max']
    cb[5 * dsq + np.arange(dmx) * (dmx + 1)] = 1.0
    cb[6 * dsq:6 * dsq + dmx] = 1.0
    return st


def random_mps_host(n, d, chi, seed):
    """Initial MPS in the reference draw order (TensorBasicModule.py:179-211):
    site 0 (1, d, chi), site n-1 (chi, d, 1), then sites 1..n-2 (chi, d, chi),
    standard normals from ``RandomState(seed)``."""
    rng = seed if hasattr(seed, 'randn') else np.random.RandomState(seed)
    host = [None] * n
    host[0] = rng.randn(1, d, chi)
    host[n - 1] = rng.randn(chi, d, 1)
    for s in range(1, n - 1):
        host[s] = rng.randn(chi, d, chi)
    return host


class Step:
    """One schedule entry of a solve: kind in {'move', 'solve', 'svdmove',
    'qmove', 'spectra', 'resume', 'solve2', 'split2', 'resume2', 'tmove',
    'gate2', 'tsplit', 'tnorm'}, site, direction (+1: centre moves right,
    -1: left).  'resume' continues the local eigensolve of a 'solve' step
    from its Ritz vector (the MPS site tensor) when one Krylov cycle did not
    converge; 'qmove' is a right QR move that keeps its R factor; 'spectra'
    (site = first bond, direction = last bond) takes the singular values of
    the kept factors.  The time-evolution steps touch no environment:
    'tmove' is a QR centre move; 'gate2' (site p) splits the previous pair
    p - dir (when dir != 0) after normalising it, forms theta = A_p A_{p+1}
    and applies the two-site gate (merged operator ``gate``); 'tsplit'
    normalises and splits the last pair; 'tnorm' normalises the centre
    tensor at the site."""
    __slots__ = ('kind', 'site', 'dir', 'sweep', 'gate')

    def __init__(self, kind, site, direction, sweep=-1, gate=0):
        self.kind, self.site, self.dir, self.sweep = kind, site, direction, sweep
        self.gate = gate

    def __repr__(self):
        return 'Step(%s,%d,%+d,sw=%d)' % (self.kind, self.site, self.dir, self.sweep)


class Solve:
    """Device state of one ground-state solve."""

    def __init__(self, engine, para, seed=0, mps=None, tag=None, stage=None, gates=0):
        """``mps``: initial site tensors (default: the seeded random MPS);
        ``stage``: its prepared pinned image (``make_stage``), when the
        caller built it already; ``gates``: merged two-site operators the
        solve reserves after its Hamiltonian's (``set_gate``)."""
        self.engine = engine
        self.dev = engine.device
        self.tag = tag
        self.para = para
        in_dims = None if mps is None else [np.shape(mps[0])[0]] + [np.shape(a)[2] for a in mps]
        self.T, dims, dm, lay = solve_geometry(para, engine.ncv, engine.uses_lowp, in_dims)
        self.n = self.T.n
        self.d = self.T.d
        self.chi = int(para['chi'])
        self.tau = float(para.get('tau', 1e-4))
        self.eigs_tol = float(para.get('eigs_tol', 1e-5))
        self.ob = int(para.get('ob_position', 0))
        n, d, chi = self.n, self.d, self.chi
        if n < 2:
            raise ValueError('the chain needs at least 2 sites')
        if not 0 <= self.ob < n:
            raise ValueError('ob_position must be in [0, l)')
        on_device = stage is None and mps is None and engine.init_rng == 'device'
        if stage is None and not on_device:
            host = (random_mps_host(n, d, initial_chi(para), seed) if mps is None
                    else [np.asarray(a, dtype=np.float64) for a in mps])
            stage = make_stage(lay, host, pin=engine.device.type != 'cpu')
        self.dims = list(dims)
        self.dmax = dm
        T = self.T
        self.nL = [len(T.left_opens[b]) for b in range(n + 1)]
        self.nR = [len(T.right_opens[b]) for b in range(n + 1)]
        self.layout = lay
        self.psi_max, self.dsq_max = lay['psi'], lay['dsq']
        self.dmax_max = lay['dmax_max']
        self.sites = dmrg_sites(para)
        self.lp_bytes = 4 if engine.uses_lowp else 0
        # one arena per solve; every buffer is a view into it
        self.arena = engine.arena_get(lay['bytes'])
        self.base = self.arena.data_ptr()
        views = {'mps': _ArenaViews(self.arena, n), 'Lenv': _ArenaViews(self.arena, n + 1),
                 'Renv': _ArenaViews(self.arena, n + 1), 'Lenv_lp': _ArenaViews(self.arena, n + 1),
                 'Renv_lp': _ArenaViews(self.arena, n + 1)}
        for name, idx, off, nb, f64 in lay['arena']:
            if idx is None:
                v = self.arena[off:off + nb]
                setattr(self, name, v.view(torch.float64) if f64 else v)
            else:
                views[name].ent[idx] = (off, nb, f64)
        self.mps, self.Lenv, self.Renv = views['mps'], views['Lenv'], views['Renv']
        if self.lp_bytes:
            self.Lenv_lp, self.Renv_lp = views['Lenv_lp'], views['Renv_lp']
        # staged prefix (initial MPS, CholeskyQR constants) from the pinned
        # image, or drawn on the device; the rest of the arena starts at zero
        if on_device:
            engine.arena_fill_device(self.arena, lay, n, d, dims, seed)
        else:
            engine.arena_fill(self.arena, stage)
        ops2 = T.ops2 if T.pairs else None
        self.gate_base = 0 if ops2 is None else len(ops2)
        if gates:
            if ops2 is None:
                raise ValueError('gates need the two-site tables (dmrg_sites = 2)')
            ops2 = np.concatenate([ops2, np.zeros((gates,) + ops2.shape[1:])])
        self.ngates = int(gates)
        self.ops_slot = engine.alloc_ops(T.ops, ops2)
        # relocation of a shared plan to this solve: arena base, operator
        # slot, coupling-coefficient values
        self.reloc = (self.base, self.ops_slot, T.coef_values)
        if T.coef_digest is None:
            T.coef_digest = hashlib.sha1(T.coef_values.tobytes()).hexdigest()
        self.reloc_key = (self.base, self.ops_slot, T.coef_digest)
        # plans depend on the structure of the term tables (which couplings,
        # operator indices), the shapes and the solver tolerances, not on the
        # coupling or operator values
        self.sig = hashlib.sha1(repr((
            self.ob, tuple(dims), tuple(dm), self.eigs_tol, self.tau, T.structure()
        )).encode()).hexdigest()
        self.d2 = d * d
        # bookkeeping
        self.center = -1
        self.queue = []
        self.sweeps_done = 0
        self.lowp_dtype = None   # dtype currently held by the *_lp copies
        self.last_theta = float('nan')
        self.thetas = []
        self.stats = {'matvecs': 0, 'steps': 0, 'unconverged': 0, 'restarts': 0}
        self._lm = [np.zeros(0) for _ in range(n - 1)]
        self._lm_pending = {}   # bond -> length of a spectrum written on the device
        self.failed = None
        # non-fatal failures of this solve's steps (an eigensolve or an SVD
        # that hit its iteration limit; the step's result is still written)
        self.warnings = []
        self._lz_site = -1    # site of the step being run (for the warnings)
        # precision of every later sweep, replacing the schedule's (the driver
        # switches a solve that converged at low precision to f64)
        self.prec_override = None
        # shared plan parts relocated to this solve, by (plan uid, phase);
        # a solve revisits the same plans every sweep, and a later solve on
        # the same arena, operator slot and couplings takes them over
        self.moved = engine._moved_pool.pop(self.reloc_key, None) or {}
        self._next = None   # (step, precision, plan) planned ahead
        self.cycles = 0       # unconverged Krylov cycles of the current local solve
        self.resume = False   # the last cycle did not converge: resume next step

    @property
    def lm(self):
        """Entanglement spectra per bond (empty until an SVD move visits the
        bond).  The SVD moves write them to the device; they are copied to
        the host in one transfer when read."""
        if self._lm_pending:
            host = self.lmv.cpu().numpy().reshape(self.n - 1, self.dmax_max)
            for b, k in self._lm_pending.items():
                self._lm[b] = host[b, :k].copy()
            self._lm_pending = {}
        return self._lm

    def set_gate(self, k, gate):
        """Write gate ``k`` (a d^2 x d^2 matrix on the pair's merged index
        s_1 d + s_2) into the solve's slot of the operator table."""
        if not 0 <= k < self.ngates:
            raise ValueError('gate %d not reserved (the solve has %d)' % (k, self.ngates))
        g = np.ascontiguousarray(gate, dtype=np.float64).reshape(-1)
        d4 = self.d2 * self.d2
        if g.size != d4:
            raise ValueError('a gate has %d elements (d^4), got %d' % (d4, g.size))
        o = self.op_index2(self.gate_base + k) * d4
        self.engine.ops[o:o + d4].copy_(torch.from_numpy(g).to(self.engine.device))

    # ----------------------------------------------------------- addresses
    def op_index(self, k):
        return self.ops_slot + k

    def op_index2(self, k):
        """Index of merged two-site operator k (d^2 x d^2).  The operator table
        keeps them in a second region at a fixed offset from the solve's slot,
        so a plan relocates them with the same slot shift as the site ones."""
        return self.ops_slot + self.engine.ops_off2 + k

    def L(self, b, m):
        return self.Lenv.ptr(b, m * self.dmax[b] ** 2)

    def Rr(self, b, m):
        return self.Renv.ptr(b, m * self.dmax[b] ** 2)

    def L_lp(self, b, m, ti):
        if ti == L.F64:
            return self.L(b, m)
        return self.Lenv_lp.ptr(b, m * self.dmax[b] ** 2, ITEMSIZE[ti])

    def R_lp(self, b, m, ti):
        if ti == L.F64:
            return self.Rr(b, m)
        return self.Renv_lp.ptr(b, m * self.dmax[b] ** 2, ITEMSIZE[ti])

    def free(self):
        if self.arena is None:
            return
        self.engine.free_ops(self.ops_slot)
        for name in ('mps', 'Lenv', 'Renv', 'Lenv_lp', 'Renv_lp'):
            if hasattr(self, name):
                setattr(self, name, None)
        for name in ('lc64', 'lc_lp', 'Tb', 'Ub', 'Yb', 'Wb', 'V', 'w', 'x', 'init', 'Rq', 'qrw',
                     'cholb', 'lmv', 'th'):
            setattr(self, name, None)
        self.engine.arena_release(self.arena)
        self.arena = None
        if self.moved:
            pool = self.engine._moved_pool
            pool[self.reloc_key] = self.moved
            while len(pool) > self.engine.moved_pool_limit:
                pool.pop(next(iter(pool)))
        self.moved = {}
        self._next = None

    def part(self, ph, key):
        """Phase ``key`` of the shared plan ``ph`` at this solve's addresses
        and coefficients (relocated once per plan, then reused)."""
        uid = ph['uid']
        got = self.moved.get((uid, key))
        if got is None:
            pk = ph['packed']
            buf = self.moved.get(uid)
            if buf is None:
                buf = pk.relocate(self.reloc)
                self.moved[uid] = buf
            got = pk.views(buf, key)
            self.moved[(uid, key)] = got
        return got

    # ------------------------------------------------------------ schedule
    def schedule_canonicalise(self):
        """``correct_orthogonal_center(ob)`` on a centre-less MPS
        (``MPSClass.py:250-279``): left moves 0..ob-1, right moves N-1..ob+1,
        each followed by the environment update through the moved site."""
        q = [Step('move', s, +1) for s in range(0, self.ob)]
        q += [Step('move', s, -1) for s in range(self.n - 1, self.ob, -1)]
        self.queue.extend(q)
        self._sched_center = self.ob

    def schedule_sweep(self, sweep):
        """One sweep of ``DMRG_anyH.py:47-64``: ob+1..N-1, N-2..0, 1..ob-1.
        With two-site sweeps, sweeps 0 .. two_site_sweeps-1 (default all)
        sweep pairs and the later ones are one-site."""
        if self.sites == 2 and sweep < int(self.para.get('two_site_sweeps', 1 << 30)):
            return self._schedule_sweep2(sweep)
        n, ob = self.n, self.ob
        sites = [(s, +1) for s in range(ob + 1, n)] + [(s, -1) for s in range(n - 2, -1, -1)] + \
            [(s, +1) for s in range(1, ob)]
        c = self._sched_center
        for s, _ in sites:
            while abs(c - s) > 1:   # multi-hop centre moves (MPSClass.py:275-277)
                dirn = 1 if s > c else -1
                self.queue.append(Step('move', c, dirn, sweep))
                c += dirn
            dirn = 1 if s > c else -1
            self.queue.append(Step('solve', s, dirn, sweep))
            c = s
        self._sched_center = c

    def two_site_pairs(self):
        """Pairs (p, p+1) of one two-site sweep from the centre ob, the same
        route as the one-site sweep: right to the end, left to the start,
        right back to ob; the turning pairs are solved once."""
        n, ob = self.n, self.ob
        right = list(range(ob, n - 1))
        left = list(range(n - 3 if right else n - 2, -1, -1))
        back = list(range(1 if left else ob, ob))
        return right + left + back

    def _schedule_sweep2(self, sweep):
        """One two-site sweep.  Step ('solve2', p, dir) first splits the
        previous pair p - dir (dir = +1: centre moves right onto p; -1: left
        onto p+1; 0: nothing to split), updates the environment through the
        new isometry, forms theta = A_p A_{p+1} and solves the pair.  The
        closing ('split2', p, dir) splits the last pair so the sweep ends with
        the centre at ob and a whole MPS."""
        if self._sched_center != self.ob:
            raise ValueError('a two-site sweep starts with the centre at ob_position')
        if self.n < 2:
            raise ValueError('two-site sweeps need at least 2 sites')
        seq = self.two_site_pairs()
        dirn = 0
        for t, p in enumerate(seq):
            self.queue.append(Step('solve2', p, dirn, sweep))
            nxt = seq[t + 1] if t + 1 < len(seq) else None
            if nxt is None:
                dirn = +1 if self.ob == p + 1 else -1
            else:
                dirn = nxt - p
        self.queue.append(Step('split2', seq[-1], dirn, sweep))
        self._sched_center = self.ob

    # ------------------------------------------------------------- planning
    def plan(self, st, prec):
        """Descriptor records of one step (dict phase -> arrays), shared by
        every solve of the engine with the same signature: device addresses
        are relative to the solve's arena and operator indices to its slot of
        the operator table (``Solve.part`` moves a phase to a solve)."""
        dims = self.dims if type(self.dims) is tuple else tuple(self.dims)
        key = (self.sig, st.kind, st.site, st.dir, dims, prec, st.gate)
        cache = self.engine.plan_cache
        hit = cache.get(key)
        if hit is not None:
            # bond dimensions after the step, shared (immutable) with the plan
            self.dims = hit['dims_after_t']
            return hit
        self.dims = list(dims)   # the builders update bond dimensions in place
        lo, hi = self.base, self.base + self.layout['bytes']
        if st.kind not in ('solve', 'solve2'):
            ph = self._build_plan(st, prec)
            R.pack_relative(ph, lo, hi, self.ops_slot)
            ph['dims_after_t'] = tuple(ph['dims_after'])
            cache[key] = ph
            return ph
        # a solve step = the gauge move (two-site: split of the previous
        # pair) and environment update, which do not depend on the matvec
        # precision (built once and shared by the plans of every precision),
        # + the precision's cast and local records
        ckey = (self.sig, st.kind + '-env', st.site, st.dir, dims)
        com = cache.get(ckey)
        if com is None:
            com = R.Plan()
            self._plan_common(com, st)
            com['dims_after'] = list(self.dims)
            R.pack_relative(com, lo, hi, self.ops_slot)
            cache[ckey] = com
        else:
            self.dims = list(com['dims_after'])
        loc = R.Plan()
        self._plan_precision(loc, st, prec)
        R.pack_relative(loc, lo, hi, self.ops_slot)
        meta = {k: loc[k] for k in ('prec', 'mv_flops')}
        meta['dims_after'] = list(self.dims)
        meta['dims_after_t'] = tuple(self.dims)
        meta['uid'] = next(_PLAN_UID)
        ph = R.merge_packed([com, loc], meta)
        cache[key] = ph
        return ph

    @staticmethod
    def _env_site(st):
        """Site whose new isometry the environment update of step ``st`` goes
        through (the new block is at bond q+1 for dir > 0, q for dir < 0)."""
        if st.kind == 'move':
            return st.site
        if st.kind == 'solve2':   # the split pair is site - dir
            return st.site - 1 if st.dir > 0 else st.site + 2
        if st.kind == 'split2':
            return st.site if st.dir > 0 else st.site + 1
        return st.site - st.dir

    def _plan_common(self, ph, st):
        """Records of a 'move', 'solve', 'solve2' or 'split2' step that do not
        depend on the matvec precision: gauge move of the previous centre
        (two-site: split of the previous pair), environment update through
        it, and for 'solve2' the pair tensor theta."""
        q = self._env_site(st)
        if st.kind in ('solve2', 'split2'):
            if st.dir:
                self._plan_split(ph, st.site - st.dir if st.kind == 'solve2' else st.site, st.dir)
                self._plan_env(ph, q, st.dir)
            if st.kind == 'solve2':
                self._plan_theta(ph, st.site)
            return
        self._plan_gauge(ph, q, st.dir, copy_back=(st.kind == 'move'))
        self._plan_env(ph, q, st.dir)

    def _plan_precision(self, ph, st, prec):
        """Records of a step that depend on the precision: low-precision copy
        of the new environment block, local eigensolve."""
        q = self._env_site(st)
        if st.dir and self.engine.uses_lowp and prec != 'fp64':
            self._plan_cast_block(ph, q + 1 if st.dir > 0 else q, st.dir > 0, PREC[prec][0])
        if st.kind in ('solve', 'solve2'):
            self._plan_local(ph, st.site, prec, pair=st.kind == 'solve2')

    def _build_plan(self, st, prec):
        self.dims = list(self.dims)   # the builders update bond dimensions in place
        ph = R.Plan()
        if st.kind in ('move', 'solve', 'solve2', 'split2'):
            self._plan_common(ph, st)
            self._plan_precision(ph, st, prec)
        elif st.kind == 'svdmove':
            self._plan_svd(ph, st.site, st.dir)
        elif st.kind == 'qmove':
            # a right QR move that keeps its R factor for the spectrum
            self._plan_common(ph, Step('move', st.site, st.dir, st.sweep))
            self._plan_save_r(ph, st.site)
        elif st.kind == 'spectra':
            self._plan_spectra(ph, st.site, st.dir)
        elif st.kind == 'tmove':
            self._plan_gauge(ph, st.site, st.dir, copy_back=True)
        elif st.kind in ('gate2', 'tsplit'):
            p = st.site - st.dir if st.kind == 'gate2' else st.site
            if st.dir:
                d2 = self.d * self.d
                ph['norm'] = np.array([(self.th.data_ptr(),
                                        self.dims[p] * d2 * self.dims[p + 2])], dtype=L.NORM_JOB)
                self._plan_split(ph, p, st.dir)
            if st.kind == 'gate2':
                self._plan_gate(ph, st.site, st.gate)
        elif st.kind == 'tnorm':
            p = st.site
            ph['norm'] = np.array([(self.mps.ptr(p), self.dims[p] * self.d * self.dims[p + 1])],
                                  dtype=L.NORM_JOB)
        elif st.kind == 'resume':
            # start vector <- the Ritz vector the last cycle wrote to the site
            # (the same vector, and so the same normalised V[0], as a restart
            # inside one Lanczos call)
            p = st.site
            mx = R.MixList()
            mx.add(self.init.data_ptr(), L.F64, 1, 1, self.dims[p] * self.d * self.dims[p + 1],
                   [(self.mps.ptr(p), L.F64, 1.0, -1)])
            ph['copy'] = mx.arrays()
            self._plan_local(ph, p, prec)
        elif st.kind == 'resume2':   # the same for a pair: theta holds the Ritz vector
            p = st.site
            mx = R.MixList()
            mx.add(self.init.data_ptr(), L.F64, 1, 1,
                   self.dims[p] * self.d * self.d * self.dims[p + 2],
                   [(self.th.data_ptr(), L.F64, 1.0, -1)])
            ph['copy'] = mx.arrays()
            self._plan_local(ph, p, prec, pair=True)
        ph['dims_after'] = list(self.dims)
        ph['uid'] = next(_PLAN_UID)
        return ph

    def _plan_cholqr(self, ph, q, m, n, right):
        """Shifted CholeskyQR3 of the tall panel M (m x n) of site q: M is
        A_q (row-major m x n) for a left move, A_q^T (A_q row-major n x m) for
        a right move.  Stages (each one launch over all solves):
        G=M^T M | chol+inv (shifted) | Q1=M Ri | G=Q1^T Q1 | chol+inv |
        Q2=Q1 Ri, R21=R2 R1 | G=Q2^T Q2 | chol+inv | Q=Q2 Ri, R=R3 R21."""
        A = self.mps.ptr(q)
        dsq = self.dsq_max
        Q1 = _p(self.qrw, 0)
        Q2 = _p(self.qrw, self.psi_max)
        G = _p(self.qrw, 2 * self.psi_max)
        Ri = _p(self.cholb, 0)
        R1, R2, R3, R21 = (_p(self.cholb, k * dsq) for k in (1, 2, 3, 4))
        st = _p(self.cholb, 6 * dsq + self.dmax_max + 1)
        shift = 11.0 * (m * n + n * (n + 1)) * 1.1102230246251565e-16

        def gemm(c, M, N, ldc, segs):
            g = R.GemmList()
            g.add(c, M, N, ldc, segs)
            return g

        def chol(r, s):
            return np.array([(G, r, Ri, st, 0, n, n, s)], dtype=L.CHOL_JOB)

        stages = []
        if not right:
            stages.append(('g', gemm(G, n, n, n, [(A, n, 1, A, n, 0, m, 1.0)])))
        else:
            stages.append(('g', gemm(G, n, n, n, [(A, m, 0, A, m, 1, m, 1.0)])))
        stages.append(('c', chol(R1, shift)))
        if not right:
            stages.append(('g', gemm(Q1, m, n, n, [(A, n, 0, Ri, n, 0, n, 1.0)])))
        else:
            stages.append(('g', gemm(Q1, m, n, n, [(A, m, 1, Ri, n, 0, n, 1.0)])))
        stages.append(('g', gemm(G, n, n, n, [(Q1, n, 1, Q1, n, 0, m, 1.0)])))
        stages.append(('c', chol(R2, 0.0)))
        g = gemm(Q2, m, n, n, [(Q1, n, 0, Ri, n, 0, n, 1.0)])
        g.add(R21, n, n, n, [(R2, n, 0, R1, n, 0, n, 1.0)])
        stages.append(('g', g))
        stages.append(('g', gemm(G, n, n, n, [(Q2, n, 1, Q2, n, 0, m, 1.0)])))
        stages.append(('c', chol(R3, 0.0)))
        if not right:
            g = gemm(A, m, n, n, [(Q2, n, 0, Ri, n, 0, n, 1.0)])
        else:
            g = gemm(A, n, m, m, [(Ri, n, 1, Q2, n, 1, n, 1.0)])
        g.add(self.Rq.data_ptr(), n, n, n, [(R3, n, 0, R21, n, 0, n, 1.0)])
        stages.append(('g', g))
        ph['cq'] = [(kind, obj.arrays() if kind == 'g' else obj) for kind, obj in stages]

    def _plan_cholqr_blocked(self, ph, q, m, n, right):
        """Shifted CholeskyQR3 for n > 128: the panel X (m x n, row-major) is
        overwritten in place by Q1, Q2, Q; each pass is the Gram GEMM, a
        blocked right-looking Cholesky (128 x 128 diagonal blocks in LDS,
        panel and trailing updates on the MFMA GEMM) and a blocked
        triangular solve Q_j = (X_j - Q_<j R_<j,j) R_jj^-1."""
        B = L.CHOL_NMAX
        nb = (n + B - 1) // B
        dsq = self.dsq_max
        A = self.mps.ptr(q)
        X = _p(self.qrw, 0) if right else A          # row-major m x n panel
        Wt = _p(self.qrw, self.psi_max)              # m x B scratch
        G = _p(self.qrw, 2 * self.psi_max)           # n x n
        Ri = _p(self.cholb, 0)                       # diagonal-block inverses
        Rs = [_p(self.cholb, k * dsq) for k in (1, 2, 3)]
        R21 = _p(self.cholb, 4 * dsq)
        I, LDI = _p(self.cholb, 5 * dsq), self.dmax_max   # identity, stride dmax
        ones = _p(self.cholb, 6 * dsq)
        tr = _p(self.cholb, 6 * dsq + self.dmax_max)
        st = tr + 8
        shift = 11.0 * (m * n + n * (n + 1)) * 1.1102230246251565e-16

        def G1(c, M, N, ldc, segs, c0=0, ldc0=0, beta=0.0):
            g = R.GemmList()
            g.add(c, M, N, ldc, segs, c0=c0, ldc0=ldc0, beta=beta)
            return ('g', g.arrays())

        stages = []
        if right:  # materialise X = A^T (A row-major n x m) with the identity
            stages.append(G1(X, m, n, n, [(A, m, 1, I, LDI, 0, n, 1.0)]))
        for ps in range(3):
            Rp = Rs[ps]
            zero = R.MixList()
            zero.add(Rp, L.F64, 1, 1, n * n, [])      # R strictly-lower blocks stay 0
            stages.append(('m', zero.arrays()))
            stages.append(G1(G, n, n, n, [(X, n, 1, X, n, 0, m, 1.0)]))
            if ps == 0:
                stages.append(G1(tr, 1, 1, 1, [(G, n + 1, 1, ones, 1, 0, n, 1.0)]))
            for k in range(nb):
                b0, bk = k * B, min(B, n - k * B)
                t = b0 + bk
                stages.append(('c', np.array([(G + 8 * (b0 * n + b0), Rp + 8 * (b0 * n + b0),
                                               Ri + 8 * (b0 * n + b0), st,
                                               tr if ps == 0 else 0, bk, n,
                                               shift if ps == 0 else 0.0)], dtype=L.CHOL_JOB)))
                if t < n:
                    stages.append(G1(Rp + 8 * (b0 * n + t), bk, n - t, n,
                                     [(Ri + 8 * (b0 * n + b0), n, 1, G + 8 * (b0 * n + t), n, 0, bk,
                                       1.0)]))
                    stages.append(G1(G + 8 * (t * n + t), n - t, n - t, n,
                                     [(Rp + 8 * (b0 * n + t), n, 1, Rp + 8 * (b0 * n + t), n, 0, bk,
                                       -1.0)], c0=G + 8 * (t * n + t), ldc0=n, beta=1.0))
            for k in range(nb):
                b0, bk = k * B, min(B, n - k * B)
                segs = [(X, n, 0, Rp + 8 * b0, n, 0, b0, -1.0)] if b0 > 0 else []
                stages.append(G1(Wt, m, bk, bk, segs, c0=X + 8 * b0, ldc0=n, beta=1.0))
                stages.append(G1(X + 8 * b0, m, bk, n, [(Wt, bk, 0, Ri + 8 * (b0 * n + b0), n, 0,
                                                         bk, 1.0)]))
        if right:  # A_q <- Q^T
            stages.append(G1(A, n, m, m, [(I, LDI, 0, X, n, 1, n, 1.0)]))
        stages.append(G1(R21, n, n, n, [(Rs[1], n, 0, Rs[0], n, 0, n, 1.0)]))
        stages.append(G1(self.Rq.data_ptr(), n, n, n, [(Rs[2], n, 0, R21, n, 0, n, 1.0)]))
        ph['cq'] = stages

    def _use_cholqr(self, m, n):
        return m >= n and 8 <= n <= L.CHOL_NMAX

    def _use_cholqr_blocked(self, m, n):
        return m >= n and n > L.CHOL_NMAX

    def _plan_gauge(self, ph, q, dirn, copy_back):
        d, dims = self.d, self.dims
        if dirn > 0:
            Dl, Dm, Dn = dims[q], dims[q + 1], dims[q + 2]
            m, nn = Dl * d, Dm
            k = min(m, nn)
            if k > L.QR_KMAX:
                raise ValueError('bond dimension %d exceeds the QR kernel limit %d' % (k, L.QR_KMAX))
            if self._use_cholqr(m, nn):
                self._plan_cholqr(ph, q, m, nn, right=False)
            elif self._use_cholqr_blocked(m, nn):
                self._plan_cholqr_blocked(ph, q, m, nn, right=False)
            else:
                ph['qr'] = np.array([(self.mps.ptr(q), self.mps.ptr(q),
                                      self.Rq.data_ptr(), self.qrw.data_ptr(), m, nn, nn, 0, 0, 0)],
                                    dtype=L.QR_JOB)
            g = R.GemmList()
            g.add(self.init.data_ptr(), k, d * Dn, d * Dn,
                  [(self.Rq.data_ptr(), nn, 0, self.mps.ptr(q + 1), d * Dn, 0, Dm, 1.0)])
            ph['absorb'] = g.arrays()
            if copy_back:
                mx = R.MixList()
                mx.add(self.mps.ptr(q + 1), L.F64, 1, 1, k * d * Dn,
                       [(self.init.data_ptr(), L.F64, 1.0, -1)])
                ph['copy'] = mx.arrays()
            self.dims[q + 1] = k
        else:
            Dp, Dm, Dr = dims[q - 1], dims[q], dims[q + 1]
            m, nn = d * Dr, Dm
            k = min(m, nn)
            if k > L.QR_KMAX:
                raise ValueError('bond dimension %d exceeds the QR kernel limit %d' % (k, L.QR_KMAX))
            if self._use_cholqr(m, nn):
                self._plan_cholqr(ph, q, m, nn, right=True)
            elif self._use_cholqr_blocked(m, nn):
                self._plan_cholqr_blocked(ph, q, m, nn, right=True)
            else:
                ph['qr'] = np.array([(self.mps.ptr(q), self.mps.ptr(q),
                                      self.Rq.data_ptr(), self.qrw.data_ptr(), m, nn, m, 1, 1, 0)],
                                    dtype=L.QR_JOB)
            g = R.GemmList()
            g.add(self.init.data_ptr(), Dp * d, k, k,
                  [(self.mps.ptr(q - 1), Dm, 0, self.Rq.data_ptr(), Dm, 1, Dm, 1.0)])
            ph['absorb'] = g.arrays()
            if copy_back:
                mx = R.MixList()
                mx.add(self.mps.ptr(q - 1), L.F64, 1, 1, Dp * d * k,
                       [(self.init.data_ptr(), L.F64, 1.0, -1)])
                ph['copy'] = mx.arrays()
            self.dims[q] = k

    def _plan_theta(self, ph, p):
        """theta = A_p A_{p+1} (row-major Dl d x d Dr) into the Lanczos start
        vector of the pair (p, p+1)."""
        d, dims = self.d, self.dims
        Dl, D, Dr = dims[p], dims[p + 1], dims[p + 2]
        g = R.GemmList()
        g.add(self.init.data_ptr(), Dl * d, d * Dr, d * Dr,
              [(self.mps.ptr(p), D, 0, self.mps.ptr(p + 1), d * Dr, 0, D, 1.0)])
        ph['absorb'] = g.arrays()

    def _plan_gate(self, ph, p, k):
        """theta = A_p A_{p+1} into the start-vector buffer, then the pair
        tensor ``th`` = G theta with G the solve's gate k acting on the
        merged physical index (``evolve_gate_tebd``, ``MPSClass.py:1416-1429``,
        for a nearest-neighbour pair: the two gate halves contracted over
        their bond are the gate itself)."""
        self._plan_theta(ph, p)
        d2, dims = self.d * self.d, self.dims
        mx = R.MixList()
        mx.add(self.th.data_ptr(), L.F64, dims[p], d2, dims[p + 2],
               [(self.init.data_ptr(), L.F64, 1.0, self.op_index2(self.gate_base + k))])
        ph['copy'] = mx.arrays()

    def _plan_split(self, ph, p, dirn):
        """Split the pair tensor theta (``th``, M = row-major m x n, m = Dl d,
        n = d Dr) into A_p (m x k) A_{p+1} (k x n) with k = min(chi, m, n):
        dir > 0 leaves A_p left-orthonormal and the centre on p+1, dir < 0
        leaves A_{p+1} right-orthonormal and the centre on p (the truncated
        two-site update of ``MPSClass.py:894-922``, on a finite chain).

        Orders N = max(m, n) >= 8: the block Jacobi SVD of X = [M | 0] or
        [M ; 0] (dir > 0) or the same padding of M^T (dir < 0), N x N, gives
        X = u^T v with the singular values sorted, so
        - dir > 0: A_p = (u[:k, :m])^T, A_{p+1} = v[:k, :n];
        - dir < 0: A_p = (v[:k, :m])^T, A_{p+1} = u[:k, :n].
        A zero padding row never rotates (its products with every row are
        0), so the kept rows of u stay inside the first m (n) coordinates
        even when M is rank-deficient.  The copies into X and the transposed
        factors are GEMMs with the identity.  Smaller pairs are never
        truncated (k = min(m, n)): an identity factor when M already has the
        kept shape, else Householder QR of M or M^T."""
        d, dims, chi = self.d, self.dims, self.chi
        Dl, Dr = dims[p], dims[p + 2]
        m, n = Dl * d, d * Dr
        k = min(chi, m, n)
        N = max(m, n)
        th = self.th.data_ptr()
        A, B = self.mps.ptr(p), self.mps.ptr(p + 1)
        dsq = self.dsq_max
        I, LDI = _p(self.cholb, 5 * dsq), self.dmax_max
        if N > L.SVDJ_NMAX:
            raise ValueError('two-site split of order %d exceeds the SVD limit %d (chi too large)'
                             % (N, L.SVDJ_NMAX))
        self.dims[p + 1] = k
        if N < 8:
            if k < min(m, n):
                raise ValueError('two-site truncation of a %d x %d pair to %d needs chi >= 4' % (m, n, k))
            stages = []
            g, mx = R.GemmList(), R.MixList()
            if dirn > 0 and m <= n:      # A_p = 1, A_{p+1} = M
                g.add(A, m, m, m, [(I, LDI, 0, I, LDI, 0, m, 1.0)])
                mx.add(B, L.F64, 1, 1, m * n, [(th, L.F64, 1.0, -1)])
            elif dirn < 0 and n <= m:    # A_{p+1} = 1, A_p = M
                g.add(B, n, n, n, [(I, LDI, 0, I, LDI, 0, n, 1.0)])
                mx.add(A, L.F64, 1, 1, m * n, [(th, L.F64, 1.0, -1)])
            elif dirn > 0:               # M = Q R: A_p = Q (m x n), A_{p+1} = R
                ph['qr'] = np.array([(th, A, B, self.qrw.data_ptr(), m, n, n, 0, 0, 0)],
                                    dtype=L.QR_JOB)
            else:                        # M^T = Q R: A_{p+1} = Q^T (m x n), A_p = R^T
                ph['qr'] = np.array([(th, B, self.Rq.data_ptr(), self.qrw.data_ptr(), n, m, n, 1,
                                      1, 0)], dtype=L.QR_JOB)
                g.add(A, m, m, m, [(I, LDI, 0, self.Rq.data_ptr(), m, 1, m, 1.0)])
            if len(g.arrays()[0]):
                stages.append(('g', g.arrays()))
            if len(mx.arrays()[0]):
                stages.append(('m', mx.arrays()))
            if stages:
                ph['cq'] = stages
            return
        X = th
        if not (dirn > 0 and m == n):
            X = _p(self.qrw, 2 * self.psi_max)
            stages = []
            if m != n:
                zero = R.MixList()
                zero.add(X, L.F64, 1, 1, N * N, [])
                stages.append(('m', zero.arrays()))
            g = R.GemmList()
            if dirn > 0:   # X[:m, :n] = M
                g.add(X, m, n, N, [(I, LDI, 0, th, n, 0, m, 1.0)])
            else:          # X[:n, :m] = M^T
                g.add(X, n, m, N, [(I, LDI, 0, th, n, 1, n, 1.0)])
            stages.append(('g', g.arrays()))
            ph['cq'] = stages
        U, V = _p(self.cholb, 3 * dsq), _p(self.cholb, 4 * dsq)
        bw, nblk = L.svdj_blocking(N)
        ph['svdj'] = np.array([(X, _p(self.cholb, 0), U, V, self.Rq.data_ptr(), N, N, bw, nblk,
                                0, 0)], dtype=L.SVDJ_JOB)
        g = R.GemmList()
        if dirn > 0:
            g.add(A, m, k, k, [(I, LDI, 0, U, N, 1, m, 1.0)])
            g.add(B, k, n, n, [(V, N, 0, I, LDI, 0, n, 1.0)])
        else:
            g.add(A, m, k, k, [(I, LDI, 0, V, N, 1, m, 1.0)])
            g.add(B, k, n, n, [(U, N, 0, I, LDI, 0, n, 1.0)])
        ph['svd_post_g'] = g.arrays()

    def _plan_env(self, ph, q, dirn):
        """Block environment through the (new) isometry at site q."""
        T, d, dims = self.T, self.d, self.dims
        site = T.sites[q]
        Qp = self.mps.ptr(q)
        lc, g1, mx, gT = R.MixList(), R.GemmList(), R.MixList(), R.GemmList()
        psi = self.psi_max
        Y = lambda t: _p(self.Yb, t * psi)
        W = lambda t: _p(self.Wb, t * psi)
        LC = lambda t: _p(self.lc64, t * self.dsq_max)
        if dirn > 0:
            Dl, k = dims[q], dims[q + 1]
            groups = site.lm
            src_opens = T.left_opens[q]
            dst_opens = T.left_opens[q + 1]
            sidx = {o: t for t, o in enumerate(src_opens)}
            for g, (opb, grp) in enumerate(groups):
                lc.add(LC(g), L.F64, Dl, 1, Dl, [(self.L(q, 1 + l), L.F64, c, -1) for l, c in grp])
            ny = 0
            yH = None
            if q > 0:
                yH = ny
                g1.add(Y(ny), Dl, d * k, d * k, [(self.L(q, 0), Dl, 0, Qp, d * k, 0, Dl, 1.0)])
                ny += 1
            yG = []
            for g in range(len(groups)):
                g1.add(Y(ny), Dl, d * k, d * k, [(LC(g), Dl, 0, Qp, d * k, 0, Dl, 1.0)])
                yG.append(ny)
                ny += 1
            terms = []
            if yH is not None:
                terms.append((Y(yH), L.F64, 1.0, -1))
            for g, (opb, _) in enumerate(groups):
                terms.append((Y(yG[g]), L.F64, 1.0, self.op_index(opb)))
            for (kop, c) in site.field:
                terms.append((Qp, L.F64, c, self.op_index(kop)))
            mx.add(W(0), L.F64, Dl, d, k, terms)
            gT.add(self.L(q + 1, 0), k, k, k, [(Qp, k, 1, W(0), k, 0, Dl * d, 1.0)])
            nw = 1
            for t, (i, a) in enumerate(dst_opens):
                if i < q:
                    g1.add(Y(ny), Dl, d * k, d * k,
                           [(self.L(q, 1 + sidx[(i, a)]), Dl, 0, Qp, d * k, 0, Dl, 1.0)])
                    gT.add(self.L(q + 1, 1 + t), k, k, k, [(Qp, k, 1, Y(ny), k, 0, Dl * d, 1.0)])
                    ny += 1
                else:
                    mx.add(W(nw), L.F64, Dl, d, k, [(Qp, L.F64, 1.0, self.op_index(a))])
                    gT.add(self.L(q + 1, 1 + t), k, k, k, [(Qp, k, 1, W(nw), k, 0, Dl * d, 1.0)])
                    nw += 1
        else:
            k, Dr = dims[q], dims[q + 1]
            groups = site.mr
            src_opens = T.right_opens[q + 1]
            dst_opens = T.right_opens[q]
            sidx = {o: t for t, o in enumerate(src_opens)}
            for g, (opa, grp) in enumerate(groups):
                lc.add(LC(g), L.F64, Dr, 1, Dr, [(self.Rr(q + 1, 1 + r), L.F64, c, -1) for r, c in grp])
            ny = 0
            yH = None
            if q < self.n - 1:
                yH = ny
                g1.add(Y(ny), k * d, Dr, Dr, [(Qp, Dr, 0, self.Rr(q + 1, 0), Dr, 1, Dr, 1.0)])
                ny += 1
            yG = []
            for g in range(len(groups)):
                g1.add(Y(ny), k * d, Dr, Dr, [(Qp, Dr, 0, LC(g), Dr, 1, Dr, 1.0)])
                yG.append(ny)
                ny += 1
            terms = []
            if yH is not None:
                terms.append((Y(yH), L.F64, 1.0, -1))
            for g, (opa, _) in enumerate(groups):
                terms.append((Y(yG[g]), L.F64, 1.0, self.op_index(opa)))
            for (kop, c) in site.field:
                terms.append((Qp, L.F64, c, self.op_index(kop)))
            mx.add(W(0), L.F64, k, d, Dr, terms)
            gT.add(self.Rr(q, 0), k, k, k, [(Qp, d * Dr, 0, W(0), d * Dr, 1, d * Dr, 1.0)])
            nw = 1
            for t, (j, b) in enumerate(dst_opens):
                if j > q:
                    g1.add(Y(ny), k * d, Dr, Dr,
                           [(Qp, Dr, 0, self.Rr(q + 1, 1 + sidx[(j, b)]), Dr, 1, Dr, 1.0)])
                    gT.add(self.Rr(q, 1 + t), k, k, k, [(Qp, d * Dr, 0, Y(ny), d * Dr, 1, d * Dr, 1.0)])
                    ny += 1
                else:
                    mx.add(W(nw), L.F64, k, d, Dr, [(Qp, L.F64, 1.0, self.op_index(b))])
                    gT.add(self.Rr(q, 1 + t), k, k, k, [(Qp, d * Dr, 0, W(nw), d * Dr, 1, d * Dr, 1.0)])
                    nw += 1
        ph['env_lc'] = lc.arrays()
        ph['env_g1'] = g1.arrays()
        ph['env_mix'] = mx.arrays()
        ph['env_gT'] = gT.arrays()

    def _plan_cast_block(self, ph, b, left, ti):
        D = self.dims[b]
        cnt = 1 + (self.nL[b] if left else self.nR[b])
        mx = R.MixList()
        for m in range(cnt):
            src = self.L(b, m) if left else self.Rr(b, m)
            dst = self.L_lp(b, m, ti) if left else self.R_lp(b, m, ti)
            mx.add(dst, ti, D, 1, D, [(src, L.F64, 1.0, -1)])
        ph['cast'] = mx.arrays()

    def cast_all_records(self, ti):
        """Mixer records copying every environment block to its ``ti`` copy,
        kept by the engine per (arena, layout, dtype, bond dimensions): a
        later solve on the same arena with the same shape reuses them."""
        key = (self.base, self.sig, ti, tuple(self.dims))
        cache = self.engine._cast_cache
        hit = cache.get(key)
        if hit is not None:
            return hit
        if len(cache) > 4096:
            cache.clear()
        mx = R.MixList()
        for b in range(self.n + 1):
            D = self.dims[b]
            for m in range(1 + self.nL[b]):
                mx.add(self.L_lp(b, m, ti), ti, D, 1, D, [(self.L(b, m), L.F64, 1.0, -1)])
            for m in range(1 + self.nR[b]):
                mx.add(self.R_lp(b, m, ti), ti, D, 1, D, [(self.Rr(b, m), L.F64, 1.0, -1)])
        cache[key] = mx.arrays()
        return cache[key]

    def _plan_local(self, ph, p, prec, pair=False):
        """Matvec records and the Lanczos entry of the local eigensolve at
        site p, or (``pair``) at the merged sites p, p+1: physical dimension
        d^2, right block at bond p+2, merged operators, result into theta."""
        ti, tc, mode1, mode2, floor_rel = PREC[prec]
        T, dims = self.T, self.dims
        if pair:
            site, d, rb, opi, out = T.pairs[p], self.d * self.d, p + 2, self.op_index2, self.th.data_ptr()
        else:
            site, d, rb, opi, out = T.sites[p], self.d, p + 1, self.op_index, self.mps.ptr(p)
        Dl, Dr = dims[p], dims[rb]
        n_el = Dl * d * Dr
        psi, dsq = self.psi_max, self.dsq_max
        isz = ITEMSIZE[ti]
        Xp = self.x.data_ptr()
        LCp = lambda t: self.lc_lp.data_ptr() + t * dsq * isz
        Tp = lambda t: self.Tb.data_ptr() + t * psi * isz
        base = self.Ub.data_ptr()
        Up = lambda t: self.Ub.data_ptr() + psi * 8 + t * psi * isz
        lc = R.MixList()
        nlm, nmr = len(site.lm), len(site.mr)
        for g, (_, grp) in enumerate(site.lm):
            lc.add(LCp(g), ti, Dl, 1, Dl, [(self.L(p, 1 + l), L.F64, c, -1) for l, c in grp])
        for g, (_, grp) in enumerate(site.mr):
            lc.add(LCp(nlm + g), ti, Dr, 1, Dr, [(self.Rr(rb, 1 + r), L.F64, c, -1) for r, c in grp])
        for u, (raw, grp) in enumerate(site.strad):
            if site.strad_left:
                lc.add(LCp(nlm + nmr + u), ti, Dr, 1, Dr,
                       [(self.Rr(rb, 1 + r), L.F64, c, -1) for r, c in grp])
            else:
                lc.add(LCp(nlm + nmr + u), ti, Dl, 1, Dl,
                       [(self.L(p, 1 + l), L.F64, c, -1) for l, c in grp])
        g1, mx, g2 = R.GemmList(), R.MixList(), R.GemmList()
        nt = 0
        tH = None
        if p > 0:
            tH = nt
            g1.add(Tp(nt), Dl, d * Dr, d * Dr, [(self.L_lp(p, 0, ti), Dl, 0, Xp, d * Dr, 0, Dl, 1.0)])
            nt += 1
        tG = []
        for g in range(nlm):
            g1.add(Tp(nt), Dl, d * Dr, d * Dr, [(LCp(g), Dl, 0, Xp, d * Dr, 0, Dl, 1.0)])
            tG.append(nt)
            nt += 1
        tS = []
        for u, (raw, _) in enumerate(site.strad):
            F = self.L_lp(p, 1 + raw, ti) if site.strad_left else LCp(nlm + nmr + u)
            g1.add(Tp(nt), Dl, d * Dr, d * Dr, [(F, Dl, 0, Xp, d * Dr, 0, Dl, 1.0)])
            tS.append(nt)
            nt += 1
        terms = []
        if tH is not None:
            terms.append((Tp(tH), ti, 1.0, -1))
        for g, (opb, _) in enumerate(site.lm):
            terms.append((Tp(tG[g]), ti, 1.0, opi(opb)))
        for (kop, c) in site.field:
            terms.append((Xp, ti, c, opi(kop)))
        mx.add(base, tc, Dl, d, Dr, terms)
        for g, (opa, _) in enumerate(site.mr):
            mx.add(Up(g), ti, Dl, d, Dr, [(Xp, ti, 1.0, opi(opa))])
        segs = []
        if rb < self.n:
            segs.append((Xp, Dr, 0, self.R_lp(rb, 0, ti), Dr, 1, Dr, 1.0))
        for g in range(nmr):
            segs.append((Up(g), Dr, 0, LCp(nlm + g), Dr, 1, Dr, 1.0))
        for u, (raw, _) in enumerate(site.strad):
            B = LCp(nlm + nmr + u) if site.strad_left else self.R_lp(rb, 1 + raw, ti)
            segs.append((Tp(tS[u]), Dr, 0, B, Dr, 1, Dr, 1.0))
        g2.add(self.w.data_ptr(), Dl * d, Dr, Dr, segs, c0=base, ldc0=Dr, beta=1.0)
        ph['mv_lc'] = lc.arrays()
        ph['mv_g1'] = g1.arrays()
        ph['mv_mix'] = mx.arrays()
        ph['mv_g2'] = g2.arrays()
        tol = self.eigs_tol if self.eigs_tol > 0 else 2.220446049250313e-16
        # the basis of a low-precision matvec is stored in f32 (half the
        # traffic of the Krylov passes; the reductions stay f64)
        vdt = L.F64 if prec == 'fp64' else L.F32
        ph['lz'] = np.array([(self.V.data_ptr(), self.w.data_ptr(), Xp, self.init.data_ptr(),
                              out, n_el, psi, min(self.engine.ncv, n_el), 0,
                              tol / self.tau, self.tau, floor_rel, ti, tc, vdt, 0)],
                            dtype=L.LANCZOS_SOLVE)
        ph['prec'] = prec
        ph['mv_flops'] = self._matvec_flops(p, pair)

    def _matvec_flops(self, p, pair=False):
        """Algorithmic flops of one matvec at site p, or at the merged pair
        p, p+1 (GEMM terms executed)."""
        rb = p + 2 if pair else p + 1
        s = self.T.pairs[p] if pair else self.T.sites[p]
        Dl, Dr, d = self.dims[p], self.dims[rb], self.d * self.d if pair else self.d
        n1 = (p > 0) + len(s.lm) + len(s.strad)
        n2 = (rb < self.n) + len(s.mr) + len(s.strad)
        return 2.0 * d * Dl * Dr * (Dl * n1 + Dr * n2)

    def _plan_svd(self, ph, q, dirn):
        """SVD centre move (``left2right/right2left_decompose_tensor`` with
        way='svd', TensorBasicModule.py:438-444,474-480): the spectrum goes
        to row ``bond`` of the device array ``lmv``.  Orders 8..1024 take
        CholeskyQR3 of the panel then the multi-workgroup block Jacobi SVD
        of its R (``_plan_svd_block``); smaller ones the single-workgroup
        Jacobi SVD of the panel."""
        d, dims = self.d, self.dims
        bond = q if dirn > 0 else q - 1
        lm_ptr = _p(self.lmv, bond * self.dmax_max)
        nn = dims[q + 1] if dirn > 0 else dims[q]
        m = dims[q] * d if dirn > 0 else d * dims[q + 1]
        if m < nn:
            raise ValueError('SVD move needs the panel to be tall (site %d: %d x %d)' % (q, m, nn))
        ph['lm_bond'] = bond
        ph['svd_n'] = nn
        if 8 <= nn <= L.SVDJ_NMAX:
            self._plan_svd_block(ph, q, dirn, m, nn, lm_ptr)
            return
        if dirn > 0:
            Dl, Dm, Dn = dims[q], dims[q + 1], dims[q + 2]
            ph['svd'] = np.array([(self.mps.ptr(q), self.mps.ptr(q), lm_ptr,
                                   _p(self.Rq, nn), self.qrw.data_ptr(), m, nn, nn, 0, 0, 0)],
                                 dtype=L.SVD_JOB)
            g = R.GemmList()
            g.add(self.init.data_ptr(), nn, d * Dn, d * Dn,
                  [(_p(self.Rq, nn), nn, 0, self.mps.ptr(q + 1), d * Dn, 0, Dm, 1.0)])
            mx = R.MixList()
            mx.add(self.mps.ptr(q + 1), L.F64, 1, 1, nn * d * Dn,
                   [(self.init.data_ptr(), L.F64, 1.0, -1)])
        else:
            Dp, Dm, Dr = dims[q - 1], dims[q], dims[q + 1]
            ph['svd'] = np.array([(self.mps.ptr(q), self.mps.ptr(q), lm_ptr,
                                   _p(self.Rq, nn), self.qrw.data_ptr(), m, nn, m, 1, 1, 0)],
                                 dtype=L.SVD_JOB)
            g = R.GemmList()
            g.add(self.init.data_ptr(), Dp * d, nn, nn,
                  [(self.mps.ptr(q - 1), Dm, 0, _p(self.Rq, nn), nn, 1, Dm, 1.0)])
            mx = R.MixList()
            mx.add(self.mps.ptr(q - 1), L.F64, 1, 1, Dp * d * nn,
                   [(self.init.data_ptr(), L.F64, 1.0, -1)])
        ph['absorb'] = g.arrays()
        ph['copy'] = mx.arrays()

    def _rsv_slot(self, b):
        return _p(self.rsv, b * rsv_stride(self.dsq_max, self.dmax_max))

    def qmove_ok(self, q):
        """A right move at site q can keep its R for the spectrum: R is square
        (the bond dimension does not shrink) and of order at least 8 (the
        block Jacobi's smallest); other bonds move by SVD."""
        n = self.dims[q + 1]
        return n >= 8 and self.dims[q] * self.d >= n

    def _plan_save_r(self, ph, q):
        """The n x n factor R of the right QR move at site q (row-major, n the
        bond dimension, ``qmove_ok``) into the slot of bond q: the bond's
        Schmidt values are its singular values (A_q = Q R, Q an isometry)."""
        n = self.dims[q + 1]
        assert self.dims[q] * self.d >= n, 'a saved R factor must be square'
        mx = R.MixList()
        mx.add(self._rsv_slot(q), L.F64, 1, 1, n * n, [(self.Rq.data_ptr(), L.F64, 1.0, -1)])
        ph['rsave'] = mx.arrays()

    def _plan_spectra(self, ph, b0, b1):
        """Singular values (values-only block Jacobi, one job per bond) of the
        R factors saved for bonds b0..b1 into the spectrum rows of those
        bonds (their orders: the bond dimensions)."""
        jobs = []
        for b in range(b0, b1 + 1):
            n = self.dims[b + 1]
            bw, nblk = L.svdj_blocking(n, values_only=True)
            slot = self._rsv_slot(b)
            jobs.append((slot, slot, 0, 0, _p(self.lmv, b * self.dmax_max), n, n, bw, nblk, 0, 0))
        ph['svdj'] = np.array(jobs, dtype=L.SVDJ_JOB)
        ph['lm_bonds'] = [(b, self.dims[b + 1]) for b in range(b0, b1 + 1)]

    def _plan_svd_block(self, ph, q, dirn, m, nn, lm_ptr):
        """A = Q R (CholeskyQR3 of the panel, Q in place of A_q), R = u^T v
        (``tnx_svd_block_jacobi``), then A_q <- Q u^T and the neighbour
        absorbs v.  Scratch in the CholeskyQR buffer below its constants:
        B, W, column norms and the scale at 0 .. 2 dsq + n + 1; u at 3 dsq; v at 4 dsq."""
        d, dims = self.d, self.dims
        if nn <= L.CHOL_NMAX:
            self._plan_cholqr(ph, q, m, nn, right=dirn < 0)
        else:
            self._plan_cholqr_blocked(ph, q, m, nn, right=dirn < 0)
        dsq = self.dsq_max
        U, V = _p(self.cholb, 3 * dsq), _p(self.cholb, 4 * dsq)
        bw, nblk = L.svdj_blocking(nn)
        ph['svdj'] = np.array([(self.Rq.data_ptr(), _p(self.cholb, 0), U, V, lm_ptr, nn, nn, bw,
                                nblk, 0, 0)], dtype=L.SVDJ_JOB)
        A = self.mps.ptr(q)
        T1, T2 = self.qrw.data_ptr(), self.init.data_ptr()
        g, mx = R.GemmList(), R.MixList()
        if dirn > 0:   # A_q (m x n) holds Q: A_q <- Q u^T, A_{q+1} <- v A_{q+1}
            Dn = dims[q + 2]
            nb = self.mps.ptr(q + 1)
            g.add(T1, m, nn, nn, [(A, nn, 0, U, nn, 1, nn, 1.0)])
            g.add(T2, nn, d * Dn, d * Dn, [(V, nn, 0, nb, d * Dn, 0, nn, 1.0)])
            mx.add(A, L.F64, 1, 1, m * nn, [(T1, L.F64, 1.0, -1)])
            mx.add(nb, L.F64, 1, 1, nn * d * Dn, [(T2, L.F64, 1.0, -1)])
        else:          # A_q (n x m) holds Q^T: A_q <- u Q^T, A_{q-1} <- A_{q-1} v^T
            Dp = dims[q - 1]
            nb = self.mps.ptr(q - 1)
            g.add(T1, nn, m, m, [(U, nn, 0, A, m, 0, nn, 1.0)])
            g.add(T2, Dp * d, nn, nn, [(nb, nn, 0, V, nn, 1, nn, 1.0)])
            mx.add(A, L.F64, 1, 1, nn * m, [(T1, L.F64, 1.0, -1)])
            mx.add(nb, L.F64, 1, 1, Dp * d * nn, [(T2, L.F64, 1.0, -1)])
        ph['svd_post_g'] = g.arrays()
        ph['svd_post_m'] = mx.arrays()

    # ----------------------------------------------------------- host view
    def host_mps(self):
        self.engine.settle()
        out = []
        for s in range(self.n):
            a = self.mps[s][:self.dims[s] * self.d * self.dims[s + 1]].cpu().numpy()
            out.append(a.reshape(self.dims[s], self.d, self.dims[s + 1]).copy())
        return out


# ------------------------------------------------------ batch descriptors
def _parts(pairs, key):
    """Phase ``key`` of each ``(solve, plan)`` that has it, at the solve's
    addresses."""
    return [s.part(ph, key) for s, ph in pairs if key in ph]


def assemble_pre(pairs):
    """Descriptor arrays (one launch per phase over the batch) and launch
    metadata of the pre-Lanczos phases of one step of every ``(solve, plan)``.
    Array order: qr, svd, 3 x (gemm problems, segments), 4 x (mix jobs,
    terms), the CholeskyQR stages, the SVD-move records, then the
    normalisations.  (The fifth mix phase saves R factors for the spectrum,
    after the CholeskyQR stages; the normalisations run first.)"""
    qr = R.concat_plain(_parts(pairs, 'qr'), L.QR_JOB)
    svd = R.concat_plain(_parts(pairs, 'svd'), L.SVD_JOB)
    gemms = [R.concat_gemm(_parts(pairs, k)) for k in ('absorb', 'env_g1', 'env_gT')]
    mixes = [R.concat_mix(_parts(pairs, k))
             for k in ('copy', 'env_lc', 'env_mix', 'cast', 'rsave')]
    arrs = [qr, svd]
    for p, sg, _ in gemms:
        arrs += [p, sg]
    for j, t, _ in mixes:
        arrs += [j, t]
    cq_meta = []
    groups_cq = collections.OrderedDict()   # same stage structure -> one launch per stage
    for c in _parts(pairs, 'cq'):
        groups_cq.setdefault(tuple(kind for kind, _ in c), []).append(c)
    for sig, cqs in groups_cq.items():
        for k, kind in enumerate(sig):
            stage = [c[k][1] for c in cqs]
            if kind == 'g':
                p, sg, nt = R.concat_gemm(stage)
                cq_meta.append(('g', len(p), nt, len(arrs), gemm_kmax(p, sg)))
                arrs += [p, sg]
            elif kind == 'm':
                j, t, nb = R.concat_mix(stage)
                cq_meta.append(('m', len(j), nb, len(arrs), 0))
                arrs += [j, t]
            else:
                jobs = R.concat_plain(stage, L.CHOL_JOB)
                cq_meta.append(('c', len(jobs), 0, len(arrs), 0))
                arrs.append(jobs)
    # block-Jacobi SVD moves: jobs (pair prefix sums over the batch), then
    # the GEMMs and copies that apply u and v
    svdj = R.concat_plain(_parts(pairs, 'svdj'), L.SVDJ_JOB)
    svdj_meta = None
    if len(svdj):
        halves = svdj['nblk'] // 2
        svdj['pair_begin'] = np.concatenate([[0], np.cumsum(halves)[:-1]])
        svdj_meta = (len(svdj), int(halves.sum()), int(svdj['nblk'].max()),
                     int(svdj['n'].max()), len(arrs))
        arrs.append(svdj)
        pg, ps, pnt = R.concat_gemm(_parts(pairs, 'svd_post_g'))
        pj, pt, pnb = R.concat_mix(_parts(pairs, 'svd_post_m'))
        svdj_meta += ((len(pg), pnt, len(arrs), gemm_kmax(pg, ps)), (len(pj), pnb, len(arrs) + 2))
        arrs += [pg, ps, pj, pt]
    norm = R.concat_plain(_parts(pairs, 'norm'), L.NORM_JOB)
    norm_meta = (len(norm), len(arrs)) if len(norm) else None
    if len(norm):
        arrs.append(norm)
    meta = {'nqr': len(qr), 'nsvd': len(svd),
            'gemm': [(len(p), nt, gemm_kmax(p, sg)) for p, sg, nt in gemms],
            'mix': [(len(j), nb) for j, _, nb in mixes], 'cq': cq_meta, 'svdj': svdj_meta,
            'norm': norm_meta}
    return arrs, meta


def _lanczos_groups(pairs):
    """The ``(solve, plan)`` pairs that run a Lanczos solve, one list per
    matvec precision (in a fixed precision order)."""
    by = {}
    for s, ph in pairs:
        if 'lz' in ph:
            by.setdefault(ph['prec'], []).append((s, ph))
    return [by[p] for p in sorted(by)]


def _split_groups(groups, nstreams, split_min):
    """The Lanczos calls of a step: all ``groups`` in one call, or (two
    streams and at least ``split_min`` solves) every precision group halved
    into two calls."""
    n = sum(len(g) for g in groups)
    if nstreams < 2 or n < split_min:
        return [groups]
    a = [g[:(len(g) + 1) // 2] for g in groups]
    b = [g[(len(g) + 1) // 2:] for g in groups]
    return [[g for g in a if g], [g for g in b if g]]


def _lanczos_key(groups):
    return ('lz',) + tuple(tuple((ph['uid'],) + s.reloc_key for s, ph in g) for g in groups)


def assemble_lanczos(groups):
    """Descriptor arrays and metadata of one batched Lanczos call over
    ``groups``, lists of ``(solve, plan)`` that share a precision: the
    linear-combination mix, the solve records (in group order, with their
    chunk prefix sums) and, per group, the three matvec phases whose
    ``active`` field is the solve's index in the whole batch.  Array order:
    mix jobs, mix terms, solve records, then 6 per group."""
    flat = [sp for g in groups for sp in g]
    lc_j, lc_t, lc_nb = R.concat_mix([s.part(ph, 'mv_lc') for s, ph in flat])
    lz = R.concat_plain([s.part(ph, 'lz') for s, ph in flat], L.LANCZOS_SOLVE)
    chunks = (lz['n'] + L.LZ_CHUNK - 1) // L.LZ_CHUNK
    lz['chunk_begin'] = np.concatenate([[0], np.cumsum(chunks)[:-1]])
    arrs = [lc_j, lc_t, lz]
    gmeta = []
    first = 0
    for g in groups:
        actives = list(range(first, first + len(g)))
        first += len(g)
        g1p, g1s, g1t = R.concat_gemm([s.part(ph, 'mv_g1') for s, ph in g], actives)
        mxj, mxt, mxb = R.concat_mix([s.part(ph, 'mv_mix') for s, ph in g], actives)
        g2p, g2s, g2t = R.concat_gemm([s.part(ph, 'mv_g2') for s, ph in g], actives)
        _, _, mode1, mode2, _ = PREC[g[0][1]['prec']]
        gmeta.append((mode1, mode2, (len(g1p), g1t, gemm_kmax(g1p, g1s)), (len(mxj), mxb),
                      (len(g2p), g2t, gemm_kmax(g2p, g2s)),
                      # algorithmic work of the three launches (all solves active)
                      (R.gemm_flops(g1p, g1s), R.mix_bytes(mxj, mxt), R.gemm_flops(g2p, g2s))))
        arrs += [g1p, g1s, mxj, mxt, g2p, g2s]
    meta = {'lc': (len(lc_j), lc_nb), 'nchunks': int(chunks.sum()), 'groups': gmeta}
    return arrs, meta


class Engine:
    """Owns the device, the stream, the operator table and the launch loop.

    With ``plan_ahead`` (default), the host work of step k+1 (planning,
    relocating the shared plans, assembling the descriptor sets) runs on a
    worker thread while step k's Lanczos call, which holds no GIL, keeps the
    GPU busy.  Every method that reads solve state first waits for that work
    (``settle``).  The interpreter's thread switch interval is lowered to
    0.2 ms the first time, so the main thread gets the GIL back promptly when
    the Lanczos call returns."""

    def __init__(self, precision_schedule=None, ncv=20, max_restarts=20, min_iters=1,
                 device='cuda:0', lowp=True, ops_capacity=4096, plan_ahead=True,
                 cycle_steps=False, autotune=True, lanczos_streams=2, cycle_restarts=2,
                 init_rng='numpy'):
        require_gpu()
        # initial random MPS: 'numpy' draws RandomState(seed) on the host in
        # the reference's order (TensorBasicModule.py:179-211); 'device' draws
        # the same shapes of standard normals on the GPU (a Philox stream
        # seeded by the solve's seed), with no host RNG or staging
        if init_rng not in ('numpy', 'device'):
            raise ValueError("init_rng must be 'numpy' or 'device'")
        self.init_rng = init_rng
        self._moved_pool = {}        # reloc key -> relocated plan parts of a freed solve
        self.moved_pool_limit = 256
        self._stage_consts = {}      # layout id -> device image of the CholeskyQR constants
        self._pending_fill = []      # initial-MPS fill jobs not yet launched
        self.lib = L.load()
        self.device = torch.device(device)
        self._dev_index = self.device.index if self.device.index is not None else \
            (torch.cuda.current_device() if self.device.type == 'cuda' else 0)
        self.ncv = int(ncv)
        if not 2 <= self.ncv <= L.LZ_NCV_MAX:
            raise ValueError('ncv must be in [2, %d]' % L.LZ_NCV_MAX)
        self.max_restarts = int(max_restarts)
        self.min_iters = int(min_iters)
        self.precision_schedule = precision_schedule or (lambda sweep, nsweeps: 'fp64')
        self.uses_lowp = bool(lowp)
        self.up = Uploader(self.device)
        self._stager = _Stager()
        self.ops_capacity = ops_capacity
        # operator table: slot s holds site operator k at element (s + k) d^2;
        # merged two-site operators (d^4 elements, d <= 2) sit in a second
        # region, operator j of slot s at element (s + ops_off2 + j) d^4
        self.ops = torch.zeros(ops_capacity * 20, dtype=torch.float64, device=self.device)
        self.ops_free = [(0, ops_capacity)]
        self.ops_d2 = None
        self.ops_off2 = None
        self.solves = []
        self.lz_states = [None, None]   # Lanczos state buffers of the (at most two) calls
        self.lanczos_streams = int(lanczos_streams)
        if self.lanczos_streams not in (1, 2):
            raise ValueError('lanczos_streams must be 1 or 2')
        self.split_min = 16   # smallest batch split over two streams
        # run_queues: a step whose eigensolves cannot be re-queued has the
        # next step's recorded pre-Lanczos launches queued by its Lanczos call
        self.chain_next = True
        self._lz_stream = None
        self._lz_pool = None
        self._resident = collections.OrderedDict()
        self._resident_bytes = 0
        self.resident_limit = 1 << 30   # descriptor sets of every distinct step of a sweep
        self.plan_cache = {}
        self._cast_cache = {}   # (arena, layout, dtype, dims) -> block cast records
        self.svd_flag = torch.zeros(64, dtype=torch.int32, device=self.device)  # rotation flag
        self._arena_pool = {}   # bytes -> freed arenas (reused by the next solve of that size)
        self.counters = {'matvec_flops': 0.0, 'matvecs': 0, 'lanczos_calls': 0, 'steps': 0,
                         'lz_time': 0.0, 't_plan': 0.0, 't_pre': 0.0, 't_lz_total': 0.0,
                         'lz_iters': 0, 't_settle': 0.0, 'ahead_used': 0, 'resumes': 0}
        self.plan_ahead = bool(plan_ahead)
        # kernel variant per grouped-GEMM shape class, timed on first use
        self.tuner = GemmTuner(self.lib) if autotune else None
        # cycle_steps: each Lanczos call runs at most 1 + cycle_restarts
        # Krylov cycles (implicit restarts inside the call); a solve that has
        # not converged then continues from its Ritz vector in the next step
        # ('resume') while the others move on, so no call waits long for the
        # batch's slowest solve
        self.cycle_steps = bool(cycle_steps)
        self.cycle_restarts = int(cycle_restarts)
        self._ahead_pool = None
        self._ahead = None     # future of the worker's next-step host work
        self._spec = None      # its speculative descriptor sets

    # ------------------------------------------------------ planning ahead
    def _prec_of(self, s, st):
        """Matvec precision of step ``st`` of solve ``s`` (moves: f64)."""
        if st.kind not in ('solve', 'resume', 'solve2', 'resume2'):
            return 'fp64'
        return self.sweep_precision(s, st.sweep)

    def sweep_precision(self, s, sweep):
        """Matvec precision of sweep ``sweep`` of solve ``s``: the schedule's,
        unless the solve has been switched to f64 (``s.prec_override``)."""
        if s.prec_override is not None:
            return s.prec_override
        return self.precision_schedule(max(sweep, 0), s.para.get('sweep_time', 1))

    def settle(self):
        """Wait for the worker's next-step host work (a no-op without it)."""
        if self._ahead is not None:
            t0 = time.perf_counter()
            fut, self._ahead = self._ahead, None
            self._spec = fut.result()   # re-raises a planning error here
            self.counters['t_settle'] += time.perf_counter() - t0

    def _start_ahead(self, cands):
        if not cands:
            return
        if self._ahead_pool is None:
            import sys
            if sys.getswitchinterval() > 2e-4:
                sys.setswitchinterval(2e-4)
            self._ahead_pool = concurrent.futures.ThreadPoolExecutor(1)
        self._ahead = self._ahead_pool.submit(self._plan_ahead, cands)

    def _plan_ahead(self, cands):
        """Worker: plan each candidate's next queued step (leaving its
        ``dims`` as they are; the step applies ``dims_after``), relocate the
        plans, and assemble the descriptor sets the next step would need if
        it runs exactly these solves."""
        pairs = []
        for s in cands:
            st = s.queue[0]
            dims0 = s.dims
            s.dims = list(dims0)
            try:
                prec = self._prec_of(s, st)
                ph = s.plan(st, prec)
            finally:
                s.dims = dims0
            for key in ph['packed'].trees:
                s.part(ph, key)
            s._next = (st, prec, ph)
            pairs.append((s, ph))
        spec = {}
        rkey = ('pre',) + tuple((ph['uid'],) + s.reloc_key for s, ph in pairs)
        if rkey not in self._resident:
            spec['pre'] = (rkey, assemble_pre(pairs))
        groups = _lanczos_groups(pairs)
        if groups:
            spec['lz'] = []
            for part in _split_groups(groups, self.lanczos_streams, self.split_min):
                rkey = _lanczos_key(part)
                if rkey not in self._resident:
                    spec['lz'].append((rkey, assemble_lanczos(part)))
        return spec

    # ------------------------------------------------------------ op table
    def alloc_ops(self, ops, ops2=None):
        """A slot of the operator table for a solve's site operators ``ops``
        (n, d, d) and, for two-site sweeps, its merged operators ``ops2``
        (n2, d^2, d^2); the slot spans max(n, n2) entries, so the merged
        regions of different slots do not overlap."""
        d2 = ops.shape[1] * ops.shape[2]
        if self.ops_d2 is None:
            self.ops_d2 = d2
            self.ops_off2 = -(-self.ops_capacity * d2 // (d2 * d2))
        if d2 != self.ops_d2:
            raise ValueError('all solves of an engine must share the physical dimension')
        cnt = ops.shape[0] if ops2 is None else max(ops.shape[0], ops2.shape[0])
        if ops2 is not None and (self.ops_off2 + self.ops_capacity) * d2 * d2 > self.ops.numel():
            raise ValueError('merged two-site operators need d <= 2')
        for t, (s, c) in enumerate(self.ops_free):
            if c >= cnt:
                self.ops_free[t] = (s + cnt, c - cnt)
                flat = torch.from_numpy(np.ascontiguousarray(ops).reshape(-1)).to(self.device)
                self.ops[s * d2:(s + ops.shape[0]) * d2].copy_(flat)
                if ops2 is not None:
                    d4 = d2 * d2
                    o = (s + self.ops_off2) * d4
                    flat = torch.from_numpy(np.ascontiguousarray(ops2).reshape(-1)).to(self.device)
                    self.ops[o:o + ops2.shape[0] * d4].copy_(flat)
                self._ops_owner = getattr(self, '_ops_owner', {})
                self._ops_owner[s] = cnt
                return s
        raise RuntimeError('operator table full')

    def free_ops(self, slot):
        """Return a slot; the free list stays sorted and coalesced, so the
        next allocation takes the lowest free slot (a batch admitted after
        another of the same shape gets the same slots back)."""
        cnt = self._ops_owner.pop(slot, 0)
        if not cnt:
            return
        free = sorted(self.ops_free + [(slot, cnt)])
        merged = [free[0]]
        for s, c in free[1:]:
            ps, pc = merged[-1]
            if ps + pc == s:
                merged[-1] = (ps, pc + c)
            else:
                merged.append((s, c))
        self.ops_free = merged

    def stream(self):
        """Torch's current stream on the engine's device, as a HIP stream
        handle (the raw handle is read directly: building a Stream object
        per launch costs more than the launch)."""
        if _RAW_STREAM is not None:
            return ctypes_void(_RAW_STREAM(self._dev_index))
        return ctypes_void(torch.cuda.current_stream().cuda_stream)

    # ------------------------------------------------------------- arenas
    def arena_get(self, nbytes):
        pool = self._arena_pool.get(nbytes)
        if pool:
            return pool.pop(0)   # first released, first reused: a repeated batch keeps its arenas
        return torch.empty(nbytes, dtype=torch.uint8, device=self.device)

    def arena_fill(self, arena, stage):
        """Arena prefix <- ``stage`` (pinned uint8 tensor, copied without a
        host sync; the pinned block is not reused before the copy ends); the
        rest of the arena <- 0.  Both are ordered on the engine's stream."""
        n = stage.numel()
        arena[:n].copy_(stage, non_blocking=True)
        arena[n:].zero_()

    def arena_fill_device(self, arena, lay, n, d, dims, seed):
        """The arena of a solve whose initial MPS is drawn on the device:
        zero, the CholeskyQR constants of the staged prefix, then every site
        (dims[s], d, dims[s+1]) filled with standard normals of the
        counter-based stream of ``seed`` (sites in the reference's draw
        order: 0, n-1, then 1..n-2), queued as jobs of one batched fill
        launch (``flush_fill``)."""
        arena.zero_()
        key = id(lay)
        got = self._stage_consts.get(key)
        if got is None or got[0] is not lay:
            o = lay['offsets'][('cholb', None)]
            img = make_stage(lay, [], pin=False)[o:lay['stage_bytes']]
            order = [0, n - 1] + list(range(1, n - 1)) if n > 1 else [0]
            offs = np.array([lay['offsets'][('mps', s)] for s in order], dtype=np.int64)
            got = (lay, o, img.to(self.device), np.array(order), offs)
            self._stage_consts[key] = got
        _, o, img, order, offs = got
        arena[o:o + img.numel()].copy_(img, non_blocking=True)
        dims = np.asarray(dims, dtype=np.int64)
        cnt = dims[order] * d * dims[order + 1]
        jobs = np.zeros(len(order), dtype=L.FILL_JOB)
        jobs['dst'] = arena.data_ptr() + offs
        jobs['n'] = cnt
        jobs['seed'] = np.uint64(int(seed) & ((1 << 64) - 1))
        jobs['offset'] = np.cumsum(cnt) - cnt
        self._pending_fill.append(jobs)

    def flush_fill(self):
        """Launch the queued initial-MPS fills (one launch for all)."""
        if not self._pending_fill:
            return
        jobs = np.concatenate(self._pending_fill)
        self._pending_fill = []
        (dj,) = self.up.upload([jobs])
        L.check(self.lib.tnx_fill_normal_f64(ctypes_void(dj), len(jobs), self.stream()),
                'tnx_fill_normal_f64')

    def arena_release(self, arena):
        self._arena_pool.setdefault(arena.numel(), []).append(arena)

    # ------------------------------------------------- resident descriptors
    def _pre_program(self, ent):
        """The launches of a step's pre-Lanczos phases over resident
        descriptor set ``ent``, in order, as tnx_run_ops records: the
        normalisations, QR / SVD gauge moves, CholeskyQR3 stages, kept R
        factors, block Jacobi SVD with its products, the absorption of R,
        the copy-back and the environment update (GEMM, mixer, GEMM), the
        low-precision copies.  GEMM variants are the tuner's at build time
        (the caller rebuilds when the tuner's table changes)."""
        ptr, meta = ent.ptrs, ent.meta
        opsp = int(self.ops.data_ptr())
        recs = []

        def gemm(a, n_p, b, nt, kmax):
            if nt:
                recs.append((L.OP_GEMM, self._gemm_mode(L.GEMM_F64, nt, kmax), n_p, nt, 0, 0,
                             a, b, 0))

        def mix(a, n_j, b, nb):
            if nb:
                recs.append((L.OP_MIX, 0, n_j, nb, 0, 0, a, b, opsp))

        def run_gemm(i):
            n_p, nt, kmax = meta['gemm'][i]
            gemm(ptr[2 + 2 * i], n_p, ptr[3 + 2 * i], nt, kmax)

        def run_mix(i):
            n_j, nb = meta['mix'][i]
            mix(ptr[8 + 2 * i], n_j, ptr[9 + 2 * i], nb)

        if meta.get('norm'):
            recs.append((L.OP_NORM, 0, meta['norm'][0], 0, 0, 0, ptr[meta['norm'][1]], 0, 0))
        if meta['nqr']:
            recs.append((L.OP_QR, 0, meta['nqr'], 0, 0, 0, ptr[0], 0, 0))
        if meta['nsvd']:
            recs.append((L.OP_SVD, 60, meta['nsvd'], 0, 0, 0, ptr[1], 0, 0))
        for (kind, cnt, nt, ai, kmax) in meta['cq']:   # CholeskyQR3 stages
            if kind == 'g':
                gemm(ptr[ai], cnt, ptr[ai + 1], nt, kmax)
            elif kind == 'm':
                mix(ptr[ai], cnt, ptr[ai + 1], nt)
            elif cnt:
                recs.append((L.OP_CHOL, 0, cnt, 0, 0, 0, ptr[ai], 0, 0))
        run_mix(4)       # R factors kept for the spectrum (qmove steps)
        if meta['svdj'] is not None:   # SVD of the CholeskyQR factors, then apply u and v
            nj, npairs, max_nblk, max_n, ai, (gp, gnt, gi, gk), (mj, mnb, mi) = meta['svdj']
            recs.append((L.OP_SVDJ, 60, nj, npairs, max_nblk, max_n, ptr[ai],
                         int(self.svd_flag.data_ptr()), 0))
            gemm(ptr[gi], gp, ptr[gi + 1], gnt, gk)
            mix(ptr[mi], mj, ptr[mi + 1], mnb)
        run_gemm(0)      # absorb R into the neighbour
        run_mix(0)       # copy back (moves)
        run_mix(1)       # linear combinations for the environment update
        run_gemm(1)      # X . Q
        run_mix(2)       # physical mixing
        run_gemm(2)      # Q^T . W
        run_mix(3)       # low-precision copies of the new block
        return np.array(recs, dtype=L.OP_REC) if recs else np.zeros(0, dtype=L.OP_REC)

    def _resident_get(self, key):
        ent = self._resident.get(key)
        if ent is not None:
            self._resident.move_to_end(key)
        return ent

    def _resident_put(self, key, arrays, meta):
        ent = _Resident(self.device, arrays, meta, self._stager)
        self._resident[key] = ent
        self._resident_bytes += ent.buf.numel()
        while self._resident_bytes > self.resident_limit and len(self._resident) > 1:
            _, old = self._resident.popitem(last=False)
            self._resident_bytes -= old.buf.numel()
        return ent

    # ------------------------------------------------------------- solves
    def add_solve(self, para, seed=0, mps=None, tag=None, stage=None, gates=0):
        """A solve of ``para``; ``stage``: its initial image from
        :meth:`prepare` (default: drawn and staged here); ``gates``: merged
        two-site operators reserved for time evolution."""
        s = Solve(self, para, seed=seed, mps=mps, tag=tag, stage=stage, gates=gates)
        self.flush_fill()
        self.solves.append(s)
        return s

    def prepare(self, para, seed=0):
        """The host work of ``add_solve(para, seed)`` that needs no device
        state: term tables and layout, the seeded initial MPS drawn (numpy's
        generator runs without the GIL) and staged in pinned memory.  Safe on
        a worker thread; the result goes to ``add_solve(..., stage=...)``.
        With ``init_rng='device'`` there is nothing to stage (None)."""
        if self.init_rng == 'device':
            return None
        T, _, _, lay = solve_geometry(para, self.ncv, self.uses_lowp)
        return make_stage(lay, random_mps_host(T.n, T.d, initial_chi(para), seed),
                          pin=self.device.type != 'cpu')

    def add_solves(self, items, threads=16):
        """Solves of ``items = [(para, seed), ...]``; the seeded initial MPS
        draws (host RNG, which runs without the GIL) go to a thread pool, or
        run on the GPU with ``init_rng='device'``."""
        if self.init_rng == 'device':   # drawn on the GPU, one fill launch for the batch
            out = []
            for para, seed in items:
                s = Solve(self, para, seed=seed)
                self.solves.append(s)
                out.append(s)
            self.flush_fill()
            return out
        geoms = [solve_geometry(para, self.ncv, self.uses_lowp) for para, _ in items]

        def draw(k):
            (para, seed), (T, _, _, lay) = items[k], geoms[k]
            return make_stage(lay, random_mps_host(T.n, T.d, initial_chi(para), seed),
                              pin=self.device.type != 'cpu')
        workers = max(1, min(threads, len(items)))
        if workers > 1:
            with concurrent.futures.ThreadPoolExecutor(workers) as ex:
                stages = list(ex.map(draw, range(len(items))))
        else:
            stages = [draw(k) for k in range(len(items))]
        out = []
        for (para, seed), st in zip(items, stages):
            s = Solve(self, para, seed=seed, stage=st)
            self.solves.append(s)
            out.append(s)
        return out

    def footprint(self, para):
        """HBM bytes a solve of ``para`` allocates (admission control)."""
        return solve_geometry(para, self.ncv, self.uses_lowp)[3]['bytes']

    def finalize(self, solves):
        """Total energies <H> of finished solves (one batched measurement)."""
        from .observe import measure
        reqs = []
        for s in solves:
            idx2 = np.asarray(s.para['index2'])
            reqs.append(([(i, 1) for i in range(s.n)] + [(i, 3) for i in range(s.n)],
                         [tuple(r) for r in idx2]))
        out = []
        for s, (vone, vtwo, _) in zip(solves, measure(self, solves, reqs)):
            e = float(np.dot(np.asarray(s.para['coeff2'], dtype=float).reshape(-1), vtwo))
            e -= s.para.get('hx', 0.0) * vone[:s.n].sum() + s.para.get('hz', 0.0) * vone[s.n:].sum()
            out.append(e)
        return out

    def remove_solve(self, s):
        self.settle()
        if s in self.solves:
            self.solves.remove(s)
        s.free()

    # ------------------------------------------------------------- launch
    def _gemm(self, mode, parts, actives=None):
        p, sg, nt = R.concat_gemm(parts, actives)
        if nt == 0:
            return
        dp, ds = self.up.upload([p, sg])
        L.check(self.lib.tnx_gemm_grouped(mode, ctypes_void(dp), len(p), ctypes_void(ds), nt, None,
                                          self.stream()), 'tnx_gemm_grouped')

    def _mix(self, parts):
        j, t, nb = R.concat_mix(parts)
        if nb == 0:
            return
        dj, dt = self.up.upload([j, t])
        L.check(self.lib.tnx_mix(ctypes_void(dj), len(j), ctypes_void(dt), nb,
                                 ctypes_void(self.ops.data_ptr()), None, self.stream()), 'tnx_mix')

    def _qr(self, jobs):
        if len(jobs) == 0:
            return
        (dj,) = self.up.upload([jobs])
        L.check(self.lib.tnx_qr_f64(ctypes_void(dj), len(jobs), self.stream()), 'tnx_qr_f64')

    def _svd(self, jobs, max_sweeps=60):
        if len(jobs) == 0:
            return
        (dj,) = self.up.upload([jobs])
        L.check(self.lib.tnx_svd_jacobi_f64(ctypes_void(dj), len(jobs), max_sweeps, self.stream()),
                'tnx_svd_jacobi_f64')

    def _tuned_gemm(self, mode, prob, seg, g):
        """Encoded mode of a matvec GEMM launch ``g = (nprob, ntiles, kmax)``
        at device descriptors ``prob``/``seg``; a new shape class is tuned
        first (a matvec GEMM does not read its own output, so the timing
        launches leave no trace)."""
        if self.tuner is None:
            return mode
        n, nt, kmax = g
        if nt:
            st = self.stream()

            def launch(m):
                L.check(self.lib.tnx_gemm_grouped(m, ctypes_void(prob), n, ctypes_void(seg), nt,
                                                  None, st), 'tnx_gemm_grouped (tuning)')
            self.tuner.tune(mode, nt, kmax, launch)
        return self.tuner.mode_for(mode, nt, kmax)

    def _gemm_mode(self, mode, ntiles, kmax):
        """Encoded mode of a launch whose shape class may have been tuned."""
        return mode if self.tuner is None else self.tuner.mode_for(mode, ntiles, kmax)

    def _lanczos(self, groups, spec=None, ahead=None, parts=None, chained=None, sites=None):
        """One batched Lanczos solve over ``groups``: lists of (solve, plan),
        one list per matvec precision (each Krylov step runs every group's
        matvec, then the shared f64 kernels).  ``spec``: descriptor sets
        assembled ahead; ``ahead``: solves whose next step the worker plans
        while the calls run.

        With ``lanczos_streams = 2`` a large batch is split in two halves
        (every precision group halved) that run as two Lanczos calls on two
        streams, each driven by its own host thread: one half's matvec
        kernels (latency bound) overlap the other half's passes over its
        Krylov basis (bandwidth bound).  A solve's arithmetic does not depend
        on the other solves of its call, so the results are the same bits.
        ``parts``: the split already made, as ``(groups, key)`` pairs.
        ``chained``: recorded launches (the next step's pre-Lanczos phases)
        that a single call queues behind its own kernels.  ``sites``: the
        eigensolve site of every solve (by ``id``), for the warnings."""
        if parts is None:
            parts = [(part, _lanczos_key(part))
                     for part in _split_groups(groups, self.lanczos_streams, self.split_min)]
        got = spec.get('lz') if spec else None
        preps = [self._lanczos_prepare(k, part, got, rkey)
                 for k, (part, rkey) in enumerate(parts)]
        if ahead:
            self._start_ahead(ahead)
        restarts = self.cycle_restarts if self.cycle_steps else self.max_restarts
        t0 = time.perf_counter()
        if len(preps) == 1:
            rcs = [self._lanczos_call(preps[0], self.stream(), restarts, chained)]
        else:
            # the second stream starts after everything queued so far
            if self._lz_stream is None:
                self._lz_stream = torch.cuda.Stream(device=self.device)
            ev = torch.cuda.Event()
            ev.record()
            self._lz_stream.wait_event(ev)
            if self._lz_pool is None:
                self._lz_pool = concurrent.futures.ThreadPoolExecutor(1)
            fut = self._lz_pool.submit(self._lanczos_call, preps[1],
                                       ctypes_void(self._lz_stream.cuda_stream), restarts)
            rc0 = self._lanczos_call(preps[0], self.stream(), restarts)
            rcs = [rc0, fut.result()]   # that call ends with its stream synchronised
        self.counters['lz_time'] += time.perf_counter() - t0
        for rc, msg in rcs:
            if rc not in (L.OK, L.ENOCONV):
                raise L.HipError('tnx_lanczos_run_multi failed (rc=%d): %s' % (rc, msg))
        self.counters['lanczos_calls'] += 1
        for prep in preps:
            # the sites of the eigensolves (a next step planned already has
            # moved _lz_site)
            prep['sites'] = ([sites[id(s)] for s, _ in prep['flat']] if sites is not None
                             else [s._lz_site for s, _ in prep['flat']])
        return preps

    def _resume_possible(self, preps):
        """Whether booking these results may re-queue an eigensolve (then the
        next step's composition is not known before the booking)."""
        return self.cycle_steps and any(bool((p['res']['converged'] == 0).any())
                                        for p in preps)

    def _lanczos_book(self, preps):
        """Per-solve booking of Lanczos results: statistics, Ritz values,
        warnings and the re-queue flag of a solve that did not converge
        within its cycle."""
        cnt = self.counters
        for prep in preps:
            res = prep['res']
            sites = prep['sites']
            iters = res['iters'].tolist()
            conv = res['converged'].tolist()
            theta = res['theta'].tolist()
            # Krylov steps the call ran in lockstep (its slowest solve)
            top = max(iters)
            cnt['lz_iters'] += top
            if top > self.ncv:
                cnt['lz_calls_restarted'] = cnt.get('lz_calls_restarted', 0) + 1
            cnt['matvecs'] += sum(iters)
            for k, (s, ph) in enumerate(prep['flat']):
                it = iters[k]
                s.stats['matvecs'] += it
                cnt['matvec_flops'] += ph['mv_flops'] * it
                if not conv[k] and self.cycle_steps and s.cycles < self.max_restarts:
                    s.cycles += 1
                    s.resume = True   # the step re-queues the eigensolve
                    cnt['resumes'] += 1
                    continue
                s.cycles = 0
                s.last_theta = theta[k]
                s.thetas.append(theta[k])
                if not conv[k]:
                    s.stats['unconverged'] += 1
                    s.warnings.append('local eigensolve at site %d did not converge '
                                      '(residual %.3g after %d matvecs)'
                                      % (sites[k], float(res['resid'][k]), it))

    def _lanczos_prepare(self, k, groups, got, rkey):
        """Resident descriptors (key ``rkey``), the linear-combination mix
        (on the engine's stream), matvec plans with tuned GEMM variants, the
        state buffer and the result array of Lanczos call ``k``."""
        flat = [sp for g in groups for sp in g]
        ent = self._resident_get(rkey)
        if ent is None:
            ahead = dict(got).get(rkey) if got else None
            ent = self._resident_put(rkey, *(ahead if ahead is not None
                                             else assemble_lanczos(groups)))
        ptrs, meta = ent.ptrs, ent.meta
        if meta['lc'][1]:
            L.check(self.lib.tnx_mix(ctypes_void(ptrs[0]), meta['lc'][0], ctypes_void(ptrs[1]),
                                     meta['lc'][1], ctypes_void(self.ops.data_ptr()), None,
                                     self.stream()), 'tnx_mix')
        if ent.host is None:
            # the plan array (GEMM variants tuned on the first build) and the
            # result array, reused every time this descriptor set runs
            plans = (L.MatvecPlan * len(groups))()
            for q, (mode1, mode2, g1, mx, g2, work) in enumerate(meta['groups']):
                b = 3 + 6 * q
                m1, m2 = self._tuned_gemm(mode1, ptrs[b], ptrs[b + 1], g1), \
                    self._tuned_gemm(mode2, ptrs[b + 4], ptrs[b + 5], g2)
                plans[q] = L.MatvecPlan(ptrs[b], ptrs[b + 1], g1[0], g1[1], ptrs[b + 2],
                                        ptrs[b + 3], mx[0], mx[1], ptrs[b + 4], ptrs[b + 5],
                                        g2[0], g2[1], self.ops.data_ptr(), m1, m2,
                                        (ctypes.c_double * 3)(*work))
            need = int(self.lib.tnx_lanczos_state_bytes(len(flat), meta['nchunks'], self.ncv))
            ent.host = (plans, np.zeros(len(flat), dtype=L.LANCZOS_RESULT), need)
        plans, res, need = ent.host
        if self.lz_states[k] is None or self.lz_states[k].numel() < need:
            self.lz_states[k] = torch.empty(max(need, 1 << 20), dtype=torch.uint8,
                                            device=self.device)
        return {'flat': flat, 'plans': plans, 'ngroups': len(groups), 'ptrs': ptrs,
                'nchunks': meta['nchunks'], 'state': self.lz_states[k], 'res': res}

    def _lanczos_call(self, prep, stream, restarts, chained=None):
        """The C Lanczos loop of one prepared call on ``stream`` (holds no
        GIL; returns when the stream's work is done, or with ``chained``
        launches queued behind it, when its results are on the host; their
        return code goes to ``prep['next_rc']``).  Returns ``(rc, msg)``:
        the library keeps its last error per thread, so the message is read
        on the thread that made the call."""
        if chained is None:
            rc = self.lib.tnx_lanczos_run_multi(prep['plans'], prep['ngroups'],
                                                ctypes_void(prep['ptrs'][2]), len(prep['flat']),
                                                prep['nchunks'], self.ncv, restarts,
                                                self.min_iters,
                                                ctypes_void(prep['state'].data_ptr()),
                                                ctypes_void(prep['res'].ctypes.data), stream)
        else:
            nrc = ctypes.c_int32(0)
            rc = self.lib.tnx_lanczos_run_multi_then(
                prep['plans'], prep['ngroups'], ctypes_void(prep['ptrs'][2]), len(prep['flat']),
                prep['nchunks'], self.ncv, restarts, self.min_iters,
                ctypes_void(prep['state'].data_ptr()), ctypes_void(prep['res'].ctypes.data),
                ctypes_void(chained.ctypes.data), len(chained), ctypes.byref(nrc), stream)
            prep['next_rc'] = nrc.value
        msg = '' if rc in (L.OK, L.ENOCONV) else self.lib.tnx_last_error().decode(errors='replace')
        return rc, msg

    def step(self, solves=None):
        """Advance every solve with a queued entry by one step."""
        return self._step(solves, None, False)[0]

    def run_queues(self, solves=None):
        """Run steps until no queue of ``solves`` holds an entry.  Once a
        step's Lanczos call has returned and no eigensolve has to be resumed,
        the next step's pre-Lanczos phases are planned and queued before the
        host books the results, so the device does not wait for that host
        work (the queued kernels follow the eigensolve on the stream).  When
        no eigensolve can be re-queued at all (``cycle_steps`` off) and the
        next step's descriptor set is resident, its launches are queued by
        the Lanczos call itself, right behind its last kernel
        (``chain_next``)."""
        pre = None
        while True:
            n, pre = self._step(solves, pre, True)
            if not n:
                return

    def _step(self, solves, pre, lookahead):
        """One step: its pre-Lanczos phases (``pre``: already queued by the
        previous step), the Lanczos call, then (``lookahead``) the next
        step's pre-Lanczos phases, then the booking.  Returns
        ``(solves stepped, the next step's queued phases or None)``."""
        if pre is None:
            pre = self._step_pre(solves)
        if pre is None:
            return 0, None
        plans, spec, host_miss, t_a, t_b, t_c = pre
        groups = _lanczos_groups([(s, ph) for s, _, ph in plans])
        preps = []
        early = chained = None
        if groups:
            # the next step's host work overlaps this Lanczos call, when this
            # step needed host work (a batch whose descriptor sets are all
            # resident repeats a known sequence and has none to overlap)
            parts = [(part, _lanczos_key(part))
                     for part in _split_groups(groups, self.lanczos_streams, self.split_min)]
            host_miss = host_miss or any(key not in self._resident for _, key in parts)
            ahead = None
            if self.plan_ahead and host_miss:
                ahead = [s for s, _, _ in plans if s.queue]
            sites = None
            if (lookahead and not host_miss and not self.cycle_steps and len(parts) == 1
                    and self.chain_next):
                sites = {id(s): st.site for s, st, _ in plans}
                # no eigensolve can be re-queued: the next step is known now.
                # A resident next step without casts has its launches queued
                # by the Lanczos call behind its own kernels, so the device
                # runs them while the host books this step.
                early = self._pre_plan(solves)
                if early is not None and early['prog'] is not None and not early['casts']:
                    chained = early['prog']
            preps = self._lanczos(groups, spec, ahead, parts, chained, sites)
        t_d = time.perf_counter()
        nxt = None
        if early is not None:
            nxt = self._pre_issue(early, preps[0]['next_rc'] if chained is not None else None)
        elif lookahead and not self._resume_possible(preps):
            nxt = self._step_pre(solves)
        self._lanczos_book(preps)
        self.counters['t_plan'] += t_b - t_a
        self.counters['t_pre'] += t_c - t_b
        self.counters['t_lz_total'] += t_d - t_c
        for s, st, ph in plans:
            s.stats['steps'] += 1
            if s.resume:
                s.resume = False
                kind = 'resume2' if st.kind in ('solve2', 'resume2') else 'resume'
                s.queue.insert(0, Step(kind, st.site, st.dir, st.sweep))
            if st.kind in ('solve', 'resume', 'solve2', 'resume2', 'gate2'):
                s.center = st.site
            elif st.kind in ('move', 'svdmove', 'qmove', 'tmove'):
                s.center = st.site + st.dir
            elif st.kind in ('split2', 'tsplit'):
                s.center = st.site + (st.dir > 0)
            if 'lm_bond' in ph:   # spectrum written to lmv on the device; read lazily
                s._lm_pending[ph['lm_bond']] = ph['svd_n']
            if 'lm_bonds' in ph:
                for b, k in ph['lm_bonds']:
                    s._lm_pending[b] = k
        self.counters['steps'] += 1
        return len(plans), nxt

    def _step_pre(self, solves):
        """Plan the next queued entry of every solve of ``solves`` with one
        and queue the step's pre-Lanczos phases (gauge moves, environment
        update, casts).  Returns the step context, or None without work."""
        ctx = self._pre_plan(solves)
        return None if ctx is None else self._pre_issue(ctx)

    def _pre_plan(self, solves):
        """The host part of ``_step_pre``: pops and plans the next queued
        entry of every solve, finds the step's resident descriptor set and
        its recorded launches (``ctx['prog']``, None before the set is
        resident).  Queues nothing on the device."""
        self.settle()
        spec, self._spec = self._spec, None
        solves = [s for s in (solves or self.solves) if s.queue and s.failed is None]
        if not solves:
            return None
        t_a = time.perf_counter()
        plans = []
        casts = []   # low-precision copies of every block, for solves changing precision
        # solves in lockstep share their step's plan: memo of (precision,
        # plan) per signature, step and bond-dimension object (the value
        # keeps that object alive, so an identical id is the same object)
        memo = {}
        for s in solves:
            st = s.queue.pop(0)
            s._lz_site = st.site
            nx, s._next = s._next, None
            if nx is not None and nx[0] is st:   # planned ahead by the worker
                prec, ph = nx[1], nx[2]
                self.counters['ahead_used'] += 1
            else:
                d = s.dims
                mk = (s.sig, st.kind, st.site, st.dir, st.gate, st.sweep, id(d),
                      s.prec_override, s.para.get('sweep_time', 1))
                hit = memo.get(mk)
                if hit is not None and hit[0] is d:
                    prec, ph = hit[1], hit[2]
                else:
                    prec, ph = self._prec_of(s, st), None
            # the casts read the bond dimensions before the step
            if self.uses_lowp and prec != 'fp64' and s.lowp_dtype != PREC[prec][0]:
                casts.append(s.cast_all_records(PREC[prec][0]))
                s.lowp_dtype = PREC[prec][0]
            if ph is None:
                ph = s.plan(st, prec)
                memo[mk] = (d, prec, ph)
            else:
                s.dims = ph['dims_after_t']
            plans.append((s, st, ph))
        t_b = time.perf_counter()
        # every descriptor of the step's pre-Lanczos phases, resident on the
        # device per combination of (plan, solve) (steps repeat every sweep)
        rkey = ('pre',) + tuple((ph['uid'],) + s.reloc_key for s, _, ph in plans)
        ent = self._resident_get(rkey)
        prog = None
        if ent is not None:
            ver = self.tuner.version if self.tuner is not None else 0
            if ent.host is None or ent.host[1] != ver:
                ent.host = (self._pre_program(ent), ver)
            prog = ent.host[0]
        return {'plans': plans, 'casts': casts, 'spec': spec, 'rkey': rkey, 'ent': ent,
                'prog': prog, 't_a': t_a, 't_b': t_b}

    def _pre_issue(self, ctx, issued_rc=None):
        """Queue the casts and the recorded pre-Lanczos launches of a planned
        step (``issued_rc``: the launches were already queued behind the
        previous Lanczos call, with this return code).  Returns the step
        context of ``_step_pre``."""
        plans, spec, rkey, ent = ctx['plans'], ctx['spec'], ctx['rkey'], ctx['ent']
        t_b = ctx['t_b'] if issued_rc is None else time.perf_counter()
        host_miss = ent is None   # a new batch composition: host work worth overlapping
        if issued_rc is None:
            if ctx['casts']:
                self._mix(ctx['casts'])   # one launch for all of them
            if ent is None:
                got = spec.get('pre') if spec else None
                ent = self._resident_put(rkey, *(got[1] if got and got[0] == rkey else
                                                 assemble_pre([(s, ph) for s, _, ph in plans])))
                ver = self.tuner.version if self.tuner is not None else 0
                ent.host = (self._pre_program(ent), ver)
            prog = ent.host[0]
            rc = self.lib.tnx_run_ops(ctypes_void(prog.ctypes.data), len(prog), self.stream())
        else:
            rc = issued_rc
        if rc == L.ENOCONV:
            # a block Jacobi SVD reached its sweep limit; the rotation flag is
            # shared by the launch: every solve with an SVD job in it is flagged
            nj = ent.meta['svdj'][0]
            self.counters['svd_unconverged'] = self.counters.get('svd_unconverged', 0) + 1
            for s, st, ph in plans:
                if 'svdj' in ph:
                    s.stats['svd_unconverged'] = s.stats.get('svd_unconverged', 0) + 1
                    s.warnings.append('block Jacobi SVD of step %r reached 60 sweeps '
                                      '(%d SVDs in the launch)' % (st, nj))
        else:
            L.check(rc, 'tnx_run_ops (pre-Lanczos phases)')
        t_c = time.perf_counter()
        # t_a..t_b: the planning; t_b..t_c: the queuing (measured after the
        # previous Lanczos call when the launches were queued by that call)
        return plans, spec, host_miss, t_b - (ctx['t_b'] - ctx['t_a']), t_b, t_c
