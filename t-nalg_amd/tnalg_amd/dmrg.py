"""``dmrg_finite_size(para)`` - the drop-in entry of the finite-size DMRG path.

Same call, keys and outputs as the reference's
``algorithms/DMRG_anyH.py:18-101``: ``ob, A, info, para =
dmrg_finite_size(para)`` with ``ob`` holding ``eb_full, mx, mz, e_per_site,
eb, corr_x, corr_z``, ``A`` an MPS object with the attributes and observation
methods of ``MpsOpenBoundaryClass`` (``library/MPSClass.py:28-108``), and
``info`` with ``convergence`` and ``t_cost``.  The sweep schedule, the
convergence test every ``dt_ob`` sweeps and the end-of-run entanglement
spectrum follow the reference; the arithmetic runs on the GPU.

Extra keyword arguments (never read from ``para``): ``seed`` (``None`` draws
the initial MPS from ``numpy.random``'s global state in the reference's
order, like the reference), ``precision`` (``'fp64'`` plain, or a mixed
schedule, see :func:`mixed_schedule`), ``ncv`` (Krylov size, scipy's default
20), ``engine`` (an existing :class:`Engine`).
"""
import copy
import time

import numpy as np

from .parameters import generate_parameters_dmrg, positions_fully_connected


def get_bond_energies(eb_full, positions, index2):
    """Sum the per-row energies per bond (``algorithms/DMRG_anyH.py:338-346``)."""
    eb_full = np.asarray(eb_full).reshape(-1)
    positions = np.asarray(positions).reshape(-1, 2)
    index2 = np.asarray(index2)
    eb = np.zeros((positions.shape[0], 1))
    if eb_full.size == 0 or positions.shape[0] == 0:
        return eb
    # every row adds to every position with its (i, j) (equal positions each get it)
    big = int(max(positions.max(), index2[:, :2].max())) + 1
    pk = positions[:, 0].astype(np.int64) * big + positions[:, 1]
    rk = index2[:eb_full.size, 0].astype(np.int64) * big + index2[:eb_full.size, 1]
    order = np.argsort(pk, kind='stable')
    lo = np.searchsorted(pk[order], rk, side='left')
    hi = np.searchsorted(pk[order], rk, side='right')
    for r in np.nonzero(hi - lo)[0]:
        eb[order[lo[r]:hi[r]], 0] += eb_full[r]
    return eb


def entanglement_entropy(lm, tol=1e-20):
    """``-2 sum lam^2 ln lam`` of the normalised spectrum (TBM:888-904)."""
    lm = np.sort(np.asarray(lm, dtype=float).reshape(-1))[::-1]
    lm = lm[lm > tol]
    lm = lm / np.linalg.norm(lm)
    return float(-2.0 * np.dot(lm ** 2, np.log(lm)))


def entanglement_entropies(lms, tol=1e-20):
    """``entanglement_entropy`` of every spectrum in ``lms`` (-1 for an
    empty one), as an ``(len(lms), 1)`` array, in one vectorised pass."""
    out = np.full((len(lms), 1), -1.0)
    full = [k for k, x in enumerate(lms) if np.size(x)]
    if not full:
        return out
    width = max(np.size(lms[k]) for k in full)
    pad = np.zeros((len(full), width))
    for r, k in enumerate(full):
        x = np.asarray(lms[k], dtype=float).reshape(-1)
        pad[r, :x.size] = x
    pad = np.where(pad > tol, pad, 0.0)
    nrm = np.sqrt(np.sum(pad ** 2, axis=1, keepdims=True))
    p = pad / np.where(nrm > 0, nrm, 1.0)
    with np.errstate(divide='ignore', invalid='ignore'):
        terms = np.where(p > 0, p ** 2 * np.log(np.where(p > 0, p, 1.0)), 0.0)
    out[full, 0] = -2.0 * terms.sum(axis=1)
    return out


def mixed_schedule(low='bf16', mid='fp32', n_fp64=1, n_mid=2):
    """Precision per sweep: ``low`` first, then ``mid`` for ``n_mid`` sweeps,
    the last ``n_fp64`` sweeps in f64."""
    def sched(sweep, nsweeps):
        if sweep >= nsweeps - n_fp64:
            return 'fp64'
        if sweep >= nsweeps - n_fp64 - n_mid:
            return mid
        return low
    return sched


def default_mixed_schedule(sweep, nsweeps):
    """bf16 for the first sweep (from the random start the local problems
    are far from converged and bf16 matvecs are the cheapest), f32 up to the
    last two, f64 for the last two (one sweep when there are fewer than
    four).  A second bf16 sweep buys little: its eigensolves stop at the
    bf16 residual floor, and the first f32 sweep then does the same work.
    Measured on MI355X, N=64, chi=128, 8 sweeps, 64 solves: one bf16 sweep
    27.7 solves/s at |dE/E| <= 1.5e-13 against f64, two bf16 sweeps 26.1;
    every 8-sweep schedule of this shape ends within 1e-11 of the f64 energy
    on 16- and 32-site chains (bf16 for six of eight sweeps does not:
    1.6e-5)."""
    n_fp64 = 2 if nsweeps >= 4 else 1
    n_low = 1 if nsweeps >= 4 else 0
    if sweep >= nsweeps - n_fp64:
        return 'fp64'
    if sweep < n_low:
        return 'bf16'
    return 'fp32'


def _schedule_from(precision):
    if precision in (None, 'fp64'):
        return lambda s, n: 'fp64'
    if precision == 'fp32':
        return mixed_schedule(low='fp32', mid='fp32', n_fp64=1, n_mid=0)
    if precision in ('bf16', 'mixed'):
        return default_mixed_schedule
    if isinstance(precision, str) and precision.startswith('mixed:'):
        # 'mixed:<n_bf16>:<n_fp64>': bf16 for the first n_bf16 sweeps, f64 for
        # the last n_fp64, f32 in between
        n_low, n_fp64 = (int(v) for v in precision.split(':')[1:3])

        def sched(sweep, nsweeps):
            if sweep >= nsweeps - n_fp64:
                return 'fp64'
            return 'bf16' if sweep < n_low else 'fp32'
        return sched
    if callable(precision):
        return precision
    raise ValueError('unknown precision %r' % (precision,))


_REBUILD_ENGINE = []   # engine that re-creates device solves of saved/cleaned MPSs


class MpsOpenBoundaryClass:
    """Result MPS of the drop-in (the reference's ``MpsOpenBoundaryClass``,
    ``library/MPSClass.py:28-108``): host copies of the tensors plus a handle
    to the device solve, which the observation methods use.

    It pickles (``save_pr``) as host data only. After loading, or after
    ``clean_to_save``, the first method that needs the device re-creates the
    device solve from the host tensors and the stored ``para``."""

    def __init__(self, solve, engine, spin='half'):
        self._solve = solve
        self._engine = engine
        self._para = solve.para
        self.spin = spin
        self.decomp_way = 'qr'
        self.refresh_from_device()
        self.ent = np.zeros((self.length - 1, 1))

    def refresh_from_device(self):
        s = self._solve
        self.length = s.n
        self.phys_dim = s.d
        self.mps = s.host_mps()
        self.virtual_dim = list(s.dims)
        self.center = s.center
        self.orthogonality = np.zeros((s.n, 1))
        self.orthogonality[:s.center] = -1
        self.orthogonality[s.center + 1:] = 1
        self.operators = [np.asarray(o) for o in s.para['op']]
        self.lm = [np.asarray(x) for x in s.lm]

    def _device(self):
        """The device solve holding this MPS (re-created from the host
        tensors when it was released or never existed in this process)."""
        if self._solve is None:
            from .engine import Engine
            if self._para is None:
                raise RuntimeError('this MPS has no parameters to rebuild its device state from')
            if not _REBUILD_ENGINE:
                _REBUILD_ENGINE.append(Engine(lowp=False, plan_ahead=False))
            eng = _REBUILD_ENGINE[0]
            s = eng.add_solve(self._para, mps=self.mps)
            s.center = self.center
            s._sched_center = self.center
            for b, x in enumerate(self.lm):   # spectra already known stay known
                s._lm[b] = np.asarray(x, dtype=np.float64).copy()
            self._solve, self._engine = s, eng
        return self._solve

    def calculate_entanglement_spectrum(self, if_fast=True):
        """``MPSClass.py:965-985``: SVD centre moves that fill every empty
        ``lm[b]`` (``if_fast``; otherwise every bond), back to the centre.
        Changes the gauge of ``mps``, not the state."""
        s = self._device()
        if not if_fast:
            for b in range(s.n - 1):
                s._lm[b] = np.zeros(0)
        schedule_spectrum(s)
        self._engine.run_queues([s])
        self.refresh_from_device()

    def calculate_entanglement_entropy(self):
        """``MPSClass.py:987-992``: ``ent[b] = -1`` where no spectrum is known."""
        for i in range(self.length - 1):
            self.ent[i] = -1 if self.lm[i].size == 0 else entanglement_entropy(self.lm[i])

    # -- observation methods (MPSClass.py:1105-1212) --------------------------
    def _measure(self, one, two):
        from .observe import measure
        s = self._device()
        return measure(self._engine, [s], [(one, two)])[0]

    def observation_s1(self, inputs):
        sn, position = inputs
        return self._measure([(position, sn)], [])[0][0]

    def observation_s1_s2(self, inputs):
        ssn, positions = inputs
        return self._measure([], [(positions[0], positions[1], ssn[0], ssn[1])])[1][0]

    def observe_magnetization(self, sn):
        v, _, _ = self._measure([(i, sn) for i in range(self.length)], [])
        return v.reshape(-1, 1)

    def observe_bond_energy(self, index2, coeff2):
        index2 = np.asarray(index2)
        _, v, _ = self._measure([], [tuple(r) for r in index2[:, [0, 1, 2, 3]]])
        return (np.asarray(coeff2, dtype=float).reshape(-1) * v).reshape(-1, 1)

    def observe_bond_energy_from_jxyz(self, pos2, jx, jy, jz, tol=1e-20):
        """``jx<SxSx> + jy Re<SySy> + jz<SzSz>`` per pair (``MPSClass.py:1184-1194``);
        ``Re<SySy>`` is measured through the real rows of
        :func:`observe.real_rows` (``Sy Sy = -(S+ - S-)(S+ - S-)/4``)."""
        pos2 = np.asarray(pos2)
        rows, coefs = [], []
        for n in range(pos2.shape[0]):
            i, j = int(pos2[n, 0]), int(pos2[n, 1])
            for op, c in ((1, jx), (2, jy), (3, jz)):
                if abs(c) > tol:
                    rows.append((i, j, op, op))
                    coefs.append((n, c))
        eb = np.zeros((pos2.shape[0], 1))
        if rows:
            _, v, _ = self._measure([], rows)
            for (n, c), x in zip(coefs, v):
                eb[n] += c * x
        return eb

    def observe_correlators_from_middle(self, op1, op2, ob_len=None):
        pairs = correlator_pairs(self.length, ob_len)
        if not pairs:
            return np.zeros(0)
        _, v, _ = self._measure([], [(a, b, op1, op2) for a, b in pairs])
        return np.asarray(v)

    def norm_mps(self):
        return float(np.sqrt(self._measure([], [])[2]))

    def full_coefficients(self):
        """Dense state vector (small chains only)."""
        v = self.mps[0].reshape(-1, self.mps[0].shape[2])
        for a in self.mps[1:]:
            v = (v @ a.reshape(a.shape[0], -1)).reshape(-1, a.shape[2])
        return v.reshape(-1)

    def clean_to_save(self):
        """Release the device state (``MPSClass.py:1396-1405`` drops caches);
        a later device method re-creates it from the host tensors."""
        if self._solve is not None and self._engine is not None:
            self._engine.remove_solve(self._solve)
        self._solve = None
        self._engine = None

    def __getstate__(self):
        st = dict(self.__dict__)
        st['_solve'] = None
        st['_engine'] = None
        return st


def correlator_pairs(length, ob_len=None):
    """Site pairs of ``observe_correlators_from_middle`` (MPSClass.py:1196-1212)."""
    if ob_len is None:
        ob_len = length
    out = []
    pos1 = int(round(length / 2))
    pos2 = pos1 + 1
    k = 0
    while pos1 > -0.1 and pos2 < length and ob_len > -0.1:
        out.append((pos1, pos2))
        if k % 2 == 0:
            pos1 -= 1
        else:
            pos2 += 1
        k += 1
        ob_len -= 1
    return out


def observe_energy(engine, solve, para):
    """``eb_full, mx, mz, e_per_site`` as in ``DMRG_anyH.py:66-70``."""
    return observe_energies(engine, [solve], [para])[0]


def observe_energies(engine, solves, paras):
    """``observe_energy`` of several solves in one batched measurement."""
    return observe_energies_end(observe_energies_begin(engine, solves, paras))


def observe_energies_begin(engine, solves, paras):
    """The launches of :func:`observe_energies` (values read by
    :func:`observe_energies_end`)."""
    from .observe import measure_begin
    reqs = []
    for s, para in zip(solves, paras):
        idx2 = np.asarray(para['index2'])
        one = [(i, 1) for i in range(s.n)] + [(i, 3) for i in range(s.n)]
        reqs.append((one, [tuple(r) for r in idx2]))
    return solves, paras, measure_begin(engine, solves, reqs,
                                        [('energies', s.sig, s.n) for s in solves])


def observe_energies_end(handle):
    from .observe import measure_end
    solves, paras, h = handle
    out = []
    for s, para, (vone, vtwo, _) in zip(solves, paras, measure_end(h)):
        n = s.n
        ob = {}
        ob['eb_full'] = (np.asarray(para['coeff2'], dtype=float).reshape(-1) * vtwo).reshape(-1, 1)
        ob['mx'] = vone[:n].reshape(-1, 1)
        ob['mz'] = vone[n:].reshape(-1, 1)
        ob['e_per_site'] = (sum(ob['eb_full']) - para['hx'] * sum(ob['mx'])
                            - para['hz'] * sum(ob['mz'])) / n
        out.append(ob)
    return out


def schedule_spectrum(solve):
    """``calculate_entanglement_spectrum(if_fast=True)`` (MPSClass.py:965-985):
    centre moves from the centre to ``p0``, to ``p1 + 1`` and back to the
    centre, where ``[p0, p1]`` spans the bonds without a spectrum.  When every
    bond has one, ``p0 = l - 1`` and ``p1 = 0`` as in the reference, so the
    spectra are recomputed (one-site sweeps move the centre by QR and leave
    earlier spectra in place, in the reference and here).

    The reference makes every move by SVD, so the bonds it crosses twice get
    their spectrum twice, the same Schmidt values of the same state.  Here
    every move is a QR: on the middle leg (``p0`` to ``p1 + 1``, every bond
    once) each move keeps its R factor (``qmove``), and one ``spectra`` step
    takes the singular values of all of them at once (the Schmidt values: the
    centre is Q R with Q an isometry); bonds whose R is not square or of
    order below 8 move by SVD instead.  The spectra are those of the
    reference's middle leg and the state is the same, up to the gauge of the
    tensors off the centre."""
    from .engine import Step
    c = solve.center
    p0, p1 = solve.n - 1, 0
    for b in range(solve.n - 1):
        if solve.lm[b].size == 0:
            p0, p1 = min(p0, b), max(p1, b)
    cur = c
    runs = []   # contiguous ranges of bonds whose R is kept
    for leg, tgt in enumerate((p0, p1 + 1, c)):
        while cur < tgt:
            if leg == 1 and solve.qmove_ok(cur):
                solve.queue.append(Step('qmove', cur, +1))
                if runs and runs[-1][1] == cur - 1:
                    runs[-1][1] = cur
                else:
                    runs.append([cur, cur])
            else:
                solve.queue.append(Step('svdmove' if leg == 1 else 'move', cur, +1))
            cur += 1
        while cur > tgt:
            solve.queue.append(Step('svdmove' if leg == 1 else 'move', cur, -1))
            cur -= 1
    for b0, b1 in runs:
        solve.queue.append(Step('spectra', b0, b1))


def dmrg_finite_size(para=None, seed=None, precision='fp64', ncv=20, engine=None, verbose=False):
    """``algorithms/DMRG_anyH.py:18-101`` on the GPU engine."""
    from .engine import Engine
    t_start = time.time()
    info = {}
    if para is None:
        para = generate_parameters_dmrg()
    eng = engine or Engine(precision_schedule=_schedule_from(precision), ncv=ncv,
                           lowp=precision not in (None, 'fp64'))
    rng = np.random if seed is None else np.random.RandomState(seed)
    s = eng.add_solve(para, seed=rng)
    s.schedule_canonicalise()
    eng.run_queues([s])
    e0 = 0.0
    info['convergence'] = 1
    info['energies'] = []
    ob = {}
    for t in range(para['sweep_time']):
        if_ob = ((t + 1) % para['dt_ob'] == 0) or t == para['sweep_time'] - 1
        s.schedule_sweep(t)
        eng.run_queues([s])
        if if_ob:
            ob = observe_energy(eng, s, para)
            e = float(np.ravel(ob['e_per_site'])[0])
            info['energies'].append(e * s.n)
            info['convergence'] = abs(ob['e_per_site'] - e0)
            if verbose:
                print('sweep %d: e_per_site = %.15g, convergence %g' % (t + 1, e, info['convergence']))
            if info['convergence'] < para['break_tol']:
                if eng.sweep_precision(s, t) == 'fp64':
                    break
                # converged with low-precision matvecs: the remaining sweeps
                # run in f64 and the test must pass again there, so a mixed
                # schedule stops at an f64-converged state like the f64 run
                s.prec_override = 'fp64'
                info['fp64_from_sweep'] = t + 1
            e0 = ob['e_per_site']
    info['sweeps'] = t + 1
    finish_outputs(eng, [s], [para], [ob])
    A = MpsOpenBoundaryClass(s, eng, spin=para.get('spin', 'half'))
    A.calculate_entanglement_entropy()
    info['t_cost'] = time.time() - t_start
    info['stats'] = dict(s.stats)
    # non-fatal failures (an eigensolve or SVD that hit its iteration limit);
    # the reference prints and carries on (DMRG_anyH.py:88-90), so does this
    info['warnings'] = list(s.warnings)
    if s.warnings:
        print('dmrg_finite_size: %d warning(s), first: %s' % (len(s.warnings), s.warnings[0]))
    return ob, A, info, para


def finish_outputs(engine, solves, paras, obs=None):
    """The end of ``dmrg_finite_size`` (``DMRG_anyH.py:66-70,91-95``) for a
    batch of solves that have run their sweeps: ``eb_full, mx, mz,
    e_per_site`` (measured here unless ``obs`` holds them), ``eb``, the
    entanglement spectrum of every bond (SVD centre moves of all solves in
    lockstep), ``ent`` and the correlators ``corr_x`` / ``corr_z`` from the
    middle (one batched measurement).  Fills ``obs`` (the keys of the
    reference's ``ob``) and returns ``(obs, spectra)`` with
    ``spectra[k] = (lm, ent)`` of solve k (``A.lm`` and ``A.ent``)."""
    from .observe import measure_begin, measure_end
    # both measurements are queued first, then the spectrum moves (which
    # change the gauge of the tensors, not the state) run behind them on the
    # stream while the host prepares them
    h_e = None if obs is not None else observe_energies_begin(engine, solves, paras)
    reqs, npairs = [], []
    for s in solves:
        pairs = correlator_pairs(s.n)
        npairs.append(len(pairs))
        reqs.append(([], [(a, b, 1, 1) for a, b in pairs] + [(a, b, 3, 3) for a, b in pairs]))
    h_c = measure_begin(engine, solves, reqs, [('corr', s.n) for s in solves])
    for s in solves:
        schedule_spectrum(s)
    engine.run_queues(solves)
    if h_e is not None:
        obs = observe_energies_end(h_e)
    fetch_spectra(solves)
    spectra = []
    for s, para, ob, k, (_, v, _) in zip(solves, paras, obs, npairs, measure_end(h_c)):
        ob['eb'] = get_bond_energies(ob['eb_full'], para['positions_h2'],
                                     np.asarray(para['index2']))
        ob['corr_x'] = np.asarray(v[:k])
        ob['corr_z'] = np.asarray(v[k:])
        lm = [np.asarray(x) for x in s.lm]
        spectra.append((lm, entanglement_entropies(lm)))
    return obs, spectra


def fetch_spectra(solves):
    """Copy the spectra written on the device of every solve to the host
    in one transfer (what reading ``solve.lm`` does for one solve)."""
    import torch
    pend = [s for s in solves if s._lm_pending]
    if len(pend) < 2:
        return
    host = torch.cat([s.lmv for s in pend]).cpu().numpy()
    o = 0
    for s in pend:
        m = s.lmv.numel()
        rows = host[o:o + m].reshape(s.n - 1, s.dmax_max)
        o += m
        for b, k in s._lm_pending.items():
            s._lm[b] = rows[b, :k].copy()
        s._lm_pending = {}


def all_pair_spin_correlations(A):
    """``<S_i . S_j>`` for every pair ``i < j`` (``observe_bond_energy_from_jxyz``
    over ``positions_fully_connected``), as an ``(N, N)`` symmetric matrix."""
    pos = positions_fully_connected(A.length)
    v = A.observe_bond_energy_from_jxyz(pos, 1.0, 1.0, 1.0).reshape(-1)
    out = np.zeros((A.length, A.length))
    out[pos[:, 0], pos[:, 1]] = v
    out[pos[:, 1], pos[:, 0]] = v
    return out
