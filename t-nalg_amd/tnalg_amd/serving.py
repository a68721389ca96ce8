"""Continuous-batching request loop for ground-state solves (no reference
counterpart; the reference runs one ``dmrg_finite_size`` at a time).

Requests (a ``para`` dict + seed) arrive over time.  The server admits a
waiting request as soon as its HBM footprint fits the budget and a slot is
free, keeps its MPS and environments resident on the device, and advances
all admitted solves together: every engine step moves each active solve by
one entry of its own schedule (solves sit at different sites, sweeps, chain
lengths, bond dimensions and precisions), and each phase of the step is one
kernel launch over all of them.  A solve that finishes its sweeps is
retired and its memory freed; the next request takes its place in the
following step.  Per request the server records arrival, admission and
completion times; ``summary()`` gives solves/s and p50/p99 latency.

Host work stays off the step loop where it can:
- a request that has arrived is prepared on a worker thread (term tables,
  the seeded initial MPS drawn and staged in pinned memory) when the engine
  offers ``prepare``;
- the energy of a finished one-site solve is the Ritz value of its last
  local eigensolve (``energy='ritz'``).  Its centre tensor is that Ritz vector
  and the MPS is in mixed canonical form, so <psi|H|psi> = x^T Heff x = theta
  to rounding.  ``energy='measure'`` (and every two-site solve, whose last
  split truncates after the last eigensolve) measures <H> on the device
  instead.

The engine is duck-typed (``add_solve``, ``remove_solve``, ``step``,
``footprint``, ``finalize``; optionally ``prepare``) so the admission logic
can be exercised without a GPU.
"""
import collections
import concurrent.futures
import time

import numpy as np


class Request:
    __slots__ = ('rid', 'para', 'seed', 'arrival', 'admitted', 'done', 'energy', 'solve', 'error',
                 'footprint', 'sweeps_left', 'prep', 'warnings')

    def __init__(self, rid, para, seed, arrival):
        self.rid, self.para, self.seed, self.arrival = rid, para, seed, arrival
        self.admitted = self.done = None
        self.energy = None
        self.solve = None
        self.error = None
        self.footprint = 0
        self.prep = None   # future of the engine's host preparation
        self.warnings = []  # non-fatal failures of the solve's steps (result still given)


class SolveServer:
    def __init__(self, engine, hbm_budget, max_active=256, clock=time.perf_counter,
                 energy='ritz', prepare_ahead=8, lowp_min_chi=64):
        if energy not in ('ritz', 'measure'):
            raise ValueError("energy must be 'ritz' or 'measure'")
        self.energy = energy
        # requests with a bond dimension below this run every sweep in f64:
        # their matvecs are latency bound, so bf16/f32 buys no time, and at
        # small chi the low-precision sweeps leave long chains measurably
        # behind the f64 trajectory after a fixed sweep count (chi = 32,
        # N ~ 80: up to 1.1e-8 relative after 8 sweeps; chi >= 64: <= 5e-10)
        self.lowp_min_chi = int(lowp_min_chi)
        self.engine = engine
        # requests prepared ahead of admission (0: prepare at admission)
        self.prepare_ahead = int(prepare_ahead) if hasattr(engine, 'prepare') else 0
        self._prep_pool = None
        self.budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.clock = clock
        self.waiting = collections.deque()
        self.active = []
        self.finished = []
        self.in_use = 0
        self.t0 = None
        self.steps = 0
        self.peak_active = 0
        self.peak_bytes = 0

    def submit(self, rid, para, seed=0, arrival=None):
        r = Request(rid, para, seed, self.clock() if arrival is None else arrival)
        r.footprint = self.engine.footprint(para)
        if r.footprint > self.budget:
            r.error = 'footprint %d B exceeds the HBM budget %d B' % (r.footprint, self.budget)
            r.done = r.arrival
            self.finished.append(r)
            return r
        self.waiting.append(r)
        return r

    def _admit(self, now):
        admitted = 0
        while self.waiting and self.waiting[0].arrival <= now and len(self.active) < self.max_active:
            r = self.waiting[0]
            if self.in_use + r.footprint > self.budget:
                break  # FIFO: wait for memory to free up
            self.waiting.popleft()
            try:
                if r.prep is not None:
                    stage, r.prep = r.prep.result(), None
                    r.solve = self.engine.add_solve(r.para, seed=r.seed, tag=r.rid, stage=stage)
                else:
                    r.solve = self.engine.add_solve(r.para, seed=r.seed, tag=r.rid)
                if int(r.para.get('chi', 0)) < self.lowp_min_chi and \
                        hasattr(r.solve, 'prec_override'):
                    r.solve.prec_override = 'fp64'
                r.solve.schedule_canonicalise()
                for t in range(int(r.para['sweep_time'])):
                    r.solve.schedule_sweep(t)
            except Exception as e:  # a bad request fails alone
                r.error = '%s: %s' % (type(e).__name__, e)
                r.done = now
                self.finished.append(r)
                continue
            r.admitted = now
            self.in_use += r.footprint
            self.active.append(r)
            admitted += 1
        self.peak_active = max(self.peak_active, len(self.active))
        self.peak_bytes = max(self.peak_bytes, self.in_use)
        return admitted

    def _prepare(self, now):
        """Start the host preparation of the next arrived requests."""
        if not self.prepare_ahead:
            return
        k = 0
        for r in self.waiting:
            if k >= self.prepare_ahead or r.arrival > now:
                break
            k += 1
            if r.prep is None:
                if self._prep_pool is None:
                    self._prep_pool = concurrent.futures.ThreadPoolExecutor(2)
                r.prep = self._prep_pool.submit(self.engine.prepare, r.para, r.seed)

    def _retire(self, now):
        still = []
        done = [r for r in self.active if not r.solve.queue or r.solve.failed is not None]
        if done:
            ok = [r for r in done if r.solve.failed is None]
            meas = []
            for r in ok:
                th = getattr(r.solve, 'last_theta', None)
                if (self.energy == 'ritz' and getattr(r.solve, 'sites', 1) == 1 and th is not None
                        and th == th):
                    r.energy = float(th)
                else:
                    meas.append(r)
            if meas:
                energies = self.engine.finalize([r.solve for r in meas])
                for r, e in zip(meas, energies):
                    r.energy = e
            t = self.clock()
            for r in done:
                if r.solve.failed is not None:
                    r.error = str(r.solve.failed)
                r.warnings = list(getattr(r.solve, 'warnings', ()))
                r.done = t
                self.engine.remove_solve(r.solve)
                r.solve = None
                self.in_use -= r.footprint
                self.finished.append(r)
        for r in self.active:
            if r.done is None:
                still.append(r)
        self.active = still

    def run(self, until_empty=True, max_steps=None, progress=None, every=5.0):
        """Serve until every submitted request has finished.  ``progress``
        (optional) is called with the summary every ``every`` seconds."""
        if self.t0 is None:
            self.t0 = self.clock()
        last = self.clock()
        while self.waiting or self.active:
            if progress is not None and self.clock() - last >= every:
                last = self.clock()
                progress(dict(self.summary(), active=len(self.active), waiting=len(self.waiting)))
            now = self.clock()
            self._admit(now)
            self._prepare(now)
            if not self.active:
                if self.waiting:
                    # idle until the next arrival (a virtual clock is advanced)
                    wait = self.waiting[0].arrival - self.clock()
                    if wait > 0:
                        if hasattr(self.clock, 'advance'):
                            self.clock.advance(wait)
                        else:
                            time.sleep(min(wait, 0.05))
                    continue
                break
            self.engine.step([r.solve for r in self.active])
            self.steps += 1
            self._retire(self.clock())
            if max_steps is not None and self.steps >= max_steps:
                break
        if not self.waiting and self._prep_pool is not None:
            self._prep_pool.shutdown(wait=True)   # every request admitted or failed
            self._prep_pool = None
        return self.summary()

    def summary(self):
        ok = [r for r in self.finished if r.error is None]
        lat = np.array([r.done - r.arrival for r in ok]) if ok else np.zeros(0)
        span = (max(r.done for r in ok) - min(r.arrival for r in ok)) if ok else 0.0
        return {
            'completed': len(ok), 'failed': len(self.finished) - len(ok),
            'with_warnings': sum(1 for r in ok if r.warnings),
            'solves_per_s': len(ok) / span if span > 0 else 0.0,
            'latency_p50_s': float(np.percentile(lat, 50)) if len(lat) else None,
            'latency_p99_s': float(np.percentile(lat, 99)) if len(lat) else None,
            'latency_mean_s': float(lat.mean()) if len(lat) else None,
            'engine_steps': self.steps, 'peak_active': self.peak_active,
            'peak_bytes': self.peak_bytes, 'span_s': span,
        }


def poisson_xxz_stream(n_req, rate, seed=0, n_range=(16, 96), chis=(32, 64, 128),
                       jz_range=(0.5, 1.5), hz_range=(0.0, 0.5), sweeps=8, eigs_tol=1e-10):
    """The serving workload: XXZ chains with N ~ U{16..96}, chi ~ U{32,64,128},
    Jz ~ U[0.5,1.5], hz ~ U[0,0.5], open BC, Poisson arrivals at ``rate``/s.
    Returns a list of ``(arrival_offset_s, para, seed)``."""
    from . import parameters as P
    rng = np.random.RandomState(seed)
    t = 0.0
    out = []
    for k in range(n_req):
        t += rng.exponential(1.0 / rate)
        para = P.parameter_dmrg_chain()
        para.update(P.common_parameters_dmrg())
        para.update(l=int(rng.randint(n_range[0], n_range[1] + 1)), chi=int(rng.choice(chis)),
                    jxy=1.0, jz=float(rng.uniform(*jz_range)), hz=float(rng.uniform(*hz_range)),
                    sweep_time=sweeps, eigs_tol=eigs_tol, tau=1e-4, dt_ob=sweeps, break_tol=0.0)
        para = P.make_consistent_parameter_dmrg(para)
        out.append((t, para, k))
    return out
