"""Admission, slot reuse and latency bookkeeping of the serving loop, driven
by a stand-in engine on a virtual clock (no GPU, no numerics)."""
import numpy as np

from tnalg_amd.serving import SolveServer, poisson_xxz_stream


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def advance(self, dt):
        self.t += dt


class FakeSolve:
    def __init__(self, para, fail_at=None, warn_at=None):
        self.para = para
        self.queue = []
        self.failed = None
        self.warnings = []
        self.fail_at = fail_at
        self.warn_at = warn_at
        self.n = para['l']

    def schedule_canonicalise(self):
        self.queue += ['c'] * 2

    def schedule_sweep(self, t):
        self.queue += ['s'] * (2 * self.n - 2)


class FakeEngine:
    """One step costs 1 ms of virtual time; footprint = chi^2 * N bytes."""

    def __init__(self, clock):
        self.clock = clock
        self.live = set()
        self.max_live_bytes = 0
        self.bytes = 0

    def footprint(self, para):
        return para['chi'] ** 2 * para['l']

    def add_solve(self, para, seed=0, tag=None):
        if para.get('bad'):
            raise ValueError('bad request')
        s = FakeSolve(para, fail_at=para.get('fail_at'), warn_at=para.get('warn_at'))
        self.live.add(s)
        self.bytes += self.footprint(para)
        self.max_live_bytes = max(self.max_live_bytes, self.bytes)
        return s

    def remove_solve(self, s):
        self.live.discard(s)
        self.bytes -= self.footprint(s.para)

    def step(self, solves):
        self.clock.t += 1e-3
        for s in solves:
            if s.queue:
                s.queue.pop(0)
                if s.fail_at is not None and len(s.queue) == s.fail_at:
                    s.failed = RuntimeError('injected device failure')
                if s.warn_at is not None and len(s.queue) == s.warn_at:
                    s.warnings.append('block Jacobi SVD reached 60 sweeps')

    def finalize(self, solves):
        return [-0.44 * s.n for s in solves]


def _para(n, chi, sweeps=2, **kw):
    p = {'l': n, 'chi': chi, 'sweep_time': sweeps}
    p.update(kw)
    return p


def test_admission_respects_budget_and_reuses_slots():
    clk = Clock()
    eng = FakeEngine(clk)
    budget = 3 * 64 * 64 * 20
    srv = SolveServer(eng, hbm_budget=budget, max_active=8, clock=clk)
    for k in range(10):
        srv.submit(k, _para(20, 64), arrival=0.0)
    summ = srv.run()
    assert summ['completed'] == 10 and summ['failed'] == 0
    assert summ['peak_active'] == 3                # budget holds exactly three
    assert eng.max_live_bytes <= budget
    assert not eng.live                             # every slot released
    # FIFO: completion order follows arrival order for identical requests
    done = [r.rid for r in sorted(srv.finished, key=lambda r: r.done)]
    assert done == sorted(done)


def test_poisson_arrivals_and_latency_percentiles():
    clk = Clock()
    eng = FakeEngine(clk)
    srv = SolveServer(eng, hbm_budget=10 ** 9, max_active=4, clock=clk)
    rng = np.random.RandomState(0)
    t = 0.0
    for k in range(40):
        t += rng.exponential(0.05)
        srv.submit(k, _para(int(rng.randint(8, 20)), 32), arrival=t)
    summ = srv.run()
    assert summ['completed'] == 40
    lat = sorted(r.done - r.arrival for r in srv.finished)
    assert abs(summ['latency_p50_s'] - float(np.percentile(lat, 50))) < 1e-12
    assert summ['latency_p99_s'] >= summ['latency_p50_s'] > 0
    for r in srv.finished:
        assert r.admitted >= r.arrival and r.done > r.admitted


def test_failures_are_isolated():
    clk = Clock()
    eng = FakeEngine(clk)
    srv = SolveServer(eng, hbm_budget=10 ** 9, max_active=4, clock=clk)
    srv.submit(0, _para(10, 16), arrival=0.0)
    srv.submit(1, _para(10, 16, bad=True), arrival=0.0)        # rejected at admission
    srv.submit(2, _para(10, 16, fail_at=5), arrival=0.0)       # fails mid-run
    srv.submit(3, _para(10, 10 ** 6), arrival=0.0)             # larger than the budget
    srv.submit(4, _para(12, 16), arrival=0.0)
    summ = srv.run()
    by = {r.rid: r for r in srv.finished}
    assert summ['completed'] == 2 and summ['failed'] == 3
    assert by[0].error is None and by[4].error is None
    assert by[0].energy == -0.44 * 10
    assert 'bad request' in by[1].error
    assert 'injected' in by[2].error
    assert 'budget' in by[3].error
    assert not eng.live


def test_stream_generator_ranges():
    reqs = poisson_xxz_stream(64, rate=5.0, seed=1)
    ns = [p['l'] for _, p, _ in reqs]
    chis = {p['chi'] for _, p, _ in reqs}
    assert min(ns) >= 16 and max(ns) <= 96
    assert chis <= {32, 64, 128}
    t = [a for a, _, _ in reqs]
    assert all(b > a for a, b in zip(t, t[1:]))
    for _, p, _ in reqs:
        jz = np.ravel(p['coeff2'])[2]
        assert 0.5 <= jz <= 1.5 and 0.0 <= p['hz'] <= 0.5


class RitzSolve(FakeSolve):
    def __init__(self, para, fail_at=None):
        super().__init__(para, fail_at)
        self.last_theta = -0.5 * self.n
        self.sites = para.get('dmrg_sites', 1)


class PreparingEngine(FakeEngine):
    """Offers ``prepare``; ``add_solve`` records which stage it was given."""

    def __init__(self, clock):
        super().__init__(clock)
        self.finalized = []
        self.stages = {}

    def prepare(self, para, seed=0):
        return ('stage', para['l'], seed)

    def add_solve(self, para, seed=0, tag=None, stage=None):
        s = super().add_solve(para, seed=seed, tag=tag)
        s.__class__ = RitzSolve
        s.last_theta = -0.5 * s.n
        s.sites = para.get('dmrg_sites', 1)
        self.stages[tag] = stage
        return s

    def finalize(self, solves):
        self.finalized += list(solves)
        return super().finalize(solves)


def test_ritz_energies_skip_the_measurement():
    """energy='ritz': one-site solves report their last Ritz value without a
    device measurement; two-site solves (truncated after the last solve) and
    energy='measure' measure."""
    for mode in ('ritz', 'measure'):
        clk = Clock()
        eng = PreparingEngine(clk)
        srv = SolveServer(eng, hbm_budget=1 << 40, max_active=4, clock=clk, energy=mode)
        srv.submit(0, _para(10, 8), seed=1, arrival=0.0)
        srv.submit(1, _para(12, 8, dmrg_sites=2), seed=2, arrival=0.0)
        srv.run()
        e = {r.rid: r.energy for r in srv.finished}
        if mode == 'ritz':
            assert e[0] == -5.0 and e[1] == -0.44 * 12
            assert [s.n for s in eng.finalized] == [12]
        else:
            assert e[0] == -0.44 * 10 and e[1] == -0.44 * 12
            assert sorted(s.n for s in eng.finalized) == [10, 12]


def test_prepared_stages_reach_admission():
    """Arrived requests are prepared ahead (engine.prepare on a worker) and
    admitted with their own stage; a request that arrives while the server
    idles is not prepared early, it is admitted at once and staged inline."""
    clk = Clock()
    eng = PreparingEngine(clk)
    srv = SolveServer(eng, hbm_budget=1 << 40, max_active=1, clock=clk)
    for k in range(3):
        srv.submit(k, _para(8 + k, 8), seed=10 + k, arrival=0.0)
    srv.submit(3, _para(9, 8), seed=13, arrival=5.0)
    srv.run()
    assert len(srv.finished) == 4 and all(r.error is None for r in srv.finished)
    assert eng.stages[1] == ('stage', 9, 11) and eng.stages[2] == ('stage', 10, 12)
    assert eng.stages[3] is None


def test_nonfatal_failures_reach_the_request():
    """A solve whose step hit an iteration limit (its result is still
    written) finishes with an energy and carries the warning; the others
    carry none."""
    clk = Clock()
    eng = FakeEngine(clk)
    srv = SolveServer(eng, hbm_budget=10 ** 9, max_active=4, clock=clk)
    srv.submit(0, _para(10, 16), arrival=0.0)
    srv.submit(1, _para(10, 16, warn_at=7), arrival=0.0)
    summ = srv.run()
    by = {r.rid: r for r in srv.finished}
    assert summ['completed'] == 2 and summ['failed'] == 0 and summ['with_warnings'] == 1
    assert by[0].warnings == [] and by[0].energy is not None
    assert by[1].warnings == ['block Jacobi SVD reached 60 sweeps']
    assert by[1].error is None and by[1].energy == -0.44 * 10
