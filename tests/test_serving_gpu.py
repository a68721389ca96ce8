"""Serving on the GPU: continuous batching of uneven requests gives every
request the same result as running it alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NCV = 8   # short Krylov cycles, so that local solves restart


def _requests():
    from tnalg_amd.serving import poisson_xxz_stream
    return poisson_xxz_stream(6, rate=1000.0, seed=3, n_range=(10, 22), chis=(8, 16),
                              sweeps=3, eigs_tol=1e-10)


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_alone_equals_in_mix(gpu, precision):
    """Each request alone (host work done in line, one request per step) and
    in a mixed stream (next-step host work planned ahead on a worker thread,
    one Lanczos call over every precision of the step) gives bit-identical
    energies, for both restart modes: implicit restarts inside the Lanczos
    call, and cycle steps (a call runs at most 1 + cycle_restarts Krylov
    cycles; an unconverged local solve resumes from its Ritz vector in the
    next step)."""
    import time
    from tnalg_amd.dmrg import _schedule_from
    from tnalg_amd.engine import Engine
    from tnalg_amd.serving import SolveServer
    reqs = _requests()
    sched = _schedule_from(precision)
    lowp = precision != 'fp64'
    for cycle in (False, True):
        alone = []
        for k, (t, para, seed) in enumerate(reqs):
            print('alone request', k, 'N', para['l'], 'chi', para['chi'], 'cycle_steps', cycle,
                  flush=True)
            eng = Engine(precision_schedule=sched, lowp=lowp, ncv=NCV, plan_ahead=False,
                         cycle_steps=cycle, cycle_restarts=0)
            srv = SolveServer(eng, hbm_budget=8 << 30, max_active=1, lowp_min_chi=0)
            srv.submit(k, para, seed=seed, arrival=0.0)
            srv.run()
            alone.append(srv.finished[0].energy)
        print('mixed stream of', len(reqs), 'cycle_steps', cycle, flush=True)
        eng = Engine(precision_schedule=sched, lowp=lowp, ncv=NCV, cycle_steps=cycle,
                     cycle_restarts=0)
        srv = SolveServer(eng, hbm_budget=8 << 30, max_active=4, lowp_min_chi=0)
        t0 = time.perf_counter()
        for k, (t, para, seed) in enumerate(reqs):
            srv.submit(k, para, seed=seed, arrival=t0 + 0.002 * k)
        summ = srv.run()
        assert summ['completed'] == len(reqs)
        assert summ['peak_active'] > 1
        assert (eng.counters['resumes'] > 0) == cycle
        mixed = {r.rid: r.energy for r in srv.finished}
        for k in range(len(reqs)):
            assert mixed[k] == alone[k], (cycle, k, mixed[k], alone[k])


def test_two_lanczos_streams_same_bits(gpu):
    """A batch large enough to be split (every precision group halved into
    two Lanczos calls on two streams, each driven by its own host thread)
    gives bit-identical energies to one call on one stream."""
    from tnalg_amd.dmrg import _schedule_from
    from tnalg_amd.engine import Engine
    from tnalg_amd.serving import poisson_xxz_stream
    reqs = poisson_xxz_stream(20, rate=1000.0, seed=7, n_range=(10, 18), chis=(8, 16),
                              sweeps=3, eigs_tol=1e-10)
    out = {}
    for streams in (1, 2):
        print('batch of', len(reqs), 'lanczos_streams', streams, flush=True)
        eng = Engine(precision_schedule=_schedule_from('mixed'), lowp=True, ncv=NCV,
                     lanczos_streams=streams)
        solves = eng.add_solves([(para, seed) for _, para, seed in reqs])
        for s, (_, para, _) in zip(solves, reqs):
            s.schedule_canonicalise()
            for t in range(int(para['sweep_time'])):
                s.schedule_sweep(t)
        eng.run_queues(solves)
        out[streams] = eng.finalize(solves)
    assert len(out[2]) == len(reqs)
    assert out[1] == out[2]


def test_ritz_energy_equals_measured(gpu):
    """The served energy of a one-site solve (the Ritz value of its last
    local eigensolve) equals the measured <H> of its final MPS; the same
    stream served with prepared stages and without gives the same bits."""
    from tnalg_amd.dmrg import _schedule_from
    from tnalg_amd.engine import Engine
    from tnalg_amd.serving import SolveServer
    reqs = _requests()
    got = {}
    for mode, ahead in (('ritz', 8), ('ritz', 0), ('measure', 8)):
        print('stream energy', mode, 'prepare_ahead', ahead, flush=True)
        eng = Engine(precision_schedule=_schedule_from('mixed'), lowp=True, ncv=NCV)
        srv = SolveServer(eng, hbm_budget=8 << 30, max_active=4, energy=mode,
                          prepare_ahead=ahead, lowp_min_chi=0)
        for k, (t, para, seed) in enumerate(reqs):
            srv.submit(k, para, seed=seed, arrival=0.0)
        summ = srv.run()
        assert summ['completed'] == len(reqs)
        got[(mode, ahead)] = {r.rid: r.energy for r in srv.finished}
    for k in range(len(reqs)):
        e_r, e_m = got[('ritz', 8)][k], got[('measure', 8)][k]
        assert got[('ritz', 0)][k] == e_r
        assert abs(e_r - e_m) < 1e-10 * abs(e_m), (k, e_r, e_m)


def test_two_site_requests_in_the_stream(gpu):
    """Two-site requests (prepared on a worker, measured at retirement) served
    together with one-site ones give the same bits as each served alone."""
    from tnalg_amd.dmrg import _schedule_from
    from tnalg_amd.engine import Engine
    from tnalg_amd.serving import SolveServer
    reqs = _requests()[:4]
    reqs = [(t, dict(para, dmrg_sites=2, chi_init=4) if k % 2 else para, seed)
            for k, (t, para, seed) in enumerate(reqs)]

    def serve(items, max_active):
        eng = Engine(precision_schedule=_schedule_from('mixed'), lowp=True, ncv=NCV)
        srv = SolveServer(eng, hbm_budget=8 << 30, max_active=max_active, lowp_min_chi=0)
        for k, (t, para, seed) in items:
            srv.submit(k, para, seed=seed, arrival=0.0)
        summ = srv.run()
        assert summ['failed'] == 0
        return {r.rid: r.energy for r in srv.finished}
    print('two-site requests: mixed stream', flush=True)
    mixed = serve(list(enumerate(reqs)), 4)
    for k, item in enumerate(reqs):
        print('two-site requests: alone', k, 'sites', item[1].get('dmrg_sites', 1), flush=True)
        assert serve([(k, item)], 1)[k] == mixed[k], k
