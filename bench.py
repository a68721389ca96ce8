"""Benchmark: DMRG solves/s + p50/p99 latency on one MI355X.

Workload (default): XXZ chain, N=64, open BC, Jxy=1, Jz=1, h=0, chi=128,
20 Krylov vectors, tol 1e-10 (tau=1e-4), 8 sweeps, seeded random initial
MPS; a *step* is one batched solve of ``--batch`` such chains (all 8 sweeps,
observables at the end), advanced as ``--pipelines`` (default 4) sub-batches,
each on its own engine, stream and host thread, so that one sub-batch's
latency-bound gauge moves overlap the others' Lanczos calls.  ``value`` = solves/s of the batched mixed-precision
engine; the plain-f64 path and the CPU baseline are reported beside it, with
the relative energy difference of every mixed solve against the f64 solve of
the same seed.

Extra JSON objects: ``roofline`` (the kernel class of the Lanczos loop with
the most estimated time in the timed batch, from the library's sampled timing
of that batch: every launch of the first Krylov step of every 4th Lanczos
call timed by its own blocks with the device wall clock; the other classes
under ``kernels``; ``traffic`` from counter profiles, MATVEC_PMC_TRAFFIC) and
``cpu_baseline`` (the oracle's restatement of the reference algorithm timed on
a bounded sample across the chain).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, 't-nalg_amd')):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

MFMA_PEAK = {'fp64': 78.6, 'fp32': 157.3, 'bf16': 2500.0}  # TFLOP/s, dense
HBM_PEAK = 8000.0  # GB/s


def c2_para(n=64, chi=128, sweeps=8, jz=1.0, hz=0.0, tol=1e-10, sites=1, two_site_sweeps=None,
            chi_init=None):
    from tnalg_amd import parameters as P
    para = P.parameter_dmrg_chain()
    para.update(P.common_parameters_dmrg())
    para.update(l=n, chi=chi, jxy=1.0, jz=jz, hz=hz, sweep_time=sweeps, eigs_tol=tol, tau=1e-4,
                dt_ob=sweeps, break_tol=0.0, dmrg_sites=sites)
    if two_site_sweeps is not None:
        para['two_site_sweeps'] = two_site_sweeps
    if chi_init is not None:
        para['chi_init'] = chi_init
    return P.make_consistent_parameter_dmrg(para)


_ENGINES = {}
_PLAN_AHEAD = [True]   # --no-plan-ahead: next-step host work in line (for comparison)
_CYCLE_STEPS = [True]  # serving: one Krylov cycle per Lanczos call (--no-cycle-steps: all)
_PIPELINES = [4]       # --pipelines k: the batch as k sub-batches on k engines, streams, threads
_LZ_STREAMS = [None]   # Lanczos calls of a step over this many streams (None: 2 for the
                       # fixed batch, 1 for serving, where one Krylov cycle per call and
                       # small mixed groups measured faster on one stream)


def engine_for(precision, ncv=20):
    """One long-lived engine per precision schedule, as a solver service keeps
    it: step plans, solve arenas and resident descriptor sets built for the
    first batch of a shape are reused by later batches of that shape."""
    from tnalg_amd.dmrg import _schedule_from
    from tnalg_amd.engine import Engine
    key = (precision, ncv)
    if key not in _ENGINES:
        _ENGINES[key] = Engine(precision_schedule=_schedule_from(precision), ncv=ncv,
                               lowp=precision not in (None, 'fp64'), plan_ahead=_PLAN_AHEAD[0],
                               lanczos_streams=_LZ_STREAMS[0] or 2, init_rng='device')
    return _ENGINES[key]


def _release_engine(precision, ncv=20):
    import torch
    _ENGINES.pop((precision, ncv), None)
    torch.cuda.empty_cache()


def _solve_batch(eng, para, seeds, keep_mps=False):
    """Every chain of ``seeds`` from its seeded random MPS through all sweeps
    in lockstep on ``eng`` (on the calling thread's current stream), then
    every output ``dmrg_finite_size`` gives (``DMRG_anyH.py:66-70,91-95``:
    eb_full, mx, mz, e_per_site, eb, the entanglement spectrum of every bond
    by SVD centre moves, the entropies, corr_x and corr_z); the solves are
    released.  Returns one dict per solve (the ob keys plus ``E``, ``lm``,
    ``ent``; with ``keep_mps`` also ``mps``, the final host tensors, whose
    copy time ``t_keep`` the caller leaves out of its timing)."""
    import torch
    from tnalg_amd.dmrg import finish_outputs
    solves = eng.add_solves([(para, s) for s in seeds])
    for s in solves:
        s.schedule_canonicalise()
    eng.run_queues(solves)
    for t in range(para['sweep_time']):
        for s in solves:
            s.schedule_sweep(t)
        eng.run_queues(solves)
    obs, spectra = finish_outputs(eng, solves, [para] * len(solves))
    torch.cuda.current_stream().synchronize()
    for s, ob, (lm, ent) in zip(solves, obs, spectra):
        ob['E'] = float(np.ravel(ob['e_per_site'])[0]) * s.n
        ob['lm'], ob['ent'] = lm, ent
        ob['warnings'] = list(s.warnings)
        if keep_mps:
            t0 = time.perf_counter()
            ob['mps'] = s.host_mps()
            ob['t_keep'] = time.perf_counter() - t0
    for s in list(solves):
        eng.remove_solve(s)
    return obs


def _freeze_caches():
    """A solver service keeps its step plans, descriptor sets and term
    tables for its lifetime: after a warm-up, move every object alive into
    the interpreter's permanent generation, so that a full cyclic
    collection (which would scan those caches and stall the host for about
    0.1 s at a random step) does not run over them."""
    import gc
    gc.collect()
    gc.freeze()


def run_batch(para, seeds, precision, ncv=20, keep_mps=False):
    """One batched solve: every chain of ``seeds`` from its seeded random MPS
    through all sweeps in lockstep, then all its outputs (``_solve_batch``);
    returns ``(seconds, outputs, engine counters, engine)``.  With ``--pipelines
    k`` the batch is cut into k sub-batches, each advanced by its own engine on
    its own stream and host thread, so one sub-batch's latency-bound gauge
    moves overlap another's Lanczos calls."""
    import torch
    if _PIPELINES[0] > 1 and len(seeds) >= 2 * _PIPELINES[0]:
        return _run_pipelined(para, seeds, precision, ncv)
    eng = engine_for(precision, ncv)
    for k in eng.counters:
        eng.counters[k] = 0.0 if isinstance(eng.counters[k], float) else 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    obs = _solve_batch(eng, para, seeds, keep_mps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0 - sum(o.get('t_keep', 0.0) for o in obs)
    return dt, obs, dict(eng.counters), eng


def _run_pipelined(para, seeds, precision, ncv):
    import threading
    import torch
    from tnalg_amd.dmrg import _schedule_from
    from tnalg_amd.engine import Engine
    k = _PIPELINES[0]
    key = ('pipelines', k, precision, ncv)
    if key not in _ENGINES:
        streams = [torch.cuda.Stream() for _ in range(k)]
        engs = []
        for st in streams:
            with torch.cuda.stream(st):
                engs.append(Engine(precision_schedule=_schedule_from(precision), ncv=ncv,
                                   init_rng='device', lowp=precision not in (None, 'fp64'),
                                   plan_ahead=_PLAN_AHEAD[0], lanczos_streams=1))
        _ENGINES[key] = (engs, streams)
    engs, streams = _ENGINES[key]
    for eng in engs:
        for c in eng.counters:
            eng.counters[c] = 0.0 if isinstance(eng.counters[c], float) else 0
    parts = [list(p) for p in np.array_split(np.asarray(seeds), k)]
    out, err = [None] * k, [None] * k

    def work(i):
        try:
            with torch.cuda.stream(streams[i]):
                out[i] = _solve_batch(engs[i], para, parts[i])
        except BaseException as e:   # re-raised on the calling thread
            err[i] = e
    threads = [threading.Thread(target=work, args=(i,)) for i in range(k)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    for e in err:
        if e is not None:
            raise e
    counters = {}
    for eng in engs:
        for c, v in eng.counters.items():
            counters[c] = counters.get(c, 0) + v
    return dt, sum(out, []), counters, engs[0]


def matvec_roofline(para, batch, precs=('fp64', 'bf16'), reps=20):
    """Average launch durations of the matvec kernels (the grouped GEMMs and
    the mixer of a mid-chain local step, all solves of the batch) at each
    precision in ``precs``, with HIP events on the engine's stream;
    algorithmic flops from the launches' own records."""
    import ctypes
    import torch
    from tnalg_amd import _lib as L
    from tnalg_amd import records as R
    from tnalg_amd.engine import PREC, Engine, Step
    # the walk to the measured site runs in f64, so the low-precision GEMM
    # kernels run only in the timed loops below
    eng = Engine(precision_schedule=lambda sweep, nsweeps: 'fp64', ncv=20,
                 lowp=any(p != 'fp64' for p in precs))
    solves = [eng.add_solve(para, seed=100 + k) for k in range(batch)]
    for s in solves:
        s.schedule_canonicalise()
    eng.run_queues(solves)
    site = para['l'] // 2
    # walk every centre to the middle with f64 steps (one batch), then plan
    # one local step there per precision
    for s in solves:
        for p in range(1, site):
            s.queue.append(Step('solve', p, +1, 0))
    eng.run_queues(solves)
    st_ = eng.stream()
    lib = eng.lib
    vp = ctypes.c_void_p

    def timed(fns):
        """Average seconds per repetition of ``fns`` (HIP events recorded on
        the stream the kernels run on, torch's current stream)."""
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            for f in fns:
                f()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps

    out = {}
    for prec in precs:
        plans = []
        for s in solves:
            dims0 = list(s.dims)
            if eng.uses_lowp and prec != 'fp64':
                eng._mix([s.cast_all_records(PREC[prec][0])])
            plans.append(s.plan(Step('solve', site, +1, para['sweep_time'] - 1), prec))
            s.dims = dims0
        ti, tc, mode1, mode2, _ = PREC[prec]
        parts = [(s, ph) for s, ph in zip(solves, plans)]
        g1p, g1s, g1t = R.concat_gemm([s.part(ph, 'mv_g1') for s, ph in parts])
        g2p, g2s, g2t = R.concat_gemm([s.part(ph, 'mv_g2') for s, ph in parts])
        mxj, mxt, mxb = R.concat_mix([s.part(ph, 'mv_mix') for s, ph in parts])
        ptr = eng.up.upload([g1p, g1s, g2p, g2s, mxj, mxt])
        flops = float(sum(ph['mv_flops'] for ph in plans))

        def gemm(i, n, t, m):
            L.check(lib.tnx_gemm_grouped(m, vp(ptr[i]), n, vp(ptr[i + 1]), t, None, st_), 'gemm')

        def mix(ptr=ptr, n=len(mxj), b=mxb):
            L.check(lib.tnx_mix(vp(ptr[4]), n, vp(ptr[5]), b, vp(eng.ops.data_ptr()), None,
                                st_), 'mix')

        res = {'flops': flops}
        best = {}
        for key, i, n, t, mode, fl in (('g1', 0, len(g1p), g1t, mode1, gemm_flops(g1p, g1s)),
                                       ('g2', 2, len(g2p), g2t, mode2, gemm_flops(g2p, g2s))):
            # every kernel variant of the mode (the engine's tuner keeps the
            # fastest); the same launches, the same bits
            tv = []
            for v in range(int(lib.tnx_gemm_variants(mode))):
                f = (lambda i=i, n=n, t=t, m=mode | (v << 8): gemm(i, n, t, m))
                for _ in range(3):
                    f()
                torch.cuda.synchronize()
                tv.append(timed([f]))
            v = int(np.argmin(tv))
            best[key] = (lambda i=i, n=n, t=t, m=mode | (v << 8): gemm(i, n, t, m))
            res[key] = {'t': tv[v], 'flops': fl, 'tiles': int(t), 'mode': mode, 'variant': v,
                        't_variants': tv}
        for _ in range(3):
            best['g1'](), mix(), best['g2']()
        torch.cuda.synchronize()
        # a spin kernel marks the start of each timed loop in a counter
        # profile (--pmc-summary): 3 warm-up launches then `reps` timed ones
        torch.cuda._sleep(1000)
        for _ in range(3):
            mix()
        res['mx'] = {'t': timed([mix]), 'bytes': R.mix_bytes(mxj, mxt), 'blocks': int(mxb)}
        torch.cuda.synchronize()
        torch.cuda._sleep(1000)
        for _ in range(3):
            best['g1'](), mix(), best['g2']()
        res['t_matvec'] = timed([best['g1'], mix, best['g2']])
        out[prec] = res
    for s in list(solves):
        eng.remove_solve(s)
    return out


GEMM_MODE_NAME = {0: 'f64', 1: 'f32', 2: 'bf16 -> f32', 3: 'bf16 -> bf16'}


def _class_peak(name):
    """(bound, peak, unit) of a sampled kernel class."""
    if name.startswith('gemm'):
        prec = 'fp64' if 'f64' in name else 'fp32' if 'f32' in name else 'bf16'
        return 'mfma', MFMA_PEAK[prec], 'TFLOP/s'
    return 'hbm', HBM_PEAK, 'GB/s'


def live_kernels(stats):
    """Per kernel class of the Lanczos loop, from the library's sampled
    timing of a timed batch (tnx_profile_begin / tnx_profile_end): average
    duration of a timed launch (first Krylov step of a sampled call, every
    solve active), algorithmic work per launch, achieved rate and fraction
    of the roofline, and the launches of the class in the batch.  ``est_s``
    (launches x average duration) ranks the classes; it overstates the
    later Krylov steps, in which converged solves are masked."""
    out = []
    for st in stats:
        if st.samples <= 0 or st.seconds <= 0:
            continue
        name = st.name.decode()
        bound, peak, unit = _class_peak(name)
        avg = st.seconds / st.samples
        work = st.work / st.samples
        ach = work / avg / (1e12 if unit == 'TFLOP/s' else 1e9)
        out.append({'kernel': name, 'bound': bound, 'achieved': ach, 'peak': peak, 'unit': unit,
                    'frac': ach / peak, 'avg_launch_s': avg,
                    ('flops_per_launch' if st.unit == 0 else 'bytes_per_launch'): work,
                    'timed_launches': int(st.samples), 'launches': int(st.launches),
                    'work_all_launches': st.work_all,
                    'est_s': avg * st.launches})
    return sorted(out, key=lambda k: -k['est_s'])


# HBM traffic per launch (bytes) of the matvec kernels at the mid-chain step
# of the 64-solve N=64, chi=128 batch, with the algorithmic bytes of the same
# launch: two rocprofv3 counter runs of `bench.py --roofline-only`
# (--pmc FETCH_SIZE, then --pmc WRITE_SIZE; KiB per dispatch averaged over
# the 23 launches after each loop marker, --pmc-summary), scaled by what the
# counters read for the calibration copy of 256 MiB (FETCH_SIZE x 2.00,
# WRITE_SIZE x 1.00).  The GEMMs are priced in flops (2.15e9 per launch at
# that step), so only the mixer has an algorithmic byte count; a GEMM
# class's traffic in the bench line is its live average launch's flops
# times the measured bytes per flop.
MATVEC_PMC_TRAFFIC = {
    'mixer matvec': {'bf16': (42.5e6, 50.3e6), 'fp32': (76.1e6, 92.3e6),
                     'fp64': (152.1e6, 184.5e6)},
    'gemm f64 matvec T = F x': {'fp64': (117.7e6, None)},
    'gemm f64 matvec w = U R^T': {'fp64': (134.4e6, None)},
    'gemm f32 matvec T = F x': {'fp32': (63.5e6, None)},
    'gemm f32 matvec w = U R^T': {'fp32': (63.5e6, None)},
    'gemm bf16 matvec T = F x': {'bf16': (29.7e6, None)},
    'gemm bf16 matvec w = U R^T': {'bf16': (42.7e6, None)},
}
MATVEC_PMC_FLOPS = 2.15e9   # flops of each GEMM launch of those counter runs


KERNEL_OF_CLASS = {'gemm f64': 'gemm_grouped_kernel<0>/<4>', 'gemm f32': 'gemm_grouped_kernel<1>',
                   'gemm bf16 matvec T': 'gemm_grouped_kernel<3>',
                   'gemm bf16 matvec w': 'gemm_grouped_kernel<2>', 'mixer': 'mix_kernel<1>',
                   'k_dots': 'k_dots', 'k_update_dots': 'k_update_dots',
                   'k_update_norm': 'k_update_norm'}


def roofline_object(kern, note):
    """``roofline`` of the bench line: the class that ranks first in
    :func:`live_kernels` (measured live on the timed batch), with the
    others under ``kernels``."""
    if not kern:
        return None
    dom = kern[0]
    hip = next((v for k, v in KERNEL_OF_CLASS.items() if dom['kernel'].startswith(k)), '?')
    pmc = MATVEC_PMC_TRAFFIC.get(dom['kernel'])
    traffic, detail = None, None
    if pmc and dom['unit'] == 'GB/s' and all(v[1] for v in pmc.values()):
        # the mixer serves every precision: weight the measured traffic per
        # algorithmic byte of each precision by that precision's share of
        # the timed matvec launches (its GEMM class's timed launches), and
        # apply it to the live average launch
        share = {}
        for k in kern:
            if k['kernel'].startswith('gemm') and 'T = F x' in k['kernel']:
                share['fp64' if 'f64' in k['kernel'] else 'fp32' if 'f32' in k['kernel']
                      else 'bf16'] = k['timed_launches']
        tot = sum(share.get(p, 0) for p in pmc) or 1
        hbm = sum(share.get(p, 0) * v[0] for p, v in pmc.items()) / tot
        alg = sum(share.get(p, 0) * v[1] for p, v in pmc.items()) / tot
        ratio = hbm / alg if alg else None
        traffic = dom['bytes_per_launch'] * ratio if ratio else None
        detail = {'hbm_per_algorithmic_byte': ratio,
                  'launch_shares': {p: share.get(p, 0) / tot for p in pmc},
                  'mid_chain_by_precision': {k: {'hbm_bytes': v[0], 'algorithmic_bytes': v[1]}
                                             for k, v in pmc.items()},
                  'source': 'rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of bench.py --roofline-only '
                            '(mid-chain step, 64 solves), scaled by the 256 MiB calibration copy; '
                            'traffic = the live average launch x the share-weighted ratio'}
    elif pmc:
        prec = next(iter(pmc))
        per_flop = pmc[prec][0] / MATVEC_PMC_FLOPS
        traffic = dom.get('flops_per_launch', 0.0) * per_flop
        detail = {'precision': prec, 'hbm_bytes_per_flop': per_flop,
                  'mid_chain_hbm_bytes': pmc[prec][0], 'mid_chain_flops': MATVEC_PMC_FLOPS,
                  'source': 'rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of bench.py --roofline-only '
                            '(mid-chain step, 64 solves), scaled by the 256 MiB calibration copy; '
                            'traffic = the live average launch\'s flops x the measured bytes per '
                            'flop'}
    return {'bound': dom['bound'], 'achieved': dom['achieved'], 'peak': dom['peak'],
            'unit': dom['unit'], 'frac': dom['frac'], 'traffic': traffic,
            'traffic_detail': detail,
            'kernel': '%s (%s)' % (hip, dom['kernel']), 'avg_launch_s': dom['avg_launch_s'],
            'measured': note, 'kernels': kern}


def output_differences(obs, ref):
    """Largest absolute difference, over the solves, of every output against
    the f64 solve of the same seed (the precision statement of every
    output): energy (relative), mx, mz, eb_full, eb, corr_x, corr_z, the
    spectra and entropies."""
    out = {'E_rel': max(abs(a['E'] - b['E']) / abs(b['E']) for a, b in zip(obs, ref))}
    for key in ('mx', 'mz', 'eb_full', 'eb', 'corr_x', 'corr_z', 'ent'):
        out[key] = max(float(np.abs(np.ravel(a[key]) - np.ravel(b[key])).max())
                       for a, b in zip(obs, ref))
    out['lm'] = max(float(np.abs(x - y[:len(x)]).max()) if len(x) == len(y) else float('inf')
                    for a, b in zip(obs, ref) for x, y in zip(a['lm'], b['lm']))
    return out


def gemm_flops(p, s):
    """Algorithmic flops of one grouped-GEMM launch: 2 M N K summed over the
    problems and their K segments."""
    fl = 0.0
    for q in p:
        k = int(s['K'][q['seg_begin']:q['seg_begin'] + q['seg_count']].sum())
        fl += 2.0 * float(q['M']) * float(q['N']) * k
    return fl


def copy_calibration(mib=256):
    """One elementwise kernel of known HBM traffic (reads and writes ``mib``
    MiB once), the scale for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    x = torch.ones(mib << 17, dtype=torch.float64, device='cuda')
    y = torch.empty_like(x)
    torch.mul(x, 2.0, out=y)
    torch.cuda.synchronize()
    return mib << 20


def pmc_summary(paths):
    """Per counter file (rocprofv3 --pmc ... --output-format csv of a
    --roofline-only run): the dispatches after each spin-kernel marker form
    a segment (per precision: the mixer alone, then the matvec GEMMs and
    mixer; last: the calibration copy); per segment and kernel, the number
    of dispatches and the mean value of each counter (KiB per dispatch for
    FETCH_SIZE / WRITE_SIZE)."""
    import csv
    out = {}
    for path in paths:
        rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r['Dispatch_Id']))
        segs, cur = [], None
        seen = set()
        for r in rows:
            name = r['Kernel_Name']
            if 'spin_kernel' in name:
                if r['Dispatch_Id'] not in seen:
                    seen.add(r['Dispatch_Id'])
                    cur = {}
                    segs.append(cur)
                continue
            if cur is None:
                continue
            short = name.replace('void ', '').replace('(anonymous namespace)::', '').split('(')[0]
            ent = cur.setdefault(short, {})
            ent.setdefault(r['Counter_Name'], []).append(float(r['Counter_Value']))
        out[path] = [{k: {c: {'n': len(v), 'mean': float(np.mean(v))} for c, v in d.items()}
                      for k, d in seg.items()} for seg in segs]
    return out


def trace_summary(path, t0=None, t1=None):
    """Busy and idle time of a rocprofv3 --kernel-trace CSV between host
    monotonic clocks t0, t1 (ns, as the timed steps print them): span,
    busy fraction, idle by gap size and by the kernel pair around each gap,
    kernel time per kernel, and the average duration of the launches of
    the first Krylov step of every Lanczos call (the launches that follow a
    call's k_init_scale up to its first k_update_norm), which is what the
    live sampled timing measures."""
    import collections
    import csv
    import re

    def short(n):
        n = n.replace('(anonymous namespace)::', '').replace('void ', '')
        return re.sub(r'\(.*', '', n)

    rows = sorted((int(r['Start_Timestamp']), int(r['End_Timestamp']), short(r['Kernel_Name']),
                   r.get('Stream_Id') or r.get('Queue_Id') or '0')
                  for r in csv.DictReader(open(path)))
    if t0 is not None:
        rows = [r for r in rows if r[0] >= t0 and r[1] <= t1]
    if not rows:
        return {'kernels': 0}
    a, b = rows[0][0], max(r[1] for r in rows)
    busy, cs, ce = 0, rows[0][0], rows[0][1]
    gaps = collections.defaultdict(lambda: [0, 0])
    hist = collections.Counter()
    prev = rows[0]
    for r in rows[1:]:
        if r[0] > ce:
            busy += ce - cs
            g = r[0] - ce
            gaps[(prev[2], r[2])][0] += g
            gaps[(prev[2], r[2])][1] += 1
            k = 0
            while (1 << k) * 1000 < g:
                k += 1
            hist[1 << k] += g
            cs, ce, prev = r[0], r[1], r
        else:
            if r[1] >= ce:
                prev = r
            ce = max(ce, r[1])
    busy += ce - cs
    tot = collections.Counter()
    for r in rows:
        tot[r[2]] += r[1] - r[0]
    first = collections.defaultdict(list)
    inside = collections.defaultdict(bool)   # per stream: inside a first Krylov step
    for r in rows:
        q = r[3]
        if r[2] == 'k_init_scale':
            inside[q] = True
        elif inside[q]:
            first[r[2]].append(r[1] - r[0])
            if r[2] == 'k_update_norm':
                inside[q] = False
    # per stream (one per sub-batch): its kernel time, the time between its
    # consecutive kernels, and its kernel time inside and outside Lanczos
    # calls (from a call's k_zero_words to its k_final)
    per_stream = {}
    for q in sorted({r[3] for r in rows}):
        rs = [r for r in rows if r[3] == q]
        kt = collections.Counter()
        lz = [0, 0]
        in_lz = False
        gap = 0
        for i, r in enumerate(rs):
            kt[r[2]] += r[1] - r[0]
            if r[2] == 'k_zero_words':
                in_lz = True
            lz[in_lz] += r[1] - r[0]
            if r[2] == 'k_final':
                in_lz = False
            if i:
                gap += max(0, r[0] - rs[i - 1][1])
        per_stream[str(q)] = {
            'kernels': len(rs), 'span_s': (rs[-1][1] - rs[0][0]) / 1e9,
            'kernel_s': sum(r[1] - r[0] for r in rs) / 1e9, 'gap_s': gap / 1e9,
            'kernel_s_in_lanczos': lz[1] / 1e9, 'kernel_s_outside_lanczos': lz[0] / 1e9,
            'top': {k: v / 1e9 for k, v in kt.most_common(10)}}
    return {'span_s': (b - a) / 1e9, 'busy_s': busy / 1e9, 'busy_frac': busy / (b - a),
            'per_stream': per_stream,
            'kernels': len(rows), 'summed_kernel_s': sum(r[1] - r[0] for r in rows) / 1e9,
            'idle_by_gap_us_le': {str(k): v / 1e9 for k, v in sorted(hist.items())},
            'idle_by_transition': [{'prev': k[0], 'next': k[1], 'idle_s': g / 1e9, 'gaps': n}
                                   for k, (g, n) in sorted(gaps.items(), key=lambda x: -x[1][0])[:20]],
            'kernel_s': {k: v / 1e9 for k, v in tot.most_common(24)},
            'first_krylov_step_avg_us': {k: {'launches': len(v), 'avg_us': float(np.mean(v)) / 1e3}
                                         for k, v in first.items()}}


def serving_run(n_req, rate, max_active=64, seed=0):
    """The serving stream: XXZ chains, N ~ U{16..96}, chi ~ U{32,64,128},
    Jz ~ U[0.5,1.5], hz ~ U[0,0.5], open BC, 8 sweeps each (fixed), Poisson
    arrivals at ``rate``/s, mixed precision, continuous batching with HBM
    admission.  Returns the server summary (solves/s, p50/p99 latency)."""
    import torch
    from tnalg_amd.dmrg import _schedule_from
    from tnalg_amd.engine import Engine
    from tnalg_amd.serving import SolveServer, poisson_xxz_stream
    reqs = poisson_xxz_stream(n_req, rate=rate, seed=seed)
    eng = Engine(precision_schedule=_schedule_from('mixed'), lowp=True, plan_ahead=_PLAN_AHEAD[0],
                 cycle_steps=_CYCLE_STEPS[0], lanczos_streams=_LZ_STREAMS[0] or 1)
    free = torch.cuda.mem_get_info()[0]
    srv = SolveServer(eng, hbm_budget=int(0.6 * free), max_active=max_active)
    t0 = time.perf_counter()
    for k, (t, para, s) in enumerate(reqs):
        srv.submit(k, para, seed=s, arrival=t0 + t)

    def progress(st):
        print('serving: %.1f s, %d done, %d active, %d waiting' % (
            time.perf_counter() - t0, st['completed'], st['active'], st['waiting']),
            file=sys.stderr, flush=True)
    summ = srv.run(progress=progress)
    c = eng.counters
    # where the wall time of the engine steps went (host planning and
    # descriptor assembly vs the Lanczos calls, which wait on the GPU)
    summ['step_time_s'] = {'plan': c['t_plan'], 'pre_phases': c['t_pre'],
                           'lanczos_total': c['t_lz_total'], 'lanczos_calls': c['lz_time'],
                           'wait_for_plan_ahead': c['t_settle'], 'engine_steps': c['steps']}
    summ['lanczos'] = {'calls': c['lanczos_calls'], 'matvecs': c['matvecs'],
                       'resumed_cycles': c['resumes'], 'cycle_steps': eng.cycle_steps,
                       'plan_ahead': eng.plan_ahead, 'streams': eng.lanczos_streams}
    summ['requests'] = n_req
    summ['arrival_rate_per_s'] = rate
    summ['max_active'] = max_active
    return summ


LARGE = {
    # dense-coupling stress: J1-J2 (J2 = 0.5) 6x6 open square lattice along the snake
    'c3': dict(width=6, height=6, j2=0.5, chi=512, sweeps=10),
    # large single solve: Heisenberg 8x8 open square lattice along the snake
    'c5': dict(width=8, height=8, j2=0.0, chi=1024, sweeps=12),
}


def large_para(name):
    from tnalg_amd import parameters as P
    w = LARGE[name]
    para = P.generate_parameters_dmrg('snakeJ1J2')
    para.update(square_width=w['width'], square_height=w['height'], j1=1.0, j2=w['j2'],
                chi=w['chi'], sweep_time=w['sweeps'], dt_ob=w['sweeps'], eigs_tol=1e-10,
                tau=1e-4, break_tol=0.0)
    return P.make_consistent_parameter_dmrg(para)


def large_solve(name, fp64_check=True):
    """One large single solve (f32 MFMA matvec sweeps, the last sweep f64;
    f64 Lanczos, gauge moves and SVD), seeded random MPS, timed end to end
    (first solve of its engine), with the f64 solve of the same seed for
    the energy difference."""
    import torch
    from tnalg_amd.dmrg import _schedule_from, dmrg_finite_size
    from tnalg_amd.engine import Engine
    para = large_para(name)
    out = {}
    for prec in (('fp32', 'fp64') if fp64_check else ('fp32',)):
        eng = Engine(precision_schedule=_schedule_from(prec), lowp=prec != 'fp64')
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ob, A, info, _ = dmrg_finite_size(para, seed=0, precision=prec, engine=eng)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        c = eng.counters
        out[prec] = {'seconds': dt, 'energy': float(np.ravel(ob['e_per_site'])[0]) * para['l'],
                     'matvecs': int(c['matvecs']),
                     'matvec_tflops': c['matvec_flops'] / max(c['lz_time'], 1e-9) / 1e12,
                     'hbm_gb': eng.footprint(para) / 1e9}
        del eng, ob, A
        torch.cuda.empty_cache()
        print('%s %s: %.1f s, E = %.12f' % (name, prec, dt, out[prec]['energy']), file=sys.stderr,
              flush=True)
    w = LARGE[name]
    res = {'metric': 'DMRG solves/sec (one large solve), 1 GPU', 'value': 1.0 / out['fp32']['seconds'],
           'unit': 'solves/s', 'n_gpus': 1, 'steps': 1, 'warmup': 0,
           'ms_per_step': 1e3 * out['fp32']['seconds'], 'higher_is_better': True,
           'scaling': 'strong', 'vs_baseline': None, 'dtype': 'f32 matvec, f64 Lanczos+QR+SVD',
           'data': 'synthetic (seeded random initial MPS)',
           'config': {'workload': '%s: %s %dx%d open square lattice along the snake, J1=1, J2=%g, '
                                  'chi=%d, %d sweeps, ncv=20, tol=1e-10' % (
                                      name, 'J1-J2' if w['j2'] else 'Heisenberg', w['width'],
                                      w['height'], w['j2'], w['chi'], w['sweeps']),
                      'batch': 1, 'precision': 'fp32 schedule (last sweep f64)'},
           'solves': out}
    if fp64_check:
        e, e64 = out['fp32']['energy'], out['fp64']['energy']
        res['dE_over_E_vs_fp64'] = {'value': abs(e - e64) / abs(e64), 'target': 1e-9,
                                    'met': abs(e - e64) <= 1e-9 * abs(e64)}
    return res


def host_description():
    """CPU model, logical cores, and the numpy / scipy / BLAS builds the CPU
    baseline runs on."""
    import platform
    import scipy
    from threadpoolctl import threadpool_info
    model = platform.processor() or 'unknown'
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    blas = ['%s %s (%s, %d threads)' % (i.get('internal_api'), i.get('version'),
                                        i.get('threading_layer', '?'), i.get('num_threads', 0))
            for i in threadpool_info() if i.get('user_api') == 'blas']
    return {'cpu_model': model, 'logical_cpus': os.cpu_count(), 'numpy': np.__version__,
            'scipy': scipy.__version__, 'blas': blas}


# sites whose local step is timed (chain of 64; the mirror image of a right
# move at site p is a left move at site 63 - p, so a sweep's cost is twice the
# sum over the sites of a right move)
_CPU_SITES = (3, 17, 31, 45, 59)


def _window_times(make, para, sites, warm=1, timed=1):
    """Seconds per local step at each of ``sites``: a fresh solver from
    ``make()`` (with its operator caches empty) is moved to ``p - warm``, runs
    ``warm`` untimed right-moving steps that fill its caches as a sweep
    would, then ``timed`` timed steps from site ``p``.  Returns
    ``(per-site seconds, ARPACK matvecs of the timed steps, seconds spent)``."""
    out, mv, t_all = [], [], time.perf_counter()
    for p in sites:
        step, counter = make(p - warm)
        for q in range(p - warm, p):
            step(q)
        m0 = counter()
        t0 = time.perf_counter()
        for q in range(p, p + timed):
            step(q)
        out.append((time.perf_counter() - t0) / timed)
        mv.append((counter() - m0) / timed)
    return out, mv, time.perf_counter() - t_all


def _sweep_seconds(sites, per_step, n):
    """One sweep's time from per-site samples (linear interpolation over the
    sites of a right move, then twice the sum: the left moves mirror it)."""
    xs = np.arange(1, n)
    return 2.0 * float(np.interp(xs, sites, per_step).sum())


def cpu_baseline(seconds=15.0, converged=None, n=64, chi=128, sweeps=8):
    """The oracle's restatement of the reference algorithm (per-term effective
    operators re-propagated as ``MPSClass.py:585-611`` does, ``opt_env``
    matvec, ARPACK ``eigsh`` ncv=20, QR gauge moves) on the N=64, chi=128 XXZ chain, timed
    on local steps sampled across the chain and extrapolated to one full
    solve; the block-environment restatement (``oracle/block_dmrg``, O(1)
    transfers per step) is timed the same way as a second number.

    Two states are sampled: the seeded random MPS (the first sweep's local
    problems) and, when given, ``converged`` (an MPS of the same chain after
    all sweeps; the later sweeps' problems, which ARPACK finishes in fewer
    matvecs).  The sweep-1 samples price sweep 1, the converged samples
    sweeps 2..8.  ``seconds`` only scales the number of sampled sites."""
    from oracle import dmrg_ref, model
    from oracle.block_dmrg import BlockDmrg
    from oracle.mps_ref import RefMps
    c = model.chain_couplings(n, 1.0, 1.0)
    p = dmrg_ref.make_para(c, chi, sweeps, tau=1e-4, eigs_tol=1e-10, break_tol=0.0, dt_ob=sweeps)
    nsites = max(3, min(len(_CPU_SITES), int(seconds / 5.0)))
    sites = [int(round(x)) for x in np.linspace(_CPU_SITES[0], _CPU_SITES[-1], nsites)]
    args = (p['index1'], p['index2'], p['coeff1'], p['coeff2'], p['tau'], True, p['eigs_tol'])

    def make_ref(state):
        def make(c0):
            A = RefMps(n, 2, chi, p['op'], np.random.RandomState(0), way='qr', is_save_op=True,
                       mps=state)
            A.correct_orthogonal_center(c0)
            return (lambda q: A.update_tensor_eigs(q, *args)), (lambda: A.n_matvec)
        return make

    def make_block(state):
        def make(c0):
            eng = BlockDmrg(p, np.random.RandomState(0))
            if state is not None:
                eng.A = RefMps(n, 2, chi, p['op'], None, way='qr', is_save_op=False, mps=state)
            eng.A.correct_orthogonal_center(c0)
            eng.initial_environments()
            return eng.solve_site, (lambda: eng.n_matvec)
        return make

    from oracle import tensor_ops as tb
    random_state, _ = tb.random_open_mps(n, 2, chi, np.random.RandomState(0))
    res = {}
    spent = 0.0
    for kind, make in (('reference', make_ref), ('block', make_block)):
        rows = {}
        for label, state in (('sweep1', random_state), ('converged', converged)):
            if state is None:
                continue
            use = sites if kind == 'reference' else [sites[len(sites) // 2]]
            t, mv, dt = _window_times(make(state), p, use)
            spent += dt
            if kind == 'block':   # O(1) transfers per step: one sampled site
                use, t, mv = [1, n - 1], [t[0], t[0]], [mv[0], mv[0]]
            rows[label] = {'sites': use, 's_per_step': t, 'matvecs_per_step': mv,
                           'sweep_s': _sweep_seconds(use, t, n)}
        later = rows.get('converged', rows['sweep1'])['sweep_s']
        solve_s = rows['sweep1']['sweep_s'] + (sweeps - 1) * later
        res[kind] = {'solve_s_extrapolated': solve_s, 'solves_per_s': 1.0 / solve_s,
                     'samples': rows}
    host = host_description()
    from threadpoolctl import threadpool_info
    threads = max([int(i.get('num_threads', 1)) for i in threadpool_info()] or [1])
    ref = res['reference']
    return {'value': ref['solves_per_s'], 'unit': 'solves/s', 'cores': threads, 'kind': 'port',
            'sample': 'oracle/mps_ref + dmrg_ref (the reference restatement: per-term effective '
                      'operators with its re-propagation, opt_env matvec, ARPACK eigsh ncv=20, '
                      'QR moves; numpy/scipy f64) on the N=64 chi=128 XXZ chain: one timed '
                      'local step (after one untimed step that fills the operator caches) at '
                      'sites %s, from the seeded random MPS (prices sweep 1) and from %s '
                      '(prices sweeps 2-%d); per-step times interpolated over the chain, '
                      'doubled for the left moves: one solve = %.0f s extrapolated, '
                      '%.1f s of sampling'
                      % (sites, 'a converged MPS of the same chain (the f64 GPU solve of seed 0)'
                         if converged is not None else 'the same random MPS', sweeps,
                         ref['solve_s_extrapolated'], spent),
            'block_env_variant': {'solves_per_s': res['block']['solves_per_s'],
                                  'solve_s_extrapolated': res['block']['solve_s_extrapolated'],
                                  'note': 'oracle/block_dmrg: the same local problems on block '
                                          'environments (O(1) transfers per step)',
                                  'samples': res['block']['samples']},
            'reference_samples': ref['samples'], 'host': host}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=1)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--batch', type=int, default=64)
    ap.add_argument('--precision', default='mixed')
    ap.add_argument('--n', type=int, default=64)
    ap.add_argument('--chi', type=int, default=128)
    ap.add_argument('--sweeps', type=int, default=8)
    ap.add_argument('--no-plain', action='store_true')
    ap.add_argument('--pipelines', type=int, default=4,
                    help='advance the batch as this many sub-batches, each on its own engine, '
                         'stream and host thread')
    ap.add_argument('--dmrg-sites', type=int, default=1, choices=[1, 2],
                    help='2: two-site sweeps with SVD truncation to chi (the roofline stays '
                         'on the one-site matvec)')
    ap.add_argument('--chi-init', type=int, default=None,
                    help='with --dmrg-sites 2: bond dimension of the random initial MPS')
    ap.add_argument('--two-site-sweeps', type=int, default=None,
                    help='with --dmrg-sites 2: only the first K sweeps are two-site')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-plan-ahead', action='store_true',
                    help='do the next step\'s host work in line instead of on a worker thread')
    ap.add_argument('--lanczos-streams', type=int, default=None, choices=[1, 2],
                    help='Lanczos calls of a step split over this many streams '
                         '(default: 2 for the batch, 1 for serving)')
    ap.add_argument('--no-cycle-steps', action='store_true',
                    help='serving: restart unconverged local solves inside the Lanczos call')
    ap.add_argument('--cpu-seconds', type=float, default=15.0)
    ap.add_argument('--serving-requests', type=int, default=0)
    ap.add_argument('--serving-rate', type=float, default=4.0)
    ap.add_argument('--serving-max-active', type=int, default=64,
                    help='most requests the server advances together')
    ap.add_argument('--workload', default='c2', choices=['c2', 'c3', 'c4', 'c5'],
                    help='c2: the batched N=64 chi=128 XXZ chain (default); c3 / c5: one large '
                         'snake-mapped square-lattice solve (chi 512 / 1024); c4: the serving '
                         'stream (--serving-requests requests at --serving-rate per second)')
    ap.add_argument('--pmc-summary', nargs='+', metavar='CSV',
                    help='summarise rocprofv3 --pmc counter_collection.csv files of a '
                         '--roofline-only run: per marked loop, the mean counter value per '
                         'dispatch of every kernel, and the copy calibration')
    ap.add_argument('--trace-summary', nargs='+', metavar='ARG',
                    help='summarise a rocprofv3 --kernel-trace CSV: CSV [t0_ns t1_ns] (the '
                         'monotonic clocks a timed step prints)')
    ap.add_argument('--roofline-only', action='store_true',
                    help='only the matvec-GEMM roofline measurement and a copy kernel of known '
                         'size (for rocprofv3 kernel-trace / counter runs)')
    args = ap.parse_args()
    _PLAN_AHEAD[0] = not args.no_plan_ahead
    _CYCLE_STEPS[0] = not args.no_cycle_steps
    _LZ_STREAMS[0] = args.lanczos_streams
    _PIPELINES[0] = max(1, args.pipelines)
    if args.gpus != 1:
        print('bench.py measures one GPU only (one-GPU build); --gpus %d is not supported'
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from tnalg_amd import _lib
    _lib.load()
    assert torch.cuda.is_available()
    if args.workload == 'c4':
        n_req = args.serving_requests or 512
        summ = serving_run(n_req, args.serving_rate, max_active=args.serving_max_active)
        print(json.dumps({
            'metric': 'DMRG solves/sec + p50/p99 latency (serving stream), 1 GPU',
            'value': summ['solves_per_s'], 'unit': 'solves/s', 'n_gpus': 1, 'steps': 1,
            'warmup': 0, 'ms_per_step': 1e3 * summ['span_s'], 'higher_is_better': True,
            'scaling': 'strong', 'vs_baseline': None, 'dtype': 'bf16/f32 matvec, f64 Lanczos+QR',
            'data': 'synthetic (Poisson arrivals, seeded random initial MPS)',
            'config': {'workload': 'c4: %d XXZ chain requests, N~U{16..96}, chi~U{32,64,128}, '
                                   'Jz~U[0.5,1.5], hz~U[0,0.5], 8 sweeps, Poisson arrivals at '
                                   '%g/s' % (n_req, args.serving_rate),
                       'batch': 'continuous', 'precision': 'mixed'},
            'latency_s': {'p50': summ['latency_p50_s'], 'p99': summ['latency_p99_s']},
            'serving': summ}))
        return
    if args.workload != 'c2':
        print(json.dumps(large_solve(args.workload, fp64_check=not args.no_plain)))
        return
    para = c2_para(args.n, args.chi, args.sweeps, sites=args.dmrg_sites,
                   two_site_sweeps=args.two_site_sweeps, chi_init=args.chi_init)
    if args.roofline_only:
        # the matvec GEMM kernel variants at a mid-chain step (HIP events
        # around each variant's launches), and the copy kernel the HBM
        # counters are scaled by
        precs = ('fp64',) if args.precision == 'fp64' else ('fp64', 'fp32', 'bf16')
        mv = matvec_roofline(para, args.batch, precs)
        torch.cuda._sleep(1000)   # marks the calibration copy for --pmc-summary
        print(json.dumps({'matvec_variants': mv, 'copy_calibration_bytes': copy_calibration()}))
        return
    if args.pmc_summary:
        print(json.dumps(pmc_summary(args.pmc_summary), indent=1))
        return
    if args.trace_summary:
        t = args.trace_summary
        print(json.dumps(trace_summary(t[0], *(int(x) for x in t[1:3])), indent=1))
        return
    seeds = list(range(args.batch))
    cold = None
    for w in range(args.warmup):
        # the same batch shape as the timed steps (other seeds): builds the
        # engine's step plans and descriptor sets
        dt, _, _, _ = run_batch(para, [1000 + s for s in seeds], args.precision)
        cold = dt if cold is None else cold
        print('warmup %d: %.3f s' % (w, dt), file=sys.stderr, flush=True)
    _freeze_caches()
    times, lat = [], []
    obs = None
    counters = None
    from tnalg_amd import _lib as L
    lib = L.load()
    prof = (L.ProfStat * L.PROF_NCLASS)()
    for k in range(args.steps):
        torch.cuda.synchronize()
        # the library times the first Krylov step of every 4th Lanczos call
        # of the first timed batch (the live roofline)
        if k == 0:
            L.check(lib.tnx_profile_begin(4, 8192), 'tnx_profile_begin')
        # host clocks at the ends of the timed step (to place it in a
        # profiler's kernel trace)
        c0 = (time.monotonic_ns(), time.clock_gettime_ns(time.CLOCK_BOOTTIME))
        dt, obs, counters, _ = run_batch(para, seeds, args.precision)
        c1 = (time.monotonic_ns(), time.clock_gettime_ns(time.CLOCK_BOOTTIME))
        if k == 0:
            nst = lib.tnx_profile_end(prof, L.PROF_NCLASS)
            if nst < 0:
                L.check(nst, 'tnx_profile_end')
        times.append(dt)
        lat += [dt] * len(seeds)
        print('step %d: %.3f s for %d solves (monotonic ns %d %d, boottime ns %d %d)'
              % (k, dt, len(seeds), c0[0], c1[0], c0[1], c1[1]), file=sys.stderr, flush=True)
    total = sum(times)
    value = args.batch * args.steps / total
    energies = np.array([o['E'] for o in obs])
    out = {'metric': 'DMRG solves/sec + p50/p99 latency at |dE/E|<=1e-9 vs FP64 plain, 1 GPU',
           'value': value, 'unit': 'solves/s', 'n_gpus': 1, 'steps': args.steps,
           'warmup': args.warmup, 'ms_per_step': 1e3 * total / args.steps,
           'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
           'dtype': 'f64' if args.precision == 'fp64' else 'bf16/f32 matvec, f64 Lanczos+QR',
           'data': 'synthetic (seeded random initial MPS drawn on the device, uniform XXZ '
                   'couplings)',
           'config': {'workload': 'XXZ chain N=%d open, Jxy=1 Jz=1 h=0, chi=%d, ncv=20, tol=1e-10, '
                                  '%d sweeps' % (args.n, args.chi, args.sweeps),
                      'batch': args.batch, 'precision': args.precision,
                      'dmrg_sites': args.dmrg_sites, 'two_site_sweeps': args.two_site_sweeps,
                      'chi_init': args.chi_init, 'pipelines': _PIPELINES[0]},
           'latency_s': {'p50': float(np.percentile(lat, 50)), 'p99': float(np.percentile(lat, 99))},
           'energy_per_site_mean': float(np.mean(energies)) / args.n,
           'cold_start_s': cold,
           'timed_outputs': 'per solve: eb_full, mx, mz, e_per_site, eb, entanglement spectrum '
                            'of every bond (SVD centre moves), ent, corr_x, corr_z',
           'solves_with_warnings': sum(1 for o in obs if o['warnings'])}
    out['engine_counters'] = {k: (float(v) if isinstance(v, float) else v)
                              for k, v in counters.items()}
    converged = None   # a converged host MPS of seed 0 (the CPU baseline's later sweeps)
    if not args.no_plain:
        # the f64 reference of every seed, batched on an engine warmed by a
        # batch of the same shape (other seeds), like the timed engine; and the
        # plain f64 path (one solve at a time, all-f64, warmed likewise)
        run_batch(para, [1000 + s for s in seeds], 'fp64')
        _freeze_caches()
        dt_b, ob64, _, _ = run_batch(para, seeds, 'fp64')
        e_b = np.array([o['E'] for o in ob64])
        rel = np.abs(energies - e_b) / np.abs(e_b)
        out['fp64_batched'] = {'solves_per_s': len(seeds) / dt_b, 'latency_s': dt_b,
                               'note': 'same batch, all-f64 schedule, warmed engine'}
        out['speedup_vs_fp64_batched'] = value * dt_b / len(seeds)
        out['dE_over_E_vs_fp64'] = {'max': float(rel.max()), 'median': float(np.median(rel)),
                                    'target': 1e-9, 'met': bool(rel.max() <= 1e-9)}
        out['max_abs_diff_vs_fp64'] = output_differences(obs, ob64)
        _release_engine('fp64')
        run_batch(para, [1000], 'fp64')     # warm the one-solve engine like the timed one
        _freeze_caches()
        dt_p, ob_p, _, _ = run_batch(para, seeds[:1], 'fp64', keep_mps=True)
        converged = ob_p[0]['mps']
        out['plain_fp64'] = {'solves_per_s': 1.0 / dt_p, 'latency_s': dt_p,
                             'energy': float(ob_p[0]['E'])}
        out['speedup_vs_plain_fp64'] = value / (1.0 / dt_p)
        print('fp64 batched %.2f solves/s, plain %.3f solves/s, max dE/E %.2e'
              % (len(seeds) / dt_b, 1.0 / dt_p, rel.max()), file=sys.stderr, flush=True)
    out['roofline'] = roofline_object(
        live_kernels(list(prof)),
        'every launch of the first Krylov step of every 4th Lanczos call of the first timed '
        'batch (%d solves) timed by its own blocks (device wall-clock stamps of the first '
        'block start and last block end, the kernel span a profiler trace shows); '
        'algorithmic work from the launches\' descriptors' % args.batch)
    if args.serving_requests > 0:
        out['serving'] = serving_run(args.serving_requests, args.serving_rate)
        print('serving: %s' % out['serving'], file=sys.stderr, flush=True)
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(args.cpu_seconds, converged, args.n, args.chi,
                                           args.sweeps)
    print(json.dumps(out))


if __name__ == '__main__':
    main()
