"""serve.py -- a seeded Poisson request stream of (graph, h, eps) jobs served by the library's job engine.

The engine (C ABI sfx_job_submit / sfx_job_wait, include/sfx.h) does the serving: a pool of workers with
their own HIP streams takes queued jobs in arrival order as workers free up, reserves each refinement
level's HBM footprint from its budget, and runs the jobs' device work concurrently.  This module is the
client: it draws the stream, submits every job at its arrival time, and reports latency percentiles
(arrival -> result) and throughput.

Stream: splitmix64 from a 64-bit seed; inter-arrival times Exp(rate); graph uniform over the chosen
set; h uniform over `hs`, eps uniform over `epss`; N = 2 ceil(48/h) + 1 (the repository default) and
top from the reference's rule (main.c:211-215).

    python -m sincfeyn_amd.serve --rate 10 --jobs 100 --seed 1 [--graphs all2,all3] [--workers 16]
    python -m sincfeyn_amd.serve --sweep 1,2,5,10,20,50 --jobs 150 --seed 7 --check tests/golden/fixtures
"""
import argparse
import json
import math
import sys
import threading
import time

from . import sfx

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.x = seed & MASK

    def next(self):
        self.x = (self.x + 0x9E3779B97F4A7C15) & MASK
        z = self.x
        z = ((z ^ (z >> 30))*0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27))*0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11)*(1.0/(1 << 53))


def default_N(h):
    return 2*math.ceil(48.0/h - 1e-9) + 1


def make_stream(seed, rate, njobs, graphs, hs=(1.0, 0.9, 0.8, 0.7), epss=(1e-10, 1e-12)):
    """[(arrival_s, graph_index, h, N, eps, top)]"""
    rng = SplitMix64(seed)
    t, out = 0.0, []
    for _ in range(njobs):
        t += -math.log(1.0 - rng.uniform())/rate
        g = graphs[int(rng.uniform()*len(graphs))]
        h = hs[int(rng.uniform()*len(hs))]
        eps = epss[int(rng.uniform()*len(epss))]
        out.append((t, g, h, default_N(h), eps, sfx.default_top(h)))
    return out


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return float('nan')
    k = (len(xs) - 1)*p/100.0
    f = int(math.floor(k))
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f])*(k - f)


def serve(stream, results=None, progress=False, wait_timeout=600.0):
    """submit each job at its arrival time; returns the report dict (and fills results[i] = result dict);
    with progress, one line per finished job on stderr (in submission order: job, graph, h, eps, queue and
    run seconds), so a long stream shows where it is.  A job not finished wait_timeout seconds after the
    wait for it began raises (SfxError naming its stage and level) instead of blocking."""
    t0 = time.perf_counter()
    ids = [None]*len(stream)
    failed = []                                 # the client thread's exception, re-raised by the waiting loop

    def client():
        try:
            for i, (a, g, h, N, eps, top) in enumerate(stream):
                dt = t0 + a - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids[i] = sfx.submit(g, h, N, eps, top)
        except BaseException as e:             # noqa: BLE001 -- handed to the main thread
            failed.append(e)

    th = threading.Thread(target=client)
    th.start()
    lat, qs, ev = [], [], 0
    names = [g['name'] for g in sfx.graphs()] if progress else None
    for i in range(len(stream)):
        while ids[i] is None:
            if failed:
                th.join()
                raise failed[0]
            time.sleep(0.0005)
        r, tq, tr = sfx.wait(ids[i], timeout=wait_timeout)
        if progress:
            a_, g, h, N, eps, top = stream[i]
            print('job %d/%d %s h=%g eps=%g: queued %.3f s, ran %.3f s, %d points' % (
                i + 1, len(stream), names[g], h, eps, tq, tr, r['nev']),
                file=sys.stderr, flush=True)
        lat.append(tq + tr)
        qs.append(tq)
        ev += r['nev']
        if results is not None:
            results[i] = dict(r, queue_s=tq, run_s=tr)
    th.join()
    wall = time.perf_counter() - t0
    return {'jobs': len(stream), 'points': ev, 'wall_s': wall, 'points_per_s': ev/wall,
            'latency_p50_s': percentile(lat, 50), 'latency_p99_s': percentile(lat, 99),
            'queue_p50_s': percentile(qs, 50), 'queue_p99_s': percentile(qs, 99),
            'offered_rate': len(stream)/stream[-1][0] if stream else 0}


def reference_records(folder):
    """{(graph, h, N, eps): {field: text}} from the "#R" result records in folder/nc_h*_N*_eps*.txt: one line
    per graph with every field of the reference's res_t (the reference built without FP contraction, with a
    one-line results printer; tests/golden/make_fixtures.py).  Where folder/refgpu_h*_N*_eps*.txt holds a
    record of the same job (the reference driver's GPU build), its counts replace the CPU build's: they
    differ for a kernel whose CPU double half rounds differently from the QREAL half that the GPU evaluates"""
    import glob
    import os
    out = {}
    for build in ('nc', 'refgpu'):
        for p in glob.glob(os.path.join(folder, build + '_h*_N*_eps*.txt')):
            _, h, N, eps = os.path.basename(p)[:-4].split('_')
            for line in open(p):
                if line.startswith('#R '):
                    t = line.split()
                    key = (t[1], float(h[1:]), int(N[1:]), float(eps[3:]))
                    rec = {t[i]: t[i + 1] for i in range(2, len(t) - 1, 2)}
                    if build == 'nc':
                        out[key] = rec
                    elif key in out:
                        out[key].update({k: rec[k] for k in ('nev', 'nq', 'kept', 'skipped')})
    return out


def check_job(r, rec, rel=1e-12):
    """a job's result against the reference's record: evaluation, fallback and kept counts equal, S_h within
    rel (the reference built without FP contraction, whose arithmetic the device reproduces)"""
    S0 = float(rec['S0'])
    d = abs(r['S'][0] - S0)/abs(S0) if S0 else abs(r['S'][0])
    same = (r['nev'], r['nq'], r['kept_top']) == (int(rec['nev']), int(rec['nq']), int(rec['kept']))
    return same and d <= rel, d


def sweep(rates, seed, graphs, names, refs=None, jobs_for=None, progress=False):
    """one stream per arrival rate, from one library instance; each job checked against refs when given.
    Returns one report per rate (with the per-job records)"""
    out = []
    for rate in rates:
        n = jobs_for(rate)
        stream = make_stream(seed, rate, n, graphs)
        res = [None]*n
        rep = serve(stream, res, progress=progress)
        jobs, worst, bad, checked = [], 0.0, 0, 0
        for (a, g, h, N, eps, top), r in zip(stream, res):
            j = {'graph': names[g], 'h': h, 'N': N, 'eps': eps, 'top': top, 'arrival_s': round(a, 4),
                 'S_h': r['S'][0], 'nev': r['nev'], 'nq': r['nq'], 'kept': r['kept_top'],
                 'queue_s': round(r['queue_s'], 5), 'run_s': round(r['run_s'], 5),
                 'coarsest_s': round(r['t_grow'], 5), 'dilation_s': round(r['t_dilate'], 5), 'refinement_s': round(r['t_refine'], 5)}
            rec = refs.get((names[g], h, N, eps)) if refs else None
            if rec is not None:
                ok, d = check_job(r, rec)
                checked += 1
                bad += not ok
                worst = max(worst, d)
                j['matches_reference'] = ok
            jobs.append(j)
        rep.update(rate=rate, seed=seed)
        if refs is not None:
            rep['reference_check'] = {'jobs_checked': checked, 'mismatches': bad, 'max_rel_value_diff': worst,
                                      'all_jobs_match_reference': checked == n and bad == 0}
        rep['per_job'] = jobs
        out.append(rep)
        print('rate %g/s: %d jobs, %.3g points/s, p50 %.3f s, p99 %.3f s%s' % (
            rate, n, rep['points_per_s'], rep['latency_p50_s'], rep['latency_p99_s'],
            '' if refs is None else ', %d/%d checked, %d mismatches' % (checked, n, bad)), file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--rate', type=float, default=10.0)
    ap.add_argument('--jobs', type=int, default=100)
    ap.add_argument('--seed', type=int, default=1)
    ap.add_argument('--graphs', default='all2,all3')
    ap.add_argument('--workers', type=int, default=16)
    ap.add_argument('--progress', action='store_true', help='one line per finished job on stderr')
    ap.add_argument('--sweep', default=None, help='comma-separated arrival rates (jobs/s), one stream each, '
                    'from one library instance; one JSON line per rate')
    ap.add_argument('--window', type=float, default=15.0,
                    help='with --sweep: the jobs of a rate are rate x window, between 20 and --jobs')
    ap.add_argument('--check', default=None, help='folder of the reference\'s "#R" result records '
                    '(tests/golden/fixtures): every job is checked against its record')
    a = ap.parse_args()
    sfx.hw_queues(a.workers)
    sfx.init(0)
    sfx.engine_start(a.workers, 0)
    graphs = sfx.select(*a.graphs.split(','))
    if a.sweep:
        names = [g['name'] for g in sfx.graphs()]
        refs = reference_records(a.check) if a.check else None
        reps = sweep([float(x) for x in a.sweep.split(',')], a.seed, graphs, names, refs,
                     lambda r: int(max(20, min(a.jobs, round(r*a.window)))), a.progress)
        for rep in reps:
            rep.update(workers=a.workers)
            print(json.dumps(rep))
        return 0 if refs is None or all(r['reference_check']['all_jobs_match_reference'] for r in reps) else 1
    rep = serve(make_stream(a.seed, a.rate, a.jobs, graphs), progress=a.progress)
    rep.update(rate=a.rate, seed=a.seed, workers=a.workers)
    print(json.dumps(rep))


if __name__ == '__main__':
    sys.exit(main())
