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
            results[i] = r
    th.join()
    wall = time.perf_counter() - t0
    return {'jobs': len(stream), 'points': ev, 'wall_s': wall, 'points_per_s': ev/wall,
            'latency_p50_s': percentile(lat, 50), 'latency_p99_s': percentile(lat, 99),
            'queue_p50_s': percentile(qs, 50), 'queue_p99_s': percentile(qs, 99),
            'offered_rate': len(stream)/stream[-1][0] if stream else 0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--rate', type=float, default=10.0)
    ap.add_argument('--jobs', type=int, default=100)
    ap.add_argument('--seed', type=int, default=1)
    ap.add_argument('--graphs', default='all2,all3')
    ap.add_argument('--workers', type=int, default=16)
    ap.add_argument('--progress', action='store_true', help='one line per finished job on stderr')
    a = ap.parse_args()
    sfx.init(0)
    sfx.engine_start(a.workers, 0)
    graphs = sfx.select(*a.graphs.split(','))
    rep = serve(make_stream(a.seed, a.rate, a.jobs, graphs), progress=a.progress)
    rep.update(rate=a.rate, seed=a.seed, workers=a.workers)
    print(json.dumps(rep))


if __name__ == '__main__':
    main()
