"""mixed.py -- a heterogeneous batch of hierarchical sums with each integral's error against an all-DD run.

The batch is every (graph, h) of the chosen graphs and steps -- by default the 46 integrals of all2 and
all3 at h = 1.0, 0.8, 0.7 (138 jobs of uneven lattice depth, D = 3..7), N = 2 ceil(48/h) + 1 and top by
the reference's rule (main.c:211-215), eps = 1e-12.  The jobs run twice on the library:

  * one after another on the library's own stream (the plain per-graph path), and
  * all submitted at once to the job engine (continuous batching over its workers),

and the results must be identical.  Then the same jobs run through the engine in the reference's all-DD
mode (G2_ABSERR=0 SINCQED_ABREL=0 SINCQED_ABSKIP=0: every point with a nonzero |.| sum is recomputed in
double-double, eval.c:36-38), and each integral's value is compared with its all-DD value.

    python -m sincfeyn_amd.mixed [--hs 1.0,0.8,0.7] [--graphs all2,all3] [--workers 16]
                                 [--dd all|none|N:SEED] [--qd] [--out FILE.jsonl]

--dd N:SEED runs the all-DD comparison on N jobs drawn with splitmix64 from SEED (the full set takes
several minutes of device time; the three-loop light-by-light integral dominates it).  --qd adds the
all-QD run (every point in quad-double) of the two-loop jobs, the error reference of both.  One JSON line per
job, then a summary line.
"""
import argparse
import json
import math
import sys
import time

from . import sfx
from .serve import SplitMix64, default_N, percentile

ALL_DD = sfx.ALL_DD
ALL_QD = sfx.ALL_QD


def job_list(graphs, hs, eps):
    return [(g, h, default_N(h), eps, sfx.default_top(h)) for h in hs for g in graphs]


def run_per_graph(jobs):
    out, t0 = [], time.perf_counter()
    for g, h, N, eps, top in jobs:
        out.append(sfx.hsum(g, h, N, eps, top))
    return out, time.perf_counter() - t0


def run_batched(jobs, **kw):
    """all jobs submitted at once; (results, wall, latencies)"""
    t0 = time.perf_counter()
    ids = [sfx.submit(g, h, N, eps, top, **kw) for g, h, N, eps, top in jobs]
    res, lat = [], []
    for j in ids:
        r, tq, tr = sfx.wait(j)
        res.append(r)
        lat.append(tq + tr)
    return res, time.perf_counter() - t0, lat


def same(a, b):
    return all((x == y) or (math.isnan(x) and math.isnan(y)) for x, y in zip(a['S'], b['S'])) and \
        (a['nev'], a['nq'], a['kept_top']) == (b['nev'], b['nq'], b['kept_top'])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--hs', default='1.0,0.8,0.7')
    ap.add_argument('--graphs', default='all2,all3')
    ap.add_argument('--eps', type=float, default=1e-12)
    ap.add_argument('--workers', type=int, default=16)
    ap.add_argument('--dd', default='all', help="all, none, or N:SEED (a seeded subset of N jobs)")
    ap.add_argument('--qd', action='store_true', help='also the all-QD run of the two-loop jobs')
    ap.add_argument('--progress', action='store_true', help='the all-DD jobs report their refinement blocks on stderr')
    ap.add_argument('--out', default=None)
    a = ap.parse_args(argv)
    info = sfx.init(0)
    G = sfx.graphs()
    graphs = sfx.select(*a.graphs.split(','))
    hs = [float(x) for x in a.hs.split(',')]
    jobs = job_list(graphs, hs, a.eps)
    out = open(a.out, 'w') if a.out else sys.stdout

    def emit(d):
        out.write(json.dumps(d) + '\n')
        out.flush()

    print('# %d jobs; per-graph pass' % len(jobs), file=sys.stderr, flush=True)
    pg, t_pg = run_per_graph(jobs)
    print('# batched pass (%d workers)' % a.workers, file=sys.stderr, flush=True)
    sfx.engine_start(a.workers, 0)
    bt, t_bt, lat = run_batched(jobs)
    ident = all(same(x, y) for x, y in zip(pg, bt))
    dd = {}
    if a.dd != 'none':
        if a.dd == 'all':
            pick = list(range(len(jobs)))
        else:
            n, seed = (int(x) for x in a.dd.split(':'))
            rng, pick = SplitMix64(seed), []
            while len(pick) < min(n, len(jobs)):
                i = int(rng.uniform()*len(jobs))
                if i not in pick:
                    pick.append(i)
            pick.sort()
        print('# all-DD pass: %d jobs' % len(pick), file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        # the longest of these jobs (the light-by-light integral, every point in double-double) run for
        # minutes: with --progress they report their refinement blocks, and every job is reported as it
        # finishes, in whatever order
        kw = dict(ALL_DD, verbose=1) if a.progress else ALL_DD
        ids = dict((sfx.submit(*jobs[i], **kw), i) for i in pick)
        while ids:
            for j in [j for j in ids if sfx.poll(j) >= 2]:
                i = ids.pop(j)
                dd[i] = sfx.wait(j)[0]
                print('#   all-DD %s h=%g done (%.0f s)' % (G[jobs[i][0]]['name'], jobs[i][1], time.perf_counter() - t0),
                      file=sys.stderr, flush=True)
            time.sleep(0.05)
        t_dd = time.perf_counter() - t0
    qd = {}
    if a.qd:
        pick = [i for i, j in enumerate(jobs) if G[j[0]]['L'] == 2]
        print('# all-QD pass: %d two-loop jobs' % len(pick), file=sys.stderr, flush=True)
        ids = [(i, sfx.submit(*jobs[i], **ALL_QD)) for i in pick]
        for i, j in ids:
            qd[i] = sfx.wait(j)[0]
    sfx.engine_stop()
    worst = 0.0
    for i, ((g, h, N, eps, top), r) in enumerate(zip(jobs, bt)):
        v, err = sfx.graph_value(G[g], r)
        d = dict(graph=G[g]['name'], D=G[g]['D'], h=h, N=N, eps=eps, top=top, value=v, err=err, nev=r['nev'],
                 nq=r['nq'], t=r['t'], latency_s=lat[i])
        if i in dd:
            vd, _ = sfx.graph_value(G[g], dd[i])
            rel = abs(v - vd)/abs(vd) if vd != 0 else abs(v - vd)
            d.update(value_all_dd=vd, nq_all_dd=dd[i]['nq'], rel_diff_all_dd=rel,
                     counts_equal_all_dd=(r['nev'], r['kept_top']) == (dd[i]['nev'], dd[i]['kept_top']))
            worst = max(worst, rel)
        if i in qd:
            vq, _ = sfx.graph_value(G[g], qd[i])
            d.update(value_all_qd=vq, rel_diff_all_qd=abs(v - vq)/abs(vq) if vq != 0 else abs(v - vq))
            if i in dd:
                d['rel_diff_all_dd_vs_all_qd'] = abs(d['value_all_dd'] - vq)/abs(vq) if vq != 0 else abs(d['value_all_dd'] - vq)
        emit(d)
    ev = sum(r['nev'] for r in bt)
    summ = dict(summary=True, jobs=len(jobs), hs=hs, eps=a.eps, points=ev, device=info,
                per_graph=dict(wall_s=t_pg, points_per_s=ev/t_pg),
                batched=dict(wall_s=t_bt, points_per_s=ev/t_bt, workers=a.workers,
                             latency_p50_s=percentile(lat, 50), latency_p99_s=percentile(lat, 99)),
                batched_over_per_graph=t_pg/t_bt, batched_identical_to_per_graph=ident)
    if dd:
        summ['all_dd'] = dict(jobs=len(dd), selection=a.dd, wall_s=t_dd,
                              points=sum(x['nev'] for x in dd.values()), dd_points=sum(x['nq'] for x in dd.values()),
                              max_rel_diff=worst)
    if qd:
        summ['all_qd'] = dict(jobs=len(qd), points=sum(x['nev'] for x in qd.values()))
    emit(summ)
    return 0


if __name__ == '__main__':
    sys.exit(main())
