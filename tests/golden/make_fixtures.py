#!/usr/bin/env python3
"""Writes the per-graph reference results of the benchmark configurations from the reference
implementation itself (oracle/_ref/sincqed_ref and sincqed_ref_nc, built from the reference sources by
oracle/Makefile), one "#R" record per (graph, h, N, eps) with every field of res_t at full precision.

Sets (N = 2 ceil(48/h) + 1, the reference's default N for h; top from main.c:211-215; SINCQED_SYM=0, the
semantics of the reference's GPU path, sum.c:571):
  cfg2   all3            h 0.7  N 139  eps 1e-12            (41 integrals)
  cfg3   all2 + all3     h 1.0 / 0.8 / 0.7, N 97 / 121 / 139, eps 1e-12   (cfg2 covers all3 at 0.7)
  serve  all2 + all3     h 0.9 N 109 eps 1e-12, and h 1.0 / 0.9 / 0.8 / 0.7 at eps 1e-10
  stream the (graph, h, eps) jobs of the serving sweep (sincfeyn_amd/serve.py --sweep 1,2,5,10,20,50
         --jobs 150 --window 15 --seed 7) that cfg2 and cfg3 do not already hold: the seeded streams are
         drawn here exactly as the sweep draws them (graphs all2 + all3 in the library's registry order)
for two builds:
  nc     sincqed_ref_nc  the reference without FP contraction (its GPU path's -fmad=false arithmetic):
                         evaluation, fallback and kept counts must equal the device's exactly
  ref    sincqed_ref     the reference's default build (FMA contraction): values within 1e-12

(g3_lbl's counts on the device follow the reference's GPU build, not its CPU build: tests/golden/make_refgpu.sh
writes that build's records for it, refgpu_*.txt.)

Each graph runs as its own process (several at once, each on one or a few OpenMP threads: the
reference's refinement sum is serial, so processes scale where threads do not).  Records are appended to
tests/golden/fixtures/<build>_h<h>_N<N>_eps<eps>.txt as they finish; graphs already there are skipped,
so an interrupted run resumes.

    python3 tests/golden/make_fixtures.py [--builds nc,ref] [--sets cfg2,cfg3,serve] [--threads 8]
"""
import argparse
import concurrent.futures as cf
import math
import os
import subprocess
import sys
import threading
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, 'fixtures')
BIN = {'nc': os.path.join(ROOT, 'oracle', '_ref', 'sincqed_ref_nc'),
       'ref': os.path.join(ROOT, 'oracle', '_ref', 'sincqed_ref')}
ALL2 = ['g2_00', 'g2_01', 'g2_02', 'g2_03', 'v2_00']
ALL3 = (['g3_%02d' % i for i in range(28)] + ['g3_lbl']
        + ['v3_%02d' % i for i in range(8)] + ['v3_p4Sc', 'v3_p4Sm', 'v3_p4Vc', 'v3_p4Vm'])
DIM = {'g2_00': 4, 'g2_01': 3, 'g2_02': 4, 'g2_03': 3, 'v2_00': 3}


def default_N(h):
    return 2*math.ceil(48/h - 1e-9) + 1


def default_top(h, ts=3.0):
    top = 1
    while top < 2 and (1 << (top + 1))*h <= ts*(1 + 1e-9):
        top += 1
    return top


SWEEP = dict(seed=7, rates=(1, 2, 5, 10, 20, 50), jobs=150, window=15.0)


def stream_jobs():
    """{(h, eps): [graph]} of the serving sweep's streams (serve.sweep: one stream per rate, rate x window
    jobs, at least 20 and at most `jobs`)"""
    sys.path.insert(0, ROOT)
    from sincfeyn_amd import serve
    out = {}
    names = ALL2 + ALL3
    for rate in SWEEP['rates']:
        n = int(max(20, min(SWEEP['jobs'], round(rate*SWEEP['window']))))
        for a, g, h, N, eps, top in serve.make_stream(SWEEP['seed'], rate, n, list(range(len(names)))):
            lst = out.setdefault((h, eps), [])
            if names[g] not in lst:
                lst.append(names[g])
    return out


def sets(names):
    out = []
    if 'cfg2' in names:
        out.append((0.7, 1e-12, ALL3))
    if 'cfg3' in names:
        out += [(1.0, 1e-12, ALL2 + ALL3), (0.8, 1e-12, ALL2 + ALL3), (0.7, 1e-12, ALL2)]
    if 'serve' in names:
        out.append((0.9, 1e-12, ALL2 + ALL3))
        out += [(h, 1e-10, ALL2 + ALL3) for h in (1.0, 0.9, 0.8, 0.7)]
    if 'stream' in names:
        for (h, eps), graphs in sorted(stream_jobs().items()):
            if eps == 1e-12 and h in (1.0, 0.8, 0.7):
                graphs = [g for g in graphs if g not in ALL3 + ALL2]      # held by cfg2 and cfg3
            if graphs:
                out.append((h, eps, graphs))
    return out


def path(build, h, eps):
    return os.path.join(OUT, '%s_h%g_N%d_eps%g.txt' % (build, h, default_N(h), eps))


def done_graphs(p):
    if not os.path.exists(p):
        return set()
    return {l.split()[1] for l in open(p) if l.startswith('#R ')}


def cost(name, h, eps):
    """a rough order of the runs, largest first: the light-by-light kernel is ~10x the others per point"""
    d = DIM.get(name, 7 if name.startswith('g3') else 5)
    c = (1.0/h)**d*(4.0 if eps < 1e-11 else 1.0)
    return c*(30 if name == 'g3_lbl' else 1)


def dedupe():
    """keep the first record (and its #X lines) of each graph in every fixture file: two runs that overlapped
    write the same graph twice (identical but for the time)"""
    import glob
    for p in sorted(glob.glob(os.path.join(OUT, '*.txt'))):
        out, seen, keep = [], set(), True
        for line in open(p):
            if line.startswith('#R '):
                g = line.split()[1]
                keep = g not in seen
                seen.add(g)
            elif not line.startswith('#X '):
                keep = True
            if keep:
                out.append(line)
        with open(p, 'w') as f:
            f.writelines(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--builds', default='nc,ref')
    ap.add_argument('--sets', default='cfg2,cfg3,serve')
    ap.add_argument('--threads', type=int, default=os.cpu_count() or 8, help='total OpenMP threads in use at once')
    ap.add_argument('--lbl-threads', type=int, default=4, help='threads of each g3_lbl run')
    ap.add_argument('--dedupe', action='store_true', help='only drop repeated records from the fixture files')
    a = ap.parse_args(argv)
    if a.dedupe:
        return dedupe()
    os.makedirs(OUT, exist_ok=True)
    tasks = []
    for b in a.builds.split(','):
        if not os.path.exists(BIN[b]):
            sys.exit('%s is not built: make -C oracle ref' % BIN[b])
        for h, eps, graphs in sets(a.sets.split(',')):
            p = path(b, h, eps)
            have = done_graphs(p)
            for g in graphs:
                if g not in have:
                    tasks.append((b, h, eps, g))
    # builds in the order given, largest runs first within a build
    order = {b: i for i, b in enumerate(a.builds.split(','))}
    tasks.sort(key=lambda t: (order[t[0]], -cost(t[3], t[1], t[2])))
    print('%d runs' % len(tasks), flush=True)
    lock = threading.Lock()
    cv = threading.Condition()
    free = [a.threads]              # OpenMP threads not in use; a run takes all of its threads at once

    def run(t):
        b, h, eps, g = t
        nth = min(a.lbl_threads, a.threads) if g == 'g3_lbl' else 1
        with cv:
            cv.wait_for(lambda: free[0] >= nth)
            free[0] -= nth
        try:
            env = dict(os.environ, SINCQED_CKPT='0', SINCQED_PROGRESS='0', SINCQED_SYM='0', OMP_NUM_THREADS=str(nth))
            env.pop('SINCQED_RESULTS', None)
            N, top = default_N(h), default_top(h)
            t0 = time.time()
            p = subprocess.run(['nice', '-n', '19', BIN[b], 'run', repr(h), str(N), repr(eps), str(top), g],
                               capture_output=True, text=True, env=env, cwd='/tmp')
            if p.returncode:
                raise RuntimeError('%s %s failed:\n%s' % (b, g, p.stderr[-2000:]))
            rec = [l for l in p.stdout.splitlines() if l.startswith('#R ')]
            ext = [l for l in p.stdout.splitlines() if 'eps extrapolation' in l]
            assert len(rec) == 1, p.stdout
            with lock:
                fp = path(b, h, eps)
                fresh = not os.path.exists(fp)
                with open(fp, 'a') as f:
                    if fresh:
                        f.write('# %s run %r %d %r %d GRAPH  (SINCQED_SYM=0 SINCQED_CKPT=0)\n' % (
                            os.path.basename(BIN[b]), h, N, eps, top))
                    f.write(rec[0] + '\n')
                    for l in ext:
                        f.write('#X ' + l.lstrip('# ') + '\n')
                print('%s h=%g eps=%g %-8s %.1fs' % (b, h, eps, g, time.time() - t0), flush=True)
        finally:
            with cv:
                free[0] += nth
                cv.notify_all()

    with cf.ThreadPoolExecutor(a.threads) as ex:
        for f in cf.as_completed([ex.submit(run, t) for t in tasks]):
            f.result()


if __name__ == '__main__':
    main()
