#!/usr/bin/env python3
"""bench.py -- Sinc-sum points/s of the MI355X evaluator on sincFeyn's three-loop benchmark.

Workload (one step): the 41 three-loop vertex integrals (`all3`: g3_00..g3_27, g3_lbl, v3_00..v3_07,
v3_p4Sc/Sm/Vc/Vm) at h = 0.7, N = 139 (the repo default 2*ceil(48/h)+1; K rounds to 72), eps = 1e-12,
top = 2, each graph's hierarchical sum run to completion on the device (one graph after another: the
plain per-graph path).  A point is one lattice-point evaluation as the reference counts it (res_t.nev).

    python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--no-cpu-baseline]

Two paths over the same jobs: the plain per-graph path (one hierarchical sum after another) and the
batched path (all jobs submitted to the library's job engine, which runs them on a pool of workers with
their own streams); `value` is the batched rate, `per_graph` the plain one, and every batched result is
checked to equal its per-graph result.

Prints one JSON line: points/s (value), per-job latency percentiles, the roofline of the dominant
kernel (FP64 operations per point of each integrand x points per launch / mean launch duration, from
HIP events on the library's stream), and the reference CPU implementation timed on a bounded sample of
the same workload on this machine's host cores.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

H, N, EPS, TOP = 0.7, 139, 1e-12, 2
PEAK_FP64 = 78.6          # TFLOP/s, MI355X FP64 vector (FMA counted as 2)
KIND = {0: 'level_eval', 1: 'level_dd', 2: 'level_sum', 3: 'nodes_eval', 4: 'nodes_dd', 5: 'level_eval_skip',
        6: 'level_split'}
KNAME = {'level_eval': 'k_eval<%s, list, AB=1>', 'level_eval_skip': 'k_eval<%s, list, AB=0>'}
# bounded CPU sample: graphs of the same workload whose reference run takes ~10-30 s on a few cores
CPU_SAMPLE = ['g3_01', 'g3_07', 'g3_09', 'g3_17', 'g3_19', 'g3_25', 'g3_27', 'v3_00', 'v3_03', 'v3_04', 'v3_06',
              'v3_07', 'v3_p4Sc', 'v3_p4Sm']


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return float('nan')
    k = (len(xs) - 1)*p/100.0
    f = math.floor(k)
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f])*(k - f)


def run_pass(S, graphs, lat, out=None):
    """the plain per-graph path: one hierarchical sum after another on the library's stream"""
    ev = 0
    for gi, name in graphs:
        t0 = time.perf_counter()
        r = S.hsum(gi, H, N, EPS, TOP)
        lat.append(time.perf_counter() - t0)
        ev += r['nev']
        if out is not None:
            out[name] = r
    return ev


def run_batched(S, graphs, lat, out=None):
    """the batched path: all jobs submitted at once to the engine (continuous batching over its workers);
    a job's latency is submit -> result"""
    ev = 0
    t0 = time.perf_counter()
    jobs = [(S.submit(gi, H, N, EPS, TOP), name) for gi, name in graphs]
    for j, name in jobs:
        r, tq, tr = S.wait(j)
        lat.append(tq + tr)
        ev += r['nev']
        if out is not None:
            out[name] = r
    return ev


def cpu_baseline():
    """the reference implementation (oracle/_ref/sincqed_ref, OpenMP on all host cores) on a bounded
    sample of the workload; symmetry fold off as on the device"""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle as O
    if not os.path.exists(O.SINCQED_REF):
        return None
    cores = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    ref, wall, _ = O.run_reference(H, N, EPS, CPU_SAMPLE, top=TOP, env={'SINCQED_SYM': '0'}, threads=cores, timeout=900)
    ev = sum(int(r['nev']) for r in ref.values())
    return {'value': ev/wall, 'unit': 'points/s', 'cores': cores, 'kind': 'reference',
            'sample': '%d of the 41 graphs (%s) at the same h, N, eps, top: %d evaluations in %.1f s' % (
                len(CPU_SAMPLE), ' '.join(CPU_SAMPLE), ev, wall)}


def roofline(S, opc):
    L = S.lib()
    ent = []
    for i in range(L.sfx_prof_count()):
        e = S.ProfEntry()
        S._check(L.sfx_prof_get(i, S.C.byref(e)))
        ent.append(dict(graph=e.graph.decode(), kind=KIND[e.kind], launches=e.launches, ms=e.ms, points=e.points,
                        skipped=e.skipped, dd=e.dd_points))
    if not ent:
        return None, []
    top = max(ent, key=lambda e: e['ms'])
    tot_ms = sum(e['ms'] for e in ent)

    def flops(e):
        oc = opc[e['graph']]
        if e['kind'] in ('level_eval', 'level_eval_skip', 'nodes_eval'):
            return e['skipped']*oc['ops0'] + (e['points'] - e['skipped'])*oc['ops1']
        return None
    fl = flops(top)
    rl = None
    if fl is not None:
        avg_s = top['ms']/top['launches']*1e-3
        ach = fl/top['launches']/avg_s/1e12
        rl = {'bound': 'fp64-valu', 'kernel': KNAME[top['kind']] % top['graph'] if top['kind'] in KNAME else top['kind'],
              'achieved': round(ach, 3), 'peak': PEAK_FP64, 'unit': 'TFLOP/s', 'frac': round(ach/PEAK_FP64, 4),
              'traffic': None, 'launches': top['launches'], 'avg_launch_ms': round(top['ms']/top['launches'], 4),
              'flops_per_launch': fl/top['launches'], 'share_of_device_time': round(top['ms']/tot_ms, 3)}
    if rl:
        # HBM bytes per launch of this kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes
        # (each counter scaled by what it reads for the library's copy kernel of known size)
        import glob
        for p in glob.glob(os.path.join(ROOT, 'profiles', 'traffic_*.json')):
            t = json.load(open(p))
            if t.get('kernel') == rl['kernel']:
                rl['traffic'] = t['traffic_bytes_per_dispatch']
                rl['traffic_source'] = os.path.relpath(p, ROOT)
    agg = [e for e in ent if e['kind'] in ('level_eval', 'level_eval_skip')]
    if agg and rl:
        f_all = sum(flops(e) for e in agg)
        ms_all = sum(e['ms'] for e in agg)
        rl['all_level_eval'] = {'achieved': round(f_all/(ms_all*1e-3)/1e12, 3), 'ms': round(ms_all, 1),
                                'share_of_device_time': round(ms_all/tot_ms, 3)}
    return rl, ent


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=1)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--graphs', default='all3', help='selector (default all3); a subset is not the benchmark')
    ap.add_argument('--profile-out', default=None, help='write the per-kernel table as JSON here')
    ap.add_argument('--workers', type=int, default=16, help='engine workers of the batched path')
    ap.add_argument('--no-batched', action='store_true',
                    help='run the per-graph pass only (one host thread: the form profiled with rocprofv3)')
    ap.add_argument('--copy-probe', type=int, default=0,
                    help='first run the library\'s plain copy kernel over this many bytes (counter calibration)')
    a = ap.parse_args()
    if a.gpus != 1:
        print('bench.py: this build measures one GPU; the Sinc sum of one graph is not split across GPUs here '
              '(--gpus %d is not supported)' % a.gpus, file=sys.stderr)
        sys.exit(2)
    from sincfeyn_amd import sfx as S
    info = S.init(0)
    opc = json.load(open(os.path.join(ROOT, 'sincfeyn_amd', '_gen', 'opcounts.json')))
    graphs = [(i, S.graphs()[i]['name']) for i in S.select(*a.graphs.split(','))]
    if a.copy_probe:
        ms = S.C.c_double()
        S._check(S.lib().sfx_copy_probe(S.C.c_long(a.copy_probe), S.C.byref(ms)))
        print('copy probe: %d bytes in %.3f ms' % (a.copy_probe, ms.value), file=sys.stderr)
    # the plain per-graph path, profiled (kernel durations from HIP events on the library's stream)
    lat_pg = []
    for _ in range(a.warmup):
        run_pass(S, graphs, [])
    S._check(S.lib().sfx_sync())
    S._check(S.lib().sfx_prof_enable(1))
    t0 = time.perf_counter()
    ev_pg = 0
    per_graph = {}
    for _ in range(a.steps):
        ev_pg += run_pass(S, graphs, lat_pg, per_graph)
    S._check(S.lib().sfx_sync())
    dt_pg = time.perf_counter() - t0
    S._check(S.lib().sfx_prof_enable(0))
    rl, ent = roofline(S, opc)
    if a.profile_out:
        with open(a.profile_out, 'w') as f:
            json.dump(ent, f, indent=1)
    if a.no_batched:
        print(json.dumps({'metric': 'Sinc-sum points/s, per-graph pass only (profiling form)', 'value': ev_pg/dt_pg,
                          'unit': 'points/s', 'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup,
                          'ms_per_step': dt_pg*1e3/a.steps, 'higher_is_better': True, 'path': 'per-graph',
                          'roofline': rl, 'device': info}))
        return
    # the batched path: the same jobs through the engine (headline value)
    S.engine_start(a.workers, 0)
    order = sorted(graphs, key=lambda g: -S.graphs()[g[0]]['D'])      # the widest lattices first
    for _ in range(a.warmup):
        run_batched(S, order, [])
    lat = []
    batched = {}
    t0 = time.perf_counter()
    ev = 0
    for _ in range(a.steps):
        ev += run_batched(S, order, lat, batched)
    dt = time.perf_counter() - t0
    same = all(batched[n]['S'] == per_graph[n]['S'] and batched[n]['nev'] == per_graph[n]['nev'] and
               batched[n]['nq'] == per_graph[n]['nq'] for _, n in graphs)
    cpu = None if a.no_cpu_baseline else cpu_baseline()
    out = {
        'metric': 'Sinc-sum points/s batched vs per-graph; p50/p99 job latency at <=1e-12 rel',
        'value': ev/dt, 'unit': 'points/s', 'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup,
        'ms_per_step': dt*1e3/a.steps, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64', 'data': 'deterministic Sinc lattice of the named integrals (no input data)',
        'config': {'workload': 'all41 three-loop integrals (h=0.7, N=139, eps=1e-12), batched through the job engine',
                   'graphs': len(graphs), 'h': H, 'N': N, 'K': S.lattice_K(N, TOP), 'eps': EPS, 'top': TOP,
                   'points_per_step': ev//a.steps, 'path': 'batched', 'workers': a.workers},
        'latency_s': {'p50': percentile(lat, 50), 'p99': percentile(lat, 99), 'jobs': len(lat)},
        'per_graph': {'value': ev_pg/dt_pg, 'ms_per_step': dt_pg*1e3/a.steps,
                      'latency_s': {'p50': percentile(lat_pg, 50), 'p99': percentile(lat_pg, 99)}},
        'batched_over_per_graph': (ev/dt)/(ev_pg/dt_pg),
        'batched_results_identical_to_per_graph': same,
        'roofline': rl,
        'cpu_baseline': cpu,
        'device': info,
    }
    print(json.dumps(out))


if __name__ == '__main__':
    main()
