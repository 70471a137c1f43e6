#!/usr/bin/env python3
"""bench.py -- Sinc-sum points/s of the MI355X evaluator on sincFeyn's batched heterogeneous workload.

Workload: the 46 two- and three-loop integrals (`all2` + `all3`: g2_00..g2_03, v2_00, g3_00..g3_27,
g3_lbl, v3_00..v3_07, v3_p4Sc/Sm/Vc/Vm) at h = 1.0, 0.8 and 0.7 (N = 97, 121, 139: the repo default
2*ceil(48/h)+1; top 1, 1, 2 by the reference's rule), eps = 1e-12 -- 138 hierarchical sums with lattice
depths and dimensions from 3 to 7, 2.66e11 lattice-point evaluations.  A point is one evaluation as the
reference counts it (res_t.nev).

A step is one tenth of the 138 jobs (the jobs sorted widest first and dealt round-robin into ten
slices; step i runs slice i mod 10), so ten steps are one pass over the whole workload.  The batched
path submits the steps' jobs to the library's job engine (a pool of workers with their own streams,
continuous batching); `value` is its rate over the timed steps.

Beside it, on the same machine:
  per_graph    one pass of the same 138 jobs as plain per-graph sums (one after another on one stream)
  plain_gpu    the reference's own host driver built as its GPU build (driver/sum.c, sumkey.h, eval.c with
               -DSQ_GPU: grow, dilation, prep, sums and double-double fallback on the host, each batch
               through the gpu.h back end -- here this library's, src/sfx_gpu.c), one graph per launch,
               on a bounded sample of the jobs; the batched and per-graph paths on the same sample
  cpu_baseline the reference's OpenMP CPU build on a bounded sample, on this machine's host cores
  roofline     the dominant kernel of a profiled per-graph pass over the h = 0.7 jobs of the three-loop
               integrals: FP64 operations per point x points per launch / mean launch duration (HIP
               events on the library's stream), algorithmic and executed (committed VALU counters)
  reference    each batched result against the reference's own results (committed fixtures written by
               tests/golden/make_fixtures.py): counts exact, values within 1e-12
  served       p50/p99 latency of a seeded Poisson stream of (graph, h, eps) jobs at 10 jobs/s through the same
               engine, each job checked against the reference's result record

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu-baseline] [--no-plain] [--no-serve]
"""
import argparse
import glob
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

EPS = 1e-12
CFG = [(1.0, 97), (0.8, 121), (0.7, 139)]
SLICES = 10
PEAK_FP64 = 78.6          # TFLOP/s, MI355X FP64 vector (FMA counted as 2)
KIND = {0: 'level_eval', 1: 'level_dd', 2: 'level_sum', 3: 'nodes_eval', 4: 'nodes_dd', 5: 'level_eval_skip',
        6: 'level_split'}
JOB_KW = {}               # run options of every job (sfx_opts_t fields), set from the command line
KNAME = {'level_eval': 'k_eval_list<%s, AB=1>', 'level_eval_skip': 'k_eval_list<%s, AB=0>'}
# bounded samples of the workload: the plain GPU path (the reference's host driver is bound by its serial
# sums, ~2.5e8 points/s on the largest jobs) and the reference CPU build.  Both hold a D = 7 integral
# (D = 7 jobs are 91% of the workload's points); the plain sample also the light-by-light integral, whose
# smallest job here (3.4e9 points of ~15k FP64 operations) takes the CPU build several minutes.
PLAIN_SAMPLE = [('g3_00', 1.0), ('g3_lbl', 1.0), ('g3_01', 0.7), ('v3_03', 0.8), ('v3_p4Sc', 1.0),
                ('g2_00', 1.0), ('g2_01', 1.0), ('g2_02', 1.0), ('g2_03', 1.0), ('v2_00', 1.0)]
CPU_SAMPLE = [j for j in PLAIN_SAMPLE if j[0] != 'g3_lbl']
# the served stream beside the throughput run: the serving sweep's stream at 10 jobs/s (150 jobs, about 15 s)
SERVE_RATE, SERVE_SEED, SERVE_JOBS = 10.0, 7, 150


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return float('nan')
    k = (len(xs) - 1)*p/100.0
    f = math.floor(k)
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f])*(k - f)


def workload(S):
    """[(graph index, name, D, h, N, top)] of the 138 jobs, and their ten slices"""
    info = S.graphs()
    gs = S.select('all2', 'all3')
    jobs = [(gi, info[gi]['name'], info[gi]['D'], h, N, S.default_top(h)) for h, N in CFG for gi in gs]
    order = sorted(jobs, key=lambda j: (-j[2], j[3], j[1]))
    return jobs, [order[i::SLICES] for i in range(SLICES)]


def run_batched(S, steps, lat, out):
    """the batched path: every job of the given steps submitted to the engine in step order (widest first
    within a step), then waited for; a job's latency is submit -> result"""
    ev = 0
    ids = [(S.submit(gi, h, N, EPS, top, **JOB_KW), (name, h)) for step in steps for gi, name, D, h, N, top in step]
    for j, key in ids:
        r, tq, tr = S.wait(j)
        lat.append(tq + tr)
        ev += r['nev']
        out.setdefault(key, r)
    return ev


def run_per_graph(S, jobs, lat, out):
    """the plain per-graph path: one hierarchical sum after another on the library's stream"""
    ev = 0
    for gi, name, D, h, N, top in jobs:
        t0 = time.perf_counter()
        r = S.hsum(gi, h, N, EPS, top, **JOB_KW)
        lat.append(time.perf_counter() - t0)
        ev += r['nev']
        out[(name, h)] = r
    return ev


def run_reference_sample(sample, gpu, cores):
    """the reference driver (GPU build over this library, or its CPU build) on the sample, one run per h"""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle as O
    binary = O.SINCQED_REF_GPU if gpu else O.SINCQED_REF
    if not os.path.exists(binary):
        return None
    ev, wall, recs = 0, 0.0, {}
    for h, N in CFG:
        names = [g for g, hh in sample if hh == h]
        if not names:
            continue
        ref, w, out = O.run_reference(h, N, EPS, names, env={'SINCQED_SYM': '0'}, threads=cores, timeout=900, gpu=gpu)
        if gpu and '# gpu off' in out:
            raise RuntimeError('the reference GPU build did not reach the device')
        ev += sum(int(r['nev']) for r in ref.values())
        wall += w
        recs.update({(g, h): r for g, r in ref.items()})
    return ev, wall, recs


def fixtures():
    """the reference's committed results without FP contraction: {(graph, h): record} at eps 1e-12, the
    counts taken from the reference driver's GPU build where one is committed (refgpu_*: g3_lbl, whose CPU
    build evaluates a differently rounded double half of the kernel)"""
    from sincfeyn_amd import serve
    return {(g, h): r for (g, h, N, eps), r in serve.reference_records(
        os.path.join(ROOT, 'tests', 'golden', 'fixtures')).items() if eps == EPS}


def check_reference(res, fx):
    n = bad = 0
    worst = 0.0
    for key, r in res.items():
        f = fx.get(key)
        if f is None:
            continue
        n += 1
        same = (r['nev'] == int(f['nev']) and r['nq'] == int(f['nq']) and r['kept_top'] == int(f['kept']))
        rel = abs(r['S'][0] - float(f['S0']))/abs(float(f['S0'])) if float(f['S0']) else abs(r['S'][0])
        worst = max(worst, rel)
        bad += not (same and rel <= 1e-12)
    return {'jobs_checked': n, 'jobs_total': len(res), 'all_match': n > 0 and bad == 0, 'mismatches': bad,
            'max_rel_value_diff': worst, 'against': 'tests/golden/fixtures/nc_*_eps1e-12.txt (reference built '
            'without FP contraction, SINCQED_SYM=0; the counts of refgpu_*.txt, the reference driver\'s GPU build, '
            'where one is committed): evaluation, fallback and kept counts exact, S_h within 1e-12'}


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
    if fl is None:
        return None, ent
    avg_s = top['ms']/top['launches']*1e-3
    ach = fl/top['launches']/avg_s/1e12
    kname = KNAME[top['kind']] % top['graph'] if top['kind'] in KNAME else top['kind']
    rl = {'bound': 'fp64-valu', 'kernel': kname, 'achieved': round(ach, 3), 'peak': PEAK_FP64, 'unit': 'TFLOP/s',
          'frac': round(ach/PEAK_FP64, 4), 'traffic': None, 'rate': 'algorithmic: the source operations of the '
          'integrand (each + - * / counted once) x valid points per launch / mean launch time',
          'launches': top['launches'], 'avg_launch_ms': round(top['ms']/top['launches'], 4),
          'points_per_launch': top['points']/top['launches'], 'flops_per_launch': fl/top['launches'],
          'share_of_device_time': round(top['ms']/tot_ms, 3)}
    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes (each counter scaled
    # by what it reads for the library's copy kernel of known size); the executed FP64 work per launch from
    # the committed SQ_INSTS_VALU_*_F64 counters: 64 lanes x (ADD + MUL + TRANS + 2 FMA)
    for p in glob.glob(os.path.join(ROOT, 'profiles', 'traffic_*.json')):
        t = json.load(open(p))
        if t.get('kernel') == kname:
            rl['traffic'] = t['traffic_bytes_per_dispatch']
            rl['traffic_source'] = os.path.relpath(p, ROOT)
    for p in glob.glob(os.path.join(ROOT, 'profiles', 'valu_*.json')):
        t = json.load(open(p))
        if t.get('kernel') == kname:
            ex = t['executed_fp64_flops_per_dispatch']
            c = t['per_dispatch']
            n64 = sum(c['SQ_INSTS_VALU_%s_F64' % k] for k in ('ADD', 'MUL', 'FMA', 'TRANS'))
            # with FP contraction off nearly every FP64 instruction is an ADD or a MUL: one operation per
            # lane per issue, so the issue limit of this instruction mix is half the FMA peak, times 1 + FMA share
            lim = PEAK_FP64/2*(1 + c['SQ_INSTS_VALU_FMA_F64']/n64)
            rl['executed'] = {'achieved': round(ex/avg_s/1e12, 3), 'frac': round(ex/avg_s/1e12/PEAK_FP64, 4),
                              'flops_per_launch': ex, 'fp64_issue_busy': t.get('fp64_issue_busy'),
                              'issue_limit_of_mix': round(lim, 2), 'frac_of_mix_limit': round(ex/avg_s/1e12/lim, 4),
                              'source': os.path.relpath(p, ROOT)}
    agg = [e for e in ent if e['kind'] in ('level_eval', 'level_eval_skip')]
    if agg:
        f_all = sum(flops(e) for e in agg)
        ms_all = sum(e['ms'] for e in agg)
        rl['all_level_eval'] = {'achieved': round(f_all/(ms_all*1e-3)/1e12, 3), 'ms': round(ms_all, 1),
                                'share_of_device_time': round(ms_all/tot_ms, 3)}
    return rl, ent


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-plain', action='store_true', help='skip the plain GPU path sample')
    ap.add_argument('--no-serve', action='store_true', help='skip the served stream')
    ap.add_argument('--workers', type=int, default=16, help='engine workers of the batched path')
    ap.add_argument('--profile-only', action='store_true',
                    help='only the profiled per-graph pass (one host thread: the form profiled with rocprofv3)')
    ap.add_argument('--graphs', default='all3', help='graphs of the profiled pass')
    ap.add_argument('--profile-h', default='0.7', help='h values (comma-separated) of the profiled pass')
    ap.add_argument('--profile-out', default=None, help='write the per-kernel table of the profiled pass as JSON here')
    ap.add_argument('--copy-probe', type=int, default=0,
                    help='first run the library\'s plain copy kernel over this many bytes (counter calibration)')
    ap.add_argument('--batched-only', action='store_true',
                    help='only the warmup and timed batched steps (for a kernel trace of the engine\'s timeline)')
    ap.add_argument('--grow-pipeline', type=int, default=None,
                    help='the grow\'s frontier loop of every job (sfx_opts_t.grow_pipeline; default: the library\'s)')
    ap.add_argument('--graph-capture', type=int, default=0,
                    help='1: every refinement level\'s block chain runs as one captured HIP graph (sfx_opts_t.graph_capture)')
    a = ap.parse_args()
    JOB_KW['graph_capture'] = a.graph_capture
    if a.grow_pipeline is not None:
        JOB_KW['grow_pipeline'] = a.grow_pipeline
    if a.gpus != 1:
        print('bench.py: this build measures one GPU; the Sinc sums are not split across GPUs here '
              '(--gpus %d is not supported)' % a.gpus, file=sys.stderr)
        sys.exit(2)
    from sincfeyn_amd import sfx as S
    S.hw_queues(a.workers)
    info = S.init(0)
    opc = json.load(open(os.path.join(ROOT, 'sincfeyn_amd', '_gen', 'opcounts.json')))
    if a.copy_probe:
        ms = S.C.c_double()
        S._check(S.lib().sfx_copy_probe(S.C.c_long(a.copy_probe), S.C.byref(ms)))
        print('copy probe: %d bytes in %.3f ms' % (a.copy_probe, ms.value), file=sys.stderr)

    def log(*x):
        print('[bench]', *x, file=sys.stderr, flush=True)

    if a.batched_only:
        jobs, slices = workload(S)
        steps = [slices[i % SLICES] for i in range(a.warmup + a.steps)]
        S.engine_start(a.workers, 0)
        run_batched(S, steps[:a.warmup], [], {})
        t0 = time.perf_counter()
        ev = run_batched(S, steps[a.warmup:], [], {})
        dt = time.perf_counter() - t0
        log('batched: %d steps, %d points in %.2f s' % (a.steps, ev, dt))
        print(json.dumps({'metric': 'Sinc-sum points/s, batched steps only', 'value': ev/dt, 'unit': 'points/s',
                          'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup, 'wall_s': dt, 'path': 'batched',
                          'workers': a.workers, 'device': info}))
        return

    # the profiled per-graph pass (kernel durations from HIP events on the library's stream)
    pgraphs = [(gi, S.graphs()[gi]['name'], S.graphs()[gi]['D'], h, N, S.default_top(h))
               for h, N in CFG if repr(h) in a.profile_h.split(',') for gi in S.select(*a.graphs.split(','))]
    S.hsum(pgraphs[0][0], pgraphs[0][3], pgraphs[0][4], EPS, pgraphs[0][5])     # first launches of the kernels
    S._check(S.lib().sfx_prof_enable(1))
    t0 = time.perf_counter()
    ev_p = run_per_graph(S, pgraphs, [], {})
    S._check(S.lib().sfx_sync())
    dt_p = time.perf_counter() - t0
    S._check(S.lib().sfx_prof_enable(0))
    rl, ent = roofline(S, opc)
    if a.profile_out:
        with open(a.profile_out, 'w') as f:
            json.dump(ent, f, indent=1)
    log('profiled per-graph pass: %d points in %.2f s' % (ev_p, dt_p))
    if a.profile_only:
        print(json.dumps({'metric': 'Sinc-sum points/s, profiled per-graph pass only', 'value': ev_p/dt_p,
                          'unit': 'points/s', 'n_gpus': 1, 'path': 'per-graph', 'roofline': rl, 'device': info}))
        return

    jobs, slices = workload(S)
    steps = [slices[i % SLICES] for i in range(a.warmup + a.steps)]
    S.engine_start(a.workers, 0)
    run_batched(S, steps[:a.warmup], [], {})
    lat, batched = [], {}
    t0 = time.perf_counter()
    ev = run_batched(S, steps[a.warmup:], lat, batched)
    dt = time.perf_counter() - t0
    log('batched: %d steps, %d points in %.2f s' % (a.steps, ev, dt))
    # the plain per-graph path over the same 138 jobs, without profiling
    lat_pg, per_graph = [], {}
    t0 = time.perf_counter()
    ev_pg = run_per_graph(S, jobs, lat_pg, per_graph)
    dt_pg = time.perf_counter() - t0
    log('per-graph: %d points in %.2f s' % (ev_pg, dt_pg))
    # where a job's time goes on the per-graph path: the coarsest level (grow, tiers), the dilation, the
    # refinement levels (each job's own timers, summed over the jobs)
    stage_s = {k: round(sum(r[f] for r in per_graph.values()), 3)
               for k, f in (('coarsest_level', 't_grow'), ('dilation', 't_dilate'), ('refinement', 't_refine'))}
    def seq(a_, b_):        # S_h, S_2h, S_4h bit for bit (NaN where a lattice has no such level)
        return all(x == y or (math.isnan(x) and math.isnan(y)) for x, y in zip(a_, b_))
    same = all(seq(r['S'], per_graph[k]['S']) and r['nev'] == per_graph[k]['nev'] and r['nq'] == per_graph[k]['nq']
               for k, r in batched.items())
    ref = check_reference(batched, fixtures())
    # served latency: the serving sweep's seeded stream at SERVE_RATE jobs/s through the same engine, every job
    # checked against the reference's result record (sincfeyn_amd/serve.py)
    served = None
    if not a.no_serve:
        from sincfeyn_amd import serve as SV
        names = [g['name'] for g in S.graphs()]
        refs = SV.reference_records(os.path.join(ROOT, 'tests', 'golden', 'fixtures'))
        rep = SV.sweep([SERVE_RATE], SERVE_SEED, S.select('all2', 'all3'), names, refs, lambda r: SERVE_JOBS)[0]
        served = {k: rep[k] for k in ('rate', 'seed', 'jobs', 'wall_s', 'points_per_s', 'latency_p50_s', 'latency_p99_s',
                                      'queue_p50_s', 'queue_p99_s', 'reference_check')}
        served['stream'] = ('Poisson arrivals at %g jobs/s (seed %d): graphs uniform over all2 + all3, h in {1.0, 0.9, 0.8, '
                            '0.7}, eps in {1e-10, 1e-12}; latency = arrival -> result' % (SERVE_RATE, SERVE_SEED))
        log('served stream: %d jobs at %g/s, p50 %.3f s, p99 %.3f s' % (rep['jobs'], SERVE_RATE, rep['latency_p50_s'],
                                                                        rep['latency_p99_s']))

    plain = None
    cores = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    if not a.no_plain:
        got = run_reference_sample(PLAIN_SAMPLE, True, cores)
        if got:
            ev_s, w_s, recs = got
            sjobs = [j for j in jobs if (j[1], j[3]) in set(PLAIN_SAMPLE)]
            t0 = time.perf_counter()
            ev_sb = run_batched(S, [sjobs], [], {})
            dt_sb = time.perf_counter() - t0
            t0 = time.perf_counter()
            samp_pg = {}
            ev_sp = run_per_graph(S, sjobs, [], samp_pg)
            dt_sp = time.perf_counter() - t0
            agree = all(r['nev'] == samp_pg[k]['nev'] and r['nq'] == samp_pg[k]['nq'] and r['kept'] == samp_pg[k]['kept_top']
                        and abs(r['S0'] - samp_pg[k]['S'][0]) <= 1e-12*abs(r['S0']) for k, r in recs.items())
            plain = {'value': ev_s/w_s, 'unit': 'points/s', 'host_threads': cores, 'wall_s': w_s, 'points': ev_s,
                     'sample': ' '.join('%s@h%g' % j for j in PLAIN_SAMPLE) + ' (%.1f%% of the workload\'s points)' % (
                         100.0*ev_s/ev*a.steps/SLICES if ev else 0),
                     'path': 'oracle/_ref/sincqed_ref_gpu: the reference driver (-DSQ_GPU) over src/sfx_gpu.c',
                     'batched_on_sample': ev_sb/dt_sb, 'per_graph_on_sample': ev_sp/dt_sp,
                     'batched_over_plain': (ev_sb/dt_sb)/(ev_s/w_s),
                     'results_agree_with_device': agree}
            log('plain GPU sample: %d points in %.2f s' % (ev_s, w_s))
    cpu = None
    if not a.no_cpu_baseline:
        got = run_reference_sample(CPU_SAMPLE, False, cores)
        if got:
            ev_c, w_c, _ = got
            cpu = {'value': ev_c/w_c, 'unit': 'points/s', 'cores': cores, 'kind': 'reference',
                   'sample': '%s at eps 1e-12: %d evaluations in %.1f s (%.2f%% of the workload\'s points; one D = 7 '
                             'integral; g3_lbl left out: its smallest job here takes minutes on the CPU)' % (
                                 ' '.join('%s@h%g' % j for j in CPU_SAMPLE), ev_c, w_c, 100.0*ev_c/ev*a.steps/SLICES if ev else 0)}
            log('cpu sample: %d points in %.2f s' % (ev_c, w_c))
            # beside the timed sample: the reference's own time for every job of the workload, as recorded in
            # the committed fixtures (one process per graph on one host thread, g3_lbl on four, in the machine
            # that wrote them -- not this one): the whole workload's CPU cost, g3_lbl included
            fx = fixtures()
            have = [(name, fx[(name, h)]) for gi, name, D, h, N, top in jobs if 't' in fx.get((name, h), {})]
            if have:
                ts = sum(float(r['t'])*(4 if name == 'g3_lbl' else 1) for name, r in have)
                pts = sum(int(r['nev']) for name, r in have)
                cpu['workload_from_fixtures'] = {
                    'jobs': len(have), 'of': len(jobs), 'points': pts, 'thread_seconds': ts,
                    'points_per_thread_second': pts/ts if ts else None,
                    'note': 'the reference build without FP contraction, SINCQED_SYM=0, one process per job on one '
                            'host thread (g3_lbl on four) in the machine that wrote tests/golden/fixtures '
                            '(make_fixtures.py), several jobs at once beside other work: the thread-seconds are '
                            'an upper bound of the CPU cost'}
    out = {
        'metric': 'Sinc-sum points/s batched vs per-graph; p50/p99 job latency at <=1e-12 rel',
        'value': ev/dt, 'unit': 'points/s', 'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup,
        'ms_per_step': dt*1e3/a.steps, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64', 'data': 'deterministic Sinc lattice of the named integrals (no input data)',
        'config': {'workload': 'batched heterogeneous launch: all2 + all3 (46 integrals) x h in {1.0, 0.8, 0.7} '
                               '(N 97/121/139, top 1/1/2), eps 1e-12: 138 jobs through the job engine; a step is one '
                               'of ten fixed slices of the jobs (widest first, dealt round-robin)',
                   'jobs': len(jobs), 'jobs_per_step': len(slices[0]), 'eps': EPS, 'points_per_step': ev//a.steps,
                   'path': 'batched', 'workers': a.workers, 'precision': 'FP64 bulk, double-double where flagged',
                   'graph_capture': a.graph_capture},
        'latency_s': {'p50': percentile(lat, 50), 'p99': percentile(lat, 99), 'jobs': len(lat),
                      'note': 'submit -> result with every timed step queued at once (a throughput run; the served '
                              'latencies are in "served")'},
        'served': served,
        'per_graph': {'value': ev_pg/dt_pg, 'wall_s': dt_pg, 'jobs': len(jobs), 'stage_s': stage_s,
                      'latency_s': {'p50': percentile(lat_pg, 50), 'p99': percentile(lat_pg, 99)}},
        'batched_over_per_graph': (ev/dt)/(ev_pg/dt_pg),
        'batched_results_identical_to_per_graph': same,
        'reference_check': ref,
        'plain_gpu': plain,
        'batched_over_plain_gpu': (ev/dt)/plain['value'] if plain else None,
        'roofline': rl,
        'roofline_pass': {'jobs': '%d jobs (%s at h %s), eps 1e-12, per graph' % (len(pgraphs), a.graphs, a.profile_h),
                          'points': ev_p, 'wall_s': dt_p},
        'cpu_baseline': cpu,
        'device': info,
    }
    print(json.dumps(out))


if __name__ == '__main__':
    main()
