"""profsum.py -- reduce rocprofv3 CSV output to the small per-kernel tables kept under profiles/.

    python -m sincfeyn_amd.profsum counters OUT.json CSV... [--match REGEX]
        mean, per dispatch, of every counter in the counter-collection CSVs (rocprofv3 --pmc ...
        --output-format csv), per kernel name matching REGEX, with the dispatch count and grid size
    python -m sincfeyn_amd.profsum stats OUT.md KERNEL_STATS_CSV [--top N] [--parts]
        the first N rows of a kernel-stats CSV (rocprofv3 --kernel-trace --stats) as a markdown table; with
        --parts, short kernel names and first a table of the device time by part of the hierarchical sum
    python -m sincfeyn_amd.profsum valu OUT.json COUNTER_CSV... --profile KERNELS_JSON [--kernel NAME]
        the FP64 and VALU issue figures of every level kernel from two counter passes, and of NAME (default
        k_eval_list<g3_lbl, 0>) also the executed FP64 operations and instructions per point: the table bench.py
        reads for its roofline's executed rate
    python -m sincfeyn_amd.profsum traffic OUT.json FETCH_CSV WRITE_CSV --profile KERNELS_JSON [--kernel NAME]
            [--copy-bytes N] [--bytes-per-point B]
        HBM bytes per dispatch of NAME from a FETCH_SIZE and a WRITE_SIZE pass, each counter scaled by what it
        reads for the library's copy kernel of N bytes (bench.py --copy-probe N), beside the algorithmic bytes
    python -m sincfeyn_amd.profsum timeline OUT.md KERNEL_TRACE_CSV [--last SECONDS] [--bins N]
        the device's timeline from a kernel trace (rocprofv3 --kernel-trace): over the last SECONDS of the
        trace (default all of it), the time some kernel runs, the time an integrand evaluation runs, the
        idle gaps, the number of kernels running at once, the occupancy per time bin, and the graphs whose
        kernels run in the last tenth

The committed tables are what bench.py reads: the executed FP64 operations per dispatch of its roofline
kernel (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64) and its HBM bytes per dispatch (FETCH_SIZE, WRITE_SIZE),
each scaled by what the counter reads for the library's copy kernel of known size.
"""
import argparse
import collections
import csv
import json
import re
import sys


def counters(out, files, match):
    rx = re.compile(match)
    acc = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    grid = collections.defaultdict(float)
    for fn in files:
        with open(fn) as f:
            for row in csv.DictReader(f):
                name = row.get('Kernel_Name', '')
                if not rx.search(name):
                    continue
                did = row.get('Dispatch_Id') or row.get('Correlation_Id')
                acc[name][row['Counter_Name']] += float(row['Counter_Value'])
                if did not in disp[name]:
                    disp[name].add(did)
                    grid[name] += float(row.get('Grid_Size', 0) or 0)
    res = {}
    for name, c in acc.items():
        n = len(disp[name])
        res[name] = {'dispatches': n, 'mean_grid_size': grid[name]/n if n else 0,
                     'per_dispatch': {k: v/n for k, v in sorted(c.items())}}
    with open(out, 'w') as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print('%d kernels -> %s' % (len(res), out))


def short_name(name):
    """a kernel's name without its parameter list, namespaces and graph-type decoration:
    k_eval_list<g3_lbl, 0>, k_dd<g3_21, true>, k_grow_cand, rocprim radix sort"""
    if 'rocprim' in name:
        return 'rocprim radix sort' if 'radix_sort' in name else 'rocprim ' + name.split('(')[0][-40:]
    n = name.split('(')[0] if not name.startswith('void ') else name[5:].split('(')[0]
    return n.replace('sfx::', '').replace('sfxk_', '').replace('::Graph', '').strip()


# the parts of a hierarchical sum, by kernel name (first match wins)
PARTS = [('refinement evaluation (k_eval_list)', re.compile(r'^k_eval_list<')),
         ('refinement sums, class split, fallback', re.compile(r'^(k_level_|k_cell_split|k_zero_acc|k_dd<.*, true>)')),
         ('coarsest level: frontier evaluation and fallback', re.compile(r'^(k_eval<|k_dd<.*, false>)')),
         ('coarsest level: grow, tiers, dilation', re.compile(r'^(k_grow_|k_front_|k_gmap_|k_tier_|k_kmap_|k_dil_|k_drop_|'
                                                            r'k_gather_keep|k_scan_u32|k_count_flags|k_scatter_flags|rocprim)')),
         ('other (copies, fills, counters)', re.compile(r'.'))]


def stats(out, fn, top, parts=False):
    with open(fn) as f:
        rows = list(csv.DictReader(f))
    lines = []
    if parts:
        tot = sum(float(r['TotalDurationNs']) for r in rows)
        by = collections.OrderedDict((p, 0.0) for p, _ in PARTS)
        for r in rows:
            n = short_name(r['Name'])
            by[next(p for p, rx in PARTS if rx.search(n))] += float(r['TotalDurationNs'])
        lines += ['Device time of all kernels: %.0f ms.' % (tot/1e6), '', '| part | ms | share |', '|---|---|---|']
        lines += ['| %s | %.0f | %.1f%% |' % (p, v/1e6, 100*v/tot) for p, v in by.items()]
        lines += ['', '| kernel | calls | total ms | avg ms | % of device time |', '|---|---|---|---|---|']
        for r in rows[:top]:
            lines.append('| `%s` | %s | %.1f | %.4f | %.2f |' % (short_name(r['Name']), r['Calls'], float(r['TotalDurationNs'])/1e6,
                                                            float(r['AverageNs'])/1e6, float(r['Percentage'])))
    else:
        keys = [k for k in ('Name', 'Calls', 'TotalDurationNs', 'AverageNs', 'Percentage', 'MinNs', 'MaxNs') if rows and k in rows[0]]
        lines += ['| ' + ' | '.join(keys) + ' |', '|' + '---|'*len(keys)]
        for r in rows[:top]:
            lines.append('| ' + ' | '.join(r[k].replace('|', '/') for k in keys) + ' |')
    with open(out, 'w') as f:
        f.write('\n'.join(lines) + '\n')
    print('%d rows -> %s' % (min(top, len(rows)), out))


def valu(out, files, profile, kernel):
    """the FP64 issue figures of the level kernels from the two counter passes (SQ_INSTS_VALU_{ADD,MUL,FMA,
    TRANS}_F64; SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU, SQ_BUSY_CYCLES, GRBM_GUI_ACTIVE), per dispatch: a wave64
    VALU instruction occupies a 16-lane SIMD for 4 cycles, GRBM_GUI_ACTIVE is summed over the 8 XCDs, the
    device has 1024 SIMDs; with the bench's per-kernel table (bench.py --profile-out) the valid points per
    dispatch of `kernel` (a short name, k_eval_list<g3_lbl, 0>) and its instructions per point"""
    acc = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(lambda: collections.defaultdict(set))
    for fn in files:
        with open(fn) as f:
            for row in csv.DictReader(f):
                n = short_name(row.get('Kernel_Name', ''))
                if not n.startswith('k_eval_list<'):
                    continue
                c = row['Counter_Name']
                acc[n][c] += float(row['Counter_Value'])
                disp[n][c].add(row.get('Dispatch_Id') or row.get('Correlation_Id'))
    per = {n: {c: v/len(disp[n][c]) for c, v in cs.items()} for n, cs in acc.items()}

    def figures(c):
        fp = sum(c.get('SQ_INSTS_VALU_%s_F64' % k, 0.0) for k in ('ADD', 'MUL', 'FMA', 'TRANS'))
        cyc = 1024*c['GRBM_GUI_ACTIVE']/8
        return {'fp64_share_of_valu': round(fp/c['SQ_INSTS_VALU'], 4), 'valu_issue_busy': round(4*c['SQ_INSTS_VALU']/cyc, 4),
                'fp64_issue_busy': round(4*fp/cyc, 4), 'fma_share_of_fp64': round(c.get('SQ_INSTS_VALU_FMA_F64', 0.0)/fp, 4)}
    need = ('SQ_INSTS_VALU', 'GRBM_GUI_ACTIVE', 'SQ_INSTS_VALU_ADD_F64', 'SQ_INSTS_VALU_MUL_F64')
    table = {n: dict(figures(c), dispatches=len(disp[n]['SQ_INSTS_VALU'])) for n, c in sorted(per.items())
             if all(k in c for k in need) and c['GRBM_GUI_ACTIVE'] > 0 and
             sum(c.get('SQ_INSTS_VALU_%s_F64' % k, 0.0) for k in ('ADD', 'MUL', 'FMA', 'TRANS')) > 0}
    c = per[kernel]
    fp = sum(c['SQ_INSTS_VALU_%s_F64' % k] for k in ('ADD', 'MUL', 'FMA', 'TRANS'))
    g, ab = re.match(r'k_eval_list<(\w+), (\d)>', kernel).groups()
    ent = [e for e in json.load(open(profile)) if e['graph'] == g and e['kind'] == ('level_eval_skip' if ab == '0' else 'level_eval')]
    pts = sum(e['points'] for e in ent)/sum(e['launches'] for e in ent)
    res = {'kernel': 'k_eval_list<%s, AB=%s>' % (g, ab), 'dispatches': table[kernel]['dispatches'],
           'valid_points_per_dispatch': pts, 'per_dispatch': c,
           'executed_fp64_flops_per_dispatch': 64*(c['SQ_INSTS_VALU_ADD_F64'] + c['SQ_INSTS_VALU_MUL_F64'] +
                                                   c.get('SQ_INSTS_VALU_TRANS_F64', 0.0) + 2*c.get('SQ_INSTS_VALU_FMA_F64', 0.0)),
           'fp64_instructions_per_point': round(64*fp/pts, 1), 'valu_instructions_per_point': round(64*c['SQ_INSTS_VALU']/pts, 1)}
    res.update(figures(c))
    res['note'] = ('executed FP64 operations = 64 lanes x (ADD + MUL + TRANS + 2 FMA) per wave instruction; issue busy = '
                   'wave instructions x 4 cycles (a wave64 VALU instruction on a 16-lane SIMD) / (1024 SIMDs x '
                   'GRBM_GUI_ACTIVE/8 cycles, the counter summed over 8 XCDs)')
    res['all_level_kernels'] = table
    with open(out, 'w') as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print('%d level kernels -> %s; %s: FP64 issue %.3f, VALU issue %.3f' % (len(table), out, kernel, res['fp64_issue_busy'],
                                                                              res['valu_issue_busy']))


def traffic(out, files, profile, kernel, copy_bytes, bytes_per_point):
    """HBM bytes per dispatch of `kernel` from a FETCH_SIZE and a WRITE_SIZE pass (KiB per dispatch), each
    scaled by what the counter reads for the library's copy kernel of copy_bytes (bench.py --copy-probe); beside
    them the algorithmic bytes, bytes_per_point x the valid points per dispatch of the bench's per-kernel table"""
    acc = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(lambda: collections.defaultdict(set))
    for fn in files:
        with open(fn) as f:
            for row in csv.DictReader(f):
                n = short_name(row.get('Kernel_Name', ''))
                c = row['Counter_Name']
                if c not in ('FETCH_SIZE', 'WRITE_SIZE') or not (n == kernel or 'copy_probe' in n):
                    continue
                n = 'copy' if 'copy_probe' in n else n
                acc[n][c] += float(row['Counter_Value'])
                disp[n][c].add(row.get('Dispatch_Id') or row.get('Correlation_Id'))
    m = {n: {c: v/len(disp[n][c]) for c, v in cs.items()} for n, cs in acc.items()}
    ff = copy_bytes/1024/m['copy']['FETCH_SIZE']; wf = copy_bytes/1024/m['copy']['WRITE_SIZE']
    g, ab = re.match(r'k_eval_list<(\w+), (\d)>', kernel).groups()
    ent = [e for e in json.load(open(profile)) if e['graph'] == g and e['kind'] == ('level_eval_skip' if ab == '0' else 'level_eval')]
    pts = sum(e['points'] for e in ent)/sum(e['launches'] for e in ent)
    k = m[kernel]
    res = {'kernel': 'k_eval_list<%s, AB=%s>' % (g, ab), 'dispatches': len(disp[kernel]['FETCH_SIZE']),
           'copy_probe': {'bytes_read': copy_bytes, 'bytes_written': copy_bytes, 'FETCH_SIZE_kib': m['copy']['FETCH_SIZE'],
                          'WRITE_SIZE_kib': m['copy']['WRITE_SIZE'], 'fetch_factor': ff, 'write_factor': wf},
           'FETCH_SIZE_kib_per_dispatch': k['FETCH_SIZE'], 'WRITE_SIZE_kib_per_dispatch': k['WRITE_SIZE'],
           'traffic_bytes_per_dispatch': (k['FETCH_SIZE']*ff + k['WRITE_SIZE']*wf)*1024,
           'algorithmic_bytes_per_dispatch': bytes_per_point*pts, 'valid_points_per_dispatch': pts}
    with open(out, 'w') as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print('%s: %.4g bytes per dispatch (algorithmic %.4g), fetch factor %.3f, write factor %.3f -> %s' % (
        kernel, res['traffic_bytes_per_dispatch'], res['algorithmic_bytes_per_dispatch'], ff, wf, out))


EVAL = re.compile(r'^(k_eval_list<|k_eval<|k_dd<|k_qd<)')


def timeline(out, fn, last=None, bins=20):
    ks = []
    with open(fn) as f:
        for r in csv.DictReader(f):
            ks.append((int(r['Start_Timestamp']), int(r['End_Timestamp']), short_name(r['Kernel_Name'])))
    if not ks:
        raise SystemExit('no dispatches in ' + fn)
    t_end = max(e for _, e, _ in ks)
    t0 = t_end - int(last*1e9) if last else min(s for s, _, _ in ks)
    ks = [(max(s, t0), e, n) for s, e, n in ks if e > t0]
    span = t_end - t0
    # sweep over start and end points: the number of kernels running, and of evaluation kernels among them
    pidx = {}

    def part(n):
        if n not in pidx:
            pidx[n] = next(i for i, (p, rx) in enumerate(PARTS) if rx.search(n))
        return pidx[n]
    ev = sorted([(s, 1, EVAL.search(n) is not None, part(n)) for s, e, n in ks] +
                [(e, -1, EVAL.search(n) is not None, part(n)) for s, e, n in ks])
    run = run_ev = 0
    prun = [0]*len(PARTS)
    other = [0.0]*len(PARTS)         # time with no evaluation running, shared among the parts running
    busy = ev_busy = 0
    conc = collections.Counter()
    gaps = []
    prev = t0
    nb = max(1, bins)
    bb = [0]*nb; be = [0]*nb

    def add(a, b, n, n_ev):
        nonlocal busy, ev_busy
        if b <= a:
            return
        conc[min(n, 4)] += b - a
        if n and not n_ev:
            for i, c in enumerate(prun):
                if c:
                    other[i] += (b - a)*c/n
        if n:
            busy += b - a
        if n_ev:
            ev_busy += b - a
        # per-bin shares
        i = min(nb - 1, (a - t0)*nb//span)
        while a < b:
            hi = t0 + (i + 1)*span//nb
            c = min(b, hi)
            if n:
                bb[i] += c - a
            if n_ev:
                be[i] += c - a
            a = c
            i = min(nb - 1, i + 1)
    for t, d, is_ev, pi in ev:
        if run == 0 and t > prev:
            gaps.append(t - prev)
        add(prev, t, run, run_ev)
        run += d
        run_ev += d if is_ev else 0
        prun[pi] += d
        prev = t
    # the graphs of the kernels in the last tenth of the window (device time by graph)
    tail0 = t_end - span//10
    tail = collections.Counter()
    for s, e, n in ks:
        if e > tail0:
            m = re.search(r'<(\w+)', n)
            tail[m.group(1) if m and not n.startswith('rocprim') else '(no graph)'] += e - max(s, tail0)
    ms = lambda x: x/1e6
    L = ['Window: %.1f ms (%s).' % (ms(span), 'the last %.3f s of the trace' % last if last else 'the whole trace'),
         '', '| | ms | share of the window |', '|---|---|---|',
         '| some kernel running | %.1f | %.1f%% |' % (ms(busy), 100*busy/span),
         '| an integrand evaluation running (k_eval_list, k_eval, k_dd, k_qd) | %.1f | %.1f%% |' % (ms(ev_busy), 100*ev_busy/span),
         '| only other kernels running | %.1f | %.1f%% |' % (ms(busy - ev_busy), 100*(busy - ev_busy)/span),
         '| no kernel running | %.1f | %.1f%% |' % (ms(span - busy), 100*(span - busy)/span),
         '', 'With no evaluation running, the kernels that run (time shared among those running at once): ' +
         ', '.join('%s %.1f ms' % (PARTS[i][0], ms(v)) for i, v in sorted(enumerate(other), key=lambda x: -x[1]) if v) + '.',
         '', 'Kernels running at once (share of the window): ' +
         ', '.join('%s: %.1f%%' % ('4+' if k == 4 else k, 100*v/span) for k, v in sorted(conc.items())) + '.',
         '', 'Idle gaps: %d, %.1f ms in all; the longest %s ms; gaps over 1 ms: %d (%.1f ms).' % (
             len(gaps), ms(sum(gaps)), ', '.join('%.2f' % ms(g) for g in sorted(gaps)[::-1][:5]),
             sum(g > 1e6 for g in gaps), ms(sum(g for g in gaps if g > 1e6))),
         '', '| bin | from ms | busy | evaluation running |', '|---|---|---|---|']
    for i in range(nb):
        w = (t0 + (i + 1)*span//nb) - (t0 + i*span//nb)
        L.append('| %d | %.0f | %.0f%% | %.0f%% |' % (i, ms(i*span//nb), 100*bb[i]/w, 100*be[i]/w))
    L += ['', 'Device time by graph in the last tenth of the window: ' +
          ', '.join('%s %.1f ms' % (g, ms(v)) for g, v in tail.most_common(8)) + '.']
    with open(out, 'w') as f:
        f.write('\n'.join(L) + '\n')
    print('%d dispatches, %.1f ms window -> %s' % (len(ks), ms(span), out))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest='cmd', required=True)
    c = sub.add_parser('counters'); c.add_argument('out'); c.add_argument('files', nargs='+'); c.add_argument('--match', default='.')
    s = sub.add_parser('stats'); s.add_argument('out'); s.add_argument('file'); s.add_argument('--top', type=int, default=40)
    s.add_argument('--parts', action='store_true')
    v = sub.add_parser('valu'); v.add_argument('out'); v.add_argument('files', nargs='+'); v.add_argument('--profile', required=True)
    v.add_argument('--kernel', default='k_eval_list<g3_lbl, 0>')
    r = sub.add_parser('traffic'); r.add_argument('out'); r.add_argument('files', nargs='+'); r.add_argument('--profile', required=True)
    r.add_argument('--kernel', default='k_eval_list<g3_lbl, 0>'); r.add_argument('--copy-bytes', type=int, default=1 << 30)
    r.add_argument('--bytes-per-point', type=float, default=9 + 20/128)
    t = sub.add_parser('timeline'); t.add_argument('out'); t.add_argument('file'); t.add_argument('--last', type=float, default=None)
    t.add_argument('--bins', type=int, default=20)
    a = ap.parse_args(argv)
    if a.cmd == 'counters':
        counters(a.out, a.files, a.match)
    elif a.cmd == 'traffic':
        traffic(a.out, a.files, a.profile, a.kernel, a.copy_bytes, a.bytes_per_point)
    elif a.cmd == 'valu':
        valu(a.out, a.files, a.profile, a.kernel)
    elif a.cmd == 'timeline':
        timeline(a.out, a.file, a.last, a.bins)
    else:
        stats(a.out, a.file, a.top, a.parts)


if __name__ == '__main__':
    sys.exit(main())
