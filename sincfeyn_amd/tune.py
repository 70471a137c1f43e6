"""tune.py -- the per-graph build choices of the refinement kernels, by measurement.

The level kernels (k_eval_list<G, AB>) of a graph are straight-line FP64 code whose speed depends on how
many waves share a SIMD (set by their register count, 512/VGPRs at most) and on the instruction
scheduler, and for the D >= 6 cell lists on how many 64-child slices a wave evaluates (points per thread:
more amortise a wave's start, but the loop can raise the register count); the best choice differs between
graphs.  Each candidate in build.VARIANTS (the compiler's own register budget, or at least 3 or 4 waves per
SIMD; the default or the bottom-up register-reducing scheduler; one or four slices per wave) is built as a
trial library of the graphs, each trial is timed on the device, and the fastest
variant of each graph goes into sincfeyn_amd/tune_table.json, which build.py applies.  Every variant is the
same instructions on the same operands: the trials' results are checked to be bit-identical.

    python -m sincfeyn_amd.tune build [--graphs g3] [--variants base,w3,...]   trial libraries (CPU)
    python -m sincfeyn_amd.tune time  [--h 1.0] --out FILE                      time them (GPU)
    python -m sincfeyn_amd.tune choose FILE                                     write tune_table.json
"""
import argparse
import json
import os
import subprocess
import sys

from . import build as B

TRIALS = os.path.join(B.PKG, '_tune')


def graph_names(sel):
    import glob
    names = sorted(os.path.basename(p)[2:-4] for p in glob.glob(os.path.join(B.GEN, 'k_*.hip')))
    out = []
    for s in sel.split(','):
        out += [n for n in names if n == s or (n.startswith(s + '_') and s in ('g2', 'g3', 'v3'))]
    return [n for n in out if n != 'g3_lbl']      # g3_lbl's unit has its own fixed choice (node split)


def cmd_build(a):
    for v in a.variants.split(','):
        d = os.path.join(TRIALS, v)
        B.build(force=False, only=set(graph_names(a.graphs)), libdir=d, objdir=os.path.join(d, 'obj'), variant=v)


TIMER = r'''
import json, sys
from sincfeyn_amd import sfx as S
S.init(0)
h = float(sys.argv[1]); N = 2*int(-(-48//h)) + 1
out = {}
for g in S.graphs():
    name = g['name']
    top = S.default_top(h)
    S.hsum(name, h, N, 1e-12, top)                      # first launches
    S._check(S.lib().sfx_prof_enable(1))
    r = S.hsum(name, h, N, 1e-12, top)
    S._check(S.lib().sfx_sync())
    S._check(S.lib().sfx_prof_enable(0))
    L = S.lib(); ms = {}
    for i in range(L.sfx_prof_count()):
        e = S.ProfEntry(); S._check(L.sfx_prof_get(i, S.C.byref(e)))
        if e.kind in (0, 5):
            ms['AB1' if e.kind == 0 else 'AB0'] = (e.ms, e.points, e.launches)
    out[name] = {'ms': ms, 'S': r['S'][:2], 'nev': r['nev'], 'nq': r['nq']}
    print(name, ms, flush=True, file=sys.stderr)
print(json.dumps(out))
'''


def cmd_time(a):
    res = {'h': a.h, 'variants': {}}
    for v in sorted(os.listdir(TRIALS)):
        lib = os.path.join(TRIALS, v, 'libsfx.so')
        if not os.path.exists(lib):
            continue
        print('== variant %s' % v, file=sys.stderr, flush=True)
        p = subprocess.run([sys.executable, '-c', TIMER, repr(a.h)], env=dict(os.environ, SFX_LIB=lib),
                           capture_output=True, text=True, timeout=900)
        sys.stderr.write(p.stderr[-3000:])
        if p.returncode:
            raise RuntimeError('variant %s failed' % v)
        res['variants'][v] = json.loads(p.stdout.splitlines()[-1])
    with open(a.out, 'w') as f:
        json.dump(res, f, indent=1)


def cmd_choose(a):
    t = json.load(open(a.file))
    V = t['variants']
    base = V['base']
    table = {}
    report = {}
    for g in base:
        tot = {}
        for v, r in V.items():
            if g not in r:
                continue
            assert r[g]['S'] == base[g]['S'] and r[g]['nev'] == base[g]['nev'], (g, v, 'results differ between builds')
            tot[v] = sum(x[0] for x in r[g]['ms'].values())
        best = min(tot, key=tot.get)
        # keep the compiler's choice unless a variant is measurably faster (2%)
        if tot[best] > 0.98*tot['base']:
            best = 'base'
        table[g] = best
        report[g] = {v: round(ms, 2) for v, ms in sorted(tot.items())}
    out = {'about': 'per-graph build variant of the refinement kernels (sincfeyn_amd/build.py VARIANTS), the fastest '
                    'sum of the two level kernels in %s (h = %g)' % (os.path.basename(a.file), t['h']),
           'graphs': table, 'ms': report}
    with open(B.TUNE_TABLE, 'w') as f:
        json.dump(out, f, indent=1, sort_keys=True)
    for g in sorted(table):
        print('%-8s %-8s %s' % (g, table[g], report[g]))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest='cmd', required=True)
    b = sub.add_parser('build'); b.add_argument('--graphs', default='g3'); b.add_argument('--variants', default=','.join(B.VARIANTS))
    t = sub.add_parser('time'); t.add_argument('--h', type=float, default=1.0); t.add_argument('--out', required=True)
    c = sub.add_parser('choose'); c.add_argument('file')
    a = ap.parse_args(argv)
    {'build': cmd_build, 'time': cmd_time, 'choose': cmd_choose}[a.cmd](a)


if __name__ == '__main__':
    main()
