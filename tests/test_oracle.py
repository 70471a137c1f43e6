"""CPU tests of the checkers: the reference build (oracle/_ref) against its committed golden outputs and
the physics closed forms; the C restatement (oracle/sq_oracle.c) bit-exact against the reference, for
both floating-point variants; the reference kernels against the committed per-point vectors."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, 'golden')
ALL2 = ['g2_00', 'g2_01', 'g2_02', 'g2_03', 'v2_00']


def _gold_records(fname):
    recs, lines = {}, {}
    for line in open(os.path.join(GOLD, fname)):
        if line.startswith('#R '):
            t = line.split()
            recs[t[1]] = {k: float(v) for k, v in zip(t[2::2], t[3::2])}
        elif line.strip() and not line.startswith('#'):
            lines[line.split()[0]] = line.rstrip('\n')
    return recs, lines


needs_ref = pytest.mark.skipif(not os.path.exists(O.SINCQED_REF), reason='oracle/_ref not built (reference sources absent)')


@needs_ref
@pytest.mark.parametrize('mode,env', [('default', {}), ('sym0', {'SINCQED_SYM': '0'})])
def test_reference_build_reproduces_golden_config1(mode, env):
    """the reference built by oracle/Makefile gives the committed config-1 results (values to all printed
    digits, every count; the time column is not compared)"""
    recs, lines = _gold_records('config1_%s.txt' % mode)
    out, wall, stdout = O.run_reference(1.0, 97, 1e-12, ['all2'], env=env, threads=4)
    for g in ALL2:
        for k in ('S0', 'S1', 'ref', 'Ah', 'prune', 'trunc', 'nev', 'nq', 'kept', 'skipped', 'unresdrop'):
            assert out[g][k] == recs[g][k], (g, k, out[g][k], recs[g][k])


def test_golden_config1_physics():
    """config 1 against the exact two-loop values (reference graphs.c:36-46) within the reference's own
    regression tolerance for this lattice (tests/run_tests.sh:20-22: the all2 sum at h = 1 within 3e-5 of
    A2 = -0.328478965579, the Pi2 insertion within 3e-6) -- the physics, not the 1e-12 parity"""
    recs, lines = _gold_records('config1_default.txt')
    tot = 0.0
    for g in ALL2:
        tok = lines[g].split()
        tot += float([t for t in tok if ('e-' in t or 'e+' in t) and len(t) > 18][0])
    assert abs(tot - (-0.328478965579)) < 3e-5
    v2 = float([t for t in lines['v2_00'].split() if ('e-' in t or 'e+' in t) and len(t) > 18][0])
    assert abs(v2 - 0.015687421859) < 3e-6


@pytest.fixture(scope='module')
def orc():
    if not O.available():
        pytest.fail('oracle not built')
    return {'ref': O.Oracle(nc=False), 'nc': O.Oracle(nc=True)}


@needs_ref
@pytest.mark.parametrize('variant', ['ref', 'nc'])
def test_oracle_bitexact_vs_reference_config1(orc, variant):
    """the restatement reproduces the reference implementation bit for bit (symmetry fold off)"""
    ref, wall, ref_out = O.run_reference(1.0, 97, 1e-12, ['all2'], env={'SINCQED_SYM': '0'}, threads=4, nc=(variant == 'nc'))
    for g in ALL2:
        r = orc[variant].hsum(g, 1.0, 97, 1e-12, 1)
        R = ref[g]
        assert r['S'][0] == R['S0'] and r['S'][1] == R['S1'], g
        assert (r['nev'], r['nq'], r['kept_top'], r['abskipped'], r['unresdrop']) == \
               (R['nev'], R['nq'], R['kept'], R['skipped'], R['unresdrop']), g
        for k in ('Ah', 'ref', 'prune', 'trunc'):
            assert r[k] == R[k], (g, k)
    _check_extrapolation(orc[variant], ref_out, 1.0, 97, 1e-12, 1, ALL2)


def _check_extrapolation(o, stdout, h, N, eps, top, graphs):
    """the shadow-tier sums against the reference's eps-extrapolation line (sum.c:806-809, printed to four
    digits): d1 = w (S_h - ES1), d2 = w (ES1 - ES2)"""
    lines = {l.split()[1]: l for l in stdout.splitlines() if 'eps extrapolation:' in l}
    for g in graphs:
        r = o.hsum(g, h, N, eps, top)
        w = o.graph(g)['weight']
        d1, d2 = w*(r['S'][0] - r['es1']), w*(r['es1'] - r['es2'])
        want = 'eps*10 would change the value by %+.3e, eps*100 by a further %+.3e' % (-d1, -d2)
        assert want in lines[g], (g, want, lines[g])


@needs_ref
@pytest.mark.parametrize('variant', ['ref', 'nc'])
@pytest.mark.parametrize('h,N,eps,graphs', [(1.5, 33, 1e-10, ['v3_00', 'v3_03', 'v3_p4Sc', 'v3_p4Vc']),
                                            (0.7, 139, 1e-12, ['v2_00', 'g2_03'])])
def test_oracle_bitexact_vs_reference_more(orc, variant, h, N, eps, graphs):
    """three-loop VP graphs (incl. the value-branching Pi4 contact and the identically zero v3_p4Vc) and
    top = 2 lattices"""
    top = O.default_top(h)
    ref, wall, ref_out = O.run_reference(h, N, eps, graphs, env={'SINCQED_SYM': '0'}, threads=4, nc=(variant == 'nc'))
    for g in graphs:
        r = orc[variant].hsum(g, h, N, eps, top)
        R = ref[g]
        assert (r['nev'], r['nq'], r['kept_top']) == (R['nev'], R['nq'], R['kept']), g
        for j, k in ((0, 'S0'), (1, 'S1'), (2, 'S2')):
            assert r['S'][j] == R[k] or (math.isnan(r['S'][j]) and math.isnan(R[k])), (g, k)
    _check_extrapolation(orc[variant], ref_out, h, N, eps, top, [g for g in graphs if ref[g]['S0'] != 0])


@needs_ref
@pytest.mark.parametrize('variant', ['ref', 'nc'])
@pytest.mark.parametrize('h,N,eps,top,graphs', [(1.5, 33, 1e-10, 1, ['g2_00', 'g2_03', 'v2_00']),
                                                 (1.5, 33, 1e-10, 2, ['g2_01', 'v3_00']),
                                                 (1.0, 97, 1e-12, 1, ['g2_02'])])
def test_oracle_scan_bitexact_vs_reference(orc, variant, h, N, eps, top, graphs):
    """SINCQED_GROW=0, the full-box scans instead of the grow (sum.c:166-191, 610-617): the reference on one
    thread (its OpenMP reduction is then the plain sum in node order) against the restatement, bit for bit"""
    ref, wall, _ = O.run_reference(h, N, eps, graphs, top=top, env={'SINCQED_SYM': '0', 'SINCQED_GROW': '0'}, threads=1,
                                   nc=(variant == 'nc'))
    for g in graphs:
        r = orc[variant].hsum(g, h, N, eps, top, grow=0)
        R = ref[g]
        assert (r['nev'], r['nq'], r['kept_top'], r['unresdrop']) == (R['nev'], R['nq'], R['kept'], R['unresdrop']), g
        for j, k in ((0, 'S0'), (1, 'S1'), (2, 'S2')):
            assert r['S'][j] == R[k] or (math.isnan(r['S'][j]) and math.isnan(R[k])), (g, k)
        for k in ('Ah', 'ref', 'prune', 'trunc', 'unres'):
            assert r[k] == R[k], (g, k)
        # a box scan sets no shadow tiers: the final level's tier sums are empty
        assert r['es1'] == 0 and r['es2'] == 0


def test_golden_points_reference_kernels(orc):
    """the reference kernels as built here reproduce the committed per-point vectors exactly (so the
    checkers' kernels are the ones the fixtures were made with)"""
    gold = json.load(open(os.path.join(GOLD, 'points.json')))
    nck = O.NcKernels()
    for g, d in gold.items():
        sf = np.array([float.fromhex(x) for x in d['sf']]).reshape(-1, d['D'])
        f, ab = orc['ref'].eval_fd(g, sf)
        assert [x.hex() for x in f] == d['fd'], g
        f64, _ = nck.fp64(g, sf)
        assert [x.hex() for x in f64] == d['fp64'], g
        dd = nck.dd(g, sf)
        assert [float(x).hex() for x in dd.ravel()] == d['dd'], g


def test_dd_against_fp64_error_bound():
    """the double-double values agree with the FP64 ones within the FP64 rounding bound the reference's
    fallback test assumes (16 eps Sum|terms|, eval.c:36), and are far tighter: a dropped or wrong lo term
    in the double-double arithmetic would show as a difference of FP64 size between dd and a better
    evaluation; here the dd value must also agree with the FMA-contracted double half within that bound"""
    gold = json.load(open(os.path.join(GOLD, 'points.json')))
    for g, d in gold.items():
        f = np.array([float.fromhex(x) for x in d['fd']])
        f64 = np.array([float.fromhex(x) for x in d['fp64']])
        ab = np.array([float.fromhex(x) for x in d['fp64_ab']])
        dd = np.array([float.fromhex(x) for x in d['dd']]).reshape(-1, 4)
        v = dd[:, 0] + dd[:, 1]
        bound = 16*np.finfo(float).eps*ab
        assert np.all(np.abs(v - f64) <= bound + 1e-300), g
        assert np.all(np.abs(v - f) <= bound + 1e-300), g
        # the pair is normalised: |lo| <= ulp(hi)/2
        assert np.all(np.abs(dd[:, 1]) <= np.spacing(np.abs(dd[:, 0]))/2 + 1e-300), g


@pytest.mark.parametrize('g', ['g2_00', 'g2_01', 'g2_02', 'g2_03', 'g3_00', 'g3_07', 'g3_21'])
def test_scale_oracle_invariants(g):
    """the restated scale test (main.c:79-116) on the reference's own point functions: (2E+1)^(n-1) - 1
    directions; at E = 1 no slow direction for these subtracted integrands, and the slowest decay of |pt x measure|
    between kappa 2/3 and kappa is the one power of rho that a single-line scaling leaves, e^(-kappa/3)
    (to 1e-4 at kappa = 24), on a direction that moves one line only -- a wrong exponent sign, kappa step
    or fixed line breaks one of these"""
    o = O.Oracle(nc=True)
    r = o.scale(g, 24.0, 1, 1)
    assert r['ndir'] == 3**(r['n'] - 1) - 1
    assert r['bad'] == 0 and r['slow'] == []
    assert abs(r['worst']/math.exp(-8.0) - 1) < 1e-4
    assert r['e_worst'][r['fixed']] == 0
    assert r['g_worst'][0] > r['g_worst'][1] > r['g_worst'][2] > 0
    # E = 2 reaches scalings where |g| sinks to the double-double floor (1e-16 .. 1e-23) and ratios are
    # roundoff; only the direction count is fixed there
    assert o.scale(g, 24.0, 1, 2)['ndir'] == 5**(r['n'] - 1) - 1


def test_scale_oracle_vp_graph_has_none():
    assert O.Oracle(nc=True).scale('v3_00') is None


@needs_ref
@pytest.mark.parametrize('variant', ['ref', 'nc'])
@pytest.mark.parametrize('h,N,eps,graphs', [(1.0, 97, 1e-12, ['g2_01', 'g2_03', 'v2_00']), (1.5, 33, 1e-12, ['g3_07'])])
def test_oracle_abcheck_vs_reference(orc, variant, h, N, eps, graphs):
    """SINCQED_ABCHECK=1 (eval.c:11-12, 35): the skip-class points are evaluated and tested in full and those
    that needed the fallback counted (ABMISS).  With SINCQED_NEARF=1 (a skip bit for every point not over
    ABSERR itself) the count is nonzero; the restatement gives the reference's values and counts exactly"""
    top = O.default_top(h)
    ref, wall, _ = O.run_reference(h, N, eps, graphs, env={'SINCQED_SYM': '0', 'SINCQED_ABCHECK': '1', 'SINCQED_NEARF': '1'},
                                   threads=4, nc=(variant == 'nc'))
    for g in graphs:
        r = orc[variant].hsum(g, h, N, eps, top, abcheck=1, nearf=1.0)
        R = ref[g]
        assert (r['nev'], r['nq'], r['kept_top'], r['abskipped'], r['abmiss']) == \
               (R['nev'], R['nq'], R['kept'], R['skipped'], R['abmiss']), g
        assert r['S'][0] == R['S0'] and r['S'][1] == R['S1'], g
    assert sum(ref[g]['abmiss'] for g in graphs) > 0


@needs_ref
@pytest.mark.parametrize('variant', ['ref', 'nc'])
@pytest.mark.parametrize('h,N,eps,graphs,env', [(1.0, 97, 1e-12, ['g2_00', 'g2_01', 'g2_03', 'v2_00'], {}),
                                                 (1.5, 33, 1e-10, ['v3_00', 'g3_07'], {}),
                                                 (1.0, 97, 1e-12, ['g2_01', 'g2_03'], {'SINCQED_GROWSKIP': '0'}),
                                                 (1.0, 97, 1e-12, ['g2_03'], {'SINCQED_GROWSKIP': '5', 'SINCQED_ABCHECK': '1'})])
def test_oracle_top0_vs_reference(orc, variant, h, N, eps, graphs, env):
    """top = 0 (sum.c:619-626, sumkey.h:165-168, 367-368): the grow at step h is the whole sum; its frontier
    points skip the |.| sums for up to SINCQED_GROWSKIP generations after a checked point, and the prune
    column is the signed sum of the never-expanded shell -- bit for bit against the reference"""
    e = {'SINCQED_SYM': '0'}
    e.update(env)
    ref, wall, _ = O.run_reference(h, N, eps, graphs, top=0, env=e, threads=4, nc=(variant == 'nc'))
    kw = {}
    if 'SINCQED_GROWSKIP' in env:
        kw['growskip'] = int(env['SINCQED_GROWSKIP'])
    if 'SINCQED_ABCHECK' in env:
        kw['abcheck'] = 1
    for g in graphs:
        r = orc[variant].hsum(g, h, N, eps, 0, **kw)
        R = ref[g]
        assert (r['nev'], r['nq'], r['kept_top'], r['abskipped'], r['abmiss']) == \
               (R['nev'], R['nq'], R['kept'], R['skipped'], R['abmiss']), g
        assert r['S'][0] == R['S0'] and math.isnan(r['S'][1]) and math.isnan(R['S1']), g
        for k in ('ref', 'prune', 'trunc', 'unres', 'Ah'):
            assert r[k] == R[k], (g, k)
