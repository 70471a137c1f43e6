"""Whole-graph parity of the device hierarchical sum against the oracle (the C restatement of the
reference's driver/sum.c, itself bit-exact against the reference): every graph value within 1e-12
relative, identical evaluation counts per level, identical kept_top (pruning decisions) and fallback
counts.  The symmetry fold is off on both sides, as in the reference's GPU build (sum.c:571)."""
import math

import pytest

ALL2 = ['g2_00', 'g2_01', 'g2_02', 'g2_03', 'v2_00']
REL = 1e-12


def _cmp(sfx, oracle, name, h, N, eps, top=None, rel=REL, **kw):
    top = sfx.default_top(h) if top is None else top
    r = sfx.hsum(name, h, N, eps, top, **kw)
    o = oracle.hsum(name, h, N, eps, top, **kw)
    print(name, h, N, eps, 'device S_h %.17g  oracle %.17g  nev %d/%d nq %d/%d kept %d/%d btuns %d/%d  t %.3fs' % (
        r['S'][0], o['S'][0], r['nev'], o['nev'], r['nq'], o['nq'], r['kept_top'], o['kept_top'], r['btuns'], o['btuns'], r['t']), flush=True)
    assert r['kept_top'] == o['kept_top']
    assert r['nev'] == o['nev']
    assert r['nev_level'] == o['nev_level']
    assert r['nq'] == o['nq']
    for j in range(3):
        a, b = r['S'][j], o['S'][j]
        if math.isnan(b):
            assert math.isnan(a)
        else:
            assert abs(a - b) <= rel*abs(b), (j, a, b)
    assert abs(r['ref'] - o['ref']) <= 1e-13*abs(o['ref'])
    assert r['abskipped'] == o['abskipped']
    # the grow's candidates that fell below the freed boundary tables (the BTBOUND count, sumkey.h:342,397)
    assert r['btuns'] == o['btuns'], (r['btuns'], o['btuns'])
    # the shadow-tier sums (the eps-extrapolation line): the same tier sets, so against the oracle of the
    # same arithmetic the sums differ by their summation error only, far below one tier-1 point (|v| >=
    # 10 thr = 1e-3 eps Ref); against a differently rounded build, by what the values' tolerance allows
    tol = 1e-15 if rel <= REL else rel
    for k in ('es1', 'es2'):
        assert abs(r[k] - o[k]) <= tol*o['Ah'] + 1e-300, (k, r[k], o[k])
    return r, o


@pytest.mark.gpu
@pytest.mark.parametrize('name', ALL2)
def test_config1_all2(sfx, oracle, name):
    """config 1: sincqed run 1.0 97 1e-12 all2"""
    _cmp(sfx, oracle, name, 1.0, 97, 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize('name', ['g2_00', 'g2_03', 'v2_00'])
def test_all2_h08_top2(sfx, oracle, name):
    _cmp(sfx, oracle, name, 0.7, 139, 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize('name', ['v3_00', 'v3_p4Sc', 'g3_07'])
def test_three_loop_small(sfx, oracle, name):
    _cmp(sfx, oracle, name, 1.5, 33, 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize('name', ['v3_01', 'g3_07'])
def test_three_loop_eps1e8(sfx, oracle, oracle_ref, name):
    """eps = 1e-8.  Against the oracle without FP contraction: same counts, values to 1e-12 (every point is
    bit-identical, only the summation order differs).  Against the reference's default build: the
    reference accepts FP64 values whose rounding bound reaches ABSERR = 1e-3 eps Ref per point
    (eval.c:36-38, sum.c:651), so two correct FP64 evaluation orders (FMA-contracted and not) differ by
    more than 1e-12 in the sum; there the value must agree to 1e-9 (the reference's own GPU-vs-CPU
    agreement at eps = 1e-8 was 7.9e-9, docs/GPU.md:77-78) and the counts still exactly on these graphs"""
    _cmp(sfx, oracle, name, 1.5, 33, 1e-8)
    _cmp(sfx, oracle_ref, name, 1.5, 33, 1e-8, rel=1e-9)


@pytest.mark.gpu
def test_lbl_against_reference_nc():
    """g3_lbl (25k FP64 operations per point): against the reference implementation itself built without
    FP contraction (multi-threaded; the single-threaded oracle takes minutes here)"""
    import oracle as O
    from sincfeyn_amd import sfx as S
    S.init(0)
    h, N, eps = 2.0, 25, 1e-10
    ref, wall, _ = O.run_reference(h, N, eps, ['g3_lbl'], top=1, env={'SINCQED_SYM': '0'}, nc=True)
    R = ref['g3_lbl']
    r = S.hsum('g3_lbl', h, N, eps, 1)
    print('g3_lbl device %.17g ref %.17g nev %d/%d nq %d/%d (reference %.1fs, device %.2fs)' % (
        r['S'][0], R['S0'], r['nev'], R['nev'], r['nq'], R['nq'], wall, r['t']), flush=True)
    assert (r['nev'], r['nq'], r['kept_top']) == (R['nev'], R['nq'], R['kept'])
    assert abs(r['S'][0] - R['S0']) <= REL*abs(R['S0'])
    assert abs(r['S'][1] - R['S1']) <= REL*abs(R['S1'])


@pytest.mark.gpu
def test_small_blocks_same_result(sfx, oracle):
    """the refinement's block size changes neither counts nor (beyond 1e-15) the sums"""
    r1 = sfx.hsum('g2_02', 1.0, 97, 1e-12, 1)
    r2 = sfx.hsum('g2_02', 1.0, 97, 1e-12, 1, block_points=1 << 12)
    assert r1['nev'] == r2['nev'] and r1['kept_top'] == r2['kept_top'] and r1['nq'] == r2['nq']
    assert abs(r1['S'][0] - r2['S'][0]) <= 1e-15*abs(r1['S'][0])


@pytest.mark.gpu
def test_all_dd_mode(sfx, oracle):
    """the reference's all-DD error reference (G2_ABSERR=0 SINCQED_ABREL=0 SINCQED_ABSKIP=0)"""
    _cmp(sfx, oracle, 'g2_01', 1.0, 97, 1e-12, abserr0=0.0, abrel=0.0, abskip=0)


@pytest.mark.gpu
@pytest.mark.parametrize('name', ALL2)
def test_config1_vs_default_reference_build(sfx, oracle_ref, name):
    """against the reference's default CPU build (FMA contraction on): the same values to 1e-12 and, on
    config 1, the same counts (measured: no near-threshold decision differs there)"""
    _cmp(sfx, oracle_ref, name, 1.0, 97, 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize('name', ALL2)
def test_all_qd_against_all_dd(sfx, name):
    """the all-QD run of a two-loop sum (every point recomputed in quad-double) against the all-DD run:
    the same points, and the sums differ by no more than the all-DD run's own round column (the
    double-double error model, eval.c:26) -- the FP64-bulk sum's distance from the all-QD one is then
    its arithmetic error"""
    h, N, eps = 1.0, 97, 1e-12
    dd = sfx.hsum(name, h, N, eps, **sfx.ALL_DD)
    qd = sfx.hsum(name, h, N, eps, **sfx.ALL_QD)
    bulk = sfx.hsum(name, h, N, eps)
    assert (qd['nev'], qd['kept_top'], qd['nq']) == (dd['nev'], dd['kept_top'], dd['nq'])
    assert abs(qd['S'][0] - dd['S'][0]) <= 3*dd['unres'] + 1e-30
    assert qd['unres'] <= 1e-50*abs(qd['S'][0]) + 1e-300
    print(name, 'bulk - QD %.2e  DD - QD %.2e  (DD round column %.2e)' % (
        bulk['S'][0] - qd['S'][0], dd['S'][0] - qd['S'][0], dd['unres']), flush=True)
    assert abs(bulk['S'][0] - qd['S'][0]) <= 1e-12*abs(qd['S'][0])


EDGE = [
    ('g2_00', 3.0, 9, 1e-3, 1),        # the smallest lattices: K = 4, a handful of cells
    ('v2_00', 2.0, 5, 1e-6, 1),        # N = 5 -> K = 4 (rounded up to a multiple of 2^(top+1))
    ('g2_02', 1.5, 33, 1e-2, 1),       # a loose eps: most of the grow's frontier is below threshold
    ('g2_01', 1.0, 97, 1e-15, 1),      # a tight eps: deep support, many fallback points
    ('g2_03', 1.0, 97, 1e-12, 3),      # top = 3: three refinement levels
    ('v3_p4Vc', 1.0, 97, 1e-12, 1),    # identically zero: box_zero reports 0 without a grow
    ('g3_07', 2.0, 25, 1e-10, 2),      # a three-loop graph on a coarse lattice at top = 2
    ('v2_00', 0.25, 249, 1e-15, 1),    # the widest lattice of 8-bit points at top = 1: K = 124, support out to the edge
    ('g2_01', 0.5, 241, 1e-12, 2),     # the widest at top = 2: K = 120, support out to the edge
]


@pytest.mark.gpu
@pytest.mark.parametrize('name,h,N,eps,top', EDGE)
def test_edge_cases(sfx, oracle, name, h, N, eps, top):
    """small, coarse, loose, tight, deep and identically-zero sums: the oracle's counts and values"""
    r, o = _cmp(sfx, oracle, name, h, N, eps, top)
    if N > 200:
        assert r['K'] >= 120 and o['trunc'] > 0 and abs(r['trunc'] - o['trunc']) <= 1e-12*o['trunc']   # the edge shells are summed
    if name == 'v3_p4Vc':
        assert r['zero'] == 1 and r['nev'] == 4097 and r['S'][0] == 0


@pytest.mark.gpu
def test_bad_arguments_are_refused(sfx):
    """errors are return codes (never exit(), gpu.cu:16): a lattice wider than 8-bit points, top out of
    range (0 .. 12), a graph index out of range"""
    L = sfx.lib()
    o = sfx.opts(); r = sfx.Result()
    assert L.sfx_hsum(sfx.find('g2_00'), 0.1, 1001, 1e-12, 1, sfx.C.byref(o), sfx.C.byref(r)) == -1
    assert b'K <= 127' in L.sfx_last_error()
    assert L.sfx_hsum(sfx.find('g2_00'), 1.0, 97, 1e-12, 13, sfx.C.byref(o), sfx.C.byref(r)) == -1
    assert L.sfx_hsum(sfx.find('g2_00'), 1.0, 97, 1e-12, -1, sfx.C.byref(o), sfx.C.byref(r)) == -1
    assert L.sfx_hsum(9999, 1.0, 97, 1e-12, 1, sfx.C.byref(o), sfx.C.byref(r)) == -1


SCAN = [
    ('g2_00', 1.5, 33, 1e-10, 1),
    ('g2_03', 1.5, 33, 1e-10, 1),
    ('v2_00', 1.5, 33, 1e-10, 1),
    ('g2_01', 1.5, 33, 1e-10, 2),
    ('v3_00', 1.5, 33, 1e-10, 2),
    ('g2_02', 1.0, 97, 1e-12, 1),     # config 1's widest support: 113k kept nodes (the list grows)
]


def _cmp_scan(sfx, oracle, name, h, N, eps, top, **kw):
    r = sfx.hsum(name, h, N, eps, top, box_scan=1, **kw)
    o = oracle.hsum(name, h, N, eps, top, grow=0)
    print(name, h, N, eps, top, 'box scan: device S_h %.17g  oracle %.17g  nev %d/%d nq %d/%d kept %d/%d  t %.3fs' % (
        r['S'][0], o['S'][0], r['nev'], o['nev'], r['nq'], o['nq'], r['kept_top'], o['kept_top'], r['t']), flush=True)
    assert (r['nev'], r['nq'], r['kept_top'], r['unresdrop']) == (o['nev'], o['nq'], o['kept_top'], o['unresdrop'])
    assert r['nev_level'] == o['nev_level']
    for j in range(3):
        a, b = r['S'][j], o['S'][j]
        if math.isnan(b):
            assert math.isnan(a)
        else:
            assert abs(a - b) <= REL*abs(b), (j, a, b)
    # the scans' sums: double-double trees here, the plain sum in node order in the oracle, whose
    # relative error on a sum of n terms of one sign is at most (n - 1) 2^-53 (prune, trunc and Ah are
    # such sums: the refinement's dropped weights and |.| sums are plain sums there too)
    for k in ('ref', 'prune', 'trunc', 'Ah'):
        assert abs(r[k] - o[k]) <= o['nev']*2.0**-53*abs(o[k]) + 1e-300, k
    assert r['esok'] == 0 and r['es1'] == 0 and r['es2'] == 0
    return r, o


@pytest.mark.gpu
@pytest.mark.parametrize('name,h,N,eps,top', SCAN)
def test_box_scan(sfx, oracle, name, h, N, eps, top):
    """SINCQED_GROW=0: the coarsest level by full-box scans (sum.c:166-191, 610-617) instead of the grow;
    the oracle's restatement is pinned bit for bit against the reference run on one thread"""
    _cmp_scan(sfx, oracle, name, h, N, eps, top)


@pytest.mark.gpu
def test_box_scan_small_blocks(sfx, oracle):
    """many scan blocks (4096 nodes each): the kept list grows across blocks and the block sums combine
    in block order -- the oracle's counts and values again"""
    _cmp_scan(sfx, oracle, 'g2_02', 1.0, 97, 1e-12, 1, block_points=1 << 12)


@pytest.mark.gpu
def test_box_scan_config1_vs_reference(sfx):
    """config 1 with SINCQED_GROW=0 against the reference's default build on several threads (its
    scan sums then have no fixed order): the same counts, values to 1e-12"""
    import oracle as O
    ref, wall, _ = O.run_reference(1.0, 97, 1e-12, ['all2'], env={'SINCQED_SYM': '0', 'SINCQED_GROW': '0'}, threads=8)
    for name in ALL2:
        r = sfx.hsum(name, 1.0, 97, 1e-12, 1, box_scan=1)
        R = ref[name]
        print(name, 'device %.17g ref %.17g' % (r['S'][0], R['S0']), flush=True)
        assert (r['nev'], r['nq'], r['kept_top']) == (R['nev'], R['nq'], R['kept']), name
        assert abs(r['S'][0] - R['S0']) <= REL*abs(R['S0'])
        assert abs(r['S'][1] - R['S1']) <= REL*abs(R['S1'])


@pytest.mark.gpu
def test_box_scan_against_grow_config2(sfx):
    """config 2's shape (h = 0.7, N = 139 -> 145, top = 2) on a D = 6 graph: 37^6 = 2.6e9 nodes in the
    fine box.  The scan covers the whole box, the grow only the support it reaches from z = 0; both sums
    approximate the same lattice sum, so they agree within the pruned weight of both runs"""
    g, h, N, eps = 'g3_01', 0.7, 139, 1e-12
    s = sfx.hsum(g, h, N, eps, box_scan=1)
    r = sfx.hsum(g, h, N, eps)
    d = abs(s['S'][0] - r['S'][0])
    print(g, 'scan %.17g (nev %d, kept %d, %.2fs)  grow %.17g (nev %d, kept %d)  |diff| %.2e  prune %.2e + %.2e' % (
        s['S'][0], s['nev'], s['kept_top'], s['t'], r['S'][0], r['nev'], r['kept_top'], d, s['prune'], r['prune']), flush=True)
    # every node of both boxes (steps 8h and 4h) once; the kept sets differ (the scan's threshold comes
    # from the coarser box's Ref, the grow's from its own A), so only the values are compared
    assert s['nev_level'][2] == 19**6 + 37**6
    assert d <= s['prune'] + r['prune']


@pytest.mark.gpu
def test_config2_full_size_vs_reference_nc(sfx):
    """config 2 at its full size (h = 0.7, N = 139 -> 145, top = 2, eps = 1e-12) for the integrals the
    reference implementation, built without FP contraction and run on the host's cores, finishes in
    seconds (three VP insertions and two three-loop vertex graphs, D = 5-6; g3_01 evaluates 4.7e8 points,
    1.6% of them in double-double): identical evaluation, fallback
    and kept counts, values to 1e-12, the shadow-tier sums through the reference's own eps-extrapolation
    line (printed to four digits)"""
    import oracle as O
    graphs = ['v3_00', 'v3_03', 'v3_04', 'g3_07', 'g3_01']
    ref, wall, out = O.run_reference(0.7, 139, 1e-12, graphs, env={'SINCQED_SYM': '0'}, nc=True)
    lines = {l.split()[1]: l for l in out.splitlines() if 'eps extrapolation:' in l}
    for g in graphs:
        r = sfx.hsum(g, 0.7, 139, 1e-12)
        R = ref[g]
        print(g, 'device %.17g ref %.17g nev %d/%d nq %d/%d kept %d/%d (reference %.1fs for all)' % (
            r['S'][0], R['S0'], r['nev'], R['nev'], r['nq'], R['nq'], r['kept_top'], R['kept'], wall), flush=True)
        assert (r['nev'], r['nq'], r['kept_top'], r['abskipped']) == (R['nev'], R['nq'], R['kept'], R['skipped']), g
        for j, k in ((0, 'S0'), (1, 'S1'), (2, 'S2')):
            assert abs(r['S'][j] - R[k]) <= REL*abs(R[k]), (g, k)
        w = sfx.graphs()[sfx.find(g)]['weight']
        d1, d2 = w*(r['S'][0] - r['es1']), w*(r['es1'] - r['es2'])
        assert 'eps*10 would change the value by %+.3e, eps*100 by a further %+.3e' % (-d1, -d2) in lines[g], (g, lines[g])


@pytest.mark.gpu
@pytest.mark.parametrize('name,h,N,eps', [('g2_01', 1.0, 97, 1e-12), ('g2_03', 1.0, 97, 1e-12), ('v2_00', 1.0, 97, 1e-12),
                                          ('g3_07', 1.5, 33, 1e-12)])
def test_abcheck(sfx, oracle, name, h, N, eps):
    """SINCQED_ABCHECK (eval.c:11-12, 35): the skip class evaluated and tested in full, its fallback points
    counted (ABMISS) -- with SINCQED_NEARF=1, where the count is nonzero (the oracle is pinned on the
    reference there); the default run's values are not changed by the check on these cases"""
    r, o = _cmp(sfx, oracle, name, h, N, eps, abcheck=1, nearf=1.0)
    assert r['abmiss'] == o['abmiss'], (r['abmiss'], o['abmiss'])
    if name in ('g2_03', 'g3_07'):
        assert r['abmiss'] > 0


TOP0 = [('g2_00', 1.0, 97, 1e-12, {}), ('g2_01', 1.0, 97, 1e-12, {}), ('g2_03', 1.0, 97, 1e-12, {}), ('v2_00', 1.0, 97, 1e-12, {}),
        ('v3_00', 1.5, 33, 1e-10, {}), ('g3_07', 1.5, 33, 1e-10, {}),
        ('g2_01', 1.0, 97, 1e-12, dict(growskip=0)), ('g2_03', 1.0, 97, 1e-12, dict(growskip=5, abcheck=1))]


@pytest.mark.gpu
@pytest.mark.parametrize('name,h,N,eps,kw', TOP0)
def test_top0(sfx, oracle, name, h, N, eps, kw):
    """top = 0: the grow at step h is the whole sum (no refinement); its frontier points skip the |.| sums
    for up to growskip generations after a checked point, and the prune column is the shell's signed sum
    (sum.c:619-626) -- the oracle's counts (evaluations, fallbacks, skipped, ABMISS) and values"""
    r, o = _cmp(sfx, oracle, name, h, N, eps, 0, **kw)
    assert r['levels'] == 1 and r['esok'] == 0 and math.isnan(r['S'][1])
    assert r['abmiss'] == o['abmiss']
    for k in ('prune', 'trunc', 'unres'):
        assert abs(r[k] - o[k]) <= 1e-12*abs(o[k]) + 1e-300, (k, r[k], o[k])
    if not kw:
        assert r['abskipped'] > 0
