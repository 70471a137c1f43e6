"""The job engine (continuous batching): jobs run concurrently on worker streams give exactly the results
they give alone, agree with the oracle, and a seeded Poisson stream is served end to end."""
import math

import pytest

MIX = [('g2_00', 1.0, 97, 1e-12), ('g2_01', 1.0, 97, 1e-12), ('g2_02', 1.0, 97, 1e-12), ('g2_03', 1.0, 97, 1e-12),
       ('v2_00', 1.0, 97, 1e-12), ('v3_00', 1.0, 97, 1e-10), ('v3_03', 0.9, 109, 1e-12), ('g3_07', 1.0, 97, 1e-10),
       ('v3_p4Sc', 0.9, 109, 1e-10), ('v3_p4Vc', 1.0, 97, 1e-12), ('g2_00', 0.7, 139, 1e-12), ('v3_01', 1.0, 97, 1e-10)]


def _same(a, b):
    return all((x == y) or (math.isnan(x) and math.isnan(y)) for x, y in zip(a['S'], b['S'])) and \
        (a['nev'], a['nq'], a['kept_top'], a['nev_level']) == (b['nev'], b['nq'], b['kept_top'], b['nev_level'])


@pytest.mark.gpu
def test_mix_equals_alone(sfx):
    alone = [sfx.hsum(g, h, N, eps) for g, h, N, eps in MIX]
    sfx.engine_stop()
    sfx.engine_start(4, 0)
    jobs = [sfx.submit(g, h, N, eps) for g, h, N, eps in MIX]
    mixed = [sfx.wait(j)[0] for j in jobs]
    for (g, h, N, eps), a, m in zip(MIX, alone, mixed):
        assert _same(a, m), (g, h, N, eps, a['S'], m['S'])
    sfx.engine_stop()


@pytest.mark.gpu
def test_engine_jobs_match_oracle(sfx, oracle):
    sfx.engine_start(3, 0)
    cases = MIX[:6]
    jobs = [sfx.submit(g, h, N, eps) for g, h, N, eps in cases]
    for (g, h, N, eps), j in zip(cases, jobs):
        r = sfx.wait(j)[0]
        o = oracle.hsum(g, h, N, eps, sfx.default_top(h))
        assert (r['nev'], r['nq'], r['kept_top']) == (o['nev'], o['nq'], o['kept_top']), g
        assert abs(r['S'][0] - o['S'][0]) <= 1e-12*abs(o['S'][0]), g
    sfx.engine_stop()


@pytest.mark.gpu
def test_serving_stream(sfx, oracle, capsys):
    """a seeded Poisson stream (rate 40/s) of cheap jobs: every job served (one progress line each), each
    result identical to the same job run alone, and each two-loop job's counts and value those of the CPU
    reference (the oracle, no contraction)"""
    from sincfeyn_amd import serve
    sfx.engine_start(8, 0)
    graphs = sfx.select('all2', 'vp3')
    stream = serve.make_stream(20261203, 40.0, 24, graphs, hs=(1.0, 0.9), epss=(1e-10, 1e-12))
    res = [None]*len(stream)
    rep = serve.serve(stream, res, progress=True)
    err = capsys.readouterr().err
    assert [l.split()[1] for l in err.splitlines() if l.startswith('job ')] == \
        ['%d/%d' % (i + 1, len(stream)) for i in range(len(stream))]
    print(rep, flush=True)
    assert rep['jobs'] == len(stream) and rep['latency_p99_s'] >= rep['latency_p50_s'] > 0
    names = {i: g['name'] for i, g in enumerate(sfx.graphs())}
    n2 = 0
    for (a, g, h, N, eps, top), r in zip(stream, res):
        assert _same(r, sfx.hsum(g, h, N, eps, top)), (g, h, eps)
        if names[g].startswith(('g2_', 'v2_')):
            o = oracle.hsum(names[g], h, N, eps, top)
            assert (r['nev'], r['nq'], r['kept_top']) == (o['nev'], o['nq'], o['kept_top']), names[g]
            assert abs(r['S'][0] - o['S'][0]) <= 1e-12*abs(o['S'][0]), names[g]
            n2 += 1
    assert n2 > 0
    sfx.engine_stop()


@pytest.mark.gpu
def test_budget_admission_serialises_when_tight(sfx):
    """with an HBM budget smaller than one level's reservation the engine still completes every job (a job
    is admitted alone when nothing else holds the budget) and the results are unchanged"""
    sfx.engine_stop()
    sfx.engine_start(4, 1e6)
    cases = MIX[:5] + [MIX[10]]          # MIX[10] has top = 2: a job reserving twice must not wait on itself
    jobs = [sfx.submit(g, h, N, eps) for g, h, N, eps in cases]
    for (g, h, N, eps), j in zip(cases, jobs):
        assert _same(sfx.wait(j)[0], sfx.hsum(g, h, N, eps)), g
    sfx.engine_stop()


@pytest.mark.gpu
def test_default_context_after_tight_engine(sfx):
    """stopping the engine lifts its budget: a top = 2 sum on the library's own context afterwards runs"""
    sfx.engine_start(2, 1e6)
    sfx.engine_stop()
    r = sfx.hsum('g2_03', 0.7, 139, 1e-12)
    assert r['nev'] > 0


@pytest.mark.gpu
def test_mix_with_box_scans_and_all_dd(sfx):
    """jobs with other options in the same mix -- full-box scans (SINCQED_GROW=0) and the all-DD error
    reference -- beside default jobs: each result identical to the same job alone"""
    cases = [('g2_02', 1.0, 97, 1e-12, dict(box_scan=1)), ('g2_00', 1.0, 97, 1e-12, {}),
             ('g2_01', 1.0, 97, 1e-12, dict(sfx.ALL_DD)), ('v3_00', 1.5, 33, 1e-10, dict(box_scan=1)),
             ('g3_07', 1.0, 97, 1e-10, {}), ('v2_00', 1.0, 97, 1e-12, dict(box_scan=1, **sfx.ALL_DD))]
    alone = [sfx.hsum(g, h, N, eps, **kw) for g, h, N, eps, kw in cases]
    sfx.engine_stop()
    sfx.engine_start(4, 0)
    jobs = [sfx.submit(g, h, N, eps, **kw) for g, h, N, eps, kw in cases]
    mixed = [sfx.wait(j)[0] for j in jobs]
    for (g, h, N, eps, kw), a, m in zip(cases, alone, mixed):
        assert _same(a, m), (g, kw, a['S'], m['S'])
        assert (a['esok'], a['es1'], a['es2']) == (m['esok'], m['es1'], m['es2']), (g, kw)
    sfx.engine_stop()


@pytest.mark.gpu
@pytest.mark.parametrize('name,h,N,eps,kw', [('g2_01', 1.0, 97, 1e-12, {}), ('v3_03', 0.9, 109, 1e-12, {}),
                                             ('g3_07', 1.0, 97, 1e-10, {'block_points': 1 << 14})])
def test_graph_capture_same_results(sfx, name, h, N, eps, kw):
    """a refinement level's block chain run as one captured HIP graph gives the directly launched result bit
    for bit (the last case: levels of many blocks)"""
    a = sfx.hsum(name, h, N, eps, **kw)
    b = sfx.hsum(name, h, N, eps, graph_capture=1, **kw)
    assert _same(a, b), (name, a['S'], b['S'])


@pytest.mark.gpu
def test_graph_capture_on_engine_workers(sfx):
    """captures on several engine workers at once (thread-local capture on each worker's stream) give the
    results the jobs give alone"""
    alone = [sfx.hsum(g, h, N, eps) for g, h, N, eps in MIX]
    sfx.engine_stop()
    sfx.engine_start(4, 0)
    jobs = [sfx.submit(g, h, N, eps, graph_capture=1) for g, h, N, eps in MIX]
    got = [sfx.wait(j, timeout=600)[0] for j in jobs]
    sfx.engine_stop()
    for (g, h, N, eps), a, m in zip(MIX, alone, got):
        assert _same(a, m), (g, h, N, eps, a['S'], m['S'])
