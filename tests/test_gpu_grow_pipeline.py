"""The grow's two frontier loops (sfx_opts_t.grow_pipeline).

The pipelined steps (2 for every frontier, 1 -- the default -- for the frontiers of at least 8192 points) take
each frontier's keep decisions on the device from rigorous bounds of the threshold eps 1e-3 A, A the
reference's serial sum of |v| pw in frontier order.  The host sums and checks every frontier one behind.  The
exact loop (0) takes the decisions on the host from the serial A itself (reference driver/sumkey.h:218-264).
Mode 3 forces every grow over to the exact loop.

All four must give the same evaluation counts, kept points, shell and boundary-table counts, and every sum
bit for bit.  The pipelined loop decides the same points as the exact one whenever no value lies between the
bounds; otherwise it hands the grow over, so its results never differ.  The cases cover top = 1 and 2 (grow,
tiers, dilation, refinement) and top = 0 (the grow is the whole sum, with the skip generations)."""
import math

import pytest

CASES = [('g2_00', 1.0, 97, 1e-12, None), ('v2_00', 0.7, 139, 1e-12, None), ('v3_03', 1.0, 97, 1e-12, None),
         ('g3_07', 1.5, 33, 1e-12, None), ('g3_00', 1.0, 97, 1e-10, None), ('g3_lbl', 2.0, 25, 1e-10, None),
         ('v3_p4Sc', 1.0, 97, 1e-12, None), ('g2_01', 1.0, 97, 1e-12, 0), ('v3_00', 1.5, 33, 1e-10, 0)]
COUNTS = ('nev', 'nq', 'kept_top', 'btuns', 'abskipped', 'unresdrop', 'shell_n', 'nev_level', 'levels', 'K', 'top')
SUMS = ('S', 'Ah', 'ref', 'prune', 'trunc', 'unres', 'es1', 'es2', 'shell_sum', 'shell_abs', 'abserr')


def _bits(x):
    return [('nan' if math.isnan(v) else v.hex()) for v in (x if isinstance(x, list) else [x])]


def _same(a, b):
    for k in COUNTS:
        assert a[k] == b[k], (k, a[k], b[k])
    for k in SUMS:
        assert _bits(a[k]) == _bits(b[k]), (k, a[k], b[k])


@pytest.mark.gpu
@pytest.mark.parametrize('name,h,N,eps,top', CASES)
def test_pipelined_grow_equals_exact_loop(sfx, name, h, N, eps, top):
    top = sfx.default_top(h) if top is None else top
    ex = sfx.hsum(name, h, N, eps, top, grow_pipeline=0)
    hy = sfx.hsum(name, h, N, eps, top, grow_pipeline=1)
    pl = sfx.hsum(name, h, N, eps, top, grow_pipeline=2)
    fo = sfx.hsum(name, h, N, eps, top, grow_pipeline=3)
    print(name, h, N, eps, top, 'nev %d kept %d S_h %.17g  retries %d / %d / %d  coarsest %.4f s (pipelined) %.4f s '
          '(by size) %.4f s (exact)' % (pl['nev'], pl['kept_top'], pl['S'][0], hy['grow_retries'], pl['grow_retries'],
                                        fo['grow_retries'], pl['t_grow'], hy['t_grow'], ex['t_grow']), flush=True)
    assert ex['grow_retries'] == 0
    assert fo['grow_retries'] == 1                 # every value between the bounds: handed over once
    assert hy['grow_retries'] in (0, 1) and pl['grow_retries'] in (0, 1)
    _same(hy, ex)
    _same(pl, ex)
    _same(fo, ex)


# the run options that change the grow's decisions or bookkeeping: ABCHECK (the skip class tested in full),
# no |.|-sum skipping, every boundary table kept (sumkey.h:446-453), no skip generations at top = 0
OPTION_CASES = [('v3_03', 1.0, 97, 1e-12, None, {'abcheck': 1}), ('g3_07', 1.5, 33, 1e-12, None, {'abskip': 0}),
                ('g2_00', 1.0, 97, 1e-12, None, {'keepbt': 1}), ('v2_00', 1.0, 97, 1e-12, 0, {'growskip': 0}),
                ('g2_03', 1.0, 97, 1e-12, 0, {'abcheck': 1})]


@pytest.mark.gpu
@pytest.mark.parametrize('name,h,N,eps,top,kw', OPTION_CASES)
def test_pipelined_grow_with_run_options(sfx, name, h, N, eps, top, kw):
    top = sfx.default_top(h) if top is None else top
    ex = sfx.hsum(name, h, N, eps, top, grow_pipeline=0, **kw)
    pl = sfx.hsum(name, h, N, eps, top, grow_pipeline=2, **kw)
    fo = sfx.hsum(name, h, N, eps, top, grow_pipeline=3, **kw)
    print(name, h, N, eps, top, kw, 'nev %d kept %d abmiss %d retries %d / %d' % (
        pl['nev'], pl['kept_top'], pl['abmiss'], pl['grow_retries'], fo['grow_retries']), flush=True)
    assert fo['grow_retries'] == 1 and pl['grow_retries'] in (0, 1)
    for r in (pl, fo):
        _same(r, ex)
        assert r['abmiss'] == ex['abmiss']


@pytest.mark.gpu
def test_pipelined_grow_in_the_engine(sfx):
    """the same jobs through the engine's workers (their own streams and buffers) as one after another"""
    jobs = [(n, h, N, eps, sfx.default_top(h) if t is None else t) for n, h, N, eps, t in CASES[:6]]
    alone = [sfx.hsum(n, h, N, eps, t, grow_pipeline=0) for n, h, N, eps, t in jobs]
    sfx.engine_stop()
    sfx.engine_start(4, 0)
    try:
        ids = [sfx.submit(n, h, N, eps, t) for n, h, N, eps, t in jobs]
        got = [sfx.wait(j)[0] for j in ids]
    finally:
        sfx.engine_stop()
    for a, b in zip(got, alone):
        _same(a, b)
