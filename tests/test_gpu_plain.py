"""The plain one-graph-per-launch GPU path: the reference's GPU back-end ABI (driver/gpu.h:12-22) served
by this library (sfx_slot_*, include/sfx.h; src/sfx_gpu.c binds gpu_setup / gpu_submit /
gpu_submit_cells / gpu_wait to it).

* Per point: what a slot returns (f w and Sum|terms| w, untested, 0 where skipped or below the box) is
  bit-identical to the library's golden per-point primitives at the same nodes and cells.
* Whole sums: the reference's own host driver built as its GPU build (oracle/_ref/sincqed_ref_gpu:
  driver/sum.c, sumkey.h, eval.c compiled with -DSQ_GPU and without FP contraction, over this back end)
  gives the same evaluation, fallback and kept counts as the reference's CPU build without contraction
  (sincqed_ref_nc), and the same values to 1e-12 (identical where the per-point values are: both sum them
  in the same order)."""
import math

import numpy as np
import pytest

MAXD = 12          # the reference's node stride (sq.h:14)


def _lattice(h, K):
    return np.array([math.exp(k*h) for k in range(-K, K + 1)])


@pytest.mark.gpu
@pytest.mark.parametrize('name', ['g2_00', 'v2_00', 'g3_00', 'g3_lbl', 'v3_p4Sc'])
def test_slot_nodes_and_cells_bitexact(sfx, name):
    gi = sfx.find(name)
    D = sfx.graphs()[gi]['D']
    h, K = 0.7, 72
    rng = np.random.default_rng(7)
    assert sfx.slot_setup(name, D, _lattice(h, K), K, h) == 1
    # nodes, in the reference's layout (int rows of stride MAXD), half of them skipping the |.| sums;
    # both slots in flight at once
    n = 5000
    k = np.zeros((2*n, MAXD), dtype=np.int32)
    k[:, :D] = rng.integers(-K, K + 1, size=(2*n, D))
    skip = (rng.random(2*n) < 0.5).astype(np.uint8)
    sfx.slot_submit(0, k[:n], skip[:n])
    sfx.slot_submit(1, k[n:], skip[n:])
    f1, a1 = sfx.slot_wait(1, n)
    f0, a0 = sfx.slot_wait(0, n)
    fw, abw = sfx.eval_nodes(gi, h, K, k[:, :D])
    assert np.array_equal(np.concatenate([f0, f1]).view(np.int64), fw.view(np.int64))
    ab = np.concatenate([a0, a1])
    assert np.all(ab[skip == 1] == 0)
    assert np.array_equal(ab[skip == 0].view(np.int64), abw[skip == 0].view(np.int64))
    # cells at step 2, some next to the lower edge of the box (children below it give 0, 0)
    nc, st = 600, 2
    cells = np.zeros((nc, MAXD), dtype=np.int32)
    cells[:, :D] = rng.integers(-K, K + 1, size=(nc, D))
    cells[:50, 0] = -K
    cskip = (rng.random(nc) < 0.5).astype(np.uint8)
    sfx.slot_submit(0, cells, cskip, cells_step=st)
    f, a = sfx.slot_wait(0, nc << D)
    fw, abw = sfx.eval_cells(gi, h, K, cells[:, :D], st)
    assert np.array_equal(f.view(np.int64), fw.view(np.int64))
    sk = np.repeat(cskip, 1 << D)
    assert np.all(a[sk == 1] == 0)
    assert np.array_equal(a[sk == 0].view(np.int64), abw[sk == 0].view(np.int64))


@pytest.mark.gpu
def test_slot_setup_unknown_graph_stays_on_cpu(sfx):
    assert sfx.slot_setup('g4_00', 8, _lattice(1.0, 48), 48, 1.0) == 0


def _plain_vs_nc(h, N, eps, graphs, top=None):
    import oracle as O
    g, wg, outg = O.run_reference(h, N, eps, graphs, top=top, env={'SINCQED_SYM': '0'}, gpu=True, threads=8)
    c, wc, _ = O.run_reference(h, N, eps, graphs, top=top, env={'SINCQED_SYM': '0'}, nc=True, threads=8)
    assert '# gpu off' not in outg and '# gpu ' in outg, 'the GPU build ran on the CPU'
    for name in graphs:
        a, b = g[name], c[name]
        print(name, 'plain GPU S_h %.17g  CPU nc %.17g  nev %d/%d nq %d/%d kept %d/%d  t %.2f/%.2f s' % (
            a['S0'], b['S0'], a['nev'], b['nev'], a['nq'], b['nq'], a['kept'], b['kept'], a['t'], b['t']), flush=True)
        for f in ('nev', 'nq', 'kept', 'skipped', 'unresdrop'):
            assert a[f] == b[f], (name, f)
        # the device evaluates each kernel's QREAL half in FP64, the CPU build its double half: bit-identical
        # for every kernel but g3_lbl, whose two halves round differently by ~1e-17 per point (a few 1e-14
        # in its sums); the bar is the north star's 1e-12
        for f in ('S0', 'S1', 'S2', 'Ah', 'ref'):
            if not math.isnan(b[f]):
                assert abs(a[f] - b[f]) <= 1e-12*abs(b[f]) + 1e-300, (name, f, a[f], b[f])
        # the dropped weight sums the |values| of points below the keep threshold, which are what is left
        # of cancellations: against the scale of the sum, not its own
        assert abs(a['prune'] - b['prune']) <= 1e-12*b['Ah'], (name, a['prune'], b['prune'])


@pytest.mark.gpu
def test_plain_path_config1():
    """config 1 (sincqed run 1.0 97 1e-12 all2) through the reference driver's GPU build"""
    _plain_vs_nc(1.0, 97, 1e-12, ['g2_00', 'g2_01', 'g2_02', 'g2_03', 'v2_00'])


@pytest.mark.gpu
def test_plain_path_three_loop():
    """three-loop graphs at top 2 (nodes at the coarser levels, cells at the last), the value-branching Pi4
    contact graph and the light-by-light graph"""
    _plain_vs_nc(1.5, 33, 1e-12, ['g3_07', 'v3_00', 'v3_p4Sc', 'g3_lbl'], top=2)
