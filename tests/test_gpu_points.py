"""Per-point parity of the device integrands (the HIP emitter's output, compiled for gfx950) against the
reference's generated kernels evaluated on the CPU with the same arithmetic: the QREAL half in plain FP64
and in the reference's double-double, both with FP contraction off (oracle/_ref/libsqk_nc.so).
Pattern of the reference's tests/check_onept.sh:18-43 (random nodes s = e^z, |z| <= 3); plus lattice
nodes of the benchmark lattices and the reference-scale weight products.  Bar: bit-exact."""
import numpy as np
import pytest

import oracle as O

GRAPHS = ['g2_00', 'g2_01', 'g2_02', 'g2_03', 'v2_00'] + ['g3_%02d' % i for i in range(28)] + ['g3_lbl'] + \
         ['v3_%02d' % i for i in range(8)] + ['v3_p4Sc', 'v3_p4Sm', 'v3_p4Vc', 'v3_p4Vm'] + \
         ['w3_p40c', 'w3_p40d', 'w3_p40m', 'w3_p41c', 'w3_p41d', 'w3_p41m']


@pytest.fixture(scope='module')
def nck():
    return O.NcKernels()


def _points(D, n, seed):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(-3.0, 3.0, size=(n, D)))


def _present(sfx, name):
    try:
        return sfx.find(name)
    except KeyError:
        pytest.skip('%s not in this build' % name)


@pytest.mark.gpu
@pytest.mark.parametrize('name', GRAPHS)
def test_fp64_points_bitexact(sfx, nck, name):
    gi = _present(sfx, name)
    D = sfx.graphs()[gi]['D']
    sf = _points(D, 200, 1234 + gi)
    dev = sfx.eval_raw(gi, sf, prec=0)
    f, ab = nck.fp64(name, sf)
    assert np.array_equal(dev[:, 0], f, equal_nan=True), (name, np.max(np.abs(dev[:, 0] - f)/np.abs(f)))
    assert np.array_equal(dev[:, 1], ab, equal_nan=True), name


@pytest.mark.gpu
@pytest.mark.parametrize('name', GRAPHS)
def test_dd_points_bitexact(sfx, nck, name):
    gi = _present(sfx, name)
    D = sfx.graphs()[gi]['D']
    sf = _points(D, 64, 99 + gi)
    dev = sfx.eval_raw(gi, sf, prec=1)
    ref = nck.dd(name, sf)
    assert np.array_equal(dev, ref, equal_nan=True), (name, np.max(np.abs(dev - ref)))


@pytest.mark.gpu
@pytest.mark.parametrize('name', ['g2_00', 'g2_03', 'v2_00', 'g3_00', 'v3_01'])
def test_eval_nodes_weighted(sfx, oracle, name):
    """sfx_eval_nodes (the reference GPU kernel's mode 0: f w and Sum|terms| w) on lattice nodes, against
    the reference's double half fd (with its FMA contraction) times the same weight: equal to the FP64
    rounding of the two evaluation orders"""
    gi = _present(sfx, name)
    D = sfx.graphs()[gi]['D']
    h, K = 1.0, 48
    rng = np.random.default_rng(7)
    k = rng.integers(-12, 13, size=(300, D)).astype(np.int16)
    fw, abw = sfx.eval_nodes(gi, h, K, k)
    EX = np.exp(np.arange(-K, K + 1)*h)
    sf = EX[k.astype(int) + K]
    w = np.ones(len(k))
    for d in range(D):
        w = w*(sf[:, d]*h)
    f, ab = oracle.eval_fd(name, sf)
    ok = np.abs(fw - f*w) <= 1e-13*ab*w + 1e-300
    assert ok.all(), (name, np.max(np.abs(fw - f*w)/(ab*w)))


@pytest.mark.gpu
def test_eval_cells_matches_nodes(sfx):
    """mode 1 (children made from the cell, k = cell - st b; below the box -> 0) equals mode 0 on the
    same children (gpu_kernel.h:27-31)"""
    gi = sfx.find('g2_02')
    D, h, K, st = 4, 1.0, 48, 1
    cells = np.array([[0, 0, 0, 0], [3, -2, 1, 5], [-48, 4, -47, 0], [10, 10, -10, -48]], dtype=np.int16)
    fw, abw = sfx.eval_cells(gi, h, K, cells, st)
    for c in range(len(cells)):
        for b in range(1 << D):
            kk = np.array([cells[c, d] - (st if (b >> d) & 1 else 0) for d in range(D)], dtype=np.int16)
            i = (c << D) | b
            if (kk < -K).any():
                assert fw[i] == 0 and abw[i] == 0
            else:
                f1, a1 = sfx.eval_nodes(gi, h, K, kk[None, :])
                assert fw[i] == f1[0] and abw[i] == a1[0]


@pytest.mark.gpu
@pytest.mark.parametrize('name,E', [('g2_00', 1), ('g2_01', 2), ('g2_03', 1), ('g3_00', 1), ('g3_07', 2),
                                    ('g3_lbl', 1)])
def test_scale_test_matches_oracle(sfx, oracle, name, E):
    """`sincqed scale` (main.c:79-116): the device's double-double point function and the host's
    direction loop give the oracle's result bit for bit -- counts, worst direction, every |g|; at E = 2 this
    includes the directions whose |g| sits at the double-double floor (g3_07 has one 'slow' one there)"""
    d = sfx.scale_test(name, 24.0, 1, E)
    o = oracle.scale(name, 24.0, 1, E)
    assert d == o, (name, E)


@pytest.mark.gpu
def test_scale_test_vp_graph(sfx):
    assert sfx.scale_test('v3_00') is None


@pytest.mark.gpu
def test_qd_ops_device_equals_host(sfx):
    """the device's quad-double operations are the host's, bit for bit (same code, no contraction)"""
    rng = np.random.default_rng(5)
    a = np.zeros((4000, 4)); b = np.zeros((4000, 4))
    for x in (a, b):
        x[:, 0] = rng.standard_normal(4000)*np.exp(rng.uniform(-20, 20, 4000))
        for k in range(1, 4):
            x[:, k] = x[:, k - 1]*2.0**-53*rng.uniform(-1, 1, 4000)
    b[:500] = -a[:500]; b[:500, 2] *= 3.0
    for op in '+-*/':
        assert (sfx.qd_ops(op, a, b, device=True) == sfx.qd_ops(op, a, b)).all(), op


@pytest.mark.gpu
@pytest.mark.parametrize('name', ['g2_00', 'g2_01', 'g2_02', 'g2_03', 'v2_00'])
def test_qd_points_against_f128(sfx, oracle, name):
    """the device's quad-double integrand against the reference's kernels built with QREAL = __float128:
    |f_qd - f_f128| <= 64 eps128 Sum|terms| (eps128 = 1.93e-34, common.h:15; measured up to 20: the
    intermediate values of the generated expressions exceed Sum|terms|), four times tighter than the
    double-double agreement, which is checked too (<= 4 QEPS Sum|terms|, measured up to 1.6)"""
    gi = sfx.find(name)
    D = sfx.graphs()[gi]['D']
    sf = np.exp(np.random.default_rng(17).uniform(-4, 4, size=(2000, D)))
    q = sfx.eval_raw(gi, sf, prec=2)
    f = oracle.eval_fq128(name, sf)
    d = sfx.eval_raw(gi, sf, prec=1)
    ab = f[:, 2]
    r128 = np.abs((q[:, 0] - f[:, 0]) + (q[:, 1] - f[:, 1]))/ab
    rdd = np.abs((q[:, 0] - d[:, 0]) + (q[:, 1] - d[:, 1]))/ab
    print(name, 'max |qd - f128|/ab %.3g  max |qd - dd|/ab %.3g' % (r128.max(), rdd.max()), flush=True)
    assert (r128 <= 64*1.93e-34).all()
    assert (rdd <= 4*4.93e-32).all()
    assert (np.abs(q[:, 2] - ab) <= 1e-30*ab).all()


@pytest.mark.gpu
def test_qd_graph_without_kernel(sfx):
    """three-loop graphs have no quad-double kernel: the mode is refused"""
    with pytest.raises(sfx.SfxError):
        sfx.hsum('g3_07', 1.0, 97, 1e-10, **sfx.ALL_QD)
