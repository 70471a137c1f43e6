"""CPU tests of the quad-double arithmetic of the all-QD mode (sincfeyn_amd/csrc/sfx_qd.hpp, host
instantiation through sfx_qd_ops) against 420-bit mpmath, and of the __float128 checker kernels
(oracle/_ref/libsqk_f128.so, the reference's DD=0 build) against the reference's double-double build."""
import os

import numpy as np
import pytest

import oracle as O

mpmath = pytest.importorskip('mpmath')


def _rand_qd(rng, n, scale, normalized=True):
    out = np.zeros((n, 4))
    out[:, 0] = rng.standard_normal(n)*np.exp(rng.uniform(-scale, scale, n))
    for k in range(1, 4):
        out[:, k] = out[:, k - 1]*2.0**-53*rng.uniform(-1, 1, n)
    if not normalized:                       # components of any size: still an exact sum of four doubles
        out[:, 1] *= 2.0**60
    return out


def _mp(q):
    return sum(mpmath.mpf(float(v)) for v in q[::-1])


@pytest.mark.parametrize('scale,normalized', [(0.0, True), (8.0, True), (30.0, True), (4.0, False)])
def test_qd_ops_against_mpmath(scale, normalized):
    """each of + - * / is correct to 1e-63 of its operands' magnitude (|a| + |b| for + -, |result| for
    * /), including exact and near cancellation and operands that are not renormalised"""
    from sincfeyn_amd import sfx
    mpmath.mp.prec = 420
    rng = np.random.default_rng(11)
    n = 600
    a, b = _rand_qd(rng, n, scale, normalized), _rand_qd(rng, n, scale, normalized)
    b[:100] = -a[:100]
    b[:100, 3] *= 1.5                        # cancellation down to the fourth component
    b[100:200, 0] = a[100:200, 0]*(1 + 2.0**-40)
    for op in '+-*/':
        r = sfx.qd_ops(op, a, b)
        for i in range(n):
            x, y, z = _mp(a[i]), _mp(b[i]), _mp(r[i])
            ex = {'+': x + y, '-': x - y, '*': x*y, '/': x/y}[op]
            mag = abs(x) + abs(y) if op in '+-' else abs(ex)
            assert abs(z - ex) <= mpmath.mpf('1e-63')*mag, (op, i, float(abs(z - ex)/mag))


def test_qd_exact_cases():
    from sincfeyn_amd import sfx
    one = np.array([[1.0, 0, 0, 0]])
    third = sfx.qd_ops('/', one, np.array([[3.0, 0, 0, 0]]))
    back = sfx.qd_ops('*', third, np.array([[3.0, 0, 0, 0]]))
    assert abs(back[0].sum() - 1.0) < 1e-60 and back[0, 0] == 1.0
    z = sfx.qd_ops('-', third, third)
    assert (z == 0).all()


@pytest.mark.skipif(not os.path.exists(O.LIB_KERNELS_F128), reason='oracle/_ref not built')
@pytest.mark.parametrize('g', ['g2_00', 'g2_01', 'g2_02', 'g2_03', 'v2_00'])
def test_f128_checker_against_reference_dd(g):
    """the __float128 build of the reference's two-loop kernels agrees with its double-double build
    (no contraction) within twice the double-double error model QEPS Sum|terms| (common.h:7; the model
    is an estimate: 1.01 QEPS is reached on g2_01)"""
    o = O.Oracle(nc=True)
    D = o.graph(g)['D']
    sf = np.exp(np.random.default_rng(3).uniform(-4, 4, size=(400, D)))
    q, d = o.eval_fq128(g, sf), o.eval_fq(g, sf)
    dif = np.abs((q[:, 0] - d[:, 0]) + (q[:, 1] - d[:, 1]))
    assert (dif <= 2*4.93e-32*q[:, 2]).all()
