"""Writes tests/golden/points.json: per-point golden vectors of every integrand, computed by the
reference's own generated kernels compiled where they lie (oracle/_ref): at 8 points s = e^z (z uniform
in [-3, 3], fixed seed) per graph, the double half fd as the reference's default build computes it (FMA
contraction on), the QREAL half in plain FP64 without contraction, and the QREAL half in the
reference's double-double without contraction (f.hi, f.lo, ab.hi, ab.lo).  Values are stored as hex
floats so that they round-trip exactly.

    python3 tests/golden/make_points.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', '..', 'oracle'))
import oracle as O  # noqa: E402

GRAPHS = ['g2_00', 'g2_01', 'g2_02', 'g2_03', 'v2_00'] + ['g3_%02d' % i for i in range(28)] + ['g3_lbl'] + \
         ['v3_%02d' % i for i in range(8)] + ['v3_p4Sc', 'v3_p4Sm', 'v3_p4Vc', 'v3_p4Vm'] + \
         ['w3_p40c', 'w3_p40d', 'w3_p40m', 'w3_p41c', 'w3_p41d', 'w3_p41m']


def hx(a):
    return [float(x).hex() for x in np.ravel(a)]


def main():
    ref = O.Oracle(nc=False)
    nck = O.NcKernels()
    out = {}
    for i, g in enumerate(GRAPHS):
        D = ref.graph(g)['D']
        rng = np.random.default_rng(20261203 + i)
        sf = np.exp(rng.uniform(-3.0, 3.0, size=(8, D)))
        f, ab = ref.eval_fd(g, sf)
        f64, ab64 = nck.fp64(g, sf)
        dd = nck.dd(g, sf)
        out[g] = {'D': D, 'sf': hx(sf), 'fd': hx(f), 'fd_ab': hx(ab), 'fp64': hx(f64), 'fp64_ab': hx(ab64), 'dd': hx(dd)}
    with open(os.path.join(HERE, 'points.json'), 'w') as fh:
        json.dump(out, fh, indent=0)
    print('wrote %d graphs' % len(out))


if __name__ == '__main__':
    main()
