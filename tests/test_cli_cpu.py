"""CPU tests of the command-line host (src/sincqed.c) that need no GPU: `sincqed compare` of two
results files (driver/main.c:118-153) and the usage line.  The h' = 1.2 results file is made by the
reference implementation (oracle/_ref/sincqed_ref); the h = 1.0 one is the committed golden output."""
import math
import os
import subprocess

import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, 'sincfeyn_amd', '_lib', 'sincqed')
GOLD = os.path.join(ROOT, 'tests', 'golden', 'config1_sym0.txt')

needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason='sincqed not built')


def _results(fn):
    """name -> (h, value, round) by the column rule of a results line (value: first %+.15e token)"""
    out = {}
    for line in open(fn):
        if line.startswith('#') or not line.strip():
            continue
        t = line.split()
        iv = next(i for i in range(1, len(t)) if t[i][0] in '+-' and len(t[i]) >= 20 and 'e' in t[i])
        out[t[0]] = (float(t[iv - 4]), float(t[iv]), float(t[iv + 5]))
    return out


@pytest.fixture(scope='module')
def coarse(tmp_path_factory):
    if not os.path.exists(O.SINCQED_REF):
        pytest.skip('oracle/_ref not built')
    d = tmp_path_factory.mktemp('cmp')
    fn = str(d / 'r12.txt')
    env = dict(os.environ, SINCQED_RESULTS=fn, SINCQED_LOG='-', SINCQED_SYM='0', OMP_NUM_THREADS='4')
    subprocess.run([O.SINCQED_REF, 'run', '1.2', '81', '1e-12', 'all2'], env=env, cwd=str(d), check=True,
                   stdout=subprocess.DEVNULL, timeout=300)
    return fn


def _compare(a, b):
    p = subprocess.run([CLI, 'compare', a, b], capture_output=True, text=True, timeout=60, check=True)
    return p.stdout.splitlines()


@needs_cli
def test_compare_two_steps(coarse):
    """each graph: the value of the smaller step, dS = S(h) - S(h'), error = |dS|/3 + round of the smaller
    step; the order of the two files does not matter; the SUM line adds values and errors"""
    fine, crs = _results(GOLD), _results(coarse)
    lines = _compare(GOLD, coarse)
    rows = [l.split() for l in lines if not l.startswith('#')]
    assert [r[0] for r in rows] == ['g2_00', 'g2_01', 'g2_02', 'g2_03', 'v2_00']
    tot, e1, e2 = 0.0, 0.0, 0.0
    for r in rows:
        hf, vf, rf = fine[r[0]]
        hc, vc, _ = crs[r[0]]
        assert (float(r[1]), float(r[2])) == (hf, hc) == (1.0, 1.2)
        assert r[3] == '%+.15e' % vf
        d = vf - vc
        e = abs(d)/3 + rf
        assert r[4] == '%+.2e' % d and r[5] == '%.2e' % e
        tot += vf; e1 += e; e2 += e*e
    assert lines[-1] == '# SUM 5 graphs  %+.12f  err %.2e (in quadrature)  %.2e (linear)' % (tot, math.sqrt(e2), e1)
    assert _compare(coarse, GOLD)[3:] == lines[3:]


@needs_cli
def test_compare_same_step_skips():
    lines = _compare(GOLD, GOLD)
    assert sum('same h in both files, skipped' in l for l in lines) == 5
    assert lines[-1].startswith('# SUM 0 graphs')


@needs_cli
def test_usage_lists_commands():
    p = subprocess.run([CLI, 'list'], capture_output=True, text=True, timeout=60, check=True)
    assert 'scale GRAPH [kappa] [seed] | compare R1 R2' in p.stdout
