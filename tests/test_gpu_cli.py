"""The command-line host end to end on the GPU: `sincqed run 1.0 97 1e-12 all2` writes the reference's
results-file format (driver/main.c:66-73, 226, 246), each line matching the reference's own output for
config 1 (tests/golden/config1_sym0.txt, the reference run with SINCQED_SYM=0)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, 'sincfeyn_amd', '_lib', 'sincqed')
GOLD = os.path.join(ROOT, 'tests', 'golden', 'config1_sym0.txt')


def _lines(fn):
    rec, tail = {}, []
    for line in open(fn):
        if line.startswith('# SUM'):
            tail.append(line.split())
        elif line.strip() and not line.startswith('#'):
            rec[line.split()[0]] = line
    return rec, tail


def _split(line):
    """(name, volkov, the numeric columns) of a results line: the Volkov string has blanks"""
    name, rest = line[:8].strip(), line[9:]
    volkov, nums = rest[:28].strip(), rest[28:].split()
    return name, volkov, nums


@pytest.mark.gpu
def test_run_config1_results_file(tmp_path):
    res, log = tmp_path / 'r.txt', tmp_path / 'log.txt'
    env = dict(os.environ, SINCQED_RESULTS=str(res), SINCQED_LOG=str(log), SINCQED_CKPT_DIR=str(tmp_path))
    p = subprocess.run([CLI, 'run', '1.0', '97', '1e-12', 'all2'], env=env, cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    print(p.stdout[-3000:])
    mine, msum = _lines(res)
    gold, gsum = _lines(GOLD)
    assert sorted(mine) == sorted(gold) == ['g2_00', 'g2_01', 'g2_02', 'g2_03', 'v2_00']
    assert open(res).readline() == open(GOLD).readline()                       # the header line
    for g in gold:
        n1, v1, a = _split(mine[g])
        n2, v2, b = _split(gold[g])
        assert (n1, v1) == (n2, v2)
        assert a[:7] == b[:7], g                                                # L D w h N eps top
        assert abs(float(a[7]) - float(b[7])) <= 1e-12*abs(float(b[7])), g     # value
        for j in range(8, 13):                                                  # err disc prune trunc round
            assert abs(float(a[j]) - float(b[j])) <= 2e-3*abs(float(b[j])) + 1e-16, (g, j, a[j], b[j])
        assert abs(float(a[13]) - float(b[13])) <= 1e-8*abs(float(b[13])), g   # S_2h (8 digits printed)
        assert a[14] == b[14], g                                                # evaluations
        assert a[16:] == b[16:], g                                              # ref and diff columns
    assert len(msum) == 1 and msum[0][:4] == gsum[0][:4]
    assert abs(float(msum[0][4]) - float(gsum[0][4])) <= 1e-12*abs(float(gsum[0][4]))
    out = p.stdout
    assert '# sum of 5 graphs' in out and 'evals 1672228' in out


@pytest.mark.gpu
def test_run_box_scan_env(tmp_path, oracle):
    """SINCQED_GROW=0 reaches the library (main.c:183): the evaluations column is the oracle's count for
    the full-box scans, which differs from the grow's"""
    res = tmp_path / 'r.txt'
    env = dict(os.environ, SINCQED_RESULTS=str(res), SINCQED_LOG='-', SINCQED_CKPT='0', SINCQED_GROW='0')
    p = subprocess.run([CLI, 'run', '1.5', '33', '1e-10', '1', 'g2_00', 'v2_00'], env=env, cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    mine, _ = _lines(res)
    for g in ('g2_00', 'v2_00'):
        o = oracle.hsum(g, 1.5, 33, 1e-10, 1, grow=0)
        _, _, a = _split(mine[g])
        assert int(a[14]) == o['nev'], g
        assert int(a[14]) != oracle.hsum(g, 1.5, 33, 1e-10, 1)['nev'], g


@pytest.mark.gpu
def test_run_batched_equals_one_at_a_time(tmp_path):
    """several graphs go through the job engine (the default) or one after another (SINCQED_BATCH=0): the
    same results lines apart from the time column, the same log lines in the same order"""
    out = {}
    for mode in ('1', '0'):
        d = tmp_path / ('b' + mode)
        d.mkdir()
        env = dict(os.environ, SINCQED_RESULTS='r.txt', SINCQED_LOG='-', SINCQED_CKPT='0', SINCQED_BATCH=mode)
        p = subprocess.run([CLI, 'run', '1.0', '97', '1e-12', 'all2', 'v3_00', 'g3_07'], env=env, cwd=str(d),
                           capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stderr
        rec, tail = _lines(d / 'r.txt')
        lines = [l for l in p.stdout.splitlines() if not l.startswith('# sincqed ')]
        out[mode] = (rec, tail, lines)
    (r1, t1, l1), (r0, t0, l0) = out['1'], out['0']
    assert sorted(r1) == sorted(r0) and len(r1) == 7
    for g in r0:
        a, b = _split(r1[g])[2], _split(r0[g])[2]
        assert a[:15] == b[:15] and a[16:] == b[16:], g                       # all but the time column
    assert t1 == t0
    strip = lambda ls: [re.sub(r'\d+\.\ds(\s|$)', 'Ts ', l) for l in ls]      # the per-graph time field
    assert strip(l1) == strip(l0)


@pytest.mark.gpu
def test_run_top0(tmp_path, oracle):
    """`sincqed run 1.0 97 1e-12 0 g2_01`: top = 0 from the command line, the shell line of sum.c:622 in the
    log, the oracle's evaluation count and value in the results line"""
    res = tmp_path / 'r.txt'
    env = dict(os.environ, SINCQED_RESULTS=str(res), SINCQED_LOG='-', SINCQED_CKPT='0')
    p = subprocess.run([CLI, 'run', '1.0', '97', '1e-12', '0', 'g2_01'], env=env, cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    assert '# g2_01: grow shell (evaluated and summed, not expanded):' in p.stdout
    o = oracle.hsum('g2_01', 1.0, 97, 1e-12, 0)
    _, _, a = _split(_lines(res)[0]['g2_01'])
    assert a[6] == '0' and int(a[14]) == o['nev']
    w = oracle.graph('g2_01')['weight']                                   # value = w S_h (main.c:46)
    assert abs(float(a[7]) - w*o['S'][0]) <= 1e-12*abs(w*o['S'][0])
