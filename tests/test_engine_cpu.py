"""The job engine's host logic (sincfeyn_amd/csrc/sfx_engine.hpp) on the CPU under ThreadSanitizer, with a
stub backend in place of the device (tests/engine_stub.cpp): the seeded rate-20 request stream of the
serving loop, the hold-and-wait admission case, stop with queued jobs, and the wait deadlines.  Any data
race or lock-order problem in the queue, the budget, the job records or the worker loop fails the test
(TSAN exits non-zero), as does a deadlock (the run's own time limit)."""
import os
import shutil
import subprocess

import pytest

from sincfeyn_amd import serve

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ALL2 = ['g2_00', 'g2_01', 'g2_02', 'g2_03', 'v2_00']
ALL3 = (['g3_%02d' % i for i in range(28)] + ['g3_lbl'] + ['v3_%02d' % i for i in range(8)]
        + ['v3_p4Sc', 'v3_p4Sm', 'v3_p4Vc', 'v3_p4Vm'])


def _dim(name):
    """the stub's job size follows the graph's Sinc dimension (roughly: two-loop 3-4, three-loop 4-7)"""
    return 4 if name[:2] in ('g2', 'v2') else 7 if name.startswith('g3') else 6


@pytest.fixture(scope='module')
def stub(tmp_path_factory):
    if not shutil.which('g++'):
        pytest.skip('g++ not available')
    exe = str(tmp_path_factory.mktemp('engine') / 'engine_stub')
    cmd = ['g++', '-std=c++17', '-O1', '-g', '-fsanitize=thread', '-pthread',
           '-I' + os.path.join(ROOT, 'sincfeyn_amd', 'csrc'), '-I' + os.path.join(ROOT, 'include'),
           os.path.join(HERE, 'engine_stub.cpp'), '-o', exe]
    subprocess.run(cmd, check=True)
    return exe


def _run(exe, *args, timeout=120):
    env = dict(os.environ, TSAN_OPTIONS='halt_on_error=1 exitcode=66 second_deadlock_stack=1')
    p = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout, env=env)
    print(p.stdout, p.stderr[-6000:])
    assert p.returncode == 0, p.stderr[-3000:]
    return p.stdout


def test_engine_serving_stream_rate20_seed7(stub, tmp_path):
    """the serving loop's seed-7 stream at 20 jobs/s (150 jobs over all2 and all3), 16 workers, a budget
    small enough that levels wait for each other: every job finishes with its own result"""
    names = ALL2 + ALL3
    st = serve.make_stream(7, 20.0, 150, names)
    f = tmp_path / 'jobs.txt'
    f.write_text(''.join('%.9f %d %r %d %r %d\n' % (a, _dim(g), h, N, eps, top) for a, g, h, N, eps, top in st))
    out = _run(stub, 'stream', f, 16, 6e10, 0.1)
    assert 'OK stream: 150 jobs' in out


def test_engine_hold_and_wait(stub):
    out = _run(stub, 'holdwait')
    assert 'OK holdwait' in out and 'admitted over cap 1' in out


def test_engine_stop_cancels_queued(stub):
    assert 'OK stop' in _run(stub, 'stop')


def test_engine_deadlines(stub):
    out = _run(stub, 'deadlines')
    assert 'OK deadlines' in out and 'stage' in out
