"""Checkpoints in the reference's SQCK4 format (driver/checkpoint.c): a device sum stopped after a
checkpoint resumes to the unbroken result, the reference implementation resumes a device checkpoint,
and the device resumes a checkpoint the reference wrote.  Counts are exact; values agree to 1e-14
relative (a resumed level continues from its partial sums rounded to one double each, as the file
holds them)."""
import os
import subprocess
import time

import pytest

import oracle as O

CASE = ('g3_07', 0.7, 139, 1e-12)          # top = 2: levels 1 and 0 are refined; ~2.7e7 evaluations


def _same_counts(a, b):
    return (a['nev'], a['nq'], a['kept_top']) == (b['nev'], b['nq'], b['kept_top'])


def _close(a, b, tol=1e-14):
    return abs(a - b) <= tol*abs(b)


@pytest.mark.gpu
@pytest.mark.parametrize('stop', [1, 7, 20, 25])
def test_device_stop_and_resume(sfx, tmp_path, stop):
    """stopped right after its stop-th checkpoint (blocks of 2^18 children, a checkpoint after every
    block and at the level end: level 1 has 20 blocks, so 20 is the level-end file and 25 one inside
    the last level), the sum resumes from the file to the unbroken result; the file is removed when the
    sum finishes"""
    g, h, N, eps = CASE
    full = sfx.hsum(g, h, N, eps, block_points=1 << 18)
    kw = dict(block_points=1 << 18, ckpt_every=1e-9, ckpt_dir=str(tmp_path))
    with pytest.raises(sfx.SfxStopped):
        sfx.hsum(g, h, N, eps, ckpt_stop_after=stop, **kw)
    f = tmp_path / ('sincqed_%s.ckpt' % g)
    assert f.exists() and f.read_bytes()[:6] == b'SQCK4\0'
    r = sfx.hsum(g, h, N, eps, **kw)
    assert r['resumed'] == 1
    assert _same_counts(r, full), (r['nev'], full['nev'])
    assert _close(r['S'][0], full['S'][0]) and _close(r['S'][1], full['S'][1])
    assert _close(r['es1'], full['es1']) and _close(r['prune'], full['prune'], 1e-12)
    assert r['abskipped'] == full['abskipped']
    assert not f.exists()


@pytest.mark.gpu
def test_other_run_checkpoint_is_ignored(sfx, tmp_path):
    """a checkpoint of another (h, N, eps, top) is not used (checkpoint.c:90-92)"""
    g, h, N, eps = CASE
    kw = dict(block_points=1 << 18, ckpt_every=1e-9, ckpt_dir=str(tmp_path))
    with pytest.raises(sfx.SfxStopped):
        sfx.hsum(g, h, N, eps, ckpt_stop_after=2, **kw)
    r = sfx.hsum(g, h, N, 1e-10, **kw)
    assert r['resumed'] == 0


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(O.SINCQED_REF_NC), reason='oracle/_ref not built')
@pytest.mark.parametrize('stop', [2, 20, 25])
def test_reference_resumes_device_checkpoint(sfx, tmp_path, stop):
    """the reference implementation (built without FP contraction, the device's arithmetic) continues a
    device checkpoint to the result of its own unbroken run"""
    g, h, N, eps = CASE
    ref, _, _ = O.run_reference(h, N, eps, [g], env={'SINCQED_SYM': '0'}, threads=8, nc=True, timeout=300)
    with pytest.raises(sfx.SfxStopped):
        sfx.hsum(g, h, N, eps, block_points=1 << 18, ckpt_every=1e-9, ckpt_dir=str(tmp_path), ckpt_stop_after=stop)
    res, _, out = O.run_reference(h, N, eps, [g], threads=8, nc=True, timeout=300,
                                  env={'SINCQED_SYM': '0', 'SINCQED_CKPT': '600', 'SINCQED_CKPT_DIR': str(tmp_path)})
    assert 'resuming %s from checkpoint' % g in out
    a, b = res[g], ref[g]
    assert (a['nev'], a['nq'], a['kept']) == (b['nev'], b['nq'], b['kept'])
    assert _close(a['S0'], b['S0']) and _close(a['S1'], b['S1'])


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(O.SINCQED_REF_NC), reason='oracle/_ref not built')
def test_device_resumes_reference_checkpoint(sfx, tmp_path):
    """the reference (no contraction, a checkpoint after every block) is killed part way; the device
    continues its file to the result of the unbroken reference run"""
    g, h, N, eps = CASE
    t = time.time()
    ref, _, _ = O.run_reference(h, N, eps, [g], env={'SINCQED_SYM': '0'}, threads=4, nc=True, timeout=300)
    t_full = time.time() - t
    f = tmp_path / ('sincqed_%s.ckpt' % g)
    env = dict(os.environ, SINCQED_SYM='0', SINCQED_CKPT='1e-9', SINCQED_CKPT_DIR=str(tmp_path),
               SINCQED_PROGRESS='0', SINCQED_RESULTS='-', SINCQED_LOG='-', OMP_NUM_THREADS='4')
    for frac in (0.6, 0.8, 0.95, 1.5):
        p = subprocess.Popen([O.SINCQED_REF_NC, 'run', repr(h), str(N), repr(eps), g], env=env,
                             stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        try:
            p.wait(timeout=frac*t_full)
        except subprocess.TimeoutExpired:
            p.kill()
            p.wait()
        if f.exists():
            break
    assert f.exists(), 'the reference wrote no checkpoint before it was stopped'
    r = sfx.hsum(g, h, N, eps, ckpt_every=600, ckpt_dir=str(tmp_path))
    assert r['resumed'] == 1
    b = ref[g]
    assert (r['nev'], r['nq'], r['kept_top']) == (b['nev'], b['nq'], b['kept'])
    assert _close(r['S'][0], b['S0']) and _close(r['S'][1], b['S1'])


@pytest.mark.gpu
@pytest.mark.parametrize('stop', [3, 21])
def test_box_scan_and_abcheck_stop_and_resume(sfx, tmp_path, stop):
    """a sum whose coarsest level came from the full-box scans, with the ABCHECK counter running: stopped
    inside level 1 or inside the last level, it resumes to the unbroken result -- the shadow sums stay off
    (the file's esok) and the skipped and ABMISS counters continue from the file (cnt[1], cnt[2])"""
    g, h, N, eps = CASE
    opt = dict(box_scan=1, abcheck=1, nearf=1.0, block_points=1 << 18)
    full = sfx.hsum(g, h, N, eps, **opt)
    kw = dict(opt, ckpt_every=1e-9, ckpt_dir=str(tmp_path))
    with pytest.raises(sfx.SfxStopped):
        sfx.hsum(g, h, N, eps, ckpt_stop_after=stop, **kw)
    r = sfx.hsum(g, h, N, eps, **kw)
    assert r['resumed'] == 1 and r['esok'] == 0 and full['esok'] == 0
    assert _same_counts(r, full)
    assert (r['abskipped'], r['abmiss']) == (full['abskipped'], full['abmiss'])
    assert _close(r['S'][0], full['S'][0]) and _close(r['S'][1], full['S'][1])
