"""CPU tests of the C ABI: the library loads, exports every function include/sfx.h declares, its registry
matches the emitted integrands and the reference's group rules, and errors come back as codes (no GPU is
touched here: every call below fails or returns before a kernel launch)."""
import ctypes as C
import json
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, 'include', 'sfx.h')


def _declared():
    txt = open(HDR).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(sfx_[a-z_0-9]+)\s*\(', txt)))


@pytest.fixture(scope='module')
def S():
    from sincfeyn_amd import sfx
    return sfx


def test_exports_every_declared_symbol(S):
    L = S.lib()
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_registry_matches_emitter(S):
    gen = os.path.join(ROOT, 'sincfeyn_amd', '_gen')
    gs = S.graphs()
    assert len(gs) == 52
    for g in gs:
        meta = json.load(open(os.path.join(gen, 'k_%s.json' % g['name'])))
        assert (meta['D'], meta['L'], meta['weight'], meta['volkov']) == (g['D'], g['L'], g['weight'], g['volkov'])


def test_groups_follow_the_reference_rules(S):
    """graphs.c:17-29: all2 = the 5 two-loop integrals, all3 = the 41 three-loop ones (w3_* excluded),
    q3 = the 28 g3_ ones without light-by-light, vp3 = the 12 v3_, lbl = g3_lbl"""
    name = {g['index']: g['name'] for g in S.graphs()}
    assert [name[i] for i in S.select('all2')] == ['g2_00', 'g2_01', 'g2_02', 'g2_03', 'v2_00']
    a3 = [name[i] for i in S.select('all3')]
    assert len(a3) == 41 and not any(n.startswith('w3_') for n in a3)
    assert len(S.select('q3')) == 28 and len(S.select('vp3')) == 12
    assert [name[i] for i in S.select('lbl')] == ['g3_lbl']
    assert [name[i] for i in S.select('2; 1-4, 3-5')] == ['g2_00']          # by Volkov string
    assert S.select('no_such_graph') == []


def test_errors_are_codes_not_exits(S):
    L = S.lib()
    r = S.Result()
    o = S.opts()
    rc = L.sfx_hsum(0, 1.0, 97, 1e-12, 1, C.byref(o), C.byref(r))
    assert rc in (-4, -2)                       # not initialised (or no device)
    assert L.sfx_graph_find(b'no_such_graph') == -1
    assert L.sfx_graph_info(999, C.byref(S.GraphInfo())) == -1
    assert L.sfx_prof_get(12345, C.byref(S.ProfEntry())) == -1
    assert L.sfx_last_error()


def test_lattice_rules(S):
    """lattice.c:16 and main.c:211-215"""
    assert S.lattice_K(97, 1) == 48 and S.lattice_K(139, 2) == 72 and S.lattice_K(121, 1) == 60
    assert S.default_top(1.0) == 1 and S.default_top(0.7) == 2 and S.default_top(0.75) == 2 and S.default_top(0.8) == 1


def test_cli_list_and_clean_failure_without_gpu():
    cli = os.path.join(ROOT, 'sincfeyn_amd', '_lib', 'sincqed')
    p = subprocess.run([cli, 'list'], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and p.stdout.startswith('52 graphs')
    assert 'g3_lbl' in p.stdout
    p = subprocess.run([cli, 'bogus'], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1


def test_run_option_defaults(S):
    """sfx_opts_default: the reference's environment defaults (main.c:176-185, sum.c:558-592) and the
    library's own (the pipelined grow steps on, graph capture off)"""
    o = S.opts()
    assert (o.abserr0, o.abrel, o.nearf, o.ddslack) == (1e-15, 1e-3, 1e-3, 1.0)
    assert (o.abskip, o.dropunres, o.keepbt, o.btmargin, o.growskip) == (1, 1, 0, 2, 2)
    assert (o.abcheck, o.box_scan, o.qd_fallback, o.graph_capture) == (0, 0, 0, 0)
    assert o.grow_pipeline == 1
