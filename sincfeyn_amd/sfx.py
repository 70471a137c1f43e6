"""sfx.py -- Python bindings of the C ABI (include/sfx.h) over ctypes.

The library is sincfeyn_amd/_lib/libsfx.so, built by `python -m sincfeyn_amd.build` (or
__graft_entry__.build()).  There is no fallback: if the library is missing or cannot initialise the GPU,
every call raises.  The functions mirror the reference's interface for this path: hsum() is
driver/sum.c's hsum() for one graph (with lattice_setup), run() is `sincqed run h N eps top GRAPH...`
(driver/main.c:205-251) returning one record per graph with the fields of its results line
(main.c:66-73), eval_nodes / eval_cells are the reference GPU back end's per-node primitive
(driver/gpu.h:17-20).
"""
import ctypes as C
import math
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# SFX_LIB: another build of the library (the tuner's trial builds of sincfeyn_amd/tune.py)
LIB_PATH = os.environ.get('SFX_LIB') or os.path.join(PKG, '_lib', 'libsfx.so')


class SfxError(RuntimeError):
    pass


class SfxStopped(SfxError):
    """the sum stopped after a checkpoint (opts ckpt_stop_after): run it again to resume from the file"""


class Opts(C.Structure):
    _fields_ = [('abserr0', C.c_double), ('abrel', C.c_double), ('nearf', C.c_double), ('ddslack', C.c_double),
                ('abskip', C.c_int), ('dropunres', C.c_int), ('keepbt', C.c_int), ('btmargin', C.c_int),
                ('block_points', C.c_long), ('verbose', C.c_int), ('ckpt_every', C.c_double),
                ('ckpt_dir', C.c_char*256), ('ckpt_stop_after', C.c_int), ('qd_fallback', C.c_int),
                ('box_scan', C.c_int), ('abcheck', C.c_int),
                ('growskip', C.c_int), ('graph_capture', C.c_int), ('grow_pipeline', C.c_int)]


class Result(C.Structure):
    _fields_ = [('S', C.c_double * 3), ('Ah', C.c_double), ('ref', C.c_double), ('prune', C.c_double),
                ('trunc', C.c_double), ('unres', C.c_double), ('t', C.c_double), ('nev', C.c_long), ('nq', C.c_long),
                ('kept_top', C.c_long), ('es1', C.c_double), ('es2', C.c_double), ('esok', C.c_int), ('zero', C.c_int),
                ('abskipped', C.c_long), ('unresdrop', C.c_long), ('btuns', C.c_long), ('nev_level', C.c_long * 16),
                ('abserr', C.c_double), ('K', C.c_int), ('top', C.c_int), ('levels', C.c_int),
                ('t_grow', C.c_double), ('t_dilate', C.c_double), ('t_refine', C.c_double),
                ('resumed', C.c_int), ('resume_lev', C.c_int), ('resume_next', C.c_long), ('resume_ncells', C.c_long),
                ('ckpts_written', C.c_int), ('abmiss', C.c_long),
                ('shell_n', C.c_long), ('shell_sum', C.c_double), ('shell_abs', C.c_double), ('grow_retries', C.c_int)]


class GraphInfo(C.Structure):
    _fields_ = [('name', C.c_char * 32), ('volkov', C.c_char * 64), ('L', C.c_int), ('D', C.c_int), ('weight', C.c_int),
                ('vgprs', C.c_int), ('sgprs', C.c_int), ('scratch_bytes', C.c_int), ('occupancy', C.c_int)]


class ScaleResult(C.Structure):
    _fields_ = [('n', C.c_int), ('fixed', C.c_int), ('ndir', C.c_long), ('bad', C.c_int), ('worst', C.c_double),
                ('worstd', C.c_long), ('e_worst', C.c_int*16), ('g_worst', C.c_double*3), ('nslow', C.c_int),
                ('e_slow', (C.c_int*16)*12), ('g_slow', (C.c_double*3)*12)]


class ProfEntry(C.Structure):
    _fields_ = [('graph', C.c_char * 32), ('kind', C.c_int), ('launches', C.c_long), ('ms', C.c_double),
                ('points', C.c_long), ('skipped', C.c_long), ('dd_points', C.c_long)]


_lib = None


def lib():
    """the loaded library (raises if it is not built)"""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SfxError('%s is not built: run python -m sincfeyn_amd.build' % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    L.sfx_init.argtypes = [C.c_int]
    L.sfx_info.restype = C.c_char_p
    L.sfx_last_error.restype = C.c_char_p
    L.sfx_opts_default.argtypes = [C.POINTER(Opts)]
    L.sfx_graph_info.argtypes = [C.c_int, C.POINTER(GraphInfo)]
    L.sfx_graph_find.argtypes = [C.c_char_p]
    L.sfx_graph_selected.argtypes = [C.c_int, C.c_char_p]
    L.sfx_eval_nodes.argtypes = [C.c_int, C.c_double, C.c_int, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.sfx_eval_cells.argtypes = [C.c_int, C.c_double, C.c_int, C.c_long, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    L.sfx_eval_raw.argtypes = [C.c_int, C.c_long, C.c_void_p, C.c_int, C.c_void_p]
    L.sfx_eval_finished.argtypes = [C.c_int, C.c_double, C.c_int, C.c_long, C.c_void_p, C.c_void_p, C.c_double,
                                    C.POINTER(Opts), C.c_void_p, C.c_void_p, C.POINTER(C.c_long)]
    L.sfx_hsum.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, C.POINTER(Opts), C.POINTER(Result)]
    L.sfx_prof_enable.argtypes = [C.c_int]
    L.sfx_prof_get.argtypes = [C.c_int, C.POINTER(ProfEntry)]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        cls = SfxStopped if rc == -6 else SfxError
        raise cls('sfx error %d: %s' % (rc, lib().sfx_last_error().decode()))


_initialised = False


def hw_queues(workers):
    """Give the job engine's workers one hardware queue each (at most 16): the HIP runtime's default of four
    makes the workers' streams share four queues, whose kernels then wait for one another in queue order.
    Past about 20 queues the device's scheduler time-slices them (24 workers on 24 queues ran the batched
    workload 22% slower).  The runtime reads GPU_MAX_HW_QUEUES once, at its first call, so this must come
    before init(); a value already in the environment is kept."""
    os.environ.setdefault('GPU_MAX_HW_QUEUES', str(max(4, min(int(workers), 16))))


def init(device=0):
    global _initialised
    if not _initialised:
        _check(lib().sfx_init(device))
        _initialised = True
    return lib().sfx_info().decode()


def opts(**kw):
    o = Opts()
    lib().sfx_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v.encode() if isinstance(v, str) else v)
    return o


def graphs():
    """[{name, volkov, L, D, weight, ...}] in registry order"""
    out = []
    for i in range(lib().sfx_graph_count()):
        g = GraphInfo()
        _check(lib().sfx_graph_info(i, C.byref(g)))
        out.append(dict(index=i, name=g.name.decode(), volkov=g.volkov.decode(), L=g.L, D=g.D, weight=g.weight,
                        vgprs=g.vgprs, scratch_bytes=g.scratch_bytes, occupancy=g.occupancy))
    return out


def find(name):
    i = lib().sfx_graph_find(name.encode())
    if i < 0:
        raise KeyError(name)
    return i


def select(*selectors):
    """graphs.c:17-29 selection, in registry order without repeats (main.c:156-168)"""
    L = lib()
    sel = []
    for s in selectors:
        for i in range(L.sfx_graph_count()):
            if L.sfx_graph_selected(i, s.encode()) and i not in sel:
                sel.append(i)
    return sel


def default_top(h, topstep=3.0):
    """main.c:211-215: 1, raised to 2 while the coarsest step 2^(top+1) h stays <= 3"""
    top = 1
    while top < 2 and (1 << (top + 1))*h <= topstep*(1 + 1e-9):
        top += 1
    return top


def lattice_K(N, top):
    """lattice.c:16"""
    m = 1 << (top + 1)
    K = (N - 1)//2
    return ((K + m - 1)//m)*m


def eval_raw(graph, sf, prec=0):
    sf = np.ascontiguousarray(sf, dtype=np.float64)
    n = sf.shape[0]
    out = np.zeros((n, 4 if prec else 2))
    _check(lib().sfx_eval_raw(graph, n, sf.ctypes.data, prec, out.ctypes.data))
    return out


def eval_nodes(graph, h, K, k, skip=None):
    k = np.ascontiguousarray(k, dtype=np.int16)
    n = k.shape[0]
    fw = np.zeros(n); abw = np.zeros(n)
    sk = None if skip is None else np.ascontiguousarray(skip, dtype=np.uint8)
    _check(lib().sfx_eval_nodes(graph, h, K, n, k.ctypes.data, None if sk is None else sk.ctypes.data,
                                fw.ctypes.data, abw.ctypes.data))
    return fw, abw


def eval_cells(graph, h, K, cells, st, skip=None):
    cells = np.ascontiguousarray(cells, dtype=np.int16)
    nc, D = cells.shape
    fw = np.zeros(nc << D); abw = np.zeros(nc << D)
    sk = None if skip is None else np.ascontiguousarray(skip, dtype=np.uint8)
    _check(lib().sfx_eval_cells(graph, h, K, nc, cells.ctypes.data, None if sk is None else sk.ctypes.data, st,
                                fw.ctypes.data, abw.ctypes.data))
    return fw, abw


def _slot_sigs():
    L = lib()
    if not getattr(L, '_slot_sigs', False):
        L.sfx_slot_setup.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_int, C.c_double]
        L.sfx_slot_submit.argtypes = [C.c_int, C.c_long, C.c_void_p, C.c_int, C.c_void_p]
        L.sfx_slot_submit_cells.argtypes = [C.c_int, C.c_long, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        L.sfx_slot_wait.argtypes = [C.c_int, C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.POINTER(C.c_double))]
        L._slot_sigs = True
    return L


def slot_setup(name, D, EX, K, H):
    """the reference back end's gpu_setup (driver/gpu.h:17): 1 if the graph has a device kernel, else 0"""
    EX = np.ascontiguousarray(EX, dtype=np.float64)
    rc = _slot_sigs().sfx_slot_setup(name.encode(), D, EX.ctypes.data, K, H)
    if rc < 0:
        _check(rc)
    return rc


def slot_submit(slot, k, skip=None, cells_step=0):
    """gpu_submit (nodes k[n, stride] int32) or, with cells_step > 0, gpu_submit_cells; returns at once.
    The arrays must stay alive until slot_wait (the library copies them before returning, so they may go)"""
    k = np.ascontiguousarray(k, dtype=np.int32)
    n, stride = k.shape
    sk = None if skip is None else np.ascontiguousarray(skip, dtype=np.uint8)
    L = _slot_sigs()
    if cells_step:
        _check(L.sfx_slot_submit_cells(slot, n, k.ctypes.data, stride, None if sk is None else sk.ctypes.data, cells_step))
    else:
        _check(L.sfx_slot_submit(slot, n, k.ctypes.data, stride, None if sk is None else sk.ctypes.data))


def slot_wait(slot, n):
    """gpu_wait: copies of (f w, Sum|terms| w) of the slot's last batch of n points"""
    fw = C.POINTER(C.c_double)(); abw = C.POINTER(C.c_double)()
    _check(_slot_sigs().sfx_slot_wait(slot, C.byref(fw), C.byref(abw)))
    if n == 0:
        return np.zeros(0), np.zeros(0)
    return np.ctypeslib.as_array(fw, (n,)).copy(), np.ctypeslib.as_array(abw, (n,)).copy()


QD_OPS = {'+': 0, '-': 1, '*': 2, '/': 3}


def qd_ops(op, a, b, device=False):
    """element-wise quad-double a op b on (n, 4) arrays of doubles (the all-QD mode's arithmetic)"""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 4)
    b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1, 4)
    out = np.zeros_like(a)
    L = lib()
    L.sfx_qd_ops.argtypes = [C.c_int, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    _check(L.sfx_qd_ops(QD_OPS[op], a.shape[0], a.ctypes.data, b.ctypes.data, out.ctypes.data, 1 if device else 0))
    return out


ALL_QD = dict(abserr0=0.0, abrel=0.0, abskip=0, qd_fallback=1)
ALL_DD = dict(abserr0=0.0, abrel=0.0, abskip=0)


def scale_test(graph, kappa=24.0, seed=1, E=1):
    """`sincqed scale` for one graph (main.c:79-116); None for a VP graph (no point function)"""
    g = find(graph) if isinstance(graph, str) else graph
    r = ScaleResult()
    L = lib()
    L.sfx_scale_test.argtypes = [C.c_int, C.c_double, C.c_uint, C.c_int, C.POINTER(ScaleResult)]
    rc = L.sfx_scale_test(g, kappa, seed, E, C.byref(r))
    if rc == -5:
        return None
    _check(rc)
    n = r.n
    return dict(n=n, fixed=r.fixed, ndir=r.ndir, bad=r.bad, worst=r.worst, worstd=r.worstd,
                e_worst=list(r.e_worst)[:n], g_worst=list(r.g_worst),
                slow=[(list(r.e_slow[k])[:n], list(r.g_slow[k])) for k in range(r.nslow)])


def eval_finished(graph, h, K, k, abserr, skip=None, **kw):
    k = np.ascontiguousarray(k, dtype=np.int16)
    n = k.shape[0]
    v = np.zeros(n); fl = np.zeros(n, dtype=np.uint8); nq = C.c_long()
    sk = None if skip is None else np.ascontiguousarray(skip, dtype=np.uint8)
    o = opts(**kw)
    _check(lib().sfx_eval_finished(graph, h, K, n, k.ctypes.data, None if sk is None else sk.ctypes.data, abserr,
                                   C.byref(o), v.ctypes.data, fl.ctypes.data, C.byref(nq)))
    return v, fl, nq.value


def result_dict(r):
    d = {k: getattr(r, k) for k, _ in r._fields_ if k not in ('S', 'nev_level')}
    d['S'] = list(r.S)
    d['nev_level'] = list(r.nev_level)
    return d


def hsum(graph, h, N, eps, top=None, **kw):
    """hsum() of one graph on the device; graph = name or index"""
    gi = find(graph) if isinstance(graph, str) else graph
    top = default_top(h) if top is None else top
    r = Result()
    o = opts(**kw)
    _check(lib().sfx_hsum(gi, h, N, eps, top, C.byref(o), C.byref(r)))
    return result_dict(r)


def graph_value(info, res):
    """main.c:46-48: value = w S_h, err = w (disc + prune + trunc + round)"""
    w = info['weight']
    S0, S1 = res['S'][0], res['S'][1]
    d2 = abs(S1 - S0) if not math.isnan(S1) else float('nan')
    disc = 2*d2*d2/res['Ah'] if res['Ah'] > 0 else 0.0
    return w*S0, w*(disc + res['prune'] + res['trunc'] + res['unres'])


# ---- job engine (continuous batching) ----
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3


def _engine_sigs():
    L = lib()
    if getattr(L, '_engine_sigs', False):
        return L
    L.sfx_engine_start.argtypes = [C.c_int, C.c_double]
    L.sfx_job_submit.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, C.POINTER(Opts), C.POINTER(C.c_long)]
    L.sfx_job_poll.argtypes = [C.c_long, C.POINTER(C.c_int)]
    L.sfx_job_wait.argtypes = [C.c_long, C.POINTER(Result), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.sfx_job_wait_for.argtypes = [C.c_long, C.c_double, C.POINTER(Result), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.sfx_job_release.argtypes = [C.c_long]
    L.sfx_engine_deadlines.argtypes = [C.c_double, C.c_double]
    L._engine_sigs = True
    return L


def engine_start(workers=0, hbm_budget_bytes=0.0):
    _check(_engine_sigs().sfx_engine_start(workers, hbm_budget_bytes))


def engine_stop():
    _check(_engine_sigs().sfx_engine_stop())


def engine_deadlines(job_wait_s=3600.0, budget_wait_s=600.0):
    """the deadlines of a job wait and of a job's wait for HBM budget (seconds, <= 0: none)"""
    _check(_engine_sigs().sfx_engine_deadlines(job_wait_s, budget_wait_s))


def submit(graph, h, N, eps, top=None, **kw):
    """queue one hierarchical sum; returns the job id at once"""
    L = _engine_sigs()
    gi = find(graph) if isinstance(graph, str) else graph
    top = default_top(h) if top is None else top
    j = C.c_long()
    o = opts(**kw)
    _check(L.sfx_job_submit(gi, h, N, eps, top, C.byref(o), C.byref(j)))
    return j.value


def poll(job):
    s = C.c_int()
    _check(_engine_sigs().sfx_job_poll(job, C.byref(s)))
    return s.value


def wait(job, release=True, timeout=None):
    """(result dict, seconds queued, seconds running) of a job, blocking until it is done -- at most
    `timeout` seconds (None: the engine's wait deadline), then SfxError (SFX_ETIMEOUT) naming the job's
    stage and level; the job keeps running and can be waited for again"""
    L = _engine_sigs()
    r = Result(); tq = C.c_double(); tr = C.c_double()
    if timeout is None:
        rc = L.sfx_job_wait(job, C.byref(r), C.byref(tq), C.byref(tr))
    else:
        rc = L.sfx_job_wait_for(job, float(timeout), C.byref(r), C.byref(tq), C.byref(tr))
    if rc == -7:                                # SFX_ETIMEOUT: the job is still there
        _check(rc)
    if release:
        L.sfx_job_release(job)
    _check(rc)
    return result_dict(r), tq.value, tr.value
