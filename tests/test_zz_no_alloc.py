"""The entry point that allocates nothing (sfx_eval_finished_noalloc): on device-resident inputs, through
the C ABI, its values, flags and fallback count equal sfx_eval_finished's on the same points -- on two
calls in a row, the second with fewer points (the shared working array keeps the first call's contents),
and with a size past its maximum refused."""
import ctypes as C
import math

import numpy as np
import pytest


class Dev:
    """device buffers through the HIP runtime the library itself links (libamdhip64.so.7), so that no
    second runtime is loaded into the process"""
    H = None

    def __init__(self, arr):
        if Dev.H is None:
            H = C.CDLL('libamdhip64.so.7')
            H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            H.hipFree.argtypes = [C.c_void_p]
            Dev.H = H
        self.host = np.ascontiguousarray(arr)
        self.p = C.c_void_p()
        assert Dev.H.hipMalloc(C.byref(self.p), max(self.host.nbytes, 8)) == 0
        assert Dev.H.hipMemcpy(self.p, self.host.ctypes.data, self.host.nbytes, 1) == 0      # host -> device

    def get(self):
        out = np.empty_like(self.host)
        assert Dev.H.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0                 # device -> host
        return out

    def __del__(self):
        if self.p:
            Dev.H.hipFree(self.p)


def _noalloc(sfx):
    L = sfx.lib()
    f = L.sfx_eval_finished_noalloc
    f.argtypes = [C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_long, C.c_void_p, C.c_void_p, C.c_double,
                  C.POINTER(sfx.Opts), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return f


def _points(rng, n, D, K, h):
    k = rng.integers(-K, K + 1, size=(n, D)).astype(np.int16)
    skip = (rng.random(n) < 0.4).astype(np.uint8)
    return k, skip


@pytest.mark.gpu
@pytest.mark.parametrize('name', ['g2_01', 'g3_07'])
def test_noalloc_matches_eval_finished(sfx, name):
    gi = sfx.find(name)
    D = sfx.graphs()[gi]['D']
    h, K = 1.0, 48
    rng = np.random.default_rng(11)
    f = _noalloc(sfx)
    EX = Dev(np.array([math.exp(k*h) for k in range(-K, K + 1)]))
    o = sfx.opts()
    # an ABSERR that sends a share of the points to the double-double fallback
    abserr = 1e-13
    for n in (200000, 37000):
        k, skip = _points(rng, n, D, K, h)
        v_ref, fl_ref, nq_ref = sfx.eval_finished(gi, h, K, k, abserr, skip=skip)
        dk, ds = Dev(k), Dev(skip)
        dv = Dev(np.full(n, np.nan)); dfl = Dev(np.full(n, 255, dtype=np.uint8)); dnq = Dev(np.zeros(1, dtype=np.uint32))
        sfx._check(f(gi, K, h, EX.p, n, dk.p, ds.p, abserr, C.byref(o), dv.p, dfl.p, dnq.p, None))
        sfx._check(sfx.lib().sfx_sync())
        v, fl, nq = dv.get(), dfl.get(), int(dnq.get()[0])
        print(name, n, 'fallback points', nq, nq_ref, flush=True)
        assert nq == nq_ref and nq > 0
        assert np.array_equal(v.view(np.int64), v_ref.view(np.int64))
        assert np.array_equal(fl, fl_ref)


@pytest.mark.gpu
def test_noalloc_refuses_past_its_maximum(sfx):
    f = _noalloc(sfx)
    n = sfx.lib().sfx_graph_count() and (1 << 18) + 1
    rc = f(sfx.find('g2_01'), 48, 1.0, 1, n, 1, None, 1e-13, None, 1, 1, 1, None)
    assert rc == -1 and b'at most 262144' in sfx.lib().sfx_last_error()


@pytest.mark.gpu
def test_noalloc_marks_points_outside_the_box(sfx):
    gi = sfx.find('g2_01')
    h, K = 1.0, 48
    f = _noalloc(sfx)
    EX = Dev(np.array([math.exp(k*h) for k in range(-K, K + 1)]))
    dk = Dev(np.array([[0, 0, 0], [K + 1, 0, 0], [1, -2, 3]], dtype=np.int16))
    dv = Dev(np.zeros(3)); dfl = Dev(np.zeros(3, dtype=np.uint8)); dnq = Dev(np.zeros(1, dtype=np.uint32))
    sfx._check(f(gi, K, h, EX.p, 3, dk.p, None, 1e-13, None, dv.p, dfl.p, dnq.p, None))
    sfx._check(sfx.lib().sfx_sync())
    fl, v = dfl.get(), dv.get()
    assert fl[1] == 128 and v[1] == 0 and fl[0] != 128 and v[0] != 0
