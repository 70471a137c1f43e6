"""oracle.py -- ctypes bindings of the CPU oracle (TEST INFRASTRUCTURE ONLY: the checker used by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg; the product never imports it).

  Oracle()                 liboracle.so (the C restatement of the reference's hsum, oracle/sq_oracle.c)
                           on top of oracle/_ref/libsqk_ref.so (the reference's generated kernels)
  run_reference(...)       oracle/_ref/sincqed_ref: the reference implementation itself, parsed
"""
import ctypes as C
import os
import re
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, '_ref')
LIB_ORACLE = os.path.join(HERE, '_build', 'liboracle.so')
LIB_KERNELS = os.path.join(REF_DIR, 'libsqk_ref.so')
LIB_KERNELS_NC = os.path.join(REF_DIR, 'libsqk_nc.so')
LIB_KERNELS_F128 = os.path.join(REF_DIR, 'libsqk_f128.so')
SINCQED_REF = os.path.join(REF_DIR, 'sincqed_ref')
SINCQED_REF_NC = os.path.join(REF_DIR, 'sincqed_ref_nc')
SINCQED_REF_GPU = os.path.join(REF_DIR, 'sincqed_ref_gpu')   # the reference driver's GPU build over libsfx


class OrcOpts(C.Structure):
    _fields_ = [('abserr0', C.c_double), ('abrel', C.c_double), ('nearf', C.c_double), ('ddslack', C.c_double),
                ('abskip', C.c_int), ('dropunres', C.c_int), ('keepbt', C.c_int), ('btmargin', C.c_int),
                ('maxpts', C.c_long), ('grow', C.c_int), ('abcheck', C.c_int),
                ('growskip', C.c_int)]


class OrcRes(C.Structure):
    _fields_ = [('S', C.c_double * 3), ('Ah', C.c_double), ('ref', C.c_double), ('prune', C.c_double),
                ('trunc', C.c_double), ('unres', C.c_double), ('nev', C.c_long), ('nq', C.c_long),
                ('kept_top', C.c_long), ('es1', C.c_double), ('es2', C.c_double), ('abskipped', C.c_long),
                ('unresdrop', C.c_long), ('btuns', C.c_long), ('nev_level', C.c_long * 16), ('zero', C.c_int),
                ('K', C.c_int), ('top', C.c_int), ('abserr', C.c_double), ('abmiss', C.c_long)]


class OrcGraph(C.Structure):
    _fields_ = [('volkov', C.c_char * 64), ('L', C.c_int), ('D', C.c_int), ('weight', C.c_int), ('has_fv', C.c_int)]


class OrcScale(C.Structure):
    _fields_ = [('n', C.c_int), ('fixed', C.c_int), ('ndir', C.c_long), ('bad', C.c_int), ('worst', C.c_double),
                ('worstd', C.c_long), ('e_worst', C.c_int*16), ('g_worst', C.c_double*3), ('nslow', C.c_int),
                ('e_slow', (C.c_int*16)*12), ('g_slow', (C.c_double*3)*12)]


def scale_dict(r):
    n = r.n
    return dict(n=n, fixed=r.fixed, ndir=r.ndir, bad=r.bad, worst=r.worst, worstd=r.worstd,
                e_worst=list(r.e_worst)[:n], g_worst=list(r.g_worst),
                slow=[(list(r.e_slow[k])[:n], list(r.g_slow[k])) for k in range(r.nslow)])


def available():
    return os.path.exists(LIB_ORACLE) and os.path.exists(LIB_KERNELS)


class Oracle:
    """nc=False: the kernels as the reference's default build compiles them (FMA contraction on, the CPU
    path the reference reports); nc=True: the same kernels and driver semantics without contraction (the
    floating-point semantics of the reference's GPU path, Makefile:188-191), which the device reproduces
    point by point."""

    def __init__(self, nc=False):
        self.kset = (LIB_KERNELS_NC, b'Q_') if nc else (LIB_KERNELS, b'G_')
        self.lib = C.CDLL(LIB_ORACLE)
        L = self.lib
        L.orc_open.argtypes = [C.c_char_p, C.c_char_p]
        L.orc_graph.argtypes = [C.c_char_p, C.POINTER(OrcGraph)]
        L.orc_eval_fd.argtypes = [C.c_char_p, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_eval_fq.argtypes = [C.c_char_p, C.c_long, C.c_void_p, C.c_void_p]
        L.orc_eval_fq128.argtypes = [C.c_char_p, C.c_long, C.c_void_p, C.c_void_p]
        L.orc_hsum.argtypes = [C.c_char_p, C.c_double, C.c_int, C.c_double, C.c_int, C.POINTER(OrcOpts), C.POINTER(OrcRes)]
        L.orc_opts_default.argtypes = [C.POINTER(OrcOpts)]
        L.orc_scale.argtypes = [C.c_char_p, C.c_double, C.c_uint, C.c_int, C.POINTER(OrcScale)]
        L.orc_error.restype = C.c_char_p
        self._select()

    def _select(self):
        if self.lib.orc_open(self.kset[0].encode(), self.kset[1]) != 0:
            raise RuntimeError(self.lib.orc_error().decode())

    def opts(self, **kw):
        o = OrcOpts()
        self.lib.orc_opts_default(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def graph(self, name):
        self._select()
        g = OrcGraph()
        if self.lib.orc_graph(name.encode(), C.byref(g)) != 0:
            raise KeyError(name)
        return dict(volkov=g.volkov.decode(), L=g.L, D=g.D, weight=g.weight, has_fv=g.has_fv)

    def eval_fd(self, name, sf):
        self._select()
        import numpy as np
        sf = np.ascontiguousarray(sf, dtype=np.float64)
        n = sf.shape[0]
        f = np.zeros(n); ab = np.zeros(n)
        if self.lib.orc_eval_fd(name.encode(), n, sf.ctypes.data, f.ctypes.data, ab.ctypes.data):
            raise RuntimeError(self.lib.orc_error().decode())
        return f, ab

    def eval_fq(self, name, sf):
        self._select()
        import numpy as np
        sf = np.ascontiguousarray(sf, dtype=np.float64)
        out = np.zeros((sf.shape[0], 4))
        if self.lib.orc_eval_fq(name.encode(), sf.shape[0], sf.ctypes.data, out.ctypes.data):
            raise RuntimeError(self.lib.orc_error().decode())
        return out

    def eval_fq128(self, name, sf):
        """the two-loop kernels built with QREAL = __float128: (n, 4) = f hi, lo, Sum|terms| hi, lo"""
        import numpy as np
        if self.lib.orc_open(LIB_KERNELS_F128.encode(), b'F_') != 0:
            raise RuntimeError(self.lib.orc_error().decode())
        try:
            sf = np.ascontiguousarray(sf, dtype=np.float64)
            out = np.zeros((sf.shape[0], 4))
            if self.lib.orc_eval_fq128(name.encode(), sf.shape[0], sf.ctypes.data, out.ctypes.data):
                raise RuntimeError(self.lib.orc_error().decode())
            return out
        finally:
            self._select()

    def hsum(self, name, h, N, eps, top, **kw):
        self._select()
        r = OrcRes()
        o = self.opts(**kw)
        if self.lib.orc_hsum(name.encode(), h, N, eps, top, C.byref(o), C.byref(r)) != 0:
            raise RuntimeError(self.lib.orc_error().decode())
        return res_dict(r)


    def scale(self, name, kappa=24.0, seed=1, E=1):
        """the reference's scale test (main.c:79-116) on the selected kernel set; None for VP graphs"""
        self._select()
        r = OrcScale()
        rc = self.lib.orc_scale(name.encode(), kappa, seed, E, C.byref(r))
        if rc == -2:
            return None
        if rc:
            raise RuntimeError(self.lib.orc_error().decode())
        return scale_dict(r)


def res_dict(r):
    d = {k: getattr(r, k) for k, _ in r._fields_ if k not in ('S', 'nev_level')}
    d['S'] = list(r.S)
    d['nev_level'] = list(r.nev_level)
    return d


def default_top(h, topstep=3.0):
    """main.c:211-215"""
    top = 1
    while top < 2 and (1 << (top + 1))*h <= topstep*(1 + 1e-9):
        top += 1
    return top


_RLINE = re.compile(r'^#R (\S+) (.*)$')


def run_reference(h, N, eps, graphs, top=None, env=None, threads=None, timeout=600, nc=False, gpu=False):
    """the reference implementation (oracle/_ref/sincqed_ref, or sincqed_ref_nc built without FP
    contraction, or with gpu sincqed_ref_gpu: that driver's GPU build over this library's gpu.h back end)
    -> ({graph: {field: value}}, wall seconds, stdout)"""
    e = dict(os.environ)
    e.update({'SINCQED_CKPT': '0', 'SINCQED_PROGRESS': '0', 'SINCQED_RESULTS': '-'})
    if threads:
        e['OMP_NUM_THREADS'] = str(threads)
    if env:
        e.update(env)
    args = [SINCQED_REF_GPU if gpu else SINCQED_REF_NC if nc else SINCQED_REF, 'run', repr(float(h)), str(N), repr(float(eps))]
    if top is not None:
        args.append(str(top))
    args += list(graphs)
    p = subprocess.run(args, env=e, capture_output=True, text=True, timeout=timeout)
    if p.returncode != 0:
        raise RuntimeError('sincqed_ref failed: %s' % p.stderr[-2000:])
    out, wall = {}, None
    for line in p.stdout.splitlines():
        m = _RLINE.match(line)
        if m:
            toks = m.group(2).split()
            d = {}
            for k, v in zip(toks[0::2], toks[1::2]):
                d[k] = float(v) if any(c in v for c in '.en') else int(v)
            out[m.group(1)] = d
        elif line.startswith('# wall'):
            wall = float(line.split()[2])
    return out, wall, p.stdout


class _Graph(C.Structure):
    """graph_t (reference driver/common.h:38)"""
    _fields_ = [('name', C.c_char_p), ('volkov', C.c_char_p), ('L', C.c_int), ('D', C.c_int), ('weight', C.c_int),
                ('n', C.c_int), ('fixed', C.c_int), ('fd', C.c_void_p), ('fq', C.c_void_p), ('ptd', C.c_void_p),
                ('ptq', C.c_void_p), ('fv', C.c_void_p)]


class _DD(C.Structure):
    _fields_ = [('hi', C.c_double), ('lo', C.c_double)]


_FD = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double))
_FQ = C.CFUNCTYPE(_DD, C.POINTER(_DD), C.POINTER(_DD))


class NcKernels:
    """the reference kernels compiled with FP contraction off (oracle/_ref/libsqk_nc.so): G_<g>.fq is the
    QREAL half evaluated in plain FP64 (the arithmetic of the device's FP64 path), Q_<g>.fq the QREAL half
    in the reference's double-double (dd.h) -- the arithmetic of the device's fallback path."""

    def __init__(self, path=LIB_KERNELS_NC):
        self.lib = C.CDLL(path)

    def fp64(self, name, sf):
        import numpy as np
        g = _Graph.in_dll(self.lib, 'G_' + name)
        fq = _FD(g.fq)
        sf = np.ascontiguousarray(sf, dtype=np.float64)
        f = np.zeros(sf.shape[0]); ab = np.zeros(sf.shape[0]); a = C.c_double()
        P = C.POINTER(C.c_double)
        for i in range(sf.shape[0]):
            f[i] = fq(sf[i].ctypes.data_as(P), C.byref(a))
            ab[i] = a.value
        return f, ab

    def dd(self, name, sf):
        import numpy as np
        g = _Graph.in_dll(self.lib, 'Q_' + name)
        fq = _FQ(g.fq)
        sf = np.ascontiguousarray(sf, dtype=np.float64)
        D = sf.shape[1]
        out = np.zeros((sf.shape[0], 4)); arr = (_DD*D)(); ab = _DD()
        for i in range(sf.shape[0]):
            for d in range(D):
                arr[d].hi = float(sf[i, d]); arr[d].lo = 0.0
            v = fq(arr, C.byref(ab))
            out[i] = (v.hi, v.lo, ab.hi, ab.lo)
        return out
