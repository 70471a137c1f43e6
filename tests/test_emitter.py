"""CPU tests of the HIP emitter (sincfeyn_amd/emit/hipgen.py) and of its operation counts."""
import json
import os

import pytest

from sincfeyn_amd.emit import hipgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SYNTH = '''#include "common.h"
static const double GLX_t[1]={0.5}, GLW_t[1]={1.0};
#define REAL double
#define FN(x) x##_d
static REAL FN(pt_t)(const REAL *x, const REAL Vx, REAL *ab){ *ab=x[0]; return x[0]*x[1]; }
#undef REAL
#undef FN
#define REAL QREAL
#define FN(x) x##_q
static void FN(bk_t_0)(const REAL *F, REAL *T){ T[0]=F[0]+F[1]; }
static REAL FN(pt_t)(const REAL *x, const REAL Vx, REAL *ab){ REAL F_[2]={x[0],x[1]}, T_[1]; FN(bk_t_0)(F_, T_);
  if(SQ_AB) *ab=RABS(T_[0]); return T_[0]*x[1]/(1+Vx); }
static REAL FN(f_t)(const REAL *sf, REAL *ab){ REAL s[2]; s[0]=sf[0]; s[1]=sf[1]; REAL a;
  REAL v=FN(pt_t)(s,0,&a); *ab=a; return v; }
#undef REAL
#undef FN
const graph_t G_t={"t","1; 1-2",2,2,1,3,1,f_t_d,f_t_q,pt_t_d,pt_t_q};
'''


def test_transform_synthetic(tmp_path):
    p = tmp_path / 't.c'
    p.write_text(SYNTH)
    out, meta = hipgen.emit(str(p), str(tmp_path))
    inc = (tmp_path / 'k_t.inc').read_text()
    hip = (tmp_path / 'k_t.hip').read_text()
    assert meta['name'] == 't' and meta['D'] == 2 and meta['L'] == 2 and meta['volkov'] == '1; 1-2'
    # a chunk: body bk_t_0_b, out-of-line copy bk_t_0_c, dispatcher bk_t_0 (inline for FP64)
    assert 'template<class REAL> SFX_FN void bk_t_0_b(const REAL *__restrict__ F, REAL *__restrict__ T){' in inc
    assert 'template<class REAL> SFX_CHUNK void bk_t_0_c(const REAL *__restrict__ F, REAL *__restrict__ T)' \
           '{ bk_t_0_b<REAL>(F, T); }' in inc
    assert 'if constexpr(SFX_CHUNK_INLINE(REAL)) bk_t_0_b<REAL>(F, T); else bk_t_0_c<REAL>(F, T); }' in inc
    assert 'template<class REAL> SFX_FN REAL pt_t(const REAL *x, const REAL Vx, REAL *ab, const int SQ_AB)' in inc
    assert 'bk_t_0<REAL>(F_, T_)' in inc                       # chunk call: no SQ_AB argument
    assert 'pt_t<REAL>(s,0,&a, SQ_AB)' in inc                  # SQ_AB passed on
    assert 'static __constant__ const double GLX_t[1] = {0.5}' in inc
    assert 'x[0]*x[1]' not in inc                              # only the QREAL half is taken
    assert 'SFX_GRAPH_ENTRY(t, sfxk_t::Graph, "1; 1-2", 2, 2, 1)' in hip


def test_rejects_malformed():
    with pytest.raises(hipgen.EmitError):
        hipgen.parse_kernel('no descriptor here')
    with pytest.raises(hipgen.EmitError):
        hipgen.parse_kernel(SYNTH.replace('#define REAL QREAL', '#define REAL QQ'))


def test_op_counts_of_emitted_integrands():
    """the per-point FP64 operation counts the build measures on the emitted templates (emit/opcount.cpp);
    for kernels without value branches they are exact integers, and they agree with an independent count
    made by compiling the reference's QREAL half against an op-counting type: g2_00 361/308, g3_00
    2137/1562, g3_lbl 25307/15082 (SQ_AB = 1 / 0), 11/19/25 divisions"""
    d = json.load(open(os.path.join(ROOT, 'sincfeyn_amd', '_gen', 'opcounts.json')))
    assert len(d) == 52
    for g, (o1, o0, dv) in {'g2_00': (361, 308, 11), 'g3_00': (2137, 1562, 19), 'g3_lbl': (25307, 15082, 25),
                            'v2_00': (126, 118, 9), 'v3_01': (723, 617, 18)}.items():
        assert (d[g]['ops1'], d[g]['ops0'], d[g]['div1']) == (o1, o0, dv), g
    for g, c in d.items():
        assert c['ops0'] <= c['ops1'] and c['ops1'] > 100, g


# a host build of an emitted integrand: REAL = double takes the node-split form (SFX_NODE_SPLIT), a
# wrapper type W with the same IEEE double operations takes the straight-line form
_HOST_DRIVER = r'''
#include <cmath>
#include <type_traits>
#define __constant__
#define SFX_FN static inline
#define SFX_CHUNK static
#define SFX_CHUNK_INLINE(R) true
#define SFX_NODE_SPLIT(R) (std::is_same<R, double>::value)
static inline void sfx_opaque(double &) {}
static inline void sfx_opaque_s(int &) {}
struct W { double v; W() = default; W(double x) : v(x) {} };
static inline W operator+(W a, W b){ return W(a.v + b.v); }
static inline W operator-(W a, W b){ return W(a.v - b.v); }
static inline W operator*(W a, W b){ return W(a.v * b.v); }
static inline W operator/(W a, W b){ return W(a.v / b.v); }
static inline W operator-(W a){ return W(-a.v); }
static inline W &operator+=(W &a, W b){ a.v += b.v; return a; }
static inline W &operator-=(W &a, W b){ a.v -= b.v; return a; }
static inline W &operator*=(W &a, W b){ a.v *= b.v; return a; }
static inline W &operator/=(W &a, W b){ a.v /= b.v; return a; }
static inline double sfx_abs(double x){ return std::fabs(x); }
static inline W sfx_abs(W x){ return W(std::fabs(x.v)); }
namespace k {
#include KINC
}
extern "C" void run(int n, int D, const double *sf, int sqab, double *out){
    for(int i = 0; i < n; i++){
        double a1 = 0; double v1 = k::KF<double>(sf + (long)i*D, &a1, sqab);
        W w[16]; for(int d = 0; d < D; d++) w[d] = W(sf[(long)i*D + d]);
        W a2(0.0); W v2 = k::KF<W>(w, &a2, sqab);
        out[4*i] = v1; out[4*i + 1] = a1; out[4*i + 2] = v2.v; out[4*i + 3] = a2.v;
    }
}
'''


def test_node_split_bit_identical(tmp_path):
    """g3_lbl's point function, emitted a second time as one switch case per forest node for the FP64
    kernels, computes bit for bit what its straight-line form computes (host build, g++ without FP
    contraction, 300 random points, SQ_AB = 0 and 1), and both equal the reference kernel compiled
    without contraction (oracle/_ref/libsqk_nc.so) -- the device point tests check the same on the GPU"""
    import ctypes as C
    import shutil
    import subprocess
    import numpy as np
    from sincfeyn_amd import build
    inc = os.path.join(ROOT, 'sincfeyn_amd', '_gen', 'k_g3_lbl.inc')
    src = os.path.join(build.KERNELS, 'g3', 'g3_lbl.c')
    if os.path.exists(src):
        _, meta = hipgen.emit(src, str(tmp_path))
        assert meta['node_split'] == 7              # six forest nodes of their own, the six small ones together
        inc = str(tmp_path / 'k_g3_lbl.inc')
    if not os.path.exists(inc) or not shutil.which('g++'):
        pytest.skip('no emitted g3_lbl integrand or no g++')
    text = open(inc).read()
    assert 'if constexpr(!SFX_NODE_SPLIT(REAL)){' in text and text.count('  case ') == 7
    drv = tmp_path / 'drv.cpp'
    drv.write_text(_HOST_DRIVER)
    so = tmp_path / 'drv.so'
    subprocess.run(['g++', '-O1', '-ffp-contract=off', '-std=c++17', '-fPIC', '-shared', '-DKINC="%s"' % inc,
                    '-DKF=f_g3_lbl', str(drv), '-o', str(so)], check=True)
    L = C.CDLL(str(so))
    rng = np.random.default_rng(7)
    n, D = 300, 7
    sf = np.exp(rng.uniform(-3, 3, size=(n, D)))
    outs = {}
    for sqab in (0, 1):
        out = np.zeros((n, 4))
        L.run(n, D, sf.ctypes.data_as(C.c_void_p), sqab, out.ctypes.data_as(C.c_void_p))
        assert np.array_equal(out[:, 0].view(np.uint64), out[:, 2].view(np.uint64)), sqab
        assert np.array_equal(out[:, 1].view(np.uint64), out[:, 3].view(np.uint64)), sqab
        outs[sqab] = out
    import oracle as O
    if os.path.exists(O.LIB_KERNELS_NC):
        f, ab = O.NcKernels().fp64('g3_lbl', sf)
        assert np.array_equal(f.view(np.uint64), outs[1][:, 0].view(np.uint64))
        assert np.array_equal(ab.view(np.uint64), outs[1][:, 1].view(np.uint64))


def test_opcount_builds_on_node_split_integrand(tmp_path):
    """the host op-count tool compiles every emitted integrand, the node-split g3_lbl included (it counts
    the straight-line form: the source's 15,082 / 25,307 operations at SQ_AB = 0 / 1)"""
    import shutil
    import subprocess
    gen = os.path.join(ROOT, 'sincfeyn_amd', '_gen')
    if not os.path.exists(os.path.join(gen, 'k_g3_lbl.inc')) or not shutil.which('g++'):
        pytest.skip('no emitted g3_lbl integrand or no g++')
    exe = str(tmp_path / 'opc')
    subprocess.run(['g++', '-O1', '-I' + gen, '-DSFX_INC="k_g3_lbl.inc"', '-DSFX_F=f_g3_lbl', '-DSFX_D=7',
                    os.path.join(ROOT, 'sincfeyn_amd', 'emit', 'opcount.cpp'), '-o', exe], check=True)
    d = json.loads(subprocess.run([exe], capture_output=True, text=True, check=True).stdout)
    assert (d['ops0'], d['ops1'], d['div0']) == (15082, 25307, 25)
