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
