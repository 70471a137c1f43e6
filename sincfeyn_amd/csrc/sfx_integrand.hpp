// sfx_integrand.hpp -- included by every emitted integrand translation unit (k_<graph>.hip).
//
// It instantiates, for one graph, the device kernels that evaluate the integrand at lattice points:
//   k_eval<G, MODE_NODES, 2>, k_eval_list<G, AB>  one thread per point: s = e^{k h} from the job's table,
//                     the weight w = prod(s_i h) in index order (reference driver/lattice.c:30-34), f in FP64,
//                     and the reference's per-point tests (driver/eval.c:33-53): the value f w (0 if not
//                     finite), the "near" flag and the fallback flag.  k_eval_list: the points are the
//                     children b = 0 .. 2^D-1 of listed cells c, at k = cell - st b (reference
//                     sum.c:405-408); a child below the box is not a lattice node (flag VALID clear, value 0).
//   k_dd<G, CELLS>    one thread per flagged point (an index list): the same point in double-double,
//                     then the reference's qvalue (eval.c:25-30): value, eps^2 term of the round column,
//                     and the "unresolved" mark.
//   k_raw<G>          golden primitive: f and Sum|terms| at nodes, in FP64 (and in DD), without the tests.
// The refinement levels run each block of cells as two launches: the cells whose children skip the |.|
// sums, with the integrand compiled for SQ_AB = 0 (those sums are dead code, far fewer registers), and
// the others with SQ_AB = 1.  Node lists choose per wave.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>
#include "sfx_dd.hpp"
#include "sfx_qd.hpp"
#include "sfx_abi_types.h"

#define SFX_FN __device__ __forceinline__
#ifndef SFX_CHUNK
#define SFX_CHUNK __device__ __noinline__     // the out-of-line forest-block functions of the largest integrands
#endif
// forest blocks are inlined into the FP64 kernels (their inputs and outputs then stay in registers instead
// of a scratch call frame) and called out of line from the double-double kernels (compile time)
#ifndef SFX_CHUNK_INLINE
#define SFX_CHUNK_INLINE(R) (sizeof(R) == sizeof(double))
#endif
// Point functions whose forest nodes are all leaves are emitted a second time for the FP64 kernels as a
// loop over the nodes with one switch case per node (or per run of small nodes), the nodes' inputs passed
// through an empty asm so the compiler neither hoists their common subexpressions out of the cases nor
// interleaves the cases: the register need is then one forest block's, not the sum of all of them held live
// together (g3_lbl: 244 VGPRs, two waves per SIMD, instead of 256 + 212 AGPRs at one).  The same + - * /
// run in the same order: the values are bit-identical to the straight-line form.
#define SFX_NODE_SPLIT(R) (sizeof(R) == sizeof(double))
__device__ __forceinline__ void sfx_opaque(double &v){ asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void sfx_opaque_s(int &v){ asm volatile("" : "+s"(v)); }

#ifndef SFX_BLOCK
#define SFX_BLOCK 256
#endif
#define SFX_QEPS 4.93e-32          /* relative precision of the double-double type (reference common.h:7) */

namespace sfx {

// coordinates of point i: node mode reads record i; cell mode derives child b of cell c
template<int D, bool CELLS>
__device__ __forceinline__ bool point_coords(const sfx_cell_t *__restrict__ pts, long i, int st, int K, int *k, unsigned *cfl){
    if(CELLS){
        long c = i >> D; int b = (int)(i & ((1 << D) - 1));
        const sfx_cell_t cell = pts[c];
        *cfl = cell.fl;
        bool ok = true;
#pragma unroll
        for(int d = 0; d < D; d++){ int kd = (int)cell.k[d] - (((b >> d) & 1) ? st : 0); k[d] = kd; ok = ok && kd >= -K; }
        return ok;
    } else {
        const sfx_cell_t p = pts[i];
        *cfl = p.fl;
#pragma unroll
        for(int d = 0; d < D; d++) k[d] = p.k[d];
        return true;
    }
}

template<int D>
__device__ __forceinline__ double lattice_node(const double *__restrict__ EX, int K, double H, const int *k, double *sf){
    double w = 1;
#pragma unroll
    for(int d = 0; d < D; d++){ double s = EX[k[d] + K]; sf[d] = s; w *= s*H; }
    return w;
}

// MODE_NODES: point i is record i, of the skip class when its flag byte is nonzero (SFX_FL_SKIPCLS in
// the output flags); with st > 0 the flag byte also carries the node's outer-shell class at step st:
// SFX_FL_SHELL1 for max_d |k_d| = K, SFX_FL_SHELL2 for K - st.  MODE_LIST: the children of the cells of a list: point i is child
// b = i mod 2^D of cell c = list[i >> D] of pts, written to slot (c << D) | b -- so a block's cells can be
// run as two launches (cells whose children skip the |.| sums, and the others) into one value array.
// With ddidx, the slots flagged for the high-precision recompute are appended to ddidx[0 .. *ddcnt), in
// no particular order.
// AB: 0 every point skips the |.| sums (the integrand is compiled with SQ_AB = 0: those sums are dead
// code and the register pressure drops), 1 every point computes them, 2 chosen per wave.
constexpr int MODE_NODES = 0, MODE_LIST = 1;

// the integrand at one point and the reference's per-point tests: value, flag byte, and the slot appended to
// the double-double list when flagged (every lane of the wave comes here: the append is wave-wide)
template<class G, int MODE, int AB>
__device__ __forceinline__ void point_out(bool in, bool valid, bool cls, const double *sf, double w, const int *k, int st, long o,
                                          const sfx_eval_params_t &P, double *__restrict__ vout, uint8_t *__restrict__ flout,
                                          uint32_t *__restrict__ ddidx, uint32_t *__restrict__ ddcnt){
    constexpr int D = G::D;
    bool skip = cls && !P.abcheck;
    double v = 0, rab = 0; unsigned fl = 0;
    if(valid){
        double ab = 0, f;
        if(AB == 0) f = G::template f<double>(sf, &ab, 0);
        else if(AB == 1) f = G::template f<double>(sf, &ab, 1);
        else {
            // the |.| sums are skipped only when every lane of the wave may skip them
            if(__all(skip)) f = G::template f<double>(sf, &ab, 0);
            else f = G::template f<double>(sf, &ab, 1);
        }
        v = f*w; fl = SFX_FL_VALID | (MODE == MODE_NODES && cls ? SFX_FL_SKIPCLS : 0u);
        if(P.raw_ab){
            // the reference GPU kernel's outputs, untested: f w, and Sum|terms| w unless skipped
            rab = skip ? 0.0 : ab*w; fl = SFX_FL_VALID;
        } else {
            if(MODE == MODE_NODES && st > 0){
                // the node's outer-shell class at step st (the grow's and the box scan's e1/e2 weights)
                int mx = 0;
#pragma unroll
                for(int d = 0; d < D; d++){ int a = k[d] < 0 ? -k[d] : k[d]; mx = a > mx ? a : mx; }
                fl |= mx == P.K ? SFX_FL_SHELL1 : mx == P.K - st ? SFX_FL_SHELL2 : 0u;
            }
            if(skip){ fl |= SFX_FL_SKIP; if(!isfinite(v)) v = 0; }
            else {
                double bnd = ab*w*16*DBL_EPSILON/P.ddslack;            // eval.c:36, same association
                if(!(bnd <= P.nearf*P.abserr)) fl |= SFX_FL_NEAR;
                if(!(bnd <= P.abserr)){ fl |= SFX_FL_DD; v = 0; }
                else if(!isfinite(v)) v = 0;
            }
        }
    }
    if(in){ vout[o] = v; flout[o] = (uint8_t)fl; if(P.raw_ab) P.raw_ab[o] = rab; }
    if(ddidx){
        // the points that need the double-double recompute, appended to ddidx (one atomic per wave; the
        // list's order is immaterial: the recompute writes each point at its own slot)
        bool q = (fl & SFX_FL_DD) != 0;
        unsigned long long m = __ballot(q);
        unsigned base = 0;
        if(m && __lane_id() == 0) base = atomicAdd(ddcnt, (unsigned)__popcll(m));
        base = __shfl(base, 0, 64);
        if(q) ddidx[base + __popcll(m & ((1ull << __lane_id()) - 1ull))] = (uint32_t)o;
    }
}

// the body of the evaluation kernels for node lists, and for the cell lists of integrals with D < 6
template<class G, int MODE, int AB>
__device__ __forceinline__ void eval_points(const sfx_cell_t *__restrict__ pts, const uint32_t *__restrict__ list,
                                            const uint32_t *__restrict__ cnt, long n, int st, const sfx_eval_params_t &P,
                                            double *__restrict__ vout, uint8_t *__restrict__ flout,
                                            uint32_t *__restrict__ ddidx, uint32_t *__restrict__ ddcnt){
    constexpr int D = G::D;
    long i = blockIdx.x*(long)blockDim.x + threadIdx.x;
    if(MODE == MODE_LIST){ long m = (long)*cnt << D; if(m < n) n = m; }
    if(blockIdx.x*(long)blockDim.x >= n) return;               // the whole block is past the list
    // the job's table e^{kh} (2K+1 <= 255 values) staged in LDS: the s lookups of every point read it
    __shared__ double sEX[256];
    for(int j = threadIdx.x; j < 2*P.K + 1; j += blockDim.x) sEX[j] = P.EX[j];
    __syncthreads();
    bool in = i < n;
    int k[D]; unsigned cfl = 0; bool valid = false; long o = i;
    if(in){
        if(MODE == MODE_LIST){
            long c = list[i >> D]; int b = (int)(i & ((1 << D) - 1));
            o = (c << D) | b;
            valid = point_coords<D, true>(pts, o, st, P.K, k, &cfl);
        } else valid = point_coords<D, false>(pts, i, st, P.K, k, &cfl);
    }
    // the skip class: ABSKIP and the parent's skip bit (cells), or a node's nonzero flag byte (the grow's
    // skip generations at top = 0, sumkey.h:165-168); with ABCHECK its points are evaluated and tested in
    // full (eval.c:35, 57) and the launch for that class uses AB = 1
    const bool cls = valid && P.abskip && (MODE == MODE_NODES ? cfl != 0u : (cfl & 1u) != 0u);
    double sf[D], w = 0;
    if(valid) w = lattice_node<D>(sEX, P.K, P.H, k, sf);
    point_out<G, MODE, AB>(in, valid, cls, sf, w, k, st, o, P, vout, flout, ddidx, ddcnt);
}

// Cell lists of the integrals with D >= 6: the 2^D children of a cell fill whole waves, so each wave takes
// 64-child slices of one cell ("tasks", SFX_LIST_TASKS of them per wave, strided over the block's waves;
// a per-graph build choice).
// The slice's cell is wave-uniform -- its list entry and coordinates come through scalar loads -- and a
// lane's child differs from the cell in the lane's own index bits (child bits 0-5) and the slice's (bits
// 6 and up): child b sits at k_d = cell_d - st for every set bit d (reference sum.c:405-408).  A lane's
// table address is the cell's minus a per-lane constant (st or 0 per dimension, fixed for the launch), so
// the coordinates cost one integer operation per dimension; a dimension whose shifted side leaves the box
// (cell_d - st < -K, uniform) marks those children invalid (value 0, flag VALID clear), and their lookups
// land in the table's padding instead.  s, w and f are the same operations on the same values as
// eval_points: the results are bit-identical.
#ifndef SFX_LIST_TASKS
#define SFX_LIST_TASKS 1          // more than one: the loop's hoisted invariants raise the register count
#endif
constexpr int SFX_EX_PAD = 128;        // table entries below k = -K: the lookups of invalid children (st <= 128)

template<class G, int AB>
__device__ __forceinline__ void eval_list_waves(const sfx_cell_t *__restrict__ pts, const uint32_t *__restrict__ list,
                                                const uint32_t *__restrict__ cnt, long n, int st, const sfx_eval_params_t &P,
                                                double *__restrict__ vout, uint8_t *__restrict__ flout,
                                                uint32_t *__restrict__ ddidx, uint32_t *__restrict__ ddcnt){
    constexpr int D = G::D, SH = D - 6;
    static_assert(D >= 6, "one cell per wave needs 2^D >= 64");
    static_assert(sizeof(sfx_cell_t) == 16 && offsetof(sfx_cell_t, fl) == 15, "a cell is read as one uint4");
    const int nwb = SFX_BLOCK/64;
    // the count, the table's address and K requested together (the empty asm keeps the compiler from
    // sinking the last two past the count's test)
    const uint32_t cn = *cnt;
    unsigned long long exa = (unsigned long long)P.EX; int K = P.K;
    asm volatile("" : "+s"(exa), "+s"(K));
    const __attribute__((address_space(1))) double *EX = (const __attribute__((address_space(1))) double*)exa;
    { long m = (long)cn << D; if(m < n) n = m; }
    const int ntask = (int)(n >> 6);
    const int tb = blockIdx.x*nwb*SFX_LIST_TASKS;               // the block's first task
    if(tb >= ntask) return;
    const int lane = __lane_id();
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // the table's entry of this thread (2K + 1 <= 255 entries, one per thread) is requested first, then the
    // wave's first slice -- its list entry and its cell, one 16-byte scalar load -- so the two latencies
    // overlap; each later slice's cell is requested before the current slice's evaluation
    static_assert(SFX_BLOCK >= 255, "one table entry per thread");
    __shared__ double sEX[SFX_EX_PAD + 256];
    const bool hx = (int)threadIdx.x < 2*K + 1;
    double ex = 0;
    if(hx) ex = EX[threadIdx.x];
    int t = tb + wv;
    uint32_t c = 0; uint4 raw = make_uint4(0, 0, 0, 0);
    if(t < ntask){ c = list[t >> SH]; raw = *(const uint4*)(pts + c); }
    if(hx) sEX[SFX_EX_PAD + threadIdx.x] = ex;
    __syncthreads();
    // per lane and dimension d < 6: the byte offset subtracted from the cell's table address
    int sub[6];
#pragma unroll
    for(int d = 0; d < 6; d++) sub[d] = ((lane >> d) & 1) ? st*8 : 0;
    for(int j = 0; j < SFX_LIST_TASKS; j++, t += nwb){
        if(t >= ntask) break;
        const uint32_t cc = c; const uint4 cr = raw;
        if(j + 1 < SFX_LIST_TASKS && t + nwb < ntask){ c = list[(t + nwb) >> SH]; raw = *(const uint4*)(pts + c); }
        const unsigned cw[4] = {cr.x, cr.y, cr.z, cr.w};
        const unsigned cfl = cr.w >> 24;                            // sfx_cell_t: k[0..13], pad, fl (byte 15)
        const int bh = (t & ((1 << SH) - 1)) << 6;                // the slice's child bits 6 and up
        const long o = ((long)cc << D) | bh | lane;
        bool valid = true;
        double sf[D], w = 1;
#pragma unroll
        for(int d = 0; d < D; d++){
            const int ks = (int)(int8_t)(cw[d >> 2] >> (8*(d & 3)));
            const bool out = ks - st < -P.K;                       // the shifted side leaves the box
            const int a0 = (SFX_EX_PAD + ks + P.K)*8;              // the cell's own table address (bytes)
            int a;
            if(d < 6){ a = a0 - sub[d < 6 ? d : 0]; valid = valid && !(out && ((lane >> d) & 1)); }
            else { const bool bit = (bh >> d) & 1; a = bit ? a0 - st*8 : a0; valid = valid && !(out && bit); }
            const double s = *(const double*)((const char*)sEX + a);
            sf[d] = s; w *= s*P.H;                                 // lattice.c:30-34, index order
        }
        const bool cls = valid && P.abskip && (cfl & 1u) != 0u;
        point_out<G, MODE_LIST, AB>(true, valid, cls, sf, w, nullptr, st, o, P, vout, flout, ddidx, ddcnt);
    }
}

// Occupancy of the refinement kernels, chosen per graph and class by measurement (sincfeyn_amd/tune.py
// writes the table that build.py turns into -DSFX_LIST_WAVES0=n / -DSFX_LIST_WAVES1=n for the graph's
// translation unit): at least n waves per SIMD, i.e. at most 512/n VGPRs; 1 leaves the compiler's choice.
#ifndef SFX_LIST_WAVES0
#define SFX_LIST_WAVES0 1
#endif
#ifndef SFX_LIST_WAVES1
#define SFX_LIST_WAVES1 1
#endif
#define SFX_LIST_ATTR __attribute__((amdgpu_waves_per_eu(AB == 0 ? SFX_LIST_WAVES0 : SFX_LIST_WAVES1)))

template<class G, int MODE, int AB>
__global__ void __launch_bounds__(SFX_BLOCK)
k_eval(const sfx_cell_t *__restrict__ pts, const uint32_t *__restrict__ list, const uint32_t *__restrict__ cnt,
       long n, int st, sfx_eval_params_t P, double *__restrict__ vout, uint8_t *__restrict__ flout,
       uint32_t *__restrict__ ddidx, uint32_t *__restrict__ ddcnt){
    eval_points<G, MODE, AB>(pts, list, cnt, n, st, P, vout, flout, ddidx, ddcnt);
}
template<class G, int AB>
__global__ void __launch_bounds__(SFX_BLOCK) SFX_LIST_ATTR
k_eval_list(const sfx_cell_t *__restrict__ pts, const uint32_t *__restrict__ list, const uint32_t *__restrict__ cnt,
            long n, int st, sfx_eval_params_t P, double *__restrict__ vout, uint8_t *__restrict__ flout,
            uint32_t *__restrict__ ddidx, uint32_t *__restrict__ ddcnt){
    if constexpr(G::D >= 6) eval_list_waves<G, AB>(pts, list, cnt, n, st, P, vout, flout, ddidx, ddcnt);
    else eval_points<G, MODE_LIST, AB>(pts, list, cnt, n, st, P, vout, flout, ddidx, ddcnt);
}

// the high-precision recompute of points idx[0 .. *cnt-1] (grid-stride, so the launch does not depend on
// the count); writes, at the point's own index i, the value, the eps^2 contribution to the round column
// and the "unresolved" flag bit (qvalue, eval.c:25-30).  Q = sfx_dd (the reference's fallback) or
// sfx_qd (the all-QD error reference of the two-loop graphs).
template<class Q> struct hp_eps;
template<> struct hp_eps<sfx_dd> { static constexpr double v = SFX_QEPS; };
template<> struct hp_eps<sfx_qd> { static constexpr double v = SFX_QD_EPS; };

template<class G, bool CELLS, class Q>
__device__ __forceinline__ void hp_recompute(const sfx_cell_t *__restrict__ pts, int st, const uint32_t *__restrict__ idx,
                                             const uint32_t *__restrict__ cnt, const sfx_eval_params_t &P,
                                             double *__restrict__ vout, double *__restrict__ e2out, uint8_t *__restrict__ flout){
    constexpr int D = G::D;
    long m = *cnt;
    for(long j = blockIdx.x*(long)blockDim.x + threadIdx.x; j < m; j += (long)gridDim.x*blockDim.x){
        long i = idx[j];
        int k[D]; unsigned cfl;
        point_coords<D, CELLS>(pts, i, st, P.K, k, &cfl);
        double sf[D]; double w = lattice_node<D>(P.EX, P.K, P.H, k, sf);
        Q sq[D];
#pragma unroll
        for(int d = 0; d < D; d++) sq[d] = Q(sf[d]);
        Q abq;
        Q vq = G::template f<Q>(sq, &abq, 1);
        double vqd = (double)vq, abqd = (double)abq;
        double e = abqd*w*hp_eps<Q>::v;                               // eval.c:26
        double e2 = (e*16 <= P.abserr) ? 0.0 : (isfinite(e) ? e*e : 0.0);
        double v = isfinite(vqd) ? vqd*w : 0;
        if(!(e <= fabs(v))){ flout[i] |= (uint8_t)SFX_FL_UNRES; if(P.dropunres) v = 0; }
        vout[i] = v; e2out[i] = e2;
    }
}

template<class G, bool CELLS>
__global__ void __launch_bounds__(SFX_BLOCK)
k_dd(const sfx_cell_t *__restrict__ pts, int st, const uint32_t *__restrict__ idx, const uint32_t *__restrict__ cnt,
     sfx_eval_params_t P, double *__restrict__ vout, double *__restrict__ e2out, uint8_t *__restrict__ flout){
    hp_recompute<G, CELLS, sfx_dd>(pts, st, idx, cnt, P, vout, e2out, flout);
}

// the same in quad-double; compiled only for the graphs that have it (G::QD: the two-loop integrands)
template<class G, bool CELLS>
__global__ void __launch_bounds__(SFX_BLOCK)
k_qd(const sfx_cell_t *__restrict__ pts, int st, const uint32_t *__restrict__ idx, const uint32_t *__restrict__ cnt,
     sfx_eval_params_t P, double *__restrict__ vout, double *__restrict__ e2out, uint8_t *__restrict__ flout){
    if constexpr(G::QD) hp_recompute<G, CELLS, sfx_qd>(pts, st, idx, cnt, P, vout, e2out, flout);
}

// golden primitive: the integrand at nodes without the tests.  prec 0: o1 = f, o2 = Sum|terms| in FP64
// (SQ_AB = 1); prec 1: o1 + o2 = f and o3 + o4 = Sum|terms| in double-double; prec 2 (graphs with G::QD):
// the same two-double split of the quad-double values
template<class G>
__global__ void __launch_bounds__(SFX_BLOCK)
k_raw(const sfx_cell_t *__restrict__ pts, long n, const double *__restrict__ sfin, int prec,
      double *o1, double *o2, double *o3, double *o4){
    constexpr int D = G::D;
    long i = blockIdx.x*(long)blockDim.x + threadIdx.x; if(i >= n) return;
    double sf[D];
#pragma unroll
    for(int d = 0; d < D; d++) sf[d] = sfin[i*D + d];
    (void)pts;
    if(prec == 0){ double ab = 0; double f = G::template f<double>(sf, &ab, 1); o1[i] = f; o2[i] = ab; }
    else if(prec == 1){
        sfx_dd sq[D];
#pragma unroll
        for(int d = 0; d < D; d++) sq[d] = sfx_dd(sf[d]);
        sfx_dd ab; sfx_dd f = G::template f<sfx_dd>(sq, &ab, 1);
        o1[i] = f.hi; o2[i] = f.lo; o3[i] = ab.hi; o4[i] = ab.lo;
    } else {
        if constexpr(G::QD){
            sfx_qd sq[D];
#pragma unroll
            for(int d = 0; d < D; d++) sq[d] = sfx_qd(sf[d]);
            sfx_qd ab; sfx_qd f = G::template f<sfx_qd>(sq, &ab, 1);
            f = sfxqd::renorm(f); ab = sfxqd::renorm(ab);
            o1[i] = f.x[0]; o2[i] = f.x[1] + (f.x[2] + f.x[3]); o3[i] = ab.x[0]; o4[i] = ab.x[1] + (ab.x[2] + ab.x[3]);
        }
    }
}

// the graph's point function (the reference's ptq, common.h:38) in double-double at n points of the
// line parameters x (NL values each, hi and lo parts): the scale test's primitive (main.c:79-116)
template<class G>
__global__ void __launch_bounds__(SFX_BLOCK)
k_pt(long n, const double *__restrict__ xhi, const double *__restrict__ xlo, double *vhi, double *vlo){
    if constexpr(G::HAS_PT){
        long i = blockIdx.x*(long)blockDim.x + threadIdx.x; if(i >= n) return;
        sfx_dd x[G::NL];
#pragma unroll
        for(int j = 0; j < G::NL; j++) x[j] = sfx_dd(xhi[i*G::NL + j], xlo[i*G::NL + j]);
        sfx_dd ab; sfx_dd v = G::template pt<sfx_dd>(x, &ab);
        vhi[i] = v.hi; vlo[i] = v.lo;
    }
}

}  // namespace sfx

// host-side launchers of one graph, reached through the registry (sfx_registry.h)
//   eval: mode MODE_NODES, n records of pts (AB chosen per wave); mode MODE_LIST, the children of the
//         cells list[0 .. *cnt) of pts (cnt a device count; n bounds *cnt << D and sizes the grid), with
//         AB = 0 or 1 fixed for the launch
struct sfx_graph_entry {
    const char *name, *volkov; int L, D, weight;
    hipError_t (*eval)(hipStream_t s, int mode, int ab, const sfx_cell_t *pts, const uint32_t *list, const uint32_t *cnt,
                       long n, int st, const sfx_eval_params_t *P, double *v, uint8_t *fl, uint32_t *ddidx,
                       uint32_t *ddcnt);
    hipError_t (*dd)(hipStream_t s, int cells, const sfx_cell_t *pts, int st, const uint32_t *idx, const uint32_t *cnt,
                     long maxcnt, const sfx_eval_params_t *P, double *v, double *e2, uint8_t *nd);
    hipError_t (*raw)(hipStream_t s, long n, const double *sf, int prec, double *o1, double *o2, double *o3, double *o4);
    hipError_t (*pt)(hipStream_t s, long n, const double *xhi, const double *xlo, double *vhi, double *vlo);
    hipError_t (*qd)(hipStream_t s, int cells, const sfx_cell_t *pts, int st, const uint32_t *idx, const uint32_t *cnt,
                     long maxcnt, const sfx_eval_params_t *P, double *v, double *e2, uint8_t *nd);
    int nl, fixed, has_pt;     // lines, the fixed line, whether the graph has a point function (not VP)
    int has_qd;                // whether the quad-double recompute exists (two-loop graphs)
    const void *kfn[5];        // the kernels (eval nodes, eval list AB 0, eval list AB 1, dd nodes, dd cells)
};

#define SFX_EVAL_ARGS const sfx_cell_t*, const uint32_t*, const uint32_t*, long, int, sfx_eval_params_t, double*, uint8_t*, \
                      uint32_t*, uint32_t*
#define SFX_DD_ARGS const sfx_cell_t*, int, const uint32_t*, const uint32_t*, sfx_eval_params_t, double*, double*, uint8_t*

// the host launchers exist only in the host compilation pass; the device pass sees just the kernels
#if defined(__HIP_DEVICE_COMPILE__)
#define SFX_GRAPH_ENTRY(NAME, G, VOLKOV, L, D, W)                                                            \
template __global__ void sfx::k_eval<G, sfx::MODE_NODES, 2>(SFX_EVAL_ARGS);                                  \
template __global__ void sfx::k_eval_list<G, 0>(SFX_EVAL_ARGS);                                              \
template __global__ void sfx::k_eval_list<G, 1>(SFX_EVAL_ARGS);                                              \
template __global__ void sfx::k_dd<G, false>(SFX_DD_ARGS);                                                   \
template __global__ void sfx::k_dd<G, true>(SFX_DD_ARGS);                                                    \
template __global__ void sfx::k_qd<G, false>(SFX_DD_ARGS);                                                   \
template __global__ void sfx::k_qd<G, true>(SFX_DD_ARGS);                                                    \
template __global__ void sfx::k_raw<G>(const sfx_cell_t*, long, const double*, int, double*, double*, double*, double*); \
template __global__ void sfx::k_pt<G>(long, const double*, const double*, double*, double*);
#else
#define SFX_GRAPH_ENTRY(NAME, G, VOLKOV, L, D, W)                                                            \
static hipError_t sfx_eval_##NAME(hipStream_t s, int mode, int ab, const sfx_cell_t *pts, const uint32_t *list, \
                                  const uint32_t *cnt, long n, int st, const sfx_eval_params_t *P, double *v, \
                                  uint8_t *fl, uint32_t *ddidx, uint32_t *ddcnt){                            \
    if(n <= 0) return hipSuccess;                                                                            \
    dim3 g((unsigned)((n + SFX_BLOCK - 1)/SFX_BLOCK)), b(SFX_BLOCK);                                         \
    if(mode != sfx::MODE_NODES && (D) >= 6){                                                                 \
        /* 64-child slices, SFX_LIST_TASKS per wave (eval_list_waves) */                                     \
        if(st > sfx::SFX_EX_PAD) return hipErrorInvalidValue;                                                \
        const long per = (SFX_BLOCK/64)*SFX_LIST_TASKS;                                                       \
        g = dim3((unsigned)(((n >> 6) + per - 1)/per));                                                      \
    }                                                                                                        \
    if(mode == sfx::MODE_NODES)                                                                              \
        hipLaunchKernelGGL((sfx::k_eval<G, sfx::MODE_NODES, 2>), g, b, 0, s, pts, list, cnt, n, st, *P, v, fl, ddidx, ddcnt); \
    else if(ab == 0)                                                                                         \
        hipLaunchKernelGGL((sfx::k_eval_list<G, 0>), g, b, 0, s, pts, list, cnt, n, st, *P, v, fl, ddidx, ddcnt);  \
    else hipLaunchKernelGGL((sfx::k_eval_list<G, 1>), g, b, 0, s, pts, list, cnt, n, st, *P, v, fl, ddidx, ddcnt);     \
    return hipGetLastError();                                                                                \
}                                                                                                            \
static hipError_t sfx_dd_##NAME(hipStream_t s, int cells, const sfx_cell_t *pts, int st, const uint32_t *idx, \
                                const uint32_t *cnt, long maxcnt, const sfx_eval_params_t *P, double *v,     \
                                double *e2, uint8_t *nd){                                                    \
    if(maxcnt <= 0) return hipSuccess;                                                                       \
    long nb = (maxcnt + SFX_BLOCK - 1)/SFX_BLOCK; if(nb > 4096) nb = 4096;                                   \
    if(cells) hipLaunchKernelGGL((sfx::k_dd<G, true>), dim3((unsigned)nb), dim3(SFX_BLOCK), 0, s, pts, st, idx, cnt, *P, v, e2, nd); \
    else hipLaunchKernelGGL((sfx::k_dd<G, false>), dim3((unsigned)nb), dim3(SFX_BLOCK), 0, s, pts, st, idx, cnt, *P, v, e2, nd);   \
    return hipGetLastError();                                                                                \
}                                                                                                            \
static hipError_t sfx_qd_##NAME(hipStream_t s, int cells, const sfx_cell_t *pts, int st, const uint32_t *idx, \
                                const uint32_t *cnt, long maxcnt, const sfx_eval_params_t *P, double *v,     \
                                double *e2, uint8_t *nd){                                                    \
    if(!G::QD) return hipErrorNotSupported;                                                                  \
    if(maxcnt <= 0) return hipSuccess;                                                                       \
    long nb = (maxcnt + SFX_BLOCK - 1)/SFX_BLOCK; if(nb > 4096) nb = 4096;                                   \
    if(cells) hipLaunchKernelGGL((sfx::k_qd<G, true>), dim3((unsigned)nb), dim3(SFX_BLOCK), 0, s, pts, st, idx, cnt, *P, v, e2, nd); \
    else hipLaunchKernelGGL((sfx::k_qd<G, false>), dim3((unsigned)nb), dim3(SFX_BLOCK), 0, s, pts, st, idx, cnt, *P, v, e2, nd);   \
    return hipGetLastError();                                                                                \
}                                                                                                            \
static hipError_t sfx_raw_##NAME(hipStream_t s, long n, const double *sf, int prec, double *o1, double *o2,  \
                                 double *o3, double *o4){                                                    \
    if(n <= 0) return hipSuccess;                                                                            \
    hipLaunchKernelGGL((sfx::k_raw<G>), dim3((unsigned)((n + SFX_BLOCK - 1)/SFX_BLOCK)), dim3(SFX_BLOCK), 0, s, \
                       (const sfx_cell_t*)nullptr, n, sf, prec, o1, o2, o3, o4);                             \
    return hipGetLastError();                                                                                \
}                                                                                                            \
static hipError_t sfx_pt_##NAME(hipStream_t s, long n, const double *xhi, const double *xlo, double *vhi,    \
                                double *vlo){                                                                \
    if(n <= 0) return hipSuccess;                                                                            \
    hipLaunchKernelGGL((sfx::k_pt<G>), dim3((unsigned)((n + SFX_BLOCK - 1)/SFX_BLOCK)), dim3(SFX_BLOCK), 0, s, \
                       n, xhi, xlo, vhi, vlo);                                                               \
    return hipGetLastError();                                                                                \
}                                                                                                            \
extern "C" const sfx_graph_entry SFX_E_##NAME = {#NAME, VOLKOV, L, D, W,                                      \
    sfx_eval_##NAME, sfx_dd_##NAME, sfx_raw_##NAME, sfx_pt_##NAME, sfx_qd_##NAME, G::NL, G::FIXED,           \
    G::HAS_PT ? 1 : 0, G::QD ? 1 : 0,                                                                        \
    {(const void*)&sfx::k_eval<G, sfx::MODE_NODES, 2>, (const void*)&sfx::k_eval_list<G, 0>,                 \
     (const void*)&sfx::k_eval_list<G, 1>,                                                                   \
     (const void*)&sfx::k_dd<G, false>, (const void*)&sfx::k_dd<G, true>}};
#endif
