// sfx_runtime.hip -- device state, graph registry, the graph-independent device kernels (ordered
// compaction, deterministic double-double reductions of a refinement block) and the per-point C ABI.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <algorithm>
#include <vector>
#include "sfx_device.hpp"

namespace sfx {

static thread_local char g_err[512];

void set_error(const char *fmt, ...){
    va_list ap; va_start(ap, fmt); vsnprintf(g_err, sizeof g_err, fmt, ap); va_end(ap);
}
int hip_fail(hipError_t e, const char *what, const char *file, int line){
    set_error("%s:%d: %s: %s", file, line, what, hipGetErrorString(e));
    return SFX_EDEVICE;
}
Lib &lib(){ static Lib L; return L; }
int require_init(){
    if(!lib().init){ set_error("sfx_init has not been called"); return SFX_ENOTINIT; }
    return SFX_OK;
}

int DevBuf::reserve(size_t bytes){
    if(bytes <= cap) return SFX_OK;
    // small buffers double (amortised growth); large ones are sized to the request in 2 MiB steps, so a
    // level's bound-sized list does not take twice its reservation
    size_t want = cap ? cap : 4096;
    while(want < bytes && want < ((size_t)64 << 20)) want *= 2;
    if(want < bytes) want = (bytes + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
    if(p){
        hipError_t e = bound ? hipFreeAsync(p, s) : hipFree(p);
        p = nullptr; cap = 0;
        if(e != hipSuccess) return hip_fail(e, bound ? "hipFreeAsync" : "hipFree", __FILE__, __LINE__);
    }
    hipError_t e = bound ? hipMallocAsync(&p, want, s) : hipMalloc(&p, want);
    if(e != hipSuccess){ p = nullptr; set_error("%s(%zu bytes): %s", bound ? "hipMallocAsync" : "hipMalloc", want, hipGetErrorString(e)); return SFX_ENOMEM; }
    cap = want;
    return SFX_OK;
}
void DevBuf::release(){ if(p) (void)(bound ? hipFreeAsync(p, s) : hipFree(p)); p = nullptr; cap = 0; }

// ---------------------------------------------------------------------------------------------------
// double-double helpers for the reductions (the sum of two doubles exactly, then renormalised)
__device__ __forceinline__ void dacc_add(double &hi, double &lo, double x){
    double s = hi + x, bb = s - hi, e = (hi - (s - bb)) + (x - bb);
    lo += e; double t = s + lo; lo = lo - (t - s); hi = t;
}
__device__ __forceinline__ void dacc_addd(double &hi, double &lo, double xh, double xl){
    double s = hi + xh, bb = s - hi, e = (hi - (s - bb)) + (xh - bb);
    lo += e + xl; double t = s + lo; lo = lo - (t - s); hi = t;
}
__device__ __forceinline__ double shfl_down(double x, int d){ return __shfl_down(x, d, 64); }

// block-wide exclusive prefix of a per-thread predicate (blockDim = 256 = 4 waves); returns the rank of
// this thread's item among the set ones, *tot = number set in the block
__device__ __forceinline__ unsigned block_rank(bool pred, unsigned *sh, unsigned *tot){
    unsigned long long m = __ballot(pred);
    int lane = __lane_id(), wid = threadIdx.x >> 6;
    unsigned r = (unsigned)__popcll(m & ((1ull << lane) - 1ull));
    if(lane == 0) sh[wid] = (unsigned)__popcll(m);
    __syncthreads();
    unsigned base = 0, all = 0;
    for(int w = 0; w < (int)(blockDim.x >> 6); w++){ unsigned c = sh[w]; if(w < wid) base += c; all += c; }
    __syncthreads();
    *tot = all;
    return base + r;
}

__global__ void __launch_bounds__(TILE_THREADS) k_count_flags(const uint8_t *fl, long n, uint8_t mask, uint32_t *blk){
    long base = (long)blockIdx.x*TILE;
    unsigned c = 0;
    for(int j = 0; j < TILE_ITEMS; j++){ long i = base + j*TILE_THREADS + threadIdx.x; if(i < n && (fl[i] & mask)) c++; }
    for(int d = 32; d > 0; d >>= 1) c += __shfl_down(c, d, 64);
    __shared__ unsigned sh[TILE_THREADS/64];
    if(__lane_id() == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if(threadIdx.x == 0){ unsigned t = 0; for(int w = 0; w < TILE_THREADS/64; w++) t += sh[w]; blk[blockIdx.x] = t; }
}

// exclusive scan of cnt[0..nb-1] into off[], total into *total (+ *base if base != null); one block
__global__ void __launch_bounds__(1024) k_scan_u32(const uint32_t *cnt, long nb, uint32_t *off, uint32_t *total,
                                                   const unsigned long long *base){
    __shared__ unsigned long long sh[1024];
    long per = (nb + blockDim.x - 1)/blockDim.x, b0 = threadIdx.x*per, b1 = b0 + per < nb ? b0 + per : nb;
    unsigned long long s = 0;
    for(long b = b0; b < b1; b++) s += cnt[b];
    sh[threadIdx.x] = s; __syncthreads();
    for(int d = 1; d < (int)blockDim.x; d <<= 1){
        unsigned long long x = threadIdx.x >= (unsigned)d ? sh[threadIdx.x - d] : 0; __syncthreads();
        sh[threadIdx.x] += x; __syncthreads();
    }
    unsigned long long ex = sh[threadIdx.x] - s + (base ? *base : 0);
    for(long b = b0; b < b1; b++){ off[b] = (uint32_t)ex; ex += cnt[b]; }
    if(threadIdx.x == blockDim.x - 1 && total) *total = (uint32_t)sh[threadIdx.x];
}

__global__ void __launch_bounds__(TILE_THREADS) k_scatter_flags(const uint8_t *fl, long n, uint8_t mask, const uint32_t *off, uint32_t *out){
    __shared__ unsigned sh[TILE_THREADS/64];
    long base = (long)blockIdx.x*TILE;
    unsigned run = off[blockIdx.x];
    for(int j = 0; j < TILE_ITEMS; j++){
        long i = base + j*TILE_THREADS + threadIdx.x;
        bool p = i < n && (fl[i] & mask);
        unsigned tot, r = block_rank(p, sh, &tot);
        if(p) out[run + r] = (uint32_t)i;
        run += tot;
    }
}

hipError_t scan_blocks(hipStream_t s, const uint32_t *cnt, long nb, uint32_t *off, uint32_t *total){
    hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, s, cnt, nb, off, total, (const unsigned long long*)nullptr);
    return hipGetLastError();
}

hipError_t compact_flags(hipStream_t s, const uint8_t *fl, long n, uint8_t mask, uint32_t *blk, uint32_t *out, uint32_t *count){
    long nb = (n + TILE - 1)/TILE;
    if(nb == 0) return hipMemsetAsync(count, 0, sizeof(uint32_t), s);
    hipLaunchKernelGGL(k_count_flags, dim3((unsigned)nb), dim3(TILE_THREADS), 0, s, fl, n, mask, blk);
    hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, s, blk, nb, blk + nb, count, (const unsigned long long*)nullptr);
    hipLaunchKernelGGL(k_scatter_flags, dim3((unsigned)nb), dim3(TILE_THREADS), 0, s, fl, n, mask, blk + nb, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------
// refinement block: partial sums per tile
__device__ __forceinline__ int tier(double f, double t){ return (f >= 10*t) + (f >= 100*t); }

// a tile's partial from every thread's: fixed reduction tree (wave shuffles, then the 4 waves in
// order), written to part[blockIdx.x] and the kept count to keepoff[blockIdx.x]
__device__ __forceinline__ void tile_store(LevelPart t, LevelPart *part, unsigned long long *keepoff){
    for(int d = 32; d > 0; d >>= 1){
        dacc_addd(t.s[0], t.s[1], shfl_down(t.s[0], d), shfl_down(t.s[1], d));
        dacc_addd(t.a[0], t.a[1], shfl_down(t.a[0], d), shfl_down(t.a[1], d));
        dacc_addd(t.b1[0], t.b1[1], shfl_down(t.b1[0], d), shfl_down(t.b1[1], d));
        dacc_addd(t.b2[0], t.b2[1], shfl_down(t.b2[0], d), shfl_down(t.b2[1], d));
        dacc_addd(t.drop[0], t.drop[1], shfl_down(t.drop[0], d), shfl_down(t.drop[1], d));
        dacc_addd(t.u[0], t.u[1], shfl_down(t.u[0], d), shfl_down(t.u[1], d));
        t.nv += __shfl_down(t.nv, d, 64); t.nskip += __shfl_down(t.nskip, d, 64); t.nkeep += __shfl_down(t.nkeep, d, 64);
        t.nq += __shfl_down(t.nq, d, 64); t.nd += __shfl_down(t.nd, d, 64); t.nmiss += __shfl_down(t.nmiss, d, 64);
    }
    __shared__ LevelPart sh[TILE_THREADS/64];
    if(__lane_id() == 0) sh[threadIdx.x >> 6] = t;
    __syncthreads();
    if(threadIdx.x == 0){
        LevelPart r = sh[0];
        for(int w = 1; w < TILE_THREADS/64; w++){
            const LevelPart &q = sh[w];
            dacc_addd(r.s[0], r.s[1], q.s[0], q.s[1]); dacc_addd(r.a[0], r.a[1], q.a[0], q.a[1]);
            dacc_addd(r.b1[0], r.b1[1], q.b1[0], q.b1[1]); dacc_addd(r.b2[0], r.b2[1], q.b2[0], q.b2[1]);
            dacc_addd(r.drop[0], r.drop[1], q.drop[0], q.drop[1]); dacc_addd(r.u[0], r.u[1], q.u[0], q.u[1]);
            r.nv += q.nv; r.nskip += q.nskip; r.nkeep += q.nkeep; r.nq += q.nq; r.nd += q.nd; r.nmiss += q.nmiss;
        }
        part[blockIdx.x] = r;
        keepoff[blockIdx.x] = r.nkeep;
    }
}

// A thread takes its tile's children four consecutive ones at a time (the flag bytes as one 32-bit load, the
// values as two 16-byte loads, the four children's cell flag once), QUAD_BATCH quads loaded together before
// their sums (one memory latency per batch); within a thread the children are added in index order.
constexpr int QUAD_BATCH = 4;
static_assert(TILE_ITEMS % (4*QUAD_BATCH) == 0, "tile items in whole quad batches");
__device__ __forceinline__ void level_item(LevelPart &t, const LevelArgs &a, long i, unsigned f, double v, unsigned cfl){
    if(!(f & SFX_FL_VALID)) return;
    const double x = v*a.pw, av = fabs(v);
    const bool sk = a.abskip && (cfl & 1u);
    t.nv++; t.nskip += sk ? 1u : 0u;
    if(f & SFX_FL_DD){ t.nq++; t.nd += (f & SFX_FL_UNRES) ? 1u : 0u; t.nmiss += sk ? 1u : 0u; dacc_add(t.u[0], t.u[1], a.e2[i]); }
    dacc_add(t.s[0], t.s[1], x); dacc_add(t.a[0], t.a[1], av);
    if(a.last){ int te = (int)(cfl >> 1); if(te >= 1) dacc_add(t.b1[0], t.b1[1], x); if(te >= 2) dacc_add(t.b2[0], t.b2[1], x); }
    else if(av >= a.thr) t.nkeep++;
    else dacc_add(t.drop[0], t.drop[1], av*a.pw);
}
__global__ void __launch_bounds__(TILE_THREADS) k_level_partials(LevelArgs a){
    const long base = (long)blockIdx.x*TILE;
    LevelPart t = {};
    for(int q0 = 0; q0 < TILE_ITEMS/4; q0 += QUAD_BATCH){
        if(base + ((long)q0*TILE_THREADS + threadIdx.x)*4 >= a.n) break;
        uint32_t fq[QUAD_BATCH]; double2 vq[QUAD_BATCH][2]; uint8_t cq[QUAD_BATCH];
        #pragma unroll
        for(int k = 0; k < QUAD_BATCH; k++){
            const long i0 = base + ((long)(q0 + k)*TILE_THREADS + threadIdx.x)*4;
            if(i0 + 4 <= a.n){
                fq[k] = *(const uint32_t*)(a.fl + i0);
                vq[k][0] = *(const double2*)(a.v + i0); vq[k][1] = *(const double2*)(a.v + i0 + 2);
            } else {
                uint32_t f = 0; double w[4] = {0, 0, 0, 0};
                for(int j = 0; j < 4; j++) if(i0 + j < a.n){ f |= (uint32_t)a.fl[i0 + j] << (8*j); w[j] = a.v[i0 + j]; }
                fq[k] = f; vq[k][0] = make_double2(w[0], w[1]); vq[k][1] = make_double2(w[2], w[3]);
            }
            cq[k] = i0 < a.n ? a.cells[i0 >> a.D].fl : 0;          // D >= 2: the four children share their cell
        }
        #pragma unroll
        for(int k = 0; k < QUAD_BATCH; k++){
            const long i0 = base + ((long)(q0 + k)*TILE_THREADS + threadIdx.x)*4;
            level_item(t, a, i0, fq[k] & 0xffu, vq[k][0].x, cq[k]);
            level_item(t, a, i0 + 1, (fq[k] >> 8) & 0xffu, vq[k][0].y, cq[k]);
            level_item(t, a, i0 + 2, (fq[k] >> 16) & 0xffu, vq[k][1].x, cq[k]);
            level_item(t, a, i0 + 3, fq[k] >> 24, vq[k][1].y, cq[k]);
        }
    }
    tile_store(t, a.part, a.keepoff);
}

// one block of CB threads: the tiles' partials in tile order, the kept counts scanned into offsets;
// everything added to the level accumulator
constexpr int CB = 256;
__global__ void __launch_bounds__(CB) k_level_combine(LevelArgs a, long nb){
    __shared__ double sh[6][2][CB];
    __shared__ unsigned long long shc[6][CB];
    long per = (nb + CB - 1)/CB, b0 = threadIdx.x*per, b1 = b0 + per < nb ? b0 + per : nb;
    double v[6][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    unsigned long long nv = 0, nsk = 0, nk = 0, nd = 0, nq = 0, nm = 0;
    for(long b = b0; b < b1; b++){
        const LevelPart &p = a.part[b];
        dacc_addd(v[0][0], v[0][1], p.s[0], p.s[1]); dacc_addd(v[1][0], v[1][1], p.a[0], p.a[1]);
        dacc_addd(v[2][0], v[2][1], p.b1[0], p.b1[1]); dacc_addd(v[3][0], v[3][1], p.b2[0], p.b2[1]);
        dacc_addd(v[4][0], v[4][1], p.drop[0], p.drop[1]); dacc_addd(v[5][0], v[5][1], p.u[0], p.u[1]);
        nv += p.nv; nsk += p.nskip; nk += p.nkeep; nq += p.nq; nd += p.nd; nm += p.nmiss;
    }
    // exclusive scan of the kept counts over the threads' tile ranges, offsets from acc->nkept
    shc[0][threadIdx.x] = nk; __syncthreads();
    for(int d = 1; d < CB; d <<= 1){
        unsigned long long x = threadIdx.x >= (unsigned)d ? shc[0][threadIdx.x - d] : 0; __syncthreads();
        shc[0][threadIdx.x] += x; __syncthreads();
    }
    {
        unsigned long long ex = shc[0][threadIdx.x] - nk + a.acc->nkept;
        for(long b = b0; b < b1; b++){ a.keepoff[b] = ex; ex += a.part[b].nkeep; }
    }
    __syncthreads();
    for(int k = 0; k < 6; k++){ sh[k][0][threadIdx.x] = v[k][0]; sh[k][1][threadIdx.x] = v[k][1]; }
    shc[0][threadIdx.x] = nv; shc[1][threadIdx.x] = nsk; shc[2][threadIdx.x] = nk; shc[3][threadIdx.x] = nd; shc[4][threadIdx.x] = nq;
    shc[5][threadIdx.x] = nm;
    __syncthreads();
    // fixed pairwise tree over the threads
    for(int d = CB/2; d > 0; d >>= 1){
        if(threadIdx.x < (unsigned)d){
            int o = threadIdx.x + d;
            for(int k = 0; k < 6; k++){
                double h = sh[k][0][threadIdx.x], l = sh[k][1][threadIdx.x];
                dacc_addd(h, l, sh[k][0][o], sh[k][1][o]);
                sh[k][0][threadIdx.x] = h; sh[k][1][threadIdx.x] = l;
            }
            for(int k = 0; k < 6; k++) shc[k][threadIdx.x] += shc[k][o];
        }
        __syncthreads();
    }
    if(threadIdx.x == 0){
        LevelAcc *A = a.acc;
        dacc_addd(A->s[0], A->s[1], sh[0][0][0], sh[0][1][0]);
        dacc_addd(A->a[0], A->a[1], sh[1][0][0], sh[1][1][0]);
        dacc_addd(A->b1[0], A->b1[1], sh[2][0][0], sh[2][1][0]);
        dacc_addd(A->b2[0], A->b2[1], sh[3][0][0], sh[3][1][0]);
        dacc_addd(A->drop[0], A->drop[1], sh[4][0][0], sh[4][1][0]);
        dacc_addd(A->u[0], A->u[1], sh[5][0][0], sh[5][1][0]);
        A->nv += shc[0][0]; A->nskip += shc[1][0]; A->nkept += shc[2][0]; A->nd += shc[3][0]; A->nq += shc[4][0];
        A->nmiss += shc[5][0];
    }
}

__global__ void __launch_bounds__(TILE_THREADS) k_level_scatter(LevelArgs a){
    __shared__ unsigned sh[TILE_THREADS/64];
    long base = (long)blockIdx.x*TILE;
    unsigned long long run = a.keepoff[blockIdx.x];
    int D = a.D, msk = (1 << D) - 1;
    for(int j = 0; j < TILE_ITEMS; j++){
        long i = base + j*TILE_THREADS + threadIdx.x;
        bool p = false; double av = 0;
        if(i < a.n && (a.fl[i] & SFX_FL_VALID)){ av = fabs(a.v[i]); p = av >= a.thr; }
        unsigned tot, r = block_rank(p, sh, &tot);
        if(p && (long)(run + r) < a.nxt_cap){
            const sfx_cell_t cell = a.cells[i >> D];
            int b = (int)(i & msk);
            sfx_cell_t o;
            for(int d = 0; d < SFX_MAXD; d++) o.k[d] = d < D ? (int8_t)(cell.k[d] - (((b >> d) & 1) ? a.st : 0)) : 0;
            o.pad = 0;
            int psk = a.abskip && (cell.fl & 1u), near = (a.fl[i] & SFX_FL_NEAR) != 0;
            int te = cell.fl >> 1, tv = tier(av, a.thr);
            o.fl = (uint8_t)((a.abskip && (psk || !near) ? 1 : 0) | ((te < tv ? te : tv) << 1));
            a.nxt[run + r] = o;
        }
        run += tot;
    }
}

hipError_t level_reduce(hipStream_t s, const LevelArgs &a){
    if(a.D < 2) return hipErrorInvalidValue;                 // k_level_partials reads four children of one cell at once
    long nb = (a.n + TILE - 1)/TILE;
    if(nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_level_partials, dim3((unsigned)nb), dim3(TILE_THREADS), 0, s, a);
    hipLaunchKernelGGL(k_level_combine, dim3(1), dim3(CB), 0, s, a, nb);
    if(!a.last) hipLaunchKernelGGL(k_level_scatter, dim3((unsigned)nb), dim3(TILE_THREADS), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------
// full-box scan block (sum.c:166-191): every node is a lattice node; the sums of |v| pw over the outer
// shells go where a refinement level keeps its shadow-tier sums
__global__ void __launch_bounds__(TILE_THREADS) k_scan_partials(ScanArgs a){
    long base = (long)blockIdx.x*TILE;
    LevelPart t = {};
    for(int j = 0; j < TILE_ITEMS; j++){
        long i = base + j*TILE_THREADS + threadIdx.x;
        if(i >= a.n) break;
        uint8_t f = a.fl[i];
        double v = a.v[i], av = fabs(v)*a.pw;
        t.nv++;
        if(f & SFX_FL_DD){ t.nq++; t.nd += (f & SFX_FL_UNRES) ? 1u : 0u; dacc_add(t.u[0], t.u[1], a.e2[i]); }
        dacc_add(t.s[0], t.s[1], v*a.pw); dacc_add(t.a[0], t.a[1], av);
        if(f & SFX_FL_SHELL1) dacc_add(t.b1[0], t.b1[1], av); else if(f & SFX_FL_SHELL2) dacc_add(t.b2[0], t.b2[1], av);
        if(a.keep){ if(fabs(v) >= a.thr) t.nkeep++; else dacc_add(t.drop[0], t.drop[1], av); }
    }
    tile_store(t, a.part, a.keepoff);
}

__global__ void __launch_bounds__(TILE_THREADS) k_scan_scatter(ScanArgs a){
    __shared__ unsigned sh[TILE_THREADS/64];
    long base = (long)blockIdx.x*TILE;
    unsigned long long run = a.keepoff[blockIdx.x];
    for(int j = 0; j < TILE_ITEMS; j++){
        long i = base + j*TILE_THREADS + threadIdx.x;
        bool p = i < a.n && fabs(a.v[i]) >= a.thr;
        unsigned tot, r = block_rank(p, sh, &tot);
        if(p && (long)(run + r) < a.out_cap){ sfx_cell_t c = a.nodes[i]; c.fl = 0; a.out[run + r] = c; }
        run += tot;
    }
}

hipError_t scan_sums(hipStream_t s, const ScanArgs &a){
    long nb = (a.n + TILE - 1)/TILE;
    if(nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scan_partials, dim3((unsigned)nb), dim3(TILE_THREADS), 0, s, a);
    LevelArgs c = {};
    c.part = a.part; c.keepoff = a.keepoff; c.acc = a.acc;       // all that the combine step reads
    hipLaunchKernelGGL(k_level_combine, dim3(1), dim3(CB), 0, s, c, nb);
    return hipGetLastError();
}

hipError_t scan_keep(hipStream_t s, const ScanArgs &a){
    long nb = (a.n + TILE - 1)/TILE;
    if(nb == 0 || !a.keep) return hipSuccess;
    hipLaunchKernelGGL(k_scan_scatter, dim3((unsigned)nb), dim3(TILE_THREADS), 0, s, a);
    return hipGetLastError();
}

}  // namespace sfx

using namespace sfx;

// ---------------------------------------------------------------------------------------------------
// The entry point that allocates nothing (sfx_eval_finished_noalloc): its working memory is this
// file-scope device array.  In host code `scratch` only names the symbol; the memory is reached through
// the device address hipGetSymbolAddress returns, taken once in sfx_init with the device current.
// Layout: point records | round terms | fallback list | flag bytes | bad-coordinate marks | counter.
static constexpr long NA_MAX = SFX_NOALLOC_MAX_POINTS;
static constexpr size_t NA_CELLS = 0, NA_E2 = NA_CELLS + (size_t)NA_MAX*sizeof(sfx_cell_t), NA_IDX = NA_E2 + (size_t)NA_MAX*8,
                        NA_FL = NA_IDX + (size_t)NA_MAX*4, NA_BAD = NA_FL + (size_t)NA_MAX, NA_CNT = NA_BAD + (size_t)NA_MAX,
                        NA_BYTES = NA_CNT + 256;
static __device__ __align__(256) unsigned char scratch[NA_BYTES];
static unsigned char *g_scratch = nullptr;          // scratch's device address on the library's device

// device int16 coordinates (D per point) and skip bytes -> point records; thread 0 also clears the
// fallback counter (a kernel, not a memset, so the sequence can be captured into a HIP graph); a point with
// a coordinate outside |k| <= K is evaluated at k = 0 and marked SFX_PT_BADK by k_na_finish
__global__ void k_na_pack(const int16_t *k, const uint8_t *skip, long n, int D, int K, sfx_cell_t *cells, uint8_t *bad, uint32_t *cnt){
    long i = blockIdx.x*(long)blockDim.x + threadIdx.x;
    if(i == 0) *cnt = 0;
    if(i >= n) return;
    sfx_cell_t c; for(int d = 0; d < SFX_MAXD; d++) c.k[d] = 0; c.pad = 0;
    bool ok = true;
    for(int d = 0; d < D; d++){ int x = k[i*D + d]; ok = ok && x >= -K && x <= K; c.k[d] = (int8_t)x; }
    if(!ok) for(int d = 0; d < SFX_MAXD; d++) c.k[d] = 0;
    c.fl = skip && skip[i] ? 1 : 0;
    cells[i] = c; bad[i] = ok ? 0 : 1;
}
// the evaluation's flag bytes -> the SFX_PT_* bits of sfx_eval_finished; the fallback count to *nq
__global__ void k_na_finish(const uint8_t *fl, const uint8_t *bad, long n, double *v, uint8_t *flags, const uint32_t *cnt, uint32_t *nq){
    long i = blockIdx.x*(long)blockDim.x + threadIdx.x;
    if(i == 0) *nq = *cnt;
    if(i >= n) return;
    const unsigned f = fl[i];
    if(bad[i]){ v[i] = 0; flags[i] = SFX_PT_BADK; return; }
    flags[i] = (uint8_t)((f & (SFX_FL_NEAR | SFX_FL_DD)) | ((f & SFX_FL_UNRES) ? SFX_PT_UNRES : 0));
}

// ---------------------------------------------------------------------------------------------------
// C ABI: library and registry
extern "C" const char *sfx_last_error(void){ return g_err; }

extern "C" int sfx_init(int device){
    Lib &L = lib();
    if(L.init){
        if(device == L.device) return SFX_OK;
        set_error("sfx_init: already initialised on device %d", L.device); return SFX_EINVAL;
    }
    int n = 0;
    SFX_HIP(hipGetDeviceCount(&n));
    if(device < 0 || device >= n){ set_error("sfx_init: device %d of %d", device, n); return SFX_EINVAL; }
    SFX_HIP(hipSetDevice(device));
    SFX_HIP(hipGetDeviceProperties(&L.prop, device));
    SFX_HIP(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    // the stream-ordered allocations of the contexts come from the device's default pool; it keeps up to
    // 16 GB of freed memory for reuse instead of returning it to the driver at every synchronisation
    {
        hipMemPool_t pool; uint64_t keep = (uint64_t)16 << 30;
        SFX_HIP(hipDeviceGetDefaultMemPool(&pool, device));
        SFX_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    }
    {
        void *p = nullptr;
        SFX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch = (unsigned char*)p;
    }
    L.device = device; L.init = true;
    char b[512];
    snprintf(b, sizeof b, "%s (%s), %d CUs, %.0f MHz, %.1f GB HBM, %d graphs", L.prop.name, L.prop.gcnArchName,
             L.prop.multiProcessorCount, L.prop.clockRate/1e3, L.prop.totalGlobalMem/1e9, SFX_NGRAPHS);
    L.info = b;
    return SFX_OK;
}

extern "C" void sfx_release_job_buffers(void);
extern "C" void sfx_shutdown(void){
    Lib &L = lib();
    if(!L.init) return;
    sfx_release_job_buffers();
    (void)hipStreamSynchronize(L.stream);
    (void)hipStreamDestroy(L.stream);
    L.stream = nullptr; L.init = false; L.device = -1;
}

extern "C" const char *sfx_info(void){ return lib().init ? lib().info.c_str() : "sfx: not initialised"; }

extern "C" void sfx_opts_default(sfx_opts_t *o){
    memset(o, 0, sizeof *o);
    o->abserr0 = 1e-15; o->abrel = 1e-3; o->nearf = 1e-3; o->ddslack = 1; o->abskip = 1; o->dropunres = 1;
    o->keepbt = 0; o->btmargin = 2; o->block_points = 0; o->verbose = 0; o->growskip = 2; o->grow_pipeline = 1;
}

extern "C" int sfx_graph_count(void){ return SFX_NGRAPHS; }

extern "C" int sfx_graph_info(int i, sfx_graph_info_t *out){
    if(i < 0 || i >= SFX_NGRAPHS || !out){ set_error("sfx_graph_info: index %d of %d", i, SFX_NGRAPHS); return SFX_EINVAL; }
    const sfx_graph_entry *g = SFX_GRAPHS[i];
    memset(out, 0, sizeof *out);
    snprintf(out->name, sizeof out->name, "%s", g->name);
    snprintf(out->volkov, sizeof out->volkov, "%s", g->volkov);
    out->L = g->L; out->D = g->D; out->weight = g->weight;
    if(lib().init){
        hipFuncAttributes fa;
        if(hipFuncGetAttributes(&fa, g->kfn[2]) == hipSuccess){
            out->vgprs = fa.numRegs; out->scratch_bytes = (int)fa.localSizeBytes;
            int occ = 0;
            if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, g->kfn[2], SFX_BLOCK, 0) == hipSuccess) out->occupancy = occ;
        }
    }
    return SFX_OK;
}

extern "C" int sfx_graph_find(const char *s){
    if(!s) return SFX_EINVAL;
    for(int i = 0; i < SFX_NGRAPHS; i++) if(!strcmp(SFX_GRAPHS[i]->name, s) || !strcmp(SFX_GRAPHS[i]->volkov, s)) return i;
    set_error("no graph %s in this build", s);
    return SFX_EINVAL;
}

// graphs.c:17-29
extern "C" int sfx_graph_selected(int i, const char *s){
    if(i < 0 || i >= SFX_NGRAPHS || !s) return 0;
    const sfx_graph_entry *g = SFX_GRAPHS[i];
    if(!strcmp(s, g->name) || !strcmp(s, g->volkov)) return 1;
    if(!strcmp(s, "all2")) return g->L == 2;
    if(!strcmp(s, "all3")) return g->L == 3 && strncmp(g->name, "w3_", 3);
    if(!strcmp(s, "all4")) return g->L == 4;
    if(!strcmp(s, "q3"))   return !strncmp(g->name, "g3_", 3) && strcmp(g->name, "g3_lbl");
    if(!strcmp(s, "vp3"))  return !strncmp(g->name, "v3_", 3);
    if(!strcmp(s, "lbl"))  return !strcmp(g->name, "g3_lbl");
    return 0;
}

// ---------------------------------------------------------------------------------------------------
// per-point primitives
static int check_graph(int graph){
    if(graph < 0 || graph >= SFX_NGRAPHS){ set_error("graph index %d out of range (%d graphs)", graph, SFX_NGRAPHS); return SFX_EINVAL; }
    return SFX_OK;
}

extern "C" int sfx_eval_finished_noalloc(int graph, int K, double h, const double *dEX, long n, const int16_t *dk,
                                         const uint8_t *dskip, double abserr, const sfx_opts_t *opts, double *dv,
                                         uint8_t *dflags, uint32_t *dnq, void *stream){
    int rc;
    if((rc = require_init()) || (rc = check_graph(graph))) return rc;
    if(n < 0 || n > NA_MAX){
        set_error("sfx_eval_finished_noalloc: %ld points; this entry point takes at most %ld (SFX_NOALLOC_MAX_POINTS)", n, NA_MAX);
        return SFX_EINVAL;
    }
    if(K < 1 || K > 127 || !(h > 0) || !dEX || !dnq || (n && (!dk || !dv || !dflags))){
        set_error("sfx_eval_finished_noalloc: bad arguments"); return SFX_EINVAL;
    }
    if(!g_scratch){ set_error("sfx_eval_finished_noalloc: no scratch address (sfx_init)"); return SFX_EDEVICE; }
    const sfx_graph_entry *g = SFX_GRAPHS[graph];
    sfx_opts_t o; if(opts) o = *opts; else sfx_opts_default(&o);
    if(o.qd_fallback && !g->has_qd){ set_error("%s: no quad-double kernel (two-loop graphs only)", g->name); return SFX_EUNSUP; }
    sfx_eval_params_t P{};
    P.EX = dEX; P.K = K; P.H = h; P.abskip = o.abskip; P.abserr = abserr; P.nearf = o.nearf; P.ddslack = o.ddslack;
    P.dropunres = o.dropunres; P.abcheck = o.abcheck; P.raw_ab = nullptr;
    hipStream_t s = stream ? (hipStream_t)stream : lib().stream;
    sfx_cell_t *cells = (sfx_cell_t*)(g_scratch + NA_CELLS);
    double *e2 = (double*)(g_scratch + NA_E2);
    uint32_t *idx = (uint32_t*)(g_scratch + NA_IDX), *cnt = (uint32_t*)(g_scratch + NA_CNT);
    uint8_t *fl = g_scratch + NA_FL, *bad = g_scratch + NA_BAD;
    const unsigned nb = (unsigned)std::max<long>((n + 255)/256, 1);
    hipLaunchKernelGGL(k_na_pack, dim3(nb), dim3(256), 0, s, dk, dskip, n, g->D, K, cells, bad, cnt);
    SFX_HIP(hipGetLastError());
    if(n){
        SFX_HIP(g->eval(s, MODE_NODES, 2, cells, nullptr, nullptr, n, 0, &P, dv, fl, idx, cnt));
        SFX_HIP((o.qd_fallback ? g->qd : g->dd)(s, 0, cells, 0, idx, cnt, n, &P, dv, e2, fl));
    }
    hipLaunchKernelGGL(k_na_finish, dim3(nb), dim3(256), 0, s, fl, bad, n, dv, dflags, cnt, dnq);
    SFX_HIP(hipGetLastError());
    return SFX_OK;
}

extern "C" int sfx_eval_raw(int graph, long n, const double *sf, int prec, double *out){
    int rc;
    if((rc = require_init()) || (rc = check_graph(graph))) return rc;
    if(n < 0 || (n && (!sf || !out)) || prec < 0 || prec > 2){ set_error("sfx_eval_raw: bad arguments"); return SFX_EINVAL; }
    if(n == 0) return SFX_OK;
    const sfx_graph_entry *g = SFX_GRAPHS[graph];
    if(prec == 2 && !g->has_qd){ set_error("sfx_eval_raw: %s has no quad-double kernel", g->name); return SFX_EUNSUP; }
    hipStream_t s = lib().stream;
    int D = g->D, no = prec ? 4 : 2;
    DevBuf dsf, dout;
    if((rc = dsf.reserve((size_t)n*D*sizeof(double))) || (rc = dout.reserve((size_t)n*4*sizeof(double)))){ dsf.release(); dout.release(); return rc; }
    double *o = dout.as<double>();
    hipError_t e = hipMemcpyAsync(dsf.p, sf, (size_t)n*D*sizeof(double), hipMemcpyHostToDevice, s);
    if(e == hipSuccess) e = g->raw(s, n, dsf.as<double>(), prec, o, o + n, o + 2*n, o + 3*n);
    std::vector<double> h((size_t)n*4);
    if(e == hipSuccess) e = hipMemcpyAsync(h.data(), o, (size_t)n*4*sizeof(double), hipMemcpyDeviceToHost, s);
    if(e == hipSuccess) e = hipStreamSynchronize(s);
    dsf.release(); dout.release();
    if(e != hipSuccess) return hip_fail(e, "sfx_eval_raw", __FILE__, __LINE__);
    for(long i = 0; i < n; i++) for(int j = 0; j < no; j++) out[i*no + j] = h[(size_t)j*n + i];
    return SFX_OK;
}

// f w and Sum|terms| w at nodes (the reference GPU kernel's mode 0/1 outputs): s and w on the host as
// lattice_node does, f and ab on the device in FP64, the products in FP64
static int eval_weighted(int graph, double h, int K, long n, const int *kk, double *fw, double *abw){
    const sfx_graph_entry *g = SFX_GRAPHS[graph];
    int D = g->D;
    std::vector<double> EX(2*K + 1);
    for(int k = -K; k <= K; k++) EX[k + K] = exp(k*h);
    std::vector<double> sf((size_t)n*D), w(n), out((size_t)n*2);
    for(long i = 0; i < n; i++){
        double ww = 1;
        for(int d = 0; d < D; d++){ double s = EX[kk[i*D + d] + K]; sf[(size_t)i*D + d] = s; ww *= s*h; }
        w[i] = ww;
    }
    int rc = sfx_eval_raw(graph, n, sf.data(), 0, out.data());
    if(rc) return rc;
    for(long i = 0; i < n; i++){ fw[i] = out[2*i]*w[i]; abw[i] = out[2*i + 1]*w[i]; }
    return SFX_OK;
}

extern "C" int sfx_eval_nodes(int graph, double h, int K, long n, const int16_t *k, const uint8_t *skip,
                              double *fw, double *abw){
    int rc;
    if((rc = require_init()) || (rc = check_graph(graph))) return rc;
    (void)skip;
    if(n < 0 || K < 1 || K > 127 || !(h > 0) || (n && (!k || !fw || !abw))){ set_error("sfx_eval_nodes: bad arguments"); return SFX_EINVAL; }
    int D = SFX_GRAPHS[graph]->D;
    std::vector<int> kk((size_t)n*D);
    for(long i = 0; i < n*D; i++){
        if(k[i] < -K || k[i] > K){ set_error("sfx_eval_nodes: coordinate %d outside |k| <= %d", k[i], K); return SFX_EINVAL; }
        kk[i] = k[i];
    }
    return n ? eval_weighted(graph, h, K, n, kk.data(), fw, abw) : SFX_OK;
}

extern "C" int sfx_eval_cells(int graph, double h, int K, long nc, const int16_t *cells, const uint8_t *skip, int st,
                              double *fw, double *abw){
    int rc;
    if((rc = require_init()) || (rc = check_graph(graph))) return rc;
    (void)skip;
    int D = SFX_GRAPHS[graph]->D;
    if(nc < 0 || K < 1 || K > 127 || st < 1 || !(h > 0) || (nc && (!cells || !fw || !abw))){ set_error("sfx_eval_cells: bad arguments"); return SFX_EINVAL; }
    long n = nc << D;
    std::vector<int> kk; std::vector<long> at; kk.reserve((size_t)n*D);
    for(long c = 0; c < nc; c++){
        for(int d = 0; d < D; d++) if(cells[c*D + d] < -K || cells[c*D + d] > K){ set_error("sfx_eval_cells: coordinate outside the box"); return SFX_EINVAL; }
        for(long b = 0; b < (1L << D); b++){
            bool ok = true; int t[SFX_MAXD];
            for(int d = 0; d < D; d++){ t[d] = cells[c*D + d] - (((b >> d) & 1) ? st : 0); if(t[d] < -K) ok = false; }
            long i = (c << D) | b;
            if(!ok){ fw[i] = 0; abw[i] = 0; continue; }
            for(int d = 0; d < D; d++) kk.push_back(t[d]);
            at.push_back(i);
        }
    }
    long m = (long)at.size();
    std::vector<double> f(m), a(m);
    if(m && (rc = eval_weighted(graph, h, K, m, kk.data(), f.data(), a.data()))) return rc;
    for(long j = 0; j < m; j++){ fw[at[j]] = f[j]; abw[at[j]] = a[j]; }
    return SFX_OK;
}

// ---------------------------------------------------------------------------------------------------
// scale test (reference main.c:79-116): the points on the host in double-double, the point function on
// the device, the measure product and the ratios on the host, directions in index order
extern "C" int sfx_scale_test(int graph, double kappa, unsigned seed, int E, sfx_scale_result_t *out){
    int rc;
    if((rc = require_init()) || (rc = check_graph(graph))) return rc;
    if(!out || E < 1 || E > 8){ set_error("sfx_scale_test: bad arguments (E in 1..8)"); return SFX_EINVAL; }
    const sfx_graph_entry *g = SFX_GRAPHS[graph];
    if(!g->has_pt){ set_error("sfx_scale_test: %s has no point function (VP graph)", g->name); return SFX_EUNSUP; }
    const int n = g->nl, fixed = g->fixed, B = 2*E + 1;
    if(n < 2 || n > 16 || fixed < 0 || fixed >= n){ set_error("sfx_scale_test: %s: %d lines", g->name, n); return SFX_EINVAL; }
    long ndir = 1;
    for(int i = 0; i < n - 1; i++){ ndir *= B; if(ndir > (1L << 24)){ set_error("sfx_scale_test: (2E+1)^(n-1) too large"); return SFX_EINVAL; } }
    memset(out, 0, sizeof *out);
    out->n = n; out->fixed = fixed; out->ndir = ndir - 1; out->worstd = -1;
    srand(seed);
    sfx_dd x0[16];
    for(int i = 0; i < n; i++) x0[i] = sfx_dd(exp(4.0*rand()/RAND_MAX - 2.0));
    x0[fixed] = sfx_dd(1.0);
    // every (direction, kappa) point: x (hi, lo) for the device, the measure kept on the host
    std::vector<double> xh, xl; std::vector<sfx_dd> meas; std::vector<long> dir;
    xh.reserve((size_t)ndir*3*n); xl.reserve((size_t)ndir*3*n);
    for(long d = 0; d < ndir; d++){
        int e[16], nz = 0; long r = d;
        for(int i = 0; i < n; i++){ if(i == fixed){ e[i] = 0; continue; } e[i] = (int)(r % B) - E; r /= B; nz += e[i] != 0; }
        if(!nz) continue;
        dir.push_back(d);
        for(int q = 0; q < 3; q++){
            double kk = kappa*(q + 1)/3.0; sfx_dd m(1.0);
            for(int i = 0; i < n; i++){
                sfx_dd x = x0[i]*sfx_dd(exp(-kk*e[i]));
                xh.push_back(x.hi); xl.push_back(x.lo);
                if(i != fixed) m *= x;
            }
            meas.push_back(m);
        }
    }
    const long np = (long)meas.size();
    DevBuf dx, dv;
    if((rc = dx.reserve((size_t)np*n*2*sizeof(double))) || (rc = dv.reserve((size_t)np*2*sizeof(double)))){ dx.release(); dv.release(); return rc; }
    hipStream_t s = lib().stream;
    double *dxh = dx.as<double>(), *dxl = dxh + (size_t)np*n, *dvh = dv.as<double>(), *dvl = dvh + np;
    std::vector<double> vh(np), vl(np);
    hipError_t e = hipMemcpyAsync(dxh, xh.data(), (size_t)np*n*sizeof(double), hipMemcpyHostToDevice, s);
    if(e == hipSuccess) e = hipMemcpyAsync(dxl, xl.data(), (size_t)np*n*sizeof(double), hipMemcpyHostToDevice, s);
    if(e == hipSuccess) e = g->pt(s, np, dxh, dxl, dvh, dvl);
    if(e == hipSuccess) e = hipMemcpyAsync(vh.data(), dvh, np*sizeof(double), hipMemcpyDeviceToHost, s);
    if(e == hipSuccess) e = hipMemcpyAsync(vl.data(), dvl, np*sizeof(double), hipMemcpyDeviceToHost, s);
    if(e == hipSuccess) e = hipStreamSynchronize(s);
    dx.release(); dv.release();
    if(e != hipSuccess) return hip_fail(e, "sfx_scale_test", __FILE__, __LINE__);
    for(size_t j = 0; j < dir.size(); j++){
        long d = dir[j]; double gv[3];
        for(int q = 0; q < 3; q++){ sfx_dd v(vh[3*j + q], vl[3*j + q]); gv[q] = fabs((double)(v*meas[3*j + q])); }
        double ratio = gv[1] > 0 ? gv[2]/gv[1] : 0;
        int ee[16]; long r = d;
        for(int i = 0; i < n; i++){ if(i == fixed){ ee[i] = 0; continue; } ee[i] = (int)(r % B) - E; r /= B; }
        if(ratio > out->worst){
            out->worst = ratio; out->worstd = d;
            for(int q = 0; q < 3; q++) out->g_worst[q] = gv[q];
            for(int i = 0; i < n; i++) out->e_worst[i] = ee[i];
        }
        if(ratio > 0.3 && gv[2] > 1e-30){
            if(out->nslow < SFX_SCALE_MAXSLOW){
                for(int q = 0; q < 3; q++) out->g_slow[out->nslow][q] = gv[q];
                for(int i = 0; i < n; i++) out->e_slow[out->nslow][i] = ee[i];
                out->nslow++;
            }
            out->bad++;
        }
    }
    return SFX_OK;
}

// ---------------------------------------------------------------------------------------------------
// quad-double operations, element-wise (the arithmetic of the all-QD mode, exposed for its tests)
static __host__ __device__ sfx_qd qd_apply(int op, const sfx_qd &a, const sfx_qd &b){
    switch(op){ case 0: return a + b; case 1: return a - b; case 2: return a*b; default: return a/b; }
}
__global__ void k_qd_ops(int op, long n, const double *a, const double *b, double *out){
    long i = blockIdx.x*(long)blockDim.x + threadIdx.x; if(i >= n) return;
    sfx_qd x(a[4*i], a[4*i + 1], a[4*i + 2], a[4*i + 3]), y(b[4*i], b[4*i + 1], b[4*i + 2], b[4*i + 3]);
    sfx_qd r = qd_apply(op, x, y);
    for(int k = 0; k < 4; k++) out[4*i + k] = r.x[k];
}
extern "C" int sfx_qd_ops(int op, long n, const double *a, const double *b, double *out, int device){
    if(op < 0 || op > 3 || n < 0 || (n && (!a || !b || !out))){ set_error("sfx_qd_ops: bad arguments"); return SFX_EINVAL; }
    if(!device){
        for(long i = 0; i < n; i++){
            sfx_qd x(a[4*i], a[4*i + 1], a[4*i + 2], a[4*i + 3]), y(b[4*i], b[4*i + 1], b[4*i + 2], b[4*i + 3]);
            sfx_qd r = qd_apply(op, x, y);
            for(int k = 0; k < 4; k++) out[4*i + k] = r.x[k];
        }
        return SFX_OK;
    }
    int rc;
    if((rc = require_init())) return rc;
    if(n == 0) return SFX_OK;
    DevBuf da, db, dout;
    size_t bytes = (size_t)n*4*sizeof(double);
    if((rc = da.reserve(bytes)) || (rc = db.reserve(bytes)) || (rc = dout.reserve(bytes))){ da.release(); db.release(); dout.release(); return rc; }
    hipStream_t s = lib().stream;
    hipError_t e = hipMemcpyAsync(da.p, a, bytes, hipMemcpyHostToDevice, s);
    if(e == hipSuccess) e = hipMemcpyAsync(db.p, b, bytes, hipMemcpyHostToDevice, s);
    if(e == hipSuccess){
        hipLaunchKernelGGL(k_qd_ops, dim3((unsigned)((n + 255)/256)), dim3(256), 0, s, op, n, da.as<const double>(), db.as<const double>(), dout.as<double>());
        e = hipGetLastError();
    }
    if(e == hipSuccess) e = hipMemcpyAsync(out, dout.p, bytes, hipMemcpyDeviceToHost, s);
    if(e == hipSuccess) e = hipStreamSynchronize(s);
    da.release(); db.release(); dout.release();
    if(e != hipSuccess) return hip_fail(e, "sfx_qd_ops", __FILE__, __LINE__);
    return SFX_OK;
}

// ---------------------------------------------------------------------------------------------------
// calibration copy: 16-byte vector loads and stores, grid-stride
__global__ void __launch_bounds__(256) k_copy_probe(const double2 *__restrict__ a, double2 *__restrict__ b, long n){
    for(long i = blockIdx.x*(long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x*blockDim.x) b[i] = a[i];
}

extern "C" int sfx_copy_probe(long bytes, double *ms){
    int rc;
    if((rc = require_init())) return rc;
    if(bytes < 16 || bytes % 16){ set_error("sfx_copy_probe: bytes must be a positive multiple of 16"); return SFX_EINVAL; }
    DevBuf a, b;
    if((rc = a.reserve(bytes)) || (rc = b.reserve(bytes))){ a.release(); b.release(); return rc; }
    hipStream_t s = lib().stream;
    hipEvent_t e0, e1;
    SFX_HIP(hipEventCreate(&e0)); SFX_HIP(hipEventCreate(&e1));
    SFX_HIP(hipMemsetAsync(a.p, 1, bytes, s));
    SFX_HIP(hipEventRecord(e0, s));
    long n = bytes/16, nb = (n + 255)/256; if(nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)nb), dim3(256), 0, s, a.as<const double2>(), b.as<double2>(), n);
    SFX_HIP(hipGetLastError());
    SFX_HIP(hipEventRecord(e1, s));
    SFX_HIP(hipStreamSynchronize(s));
    float t = 0; SFX_HIP(hipEventElapsedTime(&t, e0, e1));
    if(ms) *ms = t;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    a.release(); b.release();
    return SFX_OK;
}
