// sfx_hsum.hip -- the hierarchical Sinc sum of one graph (reference driver/sum.c:556-814 with the
// symmetry fold off, as the reference's GPU build runs it, sum.c:571).
//
//  * coarsest level (step 2^top h): the support-following grow (reference driver/sumkey.h:130-461) and the
//    dilation (sumkey.h:20-90) on the device (sfx_grow.inc), only the order-dependent frontier sums on
//    the host; the sorted top-level cells are left in HBM for the refinement;
//  * refinement levels top-1 .. 0: entirely on the device, in blocks of cells -- children made in the
//    kernel, FP64 evaluation with the per-point tests, the double-double recompute of the flagged points,
//    deterministic double-double block sums, and the ordered compaction of the kept children into the
//    next level's cell list, which never leaves HBM.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <chrono>
#include <vector>
#include <mutex>
#include <thread>
#include <condition_variable>
#include <deque>
#include <map>
#include <atomic>
#include "sfx_device.hpp"
#include "sfx_engine.hpp"

namespace sfx {

static double now(){ return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

struct JobBufs {
    DevBuf EX, pts, v, fl, ddblk, ddidx, ddcnt, e2, nd, cur, nxt, part, keepoff, acc;
    DevBuf vfe;                             // a node list's values, round terms and flag bytes, back to back
    DevBuf lsk, lfu, lcnt;                  // a refinement block's cells split by |.|-sum class; counts
    // the coarsest level (sfx_grow.inc): frontiers, map of evaluated points, claims, keep list, tiers, dilation
    DevBuf fr, nx, gkey, misc, kflag, kidx, keep, kv, src, ckey, cflag, widx, keep2, kv2, gval2, tr, q1,
           dkey, skey, skey2, sortt;
    template<class F> void each(F f){
        DevBuf *all[] = {&EX, &pts, &v, &fl, &ddblk, &ddidx, &ddcnt, &e2, &nd, &cur, &nxt, &part, &keepoff, &acc, &vfe,
                         &lsk, &lfu, &lcnt, &fr, &nx, &gkey, &misc, &kflag, &kidx, &keep, &kv, &src, &ckey, &cflag, &widx,
                         &keep2, &kv2, &gval2, &tr, &q1, &dkey, &skey, &skey2, &sortt};
        for(DevBuf *b : all) f(*b);
    }
    void release(){ each([](DevBuf &b){ b.release(); }); }
    // allocate and free every buffer in the order of stream s (released buffers only)
    void bind(hipStream_t s){ each([s](DevBuf &b){ if(!b.p){ b.s = s; b.bound = true; } }); }
    // free the buffers above `keep` bytes (a large job's cell lists and coarsest-level tables), so what a
    // worker holds between jobs stays small and the engine's HBM budget describes the device
    void trim(size_t keep){ each([keep](DevBuf &b){ if(b.cap > keep) b.release(); }); }
};
// a context: one HIP stream and its device buffers; the library's own context serves sfx_hsum and the
// per-point calls, each engine worker has its own (so jobs on different workers overlap on the device)
// pinned host staging for the coarsest level's per-frontier transfers (asynchronous copies need it)
struct PinBuf {
    void *p = nullptr; size_t cap = 0;
    int reserve(size_t bytes){
        if(bytes <= cap) return SFX_OK;
        size_t want = cap ? cap : 4096; while(want < bytes) want *= 2;
        if(p) (void)hipHostFree(p);
        p = nullptr; cap = 0;
        if(hipHostMalloc(&p, want, 0) != hipSuccess){ p = nullptr; set_error("hipHostMalloc(%zu bytes) failed", want); return SFX_ENOMEM; }
        cap = want;
        return SFX_OK;
    }
    void release(){ if(p) (void)hipHostFree(p); p = nullptr; cap = 0; }
    template<class T> T *as() const { return (T*)p; }
};
struct HostBufs {
    PinBuf vfe, vfe2, idx, misc, rb, up;  // vfe2: the other half of the pipelined grow's double buffer; rb, up:
                                          // staging of the small read-backs and uploads
    void release(){ vfe.release(); vfe2.release(); idx.release(); misc.release(); rb.release(); up.release(); }
    void trim(size_t keep){ PinBuf *all[] = {&vfe, &vfe2, &idx, &misc, &rb, &up}; for(PinBuf *b : all) if(b->cap > keep) b->release(); }
};
struct Ctx {
    hipStream_t s = nullptr; JobBufs B; HostBufs H; hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};

// Host <-> device transfers of the context's stream go through its pinned staging, never through pageable
// memory (the runtime stages such copies itself, in buffers shared by every thread of the process).
// read_back: the n small device regions into their host destinations, one stream synchronisation.
struct RB { void *dst; const void *src; size_t bytes; };
static int read_back(Ctx &X, hipStream_t s, const RB *it, int n){
    size_t tot = 0; int rc;
    for(int i = 0; i < n; i++) tot += (it[i].bytes + 15) & ~(size_t)15;
    if((rc = X.H.rb.reserve(tot))) return rc;
    char *p = X.H.rb.as<char>();
    for(size_t i = 0, off = 0; i < (size_t)n; off += (it[i].bytes + 15) & ~(size_t)15, i++)
        SFX_HIP(hipMemcpyAsync(p + off, it[i].src, it[i].bytes, hipMemcpyDeviceToHost, s));
    SFX_HIP(hipStreamSynchronize(s));
    for(size_t i = 0, off = 0; i < (size_t)n; off += (it[i].bytes + 15) & ~(size_t)15, i++) memcpy(it[i].dst, p + off, it[i].bytes);
    return SFX_OK;
}
// upload: host bytes to the device through the staging, complete when it returns (the staging is reused)
static int upload(Ctx &X, hipStream_t s, void *dst, const void *src, size_t bytes){
    if(!bytes) return SFX_OK;
    int rc; if((rc = X.H.up.reserve(bytes))) return rc;
    memcpy(X.H.up.p, src, bytes);
    SFX_HIP(hipMemcpyAsync(dst, X.H.up.p, bytes, hipMemcpyHostToDevice, s));
    SFX_HIP(hipStreamSynchronize(s));
    return SFX_OK;
}
static Ctx &default_ctx(){ static Ctx C; if(C.s != lib().stream){ C.s = lib().stream; C.B.bind(C.s); } return C; }

// ---- profiling: HIP events on the library stream around each kernel, accumulated per (graph, kind)
struct Prof {
    bool on = false;
    std::vector<sfx_prof_entry_t> table;
    std::mutex mu;
};
static Prof &prof(){ static Prof P; return P; }
static int prof_events(Ctx &c){
    for(int i = 0; i < 6; i++) if(!c.ev[i]) SFX_HIP(hipEventCreate(&c.ev[i]));
    return SFX_OK;
}
static void prof_add(const char *graph, int kind, double ms, long points, long skipped, long dd){
    Prof &P = prof();
    std::lock_guard<std::mutex> lk(P.mu);
    for(sfx_prof_entry_t &e : P.table) if(e.kind == kind && !strcmp(e.graph, graph)){
        e.launches++; e.ms += ms; e.points += points; e.skipped += skipped; e.dd_points += dd; return;
    }
    sfx_prof_entry_t e; memset(&e, 0, sizeof e);
    snprintf(e.graph, sizeof e.graph, "%s", graph); e.kind = kind;
    e.launches = 1; e.ms = ms; e.points = points; e.skipped = skipped; e.dd_points = dd;
    P.table.push_back(e);
}
static double prof_ms(Ctx &c, int a, int b){ float ms = 0; (void)hipEventElapsedTime(&ms, c.ev[a], c.ev[b]); return ms; }

// Neumaier compensated sum (reference sum.c:18-20)
struct Neu {
    double s = 0, c = 0;
    void add(double x){ double t = s + x; if(fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s; s = t; }
    double val() const { return s + c; }
};

struct Job {
    const sfx_graph_entry *g; int gi, D, K, top; double h, eps;
    sfx_opts_t o; sfx_eval_params_t P{};
    hipStream_t s; Ctx *ctx;
    JobProgress *pg = nullptr;         // where the sum is (stage, level, block), for the engine's messages
};

// ---------------------------------------------------------------------------------------------------
// evaluate a list of nodes on the device with the per-point tests and the fallback (leafx semantics,
// eval.c:33-60): v, flags (SFX_PT_*), plus the fallback count, the round column's e^2 and the
// unresolved count, accumulated in list order
__global__ void k_zero_u32(uint32_t *p, long n){ for(long i = blockIdx.x*(long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x*blockDim.x) p[i] = 0; }

// a node list's results in the job's buffer B.vfe, back to back so that one copy brings them to the host:
// values v[0..n), round terms e2[0..n) (written for the fallback points only), flag bytes fl[0..n)
static inline double *node_v(JobBufs &B){ return B.vfe.as<double>(); }
static inline double *node_e2(JobBufs &B, long n){ return B.vfe.as<double>() + n; }
static inline uint8_t *node_fl(JobBufs &B, long n){ return (uint8_t*)(B.vfe.as<double>() + 2*n); }

// the host's view of a node list's results: values, round terms (fallback points only) and flag bytes
// in the context's pinned staging, valid until its next node evaluation
struct NodeOut { const double *v, *e2; const uint8_t *fl; };

// nodes already on the device (dpts); the results stay in B.vfe (node_v, node_fl) and come back to the
// pinned staging H.vfe in one copy (with st > 0 the flag bytes also carry the outer-shell class of each
// node at that step); the caller sums what it needs in node order
// the launches of a node list's evaluation: the FP64 kernel, then the fallback recompute of the flagged points;
// the results stay in B.vfe (node_v, node_e2, node_fl).  With the timers on, the three events of the context
// bracket the two kernels.  zeroed: the fallback list's count (B.ddcnt) is already 0 (the grow's previous
// frontier cleared it), one launch fewer.
static int eval_nodes_launch(Job &J, const sfx_cell_t *dpts, long n, int st, bool zeroed = false){
    JobBufs &B = J.ctx->B; Ctx &X = *J.ctx; int rc;
    // at least 4 MiB each, so a job's growing frontiers rarely reallocate
    const size_t bytes = (size_t)n*(2*sizeof(double) + 1), want = std::max(bytes, (size_t)4 << 20);
    if((rc = B.vfe.reserve(want)) || (rc = B.ddidx.reserve(std::max<size_t>(n*sizeof(uint32_t), (size_t)1 << 20))) ||
       (rc = B.ddcnt.reserve(64))) return rc;
    hipStream_t s = J.s;
    const bool pon = prof().on;
    if(pon){ if((rc = prof_events(X))) return rc; SFX_HIP(hipEventRecord(X.ev[0], s)); }
    if(!zeroed) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, B.ddcnt.as<uint32_t>(), 1L);
    SFX_HIP(J.g->eval(s, MODE_NODES, 2, dpts, nullptr, nullptr, n, st, &J.P, node_v(B), node_fl(B, n),
                      B.ddidx.as<uint32_t>(), B.ddcnt.as<uint32_t>()));
    if(pon) SFX_HIP(hipEventRecord(X.ev[1], s));
    SFX_HIP((J.o.qd_fallback ? J.g->qd : J.g->dd)(s, 0, dpts, 0, B.ddidx.as<uint32_t>(), B.ddcnt.as<uint32_t>(), n, &J.P,
                    node_v(B), node_e2(B, n), node_fl(B, n)));
    if(pon) SFX_HIP(hipEventRecord(X.ev[2], s));
    return SFX_OK;
}
// the timers' entries of a node list evaluated by eval_nodes_launch (its events complete; fl: its flag bytes)
static void eval_nodes_prof(Job &J, long n, const uint8_t *fl){
    Ctx &X = *J.ctx;
    long m = 0; for(long i = 0; i < n; i++) m += (fl[i] & SFX_FL_DD) != 0;
    prof_add(J.g->name, SFX_K_NODES_EVAL, prof_ms(X, 0, 1), n, 0, 0); prof_add(J.g->name, SFX_K_NODES_DD, prof_ms(X, 1, 2), m, 0, m);
}
// nodes already on the device (dpts), evaluated; values, round terms and flag bytes back in pinned memory (*out)
static int eval_nodes_core(Job &J, const sfx_cell_t *dpts, long n, int st, NodeOut *out, bool zeroed = false){
    if(n <= 0) return SFX_OK;
    JobBufs &B = J.ctx->B; HostBufs &H = J.ctx->H; int rc;
    const size_t bytes = (size_t)n*(2*sizeof(double) + 1), want = std::max(bytes, (size_t)4 << 20);
    if((rc = eval_nodes_launch(J, dpts, n, st, zeroed)) || (rc = H.vfe.reserve(want))) return rc;
    hipStream_t s = J.s;
    SFX_HIP(hipMemcpyAsync(H.vfe.p, B.vfe.p, bytes, hipMemcpyDeviceToHost, s));
    SFX_HIP(hipStreamSynchronize(s));
    const double *hv = H.vfe.as<double>();
    out->v = hv; out->e2 = hv + n; out->fl = (const uint8_t*)(hv + 2*n);
    if(prof().on) eval_nodes_prof(J, n, out->fl);
    return SFX_OK;
}

// the fallback points' count, round terms (in node order) and unresolved count of a node list
static void fallback_sums(const NodeOut &o, long n, long *nq, double *u, long *nd){
    long m = 0, d = 0; double uu = *u;
    for(long i = 0; i < n; i++){
        const bool q = (o.fl[i] & SFX_FL_DD) != 0;
        m += q; d += q && (o.fl[i] & SFX_FL_UNRES); uu += q ? o.e2[i] : 0.0;     // + 0.0 leaves the sum's bits
    }
    *nq += m; *u = uu; *nd += d;
}

static int eval_nodes_finished(Job &J, const sfx_cell_t *pts, long n, double *v, uint8_t *flg, long *nq, double *u, long *nd){
    if(n <= 0) return SFX_OK;
    JobBufs &B = J.ctx->B; int rc;
    if((rc = B.pts.reserve(n*sizeof(sfx_cell_t)))) return rc;
    hipStream_t s = J.s;
    if((rc = upload(*J.ctx, s, B.pts.p, pts, n*sizeof(sfx_cell_t)))) return rc;
    NodeOut o;
    if((rc = eval_nodes_core(J, B.pts.as<sfx_cell_t>(), n, 0, &o))) return rc;
    fallback_sums(o, n, nq, u, nd);
    memcpy(v, o.v, n*sizeof(double));
    for(long i = 0; i < n; i++) flg[i] = (o.fl[i] & (SFX_FL_NEAR | SFX_FL_DD)) | ((o.fl[i] & SFX_FL_UNRES) ? SFX_PT_UNRES : 0);
    return SFX_OK;
}

// ---------------------------------------------------------------------------------------------------
struct GrowOut {
    std::vector<sfx_cell_t> keep;
    double S = 0, A = 0, e1 = 0, e2 = 0, drop = 0, u = 0; long nev = 0, nq = 0, nd = 0, btuns = 0;
    double ref = 0;                 // the reference scale Ref: A of the grow, of the coarser box for the box scans
    // host seconds of the stages (the verbose report, and t_grow / t_dilate of the result)
    double t_loop = 0, t_final = 0, t_tiers = 0, t_dilate = 0; long frontiers = 0, tier_steps = 0;
    double t_f_eval = 0, t_f_host = 0, t_f_dev = 0;   // the frontier loop: evaluation + copy back, host sums, the rest
    double gshs = 0; long gshn = 0;   // the shell: evaluated points below the running threshold, never expanded
    long retries = 0;                 // the pipelined frontier loop's hand-overs to the exact one
    long nskip = 0, nmiss = 0;        // frontier points of the skip class, and those that needed the fallback
};

// box_zero (reference sum.c:346-359): z = 0 and 4096 pseudo-random nodes of the step-2^lev lattice; the
// node sequence is the reference's (a multiplicative hash of the probe number)
static int box_zero(Job &J, int lev, int *zero, long *nev, long *unresdrop){
    const int st = 1 << lev, W = 2*(J.K/st) + 1, NP = 4096;
    std::vector<sfx_cell_t> p(NP + 1);
    for(int i = 0; i <= NP; i++){
        memset(&p[i], 0, sizeof p[i]);
        unsigned long long y = 0x9E3779B97F4A7C15ULL*(unsigned long long)(i + 1);
        if(i) for(int d = 0; d < J.D; d++){ y ^= y >> 29; y *= 0xBF58476D1CE4E5B9ULL; y ^= y >> 32; p[i].k[d] = (int8_t)(((int)(y % (unsigned long long)W) - J.K/st)*st); }
    }
    std::vector<double> v(NP + 1); std::vector<uint8_t> f(NP + 1); long q = 0, nd = 0; double u = 0;
    int rc = eval_nodes_finished(J, p.data(), NP + 1, v.data(), f.data(), &q, &u, &nd);
    if(rc) return rc;
    *zero = 1; *unresdrop += nd;          // UNRESDROP is global in the reference; the probe's q and u are local
    for(int i = 0; i <= NP; i++) if(v[i] != 0){ *zero = 0; break; }
    if(*zero) *nev += NP + 1;
    return SFX_OK;
}

static void set_params(Job &J, double abserr){
    J.P.K = J.K; J.P.H = J.h; J.P.abskip = J.o.abskip; J.P.abserr = abserr; J.P.nearf = J.o.nearf;
    J.P.ddslack = J.o.ddslack; J.P.dropunres = J.o.dropunres; J.P.abcheck = J.o.abcheck; J.P.EX = J.ctx->B.EX.as<double>();
    J.P.raw_ab = nullptr;                       // the per-point tests in the kernel
}

// grow a device buffer keeping its first `used` bytes (DevBuf::reserve does not preserve contents)
static int reserve_preserving(DevBuf &b, size_t used, size_t want, hipStream_t s){
    if(want <= b.cap) return SFX_OK;
    size_t c = b.cap ? b.cap : 4096; while(c < want) c *= 2;
    DevBuf t; t.s = b.s; t.bound = b.bound;
    int rc = t.reserve(c); if(rc) return rc;
    if(used) SFX_HIP(hipMemcpyAsync(t.p, b.p, used, hipMemcpyDeviceToDevice, s));
    SFX_HIP(hipStreamSynchronize(s));
    b.release(); b = t;
    return SFX_OK;
}

__global__ void k_zero_acc(LevelAcc *a){ if(threadIdx.x == 0 && blockIdx.x == 0) memset(a, 0, sizeof *a); }

// a block's cells split by class: lsk the cells whose children skip the |.| sums (ABSKIP and the
// cell's skip bit, sum.c:402), lfu the others; cnt[0], cnt[1] count them (zeroed beforehand).  Each
// workgroup takes SPLIT_ITEMS cells per thread and reserves its ranges with one atomic per list; the
// lists' order is immaterial (every child is written at its own slot).
constexpr int SPLIT_ITEMS = 8;
__device__ __forceinline__ unsigned wave_incl_scan(unsigned x){
    for(int d = 1; d < 64; d <<= 1){ unsigned y = __shfl_up(x, d, 64); if((int)__lane_id() >= d) x += y; }
    return x;
}
__global__ void __launch_bounds__(256) k_cell_split(const sfx_cell_t *cells, long nc, int abskip, uint32_t *lsk,
                                                    uint32_t *lfu, uint32_t *cnt){
    __shared__ unsigned ws[4], wf[4], base[2];
    const long c0 = (long)blockIdx.x*256*SPLIT_ITEMS;
    unsigned bits = 0, ns = 0, nf = 0;
    for(int j = 0; j < SPLIT_ITEMS; j++){
        long c = c0 + j*256 + threadIdx.x;
        if(c < nc){ bool sk = abskip && (cells[c].fl & 1u); bits |= (sk ? 1u : 2u) << (2*j); ns += sk; nf += !sk; }
    }
    unsigned is = wave_incl_scan(ns), jf = wave_incl_scan(nf);
    int w = threadIdx.x >> 6;
    if(__lane_id() == 63){ ws[w] = is; wf[w] = jf; }
    __syncthreads();
    unsigned os = is - ns, of = jf - nf, ts = 0, tf = 0;
    for(int v = 0; v < 4; v++){ if(v < w){ os += ws[v]; of += wf[v]; } ts += ws[v]; tf += wf[v]; }
    if(threadIdx.x == 0){ base[0] = ts ? atomicAdd(&cnt[0], ts) : 0; base[1] = tf ? atomicAdd(&cnt[1], tf) : 0; }
    __syncthreads();
    os += base[0]; of += base[1];
    for(int j = 0; j < SPLIT_ITEMS; j++){
        unsigned b = (bits >> (2*j)) & 3u;
        uint32_t c = (uint32_t)(c0 + j*256 + threadIdx.x);
        if(b == 1u) lsk[os++] = c; else if(b == 2u) lfu[of++] = c;
    }
}

#include "sfx_ckpt.inc"

// where a level starts: at cell c0 with the partial sums acc and the kept children nxt of the cells
// before it (a resumed checkpoint), or at 0 with nothing
struct LevelStart { long c0 = 0; const LevelAcc *acc = nullptr; const std::vector<sfx_cell_t> *nxt = nullptr; };

static int level_blocks(Job &J, int lev, long ncur, const LevelStart &ls, CkRun &ck, long Bc, double pw, double thr,
                        long &prof_nv, long &prof_nsk, long &prof_nq);

// one refinement level on the device: the children of cur[c0..ncur) in blocks; returns the level's sums
// (a checkpoint between blocks when one is due: ck)
static int refine_level(Job &J, int lev, long ncur, long *nnext, double ref, LevelAcc *out, const LevelStart &ls, CkRun &ck){
    JobBufs &B = J.ctx->B; Ctx &X = *J.ctx; hipStream_t s = J.s; int rc;
    const int D = J.D, st = 1 << lev; const bool last = lev == 0;
    const double pw = pow((double)st, (double)D), thr = J.eps*ref*1e-2;
    long maxpts = J.o.block_points > 0 ? J.o.block_points : (1L << 24);
    long Bc = maxpts >> D; if(Bc < 1) Bc = 1;
    long nmax = Bc << D, nbt = (nmax + TILE - 1)/TILE;
    if((rc = B.v.reserve(nmax*sizeof(double))) || (rc = B.fl.reserve(nmax)) || (rc = B.ddblk.reserve((2*nbt + 2)*sizeof(uint32_t))) ||
       (rc = B.ddidx.reserve(nmax*sizeof(uint32_t))) || (rc = B.ddcnt.reserve(64)) || (rc = B.e2.reserve(nmax*sizeof(double))) ||
       (rc = B.nd.reserve(nmax)) || (rc = B.part.reserve(nbt*sizeof(LevelPart))) || (rc = B.keepoff.reserve(nbt*sizeof(unsigned long long))) ||
       (rc = B.acc.reserve(sizeof(LevelAcc))) || (rc = B.lsk.reserve(Bc*sizeof(uint32_t))) ||
       (rc = B.lfu.reserve(Bc*sizeof(uint32_t)))) return rc;
    // the next level's list at its bound (every child kept; the job's HBM reservation covers it), so the
    // blocks run back to back without a host round trip; per block three counters (two class lists and
    // the double-double list), all zeroed here
    const long nblk = (ncur - ls.c0 + Bc - 1)/Bc, nk0 = ls.nxt ? (long)ls.nxt->size() : 0;
    if(!last && (rc = B.nxt.reserve((size_t)(nk0 + ((ncur - ls.c0) << D))*sizeof(sfx_cell_t)))) return rc;
    if((rc = B.lcnt.reserve((size_t)std::max<long>(nblk, 1)*4*sizeof(uint32_t)))) return rc;
    if(ls.acc){
        SFX_HIP(hipMemcpyAsync(B.acc.p, ls.acc, sizeof(LevelAcc), hipMemcpyHostToDevice, s));
        if(!last && nk0) SFX_HIP(hipMemcpyAsync(B.nxt.p, ls.nxt->data(), nk0*sizeof(sfx_cell_t), hipMemcpyHostToDevice, s));
        SFX_HIP(hipStreamSynchronize(s));
    } else hipLaunchKernelGGL(k_zero_acc, dim3(1), dim3(1), 0, s, B.acc.as<LevelAcc>());
    if(nblk) hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)std::min<long>((nblk*4 + 255)/256, 1024)), dim3(256), 0, s, B.lcnt.as<uint32_t>(), nblk*4);
    long prof_nv = 0, prof_nsk = 0, prof_nq = 0;
    if(ls.acc){ prof_nv = (long)ls.acc->nv; prof_nsk = (long)ls.acc->nskip; prof_nq = (long)ls.acc->nq; }
    // with graph_capture the blocks' launches are recorded on the job's stream (thread-local capture: the
    // engine's other workers keep allocating on their own streams) and run as one graph; the chain holds
    // kernels only (no memset or copy nodes), and nothing below reads the device back between blocks
    // unless the timers or checkpoints are on, which turn the capture off
    const bool cap = J.o.graph_capture && !prof().on && !ck.ck->on && nblk > 0;
    if(cap) SFX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = level_blocks(J, lev, ncur, ls, ck, Bc, pw, thr, prof_nv, prof_nsk, prof_nq);
    if(cap){
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(s, &gr);
        if(rc == SFX_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        if(rc == SFX_OK){
            hipGraphExec_t ge = nullptr;
            e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
            if(e == hipSuccess) e = hipGraphLaunch(ge, s);
            // the executable graph may be destroyed once its launch is queued (the work completes)
            if(ge) (void)hipGraphExecDestroy(ge);
            if(e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate/Launch", __FILE__, __LINE__);
        }
        if(gr) (void)hipGraphDestroy(gr);
    }
    if(rc) return rc;
    { RB rb = {out, B.acc.p, sizeof *out}; if((rc = read_back(X, s, &rb, 1))) return rc; }
    *nnext = (long)out->nkept;
    return SFX_OK;
}

// the blocks of one refinement level from cell ls.c0 on (the launches only, unless the timers or a due
// checkpoint read the device back)
static int level_blocks(Job &J, int lev, long ncur, const LevelStart &ls, CkRun &ck, long Bc, double pw, double thr,
                        long &prof_nv, long &prof_nsk, long &prof_nq){
    JobBufs &B = J.ctx->B; Ctx &X = *J.ctx; hipStream_t s = J.s; int rc;
    const int D = J.D, st = 1 << lev; const bool last = lev == 0;
    for(long c0 = ls.c0; c0 < ncur; c0 += Bc){
        long nc = ncur - c0 < Bc ? ncur - c0 : Bc, n = nc << D;
        const sfx_cell_t *cells = B.cur.as<sfx_cell_t>() + c0;
        J.pg->block.store(c0/Bc);
        const bool pon = prof().on;
        if(pon){ if((rc = prof_events(X))) return rc; SFX_HIP(hipEventRecord(X.ev[0], s)); }
        // the block's cells in two lists by class, each run by the integrand compiled for its class; the
        // evaluations append the slots that need the double-double recompute to ddidx
        uint32_t *lc = B.lcnt.as<uint32_t>() + 4*((c0 - ls.c0)/Bc);
        hipLaunchKernelGGL(k_cell_split, dim3((unsigned)((nc + 256*SPLIT_ITEMS - 1)/(256*SPLIT_ITEMS))), dim3(256), 0, s, cells, nc, J.o.abskip,
                           B.lsk.as<uint32_t>(), B.lfu.as<uint32_t>(), lc);
        if(pon) SFX_HIP(hipEventRecord(X.ev[1], s));
        // the skip class with the |.| sums compiled out; with ABCHECK evaluated and tested in full
        if(J.o.abskip) SFX_HIP(J.g->eval(s, MODE_LIST, J.o.abcheck ? 1 : 0, cells, B.lsk.as<uint32_t>(), lc, n, st, &J.P, B.v.as<double>(),
                                         B.fl.as<uint8_t>(), B.ddidx.as<uint32_t>(), lc + 2));
        if(pon) SFX_HIP(hipEventRecord(X.ev[2], s));
        SFX_HIP(J.g->eval(s, MODE_LIST, 1, cells, B.lfu.as<uint32_t>(), lc + 1, n, st, &J.P, B.v.as<double>(),
                          B.fl.as<uint8_t>(), B.ddidx.as<uint32_t>(), lc + 2));
        if(pon) SFX_HIP(hipEventRecord(X.ev[3], s));
        SFX_HIP((J.o.qd_fallback ? J.g->qd : J.g->dd)(s, 1, cells, st, B.ddidx.as<uint32_t>(), lc + 2, n, &J.P, B.v.as<double>(),
                        B.e2.as<double>(), B.fl.as<uint8_t>()));
        if(pon) SFX_HIP(hipEventRecord(X.ev[4], s));
        LevelArgs a;
        a.v = B.v.as<double>(); a.fl = B.fl.as<uint8_t>(); a.cells = cells; a.n = n; a.D = D; a.st = st; a.last = last;
        a.abskip = J.o.abskip; a.pw = pw; a.thr = thr; a.e2 = B.e2.as<double>(); a.nxt = B.nxt.as<sfx_cell_t>(); a.nxt_cap = (long)(B.nxt.cap/sizeof(sfx_cell_t));
        a.part = B.part.as<LevelPart>(); a.keepoff = B.keepoff.as<unsigned long long>(); a.acc = B.acc.as<LevelAcc>();
        SFX_HIP(level_reduce(s, a));
        if(pon){
            SFX_HIP(hipEventRecord(X.ev[5], s));
            LevelAcc now_; RB rb = {&now_, B.acc.p, sizeof now_};
            if((rc = read_back(X, s, &rb, 1))) return rc;
            long dv = (long)now_.nv - prof_nv, dsk = (long)now_.nskip - prof_nsk, dq = (long)now_.nq - prof_nq;
            prof_nv = (long)now_.nv; prof_nsk = (long)now_.nskip; prof_nq = (long)now_.nq;
            prof_add(J.g->name, SFX_K_LEVEL_SPLIT, prof_ms(X, 0, 1), nc, 0, 0);
            if(J.o.abskip) prof_add(J.g->name, SFX_K_LEVEL_EVAL_SKIP, prof_ms(X, 1, 2), dsk, dsk, 0);
            prof_add(J.g->name, SFX_K_LEVEL_EVAL, prof_ms(X, 2, 3), dv - dsk, 0, 0);
            prof_add(J.g->name, SFX_K_LEVEL_DD, prof_ms(X, 3, 4), dq, 0, dq);
            prof_add(J.g->name, SFX_K_LEVEL_SUM, prof_ms(X, 4, 5), n, 0, 0);
        }
        if(J.o.verbose) fprintf(stderr, "  level %d: cells %ld/%ld\n", lev, c0 + nc, ncur);
        if(ck.ck->due() && c0 + nc < ncur){
            // a checkpoint after this block (sum.c:782-783): the device is drained first
            LevelAcc a_; RB rb = {&a_, B.acc.p, sizeof a_};
            if((rc = read_back(X, s, &rb, 1))) return rc;
            if(ck.cur_lev != lev){ if((rc = download_cells(B.cur, ncur, s, ck.cur))) return rc; ck.cur_lev = lev; }
            std::vector<sfx_cell_t> kept;
            if(!last && (rc = download_cells(B.nxt, (long)a_.nkept, s, kept))) return rc;
            CkState c;
            c.lev = lev; c.next = c0 + nc;
            c.s = a_.s[0] + a_.s[1]; c.a = a_.a[0] + a_.a[1]; c.drop = a_.drop[0] + a_.drop[1]; c.u = a_.u[0] + a_.u[1];
            c.ne = (long)a_.nv; c.q = (long)a_.nq;
            ck_fill_run(c, ck.res, ck.Slev, ck.nl, now() - ck.t0);
            cells_to_pts(ck.cur, D, c.cur); cells_to_pts(kept, D, c.nxt);
            c.es[0] = a_.b1[0] + a_.b1[1]; c.es[1] = a_.b2[0] + a_.b2[1];
            c.cnt[1] = ck.res->abskipped + (long)a_.nskip; c.cnt[2] = ck.res->abmiss + (long)a_.nmiss;
            ck_write(*ck.ck, c);
            ck.res->ckpts_written = ck.ck->written;
            if(ck.ck->stop_after && ck.ck->written >= ck.ck->stop_after){
                set_error("%s: stopped after checkpoint %d (%s)", J.g->name, ck.ck->written, ck.ck->path.c_str());
                return SFX_ESTOPPED;
            }
        }
    }
    return SFX_OK;
}

}  // namespace sfx

using namespace sfx;

extern "C" void sfx_release_job_buffers(void){ engine_shutdown(); default_ctx().B.release(); default_ctx().H.release(); }

extern "C" int sfx_prof_enable(int on){
    Prof &P = prof();
    if(on){ P.table.clear(); P.on = true; } else P.on = false;
    return SFX_OK;
}
extern "C" int sfx_prof_count(void){ return (int)prof().table.size(); }
extern "C" int sfx_prof_get(int i, sfx_prof_entry_t *out){
    if(i < 0 || i >= (int)prof().table.size() || !out){ set_error("sfx_prof_get: index %d", i); return SFX_EINVAL; }
    *out = prof().table[i];
    return SFX_OK;
}
extern "C" int sfx_sync(void){
    int rc; if((rc = require_init())) return rc;
    SFX_HIP(hipStreamSynchronize(lib().stream));
    return SFX_OK;
}

extern "C" int sfx_eval_finished(int graph, double h, int K, long n, const int16_t *k, const uint8_t *skip,
                                 double abserr, const sfx_opts_t *opts, double *v, uint8_t *flags, long *nq){
    int rc;
    if((rc = require_init())) return rc;
    if(graph < 0 || graph >= SFX_NGRAPHS || n < 0 || K < 1 || K > 127 || !(h > 0) || (n && (!k || !v || !flags))){
        set_error("sfx_eval_finished: bad arguments"); return SFX_EINVAL;
    }
    Job J; J.gi = graph; J.g = SFX_GRAPHS[graph]; J.D = J.g->D; J.h = h; J.K = K; J.eps = 0; J.top = 0;
    J.ctx = &default_ctx(); J.s = J.ctx->s;
    if(opts) J.o = *opts; else sfx_opts_default(&J.o);
    if(J.o.qd_fallback && !J.g->has_qd){ set_error("%s: no quad-double kernel (two-loop graphs only)", J.g->name); return SFX_EUNSUP; }
    std::vector<sfx_cell_t> pts(n);
    for(long i = 0; i < n; i++){
        memset(&pts[i], 0, sizeof pts[i]);
        for(int d = 0; d < J.D; d++){
            int x = k[i*J.D + d];
            if(x < -K || x > K){ set_error("sfx_eval_finished: coordinate %d outside |k| <= %d", x, K); return SFX_EINVAL; }
            pts[i].k[d] = (int8_t)x;
        }
        pts[i].fl = skip && skip[i] ? 1 : 0;
    }
    std::vector<double> EX(2*K + 1);
    for(int x = -K; x <= K; x++) EX[x + K] = exp(x*h);
    if((rc = J.ctx->B.EX.reserve(EX.size()*sizeof(double)))) return rc;
    if((rc = upload(*J.ctx, J.s, J.ctx->B.EX.p, EX.data(), EX.size()*sizeof(double)))) return rc;
    set_params(J, abserr);
    long q = 0, nd = 0; double u = 0;
    if((rc = eval_nodes_finished(J, pts.data(), n, v, flags, &q, &u, &nd))) return rc;
    if(nq) *nq = q;
    return SFX_OK;
}

#include "sfx_grow.inc"
#include "sfx_scan.inc"

static int check_job(int graph, double h, int N, double eps, int top){
    if(graph < 0 || graph >= SFX_NGRAPHS){ set_error("bad graph index %d", graph); return SFX_EINVAL; }
    if(top < 0 || top > 12 || !(h > 0) || N < 3 || !(eps > 0)){ set_error("need h > 0, N >= 3, eps > 0, 0 <= top <= 12 (got h=%g N=%d eps=%g top=%d)", h, N, eps, top); return SFX_EINVAL; }
    int m = 1 << (top + 1), K = (N - 1)/2; K = ((K + m - 1)/m)*m;
    if(K > 127){ set_error("N = %d gives K = %d; points are stored in 8 bits, so K <= 127 (N <= 255)", N, K); return SFX_EINVAL; }
    return SFX_OK;
}

// hsum of one graph on the context's stream (arguments checked by the caller); pg (may be null) receives
// the stage and level the sum is at
static int hsum_core(Ctx &X, int graph, double h, int N, double eps, int top, const sfx_opts_t *opts, sfx_result_t *res,
                     JobProgress *pg){
    int rc;
    Job J; J.gi = graph; J.g = SFX_GRAPHS[graph]; J.D = J.g->D; J.h = h; J.eps = eps; J.top = top; J.ctx = &X; J.s = X.s;
    JobProgress pg0; J.pg = pg ? pg : &pg0;
    char who[96]; snprintf(who, sizeof who, "%s h=%g N=%d eps=%g", J.g->name, h, N, eps);
    if(opts) J.o = *opts; else sfx_opts_default(&J.o);
    if(J.o.qd_fallback && !J.g->has_qd){ set_error("%s: no quad-double kernel (two-loop graphs only)", J.g->name); return SFX_EUNSUP; }
    int m = 1 << (top + 1); J.K = (N - 1)/2; J.K = ((J.K + m - 1)/m)*m;      // lattice.c:16
    memset(res, 0, sizeof *res);
    res->K = J.K; res->top = top; res->S[1] = res->S[2] = NAN;
    JobBufs &B = X.B;
    double t0 = now();
    // the EX table on the host with exp() (lattice.c:18-19), uploaded as is
    std::vector<double> EX(2*J.K + 1);
    for(int k = -J.K; k <= J.K; k++) EX[k + J.K] = exp(k*h);
    if((rc = B.EX.reserve(EX.size()*sizeof(double)))) return rc;
    if((rc = upload(X, J.s, B.EX.p, EX.data(), EX.size()*sizeof(double)))) return rc;
    set_params(J, J.o.abserr0);
    Ckpt ck; ck_setup(ck, J, N);
    CkState C;
    const bool resumed = ck.on && ck_read(ck, C) == 1;
    double Slev[16]; int nl = 0;
    long ncur = 0;                      // the top-level cells (sorted, from the grow and the dilation) in B.cur
    int lev0 = top - 1;
    LevelStart ls0; LevelAcc acc0; std::vector<sfx_cell_t> nxt0;
    double ref = 0, t2;
    if(resumed){
        // sum.c:596-603: the run continues at level C.lev, cell C.next, with the file's sums and lists
        t0 -= C.R.t;
        res->resumed = 1; res->resume_lev = C.lev; res->resume_next = C.next; res->resume_ncells = (long)C.cur.size();
        res->ref = ref = C.R.ref; res->prune = C.R.prune; res->trunc = C.R.trunc; res->unres = C.R.unres;
        res->nev = C.R.nev; res->nq = C.R.nq; res->kept_top = C.R.kept_top; res->abskipped = C.cnt[1]; res->abmiss = C.cnt[2];
        memcpy(Slev, C.Slev, sizeof(double)*C.nl); nl = C.nl;
        std::vector<sfx_cell_t> cur; pts_to_cells(C.cur, J.D, cur);
        ncur = (long)cur.size();
        if((rc = B.cur.reserve(std::max<long>(ncur, 1)*sizeof(sfx_cell_t)))) return rc;
        if(ncur) SFX_HIP(hipMemcpyAsync(B.cur.p, cur.data(), ncur*sizeof(sfx_cell_t), hipMemcpyHostToDevice, J.s));
        SFX_HIP(hipStreamSynchronize(J.s));
        lev0 = C.lev;
        if(C.next > 0){
            memset(&acc0, 0, sizeof acc0);
            acc0.s[0] = C.s; acc0.a[0] = C.a; acc0.drop[0] = C.drop; acc0.u[0] = C.u;
            acc0.b1[0] = C.es[0]; acc0.b2[0] = C.es[1];
            acc0.nv = (unsigned long long)C.ne; acc0.nq = (unsigned long long)C.q; acc0.nkept = C.nxt.size();
            pts_to_cells(C.nxt, J.D, nxt0);
            ls0.c0 = C.next; ls0.acc = &acc0; ls0.nxt = &nxt0;
        }
        t2 = now();
    } else {
        int zero = 0;
        if((rc = box_zero(J, top, &zero, &res->nev, &res->unresdrop))) return rc;
        if(zero){ res->zero = 1; res->S[0] = res->S[1] = res->S[2] = 0; res->t = now() - t0; return SFX_OK; }
        GrowOut G;
        if(J.o.box_scan){ if((rc = scan_top(J, top, eps, G, &ncur))) return rc; }
        else { if((rc = grow_device(J, top, eps*1e-3, G, &ncur))) return rc; G.ref = G.A; }
        res->nev += G.nev; res->nq += G.nq; res->unres += G.u; res->unresdrop += G.nd; res->btuns = G.btuns;
        res->abskipped += G.nskip; res->abmiss += G.nmiss;
        res->nev_level[top] = G.nev;
        res->ref = ref = G.ref; res->prune = G.drop;
        res->shell_n = G.gshn; res->shell_sum = G.gshs; res->shell_abs = G.drop; res->grow_retries = (int)G.retries;
        // top = 0 (sum.c:619-626): every evaluated point is in the sum; the prune column is the signed sum of
        // the never-expanded shell, an estimate of what lies beyond it
        if(!J.o.box_scan && top == 0 && G.gshn) res->prune = fabs(G.gshs);
        double t1 = now(); res->t_grow = t1 - t0 - G.t_dilate; res->t_dilate = G.t_dilate;
        if(J.o.verbose)
            fprintf(stderr, "  %s coarsest level: %.2f ms (%ld frontiers %.2f ms [evaluation and copy back %.2f, host sums %.2f, "
                    "candidates %.2f], final threshold %.2f ms, tiers %ld steps %.2f ms, dilation %.2f ms)\n", J.g->name, 1e3*(t1 - t0),
                    G.frontiers, 1e3*G.t_loop, 1e3*G.t_f_eval, 1e3*G.t_f_host, 1e3*G.t_f_dev, 1e3*G.t_final, G.tier_steps,
                    1e3*G.t_tiers, 1e3*G.t_dilate);
        if(G.A == 0){ res->S[0] = res->S[1] = res->S[2] = 0; res->t = now() - t0; return SFX_OK; }
        { double r = G.e2 > 0 ? G.e1/G.e2 : 0; res->trunc = (r > 0 && r < 1) ? G.e1*r/(1 - r) : G.e1; }
        Slev[nl++] = G.S;
        res->kept_top = ncur;
        t2 = now();
    }
    double abserr = fmax(J.o.abserr0, J.o.abrel*eps*ref);          // sum.c:651
    res->abserr = abserr;
    set_params(J, abserr);
    res->esok = resumed ? C.esok : J.D < 12 && !J.o.box_scan && top > 0;   // sum.c:597, 614
    CkRun ckr; ckr.ck = &ck; ckr.res = res; ckr.Slev = Slev; ckr.t0 = t0;
    BudgetHold bh;                                            // HBM reservation of this job's cell lists
    if((rc = budget_acquire(bh, (double)ncur*sizeof(sfx_cell_t), who, J.pg))) return rc;
    for(int lev = lev0; lev >= 0; lev--){
        long nnext = 0; LevelAcc A;
        const LevelStart none, &ls = lev == lev0 ? ls0 : none;
        J.pg->set(STAGE_LEVEL, lev);
        // the bound of the next list: every child of the cells still to do kept, plus those kept already
        double want = lev > 0 ? ((double)(ncur - ls.c0)*(double)(1L << J.D) + (ls.nxt ? (double)ls.nxt->size() : 0))*sizeof(sfx_cell_t) : 0;
        if((rc = budget_acquire(bh, want, who, J.pg))){ budget_release(bh, bh.held); return rc; }
        ckr.nl = nl;
        if((rc = refine_level(J, lev, ncur, &nnext, ref, &A, ls, ckr))){ budget_release(bh, bh.held); return rc; }
        // keep the reservation of the list the next level reads
        if(lev > 0) budget_release(bh, bh.held - (double)nnext*sizeof(sfx_cell_t));
        double S = A.s[0] + A.s[1];
        res->prune += A.drop[0] + A.drop[1];
        res->nev += (long)A.nv; res->nq += (long)A.nq; res->unresdrop += (long)A.nd;
        res->abskipped += (long)A.nskip; res->abmiss += (long)A.nmiss;
        res->nev_level[lev] = (long)A.nv;
        Slev[nl++] = S;
        if(lev == 0){
            res->unres = sqrt(A.u[0] + A.u[1]); res->Ah = A.a[0] + A.a[1];
            res->es1 = A.b1[0] + A.b1[1]; res->es2 = A.b2[0] + A.b2[1];
        }
        if(lev > 0){
            std::swap(B.cur, B.nxt); ncur = nnext;
            if(ck.on){
                // the level-end checkpoint (sum.c:798): the next level's cells, nothing of it done yet
                CkState c; c.lev = lev - 1; c.next = 0;
                ck_fill_run(c, res, Slev, nl, now() - t0);
                if((rc = download_cells(B.cur, ncur, J.s, ckr.cur))) { budget_release(bh, bh.held); return rc; }
                ckr.cur_lev = lev - 1;
                cells_to_pts(ckr.cur, J.D, c.cur);
                c.cnt[1] = res->abskipped; c.cnt[2] = res->abmiss;
                ck_write(ck, c);
                res->ckpts_written = ck.written;
                if(ck.stop_after && ck.written >= ck.stop_after){
                    budget_release(bh, bh.held);
                    set_error("%s: stopped after checkpoint %d (%s)", J.g->name, ck.written, ck.path.c_str());
                    return SFX_ESTOPPED;
                }
            }
        }
    }
    budget_release(bh, bh.held);
    ck_remove(ck);
    res->levels = nl;
    res->S[0] = Slev[nl - 1]; res->S[1] = nl > 1 ? Slev[nl - 2] : NAN; res->S[2] = nl > 2 ? Slev[nl - 3] : NAN;
    res->t_refine = now() - t2;
    res->t = now() - t0;
    return SFX_OK;
}

extern "C" int sfx_hsum(int graph, double h, int N, double eps, int top, const sfx_opts_t *opts, sfx_result_t *res){
    int rc;
    if((rc = require_init())) return rc;
    if(!res){ set_error("sfx_hsum: null result"); return SFX_EINVAL; }
    if((rc = check_job(graph, h, N, eps, top))) return rc;
    return hsum_core(default_ctx(), graph, h, N, eps, top, opts, res, nullptr);
}

// ---------------------------------------------------------------------------------------------------
// job engine (sfx_engine.hpp): the library's backend runs hsum_core on the worker's own context
namespace sfx {

static void *ectx_create(){
    Ctx *c = new Ctx();
    hipError_t e = hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking);
    if(e != hipSuccess){ delete c; hip_fail(e, "hipStreamCreate", __FILE__, __LINE__); return nullptr; }
    c->B.bind(c->s);
    return c;
}
static void ectx_destroy(void *p){
    Ctx *c = (Ctx*)p;
    (void)hipStreamSynchronize(c->s); c->B.release(); (void)hipStreamSynchronize(c->s); c->H.release();
    for(hipEvent_t &v : c->ev) if(v) (void)hipEventDestroy(v);
    (void)hipStreamDestroy(c->s); delete c;
}
static int erun(void *p, EJob &j){
    Ctx &X = *(Ctx*)p;
    int rc = hsum_core(X, j.graph, j.h, j.N, j.eps, j.top, &j.o, &j.res, &j.prog);
    if(rc) snprintf(j.err, sizeof j.err, "%s", sfx_last_error());
    return rc;
}
// free the buffers above 256 MB after each job (a large job's cell lists and coarsest-level tables), so
// what a worker holds between jobs stays small and the engine's HBM budget describes the device; the
// frees are ordered on the worker's stream and do not synchronise the device
static void eafter(void *p){
    Ctx &X = *(Ctx*)p;
    X.B.trim((size_t)256 << 20);
    (void)hipStreamSynchronize(X.s);
    X.H.trim((size_t)64 << 20);
}
static void engine_bind(){ EngineBackend &b = eng().be; if(!b.run){ b.ctx_create = ectx_create; b.ctx_destroy = ectx_destroy; b.run = erun; b.after_job = eafter; } }

}  // namespace sfx

extern "C" int sfx_engine_start(int workers, double hbm_budget_bytes){
    int rc; if((rc = require_init())) return rc;
    { std::lock_guard<std::mutex> lk(eng().mu); engine_bind(); }
    return engine_start(workers, hbm_budget_bytes, (double)lib().prop.totalGlobalMem);
}
extern "C" int sfx_engine_stop(void){ engine_shutdown(); return SFX_OK; }
extern "C" int sfx_engine_deadlines(double job_wait_s, double budget_wait_s){
    { std::lock_guard<std::mutex> lk(eng().mu); eng().wait_deadline_s = job_wait_s; }
    { std::lock_guard<std::mutex> lk(budget().mu); budget().deadline_s = budget_wait_s; }
    budget().cv.notify_all();
    return SFX_OK;
}

extern "C" int sfx_job_submit(int graph, double h, int N, double eps, int top, const sfx_opts_t *opts, long *job){
    int rc;
    if((rc = require_init())) return rc;
    if(!job){ set_error("sfx_job_submit: null job pointer"); return SFX_EINVAL; }
    if((rc = check_job(graph, h, N, eps, top))) return rc;
    EJob p;
    p.graph = graph; p.h = h; p.N = N; p.eps = eps; p.top = top;
    if(opts) p.o = *opts; else sfx_opts_default(&p.o);
    snprintf(p.name, sizeof p.name, "%s", SFX_GRAPHS[graph]->name);
    { std::lock_guard<std::mutex> lk(eng().mu); engine_bind(); }
    return engine_submit(p, job, (double)lib().prop.totalGlobalMem);
}

extern "C" int sfx_job_poll(long job, int *state){ return engine_poll(job, state); }

extern "C" int sfx_job_wait(long job, sfx_result_t *res, double *t_queue, double *t_run){
    double d; { std::lock_guard<std::mutex> lk(eng().mu); d = eng().wait_deadline_s; }
    return engine_wait(job, d, res, t_queue, t_run);
}
extern "C" int sfx_job_wait_for(long job, double timeout_s, sfx_result_t *res, double *t_queue, double *t_run){
    return engine_wait(job, timeout_s, res, t_queue, t_run);
}

extern "C" int sfx_job_release(long job){ return engine_release(job); }

#include "sfx_slots.inc"
