/* sfx.h -- C ABI of the MI355X evaluator for sincFeyn's hot path: the hierarchical pruned Sinc sum over
 * the z = log s lattice and the forest-subtracted integrand kernels at its points.
 *
 * Plain C types only (pointers, sizes, PODs): the C host layer (sincqed, src/sincqed.c), the Python
 * bindings (ctypes) and any other caller bind to these symbols.  Every function returns 0 on success and
 * a negative SFX_E* code on failure (never exit(), unlike the reference's CUDA back end, gpu.cu:16);
 * sfx_last_error() describes the last failure of the calling thread.  Device memory is owned by the
 * library; result structs and host arrays are owned by the caller.  The calls that run on the library's
 * own stream (sfx_hsum, the per-point primitives, sfx_scale_test) are made from one host thread at a
 * time; the job engine's calls (sfx_job_*) may come from any thread, and its workers run jobs
 * concurrently, each on its own stream with its own buffers.
 *
 * Correspondence with the reference (reasther/sincFeyn v0.9.0):
 *   sfx_init / sfx_info          gpu_setup's device choice and gpu_info       (driver/gpu.cu:30-34, gpu.h:21)
 *   sfx_graph_*                  the graph registry graph_t / add_graph / find (driver/common.h:38,
 *                                graphs.c:8-33) and its groups (graphs.c:15-29)
 *   sfx_eval_nodes / _cells      gpu_submit + gpu_wait, and gpu_submit_cells  (driver/gpu.h:17-20,
 *                                gpu_kernel.h:22-36: f w and Sum|terms| w per node, mode 0 / mode 1)
 *   sfx_slot_*                   the same, asynchronous in two slots: gpu_setup / gpu_submit /
 *                                gpu_submit_cells / gpu_wait (driver/gpu.h:12-22, gpu.cu:54-103)
 *   sfx_eval_raw                 the kernels' fd / fq at given s (common.h:27-29), golden primitive
 *   sfx_hsum                     hsum() with lattice_setup()                  (driver/sum.c:556-814,
 *                                lattice.c:13-27), symmetry fold off as in the reference GPU build (sum.c:571)
 *   sfx_result_t                 res_t plus the counters the reference logs   (driver/sq.h:73, main.c:53-65)
 *   sfx_opts_t.ckpt_*            checkpoints in the SQCK4 file format        (driver/checkpoint.c:1-122)
 *   sfx_scale_test               `sincqed scale`                              (driver/main.c:77-116)
 *   sfx_job_* / sfx_engine_*     no counterpart: continuous batching of many sums on one device
 *   sfx_qd_ops, qd_fallback      no counterpart in the driver: the quad-double error reference of the
 *                                two-loop sums (the reference's pslq/quad programs compute such sums)
 */
#ifndef SFX_H
#define SFX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFX_OK          0
#define SFX_EINVAL     -1   /* bad argument (unknown graph, size out of range, ...) */
#define SFX_EDEVICE    -2   /* HIP runtime error */
#define SFX_ENOMEM     -3   /* device or host allocation failed */
#define SFX_ENOTINIT   -4   /* sfx_init not called */
#define SFX_EUNSUP     -5   /* a mode this build does not support */
#define SFX_ESTOPPED   -6   /* stopped after a checkpoint (sfx_opts_t.ckpt_stop_after); resume from the file */
#define SFX_ETIMEOUT   -7   /* a wait passed its deadline (job wait, HBM budget); the message names the job,
                               its stage and level */
#define SFX_ECANCELED  -8   /* the job never ran: the engine was stopped while it was queued */

typedef struct {
    char name[32];        /* "g3_00" */
    char volkov[64];      /* Volkov string, e.g. "2; 1-4, 3-6, 5-7" */
    int L, D, weight;     /* loops, Sinc dimension, mirror weight (graph_t fields, common.h:38) */
    int vgprs, sgprs, scratch_bytes, occupancy;   /* resources of the FP64 level kernel (informational) */
} sfx_graph_info_t;

/* Run options: the reference's environment knobs (main.c:176-185, sum.c:558-592), with its defaults
 * filled by sfx_opts_default().  The all-DD error reference is abserr0 = 0, abrel = 0, abskip = 0
 * (the reference's G2_ABSERR=0 SINCQED_ABREL=0 SINCQED_ABSKIP=0). */
typedef struct {
    double abserr0;       /* G2_ABSERR, default 1e-15 */
    double abrel;         /* SINCQED_ABREL, default 1e-3 */
    double nearf;         /* SINCQED_NEARF, default 1e-3 */
    double ddslack;       /* SINCQED_DDSLACK, default 1 */
    int abskip;           /* SINCQED_ABSKIP, default 1 */
    int dropunres;        /* SINCQED_DROPUNRES, default 1 */
    int keepbt;           /* SINCQED_KEEPBT: keep every grow boundary table (default 0) */
    int btmargin;         /* SINCQED_BTMARGIN, default 2 */
    long block_points;    /* children per refinement block on the device (0 = library default) */
    int verbose;          /* 1: progress lines on stderr */
    /* checkpoints (checkpoint.c:1-6): the file <ckpt_dir>/sincqed_<graph>.ckpt in the reference's SQCK4
     * format, written every ckpt_every seconds between refinement blocks and at the end of every level,
     * replaced atomically and removed when the graph finishes; read back at the start of a sum whose h,
     * N, eps and top match, so CPU (reference) and GPU runs resume each other */
    double ckpt_every;    /* SINCQED_CKPT: seconds between checkpoints, 0 = off (the library's default) */
    char ckpt_dir[256];   /* SINCQED_CKPT_DIR, "" = current directory.  The file is named after the graph
                             only (as the reference names it), so two sums of one graph that run at the
                             same time -- engine jobs at different h, say -- need different directories */
    int ckpt_stop_after;  /* stop the sum with SFX_ESTOPPED right after its n-th checkpoint (0 = never):
                             a job evicted this way resumes from its file */
    int qd_fallback;      /* 1: the fallback recompute in quad-double instead of double-double (graphs with
                             a quad-double kernel: the two-loop ones; else SFX_EUNSUP).  With abserr0 = 0,
                             abrel = 0, abskip = 0 this is the all-QD error reference of a two-loop sum. */
    int box_scan;         /* 1: SINCQED_GROW=0, the coarsest level by full-box scans instead of the grow
                             (sum.c:166-191, 610-617): every node at step 2^(top+1) h gives the reference
                             scale, every node at step 2^top h is summed and those with |f w| >=
                             1e-3 eps Ref are kept; no shadow tiers (esok = 0).  (2K/2^top + 1)^D nodes. */
    int abcheck;          /* SINCQED_ABCHECK: the points that may skip the |.| sums are evaluated and tested in
                             full anyway, and those that needed the fallback are counted (abmiss, eval.c:11-12) */
    int growskip;         /* SINCQED_GROWSKIP (default 2): at top = 0, a frontier point of the grow skips the |.|
                             sums for up to this many generations after a checked point (sumkey.h:165-168) */
    int graph_capture;    /* 1: the chain of launches of each refinement level (every block's class split,
                             evaluations, fallback and reductions) is captured on the job's stream as one HIP
                             graph, instantiated and launched once; ignored while checkpoints or the per-kernel
                             timers are on (both need the host between blocks).  Same results; 0 (default)
                             launches the kernels directly. */
    int grow_pipeline;    /* the grow's frontier loop (the library's default 1).  0: each frontier's values come
                             back to the host, whose serial sum of the reference scale A gives the keep threshold
                             before the next launches (two host round trips per frontier).  2: the device takes
                             the keep decisions from rigorous bounds of that threshold while the host sums and
                             checks each frontier one behind (one round trip); a value between the bounds, or a
                             check that fails, runs the grow again by loop 0 -- the same decisions and results
                             either way.  1: 2 for the frontiers of at least 8192 points, 0 for the smaller ones,
                             and 0 throughout while other engine jobs run or wait (the device is shared: the exact
                             loop costs it less).  3: as 2 with every value between the bounds (a test of the
                             takeover). */
} sfx_opts_t;

typedef struct {
    /* res_t (sq.h:73) */
    double S[3];          /* S_h, S_2h, S_4h (NaN where the run has no such level) */
    double Ah, ref, prune, trunc, unres, t;
    long nev, nq, kept_top;
    /* logged counters and the shadow sums (sum.c:74-81, 807-809; main.c:58-65) */
    double es1, es2;      /* final-level sums of the points of shadow tier >= 1, >= 2 */
    int esok;             /* shadow sums valid (top > 0) */
    int zero;             /* box_zero: integrand identically zero at the probe nodes (sum.c:346-359) */
    long abskipped;       /* evaluations that skipped the |.| sums */
    long unresdrop;       /* points unresolved even in double-double, counted as 0 */
    long btuns;           /* grow candidates that fell below the freed boundary tables (sumkey.h:452) */
    long nev_level[16];   /* evaluations per level (index = level, top..0), grow at index top */
    double abserr;        /* the fallback threshold used by the refinement */
    int K, top, levels;   /* lattice half-width after rounding, coarsest level, levels run */
    double t_grow, t_dilate, t_refine;   /* seconds: coarsest level (box_zero, grow or scans, tiers), dilation,
                                            refinement levels */
    int resumed;          /* the sum continued from a checkpoint: at level resume_lev, cell resume_next of */
    int resume_lev;       /*   resume_ncells (sum.c:602) */
    long resume_next, resume_ncells;
    int ckpts_written;    /* checkpoints this call wrote */
    long abmiss;          /* with opts->abcheck: skip-class points that needed the fallback (ABMISS) */
    long shell_n;         /* the grow's shell: evaluated points below the running threshold, never expanded; */
    double shell_sum, shell_abs;   /* their signed sum, and the grow's dropped |.| weight (sum.c:619-626) */
    int grow_retries;     /* times the grow's pipelined loop handed over to the exact one (opts grow_pipeline) */
} sfx_result_t;

/* ---- library ---- */
int         sfx_init(int device);
void        sfx_shutdown(void);
const char *sfx_info(void);
const char *sfx_last_error(void);
void        sfx_opts_default(sfx_opts_t *o);

/* ---- registry ---- */
int sfx_graph_count(void);
int sfx_graph_info(int index, sfx_graph_info_t *out);
/* index of the graph with this name or Volkov string, or SFX_EINVAL (graphs.c:30-33) */
int sfx_graph_find(const char *name_or_volkov);
/* graphs.c:17-29 selection rule: 1 if graph `index` is selected by `sel` (name, Volkov string or group
 * all2 all3 q3 vp3 lbl) */
int sfx_graph_selected(int index, const char *sel);

/* ---- per-point primitives (host buffers in, host buffers out; synchronous) ----
 * Lattice: K = half-width, h = step; EX[k+K] = exp(k h) is built with the host's exp().
 * sfx_eval_nodes: nodes k[i*D + d] (|k| <= K); fw[i] = f(s) w, abw[i] = Sum|terms| w with
 *   w = prod(s_d h) in index order; skip[i] != 0 evaluates without the |.| sums (abw then undefined);
 *   skip may be NULL.  This is the reference GPU kernel's mode 0 (gpu_kernel.h:32-35).
 * sfx_eval_cells: the 2^D children k = cell - st b of nc cells, child b of cell c at (c << D) | b;
 *   a child below the box gives 0, 0 (mode 1, gpu_kernel.h:27-31); skip is per cell. */
int sfx_eval_nodes(int graph, double h, int K, long n, const int16_t *k, const uint8_t *skip,
                   double *fw, double *abw);
int sfx_eval_cells(int graph, double h, int K, long nc, const int16_t *cells, const uint8_t *skip, int st,
                   double *fw, double *abw);

/* ---- the reference's GPU back end, asynchronous, in two slots (driver/gpu.h:12-22, gpu.cu:54-103) ----
 * For a host driver that keeps its own lattice loops, sums and double-double fallback and sends batches to
 * the device two at a time (the reference's driver built with SQ_GPU; src/sfx_gpu.c binds these to its
 * gpu_setup / gpu_submit / gpu_submit_cells / gpu_wait).  Called from one host thread.
 * sfx_slot_setup: the graph named `name` (with D lines) and the caller's table EX[k+K] = e^{kh},
 *   uploaded as given; returns 1, or 0 if this build has no device kernel for the graph (the caller then
 *   evaluates on the CPU), or a negative SFX_E* code.  Waits for both slots.
 * sfx_slot_submit: n nodes k[i*stride + d] (|k| <= K), skip[i] != 0 evaluates node i without the |.| sums
 *   (skip may be NULL); copies, launches and copies back, returning at once.
 * sfx_slot_submit_cells: the 2^D children k = cell - st b of nc cells (cells[c*stride + d]), child b of cell c
 *   at (c << D) | b, skip per cell; a child below the box gives 0, 0 (gpu_kernel.h:22-36, mode 1).
 * sfx_slot_wait: waits for the slot; fw[i] = f w, abw[i] = Sum|terms| w (0 where skipped), in pinned host
 *   memory valid until the slot's next submit. */
int sfx_slot_setup(const char *name, int D, const double *EX, int K, double H);
int sfx_slot_submit(int slot, long n, const int32_t *k, int stride, const uint8_t *skip);
int sfx_slot_submit_cells(int slot, long nc, const int32_t *cells, int stride, const uint8_t *skip, int st);
int sfx_slot_wait(int slot, const double **fw, const double **abw);

/* Scale test of a graph's integrand: replaces the reference's `sincqed scale` (driver/main.c:77-116,
 * scale_test).  From a point x0 (srand(seed); x0_i = exp(4 rand()/RAND_MAX - 2); x0_fixed = 1) every
 * direction e in {-E..E}^(n-1) \ {0} (the fixed line excluded; direction index d = mixed radix 2E+1,
 * first free line fastest) scales x_i -> x0_i exp(-k e_i) at k = kappa/3, 2 kappa/3, kappa, and
 * |g| = |pt(x) prod_{i != fixed} x_i| is taken in double-double (the point function on the device).
 * A direction is slow when |g(kappa)| / |g(2 kappa/3)| > 0.3 and |g(kappa)| > 1e-30.  SFX_EUNSUP for
 * the VP graphs, which have no point function. */
/* Quad-double arithmetic of the all-QD mode, element-wise on n values of four doubles each (a[4i..4i+3]):
 * op 0 a + b, 1 a - b, 2 a * b, 3 a / b; device 0 on the host (no GPU needed), 1 on the device. */
int sfx_qd_ops(int op, long n, const double *a, const double *b, double *out, int device);

#define SFX_SCALE_MAXSLOW 12
typedef struct {
    int n, fixed;                        /* lines of the graph, the fixed line */
    long ndir;                           /* directions tested, (2E+1)^(n-1) - 1 */
    int bad;                             /* slow directions */
    double worst; long worstd;           /* largest ratio and its direction index (-1: none above 0) */
    int e_worst[16]; double g_worst[3];  /* its exponents (all n lines) and |g| at the three kappas */
    int nslow;                           /* the first min(bad, 12) slow directions, in index order */
    int e_slow[SFX_SCALE_MAXSLOW][16]; double g_slow[SFX_SCALE_MAXSLOW][3];
} sfx_scale_result_t;
int sfx_scale_test(int graph, double kappa, unsigned seed, int E, sfx_scale_result_t *out);
/* the integrand at given Sinc variables sf[i*D + d], no weight and no tests.  prec 0: f and Sum|terms|
 * in FP64 into out[2i], out[2i+1]; prec 1: double-double, out[4i..4i+3] = f.hi, f.lo, ab.hi, ab.lo;
 * prec 2: quad-double (two-loop graphs, else SFX_EUNSUP), each value split into two doubles as prec 1 */
int sfx_eval_raw(int graph, long n, const double *sf, int prec, double *out);
/* the per-point values with the reference's tests and double-double fallback (eval.c:33-61) as the
 * refinement applies them: v[i] (0 if not finite), flags[i] = SFX_PT_* bits; abserr is the fallback
 * threshold.  Returns the number of fallback points in *nq. */
#define SFX_PT_NEAR 2
#define SFX_PT_DD   4
#define SFX_PT_UNRES 16
int sfx_eval_finished(int graph, double h, int K, long n, const int16_t *k, const uint8_t *skip,
                      double abserr, const sfx_opts_t *opts, double *v, uint8_t *flags, long *nq);

/* sfx_eval_finished for callers that cannot allocate: no counterpart in the reference; for code under
 * HIP graph capture, tight serving loops and harnesses that forbid hipMalloc.  Everything is
 * device-resident and the call only enqueues kernels on `stream` (NULL: the library's stream) and
 * returns: dEX[k+K] = e^{kh} as the caller computed it (2K+1 doubles), dk the n points' coordinates
 * (int16, D per point, |k| <= K), dskip the skip bytes (or NULL); out dv (values), dflags (SFX_PT_* bits;
 * SFX_PT_BADK for a point with a coordinate outside the box, its value 0) and *dnq (uint32, the fallback
 * count).  Its working memory is a fixed device array inside the library (about 7.9 MB, held until the
 * process ends): n <= SFX_NOALLOC_MAX_POINTS, larger n is refused with SFX_EINVAL.  Every call on a GPU
 * uses that one array, so calls that overlap on one device (other streams or threads) race; one call after
 * another on a stream is safe.  Same values, flags and count as sfx_eval_finished. */
#define SFX_NOALLOC_MAX_POINTS (1L << 18)
#define SFX_PT_BADK 128
int sfx_eval_finished_noalloc(int graph, int K, double h, const double *dEX, long n, const int16_t *dk,
                              const uint8_t *dskip, double abserr, const sfx_opts_t *opts, double *dv,
                              uint8_t *dflags, uint32_t *dnq, void *stream);

/* ---- the hierarchical sum ----
 * hsum(g, h, N, eps, top) for one graph on the device: K = (N-1)/2 rounded up to a multiple of
 * 2^(top+1) (lattice.c:16; K <= 127), 0 <= top <= 12.  The coarsest level (the support-following grow, or
 * the full-box scans of opts->box_scan, and the dilation) runs on the device with its tables in HBM, only
 * the grow's order-dependent frontier sums on the host; each refinement level runs on the device with
 * the cell lists resident in HBM. */
int sfx_hsum(int graph, double h, int N, double eps, int top, const sfx_opts_t *opts, sfx_result_t *res);

/* ---- job engine: continuous batching of hierarchical sums ----
 * sfx_job_submit queues a job and returns at once; a pool of worker threads (sfx_engine_start, started
 * with defaults by the first submit), each with its own HIP stream and device buffers, runs the queued
 * jobs in arrival order, so jobs that arrive while others run are taken as workers free up and their
 * device work overlaps.  Each refinement level reserves the HBM bound of its next cell list from the
 * engine's budget (default 80% of the device memory) and waits while the budget is held by others; when
 * every job that holds budget is itself waiting, the oldest waiter is admitted over the cap (no job waits
 * on another forever).  A job's result does not depend on what runs beside it.  sfx_engine_stop fails the
 * queued jobs with SFX_ECANCELED and waits for the running ones.  No counterpart in the reference, which
 * runs graphs one after another in one process (main.c:243). */
#define SFX_JOB_QUEUED  0
#define SFX_JOB_RUNNING 1
#define SFX_JOB_DONE    2
#define SFX_JOB_FAILED  3
int sfx_engine_start(int workers, double hbm_budget_bytes);      /* <= 0: 8 workers, 80% of HBM */
int sfx_engine_stop(void);
int sfx_job_submit(int graph, double h, int N, double eps, int top, const sfx_opts_t *opts, long *job);
int sfx_job_poll(long job, int *state);
/* blocks until the job has finished, at most the engine's wait deadline (default 3600 s; then
 * SFX_ETIMEOUT, the job keeps running); *t_queue = seconds from submit to start, *t_run = start to end */
int sfx_job_wait(long job, sfx_result_t *res, double *t_queue, double *t_run);
/* the same with a deadline of its own (seconds, <= 0: none) */
int sfx_job_wait_for(long job, double timeout_s, sfx_result_t *res, double *t_queue, double *t_run);
/* the deadlines of sfx_job_wait and of a job's wait for HBM budget (seconds; <= 0: none; defaults 3600
 * and 600): a wait that passes one fails with SFX_ETIMEOUT naming the job, its stage and level */
int sfx_engine_deadlines(double job_wait_s, double budget_wait_s);
int sfx_job_release(long job);

/* ---- measurement ----
 * With profiling on, every device kernel the library launches is bracketed by HIP events on the
 * library's stream, and the elapsed times are accumulated per (graph, kernel kind) together with the
 * points it processed (valid lattice points, and how many of them skipped the |.| sums). */
#define SFX_K_LEVEL_EVAL 0     /* refinement: children of the cells that compute the |.| sums, FP64 + tests */
#define SFX_K_LEVEL_DD   1     /* refinement: double-double recompute of the flagged children */
#define SFX_K_LEVEL_SUM  2     /* refinement: compaction, block sums, next-level cells */
#define SFX_K_NODES_EVAL 3     /* coarsest level: frontier nodes, FP64 + tests */
#define SFX_K_NODES_DD   4     /* coarsest level: double-double recompute */
#define SFX_K_LEVEL_EVAL_SKIP 5 /* refinement: children of the cells that skip the |.| sums, FP64 + tests */
#define SFX_K_LEVEL_SPLIT 6    /* refinement: split of a block's cells into the two classes */
typedef struct {
    char graph[32]; int kind;
    long launches; double ms; long points, skipped, dd_points;
} sfx_prof_entry_t;
int sfx_prof_enable(int on);           /* on: clears the table and starts recording; 0 stops */
int sfx_prof_count(void);
int sfx_prof_get(int i, sfx_prof_entry_t *out);
int sfx_sync(void);                    /* waits for the library's stream */
/* a plain device-to-device copy of `bytes` bytes by the kernel k_copy_probe (reads `bytes`, writes
 * `bytes`): the known-size reference that calibrates hardware traffic counters (FETCH_SIZE, WRITE_SIZE) */
int sfx_copy_probe(long bytes, double *ms);

#ifdef __cplusplus
}
#endif
#endif
