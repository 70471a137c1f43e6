/* oracle.h -- the CPU oracle: a restatement of the reference's hierarchical Sinc sum (TEST
 * INFRASTRUCTURE ONLY -- tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it as the
 * checker; the product library never links it).
 *
 * It restates, single-threaded and in plain C, reasther/sincFeyn v0.9.0:
 *   driver/sum.c:556-814    hsum() with the symmetry fold, predictor, dump and histograms off (top >= 0,
 *                           ABCHECK, the grow's skip generations at top = 0)
 *   driver/sumkey.h:20-461  grow() and dilate() (64-bit grid keys)
 *   driver/sum.c:166-191,610-617 scan(): the full-box scans of SINCQED_GROW=0 (plain sums in node order,
 *                           the reference's OpenMP reduction on one thread)
 *   driver/sum.c:346-359    box_zero()
 *   driver/eval.c:25-95     finish(), qvalue(), leafx(), eval_batch() (fd per point; fd == fv bitwise)
 *   driver/sum.c:18-20,76-81 Neumaier sums kadd/kval/kblock/kflush
 *   driver/lattice.c:13-34  the EX table and lattice_node()
 *   driver/main.c:46-48     value and error of a graph
 *   driver/main.c:79-116    scale_test() (the integrand's point function in double-double)
 * The integrand comes from the reference's own generated kernels, compiled where they lie into
 * oracle/_ref/libsqk_ref.so (graph_t descriptors G_<name>: fd in double, fq in double-double).
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double abserr0, abrel, nearf, ddslack;
    int abskip, dropunres, keepbt, btmargin;
    long maxpts;          /* nodes per refinement block: the reference CPU build's 2^18 (sum.c:643) */
    int grow;             /* SINCQED_GROW: 1 the support-following grow, 0 full-box scans (sum.c:610-617) */
    int abcheck;          /* SINCQED_ABCHECK (eval.c:11-12) */
    int growskip;         /* SINCQED_GROWSKIP: skip generations of the grow at top = 0 (sum.c:272-275), default 2 */
} orc_opts_t;

typedef struct {
    double S[3], Ah, ref, prune, trunc, unres;
    long nev, nq, kept_top;
    double es1, es2;
    long abskipped, unresdrop, btuns;
    long nev_level[16];
    int zero, K, top;
    double abserr;
    long abmiss;          /* ABMISS: with abcheck, skip-class points that needed the fallback */
} orc_res_t;

typedef struct { char volkov[64]; int L, D, weight, has_fv; } orc_graph_t;

void orc_opts_default(orc_opts_t *o);
int  orc_open(const char *kernels_so, const char *prefix);   /* 0 on success; prefix "G_" or "Q_" */
int  orc_graph(const char *name, orc_graph_t *out);
/* the reference kernel at Sinc variables sf[i*D + d]: fd (double half) -> f[i], ab[i] */
int  orc_eval_fd(const char *name, long n, const double *sf, double *f, double *ab);
/* fq (double-double half) -> out[4i..4i+3] = f.hi, f.lo, ab.hi, ab.lo */
int  orc_eval_fq(const char *name, long n, const double *sf, double *out);
/* fq of the __float128 build (libsqk_f128.so, prefix "F_"), each value split into two doubles */
int  orc_eval_fq128(const char *name, long n, const double *sf, double *out);
int  orc_hsum(const char *name, double h, int N, double eps, int top, const orc_opts_t *o, orc_res_t *r);
/* the scale test, driver/main.c:79-116 (directions in index order, one thread); -2 for a graph without
 * a point function (VP) */
typedef struct {
    int n, fixed; long ndir; int bad; double worst; long worstd;
    int e_worst[16]; double g_worst[3];
    int nslow; int e_slow[12][16]; double g_slow[12][3];
} orc_scale_t;
int  orc_scale(const char *name, double kappa, unsigned seed, int E, orc_scale_t *out);
const char *orc_error(void);

#ifdef __cplusplus
}
#endif
