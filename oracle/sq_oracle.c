/* sq_oracle.c -- the CPU oracle (TEST INFRASTRUCTURE ONLY; see oracle.h for what it restates).
 * Compiled as C++ (the reference kernels' double-double type is a C++ struct), single-threaded,
 * FP contraction off.  Each function names the reference lines it follows. */
#include "oracle.h"
#include <dlfcn.h>
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXD 12                                   /* reference driver/sq.h:13 */
#define QEPS 4.93e-32                             /* common.h:7 (QREAL = dd) */

typedef struct { double hi, lo; } ddv;            /* layout of the reference's dd (dd.h:10-16) */
typedef double (*fd_t)(const double*, double*);
typedef ddv (*fq_t)(const ddv*, ddv*);
typedef struct { const char *name, *volkov; int L, D, weight, n, fixed; void *fd, *fq, *ptd, *ptq, *fv; } graph_t;   /* common.h:38 */

static void *LIB; static int (*SETAB)(int); static char PREFIX[8] = "G_";
static char ERR[256];
const char *orc_error(void){ return ERR; }

void orc_opts_default(orc_opts_t *o){
    o->abserr0 = 1e-15; o->abrel = 1e-3; o->nearf = 1e-3; o->ddslack = 1;   /* eval.c:7,17-18 */
    o->abskip = 1; o->dropunres = 1; o->keepbt = 0; o->btmargin = 2; o->maxpts = 1L << 18; o->grow = 1; o->abcheck = 0; o->growskip = 2;
}

/* prefix "G_": the kernels as the reference builds them (libsqk_ref.so); "Q_": built without FP
 * contraction (libsqk_nc.so), fd the double half and fq the double-double half */
int orc_open(const char *so, const char *prefix){
    snprintf(PREFIX, sizeof PREFIX, "%s", prefix && *prefix ? prefix : "G_");
    LIB = dlopen(so, RTLD_NOW | RTLD_LOCAL);          /* the loader keeps one handle per file: selecting again is cheap */
    if(!LIB){ snprintf(ERR, sizeof ERR, "dlopen %s: %s", so, dlerror()); return -1; }
    SETAB = (int(*)(int))dlsym(LIB, "sqk_set_ab");
    return 0;
}
static const graph_t *findg(const char *name){
    char sym[96]; snprintf(sym, sizeof sym, "%s%s", PREFIX, name);
    const graph_t *g = LIB ? (const graph_t*)dlsym(LIB, sym) : NULL;
    if(!g) snprintf(ERR, sizeof ERR, "no kernel %s", name);
    return g;
}
int orc_graph(const char *name, orc_graph_t *o){
    const graph_t *g = findg(name); if(!g) return -1;
    snprintf(o->volkov, sizeof o->volkov, "%s", g->volkov); o->L = g->L; o->D = g->D; o->weight = g->weight; o->has_fv = g->fv != NULL;
    return 0;
}
int orc_eval_fd(const char *name, long n, const double *sf, double *f, double *ab){
    const graph_t *g = findg(name); if(!g) return -1;
    for(long i = 0; i < n; i++) f[i] = ((fd_t)g->fd)(sf + i*g->D, &ab[i]);
    return 0;
}
int orc_eval_fq(const char *name, long n, const double *sf, double *out){
    const graph_t *g = findg(name); if(!g) return -1;
    ddv q[MAXD], ab;
    for(long i = 0; i < n; i++){
        for(int d = 0; d < g->D; d++){ q[d].hi = sf[i*g->D + d]; q[d].lo = 0; }
        ddv v = ((fq_t)g->fq)(q, &ab);
        out[4*i] = v.hi; out[4*i + 1] = v.lo; out[4*i + 2] = ab.hi; out[4*i + 3] = ab.lo;
    }
    return 0;
}

/* fq of a kernel built with QREAL = __float128 (libsqk_f128.so, prefix "F_"): out[4i..4i+3] = f and
 * Sum|terms| each split into two doubles (hi = the rounded value, lo = the rounded rest) */
typedef __float128 (*fq128_t)(const __float128*, __float128*);
int orc_eval_fq128(const char *name, long n, const double *sf, double *out){
    const graph_t *g = findg(name); if(!g) return -1;
    __float128 q[MAXD], ab;
    for(long i = 0; i < n; i++){
        for(int d = 0; d < g->D; d++) q[d] = sf[i*g->D + d];
        __float128 v = ((fq128_t)g->fq)(q, &ab);
        double vh = (double)v, ah = (double)ab;
        out[4*i] = vh; out[4*i + 1] = (double)(v - (__float128)vh);
        out[4*i + 2] = ah; out[4*i + 3] = (double)(ab - (__float128)ah);
    }
    return 0;
}

/* ---- the scale test (main.c:79-116) ---- */
/* dd.h:18-20,38-40: the product of two double-double values (fma two-product, then renormalised) */
static ddv dmul(ddv a, ddv b){
    double p = a.hi*b.hi, e = fma(a.hi, b.hi, -p);
    e += a.hi*b.lo + a.lo*b.hi;
    double s = p + e; ddv r = {s, e - (s - p)}; return r;
}
typedef ddv (*ptq_t)(const ddv*, ddv, ddv*);
int orc_scale(const char *name, double kap, unsigned seed, int E, orc_scale_t *out){
    const graph_t *g = findg(name); if(!g) return -1;
    if(!g->ptq){ snprintf(ERR, sizeof ERR, "%s has no point function", name); return -2; }
    int n = g->n, fixed = g->fixed, B = 2*E + 1;
    memset(out, 0, sizeof *out); out->n = n; out->fixed = fixed; out->worstd = -1;
    srand(seed);
    ddv x0[16];
    for(int i = 0; i < n; i++){ x0[i].hi = exp(4.0*rand()/RAND_MAX - 2.0); x0[i].lo = 0; }   /* main.c:82 */
    x0[fixed].hi = 1; x0[fixed].lo = 0;
    long ndir = 1; for(int i = 0; i < n - 1; i++) ndir *= B;
    out->ndir = ndir - 1;
    for(long d = 0; d < ndir; d++){
        int e[16], nz = 0; long r = d;
        for(int i = 0; i < n; i++){ if(i == fixed){ e[i] = 0; continue; } e[i] = (int)(r % B) - E; r /= B; nz += e[i] != 0; }
        if(!nz) continue;
        double gv[3];
        for(int q = 0; q < 3; q++){
            double kk = kap*(q + 1)/3.0; ddv x[16], meas = {1, 0};
            for(int i = 0; i < n; i++){
                ddv c = {exp(-kk*e[i]), 0};
                x[i] = dmul(x0[i], c);
                if(i != fixed) meas = dmul(meas, x[i]);
            }
            ddv ab, zero = {0, 0};
            ddv v = ((ptq_t)g->ptq)(x, zero, &ab);
            ddv vm = dmul(v, meas);
            gv[q] = fabs(vm.hi + vm.lo);
        }
        double ratio = gv[1] > 0 ? gv[2]/gv[1] : 0;
        if(ratio > out->worst){
            out->worst = ratio; out->worstd = d;
            for(int q = 0; q < 3; q++) out->g_worst[q] = gv[q];
            for(int i = 0; i < n; i++) out->e_worst[i] = e[i];
        }
        if(ratio > 0.3 && gv[2] > 1e-30){
            if(out->nslow < 12){
                for(int q = 0; q < 3; q++) out->g_slow[out->nslow][q] = gv[q];
                for(int i = 0; i < n; i++) out->e_slow[out->nslow][i] = e[i];
                out->nslow++;
            }
            out->bad++;
        }
    }
    return 0;
}

/* ---- state of one run ---- */
static const graph_t *G; static int D, K; static double H, ABSERR; static double *EX; static orc_opts_t O;
static long NQ, UNRESDROP, ABMISS; static double U;

/* lattice.c:30-34 */
static double lattice_node(const int *k, double *sf){
    double w = 1;
    for(int i = 0; i < D; i++){ sf[i] = EX[k[i] + K]; w *= sf[i]*H; }
    return w;
}
/* eval.c:25-30 */
static double qvalue(double vq, double abq, double w){
    double e = abq*w*QEPS;
    if(!(e*16 <= ABSERR)) U += isfinite(e) ? e*e : 0;
    double v = isfinite(vq) ? vq*w : 0;
    if(!(e <= fabs(v))){ UNRESDROP++; if(O.dropunres) return 0; }
    return v;
}
/* eval.c:33-60: leafx = lattice_node, fd with SQ_AB = !(skip && !ABCHECK), finish */
static double leafx(const int *k, int skip, int *near){
    double sf[MAXD], ab = 0; double w = lattice_node(k, sf);
    if(SETAB) SETAB(!(skip && !O.abcheck));
    double v = ((fd_t)G->fd)(sf, &ab);
    if(SETAB) SETAB(1);
    *near = 0;
    if(skip && !O.abcheck){ v *= w; return isfinite(v) ? v : 0; }
    double bnd = ab*w*16*DBL_EPSILON/O.ddslack;
    if(!(bnd <= O.nearf*ABSERR)) *near = 1;
    if(!(bnd <= ABSERR)){
        if(skip) ABMISS++;
        ddv q[MAXD], abq; for(int i = 0; i < D; i++){ q[i].hi = sf[i]; q[i].lo = 0; }
        ddv vq = ((fq_t)G->fq)(q, &abq); NQ++;
        return qvalue(vq.hi + vq.lo, abq.hi + abq.lo, w);
    }
    v *= w; return isfinite(v) ? v : 0;
}

/* sum.c:18-20 */
typedef struct { double s, c; } ksum_t;
static void kadd(ksum_t *k, double x){ double t = k->s + x; if(fabs(k->s) >= fabs(x)) k->c += (k->s - t) + x; else k->c += (x - t) + k->s; k->s = t; }
static double kval(const ksum_t *k){ return k->s + k->c; }

/* ---- points and hash sets ---- */
typedef struct { signed char k[MAXD]; } pt_t;              /* k[MAXD-1] = PFLAG (sq.h:70) */
#define PFLAG (MAXD - 1)
typedef struct { pt_t *p; long n, cap; } vec_t;
static void vpush(vec_t *v, const pt_t *p){ if(v->n == v->cap){ v->cap = v->cap ? 2*v->cap : 1024; v->p = (pt_t*)realloc(v->p, v->cap*sizeof(pt_t)); } v->p[v->n++] = *p; }

typedef struct { uint64_t *key; long *val; long mask, n; } hmap_t;
static void hinit(hmap_t *m, long want){ long c = 1024; while(c < 2*want) c <<= 1; m->key = (uint64_t*)calloc(c, 8); m->val = (long*)calloc(c, sizeof(long)); m->mask = c - 1; m->n = 0; }
static void hfree(hmap_t *m){ free(m->key); free(m->val); m->key = NULL; m->val = NULL; }
static long hpos(const hmap_t *m, uint64_t y){ return (long)((y*0x9E3779B97F4A7C15ULL) >> 21) & m->mask; }
static long *hget(const hmap_t *m, uint64_t y){ long i = hpos(m, y); while(m->key[i]){ if(m->key[i] == y) return &m->val[i]; i = (i + 1) & m->mask; } return NULL; }
static long *hput(hmap_t *m, uint64_t y, int *fresh){
    if(2*(m->n + 1) > m->mask + 1){
        hmap_t t; hinit(&t, m->mask + 1);
        for(long i = 0; i <= m->mask; i++) if(m->key[i]) *hput(&t, m->key[i], NULL) = m->val[i];
        hfree(m); *m = t;
    }
    long i = hpos(m, y);
    while(m->key[i] && m->key[i] != y) i = (i + 1) & m->mask;
    if(!m->key[i]){ m->key[i] = y; m->val[i] = 0; m->n++; if(fresh) *fresh = 1; } else if(fresh) *fresh = 0;
    return &m->val[i];
}

/* sumkey.h:13-18: key 1 c_0 .. c_{D-1} in base W */
static int W_; static uint64_t P_[MAXD]; static int ST;
static int keyradix(int st){
    ST = st; W_ = 2*(K/st) + 1; unsigned __int128 p = 1;
    for(int i = D - 1; i >= 0; i--){ P_[i] = (uint64_t)p; p *= W_; }
    return p < ((unsigned __int128)1 << 62);
}
static uint64_t key(const pt_t *q){ uint64_t y = 1; for(int i = 0; i < D; i++) y = y*W_ + (uint64_t)(q->k[i]/ST + K/ST); return y; }

/* sumkey.h:130-461.  At top = 0 (gs > 0, sumkey.h:165-168) a frontier point skips the |.| sums when the
 * point that reached it was checked and not near the fallback, or skipped fewer than gs generations after
 * such a point (the generation count in PFLAG, sumkey.h:367-368); *gshs, *gshn: the signed sum and count
 * of the evaluated points below the running threshold (the shell, sum.c:275-276) */
static long GSKIP;              /* frontier points that skipped the |.| sums (ABSKIPPED) */
static int grow(int lev, double eps_rel, vec_t *keep, double *S, double *A, double *e1, double *e2, double *drop, long *nev, long *btuns,
                double *gshs, long *gshn){
    int st = 1 << lev, gs = lev == 0 && O.abskip && D < MAXD ? O.growskip : 0;
    *gshs = 0; *gshn = 0;
    if(!keyradix(st)){ snprintf(ERR, sizeof ERR, "grid keys do not fit"); return -1; }
    double pw = pow(st, D), a = 0, x1 = 0, x2 = 0, dr = 0; ksum_t ks = {0, 0}; long ne = 0;
    vec_t fr = {0}, nx = {0}, pk = {0};
    pt_t z; memset(&z, 0, sizeof z); vpush(&fr, &z);
    double *kv = NULL; long nkv = 0, kvcap = 0;
    hmap_t bt; hinit(&bt, 1 << 16);            /* boundary point -> frontier (all tables; freed ones masked by BTLO) */
    int lay = 0; long BTLO = 0;
    while(fr.n){
        double *v = (double*)malloc(fr.n*sizeof(double)); unsigned char *nr = (unsigned char*)malloc(fr.n), *sk = (unsigned char*)malloc(fr.n);
        for(long i = 0; i < fr.n; i++){
            int k[MAXD], n_; for(int d = 0; d < D; d++) k[d] = fr.p[i].k[d];
            sk[i] = gs && fr.p[i].k[PFLAG] > 0; GSKIP += sk[i];
            v[i] = leafx(k, sk[i], &n_); nr[i] = (unsigned char)n_;
        }
        ne += fr.n;
        for(long i = 0; i < fr.n; i++){ kadd(&ks, v[i]*pw); a += fabs(v[i])*pw; }
        double thr = eps_rel*a;
        vec_t kept = {0}; long *kfi = (long*)malloc((fr.n ? fr.n : 1)*sizeof(long)); long nkf = 0;
        for(long i = 0; i < fr.n; i++){
            int mx = 0; for(int d = 0; d < D; d++){ int x = abs(fr.p[i].k[d]); if(x > mx) mx = x; }
            double w = fabs(v[i])*pw; int kp = fabs(v[i]) >= thr;
            if(mx == K) x1 += w; else if(mx == K - st) x2 += w;
            if(!kp){ dr += w; *gshs += v[i]*pw; (*gshn)++; *hput(&bt, key(&fr.p[i]), NULL) = lay; }
            else {
                kfi[nkf++] = i;
                pt_t q = fr.p[i]; q.k[PFLAG] = (signed char)(O.abskip ? !nr[i] : 0);
                vpush(&kept, &fr.p[i]); vpush(keep, &q);
                if(nkv == kvcap){ kvcap = kvcap ? 2*kvcap : 1024; kv = (double*)realloc(kv, kvcap*sizeof(double)); }
                kv[nkv++] = fabs(v[i]);
            }
        }
        /* next frontier: (kept point, axis, sign) order, first occurrence wins */
        hmap_t hs; hinit(&hs, pk.n + kept.n*2*D);
        for(long i = 0; i < pk.n; i++) *hput(&hs, key(&pk.p[i]), NULL) = 1;
        nx.n = 0;
        for(long j = 0; j < kept.n; j++){
            const pt_t *p = &kept.p[j]; uint64_t y0 = key(p);
            int n = 0; for(int d = 0; d < D; d++) n += abs(p->k[d]);
            n /= st;
            for(int d = 0; d < D; d++) for(int sg = -1; sg <= 1; sg += 2){
                int kk = p->k[d] + st*sg;
                if(kk > K || kk < -K) continue;
                uint64_t y = sg > 0 ? y0 + P_[d] : y0 - P_[d];
                int n1 = abs(kk) > abs(p->k[d]) ? n + 1 : n - 1;
                if(n1 < BTLO) (*btuns)++;
                int hit = 0;
                for(int l = n1; l < lay && !hit; l += 2){ long *b = l >= BTLO ? hget(&bt, y) : NULL; hit = b && *b == l; }
                if(hit) continue;
                int fresh; hput(&hs, y, &fresh);
                if(!fresh) continue;
                pt_t c = *p; c.k[d] = (signed char)kk; c.k[PFLAG] = 0;
                if(gs){ long pi = kfi[j]; int pg = fr.p[pi].k[PFLAG]; c.k[PFLAG] = (signed char)(!sk[pi] ? !nr[pi] : (pg < gs ? pg + 1 : 0)); }
                vpush(&nx, &c);
            }
        }
        hfree(&hs);
        free(pk.p); pk = kept;
        if(nx.n && !O.keepbt){            /* sumkey.h:446-453 */
            long mn = 1L << 40;
            for(long i = 0; i < nx.n; i++){ long t = 0; for(int d = 0; d < D; d++) t += abs(nx.p[i].k[d]); t /= st; if(t < mn) mn = t; }
            mn -= O.btmargin;
            if(mn - 1 > BTLO) BTLO = mn - 1 < lay + 1 ? mn - 1 : lay + 1;
        }
        free(v); free(nr); free(sk); free(kfi);
        vec_t t = fr; fr = nx; nx = t; lay++;
    }
    double thr = eps_rel*a; long m = 0;                       /* sumkey.h:397-401 */
    for(long i = 0; i < keep->n; i++){ if(kv[i] >= thr){ kv[m] = kv[i]; keep->p[m] = keep->p[i]; m++; } else dr += kv[i]*pw; }
    keep->n = m;
    if(D < MAXD && m && lev > 0){                             /* sumkey.h:402-453 shadow tiers (none at top = 0) */
        hmap_t ix; hinit(&ix, m);
        for(long i = 0; i < m; i++) *hput(&ix, key(&keep->p[i]), NULL) = i;
        unsigned char *tr = (unsigned char*)calloc(m, 1); long *cu = (long*)malloc(m*sizeof(long)), *nw = (long*)malloc(m*sizeof(long));
        for(int t = 1; t <= 2; t++){
            double lim = (t == 1 ? 10 : 100)*thr; long nc = 0;
            for(long i = 0; i < m; i++){ int zz = 1; for(int d = 0; d < D; d++) if(keep->p[i].k[d]){ zz = 0; break; }
                if(zz && kv[i] >= lim){ tr[i] = (unsigned char)t; cu[nc++] = i; } }
            while(nc){
                long nn = 0;
                for(long c = 0; c < nc; c++){
                    long i = cu[c]; uint64_t y0 = key(&keep->p[i]);
                    for(int d = 0; d < D; d++) for(int sg = -1; sg <= 1; sg += 2){
                        int kk = keep->p[i].k[d] + st*sg; if(kk > K || kk < -K) continue;
                        long *q = hget(&ix, sg > 0 ? y0 + P_[d] : y0 - P_[d]);
                        if(q && kv[*q] >= lim && tr[*q] < t){ tr[*q] = (unsigned char)t; nw[nn++] = *q; }
                    }
                }
                long *x = cu; cu = nw; nw = x; nc = nn;
            }
        }
        for(long i = 0; i < m; i++) keep->p[i].k[PFLAG] = (signed char)((keep->p[i].k[PFLAG] & 1) | (tr[i] << 1));
        free(tr); free(cu); free(nw); hfree(&ix);
    }
    hfree(&bt); free(fr.p); free(nx.p); free(pk.p); free(kv);
    *S = kval(&ks); *A = a; *e1 = x1; *e2 = x2; *drop = dr; *nev += ne;
    return 0;
}

/* sum.c:166-191: every node of the step-2^lev box, id = sum_i (k_i/st + K/st) W^i (axis 0 fastest), in id
 * order with plain sums -- the reference's OpenMP reduction run on one thread; the kept nodes (|v| >= thr)
 * with their flag byte 0, in id order */
static void scan(int lev, double thr, vec_t *keep, double *S, double *A, double *e1, double *e2, double *drop, long *nev){
    int st = 1 << lev, W = 2*(K/st) + 1; long n = 1; for(int i = 0; i < D; i++) n *= W;
    double pw = pow(st, D), s = 0, a = 0, x1 = 0, x2 = 0, dr = 0;
    for(long id = 0; id < n; id++){
        int k[MAXD], mx = 0; long r = id;
        for(int i = 0; i < D; i++){ k[i] = st*((int)(r % W) - K/st); r /= W; if(abs(k[i]) > mx) mx = abs(k[i]); }
        int nr; double v = leafx(k, 0, &nr);
        s += v*pw; a += fabs(v)*pw;
        if(mx == K) x1 += fabs(v)*pw; else if(mx == K - st) x2 += fabs(v)*pw;
        if(keep){
            if(fabs(v) >= thr){ pt_t c; memset(&c, 0, sizeof c); for(int i = 0; i < D; i++) c.k[i] = (signed char)k[i]; vpush(keep, &c); }
            else dr += fabs(v)*pw;
        }
    }
    *S = s; *A = a; *e1 = x1; *e2 = x2; *drop = dr; *nev += n;
}

/* sumkey.h:20-90: neighbours merged (skip AND, tier max), sorted by key */
static int cmpu64(const void *a, const void *b){ uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b; return x < y ? -1 : x > y; }
static int dilate(vec_t *cur, int st){
    if(!keyradix(st)) return -1;
    hmap_t m; hinit(&m, cur->n*(2*D + 1));
    for(long q = 0; q < cur->n; q++){
        const pt_t *p = &cur->p[q]; int fl = p->k[PFLAG], sk = fl & 1, tr = fl >> 1; uint64_t y0 = key(p);
        for(int j = -1; j < D; j++) for(int sg = -1; sg <= 1; sg += 2){
            if(j < 0 && sg > 0) continue;
            uint64_t y = y0;
            if(j >= 0){ int kk = p->k[j] + st*sg; if(kk > K || kk < -K) continue; y = sg > 0 ? y + P_[j] : y - P_[j]; }
            int fresh; long *e = hput(&m, y, &fresh);
            int esk = fresh ? 1 : (int)(*e & 1), etr = fresh ? 0 : (int)(*e >> 1);
            if(!sk) esk = 0; if(tr > etr) etr = tr;
            *e = esk | (etr << 1);
        }
    }
    uint64_t *u = (uint64_t*)malloc((m.n ? m.n : 1)*8); long nu = 0;
    for(long i = 0; i <= m.mask; i++) if(m.key[i]) u[nu++] = m.key[i];
    qsort(u, nu, 8, cmpu64);
    cur->n = 0;
    for(long j = 0; j < nu; j++){
        pt_t c; memset(&c, 0, sizeof c); uint64_t y = u[j];
        for(int i = D - 1; i >= 0; i--){ c.k[i] = (signed char)(((int)(y % W_) - K/st)*st); y /= W_; }
        c.k[PFLAG] = (signed char)*hget(&m, u[j]);
        vpush(cur, &c);
    }
    free(u); hfree(&m);
    return 0;
}

/* sum.c:346-359 */
static int box_zero(int lev, long *nev){
    int st = 1 << lev, W = 2*(K/st) + 1, nz = 0;
    for(long i = 0; i <= 4096; i++){
        int k[MAXD] = {0}; unsigned long long y = 0x9E3779B97F4A7C15ULL*(unsigned long long)(i + 1);
        if(i) for(int d = 0; d < D; d++){ y ^= y >> 29; y *= 0xBF58476D1CE4E5B9ULL; y ^= y >> 32; k[d] = ((int)(y % (unsigned long long)W) - K/st)*st; }
        long q = NQ; double u = U; int nr;
        if(leafx(k, 0, &nr) != 0) nz++;
        NQ = q; U = u;              /* the probe's fallback count and round term are local in the reference (UNRESDROP is not) */
    }
    if(nz) return 0;
    *nev += 4097;
    return 1;
}

#define TIER(f, t) (((f) >= 10*(t)) + ((f) >= 100*(t)))

/* sum.c:556-814 with SYMON = 0, PRED = 0, no dump/hist; the refinement in blocks of maxpts >> D cells,
 * each block summed in child order with Neumaier sums, blocks combined by kblock */
int orc_hsum(const char *name, double h, int N, double eps, int top, const orc_opts_t *o, orc_res_t *R){
    G = findg(name); if(!G) return -1;
    if(top < 0){ snprintf(ERR, sizeof ERR, "top >= 0 only"); return -1; }
    orc_opts_t dflt; orc_opts_default(&dflt); O = o ? *o : dflt;
    D = G->D; H = h;
    int m = 1 << (top + 1); K = (N - 1)/2; K = ((K + m - 1)/m)*m;            /* lattice.c:16 */
    if(K > 127){ snprintf(ERR, sizeof ERR, "K > 127"); return -1; }
    free(EX); EX = (double*)malloc((2*K + 1)*sizeof(double));
    for(int k = -K; k <= K; k++) EX[k + K] = exp(k*H);
    memset(R, 0, sizeof *R); R->K = K; R->top = top;
    NQ = 0; UNRESDROP = 0; U = 0; ABMISS = 0; GSKIP = 0;
    ABSERR = O.abserr0;
    if(box_zero(top, &R->nev)){ R->zero = 1; return 0; }
    vec_t cur = {0}; double S, A, e1, e2, dr, gshs = 0; long gshn = 0;
    long nev0 = R->nev;
    if(O.grow){ if(grow(top, eps*1e-3, &cur, &S, &A, &e1, &e2, &dr, &R->nev, &R->btuns, &gshs, &gshn)) return -1; R->ref = A; }
    else {
        /* SINCQED_GROW=0 (sum.c:610-617): the reference scale from the box at step 2^(top+1) h, then the
         * box at step 2^top h with the keep threshold (its own A only decides "identically zero" below);
         * no shadow tiers */
        scan(top + 1, 0, NULL, &S, &A, &e1, &e2, &dr, &R->nev);
        R->ref = A;
        scan(top, eps*R->ref*1e-3, &cur, &S, &A, &e1, &e2, &dr, &R->nev);
    }
    R->nev_level[top] = R->nev - nev0;
    R->nq = NQ; R->prune = dr; R->unres = U; R->abskipped = GSKIP;
    /* top = 0 (sum.c:619-626): every evaluated point is in the sum; the prune column is the signed sum of
     * the never-expanded shell */
    if(O.grow && top == 0 && gshn) R->prune = fabs(gshs);
    if(A == 0){ free(cur.p); R->unresdrop = UNRESDROP; R->abmiss = ABMISS; return 0; }
    { double r = e2 > 0 ? e1/e2 : 0; R->trunc = (r > 0 && r < 1) ? e1*r/(1 - r) : e1; }
    double Slev[16]; int nl = 0; Slev[nl++] = S;
    if(top > 0 && dilate(&cur, 1 << top)) return -1;
    R->kept_top = cur.n;
    ABSERR = fmax(O.abserr0, O.abrel*eps*R->ref);                            /* sum.c:651 */
    R->abserr = ABSERR;
    double ES1 = 0, ES2 = 0, KC[3] = {0, 0, 0};
    for(int lev = top - 1; lev >= 0; lev--){
        int st = 1 << lev, last = lev == 0; double pw = pow(st, D), thr = eps*R->ref*1e-2;
        long B = O.maxpts >> D; if(B < 1) B = 1;
        double s = 0, a = 0, drop = 0; long ne = 0; double usave = U; U = 0;
        vec_t nxt = {0};
        int k[MAXD];
        for(long c0 = 0; c0 < cur.n; c0 += B){
            long c1 = c0 + B < cur.n ? c0 + B : cur.n;
            ksum_t bs = {0, 0}, b1 = {0, 0}, b2 = {0, 0};
            for(long c = c0; c < c1; c++){                                    /* prep_cell, sum.c:379-422 */
                const pt_t *pp = &cur.p[c]; int kp[MAXD]; for(int d = 0; d < D; d++) kp[d] = pp->k[d];
                int psk = O.abskip && (pp->k[PFLAG] & 1), te = pp->k[PFLAG] >> 1;
                for(int b = 0; b < (1 << D); b++){
                    int ok = 1;
                    for(int d = 0; d < D; d++){ k[d] = kp[d] - (((b >> d) & 1) ? st : 0); if(k[d] < -K) ok = 0; }
                    if(!ok) continue;
                    int nr; double v = leafx(k, psk, &nr); ne++;
                    kadd(&bs, v*pw); a += fabs(v); R->abskipped += psk;          /* sum.c:730-746 */
                    if(last){ if(te >= 1) kadd(&b1, v*pw); if(te >= 2) kadd(&b2, v*pw); }
                    else if(fabs(v) >= thr){
                        pt_t c2; memset(&c2, 0, sizeof c2); for(int d = 0; d < D; d++) c2.k[d] = (signed char)k[d];
                        int tv = TIER(fabs(v), thr);
                        c2.k[PFLAG] = (signed char)((O.abskip && (psk || !nr)) | ((te < tv ? te : tv) << 1));
                        vpush(&nxt, &c2);
                    } else drop += fabs(v)*pw;
                }
            }
            ksum_t t;                                                          /* kblock, sum.c:77-81 */
            t.s = s; t.c = KC[0]; kadd(&t, kval(&bs)); s = t.s; KC[0] = t.c;
            t.s = ES1; t.c = KC[1]; kadd(&t, kval(&b1)); ES1 = t.s; KC[1] = t.c;
            t.s = ES2; t.c = KC[2]; kadd(&t, kval(&b2)); ES2 = t.s; KC[2] = t.c;
        }
        s += KC[0]; ES1 += KC[1]; ES2 += KC[2]; KC[0] = KC[1] = KC[2] = 0;     /* kflush */
        R->prune += drop; R->nev += ne; R->nev_level[lev] = ne;
        if(last){ R->unres = sqrt(U); R->Ah = a; }
        U += usave;
        Slev[nl++] = s;
        free(cur.p); cur = nxt;
    }
    free(cur.p);
    R->nq = NQ; R->unresdrop = UNRESDROP; R->abmiss = ABMISS;
    R->S[0] = Slev[nl - 1]; R->S[1] = nl > 1 ? Slev[nl - 2] : NAN; R->S[2] = nl > 2 ? Slev[nl - 3] : NAN;
    R->es1 = ES1; R->es2 = ES2;
    return 0;
}
