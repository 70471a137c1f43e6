/* sincqed.c -- the command-line host of the MI355X evaluator: `sincqed list`,
 * `sincqed run h N [eps] [top] GRAPH...`, `sincqed scale GRAPH [kappa] [seed]` and
 * `sincqed compare R1 R2`, with the reference's arguments, environment, log lines and output formats
 * (reasther/sincFeyn driver/main.c:40-251), every graph's hierarchical sum and every integrand value
 * computed on the GPU through the C ABI of include/sfx.h.
 *
 * As in the reference's GPU build, the symmetry fold is not applied (sum.c:571): SINCQED_SYM is ignored
 * with a note.  SINCQED_GPU=d picks the device (gpu.cu:30-34).  A run of several graphs submits them all
 * to the library's job engine (SINCQED_WORKERS workers, default 16) and reports them in order as they
 * finish, with the results and lines of a run one after another; SINCQED_BATCH=0 runs them one at a time.
 */
#define _POSIX_C_SOURCE 200809L      /* setenv */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "sfx.h"

static FILE *LOGF, *RESF;

static void say(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
#include <stdarg.h>
static void say(const char *fmt, ...){
    va_list ap; va_start(ap, fmt); vprintf(fmt, ap); va_end(ap); fflush(stdout);
    if(LOGF){ va_start(ap, fmt); vfprintf(LOGF, fmt, ap); va_end(ap); fflush(LOGF); }
}

/* the reference values of graphs.c:36-46 (Volkov string -> value), for the ref and diff columns */
typedef struct { const char *volkov; double val; const char *src; } ref_t;
static const ref_t REFS[] = {
    {"3; 1-4, 2-5", -0.46764544609361365, "exact 1/6+13pi^2/36-5pi^2ln2/6+5z3/4 (crossed)"},
    {"3; 1-5, 2-4",  0.77747802228274215, "exact 11/48+pi^2/18 (ladder)"},
    {"2; 1-4, 3-5", -0.064020941344682816, "exact -55/24+pi^2/18+pi^2ln2/3-z3/2 (corner, PSLQ)"},
    {"2; 1-5, 3-4", -0.58997802228274215, "exact -1/24-pi^2/18 (SE, PSLQ)"},
    {"3; 1-6, 2-5, 4-7", -0.334695, "analytic, 6 digits"},
    {"VP: 1-loop vertex + Pi2", 0.015687421859103, "exact 119/36-pi^2/3"},
    {"VP: 2; 1-3 +Pi2(1-3)+Pi2(1-3)+Pi2(1-3)", 0.000876865858890, "1D integral (mpmath)"},
    {"LbL (6 orderings)", 0.3710052921, "analytic (electron loop)"},
    {NULL, 0, NULL}
};
static const ref_t *findref(const char *v){ for(int i = 0; REFS[i].volkov; i++) if(!strcmp(REFS[i].volkov, v)) return &REFS[i]; return NULL; }

static int isnum(const char *s){ char *e; strtod(s, &e); return e != s && *e == 0; }
static double envd(const char *n, double d){ const char *e = getenv(n); return e ? atof(e) : d; }
static int envi(const char *n, int d){ const char *e = getenv(n); return e ? atoi(e) : d; }

/* one graph's log lines and results line (main.c:40-75), from its finished sum; rc the sum's return code */
static int report_graph(int gi, double h, double eps, int top, const sfx_opts_t *o, int rc, const sfx_result_t *res,
                        double *tot, double *toterr, double *toterr2){
    sfx_graph_info_t g; sfx_graph_info(gi, &g);
    if(rc){ fprintf(stderr, "sincqed: %s: %s\n", g.name, sfx_last_error()); return rc; }
    const sfx_result_t R = *res;
    if(R.resumed)                                          /* sum.c:602 */
        say("# resuming %s from checkpoint: refine to step %dh, cell %ld of %ld\n", g.name, 1 << R.resume_lev,
            R.resume_next, R.resume_ncells);
    if(R.zero) say("# %s: integrand is exactly zero at %d probe nodes; reported as zero without a scan\n", g.name, 4097);
    if(top == 0 && !o->box_scan && R.shell_n)              /* sum.c:619-626 */
        say("# %s: grow shell (evaluated and summed, not expanded): %ld points, sum %+.3e, sum|.| %.3e; prune = |shell sum|, an estimate of what lies beyond it\n",
            g.name, R.shell_n, g.weight*R.shell_sum, g.weight*R.shell_abs);
    double w = g.weight, v = w*R.S[0], d2 = fabs(R.S[1] - R.S[0]);
    double disc = R.Ah > 0 ? 2*d2*d2/R.Ah : 0;
    double err = w*(disc + R.prune + R.trunc + R.unres);
    char lv[64] = "";
    if(isfinite(R.S[1])) snprintf(lv, sizeof lv, "  |S2h-Sh| %.1e", w*d2);
    if(isfinite(R.S[2])) snprintf(lv + strlen(lv), sizeof lv - strlen(lv), "  S_4h %+.6f", w*R.S[2]);
    say("%-7s %-22s D=%d w=%d  %+.12f  err %.2e (disc %.1e prune %.1e trunc %.1e round %.1e)%s  evals %ld quad %ld kept %ld  %.1fs",
        g.name, g.volkov, g.D, g.weight, v, err, w*disc, w*R.prune, w*R.trunc, w*R.unres, lv, R.nev, R.nq, R.kept_top, R.t);
    const ref_t *r = findref(g.volkov);
    if(r) say("  ref %+.12f (%s) diff %+.2e", r->val, r->src, v - r->val);
    say("\n");
    say("# %s: fallback when eps*sum|terms| > %.1e per point (max of G2_ABSERR and SINCQED_ABREL*eps*Ref; SINCQED_DDSLACK %g)\n",
        g.name, R.abserr, o->ddslack);
    if(o->abskip) say("# %s: %ld of %ld evaluations skipped the |.| sums (no fallback check)\n", g.name, R.abskipped, R.nev);
    if(o->abcheck) say("# %s: SINCQED_ABCHECK: %ld of the skipped points would have needed the fallback\n", g.name, R.abmiss);   /* main.c:65 */
    if(R.unresdrop)
        say("# WARNING %s: %ld points unresolved even in double-double (eps*sum|terms| > |value|, not one digit left) %s; their bound is in the round column\n",
            g.name, R.unresdrop, o->dropunres ? "counted as 0 (SINCQED_DROPUNRES=0 keeps them)" : "kept as they came out");
    if(getenv("SINCQED_TIMING") && !R.zero)                 /* sum.c:638: the stages' wall times */
        say("# %s timing: top level %.3fs, dilate+sort %.3fs, refinement %.3fs\n", g.name, R.t_grow, R.t_dilate, R.t_refine);
    if(R.esok && !R.zero && R.levels > 1){                 /* sum.c:806-809 */
        double d1 = w*(R.S[0] - R.es1), dd2 = w*(R.es1 - R.es2), rr = d1 != 0 ? dd2/d1 : 0;
        say("# %s eps extrapolation: eps*10 would change the value by %+.3e, eps*100 by a further %+.3e (ratio %.2f)", g.name, -d1, -dd2, rr);
        if(rr > 1.5) say("; pruning loss left at eps ~ %+.2e (add to the value)\n", d1/(rr - 1)); else say("\n");
    }
    if(RESF){                                              /* main.c:66-73 */
        fprintf(RESF, "%-8s %-28s %d %d %d %5.3f %4d %8.1e %d %+.15e %.3e %.2e %.2e %.2e %.2e ",
                g.name, g.volkov, g.L, g.D, g.weight, h, 2*R.K + 1, eps, top, v, err, w*disc, w*R.prune, w*R.trunc, w*R.unres);
        if(isfinite(R.S[1])) fprintf(RESF, "%+.8e", w*R.S[1]); else fprintf(RESF, "%15s", "-");
        fprintf(RESF, " %12ld %9.1f", R.nev, R.t);
        if(r) fprintf(RESF, " %+.15e %+.3e", r->val, v - r->val); else fprintf(RESF, " %22s %10s", "-", "-");
        fprintf(RESF, "\n"); fflush(RESF);
    }
    *tot += v; *toterr += err; *toterr2 += err*err;
    return 0;
}

static int run_graph(int gi, double h, int N, double eps, int top, const sfx_opts_t *o, int idx, int ng,
                     double *tot, double *toterr, double *toterr2){
    sfx_graph_info_t g; sfx_graph_info(gi, &g);
    say("  graph %d/%d %s (%s, D=%d) started on the GPU\n", idx, ng, g.name, g.volkov, g.D);
    sfx_result_t R;
    int rc = sfx_hsum(gi, h, N, eps, top, o, &R);
    return report_graph(gi, h, eps, top, o, rc, &R, tot, toterr, toterr2);
}

/* several graphs at once through the job engine (SINCQED_BATCH, default on): every graph is submitted
 * first, then the results are reported in the reference's order (main.c:243) as they finish.  Each
 * result is the one the graph gives alone; the log and results lines are those of one after another. */
static int run_batched(const int *sel, int ns, double h, int N, double eps, int top, const sfx_opts_t *o, int workers,
                       double *tot, double *toterr, double *toterr2){
    int rc = sfx_engine_start(workers, 0);
    if(rc){ fprintf(stderr, "sincqed: %s\n", sfx_last_error()); return rc; }
    long *job = (long*)malloc(ns*sizeof(long));
    for(int j = 0; j < ns; j++)
        if((rc = sfx_job_submit(sel[j], h, N, eps, top, o, &job[j]))){ fprintf(stderr, "sincqed: %s\n", sfx_last_error()); free(job); return rc; }
    int first = 0;
    for(int j = 0; j < ns; j++){
        sfx_graph_info_t g; sfx_graph_info(sel[j], &g);
        say("  graph %d/%d %s (%s, D=%d) started on the GPU\n", j + 1, ns, g.name, g.volkov, g.D);
        sfx_result_t R; double tq, tr;
        int r = sfx_job_wait(job[j], &R, &tq, &tr);
        if(!first) first = report_graph(sel[j], h, eps, top, o, r, &R, tot, toterr, toterr2);
        sfx_job_release(job[j]);
    }
    free(job);
    return first;
}

/* ---- compare (main.c:118-153): the error of each graph from two results files at different h ---- */
typedef struct { char name[32]; double h, v, rnd; } cres_t;
/* a results line holds the Volkov string (with blanks) before the numbers, so the value column is found
 * as the first signed %+.15e token; h is 4 columns before it and the round column 5 after (the format of
 * run_graph below and main.c:66-73); a later line of the same graph replaces an earlier one */
static int read_results(const char *fn, cres_t *r, int max){
    FILE *f = fopen(fn, "r"); if(!f){ perror(fn); return -1; }
    char line[4096]; int n = 0;
    while(fgets(line, sizeof line, f)){
        if(line[0] == '#') continue;
        char *tok[64]; int nt = 0;
        for(char *t = strtok(line, " \t\n"); t && nt < 64; t = strtok(NULL, " \t\n")) tok[nt++] = t;
        int iv = -1;
        for(int i = 1; i < nt && iv < 0; i++)
            if((tok[i][0] == '+' || tok[i][0] == '-') && strlen(tok[i]) >= 20 && strchr(tok[i], 'e')) iv = i;
        if(iv < 5 || iv + 5 >= nt) continue;
        int j = 0; while(j < n && strcmp(r[j].name, tok[0])) j++;
        if(j == n){ if(n == max) continue; n++; }
        snprintf(r[j].name, sizeof r[j].name, "%s", tok[0]);
        r[j].h = atof(tok[iv - 4]); r[j].v = atof(tok[iv]); r[j].rnd = atof(tok[iv + 5]);
    }
    fclose(f);
    return n;
}
static int compare(const char *fa, const char *fb){
    static cres_t A[2048], B[2048];
    int na = read_results(fa, A, 2048), nb = read_results(fb, B, 2048);
    if(na < 0 || nb < 0) return 1;
    printf("# compare %s and %s: value from the smaller step h; error = |S(h) - S(h')|/3 + round\n", fa, fb);
    printf("# (calibrated on graph 16, LbL and the Pi2 Pi2 chain with h = 1.0, h' = 1.2: true error <= |dS|/3)\n");
    printf("# %-8s %6s %6s  %22s  %10s  %10s\n", "name", "h", "h'", "value", "dS", "error");
    double tot = 0, e1 = 0, e2 = 0; int m = 0;
    for(int i = 0; i < na; i++){
        int j = 0; while(j < nb && strcmp(A[i].name, B[j].name)) j++;
        if(j == nb) continue;
        const cres_t *fine = A[i].h <= B[j].h ? &A[i] : &B[j], *coarse = A[i].h <= B[j].h ? &B[j] : &A[i];
        if(fine->h == coarse->h){ printf("  %-8s same h in both files, skipped\n", fine->name); continue; }
        double d = fine->v - coarse->v, e = fabs(d)/3 + fine->rnd;
        printf("  %-8s %6.3f %6.3f  %+22.15e  %+10.2e  %10.2e\n", fine->name, fine->h, coarse->h, fine->v, d, e);
        tot += fine->v; e1 += e; e2 += e*e; m++;
    }
    printf("# SUM %d graphs  %+.12f  err %.2e (in quadrature)  %.2e (linear)\n", m, tot, sqrt(e2), e1);
    return 0;
}

/* ---- scale (main.c:77-116): decay of integrand x measure along every scaling direction ---- */
static int scale(const char *sel, double kap, unsigned seed, int E){
    int ng = sfx_graph_count(), rc = sfx_init(envi("SINCQED_GPU", 0));
    if(rc){ fprintf(stderr, "sincqed: %s\n", sfx_last_error()); return 3; }
    for(int i = 0; i < ng; i++){
        if(!sfx_graph_selected(i, sel)) continue;
        sfx_graph_info_t g; sfx_graph_info(i, &g);
        static sfx_scale_result_t S;
        rc = sfx_scale_test(i, kap, seed, E, &S);
        if(rc == SFX_EUNSUP){ printf("%-6s %-22s (no scale test for VP graphs)\n", g.name, g.volkov); continue; }
        if(rc){ fprintf(stderr, "sincqed: %s: %s\n", g.name, sfx_last_error()); return 4; }
        for(int k = 0; k < S.nslow; k++){
            printf("  slow direction e=(");
            for(int j = 0; j < S.n; j++) if(j != S.fixed) printf("%+d", S.e_slow[k][j]);
            printf(")  |g| at kappa/3,2/3,1: %.3e %.3e %.3e\n", S.g_slow[k][0], S.g_slow[k][1], S.g_slow[k][2]);
        }
        printf("%-6s %-22s scale test kappa=%g: %ld directions, %d slow (ratio>0.3), worst ratio %.3g (dir %ld: e=(",
               g.name, g.volkov, kap, S.ndir, S.bad, S.worst, S.worstd);
        for(int j = 0; j < S.n; j++) if(j != S.fixed) printf("%+d", S.e_worst[j]);
        printf(") |g| %.2e %.2e %.2e)\n", S.g_worst[0], S.g_worst[1], S.g_worst[2]);
    }
    sfx_shutdown();
    return 0;
}

int main(int argc, char **argv){
    sfx_opts_t o; sfx_opts_default(&o);
    o.abserr0 = envd("G2_ABSERR", o.abserr0); o.abrel = envd("SINCQED_ABREL", o.abrel);
    o.abskip = envi("SINCQED_ABSKIP", o.abskip); o.nearf = envd("SINCQED_NEARF", o.nearf);
    o.ddslack = envd("SINCQED_DDSLACK", o.ddslack); o.dropunres = envi("SINCQED_DROPUNRES", o.dropunres);
    o.keepbt = getenv("SINCQED_KEEPBT") != NULL; o.btmargin = envi("SINCQED_BTMARGIN", o.btmargin);
    o.box_scan = envi("SINCQED_GROW", 1) == 0;            /* main.c:183: full-box scans instead of the grow */
    o.abcheck = envi("SINCQED_ABCHECK", 0);                 /* main.c:179 */
    o.growskip = envi("SINCQED_GROWSKIP", o.growskip);      /* sum.c:334 */
    o.graph_capture = envi("SFX_GRAPH_CAPTURE", 0);        /* no counterpart: one HIP graph per refinement level */
    o.ckpt_every = envd("SINCQED_CKPT", 600);              /* checkpoint.c:9, main.c:184 */
    { const char *d = getenv("SINCQED_CKPT_DIR"); if(d) snprintf(o.ckpt_dir, sizeof o.ckpt_dir, "%s", d); }
    int ng = sfx_graph_count();
    if(argc < 2 || !strcmp(argv[1], "list")){
        printf("%d graphs (one per mirror pair; w = mirror weight), precision fallback double-double on the GPU\n", ng);
        for(int i = 0; i < ng; i++){ sfx_graph_info_t g; sfx_graph_info(i, &g);
            printf("  %-8s L=%d  %-22s D=%d w=%d\n", g.name, g.L, g.volkov, g.D, g.weight); }
        printf("usage: %s run h N [eps] [top] GRAPH... | scale GRAPH [kappa] [seed] | compare R1 R2\n", argv[0]);
        return 0;
    }
    if(!strcmp(argv[1], "scale")){
        if(argc < 3) return 1;
        return scale(argv[2], argc > 3 ? atof(argv[3]) : 24, argc > 4 ? (unsigned)atoi(argv[4]) : 1, envi("SCALE_E", 1));
    }
    if(!strcmp(argv[1], "compare")){
        if(argc < 4){ fprintf(stderr, "compare RESULTS_h RESULTS_h2   (two results files, e.g. h = 1.0 and 1.2)\n"); return 1; }
        return compare(argv[2], argv[3]);
    }
    if(strcmp(argv[1], "run")){ fprintf(stderr, "unknown command %s\n", argv[1]); return 1; }
    if(argc < 5){ fprintf(stderr, "run h N [eps] [top] GRAPH...\n"); return 1; }
    double h = atof(argv[2]); int N = atoi(argv[3]), a = 4;
    double eps = 1e-11; int top = -1;
    if(a < argc && isnum(argv[a])) eps = atof(argv[a++]);
    if(a < argc && isnum(argv[a])) top = atoi(argv[a++]);
    if(top < 0){                                           /* main.c:211-215 */
        double ts = envd("SINCQED_TOPSTEP", 3.0);
        top = 1; while(top < 2 && (1 << (top + 1))*h <= ts*(1 + 1e-9)) top++;
    }
    const char *lf = getenv("SINCQED_LOG") ? getenv("SINCQED_LOG") : "sincqed.log";
    const char *rf = getenv("SINCQED_RESULTS") ? getenv("SINCQED_RESULTS") : "sincqed_results.txt";
    if(*lf && strcmp(lf, "-")) LOGF = fopen(lf, "a");
    if(*rf && strcmp(rf, "-")){
        FILE *t = fopen(rf, "r"); int fresh = !t; if(t) fclose(t);
        RESF = fopen(rf, "a");
        if(RESF && fresh) fprintf(RESF, "# name     volkov                       L D w      h    N      eps top  value                  err       disc     prune    trunc    round    S_2h            evals        time_s  ref                    diff\n");
    }
    int *sel = (int*)malloc((ng + 1)*sizeof(int)), ns = 0;
    for(int b = a; b < argc; b++){                         /* main.c:156-168 */
        int hit = 0;
        for(int i = 0; i < ng; i++){
            if(!sfx_graph_selected(i, argv[b])) continue;
            hit = 1;
            int dup = 0; for(int j = 0; j < ns; j++) dup |= sel[j] == i;
            if(!dup) sel[ns++] = i;
        }
        if(!hit) fprintf(stderr, "# no graph matches %s in this build\n", argv[b]);
    }
    if(ns > 1 && envi("SINCQED_BATCH", 1)){
        /* one hardware queue per engine worker's stream, at most 16 (HIP's default of four makes the workers'
           streams wait for one another in queue order); read by the runtime at its first call, so set before
           sfx_init, and a value already in the environment is kept */
        int w = envi("SINCQED_WORKERS", 16), q = w < 4 ? 4 : w > 16 ? 16 : w;
        char qs[8]; snprintf(qs, sizeof qs, "%d", q);
        setenv("GPU_MAX_HW_QUEUES", qs, 0);
    }
    int rc = sfx_init(envi("SINCQED_GPU", 0));
    if(rc){ fprintf(stderr, "sincqed: %s\n", sfx_last_error()); return 3; }
    time_t now = time(NULL); char tb[32]; strftime(tb, sizeof tb, "%Y-%m-%d %H:%M:%S", localtime(&now));
    char cmd[512] = "";
    for(int i = 0; i < argc; i++){
        size_t l = strlen(cmd), w = strlen(argv[i]);
        if(l + w + 5 >= sizeof cmd){ strcat(cmd, "..."); break; }
        memcpy(cmd + l, argv[i], w); cmd[l + w] = ' '; cmd[l + w + 1] = 0;
    }
    say("# sincqed %s: %s\n", tb, cmd);
    say("# h=%g N=%d eps=%g top=%d device=%s quad=double-double graphs=%d  log=%s  results=%s\n",
        h, N, eps, top, sfx_info(), ns, LOGF ? lf : "(none)", RESF ? rf : "(none)");
    { const char *e = getenv("SINCQED_SYM");
      if(e && strcmp(e, "0") && strcmp(e, "off")) say("# SINCQED_SYM ignored: the symmetry fold is not applied on the GPU\n"); }
    { int m = 1 << (top + 1), K0 = (N - 1)/2, K1 = ((K0 + m - 1)/m)*m;
      if(K1 != K0) say("# note: N=%d adjusted to N=%d (half-width must be a multiple of 2^(top+1)=%d); range |z| <= %g\n", N, 2*K1 + 1, m, K1*h);
      else say("# range |z| <= %g\n", K1*h);
      if(RESF) fprintf(RESF, "# %s  %s\n# window |z| <= %g\n", tb, cmd, K1*h); }
    double tot = 0, toterr = 0, toterr2 = 0;
    if(ns > 1 && envi("SINCQED_BATCH", 1))
        rc = run_batched(sel, ns, h, N, eps, top, &o, envi("SINCQED_WORKERS", 16), &tot, &toterr, &toterr2);
    else
        for(int j = 0; j < ns; j++) if((rc = run_graph(sel[j], h, N, eps, top, &o, j + 1, ns, &tot, &toterr, &toterr2))) break;
    if(ns > 1 && !rc){
        say("# sum of %d graphs: %+.12f  err %.2e (linear sum)  %.2e (in quadrature)\n", ns, tot, toterr, sqrt(toterr2));
        if(RESF) fprintf(RESF, "# SUM %d graphs  %+.15e  err %.3e (linear)  %.3e (quadrature)\n", ns, tot, toterr, sqrt(toterr2));
    }
    if(LOGF) fclose(LOGF);
    if(RESF) fclose(RESF);
    sfx_shutdown();
    free(sel);
    return rc ? 4 : 0;
}
