/* ref_main.c -- harness around the REFERENCE implementation (test infrastructure only).
 *
 * Linked with sincFeyn's own driver/{sum,eval,lattice,checkpoint,io,graphs}.c and kernels, compiled
 * unmodified from their place under the reference tree by oracle/Makefile.  It stands where the
 * reference's main.c stands for the `run` and `list` commands: the graph descriptors are found by name
 * with dlsym (the executable is linked -rdynamic) instead of through the make-generated registry, and
 * run_graph below restates main.c:40-75 (the per-graph results line, main.c:66-73) so that its output
 * has the reference's exact format.  One extra machine-readable line per graph, starting "#R", carries
 * every field of res_t (sq.h:73) at full precision for the parity tests.
 *
 *   sincqed_ref list
 *   sincqed_ref run h N [eps] [top] GRAPH...      (environment as main.c:176-185)
 */
#include "sq.h"
#include <dlfcn.h>
#ifdef SQ_GPU
#include "gpu.h"
#endif

#ifndef SQ_PREFIX
#define SQ_PREFIX "G_"          /* descriptor symbol prefix of the linked kernels */
#endif
#ifndef SQ_NAMES
#error "SQ_NAMES (the kernel names, space separated) must be defined by the build"
#endif

static void load_graphs(void){
    char names[] = SQ_NAMES, *save = NULL;
    for(char *t = strtok_r(names, " ", &save); t; t = strtok_r(NULL, " ", &save)){
        char sym[96]; snprintf(sym, sizeof sym, "%s%s", SQ_PREFIX, t);
        const graph_t *g = (const graph_t*)dlsym(RTLD_DEFAULT, sym);
        if(!g){ fprintf(stderr, "sincqed_ref: no symbol %s\n", sym); exit(2); }
        add_graph(g);
    }
}

/* main.c:40-75 restated: value = w S_h, err = w (disc + prune + trunc + round) */
static void run_graph(const graph_t *g, double h, int N, double eps, int top, double *tot, double *toterr, double *toterr2){
    CURNAME = g->name;
    lattice_setup(g, h, N, top);
    ABSKIPPED = ABMISS = 0; UNRESDROP = 0;
    res_t R = hsum(g, h, N, eps, top);
    double w = g->weight, v = w*R.S[0], d2 = fabs(R.S[1]-R.S[0]);
    double disc = R.Ah > 0 ? 2*d2*d2/R.Ah : 0;
    double err = w*(disc + R.prune + R.trunc + R.unres);
    const ref_t *r = findref(g->volkov);
    say("%-7s %-22s D=%d w=%d  %+.12f  err %.2e  evals %ld quad %ld kept %ld  %.1fs\n",
        g->name, g->volkov, g->D, g->weight, v, err, R.nev, R.nq, R.kept_top, R.t);
    if(RESF){
        fprintf(RESF, "%-8s %-28s %d %d %d %5.3f %4d %8.1e %d %+.15e %.3e %.2e %.2e %.2e %.2e ",
                g->name, g->volkov, g->L, g->D, g->weight, h, 2*K+1, eps, top, v, err, w*disc, w*R.prune, w*R.trunc, w*R.unres);
        if(isfinite(R.S[1])) fprintf(RESF, "%+.8e", w*R.S[1]); else fprintf(RESF, "%15s", "-");
        fprintf(RESF, " %12ld %9.1f", R.nev, R.t);
        if(r) fprintf(RESF, " %+.15e %+.3e", r->val, v - r->val); else fprintf(RESF, " %22s %10s", "-", "-");
        fprintf(RESF, "\n"); fflush(RESF);
    }
    printf("#R %s S0 %.17g S1 %.17g S2 %.17g Ah %.17g ref %.17g prune %.17g trunc %.17g unres %.17g nev %ld nq %ld kept %ld skipped %ld unresdrop %ld abmiss %ld t %.3f\n",
           g->name, R.S[0], R.S[1], R.S[2], R.Ah, R.ref, R.prune, R.trunc, R.unres, R.nev, R.nq, R.kept_top, ABSKIPPED, UNRESDROP, ABMISS, R.t);
    fflush(stdout);
    *tot += v; *toterr += err; *toterr2 += err*err;
}

static int isnum(const char *s){ char *e; strtod(s, &e); return e != s && *e == 0; }

int main(int argc, char **argv){
    load_graphs();
    if(getenv("G2_ABSERR")) ABSERR = ABSERR0 = atof(getenv("G2_ABSERR"));
    if(getenv("SINCQED_ABREL")) ABREL = atof(getenv("SINCQED_ABREL"));
    if(getenv("SINCQED_ABSKIP")) ABSKIP = atoi(getenv("SINCQED_ABSKIP"));
    if(getenv("SINCQED_ABCHECK")) ABCHECK = atoi(getenv("SINCQED_ABCHECK"));
    if(getenv("SINCQED_NEARF")) NEARF = atof(getenv("SINCQED_NEARF"));
    if(getenv("SINCQED_DDSLACK")) DDSLACK = atof(getenv("SINCQED_DDSLACK"));
    if(getenv("SINCQED_DROPUNRES")) DROPUNRES = atoi(getenv("SINCQED_DROPUNRES"));
    if(getenv("SINCQED_GROW")) GROW = atoi(getenv("SINCQED_GROW"));
    if(getenv("SINCQED_CKPT")) CKPT_EVERY = atof(getenv("SINCQED_CKPT"));
    if(getenv("SINCQED_PROGRESS")) PROG = atof(getenv("SINCQED_PROGRESS"));
    if(argc < 2 || !strcmp(argv[1], "list")){
        printf("%d graphs\n", NALL);
        for(int i = 0; i < NALL; i++) printf("  %-8s L=%d  %-22s D=%d w=%d\n", ALL[i]->name, ALL[i]->L, ALL[i]->volkov, ALL[i]->D, ALL[i]->weight);
        return 0;
    }
    if(strcmp(argv[1], "run") || argc < 5){ fprintf(stderr, "usage: %s run h N [eps] [top] GRAPH...\n", argv[0]); return 1; }
    double h = atof(argv[2]); int N = atoi(argv[3]), a = 4;
    double eps = 1e-11; int top = -1;
    if(a < argc && isnum(argv[a])) eps = atof(argv[a++]);
    if(a < argc && isnum(argv[a])) top = atoi(argv[a++]);
    if(top < 0){                                   /* main.c:211-215 */
        double ts = getenv("SINCQED_TOPSTEP") ? atof(getenv("SINCQED_TOPSTEP")) : 3.0;
        top = 1; while(top < 2 && (1 << (top+1))*h <= ts*(1 + 1e-9)) top++;
    }
    const char *rf = getenv("SINCQED_RESULTS") ? getenv("SINCQED_RESULTS") : "-";
    if(*rf && strcmp(rf, "-")){
        FILE *t = fopen(rf, "r"); int fresh = !t; if(t) fclose(t);
        RESF = fopen(rf, "a");
        if(RESF && fresh) fprintf(RESF, "# name     volkov                       L D w      h    N      eps top  value                  err       disc     prune    trunc    round    S_2h            evals        time_s  ref                    diff\n");
    }
    const graph_t **sel = (const graph_t**)malloc((NALL + 1)*sizeof *sel); int ns = 0;
    for(int b = a; b < argc; b++) for(int i = 0; i < NALL; i++){
        if(!selected(ALL[i], argv[b])) continue;
        int dup = 0; for(int j = 0; j < ns; j++) dup |= sel[j] == ALL[i];
        if(!dup) sel[ns++] = ALL[i];
    }
    T0 = TLAST = wtime(); CURN = ns;
    int nth = 1;
#ifdef _OPENMP
    nth = omp_get_max_threads();
#endif
    printf("# h=%g N=%d eps=%g top=%d threads=%d graphs=%d\n", h, N, eps, top, nth, ns);
    double tot = 0, toterr = 0, toterr2 = 0;
    for(int j = 0; j < ns; j++){ CURI = j + 1; run_graph(sel[j], h, N, eps, top, &tot, &toterr, &toterr2); }
    if(ns > 1 && RESF) fprintf(RESF, "# SUM %d graphs  %+.15e  err %.3e (linear)  %.3e (quadrature)\n", ns, tot, toterr, sqrt(toterr2));
    printf("# wall %.3f s\n", wtime() - T0);
#ifdef SQ_GPU
    printf("# gpu %s\n", gpu_info());          /* "off" when no graph ran on the device */
#endif
    if(RESF) fclose(RESF);
    return 0;
}
