// engine_stub.cpp -- the job engine's host logic (sincfeyn_amd/csrc/sfx_engine.hpp: queue, workers, HBM
// budget, deadlines, stop) driven on the CPU with a stub backend in place of the device, built with
// ThreadSanitizer by tests/test_engine_cpu.py.  The stub's "sum" of a job sleeps for a time set by its
// size and reserves HBM budget level by level in the pattern of hsum_core (the top cells, then each
// level's bound of its next cell list, kept down to the list the next level reads), and returns a
// result that is a function of the job alone -- so a job's result alone and in any mix must be equal.
//
//   engine_stub stream <jobs file> <workers> <budget_bytes> <time_scale>
//       jobs file: one line per job "arrival_s D h N eps top"; submits each at arrival*time_scale (the
//       stub's work is scaled by time_scale/20), waits for every job in submission order with a
//       deadline, checks each result against the job's own
//   engine_stub holdwait        two jobs that each hold their list and need more than the headroom
//   engine_stub stop            stop with queued jobs: they fail with SFX_ECANCELED, waiters wake
//   engine_stub deadlines       job-wait and budget deadlines fail with SFX_ETIMEOUT naming the job
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include "sfx_engine.hpp"

namespace sfx {
static thread_local char g_err[256];
void set_error(const char *fmt, ...){ va_list ap; va_start(ap, fmt); vsnprintf(g_err, sizeof g_err, fmt, ap); va_end(ap); }
}
using namespace sfx;

static std::atomic<long> n_ctx{0};
static double g_scale = 1e-3;          // seconds of stub work per unit of job size
static int g_barrier = 0;              // > 0: jobs meet after reserving their top cells (hold-and-wait case)
static std::atomic<int> g_held{0};

struct StubCtx { long id; double scratch[64]; };
static void *stub_create(){ StubCtx *c = new StubCtx(); c->id = n_ctx++; return c; }
static void stub_destroy(void *p){ delete (StubCtx*)p; n_ctx--; }
static void stub_after(void *p){ StubCtx *c = (StubCtx*)p; for(double &x : c->scratch) x = 0; }

static void sleep_s(double s){ if(s > 0) std::this_thread::sleep_for(std::chrono::duration<double>(s)); }

// the job's expected result: a function of (D, h, N, eps, top) only
static void expected(const EJob &j, sfx_result_t *r){
    memset(r, 0, sizeof *r);
    r->S[0] = sin(j.h*j.graph + 1.0/j.N) + log10(j.eps); r->S[1] = r->S[0]/3; r->S[2] = NAN;
    r->nev = (long)(j.graph*1000003L + j.N*17 + j.top); r->kept_top = 1000*j.graph + j.top;
    r->K = (j.N - 1)/2; r->top = j.top;
}
// graph = D here; "size" of a job grows with D and 1/h like its lattice
static int stub_run(void *p, EJob &j){
    StubCtx *c = (StubCtx*)p; (void)c;
    const int D = j.graph;
    const double size = pow(1.0/j.h, D)*(j.eps < 1e-11 ? 2.0 : 1.0);
    sleep_s(g_scale*0.2*size);                                          // coarsest level
    j.prog.set(STAGE_DILATE);
    char who[64]; snprintf(who, sizeof who, "stub D=%d h=%g", D, j.h);
    BudgetHold bh;
    double cells = 1e6*size;
    int rc = budget_acquire(bh, cells*16, who, &j.prog);
    if(rc){ snprintf(j.err, sizeof j.err, "%s", g_err); return rc; }
    if(g_barrier > 0){
        // every job of the scenario holds its top cells before any asks for its next list
        g_held++;
        while(g_held.load() < g_barrier) sleep_s(1e-4);
    }
    for(int lev = j.top - 1; lev >= 0; lev--){
        j.prog.set(STAGE_LEVEL, lev);
        double want = lev > 0 ? cells*(1 << D)*16 : 0;
        if((rc = budget_acquire(bh, want, who, &j.prog))){ budget_release(bh, bh.held); snprintf(j.err, sizeof j.err, "%s", g_err); return rc; }
        for(int b = 0; b < 4; b++){ j.prog.block.store(b); sleep_s(g_scale*0.2*size); }
        cells *= 3;
        if(lev > 0) budget_release(bh, bh.held - cells*16);
    }
    budget_release(bh, bh.held);
    expected(j, &j.res);
    return SFX_OK;
}

static void bind(){ EngineBackend &b = eng().be; b.ctx_create = stub_create; b.ctx_destroy = stub_destroy; b.run = stub_run; b.after_job = stub_after; }
static EJob proto(int D, double h, int N, double eps, int top){
    EJob p; p.graph = D; p.h = h; p.N = N; p.eps = eps; p.top = top; memset(&p.o, 0, sizeof p.o); snprintf(p.name, sizeof p.name, "D%d", D);
    return p;
}
static int fail(const char *what){ fprintf(stderr, "FAIL: %s (%s)\n", what, g_err); return 1; }
static bool same(const sfx_result_t &a, const sfx_result_t &b){
    return a.S[0] == b.S[0] && a.S[1] == b.S[1] && a.nev == b.nev && a.kept_top == b.kept_top && a.K == b.K && a.top == b.top;
}

static int scenario_stream(const char *file, int workers, double budget_bytes, double tscale){
    FILE *f = fopen(file, "r"); if(!f) return fail("jobs file");
    struct J { double at, h, eps; int D, N, top; };
    std::vector<J> js; J x;
    while(fscanf(f, "%lf %d %lf %d %lf %d", &x.at, &x.D, &x.h, &x.N, &x.eps, &x.top) == 6) js.push_back(x);
    fclose(f);
    if(engine_start(workers, budget_bytes, 1e12)) return fail("start");
    std::vector<long> ids(js.size(), 0);
    std::atomic<long> submitted{0};
    std::thread client([&]{
        const double t0 = engine_now();
        for(size_t i = 0; i < js.size(); i++){
            sleep_s(t0 + js[i].at*tscale - engine_now());
            long id; EJob p = proto(js[i].D, js[i].h, js[i].N, js[i].eps, js[i].top);
            if(engine_submit(p, &id, 1e12)){ fprintf(stderr, "submit failed: %s\n", g_err); abort(); }
            ids[i] = id; submitted.store((long)i + 1);
        }
    });
    int bad = 0;
    for(size_t i = 0; i < js.size(); i++){
        while(submitted.load() <= (long)i) sleep_s(1e-4);
        sfx_result_t r, e; double tq, tr;
        int rc = engine_wait(ids[i], 60, &r, &tq, &tr);
        if(rc){ fprintf(stderr, "job %zu: rc %d %s\n", i, rc, g_err); bad++; continue; }
        EJob p = proto(js[i].D, js[i].h, js[i].N, js[i].eps, js[i].top); expected(p, &e);
        if(!same(r, e)){ fprintf(stderr, "job %zu: result differs from the job alone\n", i); bad++; }
        if(engine_release(ids[i])) bad++;
    }
    client.join();
    engine_shutdown();
    if(n_ctx.load() != 0) return fail("contexts left after stop");
    if(budget().used != 0 || budget().holders != 0) return fail("budget not returned");
    printf("OK stream: %zu jobs, %d workers, admitted over cap %ld\n", js.size(), workers, budget().admitted_over_cap);
    return bad ? fail("stream results") : 0;
}

// two D=7 jobs at top 2: each holds its top cells, then wants 2^7 x that for its next list; the cap holds
// one job's list and one want, not two -- before rule (c) they waited for each other forever
static int scenario_holdwait(){
    g_scale = 0.02; g_barrier = 2;
    const double size = pow(1/0.7, 7)*2.0, list = 1e6*size*16, want = list*128;
    if(engine_start(2, 2*list + want*0.5, 1e12)) return fail("start");
    long a, b; EJob p = proto(7, 0.7, 139, 1e-12, 2);
    if(engine_submit(p, &a, 1e12) || engine_submit(p, &b, 1e12)) return fail("submit");
    sfx_result_t r; double tq, tr;
    if(engine_wait(a, 30, &r, &tq, &tr) || engine_wait(b, 30, &r, &tq, &tr)) return fail("hold-and-wait jobs did not finish");
    long over = budget().admitted_over_cap;
    engine_shutdown();
    g_barrier = 0;
    printf("OK holdwait: both finished, admitted over cap %ld\n", over);
    return 0;
}

static int scenario_stop(){
    g_scale = 0.01;
    if(engine_start(2, 0, 1e12)) return fail("start");
    std::vector<long> ids;
    for(int i = 0; i < 40; i++){ long id; EJob p = proto(4, 1.0, 97, 1e-12, 1); if(engine_submit(p, &id, 1e12)) return fail("submit"); ids.push_back(id); }
    // waiters on queued jobs, blocked when the engine stops
    std::atomic<int> canceled{0}, finished{0};
    std::vector<std::thread> ws;
    for(int i = 30; i < 40; i++) ws.emplace_back([&, i]{
        sfx_result_t r; double tq, tr; int rc = engine_wait(ids[i], 30, &r, &tq, &tr);
        if(rc == SFX_ECANCELED) canceled++; else if(rc == SFX_OK) finished++;
    });
    sleep_s(0.02);
    engine_shutdown();
    for(std::thread &t : ws) t.join();
    int st, nc = 0, nd = 0;
    for(long id : ids){ if(engine_poll(id, &st)) return fail("poll after stop"); nc += st == SFX_JOB_FAILED; nd += st == SFX_JOB_DONE; }
    for(long id : ids) if(engine_release(id)) return fail("release after stop");
    if(canceled.load() + finished.load() != 10) return fail("a waiter did not wake with the stop");
    if(nd + nc != 40 || nc == 0) return fail("queued jobs not canceled");
    // the engine starts again after a stop
    if(engine_start(2, 0, 1e12)) return fail("restart");
    long id; EJob p = proto(3, 1.0, 97, 1e-12, 1);
    sfx_result_t r, e; double tq, tr;
    if(engine_submit(p, &id, 1e12) || engine_wait(id, 30, &r, &tq, &tr)) return fail("job after restart");
    expected(p, &e); if(!same(r, e)) return fail("result after restart");
    engine_release(id);
    engine_shutdown();
    printf("OK stop: %d canceled, %d finished, %d waiters canceled\n", nc, nd, canceled.load());
    return 0;
}

static int scenario_deadlines(){
    g_scale = 0.05;
    if(engine_start(2, 0, 1e12)) return fail("start");
    long id; EJob p = proto(5, 0.7, 139, 1e-12, 2);
    if(engine_submit(p, &id, 1e12)) return fail("submit");
    sfx_result_t r; double tq, tr;
    int rc = engine_wait(id, 0.05, &r, &tq, &tr);
    if(rc != SFX_ETIMEOUT || !strstr(g_err, "not finished after") || !strstr(g_err, "D5")) return fail("job-wait deadline");
    printf("deadline message: %s\n", g_err);
    if(engine_wait(id, 60, &r, &tq, &tr)) return fail("job after its wait deadline");
    engine_release(id);
    engine_shutdown();
    // budget deadline: one job holds the whole budget while running; another cannot be admitted by any rule
    g_scale = 0.2;
    { std::lock_guard<std::mutex> lk(budget().mu); budget().deadline_s = 0.05; }
    const double list = 1e6*pow(1/0.7, 3)*2.0*16;
    if(engine_start(2, list*1.01, 1e12)) return fail("start 2");
    long a, b; EJob pa = proto(3, 0.7, 139, 1e-12, 2), pb = proto(3, 0.7, 139, 1e-12, 2);
    if(engine_submit(pa, &a, 1e12)) return fail("submit a");
    sleep_s(0.01);
    if(engine_submit(pb, &b, 1e12)) return fail("submit b");
    int ra = engine_wait(a, 60, &r, &tq, &tr), rb = engine_wait(b, 60, &r, &tq, &tr);
    { std::lock_guard<std::mutex> lk(budget().mu); budget().deadline_s = 600; }
    engine_shutdown();
    if(!((ra == SFX_ETIMEOUT) ^ (rb == SFX_ETIMEOUT)) || !strstr(g_err, "no HBM budget")){
        fprintf(stderr, "ra %d rb %d\n", ra, rb); return fail("budget deadline");
    }
    printf("budget message: %s\nOK deadlines\n", g_err);
    return 0;
}

int main(int argc, char **argv){
    bind();
    if(argc < 2) return fail("usage");
    std::string s = argv[1];
    if(s == "stream" && argc == 6){ g_scale = atof(argv[5])/20; return scenario_stream(argv[2], atoi(argv[3]), atof(argv[4]), atof(argv[5])); }
    if(s == "holdwait") return scenario_holdwait();
    if(s == "stop") return scenario_stop();
    if(s == "deadlines") return scenario_deadlines();
    return fail("unknown scenario");
}
