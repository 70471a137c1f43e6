// sfx_engine.hpp -- the job engine: continuous batching of hierarchical sums on one device.
//
// A pool of workers, each with its own context (a HIP stream and device buffers), takes queued (graph,
// h, N, eps, top) jobs in arrival order, so jobs that arrive while others run are taken as workers free
// up and their device work overlaps.  Admission is by HBM footprint (Budget below): before each
// refinement level that keeps cells, a job reserves the bound of its next cell list and waits while the
// budget is held by others.  No counterpart in the reference, which runs the graphs of one process one
// after another (driver/main.c:243).
//
// This header holds only host logic (threads, queue, budget, job records): the device work is reached
// through EngineBackend, so the same code is built into the library (sfx_hsum.hip, the backend runs
// hsum_core on the worker's stream) and into a CPU test under ThreadSanitizer (a stub backend).
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdio.h>
#include <string.h>
#include <thread>
#include <vector>
#include "../../include/sfx.h"

namespace sfx {

void set_error(const char *fmt, ...);

// where a running job is, for the messages of waits that pass their deadline: written by the worker
// that runs the job, read by any thread
enum { STAGE_QUEUED = 0, STAGE_COARSE, STAGE_DILATE, STAGE_LEVEL, STAGE_BUDGET, STAGE_DONE };
inline const char *stage_name(int s){
    static const char *n[] = {"queued", "coarsest level", "dilation", "refinement", "waiting for HBM budget", "done"};
    return s >= 0 && s <= STAGE_DONE ? n[s] : "?";
}
struct JobProgress {
    std::atomic<int> stage{STAGE_QUEUED}, level{-1};
    std::atomic<long> block{0};
    JobProgress() = default;
    JobProgress(const JobProgress &o){ *this = o; }
    JobProgress &operator=(const JobProgress &o){ set(o.stage.load(), o.level.load(), o.block.load()); return *this; }
    void set(int s, int lev = -1, long b = 0){ level.store(lev); block.store(b); stage.store(s); }
};

inline double engine_now(){ return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// a condition wait with a deadline in seconds.  On the system clock (pthread_cond_timedwait): libstdc++'s
// steady-clock waits go through pthread_cond_clockwait, which the ThreadSanitizer of GCC 11 does not
// intercept, so the engine's CPU test under it would see the mutex as never released
template<class P> inline bool cv_wait_for(std::condition_variable &cv, std::unique_lock<std::mutex> &lk, double s, P pred){
    auto until = std::chrono::system_clock::now() + std::chrono::duration_cast<std::chrono::system_clock::duration>(std::chrono::duration<double>(s));
    return cv.wait_until(lk, until, pred);
}

// ---- HBM budget ---------------------------------------------------------------------------------
// A job's reservation is a BudgetHold.  A waiting job is admitted when (a) its request fits beside what
// is held, or (b) nothing else is held, or (c) every job that holds memory is itself waiting -- then the
// oldest waiter goes first, over the cap: without (c) two jobs that each hold their current cell list and
// each need more than the headroom would wait for each other forever.  A wait past the budget's deadline
// fails with SFX_ETIMEOUT instead of blocking.
struct BudgetHold { double held = 0; };
struct Budget {
    std::mutex mu; std::condition_variable cv;
    double cap = 0, used = 0;          // cap <= 0: no budget (outside the engine)
    double deadline_s = 600;           // <= 0: wait without a deadline
    int holders = 0, waiting_holders = 0;
    unsigned long long next_ticket = 0;
    std::set<unsigned long long> waiters;
    long admitted_over_cap = 0;        // admissions by rule (c)
};
inline Budget &budget(){ static Budget B; return B; }

// reserve `bytes` more for the job of `h`; `who` names the job in an error message
inline int budget_acquire(BudgetHold &h, double bytes, const char *who, JobProgress *pg){
    Budget &B = budget();
    if(bytes <= 0) return SFX_OK;
    std::unique_lock<std::mutex> lk(B.mu);
    // (b) by count, not by bytes: the sums of reservations carry rounding
    auto alone = [&]{ return B.holders - (h.held > 0 ? 1 : 0) == 0; };
    if(B.cap > 0 && !(B.used + bytes <= B.cap || alone())){
        const unsigned long long t = B.next_ticket++;
        B.waiters.insert(t);
        const bool holder = h.held > 0;
        if(holder) B.waiting_holders++;
        int st = pg ? pg->stage.load() : 0;
        if(pg) pg->stage.store(STAGE_BUDGET);
        B.cv.notify_all();                 // this wait may leave every holder waiting: rule (c) for the oldest
        bool over = false;
        auto ok = [&]{
            if(B.cap <= 0 || B.used + bytes <= B.cap || alone()) return true;
            if(B.waiting_holders == B.holders && *B.waiters.begin() == t){ over = true; return true; }
            return false;
        };
        bool got;
        if(B.deadline_s > 0) got = cv_wait_for(B.cv, lk, B.deadline_s, ok);
        else { B.cv.wait(lk, ok); got = true; }
        B.waiters.erase(t);
        if(holder) B.waiting_holders--;
        if(pg) pg->stage.store(st);
        if(!got){
            set_error("%s: no HBM budget for %.3g GB after %.0f s (held by this job %.3g GB, by all jobs %.3g GB of %.3g GB; "
                      "level %d)", who, bytes/1e9, B.deadline_s, h.held/1e9, B.used/1e9, B.cap/1e9, pg ? pg->level.load() : -1);
            B.cv.notify_all();
            return SFX_ETIMEOUT;
        }
        if(over) B.admitted_over_cap++;
    }
    if(h.held <= 0) B.holders++;
    h.held += bytes; B.used += bytes;
    return SFX_OK;
}
// give back `bytes` of the job's reservation (all of it with bytes >= held)
inline void budget_release(BudgetHold &h, double bytes){
    Budget &B = budget();
    if(bytes <= 0 || h.held <= 0) return;
    {
        std::lock_guard<std::mutex> lk(B.mu);
        if(bytes > h.held) bytes = h.held;
        h.held -= bytes; B.used -= bytes;
        if(h.held <= 0){ h.held = 0; B.holders--; }
        if(B.holders == 0 || B.used < 0) B.used = 0;          // no rounding residue once nothing is held
    }
    B.cv.notify_all();
}

// ---- job records and the engine -----------------------------------------------------------------
struct EJob {
    long id = 0;
    int graph = 0, N = 0, top = 0; double h = 0, eps = 0; sfx_opts_t o;
    int state = SFX_JOB_QUEUED, rc = 0; char err[256] = "";      // guarded by Engine::mu
    sfx_result_t res;
    double t_submit = 0, t_start = 0, t_end = 0;
    char name[32] = "";
    JobProgress prog;                                             // atomics: read without the lock
};

// the device side of the engine
struct EngineBackend {
    void *(*ctx_create)();                    // a worker context (stream + buffers), nullptr on failure
    void (*ctx_destroy)(void *ctx);
    int (*run)(void *ctx, EJob &j);           // the job's sum on the context: fills j.res, on failure j.err,
                                              // returns an SFX code
    void (*after_job)(void *ctx);             // waits for the context's stream, trims its buffers
};

struct Engine {
    std::mutex mu; std::condition_variable cv, done;
    std::deque<long> queue;
    std::map<long, std::shared_ptr<EJob>> jobs;
    std::vector<std::thread> workers;
    std::vector<void*> ctx;
    EngineBackend be{};
    bool stop = false; long next = 1; int running = 0;
    // running + queued jobs, kept under mu and read without it (a job's grow takes the pipelined frontier loop
    // only while it has the device to itself)
    std::atomic<int> load{0};
    void note_load(){ load.store(running + (int)queue.size(), std::memory_order_relaxed); }
    double wait_deadline_s = 3600;            // sfx_job_wait; <= 0: no deadline
    ~Engine(){
        // at process exit: wake and join idle workers (a joinable std::thread must not be destroyed); the
        // HIP objects are left to the runtime's own teardown
        { std::lock_guard<std::mutex> lk(mu); stop = true; }
        cv.notify_all();
        for(std::thread &t : workers) if(t.joinable()) t.join();
    }
};
inline Engine &eng(){ static Engine E; return E; }

inline void engine_worker(int w){
    Engine &E = eng();
    void *X;
    { std::lock_guard<std::mutex> lk(E.mu); X = E.ctx[w]; }
    for(;;){
        std::shared_ptr<EJob> j;
        {
            std::unique_lock<std::mutex> lk(E.mu);
            E.cv.wait(lk, [&]{ return E.stop || !E.queue.empty(); });
            if(E.stop) return;
            long id = E.queue.front(); E.queue.pop_front();
            auto it = E.jobs.find(id);
            if(it == E.jobs.end()) continue;
            j = it->second; j->state = SFX_JOB_RUNNING; j->t_start = engine_now(); E.running++; E.note_load();
        }
        // the result and the message are written here, outside the lock, and read by waiters only after
        // the state below says the job has finished
        j->prog.set(STAGE_COARSE);
        memset(&j->res, 0, sizeof j->res);
        int rc = E.be.run(X, *j);
        E.be.after_job(X);
        j->prog.set(STAGE_DONE);
        if(rc && !j->err[0]) snprintf(j->err, sizeof j->err, "job %ld (%s): failed with code %d", j->id, j->name, rc);
        {
            std::lock_guard<std::mutex> lk(E.mu);
            j->rc = rc; j->t_end = engine_now();
            j->state = rc ? SFX_JOB_FAILED : SFX_JOB_DONE; E.running--; E.note_load();
        }
        E.done.notify_all();
    }
}

// under E.mu
inline int engine_start_locked(Engine &E, int workers, double budget_bytes, double hbm_bytes){
    if(!E.workers.empty()) return SFX_OK;
    if(workers <= 0) workers = 8;
    if(workers > 64) workers = 64;
    { std::lock_guard<std::mutex> lb(budget().mu); budget().cap = budget_bytes > 0 ? budget_bytes : 0.8*hbm_bytes; }
    E.stop = false;
    for(int w = 0; w < workers; w++){
        void *c = E.be.ctx_create();
        if(!c){
            for(void *x : E.ctx) E.be.ctx_destroy(x);
            E.ctx.clear();
            return SFX_EDEVICE;
        }
        E.ctx.push_back(c);
    }
    for(int w = 0; w < workers; w++) E.workers.emplace_back(engine_worker, w);
    return SFX_OK;
}
inline int engine_start(int workers, double budget_bytes, double hbm_bytes){
    Engine &E = eng();
    std::lock_guard<std::mutex> lk(E.mu);
    return engine_start_locked(E, workers, budget_bytes, hbm_bytes);
}

// stop: the queued jobs fail at once (SFX_ECANCELED; their waiters wake), the running ones finish, the
// workers are joined and their contexts destroyed.  The job records stay until released.
inline void engine_shutdown(){
    Engine &E = eng();
    std::vector<std::thread> ws;
    {
        std::lock_guard<std::mutex> lk(E.mu);
        E.stop = true;
        for(long id : E.queue){
            auto it = E.jobs.find(id);
            if(it == E.jobs.end()) continue;
            EJob &j = *it->second;
            j.state = SFX_JOB_FAILED; j.rc = SFX_ECANCELED; j.t_start = j.t_end = engine_now();
            snprintf(j.err, sizeof j.err, "job %ld (%s h=%g eps=%g): the engine stopped before it started", j.id, j.name, j.h, j.eps);
            j.prog.set(STAGE_DONE);
        }
        E.queue.clear(); E.note_load();
        ws.swap(E.workers);
    }
    E.cv.notify_all(); E.done.notify_all();
    for(std::thread &t : ws) t.join();
    std::vector<void*> cs;
    { std::lock_guard<std::mutex> lk(E.mu); cs.swap(E.ctx); E.running = 0; E.note_load(); }
    for(void *c : cs) E.be.ctx_destroy(c);
    { std::lock_guard<std::mutex> lb(budget().mu); budget().cap = 0; }   // no budget outside the engine
    budget().cv.notify_all();
}

// queue a copy of `proto` (graph, h, N, eps, top, o, name); the engine starts with its defaults if it is
// not running (hbm_bytes: the device memory the default budget is a share of)
inline int engine_submit(const EJob &proto, long *job, double hbm_bytes){
    Engine &E = eng();
    {
        std::lock_guard<std::mutex> lk(E.mu);
        int rc;
        if(E.workers.empty() && (rc = engine_start_locked(E, 0, 0, hbm_bytes))) return rc;
        long id = E.next++;
        auto j = std::make_shared<EJob>();
        j->id = id; j->graph = proto.graph; j->h = proto.h; j->N = proto.N; j->eps = proto.eps; j->top = proto.top;
        j->o = proto.o; memcpy(j->name, proto.name, sizeof j->name); j->t_submit = engine_now();
        E.jobs[id] = j;
        E.queue.push_back(id); E.note_load();
        *job = id;
    }
    E.cv.notify_one();
    return SFX_OK;
}

inline int engine_poll(long job, int *state){
    Engine &E = eng(); std::lock_guard<std::mutex> lk(E.mu);
    auto it = E.jobs.find(job);
    if(it == E.jobs.end()){ set_error("no job %ld", job); return SFX_EINVAL; }
    if(state) *state = it->second->state;
    return SFX_OK;
}

// waits until the job has finished, at most `deadline_s` seconds (<= 0: no limit)
inline int engine_wait(long job, double deadline_s, sfx_result_t *res, double *t_queue, double *t_run){
    Engine &E = eng(); std::unique_lock<std::mutex> lk(E.mu);
    auto it = E.jobs.find(job);
    if(it == E.jobs.end()){ set_error("no job %ld", job); return SFX_EINVAL; }
    std::shared_ptr<EJob> jp = it->second;          // the record outlives a concurrent release or stop
    EJob &j = *jp;
    auto fin = [&]{ return j.state == SFX_JOB_DONE || j.state == SFX_JOB_FAILED; };
    if(deadline_s > 0){
        if(!cv_wait_for(E.done, lk, deadline_s, fin)){
            int st = j.prog.stage.load();
            set_error("job %ld (%s h=%g eps=%g): not finished after %.0f s: %s, stage %s, level %d, block %ld",
                      j.id, j.name, j.h, j.eps, deadline_s, j.state == SFX_JOB_QUEUED ? "queued" : "running",
                      stage_name(st), j.prog.level.load(), j.prog.block.load());
            return SFX_ETIMEOUT;
        }
    } else E.done.wait(lk, fin);
    if(res) *res = j.res;
    if(t_queue) *t_queue = j.t_start - j.t_submit;
    if(t_run) *t_run = j.t_end - j.t_start;
    if(j.rc){ set_error("%s", j.err); return j.rc; }
    return SFX_OK;
}

inline int engine_release(long job){
    Engine &E = eng(); std::lock_guard<std::mutex> lk(E.mu);
    auto it = E.jobs.find(job);
    if(it == E.jobs.end()){ set_error("no job %ld", job); return SFX_EINVAL; }
    if(it->second->state == SFX_JOB_QUEUED || it->second->state == SFX_JOB_RUNNING){ set_error("job %ld not finished", job); return SFX_EINVAL; }
    E.jobs.erase(it);
    return SFX_OK;
}

}  // namespace sfx
