// Serving loop: many channel cases of different grid sizes, Reynolds numbers
// and models share one GPU with continuous batching.
//
// - One HBM arena is reserved at creation (a single hipMalloc of the budget);
//   a case is admitted when a contiguous region of its footprint (the exact
//   byte count its context carves, computed by a dry allocation pass) is free
//   and a slot (a HIP stream) is free.  Admission is first-come-first-served
//   in submission order once a case's arrival time has passed.
// - Admission work (field upload, projection of the initial field, capture
//   of the two step graphs) runs on a worker thread, so the scheduler keeps
//   the running cases fed meanwhile.
// - Each running case advances by launching its captured step graph on its
//   own stream, with up to `depth` steps in flight; cases of every size run
//   concurrently, and a case that arrives while others run starts as soon as
//   memory and a slot free up.  Nothing is shared between cases but the
//   arena, so a case computes the same bits alone and in a mix.
// - Step latency of a case is the GPU time between its consecutive step
//   completion events; case latency is finish minus arrival.
// - With batching on (cf_server_set_batching), cases of one size class
//   (batch_class_key: the same options but nu and the forcing, the same
//   grid) share a group: one arena region of cap member slabs at one stride
//   and one stream, stepped by the batched step (each kernel launched once
//   for all members).  A group runs in stretches of at most `chunk` steps,
//   none longer than its members' remaining steps; between stretches (the
//   group drained) finished members are read back and retired, and cases
//   that arrived for the group are prepared into free member slots, so
//   requests join a running group as members leave.  A member's step
//   latency is the group step's GPU time.  Cases outside any batchable class
//   run one per stream as above.
//
// The reference has no serving layer; one case is one run of main.f90.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/channel_flow.h"
#include "cf_internal.hpp"

namespace cf {

std::string last_error_copy();

// first-fit allocator over [0, cap) with coalescing of freed neighbours
struct Arena {
    size_t cap = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    void reset(size_t c) { cap = c; free_.clear(); if (c) free_[0] = c; }
    bool alloc(size_t n, size_t *off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const size_t rest = it->second - n, o2 = it->first + n;
            free_.erase(it);
            if (rest) free_[o2] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && off + n == nx->first) { it->second += nx->second; free_.erase(nx); }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
        }
    }
    size_t free_bytes() const {
        size_t s = 0;
        for (auto &f : free_) s += f.second;
        return s;
    }
    size_t largest() const {
        size_t s = 0;
        for (auto &f : free_) s = std::max(s, f.second);
        return s;
    }
};

static long long unique_cells(const Dims &d) { return (long long)(d.nx - 2) * (d.ny - 1) * (d.nz - 2); }

enum CaseState { PENDING = 0, PREPARING = 1, RUNNING = 2, DONE = 3, CANCELLED = 4, FAILED = 5 };

struct Group;

struct Case {
    long long id = 0;
    std::unique_ptr<Host> host;
    std::vector<double> U, V, W;  // initial field, dropped once uploaded
    int nsteps = 0;
    double arrival = 0;
    std::atomic<int> state{PENDING};
    std::atomic<bool> cancel{false};
    size_t fp = 0, off = 0, footprint = 0;
    int slot = -1;
    Slab slab;
    std::unique_ptr<CtxBase> ctx;
    std::vector<hipEvent_t> ev;
    int launched = 0, completed = 0;
    double t_admit = -1, t_ready = -1, t_done = -1;
    std::vector<float> step_ms;
    std::vector<double> Uo, Vo, Wo;
    double scal[7] = {0, 0, 0, 0, 0, 0, 0};
    std::string err;
    std::string bkey;        // batch class (empty: not batchable)
    Group *grp = nullptr;    // group and member slot while in one
    int mslot = -1;
};

// cases of one size class stepped together (see the file comment)
struct Group {
    std::string key;
    int cap = 0, slot = -1;
    size_t fp = 0, off = 0;
    hipStream_t stream = nullptr;
    std::vector<Slab> slabs;
    std::vector<std::unique_ptr<CtxBase>> ctx;  // per member slot, kept after its case leaves
    std::vector<Case *> mem;                    // case in each member slot (null: free)
    std::vector<Case *> run;                    // members stepped by the current stretch
    std::vector<hipEvent_t> ev;
    int launched = 0, completed = 0;
    std::unique_ptr<Host> proto;  // the first case's host (fills a slot that has no context)
};

struct Server {
    int dev = 0, max_active = 1, depth = 2;
    char *base = nullptr;
    Arena arena;
    std::vector<hipStream_t> streams;
    std::vector<int> free_slots;
    std::mutex mu;      // queues, case records, arena
    std::mutex gpu_mu;  // the scheduler's HIP calls vs. the worker's graph captures
    std::condition_variable cv_done, cv_work;
    std::map<long long, std::unique_ptr<Case>> cases;
    std::deque<Case *> pending;   // submission order
    std::vector<Case *> running;  // RUNNING (owned by the scheduler)
    std::deque<Case *> to_prepare, prepared;
    long long next_id = 1;
    std::thread sched, worker;
    std::atomic<bool> stop{false}, started{false};
    std::chrono::steady_clock::time_point t0;
    size_t peak_used = 0;
    int peak_running = 0;
    long long n_blocked_mem = 0;  // cases that reached the head of the queue and waited for memory
    long long last_blocked = -1;
    int batch_cap = 0, chunk = 25;  // batching: members per group (0: off), steps per stretch
    long long batch_max_cells = 0;  // largest batched case in unique cells (0: any)
    std::vector<std::unique_ptr<Group>> groups;
    long long groups_formed = 0;
    int peak_members = 0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void scheduler();
    void prepare_loop();
    void retire(Case *c, int final_state);
    bool admit_to_group(Case *c, double t);
    bool step_groups();
    int group_boundary(Group *g);
    void group_free(Group *g);
};

static int prepare_case(Server *sv, Case *c) {
    // With batching on, the scheduler captures the groups' step graphs: no
    // HIP call of this thread may overlap such a capture (a synchronous copy
    // during it is refused and invalidates it), so the whole preparation
    // holds the scheduler's lock.  Without batching only this thread
    // captures, and only the capture below takes the lock.
    std::unique_lock<std::mutex> all(sv->gpu_mu, std::defer_lock);
    if (sv->batch_cap > 0) all.lock();
    if (hipSetDevice(sv->dev) != hipSuccess) { c->err = "hipSetDevice failed"; return -1; }
    c->slab.base = sv->base + c->off;
    c->slab.cap = c->fp;
    c->slab.used = 0;
    c->ctx = make_ctx(*c->host, &c->slab, sv->streams[c->slot]);
    if (!c->ctx) { c->err = last_error_copy(); return -1; }
    CtxBase *x = c->ctx.get();
    // initialization.f90:156-160, 460: periodic/no-slip ghosts, one
    // projection of the raw field, bulk-flow target from the result
    if (x->set_state(c->U.data(), c->V.data(), c->W.data()) || x->apply_bc() || x->projection_step(0) ||
        x->apply_bc() || x->init_mass_flow()) {
        c->err = last_error_copy();
        return -1;
    }
    int rc;
    {   // no HIP call of the scheduler overlaps a stream capture
        std::unique_lock<std::mutex> g(sv->gpu_mu, std::defer_lock);
        if (!all.owns_lock()) g.lock();
        rc = x->prepare_graphs();
    }
    if (rc || x->sync()) {
        c->err = last_error_copy();
        return -1;
    }
    std::vector<double>().swap(c->U);
    std::vector<double>().swap(c->V);
    std::vector<double>().swap(c->W);
    c->ev.resize(sv->depth + 2, nullptr);
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) { c->err = "hipEventCreate failed"; return -1; }
    return 0;
}

void Server::prepare_loop() {
    hipSetDevice(dev);
    for (;;) {
        Case *c = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop.load() || !to_prepare.empty(); });
            if (stop.load() && to_prepare.empty()) return;
            c = to_prepare.front();
            to_prepare.pop_front();
        }
        const int rc = c->cancel.load() ? 1 : prepare_case(this, c);
        std::lock_guard<std::mutex> lk(mu);
        if (rc < 0) c->state = FAILED;
        c->t_ready = now();
        prepared.push_back(c);
    }
}

// release a case's slot and memory (its stream has drained); under mu
void Server::retire(Case *c, int final_state) {
    if (c->ctx) {
        if (final_state == DONE) {
            const Dims &d = c->host->d;
            c->Uo.resize((size_t)d.nx * d.nyg * d.nzg);
            c->Vo.resize((size_t)d.nxg * d.ny * d.nzg);
            c->Wo.resize((size_t)d.nxg * d.nyg * d.nz);
            if (c->ctx->get_state(c->Uo.data(), c->Vo.data(), c->Wo.data()) || c->ctx->get_scalars(c->scal)) {
                c->err = last_error_copy();
                final_state = FAILED;
            }
        }
        c->ctx->sync();
        c->ctx.reset();
    }
    for (auto &e : c->ev)
        if (e) hipEventDestroy(e);
    c->ev.clear();
    if (c->slot >= 0) { free_slots.push_back(c->slot); c->slot = -1; }
    if (c->fp && c->state != PENDING) arena.release(c->off, c->fp);
    c->fp = 0;
    c->t_done = now();
    c->state = final_state;
    cv_done.notify_all();
}

// ---- groups ----
// admission of a batchable case (under mu): a free member slot of a group of
// its class, else a new group on a free stream slot with member slots for
// the arrived cases of the class or, if more, as many as half of the largest
// free arena region holds (never above batch_cap or what that region holds),
// so that later arrivals of the class can join it.  The case is prepared at
// the group's next boundary.
bool Server::admit_to_group(Case *c, double t) {
    Group *g = nullptr;
    int m = -1;
    for (auto &gp : groups) {
        if (gp->key != c->bkey) continue;
        for (int q = 0; q < gp->cap; q++)
            if (!gp->mem[q]) { g = gp.get(); m = q; break; }
        if (g) break;
    }
    if (!g) {
        if (free_slots.empty()) return false;
        // room for the arrived cases of the class, and for later arrivals
        // to join as members: up to half of the largest free region
        int want = 0;
        for (Case *p : pending)
            if (p->bkey == c->bkey && p->arrival <= t && !p->cancel.load()) want++;
        const size_t fit = arena.largest() / c->fp;
        const size_t spare = std::max<size_t>(fit / 2, 1);
        const int cap = (int)std::min<size_t>(std::min<size_t>((size_t)batch_cap, fit),
                                              std::max<size_t>((size_t)want, spare));
        size_t off = 0;
        if (cap < 1 || !arena.alloc((size_t)cap * c->fp, &off)) {
            if (c->id != last_blocked) { n_blocked_mem++; last_blocked = c->id; }
            return false;
        }
        std::unique_ptr<Group> ng(new Group);
        ng->key = c->bkey;
        ng->cap = cap;
        ng->fp = c->fp;
        ng->off = off;
        ng->slot = free_slots.back();
        free_slots.pop_back();
        ng->stream = streams[ng->slot];
        ng->slabs.resize(cap);
        ng->ctx.resize(cap);
        ng->mem.assign(cap, nullptr);
        ng->ev.assign(chunk + 1, nullptr);
        ng->proto.reset(new Host(*c->host));
        g = ng.get();
        m = 0;
        groups.push_back(std::move(ng));
        groups_formed++;
        peak_used = std::max(peak_used, arena.cap - arena.free_bytes());
    }
    g->mem[m] = c;
    c->grp = g;
    c->mslot = m;
    c->state = PREPARING;
    c->t_admit = t;
    return true;
}

// a member's state and scalars into its case record, then out of the group
static int group_fetch(Group *g, int m, Case *c) {
    const Dims &d = c->host->d;
    c->Uo.resize((size_t)d.nx * d.nyg * d.nzg);
    c->Vo.resize((size_t)d.nxg * d.ny * d.nzg);
    c->Wo.resize((size_t)d.nxg * d.nyg * d.nz);
    if (g->ctx[m]->get_state(c->Uo.data(), c->Vo.data(), c->Wo.data()) || g->ctx[m]->get_scalars(c->scal)) {
        c->err = last_error_copy();
        return -1;
    }
    return 0;
}

// a fresh context for member slot m with host h (the slot's previous one,
// if any, is dropped), its state parity that of the running members
static int group_new_ctx(Server *sv, Group *g, int m, const Host &h) {
    g->ctx[m].reset();
    g->slabs[m].base = sv->base + g->off + (size_t)m * g->fp;
    g->slabs[m].cap = g->fp;
    g->slabs[m].used = 0;
    g->ctx[m] = make_ctx(h, &g->slabs[m], g->stream);
    if (!g->ctx[m]) return -1;
    if (g->slabs[m].used != g->fp) {
        g->ctx[m].reset();
        set_error("group member layout differs from its footprint");
        return -1;
    }
    g->ctx[m]->batch_member = true;
    for (int q = 0; q < g->cap; q++)
        if (q != m && g->ctx[q] && g->mem[q] && g->mem[q]->state == RUNNING) {
            g->ctx[m]->set_parity(g->ctx[q]->parity());
            break;
        }
    return 0;
}

// the group has drained: retire finished and cancelled members, prepare the
// admitted ones, then launch the next stretch.  Returns 1 when the group is
// empty (nothing running or admitted), -1 on a device error.
int Server::group_boundary(Group *g) {
    for (int m = 0; m < g->cap; m++) {
        Case *c = g->mem[m];
        if (!c || c->state != RUNNING) continue;
        const bool done = c->completed >= c->nsteps;
        if (!done && !c->cancel.load()) continue;
        int st = done ? DONE : CANCELLED;
        if (done && group_fetch(g, m, c)) st = FAILED;
        std::lock_guard<std::mutex> lk(mu);
        g->mem[m] = nullptr;
        c->grp = nullptr;
        c->t_done = now();
        c->state = st;
        cv_done.notify_all();
    }
    for (int m = 0; m < g->cap; m++) {
        Case *c = g->mem[m];
        if (!c || c->state != PREPARING) continue;
        int rc = c->cancel.load() ? 1 : group_new_ctx(this, g, m, *c->host);
        CtxBase *x = rc == 0 ? g->ctx[m].get() : nullptr;
        // initialization.f90:156-160, 460 on the group's stream (drained)
        if (x && (x->set_state(c->U.data(), c->V.data(), c->W.data()) || x->apply_bc() || x->projection_step(0) ||
                  x->apply_bc() || x->init_mass_flow() || x->sync()))
            rc = -1;
        std::lock_guard<std::mutex> lk(mu);
        std::vector<double>().swap(c->U);
        std::vector<double>().swap(c->V);
        std::vector<double>().swap(c->W);
        c->t_ready = now();
        if (rc != 0) {
            if (rc < 0) c->err = last_error_copy();
            g->mem[m] = nullptr;
            c->grp = nullptr;
            c->t_done = now();
            c->state = rc < 0 ? FAILED : CANCELLED;
            cv_done.notify_all();
            continue;
        }
        c->state = RUNNING;
    }
    // the stretch: members [0, hi) (a slot below hi with no context gets a
    // filler with a zero field, so that every stepped slab is initialised)
    int hi = 0, nrun = 0, steps = chunk;
    g->run.clear();
    for (int m = 0; m < g->cap; m++) {
        Case *c = g->mem[m];
        if (c && c->state == RUNNING) {
            hi = m + 1;
            nrun++;
            steps = std::min(steps, c->nsteps - c->completed);
            g->run.push_back(c);
        }
    }
    if (nrun == 0) {
        std::lock_guard<std::mutex> lk(mu);
        for (Case *c : g->mem)
            if (c) return 0;  // admitted, not yet prepared (cannot happen after the loop above)
        return 1;
    }
    for (int m = 0; m < hi; m++) {
        if (g->ctx[m]) continue;
        if (group_new_ctx(this, g, m, *g->proto)) return -1;
        const Dims &d = g->proto->d;
        std::vector<double> zu((size_t)d.nx * d.nyg * d.nzg, 0.0), zv((size_t)d.nxg * d.ny * d.nzg, 0.0),
            zw((size_t)d.nxg * d.nyg * d.nz, 0.0);
        if (g->ctx[m]->set_state(zu.data(), zv.data(), zw.data())) return -1;
    }
    peak_members = std::max(peak_members, nrun);
    // idle slots below hi follow the running members' parity (their data
    // are stepped but read by no one)
    int par = -1;
    for (int m = 0; m < hi && par < 0; m++)
        if (g->mem[m] && g->mem[m]->state == RUNNING) par = g->ctx[m]->parity();
    std::vector<CtxBase *> mem;
    for (int m = 0; m < hi; m++) {
        if (g->ctx[m]->parity() != par) g->ctx[m]->set_parity(par);
        mem.push_back(g->ctx[m].get());
    }
    for (auto &e : g->ev)
        if (!e && hipEventCreate(&e) != hipSuccess) return -1;
    g->launched = g->completed = 0;
    if (hipEventRecord(g->ev[0], g->stream) != hipSuccess) return -1;
    for (int q = 0; q < steps; q++) {
        if (mem[0]->batch_run(mem, (int64_t)g->fp, 1, true)) return -1;
        if (hipEventRecord(g->ev[q + 1], g->stream) != hipSuccess) return -1;
        g->launched++;
    }
    return 0;
}

void Server::group_free(Group *g) {
    if (g->stream) hipStreamSynchronize(g->stream);
    for (auto &x : g->ctx) x.reset();
    for (auto &e : g->ev)
        if (e) hipEventDestroy(e);
    g->ev.clear();
}

// completions of every group's stretch; at a drained group its boundary.
// Returns whether anything happened.
bool Server::step_groups() {
    bool busy = false;
    for (size_t gi = 0; gi < groups.size();) {
        Group *g = groups[gi].get();
        while (g->completed < g->launched) {
            hipEvent_t e = g->ev[g->completed + 1];
            const hipError_t st = hipEventQuery(e);
            if (st == hipErrorNotReady) break;
            float ms = 0.f;
            if (st == hipSuccess) hipEventElapsedTime(&ms, g->ev[g->completed], e);
            std::lock_guard<std::mutex> lk(mu);
            for (Case *c : g->run) {
                c->step_ms.push_back(ms);
                c->completed++;
            }
            g->completed++;
            busy = true;
        }
        if (g->completed < g->launched) { gi++; continue; }
        const int rc = group_boundary(g);
        if (rc == 0) { gi++; busy = busy || g->launched > 0; continue; }
        // empty, or a device error: fail what is left, release the group
        group_free(g);
        std::lock_guard<std::mutex> lk(mu);
        for (int m = 0; m < g->cap; m++) {
            Case *c = g->mem[m];
            if (!c) continue;
            c->err = rc < 0 ? last_error_copy() : std::string("group released");
            c->grp = nullptr;
            c->t_done = now();
            c->state = FAILED;
        }
        if (rc < 0) cv_done.notify_all();
        free_slots.push_back(g->slot);
        arena.release(g->off, (size_t)g->cap * g->fp);
        groups.erase(groups.begin() + gi);
        busy = true;
    }
    size_t nrun = running.size();
    for (auto &g : groups) nrun += g->run.size();
    peak_running = std::max(peak_running, (int)nrun);
    return busy;
}

void Server::scheduler() {
    hipSetDevice(dev);
    while (!stop.load()) {
        bool busy = false;
        // every HIP call of an iteration happens under gpu_mu, so none
        // overlaps a graph capture of the admission worker (lock order:
        // gpu_mu, then mu)
        std::unique_lock<std::mutex> gpu(gpu_mu);
        {
            std::lock_guard<std::mutex> lk(mu);
            // cases the worker has finished preparing
            while (!prepared.empty()) {
                Case *c = prepared.front();
                prepared.pop_front();
                if (c->state == FAILED) { retire(c, FAILED); continue; }
                if (c->cancel.load()) { retire(c, CANCELLED); continue; }
                c->state = RUNNING;
                running.push_back(c);
                peak_running = std::max(peak_running, (int)running.size());
            }
            // admission: submission order, once arrived, if a slot and memory are free
            const double t = now();
            while (!pending.empty()) {
                Case *c = pending.front();
                if (c->cancel.load()) { pending.pop_front(); c->state = CANCELLED; c->t_done = t; cv_done.notify_all(); continue; }
                if (c->arrival > t) break;
                if (batch_cap > 0 && !c->bkey.empty() &&
                    (batch_max_cells == 0 || unique_cells(c->host->d) <= batch_max_cells)) {
                    if (!admit_to_group(c, t)) break;
                    pending.pop_front();
                    busy = true;
                    continue;
                }
                if (free_slots.empty()) break;
                size_t off = 0;
                if (!arena.alloc(c->fp, &off)) {
                    if (c->id != last_blocked) { n_blocked_mem++; last_blocked = c->id; }
                    break;
                }
                pending.pop_front();
                c->off = off;
                c->slot = free_slots.back();
                free_slots.pop_back();
                c->state = PREPARING;
                c->t_admit = t;
                peak_used = std::max(peak_used, arena.cap - arena.free_bytes());
                to_prepare.push_back(c);
                cv_work.notify_one();
                busy = true;
            }
        }
        // completions and launches of running cases (owned by this thread)
        for (size_t q = 0; q < running.size();) {
            Case *c = running[q];
            const int R = (int)c->ev.size();
            bool failed = false;
            float got[64];
            int ngot = 0;
            while (c->completed < c->launched && ngot < 64) {
                hipEvent_t e = c->ev[(c->completed + 1) % R];
                const hipError_t st = hipEventQuery(e);
                if (st == hipErrorNotReady) break;
                if (st != hipSuccess) { c->err = hipGetErrorString(st); failed = true; break; }
                float ms = 0.f;
                hipEventElapsedTime(&ms, c->ev[c->completed % R], e);
                got[ngot++] = ms;
                c->completed++;
                busy = true;
            }
            if (ngot) {
                std::lock_guard<std::mutex> lk(mu);
                c->step_ms.insert(c->step_ms.end(), got, got + ngot);
            }
            const bool finished = c->completed == c->nsteps;
            if (failed || finished || c->cancel.load()) {
                if (!finished && !failed) hipStreamSynchronize(c->ctx->get_stream());
                std::lock_guard<std::mutex> lk(mu);
                retire(c, failed ? FAILED : (finished ? DONE : CANCELLED));
                running.erase(running.begin() + q);
                busy = true;
                continue;
            }
            while (c->launched < c->nsteps && c->launched - c->completed < depth) {
                hipStream_t s = c->ctx->get_stream();
                if (c->launched == 0 && hipEventRecord(c->ev[0], s) != hipSuccess) { failed = true; break; }
                if (c->ctx->run_steps(1, true) || hipEventRecord(c->ev[(c->launched + 1) % R], s) != hipSuccess) {
                    c->err = last_error_copy();
                    failed = true;
                    break;
                }
                c->launched++;
                busy = true;
            }
            if (failed) {
                hipStreamSynchronize(c->ctx->get_stream());
                std::lock_guard<std::mutex> lk(mu);
                retire(c, FAILED);
                running.erase(running.begin() + q);
                continue;
            }
            q++;
        }
        if (!groups.empty() && step_groups()) busy = true;
        gpu.unlock();
        if (!busy) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

}  // namespace cf

using cf::Case;
using cf::Server;
using cf::set_error;

struct cf_server {
    Server s;
};

namespace {
Case *find_case(cf_server *sv, long long id) {
    auto it = sv->s.cases.find(id);
    return it == sv->s.cases.end() ? nullptr : it->second.get();
}
double pct(std::vector<float> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double r = p / 100.0 * (double)(v.size() - 1);
    const size_t a = (size_t)r;
    const size_t b = std::min(a + 1, v.size() - 1);
    return v[a] + (r - (double)a) * (v[b] - v[a]);
}
}  // namespace

extern "C" {

int cf_server_create(int device, long long hbm_budget, int max_active, int depth, cf_server **out) {
    if (!out || hbm_budget <= 0 || max_active < 1 || depth < 1) {
        set_error("cf_server_create: bad arguments");
        return -1;
    }
    *out = nullptr;
    std::unique_ptr<cf_server> p(new cf_server);
    Server &s = p->s;
    s.dev = device;
    s.max_active = max_active;
    s.depth = depth;
    if (hipSetDevice(device) != hipSuccess) { set_error("cf_server_create: bad device"); return -1; }
    if (hipMalloc(&s.base, (size_t)hbm_budget) != hipSuccess) {
        set_error("cf_server_create: cannot reserve the HBM budget");
        return -1;
    }
    s.arena.reset((size_t)hbm_budget);
    s.streams.resize(max_active);
    for (int q = 0; q < max_active; q++) {
        if (hipStreamCreateWithFlags(&s.streams[q], hipStreamNonBlocking) != hipSuccess) {
            set_error("cf_server_create: stream creation failed");
            for (int r = 0; r < q; r++) hipStreamDestroy(s.streams[r]);
            hipFree(s.base);
            return -1;
        }
        s.free_slots.push_back(max_active - 1 - q);
    }
    *out = p.release();
    return 0;
}

int cf_server_set_batching(cf_server *sv, int max_members, int chunk, long long max_cells) {
    if (!sv || max_members < 0 || chunk < 1 || max_cells < 0) {
        set_error("cf_server_set_batching: bad arguments");
        return -1;
    }
    if (sv->s.started.load()) { set_error("cf_server_set_batching: server already started"); return -1; }
    sv->s.batch_cap = max_members;
    sv->s.chunk = chunk;
    sv->s.batch_max_cells = max_cells;
    return 0;
}

int cf_server_start(cf_server *sv) {
    if (!sv) { set_error("cf_server_start: null server"); return -1; }
    Server &s = sv->s;
    if (s.started.exchange(true)) return 0;
    s.t0 = std::chrono::steady_clock::now();
    s.worker = std::thread([&s] { s.prepare_loop(); });
    s.sched = std::thread([&s] { s.scheduler(); });
    return 0;
}

int cf_server_destroy(cf_server *sv) {
    if (!sv) return 0;
    Server &s = sv->s;
    {
        std::lock_guard<std::mutex> lk(s.mu);
        for (auto &kv : s.cases) kv.second->cancel = true;
    }
    // let the scheduler drain cancelled cases, then stop both threads
    if (s.started.load()) {
        for (int spin = 0; spin < 200000; spin++) {
            {
                std::lock_guard<std::mutex> lk(s.mu);
                if (s.running.empty() && s.to_prepare.empty() && s.prepared.empty() && s.groups.empty()) {
                    bool any = false;
                    for (auto &kv : s.cases) any = any || kv.second->state == cf::PREPARING;
                    if (!any) break;
                }
            }
            std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
        s.stop = true;
        s.cv_work.notify_all();
        if (s.sched.joinable()) s.sched.join();
        if (s.worker.joinable()) s.worker.join();
    }
    hipSetDevice(s.dev);
    for (auto &g : s.groups) s.group_free(g.get());
    s.groups.clear();
    for (Case *c : s.running) {
        if (c->ctx) c->ctx->sync();
        c->ctx.reset();
    }
    for (auto &kv : s.cases) {
        Case *c = kv.second.get();
        if (c->ctx) { c->ctx->sync(); c->ctx.reset(); }
        for (auto &e : c->ev)
            if (e) hipEventDestroy(e);
    }
    for (auto st : s.streams) hipStreamDestroy(st);
    if (s.base) hipFree(s.base);
    delete sv;
    return 0;
}

int cf_server_submit(cf_server *sv, const cf_params *p, const double *x, const double *y, const double *z,
                     const double *U, const double *V, const double *W, int nsteps, double arrival,
                     long long *id) {
    if (!sv || !p || !x || !y || !z || !U || !V || !W || !id || nsteps < 1) {
        set_error("cf_server_submit: bad arguments");
        return -1;
    }
    Server &s = sv->s;
    if (p->device != s.dev) { set_error("cf_server_submit: case device differs from the server's"); return -1; }
    std::unique_ptr<Case> c(new Case);
    c->host.reset(new cf::Host(*p, x, y, z));
    const std::string v = c->host->validate();
    if (!v.empty()) { set_error("cf_server_submit: " + v); return -1; }
    c->fp = c->footprint = cf::ctx_footprint(*c->host);
    c->bkey = cf::batch_class_key(*c->host);
    if (c->fp == 0) return -1;
    if (c->fp > s.arena.cap) { set_error("cf_server_submit: case footprint exceeds the HBM budget"); return -1; }
    const cf::Dims &d = c->host->d;
    c->U.assign(U, U + (size_t)d.nx * d.nyg * d.nzg);
    c->V.assign(V, V + (size_t)d.nxg * d.ny * d.nzg);
    c->W.assign(W, W + (size_t)d.nxg * d.nyg * d.nz);
    c->nsteps = nsteps;
    c->arrival = arrival;
    std::lock_guard<std::mutex> lk(s.mu);
    c->id = s.next_id++;
    *id = c->id;
    s.pending.push_back(c.get());
    s.cases[c->id] = std::move(c);
    return 0;
}

int cf_server_cancel(cf_server *sv, long long id) {
    if (!sv) { set_error("cf_server_cancel: null server"); return -1; }
    std::lock_guard<std::mutex> lk(sv->s.mu);
    Case *c = find_case(sv, id);
    if (!c) { set_error("cf_server_cancel: unknown case"); return -1; }
    c->cancel = true;
    return 0;
}

int cf_server_wait(cf_server *sv, double timeout_s) {
    if (!sv) { set_error("cf_server_wait: null server"); return -1; }
    Server &s = sv->s;
    if (!s.started.load()) { set_error("cf_server_wait: server not started"); return -1; }
    std::unique_lock<std::mutex> lk(s.mu);
    auto all_done = [&] {
        for (auto &kv : s.cases)
            if (kv.second->state < cf::DONE) return false;
        return true;
    };
    if (timeout_s <= 0) { s.cv_done.wait(lk, all_done); return 0; }
    return s.cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), all_done) ? 0 : 1;
}

int cf_server_status(cf_server *sv, long long id, int *state, int *steps_done) {
    if (!sv) { set_error("cf_server_status: null server"); return -1; }
    std::lock_guard<std::mutex> lk(sv->s.mu);
    Case *c = find_case(sv, id);
    if (!c) { set_error("cf_server_status: unknown case"); return -1; }
    if (state) *state = c->state.load();
    if (steps_done) *steps_done = (int)c->step_ms.size();
    return 0;
}

int cf_server_result(cf_server *sv, long long id, double *U, double *V, double *W, double *scalars) {
    if (!sv) { set_error("cf_server_result: null server"); return -1; }
    std::lock_guard<std::mutex> lk(sv->s.mu);
    Case *c = find_case(sv, id);
    if (!c) { set_error("cf_server_result: unknown case"); return -1; }
    if (c->state != cf::DONE) {
        set_error(c->state == cf::FAILED ? "cf_server_result: case failed: " + c->err
                                         : std::string("cf_server_result: case not finished"));
        return -1;
    }
    if (U) std::memcpy(U, c->Uo.data(), c->Uo.size() * sizeof(double));
    if (V) std::memcpy(V, c->Vo.data(), c->Vo.size() * sizeof(double));
    if (W) std::memcpy(W, c->Wo.data(), c->Wo.size() * sizeof(double));
    if (scalars) std::memcpy(scalars, c->scal, sizeof c->scal);
    return 0;
}

int cf_server_case_info(cf_server *sv, long long id, double *info, int n) {
    if (!sv || !info) { set_error("cf_server_case_info: null argument"); return -1; }
    std::lock_guard<std::mutex> lk(sv->s.mu);
    Case *c = find_case(sv, id);
    if (!c) { set_error("cf_server_case_info: unknown case"); return -1; }
    const cf::Dims &d = c->host->d;
    const double v[14] = {c->arrival, c->t_admit, c->t_ready, c->t_done, (double)c->step_ms.size(),
                          (double)c->nsteps, (double)c->footprint, pct(c->step_ms, 50),
                          pct(c->step_ms, 99), pct(c->step_ms, 100), (double)c->state.load(),
                          (double)(d.nxp) * (d.nyp) * (d.nzp), (double)c->host->prm.precision,
                          (double)c->host->prm.iwall_model};
    for (int q = 0; q < n && q < 14; q++) info[q] = v[q];
    return 0;
}

int cf_server_step_times(cf_server *sv, long long id, float *ms, int n) {
    if (!sv || !ms) { set_error("cf_server_step_times: null argument"); return -1; }
    std::lock_guard<std::mutex> lk(sv->s.mu);
    Case *c = find_case(sv, id);
    if (!c) { set_error("cf_server_step_times: unknown case"); return -1; }
    const int m = std::min(n, (int)c->step_ms.size());
    std::memcpy(ms, c->step_ms.data(), (size_t)m * sizeof(float));
    return m;
}

int cf_server_summary(cf_server *sv, double *out, int n) {
    if (!sv || !out) { set_error("cf_server_summary: null argument"); return -1; }
    Server &s = sv->s;
    std::lock_guard<std::mutex> lk(s.mu);
    double last = 0;
    for (auto &kv : s.cases) last = std::max(last, kv.second->t_done);
    const double v[8] = {last, (double)s.peak_used, (double)s.peak_running, (double)s.n_blocked_mem,
                         (double)s.arena.cap, (double)s.cases.size(), (double)s.groups_formed,
                         (double)s.peak_members};
    for (int q = 0; q < n && q < 8; q++) out[q] = v[q];
    return 0;
}

}  // extern "C"
