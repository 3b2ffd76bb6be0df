// Serving: continuous batching of independent projections on one device.
//
// Each slot is a full solver context (own HIP stream, own device workspace)
// driven by its own host thread; the kernels of different slots overlap on
// the device. The persistent kernels of a slot (factor wave, SpMV, pencil
// sweeps) wait on each other's workgroups and therefore need all of their
// workgroups resident; every slot's grids are capped at
// 2 * CUs / max_slots workgroups, so that all slots together never ask for
// more than two 256-thread workgroups per CU, below what any of these
// kernels can keep resident per CU. The cap depends only on the server's
// configuration, which keeps a request's arithmetic (and its reduction
// order) the same whatever else is running.
//
// Batched Krylov chunks (elq_server_opts.batched, mixed precision): each
// slot thread still runs its projection (set-up, polls, residual
// replacements, scatter and update on its own stream), but hands every Krylov
// chunk (check_every iterations) to the Batcher, and the chunks of all slots
// inside their Krylov loops advance together: one launch per kernel class
// and iteration for all of them (launch_krylov_batch), recorded once into a
// HIP graph per job table and set of sweep shapes and replayed for every
// batch after. The launch arguments never change (the grid is max_slots
// shares of the per-slot cap, the job table sits at a fixed device address
// and is uploaded before each batch), so the graph serves every request
// that ever runs. The first slot to arrive leads the batch: it waits until
// every slot inside its Krylov loop has joined (at most batch_wait_us), then
// enqueues on the batcher's stream, after each member's earlier work, the
// table upload and the graph, and makes each member's stream wait for the
// batch. A member's workgroups are those its direct launches would use, so
// its result is the same batched or not, alone or in any mix.
//
// Admission: FIFO queue; the front request is admitted when a slot is free
// and the device bytes held by all slots, with the admitted slot bounded by
// what it holds plus the request's footprint (elq_projection_footprint), fit
// the budget. Idle slots give their workspace back when that is what keeps a
// request out. A request that does not fit lets later ones pass at most
// kMaxSkips times.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/elq.h"
#include "elq_api_internal.hpp"

namespace {

constexpr int kMaxSlots = 16;
constexpr int kMaxSkips = 8;

struct Request {
    int64_t id = 0;
    elq_grid g{};
    float dt = 0.0f;
    size_t need = 0;  // footprint
    int skips = 0;
    double t_submit = 0.0;
    double t_admit = 0.0;
};

struct Slot {
    elq_ctx* ctx = nullptr;
    std::thread th;
    bool busy = false;
    Request req;
    size_t held = 0;  // device bytes this slot holds (busy: bound for the running request)
};

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

constexpr int kDefaultBatchWaitUs = 20000;

struct Batcher final : elq::KrylovBatcher {
    int nslots = 0, per_job = 0, wait_us = kDefaultBatchWaitUs;
    hipStream_t stream = nullptr;
    elq::KrylovJob* dev[2] = {nullptr, nullptr};   // job tables, one per graph family
    elq::KrylovJob* host[2] = {nullptr, nullptr};  // pinned staging of each table
    hipEvent_t uploaded[2] = {nullptr, nullptr};   // the table's last upload has run
    hipEvent_t done = nullptr;                     // the last batch has run
    std::vector<hipEvent_t> ready;                 // per slot: its work before the chunk
    std::vector<elq_ctx*> ctxs;
    std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;  // (table, shapes, n)
    std::mutex mu;
    std::condition_variable cv;
    uint64_t round = 0;
    bool leading = false;     // a leader exists for the coming batch
    bool collecting = false;  // ... and still takes members
    std::vector<char> joined, in_krylov;
    std::vector<int> rc_slot;
    std::vector<elq::KrylovJob> pending;
    int64_t batches = 0, jobs = 0;

    int init(int slots, int groups, int wait) {
        nslots = slots;
        per_job = groups;
        if (wait > 0) wait_us = wait;
        ready.assign((size_t)slots, nullptr);
        ctxs.assign((size_t)slots, nullptr);
        joined.assign((size_t)slots, 0);
        in_krylov.assign((size_t)slots, 0);
        rc_slot.assign((size_t)slots, ELQ_OK);
        pending.assign((size_t)slots, elq::KrylovJob{});
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return ELQ_ERR_DEVICE;
        for (int b = 0; b < 2; ++b) {
            if (hipMalloc((void**)&dev[b], sizeof(elq::KrylovJob) * (size_t)slots) != hipSuccess)
                return ELQ_ERR_DEVICE;
            if (hipHostMalloc((void**)&host[b], sizeof(elq::KrylovJob) * (size_t)slots, hipHostMallocDefault) !=
                hipSuccess)
                return ELQ_ERR_DEVICE;
            if (hipEventCreateWithFlags(&uploaded[b], hipEventDisableTiming) != hipSuccess) return ELQ_ERR_DEVICE;
        }
        if (hipEventCreateWithFlags(&done, hipEventDisableTiming) != hipSuccess) return ELQ_ERR_DEVICE;
        for (auto& e : ready)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return ELQ_ERR_DEVICE;
        return ELQ_OK;
    }

    // Releases everything init() created (after the slot threads have ended).
    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& kv : graphs)
            if (kv.second) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
        for (auto& e : ready)
            if (e) (void)hipEventDestroy(e);
        ready.clear();
        if (done) (void)hipEventDestroy(done);
        done = nullptr;
        for (int b = 0; b < 2; ++b) {
            if (uploaded[b]) (void)hipEventDestroy(uploaded[b]);
            if (dev[b]) (void)hipFree(dev[b]);
            if (host[b]) (void)hipHostFree(host[b]);
            uploaded[b] = nullptr;
            dev[b] = nullptr;
            host[b] = nullptr;
        }
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    // Slot k's request has left its Krylov loop (for good, or for a residual
    // replacement, after which it rejoins): no batch waits for it meanwhile.
    void leave(int k) override {
        std::lock_guard<std::mutex> lk(mu);
        in_krylov[(size_t)k] = 0;
        cv.notify_all();
    }

    bool all_joined() const {
        for (int j = 0; j < nslots; ++j)
            if (in_krylov[(size_t)j] && !joined[(size_t)j]) return false;
        return true;
    }

    // Enqueues one batch (mu not held): after each member's earlier work, the
    // table upload and the chunk's graph; each member's stream then waits.
    int launch(int buf, int n, int shapes, const std::vector<int>& members) {
        for (int j : members)
            if (hipStreamWaitEvent(stream, ready[(size_t)j], 0) != hipSuccess) return ELQ_ERR_DEVICE;
        if (hipMemcpyAsync(dev[buf], host[buf], sizeof(elq::KrylovJob) * (size_t)nslots, hipMemcpyHostToDevice,
                           stream) != hipSuccess)
            return ELQ_ERR_DEVICE;
        if (hipEventRecord(uploaded[buf], stream) != hipSuccess) return ELQ_ERR_DEVICE;
        hipGraphExec_t& g = graphs[std::make_tuple(buf, shapes, n)];
        if (!g) {
            hipGraph_t gr = nullptr;
            if (hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed) != hipSuccess) return ELQ_ERR_DEVICE;
            const hipError_t e1 = elq::launch_krylov_batch(dev[buf], nslots, per_job, n, shapes, stream);
            const hipError_t e2 = hipStreamEndCapture(stream, &gr);
            if (e1 != hipSuccess || e2 != hipSuccess || !gr) {
                if (gr) (void)hipGraphDestroy(gr);
                return ELQ_ERR_DEVICE;
            }
            const hipError_t e3 = hipGraphInstantiate(&g, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e3 != hipSuccess) {
                g = nullptr;
                return ELQ_ERR_DEVICE;
            }
        }
        if (hipGraphLaunch(g, stream) != hipSuccess) return ELQ_ERR_DEVICE;
        if (hipEventRecord(done, stream) != hipSuccess) return ELQ_ERR_DEVICE;
        for (int j : members)
            if (hipStreamWaitEvent(ctxs[(size_t)j]->stream, done, 0) != hipSuccess) return ELQ_ERR_DEVICE;
        return ELQ_OK;
    }

    int run(elq_ctx* ctx, int k, const elq::KrylovJob& job, int n) override {
        if (hipEventRecord(ready[(size_t)k], ctx->stream) != hipSuccess)
            return ctx->fail(ELQ_ERR_DEVICE, "batched Krylov chunk: hipEventRecord failed");
        std::unique_lock<std::mutex> lk(mu);
        in_krylov[(size_t)k] = 1;
        for (;;) {
            if (!leading) break;  // lead the next batch
            if (collecting) {     // join the batch being collected
                pending[(size_t)k] = job;
                joined[(size_t)k] = 1;
                const uint64_t my = round;
                cv.notify_all();  // the leader re-checks whether everyone has joined
                cv.wait(lk, [&] { return round != my; });
                const int rc = rc_slot[(size_t)k];
                lk.unlock();
                return rc == ELQ_OK ? rc : ctx->fail(rc, "batched Krylov chunk failed to launch");
            }
            cv.wait(lk, [&] { return !leading; });  // a batch is launching without this slot
        }
        leading = true;
        collecting = true;
        pending[(size_t)k] = job;
        joined[(size_t)k] = 1;
        const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(wait_us);
        cv.wait_until(lk, until, [&] { return all_joined(); });
        collecting = false;
        const int buf = (int)(round & 1);
        // the previous upload of this table (two batches ago) has run before
        // the table is rewritten
        (void)hipEventSynchronize(uploaded[buf]);
        std::vector<int> members;
        int shapes = 0;
        for (int j = 0; j < nslots; ++j) {
            if (joined[(size_t)j]) {
                members.push_back(j);
                host[buf][j] = pending[(size_t)j];
                shapes |= 1 << pending[(size_t)j].pc.shape;
                joined[(size_t)j] = 0;
            } else {
                host[buf][j] = elq::KrylovJob{};
            }
        }
        lk.unlock();
        const int rc = launch(buf, n, shapes, members);
        lk.lock();
        for (int j : members) rc_slot[(size_t)j] = rc;
        ++round;
        ++batches;
        jobs += (int64_t)members.size();
        leading = false;
        cv.notify_all();
        lk.unlock();
        return rc == ELQ_OK ? rc : ctx->fail(rc, "batched Krylov chunk failed to launch");
    }
};

}  // namespace

struct elq_server {
    int device = 0;
    elq_server_opts o{};
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> queue;
    std::deque<elq_result> done;
    std::vector<Slot> slots;
    int64_t next_id = 1;
    int running = 0;
    bool stop = false;
    // batched Krylov chunks (opts.batched = number of lockstep groups):
    // slot k belongs to group k % groups as its member k / groups
    std::vector<Batcher*> batch;

    size_t held_total() const {
        size_t n = 0;
        for (const Slot& s : slots) n += s.held;
        return n;
    }

    // Whether a request of footprint need can run on the free slot k. A slot
    // that already holds buffers may grow each of them, so it is bounded by
    // held + need; idle workspaces (this slot's first) are given back when
    // that is what keeps the request out. Called with mu held.
    bool fits(size_t need, int k) {
        auto total = [&] {
            const size_t h = slots[k].held;
            return held_total() - h + (h ? h + need : need);
        };
        if (total() <= o.hbm_budget_bytes) return true;
        for (size_t step = 0; step <= slots.size(); ++step) {
            const size_t q = step == 0 ? (size_t)k : step - 1;
            Slot& s = slots[q];
            if (s.busy || s.held == 0) continue;
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(s.ctx->stream);
            s.ctx->ws.release_all();
            s.held = 0;
            if (total() <= o.hbm_budget_bytes) return true;
        }
        return false;
    }

    void admit() {
        while (!queue.empty()) {
            int k = -1;
            for (size_t q = 0; q < slots.size(); ++q)
                if (!slots[q].busy) {
                    k = (int)q;
                    break;
                }
            if (k < 0) return;
            const size_t scan = queue.front().skips >= kMaxSkips ? 1 : queue.size();
            size_t pick = scan;
            for (size_t i = 0; i < scan; ++i)
                if (fits(queue[i].need, k)) {
                    pick = i;
                    break;
                }
            if (pick == scan) return;
            for (size_t i = 0; i < pick; ++i) queue[i].skips += 1;
            Slot& s = slots[k];
            s.req = queue[pick];
            queue.erase(queue.begin() + (long)pick);
            s.req.t_admit = ms_since(t0);
            s.held = s.held ? s.held + s.req.need : s.req.need;
            s.busy = true;
            ++running;
            cv_work.notify_all();
        }
    }

    void worker(int k) {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_work.wait(lk, [&] { return stop || slots[k].busy; });
            if (!slots[k].busy) return;  // stop with nothing assigned
            Request r = slots[k].req;
            elq_ctx* ctx = slots[k].ctx;
            lk.unlock();
            elq_stats st{};
            const int rc = elq_project(ctx, &r.g, r.dt, &o.opts, &st);
            if (!batch.empty()) batch[(size_t)k % batch.size()]->leave(k / (int)batch.size());
            const double t_done = ms_since(t0);
            const size_t bytes = ctx->ws.bytes();
            lk.lock();
            elq_result res{};
            res.id = r.id;
            res.status = rc;
            res.iterations = st.iterations;
            res.outer_iterations = st.outer_iterations;
            res.slot = k;
            res.n_unknowns = st.n_unknowns;
            res.error = st.error;
            res.true_rel_residual = st.true_rel_residual;
            res.t_submit_ms = r.t_submit;
            res.t_admit_ms = r.t_admit;
            res.t_done_ms = t_done;
            done.push_back(res);
            slots[k].busy = false;
            slots[k].held = bytes;
            --running;
            if (!stop) admit();
            cv_done.notify_all();
        }
    }
};

extern "C" {

int elq_server_create(int device, const elq_server_opts* opts, elq_server** out) try {
    if (!out) return ELQ_ERR_INVALID_ARG;
    *out = nullptr;
    if (!opts || opts->max_slots < 1 || opts->max_slots > kMaxSlots || opts->max_queue < 1)
        return ELQ_ERR_INVALID_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return ELQ_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return ELQ_ERR_DEVICE;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
        return ELQ_ERR_DEVICE;
    elq_server* S = new elq_server();
    S->device = device;
    S->o = *opts;
    if (S->o.hbm_budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete S;
            return ELQ_ERR_DEVICE;
        }
        S->o.hbm_budget_bytes = fr / 10 * 9;
    }
    const int groups = std::max(1, 2 * cus / opts->max_slots);
    S->slots.resize((size_t)opts->max_slots);
    for (auto& sl : S->slots) {
        const int rc = elq_ctx_create(device, 0, &sl.ctx);
        if (rc != ELQ_OK) {
            for (auto& t : S->slots)
                if (t.ctx) elq_ctx_destroy(t.ctx);
            delete S;
            return rc;
        }
        sl.ctx->max_groups = groups;
    }
    if (opts->batched > 0 && opts->opts.precision == ELQ_PRECISION_MIXED &&
        opts->opts.schedule != ELQ_SCHED_LEVELS) {
        const int ng = std::min(opts->batched, opts->max_slots);
        for (int gi = 0; gi < ng; ++gi) {
            Batcher* B = new Batcher();
            S->batch.push_back(B);
            const int members = (opts->max_slots - gi + ng - 1) / ng;
            const int rc = B->init(members, groups, opts->batch_wait_us);
            if (rc != ELQ_OK) {
                for (Batcher* b : S->batch) {
                    b->release();
                    delete b;
                }
                for (auto& t : S->slots) elq_ctx_destroy(t.ctx);
                delete S;
                return rc;
            }
        }
        for (int k = 0; k < opts->max_slots; ++k) {
            Batcher* B = S->batch[(size_t)k % (size_t)ng];
            B->ctxs[(size_t)(k / ng)] = S->slots[(size_t)k].ctx;
            S->slots[(size_t)k].ctx->batcher = B;
            S->slots[(size_t)k].ctx->batch_slot = k / ng;
        }
    }
    S->t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < opts->max_slots; ++k) {
        try {
            S->slots[(size_t)k].th = std::thread(&elq_server::worker, S, k);
        } catch (...) {  // no host thread for a slot: undo and report, nothing throws past the C ABI
            elq_server_destroy(S);
            return ELQ_ERR_DEVICE;
        }
    }
    *out = S;
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

void elq_server_destroy(elq_server* S) try {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->queue.clear();
        S->cv_work.notify_all();
    }
    for (auto& sl : S->slots)
        if (sl.th.joinable()) sl.th.join();
    for (Batcher* b : S->batch) {
        b->release();
        delete b;
    }
    for (auto& sl : S->slots)
        if (sl.ctx) sl.ctx->batcher = nullptr;
    S->batch.clear();
    for (auto& sl : S->slots) elq_ctx_destroy(sl.ctx);
    delete S;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
}

int elq_server_submit(elq_server* S, const elq_grid* g, float dt, int64_t* id) try {
    if (!S || !g || !g->level_set || !g->u || !g->v || !g->w || !g->u_weight || !g->v_weight || !g->w_weight)
        return ELQ_ERR_INVALID_ARG;
    const size_t need = elq_projection_footprint(g->ni, g->nj, g->nk, S->o.opts.precision, g->level_set);
    if (need == 0) return ELQ_ERR_INVALID_ARG;
    if (need > S->o.hbm_budget_bytes) return ELQ_ERR_ADMISSION;
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return ELQ_ERR_INVALID_ARG;
    if ((int)S->queue.size() >= S->o.max_queue) return ELQ_ERR_BUSY;
    Request r;
    r.id = S->next_id++;
    r.g = *g;
    r.dt = dt;
    r.need = need;
    r.t_submit = ms_since(S->t0);
    S->queue.push_back(r);
    if (id) *id = r.id;
    S->admit();
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_server_wait(elq_server* S, int timeout_ms) try {
    if (!S) return 0;
    std::unique_lock<std::mutex> lk(S->mu);
    auto ready = [&] { return !S->done.empty() || (S->queue.empty() && S->running == 0); };
    if (timeout_ms < 0)
        S->cv_done.wait(lk, ready);
    else
        S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    return (int)S->done.size();
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return 0;
}

int elq_server_poll(elq_server* S, elq_result* out, int cap) try {
    if (!S || !out || cap <= 0) return 0;
    std::lock_guard<std::mutex> lk(S->mu);
    int n = 0;
    while (n < cap && !S->done.empty()) {
        out[n++] = S->done.front();
        S->done.pop_front();
    }
    return n;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return 0;
}

int elq_server_in_flight(const elq_server* S) try {
    if (!S) return 0;
    auto* s = const_cast<elq_server*>(S);
    std::lock_guard<std::mutex> lk(s->mu);
    return (int)s->queue.size() + s->running;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return 0;
}

int elq_server_batch_stats(const elq_server* S, int64_t* batches, int64_t* jobs) try {
    if (!S || !batches || !jobs) return ELQ_ERR_INVALID_ARG;
    *batches = 0;
    *jobs = 0;
    for (Batcher* b : S->batch) {
        std::lock_guard<std::mutex> lk(b->mu);
        *batches += b->batches;
        *jobs += b->jobs;
    }
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

size_t elq_server_device_bytes(const elq_server* S) try {
    if (!S) return 0;
    auto* s = const_cast<elq_server*>(S);
    std::lock_guard<std::mutex> lk(s->mu);
    return s->held_total();
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return 0;
}

}  // extern "C"
