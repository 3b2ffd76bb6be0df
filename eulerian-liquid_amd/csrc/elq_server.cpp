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
#include <mutex>
#include <thread>
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

size_t elq_server_device_bytes(const elq_server* S) try {
    if (!S) return 0;
    auto* s = const_cast<elq_server*>(S);
    std::lock_guard<std::mutex> lk(s->mu);
    return s->held_total();
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return 0;
}

}  // extern "C"
