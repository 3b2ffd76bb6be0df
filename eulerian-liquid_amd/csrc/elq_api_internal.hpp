// Context and workspace of the C ABI (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "elq_kernels.hpp"

// Bytes of the pinned host mirror a context keeps for its polls.
constexpr size_t kHostMirrorBytes = 4096;

// Bump allocator over a fixed device region (the allocation-free entry
// point's static scratch array): carving never calls the HIP allocator, and
// a request beyond the region fails with overflow set.
struct Arena {
    uint8_t* base = nullptr;
    size_t size = 0;
    size_t off = 0;
    bool overflow = false;
};

// Device buffer that only grows (hipMalloc on growth, reused otherwise). With
// an arena it is carved from the arena instead and never freed.
// A device buffer that grows on demand. Outside an arena it is allocated and
// freed in the order of `stream` (hipMallocAsync / hipFreeAsync): growing or
// releasing it never waits for work on other streams, as hipFree would (it
// synchronises the device, which stalled every other serving slot).
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    Arena* arena = nullptr;
    hipStream_t stream = nullptr;
    hipError_t ensure(size_t bytes);
    void release();
};

// Device buffers of one projection; sized for the largest problem seen.
struct Workspace {
    // dense
    DevBuf mask, indices, pressure;
    // scans and brick metadata
    DevBuf scan_cnt, scan_off, bflag, map, coord, lvl_cnt, lvl_off, lvl_fill, info;
    DevBuf nbr, lvl_list, xch, stage;
    DevBuf pflag, pmap, bkflag, blist;
    // system
    DevBuf diag64, b64, inv_e64, inv_e32, cdiag, cax, cay, caz;
    DevBuf cscaled;  // FP64: pivot-scaled lower couplings lx, ly, lz of the forward sweep
    DevBuf uflag, umap, urow;  // uniform-brick flags (per brick, per position) and constant coefficient rows
    // Krylov vectors (storage type of the solve)
    DevBuf vx, vr, vp0, vp1, vq, vy, x64;
    DevBuf part_rr, part_rz, ctrl, nfixed;
    // staging of host inputs
    DevBuf s_phi, s_u, s_v, s_w, s_uw, s_vw, s_ww, s_pressure, s_indices;
    DevBuf export_buf;
    // extrapolation: per grid, cell stamps and two tile flag arrays
    DevBuf ext;
    // advection: copy of the advected grids taken before the step
    DevBuf adv;
    // reinitialisation: level-set copy, stamps, tile flags, list, counters
    DevBuf rei;

    int64_t n_unknowns = 0;
    int nb = 0;
    int np = 0;
    int nblk[elq::kShapeCount] = {};       // pencil blocks of each shape
    size_t blk_base[elq::kShapeCount] = {};  // offset of each shape's block list
    int setup_err = 0;
    int plan_ni = 0, plan_nj = 0, plan_nk = 0;
    int32_t* cur_indices = nullptr;
    std::vector<int32_t> lvl_host;

    template <class F>
    void for_each(F&& f) {
        for (DevBuf* b : {&mask, &indices, &pressure, &scan_cnt, &scan_off, &bflag, &map, &coord, &lvl_cnt, &lvl_off,
                          &lvl_fill, &info, &nbr, &lvl_list, &xch, &stage, &pflag, &pmap, &bkflag, &blist,
                          &diag64, &b64, &inv_e64, &inv_e32,
                          &cdiag, &cax, &cay, &caz, &cscaled, &uflag, &umap, &urow, &vx, &vr, &vp0, &vp1, &vq, &vy, &x64, &part_rr, &part_rz,
                          &ctrl, &nfixed, &s_phi, &s_u, &s_v, &s_w, &s_uw, &s_vw, &s_ww, &s_pressure, &s_indices,
                          &export_buf, &ext, &adv, &rei})
            f(*b);
    }

    size_t bytes() const {
        size_t n = 0;
        const_cast<Workspace*>(this)->for_each([&](DevBuf& b) { n += b.cap; });
        return n;
    }

    void release_all() {
        for_each([](DevBuf& b) { b.release(); });
    }
};

// Copies between pageable host arrays and device memory through a ring of
// pinned chunks: the host moves chunk q with a few threads while the DMA
// engine moves its neighbour, so a copy runs near the pinned PCIe rate rather
// than the runtime's pageable path. The ring is allocated on first use.
struct HostStage {
    static constexpr size_t kChunk = 16u << 20;
    static constexpr int kSlots = 3;
    static constexpr int kThreads = 4;
    uint8_t* pinned = nullptr;  // kSlots * kChunk bytes (hipHostMalloc)
    hipEvent_t ev[kSlots] = {};
    bool recorded[kSlots] = {};
    unsigned next = 0;
    struct Copy {
        void* dst;
        const void* src;
        size_t bytes;
    };
    hipError_t ready();
    // enqueues host -> device copies on s; returns once the host side is done
    hipError_t h2d(void* dev, const void* host, size_t bytes, hipStream_t s);
    // device -> host copies after the work already on s; returns when complete
    hipError_t d2h(const std::vector<Copy>& copies, hipStream_t s);
    void release();
};

namespace elq {
// Serving hook: runs a context's Krylov chunks in launches shared with other
// requests (set by the server on its slot contexts; see elq_server.cpp).
struct KrylovBatcher {
    virtual ~KrylovBatcher() = default;
    // n iterations of the request of slot `slot` (its solver state in job),
    // enqueued after everything on ctx->stream and before anything enqueued
    // there later; returns an elq_status, or kNotBatched when this chunk
    // must run as direct launches.
    virtual int run(elq_ctx* ctx, int slot, const KrylovJob& job, int n) = 0;
    // the request of slot `slot` has left its Krylov loop (no more chunks)
    virtual void leave(int slot) = 0;
    static constexpr int kNotBatched = -1;
};
}  // namespace elq

struct elq_ctx {
    int device = 0;
    size_t budget = 0;
    hipStream_t stream = nullptr;
    void* ctrl_host = nullptr;  // pinned host mirror (kHostMirrorBytes): control block, setup totals
    uint64_t timeout_ticks = 0;
    uint64_t clock_hz = 0;  // device wall clock (wall_clock64) rate
    int max_groups = 0;  // cap on the grid of persistent kernels (0: whole device)
    int rei_rebase = 0;  // reinitialisation tile route: passes between stamp rebases (set at creation)
    // allocation-free projection: this device's static scratch array (its
    // device address, from hipGetSymbolAddress at context creation), the
    // arena over it and the workspace carved from it for each call
    uint8_t* scratch = nullptr;
    Arena arena;
    Workspace ws_scratch;
    // Krylov chunks replayed from a captured HIP graph while their launch
    // arguments (chunk_key) stay the same
    bool graphs = false;
    hipGraphExec_t chunk_exec = nullptr;
    std::vector<uintptr_t> chunk_key;
    void drop_chunk_graph();
    // serving: batched Krylov chunks (mixed precision, wavefront sweeps)
    elq::KrylovBatcher* batcher = nullptr;
    int batch_slot = -1;
    std::string last_error;
    Workspace ws;
    HostStage stage;  // host-buffer entry points

    // per-kernel-class event timing
    bool profiling = false;
    std::string prof_filter;  // ",a,b," : only these classes are timed; empty: all
    bool prof_open = false;   // the current prof_begin recorded an event
    std::vector<hipEvent_t> events;
    size_t ev_used = 0;
    std::vector<const char*> ev_names;
    std::map<std::string, std::pair<double, int64_t>> kt;
    void prof_begin(const char* name);
    void prof_end();
    void prof_collect();

    int fail(int code, const std::string& msg);
    int hip_fail(hipError_t e, const char* what);
};

// Device bytes a projection requests on a fresh context (nb_bound: bricks
// holding unknowns, < 0 for all bricks; staged: host inputs are copied in).
size_t elq_footprint_bytes(int ni, int nj, int nk, int precision, bool staged, int64_t nb_bound);
// Bricks of the grid holding at least one cell with phi < 0 (a bound on the
// bricks that hold unknowns).
int64_t liquid_bricks(const float* phi, int ni, int nj, int nk);
