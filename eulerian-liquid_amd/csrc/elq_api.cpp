// Host orchestration of the pressure projection and the C ABI of include/elq.h.
//
// One call of elq_project_device reproduces Simulator::solve_pressure
// (simulator.cpp:96-112) on the GPU:
//   1. mask + lexicographic scan        -> indices   (projection.cpp:14-28)
//   2. brick compaction + level lists
//   3. assembly                         -> A, b      (projection.cpp:30-83)
//   4. IC(0)/MIC(0) factor (wavefront)
//   5. PCG with Eigen's ConjugateGradient semantics (simulator.cpp:102-109),
//      FP64, or FP32 inner solves under FP64 iterative refinement
//   6. scatter + velocity update                    (projection.cpp:85-134)
// All arithmetic runs in the kernels of elq_kernels.hip; the host only sizes
// buffers, launches, and polls the device-side control block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <thread>
#include <vector>

#include "../../include/elq.h"
#include "elq_api_internal.hpp"
#include "elq_frame.hpp"
#include "elq_kernels.hpp"

using namespace elq;

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// ============================================================ static scratch
// Workspace of the allocation-free projection (elq_project_device_noalloc):
// every device buffer of such a call is carved from this array. Each device
// holds its own copy (ELQ_NOALLOC_SCRATCH_BYTES) from the library's first
// use on it until the process ends. In host code `scratch` names only a
// host-side stand-in: the device address comes from hipGetSymbolAddress,
// looked up once per device when a context is created on it.
static_assert(ELQ_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size");
static __device__ __attribute__((aligned(256))) uint8_t scratch[ELQ_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_addr[kMaxDevices] = {};
hipError_t g_scratch_rc[kMaxDevices] = {};
bool g_scratch_looked_up[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be current.
hipError_t scratch_address(int device, uint8_t** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_looked_up[device]) {
        void* p = nullptr;
        g_scratch_rc[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_addr[device] = g_scratch_rc[device] == hipSuccess ? (uint8_t*)p : nullptr;
        g_scratch_looked_up[device] = true;
    }
    *out = g_scratch_addr[device];
    return g_scratch_rc[device];
}

}  // namespace

// ============================================================ context basics
hipError_t DevBuf::ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (arena) {
        const size_t need = (bytes + 255) / 256 * 256;
        if (need > arena->size - arena->off) {
            arena->overflow = true;
            return hipErrorOutOfMemory;
        }
        p = arena->base + arena->off;
        arena->off += need;
        cap = bytes;
        return hipSuccess;
    }
    if (p) (void)hipFreeAsync(p, stream);
    p = nullptr;
    cap = 0;
    if (bytes == 0) return hipSuccess;
    hipError_t e = hipMallocAsync(&p, bytes, stream);
    if (e == hipSuccess) cap = bytes;
    return e;
}

void DevBuf::release() {
    if (arena) {
        p = nullptr;
        cap = 0;
        return;
    }
    if (p) (void)hipFreeAsync(p, stream);
    p = nullptr;
    cap = 0;
}

void elq_ctx::drop_chunk_graph() {
    if (chunk_exec) (void)hipGraphExecDestroy(chunk_exec);
    chunk_exec = nullptr;
    chunk_key.clear();
}

int elq_ctx::fail(int code, const std::string& msg) {
    last_error = msg;
    return code;
}

int elq_ctx::hip_fail(hipError_t e, const char* what) {
    last_error = std::string(what) + ": " + hipGetErrorString(e);
    return ELQ_ERR_DEVICE;
}

// Splits [0, n) over HostStage::kThreads threads (the caller takes one part).
static void copy_parallel(void* dst, const void* src, size_t n) {
    constexpr int T = HostStage::kThreads;
    if (n < (1u << 20)) {
        std::memcpy(dst, src, n);
        return;
    }
    const size_t per = (n / T + 4095) & ~(size_t)4095;
    std::thread th[T - 1];
    for (int t = 1; t < T && (size_t)t * per < n; ++t) {
        const size_t a = (size_t)t * per, b = std::min(n, a + per);
        auto part = [=] { std::memcpy((uint8_t*)dst + a, (const uint8_t*)src + a, b - a); };
        try {
            th[t - 1] = std::thread(part);
        } catch (...) {  // no thread to be had: this part on the calling thread (nothing throws past the C ABI)
            part();
        }
    }
    std::memcpy(dst, src, std::min(n, per));
    for (auto& x : th)
        if (x.joinable()) x.join();
}

hipError_t HostStage::ready() {
    if (pinned) return hipSuccess;
    hipError_t e = hipHostMalloc((void**)&pinned, kSlots * kChunk, hipHostMallocDefault);
    if (e != hipSuccess) {
        pinned = nullptr;
        return e;
    }
    for (int q = 0; q < kSlots; ++q) {
        if ((e = hipEventCreateWithFlags(&ev[q], hipEventDisableTiming)) != hipSuccess) return e;
        recorded[q] = false;
    }
    return hipSuccess;
}

hipError_t HostStage::h2d(void* dev, const void* host, size_t bytes, hipStream_t s) {
    if (bytes < (1u << 20)) return hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s);
    hipError_t e = ready();
    if (e != hipSuccess) return e;
    for (size_t o = 0; o < bytes; o += kChunk) {
        const int q = (int)(next++ % kSlots);
        const size_t n = std::min(kChunk, bytes - o);
        // the chunk's previous copy out of this slot must have completed
        if (recorded[q] && (e = hipEventSynchronize(ev[q])) != hipSuccess) return e;
        copy_parallel(pinned + q * kChunk, (const uint8_t*)host + o, n);
        if ((e = hipMemcpyAsync((uint8_t*)dev + o, pinned + q * kChunk, n, hipMemcpyHostToDevice, s)) != hipSuccess)
            return e;
        if ((e = hipEventRecord(ev[q], s)) != hipSuccess) return e;
        recorded[q] = true;
    }
    return hipSuccess;
}

hipError_t HostStage::d2h(const std::vector<Copy>& copies, hipStream_t s) {
    hipError_t e = ready();
    if (e != hipSuccess) return e;
    struct Piece {
        uint8_t* dst;
        const uint8_t* src;
        size_t n;
    };
    std::vector<Piece> pcs;
    for (const Copy& c : copies)
        for (size_t o = 0; o < c.bytes; o += kChunk)
            pcs.push_back({(uint8_t*)c.dst + o, (const uint8_t*)c.src + o, std::min(kChunk, c.bytes - o)});
    // pieces go through slot (j % kSlots); work already queued on s (and any
    // earlier copy through a slot) precedes them in stream order
    auto issue = [&](size_t j) -> hipError_t {
        const int q = (int)(j % kSlots);
        hipError_t r = hipMemcpyAsync(pinned + q * kChunk, pcs[j].src, pcs[j].n, hipMemcpyDeviceToHost, s);
        if (r == hipSuccess) r = hipEventRecord(ev[q], s);
        recorded[q] = r == hipSuccess;
        return r;
    };
    for (size_t j = 0; j < pcs.size() && j < (size_t)kSlots; ++j)
        if ((e = issue(j)) != hipSuccess) return e;
    for (size_t j = 0; j < pcs.size(); ++j) {
        const int q = (int)(j % kSlots);
        if ((e = hipEventSynchronize(ev[q])) != hipSuccess) return e;
        copy_parallel(pcs[j].dst, pinned + q * kChunk, pcs[j].n);
        if (j + kSlots < pcs.size() && (e = issue(j + kSlots)) != hipSuccess) return e;
    }
    return hipStreamSynchronize(s);
}

void HostStage::release() {
    if (!pinned) return;
    for (int q = 0; q < kSlots; ++q)
        if (ev[q]) (void)hipEventDestroy(ev[q]);
    (void)hipHostFree(pinned);
    pinned = nullptr;
}

void elq_ctx::prof_begin(const char* name) {
    prof_open = false;
    if (!profiling) return;
    if (!prof_filter.empty() && prof_filter.find("," + std::string(name) + ",") == std::string::npos) return;
    if (ev_used + 2 > events.size()) {
        for (int q = 0; q < 256; ++q) {
            hipEvent_t ev;
            if (hipEventCreate(&ev) != hipSuccess) return;
            events.push_back(ev);
        }
    }
    (void)hipEventRecord(events[ev_used], stream);
    ev_names.push_back(name);
    ++ev_used;
    prof_open = true;
}

void elq_ctx::prof_end() {
    if (!profiling || !prof_open) return;
    (void)hipEventRecord(events[ev_used], stream);
    ++ev_used;
    prof_open = false;
}

void elq_ctx::prof_collect() {
    if (!profiling) return;
    (void)hipStreamSynchronize(stream);
    for (size_t q = 0; q + 1 < ev_used; q += 2) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, events[q], events[q + 1]);
        auto& e = kt[ev_names[q / 2]];
        e.first += ms;
        e.second += 1;
    }
    ev_used = 0;
    ev_names.clear();
}

// ============================================================ projection core
namespace {

struct Plan {
    int ni, nj, nk;
    int64_t C;
    int nbx, nby, nbz, nlev;
    int64_t NB;
};

Plan make_plan(int ni, int nj, int nk) {
    Plan p;
    p.ni = ni;
    p.nj = nj;
    p.nk = nk;
    p.C = (int64_t)ni * nj * nk;
    p.nbx = (ni + kBrick - 1) / kBrick;
    p.nby = (nj + kBrick - 1) / kBrick;
    p.nbz = (nk + kBrick - 1) / kBrick;
    p.nlev = p.nbx + p.nby + p.nbz - 2;
    p.NB = (int64_t)p.nbx * p.nby * p.nbz;
    return p;
}

// Alignment padding of the scratch carve (at most 255 bytes per buffer).
constexpr size_t kArenaSlack = 64 * 1024;

// Pencil blocks of the wavefront sweeps: the block lists of every shape.
size_t block_count(const Plan& p) {
    size_t n = 0;
    for (int k = 0; k < kShapeCount; ++k) n += shape_blocks(p.nby, p.nbz, k);
    return n;
}

bool valid_dims(int ni, int nj, int nk) {
    return ni >= 1 && nj >= 1 && nk >= 1 && ni <= 4092 && nj <= 4092 && nk <= 4092 &&
           (int64_t)(ni + 1) * (nj + 1) * (nk + 1) < (int64_t)INT32_MAX;
}

#define HIPCHK(expr, what)                                  \
    do {                                                    \
        hipError_t _e = (expr);                             \
        if (_e != hipSuccess) return ctx->hip_fail(_e, what); \
    } while (0)

}  // namespace

size_t elq_footprint_bytes(int ni, int nj, int nk, int precision, bool staged, int64_t nb_bound) {
    // Mirrors the allocations of a projection on a fresh context (bytes
    // requested from the allocator). nb_bound bounds the bricks that hold
    // unknowns (< 0: every brick of the grid).
    const Plan p = make_plan(ni, nj, nk);
    const int64_t nb = nb_bound < 0 ? p.NB : std::min<int64_t>(nb_bound, p.NB);
    const size_t cells = (size_t)nb * kBrickCells;
    const size_t tsz = precision == ELQ_PRECISION_MIXED ? 4 : 8;
    const size_t C = (size_t)p.C, NB = (size_t)p.NB, NP = (size_t)p.nby * p.nbz;
    const size_t G = (size_t)scan_blocks(std::max<int64_t>(p.C, p.NB));
    size_t b = 0;
    b += C * (1 + 4 + 4);                                        // mask, indices, pressure
    b += NB * (1 + 4 + 4) + 2 * G * 4 + 3 * ((size_t)p.nlev + 1) * 4 + 64;  // bricks, scans, levels, info
    b += NP * (1 + 4);                                           // pencil flags, map
    b += block_count(p) * 5;                                     // pencil blocks
    b += (size_t)nb * (8 * 4 + 4 + (size_t)kXchWords * 8 + 2 * 8);  // nbr, lvl_list, xch, partials
    b += (size_t)std::max(kRedBlocks * 4, kTailMaxGroups * 4) * 8 + sizeof(Ctrl) + 4;  // stage, ctrl, nfixed
    b += cells * (3 * 8 + 10 * tsz + 4);                         // diag64, b64, inv_e64, coefficients, vectors, inv_e32
    if (precision == ELQ_PRECISION_FP64) b += cells * 3 * 8;     // scaled lower couplings
    b += (size_t)std::max<int64_t>(nb, 1) * 4 + NB * 4 + 4 * kBrickCells * tsz;  // uniform-brick flags, constant rows
    if (precision == ELQ_PRECISION_MIXED) b += cells * 8;        // x64
    const size_t faces = (size_t)(ni + 1) * nj * nk + (size_t)ni * (nj + 1) * nk + (size_t)ni * nj * (nk + 1);
    if (staged) b += 4 * (C + 2 * faces);
    return b;
}

int64_t liquid_bricks(const float* phi, int ni, int nj, int nk) {
    const Plan p = make_plan(ni, nj, nk);
    std::vector<uint8_t> hit((size_t)p.NB, 0);
    int64_t n = 0;
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j) {
            const float* row = phi + (size_t)ni * ((size_t)j + (size_t)nj * k);
            const size_t base = (size_t)p.nbx * ((size_t)(j / kBrick) + (size_t)p.nby * (k / kBrick));
            for (int i = 0; i < ni; ++i)
                if (row[i] < 0.0f) {
                    uint8_t& h = hit[base + (size_t)(i / kBrick)];
                    n += h == 0;
                    h = 1;
                }
        }
    return n;
}

namespace {

// Phase A: mask, indices, brick compaction. Leaves N, nb, err in ctx.
int setup_phase_a(elq_ctx* ctx, const Dense& d, int32_t* indices_out) {
    const Plan p = make_plan(d.ni, d.nj, d.nk);
    Workspace& W = ctx->ws;
    const int64_t G = scan_blocks(std::max(p.C, p.NB));
    HIPCHK(W.mask.ensure((size_t)p.C), "alloc mask");
    HIPCHK(W.bflag.ensure((size_t)p.NB), "alloc bflag");
    HIPCHK(W.map.ensure((size_t)p.NB * 4), "alloc map");
    HIPCHK(W.coord.ensure((size_t)p.NB * 4), "alloc coord");
    HIPCHK(W.scan_cnt.ensure((size_t)G * 4), "alloc scan");
    HIPCHK(W.scan_off.ensure((size_t)G * 4), "alloc scan");
    HIPCHK(W.lvl_cnt.ensure((size_t)(p.nlev + 1) * 4), "alloc lvl");
    HIPCHK(W.lvl_off.ensure((size_t)(p.nlev + 1) * 4), "alloc lvl");
    HIPCHK(W.lvl_fill.ensure((size_t)(p.nlev + 1) * 4), "alloc lvl");
    const int64_t NP = (int64_t)p.nby * p.nbz;
    HIPCHK(W.pflag.ensure((size_t)NP), "alloc pencils");
    HIPCHK(W.pmap.ensure((size_t)NP * 4), "alloc pencils");
    const size_t NBLK = block_count(p);
    HIPCHK(W.bkflag.ensure(NBLK), "alloc blocks");
    HIPCHK(W.blist.ensure(NBLK * 4), "alloc blocks");
    HIPCHK(W.info.ensure(sizeof(int64_t) * (3 + kShapeCount) + sizeof(int32_t) * 2), "alloc info");
    int32_t* indices = indices_out;
    if (!indices) {
        HIPCHK(W.indices.ensure((size_t)p.C * 4), "alloc indices");
        indices = (int32_t*)W.indices.p;
    }
    W.cur_indices = indices;
    hipStream_t s = ctx->stream;
    int64_t* totals = (int64_t*)W.info.p;
    int32_t* err = (int32_t*)(totals + 3 + kShapeCount);
    HIPCHK(hipMemsetAsync(W.info.p, 0, W.info.cap, s), "memset");
    HIPCHK(hipMemsetAsync(W.lvl_cnt.p, 0, (size_t)(p.nlev + 1) * 4, s), "memset");
    ctx->prof_begin("setup_mask_scan");
    HIPCHK(launch_mask(d, (uint8_t*)W.mask.p, err, s), "k_mask");
    HIPCHK(launch_scan((const uint8_t*)W.mask.p, p.C, kFluid, indices, ELQ_UNKNOWN_INDEX, (int32_t*)W.scan_cnt.p,
                       (int32_t*)W.scan_off.p, &totals[0], s),
           "scan indices");
    HIPCHK(launch_brick_flags(d, (const uint8_t*)W.mask.p, p.nbx, p.nby, p.nbz, (uint8_t*)W.bflag.p, s),
           "k_brick_flags");
    HIPCHK(launch_scan((const uint8_t*)W.bflag.p, p.NB, 1, (int32_t*)W.map.p, -1, (int32_t*)W.scan_cnt.p,
                       (int32_t*)W.scan_off.p, &totals[1], s),
           "scan bricks");
    HIPCHK(launch_brick_meta((const int32_t*)W.map.p, p.nbx, p.nby, p.nbz, (int32_t*)W.coord.p,
                             (int32_t*)W.lvl_cnt.p, s),
           "k_brick_meta");
    HIPCHK(launch_level_scan((const int32_t*)W.lvl_cnt.p, p.nlev, (int32_t*)W.lvl_off.p, (int32_t*)W.lvl_fill.p, s),
           "k_level_scan");
    // x-pencils of the wavefront sweeps
    HIPCHK(launch_pencil_flags((const int32_t*)W.map.p, p.nbx, p.nby, p.nbz, (uint8_t*)W.pflag.p, s), "k_pencil_flags");
    HIPCHK(launch_scan((const uint8_t*)W.pflag.p, NP, 1, (int32_t*)W.pmap.p, -1, (int32_t*)W.scan_cnt.p,
                       (int32_t*)W.scan_off.p, &totals[2], s),
           "scan pencils");
    HIPCHK(launch_block_lists((const int32_t*)W.pmap.p, p.nby, p.nbz, (uint8_t*)W.bkflag.p, (int32_t*)W.blist.p,
                              &totals[3], s),
           "k_block_list");
    ctx->prof_end();
    // totals and the error word are adjacent in info: one export, one sync
    int64_t host_tot[3 + kShapeCount];
    int32_t host_err = 0;
    const size_t info_bytes = sizeof(int64_t) * (3 + kShapeCount) + sizeof(int32_t) * 2;
    static_assert(sizeof(int64_t) * (3 + kShapeCount) + sizeof(int32_t) * 2 <= kHostMirrorBytes, "info mirror");
    HIPCHK(launch_export_words(totals, info_bytes, ctx->ctrl_host, s), "k_export_words");
    HIPCHK(hipStreamSynchronize(s), "sync");
    std::memcpy(host_tot, ctx->ctrl_host, sizeof(host_tot));
    std::memcpy(&host_err, (const char*)ctx->ctrl_host + sizeof(host_tot), sizeof(host_err));
    W.n_unknowns = host_tot[0];
    W.nb = (int)host_tot[1];
    W.np = (int)host_tot[2];
    size_t base = 0;
    for (int k = 0; k < kShapeCount; ++k) {
        W.nblk[k] = (int)host_tot[3 + k];
        W.blk_base[k] = base;
        base += shape_blocks(p.nby, p.nbz, k);
    }
    W.setup_err = host_err;
    W.plan_ni = d.ni;
    W.plan_nj = d.nj;
    W.plan_nk = d.nk;
    return ELQ_OK;
}

Bricks make_bricks(elq_ctx* ctx, const Plan& p) {
    Workspace& W = ctx->ws;
    Bricks bk;
    bk.nbx = p.nbx;
    bk.nby = p.nby;
    bk.nbz = p.nbz;
    bk.nb = W.nb;
    bk.nlev = p.nlev;
    bk.map = (const int32_t*)W.map.p;
    bk.coord = (const int32_t*)W.coord.p;
    bk.nbr = (const int32_t*)W.nbr.p;
    bk.lvl_list = (const int32_t*)W.lvl_list.p;
    return bk;
}

Pencils pencils_of(elq_ctx* ctx, int shape) {
    Workspace& W = ctx->ws;
    Pencils pc;
    pc.np = W.np;
    pc.shape = shape;
    pc.nblk = W.nblk[shape];
    pc.blk_list = (const int32_t*)W.blist.p + W.blk_base[shape];
    return pc;
}

// Pencil-block shape of a sweep. FP64: 2x2 at every size (against 2x4 blocks
// for either sweep, from 64^3 to 512^3: the fastest or within 1%, except 2%
// slower at 256^3; 2x4 took 3.5% longer at 384^3, 8% at 512^3). FP32 backward
// sweep: 1x4 while its blocks fit in one round of four workgroups per CU,
// else 2x2. The FP32 forward sweep takes 1x4 at every size (fwd_shape): only
// its 1x4 kernel reads uniform bricks from the constant rows, which over
// several rounds (512^3) makes it faster than 2x2, while the backward sweep
// stays faster with 2x2 there. The choice depends only on the active
// pencils, the precision, the sweep and the context's workgroup cap.
Pencils make_pencils(elq_ctx* ctx, bool dbl) {
    if (dbl) return pencils_of(ctx, kShape2x2);
    return pencils_of(ctx, ctx->ws.nblk[kShape1x4] <= sweep_round_groups(ctx->max_groups) ? kShape1x4 : kShape2x2);
}

// Forward-sweep blocks given the backward sweep's (bwd): FP32 1x4, except in
// the batched lockstep mode, whose kernels run both sweeps on one shape.
Pencils fwd_shape(elq_ctx* ctx, bool dbl, const Pencils& bwd) {
    return dbl || ctx->batcher ? bwd : pencils_of(ctx, kShape1x4);
}

// Phase B: lists, assembly (T storage). FP64 arithmetic for diag/b always.
template <class T>
int setup_phase_b(elq_ctx* ctx, const Dense& d, float dt, bool need_levels) {
    const Plan p = make_plan(d.ni, d.nj, d.nk);
    Workspace& W = ctx->ws;
    const int nb = W.nb;
    const size_t cells = (size_t)nb * kBrickCells;
    hipStream_t s = ctx->stream;
    HIPCHK(W.nbr.ensure((size_t)nb * 8 * 4), "alloc nbr");
    HIPCHK(W.lvl_list.ensure((size_t)nb * 4), "alloc lvl_list");
    // tagged exchange words: zeroed on (re)allocation; tags (epochs) never 0
    if ((size_t)nb * kXchWords * 8 > W.xch.cap) {
        HIPCHK(W.xch.ensure((size_t)nb * kXchWords * 8), "alloc xch");
        HIPCHK(hipMemsetAsync(W.xch.p, 0, W.xch.cap, s), "memset xch");
    }
    if (!W.stage.p) {
        HIPCHK(W.stage.ensure((size_t)std::max(kRedBlocks * 4, kTailMaxGroups * 4) * 8), "alloc stage");
        HIPCHK(hipMemsetAsync(W.stage.p, 0, W.stage.cap, s), "memset stage");
    }
    HIPCHK(W.diag64.ensure(cells * 8), "alloc");
    HIPCHK(W.b64.ensure(cells * 8), "alloc");
    HIPCHK(W.inv_e64.ensure(cells * 8), "alloc");
    for (DevBuf* b : {&W.cdiag, &W.cax, &W.cay, &W.caz, &W.vx, &W.vr, &W.vp0, &W.vp1, &W.vq, &W.vy})
        HIPCHK(b->ensure(cells * sizeof(T)), "alloc");
    HIPCHK(W.inv_e32.ensure(cells * 4), "alloc");
    if (sizeof(T) == 8) HIPCHK(W.cscaled.ensure(3 * cells * 8), "alloc");
    HIPCHK(W.uflag.ensure((size_t)std::max(nb, 1) * 4), "alloc");
    HIPCHK(W.umap.ensure((size_t)p.NB * 4), "alloc");
    HIPCHK(W.urow.ensure(4 * kBrickCells * sizeof(T)), "alloc");
    HIPCHK(W.part_rr.ensure((size_t)nb * 8), "alloc");
    HIPCHK(W.part_rz.ensure((size_t)nb * 8), "alloc");
    if (!W.ctrl.p) {
        HIPCHK(W.ctrl.ensure(sizeof(Ctrl)), "alloc");
        HIPCHK(hipMemsetAsync(W.ctrl.p, 0, sizeof(Ctrl), s), "memset ctrl");
    }
    HIPCHK(W.nfixed.ensure(sizeof(int32_t)), "alloc");
    Bricks bk = make_bricks(ctx, p);
    ctx->prof_begin("setup_lists_assemble");
    HIPCHK(launch_brick_lists(bk.map, bk.coord, p.nbx, p.nby, p.nbz, nb, (const int32_t*)W.lvl_off.p,
                              (int32_t*)W.lvl_fill.p, (int32_t*)W.lvl_list.p, (int32_t*)W.nbr.p, s),
           "k_brick_lists");
    HIPCHK(launch_assemble<T>(d, dt, (const uint8_t*)W.mask.p, bk, (double*)W.diag64.p, (double*)W.b64.p,
                              (T*)W.cdiag.p, (T*)W.cax.p, (T*)W.cay.p, (T*)W.caz.p, s),
           "k_assemble");
    ctx->prof_end();
    if (need_levels) {
        W.lvl_host.resize((size_t)p.nlev + 1);
        HIPCHK(hipMemcpyAsync(W.lvl_host.data(), W.lvl_off.p, (size_t)(p.nlev + 1) * 4, hipMemcpyDeviceToHost, s),
               "d2h");
        HIPCHK(hipStreamSynchronize(s), "sync");
    }
    return ELQ_OK;
}

void ctrl_host_init(Ctrl& c, const elq_opts& o, int mixed) {
    std::memset(&c, 0, sizeof(c));
    c.tol2 = o.tol * o.tol;
    c.inner_tol2 = o.inner_tol * o.inner_tol;
    c.max_iter = o.max_iter;
    c.mixed = mixed;
}

int upload_ctrl(elq_ctx* ctx, const Ctrl& c) {
    // the epoch/ticket words persist on the device across solves
    HIPCHK(hipMemcpyAsync(ctx->ws.ctrl.p, &c, offsetof(Ctrl, epoch), hipMemcpyHostToDevice, ctx->stream), "h2d");
    HIPCHK(hipMemsetAsync((char*)ctx->ws.ctrl.p + offsetof(Ctrl, ticket), 0, 2 * sizeof(uint32_t), ctx->stream),
           "memset");
    return ELQ_OK;
}

// The control block reaches the host through a kernel that stores it into
// pinned host memory: a poll then costs one small launch and the stream
// synchronisation, without a copy-engine transfer.
int read_ctrl(elq_ctx* ctx, Ctrl& c) {
    static_assert(sizeof(Ctrl) % 4 == 0 && sizeof(Ctrl) <= kHostMirrorBytes, "control block mirror");
    HIPCHK(launch_export_words(ctx->ws.ctrl.p, sizeof(Ctrl), ctx->ctrl_host, ctx->stream), "k_export_words");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    std::memcpy(&c, ctx->ctrl_host, sizeof(Ctrl));
    return ELQ_OK;
}

template <class T>
struct Solver {
    elq_ctx* ctx;
    Plan p;
    Bricks bk;
    Pencils pc;   // forward sweep blocks (both sweeps when fused or batched)
    Pencils pcb;  // backward sweep blocks
    Coef<T> c;
    const elq_opts* o;
    bool levels;
    bool fused = false;  // forward and backward sweep of an iteration in one launch (launch_sweeps)

    Ctrl* ctrl() { return (Ctrl*)ctx->ws.ctrl.p; }
    double* prr() { return (double*)ctx->ws.part_rr.p; }
    double* prz() { return (double*)ctx->ws.part_rz.p; }
    uint64_t* stage() { return (uint64_t*)ctx->ws.stage.p; }
    T* vec(DevBuf& b) { return (T*)b.p; }

    int scalar(int mode, const double* parts, const double* parts2 = nullptr) {
        ctx->prof_begin("scalar");
        HIPCHK(launch_scalar(ctrl(), mode, parts, bk.nb, parts2, (uint64_t*)ctx->ws.stage.p, ctx->timeout_ticks,
                             ctx->stream),
               "k_scalar");
        ctx->prof_end();
        return ELQ_OK;
    }

    int fwd(int update, T* p_cur) {
        Workspace& W = ctx->ws;
        ctx->prof_begin(update ? "fwd_sweep_update" : "fwd_sweep");
        if (!levels) {
            HIPCHK(launch_fwd<T>(bk, pc, c, vec(W.vr), vec(W.vx), p_cur, vec(W.vq), vec(W.vy), prr(), ctrl(),
                                 (uint64_t*)W.xch.p, stage(), update, -1, 0, 0, ctx->timeout_ticks,
                                 ctx->max_groups, ctx->stream),
                   "k_fwd");
        } else {
            for (int l = 0; l < p.nlev; ++l) {
                const int base = W.lvl_host[l], cnt = W.lvl_host[l + 1] - base;
                HIPCHK(launch_fwd<T>(bk, pc, c, vec(W.vr), vec(W.vx), p_cur, vec(W.vq), vec(W.vy), prr(), ctrl(),
                                     (uint64_t*)W.xch.p, stage(), update, l, base, cnt, ctx->timeout_ticks,
                                     ctx->max_groups, ctx->stream),
                       "k_fwd");
            }
        }
        ctx->prof_end();
        // the level launches leave |r|^2 as per-brick partials
        return levels ? scalar(update ? SC_RR : SC_NOP, update ? prr() : nullptr) : ELQ_OK;
    }

    // Backward sweep; the r.z reduction then applies mode (SC_BETA,
    // SC_RZ_INIT or SC_RZ_REPLACE).
    int bwd(int mode) {
        Workspace& W = ctx->ws;
        ctx->prof_begin("bwd_sweep");
        if (!levels) {
            HIPCHK(launch_bwd<T>(bk, pcb, c, vec(W.vr), vec(W.vy), prz(), ctrl(), (uint64_t*)W.xch.p, stage(),
                                 mode, -1, 0, 0, ctx->timeout_ticks, ctx->max_groups, ctx->stream),
                   "k_bwd");
        } else {
            for (int l = p.nlev - 1; l >= 0; --l) {
                const int base = W.lvl_host[l], cnt = W.lvl_host[l + 1] - base;
                HIPCHK(launch_bwd<T>(bk, pcb, c, vec(W.vr), vec(W.vy), prz(), ctrl(), (uint64_t*)W.xch.p, stage(),
                                     mode, l, base, cnt, ctx->timeout_ticks, ctx->max_groups, ctx->stream),
                       "k_bwd");
            }
        }
        ctx->prof_end();
        return levels ? scalar(mode, prz()) : ELQ_OK;
    }

    int factor(double tau, double sigma, float* inv_e32) {
        Workspace& W = ctx->ws;
        int rc = scalar(SC_NOP, nullptr);
        if (rc) return rc;
        HIPCHK(hipMemsetAsync(W.nfixed.p, 0, sizeof(int32_t), ctx->stream), "memset");
        ctx->prof_begin("factor");
        if (!levels) {
            HIPCHK(launch_factor<T>(bk, (const double*)W.diag64.p, c.ax, c.ay, c.az, tau, sigma, (double*)W.inv_e64.p,
                                    inv_e32, ctrl(), (uint64_t*)W.xch.p, (int32_t*)W.nfixed.p, -1, 0, 0,
                                    ctx->timeout_ticks, ctx->max_groups, ctx->stream),
                   "k_factor");
        } else {
            for (int l = 0; l < p.nlev; ++l) {
                const int base = W.lvl_host[l], cnt = W.lvl_host[l + 1] - base;
                HIPCHK(launch_factor<T>(bk, (const double*)W.diag64.p, c.ax, c.ay, c.az, tau, sigma,
                                        (double*)W.inv_e64.p, inv_e32, ctrl(), (uint64_t*)W.xch.p,
                                        (int32_t*)W.nfixed.p, l, base, cnt, ctx->timeout_ticks, ctx->max_groups,
                                        ctx->stream),
                       "k_factor");
            }
        }
        ctx->prof_end();
        return ELQ_OK;
    }

    int64_t kiter = 0;  // iterations launched (selects the p ping-pong buffer)

    // z = M^{-1} r and rz; mode SC_RZ_INIT (beta = 0) or SC_RZ_REPLACE.
    int precondition(int rz_mode) {
        int rc;
        if ((rc = fwd(0, nullptr))) return rc;
        return bwd(rz_mode);
    }

    // n PCG iterations (Eigen's loop body), no host synchronisation. With
    // graph capture on (wavefront schedule, no event timing), the chunk of
    // 3n launches is captured once into a HIP graph and replayed while every
    // argument of those launches stays the same (the key below); the graph
    // holds kernel nodes only.
    int iterations(int n) {
        if constexpr (std::is_same<T, float>::value) {
            if (ctx->batcher && !levels && !ctx->profiling &&
                (pc.shape == kShape1x4 || pc.shape == kShape2x2)) {
                Workspace& W = ctx->ws;
                KrylovJob j{};
                j.bk = bk;
                j.pc = pc;
                j.c = c;
                j.r = vec(W.vr);
                j.x = vec(W.vx);
                j.p0 = vec(W.vp0);
                j.p1 = vec(W.vp1);
                j.q = vec(W.vq);
                j.y = vec(W.vy);
                j.ctrl = ctrl();
                j.xch = (uint64_t*)W.xch.p;
                j.stage = stage();
                j.timeout = ctx->timeout_ticks;
                krylov_groups(bk, pc, ctx->max_groups, &j.g_spmv, &j.g_fwd, &j.g_bwd);
                j.par = (int32_t)(kiter & 1);
                const int rc = ctx->batcher->run(ctx, ctx->batch_slot, j, n);
                if (rc != KrylovBatcher::kNotBatched) {
                    if (rc == ELQ_OK) kiter += n;
                    return rc;
                }
            }
        }
        if (levels || !ctx->graphs || ctx->profiling) return iterations_direct(n);
        Workspace& W = ctx->ws;
        const std::vector<uintptr_t> key = {
            (uintptr_t)n, (uintptr_t)(kiter & 1), sizeof(T), (uintptr_t)ctx->max_groups, (uintptr_t)ctx->timeout_ticks,
            (uintptr_t)bk.nbx, (uintptr_t)bk.nby, (uintptr_t)bk.nbz, (uintptr_t)bk.nb, (uintptr_t)bk.nlev,
            (uintptr_t)bk.map, (uintptr_t)bk.coord, (uintptr_t)bk.nbr, (uintptr_t)bk.lvl_list, (uintptr_t)pc.np,
            (uintptr_t)pc.shape, (uintptr_t)pc.nblk, (uintptr_t)pc.blk_list, (uintptr_t)pcb.shape,
            (uintptr_t)pcb.nblk, (uintptr_t)pcb.blk_list, (uintptr_t)c.diag, (uintptr_t)c.ax,
            (uintptr_t)c.ay, (uintptr_t)c.az, (uintptr_t)c.inv_e, (uintptr_t)c.lx, (uintptr_t)c.uflag,
            (uintptr_t)c.umap, (uintptr_t)c.urow, (uintptr_t)W.vy.p, (uintptr_t)W.vp0.p,
            (uintptr_t)W.vp1.p, (uintptr_t)W.vq.p, (uintptr_t)W.vr.p, (uintptr_t)W.vx.p, (uintptr_t)prr(),
            (uintptr_t)prz(), (uintptr_t)ctrl(), (uintptr_t)W.xch.p, (uintptr_t)stage(), (uintptr_t)fused};
        if (!ctx->chunk_exec || key != ctx->chunk_key) {
            ctx->drop_chunk_graph();
            HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
            const int64_t k0 = kiter;
            const int rc = iterations_direct(n);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            kiter = k0;  // the captured launches have not run
            if (rc != ELQ_OK || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return rc != ELQ_OK ? rc : ctx->hip_fail(e, "hipStreamEndCapture");
            }
            const hipError_t ei = hipGraphInstantiate(&ctx->chunk_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                ctx->chunk_exec = nullptr;
                return ctx->hip_fail(ei, "hipGraphInstantiate");
            }
            ctx->chunk_key = key;
        }
        HIPCHK(hipGraphLaunch(ctx->chunk_exec, ctx->stream), "hipGraphLaunch");
        kiter += n;
        return ELQ_OK;
    }

    int iterations_direct(int n) {
        Workspace& W = ctx->ws;
        int rc;
        T* P[2] = {vec(W.vp0), vec(W.vp1)};
        for (int q = 0; q < n; ++q, ++kiter) {
            ctx->prof_begin("spmv");
            HIPCHK(launch_spmv<T>(bk, c, vec(W.vy), P[kiter & 1], P[(kiter + 1) & 1], vec(W.vq), ctrl(), stage(),
                                  ctx->timeout_ticks, ctx->max_groups, ctx->stream),
                   "k_spmv");
            ctx->prof_end();
            if constexpr (std::is_same<T, float>::value) {
                if (fused) {
                    ctx->prof_begin("sweeps");
                    HIPCHK(launch_sweeps(bk, pc, c, vec(W.vr), vec(W.vx), P[(kiter + 1) & 1], vec(W.vq), vec(W.vy),
                                         ctrl(), (uint64_t*)W.xch.p, stage(), ctx->timeout_ticks, ctx->max_groups,
                                         ctx->stream),
                           "k_sweeps");
                    ctx->prof_end();
                    continue;
                }
            }
            if ((rc = fwd(1, P[(kiter + 1) & 1]))) return rc;
            if ((rc = bwd(SC_BETA))) return rc;
        }
        return ELQ_OK;
    }

    int poll(Ctrl& hc) {
        int rc;
        ctx->prof_collect();
        if ((rc = read_ctrl(ctx, hc))) return rc;
        if (hc.abort) return ctx->fail(ELQ_ERR_DEVICE, "wavefront wait exceeded its time limit");
        return ELQ_OK;
    }

    // FP64 Krylov loop with Eigen semantics, polling every check_every iterations.
    int krylov(Ctrl& hc) {
        int rc;
        if ((rc = precondition(SC_RZ_INIT))) return rc;
        const int check = std::max(1, o->check_every);
        const int64_t cap = (int64_t)std::max(o->max_iter, 0) / check + 3;
        for (int64_t chunk = 0;; ++chunk) {
            if (chunk > cap) return ctx->fail(ELQ_ERR_DEVICE, "Krylov loop exceeded its host-side bound");
            if ((rc = iterations(check))) return rc;
            if ((rc = poll(hc))) return rc;
            if (hc.done) break;
        }
        if (ctx->batcher) ctx->batcher->leave(ctx->batch_slot);
        return ELQ_OK;
    }
};

template <class T>
int solve_fp64(elq_ctx* ctx, Solver<T>& S, Ctrl& hc) {
    Workspace& W = ctx->ws;
    int rc;
    ctx->prof_begin("init");
    HIPCHK(launch_init<T>(S.bk, (const double*)W.b64.p, (T*)W.vr.p, (T*)W.vx.p, (T*)W.vp0.p, (T*)W.vp1.p, S.prr(),
                          S.ctrl(), ctx->stream),
           "k_init");
    ctx->prof_end();
    if ((rc = S.scalar(SC_INIT, S.prr()))) return rc;
    if ((rc = S.krylov(hc))) return rc;
    return ELQ_OK;
}

// Mixed precision: FP32 PCG (FP32 storage of A, M and the Krylov vectors; FP64
// dot products and scalars) whose recursive residual is replaced by the FP64
// residual r = b - A x (x accumulated in FP64) whenever it has dropped by the
// factor opts.inner_tol since the last replacement, and whenever it meets the
// stopping test. The stopping test is always decided on the FP64 residual.
int solve_mixed(elq_ctx* ctx, Solver<float>& S, Ctrl& hc) {
    const double kReplaceDrop2 = S.o->inner_tol * S.o->inner_tol;
    Workspace& W = ctx->ws;
    int rc;
    const size_t cells = (size_t)S.bk.nb * kBrickCells;
    HIPCHK(W.x64.ensure(cells * 8), "alloc x64");
    HIPCHK(hipMemsetAsync(W.x64.p, 0, cells * 8, ctx->stream), "memset");
    auto replace = [&](bool first) -> int {
        if (!first) {
            ctx->prof_begin("xupdate");
            HIPCHK(launch_xupdate(S.bk, (double*)W.x64.p, (const float*)W.vx.p, S.ctrl(), ctx->stream), "k_xupdate");
            ctx->prof_end();
        }
        ctx->prof_begin("resid64");
        HIPCHK(launch_resid64<float>(S.bk, (const double*)W.diag64.p, S.c.ax, S.c.ay, S.c.az, (const double*)W.b64.p,
                                     (const double*)W.x64.p, S.prr(), S.prz(), (float*)W.vr.p, (float*)W.vx.p,
                                     first ? (float*)W.vp0.p : nullptr, first ? (float*)W.vp1.p : nullptr, S.ctrl(),
                                     ctx->stream),
               "k_resid64");
        ctx->prof_end();
        return S.scalar(SC_REPLACE, S.prr(), S.prz());
    };
    if ((rc = replace(true))) return rc;
    if ((rc = S.poll(hc))) return rc;
    if (hc.outer_done) return ELQ_OK;
    if ((rc = S.precondition(SC_RZ_INIT))) return rc;
    const int check = std::max(1, S.o->check_every);
    const int64_t cap = (int64_t)std::max(S.o->max_iter, 0) / check + 64;
    for (int64_t chunk = 0;; ++chunk) {
        if (chunk > cap) return ctx->fail(ELQ_ERR_DEVICE, "mixed Krylov loop exceeded its host-side bound");
        if ((rc = S.iterations(check))) return rc;
        if ((rc = S.poll(hc))) return rc;
        if (hc.done || hc.rr < kReplaceDrop2 * hc.rr_ref) {
            if ((rc = replace(false))) return rc;
            if ((rc = S.poll(hc))) return rc;
            if (hc.outer_done) break;
            if ((rc = S.precondition(SC_RZ_REPLACE))) return rc;
        }
    }
    if (ctx->batcher) ctx->batcher->leave(ctx->batch_slot);
    return ELQ_OK;
}

template <class T>
int project_impl(elq_ctx* ctx, const Dense& d, float dt, const elq_opts& o, elq_stats* st, float* pressure_out,
                 int32_t* indices_out) {
    const double t0 = now_ms();
    Workspace& W = ctx->ws;
    const Plan p = make_plan(d.ni, d.nj, d.nk);
    const bool mixed = o.precision == ELQ_PRECISION_MIXED;
    const bool levels = o.schedule == ELQ_SCHED_LEVELS;
    int rc = setup_phase_a(ctx, d, indices_out);
    if (rc) return rc;
    if (st) {
        st->n_unknowns = W.n_unknowns;
        st->n_bricks = W.nb;
    }
    if (W.setup_err & kErrBoundary)
        return ctx->fail(ELQ_ERR_BOUNDARY,
                         "an unknown lies on the outer cell layer (the reference reads outside its arrays there)");
    if (W.n_unknowns == 0) return ctx->fail(ELQ_ERR_NO_FLUID, "No fluid");
    if (W.mask.arena) {
        // allocation-free call: the whole workspace, now that the active
        // bricks are known, must fit the scratch array before any more work
        const size_t need = elq_footprint_bytes(d.ni, d.nj, d.nk, o.precision, false, W.nb) + kArenaSlack;
        if (need > W.mask.arena->size)
            return ctx->fail(ELQ_ERR_ADMISSION, "projection workspace (" + std::to_string(need) +
                                                    " bytes) exceeds the static scratch array (" +
                                                    std::to_string(W.mask.arena->size) + " bytes)");
    }
    if ((rc = setup_phase_b<T>(ctx, d, dt, levels))) return rc;
    Solver<T> S;
    S.ctx = ctx;
    S.p = p;
    S.bk = make_bricks(ctx, p);
    S.pcb = make_pencils(ctx, sizeof(T) == 8);
    S.pc = fwd_shape(ctx, sizeof(T) == 8, S.pcb);
    S.o = &o;
    S.levels = levels;
    // one launch for both sweeps of an iteration: mixed precision, 1x4 blocks,
    // on grids of up to kFuseMaxBricks active bricks (where the launch and the
    // forward grid tail between the sweeps are a larger share of a sweep), or
    // whenever asked for
    if constexpr (std::is_same<T, float>::value)
        S.fused = !levels && o.schedule != ELQ_SCHED_WAVEFRONT_SPLIT && sweeps_fusable(S.pc) &&
                  sweeps_fusable(S.pcb) && (o.schedule == ELQ_SCHED_WAVEFRONT_FUSED || S.bk.nb <= kFuseMaxBricks);
    S.c.diag = (const T*)W.cdiag.p;
    S.c.ax = (const T*)W.cax.p;
    S.c.ay = (const T*)W.cay.p;
    S.c.az = (const T*)W.caz.p;
    S.c.inv_e = mixed ? (const T*)W.inv_e32.p : (const T*)W.inv_e64.p;
    S.c.uflag = (const int32_t*)W.uflag.p;
    S.c.umap = (const int32_t*)W.umap.p;
    S.c.urow = (const T*)W.urow.p;
    if constexpr (sizeof(T) == 8) {
        const size_t nc = (size_t)S.bk.nb * kBrickCells;
        S.c.lx = (const T*)W.cscaled.p;
        S.c.ly = S.c.lx + nc;
        S.c.lz = S.c.lx + 2 * nc;
    }
    Ctrl hc;
    ctrl_host_init(hc, o, mixed);
    if ((rc = upload_ctrl(ctx, hc))) return rc;
    const double tau = o.precond == ELQ_PRECOND_MIC0 ? o.tau : 0.0;
    if ((rc = S.factor(tau, o.sigma, mixed ? (float*)W.inv_e32.p : nullptr))) return rc;
    // interior coefficients: couplings -1 (full fluid faces), diagonal 6 (FP32
    // keeps the diagonal's ghost part, 0 there); FP64 also flags interior pivots
    HIPCHK(launch_brick_uniform<T>(S.bk, S.c.diag, S.c.ax, S.c.ay, S.c.az, (T)-1, sizeof(T) == 4 ? (T)0 : (T)6,
                                   (const double*)W.inv_e64.p, tau, o.sigma, (int32_t*)W.uflag.p,
                                   (int32_t*)W.umap.p, (T*)W.urow.p, ctx->stream),
           "k_brick_uniform");
    if constexpr (sizeof(T) == 8)
        HIPCHK(launch_scale_lower(S.bk, S.c.ax, S.c.ay, S.c.az, S.c.inv_e, (double*)S.c.lx, (double*)S.c.ly,
                                  (double*)S.c.lz, ctx->stream),
               "k_scale_lower");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    const double t1 = now_ms();
    if constexpr (std::is_same<T, float>::value) {
        if ((rc = solve_mixed(ctx, S, hc))) return rc;
    } else {
        if ((rc = solve_fp64<T>(ctx, S, hc))) return rc;
    }
    const double t2 = now_ms();
    const double* xsol = mixed ? (const double*)W.x64.p : (const double*)W.vx.p;
    // true residual of the final iterate (FP64)
    if (!mixed) {
        ctx->prof_begin("resid64");
        HIPCHK(launch_resid64<T>(S.bk, (const double*)W.diag64.p, S.c.ax, S.c.ay, S.c.az, (const double*)W.b64.p,
                                 xsol, S.prr(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, ctx->stream),
               "k_resid64");
        ctx->prof_end();
        if ((rc = S.scalar(SC_TRUE, S.prr()))) return rc;
    }
    float* pressure = pressure_out;
    if (!pressure) {
        HIPCHK(W.pressure.ensure((size_t)p.C * 4), "alloc pressure");
        pressure = (float*)W.pressure.p;
    }
    ctx->prof_begin("scatter_update");
    HIPCHK(launch_scatter(d, (const uint8_t*)W.mask.p, S.bk, xsol, pressure, ctx->stream), "k_scatter");
    HIPCHK(launch_update(d, dt, pressure, ctx->stream), "k_update");
    ctx->prof_end();
    ctx->prof_collect();
    if ((rc = read_ctrl(ctx, hc))) return rc;
    if (hc.abort) return ctx->fail(ELQ_ERR_DEVICE, "wavefront wait exceeded its time limit");
    const double t3 = now_ms();
    if (st) {
        st->iterations = hc.iter_total;
        st->outer_iterations = mixed ? hc.outer : 1;
        st->error = hc.last_err;
        const double rr = mixed ? hc.rr64 : hc.true_rr;
        st->true_rel_residual = hc.bnorm2 > 0 ? std::sqrt(rr / hc.bnorm2) : 0.0;
        st->converged = hc.converged;
        st->status = hc.status ? hc.status : (hc.converged ? ELQ_OK : ELQ_ERR_NOT_CONVERGED);
        st->t_setup_ms = t1 - t0;
        st->t_solve_ms = t2 - t1;
        st->t_update_ms = t3 - t2;
        st->t_total_ms = t3 - t0;
    }
    if (hc.status == ELQ_ERR_BREAKDOWN) return ctx->fail(ELQ_ERR_BREAKDOWN, "non-finite Krylov scalar");
    if (!hc.converged) {
        ctx->last_error = "max_iter reached before the tolerance";
        return ELQ_ERR_NOT_CONVERGED;
    }
    return ELQ_OK;
}

bool check_opts(const elq_opts& o) {
    return o.tol >= 0 && o.max_iter >= 0 && (o.precision == 0 || o.precision == 1) &&
           (o.precond == 0 || o.precond == 1) && (o.schedule >= 0 && o.schedule <= 3) && o.sigma >= 0 &&
           o.inner_tol > 0 && o.inner_tol < 1;
}

Dense dense_of(const elq_grid& g) {
    Dense d;
    d.ni = g.ni;
    d.nj = g.nj;
    d.nk = g.nk;
    d.dx = g.dx;
    d.phi = g.level_set;
    d.u = g.u;
    d.v = g.v;
    d.w = g.w;
    d.uw = g.u_weight;
    d.vw = g.v_weight;
    d.ww = g.w_weight;
    return d;
}

struct Sizes {
    size_t c, u, v, w;
};
Sizes sizes_of(int ni, int nj, int nk) {
    return {(size_t)ni * nj * nk, (size_t)(ni + 1) * nj * nk, (size_t)ni * (nj + 1) * nk, (size_t)ni * nj * (nk + 1)};
}

// Uploads host grids into the context's staging buffers; returns device view.
int stage_inputs(elq_ctx* ctx, const elq_grid& g, elq_grid& dg) {
    Workspace& W = ctx->ws;
    const Sizes z = sizes_of(g.ni, g.nj, g.nk);
    HIPCHK(W.s_phi.ensure(z.c * 4), "alloc");
    HIPCHK(W.s_u.ensure(z.u * 4), "alloc");
    HIPCHK(W.s_v.ensure(z.v * 4), "alloc");
    HIPCHK(W.s_w.ensure(z.w * 4), "alloc");
    HIPCHK(W.s_uw.ensure(z.u * 4), "alloc");
    HIPCHK(W.s_vw.ensure(z.v * 4), "alloc");
    HIPCHK(W.s_ww.ensure(z.w * 4), "alloc");
    hipStream_t s = ctx->stream;
    HostStage& H = ctx->stage;
    HIPCHK(H.h2d(W.s_phi.p, g.level_set, z.c * 4, s), "h2d");
    HIPCHK(H.h2d(W.s_uw.p, g.u_weight, z.u * 4, s), "h2d");
    HIPCHK(H.h2d(W.s_vw.p, g.v_weight, z.v * 4, s), "h2d");
    HIPCHK(H.h2d(W.s_ww.p, g.w_weight, z.w * 4, s), "h2d");
    HIPCHK(H.h2d(W.s_u.p, g.u, z.u * 4, s), "h2d");
    HIPCHK(H.h2d(W.s_v.p, g.v, z.v * 4, s), "h2d");
    HIPCHK(H.h2d(W.s_w.p, g.w, z.w * 4, s), "h2d");
    dg = g;
    dg.level_set = (const float*)W.s_phi.p;
    dg.u = (float*)W.s_u.p;
    dg.v = (float*)W.s_v.p;
    dg.w = (float*)W.s_w.p;
    dg.u_weight = (const float*)W.s_uw.p;
    dg.v_weight = (const float*)W.s_vw.p;
    dg.w_weight = (const float*)W.s_ww.p;
    dg.pressure = nullptr;
    dg.indices = nullptr;
    return ELQ_OK;
}

int check_grid(elq_ctx* ctx, const elq_grid* g) {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!g || !g->level_set || !g->u || !g->v || !g->w || !g->u_weight || !g->v_weight || !g->w_weight)
        return ctx->fail(ELQ_ERR_INVALID_ARG, "null grid pointer");
    if (!valid_dims(g->ni, g->nj, g->nk)) return ctx->fail(ELQ_ERR_INVALID_ARG, "grid dimensions out of range");
    if (!(g->dx > 0.0f) || !std::isfinite(g->dx)) return ctx->fail(ELQ_ERR_INVALID_ARG, "dx must be positive");
    return ELQ_OK;
}

int run_project(elq_ctx* ctx, const elq_grid& dg, float dt, const elq_opts& o, elq_stats* st) {
    if (!(dt > 0.0f) || !std::isfinite(dt)) return ctx->fail(ELQ_ERR_INVALID_ARG, "dt must be positive");
    if (!check_opts(o)) return ctx->fail(ELQ_ERR_INVALID_ARG, "invalid solver options");
    if (ctx->budget) {
        const size_t need = elq_footprint_bytes(dg.ni, dg.nj, dg.nk, o.precision, false, -1);
        if (need > ctx->budget) return ctx->fail(ELQ_ERR_ADMISSION, "projection exceeds the context's HBM budget");
    }
    const Dense d = dense_of(dg);
    if (o.precision == ELQ_PRECISION_MIXED) return project_impl<float>(ctx, d, dt, o, st, dg.pressure, dg.indices);
    return project_impl<double>(ctx, d, dt, o, st, dg.pressure, dg.indices);
}

}  // namespace

// ============================================================ C ABI
extern "C" {

void elq_default_opts(elq_opts* o) {
    if (!o) return;
    o->tol = 1e-4;
    o->max_iter = 1000;
    o->precision = ELQ_PRECISION_FP64;
    o->precond = ELQ_PRECOND_IC0;
    o->tau = 0.97;
    o->sigma = 0.25;
    o->inner_tol = 0.01;
    o->schedule = ELQ_SCHED_WAVEFRONT;
    o->check_every = 8;
}

int elq_ctx_create(int device, size_t hbm_budget_bytes, elq_ctx** out) try {
    if (!out) return ELQ_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return ELQ_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return ELQ_ERR_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return ELQ_ERR_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return ELQ_ERR_DEVICE;
    elq_ctx* ctx = new elq_ctx();
    ctx->device = device;
    ctx->budget = hbm_budget_bytes;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return ELQ_ERR_DEVICE;
    }
    // pinned host copy of the control block, written by a kernel at each poll
    if (hipHostMalloc(&ctx->ctrl_host, kHostMirrorBytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return ELQ_ERR_DEVICE;
    }
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || rate_khz <= 0)
        rate_khz = 100000;
    ctx->clock_hz = (uint64_t)rate_khz * 1000ull;
    ctx->timeout_ticks = ctx->clock_hz * 4ull;  // 4 s per wait
    ctx->rei_rebase = kReiRebase;
    ctx->ws.for_each([&](DevBuf& b) { b.stream = ctx->stream; });
    ctx->ws_scratch.for_each([&](DevBuf& b) { b.stream = ctx->stream; });
    // the static scratch array of the allocation-free projection; a failed
    // lookup only disables that entry point
    if (scratch_address(device, &ctx->scratch) != hipSuccess) ctx->scratch = nullptr;
    (void)hipGetLastError();
    *out = ctx;
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

void elq_ctx_destroy(elq_ctx* ctx) try {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->stage.release();
    ctx->drop_chunk_graph();
    ctx->ws.release_all();
    (void)hipStreamSynchronize(ctx->stream);  // the stream-ordered frees
    for (auto ev : ctx->events) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(ctx->stream);
    if (ctx->ctrl_host) (void)hipHostFree(ctx->ctrl_host);
    delete ctx;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
}

const char* elq_ctx_last_error(const elq_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int elq_set_graph_capture(elq_ctx* ctx, int enable) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    if (!enable) ctx->drop_chunk_graph();
    ctx->graphs = enable != 0;
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_set_reinit_rebase(elq_ctx* ctx, int32_t passes) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (passes < 1 || passes > kReiRebase) return ctx->fail(ELQ_ERR_INVALID_ARG, "rebase interval out of range");
    ctx->rei_rebase = passes;
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_set_wait_limit(elq_ctx* ctx, double seconds) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!(seconds >= 0.0) || seconds > 3600.0) return ctx->fail(ELQ_ERR_INVALID_ARG, "wait limit out of range");
    ctx->timeout_ticks = (uint64_t)(seconds * (double)ctx->clock_hz);
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

size_t elq_ctx_device_bytes(const elq_ctx* ctx) { return ctx ? ctx->ws.bytes() : 0; }

size_t elq_projection_footprint(int32_t ni, int32_t nj, int32_t nk, int32_t precision, const float* level_set) try {
    if (!valid_dims(ni, nj, nk)) return 0;
    const int64_t nb = level_set ? liquid_bricks(level_set, ni, nj, nk) : -1;
    return elq_footprint_bytes(ni, nj, nk, precision, true, nb);
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return 0;
}

int elq_project_device(elq_ctx* ctx, const elq_grid* g, float dt, const elq_opts* opts, elq_stats* st) try {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    elq_opts o;
    if (opts)
        o = *opts;
    else
        elq_default_opts(&o);
    if (st) std::memset(st, 0, sizeof(*st));
    (void)hipSetDevice(ctx->device);
    ctx->kt.clear();
    rc = run_project(ctx, *g, dt, o, st);
    if (st && rc != ELQ_OK && st->status == 0) st->status = rc;
    return rc;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_project_device_noalloc(elq_ctx* ctx, const elq_grid* g, float dt, const elq_opts* opts, elq_stats* st) try {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!ctx->scratch) return ctx->fail(ELQ_ERR_DEVICE, "static scratch array unavailable on this device");
    elq_opts o;
    if (opts)
        o = *opts;
    else
        elq_default_opts(&o);
    if (st) std::memset(st, 0, sizeof(*st));
    // dense arrays and brick-grid metadata alone must fit (known from the dims)
    if (valid_dims(g->ni, g->nj, g->nk) &&
        elq_footprint_bytes(g->ni, g->nj, g->nk, o.precision, false, 0) + kArenaSlack > ELQ_NOALLOC_SCRATCH_BYTES)
        return ctx->fail(ELQ_ERR_ADMISSION, "grid too large for the static scratch array (" +
                                                std::to_string((size_t)ELQ_NOALLOC_SCRATCH_BYTES) + " bytes)");
    (void)hipSetDevice(ctx->device);
    ctx->kt.clear();
    // a fresh carve of the scratch array: every buffer (control block,
    // exchange and stage words included) is cleared where it must start at zero
    ctx->arena.base = ctx->scratch;
    ctx->arena.size = ELQ_NOALLOC_SCRATCH_BYTES;
    ctx->arena.off = 0;
    ctx->arena.overflow = false;
    ctx->ws_scratch.for_each([&](DevBuf& b) {
        b.p = nullptr;
        b.cap = 0;
        b.arena = &ctx->arena;
    });
    const bool prof = ctx->profiling;  // event timing creates events on demand: off here
    ctx->profiling = false;
    std::swap(ctx->ws, ctx->ws_scratch);
    rc = run_project(ctx, *g, dt, o, st);
    std::swap(ctx->ws, ctx->ws_scratch);
    ctx->profiling = prof;
    if (rc != ELQ_OK && ctx->arena.overflow) {
        rc = ELQ_ERR_ADMISSION;
        ctx->last_error = "projection workspace exceeds the static scratch array";
    }
    if (st && rc != ELQ_OK && st->status == 0) st->status = rc;
    return rc;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_project(elq_ctx* ctx, const elq_grid* g, float dt, const elq_opts* opts, elq_stats* st) try {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    elq_opts o;
    if (opts)
        o = *opts;
    else
        elq_default_opts(&o);
    if (st) std::memset(st, 0, sizeof(*st));
    (void)hipSetDevice(ctx->device);
    ctx->kt.clear();
    elq_grid dg;
    if ((rc = stage_inputs(ctx, *g, dg))) return rc;
    if (g->pressure) {
        if (ctx->ws.s_pressure.ensure(sizes_of(g->ni, g->nj, g->nk).c * 4) != hipSuccess)
            return ctx->fail(ELQ_ERR_DEVICE, "alloc");
        dg.pressure = (float*)ctx->ws.s_pressure.p;
    }
    if (g->indices) {
        if (ctx->ws.s_indices.ensure(sizes_of(g->ni, g->nj, g->nk).c * 4) != hipSuccess)
            return ctx->fail(ELQ_ERR_DEVICE, "alloc");
        dg.indices = (int32_t*)ctx->ws.s_indices.p;
    }
    rc = run_project(ctx, dg, dt, o, st);
    if (st && rc != ELQ_OK && st->status == 0) st->status = rc;
    if (rc != ELQ_OK && rc != ELQ_ERR_NOT_CONVERGED) {
        // indices are meaningful even on NO_FLUID (all INT_MAX); copy them back
        if (rc == ELQ_ERR_NO_FLUID && g->indices) {
            const Sizes z = sizes_of(g->ni, g->nj, g->nk);
            if (hipMemcpyAsync(g->indices, dg.indices, z.c * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess)
                return ctx->fail(ELQ_ERR_DEVICE, "d2h");
        }
        return rc;
    }
    const Sizes z = sizes_of(g->ni, g->nj, g->nk);
    std::vector<HostStage::Copy> back = {{g->u, dg.u, z.u * 4}, {g->v, dg.v, z.v * 4}, {g->w, dg.w, z.w * 4}};
    if (g->pressure) back.push_back({g->pressure, dg.pressure, z.c * 4});
    if (g->indices) back.push_back({g->indices, dg.indices, z.c * 4});
    const hipError_t e = ctx->stage.d2h(back, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "copy back");
    return rc;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_assign_indices(elq_ctx* ctx, const elq_grid* g, int32_t* indices, int64_t* n_unknowns) try {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!indices) return ctx->fail(ELQ_ERR_INVALID_ARG, "null indices");
    (void)hipSetDevice(ctx->device);
    elq_grid dg;
    if ((rc = stage_inputs(ctx, *g, dg))) return rc;
    if ((rc = setup_phase_a(ctx, dense_of(dg), nullptr))) return rc;
    const Sizes z = sizes_of(g->ni, g->nj, g->nk);
    if (hipMemcpyAsync(indices, ctx->ws.cur_indices, z.c * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return ctx->fail(ELQ_ERR_DEVICE, "d2h");
    if (n_unknowns) *n_unknowns = ctx->ws.n_unknowns;
    if (ctx->ws.n_unknowns == 0) return ctx->fail(ELQ_ERR_NO_FLUID, "No fluid");
    if (ctx->ws.setup_err & kErrBoundary) return ctx->fail(ELQ_ERR_BOUNDARY, "unknown on the outer cell layer");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_build_matrix(elq_ctx* ctx, const elq_grid* g, float dt, double* diag, double* ax, double* ay, double* az,
                     double* b) try {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!diag || !ax || !ay || !az || !b) return ctx->fail(ELQ_ERR_INVALID_ARG, "null output");
    if (!(dt > 0.0f)) return ctx->fail(ELQ_ERR_INVALID_ARG, "dt must be positive");
    (void)hipSetDevice(ctx->device);
    elq_grid dg;
    if ((rc = stage_inputs(ctx, *g, dg))) return rc;
    const Dense d = dense_of(dg);
    if ((rc = setup_phase_a(ctx, d, nullptr))) return rc;
    Workspace& W = ctx->ws;
    if (W.n_unknowns == 0) return ctx->fail(ELQ_ERR_NO_FLUID, "No fluid");
    if (W.setup_err & kErrBoundary) return ctx->fail(ELQ_ERR_BOUNDARY, "unknown on the outer cell layer");
    if ((rc = setup_phase_b<double>(ctx, d, dt, false))) return rc;
    const size_t n = (size_t)W.n_unknowns;
    HIPCHK(W.export_buf.ensure(n * 8 * 5), "alloc");
    double* e = (double*)W.export_buf.p;
    const Plan p = make_plan(d.ni, d.nj, d.nk);
    Bricks bk = make_bricks(ctx, p);
    HIPCHK(launch_export<double>(d, (const uint8_t*)W.mask.p, W.cur_indices, bk, (const double*)W.diag64.p,
                                 (const double*)W.b64.p, (const double*)W.cax.p, (const double*)W.cay.p,
                                 (const double*)W.caz.p, e, e + n, e + 2 * n, e + 3 * n, e + 4 * n, ctx->stream),
           "k_export");
    double* outs[5] = {diag, ax, ay, az, b};
    for (int q = 0; q < 5; ++q)
        HIPCHK(hipMemcpyAsync(outs[q], e + q * n, n * 8, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_ic_factor(elq_ctx* ctx, const elq_grid* g, float dt, const elq_opts* opts, double* inv_e) try {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!inv_e) return ctx->fail(ELQ_ERR_INVALID_ARG, "null output");
    if (!(dt > 0.0f)) return ctx->fail(ELQ_ERR_INVALID_ARG, "dt must be positive");
    elq_opts o;
    if (opts)
        o = *opts;
    else
        elq_default_opts(&o);
    if (!check_opts(o)) return ctx->fail(ELQ_ERR_INVALID_ARG, "invalid solver options");
    (void)hipSetDevice(ctx->device);
    elq_grid dg;
    if ((rc = stage_inputs(ctx, *g, dg))) return rc;
    const Dense d = dense_of(dg);
    if ((rc = setup_phase_a(ctx, d, nullptr))) return rc;
    Workspace& W = ctx->ws;
    if (W.n_unknowns == 0) return ctx->fail(ELQ_ERR_NO_FLUID, "No fluid");
    if (W.setup_err & kErrBoundary) return ctx->fail(ELQ_ERR_BOUNDARY, "unknown on the outer cell layer");
    const bool levels = o.schedule == ELQ_SCHED_LEVELS;
    if ((rc = setup_phase_b<double>(ctx, d, dt, levels))) return rc;
    const Plan p = make_plan(d.ni, d.nj, d.nk);
    Solver<double> S;
    S.ctx = ctx;
    S.p = p;
    S.bk = make_bricks(ctx, p);
    S.pc = S.pcb = make_pencils(ctx, true);
    S.o = &o;
    S.levels = levels;
    S.c.diag = (const double*)W.cdiag.p;
    S.c.ax = (const double*)W.cax.p;
    S.c.ay = (const double*)W.cay.p;
    S.c.az = (const double*)W.caz.p;
    S.c.inv_e = (const double*)W.inv_e64.p;
    Ctrl hc;
    ctrl_host_init(hc, o, 0);
    if ((rc = upload_ctrl(ctx, hc))) return rc;
    if ((rc = S.factor(o.precond == ELQ_PRECOND_MIC0 ? o.tau : 0.0, o.sigma, nullptr))) return rc;
    const size_t n = (size_t)W.n_unknowns;
    HIPCHK(W.export_buf.ensure(n * 8), "alloc");
    HIPCHK(launch_export_cells(d, (const uint8_t*)W.mask.p, W.cur_indices, S.bk, (const double*)W.inv_e64.p,
                               (double*)W.export_buf.p, ctx->stream),
           "k_export_cells");
    if ((rc = read_ctrl(ctx, hc))) return rc;
    if (hc.abort) return ctx->fail(ELQ_ERR_DEVICE, "wavefront wait exceeded its time limit");
    HIPCHK(hipMemcpyAsync(inv_e, W.export_buf.p, n * 8, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_update_velocity(elq_ctx* ctx, const elq_grid* g, float dt) try {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!g->pressure) return ctx->fail(ELQ_ERR_INVALID_ARG, "null pressure");
    (void)hipSetDevice(ctx->device);
    elq_grid dg;
    if ((rc = stage_inputs(ctx, *g, dg))) return rc;
    const Sizes z = sizes_of(g->ni, g->nj, g->nk);
    Workspace& W = ctx->ws;
    HIPCHK(W.s_pressure.ensure(z.c * 4), "alloc");
    HIPCHK(ctx->stage.h2d(W.s_pressure.p, g->pressure, z.c * 4, ctx->stream), "h2d");
    HIPCHK(launch_update(dense_of(dg), dt, (const float*)W.s_pressure.p, ctx->stream), "k_update");
    HIPCHK(ctx->stage.d2h({{g->u, dg.u, z.u * 4}, {g->v, dg.v, z.v * 4}, {g->w, dg.w, z.w * 4}}, ctx->stream),
           "d2h");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_compute_weights_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, const float* solid_nodal,
                               float* u_weight, float* v_weight, float* w_weight) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!solid_nodal || !u_weight || !v_weight || !w_weight) return ctx->fail(ELQ_ERR_INVALID_ARG, "null pointer");
    if (!valid_dims(ni, nj, nk)) return ctx->fail(ELQ_ERR_INVALID_ARG, "grid dimensions out of range");
    (void)hipSetDevice(ctx->device);
    ctx->kt.clear();
    ctx->prof_begin("weights");
    HIPCHK(launch_weights(ni, nj, nk, solid_nodal, u_weight, v_weight, w_weight, ctx->stream), "k_weights");
    ctx->prof_end();
    ctx->prof_collect();
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_add_gravity_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float* v, float dt) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!v) return ctx->fail(ELQ_ERR_INVALID_ARG, "null pointer");
    if (!valid_dims(ni, nj, nk)) return ctx->fail(ELQ_ERR_INVALID_ARG, "grid dimensions out of range");
    (void)hipSetDevice(ctx->device);
    ctx->kt.clear();
    ctx->prof_begin("gravity");
    HIPCHK(launch_gravity(v, (size_t)ni * (size_t)(nj + 1) * (size_t)nk, dt, ctx->stream), "k_gravity");
    ctx->prof_end();
    ctx->prof_collect();
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_valid_and_zero_device(elq_ctx* ctx, const elq_grid* g) try {
    const int rc = check_grid(ctx, g);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    HIPCHK(launch_valid_zero(dense_of(*g), ctx->stream), "k_valid_zero");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_find_valid_velocity_device(elq_ctx* ctx, const elq_grid* g, uint8_t* u_valid, uint8_t* v_valid,
                                   uint8_t* w_valid) try {
    const int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!u_valid || !v_valid || !w_valid) return ctx->fail(ELQ_ERR_INVALID_ARG, "null pointer");
    (void)hipSetDevice(ctx->device);
    ctx->kt.clear();
    ctx->prof_begin("find_valid");
    HIPCHK(launch_find_valid(dense_of(*g), u_valid, v_valid, w_valid, ctx->stream), "k_find_valid");
    ctx->prof_end();
    ctx->prof_collect();
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_set_invalid_velocity_zero_device(elq_ctx* ctx, const elq_grid* g, const uint8_t* u_valid,
                                         const uint8_t* v_valid, const uint8_t* w_valid) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!g || !g->u || !g->v || !g->w || !u_valid || !v_valid || !w_valid)
        return ctx->fail(ELQ_ERR_INVALID_ARG, "null pointer");
    if (!valid_dims(g->ni, g->nj, g->nk)) return ctx->fail(ELQ_ERR_INVALID_ARG, "grid dimensions out of range");
    const Sizes z = sizes_of(g->ni, g->nj, g->nk);
    (void)hipSetDevice(ctx->device);
    ctx->kt.clear();
    ctx->prof_begin("zero_invalid");
    HIPCHK(launch_zero_invalid(g->u, u_valid, z.u, ctx->stream), "k_zero_invalid");
    HIPCHK(launch_zero_invalid(g->v, v_valid, z.v, ctx->stream), "k_zero_invalid");
    HIPCHK(launch_zero_invalid(g->w, w_valid, z.w, ctx->stream), "k_zero_invalid");
    ctx->prof_end();
    ctx->prof_collect();
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_find_valid_level_set_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, const float* solid_nodal,
                                    uint8_t* valid) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!solid_nodal || !valid) return ctx->fail(ELQ_ERR_INVALID_ARG, "null pointer");
    if (!valid_dims(ni, nj, nk)) return ctx->fail(ELQ_ERR_INVALID_ARG, "grid dimensions out of range");
    (void)hipSetDevice(ctx->device);
    ctx->kt.clear();
    ctx->prof_begin("level_set_valid");
    HIPCHK(launch_level_set_valid(ni, nj, nk, solid_nodal, valid, ctx->stream), "k_level_set_valid");
    ctx->prof_end();
    ctx->prof_collect();
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

namespace {

struct ExtArgs {
    int ni, nj, nk;
    float* f;
    const uint8_t* valid_in;
    uint8_t* valid_out;
};

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// extrapolation.cpp:5-46 over up to three grids at once. The loop's early
// return (a layer that starts with no invalid cell) needs no host check: such
// a layer adds no cell to the plane and lists nothing. Layers beyond the
// largest grid dimension change nothing either (every cell is then within
// reach of a valid cell, or there is none), so they are not launched.
int run_extrapolate(elq_ctx* ctx, const ExtArgs* a, int count, int layers) {
    ExtSet set{};
    size_t off[3][3];
    size_t total = 0;
    int64_t words = 0;
    int maxdim = 0;
    for (int q = 0; q < count; ++q) {
        const int64_t cells = (int64_t)a[q].ni * a[q].nj * a[q].nk;
        const int64_t wq = ext_words(a[q].ni, a[q].nj, a[q].nk);
        off[q][0] = total;
        total += align256((size_t)wq * 8);
        off[q][1] = total;
        total += align256((size_t)wq * 8);
        off[q][2] = total;
        total += align256((size_t)cells * 4);
        ExtGrid& g = set.g[q];
        g.f = a[q].f;
        g.valid_in = a[q].valid_in;
        g.valid_out = a[q].valid_out;
        g.ni = a[q].ni;
        g.nj = a[q].nj;
        g.nk = a[q].nk;
        g.wpr = ext_wpr(a[q].ni);
        g.slot = q;
        g.word0 = words;
        g.words = wq;
        words += (wq + kExtWordAlign - 1) / kExtWordAlign * kExtWordAlign;
        maxdim = std::max({maxdim, a[q].ni, a[q].nj, a[q].nk});
    }
    const size_t counter_off = total;
    total += 256;
    set.count = count;
    set.words = words;
    layers = std::min(layers, maxdim);
    if (layers <= 0) return ELQ_OK;
    HIPCHK(ctx->ws.ext.ensure(total), "alloc");
    uint8_t* base = (uint8_t*)ctx->ws.ext.p;
    for (int q = 0; q < count; ++q) {
        set.g[q].plane[0] = (uint64_t*)(base + off[q][0]);
        set.g[q].plane[1] = (uint64_t*)(base + off[q][1]);
        set.g[q].list = (uint32_t*)(base + off[q][2]);
    }
    set.counter = (uint32_t*)(base + counter_off);
    ctx->kt.clear();
    ctx->prof_begin("ext_init");
    HIPCHK(launch_ext_init(set, ctx->stream), "k_ext_init");
    ctx->prof_end();
    for (int l = 0; l < layers; ++l) {
        ctx->prof_begin("ext_layer");
        HIPCHK(launch_ext_layer(set, l, ctx->stream), "k_ext_layer");
        ctx->prof_end();
    }
    ctx->prof_collect();
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
}

}  // namespace

int elq_extrapolate_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float* grid, uint8_t* valid,
                           int32_t layers) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!grid || !valid) return ctx->fail(ELQ_ERR_INVALID_ARG, "null pointer");
    if (!valid_dims(ni, nj, nk)) return ctx->fail(ELQ_ERR_INVALID_ARG, "grid dimensions out of range");
    (void)hipSetDevice(ctx->device);
    const ExtArgs a{ni, nj, nk, grid, valid, valid};
    return run_extrapolate(ctx, &a, 1, layers);
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_extrapolate_const_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float* grid,
                                 const uint8_t* valid, int32_t layers) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!grid || !valid) return ctx->fail(ELQ_ERR_INVALID_ARG, "null pointer");
    if (!valid_dims(ni, nj, nk)) return ctx->fail(ELQ_ERR_INVALID_ARG, "grid dimensions out of range");
    (void)hipSetDevice(ctx->device);
    const ExtArgs a{ni, nj, nk, grid, valid, nullptr};
    return run_extrapolate(ctx, &a, 1, layers);
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_extrapolate_velocity_device(elq_ctx* ctx, const elq_grid* g, uint8_t* u_valid, uint8_t* v_valid,
                                    uint8_t* w_valid, int32_t layers) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!g || !g->u || !g->v || !g->w || !u_valid || !v_valid || !w_valid)
        return ctx->fail(ELQ_ERR_INVALID_ARG, "null pointer");
    const int ni = g->ni, nj = g->nj, nk = g->nk;
    if (!valid_dims(ni, nj, nk)) return ctx->fail(ELQ_ERR_INVALID_ARG, "grid dimensions out of range");
    (void)hipSetDevice(ctx->device);
    const ExtArgs a[3] = {{ni + 1, nj, nk, g->u, u_valid, u_valid},
                          {ni, nj + 1, nk, g->v, v_valid, v_valid},
                          {ni, nj, nk + 1, g->w, w_valid, w_valid}};
    return run_extrapolate(ctx, a, 3, layers);
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_advect_staggered_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float dx, float dt,
                                float* out_u, float* out_v, float* out_w, const float* val_u, const float* val_v,
                                const float* val_w, const float* vel_u, const float* vel_v, const float* vel_w) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!out_u || !out_v || !out_w || !val_u || !val_v || !val_w || !vel_u || !vel_v || !vel_w)
        return ctx->fail(ELQ_ERR_INVALID_ARG, "null pointer");
    if (!valid_dims(ni, nj, nk) || !(dx > 0.0f)) return ctx->fail(ELQ_ERR_INVALID_ARG, "grid dimensions out of range");
    const float* in[6] = {val_u, val_v, val_w, vel_u, vel_v, vel_w};
    for (const float* o : {out_u, out_v, out_w})
        for (const float* p : in)
            if (o == p) return ctx->fail(ELQ_ERR_INVALID_ARG, "outputs must not alias inputs");
    (void)hipSetDevice(ctx->device);
    ctx->kt.clear();
    ctx->prof_begin("advect_faces");
    HIPCHK(launch_advect_faces(ni, nj, nk, dx, dt, vel_u, vel_v, vel_w, val_u, val_v, val_w, out_u, out_v, out_w,
                               ctx->stream),
           "k_advect_faces");
    ctx->prof_end();
    ctx->prof_collect();
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_advect_collocated_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float dx, float dt, float* out,
                                 const float* val, const float* vel_u, const float* vel_v, const float* vel_w) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!out || !val || !vel_u || !vel_v || !vel_w) return ctx->fail(ELQ_ERR_INVALID_ARG, "null pointer");
    if (!valid_dims(ni, nj, nk) || !(dx > 0.0f)) return ctx->fail(ELQ_ERR_INVALID_ARG, "grid dimensions out of range");
    if (out == val || out == vel_u || out == vel_v || out == vel_w)
        return ctx->fail(ELQ_ERR_INVALID_ARG, "output must not alias inputs");
    (void)hipSetDevice(ctx->device);
    ctx->kt.clear();
    ctx->prof_begin("advect_cells");
    HIPCHK(launch_advect_cells(ni, nj, nk, dx, dt, vel_u, vel_v, vel_w, val, out, ctx->stream), "k_advect_cells");
    ctx->prof_end();
    ctx->prof_collect();
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_advect_velocity_device(elq_ctx* ctx, const elq_grid* g, float dt) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!g || !g->u || !g->v || !g->w) return ctx->fail(ELQ_ERR_INVALID_ARG, "null pointer");
    if (!valid_dims(g->ni, g->nj, g->nk) || !(g->dx > 0.0f))
        return ctx->fail(ELQ_ERR_INVALID_ARG, "grid dimensions out of range");
    const Sizes z = sizes_of(g->ni, g->nj, g->nk);
    (void)hipSetDevice(ctx->device);
    HIPCHK(ctx->ws.adv.ensure(4 * (z.u + z.v + z.w)), "alloc");
    float* t = (float*)ctx->ws.adv.p;
    float *tu = t, *tv = t + z.u, *tw = t + z.u + z.v;
    // m_velocity_temp = m_velocity (simulator.cpp:80)
    HIPCHK(launch_copy(tu, g->u, 4 * z.u, ctx->stream), "k_copy");
    HIPCHK(launch_copy(tv, g->v, 4 * z.v, ctx->stream), "k_copy");
    HIPCHK(launch_copy(tw, g->w, 4 * z.w, ctx->stream), "k_copy");
    ctx->kt.clear();
    ctx->prof_begin("advect_faces");
    HIPCHK(launch_advect_faces(g->ni, g->nj, g->nk, g->dx, dt, tu, tv, tw, tu, tv, tw, g->u, g->v, g->w, ctx->stream),
           "k_advect_faces");
    ctx->prof_end();
    ctx->prof_collect();
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_advect_level_set_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float dx, float dt,
                                float* level_set, const float* u, const float* v, const float* w) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!level_set || !u || !v || !w) return ctx->fail(ELQ_ERR_INVALID_ARG, "null pointer");
    if (!valid_dims(ni, nj, nk) || !(dx > 0.0f)) return ctx->fail(ELQ_ERR_INVALID_ARG, "grid dimensions out of range");
    const Sizes z = sizes_of(ni, nj, nk);
    (void)hipSetDevice(ctx->device);
    HIPCHK(ctx->ws.adv.ensure(4 * z.c), "alloc");
    float* t = (float*)ctx->ws.adv.p;
    // m_level_set_temp = m_level_set (simulator.cpp:85)
    HIPCHK(launch_copy(t, level_set, 4 * z.c, ctx->stream), "k_copy");
    ctx->kt.clear();
    ctx->prof_begin("advect_cells");
    HIPCHK(launch_advect_cells(ni, nj, nk, dx, dt, u, v, w, t, level_set, ctx->stream), "k_advect_cells");
    ctx->prof_end();
    ctx->prof_collect();
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_reinitialize_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float dx, float* level_set,
                            const uint8_t* valid, int64_t* unreached, int32_t* passes) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (!level_set || !valid) return ctx->fail(ELQ_ERR_INVALID_ARG, "null pointer");
    if (!valid_dims(ni, nj, nk) || !(dx > 0.0f)) return ctx->fail(ELQ_ERR_INVALID_ARG, "grid dimensions out of range");
    (void)hipSetDevice(ctx->device);
    const size_t cells = (size_t)ni * nj * nk;
    ReiSet r{};
    r.ni = ni;
    r.nj = nj;
    r.nk = nk;
    r.dx = dx;
    r.tx = (ni + kExtTX - 1) / kExtTX;
    r.ty = (nj + kExtTY - 1) / kExtTY;
    r.tz = (nk + kExtTZ - 1) / kExtTZ;
    r.tiles = r.tx * r.ty * r.tz;
    size_t off = 0;
    const size_t o_temp = off;
    off += align256(4 * cells);
    const size_t o_stamp = off;
    off += align256(2 * cells);
    const size_t o_tfar = off;
    off += align256((size_t)r.tiles);
    const size_t o_tnew = off;
    off += align256(2 * (size_t)r.tiles);
    const size_t o_list = off;
    off += align256(4 * (size_t)r.tiles);
    const size_t o_dist = off;
    off += align256(2 * cells);
    const size_t o_dcount = off;
    off += align256(4 * (size_t)(kReiDistMax + 2));
    const size_t o_dfill = off;
    off += align256(4 * (size_t)(kReiDistMax + 2));
    const size_t o_cells = off;
    off += align256(4 * cells);
    const size_t o_fixm = off;
    off += align256(cells);
    const size_t o_counter = off;
    off += 256;
    HIPCHK(ctx->ws.rei.ensure(off), "alloc");
    uint8_t* base = (uint8_t*)ctx->ws.rei.p;
    r.phi = level_set;
    r.temp = (const float*)(base + o_temp);
    r.valid = valid;
    r.stamp = (uint16_t*)(base + o_stamp);
    r.tfar = base + o_tfar;
    r.tnew = (uint16_t*)(base + o_tnew);
    r.list = (int32_t*)(base + o_list);
    r.dist = (uint16_t*)(base + o_dist);
    r.dcount = (uint32_t*)(base + o_dcount);
    r.dfill = (uint32_t*)(base + o_dfill);
    r.cells = (uint32_t*)(base + o_cells);
    r.fixm = base + o_fixm;
    r.counter = (uint32_t*)(base + o_counter);
    ctx->kt.clear();
    // grid_temp = grid (reinitialization.cpp:26)
    HIPCHK(launch_copy(base + o_temp, level_set, 4 * cells, ctx->stream), "k_copy");
    ctx->prof_begin("rei_init");
    HIPCHK(launch_rei_init(r, ctx->stream), "k_rei_init");
    ctx->prof_end();
    // Where no invalid cell lengthens a path, pass l fixes the cells at
    // taxicab distance l + 1 from the interface: one launch per pass over
    // those cells only.
    ctx->prof_begin("rei_dist");
    HIPCHK(launch_rei_dist(r, ctx->stream), "k_rei_dist");
    ctx->prof_end();
    uint32_t chk[2] = {1, 0};
    HIPCHK(hipMemcpyAsync(chk, r.counter + 6, 8, hipMemcpyDeviceToHost, ctx->stream), "read counter");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    if (chk[0] == 0) {
        ctx->prof_begin("rei_lists");
        HIPCHK(launch_rei_lists(r, ctx->stream), "k_rei_lists");
        ctx->prof_end();
        std::vector<uint32_t> offs(chk[1] + 2);
        HIPCHK(hipMemcpyAsync(offs.data(), r.dcount, offs.size() * 4, hipMemcpyDeviceToHost, ctx->stream),
               "read offsets");
        HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
        for (uint32_t t = 1; t <= chk[1]; ++t) {
            ctx->prof_begin("rei_layer");
            HIPCHK(launch_rei_layer(r, offs[t], offs[t + 1], ctx->stream), "k_rei_layer");
            ctx->prof_end();
        }
        ctx->prof_collect();
        HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
        if (unreached) *unreached = 0;
        if (passes) *passes = (int32_t)chk[1];
        return ELQ_OK;
    }
    // Passes in batches; after each batch the host reads how many cells the
    // batch's last pass fixed. A pass that fixes nothing ends the march: no
    // later pass could change anything (the reference then returns, or, with
    // cells that no path reaches, loops forever). Passes launched after the
    // end find no tile to work on.
    constexpr int kBatch = 32;
    int64_t first = 0;  // first pass of the current stamp epoch
    int64_t l = 0;
    uint32_t last_fixed = 1;
    while (last_fixed != 0 && l <= (int64_t)cells) {
        for (int q = 0; q < kBatch; ++q, ++l) {
            if (l - first == ctx->rei_rebase) {
                HIPCHK(launch_rei_rebase(r, (uint32_t)first, ctx->stream), "k_rei_rebase");
                first = l;
            }
            ctx->prof_begin("rei_pass");
            HIPCHK(launch_rei_pass(r, (int)(l - first), (int)(l & 1), ctx->stream), "k_rei_pass");
            ctx->prof_end();
        }
        HIPCHK(hipMemcpyAsync(&last_fixed, r.counter + 2 + ((l - 1) & 1), 4, hipMemcpyDeviceToHost, ctx->stream),
               "read counter");
        HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    }
    HIPCHK(launch_rei_finish(r, ctx->stream), "k_rei_finish");
    uint32_t fin[5] = {0, 0, 0, 0, 0};  // counter[4..8]
    HIPCHK(hipMemcpyAsync(fin, r.counter + 4, sizeof(fin), hipMemcpyDeviceToHost, ctx->stream), "read counter");
    ctx->prof_collect();
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    if (unreached) *unreached = fin[0];
    // the last pass that fixed cells lies in the last epoch if that epoch
    // fixed any, else before the last rebase (the batch ran on past the end)
    if (passes) *passes = fin[1] ? (int32_t)(first + fin[1]) : (int32_t)fin[4];
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_device_alloc(elq_ctx* ctx, size_t bytes, void** out) try {
    if (!ctx || !out) return ELQ_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    HIPCHK(hipMalloc(out, bytes), "hipMalloc");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_device_free(elq_ctx* ctx, void* ptr) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    HIPCHK(hipFree(ptr), "hipFree");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_memcpy_h2d(elq_ctx* ctx, void* dst, const void* src, size_t bytes) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "h2d");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_memcpy_d2h(elq_ctx* ctx, void* dst, const void* src, size_t bytes) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_memcpy_d2d(elq_ctx* ctx, void* dst, const void* src, size_t bytes) try {
    if (!ctx || (bytes & 3) || (((uintptr_t)dst | (uintptr_t)src) & 3)) return ELQ_ERR_INVALID_ARG;
    if (bytes == 0) return ELQ_OK;
    (void)hipSetDevice(ctx->device);
    HIPCHK(launch_copy(dst, src, bytes, ctx->stream), "k_copy");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

void elq_set_profiling(elq_ctx* ctx, int enable) try {
    if (ctx) ctx->profiling = enable != 0;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
}

void elq_set_profiling_filter(elq_ctx* ctx, const char* classes) try {
    if (!ctx) return;
    ctx->prof_filter.clear();
    if (classes && *classes) ctx->prof_filter = "," + std::string(classes) + ",";
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
}

int elq_set_max_groups(elq_ctx* ctx, int max_groups) try {
    if (!ctx) return ELQ_ERR_INVALID_ARG;
    if (max_groups < 0) return ctx->fail(ELQ_ERR_INVALID_ARG, "max_groups must be >= 0");
    ctx->max_groups = max_groups;
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_last_kernel_times(const elq_ctx* ctx, char* names, size_t names_cap, double* ms, int64_t* launches, int cap) try {
    if (!ctx || !names || names_cap == 0) return 0;
    std::string all;
    int n = 0;
    for (const auto& kv : ctx->kt) {
        if (n >= cap) break;
        if (!all.empty()) all += ",";
        all += kv.first;
        if (ms) ms[n] = kv.second.first;
        if (launches) launches[n] = kv.second.second;
        ++n;
    }
    std::strncpy(names, all.c_str(), names_cap - 1);
    names[names_cap - 1] = '\0';
    return n;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return 0;
}

}  // extern "C"
