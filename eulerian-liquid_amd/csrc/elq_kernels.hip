// gfx950 kernels of the pressure projection (ghost-fluid Poisson assembly,
// natural-order IC(0)/MIC(0) preconditioned CG, pressure-gradient update).
//
// Reference path (daichi-nama/eulerian-liquid):
//   assign_indices                 src/project/3d/fluid/projection.cpp:14-28
//   build_matrix                   src/project/3d/fluid/projection.cpp:30-83
//   ConjugateGradient + IC         src/project/3d/eulerian_liquid/simulator.cpp:102-109
//   build_grid_from_vector         src/project/3d/fluid/projection.cpp:85-91
//   update_velocity_using_pressure src/project/3d/fluid/projection.cpp:93-134
//
// Every kernel is HBM-bandwidth or latency bound; none uses MFMA. Built with
// -ffp-contract=off: where a result must match the reference bit for bit
// (indices, A, b, the IC pivots, the velocity update) the float/double
// operation sequence is the reference's; the Krylov kernels use explicit fma().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>
#include <cfloat>
#include <climits>
#include <tuple>
#include <utility>

#include "elq_kernels.hpp"

namespace elq {
namespace {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ uint32_t bits_of(float v) { return __float_as_uint(v); }
__device__ __forceinline__ uint64_t bits_of(double v) { return (uint64_t)__double_as_longlong(v); }

template <class T>
__device__ __forceinline__ T shfl(T v, int src) {
    return __shfl(v, src, 64);
}

// Neighbour moves inside a 4x4x4 brick (lane = lx + 4 ly + 16 lz) on the
// VALU: DPP row shifts for x and y (rows of 16 lanes), permlane swaps for z.
// Each returns the value held by the named neighbour lane; a lane without
// that neighbour inside the brick receives an unspecified value.
//   xm: lane-1 (lx > 0)   ym: lane-4 (ly > 0)   zm: lane-16 (lz > 0)
//   xp: lane+1 (lx < 3)   yp: lane+4 (ly < 3)   zp: lane+16 (lz < 3)
//   q3: lane 4q+3 of the lane's quad   q0: lane 4q of the lane's quad
enum class Mv { xm, xp, ym, yp, zm, zp, q3, q0 };

template <Mv M>
__device__ __forceinline__ uint32_t move_u32(uint32_t v) {
    if constexpr (M == Mv::xm) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x111, 0xF, 0xF, true);  // row_shr:1
    if constexpr (M == Mv::xp) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x101, 0xF, 0xF, true);  // row_shl:1
    if constexpr (M == Mv::ym) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x114, 0xF, 0xF, true);  // row_shr:4
    if constexpr (M == Mv::yp) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x104, 0xF, 0xF, true);  // row_shl:4
    if constexpr (M == Mv::q3) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xFF, 0xF, 0xF, true);  // 3333
    if constexpr (M == Mv::q0) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x00, 0xF, 0xF, true);  // 0000
    if constexpr (M == Mv::zm || M == Mv::zp) {
        // rows r0..r3 of v. a = ([r0 r0 r2 r2], [r1 r1 r3 r3]);
        // b = ([r0 r0 r1 r1], [r2 r2 r3 r3]).
        const auto a = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        const auto b = __builtin_amdgcn_permlane32_swap(a[0], a[1], false, false);
        const int row = (int)(threadIdx.x & 63) >> 4;
        if constexpr (M == Mv::zm) return row == 2 ? b[0] : a[0];  // [-  r0 r1 r2]
        else return row == 1 ? b[1] : a[1];                         // [r1 r2 r3 -]
    }
    return v;
}

template <Mv M, class T>
__device__ __forceinline__ T move(T v) {
    if constexpr (sizeof(T) == 4) {
        return __uint_as_float(move_u32<M>(__float_as_uint(v)));
    } else {
        const uint64_t u = (uint64_t)__double_as_longlong(v);
        const uint32_t lo = move_u32<M>((uint32_t)u), hi = move_u32<M>((uint32_t)(u >> 32));
        return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
    }
}

// The z move of a hyperplane step (lane - 16, or lane + 16 with UP): in FP64
// through the LDS crossbar (ds_bpermute, two per value), which takes no VALU
// issue slots; the two permlane swaps and select of move<zm/zp> per 32-bit
// half cost the FP64 sweeps more (measured 164 against 172 ms per 256^3
// solve), while in FP32, where the step waits on latency rather than issue,
// the crossbar's latency costs more than the swaps (100 against 96 ms).
template <bool UP, class T>
__device__ __forceinline__ T step_zmove(T v, int addr) {
    if constexpr (sizeof(T) == 4) {
        (void)addr;
        return move<UP ? Mv::zp : Mv::zm>(v);
    } else {
        const uint64_t u = (uint64_t)__double_as_longlong(v);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)u);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)(u >> 32));
        return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
    }
}

// One hyperplane step of a brick sweep: b0 - cx v(x) - cy v(y) - cz v(z),
// v(.) the -x/-y/-z neighbour's value (UP: +x/+y/+z). A wave issues one
// instruction per visit of its SIMD, so the step's instruction count sets a
// sweep's pace. FP32: the z neighbour takes two permlane swaps and a select;
// the second swap works in the register of the x neighbour's value, dead once
// its fma has run, so the only copy is the one the first swap needs (its two
// operands are both written). FP64: step_zmove.
template <bool UP, class T>
__device__ __forceinline__ T hyper_step(T v, T b0, T cx, T cy, T cz, int zaddr) {
    if constexpr (sizeof(T) == 4) {
        (void)zaddr;
        const uint32_t u = __float_as_uint(v);
        const T vx = move<UP ? Mv::xp : Mv::xm>(v), vy = move<UP ? Mv::yp : Mv::ym>(v);
        // a = ([r0 r0 r2 r2], [r1 r1 r3 r3]) in rows r0..r3 of v
        const auto a = __builtin_amdgcn_permlane16_swap(u, u, false, false);
        T t = fma(-cx, vx, b0);
        t = fma(-cy, vy, t);
        const int row = (int)(threadIdx.x & 63) >> 4;
        uint32_t z;
        if constexpr (!UP) {
            // [- r0 r1 r2]: rows 1, 3 of a[0]; row 2 = row 0 of a[1], moved by
            // the swap of rows 2, 3 of (vx's register) with rows 0, 1 of a[1]
            const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(vx), a[1], false, false);
            z = row == 2 ? b[0] : a[0];
        } else {
            // [r1 r2 r3 -]: rows 0, 2 of a[1]; row 1 = row 3 of a[0], moved by
            // the swap of rows 2, 3 of a[0] with rows 0, 1 of (vx's register)
            const auto b = __builtin_amdgcn_permlane32_swap(a[0], __float_as_uint(vx), false, false);
            z = row == 1 ? b[1] : a[1];
        }
        return fma(-cz, __uint_as_float(z), t);
    } else {
        const T vx = move<UP ? Mv::xp : Mv::xm>(v), vy = move<UP ? Mv::yp : Mv::ym>(v);
        const T vz = step_zmove<UP>(v, zaddr);
        T t = fma(-cx, vx, b0);
        t = fma(-cy, vy, t);
        return fma(-cz, vz, t);
    }
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int wave_sum_int(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// level_set.h:68-80
__device__ __forceinline__ float fraction_inside2(float l, float r) {
    if (l < 0.0f && r < 0.0f) return 1.0f;
    if (l < 0.0f && r >= 0.0f) return l / (l - r);
    if (l >= 0.0f && r < 0.0f) return r / (r - l);
    return 0.0f;
}

// std::max(theta, 0.01f) as <algorithm> defines it
__device__ __forceinline__ float clamp_theta(float t) { return (t < 0.01f) ? 0.01f : t; }

// Tag generations (the sweep epoch and the reduction tag) advance by one and
// skip 0: exchange and stage words start zeroed, so a generation of 0 would
// accept a word that was never published.
__device__ __forceinline__ uint32_t next_tag(uint32_t t) { return t + 1u == 0u ? 1u : t + 1u; }

__device__ __forceinline__ size_t cell(int i, int j, int k, int ni, int nj) {
    return (size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * (size_t)k);
}

// Wavefront scheduling: persistent waves walk the level-ordered brick list
// (see k_fwd_wave); each wait is bounded in wall-clock time and on expiry the
// wave raises ctrl->abort, after which every other wait returns at once.
// Face exchange between the bricks of one sweep. A brick publishes the values
// of its boundary cells as 64-bit words whose upper half is the sweep's epoch
// (one word per float, two per double), written and polled with relaxed
// device-scope atomics: a reader sees either an old tag or the whole value, so
// no cache-wide acquire/release fence (L2 write-back/invalidate) is needed.
// Word slots: brick b, face f (0 -x .. 5 +z), index idx in [0,16).
__device__ __forceinline__ size_t xslot(int b, int face, int idx) {
    return (size_t)b * kXchWords + (size_t)(face * 16 + idx) * 2;
}

__device__ __forceinline__ void xput_word(uint64_t* w, uint32_t lo, uint32_t e) {
    __hip_atomic_store(w, ((uint64_t)e << 32) | lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void xput(uint64_t* x, size_t s, float v, uint32_t e) {
    xput_word(x + s, __float_as_uint(v), e);
}

__device__ __forceinline__ void xput(uint64_t* x, size_t s, double v, uint32_t e) {
    const uint64_t bits = (uint64_t)__double_as_longlong(v);
    xput_word(x + s, (uint32_t)bits, e);
    xput_word(x + s + 1, (uint32_t)(bits >> 32), e);
}

__device__ __forceinline__ uint32_t xwait(const uint64_t* w, uint32_t e, Ctrl* c, uint64_t timeout) {
    uint64_t v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)(v >> 32) == e) return (uint32_t)v;
    const uint64_t t0 = wall_clock64();
    for (;;) {
        __builtin_amdgcn_s_sleep(1);
        v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((uint32_t)(v >> 32) == e) break;
        if (__hip_atomic_load(&c->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        if (wall_clock64() - t0 > timeout) {
            __hip_atomic_store(&c->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
    }
    return (uint32_t)v;
}

template <class T>
__device__ __forceinline__ T xget(const uint64_t* x, size_t s, uint32_t e, Ctrl* c, uint64_t timeout);

template <>
__device__ __forceinline__ float xget<float>(const uint64_t* x, size_t s, uint32_t e, Ctrl* c, uint64_t timeout) {
    return __uint_as_float(xwait(x + s, e, c, timeout));
}

template <>
__device__ __forceinline__ double xget<double>(const uint64_t* x, size_t s, uint32_t e, Ctrl* c, uint64_t timeout) {
    const uint32_t lo = xwait(x + s, e, c, timeout);
    const uint32_t hi = xwait(x + s + 1, e, c, timeout);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Gathers up to N tagged values at once (slot < 0: none, value 0). All loads
// of a round are issued before any tag is tested, so faces that are already
// published cost a single memory round trip.
template <class T, int N>
__device__ __forceinline__ void xgather(const uint64_t* x, const long (&slot)[N], uint32_t e, T (&out)[N], Ctrl* c,
                                        uint64_t timeout) {
    constexpr int W = sizeof(T) / 4;
    uint32_t lo[N][W];
    bool ok[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
        ok[q] = slot[q] < 0;
#pragma unroll
        for (int h = 0; h < W; ++h) lo[q][h] = 0u;
    }
    uint64_t t0 = 0;
    for (int round = 0;; ++round) {
        bool all = true;
#pragma unroll
        for (int q = 0; q < N; ++q) {
            if (ok[q]) continue;
            uint64_t v[W];
#pragma unroll
            for (int h = 0; h < W; ++h)
                v[h] = __hip_atomic_load(x + slot[q] + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            bool good = true;
#pragma unroll
            for (int h = 0; h < W; ++h) {
                good = good && (uint32_t)(v[h] >> 32) == e;
                lo[q][h] = (uint32_t)v[h];
            }
            ok[q] = good;
            all = all && good;
        }
        if (all) break;
        if (__hip_atomic_load(&c->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        if (round == 0) {
            t0 = wall_clock64();
        } else if (wall_clock64() - t0 > timeout) {
            __hip_atomic_store(&c->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
#pragma unroll
    for (int q = 0; q < N; ++q) {
        if (slot[q] < 0) {
            out[q] = (T)0;
        } else if constexpr (W == 1) {
            out[q] = __uint_as_float(lo[q][0]);
        } else {
            out[q] = __longlong_as_double((long long)(((uint64_t)lo[q][1] << 32) | lo[q][0]));
        }
    }
}

struct Lane {
    int lane, lx, ly, lz;
    __device__ Lane() {
        lane = lane_id();
        lx = lane & 3;
        ly = (lane >> 2) & 3;
        lz = lane >> 4;
    }
};

// Neighbour of a brick cell across a brick face: (brick id, lane) or -1.
// Directions: 0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z.
__device__ __forceinline__ int nbr_id(const Bricks& bk, int d, int b) { return bk.nbr[(size_t)b * 8 + d]; }

// ================================================================== setup
// projection.cpp:8-12 (validity) and :17 (phi < 0); flags unknowns on the
// outer cell layer, where the reference's build_matrix would index outside
// its arrays.
__global__ __launch_bounds__(256) void k_mask(Dense d, uint8_t* mask, int32_t* err) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.ni || j >= d.nj) return;
    const int ni = d.ni, nj = d.nj, nk = d.nk;
    const size_t c = cell(i, j, k, ni, nj);
    const size_t ui = cell(i, j, k, ni + 1, nj);
    const size_t vi = cell(i, j, k, ni, nj + 1);
    const bool valid = d.uw[ui] > 0.0f || d.uw[ui + 1] > 0.0f || d.vw[vi] > 0.0f || d.vw[vi + ni] > 0.0f ||
                       d.ww[c] > 0.0f || d.ww[c + (size_t)ni * nj] > 0.0f;
    const bool fluid = valid && d.phi[c] < 0.0f;
    mask[c] = (uint8_t)((valid ? kValid : 0) | (fluid ? kFluid : 0));
    if (fluid && (i == 0 || j == 0 || k == 0 || i == ni - 1 || j == nj - 1 || k == nk - 1))
        atomicOr(err, kErrBoundary);
}

// Lexicographic exclusive scan of a byte mask (three passes). Reproduces the
// serial k->j->i numbering of assign_indices exactly.
__global__ __launch_bounds__(kScanBlock) void k_scan_count(const uint8_t* f, int64_t n, uint8_t bit,
                                                           int32_t* blk_cnt) {
    __shared__ int ws[kScanBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    int cnt = 0;
#pragma unroll
    for (int m = 0; m < kScanItems; ++m) {
        const int64_t c = base + (int64_t)m * kScanBlock + threadIdx.x;
        if (c < n && (f[c] & bit)) ++cnt;
    }
    cnt = wave_sum_int(cnt);
    if (lane_id() == 0) ws[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < kScanBlock / 64; ++w) s += ws[w];
        blk_cnt[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(1024) void k_scan_blocks(const int32_t* blk_cnt, int64_t g, int32_t* blk_off,
                                                       int64_t* total) {
    __shared__ int64_t sh[1024];
    const int t = threadIdx.x;
    const int64_t per = (g + 1023) / 1024;
    const int64_t lo = t * per, hi = (lo + per < g) ? lo + per : g;
    int64_t s = 0;
    for (int64_t q = lo; q < hi; ++q) s += blk_cnt[q];
    sh[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int64_t v = (t >= o) ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    int64_t run = sh[t] - s;  // exclusive prefix of this thread's segment
    for (int64_t q = lo; q < hi; ++q) {
        blk_off[q] = (int32_t)run;
        run += blk_cnt[q];
    }
    if (t == 1023) *total = sh[1023];
}

__global__ __launch_bounds__(kScanBlock) void k_scan_write(const uint8_t* f, int64_t n, uint8_t bit,
                                                           const int32_t* blk_off, int32_t* out, int32_t fill) {
    __shared__ int ws[kScanBlock / 64];
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    int run = blk_off[blockIdx.x];
    for (int m = 0; m < kScanItems; ++m) {
        const int64_t c = base + (int64_t)m * kScanBlock + threadIdx.x;
        const bool fl = c < n && (f[c] & bit);
        const unsigned long long bal = __ballot(fl);
        const int pre = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) ws[wave] = __popcll(bal);
        __syncthreads();
        int woff = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < kScanBlock / 64; ++w) {
            woff += (w < wave) ? ws[w] : 0;
            tot += ws[w];
        }
        if (c < n) out[c] = fl ? run + woff + pre : fill;
        run += tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_brick_flags(Dense d, const uint8_t* mask, int nbx, int nby, int nbz,
                                                     uint8_t* bflag) {
    const int64_t bi = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (bi >= (int64_t)nbx * nby * nbz) return;
    const int bx = (int)(bi % nbx);
    const int by = (int)((bi / nbx) % nby);
    const int bz = (int)(bi / ((int64_t)nbx * nby));
    uint8_t any = 0;
    for (int lz = 0; lz < kBrick; ++lz)
        for (int ly = 0; ly < kBrick; ++ly)
            for (int lx = 0; lx < kBrick; ++lx) {
                const int i = bx * kBrick + lx, j = by * kBrick + ly, k = bz * kBrick + lz;
                if (i < d.ni && j < d.nj && k < d.nk && (mask[cell(i, j, k, d.ni, d.nj)] & kFluid)) any = 1;
            }
    bflag[bi] = any;
}

__global__ __launch_bounds__(256) void k_brick_meta(const int32_t* map, int nbx, int nby, int nbz,
                                                    int32_t* coord, int32_t* lvl_cnt) {
    const int64_t bi = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (bi >= (int64_t)nbx * nby * nbz) return;
    const int id = map[bi];
    if (id < 0) return;
    const int bx = (int)(bi % nbx);
    const int by = (int)((bi / nbx) % nby);
    const int bz = (int)(bi / ((int64_t)nbx * nby));
    coord[id] = bx | (by << 10) | (bz << 20);
    atomicAdd(&lvl_cnt[bx + by + bz], 1);
}

__global__ void k_level_scan(const int32_t* lvl_cnt, int nlev, int32_t* lvl_off, int32_t* lvl_fill) {
    if (threadIdx.x != 0) return;
    int s = 0;
    for (int l = 0; l < nlev; ++l) {
        lvl_off[l] = s;
        s += lvl_cnt[l];
        lvl_fill[l] = 0;
    }
    lvl_off[nlev] = s;
}

__global__ __launch_bounds__(256) void k_brick_lists(const int32_t* map, const int32_t* coord, int nbx, int nby,
                                                     int nbz, int nb, const int32_t* lvl_off, int32_t* lvl_fill,
                                                     int32_t* lvl_list, int32_t* nbr) {
    const int id = blockIdx.x * 256 + threadIdx.x;
    if (id >= nb) return;
    const int cc = coord[id];
    const int bx = cc & 1023, by = (cc >> 10) & 1023, bz = (cc >> 20) & 1023;
    auto look = [&](int x, int y, int z) -> int {
        if (x < 0 || y < 0 || z < 0 || x >= nbx || y >= nby || z >= nbz) return -1;
        return map[(size_t)x + (size_t)nbx * ((size_t)y + (size_t)nby * z)];
    };
    int32_t* nb8 = nbr + (size_t)id * 8;
    nb8[0] = look(bx - 1, by, bz);
    nb8[1] = look(bx + 1, by, bz);
    nb8[2] = look(bx, by - 1, bz);
    nb8[3] = look(bx, by + 1, bz);
    nb8[4] = look(bx, by, bz - 1);
    nb8[5] = look(bx, by, bz + 1);
    nb8[6] = -1;
    nb8[7] = -1;
    const int l = bx + by + bz;
    const int pos = atomicAdd(&lvl_fill[l], 1);
    lvl_list[lvl_off[l] + pos] = id;
}

// projection.cpp:30-83 for the cells of one brick per wave. Per unknown, the
// neighbours are visited in the reference's order -x,+x,-y,+y,-z,+z; diag is a
// double sum of float terms; b accumulates float products in double and is
// scaled by (double)(dx/dt) with dx/dt a float division done on the host.
// The FP32 copy of the diagonal holds only its ghost part g = diag + (sum of
// the row's off-diagonals) = the sum of the air faces' w / theta terms, 0 in
// the interior, for the FP32 SpMV's difference form.
template <class T>
__global__ __launch_bounds__(256) void k_assemble(Dense d, float scale, const uint8_t* mask, Bricks bk,
                                                  double* diag64, double* b64, T* diag, T* ax, T* ay, T* az) {
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= bk.nb) return;
    const Lane L;
    const int cc = bk.coord[b];
    const int i = (cc & 1023) * kBrick + L.lx, j = ((cc >> 10) & 1023) * kBrick + L.ly,
              k = ((cc >> 20) & 1023) * kBrick + L.lz;
    const int ni = d.ni, nj = d.nj, nk = d.nk;
    const size_t o = (size_t)b * kBrickCells + L.lane;
    double dg = 0.0, gh = 0.0, bb = 0.0;
    float cpl[3] = {0.0f, 0.0f, 0.0f};
    const bool inside = i < ni && j < nj && k < nk;
    const size_t c = inside ? cell(i, j, k, ni, nj) : 0;
    if (inside && (mask[c] & kFluid)) {
        const float phic = d.phi[c];
        const size_t ui = cell(i, j, k, ni + 1, nj), vi = cell(i, j, k, ni, nj + 1);
        const size_t plane = (size_t)ni * nj;
        const float vel[6] = {d.u[ui], d.u[ui + 1], d.v[vi], d.v[vi + ni], d.w[c], d.w[c + plane]};
        const float wgt[6] = {d.uw[ui], d.uw[ui + 1], d.vw[vi], d.vw[vi + ni], d.ww[c], d.ww[c + plane]};
        const int nx[6] = {i - 1, i + 1, i, i, i, i};
        const int ny[6] = {j, j, j - 1, j + 1, j, j};
        const int nz[6] = {k, k, k, k, k - 1, k + 1};
#pragma unroll
        for (int dir = 0; dir < 6; ++dir) {
            const int x = nx[dir], y = ny[dir], z = nz[dir];
            const bool in = x >= 0 && y >= 0 && z >= 0 && x < ni && y < nj && z < nk;
            const size_t cn = in ? cell(x, y, z, ni, nj) : 0;
            const uint8_t mn = in ? mask[cn] : 0;
            if (mn & kValid) {
                if (mn & kFluid) {
                    dg += (double)wgt[dir];
                    if (dir & 1) cpl[dir >> 1] = -wgt[dir];
                } else {
                    const float theta = clamp_theta(fraction_inside2(phic, d.phi[cn]));
                    dg += (double)(wgt[dir] / theta);
                    gh += (double)(wgt[dir] / theta);
                }
            }
            if (dir & 1)
                bb -= (double)(wgt[dir] * vel[dir]);
            else
                bb += (double)(wgt[dir] * vel[dir]);
        }
        bb = bb * (double)scale;
    }
    diag64[o] = dg;
    b64[o] = bb;
    diag[o] = sizeof(T) == 4 ? (T)gh : (T)dg;
    ax[o] = (T)cpl[0];
    ay[o] = (T)cpl[1];
    az[o] = (T)cpl[2];
}

// One cell's pivot inverse from its lower couplings and the lower
// neighbours' pivot inverses, in the oracle's operation order (see below).
__device__ __forceinline__ double pivot_inverse(double dg, double axm, double aym, double azm, double ay_xm,
                                                double az_xm, double ax_ym, double az_ym, double ax_zm, double ay_zm,
                                                double exm, double eym, double ezm, double tau, double sigma,
                                                int& fixed) {
    double t = dg;
    t = t - axm * axm * exm;
    t = t - aym * aym * eym;
    t = t - azm * azm * ezm;
    if (tau != 0.0) {
        double m = 0.0;
        m = m + axm * (ay_xm + az_xm) * exm;
        m = m + aym * (ax_ym + az_ym) * eym;
        m = m + azm * (ax_zm + ay_zm) * ezm;
        t = t - tau * m;
    }
    if (!(t > sigma * dg)) {
        t = dg;
        fixed = 1;
    }
    return 1.0 / t;
}

// Pivot inverse of an interior cell whose lower neighbours hold the same
// value (couplings -1, diagonal 6): the fixed point of pivot_inverse reached
// from 0, which the factorisation reaches away from the -x/-y/-z boundaries
// (0.18350341907227394 for IC(0)). Every caller computes the same value.
__device__ __forceinline__ double interior_pivot(double tau, double sigma) {
    double e = 0.0;
    int fixed = 0;
    for (int k = 0; k < 128; ++k) {
        const double en = pivot_inverse(6.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, e, e, e, tau,
                                        sigma, fixed);
        if (__double_as_longlong(en) == __double_as_longlong(e)) break;
        e = en;
    }
    return e;
}

// Natural-order IC(0)/MIC(0) pivots, one brick per wave, FP64 arithmetic in
// the oracle's operation order (bit-identical pivots):
//   t = d - a_xm^2/e_xm - a_ym^2/e_ym - a_zm^2/e_zm - tau*(...);
//   t <= sigma*d -> t = d;   inv_e = 1/t.
// Inside the brick the ten hyperplanes lx+ly+lz = h are swept in order; pivots
// of the three predecessor bricks arrive through the tagged face exchange.
template <class T, bool LEVELS>
__device__ __forceinline__ void factor_brick(int b, uint32_t epoch, const Bricks& bk, const double* diag64,
                                             const T* ax, const T* ay, const T* az, double tau, double sigma,
                                             double* inv_e64, float* inv_e32, Ctrl* ctrl, uint64_t* xch,
                                             int32_t* n_fixed, uint64_t timeout) {
    const Lane L;
    const size_t o = (size_t)b * kBrickCells + L.lane;
    const int nxm = nbr_id(bk, 0, b), nym = nbr_id(bk, 2, b), nzm = nbr_id(bk, 4, b);
    const double dg = diag64[o];
    const double axo = (double)ax[o], ayo = (double)ay[o], azo = (double)az[o];
    auto ext = [&](const T* a, int nb_, int off) -> double {
        return nb_ >= 0 ? (double)a[(size_t)nb_ * kBrickCells + L.lane + off] : 0.0;
    };
    // lower couplings a(c-x, c) = ax[c-x] etc., and for MIC the other upper
    // couplings of each lower neighbour
    const double sx_ax = shfl(axo, L.lane - 1), sx_ay = shfl(ayo, L.lane - 1), sx_az = shfl(azo, L.lane - 1);
    const double sy_ax = shfl(axo, L.lane - 4), sy_ay = shfl(ayo, L.lane - 4), sy_az = shfl(azo, L.lane - 4);
    const double sz_ax = shfl(axo, L.lane - 16), sz_ay = shfl(ayo, L.lane - 16), sz_az = shfl(azo, L.lane - 16);
    const bool ix = L.lx > 0, iy = L.ly > 0, iz = L.lz > 0;
    const double axm = ix ? sx_ax : ext(ax, nxm, 3);
    const double ay_xm = ix ? sx_ay : ext(ay, nxm, 3);
    const double az_xm = ix ? sx_az : ext(az, nxm, 3);
    const double aym = iy ? sy_ay : ext(ay, nym, 12);
    const double ax_ym = iy ? sy_ax : ext(ax, nym, 12);
    const double az_ym = iy ? sy_az : ext(az, nym, 12);
    const double azm = iz ? sz_az : ext(az, nzm, 48);
    const double ax_zm = iz ? sz_ax : ext(ax, nzm, 48);
    const double ay_zm = iz ? sz_ay : ext(ay, nzm, 48);
    double ex = 0.0, ey = 0.0, ez = 0.0;
    if (LEVELS) {
        if (!ix && nxm >= 0) ex = inv_e64[(size_t)nxm * kBrickCells + L.lane + 3];
        if (!iy && nym >= 0) ey = inv_e64[(size_t)nym * kBrickCells + L.lane + 12];
        if (!iz && nzm >= 0) ez = inv_e64[(size_t)nzm * kBrickCells + L.lane + 48];
    } else {
        const long sl[3] = {(!ix && nxm >= 0) ? (long)xslot(nxm, 1, L.ly + 4 * L.lz) : -1L,
                            (!iy && nym >= 0) ? (long)xslot(nym, 3, L.lx + 4 * L.lz) : -1L,
                            (!iz && nzm >= 0) ? (long)xslot(nzm, 5, L.lx + 4 * L.ly) : -1L};
        double got[3];
        xgather<double, 3>(xch, sl, epoch, got, ctrl, timeout);
        ex = got[0];
        ey = got[1];
        ez = got[2];
    }
    const int h_own = L.lx + L.ly + L.lz;
    double ie = 0.0;
    int fixed = 0;
#pragma unroll
    for (int h = 0; h < 10; ++h) {
        const double vx = shfl(ie, L.lane - 1), vy = shfl(ie, L.lane - 4), vz = shfl(ie, L.lane - 16);
        if (h == h_own && dg > 0.0) {
            const double exm = ix ? vx : ex, eym = iy ? vy : ey, ezm = iz ? vz : ez;
            ie = pivot_inverse(dg, axm, aym, azm, ay_xm, az_xm, ax_ym, az_ym, ax_zm, ay_zm, exm, eym, ezm, tau,
                               sigma, fixed);
        }
    }
    if (!LEVELS) {
        if (L.lx == 3) xput(xch, xslot(b, 1, L.ly + 4 * L.lz), ie, epoch);
        if (L.ly == 3) xput(xch, xslot(b, 3, L.lx + 4 * L.lz), ie, epoch);
        if (L.lz == 3) xput(xch, xslot(b, 5, L.lx + 4 * L.ly), ie, epoch);
    }
    inv_e64[o] = ie;
    if (inv_e32) inv_e32[o] = (float)ie;
    const int nf = wave_sum_int(fixed);
    if (L.lane == 0 && nf) atomicAdd(n_fixed, nf);
}

template <class T>
__global__ __launch_bounds__(64) void k_factor_level(Bricks bk, const double* diag64, const T* ax, const T* ay,
                                                     const T* az, double tau, double sigma, double* inv_e64,
                                                     float* inv_e32, Ctrl* ctrl, uint64_t* xch, int32_t* n_fixed,
                                                     int level_base, uint64_t timeout) {
    factor_brick<T, true>(bk.lvl_list[level_base + (int)blockIdx.x], 0u, bk, diag64, ax, ay, az, tau, sigma,
                          inv_e64, inv_e32, ctrl, xch, n_fixed, timeout);
}

// Persistent wavefront: wave w of W takes positions w, w+W, w+2W, ... of the
// level-ordered brick list. All waves are resident (see persistent()), so a
// brick's predecessors, which sit at smaller positions, always make progress.
template <class T>
__global__ __launch_bounds__(256) void k_factor_wave(Bricks bk, const double* diag64, const T* ax, const T* ay,
                                                     const T* az, double tau, double sigma, double* inv_e64,
                                                     float* inv_e32, Ctrl* ctrl, uint64_t* xch, int32_t* n_fixed,
                                                     uint64_t timeout) {
    const int W = (int)gridDim.x * 4;
    const uint32_t epoch = ctrl->epoch;
    for (int pos = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6); pos < bk.nb; pos += W)
        factor_brick<T, false>(bk.lvl_list[pos], epoch, bk, diag64, ax, ay, az, tau, sigma, inv_e64, inv_e32, ctrl,
                               xch, n_fixed, timeout);
}

template <class T>
__global__ __launch_bounds__(256) void k_export(Dense d, const uint8_t* mask, const int32_t* indices, Bricks bk,
                                                const double* diag64, const double* b64, const T* ax, const T* ay,
                                                const T* az, double* odiag, double* oax, double* oay, double* oaz,
                                                double* ob) {
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= bk.nb) return;
    const Lane L;
    const int cc = bk.coord[b];
    const int i = (cc & 1023) * kBrick + L.lx, j = ((cc >> 10) & 1023) * kBrick + L.ly,
              k = ((cc >> 20) & 1023) * kBrick + L.lz;
    if (i >= d.ni || j >= d.nj || k >= d.nk) return;
    const size_t c = cell(i, j, k, d.ni, d.nj);
    if (!(mask[c] & kFluid)) return;
    const int n = indices[c];
    const size_t o = (size_t)b * kBrickCells + L.lane;
    odiag[n] = diag64[o];
    oax[n] = (double)ax[o];
    oay[n] = (double)ay[o];
    oaz[n] = (double)az[o];
    ob[n] = b64[o];
}

// One FP64 value per brick cell (e.g. the IC pivots) into the reference's
// unknown numbering: out[indices[c]] = v[brick slot of c].
__global__ __launch_bounds__(256) void k_export_cells(Dense d, const uint8_t* mask, const int32_t* indices, Bricks bk,
                                                      const double* v, double* out) {
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= bk.nb) return;
    const Lane L;
    const int cc = bk.coord[b];
    const int i = (cc & 1023) * kBrick + L.lx, j = ((cc >> 10) & 1023) * kBrick + L.ly,
              k = ((cc >> 20) & 1023) * kBrick + L.lz;
    if (i >= d.ni || j >= d.nj || k >= d.nk) return;
    const size_t c = cell(i, j, k, d.ni, d.nj);
    if (!(mask[c] & kFluid)) return;
    out[indices[c]] = v[(size_t)b * kBrickCells + L.lane];
}

// ================================================================== Krylov
// x0 = 0, r = b (Eigen: residual = rhs - mat*x with x = 0), |b|^2 partials.
template <class T>
__global__ __launch_bounds__(256) void k_init(Bricks bk, const double* b64, T* r, T* x, T* p0, T* p1,
                                              double* part) {
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= bk.nb) return;
    const size_t o = (size_t)b * kBrickCells + lane_id();
    const double bv = b64[o];
    r[o] = (T)bv;
    x[o] = (T)0;
    p0[o] = (T)0;
    p1[o] = (T)0;
    const double s = wave_sum(bv * bv);
    if (lane_id() == 0) part[b] = s;
}

// Mixed precision: a residual replacement restarts the search direction when
// the FP64 |r|^2 exceeds the recursive one by more than this factor.
constexpr double kRestartGap2 = 4.0;

// Scalar recurrences of Eigen's conjugate_gradient (see the oracle's orc_pcg)
// and of the mixed-precision residual replacement. Stage 1: workgroup g sums
// the partials of a fixed contiguous segment; stage 2: workgroup 0 gathers the
// segment sums (tagged words, no cache-wide fences) and adds them in segment
// order. The result is deterministic for a given number of partials.
__device__ void scalar_update(Ctrl* c, int mode, double s, double s2) {
    c->epoch = next_tag(c->epoch);
    c->ticket = 0u;
    const bool was_done = c->done != 0;
    switch (mode) {
        case SC_INIT: {
            c->bnorm2 = s;
            c->rr = s;
            c->threshold = fmax(c->tol2 * s, DBL_MIN);
            c->iter = 0;
            c->iter_total = 0;
            c->beta = 0.0;
            c->done = 0;
            c->converged = 0;
            c->status = 0;
            if (s == 0.0) {
                c->done = 1;
                c->converged = 1;
                c->last_err = 0.0;
            } else if (s < c->threshold) {
                c->done = 1;
                c->converged = 1;
                c->last_err = sqrt(c->rr / c->bnorm2);
            } else if (c->max_iter <= 0) {
                // Eigen's loop body never runs: x = 0, iterations() = 0,
                // error() = sqrt(|r|^2 / |b|^2) = 1
                c->done = 1;
                c->last_err = sqrt(c->rr / c->bnorm2);
            }
            break;
        }
        case SC_RZ_INIT:
            if (!was_done) {
                c->rz = s;
                c->beta = 0.0;
                if (!isfinite(s)) {
                    c->done = 1;
                    c->status = 3;
                }
            }
            break;
        case SC_ALPHA:
            if (!was_done) {
                c->pq = s;
                c->alpha = c->rz / s;
                if (!isfinite(c->alpha)) {
                    c->done = 1;
                    c->status = 3;
                    c->last_err = sqrt(c->rr / c->bnorm2);
                }
            }
            break;
        case SC_RR:
            if (!was_done) {
                c->rr = s;
                if (s < c->threshold) {
                    c->done = 1;
                    c->converged = 1;
                    c->rz_div = c->rz;
                    c->last_err = sqrt(s / c->bnorm2);
                } else if (!isfinite(s)) {
                    c->done = 1;
                    c->status = 3;
                }
            }
            break;
        case SC_BETA:
            if (!was_done) {
                const double rz_new = s;
                c->rz_div = c->rz;
                c->beta = rz_new / c->rz;
                c->rz = rz_new;
                c->iter += 1;
                c->iter_total += 1;
                if (!isfinite(c->beta)) {
                    c->done = 1;
                    c->status = 3;
                } else if (c->iter_total >= c->max_iter) {
                    c->done = 1;
                    c->last_err = sqrt(c->rr / c->bnorm2);
                }
            }
            break;
        case SC_REPLACE: {
            if (c->outer == 0) c->bnorm2 = s;  // x64 = 0: r64 = b exactly
            c->rr64 = s;
            const double thr = fmax(c->tol2 * c->bnorm2, DBL_MIN);
            c->outer += 1;
            if (c->bnorm2 == 0.0) {
                c->outer_done = 1;
                c->converged = 1;
                c->done = 1;
                c->last_err = 0.0;
            } else if (s < thr) {
                c->outer_done = 1;
                c->converged = 1;
                c->done = 1;
                c->last_err = sqrt(s / c->bnorm2);
            } else if (c->iter_total >= c->max_iter || c->status != 0 || !isfinite(s)) {
                c->outer_done = 1;
                c->converged = 0;
                c->done = 1;
                c->last_err = sqrt(s / c->bnorm2);
                if (!isfinite(s)) c->status = 3;
            } else {
                c->rr_rec = c->rr;
                c->rr = s2;
                c->rr_ref = s2;
                c->threshold = thr;
                c->done = 0;
                c->converged = 0;
            }
            break;
        }
        case SC_RZ_REPLACE:
            if (!was_done) {
                // Keep the search direction (beta = rz_new / rz_old) while the
                // FP64 residual is close to the recursive one; where it is far
                // above (the recursion ran on past what FP32 can resolve), that
                // beta would make p almost the old direction, so restart.
                c->beta = c->rr64 > kRestartGap2 * c->rr_rec ? 0.0 : s / c->rz_div;
                c->rz = s;
                if (!isfinite(c->beta)) {
                    c->done = 1;
                    c->status = 3;
                }
            }
            break;
        case SC_TRUE:
            c->true_rr = s;
            break;
        default:
            break;
    }
}

// 32-bit words into host-visible memory with system-scope stores (read by
// the host after a stream synchronisation).
__global__ __launch_bounds__(64) void k_export_words(const uint32_t* src, uint32_t* out, int words) {
    for (int q = threadIdx.x; q < words; q += 64) __hip_atomic_store(out + q, src[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ __launch_bounds__(kReduceThreads) void k_scalar(Ctrl* c, int mode, const double* parts, int n,
                                                           const double* parts2, int nblk, uint64_t* stage,
                                                           uint64_t timeout) {
    __shared__ double sh[kReduceThreads];
    __shared__ double sh2[kReduceThreads];
    __shared__ double gs[kRedBlocks], gs2[kRedBlocks];
    const int t = threadIdx.x, g = blockIdx.x;
    const uint32_t tag = next_tag(c->red_epoch);
    const int seg = (n + nblk - 1) / nblk;
    const int lo = g * seg, hi = min(n, lo + seg);
    double s = 0.0, s2 = 0.0;
    if (parts)
        for (int q = lo + t; q < hi; q += kReduceThreads) s += parts[q];
    if (parts2)
        for (int q = lo + t; q < hi; q += kReduceThreads) s2 += parts2[q];
    sh[t] = s;
    sh2[t] = s2;
    __syncthreads();
    for (int o = kReduceThreads / 2; o > 0; o >>= 1) {
        if (t < o) {
            sh[t] += sh[t + o];
            sh2[t] += sh2[t + o];
        }
        __syncthreads();
    }
    if (g != 0) {
        if (t == 0) {
            xput(stage, (size_t)g * 4, sh[0], tag);
            xput(stage, (size_t)g * 4 + 2, sh2[0], tag);
        }
        return;
    }
    if (t == 0) {
        gs[0] = sh[0];
        gs2[0] = sh2[0];
    }
    if (t > 0 && t < nblk) {
        gs[t] = xget<double>(stage, (size_t)t * 4, tag, c, timeout);
        gs2[t] = xget<double>(stage, (size_t)t * 4 + 2, tag, c, timeout);
    }
    __syncthreads();
    if (t != 0) return;
    double S = 0.0, S2 = 0.0;
    for (int q = 0; q < nblk; ++q) {
        S += gs[q];
        S2 += gs2[q];
    }
    c->red_epoch = tag;
    scalar_update(c, mode, S, S2);
}

// Reduction tail of a persistent kernel (workgroup 0 waits only for
// workgroups dispatched after it, which need nothing from it, so its waits
// end). Each workgroup of NW waves adds its wave sums in wave order and
// publishes the result as a tagged word pair in stage[2g]; workgroup 0
// gathers the G sums (thread t takes g = t, t+64*NW, ... in that order, then
// a fixed tree) and applies the scalar recurrence. The result depends only on
// the grid size, not on timing. Every thread of every workgroup must call
// this; ctrl must not be read after it.
// g, G: this workgroup's index among the G workgroups of the reduction
// (the launch's blockIdx.x and gridDim.x, or a request's share of a batched
// launch).
template <int NW>
__device__ void grid_tail_n(Ctrl* c, int mode, double wave_total, uint32_t tag, uint64_t* stage, uint64_t timeout,
                            int g, int G) {
    constexpr int NT = 64 * NW;
    __shared__ double sw[NW];
    __shared__ double st[NT];
    const int t = threadIdx.x;
    if ((t & 63) == 0) sw[t >> 6] = wave_total;
    __syncthreads();
    double own = 0.0;
    if (t == 0)
        for (int w = 0; w < NW; ++w) own += sw[w];
    if (g != 0) {
        if (t == 0) xput(stage, (size_t)g * 2, own, tag);
        return;
    }
    double s = own;
    for (int k0 = 0; k0 * NT < G; k0 += 4) {
        long slot[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int gi = t + NT * (k0 + k);
            slot[k] = (gi > 0 && gi < G) ? (long)gi * 2 : -1;
        }
        double v[4];
        xgather<double, 4>(stage, slot, tag, v, c, timeout);
#pragma unroll
        for (int k = 0; k < 4; ++k) s += v[k];
    }
    st[t] = s;
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1) {
        if (t < o) st[t] += st[t + o];
        __syncthreads();
    }
    if (t == 0) {
        c->red_epoch = tag;
        scalar_update(c, mode, st[0], 0.0);
    }
}

__device__ __forceinline__ void grid_tail(Ctrl* c, int mode, double wave_total, uint32_t tag, uint64_t* stage,
                                          uint64_t timeout, int g, int G) {
    grid_tail_n<4>(c, mode, wave_total, tag, stage, timeout, g, G);
}

// The six neighbour brick ids of brick b (wave-uniform): lanes 0..5 load
// them in one access, readlane broadcasts them.
struct Nbr6 {
    int n[6];
};
__device__ __forceinline__ Nbr6 load_nbr6(const Bricks& bk, int b) {
    const int lane = lane_id();
    const int v = lane < 6 ? bk.nbr[(size_t)b * 8 + lane] : -1;
    Nbr6 r;
#pragma unroll
    for (int d = 0; d < 6; ++d) r.n[d] = __builtin_amdgcn_readlane(v, d);
    return r;
}

// Offset of a lane's neighbour cell across face d of brick b, or the lane's
// own offset when there is none (so loads stay unconditional); *has tells.
__device__ __forceinline__ size_t halo_offset(const Nbr6& nb, int d, size_t own, const int (&lc)[3], int lane,
                                              bool& has) {
    const int axis = d >> 1, up = d & 1;
    const bool face = up ? lc[axis] == 3 : lc[axis] == 0;
    const int shift = axis == 0 ? 3 : axis == 1 ? 12 : 48;  // lane offset to the opposite face
    has = face && nb.n[d] >= 0;
    return has ? (size_t)nb.n[d] * kBrickCells + (up ? lane - shift : lane + shift) : own;
}

// p = z + beta p_old (ping-pong buffers), q = A p, p.q partials. Cross-brick
// neighbour values of p are rebuilt from z and p_old; in-brick ones come
// through lane shuffles. All loads are issued unconditionally and at once.
// Accumulation order: diag (FP32: its ghost part), -x, +x, -y, +y, -z, +z.
// The FP32 SpMV issues two bricks' loads before either brick's arithmetic
// (spmv_load, spmv_compute); FP64 runs one brick after the other in one
// function (spmv_brick), whose code measured 4% faster than the same steps
// through the split form. Both take the same loads in the same order.
template <class T>
struct SpmvIn {
    size_t o;
    bool has[6];
    T zc, pi, dg, axo, ayo, azo, axe, aye, aze;
    T ze[6], pe[6];
};

// Issues every load of brick b (nothing is consumed here).
template <class T>
__device__ __forceinline__ void spmv_load(SpmvIn<T>& d, int b, const Bricks& bk, const Coef<T>& c,
                                          const T* __restrict__ z, const T* __restrict__ p_in, const Lane& L) {
    const Nbr6 nb = load_nbr6(bk, b);
    const size_t o = (size_t)b * kBrickCells + L.lane;
    d.o = o;
    const int lc[3] = {L.lx, L.ly, L.lz};
    size_t e[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) e[k] = halo_offset(nb, k, o, lc, L.lane, d.has[k]);
    // a uniform brick reads its coefficients (and the -x/-y/-z face
    // couplings, which equal them) from the 64-entry constant rows
    const bool uni = (c.uflag[b] & 1) != 0;
    const size_t oc = uni ? (size_t)L.lane : o;
    const T *cd = uni ? c.urow + kBrickCells : c.diag, *cx = uni ? c.urow : c.ax, *cy = uni ? c.urow : c.ay,
            *cz = uni ? c.urow : c.az;
    d.zc = z[o];
    d.pi = p_in[o];
    d.dg = cd[oc];
    d.axo = cx[oc];
    d.ayo = cy[oc];
    d.azo = cz[oc];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        d.ze[k] = z[e[k]];
        d.pe[k] = p_in[e[k]];
    }
    d.axe = cx[uni ? oc : e[0]];
    d.aye = cy[uni ? oc : e[2]];
    d.aze = cz[uni ? oc : e[4]];
}

template <class T>
__device__ __forceinline__ double spmv_compute(const SpmvIn<T>& d, T* __restrict__ p_out, T* __restrict__ q,
                                               T beta, const Lane& L) {
    static_assert(sizeof(T) == 4, "the split SpMV brick is the FP32 one");
    const T pc = fma(beta, d.pi, d.zc);
    p_out[d.o] = pc;
    T pn[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) pn[k] = d.has[k] ? fma(beta, d.pe[k], d.ze[k]) : (T)0;
    const T sxm = move<Mv::xm>(pc), sxp = move<Mv::xp>(pc);
    const T sym = move<Mv::ym>(pc), syp = move<Mv::yp>(pc);
    const T szm = move<Mv::zm>(pc), szp = move<Mv::zp>(pc);
    const T cxm = move<Mv::xm>(d.axo), cym = move<Mv::ym>(d.ayo), czm = move<Mv::zm>(d.azo);
    const T p_xm = L.lx > 0 ? sxm : pn[0];
    const T p_xp = L.lx < 3 ? sxp : pn[1];
    const T p_ym = L.ly > 0 ? sym : pn[2];
    const T p_yp = L.ly < 3 ? syp : pn[3];
    const T p_zm = L.lz > 0 ? szm : pn[4];
    const T p_zp = L.lz < 3 ? szp : pn[5];
    const T axm = L.lx > 0 ? cxm : (d.has[0] ? d.axe : (T)0);
    const T aym = L.ly > 0 ? cym : (d.has[2] ? d.aye : (T)0);
    const T azm = L.lz > 0 ? czm : (d.has[4] ? d.aze : (T)0);
    // difference form g p_i + sum a_ij (p_j - p_i), g the ghost part of the
    // diagonal: for a smooth p the sum of diag p_i and the off-diagonal terms
    // cancels to a small value that FP32 would round away, while the
    // differences are exact for close values
    T acc = d.dg * pc;
    acc = fma(axm, p_xm - pc, acc);
    acc = fma(d.axo, p_xp - pc, acc);
    acc = fma(aym, p_ym - pc, acc);
    acc = fma(d.ayo, p_yp - pc, acc);
    acc = fma(azm, p_zm - pc, acc);
    acc = fma(d.azo, p_zp - pc, acc);
    q[d.o] = acc;
    return (double)pc * (double)acc;
}

template <class T>
__device__ __forceinline__ double spmv_brick(int b, const Bricks& bk, const Coef<T>& c, const T* __restrict__ z,
                                             const T* __restrict__ p_in, T* __restrict__ p_out, T* __restrict__ q,
                                             T beta, const Lane& L) {
    static_assert(sizeof(T) == 8, "the one-function SpMV brick is the FP64 one");
    const Nbr6 nb = load_nbr6(bk, b);
    const size_t o = (size_t)b * kBrickCells + L.lane;
    const int lc[3] = {L.lx, L.ly, L.lz};
    bool has[6];
    size_t e[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) e[d] = halo_offset(nb, d, o, lc, L.lane, has[d]);
    // a uniform brick reads its coefficients (and the -x/-y/-z face
    // couplings, which equal them) from the 64-entry constant rows
    const bool uni = (c.uflag[b] & 1) != 0;
    const size_t oc = uni ? (size_t)L.lane : o;
    const T *cd = uni ? c.urow + kBrickCells : c.diag, *cx = uni ? c.urow : c.ax, *cy = uni ? c.urow : c.ay,
            *cz = uni ? c.urow : c.az;
    const T zc = z[o], pi = p_in[o], dg = cd[oc], axo = cx[oc], ayo = cy[oc], azo = cz[oc];
    T ze[6], pe[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        ze[d] = z[e[d]];
        pe[d] = p_in[e[d]];
    }
    const T axe = cx[uni ? oc : e[0]], aye = cy[uni ? oc : e[2]], aze = cz[uni ? oc : e[4]];
    const T pc = fma(beta, pi, zc);
    p_out[o] = pc;
    T pn[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) pn[d] = has[d] ? fma(beta, pe[d], ze[d]) : (T)0;
    const T sxm = move<Mv::xm>(pc), sxp = move<Mv::xp>(pc);
    const T sym = move<Mv::ym>(pc), syp = move<Mv::yp>(pc);
    const T szm = move<Mv::zm>(pc), szp = move<Mv::zp>(pc);
    const T cxm = move<Mv::xm>(axo), cym = move<Mv::ym>(ayo), czm = move<Mv::zm>(azo);
    const T p_xm = L.lx > 0 ? sxm : pn[0];
    const T p_xp = L.lx < 3 ? sxp : pn[1];
    const T p_ym = L.ly > 0 ? sym : pn[2];
    const T p_yp = L.ly < 3 ? syp : pn[3];
    const T p_zm = L.lz > 0 ? szm : pn[4];
    const T p_zp = L.lz < 3 ? szp : pn[5];
    const T axm = L.lx > 0 ? cxm : (has[0] ? axe : (T)0);
    const T aym = L.ly > 0 ? cym : (has[2] ? aye : (T)0);
    const T azm = L.lz > 0 ? czm : (has[4] ? aze : (T)0);
    T acc = dg * pc;
    acc = fma(axm, p_xm, acc);
    acc = fma(axo, p_xp, acc);
    acc = fma(aym, p_ym, acc);
    acc = fma(ayo, p_yp, acc);
    acc = fma(azm, p_zm, acc);
    acc = fma(azo, p_zp, acc);
    q[o] = acc;
    return (double)pc * (double)acc;
}

// Persistent: wave w of W takes bricks w, w+W, ..., two per loop step; its
// p.q terms accumulate per lane in brick order; the grid tail computes alpha.
// Workgroup g of G (see grid_tail_n).
template <class T>
__device__ __forceinline__ void spmv_body(int g, int G, const Bricks& bk, const Coef<T>& c, const T* z,
                                          const T* p_in, T* p_out, T* q, Ctrl* ctrl, uint64_t* stage,
                                          uint64_t timeout) {
    if (ctrl->done) return;
    const uint32_t tag = next_tag(ctrl->red_epoch);
    const T beta = (T)ctrl->beta;
    const Lane L;
    const int W = G * 4;
    double acc = 0.0;
    const int b0 = __builtin_amdgcn_readfirstlane(g * 4 + (int)(threadIdx.x >> 6));
    if constexpr (sizeof(T) == 4) {
        // both bricks' loads go out before either brick's arithmetic (without
        // a second brick the first is loaded twice and the copy is dropped):
        // 77 VGPRs, six workgroups per CU
        for (int b = b0; b < bk.nb; b += 2 * W) {
            const bool two = b + W < bk.nb;
            SpmvIn<T> A, B;
            spmv_load<T>(A, b, bk, c, z, p_in, L);
            spmv_load<T>(B, two ? b + W : b, bk, c, z, p_in, L);
            acc += spmv_compute<T>(A, p_out, q, beta, L);
            if (two) acc += spmv_compute<T>(B, p_out, q, beta, L);
        }
    } else {
        // FP64 would need 85 VGPRs for that and measured slower with the five
        // workgroups per CU they leave: one brick after the other
        for (int b = b0; b < bk.nb; b += 2 * W) {
            acc += spmv_brick<T>(b, bk, c, z, p_in, p_out, q, beta, L);
            if (b + W < bk.nb) acc += spmv_brick<T>(b + W, bk, c, z, p_in, p_out, q, beta, L);
        }
    }
    grid_tail(ctrl, SC_ALPHA, wave_sum(acc), tag, stage, timeout, g, G);
}

template <class T>
__global__ __launch_bounds__(256) void k_spmv(Bricks bk, Coef<T> c, const T* z, const T* p_in, T* p_out, T* q,
                                              Ctrl* ctrl, uint64_t* stage, uint64_t timeout) {
    spmv_body<T>((int)blockIdx.x, (int)gridDim.x, bk, c, z, p_in, p_out, q, ctrl, stage, timeout);
}

// Forward sweep of M^{-1} = (E+L^T)^{-1} E (E+L)^{-1}: y = (r - L y) / e, one
// brick per wave, hyperplanes lx+ly+lz = 0..9 inside the brick. With UPDATE,
// fused ahead of it: x += alpha p, r -= alpha q, |r|^2 partials.
template <class T, bool UPDATE, bool LEVELS>
__device__ __forceinline__ void fwd_brick(int b, uint32_t epoch, const Bricks& bk, const Coef<T>& c, T* r, T* x,
                                          const T* p, const T* q, T* y, double* part, Ctrl* ctrl, uint64_t* xch,
                                          uint64_t timeout) {
    const Lane L;
    const size_t o = (size_t)b * kBrickCells + L.lane;
    T rv = r[o];
    if (UPDATE) {
        const T alpha = (T)ctrl->alpha;
        const T xv = fma(alpha, p[o], x[o]);
        rv = fma(-alpha, q[o], rv);
        x[o] = xv;
        r[o] = rv;
        const double s = wave_sum((double)rv * (double)rv);
        if (L.lane == 0) part[b] = s;
    }
    const T ie = c.inv_e[o];
    const int nxm = nbr_id(bk, 0, b), nym = nbr_id(bk, 2, b), nzm = nbr_id(bk, 4, b);
    const T sx = shfl(c.ax[o], L.lane - 1), sy = shfl(c.ay[o], L.lane - 4), sz = shfl(c.az[o], L.lane - 16);
    const bool ix = L.lx > 0, iy = L.ly > 0, iz = L.lz > 0;
    const T axm = ix ? sx : (nxm >= 0 ? c.ax[(size_t)nxm * kBrickCells + L.lane + 3] : (T)0);
    const T aym = iy ? sy : (nym >= 0 ? c.ay[(size_t)nym * kBrickCells + L.lane + 12] : (T)0);
    const T azm = iz ? sz : (nzm >= 0 ? c.az[(size_t)nzm * kBrickCells + L.lane + 48] : (T)0);
    T ex = (T)0, ey = (T)0, ez = (T)0;
    if (LEVELS) {
        if (!ix && nxm >= 0) ex = y[(size_t)nxm * kBrickCells + L.lane + 3];
        if (!iy && nym >= 0) ey = y[(size_t)nym * kBrickCells + L.lane + 12];
        if (!iz && nzm >= 0) ez = y[(size_t)nzm * kBrickCells + L.lane + 48];
    } else {
        if (!ix && nxm >= 0) ex = xget<T>(xch, xslot(nxm, 1, L.ly + 4 * L.lz), epoch, ctrl, timeout);
        if (!iy && nym >= 0) ey = xget<T>(xch, xslot(nym, 3, L.lx + 4 * L.lz), epoch, ctrl, timeout);
        if (!iz && nzm >= 0) ez = xget<T>(xch, xslot(nzm, 5, L.lx + 4 * L.ly), epoch, ctrl, timeout);
    }
    const int h_own = L.lx + L.ly + L.lz;
    T yv = (T)0;
#pragma unroll
    for (int h = 0; h < 10; ++h) {
        const T vx = shfl(yv, L.lane - 1), vy = shfl(yv, L.lane - 4), vz = shfl(yv, L.lane - 16);
        if (h == h_own) {
            T t = fma(-axm, ix ? vx : ex, rv);
            t = fma(-aym, iy ? vy : ey, t);
            t = fma(-azm, iz ? vz : ez, t);
            yv = t * ie;
        }
    }
    if (!LEVELS) {
        if (L.lx == 3) xput(xch, xslot(b, 1, L.ly + 4 * L.lz), yv, epoch);
        if (L.ly == 3) xput(xch, xslot(b, 3, L.lx + 4 * L.lz), yv, epoch);
        if (L.lz == 3) xput(xch, xslot(b, 5, L.lx + 4 * L.ly), yv, epoch);
    }
    y[o] = yv;
}

template <class T, bool UPDATE>
__global__ __launch_bounds__(64) void k_fwd_level(Bricks bk, Coef<T> c, T* r, T* x, const T* p, const T* q, T* y,
                                                  double* part, Ctrl* ctrl, uint64_t* xch, int level_base,
                                                  uint64_t timeout) {
    if (ctrl->done) return;
    fwd_brick<T, UPDATE, true>(bk.lvl_list[level_base + (int)blockIdx.x], 0u, bk, c, r, x, p, q, y, part, ctrl, xch,
                               timeout);
}

// Backward sweep z = y - E^{-1} L^T z, in place over y, hyperplanes 9..0;
// fused r.z partials.
template <class T, bool LEVELS>
__device__ __forceinline__ void bwd_brick(int b, uint32_t epoch, const Bricks& bk, const Coef<T>& c, const T* r,
                                          T* y, double* part, Ctrl* ctrl, uint64_t* xch, uint64_t timeout) {
    const Lane L;
    const size_t o = (size_t)b * kBrickCells + L.lane;
    const T yv = y[o];
    const T ie = c.inv_e[o];
    const T axo = c.ax[o], ayo = c.ay[o], azo = c.az[o];
    const T rv = r[o];
    const int nxp = nbr_id(bk, 1, b), nyp = nbr_id(bk, 3, b), nzp = nbr_id(bk, 5, b);
    const bool ix = L.lx < 3, iy = L.ly < 3, iz = L.lz < 3;
    T ex = (T)0, ey = (T)0, ez = (T)0;
    if (LEVELS) {
        if (!ix && nxp >= 0) ex = y[(size_t)nxp * kBrickCells + L.lane - 3];
        if (!iy && nyp >= 0) ey = y[(size_t)nyp * kBrickCells + L.lane - 12];
        if (!iz && nzp >= 0) ez = y[(size_t)nzp * kBrickCells + L.lane - 48];
    } else {
        if (!ix && nxp >= 0) ex = xget<T>(xch, xslot(nxp, 0, L.ly + 4 * L.lz), epoch, ctrl, timeout);
        if (!iy && nyp >= 0) ey = xget<T>(xch, xslot(nyp, 2, L.lx + 4 * L.lz), epoch, ctrl, timeout);
        if (!iz && nzp >= 0) ez = xget<T>(xch, xslot(nzp, 4, L.lx + 4 * L.ly), epoch, ctrl, timeout);
    }
    const int h_own = L.lx + L.ly + L.lz;
    T zv = (T)0;
#pragma unroll
    for (int h = 9; h >= 0; --h) {
        const T vx = shfl(zv, L.lane + 1), vy = shfl(zv, L.lane + 4), vz = shfl(zv, L.lane + 16);
        if (h == h_own) {
            T t = axo * (ix ? vx : ex);
            t = fma(ayo, iy ? vy : ey, t);
            t = fma(azo, iz ? vz : ez, t);
            zv = fma(-ie, t, yv);
        }
    }
    if (!LEVELS) {
        if (L.lx == 0) xput(xch, xslot(b, 0, L.ly + 4 * L.lz), zv, epoch);
        if (L.ly == 0) xput(xch, xslot(b, 2, L.lx + 4 * L.lz), zv, epoch);
        if (L.lz == 0) xput(xch, xslot(b, 4, L.lx + 4 * L.ly), zv, epoch);
    }
    y[o] = zv;
    const double s = wave_sum((double)rv * (double)zv);
    if (L.lane == 0) part[b] = s;
}

template <class T>
__global__ __launch_bounds__(64) void k_bwd_level(Bricks bk, Coef<T> c, const T* r, T* y, double* part, Ctrl* ctrl,
                                                  uint64_t* xch, int level_base, uint64_t timeout) {
    if (ctrl->done) return;
    bwd_brick<T, true>(bk.lvl_list[level_base + (int)blockIdx.x], 0u, bk, c, r, y, part, ctrl, xch, timeout);
}

// Walks the active bricks of one pencil in x order (REVERSE: descending),
// 64 map entries per chunk held one per lane, together with the map rows of
// the two neighbouring pencils the sweep waits on (-y/-z forward, +y/+z
// backward) and the pencil's row of uniform-brick flags, so neighbour brick
// ids and flags cost no memory round trip per brick. (The flags come by
// position, loaded beside the map row: looking them up by brick id would put
// a second dependent round trip at the start of every pencil, which measured
// 10-13% slower sweeps.)
template <bool REVERSE>
struct PencilIter {
    const int32_t* row;
    const int32_t* row_y;  // nullptr when the neighbour pencil is outside the grid
    const int32_t* row_z;
    const int32_t* urow_;  // uniform flags by brick position (Coef::umap), this pencil's row
    int nbx, base;
    int mv, my, mz, mu;
    int u = 0;  // uniform flag of the brick last returned by next()
    unsigned long long mask;
    __device__ void load_chunk() {
        const int bx = base + lane_id();
        const bool in = bx >= 0 && bx < nbx;
        mv = in ? row[bx] : -1;
        my = (in && row_y) ? row_y[bx] : -1;
        mz = (in && row_z) ? row_z[bx] : -1;
        mu = in ? urow_[bx] : 0;
        mask = __ballot(mv >= 0);
    }
    __device__ PencilIter(const int32_t* r, const int32_t* ry, const int32_t* rz, int n, const int32_t* ur)
        : row(r), row_y(ry), row_z(rz), urow_(ur), nbx(n) {
        base = REVERSE ? ((n - 1) / 64) * 64 : 0;
        load_chunk();
    }
    // brick id, its x index and the ids of its y/z neighbours (-1 if none);
    // once the pencil is exhausted it returns false and leaves the outputs
    // unchanged
    __device__ bool next(int& bx, int& b, int& ny, int& nz) {
        while (mask == 0ull) {
            if (REVERSE) {
                if (base == 0) return false;
                base -= 64;
            } else {
                base += 64;
                if (base >= nbx) return false;
            }
            load_chunk();
        }
        const int l = REVERSE ? 63 - __builtin_clzll(mask) : __builtin_ctzll(mask);
        mask &= ~(1ull << l);
        bx = base + l;
        b = __builtin_amdgcn_readlane(mv, l);
        ny = __builtin_amdgcn_readlane(my, l);
        nz = __builtin_amdgcn_readlane(mz, l);
        u = __builtin_amdgcn_readlane(mu, l);
        return true;
    }
};

__device__ __forceinline__ const int32_t* map_row(const Bricks& bk, int by, int bz) {
    if (by < 0 || bz < 0 || by >= bk.nby || bz >= bk.nbz) return nullptr;
    return bk.map + (size_t)bk.nbx * ((size_t)by + (size_t)bk.nby * bz);
}

// Tagged words of a whole brick (pencil sweeps): word (lane * W + h) of
// brick b's slot, W = words per value. Consumers read the face cells they
// need from it; every lane stores and loads unconditionally (branch-free
// memory operations keep the compiler's wait counters exact).
template <class T>
__device__ __forceinline__ size_t bslot(int b, int lane) {
    return (size_t)b * kXchWords + (size_t)lane * (sizeof(T) / 4);
}

// Two tagged values per lane (the y and z neighbour cells), requested with
// unconditional loads (unused requests point at a valid dummy word and
// count as present). issue() starts a round; finish() completes it and
// re-polls, all words at once, until every requested tag matches.
constexpr int kPollSleep = 1;  // s_sleep units (64 clocks) between poll rounds

template <class T>
struct Polls2 {
    static constexpr int W = sizeof(T) / 4;
    size_t slot[2];
    bool want[2];
    uint64_t v[2][W];
    __device__ __forceinline__ void load(const uint64_t* x) {
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int h = 0; h < W; ++h)
                v[q][h] = __hip_atomic_load(x + slot[q] + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ bool ready(uint32_t e) const {
        bool ok = true;
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int h = 0; h < W; ++h) ok = ok && (!want[q] || (uint32_t)(v[q][h] >> 32) == e);
        return ok;
    }
    __device__ __forceinline__ void issue(const uint64_t* x) { load(x); }
    __device__ __forceinline__ void finish(const uint64_t* x, uint32_t e, T (&out)[2], Ctrl* c, uint64_t timeout) {
        if (!__all(ready(e))) {
            const uint64_t t0 = wall_clock64();
            for (int round = 1;; ++round) {
                __builtin_amdgcn_s_sleep(kPollSleep);
                load(x);
                if (__all(ready(e))) break;
                if ((round & 63) == 0) {
                    if (__hip_atomic_load(&c->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                    if (wall_clock64() - t0 > timeout) {
                        __hip_atomic_store(&c->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            T val;
            if constexpr (W == 1) {
                val = __uint_as_float((uint32_t)v[q][0]);
            } else {
                val = __longlong_as_double((long long)(((uint64_t)(uint32_t)v[q][1] << 32) | (uint32_t)v[q][0]));
            }
            out[q] = want[q] ? val : (T)0;
        }
    }
};

template <class T>
__device__ __forceinline__ void bput(uint64_t* x, int b, T v, uint32_t e, const Lane& L) {
    xput(x, bslot<T>(b, L.lane), v, e);
}

// Wavefront sweeps over x-pencils. A wave walks the bricks of one pencil in x
// order (reverse order for the backward sweep). The x neighbour of a brick is
// the wave's previous brick: its face values come from the registers through
// a lane move. y and z neighbours belong to other pencils and arrive as
// epoch-tagged words (or through LDS inside a pencil block, see k_fwd_block).
// Every brick a wave waits for belongs to a pencil of lower (backward:
// higher) level, i.e. an earlier position of the block list, and all waves
// are resident (see persistent()), so every wait ends.
template <class T, bool UPDATE>
struct FwdIn {
    int b, bx, ny, nz;
    T rv, xv, pv, qv, ie, axo, ayo, azo, aym_e, azm_e;
    T lx, ly, lz;  // FP64: pivot-scaled lower couplings (Coef::lx, ly, lz)
    Polls2<T> poll;
};

// Issues the loads of one brick (nothing is consumed here). Every load is
// unconditional; requests a lane does not need read a valid dummy address.
// UNI: uniform bricks read their couplings from the constant row (FP32 1x4
// blocks; the wider FP32 blocks spill registers with it at 128 VGPRs and
// measured a 36% slower forward sweep at 512^3).
template <class T, bool UPDATE, bool UNI = true>
__device__ __forceinline__ void fwd_fetch(FwdIn<T, UPDATE>& d, int b, int bx, int nym, int nzm, int u,
                                          const Coef<T>& c, const T* r, const T* x, const T* p, const T* q,
                                          const Lane& L, bool gy = true, bool gz = true) {
    d.b = b;
    d.bx = bx;
    d.ny = nym;
    d.nz = nzm;
    const size_t o = (size_t)b * kBrickCells + L.lane;
    d.rv = r[o];
    if (UPDATE) {
        d.xv = x[o];
        d.pv = p[o];
        d.qv = q[o];
    }
    const bool wy = L.ly == 0 && nym >= 0, wz = L.lz == 0 && nzm >= 0;
    const int ly_lane = wy ? L.lane + 12 : L.lane, lz_lane = wz ? L.lane + 48 : L.lane;
    const int yb = wy ? nym : b, zb = wz ? nzm : b;
    if constexpr (sizeof(T) == 8) {
        // bricks with interior coefficients and pivots: 1/e and the scaled
        // lower couplings from the constant rows E and -E
        const bool up = (u & 2) != 0;
        const size_t oc = up ? (size_t)L.lane : o;
        d.ie = (up ? c.urow + 2 * kBrickCells : c.inv_e)[oc];
        d.lx = (up ? c.urow + 3 * kBrickCells : c.lx)[oc];
        d.ly = (up ? c.urow + 3 * kBrickCells : c.ly)[oc];
        d.lz = (up ? c.urow + 3 * kBrickCells : c.lz)[oc];
    } else {
        d.ie = c.inv_e[o];
        // uniform brick: couplings, -y/-z face ones included, from the constant row
        const bool uni = UNI && (u & 1) != 0;
        const T *cx = uni ? c.urow : c.ax, *cy = uni ? c.urow : c.ay, *cz = uni ? c.urow : c.az;
        const size_t oc = uni ? (size_t)L.lane : o;
        d.axo = cx[oc];
        d.ayo = cy[oc];
        d.azo = cz[oc];
        const T ay_e = cy[uni ? oc : (size_t)yb * kBrickCells + ly_lane];
        const T az_e = cz[uni ? oc : (size_t)zb * kBrickCells + lz_lane];
        d.aym_e = wy ? ay_e : (T)0;
        d.azm_e = wz ? az_e : (T)0;
    }
    d.poll.want[0] = wy && gy;
    d.poll.want[1] = wz && gz;
    if constexpr (sizeof(T) == 8) {
        // FP64 (bricks publish their faces only): every lane polls a word of
        // the face it waits for (ly := 3, lz := 3 of the neighbour), so a poll
        // round touches only the face's cache lines; with no word to wait
        // for, all lanes read one word of their own brick. (In FP32, where
        // bricks publish all 64 words, lanes that wait for nothing reading
        // their own brick's words measured faster.)
        d.poll.slot[0] = nym >= 0 && gy ? bslot<T>(nym, L.lane | 12) : bslot<T>(b, 0);
        d.poll.slot[1] = nzm >= 0 && gz ? bslot<T>(nzm, L.lane | 48) : bslot<T>(b, 0);
    } else {
        d.poll.slot[0] = bslot<T>(yb, ly_lane);
        d.poll.slot[1] = bslot<T>(zb, lz_lane);
    }
}

template <class T>
struct FwdCarry {
    T prev_y, prev_ax;
    int prev_bx;
    double acc;  // this lane's |r|^2 terms, in the wave's brick order
};

// Part of a forward brick that needs no neighbour face: in-brick lower
// couplings, the x-neighbour values carried from the previous brick and, with
// UPDATE, the x and r updates and the brick's |r|^2 terms.
template <class T, bool UPDATE>
struct FwdPre {
    T axm, aym, azm, ex, rv;
    bool adj;
};

template <class T, bool UPDATE>
__device__ __forceinline__ FwdPre<T, UPDATE> fwd_pre(const FwdIn<T, UPDATE>& cur, FwdCarry<T>& cy, T alpha,
                                                     T* r, T* x, const Lane& L) {
    FwdPre<T, UPDATE> pr;
    const bool ix = L.lx > 0, iy = L.ly > 0, iz = L.lz > 0;
    const size_t o = (size_t)cur.b * kBrickCells + L.lane;
    T rv = cur.rv;
    if (UPDATE) {
        rv = fma(-alpha, cur.qv, rv);
        x[o] = fma(alpha, cur.pv, cur.xv);
        r[o] = rv;
        cy.acc += (double)rv * (double)rv;
    }
    pr.rv = rv;
    pr.adj = cy.prev_bx == cur.bx - 1;
    if constexpr (sizeof(T) == 4) {
        const T sx = move<Mv::xm>(cur.axo), sxp = move<Mv::q3>(cy.prev_ax);
        const T sy = move<Mv::ym>(cur.ayo), sz = move<Mv::zm>(cur.azo);
        pr.axm = ix ? sx : (pr.adj ? sxp : (T)0);
        pr.aym = iy ? sy : cur.aym_e;
        pr.azm = iz ? sz : cur.azm_e;
    }
    const T pyx = move<Mv::q3>(cy.prev_y);
    pr.ex = pr.adj ? pyx : (T)0;
    return pr;
}

// Forward brick: the ten hyperplanes, then publication.
template <class T, bool UPDATE>
__device__ __forceinline__ T fwd_compute(const FwdIn<T, UPDATE>& cur, const FwdPre<T, UPDATE>& pr,
                                         const T (&got)[2], FwdCarry<T>& cy, T* y, uint64_t* xch, uint32_t epoch,
                                         const Lane& L, bool publish = true) {
    const bool ix = L.lx > 0, iy = L.ly > 0, iz = L.lz > 0;
    // Values from outside the brick (previous brick, -y/-z faces) are fixed
    // for the whole brick: they go into the right-hand side first. The
    // in-brick couplings are scaled by the pivot, so each hyperplane step is
    // three dependent fmas. Every lane evaluates every step; a lane's value is
    // final from its own hyperplane lx+ly+lz on (its inputs are final by then
    // and it keeps recomputing the same value), so no per-step select is needed.
    // FP32 builds the scaled couplings here from ax, ay, az, which the SpMV
    // reads too: three more arrays per solve would push the working set of an
    // iteration out of the 256 MB last-level cache at 256^3. FP64, far beyond
    // that cache anyway, reads them precomputed (Coef::lx, ly, lz).
    T b0, cx, cyy, cz;
    if constexpr (sizeof(T) == 8) {
        b0 = pr.rv * cur.ie;
        b0 = ix ? b0 : fma(-cur.lx, pr.ex, b0);
        b0 = iy ? b0 : fma(-cur.ly, got[0], b0);
        b0 = iz ? b0 : fma(-cur.lz, got[1], b0);
        cx = ix ? cur.lx : (T)0;
        cyy = iy ? cur.ly : (T)0;
        cz = iz ? cur.lz : (T)0;
    } else {
        T rhs = pr.rv;
        rhs = ix ? rhs : fma(-pr.axm, pr.ex, rhs);
        rhs = iy ? rhs : fma(-pr.aym, got[0], rhs);
        rhs = iz ? rhs : fma(-pr.azm, got[1], rhs);
        b0 = rhs * cur.ie;
        cx = ix ? pr.axm * cur.ie : (T)0;
        cyy = iy ? pr.aym * cur.ie : (T)0;
        cz = iz ? pr.azm * cur.ie : (T)0;
    }
    T yv = (T)0;
    const int zaddr = ((L.lane - 16) & 63) << 2;  // ds_bpermute byte address of lane - 16
#pragma unroll
    for (int h = 0; h < 10; ++h) yv = hyper_step<false>(yv, b0, cx, cyy, cz, zaddr);
    if (publish) bput<T>(xch, cur.b, yv, epoch, L);
    y[(size_t)cur.b * kBrickCells + L.lane] = yv;
    cy.prev_y = yv;
    cy.prev_ax = cur.axo;
    cy.prev_bx = cur.bx;
    return yv;
}

template <class T>
struct BwdIn {
    int b, bx, ny, nz;
    T yv, ie, axo, ayo, azo, rv;
    Polls2<T> poll;
};

template <class T>
__device__ __forceinline__ void bwd_fetch(BwdIn<T>& d, int b, int bx, int nyp, int nzp, int u, const Coef<T>& c,
                                          const T* r, const T* y, const Lane& L, bool gy = true, bool gz = true) {
    d.b = b;
    d.bx = bx;
    d.ny = nyp;
    d.nz = nzp;
    const size_t o = (size_t)b * kBrickCells + L.lane;
    d.yv = y[o];
    const bool up = sizeof(T) == 8 && (u & 2) != 0;
    d.ie = (up ? c.urow + 2 * kBrickCells : c.inv_e)[up ? (size_t)L.lane : o];
    const bool uni = (u & 1) != 0;
    const size_t oc = uni ? (size_t)L.lane : o;
    d.axo = (uni ? c.urow : c.ax)[oc];
    d.ayo = (uni ? c.urow : c.ay)[oc];
    d.azo = (uni ? c.urow : c.az)[oc];
    d.rv = r[o];
    const bool wy = L.ly == 3 && nyp >= 0, wz = L.lz == 3 && nzp >= 0;
    d.poll.want[0] = wy && gy;
    d.poll.want[1] = wz && gz;
    if constexpr (sizeof(T) == 8) {
        // faces ly := 0 and lz := 0 of the +y and +z neighbours (see fwd_fetch)
        d.poll.slot[0] = nyp >= 0 && gy ? bslot<T>(nyp, L.lane & ~12) : bslot<T>(b, 0);
        d.poll.slot[1] = nzp >= 0 && gz ? bslot<T>(nzp, L.lane & ~48) : bslot<T>(b, 0);
    } else {
        d.poll.slot[0] = bslot<T>(wy ? nyp : b, wy ? L.lane - 12 : L.lane);
        d.poll.slot[1] = bslot<T>(wz ? nzp : b, wz ? L.lane - 48 : L.lane);
    }
}

template <class T>
struct BwdCarry {
    T prev_z;
    int prev_bx;
    double acc;  // this lane's r.z terms, in the wave's brick order
};

template <class T>
__device__ __forceinline__ T bwd_compute(const BwdIn<T>& cur, T ex, const T (&got)[2], BwdCarry<T>& cy, T* y,
                                         uint64_t* xch, uint32_t epoch, const Lane& L, bool publish = true) {
    const bool ix = L.lx < 3, iy = L.ly < 3, iz = L.lz < 3;
    // As in fwd_compute: outside values into the right-hand side, couplings
    // scaled by the pivot, every lane evaluated at every step.
    const T cx = cur.ie * cur.axo, cyy = cur.ie * cur.ayo, cz = cur.ie * cur.azo;
    T rhs = cur.yv;
    rhs = ix ? rhs : fma(-cx, ex, rhs);
    rhs = iy ? rhs : fma(-cyy, got[0], rhs);
    rhs = iz ? rhs : fma(-cz, got[1], rhs);
    const T kx = ix ? cx : (T)0, ky = iy ? cyy : (T)0, kz = iz ? cz : (T)0;
    T zv = (T)0;
    const int zaddr = ((L.lane + 16) & 63) << 2;  // ds_bpermute byte address of lane + 16
#pragma unroll
    for (int h = 9; h >= 0; --h) zv = hyper_step<true>(zv, rhs, kx, ky, kz, zaddr);
    if (publish) bput<T>(xch, cur.b, zv, epoch, L);
    y[(size_t)cur.b * kBrickCells + L.lane] = zv;
    cy.prev_z = zv;
    cy.prev_bx = cur.bx;
    cy.acc += (double)cur.rv * (double)zv;
    return zv;
}

// ---------------------------------------------------------------- pencil blocks
// A workgroup of BY x BZ waves sweeps a block of BY x BZ neighbouring
// x-pencils, wave (dy, dz) = wave dy + BY dz taking pencil (BY Y + dy,
// BZ Z + dz). Faces between pencils of the same block go through LDS rings
// (one per producer wave and direction) instead of the device-wide exchange
// buffer; only faces that leave the block are published there. Blocks are
// taken in the order of a list sorted by block level Y + Z (then Y): block
// position p goes to workgroup p mod G in round p / G, so a workgroup only
// waits for blocks at earlier positions.
//
// Ring protocol. Entry key = round * 4096 + x position + 1 (x position: bx,
// or 4094 - bx for the backward sweep, so keys grow in processing order);
// every LDS word carries (key << 32 | 32 value bits), written and read as
// single 64-bit accesses, so a reader sees a whole entry or an old one and
// needs no fence. Slot = key mod R. A consumer announces, before it waits for
// the faces of its brick with key k, that it needs nothing below k (prog =
// k - 1), and at the end of a round that it needs nothing more from it. A
// producer writes key k only once the consumer's prog >= k - R: slots of keys
// in (prog, prog + R] are distinct, so an entry is never overwritten before
// it has been read, and the producer can always write the key the consumer
// waits for (no cycle between the two waits).
template <class T>
struct Ring {
    static constexpr int W = sizeof(T) / 4;
    static constexpr int R = 8 / W;            // slots
    static constexpr int kWords = R * 16 * W;  // words per ring
};

__device__ __forceinline__ uint32_t ring_key(int round, int xpos) {
    return (uint32_t)round * 4096u + (uint32_t)xpos + 1u;
}
__device__ __forceinline__ uint32_t ring_end(int round) { return (uint32_t)round * 4096u + 4095u; }

__device__ __forceinline__ void lds_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint64_t lds_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void prog_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t prog_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Producer: waits (bounded) until the consumer no longer needs the slot of
// key, then the lanes with on_face store their face value at index f.
template <class T>
__device__ __forceinline__ void ring_put(uint64_t* ring, const uint32_t* cons_prog, uint32_t& seen, uint32_t key,
                                         bool on_face, int f, T v, Ctrl* c, uint64_t timeout) {
    constexpr int W = Ring<T>::W, R = Ring<T>::R;
    if (seen + R < key) {
        uint64_t t0 = 0;
        for (int round = 0;; ++round) {
            seen = prog_load(cons_prog);
            if (seen + R >= key) break;
            if ((round & 63) == 0) {
                if (round == 0) {
                    t0 = wall_clock64();
                } else {
                    if (__hip_atomic_load(&c->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                    if (wall_clock64() - t0 > timeout) {
                        __hip_atomic_store(&c->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (on_face) {
        uint64_t* slot = ring + (size_t)(key % R) * (16 * W) + (size_t)f * W;
        if constexpr (W == 1) {
            lds_store(slot, ((uint64_t)key << 32) | __float_as_uint(v));
        } else {
            const uint64_t bits = (uint64_t)__double_as_longlong(v);
            lds_store(slot, ((uint64_t)key << 32) | (uint32_t)bits);
            lds_store(slot + 1, ((uint64_t)key << 32) | (uint32_t)(bits >> 32));
        }
    }
}

// Consumer: every lane reads entry f of the slot of key (lanes that do not
// want it read the same words and accept anything) until all wanted words
// carry key (bounded wait).
template <class T>
__device__ __forceinline__ T ring_get(const uint64_t* ring, uint32_t key, bool want, int f, Ctrl* c,
                                      uint64_t timeout) {
    constexpr int W = Ring<T>::W, R = Ring<T>::R;
    const uint64_t* slot = ring + (size_t)(key % R) * (16 * W) + (size_t)f * W;
    uint64_t v[W];
    uint64_t t0 = 0;
    for (int round = 0;; ++round) {
        bool ok = true;
#pragma unroll
        for (int h = 0; h < W; ++h) {
            v[h] = lds_load(slot + h);
            ok = ok && (uint32_t)(v[h] >> 32) == key;
        }
        if (__all(ok || !want)) break;
        if ((round & 63) == 0) {
            if (round == 0) {
                t0 = wall_clock64();
            } else {
                if (__hip_atomic_load(&c->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                if (wall_clock64() - t0 > timeout) {
                    __hip_atomic_store(&c->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __builtin_amdgcn_s_sleep(1);
    }
    if (!want) return (T)0;
    if constexpr (W == 1) {
        return __uint_as_float((uint32_t)v[0]);
    } else {
        return __longlong_as_double((long long)(((uint64_t)(uint32_t)v[1] << 32) | (uint32_t)v[0]));
    }
}

// Shared state of a block sweep: 2 rings per wave (y and z faces it
// produces) and one progress word per wave (as a consumer).
template <class T, int NW>
struct BlockLds {
    uint64_t ring[2 * NW][Ring<T>::kWords];
    uint32_t prog[NW];
};

template <class T, int NW>
__device__ __forceinline__ void block_lds_init(BlockLds<T, NW>& s) {
    uint64_t* w = &s.ring[0][0];
    for (int q = threadIdx.x; q < 2 * NW * Ring<T>::kWords; q += 64 * NW) w[q] = 0ull;
    if (threadIdx.x < NW) s.prog[threadIdx.x] = 0u;
    __syncthreads();
}

// Forward sweep over pencil blocks. Per pencil, a wave walks its bricks in x
// order. Brick loads are issued two bricks ahead (three buffers in rotation,
// so prefetched registers are never copied: a copy would wait for the loads)
// and the -y/-z face polls one brick ahead. Per brick: everything that needs
// no neighbour face (coupling moves, x carry, x and r updates with UPDATE) is
// done first, then the polls are finished (faces inside the block come from
// the LDS rings), the next brick's polls are issued and the hyperplanes run.
// Past the pencil's last brick the fetch repeats the current brick (valid
// addresses, results unused), which keeps the loop free of memory branches.
// The wave's |r|^2 terms go to the grid tail (SC_RR with UPDATE; without
// UPDATE the tail only advances the epochs, SC_NOP).
// At most 128 VGPRs in FP32 (4 waves per SIMD: 4 workgroups of 4 waves per CU).
// Issue priority of a sweep wave from its block's place in the current
// round of the block list: blocks early in the list feed the ones after them.
__device__ __forceinline__ void sweep_priority(int pos, int round, int G) {
    const int q = (4 * (pos - round * G)) / G;
    if (q <= 0) __builtin_amdgcn_s_setprio(3);
    else if (q == 1) __builtin_amdgcn_s_setprio(2);
    else if (q == 2) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

// Workgroup g of G (see grid_tail_n).
// epoch: the exchange-word tag of this sweep; tag: its reduction tag (the
// caller reads both, and ctrl->done, before the sweep starts). UNI: uniform
// bricks read their couplings from the constant row (see fwd_fetch).
template <class T, bool UPDATE, int BY, int BZ, bool UNI = BY == 1>
__device__ __forceinline__ void fwd_block_body(int g, int G, const Bricks& bk, const Pencils& pc, const Coef<T>& c,
                                               T* r, T* x, const T* p, const T* q, T* y, Ctrl* ctrl, uint64_t* xch,
                                               uint64_t* stage, uint64_t timeout, uint32_t epoch, uint32_t tag) {
    constexpr int NW = BY * BZ;
    static_assert((NW & (NW - 1)) == 0, "grid_tail_n reduces a power-of-two number of waves");
    __shared__ BlockLds<T, NW> S;
    block_lds_init<T, NW>(S);
    const Lane L;
    const int w = (int)(threadIdx.x >> 6), dy = w % BY, dz = w / BY;
    const bool gy = dy == 0, gz = dz == 0;                    // -y/-z neighbour outside the block
    const bool out_y = dy == BY - 1, out_z = dz == BZ - 1;   // +y/+z consumer outside the block
    // FP64: only the cells other workgroups read are published (the +y face,
    // ly = 3, and the +z face, lz = 3, where their consumer is outside the
    // block); FP32: the whole brick, one word per lane (measured faster at
    // one word per value, slower at two)
    const bool publish = sizeof(T) == 8 ? (out_y && L.ly == 3) || (out_z && L.lz == 3) : out_y || out_z;
    uint64_t* ring_y = S.ring[2 * w];                          // this wave's +y face, to wave (dy+1, dz)
    uint64_t* ring_z = S.ring[2 * w + 1];                      // this wave's +z face, to wave (dy, dz+1)
    const uint64_t* in_y = gy ? S.ring[0] : S.ring[2 * (w - 1)];
    const uint64_t* in_z = gz ? S.ring[0] : S.ring[2 * (w - BY) + 1];
    const uint32_t* prog_y = &S.prog[out_y ? w : w + 1];
    const uint32_t* prog_z = &S.prog[out_z ? w : w + BY];
    uint32_t seen_y = 0, seen_z = 0;
    const int f_y = L.lx + 4 * L.lz, f_z = L.lx + 4 * L.ly;   // face index (producer and consumer side)
    const T alpha = UPDATE ? (T)ctrl->alpha : (T)0;
    double acc = 0.0;
    int round = 0;
    for (int pos = g; pos < pc.nblk; pos += G, ++round) {
        const int bc = pc.blk_list[pos];
        sweep_priority(pos, round, G);
        const int by = (bc & 1023) * BY + dy, bz = ((bc >> 10) & 1023) * BZ + dz;
        if (by < bk.nby && bz < bk.nbz) {
            PencilIter<false> it(map_row(bk, by, bz), map_row(bk, by - 1, bz), map_row(bk, by, bz - 1), bk.nbx,
                                 c.umap + (map_row(bk, by, bz) - bk.map));
            int bx, b, ny, nz;
            if (it.next(bx, b, ny, nz)) {
                FwdIn<T, UPDATE> A, B, C;
                FwdCarry<T> cy{(T)0, (T)0, -2, acc};
                fwd_fetch<T, UPDATE, UNI>(A, b, bx, ny, nz, it.u, c, r, x, p, q, L, gy, gz);
                A.poll.issue(xch);
                T got[2];
                auto step = [&](FwdIn<T, UPDATE>& cur, FwdIn<T, UPDATE>& nxt) {
                    const FwdPre<T, UPDATE> pr = fwd_pre<T, UPDATE>(cur, cy, alpha, r, x, L);
                    const uint32_t key = ring_key(round, cur.bx);
                    if (L.lane == 0) prog_store(&S.prog[w], key - 1u);
                    cur.poll.finish(xch, epoch, got, ctrl, timeout);
                    nxt.poll.issue(xch);
                    if (!gy) got[0] = ring_get<T>(in_y, key, L.ly == 0 && cur.ny >= 0, f_y, ctrl, timeout);
                    if (!gz) got[1] = ring_get<T>(in_z, key, L.lz == 0 && cur.nz >= 0, f_z, ctrl, timeout);
                    const T yv = fwd_compute<T, UPDATE>(cur, pr, got, cy, y, xch, epoch, L, publish);
                    if (!out_y) ring_put<T>(ring_y, prog_y, seen_y, key, L.ly == 3, f_y, yv, ctrl, timeout);
                    if (!out_z) ring_put<T>(ring_z, prog_z, seen_z, key, L.lz == 3, f_z, yv, ctrl, timeout);
                };
                // three buffers in rotation: a brick's loads are issued two
                // bricks ahead of its computation, its first polls one brick
                // ahead (earlier polls would mostly find stale tags)
                auto ahead = [&](FwdIn<T, UPDATE>& buf, bool prev) {
                    const bool has = prev && it.next(bx, b, ny, nz);
                    fwd_fetch<T, UPDATE, UNI>(buf, b, bx, ny, nz, it.u, c, r, x, p, q, L, gy, gz);
                    return has;
                };
                bool hB = ahead(B, true);
                for (;;) {
                    const bool hC = ahead(C, hB);
                    step(A, B);
                    if (!hB) break;
                    const bool hA = ahead(A, hC);
                    step(B, C);
                    if (!hC) break;
                    hB = ahead(B, hA);
                    step(C, A);
                    if (!hA) break;
                }
                acc = cy.acc;
            }
        }
        if (L.lane == 0) prog_store(&S.prog[w], ring_end(round));
    }
    grid_tail_n<NW>(ctrl, UPDATE ? SC_RR : SC_NOP, UPDATE ? wave_sum(acc) : 0.0, tag, stage, timeout, g, G);
}

template <class T, bool UPDATE, int BY, int BZ>
__global__ __launch_bounds__(64 * BY * BZ, sizeof(T) == 4 ? 4 : 2) void k_fwd_block(Bricks bk, Pencils pc, Coef<T> c,
                                                                                   T* r, T* x,
                                                            const T* p, const T* q, T* y, Ctrl* ctrl, uint64_t* xch,
                                                            uint64_t* stage, uint64_t timeout) {
    if (ctrl->done) return;
    fwd_block_body<T, UPDATE, BY, BZ>((int)blockIdx.x, (int)gridDim.x, bk, pc, c, r, x, p, q, y, ctrl, xch, stage,
                                      timeout, ctrl->epoch, next_tag(ctrl->red_epoch));
}

// Backward sweep over pencil blocks: blocks in reverse list order, bricks in
// descending x, +y/+z faces from LDS inside the block.
template <class T, int BY, int BZ>
__device__ __forceinline__ void bwd_block_body(int g, int G, const Bricks& bk, const Pencils& pc, const Coef<T>& c,
                                               const T* r, T* y, Ctrl* ctrl, uint64_t* xch, uint64_t* stage,
                                               int tail_mode, uint64_t timeout, uint32_t epoch, uint32_t tag) {
    constexpr int NW = BY * BZ;
    static_assert((NW & (NW - 1)) == 0, "grid_tail_n reduces a power-of-two number of waves");
    __shared__ BlockLds<T, NW> S;
    block_lds_init<T, NW>(S);
    const Lane L;
    const int w = (int)(threadIdx.x >> 6), dy = w % BY, dz = w / BY;
    const bool gy = dy == BY - 1, gz = dz == BZ - 1;   // +y/+z neighbour outside the block
    const bool out_y = dy == 0, out_z = dz == 0;       // -y/-z consumer outside the block
    // the faces other workgroups read (FP64), the whole brick (FP32), as in k_fwd_block
    const bool publish = sizeof(T) == 8 ? (out_y && L.ly == 0) || (out_z && L.lz == 0) : out_y || out_z;
    uint64_t* ring_y = S.ring[2 * w];                   // this wave's -y face, to wave (dy-1, dz)
    uint64_t* ring_z = S.ring[2 * w + 1];               // this wave's -z face, to wave (dy, dz-1)
    const uint64_t* in_y = gy ? S.ring[0] : S.ring[2 * (w + 1)];
    const uint64_t* in_z = gz ? S.ring[0] : S.ring[2 * (w + BY) + 1];
    const uint32_t* prog_y = &S.prog[out_y ? w : w - 1];
    const uint32_t* prog_z = &S.prog[out_z ? w : w - BY];
    uint32_t seen_y = 0, seen_z = 0;
    const int f_y = L.lx + 4 * L.lz, f_z = L.lx + 4 * L.ly;
    double acc = 0.0;
    int round = 0;
    for (int pos = g; pos < pc.nblk; pos += G, ++round) {
        const int bc = pc.blk_list[pc.nblk - 1 - pos];
        sweep_priority(pos, round, G);
        const int by = (bc & 1023) * BY + dy, bz = ((bc >> 10) & 1023) * BZ + dz;
        if (by < bk.nby && bz < bk.nbz) {
            PencilIter<true> it(map_row(bk, by, bz), map_row(bk, by + 1, bz), map_row(bk, by, bz + 1), bk.nbx,
                                c.umap + (map_row(bk, by, bz) - bk.map));
            int bx, b, ny, nz;
            if (it.next(bx, b, ny, nz)) {
                BwdIn<T> A, B, C;
                BwdCarry<T> cy{(T)0, -2, acc};
                bwd_fetch<T>(A, b, bx, ny, nz, it.u, c, r, y, L, gy, gz);
                A.poll.issue(xch);
                T got[2];
                auto step = [&](BwdIn<T>& cur, BwdIn<T>& nxt) {
                    const T pzx = move<Mv::q0>(cy.prev_z);
                    const T ex = cy.prev_bx == cur.bx + 1 ? pzx : (T)0;
                    const uint32_t key = ring_key(round, 4094 - cur.bx);
                    if (L.lane == 0) prog_store(&S.prog[w], key - 1u);
                    cur.poll.finish(xch, epoch, got, ctrl, timeout);
                    nxt.poll.issue(xch);
                    if (!gy) got[0] = ring_get<T>(in_y, key, L.ly == 3 && cur.ny >= 0, f_y, ctrl, timeout);
                    if (!gz) got[1] = ring_get<T>(in_z, key, L.lz == 3 && cur.nz >= 0, f_z, ctrl, timeout);
                    const T zv = bwd_compute<T>(cur, ex, got, cy, y, xch, epoch, L, publish);
                    if (!out_y) ring_put<T>(ring_y, prog_y, seen_y, key, L.ly == 0, f_y, zv, ctrl, timeout);
                    if (!out_z) ring_put<T>(ring_z, prog_z, seen_z, key, L.lz == 0, f_z, zv, ctrl, timeout);
                };
                auto ahead = [&](BwdIn<T>& buf, bool prev) {
                    const bool has = prev && it.next(bx, b, ny, nz);
                    bwd_fetch<T>(buf, b, bx, ny, nz, it.u, c, r, y, L, gy, gz);
                    return has;
                };
                bool hB = ahead(B, true);
                for (;;) {
                    const bool hC = ahead(C, hB);
                    step(A, B);
                    if (!hB) break;
                    const bool hA = ahead(A, hC);
                    step(B, C);
                    if (!hC) break;
                    hB = ahead(B, hA);
                    step(C, A);
                    if (!hA) break;
                }
                acc = cy.acc;
            }
        }
        if (L.lane == 0) prog_store(&S.prog[w], ring_end(round));
    }
    grid_tail_n<NW>(ctrl, tail_mode, wave_sum(acc), tag, stage, timeout, g, G);
}

template <class T, int BY, int BZ>
__global__ __launch_bounds__(64 * BY * BZ, sizeof(T) == 4 ? 4 : 2) void k_bwd_block(Bricks bk, Pencils pc, Coef<T> c,
                                                                                   const T* r, T* y,
                                                            Ctrl* ctrl, uint64_t* xch, uint64_t* stage, int tail_mode,
                                                            uint64_t timeout) {
    if (ctrl->done) return;
    bwd_block_body<T, BY, BZ>((int)blockIdx.x, (int)gridDim.x, bk, pc, c, r, y, ctrl, xch, stage, tail_mode, timeout,
                              ctrl->epoch, next_tag(ctrl->red_epoch));
}

// Forward sweep with the x, r updates, then the backward sweep, in one
// persistent launch (FP32, small grids): a workgroup starts its backward
// blocks as soon as its forward blocks are done, without the kernel boundary
// and the forward grid tail in between.
// - The backward phase of workgroup g takes its own forward blocks in reverse
//   order: it is the backward kernel's workgroup gb = (nblk - 1 - g) mod G,
//   whose list positions are exactly those. So a brick's backward step runs
//   in the wave that did its forward step (y and r are that wave's own
//   stores), and after every brick that consumed its forward face has
//   finished its forward step (the backward step waits for their backward
//   faces, which follow it), so its backward exchange words never replace
//   forward words still waited for. (Taking other workgroups' blocks breaks
//   both: a backward step could read a y not yet written, and a late
//   forward store could replace a backward word a consumer waits for.)
// - Tags: the backward ones follow the forward ones (next_tag), as the two
//   launches would read them, so nothing is read from ctrl after the forward
//   tail (which updates it).
// - The backward tail uses the second half of the stage buffer (a workgroup
//   may publish its backward partial before the forward ones are gathered)
//   and slot gb, with the blocks of the backward kernel's workgroup gb in
//   the same order: the sums are the two launches' bit for bit.
// - Waits end: the forward phase waits only on forward words; a workgroup
//   publishes its forward partial before its backward blocks, which wait
//   only on backward words of blocks of higher level.
// - After a convergence found by the forward tail the backward sweep still
//   runs; it only rewrites y, and its tail leaves the state alone.
template <int BY, int BZ>
__global__ __launch_bounds__(64 * BY * BZ, 4) void k_sweeps_block(Bricks bk, Pencils pc, Coef<float> c, float* r,
                                                                  float* x, const float* p, const float* q, float* y,
                                                                  Ctrl* ctrl, uint64_t* xch, uint64_t* stage,
                                                                  uint64_t timeout) {
    if (ctrl->done) return;
    const uint32_t epoch = ctrl->epoch, tag = next_tag(ctrl->red_epoch);
    const int g = (int)blockIdx.x, G = (int)gridDim.x;
    // the forward phase without the constant-row selects: with them the
    // combined kernel needs more than 128 VGPRs
    fwd_block_body<float, true, BY, BZ, false>(g, G, bk, pc, c, r, x, p, q, y, ctrl, xch, stage, timeout, epoch, tag);
    const int gb = ((pc.nblk - 1 - g) % G + G) % G;
    bwd_block_body<float, BY, BZ>(gb, G, bk, pc, c, r, y, ctrl, xch, stage + 2 * kTailMaxGroups, SC_BETA, timeout,
                                  next_tag(epoch), next_tag(tag));
}

// Uniform bricks (one wave per brick): every cell is an unknown whose
// diagonal entry (FP32: its ghost part) and six couplings, the -x/-y/-z ones
// read across the brick faces, equal the constants cref and dref bit for bit,
// as in the interior of the fluid (full faces: coupling -1, diagonal 6, ghost
// part 0). Flag bit 1 marks those; in FP64 bit 2 marks the ones whose pivot
// inverses also all equal the interior value E (interior_pivot, the same
// arithmetic as the factorisation). Workgroup 0 writes the constant rows
// urow[0, 64) = cref, [64, 128) = dref, [128, 192) = E, [192, 256) = -E (the
// pre-scaled lower coupling -1 * E), from which the SpMV and the sweeps read
// a flagged brick's coefficients (same values, a few cache lines instead of
// the arrays).
template <class T>
__global__ __launch_bounds__(256) void k_brick_uniform(Bricks bk, const T* diag, const T* ax, const T* ay,
                                                       const T* az, T cref, T dref, const double* ie64, double tau,
                                                       double sigma, int32_t* uflag, int32_t* umap, T* urow) {
    const double e_int = sizeof(T) == 8 ? interior_pivot(tau, sigma) : 0.0;
    if (blockIdx.x == 0) {
        const int t = (int)threadIdx.x, row = t >> 6;
        urow[t] = row == 0 ? cref : row == 1 ? dref : row == 2 ? (T)e_int : (T)(-e_int);
    }
    const int b = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (b >= bk.nb) return;
    const Lane L;
    const size_t o = (size_t)b * kBrickCells + L.lane;
    const int nx = nbr_id(bk, 0, b), ny = nbr_id(bk, 2, b), nz = nbr_id(bk, 4, b);
    T axm = (T)0, aym = (T)0, azm = (T)0;
    if (L.lx > 0)
        axm = ax[o - 1];
    else if (nx >= 0)
        axm = ax[(size_t)nx * kBrickCells + L.lane + 3];
    if (L.ly > 0)
        aym = ay[o - 4];
    else if (ny >= 0)
        aym = ay[(size_t)ny * kBrickCells + L.lane + 12];
    if (L.lz > 0)
        azm = az[o - 16];
    else if (nz >= 0)
        azm = az[(size_t)nz * kBrickCells + L.lane + 48];
    const auto same = [](T a, T v) { return bits_of(a) == bits_of(v); };
    const bool ok = same(diag[o], dref) && same(ax[o], cref) && same(ay[o], cref) && same(az[o], cref) &&
                    same(axm, cref) && same(aym, cref) && same(azm, cref);
    int flag = __all(ok) ? 1 : 0;
    if constexpr (sizeof(T) == 8) {
        if (flag && __all(bits_of(ie64[o]) == bits_of(e_int))) flag |= 2;
    }
    if (L.lane == 0) {
        const int cc = bk.coord[b];
        const size_t bx = (size_t)(cc & 1023), by = (size_t)((cc >> 10) & 1023), bz = (size_t)((cc >> 20) & 1023);
        const size_t pos = bx + (size_t)bk.nbx * (by + (size_t)bk.nby * bz);
        uflag[b] = flag;
        umap[pos] = flag;
    }
}

// Pivot-scaled lower couplings of the FP64 forward sweep (one wave per
// brick): lx = (coupling of the -x neighbour to this cell) * (this cell's
// 1/e), the neighbour's +x coupling read across the brick face where it lies
// in the -x brick, 0 without a -x neighbour; likewise ly, lz.
__global__ __launch_bounds__(256) void k_scale_lower(Bricks bk, const double* ax, const double* ay,
                                                     const double* az, const double* ie, double* lx, double* ly,
                                                     double* lz) {
    const int b = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (b >= bk.nb) return;
    const Lane L;
    const size_t o = (size_t)b * kBrickCells + L.lane;
    const int nx = nbr_id(bk, 0, b), ny = nbr_id(bk, 2, b), nz = nbr_id(bk, 4, b);
    double axm = 0.0, aym = 0.0, azm = 0.0;
    if (L.lx > 0)
        axm = ax[o - 1];
    else if (nx >= 0)
        axm = ax[(size_t)nx * kBrickCells + L.lane + 3];
    if (L.ly > 0)
        aym = ay[o - 4];
    else if (ny >= 0)
        aym = ay[(size_t)ny * kBrickCells + L.lane + 12];
    if (L.lz > 0)
        azm = az[o - 16];
    else if (nz >= 0)
        azm = az[(size_t)nz * kBrickCells + L.lane + 48];
    const double e = ie[o];
    lx[o] = axm * e;
    ly[o] = aym * e;
    lz[o] = azm * e;
}

// Lists of pencil blocks, one workgroup per block shape: the blocks (BY x BZ
// pencils) holding at least one active pencil, sorted by block level Y + Z,
// then Y (deterministic). Workgroup s writes flag + base[s], list + base[s]
// and count[s]. Levels: NBY + NBZ - 1 <= 2045 (nby, nbz <= 1023).
constexpr int kBlockLevels = 2048;
struct ShapeLists {
    int by[kShapeCount], bz[kShapeCount];
    int64_t base[kShapeCount];
};
__global__ __launch_bounds__(1024) void k_block_list(const int32_t* pmap, int nby, int nbz, ShapeLists sl,
                                                     uint8_t* flag_all, int32_t* list_all, int64_t* count_all) {
    __shared__ int cnt[kBlockLevels];
    __shared__ int off[kBlockLevels];
    const int BYv = sl.by[blockIdx.x], BZv = sl.bz[blockIdx.x];
    uint8_t* flag = flag_all + sl.base[blockIdx.x];
    int32_t* list = list_all + sl.base[blockIdx.x];
    int64_t* count = count_all + blockIdx.x;
    const int NBY = (nby + BYv - 1) / BYv, NBZ = (nbz + BZv - 1) / BZv;
    const int total = NBY * NBZ, nlev = NBY + NBZ - 1;
    for (int i = threadIdx.x; i < total; i += 1024) {
        const int Y = i % NBY, Z = i / NBY;
        uint8_t any = 0;
        for (int dz = 0; dz < BZv; ++dz)
            for (int dy = 0; dy < BYv; ++dy) {
                const int by = Y * BYv + dy, bz = Z * BZv + dz;
                if (by < nby && bz < nbz && pmap[(size_t)by + (size_t)nby * bz] >= 0) any = 1;
            }
        flag[i] = any;
    }
    __syncthreads();
    for (int l = threadIdx.x; l < nlev; l += 1024) {
        int n = 0;
        for (int Y = max(0, l - NBZ + 1); Y <= min(l, NBY - 1); ++Y) n += flag[(size_t)Y + (size_t)NBY * (l - Y)];
        cnt[l] = n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int o = 0;
        for (int l = 0; l < nlev; ++l) {
            off[l] = o;
            o += cnt[l];
        }
        *count = o;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < total; i += 1024) {
        if (!flag[i]) continue;
        const int Y = i % NBY, Z = i / NBY, l = Y + Z;
        int rank = 0;
        for (int Y2 = max(0, l - NBZ + 1); Y2 < Y; ++Y2) rank += flag[(size_t)Y2 + (size_t)NBY * (l - Y2)];
        list[off[l] + rank] = Y | (Z << 10);
    }
}

// ---------------------------------------------------------------- batched Krylov chunks
// Serving: one launch per kernel class advances every request of a batch
// (KrylovJob). A workgroup finds its request from its index and runs the
// same code, with the same workgroup index and count, as that request's
// direct launch; each request reduces over its own workgroups only, so its
// result does not depend on the batch it ran in.
__global__ __launch_bounds__(256) void k_spmv_batch(const KrylovJob* __restrict__ jobs, int per_job, int it) {
    const int k = (int)blockIdx.x / per_job, g = (int)blockIdx.x % per_job;
    const KrylovJob J = jobs[k];
    if (J.ctrl == nullptr || g >= J.g_spmv) return;
    const bool odd = ((J.par + it) & 1) != 0;  // p_in = P[kiter & 1], p_out = P[(kiter + 1) & 1]
    spmv_body<float>(g, J.g_spmv, J.bk, J.c, J.y, odd ? J.p1 : J.p0, odd ? J.p0 : J.p1, J.q, J.ctrl, J.stage,
                     J.timeout);
}

// One kernel per sweep shape (1x4, 2x2: both 256 threads); workgroups of
// requests of the other shape exit at once.
template <int BY, int BZ>
__global__ __launch_bounds__(256, 4) void k_fwd_batch(const KrylovJob* __restrict__ jobs, int per_job, int it) {
    constexpr int kShape = BY == 1 ? kShape1x4 : kShape2x2;
    const int k = (int)blockIdx.x / per_job, g = (int)blockIdx.x % per_job;
    const KrylovJob J = jobs[k];
    if (J.ctrl == nullptr || J.pc.shape != kShape || g >= J.g_fwd) return;
    const float* p = ((J.par + it) & 1) != 0 ? J.p0 : J.p1;  // P[(kiter + 1) & 1]
    if (J.ctrl->done) return;
    fwd_block_body<float, true, BY, BZ>(g, J.g_fwd, J.bk, J.pc, J.c, J.r, J.x, p, J.q, J.y, J.ctrl, J.xch, J.stage,
                                        J.timeout, J.ctrl->epoch, next_tag(J.ctrl->red_epoch));
}

template <int BY, int BZ>
__global__ __launch_bounds__(256, 4) void k_bwd_batch(const KrylovJob* __restrict__ jobs, int per_job) {
    constexpr int kShape = BY == 1 ? kShape1x4 : kShape2x2;
    const int k = (int)blockIdx.x / per_job, g = (int)blockIdx.x % per_job;
    const KrylovJob J = jobs[k];
    if (J.ctrl == nullptr || J.pc.shape != kShape || g >= J.g_bwd) return;
    if (J.ctrl->done) return;
    bwd_block_body<float, BY, BZ>(g, J.g_bwd, J.bk, J.pc, J.c, J.r, J.y, J.ctrl, J.xch, J.stage, SC_BETA, J.timeout,
                                  J.ctrl->epoch, next_tag(J.ctrl->red_epoch));
}

// ---------------------------------------------------------------- pencil setup
__global__ __launch_bounds__(256) void k_pencil_flags(const int32_t* bmap, int nbx, int nby, int nbz, uint8_t* pflag) {
    const int64_t pi = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (pi >= (int64_t)nby * nbz) return;
    const int32_t* row = bmap + (size_t)nbx * (size_t)pi;
    uint8_t any = 0;
    for (int bx = 0; bx < nbx; ++bx) any |= row[bx] >= 0 ? 1 : 0;
    pflag[pi] = any;
}

// r64 = b - A x64 in FP64 (order as the oracle's spmv: diag, -x, +x, -y, +y,
// -z, +z). Optionally hands float(r64) to the FP32 inner solve.
template <class TC>
__global__ __launch_bounds__(256) void k_resid64(Bricks bk, const double* diag64, const TC* ax, const TC* ay,
                                                 const TC* az, const double* b64, const double* x64, double* part,
                                                 double* part2, float* r32, float* x32, float* p0, float* p1,
                                                 const Ctrl* ctrl) {
    if (ctrl && ctrl->outer_done) return;
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= bk.nb) return;
    const Lane L;
    const size_t o = (size_t)b * kBrickCells + L.lane;
    const double xc = x64[o];
    const double axo = (double)ax[o], ayo = (double)ay[o], azo = (double)az[o];
    auto xext = [&](int d, int off) -> double {
        const int nb_ = nbr_id(bk, d, b);
        return nb_ >= 0 ? x64[(size_t)nb_ * kBrickCells + L.lane + off] : 0.0;
    };
    auto aext = [&](const TC* a, int d, int off) -> double {
        const int nb_ = nbr_id(bk, d, b);
        return nb_ >= 0 ? (double)a[(size_t)nb_ * kBrickCells + L.lane + off] : 0.0;
    };
    const double sxm = shfl(xc, L.lane - 1), sxp = shfl(xc, L.lane + 1);
    const double sym = shfl(xc, L.lane - 4), syp = shfl(xc, L.lane + 4);
    const double szm = shfl(xc, L.lane - 16), szp = shfl(xc, L.lane + 16);
    const double cxm = shfl(axo, L.lane - 1), cym = shfl(ayo, L.lane - 4), czm = shfl(azo, L.lane - 16);
    const double x_xm = L.lx > 0 ? sxm : xext(0, 3);
    const double x_xp = L.lx < 3 ? sxp : xext(1, -3);
    const double x_ym = L.ly > 0 ? sym : xext(2, 12);
    const double x_yp = L.ly < 3 ? syp : xext(3, -12);
    const double x_zm = L.lz > 0 ? szm : xext(4, 48);
    const double x_zp = L.lz < 3 ? szp : xext(5, -48);
    const double axm = L.lx > 0 ? cxm : aext(ax, 0, 3);
    const double aym = L.ly > 0 ? cym : aext(ay, 2, 12);
    const double azm = L.lz > 0 ? czm : aext(az, 4, 48);
    double acc = diag64[o] * xc;
    acc = acc + axm * x_xm;
    acc = acc + axo * x_xp;
    acc = acc + aym * x_ym;
    acc = acc + ayo * x_yp;
    acc = acc + azm * x_zm;
    acc = acc + azo * x_zp;
    const double rv = b64[o] - acc;
    const double s = wave_sum(rv * rv);
    if (L.lane == 0) part[b] = s;
    if (r32) {
        const float rf = (float)rv;
        r32[o] = rf;
        const double s2 = wave_sum((double)rf * (double)rf);
        if (L.lane == 0) part2[b] = s2;
    }
    if (x32) x32[o] = 0.0f;
    if (p0) {
        p0[o] = 0.0f;
        p1[o] = 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_xupdate(Bricks bk, double* x64, const float* x32, const Ctrl* ctrl) {
    if (ctrl->outer_done) return;
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= bk.nb) return;
    const size_t o = (size_t)b * kBrickCells + lane_id();
    x64[o] = x64[o] + (double)x32[o];
}

// ================================================================== finish
// projection.cpp:85-91 with background 0 (simulator.cpp:110).
__global__ __launch_bounds__(256) void k_scatter(Dense d, const uint8_t* mask, Bricks bk, const double* x,
                                                 float* pressure) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.ni || j >= d.nj) return;
    const size_t c = cell(i, j, k, d.ni, d.nj);
    float p = 0.0f;
    if (mask[c] & kFluid) {
        const int b = bk.map[(size_t)(i >> 2) + (size_t)bk.nbx * ((size_t)(j >> 2) + (size_t)bk.nby * (k >> 2))];
        const int l = (i & 3) + 4 * (j & 3) + 16 * (k & 3);
        p = (float)x[(size_t)b * kBrickCells + l];
    }
    pressure[c] = p;
}

// projection.cpp:93-134: one face component per launch (AXIS 0 u, 1 v, 2 w).
// Loop ranges of the reference: normal index in [1, n_axis-1], tangential
// indices in [1, n-2].
template <int AXIS>
__global__ __launch_bounds__(256) void k_update(Dense d, float dt, const float* pr) {
    const int fi = blockIdx.x * 64 + threadIdx.x;
    const int fj = blockIdx.y * 4 + threadIdx.y;
    const int fk = blockIdx.z;
    const int ni = d.ni, nj = d.nj, nk = d.nk;
    const int ei = AXIS == 0 ? ni : ni - 1;  // exclusive upper bounds
    const int ej = AXIS == 1 ? nj : nj - 1;
    const int ek = AXIS == 2 ? nk : nk - 1;
    if (fi < 1 || fj < 1 || fk < 1 || fi >= ei || fj >= ej || fk >= ek) return;
    const int di = AXIS == 0, dj = AXIS == 1, dk = AXIS == 2;
    const size_t c1 = cell(fi, fj, fk, ni, nj);
    const size_t c0 = cell(fi - di, fj - dj, fk - dk, ni, nj);
    const float ph1 = d.phi[c1], ph0 = d.phi[c0];
    if (ph1 < 0.0f || ph0 < 0.0f) {
        float theta = 1.0f;
        if (ph1 >= 0.0f || ph0 >= 0.0f) theta = fraction_inside2(ph0, ph1);
        theta = clamp_theta(theta);
        float* f;
        if (AXIS == 0)
            f = d.u + cell(fi, fj, fk, ni + 1, nj);
        else if (AXIS == 1)
            f = d.v + cell(fi, fj, fk, ni, nj + 1);
        else
            f = d.w + cell(fi, fj, fk, ni, nj);
        const float num = dt * (pr[c1] - pr[c0]);
        const float den = d.dx * theta;
        *f = *f - num / den;
    }
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

constexpr int kWaveBlocksPerCU = 2;    // persistent factor sweep: 2 workgroups (8 waves) per CU
// persistent SpMV: up to 8 workgroups (32 waves) per CU in FP32 (the
// occupancy of its two-brick loop allows six), 4 in FP64 (8 measured 2% slower)
template <class T>
constexpr int kSpmvBlocksPerCU = sizeof(T) == 4 ? 8 : 4;
constexpr int kBlockGroupsPerCU = 4;   // pencil-block sweeps: up to 4 workgroups per CU

int device_cus() {
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1;
    if (!cus[dev] && hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus[dev] = 1;
    return cus[dev] > 0 ? cus[dev] : 1;
}

// Resident workgroups per CU of a kernel at the given block size (cached per kernel).
int occupancy(const void* kern, int threads) {
    static std::mutex mu;
    static std::vector<std::pair<const void*, int>> cache;
    std::lock_guard<std::mutex> lock(mu);
    for (const auto& kv : cache)
        if (kv.first == kern) return kv.second;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, threads, 0) != hipSuccess || occ < 1) occ = 1;
    cache.emplace_back(kern, occ);
    return occ;
}

// Launch of a persistent kernel (256 threads = 4 waves per workgroup) whose
// workgroups wait on each other: grid = min(items / 4, resident workgroups),
// so on an otherwise idle device every workgroup is resident for the whole
// launch. A plain launch is used: a cooperative launch gives the same
// residency but costs about 20 us of extra dispatch latency per kernel here.
// If another workload keeps some workgroups from starting, the bounded
// waits end at the time limit and the launch reports the abort.
int persistent_grid(const void* kern, int items, int per_group, int threads, int blocks_per_cu, int max_groups) {
    const int occ = occupancy(kern, threads);
    const int per_cu = occ < blocks_per_cu ? occ : blocks_per_cu;
    int g = (items + per_group - 1) / per_group;
    if (g > per_cu * device_cus()) g = per_cu * device_cus();
    if (g > kTailMaxGroups) g = kTailMaxGroups;
    if (max_groups > 0 && g > max_groups) g = max_groups;
    if (g < 1) g = 1;
    return g;
}

template <typename... KArgs, typename... Args, size_t... I>
hipError_t persistent_impl(void (*kern)(KArgs...), int items, int per_group, int threads, int blocks_per_cu,
                           int max_groups, hipStream_t s, std::index_sequence<I...>, Args&&... args) {
    std::tuple<KArgs...> t(std::forward<Args>(args)...);
    void* ptrs[sizeof...(KArgs)] = {(void*)&std::get<I>(t)...};
    const int g = persistent_grid((const void*)kern, items, per_group, threads, blocks_per_cu, max_groups);
    return hipLaunchKernel((const void*)kern, dim3(g), dim3(threads), ptrs, 0, s);
}

// items: work items (one per wave, four waves per 256-thread workgroup);
// max_groups > 0 caps the grid (a share of the device for concurrent solves).
template <typename... KArgs, typename... Args>
hipError_t persistent(void (*kern)(KArgs...), int items, int blocks_per_cu, int max_groups, hipStream_t s,
                      Args&&... args) {
    static_assert(sizeof...(KArgs) == sizeof...(Args), "argument count");
    return persistent_impl(kern, items, 4, 256, blocks_per_cu, max_groups, s, std::index_sequence_for<KArgs...>{},
                           std::forward<Args>(args)...);
}

// Persistent launch with one work item per workgroup of `threads` threads.
template <typename... KArgs, typename... Args>
hipError_t persistent_wg(void (*kern)(KArgs...), int items, int threads, int blocks_per_cu, int max_groups,
                         hipStream_t s, Args&&... args) {
    static_assert(sizeof...(KArgs) == sizeof...(Args), "argument count");
    return persistent_impl(kern, items, 1, threads, blocks_per_cu, max_groups, s,
                           std::index_sequence_for<KArgs...>{}, std::forward<Args>(args)...);
}

}  // namespace

// ================================================================== launchers
hipError_t launch_mask(const Dense& d, uint8_t* mask, int32_t* err, hipStream_t s) {
    dim3 grid(cdiv(d.ni, 64), cdiv(d.nj, 4), d.nk), block(64, 4);
    hipLaunchKernelGGL(k_mask, grid, block, 0, s, d, mask, err);
    return hipGetLastError();
}

int64_t scan_blocks(int64_t n) { return (n + kScanChunk - 1) / kScanChunk; }

hipError_t launch_scan(const uint8_t* flags, int64_t n, uint8_t bit, int32_t* out, int32_t fill, int32_t* blk_cnt,
                       int32_t* blk_off, int64_t* total, hipStream_t s) {
    const int64_t g = scan_blocks(n);
    hipLaunchKernelGGL(k_scan_count, dim3((unsigned)g), dim3(kScanBlock), 0, s, flags, n, bit, blk_cnt);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, blk_cnt, g, blk_off, total);
    hipLaunchKernelGGL(k_scan_write, dim3((unsigned)g), dim3(kScanBlock), 0, s, flags, n, bit, blk_off, out, fill);
    return hipGetLastError();
}

hipError_t launch_brick_flags(const Dense& d, const uint8_t* mask, int nbx, int nby, int nbz, uint8_t* bflag,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_brick_flags, dim3(cdiv((int64_t)nbx * nby * nbz, 256)), dim3(256), 0, s, d, mask, nbx,
                       nby, nbz, bflag);
    return hipGetLastError();
}

hipError_t launch_brick_meta(const int32_t* map, int nbx, int nby, int nbz, int32_t* coord, int32_t* lvl_cnt,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_brick_meta, dim3(cdiv((int64_t)nbx * nby * nbz, 256)), dim3(256), 0, s, map, nbx, nby,
                       nbz, coord, lvl_cnt);
    return hipGetLastError();
}

hipError_t launch_level_scan(const int32_t* lvl_cnt, int nlev, int32_t* lvl_off, int32_t* lvl_fill,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_level_scan, dim3(1), dim3(64), 0, s, lvl_cnt, nlev, lvl_off, lvl_fill);
    return hipGetLastError();
}

hipError_t launch_brick_lists(const int32_t* map, const int32_t* coord, int nbx, int nby, int nbz, int nb,
                              const int32_t* lvl_off, int32_t* lvl_fill, int32_t* lvl_list, int32_t* nbr,
                              hipStream_t s) {
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_brick_lists, dim3(cdiv(nb, 256)), dim3(256), 0, s, map, coord, nbx, nby, nbz, nb,
                       lvl_off, lvl_fill, lvl_list, nbr);
    return hipGetLastError();
}

template <class T>
hipError_t launch_assemble(const Dense& d, float dt, const uint8_t* mask, const Bricks& bk, double* diag64,
                           double* b64, T* diag, T* ax, T* ay, T* az, hipStream_t s) {
    const float scale = d.dx / dt;  // projection.cpp:82, float division
    hipLaunchKernelGGL(k_assemble<T>, dim3(cdiv(bk.nb, 4)), dim3(256), 0, s, d, scale, mask, bk, diag64, b64,
                       diag, ax, ay, az);
    return hipGetLastError();
}

template <class T>
hipError_t launch_factor(const Bricks& bk, const double* diag64, const T* ax, const T* ay, const T* az, double tau,
                         double sigma, double* inv_e64, float* inv_e32, Ctrl* ctrl, uint64_t* xch, int32_t* n_fixed,
                         int level, int level_base, int level_count, uint64_t timeout_ticks, int max_groups,
                         hipStream_t s) {
    if (level < 0) {
        if (bk.nb == 0) return hipSuccess;
        return persistent(k_factor_wave<T>, bk.nb, kWaveBlocksPerCU, max_groups, s, bk, diag64, ax, ay, az, tau,
                          sigma, inv_e64, inv_e32, ctrl, xch,
                    n_fixed, timeout_ticks);
    }
    if (level_count > 0)
        hipLaunchKernelGGL((k_factor_level<T>), dim3(level_count), dim3(64), 0, s, bk, diag64, ax, ay, az, tau, sigma,
                           inv_e64, inv_e32, ctrl, xch, n_fixed, level_base, timeout_ticks);
    return hipGetLastError();
}

template <class T>
hipError_t launch_export(const Dense& d, const uint8_t* mask, const int32_t* indices, const Bricks& bk,
                         const double* diag64, const double* b64, const T* ax, const T* ay, const T* az,
                         double* odiag, double* oax, double* oay, double* oaz, double* ob, hipStream_t s) {
    hipLaunchKernelGGL(k_export<T>, dim3(cdiv(bk.nb, 4)), dim3(256), 0, s, d, mask, indices, bk, diag64, b64, ax,
                       ay, az, odiag, oax, oay, oaz, ob);
    return hipGetLastError();
}

hipError_t launch_export_cells(const Dense& d, const uint8_t* mask, const int32_t* indices, const Bricks& bk,
                               const double* v, double* out, hipStream_t s) {
    if (bk.nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_export_cells, dim3(cdiv(bk.nb, 4)), dim3(256), 0, s, d, mask, indices, bk, v, out);
    return hipGetLastError();
}

template <class T>
hipError_t launch_init(const Bricks& bk, const double* b64, T* r, T* x, T* p0, T* p1, double* part, Ctrl* ctrl,
                       hipStream_t s) {
    (void)ctrl;
    hipLaunchKernelGGL(k_init<T>, dim3(cdiv(bk.nb, 4)), dim3(256), 0, s, bk, b64, r, x, p0, p1, part);
    return hipGetLastError();
}

int scalar_blocks(int n) {
    const int g = (n + kRedPerBlock - 1) / kRedPerBlock;
    return g < 1 ? 1 : (g > kRedBlocks ? kRedBlocks : g);
}

hipError_t launch_export_words(const void* src, size_t bytes, void* host, hipStream_t s) {
    if (bytes % 4 != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_export_words, dim3(1), dim3(64), 0, s, (const uint32_t*)src, (uint32_t*)host, (int)(bytes / 4));
    return hipGetLastError();
}

hipError_t launch_scalar(Ctrl* ctrl, int mode, const double* parts, int n, const double* parts2, uint64_t* stage,
                         uint64_t timeout_ticks, hipStream_t s) {
    const int nblk = scalar_blocks(n);
    hipLaunchKernelGGL(k_scalar, dim3(nblk), dim3(kReduceThreads), 0, s, ctrl, mode, parts, n, parts2, nblk, stage,
                       timeout_ticks);
    return hipGetLastError();
}

template <class T>
hipError_t launch_spmv(const Bricks& bk, const Coef<T>& c, const T* z, const T* p_in, T* p_out, T* q, Ctrl* ctrl,
                       uint64_t* stage, uint64_t timeout_ticks, int max_groups, hipStream_t s) {
    return persistent(k_spmv<T>, bk.nb, kSpmvBlocksPerCU<T>, max_groups, s, bk, c, z, p_in, p_out, q, ctrl, stage,
                      timeout_ticks);
}

// Persistent sweeps of one pencil-block shape.
template <class T, int S>
hipError_t launch_fwd_shape(const Bricks& bk, const Pencils& pc, const Coef<T>& c, T* r, T* x, const T* p, const T* q,
                            T* y, Ctrl* ctrl, uint64_t* xch, uint64_t* stage, int update, uint64_t timeout_ticks,
                            int max_groups, hipStream_t s) {
    constexpr int BY = kShapeY[S], BZ = kShapeZ[S];
    if (update)
        return persistent_wg(k_fwd_block<T, true, BY, BZ>, pc.nblk, 64 * BY * BZ, kBlockGroupsPerCU, max_groups, s,
                             bk, pc, c, r, x, p, q, y, ctrl, xch, stage, timeout_ticks);
    return persistent_wg(k_fwd_block<T, false, BY, BZ>, pc.nblk, 64 * BY * BZ, kBlockGroupsPerCU, max_groups, s, bk,
                         pc, c, r, x, p, q, y, ctrl, xch, stage, timeout_ticks);
}

template <class T, int S>
hipError_t launch_bwd_shape(const Bricks& bk, const Pencils& pc, const Coef<T>& c, const T* r, T* y, Ctrl* ctrl,
                            uint64_t* xch, uint64_t* stage, int tail_mode, uint64_t timeout_ticks, int max_groups,
                            hipStream_t s) {
    constexpr int BY = kShapeY[S], BZ = kShapeZ[S];
    return persistent_wg(k_bwd_block<T, BY, BZ>, pc.nblk, 64 * BY * BZ, kBlockGroupsPerCU, max_groups, s, bk, pc, c,
                         r, y, ctrl, xch, stage, tail_mode, timeout_ticks);
}

bool sweeps_fusable(const Pencils& pc) { return pc.shape == kShape1x4; }

hipError_t launch_sweeps(const Bricks& bk, const Pencils& pc, const Coef<float>& c, float* r, float* x, const float* p,
                         const float* q, float* y, Ctrl* ctrl, uint64_t* xch, uint64_t* stage, uint64_t timeout_ticks,
                         int max_groups, hipStream_t s) {
    if (!sweeps_fusable(pc)) return hipErrorInvalidValue;
    return persistent_wg(k_sweeps_block<1, 4>, pc.nblk, 256, kBlockGroupsPerCU, max_groups, s, bk, pc, c, r, x, p, q,
                         y, ctrl, xch, stage, timeout_ticks);
}

template <class T>
hipError_t launch_fwd(const Bricks& bk, const Pencils& pc, const Coef<T>& c, T* r, T* x, const T* p, const T* q, T* y,
                      double* part, Ctrl* ctrl, uint64_t* xch, uint64_t* stage, int update, int level,
                      int level_base, int level_count, uint64_t timeout_ticks, int max_groups, hipStream_t s) {
    if (level < 0) {
        switch (pc.shape) {
        case kShape2x2:
            return launch_fwd_shape<T, kShape2x2>(bk, pc, c, r, x, p, q, y, ctrl, xch, stage, update, timeout_ticks,
                                                  max_groups, s);
        case kShape1x4:
            return launch_fwd_shape<T, kShape1x4>(bk, pc, c, r, x, p, q, y, ctrl, xch, stage, update, timeout_ticks,
                                                  max_groups, s);
        default:
            return hipErrorInvalidValue;
        }
    }
    if (level_count > 0) {
        if (update)
            hipLaunchKernelGGL((k_fwd_level<T, true>), dim3(level_count), dim3(64), 0, s, bk, c, r, x, p, q, y, part,
                               ctrl, xch, level_base, timeout_ticks);
        else
            hipLaunchKernelGGL((k_fwd_level<T, false>), dim3(level_count), dim3(64), 0, s, bk, c, r, x, p, q, y, part,
                               ctrl, xch, level_base, timeout_ticks);
    }
    return hipGetLastError();
}

template <class T>
hipError_t launch_bwd(const Bricks& bk, const Pencils& pc, const Coef<T>& c, const T* r, T* y, double* part,
                      Ctrl* ctrl, uint64_t* xch, uint64_t* stage, int tail_mode, int level, int level_base,
                      int level_count, uint64_t timeout_ticks, int max_groups, hipStream_t s) {
    if (level < 0) {
        switch (pc.shape) {
        case kShape2x2:
            return launch_bwd_shape<T, kShape2x2>(bk, pc, c, r, y, ctrl, xch, stage, tail_mode, timeout_ticks,
                                                  max_groups, s);
        case kShape1x4:
            return launch_bwd_shape<T, kShape1x4>(bk, pc, c, r, y, ctrl, xch, stage, tail_mode, timeout_ticks,
                                                  max_groups, s);
        default:
            return hipErrorInvalidValue;
        }
    }
    if (level_count > 0)
        hipLaunchKernelGGL((k_bwd_level<T>), dim3(level_count), dim3(64), 0, s, bk, c, r, y, part, ctrl, xch,
                           level_base, timeout_ticks);
    return hipGetLastError();
}

template <int BY, int BZ>
void krylov_groups_shape(const Bricks& bk, const Pencils& pc, int max_groups, int* g_spmv, int* g_fwd, int* g_bwd) {
    *g_spmv = persistent_grid((const void*)k_spmv<float>, bk.nb, 4, 256, kSpmvBlocksPerCU<float>, max_groups);
    *g_fwd = persistent_grid((const void*)k_fwd_block<float, true, BY, BZ>, pc.nblk, 1, 64 * BY * BZ,
                             kBlockGroupsPerCU, max_groups);
    *g_bwd = persistent_grid((const void*)k_bwd_block<float, BY, BZ>, pc.nblk, 1, 64 * BY * BZ, kBlockGroupsPerCU,
                             max_groups);
}

void krylov_groups(const Bricks& bk, const Pencils& pc, int max_groups, int* g_spmv, int* g_fwd, int* g_bwd) {
    if (pc.shape == kShape1x4)
        krylov_groups_shape<1, 4>(bk, pc, max_groups, g_spmv, g_fwd, g_bwd);
    else
        krylov_groups_shape<2, 2>(bk, pc, max_groups, g_spmv, g_fwd, g_bwd);
}

template <class T>
hipError_t launch_brick_uniform(const Bricks& bk, const T* diag, const T* ax, const T* ay, const T* az, T cref,
                                T dref, const double* ie64, double tau, double sigma, int32_t* uflag, int32_t* umap,
                                T* urow, hipStream_t s) {
    hipLaunchKernelGGL((k_brick_uniform<T>), dim3(std::max(1u, cdiv(bk.nb, 4))), dim3(256), 0, s, bk, diag, ax, ay,
                       az, cref, dref, ie64, tau, sigma, uflag, umap, urow);
    return hipGetLastError();
}

hipError_t launch_scale_lower(const Bricks& bk, const double* ax, const double* ay, const double* az,
                              const double* ie, double* lx, double* ly, double* lz, hipStream_t s) {
    if (bk.nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale_lower, dim3(cdiv(bk.nb, 4)), dim3(256), 0, s, bk, ax, ay, az, ie, lx, ly, lz);
    return hipGetLastError();
}

hipError_t launch_krylov_batch(const KrylovJob* jobs, int njobs, int per_job, int n, int shapes, hipStream_t s) {
    const dim3 grid((unsigned)(njobs * per_job)), block(256);
    const bool s14 = (shapes & (1 << kShape1x4)) != 0, s22 = (shapes & (1 << kShape2x2)) != 0;
    for (int it = 0; it < n; ++it) {
        hipLaunchKernelGGL(k_spmv_batch, grid, block, 0, s, jobs, per_job, it);
        if (s14) hipLaunchKernelGGL((k_fwd_batch<1, 4>), grid, block, 0, s, jobs, per_job, it);
        if (s22) hipLaunchKernelGGL((k_fwd_batch<2, 2>), grid, block, 0, s, jobs, per_job, it);
        if (s14) hipLaunchKernelGGL((k_bwd_batch<1, 4>), grid, block, 0, s, jobs, per_job);
        if (s22) hipLaunchKernelGGL((k_bwd_batch<2, 2>), grid, block, 0, s, jobs, per_job);
    }
    return hipGetLastError();
}

int sweep_round_groups(int max_groups) {
    int g = kBlockGroupsPerCU * device_cus();
    if (g > kTailMaxGroups) g = kTailMaxGroups;
    if (max_groups > 0 && g > max_groups) g = max_groups;
    return g < 1 ? 1 : g;
}

hipError_t launch_pencil_flags(const int32_t* bmap, int nbx, int nby, int nbz, uint8_t* pflag, hipStream_t s) {
    hipLaunchKernelGGL(k_pencil_flags, dim3(cdiv((int64_t)nby * nbz, 256)), dim3(256), 0, s, bmap, nbx, nby, nbz,
                       pflag);
    return hipGetLastError();
}

template <class TC>
hipError_t launch_resid64(const Bricks& bk, const double* diag64, const TC* ax, const TC* ay, const TC* az,
                          const double* b64, const double* x64, double* part, double* part2, float* r32, float* x32,
                          float* p0, float* p1, Ctrl* ctrl, hipStream_t s) {
    hipLaunchKernelGGL(k_resid64<TC>, dim3(cdiv(bk.nb, 4)), dim3(256), 0, s, bk, diag64, ax, ay, az, b64, x64,
                       part, part2, r32, x32, p0, p1, (const Ctrl*)ctrl);
    return hipGetLastError();
}

hipError_t launch_xupdate(const Bricks& bk, double* x64, const float* x32, Ctrl* ctrl, hipStream_t s) {
    hipLaunchKernelGGL(k_xupdate, dim3(cdiv(bk.nb, 4)), dim3(256), 0, s, bk, x64, x32, (const Ctrl*)ctrl);
    return hipGetLastError();
}

// ================================================================== inputs
// The steps before the projection in the reference's frame (simulator.cpp:
// 90-93, 125-178, 187-189 with boundary.cpp:6-34), on device grids. One
// thread per face, 64 x 4 threads per block over (i, j), blocks over k.

// level_set.h:83-163: area fraction of a square cell face inside the level
// set from its four corner values. The corner list (bl, br, tr, tl) is
// rotated to a canonical position; 0.5 and 0.25 are double literals in the
// reference, so those products are evaluated in double and rounded to float
// where the reference's float variables take them.
__device__ float fraction_inside4(float bl, float br, float tl, float tr) {
    const int inside = (bl < 0.0f) + (tl < 0.0f) + (br < 0.0f) + (tr < 0.0f);
    float q[4] = {bl, br, tr, tl};
    auto rotate = [&q]() {
        const float t = q[0];
        q[0] = q[1];
        q[1] = q[2];
        q[2] = q[3];
        q[3] = t;
    };
    if (inside == 4) return 1.0f;
    if (inside == 3) {
        for (int n = 0; n < 4 && q[0] < 0.0f; ++n) rotate();
        const float s0 = 1.0f - fraction_inside2(q[0], q[3]);
        const float s1 = 1.0f - fraction_inside2(q[0], q[1]);
        return (float)(1.0 - 0.5 * s0 * s1);
    }
    if (inside == 2) {
        for (int n = 0; n < 4 && (q[0] >= 0.0f || !(q[1] < 0.0f || q[2] < 0.0f)); ++n) rotate();
        if (q[1] < 0.0f) return (float)(0.5 * (fraction_inside2(q[0], q[3]) + fraction_inside2(q[1], q[2])));
        const float mid = (float)(0.25 * (q[0] + q[1] + q[2] + q[3]));
        float area = 0.0f;
        if (mid < 0.0f) {
            const float s1 = 1.0f - fraction_inside2(q[0], q[3]);
            const float s3 = 1.0f - fraction_inside2(q[2], q[3]);
            area = (float)(area + 0.5 * s1 * s3);
            const float s2 = 1.0f - fraction_inside2(q[2], q[1]);
            const float s0 = 1.0f - fraction_inside2(q[0], q[1]);
            area = (float)(area + 0.5 * s0 * s2);
            return 1.0f - area;
        }
        const float s0 = fraction_inside2(q[0], q[1]);
        const float s1 = fraction_inside2(q[0], q[3]);
        area = (float)(area + 0.5 * s0 * s1);
        const float s2 = fraction_inside2(q[2], q[1]);
        const float s3 = fraction_inside2(q[2], q[3]);
        area = (float)(area + 0.5 * s2 * s3);
        return area;
    }
    if (inside == 1) {
        for (int n = 0; n < 4 && q[0] >= 0.0f; ++n) rotate();
        const float s0 = fraction_inside2(q[0], q[3]);
        const float s1 = fraction_inside2(q[0], q[1]);
        return (float)(0.5 * s0 * s1);
    }
    return 0.0f;
}

// boundary.cpp:6-34: weight = clamp(1 - fraction of the face inside the
// solid, 0, 1) from the nodal solid level set ((ni+1)(nj+1)(nk+1) values).
template <int AXIS>
__global__ __launch_bounds__(256) void k_weights(int ni, int nj, int nk, const float* nodal, float* wgt) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    const int fi = ni + (AXIS == 0), fj = nj + (AXIS == 1);
    if (i >= fi || j >= fj) return;
    const int a = ni + 1, b = nj + 1;
    auto N = [&](int x, int y, int z) { return nodal[cell(x, y, z, a, b)]; };
    float f;
    if (AXIS == 0)
        f = fraction_inside4(N(i, j, k), N(i, j + 1, k), N(i, j, k + 1), N(i, j + 1, k + 1));
    else if (AXIS == 1)
        f = fraction_inside4(N(i, j, k), N(i, j, k + 1), N(i + 1, j, k), N(i + 1, j, k + 1));
    else
        f = fraction_inside4(N(i, j, k), N(i, j + 1, k), N(i + 1, j, k), N(i + 1, j + 1, k));
    const float v = 1.0f - f;
    wgt[cell(i, j, k, fi, fj)] = v < 0.0f ? 0.0f : (1.0f < v ? 1.0f : v);  // std::clamp(v, 0, 1)
}

// simulator.cpp:90-93: v -= gravity * dt on every v face (constants.h:5-6).
__global__ __launch_bounds__(256) void k_gravity(float* v, size_t n, float g_dt) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) v[t] = v[t] - g_dt;
}

// simulator.cpp:125-158 then 169-178: a face is invalid when its weight is 0,
// or, inside the ranges [1, n_face - 1) of each axis, when both cells beside
// it are outside the liquid; invalid faces are set to 0.
template <int AXIS>
__global__ __launch_bounds__(256) void k_valid_zero(Dense d) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    const int ni = d.ni, nj = d.nj, nk = d.nk;
    const int fi = ni + (AXIS == 0), fj = nj + (AXIS == 1), fk = nk + (AXIS == 2);
    if (i >= fi || j >= fj) return;
    const size_t f = cell(i, j, k, fi, fj);
    const float* wgt = AXIS == 0 ? d.uw : AXIS == 1 ? d.vw : d.ww;
    float* vel = AXIS == 0 ? d.u : AXIS == 1 ? d.v : d.w;
    bool invalid = wgt[f] == 0.0f;
    if (i >= 1 && j >= 1 && k >= 1 && i < fi - 1 && j < fj - 1 && k < fk - 1) {
        const size_t c1 = cell(i, j, k, ni, nj);
        const size_t c0 = cell(i - (AXIS == 0), j - (AXIS == 1), k - (AXIS == 2), ni, nj);
        if (d.phi[c0] >= 0.0f && d.phi[c1] >= 0.0f) invalid = true;
    }
    if (invalid) vel[f] = 0.0f;
}

// Plain streaming copy: 16-byte words when both pointers are 16-byte aligned,
// 4-byte words otherwise. Its HBM counters per byte copied calibrate the
// traffic counters of the solver kernels.
template <typename W>
__global__ __launch_bounds__(256) void k_copy(const W* __restrict__ src, W* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_weights(int ni, int nj, int nk, const float* nodal, float* uw, float* vw, float* ww,
                          hipStream_t s) {
    const dim3 block(64, 4);
    hipLaunchKernelGGL(k_weights<0>, dim3(cdiv(ni + 1, 64), cdiv(nj, 4), nk), block, 0, s, ni, nj, nk, nodal, uw);
    hipLaunchKernelGGL(k_weights<1>, dim3(cdiv(ni, 64), cdiv(nj + 1, 4), nk), block, 0, s, ni, nj, nk, nodal, vw);
    hipLaunchKernelGGL(k_weights<2>, dim3(cdiv(ni, 64), cdiv(nj, 4), nk + 1), block, 0, s, ni, nj, nk, nodal, ww);
    return hipGetLastError();
}

hipError_t launch_gravity(float* v, size_t n, float dt, hipStream_t s) {
    const float g_dt = 9.80665f * dt;  // constants::gravity<float> * dt
    hipLaunchKernelGGL(k_gravity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v, n, g_dt);
    return hipGetLastError();
}

hipError_t launch_valid_zero(const Dense& d, hipStream_t s) {
    const dim3 block(64, 4);
    hipLaunchKernelGGL(k_valid_zero<0>, dim3(cdiv(d.ni + 1, 64), cdiv(d.nj, 4), d.nk), block, 0, s, d);
    hipLaunchKernelGGL(k_valid_zero<1>, dim3(cdiv(d.ni, 64), cdiv(d.nj + 1, 4), d.nk), block, 0, s, d);
    hipLaunchKernelGGL(k_valid_zero<2>, dim3(cdiv(d.ni, 64), cdiv(d.nj, 4), d.nk + 1), block, 0, s, d);
    return hipGetLastError();
}

hipError_t launch_block_lists(const int32_t* pmap, int nby, int nbz, uint8_t* flag, int32_t* list, int64_t* count,
                              hipStream_t s) {
    if (nby > 1023 || nbz > 1023) return hipErrorInvalidValue;
    ShapeLists sl;
    int64_t base = 0;
    for (int k = 0; k < kShapeCount; ++k) {
        sl.by[k] = kShapeY[k];
        sl.bz[k] = kShapeZ[k];
        sl.base[k] = base;
        base += (int64_t)shape_blocks(nby, nbz, k);
    }
    hipLaunchKernelGGL(k_block_list, dim3(kShapeCount), dim3(1024), 0, s, pmap, nby, nbz, sl, flag, list, count);
    return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    const bool wide = (((uintptr_t)dst | (uintptr_t)src) & 15) == 0;
    const size_t n16 = wide ? bytes / 16 : 0;
    if (n16) {
        const unsigned blocks = (unsigned)std::min<size_t>((n16 + 255) / 256, 8192);
        hipLaunchKernelGGL(k_copy<uint4>, dim3(blocks), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n16);
    }
    const size_t n4 = (bytes - n16 * 16) / 4;
    if (n4) {
        const unsigned blocks = (unsigned)std::min<size_t>((n4 + 255) / 256, 8192);
        hipLaunchKernelGGL(k_copy<uint32_t>, dim3(blocks), dim3(256), 0, s, (const uint32_t*)src + n16 * 4,
                           (uint32_t*)dst + n16 * 4, n4);
    }
    return hipGetLastError();
}

hipError_t launch_scatter(const Dense& d, const uint8_t* mask, const Bricks& bk, const double* x, float* pressure,
                          hipStream_t s) {
    dim3 grid(cdiv(d.ni, 64), cdiv(d.nj, 4), d.nk), block(64, 4);
    hipLaunchKernelGGL(k_scatter, grid, block, 0, s, d, mask, bk, x, pressure);
    return hipGetLastError();
}

hipError_t launch_update(const Dense& d, float dt, const float* pressure, hipStream_t s) {
    dim3 block(64, 4);
    hipLaunchKernelGGL(k_update<0>, dim3(cdiv(d.ni + 1, 64), cdiv(d.nj, 4), d.nk), block, 0, s, d, dt, pressure);
    hipLaunchKernelGGL(k_update<1>, dim3(cdiv(d.ni, 64), cdiv(d.nj + 1, 4), d.nk), block, 0, s, d, dt, pressure);
    hipLaunchKernelGGL(k_update<2>, dim3(cdiv(d.ni, 64), cdiv(d.nj, 4), d.nk + 1), block, 0, s, d, dt, pressure);
    return hipGetLastError();
}

// explicit instantiations
#define ELQ_INST(T)                                                                                                 \
    template hipError_t launch_assemble<T>(const Dense&, float, const uint8_t*, const Bricks&, double*, double*, T*, \
                                           T*, T*, T*, hipStream_t);                                                \
    template hipError_t launch_factor<T>(const Bricks&, const double*, const T*, const T*, const T*, double, double, \
                                         double*, float*, Ctrl*, uint64_t*, int32_t*, int, int, int, uint64_t, int, \
                                         hipStream_t);                                                              \
    template hipError_t launch_export<T>(const Dense&, const uint8_t*, const int32_t*, const Bricks&, const double*, \
                                         const double*, const T*, const T*, const T*, double*, double*, double*,    \
                                         double*, double*, hipStream_t);                                            \
    template hipError_t launch_init<T>(const Bricks&, const double*, T*, T*, T*, T*, double*, Ctrl*, hipStream_t);  \
    template hipError_t launch_spmv<T>(const Bricks&, const Coef<T>&, const T*, const T*, T*, T*, Ctrl*, uint64_t*, \
                                       uint64_t, int, hipStream_t);                                                 \
    template hipError_t launch_fwd<T>(const Bricks&, const Pencils&, const Coef<T>&, T*, T*, const T*, const T*, T*, \
                                      double*, Ctrl*, uint64_t*, uint64_t*, int, int, int, int, uint64_t, int,      \
                                      hipStream_t);                                                                 \
    template hipError_t launch_bwd<T>(const Bricks&, const Pencils&, const Coef<T>&, const T*, T*, double*, Ctrl*,   \
                                      uint64_t*, uint64_t*, int, int, int, int, uint64_t, int, hipStream_t);        \
    template hipError_t launch_brick_uniform<T>(const Bricks&, const T*, const T*, const T*, const T*, T, T,         \
                                                const double*, double, double, int32_t*, int32_t*, T*, hipStream_t); \
    template hipError_t launch_resid64<T>(const Bricks&, const double*, const T*, const T*, const T*, const double*, \
                                          const double*, double*, double*, float*, float*, float*, float*, Ctrl*,   \
                                          hipStream_t);
ELQ_INST(float)
ELQ_INST(double)
#undef ELQ_INST

}  // namespace elq
