// gfx950 kernels of the frame steps around the projection.
//
// Reference (daichi-nama/eulerian-liquid):
//   extrapolate, extrapolate_const     src/project/3d/fluid/extrapolation.cpp:5-51
//   Simulator::extrapolate_velocity    src/project/3d/eulerian_liquid/simulator.cpp:119-123
//   Simulator::find_valid_velocity     src/project/3d/eulerian_liquid/simulator.cpp:125-158
//   Simulator::find_valid_level_set    src/project/3d/eulerian_liquid/simulator.cpp:160-167
//   Simulator::set_invalid_velocity_zero src/project/3d/eulerian_liquid/simulator.cpp:169-178
//   calc_cell_centered_solid_phi       src/project/3d/eulerian_liquid/simulator.cpp:191-200
//   reinitialize_fast_marching         src/project/3d/fluid/reinitialization.cpp:19-156
//                                      (calc_sdf include/nama/math/level_set.h:165-264,
//                                      vertex_interp include/nama/common/meshing.h:12-17)
//   advect_collocated_grid,            src/project/3d/fluid/advection.cpp:9-27
//   advect_staggered_grid              (trace_rk2 include/nama/math/ode.h:30-35, lerp
//                                      include/nama/math/interpolation.h:10-72,186-193)
//
// Most of it is byte and float streaming work, bound by HBM; the layered
// steps (extrapolation, reinitialisation) are bound by their dependent
// layers. Results are bit-exact with the reference (built with
// -ffp-contract=off; the float division is IEEE round-to-nearest).
//
// Extrapolation. In the reference each layer takes a snapshot of the flags
// and of the values, gives every invalid cell with at least one valid
// neighbour (of the 26 around it) the mean of those neighbours, and marks it
// valid. Here the flags are bit planes: the cells a layer adds are exactly
// the 3x3x3 box dilation of the plane minus the plane, so a first launch
// dilates the plane (into the second buffer) and lists the added cells, and
// a second launch computes only those cells, reading neighbours that were
// valid at the layer's start, which that launch does not write. No copy of
// the values is needed.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "elq_frame.hpp"

namespace elq {
namespace {

constexpr int kHX = kExtTX + 2, kHY = kExtTY + 2, kHZ = kExtTZ + 2;
constexpr int kHalo = kHX * kHY * kHZ;
static_assert(kExtTX == 32 && kExtTX * kExtTY == 256, "one thread per (x, y) column of a tile");

__device__ __forceinline__ size_t at(int i, int j, int k, int ni, int nj) {
    return (size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * (size_t)k);
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }


// Grid q of the set, field by field: selecting whole structs by a runtime
// index makes the compiler keep the array in scratch memory.
template <class T>
__device__ __forceinline__ T sel3(int q, T a, T b, T c) {
    return q == 2 ? c : (q == 1 ? b : a);
}
__device__ __forceinline__ ExtGrid pick_slot(const ExtSet& s, int q) {
    ExtGrid g;
#define ELQ_SEL(field) g.field = sel3(q, s.g[0].field, s.g[1].field, s.g[2].field)
    ELQ_SEL(f);
    ELQ_SEL(valid_in);
    ELQ_SEL(valid_out);
    ELQ_SEL(plane[0]);
    ELQ_SEL(plane[1]);
    ELQ_SEL(list);
    ELQ_SEL(ni);
    ELQ_SEL(nj);
    ELQ_SEL(nk);
    ELQ_SEL(wpr);
    ELQ_SEL(slot);
    ELQ_SEL(word0);
    ELQ_SEL(words);
#undef ELQ_SEL
    return g;
}

// The grid of the set that owns word gw (word0 ascending). The callers' words
// of one wave lie in one grid, so the index is made wave-uniform and the
// fields stay scalar loads from the kernel arguments (a per-lane choice
// would read them with vector loads for every word).
__device__ __forceinline__ ExtGrid pick(const ExtSet& s, int64_t gw) {
    const int q = (s.count > 2 && gw >= s.g[2].word0) ? 2 : ((s.count > 1 && gw >= s.g[1].word0) ? 1 : 0);
    return pick_slot(s, __builtin_amdgcn_readfirstlane(q));
}


struct TileXYZ {
    int t, bx, by, bz;
};

// Waves walk the words (stride: all waves of the grid); lane b reads the
// caller's flag of cell 64 x + b, a ballot makes the word.
__global__ __launch_bounds__(256) void k_ext_init(ExtSet set) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6), waves = (int64_t)gridDim.x * 4;
    if (wave == 0 && lane == 0)
        for (int q = 0; q < 6; ++q) set.counter[q] = 0;
    // eight words per step (one grid: grids start at multiples of
    // kExtWordAlign words), their loads issued together
    for (int64_t gw0 = wave * 8; gw0 < set.words; gw0 += waves * 8) {
        const ExtGrid g = pick(set, gw0);
        const int64_t w0 = gw0 - g.word0;
        bool v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int64_t w = w0 + q;
            const uint32_t row = (uint32_t)w / (uint32_t)g.wpr;
            const int x = (int)((uint32_t)w - row * (uint32_t)g.wpr) * 64 + lane;
            // unconditional load (cell 0 stands in outside the plane)
            const bool in = w < g.words && x < g.ni;
            const uint8_t b = g.valid_in[in ? (size_t)row * g.ni + x : 0];
            v[q] = in && b != 0;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint64_t bits = __ballot(v[q]);
            if (lane == 0 && w0 + q < g.words) g.plane[0][w0 + q] = bits;
        }
    }
}

// Layer l, one thread per word: the 3x3x3 box dilation of the plane (x by
// shifts with the neighbour words' edge bits, y and z by OR over rows), the
// added cells appended to the grid's list (one atomic per workgroup; a
// workgroup's words all belong to one grid).
constexpr int kExtDilateThreads = 1024;
__global__ __launch_bounds__(kExtDilateThreads) void k_ext_dilate(ExtSet set, int l) {
    __shared__ uint32_t s_tot[kExtDilateThreads / 64 + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t gw = (int64_t)blockIdx.x * kExtDilateThreads + threadIdx.x;  // set.words: whole workgroups
    const ExtGrid g = pick(set, gw);
    const int64_t w = gw - g.word0;
    // selects, not an index: a dynamic index into the by-value struct would
    // put it in scratch memory
    const uint64_t* src = (l & 1) ? g.plane[1] : g.plane[0];
    uint64_t* dst = (l & 1) ? g.plane[0] : g.plane[1];
    // padding words past the grid's plane compute on its last word and
    // write nothing; every lane takes part in the shuffles below
    const bool real = w < g.words;
    const int64_t wr = real ? w : g.words - 1;
    const uint32_t row = (uint32_t)wr / (uint32_t)g.wpr;  // planes hold fewer than 2^31 words
    const int xw = (int)((uint32_t)wr - row * (uint32_t)g.wpr);
    const int j = (int)(row % (uint32_t)g.nj), k = (int)(row / (uint32_t)g.nj);
    const int64_t cell0 = (int64_t)row * g.ni + (int64_t)xw * 64;
    const uint64_t old = src[wr];
    // the nine rows' words, all loaded before any is used
    uint64_t c[9];
    bool in[9];
    int64_t rw[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int jj = j + r % 3 - 1, kk = k + r / 3 - 1;
        in[r] = jj >= 0 && jj < g.nj && kk >= 0 && kk < g.nk;
        rw[r] = in[r] ? ((int64_t)jj + (int64_t)g.nj * kk) * g.wpr + xw : wr;
        c[r] = src[rw[r]];  // unconditional (a row outside the grid reads the centre one)
        c[r] = in[r] ? c[r] : 0ull;
    }
    // a row's neighbour words are the neighbour lanes' (lanes hold consecutive
    // words); lanes 0 and 63 load theirs when the row goes on past the wave
    uint64_t acc = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        uint64_t lft = __shfl_up(c[r], 1, 64), rgt = __shfl_down(c[r], 1, 64);
        if (lane == 0) lft = (in[r] && xw > 0) ? src[rw[r] - 1] : 0ull;      // rare: rows cross waves
        if (lane == 63) rgt = (in[r] && xw + 1 < g.wpr) ? src[rw[r] + 1] : 0ull;
        if (xw == 0) lft = 0ull;
        if (xw + 1 == g.wpr) rgt = 0ull;
        acc |= c[r] | (c[r] << 1) | (c[r] >> 1) | (lft >> 63) | (rgt << 63);
    }
    const int rem = g.ni - xw * 64;
    if (rem < 64) acc &= (1ull << rem) - 1ull;
    const uint64_t add = real ? acc & ~old : 0ull;
    if (real) dst[w] = old | add;
    const uint32_t n = (uint32_t)__popcll(add);
    uint32_t incl = n;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    // wave totals -> workgroup offsets -> one atomic
    if (lane == 63) s_tot[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int q = 0; q < kExtDilateThreads / 64; ++q) {
            const uint32_t t = s_tot[q];
            s_tot[q] = run;
            run += t;
        }
        s_tot[kExtDilateThreads / 64] = run ? atomicAdd(&set.counter[2 * g.slot + (l & 1)], run) : 0u;
    }
    __syncthreads();
    const uint32_t base = s_tot[kExtDilateThreads / 64] + s_tot[wave];
    // the wave writes one lane's cells at a time, lane b of the wave taking
    // bit b, so each store is a run of consecutive slots
    const uint32_t first = base + incl - n;
    uint64_t lanes = __ballot(add != 0ull);
    while (lanes) {
        const int src = __ffsll((unsigned long long)lanes) - 1;
        lanes &= lanes - 1ull;
        const uint64_t word = __shfl(add, src, 64);
        const uint32_t at0 = __shfl(first, src, 64);
        const int64_t c0 = __shfl(cell0, src, 64);
        if ((word >> lane) & 1ull)
            g.list[at0 + (uint32_t)__popcll(word & ((1ull << lane) - 1ull))] = (uint32_t)(c0 + lane);
    }
}

// extrapolation.cpp:20-44 for the cells layer l added (blockIdx.y = grid):
// the sum over the neighbours valid at the layer's start, in the reference's
// order kk, jj, ii = -1..1, divided by their count. The neighbours read are
// not written by this launch.
__global__ __launch_bounds__(256) void k_ext_value(ExtSet set, int l) {
    const int q = blockIdx.y;
    if (q >= set.count) return;
    const ExtGrid g = pick_slot(set, q);
    const uint32_t n = set.counter[2 * q + (l & 1)];
    if (blockIdx.x == 0 && threadIdx.x == 0) set.counter[2 * q + ((l & 1) ^ 1)] = 0;  // next layer's list
    const uint64_t* pl = (l & 1) ? g.plane[1] : g.plane[0];
    const size_t sj = (size_t)g.ni, sk = (size_t)g.ni * g.nj;
    for (uint32_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) {
        const uint32_t c = g.list[e];
        const int i = (int)(c % (uint32_t)g.ni), j = (int)((c / (uint32_t)g.ni) % (uint32_t)g.nj);
        const int k = (int)(c / ((uint32_t)g.ni * (uint32_t)g.nj));
        // row by row in the reference's order kk, jj (then ii); flags and values
        // are loaded unconditionally (neighbours outside the grid read the
        // cell itself and are masked off), so the loads do not wait on the
        // flags. +0.0f for a neighbour that is not valid leaves the sum
        // unchanged (it starts at +0.0f, so it is never -0.0f).
        const uint32_t* pw = reinterpret_cast<const uint32_t*>(pl);
        const uint32_t wpr2 = 2u * (uint32_t)g.wpr;
        const bool has_m = i > 0, has_p = i + 1 < g.ni;
        float sum = 0.0f;
        int cnt = 0;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            const int dj = r % 3 - 1, dk = r / 3 - 1;
            const int jj = j + dj, kk = k + dk;
            const bool row_in = jj >= 0 && jj < g.nj && kk >= 0 && kk < g.nk;
            const uint32_t rc = row_in ? c + (uint32_t)((int)sj * dj + (int)sk * dk) : c;
            const uint32_t rowi = row_in ? (uint32_t)jj + (uint32_t)g.nj * (uint32_t)kk
                                         : (uint32_t)j + (uint32_t)g.nj * (uint32_t)k;
            const uint32_t* rowp = pw + rowi * wpr2;
            const uint32_t im = has_m ? i - 1 : i, ip = has_p ? i + 1 : i;
            // unconditional loads, masked afterwards
            const uint32_t wm = rowp[im >> 5], w0 = rowp[(uint32_t)i >> 5], wp = rowp[ip >> 5];
            const uint32_t bm = (row_in && has_m) ? (wm >> (im & 31)) & 1u : 0u;
            const uint32_t b0 = row_in ? (w0 >> (i & 31)) & 1u : 0u;
            const uint32_t bp = (row_in && has_p) ? (wp >> (ip & 31)) & 1u : 0u;
            const float vm = g.f[has_m ? rc - 1u : rc], v0 = g.f[rc], vp = g.f[has_p ? rc + 1u : rc];
            sum += bm ? vm : 0.0f;
            sum += b0 ? v0 : 0.0f;
            sum += bp ? vp : 0.0f;
            cnt += (int)(bm + b0 + bp);
        }
        g.f[c] = sum / (float)cnt;
        if (g.valid_out) g.valid_out[c] = 1;
    }
}

// simulator.cpp:125-158 for face axis AXIS: valid iff weight != 0, then a face
// inside the loop ranges [1, n_face - 1) whose two cells both have phi >= 0
// is invalid.
template <int AXIS>
__global__ __launch_bounds__(256) void k_find_valid(Dense d, uint8_t* valid) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    const int ni = d.ni, nj = d.nj;
    const int fi = ni + (AXIS == 0), fj = nj + (AXIS == 1), fk = d.nk + (AXIS == 2);
    if (i >= fi || j >= fj) return;
    const size_t f = at(i, j, k, fi, fj);
    const float* wgt = AXIS == 0 ? d.uw : AXIS == 1 ? d.vw : d.ww;
    bool ok = wgt[f] != 0.0f;
    if (i >= 1 && j >= 1 && k >= 1 && i < fi - 1 && j < fj - 1 && k < fk - 1) {
        const size_t c1 = at(i, j, k, ni, nj);
        const size_t c0 = at(i - (AXIS == 0), j - (AXIS == 1), k - (AXIS == 2), ni, nj);
        if (d.phi[c0] >= 0.0f && d.phi[c1] >= 0.0f) ok = false;
    }
    valid[f] = ok ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_zero_invalid(float* f, const uint8_t* valid, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride)
        if (valid[c] == 0) f[c] = 0.0f;
}

// simulator.cpp:191-200 then :160-167. The eight nodes are summed in float in
// the reference's order; the product with the double literal 0.125 is
// rounded to float on assignment.
__global__ __launch_bounds__(256) void k_level_set_valid(int ni, int nj, int nk, const float* nodal,
                                                         uint8_t* valid) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= ni || j >= nj) return;
    const int mi = ni + 1, mj = nj + 1;
    const float s = nodal[at(i, j, k, mi, mj)] + nodal[at(i + 1, j, k, mi, mj)] + nodal[at(i, j + 1, k, mi, mj)] +
                    nodal[at(i + 1, j + 1, k, mi, mj)] + nodal[at(i, j, k + 1, mi, mj)] +
                    nodal[at(i + 1, j, k + 1, mi, mj)] + nodal[at(i, j + 1, k + 1, mi, mj)] +
                    nodal[at(i + 1, j + 1, k + 1, mi, mj)];
    const float centred = (float)((double)s * 0.125);
    valid[at(i, j, k, ni, nj)] = centred < 0.0f ? 0 : 1;
}

// ---------------------------------------------------------------- advection
// A grid with the reference's geometry: value (i, j, k) sits at
// ((i + 0.5) dx, (j + 0.5) dx, (k + 0.5) dx) + origin (grid.h:355-367).
struct Sampled {
    const float* __restrict__ a;
    int ni, nj, nk;
    float o0, o1, o2;
};

// The simulator's grids have origin 0; a face grid is shifted by -dx/2 on
// its own axis: origin - vector_type(0.5, 0, 0) * dx (staggered_grid.h:109-111).
__host__ __device__ inline Sampled face_grid(const float* a, int axis, int ni, int nj, int nk, float dx) {
    Sampled g;
    g.a = a;
    g.ni = ni + (axis == 0);
    g.nj = nj + (axis == 1);
    g.nk = nk + (axis == 2);
    g.o0 = 0.0f - (axis == 0 ? 0.5f : 0.0f) * dx;
    g.o1 = 0.0f - (axis == 1 ? 0.5f : 0.0f) * dx;
    g.o2 = 0.0f - (axis == 2 ? 0.5f : 0.0f) * dx;
    return g;
}

// interpolation.h:10-15
__device__ __forceinline__ int index_and_fraction(float pos, float origin, float dx, float& frac) {
    const float t = (pos - origin - 0.5f * dx) / dx;
    const int i = (int)floorf(t);
    frac = t - (float)i;
    return i;
}

// Index and fraction of one coordinate on one axis of a grid.
struct Coord {
    int i;
    float f;
};

__device__ __forceinline__ Coord coord_at(float pos, float origin, float dx) {
    Coord c;
    c.i = index_and_fraction(pos, origin, dx, c.f);
    return c;
}

__device__ __forceinline__ float lerp2(float v0, float v1, float f) { return (1.0f - f) * v0 + f * v1; }

// interpolation.h:35-72 (dim 3): indices clamped to the grid (grid.h:370-379),
// x interpolated first, then y, then z.
// Indices are 32-bit: every grid has fewer than 2^31 values. The upper
// neighbour on an axis is the lower one or the next, so its offset is a
// select, not a product.
__device__ __forceinline__ float trilerp(const Sampled& g, Coord cx, Coord cy, Coord cz) {
    const int i = cx.i, j = cy.i, k = cz.i;
    const float fx = cx.f, fy = cy.f, fz = cz.f;
    const int i0 = min(max(i, 0), g.ni - 1), i1 = min(max(i + 1, 0), g.ni - 1);
    const int j0 = min(max(j, 0), g.nj - 1), j1 = min(max(j + 1, 0), g.nj - 1);
    const int k0 = min(max(k, 0), g.nk - 1), k1 = min(max(k + 1, 0), g.nk - 1);
    const uint32_t row = (uint32_t)g.ni, plane = (uint32_t)g.ni * (uint32_t)g.nj;
    const uint32_t r00 = (uint32_t)k0 * plane + (uint32_t)j0 * row;
    const uint32_t dj = j1 > j0 ? row : 0u, dk = k1 > k0 ? plane : 0u;
    const uint32_t r10 = r00 + dj, r01 = r00 + dk, r11 = r01 + dj;
    const uint32_t a0 = (uint32_t)i0, a1 = (uint32_t)i1;
    const float v0 = g.a[r00 + a0], v1 = g.a[r00 + a1], v2 = g.a[r10 + a0], v3 = g.a[r10 + a1];
    const float v4 = g.a[r01 + a0], v5 = g.a[r01 + a1], v6 = g.a[r11 + a0], v7 = g.a[r11 + a1];
    const float b0 = lerp2(lerp2(v0, v1, fx), lerp2(v2, v3, fx), fy);
    const float b1 = lerp2(lerp2(v4, v5, fx), lerp2(v6, v7, fx), fy);
    return lerp2(b0, b1, fz);
}

__device__ __forceinline__ float trilerp_at(const Sampled& g, float x, float y, float z, float dx) {
    return trilerp(g, coord_at(x, g.o0, dx), coord_at(y, g.o1, dx), coord_at(z, g.o2, dx));
}

// The velocity's face grids (face_grid, origin 0 shifted by -dx/2 on the
// grid's own axis).
struct VelField {
    Sampled u, v, w;
};

// A point's coordinates on the velocity grids. On each axis one grid is
// shifted (origin 0 - dx/2, s[]) and the two others are not (origin 0, c[]),
// so those two share the division of interpolation.h:13.
struct StagCoords {
    Coord s[3], c[3];
};

__device__ __forceinline__ float shifted_origin(float dx) { return 0.0f - 0.5f * dx; }

__device__ __forceinline__ StagCoords stag_coords(float x, float y, float z, float dx) {
    const float os = shifted_origin(dx);
    StagCoords p;
    p.s[0] = coord_at(x, os, dx);
    p.s[1] = coord_at(y, os, dx);
    p.s[2] = coord_at(z, os, dx);
    p.c[0] = coord_at(x, 0.0f, dx);
    p.c[1] = coord_at(y, 0.0f, dx);
    p.c[2] = coord_at(z, 0.0f, dx);
    return p;
}

// lerp_staggered (interpolation.h:186-193) at the point with coordinates p.
__device__ __forceinline__ void lerp_staggered(const VelField& F, const StagCoords& p, float& vx, float& vy,
                                               float& vz) {
    vx = trilerp(F.u, p.s[0], p.c[1], p.c[2]);
    vy = trilerp(F.v, p.c[0], p.s[1], p.c[2]);
    vz = trilerp(F.w, p.c[0], p.c[1], p.s[2]);
}

// ode.h:30-35 (trace_rk2) with the step -dt (advection.cpp:13,23) from the
// point (x, y, z), whose coordinates on the velocity grids are p0.
__device__ __forceinline__ void backtrace(const VelField& F, const StagCoords& p0, float dx, float dt, float& x,
                                          float& y, float& z) {
    const float h = -dt;
    float vx, vy, vz;
    lerp_staggered(F, p0, vx, vy, vz);
    const float half = 0.5f * h;
    const float mx = x + half * vx, my = y + half * vy, mz = z + half * vz;
    float wx, wy, wz;
    lerp_staggered(F, stag_coords(mx, my, mz, dx), wx, wy, wz);
    x = x + h * wx;
    y = y + h * wy;
    z = z + h * wz;
}

struct Advect {
    int ni, nj, nk;
    float dx, dt;
    VelField field;   // the velocity that carries the values
    Sampled val[3];   // values of the advected grids (face grids, or val[0] a cell grid)
    float* out[3];
};

// advection.cpp:19-27: one thread per (i, j, k) takes the u, v and w faces
// there (those that exist), so the three backtraces gather from the same
// neighbourhood of the velocity grids. Face u starts at (cx + os, cy, cz),
// v at (cx, cy + os, cz) and w at (cx, cy, cz + os), with c* the cell
// centre's coordinates and os the shifted origin: the start points share
// their coordinates two by two, and so do their divisions.
__global__ __launch_bounds__(256) void k_advect_faces(Advect A) {
    const int i = blockIdx.x * 32 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y,
              k = blockIdx.z * 2 + threadIdx.z;
    const float dx = A.dx, os = shifted_origin(dx);
    const float c[3] = {((float)i + 0.5f) * dx, ((float)j + 0.5f) * dx, ((float)k + 0.5f) * dx};
    const float sh[3] = {c[0] + os, c[1] + os, c[2] + os};
    // coordinates of the centre (cc*) and shifted (sc*) values on both kinds of grid
    Coord cc[3][2], sc[3][2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        cc[a][0] = coord_at(c[a], os, dx);
        cc[a][1] = coord_at(c[a], 0.0f, dx);
        sc[a][0] = coord_at(sh[a], os, dx);
        sc[a][1] = coord_at(sh[a], 0.0f, dx);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const Sampled& g = A.val[d];
        if (i >= g.ni || j >= g.nj || k >= g.nk) continue;
        StagCoords p0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            p0.s[a] = a == d ? sc[a][0] : cc[a][0];
            p0.c[a] = a == d ? sc[a][1] : cc[a][1];
        }
        float x = d == 0 ? sh[0] : c[0], y = d == 1 ? sh[1] : c[1], z = d == 2 ? sh[2] : c[2];
        backtrace(A.field, p0, dx, A.dt, x, y, z);
        A.out[d][(size_t)i + (size_t)g.ni * ((size_t)j + (size_t)g.nj * k)] = trilerp_at(g, x, y, z, dx);
    }
}

// advection.cpp:9-17 on the cell grid (origin 0).
__global__ __launch_bounds__(256) void k_advect_cells(Advect A) {
    const int i = blockIdx.x * 32 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y,
              k = blockIdx.z * 2 + threadIdx.z;
    if (i >= A.ni || j >= A.nj || k >= A.nk) return;
    const float dx = A.dx;
    float x = ((float)i + 0.5f) * dx + 0.0f;
    float y = ((float)j + 0.5f) * dx + 0.0f;
    float z = ((float)k + 0.5f) * dx + 0.0f;
    backtrace(A.field, stag_coords(x, y, z, dx), dx, A.dt, x, y, z);
    A.out[0][(size_t)i + (size_t)A.ni * ((size_t)j + (size_t)A.nj * k)] = trilerp_at(A.val[0], x, y, z, dx);
}

// ---------------------------------------------------------- reinitialisation
// The reference marches from the interface one face-neighbour layer per pass:
// FAR cells next to a FIX cell become NEAR, take an upwind distance from the
// FIX neighbours' values (a snapshot, grid_temp), and become FIX. A 16-bit
// stamp per cell stands for the label and the snapshot: FIX at the start of
// pass l is stamp <= l, and a cell fixed in pass l gets l + 1, so the tile
// passes run in place over the tiles they can change. Where the pass of every
// cell is its taxicab distance to the interface (checked cell by cell), the
// passes instead run over per-distance cell lists (the distance route below).

struct P3 {
    float x, y, z;
};

__device__ __forceinline__ P3 sub3(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

// Eigen's cross() for 3-vectors
__device__ __forceinline__ P3 cross3(P3 u, P3 v) {
    return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}

// Eigen's norm() of a 3-vector: the fixed-size reduction splits the squares
// in halves, x0 + (x1 + x2)
__device__ __forceinline__ float norm3(P3 v) { return sqrtf(v.x * v.x + (v.y * v.y + v.z * v.z)); }

// meshing.h:12-17 with isolevel 0
__device__ __forceinline__ P3 vertex_interp(P3 p0, P3 p1, float l0, float l1) {
    const float mu = -(l0 - 0.0f) / (l1 - l0);
    return {p0.x + mu * (p1.x - p0.x), p0.y + mu * (p1.y - p0.y), p0.z + mu * (p1.z - p0.z)};
}

// level_set.h:212-237: distance from x0 to the zero set of the linear
// interpolant on the edges x0-x1, x0-x2, x0-x3 (1, 2 or 3 crossings).
__device__ float tet_distance(P3 x0, P3 x1, P3 x2, P3 x3, float f0, float f1, float f2, float f3) {
    if (f0 == 0.0f) return 0.0f;
    const bool b1 = f0 * f1 <= 0.0f, b2 = f0 * f2 <= 0.0f, b3 = f0 * f3 <= 0.0f;
    const int n = (int)b1 + (int)b2 + (int)b3;
    if (n == 0) return FLT_MAX;
    // edges without a crossing may divide by zero here; their points are not used
    const P3 p1 = vertex_interp(x0, x1, f0, f1), p2 = vertex_interp(x0, x2, f0, f2), p3 = vertex_interp(x0, x3, f0, f3);
    const P3 first = b1 ? p1 : (b2 ? p2 : p3);
    const P3 second = b1 ? (b2 ? p2 : p3) : p3;
    if (n == 1) return norm3(sub3(first, x0));
    if (n == 2) {  // level_set.h:172-177
        const P3 b = sub3(second, first);
        return norm3(cross3(sub3(x0, first), b)) / norm3(b);
    }
    // level_set.h:179-189
    const P3 nv = cross3(sub3(p2, p1), sub3(p3, p1));
    const float d = -(nv.x * p1.x + nv.y * p1.y + nv.z * p1.z);
    return fabsf(nv.x * x0.x + nv.y * x0.y + nv.z * x0.z + d) / sqrtf(nv.x * nv.x + nv.y * nv.y + nv.z * nv.z);
}

// reinitialization.cpp:27-62 for one valid cell: FIX if phi changes sign (or
// touches 0) towards a face neighbour (indices clamped); a FIX cell whose
// phi is not 0 takes calc_sdf over the 8 tetrahedra of its 7-point star
// (level_set.h:251-264). Star positions are Vector3f((i + 0.5) * dx, ...):
// double products rounded to float. The caller gives p and its six face
// neighbours (indices clamped to the grid).
__device__ bool rei_interface_at(const ReiSet& r, int i, int j, int k, float p, float pxm, float pxp, float pym,
                                 float pyp, float pzm, float pzp, float& out) {
    const bool flag = p * pxm <= 0.0f || p * pxp <= 0.0f || p * pym <= 0.0f || p * pyp <= 0.0f ||
                      p * pzm <= 0.0f || p * pzp <= 0.0f;
    if (!flag) return false;
    out = p;
    if (p != 0.0f) {
        const double d = (double)r.dx;
        const float xc = (float)((i + 0.5) * d), yc = (float)((j + 0.5) * d), zc = (float)((k + 0.5) * d);
        const P3 x[7] = {{xc, yc, zc},
                         {(float)(((i + 1) + 0.5) * d), yc, zc},
                         {xc, (float)(((j + 1) + 0.5) * d), zc},
                         {(float)(((i - 1) + 0.5) * d), yc, zc},
                         {xc, (float)(((j - 1) + 0.5) * d), zc},
                         {xc, yc, (float)(((k + 1) + 0.5) * d)},
                         {xc, yc, (float)(((k - 1) + 0.5) * d)}};
        const float f[7] = {p, pxp, pyp, pxm, pym, pzp, pzm};
        float l = FLT_MAX;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int a = q + 1, b = (q + 1) % 4 + 1;  // star pairs (1,2) (2,3) (3,4) (4,1)
#pragma unroll
            for (int zz = 5; zz <= 6; ++zz) {
                const float dist = tet_distance(x[0], x[a], x[b], x[zz], f[0], f[a], f[b], f[zz]);
                l = (dist < l) ? dist : l;  // std::min(l, dist)
            }
        }
        out = p < 0.0f ? -l : l;
    }
    return true;
}

__device__ __forceinline__ TileXYZ rei_tile(const ReiSet& r, int t) {
    TileXYZ T;
    T.t = t;
    T.bx = t % r.tx;
    T.by = (t / r.tx) % r.ty;
    T.bz = t / (r.tx * r.ty);
    return T;
}

// One workgroup per tile: interface cells, stamps, tfar and tnew. A thread
// takes a z column of the tile and issues the loads of all its cells first
// (flags, the column with one cell beyond each end, the four side columns;
// indices clamped, rows past the grid read row nk - 1 and are not used).
__global__ __launch_bounds__(256) void k_rei_init(ReiSet r) {
    const TileXYZ T = rei_tile(r, blockIdx.x);
    const int i = T.bx * kExtTX + (threadIdx.x & (kExtTX - 1));
    const int j = T.by * kExtTY + (threadIdx.x / kExtTX);
    bool far_any = false, fix_any = false;
    if (i < r.ni && j < r.nj) {
        const int ni = r.ni, nj = r.nj, nk = r.nk, k0 = T.bz * kExtTZ;
        const float* t = r.temp;
        const int im = max(i - 1, 0), ip = min(i + 1, ni - 1), jm = max(j - 1, 0), jp = min(j + 1, nj - 1);
        uint8_t val[kExtTZ];
        float col[kExtTZ + 2], sxm[kExtTZ], sxp[kExtTZ], sym[kExtTZ], syp[kExtTZ];
#pragma unroll
        for (int z = 0; z < kExtTZ + 2; ++z) {
            const int k = min(max(k0 + z - 1, 0), nk - 1);
            col[z] = t[at(i, j, k, ni, nj)];
        }
#pragma unroll
        for (int z = 0; z < kExtTZ; ++z) {
            const int k = min(k0 + z, nk - 1);
            val[z] = r.valid[at(i, j, k, ni, nj)];
            sxm[z] = t[at(im, j, k, ni, nj)];
            sxp[z] = t[at(ip, j, k, ni, nj)];
            sym[z] = t[at(i, jm, k, ni, nj)];
            syp[z] = t[at(i, jp, k, ni, nj)];
        }
#pragma unroll
        for (int z = 0; z < kExtTZ; ++z) {
            const int k = k0 + z;
            if (k >= nk) break;
            const size_t c = at(i, j, k, ni, nj);
            uint16_t st;
            float v;
            if (val[z] == 0) {
                st = kReiUnknown;
            } else if (rei_interface_at(r, i, j, k, col[z + 1], sxm[z], sxp[z], sym[z], syp[z], col[z], col[z + 2],
                                        v)) {
                r.phi[c] = v;
                st = 0;
                fix_any = true;
            } else {
                st = kReiFar;
                far_any = true;
            }
            r.stamp[c] = st;
        }
    }
    far_any = __syncthreads_or(far_any);
    fix_any = __syncthreads_or(fix_any);
    if (threadIdx.x == 0) {
        r.tfar[T.t] = far_any ? 1 : 0;
        r.tnew[T.t] = fix_any ? 0 : kReiFar;
        if (blockIdx.x == 0)
            for (int q = 0; q < 9; ++q) r.counter[q] = 0;
    }
    for (int q = blockIdx.x * 256 + threadIdx.x; q < kReiDistMax + 2; q += gridDim.x * 256) r.dcount[q] = 0;
}

// One thread per tile: the tiles pass l can change are listed. A cell not
// reached yet is fixed in the first pass that sees a FIX neighbour, so pass l
// can only fix cells next to cells fixed just before it (stamp == l): tiles
// with a cell not reached yet whose own tnew or a face neighbour's is l.
__global__ __launch_bounds__(256) void k_rei_select(ReiSet r, int l, int parity) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) r.counter[2 + parity] = 0;
    bool act = false;
    if (t < r.tiles && r.tfar[t] != 0) {
        const TileXYZ T = rei_tile(r, t);
        const uint16_t lv = (uint16_t)l;
        act = r.tnew[t] == lv;
        const int off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int x = T.bx + off[q][0], y = T.by + off[q][1], z = T.bz + off[q][2];
            if (x >= 0 && y >= 0 && z >= 0 && x < r.tx && y < r.ty && z < r.tz)
                act |= r.tnew[x + r.tx * (y + r.ty * z)] == lv;
        }
    }
    const uint64_t m = __ballot(act);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&r.counter[parity], (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (act) r.list[base + __popcll(m & ((1ull << lane) - 1))] = t;
}

// reinitialization.cpp:80-140 for one NEAR cell: one FIX value per axis (-x
// before +x), signs folded, sorted, then the 1-, 2- or 3-neighbour upwind
// distance with the reference's float and double steps. Asserts are off in
// the reference's release build: mixed signs proceed unflipped.
__device__ float rei_update(bool hx, float vx, bool hy, float vy, bool hz, float vz, float dx) {
    const int m = (int)hx + (int)hy + (int)hz;
    float a0 = hx ? vx : (hy ? vy : vz);
    float a1 = hx ? (hy ? vy : vz) : vz;
    float a2 = vz;
    const bool all_ge = (!hx || vx >= 0.0f) && (!hy || vy >= 0.0f) && (!hz || vz >= 0.0f);
    const bool all_le = (!hx || vx <= 0.0f) && (!hy || vy <= 0.0f) && (!hz || vz <= 0.0f);
    const bool flip = !all_ge && all_le;
    if (flip) {
        a0 = -a0;
        a1 = -a1;
        a2 = -a2;
    }
    // ascending, as an insertion sort
    if (m >= 2 && a1 < a0) {
        const float t = a0;
        a0 = a1;
        a1 = t;
    }
    if (m == 3 && a2 < a1) {
        float t = a1;
        a1 = a2;
        a2 = t;
        if (a1 < a0) {
            t = a0;
            a0 = a1;
            a1 = t;
        }
    }
    float d = a0 + dx;
    if (m >= 2 && d > a1) d = (float)(0.5 * (double)(a0 + a1 + sqrtf(2 * dx * dx - (a1 - a0) * (a1 - a0))));
    if (m == 3 && d > a2) {
        const float s = a0 + a1 + a2;
        float q = s * s - 3.0f * (a0 * a0 + a1 * a1 + a2 * a2 - dx * dx);
        q = (0.0f < q) ? q : 0.0f;  // std::max(0.0f, q)
        d = (float)(0.333333 * (double)(a0 + a1 + a2 + sqrtf(q)));
    }
    return flip ? -d : d;
}

// One pass over the listed tiles (persistent). A tile and a one-cell halo of
// stamps and values are staged in LDS; cells outside the grid count as not
// FIX, as the reference's clamped label lookup returns the cell itself.
__global__ __launch_bounds__(256) void k_rei_pass(ReiSet r, int l, int parity) {
    constexpr int kPer = (kHalo + 255) / 256;
    __shared__ uint16_t s_st[kPer * 256];
    __shared__ float s_v[kPer * 256];
    const uint32_t count = r.counter[parity];
    if (blockIdx.x == 0 && threadIdx.x == 0) r.counter[parity ^ 1] = 0;
    const int lx = threadIdx.x & (kExtTX - 1), ly = threadIdx.x / kExtTX;
    int t_next = blockIdx.x < count ? r.list[blockIdx.x] : 0;
    for (uint32_t idx = blockIdx.x; idx < count; idx += gridDim.x) {
        const int tt = t_next;
        if (idx + gridDim.x < count) t_next = r.list[idx + gridDim.x];
        const TileXYZ T = rei_tile(r, tt);
        const int i0 = T.bx * kExtTX - 1, j0 = T.by * kExtTY - 1, k0 = T.bz * kExtTZ - 1;
        uint16_t st[kPer];
        float fv[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int h = threadIdx.x + 256 * q;
            const int hx = h % kHX, hy = (h / kHX) % kHY, hz = h / (kHX * kHY);
            const int i = i0 + hx, j = j0 + hy, k = k0 + hz;
            const bool in = h < kHalo && i >= 0 && j >= 0 && k >= 0 && i < r.ni && j < r.nj && k < r.nk;
            const size_t c = in ? at(i, j, k, r.ni, r.nj) : 0;
            st[q] = in ? r.stamp[c] : kReiFar;
            fv[q] = in ? r.phi[c] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            s_st[threadIdx.x + 256 * q] = st[q];
            s_v[threadIdx.x + 256 * q] = fv[q];
        }
        __syncthreads();

        const int i = T.bx * kExtTX + lx, j = T.by * kExtTY + ly;
        bool still = false, became = false;
        uint32_t fixed = 0;
        if (i < r.ni && j < r.nj) {
            for (int z = 0; z < kExtTZ; ++z) {
                const int k = T.bz * kExtTZ + z;
                if (k >= r.nk) break;
                const int h = (lx + 1) + kHX * ((ly + 1) + kHY * (z + 1));
                if (s_st[h] != kReiFar) continue;
                const bool fxm = s_st[h - 1] <= l, fxp = s_st[h + 1] <= l;
                const bool fym = s_st[h - kHX] <= l, fyp = s_st[h + kHX] <= l;
                const bool fzm = s_st[h - kHX * kHY] <= l, fzp = s_st[h + kHX * kHY] <= l;
                if (!(fxm || fxp || fym || fyp || fzm || fzp)) {
                    still = true;
                    continue;
                }
                const float vx = fxm ? s_v[h - 1] : s_v[h + 1];
                const float vy = fym ? s_v[h - kHX] : s_v[h + kHX];
                const float vz = fzm ? s_v[h - kHX * kHY] : s_v[h + kHX * kHY];
                const size_t c = at(i, j, k, r.ni, r.nj);
                r.phi[c] = rei_update(fxm || fxp, vx, fym || fyp, vy, fzm || fzp, vz, r.dx);
                r.stamp[c] = (uint16_t)(l + 1);
                became = true;
                ++fixed;
            }
        }
        still = __syncthreads_or(still);
        became = __syncthreads_or(became);
        // cells fixed by this tile, summed over its waves
        fixed = __reduce_add_sync(~0ull, fixed);
        if ((threadIdx.x & 63) == 0 && fixed) atomicAdd(&r.counter[2 + parity], fixed);
        if (threadIdx.x == 0) r.tfar[tt] = still ? 1 : 0;
        if (became && threadIdx.x == 0) r.tnew[tt] = (uint16_t)(l + 1);
        __syncthreads();  // LDS is free for the next tile
    }
}

// ---- passes ordered by distance
// Pass l fixes exactly the valid cells whose shortest path to an interface
// cell through valid cells has l + 1 steps. Where no invalid cell lies in the
// way, that length is the taxicab distance, computed by three separable
// passes; a check proves it cell by cell (a valid non-interface cell at
// distance d has a valid neighbour at d - 1, so by induction a valid path of
// d steps exists, and no path is shorter than the taxicab distance). Then
// each pass touches only its own cells. Otherwise the tile passes run.
constexpr uint16_t kDistInf = 0xFFFF;
constexpr int kDistBig = 1 << 28;

__device__ __forceinline__ uint16_t dist_step(uint16_t prev) { return prev >= kDistInf - 1 ? kDistInf : prev + 1; }

// x: one wave per grid row, prefix and suffix minima over 64-cell chunks:
// d(i) = min(i + min_{i' <= i}(g - i'), min_{i' >= i}(g + i') - i), g = 0 at
// interface cells.
__global__ __launch_bounds__(256) void k_rei_dist_x(ReiSet r) {
    const int line = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (line >= r.nj * r.nk) return;
    const int lane = threadIdx.x & 63;
    const size_t row = (size_t)line * (size_t)r.ni;
    int carry = kDistBig;
    for (int c0 = 0; c0 < r.ni; c0 += 64) {
        const int i = c0 + lane;
        int h = (i < r.ni && r.stamp[row + i] == 0) ? -i : kDistBig;
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(h, o, 64);
            if (lane >= o) h = min(h, t);
        }
        h = min(h, carry);
        carry = __shfl(h, 63, 64);
        if (i < r.ni) r.dist[row + i] = h >= kDistBig / 2 ? kDistInf : (uint16_t)(i + h);
    }
    carry = kDistBig;
    for (int c0 = ((r.ni - 1) / 64) * 64; c0 >= 0; c0 -= 64) {
        const int i = c0 + lane;
        int h = (i < r.ni && r.stamp[row + i] == 0) ? i : kDistBig;
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_down(h, o, 64);
            if (lane + o < 64) h = min(h, t);
        }
        h = min(h, carry);
        carry = __shfl(h, 0, 64);
        if (i < r.ni && h < kDistBig / 2) {
            const uint16_t b = (uint16_t)(h - i);
            if (b < r.dist[row + i]) r.dist[row + i] = b;
        }
    }
}

// y (AXIS 1) or z (AXIS 2): one thread per column, a forward and a backward
// running minimum, loads in chunks of 16 so they overlap.
template <int AXIS>
__global__ __launch_bounds__(256) void k_rei_dist_yz(ReiSet r) {
    const int i = blockIdx.x * 64 + (int)(threadIdx.x & 63);
    const int o = blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    const int n = AXIS == 1 ? r.nj : r.nk;
    if (i >= r.ni || o >= (AXIS == 1 ? r.nk : r.nj)) return;
    const size_t stride = AXIS == 1 ? (size_t)r.ni : (size_t)r.ni * r.nj;
    const size_t base = AXIS == 1 ? (size_t)i + (size_t)r.ni * r.nj * o : (size_t)i + (size_t)r.ni * o;
    uint16_t* __restrict__ d = r.dist + base;
    constexpr int C = 16;
    uint16_t prev = kDistInf;
    for (int a0 = 0; a0 < n; a0 += C) {
        uint16_t v[C];
#pragma unroll
        for (int q = 0; q < C; ++q) v[q] = a0 + q < n ? d[(size_t)(a0 + q) * stride] : kDistInf;
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const uint16_t s1 = dist_step(prev);
            v[q] = v[q] < s1 ? v[q] : s1;
            prev = v[q];
        }
#pragma unroll
        for (int q = 0; q < C; ++q)
            if (a0 + q < n) d[(size_t)(a0 + q) * stride] = v[q];
    }
    prev = kDistInf;
    for (int a1 = n; a1 > 0; a1 -= C) {
        uint16_t v[C];
#pragma unroll
        for (int q = 0; q < C; ++q) v[q] = a1 - 1 - q >= 0 ? d[(size_t)(a1 - 1 - q) * stride] : kDistInf;
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const uint16_t s1 = dist_step(prev);
            v[q] = v[q] < s1 ? v[q] : s1;
            prev = v[q];
        }
#pragma unroll
        for (int q = 0; q < C; ++q)
            if (a1 - 1 - q >= 0) d[(size_t)(a1 - 1 - q) * stride] = v[q];
    }
}

// Whether a neighbour (in the grid, valid) lies at distance want.
// Branch-free: a neighbour outside the grid reads the cell itself (c) and is
// masked off, so the loads of several lookups go out together.
__device__ __forceinline__ bool rei_at_dist(const ReiSet& r, uint32_t c, bool inside, int32_t delta,
                                            uint16_t want) {
    const uint32_t nb = inside ? (uint32_t)((int32_t)c + delta) : c;
    const uint16_t st = r.stamp[nb], dv = r.dist[nb];
    return inside & (st != kReiUnknown) & (dv == want);
}

// The six face neighbours of cell c = (i, j, k) at distance want.
__device__ __forceinline__ uint32_t rei_fix6(const ReiSet& r, uint32_t c, int i, int j, int k, uint16_t want) {
    const int32_t sj = r.ni, sk = r.ni * r.nj;
    return (uint32_t)rei_at_dist(r, c, i > 0, -1, want) | (uint32_t)rei_at_dist(r, c, i + 1 < r.ni, 1, want) << 1 |
           (uint32_t)rei_at_dist(r, c, j > 0, -sj, want) << 2 | (uint32_t)rei_at_dist(r, c, j + 1 < r.nj, sj, want) << 3 |
           (uint32_t)rei_at_dist(r, c, k > 0, -sk, want) << 4 | (uint32_t)rei_at_dist(r, c, k + 1 < r.nk, sk, want) << 5;
}

// The cells not reached yet are bucketed by distance per chunk of
// kReiBucketCells cells (one workgroup): distances inside a chunk span less
// than kReiBins (neighbouring cells differ by at most 1, and a chunk covers at
// most a grid row plus a few cells of the next ones), so a chunk counts its
// cells per distance in LDS, with one global atomic per distinct distance.
constexpr int kReiBucketThreads = 1024;
constexpr int kReiBucketPer = 8;
constexpr int kReiBucketCells = kReiBucketThreads * kReiBucketPer;
constexpr int kReiBins = 8192;

struct ReiChunk {
    size_t c[kReiBucketPer];
    uint32_t d[kReiBucketPer];   // distance of a cell not reached yet, else 0
    uint32_t rank[kReiBucketPer];
};

// Loads a chunk's distances (0 for cells that are not FAR), its smallest
// distance (*dmin) and the span; sets LDS bins[0..span) to the per-distance
// counts and rank[] to each cell's place among equal distances. Returns the
// span, or 0 when the chunk holds no FAR cell.
__device__ int rei_chunk_bins(const ReiSet& r, ReiChunk& ch, uint32_t* bins, uint32_t* red, uint32_t& dmin) {
    const size_t n = (size_t)r.ni * r.nj * r.nk;
    const size_t c0 = (size_t)blockIdx.x * kReiBucketCells;
    uint32_t lo = 0xFFFFFFFFu, hi = 0;
#pragma unroll
    for (int q = 0; q < kReiBucketPer; ++q) {
        ch.c[q] = c0 + (size_t)q * kReiBucketThreads + threadIdx.x;
        const bool far = ch.c[q] < n && r.stamp[ch.c[q]] == kReiFar;
        ch.d[q] = far ? r.dist[ch.c[q]] : 0u;
        if (far) {
            lo = min(lo, ch.d[q]);
            hi = max(hi, ch.d[q]);
        }
    }
    lo = __reduce_min_sync(~0ull, lo);
    hi = __reduce_max_sync(~0ull, hi);
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = lo;
        red[32 + (threadIdx.x >> 6)] = hi;
    }
    __syncthreads();
    lo = 0xFFFFFFFFu;
    hi = 0;
    for (int w = 0; w < kReiBucketThreads / 64; ++w) {
        lo = min(lo, red[w]);
        hi = max(hi, red[32 + w]);
    }
    if (lo > hi) return 0;
    dmin = lo;
    const int span = (int)(hi - lo + 1);
    if (span > kReiBins) return -1;
    for (int b = threadIdx.x; b < span; b += kReiBucketThreads) bins[b] = 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kReiBucketPer; ++q)
        if (ch.d[q]) ch.rank[q] = atomicAdd(&bins[ch.d[q] - lo], 1u);
    __syncthreads();
    return span;
}

// Check, and counts per distance of the cells that pass it.
__global__ __launch_bounds__(kReiBucketThreads) void k_rei_dist_check(ReiSet r) {
    __shared__ uint32_t bins[kReiBins];
    __shared__ uint32_t red[64];
    ReiChunk ch;
    uint32_t bad = 0;
    // the check: a FAR cell needs a finite distance >= 1 and a valid neighbour one step closer
#pragma unroll
    for (int q = 0; q < kReiBucketPer; ++q) {
        const size_t n = (size_t)r.ni * r.nj * r.nk;
        const size_t c = (size_t)blockIdx.x * kReiBucketCells + (size_t)q * kReiBucketThreads + threadIdx.x;
        if (c >= n || r.stamp[c] != kReiFar) continue;
        const uint16_t dv = r.dist[c];
        bool ok = dv != 0 && dv != kDistInf;
        if (ok) {
            const uint32_t cu = (uint32_t)c, row = cu / (uint32_t)r.ni;
            const int i = (int)(cu - row * (uint32_t)r.ni), j = (int)(row % (uint32_t)r.nj),
                      k = (int)(row / (uint32_t)r.nj);
            ok = rei_fix6(r, cu, i, j, k, (uint16_t)(dv - 1)) != 0;
        }
        if (!ok) ++bad;
    }
    bad = __reduce_add_sync(~0ull, bad);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(&r.counter[6], bad);
    uint32_t dmin = 0;
    const int span = rei_chunk_bins(r, ch, bins, red, dmin);
    if (span < 0) {
        if (threadIdx.x == 0) atomicAdd(&r.counter[6], 1u);  // not expected; the tile passes run instead
        return;
    }
    for (int b = threadIdx.x; b < span; b += kReiBucketThreads)
        if (bins[b] && dmin + b <= (uint32_t)kReiDistMax) atomicAdd(&r.dcount[dmin + b], bins[b]);
}

// Exclusive scan of dcount[0..kReiDistMax + 1] in place (one workgroup);
// dfill gets the same offsets, counter[7] the largest distance counted.
__global__ __launch_bounds__(1024) void k_rei_dscan(ReiSet r) {
    constexpr int N = kReiDistMax + 2, PER = (N + 1023) / 1024;
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = t * PER + q;
        v[q] = idx < N ? r.dcount[idx] : 0u;
        sum += v[q];
    }
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t add = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    int top = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = t * PER + q;
        if (idx < N) {
            r.dcount[idx] = run;
            r.dfill[idx] = run;
        }
        if (v[q]) top = idx;
        run += v[q];
    }
    // counter[7]: the largest distance with a cell (the pass count)
    __syncthreads();
    part[t] = (uint32_t)top;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (t < o) part[t] = max(part[t], part[t + o]);
        __syncthreads();
    }
    if (t == 0) r.counter[7] = part[0];
}

// Lists the cells not reached yet by distance (after a passed check, every
// one has a distance in 1..kReiDistMax): a chunk reserves one run of slots
// per distance and keeps its cells together inside it.
__global__ __launch_bounds__(kReiBucketThreads) void k_rei_scatter(ReiSet r) {
    __shared__ uint32_t bins[kReiBins];
    __shared__ uint32_t red[64];
    ReiChunk ch;
    uint32_t dmin = 0;
    const int span = rei_chunk_bins(r, ch, bins, red, dmin);
    if (span <= 0) return;
    for (int b = threadIdx.x; b < span; b += kReiBucketThreads)
        if (bins[b]) bins[b] = atomicAdd(&r.dfill[dmin + b], bins[b]);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kReiBucketPer; ++q) {
        if (!ch.d[q]) continue;
        const uint32_t cu = (uint32_t)ch.c[q], row = cu / (uint32_t)r.ni;
        const int i = (int)(cu - row * (uint32_t)r.ni), j = (int)(row % (uint32_t)r.nj),
                  k = (int)(row / (uint32_t)r.nj);
        const uint32_t e = bins[ch.d[q] - dmin] + ch.rank[q];
        r.cells[e] = cu;
        r.fixm[e] = (uint8_t)rei_fix6(r, cu, i, j, k, (uint16_t)(ch.d[q] - 1));
    }
}

// One pass, one listed cell per thread: the cell takes its FIX neighbours
// (valid, distance target - 1, listed beside it by k_rei_scatter) exactly as
// k_rei_pass does.
__global__ __launch_bounds__(256) void k_rei_layer(ReiSet r, uint32_t begin, uint32_t end) {
    const uint32_t e = begin + blockIdx.x * 256 + threadIdx.x;
    if (e >= end) return;
    const size_t c = r.cells[e];
    const uint32_t fx = r.fixm[e];
    const bool fxm = fx & 1u, fxp = fx & 2u, fym = fx & 4u, fyp = fx & 8u, fzm = fx & 16u, fzp = fx & 32u;
    const size_t sx = 1, sy = (size_t)r.ni, sz = (size_t)r.ni * r.nj;
    const float vx = fxm ? r.phi[c - sx] : (fxp ? r.phi[c + sx] : 0.0f);
    const float vy = fym ? r.phi[c - sy] : (fyp ? r.phi[c + sy] : 0.0f);
    const float vz = fzm ? r.phi[c - sz] : (fzp ? r.phi[c + sz] : 0.0f);
    r.phi[c] = rei_update(fxm || fxp, vx, fym || fyp, vy, fzm || fzp, vz, r.dx);
}

__global__ __launch_bounds__(256) void k_rei_rebase(ReiSet r, uint32_t base) {
    const size_t n = (size_t)r.ni * r.nj * r.nk;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    uint32_t top = 0;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride) {
        const uint16_t s = r.stamp[c];
        if (s != kReiFar && s != kReiUnknown) {
            top = max(top, (uint32_t)s);
            r.stamp[c] = 0;
        }
    }
    // passes that fixed cells up to here: a cell fixed in pass a of the
    // epoch starting at pass base has stamp a - base + 1
    top = __reduce_max_sync(~0ull, top);
    if ((threadIdx.x & 63) == 0 && top) atomicMax(&r.counter[8], base + top);
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < (size_t)r.tiles; t += stride) {
        if (r.tnew[t] != kReiFar) r.tnew[t] = 0;  // every fixed cell now has stamp 0
    }
}

__global__ __launch_bounds__(256) void k_rei_finish(ReiSet r) {
    const size_t n = (size_t)r.ni * r.nj * r.nk;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    uint32_t far = 0, top = 0;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride) {
        const uint16_t s = r.stamp[c];
        if (s == kReiFar) ++far;
        else if (s != kReiUnknown) top = max(top, (uint32_t)s);
    }
    far = __reduce_add_sync(~0ull, far);
    top = __reduce_max_sync(~0ull, top);
    if ((threadIdx.x & 63) == 0) {
        if (far) atomicAdd(&r.counter[4], far);
        if (top) atomicMax(&r.counter[5], top);
    }
}

}  // namespace

hipError_t launch_ext_init(const ExtSet& set, hipStream_t s) {
    hipLaunchKernelGGL(k_ext_init, dim3(cdiv(set.words, 4 * 8)), dim3(256), 0, s, set);
    return hipGetLastError();
}

hipError_t launch_ext_layer(const ExtSet& set, int l, hipStream_t s) {
    hipLaunchKernelGGL(k_ext_dilate, dim3(cdiv(set.words, kExtDilateThreads)), dim3(kExtDilateThreads), 0, s, set, l);
    hipLaunchKernelGGL(k_ext_value, dim3(1024, (unsigned)set.count), dim3(256), 0, s, set, l);
    return hipGetLastError();
}

namespace {
VelField make_field(int ni, int nj, int nk, float dx, const float* u, const float* v, const float* w) {
    VelField F;
    F.u = face_grid(u, 0, ni, nj, nk, dx);
    F.v = face_grid(v, 1, ni, nj, nk, dx);
    F.w = face_grid(w, 2, ni, nj, nk, dx);
    return F;
}
}  // namespace

hipError_t launch_advect_faces(int ni, int nj, int nk, float dx, float dt, const float* vel_u, const float* vel_v,
                               const float* vel_w, const float* val_u, const float* val_v, const float* val_w,
                               float* out_u, float* out_v, float* out_w, hipStream_t s) {
    Advect A{};
    A.ni = ni;
    A.nj = nj;
    A.nk = nk;
    A.dx = dx;
    A.dt = dt;
    A.field = make_field(ni, nj, nk, dx, vel_u, vel_v, vel_w);
    const VelField V = make_field(ni, nj, nk, dx, val_u, val_v, val_w);
    A.val[0] = V.u;
    A.val[1] = V.v;
    A.val[2] = V.w;
    A.out[0] = out_u;
    A.out[1] = out_v;
    A.out[2] = out_w;
    hipLaunchKernelGGL(k_advect_faces, dim3(cdiv(ni + 1, 32), cdiv(nj + 1, 4), cdiv(nk + 1, 2)), dim3(32, 4, 2), 0, s,
                       A);
    return hipGetLastError();
}

hipError_t launch_advect_cells(int ni, int nj, int nk, float dx, float dt, const float* vel_u, const float* vel_v,
                               const float* vel_w, const float* val, float* out, hipStream_t s) {
    Advect A{};
    A.ni = ni;
    A.nj = nj;
    A.nk = nk;
    A.dx = dx;
    A.dt = dt;
    A.field = make_field(ni, nj, nk, dx, vel_u, vel_v, vel_w);
    A.val[0] = Sampled{val, ni, nj, nk, 0.0f, 0.0f, 0.0f};
    A.out[0] = out;
    hipLaunchKernelGGL(k_advect_cells, dim3(cdiv(ni, 32), cdiv(nj, 4), cdiv(nk, 2)), dim3(32, 4, 2), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_rei_init(const ReiSet& r, hipStream_t s) {
    hipLaunchKernelGGL(k_rei_init, dim3((unsigned)r.tiles), dim3(256), 0, s, r);
    return hipGetLastError();
}

hipError_t launch_rei_pass(const ReiSet& r, int l, int parity, hipStream_t s) {
    hipLaunchKernelGGL(k_rei_select, dim3(cdiv(r.tiles, 256)), dim3(256), 0, s, r, l, parity);
    const unsigned groups = (unsigned)std::min<int64_t>(r.tiles, kExtLayerGroups);
    hipLaunchKernelGGL(k_rei_pass, dim3(groups), dim3(256), 0, s, r, l, parity);
    return hipGetLastError();
}

hipError_t launch_rei_dist(const ReiSet& r, hipStream_t s) {
    hipLaunchKernelGGL(k_rei_dist_x, dim3(cdiv((int64_t)r.nj * r.nk, 4)), dim3(256), 0, s, r);
    hipLaunchKernelGGL(k_rei_dist_yz<1>, dim3(cdiv(r.ni, 64), cdiv(r.nk, 4)), dim3(256), 0, s, r);
    hipLaunchKernelGGL(k_rei_dist_yz<2>, dim3(cdiv(r.ni, 64), cdiv(r.nj, 4)), dim3(256), 0, s, r);
    const int64_t n = (int64_t)r.ni * r.nj * r.nk;
    hipLaunchKernelGGL(k_rei_dist_check, dim3(cdiv(n, kReiBucketCells)), dim3(kReiBucketThreads), 0, s, r);
    hipLaunchKernelGGL(k_rei_dscan, dim3(1), dim3(1024), 0, s, r);
    return hipGetLastError();
}

hipError_t launch_rei_lists(const ReiSet& r, hipStream_t s) {
    hipLaunchKernelGGL(k_rei_scatter, dim3(cdiv((int64_t)r.ni * r.nj * r.nk, kReiBucketCells)),
                       dim3(kReiBucketThreads), 0, s, r);
    return hipGetLastError();
}

hipError_t launch_rei_layer(const ReiSet& r, uint32_t begin, uint32_t end, hipStream_t s) {
    if (end <= begin) return hipSuccess;
    hipLaunchKernelGGL(k_rei_layer, dim3(cdiv((int64_t)(end - begin), 256)), dim3(256), 0, s, r, begin, end);
    return hipGetLastError();
}

hipError_t launch_rei_rebase(const ReiSet& r, uint32_t base, hipStream_t s) {
    hipLaunchKernelGGL(k_rei_rebase, dim3(2048), dim3(256), 0, s, r, base);
    return hipGetLastError();
}

hipError_t launch_rei_finish(const ReiSet& r, hipStream_t s) {
    hipLaunchKernelGGL(k_rei_finish, dim3(2048), dim3(256), 0, s, r);
    return hipGetLastError();
}

hipError_t launch_find_valid(const Dense& d, uint8_t* uv, uint8_t* vv, uint8_t* wv, hipStream_t s) {
    const dim3 block(64, 4);
    hipLaunchKernelGGL(k_find_valid<0>, dim3(cdiv(d.ni + 1, 64), cdiv(d.nj, 4), d.nk), block, 0, s, d, uv);
    hipLaunchKernelGGL(k_find_valid<1>, dim3(cdiv(d.ni, 64), cdiv(d.nj + 1, 4), d.nk), block, 0, s, d, vv);
    hipLaunchKernelGGL(k_find_valid<2>, dim3(cdiv(d.ni, 64), cdiv(d.nj, 4), d.nk + 1), block, 0, s, d, wv);
    return hipGetLastError();
}

hipError_t launch_zero_invalid(float* f, const uint8_t* valid, size_t n, hipStream_t s) {
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_zero_invalid, dim3(blocks), dim3(256), 0, s, f, valid, n);
    return hipGetLastError();
}

hipError_t launch_level_set_valid(int ni, int nj, int nk, const float* solid_nodal, uint8_t* valid,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_level_set_valid, dim3(cdiv(ni, 64), cdiv(nj, 4), nk), dim3(64, 4), 0, s, ni, nj, nk,
                       solid_nodal, valid);
    return hipGetLastError();
}

}  // namespace elq
