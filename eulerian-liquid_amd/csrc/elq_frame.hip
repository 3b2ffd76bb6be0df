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
// All of it is byte and float streaming work, bound by HBM (or, for the
// extrapolation layers, by how few tiles are active); results are bit-exact
// with the reference (built with -ffp-contract=off; the float division is
// IEEE round-to-nearest).
//
// Extrapolation. In the reference each layer takes a snapshot of the flags
// and of the values, gives every invalid cell with at least one valid
// neighbour (of the 26 around it) the mean of those neighbours, and marks it
// valid. Here each cell carries a byte stamp instead of its flag: 0 for a
// cell valid from the start, l + 1 for a cell that received its value in
// layer l, 255 for a cell still invalid. "Valid at the start of layer l" is
// stamp <= l, and a cell written in layer l gets stamp l + 1 > l, so the
// layer reads its snapshot in place: no reader in layer l looks at a value or
// a stamp that layer l changes the meaning of, and no copy of the grid is
// needed. The same holds one level up for 32x8x8 tiles: tinv says a tile
// still holds an invalid cell, tfv is the first layer at whose start it holds
// a valid one; a layer skips every tile with no invalid cell or with no valid
// cell in its 3x3x3 tile neighbourhood, which leaves such a tile unchanged in
// the reference too.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "elq_frame.hpp"

namespace elq {
namespace {

constexpr uint8_t kNever = 255;
constexpr int kHX = kExtTX + 2, kHY = kExtTY + 2, kHZ = kExtTZ + 2;
constexpr int kHalo = kHX * kHY * kHZ;
static_assert(kExtTX == 32 && kExtTX * kExtTY == 256, "one thread per (x, y) column of a tile");
static_assert(kExtRebase < kNever, "stamps of one stretch of layers fit below the invalid mark");

__device__ __forceinline__ size_t at(int i, int j, int k, int ni, int nj) {
    return (size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * (size_t)k);
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// The grid of the set that owns global tile gt (tile0 ascending), by value:
// the fields stay scalar loads from the kernel arguments.
__device__ __forceinline__ ExtGrid pick(const ExtSet& s, int gt) {
    if (s.count > 2 && gt >= s.g[2].tile0) return s.g[2];
    if (s.count > 1 && gt >= s.g[1].tile0) return s.g[1];
    return s.g[0];
}

struct TileXYZ {
    int t, bx, by, bz;
};

__device__ __forceinline__ TileXYZ tile_of(const ExtGrid& g, int gt) {
    TileXYZ r;
    r.t = gt - g.tile0;
    r.bx = r.t % g.tx;
    r.by = (r.t / g.tx) % g.ty;
    r.bz = r.t / (g.tx * g.ty);
    return r;
}

// One workgroup per tile: stamps from the caller's flags, tinv and tfv.
// Each thread takes an (x, y) column; its eight loads are issued together.
__global__ __launch_bounds__(256) void k_ext_init(ExtSet set) {
    const ExtGrid g = pick(set, blockIdx.x);
    const TileXYZ T = tile_of(g, blockIdx.x);
    const int i = T.bx * kExtTX + (threadIdx.x & (kExtTX - 1));
    const int j = T.by * kExtTY + (threadIdx.x / kExtTX);
    const bool col = i < g.ni && j < g.nj;
    uint8_t v[kExtTZ];
#pragma unroll
    for (int z = 0; z < kExtTZ; ++z) {
        const int k = T.bz * kExtTZ + z;
        v[z] = (col && k < g.nk) ? g.valid_in[at(i, j, k, g.ni, g.nj)] : 1;
    }
    bool inv = false, val = false;
#pragma unroll
    for (int z = 0; z < kExtTZ; ++z) {
        const int k = T.bz * kExtTZ + z;
        if (col && k < g.nk) {
            g.stamp[at(i, j, k, g.ni, g.nj)] = v[z] ? 0 : kNever;
            inv |= v[z] == 0;
            val |= v[z] != 0;
        }
    }
    inv = __syncthreads_or(inv);
    val = __syncthreads_or(val);
    if (threadIdx.x == 0) {
        g.tinv[T.t] = inv ? 1 : 0;
        g.tfv[T.t] = val ? 0 : kNever;
        if (blockIdx.x == 0) set.counter[0] = 0;
    }
}

// One thread per tile: tnear = 0 where a tile of the 3x3x3 block around it
// holds a valid cell at the start.
__global__ __launch_bounds__(256) void k_ext_near(ExtSet set) {
    const int gt = blockIdx.x * 256 + threadIdx.x;
    if (gt >= set.tiles) return;
    const ExtGrid g = pick(set, gt);
    const TileXYZ T = tile_of(g, gt);
    bool near = false;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const int x = T.bx + q % 3 - 1, y = T.by + (q / 3) % 3 - 1, z = T.bz + q / 9 - 1;
        if (x >= 0 && y >= 0 && z >= 0 && x < g.tx && y < g.ty && z < g.tz)
            near |= g.tfv[x + g.tx * (y + g.ty * z)] == 0;
    }
    g.tnear[T.t] = near ? 0 : kNever;
}

// One thread per tile: the tiles layer l works on (an invalid cell and a
// valid one in its 3x3x3 tile block) are appended to the list; the order of
// the list does not matter, the tiles of a layer being independent.
__global__ __launch_bounds__(256) void k_ext_select(ExtSet set, int l, int parity) {
    const int gt = blockIdx.x * 256 + threadIdx.x;
    bool act = false;
    if (gt < set.tiles) {
        const ExtGrid g = pick(set, gt);
        const int t = gt - g.tile0;
        act = g.tinv[t] != 0 && g.tnear[t] <= l;
    }
    const uint64_t m = __ballot(act);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&set.counter[parity], (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (act) set.list[base + __popcll(m & ((1ull << lane) - 1))] = gt;
}

// extrapolation.cpp:20-44, one layer over the listed tiles. A tile and its
// one-cell halo are staged in LDS: a 0/1 mask (valid at the layer's start)
// and the value of each valid cell, 0.0f for the others. Each thread walks
// one (x, y) column of the tile in z with the 3x3 neighbourhoods of three
// planes in registers, and sums all 27 masked values in the reference's
// order kk, jj, ii = -1..1. Adding +0.0f for an invalid neighbour leaves the
// float sum unchanged: the sum starts at +0.0f, and a float sum of +0.0f
// and further terms is -0.0f only when both addends are -0.0f, so the
// running sum is never -0.0f and s + 0.0f == s. The count is the integer
// sum of the masks; the cell takes sum / (float)count.
__global__ __launch_bounds__(256, 5) void k_ext_layer(ExtSet set, int l, int parity) {
    constexpr int kPer = (kHalo + 255) / 256;
    __shared__ uint8_t s_m[kPer * 256];
    __shared__ float s_v[kPer * 256];
    const uint32_t count = set.counter[parity];
    if (blockIdx.x == 0 && threadIdx.x == 0) set.counter[parity ^ 1] = 0;
    const int lx = threadIdx.x & (kExtTX - 1), ly = threadIdx.x / kExtTX;
    int gt_next = blockIdx.x < count ? set.list[blockIdx.x] : 0;
    for (uint32_t idx = blockIdx.x; idx < count; idx += gridDim.x) {
        const int gt = gt_next;
        if (idx + gridDim.x < count) gt_next = set.list[idx + gridDim.x];  // in flight during this tile
        const ExtGrid g = pick(set, gt);
        const TileXYZ T = tile_of(g, gt);
        const uint8_t tfv0 = g.tfv[T.t];
        const int i0 = T.bx * kExtTX - 1, j0 = T.by * kExtTY - 1, k0 = T.bz * kExtTZ - 1;
        // one round trip: stamps and values of the whole halo (values of
        // invalid cells are loaded too and dropped)
        uint8_t st[kPer];
        float fv[kPer];
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const int h = threadIdx.x + 256 * r;
            const int hx = h % kHX, hy = (h / kHX) % kHY, hz = h / (kHX * kHY);
            const int i = i0 + hx, j = j0 + hy, k = k0 + hz;
            const bool in = h < kHalo && i >= 0 && j >= 0 && k >= 0 && i < g.ni && j < g.nj && k < g.nk;
            const size_t c = in ? at(i, j, k, g.ni, g.nj) : 0;
            st[r] = in ? g.stamp[c] : kNever;
            fv[r] = in ? g.f[c] : 0.0f;
        }
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const bool valid = st[r] <= l;
            s_m[threadIdx.x + 256 * r] = valid ? 1 : 0;
            s_v[threadIdx.x + 256 * r] = valid ? fv[r] : 0.0f;
        }
        __syncthreads();

        const int i = T.bx * kExtTX + lx, j = T.by * kExtTY + ly;
        bool still = false, became = false;
        if (i < g.ni && j < g.nj) {
            // planes hz = z, z + 1, z + 2 of the halo hold cell planes z - 1, z, z + 1:
            // their 3x3 masked values, valid counts and centre masks
            float v[3][9];
            int cnt[3];
            bool mid[3];
            auto load_plane = [&](int p, int hz) {
                cnt[p] = 0;
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    const int h = (lx + q % 3) + kHX * ((ly + q / 3) + kHY * hz);
                    v[p][q] = s_v[h];
                    cnt[p] += s_m[h];
                }
                mid[p] = s_m[(lx + 1) + kHX * ((ly + 1) + kHY * hz)] != 0;
            };
            load_plane(0, 0);
            load_plane(1, 1);
#pragma unroll
            for (int z = 0; z < kExtTZ; ++z) {
                load_plane((z + 2) % 3, z + 2);
                const int k = T.bz * kExtTZ + z;
                if (k < g.nk && !mid[(z + 1) % 3]) {
                    const int n = cnt[0] + cnt[1] + cnt[2];
                    if (n > 0) {
                        float sum = 0.0f;
#pragma unroll
                        for (int dp = 0; dp < 3; ++dp)
#pragma unroll
                            for (int q = 0; q < 9; ++q) sum += v[(z + dp) % 3][q];
                        const size_t c = at(i, j, k, g.ni, g.nj);
                        g.f[c] = sum / (float)n;
                        g.stamp[c] = (uint8_t)(l + 1);
                        if (g.valid_out) g.valid_out[c] = 1;
                        became = true;
                    } else {
                        still = true;
                    }
                }
            }
        }
        still = __syncthreads_or(still);
        became = __syncthreads_or(became);
        if (threadIdx.x == 0) g.tinv[T.t] = still ? 1 : 0;
        // A tile's first valid cells make its 3x3x3 tile block "near" from the
        // next layer on. Writers of one layer all write l + 1, and the select
        // pass of this layer has already read tnear.
        if (became && tfv0 == kNever && threadIdx.x < 27) {
            const int q = threadIdx.x;
            const int x = T.bx + q % 3 - 1, y = T.by + (q / 3) % 3 - 1, z = T.bz + q / 9 - 1;
            if (x >= 0 && y >= 0 && z >= 0 && x < g.tx && y < g.ty && z < g.tz) {
                uint8_t* tn = g.tnear + (x + g.tx * (y + g.ty * z));
                if (*tn == kNever) *tn = (uint8_t)(l + 1);
            }
        }
        if (became && threadIdx.x == 0 && tfv0 == kNever) g.tfv[T.t] = (uint8_t)(l + 1);
        __syncthreads();  // LDS is free for the next tile
    }
}

// One workgroup per tile: stamps and tile stamps <= kExtRebase become 0.
__global__ __launch_bounds__(256) void k_ext_rebase(ExtSet set) {
    const ExtGrid g = pick(set, blockIdx.x);
    const TileXYZ T = tile_of(g, blockIdx.x);
    const int i = T.bx * kExtTX + (threadIdx.x & (kExtTX - 1));
    const int j = T.by * kExtTY + (threadIdx.x / kExtTX);
    if (i < g.ni && j < g.nj) {
        for (int z = 0; z < kExtTZ; ++z) {
            const int k = T.bz * kExtTZ + z;
            if (k >= g.nk) break;
            const size_t c = at(i, j, k, g.ni, g.nj);
            if (g.stamp[c] != kNever) g.stamp[c] = 0;
        }
    }
    if (threadIdx.x == 0) {
        if (g.tfv[T.t] != kNever) g.tfv[T.t] = 0;
        if (g.tnear[T.t] != kNever) g.tnear[T.t] = 0;
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

__device__ __forceinline__ float lerp2(float v0, float v1, float f) { return (1.0f - f) * v0 + f * v1; }

// interpolation.h:35-72 (dim 3): indices clamped to the grid (grid.h:370-379),
// x interpolated first, then y, then z.
__device__ __forceinline__ float trilerp_at(const Sampled& g, float x, float y, float z, float dx) {
    float fx, fy, fz;
    int i = index_and_fraction(x, g.o0, dx, fx);
    int j = index_and_fraction(y, g.o1, dx, fy);
    int k = index_and_fraction(z, g.o2, dx, fz);
    const int i0 = min(max(i, 0), g.ni - 1), i1 = min(max(i + 1, 0), g.ni - 1);
    const int j0 = min(max(j, 0), g.nj - 1), j1 = min(max(j + 1, 0), g.nj - 1);
    const int k0 = min(max(k, 0), g.nk - 1), k1 = min(max(k + 1, 0), g.nk - 1);
    const size_t r00 = (size_t)g.ni * ((size_t)j0 + (size_t)g.nj * k0);
    const size_t r10 = (size_t)g.ni * ((size_t)j1 + (size_t)g.nj * k0);
    const size_t r01 = (size_t)g.ni * ((size_t)j0 + (size_t)g.nj * k1);
    const size_t r11 = (size_t)g.ni * ((size_t)j1 + (size_t)g.nj * k1);
    const float v0 = g.a[r00 + i0], v1 = g.a[r00 + i1], v2 = g.a[r10 + i0], v3 = g.a[r10 + i1];
    const float v4 = g.a[r01 + i0], v5 = g.a[r01 + i1], v6 = g.a[r11 + i0], v7 = g.a[r11 + i1];
    const float b0 = lerp2(lerp2(v0, v1, fx), lerp2(v2, v3, fx), fy);
    const float b1 = lerp2(lerp2(v4, v5, fx), lerp2(v6, v7, fx), fy);
    return lerp2(b0, b1, fz);
}

struct VelField {
    Sampled u, v, w;
};

// ode.h:30-35 (trace_rk2) with the step -dt (advection.cpp:13,23); the field
// is lerp_staggered (interpolation.h:186-193).
__device__ __forceinline__ void backtrace(const VelField& F, float dx, float dt, float& x, float& y, float& z) {
    const float h = -dt;
    const float vx = trilerp_at(F.u, x, y, z, dx), vy = trilerp_at(F.v, x, y, z, dx),
                vz = trilerp_at(F.w, x, y, z, dx);
    const float half = 0.5f * h;
    const float mx = x + half * vx, my = y + half * vy, mz = z + half * vz;
    const float wx = trilerp_at(F.u, mx, my, mz, dx), wy = trilerp_at(F.v, mx, my, mz, dx),
                wz = trilerp_at(F.w, mx, my, mz, dx);
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
// neighbourhood of the velocity grids.
__global__ __launch_bounds__(256) void k_advect_faces(Advect A) {
    const int i = blockIdx.x * 32 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y,
              k = blockIdx.z * 2 + threadIdx.z;
    const float dx = A.dx;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const Sampled& g = A.val[d];
        if (i >= g.ni || j >= g.nj || k >= g.nk) continue;
        float x = ((float)i + 0.5f) * dx + g.o0;
        float y = ((float)j + 0.5f) * dx + g.o1;
        float z = ((float)k + 0.5f) * dx + g.o2;
        backtrace(A.field, dx, A.dt, x, y, z);
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
    backtrace(A.field, dx, A.dt, x, y, z);
    A.out[0][(size_t)i + (size_t)A.ni * ((size_t)j + (size_t)A.nj * k)] = trilerp_at(A.val[0], x, y, z, dx);
}

// ---------------------------------------------------------- reinitialisation
// The reference marches from the interface one face-neighbour layer per pass:
// FAR cells next to a FIX cell become NEAR, take an upwind distance from the
// FIX neighbours' values (a snapshot, grid_temp), and become FIX. As in the
// extrapolation, a 16-bit stamp per cell stands for the label and the
// snapshot: FIX at the start of pass l is stamp <= l, and a cell fixed in pass
// l gets l + 1, so each pass runs in place over the tiles it can change.

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

__device__ __forceinline__ float rei_at(const float* a, int i, int j, int k, int ni, int nj, int nk) {
    i = min(max(i, 0), ni - 1);
    j = min(max(j, 0), nj - 1);
    k = min(max(k, 0), nk - 1);
    return a[at(i, j, k, ni, nj)];
}

// reinitialization.cpp:27-62 for one valid cell: FIX if phi changes sign (or
// touches 0) towards a face neighbour (indices clamped); a FIX cell whose
// phi is not 0 takes calc_sdf over the 8 tetrahedra of its 7-point star
// (level_set.h:251-264). Star positions are Vector3f((i + 0.5) * dx, ...):
// double products rounded to float.
__device__ bool rei_interface(const ReiSet& r, int i, int j, int k, float& out) {
    const int ni = r.ni, nj = r.nj, nk = r.nk;
    const float* t = r.temp;
    const float p = t[at(i, j, k, ni, nj)];
    const float pxm = rei_at(t, i - 1, j, k, ni, nj, nk), pxp = rei_at(t, i + 1, j, k, ni, nj, nk);
    const float pym = rei_at(t, i, j - 1, k, ni, nj, nk), pyp = rei_at(t, i, j + 1, k, ni, nj, nk);
    const float pzm = rei_at(t, i, j, k - 1, ni, nj, nk), pzp = rei_at(t, i, j, k + 1, ni, nj, nk);
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

// One workgroup per tile: interface cells, stamps, tfar and tnew.
__global__ __launch_bounds__(256) void k_rei_init(ReiSet r) {
    const TileXYZ T = rei_tile(r, blockIdx.x);
    const int i = T.bx * kExtTX + (threadIdx.x & (kExtTX - 1));
    const int j = T.by * kExtTY + (threadIdx.x / kExtTX);
    bool far_any = false, fix_any = false;
    if (i < r.ni && j < r.nj) {
        for (int z = 0; z < kExtTZ; ++z) {
            const int k = T.bz * kExtTZ + z;
            if (k >= r.nk) break;
            const size_t c = at(i, j, k, r.ni, r.nj);
            uint16_t st;
            float v;
            if (r.valid[c] == 0) {
                st = kReiUnknown;
            } else if (rei_interface(r, i, j, k, v)) {
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
            for (int q = 0; q < 8; ++q) r.counter[q] = 0;
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
__device__ __forceinline__ bool rei_at_dist(const ReiSet& r, int i, int j, int k, uint16_t want) {
    if (i < 0 || j < 0 || k < 0 || i >= r.ni || j >= r.nj || k >= r.nk) return false;
    const size_t c = at(i, j, k, r.ni, r.nj);
    return r.stamp[c] != kReiUnknown && r.dist[c] == want;
}

// The cells not reached yet are bucketed by distance per chunk of
// kReiBucketCells cells (one workgroup): distances inside a chunk span less
// than kReiBins (neighbouring cells differ by at most 1, and a chunk covers at
// most a grid row plus a few cells of the next ones), so a chunk counts its
// cells per distance in LDS, with one global atomic per distinct distance.
constexpr int kReiBucketThreads = 1024;
constexpr int kReiBucketPer = 4;
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
    uint32_t bad = 0, top = 0;
    // the check: a FAR cell needs a finite distance >= 1 and a valid neighbour one step closer
#pragma unroll
    for (int q = 0; q < kReiBucketPer; ++q) {
        const size_t n = (size_t)r.ni * r.nj * r.nk;
        const size_t c = (size_t)blockIdx.x * kReiBucketCells + (size_t)q * kReiBucketThreads + threadIdx.x;
        if (c >= n || r.stamp[c] != kReiFar) continue;
        const uint16_t dv = r.dist[c];
        bool ok = dv != 0 && dv != kDistInf;
        if (ok) {
            const int i = (int)(c % r.ni), j = (int)((c / r.ni) % r.nj), k = (int)(c / ((size_t)r.ni * r.nj));
            const uint16_t w = dv - 1;
            ok = rei_at_dist(r, i - 1, j, k, w) || rei_at_dist(r, i + 1, j, k, w) ||
                 rei_at_dist(r, i, j - 1, k, w) || rei_at_dist(r, i, j + 1, k, w) ||
                 rei_at_dist(r, i, j, k - 1, w) || rei_at_dist(r, i, j, k + 1, w);
        }
        if (!ok) ++bad;
        else top = max(top, (uint32_t)dv);
    }
    bad = __reduce_add_sync(~0ull, bad);
    top = __reduce_max_sync(~0ull, top);
    if ((threadIdx.x & 63) == 0) {
        if (bad) atomicAdd(&r.counter[6], bad);
        if (top) atomicMax(&r.counter[7], top);
    }
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
// dfill gets the same offsets.
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
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = t * PER + q;
        if (idx < N) {
            r.dcount[idx] = run;
            r.dfill[idx] = run;
        }
        run += v[q];
    }
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
    for (int q = 0; q < kReiBucketPer; ++q)
        if (ch.d[q]) r.cells[bins[ch.d[q] - dmin] + ch.rank[q]] = (uint32_t)ch.c[q];
}

// One pass, one listed cell per thread: the cell takes its FIX neighbours
// (valid, distance target - 1) exactly as k_rei_pass does.
__global__ __launch_bounds__(256) void k_rei_layer(ReiSet r, uint32_t begin, uint32_t end, uint32_t target) {
    const uint32_t e = begin + blockIdx.x * 256 + threadIdx.x;
    if (e >= end) return;
    const size_t c = r.cells[e];
    const uint16_t w = (uint16_t)(target - 1);
    const int i = (int)(c % r.ni), j = (int)((c / r.ni) % r.nj), k = (int)(c / ((size_t)r.ni * r.nj));
    const bool fxm = rei_at_dist(r, i - 1, j, k, w), fxp = rei_at_dist(r, i + 1, j, k, w);
    const bool fym = rei_at_dist(r, i, j - 1, k, w), fyp = rei_at_dist(r, i, j + 1, k, w);
    const bool fzm = rei_at_dist(r, i, j, k - 1, w), fzp = rei_at_dist(r, i, j, k + 1, w);
    const size_t sx = 1, sy = (size_t)r.ni, sz = (size_t)r.ni * r.nj;
    const float vx = fxm ? r.phi[c - sx] : (fxp ? r.phi[c + sx] : 0.0f);
    const float vy = fym ? r.phi[c - sy] : (fyp ? r.phi[c + sy] : 0.0f);
    const float vz = fzm ? r.phi[c - sz] : (fzp ? r.phi[c + sz] : 0.0f);
    r.phi[c] = rei_update(fxm || fxp, vx, fym || fyp, vy, fzm || fzp, vz, r.dx);
}

__global__ __launch_bounds__(256) void k_rei_rebase(ReiSet r) {
    const size_t n = (size_t)r.ni * r.nj * r.nk;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride) {
        const uint16_t s = r.stamp[c];
        if (s != kReiFar && s != kReiUnknown) r.stamp[c] = 0;
    }
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
    hipLaunchKernelGGL(k_ext_init, dim3((unsigned)set.tiles), dim3(256), 0, s, set);
    hipLaunchKernelGGL(k_ext_near, dim3(cdiv(set.tiles, 256)), dim3(256), 0, s, set);
    return hipGetLastError();
}

hipError_t launch_ext_layer(const ExtSet& set, int l, int parity, hipStream_t s) {
    hipLaunchKernelGGL(k_ext_select, dim3(cdiv(set.tiles, 256)), dim3(256), 0, s, set, l, parity);
    const unsigned groups = (unsigned)std::min<int64_t>(set.tiles, kExtLayerGroups);
    hipLaunchKernelGGL(k_ext_layer, dim3(groups), dim3(256), 0, s, set, l, parity);
    return hipGetLastError();
}

hipError_t launch_ext_rebase(const ExtSet& set, hipStream_t s) {
    hipLaunchKernelGGL(k_ext_rebase, dim3((unsigned)set.tiles), dim3(256), 0, s, set);
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
    return hipGetLastError();
}

hipError_t launch_rei_lists(const ReiSet& r, hipStream_t s) {
    hipLaunchKernelGGL(k_rei_dscan, dim3(1), dim3(1024), 0, s, r);
    hipLaunchKernelGGL(k_rei_scatter, dim3(cdiv((int64_t)r.ni * r.nj * r.nk, kReiBucketCells)),
                       dim3(kReiBucketThreads), 0, s, r);
    return hipGetLastError();
}

hipError_t launch_rei_layer(const ReiSet& r, uint32_t begin, uint32_t end, int target, hipStream_t s) {
    if (end <= begin) return hipSuccess;
    hipLaunchKernelGGL(k_rei_layer, dim3(cdiv((int64_t)(end - begin), 256)), dim3(256), 0, s, r, begin, end,
                       (uint32_t)target);
    return hipGetLastError();
}

hipError_t launch_rei_rebase(const ReiSet& r, hipStream_t s) {
    hipLaunchKernelGGL(k_rei_rebase, dim3(2048), dim3(256), 0, s, r);
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
