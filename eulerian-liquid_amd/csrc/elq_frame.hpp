// Host-side launch wrappers for the gfx950 kernels of the frame steps around
// the projection: validity flags of the faces and of the level set, zeroing
// of invalid faces, and the layered extrapolation of extrapolation.cpp.
// Every wrapper enqueues on the given stream and returns the launch status.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "elq_kernels.hpp"

namespace elq {

// Extrapolation tiles: 32 x 8 x 8 cells per workgroup of 256 threads.
constexpr int kExtTX = 32;
constexpr int kExtTY = 8;
constexpr int kExtTZ = 8;
// Layer stamps are bytes; every kExtRebase layers they are folded back to 0.
constexpr int kExtRebase = 250;

// Workgroups of the persistent layer kernel (tiles are taken from the
// layer's active list in strides of this many).
constexpr int kExtLayerGroups = 1280;

// One grid of an extrapolation (device pointers).
struct ExtGrid {
    float* f;                 // values, updated in place
    const uint8_t* valid_in;  // caller's 0/1 flags (read by the init pass)
    uint8_t* valid_out;       // set to 1 where a cell receives a value (nullptr: left alone)
    uint8_t* stamp;           // per cell: 0 valid from the start, s: valid from layer s on, 255 invalid
    uint8_t* tinv;            // per tile: 1 while the tile holds an invalid cell
    uint8_t* tfv;             // per tile: first layer at whose start the tile holds a valid cell (255: none)
    uint8_t* tnear;           // per tile: first layer at whose start a tile of its 3x3x3 block holds one
    int ni, nj, nk;
    int tx, ty, tz;           // tiles per axis
    int tile0;                // first global tile index of this grid
};

// Up to three grids (u, v, w) extrapolated by the same launches.
struct ExtSet {
    ExtGrid g[3];
    int count;
    int tiles;         // tiles of all grids
    int32_t* list;     // active tiles of the current layer (global tile indices)
    uint32_t* counter; // two list lengths, used by alternate layers
};

inline int64_t ext_tiles(int ni, int nj, int nk) {
    return (int64_t)((ni + kExtTX - 1) / kExtTX) * ((nj + kExtTY - 1) / kExtTY) * ((nk + kExtTZ - 1) / kExtTZ);
}

// Stamps, tile flags and the first layer's neighbourhood flags from valid_in.
hipError_t launch_ext_init(const ExtSet& set, hipStream_t s);
// Layer l (0-based since the last rebase, l < kExtRebase; parity = absolute
// layer index & 1): the active-tile list, then the layer over it.
hipError_t launch_ext_layer(const ExtSet& set, int l, int parity, hipStream_t s);
// Fold every stamp <= kExtRebase back to 0 (valid from the start).
hipError_t launch_ext_rebase(const ExtSet& set, hipStream_t s);

// find_valid_velocity (simulator.cpp:125-167): 0/1 flags of the u, v, w faces.
hipError_t launch_find_valid(const Dense& d, uint8_t* uv, uint8_t* vv, uint8_t* wv, hipStream_t s);
// set_invalid_velocity_zero (simulator.cpp:169-178) on one face grid of n faces.
hipError_t launch_zero_invalid(float* f, const uint8_t* valid, size_t n, hipStream_t s);
// calc_cell_centered_solid_phi + find_valid_level_set (simulator.cpp:160-167,
// 191-200): a cell is valid unless the mean of its 8 solid nodes is < 0.
hipError_t launch_level_set_valid(int ni, int nj, int nk, const float* solid_nodal, uint8_t* valid,
                                  hipStream_t s);

// advect_staggered_grid (advection.cpp:19-27): out_{u,v,w} = val_{u,v,w}
// sampled at the RK2 backtrace (step -dt) of each face through the face
// velocities vel_{u,v,w}. Outputs must not alias any input.
hipError_t launch_advect_faces(int ni, int nj, int nk, float dx, float dt, const float* vel_u, const float* vel_v,
                               const float* vel_w, const float* val_u, const float* val_v, const float* val_w,
                               float* out_u, float* out_v, float* out_w, hipStream_t s);
// advect_collocated_grid (advection.cpp:9-17): out = val (cell grid) sampled
// at the RK2 backtrace of each cell centre. out must not alias any input.
hipError_t launch_advect_cells(int ni, int nj, int nk, float dx, float dt, const float* vel_u, const float* vel_v,
                               const float* vel_w, const float* val, float* out, hipStream_t s);

// Reinitialisation (reinitialization.cpp:19-156): same 32x8x8 tiles; cell
// stamps are 16-bit (kReiFar: not reached yet, kReiUnknown: invalid cell,
// s: fixed from the start of pass s on). Every kReiRebase passes the stamps
// are folded back to 0.
constexpr uint16_t kReiFar = 0xFFFF;
constexpr uint16_t kReiUnknown = 0xFFFE;
constexpr int kReiRebase = 60000;

struct ReiSet {
    float* phi;            // level set, updated in place
    const float* temp;     // copy of phi taken before the interface pass
    const uint8_t* valid;  // 0/1 flags
    uint16_t* stamp;
    uint8_t* tfar;         // per tile: 1 while the tile holds a cell not reached yet
    uint16_t* tnew;        // per tile: latest pass at whose start it holds cells fixed just
                           // before (stamp == that pass; kReiFar: none yet)
    int32_t* list;         // tiles of the current pass
    uint16_t* dist;        // per cell: taxicab distance to the nearest interface cell
    uint32_t* dcount;      // per distance 0..kReiDistMax: cells not reached yet, then list offsets
    uint32_t* dfill;       // per distance: next free list slot
    uint32_t* cells;       // cells not reached yet, grouped by distance
    uint32_t* counter;     // [0..1] list lengths, [2..3] cells fixed by alternate passes,
                           // [4] cells left unreached, [5] largest stamp,
                           // [6] cells where dist is not the march's pass, [7] largest dist
    int ni, nj, nk;
    int tx, ty, tz;
    int tiles;
    float dx;
};

// Interface cells (their geometric distance) and stamps; tile flags.
hipError_t launch_rei_init(const ReiSet& r, hipStream_t s);
// dist: taxicab distance transform of the interface cells (three separable
// passes), then a check that it equals the march's pass count at every
// valid cell (counter[6] = failures, counter[7] = largest distance) and a
// count of the cells at each distance (dcount).
constexpr int kReiDistMax = 3 * 4092;
hipError_t launch_rei_dist(const ReiSet& r, hipStream_t s);
// After the check passed: dcount becomes list offsets (exclusive scan over
// 0..kReiDistMax + 1) and the cells not reached yet are listed by distance.
hipError_t launch_rei_lists(const ReiSet& r, hipStream_t s);
// The cells of one pass (list entries [begin, end), all at distance target)
// from their FIX neighbours (valid, distance target - 1).
hipError_t launch_rei_layer(const ReiSet& r, uint32_t begin, uint32_t end, int target, hipStream_t s);
// Pass l (0-based since the last rebase; parity = absolute pass & 1).
hipError_t launch_rei_pass(const ReiSet& r, int l, int parity, hipStream_t s);
hipError_t launch_rei_rebase(const ReiSet& r, hipStream_t s);
// counter[4] = cells never reached, counter[5] = largest stamp (passes that fixed cells).
hipError_t launch_rei_finish(const ReiSet& r, hipStream_t s);

}  // namespace elq
