// Host-side launch wrappers for the gfx950 kernels of the frame steps around
// the projection: validity flags of the faces and of the level set, zeroing
// of invalid faces, the layered extrapolation of extrapolation.cpp, the
// semi-Lagrangian advection and the level-set reinitialisation.
// Every wrapper enqueues on the given stream and returns the launch status.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "elq_kernels.hpp"

namespace elq {

// Tiles of the reinitialisation's tile passes: 32 x 8 x 8 cells per
// workgroup of 256 threads.
constexpr int kExtTX = 32;
constexpr int kExtTY = 8;
constexpr int kExtTZ = 8;

// Workgroups of the persistent tile-pass kernel (tiles are taken from the
// pass's active list in strides of this many).
constexpr int kExtLayerGroups = 1280;

// Extrapolation (extrapolation.cpp:5-51) over bit planes: bit x % 64 of
// word (row, x / 64), row = j + nj k, is set where cell (x, j, k) is valid at
// the start of a layer. A layer dilates the plane by the 3x3x3 box and lists
// the cells it adds; a second launch gives each listed cell the mean of its
// neighbours that were valid at the layer's start.
struct ExtGrid {
    float* f;                 // values, updated in place
    const uint8_t* valid_in;  // caller's 0/1 flags (read by the init pass)
    uint8_t* valid_out;       // set to 1 where a cell receives a value (nullptr: left alone)
    uint64_t* plane[2];       // validity at the start of even / odd layers
    uint32_t* list;           // cells added by the current layer
    int ni, nj, nk;
    int wpr;                  // words per row
    int slot;                 // index of this grid in its set
    int64_t word0;            // first word in the set's word space (a multiple of kExtWordAlign)
    int64_t words;            // rows * wpr
};

// Up to three grids (u, v, w) extrapolated by the same launches.
struct ExtSet {
    ExtGrid g[3];
    int count;
    int64_t words;      // word space of all grids, each padded to a multiple of kExtWordAlign
    uint32_t* counter;  // [2 q + (layer & 1)]: length of grid q's list
};

// Grids start at multiples of this many words in a set's word space: a
// workgroup of the dilation (one thread per word) stays inside one grid.
constexpr int kExtWordAlign = 1024;

// Words per row and words of a grid's plane.
inline int ext_wpr(int ni) { return (ni + 63) / 64; }
inline int64_t ext_words(int ni, int nj, int nk) { return (int64_t)ext_wpr(ni) * nj * nk; }

// Plane of layer 0 from valid_in; list lengths cleared.
hipError_t launch_ext_init(const ExtSet& set, hipStream_t s);
// Layer l: dilation (plane l & 1 -> plane (l & 1) ^ 1, list of added cells),
// then the added cells' values.
hipError_t launch_ext_layer(const ExtSet& set, int l, hipStream_t s);

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

// Reinitialisation (reinitialization.cpp:19-156). Cell stamps are 16-bit
// (kReiFar: not reached yet, kReiUnknown: invalid cell, s: fixed from the
// start of pass s on). The tile passes (32x8x8 tiles) fold the stamps back to
// 0 every kReiRebase passes (a context can ask for fewer); the distance route
// has no stamps to fold.
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
    uint8_t* fixm;         // per listed cell: its face neighbours one step closer (bits -x +x -y +y -z +z)
    uint32_t* counter;     // [0..1] list lengths, [2..3] cells fixed by alternate passes,
                           // [4] cells left unreached, [5] largest stamp,
                           // [6] cells where dist is not the march's pass, [7] largest dist,
                           // [8] passes that fixed cells before the last rebase
    int ni, nj, nk;
    int tx, ty, tz;
    int tiles;
    float dx;
};

// Interface cells (their geometric distance) and stamps; tile flags.
hipError_t launch_rei_init(const ReiSet& r, hipStream_t s);
// dist: taxicab distance transform of the interface cells (three separable
// passes), then a check that it equals the march's pass count at every
// valid cell (counter[6] = failures), the cells counted per distance and
// the counts turned into list offsets in dcount and dfill (exclusive scan
// over 0..kReiDistMax + 1; counter[7] = largest distance).
constexpr int kReiDistMax = 3 * 4092;
hipError_t launch_rei_dist(const ReiSet& r, hipStream_t s);
// After the check passed: the cells not reached yet listed by distance, each
// with the mask of its FIX neighbours (valid, one step closer) in fixm.
hipError_t launch_rei_lists(const ReiSet& r, hipStream_t s);
// The cells of one pass (list entries [begin, end), all at one distance)
// from their FIX neighbours.
hipError_t launch_rei_layer(const ReiSet& r, uint32_t begin, uint32_t end, hipStream_t s);
// Tile route. Pass l (0-based since the last rebase; parity = absolute pass & 1).
hipError_t launch_rei_pass(const ReiSet& r, int l, int parity, hipStream_t s);
// Stamps of fixed cells back to 0; counter[8] = max(counter[8], base + the
// largest stamp folded), base being the first pass of the epoch folded.
hipError_t launch_rei_rebase(const ReiSet& r, uint32_t base, hipStream_t s);
// counter[4] = cells never reached, counter[5] = largest stamp (passes that fixed cells).
hipError_t launch_rei_finish(const ReiSet& r, hipStream_t s);

}  // namespace elq
