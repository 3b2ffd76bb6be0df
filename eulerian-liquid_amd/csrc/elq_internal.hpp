// Internal types shared by the gfx950 kernels (elq_kernels.hip) and the host
// orchestration (elq_api.cpp). Not part of the C ABI.
//
// Data layout in HBM
// ------------------
// Dense grids (level set, face velocities/weights, pressure, indices) keep the
// reference layout, i fastest (include/nama/common/array.h:285).
// The linear system lives on a brick-sparse layout: the cell grid is tiled by
// 4x4x4 bricks; only bricks holding at least one unknown are stored. Brick b
// owns the 64 consecutive entries [64*b, 64*b+64) of every per-cell array,
// lane l = lx + 4*ly + 16*lz, so one wave64 owns one brick and every brick
// load/store is a single coalesced 256 B (FP32) or 512 B (FP64) transaction.
// Bricks are numbered in lexicographic (bz, by, bx) order; lvl_list holds the
// same ids sorted by brick level bx+by+bz, a topological order of the
// natural-ordering triangular solves.
#pragma once

#include <cstdint>

namespace elq {

constexpr int kBrick = 4;
constexpr int kBrickCells = 64;
constexpr int kScanBlock = 256;               // threads per scan workgroup
constexpr int kScanItems = 16;                // cells per thread per scan block
constexpr int kScanChunk = kScanBlock * kScanItems;
constexpr int kReduceThreads = 256;           // scalar kernel workgroup width
constexpr int kRedBlocks = 128;               // max workgroups of the scalar kernel
constexpr int kTailMaxGroups = 4096;          // max workgroups of a kernel ending in a grid tail
constexpr int kFuseMaxBricks = 32768;         // mixed precision fuses the two sweeps up to this many bricks
constexpr int kRedPerBlock = 2048;            // partials per scalar workgroup (target)
constexpr int kXchWords = 192;                // face-exchange words per brick (6 x 16 x 2)

// Dense-cell mask bits
constexpr uint8_t kValid = 1;   // projection.cpp:8-12
constexpr uint8_t kFluid = 2;   // valid && phi < 0: an unknown

// Setup error bits
constexpr int kErrBoundary = 1;

// Scalar-kernel modes
enum ScalarMode : int {
    SC_NOP = 0,       // epoch/ticket bump only
    SC_INIT = 1,      // |b|^2, threshold, initial convergence test (FP64 PCG)
    SC_RZ_INIT = 2,   // rz = r.z, beta = 0
    SC_ALPHA = 3,     // alpha = rz / p.q
    SC_RR = 4,        // |r|^2 and convergence test
    SC_BETA = 5,      // beta = rz_new / rz, iteration++
    SC_REPLACE = 6,   // mixed: FP64 residual replaces the FP32 recursive one; true convergence test
    SC_TRUE = 7,      // true residual norm of the final iterate
    SC_RZ_REPLACE = 8 // mixed: rz and beta after a residual replacement
};

// Device-resident solver state. One per solve (slot).
struct alignas(64) Ctrl {
    double rz;          // r.z (current)
    double pq;          // p.Ap
    double rr;          // |r|^2 of the Krylov loop (recursive)
    double alpha;
    double beta;
    double bnorm2;      // |b|^2
    double threshold;   // Krylov stop threshold on |r|^2
    double tol2;        // tol^2
    double inner_tol2;  // mixed: inner relative tolerance^2
    double rr64;        // mixed: FP64 residual norm^2 of the current outer step
    double true_rr;     // |b - A x|^2 of the final iterate
    double last_err;    // sqrt(rr / bnorm2) at stop
    double rz_div;      // rz the next beta divides by (mixed residual replacement)
    double rr_ref;      // |r|^2 right after the last residual replacement
    double rr_rec;      // recursive |r|^2 just before the last residual replacement
    int32_t iter;       // Krylov iterations of the current (inner) solve
    int32_t iter_total; // summed over refinement steps
    int32_t max_iter;
    int32_t done;       // Krylov loop finished (converged, max_iter or breakdown)
    int32_t converged;
    int32_t outer_done;
    int32_t outer;      // residual replacements (mixed)
    int32_t status;     // elq_status
    int32_t mixed;
    uint32_t epoch;     // wavefront flag generation
    uint32_t ticket;    // wavefront work counter
    uint32_t abort;     // set by a wait that exceeded its time limit
    uint32_t red_epoch; // tag generation of the scalar kernel's partial sums
    int32_t pad[3];
};

struct SetupInfo {
    int64_t n_unknowns;
    int64_t n_bricks;
    int32_t err;
    int32_t n_levels;
};

}  // namespace elq
