// Host-side launch wrappers for the gfx950 kernels of the pressure projection.
// Every wrapper enqueues on the given stream and returns the launch status.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "elq_internal.hpp"

namespace elq {

// Dense inputs/outputs of one projection (device pointers, reference layout).
struct Dense {
    int ni, nj, nk;
    float dx;
    const float* phi;
    float* u;
    float* v;
    float* w;
    const float* uw;
    const float* vw;
    const float* ww;
};

// Brick metadata (device pointers).
struct Bricks {
    int nbx, nby, nbz;  // brick grid
    int nb;             // active bricks
    int nlev;           // brick levels (nbx + nby + nbz - 2)
    const int32_t* map;      // nbx*nby*nbz -> active id or -1
    const int32_t* coord;    // nb packed (bx | by << 10 | bz << 20)
    const int32_t* nbr;      // 8*nb: [b][-x,+x,-y,+y,-z,+z,pad,pad]
    const int32_t* lvl_list; // nb, sorted by bx+by+bz
};

// x-pencils: rows of bricks at fixed (by, bz) that hold at least one active
// brick. In the wavefront schedule one wave walks a pencil in x, so the x
// dependency of the triangular sweeps stays in registers.
struct Pencils {
    int np;                   // pencils with at least one active brick
    int shape;                // pencil-block shape (BlockShape) of the sweeps
    int nblk;                 // pencil blocks of that shape with an active pencil
    const int32_t* blk_list;  // nblk packed (Y | Z << 10), sorted by Y + Z, then Y
};

// Pencil-block shapes of the wavefront sweeps: BY x BZ pencils per workgroup,
// one wave each. Faces between the waves of a block go through LDS; faces
// between blocks through the exchange words in HBM. BY * BZ is a power of
// two (the grid tail's reduction tree over the waves of a workgroup).
enum BlockShape : int { kShape2x2 = 0, kShape1x4 = 1, kShapeCount = 2 };
constexpr int kShapeY[kShapeCount] = {2, 1};
constexpr int kShapeZ[kShapeCount] = {2, 4};

// Pencil blocks of one shape over an nby x nbz pencil grid.
inline size_t shape_blocks(int nby, int nbz, int shape) {
    return (size_t)((nby + kShapeY[shape] - 1) / kShapeY[shape]) * (size_t)((nbz + kShapeZ[shape] - 1) / kShapeZ[shape]);
}

// Block lists of every shape from the pencil map, shape k at offset
// sum_{s<k} shape_blocks(nby, nbz, s) of flag and list; count[k] = its nblk.
hipError_t launch_block_lists(const int32_t* pmap, int nby, int nbz, uint8_t* flag, int32_t* list, int64_t* count,
                              hipStream_t s);
// Workgroups of one round of a sweep at four workgroups per CU (the FP32
// residency of every shape up to 4 waves; FP64 sweeps hold two), capped by
// max_groups > 0.
int sweep_round_groups(int max_groups);

// Coefficients of the stored system (device pointers), T = storage type.
template <class T>
struct Coef {
    const T* diag = nullptr;
    const T* ax = nullptr;  // coupling of a cell to its +x neighbour (0 if none)
    const T* ay = nullptr;
    const T* az = nullptr;
    const T* inv_e = nullptr;  // IC factor pivots 1/e
    // FP64 forward sweep: lower couplings scaled by the cell's pivot inverse
    // (launch_scale_lower), lx = (coupling of the -x neighbour to the cell) / e,
    // 0 without that neighbour; likewise y, z. Unused (null) for FP32.
    const T* lx = nullptr;
    const T* ly = nullptr;
    const T* lz = nullptr;
    // per active brick: bit 1 if its coefficients are the uniform interior
    // ones, bit 2 (FP64) if its pivot inverses are also the interior value E
    // (launch_brick_uniform); the SpMV and the wavefront sweeps then read them
    // from urow: [0, 64) the coupling, [64, 128) the diagonal (FP32: its ghost
    // part), [128, 192) E, [192, 256) -E (the pre-scaled lower couplings)
    const int32_t* uflag = nullptr;
    const int32_t* umap = nullptr;  // the same flags by brick position (as Bricks::map); inactive: stale
    const T* urow = nullptr;
};

// ---------------------------------------------------------------- setup
hipError_t launch_mask(const Dense& d, uint8_t* mask, int32_t* err, hipStream_t s);
// Exclusive scan of (flags[i] & bit) != 0 in index order: out[i] = rank or fill.
hipError_t launch_scan(const uint8_t* flags, int64_t n, uint8_t bit, int32_t* out, int32_t fill,
                       int32_t* blk_cnt, int32_t* blk_off, int64_t* total, hipStream_t s);
int64_t scan_blocks(int64_t n);
hipError_t launch_brick_flags(const Dense& d, const uint8_t* mask, int nbx, int nby, int nbz,
                              uint8_t* bflag, hipStream_t s);
hipError_t launch_brick_meta(const int32_t* map, int nbx, int nby, int nbz, int32_t* coord,
                             int32_t* lvl_cnt, hipStream_t s);
hipError_t launch_level_scan(const int32_t* lvl_cnt, int nlev, int32_t* lvl_off, int32_t* lvl_fill,
                             hipStream_t s);
hipError_t launch_brick_lists(const int32_t* map, const int32_t* coord, int nbx, int nby, int nbz, int nb,
                              const int32_t* lvl_off, int32_t* lvl_fill, int32_t* lvl_list,
                              int32_t* nbr, hipStream_t s);
// Pencil flags over (by, bz) from the brick map (a scan turns them into pmap).
hipError_t launch_pencil_flags(const int32_t* bmap, int nbx, int nby, int nbz, uint8_t* pflag, hipStream_t s);
template <class T>
hipError_t launch_assemble(const Dense& d, float dt, const uint8_t* mask, const Bricks& bk, double* diag64,
                           double* b64, T* diag, T* ax, T* ay, T* az, hipStream_t s);
// IC(0)/MIC(0) factor in FP64 arithmetic; writes inv_e64 and, if inv_e32 is
// non-null, its float rounding. level < 0: wavefront launch over all bricks.
template <class T>
hipError_t launch_factor(const Bricks& bk, const double* diag64, const T* ax, const T* ay, const T* az,
                         double tau, double sigma, double* inv_e64, float* inv_e32, Ctrl* ctrl,
                         uint64_t* xch, int32_t* n_fixed, int level, int level_base, int level_count,
                         uint64_t timeout_ticks, int max_groups, hipStream_t s);
// Natural-order export of the assembled system (elq_build_matrix).
template <class T>
hipError_t launch_export(const Dense& d, const uint8_t* mask, const int32_t* indices, const Bricks& bk,
                         const double* diag64, const double* b64, const T* ax, const T* ay, const T* az,
                         double* odiag, double* oax, double* oay, double* oaz, double* ob, hipStream_t s);

// Per-brick-cell FP64 array (e.g. inv_e64) into the natural numbering.
hipError_t launch_export_cells(const Dense& d, const uint8_t* mask, const int32_t* indices, const Bricks& bk,
                               const double* v, double* out, hipStream_t s);

// ---------------------------------------------------------------- Krylov
template <class T>
hipError_t launch_init(const Bricks& bk, const double* b64, T* r, T* x, T* p0, T* p1, double* part,
                       Ctrl* ctrl, hipStream_t s);
// stage: 4*kRedBlocks tagged words (zero-initialised once, never reset).
// Copies bytes (a multiple of 4) of device memory into pinned host memory
// with a small kernel (system-scope stores), for polls that would otherwise
// wait on a copy-engine transfer.
hipError_t launch_export_words(const void* src, size_t bytes, void* host, hipStream_t s);
hipError_t launch_scalar(Ctrl* ctrl, int mode, const double* parts, int n, const double* parts2, uint64_t* stage,
                         uint64_t timeout_ticks, hipStream_t s);
template <class T>
// Persistent launches (factor wave, SpMV, pencil sweeps) take max_groups:
// 0 = as many workgroups as the device holds, > 0 = at most that many.
// Persistent launch; its grid tail sets alpha (SC_ALPHA).
hipError_t launch_spmv(const Bricks& bk, const Coef<T>& c, const T* z, const T* p_in, T* p_out, T* q, Ctrl* ctrl,
                       uint64_t* stage, uint64_t timeout_ticks, int max_groups, hipStream_t s);
// Forward sweep y = (E+L)^{-1} r, fused with x += alpha p, r -= alpha q and
// |r|^2 partials when update != 0. level < 0: one persistent launch over the
// pencil blocks (wavefront), ending in a grid tail (SC_RR with update, else
// SC_NOP); else the bricks of one brick level, with per-brick partials in part.
template <class T>
hipError_t launch_fwd(const Bricks& bk, const Pencils& pc, const Coef<T>& c, T* r, T* x, const T* p, const T* q, T* y,
                      double* part, Ctrl* ctrl, uint64_t* xch, uint64_t* stage, int update, int level,
                      int level_base, int level_count, uint64_t timeout_ticks, int max_groups, hipStream_t s);
// One request of a batched Krylov chunk (serving, mixed precision, wavefront
// sweeps of the 1x4 or 2x2 shape): its solver state and the workgroups each
// kernel class of its chunk uses (those its direct launches would use, so the
// arithmetic, reductions included, is the same). Workgroup g of a batched
// launch belongs to request g / per_job and is its workgroup g % per_job;
// requests with ctrl == nullptr, and workgroups past a request's count, exit
// at once.
struct KrylovJob {
    Bricks bk;
    Pencils pc;
    Coef<float> c;
    float* r;
    float* x;
    float* p0;
    float* p1;
    float* q;
    float* y;
    Ctrl* ctrl;
    uint64_t* xch;
    uint64_t* stage;
    uint64_t timeout;
    int32_t g_spmv, g_fwd, g_bwd;
    int32_t par;  // iterations done before the chunk, mod 2 (selects the p buffer)
};
// Workgroups of the direct launches of one Krylov iteration (SpMV, forward
// sweep with updates, backward sweep) under the cap max_groups.
void krylov_groups(const Bricks& bk, const Pencils& pc, int max_groups, int* g_spmv, int* g_fwd, int* g_bwd);
// n iterations (SpMV, forward sweep with updates, backward sweep; Eigen's
// loop body) of every request of jobs[0 .. njobs), each launch with
// njobs * per_job workgroups; shapes: bit mask (1 << BlockShape) of the sweep
// shapes the requests use (the sweeps launch once per shape).
hipError_t launch_krylov_batch(const KrylovJob* jobs, int njobs, int per_job, int n, int shapes, hipStream_t s);

// Uniform-brick flags and constant rows (Coef::uflag, urow): a brick is
// uniform when every cell's diagonal entry equals dref and its six couplings
// equal cref, bit for bit.
template <class T>
hipError_t launch_brick_uniform(const Bricks& bk, const T* diag, const T* ax, const T* ay, const T* az, T cref,
                                T dref, const double* ie64, double tau, double sigma, int32_t* uflag, int32_t* umap,
                                T* urow, hipStream_t s);
// Pivot-scaled lower couplings of the FP64 forward sweep (Coef::lx, ly, lz).
hipError_t launch_scale_lower(const Bricks& bk, const double* ax, const double* ay, const double* az,
                              const double* ie, double* lx, double* ly, double* lz, hipStream_t s);

// Forward sweep with the x, r updates and the backward sweep in one persistent
// launch (FP32, 1x4 pencil blocks): the same operations and reductions as
// launch_fwd(update) followed by launch_bwd(SC_BETA), bit for bit. The stage
// buffer holds 4 * kTailMaxGroups words (the backward tail uses the second half).
bool sweeps_fusable(const Pencils& pc);
hipError_t launch_sweeps(const Bricks& bk, const Pencils& pc, const Coef<float>& c, float* r, float* x, const float* p,
                         const float* q, float* y, Ctrl* ctrl, uint64_t* xch, uint64_t* stage, uint64_t timeout_ticks,
                         int max_groups, hipStream_t s);

// Backward sweep z = y - E^{-1} L^T z (in place over y), fused with r.z partials.
// level < 0: persistent launch over the pencil blocks ending in a grid tail
// that applies tail_mode; else one brick level with per-brick partials in part.
template <class T>
hipError_t launch_bwd(const Bricks& bk, const Pencils& pc, const Coef<T>& c, const T* r, T* y, double* part, Ctrl* ctrl,
                      uint64_t* xch, uint64_t* stage, int tail_mode, int level, int level_base, int level_count,
                      uint64_t timeout_ticks, int max_groups, hipStream_t s);
// r64 = b64 - A x64 in FP64 (couplings TC read exactly); partials |r64|^2 and,
// when r32 != nullptr, |float(r64)|^2 with r32 = float(r64); x32 and p0/p1 are
// zeroed when non-null.
template <class TC>
hipError_t launch_resid64(const Bricks& bk, const double* diag64, const TC* ax, const TC* ay, const TC* az,
                          const double* b64, const double* x64, double* part, double* part2, float* r32,
                          float* x32, float* p0, float* p1, Ctrl* ctrl, hipStream_t s);
hipError_t launch_xupdate(const Bricks& bk, double* x64, const float* x32, Ctrl* ctrl, hipStream_t s);

// ---------------------------------------------------------------- finish
hipError_t launch_scatter(const Dense& d, const uint8_t* mask, const Bricks& bk, const double* x,
                          float* pressure, hipStream_t s);
hipError_t launch_update(const Dense& d, float dt, const float* pressure, hipStream_t s);

// ---------------------------------------------------------------- inputs
// compute_staggered_grid_weight (boundary.cpp:6-34) from the nodal solid level set.
hipError_t launch_weights(int ni, int nj, int nk, const float* nodal, float* uw, float* vw, float* ww,
                          hipStream_t s);
// add_gravity (simulator.cpp:90-93) on the n faces of v.
hipError_t launch_gravity(float* v, size_t n, float dt, hipStream_t s);
// find_valid_velocity + set_invalid_velocity_zero (simulator.cpp:125-178).
hipError_t launch_valid_zero(const Dense& d, hipStream_t s);

// ---------------------------------------------------------------- utility
// Device-to-device copy of bytes (a multiple of 4) with a plain streaming kernel.
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);

}  // namespace elq
