/*
 * elq.h — C ABI of the MI355X (gfx950) pressure projection.
 *
 * Drop-in for the reference simulator's pressure projection
 * (daichi-nama/eulerian-liquid):
 *   Simulator::solve_pressure(float dt)          simulator.cpp:96-112 / simulator.h:39
 *   fluid::assign_indices(...)                   projection.cpp:14-28 / projection.h:10-11
 *   fluid::build_matrix(...)                     projection.cpp:30-83 / projection.h:13-16
 *   fluid::build_grid_from_vector(...)           projection.cpp:85-91 / projection.h:18-19
 *   fluid::update_velocity_using_pressure(...)   projection.cpp:93-134 / projection.h:21-22
 * and the other steps of the reference's frame on device memory (weights,
 * gravity, validity, extrapolation, advection, reinitialisation; each entry
 * point below cites the function it replaces).
 *
 * Plain C types only: pointers, sizes, POD structs. No exceptions cross this
 * boundary; every entry point returns an elq_status. Grids use the
 * reference's layout (include/nama/common/array.h:285): flat, index
 * i + ni*(j + nj*k); face grids u (ni+1)*nj*nk, v ni*(nj+1)*nk,
 * w ni*nj*(nk+1) (include/nama/common/staggered_grid.h:100-112).
 *
 * All computation of the projection runs on the GPU. A process without a
 * usable gfx950 device gets ELQ_ERR_DEVICE from elq_ctx_create.
 */
#ifndef ELQ_H
#define ELQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ELQ_ABI_VERSION 1
#define ELQ_UNKNOWN_INDEX 2147483647 /* projection.cpp:6 (INT_MAX) */

typedef enum {
    ELQ_OK = 0,
    ELQ_ERR_NO_FLUID = 1,       /* projection.cpp:24-26 throws "No fluid" */
    ELQ_ERR_NOT_CONVERGED = 2,  /* max_iter reached; the reference ignores this
                                   (simulator.cpp:105-109), the C++ shim does too */
    ELQ_ERR_BREAKDOWN = 3,      /* non-finite scalar inside the Krylov loop */
    ELQ_ERR_DEVICE = 4,         /* HIP error, no gfx950 device, or host memory
                                   or threads exhausted */
    ELQ_ERR_ADMISSION = 5,      /* request exceeds the context's HBM budget */
    ELQ_ERR_INVALID_ARG = 6,
    ELQ_ERR_BOUNDARY = 7,       /* an unknown lies on the outer cell layer: the
                                   reference reads outside its arrays there
                                   (projection.cpp:54 via :8-12), undefined */
    ELQ_ERR_BUSY = 8            /* serving queue full */
} elq_status;

typedef enum {
    ELQ_PRECISION_FP64 = 0,  /* plain FP64 drop-in: FP64 storage and arithmetic */
    ELQ_PRECISION_MIXED = 1  /* FP32 SpMV/preconditioner/vectors, FP64 dots and
                                scalars, FP64 outer residual + refinement */
} elq_precision;

typedef enum {
    ELQ_PRECOND_IC0 = 0, /* natural-order IC(0): the default */
    ELQ_PRECOND_MIC0 = 1 /* natural-order modified IC(0), parameter tau */
} elq_precond;

typedef enum {
    ELQ_SCHED_WAVEFRONT = 0,       /* one launch per sweep; bricks taken in level order,
                                      predecessor faces arrive as epoch-tagged words. Mixed
                                      precision on grids of up to 32768 active bricks (128^3)
                                      whose pencil blocks fit one round of the persistent
                                      grid runs the forward and backward sweeps in one launch */
    ELQ_SCHED_LEVELS = 1,          /* one launch per brick level (no device waits) */
    ELQ_SCHED_WAVEFRONT_SPLIT = 2, /* wavefront, always one launch per sweep */
    ELQ_SCHED_WAVEFRONT_FUSED = 3  /* wavefront, mixed precision: both sweeps in one launch
                                      whenever the pencil blocks fit one round of the
                                      persistent grid, at any brick count (same results as
                                      split); otherwise as ELQ_SCHED_WAVEFRONT_SPLIT */
} elq_schedule;

/* Inputs/outputs of one projection. Host or device pointers depending on the
 * entry point. pressure and indices are optional outputs (NULL to skip). */
typedef struct {
    int32_t ni, nj, nk;
    float dx;                 /* Grid3::dx(), width / ni (simulator.cpp:56) */
    const float* level_set;   /* ni*nj*nk, liquid phi (<0 inside) */
    float* u;                 /* (ni+1)*nj*nk, in/out */
    float* v;                 /* ni*(nj+1)*nk, in/out */
    float* w;                 /* ni*nj*(nk+1), in/out */
    const float* u_weight;    /* face weights, same shapes as u/v/w */
    const float* v_weight;
    const float* w_weight;
    float* pressure;          /* ni*nj*nk, out (background 0), optional */
    int32_t* indices;         /* ni*nj*nk, out (ELQ_UNKNOWN_INDEX), optional */
} elq_grid;

typedef struct {
    double tol;        /* relative residual; reference 1e-4 (simulator.cpp:106) */
    int32_t max_iter;  /* reference 1000 (simulator.cpp:105); for MIXED it bounds
                          the total number of inner iterations */
    int32_t precision; /* elq_precision */
    int32_t precond;   /* elq_precond */
    double tau;        /* MIC(0) parameter (ignored for IC0) */
    double sigma;      /* pivot safeguard: pivot t <= sigma*diag -> diag */
    double inner_tol;  /* MIXED: the FP32 recursive residual is replaced by the FP64
                          residual b - A x each time it has dropped by this factor
                          (default 0.01; checked at the polls) and whenever it meets
                          the stopping test; a replacement far above the recursive
                          residual restarts the search direction */
    int32_t schedule;  /* elq_schedule */
    int32_t check_every; /* PCG iterations between host convergence polls */
} elq_opts;

typedef struct {
    int64_t n_unknowns;        /* N (mat_size of projection.cpp:14-28) */
    int64_t n_bricks;          /* active 4x4x4 bricks */
    int32_t iterations;        /* Eigen iterations(): PCG iterations (FP64) or
                                  total inner iterations (MIXED) */
    int32_t outer_iterations;  /* MIXED: FP64 residual evaluations (residual
                                  replacements incl. the first); 1 for FP64 */
    double error;              /* Eigen error(): sqrt(|r|^2/|b|^2), recursive r */
    double true_rel_residual;  /* |b - A p|_2 / |b|_2 recomputed in FP64 */
    int32_t status;            /* elq_status of the solve */
    int32_t converged;
    double t_setup_ms;         /* indices, bricks, assembly, factorisation */
    double t_solve_ms;         /* Krylov iterations */
    double t_update_ms;        /* pressure scatter + velocity update */
    double t_total_ms;
} elq_stats;

typedef struct elq_ctx elq_ctx;

/* Fills *opts with the reference defaults (tol 1e-4, max_iter 1000, FP64,
 * IC0, sigma 0.25, wavefront schedule). */
void elq_default_opts(elq_opts* opts);

/* Context bound to one HIP device; owns device buffers that persist across
 * calls and grow on demand. hbm_budget_bytes = 0: no limit. */
int elq_ctx_create(int device, size_t hbm_budget_bytes, elq_ctx** out);
void elq_ctx_destroy(elq_ctx* ctx);
/* Last error message of this context (static storage of the context). */
const char* elq_ctx_last_error(const elq_ctx* ctx);

/* Enables (1) or disables (0, the default) graph replay of the Krylov loop:
 * the launches of one poll interval (check_every iterations: SpMV, forward
 * and backward sweep each) are captured once into a HIP graph and replayed
 * while their arguments stay the same, which takes the per-launch host cost
 * off the loop. Results are identical either way. Not used with the levels
 * schedule or while per-kernel event timing is on. No reference counterpart. */
int elq_set_graph_capture(elq_ctx* ctx, int enable);

/* Bounds every device-side wait of the persistent kernels (a workgroup waiting
 * for another's face values or partial sums) to `seconds` of device wall
 * clock; default 4. A wait that runs out raises the context's abort flag,
 * every other wait then ends at once, and the call returns ELQ_ERR_DEVICE
 * ("wavefront wait exceeded its time limit"); the next call on the context
 * starts afresh. No reference counterpart (the reference has no device
 * waits). 0 ends every wait that is not satisfied at its first check. */
int elq_set_wait_limit(elq_ctx* ctx, double seconds);

/* Folds the 16-bit pass stamps of the reinitialisation's tile route back to
 * 0 every `passes` passes (1..60000; default 60000, the most the stamps
 * hold). No reference counterpart: results do not depend on it; a smaller
 * interval exercises the rebase on short marches (tests). */
int elq_set_reinit_rebase(elq_ctx* ctx, int32_t passes);

/* Simulator::solve_pressure(dt) on host grids (simulator.cpp:96-112): copies
 * inputs to HBM, projects, copies u/v/w (and pressure/indices) back. */
int elq_project(elq_ctx* ctx, const elq_grid* host_grid, float dt, const elq_opts* opts,
                elq_stats* stats);

/* Same on grids already resident in device memory (pointers from
 * elq_device_alloc or any hipMalloc of the same device). Runs on the
 * context's stream and returns after completion. */
int elq_project_device(elq_ctx* ctx, const elq_grid* device_grid, float dt, const elq_opts* opts,
                       elq_stats* stats);

/* Allocation-free variant of elq_project_device. No reference counterpart:
 * it is for callers that cannot allocate device memory while they run (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It calls no hipMalloc, hipFree or other allocating routine;
 * every device buffer of the projection (mask, indices, bricks, system,
 * factor, Krylov vectors, exchange words, control block) is carved from one
 * static device array of ELQ_NOALLOC_SCRATCH_BYTES per device, held from the
 * library's first use on that device until the process ends. A grid whose
 * workspace does not fit is refused with ELQ_ERR_ADMISSION before any work
 * (roughly up to 64^3 cells, or larger grids with fewer liquid bricks).
 * Outputs are the same as elq_project_device's for the same inputs and
 * context. Every call on a device shares that device's array: calls that
 * overlap on one device (several contexts or threads) race and must be
 * serialised by the caller; each call clears what must start at zero. */
#define ELQ_NOALLOC_SCRATCH_BYTES (64u << 20)
int elq_project_device_noalloc(elq_ctx* ctx, const elq_grid* device_grid, float dt, const elq_opts* opts,
                               elq_stats* stats);

/* projection.cpp:14-28 on host grids: writes indices (ELQ_UNKNOWN_INDEX for
 * non-unknowns) and *n_unknowns. */
int elq_assign_indices(elq_ctx* ctx, const elq_grid* host_grid, int32_t* indices,
                       int64_t* n_unknowns);

/* projection.cpp:30-83 on host grids, stencil form in the reference's unknown
 * numbering: A(n,n) = diag[n]; A(n, +x neighbour) = ax[n] (0 when the
 * neighbour is not an unknown), same for ay/az; A is symmetric. b as the
 * reference builds it. Arrays have n_unknowns entries (from
 * elq_assign_indices). */
int elq_build_matrix(elq_ctx* ctx, const elq_grid* host_grid, float dt, double* diag, double* ax,
                     double* ay, double* az, double* b);

/* The preconditioner set-up of simulator.cpp:102-107 (solver.compute(A) of
 * Eigen's IncompleteCholesky<double>), in the form this library factors: the
 * natural-order IC(0) (opts->precond ELQ_PRECOND_MIC0: MIC(0) with opts->tau;
 * pivot safeguard opts->sigma) of the matrix elq_build_matrix exports, on
 * host grids. Writes the pivot inverses 1/e, FP64, in the reference's
 * unknown numbering (n_unknowns entries, from elq_assign_indices): the
 * factor M = (E+L) E^{-1} (E+L^T) with L the strict lower triangle of A.
 * opts NULL: defaults. */
int elq_ic_factor(elq_ctx* ctx, const elq_grid* host_grid, float dt, const elq_opts* opts, double* inv_e);

/* projection.cpp:93-134 on host grids: u/v/w updated in place from a given
 * pressure grid (grid->pressure is the input here). */
int elq_update_velocity(elq_ctx* ctx, const elq_grid* host_grid, float dt);

/* The steps before the projection in the reference's frame, on device
 * grids (same layouts), for callers that keep the whole frame in HBM:
 * compute_staggered_grid_weight (boundary.cpp:6-34) from the nodal solid
 * level set ((ni+1)*(nj+1)*(nk+1) floats, index i + (ni+1)*(j + (nj+1)*k));
 * Simulator::add_gravity (simulator.cpp:90-93) on v; find_valid_velocity +
 * set_invalid_velocity_zero (simulator.cpp:125-178), which reads level_set
 * and the weights of device_grid and zeroes u/v/w on invalid faces. Each
 * returns after completion. */
int elq_compute_weights_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, const float* solid_nodal,
                               float* u_weight, float* v_weight, float* w_weight);
int elq_add_gravity_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float* v, float dt);
int elq_valid_and_zero_device(elq_ctx* ctx, const elq_grid* device_grid);

/* The same two steps one at a time, with the flags (one byte per face, 0/1,
 * the face grids' sizes and layouts) in device memory, as the reference keeps
 * them in Simulator::m_velocity_valid for extrapolate_velocity:
 * find_valid_velocity (simulator.cpp:125-158) writes them from level_set and
 * the weights of device_grid; set_invalid_velocity_zero (simulator.cpp:169-178)
 * zeroes u/v/w where they are 0. */
int elq_find_valid_velocity_device(elq_ctx* ctx, const elq_grid* device_grid, uint8_t* u_valid,
                                   uint8_t* v_valid, uint8_t* w_valid);
int elq_set_invalid_velocity_zero_device(elq_ctx* ctx, const elq_grid* device_grid, const uint8_t* u_valid,
                                         const uint8_t* v_valid, const uint8_t* w_valid);
/* calc_cell_centered_solid_phi + find_valid_level_set (simulator.cpp:191-200,
 * 160-167): valid (ni*nj*nk bytes) is 0 where the mean of a cell's 8 solid
 * nodes is < 0, else 1. */
int elq_find_valid_level_set_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, const float* solid_nodal,
                                    uint8_t* valid);

/* The steps after the projection: extrapolation into invalid cells
 * (src/project/3d/fluid/extrapolation.cpp:5-51), on device memory,
 * bit-identical with the reference. A grid of ni*nj*nk floats (index
 * i + ni*(j + nj*k)) with one 0/1 byte per cell; for up to `layers` layers,
 * every invalid cell with a valid neighbour among the 26 around it takes the
 * float mean of those neighbours' values (summed with k slowest, i fastest)
 * and becomes valid. layers <= 0 does nothing.
 *   elq_extrapolate_device        fluid::extrapolate: grid and valid in place
 *                                 (extrapolation.cpp:5-46)
 *   elq_extrapolate_const_device  fluid::extrapolate_const: valid unchanged
 *                                 (extrapolation.cpp:48-51), e.g. the level set
 *                                 with its valid flags (simulator.cpp:115-117)
 *   elq_extrapolate_velocity_device  Simulator::extrapolate_velocity
 *                                 (simulator.cpp:119-123): u, v, w of
 *                                 device_grid with their flags, all three in
 *                                 the same launches (level_set and weights
 *                                 are not read). */
int elq_extrapolate_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float* grid, uint8_t* valid,
                           int32_t layers);
int elq_extrapolate_const_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float* grid,
                                 const uint8_t* valid, int32_t layers);
int elq_extrapolate_velocity_device(elq_ctx* ctx, const elq_grid* device_grid, uint8_t* u_valid,
                                    uint8_t* v_valid, uint8_t* w_valid, int32_t layers);

/* Semi-Lagrangian advection (src/project/3d/fluid/advection.cpp:9-27), on
 * device memory, bit-identical with the reference: each sample point is
 * traced back with the RK2 midpoint rule over -dt (include/nama/math/ode.h:30-35)
 * through the trilinearly interpolated face velocities, and the advected
 * grid is interpolated trilinearly there, indices clamped to the grid
 * (include/nama/math/interpolation.h:10-72,186-193). Cell (i,j,k) sits at
 * ((i+0.5)dx, (j+0.5)dx, (k+0.5)dx); face grids are shifted by -dx/2 on
 * their own axis. Outputs must not alias inputs.
 *   elq_advect_staggered_device   fluid::advect_staggered_grid: out_{u,v,w} =
 *                                 val_{u,v,w} carried by vel_{u,v,w}
 *   elq_advect_collocated_device  fluid::advect_collocated_grid: out (cell
 *                                 grid) = val carried by vel_{u,v,w}
 *   elq_advect_velocity_device    Simulator::advect_velocity (simulator.cpp:79-82):
 *                                 u, v, w of device_grid advected by themselves,
 *                                 in place (the library keeps the copy)
 *   elq_advect_level_set_device   Simulator::advect_level_set (simulator.cpp:84-87):
 *                                 level_set advected by u, v, w, in place */
int elq_advect_staggered_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float dx, float dt,
                                float* out_u, float* out_v, float* out_w, const float* val_u, const float* val_v,
                                const float* val_w, const float* vel_u, const float* vel_v, const float* vel_w);
int elq_advect_collocated_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float dx, float dt, float* out,
                                 const float* val, const float* vel_u, const float* vel_v, const float* vel_w);
int elq_advect_velocity_device(elq_ctx* ctx, const elq_grid* device_grid, float dt);
int elq_advect_level_set_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float dx, float dt,
                                float* level_set, const float* u, const float* v, const float* w);

/* fluid::reinitialize_fast_marching (src/project/3d/fluid/reinitialization.cpp:19-156),
 * Simulator::reinitialize (simulator.cpp:181-183), on device memory,
 * bit-identical with the reference: cells whose level set changes sign
 * towards a face neighbour take their geometric distance to the
 * interpolated zero set (include/nama/math/level_set.h:165-264); the other
 * valid cells (valid: ni*nj*nk bytes, 0/1) are reached one face-neighbour
 * layer per pass and take the first-order upwind distance from the cells
 * fixed before. level_set is updated in place. *unreached (optional): valid
 * cells no path of valid cells connects to the interface; they keep their
 * value (the reference loops forever on them). *passes (optional): marching
 * passes that fixed cells. */
int elq_reinitialize_device(elq_ctx* ctx, int32_t ni, int32_t nj, int32_t nk, float dx, float* level_set,
                            const uint8_t* valid, int64_t* unreached, int32_t* passes);

/* Device memory helpers for callers without their own allocator. */
int elq_device_alloc(elq_ctx* ctx, size_t bytes, void** out);
int elq_device_free(elq_ctx* ctx, void* ptr);
int elq_memcpy_h2d(elq_ctx* ctx, void* dst, const void* src, size_t bytes);
int elq_memcpy_d2h(elq_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy with the library's own streaming kernel; bytes and
 * both pointers must be multiples of 4. No reference counterpart: a plain
 * copy of known size, whose hardware counters calibrate traffic figures. */
int elq_memcpy_d2d(elq_ctx* ctx, void* dst, const void* src, size_t bytes);

/* Per-kernel timing of the last projection on this context, measured with
 * HIP events on the context's stream. names: comma-separated kernel classes;
 * ms[i]: total time, launches[i]: launch count. Returns the number of classes
 * written (at most cap). */
int elq_last_kernel_times(const elq_ctx* ctx, char* names, size_t names_cap, double* ms,
                          int64_t* launches, int cap);
/* Enables (1) or disables (0) per-kernel event timing (adds event records). */
void elq_set_profiling(elq_ctx* ctx, int enable);
/* Restricts event timing to the named kernel classes (comma-separated, e.g.
 * "fwd_sweep_update"); NULL or "" times every class. */
void elq_set_profiling_filter(elq_ctx* ctx, const char* classes);
/* Caps the grid of the persistent solver kernels at max_groups workgroups
 * (0: as many as the device holds), a share of the device for solves that
 * run beside other work; the serving slots use it. No reference
 * counterpart. Results depend on the cap (the reduction tree and the
 * sweeps' block shape follow the grid), never on timing. Returns
 * ELQ_ERR_INVALID_ARG for a negative cap. */
int elq_set_max_groups(elq_ctx* ctx, int max_groups);

/* ------------------------------------------------------------------ serving
 * Continuous batching of independent projections (uneven grid sizes) on one
 * device; no reference counterpart (the reference solves one grid per frame).
 * A server owns max_slots solver slots, each with its own stream and
 * device-resident workspace (grids, system, factor, Krylov vectors) that
 * persists across requests. Requests wait in a FIFO queue and are admitted
 * as soon as a slot is free and the device memory held by all slots, with
 * this request's footprint, fits hbm_budget_bytes (a request that does not
 * fit lets later, smaller ones pass at most 8 times). Admitted requests run
 * concurrently; each gets the same fixed share of the device's workgroup
 * slots, so its result does not depend on which requests share the device.
 */
typedef struct elq_server elq_server;

typedef struct {
    size_t hbm_budget_bytes; /* device bytes the slots may hold (0: 90% of free memory) */
    int32_t max_slots;       /* concurrently running requests, 1..16 */
    int32_t max_queue;       /* waiting requests; submit returns ELQ_ERR_BUSY beyond */
    elq_opts opts;           /* solver options for every request */
    int32_t batched;         /* > 0: the Krylov iterations of the running
                                requests advance together in lockstep, one
                                launch per kernel class for all of them,
                                replayed from a HIP graph captured once (mixed
                                precision, wavefront schedule; ignored
                                otherwise), in `batched` groups of slots (slot
                                k in group k % batched) on streams of their
                                own; 0: each slot launches its own */
    int32_t batch_wait_us;   /* batched: at most how long a chunk waits for the
                                other requests inside their Krylov loops to join
                                it (0: 20 ms) */
} elq_server_opts;

typedef struct {
    int64_t id;
    int32_t status;           /* elq_status of the solve */
    int32_t iterations;
    int32_t outer_iterations;
    int32_t slot;             /* slot that ran it */
    int64_t n_unknowns;
    double error;
    double true_rel_residual;
    double t_submit_ms;       /* host clock, ms since server creation */
    double t_admit_ms;
    double t_done_ms;         /* outputs written back to the request's arrays */
} elq_result;

int elq_server_create(int device, const elq_server_opts* opts, elq_server** out);
/* Waits for running requests, drops queued ones, frees every slot. */
void elq_server_destroy(elq_server* srv);
/* Queues a projection on host grids (elq_project semantics). The grid arrays
 * must stay valid and untouched until its result has been polled. Returns
 * ELQ_ERR_ADMISSION if the request alone exceeds the budget, ELQ_ERR_BUSY if
 * the queue is full. */
int elq_server_submit(elq_server* srv, const elq_grid* host_grid, float dt, int64_t* id);
/* Blocks until at least one result can be polled or timeout_ms passes
 * (< 0: no limit). Returns the number of results ready. */
int elq_server_wait(elq_server* srv, int timeout_ms);
/* Pops up to cap finished results (completion order). */
int elq_server_poll(elq_server* srv, elq_result* out, int cap);
/* Requests queued or running. */
int elq_server_in_flight(const elq_server* srv);
/* Device bytes the slots hold now. */
size_t elq_server_device_bytes(const elq_server* srv);
/* Batched servers: batched Krylov chunks launched so far and the requests
 * they advanced (jobs / batches = mean batch size); 0 and 0 otherwise. No
 * reference counterpart (serving only). */
int elq_server_batch_stats(const elq_server* srv, int64_t* batches, int64_t* jobs);

/* Device bytes a projection requests from the allocator on a fresh context,
 * host inputs staged (an upper bound; level_set NULL counts every brick,
 * otherwise bricks holding a cell with phi < 0). 0 for invalid sizes. */
size_t elq_projection_footprint(int32_t ni, int32_t nj, int32_t nk, int32_t precision, const float* level_set);
/* Device bytes a context's workspace holds now. */
size_t elq_ctx_device_bytes(const elq_ctx* ctx);

/* ------------------------------------------------------------------ scenes
 * Host-side input generation for the synthetic workloads (the simulator's own
 * set-up steps, not part of the projection): create_grid + emitters +
 * init_scene + compute_weights (simulator.cpp:52-76,187-189;
 * levelset_emitter.cpp:10-32; boundary.cpp:6-34), add_gravity
 * (simulator.cpp:90-93), find_valid_velocity + set_invalid_velocity_zero
 * (simulator.cpp:125-178). spheres: n_spheres*(cx,cy,cz,r). slab_height
 * <= -1e30 disables the pool emitter. */
int elq_scene_level_set(int32_t ni, int32_t nj, int32_t nk, float dx, float slab_height,
                        int32_t n_spheres, const float* spheres, float* level_set);
int elq_scene_weights(int32_t ni, int32_t nj, int32_t nk, float* u_weight, float* v_weight,
                      float* w_weight);
/* The room() solid (levelset_emitter.cpp:18-32) on the nodal grid, the input
 * of compute_weights: (ni+1)*(nj+1)*(nk+1) floats, i fastest. */
int elq_scene_room_nodal(int32_t ni, int32_t nj, int32_t nk, float* nodal);
int elq_scene_add_gravity(int32_t ni, int32_t nj, int32_t nk, float* v, float dt);
int elq_scene_valid_and_zero(const elq_grid* host_grid);
/* SplitMix64-driven draws: spheres (centres U[c_lo,c_hi]*L_axis, radii
 * U[r_lo,r_hi], drawn cx,cy,cz,r per sphere) and face velocities U[-amp,amp]
 * (u, then v, then w, each in linear index order). state is in/out. */
int elq_scene_random_spheres(uint64_t* state, int32_t count, double c_lo, double c_hi, double r_lo,
                             double r_hi, double lx, double ly, double lz, float* spheres);
int elq_scene_random_velocity(uint64_t* state, int32_t ni, int32_t nj, int32_t nk, double amp,
                              float* u, float* v, float* w);
double elq_uniform01(uint64_t* state);

#ifdef __cplusplus
}
#endif
#endif /* ELQ_H */
