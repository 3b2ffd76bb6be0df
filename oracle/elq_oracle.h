/*
 * CPU oracle for the pressure projection of daichi-nama/eulerian-liquid.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
 * product (eulerian-liquid_amd/libelq.so) never links, loads or calls it.
 *
 * It restates, in plain C99 with FP64 where the reference uses double:
 *   - src/project/3d/fluid/projection.cpp      (whole file)
 *   - src/project/3d/eulerian_liquid/simulator.cpp:52-76,89-112,125-178
 *   - src/project/3d/fluid/boundary.cpp:6-34
 *   - src/project/3d/fluid/levelset_emitter.cpp:10-32,74-79
 *   - include/nama/math/level_set.h:68-163
 *   - include/nama/common/array.h:281-286 (i-fastest layout), 424-472 (loops)
 *   - include/nama/common/grid.h:355-367 (cell centres)
 *   - Eigen::ConjugateGradient's documented loop (Eigen is not in the
 *     reference tree; see the header comment of orc_pcg).
 * The preconditioner is natural-order IC(0) (optionally MIC(0)), the same
 * algebra the GPU path uses; the reference's Eigen IncompleteCholesky
 * (AMD ordering + ICT) cannot be built here.
 *
 * Layout: every grid is a flat float array, index i + ni*(j + nj*k).
 * Face grids: u (ni+1)*nj*nk, v ni*(nj+1)*nk, w ni*nj*(nk+1).
 * Unknowns are numbered in the reference's lexicographic k->j->i order.
 */
#ifndef ELQ_ORACLE_H
#define ELQ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_UNKNOWN_INDEX 2147483647 /* projection.cpp:6 (INT_MAX) */

/* level_set.h:68-80 */
float orc_fraction_inside2(float phi_left, float phi_right);
/* level_set.h:83-163 */
float orc_fraction_inside4(float phi_bl, float phi_br, float phi_tl, float phi_tr);

/* levelset_emitter.cpp:18-32 through LevelSetEmitterList::apply (:74-79).
 * nodal has (ni+1)*(nj+1)*(nk+1) entries; ni,nj,nk are the CELL counts. */
void orc_room_nodal(int ni, int nj, int nk, float* nodal);

/* Liquid level set: FLT_MAX, then min with pool(slab_height) if
 * slab_height > -1e30, then min with each sphere (cx,cy,cz,r).
 * levelset_emitter.cpp:10-16,74-79; grid.h:355-367 (origin 0). */
void orc_liquid_level_set(int ni, int nj, int nk, float dx, float slab_height,
                          int n_spheres, const float* spheres, float* phi);

/* boundary.cpp:6-34 */
void orc_compute_weights(int ni, int nj, int nk, const float* nodal,
                         float* uw, float* vw, float* ww);

/* simulator.cpp:90-93 */
void orc_add_gravity(int ni, int nj, int nk, float* v, float dt);

/* simulator.cpp:125-167 (find_valid_velocity): 0/1 flags per face */
void orc_find_valid_velocity(int ni, int nj, int nk, const float* phi, const float* uw,
                             const float* vw, const float* ww, uint8_t* u_valid,
                             uint8_t* v_valid, uint8_t* w_valid);

/* simulator.cpp:169-178 (set_invalid_velocity_zero) on one face grid of n faces */
void orc_set_invalid_zero(size_t n, const uint8_t* valid, float* f);

/* simulator.cpp:125-167 then :169-178 (the flags are not returned) */
void orc_find_valid_and_zero(int ni, int nj, int nk, const float* phi,
                             const float* uw, const float* vw, const float* ww,
                             float* u, float* v, float* w);

/* advection.cpp:19-27: out_* = val_* carried by vel_* (semi-Lagrangian RK2) */
void orc_advect_staggered(int ni, int nj, int nk, float dx, float dt, const float* vel_u, const float* vel_v,
                          const float* vel_w, const float* val_u, const float* val_v, const float* val_w,
                          float* out_u, float* out_v, float* out_w);

/* advection.cpp:19-27 via simulator.cpp:79-82: semi-Lagrangian RK2
 * advection of the face grids u0, v0, w0 by themselves into u, v, w */
void orc_advect_velocity(int ni, int nj, int nk, float dx, float dt, const float* u0, const float* v0,
                         const float* w0, float* u, float* v, float* w);

/* advection.cpp:9-17 via simulator.cpp:84-87: phi0 advected by u, v, w into phi */
void orc_advect_level_set(int ni, int nj, int nk, float dx, float dt, const float* phi0, const float* u,
                          const float* v, const float* w, float* phi);

/* reinitialization.cpp:19-156: level set grid (in place) with its 0/1 flags.
 * Returns the FAR cells left when a pass changes nothing (0 whenever the
 * reference terminates); *passes = marching passes that changed labels. */
long orc_reinitialize(int ni, int nj, int nk, float dx, float* grid, const uint8_t* valid, int* passes);

/* simulator.cpp:191-200 + 160-167: level-set flags from the nodal solid phi */
void orc_find_valid_level_set(int ni, int nj, int nk, const float* nodal, uint8_t* valid);

/* extrapolation.cpp:5-46: grid and valid (0/1) of ni*nj*nk cells, in place */
void orc_extrapolate(int ni, int nj, int nk, float* grid, uint8_t* valid, int layers);

/* extrapolation.cpp:48-51: valid is left unchanged */
void orc_extrapolate_const(int ni, int nj, int nk, float* grid, const uint8_t* valid,
                           int layers);

/* projection.cpp:14-28. Returns N (>0), 0 for "No fluid", or -1 when an
 * unknown lies on the outer cell layer (the reference then reads outside the
 * arrays in build_matrix, which is undefined behaviour). */
int orc_assign_indices(int ni, int nj, int nk, const float* phi,
                       const float* uw, const float* vw, const float* ww,
                       int32_t* indices);

/* Linear system in the reference's unknown numbering, stencil form.
 * ax[n]: entry A(n, index of +x neighbour) (0 when that neighbour is not an
 * unknown); same for ay/az. Neighbour index arrays hold -1 when absent. */
typedef struct {
    int n;
    double* diag;
    double* ax;
    double* ay;
    double* az;
    int32_t* xp;
    int32_t* yp;
    int32_t* zp;
    int32_t* xm;
    int32_t* ym;
    int32_t* zm;
    double* b;
} orc_system;

/* projection.cpp:30-83 (diag/b accumulation order and float->double
 * promotion points kept). */
void orc_build_matrix(int ni, int nj, int nk, float dx, float dt,
                      const int32_t* indices, const float* phi,
                      const float* u, const float* v, const float* w,
                      const float* uw, const float* vw, const float* ww,
                      orc_system* sys);

/* CSR export of orc_system with Eigen::SparseMatrix<double,RowMajor>
 * semantics: columns sorted, explicit zeros kept (projection.cpp:55-58,81).
 * row_ptr has n+1 entries; col/val at most 7n. Returns nnz. */
int orc_to_csr(const orc_system* sys, int32_t* row_ptr, int32_t* col, double* val);

/* IC(0) (tau = 0) or MIC(0) (tau > 0) in natural order; pivot safeguard:
 * a pivot t with !(t > sigma*diag) is replaced by diag. Writes inv_e = 1/e.
 * Returns the number of safeguarded pivots. */
int orc_ic_factor(const orc_system* sys, double tau, double sigma, double* inv_e);

/* z = M^{-1} r for M = (E + L) E^{-1} (E + L^T). */
void orc_ic_apply(const orc_system* sys, const double* inv_e, const double* r, double* z);

/* y = A x */
void orc_spmv(const orc_system* sys, const double* x, double* y);
/* OpenMP threads (the scatter and update loops run in parallel, as in the
 * reference); the Krylov companion parallelises SpMV, axpys and dots. */
void orc_set_threads(int n);
int orc_max_threads(void);
void orc_set_krylov_parallel(int on);
/* every > 0: orc_pcg prints its iteration count and |r|^2/|b|^2 to stderr
 * every `every` iterations (progress of long solves); 0: silent. */
void orc_set_progress(int every);

/* Preconditioned CG with Eigen's ConjugateGradient semantics (x0 = 0,
 * recursive residual, stop when |r|^2 < max(tol^2 |b|^2, DBL_MIN)).
 * Returns 0; *iters and *err follow Eigen's iterations()/error(). */
int orc_pcg(const orc_system* sys, const double* inv_e, double tol, int max_iter,
            double* x, int* iters, double* err);

/* projection.cpp:85-91 */
void orc_build_grid_from_vector(int ni, int nj, int nk, const double* x,
                                const int32_t* indices, float background, float* pressure);

/* projection.cpp:93-134 */
void orc_update_velocity(int ni, int nj, int nk, float dx, float dt,
                         const float* pressure, const float* phi,
                         float* u, float* v, float* w);

/* Simulator::solve_pressure (simulator.cpp:96-112) with the solver replaced
 * by natural-order IC(0)/MIC(0) PCG. Outputs indices, pressure and updated
 * u,v,w. Returns N (>0), 0 (No fluid), -1 (unknown on outer layer). */
int orc_solve_pressure(int ni, int nj, int nk, float dx, float dt,
                       const float* phi, float* u, float* v, float* w,
                       const float* uw, const float* vw, const float* ww,
                       double tol, int max_iter, double tau, double sigma,
                       int32_t* indices, float* pressure,
                       int* iters, double* err, double* true_rel_res);

/* SplitMix64 stream shared with the scene generator of the GPU package. */
uint64_t orc_splitmix64(uint64_t* state);
double orc_uniform01(uint64_t* state);

#ifdef __cplusplus
}
#endif
#endif
