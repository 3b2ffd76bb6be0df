/*
 * CPU oracle (checker) for the pressure projection. TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
 * never by the product library. Every function cites the reference lines it
 * restates (paths relative to the reference repository root).
 *
 * Floating-point promotion points follow the reference exactly (float
 * products/divisions promoted to double where the reference accumulates in
 * double). Build without FMA contraction (-ffp-contract=off).
 */
#include "elq_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, k, ni, nj) ((size_t)(i) + (size_t)(ni) * ((size_t)(j) + (size_t)(nj) * (size_t)(k)))

/* std::min / std::max / std::clamp as the reference's <algorithm> defines them */
static float min_f(float a, float b) { return (b < a) ? b : a; }
static float max_f(float a, float b) { return (a < b) ? b : a; }
static float clamp_f(float v, float lo, float hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

/* level_set.h:68-80 */
float orc_fraction_inside2(float phi_left, float phi_right) {
    if (phi_left < 0.0f && phi_right < 0.0f) return 1.0f;
    if (phi_left < 0.0f && phi_right >= 0.0f) return phi_left / (phi_left - phi_right);
    if (phi_left >= 0.0f && phi_right < 0.0f) return phi_right / (phi_right - phi_left);
    return 0.0f;
}

static void cycle4(float* a) {
    float t = a[0];
    a[0] = a[1];
    a[1] = a[2];
    a[2] = a[3];
    a[3] = t;
}

/* level_set.h:83-163. The 0.5/0.25 literals are double in the reference, so
 * those products are evaluated in double and rounded to float on assignment
 * or return, exactly as below. */
float orc_fraction_inside4(float phi_bl, float phi_br, float phi_tl, float phi_tr) {
    int inside_count = (phi_bl < 0.0f ? 1 : 0) + (phi_tl < 0.0f ? 1 : 0) + (phi_br < 0.0f ? 1 : 0) +
                       (phi_tr < 0.0f ? 1 : 0);
    float list[4] = {phi_bl, phi_br, phi_tr, phi_tl};
    if (inside_count == 4) {
        return 1.0f;
    } else if (inside_count == 3) {
        while (list[0] < 0.0f) cycle4(list);
        float side0 = 1.0f - orc_fraction_inside2(list[0], list[3]);
        float side1 = 1.0f - orc_fraction_inside2(list[0], list[1]);
        return (float)(1.0 - 0.5 * side0 * side1);
    } else if (inside_count == 2) {
        while (list[0] >= 0.0f || !(list[1] < 0.0f || list[2] < 0.0f)) cycle4(list);
        if (list[1] < 0.0f) {
            float side_left = orc_fraction_inside2(list[0], list[3]);
            float side_right = orc_fraction_inside2(list[1], list[2]);
            return (float)(0.5 * (side_left + side_right));
        } else {
            float middle_point = (float)(0.25 * (list[0] + list[1] + list[2] + list[3]));
            if (middle_point < 0.0f) {
                float area = 0.0f;
                float side1 = 1.0f - orc_fraction_inside2(list[0], list[3]);
                float side3 = 1.0f - orc_fraction_inside2(list[2], list[3]);
                area = (float)(area + 0.5 * side1 * side3);
                float side2 = 1.0f - orc_fraction_inside2(list[2], list[1]);
                float side0 = 1.0f - orc_fraction_inside2(list[0], list[1]);
                area = (float)(area + 0.5 * side0 * side2);
                return 1.0f - area;
            } else {
                float area = 0.0f;
                float side0 = orc_fraction_inside2(list[0], list[1]);
                float side1 = orc_fraction_inside2(list[0], list[3]);
                area = (float)(area + 0.5 * side0 * side1);
                float side2 = orc_fraction_inside2(list[2], list[1]);
                float side3 = orc_fraction_inside2(list[2], list[3]);
                area = (float)(area + 0.5 * side2 * side3);
                return area;
            }
        }
    } else if (inside_count == 1) {
        while (list[0] >= 0.0f) cycle4(list);
        float side0 = orc_fraction_inside2(list[0], list[3]);
        float side1 = orc_fraction_inside2(list[0], list[1]);
        return (float)(0.5 * side0 * side1);
    }
    return 0.0f;
}

/* levelset_emitter.cpp:18-32 (room) applied through LevelSetEmitterList::apply
 * (:74-79: fill FLT_MAX, then min). Called in the reference with the nodal grid
 * dimensions (apps/eulerian_liquid_3d.cpp:65-66). */
void orc_room_nodal(int ni, int nj, int nk, float* nodal) {
    const int a = ni + 1, b = nj + 1, c = nk + 1;
    for (int k = 0; k < c; ++k)
        for (int j = 0; j < b; ++j)
            for (int i = 0; i < a; ++i) {
                float val;
                if (i == 0 || i == a - 1 || j == 0 || j == b - 1 || k == 0 || k == c - 1)
                    val = (float)-1.0;
                else if (i == 1 || i == a - 2 || j == 1 || j == b - 2 || k == 1 || k == c - 2)
                    val = (float)-0.01;
                else
                    val = (float)1.0;
                nodal[IDX(i, j, k, a, b)] = min_f(FLT_MAX, val);
            }
}

/* levelset_emitter.cpp:10-16 (sphere, pool), :74-79 (list apply) and the cell
 * centre grid.h:355-367: pos = (i+0.5f, j+0.5f, k+0.5f)*dx + origin(0).
 * Eigen's fixed-size norm is sqrt(d0^2 + (d1^2 + d2^2)) (unrolled halving
 * reduction). */
void orc_liquid_level_set(int ni, int nj, int nk, float dx, float slab_height, int n_spheres,
                          const float* spheres, float* phi) {
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                float px = ((float)i + 0.5f) * dx + 0.0f;
                float py = ((float)j + 0.5f) * dx + 0.0f;
                float pz = ((float)k + 0.5f) * dx + 0.0f;
                float val = FLT_MAX;
                for (int s = 0; s < n_spheres; ++s) {
                    float d0 = px - spheres[4 * s + 0];
                    float d1 = py - spheres[4 * s + 1];
                    float d2 = pz - spheres[4 * s + 2];
                    float sq = d0 * d0 + (d1 * d1 + d2 * d2);
                    val = min_f(val, sqrtf(sq) - spheres[4 * s + 3]);
                }
                if (slab_height > -1e30f) val = min_f(val, py - slab_height);
                phi[IDX(i, j, k, ni, nj)] = val;
            }
}

/* boundary.cpp:6-34 */
void orc_compute_weights(int ni, int nj, int nk, const float* nodal, float* uw, float* vw, float* ww) {
    const int a = ni + 1, b = nj + 1;
#define N(i, j, k) nodal[IDX(i, j, k, a, b)]
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni + 1; ++i) {
                float wgt = 1.0f - orc_fraction_inside4(N(i, j, k), N(i, j + 1, k), N(i, j, k + 1),
                                                        N(i, j + 1, k + 1));
                uw[IDX(i, j, k, ni + 1, nj)] = clamp_f(wgt, 0.0f, 1.0f);
            }
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj + 1; ++j)
            for (int i = 0; i < ni; ++i) {
                float wgt = 1.0f - orc_fraction_inside4(N(i, j, k), N(i, j, k + 1), N(i + 1, j, k),
                                                        N(i + 1, j, k + 1));
                vw[IDX(i, j, k, ni, nj + 1)] = clamp_f(wgt, 0.0f, 1.0f);
            }
    for (int k = 0; k < nk + 1; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                float wgt = 1.0f - orc_fraction_inside4(N(i, j, k), N(i, j + 1, k), N(i + 1, j, k),
                                                        N(i + 1, j + 1, k));
                ww[IDX(i, j, k, ni, nj)] = clamp_f(wgt, 0.0f, 1.0f);
            }
#undef N
}

/* simulator.cpp:90-93 with constants.h:5-6 (gravity<float> = 9.80665f) */
void orc_add_gravity(int ni, int nj, int nk, float* v, float dt) {
    const size_t n = (size_t)ni * (size_t)(nj + 1) * (size_t)nk;
    const float g = 9.80665f * dt;
    for (size_t c = 0; c < n; ++c) v[c] = v[c] - g;
}

/* simulator.cpp:125-158 (find_valid_velocity) then :169-178
 * (set_invalid_velocity_zero). */
/* simulator.cpp:125-167 (find_valid_velocity): every face starts valid iff
 * its weight is non-zero; then, over the reference's loop ranges [1, n-1) of
 * each face grid, a face between two cells with phi >= 0 is invalid. */
void orc_find_valid_velocity(int ni, int nj, int nk, const float* phi, const float* uw,
                             const float* vw, const float* ww, uint8_t* uv, uint8_t* vv,
                             uint8_t* wv) {
    const size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk,
                 nw = (size_t)ni * nj * (nk + 1);
    for (size_t c = 0; c < nu; ++c) uv[c] = (uw[c] == 0.0f) ? 0 : 1;
    for (size_t c = 0; c < nv; ++c) vv[c] = (vw[c] == 0.0f) ? 0 : 1;
    for (size_t c = 0; c < nw; ++c) wv[c] = (ww[c] == 0.0f) ? 0 : 1;
#define P(i, j, k) phi[IDX(i, j, k, ni, nj)]
    for (int k = 1; k < nk - 1; ++k)
        for (int j = 1; j < nj - 1; ++j)
            for (int i = 1; i < ni; ++i)
                if (P(i - 1, j, k) >= 0.0f && P(i, j, k) >= 0.0f) uv[IDX(i, j, k, ni + 1, nj)] = 0;
    for (int k = 1; k < nk - 1; ++k)
        for (int j = 1; j < nj; ++j)
            for (int i = 1; i < ni - 1; ++i)
                if (P(i, j - 1, k) >= 0.0f && P(i, j, k) >= 0.0f) vv[IDX(i, j, k, ni, nj + 1)] = 0;
    for (int k = 1; k < nk; ++k)
        for (int j = 1; j < nj - 1; ++j)
            for (int i = 1; i < ni - 1; ++i)
                if (P(i, j, k - 1) >= 0.0f && P(i, j, k) >= 0.0f) wv[IDX(i, j, k, ni, nj)] = 0;
#undef P
}

/* simulator.cpp:169-178 (set_invalid_velocity_zero), one face grid */
void orc_set_invalid_zero(size_t n, const uint8_t* valid, float* f) {
    for (size_t c = 0; c < n; ++c)
        if (valid[c] == 0) f[c] = 0.0f;
}

void orc_find_valid_and_zero(int ni, int nj, int nk, const float* phi, const float* uw,
                             const float* vw, const float* ww, float* u, float* v, float* w) {
    const size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk,
                 nw = (size_t)ni * nj * (nk + 1);
    uint8_t* uv = (uint8_t*)malloc(nu);
    uint8_t* vv = (uint8_t*)malloc(nv);
    uint8_t* wv = (uint8_t*)malloc(nw);
    orc_find_valid_velocity(ni, nj, nk, phi, uw, vw, ww, uv, vv, wv);
    orc_set_invalid_zero(nu, uv, u);
    orc_set_invalid_zero(nv, vv, v);
    orc_set_invalid_zero(nw, wv, w);
    free(uv);
    free(vv);
    free(wv);
}

/* ---------------------------------------------------------------- advection
 * A 3-D grid of floats with the reference's Grid3 geometry: value (i,j,k) at
 * ((i + 0.5) dx, (j + 0.5) dx, (k + 0.5) dx) + origin (grid.h:355-367). The
 * simulator's grids have origin 0; the face grids are shifted by -dx/2 on
 * their own axis (staggered_grid.h:109-111). */
typedef struct {
    const float* a;
    int ni, nj, nk;
    float o[3];
} orc_field;

static orc_field face_field(const float* a, int axis, int ni, int nj, int nk, float dx) {
    orc_field f;
    f.a = a;
    f.ni = ni + (axis == 0);
    f.nj = nj + (axis == 1);
    f.nk = nk + (axis == 2);
    /* origin - vector_type(0.5, 0, 0) * dx with origin 0: 0 - 0.5f * dx on the
     * axis, 0 - 0 * dx = +0 elsewhere */
    for (int c = 0; c < 3; ++c) f.o[c] = 0.0f - ((c == axis) ? 0.5f : 0.0f) * dx;
    return f;
}

/* interpolation.h:10-15 (calc_index_and_fraction) */
static int index_and_fraction(float pos, float origin, float dx, float* frac) {
    const float t = (pos - origin - 0.5f * dx) / dx;
    const int i = (int)floorf(t);
    *frac = t - (float)i;
    return i;
}

/* grid.h:370-379 (extrapolate: indices clamped to the grid) */
static float clamped(const orc_field* g, int i, int j, int k) {
    i = i < 0 ? 0 : (i > g->ni - 1 ? g->ni - 1 : i);
    j = j < 0 ? 0 : (j > g->nj - 1 ? g->nj - 1 : j);
    k = k < 0 ? 0 : (k > g->nk - 1 ? g->nk - 1 : k);
    return g->a[IDX(i, j, k, g->ni, g->nj)];
}

/* interpolation.h:17-21 */
static float lerp2(float v0, float v1, float f) { return (1.0f - f) * v0 + f * v1; }

/* interpolation.h:35-72 (dim 3) with bilerp/trilerp of :23-33: x first, then
 * y, then z */
static float trilerp_at(const orc_field* g, const float x[3], float dx) {
    float fx, fy, fz;
    const int i = index_and_fraction(x[0], g->o[0], dx, &fx);
    const int j = index_and_fraction(x[1], g->o[1], dx, &fy);
    const int k = index_and_fraction(x[2], g->o[2], dx, &fz);
    const float v0 = clamped(g, i, j, k), v1 = clamped(g, i + 1, j, k);
    const float v2 = clamped(g, i, j + 1, k), v3 = clamped(g, i + 1, j + 1, k);
    const float v4 = clamped(g, i, j, k + 1), v5 = clamped(g, i + 1, j, k + 1);
    const float v6 = clamped(g, i, j + 1, k + 1), v7 = clamped(g, i + 1, j + 1, k + 1);
    const float b0 = lerp2(lerp2(v0, v1, fx), lerp2(v2, v3, fx), fy);
    const float b1 = lerp2(lerp2(v4, v5, fx), lerp2(v6, v7, fx), fy);
    return lerp2(b0, b1, fz);
}

/* ode.h:8-19 (StaggeredGridField) + interpolation.h:186-193 (lerp_staggered) */
static void velocity_at(const orc_field vel[3], const float x[3], float dx, float out[3]) {
    for (int d = 0; d < 3; ++d) out[d] = trilerp_at(&vel[d], x, dx);
}

/* ode.h:30-35 (trace_rk2) with dt = -dt (advection.cpp:13,23) */
static void backtrace(const orc_field vel[3], const float pos[3], float dt, float dx, float out[3]) {
    const float h = -dt;
    float v[3], mid[3];
    velocity_at(vel, pos, dx, v);
    const float half = 0.5f * h;
    for (int c = 0; c < 3; ++c) mid[c] = pos[c] + half * v[c];
    velocity_at(vel, mid, dx, v);
    for (int c = 0; c < 3; ++c) out[c] = pos[c] + h * v[c];
}

/* advection.cpp:19-27 (advect_staggered_grid): out_* = val_* sampled at the
 * RK2 backtrace of each face through vel_*. */
void orc_advect_staggered(int ni, int nj, int nk, float dx, float dt, const float* vel_u, const float* vel_v,
                          const float* vel_w, const float* val_u, const float* val_v, const float* val_w,
                          float* out_u, float* out_v, float* out_w) {
    orc_field vel[3] = {face_field(vel_u, 0, ni, nj, nk, dx), face_field(vel_v, 1, ni, nj, nk, dx),
                        face_field(vel_w, 2, ni, nj, nk, dx)};
    orc_field val[3] = {face_field(val_u, 0, ni, nj, nk, dx), face_field(val_v, 1, ni, nj, nk, dx),
                        face_field(val_w, 2, ni, nj, nk, dx)};
    float* out[3] = {out_u, out_v, out_w};
    for (int d = 0; d < 3; ++d) {
        const orc_field* g = &val[d];
        for (int k = 0; k < g->nk; ++k)
            for (int j = 0; j < g->nj; ++j)
                for (int i = 0; i < g->ni; ++i) {
                    const float pos[3] = {((float)i + 0.5f) * dx + g->o[0], ((float)j + 0.5f) * dx + g->o[1],
                                          ((float)k + 0.5f) * dx + g->o[2]};
                    float back[3];
                    backtrace(vel, pos, dt, dx, back);
                    out[d][IDX(i, j, k, g->ni, g->nj)] = trilerp_at(g, back, dx);
                }
    }
}

/* Simulator::advect_velocity (simulator.cpp:79-82): the velocity field and
 * the advected values are both the copy taken before the step (u0, v0, w0);
 * u, v, w receive the result. */
void orc_advect_velocity(int ni, int nj, int nk, float dx, float dt, const float* u0, const float* v0,
                         const float* w0, float* u, float* v, float* w) {
    orc_advect_staggered(ni, nj, nk, dx, dt, u0, v0, w0, u0, v0, w0, u, v, w);
}

/* advection.cpp:9-17 (advect_collocated_grid) as Simulator::advect_level_set
 * calls it (simulator.cpp:84-87): phi0 is the copy before the step, the
 * velocity grids are the current ones; phi receives the result. */
void orc_advect_level_set(int ni, int nj, int nk, float dx, float dt, const float* phi0, const float* u,
                          const float* v, const float* w, float* phi) {
    orc_field vel[3] = {face_field(u, 0, ni, nj, nk, dx), face_field(v, 1, ni, nj, nk, dx),
                        face_field(w, 2, ni, nj, nk, dx)};
    orc_field g = {phi0, ni, nj, nk, {0.0f, 0.0f, 0.0f}};
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const float pos[3] = {((float)i + 0.5f) * dx + 0.0f, ((float)j + 0.5f) * dx + 0.0f,
                                      ((float)k + 0.5f) * dx + 0.0f};
                float back[3];
                backtrace(vel, pos, dt, dx, back);
                phi[IDX(i, j, k, ni, nj)] = trilerp_at(&g, back, dx);
            }
}

/* ------------------------------------------------------- reinitialisation
 * Eigen's Vector3f operations as the reference uses them: cross() is the
 * usual three 2x2 determinants; norm() is sqrt(squaredNorm()), and Eigen's
 * fixed-size reduction of a 3-vector splits it in halves, so the squares
 * are summed as x0 + (x1 + x2). */
static void sub3(const float a[3], const float b[3], float out[3]) {
    for (int c = 0; c < 3; ++c) out[c] = a[c] - b[c];
}

static void cross3(const float u[3], const float v[3], float out[3]) {
    out[0] = u[1] * v[2] - u[2] * v[1];
    out[1] = u[2] * v[0] - u[0] * v[2];
    out[2] = u[0] * v[1] - u[1] * v[0];
}

static float norm3(const float v[3]) { return sqrtf(v[0] * v[0] + (v[1] * v[1] + v[2] * v[2])); }

/* meshing.h:12-17 (vertex_interp) with isolevel 0 */
static void vertex_interp(const float p0[3], const float p1[3], float l0, float l1, float out[3]) {
    const float mu = -(l0 - 0.0f) / (l1 - l0);
    for (int c = 0; c < 3; ++c) out[c] = p0[c] + mu * (p1[c] - p0[c]);
}

/* level_set.h:172-177 */
static float line_distance(const float x0[3], const float x1[3], const float x2[3]) {
    float a[3], b[3], n[3];
    sub3(x0, x1, a);
    sub3(x2, x1, b);
    cross3(a, b, n);
    const float area = norm3(n);
    const float l = norm3(b);
    return area / l;
}

/* level_set.h:179-189 */
static float plane_distance(const float x0[3], const float x1[3], const float x2[3], const float x3[3]) {
    float u[3], v[3], n[3];
    sub3(x2, x1, u);
    sub3(x3, x1, v);
    cross3(u, v, n);
    const float a = n[0], b = n[1], c = n[2];
    const float d = -(n[0] * x1[0] + n[1] * x1[1] + n[2] * x1[2]);
    return fabsf(a * x0[0] + b * x0[1] + c * x0[2] + d) / sqrtf(a * a + b * b + c * c);
}

/* level_set.h:212-237 (calc_distance_from_zero, tetrahedron x0..x3) */
static float tet_distance(const float* x0, const float* x1, const float* x2, const float* x3, float phi0,
                          float phi1, float phi2, float phi3) {
    if (phi0 == 0.0f) return 0.0f;
    float xi[3][3];
    int n = 0;
    if (phi0 * phi1 <= 0.0f) vertex_interp(x0, x1, phi0, phi1, xi[n++]);
    if (phi0 * phi2 <= 0.0f) vertex_interp(x0, x2, phi0, phi2, xi[n++]);
    if (phi0 * phi3 <= 0.0f) vertex_interp(x0, x3, phi0, phi3, xi[n++]);
    if (n == 1) {
        float d[3];
        sub3(xi[0], x0, d);
        return norm3(d);
    }
    if (n == 2) return line_distance(x0, xi[0], xi[1]);
    if (n == 3) return plane_distance(x0, xi[0], xi[1], xi[2]);
    return FLT_MAX;
}

/* level_set.h:251-264 (calc_sdf, 7 points: centre, +x, +y, -x, -y, +z, -z) */
static float calc_sdf7(const float x[7][3], const float phi[7]) {
    float l = FLT_MAX;
    static const int c[5] = {1, 2, 3, 4, 1};
    for (int i = 0; i < 4; ++i)
        for (int j = 5; j <= 6; ++j) {
            const float t = tet_distance(x[0], x[c[i]], x[c[i + 1]], x[j], phi[0], phi[c[i]], phi[c[i + 1]], phi[j]);
            l = (t < l) ? t : l; /* std::min(l, t) */
        }
    return phi[0] < 0.0f ? -l : l;
}

static int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

#define RL_FIX 0
#define RL_NEAR 1
#define RL_FAR 2
#define RL_UNKNOWN 3

/* reinitialization.cpp:19-156 (reinitialize_fast_marching). grid is the
 * level set (in place), valid its 0/1 flags. The reference loops while any
 * FAR cell remains; a FAR cell with no path to the interface would keep it
 * looping forever, so this restatement also stops when a pass changes no
 * label, and returns the number of FAR cells left (0 whenever the reference
 * terminates). *passes receives the number of marching passes that changed
 * labels. */
long orc_reinitialize(int ni, int nj, int nk, float dx, float* grid, const uint8_t* valid, int* passes) {
    const size_t n = (size_t)ni * nj * nk;
    float* temp = (float*)malloc(n * sizeof(float));
    uint8_t* label = (uint8_t*)malloc(n);
    memcpy(temp, grid, n * sizeof(float));
#define T(i, j, k) temp[IDX(clampi(i, ni - 1), clampi(j, nj - 1), clampi(k, nk - 1), ni, nj)]
#define LB(i, j, k) label[IDX(clampi(i, ni - 1), clampi(j, nj - 1), clampi(k, nk - 1), ni, nj)]
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const size_t c = IDX(i, j, k, ni, nj);
                if (valid[c] == 0) {
                    label[c] = RL_UNKNOWN;
                    continue;
                }
                const float p = temp[c];
                const int flag = p * T(i - 1, j, k) <= 0.0f || p * T(i + 1, j, k) <= 0.0f ||
                                 p * T(i, j - 1, k) <= 0.0f || p * T(i, j + 1, k) <= 0.0f ||
                                 p * T(i, j, k - 1) <= 0.0f || p * T(i, j, k + 1) <= 0.0f;
                if (flag) {
                    if (grid[c] != 0.0f) {
                        /* Vector3f((i + 0.5) * dx, ...): double products rounded to float */
                        const double d = (double)dx;
                        const float xc = (float)((i + 0.5) * d), yc = (float)((j + 0.5) * d),
                                    zc = (float)((k + 0.5) * d);
                        const float xp = (float)(((i + 1) + 0.5) * d), xm = (float)(((i - 1) + 0.5) * d);
                        const float yp = (float)(((j + 1) + 0.5) * d), ym = (float)(((j - 1) + 0.5) * d);
                        const float zp = (float)(((k + 1) + 0.5) * d), zm = (float)(((k - 1) + 0.5) * d);
                        const float x[7][3] = {{xc, yc, zc}, {xp, yc, zc}, {xc, yp, zc}, {xm, yc, zc},
                                               {xc, ym, zc}, {xc, yc, zp}, {xc, yc, zm}};
                        const float phi[7] = {T(i, j, k),     T(i + 1, j, k), T(i, j + 1, k), T(i - 1, j, k),
                                              T(i, j - 1, k), T(i, j, k + 1), T(i, j, k - 1)};
                        grid[c] = calc_sdf7(x, phi);
                    }
                    label[c] = RL_FIX;
                } else {
                    label[c] = RL_FAR;
                }
            }
    int npass = 0;
    long far_left = 0;
    for (;;) {
        memcpy(temp, grid, n * sizeof(float));
        long far = 0, near = 0;
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < nj; ++j)
                for (int i = 0; i < ni; ++i) {
                    const size_t c = IDX(i, j, k, ni, nj);
                    if (label[c] != RL_FAR) continue;
                    ++far;
                    if (LB(i - 1, j, k) == RL_FIX || LB(i + 1, j, k) == RL_FIX || LB(i, j - 1, k) == RL_FIX ||
                        LB(i, j + 1, k) == RL_FIX || LB(i, j, k - 1) == RL_FIX || LB(i, j, k + 1) == RL_FIX) {
                        label[c] = RL_NEAR;
                        ++near;
                    }
                }
        if (far == 0) break;
        if (near == 0) { /* the reference would loop forever from here */
            far_left = far;
            break;
        }
        ++npass;
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < nj; ++j)
                for (int i = 0; i < ni; ++i) {
                    const size_t c = IDX(i, j, k, ni, nj);
                    if (label[c] != RL_NEAR) continue;
                    float p[3];
                    int m = 0;
                    if (LB(i - 1, j, k) == RL_FIX) p[m++] = T(i - 1, j, k);
                    else if (LB(i + 1, j, k) == RL_FIX) p[m++] = T(i + 1, j, k);
                    if (LB(i, j - 1, k) == RL_FIX) p[m++] = T(i, j - 1, k);
                    else if (LB(i, j + 1, k) == RL_FIX) p[m++] = T(i, j + 1, k);
                    if (LB(i, j, k - 1) == RL_FIX) p[m++] = T(i, j, k - 1);
                    else if (LB(i, j, k + 1) == RL_FIX) p[m++] = T(i, j, k + 1);
                    /* mode: 0 all >= 0, 1 all <= 0, else neither (asserts are off in the
                     * reference's release build, so that case proceeds like mode 0) */
                    int all_ge = 1, all_le = 1;
                    for (int q = 0; q < m; ++q) {
                        all_ge &= p[q] >= 0.0f;
                        all_le &= p[q] <= 0.0f;
                    }
                    const int mode = all_ge ? 0 : (all_le ? 1 : -1);
                    if (mode == 1)
                        for (int q = 0; q < m; ++q) p[q] = -p[q];
                    /* std::sort, ascending */
                    for (int a = 1; a < m; ++a)
                        for (int b = a; b > 0 && p[b] < p[b - 1]; --b) {
                            const float t = p[b];
                            p[b] = p[b - 1];
                            p[b - 1] = t;
                        }
                    float d = p[0] + dx;
                    if (m >= 2 && d > p[1])
                        d = (float)(0.5 * (double)(p[0] + p[1] + sqrtf(2 * dx * dx - (p[1] - p[0]) * (p[1] - p[0]))));
                    if (m == 3 && d > p[2]) {
                        const float s = p[0] + p[1] + p[2];
                        float q = s * s - 3.0f * (p[0] * p[0] + p[1] * p[1] + p[2] * p[2] - dx * dx);
                        q = (0.0f < q) ? q : 0.0f; /* std::max(0.0f, q) */
                        d = (float)(0.333333 * (double)(p[0] + p[1] + p[2] + sqrtf(q)));
                    }
                    if (mode == 1) d = -d;
                    grid[c] = d;
                }
        for (size_t c = 0; c < n; ++c)
            if (label[c] == RL_NEAR) label[c] = RL_FIX;
    }
#undef T
#undef LB
    free(temp);
    free(label);
    if (passes) *passes = npass;
    return far_left;
}

/* simulator.cpp:191-200 (calc_cell_centered_solid_phi) then :160-167
 * (find_valid_level_set). The eight nodes are summed in float, left to right;
 * the product with the double literal 0.125 is rounded to float when stored
 * in the float grid; a cell is invalid iff that value is < 0. */
void orc_find_valid_level_set(int ni, int nj, int nk, const float* nodal, uint8_t* valid) {
    const int mi = ni + 1, mj = nj + 1;
#define N(i, j, k) nodal[IDX(i, j, k, mi, mj)]
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const float s = N(i, j, k) + N(i + 1, j, k) + N(i, j + 1, k) + N(i + 1, j + 1, k) +
                                N(i, j, k + 1) + N(i + 1, j, k + 1) + N(i, j + 1, k + 1) +
                                N(i + 1, j + 1, k + 1);
                const float centred = (float)((double)s * 0.125);
                valid[IDX(i, j, k, ni, nj)] = (centred < 0.0f) ? 0 : 1;
            }
#undef N
}

/* extrapolation.cpp:5-46 (extrapolate). Each layer works from a snapshot of
 * the valid flags (old_valid) and of the values (grid; new values go to a
 * copy, grid_temp, that becomes grid at the end of the layer). An invalid
 * cell takes the float mean of its valid neighbours among the 26 around it,
 * summed in the order kk, jj, ii = -1..1 (k slowest), divided once by the
 * count; it then becomes valid. The loop returns early when a layer starts
 * with no invalid cell. valid holds 0/1. */
void orc_extrapolate(int ni, int nj, int nk, float* grid, uint8_t* valid, int layers) {
    const size_t n = (size_t)ni * nj * nk;
    float* temp = (float*)malloc(n * sizeof(float));
    uint8_t* old = (uint8_t*)malloc(n);
    memcpy(temp, grid, n * sizeof(float));
    for (int l = 0; l < layers; ++l) {
        int should_return = 1;
        memcpy(old, valid, n);
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < nj; ++j)
                for (int i = 0; i < ni; ++i) {
                    const size_t c = IDX(i, j, k, ni, nj);
                    if (old[c]) continue;
                    should_return = 0;
                    float sum = 0.0f;
                    int count = 0;
                    for (int kk = -1; kk <= 1; ++kk)
                        for (int jj = -1; jj <= 1; ++jj)
                            for (int ii = -1; ii <= 1; ++ii) {
                                const int x = i + ii, y = j + jj, z = k + kk;
                                if (x < 0 || x >= ni || y < 0 || y >= nj || z < 0 || z >= nk) continue;
                                const size_t q = IDX(x, y, z, ni, nj);
                                if (old[q]) {
                                    sum += grid[q];
                                    ++count;
                                }
                            }
                    if (count > 0) {
                        temp[c] = sum / (float)count;
                        valid[c] = 1;
                    }
                }
        if (should_return) break;
        memcpy(grid, temp, n * sizeof(float));
    }
    free(temp);
    free(old);
}

/* extrapolation.cpp:48-51 (extrapolate_const): the caller's flags stay as
 * they were. */
void orc_extrapolate_const(int ni, int nj, int nk, float* grid, const uint8_t* valid, int layers) {
    const size_t n = (size_t)ni * nj * nk;
    uint8_t* copy = (uint8_t*)malloc(n);
    memcpy(copy, valid, n);
    orc_extrapolate(ni, nj, nk, grid, copy, layers);
    free(copy);
}

/* projection.cpp:8-12 (callers guarantee in-bounds arguments) */
static int is_pressure_valid(int ni, int nj, const float* uw, const float* vw, const float* ww,
                             int i, int j, int k) {
    return uw[IDX(i, j, k, ni + 1, nj)] > 0.0f || uw[IDX(i + 1, j, k, ni + 1, nj)] > 0.0f ||
           vw[IDX(i, j, k, ni, nj + 1)] > 0.0f || vw[IDX(i, j + 1, k, ni, nj + 1)] > 0.0f ||
           ww[IDX(i, j, k, ni, nj)] > 0.0f || ww[IDX(i, j, k + 1, ni, nj)] > 0.0f;
}

/* projection.cpp:14-28, serial k->j->i loop of array.h:424-434 */
int orc_assign_indices(int ni, int nj, int nk, const float* phi, const float* uw, const float* vw,
                       const float* ww, int32_t* indices) {
    int n = 0, boundary = 0;
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                size_t c = IDX(i, j, k, ni, nj);
                if (is_pressure_valid(ni, nj, uw, vw, ww, i, j, k) && phi[c] < 0.0f) {
                    indices[c] = n++;
                    if (i == 0 || j == 0 || k == 0 || i == ni - 1 || j == nj - 1 || k == nk - 1)
                        boundary = 1;
                } else {
                    indices[c] = ORC_UNKNOWN_INDEX;
                }
            }
    if (boundary) return -1;
    return n;
}

/* projection.cpp:30-83. Per unknown, neighbours in the order -x,+x,-y,+y,-z,+z
 * (:71-76); diag is a double sum of float terms, b accumulates float products
 * in double, then b *= (double)(dx/dt) with dx/dt a float division (:82). */
void orc_build_matrix(int ni, int nj, int nk, float dx, float dt, const int32_t* indices,
                      const float* phi, const float* u, const float* v, const float* w,
                      const float* uw, const float* vw, const float* ww, orc_system* sys) {
    (void)nk;
    const float scale = dx / dt;
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const size_t c = IDX(i, j, k, ni, nj);
                const int index = indices[c];
                if (index == ORC_UNKNOWN_INDEX) continue;
                double bval = 0.0, diag = 0.0;
                const int nx[6] = {i - 1, i + 1, i, i, i, i};
                const int ny[6] = {j, j, j - 1, j + 1, j, j};
                const int nz[6] = {k, k, k, k, k - 1, k + 1};
                const float vel[6] = {u[IDX(i, j, k, ni + 1, nj)], u[IDX(i + 1, j, k, ni + 1, nj)],
                                      v[IDX(i, j, k, ni, nj + 1)], v[IDX(i, j + 1, k, ni, nj + 1)],
                                      w[IDX(i, j, k, ni, nj)],     w[IDX(i, j, k + 1, ni, nj)]};
                const float wgt[6] = {uw[IDX(i, j, k, ni + 1, nj)], uw[IDX(i + 1, j, k, ni + 1, nj)],
                                      vw[IDX(i, j, k, ni, nj + 1)], vw[IDX(i, j + 1, k, ni, nj + 1)],
                                      ww[IDX(i, j, k, ni, nj)],     ww[IDX(i, j, k + 1, ni, nj)]};
                double off[6] = {0, 0, 0, 0, 0, 0};
                int nidx[6] = {-1, -1, -1, -1, -1, -1};
                for (int d = 0; d < 6; ++d) {
                    const int x = nx[d], y = ny[d], z = nz[d];
                    if (is_pressure_valid(ni, nj, uw, vw, ww, x, y, z)) {
                        const size_t cn = IDX(x, y, z, ni, nj);
                        if (phi[cn] < 0.0f) {
                            diag += (double)wgt[d];
                            off[d] = (double)(-wgt[d]);
                            nidx[d] = indices[cn];
                        } else {
                            float theta = orc_fraction_inside2(phi[c], phi[cn]);
                            theta = max_f(theta, 0.01f);
                            diag += (double)(wgt[d] / theta);
                        }
                    }
                    if (d & 1)
                        bval -= (double)(wgt[d] * vel[d]);
                    else
                        bval += (double)(wgt[d] * vel[d]);
                }
                sys->diag[index] = diag;
                sys->b[index] = bval * (double)scale;
                sys->ax[index] = off[1];
                sys->ay[index] = off[3];
                sys->az[index] = off[5];
                sys->xm[index] = nidx[0];
                sys->xp[index] = nidx[1];
                sys->ym[index] = nidx[2];
                sys->yp[index] = nidx[3];
                sys->zm[index] = nidx[4];
                sys->zp[index] = nidx[5];
            }
}

int orc_to_csr(const orc_system* s, int32_t* row_ptr, int32_t* col, double* val) {
    int nnz = 0;
    for (int r = 0; r < s->n; ++r) {
        row_ptr[r] = nnz;
        /* lexicographic numbering: -z < -y < -x < self < +x < +y < +z */
        if (s->zm[r] >= 0) { col[nnz] = s->zm[r]; val[nnz++] = s->az[s->zm[r]]; }
        if (s->ym[r] >= 0) { col[nnz] = s->ym[r]; val[nnz++] = s->ay[s->ym[r]]; }
        if (s->xm[r] >= 0) { col[nnz] = s->xm[r]; val[nnz++] = s->ax[s->xm[r]]; }
        col[nnz] = r; val[nnz++] = s->diag[r];
        if (s->xp[r] >= 0) { col[nnz] = s->xp[r]; val[nnz++] = s->ax[r]; }
        if (s->yp[r] >= 0) { col[nnz] = s->yp[r]; val[nnz++] = s->ay[r]; }
        if (s->zp[r] >= 0) { col[nnz] = s->zp[r]; val[nnz++] = s->az[r]; }
    }
    row_ptr[s->n] = nnz;
    return nnz;
}

/* Natural-order incomplete Cholesky of the 7-point matrix, M = (E+L)E^{-1}(E+L^T):
 *   t = d - a_xm^2/e_xm - a_ym^2/e_ym - a_zm^2/e_zm
 *         - tau*(a_xm (a_y(xm)+a_z(xm))/e_xm + a_ym (a_x(ym)+a_z(ym))/e_ym
 *                + a_zm (a_x(zm)+a_y(zm))/e_zm)
 * with 1/e stored as inv_e, terms subtracted in the order written. */
int orc_ic_factor(const orc_system* s, double tau, double sigma, double* inv_e) {
    int fixed = 0;
    for (int n = 0; n < s->n; ++n) {
        const double d = s->diag[n];
        double t = d;
        const int xm = s->xm[n], ym = s->ym[n], zm = s->zm[n];
        if (xm >= 0) t -= s->ax[xm] * s->ax[xm] * inv_e[xm];
        if (ym >= 0) t -= s->ay[ym] * s->ay[ym] * inv_e[ym];
        if (zm >= 0) t -= s->az[zm] * s->az[zm] * inv_e[zm];
        if (tau != 0.0) {
            double m = 0.0;
            if (xm >= 0) m += s->ax[xm] * (s->ay[xm] + s->az[xm]) * inv_e[xm];
            if (ym >= 0) m += s->ay[ym] * (s->ax[ym] + s->az[ym]) * inv_e[ym];
            if (zm >= 0) m += s->az[zm] * (s->ax[zm] + s->ay[zm]) * inv_e[zm];
            t -= tau * m;
        }
        if (!(t > sigma * d)) {
            t = d;
            ++fixed;
        }
        inv_e[n] = 1.0 / t;
    }
    return fixed;
}

void orc_ic_apply(const orc_system* s, const double* inv_e, const double* r, double* z) {
    for (int n = 0; n < s->n; ++n) {
        double acc = r[n];
        if (s->xm[n] >= 0) acc -= s->ax[s->xm[n]] * z[s->xm[n]];
        if (s->ym[n] >= 0) acc -= s->ay[s->ym[n]] * z[s->ym[n]];
        if (s->zm[n] >= 0) acc -= s->az[s->zm[n]] * z[s->zm[n]];
        z[n] = acc * inv_e[n];
    }
    for (int n = s->n - 1; n >= 0; --n) {
        double acc = 0.0;
        if (s->xp[n] >= 0) acc += s->ax[n] * z[s->xp[n]];
        if (s->yp[n] >= 0) acc += s->ay[n] * z[s->yp[n]];
        if (s->zp[n] >= 0) acc += s->az[n] * z[s->zp[n]];
        z[n] = z[n] - inv_e[n] * acc;
    }
}

/* OpenMP threads of the loops the reference runs in parallel: the pressure
 * scatter and the velocity update (projection.cpp:85-134 through
 * parallel_loop, array.h:452-472, collapse(3)). Assembly, the factorisation
 * and the PCG stay serial as in the reference (serial_loop; Eigen's CG with a
 * Lower self-adjoint view is single-threaded) unless orc_set_krylov_parallel
 * turns on the OpenMP companion of the SpMV, axpys and dots (a timing
 * companion, not the reference's threading; the IC sweeps stay serial). */
static int g_krylov_parallel = 0;

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads(void) { return omp_get_max_threads(); }
void orc_set_krylov_parallel(int on) { g_krylov_parallel = on; }

static int g_progress = 0;
void orc_set_progress(int every) { g_progress = every > 0 ? every : 0; }

void orc_spmv(const orc_system* s, const double* x, double* y) {
#pragma omp parallel for if (g_krylov_parallel) schedule(static)
    for (int n = 0; n < s->n; ++n) {
        double acc = s->diag[n] * x[n];
        if (s->xm[n] >= 0) acc += s->ax[s->xm[n]] * x[s->xm[n]];
        if (s->xp[n] >= 0) acc += s->ax[n] * x[s->xp[n]];
        if (s->ym[n] >= 0) acc += s->ay[s->ym[n]] * x[s->ym[n]];
        if (s->yp[n] >= 0) acc += s->ay[n] * x[s->yp[n]];
        if (s->zm[n] >= 0) acc += s->az[s->zm[n]] * x[s->zm[n]];
        if (s->zp[n] >= 0) acc += s->az[n] * x[s->zp[n]];
        y[n] = acc;
    }
}

static double dot(int n, const double* a, const double* b) {
    double s = 0.0;
    if (g_krylov_parallel) {
#pragma omp parallel for reduction(+ : s) schedule(static)
        for (int i = 0; i < n; ++i) s += a[i] * b[i];
        return s;
    }
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* Eigen::ConjugateGradient (simulator.cpp:102-109). Eigen's
 * internal::conjugate_gradient loop, restated from its published source
 * (Eigen 3.4, IterativeLinearSolvers/ConjugateGradient.h; Eigen is a system
 * dependency absent from the reference tree):
 *   x = 0 (solve() without guess); r = b - A x; if |b|^2 == 0: x = 0, done;
 *   threshold = max(tol^2 |b|^2, DBL_MIN); if |r|^2 < threshold: done;
 *   p = M^{-1} r; absNew = r.p; i = 0;
 *   while i < maxIters:
 *     q = A p; alpha = absNew / p.q; x += alpha p; r -= alpha q;
 *     if |r|^2 < threshold: break
 *     z = M^{-1} r; absOld = absNew; absNew = r.z; p = z + (absNew/absOld) p;
 *     ++i
 *   iterations = i; error = sqrt(|r|^2 / |b|^2). */
int orc_pcg(const orc_system* s, const double* inv_e, double tol, int max_iter, double* x,
            int* iters, double* err) {
    const int n = s->n;
    double* r = (double*)malloc(sizeof(double) * (size_t)n);
    double* p = (double*)malloc(sizeof(double) * (size_t)n);
    double* z = (double*)malloc(sizeof(double) * (size_t)n);
    double* q = (double*)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        x[i] = 0.0;
        r[i] = s->b[i];
    }
    const double rhs2 = dot(n, s->b, s->b);
    int it = 0;
    double res2 = 0.0;
    if (rhs2 == 0.0) {
        *iters = 0;
        *err = 0.0;
        goto done;
    }
    {
        double threshold = tol * tol * rhs2;
        if (threshold < DBL_MIN) threshold = DBL_MIN;
        res2 = dot(n, r, r);
        if (res2 < threshold) {
            *iters = 0;
            *err = sqrt(res2 / rhs2);
            goto done;
        }
        orc_ic_apply(s, inv_e, r, p);
        double abs_new = dot(n, r, p);
        while (it < max_iter) {
            orc_spmv(s, p, q);
            const double alpha = abs_new / dot(n, p, q);
#pragma omp parallel for if (g_krylov_parallel) schedule(static)
            for (int i = 0; i < n; ++i) {
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
            }
            res2 = dot(n, r, r);
            if (res2 < threshold) break;
            orc_ic_apply(s, inv_e, r, z);
            const double abs_old = abs_new;
            abs_new = dot(n, r, z);
            const double beta = abs_new / abs_old;
#pragma omp parallel for if (g_krylov_parallel) schedule(static)
            for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
            ++it;
            if (g_progress > 0 && it % g_progress == 0) {
                fprintf(stderr, "# oracle pcg: iteration %d, |r|^2/|b|^2 %.3e\n", it, res2 / rhs2);
                fflush(stderr);
            }
        }
        *iters = it;
        *err = sqrt(res2 / rhs2);
    }
done:
    free(r);
    free(p);
    free(z);
    free(q);
    return 0;
}

/* projection.cpp:85-91 */
void orc_build_grid_from_vector(int ni, int nj, int nk, const double* x, const int32_t* indices,
                                float background, float* pressure) {
#pragma omp parallel for collapse(3) schedule(static)
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const size_t c = IDX(i, j, k, ni, nj);
                pressure[c] = indices[c] == ORC_UNKNOWN_INDEX ? background : (float)x[indices[c]];
            }
}

/* projection.cpp:93-134. Ranges {1, n_face - 1} per axis of each face grid. */
void orc_update_velocity(int ni, int nj, int nk, float dx, float dt, const float* pressure,
                         const float* phi, float* u, float* v, float* w) {
#define P(i, j, k) phi[IDX(i, j, k, ni, nj)]
#define PR(i, j, k) pressure[IDX(i, j, k, ni, nj)]
#pragma omp parallel for collapse(3) schedule(static)
    for (int k = 1; k < nk - 1; ++k)
        for (int j = 1; j < nj - 1; ++j)
            for (int i = 1; i < ni; ++i)
                if (P(i, j, k) < 0.0f || P(i - 1, j, k) < 0.0f) {
                    float theta = 1.0f;
                    if (P(i, j, k) >= 0.0f || P(i - 1, j, k) >= 0.0f)
                        theta = orc_fraction_inside2(P(i - 1, j, k), P(i, j, k));
                    theta = max_f(theta, 0.01f);
                    float* f = &u[IDX(i, j, k, ni + 1, nj)];
                    *f = *f - dt * (PR(i, j, k) - PR(i - 1, j, k)) / (dx * theta);
                }
#pragma omp parallel for collapse(3) schedule(static)
    for (int k = 1; k < nk - 1; ++k)
        for (int j = 1; j < nj; ++j)
            for (int i = 1; i < ni - 1; ++i)
                if (P(i, j, k) < 0.0f || P(i, j - 1, k) < 0.0f) {
                    float theta = 1.0f;
                    if (P(i, j, k) >= 0.0f || P(i, j - 1, k) >= 0.0f)
                        theta = orc_fraction_inside2(P(i, j - 1, k), P(i, j, k));
                    theta = max_f(theta, 0.01f);
                    float* f = &v[IDX(i, j, k, ni, nj + 1)];
                    *f = *f - dt * (PR(i, j, k) - PR(i, j - 1, k)) / (dx * theta);
                }
#pragma omp parallel for collapse(3) schedule(static)
    for (int k = 1; k < nk; ++k)
        for (int j = 1; j < nj - 1; ++j)
            for (int i = 1; i < ni - 1; ++i)
                if (P(i, j, k) < 0.0f || P(i, j, k - 1) < 0.0f) {
                    float theta = 1.0f;
                    if (P(i, j, k) >= 0.0f || P(i, j, k - 1) >= 0.0f)
                        theta = orc_fraction_inside2(P(i, j, k - 1), P(i, j, k));
                    theta = max_f(theta, 0.01f);
                    float* f = &w[IDX(i, j, k, ni, nj)];
                    *f = *f - dt * (PR(i, j, k) - PR(i, j, k - 1)) / (dx * theta);
                }
#undef P
#undef PR
}

/* simulator.cpp:96-112 */
int orc_solve_pressure(int ni, int nj, int nk, float dx, float dt, const float* phi, float* u,
                       float* v, float* w, const float* uw, const float* vw, const float* ww,
                       double tol, int max_iter, double tau, double sigma, int32_t* indices,
                       float* pressure, int* iters, double* err, double* true_rel_res) {
    const int n = orc_assign_indices(ni, nj, nk, phi, uw, vw, ww, indices);
    if (n <= 0) return n;
    orc_system s;
    s.n = n;
    double* dbl = (double*)malloc(sizeof(double) * (size_t)n * 8);
    int32_t* ints = (int32_t*)malloc(sizeof(int32_t) * (size_t)n * 6);
    s.diag = dbl;
    s.ax = dbl + (size_t)n;
    s.ay = dbl + 2 * (size_t)n;
    s.az = dbl + 3 * (size_t)n;
    s.b = dbl + 4 * (size_t)n;
    double* inv_e = dbl + 5 * (size_t)n;
    double* x = dbl + 6 * (size_t)n;
    double* q = dbl + 7 * (size_t)n;
    s.xp = ints;
    s.yp = ints + (size_t)n;
    s.zp = ints + 2 * (size_t)n;
    s.xm = ints + 3 * (size_t)n;
    s.ym = ints + 4 * (size_t)n;
    s.zm = ints + 5 * (size_t)n;
    orc_build_matrix(ni, nj, nk, dx, dt, indices, phi, u, v, w, uw, vw, ww, &s);
    orc_ic_factor(&s, tau, sigma, inv_e);
    orc_pcg(&s, inv_e, tol, max_iter, x, iters, err);
    orc_spmv(&s, x, q);
    double rr = 0.0, bb = 0.0;
    for (int i = 0; i < n; ++i) {
        const double d = s.b[i] - q[i];
        rr += d * d;
        bb += s.b[i] * s.b[i];
    }
    *true_rel_res = bb > 0.0 ? sqrt(rr / bb) : 0.0;
    orc_build_grid_from_vector(ni, nj, nk, x, indices, 0.0f, pressure);
    orc_update_velocity(ni, nj, nk, dx, dt, pressure, phi, u, v, w);
    free(dbl);
    free(ints);
    return n;
}

uint64_t orc_splitmix64(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double orc_uniform01(uint64_t* state) { return (double)(orc_splitmix64(state) >> 11) * 0x1.0p-53; }
