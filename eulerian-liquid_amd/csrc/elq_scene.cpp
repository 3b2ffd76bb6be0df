// Host-side input generation for the synthetic workloads: the reference
// simulator's own set-up steps that run before solve_pressure. These are the
// callers of the projection, not the projection itself; they run on the CPU
// exactly as the reference runs them.
//   create_grid / init_scene        simulator.cpp:52-76
//   sphere, pool, room emitters     levelset_emitter.cpp:10-32, list apply :74-79
//   cell centres                    grid.h:355-367
//   compute_staggered_grid_weight   boundary.cpp:6-34 with level_set.h:68-163
//   add_gravity                     simulator.cpp:90-93, constants.h:5-6
//   find_valid_velocity             simulator.cpp:125-158
//   set_invalid_velocity_zero       simulator.cpp:169-178
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <limits>
#include <thread>
#include <vector>

#include "../../include/elq.h"

namespace {

// Runs f(k) for k in [0, nk) on up to 16 host threads, contiguous k ranges
// per thread; each cell's value depends only on its own inputs, so the
// result does not depend on the split.
template <class F>
void parallel_k(int nk, F f) {
    const int hw = (int)std::thread::hardware_concurrency();
    const int nt = std::max(1, std::min({hw > 0 ? hw : 1, 16, nk / 8 > 0 ? nk / 8 : 1}));
    if (nt == 1) {
        for (int k = 0; k < nk; ++k) f(k);
        return;
    }
    std::vector<std::thread> th(nt);
    for (int t = 0; t < nt; ++t) {
        auto part = [=] {
            for (int k = (int)((int64_t)nk * t / nt); k < (int)((int64_t)nk * (t + 1) / nt); ++k) f(k);
        };
        try {
            th[t] = std::thread(part);
        } catch (...) {  // no thread to be had: this range on the calling thread
            part();
        }
    }
    for (auto& x : th)
        if (x.joinable()) x.join();
}


inline size_t at(int i, int j, int k, int ni, int nj) {
    return (size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * (size_t)k);
}

float inside2(float a, float b) {
    if (a < 0.0f && b < 0.0f) return 1.0f;
    if (a < 0.0f && b >= 0.0f) return a / (a - b);
    if (a >= 0.0f && b < 0.0f) return b / (b - a);
    return 0.0f;
}

// level_set.h:83-163 (marching-squares area of the negative region; the
// reference's 0.5/0.25 literals are double, so those products round once to
// float at assignment/return).
float inside4(float bl, float br, float tl, float tr) {
    const int count = int(bl < 0.0f) + int(tl < 0.0f) + int(br < 0.0f) + int(tr < 0.0f);
    float c[4] = {bl, br, tr, tl};
    auto rotate = [&c]() { std::rotate(c, c + 1, c + 4); };
    switch (count) {
        case 4:
            return 1.0f;
        case 3: {
            while (c[0] < 0.0f) rotate();
            const float s0 = 1.0f - inside2(c[0], c[3]);
            const float s1 = 1.0f - inside2(c[0], c[1]);
            return static_cast<float>(1.0 - 0.5 * s0 * s1);
        }
        case 2: {
            while (c[0] >= 0.0f || !(c[1] < 0.0f || c[2] < 0.0f)) rotate();
            if (c[1] < 0.0f) {
                const float left = inside2(c[0], c[3]);
                const float right = inside2(c[1], c[2]);
                return static_cast<float>(0.5 * (left + right));
            }
            const float mid = static_cast<float>(0.25 * (c[0] + c[1] + c[2] + c[3]));
            float area = 0.0f;
            if (mid < 0.0f) {
                const float s1 = 1.0f - inside2(c[0], c[3]);
                const float s3 = 1.0f - inside2(c[2], c[3]);
                area = static_cast<float>(area + 0.5 * s1 * s3);
                const float s2 = 1.0f - inside2(c[2], c[1]);
                const float s0 = 1.0f - inside2(c[0], c[1]);
                area = static_cast<float>(area + 0.5 * s0 * s2);
                return 1.0f - area;
            }
            const float s0 = inside2(c[0], c[1]);
            const float s1 = inside2(c[0], c[3]);
            area = static_cast<float>(area + 0.5 * s0 * s1);
            const float s2 = inside2(c[2], c[1]);
            const float s3 = inside2(c[2], c[3]);
            area = static_cast<float>(area + 0.5 * s2 * s3);
            return area;
        }
        case 1: {
            while (c[0] >= 0.0f) rotate();
            const float s0 = inside2(c[0], c[3]);
            const float s1 = inside2(c[0], c[1]);
            return static_cast<float>(0.5 * s0 * s1);
        }
        default:
            return 0.0f;
    }
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double unit(uint64_t& s) { return static_cast<double>(splitmix(s) >> 11) * 0x1.0p-53; }

float draw(uint64_t& s, double lo, double hi) { return static_cast<float>(lo + (hi - lo) * unit(s)); }

}  // namespace

extern "C" {

double elq_uniform01(uint64_t* state) { return state ? unit(*state) : 0.0; }

int elq_scene_level_set(int32_t ni, int32_t nj, int32_t nk, float dx, float slab_height, int32_t n_spheres,
                        const float* spheres, float* level_set) try {
    if (ni < 1 || nj < 1 || nk < 1 || !level_set || (n_spheres > 0 && !spheres)) return ELQ_ERR_INVALID_ARG;
    parallel_k(nk, [&](int k) {
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                // Grid3::pos: (i + 0.5f, j + 0.5f, k + 0.5f) * dx + origin (0)
                const float pos[3] = {(i + 0.5f) * dx + 0.0f, (j + 0.5f) * dx + 0.0f, (k + 0.5f) * dx + 0.0f};
                float phi = std::numeric_limits<float>::max();
                for (int s = 0; s < n_spheres; ++s) {
                    const float* sp = spheres + 4 * s;
                    const float d0 = pos[0] - sp[0], d1 = pos[1] - sp[1], d2 = pos[2] - sp[2];
                    const float norm = std::sqrt(d0 * d0 + (d1 * d1 + d2 * d2));
                    phi = std::min(phi, norm - sp[3]);
                }
                if (slab_height > -1e30f) phi = std::min(phi, pos[1] - slab_height);
                level_set[at(i, j, k, ni, nj)] = phi;
            }
    });
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_scene_room_nodal(int32_t ni, int32_t nj, int32_t nk, float* nodal) try {
    if (ni < 1 || nj < 1 || nk < 1 || !nodal) return ELQ_ERR_INVALID_ARG;
    // room() on the nodal grid of (ni+1)(nj+1)(nk+1) nodes, origin -dx/2
    const int a = ni + 1, b = nj + 1, c = nk + 1;
    for (int k = 0; k < c; ++k)
        for (int j = 0; j < b; ++j)
            for (int i = 0; i < a; ++i) {
                float v;
                if (i == 0 || i == a - 1 || j == 0 || j == b - 1 || k == 0 || k == c - 1)
                    v = -1.0f;
                else if (i == 1 || i == a - 2 || j == 1 || j == b - 2 || k == 1 || k == c - 2)
                    v = static_cast<float>(-0.01);
                else
                    v = 1.0f;
                nodal[at(i, j, k, a, b)] = std::min(std::numeric_limits<float>::max(), v);
            }
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_scene_weights(int32_t ni, int32_t nj, int32_t nk, float* uw, float* vw, float* ww) try {
    if (ni < 1 || nj < 1 || nk < 1 || !uw || !vw || !ww) return ELQ_ERR_INVALID_ARG;
    const int a = ni + 1, b = nj + 1, c = nk + 1;
    std::vector<float> nodal((size_t)a * b * c);
    elq_scene_room_nodal(ni, nj, nk, nodal.data());
    auto N = [&](int i, int j, int k) { return nodal[at(i, j, k, a, b)]; };
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i <= ni; ++i)
                uw[at(i, j, k, ni + 1, nj)] =
                    std::clamp(1.0f - inside4(N(i, j, k), N(i, j + 1, k), N(i, j, k + 1), N(i, j + 1, k + 1)), 0.0f,
                               1.0f);
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j <= nj; ++j)
            for (int i = 0; i < ni; ++i)
                vw[at(i, j, k, ni, nj + 1)] =
                    std::clamp(1.0f - inside4(N(i, j, k), N(i, j, k + 1), N(i + 1, j, k), N(i + 1, j, k + 1)), 0.0f,
                               1.0f);
    for (int k = 0; k <= nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i)
                ww[at(i, j, k, ni, nj)] =
                    std::clamp(1.0f - inside4(N(i, j, k), N(i, j + 1, k), N(i + 1, j, k), N(i + 1, j + 1, k)), 0.0f,
                               1.0f);
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_scene_add_gravity(int32_t ni, int32_t nj, int32_t nk, float* v, float dt) try {
    if (ni < 1 || nj < 1 || nk < 1 || !v) return ELQ_ERR_INVALID_ARG;
    const float g = 9.80665f * dt;  // constants::gravity<float> * dt
    const size_t n = (size_t)ni * (nj + 1) * nk;
    for (size_t c = 0; c < n; ++c) v[c] -= g;
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_scene_valid_and_zero(const elq_grid* g) try {
    if (!g || !g->level_set || !g->u || !g->v || !g->w || !g->u_weight || !g->v_weight || !g->w_weight)
        return ELQ_ERR_INVALID_ARG;
    const int ni = g->ni, nj = g->nj, nk = g->nk;
    const size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
    std::vector<char> uv(nu), vv(nv), wv(nw);
    for (size_t c = 0; c < nu; ++c) uv[c] = g->u_weight[c] == 0.0f ? 0 : 1;
    for (size_t c = 0; c < nv; ++c) vv[c] = g->v_weight[c] == 0.0f ? 0 : 1;
    for (size_t c = 0; c < nw; ++c) wv[c] = g->w_weight[c] == 0.0f ? 0 : 1;
    auto phi = [&](int i, int j, int k) { return g->level_set[at(i, j, k, ni, nj)]; };
    for (int k = 1; k < nk - 1; ++k)
        for (int j = 1; j < nj - 1; ++j)
            for (int i = 1; i < ni; ++i)
                if (phi(i - 1, j, k) >= 0.0f && phi(i, j, k) >= 0.0f) uv[at(i, j, k, ni + 1, nj)] = 0;
    for (int k = 1; k < nk - 1; ++k)
        for (int j = 1; j < nj; ++j)
            for (int i = 1; i < ni - 1; ++i)
                if (phi(i, j - 1, k) >= 0.0f && phi(i, j, k) >= 0.0f) vv[at(i, j, k, ni, nj + 1)] = 0;
    for (int k = 1; k < nk; ++k)
        for (int j = 1; j < nj - 1; ++j)
            for (int i = 1; i < ni - 1; ++i)
                if (phi(i, j, k - 1) >= 0.0f && phi(i, j, k) >= 0.0f) wv[at(i, j, k, ni, nj)] = 0;
    for (size_t c = 0; c < nu; ++c)
        if (!uv[c]) g->u[c] = 0.0f;
    for (size_t c = 0; c < nv; ++c)
        if (!vv[c]) g->v[c] = 0.0f;
    for (size_t c = 0; c < nw; ++c)
        if (!wv[c]) g->w[c] = 0.0f;
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_scene_random_spheres(uint64_t* state, int32_t count, double c_lo, double c_hi, double r_lo, double r_hi,
                             double lx, double ly, double lz, float* spheres) try {
    if (!state || count < 0 || (count > 0 && !spheres)) return ELQ_ERR_INVALID_ARG;
    for (int s = 0; s < count; ++s) {
        spheres[4 * s + 0] = draw(*state, c_lo * lx, c_hi * lx);
        spheres[4 * s + 1] = draw(*state, c_lo * ly, c_hi * ly);
        spheres[4 * s + 2] = draw(*state, c_lo * lz, c_hi * lz);
        spheres[4 * s + 3] = draw(*state, r_lo, r_hi);
    }
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

int elq_scene_random_velocity(uint64_t* state, int32_t ni, int32_t nj, int32_t nk, double amp, float* u, float* v,
                              float* w) try {
    if (!state || !u || !v || !w || ni < 1 || nj < 1 || nk < 1) return ELQ_ERR_INVALID_ARG;
    const size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
    for (size_t c = 0; c < nu; ++c) u[c] = draw(*state, -amp, amp);
    for (size_t c = 0; c < nv; ++c) v[c] = draw(*state, -amp, amp);
    for (size_t c = 0; c < nw; ++c) w[c] = draw(*state, -amp, amp);
    return ELQ_OK;
} catch (...) {  // host resources exhausted: reported, never thrown past the C ABI
    return ELQ_ERR_DEVICE;
}

}  // extern "C"
