// C++ host side of the projection: the reference simulator's projection step
// over the C ABI of include/elq.h.
//
// Mirrors the interface a caller of the reference uses:
//   namespace nama::_3d::fluid       (include/nama/project/3d/fluid/projection.h:10-22,
//                                     extrapolation.h:7-8)
//     assign_indices, build_matrix, build_grid_from_vector,
//     update_velocity_using_pressure, extrapolate, extrapolate_const,
//     advect_staggered_grid, advect_collocated_grid (advection.h:8-11),
//     reinitialize_fast_marching (reinitialization.h)
//   class Simulator                  (include/nama/project/3d/eulerian_liquid/simulator.h:16-73)
//     create_grid, init_scene, compute_weights, add_gravity,
//     find_valid_velocity, set_invalid_velocity_zero, solve_pressure,
//     extrapolate_velocity, find_valid_level_set, extrapolate_level_set,
//     advect_velocity, advect_level_set, reinitialize
// Grids are flat std::vector<float>/<int> in the reference layout
// (i + ni*(j + nj*k)); staggered grids are three such vectors (u, v, w).
// Errors follow the reference: solve_pressure and assign_indices throw
// std::runtime_error("No fluid"); non-convergence is not an error
// (simulator.cpp:105-109 never checks it) but is reported in last_stats().
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/elq.h"

namespace elq_host {

struct Grid3f {
    int ni = 0, nj = 0, nk = 0;
    std::vector<float> data;
    void create(int a, int b, int c, float value = 0.0f) {
        ni = a;
        nj = b;
        nk = c;
        data.assign((size_t)a * b * c, value);
    }
    float& operator()(int i, int j, int k) { return data[(size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * k)]; }
    float operator()(int i, int j, int k) const {
        return data[(size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * k)];
    }
};

struct Grid3i {
    int ni = 0, nj = 0, nk = 0;
    std::vector<int32_t> data;
    void create(int a, int b, int c, int32_t value = 0) {
        ni = a;
        nj = b;
        nk = c;
        data.assign((size_t)a * b * c, value);
    }
};

// Array3c (array.h:350): one 0/1 byte per cell
struct Grid3c {
    int ni = 0, nj = 0, nk = 0;
    std::vector<uint8_t> data;
    void create(int a, int b, int c, uint8_t value = 0) {
        ni = a;
        nj = b;
        nk = c;
        data.assign((size_t)a * b * c, value);
    }
};

// staggered_grid.h:100-112: u (ni+1)nj nk, v ni(nj+1)nk, w ni nj(nk+1)
struct StaggeredGrid3f {
    std::array<Grid3f, 3> c;
    void create(int ni, int nj, int nk, float value = 0.0f) {
        c[0].create(ni + 1, nj, nk, value);
        c[1].create(ni, nj + 1, nk, value);
        c[2].create(ni, nj, nk + 1, value);
    }
    Grid3f& u() { return c[0]; }
    Grid3f& v() { return c[1]; }
    Grid3f& w() { return c[2]; }
    const Grid3f& u() const { return c[0]; }
    const Grid3f& v() const { return c[1]; }
    const Grid3f& w() const { return c[2]; }
};

struct StaggeredGrid3c {
    std::array<Grid3c, 3> c;
    void create(int ni, int nj, int nk, uint8_t value = 0) {
        c[0].create(ni + 1, nj, nk, value);
        c[1].create(ni, nj + 1, nk, value);
        c[2].create(ni, nj, nk + 1, value);
    }
};

// Owns one device context (device buffers persist across frames).
class Device {
   public:
    explicit Device(int device = 0, size_t hbm_budget = 0);
    ~Device();
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    elq_ctx* ctx() const { return m_ctx; }

   private:
    elq_ctx* m_ctx = nullptr;
};

namespace fluid {
// projection.cpp:14-28
int assign_indices(Device& dev, Grid3i& indices, const Grid3f& level_set, const StaggeredGrid3f& weight, float dx);
// projection.cpp:30-83; A in stencil form (A(n,n) = diag[n], A(n,+x) = ax[n], ...)
struct StencilMatrix {
    std::vector<double> diag, ax, ay, az;
};
void build_matrix(Device& dev, StencilMatrix& A, std::vector<double>& b, int mat_size, const Grid3f& level_set,
                  const StaggeredGrid3f& velocity, const StaggeredGrid3f& weight, float dx, float dt);
// CSR with Eigen::SparseMatrix<double, RowMajor> semantics (columns sorted,
// explicit zeros kept), built from the stencil form and the indices grid.
void to_csr(const StencilMatrix& A, const Grid3i& indices, std::vector<int32_t>& row_ptr, std::vector<int32_t>& col,
            std::vector<double>& val);
// projection.cpp:85-91
void build_grid_from_vector(Grid3f& grid, const std::vector<double>& vector, const Grid3i& indices,
                            float background_value);
// projection.cpp:93-134
void update_velocity_using_pressure(Device& dev, StaggeredGrid3f& velocity, const Grid3f& pressure,
                                    const Grid3f& level_set, const StaggeredGrid3f& weight, float dx, float dt);
// extrapolation.cpp:5-46 (grid and valid updated) and :48-51 (valid unchanged)
void extrapolate(Device& dev, Grid3f& grid, Grid3c& valid, int layers);
void extrapolate_const(Device& dev, Grid3f& grid, const Grid3c& valid, int layers);
// advection.cpp:9-27: grid = grid_temp carried by velocity (dx: cell size)
void advect_collocated_grid(Device& dev, Grid3f& grid, const Grid3f& grid_temp, const StaggeredGrid3f& velocity,
                            float dx, float dt);
void advect_staggered_grid(Device& dev, StaggeredGrid3f& grid, const StaggeredGrid3f& grid_temp,
                           const StaggeredGrid3f& velocity, float dx, float dt);
// reinitialization.cpp:19-156 (grid_temp is scratch in the reference; here
// the library keeps its own). Throws std::runtime_error when a valid cell
// has no path to the interface: the reference would not return.
void reinitialize_fast_marching(Device& dev, Grid3f& grid, const Grid3c& valid, float dx);
}  // namespace fluid

// Device buffer owned for the lifetime of a Simulator (no per-step allocation).
class DeviceArray {
   public:
    DeviceArray() = default;
    ~DeviceArray() { reset(); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    void allocate(elq_ctx* ctx, size_t bytes);
    void reset();
    template <class T>
    T* as() const { return static_cast<T*>(m_p); }
    size_t bytes() const { return m_bytes; }

   private:
    elq_ctx* m_ctx = nullptr;
    void* m_p = nullptr;
    size_t m_bytes = 0;
};

// The reference's Simulator (simulator.h:16-73) with its grids resident in
// device memory: create_grid allocates them once, every frame step runs on
// them through the *_device entry points, and nothing is allocated or copied
// between host and device per step. The public host members mirror the
// reference's; they are written by the set-up steps and refreshed from the
// device only when the caller asks (download_*), and upload() sends host
// edits back (e.g. a velocity set by the caller).
class Simulator {
   public:
    explicit Simulator(Device& dev) : m_dev(dev) { elq_default_opts(&m_opts); }

    // simulator.cpp:52-70 (nodal solid grid is generated inside compute_weights)
    void create_grid(float width, int ni, int nj, int nk);
    // simulator.cpp:72-76: liquid = FLT_MAX, then min with the emitters
    void init_scene(float pool_height, const std::vector<std::array<float, 4>>& spheres);
    // simulator.cpp:187-189 with the room() solid of levelset_emitter.cpp:18-32
    void compute_weights();
    void add_gravity(float dt);                    // simulator.cpp:90-93
    void find_valid_velocity_and_zero_invalid();   // simulator.cpp:125-158 + 169-178
    void find_valid_velocity();                    // simulator.cpp:125-158 (m_velocity_valid)
    void set_invalid_velocity_zero();              // simulator.cpp:169-178
    void solve_pressure(float dt);                 // simulator.cpp:96-112
    void extrapolate_velocity();                   // simulator.cpp:119-123 (10 layers)
    void find_valid_level_set();                   // simulator.cpp:160-167 with :191-200
    void extrapolate_level_set();                  // simulator.cpp:115-117 (10 layers)
    void advect_velocity(float dt);                // simulator.cpp:79-82
    void advect_level_set(float dt);               // simulator.cpp:84-87
    void reinitialize();                           // simulator.cpp:181-183

    elq_opts& opts() { return m_opts; }
    const elq_stats& last_stats() const { return m_stats; }

    // device -> host mirrors, only what the caller reads
    void download_level_set();
    void download_velocity();
    void download_pressure();  // m_pressure and m_indices of the last solve_pressure
    void download_valid();     // m_velocity_valid and m_level_set_valid
    // host -> device: the level set, velocity and weights as the host members hold them
    void upload();

    float m_dx = 0.0f;
    int m_ni = 0, m_nj = 0, m_nk = 0;
    StaggeredGrid3f m_velocity;
    StaggeredGrid3f m_velocity_weight;
    Grid3f m_level_set;
    Grid3i m_indices;
    Grid3f m_pressure;
    StaggeredGrid3c m_velocity_valid;
    Grid3c m_level_set_valid;

   private:
    elq_grid device_grid();
    Device& m_dev;
    elq_opts m_opts;
    elq_stats m_stats{};
    // device-resident state
    DeviceArray d_phi, d_u, d_v, d_w, d_uw, d_vw, d_ww, d_p, d_idx, d_fu, d_fv, d_fw, d_lsv, d_nodal;
};

}  // namespace elq_host
