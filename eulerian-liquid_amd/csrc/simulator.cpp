// C++ host side of the projection over the C ABI (see simulator.hpp).
#include "simulator.hpp"

#include <cfloat>
#include <string>

namespace elq_host {

namespace {

elq_grid make_grid(int ni, int nj, int nk, float dx, const Grid3f& level_set, StaggeredGrid3f& velocity,
                   const StaggeredGrid3f& weight) {
    elq_grid g{};
    g.ni = ni;
    g.nj = nj;
    g.nk = nk;
    g.dx = dx;
    g.level_set = level_set.data.data();
    g.u = velocity.u().data.data();
    g.v = velocity.v().data.data();
    g.w = velocity.w().data.data();
    g.u_weight = weight.u().data.data();
    g.v_weight = weight.v().data.data();
    g.w_weight = weight.w().data.data();
    return g;
}

[[noreturn]] void raise(elq_ctx* ctx, int rc) {
    if (rc == ELQ_ERR_NO_FLUID) throw std::runtime_error("No fluid");  // projection.cpp:25
    throw std::runtime_error(std::string("elq: ") + elq_ctx_last_error(ctx));
}

void check(elq_ctx* ctx, int rc) {
    if (rc != ELQ_OK) raise(ctx, rc);
}

template <class G>
size_t bytes_of(const G& g) {
    return g.data.size() * sizeof(g.data[0]);
}

// Device copy of one host array for the duration of a call.
class DevArray {
   public:
    DevArray(elq_ctx* ctx, size_t bytes) : m_ctx(ctx) {
        check(ctx, elq_device_alloc(ctx, bytes < 4 ? 4 : bytes, &m_p));
        m_bytes = bytes;
    }
    template <class T>
    DevArray(elq_ctx* ctx, const std::vector<T>& host) : DevArray(ctx, host.size() * sizeof(T)) {
        check(ctx, elq_memcpy_h2d(ctx, m_p, host.data(), m_bytes));
    }
    ~DevArray() { elq_device_free(m_ctx, m_p); }
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    template <class T>
    T* as() const { return static_cast<T*>(m_p); }
    template <class T>
    void fetch(std::vector<T>& host) const {
        check(m_ctx, elq_memcpy_d2h(m_ctx, host.data(), m_p, host.size() * sizeof(T)));
    }

   private:
    elq_ctx* m_ctx;
    void* m_p = nullptr;
    size_t m_bytes = 0;
};

}  // namespace

Device::Device(int device, size_t hbm_budget) {
    const int rc = elq_ctx_create(device, hbm_budget, &m_ctx);
    if (rc != ELQ_OK) throw std::runtime_error("elq: no usable gfx950 device");
}

Device::~Device() { elq_ctx_destroy(m_ctx); }

namespace fluid {

int assign_indices(Device& dev, Grid3i& indices, const Grid3f& level_set, const StaggeredGrid3f& weight, float dx) {
    StaggeredGrid3f zero;
    zero.create(level_set.ni, level_set.nj, level_set.nk);
    elq_grid g = make_grid(level_set.ni, level_set.nj, level_set.nk, dx, level_set, zero, weight);
    indices.create(level_set.ni, level_set.nj, level_set.nk);
    int64_t n = 0;
    const int rc = elq_assign_indices(dev.ctx(), &g, indices.data.data(), &n);
    if (rc != ELQ_OK) raise(dev.ctx(), rc);
    return (int)n;
}

void build_matrix(Device& dev, StencilMatrix& A, std::vector<double>& b, int mat_size, const Grid3f& level_set,
                  const StaggeredGrid3f& velocity, const StaggeredGrid3f& weight, float dx, float dt) {
    StaggeredGrid3f vel = velocity;
    elq_grid g = make_grid(level_set.ni, level_set.nj, level_set.nk, dx, level_set, vel, weight);
    A.diag.assign(mat_size, 0.0);
    A.ax.assign(mat_size, 0.0);
    A.ay.assign(mat_size, 0.0);
    A.az.assign(mat_size, 0.0);
    b.assign(mat_size, 0.0);
    const int rc = elq_build_matrix(dev.ctx(), &g, dt, A.diag.data(), A.ax.data(), A.ay.data(), A.az.data(), b.data());
    if (rc != ELQ_OK) raise(dev.ctx(), rc);
}

void to_csr(const StencilMatrix& A, const Grid3i& indices, std::vector<int32_t>& row_ptr, std::vector<int32_t>& col,
            std::vector<double>& val) {
    const int n = (int)A.diag.size();
    const int ni = indices.ni, nj = indices.nj, nk = indices.nk;
    std::vector<int32_t> nb(6 * (size_t)n, -1);  // -x,+x,-y,+y,-z,+z
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const int32_t c = indices.data[(size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * k)];
                if (c == ELQ_UNKNOWN_INDEX) continue;
                auto at = [&](int x, int y, int z) -> int32_t {
                    if (x < 0 || y < 0 || z < 0 || x >= ni || y >= nj || z >= nk) return -1;
                    const int32_t v = indices.data[(size_t)x + (size_t)ni * ((size_t)y + (size_t)nj * z)];
                    return v == ELQ_UNKNOWN_INDEX ? -1 : v;
                };
                const int32_t q[6] = {at(i - 1, j, k), at(i + 1, j, k), at(i, j - 1, k),
                                      at(i, j + 1, k), at(i, j, k - 1), at(i, j, k + 1)};
                for (int d = 0; d < 6; ++d) nb[6 * (size_t)c + d] = q[d];
            }
    row_ptr.assign(n + 1, 0);
    col.clear();
    val.clear();
    for (int r = 0; r < n; ++r) {
        row_ptr[r] = (int32_t)col.size();
        const int32_t* q = &nb[6 * (size_t)r];
        // lexicographic numbering: -z < -y < -x < self < +x < +y < +z
        if (q[4] >= 0) { col.push_back(q[4]); val.push_back(A.az[q[4]]); }
        if (q[2] >= 0) { col.push_back(q[2]); val.push_back(A.ay[q[2]]); }
        if (q[0] >= 0) { col.push_back(q[0]); val.push_back(A.ax[q[0]]); }
        col.push_back(r);
        val.push_back(A.diag[r]);
        if (q[1] >= 0) { col.push_back(q[1]); val.push_back(A.ax[r]); }
        if (q[3] >= 0) { col.push_back(q[3]); val.push_back(A.ay[r]); }
        if (q[5] >= 0) { col.push_back(q[5]); val.push_back(A.az[r]); }
    }
    row_ptr[n] = (int32_t)col.size();
}

void build_grid_from_vector(Grid3f& grid, const std::vector<double>& vector, const Grid3i& indices,
                            float background_value) {
    grid.create(indices.ni, indices.nj, indices.nk);
    for (size_t c = 0; c < grid.data.size(); ++c)
        grid.data[c] = indices.data[c] == ELQ_UNKNOWN_INDEX ? background_value : (float)vector[indices.data[c]];
}

void update_velocity_using_pressure(Device& dev, StaggeredGrid3f& velocity, const Grid3f& pressure,
                                    const Grid3f& level_set, const StaggeredGrid3f& weight, float dx, float dt) {
    elq_grid g = make_grid(level_set.ni, level_set.nj, level_set.nk, dx, level_set, velocity, weight);
    g.pressure = const_cast<float*>(pressure.data.data());
    const int rc = elq_update_velocity(dev.ctx(), &g, dt);
    if (rc != ELQ_OK) raise(dev.ctx(), rc);
}

void extrapolate(Device& dev, Grid3f& grid, Grid3c& valid, int layers) {
    elq_ctx* c = dev.ctx();
    DevArray g(c, grid.data), v(c, valid.data);
    check(c, elq_extrapolate_device(c, grid.ni, grid.nj, grid.nk, g.as<float>(), v.as<uint8_t>(), layers));
    g.fetch(grid.data);
    v.fetch(valid.data);
}

void extrapolate_const(Device& dev, Grid3f& grid, const Grid3c& valid, int layers) {
    elq_ctx* c = dev.ctx();
    DevArray g(c, grid.data), v(c, valid.data);
    check(c, elq_extrapolate_const_device(c, grid.ni, grid.nj, grid.nk, g.as<float>(), v.as<uint8_t>(), layers));
    g.fetch(grid.data);
}

void advect_collocated_grid(Device& dev, Grid3f& grid, const Grid3f& grid_temp, const StaggeredGrid3f& velocity,
                            float dx, float dt) {
    elq_ctx* c = dev.ctx();
    DevArray src(c, grid_temp.data), u(c, velocity.u().data), v(c, velocity.v().data), w(c, velocity.w().data);
    DevArray out(c, grid_temp.data.size() * sizeof(float));
    check(c, elq_advect_collocated_device(c, grid_temp.ni, grid_temp.nj, grid_temp.nk, dx, dt, out.as<float>(),
                                          src.as<float>(), u.as<float>(), v.as<float>(), w.as<float>()));
    grid.create(grid_temp.ni, grid_temp.nj, grid_temp.nk);
    out.fetch(grid.data);
}

void advect_staggered_grid(Device& dev, StaggeredGrid3f& grid, const StaggeredGrid3f& grid_temp,
                           const StaggeredGrid3f& velocity, float dx, float dt) {
    elq_ctx* c = dev.ctx();
    const int ni = grid_temp.v().ni, nj = grid_temp.u().nj, nk = grid_temp.u().nk;
    DevArray au(c, grid_temp.u().data), av(c, grid_temp.v().data), aw(c, grid_temp.w().data);
    DevArray vu(c, velocity.u().data), vv(c, velocity.v().data), vw(c, velocity.w().data);
    DevArray ou(c, bytes_of(grid_temp.u())), ov(c, bytes_of(grid_temp.v())), ow(c, bytes_of(grid_temp.w()));
    check(c, elq_advect_staggered_device(c, ni, nj, nk, dx, dt, ou.as<float>(), ov.as<float>(), ow.as<float>(),
                                         au.as<float>(), av.as<float>(), aw.as<float>(), vu.as<float>(),
                                         vv.as<float>(), vw.as<float>()));
    grid.create(ni, nj, nk);
    ou.fetch(grid.u().data);
    ov.fetch(grid.v().data);
    ow.fetch(grid.w().data);
}

void reinitialize_fast_marching(Device& dev, Grid3f& grid, const Grid3c& valid, float dx) {
    elq_ctx* c = dev.ctx();
    DevArray g(c, grid.data), f(c, valid.data);
    int64_t unreached = 0;
    check(c, elq_reinitialize_device(c, grid.ni, grid.nj, grid.nk, dx, g.as<float>(), f.as<uint8_t>(), &unreached,
                                     nullptr));
    g.fetch(grid.data);
    if (unreached)
        throw std::runtime_error("reinitialize: " + std::to_string(unreached) +
                                 " valid cells have no path to the interface");
}

}  // namespace fluid

void DeviceArray::allocate(elq_ctx* ctx, size_t bytes) {
    reset();
    m_ctx = ctx;
    check(ctx, elq_device_alloc(ctx, bytes < 4 ? 4 : bytes, &m_p));
    m_bytes = bytes;
}

void DeviceArray::reset() {
    if (m_p) elq_device_free(m_ctx, m_p);
    m_p = nullptr;
    m_bytes = 0;
}

namespace {

template <class G>
void h2d(elq_ctx* c, DeviceArray& d, const G& g) {
    check(c, elq_memcpy_h2d(c, d.as<void>(), g.data.data(), bytes_of(g)));
}

template <class G>
void d2h(elq_ctx* c, G& g, const DeviceArray& d) {
    check(c, elq_memcpy_d2h(c, g.data.data(), d.as<void>(), bytes_of(g)));
}

}  // namespace

elq_grid Simulator::device_grid() {
    elq_grid g{};
    g.ni = m_ni;
    g.nj = m_nj;
    g.nk = m_nk;
    g.dx = m_dx;
    g.level_set = d_phi.as<float>();
    g.u = d_u.as<float>();
    g.v = d_v.as<float>();
    g.w = d_w.as<float>();
    g.u_weight = d_uw.as<float>();
    g.v_weight = d_vw.as<float>();
    g.w_weight = d_ww.as<float>();
    g.pressure = d_p.as<float>();
    g.indices = d_idx.as<int32_t>();
    return g;
}

void Simulator::create_grid(float width, int ni, int nj, int nk) {
    m_ni = ni;
    m_nj = nj;
    m_nk = nk;
    m_dx = width / ni;
    m_velocity.create(ni, nj, nk);
    m_velocity_weight.create(ni, nj, nk);
    m_level_set.create(ni, nj, nk);
    m_indices.create(ni, nj, nk);
    m_pressure.create(ni, nj, nk);
    m_velocity_valid.create(ni, nj, nk, 1);  // simulator.cpp:59,63: created with value 1
    m_level_set_valid.create(ni, nj, nk, 1);
    elq_ctx* c = m_dev.ctx();
    const size_t cells = (size_t)ni * nj * nk;
    const size_t fu = (size_t)(ni + 1) * nj * nk, fv = (size_t)ni * (nj + 1) * nk, fw = (size_t)ni * nj * (nk + 1);
    d_phi.allocate(c, cells * 4);
    d_p.allocate(c, cells * 4);
    d_idx.allocate(c, cells * 4);
    d_lsv.allocate(c, cells);
    d_u.allocate(c, fu * 4);
    d_v.allocate(c, fv * 4);
    d_w.allocate(c, fw * 4);
    d_uw.allocate(c, fu * 4);
    d_vw.allocate(c, fv * 4);
    d_ww.allocate(c, fw * 4);
    d_fu.allocate(c, fu);
    d_fv.allocate(c, fv);
    d_fw.allocate(c, fw);
    d_nodal.allocate(c, (size_t)(ni + 1) * (nj + 1) * (nk + 1) * 4);
    h2d(c, d_u, m_velocity.u());
    h2d(c, d_v, m_velocity.v());
    h2d(c, d_w, m_velocity.w());
    h2d(c, d_fu, m_velocity_valid.c[0]);
    h2d(c, d_fv, m_velocity_valid.c[1]);
    h2d(c, d_fw, m_velocity_valid.c[2]);
    h2d(c, d_lsv, m_level_set_valid);
}

void Simulator::init_scene(float pool_height, const std::vector<std::array<float, 4>>& spheres) {
    std::vector<float> flat;
    for (const auto& s : spheres) flat.insert(flat.end(), s.begin(), s.end());
    const int rc = elq_scene_level_set(m_ni, m_nj, m_nk, m_dx, pool_height, (int32_t)spheres.size(),
                                       flat.empty() ? nullptr : flat.data(), m_level_set.data.data());
    if (rc != ELQ_OK) throw std::runtime_error("elq: invalid scene");
    h2d(m_dev.ctx(), d_phi, m_level_set);
}

void Simulator::compute_weights() {
    // the room() solid on the nodal grid, once; the weights are computed on the
    // device from it and stay there (mirrored to the host member once)
    elq_ctx* c = m_dev.ctx();
    std::vector<float> nodal((size_t)(m_ni + 1) * (m_nj + 1) * (m_nk + 1));
    check(c, elq_scene_room_nodal(m_ni, m_nj, m_nk, nodal.data()));
    check(c, elq_memcpy_h2d(c, d_nodal.as<void>(), nodal.data(), nodal.size() * sizeof(float)));
    check(c, elq_compute_weights_device(c, m_ni, m_nj, m_nk, d_nodal.as<float>(), d_uw.as<float>(), d_vw.as<float>(),
                                        d_ww.as<float>()));
    d2h(c, m_velocity_weight.u(), d_uw);
    d2h(c, m_velocity_weight.v(), d_vw);
    d2h(c, m_velocity_weight.w(), d_ww);
}

void Simulator::add_gravity(float dt) {
    elq_ctx* c = m_dev.ctx();
    check(c, elq_add_gravity_device(c, m_ni, m_nj, m_nk, d_v.as<float>(), dt));
}

void Simulator::find_valid_velocity_and_zero_invalid() {
    find_valid_velocity();
    set_invalid_velocity_zero();
}

void Simulator::find_valid_velocity() {
    elq_ctx* c = m_dev.ctx();
    elq_grid g = device_grid();
    check(c, elq_find_valid_velocity_device(c, &g, d_fu.as<uint8_t>(), d_fv.as<uint8_t>(), d_fw.as<uint8_t>()));
}

void Simulator::set_invalid_velocity_zero() {
    elq_ctx* c = m_dev.ctx();
    elq_grid g = device_grid();
    check(c, elq_set_invalid_velocity_zero_device(c, &g, d_fu.as<uint8_t>(), d_fv.as<uint8_t>(), d_fw.as<uint8_t>()));
}

void Simulator::extrapolate_velocity() {
    elq_ctx* c = m_dev.ctx();
    elq_grid g = device_grid();
    check(c, elq_extrapolate_velocity_device(c, &g, d_fu.as<uint8_t>(), d_fv.as<uint8_t>(), d_fw.as<uint8_t>(), 10));
}

void Simulator::find_valid_level_set() {
    elq_ctx* c = m_dev.ctx();
    check(c, elq_find_valid_level_set_device(c, m_ni, m_nj, m_nk, d_nodal.as<float>(), d_lsv.as<uint8_t>()));
}

void Simulator::extrapolate_level_set() {
    elq_ctx* c = m_dev.ctx();
    check(c, elq_extrapolate_const_device(c, m_ni, m_nj, m_nk, d_phi.as<float>(), d_lsv.as<uint8_t>(), 10));
}

void Simulator::advect_velocity(float dt) {
    // m_velocity_temp = m_velocity (simulator.cpp:80): the library keeps the copy
    elq_ctx* c = m_dev.ctx();
    elq_grid g = device_grid();
    check(c, elq_advect_velocity_device(c, &g, dt));
}

void Simulator::advect_level_set(float dt) {
    elq_ctx* c = m_dev.ctx();
    check(c, elq_advect_level_set_device(c, m_ni, m_nj, m_nk, m_dx, dt, d_phi.as<float>(), d_u.as<float>(),
                                         d_v.as<float>(), d_w.as<float>()));
}

void Simulator::reinitialize() {
    elq_ctx* c = m_dev.ctx();
    int64_t unreached = 0;
    check(c, elq_reinitialize_device(c, m_ni, m_nj, m_nk, m_dx, d_phi.as<float>(), d_lsv.as<uint8_t>(), &unreached,
                                     nullptr));
    if (unreached)
        throw std::runtime_error("reinitialize: " + std::to_string(unreached) +
                                 " valid cells have no path to the interface");
}

void Simulator::solve_pressure(float dt) {
    elq_grid g = device_grid();
    const int rc = elq_project_device(m_dev.ctx(), &g, dt, &m_opts, &m_stats);
    if (rc != ELQ_OK && rc != ELQ_ERR_NOT_CONVERGED) raise(m_dev.ctx(), rc);
}

void Simulator::download_level_set() { d2h(m_dev.ctx(), m_level_set, d_phi); }

void Simulator::download_velocity() {
    elq_ctx* c = m_dev.ctx();
    d2h(c, m_velocity.u(), d_u);
    d2h(c, m_velocity.v(), d_v);
    d2h(c, m_velocity.w(), d_w);
}

void Simulator::download_pressure() {
    elq_ctx* c = m_dev.ctx();
    d2h(c, m_pressure, d_p);
    d2h(c, m_indices, d_idx);
}

void Simulator::download_valid() {
    elq_ctx* c = m_dev.ctx();
    d2h(c, m_velocity_valid.c[0], d_fu);
    d2h(c, m_velocity_valid.c[1], d_fv);
    d2h(c, m_velocity_valid.c[2], d_fw);
    d2h(c, m_level_set_valid, d_lsv);
}

void Simulator::upload() {
    elq_ctx* c = m_dev.ctx();
    h2d(c, d_phi, m_level_set);
    h2d(c, d_u, m_velocity.u());
    h2d(c, d_v, m_velocity.v());
    h2d(c, d_w, m_velocity.w());
    h2d(c, d_uw, m_velocity_weight.u());
    h2d(c, d_vw, m_velocity_weight.v());
    h2d(c, d_ww, m_velocity_weight.w());
}

}  // namespace elq_host
