// Sample driver in the reference's call order for a frame
// (apps/eulerian_liquid_3d.cpp:106-120, update_second_order) on the drop
// scene: slab y < 0.4 plus a sphere of radius 0.15 at (0.5, 0.7, 0.5), room()
// walls. Set-up: compute_weights, find_valid_level_set (:137-139). Per frame:
// extrapolate_velocity, extrapolate_level_set, advect_level_set,
// advect_velocity, reinitialize, add_gravity, find_valid_velocity,
// set_invalid_velocity_zero, solve_pressure, all on the GPU through the C ABI
// with the Simulator's grids resident in device memory for the whole run.
//
// usage: elq_sample3d [quality=64] [frames=3] [tol=1e-4] [mixed=0] [options]
//   --dump DIR   after each frame f, write DIR/frame<f>_{phi,u,v,w,pressure}.f32
//                (raw little-endian float32 in the reference layouts); the
//                only device-to-host copies of the run
//   --profile    time every step's kernels with HIP events as well
//                (device_ms: the sum over the step's kernels)
//
// Per frame it prints the host wall time of every step (each entry point
// returns after its work completes) and of the whole frame.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "simulator.hpp"

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

double kernel_ms(elq_ctx* ctx) {
    char names[4096];
    double ms[64];
    const int n = elq_last_kernel_times(ctx, names, sizeof(names), ms, nullptr, 64);
    double t = 0.0;
    for (int q = 0; q < n; ++q) t += ms[q];
    return t;
}

void write_f32(const std::string& path, const std::vector<float>& a) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(a.data(), sizeof(float), a.size(), f) != a.size()) {
        if (f) std::fclose(f);
        throw std::runtime_error("cannot write " + path);
    }
    std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::string> pos;
    std::string dump;
    bool profile = false;
    for (int q = 1; q < argc; ++q) {
        if (!std::strcmp(argv[q], "--dump") && q + 1 < argc)
            dump = argv[++q];
        else if (!std::strcmp(argv[q], "--profile"))
            profile = true;
        else
            pos.push_back(argv[q]);
    }
    const int q = pos.size() > 0 ? std::atoi(pos[0].c_str()) : 64;
    const int frames = pos.size() > 1 ? std::atoi(pos[1].c_str()) : 3;
    const double tol = pos.size() > 2 ? std::atof(pos[2].c_str()) : 1e-4;
    const int mixed = pos.size() > 3 ? std::atoi(pos[3].c_str()) : 0;
    try {
        elq_host::Device dev(0);
        elq_host::Simulator sim(dev);
        sim.create_grid(1.0f, q, q, q);
        sim.init_scene(0.4f, {{0.5f, 0.7f, 0.5f, 0.15f}});
        sim.compute_weights();
        sim.find_valid_level_set();
        sim.opts().tol = tol;
        sim.opts().max_iter = 10000;
        sim.opts().precision = mixed ? ELQ_PRECISION_MIXED : ELQ_PRECISION_FP64;
        elq_set_profiling(dev.ctx(), profile ? 1 : 0);
        const float dt = 0.01f;
        struct Step {
            const char* name;
            void (*run)(elq_host::Simulator&, float);
        };
        const Step steps[] = {
            {"extrapolate_velocity", [](elq_host::Simulator& s, float) { s.extrapolate_velocity(); }},
            {"extrapolate_level_set", [](elq_host::Simulator& s, float) { s.extrapolate_level_set(); }},
            {"advect_level_set", [](elq_host::Simulator& s, float t) { s.advect_level_set(t); }},
            {"advect_velocity", [](elq_host::Simulator& s, float t) { s.advect_velocity(t); }},
            {"reinitialize", [](elq_host::Simulator& s, float) { s.reinitialize(); }},
            {"add_gravity", [](elq_host::Simulator& s, float t) { s.add_gravity(t); }},
            {"find_valid_velocity", [](elq_host::Simulator& s, float) { s.find_valid_velocity(); }},
            {"set_invalid_velocity_zero", [](elq_host::Simulator& s, float) { s.set_invalid_velocity_zero(); }},
            {"solve_pressure", [](elq_host::Simulator& s, float t) { s.solve_pressure(t); }},
        };
        for (int f = 0; f < frames; ++f) {
            std::string wall, dev_ms;
            double t_dev = 0.0;
            const double t0 = now_ms();
            for (const Step& st : steps) {
                const double a = now_ms();
                st.run(sim, dt);
                const double b = now_ms();
                char buf[96];
                std::snprintf(buf, sizeof(buf), "%s%s=%.3f", wall.empty() ? "" : " ", st.name, b - a);
                wall += buf;
                if (profile) {
                    const double k = kernel_ms(dev.ctx());
                    t_dev += k;
                    std::snprintf(buf, sizeof(buf), "%s%s=%.3f", dev_ms.empty() ? "" : " ", st.name, k);
                    dev_ms += buf;
                }
            }
            const double t_frame = now_ms() - t0;
            const elq_stats& s = sim.last_stats();
            std::printf("frame %d: N=%lld iterations=%d error=%.3e true_rel_res=%.3e total=%.2f ms frame_ms=%.3f\n", f,
                        (long long)s.n_unknowns, s.iterations, s.error, s.true_rel_residual, s.t_total_ms, t_frame);
            std::printf("  wall_ms: %s\n", wall.c_str());
            if (profile) std::printf("  device_ms (sum %.3f): %s\n", t_dev, dev_ms.c_str());
            std::fflush(stdout);
            if (!dump.empty()) {
                sim.download_level_set();
                sim.download_velocity();
                sim.download_pressure();
                const std::string p = dump + "/frame" + std::to_string(f) + "_";
                write_f32(p + "phi.f32", sim.m_level_set.data);
                write_f32(p + "u.f32", sim.m_velocity.u().data);
                write_f32(p + "v.f32", sim.m_velocity.v().data);
                write_f32(p + "w.f32", sim.m_velocity.w().data);
                write_f32(p + "pressure.f32", sim.m_pressure.data);
            }
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
