// Sample driver in the reference's call order for a frame
// (apps/eulerian_liquid_3d.cpp:106-120, update_second_order) on the drop
// scene: slab y < 0.4 plus a sphere of radius 0.15 at (0.5, 0.7, 0.5), room()
// walls. Set-up: compute_weights, find_valid_level_set (:137-139). Per frame:
// extrapolate_velocity, extrapolate_level_set, advect_level_set,
// advect_velocity, reinitialize, add_gravity, find_valid_velocity,
// set_invalid_velocity_zero, solve_pressure, all on the GPU through the C ABI.
//
// usage: elq_sample3d [quality=64] [frames=3] [tol=1e-4] [mixed=0]
#include <cstdio>
#include <cstdlib>
#include <exception>

#include "simulator.hpp"

int main(int argc, char** argv) {
    const int q = argc > 1 ? std::atoi(argv[1]) : 64;
    const int frames = argc > 2 ? std::atoi(argv[2]) : 3;
    const double tol = argc > 3 ? std::atof(argv[3]) : 1e-4;
    const int mixed = argc > 4 ? std::atoi(argv[4]) : 0;
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
        const float dt = 0.01f;
        for (int f = 0; f < frames; ++f) {
            sim.extrapolate_velocity();
            sim.extrapolate_level_set();
            sim.advect_level_set(dt);
            sim.advect_velocity(dt);
            sim.reinitialize();
            sim.add_gravity(dt);
            sim.find_valid_velocity();
            sim.set_invalid_velocity_zero();
            sim.solve_pressure(dt);
            const elq_stats& s = sim.last_stats();
            std::printf("frame %d: N=%lld iterations=%d error=%.3e true_rel_res=%.3e total=%.2f ms\n", f,
                        (long long)s.n_unknowns, s.iterations, s.error, s.true_rel_residual, s.t_total_ms);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
