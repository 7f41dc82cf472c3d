// Drives the C++ facade's DamBreakScene (LagueModel, include/vfs/simulation.hpp)
// as the reference's application would (src/scenes/dam_break_scene.cpp):
// Simulation::SetScene (Init + Reset) -> Step once per frame.  Writes the final
// state for the GPU test to compare with the ctypes path:
//   u32 n, then n*3 f32 positions, n*3 f32 velocities, n f32 densities.
// Usage: facade_lague FRAMES OUT.bin
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "vfs/simulation.hpp"

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s FRAMES OUT.bin\n", argv[0]);
        return 2;
    }
    try {
        auto& simulation = vfs::Simulation::Get();
        simulation.Init();
        simulation.SetScene(std::make_unique<vfs::DamBreakScene>());
        const int frames = std::atoi(argv[1]);
        for (int f = 0; f < frames; f++) simulation.Step(nullptr);
        vfs_sim* sim = simulation.GetScene()->GetModel()->Handle();
        vfs::Check(vfs_sim_synchronize(sim), "vfs_sim_synchronize");
        const uint32_t n = (uint32_t)simulation.GetScene()->GetModel()->GetParameters().n_particles;
        std::vector<float> x(3 * (size_t)n), v(3 * (size_t)n), rho(n);
        vfs::Check(vfs_sim_get_state(sim, x.data(), v.data(), rho.data(), nullptr), "vfs_sim_get_state");
        FILE* out = std::fopen(argv[2], "wb");
        if (!out) {
            std::fprintf(stderr, "cannot write %s\n", argv[2]);
            return 1;
        }
        std::fwrite(&n, 4, 1, out);
        std::fwrite(x.data(), 4, x.size(), out);
        std::fwrite(v.data(), 4, v.size(), out);
        std::fwrite(rho.data(), 4, rho.size(), out);
        std::fclose(out);
        std::printf("facade_lague: %u particles, %d frames\n", n, frames);
        simulation.Clear();
    } catch (const vfs::Error& e) {
        std::fprintf(stderr, "vfs error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
