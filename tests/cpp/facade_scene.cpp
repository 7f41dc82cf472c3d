// Drives the C++ facade (include/vfs/simulation.hpp) the way the reference's
// application drives its classes (src/main.cpp, src/simulation.cpp):
// LoadScene -> Simulation::SetScene (Init) -> Step once per frame, then
// CopyDataBuffers into render buffers as SceneRenderer::Draw does
// (scene_renderer.cpp:52-55).  Writes the final state, read from the render
// buffers, to a binary file for the GPU test to compare with the C ABI:
//   u32 n, then n*3 f32 positions, n*3 f32 velocities, n f32 densities, n u32 ids.
// Usage: facade_scene SCENE.json FRAMES OUT.bin
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "vfs/simulation.hpp"

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s SCENE.json FRAMES OUT.bin\n", argv[0]);
        return 2;
    }
    try {
        auto& simulation = vfs::Simulation::Get();
        simulation.Init();
        simulation.SetScene(vfs::LoadScene(argv[1]));
        const int frames = std::atoi(argv[2]);
        vfs::SPHModel* model = simulation.GetScene()->GetModel();
        vfs_render_buffers render = model->CreateDataBuffers();
        for (int f = 0; f < frames; f++) {
            simulation.Step(nullptr);
            model->CopyDataBuffers(nullptr, render);
        }
        vfs_sim* sim = model->Handle();
        vfs::Check(vfs_sim_synchronize(sim), "vfs_sim_synchronize");
        vfs_params p{};
        vfs::Check(vfs_sim_get_params(sim, &p), "vfs_sim_get_params");
        const uint32_t n = (uint32_t)p.n_particles;
        std::vector<float> x(3 * (size_t)n), v(3 * (size_t)n), rho(n);
        std::vector<uint32_t> ids(n);
        const bool track = (p.flags & VFS_FLAG_TRACK_IDS) != 0;
        vfs::Check(vfs_render_buffers_read(&render, x.data(), v.data(), rho.data(), nullptr),
                   "vfs_render_buffers_read");
        if (track) vfs::Check(vfs_sim_get_state(sim, nullptr, nullptr, nullptr, ids.data()), "vfs_sim_get_state");
        vfs::Check(vfs_render_buffers_destroy(&render), "vfs_render_buffers_destroy");
        FILE* out = std::fopen(argv[3], "wb");
        if (!out) {
            std::fprintf(stderr, "cannot write %s\n", argv[3]);
            return 1;
        }
        std::fwrite(&n, 4, 1, out);
        std::fwrite(x.data(), 4, x.size(), out);
        std::fwrite(v.data(), 4, v.size(), out);
        std::fwrite(rho.data(), 4, rho.size(), out);
        std::fwrite(ids.data(), 4, ids.size(), out);
        std::fclose(out);
        std::printf("facade_scene: %u particles, %d frames (step %u)\n", n, frames, simulation.CurrentStep());
        simulation.Clear();
    } catch (const vfs::Error& e) {
        std::fprintf(stderr, "vfs error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
