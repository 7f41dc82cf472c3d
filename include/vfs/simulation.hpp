// C++ host interface kept from the reference (luihabl/VkFluidSim):
//   Simulation        src/simulation.h:17-58      (singleton facade)
//   SceneBase         src/scenes/scene.h:14-50
//   GenericScene      src/scenes/generic_scene.h   (JSON-described scene)
//   SPHModel /
//   WCSPHWithBoundaryModel  src/models/model.h:12-114, src/models/wcsph_with_boundary_model.h:14-49
//   LoadScene         src/util/scene_loader.h:8
// Vulkan command buffers are replaced by HIP streams; every device operation
// goes through the C ABI of include/vfs_c.h.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../vfs_c.h"

typedef struct ihipStream_t* hipStream_t;

namespace vfs {

using vec3 = std::array<float, 3>;
using uvec3 = std::array<uint32_t, 3>;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Throws vfs::Error when a C ABI call fails.
void Check(int status, const char* what);

struct BoundingBox {
    vec3 size{0, 0, 0};
    vec3 pos{0, 0, 0};
};

struct SimulationParameters {
    vec3 gravity{0.0f, -9.81f, 0.0f};
    float smooth_radius{0.2f};
};

class SPHModel {
public:
    struct Parameters {
        float time_scale{1.0f};
        int iterations{3};
        int n_particles{8192};
        float fixed_dt{1.0f / 60.0f};
        float target_density{100.0f};
        BoundingBox bounding_box{{10.0f, 10.0f, 10.0f}, {0.0f, 0.0f, 0.0f}};
    };

    virtual ~SPHModel() = default;
    virtual void Init() = 0;
    virtual void Step(hipStream_t stream) = 0;
    virtual void Clear() = 0;

    const Parameters& GetParameters() const { return parameters; }
    BoundingBox GetBoundingBox() const { return parameters.bounding_box; }

    // Compact lattice of count particles in box (model.cpp:35-66, :117-137).
    void SetParticlesInBox(const BoundingBox& box, uint32_t offset = 0, int64_t count = -1);
    // AoS float3 host vectors (model.cpp:139-147).
    void SetParticleState(const std::vector<vec3>& pos, const std::vector<vec3>& vel, uint32_t offset = 0,
                          int64_t count = -1);

    // SPHModel::GetDataBuffers (model.h:38-43): device addresses of the live state.
    vfs_device_buffers GetDataBuffers() const;
    // SPHModel::CreateDataBuffers (model.cpp:149-156): render buffers in the
    // reference's layout; release them with vfs_render_buffers_destroy.
    vfs_render_buffers CreateDataBuffers(bool with_accelerations = false) const;
    // SPHModel::CopyDataBuffers (model.cpp:158-170): the state after the work
    // queued so far into `dst`, ordered on `stream`.
    void CopyDataBuffers(hipStream_t stream, const vfs_render_buffers& dst) const;

    vfs_sim* Handle() const { return sim; }

protected:
    Parameters parameters;
    vfs_sim* sim{nullptr};
};

class WCSPHWithBoundaryModel final : public SPHModel {
public:
    struct Parameters {
        float wall_stiffness{1e4f};
        float stiffness{1000.0f};
        float expoent{7.0f};
        float viscosity_strenght{0.01f};
    };

    // One boundary object with its FP32 grids (BoundaryObjectInfo + grid buffers).
    struct BoundaryObject {
        std::array<float, 16> transform{};  // inverse object matrix, column-major
        std::array<float, 16> rotation{};   // rotation matrix, column-major
        BoundingBox box;
        uvec3 resolution{2, 2, 2};
        std::vector<float> sdf, volume_map;
    };

    WCSPHWithBoundaryModel(const SPHModel::Parameters& base, const Parameters& model,
                           const SimulationParameters& global, std::vector<BoundaryObject> objects,
                           float cell_size = 0.0f, uint32_t flags = 0);
    ~WCSPHWithBoundaryModel() override;

    void Init() override;
    void Step(hipStream_t stream) override;
    void Clear() override;

    const Parameters& GetModelParameters() const { return model_parameters; }
    void SetModelParameters(const Parameters& p);
    vfs_params RawParameters() const;
    // The objects as the simulation received them (GenericScene::boundary_objects
    // with their BoundaryObjectInfo, generic_scene.cpp:125-145).
    const std::vector<BoundaryObject>& GetBoundaryObjects() const { return objects; }
    // Object k as a C ABI record whose grid pointers refer to this model's host grids.
    vfs_boundary_object BoundaryObjectRecord(size_t k) const;

private:
    Parameters model_parameters;
    SimulationParameters global;
    std::vector<BoundaryObject> objects;
    float cell_size;
    uint32_t flags;
};

// LagueModel (src/models/lague_model.h:15-45): the Clavet double-density
// model on predicted positions; default parameters as lague_model.cpp:26-34.
class LagueModel final : public SPHModel {
public:
    struct Parameters {
        float wall_damping_factor{0.95f};
        float pressure_multiplier{288.0f};
        float near_pressure_multiplier{2.16f};
        float viscosity_strenght{0.03f};
    };

    LagueModel(const SPHModel::Parameters& base, const Parameters& model, const SimulationParameters& global,
               uint32_t flags = 0);
    ~LagueModel() override;

    void Init() override;
    void Step(hipStream_t stream) override;
    void Clear() override;

    const Parameters& GetModelParameters() const { return model_parameters; }
    void SetModelParameters(const Parameters& p);
    vfs_params RawParameters() const;

private:
    Parameters model_parameters;
    SimulationParameters global;
    uint32_t flags;
};

class SceneBase {
public:
    virtual ~SceneBase() = default;
    virtual void Init() = 0;
    virtual void Step(hipStream_t stream) = 0;
    virtual void Clear() = 0;
    virtual void Reset() = 0;
    SPHModel* GetModel() const { return time_step_model.get(); }

protected:
    std::unique_ptr<SPHModel> time_step_model;
};

class GenericScene : public SceneBase {
public:
    struct FluidBlock {
        uvec3 size{0, 0, 0};
        vec3 pos{0, 0, 0};
        // extension keys: lattice spacing (default cbrt(1/targetDensity)),
        // uniform jitter amplitude per axis and its splitmix64 seed
        float spacing{0.0f};
        float jitter{0.0f};
        uint64_t seed{0};
    };

    enum class ShapeKind { Box = VFS_SHAPE_BOX, Sphere = VFS_SHAPE_SPHERE, BoxWalls = VFS_SHAPE_BOX_WALLS };

    struct ObjectDef {
        std::string path;           // OBJ resource (the last shape of the file)
        bool analytic{false};
        ShapeKind kind{ShapeKind::Box};
        std::array<double, 6> shape{};  // centre + half extents (or radius)
        vec3 position{0, 0, 0};
        float angle_deg{0.0f};
        vec3 axis{0, 1, 0};
        vec3 scale{1, 1, 1};
        uvec3 resolution{2, 2, 2};
    };

    GenericScene(const SPHModel::Parameters& base, const WCSPHWithBoundaryModel::Parameters& model,
                 const SimulationParameters& global, std::vector<FluidBlock> blocks, std::vector<ObjectDef> objects,
                 float cell_size = 0.0f, uint32_t flags = 0);

    void Init() override;
    void Step(hipStream_t stream) override;
    void Clear() override;
    void Reset() override;

    const std::vector<FluidBlock>& Blocks() const { return blocks; }

private:
    SPHModel::Parameters base;
    WCSPHWithBoundaryModel::Parameters model;
    SimulationParameters global;
    std::vector<FluidBlock> blocks;
    std::vector<ObjectDef> object_defs;
    float cell_size;
    uint32_t flags;
};

// DamBreakScene (src/scenes/dam_break_scene.cpp:8-50): a 50 x 60 x 50 block of
// the Lague model at target density 630 in a 23 x 10 x 10 box, centred in z,
// dt = 1/120 over 3 iterations; gravity and smoothing radius from the
// Simulation's global parameters.
class DamBreakScene : public SceneBase {
public:
    explicit DamBreakScene(uint32_t flags = 0) : flags(flags) {}

    void Init() override;
    void Step(hipStream_t stream) override;
    void Clear() override;
    void Reset() override;

private:
    uvec3 fluid_block_size{50, 60, 50};
    uint32_t flags;
};

class Simulation {
public:
    static Simulation& Get();
    void Init();
    void Step(hipStream_t stream);
    void Clear();
    const SimulationParameters& GetGlobalParameters() const { return parameters; }
    void SetGlobalParameters(const SimulationParameters& p) { parameters = p; }
    void SetScene(std::unique_ptr<SceneBase>&& scene);
    SceneBase* GetScene() { return scene.get(); }
    uint32_t CurrentStep() const { return current_step; }

private:
    Simulation() = default;
    SimulationParameters parameters;
    std::unique_ptr<SceneBase> scene;
    uint32_t current_step{0};
};

// Parses a scene JSON file (reference schema, scene_loader.cpp:42-140, plus the
// optional extension keys) into an uninitialised GenericScene.  Updates the
// Simulation's global parameters like the reference does.
std::unique_ptr<GenericScene> LoadScene(const std::string& config_path, const std::string& resource_root = "");

// splitmix64 stream used by the seeded jitter: u_k = (mix(seed + k*golden) >> 40) * 2^-24.
float SplitMixUniform(uint64_t seed, uint64_t k);

}  // namespace vfs
