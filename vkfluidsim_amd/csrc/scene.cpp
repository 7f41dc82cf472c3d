// C++ host layer kept from the reference: JSON scene loading, GenericScene,
// the WCSPH model object and the Simulation singleton.  It talks to the
// device only through the C ABI (include/vfs_c.h).
//
// Reference: src/util/scene_loader.cpp:42-140, src/scenes/generic_scene.cpp:21-145,
// src/models/model.cpp:35-66 / :117-147, src/simulation.cpp:20-57.
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/vfs/simulation.hpp"
#include "vfs_common.h"

namespace vfs {

void Check(int status, const char* what) {
    if (status != 0) throw Error(status, std::string(what) + ": " + vfs_last_error());
}

// ------------------------------------------------------------------ JSON
namespace {

struct Json {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;

    const Json* find(const std::string& k) const {
        for (const auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    bool contains(const std::string& k) const { return type == Object && find(k) != nullptr; }
    const Json& at(const std::string& k) const {
        const Json* j = type == Object ? find(k) : nullptr;
        if (!j) throw Error(VFS_ERR_IO, "scene JSON: missing key '" + k + "'");
        return *j;
    }
    const Json& at(size_t i) const {
        if (type != Array || i >= arr.size()) throw Error(VFS_ERR_IO, "scene JSON: array index out of range");
        return arr[i];
    }
    double number() const {
        if (type != Number) throw Error(VFS_ERR_IO, "scene JSON: expected a number");
        return num;
    }
    const std::string& string() const {
        if (type != String) throw Error(VFS_ERR_IO, "scene JSON: expected a string");
        return str;
    }
};

struct Parser {
    const std::string& s;
    size_t i = 0;
    explicit Parser(const std::string& t) : s(t) {}
    [[noreturn]] void error(const std::string& m) {
        throw Error(VFS_ERR_IO, "scene JSON parse error at offset " + std::to_string(i) + ": " + m);
    }
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    bool lit(const char* w) {
        const size_t n = std::strlen(w);
        if (s.compare(i, n, w) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    Json value() {
        ws();
        if (i >= s.size()) error("unexpected end");
        Json j;
        const char c = s[i];
        if (c == '{') {
            j.type = Json::Object;
            i++;
            ws();
            if (i < s.size() && s[i] == '}') {
                i++;
                return j;
            }
            while (true) {
                ws();
                if (i >= s.size() || s[i] != '"') error("expected key");
                std::string k = string_lit();
                ws();
                if (i >= s.size() || s[i] != ':') error("expected ':'");
                i++;
                Json v = value();
                j.obj.emplace_back(std::move(k), std::move(v));
                ws();
                if (i < s.size() && s[i] == ',') {
                    i++;
                    continue;
                }
                if (i < s.size() && s[i] == '}') {
                    i++;
                    return j;
                }
                error("expected ',' or '}'");
            }
        }
        if (c == '[') {
            j.type = Json::Array;
            i++;
            ws();
            if (i < s.size() && s[i] == ']') {
                i++;
                return j;
            }
            while (true) {
                j.arr.push_back(value());
                ws();
                if (i < s.size() && s[i] == ',') {
                    i++;
                    continue;
                }
                if (i < s.size() && s[i] == ']') {
                    i++;
                    return j;
                }
                error("expected ',' or ']'");
            }
        }
        if (c == '"') {
            j.type = Json::String;
            j.str = string_lit();
            return j;
        }
        if (lit("true")) {
            j.type = Json::Bool;
            j.b = true;
            return j;
        }
        if (lit("false")) {
            j.type = Json::Bool;
            return j;
        }
        if (lit("null")) return j;
        const char* begin = s.c_str() + i;
        char* end = nullptr;
        const double v = std::strtod(begin, &end);
        if (end == begin) error("unexpected character");
        i += (size_t)(end - begin);
        j.type = Json::Number;
        j.num = v;
        return j;
    }
    std::string string_lit() {
        std::string out;
        i++;  // opening quote
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) error("bad escape");
                const char e = s[i++];
                switch (e) {
                    case 'n': c = '\n'; break;
                    case 't': c = '\t'; break;
                    case 'r': c = '\r'; break;
                    case 'b': c = '\b'; break;
                    case 'f': c = '\f'; break;
                    case 'u':
                        if (i + 4 > s.size()) error("bad \\u escape");
                        c = (char)std::strtol(s.substr(i, 4).c_str(), nullptr, 16);
                        i += 4;
                        break;
                    default: c = e;
                }
            }
            out.push_back(c);
        }
        if (i >= s.size()) error("unterminated string");
        i++;
        return out;
    }
};

Json parse_json(const std::string& text) {
    Parser p(text);
    Json j = p.value();
    p.ws();
    if (p.i != text.size()) p.error("trailing characters");
    return j;
}

vec3 get_vec3(const Json& j) {
    return {(float)j.at(0).number(), (float)j.at(1).number(), (float)j.at(2).number()};
}
uvec3 get_uvec3(const Json& j) {
    uvec3 r{};
    for (size_t a = 0; a < 3; a++) {
        const double v = j.at(a).number();
        if (v < 0 || v != std::floor(v) || v > 4294967295.0) throw Error(VFS_ERR_IO, "expected unsigned integer");
        r[a] = (uint32_t)v;
    }
    return r;
}

uint64_t splitmix64(uint64_t& state) {
    uint64_t z = (state += UINT64_C(0x9E3779B97F4A7C15));
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}

// Rotation of `deg` degrees about `axis` (glm::rotate(glm::radians(deg), axis)),
// column-major.
std::array<double, 16> rotation_matrix(float deg, const vec3& axis) {
    const double ang = (double)deg * M_PI / 180.0;
    double ax = axis[0], ay = axis[1], az = axis[2];
    const double len = std::sqrt(ax * ax + ay * ay + az * az);
    std::array<double, 16> m{};
    m[15] = 1.0;
    if (len == 0.0) {
        m[0] = m[5] = m[10] = 1.0;
        return m;
    }
    ax /= len;
    ay /= len;
    az /= len;
    const double c = std::cos(ang), s = std::sin(ang), t = 1.0 - c;
    // column 0
    m[0] = c + t * ax * ax;
    m[1] = t * ax * ay + s * az;
    m[2] = t * ax * az - s * ay;
    // column 1
    m[4] = t * ax * ay - s * az;
    m[5] = c + t * ay * ay;
    m[6] = t * ay * az + s * ax;
    // column 2
    m[8] = t * ax * az + s * ay;
    m[9] = t * ay * az - s * ax;
    m[10] = c + t * az * az;
    return m;
}

}  // namespace

float SplitMixUniform(uint64_t seed, uint64_t k) {
    uint64_t st = seed + k * UINT64_C(0x9E3779B97F4A7C15);
    const uint64_t z = splitmix64(st);
    return (float)((double)(z >> 40) * (1.0 / 16777216.0));
}

// ---------------------------------------------------------------- models
void SPHModel::SetParticleState(const std::vector<vec3>& pos, const std::vector<vec3>& vel, uint32_t offset,
                                int64_t count) {
    if (count < 0) count = (int64_t)pos.size();
    if ((int64_t)pos.size() < count || (int64_t)vel.size() < count)
        throw Error(VFS_ERR_INVALID, "SetParticleState: vectors shorter than count");
    Check(vfs_sim_set_state(sim, count ? pos[0].data() : nullptr, count ? vel[0].data() : nullptr, offset, count),
          "vfs_sim_set_state");
}

vfs_device_buffers SPHModel::GetDataBuffers() const {
    vfs_device_buffers b{};
    Check(vfs_sim_device_buffers(sim, &b), "vfs_sim_device_buffers");
    return b;
}

vfs_render_buffers SPHModel::CreateDataBuffers(bool with_accelerations) const {
    vfs_render_buffers b{};
    Check(vfs_render_buffers_create((uint32_t)parameters.n_particles, with_accelerations ? 1 : 0, &b),
          "vfs_render_buffers_create");
    return b;
}

void SPHModel::CopyDataBuffers(hipStream_t stream, const vfs_render_buffers& dst) const {
    Check(vfs_sim_copy_data_buffers(sim, &dst, stream), "vfs_sim_copy_data_buffers");
}

// SpawnCompactParticlesInBox (model.cpp:35-66) via SetParticlesInBox (:117-137).
void SPHModel::SetParticlesInBox(const BoundingBox& box, uint32_t offset, int64_t count) {
    if (count == 0) return;
    const uint32_t n = count > 0 ? (uint32_t)count : (uint32_t)parameters.n_particles;
    const float volume = 1 / parameters.target_density;
    const float diameter = std::cbrt(volume);
    const float nx = std::round(box.size[0] / diameter) - 1;
    const float ny = std::round(box.size[1] / diameter) - 1;
    const float nz = std::round(box.size[2] / diameter) - 1;
    std::vector<vec3> pos(n, vec3{0, 0, 0});
    uint32_t p = 0;
    for (uint32_t i = 0; i < nx && p < n; i++)
        for (uint32_t j = 0; j < ny && p < n; j++)
            for (uint32_t k = 0; k < nz && p < n; k++) {
                pos[p] = {diameter * (float)i + box.pos[0], diameter * (float)j + box.pos[1],
                          diameter * (float)k + box.pos[2]};
                p++;
            }
    std::vector<vec3> vel(n, vec3{0, 0, 0});
    SetParticleState(pos, vel, offset, count);
}

WCSPHWithBoundaryModel::WCSPHWithBoundaryModel(const SPHModel::Parameters& base, const Parameters& model,
                                               const SimulationParameters& g, std::vector<BoundaryObject> objs,
                                               float cs, uint32_t fl)
    : model_parameters(model), global(g), objects(std::move(objs)), cell_size(cs), flags(fl) {
    parameters = base;
}

WCSPHWithBoundaryModel::~WCSPHWithBoundaryModel() { Clear(); }

vfs_params WCSPHWithBoundaryModel::RawParameters() const {
    vfs_params p{};
    for (int a = 0; a < 3; a++) {
        p.gravity[a] = global.gravity[a];
        p.bbox_size[a] = parameters.bounding_box.size[a];
        p.bbox_pos[a] = parameters.bounding_box.pos[a];
    }
    p.smooth_radius = global.smooth_radius;
    p.time_scale = parameters.time_scale;
    p.iterations = parameters.iterations;
    p.n_particles = parameters.n_particles;
    p.fixed_dt = parameters.fixed_dt;
    p.target_density = parameters.target_density;
    p.wall_stiffness = model_parameters.wall_stiffness;
    p.stiffness = model_parameters.stiffness;
    p.expoent = model_parameters.expoent;
    p.viscosity_strenght = model_parameters.viscosity_strenght;
    p.cell_size = cell_size;
    p.flags = flags;
    return p;
}

vfs_boundary_object WCSPHWithBoundaryModel::BoundaryObjectRecord(size_t k) const {
    const auto& o = objects.at(k);
    vfs_boundary_object d{};
    std::memcpy(d.transform, o.transform.data(), sizeof(d.transform));
    std::memcpy(d.rotation, o.rotation.data(), sizeof(d.rotation));
    for (int a = 0; a < 3; a++) {
        d.box_size[a] = o.box.size[a];
        d.box_pos[a] = o.box.pos[a];
        d.resolution[a] = o.resolution[a];
    }
    d.sdf_grid = o.sdf.data();
    d.volume_map_grid = o.volume_map.data();
    return d;
}

void WCSPHWithBoundaryModel::Init() {
    Clear();
    const vfs_params p = RawParameters();
    std::vector<vfs_boundary_object> objs(objects.size());
    for (size_t k = 0; k < objects.size(); k++) objs[k] = BoundaryObjectRecord(k);
    Check(vfs_sim_create(&p, objs.empty() ? nullptr : objs.data(), (uint32_t)objs.size(), &sim), "vfs_sim_create");
}

void WCSPHWithBoundaryModel::Step(hipStream_t stream) { Check(vfs_sim_step(sim, stream), "vfs_sim_step"); }

void WCSPHWithBoundaryModel::Clear() {
    if (sim) vfs_sim_destroy(sim);
    sim = nullptr;
}

void WCSPHWithBoundaryModel::SetModelParameters(const Parameters& mp) {
    model_parameters = mp;
    if (sim) {
        const vfs_params p = RawParameters();
        Check(vfs_sim_set_params(sim, &p), "vfs_sim_set_params");
    }
}

LagueModel::LagueModel(const SPHModel::Parameters& base, const Parameters& model, const SimulationParameters& g,
                       uint32_t fl)
    : model_parameters(model), global(g), flags(fl) {
    parameters = base;
}

LagueModel::~LagueModel() { Clear(); }

vfs_params LagueModel::RawParameters() const {
    vfs_params p{};
    for (int a = 0; a < 3; a++) {
        p.gravity[a] = global.gravity[a];
        p.bbox_size[a] = parameters.bounding_box.size[a];
        p.bbox_pos[a] = parameters.bounding_box.pos[a];
    }
    p.smooth_radius = global.smooth_radius;
    p.time_scale = parameters.time_scale;
    p.iterations = parameters.iterations;
    p.n_particles = parameters.n_particles;
    p.fixed_dt = parameters.fixed_dt;
    p.target_density = parameters.target_density;
    p.flags = flags;
    return p;
}

static vfs_lague_params to_lague(const LagueModel::Parameters& m) {
    return vfs_lague_params{m.wall_damping_factor, m.pressure_multiplier, m.near_pressure_multiplier,
                            m.viscosity_strenght};
}

void LagueModel::Init() {
    Clear();
    const vfs_params p = RawParameters();
    const vfs_lague_params l = to_lague(model_parameters);
    Check(vfs_lague_create(&p, &l, &sim), "vfs_lague_create");
}

void LagueModel::Step(hipStream_t stream) { Check(vfs_sim_step(sim, stream), "vfs_sim_step"); }

void LagueModel::Clear() {
    if (sim) vfs_sim_destroy(sim);
    sim = nullptr;
}

void LagueModel::SetModelParameters(const Parameters& mp) {
    model_parameters = mp;
    if (sim) {
        const vfs_lague_params l = to_lague(model_parameters);
        Check(vfs_lague_set_params(sim, &l), "vfs_lague_set_params");
    }
}

// ---------------------------------------------------------------- scenes
void DamBreakScene::Init() {
    SPHModel::Parameters base;
    base.time_scale = 1.0f;
    base.fixed_dt = 1.0f / 120.0f;
    base.iterations = 3;
    base.n_particles = (int)(fluid_block_size[0] * fluid_block_size[1] * fluid_block_size[2]);
    base.target_density = 630.0f;
    base.bounding_box = BoundingBox{{23.0f, 10.0f, 10.0f}, {0.0f, 0.0f, 0.0f}};
    time_step_model = std::make_unique<LagueModel>(base, LagueModel::Parameters{},
                                                   Simulation::Get().GetGlobalParameters(), flags);
    time_step_model->Init();
    Reset();
}

void DamBreakScene::Step(hipStream_t stream) { time_step_model->Step(stream); }

void DamBreakScene::Clear() {
    if (time_step_model) time_step_model->Clear();
}

// DamBreakScene::Reset (dam_break_scene.cpp:34-48): a box one particle
// diameter larger than the block per axis, centred in z.
void DamBreakScene::Reset() {
    SPHModel* m = time_step_model.get();
    const BoundingBox box = m->GetBoundingBox();
    const float particle_volume = 1 / m->GetParameters().target_density;
    const float particle_diam = std::cbrt(particle_volume);
    BoundingBox block;
    for (int a = 0; a < 3; a++) {
        block.size[a] = particle_diam * (float)(fluid_block_size[a] + 1u);
        block.pos[a] = box.pos[a];
    }
    block.pos[2] += (box.size[2] - block.size[2]) / 2.0f;
    m->SetParticlesInBox(block);
}

GenericScene::GenericScene(const SPHModel::Parameters& b, const WCSPHWithBoundaryModel::Parameters& m,
                           const SimulationParameters& g, std::vector<FluidBlock> bl, std::vector<ObjectDef> od,
                           float cs, uint32_t fl)
    : base(b), model(m), global(g), blocks(std::move(bl)), object_defs(std::move(od)), cell_size(cs), flags(fl) {}

// GenericScene::Init (generic_scene.cpp:21-41) with AddBoundaryObject (:88-124)
// and CreateBoundaryObjectBuffer (:125-145): OBJ meshes (the file's last
// shape, generic_scene.cpp:93-94) or analytic shapes.
void GenericScene::Init() {
    std::vector<WCSPHWithBoundaryModel::BoundaryObject> objs;
    for (const auto& d : object_defs) {
        WCSPHWithBoundaryModel::BoundaryObject o;
        o.resolution = d.resolution;
        const size_t cnt = (size_t)d.resolution[0] * d.resolution[1] * d.resolution[2];
        o.sdf.resize(cnt);
        o.volume_map.resize(cnt);
        float bsize[3], bpos[3];
        // tolerance 0 and margin 8h (generic_scene.cpp:99)
        if (d.analytic) {
            Check(vfs_volume_map_generate((int)d.kind, d.shape.data(), d.resolution.data(), global.smooth_radius,
                                          0.0, 8.0 * global.smooth_radius, o.sdf.data(), o.volume_map.data(), bsize,
                                          bpos, 0),
                  "vfs_volume_map_generate");
        } else {
            vfs_mesh* mesh = nullptr;
            Check(vfs_mesh_load_obj(d.path.c_str(), &mesh), "vfs_mesh_load_obj");
            const int rc = vfs_mesh_volume_map_generate(mesh, d.resolution.data(), global.smooth_radius, 0.0,
                                                        8.0 * global.smooth_radius, o.sdf.data(),
                                                        o.volume_map.data(), bsize, bpos);
            vfs_mesh_destroy(mesh);
            Check(rc, "vfs_mesh_volume_map_generate");
        }
        o.box.size = {bsize[0], bsize[1], bsize[2]};
        o.box.pos = {bpos[0], bpos[1], bpos[2]};
        // object matrix M = T * R * S (transform.cpp:35-46); inverse = S^-1 R^T T^-1
        const auto R = rotation_matrix(d.angle_deg, d.axis);
        std::array<double, 16> inv{};
        for (int r = 0; r < 3; r++) {
            const double sr = 1.0 / (double)d.scale[r];
            double t = 0.0;
            for (int c = 0; c < 3; c++) {
                inv[c * 4 + r] = sr * R[r * 4 + c];  // (R^T)[r][c] = R[c][r]
                t -= R[r * 4 + c] * (double)d.position[c];
            }
            inv[12 + r] = sr * t;
        }
        inv[15] = 1.0;
        for (int e = 0; e < 16; e++) {
            o.transform[e] = (float)inv[e];
            o.rotation[e] = (float)R[e];
        }
        objs.push_back(std::move(o));
    }
    uint64_t total = 0;
    for (const auto& b : blocks) total += (uint64_t)b.size[0] * b.size[1] * b.size[2];
    if (total >= (1u << 30)) throw Error(VFS_ERR_INVALID, "too many particles");
    base.n_particles = (int)total;
    time_step_model = std::make_unique<WCSPHWithBoundaryModel>(base, model, global, std::move(objs), cell_size, flags);
    time_step_model->Init();
    Reset();
}

// GenericScene::Reset (generic_scene.cpp:42-57): one compact lattice per block.
// Extension keys: explicit spacing and seeded uniform jitter.
void GenericScene::Reset() {
    auto* m = time_step_model.get();
    const float density = m->GetParameters().target_density;
    const float particle_volume = 1 / density;
    const float particle_diam = std::cbrt(particle_volume);
    uint32_t offset = 0;
    for (const auto& b : blocks) {
        const uint32_t count = b.size[0] * b.size[1] * b.size[2];
        if (b.spacing <= 0.0f && b.jitter == 0.0f) {
            BoundingBox box;
            for (int a = 0; a < 3; a++) box.size[a] = particle_diam * (float)(b.size[a] + 1u);
            box.pos = b.pos;
            m->SetParticlesInBox(box, offset, count);
        } else {
            const float step = b.spacing > 0.0f ? b.spacing : particle_diam;
            std::vector<vec3> pos(count), vel(count, vec3{0, 0, 0});
            uint32_t p = 0;
            for (uint32_t i = 0; i < b.size[0]; i++)
                for (uint32_t j = 0; j < b.size[1]; j++)
                    for (uint32_t k = 0; k < b.size[2]; k++) {
                        const uint32_t ijk[3] = {i, j, k};
                        for (int a = 0; a < 3; a++) {
                            float x = step * (float)ijk[a] + b.pos[a];
                            if (b.jitter != 0.0f) {
                                const float u = SplitMixUniform(b.seed, 3ull * p + (uint64_t)a);
                                x += (2.0f * u - 1.0f) * b.jitter;
                            }
                            pos[p][a] = x;
                        }
                        p++;
                    }
            m->SetParticleState(pos, vel, offset, count);
        }
        offset += count;
    }
}

void GenericScene::Step(hipStream_t stream) { time_step_model->Step(stream); }

void GenericScene::Clear() {
    if (time_step_model) time_step_model->Clear();
}

// ------------------------------------------------------------ Simulation
Simulation& Simulation::Get() {
    static Simulation instance;
    return instance;
}

// Simulation::Init (simulation.cpp:20-27)
void Simulation::Init() {
    parameters = SimulationParameters{};
    parameters.gravity = {0.0f, -9.81f, 0.0f};
    parameters.smooth_radius = 0.2f;
}

void Simulation::Step(hipStream_t stream) {
    if (scene) {
        scene->Step(stream);
        current_step++;
    }
}

void Simulation::Clear() {
    if (scene) scene->Clear();
}

void Simulation::SetScene(std::unique_ptr<SceneBase>&& s) {
    if (scene) scene->Clear();
    scene = std::move(s);
    scene->Init();
}

// ------------------------------------------------------------ LoadScene
std::unique_ptr<GenericScene> LoadScene(const std::string& config_path, const std::string& resource_root) {
    std::ifstream f(config_path);
    if (!f) throw Error(VFS_ERR_IO, "cannot open scene file '" + config_path + "'");
    std::stringstream ss;
    ss << f.rdbuf();
    const Json data = parse_json(ss.str());

    auto& sim = Simulation::Get();
    SimulationParameters global = sim.GetGlobalParameters();
    const Json& sp = data.at("simulationParameters");
    if (sp.contains("gravity")) global.gravity = get_vec3(sp.at("gravity"));
    if (sp.contains("smoothRadius")) global.smooth_radius = (float)sp.at("smoothRadius").number();
    sim.SetGlobalParameters(global);

    SPHModel::Parameters base{};
    base.time_scale = (float)sp.at("timeScale").number();
    base.iterations = (int)sp.at("iterations").number();
    base.fixed_dt = (float)sp.at("dt").number();
    base.target_density = (float)sp.at("targetDensity").number();
    base.bounding_box.pos = get_vec3(sp.at("boundingBox").at("pos"));
    base.bounding_box.size = get_vec3(sp.at("boundingBox").at("size"));
    const float cell_size = sp.contains("cellSize") ? (float)sp.at("cellSize").number() : 0.0f;

    WCSPHWithBoundaryModel::Parameters mp{};
    const Json& wp = data.at("wcsphParameters");
    mp.stiffness = (float)wp.at("stiffness").number();
    mp.expoent = (float)wp.at("expoent").number();
    mp.viscosity_strenght = (float)wp.at("viscosityStrenght").number();
    mp.wall_stiffness = (float)wp.at("outerWallStiffness").number();

    std::vector<GenericScene::FluidBlock> blocks;
    const Json& fb = data.at("fluidBlocks");
    if (fb.type != Json::Array) throw Error(VFS_ERR_IO, "fluidBlocks must be an array");
    for (const auto& b : fb.arr) {
        GenericScene::FluidBlock blk;
        blk.size = get_uvec3(b.at("size"));
        blk.pos = get_vec3(b.at("pos"));
        if (b.contains("spacing")) blk.spacing = (float)b.at("spacing").number();
        if (b.contains("jitter")) blk.jitter = (float)b.at("jitter").number();
        if (b.contains("seed")) blk.seed = (uint64_t)b.at("seed").number();
        blocks.push_back(blk);
    }

    std::vector<GenericScene::ObjectDef> objects;
    const Json& bo = data.at("boundaryObjects");
    if (bo.type != Json::Array) throw Error(VFS_ERR_IO, "boundaryObjects must be an array");
    for (const auto& o : bo.arr) {
        GenericScene::ObjectDef d;
        if (o.contains("analyticShape")) {
            const Json& a = o.at("analyticShape");
            const std::string& type = a.at("type").string();
            const vec3 c = get_vec3(a.at("center"));
            d.analytic = true;
            d.shape = {c[0], c[1], c[2], 0, 0, 0};
            if (type == "sphere") {
                d.kind = GenericScene::ShapeKind::Sphere;
                d.shape[3] = a.at("radius").number();
            } else if (type == "box" || type == "boxWalls") {
                d.kind = type == "box" ? GenericScene::ShapeKind::Box : GenericScene::ShapeKind::BoxWalls;
                const vec3 e = get_vec3(a.at("halfExtents"));
                d.shape[3] = e[0];
                d.shape[4] = e[1];
                d.shape[5] = e[2];
            } else {
                throw Error(VFS_ERR_IO, "unknown analyticShape type '" + type + "'");
            }
            if (o.contains("resourcePath")) d.path = o.at("resourcePath").string();
        } else {
            d.path = o.at("resourcePath").string();
            if (!resource_root.empty()) d.path = resource_root + "/" + d.path;
        }
        if (o.contains("position")) d.position = get_vec3(o.at("position"));
        if (o.contains("rotation")) {
            d.angle_deg = (float)o.at("rotation").at("angle").number();
            d.axis = get_vec3(o.at("rotation").at("axis"));
        }
        if (o.contains("scale")) d.scale = get_vec3(o.at("scale"));
        d.resolution = get_uvec3(o.at("volumeMapResolution"));
        objects.push_back(d);
    }
    uint32_t flags = 0;
    if (data.contains("trackParticleIds") && data.at("trackParticleIds").type == Json::Bool &&
        data.at("trackParticleIds").b)
        flags |= VFS_FLAG_TRACK_IDS;
    return std::make_unique<GenericScene>(base, mp, global, std::move(blocks), std::move(objects), cell_size, flags);
}

}  // namespace vfs

// ------------------------------------------------------------ C ABI
struct vfs_scene {
    std::unique_ptr<vfs::GenericScene> scene;
};

namespace {
template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const vfs::Error& e) {
        return vfs::fail(e.code, e.what());
    } catch (const std::exception& e) {
        return vfs::fail(VFS_ERR_INVALID, e.what());
    }
}
}  // namespace

extern "C" {

int vfs_scene_load(const char* json_path, const char* resource_root, vfs_scene** out) {
    if (!json_path || !out) return vfs::fail(VFS_ERR_INVALID, "null argument");
    *out = nullptr;
    return guarded([&] {
        auto s = std::make_unique<vfs_scene>();
        s->scene = vfs::LoadScene(json_path, resource_root ? resource_root : "");
        s->scene->Init();
        *out = s.release();
    });
}

int vfs_scene_sim(vfs_scene* scene, vfs_sim** sim) {
    if (!scene || !sim || !scene->scene || !scene->scene->GetModel())
        return vfs::fail(VFS_ERR_INVALID, "null argument");
    *sim = scene->scene->GetModel()->Handle();
    return 0;
}

int vfs_scene_reset(vfs_scene* scene) {
    if (!scene || !scene->scene) return vfs::fail(VFS_ERR_INVALID, "null argument");
    return guarded([&] { scene->scene->Reset(); });
}

int vfs_scene_step(vfs_scene* scene, void* stream) {
    if (!scene || !scene->scene) return vfs::fail(VFS_ERR_INVALID, "null argument");
    return guarded([&] { scene->scene->Step(reinterpret_cast<hipStream_t>(stream)); });
}

int vfs_scene_boundary_objects(vfs_scene* scene, vfs_boundary_object* out, uint32_t capacity, uint32_t* count) {
    if (!scene || !scene->scene || !count) return vfs::fail(VFS_ERR_INVALID, "null argument");
    auto* model = dynamic_cast<vfs::WCSPHWithBoundaryModel*>(scene->scene->GetModel());
    if (!model) return vfs::fail(VFS_ERR_STATE, "the scene has no WCSPH model");
    const auto& objs = model->GetBoundaryObjects();
    *count = (uint32_t)objs.size();
    if (out)
        for (uint32_t k = 0; k < objs.size() && k < capacity; k++) out[k] = model->BoundaryObjectRecord(k);
    return 0;
}

int vfs_scene_destroy(vfs_scene* scene) {
    if (scene) {
        if (scene->scene) scene->scene->Clear();
        delete scene;
    }
    return 0;
}

}  // extern "C"
