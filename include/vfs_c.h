/*
 * vfs_c.h -- C ABI of the MI355X WCSPH substep engine (libvfs.so).
 *
 * Every entry point returns 0 on success and a negative vfs_status on failure;
 * vfs_last_error() then returns a thread-local message.  Device work is
 * stream-ordered on the caller's hipStream_t, passed as `void *stream`
 * (NULL = the library's own stream for a vfs_sim, the null stream elsewhere).
 * Host arrays are owned by the caller and copied; device buffers created by
 * the library are freed by the matching *_destroy call.
 *
 * Reference interfaces replaced (luihabl/VkFluidSim, paths relative to the
 * reference root) are cited per declaration.
 */
#ifndef VFS_C_H
#define VFS_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VFS_ABI_VERSION 2

typedef enum vfs_status {
    VFS_OK = 0,
    VFS_ERR_INVALID = -1,   /* bad argument (null pointer, size out of range, ...) */
    VFS_ERR_HIP = -2,       /* a HIP runtime call failed */
    VFS_ERR_STATE = -3,     /* object not in a state that allows the call */
    VFS_ERR_IO = -4,        /* file could not be read / parsed */
    VFS_ERR_NOMEM = -5,     /* allocation failed */
    VFS_ERR_DIVERGED = -6   /* VFS_FLAG_CHECK_STATE: the state holds non-finite or escaped particles */
} vfs_status;

/* Flags for vfs_params.flags. */
#define VFS_FLAG_TRACK_IDS 0x1u   /* carry a particle-id buffer through every reorder */
#define VFS_FLAG_NO_GRAPH 0x2u    /* launch substeps eagerly instead of replaying a hipGraph */
#define VFS_FLAG_STORE_ACCEL 0x4u /* keep the per-particle acceleration buffer up to date */
/* After the substeps of each vfs_sim_step / vfs_sim_substeps call, check the
 * state (vfs_sim_check_state with a margin of the box's largest extent) and
 * return VFS_ERR_DIVERGED when a particle has a non-finite position or
 * velocity or has left the domain.  The call then waits for its substeps.
 * Single-scene simulations only. */
#define VFS_FLAG_CHECK_STATE 0x8u

/* Neighbour-list entries kept per particle and substep (a list longer than
 * this falls back to walking the cells, with identical results). */
#define VFS_NEIGHBOUR_CAPACITY 64u

/*
 * Parameters of one scene.  Field meaning follows
 *   SimulationParameters            src/simulation.h:10-13
 *   SPHModel::Parameters            src/models/model.h:14-21
 *   WCSPHWithBoundaryModel::Parameters src/models/wcsph_with_boundary_model.h:26-33
 * (the two misspelled names `expoent` and `viscosity_strenght` are kept).
 */
typedef struct vfs_params {
    float gravity[3];
    float smooth_radius;      /* in [2^-40, 2^40] (VFS_ERR_INVALID otherwise) */
    float time_scale;
    int32_t iterations;       /* substeps per vfs_sim_step */
    int32_t n_particles;
    float fixed_dt;
    float target_density;
    float bbox_size[3];       /* gfx::BoundingBox{size, pos} (src/gfx/mesh.h:13-16) */
    float bbox_pos[3];
    float wall_stiffness;
    float stiffness;
    float expoent;
    float viscosity_strenght;
    float cell_size;          /* SpatialHash cell size; 0 selects smooth_radius (model.cpp:85-86) */
    uint32_t flags;           /* VFS_FLAG_* */
} vfs_params;

/*
 * One boundary object, as WCSPHWithBoundaryModel::BoundaryObjectInfo
 * (src/models/wcsph_with_boundary_model.h:16-24), with host grids that the
 * library uploads.  Matrices are column-major 4x4 (glm layout):
 *   transform = inverse(object matrix)   (src/scenes/generic_scene.cpp:135)
 *   rotation  = mat4_cast(rotation)      (src/scenes/generic_scene.cpp:136)
 * Grid index of node (i,j,k) is i + nx*(j + ny*k).
 */
typedef struct vfs_boundary_object {
    float transform[16];
    float rotation[16];
    float box_size[3];
    float box_pos[3];
    uint32_t resolution[3];
    uint32_t reserved;
    const float *sdf_grid;
    const float *volume_map_grid;
} vfs_boundary_object;

/* Device addresses of a simulation's particle buffers (SPHModel::GetDataBuffers,
 * src/models/model.h:38-43).  Positions and velocities are float4 per particle
 * (x, y, z, unused); densities are float; particle_ids is NULL without
 * VFS_FLAG_TRACK_IDS; accelerations is NULL without VFS_FLAG_STORE_ACCEL.
 * sorted_keys and cell_start are read-only: the next substep re-sorts
 * against them. */
typedef struct vfs_device_buffers {
    float *positions;
    float *velocities;
    float *densities;
    float *accelerations;
    uint32_t *particle_ids;
    uint32_t *sorted_keys;
    uint32_t *sorted_indices;
    uint32_t *cell_start;     /* n_particles + 1 entries: bucket k is [cell_start[k], cell_start[k+1]) */
    uint32_t n_particles;
} vfs_device_buffers;

/* Launch configuration of a simulation.  With cell_scale = 1 every setting
 * produces bit-identical results; only the speed changes.  cell_scale > 1
 * hashes with cells that many times the scene's own cell size: the same
 * neighbours are found, but summed in another order (no bit parity with the
 * reference hash). */
typedef struct vfs_tuning {
    int32_t sort_bits;   /* radix digit width of the key sort: 8, 9 or 10 */
    int32_t sort_items;  /* keys per thread of a sort tile: 8, 12 or 16 */
    int32_t block_size;  /* threads per block of the neighbour kernels: 128 or 256 */
    int32_t stage_rows;  /* 1: stage stencil rows of positions in LDS; 0: read them through L1/L2 */
    float cell_scale;    /* hash cell size / the scene's cell size, in [1, 4]; 1 = reference hash */
} vfs_tuning;

typedef struct vfs_sim vfs_sim;
typedef struct vfs_batch vfs_batch;
typedef struct vfs_spatial_hash vfs_spatial_hash;

/* ---- library ---------------------------------------------------------- */
const char *vfs_last_error(void);
int vfs_abi_version(void);
/* Number of visible HIP devices (0 without a GPU; never fails). */
int vfs_device_count(int *count);

/* ---- device memory and streams (plumbing for callers without their own allocator) */
int vfs_dev_malloc(void **ptr, size_t bytes);
int vfs_dev_free(void *ptr);
int vfs_memcpy_h2d(void *dst_dev, const void *src_host, size_t bytes);
int vfs_memcpy_d2h(void *dst_host, const void *src_dev, size_t bytes);
int vfs_memcpy_d2d(void *dst_dev, const void *src_dev, size_t bytes);
int vfs_memset_d32(void *dst_dev, uint32_t value, size_t count);
int vfs_stream_create(void **stream);
int vfs_stream_destroy(void *stream);
int vfs_stream_synchronize(void *stream);

/* ---- simulation (the substep engine) ----------------------------------- */

/* Replaces WCSPHWithBoundaryModel(params) + Init (wcsph_with_boundary_model.cpp:24-79)
 * and GenericScene::CreateBoundaryObjectBuffer (generic_scene.cpp:125-145). */
int vfs_sim_create(const vfs_params *params, const vfs_boundary_object *objects,
                   uint32_t n_objects, vfs_sim **out);
/* Replaces SPHModel::Clear / WCSPHWithBoundaryModel buffers Destroy (model.cpp:102-111). */
int vfs_sim_destroy(vfs_sim *sim);
/* Replaces SPHModel::SetParticleState (model.cpp:139-147): AoS float3 host arrays,
 * densities of the range are zeroed.  count < 0 means n_particles - offset.
 * With VFS_FLAG_TRACK_IDS the particle ids restart at the current indices. */
int vfs_sim_set_state(vfs_sim *sim, const float *positions, const float *velocities,
                      uint32_t offset, int64_t count);
/* Runtime parameter update (the ImGui sliders: model.cpp:236-264,
 * wcsph_with_boundary_model.cpp:119-141, simulation.cpp:90-104).  n_particles
 * and flags must not change. */
int vfs_sim_set_params(vfs_sim *sim, const vfs_params *params);
int vfs_sim_get_params(const vfs_sim *sim, vfs_params *params);
/* Replaces WCSPHWithBoundaryModel::Step (wcsph_with_boundary_model.cpp:81-117):
 * params.iterations substeps of dt = time_scale*fixed_dt/iterations. */
int vfs_sim_step(vfs_sim *sim, void *stream);
/* n substeps at the same dt (one substep is the unit of work of the benchmark). */
int vfs_sim_substeps(vfs_sim *sim, uint32_t n, void *stream);
/* Blocks until every substep queued on the simulation stream has finished. */
int vfs_sim_synchronize(vfs_sim *sim);
/* Host copies of the current state (any pointer may be NULL).  positions /
 * velocities are AoS float3; particle_ids requires VFS_FLAG_TRACK_IDS. */
int vfs_sim_get_state(vfs_sim *sim, float *positions, float *velocities, float *densities,
                      uint32_t *particle_ids);
/* Hash tables of the last substep (SpatialHash buffers, spatial_hash.h:23-26):
 * sorted keys, the permutation applied by that substep's reorder, and
 * reference-format offsets (first index of key k, n if empty). */
int vfs_sim_get_hash(vfs_sim *sim, uint32_t *sorted_keys, uint32_t *sorted_indices,
                     uint32_t *offsets);
/* Boundary volume psi and its gradient (float3 AoS) of the last substep
 * (CalculateBoundaryVolume outputs, wcsph_with_boundary_model.slang:102-151). */
int vfs_sim_get_boundary(vfs_sim *sim, float *boundary_density, float *boundary_gradient);
int vfs_sim_device_buffers(vfs_sim *sim, vfs_device_buffers *out);
int vfs_sim_get_tuning(const vfs_sim *sim, vfs_tuning *tuning);
int vfs_sim_set_tuning(vfs_sim *sim, const vfs_tuning *tuning);
/* Autotuner: times every vfs_tuning on `substeps_per_trial` substeps from the
 * current state (restored before each trial and at the end, so the scene does
 * not advance), keeps the fastest and returns it with its mean ms per
 * substep.  No counterpart in the reference (its launch shape is fixed: 256
 * threads, atomic counting sort). */
int vfs_sim_autotune(vfs_sim *sim, uint32_t substeps_per_trial, vfs_tuning *best, float *best_ms_per_substep);
/* Runs n substeps eagerly (advancing the state) with a HIP event pair around
 * every kernel on the simulation stream.  For each stage writes the mean time
 * per substep in ms (summed over its launches), the launches per substep
 * (total launches / n, rounded down: a stage of the first substep's full sort
 * alone, such as "reorder" or "sort_histogram", reports 0 when n > 1) and the
 * stage name (at most max_stages entries; any output array may be NULL). */
int vfs_sim_profile_stages(vfs_sim *sim, uint32_t n, float *ms, uint32_t *launches,
                           const char **names, uint32_t max_stages, uint32_t *n_stages);
/* Neighbour lists of the last substep: mean and largest entries per particle
 * (in-range neighbours, the particle itself included once per visit of its
 * own bucket) and the number of particles whose list exceeded the list
 * capacity of VFS_NEIGHBOUR_CAPACITY entries (their forces walk the cells).
 * No counterpart in the reference (it has no neighbour lists). */
int vfs_sim_neighbour_stats(vfs_sim *sim, double *mean, uint32_t *max_count, uint32_t *overflowed);
/* Counts, on the state after the work queued so far, the particles with a
 * non-finite position or velocity component and the particles farther than
 * `margin` (>= 0) outside the bounding box on any axis (the box walls are a
 * penalty force, so a stable run keeps particles within a fraction of the
 * smoothing radius of it).  Blocks until the count is done.  No counterpart in
 * the reference (it has no state check; VK_CHECK aborts on API errors only,
 * src/gfx/common.h:20-29). */
int vfs_sim_check_state(vfs_sim *sim, float margin, uint32_t *non_finite, uint32_t *outside);
/* Keys that changed bucket in the last substep: the keys its re-sort sorted
 * and merged (n when that substep sorted every key).  No counterpart in the
 * reference (it sorts every key every substep, sort.cpp:107-143). */
int vfs_sim_sort_stats(vfs_sim *sim, uint32_t *changed_keys);

/* ---- batches of independent scenes ------------------------------------ */
/* No counterpart in the reference (one scene per Simulation, src/simulation.h:49).
 * A batch steps n_scenes scenes of uneven sizes together: one launch per
 * stage serves every scene, the substep is one hipGraph, and each scene
 * evolves exactly as it would alone in a vfs_sim (bit-identical).  Scene k
 * has params[k] (all with the same flags) and objects_per_scene[k] boundary
 * objects, listed scene after scene in `objects`.  Scenes are padded to 64
 * entries inside the batch; the total must stay below 2^28 entries. */
int vfs_batch_create(const vfs_params *params, uint32_t n_scenes, const vfs_boundary_object *objects,
                     const uint32_t *objects_per_scene, vfs_batch **out);
int vfs_batch_destroy(vfs_batch *batch);
/* Scene count, particles over all scenes, and entries per array (with padding). */
int vfs_batch_info(const vfs_batch *batch, uint32_t *n_scenes, uint64_t *n_particles, uint32_t *entries);
/* Per-scene versions of the vfs_sim calls (AoS float3 host arrays of that scene). */
int vfs_batch_set_state(vfs_batch *batch, uint32_t scene, const float *positions, const float *velocities);
int vfs_batch_set_params(vfs_batch *batch, uint32_t scene, const vfs_params *params);
/* n substeps of every scene, each at its own dt. */
int vfs_batch_substeps(vfs_batch *batch, uint32_t n, void *stream);
int vfs_batch_synchronize(vfs_batch *batch);
int vfs_batch_get_state(vfs_batch *batch, uint32_t scene, float *positions, float *velocities,
                        float *densities, uint32_t *particle_ids);
/* The scene's hash tables in its own index space (as vfs_sim_get_hash). */
int vfs_batch_get_hash(vfs_batch *batch, uint32_t scene, uint32_t *sorted_keys, uint32_t *sorted_indices,
                       uint32_t *offsets);
int vfs_batch_get_boundary(vfs_batch *batch, uint32_t scene, float *boundary_density, float *boundary_gradient);
int vfs_batch_get_tuning(const vfs_batch *batch, vfs_tuning *tuning);
int vfs_batch_set_tuning(vfs_batch *batch, const vfs_tuning *tuning);
/* As vfs_sim_autotune; with allow_cell_size != 0 it also tries cell scales
 * 1.25 and 1.5 (results then differ from the reference hash's in rounding). */
int vfs_batch_autotune(vfs_batch *batch, uint32_t substeps_per_trial, int allow_cell_size, vfs_tuning *best,
                       float *best_ms_per_substep);
int vfs_batch_profile_stages(vfs_batch *batch, uint32_t n, float *ms, uint32_t *launches,
                             const char **names, uint32_t max_stages, uint32_t *n_stages);
int vfs_batch_neighbour_stats(vfs_batch *batch, double *mean, uint32_t *max_count, uint32_t *overflowed);

/* ---- differentiable substeps ------------------------------------------- */
/* No counterpart in the reference (it has no differentiation).  A hand-written
 * adjoint of density, pressure, viscosity, boundary volume and integration,
 * with the neighbour lists of each substep frozen.  The loss is the mean
 * squared difference between the density field sampled on a grid,
 * rho(g) = sum_j m W(|g - x_j|) over the particles within h of g, and a
 * target field.  The forward pass keeps (x, v) every checkpoint_every
 * substeps; the backward pass restores each checkpoint, recomputes its
 * substeps (hash, sort, lists) and reverses them.  Single-scene simulations
 * with the scene's own cell size only; every list must fit
 * VFS_NEIGHBOUR_CAPACITY. */
typedef struct vfs_density_grid {
    float origin[3];   /* position of grid point (0, 0, 0) */
    float spacing[3];
    uint32_t dims[3];  /* point (i, j, k) is origin + (i, j, k) * spacing, index i + nx (j + ny k) */
} vfs_density_grid;

typedef struct vfs_adjoint vfs_adjoint;
int vfs_adjoint_create(vfs_sim *sim, uint32_t n_substeps, uint32_t checkpoint_every,
                       const vfs_density_grid *grid, vfs_adjoint **out);
int vfs_adjoint_destroy(vfs_adjoint *adj);
/* Sets the simulation's state from device AoS float3 arrays (n_particles x 3),
 * runs n_substeps substeps and returns the loss.  target_dev: grid values on
 * the device (NULL: zero target); density_dev: optional device output of the
 * sampled density field. */
int vfs_adjoint_forward(vfs_adjoint *adj, const float *x0_dev, const float *v0_dev, const float *target_dev,
                        float *density_dev, double *loss);
/* Gradients of the last forward's loss with respect to the initial positions
 * and velocities (device AoS float3 outputs, either may be NULL) and to the
 * viscosity parameter (viscosity_strenght).  The simulation is left at the
 * forward's final positions and velocities. */
int vfs_adjoint_backward(vfs_adjoint *adj, float *grad_x0_dev, float *grad_v0_dev, double *grad_viscosity);
/* Runs the reverse pass of the last forward once more (the gradients are not
 * returned; the simulation is left as by vfs_adjoint_backward) with a HIP
 * event pair around every launch group on the simulation stream, and reports
 * per stage the total time in ms over the whole pass, the number of timed
 * groups and the stage name (at most max_stages entries; any output array may
 * be NULL): the loss gradient, the recomputed substeps' sort stages, density
 * and forces_integrate, the list-capacity check, and the reverse kernels
 * adj_integrate, adj_forces, adj_viscosity_sum, adj_boundary, adj_density.
 * No counterpart in the reference (it has no adjoint). */
int vfs_adjoint_profile_backward(vfs_adjoint *adj, float *ms, uint32_t *launches, const char **names,
                                 uint32_t max_stages, uint32_t *n_stages);

/* ---- compute layer drop-ins (reference src/compute headers) ------------ */

/* SpatialHash (src/compute/spatial_hash.h:9-34): Init(n, cell_size), Run(positions),
 * buffers.  positions_dev is float3 AoS (12-byte stride) or float4 (16-byte
 * stride) as selected by `stride_floats` (3 or 4). */
int vfs_spatial_hash_create(uint32_t n, float cell_size, vfs_spatial_hash **out);
int vfs_spatial_hash_set_cell_size(vfs_spatial_hash *h, float cell_size);
int vfs_spatial_hash_run(vfs_spatial_hash *h, const float *positions_dev, uint32_t stride_floats,
                         void *stream);
int vfs_spatial_hash_buffers(vfs_spatial_hash *h, uint32_t **spatial_keys,
                             uint32_t **spatial_indices, uint32_t **spatial_offsets);
int vfs_spatial_hash_destroy(vfs_spatial_hash *h);

/* GPUCountSort::Run (src/compute/sort.cpp:107-143): sorts (items, keys) by key
 * in place, keys <= max_value.  Stable: equal keys keep their input order.
 * The caller provides workspace of vfs_sort_workspace_size bytes. */
int vfs_sort_workspace_size(uint32_t n, uint32_t max_value, size_t *bytes);
int vfs_sort_pairs(uint32_t *items_dev, uint32_t *keys_dev, uint32_t n, uint32_t max_value,
                   void *workspace, size_t workspace_bytes, void *stream);
/* vfs_sort_pairs without a workspace argument and without any allocation: the
 * workspace is a fixed device array of the library (26 MiB per GPU, held until
 * the process ends), so n is limited to 2^20 keys (VFS_ERR_INVALID beyond).
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under stream capture, serving loops, harnesses that
 * forbid hipMalloc).  All calls on one GPU share that array: calls that
 * overlap on one GPU race, so the caller serialises them. */
#define VFS_SORT_NOALLOC_MAX_KEYS (1u << 20)
int vfs_sort_pairs_noalloc(uint32_t *items_dev, uint32_t *keys_dev, uint32_t n, uint32_t max_value,
                           void *stream);
/* GPUScan::Run (src/compute/sort.cpp:57-86): exclusive prefix sum in place. */
int vfs_scan_workspace_size(uint32_t n, size_t *bytes);
int vfs_scan_exclusive(uint32_t *elements_dev, uint32_t n, void *workspace, size_t workspace_bytes,
                       void *stream);
/* SpatialOffset::Run (src/compute/sort.cpp:155-176). */
int vfs_spatial_offsets(const uint32_t *sorted_keys_dev, uint32_t *offsets_dev, uint32_t n,
                        int init, void *stream);
/* BufferReorder::Reorder (src/compute/sort.cpp:206-219), one float3 AoS buffer. */
int vfs_reorder_f32x3(const uint32_t *sorted_indices_dev, const float *src_dev, float *dst_dev,
                      uint32_t n, void *stream);

/* ---- Lague (Clavet double-density) model ------------------------------
 * LagueModel (src/models/lague_model.h:17-24, lague_model.cpp:87-125): per
 * substep, external forces with predicted positions x + v / 120, the spatial
 * hash of the predictions, density and near density (spiky kernels),
 * pressure, near pressure and Poly6 viscosity, integration and wall
 * collisions with damping.  The handle is a vfs_sim: vfs_sim_step /
 * substeps / set_state / get_state / get_hash / set_params / destroy apply
 * (vfs_params' WCSPH fields are unused).  One reading differs from the
 * reference: its pressure kernel reads neighbour velocities while other
 * invocations of the same dispatch update theirs; here every read sees the
 * velocities from before the dispatch. */
typedef struct vfs_lague_params {
    float wall_damping_factor;
    float pressure_multiplier;
    float near_pressure_multiplier;
    float viscosity_strenght;
} vfs_lague_params;
/* LagueModel(base, par) + Init; lague == NULL selects the reference's
 * defaults {0.95, 288, 2.16, 0.03} (lague_model.cpp:26-34). */
int vfs_lague_create(const vfs_params *params, const vfs_lague_params *lague, vfs_sim **out);
/* The parameter sliders of LagueModel::DrawDebugUI (lague_model.cpp:131-154). */
int vfs_lague_set_params(vfs_sim *sim, const vfs_lague_params *lague);
int vfs_lague_get_params(const vfs_sim *sim, vfs_lague_params *lague);

/* ---- boundary grids (host precompute, src/util/volume_map.cpp:9-54) --- */

/* Analytic shapes whose signed distance the grid generator evaluates. */
typedef enum vfs_shape_kind {
    VFS_SHAPE_BOX = 0,        /* solid box: params = {cx, cy, cz, hx, hy, hz} (centre, half extents) */
    VFS_SHAPE_SPHERE = 1,     /* solid sphere: params = {cx, cy, cz, r} */
    VFS_SHAPE_BOX_WALLS = 2   /* box container (solid outside): params = {cx, cy, cz, hx, hy, hz} */
} vfs_shape_kind;

/* Fills sdf and volume map grids (res[0]*res[1]*res[2] floats each) for an
 * analytic shape, with the reference's domain (shape AABB grown by
 * tolerance + margin, mesh_sdf.cpp:28-45) and volume-map integrand
 * (16^3 Gauss-Legendre, volume_map.cpp:9-54).  box_size/box_pos receive the
 * grid domain.  Evaluated on the GPU in FP64, one grid node per thread, and
 * copied to the host arrays; `threads` is accepted for source compatibility
 * and ignored. */
int vfs_volume_map_generate(int shape_kind, const double *shape_params, const uint32_t *res,
                            double support_radius, double tolerance, double margin,
                            float *sdf_out, float *vm_out, float *box_size, float *box_pos,
                            int threads);

/* ---- triangle meshes as boundary objects -------------------------------
 * A mesh holds positions and triangles (host), their pseudonormals
 * (MeshPseudonormals::Build, src/util/mesh_pseudonormals.cpp:11-103) and a
 * BVH; signed distances (SignedDistanceToMesh, src/util/geometry.cpp:239-286:
 * closest point in double, sign from the pseudonormal of the closest feature)
 * are evaluated on the GPU.  The device copy is made on first GPU use. */
typedef struct vfs_mesh vfs_mesh;
/* LoadObjMesh (src/util/mesh_loader.cpp:20-113) as generic_scene.cpp:93-94
 * uses it: the last shape of the file, polygons triangulated. */
int vfs_mesh_load_obj(const char *path, vfs_mesh **out);
/* A mesh from positions (3 floats each) and triangles (3 indices each). */
int vfs_mesh_create(const float *vertices, uint32_t n_vertices, const uint32_t *triangles,
                    uint32_t n_triangles, vfs_mesh **out);
/* Counts and the box of the triangles (the BVH root box of mesh_sdf.cpp:33). */
int vfs_mesh_info(const vfs_mesh *mesh, uint32_t *n_vertices, uint32_t *n_triangles,
                  float *aabb_min, float *aabb_max);
/* closed: every edge shared by two triangles in opposite directions; acute:
 * no triangle angle above 90 degrees.  When both hold, the volume-map
 * generator skips the searches whose outcome the distance at the grid node
 * already fixes (the signed distance is then 1-Lipschitz); the grids are the
 * same either way.  No counterpart in the reference. */
int vfs_mesh_properties(const vfs_mesh *mesh, int *closed, int *acute);
/* Copies positions (3 * n_vertices floats) and triangles (3 * n_triangles). */
int vfs_mesh_get(const vfs_mesh *mesh, float *vertices, uint32_t *triangles);
/* SignedDistanceToMesh for n host points (3 floats each) into out (n doubles). */
int vfs_mesh_signed_distance(vfs_mesh *mesh, const float *points, uint32_t n, double *out);
/* MeshSDF::Build (src/util/mesh_sdf.cpp:28-45) + GenerateVolumeMap
 * (src/util/volume_map.cpp:9-54): SDF and volume-map grids
 * (res[0]*res[1]*res[2] floats each, host) over the mesh box grown by
 * tolerance + margin (the scene loader passes 0 and 8h); box_size/box_pos
 * receive the grid domain.  One grid node per GPU thread, FP64. */
int vfs_mesh_volume_map_generate(vfs_mesh *mesh, const uint32_t *res, double support_radius,
                                 double tolerance, double margin, float *sdf_out, float *vm_out,
                                 float *box_size, float *box_pos);
int vfs_mesh_destroy(vfs_mesh *mesh);

/* ---- trajectory export ----------------------------------------------------
 * The "record particle and object positions, and export to a file" item of
 * the reference's to-do list (README.md:33; no counterpart in its code).
 * File layout, little endian:
 *   header, 64 bytes: char magic[8] = "VFSTRAJ1"; uint32 version = 1;
 *     uint32 n_particles; uint32 fields (VFS_TRAJ_* bits); uint32 n_objects;
 *     uint64 n_frames (written by vfs_trajectory_close); float box_pos[3];
 *     float box_size[3]; uint32 reserved[2];
 *   n_objects x float[16]: each boundary object's world-to-object matrix
 *     (column-major, as vfs_boundary_object.transform);
 *   frames: uint64 substep; double time; then, for each field present in
 *     this order, positions float[3 n], velocities float[3 n], densities
 *     float[n], particle ids uint32[n] (particles in the simulation's current
 *     order; the ids map them to their initial index). */
enum {
    VFS_TRAJ_POSITIONS = 1u,
    VFS_TRAJ_VELOCITIES = 2u,
    VFS_TRAJ_DENSITIES = 4u,
    VFS_TRAJ_IDS = 8u /* needs VFS_FLAG_TRACK_IDS */
};
typedef struct vfs_trajectory vfs_trajectory;
int vfs_trajectory_open(vfs_sim *sim, const char *path, uint32_t fields, vfs_trajectory **out);
/* Appends the simulation's current state (after the work queued so far). */
int vfs_trajectory_record(vfs_trajectory *trajectory, uint64_t substep, double time);
/* Writes the frame count into the header and closes the file. */
int vfs_trajectory_close(vfs_trajectory *trajectory);

/* ---- render-buffer handoff ---------------------------------------------
 * Render buffers in the reference's data-buffer layout
 * (SPHModel::CreateDataBuffers, src/models/model.cpp:149-156): positions,
 * velocities and accelerations packed float3 (12 bytes per particle),
 * densities float; particles in the simulation's current order.  The
 * renderer reads them after vfs_sim_copy_data_buffers, in this process (its
 * stream waits on `ready`, a HIP event recorded after each copy) or, through
 * vfs_render_buffers_export / _open, in another one (zero copy: the other
 * process maps the same device memory; the producer calls
 * vfs_render_buffers_synchronize after the copy and then signals the renderer
 * over its own channel). */
typedef struct vfs_render_buffers {
    float *positions;
    float *velocities;
    float *densities;
    float *accelerations; /* NULL: accelerations are not handed over */
    void *ready;          /* hipEvent_t; NULL in opened buffers */
    uint32_t n_particles;
    uint32_t opened;      /* 1: mapped from another process by vfs_render_buffers_open */
} vfs_render_buffers;

/* Handles of a set of render buffers for another process (64-byte HIP IPC
 * handles of the four buffers); plain bytes, pass them over any channel.
 * Every buffer must be smaller than 2 GiB (n_particles below 178,956,971). */
typedef struct vfs_render_ipc {
    unsigned char buffers[4][64];
    uint32_t n_particles;
    uint32_t has_accelerations;
} vfs_render_ipc;

/* SPHModel::CreateDataBuffers (src/models/model.cpp:149-156), device memory
 * of the current device. */
int vfs_render_buffers_create(uint32_t n_particles, int with_accelerations, vfs_render_buffers *out);
/* Frees created buffers, or unmaps opened ones. */
int vfs_render_buffers_destroy(vfs_render_buffers *buffers);
/* SPHModel::CopyDataBuffers (src/models/model.cpp:158-170): the state after
 * the work queued so far, converted to the render layout, ordered on `stream`
 * like vfs_sim_step; then records buffers->ready.  Accelerations need
 * VFS_FLAG_STORE_ACCEL when buffers->accelerations is set. */
int vfs_sim_copy_data_buffers(vfs_sim *sim, const vfs_render_buffers *buffers, void *stream);
/* Handles of created buffers for a renderer in another process. */
int vfs_render_buffers_export(const vfs_render_buffers *buffers, vfs_render_ipc *out);
/* In the renderer's process: maps the buffers of `handles` (read them only). */
int vfs_render_buffers_open(const vfs_render_ipc *handles, vfs_render_buffers *out);
/* Makes `stream` wait for the last copy into created buffers. */
int vfs_render_buffers_wait(const vfs_render_buffers *buffers, void *stream);
/* Blocks until the last copy into created buffers has finished (no-op for
 * opened buffers). */
int vfs_render_buffers_synchronize(const vfs_render_buffers *buffers);
/* Host copies of the buffers (any pointer may be NULL), after the last copy. */
int vfs_render_buffers_read(const vfs_render_buffers *buffers, float *positions, float *velocities,
                            float *densities, float *accelerations);

/* ---- scenes (src/util/scene_loader.cpp:113-140, src/scenes/generic_scene.cpp) */
typedef struct vfs_scene vfs_scene;
/* Parses a scene JSON (reference schema + optional extension keys) and builds
 * the simulation with its boundary objects and seeded particles. */
int vfs_scene_load(const char *json_path, const char *resource_root, vfs_scene **out);
int vfs_scene_sim(vfs_scene *scene, vfs_sim **sim);
int vfs_scene_reset(vfs_scene *scene);
int vfs_scene_step(vfs_scene *scene, void *stream);
/* The scene's boundary objects as its simulation received them (the
 * BoundaryObjectInfo records of GenericScene::CreateBoundaryObjectBuffer,
 * generic_scene.cpp:125-145: world-to-object matrix inverse(T R S), rotation,
 * grid box and resolution; grid pointers refer to the scene's host grids,
 * valid until vfs_scene_destroy).  Writes min(capacity, count) records to
 * out (may be NULL) and the object count to count. */
int vfs_scene_boundary_objects(vfs_scene *scene, vfs_boundary_object *out, uint32_t capacity, uint32_t *count);
int vfs_scene_destroy(vfs_scene *scene);

#ifdef __cplusplus
}
#endif

#endif /* VFS_C_H */
