/*
 * se8.h - C-ABI of the MI355X-native single-resolution TSDF pipeline (libse8.so).
 *
 * Every entry point is extern "C", takes plain pointers, sizes and POD structs, and returns an
 * int status (SE8_OK = 0, negative on error).  No C++ or torch types cross this boundary.
 *
 * Memory convention: arguments named *_dev are device (HBM) pointers, everything else is host
 * memory.  Poses are row-major 4x4 float matrices (rigid transforms, last row 0 0 0 1).
 * Work is enqueued on the given HIP stream (NULL = the default stream); entry points that
 * return host-visible results (counts, poses) synchronise that stream before returning.
 *
 * Each declaration names the reference interface it replaces (paths relative to the
 * supereight2 source tree).  The header-only C++ facade in include/se8/supereight.hpp maps
 * the reference's templated C++ API onto these functions.
 */
#ifndef SE8_H
#define SE8_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SE8_OK 0
#define SE8_ERR_INVALID -1  /* bad argument (null pointer, size mismatch, unsupported config) */
#define SE8_ERR_OOM -2      /* device allocation failed */
#define SE8_ERR_HIP -3      /* a HIP runtime call failed */
#define SE8_ERR_CAPACITY -4 /* a fixed capacity is exhausted.  The block/node/brick pool (fixed at map
                             * creation): reported by the next integration, after which the map keeps
                             * the blocks that fit, stays readable and refuses further integration.
                             * se8_mesh: the caller's face buffer is too small (nothing written). */
#define SE8_ERR_IO -5       /* a file could not be written */

#define SE8_STORAGE_F32 0 /* float tsdf + float weight, 8 B/voxel (reference layout minus padding) */
#define SE8_STORAGE_F16 1 /* half tsdf + uint16 weight, 4 B/voxel */
#define SE8_STORAGE_MULTIRES 2 /* multi-resolution TSDF (se::TSDFMap<se::Res::Multi>, block.hpp:66-166):
                                * per block the data and the past data of scales 0..3 (512 + 64 + 8 + 1
                                * entries of float tsdf + float weight each) and its current and minimum
                                * scale; integrated by the multi-resolution updater */

typedef void* se8_stream; /* a hipStream_t */
typedef struct se8_map se8_map;
typedef struct se8_tracker se8_tracker;
typedef struct se8_lidar se8_lidar;
typedef struct se8_reader se8_reader;
typedef struct se8_frame_graph se8_frame_graph;

/* se::Map::Config (include/se/map/map.hpp:42-71).  T_MW is a pure translation (the reference's
 * raycaster assumes no rotation, include/se/map/impl/raycaster_impl.hpp:465); the default
 * t_MW is dim/2. */
typedef struct {
    float dim[3];
    float res;
    float t_MW[3];
} se8_map_config;

/* se::FieldData<Field::TSDF>::Config (include/se/map/data_field.hpp:76-78) plus the storage
 * precision and the block-pool capacity (blocks of 8^3 voxels) reserved up front. */
typedef struct {
    float truncation_boundary_factor; /* default 8 */
    float max_weight;                 /* default 100 */
    int storage;                      /* SE8_STORAGE_F32 or SE8_STORAGE_F16 */
    int64_t max_blocks;               /* <= 0: a default derived from the map size */
} se8_tsdf_config;

/* se::PinholeCamera::Config (include/se/sensor/pinhole_camera.hpp:24-40, sensor.hpp:30-45);
 * T_BS is folded into the poses passed by the caller. */
typedef struct {
    int width, height;
    float fx, fy, cx, cy;
    float near_plane, far_plane;
} se8_pinhole;

/* se::TrackerConfig (include/se/tracker/tracker.hpp:22-33).  iterations[0..num_levels) are the
 * Gauss-Newton iterations per pyramid level, finest first (the reference's std::vector
 * iterations, whose size is the number of levels, tracker_impl.hpp:50-84); 1 <= num_levels <=
 * SE8_TRACKER_MAX_LEVELS, and the coarsest level must keep at least one pixel in each direction
 * (se8_tracker_create returns SE8_ERR_INVALID otherwise).  compensated = 1 selects compensated
 * per-workgroup partials (fp32 sums within each wave, TwoSum sum-and-error pairs across a
 * workgroup's waves), an FP64 cross-workgroup reduction and an FP64 LDLT solve; 0 is the plain
 * fp32 reduction + fp32 Cholesky like the reference. */
#define SE8_TRACKER_MAX_LEVELS 6
typedef struct {
    int iterations[SE8_TRACKER_MAX_LEVELS];
    int num_levels;
    float dist_threshold;
    float normal_threshold;
    float track_threshold;
    float icp_threshold;
    int compensated;
} se8_tracker_config;

typedef struct {
    int octree_size;    /* voxels per side (power of two) */
    int max_scale;      /* log2(octree_size) */
    float res;
    float dim;          /* octree_size * res */
    int storage;
    int64_t num_blocks;
    int64_t num_nodes;  /* root included */
    int64_t max_blocks;
    int aabb_valid;
    int aabb_min[3], aabb_max[3]; /* voxel AABB of allocated blocks (octree_impl.hpp:267-273) */
} se8_map_info;

/* Per-stage device times (HIP events on the caller's stream), in milliseconds, using the
 * reference's stage names (map_integrator_impl.hpp:63-75, raycast_carver_impl.hpp:58-126,
 * main.cpp:101-217): the last completed call of each kind, and running sums over completed calls
 * since creation or se8_stage_times_reset.  A call's events are harvested when the next call of
 * the same kind starts (if they have completed by then) or by se8_stage_times_get. */
typedef struct {
    float fetch_frustum, create_list, allocate_list, update, raycast, tracking;
    int64_t last_frustum_blocks, last_new_blocks, last_candidates;
    double sum_fetch_frustum, sum_create_list, sum_allocate_list, sum_update, sum_raycast, sum_tracking;
    int64_t n_integrations, n_raycasts, n_tracks;
    int64_t sum_update_blocks; /* sum over integrations of (frustum + new) blocks updated */
    int64_t sum_candidates;
    /* SE8_TIMING_LEVELS only: per pyramid level (finest first), the summed device time from the
     * first Gauss-Newton launch of the level to the first launch of the next (the last level: to the
     * end of its last launch), and the number of launches it covers */
    double sum_track_level[SE8_TRACKER_MAX_LEVELS];
    int64_t n_track_level_launches[SE8_TRACKER_MAX_LEVELS];
} se8_stage_times;

/* -------- map lifetime: se::Map ctor/dtor (include/se/map/impl/map_impl.hpp:37-55) -------- */
int se8_default_configs(se8_map_config* map_cfg, se8_tsdf_config* tsdf_cfg, se8_pinhole* pinhole,
                        se8_tracker_config* tracker_cfg);
int se8_map_create(const se8_map_config* map_cfg, const se8_tsdf_config* tsdf_cfg, se8_map** out);
int se8_map_destroy(se8_map* map);
int se8_map_get_info(const se8_map* map, se8_map_info* out);

/* -------- integration: se::MapIntegrator::integrateDepth
 * (include/se/integrator/map_integrator.hpp:78-81, impl/map_integrator_impl.hpp:52-83):
 * RaycastCarver allocation + single-res Updater.  depth_dev is width*height floats, row-major. */
int se8_integrate_depth(se8_map* map, int timestamp, const float* depth_dev, const se8_pinhole* sensor,
                        const float T_WC[16], se8_stream stream);

/* se::OusterLidar::Config (include/se/sensor/ouster_lidar.hpp:24-44, sensor.hpp:30-45): one
 * elevation and one azimuth offset per beam (height entries, degrees), elevations strictly
 * descending (row 0 the highest beam).  The beam model (ouster_lidar.cpp:59-253) is evaluated by
 * the library; its direction table is built once at creation. */
typedef struct {
    int width, height;
    float near_plane, far_plane;
    const float* beam_elevation_deg;
    const float* beam_azimuth_deg;
} se8_lidar_config;
int se8_lidar_create(const se8_lidar_config* cfg, se8_lidar** out);
int se8_lidar_destroy(se8_lidar* lidar);
/* MapIntegrator::integrateDepth with an OusterLidar (map_integrator_impl.hpp:52-83; the sensor
 * hooks of ouster_lidar_impl.hpp:15-32): range_dev is width*height floats of range |p| (metres,
 * 0 = no return), row-major; T_WS is the sensor pose. */
int se8_integrate_range(se8_map* map, int timestamp, const float* range_dev, const se8_lidar* lidar,
                        const float T_WS[16], se8_stream stream);
/* The beam model's project, as the updater calls it (sensor.model.project,
 * singleres_tsdf_updater_impl.hpp:65; srl::projection::OusterLidar, not vendored), for n
 * sensor-frame points (host arrays, n*3 floats in, n*2 floats out): in_image[i] = 1 where the
 * projection succeeds; uv[i] holds the continuous image coordinates. */
int se8_lidar_project(const se8_lidar* lidar, const float* points, int64_t n, float* uv, uint8_t* in_image);

/* -------- raycast: se::raycaster::raycast_volume (include/se/map/raycaster.hpp:56-64,
 * impl/raycaster_impl.hpp:537-627).  points/normals are width*height*3 floats (x,y,z per pixel);
 * scale_dev (width*height int8) may be NULL; misses write point 0 and normal 0 and leave scale. */
int se8_raycast(const se8_map* map, const se8_pinhole* sensor, const float T_WS[16], float* points_dev,
                float* normals_dev, int8_t* scale_dev, se8_stream stream);

/* -------- tracking: se::Tracker (include/se/tracker/tracker.hpp:42-73,
 * impl/tracker_impl.hpp:17-116) -------- */
int se8_tracker_create(const se8_pinhole* sensor, const se8_tracker_config* cfg, se8_tracker** out);
int se8_tracker_destroy(se8_tracker* tracker);
/* Tracker::track(depth, T_WS, points_W, normals_W): T_WS is in/out; *ok receives the return
 * value; sums32 (may be NULL) receives the last reduction (icp.cpp:45-204 layout). */
int se8_track(se8_tracker* tracker, const float* depth_dev, float T_WS[16], const float* ref_points_dev,
              const float* ref_normals_dev, int* ok, float* sums32, se8_stream stream);
/* The tracker's input pyramid of the last track call at one level (0 = full resolution):
 * input_point_cloud_S and input_normals_S of tracker_impl.hpp:59-75 (depth_to_point_cloud of the
 * robustly half-sampled depth, point_cloud_to_normal), (W >> level) * (H >> level) * 3 floats each
 * into host arrays (either may be NULL).  Waits for the device. */
int se8_tracker_read_pyramid(const se8_tracker* tracker, int level, float* points, float* normals);
/* Tracker::track(depth, T_WS): raycasts map at T_WS first (tracker_impl.hpp:17-26). */
int se8_track_raycast(se8_tracker* tracker, const se8_map* map, const float* depth_dev, float T_WS[16], int* ok,
                      se8_stream stream);

/* -------- differentiation (growth pattern; no counterpart in the reference) --------
 * Batched point-to-plane cost and its adjoint for n candidate poses (poses16: n row-major 4x4,
 * host).  cost_k = 0.5 * sum e_i^2 over the inliers of trackKernel (icp_impl.hpp:16-102) against
 * the reference maps raycast at T_ref (full resolution, level 0); grad6_k = d cost / d xi for the
 * left twist T <- exp(xi) T, xi = [v; w] (math_util_impl.hpp:148-174), at xi = 0, which equals
 * -sum e_i J_i = -b of the reference's reduction (icp.cpp:45-204); jtj21_k (may be NULL) is J^T J
 * in the same upper-triangular layout; inliers_k (may be NULL) the inlier count. */
int se8_icp_cost_grad_batched(se8_tracker* tracker, const float* depth_dev, const float* ref_points_dev,
                              const float* ref_normals_dev, const float T_ref[16], int n_poses, const float* poses16,
                              float* cost, float* grad6, float* jtj21, int* inliers, se8_stream stream);
/* No counterpart in the reference.  se8_icp_cost_grad_batched for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops that forbid hipMalloc): the same
 * kernels and results (bit for bit), with the candidate poses (poses16_dev: n row-major 4x4) and
 * the results in device memory (cost_dev n, grad6_dev n*6, jtj21_dev n*21 or NULL, inliers_dev n
 * or NULL); T_ref is read on the host at call time.  Allocates nothing and never waits: the
 * per-workgroup partials live in a fixed array of the library (at most 64 poses per call;
 * n_poses above that returns SE8_ERR_INVALID), one per device, held until the process ends.
 * Every call on a device shares that array and the tracker's level-0 maps, so calls that overlap
 * on one device (other streams, other threads) race. */
int se8_icp_cost_grad_batched_noalloc(se8_tracker* tracker, const float* depth_dev, const float* ref_points_dev,
                                      const float* ref_normals_dev, const float T_ref[16], int n_poses,
                                      const float* poses16_dev, float* cost_dev, float* grad6_dev, float* jtj21_dev,
                                      int* inliers_dev, se8_stream stream);
/* Batched raycast depth and its adjoint for n poses (poses16 host, row-major 4x4): per pixel the
 * z-depth of the root of the trilinear TSDF along the ray, refined inside the bracket of the
 * reference march (raycaster_impl.hpp:456-533), and d depth / d xi (6 per pixel) by the
 * implicit-function theorem.  depth_dev n*H*W, grad_dev n*H*W*6, valid_dev n*H*W (0 where the
 * ray misses or the field is not interpolable at the root). */
int se8_raycast_depth_grad_batched(const se8_map* map, const se8_pinhole* sensor, int n_poses, const float* poses16,
                                   float* depth_dev, float* grad_dev, uint8_t* valid_dev, se8_stream stream);

/* -------- device-resident poses (no counterpart in the reference: its poses live on the host).
 * The same calls with the pose as 16 floats (row-major 4x4) in device memory, read by the first
 * kernel of the call, so that a pose one call produces on the GPU feeds the next without a host
 * round trip: se8_track_dev writes the tracked pose (and, if ok_dev is set, the check result) to
 * device memory and returns without waiting; T_in_dev may equal T_out_dev.  Results are identical
 * to the host-pose calls given the same pose. -------- */
int se8_integrate_depth_dev(se8_map* map, int timestamp, const float* depth_dev, const se8_pinhole* sensor,
                            const float* T_WC_dev, se8_stream stream);
int se8_integrate_range_dev(se8_map* map, int timestamp, const float* range_dev, const se8_lidar* lidar,
                            const float* T_WS_dev, se8_stream stream);
int se8_raycast_dev(const se8_map* map, const se8_pinhole* sensor, const float* T_WS_dev, float* points_dev,
                    float* normals_dev, int8_t* scale_dev, se8_stream stream);
int se8_track_dev(se8_tracker* tracker, const float* depth_dev, const float* T_in_dev, float* T_out_dev,
                  const float* ref_points_dev, const float* ref_normals_dev, int* ok_dev, se8_stream stream);
/* se8_raycast_dev, and in the same launch the input side of the tracker's next call -- the pyramid
 * of Tracker::track (tracker_impl.hpp:50-75) for next_depth_dev, starting from the pose at T_WS_dev
 * -- for the loop of app/src/main.cpp:98-162, whose next track starts from the pose just raycast
 * (no counterpart in the reference).  The next se8_track_dev on this tracker with depth_dev ==
 * next_depth_dev and T_in_dev == T_WS_dev skips its own pyramid; results are identical to
 * se8_raycast_dev followed by that se8_track_dev.  The depth image and the pose at T_WS_dev must
 * not change in between; any other pyramid of the tracker (another track, a frame graph replay, a
 * batched cost evaluation) discards the prepared one. */
/* se8_track_dev, whose last launch also starts the next integration of map at the tracked pose (the
 * per-frame state that integrateDepth's first step sets up, map_integrator_impl.hpp:52-83): the next
 * se8_integrate_depth_dev on map with T_WC_dev == T_out_dev skips that launch; results are
 * identical to se8_track_dev followed by that integration (no counterpart in the reference).  Any
 * other integration of map discards the prepared start. */
int se8_track_dev_prepare_integrate(se8_tracker* tracker, const float* depth_dev, const float* T_in_dev,
                                    float* T_out_dev, const float* ref_points_dev, const float* ref_normals_dev,
                                    int* ok_dev, se8_map* map, se8_stream stream);
int se8_raycast_dev_prepare_track(const se8_map* map, const se8_pinhole* sensor, const float* T_WS_dev,
                                  float* points_dev, float* normals_dev, int8_t* scale_dev, se8_tracker* tracker,
                                  const float* next_depth_dev, se8_stream stream);

/* -------- frame graphs (no counterpart in the reference; for callers that run the loop of
 * app/src/main.cpp:98-162 at device speed).  One frame of that loop -- Tracker::track from the
 * previous pose against the previous raycast (tracker_impl.hpp:17-116), integrateDepth at the
 * tracked pose (map_integrator_impl.hpp:52-83), raycast_volume at it (raycaster_impl.hpp:537-627)
 * -- captured once as a HIP graph of kernel nodes and replayed per frame: results are identical to
 * se8_track_dev + se8_integrate_depth_dev + se8_raycast_dev with the same arguments.
 * se8_frame_graph_create captures it for one tracker, map (at most 4096 voxels a side,
 * single-resolution) and sensor, with stage timing off, and two vertex/normal map pairs (W*H*3
 * floats each): launch i (counting from 0) tracks against pair (i + 1) % 2, then raycasts into pair
 * i % 2, so the first launch's reference is pair 1 (raycast it with se8_raycast_dev first).
 * se8_frame_graph_launch replays one frame on the stream: depth_dev is W*H floats, T_in_dev the
 * previous pose and T_out_dev the tracked pose (16 floats each, device memory, may be equal),
 * ok_dev (may be NULL) the pose check.  The host returns without waiting for this frame; it waits
 * for the frame two launches back before reusing its graph.  Other calls on the same map and
 * tracker may be interleaved in stream order. -------- */
int se8_frame_graph_create(se8_tracker* tracker, se8_map* map, const se8_pinhole* sensor, float* points0_dev,
                           float* normals0_dev, float* points1_dev, float* normals1_dev, se8_frame_graph** out);
int se8_frame_graph_launch(se8_frame_graph* graph, int timestamp, const float* depth_dev, const float* T_in_dev,
                           float* T_out_dev, int* ok_dev, se8_stream stream);
int se8_frame_graph_destroy(se8_frame_graph* graph);

/* -------- map queries (host arrays) -------- */
/* se::allocator::blocks(voxel_coords, octree, root, only_allocated)
 * (include/se/map/octree/impl/allocator_impl.hpp:40-116): voxel_coords is n*3 ints; the block
 * coords of the returned blocks go to out_block_coords (n*3 ints, sorted by Morton key). */
int se8_allocate_blocks(se8_map* map, const int* voxel_coords, int64_t n, int only_new, int* out_block_coords,
                        int64_t* out_n);
/* Write/read the 512 voxels (x fastest, BlockSingleRes::data, block_impl.hpp:45-65) of one block. */
int se8_write_block(se8_map* map, const int block_coord[3], const float* tsdf512, const float* weight512);
int se8_read_block(const se8_map* map, const int block_coord[3], float* tsdf512, float* weight512);
/* The blocks the last integration updated (raycast_carver_impl.hpp:55-128: the frustum blocks and
 * the blocks it allocated), the octants integrateDepth reports in updated_octants
 * (map_integrator.hpp:78-81, impl/map_integrator_impl.hpp:77-80), as keys (code<<5|3) in no
 * particular order.  keys NULL only counts; cap < the count returns SE8_ERR_CAPACITY with *out_n
 * set.  Waits for the device. */
int se8_read_update_list(const se8_map* map, uint64_t* keys, int64_t cap, int64_t* out_n);
/* All allocated blocks sorted by Morton key (key_util_impl.hpp:74-101 layout code<<5|3), with
 * their timestamps; tsdf/weight (512 per block) may be NULL.  cap = capacity of the arrays. */
int se8_dump_blocks(const se8_map* map, uint64_t* keys, int* timestamps, float* tsdf, float* weight, int64_t cap,
                    int64_t* out_n);
/* The blocks of a multi-resolution map (SE8_STORAGE_MULTIRES) sorted by key: their current and
 * minimum scale (2 ints per block, -1 = none), data and past data (585 entries of {tsdf, weight} per
 * block: scale 0 (512, x fastest), 1 (64), 2 (8), 3 (1); block.hpp:66-166); any output may be NULL. */
int se8_dump_blocks_multires(const se8_map* map, uint64_t* keys, int* scales, float* data, float* past_data, int64_t cap,
                             int64_t* out_n);
/* All octree nodes (root included) sorted by key, with child masks (octant.hpp:19-72) and
 * timestamps (propagator_impl.hpp:146-197). */
int se8_dump_nodes(const se8_map* map, uint64_t* keys, uint8_t* child_masks, int* timestamps, int64_t cap,
                   int64_t* out_n);
/* Map::getData / getFieldInterp / getFieldGrad (map_impl.hpp:68-86, 154-172, 239-268) for n
 * world points (n*3 floats).  valid[i] = 0 where the reference returns std::nullopt/invalid. */
int se8_get_data(const se8_map* map, const float* points_W, int64_t n, float* tsdf, float* weight);
int se8_field_interp(const se8_map* map, const float* points_W, int64_t n, float* out, uint8_t* valid);
int se8_field_grad(const se8_map* map, const float* points_W, int64_t n, float* out3, uint8_t* valid);

/* -------- surface mesh: se::Map::mesh / saveMesh (include/se/map/map.hpp:396-413,
 * impl/map_impl.hpp:452-469) through the single-resolution marching cubes
 * (include/se/map/algorithms/impl/marching_cube_impl.hpp:832-889, 1175-1203).
 * se8_mesh writes each face as 9 floats (its three vertices x,y,z in frame O, metres; T_OW NULL =
 * identity) into faces_dev (device memory for max_faces faces) and the number of faces to
 * *num_faces.  max_faces = 0 only counts.  A mesh larger than max_faces writes nothing and returns
 * SE8_ERR_CAPACITY with *num_faces set.  The call waits for the stream (the count is read back);
 * the faces are then written asynchronously on it.  Face order: block slot, then cell, then the
 * case's triangle order (fixed for a given map; the reference's order follows its OpenMP
 * schedule).  The reference's min_desired_scale is unused for single-resolution maps. */
int se8_mesh(const se8_map* map, const float T_OW[16], float* faces_dev, int64_t max_faces, int64_t* num_faces,
             se8_stream stream);
/* se::io::save_mesh (include/se/map/io/mesh_io.hpp:44-46, impl/mesh_io_impl.hpp:14-230): writes n
 * faces of nv = 3 (surface) or 4 (octree structure) vertices (host memory, 3*nv floats per face)
 * as a .ply, .vtk or .obj file chosen by the extension, every vertex mapped by T_OM (NULL =
 * identity), each face coloured by its scale (scales: n int8, NULL = all 0; colour_utils.hpp:23-30).
 * SE8_ERR_INVALID for another extension (the reference throws std::invalid_argument), SE8_ERR_IO if
 * the file cannot be written. */
int se8_save_mesh(const float* vertices, int64_t n, int nv, const int8_t* scales, const char* filename,
                  const float T_OM[16]);
/* se::io::has_supported_mesh_extension (include/se/map/io/mesh_io.hpp:24-29): 1 for .ply/.vtk/.obj */
int se8_has_supported_mesh_extension(const char* filename);

/* -------- debug views -------- */
/* Map::saveFieldSlices (include/se/map/impl/map_impl.hpp:272-308) with io::save_3d_slice_vtk
 * (include/se/map/io/impl/octree_io_impl.hpp:19-80): the tsdf (get_field of getData; 1 where no
 * block is allocated) over the octree's three axis-aligned planes through the voxel of point_W, each
 * as a VTK rectilinear grid; a NULL or empty filename skips that plane.  SE8_ERR_IO if a file cannot
 * be written (the reference prints an error and goes on). */
int se8_save_field_slices(const se8_map* map, const char* filename_x, const char* filename_y, const char* filename_z,
                          const float point_W[3]);
/* One plane of the above on the device: axis 0/1/2 at voxel coordinate index (outside the octree:
 * all 1); out_dev holds size*size floats in save_3d_slice_vtk's order (the slower free coordinate
 * outer).  Asynchronous on the stream. */
int se8_field_slice(const se8_map* map, int axis, int index, float* out_dev, se8_stream stream);
/* Map::structure (map_impl.hpp:473-479) -> octree_structure_mesh
 * (include/se/map/algorithms/impl/structure_meshing_impl.hpp:15-60): six quads per octant (every node
 * and block, or with only_leaves the octants without children), 12 floats per quad (4 vertices in
 * frame O, metres; T_OW NULL = identity) into quads_dev and the face scale (blocks 0, nodes 7) into
 * scales_dev (may be NULL), both device memory for max_quads quads; *num_quads = the count, and
 * max_quads = 0 only counts.  Too small a buffer: SE8_ERR_CAPACITY, nothing written.  Order: nodes,
 * then blocks (the reference's follows its octree iterator). */
int se8_structure(const se8_map* map, const float T_OW[16], int only_leaves, float* quads_dev, int8_t* scales_dev,
                  int64_t max_quads, int64_t* num_quads, se8_stream stream);

/* -------- input side of the frame loop (app/src/main.cpp:104-120) -------- */
/* se::preprocessor::downsample_depth (src/map/preprocessor.cpp:19-84): each output pixel is the
 * median of its ratio x ratio input region's depths >= 1e-5 (region[n / 2] of the ascending order),
 * 0 where there are none; the input must be the same integer multiple of the output in both
 * directions.  map_dev (NULL, or w_out*h_out uint64) receives the input pixel index x + y*w_in the
 * median came from (the region's first pixel where none is valid; among equal depths the first in
 * the region's row-major order).  Device buffers, asynchronous on the stream. */
int se8_downsample_depth(const float* in_dev, int w_in, int h_in, float* out_dev, int w_out, int h_out,
                         uint64_t* map_dev, se8_stream stream);
/* The same over depth as a reader stores it: raw uint16 converted to metres first as the readers
 * do, raw / scale (divide = 1: RAW millimetres, scale 1000, reader_raw.cpp:116-121) or raw * scale
 * (divide = 0: TUM inverse scale, reader_tum.cpp nextDepth).  w_out = w_in only converts. */
int se8_downsample_depth_u16(const uint16_t* in_dev, int w_in, int h_in, float scale, int divide, float* out_dev,
                             int w_out, int h_out, uint64_t* map_dev, se8_stream stream);

/* Dataset readers (app/include/reader_base.hpp:24-62, reader_raw.hpp, reader_tum.hpp; host code):
 * the RAW format (one file: per frame [u32 w][u32 h][w*h u16 depth, mm][u32 w][u32 h][w*h RGB8]) and
 * the TUM RGB-D directory (depth.txt, rgb.txt, depth/ 16-bit PNGs; colour is not read).  Ground
 * truth: TUM groundtruth.txt interpolated at the depth timestamps (TUM), or one line per frame
 * whose last 7 columns are tx ty tz qx qy qz qw (RAW). */
#define SE8_READER_RAW 0
#define SE8_READER_TUM 1
#define SE8_READER_OK 0    /* se::ReaderStatus::ok */
#define SE8_READER_SKIP 1  /* ::skip (invalid pose of this frame) */
#define SE8_READER_EOF 2   /* ::eof */
#define SE8_READER_ERROR 3 /* ::error */
typedef struct {
    int type;                      /* SE8_READER_RAW or SE8_READER_TUM */
    const char* sequence_path;     /* the .raw file, or the TUM directory */
    const char* ground_truth_file; /* NULL or "" for none */
    float inverse_scale;           /* TUM metres per stored unit; 0 = 1/5000 (reader_tum.hpp:46) */
} se8_reader_config;
/* SE8_ERR_IO (with the reason in se8_last_error) where the reference's reader would start in the
 * error status. */
int se8_reader_create(const se8_reader_config* cfg, se8_reader** out);
int se8_reader_destroy(se8_reader* reader);
int se8_reader_info(const se8_reader* reader, int* width, int* height, int64_t* num_frames, int* has_ground_truth);
/* Reader::nextData (reader_base.cpp:536-572): the next frame's depth as stored (width*height uint16,
 * host) with the conversion to metres (*scale, *divide as in se8_downsample_depth_u16) and, if T_WB
 * is not NULL, its ground-truth pose; *status is the merged reader status. */
int se8_reader_next_raw(se8_reader* reader, uint16_t* depth, float* scale, int* divide, float T_WB[16], int* status);
/* The same with the depth converted to metres on the host, as the reference returns it. */
int se8_reader_next(se8_reader* reader, float* depth, float T_WB[16], int* status);
int se8_reader_restart(se8_reader* reader);

/* Stage timing of a map and/or a tracker (either may be NULL).  The reference's TICK/TOCK
 * (perfstats_impl.hpp:352-388) are always on; here they are a mode:
 *   SE8_TIMING_STAGES (the default): HIP events around each stage; a call waits on the host for the
 *     previous call of its kind to finish (its events are read then), so the host runs at most about
 *     one frame ahead.
 *   SE8_TIMING_LEVELS: also events between the pyramid levels of each track call
 *     (se8_stage_times.sum_track_level).
 *   SE8_TIMING_OFF: no events and no host waits inside integrate / raycast / track: the *_dev calls
 *     of a frame only enqueue kernels (they may be captured in a HIP graph once the map's and
 *     tracker's buffers exist), and the host runs ahead freely.  A capacity overflow is then reported
 *     by se8_map_get_info, and by the first integration after timing is switched back on.
 * Switching waits for the stage events still pending. */
#define SE8_TIMING_OFF 0
#define SE8_TIMING_STAGES 1
#define SE8_TIMING_LEVELS 2
int se8_set_stage_timing(se8_map* map, se8_tracker* tracker, int mode);
int se8_stage_times_get(const se8_map* map, const se8_tracker* tracker, se8_stage_times* out);
int se8_stage_times_reset(se8_map* map, se8_tracker* tracker);

/* -------- device staging for host-image callers (the reference API takes host se::Image<T>,
 * include/se/image/image.hpp:18-111); the C++ facade stages through these. -------- */
#define SE8_COPY_H2D 0
#define SE8_COPY_D2H 1
#define SE8_COPY_D2D 2
int se8_device_alloc(size_t bytes, void** out_dev);
int se8_device_free(void* dev);
int se8_copy(void* dst, const void* src, size_t bytes, int direction, se8_stream stream);
int se8_stream_synchronize(se8_stream stream);

/* Human-readable text of the last error on this thread. */
const char* se8_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SE8_H */
