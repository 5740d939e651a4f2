// Map object, allocation pipeline and TSDF update of libse8.so.
//
// integrateDepth (include/se/integrator/impl/map_integrator_impl.hpp:52-83) becomes, per frame:
//   allocation     allocate_frame (se8_alloc.hpp): fetch-frustum, create-list and allocate-list
//   k_update       Updater<Map<TSDF,Single>> (singleres_tsdf_updater_impl.hpp:14-132), one 8^3
//                  block per 512-thread workgroup
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "se8_alloc.hpp"
#include "se8_internal.hpp"

namespace se8 {

static thread_local std::string g_last_error;

void set_error(const std::string& msg)
{
    g_last_error = msg;
}

int dev_alloc(DeviceBuffer& b, size_t bytes)
{
    if (b.ptr && b.bytes >= bytes)
        return SE8_OK;
    dev_free(b);
    if (bytes == 0)
        bytes = 16;
    if (hipMalloc(&b.ptr, bytes) != hipSuccess) {
        b.ptr = nullptr;
        b.bytes = 0;
        (void) hipGetLastError();
        set_error("hipMalloc of " + std::to_string(bytes) + " bytes failed");
        return SE8_ERR_OOM;
    }
    b.bytes = bytes;
    return SE8_OK;
}

void dev_free(DeviceBuffer& b)
{
    if (b.ptr)
        (void) hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

// ----------------------------------------------------------------------------------------------
// small utility kernels
// ----------------------------------------------------------------------------------------------
__global__ void k_fill_i32(int* p, int64_t n, int v)
{
    for (int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_publish_counters(const int* counters, int* host)
{
    if (threadIdx.x < kNumCounters)
        host[threadIdx.x] = counters[threadIdx.x];
}

static int fill_i32(DeviceBuffer& b, int64_t n, int v, hipStream_t s)
{
    if (n <= 0)
        return SE8_OK;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    k_fill_i32<<<(unsigned) blocks, 256, 0, s>>>((int*) b.ptr, n, v);
    SE8_HIP_CHECK(hipGetLastError());
    return SE8_OK;
}

// ----------------------------------------------------------------------------------------------
// update: one block (512 voxels) per workgroup of 8 wave64s
// ----------------------------------------------------------------------------------------------
// Raw voxel word of a storage type: float2 {tsdf, weight} (FP32) or packed half + uint16 (FP16).
template<int Storage>
struct VoxelRaw {
    float2 v;
    __device__ __forceinline__ void load(const MapView& m, size_t vi) { v = reinterpret_cast<const float2*>(m.voxels)[vi]; }
    __device__ __forceinline__ void clear() { v = make_float2(1.f, 0.f); }
    __device__ __forceinline__ void get(float& tsdf, float& weight) const
    {
        tsdf = v.x;
        weight = v.y;
    }
    __device__ __forceinline__ static void store(const MapView& m, size_t vi, float tsdf, float weight)
    {
        reinterpret_cast<float2*>(m.voxels)[vi] = make_float2(tsdf, weight);
    }
};

template<>
struct VoxelRaw<SE8_STORAGE_F16> {
    uint32_t v;
    __device__ __forceinline__ void load(const MapView& m, size_t vi) { v = reinterpret_cast<const uint32_t*>(m.voxels)[vi]; }
    __device__ __forceinline__ void clear() { v = pack_f16(1.f, 0.f); }
    __device__ __forceinline__ void get(float& tsdf, float& weight) const { unpack_f16(v, tsdf, weight); }
    __device__ __forceinline__ static void store(const MapView& m, size_t vi, float tsdf, float weight)
    {
        reinterpret_cast<uint32_t*>(m.voxels)[vi] = pack_f16(tsdf, weight);
    }
};

constexpr int kUpdChunk = 512; // blocks of a workgroup whose list entries are staged in LDS at once
constexpr int kUpdK = 2;       // blocks a workgroup updates together (their loads in flight together)

// What the update of one voxel needs from its position: the measurement along the sensor axis
// (z for the pinhole, |p| for the lidar), the distance |p| and the depth sample at its pixel (ok =
// the voxel is within range and projects into the image).
struct VoxelSample {
    float mz, nrm, d;
    bool ok;
};

// Per voxel as singleres_tsdf_updater_impl.hpp:51-71: the voxel centre in the sensor frame, the far
// test, the projection and the depth sample (issued, not awaited).  The load is issued on every lane
// (pixel 0 where the voxel does not project) so that the pipeline's loads are straight-line code and
// each wait covers only the loads it needs; `live` = false marks a lane with no block.
template<class Sensor>
__device__ __forceinline__ VoxelSample sample_voxel(const Iso& T_CV, const Sensor& sensor, const float* depth,
                                                    int4 c, int vx, int vy, int vz, bool live)
{
    const float v[3] = {(float) (c.x + vx), (float) (c.y + vy), (float) (c.z + vz)};
    float pC[3];
    iso_apply(T_CV, v, pC);
    VoxelSample o;
    o.nrm = Sensor::kRejectsNearOrigin ? sqrt_rn_large(sqnorm3(pC)) : sqrtf(sqnorm3(pC));
    o.mz = sensor.measurement(pC);
    float uv[2] = {0.f, 0.f};
    o.ok = live && !sensor.beyond_far(pC, o.nrm) && sensor.project(pC, uv);
    const int px = (int) (uv[0] + 0.5f), py = (int) (uv[1] + 0.5f);
    o.d = depth[o.ok ? px + (size_t) py * sensor.width() : 0];
    return o;
}

// Per voxel as singleres_tsdf_updater_impl.hpp:44-128.  A workgroup owns the list entries
// blockIdx.x, blockIdx.x + gridDim.x, ...: it stages up to kUpdChunk of them (entry, block
// coordinates, each key decoded once by the thread that loads it) in LDS with one round of loads
// (stamping the block and parent timestamps at the same time), then walks
// them kUpdK blocks at a time as a two-stage pipeline: while a group is updated, the next group's
// voxels and depth samples are already loading (its projections are computed first), so neither
// load waits on the arithmetic it feeds.
// amdgpu_waves_per_eu(8): the scalar registers must stay within the budget of 8 waves per SIMD
// (96 allocated), or the one resident wave of workgroups shrinks from 4 to 3 per CU.
// With link set (maps with the dense tables), the kernel first links the octants the carver
// created (link_part, which also stamps the new blocks' parents); nothing here reads those links,
// so no workgroup waits for another.
template<int Storage, class Sensor>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_update(
    MapView m, const int* update_list, const FramePoses* fp, Sensor sensor, const float* depth, int ts, int link)
{
    if (link)
        link_part(m, update_list, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, ts);
    __shared__ int s_entry[kUpdChunk];
    __shared__ int4 s_coord[kUpdChunk];
    const Iso T_CV = fp->T_CV;
    const int n = m.counters[kCntUpdate];
    const int t = threadIdx.x;
    const int vx = t & 7, vy = (t >> 3) & 7, vz = t >> 6;
    const int G = gridDim.x;
    const int iters = n > (int) blockIdx.x ? (n - 1 - (int) blockIdx.x) / G + 1 : 0;
    for (int c0 = 0; c0 < iters; c0 += kUpdChunk) {
        const int cnt = min(kUpdChunk, iters - c0);
        if (c0 > 0)
            __syncthreads(); // the previous chunk's entries are no longer read
        if (t < cnt) {
            const int entry = update_list[blockIdx.x + (size_t) (c0 + t) * G];
            const int slot = entry & 0x7fffffff;
            s_entry[t] = entry;
            int c[3];
            key_to_coord(m.block_key[slot], c);
            s_coord[t] = make_int4(c[0], c[1], c[2], 0);
            // block timestamp and one-level parent propagation (propagator_impl.hpp:146-197); a
            // block new this frame has its parent stamped by link_part when the links are made here
            m.block_ts[slot] = ts;
            if (!(link && entry < 0))
                atomicMax(&m.node_ts[m.block_parent[slot]], ts);
        }
        __syncthreads();
        // two register sets take turns: while the group in one is updated, the next group's voxels
        // and depth samples load into the other (the loop body runs twice with the roles swapped,
        // so no registers are copied between iterations)
        // every lane issues every load (block slot 0 and pixel 0 stand in where there is nothing to
        // read; update_group ignores those values), so the waits count exactly the loads they need
        auto load_group = [&](int j, VoxelRaw<Storage> v[kUpdK], VoxelSample smp[kUpdK]) {
#pragma unroll
            for (int k = 0; k < kUpdK; k++) {
                const int e = s_entry[min(j + k, cnt - 1)];
                v[k].load(m, (size_t) (j + k < cnt && e >= 0 ? e : 0) * 512 + t);
            }
#pragma unroll
            for (int k = 0; k < kUpdK; k++)
                smp[k] = sample_voxel(T_CV, sensor, depth, s_coord[min(j + k, cnt - 1)], vx, vy, vz, j + k < cnt);
        };
        auto update_group = [&](int i, VoxelRaw<Storage> v[kUpdK], const VoxelSample smp[kUpdK]) {
#pragma unroll
            for (int k = 0; k < kUpdK; k++) {
                if (i + k >= cnt)
                    break;
                const int entry = s_entry[i + k];
                // a block new this frame starts from FieldData<TSDF>{} (its slot was not read)
                if (entry < 0)
                    v[k].clear();
                float tsdf, weight;
                v[k].get(tsdf, weight);
                const bool changed = tsdf_update_voxel(m, smp[k].d, smp[k].ok, sensor.near_plane(), smp[k].mz,
                                                       smp[k].nrm, tsdf, weight);
                if (changed || entry < 0)
                    VoxelRaw<Storage>::store(m, (size_t) (entry & 0x7fffffff) * 512 + t, tsdf, weight);
            }
        };
        VoxelRaw<Storage> va[kUpdK], vb[kUpdK];
        VoxelSample sa[kUpdK], sb[kUpdK];
        load_group(0, va, sa);
        for (int i = 0; i < cnt; i += 2 * kUpdK) {
            load_group(i + kUpdK, vb, sb);
            update_group(i, va, sa);
            if (i + kUpdK >= cnt)
                break;
            load_group(i + 2 * kUpdK, va, sa);
            update_group(i + kUpdK, vb, sb);
        }
    }
}

static int read_counters(const se8_map* map, int* host, hipStream_t s)
{
    SE8_HIP_CHECK(hipMemcpyAsync(host, map->counters.ptr, sizeof(int) * kNumCounters, hipMemcpyDeviceToHost, s));
    SE8_HIP_CHECK(hipStreamSynchronize(s));
    return SE8_OK;
}

// Compute units of the current device (cached).
int device_count_sm()
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        (void) hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    return cus;
}

// Workgroups of `threads` threads of kernel k that fit on the device at once.
template<class K>
static int resident_grid(K k, int threads)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, threads, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    return per_cu * device_count_sm();
}


} // namespace se8

using namespace se8;

// ==============================================================================================
// C-ABI
// ==============================================================================================
extern "C" const char* se8_last_error(void)
{
    return g_last_error.c_str();
}

extern "C" int se8_default_configs(se8_map_config* map_cfg, se8_tsdf_config* tsdf_cfg, se8_pinhole* pinhole,
                                   se8_tracker_config* tracker_cfg)
{
    if (map_cfg) {
        // Map::Config defaults (map.hpp:45-53)
        for (int i = 0; i < 3; i++) {
            map_cfg->dim[i] = 10.f;
            map_cfg->t_MW[i] = 5.f;
        }
        map_cfg->res = 0.1f;
    }
    if (tsdf_cfg) {
        tsdf_cfg->truncation_boundary_factor = 8.f;
        tsdf_cfg->max_weight = 100.f;
        tsdf_cfg->storage = SE8_STORAGE_F32;
        tsdf_cfg->max_blocks = 0;
    }
    if (pinhole) {
        // SensorBase::Config defaults (sensor.hpp:30-45); intrinsics have no defaults upstream
        pinhole->width = 640;
        pinhole->height = 480;
        pinhole->fx = pinhole->fy = 525.f;
        pinhole->cx = 319.5f;
        pinhole->cy = 239.5f;
        pinhole->near_plane = 0.01f;
        pinhole->far_plane = 10.f;
    }
    if (tracker_cfg) {
        // TrackerConfig defaults (tracker.hpp:22-27)
        for (int l = 0; l < SE8_TRACKER_MAX_LEVELS; l++)
            tracker_cfg->iterations[l] = 0;
        tracker_cfg->iterations[0] = 10;
        tracker_cfg->iterations[1] = 5;
        tracker_cfg->iterations[2] = 4;
        tracker_cfg->num_levels = 3;
        tracker_cfg->dist_threshold = 0.1f;
        tracker_cfg->normal_threshold = 0.8f;
        tracker_cfg->track_threshold = 0.15f;
        tracker_cfg->icp_threshold = 0.00001f;
        tracker_cfg->compensated = 0;
    }
    return SE8_OK;
}

extern "C" int se8_map_create(const se8_map_config* cfg, const se8_tsdf_config* tcfg, se8_map** out)
{
    SE8_REQUIRE(cfg && tcfg && out, "se8_map_create: null argument");
    SE8_REQUIRE(cfg->res > 0.f, "se8_map_create: res must be positive");
    SE8_REQUIRE(tcfg->storage == SE8_STORAGE_F32 || tcfg->storage == SE8_STORAGE_F16 ||
                    tcfg->storage == SE8_STORAGE_MULTIRES,
                "se8_map_create: unknown storage");
    SE8_REQUIRE(tcfg->truncation_boundary_factor > 0.f && tcfg->max_weight >= 1.f,
                "se8_map_create: invalid TSDF config");
    SE8_REQUIRE(tcfg->storage != SE8_STORAGE_F16 || tcfg->max_weight <= 65535.f,
                "se8_map_create: max_weight does not fit uint16 weights");
    *out = nullptr;
    se8_map* m = new se8_map();
    m->cfg = *cfg;
    m->tsdf = *tcfg;
    MapView& v = m->view;
    int64_t node_dense_total = 0;
    if (!plan_map_view(*cfg, *tcfg, v, node_dense_total)) {
        delete m;
        set_error("se8_map_create: octree deeper than the 19-scale key limit");
        return SE8_ERR_INVALID;
    }
    const int p = v.max_scale, size = v.size;
    const int brick_scale = v.brick_shift + kBlockScale;
    const int64_t max_blocks = v.max_blocks;
    const int64_t bricks_side3 = (int64_t) v.bricks_side * v.bricks_side * v.bricks_side;
    const size_t block_bytes = tcfg->storage == SE8_STORAGE_F32 ? 8 * 512
                               : tcfg->storage == SE8_STORAGE_F16 ? 4 * 512 : 8 * (size_t) kMrBlockFloat2;
    int rc = SE8_OK;
    rc = rc ? rc : dev_alloc(m->brick_top, sizeof(int) * bricks_side3);
    rc = rc ? rc : dev_alloc(m->brick_slots, sizeof(int) * 512 * v.max_bricks);
    // Dense tables (maps of at most kDenseMaxSize voxels a side): one block slot per block position
    // (a lookup is one load) and one node index per octant of every scale in [4, max_scale)
    // (allocation claims octants in place).  They are an acceleration only: if either allocation
    // fails, the map uses the brick index and the sorted-key allocation instead.
    const int64_t dense_side = size / 8;
    if (rc == SE8_OK && size <= kDenseMaxSize) {
        if (dev_alloc(m->block_dense, sizeof(int) * dense_side * dense_side * dense_side) != SE8_OK
            || dev_alloc(m->node_dense, sizeof(int) * std::max<int64_t>(node_dense_total, 1)) != SE8_OK) {
            dev_free(m->block_dense);
            dev_free(m->node_dense);
            set_error("");
        }
    }
    rc = rc ? rc : dev_alloc(m->node_child, sizeof(int) * 8 * v.max_nodes);
    rc = rc ? rc : dev_alloc(m->node_key, sizeof(uint64_t) * v.max_nodes);
    rc = rc ? rc : dev_alloc(m->node_ts, sizeof(int) * v.max_nodes);
    rc = rc ? rc : dev_alloc(m->block_key, sizeof(uint64_t) * max_blocks);
    rc = rc ? rc : dev_alloc(m->block_parent, sizeof(int) * max_blocks);
    rc = rc ? rc : dev_alloc(m->block_ts, sizeof(int) * max_blocks);
    rc = rc ? rc : dev_alloc(m->voxels, block_bytes * (size_t) max_blocks);
    if (tcfg->storage == SE8_STORAGE_MULTIRES)
        rc = rc ? rc : dev_alloc(m->block_scale, sizeof(int) * 2 * (size_t) max_blocks);
    rc = rc ? rc : dev_alloc(m->counters, sizeof(int) * kNumCounters);
    rc = rc ? rc : dev_alloc(m->update_list, sizeof(int) * max_blocks);
    rc = rc ? rc : dev_alloc(m->frame_poses, sizeof(FramePoses));
    if (rc != SE8_OK) {
        se8_map_destroy(m);
        return rc;
    }
    v.brick_top = (int*) m->brick_top.ptr;
    v.brick_slots = (int*) m->brick_slots.ptr;
    v.block_dense = (int*) m->block_dense.ptr;
    v.node_dense = (int*) m->node_dense.ptr;
    v.node_child = (int*) m->node_child.ptr;
    v.node_key = (uint64_t*) m->node_key.ptr;
    v.node_ts = (int*) m->node_ts.ptr;
    v.block_key = (uint64_t*) m->block_key.ptr;
    v.block_parent = (int*) m->block_parent.ptr;
    v.block_ts = (int*) m->block_ts.ptr;
    v.voxels = m->voxels.ptr;
    v.counters = (int*) m->counters.ptr;
    v.block_scale = (int*) m->block_scale.ptr;
    hipStream_t s = nullptr;
    rc = rc ? rc : fill_i32(m->brick_top, bricks_side3, -1, s);
    rc = rc ? rc : fill_i32(m->brick_slots, 512 * v.max_bricks, -1, s);
    if (m->block_dense.ptr) {
        rc = rc ? rc : fill_i32(m->block_dense, dense_side * dense_side * dense_side, -1, s);
        rc = rc ? rc : fill_i32(m->node_dense, std::max<int64_t>(node_dense_total, 1), -1, s);
    }
    rc = rc ? rc : fill_i32(m->node_child, 8 * v.max_nodes, -1, s);
    if (rc != SE8_OK) {
        se8_map_destroy(m);
        return rc;
    }
    int host[kNumCounters];
    std::memset(host, 0, sizeof(host));
    host[kCntBlocks] = 0;
    host[kCntNodes] = 1; // the root
    host[kCntBricks] = 0;
    host[kCntAabb + 0] = host[kCntAabb + 1] = host[kCntAabb + 2] = 0x7fffffff;
    host[kCntAabb + 3] = host[kCntAabb + 4] = host[kCntAabb + 5] = -1;
    const uint64_t root_key = (uint64_t) p; // code 0, scale max
    const int root_ts = -1;
    if (brick_scale == p) {
        host[kCntBricks] = 1; // the root is the only brick
        const int zero = 0;
        if (hipMemcpy(m->brick_top.ptr, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("se8_map_create: hipMemcpy failed");
            se8_map_destroy(m);
            return SE8_ERR_HIP;
        }
    }
    if (hipMemcpy(m->counters.ptr, host, sizeof(host), hipMemcpyHostToDevice) != hipSuccess
        || hipMemcpy(m->node_key.ptr, &root_key, sizeof(root_key), hipMemcpyHostToDevice) != hipSuccess
        || hipMemcpy(m->node_ts.ptr, &root_ts, sizeof(root_ts), hipMemcpyHostToDevice) != hipSuccess
        || hipDeviceSynchronize() != hipSuccess) {
        set_error("se8_map_create: initialisation copy failed");
        se8_map_destroy(m);
        return SE8_ERR_HIP;
    }
    // coherent and mapped: the end of each integration writes the counters here from a kernel
    if (hipHostMalloc((void**) &m->host_counters, sizeof(int) * kNumCounters,
                      hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void**) &m->host_counters_dev, m->host_counters, 0) != hipSuccess) {
        set_error("se8_map_create: hipHostMalloc failed");
        se8_map_destroy(m);
        return SE8_ERR_OOM;
    }
    std::memcpy(m->host_counters, host, sizeof(host));
    for (int i = 0; i < 8; i++) {
        if (hipEventCreate(&m->ev[i]) != hipSuccess) {
            set_error("se8_map_create: hipEventCreate failed");
            se8_map_destroy(m);
            return SE8_ERR_HIP;
        }
    }
    *out = m;
    return SE8_OK;
}

extern "C" int se8_map_destroy(se8_map* m)
{
    if (!m)
        return SE8_OK;
    (void) hipDeviceSynchronize();
    DeviceBuffer* bufs[] = {&m->brick_top,    &m->brick_slots, &m->node_child, &m->node_key,   &m->node_ts,
                            &m->block_key,    &m->block_parent, &m->block_ts,  &m->voxels,     &m->counters,
                            &m->update_list,  &m->candidates,   &m->unique_keys, &m->cur_node,
                            &m->frame_poses,  &m->mesh_offsets, &m->block_scale, &m->mr_new_scale,
                            &m->block_dense,  &m->node_dense};
    for (DeviceBuffer* b : bufs)
        dev_free(*b);
    m->sorter.release();
    for (int i = 0; i < 8; i++)
        if (m->ev[i])
            (void) hipEventDestroy(m->ev[i]);
    if (m->host_counters)
        (void) hipHostFree(m->host_counters);
    delete m;
    return SE8_OK;
}

extern "C" int se8_map_get_info(const se8_map* m, se8_map_info* out)
{
    SE8_REQUIRE(m && out, "se8_map_get_info: null argument");
    int host[kNumCounters];
    int rc = read_counters(m, host, nullptr);
    if (rc != SE8_OK)
        return rc;
    out->octree_size = m->view.size;
    out->max_scale = m->view.max_scale;
    out->res = m->view.res;
    out->dim = m->view.dim;
    out->storage = m->view.storage;
    out->num_blocks = std::min<int64_t>(host[kCntBlocks], m->view.max_blocks);
    out->num_nodes = std::min<int64_t>(host[kCntNodes], m->view.max_nodes);
    out->max_blocks = m->view.max_blocks;
    out->aabb_valid = host[kCntAabb + 3] >= 0;
    for (int i = 0; i < 3; i++) {
        out->aabb_min[i] = host[kCntAabb + i];
        out->aabb_max[i] = host[kCntAabb + 3 + i];
    }
    return host[kCntOverflow] ? SE8_ERR_CAPACITY : SE8_OK;
}

namespace se8 {

// integrateDepth (map_integrator_impl.hpp:52-83) for either sensor; the pose is T_WC16 (host) or,
// when T16_dev is set, 16 floats in device memory read by the first kernel.  `what` names the
// entry point in error messages.
template<class Sensor>
static int integrate_impl(se8_map* m, int timestamp, const float* depth_dev, const Sensor& sensor,
                          const float* T_WC16, const float* T16_dev, hipStream_t s, const char* what)
{
    const Iso T_arg = T_WC16 ? iso_from16(T_WC16) : Iso{};
    // maps with the dense block and node tables allocate in place (k_carve_claim + k_link_new);
    // others sort the candidate keys and build the octree level by level
    if (!(m->view.block_dense && m->view.node_dense)) {
        const float band = 2.0f * m->view.res * m->tsdf.truncation_boundary_factor;
        const int num_steps = (int) std::ceil(band / (2.f * m->view.res));
        const int rc = ensure_candidates(m, (size_t) sensor.width() * sensor.height() * num_steps);
        if (rc != SE8_OK)
            return rc;
    }

    // Capacity overflow of the previous frame is reported now: its counters were copied to pinned
    // memory at the end of that call (the copy is normally complete by this point).  The map keeps
    // the blocks that fit, stays readable, and refuses further integration.
    if (m->counters_pending) {
        SE8_HIP_CHECK(hipEventSynchronize(m->ev[7]));
        m->counters_pending = false;
        harvest_integration(m);
        if (m->host_counters[kCntOverflow])
            m->overflowed = true;
    }
    if (m->overflowed) {
        set_error(std::string(what) + ": map capacity exhausted (raise se8_tsdf_config.max_blocks)");
        return SE8_ERR_CAPACITY;
    }
    // with timing off no event is recorded and nothing here waits for the device
    const bool timed = m->timing != SE8_TIMING_OFF;
    hipEvent_t ev_upd0 = timed ? m->ev[3] : nullptr, ev_upd1 = timed ? m->ev[4] : nullptr;
    FramePoses* fp = (FramePoses*) m->frame_poses.ptr;
    if (timed)
        SE8_HIP_CHECK(hipEventRecord(m->ev[0], s));
    // single-resolution maps with the dense tables link the new octants inside the update kernel
    const int link = m->view.block_dense && m->view.node_dense && !is_multires(m) ? 1 : 0;
    int rc = allocate_frame(m, T_arg, T16_dev, depth_dev, sensor, timed ? m->ev[1] : nullptr,
                            timed ? m->ev[2] : nullptr, s, link != 0);
    if (rc != SE8_OK)
        return rc;
    // ev[3] / ev[4] are stamped at the start and end of the update kernel itself
    const int* ulist = (const int*) m->update_list.ptr;
    if (is_multires(m)) {
        if constexpr (std::is_same<Sensor, PinholeSensor>::value) {
            if (timed)
                SE8_HIP_CHECK(hipEventRecord(m->ev[3], s));
            rc = launch_multires_update(m, fp, sensor, depth_dev, timestamp, s);
            if (rc != SE8_OK)
                return rc;
            if (timed)
                SE8_HIP_CHECK(hipEventRecord(m->ev[4], s));
        }
        else {
            set_error(std::string(what) + ": the multi-resolution updater takes a pinhole camera");
            return SE8_ERR_INVALID;
        }
    }
    else {
        // one resident wave of workgroups: the grid-stride loop then never leaves a second, partial
        // round of workgroups waiting for the first to finish.  Timed launches carry the stage
        // events in the dispatch; untimed ones are plain launches (capturable in a frame graph).
        auto* ku = m->view.storage == SE8_STORAGE_F32 ? k_update<SE8_STORAGE_F32, Sensor> : k_update<SE8_STORAGE_F16, Sensor>;
        const dim3 grid(resident_grid(ku, 512));
        if (timed)
            hipExtLaunchKernelGGL(ku, grid, dim3(512), 0, s, ev_upd0, ev_upd1, 0, m->view, ulist,
                                  (const FramePoses*) fp, sensor, depth_dev, timestamp, link);
        else
            hipLaunchKernelGGL(ku, grid, dim3(512), 0, s, m->view, ulist, (const FramePoses*) fp, sensor, depth_dev,
                               timestamp, link);
        SE8_HIP_CHECK(hipGetLastError());
        graph_note_launch(kNodeUpdate, s);
    }
    SE8_HIP_CHECK(hipGetLastError());
    if (!timed)
        return SE8_OK;
    // the counters reach the pinned host copy from a one-wave kernel (a copy engine transfer costs
    // more than the kernel at 80 bytes); the host reads them after event 7
    k_publish_counters<<<1, 64, 0, s>>>(m->view.counters, m->host_counters_dev);
    SE8_HIP_CHECK(hipGetLastError());
    SE8_HIP_CHECK(hipEventRecord(m->ev[7], s));
    m->counters_pending = true;
    m->integ_pending = true;
    return SE8_OK;
}

// Direction table of the lidar model: evaluated in double, rounded to float once.
static void lidar_directions(const se8_lidar_config& c, std::vector<float>& dir4)
{
    dir4.assign((size_t) 4 * c.width * c.height, 0.f);
    const double pi = 3.14159265358979323846;
    for (int v = 0; v < c.height; v++) {
        const double e = (double) c.beam_elevation_deg[v] * (pi / 180.0), a = (double) c.beam_azimuth_deg[v] * (pi / 180.0);
        for (int u = 0; u < c.width; u++) {
            const double phi = 2.0 * pi * (double) u / (double) c.width - pi + a;
            float* d = dir4.data() + 4 * ((size_t) v * c.width + u);
            d[0] = (float) (cos(e) * cos(phi));
            d[1] = (float) (cos(e) * sin(phi));
            d[2] = (float) sin(e);
        }
    }
}

__global__ void k_lidar_project(LidarModel L, const float* pts, int64_t n, float* uv, uint8_t* ok)
{
    for (int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        const float p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        float q[2] = {0.f, 0.f};
        ok[i] = lidar_project(L, p, q) ? 1 : 0;
        uv[2 * i] = q[0];
        uv[2 * i + 1] = q[1];
    }
}

} // namespace se8

extern "C" int se8_integrate_depth(se8_map* m, int timestamp, const float* depth_dev, const se8_pinhole* sensor,
                                   const float T_WC16[16], se8_stream stream)
{
    SE8_REQUIRE(m && depth_dev && sensor && T_WC16, "se8_integrate_depth: null argument");
    SE8_REQUIRE(sensor->width > 0 && sensor->height > 0, "se8_integrate_depth: empty image");
    const PinholeSensor ps{make_cam(*sensor, 1, false)};
    return integrate_impl(m, timestamp, depth_dev, ps, T_WC16, nullptr, (hipStream_t) stream, "se8_integrate_depth");
}

extern "C" int se8_integrate_depth_dev(se8_map* m, int timestamp, const float* depth_dev, const se8_pinhole* sensor,
                                       const float* T_WC_dev, se8_stream stream)
{
    SE8_REQUIRE(m && depth_dev && sensor && T_WC_dev, "se8_integrate_depth_dev: null argument");
    SE8_REQUIRE(sensor->width > 0 && sensor->height > 0, "se8_integrate_depth_dev: empty image");
    const PinholeSensor ps{make_cam(*sensor, 1, false)};
    return integrate_impl(m, timestamp, depth_dev, ps, nullptr, T_WC_dev, (hipStream_t) stream,
                          "se8_integrate_depth_dev");
}

extern "C" int se8_lidar_create(const se8_lidar_config* cfg, se8_lidar** out)
{
    SE8_REQUIRE(cfg && out && cfg->beam_elevation_deg && cfg->beam_azimuth_deg, "se8_lidar_create: null argument");
    SE8_REQUIRE(cfg->width > 0 && cfg->height >= 2, "se8_lidar_create: need width >= 1 and height >= 2");
    SE8_REQUIRE(cfg->near_plane > 0.f && cfg->far_plane > cfg->near_plane, "se8_lidar_create: bad near/far planes");
    for (int v = 1; v < cfg->height; v++)
        SE8_REQUIRE(cfg->beam_elevation_deg[v] < cfg->beam_elevation_deg[v - 1],
                    "se8_lidar_create: beam elevations must be strictly descending");
    *out = nullptr;
    se8_lidar* l = new se8_lidar();
    l->width = cfg->width;
    l->height = cfg->height;
    const float deg_to_rad = (float) (3.14159265358979323846 / 180.0); // ouster_lidar.cpp:80
    std::vector<float> elev(cfg->height), az(cfg->height), dir4;
    float mn = cfg->beam_elevation_deg[0], mx = cfg->beam_elevation_deg[0];
    for (int v = 0; v < cfg->height; v++) {
        elev[v] = cfg->beam_elevation_deg[v] * deg_to_rad;
        az[v] = cfg->beam_azimuth_deg[v] * deg_to_rad;
        mn = std::min(mn, cfg->beam_elevation_deg[v]);
        mx = std::max(mx, cfg->beam_elevation_deg[v]);
    }
    lidar_directions(*cfg, dir4);
    // search hint: for bucket b (elevations below elev[0] - b * span / nb), the row of its top edge
    const int nb = 4 * cfg->height;
    const float span = elev[0] - elev[cfg->height - 1];
    std::vector<int> bucket(nb);
    for (int b = 0, lo = 0; b < nb; b++) {
        const float top = elev[0] - span * ((float) b / (float) nb);
        while (lo < cfg->height - 2 && elev[lo + 1] >= top)
            lo++;
        bucket[b] = lo;
    }
    int rc = dev_alloc(l->elev, sizeof(float) * elev.size());
    rc = rc ? rc : dev_alloc(l->az, sizeof(float) * az.size());
    rc = rc ? rc : dev_alloc(l->dir, sizeof(float) * dir4.size());
    rc = rc ? rc : dev_alloc(l->bucket, sizeof(int) * bucket.size());
    if (rc == SE8_OK && (hipMemcpy(l->elev.ptr, elev.data(), sizeof(float) * elev.size(), hipMemcpyHostToDevice)
                             != hipSuccess
                         || hipMemcpy(l->az.ptr, az.data(), sizeof(float) * az.size(), hipMemcpyHostToDevice) != hipSuccess
                         || hipMemcpy(l->dir.ptr, dir4.data(), sizeof(float) * dir4.size(), hipMemcpyHostToDevice)
                             != hipSuccess
                         || hipMemcpy(l->bucket.ptr, bucket.data(), sizeof(int) * bucket.size(), hipMemcpyHostToDevice)
                             != hipSuccess)) {
        set_error("se8_lidar_create: upload failed");
        rc = SE8_ERR_HIP;
    }
    if (rc != SE8_OK) {
        se8_lidar_destroy(l);
        return rc;
    }
    LidarModel& L = l->model;
    L.w = cfg->width;
    L.h = cfg->height;
    L.near_plane = cfg->near_plane;
    L.far_plane = cfg->far_plane;
    L.min_elev = mn * deg_to_rad;
    L.max_elev = mx * deg_to_rad;
    L.cols_per_rad = (float) ((double) cfg->width / (2.0 * 3.14159265358979323846));
    L.elev = (const float*) l->elev.ptr;
    L.az = (const float*) l->az.ptr;
    L.dir = (const float4*) l->dir.ptr;
    L.bucket = (const int*) l->bucket.ptr;
    L.nbucket = nb;
    L.bucket_scale = (float) nb / span;
    *out = l;
    return SE8_OK;
}

extern "C" int se8_lidar_destroy(se8_lidar* l)
{
    if (!l)
        return SE8_OK;
    (void) hipDeviceSynchronize();
    dev_free(l->elev);
    dev_free(l->az);
    dev_free(l->dir);
    dev_free(l->bucket);
    delete l;
    return SE8_OK;
}

extern "C" int se8_integrate_range(se8_map* m, int timestamp, const float* range_dev, const se8_lidar* lidar,
                                   const float T_WS16[16], se8_stream stream)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_integrate_range: not available for a multi-resolution map");
    SE8_REQUIRE(m && range_dev && lidar && T_WS16, "se8_integrate_range: null argument");
    const LidarSensor ls{lidar->model};
    return integrate_impl(m, timestamp, range_dev, ls, T_WS16, nullptr, (hipStream_t) stream, "se8_integrate_range");
}

extern "C" int se8_integrate_range_dev(se8_map* m, int timestamp, const float* range_dev, const se8_lidar* lidar,
                                       const float* T_WS_dev, se8_stream stream)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_integrate_range_dev: not available for a multi-resolution map");
    SE8_REQUIRE(m && range_dev && lidar && T_WS_dev, "se8_integrate_range_dev: null argument");
    const LidarSensor ls{lidar->model};
    return integrate_impl(m, timestamp, range_dev, ls, nullptr, T_WS_dev, (hipStream_t) stream,
                          "se8_integrate_range_dev");
}

extern "C" int se8_lidar_project(const se8_lidar* lidar, const float* points, int64_t n, float* uv,
                                 uint8_t* in_image)
{
    SE8_REQUIRE(lidar && (n == 0 || (points && uv && in_image)) && n >= 0, "se8_lidar_project: bad argument");
    if (n == 0)
        return SE8_OK;
    DeviceBuffer dp, du, dk;
    int rc = dev_alloc(dp, sizeof(float) * 3 * n);
    rc = rc ? rc : dev_alloc(du, sizeof(float) * 2 * n);
    rc = rc ? rc : dev_alloc(dk, (size_t) n);
    if (rc == SE8_OK) {
        if (hipMemcpy(dp.ptr, points, sizeof(float) * 3 * n, hipMemcpyHostToDevice) != hipSuccess)
            rc = SE8_ERR_HIP;
        else {
            const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
            k_lidar_project<<<(unsigned) blocks, 256>>>(lidar->model, (const float*) dp.ptr, n, (float*) du.ptr,
                                                        (uint8_t*) dk.ptr);
            if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess
                || hipMemcpy(uv, du.ptr, sizeof(float) * 2 * n, hipMemcpyDeviceToHost) != hipSuccess
                || hipMemcpy(in_image, dk.ptr, (size_t) n, hipMemcpyDeviceToHost) != hipSuccess)
                rc = SE8_ERR_HIP;
        }
        if (rc == SE8_ERR_HIP)
            set_error("se8_lidar_project: device error");
    }
    dev_free(dp);
    dev_free(du);
    dev_free(dk);
    return rc;
}

// ---------------------------------------------------------------------------------------------
// explicit allocation and block access (tests, tools)
// ---------------------------------------------------------------------------------------------
namespace se8 {
__global__ void k_lookup_blocks(MapView m, const int* coords, int n, int* slots)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
    if (x < 0 || y < 0 || z < 0 || x >= m.size || y >= m.size || z >= m.size) {
        slots[i] = -1;
        return;
    }
    slots[i] = lookup_block(m, x, y, z);
}

int lookup_slots(const se8_map* m, const std::vector<int>& coords, std::vector<int>& slots)
{
    const int n = (int) (coords.size() / 3);
    slots.assign(n, -1);
    if (n == 0)
        return SE8_OK;
    DeviceBuffer dc, ds;
    int rc = dev_alloc(dc, sizeof(int) * coords.size());
    rc = rc ? rc : dev_alloc(ds, sizeof(int) * n);
    if (rc == SE8_OK) {
        if (hipMemcpy(dc.ptr, coords.data(), sizeof(int) * coords.size(), hipMemcpyHostToDevice) != hipSuccess)
            rc = SE8_ERR_HIP;
        else {
            k_lookup_blocks<<<(n + 255) / 256, 256>>>(m->view, (const int*) dc.ptr, n, (int*) ds.ptr);
            if (hipMemcpy(slots.data(), ds.ptr, sizeof(int) * n, hipMemcpyDeviceToHost) != hipSuccess)
                rc = SE8_ERR_HIP;
        }
    }
    if (rc == SE8_ERR_HIP)
        set_error("lookup_slots: HIP copy failed");
    dev_free(dc);
    dev_free(ds);
    return rc;
}
} // namespace se8

extern "C" int se8_allocate_blocks(se8_map* m, const int* voxel_coords, int64_t n, int only_new,
                                   int* out_block_coords, int64_t* out_n)
{
    SE8_REQUIRE(m && (n == 0 || voxel_coords) && out_n, "se8_allocate_blocks: null argument");
    std::vector<uint64_t> keys;
    keys.reserve((size_t) n);
    for (int64_t i = 0; i < n; i++) {
        const int x = voxel_coords[3 * i], y = voxel_coords[3 * i + 1], z = voxel_coords[3 * i + 2];
        SE8_REQUIRE(x >= 0 && y >= 0 && z >= 0 && x < m->view.size && y < m->view.size && z < m->view.size,
                    "se8_allocate_blocks: voxel outside the octree");
        keys.push_back(block_key_of_voxel(x, y, z));
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<int> coords;
    for (uint64_t k : keys) {
        int c[3];
        key_to_coord(k, c);
        coords.insert(coords.end(), c, c + 3);
    }
    std::vector<int> slots;
    int rc = lookup_slots(m, coords, slots);
    if (rc != SE8_OK)
        return rc;
    std::vector<uint64_t> missing;
    for (size_t i = 0; i < keys.size(); i++)
        if (slots[i] < 0)
            missing.push_back(keys[i]);
    if (!missing.empty()) {
        rc = ensure_candidates(m, missing.size());
        if (rc != SE8_OK)
            return rc;
        SE8_HIP_CHECK(hipMemcpy(m->candidates.ptr, missing.data(), sizeof(uint64_t) * missing.size(),
                                hipMemcpyHostToDevice));
        const int n_missing = (int) missing.size();
        SE8_HIP_CHECK(hipMemcpy((int*) m->counters.ptr + kCntCandidates, &n_missing, sizeof(int),
                                hipMemcpyHostToDevice));
        rc = build_from_candidates(m, (uint64_t*) m->candidates.ptr, m->view.counters + kCntCandidates, n_missing,
                                   nullptr, 0, nullptr);
        if (rc != SE8_OK)
            return rc;
        int host[kNumCounters];
        rc = read_counters(m, host, nullptr);
        if (rc != SE8_OK)
            return rc;
        if (host[kCntOverflow]) {
            m->overflowed = true;
            set_error("se8_allocate_blocks: map capacity exhausted");
            return SE8_ERR_CAPACITY;
        }
    }
    const std::vector<uint64_t>& ret = only_new ? missing : keys;
    *out_n = (int64_t) ret.size();
    if (out_block_coords) {
        for (size_t i = 0; i < ret.size(); i++)
            key_to_coord(ret[i], out_block_coords + 3 * i);
    }
    return SE8_OK;
}

extern "C" int se8_write_block(se8_map* m, const int bc[3], const float* tsdf512, const float* weight512)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_write_block: not available for a multi-resolution map");
    SE8_REQUIRE(m && bc && tsdf512 && weight512, "se8_write_block: null argument");
    std::vector<int> coords(bc, bc + 3), slots;
    int rc = lookup_slots(m, coords, slots);
    if (rc != SE8_OK)
        return rc;
    SE8_REQUIRE(slots[0] >= 0, "se8_write_block: block not allocated");
    const size_t off = (size_t) slots[0] * 512;
    if (m->view.storage == SE8_STORAGE_F32) {
        std::vector<float> buf(1024);
        for (int i = 0; i < 512; i++) {
            buf[2 * i] = tsdf512[i];
            buf[2 * i + 1] = weight512[i];
        }
        SE8_HIP_CHECK(hipMemcpy((float2*) m->voxels.ptr + off, buf.data(), 4096, hipMemcpyHostToDevice));
    }
    else {
        std::vector<uint32_t> buf(512);
        for (int i = 0; i < 512; i++)
            buf[i] = pack_f16(tsdf512[i], weight512[i]);
        SE8_HIP_CHECK(hipMemcpy((uint32_t*) m->voxels.ptr + off, buf.data(), 2048, hipMemcpyHostToDevice));
    }
    return SE8_OK;
}

extern "C" int se8_read_block(const se8_map* m, const int bc[3], float* tsdf512, float* weight512)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_read_block: not available for a multi-resolution map");
    SE8_REQUIRE(m && bc && tsdf512 && weight512, "se8_read_block: null argument");
    std::vector<int> coords(bc, bc + 3), slots;
    int rc = lookup_slots(m, coords, slots);
    if (rc != SE8_OK)
        return rc;
    SE8_REQUIRE(slots[0] >= 0, "se8_read_block: block not allocated");
    const size_t off = (size_t) slots[0] * 512;
    if (m->view.storage == SE8_STORAGE_F32) {
        std::vector<float> buf(1024);
        SE8_HIP_CHECK(hipMemcpy(buf.data(), (float2*) m->voxels.ptr + off, 4096, hipMemcpyDeviceToHost));
        for (int i = 0; i < 512; i++) {
            tsdf512[i] = buf[2 * i];
            weight512[i] = buf[2 * i + 1];
        }
    }
    else {
        std::vector<uint32_t> buf(512);
        SE8_HIP_CHECK(hipMemcpy(buf.data(), (uint32_t*) m->voxels.ptr + off, 2048, hipMemcpyDeviceToHost));
        for (int i = 0; i < 512; i++)
            unpack_f16(buf[i], tsdf512[i], weight512[i]);
    }
    return SE8_OK;
}

extern "C" int se8_read_update_list(const se8_map* m, uint64_t* keys, int64_t cap, int64_t* out_n)
{
    SE8_REQUIRE(m && out_n, "se8_read_update_list: null argument");
    int host[kNumCounters];
    int rc = read_counters(m, host, nullptr);
    if (rc != SE8_OK)
        return rc;
    const int64_t n = std::min<int64_t>(host[kCntUpdate], m->view.max_blocks);
    const int64_t nb = std::min<int64_t>(host[kCntBlocks], m->view.max_blocks);
    *out_n = n;
    if (!keys)
        return SE8_OK;
    if (cap < n) {
        set_error("se8_read_update_list: capacity below the list length");
        return SE8_ERR_CAPACITY;
    }
    std::vector<int> entries((size_t) n);
    std::vector<uint64_t> block_keys((size_t) nb);
    if (n > 0 && (hipMemcpy(entries.data(), m->update_list.ptr, sizeof(int) * n, hipMemcpyDeviceToHost) != hipSuccess
                  || hipMemcpy(block_keys.data(), m->block_key.ptr, sizeof(uint64_t) * nb, hipMemcpyDeviceToHost)
                         != hipSuccess)) {
        set_error("se8_read_update_list: copy failed");
        return SE8_ERR_HIP;
    }
    for (int64_t i = 0; i < n; i++) {
        const int slot = entries[i] & 0x7fffffff;
        keys[i] = slot < nb ? block_keys[slot] : 0;
    }
    return SE8_OK;
}

extern "C" int se8_dump_blocks(const se8_map* m, uint64_t* keys, int* timestamps, float* tsdf, float* weight,
                               int64_t cap, int64_t* out_n)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_dump_blocks: not available for a multi-resolution map");
    SE8_REQUIRE(m && out_n, "se8_dump_blocks: null argument");
    int host[kNumCounters];
    int rc = read_counters(m, host, nullptr);
    if (rc != SE8_OK)
        return rc;
    const int64_t n = std::min<int64_t>(host[kCntBlocks], m->view.max_blocks);
    *out_n = n;
    if (!keys)
        return SE8_OK;
    SE8_REQUIRE(cap >= n, "se8_dump_blocks: output capacity too small");
    std::vector<uint64_t> k(n);
    std::vector<int> ts(n);
    if (n > 0) {
        SE8_HIP_CHECK(hipMemcpy(k.data(), m->block_key.ptr, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
        SE8_HIP_CHECK(hipMemcpy(ts.data(), m->block_ts.ptr, sizeof(int) * n, hipMemcpyDeviceToHost));
    }
    std::vector<int64_t> order(n);
    for (int64_t i = 0; i < n; i++)
        order[i] = i;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return k[a] < k[b]; });
    std::vector<uint32_t> vox16;
    std::vector<float> vox32;
    if ((tsdf || weight) && n > 0) {
        if (m->view.storage == SE8_STORAGE_F32) {
            vox32.resize((size_t) n * 1024);
            SE8_HIP_CHECK(hipMemcpy(vox32.data(), m->voxels.ptr, sizeof(float) * vox32.size(), hipMemcpyDeviceToHost));
        }
        else {
            vox16.resize((size_t) n * 512);
            SE8_HIP_CHECK(
                hipMemcpy(vox16.data(), m->voxels.ptr, sizeof(uint32_t) * vox16.size(), hipMemcpyDeviceToHost));
        }
    }
    for (int64_t i = 0; i < n; i++) {
        const int64_t s = order[i];
        keys[i] = k[s];
        if (timestamps)
            timestamps[i] = ts[s];
        for (int v = 0; v < 512 && (tsdf || weight); v++) {
            float t, w;
            if (m->view.storage == SE8_STORAGE_F32) {
                t = vox32[(size_t) s * 1024 + 2 * v];
                w = vox32[(size_t) s * 1024 + 2 * v + 1];
            }
            else
                unpack_f16(vox16[(size_t) s * 512 + v], t, w);
            if (tsdf)
                tsdf[(size_t) i * 512 + v] = t;
            if (weight)
                weight[(size_t) i * 512 + v] = w;
        }
    }
    return SE8_OK;
}

extern "C" int se8_dump_nodes(const se8_map* m, uint64_t* keys, uint8_t* masks, int* timestamps, int64_t cap,
                              int64_t* out_n)
{
    SE8_REQUIRE(m && out_n, "se8_dump_nodes: null argument");
    int host[kNumCounters];
    int rc = read_counters(m, host, nullptr);
    if (rc != SE8_OK)
        return rc;
    const int64_t n = std::min<int64_t>(host[kCntNodes], m->view.max_nodes);
    *out_n = n;
    if (!keys)
        return SE8_OK;
    SE8_REQUIRE(cap >= n, "se8_dump_nodes: output capacity too small");
    std::vector<uint64_t> k(n);
    std::vector<int> ch((size_t) n * 8), ts(n);
    SE8_HIP_CHECK(hipMemcpy(k.data(), m->node_key.ptr, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    SE8_HIP_CHECK(hipMemcpy(ch.data(), m->node_child.ptr, sizeof(int) * 8 * n, hipMemcpyDeviceToHost));
    SE8_HIP_CHECK(hipMemcpy(ts.data(), m->node_ts.ptr, sizeof(int) * n, hipMemcpyDeviceToHost));
    std::vector<int64_t> order(n);
    for (int64_t i = 0; i < n; i++)
        order[i] = i;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return k[a] < k[b]; });
    for (int64_t i = 0; i < n; i++) {
        const int64_t s = order[i];
        keys[i] = k[s];
        uint8_t mk = 0;
        for (int c = 0; c < 8; c++)
            if (ch[s * 8 + c] >= 0)
                mk |= (uint8_t) (1u << c);
        if (masks)
            masks[i] = mk;
        if (timestamps)
            timestamps[i] = ts[s];
    }
    return SE8_OK;
}

namespace se8 {
static float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void) hipGetLastError();
        return 0.f;
    }
    return ms;
}

// Folds the events of the last integration into the stage times (requires ev[7] complete).
void harvest_integration(se8_map* m)
{
    if (!m->integ_pending)
        return;
    se8_stage_times& t = m->times;
    t.fetch_frustum = elapsed(m->ev[0], m->ev[1]);
    t.create_list = elapsed(m->ev[1], m->ev[2]);
    t.allocate_list = elapsed(m->ev[2], m->ev[3]);
    t.update = elapsed(m->ev[3], m->ev[4]);
    const int* h = m->host_counters;
    t.last_frustum_blocks = h[kCntFrustum];
    t.last_new_blocks = h[kCntUpdate] - h[kCntFrustum];
    t.last_candidates = h[kCntCandidates];
    t.sum_fetch_frustum += t.fetch_frustum;
    t.sum_create_list += t.create_list;
    t.sum_allocate_list += t.allocate_list;
    t.sum_update += t.update;
    t.sum_update_blocks += h[kCntUpdate];
    t.sum_candidates += h[kCntCandidates];
    t.n_integrations += 1;
    m->integ_pending = false;
}

// Folds the events of the last raycast into the stage times; without blocking it only does so
// if they have completed.
void harvest_raycast(se8_map* m, bool block)
{
    if (!m->ray_pending)
        return;
    if (block)
        (void) hipEventSynchronize(m->ev[6]);
    if (hipEventQuery(m->ev[6]) != hipSuccess) {
        (void) hipGetLastError();
        if (!block)
            m->ray_pending = false; // about to be re-recorded: the sample is dropped
        return;
    }
    m->times.raycast = elapsed(m->ev[5], m->ev[6]);
    m->times.sum_raycast += m->times.raycast;
    m->times.n_raycasts += 1;
    m->ray_pending = false;
}

void harvest_tracking(se8_tracker* t)
{
    if (!t->track_pending)
        return;
    (void) hipEventSynchronize(t->ev[1]);
    t->times.tracking = elapsed(t->ev[0], t->ev[1]);
    t->times.sum_tracking += t->times.tracking;
    t->times.n_tracks += 1;
    t->track_pending = false;
    if (t->lvl_pending) {
        // levels run coarsest first: level l lasts until the first launch of level l - 1
        for (int l = t->cfg.num_levels - 1; l >= 0; l--) {
            hipEvent_t end = l > 0 ? t->ev_lvl[l - 1] : t->ev_lvl[SE8_TRACKER_MAX_LEVELS];
            t->times.sum_track_level[l] += elapsed(t->ev_lvl[l], end);
            t->times.n_track_level_launches[l] += t->cfg.iterations[l];
        }
        t->lvl_pending = false;
    }
}
} // namespace se8

extern "C" int se8_stage_times_get(const se8_map* m, const se8_tracker* tr, se8_stage_times* out);

extern "C" int se8_set_stage_timing(se8_map* m, se8_tracker* tr, int mode)
{
    SE8_REQUIRE(mode == SE8_TIMING_OFF || mode == SE8_TIMING_STAGES || mode == SE8_TIMING_LEVELS,
                "se8_set_stage_timing: unknown mode");
    se8_stage_times tmp;
    int rc = se8_stage_times_get(m, tr, &tmp); // harvest what is pending under the old mode
    if (rc != SE8_OK)
        return rc;
    if (m) {
        if (m->timing == SE8_TIMING_OFF && mode != SE8_TIMING_OFF) {
            // an overflow during untimed integrations is reported from now on
            int host[kNumCounters];
            rc = read_counters(m, host, nullptr);
            if (rc != SE8_OK)
                return rc;
            if (host[kCntOverflow])
                m->overflowed = true;
        }
        m->timing = mode;
    }
    if (tr) {
        if (mode == SE8_TIMING_LEVELS)
            for (int i = 0; i <= SE8_TRACKER_MAX_LEVELS; i++)
                if (!tr->ev_lvl[i])
                    SE8_HIP_CHECK(hipEventCreate(&tr->ev_lvl[i]));
        tr->timing = mode;
    }
    return SE8_OK;
}

extern "C" int se8_stage_times_get(const se8_map* m, const se8_tracker* tr, se8_stage_times* out)
{
    SE8_REQUIRE(out, "se8_stage_times_get: null argument");
    std::memset(out, 0, sizeof(*out));
    if (m) {
        se8_map* mm = const_cast<se8_map*>(m);
        if (mm->counters_pending) {
            SE8_HIP_CHECK(hipEventSynchronize(mm->ev[7]));
            mm->counters_pending = false;
        }
        harvest_integration(mm);
        harvest_raycast(mm, true);
        *out = mm->times;
    }
    if (tr) {
        se8_tracker* tt = const_cast<se8_tracker*>(tr);
        harvest_tracking(tt);
        out->tracking = tt->times.tracking;
        out->sum_tracking = tt->times.sum_tracking;
        out->n_tracks = tt->times.n_tracks;
        for (int l = 0; l < SE8_TRACKER_MAX_LEVELS; l++) {
            out->sum_track_level[l] = tt->times.sum_track_level[l];
            out->n_track_level_launches[l] = tt->times.n_track_level_launches[l];
        }
    }
    return SE8_OK;
}

extern "C" int se8_stage_times_reset(se8_map* m, se8_tracker* tr)
{
    se8_stage_times out;
    int rc = se8_stage_times_get(m, tr, &out); // harvest everything pending first
    if (rc != SE8_OK)
        return rc;
    if (m)
        std::memset(&m->times, 0, sizeof(m->times));
    if (tr)
        std::memset(&tr->times, 0, sizeof(tr->times));
    return SE8_OK;
}

extern "C" int se8_device_alloc(size_t bytes, void** out_dev)
{
    SE8_REQUIRE(out_dev, "se8_device_alloc: null argument");
    *out_dev = nullptr;
    if (hipMalloc(out_dev, bytes ? bytes : 16) != hipSuccess) {
        (void) hipGetLastError();
        *out_dev = nullptr;
        set_error("se8_device_alloc: hipMalloc of " + std::to_string(bytes) + " bytes failed");
        return SE8_ERR_OOM;
    }
    return SE8_OK;
}

extern "C" int se8_device_free(void* dev)
{
    if (dev)
        SE8_HIP_CHECK(hipFree(dev));
    return SE8_OK;
}

extern "C" int se8_copy(void* dst, const void* src, size_t bytes, int direction, se8_stream stream)
{
    SE8_REQUIRE((dst && src) || bytes == 0, "se8_copy: null argument");
    SE8_REQUIRE(direction >= SE8_COPY_H2D && direction <= SE8_COPY_D2D, "se8_copy: bad direction");
    const hipMemcpyKind kind = direction == SE8_COPY_H2D ? hipMemcpyHostToDevice
        : direction == SE8_COPY_D2H                       ? hipMemcpyDeviceToHost
                                                          : hipMemcpyDeviceToDevice;
    if (bytes)
        SE8_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, kind, (hipStream_t) stream));
    return SE8_OK;
}

extern "C" int se8_stream_synchronize(se8_stream stream)
{
    SE8_HIP_CHECK(hipStreamSynchronize((hipStream_t) stream));
    return SE8_OK;
}
