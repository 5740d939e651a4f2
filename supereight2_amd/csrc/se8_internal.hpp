// Internal declarations shared by the HIP translation units of libse8.so.
//
// Floating-point evaluation order of every fp32 expression on the allocation/update/raycast path
// (kept identical in host and device code so allocation is bit-exact):
//   dot3(a,b)     = a0*b0 + (a1*b1 + a2*b2)
//   (R*v)_i       = R_i0*v0 + (R_i1*v1 + R_i2*v2)
//   (T*p)_i       = (R*p)_i + t_i
//   plane.[c;1]   = (n0*c0 + n1*c1) + (n2*c2 + n3)
//   project(p)    = (fx*(x*(1/z)) + cx, fy*(y*(1/z)) + cy)
//   backProject   = ((u-cx)*(1/fx), (v-cy)*(1/fy), 1)
// All code is compiled with -ffp-contract=off and IEEE division/sqrt.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/se8.h"

#define SE8_HD __host__ __device__ __forceinline__

namespace se8 {

// ----------------------------------------------------------------------------------------------
// error handling
// ----------------------------------------------------------------------------------------------
void set_error(const std::string& msg);

#define SE8_HIP_CHECK(expr)                                                                        \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            se8::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                     \
            return SE8_ERR_HIP;                                                                    \
        }                                                                                          \
    } while (0)

#define SE8_REQUIRE(cond, msg)                                                                     \
    do {                                                                                           \
        if (!(cond)) {                                                                             \
            se8::set_error(msg);                                                                   \
            return SE8_ERR_INVALID;                                                                \
        }                                                                                          \
    } while (0)

// ----------------------------------------------------------------------------------------------
// rigid transforms in the documented operation order
// ----------------------------------------------------------------------------------------------
struct Iso {
    float R[3][3];
    float t[3];
};

SE8_HD float dot3(const float a[3], const float b[3])
{
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

SE8_HD void lin_apply(const Iso& T, const float v[3], float o[3])
{
    const float x = T.R[0][0] * v[0] + (T.R[0][1] * v[1] + T.R[0][2] * v[2]);
    const float y = T.R[1][0] * v[0] + (T.R[1][1] * v[1] + T.R[1][2] * v[2]);
    const float z = T.R[2][0] * v[0] + (T.R[2][1] * v[1] + T.R[2][2] * v[2]);
    o[0] = x;
    o[1] = y;
    o[2] = z;
}

SE8_HD void iso_apply(const Iso& T, const float v[3], float o[3])
{
    const float x = (T.R[0][0] * v[0] + (T.R[0][1] * v[1] + T.R[0][2] * v[2])) + T.t[0];
    const float y = (T.R[1][0] * v[0] + (T.R[1][1] * v[1] + T.R[1][2] * v[2])) + T.t[1];
    const float z = (T.R[2][0] * v[0] + (T.R[2][1] * v[1] + T.R[2][2] * v[2])) + T.t[2];
    o[0] = x;
    o[1] = y;
    o[2] = z;
}

SE8_HD Iso iso_mul(const Iso& A, const Iso& B)
{
    Iso o;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            o.R[i][j] = A.R[i][0] * B.R[0][j] + (A.R[i][1] * B.R[1][j] + A.R[i][2] * B.R[2][j]);
        o.t[i] = (A.R[i][0] * B.t[0] + (A.R[i][1] * B.t[1] + A.R[i][2] * B.t[2])) + A.t[i];
    }
    return o;
}

SE8_HD Iso iso_inv(const Iso& A)
{
    Iso o;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            o.R[i][j] = A.R[j][i];
    for (int i = 0; i < 3; i++)
        o.t[i] = -(o.R[i][0] * A.t[0] + (o.R[i][1] * A.t[1] + o.R[i][2] * A.t[2]));
    return o;
}

inline Iso iso_from16(const float* m)
{
    Iso T;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T.R[i][j] = m[4 * i + j];
        T.t[i] = m[4 * i + 3];
    }
    return T;
}

inline void iso_to16(const Iso& T, float* m)
{
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            m[4 * i + j] = T.R[i][j];
        m[4 * i + 3] = T.t[i];
    }
    m[12] = m[13] = m[14] = 0.f;
    m[15] = 1.f;
}

SE8_HD float sqnorm3(const float v[3])
{
    return dot3(v, v);
}

SE8_HD void normalized3(const float v[3], float o[3])
{
    const float z = sqnorm3(v);
    if (z > 0.f) {
        const float n = sqrtf(z);
        o[0] = v[0] / n;
        o[1] = v[1] / n;
        o[2] = v[2] / n;
    }
    else {
        o[0] = v[0];
        o[1] = v[1];
        o[2] = v[2];
    }
}

SE8_HD void cross3(const float a[3], const float b[3], float o[3])
{
    const float x = a[1] * b[2] - a[2] * b[1];
    const float y = a[2] * b[0] - a[0] * b[2];
    const float z = a[0] * b[1] - a[1] * b[0];
    o[0] = x;
    o[1] = y;
    o[2] = z;
}

// ----------------------------------------------------------------------------------------------
// pinhole camera (src/sensor/pinhole_camera.cpp:29-111,189-310; pinhole_camera_impl.hpp:15-32)
// ----------------------------------------------------------------------------------------------
struct Cam {
    int w, h;
    float fx, fy, cx, cy, inv_fx, inv_fy, near_plane, far_plane;
    float planes[6][4];
    int left_hand_frame;
};

// scaled = 1 reproduces PinholeCamera(sensor, dsf), which applies the downsampling formulas
// even for dsf = 1.
Cam make_cam(const se8_pinhole& p, int dsf, bool scaled);

SE8_HD void cam_backproject(const Cam& c, float u, float v, float o[3])
{
    o[0] = (u - c.cx) * c.inv_fx;
    o[1] = (v - c.cy) * c.inv_fy;
    o[2] = 1.f;
}

SE8_HD bool cam_project(const Cam& c, const float p[3], float uv[2])
{
    if (fabs((double) p[2]) < 1.0e-12)
        return false;
    const float rz = 1.f / p[2];
    const float x = p[0] * rz, y = p[1] * rz;
    uv[0] = c.fx * x + c.cx;
    uv[1] = c.fy * y + c.cy;
    if (uv[0] < -0.5f || uv[1] < -0.5f)
        return false;
    if (uv[0] >= (float) c.w - 0.5f || uv[1] >= (float) c.h - 0.5f)
        return false;
    return p[2] > 0.f;
}

SE8_HD bool cam_sphere_in_frustum(const Cam& c, const float ctr[3], float radius)
{
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const float d = (c.planes[i][0] * ctr[0] + c.planes[i][1] * ctr[1])
            + (c.planes[i][2] * ctr[2] + c.planes[i][3]);
        if (d < -radius)
            return false;
    }
    return true;
}

SE8_HD float cam_far_dist(const Cam& c, const float ray[3])
{
    float n[3];
    normalized3(ray, n);
    return c.far_plane / n[2];
}

SE8_HD float cam_near_dist(const Cam& c, const float ray[3])
{
    float n[3];
    normalized3(ray, n);
    return c.near_plane / n[2];
}

// ----------------------------------------------------------------------------------------------
// Morton keys (include/se/map/utils/impl/key_util_impl.hpp:44-247)
// ----------------------------------------------------------------------------------------------
SE8_HD uint64_t expand_bits(uint64_t v)
{
    uint64_t x = v & 0x7FFFFull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

SE8_HD uint32_t compact_bits(uint64_t v)
{
    uint64_t x = v & 0x49249249249249ull;
    x = (x | x >> 2) & 0x10c30c30c30c30c3ull;
    x = (x | x >> 4) & 0x100f00f00f00f00full;
    x = (x | x >> 8) & 0x1f0000ff0000ffull;
    x = (x | x >> 16) & 0x1f00000000ffffull;
    x = (x | x >> 32) & 0x1fffffull;
    return (uint32_t) x;
}

SE8_HD uint64_t morton_code(int x, int y, int z)
{
    return expand_bits((uint64_t) x) | (expand_bits((uint64_t) y) << 1) | (expand_bits((uint64_t) z) << 2);
}

constexpr int kScaleOffset = 5;
constexpr int kBlockScale = 3;
constexpr int kBlockVoxels = 512;

// Key of the block (scale 3) containing voxel (x,y,z).
SE8_HD uint64_t block_key_of_voxel(int x, int y, int z)
{
    return ((morton_code(x, y, z) & ~0x1FFull) << kScaleOffset) | (uint64_t) kBlockScale;
}

SE8_HD void key_to_coord(uint64_t key, int c[3])
{
    const uint64_t code = key >> kScaleOffset;
    c[0] = (int) compact_bits(code);
    c[1] = (int) compact_bits(code >> 1);
    c[2] = (int) compact_bits(code >> 2);
}

// ----------------------------------------------------------------------------------------------
// voxel storage
// ----------------------------------------------------------------------------------------------
// F32: float2 {tsdf, weight} per voxel.  F16: uint32 {half tsdf (low 16 bits), uint16 weight}.
SE8_HD uint32_t pack_f16(float tsdf, float weight)
{
    const __half h = __float2half_rn(tsdf);
    const uint16_t hb = *reinterpret_cast<const uint16_t*>(&h);
    return (uint32_t) hb | ((uint32_t) (uint16_t) weight << 16);
}

SE8_HD void unpack_f16(uint32_t v, float& tsdf, float& weight)
{
    const uint16_t hb = (uint16_t) (v & 0xFFFFu);
    __half h;
    *reinterpret_cast<uint16_t*>(&h) = hb;
    tsdf = __half2float(h);
    weight = (float) (v >> 16);
}

// Everything a kernel needs to address the map, passed by value.
struct MapView {
    int size;         // octree size in voxels
    int max_scale;    // log2(size)
    int brick_shift;  // log2(blocks per brick side): bricks are octants of scale 3 + brick_shift
    int bricks_side;  // size >> (3 + brick_shift)
    float res;
    float inv_dim_unused;
    float dim;        // size * res
    float mu;         // res * truncation factor
    float max_weight;
    float t_MW[3];
    int storage;
    int* brick_top;     // bricks_side^3 brick ids (-1 = none)
    int* brick_slots;   // per brick 512 block slots (-1 = none), index (bx&m) + 8(by&m) + 64(bz&m)
    int* node_child;    // per node 8 children: node index, or block slot below scale-4 nodes
    uint64_t* node_key; // per node key
    int* node_ts;       // per node timestamp
    uint64_t* block_key;
    int* block_parent;  // scale-4 node owning each block
    int* block_ts;
    void* voxels;       // max_blocks * 512 voxels
    int* counters;      // [0] blocks, [1] nodes, [2] bricks, [3] overflow flag, [4..9] aabb
    int64_t max_blocks, max_nodes, max_bricks;
};

enum Counter {
    kCntBlocks = 0,
    kCntNodes = 1,
    kCntBricks = 2,
    kCntOverflow = 3,
    kCntAabb = 4, // 6 ints: min xyz, max xyz
    kCntFrustum = 10,
    kCntUpdate = 11,
    kCntCandidates = 12,
    kCntUnique = 13,
    kNumCounters = 16
};

// Block slot of the block containing voxel (x,y,z); -1 if unallocated.  Coordinates are taken
// modulo the octree size, as the reference's top-down fetcher ignores bits at and above the
// root size (fetcher_impl.hpp:107-126); a world point within an ulp of the map's upper bound can
// round to voxel index == size.
__device__ __forceinline__ int lookup_block(const MapView& m, int x, int y, int z)
{
    const int wrap = m.size - 1;
    const int bx = (x & wrap) >> 3, by = (y & wrap) >> 3, bz = (z & wrap) >> 3;
    const int s = m.brick_shift;
    const int brick = m.brick_top[((bz >> s) * m.bricks_side + (by >> s)) * m.bricks_side + (bx >> s)];
    if (brick < 0)
        return -1;
    const int msk = (1 << s) - 1;
    return m.brick_slots[brick * 512 + (bx & msk) + ((by & msk) << 3) + ((bz & msk) << 6)];
}

// lookup_block without a branch on the brick: both loads are always issued (the second from brick
// 0 when the brick is missing, whose result is then discarded), so a caller that looks up several
// blocks gets all their loads in flight together.
__device__ __forceinline__ int lookup_block_nb(const MapView& m, int x, int y, int z)
{
    const int wrap = m.size - 1;
    const int bx = (x & wrap) >> 3, by = (y & wrap) >> 3, bz = (z & wrap) >> 3;
    const int s = m.brick_shift;
    const int brick = m.brick_top[((bz >> s) * m.bricks_side + (by >> s)) * m.bricks_side + (bx >> s)];
    const int msk = (1 << s) - 1;
    const int slot = m.brick_slots[max(brick, 0) * 512 + (bx & msk) + ((by & msk) << 3) + ((bz & msk) << 6)];
    return brick < 0 ? -1 : slot;
}

__device__ __forceinline__ void load_voxel(const MapView& m, int slot, int vi, float& tsdf, float& weight)
{
    if (m.storage == SE8_STORAGE_F32) {
        const float2 v = reinterpret_cast<const float2*>(m.voxels)[(size_t) slot * 512 + vi];
        tsdf = v.x;
        weight = v.y;
    }
    else {
        const uint32_t v = reinterpret_cast<const uint32_t*>(m.voxels)[(size_t) slot * 512 + vi];
        unpack_f16(v, tsdf, weight);
    }
}

// getData semantics: data of the voxel if its block is allocated, otherwise {1, 0}.
__device__ __forceinline__ void get_voxel(const MapView& m, int x, int y, int z, float& tsdf, float& weight)
{
    const int slot = lookup_block(m, x, y, z);
    if (slot < 0) {
        tsdf = 1.f;
        weight = 0.f;
        return;
    }
    load_voxel(m, slot, (x & 7) + ((y & 7) << 3) + ((z & 7) << 6), tsdf, weight);
}

// ----------------------------------------------------------------------------------------------
// host-side map object
// ----------------------------------------------------------------------------------------------
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

int dev_alloc(DeviceBuffer& b, size_t bytes);
void dev_free(DeviceBuffer& b);

struct RadixSorter {
    DeviceBuffer keys_alt, tile_hist, tile_flags;
    size_t cap = 0;
    int reserve(size_t n);
    void release();
};

} // namespace se8

struct se8_map {
    se8_map_config cfg;
    se8_tsdf_config tsdf;
    se8::MapView view;
    se8::DeviceBuffer brick_top, brick_slots, node_child, node_key, node_ts, block_key, block_parent, block_ts,
        voxels, counters;
    // per-frame scratch
    se8::DeviceBuffer update_list, candidates, unique_keys, cur_node;
    se8::RadixSorter sorter;
    size_t candidates_cap = 0;
    int* host_counters = nullptr; // pinned copy of counters taken at the end of each integration
    bool counters_pending = false; // a copy into host_counters is in flight (event ev[7])
    bool integ_pending = false;    // events ev[0..4] of the last integration not yet harvested
    bool ray_pending = false;      // events ev[5..6] of the last raycast not yet harvested
    se8_stage_times times{};
    hipEvent_t ev[8]{};
};

struct se8_tracker {
    se8_pinhole sensor;
    se8_tracker_config cfg;
    se8::DeviceBuffer vert_lvl[3], norm_lvl[3]; // float4 per pixel of each pyramid level
    se8::DeviceBuffer partials; // per-workgroup reduction partials
    se8::DeviceBuffer state;    // counters, flags, final sums
    se8::DeviceBuffer ref_points, ref_normals; // used by se8_track_raycast
    se8::DeviceBuffer adj_poses, adj_partials, adj_out; // batched cost/gradient evaluation
    int num_wg[3] = {0, 0, 0};    // Gauss-Newton grid per level
    int icp_ppt[3] = {1, 1, 1};   // pixels per thread per tile
    int icp_tiles[3] = {0, 0, 0}; // tiles of 256 * ppt pixels
    void* host_state = nullptr; // pinned readback of the device tracking state
    bool track_pending = false;
    se8_stage_times times{};
    hipEvent_t ev[2]{};
};

namespace se8 {
// stage-time bookkeeping (se8_map.hip)
void harvest_integration(se8_map* m);
void harvest_raycast(se8_map* m, bool block);
void harvest_tracking(se8_tracker* t);
} // namespace se8
