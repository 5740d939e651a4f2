// Map object, allocation pipeline and TSDF update of libse8.so.
//
// integrateDepth (include/se/integrator/impl/map_integrator_impl.hpp:52-83) becomes, per frame:
//   k_frame_begin  per-frame counters cleared; T_WS, its inverse and T_CV derived from the pose
//   k_frustum      FrustumIterator over the allocated blocks (raycast_carver_impl.hpp:17-32,
//                  iterator.hpp:197-238) as a flat per-block test of the block and its ancestors
//   k_alloc_rays   the RaycastCarver band march (raycast_carver_impl.hpp:74-111): Morton keys of
//                  unallocated blocks appended to a candidate list
//   radix sort     LSD sort of the candidate keys (11-bit digits over the significant key bits)
//   k_build_level  one launch per octree scale: segment heads of the sorted keys create missing
//                  nodes (allocator_impl.hpp:68-147, octree_impl.hpp:171-199); duplicate keys
//                  fall in the same segment, so the unique step is the segment-head flag
//   k_build_blocks block slots, brick index, AABB and the "new" entries of the update list
//   k_update       Updater<Map<TSDF,Single>> (singleres_tsdf_updater_impl.hpp:14-132), one 8^3
//                  block per 512-thread workgroup
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

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

static void plane_normal(const float p1[3], const float p2[3], const float p3[3], float out[3])
{
    float t1[3], t2[3], n[3];
    for (int i = 0; i < 3; i++) {
        t1[i] = p2[i] - p1[i];
        t2[i] = p3[i] - p2[i];
    }
    cross3(t1, t2, n);
    normalized3(n, out);
}

Cam make_cam(const se8_pinhole& p, int dsf, bool scaled)
{
    Cam c;
    const float s = (float) dsf;
    c.w = (int) ((float) p.width / s);
    c.h = (int) ((float) p.height / s);
    if (!scaled) {
        c.fx = p.fx;
        c.fy = p.fy;
        c.cx = p.cx;
        c.cy = p.cy;
    }
    else {
        c.fx = p.fx / s;
        c.fy = p.fy / s;
        c.cx = (p.cx + 0.5f) / s - 0.5f;
        c.cy = (p.cy + 0.5f) / s - 0.5f;
    }
    c.inv_fx = 1.f / c.fx;
    c.inv_fy = 1.f / c.fy;
    c.near_plane = p.near_plane;
    c.far_plane = p.far_plane;
    c.left_hand_frame = (p.fx < 0.0f) ^ (p.fy < 0.0f);
    float tl[3], tr[3], br[3], bl[3];
    cam_backproject(c, 0.f, 0.f, tl);
    cam_backproject(c, (float) c.w, 0.f, tr);
    cam_backproject(c, (float) c.w, (float) c.h, br);
    cam_backproject(c, 0.f, (float) c.h, bl);
    float v[8][3]; // near TL TR BR BL, far TL TR BR BL
    for (int i = 0; i < 3; i++) {
        v[0][i] = c.near_plane * tl[i];
        v[1][i] = c.near_plane * tr[i];
        v[2][i] = c.near_plane * br[i];
        v[3][i] = c.near_plane * bl[i];
        v[4][i] = c.far_plane * tl[i];
        v[5][i] = c.far_plane * tr[i];
        v[6][i] = c.far_plane * br[i];
        v[7][i] = c.far_plane * bl[i];
    }
    plane_normal(v[4], v[0], v[3], c.planes[0]); // left
    plane_normal(v[1], v[5], v[6], c.planes[1]); // right
    plane_normal(v[7], v[3], v[2], c.planes[2]); // bottom
    plane_normal(v[5], v[1], v[0], c.planes[3]); // top
    for (int i = 0; i < 4; i++)
        c.planes[i][3] = 0.f;
    const float nearp[4] = {0.f, 0.f, 1.f, -c.near_plane};
    const float farp[4] = {0.f, 0.f, -1.f, c.far_plane};
    for (int j = 0; j < 4; j++) {
        c.planes[4][j] = nearp[j];
        c.planes[5][j] = farp[j];
    }
    float sign = c.fx * c.fy;
    sign = sign / fabsf(sign);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 3; j++)
            c.planes[i][j] *= sign;
    return c;
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

// Start of an integration: clears the per-frame counters and derives the frame's poses from T_WS,
// read from device memory when T16_dev is set, else taken from the argument.
__global__ void k_frame_begin(MapView m, Iso T_arg, const float* T16_dev, FramePoses* fp)
{
    if (threadIdx.x == 0) {
        m.counters[kCntFrustum] = 0;
        m.counters[kCntUpdate] = 0;
        m.counters[kCntCandidates] = 0;
        m.counters[kCntUnique] = 0;
        m.counters[kCntMrNew] = 0;
        m.counters[kCntMrDown] = 0;
        m.counters[kCntNodesStart] = m.counters[kCntNodes];
        for (int i = 0; i < 4; i++)
            m.counters[kCntMrScale + i] = 0;
        const Iso T = T16_dev ? iso_from16(T16_dev) : T_arg;
        fp->T_WS = T;
        fp->T_SW = iso_inv(T);
        fp->T_CV = compute_T_CV(m, T);
    }
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
// fetch-frustum: a block is listed iff its own sphere and those of all its ancestors below the
// root pass sphereInFrustum (the DFS of iterator_impl.hpp:86-156 never prunes at the root).
// ----------------------------------------------------------------------------------------------
template<class Sensor>
__global__ void k_frustum(MapView m, const FramePoses* fp, Sensor sensor, int* update_list)
{
    // the block counter runs past the pool when an allocation overflows; only slots below the
    // capacity hold blocks
    const int n = (int) min((int64_t) m.counters[kCntBlocks], m.max_blocks);
    const Iso T_SW = fp->T_SW;
    const float r_unit = sqrtf(3.0f) / 2.0f * m.res;
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const int slot = base + threadIdx.x;
        bool in = false;
        if (slot < n) {
            int c[3];
            key_to_coord(m.block_key[slot], c);
            in = true;
            for (int s = kBlockScale; s < m.max_scale && in; s++) {
                const int size = 1 << s;
                float cW[3];
                for (int i = 0; i < 3; i++) {
                    const int oc = c[i] & ~(size - 1);
                    const float pM = ((float) oc + (float) size * 0.5f) * m.res;
                    cW[i] = pM + (-m.t_MW[i]);
                }
                float cS[3];
                iso_apply(T_SW, cW, cS);
                in = sensor.sphere_in_frustum(cS, r_unit * (float) size);
            }
        }
        const int pos = wave_append_slot(in, &m.counters[kCntUpdate]);
        if (in)
            update_list[pos] = slot;
    }
    // the frustum part of the list is complete when this kernel ends; the carver kernel records
    // its size
}

// ----------------------------------------------------------------------------------------------
// create-list: band march of RaycastCarver (raycast_carver_impl.hpp:74-111)
// ----------------------------------------------------------------------------------------------
template<class Sensor>
__global__ void k_alloc_rays(MapView m, Sensor sensor, const FramePoses* fp, const float* depth, float band,
                             int num_steps, uint64_t* cand, int64_t cand_cap)
{
    const Iso T_WS = fp->T_WS;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    // the frustum part of the update list is complete (the frustum kernel has ended): record its
    // size; nothing in this kernel touches the update counter
    if (x == 0 && y == 0)
        m.counters[kCntFrustum] = m.counters[kCntUpdate];
    bool live = x < sensor.width();
    float pos[3] = {0.f, 0.f, 0.f}, step[3] = {0.f, 0.f, 0.f};
    if (live) {
        const float d = depth[x + (size_t) y * sensor.width()];
        if (d < sensor.near_plane() || d > (sensor.far_plane() + band * 0.5f))
            live = false;
        else {
            float dir[3], pC[3], pW[3];
            sensor.backproject(x, y, dir);
            for (int i = 0; i < 3; i++)
                pC[i] = d * dir[i];
            iso_apply(T_WS, pC, pW);
            float diff[3], rev[3];
            for (int i = 0; i < 3; i++)
                diff[i] = T_WS.t[i] - pW[i];
            normalized3(diff, rev);
            for (int i = 0; i < 3; i++) {
                pos[i] = pW[i] - (band * 0.5f) * rev[i];
                step[i] = (rev[i] * band) / (float) num_steps;
            }
        }
    }
    // The band samples are taken in batches: every sample's voxel and both lookup loads of the
    // batch are in flight together, then the keys of unallocated blocks are emitted in sample
    // order (consecutive repeats along the ray dropped), with one atomic per wave for the batch.
    constexpr int kBatch = 8;
    uint64_t prev = ~0ull;
    for (int s0 = 0; s0 < num_steps; s0 += kBatch) {
        int vox[kBatch][3];
        bool inb[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            inb[j] = false;
            vox[j][0] = vox[j][1] = vox[j][2] = 0;
            if (live && s0 + j < num_steps) {
                float pM[3];
                for (int i = 0; i < 3; i++)
                    pM[i] = pos[i] + m.t_MW[i];
                if (pM[0] >= 0.f && pM[0] < m.dim && pM[1] >= 0.f && pM[1] < m.dim && pM[2] >= 0.f && pM[2] < m.dim) {
                    const int wrap = m.size - 1; // see lookup_block
                    vox[j][0] = ((int) (pM[0] / m.res)) & wrap;
                    vox[j][1] = ((int) (pM[1] / m.res)) & wrap;
                    vox[j][2] = ((int) (pM[2] / m.res)) & wrap;
                    inb[j] = true;
                }
                for (int i = 0; i < 3; i++)
                    pos[i] = pos[i] + step[i];
            }
        }
        int slot[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; j++)
            slot[j] = lookup_block_nb(m, vox[j][0], vox[j][1], vox[j][2]);
        uint64_t keys[kBatch];
        bool emit[kBatch];
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            keys[j] = block_key_of_voxel(vox[j][0], vox[j][1], vox[j][2]);
            emit[j] = false;
            if (inb[j] && slot[j] < 0) {
                emit[j] = keys[j] != prev;
                prev = keys[j];
            }
            cnt += emit[j] ? 1 : 0;
        }
        // exclusive prefix of the per-lane counts (at most kBatch, 4 bits) over the wave
        const uint64_t lt = (1ull << __lane_id()) - 1ull;
        int before = 0, total = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const uint64_t mb = __ballot((cnt >> b) & 1);
            before += __popcll(mb & lt) << b;
            total += __popcll(mb) << b;
        }
        if (total == 0)
            continue;
        const int leader = __ffsll((unsigned long long) __ballot(cnt > 0)) - 1;
        int base = 0;
        if (__lane_id() == leader)
            base = atomicAdd(&m.counters[kCntCandidates], total);
        base = __shfl(base, leader);
        int64_t at = (int64_t) base + before;
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            if (emit[j]) {
                if (at < cand_cap)
                    cand[at] = keys[j];
                at++;
            }
        }
    }
}

// Exclusive prefix over the wave of per-lane counts in [0, 16) and the wave's total, by ballots of
// the count's bits.
__device__ __forceinline__ void wave_prefix_small(int cnt, int& before, int& total)
{
    const uint64_t lt = (1ull << __lane_id()) - 1ull;
    before = total = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const uint64_t mb = __ballot((cnt >> b) & 1);
        before += __popcll(mb & lt) << b;
        total += __popcll(mb) << b;
    }
}

// Per lane, `total` consecutive positions of *counter for the wave (one atomic, by the first lane
// with a non-zero count); returns this lane's first position.  Every lane of the wave must call it.
__device__ __forceinline__ int wave_reserve(int cnt, int* counter)
{
    int before, total;
    wave_prefix_small(cnt, before, total);
    if (total == 0)
        return 0;
    const int leader = __ffsll((unsigned long long) __ballot(cnt > 0)) - 1;
    int base = 0;
    if (__lane_id() == leader)
        base = atomicAdd(counter, total);
    return __shfl(base, leader) + before;
}

constexpr int kClaimed = -2; // dense-table entry of an octant a thread is creating in this kernel

// Creates the missing ancestors of a block this thread has just created, bottom-up (Octree::allocate
// creates nodes top-down, octree_impl.hpp:171-199; the resulting node set is the same): each node is
// claimed in its scale's dense table, so only one thread creates an octant, then takes a pool
// index.  A node that exists, or that another thread has claimed, ends the walk: its ancestors exist
// or are being created by that thread.  Child links are written by k_link_new, once every index of
// the frame is final.
__device__ void claim_ancestors(const MapView& m, const int c[3], uint64_t code)
{
    const int brick_scale = m.brick_shift + kBlockScale;
    for (int s = kBlockScale + 1; s < m.max_scale; s++) {
        int* e = m.node_dense + m.node_dense_off[s] + dense_index(m, c, s);
        if (atomicCAS(e, -1, kClaimed) != -1)
            return;
        const int nn = atomicAdd(&m.counters[kCntNodes], 1);
        if (nn >= m.max_nodes) {
            atomicExch(&m.counters[kCntOverflow], 1);
            atomicExch(e, -1);
            return;
        }
        const uint64_t mask = (0x1ffffffffffffffull >> (3 * s)) << (3 * s);
        m.node_key[nn] = ((code & mask) << kScaleOffset) | (uint64_t) s;
        m.node_ts[nn] = -1;
        if (s == brick_scale) {
            const int br = atomicAdd(&m.counters[kCntBricks], 1);
            if (br >= m.max_bricks)
                atomicExch(&m.counters[kCntOverflow], 1);
            else
                m.brick_top[((c[2] >> s) * m.bricks_side + (c[1] >> s)) * m.bricks_side + (c[0] >> s)] = br;
        }
        atomicExch(e, nn);
    }
}

// create-list + allocate-list for maps with the dense tables: the band march of RaycastCarver
// (raycast_carver_impl.hpp:74-111), as k_alloc_rays samples it, with the allocation of
// allocator::blocks / Octree::allocate (allocator_impl.hpp:68-147, octree_impl.hpp:171-199) done in
// place: the first thread to reach an unallocated block claims its dense-table entry, creates the
// block (slot, key, timestamp, an update-list entry marked new) and its missing ancestors.  The set
// of blocks and nodes is that of the sorted-key path; slot and node numbering follow the schedule.
// kCntCandidates counts the samples' block keys of unallocated blocks (consecutive repeats dropped).
template<class Sensor>
__global__ void k_carve_claim(MapView m, Sensor sensor, const FramePoses* fp, const float* depth, float band,
                              int num_steps, int* update_list)
{
    const Iso T_WS = fp->T_WS;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    bool live = x < sensor.width();
    float pos[3] = {0.f, 0.f, 0.f}, step[3] = {0.f, 0.f, 0.f};
    if (live) {
        const float d = depth[x + (size_t) y * sensor.width()];
        if (d < sensor.near_plane() || d > (sensor.far_plane() + band * 0.5f))
            live = false;
        else {
            float dir[3], pC[3], pW[3];
            sensor.backproject(x, y, dir);
            for (int i = 0; i < 3; i++)
                pC[i] = d * dir[i];
            iso_apply(T_WS, pC, pW);
            float diff[3], rev[3];
            for (int i = 0; i < 3; i++)
                diff[i] = T_WS.t[i] - pW[i];
            normalized3(diff, rev);
            for (int i = 0; i < 3; i++) {
                pos[i] = pW[i] - (band * 0.5f) * rev[i];
                step[i] = (rev[i] * band) / (float) num_steps;
            }
        }
    }
    const int side = m.size >> 3;
    constexpr int kBatch = 8;
    uint64_t prev = ~0ull;
    for (int s0 = 0; s0 < num_steps; s0 += kBatch) {
        int vox[kBatch][3];
        bool inb[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            inb[j] = false;
            vox[j][0] = vox[j][1] = vox[j][2] = 0;
            if (live && s0 + j < num_steps) {
                float pM[3];
                for (int i = 0; i < 3; i++)
                    pM[i] = pos[i] + m.t_MW[i];
                if (pM[0] >= 0.f && pM[0] < m.dim && pM[1] >= 0.f && pM[1] < m.dim && pM[2] >= 0.f && pM[2] < m.dim) {
                    const int wrap = m.size - 1; // see lookup_block
                    vox[j][0] = ((int) (pM[0] / m.res)) & wrap;
                    vox[j][1] = ((int) (pM[1] / m.res)) & wrap;
                    vox[j][2] = ((int) (pM[2] / m.res)) & wrap;
                    inb[j] = true;
                }
                for (int i = 0; i < 3; i++)
                    pos[i] = pos[i] + step[i];
            }
        }
        int slot[kBatch];
        int64_t bi[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            bi[j] = ((int64_t) (vox[j][2] >> 3) * side + (vox[j][1] >> 3)) * side + (vox[j][0] >> 3);
            slot[j] = m.block_dense[bi[j]];
        }
        uint64_t keys[kBatch];
        bool won[kBatch];
        int cnt = 0, wins = 0;
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            keys[j] = block_key_of_voxel(vox[j][0], vox[j][1], vox[j][2]);
            won[j] = false;
            if (inb[j] && slot[j] < 0) {
                if (keys[j] != prev) {
                    cnt++;
                    // an entry already claimed in this kernel is being created by its claimant
                    if (slot[j] == -1)
                        won[j] = atomicCAS(&m.block_dense[bi[j]], -1, kClaimed) == -1;
                }
                prev = keys[j];
            }
            wins += won[j] ? 1 : 0;
        }
        const int cand0 = wave_reserve(cnt, &m.counters[kCntCandidates]);
        (void) cand0;
        // slots for the wave's new blocks, then update-list positions for those that fit the pool
        const int sbase = wave_reserve(wins, &m.counters[kCntBlocks]);
        int made = 0;
#pragma unroll
        for (int j = 0, k = 0; j < kBatch; j++)
            if (won[j])
                made += (sbase + k++ < m.max_blocks) ? 1 : 0;
        const int ubase = wave_reserve(made, &m.counters[kCntUpdate]);
        (void) wave_reserve(made, &m.counters[kCntUnique]);
#pragma unroll
        for (int j = 0, k = 0, u = 0; j < kBatch; j++) {
            if (!won[j])
                continue;
            const int sl = sbase + k++;
            if (sl >= m.max_blocks) {
                atomicExch(&m.counters[kCntOverflow], 1);
                atomicExch(&m.block_dense[bi[j]], -1);
                continue;
            }
            m.block_key[sl] = keys[j];
            m.block_ts[sl] = -1;
            update_list[ubase + u++] = (int) ((uint32_t) sl | 0x80000000u);
            atomicExch(&m.block_dense[bi[j]], sl);
            const int c[3] = {vox[j][0] & ~7, vox[j][1] & ~7, vox[j][2] & ~7};
            claim_ancestors(m, c, keys[j] >> kScaleOffset);
        }
    }
}

// The links of the octants k_carve_claim created: each new node becomes its parent's child (the root
// for the top scale), each new block its scale-4 parent's child (block_parent), its brick's slot
// entry and part of the AABB (octree_impl.hpp:267-273); the frustum part of the update list is the
// rest of it.  Loops run over whole waves (the AABB is aggregated by ballots).
__global__ void k_link_new(MapView m, const int* update_list)
{
    const int n_upd = m.counters[kCntUpdate];
    if (blockIdx.x == 0 && threadIdx.x == 0)
        m.counters[kCntFrustum] = n_upd - m.counters[kCntUnique];
    const int stride = gridDim.x * blockDim.x;
    const int n0 = m.counters[kCntNodesStart];
    const int n1 = (int) min((int64_t) m.counters[kCntNodes], m.max_nodes);
    for (int i = n0 + blockIdx.x * blockDim.x + threadIdx.x; i < n1; i += stride) {
        const uint64_t key = m.node_key[i];
        const int s = (int) (key & 0x1f);
        const uint64_t code = key >> kScaleOffset;
        int c[3];
        key_to_coord(key, c);
        const int parent = s + 1 >= m.max_scale ? 0 : m.node_dense[m.node_dense_off[s + 1] + dense_index(m, c, s + 1)];
        if (parent >= 0)
            m.node_child[parent * 8 + (int) ((code >> (3 * s)) & 7)] = i;
    }
    for (int base = blockIdx.x * blockDim.x; base < n_upd; base += stride) {
        const int i = base + threadIdx.x;
        bool made = false;
        int c[3] = {0, 0, 0};
        if (i < n_upd) {
            const int entry = update_list[i];
            if (entry < 0) {
                made = true;
                const int slot = entry & 0x7fffffff;
                const uint64_t key = m.block_key[slot];
                const uint64_t code = key >> kScaleOffset;
                key_to_coord(key, c);
                const int parent = m.max_scale > kBlockScale + 1
                    ? m.node_dense[m.node_dense_off[kBlockScale + 1] + dense_index(m, c, kBlockScale + 1)]
                    : 0;
                m.block_parent[slot] = parent;
                if (parent >= 0)
                    m.node_child[parent * 8 + (int) ((code >> 9) & 7)] = slot;
                const int bs = m.brick_shift + kBlockScale;
                const int brick = m.brick_top[((c[2] >> bs) * m.bricks_side + (c[1] >> bs)) * m.bricks_side + (c[0] >> bs)];
                const int msk = (1 << m.brick_shift) - 1;
                const int bx = c[0] >> 3, by = c[1] >> 3, bz = c[2] >> 3;
                if (brick >= 0)
                    m.brick_slots[brick * 512 + (bx & msk) + ((by & msk) << 3) + ((bz & msk) << 6)] = slot;
            }
        }
        if (__ballot(made)) {
            const int lo[3] = {made ? c[0] : 0x7fffffff, made ? c[1] : 0x7fffffff, made ? c[2] : 0x7fffffff};
            const int hi[3] = {made ? c[0] + 7 : -1, made ? c[1] + 7 : -1, made ? c[2] + 7 : -1};
            int wlo[3], whi[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                wlo[a] = lo[a];
                whi[a] = hi[a];
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) {
                    wlo[a] = min(wlo[a], __shfl_xor(wlo[a], o));
                    whi[a] = max(whi[a], __shfl_xor(whi[a], o));
                }
            }
            if (__lane_id() == 0) {
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    atomicMin(&m.counters[kCntAabb + a], wlo[a]);
                    atomicMax(&m.counters[kCntAabb + 3 + a], whi[a]);
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------------------------
// LSD radix sort of 64-bit keys, 11-bit digits, stable (tile = 256 threads x 8 keys, striped).
// 11-bit digits cover the 21 significant key bits of a 1024^3 map in two passes (three at 16384^3).
// ----------------------------------------------------------------------------------------------
constexpr int kRsThreads = 256;
constexpr int kRsItems = 8;
constexpr int kRsTile = kRsThreads * kRsItems;
constexpr int kRsBits = 11;
constexpr int kRsBins = 1 << kRsBits;
constexpr int kRsBinsPerThread = kRsBins / kRsThreads;

// The key count lives in device memory (n = min(*n_dev, cap)), so the sort is enqueued without a
// host round trip; the grid covers `cap` keys and workgroups past the live tiles return at once.
__device__ __forceinline__ int live_count(const int* n_dev, int cap)
{
    return min(*n_dev, cap);
}

__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint64_t* keys, const int* n_dev, int cap, int shift,
                                                        uint32_t* tile_hist)
{
    const int n = live_count(n_dev, cap);
    const int n_tiles = (n + kRsTile - 1) / kRsTile;
    if ((int) blockIdx.x >= n_tiles)
        return;
    __shared__ uint32_t h[kRsBins];
    for (int j = 0; j < kRsBinsPerThread; j++)
        h[j * kRsThreads + threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRsTile;
    for (int i = 0; i < kRsItems; i++) {
        const int idx = base + i * kRsThreads + threadIdx.x;
        if (idx < n)
            atomicAdd(&h[(keys[idx] >> shift) & (kRsBins - 1)], 1u);
    }
    __syncthreads();
    for (int j = 0; j < kRsBinsPerThread; j++) {
        const int d = j * kRsThreads + threadIdx.x;
        tile_hist[(size_t) d * n_tiles + blockIdx.x] = h[d];
    }
}

// In-place exclusive scan of the n_tiles * kRsBins tile histograms (digit-major) with one
// 1024-thread workgroup.
__global__ __launch_bounds__(1024) void k_scan_single(uint32_t* data, const int* n_dev, int cap)
{
    __shared__ uint32_t sums[1024];
    const int n = ((live_count(n_dev, cap) + kRsTile - 1) / kRsTile) * kRsBins;
    const int t = threadIdx.x;
    const int chunk = (n + 1023) / 1024;
    const int start = t * chunk;
    const int end = min(start + chunk, n);
    // each thread walks its contiguous chunk 16 elements at a time, all 16 loads in flight
    constexpr int kBatch = 16;
    uint32_t s = 0;
    int i = start;
    for (; i + kBatch <= end; i += kBatch) {
        uint32_t v[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; j++)
            v[j] = data[i + j];
#pragma unroll
        for (int j = 0; j < kBatch; j++)
            s += v[j];
    }
    for (; i < end; i++)
        s += data[i];
    sums[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? sums[t - off] : 0u;
        __syncthreads();
        sums[t] += v;
        __syncthreads();
    }
    uint32_t run = sums[t] - s;
    i = start;
    for (; i + kBatch <= end; i += kBatch) {
        uint32_t v[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; j++)
            v[j] = data[i + j];
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            data[i + j] = run;
            run += v[j];
        }
    }
    for (; i < end; i++) {
        const uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
}

// Scatter of one tile: per round of 256 keys, lanes with equal digits are ranked by ballots, the
// per-wave digit counts go through LDS, and only the entries a round wrote are cleared again.
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const uint64_t* in, uint64_t* out, const int* n_dev, int cap,
                                                           int shift, const uint32_t* tile_off)
{
    const int n = live_count(n_dev, cap);
    const int n_tiles = (n + kRsTile - 1) / kRsTile;
    if ((int) blockIdx.x >= n_tiles)
        return;
    __shared__ uint32_t base[kRsBins];
    __shared__ uint32_t wave_cnt[kRsThreads / 64][kRsBins];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    for (int j = 0; j < kRsBinsPerThread; j++) {
        const int d = j * kRsThreads + t;
        base[d] = tile_off[(size_t) d * n_tiles + blockIdx.x];
        for (int w = 0; w < kRsThreads / 64; w++)
            wave_cnt[w][d] = 0;
    }
    __syncthreads();
    for (int r = 0; r < kRsItems; r++) {
        const int idx = blockIdx.x * kRsTile + r * kRsThreads + t;
        const bool valid = idx < n;
        const uint64_t key = valid ? in[idx] : 0ull;
        const uint32_t d = (uint32_t) (key >> shift) & (kRsBins - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRsBits; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t mb = __ballot(bit);
            peers &= bit ? mb : ~mb;
        }
        const int rank = __popcll(peers & ((1ull << lane) - 1ull));
        const uint32_t cnt = (uint32_t) __popcll(peers);
        const bool leader = valid && rank == 0;
        if (leader)
            wave_cnt[wave][d] = cnt;
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + (uint32_t) rank;
            for (int w = 0; w < wave; w++)
                pos += wave_cnt[w][d];
            out[pos] = key;
        }
        __syncthreads();
        if (leader) {
            atomicAdd(&base[d], cnt);
            wave_cnt[wave][d] = 0;
        }
        __syncthreads();
    }
}

int RadixSorter::reserve(size_t n)
{
    if (n <= cap)
        return SE8_OK;
    const size_t tiles = (n + kRsTile - 1) / kRsTile;
    int rc = dev_alloc(keys_alt, n * sizeof(uint64_t));
    if (rc == SE8_OK)
        rc = dev_alloc(tile_hist, tiles * kRsBins * sizeof(uint32_t));
    if (rc != SE8_OK)
        return rc;
    cap = n;
    return SE8_OK;
}

void RadixSorter::release()
{
    dev_free(keys_alt);
    dev_free(tile_hist);
    cap = 0;
}

// Sorts keys[0..min(*n_dev, cap)) on bit range [lo_bit, hi_bit); returns the buffer holding the
// result.  The number of passes is fixed by the bit range, so the result buffer is known here.
static uint64_t* radix_sort(RadixSorter& rs, uint64_t* keys, const int* n_dev, int cap, int lo_bit, int hi_bit,
                            hipStream_t s)
{
    const int max_tiles = (cap + kRsTile - 1) / kRsTile;
    uint64_t* a = keys;
    uint64_t* b = (uint64_t*) rs.keys_alt.ptr;
    for (int shift = lo_bit; shift < hi_bit; shift += kRsBits) {
        k_rs_hist<<<max_tiles, kRsThreads, 0, s>>>(a, n_dev, cap, shift, (uint32_t*) rs.tile_hist.ptr);
        k_scan_single<<<1, 1024, 0, s>>>((uint32_t*) rs.tile_hist.ptr, n_dev, cap);
        k_rs_scatter<<<max_tiles, kRsThreads, 0, s>>>(a, b, n_dev, cap, shift, (const uint32_t*) rs.tile_hist.ptr);
        std::swap(a, b);
    }
    return a;
}

// ----------------------------------------------------------------------------------------------
// allocate-list: level-by-level octree build from the sorted candidate keys
// ----------------------------------------------------------------------------------------------
__device__ void build_level_one(const MapView& m, const uint64_t* keys, int i, int* cur_node, int scale,
                                int brick_scale)
{
    const uint64_t key = keys[i];
    const uint64_t code = key >> kScaleOffset;
    int cur;
    if (scale == m.max_scale - 1)
        cur = 0;
    else {
        const int prev = cur_node[i];
        cur = prev < 0 ? -1 : m.node_child[prev * 8 + (int) ((code >> (3 * (scale + 1))) & 7)];
    }
    cur_node[i] = cur;
    if (cur < 0)
        return; // only after a node-pool overflow
    const bool head = (i == 0) || (((key ^ keys[i - 1]) >> (kScaleOffset + 3 * scale)) != 0);
    if (!head)
        return;
    const int idx = (int) ((code >> (3 * scale)) & 7);
    if (m.node_child[cur * 8 + idx] >= 0)
        return;
    const int nn = atomicAdd(&m.counters[kCntNodes], 1);
    if (nn >= m.max_nodes) {
        atomicExch(&m.counters[kCntOverflow], 1);
        return;
    }
    const uint64_t mask = (0x1ffffffffffffffull >> (3 * scale)) << (3 * scale);
    m.node_key[nn] = ((code & mask) << kScaleOffset) | (uint64_t) scale;
    m.node_ts[nn] = -1;
    if (m.node_dense) {
        int c[3];
        key_to_coord(key, c);
        m.node_dense[m.node_dense_off[scale] + dense_index(m, c, scale)] = nn;
    }
    if (scale == brick_scale) {
        const int br = atomicAdd(&m.counters[kCntBricks], 1);
        if (br >= m.max_bricks) {
            atomicExch(&m.counters[kCntOverflow], 1);
            return;
        }
        int c[3];
        key_to_coord(key, c);
        const int bs = brick_scale;
        m.brick_top[((c[2] >> bs) * m.bricks_side + (c[1] >> bs)) * m.bricks_side + (c[0] >> bs)] = br;
    }
    m.node_child[cur * 8 + idx] = nn;
}

// One launch per scale; keys[0..min(*n_dev, cap)) are sorted.  Grid-stride over the keys.
__global__ void k_build_level(MapView m, const uint64_t* keys, const int* n_dev, int cap, int* cur_node, int scale,
                              int brick_scale)
{
    const int n = min(*n_dev, cap);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        build_level_one(m, keys, i, cur_node, scale, brick_scale);
}

// Minimum / maximum over the lanes of a wave (every lane active).
__device__ __forceinline__ int wave_min_i32(int v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        v = min(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ int wave_max_i32(int v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        v = max(v, __shfl_xor(v, o));
    return v;
}

// Blocks of the sorted, deduplicated candidate keys that are not allocated yet: each gets a slot,
// its parent's child entry, its brick entry, the AABB and (integration) an update-list entry marked
// new, or (explicit allocation) its voxels initialised.  The slot, AABB and update-list atomics are
// aggregated per wave, so the loop runs over whole waves (every lane reaches each ballot).
__global__ void k_build_blocks(MapView m, const uint64_t* keys, const int* n_dev, int cap, const int* cur_node,
                               int* update_list, int mark_new)
{
    const int n = min(*n_dev, cap);
    const int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const int i = base + threadIdx.x;
        bool want = false;
        uint64_t key = 0;
        int parent = -1, idx = 0;
        if (i < n) {
            key = keys[i];
            if (i == 0 || keys[i - 1] != key) {
                const uint64_t code = key >> kScaleOffset;
                parent = 0;
                if (m.max_scale > 4) {
                    const int prev = cur_node[i];
                    parent = prev < 0 ? -1 : m.node_child[prev * 8 + (int) ((code >> 12) & 7)];
                }
                idx = (int) ((code >> 9) & 7);
                // a missing parent only follows a node-pool overflow; an existing child is an
                // allocated block (explicit allocation requests may repeat blocks)
                want = parent >= 0 && m.node_child[parent * 8 + idx] < 0;
            }
        }
        const int slot = wave_append_slot(want, &m.counters[kCntBlocks]);
        const bool made = want && slot < m.max_blocks;
        if (want && !made)
            atomicExch(&m.counters[kCntOverflow], 1);
        int c[3] = {0, 0, 0};
        if (made) {
            key_to_coord(key, c);
            m.block_key[slot] = key;
            m.block_parent[slot] = parent;
            m.block_ts[slot] = -1;
            m.node_child[parent * 8 + idx] = slot;
            const int bs = m.brick_shift + kBlockScale;
            const int brick =
                m.brick_top[((c[2] >> bs) * m.bricks_side + (c[1] >> bs)) * m.bricks_side + (c[0] >> bs)];
            const int msk = (1 << m.brick_shift) - 1;
            const int bx = c[0] >> 3, by = c[1] >> 3, bz = c[2] >> 3;
            if (brick >= 0)
                m.brick_slots[brick * 512 + (bx & msk) + ((by & msk) << 3) + ((bz & msk) << 6)] = slot;
            if (m.block_dense) {
                const int side = m.size >> 3;
                m.block_dense[(bz * side + by) * side + bx] = slot;
            }
        }
        // AABB of the allocated blocks (octree_impl.hpp:267-273): one atomic per wave and bound
        if (__ballot(made)) {
            const int lo[3] = {made ? c[0] : 0x7fffffff, made ? c[1] : 0x7fffffff, made ? c[2] : 0x7fffffff};
            const int hi[3] = {made ? c[0] + 7 : -1, made ? c[1] + 7 : -1, made ? c[2] + 7 : -1};
            int wlo[3], whi[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                wlo[a] = wave_min_i32(lo[a]);
                whi[a] = wave_max_i32(hi[a]);
            }
            if (__lane_id() == 0) {
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    atomicMin(&m.counters[kCntAabb + a], wlo[a]);
                    atomicMax(&m.counters[kCntAabb + 3 + a], whi[a]);
                }
            }
        }
        if (update_list) {
            const int pos = wave_append_slot(made, &m.counters[kCntUpdate]);
            if (made)
                update_list[pos] = mark_new ? (int) ((uint32_t) slot | 0x80000000u) : slot;
        }
        else if (made) {
            // explicit allocation: initialise the voxels here (integration marks new blocks instead)
            if (m.storage == SE8_STORAGE_MULTIRES) {
                for (int v = 0; v < kMrBlockFloat2; v++)
                    reinterpret_cast<float2*>(m.voxels)[(size_t) slot * kMrBlockFloat2 + v] = make_float2(1.f, 0.f);
                m.block_scale[2 * slot] = -1;
                m.block_scale[2 * slot + 1] = -1;
            }
            else
                for (int v = 0; v < 512; v++) {
                    if (m.storage == SE8_STORAGE_F32)
                        reinterpret_cast<float2*>(m.voxels)[(size_t) slot * 512 + v] = make_float2(1.f, 0.f);
                    else
                        reinterpret_cast<uint32_t*>(m.voxels)[(size_t) slot * 512 + v] = pack_f16(1.f, 0.f);
                }
        }
    }
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

// Per voxel as singleres_tsdf_updater_impl.hpp:44-128.  A workgroup owns the list entries
// blockIdx.x, blockIdx.x + gridDim.x, ...: it stages up to kUpdChunk of them (entry, key) in LDS
// with one round of loads (stamping the block and parent timestamps at the same time), then walks
// them kUpdK blocks at a time: the voxels of the next group are loaded while the current group
// projects, and the group's depth samples are all in flight together.
// amdgpu_waves_per_eu(8): the scalar registers must stay within the budget of 8 waves per SIMD
// (96 allocated), or the one resident wave of workgroups shrinks from 4 to 3 per CU.
template<int Storage, class Sensor>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_update(
    MapView m, const int* update_list, const FramePoses* fp, Sensor sensor, const float* depth, int ts)
{
    __shared__ int s_entry[kUpdChunk];
    __shared__ uint64_t s_key[kUpdChunk];
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
            s_key[t] = m.block_key[slot];
            // block timestamp and one-level parent propagation (propagator_impl.hpp:146-197)
            m.block_ts[slot] = ts;
            atomicMax(&m.node_ts[m.block_parent[slot]], ts);
        }
        __syncthreads();
        VoxelRaw<Storage> cur[kUpdK];
#pragma unroll
        for (int k = 0; k < kUpdK; k++) {
            cur[k].clear();
            if (k < cnt && s_entry[k] >= 0)
                cur[k].load(m, (size_t) s_entry[k] * 512 + t);
        }
        for (int i = 0; i < cnt; i += kUpdK) {
            // the next group's voxels, issued before this group's depth samples
            VoxelRaw<Storage> nxt[kUpdK];
#pragma unroll
            for (int k = 0; k < kUpdK; k++) {
                nxt[k].clear();
                const int j = i + kUpdK + k;
                if (j < cnt && s_entry[j] >= 0)
                    nxt[k].load(m, (size_t) s_entry[j] * 512 + t);
            }
            float pC[kUpdK][3], nrm[kUpdK], dval[kUpdK];
            bool d_ok[kUpdK];
#pragma unroll
            for (int k = 0; k < kUpdK; k++) {
                d_ok[k] = false;
                dval[k] = 0.f;
                nrm[k] = 0.f;
                pC[k][0] = pC[k][1] = pC[k][2] = 0.f;
                if (i + k < cnt) {
                    int c[3];
                    key_to_coord(s_key[i + k], c);
                    const float v[3] = {(float) (c[0] + vx), (float) (c[1] + vy), (float) (c[2] + vz)};
                    iso_apply(T_CV, v, pC[k]);
                    nrm[k] = sqrtf(sqnorm3(pC[k]));
                    float uv[2];
                    if (!sensor.beyond_far(pC[k], nrm[k]) && sensor.project(pC[k], uv)) {
                        const int px = (int) (uv[0] + 0.5f), py = (int) (uv[1] + 0.5f);
                        dval[k] = depth[px + (size_t) py * sensor.width()];
                        d_ok[k] = true;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < kUpdK; k++) {
                if (i + k >= cnt)
                    break;
                const int entry = s_entry[i + k];
                float tsdf, weight;
                cur[k].get(tsdf, weight);
                // measurementFromPoint: z for the pinhole, |p| for the lidar
                const bool changed = tsdf_update_voxel(m, dval[k], d_ok[k], sensor.near_plane(),
                                                       sensor.measurement(pC[k]), nrm[k], tsdf, weight);
                if (changed || entry < 0)
                    VoxelRaw<Storage>::store(m, (size_t) (entry & 0x7fffffff) * 512 + t, tsdf, weight);
            }
#pragma unroll
            for (int k = 0; k < kUpdK; k++)
                cur[k] = nxt[k];
        }
    }
}

static int read_counters(const se8_map* map, int* host, hipStream_t s)
{
    SE8_HIP_CHECK(hipMemcpyAsync(host, map->counters.ptr, sizeof(int) * kNumCounters, hipMemcpyDeviceToHost, s));
    SE8_HIP_CHECK(hipStreamSynchronize(s));
    return SE8_OK;
}

static int ensure_candidates(se8_map* map, size_t n)
{
    if (n <= map->candidates_cap)
        return SE8_OK;
    int rc = dev_alloc(map->candidates, n * sizeof(uint64_t));
    if (rc == SE8_OK)
        rc = dev_alloc(map->cur_node, n * sizeof(int));
    if (rc == SE8_OK)
        rc = map->sorter.reserve(n);
    if (rc != SE8_OK)
        return rc;
    map->candidates_cap = n;
    return SE8_OK;
}

// Sort the candidate keys and build nodes/blocks.  n = number of candidates (host-known).
static int device_count_sm()
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


// Sorts the candidate keys and builds nodes and blocks.  The live count is *n_dev (device memory),
// at most cap; nothing here waits for the device.
static int build_from_candidates(se8_map* map, uint64_t* cand, const int* n_dev, int cap, int* update_list,
                                 int mark_new, hipStream_t s)
{
    if (cap <= 0)
        return SE8_OK;
    const int hi_bit = kScaleOffset + 3 * map->view.max_scale;
    const int lo_bit = kScaleOffset + 3 * kBlockScale;
    uint64_t* sorted = radix_sort(map->sorter, cand, n_dev, cap, lo_bit, hi_bit, s);
    const int brick_scale = map->view.brick_shift + kBlockScale;
    const int grid = std::min((cap + 255) / 256, device_count_sm() * 4);
    for (int scale = map->view.max_scale - 1; scale > kBlockScale; scale--) {
        k_build_level<<<grid, 256, 0, s>>>(map->view, sorted, n_dev, cap, (int*) map->cur_node.ptr, scale,
                                            brick_scale);
    }
    k_build_blocks<<<grid, 256, 0, s>>>(map->view, sorted, n_dev, cap, (const int*) map->cur_node.ptr, update_list,
                                        mark_new);
    SE8_HIP_CHECK(hipGetLastError());
    return SE8_OK;
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
    // Octree(ceil(max dim / res)) -> power_two_up(max(size, 2*BlockSize)) (octree_impl.hpp:14-19)
    const float maxdim = std::max(cfg->dim[0], std::max(cfg->dim[1], cfg->dim[2]));
    int size = (int) std::ceil(maxdim / cfg->res);
    size = std::max(size, 16);
    const int p = (int) std::ceil(std::log2((float) size));
    size = 1 << p;
    SE8_REQUIRE(p <= 19, "se8_map_create: octree deeper than the 19-scale key limit");
    MapView& v = m->view;
    std::memset(&v, 0, sizeof(v));
    v.size = size;
    v.max_scale = p;
    const int brick_scale = std::min(6, p);
    v.brick_shift = brick_scale - kBlockScale;
    v.bricks_side = size >> brick_scale;
    v.res = cfg->res;
    v.dim = (float) size * cfg->res;
    v.mu = cfg->res * tcfg->truncation_boundary_factor;
    v.max_weight = tcfg->max_weight;
    for (int i = 0; i < 3; i++)
        v.t_MW[i] = cfg->t_MW[i];
    v.storage = tcfg->storage;
    int64_t max_blocks = tcfg->max_blocks;
    if (max_blocks <= 0) {
        const int64_t blocks_side = size / 8;
        max_blocks = std::min<int64_t>(blocks_side * blocks_side * blocks_side, (int64_t) 1 << 18);
    }
    v.max_blocks = max_blocks;
    v.max_nodes = 2 * max_blocks + 64;
    const int64_t bricks_side3 = (int64_t) v.bricks_side * v.bricks_side * v.bricks_side;
    v.max_bricks = std::min<int64_t>(bricks_side3, max_blocks / 4 + 64);
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
    int64_t node_dense_total = 0;
    for (int sc = kBlockScale + 1; sc < p; sc++) {
        v.node_dense_off[sc] = node_dense_total;
        const int64_t sd = (int64_t) size >> sc;
        node_dense_total += sd * sd * sd;
    }
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
    const float band = 2.0f * m->view.res * m->tsdf.truncation_boundary_factor;
    const int num_steps = (int) std::ceil(band / (2.f * m->view.res));
    const size_t cand_cap = (size_t) sensor.width() * sensor.height() * num_steps;
    // maps with the dense block and node tables allocate in place (k_carve_claim + k_link_new);
    // others sort the candidate keys and build the octree level by level
    const bool claim = m->view.block_dense && m->view.node_dense;
    int rc = claim ? SE8_OK : ensure_candidates(m, cand_cap);
    if (rc != SE8_OK)
        return rc;

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
    k_frame_begin<<<1, 64, 0, s>>>(m->view, T_arg, T16_dev, fp);
    const int cus = device_count_sm();
    k_frustum<Sensor><<<cus * 4, 256, 0, s>>>(m->view, fp, sensor, (int*) m->update_list.ptr);
    if (timed)
        SE8_HIP_CHECK(hipEventRecord(m->ev[1], s));
    dim3 grid((sensor.width() + 63) / 64, sensor.height());
    if (claim) {
        k_carve_claim<Sensor><<<grid, 64, 0, s>>>(m->view, sensor, fp, depth_dev, band, num_steps,
                                                  (int*) m->update_list.ptr);
        SE8_HIP_CHECK(hipGetLastError());
        if (timed)
            SE8_HIP_CHECK(hipEventRecord(m->ev[2], s));
        k_link_new<<<cus, 256, 0, s>>>(m->view, (const int*) m->update_list.ptr);
        SE8_HIP_CHECK(hipGetLastError());
    }
    else {
        k_alloc_rays<Sensor><<<grid, 64, 0, s>>>(m->view, sensor, fp, depth_dev, band, num_steps,
                                                 (uint64_t*) m->candidates.ptr, (int64_t) cand_cap);
        SE8_HIP_CHECK(hipGetLastError());
        if (timed)
            SE8_HIP_CHECK(hipEventRecord(m->ev[2], s));
        rc = build_from_candidates(m, (uint64_t*) m->candidates.ptr, m->view.counters + kCntCandidates,
                                   (int) cand_cap, (int*) m->update_list.ptr, 1, s);
        if (rc != SE8_OK)
            return rc;
    }
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
    else if (m->view.storage == SE8_STORAGE_F32) {
        // one resident wave of workgroups: the grid-stride loop then never leaves a second, partial
        // round of workgroups waiting for the first to finish
        hipExtLaunchKernelGGL(k_update<SE8_STORAGE_F32, Sensor>,
                              dim3(resident_grid(k_update<SE8_STORAGE_F32, Sensor>, 512)), dim3(512), 0, s, ev_upd0,
                              ev_upd1, 0, m->view, ulist, (const FramePoses*) fp, sensor, depth_dev, timestamp);
    }
    else {
        hipExtLaunchKernelGGL(k_update<SE8_STORAGE_F16, Sensor>,
                              dim3(resident_grid(k_update<SE8_STORAGE_F16, Sensor>, 512)), dim3(512), 0, s, ev_upd0,
                              ev_upd1, 0, m->view, ulist, (const FramePoses*) fp, sensor, depth_dev, timestamp);
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
