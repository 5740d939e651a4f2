// Allocation pipeline of integrateDepth (map_integrator_impl.hpp:62-70: fetch-frustum, create-list,
// allocate-list) and of allocator::blocks, as kernels and the host code that enqueues them.
// Included by se8_map.hip only; launches go through hipLaunchKernelGGL so the same source also
// compiles as host code for the CPU bounds checks of the tests.
//
//   k_frame_begin  per-frame counters cleared; T_WS, its inverse and T_CV derived from the pose
//   k_frustum      FrustumIterator over the allocated blocks (raycast_carver_impl.hpp:17-32,
//                  iterator.hpp:197-238) as a flat per-block test of the block and its ancestors
//   dense tables (maps of at most 4096 voxels a side):
//     k_carve_claim  the RaycastCarver band march (raycast_carver_impl.hpp:74-111) that creates
//                    each missing block and its ancestors in place by claiming dense-table entries
//     k_link_new     child links, brick entries and the AABB of the frame's new octants
//   otherwise:
//     k_alloc_rays   the band march appending the keys of unallocated blocks to a candidate list
//     radix sort     LSD sort of the candidate keys (11-bit digits over the significant key bits)
//     k_build_level  one launch per octree scale: segment heads of the sorted keys create missing
//                    nodes (allocator_impl.hpp:68-147, octree_impl.hpp:171-199)
//     k_build_blocks block slots, brick index, AABB and the "new" entries of the update list
#pragma once

#include <algorithm>
#include <cmath>

#include "se8_internal.hpp"

namespace se8 {

// Start of an integration: clears the per-frame counters and derives the frame's poses from T_WS,
// read from device memory when T16_dev is set, else taken from the argument.
__global__ void k_frame_begin(MapView m, Iso T_arg, const float* T16_dev, FramePoses* fp)
{
    if (threadIdx.x == 0)
        frame_begin_state(m.counters, fp, m.t_MW, m.res, T16_dev ? iso_from16(T16_dev) : T_arg);
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

// ----------------------------------------------------------------------------------------------
// fetch-frustum: a block is listed iff its own sphere and those of all its ancestors below the
// root pass sphereInFrustum (the DFS of iterator_impl.hpp:86-156 never prunes at the root).
// ----------------------------------------------------------------------------------------------
// The blocks allocated before this integration (slots below kCntBlocksStart; slots above it may be
// in creation while this runs), `wg` of `n_wg` workgroups of 64-thread waves grid-striding over them,
// kFrustumBatch blocks per thread per round: the round's keys are loaded together and its blocks in
// the frustum take their list positions with one atomic per wave.
constexpr int kFrustumBatch = 4;

template<class Sensor>
__device__ void frustum_part(const MapView& m, const FramePoses* fp, const Sensor& sensor, int* update_list, int wg,
                             int n_wg)
{
    // the block counter runs past the pool when an allocation overflows; only slots below the
    // capacity hold blocks
    const int n = (int) min((int64_t) m.counters[kCntBlocksStart], m.max_blocks);
    const Iso T_SW = fp->T_SW;
    const float r_unit = sqrtf(3.0f) / 2.0f * m.res;
    const int stride = n_wg * (int) blockDim.x;
    // base is the same in every lane, so every lane of a wave takes the reservation together
    for (int base = wg * (int) blockDim.x; base < n; base += kFrustumBatch * stride) {
        const int slot0 = base + (int) threadIdx.x;
        uint64_t key[kFrustumBatch];
#pragma unroll
        for (int j = 0; j < kFrustumBatch; j++) {
            const int slot = slot0 + j * stride;
            key[j] = slot < n ? m.block_key[slot] : 0;
        }
        bool in[kFrustumBatch];
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < kFrustumBatch; j++) {
            in[j] = slot0 + j * stride < n;
            if (in[j]) {
                int c[3];
                key_to_coord(key[j], c);
                for (int s = kBlockScale; s < m.max_scale && in[j]; s++) {
                    const int size = 1 << s;
                    float cW[3];
                    for (int i = 0; i < 3; i++) {
                        const int oc = c[i] & ~(size - 1);
                        const float pM = ((float) oc + (float) size * 0.5f) * m.res;
                        cW[i] = pM + (-m.t_MW[i]);
                    }
                    float cS[3];
                    iso_apply(T_SW, cW, cS);
                    in[j] = sensor.sphere_in_frustum(cS, r_unit * (float) size);
                }
            }
            cnt += in[j] ? 1 : 0;
        }
        int pos = wave_reserve(cnt, &m.counters[kCntUpdate]);
#pragma unroll
        for (int j = 0; j < kFrustumBatch; j++)
            if (in[j])
                update_list[pos++] = slot0 + j * stride;
    }
}

template<class Sensor>
__global__ void k_frustum(MapView m, const FramePoses* fp, Sensor sensor, int* update_list)
{
    frustum_part(m, fp, sensor, update_list, blockIdx.x, gridDim.x);
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
                    // pM / res as IEEE division gives it for pM >= 2^-60 (div_by_recip); below
                    // that both truncate to voxel 0
                    const int wrap = m.size - 1; // see lookup_block
                    vox[j][0] = ((int) div_by_recip(pM[0], m.res, m.inv_res)) & wrap;
                    vox[j][1] = ((int) div_by_recip(pM[1], m.res, m.inv_res)) & wrap;
                    vox[j][2] = ((int) div_by_recip(pM[2], m.res, m.inv_res)) & wrap;
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

// The carver's band march for the pixels [64 * bx, +64) of row y (see k_frustum_carve).
template<class Sensor>
__device__ void carve_part(const MapView& m, const Sensor& sensor, const FramePoses* fp, const float* depth, float band,
                           int num_steps, int* update_list, int bx, int y)
{
    const Iso T_WS = fp->T_WS;
    const int x = bx * blockDim.x + threadIdx.x;
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
    int prev = -1; // dense index of the previous unallocated sample's block (repeats are dropped)
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
                    // pM / res as IEEE division gives it for pM >= 2^-60 (div_by_recip); below
                    // that both truncate to voxel 0
                    const int wrap = m.size - 1; // see lookup_block
                    vox[j][0] = ((int) div_by_recip(pM[0], m.res, m.inv_res)) & wrap;
                    vox[j][1] = ((int) div_by_recip(pM[1], m.res, m.inv_res)) & wrap;
                    vox[j][2] = ((int) div_by_recip(pM[2], m.res, m.inv_res)) & wrap;
                    inb[j] = true;
                }
                for (int i = 0; i < 3; i++)
                    pos[i] = pos[i] + step[i];
            }
        }
        // block positions as 32-bit dense indices (at most 512^3 on a dense map) and keys only for
        // the blocks this thread creates: the fewer registers, the more of the carver's waves run
        int slot[kBatch];
        int bi[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            bi[j] = ((vox[j][2] >> 3) * side + (vox[j][1] >> 3)) * side + (vox[j][0] >> 3);
            slot[j] = m.block_dense[bi[j]];
        }
        // the claims of the batch are issued together and awaited once: which samples claim depends
        // only on the lookups (an entry already claimed in this kernel is being created by its
        // claimant), so every compare-and-swap is in flight before any result is read
        bool need[kBatch];
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            need[j] = false;
            if (inb[j] && slot[j] < 0) {
                if (bi[j] != prev) {
                    cnt++;
                    need[j] = slot[j] == -1;
                }
                prev = bi[j];
            }
        }
        int cas[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            cas[j] = 0;
            if (need[j])
                cas[j] = atomicCAS(&m.block_dense[bi[j]], -1, kClaimed);
        }
        // every result in its own register past this point, so no result is tested (and awaited)
        // inside the branch that issued it
#pragma unroll
        for (int j = 0; j < kBatch; j++)
            SE8_REG_BARRIER(cas[j]);
        bool won[kBatch];
        int wins = 0;
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            won[j] = need[j] && cas[j] == -1;
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
            const uint64_t key = block_key_of_voxel(vox[j][0], vox[j][1], vox[j][2]);
            m.block_key[sl] = key;
            m.block_ts[sl] = -1;
            update_list[ubase + u++] = (int) ((uint32_t) sl | 0x80000000u);
            atomicExch(&m.block_dense[bi[j]], sl);
            const int c[3] = {vox[j][0] & ~7, vox[j][1] & ~7, vox[j][2] & ~7};
            claim_ancestors(m, c, key >> kScaleOffset);
        }
    }
}

// create-list + allocate-list for maps with the dense tables: the band march of RaycastCarver
// (raycast_carver_impl.hpp:74-111), as k_alloc_rays samples it, with the allocation of
// allocator::blocks / Octree::allocate (allocator_impl.hpp:68-147, octree_impl.hpp:171-199) done in
// place: the first thread to reach an unallocated block claims its dense-table entry, creates the
// block (slot, key, timestamp, an update-list entry marked new) and its missing ancestors.  The set
// of blocks and nodes is that of the sorted-key path; slot and node numbering follow the schedule.
// kCntCandidates counts the samples' block keys of unallocated blocks (consecutive repeats dropped).
// The first n_frustum_wg workgroups run fetch-frustum over the blocks that existed before, in the
// same launch: their list entries and the carver's interleave (the new ones carry bit 31).
// 64-thread workgroups: carver workgroup b - n_frustum_wg covers pixels
// [64 * (b' % tiles_x), +64) of row b' / tiles_x.
template<class Sensor>
__global__ void k_frustum_carve(MapView m, Sensor sensor, const FramePoses* fp, const float* depth, float band,
                                int num_steps, int* update_list, int n_frustum_wg, int tiles_x)
{
    const int b = blockIdx.x;
    if (b < n_frustum_wg) {
        frustum_part(m, fp, sensor, update_list, b, n_frustum_wg);
        return;
    }
    const int c = b - n_frustum_wg;
    carve_part(m, sensor, fp, depth, band, num_steps, update_list, c % tiles_x, c / tiles_x);
}

// The links of the octants the carver created: each new node becomes its parent's child (the root
// for the top scale), each new block its scale-4 parent's child (block_parent, whose timestamp it
// raises to ts when ts >= 0), its brick's slot entry and part of the AABB (octree_impl.hpp:267-273);
// the frustum part of the update list is the rest of it.  Thread gt of gn threads in whole waves
// (the AABB is aggregated by ballots).
__device__ void link_part(const MapView& m, const int* update_list, int gt, int gn, int ts)
{
    const int n_upd = m.counters[kCntUpdate];
    if (gt == 0)
        m.counters[kCntFrustum] = n_upd - m.counters[kCntUnique];
    const int n0 = m.counters[kCntNodesStart];
    const int n1 = (int) min((int64_t) m.counters[kCntNodes], m.max_nodes);
    for (int i = n0 + gt; i < n1; i += gn) {
        const uint64_t key = m.node_key[i];
        const int s = (int) (key & 0x1f);
        const uint64_t code = key >> kScaleOffset;
        int c[3];
        key_to_coord(key, c);
        const int parent = s + 1 >= m.max_scale ? 0 : m.node_dense[m.node_dense_off[s + 1] + dense_index(m, c, s + 1)];
        if (parent >= 0)
            m.node_child[parent * 8 + (int) ((code >> (3 * s)) & 7)] = i;
    }
    const int lane0 = gt - (int) (threadIdx.x & 63);
    for (int base = lane0; base < n_upd; base += gn) {
        const int i = base + (int) (threadIdx.x & 63);
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
                if (parent >= 0) {
                    m.node_child[parent * 8 + (int) ((code >> 9) & 7)] = slot;
                    if (ts >= 0)
                        atomicMax(&m.node_ts[parent], ts);
                }
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

// link_part as its own launch (multi-resolution maps, whose updater reads block_parent); the update
// kernel of a single-resolution map runs link_part itself
__global__ void k_link_new(MapView m, const int* update_list)
{
    link_part(m, update_list, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, -1);
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

inline int RadixSorter::reserve(size_t n)
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

inline void RadixSorter::release()
{
    dev_free(keys_alt);
    dev_free(tile_hist);
    cap = 0;
}

// Sorts keys[0..min(*n_dev, cap)) on bit range [lo_bit, hi_bit); returns the buffer holding the
// result.  The number of passes is fixed by the bit range, so the result buffer is known here.
inline uint64_t* radix_sort(RadixSorter& rs, uint64_t* keys, const int* n_dev, int cap, int lo_bit, int hi_bit,
                            hipStream_t s)
{
    const int max_tiles = (cap + kRsTile - 1) / kRsTile;
    uint64_t* a = keys;
    uint64_t* b = (uint64_t*) rs.keys_alt.ptr;
    for (int shift = lo_bit; shift < hi_bit; shift += kRsBits) {
        hipLaunchKernelGGL(k_rs_hist, dim3(max_tiles), dim3(kRsThreads), 0, s, (const uint64_t*) a, n_dev, cap, shift,
                           (uint32_t*) rs.tile_hist.ptr);
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, s, (uint32_t*) rs.tile_hist.ptr, n_dev, cap);
        hipLaunchKernelGGL(k_rs_scatter, dim3(max_tiles), dim3(kRsThreads), 0, s, (const uint64_t*) a, b, n_dev, cap,
                           shift, (const uint32_t*) rs.tile_hist.ptr);
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

inline int ensure_candidates(se8_map* map, size_t n)
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

// Sorts the candidate keys and builds nodes and blocks.  The live count is *n_dev (device memory),
// at most cap; nothing here waits for the device.
inline int build_from_candidates(se8_map* map, uint64_t* cand, const int* n_dev, int cap, int* update_list,
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
        hipLaunchKernelGGL(k_build_level, dim3(grid), dim3(256), 0, s, map->view, (const uint64_t*) sorted, n_dev, cap,
                           (int*) map->cur_node.ptr, scale, brick_scale);
    }
    hipLaunchKernelGGL(k_build_blocks, dim3(grid), dim3(256), 0, s, map->view, (const uint64_t*) sorted, n_dev, cap,
                       (const int*) map->cur_node.ptr, update_list, mark_new);
    SE8_HIP_CHECK(hipGetLastError());
    return SE8_OK;
}

// Enqueues the allocation stages of one integration on stream s: k_frame_begin, k_frustum, then
// the in-place claim path (maps with dense tables) or the band march + sort + level build.  The
// update list then holds the frame's frustum blocks and its new blocks (marked with bit 31).
// ev_frustum / ev_carve (may be null) are recorded after the frustum and the band march.
template<class Sensor>
int allocate_frame(se8_map* m, const Iso& T_arg, const float* T16_dev, const float* depth_dev, const Sensor& sensor,
                   hipEvent_t ev_frustum, hipEvent_t ev_carve, hipStream_t s, bool link_in_update = false)
{
    const float band = 2.0f * m->view.res * m->tsdf.truncation_boundary_factor;
    const int num_steps = (int) std::ceil(band / (2.f * m->view.res));
    const size_t cand_cap = (size_t) sensor.width() * sensor.height() * num_steps;
    const bool claim = m->view.block_dense && m->view.node_dense;
    FramePoses* fp = (FramePoses*) m->frame_poses.ptr;
    int* ulist = (int*) m->update_list.ptr;
    // a tracker that ended at this pose pointer already started the integration
    // (se8_track_dev_prepare_integrate); otherwise k_frame_begin does
    const bool prepared = T16_dev && T16_dev == m->prepared_T;
    m->prepared_T = nullptr;
    if (!prepared) {
        hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(64), 0, s, m->view, T_arg, T16_dev, fp);
        graph_note_launch(kNodeFrameBegin, s);
    }
    const int cus = device_count_sm();
    if (claim) {
        // fetch-frustum and the carver in one launch (ev_frustum then marks the same point as
        // ev_carve's start)
        if (ev_frustum)
            SE8_HIP_CHECK(hipEventRecord(ev_frustum, s));
        const int tiles_x = (sensor.width() + 63) / 64, n_frustum_wg = cus;
        hipLaunchKernelGGL(k_frustum_carve<Sensor>, dim3(n_frustum_wg + tiles_x * sensor.height()), dim3(64), 0, s,
                           m->view, sensor, (const FramePoses*) fp, depth_dev, band, num_steps, ulist, n_frustum_wg,
                           tiles_x);
        SE8_HIP_CHECK(hipGetLastError());
        graph_note_launch(kNodeFrustumCarve, s);
        if (ev_carve)
            SE8_HIP_CHECK(hipEventRecord(ev_carve, s));
        if (!link_in_update) {
            hipLaunchKernelGGL(k_link_new, dim3(cus), dim3(256), 0, s, m->view, (const int*) ulist);
            SE8_HIP_CHECK(hipGetLastError());
        }
        return SE8_OK;
    }
    hipLaunchKernelGGL(k_frustum<Sensor>, dim3(cus * 4), dim3(256), 0, s, m->view, (const FramePoses*) fp, sensor,
                       ulist);
    if (ev_frustum)
        SE8_HIP_CHECK(hipEventRecord(ev_frustum, s));
    const dim3 grid((sensor.width() + 63) / 64, sensor.height());
    hipLaunchKernelGGL(k_alloc_rays<Sensor>, grid, dim3(64), 0, s, m->view, sensor, (const FramePoses*) fp, depth_dev,
                       band, num_steps, (uint64_t*) m->candidates.ptr, (int64_t) cand_cap);
    SE8_HIP_CHECK(hipGetLastError());
    if (ev_carve)
        SE8_HIP_CHECK(hipEventRecord(ev_carve, s));
    return build_from_candidates(m, (uint64_t*) m->candidates.ptr, m->view.counters + kCntCandidates, (int) cand_cap,
                                 ulist, 1, s);
}

} // namespace se8
