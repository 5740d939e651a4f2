// Multi-resolution TSDF updater: Updater<Map<TSDF, Res::Multi>, PinholeCamera>
// (include/se/integrator/updater/impl/multires_tsdf_updater_impl.hpp:14-254) over the blocks of the
// update list (the same allocation and frustum fetch as the single-resolution map).
//
// Per block the reference: stamps the timestamp; picks the integration scale from the block
// centre's pixel footprint (PinholeCamera::computeIntegrationScaleImpl, pinhole_camera.cpp:115-152,
// with 0.25 m of hysteresis once the block has a minimum scale), never more than one scale finer
// than the last; when finer, propagates the data down (propagator_impl.hpp:72-117): an observed
// child takes the parent's change since its past data (delta update), an unobserved one the
// multi-resolution trilinear interpolation of the map at its scale (visitor_impl.hpp:399-460,
// 905-971); updates the voxels of the new scale as the single-resolution updater does (plus one
// past-data weight per updated voxel); and propagates up the mean of the valid children
// (propagator_impl.hpp:18-68, data_impl.hpp:14-53).
//
// The reference runs the blocks in an OpenMP loop, and a block's interpolation may read a
// neighbour block that another thread is updating in the same frame, so its result depends on
// the schedule.  This implementation fixes one legal schedule: every block's scale choice and
// down-propagation read the map as it stood before the frame's updates (the new scales are
// committed afterwards), then every block updates and propagates up.  Four launches:
//   k_mr_scales        one thread per list entry: the integration scale (kept aside), and the
//                      entries of new blocks and of blocks moving to a finer scale appended to two
//                      short lists
//   k_mr_init_new      one workgroup per new block: default data, past data and scales
//   k_mr_down          one workgroup per block moving finer: down-propagation level by level
//   k_mr_update_up     one workgroup per list entry: the entries of the block's current and coarser
//                      scales (a contiguous suffix of its data and of its past data) staged in LDS,
//                      the scale committed, the voxels of that scale updated, the coarser scales
//                      recomputed up, the suffix written back; block and parent timestamps
#include "se8_internal.hpp"

#include <algorithm>
#include <vector>

using namespace se8;

namespace {

constexpr int kMrThreads = 256;

__device__ __forceinline__ float2* mr_block(const MapView& m, int slot)
{
    return reinterpret_cast<float2*>(m.voxels) + (size_t) slot * kMrBlockFloat2;
}

__device__ __forceinline__ int cur_scale(const MapView& m, int slot)
{
    return m.block_scale[2 * slot];
}

// pixel_voxel_ratio_per_scale = {1.5, 3, 6} (sensor.hpp:56)
__device__ __forceinline__ int scale_from_depth(float z, float scaled_pixel, float res, int max_scale)
{
    const float pv = z * scaled_pixel / res;
    int scale = 0;
    if (!(pv < 1.5f)) {
        scale = 1;
        if (!(pv < 3.0f)) {
            scale = 2;
            if (!(pv < 6.0f))
                scale = 3;
        }
    }
    return min(scale, max_scale);
}

// PinholeCamera::computeIntegrationScaleImpl with its one-level hysteresis recursion.
__device__ __forceinline__ int integration_scale(float z, float scaled_pixel, float res, int last_scale,
                                                 int min_scale)
{
    const int scale = scale_from_depth(z, scaled_pixel, res, 3);
    if (scale > last_scale && min_scale != -1)
        return scale_from_depth(z - 0.25f, scaled_pixel, res, 3);
    if (scale < last_scale && min_scale != -1)
        return scale_from_depth(z + 0.25f, scaled_pixel, res, 3);
    return scale;
}

// visitor::getFieldInterp(octree, voxel_coord_f, desired_scale) for a multi-resolution map
// (visitor_impl.hpp:905-971 with get_neighbours 399-460), reading the scales in block_scale.
__device__ bool field_interp_mr(const MapView& m, const float cf[3], int desired_scale, float& out)
{
    const int leaf = lookup_block(m, (int) cf[0], (int) cf[1], (int) cf[2]);
    // a node leaf starts at scale 0 (its data is the default, so any interpolation over it fails)
    const int init_scale = leaf >= 0 ? max(cur_scale(m, leaf), desired_scale) : 0;
    for (int scale = init_scale; scale <= 3; scale++) {
        const int stride = 1 << scale;
        const float inv = 1.0f / (float) stride;
        float bf[3];
        int b[3];
        bool in_box = true, in_map = true;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            bf[i] = inv * cf[i] - 0.5f;
            b[i] = stride * (int) bf[i];
            in_box = in_box && m.counters[kCntAabb + i] <= b[i] && b[i] <= m.counters[kCntAabb + 3 + i];
            in_map = in_map && b[i] >= 0 && b[i] + stride < m.size;
        }
        if (!in_box)
            return false;
        if (!in_map)
            continue;
        float d[8];
        bool coarser = false, valid = true;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int c[3] = {b[0] + stride * (k & 1), b[1] + stride * ((k >> 1) & 1), b[2] + stride * ((k >> 2) & 1)};
            const int slot = lookup_block(m, c[0], c[1], c[2]);
            d[k] = 1.f;
            if (slot < 0) {
                valid = false;
                continue;
            }
            if (cur_scale(m, slot) > scale) {
                coarser = true;
                continue;
            }
            const float2 v = mr_block(m, slot)[mr_index(c[0] & 7, c[1] & 7, c[2] & 7, scale)];
            d[k] = v.x;
            valid = valid && v.y != 0.f;
        }
        if (coarser)
            continue;
        if (!valid)
            return false;
        float t[3], tc[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            t[i] = bf[i] - floorf(bf[i]);
            tc[i] = 1.f - t[i];
        }
        out = ((d[0] * tc[0] + d[1] * t[0]) * tc[1] + (d[2] * tc[0] + d[3] * t[0]) * t[1]) * tc[2] +
              ((d[4] * tc[0] + d[5] * t[0]) * tc[1] + (d[6] * tc[0] + d[7] * t[0]) * t[1]) * t[2];
        return true;
    }
    return false;
}

// The six lists of k_mr_scales: new blocks, blocks moving finer, entries of scale 0..3; their
// counters are consecutive from kCntMrNew.
constexpr int kMrLists = 6;

__global__ __launch_bounds__(256) void k_mr_scales(MapView m, const int* update_list, const FramePoses* fp,
                                                   float scaled_pixel, int* new_scale, int* lists, int list_cap)
{
    __shared__ int s_cnt[kMrLists], s_base[kMrLists];
    const int n = m.counters[kCntUpdate];
    const int lane = __lane_id();
    const int iters = (n + (int) (gridDim.x * blockDim.x) - 1) / (int) (gridDim.x * blockDim.x);
    for (int it = 0; it < iters; it++) {
        const int e = (it * (int) gridDim.x + (int) blockIdx.x) * (int) blockDim.x + (int) threadIdx.x;
        bool is_new = false, finer = false;
        int s = -1;
        if (e < n) {
            const int entry = update_list[e];
            const int slot = entry & 0x7fffffff;
            is_new = entry < 0;
            int bc[3];
            key_to_coord(m.block_key[slot], bc);
            // block centre: T_WM * ((coord + 8 * 0.5) * res), in the camera frame
            float cW[3], cC[3];
            for (int i = 0; i < 3; i++)
                cW[i] = ((float) bc[i] + 4.0f) * m.res - m.t_MW[i];
            iso_apply(fp->T_SW, cW, cC);
            const int last = is_new ? -1 : m.block_scale[2 * slot];
            const int mins = is_new ? -1 : m.block_scale[2 * slot + 1];
            s = max(integration_scale(cC[2], scaled_pixel, m.res, last, mins), last - 1);
            new_scale[e] = s;
            finer = s < last;
        }
        const bool in_list[kMrLists] = {is_new, finer, s == 0, s == 1, s == 2, s == 3};
        // positions: lane rank within the wave, wave offset within the workgroup (LDS atomics),
        // workgroup base (one global atomic per list and workgroup)
        if (threadIdx.x < kMrLists)
            s_cnt[threadIdx.x] = 0;
        __syncthreads();
        int pos[kMrLists];
#pragma unroll
        for (int j = 0; j < kMrLists; j++) {
            const uint64_t mask = __ballot(in_list[j]);
            int off = 0;
            if (mask != 0) {
                const int leader = __ffsll((unsigned long long) mask) - 1;
                if (lane == leader)
                    off = atomicAdd(&s_cnt[j], __popcll(mask));
                off = __shfl(off, leader);
            }
            pos[j] = off + __popcll(mask & ((1ull << lane) - 1ull));
        }
        __syncthreads();
        if (threadIdx.x < kMrLists)
            s_base[threadIdx.x] = s_cnt[threadIdx.x] > 0 ? atomicAdd(&m.counters[kCntMrNew + threadIdx.x], s_cnt[threadIdx.x]) : 0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kMrLists; j++)
            if (in_list[j])
                lists[(size_t) j * list_cap + s_base[j] + pos[j]] = e;
        __syncthreads(); // s_cnt and s_base are reused by the next iteration
    }
}

__global__ __launch_bounds__(kMrThreads) void k_mr_init_new(MapView m, const int* update_list, const int* new_list)
{
    const int n = m.counters[kCntMrNew];
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int slot = update_list[new_list[i]] & 0x7fffffff;
        float2* blk = mr_block(m, slot);
        for (int j = threadIdx.x; j < kMrBlockFloat2; j += kMrThreads)
            blk[j] = make_float2(1.f, 0.f);
        if (threadIdx.x == 0) {
            m.block_scale[2 * slot] = -1;
            m.block_scale[2 * slot + 1] = -1;
        }
    }
}

__global__ __launch_bounds__(kMrThreads) void k_mr_down(MapView m, const int* update_list, const int* down_list,
                                                         const int* new_scale)
{
    const int n = m.counters[kCntMrDown];
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int e = down_list[i];
        const int slot = update_list[e] & 0x7fffffff;
        const int last = m.block_scale[2 * slot], curr = new_scale[e];
        int bc[3];
        key_to_coord(m.block_key[slot], bc);
        float2* blk = mr_block(m, slot);
        float2* past = blk + kMrEntries;
        for (int ps = last; ps > curr; ps--) {
            const int cs = ps - 1, cn = 8 >> cs, cstride = 1 << cs;
            // children of every parent of this level
            for (int j = threadIdx.x; j < cn * cn * cn; j += kMrThreads) {
                const int lx = (j % cn) * cstride, ly = ((j / cn) % cn) * cstride, lz = (j / (cn * cn)) * cstride;
                const int ci = mr_index(lx, ly, lz, cs), pi = mr_index(lx, ly, lz, ps);
                const float2 p = blk[pi], pp = past[pi];
                float2 c = blk[ci], cp = past[ci];
                if (c.y != 0.f) {
                    // delta update: std::max(tsdf + delta, -1) keeps a NaN, fminf does not
                    const float v = c.x + (p.x - pp.x);
                    c.x = v < -1.f ? -1.f : v;
                    c.y = fminf(c.y + pp.y, m.max_weight);
                    cp.y = pp.y;
                }
                else {
                    const float cf[3] = {(float) (bc[0] + lx) + 0.5f * (float) cstride,
                                         (float) (bc[1] + ly) + 0.5f * (float) cstride,
                                         (float) (bc[2] + lz) + 0.5f * (float) cstride};
                    float f;
                    if (field_interp_mr(m, cf, cs, f)) {
                        c.x = f;
                        c.y = p.y;
                        cp = make_float2(f, 0.f);
                    }
                }
                blk[ci] = c;
                past[ci] = cp;
            }
            __syncthreads();
            // parent_down_funct: the parent's past data restarts from its current value
            const int pn = 8 >> ps;
            for (int j = threadIdx.x; j < pn * pn * pn; j += kMrThreads) {
                const int pi = mr_offset(ps) + j;
                past[pi] = make_float2(blk[pi].x, 0.f);
            }
            __syncthreads();
        }
    }
}

// One list entry of scale S per workgroup of T threads: the suffix of entries of scales S..3 (data
// and past data) in LDS, the (8 >> S)^3 voxels of scale S updated (all depth samples of a thread in
// flight together), the coarser scales recomputed up, the suffix written back.
template <int S, int T>
__global__ __launch_bounds__(T) void k_mr_update_up(MapView m, const int* update_list, const int* list,
                                                    const FramePoses* fp, PinholeSensor sensor, const float* depth,
                                                    int ts)
{
    constexpr int kOff = S == 0 ? 0 : (S == 1 ? 512 : (S == 2 ? 576 : 584));
    constexpr int kLen = kMrEntries - kOff;
    constexpr int kSn = 8 >> S, kVox = kSn * kSn * kSn, kStride = 1 << S;
    constexpr int kPer = (kVox + T - 1) / T;
    __shared__ float2 s_data[kLen], s_past[kLen];
    const Iso T_CV = fp->T_CV;
    const int n = m.counters[kCntMrScale + S];
    for (int li = blockIdx.x; li < n; li += gridDim.x) {
        const int e = list[li];
        const int slot = update_list[e] & 0x7fffffff;
        int bc[3];
        key_to_coord(m.block_key[slot], bc);
        float2* blk = mr_block(m, slot);
        float2* past = blk + kMrEntries;
        __syncthreads(); // the previous entry's suffix has been written back
        for (int i = threadIdx.x; i < kLen; i += T) {
            s_data[i] = blk[kOff + i];
            s_past[i] = past[kOff + i];
        }
        if (threadIdx.x == 0) {
            m.block_ts[slot] = ts;
            atomicMax(&m.node_ts[m.block_parent[slot]], ts);
            const int mins = m.block_scale[2 * slot + 1];
            m.block_scale[2 * slot] = S;
            m.block_scale[2 * slot + 1] = mins < 0 ? S : min(mins, S);
        }
        float d[kPer], mz[kPer], nrm[kPer];
        bool ok[kPer];
#pragma unroll
        for (int h = 0; h < kPer; h++) {
            const int i = threadIdx.x + h * T;
            ok[h] = false;
            d[h] = mz[h] = nrm[h] = 0.f;
            if (i < kVox) {
                const int lx = (i % kSn) * kStride, ly = ((i / kSn) % kSn) * kStride, lz = (i / (kSn * kSn)) * kStride;
                const float v[3] = {(float) (bc[0] + lx), (float) (bc[1] + ly), (float) (bc[2] + lz)};
                float pC[3];
                iso_apply(T_CV, v, pC);
                nrm[h] = sqrtf(sqnorm3(pC));
                float uv[2];
                if (!sensor.beyond_far(pC, nrm[h]) && sensor.project(pC, uv)) {
                    const int px = (int) (uv[0] + 0.5f), py = (int) (uv[1] + 0.5f);
                    d[h] = depth[px + (size_t) py * sensor.width()];
                    mz[h] = sensor.measurement(pC);
                    ok[h] = true;
                }
            }
        }
        __syncthreads(); // the suffix is in LDS
#pragma unroll
        for (int h = 0; h < kPer; h++) {
            const int i = threadIdx.x + h * T;
            if (!ok[h] || d[h] < sensor.near_plane())
                continue;
            // the entries of scale S are the suffix's first kVox, in the same x-fastest order
            float2 c = s_data[i];
            tsdf_update_voxel(m, d[h], true, sensor.near_plane(), mz[h], nrm[h], c.x, c.y);
            s_data[i] = c;
            s_past[i].y = s_past[i].y + 1.f;
        }
        __syncthreads();
        // up-propagation: the mean of the valid children, level by level
#pragma unroll
        for (int cs = S; cs < 3; cs++) {
            const int ps = cs + 1, pn = 8 >> ps, cstride = 1 << cs;
            for (int i = threadIdx.x; i < pn * pn * pn; i += T) {
                const int px0 = (i % pn) << ps, py0 = ((i / pn) % pn) << ps, pz0 = (i / (pn * pn)) << ps;
                float ts_sum = 0.f, w_sum = 0.f;
                int nv = 0;
                for (int k = 0; k < 8; k++) {
                    const float2 c = s_data[mr_index(px0 + cstride * (k & 1), py0 + cstride * ((k >> 1) & 1),
                                                     pz0 + cstride * ((k >> 2) & 1), cs) - kOff];
                    if (c.y != 0.f) {
                        ts_sum += c.x;
                        w_sum += c.y;
                        nv++;
                    }
                }
                const int pi = mr_index(px0, py0, pz0, ps) - kOff;
                if (nv > 0) {
                    const float t = ts_sum / (float) nv;
                    s_data[pi] = make_float2(t, ceilf(w_sum / (float) nv));
                    s_past[pi] = make_float2(t, 0.f);
                }
                else
                    s_past[pi] = make_float2(1.f, 0.f);
            }
            __syncthreads();
        }
        for (int i = threadIdx.x; i < kLen; i += T) {
            blk[kOff + i] = s_data[i];
            past[kOff + i] = s_past[i];
        }
    }
}

} // namespace

int se8::launch_multires_update(se8_map* m, const FramePoses* fp, const PinholeSensor& sensor, const float* depth,
                                int ts, hipStream_t s)
{
    // per list entry: the new scale; the lists of new blocks, of blocks moving finer, and of the
    // entries of each scale
    int rc = dev_alloc(m->mr_new_scale, sizeof(int) * (1 + kMrLists) * (size_t) m->view.max_blocks);
    if (rc != SE8_OK)
        return rc;
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        SE8_HIP_CHECK(hipGetDevice(&dev));
        SE8_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    const int* ulist = (const int*) m->update_list.ptr;
    int* ns = (int*) m->mr_new_scale.ptr;
    int* new_list = ns + m->view.max_blocks;
    int* down_list = new_list + m->view.max_blocks;
    int* scale_lists = down_list + m->view.max_blocks;
    const int cap = (int) m->view.max_blocks;
    const unsigned grid = (unsigned) (cus * 8);
    k_mr_scales<<<grid, 256, 0, s>>>(m->view, ulist, fp, sensor.c.inv_fx, ns, new_list, cap);
    k_mr_init_new<<<grid, kMrThreads, 0, s>>>(m->view, ulist, new_list);
    k_mr_down<<<grid, kMrThreads, 0, s>>>(m->view, ulist, down_list, ns);
    // scale 0: 512 voxels per block, 256 threads; coarser scales: one wave per block, 4x as many
    // blocks in flight per CU
    k_mr_update_up<0, 256><<<grid, 256, 0, s>>>(m->view, ulist, scale_lists, fp, sensor, depth, ts);
    k_mr_update_up<1, 64><<<grid * 4, 64, 0, s>>>(m->view, ulist, scale_lists + cap, fp, sensor, depth, ts);
    k_mr_update_up<2, 64><<<grid * 4, 64, 0, s>>>(m->view, ulist, scale_lists + 2 * (size_t) cap, fp, sensor, depth, ts);
    k_mr_update_up<3, 64><<<grid * 4, 64, 0, s>>>(m->view, ulist, scale_lists + 3 * (size_t) cap, fp, sensor, depth, ts);
    SE8_HIP_CHECK(hipGetLastError());
    return SE8_OK;
}

extern "C" int se8_dump_blocks_multires(const se8_map* m, uint64_t* keys, int* scales, float* data, float* past_data,
                                        int64_t cap, int64_t* out_n)
{
    SE8_REQUIRE(m && out_n, "se8_dump_blocks_multires: null argument");
    SE8_REQUIRE(m->view.storage == SE8_STORAGE_MULTIRES, "se8_dump_blocks_multires: not a multi-resolution map");
    int host[kNumCounters];
    SE8_HIP_CHECK(hipMemcpy(host, m->counters.ptr, sizeof(host), hipMemcpyDeviceToHost));
    const int64_t n = std::min<int64_t>(host[kCntBlocks], m->view.max_blocks);
    *out_n = n;
    if (!keys)
        return SE8_OK;
    SE8_REQUIRE(cap >= n, "se8_dump_blocks_multires: output capacity too small");
    std::vector<uint64_t> k(n);
    std::vector<int> sc(2 * n);
    std::vector<float> vox((size_t) n * 2 * kMrBlockFloat2);
    if (n > 0) {
        SE8_HIP_CHECK(hipMemcpy(k.data(), m->block_key.ptr, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
        SE8_HIP_CHECK(hipMemcpy(sc.data(), m->block_scale.ptr, sizeof(int) * 2 * n, hipMemcpyDeviceToHost));
        SE8_HIP_CHECK(hipMemcpy(vox.data(), m->voxels.ptr, sizeof(float) * vox.size(), hipMemcpyDeviceToHost));
    }
    std::vector<int64_t> order(n);
    for (int64_t i = 0; i < n; i++)
        order[i] = i;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return k[a] < k[b]; });
    for (int64_t i = 0; i < n; i++) {
        const int64_t s = order[i];
        keys[i] = k[s];
        if (scales) {
            scales[2 * i] = sc[2 * s];
            scales[2 * i + 1] = sc[2 * s + 1];
        }
        const float* src = vox.data() + (size_t) s * 2 * kMrBlockFloat2;
        if (data)
            std::copy(src, src + 2 * kMrEntries, data + (size_t) i * 2 * kMrEntries);
        if (past_data)
            std::copy(src + 2 * kMrEntries, src + 4 * kMrEntries, past_data + (size_t) i * 2 * kMrEntries);
    }
    return SE8_OK;
}
