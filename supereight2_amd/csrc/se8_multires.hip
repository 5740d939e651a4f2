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
// committed afterwards), then every block updates and propagates up.  Three launches, one
// workgroup per list entry at a time (grid-stride):
//   k_mr_init_new      default data, past data and scales of the blocks allocated this frame
//   k_mr_scale_down    integration scale (kept aside), down-propagation level by level
//   k_mr_update_up     commit the scale, update the voxels of that scale, up-propagation,
//                      block and parent timestamps
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

__global__ __launch_bounds__(kMrThreads) void k_mr_init_new(MapView m, const int* update_list)
{
    const int n = m.counters[kCntUpdate];
    for (int e = blockIdx.x; e < n; e += gridDim.x) {
        const int entry = update_list[e];
        if (entry >= 0)
            continue; // blocks from earlier frames keep their data
        const int slot = entry & 0x7fffffff;
        float2* blk = mr_block(m, slot);
        for (int i = threadIdx.x; i < kMrBlockFloat2; i += kMrThreads)
            blk[i] = make_float2(1.f, 0.f);
        if (threadIdx.x == 0) {
            m.block_scale[2 * slot] = -1;
            m.block_scale[2 * slot + 1] = -1;
        }
    }
}

__global__ __launch_bounds__(kMrThreads) void k_mr_scale_down(MapView m, const int* update_list, const FramePoses* fp,
                                                               float scaled_pixel, int* new_scale)
{
    __shared__ int s_scale[2];
    const int n = m.counters[kCntUpdate];
    for (int e = blockIdx.x; e < n; e += gridDim.x) {
        const int slot = update_list[e] & 0x7fffffff;
        int bc[3];
        key_to_coord(m.block_key[slot], bc);
        __syncthreads(); // s_scale of the previous entry is no longer read
        if (threadIdx.x == 0) {
            // block centre: T_WM * ((coord + 8 * 0.5) * res), in the camera frame
            float cW[3], cC[3];
            for (int i = 0; i < 3; i++)
                cW[i] = ((float) bc[i] + 4.0f) * m.res - m.t_MW[i];
            iso_apply(fp->T_SW, cW, cC);
            const int last = m.block_scale[2 * slot], mins = m.block_scale[2 * slot + 1];
            const int s = max(integration_scale(cC[2], scaled_pixel, m.res, last, mins), last - 1);
            new_scale[e] = s;
            s_scale[0] = last;
            s_scale[1] = s;
        }
        __syncthreads();
        const int last = s_scale[0], curr = s_scale[1];
        if (curr >= last)
            continue;
        float2* blk = mr_block(m, slot);
        float2* past = blk + kMrEntries;
        for (int ps = last; ps > curr; ps--) {
            const int cs = ps - 1, cn = 8 >> cs, cstride = 1 << cs;
            // children of every parent of this level
            for (int i = threadIdx.x; i < cn * cn * cn; i += kMrThreads) {
                const int lx = (i % cn) * cstride, ly = ((i / cn) % cn) * cstride, lz = (i / (cn * cn)) * cstride;
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
            for (int i = threadIdx.x; i < pn * pn * pn; i += kMrThreads) {
                const int pi = mr_offset(ps) + i;
                past[pi] = make_float2(blk[pi].x, 0.f);
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(kMrThreads) void k_mr_update_up(MapView m, const int* update_list, const FramePoses* fp,
                                                              PinholeSensor sensor, const float* depth, int ts,
                                                              const int* new_scale)
{
    const Iso T_CV = fp->T_CV;
    const int n = m.counters[kCntUpdate];
    for (int e = blockIdx.x; e < n; e += gridDim.x) {
        const int slot = update_list[e] & 0x7fffffff;
        const int curr = new_scale[e];
        int bc[3];
        key_to_coord(m.block_key[slot], bc);
        __syncthreads(); // the previous entry's block is complete
        if (threadIdx.x == 0) {
            m.block_ts[slot] = ts;
            atomicMax(&m.node_ts[m.block_parent[slot]], ts);
            const int mins = m.block_scale[2 * slot + 1];
            m.block_scale[2 * slot] = curr;
            m.block_scale[2 * slot + 1] = mins < 0 ? curr : min(mins, curr);
        }
        float2* blk = mr_block(m, slot);
        float2* past = blk + kMrEntries;
        const int sn = 8 >> curr, stride = 1 << curr;
        for (int i = threadIdx.x; i < sn * sn * sn; i += kMrThreads) {
            const int lx = (i % sn) * stride, ly = ((i / sn) % sn) * stride, lz = (i / (sn * sn)) * stride;
            const float v[3] = {(float) (bc[0] + lx), (float) (bc[1] + ly), (float) (bc[2] + lz)};
            float pC[3];
            iso_apply(T_CV, v, pC);
            const float nrm = sqrtf(sqnorm3(pC));
            float uv[2];
            if (sensor.beyond_far(pC, nrm) || !sensor.project(pC, uv))
                continue;
            const int px = (int) (uv[0] + 0.5f), py = (int) (uv[1] + 0.5f);
            const float d = depth[px + (size_t) py * sensor.width()];
            if (d < sensor.near_plane())
                continue;
            const int vi = mr_index(lx, ly, lz, curr);
            float2 c = blk[vi];
            tsdf_update_voxel(m, d, true, sensor.near_plane(), sensor.measurement(pC), nrm, c.x, c.y);
            blk[vi] = c;
            past[vi].y = past[vi].y + 1.f;
        }
        __syncthreads();
        // up-propagation: the mean of the valid children, level by level
        for (int cs = curr; cs < 3; cs++) {
            const int ps = cs + 1, pn = 8 >> ps, cstride = 1 << cs;
            for (int i = threadIdx.x; i < pn * pn * pn; i += kMrThreads) {
                const int px0 = (i % pn) << ps, py0 = ((i / pn) % pn) << ps, pz0 = (i / (pn * pn)) << ps;
                float ts_sum = 0.f, w_sum = 0.f;
                int nv = 0;
                for (int k = 0; k < 8; k++) {
                    const float2 c = blk[mr_index(px0 + cstride * (k & 1), py0 + cstride * ((k >> 1) & 1),
                                                  pz0 + cstride * ((k >> 2) & 1), cs)];
                    if (c.y != 0.f) {
                        ts_sum += c.x;
                        w_sum += c.y;
                        nv++;
                    }
                }
                const int pi = mr_index(px0, py0, pz0, ps);
                if (nv > 0) {
                    const float t = ts_sum / (float) nv;
                    blk[pi] = make_float2(t, ceilf(w_sum / (float) nv));
                    past[pi] = make_float2(t, 0.f);
                }
                else
                    past[pi] = make_float2(1.f, 0.f);
            }
            __syncthreads();
        }
    }
}

} // namespace

int se8::launch_multires_update(se8_map* m, const FramePoses* fp, const PinholeSensor& sensor, const float* depth,
                                int ts, hipStream_t s)
{
    int rc = dev_alloc(m->mr_new_scale, sizeof(int) * (size_t) m->view.max_blocks);
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
    const unsigned grid = (unsigned) (cus * 8);
    k_mr_init_new<<<grid, kMrThreads, 0, s>>>(m->view, ulist);
    k_mr_scale_down<<<grid, kMrThreads, 0, s>>>(m->view, ulist, fp, sensor.c.inv_fx, ns);
    k_mr_update_up<<<grid, kMrThreads, 0, s>>>(m->view, ulist, fp, sensor, depth, ts, ns);
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
