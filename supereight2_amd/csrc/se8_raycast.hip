// Raycast and map queries of libse8.so.
//
// raycast_volume (include/se/map/impl/raycaster_impl.hpp:537-627): one thread per pixel in 16x8
// tiles.  Per ray:
//   1. first allocated block along the ray and its entry t, by a stack-based octree traversal in
//      the scaled [1,2]^3 octree space (Laine & Karras 2010, "Efficient Sparse Voxel Octrees"),
//      with the same float operations as VoxelBlockRayIterator
//      (voxel_block_ray_iterator_impl.hpp:19-249) so t_min is bit-identical; the traversal stack
//      lives in LDS;
//   2. the TSDF march of raycast<TSDF> (raycaster_impl.hpp:456-533) with O(1) voxel lookups
//      through the brick index instead of octree descents;
//   3. the normal from the single-res 32-sample gradient (visitor_impl.hpp:1094-1251).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "se8_internal.hpp"
#include "se8_pyramid.hpp"

namespace se8 {

// The raycast kernels and their helpers are instantiated per map layout: ST = storage
// (SE8_STORAGE_F32 = 0 or SE8_STORAGE_F16 = 1) + 2 if the map has the dense block table.
template <typename K>
K pick_kernel(const MapView& m, K f32_brick, K f16_brick, K f32_dense, K f16_dense)
{
    const bool f32 = m.storage == SE8_STORAGE_F32;
    return m.block_dense ? (f32 ? f32_dense : f16_dense) : (f32 ? f32_brick : f16_brick);
}

constexpr int kStackDepth = 23;  // CAST_STACK_DEPTH (voxel_block_ray_iterator.hpp:15)
constexpr int kMaxOctreeLevels = 16;   // log2(size / 8) for the largest supported octree (2^19)
constexpr int kRayThreads = 128; // 16 x 8 pixel tile

// Dynamic LDS of a raycast workgroup: the traversal stack of every thread.
inline size_t ray_stack_bytes(const MapView& m)
{
    return (size_t) (m.max_scale - kBlockScale) * kRayThreads * (sizeof(int) + sizeof(float));
}

__device__ __forceinline__ int f2i(float f)
{
    return __float_as_int(f);
}

__device__ __forceinline__ float i2f(int i)
{
    return __int_as_float(i);
}

struct Span {
    float t_min;  // tcmin(): entry t of the first allocated block (metres)
    float t_max;  // tmax(): t_max_init (metres)
};

// st_parent / st_tmax: this thread's stack slots in LDS, stride = threads per workgroup.
__device__ Span first_block(const MapView& m, const float oW[3], const float dW[3], float near_t, float far_t,
                            int* st_parent, float* st_tmax, int stride)
{
    const float scaling = (float) m.size * m.res;
    const int lg = m.max_scale - kBlockScale;
    const int min_scale = kStackDepth - lg;
    const float eps = exp2f((float) -m.max_scale);
    float tc[3], tb[3];
    int octant_mask = 7;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float d = fabsf(dW[i]) < eps ? copysignf(eps, dW[i]) : dW[i];
        const float pM = oW[i] + m.t_MW[i];
        const float o = div_rn(pM, m.res, m.inv_res) / (float) m.size + 1.0f;
        tc[i] = -1.f * (1.0f / fabsf(d));
        tb[i] = tc[i] * o;
        if (d > 0.0f) {
            octant_mask ^= 1 << i;
            tb[i] = 3.0f * tc[i] - tb[i];
        }
    }
    float tmin_init = fmaxf(fmaxf(2.0f * tc[0] - tb[0], 2.0f * tc[1] - tb[1]), 2.0f * tc[2] - tb[2]);
    float tmax_init = fminf(fminf(tc[0] - tb[0], tc[1] - tb[1]), tc[2] - tb[2]);
    float h = tmax_init;
    tmin_init = fmaxf(tmin_init, near_t / scaling);
    tmax_init = fminf(tmax_init, far_t / scaling);
    float t_min = tmin_init, t_max = tmax_init;
    float pos[3] = {1.f, 1.f, 1.f};
    int idx = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (1.5f * tc[i] - tb[i] > t_min) {
            idx ^= 1 << i;
            pos[i] = 1.5f;
        }
    }
    for (int l = 0; l < lg; l++)
        st_parent[l * stride] = -1;
    int parent = 0;
    // the 8 child entries of `parent` (one 32-byte row), reloaded only when parent changes, so
    // advancing between siblings costs no memory access
    int4 row_lo, row_hi;
    auto load_row = [&](int p) {
        const int4* r = reinterpret_cast<const int4*>(m.node_child + (size_t) p * 8);
        row_lo = r[0];
        row_hi = r[1];
    };
    load_row(parent);
    int scale = kStackDepth - 1;
    float scale_exp2 = 0.5f;
    // the traversal visits each octant at most once; the cap only guards against a corrupt octree.
    // A descent and a pop both end with the row of the new parent; that load is issued from one
    // place at the end of the iteration, so a wave whose lanes take both paths issues one gather.
    for (int guard = 0; scale < kStackDepth && guard < (1 << 20); guard++) {
        bool reload = false;
        float tcorner[3];
#pragma unroll
        for (int i = 0; i < 3; i++)
            tcorner[i] = pos[i] * tc[i] - tb[i];
        const float tc_max = fminf(fminf(tcorner[0], tcorner[1]), tcorner[2]);
        const int ci = idx ^ octant_mask ^ 7;
        const int4 half_row = (ci & 4) ? row_hi : row_lo;
        const int child = (ci & 2) ? ((ci & 1) ? half_row.w : half_row.z) : ((ci & 1) ? half_row.y : half_row.x);
        if (scale == min_scale && child >= 0)
            break;
        if (child >= 0 && t_min <= t_max) {
            const float tv_max = fminf(t_max, tc_max);
            const float half = scale_exp2 * 0.5f;
            float tcen[3];
#pragma unroll
            for (int i = 0; i < 3; i++)
                tcen[i] = half * tc[i] + tcorner[i];
            if (tc_max < h) {
                const int l = kStackDepth - 1 - scale;
                st_parent[l * stride] = parent;
                st_tmax[l * stride] = t_max;
            }
            h = tc_max;
            parent = child;
            reload = true;
            idx = 0;
            scale--;
            scale_exp2 = half;
#pragma unroll
            for (int i = 0; i < 3; i++)
                if (tcen[i] > t_min)
                    idx ^= 1 << i;
            // pos + scale_exp2 * bit, as a select (adding 0 leaves pos unchanged)
#pragma unroll
            for (int i = 0; i < 3; i++)
                pos[i] = ((idx >> i) & 1) ? pos[i] + scale_exp2 : pos[i];
            t_max = tv_max;
        }
        else {
            const int step_mask =
                (tcorner[0] <= tc_max) | ((tcorner[1] <= tc_max) << 1) | ((tcorner[2] <= tc_max) << 2);
#pragma unroll
            for (int i = 0; i < 3; i++)
                pos[i] = ((step_mask >> i) & 1) ? pos[i] - scale_exp2 : pos[i];
            t_min = tc_max;
            idx ^= step_mask;
            if ((idx & step_mask) != 0) {
                unsigned int diff = 0;
#pragma unroll
                for (int i = 0; i < 3; i++)
                    if (step_mask & (1 << i))
                        diff |= (unsigned) (f2i(pos[i]) ^ f2i(pos[i] + scale_exp2));
                scale = (f2i((float) diff) >> 23) - 127;
                if (scale >= kStackDepth)
                    break;
                const int l = kStackDepth - 1 - scale;
                if (l < 0 || l >= lg || st_parent[l * stride] < 0) {
                    scale = kStackDepth; // unreachable for a consistent octree
                    break;
                }
                scale_exp2 = i2f((scale - kStackDepth + 127) << 23);
                parent = st_parent[l * stride];
                reload = true;
                t_max = st_tmax[l * stride];
                int sh[3];
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    sh[i] = f2i(pos[i]) >> scale;
                    pos[i] = i2f(sh[i] << scale);
                }
                idx = (sh[0] & 1) | ((sh[1] & 1) << 1) | ((sh[2] & 1) << 2);
                h = 0.0f;
            }
        }
        if (reload)
            load_row(parent);
    }
    Span s;
    s.t_min = t_min * scaling;
    s.t_max = tmax_init * scaling;
    return s;
}

// pointToVoxel<Safe::On> before the truncation: the voxel position of a world point (point_M /
// res, equal to IEEE division), false outside the map.  Inside the map 0 <= point_M < dim, so
// div_by_recip is exact unless a coordinate is below 2^-60; the bounds are tested without branches
// and the quotients taken for every lane.
__device__ __forceinline__ bool point_to_voxel_f(const MapView& m, const float p[3], float v[3])
{
    float pM[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
        pM[i] = p[i] + m.t_MW[i];
    const bool in = (pM[0] >= 0.f) & (pM[0] < m.dim) & (pM[1] >= 0.f) & (pM[1] < m.dim) & (pM[2] >= 0.f)
        & (pM[2] < m.dim);
#pragma unroll
    for (int i = 0; i < 3; i++)
        v[i] = div_by_recip(pM[i], m.res, m.inv_res);
    if (in && !(fminf(fminf(pM[0], pM[1]), pM[2]) >= 0x1p-60f)) {
#pragma unroll
        for (int i = 0; i < 3; i++)
            v[i] = pM[i] / m.res;
    }
    return in;
}

// pointToVoxel<Safe::On>: voxel of a world point, false outside the map
__device__ __forceinline__ bool point_to_voxel(const MapView& m, const float p[3], int v[3])
{
    float vf[3];
    if (!point_to_voxel_f(m, p, vf))
        return false;
#pragma unroll
    for (int i = 0; i < 3; i++)
        v[i] = (int) vf[i];
    return true;
}

// Map::getData<Safe::On>
template <int ST>
__device__ __forceinline__ void get_data(const MapView& m, const float p[3], float& tsdf, float& weight)
{
    int v[3];
    if (!point_to_voxel(m, p, v)) {
        tsdf = 1.f;
        weight = 0.f;
        return;
    }
    const int slot = lookup_block_t<(ST >> 1) != 0>(m, v[0], v[1], v[2]);
    if (slot < 0) {
        tsdf = 1.f;
        weight = 0.f;
        return;
    }
    load_voxel_t<(ST & 1)>(m, slot, (v[0] & 7) + ((v[1] & 7) << 3) + ((v[2] & 7) << 6), tsdf, weight);
}

// Field values of the 8 voxels base + {0,1}^3 (x fastest) and whether all are allocated and
// observed (weight != 0).  Every block lookup and voxel load is issued before any result is
// tested, so the set costs the round trips of one lookup, not eight in sequence.
template <int ST>
__device__ __forceinline__ bool gather_corners(const MapView& m, const int base[3], float d[8])
{
    int slot[8];
#pragma unroll
    for (int k = 0; k < 8; k++)
        slot[k] = lookup_block_nb_t<(ST >> 1) != 0>(m, base[0] + (k & 1), base[1] + ((k >> 1) & 1), base[2] + ((k >> 2) & 1));
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int x = base[0] + (k & 1), y = base[1] + ((k >> 1) & 1), z = base[2] + ((k >> 2) & 1);
        float w;
        load_voxel_t<(ST & 1)>(m, max(slot[k], 0), (x & 7) + ((y & 7) << 3) + ((z & 7) << 6), d[k], w);
        ok = ok && slot[k] >= 0 && w != 0.f;
    }
    return ok;
}

// getData<Safe::On> along a march, from the sample's voxel position (point_to_voxel_f; in = inside
// the map): consecutive samples usually fall in the same block, so the block slot of the previous
// sample is kept and the brick/slot loads are skipped when its block coordinates repeat.
struct SlotCache {
    int bx = -1, by = -1, bz = -1;
    int slot = -1;
};

template <int ST>
__device__ __forceinline__ void get_data_cached(const MapView& m, bool in, const float vf[3], float& tsdf,
                                                float& weight, SlotCache& cache)
{
    if (!in) {
        tsdf = 1.f;
        weight = 0.f;
        return;
    }
    const int v[3] = {(int) vf[0], (int) vf[1], (int) vf[2]};
    const int wrap = m.size - 1;
    const int bx = (v[0] & wrap) >> 3, by = (v[1] & wrap) >> 3, bz = (v[2] & wrap) >> 3;
    if (bx != cache.bx || by != cache.by || bz != cache.bz) {
        cache.bx = bx;
        cache.by = by;
        cache.bz = bz;
        cache.slot = lookup_block_t<(ST >> 1) != 0>(m, v[0], v[1], v[2]);
    }
    if (cache.slot < 0) {
        tsdf = 1.f;
        weight = 0.f;
        return;
    }
    load_voxel_t<(ST & 1)>(m, cache.slot, (v[0] & 7) + ((v[1] & 7) << 3) + ((v[2] & 7) << 6), tsdf, weight);
}

// single-res getInterp (visitor_impl.hpp:976-1024) from voxel coordinates.  The reference's
// octree-AABB test is implied by the requirement that the block of base be allocated.
template <int ST>
__device__ bool field_interp_vox(const MapView& m, const float vf[3], float& out)
{
    float bf[3];
    int base[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        bf[i] = vf[i] - 0.5f;
        base[i] = (int) bf[i];
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
        if (base[i] + 1 >= m.size || base[i] < 0)
            return false;
    float d[8];
    if (!gather_corners<ST>(m, base, d))
        return false;
    float t[3], tcm[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        t[i] = bf[i] - floorf(bf[i]);
        tcm[i] = 1.f - t[i];
    }
    out = ((d[0] * tcm[0] + d[1] * t[0]) * tcm[1] + (d[2] * tcm[0] + d[3] * t[0]) * t[1]) * tcm[2]
        + ((d[4] * tcm[0] + d[5] * t[0]) * tcm[1] + (d[6] * tcm[0] + d[7] * t[0]) * t[1]) * t[2];
    return true;
}

// single-res getFieldGrad (visitor_impl.hpp:1094-1251): trilinear blend of the central differences
// at the 8 voxels around the query; x0.5 (not yet divided by res).
template <int ST>
__device__ bool field_grad_vox(const MapView& m, const float vf[3], float g[3])
{
    float fac[3];
    int base[3];
    int c[3][4];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float s = vf[i] - 0.5f;
        fac[i] = s - floorf(s);
        base[i] = (int) s;
        c[i][0] = max(base[i] - 1, 0);
        c[i][1] = max(base[i], 0);
        c[i][2] = min(base[i] + 1, m.size - 1);
        c[i][3] = min(base[i] + 2, m.size - 1);
    }
    // The central differences read the 32 voxels c[0][i], c[1][j], c[2][k] of the 4x4x4
    // neighbourhood with at most one of i, j, k on an outer layer (0 or 3).  The 4 coordinates of
    // an axis are consecutive (or clamped) voxels, so they lie in at most two blocks: hi[a][n]
    // says whether c[a][n] lies in the block of c[a][3] rather than that of c[a][0].  The slots of
    // those 2x2x2 blocks are looked up once and each voxel selects its own.  The result is
    // undefined if any of the 32 voxels is invalid or the block of base is missing, as in the
    // reference.
    bool hi[3][4];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int n = 0; n < 4; n++)
            hi[a][n] = (c[a][n] >> 3) != (c[a][0] >> 3);
    // The 8 slots are separate values and the columns are expanded by a macro, not a lambda: an
    // array element selected by a runtime condition, or a variable a lambda captures by reference,
    // can be left in private (scratch) memory, and scratch use limits the kernel's occupancy.
#define SE8_GRAD_SLOT(kk) \
    lookup_block_nb_t<(ST >> 1) != 0>(m, c[0][((kk) & 1) ? 3 : 0], c[1][((kk) & 2) ? 3 : 0], c[2][((kk) & 4) ? 3 : 0])
    const int bs0 = SE8_GRAD_SLOT(0), bs1 = SE8_GRAD_SLOT(1), bs2 = SE8_GRAD_SLOT(2), bs3 = SE8_GRAD_SLOT(3),
              bs4 = SE8_GRAD_SLOT(4), bs5 = SE8_GRAD_SLOT(5), bs6 = SE8_GRAD_SLOT(6), bs7 = SE8_GRAD_SLOT(7);
#undef SE8_GRAD_SLOT
    const int base_slot = lookup_block_nb_t<(ST >> 1) != 0>(m, base[0], base[1], base[2]);
    constexpr int kOuter = 0x9; // bit i set: index i is an outer layer
    float F[4][4][4];
    int ok = base_slot >= 0;
    // the voxels of one (y, z) column: all four x for an inner (y, z), the inner two otherwise
#define SE8_GRAD_COLUMN(j, k)                                                                              \
    {                                                                                                      \
        const int lo = hi[2][k] ? (hi[1][j] ? bs6 : bs4) : (hi[1][j] ? bs2 : bs0);                          \
        const int up = hi[2][k] ? (hi[1][j] ? bs7 : bs5) : (hi[1][j] ? bs3 : bs1);                          \
        const bool inner = ((kOuter >> (j)) & 1) + ((kOuter >> (k)) & 1) == 0;                             \
        _Pragma("unroll") for (int i = 0; i < 4; i++) if (inner || !((kOuter >> i) & 1))                   \
        {                                                                                                  \
            const int sl = hi[0][i] ? up : lo;                                                             \
            float w;                                                                                       \
            load_voxel_t<(ST & 1)>(m, max(sl, 0), (c[0][i] & 7) + ((c[1][j] & 7) << 3) + ((c[2][k] & 7) << 6), \
                                   F[i][j][k], w);                                                         \
            ok &= (sl >= 0) & (w != 0.f);                                                                  \
        }                                                                                                  \
    }
    // Three batches of loads (inner y and z, outer y, outer z) keep at most 16 voxel loads in
    // flight.  Each axis's blend is taken as soon as its voxels are in: x needs the first batch,
    // y the second, z the third, so the voxels only one axis reads are released early.  The
    // validity of each batch is folded in as it arrives, so its weights are not kept either.
    const float rx = 1.f - fac[0], ry = 1.f - fac[1], rz = 1.f - fac[2];
#define SE8_GRAD_AXIS(a)                                                                                   \
    {                                                                                                      \
        float cd[8];                                                                                       \
        _Pragma("unroll") for (int k = 0; k < 8; k++)                                                      \
        {                                                                                                  \
            const int ck[3] = {k & 1, (k >> 1) & 1, (k >> 2) & 1};                                         \
            int ip[3], im[3];                                                                              \
            _Pragma("unroll") for (int i = 0; i < 3; i++)                                                  \
            {                                                                                              \
                ip[i] = 1 + ck[i];                                                                         \
                im[i] = 1 + ck[i];                                                                         \
            }                                                                                              \
            ip[a] = 2 + ck[a];                                                                             \
            im[a] = ck[a];                                                                                 \
            cd[k] = F[ip[0]][ip[1]][ip[2]] - F[im[0]][im[1]][im[2]];                                       \
        }                                                                                                  \
        const float v = ((cd[0] * rx + cd[1] * fac[0]) * ry + (cd[2] * rx + cd[3] * fac[0]) * fac[1]) * rz \
            + ((cd[4] * rx + cd[5] * fac[0]) * ry + (cd[6] * rx + cd[7] * fac[0]) * fac[1]) * fac[2];      \
        g[a] = 0.5f * v;                                                                                   \
    }
    SE8_GRAD_COLUMN(1, 1)
    SE8_GRAD_COLUMN(2, 1)
    // FP32 voxels are two words each: the first batch is split in two so that its loads fit the
    // registers of 5 resident waves per SIMD (every wave of a 640x480 frame)
    if constexpr ((ST & 1) == SE8_STORAGE_F32)
        asm volatile("" : "+v"(F[1][1][1]), "+v"(F[2][1][1]), "+v"(F[1][2][1]), "+v"(F[2][2][1]), "+v"(ok) : : "memory");
    SE8_GRAD_COLUMN(1, 2)
    SE8_GRAD_COLUMN(2, 2)
    SE8_GRAD_AXIS(0)
    // each axis is complete before the next batch of loads is issued (an empty asm statement that
    // takes the axis's value and orders memory), so that only one batch is in flight at a time
    asm volatile("" : "+v"(g[0]), "+v"(ok) : : "memory");
    SE8_GRAD_COLUMN(0, 1)
    SE8_GRAD_COLUMN(3, 1)
    SE8_GRAD_COLUMN(0, 2)
    SE8_GRAD_COLUMN(3, 2)
    SE8_GRAD_AXIS(1)
    asm volatile("" : "+v"(g[1]), "+v"(ok) : : "memory");
    SE8_GRAD_COLUMN(1, 0)
    SE8_GRAD_COLUMN(2, 0)
    SE8_GRAD_COLUMN(1, 3)
    SE8_GRAD_COLUMN(2, 3)
    SE8_GRAD_AXIS(2)
#undef SE8_GRAD_AXIS
#undef SE8_GRAD_COLUMN
    return ok != 0;
}

// raycast<TSDF> (raycaster_impl.hpp:456-533).  On a hit, bracket (if given) receives the ray
// parameters of the last two samples, whose field values straddle zero.
template <int ST>
__device__ bool march(const MapView& m, const float o[3], const float d[3], float t_near, float t_far, int* st_parent,
                      float* st_tmax, int stride, float hit[3], float* bracket = nullptr)
{
    const Span span = first_block(m, o, d, t_near, t_far, st_parent, st_tmax, stride);
    const float step = m.res;
    const float largestep = 8.f * step;
    if (span.t_min <= 0.f)
        return false;
    float t = span.t_min;
    if (!(t_near < span.t_max))
        return false;
    float stepsize = largestep;
    float p[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
        p[i] = o[i] + d[i] * t;
    float f_t, w, vf[3];
    SlotCache cache;
    get_data_cached<ST>(m, point_to_voxel_f(m, p, vf), vf, f_t, w, cache);
    float f_tt = 0.f;
    if (!(f_t >= 0.f))
        return false;
    for (; t < t_far; t += stepsize) {
        float f_new;
        get_data_cached<ST>(m, point_to_voxel_f(m, p, vf), vf, f_new, w, cache);
        if (w == 0.f) {
            stepsize = largestep;
#pragma unroll
            for (int i = 0; i < 3; i++)
                p[i] = p[i] + stepsize * d[i];
            continue;
        }
        f_tt = f_new;
        // a sample with a weight lies inside the map, and vf is its voxel position
        if (f_tt <= 0.1f && f_tt >= -0.5f) {
            float fi;
            if (field_interp_vox<ST>(m, vf, fi))
                f_tt = fi;
        }
        if (f_tt < 0.f)
            break;
        // std::max(f_tt * mu, step) (raycaster_impl.hpp:515): a NaN first argument is returned
        const float fs = f_tt * m.mu;
        stepsize = fs < step ? step : fs;
#pragma unroll
        for (int i = 0; i < 3; i++)
            p[i] = p[i] + stepsize * d[i];
        f_t = f_tt;
    }
    if (f_tt < 0.f) {
        if (bracket) {
            bracket[0] = t - stepsize;
            bracket[1] = t;
        }
        t = t - stepsize * (f_tt - 0.f) / (f_tt - f_t);
#pragma unroll
        for (int i = 0; i < 3; i++)
            hit[i] = o[i] + d[i] * t;
        return true;
    }
    return false;
}

// Trilinear field value and its spatial gradient (per metre) at a world point; false where the
// reference's interpolation is undefined (visitor_impl.hpp:976-1024).
template <int ST>
__device__ bool interp_and_grad(const MapView& m, const float pW[3], float& f, float g[3])
{
    float vf[3];
    if (!point_to_voxel_f(m, pW, vf))
        return false;
    float bf[3], t[3], tc[3];
    int base[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        bf[i] = vf[i] - 0.5f;
        base[i] = (int) bf[i];
        if (base[i] + 1 >= m.size || base[i] < 0)
            return false;
    }
    float d[8];
    if (!gather_corners<ST>(m, base, d))
        return false;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        t[i] = bf[i] - floorf(bf[i]);
        tc[i] = 1.f - t[i];
    }
    f = ((d[0] * tc[0] + d[1] * t[0]) * tc[1] + (d[2] * tc[0] + d[3] * t[0]) * t[1]) * tc[2]
        + ((d[4] * tc[0] + d[5] * t[0]) * tc[1] + (d[6] * tc[0] + d[7] * t[0]) * t[1]) * t[2];
    g[0] = div_rn(((d[1] - d[0]) * tc[1] + (d[3] - d[2]) * t[1]) * tc[2] + ((d[5] - d[4]) * tc[1] + (d[7] - d[6]) * t[1]) * t[2],
                  m.res, m.inv_res);
    g[1] = div_rn(((d[2] - d[0]) * tc[0] + (d[3] - d[1]) * t[0]) * tc[2] + ((d[6] - d[4]) * tc[0] + (d[7] - d[5]) * t[0]) * t[2],
                  m.res, m.inv_res);
    g[2] = div_rn(((d[4] - d[0]) * tc[0] + (d[5] - d[1]) * t[0]) * tc[1] + ((d[6] - d[2]) * tc[0] + (d[7] - d[3]) * t[0]) * t[1],
                  m.res, m.inv_res);
    return true;
}

// Root of the trilinear field on [ta, tb] (f(ta) >= 0 > f(tb)) by Illinois regula falsi.
template <int ST>
__device__ bool refine_root(const MapView& m, const float o[3], const float d[3], float ta, float tb, int iters,
                            float& t_out)
{
    float p[3], fa, fb, g[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
        p[i] = o[i] + d[i] * ta;
    if (!interp_and_grad<ST>(m, p, fa, g))
        return false;
#pragma unroll
    for (int i = 0; i < 3; i++)
        p[i] = o[i] + d[i] * tb;
    if (!interp_and_grad<ST>(m, p, fb, g))
        return false;
    if (!(fa >= 0.f && fb < 0.f))
        return false;
    int side = 0;
    float t = tb;
    for (int it = 0; it < iters; it++) {
        t = (ta * fb - tb * fa) / (fb - fa);
        float f;
#pragma unroll
        for (int i = 0; i < 3; i++)
            p[i] = o[i] + d[i] * t;
        if (!interp_and_grad<ST>(m, p, f, g))
            return false;
        if (f < 0.f) {
            tb = t;
            fb = f;
            if (side == -1)
                fa *= 0.5f;
            side = -1;
        }
        else {
            ta = t;
            fa = f;
            if (side == 1)
                fb *= 0.5f;
            side = 1;
        }
        if (tb - ta <= 0.f)
            break;
    }
    t_out = t;
    return true;
}

// Batched differentiable depth: grid (tiles, poses).  d depth / d xi = dn_z * dt/dxi with
// dt/dxi = -[g, x cross g] / (g . d) for the left twist xi = [v; w].
template <int ST>
__global__ __launch_bounds__(kRayThreads) void k_raycast_depth_grad(MapView m, Cam cam, const float* poses12,
                                                                     float* depth, float* grad, uint8_t* valid)
{
    // traversal stack: (max_scale - 3) levels of {parent, t_max} per thread, sized at launch
    extern __shared__ int ray_stack[];
    int* st_parent = ray_stack;
    float* st_tmax = reinterpret_cast<float*>(ray_stack + (m.max_scale - kBlockScale) * kRayThreads);
    const int k = blockIdx.z;
    const int t = threadIdx.x;
    const int x = blockIdx.x * 16 + (t & 15);
    const int y = blockIdx.y * 8 + (t >> 4);
    if (x >= cam.w || y >= cam.h)
        return;
    Iso T;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T.R[i][j] = poses12[12 * k + 4 * i + j];
        T.t[i] = poses12[12 * k + 4 * i + 3];
    }
    const size_t idx = (size_t) k * cam.w * cam.h + x + (size_t) y * cam.w;
    float dS[3], dn[3], dW[3], hit[3], br[2];
    cam_backproject(cam, (float) x, (float) y, dS);
    normalized3(dS, dn);
    lin_apply(T, dn, dW);
    float out_d = 0.f, out_g[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    uint8_t ok = 0;
    if (march<ST>(m, T.t, dW, cam_near_dist(cam, dS), cam_far_dist(cam, dS), st_parent + t, st_tmax + t, kRayThreads, hit,
              br)) {
        float tr;
        if (refine_root<ST>(m, T.t, dW, br[0], br[1], 24, tr)) {
            float xr[3], f, g[3];
#pragma unroll
            for (int i = 0; i < 3; i++)
                xr[i] = T.t[i] + dW[i] * tr;
            if (interp_and_grad<ST>(m, xr, f, g)) {
                const float gd = dot3(g, dW);
                if (gd != 0.f) {
                    float xg[3];
                    cross3(xr, g, xg);
                    out_d = tr * dn[2];
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        out_g[i] = -g[i] / gd * dn[2];
                        out_g[3 + i] = -xg[i] / gd * dn[2];
                    }
                    ok = 1;
                }
            }
        }
    }
    depth[idx] = out_d;
#pragma unroll
    for (int i = 0; i < 6; i++)
        grad[6 * idx + i] = out_g[i];
    valid[idx] = ok;
}

// The rays of the 16x8 pixel tile (bx, by); ray_stack is the workgroup's dynamic LDS.
template <int ST>
__device__ void raycast_tile(const MapView& m, const Cam& cam, const Iso& T_arg, const float* T16_dev, float* points,
                             float* normals, int8_t* scale_out, int bx, int by, int* ray_stack)
{
    // the pose comes from device memory when T16_dev is set (a pose a previous kernel produced)
    const Iso T_WS = T16_dev ? iso_from16(T16_dev) : T_arg;
    // traversal stack: (max_scale - 3) levels of {parent, t_max} per thread, sized at launch
    int* st_parent = ray_stack;
    float* st_tmax = reinterpret_cast<float*>(ray_stack + (m.max_scale - kBlockScale) * kRayThreads);
    const int t = threadIdx.x;
    const int x = bx * 16 + (t & 15);
    const int y = by * 8 + (t >> 4);
    if (x >= cam.w || y >= cam.h)
        return;
    const size_t idx = x + (size_t) y * cam.w;
    float dS[3], dn[3], dW[3];
    cam_backproject(cam, (float) x, (float) y, dS);
    normalized3(dS, dn);
    lin_apply(T_WS, dn, dW);
    float hit[3];
    float nrm[3] = {0.f, 0.f, 0.f};
    const bool ok = march<ST>(m, T_WS.t, dW, cam_near_dist(cam, dS), cam_far_dist(cam, dS), st_parent + t,
                          st_tmax + t, kRayThreads, hit);
    if (ok) {
        if (scale_out)
            scale_out[idx] = 0;
        float vf[3], g[3];
        if (point_to_voxel_f(m, hit, vf) && field_grad_vox<ST>(m, vf, g)) {
#pragma unroll
            for (int i = 0; i < 3; i++)
                g[i] = div_rn(g[i], m.res, m.inv_res);
            normalized3(g, nrm);
        }
    }
    else {
        hit[0] = hit[1] = hit[2] = 0.f;
    }
    points[3 * idx + 0] = hit[0];
    points[3 * idx + 1] = hit[1];
    points[3 * idx + 2] = hit[2];
    normals[3 * idx + 0] = nrm[0];
    normals[3 * idx + 1] = nrm[1];
    normals[3 * idx + 2] = nrm[2];
}

template <int ST>
__global__ __launch_bounds__(kRayThreads) void k_raycast(MapView m, Cam cam, Iso T_arg, const float* T16_dev,
                                                         float* points, float* normals, int8_t* scale_out)
{
    extern __shared__ int ray_stack[];
    raycast_tile<ST>(m, cam, T_arg, T16_dev, points, normals, scale_out, blockIdx.x, blockIdx.y, ray_stack);
}

// The raycast and, in the same launch, the pyramid of a tracker's next call (pyramid_tile over
// 128-thread workgroups): workgroups below n_ray take the raycast's tiles (n_ray_x per row), the
// rest the pyramid's.  The pyramid's workgroups are dispatched after the raycast's and run while
// the raycast's slowest rays finish, on a device the raycast leaves mostly idle by then.
template <int ST>
__global__ __launch_bounds__(kRayThreads) void k_raycast_pyramid(MapView m, Cam cam, const float* T16_dev,
                                                                 float* points, float* normals, int8_t* scale_out,
                                                                 int n_ray_x, int n_ray, PyrArgs pa)
{
    extern __shared__ int lds[];
    const int b = blockIdx.x;
    if (b < n_ray)
        raycast_tile<ST>(m, cam, Iso{}, T16_dev, points, normals, scale_out, b % n_ray_x, b / n_ray_x, lds);
    else
        pyramid_tile<kRayThreads>(pa, b - n_ray, reinterpret_cast<float*>(lds));
}

// ----------------------------------------------------------------------------------------------
// point queries
// ----------------------------------------------------------------------------------------------
template <int ST>
__global__ void k_query(MapView m, const float* pts, int n, int mode, float* out, uint8_t* valid)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    if (mode == 0) {
        float t, w;
        get_data<ST>(m, p, t, w);
        out[2 * i] = t;
        out[2 * i + 1] = w;
        return;
    }
    float vf[3];
    if (!point_to_voxel_f(m, p, vf)) {
        valid[i] = 0;
        return;
    }
    if (mode == 1) {
        float f;
        const bool ok = field_interp_vox<ST>(m, vf, f);
        valid[i] = ok;
        out[i] = ok ? f : 0.f;
    }
    else {
        float g[3] = {0.f, 0.f, 0.f};
        const bool ok = field_grad_vox<ST>(m, vf, g);
        valid[i] = ok;
        for (int k = 0; k < 3; k++)
            out[3 * i + k] = ok ? div_rn(g[k], m.res, m.inv_res) : 0.f;
    }
}

static int run_query(const se8_map* m, const float* pts, int64_t n, int mode, std::vector<float>& out,
                     std::vector<uint8_t>& valid)
{
    const int width = mode == 0 ? 2 : (mode == 1 ? 1 : 3);
    out.assign((size_t) n * width, 0.f);
    valid.assign((size_t) n, 0);
    if (n == 0)
        return SE8_OK;
    DeviceBuffer dp, dout, dval;
    int rc = dev_alloc(dp, sizeof(float) * 3 * n);
    rc = rc ? rc : dev_alloc(dout, sizeof(float) * width * n);
    rc = rc ? rc : dev_alloc(dval, (size_t) n);
    if (rc == SE8_OK) {
        if (hipMemcpy(dp.ptr, pts, sizeof(float) * 3 * n, hipMemcpyHostToDevice) != hipSuccess)
            rc = SE8_ERR_HIP;
        else {
            auto* kq = pick_kernel(m->view, k_query<0>, k_query<1>, k_query<2>, k_query<3>);
            kq<<<(unsigned) ((n + 255) / 256), 256>>>(m->view, (const float*) dp.ptr, (int) n, mode,
                                                           (float*) dout.ptr, (uint8_t*) dval.ptr);
            if (hipMemcpy(out.data(), dout.ptr, sizeof(float) * width * n, hipMemcpyDeviceToHost) != hipSuccess
                || hipMemcpy(valid.data(), dval.ptr, (size_t) n, hipMemcpyDeviceToHost) != hipSuccess)
                rc = SE8_ERR_HIP;
        }
    }
    if (rc == SE8_ERR_HIP)
        set_error("map query: HIP call failed");
    dev_free(dp);
    dev_free(dout);
    dev_free(dval);
    return rc;
}

// ev_start / ev_stop, when given, are stamped at the start and end of the kernel itself (not at
// enqueue time), so their difference is the kernel's execution time.
int launch_raycast(const se8_map* m, const Cam& cam, const Iso& T_WS, const float* T16_dev, float* points,
                   float* normals, int8_t* scale, hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_stop)
{
    const dim3 grid((cam.w + 15) / 16, (cam.h + 7) / 8);
    auto* kr = pick_kernel(m->view, k_raycast<0>, k_raycast<1>, k_raycast<2>, k_raycast<3>);
    const uint32_t lds = (uint32_t) ray_stack_bytes(m->view);
    if (ev_start || ev_stop)
        hipExtLaunchKernelGGL(kr, grid, dim3(kRayThreads), lds, s, ev_start, ev_stop, 0, m->view, cam, T_WS, T16_dev,
                              points, normals, scale);
    else
        hipLaunchKernelGGL(kr, grid, dim3(kRayThreads), lds, s, m->view, cam, T_WS, T16_dev, points, normals, scale);
    SE8_HIP_CHECK(hipGetLastError());
    graph_note_launch(kNodeRaycast, s);
    return SE8_OK;
}

} // namespace se8

using namespace se8;

extern "C" int se8_raycast(const se8_map* m, const se8_pinhole* sensor, const float T_WS16[16], float* points_dev,
                           float* normals_dev, int8_t* scale_dev, se8_stream stream)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_raycast: not available for a multi-resolution map");
    SE8_REQUIRE(m && sensor && T_WS16 && points_dev && normals_dev, "se8_raycast: null argument");
    SE8_REQUIRE(sensor->width > 0 && sensor->height > 0, "se8_raycast: empty image");
    SE8_REQUIRE(m->view.max_scale - kBlockScale <= kMaxOctreeLevels, "se8_raycast: octree too deep");
    hipStream_t s = (hipStream_t) stream;
    se8_map* mm = const_cast<se8_map*>(m);
    const Cam cam = make_cam(*sensor, 1, false);
    const bool timed = mm->timing != SE8_TIMING_OFF;
    harvest_raycast(mm, true);
    int rc = launch_raycast(m, cam, iso_from16(T_WS16), nullptr, points_dev, normals_dev, scale_dev, s,
                            timed ? mm->ev[5] : nullptr, timed ? mm->ev[6] : nullptr);
    mm->ray_pending = timed;
    return rc;
}

extern "C" int se8_raycast_dev(const se8_map* m, const se8_pinhole* sensor, const float* T_WS_dev, float* points_dev,
                               float* normals_dev, int8_t* scale_dev, se8_stream stream)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_raycast_dev: not available for a multi-resolution map");
    SE8_REQUIRE(m && sensor && T_WS_dev && points_dev && normals_dev, "se8_raycast_dev: null argument");
    SE8_REQUIRE(sensor->width > 0 && sensor->height > 0, "se8_raycast_dev: empty image");
    SE8_REQUIRE(m->view.max_scale - kBlockScale <= kMaxOctreeLevels, "se8_raycast_dev: octree too deep");
    hipStream_t s = (hipStream_t) stream;
    se8_map* mm = const_cast<se8_map*>(m);
    const Cam cam = make_cam(*sensor, 1, false);
    const bool timed = mm->timing != SE8_TIMING_OFF;
    harvest_raycast(mm, true);
    int rc = launch_raycast(m, cam, Iso{}, T_WS_dev, points_dev, normals_dev, scale_dev, s,
                            timed ? mm->ev[5] : nullptr, timed ? mm->ev[6] : nullptr);
    mm->ray_pending = timed;
    return rc;
}

extern "C" int se8_raycast_dev_prepare_track(const se8_map* m, const se8_pinhole* sensor, const float* T_WS_dev,
                                            float* points_dev, float* normals_dev, int8_t* scale_dev,
                                            se8_tracker* tracker, const float* next_depth_dev, se8_stream stream)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_raycast_dev_prepare_track: not available for a multi-resolution map");
    SE8_REQUIRE(m && sensor && T_WS_dev && points_dev && normals_dev && tracker && next_depth_dev,
                "se8_raycast_dev_prepare_track: null argument");
    SE8_REQUIRE(sensor->width > 0 && sensor->height > 0, "se8_raycast_dev_prepare_track: empty image");
    SE8_REQUIRE(m->view.max_scale - kBlockScale <= kMaxOctreeLevels, "se8_raycast_dev_prepare_track: octree too deep");
    hipStream_t s = (hipStream_t) stream;
    se8_map* mm = const_cast<se8_map*>(m);
    const Cam cam = make_cam(*sensor, 1, false);
    const bool timed = mm->timing != SE8_TIMING_OFF;
    harvest_raycast(mm, true);
    PyrArgs pa{};
    int n_pyr = 0;
    int rc = tracker_prepare_pyramid(tracker, next_depth_dev, T_WS_dev, &pa, &n_pyr);
    if (rc != SE8_OK)
        return rc;
    const int n_ray_x = (cam.w + 15) / 16, n_ray = n_ray_x * ((cam.h + 7) / 8);
    auto* kr = pick_kernel(m->view, k_raycast_pyramid<0>, k_raycast_pyramid<1>, k_raycast_pyramid<2>,
                           k_raycast_pyramid<3>);
    const uint32_t lds = (uint32_t) std::max(ray_stack_bytes(m->view), pyramid_lds_bytes(pa.levels));
    const dim3 grid((unsigned) (n_ray + n_pyr));
    if (timed)
        hipExtLaunchKernelGGL(kr, grid, dim3(kRayThreads), lds, s, mm->ev[5], mm->ev[6], 0, m->view, cam, T_WS_dev,
                              points_dev, normals_dev, scale_dev, n_ray_x, n_ray, pa);
    else
        hipLaunchKernelGGL(kr, grid, dim3(kRayThreads), lds, s, m->view, cam, T_WS_dev, points_dev, normals_dev,
                           scale_dev, n_ray_x, n_ray, pa);
    SE8_HIP_CHECK(hipGetLastError());
    mm->ray_pending = timed;
    return SE8_OK;
}

extern "C" int se8_get_data(const se8_map* m, const float* pts, int64_t n, float* tsdf, float* weight)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_get_data: not available for a multi-resolution map");
    SE8_REQUIRE(m && (n == 0 || (pts && tsdf && weight)), "se8_get_data: null argument");
    std::vector<float> out;
    std::vector<uint8_t> valid;
    int rc = run_query(m, pts, n, 0, out, valid);
    if (rc != SE8_OK)
        return rc;
    for (int64_t i = 0; i < n; i++) {
        tsdf[i] = out[2 * i];
        weight[i] = out[2 * i + 1];
    }
    return SE8_OK;
}

extern "C" int se8_field_interp(const se8_map* m, const float* pts, int64_t n, float* outv, uint8_t* valid)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_field_interp: not available for a multi-resolution map");
    SE8_REQUIRE(m && (n == 0 || (pts && outv && valid)), "se8_field_interp: null argument");
    std::vector<float> out;
    std::vector<uint8_t> val;
    int rc = run_query(m, pts, n, 1, out, val);
    if (rc != SE8_OK)
        return rc;
    std::memcpy(outv, out.data(), sizeof(float) * n);
    std::memcpy(valid, val.data(), (size_t) n);
    return SE8_OK;
}

extern "C" int se8_field_grad(const se8_map* m, const float* pts, int64_t n, float* out3, uint8_t* valid)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_field_grad: not available for a multi-resolution map");
    SE8_REQUIRE(m && (n == 0 || (pts && out3 && valid)), "se8_field_grad: null argument");
    std::vector<float> out;
    std::vector<uint8_t> val;
    int rc = run_query(m, pts, n, 2, out, val);
    if (rc != SE8_OK)
        return rc;
    std::memcpy(out3, out.data(), sizeof(float) * 3 * n);
    std::memcpy(valid, val.data(), (size_t) n);
    return SE8_OK;
}

extern "C" int se8_raycast_depth_grad_batched(const se8_map* m, const se8_pinhole* sensor, int n_poses,
                                              const float* poses16, float* depth_dev, float* grad_dev,
                                              uint8_t* valid_dev, se8_stream stream)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_raycast_depth_grad_batched: not available for a multi-resolution map");
    SE8_REQUIRE(m && sensor && poses16 && depth_dev && grad_dev && valid_dev,
                "se8_raycast_depth_grad_batched: null argument");
    SE8_REQUIRE(n_poses > 0 && n_poses <= 65535, "se8_raycast_depth_grad_batched: n_poses out of range");
    SE8_REQUIRE(m->view.max_scale - kBlockScale <= kMaxOctreeLevels, "se8_raycast_depth_grad_batched: octree too deep");
    hipStream_t s = (hipStream_t) stream;
    const Cam cam = make_cam(*sensor, 1, false);
    std::vector<float> p12((size_t) 12 * n_poses);
    for (int k = 0; k < n_poses; k++)
        for (int i = 0; i < 12; i++)
            p12[(size_t) 12 * k + i] = poses16[(size_t) 16 * k + i];
    DeviceBuffer dp;
    int rc = dev_alloc(dp, sizeof(float) * p12.size());
    if (rc != SE8_OK)
        return rc;
    if (hipMemcpyAsync(dp.ptr, p12.data(), sizeof(float) * p12.size(), hipMemcpyHostToDevice, s) != hipSuccess) {
        dev_free(dp);
        set_error("se8_raycast_depth_grad_batched: pose upload failed");
        return SE8_ERR_HIP;
    }
    dim3 grid((cam.w + 15) / 16, (cam.h + 7) / 8, n_poses);
    auto* kd = pick_kernel(m->view, k_raycast_depth_grad<0>, k_raycast_depth_grad<1>, k_raycast_depth_grad<2>,
                           k_raycast_depth_grad<3>);
    kd<<<grid, kRayThreads, ray_stack_bytes(m->view), s>>>(m->view, cam, (const float*) dp.ptr, depth_dev, grad_dev,
                                                      valid_dev);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipStreamSynchronize(s);
    dev_free(dp);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        set_error(std::string("se8_raycast_depth_grad_batched: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2));
        return SE8_ERR_HIP;
    }
    return SE8_OK;
}
