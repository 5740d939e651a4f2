// Pyramid of Tracker::track (tracker_impl.hpp:50-75) for the tracker (se8_track.hip) and, fused with
// the raycast, for the raycast launch that prepares a tracker's next call (se8_raycast.hip); the
// device state the pyramid initialises for the first Gauss-Newton launch.
#pragma once

#include "se8_internal.hpp"

namespace se8 {

constexpr int kPyrThreads = 256;
constexpr int kMaxLevels = SE8_TRACKER_MAX_LEVELS;
// Hand-over between Gauss-Newton launches: workgroup i adds its partial row, in 64-bit fixed point,
// into bin i % kBins of the launch's round (rounds rotate over three buffers: launch k adds into
// round k % 3, reads round (k + 2) % 3 and clears round (k + 1) % 3).  Integer addition is exact and
// associative, so the sum does not depend on the order of the atomics, and every workgroup of the
// next launch reads kBins rows (kBinWords 64-bit words per thread) instead of one row per
// workgroup.  More bins spread the atomics of a launch over more addresses.
constexpr int kBins = 16;
constexpr int kBinWords = kBins * 32 / 256;
static_assert(kBinWords * 256 == kBins * 32, "whole bin entries per thread of a 256-thread workgroup");
// Level-0 tile side of a k_pyramid workgroup: 16 (2^4), or 2^(L-1) for deeper pyramids, so that every
// level keeps at least one tile pixel.  Small tiles give many light workgroups, which hide the latency
// of the per-pixel division chains better than few heavy ones.
constexpr int kPyrTileLog2Min = 4;
inline int pyramid_tile_log2(int levels)
{
    return levels - 1 > kPyrTileLog2Min ? levels - 1 : kPyrTileLog2Min;
}

// State handed from one Gauss-Newton launch to the next, double-buffered by launch index k:
// launch k reads slot (k + 1) & 1 and its workgroup 0 writes slot k & 1.
struct IcpSlot {
    float T[12];       // pose of this launch's pixel pass, row-major 3x4
    float sums[32];    // last solved reduction
    double sums_d[32]; // last solved reduction (compensated path)
    int done[kMaxLevels]; // per-level convergence flags
    int did_pixels;    // this launch ran its pixel pass: its partial rows are valid
    int level;         // level of this launch
    int rows;          // partial rows written (the launch's grid size)
    int iterations;    // Gauss-Newton steps solved so far
};

// Result of a track call, read back by the host.
struct IcpState {
    float T[12];
    float sums[32];
    double sums_d[32];
    int ok;
    int iterations;
    int pad[2];
};

struct IcpShared {
    IcpState result;
    IcpSlot slot[2];
    Iso T_ref;     // pose at the start of the call (restored by a failed check)
    Iso T_ref_inv; // its inverse: projection of world points into the reference maps
};

// Backprojection of pyramid level l: PinholeCamera(sensor, 2^l) (pinhole_camera.cpp:29-60).
struct PyrLevel {
    int w, h;
    float cx, cy, inv_fx, inv_fy;
    float4* vert; // one float4 {x, y, z, 0} per pixel
    float4* norm;
};

struct PyrArgs {
    const float* d0;    // full-resolution depth, w0 x h0
    float e_d;          // e_delta * 3 (tracker.hpp:20, tracker_impl.hpp:66-67)
    PyrLevel lv[kMaxLevels];
    int levels;
    int neg_y;          // point_cloud_to_normal<true> (left-handed sensor frame)
    int tile_log2;      // log2 of the level-0 tile side
    int tiles_x;        // level-0 tiles per row
    IcpShared* init;    // when not null, thread 0 writes the state that the first launch reads
    long long* bins;    // when not null, the round the first launch reads is cleared
    Iso T;              // start pose, unless T16_dev is set
    const float* T16_dev;
};

// LDS floats of one k_pyramid workgroup: the level-0 region of side tile + 2 * 2^(L-1) (the tile and
// a ring of 2^(L-1) pixels) and each coarser level's region of half the side.
inline size_t pyramid_lds_bytes(int levels)
{
    size_t n = 0;
    for (int l = 0, side = (1 << pyramid_tile_log2(levels)) + (2 << (levels - 1)); l < levels; l++, side >>= 1)
        n += (size_t) side * side;
    return n * sizeof(float);
}

// The whole input side of Tracker::track (tracker_impl.hpp:50-75) in one launch.  A workgroup owns
// an aligned tile x tile square of the full-resolution image.  A pixel of level l is the
// robust half-sample of an aligned 2^l x 2^l square of level 0 (half_sample_robust_image,
// preprocessor.cpp:156-195: its index clamp never moves a pixel of the image), so the workgroup
// stages its tile plus a ring of 2^(L-1) pixels in LDS, half-samples level by level in LDS (each
// level keeps a ring of at least one pixel, which the normals of the tile need), then writes the
// tile's vertices (depth_to_point_cloud, preprocessor_impl.hpp:18-40) and normals
// (point_cloud_to_normal, preprocessor.cpp:109-152) at every level.  Pixels of a region outside the
// image are never read by pixels inside it (the normals clamp their neighbours to the image).
template <int NT>
__device__ void pyramid_tile(const PyrArgs& a, int tile_index, float* pyr)
{
    if (a.init && tile_index == 0 && threadIdx.x == 0) {
        // the first Gauss-Newton launch (k = 0) reads slot 1
        IcpSlot* st = &a.init->slot[1];
        const Iso T = a.T16_dev ? iso_from16(a.T16_dev) : a.T;
        a.init->T_ref = T;
        a.init->T_ref_inv = iso_inv(T);
        for (int r = 0; r < 3; r++) {
            for (int c = 0; c < 3; c++)
                st->T[4 * r + c] = T.R[r][c];
            st->T[4 * r + 3] = T.t[r];
        }
        for (int k = 0; k < 32; k++) {
            st->sums[k] = 0.f;
            st->sums_d[k] = 0.0;
        }
        for (int l = 0; l < kMaxLevels; l++)
            st->done[l] = 0;
        st->did_pixels = 0;
        st->level = 0;
        st->rows = 0;
        st->iterations = 0;
    }
    if (a.bins && tile_index == 0)
        for (int i = threadIdx.x; i < 3 * kBins * 32; i += NT)
            a.bins[i] = 0;
    const int L = a.levels;
    const int S = 1 << (L - 1);                 // ring at level 0
    const int tile = 1 << a.tile_log2;
    const int tx = tile_index % a.tiles_x, ty = tile_index / a.tiles_x;
    const int ox = tx * tile - S, oy = ty * tile - S; // level-0 region origin, a multiple of 2^(L-1)
    int side = tile + 2 * S;
    {
        // every load of a batch is issued before any LDS store, so the region arrives in one round trip
        // (two at most for deep pyramids) rather than one per element
        constexpr int kBatch = 8;
        const int w0 = a.lv[0].w, h0 = a.lv[0].h, n0 = side * side;
        for (int base = 0; base < n0; base += kBatch * NT) {
            float v[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; j++) {
                const int i = base + j * NT + threadIdx.x;
                const int gx = ox + i % side, gy = oy + i / side;
                v[j] = (i < n0 && gx >= 0 && gx < w0 && gy >= 0 && gy < h0) ? a.d0[gx + (size_t) gy * w0] : 0.f;
            }
#pragma unroll
            for (int j = 0; j < kBatch; j++) {
                const int i = base + j * NT + threadIdx.x;
                if (i < n0)
                    pyr[i] = v[j];
            }
        }
    }
    int off = 0;
    for (int l = 1; l < L; l++) {
        __syncthreads();
        const int poff = off, pside = side;
        off += side * side;
        side >>= 1;
        const int lox = ox >> l, loy = oy >> l; // exact: ox, oy are multiples of 2^l
        for (int i = threadIdx.x; i < side * side; i += NT) {
            const int rx = i % side, ry = i / side;
            const int gx = lox + rx, gy = loy + ry;
            float out = 0.f;
            if (gx >= 0 && gx < a.lv[l].w && gy >= 0 && gy < a.lv[l].h) {
                const float* in = pyr + poff + 2 * ry * pside + 2 * rx;
                const float c = in[0];
                float count = 0.f, sum = 0.f;
#pragma unroll
                for (int ii = 0; ii <= 1; ii++) {
#pragma unroll
                    for (int jj = 0; jj <= 1; jj++) {
                        const float v = in[ii * pside + jj];
                        if (fabsf(v - c) < a.e_d) {
                            count += 1.f;
                            sum += v;
                        }
                    }
                }
                out = sum / count;
            }
            pyr[off + i] = out;
        }
    }
    __syncthreads();
    // vertices and normals of the tile's pixels, level by level (the level is uniform in the
    // workgroup, so its parameters are scalar loads): level l has (tile >> l)^2 of them
    for (int l = 0, loff = 0, lside = tile + 2 * S; l < L; loff += lside * lside, lside >>= 1, l++) {
        const PyrLevel P = a.lv[l];
        const int lg = a.tile_log2 - l; // log2(tile >> l)
        const int tside = tile >> l, ring = S >> l;
        for (int j = threadIdx.x; j < tside * tside; j += NT) {
            const int gx = (ox >> l) + ring + (j & (tside - 1)), gy = (oy >> l) + ring + (j >> lg);
            if (gx >= P.w || gy >= P.h)
                continue;
            const float* D = pyr + loff;
            const int rox = ox >> l, roy = oy >> l; // region origin at level l
            auto point = [&](int x, int y, float p[3]) {
                const float d = D[(x - rox) + (y - roy) * lside];
                if (d > 0.f) {
                    const float dir[3] = {((float) x - P.cx) * P.inv_fx, ((float) y - P.cy) * P.inv_fy, 1.f};
                    p[0] = d * dir[0];
                    p[1] = d * dir[1];
                    p[2] = d * dir[2];
                }
                else
                    p[0] = p[1] = p[2] = 0.f;
            };
            float p[3];
            point(gx, gy, p);
            float n[3] = {0.f, 0.f, 0.f};
            if (p[2] != 0.f) {
                const int xl = max(gx - 1, 0), xr = min(gx + 1, P.w - 1);
                const int ym = max(gy - 1, 0), yp = min(gy + 1, P.h - 1);
                const int yu = a.neg_y ? ym : yp, yd = a.neg_y ? yp : ym;
                float pl[3], pr[3], pu[3], pd[3];
                point(xl, gy, pl);
                point(xr, gy, pr);
                point(gx, yu, pu);
                point(gx, yd, pd);
                if (pl[2] != 0.f && pr[2] != 0.f && pu[2] != 0.f && pd[2] != 0.f) {
                    float dx[3], dy[3], cr[3];
                    for (int k = 0; k < 3; k++) {
                        dx[k] = pr[k] - pl[k];
                        dy[k] = pu[k] - pd[k];
                    }
                    cross3(dx, dy, cr);
                    normalized3(cr, n);
                }
            }
            const size_t o = gx + (size_t) gy * P.w;
            P.vert[o] = make_float4(p[0], p[1], p[2], 0.f);
            P.norm[o] = make_float4(n[0], n[1], n[2], 0.f);
        }
    }
}

// Arguments of the pyramid that starts the tracker's next call from depth_dev and the pose at
// T_in_dev (device pointers), with its tile count; the call with the same two pointers then skips
// its pyramid launch (se8_track.hip).
int tracker_prepare_pyramid(se8_tracker* tr, const float* depth_dev, const float* T_in_dev, PyrArgs* out,
                            int* n_tiles);

} // namespace se8
