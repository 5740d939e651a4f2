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

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/se8.h"

#define SE8_HD __host__ __device__ __forceinline__

// Pins v to a register of its own at this point of the program (an empty asm statement that reads
// and writes it), so that its uses stay after this point: the results of atomics issued in separate
// branches are then awaited together here, not each inside the branch that issued it.  The host
// build of the allocation kernels (the CPU bounds checks) defines it away.
#ifndef SE8_REG_BARRIER
#define SE8_REG_BARRIER(v) asm volatile("" : "+v"(v))
#endif

namespace se8 {

// ----------------------------------------------------------------------------------------------
// error handling
// ----------------------------------------------------------------------------------------------
void set_error(const std::string& msg);

// ----------------------------------------------------------------------------------------------
// frame graphs (se8_graph.hip): while a frame is being captured, the launches whose arguments change
// from frame to frame (input depth, pose pointers, timestamp) report the graph node they became, so
// that each replay can set those arguments.  A no-op unless this thread is capturing a frame graph.
// ----------------------------------------------------------------------------------------------
enum GraphNodeKind {
    kNodePyramid,      // k_pyramid(PyrArgs): d0, T16_dev
    kNodeIcpFinish,    // k_icp_finish(IcpArgs): T_out_dev, ok_dev
    kNodeFrameBegin,   // k_frame_begin(m, T_arg, T16_dev, fp): argument 2
    kNodeFrustumCarve, // k_frustum_carve(m, sensor, fp, depth, ...): argument 3
    kNodeUpdate,       // k_update(m, list, fp, sensor, depth, ts, link): arguments 4 and 5
    kNodeRaycast,      // k_raycast(m, cam, T_arg, T16_dev, ...): argument 3
    kNumGraphNodes
};
void graph_note_launch(int kind, hipStream_t s);

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

SE8_HD Iso iso_from16(const float* m)
{
    Iso T;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T.R[i][j] = m[4 * i + j];
        T.t[i] = m[4 * i + 3];
    }
    return T;
}

SE8_HD void iso_to16(const Iso& T, float* m)
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

// x / b rounded to nearest even, equal to IEEE division, from inv_b = 1 / b rounded to nearest.
// q0 = x * inv_b lies within 1.5 ulp of x / b; the correction q0 + (x - b q0) inv_b, whose
// residual one fma gives exactly, brings it within one ulp, and a second correction then gives the
// correctly rounded quotient by Markstein's theorem (reciprocal within half an ulp of 1 / b,
// quotient within one ulp: P. Markstein, "Computation of elementary functions on the IBM RISC
// System/6000 processor", IBM J. Res. Dev. 34(1), 1990).  The residuals are exact while they stay
// clear of underflow: 2^-60 <= |x| <= 2^60 with x / b a normal number (any map resolution between
// 2^-40 and 2^40 m).  Five instructions instead of the division's twelve on the device.
SE8_HD float div_by_recip(float x, float b, float inv_b)
{
    float q = x * inv_b;
    q = fmaf(fmaf(-q, b, x), inv_b, q);
    return fmaf(fmaf(-q, b, x), inv_b, q);
}

// x / b for any x: div_by_recip where it is exact, the division itself for smaller, zero and
// non-finite x
SE8_HD float div_rn(float x, float b, float inv_b)
{
    if (!(fabsf(x) >= 0x1p-60f && fabsf(x) <= 0x1p60f))
        return x / b;
    return div_by_recip(x, b, inv_b);
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


SE8_HD void cam_backproject(const Cam& c, float u, float v, float o[3])
{
    o[0] = (u - c.cx) * c.inv_fx;
    o[1] = (v - c.cy) * c.inv_fy;
    o[2] = 1.f;
}

SE8_HD bool cam_project(const Cam& c, const float p[3], float uv[2])
{
    // |z| < 1e-12 compared in double: 0x1.19799ap-40 is the least float above 1e-12, so the float
    // comparison below selects exactly the same z
    if (fabsf(p[2]) < 0x1.19799ap-40f)
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
// Ouster lidar (src/sensor/ouster_lidar.cpp:59-253, ouster_lidar_impl.hpp:15-32).  The beam model
// of srl::projection::OusterLidar is not part of the reference's sources; this is the library's
// own definition:
//   beam v: elevation e_v (rows strictly descending), azimuth offset a_v; column u looks along
//   phi = 2 pi u / W - pi + a_v; backProject(u, v) = (cos e_v cos phi, cos e_v sin phi, sin e_v),
//   evaluated in double on the host and rounded to float once (a per-pixel table);
//   project(p): e = atan2(p_z, |p_xy|) interpolated linearly between the beam elevations to a
//   fractional row (extrapolated by the end gaps), accepted for rows in [-0.5, H - 0.5);
//   u = (atan2(p_y, p_x) - a_row + pi) * W / (2 pi), wrapped into [-0.5, W - 0.5).
// atan and sin are fixed polynomials so that host and device round identically: atan(t) = t P(t^2)
// on [0, 1] (least-squares fit, max error 8.2e-8) and the Taylor series of sin to x^17 on [0, pi].
// ----------------------------------------------------------------------------------------------
constexpr float kPiF = (float) 3.14159265358979323846;
constexpr float kHalfPiF = (float) (3.14159265358979323846 / 2);

SE8_HD float atan_unit(float t)
{
    const float c[9] = {0.9999998807907104f,  -0.33332598209381104f, 0.19985906779766083f,
                        -0.141612246632576f,  0.10498932003974915f,  -0.07234828174114227f,
                        0.0397808775305748f,  -0.014401142485439777f, 0.0024566699285060167f};
    const float z = t * t;
    float r = c[8];
#pragma unroll
    for (int k = 7; k >= 0; k--)
        r = r * z + c[k];
    return t * r;
}

SE8_HD float atan2_det(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    if (mx == 0.f)
        return 0.f;
    float a = atan_unit(mn / mx);
    if (ay > ax)
        a = kHalfPiF - a;
    if (x < 0.f)
        a = kPiF - a;
    if (y < 0.f)
        a = -a;
    return a;
}

SE8_HD float sin_det(float x) // x in [0, pi]
{
    const float c[9] = {1.0f,
                        -0.1666666716337204f,
                        0.008333333767950535f,
                        -0.00019841270113829523f,
                        2.7557318844628753e-06f,
                        -2.5052107943679403e-08f,
                        1.6059044372074283e-10f,
                        -7.647163609812713e-13f,
                        2.8114573589663704e-15f};
    const float z = x * x;
    float r = c[8];
#pragma unroll
    for (int k = 7; k >= 0; k--)
        r = r * z + c[k];
    return x * r;
}

struct LidarModel {
    int w, h;
    float near_plane, far_plane;
    float min_elev, max_elev; // radians, as OusterLidar's constructor computes them
    float cols_per_rad;       // W / (2 pi)
    const float* elev;        // h beam elevations (radians, descending), device
    const float* az;          // h beam azimuth offsets (radians), device
    const float4* dir;        // h * w back-projection directions, row-major, device
    // Search hint for the elevation table: bucket b of nbucket equal slices of [elev[h-1], elev[0]]
    // (counted from the top) starts its search at row bucket[b].  Only the speed of the search
    // depends on it: the row found is the unique one with elev[row] >= e > elev[row + 1].
    const int* bucket;
    int nbucket;
    float bucket_scale; // nbucket / (elev[0] - elev[h - 1])
};

SE8_HD bool lidar_project(const LidarModel& L, const float p[3], float uv[2])
{
    const float hxy = sqrtf(p[0] * p[0] + p[1] * p[1]);
    if (hxy == 0.f && p[2] == 0.f)
        return false;
    const float e = atan2_det(p[2], hxy);
    const int h = L.h;
    float vf;
    if (e >= L.elev[0])
        vf = (L.elev[0] - e) / (L.elev[0] - L.elev[1]);
    else if (e <= L.elev[h - 1])
        vf = (float) (h - 1) + (L.elev[h - 1] - e) / (L.elev[h - 2] - L.elev[h - 1]);
    else if (L.bucket) {
        // elev[0] > e > elev[h - 1] (a NaN e fails both tests above and the range test below):
        // start at the bucket's row and walk to the unique lo with elev[lo] >= e > elev[lo + 1]
        const float u = (L.elev[0] - e) * L.bucket_scale;
        const int b = u >= 0.f ? min((int) u, L.nbucket - 1) : 0;
        int lo = L.bucket[b];
        while (lo > 0 && L.elev[lo] < e)
            lo--;
        while (lo < h - 2 && L.elev[lo + 1] >= e)
            lo++;
        vf = (float) lo + (L.elev[lo] - e) / (L.elev[lo] - L.elev[lo + 1]);
    }
    else {
        int lo = 0, hi = h - 1; // elev[lo] > e > elev[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (L.elev[mid] >= e)
                lo = mid;
            else
                hi = mid;
        }
        vf = (float) lo + (L.elev[lo] - e) / (L.elev[lo] - L.elev[lo + 1]);
    }
    if (!(vf >= -0.5f && vf < (float) h - 0.5f))
        return false;
    const int row = (int) (vf + 0.5f);
    float uf = ((atan2_det(p[1], p[0]) - L.az[row]) + kPiF) * L.cols_per_rad;
    for (int k = 0; k < 2 && uf >= (float) L.w - 0.5f; k++)
        uf = uf - (float) L.w;
    for (int k = 0; k < 2 && uf < -0.5f; k++)
        uf = uf + (float) L.w;
    if (!(uf >= -0.5f && uf < (float) L.w - 0.5f))
        return false;
    uv[0] = uf;
    uv[1] = vf;
    return true;
}

// OusterLidar::sphereInFrustumImpl (ouster_lidar.cpp:195-224), asin(c_z / |c|) as atan2(c_z, |c_xy|)
SE8_HD bool lidar_sphere_in_frustum(const LidarModel& L, const float c[3], float radius)
{
    const float n = sqrtf(sqnorm3(c));
    if (n - radius > L.far_plane)
        return false;
    if (n + radius < L.near_plane)
        return false;
    const float e = atan2_det(c[2], sqrtf(c[0] * c[0] + c[1] * c[1]));
    if (e < L.min_elev && sin_det(fabsf(e - L.min_elev)) * n > radius)
        return false;
    if (e > L.max_elev && sin_det(fabsf(e - L.max_elev)) * n > radius)
        return false;
    return true;
}

inline void plane_normal(const float p1[3], const float p2[3], const float p3[3], float out[3])
{
    float t1[3], t2[3], n[3];
    for (int i = 0; i < 3; i++) {
        t1[i] = p2[i] - p1[i];
        t2[i] = p3[i] - p2[i];
    }
    cross3(t1, t2, n);
    normalized3(n, out);
}

// scaled = 1 reproduces PinholeCamera(sensor, dsf), which applies the downsampling formulas
// even for dsf = 1.
inline Cam make_cam(const se8_pinhole& p, int dsf, bool scaled)
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

// The sensor interface of the reference's templated integrator (SensorBase: backProject, project,
// farDist, near_plane, measurementFromPoint, sphereInFrustum) for the two sensor types.
// sqrtf(x) for x >= 2^-96, +inf and NaN: the hardware root corrected by the signs of its two
// neighbours' residuals, the correctly rounded sequence without the scaling of small inputs and the
// selects for zero and infinity (below 2^-96 the residuals can underflow and the result is not the
// rounded root).
__device__ __forceinline__ float sqrt_rn_large(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float s_dn = __int_as_float(__float_as_int(s) - 1);
    const float s_up = __int_as_float(__float_as_int(s) + 1);
    float r = fmaf(-s_dn, s, x) <= 0.f ? s_dn : s;
    r = fmaf(-s_up, s, x) > 0.f ? s_up : r;
    return r;
}

struct PinholeSensor {
    Cam c;
    SE8_HD int width() const { return c.w; }
    SE8_HD int height() const { return c.h; }
    SE8_HD float near_plane() const { return c.near_plane; }
    SE8_HD float far_plane() const { return c.far_plane; }
    SE8_HD void backproject(int x, int y, float dir[3]) const { cam_backproject(c, (float) x, (float) y, dir); }
    SE8_HD bool project(const float p[3], float uv[2]) const { return cam_project(c, p, uv); }
    SE8_HD float far_dist(const float p[3]) const { return cam_far_dist(c, p); }
    // nrm > far_dist(p) with nrm = |p| (the updater's far test), without the two divisions in the
    // common case.  For z >= +0 (sign bit clear), z / nrm <= 1 rounds to at most 1, so far_dist >= far
    // and a point with nrm <= far is never beyond it.  Otherwise far_dist = far / (z / nrm) is
    // bracketed by an estimate with a relative error below 1e-6 (a hardware reciprocal and three
    // products), and only a point inside that bracket takes the exact path, so the result is that of
    // the exact expression.
    __device__ bool beyond_far(const float p[3], float nrm) const
    {
        if (__float_as_int(p[2]) >= 0 && nrm <= c.far_plane)
            return false;
        if (p[2] < 0.f)
            return true; // z / nrm < 0, so far_dist < 0 <= nrm
        if (p[2] > 0.f && nrm > 0.f) {
            const float est = c.far_plane * nrm * __builtin_amdgcn_rcpf(p[2]);
            const float tol = est * 1e-5f;
            if (nrm > est + tol)
                return true;
            if (nrm < est - tol)
                return false;
        }
        return nrm > cam_far_dist(c, p);
    }
    SE8_HD float measurement(const float p[3]) const { return p[2]; }
    SE8_HD bool sphere_in_frustum(const float ctr[3], float r) const { return cam_sphere_in_frustum(c, ctr, r); }
    // project() rejects every point with |p|^2 < 2^-96 (then |z| < 2^-48 < 1e-12), so the updater's
    // |p| only matters at or above 2^-96, where sqrt_rn_large is the rounded root
    static constexpr bool kRejectsNearOrigin = true;
};

struct LidarSensor {
    LidarModel L;
    SE8_HD int width() const { return L.w; }
    SE8_HD int height() const { return L.h; }
    SE8_HD float near_plane() const { return L.near_plane; }
    SE8_HD float far_plane() const { return L.far_plane; }
    __device__ void backproject(int x, int y, float dir[3]) const
    {
        const float4 d = L.dir[(size_t) y * L.w + x];
        dir[0] = d.x;
        dir[1] = d.y;
        dir[2] = d.z;
    }
    SE8_HD bool project(const float p[3], float uv[2]) const { return lidar_project(L, p, uv); }
    SE8_HD float far_dist(const float*) const { return L.far_plane; }
    __device__ bool beyond_far(const float*, float nrm) const { return nrm > L.far_plane; }
    SE8_HD float measurement(const float p[3]) const { return sqrtf(sqnorm3(p)); }
    SE8_HD bool sphere_in_frustum(const float ctr[3], float r) const { return lidar_sphere_in_frustum(L, ctr, r); }
    static constexpr bool kRejectsNearOrigin = false;
};

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

// Maps of at most this many voxels a side also keep a dense block table (at most 512 MiB): a
// block lookup is then one load instead of the brick id and the brick's slot.
constexpr int kDenseMaxSize = 4096;

// Everything a kernel needs to address the map, passed by value.
struct MapView {
    int size;         // octree size in voxels
    int max_scale;    // log2(size)
    int brick_shift;  // log2(blocks per brick side): bricks are octants of scale 3 + brick_shift
    int bricks_side;  // size >> (3 + brick_shift)
    float res;
    float inv_res;    // 1 / res rounded to nearest (div_rn)
    float dim;        // size * res
    float mu;         // res * truncation factor
    float max_weight;
    float t_MW[3];
    int storage;
    int* brick_top;     // bricks_side^3 brick ids (-1 = none)
    int* brick_slots;   // per brick 512 block slots (-1 = none), index (bx&m) + 8(by&m) + 64(bz&m)
    int* block_dense;   // maps of at most kDenseMaxSize voxels a side: the slot of every block
                        // (-1 = none), index (bz * S + by) * S + bx, S = size / 8; null otherwise
    int* node_dense;    // with block_dense: per scale s in [4, max_scale), the node index of every
                        // octant of that scale (-1 = none), at node_dense_off[s] + dense_index(s)
    int64_t node_dense_off[20];
    int* node_child;    // per node 8 children: node index, or block slot below scale-4 nodes
    uint64_t* node_key; // per node key
    int* node_ts;       // per node timestamp
    uint64_t* block_key;
    int* block_parent;  // scale-4 node owning each block
    int* block_ts;
    void* voxels;       // max_blocks * 512 voxels
    int* counters;      // [0] blocks, [1] nodes, [2] bricks, [3] overflow flag, [4..9] aabb
    int* block_scale;   // multi-resolution maps: per block {current scale, minimum scale}, -1 = none
    int64_t max_blocks, max_nodes, max_bricks;
};

// Layout of a map (se8_map_create): the octree size is power_two_up(max(ceil(max dim / res), 16))
// (octree_impl.hpp:14-19), bricks are octants of scale min(6, max_scale), the block pool holds
// tsdf_config.max_blocks blocks (default: every block of the octree, at most 2^18), the node pool
// 2 blocks + 64 nodes, the brick pool a quarter of the blocks + 64 bricks, and the dense node table
// one entry per octant of every scale in [4, max_scale), at node_dense_off[s] (node_dense_total
// entries in all).  Pointers are left null.  Returns false for an octree deeper than the 19 scales
// a key holds.
inline bool plan_map_view(const se8_map_config& cfg, const se8_tsdf_config& tcfg, MapView& v,
                          int64_t& node_dense_total)
{
    const float maxdim = std::max(cfg.dim[0], std::max(cfg.dim[1], cfg.dim[2]));
    int size = std::max((int) std::ceil(maxdim / cfg.res), 16);
    const int p = (int) std::ceil(std::log2((float) size));
    if (p > 19)
        return false;
    size = 1 << p;
    std::memset(&v, 0, sizeof(v));
    v.size = size;
    v.max_scale = p;
    const int brick_scale = std::min(6, p);
    v.brick_shift = brick_scale - kBlockScale;
    v.bricks_side = size >> brick_scale;
    v.res = cfg.res;
    v.inv_res = 1.0f / cfg.res;
    v.dim = (float) size * cfg.res;
    v.mu = cfg.res * tcfg.truncation_boundary_factor;
    v.max_weight = tcfg.max_weight;
    for (int i = 0; i < 3; i++)
        v.t_MW[i] = cfg.t_MW[i];
    v.storage = tcfg.storage;
    int64_t max_blocks = tcfg.max_blocks;
    if (max_blocks <= 0) {
        const int64_t blocks_side = size / 8;
        max_blocks = std::min<int64_t>(blocks_side * blocks_side * blocks_side, (int64_t) 1 << 18);
    }
    v.max_blocks = max_blocks;
    v.max_nodes = 2 * max_blocks + 64;
    const int64_t bricks_side3 = (int64_t) v.bricks_side * v.bricks_side * v.bricks_side;
    v.max_bricks = std::min<int64_t>(bricks_side3, max_blocks / 4 + 64);
    node_dense_total = 0;
    for (int sc = kBlockScale + 1; sc < p; sc++) {
        v.node_dense_off[sc] = node_dense_total;
        const int64_t sd = (int64_t) size >> sc;
        node_dense_total += sd * sd * sd;
    }
    return true;
}

// Per-integration poses derived from T_WS (in device memory, so that the pose itself may come
// from device memory): the sensor pose, its inverse for the frustum test and T_CV of the updater.
struct FramePoses {
    Iso T_WS;
    Iso T_SW;
    Iso T_CV;
};

// T_CV = T_CW * T_WM * Scaling(res) * Translation(0.5) (singleres_tsdf_updater_impl.hpp:37-42)
SE8_HD Iso compute_T_CV_from(const float t_MW[3], float res, const Iso& T_WC)
{
    const Iso T_CW = iso_inv(T_WC);
    Iso T_WM;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T_WM.R[i][j] = i == j ? 1.f : 0.f;
        T_WM.t[i] = -t_MW[i];
    }
    Iso A = iso_mul(T_CW, T_WM);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            A.R[i][j] = A.R[i][j] * res;
    for (int i = 0; i < 3; i++)
        A.t[i] = A.t[i] + (A.R[i][0] * 0.5f + (A.R[i][1] * 0.5f + A.R[i][2] * 0.5f));
    return A;
}

SE8_HD Iso compute_T_CV(const MapView& v, const Iso& T_WC)
{
    return compute_T_CV_from(v.t_MW, v.res, T_WC);
}


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
    kCntMrNew = 14,  // multi-resolution update: blocks allocated this frame
    kCntMrDown = 15, // multi-resolution update: blocks moving to a finer scale
    kCntMrScale = 16, // 4 counters: multi-resolution update list entries at scale 0..3
    kCntNodesStart = 20, // node count at the start of the integration in flight (its nodes are new)
    kCntBlocksStart = 21, // block count at the start of the integration in flight
    kNumCounters = 22
};

// The start of an integration (one thread): clears the per-frame counters, records where the frame's
// new nodes and blocks begin, and derives the frame's poses from T (k_frame_begin, or the tracker's
// last launch when it prepares the integration that follows it).
SE8_HD void frame_begin_state(int* counters, FramePoses* fp, const float t_MW[3], float res, const Iso& T)
{
    counters[kCntFrustum] = 0;
    counters[kCntUpdate] = 0;
    counters[kCntCandidates] = 0;
    counters[kCntUnique] = 0;
    counters[kCntMrNew] = 0;
    counters[kCntMrDown] = 0;
    counters[kCntNodesStart] = counters[kCntNodes];
    counters[kCntBlocksStart] = counters[kCntBlocks];
    for (int i = 0; i < 4; i++)
        counters[kCntMrScale + i] = 0;
    fp->T_WS = T;
    fp->T_SW = iso_inv(T);
    fp->T_CV = compute_T_CV_from(t_MW, res, T);
}

// Wave-aggregated append: every lane with pred gets the next free position of *counter (one
// atomic per wave); -1 for the other lanes.
__device__ __forceinline__ int wave_append_slot(bool pred, int* counter)
{
    const uint64_t mask = __ballot(pred);
    if (mask == 0)
        return -1;
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long) mask) - 1;
    int base = 0;
    if (lane == leader)
        base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    if (!pred)
        return -1;
    return base + __popcll(mask & ((1ull << lane) - 1ull));
}

// Index of the octant of scale s containing voxel c in a dense table of (size >> s)^3 entries.
__device__ __forceinline__ int64_t dense_index(const MapView& m, const int c[3], int s)
{
    const int64_t side = m.size >> s;
    return ((int64_t) (c[2] >> s) * side + (c[1] >> s)) * side + (c[0] >> s);
}

// Block slot of the block containing voxel (x,y,z); -1 if unallocated.  Coordinates are taken
// modulo the octree size, as the reference's top-down fetcher ignores bits at and above the
// root size (fetcher_impl.hpp:107-126); a world point within an ulp of the map's upper bound can
// round to voxel index == size.
__device__ __forceinline__ int lookup_block(const MapView& m, int x, int y, int z)
{
    const int wrap = m.size - 1;
    const int bx = (x & wrap) >> 3, by = (y & wrap) >> 3, bz = (z & wrap) >> 3;
    if (m.block_dense) {
        const int side = m.size >> 3;
        return m.block_dense[(bz * side + by) * side + bx];
    }
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
    if (m.block_dense) {
        const int side = m.size >> 3;
        return m.block_dense[(bz * side + by) * side + bx];
    }
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

// lookup_block / lookup_block_nb with the choice of block table fixed at compile time (Dense: the
// map has block_dense), for kernels instantiated per table.
template <bool Dense>
__device__ __forceinline__ int lookup_block_t(const MapView& m, int x, int y, int z)
{
    if constexpr (Dense) {
        const int wrap = m.size - 1, side = m.size >> 3;
        return m.block_dense[((((z & wrap) >> 3) * side) + ((y & wrap) >> 3)) * side + ((x & wrap) >> 3)];
    }
    else {
        const int wrap = m.size - 1;
        const int bx = (x & wrap) >> 3, by = (y & wrap) >> 3, bz = (z & wrap) >> 3;
        const int s = m.brick_shift;
        const int brick = m.brick_top[((bz >> s) * m.bricks_side + (by >> s)) * m.bricks_side + (bx >> s)];
        if (brick < 0)
            return -1;
        const int msk = (1 << s) - 1;
        return m.brick_slots[brick * 512 + (bx & msk) + ((by & msk) << 3) + ((bz & msk) << 6)];
    }
}
template <bool Dense>
__device__ __forceinline__ int lookup_block_nb_t(const MapView& m, int x, int y, int z)
{
    if constexpr (Dense) {
        return lookup_block_t<true>(m, x, y, z);
    }
    else {
        const int wrap = m.size - 1;
        const int bx = (x & wrap) >> 3, by = (y & wrap) >> 3, bz = (z & wrap) >> 3;
        const int s = m.brick_shift;
        const int brick = m.brick_top[((bz >> s) * m.bricks_side + (by >> s)) * m.bricks_side + (bx >> s)];
        const int msk = (1 << s) - 1;
        const int slot = m.brick_slots[max(brick, 0) * 512 + (bx & msk) + ((by & msk) << 3) + ((bz & msk) << 6)];
        return brick < 0 ? -1 : slot;
    }
}

// load_voxel for a storage fixed at compile time: kernels that load many voxels per thread are
// instantiated per storage so that no load sits behind a branch on the storage type.
template <int ST>
__device__ __forceinline__ void load_voxel_t(const MapView& m, int slot, int vi, float& tsdf, float& weight)
{
    static_assert(ST == SE8_STORAGE_F32 || ST == SE8_STORAGE_F16, "single-resolution storage");
    if constexpr (ST == SE8_STORAGE_F32) {
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

// Multi-resolution block layout (block.hpp:66-166): entries of scale s start at kMrOffset[s], x
// fastest, (8 >> s) per side; the data of the 585 entries, then their past data, float2 {tsdf,
// weight} each.
constexpr int kMrEntries = 585;
constexpr int kMrBlockFloat2 = 2 * kMrEntries;
SE8_HD int mr_offset(int s)
{
    return s == 0 ? 0 : (s == 1 ? 512 : (s == 2 ? 576 : 584));
}
// entry of the voxel with block-local coordinates (x, y, z) at scale s
SE8_HD int mr_index(int x, int y, int z, int s)
{
    const int n = 8 >> s;
    return mr_offset(s) + (x >> s) + (y >> s) * n + (z >> s) * n * n;
}

// FieldData<TSDF>::update (data_field_impl.hpp:39-54) of one voxel from a depth sample d (skipped
// unless d_ok and d >= near_plane); returns whether the voxel changed.
__device__ __forceinline__ bool tsdf_update_voxel(const MapView& m, float d, bool d_ok, float near_plane, float mz,
                                                  float nrm, float& tsdf, float& weight)
{
    if (!d_ok || d < near_plane)
        return false;
    const float sdf = (d - mz) / mz * nrm;
    if (sdf < -m.mu)
        return false;
    // std::min(sdf / mu, 1) (data_field_impl.hpp:45), which keeps a NaN quotient; sdf >= mu > 0
    // implies sdf / mu >= 1 (rounded division is monotone), so the division is only needed below mu
    float u = 1.f;
    if (!(sdf >= m.mu)) {
        const float q = sdf / m.mu;
        u = 1.f < q ? 1.f : q;
    }
    if (weight < m.max_weight)
        weight = weight + 1.f;
    tsdf = (tsdf * (weight - 1.f) + u) / weight;
    return true;
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
int device_count_sm(); // compute units of the current device

struct RadixSorter {
    DeviceBuffer keys_alt, tile_hist;
    size_t cap = 0;
    int reserve(size_t n);
    void release();
};

} // namespace se8

struct se8_map {
    se8_map_config cfg;
    se8_tsdf_config tsdf;
    se8::MapView view;
    se8::DeviceBuffer block_dense, node_dense;
    se8::DeviceBuffer brick_top, brick_slots, node_child, node_key, node_ts, block_key, block_parent, block_ts,
        voxels, counters, block_scale;
    // per-frame scratch
    se8::DeviceBuffer update_list, candidates, unique_keys, cur_node;
    se8::DeviceBuffer frame_poses; // FramePoses of the integration in flight
    se8::RadixSorter sorter;
    size_t candidates_cap = 0;
    int* host_counters = nullptr; // pinned copy of counters taken at the end of each integration
    int* host_counters_dev = nullptr; // the same memory as the device addresses it
    bool counters_pending = false; // a copy into host_counters is in flight (event ev[7])
    bool overflowed = false;       // a capacity overflow was reported: integration is refused from then on
    bool integ_pending = false;    // events ev[0..4] of the last integration not yet harvested
    bool ray_pending = false;      // events ev[5..6] of the last raycast not yet harvested
    // se8_track_dev_prepare_integrate started the next integration at the pose at this device
    // address (k_frame_begin's work): an integration at the same pose pointer skips that launch
    const float* prepared_T = nullptr;
    se8_stage_times times{};
    hipEvent_t ev[8]{};
    int timing = SE8_TIMING_STAGES; // se8_set_stage_timing
    // marching cubes: per-block face counts, scanned in place into offsets (+ the total); grown on
    // demand by the (logically const) se8_mesh
    mutable se8::DeviceBuffer mesh_offsets;
    se8::DeviceBuffer mr_new_scale; // multi-resolution update: each list entry's new scale
};

namespace se8 {
// Multi-resolution TSDF update of the frame's update list (pinhole sensor).
int launch_multires_update(se8_map* m, const FramePoses* fp, const PinholeSensor& sensor, const float* depth, int ts,
                           hipStream_t s);

// Entry points that read or write single-resolution voxels refuse a multi-resolution map.
inline bool is_multires(const se8_map* m);
} // namespace se8

struct se8_tracker {
    se8_pinhole sensor;
    se8_tracker_config cfg;
    se8::DeviceBuffer vert_lvl[SE8_TRACKER_MAX_LEVELS], norm_lvl[SE8_TRACKER_MAX_LEVELS]; // float4 per pixel, per level
    se8::DeviceBuffer bins;     // fixed-point partial sums of the Gauss-Newton launches, cleared by each track call
    se8::DeviceBuffer state;    // counters, flags, final sums
    se8::DeviceBuffer ref_points, ref_normals; // used by se8_track_raycast
    se8::DeviceBuffer adj_poses, adj_partials, adj_out; // batched cost/gradient evaluation
    int num_wg[SE8_TRACKER_MAX_LEVELS] = {};    // Gauss-Newton grid per level
    int icp_ppt[SE8_TRACKER_MAX_LEVELS] = {};   // pixels per thread per tile
    int icp_tiles[SE8_TRACKER_MAX_LEVELS] = {}; // tiles of 256 * ppt pixels
    int adj_wg = 0;                             // batched adjoint grid over level 0
    void* host_state = nullptr; // pinned readback of the device tracking state
    bool track_pending = false;
    se8_stage_times times{};
    hipEvent_t ev[2]{};
    int timing = SE8_TIMING_STAGES; // se8_set_stage_timing
    // SE8_TIMING_LEVELS: an event before each level's first launch (index = level) and one after the
    // last launch of the call (index SE8_TRACKER_MAX_LEVELS); created on first use
    hipEvent_t ev_lvl[SE8_TRACKER_MAX_LEVELS + 1]{};
    bool lvl_pending = false;
    // se8_raycast_dev_prepare_track built the pyramid of the next call for this depth and start pose
    // (device pointers): a track call with the same two skips its own pyramid launch
    const float* prepared_depth = nullptr;
    const float* prepared_T = nullptr;
};

struct se8_lidar {
    int width = 0, height = 0;
    se8::LidarModel model{}; // device pointers into elev / az / dir / bucket
    se8::DeviceBuffer elev, az, dir, bucket;
};

namespace se8 {
// stage-time bookkeeping (se8_map.hip)
void harvest_integration(se8_map* m);
void harvest_raycast(se8_map* m, bool block);
void harvest_tracking(se8_tracker* t);
} // namespace se8

inline bool se8::is_multires(const se8_map* m)
{
    return m->view.storage == SE8_STORAGE_MULTIRES;
}
