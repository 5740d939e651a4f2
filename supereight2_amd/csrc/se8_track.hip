// Projective point-to-plane ICP of libse8.so: se::Tracker::track
// (include/se/tracker/impl/tracker_impl.hpp:17-116).
//
// Per call: one launch builds the vertex and normal maps of all the pyramid levels (robust
// half-sampling preprocessor.cpp:156-195, preprocessor_impl.hpp:18-40, preprocessor.cpp:109-152;
// se8_pyramid.hpp) and initialises the device tracking state, unless a raycast launch already did
// (se8_raycast_dev_prepare_track); then one launch per Gauss-Newton iteration.  Launch k first
// solves the system of launch k - 1 in every workgroup (the exact sum of its fixed-point bins, 6x6
// solve across one wave, T <- exp(x) T and convergence test, icp.cpp:15-41, 208-223, so all
// workgroups agree bit for bit), then runs its own pixel pass: trackKernel (icp_impl.hpp:16-102)
// fused with the 32-term sums of reduceKernel (icp.cpp:45-204), per-thread accumulation over a
// tile of pixels whose loads are issued together, a halving wave shuffle and an LDS reduction to
// one partial row per workgroup, added into the launch's fixed-point bins.  Workgroups never wait
// on each other: the kernel boundary hands the sums and the pose to the next launch, so there is no
// device-wide fence.  Passes of a converged level are skipped, so the host never waits inside the
// pyramid loop.  A final one-workgroup launch solves the last pass and applies checkPoseKernel
// (icp.cpp:227-246).
//
// compensated = 0: fp32 sums within a workgroup, exact fixed-point sum across workgroups, fp32
// Cholesky (the reference's precision).
// compensated = 1: fp32 sums over a thread's few pixels and a fp32 pairwise tree within each wave,
// compensated (TwoSum sum and error pairs) combination of the workgroup's waves into its partial,
// exact fixed-point sum across workgroups and FP64 LDL^T solve.
//
// A launch after its level has converged (the reference's break, icp.cpp:208-223) with nothing left
// to solve only hands the state on, without waiting for the bins or reading its tile.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "se8_internal.hpp"
#include "se8_pyramid.hpp"

namespace se8 {

constexpr int kIcpThreads = 256;
// Cap on the batched adjoint's grid (its partials, kIcpMaxWg per pose); levels larger than that many
// tiles of kIcpThreads * 4 pixels loop over tiles.
constexpr int kIcpMaxWg = 320;
// Cap on a Gauss-Newton launch's grid: the workgroups hand over through the fixed-point bins, so a
// larger grid costs the next launch nothing to read.
constexpr int kIcpTrackMaxWg = 640;

struct IcpArgs {
    const float4* vert;    // level vertex map, one float4 {x, y, z, 0} per pixel
    const float4* norm;    // level normal map, one float4 {x, y, z, 0} per pixel
    const float* ref_pts;  // full-res raycast points (w*h*3, the caller's layout)
    const float* ref_nrm;  // full-res raycast normals
    long long* bins;       // three rounds of kBins x 32 fixed-point sums, by launch index mod 3
    IcpShared* sh;
    Iso T_ref_inv;         // batched adjoint only; the tracker reads sh->T_ref_inv
    Cam cam;               // full-resolution sensor (projection into the reference maps)
    int w, h;              // level size
    int n_tiles;           // tiles of kIcpThreads * PPT pixels
    float dist_th, normal_th, icp_th, track_th;
    int level;
    int k;                 // launch index within the call
    float* T_out_dev;      // k_icp_finish: the tracked pose also goes here (16 floats) when set
    int* ok_dev;           // k_icp_finish: the check result also goes here when set
    // k_icp_finish, when fb_counters is set: the start of the integration that follows at the
    // tracked pose (frame_begin_state on that map's counters and frame poses)
    int* fb_counters;
    FramePoses* fb_fp;
    float fb_t_MW[3];
    float fb_res;
};

// The pyramid over one 256-thread workgroup per tile (pyramid_tile, se8_pyramid.hpp).
__global__ __launch_bounds__(kPyrThreads) void k_pyramid(PyrArgs a)
{
    extern __shared__ float pyr[];
    pyramid_tile<kPyrThreads>(a, blockIdx.x, pyr);
}

__device__ __forceinline__ Iso pose_of(const float* T12)
{
    Iso T;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T.R[i][j] = T12[4 * i + j];
        T.t[i] = T12[4 * i + 3];
    }
    return T;
}

__device__ __forceinline__ void store_pose(const Iso& T, float* T12)
{
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T12[4 * i + j] = T.R[i][j];
        T12[4 * i + 3] = T.t[i];
    }
}

// trackKernel (icp_impl.hpp:16-102) for one pixel, split so that a batch of pixels can issue its
// loads together: icp_associate transforms and projects the input vertex (result -1 / -2, or 1 with
// the reference-map index), icp_residual applies the remaining tests to the gathered reference
// point and normal (result -3 / -4 / -5, or 1 with e and J).
struct IcpAssoc {
    float pW[3];
    float n_in[3];
    int ri;  // reference-map index (0 when the pixel failed association)
    int res; // 1: associated; -1: no input normal; -2: projection failed
};

__device__ __forceinline__ IcpAssoc icp_associate(const IcpArgs& a, const Iso& T_ref_inv, const Iso& T, float4 v4,
                                                  float4 n4)
{
    IcpAssoc o;
    o.n_in[0] = n4.x;
    o.n_in[1] = n4.y;
    o.n_in[2] = n4.z;
    const float v[3] = {v4.x, v4.y, v4.z};
    float pR[3], uv[2] = {0.f, 0.f};
    iso_apply(T, v, o.pW);
    iso_apply(T_ref_inv, o.pW, pR);
    const bool proj = cam_project(a.cam, pR, uv);
    const int px = proj ? (int) (uv[0] + 0.5f) : 0, py = proj ? (int) (uv[1] + 0.5f) : 0;
    o.ri = px + py * a.cam.w;
    o.res = (n4.x == 0.f && n4.y == 0.f && n4.z == 0.f) ? -1 : (proj ? 1 : -2);
    return o;
}

__device__ __forceinline__ int icp_residual(const IcpArgs& a, const Iso& T, const IcpAssoc& s, const float rp[3],
                                            const float rn_raw[3], float& e, float J[6])
{
    if (s.res != 1)
        return s.res;
    const float rn[3] = {-rn_raw[0], -rn_raw[1], -rn_raw[2]};
    if (rn[0] == 0.f && rn[1] == 0.f && rn[2] == 0.f)
        return -3;
    float diff[3], nW[3];
    for (int k = 0; k < 3; k++)
        diff[k] = rp[k] - s.pW[k];
    lin_apply(T, s.n_in, nW);
    if (sqrtf(sqnorm3(diff)) > a.dist_th)
        return -4;
    if (dot3(nW, rn) < a.normal_th)
        return -5;
    e = dot3(rn, diff);
    J[0] = rn[0];
    J[1] = rn[1];
    J[2] = rn[2];
    float cp[3];
    cross3(s.pW, rn, cp);
    J[3] = cp[0];
    J[4] = cp[1];
    J[5] = cp[2];
    return 1;
}

__device__ __forceinline__ void gather_ref(const IcpArgs& a, int ri, float rp[3], float rn[3])
{
    for (int k = 0; k < 3; k++) {
        rp[k] = a.ref_pts[3 * (size_t) ri + k];
        rn[k] = a.ref_nrm[3 * (size_t) ri + k];
    }
}

// One pixel of trackKernel; returns the result code and fills e, J on success.
__device__ __forceinline__ int icp_pixel(const IcpArgs& a, const Iso& T, int p, float& e, float J[6])
{
    const IcpAssoc s = icp_associate(a, a.T_ref_inv, T, a.vert[p], a.norm[p]);
    float rp[3], rn[3];
    gather_ref(a, s.ri, rp, rn);
    return icp_residual(a, T, s, rp, rn, e, J);
}

__device__ __forceinline__ void row_terms(int res, float e, const float J[6], float v[32])
{
    for (int k = 0; k < 32; k++)
        v[k] = 0.f;
    if (res < 1) {
        v[29] = res == -4 ? 1.f : 0.f;
        v[30] = res == -5 ? 1.f : 0.f;
        v[31] = res > -4 ? 1.f : 0.f;
        return;
    }
    v[0] = e * e;
    for (int k = 0; k < 6; k++)
        v[1 + k] = e * J[k];
    int c = 7;
    for (int i = 0; i < 6; i++)
        for (int j = i; j < 6; j++)
            v[c++] = J[i] * J[j];
    v[28] = 1.f;
}

// error-free sum: s + err == a + b exactly
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& err)
{
    s = a + b;
    const float bb = s - a;
    err = (a - (s - bb)) + (b - bb);
}

// The solve and the exponential map run in one wave at the start of every Gauss-Newton launch, so
// they sit on the launch chain's critical path.  They use the hardware reciprocal, reciprocal
// square root and square root (about 1 ulp) instead of the correctly rounded sequences, and
// polynomial sine and cosine for the small angles of a step: the tracked pose is compared with the
// reference to a tolerance (the sums' order already differs), and these differ from the rounded
// operations by an ulp.
__device__ __forceinline__ float fast_rcp(float x)
{
    return __builtin_amdgcn_rcpf(x);
}

// sin and cos: Taylor polynomials to x^9 / x^10 on |x| <= pi/4 (truncation below 2e-9, under half
// an ulp of the result), libm beyond
__device__ __forceinline__ void sincos_step(float x, float& sn, float& cs)
{
    if (fabsf(x) <= 0.785f) {
        const float x2 = x * x;
        sn = x * __builtin_fmaf(x2, __builtin_fmaf(x2, __builtin_fmaf(x2, __builtin_fmaf(x2, 1.f / 362880.f, -1.f / 5040.f),
                                                                      1.f / 120.f), -1.f / 6.f), 1.f);
        cs = __builtin_fmaf(x2, __builtin_fmaf(x2, __builtin_fmaf(x2, __builtin_fmaf(x2, __builtin_fmaf(x2, -1.f / 3628800.f,
                                                                                                         1.f / 40320.f),
                                                                                     -1.f / 720.f), 1.f / 24.f), -0.5f), 1.f);
    }
    else {
        sn = sinf(x);
        cs = cosf(x);
    }
}

// reciprocal in FP64: the hardware estimate refined by two Newton steps (error below an ulp)
__device__ __forceinline__ double rcp_f64(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
    return __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
}

// ---- the per-launch solve across the lanes of one wave ----
// The 6x6 solve (icp.cpp:15-41), the exponential map (math::exp, math_util_impl.hpp:117-174) and the
// pose product T <- exp(x) T (icp.cpp:208-223) sit on every Gauss-Newton launch's critical path.  Wave
// 0 runs them with the rows of the system spread over its lanes: the factorisation is right-looking
// (step k takes the pivot from lane k and updates the remaining rows in their own lanes), which
// performs every entry's subtractions in the same order as a row-by-row factorisation; the
// exponential map and the product give one output entry per lane.
__device__ __forceinline__ float bcast_f(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ double bcast_d(double v, int lane)
{
    const unsigned long long u = (unsigned long long) __double_as_longlong(v);
    const unsigned lo = (unsigned) __builtin_amdgcn_readlane((int) (unsigned) u, lane);
    const unsigned hi = (unsigned) __builtin_amdgcn_readlane((int) (unsigned) (u >> 32), lane);
    return __longlong_as_double((long long) (((unsigned long long) hi << 32) | lo));
}

// index of J^T J entry (i, j), i <= j, in the row-major upper triangle that follows the 6 J^T r terms:
// row i starts at i * (11 - i) / 2 = 0, 5, 9, 12, 14, 15 (a byte table in one 64-bit constant)
__device__ __forceinline__ int tri6(int i, int j)
{
    return (int) ((0x0F0E0C090500ull >> (8 * i)) & 0xff) + j;
}

// 6x6 Cholesky in fp32 (icp.cpp:33-41) across the lanes of one wave, every lane returning x; x = 0 if
// the factorisation fails.  Lane r holds row r of A (lanes above 5 repeat row 5); only the reciprocal
// of each pivot's square root is needed (one hardware rsq per column).
__device__ void solve_f32_wave(const float* vals, float x[6])
{
    const int r = min((int) (threadIdx.x & 63), 5);
    float A[6];
#pragma unroll
    for (int c = 0; c < 6; c++)
        A[c] = vals[6 + tri6(min(r, c), max(r, c))];
    float b = vals[r];
    float inv[6], Lrow[6], Lcol[6][6]; // Lcol[k][j] = L[j][k] (j > k), the same in every lane
    // straight-line code (a failed pivot is remembered, the result discarded at the end), so that the
    // loads and broadcasts of the steps can be scheduled together
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const float d = bcast_f(A[k], k); // A[k][k] less its k subtractions
        ok = ok && d > 0.f;
        inv[k] = __builtin_amdgcn_rsqf(d);
        const float lk = A[k] * inv[k]; // L[r][k] in lanes r > k
        Lrow[k] = lk;
#pragma unroll
        for (int j = k + 1; j < 6; j++)
            Lcol[k][j] = bcast_f(lk, j);
#pragma unroll
        for (int j = k + 1; j < 6; j++) {
            const float u = A[j] - lk * Lcol[k][j];
            A[j] = j <= r ? u : A[j];
        }
    }
    float y[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        y[k] = bcast_f(b, k) * inv[k];
        const float u = b - Lrow[k] * y[k];
        b = r > k ? u : b;
    }
#pragma unroll
    for (int i = 5; i >= 0; i--) {
        float v = y[i];
#pragma unroll
        for (int j = i + 1; j < 6; j++)
            v -= Lcol[i][j] * x[j];
        x[i] = v * inv[i];
    }
#pragma unroll
    for (int i = 0; i < 6; i++)
        x[i] = ok ? x[i] : 0.f;
}

// 6x6 LDL^T in FP64 (no pivoting: J^T J is symmetric positive semi-definite) across the lanes of one
// wave, as solve_f32_wave; x = 0 on failure.
__device__ void solve_f64_wave(const double* vals, float x[6])
{
    const int r = min((int) (threadIdx.x & 63), 5);
    double A[6];
#pragma unroll
    for (int c = 0; c < 6; c++)
        A[c] = vals[6 + tri6(min(r, c), max(r, c))];
    double b = vals[r];
    double D[6], rD[6], Lrow[6], Lcol[6][6];
    bool ok = true; // as in solve_f32_wave
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const double d = bcast_d(A[k], k);
        ok = ok && d > 0.0;
        D[k] = d;
        rD[k] = rcp_f64(d);
        const double lk = A[k] * rD[k];
        Lrow[k] = lk;
#pragma unroll
        for (int j = k + 1; j < 6; j++)
            Lcol[k][j] = bcast_d(lk, j);
#pragma unroll
        for (int j = k + 1; j < 6; j++) {
            const double u = A[j] - lk * Lcol[k][j] * D[k];
            A[j] = j <= r ? u : A[j];
        }
    }
    double y[6], z[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        y[k] = bcast_d(b, k);
        const double u = b - Lrow[k] * y[k];
        b = r > k ? u : b;
    }
#pragma unroll
    for (int i = 0; i < 6; i++)
        y[i] *= rD[i];
#pragma unroll
    for (int i = 5; i >= 0; i--) {
        double v = y[i];
#pragma unroll
        for (int k = i + 1; k < 6; k++)
            v -= Lcol[i][k] * z[k];
        z[i] = v;
    }
#pragma unroll
    for (int i = 0; i < 6; i++)
        x[i] = ok ? (float) z[i] : 0.f;
}

// T <- exp(a) T across the lanes of one wave, for the twist a = [v; w] with the reference's formulas
// (math_util_impl.hpp:117-174: quaternion of the half angle; V = I + c1 W + c2 W^2): lane e < 12 writes
// entry e of the row-major 3x4 pose (row e >> 2; column e & 3, the last one the translation).  Every
// lane reads the old pose before any lane writes it (one wave, in order).
__device__ void exp_update_wave(const float a[6], float* T12)
{
    const int e = (int) (threadIdx.x & 63);
    const int i = min(e >> 2, 2), j = e & 3;
    const float w[3] = {a[3], a[4], a[5]};
    const float theta_sq = dot3(w, w);
    const float theta = __builtin_amdgcn_sqrtf(theta_sq);
    const float half_theta = 0.5f * theta;
    float imag, re, s_t = 0.f, c_t = 1.f;
    const bool small = theta < 1e-10f;
    if (small) {
        const float t4 = theta_sq * theta_sq;
        imag = 0.5f - (1.0f / 48.0f) * theta_sq + (1.0f / 3840.0f) * t4;
        re = 1.f - 0.5f * theta_sq + (1.f / 384.f) * t4;
    }
    else {
        float s_h;
        sincos_step(half_theta, s_h, re);
        sincos_step(theta, s_t, c_t);
        imag = s_h * fast_rcp(theta);
    }
    float q[4] = {re, imag * w[0], imag * w[1], imag * w[2]};
    const float rq = __builtin_amdgcn_rsqf(q[0] * q[0] + (q[1] * q[1] + (q[2] * q[2] + q[3] * q[3])));
    for (int k = 0; k < 4; k++)
        q[k] = q[k] * rq;
    const float qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    const float tx = 2.f * qx, ty = 2.f * qy, tz = 2.f * qz;
    const float twx = tx * qw, twy = ty * qw, twz = tz * qw;
    const float txx = tx * qx, txy = ty * qx, txz = tz * qx;
    const float tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
    // row i of E.R
    float Ri[3];
    Ri[0] = i == 0 ? 1.f - (tyy + tzz) : (i == 1 ? txy + twz : txz - twy);
    Ri[1] = i == 0 ? txy - twz : (i == 1 ? 1.f - (txx + tzz) : tyz + twx);
    Ri[2] = i == 0 ? txz + twy : (i == 1 ? tyz - twx : 1.f - (txx + tyy));
    // E.t[i] = V[i] . v, V = R for small angles, else I + c1 W + c2 W^2
    float Vi[3];
    if (small) {
        for (int c = 0; c < 3; c++)
            Vi[c] = Ri[c];
    }
    else {
        const float W[3][3] = {{0.f, -w[2], w[1]}, {w[2], 0.f, -w[0]}, {-w[1], w[0], 0.f}};
        const float Wi[3] = {i == 0 ? W[0][0] : (i == 1 ? W[1][0] : W[2][0]), i == 0 ? W[0][1] : (i == 1 ? W[1][1] : W[2][1]),
                             i == 0 ? W[0][2] : (i == 1 ? W[1][2] : W[2][2])};
        const float r2 = fast_rcp(theta_sq);
        const float c1 = (1.f - c_t) * r2;
        const float c2 = (theta - s_t) * (r2 * fast_rcp(theta));
        for (int c = 0; c < 3; c++) {
            const float w2 = Wi[0] * W[0][c] + (Wi[1] * W[1][c] + Wi[2] * W[2][c]);
            Vi[c] = ((i == c ? 1.f : 0.f) + c1 * Wi[c]) + c2 * w2;
        }
    }
    const float ti = Vi[0] * a[0] + (Vi[1] * a[1] + Vi[2] * a[2]);
    // column j of the old pose (R column, or t for j = 3)
    const float B0 = T12[j], B1 = T12[4 + j], B2 = T12[8 + j];
    float out = Ri[0] * B0 + (Ri[1] * B1 + Ri[2] * B2);
    if (j == 3)
        out = out + ti;
    if (e < 12)
        T12[e] = out;
}

// checkPoseKernel (icp.cpp:227-246) on the last reduction: fail if the RMS residual exceeds 2 cm or
// the inlier fraction of the full image is below track_th, and then restore the pose of the call.
__device__ void check_pose(IcpState* st, const Iso& T_ref, int W, int H, float track_th, int comp)
{
    bool fail;
    if (comp) {
        fail = (sqrt(st->sums_d[0] / st->sums_d[28]) > 2e-2) || (st->sums_d[28] / (double) (W * H) < track_th);
    }
    else {
        fail = ((double) sqrtf(st->sums[0] / st->sums[28]) > 2e-2) || (st->sums[28] / (float) (W * H) < track_th);
    }
    if (fail) {
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++)
                st->T[4 * i + j] = T_ref.R[i][j];
            st->T[4 * i + 3] = T_ref.t[i];
        }
    }
    st->ok = fail ? 0 : 1;
}

// Adds (b, cb) into (s, cs): plain fp32 add, or an error-free TwoSum whose rounding error joins the
// compensation term.
template<bool Comp>
__device__ __forceinline__ void pair_add(float& s, float& cs, float b, float cb)
{
    if (Comp) {
        float sum, err;
        two_sum(s, b, sum, err);
        s = sum;
        cs = cs + cb + err;
    }
    else {
        s += b;
    }
}

// Sums the 32 terms s[k] over the 64 lanes of a wave by halving exchanges: at each step a lane
// keeps half of its terms and receives its partner's copy of that half (16 + 8 + 4 + 2 + 1 + 1 = 32
// exchanges in place of 6 x 32).  Afterwards s[0] of lanes 2t and 2t+1 holds the wave total of term
// t.  Every exchange is a VALU lane permutation, no LDS traffic: the two cross-row steps use the
// permlane swaps (partner = lane ^ 32, then lane ^ 16), the in-row steps DPP (row mirror: lane ^ 15,
// half-row mirror: lane ^ 7, quad permutations: lane ^ 2, lane ^ 1).  Each step's partner differs
// from the lane in exactly the bit that decides which half it keeps, and the masks 32, 16, 15, 7,
// 2, 1 are independent, so every lane's total covers all 64 lanes once.  The order of the
// additions is fixed, so results do not depend on scheduling.

// Cross-row step (O = 32 or 16): for each pair (s[i], s[i + O/2]) one swap leaves, in every lane,
// its own copy of the term it keeps and its partner's copy (in either order: the fp32 sum and the
// TwoSum error term are symmetric).
template<bool Comp, int O>
__device__ __forceinline__ void swap_step(float s[32], float c[32])
{
#pragma unroll
    for (int i = 0; i < O / 2; i++) {
        const auto r = O == 32 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(s[i]), __float_as_uint(s[i + O / 2]),
                                                                 false, false)
                               : __builtin_amdgcn_permlane16_swap(__float_as_uint(s[i]), __float_as_uint(s[i + O / 2]),
                                                                  false, false);
        float a = __uint_as_float(r[0]), ca = 0.f, cb = 0.f;
        const float b = __uint_as_float(r[1]);
        if (Comp) {
            const auto rc = O == 32 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(c[i]),
                                                                      __float_as_uint(c[i + O / 2]), false, false)
                                    : __builtin_amdgcn_permlane16_swap(__float_as_uint(c[i]),
                                                                       __float_as_uint(c[i + O / 2]), false, false);
            ca = __uint_as_float(rc[0]);
            cb = __uint_as_float(rc[1]);
        }
        pair_add<Comp>(a, ca, b, cb);
        s[i] = a;
        c[i] = ca;
    }
}

template<int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// In-row step over the DPP permutation CTRL (an involution pairing the lanes with `up` clear with
// those with `up` set): the O/2 surviving terms are compile-time register indices.
template<bool Comp, int O, int CTRL>
__device__ __forceinline__ void dpp_step(float s[32], float c[32], bool up)
{
#pragma unroll
    for (int i = 0; i < O / 2; i++) {
        const float keep = up ? s[i + O / 2] : s[i];
        const float recv = dpp_f<CTRL>(up ? s[i] : s[i + O / 2]);
        float keepc = 0.f, recvc = 0.f;
        if (Comp) {
            keepc = up ? c[i + O / 2] : c[i];
            recvc = dpp_f<CTRL>(up ? c[i] : c[i + O / 2]);
        }
        s[i] = keep;
        c[i] = keepc;
        pair_add<Comp>(s[i], c[i], recv, recvc);
    }
}

constexpr int kDppRowMirror = 0x140;     // lane ^ 15 within a row of 16
constexpr int kDppRowHalfMirror = 0x141; // lane ^ 7 within a half-row of 8
constexpr int kDppQuad2301 = 0x4e;       // lane ^ 2
constexpr int kDppQuad1032 = 0xb1;       // lane ^ 1

template<bool Comp>
__device__ __forceinline__ void wave_reduce32(float s[32], float c[32])
{
    const int lane = threadIdx.x & 63;
    swap_step<Comp, 32>(s, c);
    swap_step<Comp, 16>(s, c);
    dpp_step<Comp, 8, kDppRowMirror>(s, c, (lane & 8) != 0);
    dpp_step<Comp, 4, kDppRowHalfMirror>(s, c, (lane & 4) != 0);
    dpp_step<Comp, 2, kDppQuad2301>(s, c, (lane & 2) != 0);
    const float recv = dpp_f<kDppQuad1032>(s[0]);
    const float recvc = Comp ? dpp_f<kDppQuad1032>(c[0]) : 0.f;
    pair_add<Comp>(s[0], c[0], recv, recvc);
}

// Input vertex/normal of the PPT pixels of one tile (independent of the pose, so they can be in
// flight while the previous system is solved).
template<int PPT>
__device__ __forceinline__ void load_tile(const IcpArgs& a, int tile, int n, float4 v4[PPT], float4 n4[PPT])
{
    const int base = tile * (kIcpThreads * PPT) + threadIdx.x;
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        const int p = min(base + j * kIcpThreads, n - 1);
        v4[j] = a.vert[p];
        n4[j] = a.norm[p];
    }
}

// Solving the previous launch's pass is split so that its memory round trips overlap: load_bins
// issues the load of the previous launch's fixed-point sums without waiting, slot_word /
// put_slot_word bring the previous slot into LDS, gather_sums adds the bins and solve_step solves.
// Identical in every workgroup.
// The previous launch's slot is copied to LDS as whole 32-bit words, one per thread, so the copy
// costs a single load per thread: slot_word issues it, put_slot_word stores it (the caller
// synchronises before reading `cur`).
constexpr int kSlotWords = (int) (sizeof(IcpSlot) / 4);
static_assert(sizeof(IcpSlot) % 8 == 0 && kSlotWords <= kIcpThreads, "slot copy: one word per thread");

__device__ __forceinline__ uint32_t slot_word(const IcpSlot* prev)
{
    const int t = threadIdx.x;
    return t < kSlotWords ? reinterpret_cast<const uint32_t*>(prev)[t] : 0u;
}

__device__ __forceinline__ void put_slot_word(IcpSlot& cur, uint32_t w)
{
    const int t = threadIdx.x;
    if (t < kSlotWords)
        reinterpret_cast<uint32_t*>(&cur)[t] = w;
}

// Fixed-point scale of term t of a row: 2^40 for e^2 and e J (t < 7), 2^24 for J^T J (t < 28), 1 for
// the counts.  |sum| stays below 2^23 resp. 2^39 for any frame of up to 2^20 pixels at ranges up to
// 100 m; out-of-range values (and NaN) saturate.
__device__ __forceinline__ double fixed_scale(int t)
{
    return t < 7 ? 1099511627776.0 : (t < 28 ? 16777216.0 : 1.0);
}

__device__ __forceinline__ unsigned long long to_fixed(double v, int t)
{
    const double lim = 4.0e18;
    const double x = fmin(fmax(v * fixed_scale(t), -lim), lim);
    return (unsigned long long) (long long) rint(x);
}

// The previous launch's round, kBinWords entries per thread (entry t + 256 i: bin (t >> 5) + 8 i,
// term t & 31); workgroup 0 also clears the round the next launch adds into.  Issues the loads
// without waiting.
struct BinWords {
    long long v[kBinWords];
};

__device__ __forceinline__ BinWords load_bins(const IcpArgs& a)
{
    const int t = threadIdx.x;
    BinWords w;
#pragma unroll
    for (int i = 0; i < kBinWords; i++)
        w.v[i] = a.bins[((a.k + 2) % 3) * (kBins * 32) + t + 256 * i];
    if (blockIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < kBinWords; i++)
            a.bins[((a.k + 1) % 3) * (kBins * 32) + t + 256 * i] = 0;
    }
    return w;
}

// Sums of the previous launch's pass into cur.sums / cur.sums_d: the kBins fixed-point entries of
// each term added exactly, the total converted to FP64 (and rounded to fp32 for the plain solve).
// Ends with a barrier.
__device__ void gather_sums(long long* lb, const BinWords& bw, IcpSlot& cur)
{
    const int t = threadIdx.x;
    long long own = 0;
#pragma unroll
    for (int i = 0; i < kBinWords; i++)
        own += bw.v[i];
    lb[t] = own;
    __syncthreads();
    if (t < 32) {
        long long acc = 0;
        for (int b = 0; b < 8; b++)
            acc += lb[b * 32 + t];
        const double v = (double) acc / fixed_scale(t);
        cur.sums_d[t] = v;
        cur.sums[t] = (float) v;
    }
    __syncthreads();
}

// Solves the system of cur.sums (plain: fp32 Cholesky; compensated: FP64 LDL^T), applies
// T <- exp(x) T and raises the level's convergence flag (icp.cpp:15-41, 208-223).  Ends with a
// barrier.
template<bool Comp>
__device__ __forceinline__ void solve_step(const IcpArgs& a, IcpSlot& cur)
{
    if (threadIdx.x < 64) {
        float x[6];
        if (Comp)
            solve_f64_wave(cur.sums_d + 1, x);
        else
            solve_f32_wave(cur.sums + 1, x);
        exp_update_wave(x, cur.T);
        if (threadIdx.x == 0) {
            const float xn = sqrtf(x[0] * x[0] + (x[1] * x[1] + (x[2] * x[2] + (x[3] * x[3] + (x[4] * x[4] + x[5] * x[5])))));
            if (xn < a.icp_th)
                cur.done[cur.level] = 1;
            cur.iterations += 1;
        }
    }
    __syncthreads();
}

// Launch k of the Gauss-Newton sequence of Tracker::track (tracker_impl.hpp:77-108): solve the
// pass of launch k - 1 (slot_word, load_bins, gather_sums, solve_step), then, unless this level has
// converged, the pixel pass of iteration k at this level: trackKernel (icp_impl.hpp:16-102) fused
// with the 32-term reduceKernel sums (icp.cpp:45-204) over tiles of PPT pixels per thread, a halving
// wave reduction and an LDS reduction to one partial row per workgroup, added into the launch's
// fixed-point bins.  Workgroup 0 publishes the pose and flags for launch k + 1.  No workgroup waits
// on another, so no device-wide fence is needed; the kernel boundary orders the hand-over.
template<bool Comp, int PPT>
__global__ __launch_bounds__(kIcpThreads) void k_icp_iter(IcpArgs a)
{
    __shared__ IcpSlot cur;
    __shared__ long long lb[256];
    __shared__ float red[kIcpThreads / 64][64];
    const IcpSlot* prev = &a.sh->slot[(a.k + 1) & 1];
    const int n = a.w * a.h;
    // issue order = wait order: the slot word first, then the bins; the slot is stored (waiting for
    // it alone) and read before the bins are awaited
    const uint32_t sw = slot_word(prev);
    const BinWords bv = load_bins(a);
    put_slot_word(cur, sw);
    __syncthreads();
    // a launch after this level converged with nothing left to solve only hands the slot on; it
    // does not wait for the bins (their load was issued ahead, in case it was needed) and does not
    // read its tile.  Any other launch issues its tile's loads now: they arrive while the previous
    // pass is solved.
    const bool noop = cur.did_pixels == 0 && cur.done[a.level] != 0;
    float4 v4[PPT], n4[PPT];
    if (!noop)
        load_tile<PPT>(a, blockIdx.x, n, v4, n4);
    bool run = false;
    if (!noop) {
        if (cur.did_pixels) {
            gather_sums(lb, bv, cur);
            solve_step<Comp>(a, cur);
        }
        run = cur.done[a.level] == 0;
    }
    if (run) {
        float s[32], c[32];
        for (int k = 0; k < 32; k++) {
            s[k] = 0.f;
            c[k] = 0.f;
        }
        const Iso T = pose_of(cur.T);
        const Iso T_ref_inv = a.sh->T_ref_inv;
        for (int tile = blockIdx.x; tile < a.n_tiles; tile += gridDim.x) {
            if (tile != (int) blockIdx.x)
                load_tile<PPT>(a, tile, n, v4, n4);
            const int base = tile * (kIcpThreads * PPT) + threadIdx.x;
            IcpAssoc as[PPT];
            float rp[PPT][3], rn[PPT][3];
#pragma unroll
            for (int j = 0; j < PPT; j++)
                as[j] = icp_associate(a, T_ref_inv, T, v4[j], n4[j]);
#pragma unroll
            for (int j = 0; j < PPT; j++)
                gather_ref(a, as[j].ri, rp[j], rn[j]);
#pragma unroll
            for (int j = 0; j < PPT; j++) {
                if (base + j * kIcpThreads >= n)
                    continue;
                float e = 0.f, J[6];
                const int res = icp_residual(a, T, as[j], rp[j], rn[j], e, J);
                float v[32];
                row_terms(res, e, J, v);
                // a thread sums only the pixels of its tiles (PPT per tile; one tile per workgroup
                // at 640x480), a few fp32 roundings
                for (int k = 0; k < 32; k++)
                    s[k] += v[k];
            }
        }
        // within a wave a plain fp32 pairwise tree (6 levels); the compensated path starts at the
        // combination of the workgroup's waves, then FP64 across workgroups
        wave_reduce32<false>(s, c);
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if ((lane & 1) == 0) {
            red[wave][lane >> 1] = s[0];
            red[wave][32 + (lane >> 1)] = Comp ? c[0] : 0.f;
        }
        __syncthreads();
        if (threadIdx.x < 32) {
            const int k = threadIdx.x;
            float ss = red[0][k], cc = red[0][32 + k];
            for (int w = 1; w < kIcpThreads / 64; w++)
                pair_add<Comp>(ss, cc, red[w][k], red[w][32 + k]);
            unsigned long long* bin = reinterpret_cast<unsigned long long*>(a.bins)
                + (a.k % 3) * (kBins * 32) + (blockIdx.x % kBins) * 32;
            atomicAdd(bin + k, to_fixed(Comp ? (double) ss + (double) cc : (double) ss, k));
        }
    }
    if (blockIdx.x == 0) {
        IcpSlot* out = &a.sh->slot[a.k & 1];
        const int t = threadIdx.x;
        if (t < 12)
            out->T[t] = cur.T[t];
        if (t < 32) {
            out->sums[t] = cur.sums[t];
            out->sums_d[t] = cur.sums_d[t];
        }
        if (t < kMaxLevels)
            out->done[t] = cur.done[t];
        if (t == 0) {
            out->did_pixels = run ? 1 : 0;
            out->level = a.level;
            out->rows = gridDim.x;
            out->iterations = cur.iterations;
        }
    }
}

// After the last Gauss-Newton launch (a.k = number of launches): solve its pass, then
// checkPoseKernel (icp.cpp:227-246) on the last reduction, into the result the host reads.
template<bool Comp>
__global__ __launch_bounds__(kIcpThreads) void k_icp_finish(IcpArgs a)
{
    __shared__ IcpSlot cur;
    __shared__ long long lb[256];
    const uint32_t sw = slot_word(&a.sh->slot[(a.k + 1) & 1]);
    const BinWords bv = load_bins(a);
    put_slot_word(cur, sw);
    __syncthreads();
    if (cur.did_pixels) {
        gather_sums(lb, bv, cur);
        solve_step<Comp>(a, cur);
    }
    if (threadIdx.x == 0) {
        IcpState* r = &a.sh->result;
        for (int i = 0; i < 12; i++)
            r->T[i] = cur.T[i];
        for (int k = 0; k < 32; k++) {
            r->sums[k] = cur.sums[k];
            r->sums_d[k] = cur.sums_d[k];
        }
        r->iterations = cur.iterations;
        check_pose(r, a.sh->T_ref, a.cam.w, a.cam.h, a.track_th, Comp);
        if (a.T_out_dev) {
            for (int q = 0; q < 12; q++)
                a.T_out_dev[q] = r->T[q];
            a.T_out_dev[12] = a.T_out_dev[13] = a.T_out_dev[14] = 0.f;
            a.T_out_dev[15] = 1.f;
        }
        if (a.ok_dev)
            *a.ok_dev = r->ok;
        if (a.fb_counters)
            frame_begin_state(a.fb_counters, a.fb_fp, a.fb_t_MW, a.fb_res, iso_from16(r->T));
    }
}

// Vertex and normal maps of the first `levels` pyramid levels in one launch; writes the state the
// first Gauss-Newton launch reads (pose T) when init is not null.
static PyrArgs pyramid_args(se8_tracker* tr, const float* depth, int levels, IcpShared* init, const Iso& T,
                            const float* T16_dev, long long* bins, int* n_tiles)
{
    const se8_pinhole& P = tr->sensor;
    PyrArgs pa{};
    pa.d0 = depth;
    pa.e_d = 0.1f * 3; // e_delta * 3 (tracker.hpp:20, tracker_impl.hpp:66-67)
    for (int l = 0; l < levels; l++) {
        const Cam c = make_cam(P, 1 << l, true);
        pa.lv[l].w = c.w;
        pa.lv[l].h = c.h;
        pa.lv[l].cx = c.cx;
        pa.lv[l].cy = c.cy;
        pa.lv[l].inv_fx = c.inv_fx;
        pa.lv[l].inv_fy = c.inv_fy;
        pa.lv[l].vert = (float4*) tr->vert_lvl[l].ptr;
        pa.lv[l].norm = (float4*) tr->norm_lvl[l].ptr;
    }
    pa.levels = levels;
    pa.neg_y = make_cam(P, 1, false).left_hand_frame;
    pa.tile_log2 = pyramid_tile_log2(levels);
    const int tile = 1 << pa.tile_log2;
    pa.tiles_x = (P.width + tile - 1) / tile;
    const int tiles_y = (P.height + tile - 1) / tile;
    pa.init = init;
    pa.bins = bins;
    pa.T = T;
    pa.T16_dev = T16_dev;
    *n_tiles = pa.tiles_x * tiles_y;
    return pa;
}

static void launch_pyramid(se8_tracker* tr, const float* depth, int levels, IcpShared* init, const Iso& T,
                           const float* T16_dev, hipStream_t s, long long* bins = nullptr)
{
    int n_tiles = 0;
    const PyrArgs pa = pyramid_args(tr, depth, levels, init, T, T16_dev, bins, &n_tiles);
    tr->prepared_depth = tr->prepared_T = nullptr; // the prepared pyramid's buffers are overwritten
    k_pyramid<<<n_tiles, kPyrThreads, pyramid_lds_bytes(levels), s>>>(pa);
    graph_note_launch(kNodePyramid, s);
}

// The arguments of the pyramid that starts a track call from depth_dev and the pose at T_in_dev,
// for a launch that runs it ahead of the call; the call that follows with the same two pointers
// then skips its own pyramid.
int tracker_prepare_pyramid(se8_tracker* tr, const float* depth_dev, const float* T_in_dev, PyrArgs* out,
                            int* n_tiles)
{
    *out = pyramid_args(tr, depth_dev, tr->cfg.num_levels, (IcpShared*) tr->state.ptr, Iso{}, T_in_dev,
                        (long long*) tr->bins.ptr, n_tiles);
    tr->prepared_depth = depth_dev;
    tr->prepared_T = T_in_dev;
    return SE8_OK;
}

template<bool Comp>
static void launch_icp_iter(int ppt, int grid, const IcpArgs& a, hipStream_t s)
{
    switch (ppt) {
    case 1:
        k_icp_iter<Comp, 1><<<grid, kIcpThreads, 0, s>>>(a);
        break;
    case 2:
        k_icp_iter<Comp, 2><<<grid, kIcpThreads, 0, s>>>(a);
        break;
    default:
        k_icp_iter<Comp, 4><<<grid, kIcpThreads, 0, s>>>(a);
        break;
    }
}

// Enqueues Tracker::track (tracker_impl.hpp:17-116): the pyramid, every Gauss-Newton launch and the
// final check.  The start pose is T16_host, or 16 floats at T_in_dev; the result lands in
// IcpShared::result, and also at T_out_dev / ok_dev when they are set.  Nothing waits on the host.
static int enqueue_track(se8_tracker* tr, const float* depth, const float* T16_host, const float* T_in_dev,
                         float* T_out_dev, int* ok_dev, const float* ref_pts, const float* ref_nrm, hipStream_t s,
                         se8_map* begin_map = nullptr)
{
    const se8_pinhole& P = tr->sensor;
    IcpShared* sh = (IcpShared*) tr->state.ptr;
    // the previous call's stage events, before ev[0] / ev[1] are recorded again (with timing off,
    // nothing is recorded and nothing waits)
    const bool timed = tr->timing != SE8_TIMING_OFF, levels = tr->timing == SE8_TIMING_LEVELS;
    harvest_tracking(tr);
    if (timed)
        SE8_HIP_CHECK(hipEventRecord(tr->ev[0], s));
    long long* bins = (long long*) tr->bins.ptr;
    // a pyramid already run for this depth and start pose (se8_raycast_dev_prepare_track) is not
    // run again
    const bool prepared = !T16_host && depth == tr->prepared_depth && T_in_dev == tr->prepared_T;
    tr->prepared_depth = tr->prepared_T = nullptr;
    if (!prepared)
        launch_pyramid(tr, depth, tr->cfg.num_levels, sh, T16_host ? iso_from16(T16_host) : Iso{}, T_in_dev, s, bins);
    IcpArgs a{};
    a.bins = bins;
    a.ref_pts = ref_pts;
    a.ref_nrm = ref_nrm;
    a.sh = sh;
    a.cam = make_cam(P, 1, false);
    a.dist_th = tr->cfg.dist_threshold;
    a.normal_th = tr->cfg.normal_threshold;
    a.icp_th = tr->cfg.icp_threshold;
    a.track_th = tr->cfg.track_threshold;
    a.k = 0;
    for (int level = tr->cfg.num_levels - 1; level >= 0; level--) {
        a.vert = (const float4*) tr->vert_lvl[level].ptr;
        a.norm = (const float4*) tr->norm_lvl[level].ptr;
        // reduction resolution is the full resolution / 2^level (tracker_impl.hpp:85-86)
        a.w = P.width / (1 << level);
        a.h = P.height / (1 << level);
        a.level = level;
        a.n_tiles = tr->icp_tiles[level];
        const int nwg = tr->num_wg[level];
        if (levels)
            SE8_HIP_CHECK(hipEventRecord(tr->ev_lvl[level], s));
        for (int it = 0; it < tr->cfg.iterations[level]; it++, a.k++) {
            if (tr->cfg.compensated)
                launch_icp_iter<true>(tr->icp_ppt[level], nwg, a, s);
            else
                launch_icp_iter<false>(tr->icp_ppt[level], nwg, a, s);
        }
    }
    if (levels)
        SE8_HIP_CHECK(hipEventRecord(tr->ev_lvl[kMaxLevels], s));
    a.T_out_dev = T_out_dev;
    a.ok_dev = ok_dev;
    if (begin_map) {
        // the finish kernel also starts begin_map's next integration at the tracked pose
        a.fb_counters = begin_map->view.counters;
        a.fb_fp = (FramePoses*) begin_map->frame_poses.ptr;
        for (int i = 0; i < 3; i++)
            a.fb_t_MW[i] = begin_map->view.t_MW[i];
        a.fb_res = begin_map->view.res;
    }
    if (tr->cfg.compensated)
        k_icp_finish<true><<<1, kIcpThreads, 0, s>>>(a);
    else
        k_icp_finish<false><<<1, kIcpThreads, 0, s>>>(a);
    SE8_HIP_CHECK(hipGetLastError());
    graph_note_launch(kNodeIcpFinish, s);
    if (timed) {
        SE8_HIP_CHECK(hipEventRecord(tr->ev[1], s));
        tr->track_pending = true;
        tr->lvl_pending = levels;
    }
    return SE8_OK;
}

// Tracker::track with a host pose: enqueue, then wait for the result and write it back.
static int run_track(se8_tracker* tr, const float* depth, float T16[16], const float* ref_pts, const float* ref_nrm,
                     int* ok, float* sums32, hipStream_t s)
{
    int rc = enqueue_track(tr, depth, T16, nullptr, nullptr, nullptr, ref_pts, ref_nrm, s);
    if (rc != SE8_OK)
        return rc;
    IcpShared* sh = (IcpShared*) tr->state.ptr;
    IcpState& host = *(IcpState*) tr->host_state;
    SE8_HIP_CHECK(hipMemcpyAsync(&host, &sh->result, sizeof(IcpState), hipMemcpyDeviceToHost, s));
    SE8_HIP_CHECK(hipStreamSynchronize(s));
    harvest_tracking(tr);
    iso_to16(iso_from16(host.T), T16);
    if (ok)
        *ok = host.ok;
    if (sums32)
        std::memcpy(sums32, host.sums, sizeof(float) * 32);
    return SE8_OK;
}

int launch_raycast(const se8_map* m, const Cam& cam, const Iso& T_WS, const float* T16_dev, float* points,
                   float* normals, int8_t* scale, hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_stop);

// ----------------------------------------------------------------------------------------------
// Batched point-to-plane cost and its adjoint (differentiation growth pattern).
// For pose k: cost_k = 0.5 sum_i e_i^2 with e_i the trackKernel residual against fixed reference
// maps raycast at T_ref; grad_k = d cost / d xi (left twist T <- exp(xi) T, xi = [v; w]) at xi = 0,
// which is -sum_i e_i J_i (de/dxi = -J); J^T J (21 terms) is returned for Gauss-Newton steps.
// Grid (pixel workgroups, poses); per-workgroup partials, then one workgroup per pose reduces in
// FP64 in a fixed order.
// ----------------------------------------------------------------------------------------------
constexpr int kAdjTerms = 29; // e^2, e J (6), J^T J (21), inliers

// poses: pose k's rows R|t start at poses[pose_stride * k] (12: packed 3x4, 16: row-major 4x4)
__global__ __launch_bounds__(kIcpThreads) void k_adj_icp(IcpArgs a, const float* poses, int pose_stride, int n_wg,
                                                        float* partials)
{
    const int k = blockIdx.y;
    Iso T;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T.R[i][j] = poses[(size_t) pose_stride * k + 4 * i + j];
        T.t[i] = poses[(size_t) pose_stride * k + 4 * i + 3];
    }
    float s[kAdjTerms];
    for (int q = 0; q < kAdjTerms; q++)
        s[q] = 0.f;
    const int n = a.w * a.h;
    for (int p = blockIdx.x * kIcpThreads + threadIdx.x; p < n; p += n_wg * kIcpThreads) {
        float e = 0.f, J[6];
        if (icp_pixel(a, T, p, e, J) == 1) {
            s[0] += e * e;
            for (int q = 0; q < 6; q++)
                s[1 + q] += e * J[q];
            int c = 7;
            for (int i = 0; i < 6; i++)
                for (int j = i; j < 6; j++)
                    s[c++] += J[i] * J[j];
            s[28] += 1.f;
        }
    }
    __shared__ float red[kIcpThreads / 64][kAdjTerms];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < kAdjTerms; q++)
            s[q] += __shfl_xor(s[q], off);
    if (lane == 0)
        for (int q = 0; q < kAdjTerms; q++)
            red[wave][q] = s[q];
    __syncthreads();
    if (threadIdx.x < kAdjTerms) {
        float v = 0.f;
        for (int w = 0; w < kIcpThreads / 64; w++)
            v += red[w][threadIdx.x];
        partials[((size_t) k * n_wg + blockIdx.x) * kAdjTerms + threadIdx.x] = v;
    }
}

// Per pose, the FP64 sum of the workgroup partials in workgroup order.  The sums go to out (29
// doubles per pose) when it is set; otherwise the results are written in their final form, as the
// host converts them: cost = 0.5 e^2 sum, grad = -(e J sum), J^T J, inlier count.
__global__ void k_adj_icp_reduce(const float* partials, int n_wg, double* out, float* cost, float* grad6, float* jtj21,
                                 int* inliers)
{
    const int k = blockIdx.x;
    const int q = threadIdx.x;
    if (q >= kAdjTerms)
        return;
    double acc = 0.0;
    for (int b = 0; b < n_wg; b++)
        acc += (double) partials[((size_t) k * n_wg + b) * kAdjTerms + q];
    if (out) {
        out[(size_t) k * kAdjTerms + q] = acc;
        return;
    }
    if (q == 0)
        cost[k] = (float) (0.5 * acc);
    else if (q < 7)
        grad6[6 * k + q - 1] = (float) (-acc);
    else if (q < 28) {
        if (jtj21)
            jtj21[21 * k + q - 7] = (float) acc;
    }
    else if (inliers)
        inliers[k] = (int) acc;
}

// Scratch of se8_icp_cost_grad_batched_noalloc: the per-workgroup partials of at most
// kNoAllocMaxPoses poses (kIcpMaxWg workgroups each), 2.4 MB held by every device the library
// runs on until the process ends.  Host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (noalloc_scratch_ptr).
constexpr int kNoAllocMaxPoses = 64;
static __device__ __align__(256) float scratch[(size_t) kNoAllocMaxPoses * kIcpMaxWg * kAdjTerms];

// Device address of `scratch` on the current device, looked up once per device.
constexpr int kNoAllocMaxDevices = 64;
static int noalloc_scratch_ptr(float** out)
{
    static std::mutex mu;
    static float* ptrs[kNoAllocMaxDevices] = {};
    int dev = 0;
    SE8_HIP_CHECK(hipGetDevice(&dev));
    SE8_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, "se8_icp_cost_grad_batched_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[dev]) {
        void* p = nullptr;
        SE8_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptrs[dev] = (float*) p;
    }
    *out = ptrs[dev];
    return SE8_OK;
}

// Frame graphs (se8_graph.hip): size of the argument struct of the pyramid launch (kNodePyramid)
// or the final ICP launch (kNodeIcpFinish), and the addresses of its per-frame fields inside a copy
// of it: the input depth and the start pose (d0, T16_dev), or the result's destinations (T_out_dev,
// ok_dev).
size_t graph_track_args_bytes(int kind)
{
    return kind == kNodePyramid ? sizeof(PyrArgs) : sizeof(IcpArgs);
}

void graph_track_args_fields(int kind, void* args, void** f0, void** f1)
{
    if (kind == kNodePyramid) {
        PyrArgs* a = (PyrArgs*) args;
        *f0 = &a->d0;
        *f1 = &a->T16_dev;
    }
    else {
        IcpArgs* a = (IcpArgs*) args;
        *f0 = &a->T_out_dev;
        *f1 = &a->ok_dev;
    }
}

} // namespace se8

using namespace se8;

extern "C" int se8_tracker_create(const se8_pinhole* sensor, const se8_tracker_config* cfg, se8_tracker** out)
{
    SE8_REQUIRE(sensor && cfg && out, "se8_tracker_create: null argument");
    SE8_REQUIRE(sensor->width >= 1 && sensor->height >= 1, "se8_tracker_create: empty image");
    SE8_REQUIRE(cfg->num_levels >= 1 && cfg->num_levels <= kMaxLevels,
                "se8_tracker_create: num_levels must be in [1, SE8_TRACKER_MAX_LEVELS]");
    SE8_REQUIRE((sensor->width >> (cfg->num_levels - 1)) >= 1 && (sensor->height >> (cfg->num_levels - 1)) >= 1,
                "se8_tracker_create: the coarsest pyramid level has no pixels");
    for (int l = 0; l < cfg->num_levels; l++)
        SE8_REQUIRE(cfg->iterations[l] >= 0 && cfg->iterations[l] < 1000, "se8_tracker_create: bad iterations");
    *out = nullptr;
    se8_tracker* t = new se8_tracker();
    t->sensor = *sensor;
    t->cfg = *cfg;
    for (int l = cfg->num_levels; l < kMaxLevels; l++)
        t->cfg.iterations[l] = 0;
    int rc = SE8_OK;
    for (int l = 0; l < cfg->num_levels; l++) {
        const Cam cl = make_cam(*sensor, 1 << l, true);
        const size_t n = (size_t) cl.w * cl.h;
        rc = rc ? rc : dev_alloc(t->vert_lvl[l], sizeof(float4) * n);
        rc = rc ? rc : dev_alloc(t->norm_lvl[l], sizeof(float4) * n);
        // smallest pixels-per-thread whose tiles fit one grid of at most kIcpTrackMaxWg workgroups:
        // more, lighter workgroups keep each wave's row arithmetic short (320x240 and 160x120 are
        // faster in 300 and 75 workgroups of 256 pixels than in 75 and 19 of 1024)
        int ppt = 1;
        while (ppt < 4 && (n + kIcpThreads * ppt - 1) / (kIcpThreads * ppt) > (size_t) kIcpTrackMaxWg)
            ppt *= 2;
        t->icp_ppt[l] = ppt;
        t->icp_tiles[l] = (int) std::max<size_t>(1, (n + kIcpThreads * ppt - 1) / (kIcpThreads * ppt));
        t->num_wg[l] = std::min(t->icp_tiles[l], kIcpTrackMaxWg);
    }
    {
        // the batched adjoint's grid over level 0: at most kIcpMaxWg workgroups (its scratch)
        const size_t n = (size_t) t->sensor.width * t->sensor.height;
        int ppt = 1;
        while (ppt < 4 && (n + kIcpThreads * ppt - 1) / (kIcpThreads * ppt) > (size_t) kIcpMaxWg)
            ppt *= 2;
        t->adj_wg = (int) std::min<size_t>(std::max<size_t>(1, (n + kIcpThreads * ppt - 1) / (kIcpThreads * ppt)),
                                           (size_t) kIcpMaxWg);
    }
    rc = rc ? rc : dev_alloc(t->bins, sizeof(long long) * 3 * kBins * 32);
    rc = rc ? rc : dev_alloc(t->state, sizeof(IcpShared));
    if (rc == SE8_OK && hipHostMalloc(&t->host_state, sizeof(IcpState), hipHostMallocDefault) != hipSuccess) {
        t->host_state = nullptr;
        rc = SE8_ERR_OOM;
    }
    for (int i = 0; i < 2 && rc == SE8_OK; i++)
        if (hipEventCreate(&t->ev[i]) != hipSuccess)
            rc = SE8_ERR_HIP;
    if (rc != SE8_OK) {
        se8_tracker_destroy(t);
        return rc;
    }
    *out = t;
    return SE8_OK;
}

extern "C" int se8_tracker_destroy(se8_tracker* t)
{
    if (!t)
        return SE8_OK;
    (void) hipDeviceSynchronize();
    for (int l = 0; l < kMaxLevels; l++) {
        dev_free(t->vert_lvl[l]);
        dev_free(t->norm_lvl[l]);
    }
    dev_free(t->bins);
    dev_free(t->state);
    if (t->host_state)
        (void) hipHostFree(t->host_state);
    dev_free(t->ref_points);
    dev_free(t->ref_normals);
    dev_free(t->adj_poses);
    dev_free(t->adj_partials);
    dev_free(t->adj_out);
    for (int i = 0; i < 2; i++)
        if (t->ev[i])
            (void) hipEventDestroy(t->ev[i]);
    for (int i = 0; i <= kMaxLevels; i++)
        if (t->ev_lvl[i])
            (void) hipEventDestroy(t->ev_lvl[i]);
    delete t;
    return SE8_OK;
}

extern "C" int se8_track(se8_tracker* t, const float* depth_dev, float T_WS[16], const float* ref_points_dev,
                         const float* ref_normals_dev, int* ok, float* sums32, se8_stream stream)
{
    SE8_REQUIRE(t && depth_dev && T_WS && ref_points_dev && ref_normals_dev, "se8_track: null argument");
    return run_track(t, depth_dev, T_WS, ref_points_dev, ref_normals_dev, ok, sums32, (hipStream_t) stream);
}

extern "C" int se8_track_dev(se8_tracker* t, const float* depth_dev, const float* T_in_dev, float* T_out_dev,
                             const float* ref_points_dev, const float* ref_normals_dev, int* ok_dev, se8_stream stream)
{
    SE8_REQUIRE(t && depth_dev && T_in_dev && T_out_dev && ref_points_dev && ref_normals_dev,
                "se8_track_dev: null argument");
    return enqueue_track(t, depth_dev, nullptr, T_in_dev, T_out_dev, ok_dev, ref_points_dev, ref_normals_dev,
                         (hipStream_t) stream);
}

extern "C" int se8_track_dev_prepare_integrate(se8_tracker* t, const float* depth_dev, const float* T_in_dev,
                                               float* T_out_dev, const float* ref_points_dev,
                                               const float* ref_normals_dev, int* ok_dev, se8_map* map,
                                               se8_stream stream)
{
    SE8_REQUIRE(t && depth_dev && T_in_dev && T_out_dev && ref_points_dev && ref_normals_dev && map,
                "se8_track_dev_prepare_integrate: null argument");
    const int rc = enqueue_track(t, depth_dev, nullptr, T_in_dev, T_out_dev, ok_dev, ref_points_dev,
                                 ref_normals_dev, (hipStream_t) stream, map);
    if (rc == SE8_OK)
        map->prepared_T = T_out_dev;
    return rc;
}

extern "C" int se8_tracker_read_pyramid(const se8_tracker* t, int level, float* points, float* normals)
{
    SE8_REQUIRE(t, "se8_tracker_read_pyramid: null tracker");
    SE8_REQUIRE(level >= 0 && level < t->cfg.num_levels, "se8_tracker_read_pyramid: level out of range");
    const Cam c = make_cam(t->sensor, 1 << level, true);
    const size_t n = (size_t) c.w * c.h;
    std::vector<float4> buf(n);
    const DeviceBuffer* src[2] = {&t->vert_lvl[level], &t->norm_lvl[level]};
    float* dst[2] = {points, normals};
    for (int k = 0; k < 2; k++) {
        if (!dst[k])
            continue;
        SE8_HIP_CHECK(hipDeviceSynchronize());
        SE8_HIP_CHECK(hipMemcpy(buf.data(), src[k]->ptr, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) {
            dst[k][3 * i] = buf[i].x;
            dst[k][3 * i + 1] = buf[i].y;
            dst[k][3 * i + 2] = buf[i].z;
        }
    }
    return SE8_OK;
}

extern "C" int se8_track_raycast(se8_tracker* t, const se8_map* m, const float* depth_dev, float T_WS[16], int* ok,
                                 se8_stream stream)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_track_raycast: not available for a multi-resolution map");
    SE8_REQUIRE(t && m && depth_dev && T_WS, "se8_track_raycast: null argument");
    const size_t n = (size_t) t->sensor.width * t->sensor.height;
    int rc = dev_alloc(t->ref_points, sizeof(float) * 3 * n);
    rc = rc ? rc : dev_alloc(t->ref_normals, sizeof(float) * 3 * n);
    if (rc != SE8_OK)
        return rc;
    hipStream_t s = (hipStream_t) stream;
    const Cam cam = make_cam(t->sensor, 1, false);
    rc = launch_raycast(m, cam, iso_from16(T_WS), nullptr, (float*) t->ref_points.ptr, (float*) t->ref_normals.ptr,
                        nullptr, s, nullptr, nullptr);
    if (rc != SE8_OK)
        return rc;
    return run_track(t, depth_dev, T_WS, (const float*) t->ref_points.ptr, (const float*) t->ref_normals.ptr, ok,
                     nullptr, s);
}

extern "C" int se8_icp_cost_grad_batched(se8_tracker* t, const float* depth_dev, const float* ref_points_dev,
                                         const float* ref_normals_dev, const float T_ref16[16], int n_poses,
                                         const float* poses16, float* cost, float* grad6, float* jtj21,
                                         int* inliers, se8_stream stream)
{
    SE8_REQUIRE(t && depth_dev && ref_points_dev && ref_normals_dev && T_ref16 && poses16 && cost && grad6,
                "se8_icp_cost_grad_batched: null argument");
    SE8_REQUIRE(n_poses > 0 && n_poses <= 4096, "se8_icp_cost_grad_batched: n_poses out of range");
    hipStream_t s = (hipStream_t) stream;
    const se8_pinhole& P = t->sensor;
    const int n_wg = t->adj_wg;
    int rc = dev_alloc(t->adj_poses, sizeof(float) * 12 * n_poses);
    rc = rc ? rc : dev_alloc(t->adj_partials, sizeof(float) * kAdjTerms * n_wg * n_poses);
    rc = rc ? rc : dev_alloc(t->adj_out, sizeof(double) * kAdjTerms * n_poses);
    if (rc != SE8_OK)
        return rc;
    std::vector<float> p12((size_t) 12 * n_poses);
    for (int k = 0; k < n_poses; k++)
        for (int i = 0; i < 12; i++)
            p12[(size_t) 12 * k + i] = poses16[(size_t) 16 * k + i];
    SE8_HIP_CHECK(hipMemcpyAsync(t->adj_poses.ptr, p12.data(), sizeof(float) * p12.size(), hipMemcpyHostToDevice, s));
    launch_pyramid(t, depth_dev, 1, nullptr, Iso{}, nullptr, s);
    IcpArgs a{};
    a.vert = (const float4*) t->vert_lvl[0].ptr;
    a.norm = (const float4*) t->norm_lvl[0].ptr;
    a.ref_pts = ref_points_dev;
    a.ref_nrm = ref_normals_dev;
    a.sh = nullptr;
    a.T_ref_inv = iso_inv(iso_from16(T_ref16));
    a.cam = make_cam(P, 1, false);
    a.w = P.width;
    a.h = P.height;
    a.dist_th = t->cfg.dist_threshold;
    a.normal_th = t->cfg.normal_threshold;
    a.icp_th = t->cfg.icp_threshold;
    a.level = 0;
    k_adj_icp<<<dim3(n_wg, n_poses), kIcpThreads, 0, s>>>(a, (const float*) t->adj_poses.ptr, 12, n_wg,
                                                          (float*) t->adj_partials.ptr);
    k_adj_icp_reduce<<<n_poses, 64, 0, s>>>((const float*) t->adj_partials.ptr, n_wg, (double*) t->adj_out.ptr,
                                            nullptr, nullptr, nullptr, nullptr);
    SE8_HIP_CHECK(hipGetLastError());
    std::vector<double> out((size_t) kAdjTerms * n_poses);
    SE8_HIP_CHECK(hipMemcpyAsync(out.data(), t->adj_out.ptr, sizeof(double) * out.size(), hipMemcpyDeviceToHost, s));
    SE8_HIP_CHECK(hipStreamSynchronize(s));
    for (int k = 0; k < n_poses; k++) {
        const double* o = out.data() + (size_t) kAdjTerms * k;
        cost[k] = (float) (0.5 * o[0]);
        for (int q = 0; q < 6; q++)
            grad6[6 * k + q] = (float) (-o[1 + q]);
        if (jtj21)
            for (int q = 0; q < 21; q++)
                jtj21[21 * k + q] = (float) o[7 + q];
        if (inliers)
            inliers[k] = (int) o[28];
    }
    return SE8_OK;
}

extern "C" int se8_icp_cost_grad_batched_noalloc(se8_tracker* t, const float* depth_dev, const float* ref_points_dev,
                                                 const float* ref_normals_dev, const float T_ref16[16], int n_poses,
                                                 const float* poses16_dev, float* cost_dev, float* grad6_dev,
                                                 float* jtj21_dev, int* inliers_dev, se8_stream stream)
{
    SE8_REQUIRE(t && depth_dev && ref_points_dev && ref_normals_dev && T_ref16 && poses16_dev && cost_dev && grad6_dev,
                "se8_icp_cost_grad_batched_noalloc: null argument");
    SE8_REQUIRE(n_poses > 0 && n_poses <= kNoAllocMaxPoses,
                "se8_icp_cost_grad_batched_noalloc: n_poses must be in [1, 64] (the size of the library's scratch)");
    const int n_wg = t->adj_wg;
    SE8_REQUIRE(n_wg > 0 && n_wg <= kIcpMaxWg, "se8_icp_cost_grad_batched_noalloc: tracker grid out of range");
    float* partials = nullptr;
    int rc = noalloc_scratch_ptr(&partials);
    if (rc != SE8_OK)
        return rc;
    hipStream_t s = (hipStream_t) stream;
    const se8_pinhole& P = t->sensor;
    // the level-0 vertex and normal maps go to the tracker's own pyramid buffers (allocated with
    // the tracker)
    launch_pyramid(t, depth_dev, 1, nullptr, Iso{}, nullptr, s);
    IcpArgs a{};
    a.vert = (const float4*) t->vert_lvl[0].ptr;
    a.norm = (const float4*) t->norm_lvl[0].ptr;
    a.ref_pts = ref_points_dev;
    a.ref_nrm = ref_normals_dev;
    a.T_ref_inv = iso_inv(iso_from16(T_ref16));
    a.cam = make_cam(P, 1, false);
    a.w = P.width;
    a.h = P.height;
    a.dist_th = t->cfg.dist_threshold;
    a.normal_th = t->cfg.normal_threshold;
    a.icp_th = t->cfg.icp_threshold;
    a.level = 0;
    k_adj_icp<<<dim3(n_wg, n_poses), kIcpThreads, 0, s>>>(a, poses16_dev, 16, n_wg, partials);
    k_adj_icp_reduce<<<n_poses, 64, 0, s>>>(partials, n_wg, nullptr, cost_dev, grad6_dev, jtj21_dev, inliers_dev);
    SE8_HIP_CHECK(hipGetLastError());
    return SE8_OK;
}
