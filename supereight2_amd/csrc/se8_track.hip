// Projective point-to-plane ICP of libse8.so: se::Tracker::track
// (include/se/tracker/impl/tracker_impl.hpp:17-116).
//
// Per call: two robust half-sampling launches (preprocessor.cpp:156-195), one vertex+normal
// launch per pyramid level (preprocessor_impl.hpp:18-40, preprocessor.cpp:109-152), then one
// launch per Gauss-Newton iteration that fuses trackKernel (icp_impl.hpp:16-102) with the
// 32-term reduction of reduceKernel (icp.cpp:45-204): per-thread accumulation, wave-shuffle and
// LDS reduction to per-workgroup partials, and the last workgroup to finish reduces the
// partials in a fixed order, solves the 6x6 system, applies T <- exp(x) T and raises the
// level's convergence flag (icp.cpp:15-41, 208-223).  Later iterations of a converged level
// return at once, so the host never waits inside the pyramid loop.  A final launch applies
// checkPoseKernel (icp.cpp:227-246).
//
// compensated = 0: fp32 sums, fixed-order fp32 cross-workgroup sum, fp32 Cholesky (reference).
// compensated = 1: Kahan-compensated fp32 per-thread sums, error-free (TwoSum) pair reduction in
// the workgroup, FP64 cross-workgroup sum and FP64 LDL^T solve.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "se8_internal.hpp"

namespace se8 {

constexpr int kIcpThreads = 256;
constexpr int kIcpPxPerThread = 4;

struct IcpState {
    float T[12];       // current pose, row-major 3x4
    float sums[32];    // last reduction
    double sums_d[32]; // last reduction (compensated path)
    int done[4];       // per-level convergence flags
    unsigned ticket;   // last-workgroup counter
    int ok;
    int iterations;    // Gauss-Newton iterations executed
    int pad;
};

struct IcpArgs {
    const float* vert;     // level vertex map (w*h*3)
    const float* norm;     // level normal map (w*h*3)
    const float* ref_pts;  // full-res raycast points
    const float* ref_nrm;  // full-res raycast normals
    float* partials;       // nwg * 64
    IcpState* st;
    Iso T_ref_inv;
    Cam cam;               // full-resolution sensor (projection into the reference maps)
    int w, h;              // level size
    float dist_th, normal_th, icp_th;
    int level;
};

__global__ void k_half_sample(const float* in, int w, int h, float* out, float e_d)
{
    const int ow = w / 2, oh = h / 2;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= ow || y >= oh)
        return;
    const int ix = 2 * x, iy = 2 * y;
    const float c = in[ix + (size_t) iy * w];
    float count = 0.f, sum = 0.f;
#pragma unroll
    for (int i = 0; i <= 1; i++) {
#pragma unroll
        for (int j = 0; j <= 1; j++) {
            const int px = min(ix + j, 2 * ow - 1), py = min(iy + i, 2 * oh - 1);
            const float v = in[px + (size_t) py * w];
            if (fabsf(v - c) < e_d) {
                count += 1.f;
                sum += v;
            }
        }
    }
    out[x + (size_t) y * ow] = sum / count;
}

__device__ __forceinline__ void depth_point(const float* depth, const Cam& c, int x, int y, float p[3])
{
    const float d = depth[x + (size_t) y * c.w];
    if (d > 0.f) {
        float dir[3];
        cam_backproject(c, (float) x, (float) y, dir);
        p[0] = d * dir[0];
        p[1] = d * dir[1];
        p[2] = d * dir[2];
    }
    else {
        p[0] = p[1] = p[2] = 0.f;
    }
}

// depth_to_point_cloud + point_cloud_to_normal<NegY> fused: neighbour points are recomputed from
// depth, which gives the same values as reading them back.
__global__ void k_vertex_normal(const float* depth, Cam c, int neg_y, float* vert, float* norm)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= c.w || y >= c.h)
        return;
    const size_t i = x + (size_t) y * c.w;
    float p[3];
    depth_point(depth, c, x, y, p);
    vert[3 * i] = p[0];
    vert[3 * i + 1] = p[1];
    vert[3 * i + 2] = p[2];
    float n[3] = {0.f, 0.f, 0.f};
    if (p[2] != 0.f) {
        const int xl = max(x - 1, 0), xr = min(x + 1, c.w - 1);
        const int ym = max(y - 1, 0), yp = min(y + 1, c.h - 1);
        const int yu = neg_y ? ym : yp, yd = neg_y ? yp : ym;
        float l[3], r[3], u[3], d[3];
        depth_point(depth, c, xl, y, l);
        depth_point(depth, c, xr, y, r);
        depth_point(depth, c, x, yu, u);
        depth_point(depth, c, x, yd, d);
        if (l[2] != 0.f && r[2] != 0.f && u[2] != 0.f && d[2] != 0.f) {
            float dx[3], dy[3], cr[3];
            for (int k = 0; k < 3; k++) {
                dx[k] = r[k] - l[k];
                dy[k] = u[k] - d[k];
            }
            cross3(dx, dy, cr);
            normalized3(cr, n);
        }
    }
    norm[3 * i] = n[0];
    norm[3 * i + 1] = n[1];
    norm[3 * i + 2] = n[2];
}

__global__ void k_icp_init(IcpState* st, Iso T)
{
    if (threadIdx.x == 0) {
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++)
                st->T[4 * i + j] = T.R[i][j];
            st->T[4 * i + 3] = T.t[i];
        }
        for (int k = 0; k < 32; k++) {
            st->sums[k] = 0.f;
            st->sums_d[k] = 0.0;
        }
        for (int l = 0; l < 4; l++)
            st->done[l] = 0;
        st->ticket = 0;
        st->ok = 0;
        st->iterations = 0;
    }
}

__device__ __forceinline__ Iso load_pose(const IcpState* st)
{
    Iso T;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T.R[i][j] = st->T[4 * i + j];
        T.t[i] = st->T[4 * i + 3];
    }
    return T;
}

// One pixel of trackKernel; returns the result code and fills e, J on success.
__device__ __forceinline__ int icp_pixel(const IcpArgs& a, const Iso& T, int p, float& e, float J[6])
{
    const float n_in[3] = {a.norm[3 * p], a.norm[3 * p + 1], a.norm[3 * p + 2]};
    if (n_in[0] == 0.f && n_in[1] == 0.f && n_in[2] == 0.f)
        return -1;
    const float v[3] = {a.vert[3 * p], a.vert[3 * p + 1], a.vert[3 * p + 2]};
    float pW[3], pR[3], uv[2];
    iso_apply(T, v, pW);
    iso_apply(a.T_ref_inv, pW, pR);
    if (!cam_project(a.cam, pR, uv))
        return -2;
    const int px = (int) (uv[0] + 0.5f), py = (int) (uv[1] + 0.5f);
    const size_t ri = px + (size_t) py * a.cam.w;
    const float rn[3] = {-a.ref_nrm[3 * ri], -a.ref_nrm[3 * ri + 1], -a.ref_nrm[3 * ri + 2]};
    if (rn[0] == 0.f && rn[1] == 0.f && rn[2] == 0.f)
        return -3;
    float diff[3], nW[3];
    for (int k = 0; k < 3; k++)
        diff[k] = a.ref_pts[3 * ri + k] - pW[k];
    lin_apply(T, n_in, nW);
    if (sqrtf(sqnorm3(diff)) > a.dist_th)
        return -4;
    if (dot3(nW, rn) < a.normal_th)
        return -5;
    e = dot3(rn, diff);
    J[0] = rn[0];
    J[1] = rn[1];
    J[2] = rn[2];
    float cp[3];
    cross3(pW, rn, cp);
    J[3] = cp[0];
    J[4] = cp[1];
    J[5] = cp[2];
    return 1;
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

// 6x6 Cholesky in fp32 (icp.cpp:33-41): x = 0 if the factorisation fails
__device__ void solve_f32(const float* vals, float x[6])
{
    float A[6][6], L[6][6];
    int c = 6;
    for (int i = 0; i < 6; i++)
        for (int j = i; j < 6; j++) {
            A[i][j] = vals[c++];
            A[j][i] = A[i][j];
        }
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++)
            L[i][j] = 0.f;
    for (int k = 0; k < 6; k++) {
        float s = A[k][k];
        for (int j = 0; j < k; j++)
            s -= L[k][j] * L[k][j];
        if (!(s > 0.f)) {
            for (int i = 0; i < 6; i++)
                x[i] = 0.f;
            return;
        }
        L[k][k] = sqrtf(s);
        for (int i = k + 1; i < 6; i++) {
            float v = A[i][k];
            for (int j = 0; j < k; j++)
                v -= L[i][j] * L[k][j];
            L[i][k] = v / L[k][k];
        }
    }
    float y[6];
    for (int i = 0; i < 6; i++) {
        float v = vals[i];
        for (int j = 0; j < i; j++)
            v -= L[i][j] * y[j];
        y[i] = v / L[i][i];
    }
    for (int i = 5; i >= 0; i--) {
        float v = y[i];
        for (int j = i + 1; j < 6; j++)
            v -= L[j][i] * x[j];
        x[i] = v / L[i][i];
    }
}

// 6x6 LDL^T in FP64 (no pivoting: J^T J is symmetric positive semi-definite); x = 0 on failure
__device__ void solve_f64(const double* vals, float x[6])
{
    double A[6][6], L[6][6], D[6];
    int c = 6;
    for (int i = 0; i < 6; i++)
        for (int j = i; j < 6; j++) {
            A[i][j] = vals[c++];
            A[j][i] = A[i][j];
        }
    for (int j = 0; j < 6; j++) {
        double dj = A[j][j];
        for (int k = 0; k < j; k++)
            dj -= L[j][k] * L[j][k] * D[k];
        if (!(dj > 0.0)) {
            for (int i = 0; i < 6; i++)
                x[i] = 0.f;
            return;
        }
        D[j] = dj;
        L[j][j] = 1.0;
        for (int i = j + 1; i < 6; i++) {
            double v = A[i][j];
            for (int k = 0; k < j; k++)
                v -= L[i][k] * L[j][k] * D[k];
            L[i][j] = v / dj;
        }
    }
    double y[6], z[6];
    for (int i = 0; i < 6; i++) {
        double v = vals[i];
        for (int k = 0; k < i; k++)
            v -= L[i][k] * y[k];
        y[i] = v;
    }
    for (int i = 0; i < 6; i++)
        y[i] /= D[i];
    for (int i = 5; i >= 0; i--) {
        double v = y[i];
        for (int k = i + 1; k < 6; k++)
            v -= L[k][i] * z[k];
        z[i] = v;
    }
    for (int i = 0; i < 6; i++)
        x[i] = (float) z[i];
}

// math::exp (math_util_impl.hpp:117-174) for the twist a = [v; w]
__device__ Iso se3_exp(const float a[6])
{
    const float w[3] = {a[3], a[4], a[5]};
    const float theta_sq = dot3(w, w);
    const float theta = sqrtf(theta_sq);
    const float half_theta = 0.5f * theta;
    float imag, re;
    if (theta < 1e-10f) {
        const float t4 = theta_sq * theta_sq;
        imag = 0.5f - (1.0f / 48.0f) * theta_sq + (1.0f / 3840.0f) * t4;
        re = 1.f - 0.5f * theta_sq + (1.f / 384.f) * t4;
    }
    else {
        imag = sinf(half_theta) / theta;
        re = cosf(half_theta);
    }
    float q[4] = {re, imag * w[0], imag * w[1], imag * w[2]};
    const float qn = sqrtf(q[0] * q[0] + (q[1] * q[1] + (q[2] * q[2] + q[3] * q[3])));
    for (int i = 0; i < 4; i++)
        q[i] = q[i] / qn;
    const float qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    const float tx = 2.f * qx, ty = 2.f * qy, tz = 2.f * qz;
    const float twx = tx * qw, twy = ty * qw, twz = tz * qw;
    const float txx = tx * qx, txy = ty * qx, txz = tz * qx;
    const float tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
    Iso E;
    E.R[0][0] = 1.f - (tyy + tzz);
    E.R[0][1] = txy - twz;
    E.R[0][2] = txz + twy;
    E.R[1][0] = txy + twz;
    E.R[1][1] = 1.f - (txx + tzz);
    E.R[1][2] = tyz - twx;
    E.R[2][0] = txz - twy;
    E.R[2][1] = tyz + twx;
    E.R[2][2] = 1.f - (txx + tyy);
    float V[3][3];
    if (theta < 1e-10f) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                V[i][j] = E.R[i][j];
    }
    else {
        const float W[3][3] = {{0.f, -w[2], w[1]}, {w[2], 0.f, -w[0]}, {-w[1], w[0], 0.f}};
        const float c1 = (1.f - cosf(theta)) / theta_sq;
        const float c2 = (theta - sinf(theta)) / (theta_sq * theta);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const float w2 = W[i][0] * W[0][j] + (W[i][1] * W[1][j] + W[i][2] * W[2][j]);
                V[i][j] = ((i == j ? 1.f : 0.f) + c1 * W[i][j]) + c2 * w2;
            }
    }
    for (int i = 0; i < 3; i++)
        E.t[i] = V[i][0] * a[0] + (V[i][1] * a[1] + V[i][2] * a[2]);
    return E;
}

template<bool Comp>
__global__ __launch_bounds__(kIcpThreads) void k_icp_iter(IcpArgs a)
{
    IcpState* st = a.st;
    if (st->done[a.level])
        return;
    __shared__ float red[kIcpThreads / 64][64];
    __shared__ int is_last;
    const Iso T = load_pose(st);
    const int n = a.w * a.h;
    float s[32], c[32];
    for (int k = 0; k < 32; k++) {
        s[k] = 0.f;
        c[k] = 0.f;
    }
    const int stride = gridDim.x * kIcpThreads;
    for (int p = blockIdx.x * kIcpThreads + threadIdx.x; p < n; p += stride) {
        const int x = p % a.w, y = p / a.w;
        float e = 0.f, J[6];
        const int res = icp_pixel(a, T, x + y * a.w, e, J);
        float v[32];
        row_terms(res, e, J, v);
        for (int k = 0; k < 32; k++) {
            if (Comp) {
                // Kahan summation
                const float yk = v[k] - c[k];
                const float tk = s[k] + yk;
                c[k] = (tk - s[k]) - yk;
                s[k] = tk;
            }
            else {
                s[k] += v[k];
            }
        }
    }
    if (Comp) {
        // (s, -c) is the compensated pair; keep err = -c
        for (int k = 0; k < 32; k++)
            c[k] = -c[k];
    }
    // wave reduction (64 lanes)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        for (int k = 0; k < 32; k++) {
            const float os = __shfl_xor(s[k], off);
            if (Comp) {
                const float oc = __shfl_xor(c[k], off);
                float sum, err;
                two_sum(s[k], os, sum, err);
                s[k] = sum;
                c[k] = c[k] + oc + err;
            }
            else {
                s[k] += os;
            }
        }
    }
    if (lane == 0) {
        for (int k = 0; k < 32; k++) {
            red[wave][k] = s[k];
            red[wave][32 + k] = Comp ? c[k] : 0.f;
        }
    }
    __syncthreads();
    if (threadIdx.x < 32) {
        const int k = threadIdx.x;
        float ss = red[0][k], cc = red[0][32 + k];
        for (int w = 1; w < kIcpThreads / 64; w++) {
            if (Comp) {
                float sum, err;
                two_sum(ss, red[w][k], sum, err);
                ss = sum;
                cc = cc + red[w][32 + k] + err;
            }
            else {
                ss += red[w][k];
            }
        }
        a.partials[blockIdx.x * 64 + k] = ss;
        a.partials[blockIdx.x * 64 + 32 + k] = cc;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0)
        is_last = atomicAdd(&st->ticket, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!is_last)
        return;
    __threadfence();
    // last workgroup: fixed-order cross-workgroup reduction over all 256 threads (32 terms x 8
    // strided groups of workgroup partials, then the 8 group sums in order), solve, pose update.
    // The fence above invalidated this CU's L1, so plain loads see the other workgroups' partials.
    __shared__ float fin[32];
    __shared__ double fin_d[32];
    __shared__ double grp_d[8][32];
    __shared__ float grp[8][32];
    {
        const int k = threadIdx.x & 31, g = threadIdx.x >> 5;
        const float* pp = a.partials;
        if (Comp) {
            double acc = 0.0;
            for (int b = g; b < (int) gridDim.x; b += 8)
                acc += (double) pp[b * 64 + k] + (double) pp[b * 64 + 32 + k];
            grp_d[g][k] = acc;
        }
        else {
            float acc = 0.f;
            for (int b = g; b < (int) gridDim.x; b += 8)
                acc += pp[b * 64 + k];
            grp[g][k] = acc;
        }
    }
    __syncthreads();
    if (threadIdx.x < 32) {
        const int k = threadIdx.x;
        if (Comp) {
            double acc = grp_d[0][k];
            for (int g = 1; g < 8; g++)
                acc += grp_d[g][k];
            fin_d[k] = acc;
            fin[k] = (float) acc;
        }
        else {
            float acc = grp[0][k];
            for (int g = 1; g < 8; g++)
                acc += grp[g][k];
            fin[k] = acc;
            fin_d[k] = (double) acc;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float x[6];
        if (Comp)
            solve_f64(fin_d + 1, x);
        else
            solve_f32(fin + 1, x);
        const Iso D = se3_exp(x);
        const Iso Tn = iso_mul(D, T);
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++)
                st->T[4 * i + j] = Tn.R[i][j];
            st->T[4 * i + 3] = Tn.t[i];
        }
        for (int k = 0; k < 32; k++) {
            st->sums[k] = fin[k];
            st->sums_d[k] = fin_d[k];
        }
        const float xn = sqrtf(x[0] * x[0] + (x[1] * x[1] + (x[2] * x[2] + (x[3] * x[3] + (x[4] * x[4] + x[5] * x[5])))));
        if (xn < a.icp_th)
            st->done[a.level] = 1;
        st->iterations += 1;
        st->ticket = 0;
        __threadfence();
    }
}

__global__ void k_icp_check(IcpState* st, Iso T_ref, int W, int H, float track_th, int comp)
{
    if (threadIdx.x != 0)
        return;
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

static int run_track(se8_tracker* tr, const float* depth, float T16[16], const float* ref_pts, const float* ref_nrm,
                     int* ok, float* sums32, hipStream_t s)
{
    const se8_pinhole& P = tr->sensor;
    const Cam cam = make_cam(P, 1, false);
    const Iso T_ref = iso_from16(T16);
    IcpState* st = (IcpState*) tr->state.ptr;
    SE8_HIP_CHECK(hipEventRecord(tr->ev[0], s));
    k_icp_init<<<1, 64, 0, s>>>(st, T_ref);
    const float* lvl_depth[3] = {depth, (const float*) tr->depth_lvl[1].ptr, (const float*) tr->depth_lvl[2].ptr};
    int lw[3], lh[3];
    for (int l = 0; l < 3; l++) {
        lw[l] = P.width >> l;
        lh[l] = P.height >> l;
    }
    const float e_d = 0.1f * 3;
    for (int l = 1; l < 3; l++) {
        dim3 g((lw[l] + 63) / 64, lh[l]);
        k_half_sample<<<g, 64, 0, s>>>(lvl_depth[l - 1], lw[l - 1], lh[l - 1], (float*) tr->depth_lvl[l].ptr, e_d);
    }
    for (int l = 0; l < 3; l++) {
        const Cam cl = make_cam(P, 1 << l, true);
        dim3 g((cl.w + 63) / 64, cl.h);
        k_vertex_normal<<<g, 64, 0, s>>>(lvl_depth[l], cl, cam.left_hand_frame, (float*) tr->vert_lvl[l].ptr,
                                         (float*) tr->norm_lvl[l].ptr);
    }
    IcpArgs a;
    a.ref_pts = ref_pts;
    a.ref_nrm = ref_nrm;
    a.partials = (float*) tr->partials.ptr;
    a.st = st;
    a.T_ref_inv = iso_inv(T_ref);
    a.cam = cam;
    a.dist_th = tr->cfg.dist_threshold;
    a.normal_th = tr->cfg.normal_threshold;
    a.icp_th = tr->cfg.icp_threshold;
    for (int level = 2; level >= 0; level--) {
        a.vert = (const float*) tr->vert_lvl[level].ptr;
        a.norm = (const float*) tr->norm_lvl[level].ptr;
        // reduction resolution is the full resolution / 2^level (tracker_impl.hpp:85-86)
        a.w = P.width / (1 << level);
        a.h = P.height / (1 << level);
        a.level = level;
        const int nwg = tr->num_wg[level];
        for (int it = 0; it < tr->cfg.iterations[level]; it++) {
            if (tr->cfg.compensated)
                k_icp_iter<true><<<nwg, kIcpThreads, 0, s>>>(a);
            else
                k_icp_iter<false><<<nwg, kIcpThreads, 0, s>>>(a);
        }
    }
    k_icp_check<<<1, 64, 0, s>>>(st, T_ref, P.width, P.height, tr->cfg.track_threshold, tr->cfg.compensated);
    SE8_HIP_CHECK(hipGetLastError());
    SE8_HIP_CHECK(hipEventRecord(tr->ev[1], s));
    IcpState& host = *(IcpState*) tr->host_state;
    SE8_HIP_CHECK(hipMemcpyAsync(&host, st, sizeof(IcpState), hipMemcpyDeviceToHost, s));
    SE8_HIP_CHECK(hipStreamSynchronize(s));
    tr->track_pending = true;
    harvest_tracking(tr);
    Iso T;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T.R[i][j] = host.T[4 * i + j];
        T.t[i] = host.T[4 * i + 3];
    }
    iso_to16(T, T16);
    if (ok)
        *ok = host.ok;
    if (sums32)
        std::memcpy(sums32, host.sums, sizeof(float) * 32);
    return SE8_OK;
}

int launch_raycast(const se8_map* m, const Cam& cam, const Iso& T_WS, float* points, float* normals, int8_t* scale,
                   hipStream_t s);

// ----------------------------------------------------------------------------------------------
// Batched point-to-plane cost and its adjoint (differentiation growth pattern).
// For pose k: cost_k = 0.5 sum_i e_i^2 with e_i the trackKernel residual against fixed reference
// maps raycast at T_ref; grad_k = d cost / d xi (left twist T <- exp(xi) T, xi = [v; w]) at xi = 0,
// which is -sum_i e_i J_i (de/dxi = -J); J^T J (21 terms) is returned for Gauss-Newton steps.
// Grid (pixel workgroups, poses); per-workgroup partials, then one workgroup per pose reduces in
// FP64 in a fixed order.
// ----------------------------------------------------------------------------------------------
constexpr int kAdjTerms = 29; // e^2, e J (6), J^T J (21), inliers

__global__ __launch_bounds__(kIcpThreads) void k_adj_icp(IcpArgs a, const float* poses12, int n_wg, float* partials)
{
    const int k = blockIdx.y;
    Iso T;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T.R[i][j] = poses12[12 * k + 4 * i + j];
        T.t[i] = poses12[12 * k + 4 * i + 3];
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

__global__ void k_adj_icp_reduce(const float* partials, int n_wg, double* out)
{
    const int k = blockIdx.x;
    const int q = threadIdx.x;
    if (q >= kAdjTerms)
        return;
    double acc = 0.0;
    for (int b = 0; b < n_wg; b++)
        acc += (double) partials[((size_t) k * n_wg + b) * kAdjTerms + q];
    out[(size_t) k * kAdjTerms + q] = acc;
}

} // namespace se8

using namespace se8;

extern "C" int se8_tracker_create(const se8_pinhole* sensor, const se8_tracker_config* cfg, se8_tracker** out)
{
    SE8_REQUIRE(sensor && cfg && out, "se8_tracker_create: null argument");
    SE8_REQUIRE(sensor->width >= 4 && sensor->height >= 4, "se8_tracker_create: image too small");
    for (int l = 0; l < 3; l++)
        SE8_REQUIRE(cfg->iterations[l] >= 0 && cfg->iterations[l] < 1000, "se8_tracker_create: bad iterations");
    *out = nullptr;
    se8_tracker* t = new se8_tracker();
    t->sensor = *sensor;
    t->cfg = *cfg;
    int rc = SE8_OK;
    int maxwg = 1;
    for (int l = 0; l < 3; l++) {
        const size_t w = sensor->width >> l, h = sensor->height >> l;
        if (l > 0)
            rc = rc ? rc : dev_alloc(t->depth_lvl[l], sizeof(float) * w * h);
        rc = rc ? rc : dev_alloc(t->vert_lvl[l], sizeof(float) * 3 * w * h);
        rc = rc ? rc : dev_alloc(t->norm_lvl[l], sizeof(float) * 3 * w * h);
        const size_t n = (size_t) (sensor->width / (1 << l)) * (sensor->height / (1 << l));
        t->num_wg[l] = (int) std::max<size_t>(1, (n + kIcpThreads * kIcpPxPerThread - 1) / (kIcpThreads * kIcpPxPerThread));
        maxwg = std::max(maxwg, t->num_wg[l]);
    }
    rc = rc ? rc : dev_alloc(t->partials, sizeof(float) * 64 * maxwg);
    rc = rc ? rc : dev_alloc(t->state, sizeof(IcpState));
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
    for (int l = 0; l < 3; l++) {
        dev_free(t->depth_lvl[l]);
        dev_free(t->vert_lvl[l]);
        dev_free(t->norm_lvl[l]);
    }
    dev_free(t->partials);
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
    delete t;
    return SE8_OK;
}

extern "C" int se8_track(se8_tracker* t, const float* depth_dev, float T_WS[16], const float* ref_points_dev,
                         const float* ref_normals_dev, int* ok, float* sums32, se8_stream stream)
{
    SE8_REQUIRE(t && depth_dev && T_WS && ref_points_dev && ref_normals_dev, "se8_track: null argument");
    return run_track(t, depth_dev, T_WS, ref_points_dev, ref_normals_dev, ok, sums32, (hipStream_t) stream);
}

extern "C" int se8_track_raycast(se8_tracker* t, const se8_map* m, const float* depth_dev, float T_WS[16], int* ok,
                                 se8_stream stream)
{
    SE8_REQUIRE(t && m && depth_dev && T_WS, "se8_track_raycast: null argument");
    const size_t n = (size_t) t->sensor.width * t->sensor.height;
    int rc = dev_alloc(t->ref_points, sizeof(float) * 3 * n);
    rc = rc ? rc : dev_alloc(t->ref_normals, sizeof(float) * 3 * n);
    if (rc != SE8_OK)
        return rc;
    hipStream_t s = (hipStream_t) stream;
    const Cam cam = make_cam(t->sensor, 1, false);
    rc = launch_raycast(m, cam, iso_from16(T_WS), (float*) t->ref_points.ptr, (float*) t->ref_normals.ptr, nullptr, s);
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
    const int n_wg = t->num_wg[0];
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
    const Cam c0 = make_cam(P, 1, true);
    dim3 g((c0.w + 63) / 64, c0.h);
    k_vertex_normal<<<g, 64, 0, s>>>(depth_dev, c0, c0.left_hand_frame, (float*) t->vert_lvl[0].ptr,
                                     (float*) t->norm_lvl[0].ptr);
    IcpArgs a;
    a.vert = (const float*) t->vert_lvl[0].ptr;
    a.norm = (const float*) t->norm_lvl[0].ptr;
    a.ref_pts = ref_points_dev;
    a.ref_nrm = ref_normals_dev;
    a.partials = nullptr;
    a.st = nullptr;
    a.T_ref_inv = iso_inv(iso_from16(T_ref16));
    a.cam = make_cam(P, 1, false);
    a.w = P.width;
    a.h = P.height;
    a.dist_th = t->cfg.dist_threshold;
    a.normal_th = t->cfg.normal_threshold;
    a.icp_th = t->cfg.icp_threshold;
    a.level = 0;
    k_adj_icp<<<dim3(n_wg, n_poses), kIcpThreads, 0, s>>>(a, (const float*) t->adj_poses.ptr, n_wg,
                                                          (float*) t->adj_partials.ptr);
    k_adj_icp_reduce<<<n_poses, 64, 0, s>>>((const float*) t->adj_partials.ptr, n_wg, (double*) t->adj_out.ptr);
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
