/*
 * CPU oracle for the single-resolution TSDF path (allocation, update, raycast, ICP).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the reference
 * algorithm, used by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg
 * of bench.py, to check results.  The product (libse8.so) never links, loads or calls it.
 *
 * It is compiled twice by oracle/Makefile:
 *   liboracle_f32.so  (SE8_REAL=float)  - same fp32 operation order as the GPU path
 *   liboracle_f64.so  (SE8_REAL=double) - FP64 twin used to report precision errors
 * with -O2 -ffp-contract=off -fno-fast-math and OpenMP loops where the reference has them.
 *
 * Floating-point evaluation order (shared with the GPU kernels):
 *   dot3(a,b)       = a0*b0 + (a1*b1 + a2*b2)         (Eigen's 3-term redux split)
 *   (R*v)_i         = R_i0*v0 + (R_i1*v1 + R_i2*v2)
 *   (T*p)_i         = (R*p)_i + t_i
 *   plane . [c;1]   = (n0*c0 + n1*c1) + (n2*c2 + n3)
 *   project(p)      = (fx*(x*(1/z)) + cx, fy*(y*(1/z)) + cy), success iff
 *                     |z| >= 1e-12, u,v in [-0.5, W-0.5) x [-0.5, H-0.5), z > 0
 *   backProject(u,v)= ((u-cx)*(1/fx), (v-cy)*(1/fy), 1)
 * srl_projection (the reference's camera library) is not vendored, so the last two
 * lines are the convention implied by the reference's call sites; they are
 * "parity unpinned" against the real library.
 */
#include <tgmath.h> /* type-generic sqrt/sin/cos/floor follow the real type */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef SE8_REAL
#define SE8_REAL float
#endif
typedef SE8_REAL real;

#define OR_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* Key operations: include/se/map/utils/impl/key_util_impl.hpp:44-247,406-436 */
/* masks: include/se/map/utils/key_util.hpp:19-58                            */
/* ------------------------------------------------------------------------- */
#define OR_SCALE_OFFSET 5
#define OR_SCALE_MASK 0x1Full

static uint64_t or_code_mask(int scale)
{
    /* bit pattern of CODE_MASK[scale]: 57-bit all-ones with the low 3*scale bits cleared */
    const uint64_t full = 0x1ffffffffffffffull;
    return (full >> (3 * scale)) << (3 * scale);
}

OR_API uint64_t or_expand(uint64_t v)
{
    uint64_t x = v & 0x7FFFFull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

OR_API uint64_t or_compact(uint64_t v)
{
    uint64_t x = v & 0x49249249249249ull;
    x = (x | x >> 2) & 0x10c30c30c30c30c3ull;
    x = (x | x >> 4) & 0x100f00f00f00f00full;
    x = (x | x >> 8) & 0x1f0000ff0000ffull;
    x = (x | x >> 16) & 0x1f00000000ffffull;
    x = (x | x >> 32) & 0x1fffffull;
    return x;
}

OR_API uint64_t or_encode_code(int x, int y, int z)
{
    return or_expand((uint64_t) x) | (or_expand((uint64_t) y) << 1) | (or_expand((uint64_t) z) << 2);
}

OR_API void or_decode_code(uint64_t code, int out[3])
{
    out[0] = (int) or_compact(code);
    out[1] = (int) or_compact(code >> 1);
    out[2] = (int) or_compact(code >> 2);
}

/* returns 1 when no coordinate detail is lost at this scale (key_util_impl.hpp:74-86) */
OR_API int or_encode_key(const int c[3], int scale, uint64_t* key)
{
    const uint64_t detailed = or_encode_code(c[0], c[1], c[2]);
    const uint64_t code = detailed & or_code_mask(scale);
    *key = code << OR_SCALE_OFFSET | (uint64_t) scale;
    return detailed == code;
}

OR_API void or_decode_key(uint64_t key, int out[3], int* scale)
{
    *scale = (int) (key & OR_SCALE_MASK);
    or_decode_code(key >> OR_SCALE_OFFSET, out);
}

OR_API int or_code_to_child_idx(uint64_t code, int child_scale)
{
    return (int) ((code >> (child_scale * 3)) & 7);
}

/* key_util_impl.hpp:212-219 */
OR_API int or_key_at_scale(uint64_t key, int scale, uint64_t* out)
{
    *out = (key & (or_code_mask(scale) << OR_SCALE_OFFSET)) | (uint64_t) scale;
    return (int) (key & OR_SCALE_MASK) <= scale;
}

OR_API uint64_t or_parent_key(uint64_t key)
{
    const int ps = (int) (key & OR_SCALE_MASK) + 1;
    return (key & (or_code_mask(ps) << OR_SCALE_OFFSET)) | (uint64_t) ps;
}

OR_API uint64_t or_block_key(uint64_t key, int max_block_scale)
{
    return key & ((or_code_mask(max_block_scale) << OR_SCALE_OFFSET) | OR_SCALE_MASK);
}

OR_API int or_is_child(uint64_t parent, uint64_t child)
{
    const int ps = (int) (parent & OR_SCALE_MASK), cs = (int) (child & OR_SCALE_MASK);
    if (cs >= ps)
        return 0;
    return ((parent >> OR_SCALE_OFFSET) ^ ((child >> OR_SCALE_OFFSET) & or_code_mask(ps))) == 0;
}

OR_API int or_is_siblings(uint64_t a, uint64_t b)
{
    const int sa = (int) (a & OR_SCALE_MASK), sb = (int) (b & OR_SCALE_MASK);
    if (sa != sb)
        return 0;
    const uint64_t m = or_code_mask(sa + 1);
    return (((a >> OR_SCALE_OFFSET) & m) ^ ((b >> OR_SCALE_OFFSET) & m)) == 0;
}

static int cmp_u64(const void* a, const void* b)
{
    const uint64_t x = *(const uint64_t*) a, y = *(const uint64_t*) b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

OR_API void or_sort_keys(uint64_t* keys, long n)
{
    qsort(keys, (size_t) n, sizeof(uint64_t), cmp_u64);
}

/* unique_keys (key_util_impl.hpp:322-340): keys sorted in place, returns count */
OR_API long or_unique_keys(uint64_t* keys, long n, uint64_t* out)
{
    if (n == 0)
        return 0;
    or_sort_keys(keys, n);
    long m = 0;
    out[m++] = keys[0];
    for (long i = 0; i < n; i++)
        if (out[m - 1] != keys[i])
            out[m++] = keys[i];
    return m;
}

/* unique_codes (key_util_impl.hpp:344-366) */
OR_API long or_unique_codes(uint64_t* keys, long n, uint64_t* out)
{
    if (n == 0)
        return 0;
    or_sort_keys(keys, n);
    long m = 0;
    out[m++] = keys[0];
    for (long i = 0; i < n; i++) {
        if ((out[m - 1] >> OR_SCALE_OFFSET) != (keys[i] >> OR_SCALE_OFFSET))
            out[m++] = keys[i];
        else if ((out[m - 1] & OR_SCALE_MASK) > (keys[i] & OR_SCALE_MASK))
            out[m - 1] = keys[i];
    }
    return m;
}

/* unique_allocation (key_util_impl.hpp:370-396) */
OR_API long or_unique_allocation(uint64_t* keys, long n, int max_block_scale, uint64_t* out)
{
    if (n == 0)
        return 0;
    or_sort_keys(keys, n);
    long m = 0;
    out[m++] = or_block_key(keys[0], max_block_scale);
    for (long i = 0; i < n; i++) {
        const uint64_t bk = or_block_key(keys[i], max_block_scale);
        if (or_is_child(out[m - 1], bk))
            out[m - 1] = bk;
        else if (out[m - 1] != bk)
            out[m++] = bk;
    }
    return m;
}

/* unique_at_scale (key_util_impl.hpp:406-436) */
OR_API long or_unique_at_scale(uint64_t* keys, long n, int scale, uint64_t* out)
{
    if (n == 0)
        return 0;
    or_sort_keys(keys, n);
    long m = 0, restart = 1;
    uint64_t k;
    for (long i = 0; i < n; i++) {
        if (or_key_at_scale(keys[i], scale, &k)) {
            out[m++] = k;
            break;
        }
        restart++;
    }
    for (long i = restart; i < n; i++)
        if (or_key_at_scale(keys[i], scale, &k) && out[m - 1] != k)
            out[m++] = k;
    return m;
}

/* ------------------------------------------------------------------------- */
/* Small fixed-size linear algebra in the documented evaluation order.        */
/* ------------------------------------------------------------------------- */
typedef struct {
    real R[3][3];
    real t[3];
} or_iso;

static inline real dot3(const real a[3], const real b[3])
{
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

static inline void lin_apply(const or_iso* T, const real v[3], real out[3])
{
    real o[3];
    for (int i = 0; i < 3; i++)
        o[i] = T->R[i][0] * v[0] + (T->R[i][1] * v[1] + T->R[i][2] * v[2]);
    memcpy(out, o, sizeof(o));
}

static inline void iso_apply(const or_iso* T, const real v[3], real out[3])
{
    real o[3];
    for (int i = 0; i < 3; i++)
        o[i] = (T->R[i][0] * v[0] + (T->R[i][1] * v[1] + T->R[i][2] * v[2])) + T->t[i];
    memcpy(out, o, sizeof(o));
}

/* Isometry product (Eigen transform_transform_product_impl): linear = A.R*B.R,
 * translation = A.R*B.t + A.t */
static void iso_mul(const or_iso* A, const or_iso* B, or_iso* out)
{
    or_iso o;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            o.R[i][j] = A->R[i][0] * B->R[0][j] + (A->R[i][1] * B->R[1][j] + A->R[i][2] * B->R[2][j]);
        o.t[i] = (A->R[i][0] * B->t[0] + (A->R[i][1] * B->t[1] + A->R[i][2] * B->t[2])) + A->t[i];
    }
    *out = o;
}

/* Isometry inverse: R' = R^T, t' = -(R' t) */
static void iso_inv(const or_iso* A, or_iso* out)
{
    or_iso o;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            o.R[i][j] = A->R[j][i];
    for (int i = 0; i < 3; i++)
        o.t[i] = -(o.R[i][0] * A->t[0] + (o.R[i][1] * A->t[1] + o.R[i][2] * A->t[2]));
    *out = o;
}

static void iso_from16(const real* m, or_iso* T)
{
    /* row-major 4x4 */
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T->R[i][j] = m[4 * i + j];
        T->t[i] = m[4 * i + 3];
    }
}

static void iso_to16(const or_iso* T, real* m)
{
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            m[4 * i + j] = T->R[i][j];
        m[4 * i + 3] = T->t[i];
    }
    m[12] = 0;
    m[13] = 0;
    m[14] = 0;
    m[15] = 1;
}

static inline real sqnorm3(const real v[3])
{
    return dot3(v, v);
}

static inline real norm3(const real v[3])
{
    return sqrt(sqnorm3(v));
}

static inline void normalized3(const real v[3], real out[3])
{
    const real z = sqnorm3(v);
    if (z > 0) {
        const real n = sqrt(z);
        out[0] = v[0] / n;
        out[1] = v[1] / n;
        out[2] = v[2] / n;
    }
    else {
        out[0] = v[0];
        out[1] = v[1];
        out[2] = v[2];
    }
}

static inline void cross3(const real a[3], const real b[3], real out[3])
{
    real o[3];
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
    memcpy(out, o, sizeof(o));
}

/* ------------------------------------------------------------------------- */
/* Pinhole camera: src/sensor/pinhole_camera.cpp:29-111,189-310,             */
/* include/se/sensor/impl/pinhole_camera_impl.hpp:15-32                      */
/* ------------------------------------------------------------------------- */
typedef struct {
    int width, height;
    float fx, fy, cx, cy;
    float near_plane, far_plane;
} or_pinhole; /* same layout as se8_pinhole in include/se8.h */

typedef struct {
    int w, h;
    real fx, fy, cx, cy, inv_fx, inv_fy, near_plane, far_plane;
    real planes[6][4];
    int left_hand_frame;
} or_cam;

static void cam_backproject(const or_cam* c, real u, real v, real out[3])
{
    out[0] = (u - c->cx) * c->inv_fx;
    out[1] = (v - c->cy) * c->inv_fy;
    out[2] = 1;
}

static int cam_project(const or_cam* c, const real p[3], real uv[2])
{
    if (fabs((double) p[2]) < 1.0e-12)
        return 0;
    const real rz = (real) 1 / p[2];
    const real x = p[0] * rz, y = p[1] * rz;
    uv[0] = c->fx * x + c->cx;
    uv[1] = c->fy * y + c->cy;
    if (uv[0] < (real) -0.5 || uv[1] < (real) -0.5)
        return 0;
    if (uv[0] >= (real) c->w - (real) 0.5 || uv[1] >= (real) c->h - (real) 0.5)
        return 0;
    return p[2] > 0;
}

static void plane_normal(const real p1[3], const real p2[3], const real p3[3], real out[3])
{
    real t1[3], t2[3], n[3];
    for (int i = 0; i < 3; i++) {
        t1[i] = p2[i] - p1[i];
        t2[i] = p3[i] - p2[i];
    }
    cross3(t1, t2, n);
    normalized3(n, out);
}

/* scaled = 0: the sensor itself; scaled = 1: PinholeCamera(sensor, dsf) as built for the tracker
 * pyramid, which applies the downsampling formulas even for dsf = 1 (pinhole_camera.cpp:95-111) */
static void cam_init_s(or_cam* c, const or_pinhole* p, int dsf, int scaled)
{
    const real s = (real) dsf;
    c->w = (int) ((real) p->width / s);
    c->h = (int) ((real) p->height / s);
    if (!scaled) {
        c->fx = p->fx;
        c->fy = p->fy;
        c->cx = p->cx;
        c->cy = p->cy;
    }
    else {
        c->fx = (real) p->fx / s;
        c->fy = (real) p->fy / s;
        c->cx = ((real) p->cx + (real) 0.5) / s - (real) 0.5;
        c->cy = ((real) p->cy + (real) 0.5) / s - (real) 0.5;
    }
    c->inv_fx = (real) 1 / c->fx;
    c->inv_fy = (real) 1 / c->fy;
    c->near_plane = p->near_plane;
    c->far_plane = p->far_plane;
    c->left_hand_frame = (p->fx < 0.0f) ^ (p->fy < 0.0f);

    /* frustum vertices: corners (0,0) (W,0) (W,H) (0,H) at near and far */
    real tl[3], tr[3], br[3], bl[3];
    cam_backproject(c, 0, 0, tl);
    cam_backproject(c, (real) c->w, 0, tr);
    cam_backproject(c, (real) c->w, (real) c->h, br);
    cam_backproject(c, 0, (real) c->h, bl);
    real v[8][3]; /* TLN TRN BRN BLN TLF TRF BRF BLF */
    for (int i = 0; i < 3; i++) {
        v[0][i] = c->near_plane * tl[i];
        v[4][i] = c->far_plane * tl[i];
        v[1][i] = c->near_plane * tr[i];
        v[5][i] = c->far_plane * tr[i];
        v[2][i] = c->near_plane * br[i];
        v[6][i] = c->far_plane * br[i];
        v[3][i] = c->near_plane * bl[i];
        v[7][i] = c->far_plane * bl[i];
    }
    plane_normal(v[4], v[0], v[3], c->planes[0]); /* left   */
    plane_normal(v[1], v[5], v[6], c->planes[1]); /* right  */
    plane_normal(v[7], v[3], v[2], c->planes[2]); /* bottom */
    plane_normal(v[5], v[1], v[0], c->planes[3]); /* top    */
    for (int i = 0; i < 4; i++)
        c->planes[i][3] = 0;
    c->planes[4][0] = 0;
    c->planes[4][1] = 0;
    c->planes[4][2] = 1;
    c->planes[4][3] = -c->near_plane;
    c->planes[5][0] = 0;
    c->planes[5][1] = 0;
    c->planes[5][2] = -1;
    c->planes[5][3] = c->far_plane;
    real sign = c->fx * c->fy;
    sign = sign / (real) fabs((double) sign);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 3; j++)
            c->planes[i][j] *= sign;
}

static void cam_init(or_cam* c, const or_pinhole* p, int dsf)
{
    cam_init_s(c, p, dsf, dsf != 1);
}

static int cam_sphere_in_frustum(const or_cam* c, const real ctr[3], real radius)
{
    for (int i = 0; i < 6; i++) {
        const real* n = c->planes[i];
        const real d = (n[0] * ctr[0] + n[1] * ctr[1]) + (n[2] * ctr[2] + n[3]);
        if (d < -radius)
            return 0;
    }
    return 1;
}

static inline real cam_far_dist(const or_cam* c, const real ray[3])
{
    real n[3];
    normalized3(ray, n);
    return c->far_plane / n[2];
}

static inline real cam_near_dist(const or_cam* c, const real ray[3])
{
    real n[3];
    normalized3(ray, n);
    return c->near_plane / n[2];
}

/* round_pixel: include/se/common/impl/image_utils_impl.hpp:12-15 */
static inline void round_pixel(const real uv[2], int px[2])
{
    px[0] = (int) (uv[0] + (real) 0.5);
    px[1] = (int) (uv[1] + (real) 0.5);
}

/* ------------------------------------------------------------------------- */
/* Ouster lidar: src/sensor/ouster_lidar.cpp:59-253,                          */
/* include/se/sensor/impl/ouster_lidar_impl.hpp:15-32.  The projection model   */
/* itself (srl::projection::OusterLidar) is not vendored with the reference;   */
/* the model below is this restatement's own definition, shared operation for */
/* operation with the GPU kernels (parity unpinned against srl-projection):   */
/*   beam v has elevation e_v (strictly descending rows) and azimuth offset    */
/*   a_v; column u looks along phi = 2 pi u / W - pi + a_v;                    */
/*   backProject(u, v) = (cos e_v cos phi, cos e_v sin phi, sin e_v), computed */
/*   in double and rounded to float once, into a per-pixel table;              */
/*   project(p): e = atan2(p_z, |p_xy|) interpolated linearly between beam     */
/*   elevations to a fractional row (extrapolated by the end gaps), accepted   */
/*   for rows in [-0.5, H - 0.5); u = (atan2(p_y, p_x) - a_row + pi) W / 2 pi, */
/*   wrapped into [-0.5, W - 0.5).                                             */
/* atan and sin are fixed polynomials so that the host and the GPU round       */
/* identically: atan(t) = t P(t^2) on [0, 1] (a least-squares fit, max error  */
/* 8.2e-8) and the Taylor series of sin to x^17 on [0, pi].                    */
/* ------------------------------------------------------------------------- */
#define OR_PI 3.14159265358979323846

static inline real atan_unit(real t)
{
    static const real c[9] = {(real) 0.9999998807907104, (real) -0.33332598209381104, (real) 0.19985906779766083,
                              (real) -0.141612246632576,  (real) 0.10498932003974915, (real) -0.07234828174114227,
                              (real) 0.0397808775305748,  (real) -0.014401142485439777, (real) 0.0024566699285060167};
    const real z = t * t;
    real r = c[8];
    for (int k = 7; k >= 0; k--)
        r = r * z + c[k];
    return t * r;
}

static inline real atan2_det(real y, real x)
{
    const real ax = (real) fabs((double) x), ay = (real) fabs((double) y);
    const real mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    if (mx == 0)
        return 0;
    real a = atan_unit(mn / mx);
    if (ay > ax)
        a = (real) (OR_PI / 2) - a;
    if (x < 0)
        a = (real) OR_PI - a;
    if (y < 0)
        a = -a;
    return a;
}

static inline real sin_det(real x) /* x in [0, pi] */
{
    static const real c[9] = {(real) 1.0,
                              (real) -0.1666666716337204,
                              (real) 0.008333333767950535,
                              (real) -0.00019841270113829523,
                              (real) 2.7557318844628753e-06,
                              (real) -2.5052107943679403e-08,
                              (real) 1.6059044372074283e-10,
                              (real) -7.647163609812713e-13,
                              (real) 2.8114573589663704e-15};
    const real z = x * x;
    real r = c[8];
    for (int k = 7; k >= 0; k--)
        r = r * z + c[k];
    return x * r;
}

typedef struct {
    int w, h;
    real near_plane, far_plane;
    real min_elev, max_elev; /* radians, as OusterLidar's constructor computes them (float) */
    real* elev;              /* h beam elevations, radians, strictly descending */
    real* az;                /* h beam azimuth offsets, radians */
    real* dir;               /* h * w * 3 back-projection directions, row-major */
    real cols_per_rad;       /* W / (2 pi) */
} or_lidar;

/* Returns NULL unless h >= 2 and the elevations are strictly descending. */
OR_API void* or_lidar_create(int w, int h, float near_plane, float far_plane, const float* elev_deg, const float* az_deg)
{
    if (w < 1 || h < 2)
        return NULL;
    for (int v = 1; v < h; v++)
        if (!(elev_deg[v] < elev_deg[v - 1]))
            return NULL;
    or_lidar* L = (or_lidar*) calloc(1, sizeof(or_lidar));
    L->w = w;
    L->h = h;
    L->near_plane = near_plane;
    L->far_plane = far_plane;
    L->elev = (real*) malloc(sizeof(real) * h);
    L->az = (real*) malloc(sizeof(real) * h);
    L->dir = (real*) malloc(sizeof(real) * 3 * (size_t) w * h);
    const float deg_to_rad = (float) (OR_PI / 180.0); /* ouster_lidar.cpp:80 */
    float mn = elev_deg[0], mx = elev_deg[0];
    for (int v = 0; v < h; v++) {
        L->elev[v] = (real) (elev_deg[v] * deg_to_rad);
        L->az[v] = (real) (az_deg[v] * deg_to_rad);
        mn = elev_deg[v] < mn ? elev_deg[v] : mn;
        mx = elev_deg[v] > mx ? elev_deg[v] : mx;
        const double e = (double) elev_deg[v] * (OR_PI / 180.0), a = (double) az_deg[v] * (OR_PI / 180.0);
        for (int u = 0; u < w; u++) {
            const double phi = 2.0 * OR_PI * (double) u / (double) w - OR_PI + a;
            real* d = L->dir + 3 * ((size_t) v * w + u);
            d[0] = (real) (float) (cos(e) * cos(phi));
            d[1] = (real) (float) (cos(e) * sin(phi));
            d[2] = (real) (float) sin(e);
        }
    }
    L->min_elev = (real) (mn * deg_to_rad);
    L->max_elev = (real) (mx * deg_to_rad);
    L->cols_per_rad = (real) (float) ((double) w / (2.0 * OR_PI));
    return L;
}

OR_API void or_lidar_destroy(void* vl)
{
    or_lidar* L = (or_lidar*) vl;
    if (!L)
        return;
    free(L->elev);
    free(L->az);
    free(L->dir);
    free(L);
}

static int lidar_project(const or_lidar* L, const real p[3], real uv[2])
{
    const real hxy = (real) sqrt((double) (p[0] * p[0] + p[1] * p[1]));
    if (hxy == 0 && p[2] == 0)
        return 0;
    const real e = atan2_det(p[2], hxy);
    const int h = L->h;
    real vf;
    if (e >= L->elev[0])
        vf = (L->elev[0] - e) / (L->elev[0] - L->elev[1]);
    else if (e <= L->elev[h - 1])
        vf = (real) (h - 1) + (L->elev[h - 1] - e) / (L->elev[h - 2] - L->elev[h - 1]);
    else {
        int lo = 0, hi = h - 1; /* elev[lo] > e > elev[hi] */
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (L->elev[mid] >= e)
                lo = mid;
            else
                hi = mid;
        }
        vf = (real) lo + (L->elev[lo] - e) / (L->elev[lo] - L->elev[lo + 1]);
    }
    if (!(vf >= (real) -0.5 && vf < (real) h - (real) 0.5))
        return 0;
    const int row = (int) (vf + (real) 0.5);
    real uf = ((atan2_det(p[1], p[0]) - L->az[row]) + (real) OR_PI) * L->cols_per_rad;
    for (int k = 0; k < 2 && uf >= (real) L->w - (real) 0.5; k++)
        uf = uf - (real) L->w;
    for (int k = 0; k < 2 && uf < (real) -0.5; k++)
        uf = uf + (real) L->w;
    if (!(uf >= (real) -0.5 && uf < (real) L->w - (real) 0.5))
        return 0;
    uv[0] = uf;
    uv[1] = vf;
    return 1;
}

/* OusterLidar::sphereInFrustumImpl (ouster_lidar.cpp:195-224); asin(c_z / |c|) is evaluated as
 * atan2(c_z, |c_xy|), equal in exact arithmetic. */
static int lidar_sphere_in_frustum(const or_lidar* L, const real c[3], real radius)
{
    const real n = norm3(c);
    if (n - radius > L->far_plane)
        return 0;
    if (n + radius < L->near_plane)
        return 0;
    const real e = atan2_det(c[2], (real) sqrt((double) (c[0] * c[0] + c[1] * c[1])));
    if (e < L->min_elev) {
        const real de = (real) fabs((double) (e - L->min_elev));
        if (sin_det(de) * n > radius)
            return 0;
    }
    if (e > L->max_elev) {
        const real de = (real) fabs((double) (e - L->max_elev));
        if (sin_det(de) * n > radius)
            return 0;
    }
    return 1;
}

OR_API int or_lidar_project(void* vl, const real p[3], real uv[2])
{
    return lidar_project((const or_lidar*) vl, p, uv);
}

OR_API void or_lidar_project_n(void* vl, const real* pts, long n, real* uv, unsigned char* ok)
{
    for (long i = 0; i < n; i++)
        ok[i] = (unsigned char) lidar_project((const or_lidar*) vl, pts + 3 * i, uv + 2 * i);
}

OR_API void or_lidar_backproject(void* vl, int u, int v, real out[3])
{
    const or_lidar* L = (const or_lidar*) vl;
    const real* d = L->dir + 3 * ((size_t) v * L->w + u);
    out[0] = d[0];
    out[1] = d[1];
    out[2] = d[2];
}

OR_API int or_lidar_sphere_in_frustum(void* vl, const real c[3], real radius)
{
    return lidar_sphere_in_frustum((const or_lidar*) vl, c, radius);
}

OR_API real or_atan2_det(real y, real x)
{
    return atan2_det(y, x);
}

OR_API real or_sin_det(real x)
{
    return sin_det(x);
}

/* The sensor seen by integration: a pinhole camera or a lidar, through the interface the
 * reference's templated integrator uses (backProject, project, far/near, measurementFromPoint,
 * sphereInFrustum). */
typedef struct {
    const or_cam* cam;   /* pinhole when not NULL */
    const or_lidar* lid; /* otherwise */
    int w, h;
    real near_plane, far_plane;
} or_sensor;

static inline void sensor_backproject(const or_sensor* s, int x, int y, real dir[3])
{
    if (s->cam)
        cam_backproject(s->cam, (real) x, (real) y, dir);
    else
        or_lidar_backproject((void*) s->lid, x, y, dir);
}

static inline int sensor_project(const or_sensor* s, const real p[3], real uv[2])
{
    return s->cam ? cam_project(s->cam, p, uv) : lidar_project(s->lid, p, uv);
}

static inline real sensor_far_dist(const or_sensor* s, const real p[3])
{
    return s->cam ? cam_far_dist(s->cam, p) : s->far_plane;
}

static inline real sensor_measurement(const or_sensor* s, const real p[3])
{
    return s->cam ? p[2] : norm3(p);
}

static inline int sensor_sphere_in_frustum(const or_sensor* s, const real c[3], real radius)
{
    return s->cam ? cam_sphere_in_frustum(s->cam, c, radius) : lidar_sphere_in_frustum(s->lid, c, radius);
}

/* ------------------------------------------------------------------------- */
/* Octree + map: include/se/map/octree/impl/octree_impl.hpp:14-19,171-199,    */
/* include/se/map/impl/map_impl.hpp:37-64,494-590                             */
/* ------------------------------------------------------------------------- */
#define OR_BLOCK 8
#define OR_BLOCK_SCALE 3

typedef struct or_octant {
    struct or_octant* parent;
    struct or_octant* child[8];
    int coord[3];
    int size;
    int is_block;
    int timestamp;
    unsigned char child_mask;
    real* tsdf;   /* 512, x fastest, blocks only */
    real* weight; /* 512 */
    /* multi-resolution blocks: 585 entries (scales 0..3: 512, 64, 8, 1, x fastest) of {tsdf, weight},
     * then their past data; current / minimum scale (-1 = none) and the scale chosen this frame */
    real* mr;
    int cur_scale, min_scale, new_scale;
    long slot; /* blocks: index in or_map.blocks */
} or_octant;

typedef struct {
    int size, max_scale;
    real res, mu, max_weight, mu_factor;
    real dim;       /* octree size * res */
    real t_MW[3];   /* map-from-world translation (rotation is identity) */
    real t_WM[3];
    or_octant* root;
    or_octant** blocks;
    long n_blocks, cap_blocks;
    long n_nodes;
    int aabb_min[3], aabb_max[3], aabb_empty;
    int multires; /* TSDFMap<Res::Multi>: blocks carry the multi-resolution data */
    int storage_f16; /* tsdf stored as IEEE binary16: rounded to the nearest half after each update */
    /* while set (or_map_touch_begin .. _end): one bit per voxel of every block, raised by each read
     * of the voxel's data through getData, the interpolation corners and the gradient samples */
    uint64_t* touch;
    long touch_blocks;
    /* keys of the blocks the last integration updated (allocated + frustum), in update order */
    uint64_t* last_update;
    long n_last_update;
} or_map;

/* Per-ray march statistics of or_raycast_stats (test infrastructure): the ray being cast on this
 * thread accumulates into 7 counters: [0] octree traversal iterations of the first-block search,
 * [1] samples in a brick (scale-6 octant) that does not exist, [2] samples in an unallocated block
 * of an existing brick, [3] samples in an allocated block at an unobserved voxel, [4] samples at an
 * observed voxel, [5] interpolations, [6] the samples of [2] whose 32-voxel octant does not exist. */
static __thread int* or_cur_stats;

/* records a read of voxel i of block b (no-op unless recording) */
static inline void touch_voxel(const or_map* m, const or_octant* b, int i)
{
    if (m->touch && b->slot < m->touch_blocks)
        __atomic_fetch_or(&m->touch[b->slot * 8 + (i >> 6)], 1ull << (i & 63), __ATOMIC_RELAXED);
}

static unsigned or_power_two_up(float x)
{
    const int p = (int) ceil(log2((double) x));
    return 1u << p;
}

static or_octant* new_octant(or_octant* parent, const int coord[3], int size, int is_block)
{
    or_octant* o = (or_octant*) calloc(1, sizeof(or_octant));
    o->parent = parent;
    memcpy(o->coord, coord, sizeof(int) * 3);
    o->size = size;
    o->is_block = is_block;
    o->timestamp = -1;
    if (is_block) {
        o->tsdf = (real*) malloc(sizeof(real) * 512);
        o->weight = (real*) malloc(sizeof(real) * 512);
        for (int i = 0; i < 512; i++) {
            o->tsdf[i] = 1; /* FieldData<TSDF>{} = {1, 0}, data_field.hpp:86-87 */
            o->weight[i] = 0;
        }
    }
    return o;
}

OR_API void* or_map_create(float dim, float res, float mu_factor, float max_weight)
{
    or_map* m = (or_map*) calloc(1, sizeof(or_map));
    int size = (int) ceil((double) (dim / res));
    if (size < 2 * OR_BLOCK)
        size = 2 * OR_BLOCK;
    m->size = (int) or_power_two_up((float) size);
    int s = 0;
    while ((1 << s) < m->size)
        s++;
    m->max_scale = s;
    m->res = res;
    m->mu = (real) res * (real) mu_factor;
    m->mu_factor = mu_factor;
    m->max_weight = max_weight;
    m->dim = (real) m->size * (real) res;
    /* default T_MW = Translation(dim/2) with the user's dim (map_impl.hpp:697) */
    for (int i = 0; i < 3; i++) {
        m->t_MW[i] = (real) (dim / 2.0f);
        m->t_WM[i] = -m->t_MW[i];
    }
    const int zero[3] = {0, 0, 0};
    m->root = new_octant(NULL, zero, m->size, 0);
    m->n_nodes = 1;
    m->cap_blocks = 1024;
    m->blocks = (or_octant**) malloc(sizeof(or_octant*) * m->cap_blocks);
    m->aabb_empty = 1;
    return m;
}

static void free_octant(or_octant* o)
{
    if (!o)
        return;
    for (int i = 0; i < 8; i++)
        free_octant(o->child[i]);
    free(o->tsdf);
    free(o->weight);
    free(o->mr);
    free(o);
}

OR_API void or_map_destroy(void* vm)
{
    or_map* m = (or_map*) vm;
    free_octant(m->root);
    free(m->blocks);
    free(m->touch);
    free(m->last_update);
    free(m);
}

OR_API int or_map_size(void* vm)
{
    return ((or_map*) vm)->size;
}

OR_API long or_map_num_blocks(void* vm)
{
    return ((or_map*) vm)->n_blocks;
}

OR_API long or_map_num_nodes(void* vm)
{
    return ((or_map*) vm)->n_nodes;
}

/* Map::contains (map_impl.hpp:58-64) with an identity-rotation T_MW */
static inline int map_contains(const or_map* m, const real pW[3], real pM[3])
{
    for (int i = 0; i < 3; i++)
        pM[i] = pW[i] + m->t_MW[i];
    return pM[0] >= 0 && pM[0] < m->dim && pM[1] >= 0 && pM[1] < m->dim && pM[2] >= 0 && pM[2] < m->dim;
}

/* pointToVoxel<Safe::On> (map_impl.hpp:538-550) */
static inline int map_point_to_voxel(const or_map* m, const real pW[3], int v[3])
{
    real pM[3];
    if (!map_contains(m, pW, pM)) {
        v[0] = v[1] = v[2] = -1;
        return 0;
    }
    /* a point within an ulp of the upper bound can round to index == size; wrap it as the
     * top-down fetcher does (it ignores bits at and above the root size) */
    for (int i = 0; i < 3; i++)
        v[i] = ((int) (pM[i] / m->res)) & (m->size - 1);
    return 1;
}

static inline int map_point_to_voxel_f(const or_map* m, const real pW[3], real v[3])
{
    real pM[3];
    if (!map_contains(m, pW, pM)) {
        v[0] = v[1] = v[2] = -1;
        return 0;
    }
    for (int i = 0; i < 3; i++)
        v[i] = pM[i] / m->res;
    return 1;
}

/* fetcher::block (fetcher_impl.hpp:107-126) */
static or_octant* fetch_block(const or_map* m, const int c[3])
{
    or_octant* p = m->root;
    or_octant* child = NULL;
    for (int cs = p->size >> 1; cs >= OR_BLOCK; cs >>= 1) {
        const int idx = ((c[0] & cs) > 0) + 2 * ((c[1] & cs) > 0) + 4 * ((c[2] & cs) > 0);
        child = p->child[idx];
        if (!child)
            return NULL;
        p = child;
    }
    return child;
}

static inline int block_voxel_idx(const or_octant* b, const int v[3])
{
    return (v[0] - b->coord[0]) + (v[1] - b->coord[1]) * 8 + (v[2] - b->coord[2]) * 64;
}

static void push_block(or_map* m, or_octant* b)
{
    if (m->n_blocks == m->cap_blocks) {
        m->cap_blocks *= 2;
        m->blocks = (or_octant**) realloc(m->blocks, sizeof(or_octant*) * m->cap_blocks);
    }
    b->slot = m->n_blocks;
    m->blocks[m->n_blocks++] = b;
}

/* Octree::allocate (octree_impl.hpp:171-199) + aabbExtend (:267-273) */
static int octree_allocate(or_map* m, or_octant* parent, int idx, or_octant** child)
{
    *child = parent->child[idx];
    if (*child)
        return 0;
    const int cs = parent->size / 2;
    const int coord[3] = {parent->coord[0] + cs * ((idx & 1) > 0),
                          parent->coord[1] + cs * ((idx & 2) > 0),
                          parent->coord[2] + cs * ((idx & 4) > 0)};
    if (parent->size == 2 * OR_BLOCK) {
        *child = new_octant(parent, coord, OR_BLOCK, 1);
        push_block(m, *child);
        for (int i = 0; i < 3; i++) {
            if (m->aabb_empty || coord[i] < m->aabb_min[i])
                m->aabb_min[i] = coord[i];
            if (m->aabb_empty || coord[i] + OR_BLOCK - 1 > m->aabb_max[i])
                m->aabb_max[i] = coord[i] + OR_BLOCK - 1;
        }
        if (m->aabb_empty) {
            for (int i = 0; i < 3; i++) {
                m->aabb_min[i] = coord[i];
                m->aabb_max[i] = coord[i] + OR_BLOCK - 1;
            }
        }
        m->aabb_empty = 0;
    }
    else {
        *child = new_octant(parent, coord, cs, 0);
        m->n_nodes++;
    }
    parent->child[idx] = *child;
    parent->child_mask |= (unsigned char) (1u << idx);
    return 1;
}

/* allocator::detail::allocate_key (allocator_impl.hpp:124-147) */
static int allocate_key(or_map* m, uint64_t key, or_octant** out)
{
    or_octant* parent = m->root;
    or_octant* child = NULL;
    int cs = m->max_scale - 1;
    const uint64_t code = key >> OR_SCALE_OFFSET;
    int ks = (int) (key & OR_SCALE_MASK);
    const int min_scale = ks > OR_BLOCK_SCALE ? ks : OR_BLOCK_SCALE;
    int did = 0;
    for (; cs >= min_scale; --cs) {
        did = octree_allocate(m, parent, or_code_to_child_idx(code, cs), &child);
        parent = child;
    }
    *out = child;
    return did;
}

/* allocator::blocks(keys, ...) (allocator_impl.hpp:68-116).  Keys must be unique voxel keys
 * at scale <= block scale.  Returns the allocated (or all, when only_allocated == 0) blocks in
 * sorted key order. */
static long allocator_blocks(or_map* m, uint64_t* keys, long n, int only_allocated, or_octant** out)
{
    uint64_t* tmp = (uint64_t*) malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
    uint64_t* uk = (uint64_t*) malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
    for (int scale = m->max_scale; scale > OR_BLOCK_SCALE; scale--) {
        memcpy(tmp, keys, sizeof(uint64_t) * n);
        const long nu = or_unique_at_scale(tmp, n, scale, uk);
        for (long i = 0; i < nu; i++) {
            or_octant* unused;
            allocate_key(m, uk[i], &unused);
        }
    }
    free(tmp);
    free(uk);
    long k = 0;
    for (long i = 0; i < n; i++) {
        or_octant* b;
        const int did = allocate_key(m, keys[i], &b);
        if (!(only_allocated && !did))
            out[k++] = b;
    }
    return k;
}

/* allocator::blocks(voxel_coords, ...) (allocator_impl.hpp:40-64): returns the block coords of
 * the returned blocks in out_coords (3 ints each), count as return value. */
OR_API long or_allocate_blocks(void* vm, const int* voxel_coords, long n, int only_new, int* out_coords)
{
    or_map* m = (or_map*) vm;
    uint64_t* keys = (uint64_t*) malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
    uint64_t* uk = (uint64_t*) malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
    for (long i = 0; i < n; i++)
        or_encode_key(voxel_coords + 3 * i, OR_BLOCK_SCALE, &keys[i]);
    const long nu = or_unique_keys(keys, n, uk);
    or_octant** bl = (or_octant**) malloc(sizeof(or_octant*) * (nu > 0 ? nu : 1));
    const long k = allocator_blocks(m, uk, nu, only_new, bl);
    for (long i = 0; i < k; i++)
        memcpy(out_coords + 3 * i, bl[i]->coord, sizeof(int) * 3);
    free(keys);
    free(uk);
    free(bl);
    return k;
}

OR_API int or_set_block_data(void* vm, const int block_coord[3], const real* tsdf, const real* weight)
{
    or_map* m = (or_map*) vm;
    or_octant* b = fetch_block(m, block_coord);
    if (!b)
        return -1;
    memcpy(b->tsdf, tsdf, sizeof(real) * 512);
    memcpy(b->weight, weight, sizeof(real) * 512);
    return 0;
}

OR_API int or_map_aabb(void* vm, int out[6])
{
    or_map* m = (or_map*) vm;
    for (int i = 0; i < 3; i++) {
        out[i] = m->aabb_min[i];
        out[3 + i] = m->aabb_max[i];
    }
    return !m->aabb_empty;
}

static int cmp_octant_key(const void* a, const void* b)
{
    const or_octant* x = *(or_octant* const*) a;
    const or_octant* y = *(or_octant* const*) b;
    const uint64_t kx = or_encode_code(x->coord[0], x->coord[1], x->coord[2]);
    const uint64_t ky = or_encode_code(y->coord[0], y->coord[1], y->coord[2]);
    return kx < ky ? -1 : (kx > ky ? 1 : 0);
}

/* Keys of the blocks the last integration updated; keys NULL only counts. */
OR_API long or_map_last_update(void* vm, uint64_t* keys)
{
    or_map* m = (or_map*) vm;
    if (keys && m->n_last_update > 0)
        memcpy(keys, m->last_update, sizeof(uint64_t) * (size_t) m->n_last_update);
    return m->n_last_update;
}

/* Blocks sorted by Morton key; tsdf/weight as 512-arrays per block (x fastest). */
OR_API void or_map_dump_blocks(void* vm, uint64_t* keys, int* timestamps, real* tsdf, real* weight)
{
    or_map* m = (or_map*) vm;
    or_octant** tmp = (or_octant**) malloc(sizeof(or_octant*) * (m->n_blocks + 1));
    memcpy(tmp, m->blocks, sizeof(or_octant*) * m->n_blocks);
    qsort(tmp, (size_t) m->n_blocks, sizeof(or_octant*), cmp_octant_key);
    for (long i = 0; i < m->n_blocks; i++) {
        or_encode_key(tmp[i]->coord, OR_BLOCK_SCALE, &keys[i]);
        if (timestamps)
            timestamps[i] = tmp[i]->timestamp;
        if (tsdf)
            memcpy(tsdf + 512 * i, tmp[i]->tsdf, sizeof(real) * 512);
        if (weight)
            memcpy(weight + 512 * i, tmp[i]->weight, sizeof(real) * 512);
    }
    free(tmp);
}

static void collect_nodes(or_octant* o, int scale, uint64_t* keys, unsigned char* masks, int* ts, long* k)
{
    if (!o || o->is_block)
        return;
    or_encode_key(o->coord, scale, &keys[*k]);
    masks[*k] = o->child_mask;
    if (ts)
        ts[*k] = o->timestamp;
    (*k)++;
    for (int i = 0; i < 8; i++)
        collect_nodes(o->child[i], scale - 1, keys, masks, ts, k);
}

/* All nodes (root included) as (key, child_mask, timestamp), unsorted (DFS order). */
OR_API long or_map_dump_nodes(void* vm, uint64_t* keys, unsigned char* masks, int* timestamps)
{
    or_map* m = (or_map*) vm;
    long k = 0;
    collect_nodes(m->root, m->max_scale, keys, masks, timestamps, &k);
    return k;
}

/* ------------------------------------------------------------------------- */
/* Frustum fetch: raycast_carver_impl.hpp:17-32, iterator.hpp:197-238,        */
/* iterator_impl.hpp:86-156.  Root children are always visited (the root's own */
/* test does not prune), every other octant is pruned by its sphere test.      */
/* ------------------------------------------------------------------------- */
static int octant_ignored(const or_map* m, const or_sensor* sen, const or_iso* T_SW, const or_octant* o)
{
    real cW[3], cS[3];
    for (int i = 0; i < 3; i++) {
        const real pM = ((real) o->coord[i] + (real) o->size * (real) 0.5) * m->res;
        cW[i] = pM + m->t_WM[i];
    }
    iso_apply(T_SW, cW, cS);
    const real r = (real) (sqrtf(3.0f) / 2.0f) * m->res * (real) o->size;
    return !sensor_sphere_in_frustum(sen, cS, r);
}

static long fetch_frustum(const or_map* m, const or_sensor* sen, const or_iso* T_WS, or_octant*** out)
{
    or_iso T_SW;
    iso_inv(T_WS, &T_SW);
    long cap = 1024, n = 0;
    or_octant** list = (or_octant**) malloc(sizeof(or_octant*) * cap);
    long scap = 1024, sn = 0;
    or_octant** stack = (or_octant**) malloc(sizeof(or_octant*) * scap);
    for (int i = 0; i < 8; i++)
        if (m->root->child[i])
            stack[sn++] = m->root->child[i];
    while (sn > 0) {
        or_octant* o = stack[--sn];
        if (octant_ignored(m, sen, &T_SW, o))
            continue;
        if (!o->is_block) {
            for (int i = 0; i < 8; i++) {
                if (o->child[i]) {
                    if (sn == scap) {
                        scap *= 2;
                        stack = (or_octant**) realloc(stack, sizeof(or_octant*) * scap);
                    }
                    stack[sn++] = o->child[i];
                }
            }
        }
        else {
            if (n == cap) {
                cap *= 2;
                list = (or_octant**) realloc(list, sizeof(or_octant*) * cap);
            }
            list[n++] = o;
        }
    }
    free(stack);
    *out = list;
    return n;
}

/* ------------------------------------------------------------------------- */
/* TSDF integration: map_integrator_impl.hpp:52-83, raycast_carver_impl.hpp:55-128, */
/* singleres_tsdf_updater_impl.hpp:14-132, data_field_impl.hpp:39-54,        */
/* propagator_impl.hpp:146-197                                               */
/* ------------------------------------------------------------------------- */

/* The nearest IEEE binary16 value of x (ties to even; overflow to infinity; infinities and NaN
 * kept), as a float: the storage rounding of an FP16 TSDF map.  x = m 2^e with m in [0.5, 1); a
 * normal half keeps 11 significant bits, so the quantum is 2^(e - 11), and 2^-24 below the
 * smallest normal half (2^-14).  Dividing by a power of two is exact and rintf rounds to the
 * nearest integer, ties to even. */
static float half_round(float x)
{
    if (isnan(x) || isinf(x))
        return x;
    const float ax = fabsf(x);
    if (ax >= 65520.f) /* halfway between the largest half (65504) and 2^16 rounds up */
        return x < 0.f ? -INFINITY : INFINITY;
    int e = 0;
    (void) frexpf(ax, &e);
    const int qe = e - 11 < -24 ? -24 : e - 11;
    const float q = ldexpf(1.0f, qe);
    const float r = rintf(ax / q) * q;
    return signbit(x) ? -r : r;
}

OR_API float or_half_round(float x)
{
    return half_round(x);
}

/* One voxel update; returns 1 if the voxel changed. */
static inline int field_update(real* tsdf, real* weight, real sdf, real mu, real max_w)
{
    if (sdf < -mu)
        return 0;
    real u = sdf / mu;
    if (u > (real) 1)
        u = (real) 1;
    if (*weight < max_w)
        *weight = *weight + (real) 1;
    *tsdf = (*tsdf * (*weight - (real) 1) + u) / *weight;
    return 1;
}

static void update_block(const or_map* m, or_octant* b, const or_iso* T_CV, const or_sensor* sen,
                         const float* depth, int ts)
{
    b->timestamp = ts;
    for (int x = 0; x < 8; x++) {
        for (int y = 0; y < 8; y++) {
            for (int z = 0; z < 8; z++) {
                const real v[3] = {(real) (b->coord[0] + x), (real) (b->coord[1] + y), (real) (b->coord[2] + z)};
                real pC[3];
                iso_apply(T_CV, v, pC);
                const real nrm = norm3(pC);
                if (nrm > sensor_far_dist(sen, pC))
                    continue;
                real uv[2];
                if (!sensor_project(sen, pC, uv))
                    continue;
                int px[2];
                round_pixel(uv, px);
                const real d = depth[px[0] + px[1] * sen->w];
                if (d < sen->near_plane)
                    continue;
                /* measurementFromPoint: z for the pinhole, |p| for the lidar */
                const real mz = sensor_measurement(sen, pC);
                const real sdf = (d - mz) / mz * norm3(pC);
                const int vi = x + 8 * y + 64 * z;
                if (field_update(&b->tsdf[vi], &b->weight[vi], sdf, m->mu, m->max_weight) && m->storage_f16)
                    b->tsdf[vi] = (real) half_round((float) b->tsdf[vi]);
            }
        }
    }
}

/* The allocation band march (raycast_carver_impl.hpp:74-111) for one pixel.  Writes the
 * keys of unallocated blocks hit by the samples into keys (at most num_steps) and returns
 * their count. */
static int carve_pixel(const or_map* m, const or_sensor* sen, const or_iso* T_WS, const float* depth,
                       int x, int y, real band, int num_steps, uint64_t* keys)
{
    const real d = depth[x + y * sen->w];
    if (d < sen->near_plane || d > (sen->far_plane + band * (real) 0.5))
        return 0;
    real dir[3], pC[3], pW[3];
    sensor_backproject(sen, x, y, dir);
    for (int i = 0; i < 3; i++)
        pC[i] = d * dir[i];
    iso_apply(T_WS, pC, pW);
    real rev[3], diff[3];
    for (int i = 0; i < 3; i++)
        diff[i] = T_WS->t[i] - pW[i];
    normalized3(diff, rev);
    real pos[3], step[3];
    for (int i = 0; i < 3; i++) {
        pos[i] = pW[i] - (band * (real) 0.5) * rev[i];
        step[i] = (rev[i] * band) / (real) num_steps;
    }
    int k = 0;
    for (int s = 0; s < num_steps; s++) {
        int vc[3];
        if (map_point_to_voxel(m, pos, vc)) {
            if (fetch_block(m, vc) == NULL)
                or_encode_key(vc, OR_BLOCK_SCALE, &keys[k++]);
        }
        for (int i = 0; i < 3; i++)
            pos[i] = pos[i] + step[i];
    }
    return k;
}

/* T_CV = T_CW * T_WM * Scaling(res) * Translation(0.5) (singleres_tsdf_updater_impl.hpp:37-42) */
static void compute_T_CV(const or_map* m, const or_iso* T_WC, or_iso* T_CV)
{
    or_iso T_CW, T_WM, A;
    iso_inv(T_WC, &T_CW);
    memset(&T_WM, 0, sizeof(T_WM));
    for (int i = 0; i < 3; i++) {
        T_WM.R[i][i] = 1;
        T_WM.t[i] = m->t_WM[i];
    }
    iso_mul(&T_CW, &T_WM, &A);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            A.R[i][j] = A.R[i][j] * m->res;
    const real h[3] = {(real) 0.5, (real) 0.5, (real) 0.5};
    for (int i = 0; i < 3; i++)
        A.t[i] = A.t[i] + (A.R[i][0] * h[0] + (A.R[i][1] * h[1] + A.R[i][2] * h[2]));
    *T_CV = A;
}

OR_API void or_compute_T_CV(void* vm, const real T_WC16[16], real out16[16])
{
    or_iso T_WC, T_CV;
    iso_from16(T_WC16, &T_WC);
    compute_T_CV((or_map*) vm, &T_WC, &T_CV);
    iso_to16(&T_CV, out16);
}

/* ------------------------------------------------------------------------- */
/* Multi-resolution TSDF updater: multires_tsdf_updater_impl.hpp:14-254,     */
/* PinholeCamera::computeIntegrationScaleImpl (pinhole_camera.cpp:115-152),   */
/* propagateBlockDown / propagateBlockUp (propagator_impl.hpp:18-117),        */
/* up_prop_mean (data_impl.hpp:14-53), multi-resolution getFieldInterp        */
/* (visitor_impl.hpp:399-460, 905-971).  The reference's OpenMP loop lets a   */
/* block's interpolation read a neighbour updated in the same frame, so its   */
/* result depends on the schedule; this restatement takes the schedule in     */
/* which every scale choice and down-propagation reads the map as it was      */
/* before the frame's updates, then every block updates and propagates up.    */
/* ------------------------------------------------------------------------- */
static const int or_mr_offsets[4] = {0, 512, 576, 584};

static inline int mr_idx(int x, int y, int z, int s)
{
    const int n = 8 >> s;
    return or_mr_offsets[s] + (x >> s) + (y >> s) * n + (z >> s) * n * n;
}

/* entry i: data {mr[2i], mr[2i+1]}, past data {mr[1170 + 2i], mr[1171 + 2i]} */
static void mr_init(or_octant* b)
{
    if (b->mr)
        return;
    b->mr = (real*) malloc(sizeof(real) * 4 * 585);
    for (int i = 0; i < 2 * 585; i++) {
        b->mr[2 * i] = 1;
        b->mr[2 * i + 1] = 0;
    }
    b->cur_scale = -1;
    b->min_scale = -1;
}

/* pixel_voxel_ratio_per_scale = {1.5, 3, 6} (sensor.hpp:56) */
static int mr_scale_from_z(real z, real scaled_pixel, real res)
{
    const real pv = z * scaled_pixel / res;
    int scale = 0;
    const real ratios[3] = {(real) 1.5, (real) 3, (real) 6};
    for (int i = 0; i < 3; i++) {
        if (pv < ratios[i])
            break;
        scale++;
    }
    return scale < 3 ? scale : 3;
}

static int mr_integration_scale(real z, real scaled_pixel, real res, int last, int min_scale)
{
    const int scale = mr_scale_from_z(z, scaled_pixel, res);
    if (scale > last && min_scale != -1)
        return mr_scale_from_z(z - (real) 0.25, scaled_pixel, res);
    if (scale < last && min_scale != -1)
        return mr_scale_from_z(z + (real) 0.25, scaled_pixel, res);
    return scale;
}

static int mr_interp(const or_map* m, const real cf[3], int desired, real* out)
{
    const int ic[3] = {(int) cf[0], (int) cf[1], (int) cf[2]};
    const or_octant* leaf = fetch_block(m, ic);
    const int leaf_scale = leaf ? (leaf->mr ? leaf->cur_scale : -1) : 0;
    const int init = leaf ? (leaf_scale > desired ? leaf_scale : desired) : 0;
    for (int scale = init; scale <= 3; scale++) {
        const int stride = 1 << scale;
        const real inv = (real) 1 / (real) stride;
        real bf[3];
        int bi[3];
        int in_box = !m->aabb_empty, in_map = 1;
        for (int i = 0; i < 3; i++) {
            bf[i] = inv * cf[i] - (real) 0.5;
            bi[i] = stride * (int) bf[i];
            in_box = in_box && m->aabb_min[i] <= bi[i] && bi[i] <= m->aabb_max[i];
            in_map = in_map && bi[i] >= 0 && bi[i] + stride < m->size;
        }
        if (!in_box)
            return 0;
        if (!in_map)
            continue;
        real d[8];
        int coarser = 0, valid = 1;
        for (int k = 0; k < 8; k++) {
            const int c[3] = {bi[0] + stride * (k & 1), bi[1] + stride * ((k >> 1) & 1), bi[2] + stride * ((k >> 2) & 1)};
            const or_octant* nb = fetch_block(m, c);
            d[k] = 1;
            if (!nb) {
                valid = 0;
                continue;
            }
            if (nb->mr && nb->cur_scale > scale) {
                coarser = 1;
                continue;
            }
            if (!nb->mr) {
                valid = 0;
                continue;
            }
            const int e = mr_idx(c[0] - nb->coord[0], c[1] - nb->coord[1], c[2] - nb->coord[2], scale);
            d[k] = nb->mr[2 * e];
            valid = valid && nb->mr[2 * e + 1] != 0;
        }
        if (coarser)
            continue;
        if (!valid)
            return 0;
        real t[3], tc[3];
        for (int i = 0; i < 3; i++) {
            t[i] = bf[i] - floor(bf[i]);
            tc[i] = (real) 1 - t[i];
        }
        *out = ((d[0] * tc[0] + d[1] * t[0]) * tc[1] + (d[2] * tc[0] + d[3] * t[0]) * t[1]) * tc[2] +
               ((d[4] * tc[0] + d[5] * t[0]) * tc[1] + (d[6] * tc[0] + d[7] * t[0]) * t[1]) * t[2];
        return 1;
    }
    return 0;
}

/* scale choice and down-propagation (reads only pre-frame data of other blocks) */
static void mr_scale_down(const or_map* m, or_octant* b, const or_iso* T_CW, real scaled_pixel)
{
    real cW[3], cC[3];
    for (int i = 0; i < 3; i++)
        cW[i] = ((real) b->coord[i] + (real) 4) * m->res - m->t_MW[i];
    iso_apply(T_CW, cW, cC);
    const int last = b->cur_scale;
    int s = mr_integration_scale(cC[2], scaled_pixel, m->res, last, b->min_scale);
    if (s < last - 1)
        s = last - 1;
    b->new_scale = s;
    real* data = b->mr;
    real* past = b->mr + 2 * 585;
    for (int ps = last; ps > s; ps--) {
        const int cs = ps - 1, cn = 8 >> cs, cstride = 1 << cs;
        for (int i = 0; i < cn * cn * cn; i++) {
            const int lx = (i % cn) * cstride, ly = ((i / cn) % cn) * cstride, lz = (i / (cn * cn)) * cstride;
            const int ci = mr_idx(lx, ly, lz, cs), pi = mr_idx(lx, ly, lz, ps);
            if (data[2 * ci + 1] != 0) {
                const real v = data[2 * ci] + (data[2 * pi] - past[2 * pi]);
                data[2 * ci] = v < (real) -1 ? (real) -1 : v; /* std::max(v, -1) keeps a NaN */
                data[2 * ci + 1] = fmin(data[2 * ci + 1] + past[2 * pi + 1], m->max_weight);
                past[2 * ci + 1] = past[2 * pi + 1];
            }
            else {
                const real cf[3] = {(real) (b->coord[0] + lx) + (real) 0.5 * (real) cstride,
                                    (real) (b->coord[1] + ly) + (real) 0.5 * (real) cstride,
                                    (real) (b->coord[2] + lz) + (real) 0.5 * (real) cstride};
                real f;
                if (mr_interp(m, cf, cs, &f)) {
                    data[2 * ci] = f;
                    data[2 * ci + 1] = data[2 * pi + 1];
                    past[2 * ci] = f;
                    past[2 * ci + 1] = 0;
                }
            }
        }
        const int pn = 8 >> ps;
        for (int i = 0; i < pn * pn * pn; i++) {
            const int pi = or_mr_offsets[ps] + i;
            past[2 * pi] = data[2 * pi];
            past[2 * pi + 1] = 0;
        }
    }
}

/* commit the scale, update the voxels of that scale, up-propagate */
static void mr_update_up(const or_map* m, or_octant* b, const or_iso* T_CV, const or_sensor* sen, const float* depth,
                         int ts)
{
    b->timestamp = ts;
    const int curr = b->new_scale;
    b->cur_scale = curr;
    b->min_scale = b->min_scale < 0 ? curr : (b->min_scale < curr ? b->min_scale : curr);
    real* data = b->mr;
    real* past = b->mr + 2 * 585;
    const int stride = 1 << curr;
    for (int x = 0; x < 8; x += stride)
        for (int y = 0; y < 8; y += stride)
            for (int z = 0; z < 8; z += stride) {
                const real v[3] = {(real) (b->coord[0] + x), (real) (b->coord[1] + y), (real) (b->coord[2] + z)};
                real pC[3];
                iso_apply(T_CV, v, pC);
                if (norm3(pC) > sensor_far_dist(sen, pC))
                    continue;
                real uv[2];
                if (!sensor_project(sen, pC, uv))
                    continue;
                int px[2];
                round_pixel(uv, px);
                const real d = depth[px[0] + px[1] * sen->w];
                if (d < sen->near_plane)
                    continue;
                const real mz = sensor_measurement(sen, pC);
                const real sdf = (d - mz) / mz * norm3(pC);
                const int vi = mr_idx(x, y, z, curr);
                field_update(&data[2 * vi], &data[2 * vi + 1], sdf, m->mu, m->max_weight);
                past[2 * vi + 1] = past[2 * vi + 1] + (real) 1;
            }
    for (int cs = curr; cs < 3; cs++) {
        const int ps = cs + 1, pn = 8 >> ps, cstride = 1 << cs;
        for (int i = 0; i < pn * pn * pn; i++) {
            const int x0 = (i % pn) << ps, y0 = ((i / pn) % pn) << ps, z0 = (i / (pn * pn)) << ps;
            real tsum = 0, wsum = 0;
            int nv = 0;
            for (int k = 0; k < 8; k++) {
                const int ci = mr_idx(x0 + cstride * (k & 1), y0 + cstride * ((k >> 1) & 1), z0 + cstride * ((k >> 2) & 1), cs);
                if (data[2 * ci + 1] != 0) {
                    tsum += data[2 * ci];
                    wsum += data[2 * ci + 1];
                    nv++;
                }
            }
            const int pi = mr_idx(x0, y0, z0, ps);
            if (nv > 0) {
                data[2 * pi] = tsum / (real) nv;
                data[2 * pi + 1] = ceil(wsum / (real) nv);
                past[2 * pi] = data[2 * pi];
                past[2 * pi + 1] = 0;
            }
            else {
                past[2 * pi] = 1;
                past[2 * pi + 1] = 0;
            }
        }
    }
}

/* Sets an allocated block's multi-resolution state (test infrastructure): current and minimum
 * scale, 585 x {tsdf, weight} data and past data as or_map_dump_multires returns them. */
OR_API int or_set_multires_block(void* vm, const int block_coord[3], int cur_scale, int min_scale, const real* data,
                                 const real* past)
{
    or_map* m = (or_map*) vm;
    or_octant* b = fetch_block(m, block_coord);
    if (!b || !m->multires)
        return -1;
    mr_init(b);
    b->cur_scale = cur_scale;
    b->min_scale = min_scale;
    memcpy(b->mr, data, sizeof(real) * 2 * 585);
    memcpy(b->mr + 2 * 585, past, sizeof(real) * 2 * 585);
    return 0;
}

OR_API void or_map_set_storage_f16(void* vm)
{
    ((or_map*) vm)->storage_f16 = 1;
}

OR_API void or_map_set_multires(void* vm)
{
    ((or_map*) vm)->multires = 1;
}

/* blocks sorted by key: scales (current, minimum), 585 x {tsdf, weight} data and past data */
OR_API void or_map_dump_multires(void* vm, uint64_t* keys, int* scales, real* data, real* past)
{
    or_map* m = (or_map*) vm;
    or_octant** tmp = (or_octant**) malloc(sizeof(or_octant*) * (m->n_blocks + 1));
    memcpy(tmp, m->blocks, sizeof(or_octant*) * m->n_blocks);
    qsort(tmp, (size_t) m->n_blocks, sizeof(or_octant*), cmp_octant_key);
    for (long i = 0; i < m->n_blocks; i++) {
        or_octant* b = tmp[i];
        mr_init(b);
        or_encode_key(b->coord, OR_BLOCK_SCALE, &keys[i]);
        scales[2 * i] = b->cur_scale;
        scales[2 * i + 1] = b->min_scale;
        memcpy(data + 2 * 585 * i, b->mr, sizeof(real) * 2 * 585);
        memcpy(past + 2 * 585 * i, b->mr + 2 * 585, sizeof(real) * 2 * 585);
    }
    free(tmp);
}

/* stats: [0] frustum blocks, [1] new blocks, [2] candidate keys (with per-pixel repeats),
 *        [3] valid pixels */
static void integrate_sensor(or_map* m, int ts, const float* depth, const or_sensor* sen, const real T_WC16[16],
                             long stats[4])
{
    or_iso T_WC;
    iso_from16(T_WC16, &T_WC);
    /* band_ = 2 * res * truncation_boundary_factor (raycast_carver_impl.hpp:50) */
    const real band = (real) 2 * m->res * m->mu_factor;

    /* fetch-frustum (before any allocation) */
    or_octant** frustum = NULL;
    const long n_frustum = fetch_frustum(m, sen, &T_WC, &frustum);

    /* create-list */
    const int num_steps = (int) ceil((double) (band / ((real) 2 * m->res)));
    const long npx = (long) sen->w * sen->h;
    uint64_t* cand = (uint64_t*) malloc(sizeof(uint64_t) * (size_t) npx * (size_t) num_steps);
    int* cnt = (int*) calloc((size_t) npx, sizeof(int));
    long valid = 0;
#pragma omp parallel for reduction(+ : valid) schedule(static)
    for (int x = 0; x < sen->w; x++) {
        for (int y = 0; y < sen->h; y++) {
            const long p = x + (long) y * sen->w;
            const real d = depth[p];
            if (!(d < sen->near_plane || d > (sen->far_plane + band * (real) 0.5)))
                valid++;
            cnt[p] = carve_pixel(m, sen, &T_WC, depth, x, y, band, num_steps, cand + p * num_steps);
        }
    }
    long nk = 0;
    for (long p = 0; p < npx; p++)
        for (int s = 0; s < cnt[p]; s++)
            cand[nk++] = cand[p * num_steps + s];
    uint64_t* uk = (uint64_t*) malloc(sizeof(uint64_t) * (nk > 0 ? nk : 1));
    const long nu = or_unique_keys(cand, nk, uk);
    free(cand);
    free(cnt);

    /* allocate-list */
    or_octant** newb = (or_octant**) malloc(sizeof(or_octant*) * (nu > 0 ? nu : 1));
    const long n_new = allocator_blocks(m, uk, nu, 1, newb);
    free(uk);

    /* update over new + frustum blocks */
    const long nb = n_new + n_frustum;
    or_octant** list = (or_octant**) malloc(sizeof(or_octant*) * (nb > 0 ? nb : 1));
    memcpy(list, newb, sizeof(or_octant*) * n_new);
    memcpy(list + n_new, frustum, sizeof(or_octant*) * n_frustum);
    or_iso T_CV;
    compute_T_CV(m, &T_WC, &T_CV);
    if (m->multires && sen->cam) {
        or_iso T_CW;
        iso_inv(&T_WC, &T_CW);
        for (long i = 0; i < nb; i++)
            mr_init(list[i]);
#pragma omp parallel for schedule(dynamic, 16)
        for (long i = 0; i < nb; i++)
            mr_scale_down(m, list[i], &T_CW, sen->cam->inv_fx);
#pragma omp parallel for schedule(dynamic, 16)
        for (long i = 0; i < nb; i++)
            mr_update_up(m, list[i], &T_CV, sen, depth, ts);
    }
    else {
#pragma omp parallel for schedule(dynamic, 16)
        for (long i = 0; i < nb; i++)
            update_block(m, list[i], &T_CV, sen, depth, ts);
    }
    /* the updated_octants integrateDepth reports (map_integrator_impl.hpp:77-80) */
    free(m->last_update);
    m->last_update = (uint64_t*) malloc(sizeof(uint64_t) * (nb > 0 ? nb : 1));
    m->n_last_update = nb;
    for (long i = 0; i < nb; i++)
        or_encode_key(list[i]->coord, OR_BLOCK_SCALE, &m->last_update[i]);
    /* propagateTimeStampToRoot: only the blocks' direct parents are reached */
    for (long i = 0; i < nb; i++) {
        or_octant* p = list[i]->parent;
        if (list[i]->timestamp > p->timestamp)
            p->timestamp = list[i]->timestamp;
    }
    if (stats) {
        stats[0] = n_frustum;
        stats[1] = n_new;
        stats[2] = nk;
        stats[3] = valid;
    }
    free(list);
    free(newb);
    free(frustum);
}

OR_API void or_integrate(void* vm, int ts, const float* depth, const or_pinhole* pin, const real T_WC16[16],
                         long stats[4])
{
    or_cam cam;
    cam_init(&cam, pin, 1);
    const or_sensor sen = {&cam, NULL, cam.w, cam.h, cam.near_plane, cam.far_plane};
    integrate_sensor((or_map*) vm, ts, depth, &sen, T_WC16, stats);
}

/* integrateDepth with an OusterLidar: range image (H rows of W columns, metres) */
OR_API void or_integrate_lidar(void* vm, int ts, const float* range, void* vl, const real T_WS16[16], long stats[4])
{
    const or_lidar* L = (const or_lidar*) vl;
    const or_sensor sen = {NULL, L, L->w, L->h, L->near_plane, L->far_plane};
    integrate_sensor((or_map*) vm, ts, range, &sen, T_WS16, stats);
}

/* ------------------------------------------------------------------------- */
/* Queries: visitor_impl.hpp:67-302,704-737,841-864,976-1024,1094-1251,       */
/* map_impl.hpp:68-86,154-172,239-268                                        */
/* ------------------------------------------------------------------------- */
static inline int get_voxel(const or_map* m, const int v[3], real* f)
{
    /* getData via fetcher::leaf: data of the block if allocated, otherwise invalid */
    if (v[0] < 0 || v[1] < 0 || v[2] < 0 || v[0] >= m->size || v[1] >= m->size || v[2] >= m->size)
        return 0;
    const or_octant* b = fetch_block(m, v);
    if (!b)
        return 0;
    const int i = block_voxel_idx(b, v);
    touch_voxel(m, b, i);
    if (b->weight[i] == (real) 0)
        return 0;
    *f = b->tsdf[i];
    return 1;
}

static int field_interp_vox(const or_map* m, const real vf[3], real* out)
{
    real bf[3];
    int base[3];
    for (int i = 0; i < 3; i++) {
        bf[i] = vf[i] - (real) 0.5;
        base[i] = (int) bf[i];
    }
    if (m->aabb_empty)
        return 0;
    for (int i = 0; i < 3; i++)
        if (base[i] < m->aabb_min[i] || base[i] > m->aabb_max[i])
            return 0;
    for (int i = 0; i < 3; i++)
        if (base[i] + 1 >= m->size || base[i] < 0)
            return 0;
    real d[8];
    for (int k = 0; k < 8; k++) {
        const int c[3] = {base[0] + (k & 1), base[1] + ((k >> 1) & 1), base[2] + ((k >> 2) & 1)};
        const or_octant* b = fetch_block(m, c);
        if (!b)
            return 0;
        const int i = block_voxel_idx(b, c);
        touch_voxel(m, b, i);
        if (b->weight[i] == (real) 0)
            return 0;
        d[k] = b->tsdf[i];
    }
    real t[3], tc[3];
    for (int i = 0; i < 3; i++) {
        t[i] = bf[i] - floor(bf[i]);
        tc[i] = (real) 1 - t[i];
    }
    *out = ((d[0] * tc[0] + d[1] * t[0]) * tc[1] + (d[2] * tc[0] + d[3] * t[0]) * t[1]) * tc[2]
        + ((d[4] * tc[0] + d[5] * t[0]) * tc[1] + (d[6] * tc[0] + d[7] * t[0]) * t[1]) * t[2];
    return 1;
}

OR_API int or_get_field_interp(void* vm, const real pW[3], real* out)
{
    const or_map* m = (const or_map*) vm;
    real vf[3];
    if (!map_point_to_voxel_f(m, pW, vf))
        return 0;
    return field_interp_vox(m, vf, out);
}

/* Single-res getFieldGrad: trilinear blend of central differences at the 8 voxels around the
 * query, all 32 samples must be valid (visitor_impl.hpp:1094-1251), then x0.5 and /res
 * (map_impl.hpp:264-266). */
static int field_grad_vox(const or_map* m, const real vf[3], real g[3])
{
    real s[3], fac[3];
    int base[3];
    for (int i = 0; i < 3; i++) {
        s[i] = vf[i] - (real) 0.5;
        fac[i] = s[i] - floor(s[i]);
        base[i] = (int) s[i];
    }
    int lo[3][4]; /* per axis: b-1, b, b+1, b+2 clamped */
    for (int i = 0; i < 3; i++) {
        lo[i][0] = base[i] - 1 > 0 ? base[i] - 1 : 0;
        lo[i][1] = base[i] > 0 ? base[i] : 0;
        lo[i][2] = base[i] + 1 < m->size - 1 ? base[i] + 1 : m->size - 1;
        lo[i][3] = base[i] + 2 < m->size - 1 ? base[i] + 2 : m->size - 1;
    }
    if (!fetch_block(m, base))
        return 0;
    /* central difference along axis a at corner (cx,cy,cz) in {0,1}^3 */
    real cd[3][8];
    for (int a = 0; a < 3; a++) {
        for (int k = 0; k < 8; k++) {
            const int c[3] = {k & 1, (k >> 1) & 1, (k >> 2) & 1};
            int vp[3], vm_[3];
            for (int i = 0; i < 3; i++) {
                vp[i] = lo[i][1 + c[i]];
                vm_[i] = lo[i][1 + c[i]];
            }
            vp[a] = lo[a][2 + c[a]];
            vm_[a] = lo[a][c[a]];
            real fp, fm;
            if (!get_voxel(m, vp, &fp) || !get_voxel(m, vm_, &fm))
                return 0;
            cd[a][k] = fp - fm;
        }
    }
    const real rx = (real) 1 - fac[0], ry = (real) 1 - fac[1], rz = (real) 1 - fac[2];
    for (int a = 0; a < 3; a++) {
        const real* c = cd[a];
        const real v = ((c[0] * rx + c[1] * fac[0]) * ry + (c[2] * rx + c[3] * fac[0]) * fac[1]) * rz
            + ((c[4] * rx + c[5] * fac[0]) * ry + (c[6] * rx + c[7] * fac[0]) * fac[1]) * fac[2];
        g[a] = (real) 0.5 * v;
    }
    return 1;
}

OR_API int or_get_field_grad(void* vm, const real pW[3], real out[3])
{
    const or_map* m = (const or_map*) vm;
    real vf[3];
    if (!map_point_to_voxel_f(m, pW, vf))
        return 0;
    if (!field_grad_vox(m, vf, out))
        return 0;
    for (int i = 0; i < 3; i++)
        out[i] = out[i] / m->res;
    return 1;
}

OR_API int or_get_data(void* vm, const real pW[3], real* tsdf, real* weight)
{
    const or_map* m = (const or_map*) vm;
    int v[3];
    *tsdf = 1;
    *weight = 0;
    if (!map_point_to_voxel(m, pW, v))
        return 0;
    const or_octant* b = fetch_block(m, v);
    if (!b)
        return 0;
    const int i = block_voxel_idx(b, v);
    touch_voxel(m, b, i);
    *tsdf = b->tsdf[i];
    *weight = b->weight[i];
    return 1;
}

/* Distinct voxel reads of the raycast (the algorithmic bytes of raycast_volume): between begin and
 * end, every voxel whose data a query reads is marked; end returns the number of distinct voxels
 * (out[0]) and of distinct blocks with at least one of them (out[1]). */
OR_API void or_map_touch_begin(void* vm)
{
    or_map* m = (or_map*) vm;
    free(m->touch);
    m->touch_blocks = m->n_blocks;
    m->touch = (uint64_t*) calloc((size_t) (m->n_blocks > 0 ? m->n_blocks : 1) * 8, sizeof(uint64_t));
}

OR_API void or_map_touch_end(void* vm, long out[2], uint64_t* bits)
{
    or_map* m = (or_map*) vm;
    long nv = 0, nb = 0;
    if (bits && m->touch)
        memcpy(bits, m->touch, sizeof(uint64_t) * 8 * (size_t) m->touch_blocks);
    for (long b = 0; m->touch && b < m->touch_blocks; b++) {
        long c = 0;
        for (int w = 0; w < 8; w++)
            c += __builtin_popcountll(m->touch[b * 8 + w]);
        nv += c;
        nb += c > 0;
    }
    free(m->touch);
    m->touch = NULL;
    out[0] = nv;
    out[1] = nb;
}

/* ------------------------------------------------------------------------- */
/* First allocated block along a ray, in the scaled [1,2]^3 octree space of   */
/* voxel_block_ray_iterator_impl.hpp:19-249 (ESVO, Laine & Karras 2010).  The  */
/* traversal runs in fp32 in both builds; only t_min/t_max leave it.          */
/* ------------------------------------------------------------------------- */
#define OR_STACK 23

static inline int f2i(float f)
{
    int i;
    memcpy(&i, &f, 4);
    return i;
}

static inline float i2f(int i)
{
    float f;
    memcpy(&f, &i, 4);
    return f;
}

typedef struct {
    float t_min_ray, t_max_ray; /* tcmin, tmax in metres */
} or_ray_span;

static void esvo_first_block(const or_map* m, const float o_W[3], const float d_W[3], float near_t, float far_t,
                             or_ray_span* out)
{
    const int size = m->size;
    const float scaling = (float) size * (float) m->res;
    int lg = 0;
    while ((1 << lg) < size / OR_BLOCK)
        lg++;
    const int min_scale = OR_STACK - lg;
    int lgs = 0;
    while ((1 << lgs) < size)
        lgs++;
    const float eps = exp2f((float) -lgs);
    float d[3], o[3], tc[3], tb[3];
    for (int i = 0; i < 3; i++) {
        d[i] = fabsf(d_W[i]) < eps ? copysignf(eps, d_W[i]) : d_W[i];
        /* pointToVoxel<Safe::Off> then /size + 1 */
        const float pM = o_W[i] + (float) m->t_MW[i];
        const float vox = pM / (float) m->res;
        o[i] = vox / (float) size + 1.0f;
        tc[i] = -1.f * (1.0f / fabsf(d[i]));
        tb[i] = tc[i] * o[i];
    }
    int octant_mask = 7;
    for (int i = 0; i < 3; i++) {
        if (d[i] > 0.0f) {
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
    for (int i = 0; i < 3; i++) {
        if (1.5f * tc[i] - tb[i] > t_min) {
            idx ^= 1 << i;
            pos[i] = 1.5f;
        }
    }
    const or_octant* stack_parent[OR_STACK];
    float stack_tmax[OR_STACK];
    memset(stack_parent, 0, sizeof(stack_parent));
    memset(stack_tmax, 0, sizeof(stack_tmax));
    const or_octant* parent = m->root;
    int scale = OR_STACK - 1;
    float scale_exp2 = 0.5f;
    float tcorner[3];
    float tc_max = 0.f;
    while (scale < OR_STACK) {
        if (or_cur_stats)
            or_cur_stats[0]++;
        for (int i = 0; i < 3; i++)
            tcorner[i] = pos[i] * tc[i] - tb[i];
        tc_max = fminf(fminf(tcorner[0], tcorner[1]), tcorner[2]);
        const or_octant* child = parent->child[idx ^ octant_mask ^ 7];
        if (scale == min_scale && child != NULL)
            break; /* found the first allocated block */
        if (child != NULL && t_min <= t_max) {
            /* descend */
            const float tv_max = fminf(t_max, tc_max);
            const float half = scale_exp2 * 0.5f;
            float tcen[3];
            for (int i = 0; i < 3; i++)
                tcen[i] = half * tc[i] + tcorner[i];
            if (tc_max < h) {
                stack_parent[scale] = parent;
                stack_tmax[scale] = t_max;
            }
            h = tc_max;
            parent = child;
            idx = 0;
            scale--;
            scale_exp2 = half;
            for (int i = 0; i < 3; i++) {
                if (tcen[i] > t_min)
                    idx ^= 1 << i;
            }
            for (int i = 0; i < 3; i++)
                pos[i] = pos[i] + scale_exp2 * (float) ((idx >> i) & 1);
            t_max = tv_max;
            continue;
        }
        /* advance */
        const int step_mask = (tcorner[0] <= tc_max) | ((tcorner[1] <= tc_max) << 1) | ((tcorner[2] <= tc_max) << 2);
        for (int i = 0; i < 3; i++)
            pos[i] = pos[i] - scale_exp2 * (float) ((step_mask >> i) & 1);
        t_min = tc_max;
        idx ^= step_mask;
        if ((idx & step_mask) != 0) {
            unsigned int diff = 0;
            for (int i = 0; i < 3; i++)
                if (step_mask & (1 << i))
                    diff |= (unsigned) (f2i(pos[i]) ^ f2i(pos[i] + scale_exp2));
            scale = (f2i((float) diff) >> 23) - 127;
            if (scale >= OR_STACK)
                break; /* left the root: no block */
            scale_exp2 = i2f((scale - OR_STACK + 127) << 23);
            parent = stack_parent[scale];
            t_max = stack_tmax[scale];
            int sh[3];
            for (int i = 0; i < 3; i++) {
                sh[i] = f2i(pos[i]) >> scale;
                pos[i] = i2f(sh[i] << scale);
            }
            idx = (sh[0] & 1) | ((sh[1] & 1) << 1) | ((sh[2] & 1) << 2);
            h = 0.0f;
        }
    }
    out->t_min_ray = t_min * scaling;
    out->t_max_ray = tmax_init * scaling;
}

/* ------------------------------------------------------------------------- */
/* TSDF raycast: raycaster_impl.hpp:456-627                                  */
/* ------------------------------------------------------------------------- */
static int raycast_ray_bracket(const or_map* m, const real o[3], const real d[3], real t_near, real t_far,
                               real hit[3], real* t_prev, real* t_hit);

static int raycast_ray(const or_map* m, const real o[3], const real d[3], real t_near, real t_far, real hit[3])
{
    return raycast_ray_bracket(m, o, d, t_near, t_far, hit, NULL, NULL);
}

/* The march; on a hit also returns the bracket [t_prev, t_hit] of the last two samples, whose
 * field values straddle zero. */
static int raycast_ray_bracket(const or_map* m, const real o[3], const real d[3], real t_near, real t_far,
                               real hit[3], real* t_prev, real* t_hit)
{
    or_ray_span span;
    const float of[3] = {(float) o[0], (float) o[1], (float) o[2]};
    const float df[3] = {(float) d[0], (float) d[1], (float) d[2]};
    esvo_first_block(m, of, df, (float) t_near, (float) t_far, &span);
    const real step = m->res;
    const real largestep = (real) 8 * step;
    const real t_min = span.t_min_ray;
    if (t_min <= (real) 0)
        return 0;
    const real t_max = span.t_max_ray;
    real t = t_min;
    const real mu = m->mu;
    if (!(t_near < t_max))
        return 0;
    real stepsize = largestep;
    real p[3];
    for (int i = 0; i < 3; i++)
        p[i] = o[i] + d[i] * t;
    real f_t = 1, f_tt = 0;
    {
        real tsdf, w;
        or_get_data((void*) m, p, &tsdf, &w);
        f_t = tsdf;
    }
    if (!(f_t >= (real) 0))
        return 0;
    for (; t < t_far; t += stepsize) {
        real tsdf, w;
        const int allocated = or_get_data((void*) m, p, &tsdf, &w);
        if (or_cur_stats) {
            int v[3];
            int cat = 4;
            if (!allocated) {
                cat = 2;
                if (map_point_to_voxel(m, p, v)) {
                    /* descend to the brick octant (scale 6, or the root for smaller maps) */
                    const or_octant* o = m->root;
                    const int bsize = m->size < 64 ? m->size : 64;
                    while (o && o->size > bsize) {
                        const int half = o->size / 2;
                        const int ci = ((v[0] & half) ? 1 : 0) | ((v[1] & half) ? 2 : 0) | ((v[2] & half) ? 4 : 0);
                        o = o->child[ci];
                    }
                    if (!o)
                        cat = 1;
                    else {
                        /* present brick: is the 32-voxel octant (scale 5) absent too? */
                        while (o && o->size > 32) {
                            const int half = o->size / 2;
                            const int ci = ((v[0] & half) ? 1 : 0) | ((v[1] & half) ? 2 : 0) | ((v[2] & half) ? 4 : 0);
                            o = o->child[ci];
                        }
                        if (!o)
                            or_cur_stats[6]++;
                    }
                }
            }
            else if (w == (real) 0)
                cat = 3;
            or_cur_stats[cat]++;
        }
        if (w == (real) 0) {
            stepsize = largestep;
            for (int i = 0; i < 3; i++)
                p[i] = p[i] + stepsize * d[i];
            continue;
        }
        f_tt = tsdf;
        if (f_tt <= (real) 0.1 && f_tt >= (real) -0.5) {
            real fi;
            if (or_cur_stats)
                or_cur_stats[5]++;
            if (or_get_field_interp((void*) m, p, &fi))
                f_tt = fi;
        }
        if (f_tt < (real) 0)
            break;
        /* std::max(f_tt * mu, step): a NaN first argument is returned */
        stepsize = f_tt * mu < step ? step : f_tt * mu;
        for (int i = 0; i < 3; i++)
            p[i] = p[i] + stepsize * d[i];
        f_t = f_tt;
    }
    if (f_tt < (real) 0) {
        if (t_prev)
            *t_prev = t - stepsize;
        if (t_hit)
            *t_hit = t;
        t = t - stepsize * (f_tt - (real) 0) / (f_tt - f_t);
        for (int i = 0; i < 3; i++)
            hit[i] = o[i] + d[i] * t;
        return 1;
    }
    return 0;
}

static void raycast_impl(void* vm, const or_pinhole* pin, const real T_WS16[16], real* points, real* normals,
                         signed char* scale, int* stats);

OR_API void or_raycast(void* vm, const or_pinhole* pin, const real T_WS16[16], real* points, real* normals,
                       signed char* scale)
{
    raycast_impl(vm, pin, T_WS16, points, normals, scale, NULL);
}

/* or_raycast with the per-ray march statistics above (7 ints per pixel) */
OR_API void or_raycast_stats(void* vm, const or_pinhole* pin, const real T_WS16[16], real* points, real* normals,
                             int* stats)
{
    raycast_impl(vm, pin, T_WS16, points, normals, NULL, stats);
}

static void raycast_impl(void* vm, const or_pinhole* pin, const real T_WS16[16], real* points, real* normals,
                         signed char* scale, int* stats)
{
    const or_map* m = (const or_map*) vm;
    or_cam cam;
    cam_init(&cam, pin, 1);
    or_iso T_WS;
    iso_from16(T_WS16, &T_WS);
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < cam.h; y++) {
        for (int x = 0; x < cam.w; x++) {
            const long idx = x + (long) y * cam.w;
            real dS[3], dn[3], dW[3];
            cam_backproject(&cam, (real) x, (real) y, dS);
            normalized3(dS, dn);
            lin_apply(&T_WS, dn, dW);
            real hit[3];
            or_cur_stats = stats ? stats + 7 * idx : NULL;
            const int hit_ok = raycast_ray(m, T_WS.t, dW, cam_near_dist(&cam, dS), cam_far_dist(&cam, dS), hit);
            or_cur_stats = NULL;
            if (hit_ok) {
                if (scale)
                    scale[idx] = 0;
                memcpy(points + 3 * idx, hit, sizeof(real) * 3);
                real g[3];
                if (or_get_field_grad((void*) m, hit, g)) {
                    normalized3(g, normals + 3 * idx);
                }
                else {
                    normals[3 * idx] = normals[3 * idx + 1] = normals[3 * idx + 2] = 0;
                }
            }
            else {
                points[3 * idx] = points[3 * idx + 1] = points[3 * idx + 2] = 0;
                normals[3 * idx] = normals[3 * idx + 1] = normals[3 * idx + 2] = 0;
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Preprocessing: src/map/preprocessor.cpp:109-195,                          */
/* include/se/map/impl/preprocessor_impl.hpp:18-40                           */
/* ------------------------------------------------------------------------- */
OR_API void or_half_sample(const float* in, int w, int h, float* out, float e_d)
{
    const int ow = w / 2, oh = h / 2;
#pragma omp parallel for
    for (int y = 0; y < oh; y++) {
        for (int x = 0; x < ow; x++) {
            const int ix = 2 * x, iy = 2 * y;
            float count = 0.0f, sum = 0.0f;
            const float c = in[ix + iy * w];
            for (int i = 0; i <= 1; i++) {
                for (int j = 0; j <= 1; j++) {
                    int px = ix + j, py = iy + i;
                    if (px > 2 * ow - 1)
                        px = 2 * ow - 1;
                    if (py > 2 * oh - 1)
                        py = 2 * oh - 1;
                    const float v = in[px + py * w];
                    if (fabsf(v - c) < e_d) {
                        count += 1.0f;
                        sum += v;
                    }
                }
            }
            out[x + y * ow] = sum / count;
        }
    }
}

OR_API void or_depth_to_points(const float* depth, const or_pinhole* pin, int level, real* points)
{
    or_cam cam;
    cam_init_s(&cam, pin, 1 << level, 1);
#pragma omp parallel for
    for (int y = 0; y < cam.h; y++) {
        for (int x = 0; x < cam.w; x++) {
            const long i = x + (long) y * cam.w;
            if (depth[i] > 0) {
                real dir[3];
                cam_backproject(&cam, (real) x, (real) y, dir);
                for (int k = 0; k < 3; k++)
                    points[3 * i + k] = (real) depth[i] * dir[k];
            }
            else {
                points[3 * i] = points[3 * i + 1] = points[3 * i + 2] = 0;
            }
        }
    }
}

OR_API void or_points_to_normals(const real* pts, int w, int h, int neg_y, real* normals)
{
#pragma omp parallel for
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            const long i = x + (long) y * w;
            real* n = normals + 3 * i;
            if (pts[3 * i + 2] == 0) {
                n[0] = n[1] = n[2] = 0;
                continue;
            }
            const int xl = x - 1 > 0 ? x - 1 : 0, xr = x + 1 < w - 1 ? x + 1 : w - 1;
            const int ym = y - 1 > 0 ? y - 1 : 0, yp = y + 1 < h - 1 ? y + 1 : h - 1;
            const int yu = neg_y ? ym : yp, yd = neg_y ? yp : ym;
            const real* l = pts + 3 * (xl + (long) y * w);
            const real* r = pts + 3 * (xr + (long) y * w);
            const real* u = pts + 3 * (x + (long) yu * w);
            const real* d = pts + 3 * (x + (long) yd * w);
            if (l[2] == 0 || r[2] == 0 || u[2] == 0 || d[2] == 0) {
                n[0] = n[1] = n[2] = 0;
                continue;
            }
            real dx[3], dy[3], c[3];
            for (int k = 0; k < 3; k++) {
                dx[k] = r[k] - l[k];
                dy[k] = u[k] - d[k];
            }
            cross3(dx, dy, c);
            normalized3(c, n);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* ICP: include/se/tracker/impl/icp_impl.hpp:16-102, src/tracker/icp.cpp:15-246, */
/* include/se/tracker/impl/tracker_impl.hpp:30-116,                          */
/* include/se/common/impl/math_util_impl.hpp:117-174                         */
/* ------------------------------------------------------------------------- */
typedef struct {
    int result;
    real error;
    real J[6];
} or_icp_row;

static void icp_track(or_icp_row* rows, int ref_w, const real* in_pts, const real* in_nrm, int w, int h,
                      const real* ref_pts, const real* ref_nrm, const or_iso* T_WS, const or_iso* T_WS_ref,
                      const or_cam* cam_full, real dist_th, real normal_th)
{
    or_iso T_refW;
    iso_inv(T_WS_ref, &T_refW);
#pragma omp parallel for
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            or_icp_row* row = rows + x + (long) y * ref_w;
            const long i = x + (long) y * w;
            const real* n_in = in_nrm + 3 * i;
            if (n_in[0] == 0 && n_in[1] == 0 && n_in[2] == 0) {
                row->result = -1;
                continue;
            }
            real pW[3], pR[3], uv[2];
            iso_apply(T_WS, in_pts + 3 * i, pW);
            iso_apply(&T_refW, pW, pR);
            if (!cam_project(cam_full, pR, uv)) {
                row->result = -2;
                continue;
            }
            int px[2];
            round_pixel(uv, px);
            const long ri = px[0] + (long) px[1] * ref_w;
            const real rn[3] = {-ref_nrm[3 * ri], -ref_nrm[3 * ri + 1], -ref_nrm[3 * ri + 2]};
            if (rn[0] == 0 && rn[1] == 0 && rn[2] == 0) {
                row->result = -3;
                continue;
            }
            real diff[3], nW[3];
            for (int k = 0; k < 3; k++)
                diff[k] = ref_pts[3 * ri + k] - pW[k];
            lin_apply(T_WS, n_in, nW);
            if (norm3(diff) > dist_th) {
                row->result = -4;
                continue;
            }
            if (dot3(nW, rn) < normal_th) {
                row->result = -5;
                continue;
            }
            row->result = 1;
            row->error = dot3(rn, diff);
            row->J[0] = rn[0];
            row->J[1] = rn[1];
            row->J[2] = rn[2];
            real cp[3];
            cross3(pW, rn, cp);
            row->J[3] = cp[0];
            row->J[4] = cp[1];
            row->J[5] = cp[2];
        }
    }
}

/* reduceKernel: 8 row-interleaved partials (each summed in row order here), then summed into
 * row 0 (icp.cpp:45-204).  The reference's OpenMP reduction order is unspecified. */
static void icp_reduce(real out[32], const or_icp_row* rows, int ref_w, int w, int h)
{
    real part[8][32];
    memset(part, 0, sizeof(part));
    for (int b = 0; b < 8; b++) {
        real* s = part[b];
        for (int y = b; y < h; y += 8) {
            for (int x = 0; x < w; x++) {
                const or_icp_row* r = rows + x + (long) y * ref_w;
                if (r->result < 1) {
                    s[29] += r->result == -4 ? 1 : 0;
                    s[30] += r->result == -5 ? 1 : 0;
                    s[31] += r->result > -4 ? 1 : 0;
                    continue;
                }
                s[0] += r->error * r->error;
                for (int k = 0; k < 6; k++)
                    s[1 + k] += r->error * r->J[k];
                int c = 7;
                for (int a = 0; a < 6; a++)
                    for (int bb = a; bb < 6; bb++)
                        s[c++] += r->J[a] * r->J[bb];
                s[28] += 1;
            }
        }
    }
    for (int k = 0; k < 32; k++) {
        real v = part[0][k];
        for (int b = 1; b < 8; b++)
            v += part[b][k];
        out[k] = v;
    }
}

/* solve (icp.cpp:33-41): 6x6 Cholesky of JTJ, x = 0 if it fails */
static void icp_solve(const real vals[27], real x[6])
{
    real A[6][6];
    int c = 6;
    for (int i = 0; i < 6; i++)
        for (int j = i; j < 6; j++) {
            A[i][j] = vals[c++];
            A[j][i] = A[i][j];
        }
    real L[6][6];
    memset(L, 0, sizeof(L));
    for (int k = 0; k < 6; k++) {
        real s = A[k][k];
        for (int j = 0; j < k; j++)
            s -= L[k][j] * L[k][j];
        if (!(s > 0)) {
            for (int i = 0; i < 6; i++)
                x[i] = 0;
            return;
        }
        L[k][k] = sqrt(s);
        for (int i = k + 1; i < 6; i++) {
            real v = A[i][k];
            for (int j = 0; j < k; j++)
                v -= L[i][j] * L[k][j];
            L[i][k] = v / L[k][k];
        }
    }
    real yv[6];
    for (int i = 0; i < 6; i++) {
        real v = vals[i];
        for (int j = 0; j < i; j++)
            v -= L[i][j] * yv[j];
        yv[i] = v / L[i][i];
    }
    for (int i = 5; i >= 0; i--) {
        real v = yv[i];
        for (int j = i + 1; j < 6; j++)
            v -= L[j][i] * x[j];
        x[i] = v / L[i][i];
    }
}

/* math::exp (math_util_impl.hpp:117-174): twist a = [v; w] */
OR_API void or_se3_exp(const real a[6], real T16[16])
{
    const real w[3] = {a[3], a[4], a[5]};
    const real theta_sq = dot3(w, w);
    const real theta = sqrt(theta_sq);
    const real half_theta = (real) 0.5 * theta;
    real imag, re;
    if (theta < (real) 1e-10) {
        const real t4 = theta_sq * theta_sq;
        imag = (real) 0.5 - (real) (1.0 / 48.0) * theta_sq + (real) (1.0 / 3840.0) * t4;
        re = (real) 1 - (real) 0.5 * theta_sq + (real) (1.0 / 384.0) * t4;
    }
    else {
        imag = sin(half_theta) / theta;
        re = cos(half_theta);
    }
    real q[4] = {re, imag * w[0], imag * w[1], imag * w[2]}; /* w x y z */
    const real qn = sqrt(q[0] * q[0] + (q[1] * q[1] + (q[2] * q[2] + q[3] * q[3])));
    for (int i = 0; i < 4; i++)
        q[i] = q[i] / qn;
    const real qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    /* Quaternion::toRotationMatrix term order */
    const real tx = (real) 2 * qx, ty = (real) 2 * qy, tz = (real) 2 * qz;
    const real twx = tx * qw, twy = ty * qw, twz = tz * qw;
    const real txx = tx * qx, txy = ty * qx, txz = tz * qx;
    const real tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
    real R[3][3];
    R[0][0] = (real) 1 - (tyy + tzz);
    R[0][1] = txy - twz;
    R[0][2] = txz + twy;
    R[1][0] = txy + twz;
    R[1][1] = (real) 1 - (txx + tzz);
    R[1][2] = tyz - twx;
    R[2][0] = txz - twy;
    R[2][1] = tyz + twx;
    R[2][2] = (real) 1 - (txx + tyy);
    real V[3][3];
    if (theta < (real) 1e-10) {
        memcpy(V, R, sizeof(V));
    }
    else {
        const real W[3][3] = {{0, -w[2], w[1]}, {w[2], 0, -w[0]}, {-w[1], w[0], 0}};
        real W2[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                W2[i][j] = W[i][0] * W[0][j] + (W[i][1] * W[1][j] + W[i][2] * W[2][j]);
        const real c1 = ((real) 1 - cos(theta)) / theta_sq;
        const real c2 = (theta - sin(theta)) / (theta_sq * theta);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                V[i][j] = ((i == j ? (real) 1 : (real) 0) + c1 * W[i][j]) + c2 * W2[i][j];
    }
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T16[4 * i + j] = R[i][j];
        T16[4 * i + 3] = V[i][0] * a[0] + (V[i][1] * a[1] + V[i][2] * a[2]);
    }
    T16[12] = T16[13] = T16[14] = 0;
    T16[15] = 1;
}

#define OR_MAX_LEVELS 6
typedef struct {
    int iterations[OR_MAX_LEVELS]; /* per level, finest first; num_levels = iterations.size() */
    int num_levels;
    float dist_threshold, normal_threshold, track_threshold, icp_threshold;
} or_tracker_config;

/* Tracker::track(depth, T_WS, ref points, ref normals) (tracker_impl.hpp:30-116).
 * T_WS16 is in/out.  sums_out receives the last reduction (32 values).  Returns 1 on success. */
OR_API int or_track(const float* depth, const or_pinhole* pin, const or_tracker_config* cfg, real T_WS16[16],
                    const real* ref_pts, const real* ref_nrm, real sums_out[32], int* iters_out,
                    int* level_iters_out)
{
    or_cam cam;
    cam_init(&cam, pin, 1);
    const int W = cam.w, H = cam.h;
    const int nl = cfg->num_levels < 1 ? 1 : (cfg->num_levels > OR_MAX_LEVELS ? OR_MAX_LEVELS : cfg->num_levels);
    /* tracker_impl.hpp:50-84: one level per iterations entry */
    float* dimg[OR_MAX_LEVELS] = {0};
    real* pts[OR_MAX_LEVELS] = {0};
    real* nrm[OR_MAX_LEVELS] = {0};
    int lw[OR_MAX_LEVELS], lh[OR_MAX_LEVELS];
    for (int l = 0; l < nl; l++) {
        lw[l] = W >> l;
        lh[l] = H >> l;
        dimg[l] = (float*) calloc((size_t) lw[l] * lh[l], sizeof(float));
        pts[l] = (real*) calloc((size_t) lw[l] * lh[l] * 3, sizeof(real));
        nrm[l] = (real*) calloc((size_t) lw[l] * lh[l] * 3, sizeof(real));
    }
    memcpy(dimg[0], depth, sizeof(float) * W * H);
    const float e_d = 0.1f * 3;
    for (int l = 1; l < nl; l++)
        or_half_sample(dimg[l - 1], lw[l - 1], lh[l - 1], dimg[l], e_d);
    for (int l = 0; l < nl; l++) {
        or_depth_to_points(dimg[l], pin, l, pts[l]);
        or_points_to_normals(pts[l], lw[l], lh[l], cam.left_hand_frame, nrm[l]);
    }
    or_iso T, T_ref;
    iso_from16(T_WS16, &T);
    T_ref = T;
    or_icp_row* rows = (or_icp_row*) calloc((size_t) W * H, sizeof(or_icp_row));
    real sums[32];
    memset(sums, 0, sizeof(sums));
    int total_iters = 0;
    for (int level = nl - 1; level >= 0; level--) {
        for (int it = 0; it < cfg->iterations[level]; it++) {
            icp_track(rows, W, pts[level], nrm[level], lw[level], lh[level], ref_pts, ref_nrm, &T, &T_ref, &cam,
                      cfg->dist_threshold, cfg->normal_threshold);
            icp_reduce(sums, rows, W, W / (1 << level), H / (1 << level));
            total_iters++;
            if (level_iters_out)
                level_iters_out[level]++;
            real x[6];
            icp_solve(sums + 1, x);
            real D16[16];
            or_se3_exp(x, D16);
            or_iso D, Tn;
            iso_from16(D16, &D);
            iso_mul(&D, &T, &Tn);
            T = Tn;
            real xn = sqrt(x[0] * x[0] + (x[1] * x[1] + (x[2] * x[2] + (x[3] * x[3] + (x[4] * x[4] + x[5] * x[5])))));
            if (xn < cfg->icp_threshold)
                break;
        }
    }
    int ok = 1;
    if (((double) sqrt(sums[0] / sums[28]) > 2e-2) || (sums[28] / (real) (W * H) < cfg->track_threshold)) {
        T = T_ref;
        ok = 0;
    }
    iso_to16(&T, T_WS16);
    if (sums_out)
        memcpy(sums_out, sums, sizeof(sums));
    if (iters_out)
        *iters_out = total_iters;
    for (int l = 0; l < nl; l++) {
        free(dimg[l]);
        free(pts[l]);
        free(nrm[l]);
    }
    free(rows);
    return ok;
}

/* One trackKernel + reduceKernel pass at one level, for unit checks of the GPU kernels. */
OR_API void or_icp_track_reduce(const real* in_pts, const real* in_nrm, int w, int h, const real* ref_pts,
                                const real* ref_nrm, const or_pinhole* pin, const real T_WS16[16],
                                const real T_ref16[16], float dist_th, float normal_th, int* results,
                                real* errors, real* J, real sums[32])
{
    or_cam cam;
    cam_init(&cam, pin, 1);
    or_iso T, Tr;
    iso_from16(T_WS16, &T);
    iso_from16(T_ref16, &Tr);
    or_icp_row* rows = (or_icp_row*) calloc((size_t) cam.w * cam.h, sizeof(or_icp_row));
    icp_track(rows, cam.w, in_pts, in_nrm, w, h, ref_pts, ref_nrm, &T, &Tr, &cam, dist_th, normal_th);
    icp_reduce(sums, rows, cam.w, w, h);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const or_icp_row* r = rows + x + (long) y * cam.w;
            const long i = x + (long) y * w;
            if (results)
                results[i] = r->result;
            if (errors)
                errors[i] = r->error;
            if (J)
                memcpy(J + 6 * i, r->J, sizeof(real) * 6);
        }
    free(rows);
}

OR_API void or_icp_solve(const real vals[27], real x[6])
{
    icp_solve(vals, x);
}

OR_API int or_real_size(void)
{
    return (int) sizeof(real);
}

OR_API void or_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void) n;
#endif
}

/* Entry t of the first allocated block along a ray and t_max (tcmin/tmax of the reference's
 * VoxelBlockRayIterator), for checks against a brute-force ray/box search. */
OR_API void or_first_block(void* vm, const float o_W[3], const float d_W[3], float near_t, float far_t, float out[2])
{
    or_ray_span s;
    esvo_first_block((const or_map*) vm, o_W, d_W, near_t, far_t, &s);
    out[0] = s.t_min_ray;
    out[1] = s.t_max_ray;
}

/* ------------------------------------------------------------------------- */
/* Differentiable raycast depth (growth pattern, no counterpart in the         */
/* reference).  The reference march (raycaster_impl.hpp:456-533) brackets the  */
/* zero crossing; here the root of the trilinear field along the ray is       */
/* refined inside that bracket (Illinois regula falsi), so depth is a smooth   */
/* function of the pose within a voxel cell, and                             */
/*   dt/dxi = -[g, x cross g] / (g . d)   (implicit-function theorem)          */
/* for the left twist T <- exp(xi) T, xi = [v; w], with g the gradient of the  */
/* trilinear field at the root x = o + t d.  depth = t * dn_z.                */
/* ------------------------------------------------------------------------- */
static int interp_and_grad(const or_map* m, const real pW[3], real* f, real g[3])
{
    real vf[3];
    if (!map_point_to_voxel_f(m, pW, vf))
        return 0;
    real bf[3], t[3], tc[3];
    int base[3];
    for (int i = 0; i < 3; i++) {
        bf[i] = vf[i] - (real) 0.5;
        base[i] = (int) bf[i];
        if (base[i] + 1 >= m->size || base[i] < 0)
            return 0;
    }
    real d[8];
    for (int k = 0; k < 8; k++) {
        const int c[3] = {base[0] + (k & 1), base[1] + ((k >> 1) & 1), base[2] + ((k >> 2) & 1)};
        const or_octant* b = fetch_block(m, c);
        if (!b)
            return 0;
        const int i = block_voxel_idx(b, c);
        if (b->weight[i] == (real) 0)
            return 0;
        d[k] = b->tsdf[i];
    }
    for (int i = 0; i < 3; i++) {
        t[i] = bf[i] - floor(bf[i]);
        tc[i] = (real) 1 - t[i];
    }
    *f = ((d[0] * tc[0] + d[1] * t[0]) * tc[1] + (d[2] * tc[0] + d[3] * t[0]) * t[1]) * tc[2]
        + ((d[4] * tc[0] + d[5] * t[0]) * tc[1] + (d[6] * tc[0] + d[7] * t[0]) * t[1]) * t[2];
    if (g) {
        const real gx = ((d[1] - d[0]) * tc[1] + (d[3] - d[2]) * t[1]) * tc[2]
            + ((d[5] - d[4]) * tc[1] + (d[7] - d[6]) * t[1]) * t[2];
        const real gy = ((d[2] - d[0]) * tc[0] + (d[3] - d[1]) * t[0]) * tc[2]
            + ((d[6] - d[4]) * tc[0] + (d[7] - d[5]) * t[0]) * t[2];
        const real gz = ((d[4] - d[0]) * tc[0] + (d[5] - d[1]) * t[0]) * tc[1]
            + ((d[6] - d[2]) * tc[0] + (d[7] - d[3]) * t[0]) * t[1];
        g[0] = gx / m->res;
        g[1] = gy / m->res;
        g[2] = gz / m->res;
    }
    return 1;
}

/* Refined root in [ta, tb] of f(o + t d) (f(ta) >= 0 > f(tb)); returns 0 if the field is not
 * interpolable somewhere on the way. */
static int refine_root(const or_map* m, const real o[3], const real d[3], real ta, real tb, int iters, real* t_out)
{
    real p[3], fa, fb;
    for (int i = 0; i < 3; i++)
        p[i] = o[i] + d[i] * ta;
    if (!interp_and_grad(m, p, &fa, NULL))
        return 0;
    for (int i = 0; i < 3; i++)
        p[i] = o[i] + d[i] * tb;
    if (!interp_and_grad(m, p, &fb, NULL))
        return 0;
    if (!(fa >= 0 && fb < 0))
        return 0;
    int side = 0;
    real t = tb;
    for (int it = 0; it < iters; it++) {
        t = (ta * fb - tb * fa) / (fb - fa);
        real f;
        for (int i = 0; i < 3; i++)
            p[i] = o[i] + d[i] * t;
        if (!interp_and_grad(m, p, &f, NULL))
            return 0;
        if (f < 0) {
            tb = t;
            fb = f;
            if (side == -1)
                fa *= (real) 0.5;
            side = -1;
        }
        else {
            ta = t;
            fa = f;
            if (side == 1)
                fb *= (real) 0.5;
            side = 1;
        }
        if (tb - ta < (real) 1e-12)
            break;
    }
    *t_out = t;
    return 1;
}

/* Per pixel: depth (z) of the refined root and d depth / d xi (6); valid = 0 where the march
 * misses or the field is not interpolable around the root. */
OR_API void or_raycast_depth_grad(void* vm, const or_pinhole* pin, const real T_WS16[16], int iters, real* depth,
                                  real* grad6, unsigned char* valid)
{
    const or_map* m = (const or_map*) vm;
    or_cam cam;
    cam_init(&cam, pin, 1);
    or_iso T_WS;
    iso_from16(T_WS16, &T_WS);
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < cam.h; y++) {
        for (int x = 0; x < cam.w; x++) {
            const long idx = x + (long) y * cam.w;
            real dS[3], dn[3], dW[3], hit[3];
            cam_backproject(&cam, (real) x, (real) y, dS);
            normalized3(dS, dn);
            lin_apply(&T_WS, dn, dW);
            depth[idx] = 0;
            valid[idx] = 0;
            for (int k = 0; k < 6; k++)
                grad6[6 * idx + k] = 0;
            real ta, tb;
            if (!raycast_ray_bracket(m, T_WS.t, dW, cam_near_dist(&cam, dS), cam_far_dist(&cam, dS), hit, &ta, &tb))
                continue;
            real t;
            if (!refine_root(m, T_WS.t, dW, ta, tb, iters, &t))
                continue;
            real xr[3], f, g[3];
            for (int i = 0; i < 3; i++)
                xr[i] = T_WS.t[i] + dW[i] * t;
            if (!interp_and_grad(m, xr, &f, g))
                continue;
            const real gd = dot3(g, dW);
            if (gd == 0)
                continue;
            real xg[3];
            cross3(xr, g, xg);
            depth[idx] = t * dn[2];
            for (int i = 0; i < 3; i++) {
                grad6[6 * idx + i] = -g[i] / gd * dn[2];
                grad6[6 * idx + 3 + i] = -xg[i] / gd * dn[2];
            }
            valid[idx] = 1;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Single-res marching cubes: marching_cube_impl.hpp:83-95 (compute_intersection), */
/* 97-127 (edge_to_corners), 143-158 (gather_data), 20-38 (isosurface::occupied),  */
/* 814-818 (checkVertex), 832-889 (marching_cube_kernel); the vertices mapped by   */
/* T_OV = T_OW * T_WM * Scaling(res) (map_impl.hpp:452-458, octree_impl.hpp:121-133). */
/* The case table is the one the library meshes with; it is pinned by its own     */
/* properties in the tests.                                                      */
/* ------------------------------------------------------------------------- */
#include "../supereight2_amd/csrc/se8_mc_table.h"

/* getData: the voxel of an allocated block, else the default {1, 0} (visitor_impl.hpp:67-81) */
static inline void mc_get_data(const or_map* m, const int v[3], real* tsdf, real* weight)
{
    const or_octant* b = fetch_block(m, v);
    if (!b) {
        *tsdf = 1;
        *weight = 0;
        return;
    }
    const int i = block_voxel_idx(b, v);
    *tsdf = b->tsdf[i];
    *weight = b->weight[i];
}

static inline void mc_corner(int x, int y, int z, int k, int c[3])
{
    c[0] = x + (int) ((SE8_MC_CORNER_DX >> k) & 1);
    c[1] = y + (int) ((SE8_MC_CORNER_DY >> k) & 1);
    c[2] = z + (int) ((SE8_MC_CORNER_DZ >> k) & 1);
}

/* Triangles of the cell (x,y,z) into tri[] (3 vertices each, map voxel units); returns the count. */
static int mc_cell(const or_map* m, int x, int y, int z, real tri[5][3][3])
{
    unsigned idx = 0;
    for (int k = 0; k < 8; k++) {
        int c[3];
        real f, w;
        mc_corner(x, y, z, k, c);
        mc_get_data(m, c, &f, &w);
        if (w == 0)
            return 0; /* an unknown corner: no face (isosurface::occupied) */
        if (f < 0)
            idx |= 1u << k;
    }
    const uint64_t word = se8_mc_tri[idx];
    const int ntri = (int) (word >> 60);
    int n = 0;
    for (int j = 0; j < ntri; j++) {
        int skip = 0;
        for (int v = 0; v < 3 && !skip; v++) {
            const int e = (int) ((word >> (4 * (3 * j + v))) & 0xF);
            const int a = (int) ((SE8_MC_EDGE_A >> (4 * e)) & 0xF);
            const int b = (int) ((SE8_MC_EDGE_B >> (4 * e)) & 0xF);
            int c0[3], c1[3];
            mc_corner(x, y, z, a, c0);
            mc_corner(x, y, z, b, c1);
            real v0, v1, w;
            mc_get_data(m, c0, &v0, &w);
            mc_get_data(m, c1, &v1, &w);
            const real t = ((real) 0 - v0) / (v1 - v0);
            for (int i = 0; i < 3; i++) {
                const real p0 = (real) c0[i] + (real) 0.5;
                const real p1 = (real) c1[i] + (real) 0.5;
                tri[n][v][i] = p0 + t * (p1 - p0);
            }
            const real* p = tri[n][v];
            const real dim = (real) m->size;
            if (p[0] <= 0 || p[1] <= 0 || p[2] <= 0 || p[0] > dim || p[1] > dim || p[2] > dim)
                skip = 1; /* checkVertex */
        }
        if (!skip)
            n++;
    }
    return n;
}

/* Blocks in m->blocks order, cells x, y, z nested as the reference loops.  faces: cap * 9 reals
 * (may be NULL with cap 0 to count).  Returns the number of faces. */
OR_API long or_mesh(void* vm, const real T_OW16[16], real* faces, long cap)
{
    const or_map* m = (const or_map*) vm;
    or_iso T_OW, T_WM, T_OV;
    iso_from16(T_OW16, &T_OW);
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T_WM.R[i][j] = i == j ? 1 : 0;
        T_WM.t[i] = -m->t_MW[i];
    }
    iso_mul(&T_OW, &T_WM, &T_OV);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            T_OV.R[i][j] = T_OV.R[i][j] * m->res;
    const long nb = m->n_blocks;
    long* count = (long*) calloc((size_t) nb + 1, sizeof(long));
    for (int pass = 0; pass < 2; pass++) {
#pragma omp parallel for schedule(dynamic, 16)
        for (long bi = 0; bi < nb; bi++) {
            const or_octant* b = m->blocks[bi];
            int last[3];
            for (int i = 0; i < 3; i++)
                last[i] = b->coord[i] + OR_BLOCK < m->size - 1 ? b->coord[i] + OR_BLOCK : m->size - 1;
            long out = pass ? count[bi] : 0;
            for (int x = b->coord[0]; x < last[0]; x++)
                for (int y = b->coord[1]; y < last[1]; y++)
                    for (int z = b->coord[2]; z < last[2]; z++) {
                        real tri[5][3][3];
                        const int n = mc_cell(m, x, y, z, tri);
                        if (pass && out + n <= cap)
                            for (int j = 0; j < n; j++)
                                for (int v = 0; v < 3; v++) {
                                    real o[3];
                                    iso_apply(&T_OV, tri[j][v], o);
                                    memcpy(faces + 9 * (out + j) + 3 * v, o, sizeof(o));
                                }
                        out += n;
                    }
            if (!pass)
                count[bi] = out;
        }
        if (!pass) {
            long run = 0;
            for (long bi = 0; bi < nb; bi++) {
                const long c = count[bi];
                count[bi] = run;
                run += c;
            }
            count[nb] = run;
            if (!faces || cap < run)
                break;
        }
    }
    const long total = count[nb];
    free(count);
    return total;
}

/* ------------------------------------------------------------------------- */
/* Input side: preprocessor::downsample_depth (src/map/preprocessor.cpp:19-84) */
/* and the readers' uint16 -> metres conversion (app/src/reader_raw.cpp:116-121, */
/* app/src/reader_tum.cpp nextDepth).  The region's valid depths are sorted   */
/* stably (equal depths keep the region's row-major order; std::sort leaves it */
/* unspecified) and region[n / 2] is taken.                                   */
/* ------------------------------------------------------------------------- */
OR_API void or_depth_from_u16(const uint16_t* raw, long n, float scale, int divide, float* out)
{
    for (long i = 0; i < n; i++)
        out[i] = divide ? (float) raw[i] / scale : (float) raw[i] * scale;
}

OR_API int or_downsample_depth(const float* in, int w_in, int h_in, float* out, int w_out, int h_out, uint64_t* map)
{
    if (w_out <= 0 || h_out <= 0 || w_in % w_out || h_in % h_out || w_in / w_out != h_in / h_out)
        return -1;
    const int ratio = w_in / w_out;
    float* dv = (float*) malloc(sizeof(float) * ratio * ratio);
    int* ix = (int*) malloc(sizeof(int) * ratio * ratio);
    for (int y_out = 0; y_out < h_out; y_out++)
        for (int x_out = 0; x_out < w_out; x_out++) {
            const int xb = x_out * ratio, yb = y_out * ratio;
            int n = 0;
            for (int yr = 0; yr < ratio; yr++)
                for (int xr = 0; xr < ratio; xr++) {
                    const float d = in[(size_t) (yb + yr) * w_in + xb + xr];
                    if (d >= 1e-5f) {
                        /* insertion keeps equal depths in arrival order */
                        int k = n++;
                        while (k > 0 && dv[k - 1] > d) {
                            dv[k] = dv[k - 1];
                            ix[k] = ix[k - 1];
                            k--;
                        }
                        dv[k] = d;
                        ix[k] = (xb + xr) + (yb + yr) * w_in;
                    }
                }
            if (n == 0) {
                out[(size_t) y_out * w_out + x_out] = 0.0f;
                if (map)
                    map[(size_t) y_out * w_out + x_out] = (uint64_t) xb + (uint64_t) yb * (uint64_t) w_in;
            }
            else {
                out[(size_t) y_out * w_out + x_out] = dv[n / 2];
                if (map)
                    map[(size_t) y_out * w_out + x_out] = (uint64_t) ix[n / 2];
            }
        }
    free(dv);
    free(ix);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Debug views: Map::saveFieldSlices' planes (map_impl.hpp:272-308,           */
/* octree_io_impl.hpp:19-80: z outer, y, x inner) and octree_structure_mesh   */
/* (structure_meshing_impl.hpp:15-60) in the octree iterator's depth-first    */
/* order, vertices mapped by T_OV = T_OW * T_WM * Scaling(res).               */
/* ------------------------------------------------------------------------- */
OR_API void or_field_slice(void* vm, int axis, int index, real* out)
{
    const or_map* m = (const or_map*) vm;
    const int n = m->size;
    for (int outer = 0; outer < n; outer++)
        for (int inner = 0; inner < n; inner++) {
            int v[3];
            if (axis == 0)
                v[0] = index, v[1] = inner, v[2] = outer;
            else if (axis == 1)
                v[0] = inner, v[1] = index, v[2] = outer;
            else
                v[0] = inner, v[1] = outer, v[2] = index;
            real t = 1, w = 0;
            if (index >= 0 && index < n)
                mc_get_data(m, v, &t, &w);
            out[(size_t) outer * n + inner] = t;
        }
}

static const int or_quad_corners[6][4] = {{0, 2, 3, 1}, {1, 3, 7, 5}, {5, 7, 6, 4}, {0, 4, 6, 2}, {0, 1, 5, 4}, {2, 6, 7, 3}};

static long structure_rec(const or_octant* o, const or_iso* T, int only_leaves, real* quads, signed char* scales,
                          long n, long cap)
{
    int leaf = 1;
    for (int k = 0; k < 8 && !o->is_block; k++)
        leaf = leaf && !o->child[k];
    if (!only_leaves || leaf) {
        if (quads && n + 6 <= cap) {
            real corner[8][3];
            for (int k = 0; k < 8; k++) {
                const real p[3] = {(real) (o->coord[0] + ((k & 1) ? o->size : 0)),
                                   (real) (o->coord[1] + ((k & 2) ? o->size : 0)),
                                   (real) (o->coord[2] + ((k & 4) ? o->size : 0))};
                iso_apply(T, p, corner[k]);
            }
            for (int f = 0; f < 6; f++) {
                for (int v = 0; v < 4; v++)
                    memcpy(quads + 12 * (n + f) + 3 * v, corner[or_quad_corners[f][v]], sizeof(real) * 3);
                scales[n + f] = o->is_block ? 0 : 7;
            }
        }
        n += 6;
    }
    for (int k = 0; k < 8 && !o->is_block; k++)
        if (o->child[k])
            n = structure_rec(o->child[k], T, only_leaves, quads, scales, n, cap);
    return n;
}

OR_API long or_structure(void* vm, const real T_OW16[16], int only_leaves, real* quads, signed char* scales, long cap)
{
    const or_map* m = (const or_map*) vm;
    or_iso T_OW, T_WM, T_OV;
    iso_from16(T_OW16, &T_OW);
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T_WM.R[i][j] = i == j ? 1 : 0;
        T_WM.t[i] = -m->t_MW[i];
    }
    iso_mul(&T_OW, &T_WM, &T_OV);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            T_OV.R[i][j] = T_OV.R[i][j] * m->res;
    return structure_rec(m->root, &T_OV, only_leaves, quads, scales, 0, cap);
}
