// Device math of the MLS-MPM engine for gfx950: quaternions, the register
// 3x3 SVD, tool SDFs and their adjoints, grid collision, primitive kinematics.
//
// Everything here is fp32 (the reference runs ti.f32, plb/config/default_config.py:15).
// Adjoints follow Taichi 0.7.26 reverse-mode rules (the dependency that
// generated the reference's .grad kernels): only the taken branch is
// differentiated; max(a, b) gives its adjoint to a iff b < a, else to b;
// min(a, b) gives it to a iff a < b, else to b; abs'(x) = sgn(x), sgn(0) = 0.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dsk {

#define DEV __device__ __forceinline__

// ------------------------------------------------------------------ tables
struct SceneC {
    int n_grid, nb, p_off, p_cnt;
    int blk_off, prim_off, n_prims, gf_mode;   // gf_mode: 0 slip walls, 1 Coulomb ground, 2 sticky ground
    int m_off, chunk_off, n_chunks, bm_off;    // bm_off: first word of the scene's node-block bitmap
    long long g_off;                            // first node (block-major layout)
    float dx, inv_dx, dt, p_mass;
    float mu, lam, ys, stress_k;
    float c4, lower, upper, gf;
    float g30[4];
    double p_mass_d;                            // p_mass before rounding to fp32
};

struct PrimC {
    int type, action_dim, unit_shape, n_units;
    float friction, softness, h, r;
    float size[4];
    float ph[4];
    float prot[4];
    float lo[4];
    float hi[4];
    float min_gap, max_gap, pad0, pad1;
    float scale[8];
};

enum { P_SPHERE = 0, P_CAPSULE = 1, P_ROLLINGPIN = 2, P_ROLLINGPIN_EXT = 3, P_BOX = 4, P_GRIPPER = 5,
       P_PRISM = 6, P_KNIFE = 7, P_CHOPSTICKS = 8, P_GRIPPER2 = 9, P_CYLINDER = 10, P_TORUS = 11 };
enum { SH_SPHERE = 0, SH_CAPSULE = 1, SH_BOX = 2, SH_PRISM = 3, SH_KNIFE = 4, SH_CHOP = 5, SH_CYL = 6, SH_TORUS = 7 };

constexpr float LEN_EPS_SHAPE = 1e-14f;  // primitives.py:14-15
constexpr float LEN_EPS_BASE = 1e-8f;    // primitive/utils.py:5-6

// ------------------------------------------------------------------ vec3
struct v3 { float x, y, z; };
DEV v3 mk(float a, float b, float c) { v3 r; r.x = a; r.y = b; r.z = c; return r; }
DEV v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
DEV v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
DEV v3 operator-(v3 a) { return mk(-a.x, -a.y, -a.z); }
DEV v3 operator*(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
DEV v3 operator*(v3 a, float s) { return mk(s * a.x, s * a.y, s * a.z); }
DEV v3& operator+=(v3& a, v3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
DEV v3& operator-=(v3& a, v3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
DEV float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DEV v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
DEV float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
DEV void setc(v3& a, int i, float v) { if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v; }
DEV float tmax(float a, float b) { return b < a ? a : b; }
DEV float tmin(float a, float b) { return a < b ? a : b; }
DEV float sgnf(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }

struct q4 { float w, x, y, z; };
DEV q4 mkq(float w, float x, float y, float z) { q4 q; q.w = w; q.x = x; q.y = y; q.z = z; return q; }
DEV v3 qv(q4 q) { return mk(q.x, q.y, q.z); }

// qrot (plb/engine/primitive/utils.py:9-15)
DEV v3 qrot(q4 q, v3 v) {
    v3 u = qv(q);
    v3 uv = cross(u, v);
    v3 uuv = cross(u, uv);
    return v + 2.f * (q.w * uv + uuv);
}
// adjoint of qrot: adds to gq (if non-null) and returns gv
DEV v3 qrot_rev(q4 q, v3 v, v3 go, q4* gq) {
    v3 u = qv(q);
    v3 uv = cross(u, v);
    v3 guuv = 2.f * go;
    v3 guv = (2.f * q.w) * go;
    v3 gu = cross(uv, guuv);
    guv += cross(guuv, u);
    gu += cross(v, guv);
    if (gq) { gq->w += 2.f * dot(go, uv); gq->x += gu.x; gq->y += gu.y; gq->z += gu.z; }
    return cross(guv, u) + go;
}
// conjugate, normalized without epsilon (Vector.normalized(), utils.py:52)
DEV q4 conj_n(q4 q) {
    float n = sqrtf(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    return mkq(q.w / n, -q.x / n, -q.y / n, -q.z / n);
}
DEV void conj_n_rev(q4 q, q4 g, q4* gq) {
    q4 c = mkq(q.w, -q.x, -q.y, -q.z);
    float n = sqrtf(c.w * c.w + c.x * c.x + c.y * c.y + c.z * c.z);
    float d = c.w * g.w + c.x * g.x + c.y * g.y + c.z * g.z;
    float n3 = n * n * n;
    gq->w += g.w / n - c.w * d / n3;
    gq->x -= g.x / n - c.x * d / n3;
    gq->y -= g.y / n - c.y * d / n3;
    gq->z -= g.z / n - c.z * d / n3;
}
// rotation matrix (row-major) <-> unit quaternion (w, x, y, z), the storage form
// of the stored SVD's U and V: Shepperd's method divides by the largest of
// 4w^2, 4x^2, 4y^2, 4z^2; the matrix of a quaternion is normalized by |q|^2
DEV q4 rot_to_quat(const float* R) {
    // t = (4w^2, 4x^2, 4y^2, 4z^2); k = the largest; every lane takes the same
    // instruction path (selects, one square root, one reciprocal)
    float t0 = 1.f + R[0] + R[4] + R[8], t1 = 1.f + R[0] - R[4] - R[8];
    float t2 = 1.f - R[0] + R[4] - R[8], t3 = 1.f - R[0] - R[4] + R[8];
    float a = R[7] - R[5], b = R[2] - R[6], c = R[3] - R[1];   // 4wx, 4wy, 4wz
    float d = R[1] + R[3], e = R[2] + R[6], g = R[5] + R[7];   // 4xy, 4xz, 4yz
    int k = 0;
    float tm = t0;
    if (t1 > tm) { k = 1; tm = t1; }
    if (t2 > tm) { k = 2; tm = t2; }
    if (t3 > tm) { k = 3; tm = t3; }
    float s = sqrtf(tm) * 2.f;   // 4 |q_k|
    float h = 0.25f * s, inv = __builtin_amdgcn_rcpf(s);
    if (k == 0) return mkq(h, a * inv, b * inv, c * inv);
    if (k == 1) return mkq(a * inv, h, d * inv, e * inv);
    if (k == 2) return mkq(b * inv, d * inv, h, g * inv);
    return mkq(c * inv, e * inv, g * inv, h);
}
DEV void quat_to_rot(q4 q, float* R) {
    float s = 2.f / (q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z, wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    R[0] = 1.f - s * (yy + zz); R[1] = s * (xy - wz);        R[2] = s * (xz + wy);
    R[3] = s * (xy + wz);        R[4] = 1.f - s * (xx + zz); R[5] = s * (yz - wx);
    R[6] = s * (xz - wy);        R[7] = s * (yz + wx);        R[8] = 1.f - s * (xx + yy);
}
// qmul (utils.py:18-25): Hamilton product, renormalized
DEV q4 qmul_raw(q4 q, q4 r) {
    return mkq(r.w * q.w - r.x * q.x - r.y * q.y - r.z * q.z,
               r.w * q.x + r.x * q.w - r.y * q.z + r.z * q.y,
               r.w * q.y + r.x * q.z + r.y * q.w - r.z * q.x,
               r.w * q.z - r.x * q.y + r.y * q.x + r.z * q.w);
}
DEV q4 qmul(q4 q, q4 r) {
    q4 o = qmul_raw(q, r);
    float n = sqrtf(o.w * o.w + o.x * o.x + o.y * o.y + o.z * o.z);
    return mkq(o.w / n, o.x / n, o.y / n, o.z / n);
}
DEV void qmul_rev(q4 q, q4 r, q4 go, q4* gq, q4* gr) {
    q4 o = qmul_raw(q, r);
    float n = sqrtf(o.w * o.w + o.x * o.x + o.y * o.y + o.z * o.z);
    float d = o.w * go.w + o.x * go.x + o.y * go.y + o.z * go.z;
    float n3 = n * n * n;
    q4 g = mkq(go.w / n - o.w * d / n3, go.x / n - o.x * d / n3, go.y / n - o.y * d / n3, go.z / n - o.z * d / n3);
    gq->w += g.w * r.w + g.x * r.x + g.y * r.y + g.z * r.z;
    gq->x += -g.w * r.x + g.x * r.w - g.y * r.z + g.z * r.y;
    gq->y += -g.w * r.y + g.x * r.z + g.y * r.w - g.z * r.x;
    gq->z += -g.w * r.z - g.x * r.y + g.y * r.x + g.z * r.w;
    gr->w += g.w * q.w + g.x * q.x + g.y * q.y + g.z * q.z;
    gr->x += -g.w * q.x + g.x * q.w + g.y * q.z - g.z * q.y;
    gr->y += -g.w * q.y - g.x * q.z + g.y * q.w + g.z * q.x;
    gr->z += -g.w * q.z + g.x * q.y - g.y * q.x + g.z * q.w;
}
// w2quat (utils.py:28-40)
DEV q4 w2quat(v3 a) {
    float w = sqrtf(dot(a, a) + 1e-16f);
    q4 o = mkq(1.f, 0.f, 0.f, 0.f);
    if (w > 1e-9f) {
        float s = sinf(w / 2.f);
        o = mkq(cosf(w / 2.f), (a.x / w) * s, (a.y / w) * s, (a.z / w) * s);
    }
    return o;
}
DEV v3 w2quat_rev(v3 a, q4 go) {
    float w = sqrtf(dot(a, a) + 1e-16f);
    if (!(w > 1e-9f)) return mk(0.f, 0.f, 0.f);
    float s = sinf(w / 2.f), c = cosf(w / 2.f);
    float gw = -s / 2.f * go.w;
    v3 gv = qv(go);
    gw += dot(gv, a) * (c / (2.f * w) - s / (w * w));
    return (s / w) * gv + (gw / w) * a;
}

// ------------------------------------------------------------------ 3x3 matrices
struct m3 { float a[9]; };
DEV m3 mzero() { m3 r; for (int i = 0; i < 9; i++) r.a[i] = 0.f; return r; }
DEV m3 mmul(const m3& A, const m3& B) {
    m3 C;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) C.a[i * 3 + j] = A.a[i * 3] * B.a[j] + A.a[i * 3 + 1] * B.a[3 + j] + A.a[i * 3 + 2] * B.a[6 + j];
    return C;
}
DEV m3 mmul_tn(const m3& A, const m3& B) {  // A^T B
    m3 C;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) C.a[i * 3 + j] = A.a[i] * B.a[j] + A.a[3 + i] * B.a[3 + j] + A.a[6 + i] * B.a[6 + j];
    return C;
}
DEV m3 mmul_nt(const m3& A, const m3& B) {  // A B^T
    m3 C;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) C.a[i * 3 + j] = A.a[i * 3] * B.a[j * 3] + A.a[i * 3 + 1] * B.a[j * 3 + 1] + A.a[i * 3 + 2] * B.a[j * 3 + 2];
    return C;
}
DEV float mdet(const m3& F) {
    const float* f = F.a;
    return f[0] * (f[4] * f[8] - f[5] * f[7]) - f[1] * (f[3] * f[8] - f[5] * f[6]) + f[2] * (f[3] * f[7] - f[4] * f[6]);
}
DEV m3 mcof(const m3& F) {
    const float* f = F.a;
    m3 c;
    c.a[0] = f[4] * f[8] - f[5] * f[7]; c.a[1] = f[5] * f[6] - f[3] * f[8]; c.a[2] = f[3] * f[7] - f[4] * f[6];
    c.a[3] = f[2] * f[7] - f[1] * f[8]; c.a[4] = f[0] * f[8] - f[2] * f[6]; c.a[5] = f[1] * f[6] - f[0] * f[7];
    c.a[6] = f[1] * f[5] - f[2] * f[4]; c.a[7] = f[2] * f[3] - f[0] * f[5]; c.a[8] = f[0] * f[4] - f[1] * f[3];
    return c;
}

// ------------------------------------------------------------------ register SVD
// Own implementation following the structure of McAdams et al. 2011,
// "Computing the Singular Value Decomposition of 3x3 matrices with minimal
// branching" (UW-Madison CS TR1690): cyclic Jacobi eigen-decomposition of
// A^T A (here with exact Givens angles) gives V; B = A V; columns sorted by
// norm, descending; Givens QR of B gives U and R ~ diag(sigma).  U and V are
// rotations, sigma_1 >= sigma_2 >= |sigma_3|, and sigma_3 carries the sign of
// det(A) -- the convention of the reference's ti.svd (mpm_simulator.py:129).
// 1/x and 1/sqrt(x) for the rotation angles: in fp64 the hardware estimates
// (v_rcp_f64, v_rsq_f64) refined by two Newton steps (y <- y + y (1 - x y) and
// y <- y + y (1/2 - x y^2 / 2); each step squares the relative error), a few
// ulp from the correctly rounded division and square root at a fraction of
// their instructions.  Inputs here are normal and positive (or nonzero).
template <typename T> DEV T rot_rcp(T x) { return T(1) / x; }
template <typename T> DEV T rot_rsq(T x) { return T(1) / sqrt(x); }
template <> DEV double rot_rcp<double>(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = fma(y, fma(-x, y, 1.0), y);
    return fma(y, fma(-x, y, 1.0), y);
}
template <> DEV double rot_rsq<double>(double x) {
    double h = 0.5 * x, y = __builtin_amdgcn_rsq(x);
    y = fma(y, fma(-h * y, y, 0.5), y);
    return fma(y, fma(-h * y, y, 0.5), y);
}
template <typename T>
DEV void jacobi_rot(T& s_pp, T& s_qq, T& s_pq, T& s_rp, T& s_rq, T& vp0, T& vp1, T& vp2, T& vq0, T& vq1, T& vq2) {
    if (fabs(s_pq) < T(1e-30)) return;
    T th = (s_qq - s_pp) * rot_rcp(T(2) * s_pq);
    // t = sgn(th) / (|th| + sqrt(th^2 + 1)); past |th| = 1e15 that is
    // 1 / (2 th) to the last bit in either precision, and th^2 would
    // overflow fp32 (|th| reaches ~1e30 at |s_pq| ~ 1e-30)
    T h2 = fma(th, th, T(1));
    T t = fabs(th) > T(1e15) ? T(0.5) * rot_rcp(th)
                             : (th >= T(0) ? T(1) : T(-1)) * rot_rcp(fabs(th) + h2 * rot_rsq(h2));
    T c = rot_rsq(fma(t, t, T(1)));
    T s = t * c;
    s_pp -= t * s_pq;
    s_qq += t * s_pq;
    s_pq = T(0);
    T rp = s_rp, rq = s_rq;
    s_rp = c * rp - s * rq;
    s_rq = s * rp + c * rq;
    T a0 = vp0, a1 = vp1, a2 = vp2;
    vp0 = c * a0 - s * vq0; vp1 = c * a1 - s * vq1; vp2 = c * a2 - s * vq2;
    vq0 = s * a0 + c * vq0; vq1 = s * a1 + c * vq1; vq2 = s * a2 + c * vq2;
}
template <typename T>
DEV void givens_qr(T& a, T& b, T& c, T& s) {
    T r2 = a * a + b * b;
    if (r2 > T(1e-60)) {
        T y = rot_rsq(r2);
        c = a * y;
        s = b * y;
    } else {
        c = T(1);
        s = T(0);
    }
}
// Jacobi on A^T A (at most 5 cyclic sweeps, fewer once the off-diagonal
// part is below 1e-15 of the diagonal) for V, column sort, Givens QR of A V
// for U and the singular values; computed in T, returned in fp32.
template <typename T>
DEV void svd3_t(const m3& Af, m3& U, float sig[3], m3& V) {
    T A[9];
#pragma unroll
    for (int i = 0; i < 9; i++) A[i] = (T)Af.a[i];
    // S = A^T A
    T s00 = A[0] * A[0] + A[3] * A[3] + A[6] * A[6];
    T s11 = A[1] * A[1] + A[4] * A[4] + A[7] * A[7];
    T s22 = A[2] * A[2] + A[5] * A[5] + A[8] * A[8];
    T s01 = A[0] * A[1] + A[3] * A[4] + A[6] * A[7];
    T s02 = A[0] * A[2] + A[3] * A[5] + A[6] * A[8];
    T s12 = A[1] * A[2] + A[4] * A[5] + A[7] * A[8];
    // V columns c0, c1, c2 stored as v[row][col] -> keep per column
    T v00 = 1, v10 = 0, v20 = 0, v01 = 0, v11 = 1, v21 = 0, v02 = 0, v12 = 0, v22 = 1;
#pragma unroll
    for (int sweep = 0; sweep < 5; sweep++) {
        // (p,q) = (0,1), r = 2
        jacobi_rot(s00, s11, s01, s02, s12, v00, v10, v20, v01, v11, v21);
        // (p,q) = (0,2), r = 1
        jacobi_rot(s00, s22, s02, s01, s12, v00, v10, v20, v02, v12, v22);
        // (p,q) = (1,2), r = 0
        jacobi_rot(s11, s22, s12, s01, s02, v01, v11, v21, v02, v12, v22);
        // converged: further rotations change nothing at this precision
        T off = s01 * s01 + s02 * s02 + s12 * s12, on = s00 * s00 + s11 * s11 + s22 * s22;
        if (off <= T(1e-30) * on) break;
    }
    // B = A V
    T b[9];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        b[i * 3 + 0] = A[i * 3] * v00 + A[i * 3 + 1] * v10 + A[i * 3 + 2] * v20;
        b[i * 3 + 1] = A[i * 3] * v01 + A[i * 3 + 1] * v11 + A[i * 3 + 2] * v21;
        b[i * 3 + 2] = A[i * 3] * v02 + A[i * 3 + 1] * v12 + A[i * 3 + 2] * v22;
    }
    T vv[9] = {v00, v01, v02, v10, v11, v12, v20, v21, v22};
    T rho[3];
#pragma unroll
    for (int j = 0; j < 3; j++) rho[j] = b[j] * b[j] + b[3 + j] * b[3 + j] + b[6 + j] * b[6 + j];
    // sort columns descending; each swap negates one column to keep det V = +1
#define DSK_SWAP(i, j)                                                                        \
    if (rho[i] < rho[j]) {                                                                    \
        T t_ = rho[i]; rho[i] = rho[j]; rho[j] = t_;                                      \
        _Pragma("unroll") for (int r_ = 0; r_ < 3; r_++) {                                    \
            T tb = b[r_ * 3 + i]; b[r_ * 3 + i] = b[r_ * 3 + j]; b[r_ * 3 + j] = -tb;     \
            T tv = vv[r_ * 3 + i]; vv[r_ * 3 + i] = vv[r_ * 3 + j]; vv[r_ * 3 + j] = -tv; \
        }                                                                                     \
    }
    DSK_SWAP(0, 1)
    DSK_SWAP(0, 2)
    DSK_SWAP(1, 2)
#undef DSK_SWAP
    // Givens QR: zero b10, b20, b21 ; U accumulates the rotations
    T u[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    T c, s;
    // rows 0,1 -> zero b[1][0]
    givens_qr(b[0], b[3], c, s);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        T r0 = b[j], r1 = b[3 + j];
        b[j] = c * r0 + s * r1; b[3 + j] = -s * r0 + c * r1;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {  // U = U * G^T, G rows (c s; -s c)
        T u0 = u[i * 3], u1 = u[i * 3 + 1];
        u[i * 3] = c * u0 + s * u1; u[i * 3 + 1] = -s * u0 + c * u1;
    }
    // rows 0,2 -> zero b[2][0]
    givens_qr(b[0], b[6], c, s);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        T r0 = b[j], r2 = b[6 + j];
        b[j] = c * r0 + s * r2; b[6 + j] = -s * r0 + c * r2;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        T u0 = u[i * 3], u2 = u[i * 3 + 2];
        u[i * 3] = c * u0 + s * u2; u[i * 3 + 2] = -s * u0 + c * u2;
    }
    // rows 1,2 -> zero b[2][1]
    givens_qr(b[4], b[7], c, s);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        T r1 = b[3 + j], r2 = b[6 + j];
        b[3 + j] = c * r1 + s * r2; b[6 + j] = -s * r1 + c * r2;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        T u1 = u[i * 3 + 1], u2 = u[i * 3 + 2];
        u[i * 3 + 1] = c * u1 + s * u2; u[i * 3 + 2] = -s * u1 + c * u2;
    }
    sig[0] = (float)b[0]; sig[1] = (float)b[4]; sig[2] = (float)b[8];
#pragma unroll
    for (int i = 0; i < 9; i++) { U.a[i] = (float)u[i]; V.a[i] = (float)vv[i]; }
}
// The SVD runs in fp64 (fp32 in and out): the singular-value gaps of nearly
// isotropic deformations (all sigma ~ 1, the elastic regime) are far below
// fp32 resolution, and backward_svd divides by them; an fp32 Jacobi doubles
// the gradient error of long horizons against an fp64 trajectory.
#ifndef DSK_SVD_REAL
#define DSK_SVD_REAL double
#endif
DEV void svd3(const m3& A, m3& U, float sig[3], m3& V) { svd3_t<DSK_SVD_REAL>(A, U, sig, V); }


// ------------------------------------------------------------------ shapes (local frame)
// Box._sdf (primitives.py:369-374); g receives d sdf / d p
DEV float box_sdf(v3 s, v3 p, v3* g) {
    float q0 = fabsf(p.x) - s.x, q1 = fabsf(p.y) - s.y, q2 = fabsf(p.z) - s.z;
    float a0 = tmax(q0, 0.f), a1 = tmax(q1, 0.f), a2 = tmax(q2, 0.f);
    float L = sqrtf(a0 * a0 + a1 * a1 + a2 * a2 + LEN_EPS_SHAPE);
    float m12 = tmax(q1, q2), m = tmax(q0, m12);
    if (g) {
        float g0 = (0.f < q0) ? a0 / L : 0.f, g1 = (0.f < q1) ? a1 / L : 0.f, g2 = (0.f < q2) ? a2 / L : 0.f;
        if (m < 0.f) {
            if (m12 < q0) g0 += 1.f;
            else if (q2 < q1) g1 += 1.f;
            else g2 += 1.f;
        }
        *g = mk(sgnf(p.x) * g0, sgnf(p.y) * g1, sgnf(p.z) * g2);
    }
    return L + tmin(m, 0.f);
}
// Capsule (primitives.py:54-66): local y is the segment axis, h = full length
DEV v3 capsule_p2(float h, v3 p, float* dy) {
    float y = p.y + h / 2.f;
    float mx = tmax(y, 0.f);
    float c = tmin(mx, h);
    *dy = ((0.f < y) && (mx < h)) ? 0.f : 1.f;
    return mk(p.x, y - c, p.z);
}
DEV float capsule_sdf(float h, float r, v3 p, v3* g) {
    float dy;
    v3 v = capsule_p2(h, p, &dy);
    float L = sqrtf(dot(v, v) + LEN_EPS_SHAPE);
    if (g) *g = mk(v.x / L, v.y / L * dy, v.z / L);
    return L - r;
}
DEV v3 capsule_normal(float h, v3 p) {
    float dy;
    v3 v = capsule_p2(h, p, &dy);
    float L = sqrtf(dot(v, v) + LEN_EPS_SHAPE);
    return mk(v.x / L, v.y / L, v.z / L);
}
DEV v3 capsule_normal_rev(float h, v3 p, v3 gn) {
    float dy;
    v3 v = capsule_p2(h, p, &dy);
    float L = sqrtf(dot(v, v) + LEN_EPS_SHAPE);
    float d = dot(v, gn);
    float L3 = L * L * L;
    v3 gv = mk(gn.x / L - v.x * d / L3, gn.y / L - v.y * d / L3, gn.z / L - v.z * d / L3);
    return mk(gv.x, gv.y * dy, gv.z);
}
// Prism._sdf (primitives.py:711-719)
DEV float prism_sdf(const PrimC& P, v3 p, v3* g) {
    q4 iq = conj_n(mkq(P.prot[0], P.prot[1], P.prot[2], P.prot[3]));
    v3 gg = qrot(iq, p);
    float q0 = fabsf(gg.x), q2 = fabsf(gg.z);
    float t1 = q2 - P.ph[1];
    float ta = q0 * 0.866025f + gg.y * 0.5f;
    float tb = -gg.y;
    float t2 = tmax(ta, tb);
    float t3 = t2 - P.ph[0] * 0.5f;
    if (g) {
        v3 ggg = mk(0.f, 0.f, 0.f);
        if (t3 < t1) ggg.z = sgnf(gg.z);
        else if (tb < ta) { ggg.x = sgnf(gg.x) * 0.866025f; ggg.y = 0.5f; }
        else ggg.y = -1.f;
        *g = qrot_rev(iq, p, ggg, nullptr);
    }
    return tmax(t1, t3);
}
DEV float chop_sdf(const PrimC& P, float gap, v3 p, v3* g, float* ggap) {
    v3 pp = mk(p.x, p.y + P.h / 2.f, p.z);
    v3 pa = mk(pp.x - gap / 2.f, pp.y, pp.z), pb = mk(pp.x + gap / 2.f, pp.y, pp.z);
    v3 ga, gb;
    float a = capsule_sdf(P.h, P.r, pa, g ? &ga : nullptr), b = capsule_sdf(P.h, P.r, pb, g ? &gb : nullptr);
    if (g) {
        if (a < b) { *g = ga; if (ggap) *ggap = -ga.x / 2.f; }
        else { *g = gb; if (ggap) *ggap = gb.x / 2.f; }
    }
    return tmin(a, b);
}
// VJP of v / length(v) (length with the 1e-14 epsilon), 2-D and 3-D
DEV void nrm2_rev(float v0, float v1, float go0, float go1, float* gv0, float* gv1) {
    float L = sqrtf(v0 * v0 + v1 * v1 + LEN_EPS_SHAPE), d = v0 * go0 + v1 * go1, L3 = L * L * L;
    *gv0 += go0 / L - v0 * d / L3;
    *gv1 += go1 / L - v1 * d / L3;
}
DEV v3 nrm3_rev(v3 v, v3 go) {
    float L = sqrtf(dot(v, v) + LEN_EPS_SHAPE), d = dot(v, go), L3 = L * L * L;
    return mk(go.x / L - v.x * d / L3, go.y / L - v.y * d / L3, go.z / L - v.z * d / L3);
}
// Cylinder._sdf (primitives.py:300-304): 2-D box SDF of (|(x, z)|, |y|)
// against (h, r): h is the disk radius, r the half height along y
DEV float cylinder_sdf(float h, float r, v3 p, v3* g) {
    float a = sqrtf(p.x * p.x + p.z * p.z + LEN_EPS_SHAPE);
    float d0 = fabsf(a) - h, d1 = fabsf(p.y) - r;
    float m = tmax(d0, d1);
    float e0 = tmax(d0, 0.f), e1 = tmax(d1, 0.f);
    float L = sqrtf(e0 * e0 + e1 * e1 + LEN_EPS_SHAPE);
    if (g) {
        float gd0 = (0.f < d0) ? e0 / L : 0.f, gd1 = (0.f < d1) ? e1 / L : 0.f;
        if (m < 0.f) { if (d1 < d0) gd0 += 1.f; else gd1 += 1.f; }
        float ga = gd0 * sgnf(a);
        *g = mk(ga * p.x / a, gd1 * sgnf(p.y), ga * p.z / a);
    }
    return tmin(m, 0.f) + L;
}
// Cylinder._normal (primitives.py:306-320), analytic; the forward values the
// VJP needs are recomputed by cyl_normal_tr
struct CylTr { float l, d0, d1, n20, n21, m0, m1, q0, q1, s; v3 n3; };
DEV CylTr cyl_normal_tr(float h, float r, v3 p) {
    CylTr t;
    t.l = sqrtf(p.x * p.x + p.z * p.z + LEN_EPS_SHAPE);
    t.d0 = t.l - h;
    t.d1 = fabsf(p.y) - r;
    float f = t.d0 > t.d1 ? 1.f : 0.f, c = tmax(t.d0, t.d1) <= 0.f ? 1.f : 0.f;
    t.n20 = tmax(t.d0, 0.f) + c * f;
    t.n21 = tmax(t.d1, 0.f) + c * (1.f - f);
    float L2 = sqrtf(t.n20 * t.n20 + t.n21 * t.n21 + LEN_EPS_SHAPE);
    t.m0 = t.n20 / L2;
    t.m1 = t.n21 / L2;
    t.q0 = p.x / t.l;
    t.q1 = p.z / t.l;
    t.s = p.y >= 0.f ? 1.f : -1.f;
    t.n3 = mk(t.q0 * t.m0, t.m1 * t.s, t.q1 * t.m0);
    return t;
}
DEV v3 cylinder_normal(float h, float r, v3 p) {
    CylTr t = cyl_normal_tr(h, r, p);
    float L = sqrtf(dot(t.n3, t.n3) + LEN_EPS_SHAPE);
    return mk(t.n3.x / L, t.n3.y / L, t.n3.z / L);
}
DEV v3 cylinder_normal_rev(float h, float r, v3 p, v3 gn) {
    CylTr t = cyl_normal_tr(h, r, p);
    v3 gn3 = nrm3_rev(t.n3, gn);
    float gq0 = gn3.x * t.m0, gq1 = gn3.z * t.m0;
    float gm0 = gn3.x * t.q0 + gn3.z * t.q1, gm1 = gn3.y * t.s;
    float gn20 = 0.f, gn21 = 0.f;
    nrm2_rev(t.n20, t.n21, gm0, gm1, &gn20, &gn21);
    float gd0 = (0.f < t.d0) ? gn20 : 0.f, gd1 = (0.f < t.d1) ? gn21 : 0.f;
    // q = (x, z) / l, d0 = l - h, d1 = |y| - r
    float gl = gd0 - (gq0 * p.x + gq1 * p.z) / (t.l * t.l);
    return mk(gq0 / t.l + gl * p.x / t.l, gd1 * sgnf(p.y), gq1 / t.l + gl * p.z / t.l);
}
// Torus._sdf (primitives.py:336-339): tx the ring radius, ty the tube radius
DEV float torus_sdf(float tx, float ty, v3 p, v3* g) {
    float a = sqrtf(p.x * p.x + p.z * p.z + LEN_EPS_SHAPE);
    float q0 = a - tx, q1 = p.y;
    float L = sqrtf(q0 * q0 + q1 * q1 + LEN_EPS_SHAPE);
    if (g) {
        float ga = q0 / L;
        *g = mk(ga * p.x / a, q1 / L, ga * p.z / a);
    }
    return L - ty;
}
// Torus._normal (primitives.py:341-350), analytic
DEV v3 torus_normal_n3(float tx, v3 p, float* l, float* q0, float* m0, float* m1) {
    *l = sqrtf(p.x * p.x + p.z * p.z + LEN_EPS_SHAPE);
    *q0 = *l - tx;
    float L = sqrtf(*q0 * *q0 + p.y * p.y + LEN_EPS_SHAPE);
    *m0 = *q0 / L;
    *m1 = p.y / L;
    return mk(p.x / *l * *m0, *m1, p.z / *l * *m0);
}
DEV v3 torus_normal(float tx, v3 p) {
    float l, q0, m0, m1;
    v3 n3 = torus_normal_n3(tx, p, &l, &q0, &m0, &m1);
    float L = sqrtf(dot(n3, n3) + LEN_EPS_SHAPE);
    return mk(n3.x / L, n3.y / L, n3.z / L);
}
DEV v3 torus_normal_rev(float tx, v3 p, v3 gn) {
    float l, q0, m0, m1;
    v3 n3 = torus_normal_n3(tx, p, &l, &q0, &m0, &m1);
    v3 gn3 = nrm3_rev(n3, gn);
    float gx0 = gn3.x * m0, gx1 = gn3.z * m0;
    float gm0 = gn3.x * p.x / l + gn3.z * p.z / l, gm1 = gn3.y;
    float gq0 = 0.f, gy = 0.f;
    nrm2_rev(q0, p.y, gm0, gm1, &gq0, &gy);
    float gl = gq0 - (gx0 * p.x + gx1 * p.z) / (l * l);
    return mk(gx0 / l + gl * p.x / l, gy, gx1 / l + gl * p.z / l);
}
DEV float local_sdf(const PrimC& P, int sh, float gap, v3 p, v3* g, float* ggap) {
    if (ggap) *ggap = 0.f;
    switch (sh) {
    case SH_CAPSULE: return capsule_sdf(P.h, P.r, p, g);
    case SH_BOX: return box_sdf(mk(P.size[0], P.size[1], P.size[2]), p, g);
    case SH_PRISM: return prism_sdf(P, p, g);
    case SH_KNIFE: {  // Knife._sdf = max(prism, box) (primitives.py:769-773)
        v3 g1, g2;
        float a = prism_sdf(P, p, g ? &g1 : nullptr);
        float b = box_sdf(mk(P.size[0], P.size[1], P.size[2]), p, g ? &g2 : nullptr);
        if (g) *g = (b < a) ? g1 : g2;
        return tmax(a, b);
    }
    case SH_CHOP: return chop_sdf(P, gap, p, g, ggap);
    case SH_CYL: return cylinder_sdf(P.h, P.r, p, g);
    case SH_TORUS: return torus_sdf(P.h, P.r, p, g);
    default: {  // Sphere.sdf (primitives.py:28-30), world offset p = x - position
        float L = sqrtf(dot(p, p) + LEN_EPS_SHAPE);
        if (g) *g = mk(p.x / L, p.y / L, p.z / L);
        return L - P.r;
    }
    }
}
// ---- fp64 evaluation of the central-difference normals (Box / Prism / Knife).
// In fp32 the shifted positions p +- 1e-4 are rounded to ~3e-8, a ~3e-4
// relative error in the difference quotient that dominates the adjoint error
// of tools with FD normals; the quotient and its adjoint are therefore
// evaluated in fp64 (per grid node only).
struct d3 { double x, y, z; };
DEV double dcomp(const d3& a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
DEV void dsetc(d3& a, int i, double v) { if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v; }
DEV double dmax(double a, double b) { return b < a ? a : b; }
DEV double dmin(double a, double b) { return a < b ? a : b; }
DEV double dsgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }
// (the FD normals' fp64 square roots and divisions take the Newton-refined
// reciprocal square root, a few ulp from correct rounding: the normal is
// rounded to fp32 after a 2e-4 difference, so those ulps never reach it;
// the argument is at least LEN_EPS_SHAPE)
DEV double box_sdf_d(const PrimC& P, d3 p, d3* g) {
    double q0 = fabs(p.x) - (double)P.size[0], q1 = fabs(p.y) - (double)P.size[1], q2 = fabs(p.z) - (double)P.size[2];
    double a0 = dmax(q0, 0.0), a1 = dmax(q1, 0.0), a2 = dmax(q2, 0.0);
    double L2 = a0 * a0 + a1 * a1 + a2 * a2 + (double)LEN_EPS_SHAPE;
    double iL = rot_rsq(L2);
    double L = L2 * iL;
    double m12 = dmax(q1, q2), m = dmax(q0, m12);
    if (g) {
        double g0 = (0.0 < q0) ? a0 * iL : 0.0, g1 = (0.0 < q1) ? a1 * iL : 0.0, g2 = (0.0 < q2) ? a2 * iL : 0.0;
        if (m < 0.0) {
            if (m12 < q0) g0 += 1.0;
            else if (q2 < q1) g1 += 1.0;
            else g2 += 1.0;
        }
        g->x = dsgn(p.x) * g0; g->y = dsgn(p.y) * g1; g->z = dsgn(p.z) * g2;
    }
    return L + dmin(m, 0.0);
}
// qrot with the (fp32) prism rotation, in fp64
DEV d3 qrot_d(double w, double qx, double qy, double qz, d3 v) {
    d3 uv = {qy * v.z - qz * v.y, qz * v.x - qx * v.z, qx * v.y - qy * v.x};
    d3 uuv = {qy * uv.z - qz * uv.y, qz * uv.x - qx * uv.z, qx * uv.y - qy * uv.x};
    d3 o = {v.x + 2.0 * (w * uv.x + uuv.x), v.y + 2.0 * (w * uv.y + uuv.y), v.z + 2.0 * (w * uv.z + uuv.z)};
    return o;
}
DEV double prism_sdf_d(const PrimC& P, d3 p, d3* g) {
    q4 iq = conj_n(mkq(P.prot[0], P.prot[1], P.prot[2], P.prot[3]));
    d3 gg = qrot_d(iq.w, iq.x, iq.y, iq.z, p);
    double q0 = fabs(gg.x), q2 = fabs(gg.z);
    double t1 = q2 - (double)P.ph[1];
    double ta = q0 * (double)0.866025f + gg.y * 0.5;
    double tb = -gg.y;
    double t2 = dmax(ta, tb);
    double t3 = t2 - (double)P.ph[0] * 0.5;
    if (g) {
        d3 ggg = {0.0, 0.0, 0.0};
        if (t3 < t1) ggg.z = dsgn(gg.z);
        else if (tb < ta) { ggg.x = dsgn(gg.x) * (double)0.866025f; ggg.y = 0.5; }
        else ggg.y = -1.0;
        // transpose of the rotation: qrot by the conjugate of the (unit) quaternion
        *g = qrot_d(iq.w, -iq.x, -iq.y, -iq.z, ggg);
    }
    return dmax(t1, t3);
}
DEV double fd_sdf_d(const PrimC& P, int sh, d3 p, d3* g) {
    if (sh == SH_BOX) return box_sdf_d(P, p, g);
    if (sh == SH_PRISM) return prism_sdf_d(P, p, g);
    d3 g1, g2;  // SH_KNIFE: max(prism, box)
    double a = prism_sdf_d(P, p, g ? &g1 : nullptr), b = box_sdf_d(P, p, g ? &g2 : nullptr);
    if (g) *g = (b < a) ? g1 : g2;
    return dmax(a, b);
}
// the normalised central difference; *Lout (when given) receives the
// difference vector's length L, which the reverse needs
DEV v3 fd_normal_d(const PrimC& P, int sh, v3 pf, float* Lout = nullptr) {
    const double d = (double)1e-4f;
    d3 p = {pf.x, pf.y, pf.z}, m;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        d3 inc = p, dec = p;
        dsetc(inc, i, dcomp(p, i) + d);
        dsetc(dec, i, dcomp(p, i) - d);
        dsetc(m, i, (0.5 / d) * (fd_sdf_d(P, sh, inc, nullptr) - fd_sdf_d(P, sh, dec, nullptr)));
    }
    double L2 = m.x * m.x + m.y * m.y + m.z * m.z + (double)LEN_EPS_SHAPE;
    double iL = rot_rsq(L2);
    if (Lout) *Lout = (float)(L2 * iL);
    return mk((float)(m.x * iL), (float)(m.y * iL), (float)(m.z * iL));
}
// adjoint of fd_normal_d from its forward result n = m / L and L:
// dn/dm = (I - n n^T) / L, and dm_i/dp = (grad sdf(p + d e_i) - grad sdf(p - d e_i)) * 0.5 / d,
// the sdf gradients in fp64 (their difference cancels to O(d)), one axis at a time
DEV v3 fd_normal_rev_d(const PrimC& P, int sh, v3 pf, v3 n, float L, v3 gnf) {
    const double d = (double)1e-4f;
    d3 p = {pf.x, pf.y, pf.z};
    double nd = (double)dot(n, gnf);
    double inv = (0.5 / d) / (double)L;
    double gp[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        d3 inc = p, dec = p, gi, gd;
        dsetc(inc, i, dcomp(p, i) + d);
        dsetc(dec, i, dcomp(p, i) - d);
        fd_sdf_d(P, sh, inc, &gi);
        fd_sdf_d(P, sh, dec, &gd);
        double gm = ((double)comp(gnf, i) - (double)comp(n, i) * nd) * inv;
        gp[0] += gm * (gi.x - gd.x);
        gp[1] += gm * (gi.y - gd.y);
        gp[2] += gm * (gi.z - gd.z);
    }
    return mk((float)gp[0], (float)gp[1], (float)gp[2]);
}

// normals: analytic for sphere / capsule / chopsticks / cylinder / torus,
// central FD with the float32 step 1e-4 for box / prism / knife
// (primitives.py:382-393,720-731,796-807)
// *L (when given): the FD length of box-like shapes (1 for the others)
DEV v3 local_normal(const PrimC& P, int sh, float gap, v3 p, float* L = nullptr) {
    if (L) *L = 1.f;
    if (sh == SH_CAPSULE) return capsule_normal(P.h, p);
    if (sh == SH_CYL) return cylinder_normal(P.h, P.r, p);
    if (sh == SH_TORUS) return torus_normal(P.h, p);
    if (sh == SH_SPHERE) { float L = sqrtf(dot(p, p) + LEN_EPS_SHAPE); return mk(p.x / L, p.y / L, p.z / L); }
    if (sh == SH_CHOP) {
        v3 pp = mk(p.x, p.y + P.h / 2.f, p.z);
        v3 pa = mk(pp.x - gap / 2.f, pp.y, pp.z), pb = mk(pp.x + gap / 2.f, pp.y, pp.z);
        float a = capsule_sdf(P.h, P.r, pa, nullptr), b = capsule_sdf(P.h, P.r, pb, nullptr);
        return (a <= b) ? capsule_normal(P.h, pa) : capsule_normal(P.h, pb);
    }
    return fd_normal_d(P, sh, p, L);
}
// n, L: local_normal's result and FD length at p
DEV v3 local_normal_rev(const PrimC& P, int sh, float gap, v3 p, v3 gn, float* ggap, v3 n, float L) {
    if (sh == SH_CAPSULE) return capsule_normal_rev(P.h, p, gn);
    if (sh == SH_CYL) return cylinder_normal_rev(P.h, P.r, p, gn);
    if (sh == SH_TORUS) return torus_normal_rev(P.h, p, gn);
    if (sh == SH_SPHERE) {
        float L = sqrtf(dot(p, p) + LEN_EPS_SHAPE), dd = dot(p, gn), L3 = L * L * L;
        return mk(gn.x / L - p.x * dd / L3, gn.y / L - p.y * dd / L3, gn.z / L - p.z * dd / L3);
    }
    if (sh == SH_CHOP) {
        v3 pp = mk(p.x, p.y + P.h / 2.f, p.z);
        v3 pa = mk(pp.x - gap / 2.f, pp.y, pp.z), pb = mk(pp.x + gap / 2.f, pp.y, pp.z);
        float a = capsule_sdf(P.h, P.r, pa, nullptr), b = capsule_sdf(P.h, P.r, pb, nullptr);
        v3 g3;
        if (a <= b) { g3 = capsule_normal_rev(P.h, pa, gn); *ggap += -g3.x / 2.f; }
        else { g3 = capsule_normal_rev(P.h, pb, gn); *ggap += g3.x / 2.f; }
        return g3;
    }
    return fd_normal_rev_d(P, sh, p, n, L, gn);
}

// ------------------------------------------------------------------ primitive frames
struct PS { v3 pos; q4 rot; float gap; };
struct PSg { v3 pos; q4 rot; float gap; };
DEV PS load_ps(const float* p) {
    PS s;
    s.pos = mk(p[0], p[1], p[2]);
    s.rot = mkq(p[3], p[4], p[5], p[6]);
    s.gap = p[7];
    return s;
}
DEV PSg zero_psg() { PSg g; g.pos = mk(0.f, 0.f, 0.f); g.rot = mkq(0.f, 0.f, 0.f, 0.f); g.gap = 0.f; return g; }

// a collision unit: the primitive itself (jaw 0) or a Gripper / Gripper2 jaw
// at get_pos(f, flag) = pos + qrot(rot, (gap / 2 * flag, 0, 0)) (primitives.py:471-473)
struct Unit { int jaw, sh, norot; float len_eps; };

DEV v3 unit_center(const PS& s, int jaw) {
    if (!jaw) return s.pos;
    return s.pos + qrot(s.rot, mk(s.gap / 2.f * (float)jaw, 0.f, 0.f));
}
DEV void unit_center_rev(const PS& s, int jaw, v3 gc, PSg* g) {
    g->pos += gc;
    if (jaw) {
        v3 go = qrot_rev(s.rot, mk(s.gap / 2.f * (float)jaw, 0.f, 0.f), gc, &g->rot);
        g->gap += go.x * (float)jaw / 2.f;
    }
}
// inv_trans (utils.py:50-54)
DEV v3 to_local(const PS& s, const Unit& u, v3 x) {
    v3 d = x - unit_center(s, u.jaw);
    if (u.norot) return d;
    return qrot(conj_n(s.rot), d);
}
DEV void to_local_rev(const PS& s, const Unit& u, v3 x, v3 gp, PSg* g) {
    v3 d = x - unit_center(s, u.jaw);
    v3 gd;
    if (u.norot) gd = gp;
    else {
        q4 giq = mkq(0.f, 0.f, 0.f, 0.f);
        gd = qrot_rev(conj_n(s.rot), d, gp, &giq);
        conj_n_rev(s.rot, giq, &g->rot);
    }
    unit_center_rev(s, u.jaw, -gd, g);
}

DEV float unit_sdf(const PrimC& P, const Unit& u, const PS& s, v3 x) {
    return local_sdf(P, u.sh, s.gap, to_local(s, u, x), nullptr, nullptr);
}
DEV void unit_sdf_rev(const PrimC& P, const Unit& u, const PS& s, v3 x, float gd, PSg* g) {
    v3 p = to_local(s, u, x), gp;
    float ggap;
    local_sdf(P, u.sh, s.gap, p, &gp, &ggap);
    if (u.sh == SH_CHOP) g->gap += ggap * gd;
    to_local_rev(s, u, x, gd * gp, g);
}
// adjoint of unit_sdf w.r.t. both the primitive state (into g) and the world
// point x (returned): x enters as x - centre, so its adjoint is the offset's
DEV v3 unit_sdf_rev_x(const PrimC& P, const Unit& u, const PS& s, v3 x, float gd, PSg* g) {
    v3 p = to_local(s, u, x), gp;
    float ggap;
    local_sdf(P, u.sh, s.gap, p, &gp, &ggap);
    if (u.sh == SH_CHOP) g->gap += ggap * gd;
    v3 gl = gd * gp;
    v3 gx;
    if (u.norot) gx = gl;
    else {
        q4 giq = mkq(0.f, 0.f, 0.f, 0.f);
        gx = qrot_rev(conj_n(s.rot), x - unit_center(s, u.jaw), gl, &giq);
    }
    to_local_rev(s, u, x, gl, g);
    return gx;
}
// world normal qrot(rot, _normal(local)) (primive_base.py:81-85)
// *nl, *L (when given): the local normal and its FD length, for the reverse
DEV v3 unit_normal(const PrimC& P, const Unit& u, const PS& s, v3 x, v3* nl = nullptr, float* L = nullptr) {
    v3 n = local_normal(P, u.sh, s.gap, to_local(s, u, x), L);
    if (nl) *nl = n;
    return u.norot ? n : qrot(s.rot, n);
}
// nl, L: unit_normal's local normal and FD length at x
DEV void unit_normal_rev(const PrimC& P, const Unit& u, const PS& s, v3 x, v3 gD, PSg* g, v3 nl, float L) {
    v3 p = to_local(s, u, x);
    v3 gn;
    if (u.norot) gn = gD;
    else gn = qrot_rev(s.rot, nl, gD, &g->rot);
    float ggap = 0.f;
    v3 gp = local_normal_rev(P, u.sh, s.gap, p, gn, &ggap, nl, L);
    if (u.sh == SH_CHOP) g->gap += ggap;
    to_local_rev(s, u, x, gp, g);
}
DEV void unit_normal_rev(const PrimC& P, const Unit& u, const PS& s, v3 x, v3 gD, PSg* g) {
    float L;
    v3 nl = local_normal(P, u.sh, s.gap, to_local(s, u, x), &L);
    unit_normal_rev(P, u, s, x, gD, g, nl, L);
}
// collider_v (primive_base.py:87-94, primitives.py:497-502)
DEV v3 unit_collider_v(const Unit& u, const PS& s0, const PS& s1, v3 x, float dt) {
    v3 rel = qrot(conj_n(s0.rot), x - unit_center(s0, u.jaw));
    v3 nw = qrot(s1.rot, rel) + unit_center(s1, u.jaw);
    return mk((nw.x - x.x) / dt, (nw.y - x.y) / dt, (nw.z - x.z) / dt);
}
DEV void unit_collider_v_rev(const Unit& u, const PS& s0, const PS& s1, v3 x, float dt, v3 gvc, PSg* g0, PSg* g1) {
    v3 d = x - unit_center(s0, u.jaw);
    q4 iq = conj_n(s0.rot);
    v3 rel = qrot(iq, d);
    v3 gn = mk(gvc.x / dt, gvc.y / dt, gvc.z / dt);
    unit_center_rev(s1, u.jaw, gn, g1);
    v3 grel = qrot_rev(s1.rot, rel, gn, &g1->rot);
    q4 giq = mkq(0.f, 0.f, 0.f, 0.f);
    v3 gd = qrot_rev(iq, d, grel, &giq);
    conj_n_rev(s0.rot, giq, &g0->rot);
    unit_center_rev(s0, u.jaw, -gd, g0);
}

// the number of collision units of a primitive, and unit j by value (no
// array: a dynamically indexed local array would be placed in LDS, indexed by
// a flat work-item id read from the dispatch packet)
DEV int prim_unit_count(const PrimC& P) { return (P.type == P_GRIPPER || P.type == P_GRIPPER2) ? 2 : 1; }
DEV Unit prim_unit(const PrimC& P, int j) {
    Unit u;
    u.sh = P.unit_shape;
    if (P.type == P_GRIPPER || P.type == P_GRIPPER2) {
        u.jaw = j == 0 ? -1 : 1; u.norot = 0; u.len_eps = LEN_EPS_SHAPE;
    } else {
        u.jaw = 0; u.norot = (P.unit_shape == SH_SPHERE); u.len_eps = LEN_EPS_BASE;
    }
    return u;
}

// Primitive.collide (primive_base.py:96-120) / Gripper.collide2 (primitives.py:508-536)
struct CollTr {
    int branch, flag;
    float dist, ex, infl, nc, mn, vtn, mx, Ln;
    v3 D, vc, u, vt, vt2, nl;   // nl, Ln: the local normal and its FD length
};
DEV v3 unit_collide(const PrimC& P, const Unit& un, const PS& s0, const PS& s1, v3 x, float dt, v3 v, CollTr* tr) {
    CollTr t;
    t.dist = unit_sdf(P, un, s0, x);
    t.ex = expf(-t.dist * P.softness);
    t.infl = tmin(t.ex, 1.f);
    t.branch = (P.softness > 0.f && t.infl > 0.1f) || t.dist <= 0.f;
    v3 out = v;
    if (t.branch) {
        t.D = unit_normal(P, un, s0, x, &t.nl, &t.Ln);
        t.vc = unit_collider_v(un, s0, s1, x, dt);
        t.u = v - t.vc;
        t.nc = dot(t.u, t.D);
        t.mn = tmin(t.nc, 0.f);
        t.vt = t.u - t.mn * t.D;
        t.vtn = sqrtf(dot(t.vt, t.vt) + un.len_eps);
        t.mx = tmax(0.f, t.vtn + t.nc * P.friction);
        t.flag = (t.nc < 0.f) && (sqrtf(dot(t.vt, t.vt)) > 1e-30f);
        float fl = t.flag ? 1.f : 0.f;
        v3 vtf = mk(t.vt.x / t.vtn * t.mx, t.vt.y / t.vtn * t.mx, t.vt.z / t.vtn * t.mx);
        t.vt2 = vtf * fl + t.vt * (1.f - fl);
        out = t.vc + t.u * (1.f - t.infl) + t.vt2 * t.infl;
    }
    if (tr) *tr = t;
    return out;
}
DEV v3 unit_collide_rev(const PrimC& P, const Unit& un, const PS& s0, const PS& s1, v3 x, float dt, v3 v, v3 gout,
                        PSg* g0, PSg* g1) {
    CollTr t;
    unit_collide(P, un, s0, s1, x, dt, v, &t);
    if (!t.branch) return gout;
    v3 gvc = gout;
    v3 gu = gout * (1.f - t.infl);
    float ginfl = dot(gout, t.vt2 - t.u);
    v3 gvt2 = gout * t.infl;
    v3 gvt = mk(0.f, 0.f, 0.f);
    float gvtn = 0.f, gnc = 0.f;
    if (t.flag) {
        v3 gq = gvt2 * t.mx;
        float gmx = dot(gvt2, mk(t.vt.x / t.vtn, t.vt.y / t.vtn, t.vt.z / t.vtn));
        gvt += mk(gq.x / t.vtn, gq.y / t.vtn, gq.z / t.vtn);
        gvtn -= dot(gq, t.vt) / (t.vtn * t.vtn);
        if (!((t.vtn + t.nc * P.friction) < 0.f)) { gvtn += gmx; gnc += gmx * P.friction; }
    } else {
        gvt += gvt2;
    }
    gvt += (gvtn / t.vtn) * t.vt;
    gu += gvt;
    float gmn = -dot(gvt, t.D);
    v3 gD = (-t.mn) * gvt;
    if (t.nc < 0.f) gnc += gmn;
    gu += gnc * t.D;
    gD += gnc * t.u;
    v3 gv = gu;
    gvc -= gu;
    unit_normal_rev(P, un, s0, x, gD, g0, t.nl, t.Ln);
    unit_collider_v_rev(un, s0, s1, x, dt, gvc, g0, g1);
    if (t.ex < 1.f) {
        float gdist = ginfl * t.ex * (-P.softness);
        if (gdist != 0.f) unit_sdf_rev(P, un, s0, x, gdist, g0);
    }
    return gv;
}

// ------------------------------------------------------------------ primitive kinematics
DEV float clampv(float a, float lo, float hi) { return tmax(tmin(a, hi), lo); }
DEV float clampv_grad(float a, float lo, float hi) { return (a < hi && lo < tmin(a, hi)) ? 1.f : 0.f; }

// forward_kinematics: primive_base.py:152-156; RollingPin primitives.py:103-117;
// RollingPinExt :121-136; Chopsticks :230-235; Gripper :456-460; Gripper2 :588-592
DEV void prim_fk(const PrimC& P, const float* s, const float* pv, float* o) {
    PS st = load_ps(s);
    v3 pos;
    q4 rot;
    float gap = st.gap;
    if (P.type == P_ROLLINGPIN || P.type == P_ROLLINGPIN_EXT) {
        float dw = pv[0], dth = pv[1], dy = pv[2];
        v3 ydir = qrot(st.rot, mk(0.f, -1.f, 0.f));
        float k = (P.type == P_ROLLINGPIN) ? dw * 0.03f : dw * 0.03f + pv[3];
        v3 xdir = mk(ydir.z * k, dy, -ydir.x * k);
        rot = qmul(w2quat(mk(0.f, -dth, 0.f)), qmul(st.rot, w2quat(mk(0.f, dw, 0.f))));
        pos = mk(clampv(st.pos.x + xdir.x, P.lo[0], P.hi[0]), clampv(st.pos.y + xdir.y, P.lo[1], P.hi[1]),
                 clampv(st.pos.z + xdir.z, P.lo[2], P.hi[2]));
    } else {
        pos = mk(clampv(st.pos.x + pv[0], P.lo[0], P.hi[0]), clampv(st.pos.y + pv[1], P.lo[1], P.hi[1]),
                 clampv(st.pos.z + pv[2], P.lo[2], P.hi[2]));
        q4 wq = w2quat(mk(pv[3], pv[4], pv[5]));
        if (P.type == P_GRIPPER || P.type == P_GRIPPER2 || P.type == P_CHOPSTICKS) {
            rot = qmul(st.rot, wq);
            float g = tmax(st.gap - pv[6], P.min_gap);
            gap = (P.type == P_CHOPSTICKS) ? g : tmin(g, P.max_gap);
        } else {
            rot = qmul(wq, st.rot);
        }
    }
    o[0] = pos.x; o[1] = pos.y; o[2] = pos.z;
    o[3] = rot.w; o[4] = rot.x; o[5] = rot.y; o[6] = rot.z;
    o[7] = gap;
}
// reverse: go = adjoint of frame f+1 state; adds into gs (frame f) and gpv
DEV void prim_fk_rev(const PrimC& P, const float* s, const float* pv, const float* go, float* gs, float* gpv) {
    PS st = load_ps(s);
    q4 gq = mkq(0.f, 0.f, 0.f, 0.f);
    q4 gro = mkq(go[3], go[4], go[5], go[6]);
    if (P.type == P_ROLLINGPIN || P.type == P_ROLLINGPIN_EXT) {
        float dw = pv[0], dth = pv[1], dy = pv[2];
        v3 ny = mk(0.f, -1.f, 0.f);
        v3 ydir = qrot(st.rot, ny);
        float k = (P.type == P_ROLLINGPIN) ? dw * 0.03f : dw * 0.03f + pv[3];
        v3 xdir = mk(ydir.z * k, dy, -ydir.x * k);
        float gx0 = go[0] * clampv_grad(st.pos.x + xdir.x, P.lo[0], P.hi[0]);
        float gx1 = go[1] * clampv_grad(st.pos.y + xdir.y, P.lo[1], P.hi[1]);
        float gx2 = go[2] * clampv_grad(st.pos.z + xdir.z, P.lo[2], P.hi[2]);
        gs[0] += gx0; gs[1] += gx1; gs[2] += gx2;
        float gk = gx0 * ydir.z - gx2 * ydir.x;
        gpv[2] += gx1;
        qrot_rev(st.rot, ny, mk(-gx2 * k, 0.f, gx0 * k), &gq);
        gpv[0] += gk * 0.03f;
        if (P.type == P_ROLLINGPIN_EXT) gpv[3] += gk;
        v3 a1 = mk(0.f, -dth, 0.f), a2 = mk(0.f, dw, 0.f);
        q4 q1 = w2quat(a1), q2 = w2quat(a2);
        q4 t = qmul(st.rot, q2);
        q4 gq1 = mkq(0.f, 0.f, 0.f, 0.f), gt = mkq(0.f, 0.f, 0.f, 0.f), gq2 = mkq(0.f, 0.f, 0.f, 0.f);
        qmul_rev(q1, t, gro, &gq1, &gt);
        qmul_rev(st.rot, q2, gt, &gq, &gq2);
        v3 ga1 = w2quat_rev(a1, gq1), ga2 = w2quat_rev(a2, gq2);
        gpv[1] -= ga1.y;
        gpv[0] += ga2.y;
        gs[7] += go[7];
    } else {
        float g0 = go[0] * clampv_grad(st.pos.x + pv[0], P.lo[0], P.hi[0]);
        float g1 = go[1] * clampv_grad(st.pos.y + pv[1], P.lo[1], P.hi[1]);
        float g2 = go[2] * clampv_grad(st.pos.z + pv[2], P.lo[2], P.hi[2]);
        gs[0] += g0; gs[1] += g1; gs[2] += g2;
        gpv[0] += g0; gpv[1] += g1; gpv[2] += g2;
        v3 w = mk(pv[3], pv[4], pv[5]);
        q4 wq = w2quat(w), gwq = mkq(0.f, 0.f, 0.f, 0.f);
        if (P.type == P_GRIPPER || P.type == P_GRIPPER2 || P.type == P_CHOPSTICKS) {
            qmul_rev(st.rot, wq, gro, &gq, &gwq);
            float a = st.gap - pv[6];
            float g1g = tmax(a, P.min_gap);
            float gg = go[7];
            if (P.type != P_CHOPSTICKS) gg = (g1g < P.max_gap) ? gg : 0.f;
            gg = (P.min_gap < a) ? gg : 0.f;
            gs[7] += gg;
            gpv[6] -= gg;
        } else {
            qmul_rev(wq, st.rot, gro, &gwq, &gq);
            gs[7] += go[7];
        }
        v3 gw = w2quat_rev(w, gwq);
        gpv[3] += gw.x; gpv[4] += gw.y; gpv[5] += gw.z;
    }
    gs[3] += gq.w; gs[4] += gq.x; gs[5] += gq.y; gs[6] += gq.z;
}

}  // namespace dsk
