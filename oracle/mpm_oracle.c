/*
 * CPU oracle for the differentiable MLS-MPM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the parity tests, the
 * smoke check and the CPU-baseline leg of the benchmark load it.  No product
 * code path may call it.
 *
 * It restates, in plain C, the algorithm of the reference simulator
 *   plb/engine/mpm_simulator.py           (MPMSimulator kernels)
 *   plb/engine/primitive/primive_base.py  (Primitive.collide / collider_v / FK)
 *   plb/engine/primitive/primitives.py    (shape SDFs, normals, kinematics)
 *   plb/engine/primitive/utils.py         (qrot / qmul / w2quat / inv_trans)
 * together with a hand-derived reverse-mode adjoint that follows the
 * semantics of Taichi 0.7.26 reverse-mode autodiff (the third-party
 * dependency that generates the reference's .grad kernels): only the taken
 * branch of an `if` is differentiated, ti.max(a, b) sends its adjoint to a
 * iff b < a (else to b), ti.min(a, b) sends it to a iff a < b (else to b),
 * ti.abs'(x) = sgn(x) with sgn(0) = 0, and integer casts carry no gradient.
 *
 * Build twice: ORC_REAL=double (liborc64.so, the parity oracle) and
 * ORC_REAL=float (liborc32.so, fp32 like the reference's ti.f32 and used as
 * the timed CPU baseline).  Scatter phases use OpenMP atomics like the
 * reference's implicit atomic `+=`; orc_set_threads(1) gives a deterministic
 * fixed-order scatter.
 *
 * Particle state layout (one frame): n x 24 reals, per particle
 *   x[3] v[3] C[3][3] F[3][3]   (row-major matrices)
 * Primitive frame state: 8 reals  pos[3] rot[4] (w,x,y,z) gap
 * Primitive per-substep velocity: 7 reals  v[3] w[3] gap_vel
 * Grid: n_grid^3 nodes, node (i,j,k) at ((i*n + j)*n + k), world pos I*dx.
 */
#include <tgmath.h>
#undef I /* complex.h defines I; this file uses it as the grid-node index */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;

#define NPS 8   /* primitive frame-state reals */
#define NPV 7   /* primitive velocity reals */
#define NST 24  /* particle state reals */

/* ------------------------------------------------------------------ */
/* configuration structs (host doubles, mirrored by ctypes)            */
/* ------------------------------------------------------------------ */
typedef struct {
    int n_grid;
    int n_prims;
    double dx, inv_dx, dt, p_vol, p_mass;
    double mu, lam, yield_stress;
    double gravity[3];
    double ground_friction;
    double lower_bound;
} orc_sim;

enum {
    ORC_SPHERE = 0, ORC_CAPSULE = 1, ORC_ROLLINGPIN = 2, ORC_ROLLINGPIN_EXT = 3,
    ORC_BOX = 4, ORC_GRIPPER = 5, ORC_PRISM = 6, ORC_KNIFE = 7,
    ORC_CHOPSTICKS = 8, ORC_GRIPPER2 = 9, ORC_CYLINDER = 10, ORC_TORUS = 11
};

typedef struct {
    int type;
    int action_dim;
    double friction, softness;
    double size[3];   /* Box / Gripper half extents; Knife box size */
    double h, r;      /* Capsule family: segment length h, radius r; Sphere radius r;
                         Cylinder h (disk radius), r (half height); Torus tx in h, ty in r */
    double ph[2];     /* Prism / Knife prism h */
    double prot[4];   /* Prism / Knife prism rotation (w,x,y,z) */
    double lo[3], hi[3];  /* xyz_limit */
    double min_gap, max_gap;
    double action_scale[7];
} orc_prim;

/* tool-tool collision pairs of one scene and their per-substep records */
typedef struct {
    int n_pairs, n_points;
    const int *pairs;   /* n_pairs x 2: (i, j), primitive i projected out of box j */
    const real *rand;   /* n_pairs x n_points x 3 */
    int *idx;           /* n_pairs: collision_idx (-1: none) */
    real *pt;           /* n_pairs x 3: the chosen surface point */
    real *pre;          /* n_pairs x 3: position of i before the pair's projection */
} orc_coll;

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_real_size(void) { return (int)sizeof(real); }

/* ------------------------------------------------------------------ */
/* small vector algebra                                                */
/* ------------------------------------------------------------------ */
static inline real dot3(const real *a, const real *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const real *a, const real *b, real *c) {
    real t0 = a[1] * b[2] - a[2] * b[1], t1 = a[2] * b[0] - a[0] * b[2], t2 = a[0] * b[1] - a[1] * b[0];
    c[0] = t0; c[1] = t1; c[2] = t2;
}
static inline real sgn(real x) { return x > 0 ? (real)1 : (x < 0 ? (real)-1 : (real)0); }
/* Taichi max/min value (ties give the same value either way) */
static inline real tmax(real a, real b) { return b < a ? a : b; }
static inline real tmin(real a, real b) { return a < b ? a : b; }

static void matmul3(const real *A, const real *B, real *C) {
    real T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            T[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
    memcpy(C, T, sizeof(T));
}
static void matmul3_tn(const real *A, const real *B, real *C) { /* A^T B */
    real T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            T[i * 3 + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
    memcpy(C, T, sizeof(T));
}
static void matmul3_nt(const real *A, const real *B, real *C) { /* A B^T */
    real T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            T[i * 3 + j] = A[i * 3] * B[j * 3] + A[i * 3 + 1] * B[j * 3 + 1] + A[i * 3 + 2] * B[j * 3 + 2];
    memcpy(C, T, sizeof(T));
}
static real det3(const real *F) {
    return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) + F[2] * (F[3] * F[7] - F[4] * F[6]);
}
static void cof3(const real *F, real *Cf) { /* d det / dF */
    Cf[0] = F[4] * F[8] - F[5] * F[7];
    Cf[1] = F[5] * F[6] - F[3] * F[8];
    Cf[2] = F[3] * F[7] - F[4] * F[6];
    Cf[3] = F[2] * F[7] - F[1] * F[8];
    Cf[4] = F[0] * F[8] - F[2] * F[6];
    Cf[5] = F[1] * F[6] - F[0] * F[7];
    Cf[6] = F[1] * F[5] - F[2] * F[4];
    Cf[7] = F[2] * F[3] - F[0] * F[5];
    Cf[8] = F[0] * F[4] - F[1] * F[3];
}

/* ------------------------------------------------------------------ */
/* quaternion helpers: plb/engine/primitive/utils.py:9-54              */
/* ------------------------------------------------------------------ */
/* qrot (utils.py:9-15): v + 2 (w * (q_v x v) + q_v x (q_v x v)) */
static void qrot(const real *q, const real *v, real *out) {
    real qv[3] = {q[1], q[2], q[3]}, uv[3], uuv[3];
    cross3(qv, v, uv);
    cross3(qv, uv, uuv);
    for (int i = 0; i < 3; i++) out[i] = v[i] + 2 * (q[0] * uv[i] + uuv[i]);
}
static void qrot_rev(const real *q, const real *v, const real *go, real *gq, real *gv) {
    real qv[3] = {q[1], q[2], q[3]}, uv[3], guv[3], guuv[3], t[3], gqv[3] = {0, 0, 0};
    cross3(qv, v, uv);
    for (int i = 0; i < 3; i++) { guuv[i] = 2 * go[i]; guv[i] = 2 * q[0] * go[i]; }
    if (gq) gq[0] += 2 * dot3(go, uv);
    cross3(uv, guuv, t); for (int i = 0; i < 3; i++) gqv[i] += t[i];
    cross3(guuv, qv, t); for (int i = 0; i < 3; i++) guv[i] += t[i];
    cross3(v, guv, t); for (int i = 0; i < 3; i++) gqv[i] += t[i];
    if (gv) {
        cross3(guv, qv, t);
        for (int i = 0; i < 3; i++) gv[i] += t[i] + go[i];
    }
    if (gq) for (int i = 0; i < 3; i++) gq[i + 1] += gqv[i];
}
/* Taichi Vector.normalized(): v / sqrt(v.v) of the conjugate (utils.py:50-54, primive_base.py:89) */
static void conj_normalized(const real *q, real *iq) {
    real c[4] = {q[0], -q[1], -q[2], -q[3]};
    real n = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2] + c[3] * c[3]);
    for (int i = 0; i < 4; i++) iq[i] = c[i] / n;
}
static void conj_normalized_rev(const real *q, const real *giq, real *gq) {
    real c[4] = {q[0], -q[1], -q[2], -q[3]};
    real n = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2] + c[3] * c[3]);
    real d = 0;
    for (int i = 0; i < 4; i++) d += c[i] * giq[i];
    real gc[4];
    for (int i = 0; i < 4; i++) gc[i] = giq[i] / n - c[i] * d / (n * n * n);
    gq[0] += gc[0]; gq[1] -= gc[1]; gq[2] -= gc[2]; gq[3] -= gc[3];
}
/* qmul (utils.py:18-25): Hamilton product q (x) r, renormalised */
static void qmul(const real *q, const real *r, real *out) {
    real o[4];
    o[0] = r[0] * q[0] - r[1] * q[1] - r[2] * q[2] - r[3] * q[3];
    o[1] = r[0] * q[1] + r[1] * q[0] - r[2] * q[3] + r[3] * q[2];
    o[2] = r[0] * q[2] + r[1] * q[3] + r[2] * q[0] - r[3] * q[1];
    o[3] = r[0] * q[3] - r[1] * q[2] + r[2] * q[1] + r[3] * q[0];
    real n = sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2] + o[3] * o[3]);
    for (int i = 0; i < 4; i++) out[i] = o[i] / n;
}
static void qmul_rev(const real *q, const real *r, const real *gout, real *gq, real *gr) {
    real o[4];
    o[0] = r[0] * q[0] - r[1] * q[1] - r[2] * q[2] - r[3] * q[3];
    o[1] = r[0] * q[1] + r[1] * q[0] - r[2] * q[3] + r[3] * q[2];
    o[2] = r[0] * q[2] + r[1] * q[3] + r[2] * q[0] - r[3] * q[1];
    o[3] = r[0] * q[3] - r[1] * q[2] + r[2] * q[1] + r[3] * q[0];
    real n = sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2] + o[3] * o[3]);
    real d = 0;
    for (int i = 0; i < 4; i++) d += o[i] * gout[i];
    real g[4];
    for (int i = 0; i < 4; i++) g[i] = gout[i] / n - o[i] * d / (n * n * n);
    /* o is bilinear: d o_k / d q_j and d o_k / d r_j */
    gq[0] += g[0] * r[0] + g[1] * r[1] + g[2] * r[2] + g[3] * r[3];
    gq[1] += -g[0] * r[1] + g[1] * r[0] - g[2] * r[3] + g[3] * r[2];
    gq[2] += -g[0] * r[2] + g[1] * r[3] + g[2] * r[0] - g[3] * r[1];
    gq[3] += -g[0] * r[3] - g[1] * r[2] + g[2] * r[1] + g[3] * r[0];
    gr[0] += g[0] * q[0] + g[1] * q[1] + g[2] * q[2] + g[3] * q[3];
    gr[1] += -g[0] * q[1] + g[1] * q[0] + g[2] * q[3] - g[3] * q[2];
    gr[2] += -g[0] * q[2] - g[1] * q[3] + g[2] * q[0] + g[3] * q[1];
    gr[3] += -g[0] * q[3] + g[1] * q[2] - g[2] * q[1] + g[3] * q[0];
}
/* w2quat (utils.py:28-40) */
static void w2quat(const real *a, real *out) {
    real w = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + (real)1e-16);
    out[0] = 1; out[1] = out[2] = out[3] = 0;
    if (w > (real)1e-9) {
        real s = sin(w / 2);
        for (int i = 0; i < 3; i++) out[i + 1] = (a[i] / w) * s;
        out[0] = cos(w / 2);
    }
}
static void w2quat_rev(const real *a, const real *gout, real *ga) {
    real w = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + (real)1e-16);
    if (!(w > (real)1e-9)) return;
    real s = sin(w / 2), c = cos(w / 2);
    real gw = -s / 2 * gout[0];
    real ga_dot = 0;
    for (int i = 0; i < 3; i++) {
        ga[i] += gout[i + 1] * s / w;
        ga_dot += gout[i + 1] * a[i];
    }
    /* d/dw of (a_i / w) * s = -a_i s / w^2 + a_i / w * c / 2 */
    gw += ga_dot * (c / (2 * w) - s / (w * w));
    for (int i = 0; i < 3; i++) ga[i] += gw * a[i] / w;
}

/* ------------------------------------------------------------------ */
/* shape SDFs in the primitive's local frame, with d sdf / d p         */
/* plb/engine/primitive/primitives.py                                  */
/* ------------------------------------------------------------------ */
#define LEN_EPS_SHAPE ((real)1e-14)   /* primitives.py:14-15 */
#define LEN_EPS_BASE ((real)1e-8)     /* utils.py:5-6 */

/* Box._sdf (primitives.py:369-374) */
static real box_sdf(const real *s, const real *p, real *g) {
    real q[3], a[3];
    for (int i = 0; i < 3; i++) { q[i] = fabs(p[i]) - s[i]; a[i] = tmax(q[i], 0); }
    real L = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + LEN_EPS_SHAPE);
    real m12 = tmax(q[1], q[2]);
    real m = tmax(q[0], m12);
    real t = tmin(m, 0);
    if (g) {
        real gq[3];
        for (int i = 0; i < 3; i++) gq[i] = (0 < q[i]) ? a[i] / L : 0;
        if (m < 0) {
            if (m12 < q[0]) gq[0] += 1;
            else if (q[2] < q[1]) gq[1] += 1;
            else gq[2] += 1;
        }
        for (int i = 0; i < 3; i++) g[i] = sgn(p[i]) * gq[i];
    }
    return L + t;
}
/* Capsule._sdf (primitives.py:54-59); local y is the segment axis */
static void capsule_p2(real h, const real *p, real *v, real *dy) {
    real y = p[1] + h / 2;
    real mx = tmax(y, 0);
    real c = tmin(mx, h);
    v[0] = p[0]; v[1] = y - c; v[2] = p[2];
    *dy = 1 - ((0 < y) && (mx < h) ? 1 : 0);
}
static real capsule_sdf(real h, real r, const real *p, real *g) {
    real v[3], dy;
    capsule_p2(h, p, v, &dy);
    real L = sqrt(dot3(v, v) + LEN_EPS_SHAPE);
    if (g) { g[0] = v[0] / L; g[1] = v[1] / L * dy; g[2] = v[2] / L; }
    return L - r;
}
/* Capsule._normal (primitives.py:61-66) and its VJP */
static void capsule_normal(real h, const real *p, real *n) {
    real v[3], dy;
    capsule_p2(h, p, v, &dy);
    real L = sqrt(dot3(v, v) + LEN_EPS_SHAPE);
    for (int i = 0; i < 3; i++) n[i] = v[i] / L;
}
static void capsule_normal_rev(real h, const real *p, const real *gn, real *gp) {
    real v[3], dy;
    capsule_p2(h, p, v, &dy);
    real L = sqrt(dot3(v, v) + LEN_EPS_SHAPE);
    real d = dot3(v, gn);
    real gv[3];
    for (int i = 0; i < 3; i++) gv[i] = gn[i] / L - v[i] * d / (L * L * L);
    gp[0] += gv[0]; gp[1] += gv[1] * dy; gp[2] += gv[2];
}
/* Prism._sdf (primitives.py:711-719) */
static real prism_sdf(const real *ph, const real *prot, const real *p, real *g) {
    real iq[4], gg[3];
    conj_normalized(prot, iq);
    qrot(iq, p, gg);
    real q0 = fabs(gg[0]), q2 = fabs(gg[2]);
    real t1 = q2 - ph[1];
    real ta = q0 * (real)0.866025 + gg[1] * (real)0.5;
    real tb = -gg[1];
    real t2 = tmax(ta, tb);
    real t3 = t2 - ph[0] * (real)0.5;
    real out = tmax(t1, t3);
    if (g) {
        real ggg[3] = {0, 0, 0};
        if (t3 < t1) ggg[2] += sgn(gg[2]);
        else if (tb < ta) { ggg[0] += sgn(gg[0]) * (real)0.866025; ggg[1] += (real)0.5; }
        else ggg[1] -= 1;
        g[0] = g[1] = g[2] = 0;
        qrot_rev(iq, p, ggg, NULL, g);
    }
    return out;
}
/* Knife._sdf (primitives.py:769-773): max(prism, box) */
static real knife_sdf(const orc_prim *P, const real *p, real *g) {
    real s[3] = {P->size[0], P->size[1], P->size[2]};
    real ph[2] = {P->ph[0], P->ph[1]};
    real prot[4] = {P->prot[0], P->prot[1], P->prot[2], P->prot[3]};
    real g1[3], g2[3];
    real a = prism_sdf(ph, prot, p, g ? g1 : NULL);
    real b = box_sdf(s, p, g ? g2 : NULL);
    if (g) { if (b < a) memcpy(g, g1, sizeof(g1)); else memcpy(g, g2, sizeof(g2)); }
    return tmax(a, b);
}

/* v / length(v), length with the 1e-14 epsilon, and its VJP (adds into gv) */
static void nrm2(const real *v, real *o) {
    real L = sqrt(v[0] * v[0] + v[1] * v[1] + LEN_EPS_SHAPE);
    o[0] = v[0] / L; o[1] = v[1] / L;
}
static void nrm2_rev(const real *v, const real *go, real *gv) {
    real L = sqrt(v[0] * v[0] + v[1] * v[1] + LEN_EPS_SHAPE), d = v[0] * go[0] + v[1] * go[1];
    for (int i = 0; i < 2; i++) gv[i] += go[i] / L - v[i] * d / (L * L * L);
}
static void nrm3_rev(const real *v, const real *go, real *gv) {
    real L = sqrt(dot3(v, v) + LEN_EPS_SHAPE), d = dot3(v, go);
    for (int i = 0; i < 3; i++) gv[i] += go[i] / L - v[i] * d / (L * L * L);
}

/* Cylinder._sdf (primitives.py:300-304): the 2-D box SDF of (|(x, z)|, |y|)
 * against (h, r) -- h is the disk radius, r the half height along y */
static real cylinder_sdf(real h, real r, const real *p, real *g) {
    real a = sqrt(p[0] * p[0] + p[2] * p[2] + LEN_EPS_SHAPE);
    real d0 = fabs(a) - h, d1 = fabs(p[1]) - r;
    real m = tmax(d0, d1), t = tmin(m, 0);
    real e0 = tmax(d0, 0), e1 = tmax(d1, 0);
    real L = sqrt(e0 * e0 + e1 * e1 + LEN_EPS_SHAPE);
    if (g) {
        real gd0 = (0 < d0) ? e0 / L : 0, gd1 = (0 < d1) ? e1 / L : 0;
        if (m < 0) { if (d1 < d0) gd0 += 1; else gd1 += 1; }
        real ga = gd0 * sgn(a);
        g[0] = ga * p[0] / a; g[1] = gd1 * sgn(p[1]); g[2] = ga * p[2] / a;
    }
    return t + L;
}
/* Cylinder._normal (primitives.py:306-320), analytic */
typedef struct { real pp[2], l, d0, d1, n2[2], n2n[2], p2[2], s, n3[3]; } cyl_tr;
static void cylinder_normal_tr(real h, real r, const real *p, cyl_tr *t) {
    t->pp[0] = p[0]; t->pp[1] = p[2];
    t->l = sqrt(p[0] * p[0] + p[2] * p[2] + LEN_EPS_SHAPE);
    t->d0 = t->l - h; t->d1 = fabs(p[1]) - r;
    real f = t->d0 > t->d1 ? 1 : 0, c = tmax(t->d0, t->d1) <= 0 ? 1 : 0;
    t->n2[0] = tmax(t->d0, 0) + c * f; t->n2[1] = tmax(t->d1, 0) + c * (1 - f);
    nrm2(t->n2, t->n2n);
    t->p2[0] = t->pp[0] / t->l; t->p2[1] = t->pp[1] / t->l;
    t->s = p[1] >= 0 ? 1 : -1;
    t->n3[0] = t->p2[0] * t->n2n[0]; t->n3[1] = t->n2n[1] * t->s; t->n3[2] = t->p2[1] * t->n2n[0];
}
static void cylinder_normal(real h, real r, const real *p, real *n) {
    cyl_tr t;
    cylinder_normal_tr(h, r, p, &t);
    real L = sqrt(dot3(t.n3, t.n3) + LEN_EPS_SHAPE);
    for (int i = 0; i < 3; i++) n[i] = t.n3[i] / L;
}
static void cylinder_normal_rev(real h, real r, const real *p, const real *gn, real *gp) {
    cyl_tr t;
    cylinder_normal_tr(h, r, p, &t);
    real gn3[3] = {0, 0, 0}, gn2n[2], gn2[2] = {0, 0}, gp2[2];
    nrm3_rev(t.n3, gn, gn3);
    gp2[0] = gn3[0] * t.n2n[0]; gp2[1] = gn3[2] * t.n2n[0];
    gn2n[0] = gn3[0] * t.p2[0] + gn3[2] * t.p2[1]; gn2n[1] = gn3[1] * t.s;
    nrm2_rev(t.n2, gn2n, gn2);
    real gd0 = (0 < t.d0) ? gn2[0] : 0, gd1 = (0 < t.d1) ? gn2[1] : 0;
    /* p2 = pp / l; d0 = l - h; d1 = |y| - r */
    real gl = gd0 - (gp2[0] * t.pp[0] + gp2[1] * t.pp[1]) / (t.l * t.l);
    real gpp0 = gp2[0] / t.l + gl * t.pp[0] / t.l, gpp1 = gp2[1] / t.l + gl * t.pp[1] / t.l;
    gp[0] += gpp0; gp[1] += gd1 * sgn(p[1]); gp[2] += gpp1;
}
/* Torus._sdf (primitives.py:336-339): tx the ring radius, ty the tube radius */
static real torus_sdf(real tx, real ty, const real *p, real *g) {
    real a = sqrt(p[0] * p[0] + p[2] * p[2] + LEN_EPS_SHAPE);
    real q[2] = {a - tx, p[1]};
    real L = sqrt(q[0] * q[0] + q[1] * q[1] + LEN_EPS_SHAPE);
    if (g) {
        real ga = q[0] / L;
        g[0] = ga * p[0] / a; g[1] = q[1] / L; g[2] = ga * p[2] / a;
    }
    return L - ty;
}
/* Torus._normal (primitives.py:341-350), analytic */
static void torus_normal_tr(real tx, const real *p, real *x, real *l, real *q, real *n2, real *x2, real *n3) {
    x[0] = p[0]; x[1] = p[2];
    *l = sqrt(x[0] * x[0] + x[1] * x[1] + LEN_EPS_SHAPE);
    q[0] = *l - tx; q[1] = p[1];
    nrm2(q, n2);
    x2[0] = x[0] / *l; x2[1] = x[1] / *l;
    n3[0] = x2[0] * n2[0]; n3[1] = n2[1]; n3[2] = x2[1] * n2[0];
}
static void torus_normal(real tx, const real *p, real *n) {
    real x[2], l, q[2], n2[2], x2[2], n3[3];
    torus_normal_tr(tx, p, x, &l, q, n2, x2, n3);
    real L = sqrt(dot3(n3, n3) + LEN_EPS_SHAPE);
    for (int i = 0; i < 3; i++) n[i] = n3[i] / L;
}
static void torus_normal_rev(real tx, const real *p, const real *gn, real *gp) {
    real x[2], l, q[2], n2[2], x2[2], n3[3];
    torus_normal_tr(tx, p, x, &l, q, n2, x2, n3);
    real gn3[3] = {0, 0, 0}, gx2[2], gn2[2], gq[2] = {0, 0};
    nrm3_rev(n3, gn, gn3);
    gx2[0] = gn3[0] * n2[0]; gx2[1] = gn3[2] * n2[0];
    gn2[0] = gn3[0] * x2[0] + gn3[2] * x2[1]; gn2[1] = gn3[1];
    nrm2_rev(q, gn2, gq);
    real gl = gq[0] - (gx2[0] * x[0] + gx2[1] * x[1]) / (l * l);
    gp[0] += gx2[0] / l + gl * x[0] / l; gp[1] += gq[1]; gp[2] += gx2[1] / l + gl * x[1] / l;
}

/* local SDF of one "collision unit" shape (the shape whose SDF the unit tests) */
enum { SH_SPHERE, SH_CAPSULE, SH_BOX, SH_PRISM, SH_KNIFE, SH_CHOP, SH_CYL, SH_TORUS };
static int unit_shape(int type) {
    switch (type) {
    case ORC_SPHERE: return SH_SPHERE;
    case ORC_CAPSULE: case ORC_ROLLINGPIN: case ORC_ROLLINGPIN_EXT: case ORC_GRIPPER2: return SH_CAPSULE;
    case ORC_BOX: case ORC_GRIPPER: return SH_BOX;
    case ORC_PRISM: return SH_PRISM;
    case ORC_KNIFE: return SH_KNIFE;
    case ORC_CHOPSTICKS: return SH_CHOP;
    case ORC_CYLINDER: return SH_CYL;
    case ORC_TORUS: return SH_TORUS;
    }
    return SH_BOX;
}

/* Chopsticks._sdf (primitives.py:248-254): min of two capsules offset by gap/2 */
static real chop_sdf(const orc_prim *P, real gap, const real *p, real *g, real *ggap) {
    real h = P->h, r = P->r;
    real pp[3] = {p[0], p[1] + h / 2, p[2]};
    real pa[3] = {pp[0] - gap / 2, pp[1], pp[2]}, pb[3] = {pp[0] + gap / 2, pp[1], pp[2]};
    real ga[3], gb[3];
    real a = capsule_sdf(h, r, pa, ga), b = capsule_sdf(h, r, pb, gb);
    if (g) {
        if (a < b) { memcpy(g, ga, sizeof(ga)); if (ggap) *ggap = -ga[0] / 2; }
        else { memcpy(g, gb, sizeof(gb)); if (ggap) *ggap = gb[0] / 2; }
    }
    return tmin(a, b);
}

static real local_sdf(const orc_prim *P, int sh, real gap, const real *p, real *g, real *ggap) {
    if (ggap) *ggap = 0;
    switch (sh) {
    case SH_CAPSULE: return capsule_sdf(P->h, P->r, p, g);
    case SH_BOX: { real s[3] = {P->size[0], P->size[1], P->size[2]}; return box_sdf(s, p, g); }
    case SH_PRISM: {
        real ph[2] = {P->ph[0], P->ph[1]};
        real prot[4] = {P->prot[0], P->prot[1], P->prot[2], P->prot[3]};
        return prism_sdf(ph, prot, p, g);
    }
    case SH_KNIFE: return knife_sdf(P, p, g);
    case SH_CHOP: return chop_sdf(P, gap, p, g, ggap);
    case SH_CYL: return cylinder_sdf(P->h, P->r, p, g);
    case SH_TORUS: return torus_sdf(P->h, P->r, p, g);
    case SH_SPHERE: { /* Sphere.sdf (primitives.py:28-30), world-frame offset */
        real L = sqrt(dot3(p, p) + LEN_EPS_SHAPE);
        if (g) for (int i = 0; i < 3; i++) g[i] = p[i] / L;
        return L - P->r;
    }
    }
    return 0;
}

/* normals: analytic (Sphere, Capsule, Chopsticks, Cylinder, Torus) or central FD with a
 * float32 step 1e-4 (Box/Prism/Knife, primitives.py:382-393,720-731,796-807) */
static void local_normal(const orc_prim *P, int sh, real gap, const real *p, real *n) {
    if (sh == SH_CAPSULE) { capsule_normal(P->h, p, n); return; }
    if (sh == SH_CYL) { cylinder_normal(P->h, P->r, p, n); return; }
    if (sh == SH_TORUS) { torus_normal(P->h, p, n); return; }
    if (sh == SH_SPHERE) { /* Sphere.normal = normalize(grid_pos - position) */
        real L = sqrt(dot3(p, p) + LEN_EPS_SHAPE);
        for (int i = 0; i < 3; i++) n[i] = p[i] / L;
        return;
    }
    if (sh == SH_CHOP) { /* primitives.py:256-265 */
        real h = P->h, r = P->r;
        real pp[3] = {p[0], p[1] + h / 2, p[2]};
        real pa[3] = {pp[0] - gap / 2, pp[1], pp[2]}, pb[3] = {pp[0] + gap / 2, pp[1], pp[2]};
        real a = capsule_sdf(h, r, pa, NULL), b = capsule_sdf(h, r, pb, NULL);
        if (a <= b) capsule_normal(h, pa, n); else capsule_normal(h, pb, n);
        return;
    }
    const real d = (real)1e-4f;
    real m[3];
    for (int i = 0; i < 3; i++) {
        real inc[3] = {p[0], p[1], p[2]}, dec[3] = {p[0], p[1], p[2]};
        inc[i] += d; dec[i] -= d;
        m[i] = ((real)0.5 / d) * (local_sdf(P, sh, gap, inc, NULL, NULL) - local_sdf(P, sh, gap, dec, NULL, NULL));
    }
    real L = sqrt(dot3(m, m) + LEN_EPS_SHAPE);
    for (int i = 0; i < 3; i++) n[i] = m[i] / L;
}
static void local_normal_rev(const orc_prim *P, int sh, real gap, const real *p, const real *gn, real *gp, real *ggap) {
    if (sh == SH_CAPSULE) { capsule_normal_rev(P->h, p, gn, gp); return; }
    if (sh == SH_CYL) { cylinder_normal_rev(P->h, P->r, p, gn, gp); return; }
    if (sh == SH_TORUS) { torus_normal_rev(P->h, p, gn, gp); return; }
    if (sh == SH_SPHERE) {
        real L = sqrt(dot3(p, p) + LEN_EPS_SHAPE), d = dot3(p, gn);
        for (int i = 0; i < 3; i++) gp[i] += gn[i] / L - p[i] * d / (L * L * L);
        return;
    }
    if (sh == SH_CHOP) {
        real h = P->h, r = P->r;
        real pp[3] = {p[0], p[1] + h / 2, p[2]};
        real pa[3] = {pp[0] - gap / 2, pp[1], pp[2]}, pb[3] = {pp[0] + gap / 2, pp[1], pp[2]};
        real a = capsule_sdf(h, r, pa, NULL), b = capsule_sdf(h, r, pb, NULL);
        real g3[3] = {0, 0, 0};
        if (a <= b) { capsule_normal_rev(h, pa, gn, g3); if (ggap) *ggap += -g3[0] / 2; }
        else { capsule_normal_rev(h, pb, gn, g3); if (ggap) *ggap += g3[0] / 2; }
        for (int i = 0; i < 3; i++) gp[i] += g3[i];
        return;
    }
    const real d = (real)1e-4f;
    real m[3], gi[3][3], gd[3][3];
    for (int i = 0; i < 3; i++) {
        real inc[3] = {p[0], p[1], p[2]}, dec[3] = {p[0], p[1], p[2]};
        inc[i] += d; dec[i] -= d;
        real a = local_sdf(P, sh, gap, inc, gi[i], NULL);
        real b = local_sdf(P, sh, gap, dec, gd[i], NULL);
        m[i] = ((real)0.5 / d) * (a - b);
    }
    real L = sqrt(dot3(m, m) + LEN_EPS_SHAPE), dd = dot3(m, gn);
    for (int i = 0; i < 3; i++) {
        real gm = gn[i] / L - m[i] * dd / (L * L * L);
        for (int k = 0; k < 3; k++) gp[k] += gm * ((real)0.5 / d) * (gi[i][k] - gd[i][k]);
    }
}

/* ------------------------------------------------------------------ */
/* world <-> local frame of a collision unit                            */
/* jaw = 0: the primitive itself; jaw = -1/+1: Gripper/Gripper2 jaw at  */
/* get_pos(f, flag) = pos + qrot(rot, (gap/2 * flag, 0, 0))            */
/* (primitives.py:471-473).  Sphere uses no rotation.                  */
/* ------------------------------------------------------------------ */
static void unit_center(const real *ps, int jaw, real *c) {
    c[0] = ps[0]; c[1] = ps[1]; c[2] = ps[2];
    if (jaw) {
        real o[3] = {ps[7] / 2 * jaw, 0, 0}, t[3];
        qrot(ps + 3, o, t);
        for (int i = 0; i < 3; i++) c[i] += t[i];
    }
}
static void unit_center_rev(const real *ps, int jaw, const real *gc, real *gps) {
    for (int i = 0; i < 3; i++) gps[i] += gc[i];
    if (jaw) {
        real o[3] = {ps[7] / 2 * jaw, 0, 0}, go[3] = {0, 0, 0};
        qrot_rev(ps + 3, o, gc, gps + 3, go);
        gps[7] += go[0] * jaw / 2;
    }
}
/* inv_trans (utils.py:50-54) */
static void to_local(const real *ps, int jaw, int norot, const real *x, real *p) {
    real c[3], d[3];
    unit_center(ps, jaw, c);
    for (int i = 0; i < 3; i++) d[i] = x[i] - c[i];
    if (norot) { memcpy(p, d, sizeof(d)); return; }
    real iq[4];
    conj_normalized(ps + 3, iq);
    qrot(iq, d, p);
}
static void to_local_rev(const real *ps, int jaw, int norot, const real *x, const real *gp, real *gps) {
    real c[3], d[3], gd[3] = {0, 0, 0};
    unit_center(ps, jaw, c);
    for (int i = 0; i < 3; i++) d[i] = x[i] - c[i];
    if (norot) memcpy(gd, gp, sizeof(gd));
    else {
        real iq[4], giq[4] = {0, 0, 0, 0};
        conj_normalized(ps + 3, iq);
        qrot_rev(iq, d, gp, giq, gd);
        conj_normalized_rev(ps + 3, giq, gps + 3);
    }
    real gc[3] = {-gd[0], -gd[1], -gd[2]};
    unit_center_rev(ps, jaw, gc, gps);
}

typedef struct { const orc_prim *P; int jaw; int sh; int norot; real len_eps; } unit_t;

static int prim_units(const orc_prim *P, unit_t *u) {
    int sh = unit_shape(P->type);
    if (P->type == ORC_GRIPPER || P->type == ORC_GRIPPER2) {
        /* Gripper.collide: collide2(-1) then collide2(+1) (primitives.py:503-506) */
        u[0].P = P; u[0].jaw = -1; u[0].sh = sh; u[0].norot = 0; u[0].len_eps = LEN_EPS_SHAPE;
        u[1] = u[0]; u[1].jaw = 1;
        return 2;
    }
    u[0].P = P; u[0].jaw = 0; u[0].sh = sh; u[0].norot = (sh == SH_SPHERE); u[0].len_eps = LEN_EPS_BASE;
    return 1;
}

static real unit_sdf(const unit_t *u, const real *ps, const real *x) {
    real p[3];
    to_local(ps, u->jaw, u->norot, x, p);
    return local_sdf(u->P, u->sh, ps[7], p, NULL, NULL);
}
static void unit_sdf_rev(const unit_t *u, const real *ps, const real *x, real gd, real *gps) {
    real p[3], g[3], ggap;
    to_local(ps, u->jaw, u->norot, x, p);
    local_sdf(u->P, u->sh, ps[7], p, g, &ggap);
    for (int i = 0; i < 3; i++) g[i] *= gd;
    if (u->sh == SH_CHOP) gps[7] += ggap * gd;
    to_local_rev(ps, u->jaw, u->norot, x, g, gps);
}
/* world normal: qrot(rot, _normal(local)) (primive_base.py:81-85) */
static void unit_normal(const unit_t *u, const real *ps, const real *x, real *D) {
    real p[3], n[3];
    to_local(ps, u->jaw, u->norot, x, p);
    local_normal(u->P, u->sh, ps[7], p, n);
    if (u->norot) memcpy(D, n, sizeof(n));
    else qrot(ps + 3, n, D);
}
static void unit_normal_rev(const unit_t *u, const real *ps, const real *x, const real *gD, real *gps) {
    real p[3], n[3], gn[3] = {0, 0, 0}, gp[3] = {0, 0, 0}, ggap = 0;
    to_local(ps, u->jaw, u->norot, x, p);
    local_normal(u->P, u->sh, ps[7], p, n);
    if (u->norot) memcpy(gn, gD, sizeof(gn));
    else qrot_rev(ps + 3, n, gD, gps + 3, gn);
    local_normal_rev(u->P, u->sh, ps[7], p, gn, gp, &ggap);
    if (u->sh == SH_CHOP) gps[7] += ggap;
    to_local_rev(ps, u->jaw, u->norot, x, gp, gps);
}
/* collider_v (primive_base.py:87-94, primitives.py:497-502) */
static void unit_collider_v(const unit_t *u, const real *ps0, const real *ps1, const real *x, real dt, real *vc) {
    real rel[3], nw[3], c1[3];
    real iq[4], d[3], c0[3];
    unit_center(ps0, u->jaw, c0);
    for (int i = 0; i < 3; i++) d[i] = x[i] - c0[i];
    conj_normalized(ps0 + 3, iq);
    qrot(iq, d, rel);
    qrot(ps1 + 3, rel, nw);
    unit_center(ps1, u->jaw, c1);
    for (int i = 0; i < 3; i++) vc[i] = (nw[i] + c1[i] - x[i]) / dt;
}
static void unit_collider_v_rev(const unit_t *u, const real *ps0, const real *ps1, const real *x, real dt,
                                const real *gvc, real *gps0, real *gps1) {
    real rel[3], iq[4], d[3], c0[3];
    unit_center(ps0, u->jaw, c0);
    for (int i = 0; i < 3; i++) d[i] = x[i] - c0[i];
    conj_normalized(ps0 + 3, iq);
    qrot(iq, d, rel);
    real gn[3];
    for (int i = 0; i < 3; i++) gn[i] = gvc[i] / dt;
    unit_center_rev(ps1, u->jaw, gn, gps1);
    real grel[3] = {0, 0, 0};
    qrot_rev(ps1 + 3, rel, gn, gps1 + 3, grel);
    real giq[4] = {0, 0, 0, 0}, gd[3] = {0, 0, 0};
    qrot_rev(iq, d, grel, giq, gd);
    conj_normalized_rev(ps0 + 3, giq, gps0 + 3);
    real gc[3] = {-gd[0], -gd[1], -gd[2]};
    unit_center_rev(ps0, u->jaw, gc, gps0);
}

/* Primitive.collide (primive_base.py:96-120) / Gripper.collide2 (primitives.py:508-536) */
typedef struct {
    int branch, flag;
    real dist, ex, infl, D[3], vc[3], u[3], nc, mn, vt[3], vtn, mx, vt2[3];
} collide_tr;

static void unit_collide(const unit_t *u, const real *ps0, const real *ps1, const real *x, real dt,
                         const real *v, real *vout, collide_tr *tr) {
    const orc_prim *P = u->P;
    real soft = (real)P->softness, fr = (real)P->friction;
    collide_tr t;
    memset(&t, 0, sizeof(t));
    t.dist = unit_sdf(u, ps0, x);
    t.ex = exp(-t.dist * soft);
    t.infl = tmin(t.ex, 1);
    t.branch = (soft > 0 && t.infl > (real)0.1) || t.dist <= 0;
    if (t.branch) {
        unit_normal(u, ps0, x, t.D);
        unit_collider_v(u, ps0, ps1, x, dt, t.vc);
        for (int i = 0; i < 3; i++) t.u[i] = v[i] - t.vc[i];
        t.nc = dot3(t.u, t.D);
        t.mn = tmin(t.nc, 0);
        for (int i = 0; i < 3; i++) t.vt[i] = t.u[i] - t.mn * t.D[i];
        t.vtn = sqrt(dot3(t.vt, t.vt) + u->len_eps);
        t.mx = tmax(0, t.vtn + t.nc * fr);
        t.flag = (t.nc < 0) && (sqrt(dot3(t.vt, t.vt)) > (real)1e-30);
        for (int i = 0; i < 3; i++) {
            real vtf = t.vt[i] / t.vtn * t.mx;
            real fl = t.flag ? 1 : 0;
            t.vt2[i] = vtf * fl + t.vt[i] * (1 - fl);
            vout[i] = t.vc[i] + t.u[i] * (1 - t.infl) + t.vt2[i] * t.infl;
        }
    } else {
        for (int i = 0; i < 3; i++) vout[i] = v[i];
    }
    if (tr) *tr = t;
}
static void unit_collide_rev(const unit_t *u, const real *ps0, const real *ps1, const real *x, real dt,
                             const real *v, const real *gvout, real *gv, real *gps0, real *gps1) {
    collide_tr t;
    real vout[3];
    unit_collide(u, ps0, ps1, x, dt, v, vout, &t);
    if (!t.branch) { for (int i = 0; i < 3; i++) gv[i] += gvout[i]; return; }
    const orc_prim *P = u->P;
    real soft = (real)P->softness, fr = (real)P->friction;
    real gvc[3], gu[3], gvt2[3], ginfl = 0, gvt[3] = {0, 0, 0}, gD[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++) {
        gvc[i] = gvout[i];
        gu[i] = gvout[i] * (1 - t.infl);
        ginfl += -gvout[i] * t.u[i] + gvout[i] * t.vt2[i];
        gvt2[i] = gvout[i] * t.infl;
    }
    real gvtn = 0, gnc = 0;
    if (t.flag) {
        /* vtf = (vt / vtn) * mx */
        real gmx = 0, gq[3];
        for (int i = 0; i < 3; i++) { gq[i] = gvt2[i] * t.mx; gmx += gvt2[i] * t.vt[i] / t.vtn; }
        for (int i = 0; i < 3; i++) { gvt[i] += gq[i] / t.vtn; gvtn -= gq[i] * t.vt[i] / (t.vtn * t.vtn); }
        if (!((t.vtn + t.nc * fr) < 0)) { gvtn += gmx; gnc += gmx * fr; }
    } else {
        for (int i = 0; i < 3; i++) gvt[i] += gvt2[i];
    }
    for (int i = 0; i < 3; i++) gvt[i] += gvtn * t.vt[i] / t.vtn;
    real gmn = 0;
    for (int i = 0; i < 3; i++) { gu[i] += gvt[i]; gmn -= gvt[i] * t.D[i]; gD[i] -= t.mn * gvt[i]; }
    if (t.nc < 0) gnc += gmn;
    for (int i = 0; i < 3; i++) { gu[i] += gnc * t.D[i]; gD[i] += gnc * t.u[i]; }
    for (int i = 0; i < 3; i++) { gv[i] += gu[i]; gvc[i] -= gu[i]; }
    real gdist = 0;
    if (t.ex < 1) gdist = ginfl * t.ex * (-soft);
    unit_normal_rev(u, ps0, x, gD, gps0);
    unit_collider_v_rev(u, ps0, ps1, x, dt, gvc, gps0, gps1);
    if (gdist != 0) unit_sdf_rev(u, ps0, x, gdist, gps0);
}

/* ------------------------------------------------------------------ */
/* primitive kinematics                                                */
/* ------------------------------------------------------------------ */
static real clampv(real a, real lo, real hi) { return tmax(tmin(a, hi), lo); }
static real clampv_grad(real a, real lo, real hi) { real m = tmin(a, hi); return (a < hi && lo < m) ? 1 : 0; }

/* forward_kinematics: primive_base.py:152-156, RollingPin primitives.py:103-117,
 * RollingPinExt :121-136, Chopsticks :230-235, Gripper :456-460, Gripper2 :588-592 */
static void prim_fk(const orc_prim *P, const real *s, const real *pv, real *o) {
    real lo[3] = {P->lo[0], P->lo[1], P->lo[2]}, hi[3] = {P->hi[0], P->hi[1], P->hi[2]};
    o[7] = s[7];
    if (P->type == ORC_ROLLINGPIN || P->type == ORC_ROLLINGPIN_EXT) {
        real dw = pv[0], dth = pv[1], dy = pv[2];
        real ny[3] = {0, -1, 0}, ydir[3];
        qrot(s + 3, ny, ydir);
        real k = (P->type == ORC_ROLLINGPIN) ? dw * (real)0.03 : dw * (real)0.03 + pv[3];
        real xdir[3] = {ydir[2] * k, dy, -ydir[0] * k};
        real a1[3] = {0, -dth, 0}, a2[3] = {0, dw, 0}, q1[4], q2[4], t[4];
        w2quat(a1, q1); w2quat(a2, q2);
        qmul(s + 3, q2, t);
        qmul(q1, t, o + 3);
        for (int i = 0; i < 3; i++) o[i] = clampv(s[i] + xdir[i], lo[i], hi[i]);
        return;
    }
    for (int i = 0; i < 3; i++) o[i] = clampv(s[i] + pv[i], lo[i], hi[i]);
    real wq[4];
    w2quat(pv + 3, wq);
    if (P->type == ORC_GRIPPER || P->type == ORC_GRIPPER2 || P->type == ORC_CHOPSTICKS) {
        qmul(s + 3, wq, o + 3);
        real g = tmax(s[7] - pv[6], P->min_gap);
        o[7] = (P->type == ORC_CHOPSTICKS) ? g : tmin(g, P->max_gap);
    } else {
        qmul(wq, s + 3, o + 3);
    }
}
/* reverse of prim_fk: go = adjoint of frame f+1 state; adds to gs (frame f), gpv */
static void prim_fk_rev(const orc_prim *P, const real *s, const real *pv, const real *go, real *gs, real *gpv) {
    real lo[3] = {P->lo[0], P->lo[1], P->lo[2]}, hi[3] = {P->hi[0], P->hi[1], P->hi[2]};
    if (P->type == ORC_ROLLINGPIN || P->type == ORC_ROLLINGPIN_EXT) {
        real dw = pv[0], dth = pv[1], dy = pv[2];
        real ny[3] = {0, -1, 0}, ydir[3];
        qrot(s + 3, ny, ydir);
        real k = (P->type == ORC_ROLLINGPIN) ? dw * (real)0.03 : dw * (real)0.03 + pv[3];
        real xdir[3] = {ydir[2] * k, dy, -ydir[0] * k};
        real gx[3];
        for (int i = 0; i < 3; i++) {
            real g = go[i] * clampv_grad(s[i] + xdir[i], lo[i], hi[i]);
            gs[i] += g; gx[i] = g;
        }
        real gk = gx[0] * ydir[2] - gx[2] * ydir[0];
        real gydir[3] = {-gx[2] * k, 0, gx[0] * k};
        gpv[2] += gx[1];
        qrot_rev(s + 3, ny, gydir, gs + 3, NULL);
        gpv[0] += gk * (real)0.03;
        if (P->type == ORC_ROLLINGPIN_EXT) gpv[3] += gk;
        real a1[3] = {0, -dth, 0}, a2[3] = {0, dw, 0}, q1[4], q2[4], t[4];
        w2quat(a1, q1); w2quat(a2, q2);
        qmul(s + 3, q2, t);
        real gq1[4] = {0, 0, 0, 0}, gt[4] = {0, 0, 0, 0}, gq2[4] = {0, 0, 0, 0};
        qmul_rev(q1, t, go + 3, gq1, gt);
        qmul_rev(s + 3, q2, gt, gs + 3, gq2);
        real ga1[3] = {0, 0, 0}, ga2[3] = {0, 0, 0};
        w2quat_rev(a1, gq1, ga1);
        w2quat_rev(a2, gq2, ga2);
        gpv[1] -= ga1[1];
        gpv[0] += ga2[1];
        gs[7] += go[7];
        return;
    }
    for (int i = 0; i < 3; i++) {
        real g = go[i] * clampv_grad(s[i] + pv[i], lo[i], hi[i]);
        gs[i] += g; gpv[i] += g;
    }
    real wq[4], gwq[4] = {0, 0, 0, 0};
    w2quat(pv + 3, wq);
    if (P->type == ORC_GRIPPER || P->type == ORC_GRIPPER2 || P->type == ORC_CHOPSTICKS) {
        qmul_rev(s + 3, wq, go + 3, gs + 3, gwq);
        real a = s[7] - pv[6];
        real g1 = tmax(a, P->min_gap);
        real gg = go[7];
        if (P->type != ORC_CHOPSTICKS) gg = (g1 < P->max_gap) ? gg : 0;
        gg = ((real)P->min_gap < a) ? gg : 0;
        gs[7] += gg; gpv[6] -= gg;
    } else {
        qmul_rev(wq, s + 3, go + 3, gwq, gs + 3);
        gs[7] += go[7];
    }
    w2quat_rev(pv + 3, gwq, gpv + 3);
}

void orc_fk(const orc_prim *prims, int n_prims, const real *ps_f, const real *pv_f, real *ps_f1) {
    for (int i = 0; i < n_prims; i++) prim_fk(&prims[i], ps_f + i * NPS, pv_f + i * NPV, ps_f1 + i * NPS);
}

/* set_velocity (primive_base.py:260-268, primitives.py:238-245,462-469,594-601):
 * per-substep displacement v = a * scale / n_substeps */
void orc_set_velocity(const orc_prim *P, const real *a, int n_substeps, real *pv) {
    for (int k = 0; k < NPV; k++) pv[k] = 0;
    int ad = P->action_dim;
    for (int k = 0; k < 3 && k < ad; k++) pv[k] = a[k] * (real)P->action_scale[k] / (real)n_substeps;
    if (ad > 3) for (int k = 0; k < 3; k++) pv[3 + k] = a[3 + k] * (real)P->action_scale[3 + k] / (real)n_substeps;
    if (ad > 6) pv[6] = a[6] * (real)P->action_scale[6] / (real)n_substeps;
}

/* ------------------------------------------------------------------ */
/* 3x3 SVD (one-sided Jacobi, run to convergence), convention of       */
/* ti.svd: U, V rotations, sigma sorted descending, sign of det(F) on  */
/* the last singular value.                                            */
/* ------------------------------------------------------------------ */
static void svd3(const real *A, real *U, real *S, real *V) {
    double B[9], W[9];
    for (int i = 0; i < 9; i++) { B[i] = A[i]; W[i] = (i % 4 == 0) ? 1.0 : 0.0; }
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0;
        for (int pq = 0; pq < 3; pq++) {
            int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
            double al = 0, be = 0, ga = 0;
            for (int r = 0; r < 3; r++) {
                al += B[r * 3 + p] * B[r * 3 + p];
                be += B[r * 3 + q] * B[r * 3 + q];
                ga += B[r * 3 + p] * B[r * 3 + q];
            }
            if (fabs(ga) <= 1e-300 || fabs(ga) <= 1e-17 * sqrt(al * be)) continue;
            off += fabs(ga) / sqrt(al * be);
            double zeta = (be - al) / (2 * ga);
            double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
            double c = 1 / sqrt(1 + t * t), s = c * t;
            for (int r = 0; r < 3; r++) {
                double bp = B[r * 3 + p], bq = B[r * 3 + q];
                B[r * 3 + p] = c * bp - s * bq;
                B[r * 3 + q] = s * bp + c * bq;
                double wp = W[r * 3 + p], wq = W[r * 3 + q];
                W[r * 3 + p] = c * wp - s * wq;
                W[r * 3 + q] = s * wp + c * wq;
            }
        }
        if (off < 1e-15) break;
    }
    double sig[3];
    for (int j = 0; j < 3; j++) sig[j] = sqrt(B[j] * B[j] + B[3 + j] * B[3 + j] + B[6 + j] * B[6 + j]);
    int ord[3] = {0, 1, 2};
    for (int a = 0; a < 3; a++)
        for (int b = a + 1; b < 3; b++)
            if (sig[ord[b]] > sig[ord[a]]) { int t = ord[a]; ord[a] = ord[b]; ord[b] = t; }
    double Uo[9], Vo[9], so[3];
    for (int j = 0; j < 3; j++) {
        int k = ord[j];
        so[j] = sig[k];
        for (int r = 0; r < 3; r++) {
            Vo[r * 3 + j] = W[r * 3 + k];
            Uo[r * 3 + j] = so[j] > 0 ? B[r * 3 + k] / so[j] : 0;
        }
    }
    /* complete U for zero singular values */
    if (!(so[1] > 0)) {
        double e[3] = {0, 0, 0};
        int m = fabs(Uo[0]) < 0.6 ? 0 : (fabs(Uo[3]) < 0.6 ? 1 : 2);
        e[m] = 1;
        if (!(so[0] > 0)) { Uo[0] = 1; Uo[3] = 0; Uo[6] = 0; e[0] = 0; e[1] = 1; e[2] = 0; }
        double d = e[0] * Uo[0] + e[1] * Uo[3] + e[2] * Uo[6];
        double u1[3] = {e[0] - d * Uo[0], e[1] - d * Uo[3], e[2] - d * Uo[6]};
        double nn = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
        for (int r = 0; r < 3; r++) Uo[r * 3 + 1] = u1[r] / nn;
    }
    if (!(so[2] > 0)) {
        Uo[2] = Uo[3] * Uo[7] - Uo[6] * Uo[4];
        Uo[5] = Uo[6] * Uo[1] - Uo[0] * Uo[7];
        Uo[8] = Uo[0] * Uo[4] - Uo[3] * Uo[1];
    }
    double dV = Vo[0] * (Vo[4] * Vo[8] - Vo[5] * Vo[7]) - Vo[1] * (Vo[3] * Vo[8] - Vo[5] * Vo[6]) + Vo[2] * (Vo[3] * Vo[7] - Vo[4] * Vo[6]);
    if (dV < 0) for (int r = 0; r < 3; r++) { Vo[r * 3 + 2] = -Vo[r * 3 + 2]; Uo[r * 3 + 2] = -Uo[r * 3 + 2]; }
    double dU = Uo[0] * (Uo[4] * Uo[8] - Uo[5] * Uo[7]) - Uo[1] * (Uo[3] * Uo[8] - Uo[5] * Uo[6]) + Uo[2] * (Uo[3] * Uo[7] - Uo[4] * Uo[6]);
    if (dU < 0) { for (int r = 0; r < 3; r++) Uo[r * 3 + 2] = -Uo[r * 3 + 2]; so[2] = -so[2]; }
    for (int i = 0; i < 9; i++) { U[i] = (real)Uo[i]; V[i] = (real)Vo[i]; S[i] = 0; }
    S[0] = (real)so[0]; S[4] = (real)so[1]; S[8] = (real)so[2];
}
void orc_svd3(const real *A, real *U, real *S, real *V) { svd3(A, U, S, V); }

/* clamp (mpm_simulator.py:184-192) */
static real gap_clamp(real a) { return a >= 0 ? tmax(a, (real)1e-6) : tmin(a, (real)-1e-6); }

/* backward_svd (mpm_simulator.py:136-156) */
static void backward_svd(const real *gu, const real *gs, const real *gv, const real *u, const real *sig, const real *v, real *out) {
    real s[3] = {sig[0] * sig[0], sig[4] * sig[4], sig[8] * sig[8]};
    real Fm[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Fm[i * 3 + j] = (i == j) ? 0 : (real)1 / gap_clamp(s[j] - s[i]);
    real t1[9], t2[9], M[9], tmp[9];
    real sigma_term[9];
    matmul3(u, gs, tmp);
    matmul3_nt(tmp, v, sigma_term);
    /* u_term = u @ ((F * (ut @ gu - gu^T @ u)) @ sig) @ vt */
    matmul3_tn(u, gu, t1);
    matmul3_tn(gu, u, t2);
    for (int i = 0; i < 9; i++) M[i] = Fm[i] * (t1[i] - t2[i]);
    matmul3(M, sig, tmp);
    real ut[9];
    matmul3(u, tmp, ut);
    real u_term[9];
    matmul3_nt(ut, v, u_term);
    /* v_term = u @ (sig @ ((F * (vt @ gv - gv^T @ v)) @ vt)) */
    matmul3_tn(v, gv, t1);
    matmul3_tn(gv, v, t2);
    for (int i = 0; i < 9; i++) M[i] = Fm[i] * (t1[i] - t2[i]);
    matmul3_nt(M, v, tmp);
    real t3[9];
    matmul3(sig, tmp, t3);
    real v_term[9];
    matmul3(u, t3, v_term);
    for (int i = 0; i < 9; i++) out[i] = u_term[i] + v_term[i] + sigma_term[i];
}

/* ------------------------------------------------------------------ */
/* particle-side math                                                  */
/* ------------------------------------------------------------------ */
typedef struct {
    real dx, inv_dx, dt, p_vol, p_mass, mu, lam, ys, stress_k, g30[3], lower, upper, c4;
    int n;
} consts_t;

static void make_consts(const orc_sim *S, consts_t *k) {
    k->n = S->n_grid;
    k->dx = (real)S->dx; k->inv_dx = (real)S->inv_dx; k->dt = (real)S->dt;
    k->p_vol = (real)S->p_vol; k->p_mass = (real)S->p_mass;
    k->mu = (real)S->mu; k->lam = (real)S->lam; k->ys = (real)S->yield_stress;
    /* (-dt * p_vol * 4 * inv_dx * inv_dx) is folded in double by the reference's Python (mpm_simulator.py:214) */
    k->stress_k = (real)(-S->dt * S->p_vol * 4 * S->inv_dx * S->inv_dx);
    for (int i = 0; i < 3; i++) k->g30[i] = k->dt * (real)S->gravity[i] * (real)30;
    k->upper = (real)(1. - 3 * S->dx);
    k->lower = (real)(S->lower_bound * S->dx);
    k->c4 = (real)(4 * S->inv_dx);
}

/* quadratic B-spline stencil (mpm_simulator.py:201-204): base uses cast(int) = truncation */
static void stencil(const consts_t *k, const real *x, int *base, real w[3][3], real dw[3][3], real *fx) {
    for (int d = 0; d < 3; d++) {
        real xs = x[d] * k->inv_dx;
        base[d] = (int)(xs - (real)0.5);
        fx[d] = xs - (real)base[d];
        w[0][d] = (real)0.5 * ((real)1.5 - fx[d]) * ((real)1.5 - fx[d]);
        w[1][d] = (real)0.75 - (fx[d] - 1) * (fx[d] - 1);
        w[2][d] = (real)0.5 * (fx[d] - (real)0.5) * (fx[d] - (real)0.5);
        dw[0][d] = fx[d] - (real)1.5;
        dw[1][d] = -2 * (fx[d] - 1);
        dw[2][d] = fx[d] - (real)0.5;
    }
}
static inline int in_grid(int n, const int *I) {
    return I[0] >= 0 && I[0] < n && I[1] >= 0 && I[1] < n && I[2] >= 0 && I[2] < n;
}

typedef struct {
    real Ftmp[9], U[9], S[9], V[9], Fn[9];
    real Sc[3], eps[3], eh[3], ehn, dg, e[3];
    int yield;
    real J, R[9], stress[9], affine[9];
} ptrace_t;

/* compute_F_tmp + svd + compute_von_mises + stress (mpm_simulator.py:121-129,165-182,198-216) */
static void particle_stress(const consts_t *k, const real *st, ptrace_t *t) {
    const real *C = st + 6, *F = st + 15;
    real A[9];
    for (int i = 0; i < 9; i++) A[i] = ((i % 4 == 0) ? 1 : 0) + k->dt * C[i];
    matmul3(A, F, t->Ftmp);
    svd3(t->Ftmp, t->U, t->S, t->V);
    for (int i = 0; i < 3; i++) {
        t->Sc[i] = tmax(t->S[i * 4], (real)0.05);
        t->eps[i] = log(t->Sc[i]);
    }
    real mean = (t->eps[0] + t->eps[1] + t->eps[2]) / 3;
    for (int i = 0; i < 3; i++) t->eh[i] = t->eps[i] - mean;
    t->ehn = sqrt(dot3(t->eh, t->eh) + (real)1e-8);
    t->dg = t->ehn - k->ys / (2 * k->mu);
    t->yield = t->dg > 0;
    if (t->yield) {
        real D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, tmp[9];
        for (int i = 0; i < 3; i++) {
            t->e[i] = exp(t->eps[i] - (t->dg / t->ehn) * t->eh[i]);
            D[i * 4] = t->e[i];
        }
        matmul3(t->U, D, tmp);
        matmul3_nt(tmp, t->V, t->Fn);
    } else {
        memcpy(t->Fn, t->Ftmp, sizeof(t->Fn));
    }
    t->J = det3(t->Fn);
    matmul3_nt(t->U, t->V, t->R);
    real FmR[9], P1[9];
    for (int i = 0; i < 9; i++) FmR[i] = t->Fn[i] - t->R[i];
    matmul3_nt(FmR, t->Fn, P1);
    for (int i = 0; i < 9; i++) {
        real s = 2 * k->mu * P1[i] + ((i % 4 == 0) ? k->lam * t->J * (t->J - 1) : 0);
        t->stress[i] = k->stress_k * s;
        t->affine[i] = t->stress[i] + k->p_mass * C[i];
    }
}

/* ------------------------------------------------------------------ */
/* forward phases                                                      */
/* ------------------------------------------------------------------ */
static void p2g(const consts_t *k, int np, const real *st, real *st_out, real *grid) {
    int n = k->n;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int p = 0; p < np; p++) {
        const real *s = st + (size_t)p * NST;
        ptrace_t t;
        particle_stress(k, s, &t);
        if (st_out) memcpy(st_out + (size_t)p * NST + 15, t.Fn, sizeof(t.Fn));
        int base[3];
        real w[3][3], dw[3][3], fx[3];
        stencil(k, s, base, w, dw, fx);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 3; c++) {
                    int I[3] = {base[0] + a, base[1] + b, base[2] + c};
                    if (!in_grid(n, I)) continue;
                    real wt = w[a][0] * w[b][1] * w[c][2];
                    real dpos[3] = {((real)a - fx[0]) * k->dx, ((real)b - fx[1]) * k->dx, ((real)c - fx[2]) * k->dx};
                    real *g = grid + ((size_t)(I[0] * n + I[1]) * n + I[2]) * 4;
                    for (int i = 0; i < 3; i++) {
                        real val = wt * (k->p_mass * s[3 + i] + t.affine[i * 3] * dpos[0] + t.affine[i * 3 + 1] * dpos[1] + t.affine[i * 3 + 2] * dpos[2]);
#pragma omp atomic
                        g[i] += val;
                    }
                    real mv = wt * k->p_mass;
#pragma omp atomic
                    g[3] += mv;
                }
    }
}

typedef struct {
    real v0[3], v1[3];
    real vu[16][3];      /* input velocity of each collision unit */
    real vw[3][3];       /* velocity before wall step d */
    int nunits;
} node_tr;

/* grid_op (mpm_simulator.py:230-262) for one node with m > 1e-12 */
static void grid_node(const consts_t *k, const orc_sim *S, const orc_prim *prims, const real *ps0, const real *ps1,
                      const int *I, const real *pm, real *vout, node_tr *tr) {
    real inv_m = 1 / pm[3];
    real v[3];
    for (int i = 0; i < 3; i++) { v[i] = inv_m * pm[i]; if (tr) tr->v0[i] = v[i]; v[i] += k->g30[i]; }
    real x[3] = {(real)I[0] * k->dx, (real)I[1] * k->dx, (real)I[2] * k->dx};
    int nu = 0;
    for (int q = 0; q < S->n_prims; q++) {
        unit_t u[2];
        int m = prim_units(&prims[q], u);
        for (int j = 0; j < m; j++) {
            if (tr) memcpy(tr->vu[nu], v, sizeof(v));
            real vo[3];
            unit_collide(&u[j], ps0 + q * NPS, ps1 + q * NPS, x, k->dt, v, vo, NULL);
            memcpy(v, vo, sizeof(v));
            nu++;
        }
    }
    if (tr) tr->nunits = nu;
    real gf = (real)S->ground_friction;
    for (int d = 0; d < 3; d++) {
        if (tr) memcpy(tr->vw[d], v, sizeof(v));
        if (I[d] < 3 && v[d] < 0) {
            if (d != 1 || gf == 0) v[d] = 0;
            else if (gf < 10) {
                real lin = v[1] + (real)1e-30;
                real vit[3];
                for (int i = 0; i < 3; i++) vit[i] = v[i] - (i == 1 ? lin : 0) - (real)I[i] * (real)1e-30;
                real lit = sqrt(dot3(vit, vit) + (real)1e-8);
                real c = tmax(1 + gf * lin / lit, 0);
                for (int i = 0; i < 3; i++) v[i] = c * (vit[i] + (real)I[i] * (real)1e-30);
                v[1] = 0;
            } else {
                v[0] = v[1] = v[2] = 0;
            }
        }
        if (I[d] > k->n - 3 && v[d] > 0) v[d] = 0;
    }
    memcpy(vout, v, sizeof(v));
}

static void grid_node_rev(const consts_t *k, const orc_sim *S, const orc_prim *prims, const real *ps0, const real *ps1,
                          const int *I, const real *pm, const real *gvout, real *gpm, real *gps0, real *gps1) {
    node_tr tr;
    real vout[3];
    grid_node(k, S, prims, ps0, ps1, I, pm, vout, &tr);
    real g[3] = {gvout[0], gvout[1], gvout[2]};
    real gf = (real)S->ground_friction;
    for (int d = 2; d >= 0; d--) {
        /* replay step d forward from tr.vw[d] to know the branch taken */
        real v[3];
        memcpy(v, tr.vw[d], sizeof(v));
        int lower = I[d] < 3 && v[d] < 0;
        real vmid[3];
        memcpy(vmid, v, sizeof(v));
        int mode = 0; /* 0 none, 1 zero comp, 2 coulomb, 3 zero all */
        real lin = 0, vit[3] = {0, 0, 0}, lit = 0, c = 0, sarg = 0;
        if (lower) {
            if (d != 1 || gf == 0) { mode = 1; vmid[d] = 0; }
            else if (gf < 10) {
                mode = 2;
                lin = v[1] + (real)1e-30;
                for (int i = 0; i < 3; i++) vit[i] = v[i] - (i == 1 ? lin : 0) - (real)I[i] * (real)1e-30;
                lit = sqrt(dot3(vit, vit) + (real)1e-8);
                sarg = 1 + gf * lin / lit;
                c = tmax(sarg, 0);
                for (int i = 0; i < 3; i++) vmid[i] = c * (vit[i] + (real)I[i] * (real)1e-30);
                vmid[1] = 0;
            } else { mode = 3; vmid[0] = vmid[1] = vmid[2] = 0; }
        }
        if (I[d] > k->n - 3 && vmid[d] > 0) g[d] = 0;
        if (mode == 1) g[d] = 0;
        else if (mode == 3) { g[0] = g[1] = g[2] = 0; }
        else if (mode == 2) {
            g[1] = 0;
            real uu[3], gc = 0, gvit[3];
            for (int i = 0; i < 3; i++) { uu[i] = vit[i] + (real)I[i] * (real)1e-30; gc += g[i] * uu[i]; gvit[i] = c * g[i]; }
            real gs = (0 < sarg) ? gc : 0;
            real glin = gs * gf / lit;
            real glit = -gs * gf * lin / (lit * lit);
            for (int i = 0; i < 3; i++) gvit[i] += glit * vit[i] / lit;
            real gv[3] = {gvit[0], gvit[1], gvit[2]};
            glin -= gvit[1];
            gv[1] += glin;
            memcpy(g, gv, sizeof(gv));
        }
    }
    real x[3] = {(real)I[0] * k->dx, (real)I[1] * k->dx, (real)I[2] * k->dx};
    int uidx = tr.nunits;
    for (int q = S->n_prims - 1; q >= 0; q--) {
        unit_t u[2];
        int m = prim_units(&prims[q], u);
        for (int j = m - 1; j >= 0; j--) {
            uidx--;
            real gv[3] = {0, 0, 0};
            unit_collide_rev(&u[j], ps0 + q * NPS, ps1 + q * NPS, x, k->dt, tr.vu[uidx], g, gv, gps0 + q * NPS, gps1 + q * NPS);
            memcpy(g, gv, sizeof(gv));
        }
    }
    real inv_m = 1 / pm[3];
    real ginv = 0;
    for (int i = 0; i < 3; i++) { gpm[i] += inv_m * g[i]; ginv += g[i] * pm[i]; }
    gpm[3] += -ginv / (pm[3] * pm[3]);
}

static void grid_op(const consts_t *k, const orc_sim *S, const orc_prim *prims, const real *ps0, const real *ps1,
                    const real *grid, real *gout) {
    int n = k->n;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int l = 0; l < n; l++) {
                size_t id = ((size_t)(i * n + j)) * n + l;
                const real *pm = grid + id * 4;
                real *vo = gout + id * 3;
                vo[0] = vo[1] = vo[2] = 0;
                if (pm[3] > (real)1e-12) {
                    int I[3] = {i, j, l};
                    grid_node(k, S, prims, ps0, ps1, I, pm, vo, NULL);
                }
            }
}

/* g2p (mpm_simulator.py:264-283) */
static void g2p(const consts_t *k, int np, const real *st, const real *gout, real *st_out) {
    int n = k->n;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int p = 0; p < np; p++) {
        const real *s = st + (size_t)p * NST;
        real *o = st_out + (size_t)p * NST;
        int base[3];
        real w[3][3], dw[3][3], fx[3];
        stencil(k, s, base, w, dw, fx);
        real nv[3] = {0, 0, 0}, nC[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 3; c++) {
                    int I[3] = {base[0] + a, base[1] + b, base[2] + c};
                    if (!in_grid(n, I)) continue;
                    real wt = w[a][0] * w[b][1] * w[c][2];
                    real dpos[3] = {(real)a - fx[0], (real)b - fx[1], (real)c - fx[2]};
                    const real *g = gout + ((size_t)(I[0] * n + I[1]) * n + I[2]) * 3;
                    for (int i = 0; i < 3; i++) {
                        nv[i] += wt * g[i];
                        for (int j = 0; j < 3; j++) nC[i * 3 + j] += k->c4 * wt * g[i] * dpos[j];
                    }
                }
        for (int i = 0; i < 3; i++) {
            o[3 + i] = nv[i];
            o[i] = tmax(tmin(s[i] + k->dt * nv[i], k->upper), k->lower);
        }
        memcpy(o + 6, nC, sizeof(nC));
    }
}

/* ------------------------------------------------------------------ */
/* tool-tool collision projection (mpm_simulator.py:286-305,316-320)   */
/* ------------------------------------------------------------------ */
/* Primitive.sdf and Primitive.normal of a whole primitive at world point x:
 * a two-jaw gripper takes the smaller jaw SDF and the normal of jaw -1 when
 * a <= b (primitives.py:486-496,618-628).  *jsdf is the jaw that receives
 * the SDF's adjoint: Taichi's min sends it to the left operand only when it
 * is strictly smaller. */
static real prim_sdf_normal(const orc_prim *P, const real *ps, const real *x, real *n, int *jsdf, int *jn) {
    unit_t u[2];
    int m = prim_units(P, u);
    if (m == 1) {
        *jsdf = *jn = 0;
        if (n) unit_normal(&u[0], ps, x, n);
        return unit_sdf(&u[0], ps, x);
    }
    real a = unit_sdf(&u[0], ps, x), b = unit_sdf(&u[1], ps, x);
    *jsdf = a < b ? 0 : 1;
    *jn = a <= b ? 0 : 1;
    if (n) unit_normal(&u[*jn], ps, x, n);
    return tmin(a, b);
}

/* All pairs' surface points and collision indices are taken from the poses
 * after forward kinematics, then the projections are applied in pair order:
 *  - set_surface_points (mpm_simulator.py:286-292): point k of pair (i, j) is
 *    rand[k] * size_j clamped to the box (Box._project, primitives.py:369-372)
 *    and moved to the world by j's pose (Primitive.get_surface_pos,
 *    primive_base.py:284-287);
 *  - set_collision_idx (mpm_simulator.py:294-298, primive_base.py:135-143):
 *    the point of smallest negative SDF of i, the first one on ties (what the
 *    reference's loop computes run serially), -1 if none is inside;
 *  - apply_collision_projection (mpm_simulator.py:300-305,
 *    primive_base.py:145-150): position_i += normalized(normal_i) * sdf_i.
 * pre[p] records position_i before pair p's projection. */
static void tool_collide(const orc_prim *prims, const orc_coll *c, real *ps1) {
    for (int p = 0; p < c->n_pairs; p++) {
        int i = c->pairs[2 * p], j = c->pairs[2 * p + 1];
        const orc_prim *Pi = &prims[i], *Pj = &prims[j];
        const real *psi = ps1 + i * NPS, *psj = ps1 + j * NPS;
        real best = 0, bw[3] = {0, 0, 0};
        int bk = -1;
        for (int k = 0; k < c->n_points; k++) {
            const real *r = c->rand + ((size_t)p * c->n_points + k) * 3;
            real loc[3], w[3];
            for (int d = 0; d < 3; d++) loc[d] = clampv(r[d] * Pj->size[d], -Pj->size[d], Pj->size[d]);
            qrot(psj + 3, loc, w);
            for (int d = 0; d < 3; d++) w[d] += psj[d];
            int js, jn;
            real dist = prim_sdf_normal(Pi, psi, w, NULL, &js, &jn);
            if (dist < best) { best = dist; bk = k; memcpy(bw, w, sizeof(bw)); }
        }
        c->idx[p] = bk;
        memcpy(c->pt + 3 * p, bw, sizeof(bw));
    }
    for (int p = 0; p < c->n_pairs; p++) {
        int i = c->pairs[2 * p];
        real *psi = ps1 + i * NPS;
        memcpy(c->pre + 3 * p, psi, 3 * sizeof(real));
        if (c->idx[p] < 0) continue;
        real n[3];
        int js, jn;
        real dist = prim_sdf_normal(&prims[i], psi, c->pt + 3 * p, n, &js, &jn);
        real L = sqrt(dot3(n, n));
        for (int d = 0; d < 3; d++) psi[d] += n[d] / L * dist;
    }
}

/* Adjoint of the projections, pairs in reverse order.  The surface points
 * carry no gradient (rand_points has no needs_grad, mpm_simulator.py:65), so
 * only primitive i's state receives one: position_i' = position_i +
 * n^(x_i) d(x_i) gives g_pos += J_pos^T g_pos and the rotation / gap
 * adjoints J^T g_pos.  J is taken at the pose the projection started from
 * (pre[p], rotation and gap unchanged by projections); the reference's
 * in-place update makes its grad kernel read the projected pose instead,
 * where the point lies on the tool surface and a box SDF's gradient is
 * decided by rounding. */
static void tool_collide_rev(const orc_prim *prims, const orc_coll *c, const real *ps1, real *gps1) {
    for (int p = c->n_pairs - 1; p >= 0; p--) {
        if (c->idx[p] < 0) continue;
        int i = c->pairs[2 * p];
        const orc_prim *Pi = &prims[i];
        real ps[NPS];
        memcpy(ps, ps1 + i * NPS, sizeof(ps));
        memcpy(ps, c->pre + 3 * p, 3 * sizeof(real));
        const real *x = c->pt + 3 * p;
        real n[3];
        int js, jn;
        real dist = prim_sdf_normal(Pi, ps, x, n, &js, &jn);
        real L = sqrt(dot3(n, n)), nh[3], g[3], gnh[3], gn[3];
        real *gi = gps1 + i * NPS;
        for (int d = 0; d < 3; d++) { nh[d] = n[d] / L; g[d] = gi[d]; gnh[d] = dist * g[d]; }
        real gd = dot3(nh, g), t = dot3(nh, gnh);
        for (int d = 0; d < 3; d++) gn[d] = (gnh[d] - nh[d] * t) / L;
        unit_t u[2];
        prim_units(Pi, u);
        unit_sdf_rev(&u[js], ps, x, gd, gi);
        unit_normal_rev(&u[jn], ps, x, gn, gi);
    }
}

/* the collision pass alone on one frame's primitive states (in place) */
void orc_tool_collide(const orc_prim *prims, const orc_coll *c, real *ps1) { tool_collide(prims, c, ps1); }
void orc_tool_collide_rev(const orc_prim *prims, const orc_coll *c, const real *ps1, real *gps1) {
    tool_collide_rev(prims, c, ps1, gps1);
}

/* one substep s -> s+1 (mpm_simulator.py:307-323).  grid: n^3*4, gout: n^3*3
 * scratch; c (may be NULL) the tool-tool collision pairs and their records */
void orc_substep_c(const orc_sim *S, const orc_prim *prims, int np, const real *st, real *st_out,
                   const real *ps_f, const real *pv_f, real *ps_f1, real *grid, real *gout, const orc_coll *c) {
    consts_t k;
    make_consts(S, &k);
    size_t G = (size_t)k.n * k.n * k.n;
    memset(grid, 0, G * 4 * sizeof(real));
    p2g(&k, np, st, st_out, grid);
    orc_fk(prims, S->n_prims, ps_f, pv_f, ps_f1);
    if (c && c->n_pairs > 0) tool_collide(prims, c, ps_f1);
    grid_op(&k, S, prims, ps_f, ps_f1, grid, gout);
    g2p(&k, np, st, gout, st_out);
}
void orc_substep(const orc_sim *S, const orc_prim *prims, int np, const real *st, real *st_out,
                 const real *ps_f, const real *pv_f, real *ps_f1, real *grid, real *gout) {
    orc_substep_c(S, prims, np, st, st_out, ps_f, pv_f, ps_f1, grid, gout, NULL);
}

/* ------------------------------------------------------------------ */
/* adjoint phases                                                      */
/* ------------------------------------------------------------------ */
/* reverse of g2p: adds to adj_f (x), returns grid adjoint of grid_v_out */
static void g2p_rev(const consts_t *k, int np, const real *st, const real *gout, const real *adj1, real *adj0, real *ggout) {
    int n = k->n;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int p = 0; p < np; p++) {
        const real *s = st + (size_t)p * NST;
        const real *a1 = adj1 + (size_t)p * NST;
        real *a0 = adj0 + (size_t)p * NST;
        int base[3];
        real w[3][3], dw[3][3], fx[3];
        stencil(k, s, base, w, dw, fx);
        real nv[3] = {0, 0, 0};
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 3; c++) {
                    int I[3] = {base[0] + a, base[1] + b, base[2] + c};
                    if (!in_grid(n, I)) continue;
                    real wt = w[a][0] * w[b][1] * w[c][2];
                    const real *g = gout + ((size_t)(I[0] * n + I[1]) * n + I[2]) * 3;
                    for (int i = 0; i < 3; i++) nv[i] += wt * g[i];
                }
        real gnv[3], gx[3];
        for (int i = 0; i < 3; i++) {
            real xn = s[i] + k->dt * nv[i];
            real gxn = a1[i] * clampv_grad(xn, k->lower, k->upper);
            gx[i] = gxn;
            gnv[i] = a1[3 + i] + k->dt * gxn;
        }
        const real *gC = a1 + 6;
        real gfx[3] = {0, 0, 0};
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 3; c++) {
                    int I[3] = {base[0] + a, base[1] + b, base[2] + c};
                    if (!in_grid(n, I)) continue;
                    int off[3] = {a, b, c};
                    real wt = w[a][0] * w[b][1] * w[c][2];
                    real dpos[3] = {(real)a - fx[0], (real)b - fx[1], (real)c - fx[2]};
                    size_t id = ((size_t)(I[0] * n + I[1]) * n + I[2]);
                    const real *g = gout + id * 3;
                    real Cd[3], gw = 0, CTg[3];
                    for (int i = 0; i < 3; i++) {
                        Cd[i] = gC[i * 3] * dpos[0] + gC[i * 3 + 1] * dpos[1] + gC[i * 3 + 2] * dpos[2];
                        CTg[i] = gC[i] * g[0] + gC[3 + i] * g[1] + gC[6 + i] * g[2];
                        gw += gnv[i] * g[i] + k->c4 * g[i] * Cd[i];
                    }
                    real *gg = ggout + id * 3;
                    for (int i = 0; i < 3; i++) {
                        real val = wt * gnv[i] + k->c4 * wt * Cd[i];
#pragma omp atomic
                        gg[i] += val;
                    }
                    for (int d = 0; d < 3; d++) {
                        real dwd = dw[off[d]][d];
                        for (int e = 0; e < 3; e++) if (e != d) dwd *= w[off[e]][e];
                        gfx[d] += gw * dwd - k->c4 * wt * CTg[d];
                    }
                }
        for (int i = 0; i < 3; i++) a0[i] += gx[i] + k->inv_dx * gfx[i];
    }
}

static void grid_op_rev(const consts_t *k, const orc_sim *S, const orc_prim *prims, const real *ps0, const real *ps1,
                        const real *grid, const real *ggout, real *ggrid, real *gps0, real *gps1) {
    int n = k->n, npr = S->n_prims;
    int nthr = g_threads;
    real *acc0 = (real *)calloc((size_t)nthr * npr * NPS + 1, sizeof(real));
    real *acc1 = (real *)calloc((size_t)nthr * npr * NPS + 1, sizeof(real));
#pragma omp parallel num_threads(nthr)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        real *l0 = acc0 + (size_t)tid * npr * NPS, *l1 = acc1 + (size_t)tid * npr * NPS;
#pragma omp for schedule(static)
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
                for (int l = 0; l < n; l++) {
                    size_t id = ((size_t)(i * n + j)) * n + l;
                    const real *pm = grid + id * 4;
                    if (!(pm[3] > (real)1e-12)) continue;
                    const real *g = ggout + id * 3;
                    if (g[0] == 0 && g[1] == 0 && g[2] == 0) continue;
                    int I[3] = {i, j, l};
                    grid_node_rev(k, S, prims, ps0, ps1, I, pm, g, ggrid + id * 4, l0, l1);
                }
    }
    for (int t = 0; t < nthr; t++)
        for (int i = 0; i < npr * NPS; i++) { gps0[i] += acc0[(size_t)t * npr * NPS + i]; gps1[i] += acc1[(size_t)t * npr * NPS + i]; }
    free(acc0); free(acc1);
}

/* reverse of p2g + svd_grad + compute_F_tmp.grad (mpm_simulator.py:121-156,198-225) */
static void p2g_rev(const consts_t *k, int np, const real *st, const real *ggrid, const real *adj1, real *adj0) {
    int n = k->n;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int p = 0; p < np; p++) {
        const real *s = st + (size_t)p * NST;
        const real *a1 = adj1 + (size_t)p * NST;
        real *a0 = adj0 + (size_t)p * NST;
        ptrace_t t;
        particle_stress(k, s, &t);
        int base[3];
        real w[3][3], dw[3][3], fx[3];
        stencil(k, s, base, w, dw, fx);
        real gv[3] = {0, 0, 0}, gaff[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gfx[3] = {0, 0, 0};
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 3; c++) {
                    int I[3] = {base[0] + a, base[1] + b, base[2] + c};
                    if (!in_grid(n, I)) continue;
                    int off[3] = {a, b, c};
                    real wt = w[a][0] * w[b][1] * w[c][2];
                    real dpos[3] = {((real)a - fx[0]) * k->dx, ((real)b - fx[1]) * k->dx, ((real)c - fx[2]) * k->dx};
                    const real *gg = ggrid + ((size_t)(I[0] * n + I[1]) * n + I[2]) * 4;
                    real contrib[3], gw = 0, gdpos[3] = {0, 0, 0};
                    for (int i = 0; i < 3; i++) {
                        contrib[i] = k->p_mass * s[3 + i] + t.affine[i * 3] * dpos[0] + t.affine[i * 3 + 1] * dpos[1] + t.affine[i * 3 + 2] * dpos[2];
                        gw += gg[i] * contrib[i];
                        gv[i] += wt * k->p_mass * gg[i];
                        for (int j = 0; j < 3; j++) {
                            gaff[i * 3 + j] += wt * gg[i] * dpos[j];
                            gdpos[j] += wt * gg[i] * t.affine[i * 3 + j];
                        }
                    }
                    gw += gg[3] * k->p_mass;
                    for (int d = 0; d < 3; d++) {
                        real dwd = dw[off[d]][d];
                        for (int e = 0; e < 3; e++) if (e != d) dwd *= w[off[e]][e];
                        gfx[d] += gw * dwd - gdpos[d] * k->dx;
                    }
                }
        /* affine = stress_k * stress + p_mass * C */
        real gC[9], gst[9];
        for (int i = 0; i < 9; i++) { gC[i] = k->p_mass * gaff[i]; gst[i] = k->stress_k * gaff[i]; }
        /* stress = 2 mu (F - R) F^T + lam J (J - 1) I */
        real gF[9], tmp[9], tmp2[9], FmR[9];
        for (int i = 0; i < 9; i++) FmR[i] = t.Fn[i] - t.R[i];
        matmul3(gst, t.Fn, tmp);
        matmul3_tn(gst, FmR, tmp2);
        for (int i = 0; i < 9; i++) gF[i] = 2 * k->mu * (tmp[i] + tmp2[i]) + a1[15 + i];
        real gR[9];
        for (int i = 0; i < 9; i++) gR[i] = -2 * k->mu * tmp[i];
        real trs = gst[0] + gst[4] + gst[8];
        real gJ = k->lam * (2 * t.J - 1) * trs;
        real cf[9];
        cof3(t.Fn, cf);
        for (int i = 0; i < 9; i++) gF[i] += gJ * cf[i];
        /* R = U V^T */
        real gU[9], gV[9], gS[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gFt[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        matmul3(gR, t.V, gU);
        matmul3_tn(gR, t.U, gV);
        /* von Mises return map */
        if (t.yield) {
            real D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int i = 0; i < 3; i++) D[i * 4] = t.e[i];
            real gFV[9], gFtU[9], M[9];
            matmul3(gF, t.V, gFV);
            matmul3(gFV, D, tmp);
            for (int i = 0; i < 9; i++) gU[i] += tmp[i];
            matmul3_tn(gF, t.U, gFtU);
            matmul3(gFtU, D, tmp);
            for (int i = 0; i < 9; i++) gV[i] += tmp[i];
            matmul3_tn(t.U, gF, tmp);
            matmul3(tmp, t.V, M);
            real geps[3], gep[3];
            for (int i = 0; i < 3; i++) { gep[i] = M[i * 4] * t.e[i]; geps[i] = gep[i]; }
            real r = t.dg / t.ehn;
            real geh[3], gr = 0;
            for (int i = 0; i < 3; i++) { geh[i] = -r * gep[i]; gr -= gep[i] * t.eh[i]; }
            real gdg = gr / t.ehn;
            real gehn = -gr * t.dg / (t.ehn * t.ehn) + gdg;
            for (int i = 0; i < 3; i++) geh[i] += gehn * t.eh[i] / t.ehn;
            real mg = (geh[0] + geh[1] + geh[2]) / 3;
            for (int i = 0; i < 3; i++) geps[i] += geh[i] - mg;
            for (int i = 0; i < 3; i++)
                if ((real)0.05 < t.S[i * 4]) gS[i * 4] += geps[i] / t.Sc[i];
        } else {
            for (int i = 0; i < 9; i++) gFt[i] += gF[i];
        }
        real bs[9];
        backward_svd(gU, gS, gV, t.U, t.S, t.V, bs);
        for (int i = 0; i < 9; i++) gFt[i] += bs[i];
        /* F_tmp = (I + dt C) F */
        const real *C = s + 6, *F = s + 15;
        real A[9];
        for (int i = 0; i < 9; i++) A[i] = ((i % 4 == 0) ? 1 : 0) + k->dt * C[i];
        matmul3_nt(gFt, F, tmp);
        matmul3_tn(A, gFt, tmp2);
        for (int i = 0; i < 3; i++) { a0[i] += k->inv_dx * gfx[i]; a0[3 + i] += gv[i]; }
        for (int i = 0; i < 9; i++) { a0[6 + i] += gC[i] + k->dt * tmp[i]; a0[15 + i] += tmp2[i]; }
    }
}

/* substep_grad (mpm_simulator.py:325-345).  adj1: adjoint of frame s+1
 * (read), adj0: adjoint of frame s (accumulated).  gps0/gps1: adjoints of
 * the primitive states at s and s+1 (gps1 must already hold every other
 * contribution, as in the reference's LIFO order); gpv: adjoint of the
 * per-substep primitive velocity at s.  Scratch: grid n^3*4, gout n^3*3,
 * ggout n^3*3, ggrid n^3*4. */
void orc_substep_grad_c(const orc_sim *S, const orc_prim *prims, int np, const real *st, const real *adj1, real *adj0,
                        const real *ps_f, const real *pv_f, const real *ps_f1, real *gps_f, real *gps_f1, real *gpv_f,
                        real *grid, real *gout, real *ggout, real *ggrid, const orc_coll *c) {
    consts_t k;
    make_consts(S, &k);
    size_t G = (size_t)k.n * k.n * k.n;
    memset(grid, 0, G * 4 * sizeof(real));
    memset(ggout, 0, G * 3 * sizeof(real));
    memset(ggrid, 0, G * 4 * sizeof(real));
    p2g(&k, np, st, NULL, grid);
    grid_op(&k, S, prims, ps_f, ps_f1, grid, gout);
    g2p_rev(&k, np, st, gout, adj1, adj0, ggout);
    grid_op_rev(&k, S, prims, ps_f, ps_f1, grid, ggout, ggrid, gps_f, gps_f1);
    if (c && c->n_pairs > 0) tool_collide_rev(prims, c, ps_f1, gps_f1);
    for (int q = S->n_prims - 1; q >= 0; q--)
        prim_fk_rev(&prims[q], ps_f + q * NPS, pv_f + q * NPV, gps_f1 + q * NPS, gps_f + q * NPS, gpv_f + q * NPV);
    p2g_rev(&k, np, st, ggrid, adj1, adj0);
}
void orc_substep_grad(const orc_sim *S, const orc_prim *prims, int np, const real *st, const real *adj1, real *adj0,
                      const real *ps_f, const real *pv_f, const real *ps_f1, real *gps_f, real *gps_f1, real *gpv_f,
                      real *grid, real *gout, real *ggout, real *ggrid) {
    orc_substep_grad_c(S, prims, np, st, adj1, adj0, ps_f, pv_f, ps_f1, gps_f, gps_f1, gpv_f, grid, gout, ggout, ggrid, NULL);
}

/* compute_grid_m_kernel (mpm_simulator.py:456-466): adds into m (n^3) */
void orc_grid_mass(const orc_sim *S, int np, const real *st, real *m) {
    consts_t k;
    make_consts(S, &k);
    int n = k.n;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int p = 0; p < np; p++) {
        const real *s = st + (size_t)p * NST;
        int base[3];
        real w[3][3], dw[3][3], fx[3];
        stencil(&k, s, base, w, dw, fx);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 3; c++) {
                    int I[3] = {base[0] + a, base[1] + b, base[2] + c};
                    if (!in_grid(n, I)) continue;
                    real val = w[a][0] * w[b][1] * w[c][2] * k.p_mass;
#pragma omp atomic
                    m[((size_t)(I[0] * n + I[1]) * n + I[2])] += val;
                }
    }
}
/* compute_grid_m_kernel.grad: adds d(sum gm * m)/dx into adj[:, 0:3] */
void orc_grid_mass_grad(const orc_sim *S, int np, const real *st, const real *gm, real *adj) {
    consts_t k;
    make_consts(S, &k);
    int n = k.n;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int p = 0; p < np; p++) {
        const real *s = st + (size_t)p * NST;
        int base[3];
        real w[3][3], dw[3][3], fx[3];
        stencil(&k, s, base, w, dw, fx);
        real gfx[3] = {0, 0, 0};
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 3; c++) {
                    int I[3] = {base[0] + a, base[1] + b, base[2] + c};
                    if (!in_grid(n, I)) continue;
                    int off[3] = {a, b, c};
                    real g = gm[((size_t)(I[0] * n + I[1]) * n + I[2])] * k.p_mass;
                    for (int d = 0; d < 3; d++) {
                        real dwd = dw[off[d]][d];
                        for (int e = 0; e < 3; e++) if (e != d) dwd *= w[off[e]][e];
                        gfx[d] += g * dwd;
                    }
                }
        for (int d = 0; d < 3; d++) adj[(size_t)p * NST + d] += k.inv_dx * gfx[d];
    }
}

/* SDF of a primitive at world points (Primitive.sdf / Gripper.sdf: min over jaws) */
void orc_prim_sdf(const orc_prim *P, const real *ps, int npts, const real *x, real *out) {
    unit_t u[2];
    int m = prim_units(P, u);
    for (int i = 0; i < npts; i++) {
        real d = unit_sdf(&u[0], ps, x + 3 * i);
        if (m == 2) d = tmin(d, unit_sdf(&u[1], ps, x + 3 * i));
        out[i] = d;
    }
}
void orc_prim_normal(const orc_prim *P, const real *ps, int npts, const real *x, real *out) {
    unit_t u[2];
    int m = prim_units(P, u);
    for (int i = 0; i < npts; i++) {
        if (m == 2) {
            real a = unit_sdf(&u[0], ps, x + 3 * i), b = unit_sdf(&u[1], ps, x + 3 * i);
            unit_normal(&u[a <= b ? 0 : 1], ps, x + 3 * i, out + 3 * i);
        } else unit_normal(&u[0], ps, x + 3 * i, out + 3 * i);
    }
}
/* compute_min_dist (function.py:79-88): per particle, the SDF of every
 * primitive (Primitive.sdf), or for the two-jaw grippers both jaw SDFs
 * (sdf_2 with flag -1 then +1), columns in primitive order:
 * out[col * np + p].  Returns the number of columns. */
int orc_min_dist_cols(const orc_prim *prims, int n_prims) {
    int c = 0;
    unit_t u[2];
    for (int q = 0; q < n_prims; q++) c += prim_units(&prims[q], u);
    return c;
}
void orc_min_dist(const orc_prim *prims, int n_prims, int np, const real *st, const real *ps, real *out) {
    int col = 0;
    for (int q = 0; q < n_prims; q++) {
        unit_t u[2];
        int m = prim_units(&prims[q], u);
        for (int j = 0; j < m; j++, col++)
            for (int p = 0; p < np; p++) out[(size_t)col * np + p] = unit_sdf(&u[j], ps + q * NPS, st + (size_t)p * NST);
    }
}
/* compute_min_dist.grad (function.py:154-160): gout[col * np + p] is the
 * adjoint of each distance; adds d/dx into adj (n x 24, x rows) and the
 * primitive-state adjoint into gps (n_prims x 8). */
void orc_min_dist_grad(const orc_prim *prims, int n_prims, int np, const real *st, const real *ps, const real *gout, real *adj,
                       real *gps) {
    int col = 0;
    for (int q = 0; q < n_prims; q++) {
        unit_t u[2];
        int m = prim_units(&prims[q], u);
        for (int j = 0; j < m; j++, col++)
            for (int p = 0; p < np; p++) {
                real gd = gout[(size_t)col * np + p];
                if (gd == 0) continue;
                const real *x = st + (size_t)p * NST;
                const real *s = ps + q * NPS;
                real lp[3], g[3], ggap;
                to_local(s, u[j].jaw, u[j].norot, x, lp);
                local_sdf(u[j].P, u[j].sh, s[7], lp, g, &ggap);
                for (int i = 0; i < 3; i++) g[i] *= gd;
                if (u[j].sh == SH_CHOP) gps[q * NPS + 7] += ggap * gd;
                /* x enters as x - centre: its adjoint is that of the offset */
                real c[3], d[3], gx[3] = {0, 0, 0};
                unit_center(s, u[j].jaw, c);
                for (int i = 0; i < 3; i++) d[i] = x[i] - c[i];
                if (u[j].norot) memcpy(gx, g, sizeof(gx));
                else {
                    real iq[4], giq[4] = {0, 0, 0, 0};
                    conj_normalized(s + 3, iq);
                    qrot_rev(iq, d, g, giq, gx);
                }
                for (int i = 0; i < 3; i++) adj[(size_t)p * NST + i] += gx[i];
                to_local_rev(s, u[j].jaw, u[j].norot, x, g, gps + q * NPS);
            }
    }
}

/* one node of grid_op, for unit tests of the boundary / collision logic */
void orc_grid_node(const orc_sim *S, const orc_prim *prims, const real *ps0, const real *ps1, const int *I,
                   const real *pm, real *vout) {
    consts_t k;
    make_consts(S, &k);
    vout[0] = vout[1] = vout[2] = 0;
    if (pm[3] > (real)1e-12) grid_node(&k, S, prims, ps0, ps1, I, pm, vout, NULL);
}
void orc_grid_node_rev(const orc_sim *S, const orc_prim *prims, const real *ps0, const real *ps1, const int *I,
                       const real *pm, const real *gvout, real *gpm, real *gps0, real *gps1) {
    consts_t k;
    make_consts(S, &k);
    if (pm[3] > (real)1e-12) grid_node_rev(&k, S, prims, ps0, ps1, I, pm, gvout, gpm, gps0, gps1);
}
void orc_fk_rev(const orc_prim *prims, int n_prims, const real *ps_f, const real *pv_f, const real *gps_f1, real *gps_f, real *gpv_f) {
    for (int q = n_prims - 1; q >= 0; q--)
        prim_fk_rev(&prims[q], ps_f + q * NPS, pv_f + q * NPV, gps_f1 + q * NPS, gps_f + q * NPS, gpv_f + q * NPV);
}
/* F_new and stress of single particles (return-map unit tests) */
void orc_particle_stress(const orc_sim *S, int np, const real *st, real *Fn, real *stress) {
    consts_t k;
    make_consts(S, &k);
    for (int p = 0; p < np; p++) {
        ptrace_t t;
        particle_stress(&k, st + (size_t)p * NST, &t);
        memcpy(Fn + 9 * p, t.Fn, sizeof(t.Fn));
        memcpy(stress + 9 * p, t.stress, sizeof(t.stress));
    }
}
/* backward_svd of one matrix (unit tests against autograd of a library SVD) */
void orc_backward_svd(const real *gu, const real *gs, const real *gv, const real *u, const real *sig, const real *v, real *out) {
    backward_svd(gu, gs, gv, u, sig, v, out);
}
