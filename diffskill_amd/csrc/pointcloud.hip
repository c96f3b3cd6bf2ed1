// Point-cloud nearest neighbours on gfx950: the search under the Chamfer
// losses of the reference (plb/engine/losses/loss.py:183-204, and the
// chamferdist.ChamferDistance loss of plb/engine/taichi_env.py:28-32).
//
// Brute force over all (query, target) pairs, exact in fp32: a workgroup
// takes 256 queries (one per lane) and one chunk of targets staged in LDS;
// every lane scans the chunk (LDS broadcast reads) keeping its smallest
// squared distance and the first index reaching it.  Chunks meet through one
// 64-bit atomicMin per query on (distance bits << 32 | index): distances are
// non-negative, so their bit patterns order like the values, and equal
// distances keep the smaller index.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/diffskill_pointcloud.h"

#define DEV_PC __device__ __forceinline__

namespace {

constexpr int PC_WG = 256;     // queries per workgroup
constexpr int PC_CHUNK = 2048; // targets per workgroup (24 KB of LDS)
constexpr unsigned long long PC_EMPTY = 0xffffffffffffffffull;

thread_local std::string g_pc_err;
int pc_fail(const std::string& m) {
    g_pc_err = m;
    return -1;
}

// device address of `scratch` on each GPU, looked up once per device (with
// that device current) on the first pc_nearest_noalloc call there
constexpr int PC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch_dev[PC_MAX_DEVICES] = {};
int noalloc_scratch(unsigned long long** out);

__global__ void __launch_bounds__(PC_WG) k_pc_clear(unsigned long long* __restrict__ best, int n) {
    int i = blockIdx.x * PC_WG + threadIdx.x;
    if (i < n) best[i] = PC_EMPTY;
}

__global__ void __launch_bounds__(PC_WG) k_pc_nearest(const float* __restrict__ q, int nq, const float* __restrict__ t, int nt,
                                                     unsigned long long* __restrict__ best) {
    __shared__ float sx[PC_CHUNK], sy[PC_CHUNK], sz[PC_CHUNK];
    int t0 = blockIdx.y * PC_CHUNK;
    int nc = min(PC_CHUNK, nt - t0);
    for (int k = threadIdx.x; k < nc; k += PC_WG) {
        const float* p = t + (size_t)(t0 + k) * 3;
        sx[k] = p[0];
        sy[k] = p[1];
        sz[k] = p[2];
    }
    __syncthreads();
    int i = blockIdx.x * PC_WG + threadIdx.x;
    if (i >= nq) return;
    float qx = q[(size_t)i * 3], qy = q[(size_t)i * 3 + 1], qz = q[(size_t)i * 3 + 2];
    float bd = __int_as_float(0x7f800000);
    int bk = -1;
    int k = 0;
    // four independent distances per step, compared in index order
    for (; k + 4 <= nc; k += 4) {
        float d[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            float ex = qx - sx[k + u], ey = qy - sy[k + u], ez = qz - sz[k + u];
            d[u] = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (d[u] < bd) { bd = d[u]; bk = k + u; }
    }
    for (; k < nc; k++) {
        float ex = qx - sx[k], ey = qy - sy[k], ez = qz - sz[k];
        float d = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
        if (d < bd) { bd = d; bk = k; }
    }
    if (bk >= 0)
        atomicMin(&best[i], ((unsigned long long)__float_as_uint(bd) << 32) | (unsigned)(t0 + bk));
}

// Scratch of pc_nearest_noalloc: the per-query (distance, index) words that
// pc_nearest takes from its workspace argument, as a file-scope device array
// (one copy per GPU, held until the process ends).  Host code reaches it only
// through the address hipGetSymbolAddress returns for the current device.
constexpr int PC_NOALLOC_MAX_QUERIES = 1 << 20;   // 8 MiB per GPU
static __device__ __align__(256) unsigned long long scratch[PC_NOALLOC_MAX_QUERIES];

int noalloc_scratch(unsigned long long** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return pc_fail(std::string("pc_nearest_noalloc: hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= PC_MAX_DEVICES) return pc_fail("pc_nearest_noalloc: device index outside [0, 64)");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return pc_fail(std::string("pc_nearest_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<unsigned long long*>(p);
    }
    *out = g_scratch_dev[dev];
    return 0;
}

__global__ void __launch_bounds__(PC_WG) k_pc_unpack(const unsigned long long* __restrict__ best, int n, int* __restrict__ idx,
                                                    float* __restrict__ d2) {
    int i = blockIdx.x * PC_WG + threadIdx.x;
    if (i >= n) return;
    unsigned long long b = best[i];
    bool none = b == PC_EMPTY;
    idx[i] = none ? -1 : (int)(unsigned)(b & 0xffffffffull);
    d2[i] = none ? __int_as_float(0x7f800000) : __uint_as_float((unsigned)(b >> 32));
}

// ---------------------------------------------------------------- Sinkhorn
// The soft-min of the Sinkhorn divergence (geomloss SamplesLoss('sinkhorn'),
// the EMD loss of plb/engine/taichi_env.py:23-26) and the gradient of its
// transport plan, over all (row, column) pairs:
//   softmin_i = -eps log sum_j exp(h_j - C(x_i, y_j) / eps)
//   grad_i    = sum_j exp(h_j - C(x_i, y_j) / eps + r_i) dC/dx_i
// with C = |x - y| (p = 1) or |x - y|^2 / 2 (p = 2).  Base-2 arithmetic: the
// host passes h and r in log2 units (h * log2 e), the kernels scale the cost
// by log2(e) / eps and exponentiate with v_exp_f32.  A workgroup takes 256
// rows (one per lane) and one column range (a "split") staged through LDS;
// each lane keeps a running (max, scaled sum) per block of 4 columns, and the
// splits' partials are combined in split order by a second kernel, so the
// result does not depend on scheduling.
constexpr int SK_CHUNK = 2048;   // columns per LDS pass (32 KB)

DEV_PC float sk_cost(float dx, float dy, float dz, int p) {
    float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    return p == 1 ? sqrtf(d2) : 0.5f * d2;
}

__global__ void __launch_bounds__(PC_WG) k_sk_softmin(const float* __restrict__ x, int n, const float* __restrict__ y, int m,
                                                     const float* __restrict__ h2, float k2, int p, int per_split,
                                                     float2* __restrict__ part) {
    __shared__ float sx[SK_CHUNK], sy[SK_CHUNK], sz[SK_CHUNK], sh[SK_CHUNK];
    int i = blockIdx.x * PC_WG + threadIdx.x;
    int c0 = blockIdx.y * per_split, c1 = min(m, c0 + per_split);
    float xi = 0.f, yi = 0.f, zi = 0.f;
    if (i < n) { xi = x[(size_t)i * 3]; yi = x[(size_t)i * 3 + 1]; zi = x[(size_t)i * 3 + 2]; }
    float mx = -INFINITY, sm = 0.f;
    for (int b0 = c0; b0 < c1; b0 += SK_CHUNK) {
        int nc = min(SK_CHUNK, c1 - b0);
        __syncthreads();
        for (int k = threadIdx.x; k < nc; k += PC_WG) {
            const float* q = y + (size_t)(b0 + k) * 3;
            sx[k] = q[0]; sy[k] = q[1]; sz[k] = q[2]; sh[k] = h2[b0 + k];
        }
        __syncthreads();
        int k = 0;
        for (; k + 4 <= nc; k += 4) {
            float z[4];
#pragma unroll
            for (int u = 0; u < 4; u++) z[u] = sh[k + u] - k2 * sk_cost(xi - sx[k + u], yi - sy[k + u], zi - sz[k + u], p);
            float bm = fmaxf(fmaxf(z[0], z[1]), fmaxf(z[2], z[3]));
            float nm = fmaxf(mx, bm);
            float acc = 0.f;
#pragma unroll
            for (int u = 0; u < 4; u++) acc += __builtin_amdgcn_exp2f(z[u] - nm);
            sm = sm * __builtin_amdgcn_exp2f(mx - nm) + acc;
            mx = nm;
        }
        for (; k < nc; k++) {
            float z = sh[k] - k2 * sk_cost(xi - sx[k], yi - sy[k], zi - sz[k], p);
            float nm = fmaxf(mx, z);
            sm = sm * __builtin_amdgcn_exp2f(mx - nm) + __builtin_amdgcn_exp2f(z - nm);
            mx = nm;
        }
    }
    if (i < n) part[(size_t)blockIdx.y * n + i] = make_float2(mx, sm);
}

// combine the splits: softmin_i = -eps ln2 (M + log2 sum_s s_s 2^(m_s - M))
__global__ void __launch_bounds__(PC_WG) k_sk_softmin_combine(const float2* __restrict__ part, int n, int splits, float eps,
                                                             float* __restrict__ out) {
    int i = blockIdx.x * PC_WG + threadIdx.x;
    if (i >= n) return;
    float M = -INFINITY;
    for (int s = 0; s < splits; s++) M = fmaxf(M, part[(size_t)s * n + i].x);
    float S = 0.f;
    for (int s = 0; s < splits; s++) {
        float2 v = part[(size_t)s * n + i];
        S += v.y * __builtin_amdgcn_exp2f(v.x - M);
    }
    out[i] = -eps * 0.69314718055994531f * (M + __builtin_amdgcn_logf(S));
}

__global__ void __launch_bounds__(PC_WG) k_sk_plan_grad(const float* __restrict__ x, int n, const float* __restrict__ y, int m,
                                                       const float* __restrict__ h2, const float* __restrict__ r2, float k2, int p,
                                                       int per_split, float4* __restrict__ part) {
    __shared__ float sx[SK_CHUNK], sy[SK_CHUNK], sz[SK_CHUNK], sh[SK_CHUNK];
    int i = blockIdx.x * PC_WG + threadIdx.x;
    int c0 = blockIdx.y * per_split, c1 = min(m, c0 + per_split);
    float xi = 0.f, yi = 0.f, zi = 0.f, ri = 0.f;
    if (i < n) { xi = x[(size_t)i * 3]; yi = x[(size_t)i * 3 + 1]; zi = x[(size_t)i * 3 + 2]; ri = r2[i]; }
    float gx = 0.f, gy = 0.f, gz = 0.f;
    for (int b0 = c0; b0 < c1; b0 += SK_CHUNK) {
        int nc = min(SK_CHUNK, c1 - b0);
        __syncthreads();
        for (int k = threadIdx.x; k < nc; k += PC_WG) {
            const float* q = y + (size_t)(b0 + k) * 3;
            sx[k] = q[0]; sy[k] = q[1]; sz[k] = q[2]; sh[k] = h2[b0 + k];
        }
        __syncthreads();
        for (int k = 0; k < nc; k++) {
            float dx = xi - sx[k], dy = yi - sy[k], dz = zi - sz[k];
            float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            float c, w;
            if (p == 1) {
                float d = sqrtf(d2);
                c = d;
                w = d > 0.f ? 1.f / d : 0.f;   // dC/dx = (x - y) / |x - y|, 0 at coincident points
            } else {
                c = 0.5f * d2;
                w = 1.f;
            }
            float e = __builtin_amdgcn_exp2f(sh[k] + ri - k2 * c) * w;
            gx = fmaf(e, dx, gx); gy = fmaf(e, dy, gy); gz = fmaf(e, dz, gz);
        }
    }
    if (i < n) part[(size_t)blockIdx.y * n + i] = make_float4(gx, gy, gz, 0.f);
}

__global__ void __launch_bounds__(PC_WG) k_sk_grad_combine(const float4* __restrict__ part, int n, int splits, float* __restrict__ g) {
    int i = blockIdx.x * PC_WG + threadIdx.x;
    if (i >= n) return;
    float gx = 0.f, gy = 0.f, gz = 0.f;
    for (int s = 0; s < splits; s++) {
        float4 v = part[(size_t)s * n + i];
        gx += v.x; gy += v.y; gz += v.z;
    }
    g[(size_t)i * 3] = gx; g[(size_t)i * 3 + 1] = gy; g[(size_t)i * 3 + 2] = gz;
}

// One annealing step of the divergence, all four soft-mins in one launch
// (blockIdx.z = problem) and their combination in a second: the potentials
// live in one device array pot = [a_x, b_x, a_y, b_y, pa_x, pb_x, pa_y, pb_y]
// (lengths n, n, m, m, n, n, m, m), and the column weights h are formed while
// the columns are staged (uniform log weights plus potential / eps):
//   problem 0: a_y <- softmin over x of (log 1/n + b_x / eps)   rows y, cols x
//   problem 1: b_x <- softmin over y of (log 1/m + a_y / eps)   rows x, cols y
//   problem 2: a_x <- softmin over x of (log 1/n + a_x / eps)   rows x, cols x
//   problem 3: b_y <- softmin over y of (log 1/m + b_y / eps)   rows y, cols y
// mode 0 (first eps): h = log weights only, the result replaces the potential;
// mode 1: symmetric update, potential <- (potential + result) / 2;
// mode 2 (last extrapolation): the potentials move to pa_x.. and the results
// replace them.
struct SkProb { const float* rows; int nr; const float* cols; int nc; const float* hsrc; float lw; float* dst; float* prev; };
DEV_PC SkProb sk_problem(int z, const float* x, int n, const float* y, int m, float* pot) {
    float* a_x = pot; float* b_x = pot + n; float* a_y = pot + 2 * n; float* b_y = pot + 2 * n + m;
    float* pa_x = pot + 2 * n + 2 * m; float* pb_x = pa_x + n; float* pa_y = pa_x + 2 * n; float* pb_y = pa_y + m;
    SkProb q;
    if (z == 0) { q.rows = y; q.nr = m; q.cols = x; q.nc = n; q.hsrc = b_x; q.dst = a_y; q.prev = pa_y; }
    else if (z == 1) { q.rows = x; q.nr = n; q.cols = y; q.nc = m; q.hsrc = a_y; q.dst = b_x; q.prev = pb_x; }
    else if (z == 2) { q.rows = x; q.nr = n; q.cols = x; q.nc = n; q.hsrc = a_x; q.dst = a_x; q.prev = pa_x; }
    else { q.rows = y; q.nr = m; q.cols = y; q.nc = m; q.hsrc = b_y; q.dst = b_y; q.prev = pb_y; }
    q.lw = -logf((float)q.nc);
    return q;
}

__global__ void __launch_bounds__(PC_WG) k_sk_step(const float* __restrict__ x, int n, const float* __restrict__ y, int m,
                                                  float* __restrict__ pot, float eps, int p, int mode, int splits_xy, int splits_xx,
                                                  int splits_yy, float2* __restrict__ part, size_t part_stride) {
    __shared__ float sx[SK_CHUNK], sy[SK_CHUNK], sz[SK_CHUNK], sh[SK_CHUNK];
    int z = blockIdx.z;
    SkProb q = sk_problem(z, x, n, y, m, pot);
    int splits = z < 2 ? splits_xy : (z == 2 ? splits_xx : splits_yy);
    if ((int)blockIdx.y >= splits || (int)blockIdx.x * PC_WG >= q.nr) return;
    int per = (q.nc + splits - 1) / splits;
    int i = blockIdx.x * PC_WG + threadIdx.x;
    int c0 = blockIdx.y * per, c1 = min(q.nc, c0 + per);
    float k2 = 1.4426950408889634f / eps, inv = 1.f / eps;
    float xi = 0.f, yi = 0.f, zi = 0.f;
    if (i < q.nr) { xi = q.rows[(size_t)i * 3]; yi = q.rows[(size_t)i * 3 + 1]; zi = q.rows[(size_t)i * 3 + 2]; }
    float mx = -INFINITY, sm = 0.f;
    for (int b0 = c0; b0 < c1; b0 += SK_CHUNK) {
        int nc = min(SK_CHUNK, c1 - b0);
        __syncthreads();
        for (int k = threadIdx.x; k < nc; k += PC_WG) {
            const float* c = q.cols + (size_t)(b0 + k) * 3;
            sx[k] = c[0]; sy[k] = c[1]; sz[k] = c[2];
            float h = q.lw + (mode == 0 ? 0.f : q.hsrc[b0 + k] * inv);
            sh[k] = h * 1.4426950408889634f;
        }
        __syncthreads();
        int k = 0;
        for (; k + 4 <= nc; k += 4) {
            float zz[4];
#pragma unroll
            for (int u = 0; u < 4; u++) zz[u] = sh[k + u] - k2 * sk_cost(xi - sx[k + u], yi - sy[k + u], zi - sz[k + u], p);
            float bm = fmaxf(fmaxf(zz[0], zz[1]), fmaxf(zz[2], zz[3]));
            float nm = fmaxf(mx, bm);
            float acc = 0.f;
#pragma unroll
            for (int u = 0; u < 4; u++) acc += __builtin_amdgcn_exp2f(zz[u] - nm);
            sm = sm * __builtin_amdgcn_exp2f(mx - nm) + acc;
            mx = nm;
        }
        for (; k < nc; k++) {
            float zz = sh[k] - k2 * sk_cost(xi - sx[k], yi - sy[k], zi - sz[k], p);
            float nm = fmaxf(mx, zz);
            sm = sm * __builtin_amdgcn_exp2f(mx - nm) + __builtin_amdgcn_exp2f(zz - nm);
            mx = nm;
        }
    }
    if (i < q.nr) part[(size_t)z * part_stride + (size_t)blockIdx.y * q.nr + i] = make_float2(mx, sm);
}

__global__ void __launch_bounds__(PC_WG) k_sk_step_combine(const float* __restrict__ x, int n, const float* __restrict__ y, int m,
                                                          float* __restrict__ pot, float eps, int mode, int splits_xy, int splits_xx,
                                                          int splits_yy, const float2* __restrict__ part, size_t part_stride) {
    int z = blockIdx.y;
    SkProb q = sk_problem(z, x, n, y, m, pot);
    int splits = z < 2 ? splits_xy : (z == 2 ? splits_xx : splits_yy);
    int i = blockIdx.x * PC_WG + threadIdx.x;
    if (i >= q.nr) return;
    const float2* pp = part + (size_t)z * part_stride;
    float M = -INFINITY;
    for (int s = 0; s < splits; s++) M = fmaxf(M, pp[(size_t)s * q.nr + i].x);
    float S = 0.f;
    for (int s = 0; s < splits; s++) {
        float2 v = pp[(size_t)s * q.nr + i];
        S += v.y * __builtin_amdgcn_exp2f(v.x - M);
    }
    float r = -eps * 0.69314718055994531f * (M + __builtin_amdgcn_logf(S));
    if (mode == 1) r = 0.5f * (q.dst[i] + r);
    if (mode == 2) q.prev[i] = q.dst[i];
    q.dst[i] = r;
}

// column splits: enough workgroups to fill the GPU (about 2048), at least
// 256 columns each
int sk_splits(int n, int m) {
    int gx = (n + PC_WG - 1) / PC_WG;
    int s = (2048 + gx - 1) / gx;
    s = std::min(s, (m + 255) / 256);
    return std::max(1, std::min(s, 65535));
}

}  // namespace

extern "C" {

long long pc_sinkhorn_work_bytes(int n, int m) {
    if (n < 0 || m < 0) return -1;
    return (long long)sk_splits(n, m) * std::max(n, 1) * (long long)sizeof(float4);
}

long long pc_sinkhorn_step_work_bytes(int n, int m) {
    if (n <= 0 || m <= 0) return -1;
    long long rows = std::max(n, m);
    int s = std::max(std::max(sk_splits(n, m), sk_splits(m, n)), std::max(sk_splits(n, n), sk_splits(m, m)));
    return 4LL * s * rows * (long long)sizeof(float2);
}

int pc_sinkhorn_step(const float* x, int n, const float* y, int m, float* pot, float eps, int p, int mode, void* work, void* stream) {
    if (n <= 0 || m <= 0) return pc_fail("pc_sinkhorn_step: both clouds must be non-empty");
    if (!(eps > 0.f)) return pc_fail("pc_sinkhorn_step: eps must be positive");
    if (p != 1 && p != 2) return pc_fail("pc_sinkhorn_step: p must be 1 or 2");
    if (mode < 0 || mode > 2) return pc_fail("pc_sinkhorn_step: mode must be 0, 1 or 2");
    if (!x || !y || !pot || !work) return pc_fail("pc_sinkhorn_step: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // splits per problem (rows, cols): y-x and x-y share max(splits(m, n), splits(n, m)) slots
    int sxy = std::max(sk_splits(m, n), sk_splits(n, m)), sxx = sk_splits(n, n), syy = sk_splits(m, m);
    int rows = std::max(n, m);
    size_t stride = (size_t)std::max(std::max(sxy, sxx), syy) * rows;
    int gx = (rows + PC_WG - 1) / PC_WG;
    int gy = std::max(std::max(sxy, sxx), syy);
    hipLaunchKernelGGL(k_sk_step, dim3(gx, gy, 4), dim3(PC_WG), 0, st, x, n, y, m, pot, eps, p, mode, sxy, sxx, syy, (float2*)work,
                       stride);
    hipLaunchKernelGGL(k_sk_step_combine, dim3(gx, 4), dim3(PC_WG), 0, st, x, n, y, m, pot, eps, mode, sxy, sxx, syy,
                       (const float2*)work, stride);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return pc_fail(std::string("pc_sinkhorn_step: ") + hipGetErrorString(e));
    return 0;
}

int pc_softmin(const float* x, int n, const float* y, int m, const float* h2, float eps, int p, void* work, float* out, void* stream) {
    if (n < 0 || m <= 0) return pc_fail("pc_softmin: need n >= 0 rows and m > 0 columns");
    if (!(eps > 0.f)) return pc_fail("pc_softmin: eps must be positive");
    if (p != 1 && p != 2) return pc_fail("pc_softmin: p must be 1 or 2");
    if (n == 0) return 0;
    if (!x || !y || !h2 || !work || !out) return pc_fail("pc_softmin: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    int splits = sk_splits(n, m);
    int per = (m + splits - 1) / splits;
    int gx = (n + PC_WG - 1) / PC_WG;
    float k2 = 1.4426950408889634f / eps;
    hipLaunchKernelGGL(k_sk_softmin, dim3(gx, splits), dim3(PC_WG), 0, st, x, n, y, m, h2, k2, p, per, (float2*)work);
    hipLaunchKernelGGL(k_sk_softmin_combine, dim3(gx), dim3(PC_WG), 0, st, (const float2*)work, n, splits, eps, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return pc_fail(std::string("pc_softmin: ") + hipGetErrorString(e));
    return 0;
}

int pc_plan_grad(const float* x, int n, const float* y, int m, const float* h2, const float* r2, float eps, int p, void* work,
                 float* g, void* stream) {
    if (n < 0 || m <= 0) return pc_fail("pc_plan_grad: need n >= 0 rows and m > 0 columns");
    if (!(eps > 0.f)) return pc_fail("pc_plan_grad: eps must be positive");
    if (p != 1 && p != 2) return pc_fail("pc_plan_grad: p must be 1 or 2");
    if (n == 0) return 0;
    if (!x || !y || !h2 || !r2 || !work || !g) return pc_fail("pc_plan_grad: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    int splits = sk_splits(n, m);
    int per = (m + splits - 1) / splits;
    int gx = (n + PC_WG - 1) / PC_WG;
    float k2 = 1.4426950408889634f / eps;
    hipLaunchKernelGGL(k_sk_plan_grad, dim3(gx, splits), dim3(PC_WG), 0, st, x, n, y, m, h2, r2, k2, p, per, (float4*)work);
    hipLaunchKernelGGL(k_sk_grad_combine, dim3(gx), dim3(PC_WG), 0, st, (const float4*)work, n, splits, g);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return pc_fail(std::string("pc_plan_grad: ") + hipGetErrorString(e));
    return 0;
}


const char* pc_last_error(void) { return g_pc_err.c_str(); }

int pc_nearest(const float* q, int nq, const float* t, int nt, unsigned long long* work, int* idx, float* d2, void* stream) {
    if (nq < 0 || nt < 0) return pc_fail("pc_nearest: negative point count");
    if (nq == 0) return 0;
    if (!q || !work || !idx || !d2 || (nt > 0 && !t)) return pc_fail("pc_nearest: null argument");
    int gc = (nt + PC_CHUNK - 1) / PC_CHUNK;
    if (gc > 65535) return pc_fail("pc_nearest: too many targets (at most 65535 x 2048)");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    int gq = (nq + PC_WG - 1) / PC_WG;
    hipLaunchKernelGGL(k_pc_clear, dim3(gq), dim3(PC_WG), 0, st, work, nq);
    if (gc > 0) hipLaunchKernelGGL(k_pc_nearest, dim3(gq, gc), dim3(PC_WG), 0, st, q, nq, t, nt, work);
    hipLaunchKernelGGL(k_pc_unpack, dim3(gq), dim3(PC_WG), 0, st, (const unsigned long long*)work, nq, idx, d2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return pc_fail(std::string("pc_nearest: ") + hipGetErrorString(e));
    return 0;
}

int pc_nearest_noalloc(const float* q, int nq, const float* t, int nt, int* idx, float* d2, void* stream) {
    if (nq < 0 || nt < 0) return pc_fail("pc_nearest_noalloc: negative point count");
    if (nq > PC_NOALLOC_MAX_QUERIES)
        return pc_fail("pc_nearest_noalloc: " + std::to_string(nq) + " queries exceed the static scratch of " +
                       std::to_string(PC_NOALLOC_MAX_QUERIES) + " (use pc_nearest with a workspace)");
    if (nq == 0) return 0;
    unsigned long long* work = nullptr;
    if (noalloc_scratch(&work)) return -1;
    // the scratch keeps the previous call's words; pc_nearest clears the
    // first nq of them (a fill kernel) before the search
    return pc_nearest(q, nq, t, nt, work, idx, d2, stream);
}

}  // extern "C"
