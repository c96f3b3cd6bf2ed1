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
#include <cstdint>
#include <string>

#include "../../include/diffskill_pointcloud.h"

namespace {

constexpr int PC_WG = 256;     // queries per workgroup
constexpr int PC_CHUNK = 2048; // targets per workgroup (24 KB of LDS)
constexpr unsigned long long PC_EMPTY = 0xffffffffffffffffull;

thread_local std::string g_pc_err;
int pc_fail(const std::string& m) {
    g_pc_err = m;
    return -1;
}

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

__global__ void __launch_bounds__(PC_WG) k_pc_unpack(const unsigned long long* __restrict__ best, int n, int* __restrict__ idx,
                                                    float* __restrict__ d2) {
    int i = blockIdx.x * PC_WG + threadIdx.x;
    if (i >= n) return;
    unsigned long long b = best[i];
    bool none = b == PC_EMPTY;
    idx[i] = none ? -1 : (int)(unsigned)(b & 0xffffffffull);
    d2[i] = none ? __int_as_float(0x7f800000) : __uint_as_float((unsigned)(b >> 32));
}

}  // namespace

extern "C" {

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

}  // extern "C"
