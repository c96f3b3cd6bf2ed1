// MI355X (gfx950) differentiable MLS-MPM engine: kernels and the C ABI.
//
// One context packs any number of scenes (each with its own grid and tool
// primitives) into shared device arrays, so every phase of a substep is ONE
// launch over all scenes:
//   forward  : mark(+primitive FK) -> compact -> p2g -> grid_op -> g2p
//   adjoint  : mark -> compact -> p2g -> grid_op(keep) -> g2p_bwd ->
//              grid_op_bwd -> p2g_bwd(+primitive FK adjoint)
// The reference runs the same sequence as ~7 + n_prim Taichi launches per
// scene and substep (plb/engine/mpm_simulator.py:307-345).
//
// Layouts in HBM:
//   particles  SoA, 24 rows of `stride` floats per frame (x, v, C, F), each
//              scene a 64-aligned column range; frames live in checkpoint
//              slots (every K-th frame) plus a ring of K-1 live frames.
//   grid       block-major: 4x4x4 node blocks of 64 float4 (1 KiB); only
//              blocks touched by particles ("active") are visited.
//   primitives [frame][prim][8] (pos, quat wxyz, gap), velocities [frame][prim][8].
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/diffskill_mpm.h"
#include "mpm_device.h"

using namespace dsk;

namespace {

thread_local std::string g_err;
int fail(const std::string& m) { g_err = m; return -1; }
#define HIPCHK(x)                                                              \
    do {                                                                       \
        hipError_t e_ = (x);                                                   \
        if (e_ != hipSuccess) return fail(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int NST = 24;      // particle state rows
constexpr int PPB = 256;     // particles per chunk / threads per particle block
constexpr int NODES_PER_BLK = 64;
constexpr int MAX_UNITS = 8;

// ------------------------------------------------------------------ device helpers
DEV long long node_id(const SceneC& S, int i, int j, int k) {
    int b = ((i >> 2) * S.nb + (j >> 2)) * S.nb + (k >> 2);
    return S.g_off + (long long)b * NODES_PER_BLK + (((i & 3) << 4) | ((j & 3) << 2) | (k & 3));
}
DEV bool in_grid(const SceneC& S, int i, int j, int k) {
    return (unsigned)i < (unsigned)S.n_grid && (unsigned)j < (unsigned)S.n_grid && (unsigned)k < (unsigned)S.n_grid;
}
struct Stencil {
    int base[3];
    float fx[3];
    float w[3][3];   // w[offset][dim]
    float dw[3][3];
};
// quadratic B-spline weights; base uses truncation (cast(int)), mpm_simulator.py:201-204
DEV void stencil(const SceneC& S, v3 x, Stencil& st) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
        float xs = comp(x, d) * S.inv_dx;
        int b = (int)(xs - 0.5f);
        float f = xs - (float)b;
        st.base[d] = b;
        st.fx[d] = f;
        st.w[0][d] = 0.5f * (1.5f - f) * (1.5f - f);
        st.w[1][d] = 0.75f - (f - 1.f) * (f - 1.f);
        st.w[2][d] = 0.5f * (f - 0.5f) * (f - 0.5f);
        st.dw[0][d] = f - 1.5f;
        st.dw[1][d] = -2.f * (f - 1.f);
        st.dw[2][d] = f - 0.5f;
    }
}
DEV v3 load3(const float* st, int stride, int row, int slot) {
    return mk(st[(size_t)row * stride + slot], st[(size_t)(row + 1) * stride + slot], st[(size_t)(row + 2) * stride + slot]);
}
DEV m3 load9(const float* st, int stride, int row, int slot) {
    m3 m;
#pragma unroll
    for (int i = 0; i < 9; i++) m.a[i] = st[(size_t)(row + i) * stride + slot];
    return m;
}
DEV void store3(float* st, int stride, int row, int slot, v3 v) {
    st[(size_t)row * stride + slot] = v.x;
    st[(size_t)(row + 1) * stride + slot] = v.y;
    st[(size_t)(row + 2) * stride + slot] = v.z;
}
DEV void store9(float* st, int stride, int row, int slot, const m3& m) {
#pragma unroll
    for (int i = 0; i < 9; i++) st[(size_t)(row + i) * stride + slot] = m.a[i];
}

// per-particle constitutive update: F_tmp, SVD, von Mises return map, stress
// (mpm_simulator.py:121-129,165-182,205-216)
struct PTrace {
    m3 Ftmp, U, V, Fn, R, affine;
    float sig[3], sc[3], eh[3], ehn, dg, e[3], J;
    bool yield;
};
// F_tmp = (I + dt C) F (mpm_simulator.py:121-124)
DEV m3 f_tmp(const SceneC& S, const m3& C, const m3& F) {
    m3 A;
#pragma unroll
    for (int i = 0; i < 9; i++) A.a[i] = ((i % 4 == 0) ? 1.f : 0.f) + S.dt * C.a[i];
    return mmul(A, F);
}
// everything after the SVD: return map and stress, with t.Ftmp, t.U, t.sig, t.V set
DEV void stress_from_svd(const SceneC& S, const m3& C, PTrace& t) {
    float eps[3];
#pragma unroll
    for (int i = 0; i < 3; i++) { t.sc[i] = tmax(t.sig[i], 0.05f); eps[i] = logf(t.sc[i]); }
    float mean = (eps[0] + eps[1] + eps[2]) / 3.f;
#pragma unroll
    for (int i = 0; i < 3; i++) t.eh[i] = eps[i] - mean;
    t.ehn = sqrtf(t.eh[0] * t.eh[0] + t.eh[1] * t.eh[1] + t.eh[2] * t.eh[2] + 1e-8f);
    t.dg = t.ehn - S.ys / (2.f * S.mu);
    t.yield = t.dg > 0.f;
    if (t.yield) {
        m3 D = mzero();
#pragma unroll
        for (int i = 0; i < 3; i++) { t.e[i] = expf(eps[i] - (t.dg / t.ehn) * t.eh[i]); D.a[i * 4] = t.e[i]; }
        t.Fn = mmul_nt(mmul(t.U, D), t.V);
    } else {
        t.Fn = t.Ftmp;
    }
    t.J = mdet(t.Fn);
    t.R = mmul_nt(t.U, t.V);
    m3 FmR;
#pragma unroll
    for (int i = 0; i < 9; i++) FmR.a[i] = t.Fn.a[i] - t.R.a[i];
    m3 P1 = mmul_nt(FmR, t.Fn);
#pragma unroll
    for (int i = 0; i < 9; i++) {
        float s = 2.f * S.mu * P1.a[i] + ((i % 4 == 0) ? S.lam * t.J * (t.J - 1.f) : 0.f);
        t.affine.a[i] = S.stress_k * s + S.p_mass * C.a[i];
    }
}
DEV void particle_stress(const SceneC& S, const m3& C, const m3& F, PTrace& t) {
    t.Ftmp = f_tmp(S, C, F);
    svd3(t.Ftmp, t.U, t.sig, t.V);
    stress_from_svd(S, C, t);
}
// rows of the stored SVD of F_tmp per frame slot: U as a unit quaternion (U and
// V are proper rotations: products of Givens / Jacobi rotations, column swaps
// with a sign flip), sigma, V as a unit quaternion
constexpr int USV_ROWS = 11;
DEV void store_usv(float* usv, int stride, int p, const PTrace& t) {
    q4 qu = rot_to_quat(t.U.a), qv = rot_to_quat(t.V.a);
    float r[USV_ROWS] = {qu.w, qu.x, qu.y, qu.z, t.sig[0], t.sig[1], t.sig[2], qv.w, qv.x, qv.y, qv.z};
#pragma unroll
    for (int k = 0; k < USV_ROWS; k++) usv[(size_t)k * stride + p] = r[k];
}
DEV void load_usv(const float* usv, int stride, int p, PTrace& t) {
    float r[USV_ROWS];
#pragma unroll
    for (int k = 0; k < USV_ROWS; k++) r[k] = usv[(size_t)k * stride + p];
    quat_to_rot(mkq(r[0], r[1], r[2], r[3]), t.U.a);
    t.sig[0] = r[4]; t.sig[1] = r[5]; t.sig[2] = r[6];
    quat_to_rot(mkq(r[7], r[8], r[9], r[10]), t.V.a);
}

// grid_op for one node with m > 1e-12 (mpm_simulator.py:230-262)
// the node chain's intermediate velocities for its adjoint: before each
// collision unit (in LDS, vu[(unit * 3 + c) * 64 + lane]: indexed by the unit
// at run time, which in registers would be a scratch array in memory) and
// before each wall stage (registers)
struct NodeTr { float* vu; v3 vw[3]; };
DEV v3 grid_node(const SceneC& S, const PrimC* prims, const float* ps0, const float* ps1, int I0, int I1, int I2,
                 float4 pm, NodeTr* tr) {
    float inv_m = 1.f / pm.w;
    v3 v = mk(inv_m * pm.x, inv_m * pm.y, inv_m * pm.z);
    v += mk(S.g30[0], S.g30[1], S.g30[2]);
    v3 x = mk((float)I0 * S.dx, (float)I1 * S.dx, (float)I2 * S.dx);
    int nu = 0;
    for (int q = 0; q < S.n_prims; q++) {
        const PrimC& P = prims[q];
        PS s0 = load_ps(ps0 + q * 8), s1 = load_ps(ps1 + q * 8);
        int m = prim_unit_count(P);
        for (int j = 0; j < m; j++) {
            if (tr && nu < MAX_UNITS) {
                int lane = threadIdx.x & 63;
                tr->vu[(nu * 3 + 0) * 64 + lane] = v.x;
                tr->vu[(nu * 3 + 1) * 64 + lane] = v.y;
                tr->vu[(nu * 3 + 2) * 64 + lane] = v.z;
            }
            v = unit_collide(P, prim_unit(P, j), s0, s1, x, S.dt, v, nullptr);
            nu++;
        }
    }
    int I[3] = {I0, I1, I2};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (tr) tr->vw[d] = v;
        if (I[d] < 3 && comp(v, d) < 0.f) {
            if (d != 1 || S.gf_mode == 0) setc(v, d, 0.f);
            else if (S.gf_mode == 1) {
                float lin = v.y + 1e-30f;
                v3 vit = mk(v.x - (float)I0 * 1e-30f, v.y - lin - (float)I1 * 1e-30f, v.z - (float)I2 * 1e-30f);
                float lit = sqrtf(dot(vit, vit) + 1e-8f);
                float c = tmax(1.f + S.gf * lin / lit, 0.f);
                v = mk(c * (vit.x + (float)I0 * 1e-30f), 0.f, c * (vit.z + (float)I2 * 1e-30f));
            } else {
                v = mk(0.f, 0.f, 0.f);
            }
        }
        if (I[d] > S.n_grid - 3 && comp(v, d) > 0.f) setc(v, d, 0.f);
    }
    return v;
}

// ------------------------------------------------------------------ kernels
struct ChunkArgs {
    const SceneC* scenes;
    const int4* chunks;   // (scene, first slot, count, -)
    int n_chunks;
    int stride;
};

// ------------------------------------------------------------------ binning
// Each substep the particles of a frame are counting-sorted by (4x4x4 cell
// block, cell): per-cell counts with wave-aggregated atomics, per-block scans,
// then an index scatter.  A block's particles form one or more work items of
// at most `wi` particles (a runtime launch parameter, a multiple of WG
// between WI_MIN and WI_MAX, default WI); the particle phases run one
// workgroup per work item
// over a 6x6x6-node LDS tile (the stencil footprint of a cell block).
constexpr int WI = 1024;     // default particles per work item
constexpr int WI_MIN = 256;  // the work-item capacity of a context is sized for this
constexpr int WI_MAX = 8192;
constexpr int TILE = 216;    // 6x6x6 nodes
constexpr int WG = 256;      // threads per work-item workgroup
constexpr int MAXSEG = 4;    // distinct cells per wave the mass grid reduces through the transpose tree

// per bin slot counters: CNT_X0 + x counts the work items of XCD x, CNT_XB0 + x
// is the first entry of its list in xlist
constexpr int NXCD = 8;        // MI355X: 8 XCDs, workgroup b dispatched to XCD b % 8
constexpr int DET_GMAX = 16;   // record-array entries per cell (fixed-order ranks)
constexpr int FB_LEAD = 64;    // leading workgroups that carry an overflow fallback
constexpr int RESORT_MIN = 8;  // shortest slot re-sort interval (substeps)
constexpr long long FORK_MIN_PARTICLES = 1 << 17;  // contexts this large fork the constitutive update
enum { CNT_NBLK = 1, CNT_PART = 2, CNT_ITEMS = 3, CNT_STAMP = 4, CNT_X0 = 8, CNT_XB0 = 16 };
// global binning state: the pass stamp base and the occupied-block count of
// the pass in flight (both advanced/reset by the pass's last kernel)
enum { G_STAMP = 0, G_PBLK = 1 };

struct BinArgs {
    const SceneC* scenes;
    const int4* chunks;   // (scene, first slot, count, -) per PPB particles
    const int* blk_scene;
    int stride;
    int* cell_count;
    int* cell_start;
    int* bin_q;
    int* bin_rank;
    int* blk_flag;
    int* node_flag;
    int* counters;   // this bin slot's counters
    int* gstamp;     // global binning state (G_STAMP, G_PBLK)
    int* pblocks;
    int* blk_stamp;
    int* blk_item0;
    int* blk_nitems;
    int4* items;
    int* order;
    int wi;          // particles per work item
    // fixed-order mode: each (cell, wave) group of a pass records its wave and
    // count, so a particle's rank is its rank within its wave plus the counts
    // of its cell's groups in lower waves -- slot order, not the order in which
    // the groups' atomics landed.  A group claims its index among its cell's
    // groups of the pass with an atomic on cell_gcnt (zero between passes:
    // k_blocks moves each occupied cell's total to cell_ng and clears it); the
    // first DET_GMAX groups take the entries of the cell's record array
    // (cell_grp: (wave << 7) | count), the rest are listed (cell_head: pass
    // stamp and newest group, grp_next / grp_cnt per group leader's slot)
    int det;
    int* cell_gcnt;
    int* cell_ng;
    unsigned* cell_grp;
    unsigned long long* cell_head;
    int* grp_next;
    int* grp_cnt;
    // XCD-aware work: blk_rep[block] = the slot of the particle whose group
    // first reached the block this pass; the block's work items are tagged
    // (item.w) with the XCD that owns that slot's eighth of the particle
    // columns, so items whose particles share cache lines of the SoA rows run
    // on one XCD; xlist lists the items grouped by tag, each entry the item
    // itself with .w its index (a tile kernel then reads one entry, not an
    // index and then the item)
    int* blk_rep;
    int4* xlist;
    int* blk_tot;    // particles of each occupied block this pass (slot re-sort on), or null
};

DEV unsigned long long lanemask_lt(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

DEV void block_origin(const SceneC& S, int gb, int& bi, int& bj, int& bk) {
    int lb = gb - S.blk_off;
    bk = lb % S.nb;
    int t = lb / S.nb;
    bj = t % S.nb;
    bi = t / S.nb;
}

// bin frame f's particles by cell; optionally run primitive forward kinematics
// f -> f+1 (forward_kinematics, primive_base.py:152-156).  A pass is stamped
// gstamp[G_STAMP] + 1; the first group to reach a cell block (exchange of the
// block's flag with the stamp) appends it to the occupied-block list.
__global__ void __launch_bounds__(PPB) k_bin(BinArgs a, const float* __restrict__ st, int do_fk, const PrimC* __restrict__ prims,
                                             int n_prims_total, float* __restrict__ pstate, const float* __restrict__ pvel, int f) {
    int stamp = a.gstamp[G_STAMP] + 1;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.counters[CNT_NBLK] = 0;
        a.counters[CNT_PART] = 0;
        a.counters[CNT_ITEMS] = 0;
        a.counters[CNT_STAMP] = stamp;

    }
    if (do_fk && blockIdx.x == 0)
        for (int q = threadIdx.x; q < n_prims_total; q += blockDim.x)
            prim_fk(prims[q], pstate + ((size_t)f * n_prims_total + q) * 8, pvel + ((size_t)f * n_prims_total + q) * 8,
                    pstate + ((size_t)(f + 1) * n_prims_total + q) * 8);
    int4 ch = a.chunks[blockIdx.x];
    int i = threadIdx.x;
    int lane = i & 63;
    bool valid = i < ch.z;
    int slot = ch.y + i;
    int q = -1;
    if (valid) {
        const SceneC& S = a.scenes[ch.x];
        v3 x = load3(st, a.stride, 0, slot);
        int b0 = min(max((int)(x.x * S.inv_dx - 0.5f), 0), S.n_grid - 1);
        int b1 = min(max((int)(x.y * S.inv_dx - 0.5f), 0), S.n_grid - 1);
        int b2 = min(max((int)(x.z * S.inv_dx - 0.5f), 0), S.n_grid - 1);
        int gb = S.blk_off + ((b0 >> 2) * S.nb + (b1 >> 2)) * S.nb + (b2 >> 2);
        q = gb * 64 + (((b0 & 3) << 4) | ((b1 & 3) << 2) | (b2 & 3));
    }
    // group the wave's lanes by cell (ballots only), then every group leader
    // claims its group's ranks with one atomic, all in a single instruction
    unsigned long long act = __ballot(valid);
    int rank = 0, cnt = 0, leader = lane;
    while (act) {
        int lead = __ffsll((long long)act) - 1;
        int q0 = __shfl(q, lead);
        unsigned long long m = __ballot(valid && q == q0);
        if (valid && q == q0) { rank = __popcll(m & lanemask_lt(lane)); leader = lead; }
        if (lane == lead) cnt = __popcll(m);
        act &= ~m;
    }
    int base = 0;
    bool fresh = false;
    if (valid && lane == leader) {
        if (a.det) {
            atomicAdd(&a.cell_count[q], cnt);
            int gi = atomicAdd(&a.cell_gcnt[q], 1);   // this group's index among the cell's groups of the pass
            if (gi < DET_GMAX) {
                a.cell_grp[(size_t)q * DET_GMAX + gi] = ((unsigned)(slot >> 6) << 7) | (unsigned)cnt;
            } else {
                unsigned long long mine = ((unsigned long long)(unsigned)stamp << 32) | (unsigned)slot;
                unsigned long long old = atomicExch(&a.cell_head[q], mine);
                a.grp_next[slot] = ((int)(old >> 32) == stamp) ? (int)(unsigned)(old & 0xffffffffull) : -1;
                a.grp_cnt[slot] = cnt;
            }
        } else {
            base = atomicAdd(&a.cell_count[q], cnt);
        }
        int gb = q >> 6;
        fresh = a.blk_flag[gb] != stamp && atomicExch(&a.blk_flag[gb], stamp) != stamp;
        if (fresh) a.blk_rep[gb] = slot;
    }
    rank += __shfl(base, leader);
    unsigned long long nm = __ballot(fresh);
    if (nm) {
        int lead = __ffsll((long long)nm) - 1;
        int b0 = 0;
        if (lane == lead) b0 = atomicAdd(&a.gstamp[G_PBLK], __popcll(nm));
        b0 = __shfl(b0, lead);
        if (fresh) a.pblocks[b0 + __popcll(nm & lanemask_lt(lane))] = q >> 6;
    }
    if (valid) {
        a.bin_q[slot] = q;
        a.bin_rank[slot] = rank;
    }
}

// per occupied cell block (one wave, lane = cell): scan the cell counts, claim
// a particle range and work items, stamp the block, and append the node blocks
// its 6x6x6 footprint covers to the active node-block list (first setter of a
// node block's flag, as in k_bin) with their list positions in nmap.  A
// workgroup takes 4 blocks at a time and claims their ranges with one 64-bit
// atomic on the (particles, items) counter pair and one on the node count.
// Node-block records: for active node block e < rec_cap, NREC ints
// {gb, scene, scene's first primitive, its primitive count, item0[8], nitems[8]}
// with the work-item ranges of the 8 cell
// blocks whose 6x6x6 footprint covers it (cell block B - (d>>2&1, d>>1&1, d&1);
// 0 items when it is not occupied this pass), so the grid phases read one record
// instead of chasing block list -> scene -> block stamps -> item ranges.
// k_blocks writes gb, scene and the 8 covering block ids (-1: outside the grid)
// into slots 4..11; k_scatter, once every block is stamped, replaces those with
// the item ranges.
constexpr int NREC = 20;
__global__ void __launch_bounds__(256) k_blocks(BinArgs a, int* __restrict__ nblocks, int* __restrict__ nmap, int* __restrict__ nrec,
                                                int rec_cap) {
    __shared__ unsigned long long s_tot[4];
    __shared__ unsigned long long s_base;
    __shared__ int s_nn[4];
    __shared__ int s_nbase;

    int lane = threadIdx.x & 63;
    int wave = threadIdx.x >> 6;
    int n = a.gstamp[G_PBLK];
    int stamp = a.counters[CNT_STAMP];
    unsigned long long* pair = reinterpret_cast<unsigned long long*>(a.counters + CNT_PART);  // low: particles, high: items
    for (int e0 = blockIdx.x * 4; e0 < n; e0 += gridDim.x * 4) {
        int e = e0 + wave;
        bool has = e < n;
        int gb = has ? a.pblocks[e] : 0;
        int cnt = has ? a.cell_count[gb * 64 + lane] : 0;
        if (has) a.cell_count[gb * 64 + lane] = 0;
        if (has && a.det) {
            a.cell_ng[gb * 64 + lane] = a.cell_gcnt[gb * 64 + lane];
            a.cell_gcnt[gb * 64 + lane] = 0;
        }
        int incl = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            int t = __shfl_up(incl, d);
            if (lane >= d) incl += t;
        }
        int total = __shfl(incl, 63);
        int nit = (total + a.wi - 1) / a.wi;
        int xcd = has ? min(NXCD - 1, (int)(((long long)a.blk_rep[gb] * NXCD) / a.stride)) : 0;
        // node blocks of the footprint reached first by this pass
        int nb = -1, sc = 0, ni = 0, nj = 0, nk = 0;
        if (has && lane < 8) {
            sc = a.blk_scene[gb];
            const SceneC& S = a.scenes[sc];
            int bi, bj, bk;
            block_origin(S, gb, bi, bj, bk);
            ni = bi + ((lane >> 2) & 1), nj = bj + ((lane >> 1) & 1), nk = bk + (lane & 1);
            if (ni < S.nb && nj < S.nb && nk < S.nb) nb = S.blk_off + (ni * S.nb + nj) * S.nb + nk;
        }
        bool fresh = nb >= 0 && a.node_flag[nb] != stamp && atomicExch(&a.node_flag[nb], stamp) != stamp;
        unsigned long long fm = __ballot(fresh);
        if (lane == 0) {
            s_tot[wave] = ((unsigned long long)nit << 32) | (unsigned)total;
            s_nn[wave] = __popcll(fm);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            s_base = atomicAdd(pair, s_tot[0] + s_tot[1] + s_tot[2] + s_tot[3]);
            s_nbase = atomicAdd(&a.counters[CNT_NBLK], s_nn[0] + s_nn[1] + s_nn[2] + s_nn[3]);
        }
        __syncthreads();
        unsigned long long off = s_base;
        int noff = s_nbase;
        for (int w = 0; w < wave; w++) { off += s_tot[w]; noff += s_nn[w]; }
        int start = (int)(off & 0xffffffffull), ib = (int)(off >> 32);
        __syncthreads();
        if (fresh) {
            int pos = noff + __popcll(fm & lanemask_lt(lane));
            nblocks[pos] = nb;
            nmap[nb] = pos;
            if (pos < rec_cap) {
                const SceneC& S = a.scenes[sc];
                int* r = nrec + (size_t)pos * NREC;
                r[0] = nb;
                r[1] = sc;
                r[2] = S.prim_off;
                r[3] = S.n_prims;
#pragma unroll
                for (int d = 0; d < 8; d++) {
                    int P0 = ni - ((d >> 2) & 1), P1 = nj - ((d >> 1) & 1), P2 = nk - (d & 1);
                    r[4 + d] = (P0 < 0 || P1 < 0 || P2 < 0) ? -1 : S.blk_off + (P0 * S.nb + P1) * S.nb + P2;
                }
            }
        }
        if (!has) continue;
        a.cell_start[gb * 64 + lane] = start + incl - cnt;
        for (int it = lane; it < nit; it += 64) a.items[ib + it] = make_int4(gb, start + it * a.wi, min(a.wi, total - it * a.wi), xcd);
        if (lane == 0) {
            a.blk_stamp[gb] = stamp;
            a.blk_item0[gb] = ib;
            a.blk_nitems[gb] = nit;
            if (a.blk_tot) a.blk_tot[gb] = total;
        }
    }
}

// index scatter; the pass's last kernel also advances the stamp base and
// resets the occupied-block count for the next pass
__global__ void __launch_bounds__(PPB) k_scatter(BinArgs a, int* __restrict__ nrec, int rec_cap) {
    __shared__ int s_cnt[PPB][NXCD];
    __shared__ int s_wtot[PPB / 64][NXCD];
    __shared__ int s_base[NXCD];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.gstamp[G_STAMP] += 1;
        a.gstamp[G_PBLK] = 0;
    }
    if (blockIdx.x == 0) {
        // the work items grouped by XCD tag, each group in item order: thread t
        // counts a contiguous range of items per tag, the counts are scanned
        // over the threads (wave scans, then the waves' totals), then every
        // item is placed
        int t = threadIdx.x, lane = t & 63, wave = t >> 6;
        int n = a.counters[CNT_ITEMS];
        int per = (n + PPB - 1) / PPB;
        int e0 = min(n, t * per), e1 = min(n, e0 + per);
#pragma unroll
        for (int x = 0; x < NXCD; x++) s_cnt[t][x] = 0;
        for (int e = e0; e < e1; e++) s_cnt[t][a.items[e].w] += 1;
        int c[NXCD], incl[NXCD];
#pragma unroll
        for (int x = 0; x < NXCD; x++) { c[x] = s_cnt[t][x]; incl[x] = c[x]; }
#pragma unroll
        for (int d = 1; d < 64; d <<= 1)
#pragma unroll
            for (int x = 0; x < NXCD; x++) {
                int u = __shfl_up(incl[x], d);
                if (lane >= d) incl[x] += u;
            }
        if (lane == 63)
#pragma unroll
            for (int x = 0; x < NXCD; x++) s_wtot[wave][x] = incl[x];
        __syncthreads();
        if (t == 0) {
            int acc = 0;
            for (int x = 0; x < NXCD; x++) {
                int tot = 0;
                for (int w = 0; w < PPB / 64; w++) tot += s_wtot[w][x];
                a.counters[CNT_X0 + x] = tot;
                a.counters[CNT_XB0 + x] = acc;
                s_base[x] = acc;
                acc += tot;
            }
        }
        __syncthreads();
        int off[NXCD];
#pragma unroll
        for (int x = 0; x < NXCD; x++) {
            int b = s_base[x] + incl[x] - c[x];
            for (int w = 0; w < wave; w++) b += s_wtot[w][x];
            off[x] = b;
        }
#pragma unroll
        for (int x = 0; x < NXCD; x++) s_cnt[t][x] = off[x];
        for (int e = e0; e < e1; e++) {
            int x = a.items[e].w;
            int4 it = a.items[e];
            a.xlist[s_cnt[t][x]] = make_int4(it.x, it.y, it.z, e);
            s_cnt[t][x] += 1;
        }
    }
    // node-block records: covering block ids -> their item ranges this pass
    int n_rec = min(a.counters[CNT_NBLK], rec_cap);
    int stamp = a.counters[CNT_STAMP];
    for (int t = blockIdx.x * PPB + threadIdx.x; t < n_rec; t += gridDim.x * PPB) {
        int* r = nrec + (size_t)t * NREC;
        int pgb[8], bst[8], i0[8], nit[8];
#pragma unroll
        for (int d = 0; d < 8; d++) pgb[d] = r[4 + d];
#pragma unroll
        for (int d = 0; d < 8; d++) {
            int q = max(pgb[d], 0);
            bst[d] = a.blk_stamp[q];
            i0[d] = a.blk_item0[q];
            nit[d] = a.blk_nitems[q];
        }
#pragma unroll
        for (int d = 0; d < 8; d++) {
            bool hit = pgb[d] >= 0 && bst[d] == stamp;
            r[4 + d] = hit ? i0[d] : 0;
            r[12 + d] = hit ? nit[d] : 0;
        }
    }
    int4 ch = a.chunks[blockIdx.x];
    int i = threadIdx.x;
    if (i >= ch.z) return;
    int slot = ch.y + i;
    int q = a.bin_q[slot];
    int r = a.bin_rank[slot];
    if (a.det) {
        // the cell's groups of this pass in lower waves (slots below this
        // wave's) precede this particle: the record array, then the list of
        // the groups past DET_GMAX, newest first
        int ng = a.cell_ng[q];
        unsigned w = (unsigned)(slot >> 6);
        const uint4* ra = reinterpret_cast<const uint4*>(a.cell_grp + (size_t)q * DET_GMAX);
#pragma unroll
        for (int k = 0; k < DET_GMAX / 4; k++) {
            if (4 * k < ng) {
                uint4 v = ra[k];
                unsigned e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (4 * k + j < ng && (e[j] >> 7) < w) r += (int)(e[j] & 127u);
            }
        }
        if (ng > DET_GMAX) {
            int rec = (int)(unsigned)(a.cell_head[q] & 0xffffffffull);
            while (rec >= 0) {
                if ((unsigned)(rec >> 6) < w) r += a.grp_cnt[rec];
                rec = a.grp_next[rec];
            }
        }
    }
    a.order[a.cell_start[q] + r] = slot;
}

// ------------------------------------------------------------------ wave transpose reduction
// Sums 32 per-lane values over the 64 lanes of a wave in 6 butterfly steps,
// entirely in VALU (no LDS): bit 5 and bit 4 with gfx950's v_permlane32_swap /
// v_permlane16_swap (one swap + one add per value), bits 3..0 with DPP
// row_mirror, row_half_mirror and quad permutations.  At each step a lane keeps
// the half of its values selected by its lane bit and adds the partner's copy
// of that half; afterwards lanes 2j and 2j+1 hold the total of value j.
DEV float f_of(unsigned u) { return __uint_as_float(u); }
DEV unsigned u_of(float f) { return __float_as_uint(f); }
// a lane's value read from the lane that the DPP control selects (every source
// lane valid for the controls used here); written as update_dpp with a zero
// old value and bound_ctrl so that the compiler folds it into the consuming
// add (v_add_f32_dpp) instead of issuing a separate move
template <int DPP>
DEV float dpp_get(unsigned u) { return f_of(__builtin_amdgcn_update_dpp(0u, u, DPP, 0xf, 0xf, true)); }
template <int H>
DEV void tstep_swap32(float* r) {
#pragma unroll
    for (int j = 0; j < H; j++) {
        auto sw = __builtin_amdgcn_permlane32_swap(u_of(r[j]), u_of(r[j + H]), false, false);
        r[j] = f_of(sw[0]) + f_of(sw[1]);
    }
}
template <int H>
DEV void tstep_swap16(float* r) {
#pragma unroll
    for (int j = 0; j < H; j++) {
        auto sw = __builtin_amdgcn_permlane16_swap(u_of(r[j]), u_of(r[j + H]), false, false);
        r[j] = f_of(sw[0]) + f_of(sw[1]);
    }
}
template <int H, int MASK, int DPP>
DEV void tstep_dpp(float* r, int lane) {
    bool up = (lane & MASK) != 0;
#pragma unroll
    for (int j = 0; j < H; j++) {
        float lo = r[j], hi = r[j + H];
        float send = up ? lo : hi;
        float keep = up ? hi : lo;
        r[j] = keep + dpp_get<DPP>(u_of(send));
    }
}
// The mirrors pair lane i with i ^ 15 (resp. i ^ 7): the partner has the opposite
// split bit and the same higher bits, so it holds the same value set and the
// lane sets summed at each step stay disjoint (their xor span doubles).
constexpr int DPP_ROW_MIRROR = 0x140;       // lane i <- lane 15 - i within a row of 16
constexpr int DPP_ROW_HALF_MIRROR = 0x141;  // lane i <- lane 7 - i within a half row
constexpr int DPP_QUAD_XOR2 = 0x4E;         // quad_perm [2,3,0,1]
constexpr int DPP_QUAD_XOR1 = 0xB1;         // quad_perm [1,0,3,2]
DEV float tsum32(float* r, int lane) {
    tstep_swap32<16>(r);
    tstep_swap16<8>(r);
    tstep_dpp<4, 8, DPP_ROW_MIRROR>(r, lane);
    tstep_dpp<2, 4, DPP_ROW_HALF_MIRROR>(r, lane);
    tstep_dpp<1, 2, DPP_QUAD_XOR2>(r, lane);
    return r[0] + dpp_get<DPP_QUAD_XOR1>(u_of(r[0]));
}
// 16 values: four halving steps (bits 5..2), then the quad sum; afterwards
// every lane of quad j holds the total of value j
DEV float tsum16(float* r, int lane) {
    tstep_swap32<8>(r);
    tstep_swap16<4>(r);
    tstep_dpp<2, 8, DPP_ROW_MIRROR>(r, lane);
    tstep_dpp<1, 4, DPP_ROW_HALF_MIRROR>(r, lane);
    float s = r[0] + dpp_get<DPP_QUAD_XOR2>(u_of(r[0]));
    return s + dpp_get<DPP_QUAD_XOR1>(u_of(s));
}
// 4 values: two halving steps (bits 5, 4), then the sum over each 16-lane
// group; afterwards every lane of group g = lane >> 4 holds the total of value g
DEV float tsum4(float* r) {
    tstep_swap32<2>(r);
    tstep_swap16<1>(r);
    float s = r[0] + dpp_get<DPP_ROW_MIRROR>(u_of(r[0]));
    s += dpp_get<DPP_ROW_HALF_MIRROR>(u_of(s));
    s += dpp_get<DPP_QUAD_XOR2>(u_of(s));
    return s + dpp_get<DPP_QUAD_XOR1>(u_of(s));
}
DEV int tile_index(int l0, int l1, int l2, int n) { return ((l0 + n / 9) * 6 + (l1 + (n / 3) % 3)) * 6 + (l2 + n % 3); }
// node v = (a * 3 + b) * 4 + z of the 3x3x4 nodes two z-adjacent cells cover
DEV int tile_index4(int l0, int l1, int l2, int v) { return ((l0 + v / 12) * 6 + (l1 + (v / 4) % 3)) * 6 + (l2 + v % 4); }

// Segmented scatter of 27 x NCOMP per-particle node values into the calling
// wave's own LDS tile (component-major, TILE floats per component): the lanes
// of each distinct cell of the wave, in order of their first lane, are summed
// through the transpose tree and the 27 node sums added into the tile.  Every
// cell takes the tree, however many a wave holds (adding the lanes' own values
// with LDS atomics past a few cells measured slower, and its order within an
// instruction is not fixed).  When the rest of the wave is one more cell, the
// z-neighbour of the first (the common case in cell order: a wave of 64
// sorted particles spans one or two cells), both go through one tree pass
// over the 3x3x4 nodes they cover together (36 values: the 32-value tree and
// a 4-value one) instead of two 27-value passes.  CHAN(c, r) writes the
// lane's 27 node values of component c into r[0..26] (node n = (a * 3 + b) *
// 3 + cc); one transpose pass per component, so a caller forms a pass's 27
// values incrementally along the stencil.
template <int NCOMP, class ChanFn>
DEV void scatter_tile_ch(float* tile, bool valid, int key, int l0, int l1, int l2, int lane, ChanFn chan) {
    unsigned long long act = __ballot(valid);
    while (act) {
        int lead = __ffsll((long long)act) - 1;
        int key0 = __shfl(key, lead);
        unsigned long long m = __ballot(valid && key == key0);
        bool in = (m >> lane) & 1ull;
        int s0 = __shfl(l0, lead), s1 = __shfl(l1, lead), s2 = __shfl(l2, lead);
        unsigned long long mz = (key0 & 3) != 3 ? __ballot(valid && key == key0 + 1) : 0ull;
        if (mz != 0ull && mz == (act & ~m)) {
            bool inz = (mz >> lane) & 1ull;
#pragma unroll
            for (int c = 0; c < NCOMP; c++) {
                float r[32], u[32], u4[4];
                chan(c, r);
#pragma unroll
                for (int v = 0; v < 36; v++) {
                    int ab = v / 4, z = v % 4;
                    float fa = z < 3 ? r[ab * 3 + z] : 0.f;
                    float fb = z > 0 ? r[ab * 3 + z - 1] : 0.f;
                    float val = in ? fa : (inz ? fb : 0.f);
                    if (v < 32) u[v] = val;
                    else u4[v - 32] = val;
                }
                float s = tsum32(u, lane);
                float t = tsum4(u4);
                if (!(lane & 1)) atomicAdd(&tile[c * TILE + tile_index4(s0, s1, s2, lane >> 1)], s);
                if (!(lane & 15)) atomicAdd(&tile[c * TILE + tile_index4(s0, s1, s2, 32 + (lane >> 4))], t);
            }
            break;
        }
#pragma unroll
        for (int c = 0; c < NCOMP; c++) {
            float r[32];
            chan(c, r);
#pragma unroll
            for (int j = 0; j < 32; j++) r[j] = (in && j < 27) ? r[j] : 0.f;
            float s = tsum32(r, lane);
            int n = lane >> 1;
            if (!(lane & 1) && n < 27) atomicAdd(&tile[c * TILE + tile_index(s0, s1, s2, n)], s);
        }
        act &= ~m;
    }
}

struct TileArgs {
    const SceneC* scenes;
    const int* blk_scene;
    const int4* items;
    const int* counters;
    const int* order;
    int stride;
    // grid kept from the forward pass: node block e of the active list at
    // keep[e * 64 + node], nmap[block] = e; valid while the list fits keep_cap
    const float4* keep_v;
    const int* nmap;
    int keep_cap;
    const int4* xlist;     // work items grouped by XCD, .w the item index (see BinArgs)
};

// The work items a workgroup of a tile kernel takes: those of its own XCD's
// list (workgroup b runs on XCD b % NXCD), entry j going to that XCD's
// workgroup j mod (its workgroup count); with fewer workgroups than XCDs,
// every item in turn.  The item workgroups are the launch's workgroups
// [lead, lead + nwg) (lead a multiple of NXCD; nwg <= 0: the rest of the
// launch); the others do other work.
struct ItemIter { const int4* list; int j, step, n; };
DEV ItemIter item_iter(const TileArgs& ta, int lead = 0, int nwg = 0) {
    ItemIter r;
    int b = (int)blockIdx.x - lead, nb = nwg > 0 ? nwg : (int)gridDim.x - lead;
    if (nb >= NXCD) {
        int x = b % NXCD;
        r.list = ta.xlist + ta.counters[CNT_XB0 + x];
        r.n = ta.counters[CNT_X0 + x];
        r.j = b / NXCD;
        r.step = (nb - x + NXCD - 1) / NXCD;
    } else {
        r.list = nullptr;
        r.n = ta.counters[CNT_ITEMS];
        r.j = b;
        r.step = nb;
    }
    return r;
}
// entry j of a workgroup's items: the item and its index e
DEV int4 item_at(const TileArgs& ta, const ItemIter& ir, int j, int& e) {
    if (ir.list) {
        int4 v = ir.list[j];
        e = v.w;
        return v;
    }
    e = j;
    return ta.items[j];
}

struct GridArgs {
    const SceneC* scenes;
    const PrimC* prims;
    const int* blk_scene;
    const int* nblocks;
    const int* counters;
    const int* blk_stamp;
    const int* blk_item0;
    const int* blk_nitems;
    const int* nrec;       // node-block records (NREC ints each) for e < rec_cap
    int rec_cap;
    const float* pstate;   // [frame][n_prims_total][8]
    int n_prims_total;
    int f;
    float4* keep_pm;       // kept grid (see TileArgs)
    float4* keep_v;
    int keep_cap;
    int keep_write;        // grid_op stores its result into the kept grid
    // fixed-order mode: the grid adjoint writes each node block's primitive
    // adjoints (per collision unit, 16 values: frames f and f+1) to
    // padj[block] and flags the block in its scene's bitmap; a reduction
    // sums the flagged blocks in block order
    int det;
    float* padj;
    unsigned* pbits;
};

// true when the active node-block list of this pass fits the kept grid
DEV bool kept_fits(const int* counters, int keep_cap) { return counters[CNT_NBLK] <= keep_cap; }

DEV void node_coords(const SceneC& S, int gb, int lane, int& B0, int& B1, int& B2, int& I0, int& I1, int& I2, long long& node) {
    block_origin(S, gb, B0, B1, B2);
    I0 = B0 * 4 + (lane >> 4);
    I1 = B1 * 4 + ((lane >> 2) & 3);
    I2 = B2 * 4 + (lane & 3);
    node = S.g_off + (long long)(gb - S.blk_off) * NODES_PER_BLK + lane;
}

DEV int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The active node block e of this pass: its global block id, scene and the
// work-item ranges (i0, ni) of the 8 covering cell blocks -- from its record
// when e < rec_cap, else from the block list and the block stamps.  All of it
// is wave-uniform (scalar loads).
struct BlockMeta { int gb, sc, poff, np, i0[8], ni[8]; };
DEV void read_rec(const GridArgs& ga, int e, BlockMeta& m) {
    const int* r = ga.nrec + (size_t)e * NREC;
    m.gb = uniform(r[0]);
    m.sc = uniform(r[1]);
    m.poff = uniform(r[2]);
    m.np = uniform(r[3]);
#pragma unroll
    for (int d = 0; d < 8; d++) { m.i0[d] = uniform(r[4 + d]); m.ni[d] = uniform(r[12 + d]); }
}
DEV void block_meta(const GridArgs& ga, int e, int stamp, BlockMeta& m) {
    if (e < ga.rec_cap) {
        read_rec(ga, e, m);
        return;
    }
    m.gb = uniform(ga.nblocks[e]);
    m.sc = uniform(ga.blk_scene[m.gb]);
    const SceneC& S = ga.scenes[m.sc];
    m.poff = uniform(S.prim_off);
    m.np = uniform(S.n_prims);
    int B0, B1, B2;
    block_origin(S, m.gb, B0, B1, B2);
    int pgb[8], st[8];
#pragma unroll
    for (int d = 0; d < 8; d++) {
        int P0 = B0 - ((d >> 2) & 1), P1 = B1 - ((d >> 1) & 1), P2 = B2 - (d & 1);
        pgb[d] = (P0 < 0 || P1 < 0 || P2 < 0) ? -1 : S.blk_off + (P0 * S.nb + P1) * S.nb + P2;
    }
#pragma unroll
    for (int d = 0; d < 8; d++) st[d] = pgb[d] >= 0 ? ga.blk_stamp[pgb[d]] : stamp - 1;
#pragma unroll
    for (int d = 0; d < 8; d++) {
        bool hit = st[d] == stamp;
        m.i0[d] = uniform(hit ? ga.blk_item0[pgb[d]] : 0);
        m.ni[d] = uniform(hit ? ga.blk_nitems[pgb[d]] : 0);
    }
}

// A scene's constants, its primitives and their states at frames f and f+1,
// staged in LDS by one wave for the node blocks of that scene: the node update
// reads them many times per node and under divergent branches, where they
// would otherwise be vector loads from memory, each a full memory latency on
// the wave's critical path.
constexpr int SCENE_DW = sizeof(SceneC) / 4, PRIM_DW = sizeof(PrimC) / 4;
static_assert(SCENE_DW * 4 == sizeof(SceneC) && SCENE_DW <= 64 && PRIM_DW * 4 == sizeof(PrimC), "staged as whole dwords");
struct SceneStage {
    SceneC sc;
    PrimC pr[MAX_UNITS];
    float ps[2][MAX_UNITS * 8];
};
// (the block's record carries the scene's primitive range, so every load here
// is issued at once)
DEV void stage_scene(SceneStage& st, const GridArgs& ga, const BlockMeta& m, int lane) {
    // the previous block's reads of the stage are issued before it is rewritten
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const int* src = reinterpret_cast<const int*>(ga.scenes + m.sc);
    int v = lane < SCENE_DW ? src[lane] : 0;
    int poff = m.poff, np = m.np;
    if (lane < SCENE_DW) reinterpret_cast<int*>(&st.sc)[lane] = v;
    const int* ps = reinterpret_cast<const int*>(ga.prims + poff);
    int* dp = reinterpret_cast<int*>(st.pr);
    for (int k = lane; k < np * PRIM_DW; k += 64) dp[k] = ps[k];
    const float* s0 = ga.pstate + ((size_t)ga.f * ga.n_prims_total + poff) * 8;
    const float* s1 = s0 + (size_t)ga.n_prims_total * 8;
    for (int k = lane; k < np * 8; k += 64) {
        st.ps[0][k] = s0[k];
        st.ps[1][k] = s1[k];
    }
    // LDS operations of one wave complete in order: the stage is visible to
    // every lane of the wave once the compiler keeps the reads after the writes
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// owner-computes gather: the node's value is the sum, in a fixed order, of the
// tiles of the work items of the (up to 8) cell blocks whose footprint covers
// it, loaded BATCH at a time across blocks; [kb, ke) restricts the sum to
// those pairs (ke < 0: to the last)
template <int BATCH>
DEV float4 gather_tiles(const BlockMeta& m, int lane, const float4* __restrict__ tiles, int kb = 0, int ke = -1) {
    int l0 = lane >> 4, l1 = (lane >> 2) & 3, l2 = lane & 3;
    const int* i0 = m.i0;
    const int* ni = m.ni;
    // the (block d, item j) pairs in (d, j) order, flattened: pair k belongs to
    // the block d with pre[d] <= k < pre[d + 1].  BATCH tiles are loaded before
    // they are summed, in pair order, so at most a few load latencies chain.
    int pre[8];
    pre[0] = 0;
#pragma unroll
    for (int d = 1; d < 8; d++) pre[d] = pre[d - 1] + ni[d - 1];
    int total = ke < 0 ? pre[7] + ni[7] : ke;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int k0 = kb; k0 < total; k0 += BATCH) {
        float4 u[BATCH];
#pragma unroll
        for (int s = 0; s < BATCH; s++) {
            int k = k0 + s;
            int d = 0;
#pragma unroll
            for (int q = 1; q < 8; q++) d += (k >= pre[q]) ? 1 : 0;
            int item = i0[0] + k;
#pragma unroll
            for (int q = 1; q < 8; q++)
                if (d == q) item = i0[q] + (k - pre[q]);
            d = uniform(d);
            item = uniform(item);
            int t0 = l0 + 4 * ((d >> 2) & 1), t1 = l1 + 4 * ((d >> 1) & 1), t2 = l2 + 4 * (d & 1);
            bool in = k < total && t0 < 6 && t1 < 6 && t2 < 6;
            u[s] = in ? tiles[(size_t)item * TILE + (t0 * 6 + t1) * 6 + t2] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int s = 0; s < BATCH; s++) { acc.x += u[s].x; acc.y += u[s].y; acc.z += u[s].z; acc.w += u[s].w; }
    }
    return acc;
}

// grid_op over the active node blocks (mpm_simulator.py:230-262), one block
// per workgroup.  Writes the gathered momentum/mass (kept for the adjoint) and
// the velocity.  The covering tiles are summed in two halves, pairs [0, h) and
// [h, total) with h = (total + 1) / 2, each in pair order, then added: with
// WPG = 2 a second wave sums the second half at the same time and hands it
// over through LDS (part), halving the gather of the densest blocks (40+
// tiles), the operator's longest waves; with WPG = 1 the one wave sums both
// halves in turn -- the same bits either way.
// With only_overflow set it runs only when the kept grid overflowed (the
// adjoint then recomputes the grid instead of reading the kept one).
template <int WPG>
DEV void grid_op_body(const GridArgs& ga, const float4* __restrict__ tiles, float4* __restrict__ g_pm, float4* __restrict__ g_v,
                      SceneStage& stg, float4* part, int only_overflow, int lead = 0, int nwg = 0) {
    static_assert(WPG == 1 || WPG == 2, "one or two waves per block");
    int lane = threadIdx.x & 63;
    int wave = WPG == 2 ? (int)(threadIdx.x >> 6) : 0;
    int nb = nwg > 0 ? nwg : (int)gridDim.x - lead;
    int e = uniform((int)blockIdx.x - lead);
    // the first block's record is read together with the counters (a record
    // beyond the active count is stale and left unused; the read is
    // unconditional, clamped to the last record, so that both issue at once)
    BlockMeta m;
    bool have = e < ga.rec_cap;
    read_rec(ga, min(e, ga.rec_cap - 1), m);
    if (only_overflow && kept_fits(ga.counters, ga.keep_cap)) return;
    int n_active = ga.counters[CNT_NBLK];
    int stamp = ga.counters[CNT_STAMP];
    for (; e < n_active; e += nb) {
        if (!have) block_meta(ga, e, stamp, m);
        have = false;
        int total = 0;
#pragma unroll
        for (int d = 0; d < 8; d++) total += m.ni[d];
        int h = (total + 1) / 2;
        if (WPG == 2 && wave == 1) {
            part[lane] = gather_tiles<8>(m, lane, tiles, h, total);
            __syncthreads();   // the second half is in LDS
            __syncthreads();   // and read before the next block's
            continue;
        }
        stage_scene(stg, ga, m, lane);
        const SceneC& S = stg.sc;
        int B0, B1, B2, I0, I1, I2;
        long long node;
        node_coords(S, m.gb, lane, B0, B1, B2, I0, I1, I2, node);
        float4 pm = gather_tiles<8>(m, lane, tiles, 0, h);
        float4 b;
        if (WPG == 2) {
            __syncthreads();
            b = part[lane];
            __syncthreads();
        } else {
            b = gather_tiles<8>(m, lane, tiles, h, total);
        }
        pm.x += b.x; pm.y += b.y; pm.z += b.z; pm.w += b.w;
        v3 v = mk(0.f, 0.f, 0.f);
        if (pm.w > 1e-12f && in_grid(S, I0, I1, I2)) v = grid_node(S, stg.pr, stg.ps[0], stg.ps[1], I0, I1, I2, pm, nullptr);
        float4 v4 = make_float4(v.x, v.y, v.z, 0.f);
        g_pm[node] = pm;
        g_v[node] = v4;
        if (ga.keep_write && e < ga.keep_cap) {
            ga.keep_pm[(size_t)e * NODES_PER_BLK + lane] = pm;
            ga.keep_v[(size_t)e * NODES_PER_BLK + lane] = v4;
        }
    }
}
// the forward substep's grid operator (two waves per node block)
__global__ void __launch_bounds__(128) k_grid_op(GridArgs ga, const float4* __restrict__ tiles, float4* __restrict__ g_pm,
                                                 float4* __restrict__ g_v) {
    __shared__ SceneStage stage;
    __shared__ float4 part[64];
    grid_op_body<2>(ga, tiles, g_pm, g_v, stage, part, 0);
}
// the same operator rebuilding the grid of a substep for its adjoint (the
// grid was not kept, or, with only_overflow, its active list outgrew the
// kept grid): a kernel of its own so that profiles separate the two
__global__ void __launch_bounds__(128) k_grid_op_rebuild(GridArgs ga, const float4* __restrict__ tiles, float4* __restrict__ g_pm,
                                                         float4* __restrict__ g_v, int only_overflow) {
    __shared__ SceneStage stage;
    __shared__ float4 part[64];
    grid_op_body<2>(ga, tiles, g_pm, g_v, stage, part, only_overflow);
}

// load the 6x6x6 node tile of a cell block from a node grid into LDS
DEV void load_tile(float4* tv, const SceneC& S, int bi, int bj, int bk, const float4* __restrict__ g) {
    for (int k = threadIdx.x; k < TILE; k += WG) {
        int t0 = k / 36, t1 = (k / 6) % 6, t2 = k % 6;
        int I0 = bi * 4 + t0, I1 = bj * 4 + t1, I2 = bk * 4 + t2;
        tv[k] = in_grid(S, I0, I1, I2) ? g[node_id(S, I0, I1, I2)] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// the same tile from the kept (compact) grid
DEV void load_tile_kept(float4* tv, const SceneC& S, int bi, int bj, int bk, const float4* __restrict__ keep,
                        const int* __restrict__ nmap, int keep_cap) {
    for (int k = threadIdx.x; k < TILE; k += WG) {
        int t0 = k / 36, t1 = (k / 6) % 6, t2 = k % 6;
        int I0 = bi * 4 + t0, I1 = bj * 4 + t1, I2 = bk * 4 + t2;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (in_grid(S, I0, I1, I2)) {
            int gb = S.blk_off + ((I0 >> 2) * S.nb + (I1 >> 2)) * S.nb + (I2 >> 2);
            int e = nmap[gb];
            if ((unsigned)e < (unsigned)keep_cap)
                v = keep[(size_t)e * NODES_PER_BLK + (((I0 & 3) << 4) | ((I1 & 3) << 2) | (I2 & 3))];
        }
        tv[k] = v;
    }
}

// g2p over work items: 27-node gather from the LDS tile, APIC C, advection
// with clamp (mpm_simulator.py:264-283).  With perm given (a substep that ends
// at a re-sort boundary) particle p's outputs go to slot perm[p] of frame f+1,
// and its F of frame f+1, which the constitutive kernel left at slot p of
// fstage, moves there too.
__global__ void __launch_bounds__(WG) k_g2p_tile(TileArgs ta, const float* __restrict__ st, float* __restrict__ st_out,
                                                 const float4* __restrict__ g_v, const int* __restrict__ perm,
                                                 const float* __restrict__ fstage) {
    __shared__ float4 tv[TILE];
    ItemIter ir = item_iter(ta);
    int stride = ta.stride;
    for (int j = ir.j; j < ir.n; j += ir.step) {
        int e;
        int4 it = item_at(ta, ir, j, e);
        // the first batch's positions do not depend on the tile: their loads
        // are issued first, so both latencies overlap; every later batch's
        // are issued one batch ahead
        int k = threadIdx.x;
        int p = ta.order[it.y + min(k, it.z - 1)];
        v3 x = load3(st, stride, 0, p);
        const SceneC& S = ta.scenes[ta.blk_scene[it.x]];
        int bi, bj, bk;
        block_origin(S, it.x, bi, bj, bk);
        load_tile(tv, S, bi, bj, bk, g_v);
        __syncthreads();
        for (; k < it.z; k += WG) {
            int pn = p;
            v3 xn = x;
            if (k + WG < it.z) {
                pn = ta.order[it.y + k + WG];
                xn = load3(st, stride, 0, pn);
            }
            Stencil sw;
            stencil(S, x, sw);
            int l0 = sw.base[0] - 4 * bi, l1 = sw.base[1] - 4 * bj, l2 = sw.base[2] - 4 * bk;
            bool in_tile = (unsigned)l0 < 4u && (unsigned)l1 < 4u && (unsigned)l2 < 4u;
            v3 nv = mk(0.f, 0.f, 0.f);
            m3 nC = mzero();
            if (in_tile) {
                // new_v = sum_n wt g and new_C[:, j] = c4 sum_n wt (n_j - fx_j) g:
                // the weights factor over the axes, so the sums are taken axis by
                // axis (pw = w (n - fx) along one axis, w along the others).
                float pw[3][3];
#pragma unroll
                for (int o = 0; o < 3; o++)
#pragma unroll
                    for (int d = 0; d < 3; d++) pw[o][d] = sw.w[o][d] * ((float)o - sw.fx[d]);
                v3 G0 = mk(0.f, 0.f, 0.f), Gx = G0, Gy = G0, Gz = G0;
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    v3 J = mk(0.f, 0.f, 0.f), Jy = J, Jz = J;
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        v3 H = mk(0.f, 0.f, 0.f), Hz = H;
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            float4 g4 = tv[((l0 + a) * 6 + (l1 + b)) * 6 + (l2 + c)];
                            v3 g = mk(g4.x, g4.y, g4.z);
                            H += sw.w[c][2] * g;
                            Hz += pw[c][2] * g;
                        }
                        J += sw.w[b][1] * H;
                        Jy += pw[b][1] * H;
                        Jz += sw.w[b][1] * Hz;
                    }
                    G0 += sw.w[a][0] * J;
                    Gx += pw[a][0] * J;
                    Gy += sw.w[a][0] * Jy;
                    Gz += sw.w[a][0] * Jz;
                }
                float c4 = S.c4;
                nv = G0;
                nC.a[0] = c4 * Gx.x, nC.a[1] = c4 * Gy.x, nC.a[2] = c4 * Gz.x;
                nC.a[3] = c4 * Gx.y, nC.a[4] = c4 * Gy.y, nC.a[5] = c4 * Gz.y;
                nC.a[6] = c4 * Gx.z, nC.a[7] = c4 * Gy.z, nC.a[8] = c4 * Gz.z;
            }
            int o = p;
            if (perm) {
                o = perm[p];
                store9(st_out, stride, 15, o, load9(fstage, stride, 15, p));
            }
            store3(st_out, stride, 3, o, nv);
            store9(st_out, stride, 6, o, nC);
            v3 xo = mk(tmax(tmin(x.x + S.dt * nv.x, S.upper), S.lower), tmax(tmin(x.y + S.dt * nv.y, S.upper), S.lower),
                       tmax(tmin(x.z + S.dt * nv.z, S.upper), S.lower));
            store3(st_out, stride, 0, o, xo);
            p = pn;
            x = xn;
        }
        __syncthreads();
    }
}

// The kept-grid overflow fallback of the next adjoint substep (the grid of
// substep f rebuilt when its active list outgrew the kept grid), run by the
// leading workgroups of two launches of substep f+1's adjoint instead of two
// launches of its own that exit at once on every substep whose grid fits:
// the p2g rebuild beside the g2p adjoint, the grid rebuild beside the grid
// adjoint (lead 0: not here).
struct P2GRebuild { TileArgs ta; const float* st; const float* aff; float4* tiles; int lead; };
DEV void p2g_scatter_body(float* tile, const TileArgs& ta, const float* __restrict__ st, const float* __restrict__ aff,
                          float4* __restrict__ tiles, int only_overflow, int lead = 0, int nwg = 0);
struct GridRebuild { GridArgs ga; const float4* tiles; float4* g_pm; float4* g_v; int lead; };

// reverse of g2p over work items: scatter the adjoint of grid_v_out into the
// item's tile buffer, write the first part of the adjoint of x (frame f).
// a1 = adjoint of frame f+1, st1 = frame f+1.
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(3))) k_g2p_bwd_tile(TileArgs ta, const float* __restrict__ st, const float* __restrict__ st1,
                                                     const float* __restrict__ a1, float* __restrict__ a0,
                                                     const float4* __restrict__ g_v, float4* __restrict__ tiles_v, int use_kept,
                                                     P2GRebuild rb, const int* __restrict__ perm1) {
    // the velocity tile and the wave tiles, or the p2g tiles of a rebuild
    __shared__ float4 smem[4 * TILE];
    float4* tv = smem;
    float* acc = reinterpret_cast<float*>(smem + TILE);
    if ((int)blockIdx.x < rb.lead) {
        p2g_scatter_body(reinterpret_cast<float*>(smem), rb.ta, rb.st, rb.aff, rb.tiles, 1, 0, rb.lead);
        return;
    }
    ItemIter ir = item_iter(ta, rb.lead);
    bool kept = use_kept && kept_fits(ta.counters, ta.keep_cap);
    int lane = threadIdx.x & 63;
    int stride = ta.stride;
    // every wave scatters into a tile of its own (no LDS atomics between
    // waves), the workgroup's tiles are summed in wave order at the flush
    int n_acc = (WG / 64) * 3 * TILE;
    float* wacc = acc + (threadIdx.x >> 6) * 3 * TILE;
    for (int j = ir.j; j < ir.n; j += ir.step) {
        int e;
        int4 it = item_at(ta, ir, j, e);
        // the first batch's particle data does not depend on the tile: its
        // loads are issued first, so both latencies overlap
        int p = ta.order[it.y + min((int)threadIdx.x, it.z - 1)];
        v3 x = load3(st, stride, 0, p);
        // frame f+1 and its adjoint hold particle p at slot perm1[p] when the
        // frame was re-sorted
        int p1 = perm1 ? perm1[p] : p;
        v3 x1 = load3(st1, stride, 0, p1);
        v3 gx1 = load3(a1, stride, 0, p1);
        v3 gv1 = load3(a1, stride, 3, p1);
        m3 gC = load9(a1, stride, 6, p1);
        const SceneC& S = ta.scenes[ta.blk_scene[it.x]];
        int bi, bj, bk;
        block_origin(S, it.x, bi, bj, bk);
        if (kept) load_tile_kept(tv, S, bi, bj, bk, ta.keep_v, ta.nmap, ta.keep_cap);
        else load_tile(tv, S, bi, bj, bk, g_v);
        for (int k = threadIdx.x; k < n_acc; k += WG) acc[k] = 0.f;
        __syncthreads();
        for (int k0 = 0; k0 < it.z; k0 += WG) {
            int k = k0 + threadIdx.x;
            bool valid = k < it.z;
            if (k0 > 0) {
                p = valid ? ta.order[it.y + k] : ta.order[it.y];
                x = load3(st, stride, 0, p);
                p1 = perm1 ? perm1[p] : p;
                x1 = load3(st1, stride, 0, p1);
                gx1 = load3(a1, stride, 0, p1);
                gv1 = load3(a1, stride, 3, p1);
                gC = load9(a1, stride, 6, p1);
            }
            // clamp mask from the stored output position: the adjoint passes iff lower < x' < upper
            v3 gxn = mk((S.lower < x1.x && x1.x < S.upper) ? gx1.x : 0.f, (S.lower < x1.y && x1.y < S.upper) ? gx1.y : 0.f,
                        (S.lower < x1.z && x1.z < S.upper) ? gx1.z : 0.f);
            v3 gnv = gv1 + S.dt * gxn;
            Stencil sw;
            stencil(S, x, sw);
            int l0 = sw.base[0] - 4 * bi, l1 = sw.base[1] - 4 * bj, l2 = sw.base[2] - 4 * bk;
            bool in_tile = (unsigned)l0 < 4u && (unsigned)l1 < 4u && (unsigned)l2 < 4u;
            float gfx[3] = {0.f, 0.f, 0.f};
            // Per node n = (a, b, c) with dp = n - fx, the adjoint of the node's
            // velocity is wt u, u = gnv + c4 gC dp = u0 + a K0 + b K1 + c K2
            // (Kj = c4 gC[:, j], u0 = gnv - c4 gC fx).
            float c4 = S.c4;
            v3 K0 = mk(c4 * gC.a[0], c4 * gC.a[3], c4 * gC.a[6]);
            v3 K1 = mk(c4 * gC.a[1], c4 * gC.a[4], c4 * gC.a[7]);
            v3 K2 = mk(c4 * gC.a[2], c4 * gC.a[5], c4 * gC.a[8]);
            v3 u0 = gnv - (sw.fx[0] * K0 + sw.fx[1] * K1 + sw.fx[2] * K2);
            if (in_tile) {
                // The weight's adjoint is gw = g . u, and gfx[d] sums gw times the
                // derivative of the weight along d minus c4 (gC^T sum_n wt g)[d].
                // Both factor over the three axes, so they are accumulated axis by axis.
                v3 ua = u0;
                v3 G0 = mk(0.f, 0.f, 0.f);
                float X0 = 0.f, X1 = 0.f, X2 = 0.f;
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    v3 ub = ua, Ja = mk(0.f, 0.f, 0.f);
                    float Ra = 0.f, Sa = 0.f, Ta = 0.f;
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        v3 u = ub, H = mk(0.f, 0.f, 0.f);
                        float P = 0.f, Q = 0.f;
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            float4 g4 = tv[((l0 + a) * 6 + (l1 + b)) * 6 + (l2 + c)];
                            v3 g = mk(g4.x, g4.y, g4.z);
                            float gw = dot(g, u);
                            P += sw.w[c][2] * gw;
                            Q += sw.dw[c][2] * gw;
                            H += sw.w[c][2] * g;
                            u += K2;
                        }
                        Ja += sw.w[b][1] * H;
                        Ra += sw.w[b][1] * P;
                        Sa += sw.dw[b][1] * P;
                        Ta += sw.w[b][1] * Q;
                        ub += K1;
                    }
                    G0 += sw.w[a][0] * Ja;
                    X0 += sw.dw[a][0] * Ra;
                    X1 += sw.w[a][0] * Sa;
                    X2 += sw.w[a][0] * Ta;
                    ua += K0;
                }
                gfx[0] = X0 - c4 * (gC.a[0] * G0.x + gC.a[3] * G0.y + gC.a[6] * G0.z);
                gfx[1] = X1 - c4 * (gC.a[1] * G0.x + gC.a[4] * G0.y + gC.a[7] * G0.z);
                gfx[2] = X2 - c4 * (gC.a[2] * G0.x + gC.a[5] * G0.y + gC.a[8] * G0.z);
            }
            if (valid) store3(a0, stride, 0, p, gxn + S.inv_dx * mk(gfx[0], gfx[1], gfx[2]));
            int key = (l0 << 4) | (l1 << 2) | l2;
            // the node's velocity-adjoint contribution is wt u (u as above),
            // formed per component along the stencil
            auto chan = [&](int c, float* r) {
                float k0 = comp(K0, c), k1 = comp(K1, c), k2 = comp(K2, c);
                float ua = comp(u0, c);
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    float ub = ua;
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        float wab = sw.w[a][0] * sw.w[b][1];
                        float u = ub;
#pragma unroll
                        for (int cc = 0; cc < 3; cc++) {
                            r[(a * 3 + b) * 3 + cc] = wab * (sw.w[cc][2] * u);
                            u += k2;
                        }
                        ub += k1;
                    }
                    ua += k0;
                }
            };
            scatter_tile_ch<3>(wacc, valid && in_tile, key, l0, l1, l2, lane, chan);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < TILE; k += WG) {
            float4 r = make_float4(acc[k], acc[TILE + k], acc[2 * TILE + k], 0.f);
#pragma unroll
            for (int w = 1; w < WG / 64; w++) {
                const float* t = acc + w * 3 * TILE;
                r.x += t[k]; r.y += t[TILE + k]; r.z += t[2 * TILE + k];
            }
            tiles_v[(size_t)e * TILE + k] = r;
        }
        __syncthreads();
    }
}

// reverse of grid_op: adjoint of (momentum, mass) per node, plus the adjoint
// of the primitive states of frames f and f+1 (wave-reduced, then atomics).
// One wave per workgroup and per node block: at two waves per SIMD (its
// register count), a wave slot frees as soon as its own block is done.
__global__ void __launch_bounds__(64) k_grid_op_bwd(GridArgs ga, const float4* __restrict__ g_pm, const float4* __restrict__ tiles_v,
                                                    float4* __restrict__ g_adjpm, float* __restrict__ gpstate, int use_kept,
                                                    GridRebuild rb) {
    __shared__ SceneStage stage[1];
    __shared__ float s_vu[MAX_UNITS * 3 * 64];
    if ((int)blockIdx.x < rb.lead) {
        grid_op_body<1>(rb.ga, rb.tiles, rb.g_pm, rb.g_v, stage[0], nullptr, 1, 0, rb.lead);
        return;
    }
    int lane = threadIdx.x & 63;
    int nb = (int)gridDim.x - rb.lead;
    int e = uniform((int)blockIdx.x - rb.lead);
    // the first block's record and kept (momentum, mass) are read together
    // with the counters (a record or kept block beyond the active count is
    // stale and left unused)
    float4 pm0 = (use_kept && e < ga.keep_cap) ? ga.keep_pm[(size_t)e * NODES_PER_BLK + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    BlockMeta m;
    bool have = e < ga.rec_cap;
    read_rec(ga, min(e, ga.rec_cap - 1), m);
    int n_active = ga.counters[CNT_NBLK];
    bool kept = use_kept && kept_fits(ga.counters, ga.keep_cap);
    int stamp = ga.counters[CNT_STAMP];
    bool first = true;
    for (; e < n_active; e += nb) {
        float4 pm = kept ? (first ? pm0 : ga.keep_pm[(size_t)e * NODES_PER_BLK + lane]) : make_float4(0.f, 0.f, 0.f, 0.f);
        first = false;
        if (!have) block_meta(ga, e, stamp, m);
        have = false;
        stage_scene(stage[0], ga, m, lane);
        const SceneC& S = stage[0].sc;
        int B0, B1, B2, I0, I1, I2;
        long long node;
        node_coords(S, m.gb, lane, B0, B1, B2, I0, I1, I2, node);
        if (!kept) pm = g_pm[node];
        float4 gv4 = gather_tiles<8>(m, lane, tiles_v);
        bool act = pm.w > 1e-12f && in_grid(S, I0, I1, I2);
        const PrimC* prims = stage[0].pr;
        const float* ps0 = stage[0].ps[0];
        const float* ps1 = stage[0].ps[1];
        float* gps0 = gpstate + ((size_t)ga.f * ga.n_prims_total + S.prim_off) * 8;
        float* gps1 = gpstate + ((size_t)(ga.f + 1) * ga.n_prims_total + S.prim_off) * 8;
        NodeTr tr;
        tr.vu = s_vu;
        v3 g = mk(gv4.x, gv4.y, gv4.z);
        if (act) {
            grid_node(S, prims, ps0, ps1, I0, I1, I2, pm, &tr);
            int Iv[3] = {I0, I1, I2};
            for (int d = 2; d >= 0; d--) {
                v3 v = tr.vw[d];
                v3 vmid = v;
                int mode = 0;
                float lin = 0.f, lit = 1.f, c = 0.f, sarg = 0.f;
                v3 vit = mk(0.f, 0.f, 0.f);
                if (Iv[d] < 3 && comp(v, d) < 0.f) {
                    if (d != 1 || S.gf_mode == 0) { mode = 1; setc(vmid, d, 0.f); }
                    else if (S.gf_mode == 1) {
                        mode = 2;
                        lin = v.y + 1e-30f;
                        vit = mk(v.x - (float)I0 * 1e-30f, v.y - lin - (float)I1 * 1e-30f, v.z - (float)I2 * 1e-30f);
                        lit = sqrtf(dot(vit, vit) + 1e-8f);
                        sarg = 1.f + S.gf * lin / lit;
                        c = tmax(sarg, 0.f);
                        vmid = mk(c * (vit.x + (float)I0 * 1e-30f), 0.f, c * (vit.z + (float)I2 * 1e-30f));
                    } else { mode = 3; vmid = mk(0.f, 0.f, 0.f); }
                }
                if (Iv[d] > S.n_grid - 3 && comp(vmid, d) > 0.f) setc(g, d, 0.f);
                if (mode == 1) setc(g, d, 0.f);
                else if (mode == 3) g = mk(0.f, 0.f, 0.f);
                else if (mode == 2) {
                    g.y = 0.f;
                    v3 uu = mk(vit.x + (float)I0 * 1e-30f, vit.y + (float)I1 * 1e-30f, vit.z + (float)I2 * 1e-30f);
                    float gc = dot(g, uu);
                    v3 gvit = c * g;
                    float gs = (0.f < sarg) ? gc : 0.f;
                    float glin = gs * S.gf / lit;
                    float glit = -gs * S.gf * lin / (lit * lit);
                    gvit += (glit / lit) * vit;
                    glin -= gvit.y;
                    g = gvit;
                    g.y += glin;
                }
            }
        }
        // collision units in reverse order; each unit's primitive adjoint is wave-reduced
        v3 x = mk((float)I0 * S.dx, (float)I1 * S.dx, (float)I2 * S.dx);
        const int blk_gb = m.gb;
        bool blk_any = false;
        int uidx = 0;
        for (int q = 0; q < S.n_prims; q++) uidx += prim_unit_count(prims[q]);
        for (int q = S.n_prims - 1; q >= 0; q--) {
            const PrimC& P = prims[q];
            PS s0 = load_ps(ps0 + q * 8), s1 = load_ps(ps1 + q * 8);
            int m = prim_unit_count(P);
            for (int j = m - 1; j >= 0; j--) {
                uidx--;
                PSg g0 = zero_psg(), g1 = zero_psg();
                if (act && uidx < MAX_UNITS) {
                    v3 vin = mk(tr.vu[(uidx * 3 + 0) * 64 + lane], tr.vu[(uidx * 3 + 1) * 64 + lane], tr.vu[(uidx * 3 + 2) * 64 + lane]);
                    g = unit_collide_rev(P, prim_unit(P, j), s0, s1, x, S.dt, vin, g, &g0, &g1);
                }
                float vals[16] = {g0.pos.x, g0.pos.y, g0.pos.z, g0.rot.w, g0.rot.x, g0.rot.y, g0.rot.z, g0.gap,
                                  g1.pos.x, g1.pos.y, g1.pos.z, g1.rot.w, g1.rot.x, g1.rot.y, g1.rot.z, g1.gap};
                bool any = false;
#pragma unroll
                for (int k = 0; k < 16; k++) any |= (vals[k] != 0.f);
                int k = lane >> 2;
                if (__ballot(any) == 0ull) {
                    if (ga.det && (lane & 3) == 0) ga.padj[((size_t)blk_gb * MAX_UNITS + uidx) * 16 + k] = 0.f;
                    continue;
                }
                blk_any = true;
                // the 16 wave sums at once; quad k adds value k
                float s = tsum16(vals, lane);
                if (ga.det) {
                    if ((lane & 3) == 0) ga.padj[((size_t)blk_gb * MAX_UNITS + uidx) * 16 + k] = s;
                } else if ((lane & 3) == 0 && s != 0.f) {
                    unsafeAtomicAdd((k < 8 ? gps0 : gps1) + q * 8 + (k & 7), s);
                }
            }
        }
        // (blocks whose units all have zero adjoint add nothing and are not flagged)
        if (ga.det && lane == 0 && blk_any) {
            int lb = blk_gb - S.blk_off;
            atomicOr(&ga.pbits[S.bm_off + (lb >> 5)], 1u << (lb & 31));
        }
        float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
        if (act) {
            float inv_m = 1.f / pm.w;
            out.x = inv_m * g.x; out.y = inv_m * g.y; out.z = inv_m * g.z;
            out.w = -(g.x * pm.x + g.y * pm.y + g.z * pm.z) / (pm.w * pm.w);
        }
        g_adjpm[node] = out;
    }
}

// 1/x from the hardware estimate (v_rcp_f32) and one Newton step
// (y <- y + y (1 - x y)), for the adjoint's divisions by nonzero values
DEV float frcp(float x) {
    float y = __builtin_amdgcn_rcpf(x);
    return fmaf(y, fmaf(-x, y, 1.f), y);
}

// backward_svd (mpm_simulator.py:136-156) with the gap clamp of :184-192
DEV float gap_clamp(float a) { return a >= 0.f ? tmax(a, 1e-6f) : tmin(a, -1e-6f); }
DEV m3 backward_svd(const m3& gu, const float gs[3], const m3& gv, const m3& u, const float sig[3], const m3& v) {
    float s[3] = {sig[0] * sig[0], sig[1] * sig[1], sig[2] * sig[2]};
    m3 Fm;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Fm.a[i * 3 + j] = (i == j) ? 0.f : frcp(gap_clamp(s[j] - s[i]));
    m3 S = mzero(), GS = mzero();
#pragma unroll
    for (int i = 0; i < 3; i++) { S.a[i * 4] = sig[i]; GS.a[i * 4] = gs[i]; }
    m3 sigma_term = mmul_nt(mmul(u, GS), v);
    m3 t1 = mmul_tn(u, gu), t2 = mmul_tn(gu, u), M;
#pragma unroll
    for (int i = 0; i < 9; i++) M.a[i] = Fm.a[i] * (t1.a[i] - t2.a[i]);
    m3 u_term = mmul_nt(mmul(u, mmul(M, S)), v);
    t1 = mmul_tn(v, gv);
    t2 = mmul_tn(gv, v);
#pragma unroll
    for (int i = 0; i < 9; i++) M.a[i] = Fm.a[i] * (t1.a[i] - t2.a[i]);
    m3 v_term = mmul(u, mmul(S, mmul_nt(M, v)));
    m3 out;
#pragma unroll
    for (int i = 0; i < 9; i++) out.a[i] = u_term.a[i] + v_term.a[i] + sigma_term.a[i];
    return out;
}

// constitutive update in particle-slot order (coalesced): F_tmp, register SVD,
// return map, stress (mpm_simulator.py:121-129,165-182,205-216).  Writes the
// affine matrix stress_k * stress + p_mass * C of frame f, the SVD of F_tmp
// (U, sigma, V: what the adjoint needs instead of a second SVD) and, when
// st_out is given, F of frame f+1.
__global__ void __launch_bounds__(PPB) k_stress(ChunkArgs ca, const float* __restrict__ st, float* __restrict__ st_out,
                                                float* __restrict__ aff, float* __restrict__ usv) {
    int4 ch = ca.chunks[blockIdx.x];
    int i = threadIdx.x;
    if (i >= ch.z) return;
    const SceneC& S = ca.scenes[ch.x];
    int slot = ch.y + i;
    m3 C = load9(st, ca.stride, 6, slot);
    m3 F = load9(st, ca.stride, 15, slot);
    PTrace t;
    particle_stress(S, C, F, t);
    if (st_out) store9(st_out, ca.stride, 15, slot, t.Fn);
    store9(aff, ca.stride, 0, slot, t.affine);
    store_usv(usv, ca.stride, slot, t);
}

// p2g scatter over work items (mpm_simulator.py:217-225): momentum and mass of
// every particle of the item into the block's LDS tile, flushed to the item's
// tile buffer.  Reads x, v and the stored affine matrix.
DEV void p2g_scatter_body(float* tile, const TileArgs& ta, const float* __restrict__ st, const float* __restrict__ aff,
                          float4* __restrict__ tiles, int only_overflow, int lead, int nwg) {
    if (only_overflow && kept_fits(ta.counters, ta.keep_cap)) return;
    ItemIter ir = item_iter(ta, lead, nwg);
    int lane = threadIdx.x & 63;
    int stride = ta.stride;
    // every wave scatters into a tile of its own (no LDS atomics between
    // waves), the workgroup's tiles are summed in wave order at the flush
    int n_tile = (WG / 64) * 4 * TILE;
    float* wtile = tile + (threadIdx.x >> 6) * 4 * TILE;
    for (int j = ir.j; j < ir.n; j += ir.step) {
        int e;
        int4 it = item_at(ta, ir, j, e);
        const SceneC& S = ta.scenes[ta.blk_scene[it.x]];
        int bi, bj, bk;
        block_origin(S, it.x, bi, bj, bk);
        for (int k = threadIdx.x; k < n_tile; k += WG) tile[k] = 0.f;
        __syncthreads();
        for (int k0 = 0; k0 < it.z; k0 += WG) {
            int k = k0 + threadIdx.x;
            bool valid = k < it.z;
            int p = valid ? ta.order[it.y + k] : ta.order[it.y];
            v3 x = load3(st, stride, 0, p);
            v3 v = load3(st, stride, 3, p);
            m3 Am = load9(aff, stride, 0, p);
            Stencil sw;
            stencil(S, x, sw);
            int l0 = sw.base[0] - 4 * bi, l1 = sw.base[1] - 4 * bj, l2 = sw.base[2] - 4 * bk;
            bool in_tile = (unsigned)l0 < 4u && (unsigned)l1 < 4u && (unsigned)l2 < 4u;
            int key = (l0 << 4) | (l1 << 2) | l2;
            v3 mv = S.p_mass * v;
            const float* A = Am.a;
            float dx = S.dx, pm = S.p_mass;
            // momentum component c at node n = (a, b, cc): wt (mv + A dpos)[c]
            // with dpos = (n - fx) dx, formed along the stencil as u0 + a k0 +
            // b k1 + cc k2 (kj = dx A[c][j]); component 3 is the mass wt pm
            auto chan = [&](int c, float* r) {
                float k0 = c < 3 ? dx * A[c * 3] : 0.f, k1 = c < 3 ? dx * A[c * 3 + 1] : 0.f, k2 = c < 3 ? dx * A[c * 3 + 2] : 0.f;
                float ua = c < 3 ? comp(mv, c) - (sw.fx[0] * k0 + sw.fx[1] * k1 + sw.fx[2] * k2) : pm;
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    float ub = ua;
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        float wab = sw.w[a][0] * sw.w[b][1];
                        float u = ub;
#pragma unroll
                        for (int cc = 0; cc < 3; cc++) {
                            r[(a * 3 + b) * 3 + cc] = wab * (sw.w[cc][2] * u);
                            u += k2;
                        }
                        ub += k1;
                    }
                    ua += k0;
                }
            };
            scatter_tile_ch<4>(wtile, valid && in_tile, key, l0, l1, l2, lane, chan);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < TILE; k += WG) {
            float4 r = make_float4(tile[k], tile[TILE + k], tile[2 * TILE + k], tile[3 * TILE + k]);
#pragma unroll
            for (int w = 1; w < WG / 64; w++) {
                const float* t = tile + w * 4 * TILE;
                r.x += t[k]; r.y += t[TILE + k]; r.z += t[2 * TILE + k]; r.w += t[3 * TILE + k];
            }
            tiles[(size_t)e * TILE + k] = r;
        }
        __syncthreads();
    }
}
// the forward substep's p2g
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(3))) k_p2g_scatter(TileArgs ta, const float* __restrict__ st, const float* __restrict__ aff,
                                                    float4* __restrict__ tiles) {
    __shared__ float tile[(WG / 64) * 4 * TILE];
    p2g_scatter_body(tile, ta, st, aff, tiles, 0);
}
// the same scatter rebuilding a substep's grid for its adjoint (see
// k_grid_op_rebuild)
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(3))) k_p2g_scatter_rebuild(TileArgs ta, const float* __restrict__ st, const float* __restrict__ aff,
                                                    float4* __restrict__ tiles, int only_overflow) {
    __shared__ float tile[(WG / 64) * 4 * TILE];
    p2g_scatter_body(tile, ta, st, aff, tiles, only_overflow);
}

// Fixed-order mode: the primitive adjoints of substep f from the grid
// adjoint's per-block partials (written for the node blocks whose bit the grid
// adjoint set: those where some collision unit's adjoint is nonzero).
struct PrimRedArgs {
    const SceneC* scenes;
    const PrimC* prims;
    const float* padj;
    unsigned* pbits;
    float* gpstate;
    int n_prims_total, f, n_scenes;
};
constexpr int PR_LIST = 64 * 32;   // flagged node blocks listed per round (64 bitmap words)

// Scene sc, by the first wave of the calling workgroup (every thread of the
// workgroup calls this).  In rounds of 64 bitmap words, the flagged blocks
// are listed in block order (the words' popcounts scanned over the wave),
// lane l sums the l-th 64th of the list in list order, the lanes' sums are
// combined by the fixed butterfly of tsum16 and added to the running totals
// in round order; then each primitive's units are added in unit order into the
// adjoint of frames f (values 0-7) and f+1 (8-15).  The order depends only on
// which blocks are flagged, never on timing or launch shape.  The bitmap
// words are cleared as they are read.
DEV void prim_adj_reduce_scene(const PrimRedArgs& pr, int sc, int* lst, float (*tot)[16]) {
    const SceneC& S = pr.scenes[sc];
    int lane = threadIdx.x & 63;
    bool wk = threadIdx.x < 64;
    int nbk = S.nb * S.nb * S.nb;
    int nw = (nbk + 31) >> 5;
    int n_units = 0;
    for (int q = 0; q < S.n_prims; q++) n_units += pr.prims[S.prim_off + q].n_units;
    for (int w0 = 0; w0 < nw; w0 += 64) {
        int wi = w0 + lane;
        unsigned wd = (wk && wi < nw) ? pr.pbits[S.bm_off + wi] : 0u;
        int c = __popc(wd), incl = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            int u = __shfl_up(incl, d);
            if (lane >= d) incl += u;
        }
        int n = __shfl(incl, 63);
        if (wd) {
            int o = incl - c;
            for (unsigned w = wd; w; w &= w - 1) lst[o++] = wi * 32 + __ffs(w) - 1;
            pr.pbits[S.bm_off + wi] = 0u;
        }
        __syncthreads();
        int i0 = (n * lane) >> 6, i1 = wk ? (n * (lane + 1)) >> 6 : i0;
        for (int u = 0; u < n_units; u++) {
            float a[16];
#pragma unroll
            for (int k = 0; k < 16; k++) a[k] = 0.f;
            for (int i = i0; i < i1; i++) {
                const float4* r = reinterpret_cast<const float4*>(pr.padj + ((size_t)(S.blk_off + lst[i]) * MAX_UNITS + u) * 16);
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    float4 v = r[j];
                    a[4 * j] += v.x; a[4 * j + 1] += v.y; a[4 * j + 2] += v.z; a[4 * j + 3] += v.w;
                }
            }
            // quad k of the wave holds the total of value k
            float t = tsum16(a, lane);
            if (wk && (lane & 3) == 0) tot[u][lane >> 2] = (w0 == 0 ? 0.f : tot[u][lane >> 2]) + t;
        }
        __syncthreads();
    }
    if (wk && lane < 16) {
        int u = 0;
        for (int q = 0; q < S.n_prims; q++) {
            int nu = pr.prims[S.prim_off + q].n_units;
            float v = 0.f;
            for (int j = 0; j < nu; j++, u++) v += tot[u][lane];
            int fr = lane < 8 ? pr.f : pr.f + 1;
            pr.gpstate[((size_t)fr * pr.n_prims_total + S.prim_off + q) * 8 + (lane & 7)] += v;
        }
    }
}

// first half of the p2g adjoint over work items: gathers the node adjoints of
// (momentum, mass) from the block's LDS tile and produces the adjoints of v,
// of the affine matrix (to gaff) and the weight/position part of x.
__global__ void __launch_bounds__(WG) k_p2g_bwd_gather(TileArgs ta, const float* __restrict__ st, const float* __restrict__ aff,
                                                       float* __restrict__ a0, float* __restrict__ gaff_out,
                                                       const float4* __restrict__ g_adjpm, PrimRedArgs pr, int lead) {
    __shared__ float4 tg[TILE];
    __shared__ int lst[PR_LIST];
    __shared__ float tot[MAX_UNITS][16];
    // the fixed-order mode's primitive-adjoint reduction of this substep, one
    // leading workgroup per scene, beside the work items
    if ((int)blockIdx.x < lead) {
        if ((int)blockIdx.x < pr.n_scenes && pr.scenes[blockIdx.x].n_prims > 0) prim_adj_reduce_scene(pr, blockIdx.x, lst, tot);
        return;
    }
    ItemIter ir = item_iter(ta, lead);
    int stride = ta.stride;
    for (int j = ir.j; j < ir.n; j += ir.step) {
        int e;
        int4 it = item_at(ta, ir, j, e);
        // the first batch's particle data does not depend on the tile: its
        // loads are issued first, so both latencies overlap
        int k = threadIdx.x;
        int p = ta.order[it.y + min(k, it.z - 1)];
        v3 x = load3(st, stride, 0, p);
        v3 v = load3(st, stride, 3, p);
        m3 Am = load9(aff, stride, 0, p);
        const SceneC& S = ta.scenes[ta.blk_scene[it.x]];
        int bi, bj, bk;
        block_origin(S, it.x, bi, bj, bk);
        load_tile(tg, S, bi, bj, bk, g_adjpm);
        __syncthreads();
        for (; k < it.z; k += WG) {
            // the next batch's particle data, one batch ahead
            int pn = p;
            v3 xn = x, vn = v;
            m3 An = Am;
            if (k + WG < it.z) {
                pn = ta.order[it.y + k + WG];
                xn = load3(st, stride, 0, pn);
                vn = load3(st, stride, 3, pn);
                An = load9(aff, stride, 0, pn);
            }
            const float* A = Am.a;
            Stencil sw;
            stencil(S, x, sw);
            int l0 = sw.base[0] - 4 * bi, l1 = sw.base[1] - 4 * bj, l2 = sw.base[2] - 4 * bk;
            bool in_tile = (unsigned)l0 < 4u && (unsigned)l1 < 4u && (unsigned)l2 < 4u;
            v3 gv = mk(0.f, 0.f, 0.f);
            m3 gaff = mzero();
            float gfx[3] = {0.f, 0.f, 0.f};
            v3 mv = S.p_mass * v;
            if (in_tile) {
                // Per node n = (a, b, c) with dpos = dx (n - fx): the node's
                // momentum is wt u, u = mv + A dpos = u0 + a K0 + b K1 + c K2
                // (Kj = dx A[:, j]), and the weight's adjoint gw = g . u + g_m pm.
                // The adjoints are sums over n of wt g (v), wt g dpos^T (affine),
                // gw times the weight's derivative along d and A^T wt g (x); the
                // weights factor over the axes, so all of them are accumulated axis
                // by axis (pw = w (n - fx) along one axis).
                float dx = S.dx, pm = S.p_mass;
                float pw[3][3];
#pragma unroll
                for (int o = 0; o < 3; o++)
#pragma unroll
                    for (int d = 0; d < 3; d++) pw[o][d] = sw.w[o][d] * ((float)o - sw.fx[d]);
                v3 K0 = mk(dx * A[0], dx * A[3], dx * A[6]);
                v3 K1 = mk(dx * A[1], dx * A[4], dx * A[7]);
                v3 K2 = mk(dx * A[2], dx * A[5], dx * A[8]);
                v3 ua = mv - (sw.fx[0] * K0 + sw.fx[1] * K1 + sw.fx[2] * K2);
                v3 G0 = mk(0.f, 0.f, 0.f), Gx = G0, Gy = G0, Gz = G0;
                float X0 = 0.f, X1 = 0.f, X2 = 0.f;
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    v3 ub = ua, J = mk(0.f, 0.f, 0.f), Jy = J, Jz = J;
                    float Ra = 0.f, Sa = 0.f, Ta = 0.f;
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        v3 u = ub, H = mk(0.f, 0.f, 0.f), Hz = H;
                        float P = 0.f, Q = 0.f;
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            float4 gg = tg[((l0 + a) * 6 + (l1 + b)) * 6 + (l2 + c)];
                            v3 g = mk(gg.x, gg.y, gg.z);
                            float gw = dot(g, u) + gg.w * pm;
                            P += sw.w[c][2] * gw;
                            Q += sw.dw[c][2] * gw;
                            H += sw.w[c][2] * g;
                            Hz += pw[c][2] * g;
                            u += K2;
                        }
                        J += sw.w[b][1] * H;
                        Jy += pw[b][1] * H;
                        Jz += sw.w[b][1] * Hz;
                        Ra += sw.w[b][1] * P;
                        Sa += sw.dw[b][1] * P;
                        Ta += sw.w[b][1] * Q;
                        ub += K1;
                    }
                    G0 += sw.w[a][0] * J;
                    Gx += pw[a][0] * J;
                    Gy += sw.w[a][0] * Jy;
                    Gz += sw.w[a][0] * Jz;
                    X0 += sw.dw[a][0] * Ra;
                    X1 += sw.w[a][0] * Sa;
                    X2 += sw.w[a][0] * Ta;
                    ua += K0;
                }
                gv = pm * G0;
                gaff.a[0] = dx * Gx.x, gaff.a[1] = dx * Gy.x, gaff.a[2] = dx * Gz.x;
                gaff.a[3] = dx * Gx.y, gaff.a[4] = dx * Gy.y, gaff.a[5] = dx * Gz.y;
                gaff.a[6] = dx * Gx.z, gaff.a[7] = dx * Gy.z, gaff.a[8] = dx * Gz.z;
                gfx[0] = X0 - dx * (A[0] * G0.x + A[3] * G0.y + A[6] * G0.z);
                gfx[1] = X1 - dx * (A[1] * G0.x + A[4] * G0.y + A[7] * G0.z);
                gfx[2] = X2 - dx * (A[2] * G0.x + A[5] * G0.y + A[8] * G0.z);
            }
            v3 gx0 = load3(a0, stride, 0, p) + S.inv_dx * mk(gfx[0], gfx[1], gfx[2]);
            store3(a0, stride, 0, p, gx0);
            store3(a0, stride, 3, p, gv);
            store9(gaff_out, stride, 0, p, gaff);
            p = pn;
            x = xn;
            v = vn;
            Am = An;
        }
        __syncthreads();
    }
}

// second half of the p2g adjoint in particle-slot order: stress, return map,
// backward_svd, F_tmp (mpm_simulator.py:121-156,205-216); writes the adjoints
// of C and F of frame f.  The SVD of F_tmp is the one the forward stored for
// frame f.  Workgroup 0 also runs the primitive FK adjoint of substep f.
__global__ void __launch_bounds__(PPB) k_stress_bwd(ChunkArgs ca, const float* __restrict__ st, const float* __restrict__ gaff_in,
                                                    const float* __restrict__ a1, float* __restrict__ a0,
                                                    const float* __restrict__ usv,
                                                    const PrimC* __restrict__ prims, int n_prims_total,
                                                    const float* __restrict__ pstate, const float* __restrict__ pvel,
                                                    float* __restrict__ gpstate, float* __restrict__ gpvel, int f,
                                                    const int* __restrict__ perm1) {
    if (blockIdx.x == 0)
        for (int q = threadIdx.x; q < n_prims_total; q += blockDim.x) {
            size_t o0 = ((size_t)f * n_prims_total + q) * 8, o1 = ((size_t)(f + 1) * n_prims_total + q) * 8;
            prim_fk_rev(prims[q], pstate + o0, pvel + o0, gpstate + o1, gpstate + o0, gpvel + o0);
        }
    int4 ch = ca.chunks[blockIdx.x];
    int i = threadIdx.x;
    if (i >= ch.z) return;
    const SceneC& S = ca.scenes[ch.x];
    int p = ch.y + i;
    int stride = ca.stride;
    m3 C = load9(st, stride, 6, p);
    m3 F = load9(st, stride, 15, p);
    PTrace t;
    t.Ftmp = f_tmp(S, C, F);
    load_usv(usv, stride, p, t);
    stress_from_svd(S, C, t);
    m3 gaff = load9(gaff_in, stride, 0, p);
    m3 gC, gst;
#pragma unroll
    for (int kk = 0; kk < 9; kk++) { gC.a[kk] = S.p_mass * gaff.a[kk]; gst.a[kk] = S.stress_k * gaff.a[kk]; }
    m3 FmR;
#pragma unroll
    for (int kk = 0; kk < 9; kk++) FmR.a[kk] = t.Fn.a[kk] - t.R.a[kk];
    m3 sF = mmul(gst, t.Fn), sT = mmul_tn(gst, FmR);
    m3 gF1 = load9(a1, stride, 15, perm1 ? perm1[p] : p);   // frame f+1 re-sorted: slot perm1[p]
    m3 gF, gR;
#pragma unroll
    for (int kk = 0; kk < 9; kk++) { gF.a[kk] = 2.f * S.mu * (sF.a[kk] + sT.a[kk]) + gF1.a[kk]; gR.a[kk] = -2.f * S.mu * sF.a[kk]; }
    float gJ = S.lam * (2.f * t.J - 1.f) * (gst.a[0] + gst.a[4] + gst.a[8]);
    m3 cf = mcof(t.Fn);
#pragma unroll
    for (int kk = 0; kk < 9; kk++) gF.a[kk] += gJ * cf.a[kk];
    m3 gU = mmul(gR, t.V), gV = mmul_tn(gR, t.U), gFt = mzero();
    float gS[3] = {0.f, 0.f, 0.f};
    if (t.yield) {
        m3 D = mzero();
#pragma unroll
        for (int kk = 0; kk < 3; kk++) D.a[kk * 4] = t.e[kk];
        m3 aa = mmul(mmul(gF, t.V), D), bb = mmul(mmul_tn(gF, t.U), D);
#pragma unroll
        for (int kk = 0; kk < 9; kk++) { gU.a[kk] += aa.a[kk]; gV.a[kk] += bb.a[kk]; }
        m3 M = mmul(mmul_tn(t.U, gF), t.V);
        float gep[3], geps[3], geh[3];
        // ehn > 0 on this branch (it exceeds the yield offset dg > 0)
        float inv_ehn = frcp(t.ehn);
        float r = t.dg * inv_ehn, gr = 0.f;
#pragma unroll
        for (int kk = 0; kk < 3; kk++) { gep[kk] = M.a[kk * 4] * t.e[kk]; geps[kk] = gep[kk]; geh[kk] = -r * gep[kk]; gr -= gep[kk] * t.eh[kk]; }
        float gdg = gr * inv_ehn;
        float gehn = -gr * t.dg * (inv_ehn * inv_ehn) + gdg;
#pragma unroll
        for (int kk = 0; kk < 3; kk++) geh[kk] += gehn * t.eh[kk] * inv_ehn;
        float mg = (geh[0] + geh[1] + geh[2]) * (1.f / 3.f);
#pragma unroll
        for (int kk = 0; kk < 3; kk++) {
            geps[kk] += geh[kk] - mg;
            if (0.05f < t.sig[kk]) gS[kk] += geps[kk] * frcp(t.sc[kk]);
        }
    } else {
        gFt = gF;
    }
    m3 bs = backward_svd(gU, gS, gV, t.U, t.sig, t.V);
#pragma unroll
    for (int kk = 0; kk < 9; kk++) gFt.a[kk] += bs.a[kk];
    m3 Am;
#pragma unroll
    for (int kk = 0; kk < 9; kk++) Am.a[kk] = ((kk % 4 == 0) ? 1.f : 0.f) + S.dt * C.a[kk];
    m3 gCa = mmul_nt(gFt, F), gF0 = mmul_tn(Am, gFt);
#pragma unroll
    for (int kk = 0; kk < 9; kk++) gC.a[kk] += S.dt * gCa.a[kk];
    store9(a0, stride, 6, p, gC);
    store9(a0, stride, 15, p, gF0);
}

// one workgroup per scene; used where the reduction must finish before the
// tool-tool collision adjoint (otherwise the p2g adjoint's leading workgroups
// run it beside their own work)
__global__ void __launch_bounds__(64) k_prim_adj_reduce(PrimRedArgs pr) {
    __shared__ int lst[PR_LIST];
    __shared__ float tot[MAX_UNITS][16];
    if (pr.scenes[blockIdx.x].n_prims == 0) return;
    prim_adj_reduce_scene(pr, blockIdx.x, lst, tot);
}

// Fixed-order mode of compute_grid_m_kernel: the wave-reduced node masses are
// added as integers (units of 2^-40 p_mass, exact in any order) into macc,
// then converted and added into the output once all particles are in.
constexpr double MASS_FIX = 1099511627776.0;   // 2^40
__global__ void k_grid_mass_convert(const SceneC* __restrict__ scenes, unsigned long long* __restrict__ macc, float* __restrict__ out) {
    const SceneC& S = scenes[blockIdx.y];
    long long n3 = (long long)S.n_grid * S.n_grid * S.n_grid;
    double scale = (double)S.p_mass_d / MASS_FIX;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long long)gridDim.x * blockDim.x) {
        unsigned long long v = macc[S.m_off + i];
        if (v) {
            out[S.m_off + i] += (float)((double)(long long)v * scale);
            macc[S.m_off + i] = 0ull;
        }
    }
}
DEV void mass_add(float* o, unsigned long long* macc, size_t id, float val, double inv_pmass, bool det) {
    if (det) atomicAdd(macc + id, (unsigned long long)__double2ll_rn((double)val * inv_pmass * MASS_FIX));
    else unsafeAtomicAdd(o + id, val);
}

// compute_grid_m_kernel (mpm_simulator.py:456-466) into a dense per-scene
// grid.  The lanes of a wave that share a stencil base are summed through the
// wave transpose tree and add their 27 node masses with one atomic each (lanes
// past MAXSEG groups add their own), so the kernel wants the lanes of a wave
// to hold particles of few cells: particle p of scene S at lane `valid`.
DEV void grid_mass_wave(const SceneC& S, const float* __restrict__ st, int stride, int p, bool valid, float* __restrict__ out,
                        unsigned long long* __restrict__ macc, int det) {
    int lane = threadIdx.x & 63;
    v3 x = load3(st, stride, 0, p);
    Stencil sw;
    stencil(S, x, sw);
    float* o = out + (size_t)S.m_off;
    unsigned long long* ma = macc + (size_t)S.m_off;
    double inv_pm = 1.0 / S.p_mass_d;
    int n = S.n_grid;
    unsigned long long act = __ballot(valid);
    int seg = 0;
    while (act) {
        if (seg == MAXSEG) {
            if ((act >> lane) & 1ull) {
#pragma unroll
                for (int j = 0; j < 27; j++) {
                    int I0 = sw.base[0] + j / 9, I1 = sw.base[1] + (j / 3) % 3, I2 = sw.base[2] + j % 3;
                    if (in_grid(S, I0, I1, I2))
                        mass_add(o, ma, ((size_t)I0 * n + I1) * n + I2, sw.w[j / 9][0] * sw.w[(j / 3) % 3][1] * sw.w[j % 3][2] * S.p_mass,
                                 inv_pm, det);
                }
            }
            break;
        }
        int lead = __ffsll((long long)act) - 1;
        int s0 = __shfl(sw.base[0], lead), s1 = __shfl(sw.base[1], lead), s2 = __shfl(sw.base[2], lead);
        unsigned long long m = __ballot(valid && sw.base[0] == s0 && sw.base[1] == s1 && sw.base[2] == s2);
        bool in = (m >> lane) & 1ull;
        float r[32];
#pragma unroll
        for (int j = 0; j < 32; j++)
            r[j] = (in && j < 27) ? sw.w[j / 9][0] * sw.w[(j / 3) % 3][1] * sw.w[j % 3][2] * S.p_mass : 0.f;
        float sum = tsum32(r, lane);
        int j = lane >> 1;
        if (!(lane & 1) && j < 27) {
            int I0 = s0 + j / 9, I1 = s1 + (j / 3) % 3, I2 = s2 + j % 3;
            if (in_grid(S, I0, I1, I2)) mass_add(o, ma, ((size_t)I0 * n + I1) * n + I2, sum, inv_pm, det);
        }
        act &= ~m;
        seg++;
    }
}
// particles in slot order, one chunk per workgroup (stored in cell order at
// set_state, they drift from it as the material deforms)
__global__ void __launch_bounds__(PPB) k_grid_mass(ChunkArgs ca, const float* __restrict__ st, float* __restrict__ out,
                                                   unsigned long long* __restrict__ macc, int det) {
    int4 ch = ca.chunks[blockIdx.x];
    bool valid = (int)threadIdx.x < ch.z;
    grid_mass_wave(ca.scenes[ch.x], st, ca.stride, ch.y + (valid ? (int)threadIdx.x : 0), valid, out, macc, det);
}
// particles in the cell order of a kept binning of this frame or the one
// before (its work items), so a wave's lanes share one or two stencil bases
__global__ void __launch_bounds__(WG) k_grid_mass_items(TileArgs ta, const float* __restrict__ st, float* __restrict__ out,
                                                        unsigned long long* __restrict__ macc, int det) {
    ItemIter ir = item_iter(ta);
    for (int j = ir.j; j < ir.n; j += ir.step) {
        int e;
        int4 it = item_at(ta, ir, j, e);
        const SceneC& S = ta.scenes[ta.blk_scene[it.x]];
        for (int k0 = 0; k0 < it.z; k0 += WG) {
            int k = k0 + (int)threadIdx.x;
            bool valid = k < it.z;
            int p = ta.order[it.y + (valid ? k : 0)];
            grid_mass_wave(S, st, ta.stride, p, valid, out, macc, det);
        }
    }
}
// adjoint of compute_grid_m_kernel: gm is the adjoint of the mass grid, or,
// with tgt given, the mass grid itself, whose adjoint under the weighted L1
// density loss (loss.py:173-175, :213) wt * sign(m - tgt) is formed per node
// here instead of in a grid of its own.
__global__ void __launch_bounds__(PPB) k_grid_mass_bwd(ChunkArgs ca, const float* __restrict__ st, const float* __restrict__ gm,
                                                       const float* __restrict__ tgt, float wt, float* __restrict__ adj) {
    int4 ch = ca.chunks[blockIdx.x];
    int i = threadIdx.x;
    if (i >= ch.z) return;
    const SceneC& S = ca.scenes[ch.x];
    int slot = ch.y + i;
    v3 x = load3(st, ca.stride, 0, slot);
    Stencil sw;
    stencil(S, x, sw);
    const float* g = gm + (size_t)S.m_off;
    const float* t = tgt ? tgt + (size_t)S.m_off : nullptr;
    int n = S.n_grid;
    float gfx[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                int I0 = sw.base[0] + a, I1 = sw.base[1] + b, I2 = sw.base[2] + c;
                if (!in_grid(S, I0, I1, I2)) continue;
                size_t id = ((size_t)I0 * n + I1) * n + I2;
                float gn = g[id];
                if (t) {
                    float d = gn - t[id];
                    gn = d > 0.f ? wt : (d < 0.f ? -wt : 0.f);
                }
                float gg = gn * S.p_mass;
                int off[3] = {a, b, c};
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    float dwd = sw.dw[off[d]][d];
#pragma unroll
                    for (int e2 = 0; e2 < 3; e2++)
                        if (e2 != d) dwd *= sw.w[off[e2]][e2];
                    gfx[d] += gg * dwd;
                }
            }
    v3 gx = load3(adj, ca.stride, 0, slot) + S.inv_dx * mk(gfx[0], gfx[1], gfx[2]);
    store3(adj, ca.stride, 0, slot, gx);
}

// reset_kernel (mpm_simulator.py:402-410)
// compute_min_dist (function.py:79-88): distance of every particle of frame f
// to every collision unit of its scene's primitives (both jaws of a two-jaw
// gripper), out[col][stride]; columns past a scene's own count are untouched
__global__ void __launch_bounds__(PPB) k_min_dist(ChunkArgs ca, const float* __restrict__ st, const PrimC* __restrict__ prims,
                                                  const float* __restrict__ ps, float* __restrict__ out) {
    int4 ch = ca.chunks[blockIdx.x];
    int i = threadIdx.x;
    if (i >= ch.z) return;
    const SceneC& S = ca.scenes[ch.x];
    int slot = ch.y + i;
    v3 x = load3(st, ca.stride, 0, slot);
    int col = 0;
    for (int q = 0; q < S.n_prims; q++) {
        const PrimC& P = prims[S.prim_off + q];
        PS s = load_ps(ps + (size_t)(S.prim_off + q) * 8);
        int m = prim_unit_count(P);
        for (int j = 0; j < m; j++, col++) out[(size_t)col * ca.stride + slot] = unit_sdf(P, prim_unit(P, j), s, x);
    }
}

// compute_min_dist.grad (function.py:154-160): gout[col][stride] adjoints of
// the distances; adds d/dx into the particle adjoint (x rows) and the
// primitive-state adjoint of frame f (wave-reduced, one atomic per wave; in
// the fixed-order mode each wave's sums go to pmd[wave][prim] instead and
// k_min_dist_reduce adds them in wave order)
__global__ void __launch_bounds__(PPB) k_min_dist_bwd(ChunkArgs ca, const float* __restrict__ st, const PrimC* __restrict__ prims,
                                                      const float* __restrict__ ps, const float* __restrict__ gout,
                                                      float* __restrict__ adj, float* __restrict__ gps, float* __restrict__ pmd) {
    int4 ch = ca.chunks[blockIdx.x];
    int i = threadIdx.x;
    int lane = i & 63;
    bool valid = i < ch.z;
    const SceneC& S = ca.scenes[ch.x];
    int slot = ch.y + (valid ? i : 0);
    v3 x = load3(st, ca.stride, 0, slot);
    v3 gx = mk(0.f, 0.f, 0.f);
    int col = 0;
    for (int q = 0; q < S.n_prims; q++) {
        const PrimC& P = prims[S.prim_off + q];
        PS s = load_ps(ps + (size_t)(S.prim_off + q) * 8);
        int m = prim_unit_count(P);
        PSg g = zero_psg();
        for (int j = 0; j < m; j++, col++) {
            float gd = valid ? gout[(size_t)col * ca.stride + slot] : 0.f;
            if (gd != 0.f) gx += unit_sdf_rev_x(P, prim_unit(P, j), s, x, gd, &g);
        }
        float vals[8] = {g.pos.x, g.pos.y, g.pos.z, g.rot.w, g.rot.x, g.rot.y, g.rot.z, g.gap};
        bool any = false;
#pragma unroll
        for (int k = 0; k < 8; k++) any |= (vals[k] != 0.f);
        float* wp = pmd ? pmd + ((size_t)((ch.y + (i & ~63)) >> 6) * MAX_UNITS + q) * 8 : nullptr;
        if (__ballot(any) == 0ull) {
            if (wp && lane < 8) wp[lane] = 0.f;
            continue;
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            float v = vals[k];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (wp) {
                if (lane == 0) wp[k] = v;
            } else if (lane == 0 && v != 0.f) {
                unsafeAtomicAdd(gps + (size_t)(S.prim_off + q) * 8 + k, v);
            }
        }
    }
    if (valid) store3(adj, ca.stride, 0, slot, load3(adj, ca.stride, 0, slot) + gx);
}

// Fixed-order mode of the distance adjoint: one wave per scene adds its
// particle waves' per-primitive sums in wave order (lane l takes waves l,
// l + 64, ..., then a fixed butterfly) into the primitive-state adjoint.
__global__ void __launch_bounds__(64) k_min_dist_reduce(const SceneC* __restrict__ scenes, const float* __restrict__ pmd,
                                                        float* __restrict__ gps) {
    const SceneC& S = scenes[blockIdx.x];
    int lane = threadIdx.x;
    int w0 = S.p_off >> 6, nw = (S.p_cnt + 63) >> 6;
    for (int q = 0; q < S.n_prims; q++) {
        float a[8];
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] = 0.f;
        for (int w = lane; w < nw; w += 64) {
            const float* r = pmd + ((size_t)(w0 + w) * MAX_UNITS + q) * 8;
#pragma unroll
            for (int k = 0; k < 8; k++) a[k] += r[k];
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) a[k] += __shfl_xor(a[k], o);
        if (lane == 0)
#pragma unroll
            for (int k = 0; k < 8; k++) gps[(size_t)(S.prim_off + q) * 8 + k] += a[k];
    }
}

// ------------------------------------------------------------------ tool-tool collision
// Pairs (i, j) of one scene: after forward kinematics primitive i is moved
// out of box j (set_surface_points / set_collision_idx /
// apply_collision_projection, mpm_simulator.py:286-305,316-320).
struct CollArgs {
    const int4* pairs;    // (i, j, 0, 0): absolute primitive indices
    const int2* groups;   // per scene with pairs: (first pair, pair count)
    const float4* rand;   // [pair][point] rand_num (xyz)
    int* idx;             // [frame][pair] collision_idx, -1: no contact
    float4* pt;           // [frame][pair] the chosen surface point
    float4* pre;          // [frame][pair] position of i before the pair's move
    int n_pairs, n_points;
};

// Primitive.sdf / normal of a whole primitive at world point x: a two-jaw
// gripper takes the smaller jaw SDF and the normal of jaw -1 when a <= b
// (primitives.py:486-496,618-628).  *js is the jaw receiving the SDF's
// adjoint (Taichi's min gives it to the left operand only when strictly
// smaller), *jn the jaw whose normal is used.
DEV float prim_sdf_jaws(const PrimC& P, const PS& s, v3 x, int* js, int* jn) {
    int m = prim_unit_count(P);
    float a = unit_sdf(P, prim_unit(P, 0), s, x);
    if (m == 1) { *js = *jn = 0; return a; }
    float b = unit_sdf(P, prim_unit(P, 1), s, x);
    *js = a < b ? 0 : 1;
    *jn = a <= b ? 0 : 1;
    return tmin(a, b);
}

// point k of pair p on box j: rand * size clamped to the box (Box._project,
// primitives.py:369-372) in j's frame (get_surface_pos, primive_base.py:284-287)
DEV v3 coll_surface_point(const CollArgs& ca, int p, int k, const PrimC& Pj, const PS& sj) {
    float4 r = ca.rand[(size_t)p * ca.n_points + k];
    v3 loc = mk(clampv(r.x * Pj.size[0], -Pj.size[0], Pj.size[0]), clampv(r.y * Pj.size[1], -Pj.size[1], Pj.size[1]),
                clampv(r.z * Pj.size[2], -Pj.size[2], Pj.size[2]));
    return qrot(sj.rot, loc) + sj.pos;
}

// One workgroup per scene with pairs.  Each pair's points are scored in
// parallel against the poses after kinematics; the collision index is the
// first point of the strictly smallest negative SDF (what the reference's
// loop computes run serially).  Lane 0 then applies the moves in pair order.
__global__ void __launch_bounds__(256) k_tool_collide(CollArgs ca, const PrimC* __restrict__ prims, float* __restrict__ pstate,
                                                      int n_prims_total, int f1) {
    int2 gr = ca.groups[blockIdx.x];
    float* ps1 = pstate + (size_t)f1 * n_prims_total * 8;
    size_t rec0 = (size_t)f1 * ca.n_pairs;
    __shared__ float s_d[4];
    __shared__ int s_k[4];
    int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int p = gr.x; p < gr.x + gr.y; p++) {
        int4 pr = ca.pairs[p];
        const PrimC& Pi = prims[pr.x];
        const PrimC& Pj = prims[pr.y];
        PS si = load_ps(ps1 + (size_t)pr.x * 8), sj = load_ps(ps1 + (size_t)pr.y * 8);
        float best = 0.f;
        int bk = INT_MAX;
        for (int k = t; k < ca.n_points; k += 256) {
            int js, jn;
            float d = prim_sdf_jaws(Pi, si, coll_surface_point(ca, p, k, Pj, sj), &js, &jn);
            if (d < best) { best = d; bk = k; }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            float od = __shfl_xor(best, o);
            int ok = __shfl_xor(bk, o);
            if (od < best || (od == best && ok < bk)) { best = od; bk = ok; }
        }
        if (lane == 0) { s_d[wv] = best; s_k[wv] = bk; }
        __syncthreads();
        if (t == 0) {
            for (int w = 1; w < 4; w++)
                if (s_d[w] < best || (s_d[w] == best && s_k[w] < bk)) { best = s_d[w]; bk = s_k[w]; }
            int k = bk == INT_MAX ? -1 : bk;
            v3 x = k >= 0 ? coll_surface_point(ca, p, k, Pj, sj) : mk(0.f, 0.f, 0.f);
            ca.idx[rec0 + p] = k;
            ca.pt[rec0 + p] = make_float4(x.x, x.y, x.z, 0.f);
        }
        __syncthreads();
    }
    if (t != 0) return;
    for (int p = gr.x; p < gr.x + gr.y; p++) {
        int i = ca.pairs[p].x;
        float* psi = ps1 + (size_t)i * 8;
        ca.pre[rec0 + p] = make_float4(psi[0], psi[1], psi[2], 0.f);
        if (ca.idx[rec0 + p] < 0) continue;
        const PrimC& Pi = prims[i];
        PS si = load_ps(psi);
        float4 x4 = ca.pt[rec0 + p];
        v3 x = mk(x4.x, x4.y, x4.z);
        int js, jn;
        float d = prim_sdf_jaws(Pi, si, x, &js, &jn);
        v3 n = unit_normal(Pi, prim_unit(Pi, jn), si, x);
        float L = sqrtf(dot(n, n));
        psi[0] += n.x / L * d;
        psi[1] += n.y / L * d;
        psi[2] += n.z / L * d;
    }
}

// Adjoint of the moves, pairs in reverse order, one lane per scene: with
// position_i' = position_i + n^ d, g_pos += J_pos^T g_pos and the rotation /
// gap of i get J^T g_pos.  J is taken at the pose each move started from
// (the recorded position; rotation and gap are not moved).  The points are
// constants (rand_points has no gradient, mpm_simulator.py:65).
__global__ void __launch_bounds__(64) k_tool_collide_bwd(CollArgs ca, const PrimC* __restrict__ prims,
                                                         const float* __restrict__ pstate, float* __restrict__ gpstate,
                                                         int n_prims_total, int f1) {
    if (threadIdx.x != 0) return;
    int2 gr = ca.groups[blockIdx.x];
    size_t rec0 = (size_t)f1 * ca.n_pairs;
    for (int p = gr.x + gr.y - 1; p >= gr.x; p--) {
        if (ca.idx[rec0 + p] < 0) continue;
        int i = ca.pairs[p].x;
        const PrimC& Pi = prims[i];
        size_t o = ((size_t)f1 * n_prims_total + i) * 8;
        PS s = load_ps(pstate + o);
        float4 q4 = ca.pre[rec0 + p], x4 = ca.pt[rec0 + p];
        s.pos = mk(q4.x, q4.y, q4.z);
        v3 x = mk(x4.x, x4.y, x4.z);
        int js, jn;
        float d = prim_sdf_jaws(Pi, s, x, &js, &jn);
        v3 n = unit_normal(Pi, prim_unit(Pi, jn), s, x);
        float L = sqrtf(dot(n, n));
        v3 nh = mk(n.x / L, n.y / L, n.z / L);
        float* gi = gpstate + o;
        v3 g = mk(gi[0], gi[1], gi[2]);
        v3 gnh = d * g;
        float tt = dot(nh, gnh);
        v3 gn = mk((gnh.x - nh.x * tt) / L, (gnh.y - nh.y * tt) / L, (gnh.z - nh.z * tt) / L);
        PSg acc = zero_psg();
        unit_sdf_rev(Pi, prim_unit(Pi, js), s, x, dot(nh, g), &acc);
        unit_normal_rev(Pi, prim_unit(Pi, jn), s, x, gn, &acc);
        gi[0] += acc.pos.x; gi[1] += acc.pos.y; gi[2] += acc.pos.z;
        gi[3] += acc.rot.w; gi[4] += acc.rot.x; gi[5] += acc.rot.y; gi[6] += acc.rot.z;
        gi[7] += acc.gap;
    }
}

__global__ void k_reset_frame0(int stride, const float* __restrict__ x, float* __restrict__ st) {
    int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= stride) return;
    for (int r = 0; r < 3; r++) st[(size_t)r * stride + slot] = x[(size_t)r * stride + slot];
    for (int r = 3; r < NST; r++) {
        int k = r - 15;
        st[(size_t)r * stride + slot] = (r >= 15 && (k == 0 || k == 4 || k == 8)) ? 1.f : 0.f;
    }
}
// GradModel._set_obs_grad (function.py:130-142)
__global__ void k_obs_grad(int stride, int n_prims_total, const float* __restrict__ xv, const float* __restrict__ pg,
                           float* __restrict__ adj, float* __restrict__ gps) {
    int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot < stride && xv) {
        for (int r = 0; r < 6; r++) adj[(size_t)r * stride + slot] += xv[(size_t)slot * 6 + r];
    }
    if (pg && slot < n_prims_total * 8) gps[slot] += pg[slot];
}
// set_action + set_velocity for env step s (primive_base.py:260-274)
__global__ void k_set_actions(const PrimC* __restrict__ prims, int n_prims_total, int s, int n_sub, const float* __restrict__ act,
                              float* __restrict__ pvel) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int q = t % max(n_prims_total, 1), j = t / max(n_prims_total, 1);
    if (n_prims_total == 0 || j >= n_sub) return;
    const PrimC& P = prims[q];
    float* o = pvel + ((size_t)(s * n_sub + j) * n_prims_total + q) * 8;
    if (P.action_dim <= 0) return;
    const float* a = act + q * 7;
    float ns = (float)n_sub;
    for (int k = 0; k < 3; k++) o[k] = a[k] * P.scale[k] / ns;
    if (P.action_dim > 3)
        for (int k = 3; k < 6; k++) o[k] = a[k] * P.scale[k] / ns;
    if (P.action_dim > 6) o[6] = a[6] * P.scale[6] / ns;
}
// set_velocity.grad + get_action_grad (function.py:180-190)
__global__ void k_action_grad(const PrimC* __restrict__ prims, int n_prims_total, int s, int n_sub, const float* __restrict__ gpvel,
                              float* __restrict__ out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_prims_total * 7) return;
    int q = t / 7, k = t % 7;
    const PrimC& P = prims[q];
    float acc = 0.f;
    bool used = (k < 3) || (k < 6 && P.action_dim > 3) || (k == 6 && P.action_dim > 6);
    if (P.action_dim > 0 && used)
        for (int j = 0; j < n_sub; j++) acc += gpvel[((size_t)(s * n_sub + j) * n_prims_total + q) * 8 + k];
    out[t] = (P.action_dim > 0 && used) ? acc * P.scale[k] / (float)n_sub : 0.f;
}
__global__ void k_fill(float* __restrict__ p, size_t n, float v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ------------------------------------------------------------------ slot re-sort
// Every R substeps the particles' slots are re-sorted into the cell order of
// the binning of the substep that ends at the boundary frame (a relabelling
// of slots within each scene: gathers and scatters of the tile kernels then
// touch whole cache lines again as the material deforms).  The new slot of a
// particle is its scene's first column + the particles of the scene's
// occupied blocks below its block (block-index order) + its position within
// its block's range of the sorted list: a function of the binning alone, so
// the fixed-order mode's binning gives a fixed-order re-sort.
// Pass 1, one workgroup per scene: exclusive sums of the occupied blocks'
// particle counts in block-index order.
__global__ void __launch_bounds__(1024) k_resort_blocks(BinArgs a, int* __restrict__ blk_pre) {
    __shared__ int s_w[16];
    const SceneC& S = a.scenes[blockIdx.x];
    int nbk = S.nb * S.nb * S.nb;
    int stamp = a.counters[CNT_STAMP];
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int run = 0;
    for (int b0 = 0; b0 < nbk; b0 += 1024) {
        int b = b0 + (int)threadIdx.x;
        int gb = S.blk_off + b;
        int c = (b < nbk && a.blk_stamp[gb] == stamp) ? a.blk_tot[gb] : 0;
        int incl = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            int t = __shfl_up(incl, d);
            if (lane >= d) incl += t;
        }
        if (lane == 63) s_w[wave] = incl;
        __syncthreads();
        int off = run;
        for (int w = 0; w < wave; w++) off += s_w[w];
        if (b < nbk) blk_pre[gb] = off + incl - c;
        for (int w = 0; w < 16; w++) run += s_w[w];
        __syncthreads();
    }
}
// Pass 2 over the work items: perm[slot] = new slot of the particle.
__global__ void __launch_bounds__(WG) k_resort_define(TileArgs ta, const int* __restrict__ cell_start, const int* __restrict__ blk_pre,
                                                      int* __restrict__ perm) {
    ItemIter ir = item_iter(ta);
    for (int j = ir.j; j < ir.n; j += ir.step) {
        int e;
        int4 it = item_at(ta, ir, j, e);
        const SceneC& S = ta.scenes[ta.blk_scene[it.x]];
        int base = S.p_off + blk_pre[it.x] - cell_start[it.x * 64];
        for (int k = threadIdx.x; k < it.z; k += WG) {
            int i = it.y + k;
            perm[ta.order[i]] = base + i;
        }
    }
}
// caller's view: map[c] (slot of frame 0's particle c) through perm
__global__ void k_resort_compose(int* __restrict__ map, const int* __restrict__ perm, const SceneC* __restrict__ scenes) {
    const SceneC& S = scenes[blockIdx.y];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < S.p_cnt; i += gridDim.x * blockDim.x) {
        int c = S.p_off + i;
        map[c] = perm[map[c]];
    }
}

}  // namespace

// ==================================================================== host side
struct mpm_ctx {
    int n_scenes = 0, n_prims = 0, max_frames = 0, K = 0, n_ckpt = 0, ring = 0;
    int stride = 0, n_chunks = 0, n_blocks = 0;
    long long n_nodes = 0, n_particles = 0;
    std::vector<SceneC> scenes;
    std::vector<PrimC> prims;
    std::vector<int> p_off;
    SceneC* d_scenes = nullptr;
    PrimC* d_prims = nullptr;
    int4* d_chunks = nullptr;
    int* d_blk_scene = nullptr;
    float* ckpt = nullptr;
    float* ringbuf = nullptr;
    float* aff = nullptr;   // affine matrix (9 rows) per frame slot: checkpoints then ring
    float* usv = nullptr;   // SVD of F_tmp (quaternion of U, sigma, quaternion of V: 11 rows) per frame slot, valid with aff
    float* gaff = nullptr;  // adjoint of the affine matrix (9 rows), one frame
    std::vector<int> aff_of; // frame whose affine matrix each slot holds (-1: none)
    float* adj[2] = {nullptr, nullptr};
    // dense grids (momentum/mass, velocity) of substep f at gpm_[f & 1],
    // gv_[f & 1], so that a rebuild of substep f can run while substep f+1's
    // grid adjoint still reads its own; one grid adjoint
    float4 *gpm_[2] = {nullptr, nullptr}, *gv_[2] = {nullptr, nullptr}, *g_adjpm = nullptr;
    float4* g_pm(int f) const { return gpm_[f & 1]; }
    float4* g_v(int f) const { return gv_[f & 1]; }
    // substep whose overflow fallback ran beside the adjoint of substep f+1
    // (-1: none pending)
    int fb_done = -1;
    // slot re-sort every `resort` substeps (0: off): frame s with s % resort
    // == 0 (0 < s < max_frames - 1) is written in the cell order of substep
    // s-1's binning; perm_of(s)[slot in frame s-1] = slot in frame s
    int resort = 0;
    int* rs_perm = nullptr;
    float* rs_fstage = nullptr;
    int *blk_tot = nullptr, *blk_pre = nullptr;
    bool is_bnd(int s) const { return resort > 0 && s > 0 && s % resort == 0 && s + 1 < max_frames; }
    int* perm_of(int s) const { return rs_perm + (size_t)(s / resort - 1) * stride; }
    int epoch_of(int f) const { return resort > 0 ? std::max(0, std::min(f, max_frames - 2) / resort) : 0; }
    int *blk_flag = nullptr, *node_flag = nullptr, *pblocks = nullptr, *gstamp = nullptr;
    int *cell_count = nullptr, *cell_start = nullptr, *bin_q = nullptr, *bin_rank = nullptr, *blk_rep = nullptr;
    // Bin slots: the binning of a substep (index list, work items, per-block
    // item ranges and stamps, active node-block list and map, counters) and,
    // when the grid is kept, that substep's grid (momentum/mass and velocity of
    // its active node blocks).  With keeping on there is one bin slot per frame
    // slot, so the adjoint of a substep the forward ran reuses both instead of
    // binning and rebuilding the grid again; otherwise one shared slot.
    struct BinSlot {
        int* order; int4* items; int* blk_stamp; int* blk_item0; int* blk_nitems; int* nblocks; int* nmap; int* counters; int* nrec; int4* xlist;
        float4* keep_pm; float4* keep_v;
    };
    std::vector<BinSlot> bslots;
    std::vector<int> kept_of;  // substep whose bins + grid each bin slot holds (-1: none)
    bool keep = false;
    int keep_cap = 0;          // node blocks a kept grid can hold
    int rec_cap = 0;           // node-block records per bin slot
    float4 *tiles_pm = nullptr, *tiles_v = nullptr;
    int max_items = 0;
    int wi = WI;          // particles per work item
    int tile_ctas = 1024; // persistent work-item workgroups
    float *pstate = nullptr, *pstate_g = nullptr, *pvel = nullptr, *pvel_g = nullptr;
    int valid_hi = -1;    // frames > valid_hi are stale
    int seg_valid = -1;   // segment whose ring frames are resident
    int seg_lo = -1;      // lowest resident ring frame of that segment
    int seg_hi = -1;      // highest resident ring frame of that segment
    int adj_top = -1;     // frame whose adjoint is current
    int grid_ctas = 1024; // persistent grid-phase workgroups
    // fixed-order mode (mpm_set_deterministic, on by default): slot-ordered
    // cell ranks, block-ordered primitive adjoints, wave-ordered distance
    // adjoints, integer mass grid
    bool det = true;
    int *cell_gcnt = nullptr, *cell_ng = nullptr;
    unsigned* cell_grp = nullptr;
    unsigned long long* cell_head = nullptr;
    int *grp_next = nullptr, *grp_cnt = nullptr;
    float* padj = nullptr;
    unsigned* pbits = nullptr;
    unsigned long long* macc = nullptr;
    float* pmd = nullptr;
    // the constitutive update of a forward substep does not depend on its
    // binning: it runs on a second stream (default priority) between an
    // event fork and join, so the two overlap (parallel branches in a graph)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // tool-tool collision pairs (mpm_set_collisions); tables and per-frame
    // records live in a caller-owned device buffer
    int n_pairs = 0, n_cgroups = 0;
    CollArgs coll{};
    // per-kernel timing with HIP events on the launch stream (mpm_set_profiling)
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    struct Rec { int phase; hipEvent_t a, b; };
    std::vector<Rec> recs;
    double prof_ms[MPM_N_PHASES] = {0};
    long long prof_n[MPM_N_PHASES] = {0};

    size_t frame_floats() const { return (size_t)NST * stride; }
    float* frame_ptr(int f) const {
        if (f % K == 0) return ckpt + (size_t)(f / K) * frame_floats();
        return ringbuf + (size_t)((f % K) - 1) * frame_floats();
    }
    float* adj_ptr(int f) const { return adj[f & 1]; }
    int slot_of(int f) const { return (f % K == 0) ? f / K : n_ckpt + (f % K) - 1; }
    float* aff_ptr(int f) const { return aff + (size_t)slot_of(f) * 9 * stride; }
    float* usv_ptr(int f) const { return usv + (size_t)slot_of(f) * USV_ROWS * stride; }
    int bslot_of(int f) const { return keep ? slot_of(f) : 0; }
    void drop_kept() { for (auto& k : kept_of) k = -1; fb_done = -1; }
    // frame f's state is in memory: a checkpoint, or a ring frame of the
    // resident segment
    bool frame_resident(int f) const {
        if (f < 0 || f > valid_hi) return false;
        if (f % K == 0) return true;
        return seg_valid == f / K && seg_lo <= f && f <= seg_hi;
    }
    ChunkArgs cargs() const { ChunkArgs a; a.scenes = d_scenes; a.chunks = d_chunks; a.n_chunks = n_chunks; a.stride = stride; return a; }
    BinArgs bargs(int bs) const {
        const BinSlot& b = bslots[bs];
        BinArgs a;
        a.scenes = d_scenes; a.chunks = d_chunks; a.blk_scene = d_blk_scene; a.stride = stride;
        a.cell_count = cell_count; a.cell_start = cell_start; a.bin_q = bin_q; a.bin_rank = bin_rank;
        a.blk_flag = blk_flag; a.node_flag = node_flag; a.counters = b.counters; a.gstamp = gstamp; a.pblocks = pblocks;
        a.blk_stamp = b.blk_stamp; a.blk_item0 = b.blk_item0; a.blk_nitems = b.blk_nitems; a.items = b.items; a.order = b.order;
        a.wi = wi;
        a.det = det ? 1 : 0;
        a.cell_gcnt = cell_gcnt; a.cell_ng = cell_ng; a.cell_grp = cell_grp;
        a.cell_head = cell_head; a.grp_next = grp_next; a.grp_cnt = grp_cnt;
        a.blk_rep = blk_rep; a.xlist = b.xlist; a.blk_tot = blk_tot;
        return a;
    }
    TileArgs targs(int bs) const {
        const BinSlot& b = bslots[bs];
        TileArgs t;
        t.scenes = d_scenes; t.blk_scene = d_blk_scene; t.items = b.items; t.counters = b.counters; t.order = b.order; t.stride = stride;
        t.keep_v = b.keep_v; t.nmap = b.nmap; t.keep_cap = keep_cap;
        t.xlist = b.xlist;
        return t;
    }
    // write_keep: the forward grid_op also stores the grid into the bin slot
    GridArgs gargs(int f, int bs, bool write_keep) const {
        const BinSlot& b = bslots[bs];
        GridArgs g;
        g.scenes = d_scenes; g.prims = d_prims; g.blk_scene = d_blk_scene; g.nblocks = b.nblocks; g.counters = b.counters;
        g.blk_stamp = b.blk_stamp; g.blk_item0 = b.blk_item0; g.blk_nitems = b.blk_nitems;
        g.nrec = b.nrec; g.rec_cap = rec_cap;
        g.pstate = pstate; g.n_prims_total = n_prims; g.f = f;
        g.keep_pm = b.keep_pm;
        g.keep_v = b.keep_v;
        g.keep_cap = b.keep_pm ? keep_cap : 0;
        g.keep_write = (write_keep && b.keep_pm) ? 1 : 0;
        g.det = det ? 1 : 0;
        g.padj = padj;
        g.pbits = pbits;

        return g;
    }
};

namespace {

hipEvent_t take_event(mpm_ctx* c) {
    if (!c->ev_pool.empty()) { hipEvent_t e = c->ev_pool.back(); c->ev_pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
// wraps one kernel launch in a pair of events when profiling is on
struct PhaseTimer {
    mpm_ctx* c; int phase; hipStream_t st; hipEvent_t a = nullptr, b = nullptr;
    PhaseTimer(mpm_ctx* c_, int ph, hipStream_t s) : c(c_), phase(ph), st(s) {
        if (c->prof) { a = take_event(c); b = take_event(c); if (a) hipEventRecord(a, st); }
    }
    ~PhaseTimer() {
        if (c->prof && a && b) { hipEventRecord(b, st); c->recs.push_back({phase, a, b}); }
    }
};

struct Layout {
    size_t off_scenes, off_prims, off_chunks, off_blk_scene, off_ckpt, off_ring, off_aff, off_usv, off_adj, off_grid, off_blkints, off_cells,
        off_pints, off_bslots, off_keep, off_tiles, off_counters, off_prim_bufs, off_det, total;
    // fixed-order mode buffers (within off_det): cell list heads, group
    // records, per-block primitive adjoints, block bitmaps, integer mass grid
    size_t det_gn, det_garr, det_head, det_grp, det_padj, det_pbits, det_macc, det_pmd;

    int n_bm_words;
    long long n_mass;
    // one bin slot: order, items, blk_stamp, blk_item0, blk_nitems, nblocks, nmap, counters
    size_t bs_order, bs_items, bs_stamp, bs_item0, bs_nitems, bs_nblocks, bs_nmap, bs_counters, bs_nrec, bs_xlist, bs_bytes, keep_bytes;
    int stride, n_chunks, n_blocks, n_ckpt, ring, K, max_items, n_bslots, keep_cap, rec_cap;
    long long n_nodes;
    int n_prims;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// keep_blocks: < 0 default capacity, 0 no kept grids, > 0 capacity in node blocks
Layout make_layout(const mpm_scene_desc* sc, int n, int max_frames, int ckpt_k, int keep_blocks) {
    Layout L;
    memset(&L, 0, sizeof(L));
    int stride = 0, n_chunks = 0, n_blocks = 0, n_prims = 0, bm_words = 0;
    long long nodes = 0, mass = 0;
    for (int i = 0; i < n; i++) {
        int np = std::max(sc[i].n_particles, 0);
        stride += (np + 63) & ~63;
        n_chunks += (np + PPB - 1) / PPB;
        int nb = (sc[i].n_grid + 3) / 4;
        n_blocks += nb * nb * nb;
        bm_words += (nb * nb * nb + 31) / 32;
        nodes += (long long)nb * nb * nb * NODES_PER_BLK;
        mass += (long long)sc[i].n_grid * sc[i].n_grid * sc[i].n_grid;
        n_prims += sc[i].n_prims;
    }
    stride = std::max(64, (stride + 255) & ~255);
    int K = (ckpt_k <= 0 || ckpt_k >= max_frames) ? max_frames : ckpt_k;
    L.K = K;
    L.n_ckpt = (max_frames - 1) / K + 1;
    L.ring = std::max(1, std::min(K - 1, max_frames - 1));
    L.stride = stride;
    L.n_chunks = std::max(1, n_chunks);
    L.n_blocks = std::max(1, n_blocks);
    L.n_nodes = std::max(64LL, nodes);
    L.n_prims = n_prims;
    size_t o = 0;
    L.off_scenes = o; o = align256(o + sizeof(SceneC) * std::max(1, n));
    L.off_prims = o; o = align256(o + sizeof(PrimC) * std::max(1, n_prims));
    L.off_chunks = o; o = align256(o + sizeof(int4) * L.n_chunks);
    L.off_blk_scene = o; o = align256(o + sizeof(int) * L.n_blocks);
    size_t ff = (size_t)NST * stride * sizeof(float);
    L.off_ckpt = o; o = align256(o + ff * L.n_ckpt);
    L.off_ring = o; o = align256(o + ff * L.ring);
    L.off_aff = o; o = align256(o + sizeof(float) * 9 * (size_t)stride * (L.n_ckpt + L.ring + 1));
    L.off_usv = o; o = align256(o + sizeof(float) * USV_ROWS * (size_t)stride * (L.n_ckpt + L.ring));
    L.off_adj = o; o = align256(o + ff * 2);
    L.off_grid = o; o = align256(o + sizeof(float4) * (size_t)L.n_nodes * 5);
    L.off_blkints = o; o = align256(o + sizeof(int) * (size_t)L.n_blocks * 4);   // block flags, node flags, occupied list, representative slots
    L.off_cells = o; o = align256(o + sizeof(int) * (size_t)L.n_blocks * 64 * 2); // cell counts, cell starts
    L.off_pints = o; o = align256(o + sizeof(int) * (size_t)stride * 2);         // bin key, bin rank
    // work items: at most one partial item per occupied cell block plus
    // stride / wi full ones, for the smallest allowed wi
    L.max_items = stride / WI_MIN + std::min(L.n_blocks, stride) + 1;
    // default kept-grid capacity: generous for particle clouds of at least a few
    // hundred particles per occupied cell block; a pass whose active node-block
    // list does not fit falls back to recomputing its grid in the adjoint
    int cap = keep_blocks < 0 ? stride / 64 + 64 * n : keep_blocks;
    L.keep_cap = std::min(cap, L.n_blocks);
    // node-block records: the same generous count, or a caller-given kept-grid
    // capacity (passes with more active node blocks look the rest up from the
    // block stamps, the path a small capacity exercises)
    L.rec_cap = std::min(keep_blocks > 0 ? keep_blocks : stride / 64 + 64 * n, L.n_blocks);
    L.n_bslots = L.keep_cap > 0 ? L.n_ckpt + L.ring : 1;
    size_t b = 0;
    L.bs_order = b; b = align256(b + sizeof(int) * (size_t)stride);
    L.bs_items = b; b = align256(b + sizeof(int4) * (size_t)L.max_items);
    L.bs_stamp = b; b = align256(b + sizeof(int) * (size_t)L.n_blocks);
    L.bs_item0 = b; b = align256(b + sizeof(int) * (size_t)L.n_blocks);
    L.bs_nitems = b; b = align256(b + sizeof(int) * (size_t)L.n_blocks);
    L.bs_nblocks = b; b = align256(b + sizeof(int) * (size_t)L.n_blocks);
    L.bs_nmap = b; b = align256(b + sizeof(int) * (size_t)L.n_blocks);
    L.bs_counters = b; b = align256(b + sizeof(int) * 64);
    L.bs_nrec = b; b = align256(b + sizeof(int) * NREC * (size_t)L.rec_cap);
    L.bs_xlist = b; b = align256(b + sizeof(int4) * (size_t)L.max_items);
    L.bs_bytes = b;
    L.keep_bytes = (size_t)L.keep_cap * NODES_PER_BLK * sizeof(float4) * 2;
    L.off_bslots = o; o = align256(o + L.bs_bytes * L.n_bslots);
    L.off_keep = o; o = align256(o + L.keep_bytes * L.n_bslots);
    L.off_tiles = o; o = align256(o + sizeof(float4) * (size_t)L.max_items * TILE * 2);
    L.off_counters = o; o = align256(o + sizeof(int) * 64);
    L.off_prim_bufs = o; o = align256(o + sizeof(float) * 8 * (size_t)max_frames * std::max(1, n_prims) * 4);
    L.n_bm_words = std::max(1, bm_words);
    L.n_mass = std::max(1LL, mass);
    L.off_det = o;
    size_t d = 0;
    L.det_gn = d; d = align256(d + sizeof(int) * 2 * (size_t)L.n_blocks * 64);
    L.det_garr = d; d = align256(d + sizeof(unsigned) * DET_GMAX * (size_t)L.n_blocks * 64);
    L.det_head = d; d = align256(d + sizeof(unsigned long long) * (size_t)L.n_blocks * 64);
    L.det_grp = d; d = align256(d + sizeof(int) * (size_t)stride * 2);
    L.det_padj = d; d = align256(d + sizeof(float) * 16 * MAX_UNITS * (size_t)L.n_blocks);
    L.det_pbits = d; d = align256(d + sizeof(unsigned) * (size_t)L.n_bm_words);
    L.det_macc = d; d = align256(d + sizeof(unsigned long long) * (size_t)L.n_mass);
    // one record per particle wave of any chunk: a chunk's last waves may
    // lie past the stride (its first slot plus 256)
    L.det_pmd = d; d = align256(d + sizeof(float) * 8 * MAX_UNITS * (size_t)(stride / 64 + PPB / 64));
    o = align256(o + d);
    L.total = o;
    return L;
}

hipStream_t S_(void* s) { return reinterpret_cast<hipStream_t>(s); }

int launch_fill(float* p, size_t n, hipStream_t st) {
    if (n == 0) return 0;
    int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, st, p, n, 0.f);
    HIPCHK(hipGetLastError());
    return 0;
}

// counting-sort frame f's particles into cell blocks and work items
// (stage 1: cell keys and ranks; stage 2: blocks, items, the sorted list)
int run_bin_keys(mpm_ctx* c, int f, int do_fk, hipStream_t st) {
    BinArgs ba = c->bargs(c->bslot_of(f));
    { PhaseTimer t(c, MPM_PH_BIN, st);
      hipLaunchKernelGGL(k_bin, dim3(c->n_chunks), dim3(PPB), 0, st, ba, (const float*)c->frame_ptr(f), do_fk, c->d_prims, c->n_prims,
                         c->pstate, c->pvel, f); }
    HIPCHK(hipGetLastError());
    return 0;
}
int run_bin_lists(mpm_ctx* c, int f, hipStream_t st) {
    int bs = c->bslot_of(f);
    const mpm_ctx::BinSlot& B = c->bslots[bs];
    BinArgs ba = c->bargs(bs);
    { PhaseTimer t(c, MPM_PH_BLOCKS, st);
      hipLaunchKernelGGL(k_blocks, dim3(c->grid_ctas), dim3(256), 0, st, ba, B.nblocks, B.nmap, B.nrec, c->rec_cap); }
    { PhaseTimer t(c, MPM_PH_SCATTER, st);
      hipLaunchKernelGGL(k_scatter, dim3(c->n_chunks), dim3(PPB), 0, st, ba, B.nrec, c->rec_cap); }
    HIPCHK(hipGetLastError());
    return 0;
}
int run_bin(mpm_ctx* c, int f, int do_fk, hipStream_t st) {
    if (run_bin_keys(c, f, do_fk, st)) return -1;
    return run_bin_lists(c, f, st);
}

// forward substep f -> f+1 (writes frame f+1 wherever its slot is).  When
// frame f+1 is a re-sort boundary, its slot order is defined from this
// substep's binning, or, recomputing (a checkpoint segment), the order the
// first pass defined is applied again.
int run_fwd(mpm_ctx* c, int f, hipStream_t st, bool recompute = false) {
    const float* in = c->frame_ptr(f);
    float* out = c->frame_ptr(f + 1);
    bool bnd = c->is_bnd(f + 1);
    int bs = c->bslot_of(f);
    TileArgs ta = c->targs(bs);
    float* aff = c->aff_ptr(f);
    // The binning chain is the longer branch: its first kernel runs alone,
    // then the constitutive update (which needs no binning) forks onto the
    // second stream beside the rest of the chain.  Profiling serialises the
    // two branches so that every kernel's event interval is its own.  A
    // small context runs them in line too: its kernels are a few µs each,
    // less than the fork and join cost in a graph (a 5k-particle scene:
    // 43.2 against 47.4 ms per 500-substep forward pass).
    bool fork = !c->prof && c->n_particles >= FORK_MIN_PARTICLES;
    hipStream_t sst = fork ? c->side : st;
    if (run_bin_keys(c, f, 1, st)) return -1;
    if (fork) {
        HIPCHK(hipEventRecord(c->ev_fork, st));
        HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    }
    { PhaseTimer t(c, MPM_PH_STRESS, sst);
      hipLaunchKernelGGL(k_stress, dim3(c->n_chunks), dim3(PPB), 0, sst, c->cargs(), in, bnd ? c->rs_fstage : out, aff, c->usv_ptr(f)); }
    if (fork) HIPCHK(hipEventRecord(c->ev_join, c->side));
    c->aff_of[c->slot_of(f)] = f;
    if (run_bin_lists(c, f, st)) return -1;
    if (bnd && !recompute) {
        hipLaunchKernelGGL(k_resort_blocks, dim3(c->n_scenes), dim3(1024), 0, st, c->bargs(bs), c->blk_pre);
        hipLaunchKernelGGL(k_resort_define, dim3(c->tile_ctas), dim3(WG), 0, st, ta, (const int*)c->cell_start, (const int*)c->blk_pre,
                           c->perm_of(f + 1));
    }
    if (c->n_pairs > 0) {
        PhaseTimer t(c, MPM_PH_COLLIDE, st);
        hipLaunchKernelGGL(k_tool_collide, dim3(c->n_cgroups), dim3(256), 0, st, c->coll, c->d_prims, c->pstate, c->n_prims, f + 1);
    }
    if (fork) HIPCHK(hipStreamWaitEvent(st, c->ev_join, 0));
    { PhaseTimer t(c, MPM_PH_P2G, st);
      hipLaunchKernelGGL(k_p2g_scatter, dim3(c->tile_ctas), dim3(WG), 0, st, ta, in, (const float*)aff, c->tiles_pm); }
    { PhaseTimer t(c, MPM_PH_GRID, st);
      hipLaunchKernelGGL(k_grid_op, dim3(c->grid_ctas), dim3(128), 0, st, c->gargs(f, bs, c->keep), c->tiles_pm, c->g_pm(f),
                         c->g_v(f)); }
    { PhaseTimer t(c, MPM_PH_G2P, st);
      hipLaunchKernelGGL(k_g2p_tile, dim3(c->tile_ctas), dim3(WG), 0, st, ta, in, out, c->g_v(f),
                         bnd ? (const int*)c->perm_of(f + 1) : (const int*)nullptr, (const float*)c->rs_fstage); }
    HIPCHK(hipGetLastError());
    c->kept_of[bs] = c->keep ? f : -1;
    c->fb_done = -1;   // the forward reuses the tile buffer and the dense grids
    return 0;
}

int ensure_frame(mpm_ctx* c, int f, hipStream_t st) {
    if (f < 0 || f >= c->max_frames) return fail("frame " + std::to_string(f) + " outside [0, max_frames)");
    if (f > c->valid_hi) return fail("frame " + std::to_string(f) + " has not been simulated");
    if (f % c->K == 0) return 0;
    int a = f / c->K;
    if (c->seg_valid == a && c->seg_lo <= f && f <= c->seg_hi) return 0;
    // recompute the segment from its checkpoint (never rewrites a checkpoint)
    int start = a * c->K;
    for (int s = start; s < f; s++)
        if (run_fwd(c, s, st, true)) return -1;
    c->seg_valid = a;
    c->seg_lo = start + 1;
    c->seg_hi = f;
    return 0;
}

// the kept-grid overflow fallbacks of substep f (p2g then grid_op, each a
// no-op unless the pass's active node blocks outgrew the kept grid)
int launch_fallbacks(mpm_ctx* c, int f, hipStream_t s) {
    int bs = c->bslot_of(f);
    TileArgs ta = c->targs(bs);
    GridArgs ga = c->gargs(f, bs, false);
    { PhaseTimer t(c, MPM_PH_P2G_GRID, s);
      hipLaunchKernelGGL(k_p2g_scatter_rebuild, dim3(64), dim3(WG), 0, s, ta, (const float*)c->frame_ptr(f), (const float*)c->aff_ptr(f),
                         c->tiles_pm, 1); }
    { PhaseTimer t(c, MPM_PH_GRID_KEEP, s);
      hipLaunchKernelGGL(k_grid_op_rebuild, dim3(256), dim3(128), 0, s, ga, c->tiles_pm, c->g_pm(f), c->g_v(f), 1); }
    HIPCHK(hipGetLastError());
    return 0;
}

int run_bwd(mpm_ctx* c, int f, hipStream_t st) {
    if (ensure_frame(c, f + 1, st)) return -1;
    if (ensure_frame(c, f, st)) return -1;
    const float* in = c->frame_ptr(f);
    const float* in1 = c->frame_ptr(f + 1);
    const float* a1 = c->adj_ptr(f + 1);
    // frame f+1 re-sorted: its state and adjoint read through the permutation
    const int* perm1 = c->is_bnd(f + 1) ? c->perm_of(f + 1) : nullptr;
    float* a0 = c->adj_ptr(f);
    int bs = c->bslot_of(f);
    // bins and grid of substep f kept from its forward run, or rebuilt here
    bool kept = c->keep && c->kept_of[bs] == f;
    if (!kept) {
        c->kept_of[bs] = -1;
        if (run_bin(c, f, 0, st)) return -1;
    }
    TileArgs ta = c->targs(bs);
    GridArgs ga = c->gargs(f, bs, false);
    float* aff = c->aff_ptr(f);
    if (c->aff_of[c->slot_of(f)] != f) {
        // the affine matrix and SVD of frame f are not stored (recomputed segment ended at f)
        PhaseTimer t(c, MPM_PH_STRESS, st);
        hipLaunchKernelGGL(k_stress, dim3(c->n_chunks), dim3(PPB), 0, st, c->cargs(), in, (float*)nullptr, aff, c->usv_ptr(f));
        c->aff_of[c->slot_of(f)] = f;
    }
    // rebuild the grid when it was not kept; a kept grid whose active list
    // overflowed the capacity is rebuilt on the device (only_overflow launches,
    // small persistent grids so that their no-op case dispatches quickly).
    // They stay on this stream: moved to the side stream (overlapping the rest
    // of substep f+1) the graph's cross-stream fork and join cost more idle
    // time than the two no-op launches themselves (80.4 vs 79.2 ms per pass).
    if (!kept) {
        { PhaseTimer t(c, MPM_PH_P2G_GRID, st);
          hipLaunchKernelGGL(k_p2g_scatter_rebuild, dim3(c->tile_ctas), dim3(WG), 0, st, ta, in, (const float*)aff, c->tiles_pm, 0); }
        { PhaseTimer t(c, MPM_PH_GRID_KEEP, st);
          hipLaunchKernelGGL(k_grid_op_rebuild, dim3(c->grid_ctas), dim3(128), 0, st, ga, c->tiles_pm, c->g_pm(f), c->g_v(f), 0); }
    } else if (c->keep_cap < c->n_blocks && c->fb_done != f) {
        if (launch_fallbacks(c, f, st)) return -1;
    }
    c->fb_done = -1;
    // the next adjoint substep's overflow fallback rides on this one's g2p and
    // grid adjoint launches when its binning, grid and frame are already here
    // (and not under profiling, which times every launch on its own)
    int g = f - 1;
    bool fuse = g >= 0 && !c->prof && c->keep && c->keep_cap < c->n_blocks && c->kept_of[c->bslot_of(g)] == g &&
                c->frame_resident(g) && c->aff_of[c->slot_of(g)] == g;
    P2GRebuild prb{};
    GridRebuild grb{};
    if (fuse) {
        int bg = c->bslot_of(g);
        prb = P2GRebuild{c->targs(bg), c->frame_ptr(g), c->aff_ptr(g), c->tiles_pm, FB_LEAD};
        // (the grid adjoint runs one-wave workgroups: 4 * FB_LEAD of them lead)
        grb = GridRebuild{c->gargs(g, bg, false), c->tiles_pm, c->g_pm(g), c->g_v(g), 4 * FB_LEAD};
    }
    int use_kept = kept ? 1 : 0;
    { PhaseTimer t(c, MPM_PH_G2P_BWD, st);
      hipLaunchKernelGGL(k_g2p_bwd_tile, dim3(c->tile_ctas + prb.lead), dim3(WG), 0, st, ta, in, in1, a1, a0, c->g_v(f), c->tiles_v,
                         use_kept, prb, perm1); }
    { PhaseTimer t(c, MPM_PH_GRID_BWD, st);
      hipLaunchKernelGGL(k_grid_op_bwd, dim3(2 * c->grid_ctas + grb.lead), dim3(64), 0, st, ga, c->g_pm(f), c->tiles_v, c->g_adjpm,
                         c->pstate_g, use_kept, grb); }
    // fixed-order mode: the primitive adjoints are reduced by their own launch
    // when the collision adjoint needs them, else by the p2g adjoint's leading
    // workgroups (a multiple of NXCD, so the item workgroups keep their XCDs)
    PrimRedArgs pr{c->d_scenes, c->d_prims, (const float*)c->padj, c->pbits, c->pstate_g, c->n_prims, f, c->n_scenes};
    int pr_lead = 0;
    if (c->det && c->n_prims > 0) {
        if (c->n_pairs > 0) {
            PhaseTimer t(c, MPM_PH_PRIM_REDUCE, st);
            hipLaunchKernelGGL(k_prim_adj_reduce, dim3(c->n_scenes), dim3(64), 0, st, pr);
        } else {
            pr_lead = (c->n_scenes + NXCD - 1) / NXCD * NXCD;
        }
    }
    if (c->n_pairs > 0) {
        // between the grid adjoint (which adds into frame f+1's primitive
        // adjoint) and the kinematics adjoint (which consumes it)
        PhaseTimer t(c, MPM_PH_COLLIDE_BWD, st);
        hipLaunchKernelGGL(k_tool_collide_bwd, dim3(c->n_cgroups), dim3(64), 0, st, c->coll, c->d_prims, c->pstate, c->pstate_g,
                           c->n_prims, f + 1);
    }
    { PhaseTimer t(c, MPM_PH_P2G_BWD, st);
      hipLaunchKernelGGL(k_p2g_bwd_gather, dim3(c->tile_ctas + pr_lead), dim3(WG), 0, st, ta, in, (const float*)aff, a0, c->gaff,
                         c->g_adjpm, pr, pr_lead); }
    { PhaseTimer t(c, MPM_PH_STRESS_BWD, st);
      hipLaunchKernelGGL(k_stress_bwd, dim3(c->n_chunks), dim3(PPB), 0, st, c->cargs(), in, (const float*)c->gaff, a1, a0,
                         (const float*)c->usv_ptr(f), c->d_prims, c->n_prims, c->pstate, c->pvel, c->pstate_g, c->pvel_g, f, perm1); }
    HIPCHK(hipGetLastError());
    if (fuse) c->fb_done = g;
    return 0;
}

PrimC to_primc(const mpm_prim_desc& d) {
    PrimC p;
    memset(&p, 0, sizeof(p));
    p.type = d.type;
    p.action_dim = d.action_dim;
    switch (d.type) {
    case MPM_SPHERE: p.unit_shape = SH_SPHERE; break;
    case MPM_CAPSULE: case MPM_ROLLINGPIN: case MPM_ROLLINGPIN_EXT: case MPM_GRIPPER2: p.unit_shape = SH_CAPSULE; break;
    case MPM_BOX: case MPM_GRIPPER: p.unit_shape = SH_BOX; break;
    case MPM_PRISM: p.unit_shape = SH_PRISM; break;
    case MPM_KNIFE: p.unit_shape = SH_KNIFE; break;
    case MPM_CHOPSTICKS: p.unit_shape = SH_CHOP; break;
    case MPM_CYLINDER: p.unit_shape = SH_CYL; break;
    case MPM_TORUS: p.unit_shape = SH_TORUS; break;
    default: p.unit_shape = SH_BOX;
    }
    p.n_units = (d.type == MPM_GRIPPER || d.type == MPM_GRIPPER2) ? 2 : 1;
    p.friction = (float)d.friction;
    p.softness = (float)d.softness;
    p.h = (float)d.h;
    p.r = (float)d.r;
    for (int i = 0; i < 3; i++) { p.size[i] = (float)d.size[i]; p.lo[i] = (float)d.lo[i]; p.hi[i] = (float)d.hi[i]; }
    for (int i = 0; i < 2; i++) p.ph[i] = (float)d.ph[i];
    for (int i = 0; i < 4; i++) p.prot[i] = (float)d.prot[i];
    p.min_gap = (float)d.min_gap;
    p.max_gap = (float)d.max_gap;
    for (int i = 0; i < 7; i++) p.scale[i] = (float)d.action_scale[i];
    return p;
}

}  // namespace

extern "C" {

const char* mpm_last_error(void) { return g_err.c_str(); }

int64_t mpm_workspace_bytes(const mpm_scene_desc* scenes, int n_scenes, int max_frames, int ckpt_k, int keep_blocks) {
    if (!scenes || n_scenes <= 0 || max_frames < 2) return -1;
    return (int64_t)make_layout(scenes, n_scenes, max_frames, ckpt_k, keep_blocks).total;
}

int mpm_create(const mpm_scene_desc* scenes, int n_scenes, const mpm_prim_desc* prims, int max_frames, int ckpt_k, int keep_blocks,
               void* workspace, int64_t workspace_bytes, void* stream, mpm_ctx** out) {
    if (!out) return fail("mpm_create: out is null");
    *out = nullptr;
    if (!scenes || n_scenes <= 0) return fail("mpm_create: no scenes");
    if (max_frames < 2) return fail("mpm_create: max_frames must be >= 2");
    for (int i = 0; i < n_scenes; i++) {
        if (scenes[i].n_grid < 8 || scenes[i].n_grid > 1024) return fail("mpm_create: n_grid outside [8, 1024]");
        if (scenes[i].n_particles < 0) return fail("mpm_create: negative particle count");
        if (scenes[i].n_prims < 0) return fail("mpm_create: negative primitive count");
    }
    if (!prims) {
        for (int i = 0; i < n_scenes; i++)
            if (scenes[i].n_prims > 0) return fail("mpm_create: primitives are null");
    }
    // every collision unit of a scene (one per primitive, two for the
    // two-jaw grippers) carries its pre-collide velocity through the grid
    // adjoint in registers, MAX_UNITS of them
    for (int i = 0, q = 0; i < n_scenes; q += scenes[i].n_prims, i++) {
        int units = 0;
        for (int j = 0; j < scenes[i].n_prims; j++) {
            int t = prims[q + j].type;
            if (t < MPM_SPHERE || t > MPM_TORUS) return fail("mpm_create: unknown primitive type " + std::to_string(t));
            units += (t == MPM_GRIPPER || t == MPM_GRIPPER2) ? 2 : 1;
        }
        if (units > MAX_UNITS)
            return fail("mpm_create: scene " + std::to_string(i) + " has " + std::to_string(units) +
                        " collision units (grippers count two); at most " + std::to_string(MAX_UNITS));
    }
    Layout L = make_layout(scenes, n_scenes, max_frames, ckpt_k, keep_blocks);
    if (!workspace || workspace_bytes < (int64_t)L.total)
        return fail("mpm_create: workspace smaller than mpm_workspace_bytes()");
    hipStream_t st = S_(stream);
    mpm_ctx* c = new mpm_ctx();
    c->n_scenes = n_scenes;
    c->n_prims = L.n_prims;
    c->max_frames = max_frames;
    c->K = L.K;
    c->n_ckpt = L.n_ckpt;
    c->ring = L.ring;
    c->stride = L.stride;
    c->n_chunks = L.n_chunks;
    c->n_blocks = L.n_blocks;
    c->n_nodes = L.n_nodes;
    char* w = static_cast<char*>(workspace);
    c->d_scenes = (SceneC*)(w + L.off_scenes);
    c->d_prims = (PrimC*)(w + L.off_prims);
    c->d_chunks = (int4*)(w + L.off_chunks);
    c->d_blk_scene = (int*)(w + L.off_blk_scene);
    c->ckpt = (float*)(w + L.off_ckpt);
    c->ringbuf = (float*)(w + L.off_ring);
    c->aff = (float*)(w + L.off_aff);
    c->gaff = c->aff + (size_t)9 * L.stride * (L.n_ckpt + L.ring);
    c->aff_of.assign(L.n_ckpt + L.ring, -1);
    c->usv = (float*)(w + L.off_usv);
    c->adj[0] = (float*)(w + L.off_adj);
    c->adj[1] = c->adj[0] + c->frame_floats();
    float4* g = (float4*)(w + L.off_grid);
    c->gpm_[0] = g;
    c->gv_[0] = g + L.n_nodes;
    c->g_adjpm = g + 2 * L.n_nodes;
    c->gpm_[1] = g + 3 * L.n_nodes;
    c->gv_[1] = g + 4 * L.n_nodes;
    int* bi = (int*)(w + L.off_blkints);
    c->blk_flag = bi;
    c->node_flag = bi + (size_t)L.n_blocks;
    c->pblocks = bi + (size_t)L.n_blocks * 2;
    c->blk_rep = bi + (size_t)L.n_blocks * 3;
    c->cell_count = (int*)(w + L.off_cells);
    c->cell_start = c->cell_count + (size_t)L.n_blocks * 64;
    c->bin_q = (int*)(w + L.off_pints);
    c->bin_rank = c->bin_q + L.stride;
    c->keep = L.keep_cap > 0;
    c->keep_cap = L.keep_cap;
    c->rec_cap = L.rec_cap;
    for (int k = 0; k < L.n_bslots; k++) {
        char* b = w + L.off_bslots + (size_t)k * L.bs_bytes;
        mpm_ctx::BinSlot S;
        S.order = (int*)(b + L.bs_order);
        S.items = (int4*)(b + L.bs_items);
        S.blk_stamp = (int*)(b + L.bs_stamp);
        S.blk_item0 = (int*)(b + L.bs_item0);
        S.blk_nitems = (int*)(b + L.bs_nitems);
        S.nblocks = (int*)(b + L.bs_nblocks);
        S.nmap = (int*)(b + L.bs_nmap);
        S.counters = (int*)(b + L.bs_counters);
        S.nrec = (int*)(b + L.bs_nrec);
        S.xlist = (int4*)(b + L.bs_xlist);
        S.keep_pm = S.keep_v = nullptr;
        if (c->keep) {
            S.keep_pm = (float4*)(w + L.off_keep + (size_t)k * L.keep_bytes);
            S.keep_v = S.keep_pm + (size_t)L.keep_cap * NODES_PER_BLK;
        }
        c->bslots.push_back(S);
    }
    c->kept_of.assign(L.n_bslots, -1);
    c->max_items = L.max_items;
    c->tiles_pm = (float4*)(w + L.off_tiles);
    c->tiles_v = c->tiles_pm + (size_t)L.max_items * TILE;
    c->gstamp = (int*)(w + L.off_counters);
    size_t pb = (size_t)8 * max_frames * std::max(1, L.n_prims);
    c->pstate = (float*)(w + L.off_prim_bufs);
    c->pstate_g = c->pstate + pb;
    c->pvel = c->pstate + 2 * pb;
    c->pvel_g = c->pstate + 3 * pb;
    c->cell_gcnt = (int*)(w + L.off_det + L.det_gn);
    c->cell_ng = c->cell_gcnt + (size_t)L.n_blocks * 64;
    c->cell_grp = (unsigned*)(w + L.off_det + L.det_garr);
    c->cell_head = (unsigned long long*)(w + L.off_det + L.det_head);
    c->grp_next = (int*)(w + L.off_det + L.det_grp);
    c->grp_cnt = c->grp_next + L.stride;
    c->padj = (float*)(w + L.off_det + L.det_padj);
    c->pbits = (unsigned*)(w + L.off_det + L.det_pbits);
    c->macc = (unsigned long long*)(w + L.off_det + L.det_macc);
    c->pmd = (float*)(w + L.off_det + L.det_pmd);

    // host tables
    std::vector<int4> chunks;
    std::vector<int> blk_scene(L.n_blocks, 0);
    int poff = 0, boff = 0, prim_off = 0, moff = 0, bmoff = 0;
    long long goff = 0;
    for (int i = 0; i < n_scenes; i++) {
        const mpm_scene_desc& d = scenes[i];
        SceneC s;
        memset(&s, 0, sizeof(s));
        s.n_grid = d.n_grid;
        s.nb = (d.n_grid + 3) / 4;
        s.p_off = poff;
        s.p_cnt = d.n_particles;
        s.blk_off = boff;
        s.prim_off = prim_off;
        s.n_prims = d.n_prims;
        s.gf_mode = d.ground_friction == 0 ? 0 : (d.ground_friction < 10 ? 1 : 2);
        s.m_off = moff;
        s.bm_off = bmoff;
        s.chunk_off = (int)chunks.size();
        s.g_off = goff;
        s.dx = (float)d.dx;
        s.inv_dx = (float)d.inv_dx;
        s.dt = (float)d.dt;
        s.p_mass = (float)d.p_mass;
        s.p_mass_d = d.p_mass;
        s.mu = (float)d.mu;
        s.lam = (float)d.lam;
        s.ys = (float)d.yield_stress;
        s.stress_k = (float)(-d.dt * d.p_vol * 4 * d.inv_dx * d.inv_dx);
        s.c4 = (float)(4 * d.inv_dx);
        s.lower = (float)(d.lower_bound * d.dx);
        s.upper = (float)(1. - 3 * d.dx);
        s.gf = (float)d.ground_friction;
        for (int k = 0; k < 3; k++) s.g30[k] = (float)d.dt * (float)d.gravity[k] * 30.f;
        for (int p = 0; p < d.n_particles; p += PPB) {
            int4 ch;
            ch.x = i; ch.y = poff + p; ch.z = std::min(PPB, d.n_particles - p); ch.w = 0;
            chunks.push_back(ch);
        }
        s.n_chunks = (int)chunks.size() - s.chunk_off;
        int nbk = s.nb * s.nb * s.nb;
        for (int b = 0; b < nbk; b++) blk_scene[boff + b] = i;
        c->scenes.push_back(s);
        c->p_off.push_back(poff);
        c->n_particles += d.n_particles;
        poff += (d.n_particles + 63) & ~63;
        boff += nbk;
        bmoff += (nbk + 31) / 32;
        goff += (long long)nbk * NODES_PER_BLK;
        prim_off += d.n_prims;
        moff += d.n_grid * d.n_grid * d.n_grid;
    }
    if (chunks.empty()) { int4 z; z.x = 0; z.y = 0; z.z = 0; z.w = 0; chunks.push_back(z); }
    for (int q = 0; q < L.n_prims; q++) c->prims.push_back(to_primc(prims[q]));
    if (c->prims.empty()) { PrimC z; memset(&z, 0, sizeof(z)); c->prims.push_back(z); }
    c->n_chunks = (int)chunks.size();
    hipError_t e = hipSuccess;
    e = (e == hipSuccess) ? hipMemsetAsync(workspace, 0, L.total, st) : e;
    e = (e == hipSuccess) ? hipMemcpyAsync(c->d_scenes, c->scenes.data(), sizeof(SceneC) * c->scenes.size(), hipMemcpyHostToDevice, st) : e;
    e = (e == hipSuccess) ? hipMemcpyAsync(c->d_prims, c->prims.data(), sizeof(PrimC) * c->prims.size(), hipMemcpyHostToDevice, st) : e;
    e = (e == hipSuccess) ? hipMemcpyAsync(c->d_chunks, chunks.data(), sizeof(int4) * chunks.size(), hipMemcpyHostToDevice, st) : e;
    e = (e == hipSuccess) ? hipMemcpyAsync(c->d_blk_scene, blk_scene.data(), sizeof(int) * blk_scene.size(), hipMemcpyHostToDevice, st) : e;
    e = (e == hipSuccess) ? hipStreamSynchronize(st) : e;
    if (e != hipSuccess) {
        delete c;
        return fail(std::string("mpm_create: ") + hipGetErrorString(e));
    }
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
    }
    c->grid_ctas = std::max(64, ncu * 8);
    c->tile_ctas = std::max(64, ncu * 4);
    // the fork and join events order work between two streams of this device
    // only: a device-scope fence suffices (no system-scope cache write-back at
    // each record; 0.75 ms less per eager 200-substep forward + adjoint pass)
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess) {
        mpm_destroy(c);
        return fail("mpm_create: stream/event creation failed");
    }
    *out = c;
    return 0;
}

int mpm_destroy(mpm_ctx* ctx) {
    if (!ctx) return 0;
    if (ctx->side) { hipStreamSynchronize(ctx->side); hipStreamDestroy(ctx->side); }
    if (ctx->ev_fork) hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) hipEventDestroy(ctx->ev_join);
    for (auto& r : ctx->recs) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
    for (auto e : ctx->ev_pool) hipEventDestroy(e);
    delete ctx;
    return 0;
}

int mpm_set_launch(mpm_ctx* ctx, int particles_per_item, int tile_ctas, int grid_ctas) {
    if (!ctx) return fail("mpm_set_launch: null context");
    if (particles_per_item != 0 &&
        (particles_per_item < WI_MIN || particles_per_item > WI_MAX || particles_per_item % WG != 0))
        return fail("mpm_set_launch: particles_per_item must be a multiple of " + std::to_string(WG) + " in [" +
                    std::to_string(WI_MIN) + ", " + std::to_string(WI_MAX) + "]");
    if (tile_ctas < 0 || tile_ctas > 65536 || grid_ctas < 0 || grid_ctas > 65536)
        return fail("mpm_set_launch: workgroup counts must be in [1, 65536] (0 keeps the current value)");
    if (particles_per_item) ctx->wi = particles_per_item;
    if (tile_ctas) ctx->tile_ctas = tile_ctas;
    if (grid_ctas) ctx->grid_ctas = grid_ctas;
    ctx->drop_kept();  // kept binnings were made with the previous item size
    return 0;
}

int mpm_get_launch(mpm_ctx* ctx, int* particles_per_item, int* tile_ctas, int* grid_ctas) {
    if (!ctx) return fail("mpm_get_launch: null context");
    if (particles_per_item) *particles_per_item = ctx->wi;
    if (tile_ctas) *tile_ctas = ctx->tile_ctas;
    if (grid_ctas) *grid_ctas = ctx->grid_ctas;
    return 0;
}

int mpm_set_deterministic(mpm_ctx* ctx, int on) {
    if (!ctx) return fail("mpm_set_deterministic: null context");
    ctx->det = on != 0;
    ctx->drop_kept();  // kept binnings were ranked the other way
    return 0;
}

int mpm_get_deterministic(mpm_ctx* ctx) { return ctx ? (ctx->det ? 1 : 0) : fail("mpm_get_deterministic: null context"); }

// the caller-owned buffer of a slot re-sort every `substeps` substeps: the
// permutation of every boundary frame (one int per slot each), a staging
// frame for F of a boundary frame in the old order, per-block particle
// counts and their per-scene sums
struct ResortLayout { size_t perm, fstage, blk, total; };
static ResortLayout resort_layout(const mpm_ctx* c, int substeps) {
    ResortLayout r;
    size_t n_bnd = (size_t)std::max(0, (c->max_frames - 2) / substeps);
    size_t o = 0;
    r.perm = o; o = align256(o + sizeof(int) * n_bnd * (size_t)c->stride);
    r.fstage = o; o = align256(o + c->frame_floats() * sizeof(float));
    r.blk = o; o = align256(o + sizeof(int) * 2 * (size_t)c->n_blocks);
    r.total = o;
    return r;
}

int64_t mpm_resort_bytes(mpm_ctx* ctx, int substeps) {
    if (!ctx) return fail("mpm_resort_bytes: null context");
    if (substeps < RESORT_MIN)
        return fail("mpm_resort_bytes: interval must be at least " + std::to_string(RESORT_MIN) + " substeps");
    return (int64_t)resort_layout(ctx, substeps).total;
}

int mpm_set_resort(mpm_ctx* ctx, int substeps, void* buffer, int64_t buffer_bytes) {
    if (!ctx) return fail("mpm_set_resort: null context");
    if (substeps != 0 && substeps < RESORT_MIN)
        return fail("mpm_set_resort: interval must be 0 (off) or at least " + std::to_string(RESORT_MIN) + " substeps");
    if (substeps != 0) {
        ResortLayout r = resort_layout(ctx, substeps);
        if (!buffer || buffer_bytes < (int64_t)r.total)
            return fail("mpm_set_resort: buffer of " + std::to_string(buffer_bytes) + " bytes, " + std::to_string(r.total) + " needed");
        char* b = static_cast<char*>(buffer);
        ctx->rs_perm = (int*)(b + r.perm);
        ctx->rs_fstage = (float*)(b + r.fstage);
        ctx->blk_tot = (int*)(b + r.blk);
        ctx->blk_pre = ctx->blk_tot + ctx->n_blocks;
    } else {
        ctx->rs_perm = nullptr;
        ctx->rs_fstage = nullptr;
        ctx->blk_tot = ctx->blk_pre = nullptr;
    }
    ctx->resort = substeps;
    // frames after 0 were written in the previous slot orders
    ctx->valid_hi = std::min(ctx->valid_hi, 0);
    ctx->seg_valid = -1;
    ctx->adj_top = -1;
    ctx->drop_kept();
    return 0;
}

int mpm_get_resort(mpm_ctx* ctx) { return ctx ? ctx->resort : fail("mpm_get_resort: null context"); }

int mpm_slot_order(mpm_ctx* ctx, int f, int* map, void* stream) {
    if (!ctx || !map) return fail("mpm_slot_order: null argument");
    if (f < 0 || f >= ctx->max_frames) return fail("mpm_slot_order: frame " + std::to_string(f) + " outside [0, max_frames)");
    if (f > ctx->valid_hi) return fail("mpm_slot_order: frame " + std::to_string(f) + " has not been simulated");
    hipStream_t st = S_(stream);
    std::vector<int> id(ctx->stride);
    for (int i = 0; i < ctx->stride; i++) id[i] = i;
    HIPCHK(hipMemcpyAsync(map, id.data(), sizeof(int) * (size_t)ctx->stride, hipMemcpyHostToDevice, st));
    for (int s = 1; s <= f; s++)
        if (ctx->is_bnd(s))
            hipLaunchKernelGGL(k_resort_compose, dim3(64, ctx->n_scenes), dim3(256), 0, st, map, (const int*)ctx->perm_of(s), ctx->d_scenes);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int mpm_set_profiling(mpm_ctx* ctx, int on) {
    if (!ctx) return fail("mpm_set_profiling: null context");
    ctx->prof = on != 0;
    return 0;
}

int mpm_profile_read(mpm_ctx* ctx, double* ms, long long* counts, int reset) {
    if (!ctx) return fail("mpm_profile_read: null context");
    for (auto& r : ctx->recs) {
        HIPCHK(hipEventSynchronize(r.b));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
        ctx->prof_ms[r.phase] += t;
        ctx->prof_n[r.phase] += 1;
        ctx->ev_pool.push_back(r.a);
        ctx->ev_pool.push_back(r.b);
    }
    ctx->recs.clear();
    for (int i = 0; i < MPM_N_PHASES; i++) {
        if (ms) ms[i] = ctx->prof_ms[i];
        if (counts) counts[i] = ctx->prof_n[i];
        if (reset) { ctx->prof_ms[i] = 0; ctx->prof_n[i] = 0; }
    }
    return 0;
}

int mpm_stride(mpm_ctx* ctx) { return ctx ? ctx->stride : -1; }
int mpm_particle_offset(mpm_ctx* ctx, int scene) {
    if (!ctx || scene < 0 || scene >= ctx->n_scenes) return fail("mpm_particle_offset: bad scene");
    return ctx->p_off[scene];
}

int mpm_frame(mpm_ctx* ctx, int f, void* stream, float** ptr) {
    if (!ctx || !ptr) return fail("mpm_frame: null argument");
    if (ensure_frame(ctx, f, S_(stream))) return -1;
    *ptr = ctx->frame_ptr(f);
    return 0;
}

int mpm_adjoint(mpm_ctx* ctx, int f, float** ptr) {
    if (!ctx || !ptr) return fail("mpm_adjoint: null argument");
    if (f != ctx->adj_top) return fail("mpm_adjoint: frame " + std::to_string(f) + " is not the current adjoint frame " + std::to_string(ctx->adj_top));
    *ptr = ctx->adj_ptr(f);
    return 0;
}

int mpm_prim_buffers(mpm_ctx* ctx, float** state, float** state_grad, float** vel, float** vel_grad) {
    if (!ctx) return fail("mpm_prim_buffers: null context");
    if (state) *state = ctx->pstate;
    if (state_grad) *state_grad = ctx->pstate_g;
    if (vel) *vel = ctx->pvel;
    if (vel_grad) *vel_grad = ctx->pvel_g;
    return 0;
}

// frame f was produced by a forward substep from frame f - 1
static void mark_simulated(mpm_ctx* c, int f) {
    c->valid_hi = f;
    if (f % c->K == 0) return;
    if (c->seg_valid == f / c->K && c->seg_hi == f - 1) c->seg_hi = f;
    else { c->seg_valid = f / c->K; c->seg_lo = f; c->seg_hi = f; }
}
// frame f was written directly: the trajectory restarts there
static void mark_set(mpm_ctx* c, int f) {
    c->valid_hi = f;
    for (auto& a : c->aff_of) a = -1;
    c->drop_kept();
    if (f % c->K == 0) { c->seg_valid = -1; c->seg_lo = c->seg_hi = -1; }
    else { c->seg_valid = f / c->K; c->seg_lo = f; c->seg_hi = f; }
}

// an entry point that adds into the adjoint of frame f: f must be a simulated
// frame of the context and, once the adjoint chain has a top, that top
static int adj_frame_check(mpm_ctx* c, int f, const char* who) {
    if (f < 0 || f >= c->max_frames) return fail(std::string(who) + ": frame " + std::to_string(f) + " outside [0, max_frames)");
    if (f > c->valid_hi) return fail(std::string(who) + ": frame " + std::to_string(f) + " has not been simulated");
    if (c->adj_top >= 0 && f != c->adj_top)
        return fail(std::string(who) + ": frame " + std::to_string(f) + " is not the current adjoint frame " +
                    std::to_string(c->adj_top));
    return 0;
}

int mpm_set_frame(mpm_ctx* ctx, int f, const float* src, void* stream) {
    if (!ctx || !src) return fail("mpm_set_frame: null argument");
    if (f < 0 || f >= ctx->max_frames) return fail("mpm_set_frame: frame out of range");
    HIPCHK(hipMemcpyAsync(ctx->frame_ptr(f), src, ctx->frame_floats() * sizeof(float), hipMemcpyDeviceToDevice, S_(stream)));
    mark_set(ctx, f);
    return 0;
}

int mpm_get_frame(mpm_ctx* ctx, int f, float* dst, void* stream) {
    if (!ctx || !dst) return fail("mpm_get_frame: null argument");
    if (ensure_frame(ctx, f, S_(stream))) return -1;
    HIPCHK(hipMemcpyAsync(dst, ctx->frame_ptr(f), ctx->frame_floats() * sizeof(float), hipMemcpyDeviceToDevice, S_(stream)));
    return 0;
}

int mpm_reset_frame0(mpm_ctx* ctx, const float* x, void* stream) {
    if (!ctx || !x) return fail("mpm_reset_frame0: null argument");
    hipLaunchKernelGGL(k_reset_frame0, dim3((ctx->stride + 255) / 256), dim3(256), 0, S_(stream), ctx->stride, x, ctx->frame_ptr(0));
    HIPCHK(hipGetLastError());
    mark_set(ctx, 0);
    return 0;
}

int mpm_set_prim_frame(mpm_ctx* ctx, int f, const float* src, void* stream) {
    if (!ctx || !src) return fail("mpm_set_prim_frame: null argument");
    if (f < 0 || f >= ctx->max_frames) return fail("mpm_set_prim_frame: frame out of range");
    if (ctx->n_prims == 0) return 0;
    HIPCHK(hipMemcpyAsync(ctx->pstate + (size_t)f * ctx->n_prims * 8, src, sizeof(float) * 8 * ctx->n_prims, hipMemcpyDeviceToDevice,
                          S_(stream)));
    ctx->drop_kept();  // kept grids were built with the old primitive states
    return 0;
}

int mpm_get_prim_frame(mpm_ctx* ctx, int f, float* dst, void* stream) {
    if (!ctx || !dst) return fail("mpm_get_prim_frame: null argument");
    if (f < 0 || f >= ctx->max_frames) return fail("mpm_get_prim_frame: frame out of range");
    if (ctx->n_prims == 0) return 0;
    HIPCHK(hipMemcpyAsync(dst, ctx->pstate + (size_t)f * ctx->n_prims * 8, sizeof(float) * 8 * ctx->n_prims, hipMemcpyDeviceToDevice,
                          S_(stream)));
    return 0;
}

int mpm_set_actions(mpm_ctx* ctx, int s, int n_substeps, const float* actions, void* stream) {
    if (!ctx || !actions) return fail("mpm_set_actions: null argument");
    if (n_substeps <= 0 || s < 0 || (long long)(s + 1) * n_substeps >= ctx->max_frames)
        return fail("mpm_set_actions: env step outside the frame capacity");
    if (ctx->n_prims == 0) return 0;
    int n = ctx->n_prims * n_substeps;
    hipLaunchKernelGGL(k_set_actions, dim3((n + 63) / 64), dim3(64), 0, S_(stream), ctx->d_prims, ctx->n_prims, s, n_substeps, actions,
                       ctx->pvel);
    HIPCHK(hipGetLastError());
    return 0;
}

int mpm_substep_fwd(mpm_ctx* ctx, int s, int count, void* stream) {
    if (!ctx) return fail("mpm_substep_fwd: null context");
    if (count < 0 || s < 0 || s + count >= ctx->max_frames) return fail("mpm_substep_fwd: frames outside [0, max_frames)");
    hipStream_t st = S_(stream);
    if (count == 0) return 0;
    if (ensure_frame(ctx, s, st)) return -1;
    for (int f = s; f < s + count; f++) {
        if (run_fwd(ctx, f, st)) return -1;
        mark_simulated(ctx, f + 1);
    }
    return 0;
}

int mpm_substep_bwd(mpm_ctx* ctx, int s, int count, void* stream) {
    if (!ctx) return fail("mpm_substep_bwd: null context");
    if (count <= 0) return 0;
    if (s < 0 || s + count > ctx->valid_hi) return fail("mpm_substep_bwd: substeps not simulated");
    if (ctx->adj_top < 0) ctx->adj_top = s + count;
    if (ctx->adj_top != s + count)
        return fail("mpm_substep_bwd: adjoint order violated (current adjoint frame " + std::to_string(ctx->adj_top) + ", expected " +
                    std::to_string(s + count) + ")");
    hipStream_t st = S_(stream);
    for (int f = s + count - 1; f >= s; f--) {
        if (run_bwd(ctx, f, st)) return -1;
        ctx->adj_top = f;
    }
    return 0;
}

int mpm_grid_mass(mpm_ctx* ctx, int f, float* out, void* stream) {
    if (!ctx || !out) return fail("mpm_grid_mass: null argument");
    hipStream_t st = S_(stream);
    if (ensure_frame(ctx, f, st)) return -1;
    // visit the particles in the cell order of frame f's binning, or of frame
    // f-1's (one substep stale, nearly the same cells), when one is kept
    int g = -1;
    for (int h = f; h >= std::max(0, f - 1); h--)
        if (g < 0 && ctx->keep && ctx->kept_of[ctx->bslot_of(h)] == h && ctx->epoch_of(h) == ctx->epoch_of(f)) g = h;
    if (g >= 0)
        hipLaunchKernelGGL(k_grid_mass_items, dim3(ctx->tile_ctas), dim3(WG), 0, st, ctx->targs(ctx->bslot_of(g)), ctx->frame_ptr(f), out,
                           ctx->macc, ctx->det ? 1 : 0);
    else
        hipLaunchKernelGGL(k_grid_mass, dim3(ctx->n_chunks), dim3(PPB), 0, st, ctx->cargs(), ctx->frame_ptr(f), out, ctx->macc,
                           ctx->det ? 1 : 0);
    if (ctx->det)
        hipLaunchKernelGGL(k_grid_mass_convert, dim3(256, ctx->n_scenes), dim3(256), 0, st, ctx->d_scenes, ctx->macc, out);
    HIPCHK(hipGetLastError());
    return 0;
}

int mpm_grid_mass_bwd(mpm_ctx* ctx, int f, const float* g, void* stream) {
    if (!ctx || !g) return fail("mpm_grid_mass_bwd: null argument");
    if (adj_frame_check(ctx, f, "mpm_grid_mass_bwd")) return -1;
    if (ctx->adj_top < 0) ctx->adj_top = f;
    hipStream_t st = S_(stream);
    if (ensure_frame(ctx, f, st)) return -1;
    hipLaunchKernelGGL(k_grid_mass_bwd, dim3(ctx->n_chunks), dim3(PPB), 0, st, ctx->cargs(), ctx->frame_ptr(f), g, (const float*)nullptr,
                       0.f, ctx->adj_ptr(f));
    HIPCHK(hipGetLastError());
    return 0;
}

int mpm_density_loss_bwd(mpm_ctx* ctx, int f, const float* mass, const float* target, float weight, void* stream) {
    if (!ctx || !mass || !target) return fail("mpm_density_loss_bwd: null argument");
    if (adj_frame_check(ctx, f, "mpm_density_loss_bwd")) return -1;
    if (ctx->adj_top < 0) ctx->adj_top = f;
    hipStream_t st = S_(stream);
    if (ensure_frame(ctx, f, st)) return -1;
    hipLaunchKernelGGL(k_grid_mass_bwd, dim3(ctx->n_chunks), dim3(PPB), 0, st, ctx->cargs(), ctx->frame_ptr(f), mass, target, weight,
                       ctx->adj_ptr(f));
    HIPCHK(hipGetLastError());
    return 0;
}

int mpm_add_obs_grad(mpm_ctx* ctx, int f, const float* xv_grad, const float* prim_grad, void* stream) {
    if (!ctx) return fail("mpm_add_obs_grad: null context");
    if (adj_frame_check(ctx, f, "mpm_add_obs_grad")) return -1;
    if (ctx->adj_top < 0) ctx->adj_top = f;
    int n = std::max(ctx->stride, ctx->n_prims * 8);
    hipLaunchKernelGGL(k_obs_grad, dim3((n + 255) / 256), dim3(256), 0, S_(stream), ctx->stride, ctx->n_prims, xv_grad, prim_grad,
                       ctx->adj_ptr(f), ctx->pstate_g + (size_t)f * ctx->n_prims * 8);
    HIPCHK(hipGetLastError());
    return 0;
}

int mpm_min_dist_cols(mpm_ctx* ctx, int scene) {
    if (!ctx) return fail("mpm_min_dist_cols: null context");
    int lo = 0, hi = ctx->n_scenes;
    if (scene >= 0) {
        if (scene >= ctx->n_scenes) return fail("mpm_min_dist_cols: bad scene");
        lo = scene; hi = scene + 1;
    }
    int best = 0;
    for (int k = lo; k < hi; k++) {
        const SceneC& S = ctx->scenes[k];
        int c = 0;
        for (int q = 0; q < S.n_prims; q++) {
            int t = ctx->prims[S.prim_off + q].type;
            c += (t == MPM_GRIPPER || t == MPM_GRIPPER2) ? 2 : 1;
        }
        best = std::max(best, c);
    }
    return best;
}

int mpm_min_dist(mpm_ctx* ctx, int f, float* out, int n_cols, void* stream) {
    if (!ctx || !out) return fail("mpm_min_dist: null argument");
    if (n_cols < mpm_min_dist_cols(ctx, -1)) return fail("mpm_min_dist: output has fewer columns than the scenes need");
    hipStream_t st = S_(stream);
    if (ensure_frame(ctx, f, st)) return -1;
    if (ctx->n_prims == 0) return 0;
    hipLaunchKernelGGL(k_min_dist, dim3(ctx->n_chunks), dim3(PPB), 0, st, ctx->cargs(), ctx->frame_ptr(f), ctx->d_prims,
                       ctx->pstate + (size_t)f * ctx->n_prims * 8, out);
    HIPCHK(hipGetLastError());
    return 0;
}

int mpm_min_dist_bwd(mpm_ctx* ctx, int f, const float* gout, int n_cols, void* stream) {
    if (!ctx || !gout) return fail("mpm_min_dist_bwd: null argument");
    if (n_cols < mpm_min_dist_cols(ctx, -1)) return fail("mpm_min_dist_bwd: adjoint has fewer columns than the scenes need");
    if (adj_frame_check(ctx, f, "mpm_min_dist_bwd")) return -1;
    if (ctx->adj_top < 0) ctx->adj_top = f;
    hipStream_t st = S_(stream);
    if (ensure_frame(ctx, f, st)) return -1;
    if (ctx->n_prims == 0) return 0;
    float* gps = ctx->pstate_g + (size_t)f * ctx->n_prims * 8;
    hipLaunchKernelGGL(k_min_dist_bwd, dim3(ctx->n_chunks), dim3(PPB), 0, st, ctx->cargs(), ctx->frame_ptr(f), ctx->d_prims,
                       ctx->pstate + (size_t)f * ctx->n_prims * 8, gout, ctx->adj_ptr(f), gps, ctx->det ? ctx->pmd : (float*)nullptr);
    if (ctx->det) hipLaunchKernelGGL(k_min_dist_reduce, dim3(ctx->n_scenes), dim3(64), 0, st, ctx->d_scenes, (const float*)ctx->pmd, gps);
    HIPCHK(hipGetLastError());
    return 0;
}

static size_t coll_layout(int n_pairs, int n_points, int n_scenes, int max_frames, size_t off[6]) {
    size_t o = 0;
    off[0] = o; o = align256(o + sizeof(int4) * (size_t)n_pairs);
    off[1] = o; o = align256(o + sizeof(int2) * (size_t)n_scenes);
    off[2] = o; o = align256(o + sizeof(float4) * (size_t)n_pairs * n_points);
    off[3] = o; o = align256(o + sizeof(int) * (size_t)max_frames * n_pairs);
    off[4] = o; o = align256(o + sizeof(float4) * (size_t)max_frames * n_pairs);
    off[5] = o; o = align256(o + sizeof(float4) * (size_t)max_frames * n_pairs);
    return o;
}

static const int COLL_MAX_POINTS = 1 << 16;
static const int COLL_MAX_PAIRS = 1 << 12;

int64_t mpm_collision_bytes(mpm_ctx* ctx, int n_pairs, int n_points) {
    if (!ctx || n_pairs < 0 || n_pairs > COLL_MAX_PAIRS || n_points < 1 || n_points > COLL_MAX_POINTS) return -1;
    size_t off[6];
    return (int64_t)coll_layout(n_pairs, n_points, ctx->n_scenes, ctx->max_frames, off);
}

int mpm_set_collisions(mpm_ctx* ctx, int n_pairs, const int* pairs, int n_points, const float* rand_num, void* buffer,
                       int64_t buffer_bytes, void* stream) {
    if (!ctx) return fail("mpm_set_collisions: null context");
    hipStream_t st = S_(stream);
    if (n_pairs == 0) {
        ctx->n_pairs = 0;
        ctx->n_cgroups = 0;
        ctx->coll = CollArgs{};
        return 0;
    }
    if (n_pairs < 0 || n_pairs > COLL_MAX_PAIRS) return fail("mpm_set_collisions: pair count outside [0, 4096]");
    if (n_points < 1 || n_points > COLL_MAX_POINTS) return fail("mpm_set_collisions: point count outside [1, 65536]");
    if (!pairs || !rand_num || !buffer) return fail("mpm_set_collisions: null argument");
    size_t off[6];
    size_t need = coll_layout(n_pairs, n_points, ctx->n_scenes, ctx->max_frames, off);
    if (buffer_bytes < (int64_t)need) return fail("mpm_set_collisions: buffer smaller than mpm_collision_bytes");
    if (((uintptr_t)buffer & 255) != 0) return fail("mpm_set_collisions: buffer not 256-byte aligned");
    // scene of each primitive
    std::vector<int> prim_scene(ctx->n_prims, -1);
    for (int k = 0; k < ctx->n_scenes; k++)
        for (int q = 0; q < ctx->scenes[k].n_prims; q++) prim_scene[ctx->scenes[k].prim_off + q] = k;
    std::vector<int4> hp(n_pairs);
    std::vector<int2> hg;
    int last_scene = -1;
    for (int p = 0; p < n_pairs; p++) {
        int i = pairs[2 * p], j = pairs[2 * p + 1];
        if (i < 0 || i >= ctx->n_prims || j < 0 || j >= ctx->n_prims)
            return fail("mpm_set_collisions: pair " + std::to_string(p) + " names a primitive outside the context");
        if (prim_scene[i] != prim_scene[j]) return fail("mpm_set_collisions: pair " + std::to_string(p) + " spans two scenes");
        int tj = ctx->prims[j].type;
        if (tj != P_BOX && tj != P_GRIPPER)
            return fail("mpm_set_collisions: pair " + std::to_string(p) + ": the sampled primitive must be a Box or Gripper");
        int sc = prim_scene[i];
        if (sc < last_scene) return fail("mpm_set_collisions: pairs are not grouped by ascending scene");
        if (sc != last_scene) { hg.push_back(make_int2(p, 0)); last_scene = sc; }
        hg.back().y++;
        hp[p] = make_int4(i, j, 0, 0);
    }
    std::vector<float4> hr((size_t)n_pairs * n_points);
    for (size_t k = 0; k < hr.size(); k++) hr[k] = make_float4(rand_num[3 * k], rand_num[3 * k + 1], rand_num[3 * k + 2], 0.f);
    char* b = (char*)buffer;
    CollArgs ca{};
    ca.pairs = (const int4*)(b + off[0]);
    ca.groups = (const int2*)(b + off[1]);
    ca.rand = (const float4*)(b + off[2]);
    ca.idx = (int*)(b + off[3]);
    ca.pt = (float4*)(b + off[4]);
    ca.pre = (float4*)(b + off[5]);
    ca.n_pairs = n_pairs;
    ca.n_points = n_points;
    HIPCHK(hipMemcpyAsync(b + off[0], hp.data(), sizeof(int4) * hp.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b + off[1], hg.data(), sizeof(int2) * hg.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b + off[2], hr.data(), sizeof(float4) * hr.size(), hipMemcpyHostToDevice, st));
    // no frame has records yet: every collision index reads -1 (all bytes 0xff)
    HIPCHK(hipMemsetAsync(b + off[3], 0xff, sizeof(int) * (size_t)ctx->max_frames * n_pairs, st));
    HIPCHK(hipMemsetAsync(b + off[4], 0, need - off[4], st));
    HIPCHK(hipStreamSynchronize(st));
    ctx->coll = ca;
    ctx->n_pairs = n_pairs;
    ctx->n_cgroups = (int)hg.size();
    return 0;
}

int mpm_get_collisions(mpm_ctx* ctx, int f, int* idx, float* points, float* pre, void* stream) {
    if (!ctx) return fail("mpm_get_collisions: null context");
    if (f < 1 || f >= ctx->max_frames) return fail("mpm_get_collisions: frame outside [1, max_frames)");
    if (ctx->n_pairs == 0) return 0;
    hipStream_t st = S_(stream);
    int n = ctx->n_pairs;
    size_t r0 = (size_t)f * n;
    std::vector<float4> tmp(n);
    if (idx) HIPCHK(hipMemcpyAsync(idx, ctx->coll.idx + r0, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    if (points) {
        HIPCHK(hipMemcpyAsync(tmp.data(), ctx->coll.pt + r0, sizeof(float4) * n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int p = 0; p < n; p++) { points[3 * p] = tmp[p].x; points[3 * p + 1] = tmp[p].y; points[3 * p + 2] = tmp[p].z; }
    }
    if (pre) {
        HIPCHK(hipMemcpyAsync(tmp.data(), ctx->coll.pre + r0, sizeof(float4) * n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int p = 0; p < n; p++) { pre[3 * p] = tmp[p].x; pre[3 * p + 1] = tmp[p].y; pre[3 * p + 2] = tmp[p].z; }
    }
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int mpm_action_grad(mpm_ctx* ctx, int s, int n_substeps, float* out, void* stream) {
    if (!ctx || !out) return fail("mpm_action_grad: null argument");
    if (n_substeps <= 0 || s < 0 || (long long)(s + 1) * n_substeps >= ctx->max_frames) return fail("mpm_action_grad: bad env step");
    if (ctx->n_prims == 0) return 0;
    int n = ctx->n_prims * 7;
    hipLaunchKernelGGL(k_action_grad, dim3((n + 63) / 64), dim3(64), 0, S_(stream), ctx->d_prims, ctx->n_prims, s, n_substeps,
                       ctx->pvel_g, out);
    HIPCHK(hipGetLastError());
    return 0;
}

int mpm_update_prims(mpm_ctx* ctx, const mpm_prim_desc* prims, void* stream) {
    if (!ctx || (!prims && ctx->n_prims > 0)) return fail("mpm_update_prims: null argument");
    for (int q = 0; q < ctx->n_prims; q++) {
        if (prims[q].type != ctx->prims[q].type) return fail("mpm_update_prims: primitive types may not change");
        ctx->prims[q] = to_primc(prims[q]);
    }
    ctx->drop_kept();  // kept grids were built with the old primitive parameters
    if (ctx->n_prims == 0) return 0;
    HIPCHK(hipMemcpyAsync(ctx->d_prims, ctx->prims.data(), sizeof(PrimC) * ctx->n_prims, hipMemcpyHostToDevice, S_(stream)));
    HIPCHK(hipStreamSynchronize(S_(stream)));
    return 0;
}

int mpm_copy_frame(mpm_ctx* ctx, int src, int dst, void* stream) {
    if (!ctx) return fail("mpm_copy_frame: null context");
    if (dst < 0 || dst >= ctx->max_frames) return fail("mpm_copy_frame: frame out of range");
    hipStream_t st = S_(stream);
    if (ctx->epoch_of(src) != ctx->epoch_of(dst))
        return fail("mpm_copy_frame: frames " + std::to_string(src) + " and " + std::to_string(dst) +
                    " have different slot orders (slot re-sort is on: mpm_set_resort)");
    if (ensure_frame(ctx, src, st)) return -1;
    if (src != dst) {
        HIPCHK(hipMemcpyAsync(ctx->frame_ptr(dst), ctx->frame_ptr(src), ctx->frame_floats() * sizeof(float), hipMemcpyDeviceToDevice, st));
        if (ctx->n_prims > 0)
            HIPCHK(hipMemcpyAsync(ctx->pstate + (size_t)dst * ctx->n_prims * 8, ctx->pstate + (size_t)src * ctx->n_prims * 8,
                                  sizeof(float) * 8 * ctx->n_prims, hipMemcpyDeviceToDevice, st));
    }
    mark_set(ctx, dst);
    return 0;
}

int mpm_clear_grads(mpm_ctx* ctx, int f, void* stream) {
    if (!ctx) return fail("mpm_clear_grads: null context");
    if (f < -1 || f >= ctx->max_frames) return fail("mpm_clear_grads: frame out of range");
    hipStream_t st = S_(stream);
    if (launch_fill(ctx->adj[0], 2 * ctx->frame_floats(), st)) return -1;
    size_t pb = (size_t)8 * ctx->max_frames * std::max(1, ctx->n_prims);
    if (launch_fill(ctx->pstate_g, pb, st)) return -1;
    if (launch_fill(ctx->pvel_g, pb, st)) return -1;
    ctx->adj_top = f;
    return 0;
}

}  // extern "C"
