// Mixed-precision CG2D on gfx950: FP32 PCG with FP64 residual replacement, batched over members.
//
// The system is the one cg2d_fortran solves (reference core/cg2d.py:205-269): the same normalisation (b * cg2dNorm,
// rhsNorm = 1/max|b|, normalised first guess with exchanged halos), the same 5-point operator and preconditioner,
// the same stopping rule err < tol^2 on the normalised residual. Per member:
//   r = bn - A x0, err = sum r^2                     (FP64: true residual, FP64 operator); stop when err < tol^2
//   sigma = sqrt(err); r32 = r / sigma; d = 0        (FP32 residual in units of sigma, kept for the whole solve)
//   PCG on A32 d = r32 in FP32 with FP32 operator/preconditioner copies; every dot product is formed and summed
//   in FP64 in a fixed order (below) and alpha/beta are FP64 ratios. When the recursive residual has fallen by
//   inner_tolerance since the last replacement (or below tol^2 / sigma^2, or after inner_max_iters iterations):
//     x = x + sigma * d (FP64); r = bn - A x (FP64); stop when err < tol^2; else the recursive residual is
//     replaced: r32 = r / sigma, d = 0, z = M32 r32, and PCG goes on from the same search direction with
//     beta = (z, r32) / eta of the previous iteration ("reliable updates": the Krylov space is kept, so the
//     iteration count stays that of the FP64 PCG instead of growing with every restart).
// max_iters caps the total inner iterations and max_outer the replacements (the first residual counted); a member
// that reaches a cap finishes with converged = 0. Un-normalisation and the halos of x are as cg2d_fortran leaves
// them.
//
// Work decomposition as the FP64 solve: one block per (member, tile), rings computed from their interior sources,
// per-tile partials summed in tile order. The FP32 work vectors (r32, s, d, z, q) are the solver's own and live in a
// compact layout (mx_geo: per member and tile the interior rows padded to a multiple of 4 floats, then copies of the
// west and east columns). A tile is NS float4 "slots" (row j, columns 4c..4c+3); every phase works on slots: the
// pointwise updates stream them flat, the stencil operand goes to an LDS tile (ring computed from the sources'
// compact values), and the stencil gives each lane one slot per 64-slot block, reading its neighbours from LDS and
// its FP32 coefficients (own 4-aligned copies, mx_geo) as float4 loads issued before the pointwise phase.
//
// Members move through their own state machine, one state per launch ("step"): every block of a member derives the
// member's action for the step from the member's record of the previous step and the partials, so all tiles of a
// member act alike; tile 0 writes the record of the step (records are double-buffered by step parity). States:
//   RST   err from the true-residual partials; finish, or (re)place the recursive residual (r32 = r / sigma,
//         d = 0, z, eta(n)); the first one starts the solve (sigma = sqrt(err), n = 1)
//   IB    inner phase B(n): d += alpha_{n-1} s_{n-1} (the solution update of iteration n-1, moved here where
//         s_{n-1} is read anyway; not right after a replacement, which absorbed it into x); replacement due ->
//         x += sigma d (next: TRUE); else s = z + beta s (s_1 = z), q = A32 s, sigma_n = s.q
//   IC    inner phase C(n): r32 -= alpha q; z = M32 r32; err(n) = r32.r32, eta(n+1) = z.r32
// (z and q are kept as FP32 vectors, so each inner phase applies one stencil and reads one coefficient set.)
//   TRUE  x's ring from its sources; r = bn - A x; err partials
//
// Inner dot products (FP64): per tile, the slots in blocks of 64 consecutive slots (row-major); each slot's four
// products are added in column order, a block's 64 slot sums are reduced by the wave tree of the "tile" order
// (wave_tree, slot k of the block in lane k), and the block sums are added in block order. The order depends on the
// tile shape only, not on the launch geometry. The true-residual sums (FP64 refinement) use the "tile" order itself.
#include <algorithm>

#include "mj_cg_sweep.h"
#include "mj_common.h"
#include "mj_internal.h"

namespace mj {

struct MxRec {
    int mode, n, iters, outer;  // state, inner iteration, inner iterations done, residuals placed
    int n0, fresh;              // n of the last replacement; 1 until the IB step after a replacement has run
    int def, pad_;              // 1: the update alpha_{n-2} s_{n-2} is deferred to B(n) (alpha_{n-2} in ap)
    double sigma;               // units of r32: the norm of the first true residual
    double elast;               // err of the last replacement in units of sigma^2
    double ap;                  // the deferred alpha
};
enum { MX_RST = 0, MX_IB = 1, MX_IC = 2, MX_TRUE = 3 };
// AX_IBF: B(n) with no update pending (after a replacement); AX_IBD: B(n) deferring alpha_{n-1} s_{n-1}; AX_IB:
// B(n) applying the deferred and the pending update
enum { AX_FINISH = 1, AX_RESTART, AX_UPD, AX_IB, AX_IBF, AX_IBD, AX_IC, AX_TRUE };

struct MxDev {
    int B, T, sNx, sNy, OL, nx, ny, plane;
    int64_t ms, oms, cms;  // member stride; FP64 operator slot stride; FP32 coefficient slot stride (T * G.TSc)
    const double *aW, *aS, *aC;
    const float *aW32, *aS32, *aC32, *pW32, *pS32, *pC32;
    const double *b, *x0;
    double* x;
    const double* norm;
    double tol_sq, itol_sq;
    int max_iters, inner_max, max_outer;
    const int32_t *full_off, *full_dst, *full_src;
    const int8_t* full_sign;
    const int32_t *ring_off, *ring_dst, *ring_src;
    const int8_t* ring_sign;
    const int4* cring;  // ring entries: {compact source or -1, padded source or -1, LDS offset | neg << 30, 0}
    double* r64;
    MxGeo G;     // compact geometry of the work vectors and coefficient copies
    int64_t cs;  // compact member stride = T * G.TS
    float *r0, *r1, *s0, *s1, *d, *z, *q;  // compact
    double *pmax, *psum, *perr64, *perr, *peta, *psig;  // [B*T] x3, [2][B*T] x3
    const int32_t* op_index;
    MxRec* rec;  // [2][B]
    int* nactive;
    int32_t* mdone;  // [B] step at which a member finished, -1 while active
    mj_cg2d_diag* diag;
};

#ifndef MJ_MX_THREADS
#define MJ_MX_THREADS 1024
#endif  // largest block
// Block when the config leaves it open: 1024 threads while the (member, tile) blocks fit one per CU (a small batch
// is latency-bound: one member's inner iteration takes ~29-36 us at 1024 threads against ~39-50 us at 256), else 512
constexpr int kMxSmallThreads = 1024, kMxLargeThreads = 512;

__device__ __forceinline__ int slot_of(const MxDev& a, int m) { return a.op_index ? a.op_index[m] : m; }
__device__ __forceinline__ float* R32(const MxDev& a, int n) { return n & 1 ? a.r1 : a.r0; }
__device__ __forceinline__ float* S32(const MxDev& a, int n) { return n & 1 ? a.s1 : a.s0; }
// compact offset of member m, tile t
__device__ __forceinline__ int64_t cbase(const MxDev& a, int m, int t) { return m * a.cs + (int64_t)t * a.G.TS; }
// FP32 coefficient copies of member m's operator slot at tile t
__device__ __forceinline__ int64_t coef_base(const MxDev& a, int m, int t) {
    return (int64_t)slot_of(a, m) * a.cms + (int64_t)t * a.G.TSc;
}
__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ void st4(float* p, float4 v) { *reinterpret_cast<float4*>(p) = v; }
// Row of slot f of a tile (exact: f < 2^14, and the fraction of (f + 0.5) / NC stays >= 1/64 away from an integer).
__device__ __forceinline__ int srow(int f, float inv_nc) { return (int)((float(f) + 0.5f) * inv_nc); }
// The components of a slot at columns >= sNx (padding) set to 0.
__device__ __forceinline__ float4 mask_pad(float4 v, int c4, int sNx) {
    if (c4 + 4 <= sNx) return v;
    v.w = 0.0f;
    if (c4 + 2 >= sNx) v.z = 0.0f;
    if (c4 + 1 >= sNx) v.y = 0.0f;
    return v;
}
// A slot's values into the LDS tile (its columns < sNx only: column sNx belongs to the ring).
__device__ __forceinline__ void lds_put(float* tile, const MxGeo& G, int j, int c4, float4 v, int sNx) {
    float* p = tile + G.lds(j, c4);
    if (c4 + 4 <= sNx) {
        st4(p, v);
        return;
    }
    p[0] = v.x;
    if (c4 + 1 < sNx) p[1] = v.y;
    if (c4 + 2 < sNx) p[2] = v.z;
}
// The west (then east) column of a stencil result, gathered in LDS and written as the compact vector's contiguous
// column copies after the sweep (coalesced, instead of a partial store per slot).
static __shared__ float g_edge[256];
__device__ __forceinline__ void edge_stash(int j, int c4, float4 v, int sNx, int sNy) {
    if (c4 == 0) g_edge[j] = v.x;
    const int k = sNx - 1 - c4;
    if (k >= 0 && k < 4) g_edge[sNy + j] = k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}
// The west and east column copies of a tile's compact vector (base: the tile's compact start) from the columns 0
// and sNx - 1 of the LDS tile (after the barrier that completes the tile).
__device__ __forceinline__ void edges_from_tile(const MxDev& a, const float* tile, float* base) {
    const int sNy = a.sNy;
    for (int e = threadIdx.x; e < 2 * sNy; e += blockDim.x) {
        const bool east = e >= sNy;
        base[sNy * a.G.CP + e] = tile[a.G.lds(east ? e - sNy : e, east ? a.sNx - 1 : 0)];
    }
}

// The FP32 coefficients one stencil slot reads: W (aW or pW at columns c4..c4+3), S0 / S1 (aS or pS of rows j and
// j + 1), C (aC or pC), and wr = W at column c4 + 4 for the lanes that cannot take it from the next lane (lane 63,
// and the last chunk of a row when the row has no padding).
struct MxCoef {
    float4 W, S0, S1, C;
    float wr;
};
__device__ __forceinline__ bool own_wr(int lane, int c4, const MxGeo& G, int sNx) {
    return (lane == 63 || c4 + 4 == G.CP) && c4 + 4 <= sNx;
}
__device__ __forceinline__ MxCoef coef_load(const MxGeo& G, int sNx, const float* __restrict__ cW,
                                            const float* __restrict__ cS, const float* __restrict__ cC, int f,
                                            float inv) {
    const int lane = threadIdx.x & 63;
    const int fi = f < G.NS ? f : G.NS - 1;
    const int j = srow(fi, inv), c4 = 4 * (fi - j * G.NC);
    const int qc = j * G.CPc + c4;
    MxCoef k;
    k.W = ld4(cW + qc);
    k.S0 = ld4(cS + qc);
    k.S1 = ld4(cS + qc + G.CPc);
    k.C = ld4(cC + qc);
    k.wr = own_wr(lane, c4, G, sNx) ? cW[qc + 4] : 0.0f;
    return k;
}

// IS_A: apply_operator (core/cg2d.py:181-188) term order; else apply_preconditioner (:191-198), on four columns.
template <bool IS_A>
__device__ __forceinline__ float4 stencil4(const MxCoef& k, float wr, float4 Xm, float4 Xc, float4 Xp, float xl,
                                           float xr) {
    float4 o;
    if (IS_A) {
        o.x = k.W.x * xl + k.W.y * Xc.y + k.S0.x * Xm.x + k.S1.x * Xp.x + k.C.x * Xc.x;
        o.y = k.W.y * Xc.x + k.W.z * Xc.z + k.S0.y * Xm.y + k.S1.y * Xp.y + k.C.y * Xc.y;
        o.z = k.W.z * Xc.y + k.W.w * Xc.w + k.S0.z * Xm.z + k.S1.z * Xp.z + k.C.z * Xc.z;
        o.w = k.W.w * Xc.z + wr * xr + k.S0.w * Xm.w + k.S1.w * Xp.w + k.C.w * Xc.w;
    } else {
        o.x = k.C.x * Xc.x + k.W.x * xl + k.W.y * Xc.y + k.S0.x * Xm.x + k.S1.x * Xp.x;
        o.y = k.C.y * Xc.y + k.W.y * Xc.x + k.W.z * Xc.z + k.S0.y * Xm.y + k.S1.y * Xp.y;
        o.z = k.C.z * Xc.z + k.W.z * Xc.y + k.W.w * Xc.w + k.S0.z * Xm.z + k.S1.z * Xp.z;
        o.w = k.C.w * Xc.w + k.W.w * Xc.z + wr * xr + k.S0.w * Xm.w + k.S1.w * Xp.w;
    }
    return o;
}
// FP64 dot-product term of one slot: the four products in column order.
__device__ __forceinline__ double dot4(float4 u, float4 v) {
    const double p0 = (double)u.x * (double)v.x, p1 = (double)u.y * (double)v.y;
    const double p2 = (double)u.z * (double)v.z, p3 = (double)u.w * (double)v.w;
    return ((p0 + p1) + p2) + p3;
}

// Stencil over the LDS tile (IS_A: A32 = (W, S, C) = (aW, aS, aC); else M32 = (pW, pS, pC); coefficient pointers
// at the slot's tile). Wave w takes the 64-slot blocks w, w + nwaves, ...; the coefficients of the next block are
// loaded while the current one is computed (the first block's, `pre`, before the caller's pointwise phase). The
// result of each slot goes to out (compact tile) with its column copies; prods(X, o, pv) gives the slot's NR
// dot-product terms (X the operand with padding zeroed, o the result), reduced in the inner dot-product order; the
// tile sum r is written to *dst[r].
template <bool IS_A, int NR, class F>
__device__ __forceinline__ void mx_stencil(const MxDev& a, const float* __restrict__ cW, const float* __restrict__ cS,
                                           const float* __restrict__ cC, const float* tile, float* __restrict__ out,
                                           MxCoef pre, F&& prods, double* const (&dst)[NR]) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const MxGeo G = a.G;
    const int sNx = a.sNx, sNy = a.sNy;
    const float inv = 1.0f / (float)G.NC;
    const int nblk = (G.NS + 63) / 64;
    MxCoef cur = pre;
    for (int blk = wave; blk < nblk; blk += nwaves) {
        const int bn = blk + nwaves;
        MxCoef nxt = cur;
        if (bn < nblk) nxt = coef_load(G, sNx, cW, cS, cC, bn * 64 + lane, inv);
        const int f = blk * 64 + lane;
        const bool valid = f < G.NS;
        const int fi = valid ? f : G.NS - 1;
        const int j = srow(fi, inv), c4 = 4 * (fi - j * G.NC);
        float wr = __shfl_down(cur.W.x, 1, 64);
        if (own_wr(lane, c4, G, sNx)) wr = cur.wr;
        const int q0 = G.lds(j, c4);
        const float4 Xc = ld4(tile + q0), Xm = ld4(tile + q0 - G.W), Xp = ld4(tile + q0 + G.W);
        const float xl = tile[q0 - 1], xr = tile[q0 + 4];
        const float4 o = mask_pad(stencil4<IS_A>(cur, wr, Xm, Xc, Xp, xl, xr), c4, sNx);
        if (valid) {
            st4(out + j * G.CP + c4, o);
            edge_stash(j, c4, o, sNx, sNy);
        }
        double pv[NR];
        prods(mask_pad(Xc, c4, sNx), o, pv);
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const double w = wave_tree(valid ? pv[k] : 0.0, 0.0, false);
            if (lane == 0) g_rows[k * 128 + blk] = w;
        }
        cur = nxt;
    }
    __syncthreads();
    for (int e = tid; e < 2 * sNy; e += blockDim.x) out[sNy * G.CP + e] = g_edge[e];
    if (tid < NR) {
        double sum = g_rows[tid * 128];
        for (int b = 1; b < nblk; ++b) sum = sum + g_rows[tid * 128 + b];
        *dst[tid] = sum;
    }
    __syncthreads();
}

// This thread's first ring entry of tile t (LDS offset, compact or padded source, sign), loaded ahead of the
// pointwise phase; the rest of the ring (blocks smaller than the ring) is read as it is filled.
struct MxRing {
    int k, src, lds;
    bool neg;
};
// (compact: the compact source index, else the padded one)
__device__ __forceinline__ MxRing ring_entry(const int4 e, bool compact) {
    MxRing r;
    r.src = compact ? e.x : e.y;
    r.lds = e.z & ((1 << 30) - 1);
    r.neg = (e.z >> 30) & 1;
    return r;
}
__device__ __forceinline__ MxRing ring_first(const MxDev& a, int t, bool compact) {
    const int k = a.ring_off[t] + (int)threadIdx.x;
    MxRing r = ring_entry(k < a.ring_off[t + 1] ? a.cring[k] : make_int4(-1, -1, 0, 0), compact);
    r.k = k;
    return r;
}
// Fills the ring of the LDS tile: value(s) of each entry's source (0 without one), signed.
template <class V>
__device__ __forceinline__ void ring_fill(const MxDev& a, int t, const MxRing& r0, bool compact, float* tile,
                                          V&& value) {
    if (r0.k < a.ring_off[t + 1]) {
        const float v = r0.src < 0 ? 0.0f : value(r0.src);
        tile[r0.lds] = r0.neg ? -v : v;
    }
    for (int k = r0.k + blockDim.x; k < a.ring_off[t + 1]; k += blockDim.x) {
        const MxRing r = ring_entry(a.cring[k], compact);
        const float v = r.src < 0 ? 0.0f : value(r.src);
        tile[r.lds] = r.neg ? -v : v;
    }
}

// I1: x = normalised first guess with exchanged halos; r = bn - A x (FP64); partials of err and sumRHS; the
// member's record starts in RST.
__device__ __forceinline__ void mx_I1(const MxDev& a, int m, int t) {
    MJ_UNIT_IDS
    const double rhsNorm = member_rhs_norm(a, m);
    const double nm = a.norm[m];
    const double* b = a.b + m * a.ms;
    double* x = a.x + m * a.ms;
    double* r = a.r64 + m * a.ms;
    const int64_t ob = (int64_t)slot_of(a, m) * a.oms;
    first_guess(a, m, t, rhsNorm);
    MJ_ROWS({
        const double bn0 = b[p0] * nm * rhsNorm;
        const double r_0 = bn0 - op_A(a.aW + ob, a.aS + ob, a.aC + ob, x, p0, nx);
        r[p0] = r_0;
        double bn1 = 0.0, r_1 = 0.0;
        if (has1) {
            bn1 = b[p1] * nm * rhsNorm;
            r_1 = bn1 - op_A(a.aW + ob, a.aS + ob, a.aC + ob, x, p1, nx);
            r[p1] = r_1;
        }
        put_row<MJ_SUM_TREE>(0, jl, lane, r_0 * r_0, r_1 * r_1, has1, nullptr, sNx, sNy);
        put_row<MJ_SUM_TREE>(1, jl, lane, bn0, bn1, has1, nullptr, sNx, sNy);
    })
    store_tile_sums<MJ_SUM_TREE>(a.perr64, a.psum, m * a.T + t, lane, wave, nwaves, nullptr, sNx, sNy);
    if (t == 0 && tid == 0) {
        a.rec[m] = MxRec{MX_RST, 0, 0, 0, 0, 0, 0, 0, 0.0, 0.0, 0.0};
        a.mdone[m] = -1;
    }
}

// Finish: x /= rhsNorm on the interior; the halo points the true-residual steps used for x's ring get back the
// exchanged normalised first guess.
__device__ __forceinline__ void mx_finish(const MxDev& a, int m, int t) {
    MJ_UNIT_IDS
    const double rhsNorm = member_rhs_norm(a, m);
    double* x = a.x + m * a.ms;
    MJ_ROWS({
        x[p0] = x[p0] / rhsNorm;
        if (has1) x[p1] = x[p1] / rhsNorm;
    })
    exchanged_first_guess(a, m, t, rhsNorm);
}

// Residual replacement before inner phase B(n): r32_{n-1} = r / sigma on the interior (compact, padding 0) and the
// ring (LDS, from the sources' r), d = 0, z = M32 r32, partial of eta(n) = z.r32.
__device__ __forceinline__ void mx_restart(const MxDev& a, int m, int t, double sigma, int n) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sNx = a.sNx, sNy = a.sNy, nx = a.nx, OL = a.OL, tbase = t * a.plane;
    float* tile = reinterpret_cast<float*>(mj_lds);
    const MxGeo G = a.G;
    const float inv = 1.0f / (float)G.NC;
    const double* r = a.r64 + m * a.ms;
    const int64_t cb = cbase(a, m, t), oc = coef_base(a, m, t);
    const float *cW = a.pW32 + oc, *cS = a.pS32 + oc, *cC = a.pC32 + oc;
    const MxCoef pre = coef_load(G, sNx, cW, cS, cC, wave * 64 + lane, inv);
    const MxRing rr = ring_first(a, t, false);
    float* r32 = R32(a, n - 1) + cb;
    float* d = a.d + cb;
    for (int f = tid; f < G.NS; f += blockDim.x) {
        const int e = 4 * f, j = srow(f, inv), c4 = e - j * G.CP;
        const int p = MJ_ROW(j) + c4;
        const double2 R0 = ld2(r + p), R1 = ld2(r + p + 2);
        const float4 v = mask_pad(make_float4((float)(R0.x / sigma), (float)(R0.y / sigma), (float)(R1.x / sigma),
                                              (float)(R1.y / sigma)), c4, sNx);
        st4(r32 + e, v);
        lds_put(tile, G, j, c4, v, sNx);
        st4(d + e, make_float4(0.f, 0.f, 0.f, 0.f));
    }
    ring_fill(a, t, rr, false, tile, [&](int s) { return (float)(r[s] / sigma); });
    __syncthreads();
    edges_from_tile(a, tile, r32);
    double* const dst[1] = {a.peta + (n & 1) * a.B * a.T + m * a.T + t};  // eta(n)
    mx_stencil<false, 1>(
        a, cW, cS, cC, tile, a.z + cb, pre, [](float4 R, float4 zz, double(&pv)[1]) { pv[0] = dot4(zz, R); }, dst);
}

// Solution update at a replacement due in B(n): the deferred d += alpha_{n-2} s_{n-2} (def; s_{n-2} is still in
// s_n's buffer) and the pending d += alpha_{n-1} s_{n-1} (n > 1), then x = x + sigma * d on the interior (FP64).
__device__ __forceinline__ void mx_update(const MxDev& a, int m, int t, int n, double sigma, double alpha64, bool def,
                                          double ap64) {
    const int tid = threadIdx.x;
    const int sNx = a.sNx, nx = a.nx, OL = a.OL, tbase = t * a.plane;
    const MxGeo G = a.G;
    const float inv = 1.0f / (float)G.NC;
    double* x = a.x + m * a.ms;
    const int64_t cb = cbase(a, m, t);
    const float* d = a.d + cb;
    const float* so = S32(a, n - 1) + cb;
    const float* so2 = S32(a, n) + cb;
    const float alpha = (float)alpha64, ap = (float)ap64;
    for (int f = tid; f < G.NS; f += blockDim.x) {
        const int e = 4 * f, j = srow(f, inv), c4 = e - j * G.CP;
        float4 D = ld4(d + e);
        if (def) {
            const float4 S2 = ld4(so2 + e);
            D.x = D.x + ap * S2.x;
            D.y = D.y + ap * S2.y;
            D.z = D.z + ap * S2.z;
            D.w = D.w + ap * S2.w;
        }
        if (n > 1) {
            const float4 S = ld4(so + e);
            D.x = D.x + alpha * S.x;
            D.y = D.y + alpha * S.y;
            D.z = D.z + alpha * S.z;
            D.w = D.w + alpha * S.w;
        }
        const int p = MJ_ROW(j) + c4;
        const double2 X0 = ld2(x + p);
        st2(x + p, make_double2(X0.x + sigma * (double)D.x, X0.y + sigma * (double)D.y));
        if (c4 + 2 < sNx) {  // sNx is even: a slot holds 2 or 4 interior columns
            const double2 X1 = ld2(x + p + 2);
            st2(x + p + 2, make_double2(X1.x + sigma * (double)D.z, X1.y + sigma * (double)D.w));
        }
    }
}

// True residual: x's ring from its sources (the updated interior of other tiles), r = bn - A x (FP64), partials
// of err.
__device__ __forceinline__ void mx_true(const MxDev& a, int m, int t) {
    MJ_UNIT_IDS
    MJ_STRIP
    const double rhsNorm = member_rhs_norm(a, m);
    const double nm = a.norm[m];
    const double* b = a.b + m * a.ms;
    double* x = a.x + m * a.ms;
    double* r = a.r64 + m * a.ms;
    const int64_t ob = (int64_t)slot_of(a, m) * a.oms;
    const double *__restrict__ aW = a.aW + ob, *__restrict__ aS = a.aS + ob, *__restrict__ aC = a.aC + ob;
    for (int k = a.ring_off[t] + tid; k < a.ring_off[t + 1]; k += blockDim.x) {
        const int s = a.ring_src[k];
        if (s < 0) continue;  // no source: x keeps its halo value
        const double v = x[s];
        x[a.ring_dst[k]] = a.ring_sign[k] < 0 ? -v : v;
    }
    __syncthreads();
    if (j0 < j1) {
        int p = MJ_ROW(j0) + L.c;
        double2 Xm = ld2(x + p - nx), Xc = ld2(x + p), AS0 = ld2(aS + p);
        MJ_UNROLL(MJ_SWEEP_UNROLL) for (int j = j0; j < j1; ++j, p += nx) {
            const int prow = MJ_ROW(j);
            const double2 Xp = ld2(x + p + nx), AS1 = ld2(aS + p + nx), AW = ld2(aW + p), AC = ld2(aC + p);
            const double2 Bv = ld2(b + p);
            const double hx = halo_of(x, prow, L, sNx), hw = halo_of(aW, prow, L, sNx);
            const double2 q = pair_A(AW, right_of(AW, hw, L), AS0, AS1, AC, Xm, Xc, Xp, left_of(Xc, hx, L),
                                     right_of(Xc, hx, L));
            double2 rv;
            rv.x = Bv.x * nm * rhsNorm - q.x;
            rv.y = Bv.y * nm * rhsNorm - q.y;
            if (L.act) st2(r + p, rv);
            double2 e = make_double2(rv.x * rv.x, rv.y * rv.y);
            if (!L.act) e = make_double2(0.0, 0.0);
            put_row2<MJ_SUM_TREE>(0, j, L, e, nullptr, sNx, sNy);
            Xm = Xc;
            Xc = Xp;
            AS0 = AS1;
        }
    }
    store_tile_sums<MJ_SUM_TREE>(a.perr64, nullptr, m * a.T + t, lane, wave, nwaves, nullptr, sNx, sNy);
}

// Inner phase B(n): the solution updates, every other iteration two at once (d is read and written half as
// often; the same sums in the same order): with both (both, after a deferring B(n-1)) d += alpha_{n-2} s_{n-2}
// (s_{n-2} read from s_n's buffer before s_n overwrites it), then d += alpha_{n-1} s_{n-1}; otherwise the update of
// iteration n-1 is deferred to B(n+1), or there is none (after a replacement). Then s_n = z + beta s_{n-1} (s_1 = z)
// on the interior (compact, flat), s_n on the ring (LDS), q = A32 s_n, sigma_n partial (s_n, q).
__device__ __forceinline__ void mx_IB(const MxDev& a, int m, int t, int n, double beta64, double alpha64, bool both,
                                      double ap64) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sNx = a.sNx, sNy = a.sNy;
    float* tile = reinterpret_cast<float*>(mj_lds);
    const MxGeo G = a.G;
    const float inv = 1.0f / (float)G.NC;
    const int64_t cm = m * a.cs, cb = cbase(a, m, t), ot = (int64_t)t * G.TS, oc = coef_base(a, m, t);
    const float *cW = a.aW32 + oc, *cS = a.aS32 + oc, *cC = a.aC32 + oc;
    const MxCoef pre = coef_load(G, sNx, cW, cS, cC, wave * 64 + lane, inv);
    const MxRing rr = ring_first(a, t, true);
    const float* __restrict__ z = a.z + cm;
    const float* __restrict__ so = S32(a, n - 1) + cm;
    float* __restrict__ sn = S32(a, n) + cb;
    float* __restrict__ d = a.d + cb;
    const float beta = (float)beta64, alpha = (float)alpha64, ap = (float)ap64;
    const bool first = n == 1;
    for (int f = tid; f < G.NS; f += blockDim.x) {
        const int e = 4 * f;
        float4 v = ld4(z + ot + e);
        if (!first) {
            const float4 O = ld4(so + ot + e);
            if (both) {
                const float4 D = ld4(d + e), S2 = ld4(sn + e);
                float4 dn;
                dn.x = D.x + ap * S2.x;
                dn.y = D.y + ap * S2.y;
                dn.z = D.z + ap * S2.z;
                dn.w = D.w + ap * S2.w;
                dn.x = dn.x + alpha * O.x;
                dn.y = dn.y + alpha * O.y;
                dn.z = dn.z + alpha * O.z;
                dn.w = dn.w + alpha * O.w;
                st4(d + e, dn);
            }
            v.x = v.x + beta * O.x;
            v.y = v.y + beta * O.y;
            v.z = v.z + beta * O.z;
            v.w = v.w + beta * O.w;
        }
        st4(sn + e, v);
        const int j = srow(f, inv), c4 = e - j * G.CP;
        lds_put(tile, G, j, c4, v, sNx);
    }
    ring_fill(a, t, rr, true, tile, [&](int s) { return first ? z[s] : z[s] + beta * so[s]; });
    __syncthreads();
    edges_from_tile(a, tile, sn);
    double* const dst[1] = {a.psig + (n & 1) * a.B * a.T + m * a.T + t};
    mx_stencil<true, 1>(
        a, cW, cS, cC, tile, a.q + cb, pre, [](float4 X, float4 q, double(&pv)[1]) { pv[0] = dot4(X, q); }, dst);
}

// Inner phase C(n): r32_n = r32_{n-1} - alpha q on the interior (compact, flat) and the ring (LDS); z = M32 r32_n;
// partials of err(n) = (r32_n, r32_n) and eta(n+1) = (z, r32_n).
__device__ __forceinline__ void mx_IC(const MxDev& a, int m, int t, int n, double alpha64) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sNx = a.sNx, sNy = a.sNy;
    float* tile = reinterpret_cast<float*>(mj_lds);
    const MxGeo G = a.G;
    const float inv = 1.0f / (float)G.NC;
    const int64_t cm = m * a.cs, cb = cbase(a, m, t), ot = (int64_t)t * G.TS, oc = coef_base(a, m, t);
    const float *cW = a.pW32 + oc, *cS = a.pS32 + oc, *cC = a.pC32 + oc;
    const MxCoef pre = coef_load(G, sNx, cW, cS, cC, wave * 64 + lane, inv);
    const MxRing rr = ring_first(a, t, true);
    const float* __restrict__ qv = a.q + cm;
    const float* __restrict__ rin = R32(a, n - 1) + cm;
    float* __restrict__ rout = R32(a, n) + cb;
    const float alpha = (float)alpha64;
    for (int f = tid; f < G.NS; f += blockDim.x) {
        const int e = 4 * f;
        const float4 Q = ld4(qv + ot + e), R = ld4(rin + ot + e);
        float4 rn;
        rn.x = R.x - alpha * Q.x;
        rn.y = R.y - alpha * Q.y;
        rn.z = R.z - alpha * Q.z;
        rn.w = R.w - alpha * Q.w;
        st4(rout + e, rn);
        const int j = srow(f, inv), c4 = e - j * G.CP;
        lds_put(tile, G, j, c4, rn, sNx);
    }
    ring_fill(a, t, rr, true, tile, [&](int s) { return rin[s] - alpha * qv[s]; });
    __syncthreads();
    edges_from_tile(a, tile, rout);
    const int BT = a.B * a.T;
    double* const dst[2] = {a.perr + (n & 1) * BT + m * a.T + t, a.peta + ((n + 1) & 1) * BT + m * a.T + t};
    mx_stencil<false, 2>(
        a, cW, cS, cC, tile, a.z + cb, pre,
        [](float4 R, float4 zz, double(&pv)[2]) {
            pv[0] = dot4(R, R);
            pv[1] = dot4(zz, R);
        },
        dst);
}

__global__ void __launch_bounds__(MJ_MX_THREADS) mx_init_kernel(MxDev a, int ph) {
    int m, t;
    if (!phase_unit(a.B, a.T, m, t)) return;
    if (ph == 0)
        unit_I0(a, m, t);
    else
        mx_I1(a, m, t);
}

// The record of inner phase B(n + 1) after C(n).
__device__ __forceinline__ MxRec mx_next_ib(const MxRec& r) {
    MxRec nr = r;
    nr.mode = MX_IB;
    nr.n = r.n + 1;
    nr.iters = r.iters + 1;
    nr.fresh = 0;
    return nr;
}
// The record of C(n) after B(n) (stop: a replacement instead): a pending update (pend) alternately deferred
// (alpha_{n-1} kept in the record) and applied with the deferred one.
__device__ __forceinline__ MxRec mx_after_ib(const MxRec& r, bool stop, bool pend, double alpha_prev) {
    MxRec nr = r;
    nr.mode = stop ? MX_TRUE : MX_IC;
    nr.fresh = 0;
    if (!stop && pend) {
        nr.def = r.def ? 0 : 1;
        nr.ap = r.def ? 0.0 : alpha_prev;
    }
    return nr;
}
// Whether B(n) replaces the recursive residual instead (err(n-1) the recursive err in units of sigma^2): a cap
// reached, the recursive residual fallen by inner_tolerance since the last replacement, or below the target.
__device__ __forceinline__ bool mx_replace_due(const MxDev& a, const MxRec& r, double errp) {
    if (r.iters >= a.max_iters || r.n - r.n0 >= a.inner_max) return true;
    return errp < fmax(a.itol_sq * r.elast, a.tol_sq / (r.sigma * r.sigma));
}

// Tile-order sum of T <= 64 partials held by the lanes of one wave (lane t: tile t); every lane gets it.
__device__ __forceinline__ double lane_tile_sum(double v, int T) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    double sum = 0.0;
    for (int i = 0; i < T; ++i)
        sum = sum + __hiloint2double(__builtin_amdgcn_readlane(hi, i), __builtin_amdgcn_readlane(lo, i));
    return sum;
}

// One step of every member's state machine.
#ifndef MJ_PH_WAVES_PER_EU
#define MJ_PH_WAVES_PER_EU 7
#endif  // 72 VGPRs (one 8-byte spill outside the inner phases); 4 resident blocks of up to 448 threads per CU
__global__ void __launch_bounds__(MJ_MX_THREADS) __attribute__((amdgpu_waves_per_eu(MJ_PH_WAVES_PER_EU)))
mx_step_kernel(MxDev a, int k) {
    int m, t;
    if (!phase_unit(a.B, a.T, m, t)) return;
    const int st = a.mdone[m];
    if (st >= 0 && st < k) return;  // finished in an earlier step
    const int BT = a.B * a.T;
    const MxRec r0 = a.rec[((k - 1) & 1) * a.B + m];
    if ((r0.mode == MX_IB || r0.mode == MX_IC) && a.T <= 64) {
        // inner phases: every wave derives the action itself from the partials its lanes load (the same loads and
        // sums in every wave, so every wave takes the same branch) and goes on without a barrier
        const int lane = threadIdx.x & 63, o = m * a.T, n = r0.n, cp = n & 1, pp = (n - 1) & 1;
        const bool tl = lane < a.T;
        if (r0.mode == MX_IC) {
            const double E = tl ? a.peta[cp * BT + o + lane] : 0.0, S = tl ? a.psig[cp * BT + o + lane] : 0.0;
            const double alpha = lane_tile_sum(E, a.T) / lane_tile_sum(S, a.T);
            if (t == 0 && threadIdx.x == 0) a.rec[(k & 1) * a.B + m] = mx_next_ib(r0);
            mx_IC(a, m, t, n, alpha);
            return;
        }
        const bool pend = !r0.fresh;  // n > 1 whenever an update is pending
        bool stop = false;
        double beta = 0.0, alpha_prev = 0.0;
        if (n > 1) {
            // eta(n), eta(n-1) (beta); with an update pending also err(n-1) (replacement rule) and sigma(n-1)
            // (alpha_{n-1} of the pending solution update: the value C(n-1) used)
            const double ER = tl ? a.perr[pp * BT + o + lane] : 0.0, EN = tl ? a.peta[cp * BT + o + lane] : 0.0;
            const double EP = tl ? a.peta[pp * BT + o + lane] : 0.0, SP = tl ? a.psig[pp * BT + o + lane] : 0.0;
            const double etan = lane_tile_sum(EN, a.T), etap = lane_tile_sum(EP, a.T);
            beta = etan / etap;
            if (pend) {
                const double errp = lane_tile_sum(ER, a.T), sigp = lane_tile_sum(SP, a.T);
                stop = mx_replace_due(a, r0, errp);
                alpha_prev = etap / sigp;
            }
        }
        if (t == 0 && threadIdx.x == 0) a.rec[(k & 1) * a.B + m] = mx_after_ib(r0, stop, pend, alpha_prev);
        if (stop)
            mx_update(a, m, t, n, r0.sigma, alpha_prev, r0.def, r0.ap);
        else
            mx_IB(a, m, t, n, beta, alpha_prev, pend && r0.def, r0.ap);
        return;
    }
    // wave 0 derives the member's action (every lane alike), lane 0 publishes it. The member's record and every
    // partial a decision may need are loaded at once (lane t: tile t), so the decision costs one memory latency;
    // the tile-order sums are then formed from lane reads.
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const MxRec r = a.rec[((k - 1) & 1) * a.B + m];
        const bool lanes = a.T <= 64;  // partials held by the lanes (else summed from memory)
        double P[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // perr64, psum, perr[0..1], peta[0..1], psig[0..1]
        if (lanes && lane < a.T) {
            const int i = m * a.T + lane;
            P[0] = a.perr64[i];
            P[1] = a.psum[i];
            P[2] = a.perr[i];
            P[3] = a.perr[BT + i];
            P[4] = a.peta[i];
            P[5] = a.peta[BT + i];
            P[6] = a.psig[i];
            P[7] = a.psig[BT + i];
        }
        // tile-order sum of partial array `p` (one of the above, v its lane values)
        auto tsum = [&](double v, const double* p) {
            if (!lanes) return tile_order_sum(p, a.T);
            double sum = 0.0;
            const int lo = __double2loint(v), hi = __double2hiint(v);
            for (int i = 0; i < a.T; ++i)
                sum = sum + __hiloint2double(__builtin_amdgcn_readlane(hi, i), __builtin_amdgcn_readlane(lo, i));
            return sum;
        };
        MxRec nr = r;
        int act = 0;
        double c = 0.0, alpha_prev = 0.0;
        const int o = m * a.T;
        switch (r.mode) {
            case MX_RST: {
                const double e = tsum(P[0], a.perr64 + o);
                c = e;
                if (t == 0 && r.outer == 0) {
                    const double sr = tsum(P[1], a.psum + o);
                    if (lane == 0) {
                        a.diag[m].firstResidual = sqrt(e);
                        a.diag[m].sumRHS = sr;
                        a.diag[m].rhsMax = member_rhs_max(a, m);
                        a.diag[m].launches = 0;
                    }
                }
                if (e < a.tol_sq || r.iters >= a.max_iters || r.outer >= a.max_outer) {
                    act = AX_FINISH;
                } else {
                    // the first residual sets the units of r32 and starts at n = 1; a replacement keeps both
                    act = AX_RESTART;
                    const double sg = r.outer == 0 ? sqrt(e) : r.sigma;
                    const int n1 = r.outer == 0 ? 1 : r.n;
                    nr = MxRec{MX_IB, n1, r.iters, r.outer + 1, n1, 1, 0, 0, sg, e / (sg * sg), 0.0};
                }
                break;
            }
            case MX_IB: {
                bool stop = false;
                if (r.n > 1) {
                    // eta(n), eta(n-1) (beta); with an update pending err(n-1) (replacement rule) and sigma(n-1)
                    // (alpha_{n-1} of the pending solution update: the value C(n-1) used)
                    const int cp = r.n & 1, pp = (r.n - 1) & 1;
                    const double etan = tsum(cp ? P[5] : P[4], a.peta + cp * BT + o);
                    const double etap = tsum(pp ? P[5] : P[4], a.peta + pp * BT + o);
                    c = etan / etap;
                    if (!r.fresh) {
                        const double errp = tsum(pp ? P[3] : P[2], a.perr + pp * BT + o);
                        const double sigp = tsum(pp ? P[7] : P[6], a.psig + pp * BT + o);
                        stop = mx_replace_due(a, r, errp);
                        alpha_prev = etap / sigp;
                    }
                }
                act = stop ? AX_UPD : r.fresh ? AX_IBF : r.def ? AX_IB : AX_IBD;
                nr = mx_after_ib(r, stop, !r.fresh, alpha_prev);
                break;
            }
            case MX_IC: {
                const int cp = r.n & 1;
                const double etan = tsum(cp ? P[5] : P[4], a.peta + cp * BT + o);
                const double sign = tsum(cp ? P[7] : P[6], a.psig + cp * BT + o);
                act = AX_IC;
                c = etan / sign;
                nr = mx_next_ib(r);
                break;
            }
            default:
                act = AX_TRUE;
                nr.mode = MX_RST;
                break;
        }
        if (threadIdx.x == 0) {
            g_bcast[3] = alpha_prev;
            g_ibcast = act;
            g_bcast[0] = c;
            g_bcast[1] = nr.sigma;
            g_bcast[2] = (double)(act == AX_RESTART ? nr.n : r.n);  // the step's n
            g_bcast[4] = r.ap;                                      // a deferred alpha (AX_IB, AX_UPD)
            g_bcast[5] = (double)r.def;
            if (t == 0) {
                a.rec[(k & 1) * a.B + m] = nr;
                if (act == AX_FINISH) {
                    a.diag[m].lastResidual = sqrt(c);
                    a.diag[m].numIters = r.iters;
                    a.diag[m].outerIters = r.outer;
                    a.diag[m].converged = c < a.tol_sq ? 1 : 0;
                    a.mdone[m] = k;
                    atomicSub(a.nactive, 1);
                }
            }
        }
    }
    __syncthreads();
    const int act = g_ibcast;
    const double c = g_bcast[0], sigma = g_bcast[1], alpha_prev = g_bcast[3], ap = g_bcast[4];
    const int n = (int)g_bcast[2];
    const bool def = g_bcast[5] != 0.0;
    switch (act) {
        case AX_FINISH: mx_finish(a, m, t); break;
        case AX_RESTART: mx_restart(a, m, t, sigma, n); break;
        case AX_UPD: mx_update(a, m, t, n, sigma, alpha_prev, def, ap); break;
        case AX_IB: mx_IB(a, m, t, n, c, alpha_prev, true, ap); break;
        case AX_IBF:
        case AX_IBD: mx_IB(a, m, t, n, c, alpha_prev, false, 0.0); break;
        case AX_IC: mx_IC(a, m, t, n, c); break;
        default: mx_true(a, m, t); break;
    }
}

// FP32 copies of the six coefficient arrays of every slot in the coefficient layout of mx_geo:
// dst[k][slot][t][j * CPc + i] = src_k[slot][t][OL + j][OL + i] for j <= sNy, i <= sNx, 0 in the padding columns.
__global__ void mx_convert_kernel(const double* aW, const double* aS, const double* aC, const double* pW,
                                  const double* pS, const double* pC, int64_t oms, int nslots, int T, int OL, int nx,
                                  int plane, int sNx, MxGeo G, float* dst) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    const int64_t per_slot = (int64_t)T * G.TSc;
    if (q >= per_slot) return;
    const int t = (int)(q / G.TSc), r = (int)(q - (int64_t)t * G.TSc), j = r / G.CPc, i = r - j * G.CPc;
    const bool inside = i <= sNx;
    const int64_t src = s * oms + (int64_t)t * plane + (int64_t)(OL + j) * nx + OL + (inside ? i : 0);
    const double* from[6] = {aW, aS, aC, pW, pS, pC};
#pragma unroll
    for (int k = 0; k < 6; ++k) dst[((int64_t)k * nslots + s) * per_slot + q] = inside ? (float)from[k][src] : 0.0f;
}

__global__ void mx_reset_kernel(int* nactive, int B) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *nactive = B;
}

// The launch count of the step loop into every member's record (written after the loop).
__global__ void mx_launches_kernel(mj_cg2d_diag* diag, int B, int launches) {
    for (int m = threadIdx.x; m < B; m += blockDim.x) diag[m].launches = launches;
}

}  // namespace mj

using namespace mj;

namespace {

struct MxLayout {
    size_t r64, r0, r1, s0, s1, d, z, q, coef, pmax, psum, perr64, perr, peta, psig, rec, mdone, misc, total;
};

int nslots_of(const mj_cg2d_config* c, const mj_cg2d_ops* ops, const mj_cg2d_mixed_config* mc) {
    if (ops && ops->member_stride == 0) return 1;
    return mc->op_slots > 0 ? mc->op_slots : c->batch;
}

MxLayout mx_layout(const mj_cg2d_config* c, int nslots) {
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    MxLayout L{};
    const size_t n = (size_t)c->batch * c->member_stride;
    const MxGeo G = mx_geo(c->sNx, c->sNy);
    const size_t nc = (size_t)c->batch * c->ntiles * G.TS;  // compact vectors
    const size_t tp = (size_t)c->ntiles * G.TSc;             // one coefficient copy of one slot
    const size_t bt = al((size_t)c->batch * c->ntiles * 8);
    size_t o = 0;
    L.r64 = o; o += al(n * 8);
    L.r0 = o; o += al(nc * 4);
    L.r1 = o; o += al(nc * 4);
    L.s0 = o; o += al(nc * 4);
    L.s1 = o; o += al(nc * 4);
    L.d = o; o += al(nc * 4);
    L.z = o; o += al(nc * 4);
    L.q = o; o += al(nc * 4);
    L.coef = o; o += al(6 * (size_t)nslots * tp * 4);
    L.pmax = o; o += bt;
    L.psum = o; o += bt;
    L.perr64 = o; o += bt;
    L.perr = o; o += 2 * bt;
    L.peta = o; o += 2 * bt;
    L.psig = o; o += 2 * bt;
    L.rec = o; o += al(2 * (size_t)c->batch * sizeof(MxRec));
    L.mdone = o; o += al((size_t)c->batch * 4);
    L.misc = o; o += 256;
    L.total = o;
    return L;
}

int check_mixed(const mj_cg2d_config* c, const mj_cg2d_mixed_config* mc) {
    MJ_REQUIRE(c && mc, "cg2d mixed: null config");
    MJ_REQUIRE(c->batch >= 1 && c->ntiles >= 1, "cg2d mixed: bad batch/layout");
    MJ_REQUIRE(c->sNx >= 64 && c->sNx <= 128 && c->sNy >= 64 && c->sNy <= 128 && c->sNx % 2 == 0 &&
                   c->sNy % 2 == 0 && c->OL >= 2 && c->OL % 2 == 0,
               "cg2d mixed: sNx, sNy must be even in 64..128, OL even and >= 2");
    const int64_t plane = (int64_t)(c->sNx + 2 * c->OL) * (c->sNy + 2 * c->OL);
    MJ_REQUIRE(c->member_stride >= c->ntiles * plane && c->member_stride % 2 == 0,
               "cg2d mixed: member_stride must be even and >= ntiles * ny * nx");
    MJ_REQUIRE((int64_t)c->batch * c->member_stride < ((int64_t)1 << 31), "cg2d mixed: batch * member_stride too large");
    MJ_REQUIRE(c->sum_order == MJ_SUM_TREE, "cg2d mixed: only MJ_SUM_TREE dot products");
    MJ_REQUIRE(c->max_iters >= 0, "cg2d mixed: max_iters < 0");
    MJ_REQUIRE(c->tolerance > 0.0, "cg2d mixed: tolerance must be > 0");
    MJ_REQUIRE(c->launch_mode == MJ_CG_LAUNCH_AUTO || c->launch_mode == MJ_CG_LAUNCH_PHASES ||
                   c->launch_mode == MJ_CG_LAUNCH_PHASES_FIXED,
               "cg2d mixed: launch_mode must be AUTO, PHASES or PHASES_FIXED");
    MJ_REQUIRE(c->block_threads == 0 || (c->block_threads % 64 == 0 && c->block_threads >= 64 &&
                                         c->block_threads <= MJ_MX_THREADS),
               "cg2d mixed: block_threads must be a multiple of 64 in 64..%d", MJ_MX_THREADS);
    MJ_REQUIRE(c->poll_iters >= 0, "cg2d mixed: poll_iters < 0");
    MJ_REQUIRE(mc->inner_tolerance > 0.0 && mc->inner_tolerance < 1.0, "cg2d mixed: inner_tolerance must be in (0, 1)");
    MJ_REQUIRE(mc->inner_max_iters >= 1 && mc->max_outer >= 1, "cg2d mixed: inner_max_iters, max_outer must be >= 1");
    MJ_REQUIRE(mc->op_slots >= 0, "cg2d mixed: op_slots < 0");
    return MJ_OK;
}

}  // namespace

extern "C" size_t mj_cg2d_mixed_workspace_bytes(const mj_cg2d_config* cfg, const mj_cg2d_mixed_config* mcfg,
                                                int64_t ops_member_stride) {
    if (check_mixed(cfg, mcfg)) return 0;
    mj_cg2d_ops o{};
    o.member_stride = ops_member_stride;
    return mx_layout(cfg, nslots_of(cfg, &o, mcfg)).total;
}

extern "C" int mj_cg2d_solve_mixed(const mj_halo_plan* plan_T, const mj_cg2d_ops* ops, const double* b,
                                   const double* x0, double* x, const double* cg2dNorm, const mj_cg2d_config* cfg,
                                   const mj_cg2d_mixed_config* mcfg, void* workspace, size_t workspace_bytes,
                                   mj_cg2d_diag* diag, void* stream) {
    int rc = check_mixed(cfg, mcfg);
    if (rc) return rc;
    MJ_REQUIRE(plan_T && ops && b && x0 && x && cg2dNorm && workspace && diag, "mj_cg2d_solve_mixed: null argument");
    MJ_REQUIRE(ops->aW2d && ops->aS2d && ops->aC2d && ops->pW && ops->pS && ops->pC,
               "mj_cg2d_solve_mixed: null operator");
    MJ_REQUIRE(x != b && x != x0, "mj_cg2d_solve_mixed: x must not alias b or x0");
    MJ_REQUIRE(((uintptr_t)workspace & 255) == 0, "mj_cg2d_solve_mixed: workspace must be 256-byte aligned");
    MJ_REQUIRE(ops->member_stride % 2 == 0, "mj_cg2d_solve_mixed: operator member_stride must be even");
    MJ_REQUIRE(ops->member_stride == 0 || ops->op_index == nullptr || mcfg->op_slots > 0,
               "mj_cg2d_solve_mixed: op_slots (the number of operator slots) is required with op_index");
    for (const void* q : {(const void*)x, (const void*)b, (const void*)ops->aW2d, (const void*)ops->aS2d,
                          (const void*)ops->aC2d})
        MJ_REQUIRE(((uintptr_t)q & 15) == 0, "mj_cg2d_solve_mixed: x, b and the operator must be 16-byte aligned");
    const int nslots = nslots_of(cfg, ops, mcfg);
    const MxLayout L = mx_layout(cfg, nslots);
    MJ_REQUIRE(workspace_bytes >= L.total, "mj_cg2d_solve_mixed: workspace %zu < %zu bytes", workspace_bytes,
               L.total);
    mj_halo_plan* P = const_cast<mj_halo_plan*>(plan_T);
    MJ_REQUIRE(P->ntiles == cfg->ntiles, "mj_cg2d_solve_mixed: plan has %d tiles, config %d", P->ntiles,
               cfg->ntiles);
    if ((rc = ensure_cg_tables(P, cfg->sNx, cfg->sNy, cfg->OL))) return rc;

    char* w = (char*)workspace;
    MxDev a{};
    a.B = cfg->batch;
    a.T = cfg->ntiles;
    a.sNx = cfg->sNx;
    a.sNy = cfg->sNy;
    a.OL = cfg->OL;
    a.nx = cfg->sNx + 2 * cfg->OL;
    a.ny = cfg->sNy + 2 * cfg->OL;
    a.plane = a.nx * a.ny;
    a.ms = cfg->member_stride;
    a.oms = ops->member_stride;
    a.G = mx_geo(a.sNx, a.sNy);
    const int64_t tp = (int64_t)a.T * a.G.TSc;  // one FP32 coefficient copy of one slot
    a.cms = ops->member_stride == 0 ? 0 : tp;
    a.aW = ops->aW2d;
    a.aS = ops->aS2d;
    a.aC = ops->aC2d;
    float* coef = (float*)(w + L.coef);
    a.aW32 = coef;
    a.aS32 = coef + 1 * nslots * tp;
    a.aC32 = coef + 2 * nslots * tp;
    a.pW32 = coef + 3 * nslots * tp;
    a.pS32 = coef + 4 * nslots * tp;
    a.pC32 = coef + 5 * nslots * tp;
    a.b = b;
    a.x0 = x0;
    a.x = x;
    a.norm = cg2dNorm;
    a.tol_sq = cfg->tolerance * cfg->tolerance;
    a.itol_sq = mcfg->inner_tolerance * mcfg->inner_tolerance;
    a.max_iters = cfg->max_iters;
    a.inner_max = mcfg->inner_max_iters;
    a.max_outer = mcfg->max_outer;
    a.full_off = P->cg_full_off;
    a.full_dst = P->cg_full_dst;
    a.full_src = P->cg_full_src;
    a.full_sign = P->cg_full_sign;
    a.ring_off = P->cg_ring_off;
    a.ring_dst = P->cg_ring_dst;
    a.ring_src = P->cg_ring_src;
    a.ring_sign = P->cg_ring_sign;
    a.cring = P->cg_cring;
    a.cs = (int64_t)a.T * a.G.TS;
    a.r64 = (double*)(w + L.r64);
    a.r0 = (float*)(w + L.r0);
    a.r1 = (float*)(w + L.r1);
    a.s0 = (float*)(w + L.s0);
    a.s1 = (float*)(w + L.s1);
    a.d = (float*)(w + L.d);
    a.z = (float*)(w + L.z);
    a.q = (float*)(w + L.q);
    a.pmax = (double*)(w + L.pmax);
    a.psum = (double*)(w + L.psum);
    a.perr64 = (double*)(w + L.perr64);
    a.perr = (double*)(w + L.perr);
    a.peta = (double*)(w + L.peta);
    a.psig = (double*)(w + L.psig);
    a.op_index = ops->op_index;
    a.rec = (MxRec*)(w + L.rec);
    a.mdone = (int32_t*)(w + L.mdone);
    a.nactive = (int*)(w + L.misc);
    a.diag = diag;

    hipStream_t st = as_stream(stream);
    int dev = 0, ncu = 0;
    MJ_HIP(hipGetDevice(&dev));
    MJ_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int threads = cfg->block_threads ? cfg->block_threads
                                           : ((int64_t)a.B * a.T <= ncu ? kMxSmallThreads : kMxLargeThreads);
    const int nunits = a.B * a.T;
    const dim3 grid((unsigned)((nunits + kXcds - 1) / kXcds * kXcds)), block((unsigned)threads);
    // the step kernel's LDS tile of the stencil operand (FP32)
    const size_t lds = mx_lds_bytes(cfg->sNx, cfg->sNy);
    if (lds > 65536)
        MJ_HIP(hipFuncSetAttribute((const void*)mx_step_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(mx_reset_kernel, dim3(1), dim3(64), 0, st, a.nactive, a.B);
    MJ_HIP(hipGetLastError());
    hipLaunchKernelGGL(mx_convert_kernel, dim3((unsigned)((tp + 255) / 256), (unsigned)nslots), dim3(256), 0, st,
                       ops->aW2d, ops->aS2d, ops->aC2d, ops->pW, ops->pS, ops->pC, ops->member_stride, nslots, a.T,
                       a.OL, a.nx, a.plane, a.sNx, a.G, coef);
    MJ_HIP(hipGetLastError());
    for (int ph = 0; ph < 2; ++ph) {
        hipLaunchKernelGGL(mx_init_kernel, grid, block, 0, st, a, ph);
        MJ_HIP(hipGetLastError());
    }
    // steps: every inner iteration takes two, every refinement three (RST, the stopping IB, TRUE), the last RST one
    const int64_t kmax = 2 * (int64_t)cfg->max_iters + 3 * (int64_t)mcfg->max_outer + 2;
    const int chunk = 2 * (cfg->poll_iters > 0 ? cfg->poll_iters : 64);
    int64_t k = 1;
    while (k <= kmax) {
        const int64_t kend = cfg->launch_mode == MJ_CG_LAUNCH_PHASES_FIXED ? kmax + 1 : std::min(kmax + 1, k + chunk);
        for (; k < kend; ++k) {
            hipLaunchKernelGGL(mx_step_kernel, grid, block, lds, st, a, (int)k);
            MJ_HIP(hipGetLastError());
        }
        if (k > kmax) break;
        int h = 0;
        MJ_HIP(hipMemcpyAsync(&h, a.nactive, sizeof(int), hipMemcpyDeviceToHost, st));
        MJ_HIP(hipStreamSynchronize(st));
        if (h == 0) break;
    }
    hipLaunchKernelGGL(mx_launches_kernel, dim3(1), dim3(256), 0, st, a.diag, a.B, (int)(k - 1));
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}
