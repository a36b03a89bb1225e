// Mixed-precision CG2D on gfx950: FP32 inner PCG with FP64 iterative refinement, batched over members.
//
// The system is the one cg2d_fortran solves (reference core/cg2d.py:205-269): the same normalisation (b * cg2dNorm,
// rhsNorm = 1/max|b|, normalised first guess with exchanged halos), the same 5-point operator and preconditioner,
// the same stopping rule err < tol^2 on the normalised residual. Per member:
//   r = bn - A x                                     (FP64: true residual, FP64 operator)
//   err = sum r^2                                    (FP64); stop when err < tol^2
//   sigma = sqrt(err); r32 = r / sigma               (inner right-hand side of unit norm)
//   PCG on A32 d = r32 from d = 0 in FP32 with FP32 operator/preconditioner copies; every dot product is formed
//   and summed in FP64 in a fixed order (below) and alpha/beta are FP64 ratios; the inner solve stops when its
//   recursive residual falls below inner_tolerance (or after inner_max_iters)
//   x = x + sigma * d                                (FP64 solution update)
// and repeat. max_iters caps the total inner iterations and max_outer the refinements; a member that reaches a
// cap finishes with converged = 0. Un-normalisation and the halos of x are as cg2d_fortran leaves them.
//
// Work decomposition as the FP64 solve: one block per (member, tile), rings computed from their interior sources,
// per-tile partials summed in tile order. The FP32 work vectors (r32, s, d, z, q) are the solver's own and live in a
// compact layout: per member and tile the interior only, row-major sNy x sNx, tiles 128-byte aligned. The pointwise
// updates stream them as flat 16-byte accesses; the stencil operand is staged in an LDS tile whose ring is computed
// from the sources' compact values; the stencils stream the FP32 coefficient rows down each wave's strip.
//
// Members move through their own state machine, one state per launch ("step"): every block of a member derives the
// member's action for the step from the member's record of the previous step and the partials, so all tiles of a
// member act alike; tile 0 writes the record of the step (records are double-buffered by step parity). States:
//   RST   err from the true-residual partials; finish, or restart the inner solve (r32, d = 0, eta(1))
//   IB    inner phase B(n): d += alpha_{n-1} s_{n-1} (the solution update of iteration n-1, moved here where
//         s_{n-1} is read anyway); inner stop -> x += sigma d (next: TRUE); else s = z + beta s, q = A32 s,
//         sigma_n = s.q
//   IC    inner phase C(n): r32 -= alpha q; z = M32 r32; err(n) = r32.r32, eta(n+1) = z.r32
// (z and q are kept as FP32 vectors, so each inner phase applies one stencil and reads one coefficient set.)
//   TRUE  x's ring from its sources; r = bn - A x; err partials
//
// Inner dot products (FP64): per tile, the rows in fixed groups of kRowGroup consecutive rows; each column pair is
// accumulated down the group's rows in row order, the group's column sums are reduced by the row tree of the
// "tile" order (pair_tree), and the group sums are added in group order. The order depends on the tile shape only,
// not on the launch geometry. The true-residual sums (FP64 refinement) use the "tile" order itself.
#include <algorithm>

#include "mj_cg_sweep.h"
#include "mj_common.h"
#include "mj_internal.h"

namespace mj {

struct MxRec {
    int mode, n, iters, outer;  // state, inner iteration, inner iterations done, refinements started
    double sigma;               // norm of the residual the current inner solve started from
};
enum { MX_RST = 0, MX_IB = 1, MX_IC = 2, MX_TRUE = 3 };
enum { AX_FINISH = 1, AX_RESTART, AX_UPD, AX_IB, AX_IC, AX_TRUE };

struct MxDev {
    int B, T, sNx, sNy, OL, nx, ny, plane;
    int64_t ms, oms, cms;  // member stride; FP64 operator slot stride; FP32 coefficient slot stride
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
    const int32_t *cring_lds, *cring_src;  // ring entries in compact terms (LDS offset, compact source or -1)
    double* r64;
    int TS;      // compact tile stride (elements)
    int64_t cs;  // compact member stride = T * TS
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
#ifndef MJ_MX_DEFAULT_THREADS
#define MJ_MX_DEFAULT_THREADS 256
#endif  // block when the config leaves it open

__device__ __forceinline__ int slot_of(const MxDev& a, int m) { return a.op_index ? a.op_index[m] : m; }
__device__ __forceinline__ float* R32(const MxDev& a, int n) { return n & 1 ? a.r1 : a.r0; }
__device__ __forceinline__ float* S32(const MxDev& a, int n) { return n & 1 ? a.s1 : a.s0; }
// compact offset of member m, tile t
__device__ __forceinline__ int64_t cbase(const MxDev& a, int m, int t) { return m * a.cs + (int64_t)t * a.TS; }

__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ void st4(float* p, float4 v) { *reinterpret_cast<float4*>(p) = v; }
// Row and column of compact element e of a tile (exact: e < 2^14, the fraction of (e + 0.5) / sNx stays >= 1/256
// away from an integer).
__device__ __forceinline__ int crow(int e, float inv_sNx) { return (int)((float(e) + 0.5f) * inv_sNx); }
// The four compact elements e .. e+3 (e a multiple of 4, sNx even) into the LDS tile: two column pairs, each within
// one row.
__device__ __forceinline__ void lds_put4(float* tile, const LdsGeo& G, int e, float4 v, int sNx, float inv_sNx) {
    const int j0 = crow(e, inv_sNx), j1 = crow(e + 2, inv_sNx);
    st2(tile + G.at(j0, e - j0 * sNx), make_float2(v.x, v.y));
    st2(tile + G.at(j1, e + 2 - j1 * sNx), make_float2(v.z, v.w));
}

// The west/east column copies of a tile's compact vector (base: the tile's compact start) for the four elements
// e .. e+3 (two column pairs, each within one row; column 0 is the first of a pair, column sNx-1 the second).
__device__ __forceinline__ void edge_put4(float* base, int e, float4 v, int sNx, int sNy, float inv_sNx) {
    if (!MJ_MX_EDGES) return;
    const int j0 = crow(e, inv_sNx), j1 = crow(e + 2, inv_sNx);
    const int c0 = e - j0 * sNx, c1 = e + 2 - j1 * sNx;
    float* ed = base + sNx * sNy;
    if (c0 == 0) ed[j0] = v.x;
    if (c0 + 1 == sNx - 1) ed[sNy + j0] = v.y;
    if (c1 == 0) ed[j1] = v.z;
    if (c1 + 1 == sNx - 1) ed[sNy + j1] = v.w;
}
__device__ __forceinline__ void edge_put2(float* base, int j, int c, float2 v, int sNx, int sNy) {
    if (!MJ_MX_EDGES) return;
    float* ed = base + sNx * sNy;
    if (c == 0) ed[j] = v.x;
    if (c + 1 == sNx - 1) ed[sNy + j] = v.y;
}

// Stencil over the LDS tile: IS_A: A32 = (W, S, C) = (aW, aS, aC); else M32 = (pW, pS, pC). Rows go in groups of
// kRowGroup, group g to wave g % nwaves; the result of each row goes to out (compact, row j at j * sNx) and
// prods(Xc, v, pv) gives the row's NR per-point products, reduced in the inner dot-product order; tile sum r of
// the block is written to *dst[r].
#ifndef MJ_ROW_GROUP
#define MJ_ROW_GROUP 6
#endif
constexpr int kRowGroup = MJ_ROW_GROUP;
template <bool IS_A, int NR, class F>
__device__ __forceinline__ void mx_stencil(const MxDev& a, int t, const float* __restrict__ cW,
                                           const float* __restrict__ cS, const float* __restrict__ cC,
                                           const float* tile, float* __restrict__ out, F&& prods,
                                           double* const (&dst)[NR]) {
    MJ_UNIT_IDS
    const Lanes L = lanes_of(lane, sNx);
    const LdsGeo G{sNx + 4};
    const int ce = east_col(L, sNx);
    const int ng = (sNy + kRowGroup - 1) / kRowGroup;
    for (int g = wave; g < ng; g += nwaves) {
        const int j0 = g * kRowGroup, j1 = min(j0 + kRowGroup, sNy);
        double2 acc[NR];
        int pr = MJ_ROW(j0);
        // the coefficient rows of row j + 1 are loaded while row j is computed
        float2 S0 = ld2(cS + pr + L.c), S1n = ld2(cS + pr + nx + L.c), Cn = ld2(cC + pr + L.c), Wn = ld2(cW + pr + ce);
        for (int j = j0; j < j1; ++j, pr += nx) {
            const int p = pr + L.c;
            const float2 S1 = S1n, C = Cn, Wc = Wn;
            if (j + 1 < j1) {
                S1n = ld2(cS + p + 2 * nx);
                Cn = ld2(cC + p + nx);
                Wn = ld2(cW + pr + nx + ce);
            }
            float wr = __shfl_down(Wc.x, 1, 64);
            if (L.last == 63 && lane == 63) wr = cW[pr + sNx];
            const int q0 = G.at(j, L.c);
            const float2 Xc = ld2(tile + q0), Xm = ld2(tile + q0 - G.W), Xp = ld2(tile + q0 + G.W);
            const float xl = tile[q0 - 1], xr = tile[q0 + 2];
            const float2 v = IS_A ? pair_A(Wc, wr, S0, S1, C, Xm, Xc, Xp, xl, xr)
                                  : pair_M(C, Wc, wr, S0, S1, Xm, Xc, Xp, xl, xr);
            if (L.act) {
                st2(out + j * sNx + L.c, v);
                edge_put2(out, j, L.c, v, sNx, sNy);
            }
            double2 pv[NR];
            prods(Xc, v, pv);
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                if (!L.act) pv[k] = make_double2(0.0, 0.0);
                if (j == j0) {
                    acc[k] = pv[k];
                } else {
                    acc[k].x = acc[k].x + pv[k].x;
                    acc[k].y = acc[k].y + pv[k].y;
                }
            }
            S0 = S1;
        }
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const double w = pair_tree(acc[k], sNx, lane);
            if (lane == 0) g_rows[k * 128 + g] = w;
        }
    }
    __syncthreads();
    if (tid < NR) {
        double sum = g_rows[tid * 128];
        for (int g = 1; g < ng; ++g) sum = sum + g_rows[tid * 128 + g];
        *dst[tid] = sum;
    }
    __syncthreads();
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
        a.rec[m] = MxRec{MX_RST, 0, 0, 0, 0.0};
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

// Restart: r32 = r / sigma on the interior (compact) and the ring (LDS, from the sources' r), d = 0, z = M32 r32,
// partial of eta(1) = z.r32.
__device__ __forceinline__ void mx_restart(const MxDev& a, int m, int t, double sigma) {
    MJ_UNIT_IDS
    MJ_STRIP
    float* tile = reinterpret_cast<float*>(mj_lds);
    const LdsGeo G{sNx + 4};
    const double* r = a.r64 + m * a.ms;
    const int64_t cb = cbase(a, m, t);
    float* r32 = R32(a, 0) + cb;
    float* d = a.d + cb;
    for (int j = j0; j < j1; ++j) {
        const double2 R = ld2(r + MJ_ROW(j) + L.c);
        const float2 v = make_float2((float)(R.x / sigma), (float)(R.y / sigma));
        if (L.act) {
            st2(r32 + j * sNx + L.c, v);
            edge_put2(r32, j, L.c, v, sNx, sNy);
            st2(tile + G.at(j, L.c), v);
        }
    }
    const int nf4 = sNx * sNy / 4;
    for (int f = tid; f < nf4; f += blockDim.x) st4(d + 4 * f, make_float4(0.f, 0.f, 0.f, 0.f));
    for (int k = a.ring_off[t] + tid; k < a.ring_off[t + 1]; k += blockDim.x) {
        const int s = a.ring_src[k];
        const float v = s < 0 ? 0.0f : (float)(r[s] / sigma);
        tile[a.cring_lds[k]] = a.ring_sign[k] < 0 ? -v : v;
    }
    __syncthreads();
    const int64_t oc = (int64_t)slot_of(a, m) * a.cms;
    double* const dst[1] = {a.peta + a.B * a.T + m * a.T + t};  // eta(1)
    mx_stencil<false, 1>(
        a, t, a.pW32 + oc, a.pS32 + oc, a.pC32 + oc, tile, a.z + cb,
        [](float2 Rc, float2 q, double2(&pv)[1]) {
            pv[0] = make_double2((double)q.x * (double)Rc.x, (double)q.y * (double)Rc.y);
        },
        dst);
}

// Solution update of a finished inner solve at B(n): the pending d += alpha_{n-1} s_{n-1} (n > 1), then
// x = x + sigma * d on the interior (FP64).
__device__ __forceinline__ void mx_update(const MxDev& a, int m, int t, int n, double sigma, double alpha64) {
    MJ_UNIT_IDS
    MJ_STRIP
    double* x = a.x + m * a.ms;
    const int64_t cb = cbase(a, m, t);
    const float* d = a.d + cb;
    const float* so = S32(a, n - 1) + cb;
    const float alpha = (float)alpha64;
    for (int j = j0; j < j1; ++j) {
        const int p = MJ_ROW(j) + L.c;
        const double2 X = ld2(x + p);
        float2 D = ld2(d + j * sNx + L.c);
        if (n > 1) {
            const float2 S = ld2(so + j * sNx + L.c);
            D.x = D.x + alpha * S.x;
            D.y = D.y + alpha * S.y;
        }
        if (L.act) st2(x + p, make_double2(X.x + sigma * (double)D.x, X.y + sigma * (double)D.y));
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

// Inner phase B(n): d += alpha_{n-1} s_{n-1} and s_n = z + beta s_{n-1} (s_1 = z) on the interior (compact, flat),
// s_n on the ring (LDS), q = A32 s_n, sigma_n partial (s_n, q).
__device__ __forceinline__ void mx_IB(const MxDev& a, int m, int t, int n, double beta64, double alpha64) {
    MJ_UNIT_IDS
    float* tile = reinterpret_cast<float*>(mj_lds);
    const LdsGeo G{sNx + 4};
    const float inv = 1.0f / (float)sNx;
    const int64_t cm = m * a.cs, cb = cbase(a, m, t), ot = (int64_t)t * a.TS;
    const float *z = a.z + cm, *so = S32(a, n - 1) + cm;
    float *sn = S32(a, n) + cb, *d = a.d + cb;
    const float beta = (float)beta64, alpha = (float)alpha64;
    const bool first = n == 1;
    const int nf4 = sNx * sNy / 4;
    for (int f = tid; f < nf4; f += blockDim.x) {
        const int e = 4 * f;
        float4 v = ld4(z + ot + e);
        if (!first) {
            const float4 O = ld4(so + ot + e), D = ld4(d + e);
            float4 dn;
            dn.x = D.x + alpha * O.x;
            dn.y = D.y + alpha * O.y;
            dn.z = D.z + alpha * O.z;
            dn.w = D.w + alpha * O.w;
            st4(d + e, dn);
            v.x = v.x + beta * O.x;
            v.y = v.y + beta * O.y;
            v.z = v.z + beta * O.z;
            v.w = v.w + beta * O.w;
        }
        st4(sn + e, v);
        edge_put4(sn, e, v, sNx, sNy, inv);
        lds_put4(tile, G, e, v, sNx, inv);
    }
    for (int k = a.ring_off[t] + tid; k < a.ring_off[t + 1]; k += blockDim.x) {
        const int s = a.cring_src[k];
        float v = 0.0f;
        if (s >= 0) v = first ? z[s] : z[s] + beta * so[s];
        tile[a.cring_lds[k]] = a.ring_sign[k] < 0 ? -v : v;
    }
    __syncthreads();
    const int64_t oc = (int64_t)slot_of(a, m) * a.cms;
    double* const dst[1] = {a.psig + (n & 1) * a.B * a.T + m * a.T + t};
    mx_stencil<true, 1>(
        a, t, a.aW32 + oc, a.aS32 + oc, a.aC32 + oc, tile, a.q + cb,
        [](float2 Xc, float2 q, double2(&pv)[1]) {
            pv[0] = make_double2((double)Xc.x * (double)q.x, (double)Xc.y * (double)q.y);
        },
        dst);
}

// Inner phase C(n): r32_n = r32_{n-1} - alpha q on the interior (compact, flat) and the ring (LDS); z = M32 r32_n;
// partials of err(n) = (r32_n, r32_n) and eta(n+1) = (z, r32_n).
__device__ __forceinline__ void mx_IC(const MxDev& a, int m, int t, int n, double alpha64) {
    MJ_UNIT_IDS
    float* tile = reinterpret_cast<float*>(mj_lds);
    const LdsGeo G{sNx + 4};
    const float inv = 1.0f / (float)sNx;
    const int64_t cm = m * a.cs, cb = cbase(a, m, t), ot = (int64_t)t * a.TS;
    const float *qv = a.q + cm, *rin = R32(a, n - 1) + cm;
    float* rout = R32(a, n) + cb;
    const float alpha = (float)alpha64;
    const int nf4 = sNx * sNy / 4;
    for (int f = tid; f < nf4; f += blockDim.x) {
        const int e = 4 * f;
        const float4 Q = ld4(qv + ot + e), R = ld4(rin + ot + e);
        float4 rn;
        rn.x = R.x - alpha * Q.x;
        rn.y = R.y - alpha * Q.y;
        rn.z = R.z - alpha * Q.z;
        rn.w = R.w - alpha * Q.w;
        st4(rout + e, rn);
        edge_put4(rout, e, rn, sNx, sNy, inv);
        lds_put4(tile, G, e, rn, sNx, inv);
    }
    for (int k = a.ring_off[t] + tid; k < a.ring_off[t + 1]; k += blockDim.x) {
        const int s = a.cring_src[k];
        const float v = s < 0 ? 0.0f : rin[s] - alpha * qv[s];
        tile[a.cring_lds[k]] = a.ring_sign[k] < 0 ? -v : v;
    }
    __syncthreads();
    const int64_t oc = (int64_t)slot_of(a, m) * a.cms;
    const int BT = a.B * a.T;
    double* const dst[2] = {a.perr + (n & 1) * BT + m * a.T + t, a.peta + ((n + 1) & 1) * BT + m * a.T + t};
    mx_stencil<false, 2>(
        a, t, a.pW32 + oc, a.pS32 + oc, a.pC32 + oc, tile, a.z + cb,
        [](float2 Rc, float2 zz, double2(&pv)[2]) {
            pv[0] = make_double2((double)Rc.x * (double)Rc.x, (double)Rc.y * (double)Rc.y);
            pv[1] = make_double2((double)zz.x * (double)Rc.x, (double)zz.y * (double)Rc.y);
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
#define MJ_PH_WAVES_PER_EU 8
#endif
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
            if (t == 0 && threadIdx.x == 0) a.rec[(k & 1) * a.B + m] = MxRec{MX_IB, n + 1, r0.iters + 1, r0.outer, r0.sigma};
            mx_IC(a, m, t, n, alpha);
            return;
        }
        bool stop = r0.iters >= a.max_iters || n - 1 >= a.inner_max;
        double beta = 0.0, alpha_prev = 0.0;
        if (n > 1) {
            // err(n-1) (inner stopping rule), eta(n), eta(n-1) (beta), sigma(n-1) (alpha_{n-1} of the pending
            // solution update: the value C(n-1) used)
            const double ER = tl ? a.perr[pp * BT + o + lane] : 0.0, EN = tl ? a.peta[cp * BT + o + lane] : 0.0;
            const double EP = tl ? a.peta[pp * BT + o + lane] : 0.0, SP = tl ? a.psig[pp * BT + o + lane] : 0.0;
            const double errp = lane_tile_sum(ER, a.T), etan = lane_tile_sum(EN, a.T);
            const double etap = lane_tile_sum(EP, a.T), sigp = lane_tile_sum(SP, a.T);
            if (!stop) stop = errp < a.itol_sq;
            alpha_prev = etap / sigp;
            beta = etan / etap;
        }
        if (t == 0 && threadIdx.x == 0) {
            MxRec nr = r0;
            nr.mode = stop ? MX_TRUE : MX_IC;
            a.rec[(k & 1) * a.B + m] = nr;
        }
        if (stop)
            mx_update(a, m, t, n, r0.sigma, alpha_prev);
        else
            mx_IB(a, m, t, n, beta, alpha_prev);
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
                    act = AX_RESTART;
                    nr = MxRec{MX_IB, 1, r.iters, r.outer + 1, sqrt(e)};
                }
                break;
            }
            case MX_IB: {
                bool stop = r.iters >= a.max_iters || r.n - 1 >= a.inner_max;
                if (r.n > 1) {
                    // err(n-1) (inner stopping rule), eta(n), eta(n-1) (beta), sigma(n-1) (alpha_{n-1} of the
                    // pending solution update: the value C(n-1) used)
                    const int cp = r.n & 1, pp = (r.n - 1) & 1;
                    const double errp = tsum(pp ? P[3] : P[2], a.perr + pp * BT + o);
                    const double etan = tsum(cp ? P[5] : P[4], a.peta + cp * BT + o);
                    const double etap = tsum(pp ? P[5] : P[4], a.peta + pp * BT + o);
                    const double sigp = tsum(pp ? P[7] : P[6], a.psig + pp * BT + o);
                    if (!stop) stop = errp < a.itol_sq;
                    alpha_prev = etap / sigp;
                    c = etan / etap;
                }
                if (stop) {
                    act = AX_UPD;
                    nr.mode = MX_TRUE;
                } else {
                    act = AX_IB;
                    nr.mode = MX_IC;
                }
                if (r.n == 1) c = 0.0;
                break;
            }
            case MX_IC: {
                const int cp = r.n & 1;
                const double etan = tsum(cp ? P[5] : P[4], a.peta + cp * BT + o);
                const double sign = tsum(cp ? P[7] : P[6], a.psig + cp * BT + o);
                act = AX_IC;
                c = etan / sign;
                nr = MxRec{MX_IB, r.n + 1, r.iters + 1, r.outer, r.sigma};
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
            g_bcast[1] = act == AX_RESTART ? nr.sigma : r.sigma;
            g_bcast[2] = (double)r.n;
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
    const int act = g_ibcast, n = (int)g_bcast[2];
    const double c = g_bcast[0], sigma = g_bcast[1], alpha_prev = g_bcast[3];
    switch (act) {
        case AX_FINISH: mx_finish(a, m, t); break;
        case AX_RESTART: mx_restart(a, m, t, sigma); break;
        case AX_UPD: mx_update(a, m, t, n, sigma, alpha_prev); break;
        case AX_IB: mx_IB(a, m, t, n, c, alpha_prev); break;
        case AX_IC: mx_IC(a, m, t, n, c); break;
        default: mx_true(a, m, t); break;
    }
}

// FP32 copies of the six coefficient arrays of every slot: dst[k][slot][q].
__global__ void mx_convert_kernel(const double* aW, const double* aS, const double* aC, const double* pW,
                                  const double* pS, const double* pC, int64_t oms, int nslots, int64_t tp,
                                  float* dst) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (q >= tp) return;
    const double* src[6] = {aW, aS, aC, pW, pS, pC};
#pragma unroll
    for (int k = 0; k < 6; ++k) dst[((int64_t)k * nslots + s) * tp + q] = (float)src[k][s * oms + q];
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
    const size_t nc = (size_t)c->batch * c->ntiles * compact_tile_stride(c->sNx, c->sNy);  // compact vectors
    const size_t tp = (size_t)c->ntiles * (c->sNx + 2 * c->OL) * (c->sNy + 2 * c->OL);
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
    const int64_t tp = (int64_t)a.T * a.plane;
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
    a.cring_lds = P->cg_cring_lds;
    a.cring_src = P->cg_cring_src;
    a.TS = compact_tile_stride(a.sNx, a.sNy);
    a.cs = (int64_t)a.T * a.TS;
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
    const int threads = cfg->block_threads ? cfg->block_threads : MJ_MX_DEFAULT_THREADS;
    const int nunits = a.B * a.T;
    const dim3 grid((unsigned)((nunits + kXcds - 1) / kXcds * kXcds)), block((unsigned)threads);
    // the step kernel's LDS tile of the stencil operand (FP32)
    const size_t lds = lds_tile_bytes(cfg->sNx, cfg->sNy, (int)sizeof(float));
    if (lds > 65536)
        MJ_HIP(hipFuncSetAttribute((const void*)mx_step_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(mx_reset_kernel, dim3(1), dim3(64), 0, st, a.nactive, a.B);
    MJ_HIP(hipGetLastError());
    hipLaunchKernelGGL(mx_convert_kernel, dim3((unsigned)((tp + 255) / 256), (unsigned)nslots), dim3(256), 0, st,
                       ops->aW2d, ops->aS2d, ops->aC2d, ops->pW, ops->pS, ops->pC, ops->member_stride, nslots, tp,
                       coef);
    MJ_HIP(hipGetLastError());
    for (int ph = 0; ph < 2; ++ph) {
        hipLaunchKernelGGL(mx_init_kernel, grid, block, 0, st, a, ph);
        MJ_HIP(hipGetLastError());
    }
    // steps: every inner iteration takes two, every refinement three (RST, the stopping IB, TRUE), the last RST one
    const int64_t kmax = 2 * (int64_t)cfg->max_iters + 3 * (int64_t)mcfg->max_outer + 2;
    const int chunk = cfg->poll_iters > 0 ? 2 * cfg->poll_iters : 64;
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
