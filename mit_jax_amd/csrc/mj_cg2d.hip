// CG2D on gfx950: one persistent cooperative launch per (batched) solve, two grid barriers per CG iteration.
//
// Algorithm: the literal cg2d_fortran of the reference (core/cg2d.py:205-269, cg2d.F:110-395). Per member:
//   b <- cg2d_b * cg2dNorm; rhsMax = max|b|; rhsNorm = 1/rhsMax (1 if 0); b <- b * rhsNorm; x <- x0 * rhsNorm on
//   the interior, then EXCH_XY; r = b - A x; err = sum r^2; r exchanged; loop while it < max_iters and
//   !(err < tol^2): q = M r; eta = sum q r; beta = eta/eta_prev; s = q + beta s; exch s; q = A s;
//   alpha = eta / sum s q; x += alpha s; r -= alpha q; err = sum r^2; exch r. x /= rhsNorm on the interior.
//
// Work decomposition: a unit is one (member, tile); a block owns whole units, its waves own rows (one 90-point
// row per wave: lane l holds i = l and i = l + 64), so the fixed row/tile reduction trees of MJ_SUM_TREE (or the
// sequential MJ_SUM_FORTRAN sum, staged per block) are computed inside one block. Global sums are the fixed tile-order
// sum of the per-tile partials, recomputed identically by every block that needs them: no floating-point atomics,
// results independent of the launch geometry.
//
// Exchanges: the stencils read only the width-1 ring (i = 0, sNx+1 for j = 1..sNy; j = 0, sNy+1 for i = 1..sNx),
// whose written points are all sourced from interior points of other tiles. Instead of exchanging after each
// update (a third barrier), a block computes the ring values of its own tile directly from the sources with the
// same arithmetic the source's owner uses, which gives the same bits as the exchanged copy. s and r are ping-pong
// buffers so a source is never overwritten while another block reads it. Per iteration:
//   phase B(n): beta from eta(n), eta(n-1); s_n = M r_{n-1} + beta s_{n-1} on interior + ring; sigma = sum s_n A s_n
//   phase C(n): alpha = eta(n)/sigma; x += alpha s_n; r_n = r_{n-1} - alpha A s_n on interior + ring;
//               err(n) = sum r_n^2; eta(n+1) = sum (M r_n) r_n
// which is the reference recurrence with eta(n+1) computed one phase early (the value is the same).
#include <algorithm>

#include "mj_common.h"
#include "mj_internal.h"

namespace mj {

struct CgDev {
    int B, T, sNx, sNy, OL, nx, ny, plane;
    int64_t ms, oms;
    const double *aW, *aS, *aC, *pW, *pS, *pC;
    const double *b, *x0;
    double* x;
    const double* norm;
    double tol_sq;
    int max_iters;
    const int32_t *full_off, *full_dst, *full_src;
    const int8_t* full_sign;
    const int32_t *ring_off, *ring_dst, *ring_src;
    const int8_t* ring_sign;
    double *r0, *r1, *s0, *s1;
    double *pmax, *psum, *perr, *peta, *psig;  // partials: [B*T], [B*T], [2][B*T] x3
    const int32_t* op_index;                   // operator slot per member (NULL: slot = member)
    int* nactive;
    int32_t* mdone;  // [B] iterations at which a member finished, -1 while active
    mj_cg2d_diag* diag;
};

// ------------------------------------------------------------------------------------------------ stencils
// apply_operator (core/cg2d.py:181-188): aW(i)x(i-1) + aW(i+1)x(i+1) + aS(j)x(j-1) + aS(j+1)x(j+1) + aC x
__device__ __forceinline__ double op_A(const double* aW, const double* aS, const double* aC, const double* x, int p,
                                       int nx) {
    return aW[p] * x[p - 1] + aW[p + 1] * x[p + 1] + aS[p] * x[p - nx] + aS[p + nx] * x[p + nx] + aC[p] * x[p];
}
// apply_preconditioner (core/cg2d.py:191-198): pC r + pW(i)r(i-1) + pW(i+1)r(i+1) + pS(j)r(j-1) + pS(j+1)r(j+1)
__device__ __forceinline__ double op_M(const double* pW, const double* pS, const double* pC, const double* r, int p,
                                       int nx) {
    return pC[p] * r[p] + pW[p] * r[p - 1] + pW[p + 1] * r[p + 1] + pS[p] * r[p - nx] + pS[p + nx] * r[p + nx];
}

// ------------------------------------------------------------------------------------------------ LDS tile staging
// A unit's (member, tile) interior plus its width-1 ring, (sNy+2) x (sNx+2) points, is staged in LDS for the
// stencils. LDS index of padded (j, i): (j - OL + 1) * RW + (i - OL + 1), RW = sNx + 2.
constexpr int kRingMax = 92;                    // sNx, sNy <= 90 (LLC90)
constexpr int kTileLds = kRingMax * kRingMax;   // 8464 doubles = 67.7 KB per staged tile

// Reduction of a per-point array over the tile interior: P(jl, il) read through `get`, the tree or sequential
// order of MODE. Entry: P complete (after a __syncthreads). Returns the tile partial in thread 0. Exit with a
// __syncthreads (P and rowbuf reusable).
template <int MODE, class Get>
__device__ __forceinline__ double tile_reduce(Get get, double* rowbuf, int sNx, int sNy, int tid) {
    const int lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    double res = 0.0;
    if (MODE == MJ_SUM_TREE) {
        const bool hx = lane + 64 < sNx;
        for (int jl = wave; jl < sNy; jl += nwaves) {
            const double w = wave_tree(get(jl, lane), hx ? get(jl, lane + 64) : 0.0, hx);
            if (lane == 0) rowbuf[jl] = w;
        }
        __syncthreads();
        if (wave == 0) {
            const bool hy = lane + 64 < sNy;
            res = wave_tree(rowbuf[lane], hy ? rowbuf[lane + 64] : 0.0, hy);
        }
    } else {
        if (tid == 0)
            for (int jl = 0; jl < sNy; ++jl)
                for (int il = 0; il < sNx; ++il) res = res + get(jl, il);
    }
    __syncthreads();
    return res;
}

// Stage the ring region of a padded member array into LDS, 8 loads per thread in flight.
__device__ __forceinline__ void stage_tile(double* L, const double* __restrict__ g, int tbase, int nx, int OL, int RW,
                                           int RH, int tid) {
    const int n = RW * RH, nt = blockDim.x;
    for (int e0 = 0; e0 < n; e0 += 8 * nt) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int e = e0 + k * nt + tid;
            if (e < n) {
                const int jr = e / RW, ir = e - jr * RW;
                v[k] = g[tbase + (jr + OL - 1) * nx + (ir + OL - 1)];
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int e = e0 + k * nt + tid;
            if (e < n) L[e] = v[k];
        }
    }
}

// The thread index through an empty asm the compiler cannot see through: values derived from it inside a
// phase are computed there, not hoisted out of the persistent iteration loop (where they would all stay live
// at once and spill).
__device__ __forceinline__ int opaque_tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// Stencils on an LDS ring tile at LDS index e (row stride RW), coefficients at padded index p.
__device__ __forceinline__ double lds_A(const double* __restrict__ aW, const double* __restrict__ aS,
                                        const double* __restrict__ aC, const double* L, int p, int nx, int e, int RW) {
    return aW[p] * L[e - 1] + aW[p + 1] * L[e + 1] + aS[p] * L[e - RW] + aS[p + nx] * L[e + RW] + aC[p] * L[e];
}
__device__ __forceinline__ double lds_M(const double* __restrict__ pW, const double* __restrict__ pS,
                                        const double* __restrict__ pC, const double* L, int p, int nx, int e, int RW) {
    return pC[p] * L[e] + pW[p] * L[e - 1] + pW[p + 1] * L[e + 1] + pS[p] * L[e - RW] + pS[p + nx] * L[e + RW];
}

// Interior point loop: e over sNy*sNx; (jl, il), padded index p, LDS ring-tile index q.
#define MJ_INTERIOR(...)                                                                                  \
    for (int e_ = tid; e_ < a.sNy * a.sNx; e_ += blockDim.x) {                                            \
        const int jl = e_ / a.sNx, il = e_ - jl * a.sNx;                                                  \
        const int p = tbase + (a.OL + jl) * a.nx + a.OL + il;                                             \
        const int q = (jl + 1) * RW + il + 1;                                                             \
        __VA_ARGS__                                                                                       \
    }

// Global-memory row loop of the once-per-solve phases: wave w takes rows jl = w, w + nwaves, ...; lane l holds
// i = l and i = l + 64.
#define MJ_ROWS(...)                                                                                      \
    for (int jl = wave; jl < a.sNy; jl += nwaves) {                                                       \
        const int prow = tbase + (a.OL + jl) * a.nx + a.OL;                                               \
        const int p0 = prow + lane;                                                                       \
        const bool has1 = lane + 64 < a.sNx;                                                              \
        const int p1 = prow + lane + 64;                                                                  \
        __VA_ARGS__                                                                                       \
    }

__device__ __forceinline__ double member_rhs_max(const CgDev& a, int m) {
    double mx = 0.0;
    for (int t = 0; t < a.T; ++t) mx = fmax(mx, a.pmax[m * a.T + t]);
    return mx;
}
__device__ __forceinline__ double member_rhs_norm(const CgDev& a, int m) {
    const double mx = member_rhs_max(a, m);
    return mx != 0.0 ? 1.0 / mx : 1.0;
}

#ifndef MJ_CG_THREADS
#define MJ_CG_THREADS 1024
#endif  // block size of the fused solve (one block per CU: its LDS tiles take 136 KB)
#ifndef MJ_BATCH
#define MJ_BATCH 2
#endif  // interior points per thread whose loads are issued together
constexpr int kMaxMembers = 2048;  // members of one batched solve (LDS member lists)
constexpr int kXcds = 8;           // MI355X: 8 XCDs, workgroups dispatched round-robin over them

// The LDS tiles of the fused kernel, at file scope so the per-unit phase functions below address them directly.
__shared__ double g_R[kTileLds];
__shared__ double g_S[kTileLds];
__shared__ double g_rowbuf[128];

// Interior loop of the phase functions: MJ_BATCH points per thread, all loads first (registers c0..c6[k]), then
// the arithmetic.
#define MJ_UNIT_INTERIOR(LOADS, COMPUTE)                                                                  \
    for (int base_ = 0; base_ < sNy * sNx; base_ += MJ_BATCH * nt) {                                      \
        double c0[MJ_BATCH], c1[MJ_BATCH], c2[MJ_BATCH], c3[MJ_BATCH], c4[MJ_BATCH], c5[MJ_BATCH],         \
            c6[MJ_BATCH];                                                                                 \
        int pk[MJ_BATCH], qk[MJ_BATCH], ek[MJ_BATCH];                                                     \
        _Pragma("unroll") for (int k = 0; k < MJ_BATCH; ++k) {                                            \
            const int e_ = base_ + k * nt + tid;                                                          \
            ek[k] = e_;                                                                                   \
            if (e_ < sNy * sNx) {                                                                         \
                const int jl = e_ / sNx, il = e_ - jl * sNx;                                              \
                const int p = tbase + (OL + jl) * nx + OL + il;                                           \
                pk[k] = p;                                                                                \
                qk[k] = (jl + 1) * RW + il + 1;                                                           \
                LOADS                                                                                     \
            }                                                                                             \
        }                                                                                                 \
        _Pragma("unroll") for (int k = 0; k < MJ_BATCH; ++k) {                                            \
            const int e_ = ek[k];                                                                         \
            if (e_ < sNy * sNx) {                                                                         \
                const int p = pk[k], q = qk[k];                                                           \
                (void)p;                                                                                  \
                COMPUTE                                                                                   \
            }                                                                                             \
        }                                                                                                 \
    }

// Phase B(n) of one unit: s_new = M r + beta s_old on interior and ring (ring from its interior sources), then
// sigma = sum s_new (A s_new). Returns the tile partial of sigma (thread 0).
template <int MODE>
__device__ __noinline__ double unit_phase_B(const double* __restrict__ rin, const double* __restrict__ sold,
                                            double* __restrict__ snew, const double* __restrict__ pW,
                                            const double* __restrict__ pS, const double* __restrict__ pC,
                                            const double* __restrict__ aW, const double* __restrict__ aS,
                                            const double* __restrict__ aC, const int32_t* __restrict__ ring_dst,
                                            const int32_t* __restrict__ ring_src, const int8_t* __restrict__ ring_sign,
                                            int rk0, int rk1, double beta, int tbase, int nx, int OL, int sNx,
                                            int sNy) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int RW = sNx + 2, RH = sNy + 2;
    double* R = g_R;
    double* S = g_S;
    stage_tile(R, rin, tbase, nx, OL, RW, RH, tid);
    for (int e = tid; e < RW * RH; e += nt) S[e] = 0.0;
    __syncthreads();
    MJ_UNIT_INTERIOR(
        {
            c0[k] = pC[p];
            c1[k] = pW[p];
            c2[k] = pW[p + 1];
            c3[k] = pS[p];
            c4[k] = pS[p + nx];
            c5[k] = sold[p];
        },
        {
            const double v = c0[k] * R[q] + c1[k] * R[q - 1] + c2[k] * R[q + 1] + c3[k] * R[q - RW] +
                             c4[k] * R[q + RW] + beta * c5[k];
            snew[p] = v;
            S[q] = v;
        })
    for (int k2 = rk0 + tid; k2 < rk1; k2 += nt) {
        const int s = ring_src[k2];
        double v = op_M(pW, pS, pC, rin, s, nx) + beta * sold[s];
        v = ring_sign[k2] < 0 ? -v : v;
        const int d = ring_dst[k2];
        snew[d] = v;
        const int dq = d - tbase, dj = dq / nx, di = dq - dj * nx;
        S[(dj - OL + 1) * RW + (di - OL + 1)] = v;
    }
    __syncthreads();
    MJ_UNIT_INTERIOR(
        {
            c0[k] = aW[p];
            c1[k] = aW[p + 1];
            c2[k] = aS[p];
            c3[k] = aS[p + nx];
            c4[k] = aC[p];
        },
        {
            R[e_] = S[q] * (c0[k] * S[q - 1] + c1[k] * S[q + 1] + c2[k] * S[q - RW] + c3[k] * S[q + RW] +
                            c4[k] * S[q]);
        })
    __syncthreads();
    return tile_reduce<MODE>([&](int jl, int il) { return R[jl * sNx + il]; }, g_rowbuf, sNx, sNy, tid);
}

// Phase C(n) of one unit: x += alpha s_new; r_new = r - alpha A s_new on interior and ring; then the tile
// partials of err = sum r_new^2 (returned in .x) and eta(n+1) = sum (M r_new) r_new (in .y), thread 0.
template <int MODE>
__device__ __noinline__ double2 unit_phase_C(const double* __restrict__ snew, const double* __restrict__ rin,
                                             double* __restrict__ rout, double* __restrict__ x,
                                             const double* __restrict__ pW, const double* __restrict__ pS,
                                             const double* __restrict__ pC, const double* __restrict__ aW,
                                             const double* __restrict__ aS, const double* __restrict__ aC,
                                             const int32_t* __restrict__ ring_dst,
                                             const int32_t* __restrict__ ring_src,
                                             const int8_t* __restrict__ ring_sign, int rk0, int rk1, double alpha,
                                             int tbase, int nx, int OL, int sNx, int sNy) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int RW = sNx + 2, RH = sNy + 2;
    double* R = g_R;
    double* S = g_S;
    stage_tile(S, snew, tbase, nx, OL, RW, RH, tid);
    for (int e = tid; e < RW * RH; e += nt) R[e] = 0.0;
    __syncthreads();
    MJ_UNIT_INTERIOR(
        {
            c0[k] = aW[p];
            c1[k] = aW[p + 1];
            c2[k] = aS[p];
            c3[k] = aS[p + nx];
            c4[k] = aC[p];
            c5[k] = x[p];
            c6[k] = rin[p];
        },
        {
            x[p] = c5[k] + alpha * S[q];
            const double v = c6[k] - alpha * (c0[k] * S[q - 1] + c1[k] * S[q + 1] + c2[k] * S[q - RW] +
                                              c3[k] * S[q + RW] + c4[k] * S[q]);
            rout[p] = v;
            R[q] = v;
        })
    for (int k2 = rk0 + tid; k2 < rk1; k2 += nt) {
        const int s = ring_src[k2];
        double v = rin[s] - alpha * op_A(aW, aS, aC, snew, s, nx);
        v = ring_sign[k2] < 0 ? -v : v;
        const int d = ring_dst[k2];
        rout[d] = v;
        const int dq = d - tbase, dj = dq / nx, di = dq - dj * nx;
        R[(dj - OL + 1) * RW + (di - OL + 1)] = v;
    }
    __syncthreads();
    MJ_UNIT_INTERIOR(
        {
            c0[k] = pC[p];
            c1[k] = pW[p];
            c2[k] = pW[p + 1];
            c3[k] = pS[p];
            c4[k] = pS[p + nx];
        },
        {
            S[e_] = (c0[k] * R[q] + c1[k] * R[q - 1] + c2[k] * R[q + 1] + c3[k] * R[q - RW] + c4[k] * R[q + RW]) *
                    R[q];
        })
    __syncthreads();
    double2 out;
    out.x = tile_reduce<MODE>(
        [&](int jl, int il) {
            const double v = R[(jl + 1) * RW + il + 1];
            return v * v;
        },
        g_rowbuf, sNx, sNy, tid);
    out.y = tile_reduce<MODE>([&](int jl, int il) { return S[jl * sNx + il]; }, g_rowbuf, sNx, sNy, tid);
    return out;
}

template <int MODE>
__global__ void __launch_bounds__(MJ_CG_THREADS) cg2d_fused_kernel(CgDev a, GridSync gs) {
    double* R = g_R;
    double* S = g_S;
    double* rowbuf = g_rowbuf;
    __shared__ double redbuf[16];
    __shared__ int wcnt[32];
    __shared__ int s_nact, s_nfin;
    __shared__ short alist[kMaxMembers], flist[kMaxMembers];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const int nunits = a.B * a.T;
    const int plane = a.plane;
    const int RW = a.sNx + 2, RH = a.sNy + 2;
    const int sNx = a.sNx, sNy = a.sNy;
    auto opbase = [&](int m) { return (int64_t)(a.op_index ? a.op_index[m] : m) * a.oms; };
    for (int m = blockIdx.x * blockDim.x + tid; m < a.B; m += gridDim.x * blockDim.x) a.mdone[m] = -1;

    // ---------------------------------------------------------------- I0: per-tile max |b * cg2dNorm|
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int m = u / a.T, t = u - m * a.T, tbase = t * plane;
        const double* b = a.b + m * a.ms;
        const double nm = a.norm[m];
        double mx = 0.0;
        MJ_ROWS({
            mx = fmax(mx, fabs(b[p0] * nm));
            if (has1) mx = fmax(mx, fabs(b[p1] * nm));
        })
        for (int o = 32; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
        if (lane == 0) redbuf[wave] = mx;
        __syncthreads();
        if (tid == 0) {
            double v = 0.0;
            for (int w = 0; w < nwaves; ++w) v = fmax(v, redbuf[w]);
            a.pmax[u] = v;
        }
        __syncthreads();
    }
    if (!grid_sync(gs)) goto done;

    // ---------------------------------------------------------------- I1: x (normalised, exchanged), r = b - A x
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int m = u / a.T, t = u - m * a.T, tbase = t * plane;
        const double rhsNorm = member_rhs_norm(a, m);
        const double nm = a.norm[m];
        const double* x0 = a.x0 + m * a.ms;
        const double* b = a.b + m * a.ms;
        double* x = a.x + m * a.ms;
        double *r0 = a.r0 + m * a.ms, *r1 = a.r1 + m * a.ms, *s0 = a.s0 + m * a.ms, *s1 = a.s1 + m * a.ms;
        const int64_t ob = opbase(m);
        // halo points: the first guess itself; the work vectors: 0 (s0 entirely)
        for (int qq = tid; qq < plane; qq += blockDim.x) {
            const int j = qq / a.nx, i = qq - j * a.nx, p = tbase + qq;
            const bool inner = j >= a.OL && j < a.OL + sNy && i >= a.OL && i < a.OL + sNx;
            s0[p] = 0.0;
            if (!inner) {
                x[p] = x0[p];
                r0[p] = 0.0;
                r1[p] = 0.0;
                s1[p] = 0.0;
            } else {
                x[p] = x0[p] * rhsNorm;
            }
        }
        __syncthreads();
        // exchanged points: the pre-exchange value at the source (normalised where the source is interior)
        for (int k = a.full_off[t] + tid; k < a.full_off[t + 1]; k += blockDim.x) {
            const int s = a.full_src[k];
            const int sq = s % plane, sj = sq / a.nx, si = sq - sj * a.nx;
            const bool inner = sj >= a.OL && sj < a.OL + sNy && si >= a.OL && si < a.OL + sNx;
            const double v = inner ? x0[s] * rhsNorm : x0[s];
            x[a.full_dst[k]] = a.full_sign[k] < 0 ? -v : v;
        }
        __syncthreads();
        // r = b - A x on the interior; per-point r^2 into R, b into S (as sNy x sNx arrays)
        MJ_INTERIOR({
            const double bn = b[p] * nm * rhsNorm;
            const double rr = bn - op_A(a.aW + ob, a.aS + ob, a.aC + ob, x, p, a.nx);
            r0[p] = rr;
            R[e_] = rr * rr;
            S[e_] = bn;
        })
        __syncthreads();
        const double e0 = tile_reduce<MODE>([&](int jl, int il) { return R[jl * sNx + il]; }, rowbuf, sNx, sNy, tid);
        const double sr = tile_reduce<MODE>([&](int jl, int il) { return S[jl * sNx + il]; }, rowbuf, sNx, sNy, tid);
        if (tid == 0) {
            a.perr[u] = e0;  // slot 0: err(0)
            a.psum[u] = sr;
        }
    }
    if (!grid_sync(gs)) goto done;

    // ---------------------------------------------------------------- I2: exchange r, eta(1) = sum (M r) r
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int m = u / a.T, t = u - m * a.T, tbase = t * plane;
        double* r0 = a.r0 + m * a.ms;
        const int64_t ob = opbase(m);
        for (int k = a.ring_off[t] + tid; k < a.ring_off[t + 1]; k += blockDim.x) {
            const double v = r0[a.ring_src[k]];
            r0[a.ring_dst[k]] = a.ring_sign[k] < 0 ? -v : v;
        }
        __syncthreads();
        stage_tile(R, r0, tbase, a.nx, a.OL, RW, RH, tid);
        __syncthreads();
        MJ_INTERIOR({ S[e_] = lds_M(a.pW + ob, a.pS + ob, a.pC + ob, R, p, a.nx, q, RW) * R[q]; })
        __syncthreads();
        const double e1 = tile_reduce<MODE>([&](int jl, int il) { return S[jl * sNx + il]; }, rowbuf, sNx, sNy, tid);
        if (tid == 0) a.peta[a.B * a.T + u] = e1;  // slot 1: eta(1)
        if (t == 0 && tid == 0) {
            const double err0 = tile_order_sum(a.perr + m * a.T, a.T);
            a.diag[m].firstResidual = sqrt(err0);
            a.diag[m].sumRHS = tile_order_sum(a.psum + m * a.T, a.T);
            a.diag[m].rhsMax = member_rhs_max(a, m);
            a.diag[m].outerIters = 0;
            a.diag[m].pad_ = 0;
        }
    }
    if (!grid_sync(gs)) goto done;

    // ---------------------------------------------------------------- iterations
    // Every block derives the same member status from the same partials at the start of B(n): finished earlier
    // (mdone[m] < n-1), finishing now (stopping rule on err(n-1)), or active. Active members are compacted into
    // alist and their units are spread evenly over the grid, tile-major and in contiguous chunks per XCD (block b
    // runs on XCD b % kXcds), so the members of one tile share its coefficients in that XCD's L2.
    for (int n = 1;; ++n) {
        const int BT = a.B * a.T;
        const double* peta_n = a.peta + (n & 1) * BT;
        const double* peta_m = a.peta + ((n - 1) & 1) * BT;
        const double* perr_m = a.perr + ((n - 1) & 1) * BT;
        double* psig_n = a.psig + (n & 1) * BT;
        // ------------------------------------------------------------ member status and compaction
        for (int c0 = 0; c0 < a.B; c0 += blockDim.x) {
            const int m = c0 + tid;
            int code = 0;  // 0 finished earlier, 1 finishing now, 2 active
            if (m < a.B) {
                const int st = a.mdone[m];
                if (!(st >= 0 && st < n - 1)) {
                    const double errp = tile_order_sum(perr_m + m * a.T, a.T);
                    code = (!(n - 1 < a.max_iters) || errp < a.tol_sq) ? 1 : 2;
                }
            }
            const unsigned long long ma = __ballot(code == 2), mf = __ballot(code == 1);
            const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
            if (lane == 0) {
                wcnt[wave] = __popcll(ma);
                wcnt[16 + wave] = __popcll(mf);
            }
            __syncthreads();
            int oa = c0 == 0 ? 0 : s_nact, of = c0 == 0 ? 0 : s_nfin;
            for (int w = 0; w < wave; ++w) {
                oa += wcnt[w];
                of += wcnt[16 + w];
            }
            if (code == 2) alist[oa + __popcll(ma & below)] = (short)m;
            if (code == 1) flist[of + __popcll(mf & below)] = (short)m;
            __syncthreads();
            if (tid == 0) {
                int ta = c0 == 0 ? 0 : s_nact, tf = c0 == 0 ? 0 : s_nfin;
                for (int w = 0; w < nwaves; ++w) {
                    ta += wcnt[w];
                    tf += wcnt[16 + w];
                }
                s_nact = ta;
                s_nfin = tf;
            }
            __syncthreads();
        }
        const int nact = s_nact, nfin = s_nfin;
        // ------------------------------------------------------------ finishing members (cg2d.F:380-395)
        for (int it = blockIdx.x; it < nfin * a.T; it += gridDim.x) {
            const int tid = opaque_tid();
            const int m = flist[it / a.T], t = it % a.T, tbase = t * plane;
            const double rhsNorm = member_rhs_norm(a, m);
            double* x = a.x + m * a.ms;
            MJ_INTERIOR({ x[p] = x[p] / rhsNorm; })
            if (t == 0 && tid == 0) {
                const double errp = tile_order_sum(perr_m + m * a.T, a.T);
                a.diag[m].lastResidual = sqrt(errp);
                a.diag[m].numIters = n - 1;
                a.diag[m].converged = errp < a.tol_sq ? 1 : 0;
            }
        }
        if (blockIdx.x == 0)
            for (int k = tid; k < nfin; k += blockDim.x) a.mdone[flist[k]] = n - 1;
        if (nact == 0) break;
        // XCD-contiguous chunk of the nact*T active units this block works on
        const int nitems = nact * a.T;
        const int xcd = blockIdx.x % kXcds, lb = blockIdx.x / kXcds;
        const int gx = (gridDim.x - xcd + kXcds - 1) / kXcds;
        const int i0 = (int)((int64_t)nitems * xcd / kXcds), i1 = (int)((int64_t)nitems * (xcd + 1) / kXcds);
        // ------------------------------------------------------------ phase B(n)
        for (int it = i0 + lb; it < i1; it += gx) {
            const int t = it / nact, m = alist[it - t * nact], u = m * a.T + t, tbase = t * plane;
            const double eta = tile_order_sum(peta_n + m * a.T, a.T);
            const double etaNM1 = n == 1 ? 1.0 : tile_order_sum(peta_m + m * a.T, a.T);
            const double beta = eta / etaNM1;
            const double* rin = ((n - 1) & 1 ? a.r1 : a.r0) + m * a.ms;
            const double* sold = ((n - 1) & 1 ? a.s1 : a.s0) + m * a.ms;
            double* snew = (n & 1 ? a.s1 : a.s0) + m * a.ms;
            const int64_t ob = opbase(m);
            const double sg = unit_phase_B<MODE>(rin, sold, snew, a.pW + ob, a.pS + ob, a.pC + ob, a.aW + ob,
                                                 a.aS + ob, a.aC + ob, a.ring_dst, a.ring_src, a.ring_sign,
                                                 a.ring_off[t], a.ring_off[t + 1], beta, tbase, a.nx, a.OL, sNx, sNy);
            if (tid == 0) psig_n[u] = sg;
        }
        if (!grid_sync(gs)) goto done;
        // ------------------------------------------------------------ phase C(n)
        double* perr_n = a.perr + (n & 1) * BT;
        double* peta_p = a.peta + ((n + 1) & 1) * BT;
        for (int it = i0 + lb; it < i1; it += gx) {
            const int t = it / nact, m = alist[it - t * nact], u = m * a.T + t, tbase = t * plane;
            const double eta = tile_order_sum(peta_n + m * a.T, a.T);
            const double alpha = eta / tile_order_sum(psig_n + m * a.T, a.T);
            const double* snew = (n & 1 ? a.s1 : a.s0) + m * a.ms;
            const double* rin = ((n - 1) & 1 ? a.r1 : a.r0) + m * a.ms;
            double* rout = (n & 1 ? a.r1 : a.r0) + m * a.ms;
            const int64_t ob = opbase(m);
            const double2 eq = unit_phase_C<MODE>(snew, rin, rout, a.x + m * a.ms, a.pW + ob, a.pS + ob, a.pC + ob,
                                                  a.aW + ob, a.aS + ob, a.aC + ob, a.ring_dst, a.ring_src,
                                                  a.ring_sign, a.ring_off[t], a.ring_off[t + 1], alpha, tbase, a.nx,
                                                  a.OL, sNx, sNy);
            const double e = eq.x, qv = eq.y;
            if (tid == 0) {
                perr_n[u] = e;
                peta_p[u] = qv;
            }
        }
        if (!grid_sync(gs)) goto done;
    }
done:
    // a timed-out barrier: mark every member's diag (converged = -2) so the host sees the failure
    if (__hip_atomic_load(gs.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 && tid == 0) {
        for (int u = blockIdx.x; u < nunits; u += gridDim.x)
            if (u % a.T == 0) a.diag[u / a.T].converged = -2;
    }
}

__global__ void cg2d_init_kernel(unsigned int* counter, int* error, int* nactive, int B) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *counter = 0u;
        *error = 0;
        *nactive = B;
    }
}

// (A / cg2dNorm) v on the interior, 0 on halos, v already exchanged (core/cg2d.py:292-298)
__global__ void cg2d_matvec_kernel(const double* aW, const double* aS, const double* aC, int64_t oms,
                                   const int32_t* op_index, const double* v, double* y, const double* norm, int64_t ms,
                                   int T, int ny, int nx, int OL, int sNx, int sNy) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    const int plane = ny * nx;
    if (q >= (int64_t)T * plane) return;
    const int r = (int)(q % plane), j = r / nx, i = r - j * nx;
    const bool inner = j >= OL && j < OL + sNy && i >= OL && i < OL + sNx;
    const int p = (int)q;
    const int64_t ob = (int64_t)(op_index ? op_index[m] : m) * oms;
    y[m * ms + p] = inner ? op_A(aW + ob, aS + ob, aC + ob, v + m * ms, p, nx) / norm[m] : 0.0;
}

__global__ void zero_halo_copy_kernel(const double* v, double* w, int64_t ms, int T, int ny, int nx, int OL, int sNx,
                                      int sNy) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    const int plane = ny * nx;
    if (q >= (int64_t)T * plane) return;
    const int r = (int)(q % plane), j = r / nx, i = r - j * nx;
    const bool inner = j >= OL && j < OL + sNy && i >= OL && i < OL + sNx;
    w[m * ms + q] = inner ? v[m * ms + q] : 0.0;
}

}  // namespace mj

using namespace mj;

// ------------------------------------------------------------------------------------------------ host side
namespace {

struct WsLayout {
    size_t r0, r1, s0, s1, pmax, psum, perr, peta, psig, mdone, misc, total;
};

WsLayout ws_layout(const mj_cg2d_config* c) {
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    WsLayout L{};
    const size_t vec = al((size_t)c->batch * c->member_stride * 8);
    const size_t bt = al((size_t)c->batch * c->ntiles * 8);
    size_t o = 0;
    L.r0 = o; o += vec;
    L.r1 = o; o += vec;
    L.s0 = o; o += vec;
    L.s1 = o; o += vec;
    L.pmax = o; o += bt;
    L.psum = o; o += bt;
    L.perr = o; o += 2 * bt;
    L.peta = o; o += 2 * bt;
    L.psig = o; o += 2 * bt;
    L.mdone = o; o += al((size_t)c->batch * 4);
    L.misc = o; o += 256;  // barrier counter, error flag, nactive
    L.total = o;
    return L;
}

int check_cfg(const mj_cg2d_config* c) {
    MJ_REQUIRE(c, "cg2d: null config");
    MJ_REQUIRE(c->batch >= 1 && c->batch <= kMaxMembers, "cg2d: batch %d out of range 1..%d", c->batch,
               kMaxMembers);
    MJ_REQUIRE(c->ntiles >= 1 && c->OL >= 1, "cg2d: bad layout");
    MJ_REQUIRE(c->sNx >= 64 && c->sNx <= kRingMax - 2 && c->sNy >= 64 && c->sNy <= kRingMax - 2,
               "cg2d: sNx, sNy must be in 64..%d (got %d, %d)", kRingMax - 2, c->sNx, c->sNy);
    const int64_t plane = (int64_t)(c->sNx + 2 * c->OL) * (c->sNy + 2 * c->OL);
    MJ_REQUIRE(c->member_stride >= c->ntiles * plane, "cg2d: member_stride too small");
    MJ_REQUIRE(c->sum_order == MJ_SUM_FORTRAN || c->sum_order == MJ_SUM_TREE, "cg2d: bad sum_order %d", c->sum_order);
    MJ_REQUIRE(c->max_iters >= 0, "cg2d: max_iters < 0");
    MJ_REQUIRE(c->block_threads == 0 || (c->block_threads % 64 == 0 && c->block_threads >= 64 &&
                                         c->block_threads <= MJ_CG_THREADS),
               "cg2d: block_threads must be a multiple of 64 in 64..%d", MJ_CG_THREADS);
    return MJ_OK;
}

// Per-tile tables of the T plan for the fused kernel (see mj_halo_plan::cg_*).
int ensure_cg_tables(mj_halo_plan* P, int sNx, int sNy, int OL) {
    std::lock_guard<std::mutex> lk(P->cg_mu);
    if (P->cg_full_off && P->cg_sNx == sNx && P->cg_sNy == sNy && P->cg_OL == OL) return MJ_OK;
    MJ_REQUIRE(!P->cg_full_off, "cg2d: the T plan was already used with another layout");
    MJ_REQUIRE(P->nmaps == 1, "cg2d: plan_T must be a scalar (T) plan");
    const int nx = P->nx, ny = P->ny, T = P->ntiles, plane = nx * ny;
    MJ_REQUIRE(nx == sNx + 2 * OL && ny == sNy + 2 * OL, "cg2d: plan layout %dx%d does not match sNx/OL", ny, nx);
    std::vector<int32_t> foff(T + 1, 0), roff(T + 1, 0), fdst, fsrc, rdst, rsrc;
    std::vector<int8_t> fsg, rsg;
    auto interior = [&](int q) {
        const int r = q % plane, j = r / nx, i = r % nx;
        return j >= OL && j < OL + sNy && i >= OL && i < OL + sNx;
    };
    for (int t = 0; t < T; ++t) {
        for (int k = 0; k < P->npts[0]; ++k) {
            const int d = P->h_dst[0][k];
            if (d / plane != t) continue;
            fdst.push_back(d);
            fsrc.push_back(P->h_src[0][k]);
            fsg.push_back(P->h_sign[0][k]);
            const int r = d % plane, j = r / nx, i = r % nx;
            const bool ring = ((j == OL - 1 || j == OL + sNy) && i >= OL && i < OL + sNx) ||
                              ((i == OL - 1 || i == OL + sNx) && j >= OL && j < OL + sNy);
            if (ring) {
                // the fused kernel computes ring values from their sources: they must be interior points
                MJ_REQUIRE(interior(P->h_src[0][k]), "cg2d: ring point %d has a non-interior source %d", d,
                           P->h_src[0][k]);
                rdst.push_back(d);
                rsrc.push_back(P->h_src[0][k]);
                rsg.push_back(P->h_sign[0][k]);
            }
        }
        foff[t + 1] = (int32_t)fdst.size();
        roff[t + 1] = (int32_t)rdst.size();
        P->cg_ring_max = std::max(P->cg_ring_max, roff[t + 1] - roff[t]);
    }
    auto up = [](const void* h, size_t n, void** d) -> int {
        MJ_HIP(hipMalloc(d, n ? n : 4));
        if (n) MJ_HIP(hipMemcpy(*d, h, n, hipMemcpyHostToDevice));
        return MJ_OK;
    };
    int rc;
    if ((rc = up(foff.data(), foff.size() * 4, (void**)&P->cg_full_off)) ||
        (rc = up(fdst.data(), fdst.size() * 4, (void**)&P->cg_full_dst)) ||
        (rc = up(fsrc.data(), fsrc.size() * 4, (void**)&P->cg_full_src)) ||
        (rc = up(fsg.data(), fsg.size(), (void**)&P->cg_full_sign)) ||
        (rc = up(roff.data(), roff.size() * 4, (void**)&P->cg_ring_off)) ||
        (rc = up(rdst.data(), rdst.size() * 4, (void**)&P->cg_ring_dst)) ||
        (rc = up(rsrc.data(), rsrc.size() * 4, (void**)&P->cg_ring_src)) ||
        (rc = up(rsg.data(), rsg.size(), (void**)&P->cg_ring_sign)))
        return rc;
    P->cg_sNx = sNx;
    P->cg_sNy = sNy;
    P->cg_OL = OL;
    return MJ_OK;
}

}  // namespace

extern "C" size_t mj_cg2d_workspace_bytes(const mj_cg2d_config* cfg) {
    if (check_cfg(cfg)) return 0;
    return ws_layout(cfg).total;
}

extern "C" int mj_cg2d_solve(const mj_halo_plan* plan_T, const mj_cg2d_ops* ops, const double* b, const double* x0,
                             double* x, const double* cg2dNorm, const mj_cg2d_config* cfg, void* workspace,
                             size_t workspace_bytes, mj_cg2d_diag* diag, void* stream) {
    int rc = check_cfg(cfg);
    if (rc) return rc;
    MJ_REQUIRE(plan_T && ops && b && x0 && x && cg2dNorm && workspace && diag, "mj_cg2d_solve: null argument");
    MJ_REQUIRE(ops->aW2d && ops->aS2d && ops->aC2d && ops->pW && ops->pS && ops->pC, "mj_cg2d_solve: null operator");
    MJ_REQUIRE(x != b && x != x0, "mj_cg2d_solve: x must not alias b or x0");
    MJ_REQUIRE(((uintptr_t)workspace & 255) == 0, "mj_cg2d_solve: workspace must be 256-byte aligned");
    const WsLayout L = ws_layout(cfg);
    MJ_REQUIRE(workspace_bytes >= L.total, "mj_cg2d_solve: workspace %zu < %zu bytes", workspace_bytes, L.total);
    mj_halo_plan* P = const_cast<mj_halo_plan*>(plan_T);
    MJ_REQUIRE(P->ntiles == cfg->ntiles, "mj_cg2d_solve: plan has %d tiles, config %d", P->ntiles, cfg->ntiles);
    if ((rc = ensure_cg_tables(P, cfg->sNx, cfg->sNy, cfg->OL))) return rc;

    char* w = (char*)workspace;
    CgDev a{};
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
    a.aW = ops->aW2d;
    a.aS = ops->aS2d;
    a.aC = ops->aC2d;
    a.pW = ops->pW;
    a.pS = ops->pS;
    a.pC = ops->pC;
    a.b = b;
    a.x0 = x0;
    a.x = x;
    a.norm = cg2dNorm;
    a.tol_sq = cfg->tolerance * cfg->tolerance;  // cg2d.F:111
    a.max_iters = cfg->max_iters;
    a.full_off = P->cg_full_off;
    a.full_dst = P->cg_full_dst;
    a.full_src = P->cg_full_src;
    a.full_sign = P->cg_full_sign;
    a.ring_off = P->cg_ring_off;
    a.ring_dst = P->cg_ring_dst;
    a.ring_src = P->cg_ring_src;
    a.ring_sign = P->cg_ring_sign;
    a.r0 = (double*)(w + L.r0);
    a.r1 = (double*)(w + L.r1);
    a.s0 = (double*)(w + L.s0);
    a.s1 = (double*)(w + L.s1);
    a.pmax = (double*)(w + L.pmax);
    a.psum = (double*)(w + L.psum);
    a.perr = (double*)(w + L.perr);
    a.peta = (double*)(w + L.peta);
    a.psig = (double*)(w + L.psig);
    a.op_index = ops->op_index;
    a.mdone = (int32_t*)(w + L.mdone);
    unsigned int* counter = (unsigned int*)(w + L.misc);
    int* error = (int*)(w + L.misc + 64);
    a.nactive = (int*)(w + L.misc + 128);
    a.diag = diag;

    int dev = 0;
    MJ_HIP(hipGetDevice(&dev));
    int ncu = 0, coop = 0, clk_khz = 0;
    MJ_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    MJ_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    MJ_HIP(hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeWallClockRate, dev));
    MJ_REQUIRE(coop, "mj_cg2d_solve: device does not support cooperative launches");
    const int nunits = a.B * a.T;
    int threads = cfg->block_threads;
    if (threads == 0) threads = MJ_CG_THREADS;
    const size_t shmem = 0;
    const void* kfn = cfg->sum_order == MJ_SUM_FORTRAN ? (const void*)cg2d_fused_kernel<MJ_SUM_FORTRAN>
                                                       : (const void*)cg2d_fused_kernel<MJ_SUM_TREE>;
    int per_cu = 0;
    MJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, threads, shmem));
    MJ_REQUIRE(per_cu >= 1, "mj_cg2d_solve: kernel cannot be resident with %d threads", threads);
    int grid = std::min(nunits, per_cu * ncu);
    if (cfg->grid_blocks > 0) grid = std::min(grid, cfg->grid_blocks);
    MJ_REQUIRE(grid >= kXcds || grid == nunits, "mj_cg2d_solve: grid of %d blocks is too small", grid);
    GridSync gs{counter, error, (unsigned)grid, 0u,
                (unsigned long long)(clk_khz > 0 ? clk_khz : 100000) * 1000ull * 4ull /* 4 s */};
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(cg2d_init_kernel, dim3(1), dim3(64), 0, st, counter, error, a.nactive, a.B);
    MJ_HIP(hipGetLastError());
    void* args[] = {(void*)&a, (void*)&gs};
    MJ_HIP(hipLaunchCooperativeKernel(kfn, dim3(grid), dim3(threads), args, (unsigned)shmem, st));
    return MJ_OK;
}

extern "C" int mj_cg2d_matvec(const mj_halo_plan* plan_T, const mj_cg2d_ops* ops, const double* v, double* y,
                              const double* cg2dNorm, const mj_cg2d_config* cfg, double* work, void* stream) {
    int rc = check_cfg(cfg);
    if (rc) return rc;
    MJ_REQUIRE(plan_T && ops && v && y && cg2dNorm && work, "mj_cg2d_matvec: null argument");
    MJ_REQUIRE(plan_T->nmaps == 1, "mj_cg2d_matvec: plan_T must be scalar");
    const int nx = cfg->sNx + 2 * cfg->OL, ny = cfg->sNy + 2 * cfg->OL;
    const int64_t n = (int64_t)cfg->ntiles * nx * ny;
    hipStream_t st = as_stream(stream);
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)cfg->batch);
    hipLaunchKernelGGL(zero_halo_copy_kernel, grid, dim3(256), 0, st, v, work, cfg->member_stride, cfg->ntiles, ny,
                       nx, cfg->OL, cfg->sNx, cfg->sNy);
    MJ_HIP(hipGetLastError());
    mj_halo_geom g{cfg->batch, cfg->member_stride, (int64_t)nx * ny, 1, 0};
    if ((rc = mj_halo_exchange(plan_T, work, work, &g, stream))) return rc;
    hipLaunchKernelGGL(cg2d_matvec_kernel, grid, dim3(256), 0, st, ops->aW2d, ops->aS2d, ops->aC2d,
                       ops->member_stride, ops->op_index, (const double*)work, y, cg2dNorm, cfg->member_stride, cfg->ntiles, ny, nx,
                       cfg->OL, cfg->sNx, cfg->sNy);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}
