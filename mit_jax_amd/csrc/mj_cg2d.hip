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
    double* prod;  // MJ_SUM_FORTRAN: per block [2][sNy*sNx] staging of the per-point values
    int32_t* niter;
    int* nactive;
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

// ------------------------------------------------------------------------------------------------ row reductions
// rowbuf[k][128] in LDS (tree order row sums); prod[k][sNy*sNx] in the block's global staging area (fortran
// order per-point values, summed sequentially by one thread).
template <int MODE>
struct Red {
    double* rowbuf;  // [2][128]
    double* prod;    // [2][sNy*sNx] (MODE == MJ_SUM_FORTRAN only), global memory
    int sNx, sNy;

    __device__ __forceinline__ void put(int k, int jl, int lane, double v0, double v1, bool has1) {
        if (MODE == MJ_SUM_TREE) {
            const double w = wave_tree(v0, v1, has1);
            if (lane == 0) rowbuf[k * 128 + jl] = w;
        } else {
            double* pr = prod + (size_t)k * sNy * sNx + jl * sNx;
            pr[lane] = v0;
            if (has1) pr[lane + 64] = v1;
        }
    }
    // After __syncthreads: the tile partial of slot k, valid in thread 0 (called by wave 0 only).
    __device__ __forceinline__ double tile(int k, int lane) {
        if (MODE == MJ_SUM_TREE) {
            const double* rb = rowbuf + k * 128;
            const bool has1 = lane + 64 < sNy;
            return wave_tree(rb[lane], has1 ? rb[lane + 64] : 0.0, has1);
        } else {
            double s = 0.0;
            if (lane == 0) {
                const double* pr = prod + (size_t)k * sNy * sNx;
                const int n = sNy * sNx;
                for (int q = 0; q < n; ++q) s = s + pr[q];
            }
            return s;
        }
    }
};

// Per row helper: wave w of the block processes rows jl = w, w + nwaves, ...; lane l holds i = l, l + 64.
#define MJ_ROWS(...)                                                                                      \
    for (int jl = wave; jl < a.sNy; jl += nwaves) {                                                       \
        const int prow = tbase + (a.OL + jl) * a.nx + a.OL;                                               \
        const int p0 = prow + lane;                                                                       \
        const bool has1 = lane + 64 < a.sNx;                                                              \
        const int p1 = prow + lane + 64;                                                                  \
        __VA_ARGS__                                                                                       \
    }

__device__ __forceinline__ double member_rhs_norm(const CgDev& a, int m) {
    double mx = 0.0;
    for (int t = 0; t < a.T; ++t) mx = fmax(mx, a.pmax[m * a.T + t]);
    return mx != 0.0 ? 1.0 / mx : 1.0;
}
__device__ __forceinline__ double member_rhs_max(const CgDev& a, int m) {
    double mx = 0.0;
    for (int t = 0; t < a.T; ++t) mx = fmax(mx, a.pmax[m * a.T + t]);
    return mx;
}

template <int MODE>
__global__ void __launch_bounds__(1024) cg2d_fused_kernel(CgDev a, GridSync gs) {
    __shared__ double rowbuf[2 * 128];
    __shared__ double redbuf[16];
    __shared__ int s_flag;
    Red<MODE> red{rowbuf, MODE == MJ_SUM_FORTRAN ? a.prod + (size_t)blockIdx.x * 2 * a.sNx * a.sNy : nullptr, a.sNx,
                  a.sNy};
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const int nunits = a.B * a.T;
    const int plane = a.plane;

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
        // halo points: the first guess itself; the work vectors: 0 (s0 entirely)
        for (int q = tid; q < plane; q += blockDim.x) {
            const int j = q / a.nx, i = q - j * a.nx, p = tbase + q;
            const bool inner = j >= a.OL && j < a.OL + a.sNy && i >= a.OL && i < a.OL + a.sNx;
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
            const bool inner = sj >= a.OL && sj < a.OL + a.sNy && si >= a.OL && si < a.OL + a.sNx;
            const double v = inner ? x0[s] * rhsNorm : x0[s];
            x[a.full_dst[k]] = a.full_sign[k] < 0 ? -v : v;
        }
        __syncthreads();
        MJ_ROWS({
            const double bn0 = b[p0] * nm * rhsNorm;
            const double r_0 = bn0 - op_A(a.aW + m * a.oms, a.aS + m * a.oms, a.aC + m * a.oms, x, p0, a.nx);
            r0[p0] = r_0;
            double bn1 = 0.0, r_1 = 0.0;
            if (has1) {
                bn1 = b[p1] * nm * rhsNorm;
                r_1 = bn1 - op_A(a.aW + m * a.oms, a.aS + m * a.oms, a.aC + m * a.oms, x, p1, a.nx);
                r0[p1] = r_1;
            }
            red.put(0, jl, lane, r_0 * r_0, r_1 * r_1, has1);
            red.put(1, jl, lane, bn0, bn1, has1);
        })
        __syncthreads();
        if (wave == 0) {
            const double e = red.tile(0, lane);
            const double sr = red.tile(1, lane);
            if (lane == 0) {
                a.perr[u] = e;  // slot 0: err(0)
                a.psum[u] = sr;
            }
        }
        __syncthreads();
    }
    if (!grid_sync(gs)) goto done;

    // ---------------------------------------------------------------- I2: exchange r, eta(1) = sum (M r) r
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int m = u / a.T, t = u - m * a.T, tbase = t * plane;
        double* r0 = a.r0 + m * a.ms;
        const double *pW = a.pW + m * a.oms, *pS = a.pS + m * a.oms, *pC = a.pC + m * a.oms;
        for (int k = a.ring_off[t] + tid; k < a.ring_off[t + 1]; k += blockDim.x) {
            const double v = r0[a.ring_src[k]];
            r0[a.ring_dst[k]] = a.ring_sign[k] < 0 ? -v : v;
        }
        __syncthreads();
        MJ_ROWS({
            const double q0 = op_M(pW, pS, pC, r0, p0, a.nx);
            const double q1 = has1 ? op_M(pW, pS, pC, r0, p1, a.nx) : 0.0;
            red.put(0, jl, lane, q0 * r0[p0], has1 ? q1 * r0[p1] : 0.0, has1);
        })
        __syncthreads();
        if (wave == 0) {
            const double e = red.tile(0, lane);
            if (lane == 0) a.peta[a.B * a.T + u] = e;  // slot 1: eta(1)
        }
        if (t == 0 && tid == 0) {
            const double err0 = tile_order_sum(a.perr + m * a.T, a.T);
            a.diag[m].firstResidual = sqrt(err0);
            a.diag[m].sumRHS = tile_order_sum(a.psum + m * a.T, a.T);
            a.diag[m].rhsMax = member_rhs_max(a, m);
            a.diag[m].outerIters = 0;
            a.diag[m].pad_ = 0;
            a.niter[m] = -1;
        }
        __syncthreads();
    }
    if (!grid_sync(gs)) goto done;

    // ---------------------------------------------------------------- iterations
    for (int n = 1;; ++n) {
        const int BT = a.B * a.T;
        const double* peta_n = a.peta + (n & 1) * BT;
        const double* peta_m = a.peta + ((n - 1) & 1) * BT;
        const double* perr_m = a.perr + ((n - 1) & 1) * BT;
        double* psig_n = a.psig + (n & 1) * BT;
        // ------------------------------------------------------------ phase B(n)
        for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
            const int m = u / a.T, t = u - m * a.T, tbase = t * plane;
            const int ni = a.niter[m];
            if (ni >= 0 && ni < n - 1) continue;  // finished in an earlier iteration
            const double errp = tile_order_sum(perr_m + m * a.T, a.T);
            if (!(n - 1 < a.max_iters) || errp < a.tol_sq) {
                // member done after n-1 iterations (cg2d.F:213/:346 stopping rule)
                const double rhsNorm = member_rhs_norm(a, m);
                double* x = a.x + m * a.ms;
                MJ_ROWS({
                    x[p0] = x[p0] / rhsNorm;
                    if (has1) x[p1] = x[p1] / rhsNorm;
                })
                if (t == 0 && tid == 0) {
                    a.niter[m] = n - 1;
                    a.diag[m].lastResidual = sqrt(errp);
                    a.diag[m].numIters = n - 1;
                    a.diag[m].converged = errp < a.tol_sq ? 1 : 0;
                    atomicSub(a.nactive, 1);
                }
                __syncthreads();
                continue;
            }
            const double eta = tile_order_sum(peta_n + m * a.T, a.T);
            const double etaNM1 = n == 1 ? 1.0 : tile_order_sum(peta_m + m * a.T, a.T);
            const double beta = eta / etaNM1;
            const double* rin = ((n - 1) & 1 ? a.r1 : a.r0) + m * a.ms;
            const double* sold = ((n - 1) & 1 ? a.s1 : a.s0) + m * a.ms;
            double* snew = (n & 1 ? a.s1 : a.s0) + m * a.ms;
            const double *pW = a.pW + m * a.oms, *pS = a.pS + m * a.oms, *pC = a.pC + m * a.oms;
            const double *aW = a.aW + m * a.oms, *aS = a.aS + m * a.oms, *aC = a.aC + m * a.oms;
            MJ_ROWS({
                snew[p0] = op_M(pW, pS, pC, rin, p0, a.nx) + beta * sold[p0];
                if (has1) snew[p1] = op_M(pW, pS, pC, rin, p1, a.nx) + beta * sold[p1];
            })
            for (int k = a.ring_off[t] + tid; k < a.ring_off[t + 1]; k += blockDim.x) {
                const int s = a.ring_src[k];
                const double v = op_M(pW, pS, pC, rin, s, a.nx) + beta * sold[s];
                snew[a.ring_dst[k]] = a.ring_sign[k] < 0 ? -v : v;
            }
            __syncthreads();
            MJ_ROWS({
                const double v0 = snew[p0] * op_A(aW, aS, aC, snew, p0, a.nx);
                const double v1 = has1 ? snew[p1] * op_A(aW, aS, aC, snew, p1, a.nx) : 0.0;
                red.put(0, jl, lane, v0, v1, has1);
            })
            __syncthreads();
            if (wave == 0) {
                const double sg = red.tile(0, lane);
                if (lane == 0) psig_n[u] = sg;
            }
            __syncthreads();
        }
        if (!grid_sync(gs)) goto done;
        if (tid == 0) s_flag = __hip_atomic_load(a.nactive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        if (s_flag == 0) break;
        // ------------------------------------------------------------ phase C(n)
        double* perr_n = a.perr + (n & 1) * BT;
        double* peta_p = a.peta + ((n + 1) & 1) * BT;
        for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
            const int m = u / a.T, t = u - m * a.T, tbase = t * plane;
            if (a.niter[m] >= 0) continue;
            const double eta = tile_order_sum(peta_n + m * a.T, a.T);
            const double alpha = eta / tile_order_sum(psig_n + m * a.T, a.T);
            const double* snew = (n & 1 ? a.s1 : a.s0) + m * a.ms;
            const double* rin = ((n - 1) & 1 ? a.r1 : a.r0) + m * a.ms;
            double* rout = (n & 1 ? a.r1 : a.r0) + m * a.ms;
            double* x = a.x + m * a.ms;
            const double *pW = a.pW + m * a.oms, *pS = a.pS + m * a.oms, *pC = a.pC + m * a.oms;
            const double *aW = a.aW + m * a.oms, *aS = a.aS + m * a.oms, *aC = a.aC + m * a.oms;
            MJ_ROWS({
                x[p0] = x[p0] + alpha * snew[p0];
                rout[p0] = rin[p0] - alpha * op_A(aW, aS, aC, snew, p0, a.nx);
                if (has1) {
                    x[p1] = x[p1] + alpha * snew[p1];
                    rout[p1] = rin[p1] - alpha * op_A(aW, aS, aC, snew, p1, a.nx);
                }
            })
            for (int k = a.ring_off[t] + tid; k < a.ring_off[t + 1]; k += blockDim.x) {
                const int s = a.ring_src[k];
                const double v = rin[s] - alpha * op_A(aW, aS, aC, snew, s, a.nx);
                rout[a.ring_dst[k]] = a.ring_sign[k] < 0 ? -v : v;
            }
            __syncthreads();
            MJ_ROWS({
                const double e0 = rout[p0] * rout[p0];
                const double e1 = has1 ? rout[p1] * rout[p1] : 0.0;
                red.put(0, jl, lane, e0, e1, has1);
                const double q0 = op_M(pW, pS, pC, rout, p0, a.nx) * rout[p0];
                const double q1 = has1 ? op_M(pW, pS, pC, rout, p1, a.nx) * rout[p1] : 0.0;
                red.put(1, jl, lane, q0, q1, has1);
            })
            __syncthreads();
            if (wave == 0) {
                const double e = red.tile(0, lane);
                const double q = red.tile(1, lane);
                if (lane == 0) {
                    perr_n[u] = e;
                    peta_p[u] = q;
                }
            }
            __syncthreads();
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
                                   const double* v, double* y, const double* norm, int64_t ms, int T, int ny, int nx,
                                   int OL, int sNx, int sNy) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    const int plane = ny * nx;
    if (q >= (int64_t)T * plane) return;
    const int r = (int)(q % plane), j = r / nx, i = r - j * nx;
    const bool inner = j >= OL && j < OL + sNy && i >= OL && i < OL + sNx;
    const int p = (int)q;
    y[m * ms + p] = inner ? op_A(aW + m * oms, aS + m * oms, aC + m * oms, v + m * ms, p, nx) / norm[m] : 0.0;
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

constexpr int kFortranMaxBlocks = 1024;  // grid cap of the fortran sum order (bounds its staging area)

struct WsLayout {
    size_t r0, r1, s0, s1, pmax, psum, perr, peta, psig, niter, misc, prod, total;
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
    L.niter = o; o += al((size_t)c->batch * 4);
    L.misc = o; o += 256;  // barrier counter, error flag, nactive
    L.prod = o;
    if (c->sum_order == MJ_SUM_FORTRAN) o += al((size_t)kFortranMaxBlocks * 2 * c->sNx * c->sNy * 8);
    L.total = o;
    return L;
}

int check_cfg(const mj_cg2d_config* c) {
    MJ_REQUIRE(c, "cg2d: null config");
    MJ_REQUIRE(c->batch >= 1 && c->batch <= 65535, "cg2d: batch %d out of range", c->batch);
    MJ_REQUIRE(c->ntiles >= 1 && c->OL >= 1, "cg2d: bad layout");
    MJ_REQUIRE(c->sNx >= 64 && c->sNx <= 128 && c->sNy >= 64 && c->sNy <= 128,
               "cg2d: sNx, sNy must be in 64..128 (got %d, %d)", c->sNx, c->sNy);
    const int64_t plane = (int64_t)(c->sNx + 2 * c->OL) * (c->sNy + 2 * c->OL);
    MJ_REQUIRE(c->member_stride >= c->ntiles * plane, "cg2d: member_stride too small");
    MJ_REQUIRE(c->sum_order == MJ_SUM_FORTRAN || c->sum_order == MJ_SUM_TREE, "cg2d: bad sum_order %d", c->sum_order);
    MJ_REQUIRE(c->max_iters >= 0, "cg2d: max_iters < 0");
    MJ_REQUIRE(c->block_threads == 0 || (c->block_threads % 64 == 0 && c->block_threads >= 64 &&
                                         c->block_threads <= 1024),
               "cg2d: block_threads must be a multiple of 64 in 64..1024");
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
    a.niter = (int32_t*)(w + L.niter);
    a.prod = (double*)(w + L.prod);
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
    if (threads == 0) threads = nunits >= 2 * ncu ? 256 : 1024;
    const size_t shmem = 0;
    const void* kfn = cfg->sum_order == MJ_SUM_FORTRAN ? (const void*)cg2d_fused_kernel<MJ_SUM_FORTRAN>
                                                       : (const void*)cg2d_fused_kernel<MJ_SUM_TREE>;
    int per_cu = 0;
    MJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, threads, shmem));
    MJ_REQUIRE(per_cu >= 1, "mj_cg2d_solve: kernel cannot be resident with %d threads", threads);
    int grid = std::min(nunits, per_cu * ncu);
    if (cfg->grid_blocks > 0) grid = std::min(grid, cfg->grid_blocks);
    if (cfg->sum_order == MJ_SUM_FORTRAN) grid = std::min(grid, kFortranMaxBlocks);

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
                       ops->member_stride, (const double*)work, y, cg2dNorm, cfg->member_stride, cfg->ntiles, ny, nx,
                       cfg->OL, cfg->sNx, cfg->sNy);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}
