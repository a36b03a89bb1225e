// CG2D on gfx950: the batched preconditioned conjugate-gradient solve of the free-surface equation.
//
// Algorithm: the literal cg2d_fortran of the reference (core/cg2d.py:205-269, cg2d.F:110-395). Per member:
//   b <- cg2d_b * cg2dNorm; rhsMax = max|b|; rhsNorm = 1/rhsMax (1 if 0); b <- b * rhsNorm; x <- x0 * rhsNorm on
//   the interior, then EXCH_XY; r = b - A x; err = sum r^2; r exchanged; loop while it < max_iters and
//   !(err < tol^2): q = M r; eta = sum q r; beta = eta/eta_prev; s = q + beta s; exch s; q = A s;
//   alpha = eta / sum s q; x += alpha s; r -= alpha q; err = sum r^2; exch r. x /= rhsNorm on the interior.
//
// Work decomposition: a unit is one (member, tile), processed by one block whose waves own rows (one 90-point row
// per wave: lane l holds i = l and i = l + 64), so the fixed row/tile reduction trees of MJ_SUM_TREE (or the
// sequential MJ_SUM_FORTRAN sum, staged per unit) are computed inside one block. Global sums are the fixed
// tile-order sum of the per-tile partials, recomputed identically by every block that needs them: no
// floating-point atomics, results independent of the launch geometry.
//
// Exchanges: the stencils read only the width-1 ring (i = 0, sNx+1 for j = 1..sNy; j = 0, sNy+1 for i = 1..sNx),
// whose written points are all sourced from interior points of other tiles. Instead of exchanging after each
// update, a block computes the ring values of its own tile directly from the sources with the same arithmetic the
// source's owner uses, which gives the same bits as the exchanged copy. s and r are ping-pong buffers so a source
// is never overwritten while another block reads it. Per iteration two phases:
//   phase B(n): stopping rule on err(n-1) (finished members un-normalise x and leave); beta = eta(n)/eta(n-1);
//               s_n = z + beta s_{n-1} on interior + ring; q = A s_n; sigma = sum s_n q
//   phase C(n): alpha = eta(n)/sigma; x += alpha s_n; r_n = r_{n-1} - alpha q on interior + ring;
//               z = M r_n; err(n) = sum r_n^2; eta(n+1) = sum z r_n
// which is the reference recurrence with z = M r_n and eta(n+1) computed one phase early (the values are the same).
// z and q are kept as vectors, so each phase applies one stencil and reads one coefficient set.
//
// Two launch shapes run the same unit functions, so their results are bitwise identical:
//   persistent: one cooperative launch, phases separated by grid barriers (small batches: no launch gaps, no host
//               round trips, fully asynchronous);
//   phases:     one ordinary launch per phase with one block per unit (large batches: high occupancy, the block
//               scheduler balances the units of active members; units of finished members exit at once). The host
//               enqueues iterations in chunks and reads the count of active members between chunks, or enqueues
//               exactly max_iters iterations (no host read, capturable in a graph).
#include <algorithm>

#include "mj_cg_sweep.h"
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
    double *z, *q;                             // M r and A s of the current iteration (interior points)
    double *pmax, *psum, *perr, *peta, *psig;  // partials [m * T + t]: pmax, psum [B*T]; perr, peta, psig [2][B*T]
    const int32_t* op_index;                   // operator slot per member (NULL: slot = member)
    double* prod;                              // MJ_SUM_FORTRAN staging: [B*T][2][sNy*sNx]
    int* nactive;                              // members not yet finished
    int32_t* mdone;                            // [B] iterations at which a member finished, -1 while active
    int32_t* grp;                              // [B][kGroup] member groups (build_groups_kernel), -1 = empty
    int* ngroups;
    mj_cg2d_diag* diag;
};

#ifndef MJ_CG_THREADS
#define MJ_CG_THREADS 1024
#endif  // largest block of the persistent launch
#ifndef MJ_PH_THREADS
#define MJ_PH_THREADS 1024
#endif  // largest block of the phase launches (the group phases: MJ_PH_THREADS * kGroup)
#ifndef MJ_PH_DEFAULT_THREADS
#define MJ_PH_DEFAULT_THREADS 1024
#endif  // block of the phase launches when the config leaves it open (the fastest of 128-1024 for 128 members)
#ifndef MJ_PH_WAVES_PER_EU
#define MJ_PH_WAVES_PER_EU 8
#endif  // occupancy target of the phase launches (waves per SIMD): bounds their register use (a group's rows)
__device__ __forceinline__ int64_t op_base(const CgDev& a, int m) {
    return (int64_t)(a.op_index ? a.op_index[m] : m) * a.oms;
}
__device__ __forceinline__ double* stage_of(const CgDev& a, int m, int t) {
    return a.prod + ((size_t)m * a.T + t) * 2 * a.sNx * a.sNy;
}
__device__ __forceinline__ double* R(const CgDev& a, int n) { return n & 1 ? a.r1 : a.r0; }
__device__ __forceinline__ double* S(const CgDev& a, int n) { return n & 1 ? a.s1 : a.s0; }

// ------------------------------------------------------------------------------------------------ unit functions
// I1: x = normalised first guess (halo points of the tile: the exchanged value), r = b - A x on the interior,
// tile partials of err(0) = sum r^2 and sumRHS = sum b. Work vectors start at 0 (s0 entirely).
template <int MODE>
__device__ __forceinline__ void unit_I1(const CgDev& a, int m, int t) {
    MJ_UNIT_IDS
    const int plane = a.plane;
    const double rhsNorm = member_rhs_norm(a, m);
    const double nm = a.norm[m];
    const double* b = a.b + m * a.ms;
    double* x = a.x + m * a.ms;
    double *r0 = a.r0 + m * a.ms, *r1 = a.r1 + m * a.ms, *s0 = a.s0 + m * a.ms, *s1 = a.s1 + m * a.ms;
    double* stage = stage_of(a, m, t);
    const int64_t ob = op_base(a, m);
    for (int qq = tid; qq < plane; qq += blockDim.x) {
        const int j = qq / nx, i = qq - j * nx, p = tbase + qq;
        const bool inner = j >= OL && j < OL + sNy && i >= OL && i < OL + sNx;
        s0[p] = 0.0;
        if (!inner) {
            r0[p] = 0.0;
            r1[p] = 0.0;
            s1[p] = 0.0;
        }
    }
    first_guess(a, m, t, rhsNorm);
    MJ_ROWS({
        const double bn0 = b[p0] * nm * rhsNorm;
        const double r_0 = bn0 - op_A(a.aW + ob, a.aS + ob, a.aC + ob, x, p0, nx);
        r0[p0] = r_0;
        double bn1 = 0.0, r_1 = 0.0;
        if (has1) {
            bn1 = b[p1] * nm * rhsNorm;
            r_1 = bn1 - op_A(a.aW + ob, a.aS + ob, a.aC + ob, x, p1, nx);
            r0[p1] = r_1;
        }
        put_row<MODE>(0, jl, lane, r_0 * r_0, r_1 * r_1, has1, stage, sNx, sNy);
        put_row<MODE>(1, jl, lane, bn0, bn1, has1, stage, sNx, sNy);
    })
    store_tile_sums<MODE>(a.perr, a.psum, m * a.T + t, lane, wave, nwaves, stage, sNx, sNy);  // perr slot 0: err(0)
}

// I2: exchange r (ring), z = M r, tile partial of eta(1) = sum z r; tile 0 initialises the member's diag and status.
template <int MODE>
__device__ __forceinline__ void unit_I2(const CgDev& a, int m, int t) {
    MJ_UNIT_IDS
    double* r0 = a.r0 + m * a.ms;
    double* z = a.z + m * a.ms;
    double* stage = stage_of(a, m, t);
    const int64_t ob = op_base(a, m);
    for (int k = a.ring_off[t] + tid; k < a.ring_off[t + 1]; k += blockDim.x) {
        const int s = a.ring_src[k];
        const double v = s < 0 ? 0.0 : r0[s];
        r0[a.ring_dst[k]] = a.ring_sign[k] < 0 ? -v : v;
    }
    __syncthreads();
    MJ_ROWS({
        const double z0 = op_M(a.pW + ob, a.pS + ob, a.pC + ob, r0, p0, nx);
        z[p0] = z0;
        double z1 = 0.0;
        if (has1) {
            z1 = op_M(a.pW + ob, a.pS + ob, a.pC + ob, r0, p1, nx);
            z[p1] = z1;
        }
        put_row<MODE>(0, jl, lane, z0 * r0[p0], has1 ? z1 * r0[p1] : 0.0, has1, stage, sNx, sNy);
    })
    store_tile_sums<MODE>(a.peta + a.B * a.T, nullptr, m * a.T + t, lane, wave, nwaves, stage, sNx, sNy);  // eta(1)
    if (t == 0 && tid == 0) {
        const double err0 = tile_order_sum(a.perr + m * a.T, a.T);
        a.diag[m].firstResidual = sqrt(err0);
        a.diag[m].sumRHS = tile_order_sum(a.psum + m * a.T, a.T);
        a.diag[m].rhsMax = member_rhs_max(a, m);
        a.diag[m].outerIters = 0;
        a.diag[m].launches = 0;
        a.mdone[m] = -1;
    }
}

__device__ __forceinline__ SweepGeo geo_of(const CgDev& a, int t) {
    return SweepGeo{t, t * a.plane, a.sNx, a.sNy, a.OL, a.nx, a.ring_off, a.ring_dst, a.ring_src, a.ring_sign};
}

// Finishing member at B(n): un-normalise x on the tile (cg2d.F:380-395); tile 0 completes the diag.
__device__ __forceinline__ void unit_finish(const CgDev& a, int m, int t, int n, double errp) {
    MJ_UNIT_IDS
    const double rhsNorm = member_rhs_norm(a, m);
    double* x = a.x + m * a.ms;
    MJ_ROWS({
        x[p0] = x[p0] / rhsNorm;
        if (has1) x[p1] = x[p1] / rhsNorm;
    })
    if (t == 0 && tid == 0) {
        a.diag[m].lastResidual = sqrt(errp);
        a.diag[m].numIters = n - 1;
        a.diag[m].converged = errp < a.tol_sq ? 1 : 0;
        a.mdone[m] = n - 1;
        atomicSub(a.nactive, 1);
    }
}

// Phase B(n) of a group (tile t): per member the status at the start of B(n), the same in every block — finished
// earlier, finishing now (stopping rule on err(n-1), cg2d.F:245-250: x un-normalised), or active; the active members
// then sweep together: s_n = z + beta s_{n-1} on interior and ring, q = A s_n, partial of sigma = sum s_n q.
template <int MODE, bool LDS>
__device__ __forceinline__ void group_phase_B(const CgDev& a, int gi, int t, int n) {
    __shared__ int s_m[kGroup], s_code[kGroup], s_act[kGroup], s_nm;
    __shared__ double s_beta[kGroup], s_err[kGroup];
    const int tid = threadIdx.x, BT = a.B * a.T;
    // wave 0 derives each member's status (the partial sums loaded by the lanes in parallel), lane 0 records it
    for (int k = 0; k < kGroup && tid < 64; ++k) {
        const int m = a.grp[gi * kGroup + k];
        int code = 0;
        double beta = 0.0, errp = 0.0;
        if (m >= 0) {
            const int st = a.mdone[m];
            if (!(st >= 0 && st < n - 1)) {
                const double* pe = a.peta + m * a.T;
                const double* const src[3] = {a.perr + ((n - 1) & 1) * BT + m * a.T, pe + (n & 1) * BT,
                                              pe + ((n - 1) & 1) * BT};
                double v[3];
                wave_tile_sums<3>(src, a.T, v);
                errp = v[0];
                code = (!(n - 1 < a.max_iters) || errp < a.tol_sq) ? 1 : 2;
                if (code == 2) {
                    const double etaNM1 = n == 1 ? 1.0 : v[2];
                    beta = v[1] / etaNM1;  // cg2d.F:268-271
                }
            }
        }
        if (tid == 0) {
            s_m[k] = m;
            s_code[k] = code;
            s_beta[k] = beta;
            s_err[k] = errp;
        }
    }
    __syncthreads();
    if (tid == 0) {
        int c = 0;
        for (int k = 0; k < kGroup; ++k)
            if (s_code[k] == 2) s_act[c++] = k;
        s_nm = c;
    }
    __syncthreads();
    for (int k = 0; k < kGroup; ++k)
        if (s_code[k] == 1) unit_finish(a, s_m[k], t, n, s_err[k]);
    const int nm = s_nm;
    if (nm > 0) {
        MemberB<double> mb[kGroup];
#pragma unroll
        for (int k = 0; k < kGroup; ++k) {
            const int m = s_m[s_act[k < nm ? k : 0]];
            mb[k] = MemberB<double>{a.z + m * a.ms, S(a, n - 1) + m * a.ms, S(a, n) + m * a.ms, a.q + m * a.ms,
                                    s_beta[s_act[k < nm ? k : 0]], false, a.psig + (n & 1) * BT + m * a.T + t,
                                    stage_of(a, m, t)};
        }
        const int64_t ob = op_base(a, s_m[s_act[0]]);
        if constexpr (LDS)
            lds_B<double, MODE>(geo_of(a, t), a.aW + ob, a.aS + ob, a.aC + ob, mb[0]);
        else
            group_B<double, kGroup, MODE>(geo_of(a, t), a.aW + ob, a.aS + ob, a.aC + ob, mb, nm);
    }
    __syncthreads();
}

// Phase C(n) of a group (tile t): the members still active sweep together: x += alpha s_n;
// r_n = r_{n-1} - alpha q on interior and ring; z = M r_n; partials of err(n) = sum r_n^2 and eta(n+1) = sum z r_n.
template <int MODE, bool LDS>
__device__ __forceinline__ void group_phase_C(const CgDev& a, int gi, int t, int n) {
    __shared__ int s_m[kGroup], s_act[kGroup], s_nm;
    __shared__ double s_alpha[kGroup];
    const int tid = threadIdx.x, BT = a.B * a.T;
    for (int k = 0; k < kGroup && tid < 64; ++k) {
        const int m = a.grp[gi * kGroup + k];
        const bool act = m >= 0 && a.mdone[m] < 0;
        double alpha = 0.0;
        if (act) {
            const double* const src[2] = {a.peta + (n & 1) * BT + m * a.T, a.psig + (n & 1) * BT + m * a.T};
            double v[2];
            wave_tile_sums<2>(src, a.T, v);
            alpha = v[0] / v[1];
        }
        if (tid == 0) {
            s_m[k] = act ? m : -1;
            s_alpha[k] = alpha;
        }
    }
    __syncthreads();
    if (tid == 0) {
        int c = 0;
        for (int k = 0; k < kGroup; ++k)
            if (s_m[k] >= 0) s_act[c++] = k;
        s_nm = c;
    }
    __syncthreads();
    const int nm = s_nm;
    if (nm > 0) {
        MemberC<double> mc[kGroup];
#pragma unroll
        for (int k = 0; k < kGroup; ++k) {
            const int kk = s_act[k < nm ? k : 0], m = s_m[kk];
            mc[k] = MemberC<double>{S(a, n) + m * a.ms, a.q + m * a.ms, R(a, n - 1) + m * a.ms, R(a, n) + m * a.ms,
                                    a.x + m * a.ms, a.z + m * a.ms, s_alpha[kk], a.perr + (n & 1) * BT + m * a.T + t,
                                    a.peta + ((n + 1) & 1) * BT + m * a.T + t, stage_of(a, m, t)};
        }
        const int64_t ob = op_base(a, s_m[s_act[0]]);
        if constexpr (LDS)
            lds_C<double, MODE>(geo_of(a, t), a.pW + ob, a.pS + ob, a.pC + ob, mc[0]);
        else
            group_C<double, kGroup, MODE>(geo_of(a, t), a.pW + ob, a.pS + ob, a.pC + ob, mc, nm);
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------------------------ persistent
// Units u = blockIdx.x, blockIdx.x + gridDim.x, ...; tile-major (t = u / B), so concurrent blocks share a tile's
// coefficients.
template <int MODE>
__global__ void __launch_bounds__(MJ_CG_THREADS) cg2d_persistent_kernel(CgDev a, GridSync gs) {
    const int nunits = a.B * a.T;
#define MJ_FOR_UNITS(...)                                                   \
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {                  \
        const int t = u / a.B, m = u - t * a.B;                             \
        __VA_ARGS__                                                         \
    }
    MJ_FOR_UNITS(unit_I0(a, m, t);)
    if (!grid_sync(gs)) goto done;
    MJ_FOR_UNITS(unit_I1<MODE>(a, m, t);)
    if (!grid_sync(gs)) goto done;
    MJ_FOR_UNITS(unit_I2<MODE>(a, m, t); __syncthreads();)
    if (!grid_sync(gs)) goto done;
    for (int n = 1;; ++n) {
        const int ng = *a.ngroups, ngu = ng * a.T;
        for (int u = blockIdx.x; u < ngu; u += gridDim.x) group_phase_B<MODE, false>(a, u % ng, u / ng, n);
        if (!grid_sync(gs)) goto done;
        // nactive only changes in B phases: every block reads the same value here
        if (__hip_atomic_load(a.nactive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
        for (int u = blockIdx.x; u < ngu; u += gridDim.x) group_phase_C<MODE, false>(a, u % ng, u / ng, n);
        if (!grid_sync(gs)) goto done;
    }
#undef MJ_FOR_UNITS
done:
    // a timed-out barrier: mark every member's diag (converged = -2) so the host sees the failure
    if (__hip_atomic_load(gs.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 && threadIdx.x == 0) {
        for (int u = blockIdx.x; u < nunits; u += gridDim.x)
            if (u < a.B) a.diag[u].converged = -2;
    }
}

// ------------------------------------------------------------------------------------------------ phases
enum { PH_I0 = 0, PH_I1 = 1, PH_I2 = 2, PH_B = 3, PH_C = 4 };
// the iteration phases of the phase launches stage the stencil operand in LDS (one member per block)
constexpr bool kPhaseLds = MJ_LDS_SWEEPS && kGroup == 1;

template <int MODE, int PH>
__global__ void __launch_bounds__(MJ_PH_THREADS * kGroup) __attribute__((amdgpu_waves_per_eu(MJ_PH_WAVES_PER_EU)))
cg2d_phase_kernel(CgDev a, int n) {
    int m, t;
    if (PH == PH_B || PH == PH_C) {
        if (!group_unit(*a.ngroups, a.T, m, t)) return;  // m: group index
        if (PH == PH_B) group_phase_B<MODE, kPhaseLds>(a, m, t, n);
        if (PH == PH_C) group_phase_C<MODE, kPhaseLds>(a, m, t, n);
        return;
    }
    if (!phase_unit(a.B, a.T, m, t)) return;
    if (PH == PH_I0) unit_I0(a, m, t);
    if (PH == PH_I1) unit_I1<MODE>(a, m, t);
    if (PH == PH_I2) unit_I2<MODE>(a, m, t);
}

// The launch count of the iteration loop into every member's record (written after the loop).
__global__ void diag_launches_kernel(mj_cg2d_diag* diag, int B, int launches) {
    for (int m = threadIdx.x; m < B; m += blockDim.x) diag[m].launches = launches;
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

// Per-tile tables of the T plan for the solve (see mj_halo_plan::cg_*).
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
                // the solve computes ring values from their sources: they must be interior points
                MJ_REQUIRE(interior(P->h_src[0][k]), "cg2d: ring point %d has a non-interior source %d", d,
                           P->h_src[0][k]);
                rdst.push_back(d);
                rsrc.push_back(P->h_src[0][k]);
                rsg.push_back(P->h_sign[0][k]);
            }
        }
        // ring points the exchange leaves unwritten (open facet edges): entries with no source (value 0), so that
        // every ring point of the tile has an entry
        std::vector<char> seen((size_t)plane, 0);
        for (int k = roff[t]; k < (int)rdst.size(); ++k) seen[rdst[k] - t * plane] = 1;
        auto add = [&](int j, int i) {
            const int r = j * nx + i;
            if (seen[r]) return;
            rdst.push_back(t * plane + r);
            rsrc.push_back(-1);
            rsg.push_back(1);
        };
        for (int i = OL; i < OL + sNx; ++i) {
            add(OL - 1, i);
            add(OL + sNy, i);
        }
        for (int j = OL; j < OL + sNy; ++j) {
            add(j, OL - 1);
            add(j, OL + sNx);
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
    // the mixed solve's form, one 16-byte entry per ring point: compact source index (mx_geo; -1: none), padded
    // source, LDS offset with the sign in bit 30
    const MxGeo G = mx_geo(sNx, sNy);
    std::vector<int4> cring(rdst.size());
    for (size_t k = 0; k < rdst.size(); ++k) {
        const int r = rdst[k] % plane, jh = r / nx, ih = r % nx;
        const int lds = G.lds(jh - OL, ih - OL);
        MJ_REQUIRE(lds >= 0 && lds < (1 << 30), "cg2d: LDS offset out of range");
        const int s = rsrc[k];
        int csrc = -1;
        if (s >= 0) {
            const int ts = s / plane, rs = s % plane, js = rs / nx - OL, is = rs % nx - OL;
            csrc = ts * G.TS + js * G.CP + is;
            // a source in the west or east column: its contiguous copy after the tile's rows
            if (is == 0) csrc = ts * G.TS + sNy * G.CP + js;
            if (is == sNx - 1) csrc = ts * G.TS + sNy * G.CP + sNy + js;
        }
        cring[k] = make_int4(csrc, s, lds | (rsg[k] < 0 ? 1 << 30 : 0), 0);
    }
    if ((rc = up(cring.data(), cring.size() * sizeof(int4), (void**)&P->cg_cring))) return rc;
    P->cg_sNx = sNx;
    P->cg_sNy = sNy;
    P->cg_OL = OL;
    return MJ_OK;
}

}  // namespace mj

using namespace mj;

// ------------------------------------------------------------------------------------------------ host side
namespace {

struct WsLayout {
    size_t r0, r1, s0, s1, z, q, pmax, psum, perr, peta, psig, mdone, grp, misc, prod, total;
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
    L.z = o; o += vec;
    L.q = o; o += vec;
    L.pmax = o; o += bt;
    L.psum = o; o += bt;
    L.perr = o; o += 2 * bt;
    L.peta = o; o += 2 * bt;
    L.psig = o; o += 2 * bt;
    L.mdone = o; o += al((size_t)c->batch * 4);
    L.grp = o; o += al((size_t)c->batch * kGroup * 4);
    L.misc = o; o += 256;  // barrier counter, error flag, nactive, ngroups
    L.prod = o;
    if (c->sum_order == MJ_SUM_FORTRAN) o += al((size_t)c->batch * c->ntiles * 2 * c->sNx * c->sNy * 8);
    L.total = o;
    return L;
}

int launch_mode_of(const mj_cg2d_config* c) {
    if (c->launch_mode != MJ_CG_LAUNCH_AUTO) return c->launch_mode;
    // one-tile-per-block phases fill the chip once the batch has a few hundred units; below that the persistent
    // launch (no launch gaps, no host round trips) is faster
    return c->batch * c->ntiles >= 256 ? MJ_CG_LAUNCH_PHASES : MJ_CG_LAUNCH_PERSISTENT;
}

int block_threads_of(const mj_cg2d_config* c, int mode) {
    const int limit = mode == MJ_CG_LAUNCH_PERSISTENT ? MJ_CG_THREADS : MJ_PH_THREADS * kGroup;
    if (c->block_threads == 0) return mode == MJ_CG_LAUNCH_PERSISTENT ? limit : MJ_PH_DEFAULT_THREADS * kGroup;
    MJ_REQUIRE(c->block_threads % 64 == 0 && c->block_threads >= 64 && c->block_threads <= limit,
               "cg2d: block_threads must be a multiple of 64 in 64..%d for this launch mode", limit);
    return c->block_threads;
}

int check_cfg(const mj_cg2d_config* c) {
    MJ_REQUIRE(c, "cg2d: null config");
    MJ_REQUIRE(c->batch >= 1 && c->batch <= (1 << 20), "cg2d: batch %d out of range", c->batch);
    MJ_REQUIRE(c->ntiles >= 1 && c->OL >= 1, "cg2d: bad layout");
    MJ_REQUIRE(c->sNx >= 64 && c->sNx <= 128 && c->sNy >= 64 && c->sNy <= 128,
               "cg2d: sNx, sNy must be in 64..128 (got %d, %d)", c->sNx, c->sNy);
    // the row sweeps move column pairs as 16-byte accesses: every row start must be an even element offset
    MJ_REQUIRE(c->sNx % 2 == 0 && c->OL % 2 == 0 && c->member_stride % 2 == 0,
               "cg2d: sNx, OL and member_stride must be even (got %d, %d, %lld)", c->sNx, c->OL,
               (long long)c->member_stride);
    const int64_t plane = (int64_t)(c->sNx + 2 * c->OL) * (c->sNy + 2 * c->OL);
    MJ_REQUIRE(c->member_stride >= c->ntiles * plane, "cg2d: member_stride too small");
    MJ_REQUIRE((int64_t)c->batch * c->member_stride < ((int64_t)1 << 31), "cg2d: batch * member_stride exceeds int32");
    MJ_REQUIRE(c->sum_order == MJ_SUM_FORTRAN || c->sum_order == MJ_SUM_TREE, "cg2d: bad sum_order %d", c->sum_order);
    MJ_REQUIRE(c->max_iters >= 0, "cg2d: max_iters < 0");
    MJ_REQUIRE(c->launch_mode >= MJ_CG_LAUNCH_AUTO && c->launch_mode <= MJ_CG_LAUNCH_PHASES_FIXED,
               "cg2d: bad launch_mode %d", c->launch_mode);
    MJ_REQUIRE(c->poll_iters >= 0, "cg2d: poll_iters < 0");
    const int bt = block_threads_of(c, launch_mode_of(c));
    if (bt < 0) return bt;
    return MJ_OK;
}

template <int MODE>
int launch_phases(const CgDev& a, const mj_cg2d_config* cfg, int threads, int mode, hipStream_t st) {
    const int nunits = a.B * a.T;
    const dim3 grid((unsigned)((nunits + kXcds - 1) / kXcds * kXcds)), block((unsigned)threads);
    const dim3 iblock((unsigned)std::min(threads, MJ_PH_THREADS));
    const size_t lds = kPhaseLds ? lds_tile_bytes(a.sNx, a.sNy, (int)sizeof(double)) : 0;
    if (lds > 65536) {
        MJ_HIP(hipFuncSetAttribute((const void*)cg2d_phase_kernel<MODE, PH_B>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        MJ_HIP(hipFuncSetAttribute((const void*)cg2d_phase_kernel<MODE, PH_C>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    auto go = [&](int ph, int n) -> int {
        switch (ph) {
            case PH_I0: hipLaunchKernelGGL((cg2d_phase_kernel<MODE, PH_I0>), grid, iblock, 0, st, a, n); break;
            case PH_I1: hipLaunchKernelGGL((cg2d_phase_kernel<MODE, PH_I1>), grid, iblock, 0, st, a, n); break;
            case PH_I2: hipLaunchKernelGGL((cg2d_phase_kernel<MODE, PH_I2>), grid, iblock, 0, st, a, n); break;
            case PH_B: hipLaunchKernelGGL((cg2d_phase_kernel<MODE, PH_B>), grid, block, lds, st, a, n); break;
            default: hipLaunchKernelGGL((cg2d_phase_kernel<MODE, PH_C>), grid, block, lds, st, a, n); break;
        }
        MJ_HIP(hipGetLastError());
        return MJ_OK;
    };
    int rc;
    if ((rc = go(PH_I0, 0)) || (rc = go(PH_I1, 0)) || (rc = go(PH_I2, 0))) return rc;
    // B(n) for n = 1 .. max_iters + 1 (the last one only finishes members), C(n) for n <= max_iters
    const int nlast = cfg->max_iters + 1;
    const int chunk = cfg->poll_iters > 0 ? cfg->poll_iters : 32;
    int n = 1, launches = 0;
    while (n <= nlast) {
        const int nend = mode == MJ_CG_LAUNCH_PHASES_FIXED ? nlast + 1 : std::min(nlast + 1, n + chunk);
        for (; n < nend; ++n) {
            if ((rc = go(PH_B, n))) return rc;
            ++launches;
            if (n <= cfg->max_iters) {
                if ((rc = go(PH_C, n))) return rc;
                ++launches;
            }
        }
        if (n > nlast) break;
        int h = 0;
        MJ_HIP(hipMemcpyAsync(&h, a.nactive, sizeof(int), hipMemcpyDeviceToHost, st));
        MJ_HIP(hipStreamSynchronize(st));
        if (h == 0) break;
    }
    hipLaunchKernelGGL(diag_launches_kernel, dim3(1), dim3(256), 0, st, a.diag, a.B, launches);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}

}  // namespace

extern "C" size_t mj_cg2d_workspace_bytes(const mj_cg2d_config* cfg) {
    if (check_cfg(cfg)) return 0;
    return ws_layout(cfg).total;
}

namespace {

// The solve behind mj_cg2d_solve and mj_cg2d_solve_noalloc. With may_alloc false it allocates nothing: the T plan's
// CG tables must already exist (built by an earlier call that may allocate).
int solve_impl(const mj_halo_plan* plan_T, const mj_cg2d_ops* ops, const double* b, const double* x0, double* x,
               const double* cg2dNorm, const mj_cg2d_config* cfg, void* workspace, size_t workspace_bytes,
               mj_cg2d_diag* diag, void* stream, bool may_alloc) {
    int rc = check_cfg(cfg);
    if (rc) return rc;
    MJ_REQUIRE(plan_T && ops && b && x0 && x && cg2dNorm && workspace && diag, "mj_cg2d_solve: null argument");
    MJ_REQUIRE(ops->aW2d && ops->aS2d && ops->aC2d && ops->pW && ops->pS && ops->pC, "mj_cg2d_solve: null operator");
    MJ_REQUIRE(x != b && x != x0, "mj_cg2d_solve: x must not alias b or x0");
    MJ_REQUIRE(ops->member_stride % 2 == 0, "mj_cg2d_solve: operator member_stride must be even");
    for (const void* q : {(const void*)x, (const void*)ops->aW2d, (const void*)ops->aS2d, (const void*)ops->aC2d,
                          (const void*)ops->pW, (const void*)ops->pS, (const void*)ops->pC})
        MJ_REQUIRE(((uintptr_t)q & 15) == 0, "mj_cg2d_solve: x and the operator arrays must be 16-byte aligned");
    MJ_REQUIRE(((uintptr_t)workspace & 255) == 0, "mj_cg2d_solve: workspace must be 256-byte aligned");
    const WsLayout L = ws_layout(cfg);
    MJ_REQUIRE(workspace_bytes >= L.total, "mj_cg2d_solve: workspace %zu < %zu bytes", workspace_bytes, L.total);
    mj_halo_plan* P = const_cast<mj_halo_plan*>(plan_T);
    MJ_REQUIRE(P->ntiles == cfg->ntiles, "mj_cg2d_solve: plan has %d tiles, config %d", P->ntiles, cfg->ntiles);
    if (may_alloc) {
        if ((rc = ensure_cg_tables(P, cfg->sNx, cfg->sNy, cfg->OL))) return rc;
    } else {
        std::lock_guard<std::mutex> lk(P->cg_mu);
        MJ_REQUIRE(P->cg_full_off && P->cg_sNx == cfg->sNx && P->cg_sNy == cfg->sNy && P->cg_OL == cfg->OL,
                   "mj_cg2d_solve_noalloc: the T plan has no CG tables for this layout yet (they are built, with "
                   "device allocations, by the plan's first mj_cg2d_solve / mj_cg2d_solve_mixed call)");
    }
    const int mode = launch_mode_of(cfg);
    const int threads = block_threads_of(cfg, mode);
    if (threads < 0) return threads;

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
    a.z = (double*)(w + L.z);
    a.q = (double*)(w + L.q);
    a.pmax = (double*)(w + L.pmax);
    a.psum = (double*)(w + L.psum);
    a.perr = (double*)(w + L.perr);
    a.peta = (double*)(w + L.peta);
    a.psig = (double*)(w + L.psig);
    a.op_index = ops->op_index;
    a.mdone = (int32_t*)(w + L.mdone);
    a.prod = (double*)(w + L.prod);
    unsigned int* counter = (unsigned int*)(w + L.misc);
    int* error = (int*)(w + L.misc + 64);
    a.nactive = (int*)(w + L.misc + 128);
    a.ngroups = (int*)(w + L.misc + 192);
    a.grp = (int32_t*)(w + L.grp);
    a.diag = diag;

    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(cg2d_init_kernel, dim3(1), dim3(64), 0, st, counter, error, a.nactive, a.B);
    MJ_HIP(hipGetLastError());
    hipLaunchKernelGGL(build_groups_kernel, dim3(1), dim3(1024), 0, st, a.op_index, a.oms, a.B, a.grp, a.ngroups);
    MJ_HIP(hipGetLastError());
    if (mode != MJ_CG_LAUNCH_PERSISTENT)
        return cfg->sum_order == MJ_SUM_FORTRAN ? launch_phases<MJ_SUM_FORTRAN>(a, cfg, threads, mode, st)
                                                : launch_phases<MJ_SUM_TREE>(a, cfg, threads, mode, st);

    int dev = 0;
    MJ_HIP(hipGetDevice(&dev));
    int ncu = 0, coop = 0, clk_khz = 0;
    MJ_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    MJ_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    MJ_HIP(hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeWallClockRate, dev));
    MJ_REQUIRE(coop, "mj_cg2d_solve: device does not support cooperative launches");
    const void* kfn = cfg->sum_order == MJ_SUM_FORTRAN ? (const void*)cg2d_persistent_kernel<MJ_SUM_FORTRAN>
                                                       : (const void*)cg2d_persistent_kernel<MJ_SUM_TREE>;
    int per_cu = 0;
    MJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, threads, 0));
    MJ_REQUIRE(per_cu >= 1, "mj_cg2d_solve: kernel cannot be resident with %d threads", threads);
    int grid = std::min(a.B * a.T, per_cu * ncu);
    if (cfg->grid_blocks > 0) grid = std::min(grid, cfg->grid_blocks);
    GridSync gs{counter, error, (unsigned)grid, 0u,
                (unsigned long long)(clk_khz > 0 ? clk_khz : 100000) * 1000ull * 4ull /* 4 s */};
    void* args[] = {(void*)&a, (void*)&gs};
    MJ_HIP(hipLaunchCooperativeKernel(kfn, dim3(grid), dim3(threads), args, 0u, st));
    return MJ_OK;
}

}  // namespace

// Scratch of mj_cg2d_solve_noalloc: one copy per device, held until the process ends. Host code never uses this
// name as an address: it passes the device address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[MJ_CG2D_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)MJ_CG2D_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch` (looked up once per device, with that device current).
int scratch_of_current_device(void** out) {
    int dev = 0;
    MJ_HIP(hipGetDevice(&dev));
    MJ_REQUIRE(dev >= 0 && dev < kMaxDevices, "mj_cg2d_solve_noalloc: device index %d >= %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        MJ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        MJ_REQUIRE(p != nullptr, "mj_cg2d_solve_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MJ_OK;
}

}  // namespace

extern "C" int mj_cg2d_solve(const mj_halo_plan* plan_T, const mj_cg2d_ops* ops, const double* b, const double* x0,
                             double* x, const double* cg2dNorm, const mj_cg2d_config* cfg, void* workspace,
                             size_t workspace_bytes, mj_cg2d_diag* diag, void* stream) {
    return solve_impl(plan_T, ops, b, x0, x, cg2dNorm, cfg, workspace, workspace_bytes, diag, stream, true);
}

extern "C" int mj_cg2d_solve_noalloc(const mj_halo_plan* plan_T, const mj_cg2d_ops* ops, const double* b,
                                     const double* x0, double* x, const double* cg2dNorm, const mj_cg2d_config* cfg,
                                     mj_cg2d_diag* diag, void* stream) {
    int rc = check_cfg(cfg);
    if (rc) return rc;
    const size_t need = ws_layout(cfg).total;
    MJ_REQUIRE(need <= kNoAllocScratchBytes,
               "mj_cg2d_solve_noalloc: this configuration needs %zu bytes of scratch, more than the %zu the library "
               "holds per device (use mj_cg2d_solve with a workspace)", need, kNoAllocScratchBytes);
    void* ws = nullptr;
    if ((rc = scratch_of_current_device(&ws))) return rc;
    return solve_impl(plan_T, ops, b, x0, x, cg2dNorm, cfg, ws, kNoAllocScratchBytes, diag, stream, false);
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
                       ops->member_stride, ops->op_index, (const double*)work, y, cg2dNorm, cfg->member_stride,
                       cfg->ntiles, ny, nx, cfg->OL, cfg->sNx, cfg->sNy);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}
