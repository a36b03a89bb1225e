// GAD_IMPLICIT_R on gfx950: the implicit vertical advection (3rd-order upwind) + diffusion solve of a tracer
// (reference core/tracer_implicit.py:34-173; gad_implicit_r.F, gad_u3c4_impl_r.F, solve_pentadiagonal.F).
//
// One thread per water column (every padded column of every member; halo columns carry identity rows, as the
// reference's matrix does). The thread walks k = 1..Nr once: it builds row k of the pentadiagonal matrix from the
// level data of k-1..k+2 (diffusion from kappaRX and the masks; advection from the U3 coefficients of the interface
// below level k, computed here, and of the interface above it, kept from row k-1) and eliminates the row at once
// (the forward sweep of SOLVE_DIAGONAL_KINNER), keeping the normalised super-diagonals d', e' in LDS and y' in the
// output column; the back substitution then walks k = Nr..1. Every expression is the reference's, in its order
// (-ffp-contract=off), so the result is bitwise the reference's. Level data are loaded kPf levels ahead of the row
// being built (a register ring with compile-time slots), so each lane keeps up to 5 x kPf loads in flight; the
// columns of a wave are consecutive points of a row (coalesced). The per-level constants sit in LDS and the rows
// between the first two and the last two run branch-free, so that no wait for a load is forced before its use.
#include "mj_common.h"

namespace mj {
namespace {

#ifndef MJ_IMPR_REG50
#define MJ_IMPR_REG50 1
#endif
constexpr int kImprThreads = 64;  // one wave per workgroup (LDS: 2 * nr * 64 + 3 * nr doubles)
#ifndef MJ_IMPR_PF
#define MJ_IMPR_PF 16
#endif
constexpr int kPf = MJ_IMPR_PF;  // prefetch distance in levels
static_assert(kPf >= 3, "the ring must hold levels m .. m+2");
constexpr double kSixth = 1.0 / 6.0;

struct ImprArgs {
    int T, nx, ny, sNx, sNy, OL, nr, adv;
    int64_t plane, ms3, mms;
    const double *maskC, *rA, *recip_rA, *dt, *rdrF, *rdrC;
    double rkSign;
    const double *kap, *rh, *w, *y;
    double* out;
};

// U3 coefficients of interface K (0-based, between levels K-1 and K; Fortran k = K+1), reference
// tracer_implicit.py:71-86: mCup = maskC(K-1), mCkm2 = maskC(max(0, K-2)), mCkp1 = maskC(min(nr-1, K+1)).
struct Iface {
    double rC, rUp, km, kp;
};
__device__ __forceinline__ Iface iface(const ImprArgs& a, int K, double w, double rA, double rrA, double mCup,
                                       double mCkm2, double mCkp1) {
    Iface f;
    const double rTrans = (w * rA) * mCup;
    f.rC = ((0.5 * rTrans) * rrA) * a.rkSign;
    const double mskM = mCkm2 * (K <= 1 ? 0.0 : 1.0);
    const double mskP = mCkp1 * (K >= a.nr - 1 ? 0.0 : 1.0);
    const double ab = fabs(f.rC);
    f.rUp = (2.0 * kSixth) * ab;
    f.km = (kSixth * (f.rC + ab)) * mskM;
    f.kp = (kSixth * (f.rC - ab)) * mskP;
    return f;
}

// solve_pentadiagonal.F:256-267 (reference tracer_implicit.py:120-124)
__device__ __forceinline__ void norm3(double cp, double& dp, double& ep, double& yp) {
    const bool nz = cp != 0.0;
    const double rec = 1.0 / (nz ? cp : 1.0);
    dp = nz ? dp * rec : 0.0;
    ep = nz ? ep * rec : 0.0;
    yp = nz ? yp * rec : 0.0;
}

// Per-thread state of the sweep down a column.
struct ColState {
    double mC1;            // maskC of level m-1
    Iface up;              // interface m (above level m), built by row m-1
    double dp1, ep1, yp1;  // row m-1 after elimination
    double dp2, ep2, yp2;  // row m-2
};

// Row m of the matrix and its elimination. GEN: any m (the branches of the first and last rows); otherwise
// 2 <= m <= nr-3, branch-free, so that the loads of the ring stay in flight across rows. Level data: mC0, rh0, kap0,
// y0 of level m; mCn, kapn, wn of level m+1; mCk = maskC(min(nr-1, m+2)). Halo columns (in = false) get identity
// rows; their level data are loaded like the others' and ignored.
// STORE: write d', e' to LDS and y' to the output column (else the caller keeps them, from c); NRC: the level count
// when known at compile time (0: a.nr).
template <bool GEN, bool ADV, bool STORE = true, int NRC = 0>
__device__ __forceinline__ void impr_row(const ImprArgs& a, int m, bool in, double rA, double rrA, double mC0,
                                         double rh0, double kap0, double y0, double mCn, double kapn, double wn,
                                         double mCk, ColState& c, double* dpL, double* epL, double* out,
                                         const double* lv) {
    const int nr = NRC ? NRC : a.nr;
    const double dtm = lv[m], rdf = lv[nr + m];
    const double* rdrC = lv + 2 * nr;
    const bool hasU = !GEN || m >= 1, hasL = !GEN || m + 1 < nr;
    // diffusion, reference tracer_implicit.py:54-64
    double rb = 0.0, rd = 0.0, ra = 0.0, re = 0.0;
    if (hasU) rb = -(((((dtm * c.mC1) * rh0) * rdf) * kap0) * rdrC[m]);
    if (hasL) rd = -(((((dtm * mCn) * rh0) * rdf) * kapn) * rdrC[m + 1]);
    double rc = 1.0 - (rb + rd);
    if (ADV) {
        // U3 advection, :95-105: level m takes the updates of interface m+1 (its bottom) first, then those of
        // interface m (its top)
        Iface lo{0.0, 0.0, 0.0, 0.0};
        if (hasL) {
            lo = iface(a, m + 1, wn, rA, rrA, mC0, (!GEN || m >= 1) ? c.mC1 : mC0, mCk);
            rb = rb - ((lo.km * dtm) * rh0) * rdf;
            rc = rc + (((((lo.rC + lo.rUp) + lo.km) * dtm) * rh0) * rdf);
            rd = rd + (((((lo.rC - lo.rUp) + lo.kp) * dtm) * rh0) * rdf);
            re = re - ((lo.kp * dtm) * rh0) * rdf;
        }
        if (hasU) {
            ra = ra + ((c.up.km * dtm) * rh0) * rdf;
            rb = rb - (((((c.up.rC + c.up.rUp) + c.up.km) * dtm) * rh0) * rdf);
            rc = rc - (((((c.up.rC - c.up.rUp) + c.up.kp) * dtm) * rh0) * rdf);
            rd = rd + ((c.up.kp * dtm) * rh0) * rdf;
        }
        c.up = lo;
    }
    if (!in) ra = 0.0, rb = 0.0, rc = 1.0, rd = 0.0, re = 0.0;
    // forward elimination, reference tracer_implicit.py:126-143 (rows 0 and 1 through their own formulas)
    double dq, eq = re, yq;
    if (GEN && m == 0) {
        dq = rd;
        yq = y0;
        norm3(rc, dq, eq, yq);
    } else if (GEN && m == 1) {
        const double c2 = rc - rb * c.dp1;
        dq = rd - rb * c.ep1;
        yq = y0 - rb * c.yp1;
        norm3(c2, dq, eq, yq);
    } else {
        const double bb = rb - ra * c.dp2;
        const double cq = (rc - ra * c.ep2) - bb * c.dp1;
        dq = rd - bb * c.ep1;
        yq = (y0 - ra * c.yp2) - bb * c.yp1;
        norm3(cq, dq, eq, yq);
    }
    if (STORE) {
        dpL[(int64_t)m * kImprThreads] = dq;
        epL[(int64_t)m * kImprThreads] = eq;
        out[(int64_t)m * a.plane] = yq;
    }
    c.dp2 = c.dp1, c.ep2 = c.ep1, c.yp2 = c.yp1;
    c.dp1 = dq, c.ep1 = eq, c.yp1 = yq;
    c.mC1 = mC0;
}

template <bool ADV>
__global__ void __launch_bounds__(kImprThreads) impr_kernel(ImprArgs a) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int64_t col = (int64_t)blockIdx.x * kImprThreads + lane;
    const int mem = blockIdx.y;
    // the per-level constants (deltaT, recip_drF, recip_drC) in LDS: read per row without vector-memory waits
    double* lv = lds + (int64_t)2 * a.nr * kImprThreads;
    for (int k = lane; k < a.nr; k += kImprThreads) {
        lv[k] = a.dt[k];
        lv[a.nr + k] = a.rdrF[k];
        lv[2 * a.nr + k] = a.rdrC[k];
    }
    __syncthreads();
    if (col >= (int64_t)a.T * a.plane) return;  // no barriers below
    const int t = (int)(col / a.plane), q = (int)(col - (int64_t)t * a.plane);
    const int j = q / a.nx, i = q - j * a.nx;
    const bool in = j >= a.OL && j < a.OL + a.sNy && i >= a.OL && i < a.OL + a.sNx;
    const int nr = a.nr;
    const int64_t o3 = mem * a.ms3 + (int64_t)t * nr * a.plane + q;
    const int64_t om = mem * a.mms + (int64_t)t * nr * a.plane + q;
    const double* y = a.y + o3;
    double* out = a.out + o3;
    double* dpL = lds + lane;
    double* epL = lds + (int64_t)nr * kImprThreads + lane;
    const int64_t o2 = (int64_t)t * a.plane + q;
    const double rA = ADV ? a.rA[o2] : 0.0;
    const double rrA = ADV ? a.recip_rA[o2] : 0.0;
    // as in impr_reg_kernel: halo columns re-read the tracer's lines
    const double* mCp = in ? a.maskC + om : y;
    const double *rhp = in ? a.rh + o3 : y, *kp = in ? a.kap + o3 : y, *wp = in ? a.w + o3 : y;

    // ring of level data: slot s holds a level m with m % kPf == s; loads are unconditional (levels clamped to
    // nr-1) so that no branch separates a load from its use
    double rmC[kPf], rrh[kPf], rkap[kPf], rw[kPf], ry[kPf];
    auto load = [&](int s, int lev) {
        const int64_t ok = (int64_t)min(lev, nr - 1) * a.plane;
        ry[s] = y[ok];
        rmC[s] = mCp[ok];
        rrh[s] = rhp[ok];
        rkap[s] = kp[ok];
        rw[s] = wp[ok];
    };
#pragma unroll
    for (int s = 0; s < kPf; ++s) load(s, s);
    ColState c{0.0, {0.0, 0.0, 0.0, 0.0}, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    auto mck = [&](int m, int s2, int s1) { return m + 2 < nr ? rmC[s2] : rmC[s1]; };
    // rows 0 and 1
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (s < nr) {
            const int s1 = (s + 1) % kPf, s2 = (s + 2) % kPf;
            impr_row<true, ADV>(a, s, in, rA, rrA, rmC[s], rrh[s], rkap[s], ry[s], rmC[s1], rkap[s1], rw[s1],
                           mck(s, s2, s1), c, dpL, epL, out, lv);
            load(s, s + kPf);
        }
    }
    // rows 2 .. nr-3 in laps of kPf (a lap starts at m = 2 mod kPf, so slot (2 + s) % kPf is level m0 + s)
    int m0 = 2;
    for (; m0 + kPf - 1 <= nr - 3; m0 += kPf) {
#pragma unroll
        for (int s = 0; s < kPf; ++s) {
            const int m = m0 + s;
            const int s0 = (2 + s) % kPf, s1 = (3 + s) % kPf, s2 = (4 + s) % kPf;
            impr_row<false, ADV>(a, m, in, rA, rrA, rmC[s0], rrh[s0], rkap[s0], ry[s0], rmC[s1], rkap[s1], rw[s1],
                            rmC[s2], c, dpL, epL, out, lv);
            load(s0, m + kPf);
        }
    }
    // the remaining rows (at most kPf + 1)
#pragma unroll
    for (int s = 0; s < kPf + 2; ++s) {
        const int m = m0 + s;
        if (m < nr) {
            const int s0 = (2 + s) % kPf, s1 = (3 + s) % kPf, s2 = (4 + s) % kPf;
            impr_row<true, ADV>(a, m, in, rA, rrA, rmC[s0], rrh[s0], rkap[s0], ry[s0], rmC[s1], rkap[s1], rw[s1],
                           mck(m, s2, s1), c, dpL, epL, out, lv);
            load(s0, m + kPf);
        }
    }
    // back substitution, reference tracer_implicit.py:150-160: u(nr-1) = y'(nr-1) is in place already
    if (nr < 2) return;
    double u1 = c.yp1;  // u(m+1)
    double u0 = c.yp2 - u1 * dpL[(int64_t)(nr - 2) * kImprThreads];
    out[(int64_t)(nr - 2) * a.plane] = u0;
    double u2 = u1;
    u1 = u0;
    // y'(m) for m = nr-3 .. 0 are re-read from the output column, kPf levels ahead (clamped to level 0)
    double ryp[kPf];
#pragma unroll
    for (int s = 0; s < kPf; ++s) ryp[s] = out[(int64_t)max(nr - 3 - s, 0) * a.plane];
    for (int mb = nr - 3; mb >= 0; mb -= kPf) {
#pragma unroll
        for (int s = 0; s < kPf; ++s) {
            const int m = mb - s;
            if (m >= 0) {
                const double u = (ryp[s] - u1 * dpL[(int64_t)m * kImprThreads]) - u2 * epL[(int64_t)m * kImprThreads];
                out[(int64_t)m * a.plane] = u;
                u2 = u1;
                u1 = u;
                ryp[s] = out[(int64_t)max(m - kPf, 0) * a.plane];
            }
        }
    }
}

// The same sweep for a level count known at compile time: the rows are unrolled, so y' of every level stays in
// registers (d', e' in LDS as before) and makes no round trip through the output column: each column is read once
// and written once. Bitwise the generic kernel's arithmetic.
constexpr int kPfR = 8;
template <bool ADV, int NR>
__global__ void __launch_bounds__(kImprThreads) impr_reg_kernel(ImprArgs a) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int64_t col = (int64_t)blockIdx.x * kImprThreads + lane;
    const int mem = blockIdx.y;
    double* lv = lds;
    for (int k = lane; k < NR; k += kImprThreads) {
        lv[k] = a.dt[k];
        lv[NR + k] = a.rdrF[k];
        lv[2 * NR + k] = a.rdrC[k];
    }
    __syncthreads();
    if (col >= (int64_t)a.T * a.plane) return;  // no barriers below
    const int t = (int)(col / a.plane), q = (int)(col - (int64_t)t * a.plane);
    const int j = q / a.nx, i = q - j * a.nx;
    const bool in = j >= a.OL && j < a.OL + a.sNy && i >= a.OL && i < a.OL + a.sNx;
    const int64_t o3 = mem * a.ms3 + (int64_t)t * NR * a.plane + q;
    const int64_t om = mem * a.mms + (int64_t)t * NR * a.plane + q;
    const double* y = a.y + o3;
    double* out = a.out + o3;
    const int64_t o2 = (int64_t)t * a.plane + q;
    const double rA = ADV ? a.rA[o2] : 0.0;
    const double rrA = ADV ? a.recip_rA[o2] : 0.0;
    // a halo column only needs the tracer (identity rows): its other loads read the tracer's lines again, which
    // the wave loads anyway (no HBM traffic)
    const double* mCp = in ? a.maskC + om : y;
    const double *rhp = in ? a.rh + o3 : y, *kp = in ? a.kap + o3 : y, *wp = in ? a.w + o3 : y;
    double rmC[kPfR], rrh[kPfR], rkap[kPfR], rw[kPfR], ry[kPfR];
    auto load = [&](int s, int lev) {
        const int64_t ok = (int64_t)lev * a.plane;
        ry[s] = y[ok];
        rmC[s] = mCp[ok];
        rrh[s] = rhp[ok];
        rkap[s] = kp[ok];
        rw[s] = wp[ok];
    };
#pragma unroll
    for (int s = 0; s < kPfR; ++s) load(s, s < NR ? s : NR - 1);
    ColState c{0.0, {0.0, 0.0, 0.0, 0.0}, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double* dpL = lds + 3 * NR + lane;
    double* epL = dpL + NR * kImprThreads;
    double ypR[NR];
#pragma unroll
    for (int m = 0; m < NR; ++m) {
        const int s = m % kPfR, s1 = (m + 1) % kPfR, s2 = (m + 2) % kPfR;
        const double mCn = m + 1 < NR ? rmC[s1] : 0.0;
        const double mCk = m + 2 < NR ? rmC[s2] : rmC[s1];
        if (m <= 1 || m >= NR - 2)
            impr_row<true, ADV, false, NR>(a, m, in, rA, rrA, rmC[s], rrh[s], rkap[s], ry[s], mCn, rkap[s1], rw[s1],
                                           mCk, c, nullptr, nullptr, nullptr, lv);
        else
            impr_row<false, ADV, false, NR>(a, m, in, rA, rrA, rmC[s], rrh[s], rkap[s], ry[s], mCn, rkap[s1], rw[s1],
                                            mCk, c, nullptr, nullptr, nullptr, lv);
        dpL[m * kImprThreads] = c.dp1;
        epL[m * kImprThreads] = c.ep1;
        ypR[m] = c.yp1;
        if (m + kPfR < NR) load(s, m + kPfR);
    }
    double u1 = ypR[NR - 1];
    out[(int64_t)(NR - 1) * a.plane] = u1;
    double u0 = ypR[NR - 2] - u1 * dpL[(NR - 2) * kImprThreads];
    out[(int64_t)(NR - 2) * a.plane] = u0;
    double u2 = u1;
    u1 = u0;
#pragma unroll
    for (int m = NR - 3; m >= 0; --m) {
        const double u = (ypR[m] - u1 * dpL[m * kImprThreads]) - u2 * epL[m * kImprThreads];
        out[(int64_t)m * a.plane] = u;
        u2 = u1;
        u1 = u;
    }
}

}  // namespace
}  // namespace mj

using namespace mj;

extern "C" int mj_gad_implicit_r(const mj_impr_grid* g, const mj_impr_config* c, const double* kappaRX,
                                 const double* recip_hFac, const double* wFld, const double* gTracer, double* out,
                                 void* stream) {
    MJ_REQUIRE(g && c, "mj_gad_implicit_r: null grid/config");
    MJ_REQUIRE(c->batch >= 1 && c->batch <= 65535 && c->ntiles >= 1 && c->nr >= 1 && c->nr <= 150,
               "mj_gad_implicit_r: batch in 1..65535, ntiles >= 1, nr in 1..150");
    MJ_REQUIRE(c->sNx >= 1 && c->sNy >= 1 && c->OL >= 0, "mj_gad_implicit_r: bad tile layout");
    MJ_REQUIRE(gTracer && out, "mj_gad_implicit_r: null gTracer/out");
    const int nx = c->sNx + 2 * c->OL, ny = c->sNy + 2 * c->OL;
    const int64_t plane = (int64_t)nx * ny, member = (int64_t)c->ntiles * c->nr * plane;
    MJ_REQUIRE(c->batch == 1 || (c->member_stride3d >= member && (c->mask_member_stride == 0 ||
                                                                  c->mask_member_stride >= member)),
               "mj_gad_implicit_r: member strides smaller than one member's arrays");
    MJ_REQUIRE(c->nr == 1 || (kappaRX && recip_hFac && g->maskC && g->dTtracerLev && g->recip_drF && g->recip_drC),
               "mj_gad_implicit_r: null kappaRX/recip_hFac or grid field");
    MJ_REQUIRE(!c->implicitAdvection || (wFld && g->rA && g->recip_rA),
               "mj_gad_implicit_r: implicit advection needs wFld, rA, recip_rA");
    const double* ins[] = {kappaRX, recip_hFac, wFld};
    for (const double* p : ins) MJ_REQUIRE(p != out, "mj_gad_implicit_r: out aliases an input other than gTracer");
    hipStream_t st = as_stream(stream);
    const int64_t cols = (int64_t)c->ntiles * plane;
    if (c->nr == 1) {  // reference tracer_implicit.py:169-170: returned unchanged
        if (out != gTracer)
            MJ_HIP(hipMemcpy2DAsync(out, (size_t)(c->batch == 1 ? cols : c->member_stride3d) * sizeof(double),
                                    gTracer, (size_t)(c->batch == 1 ? cols : c->member_stride3d) * sizeof(double),
                                    (size_t)cols * sizeof(double), (size_t)c->batch, hipMemcpyDeviceToDevice, st));
        return MJ_OK;
    }
    ImprArgs a;
    a.T = c->ntiles;
    a.nx = nx;
    a.ny = ny;
    a.sNx = c->sNx;
    a.sNy = c->sNy;
    a.OL = c->OL;
    a.nr = c->nr;
    a.adv = c->implicitAdvection ? 1 : 0;
    a.plane = plane;
    a.ms3 = c->member_stride3d;
    a.mms = c->mask_member_stride;
    a.maskC = g->maskC;
    a.rA = g->rA;
    a.recip_rA = g->recip_rA;
    a.dt = g->dTtracerLev;
    a.rdrF = g->recip_drF;
    a.rdrC = g->recip_drC;
    a.rkSign = g->rkSign;
    a.kap = kappaRX;
    a.rh = recip_hFac;
    a.w = c->implicitAdvection ? wFld : gTracer;  // without advection the w loads read a valid dummy
    a.y = gTracer;
    a.out = out;
    const size_t lds = ((size_t)2 * c->nr * kImprThreads + 3 * (size_t)c->nr) * sizeof(double);
    const dim3 grid((unsigned)((cols + kImprThreads - 1) / kImprThreads), (unsigned)c->batch);
#if MJ_IMPR_REG50
    if (c->nr == 50) {  // the LLC90 level count: the register-resident sweep
        const size_t lds50 = (3 * 50 + 2 * 50 * kImprThreads) * sizeof(double);
        if (a.adv)
            hipLaunchKernelGGL((impr_reg_kernel<true, 50>), grid, dim3(kImprThreads), lds50, st, a);
        else
            hipLaunchKernelGGL((impr_reg_kernel<false, 50>), grid, dim3(kImprThreads), lds50, st, a);
        MJ_HIP(hipGetLastError());
        return MJ_OK;
    }
#endif
    const void* kern = a.adv ? (const void*)impr_kernel<true> : (const void*)impr_kernel<false>;
    if (lds > 65536) MJ_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (a.adv)
        hipLaunchKernelGGL(impr_kernel<true>, grid, dim3(kImprThreads), lds, st, a);
    else
        hipLaunchKernelGGL(impr_kernel<false>, grid, dim3(kImprThreads), lds, st, a);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}

// ---------------------------------------------------------------------------------------------------- reverse mode
// u = A(kappaRX, wFld)^-1 y per column. Given ubar: lam = A^-T ubar by the same guarded elimination on the
// transposed diagonals (row m of A^T = (e(m-2), d(m-1), c(m), b(m+1), a(m+2))), ybar = lam, and the cotangents of
// kappaRX and wFld from the derivatives of the rows in those parameters (diffusion: b(k), d(k-1), c in kappaRX(k);
// U3: rows K-1 and K in the rCenter of interface K). One thread per column; rows are built two ahead of the
// transposed elimination, from level data read as needed (the reverse mode is not tuned).
namespace mj {
namespace {

struct Row {
    double a, b, c, d, e;
};

template <bool ADV>
__device__ Row build_row(const ImprArgs& a, const double* lv, int m, bool in, double rA, double rrA, const double* mCp,
                         const double* rhp, const double* kp, const double* wp, Iface& up) {
    const int nr = a.nr;
    Row r{0.0, 0.0, 1.0, 0.0, 0.0};
    if (!in) return r;
    const int64_t P = a.plane;
    const double dtm = lv[m], rdf = lv[nr + m], rh0 = rhp[m * P];
    const double* rdrC = lv + 2 * nr;
    const double mC0 = mCp[m * P], mC1 = m >= 1 ? mCp[(m - 1) * P] : 0.0;
    const double mCn = m + 1 < nr ? mCp[(m + 1) * P] : 0.0;
    if (m >= 1) r.b = -(((((dtm * mC1) * rh0) * rdf) * kp[m * P]) * rdrC[m]);
    if (m + 1 < nr) r.d = -(((((dtm * mCn) * rh0) * rdf) * kp[(m + 1) * P]) * rdrC[m + 1]);
    r.c = 1.0 - (r.b + r.d);
    if (ADV) {
        Iface lo{0.0, 0.0, 0.0, 0.0};
        if (m + 1 < nr) {
            const double mCk = m + 2 < nr ? mCp[(m + 2) * P] : mCn;
            lo = iface(a, m + 1, wp[(m + 1) * P], rA, rrA, mC0, m >= 1 ? mC1 : mC0, mCk);
            r.b = r.b - ((lo.km * dtm) * rh0) * rdf;
            r.c = r.c + (((((lo.rC + lo.rUp) + lo.km) * dtm) * rh0) * rdf);
            r.d = r.d + (((((lo.rC - lo.rUp) + lo.kp) * dtm) * rh0) * rdf);
            r.e = r.e - ((lo.kp * dtm) * rh0) * rdf;
        }
        if (m >= 1) {
            r.a = r.a + ((up.km * dtm) * rh0) * rdf;
            r.b = r.b - (((((up.rC + up.rUp) + up.km) * dtm) * rh0) * rdf);
            r.c = r.c - (((((up.rC - up.rUp) + up.kp) * dtm) * rh0) * rdf);
            r.d = r.d + ((up.kp * dtm) * rh0) * rdf;
        }
        up = lo;
    }
    return r;
}

template <bool ADV>
__global__ void __launch_bounds__(kImprThreads)
    impr_adjoint_kernel(ImprArgs a, const double* __restrict__ u, const double* __restrict__ ubar, double* ybar,
                        double* kbar, double* wbar) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int64_t col = (int64_t)blockIdx.x * kImprThreads + lane;
    const int mem = blockIdx.y;
    double* lv = lds + (int64_t)2 * a.nr * kImprThreads;
    for (int k = lane; k < a.nr; k += kImprThreads) {
        lv[k] = a.dt[k];
        lv[a.nr + k] = a.rdrF[k];
        lv[2 * a.nr + k] = a.rdrC[k];
    }
    __syncthreads();
    if (col >= (int64_t)a.T * a.plane) return;  // no barriers below
    const int t = (int)(col / a.plane), q = (int)(col - (int64_t)t * a.plane);
    const int j = q / a.nx, i = q - j * a.nx;
    const bool in = j >= a.OL && j < a.OL + a.sNy && i >= a.OL && i < a.OL + a.sNx;
    const int nr = a.nr;
    const int64_t P = a.plane;
    const int64_t o3 = mem * a.ms3 + (int64_t)t * nr * P + q;
    const int64_t om = mem * a.mms + (int64_t)t * nr * P + q;
    const int64_t o2 = (int64_t)t * P + q;
    const double rA = ADV && in ? a.rA[o2] : 0.0;
    const double rrA = ADV && in ? a.recip_rA[o2] : 0.0;
    const double *mCp = a.maskC + om, *rhp = a.rh + o3, *kp = a.kap + o3, *wp = a.w + o3;
    const double *uc = u + o3, *ub = ubar + o3;
    double* yb = ybar + o3;
    double* dpL = lds + lane;
    double* epL = lds + (int64_t)nr * kImprThreads + lane;
    // transposed elimination, rows of A built two ahead
    Iface up{0.0, 0.0, 0.0, 0.0};
    const Row zero{0.0, 0.0, 0.0, 0.0, 0.0};
    Row rM2 = zero, rM1 = zero, r0 = zero, rP1 = zero, rP2 = zero;
    double dp1 = 0.0, ep1 = 0.0, yp1 = 0.0, dp2 = 0.0, ep2 = 0.0, yp2 = 0.0;
    for (int r = 0; r < nr + 2; ++r) {
        rM2 = rM1, rM1 = r0, r0 = rP1, rP1 = rP2;
        rP2 = r < nr ? build_row<ADV>(a, lv, r, in, rA, rrA, mCp, rhp, kp, wp, up) : zero;
        const int m = r - 2;
        if (m < 0) continue;
        const double ra = rM2.e, rb = rM1.d, rc = r0.c, rd = rP1.b, re = rP2.a;  // row m of A^T
        double dq, eq = re, yq;
        const double y0 = ub[m * P];
        if (m == 0) {
            dq = rd;
            yq = y0;
            norm3(rc, dq, eq, yq);
        } else if (m == 1) {
            const double c2 = rc - rb * dp1;
            dq = rd - rb * ep1;
            yq = y0 - rb * yp1;
            norm3(c2, dq, eq, yq);
        } else {
            const double bb = rb - ra * dp2;
            const double cq = (rc - ra * ep2) - bb * dp1;
            dq = rd - bb * ep1;
            yq = (y0 - ra * yp2) - bb * yp1;
            norm3(cq, dq, eq, yq);
        }
        dpL[(int64_t)m * kImprThreads] = dq;
        epL[(int64_t)m * kImprThreads] = eq;
        yb[m * P] = yq;
        dp2 = dp1, ep2 = ep1, yp2 = yp1;
        dp1 = dq, ep1 = eq, yp1 = yq;
    }
    if (nr >= 2) {
        double u1 = yp1;
        double u0 = yp2 - u1 * dpL[(int64_t)(nr - 2) * kImprThreads];
        yb[(int64_t)(nr - 2) * P] = u0;
        double u2 = u1;
        u1 = u0;
        for (int m = nr - 3; m >= 0; --m) {
            const double v = (yb[m * P] - u1 * dpL[(int64_t)m * kImprThreads]) - u2 * epL[(int64_t)m * kImprThreads];
            yb[m * P] = v;
            u2 = u1;
            u1 = v;
        }
    }
    // parameter cotangents (interior columns; level 0 and the surface interface have none)
    double* kb = kbar + o3;
    double* wbp = ADV ? wbar + o3 : nullptr;
    kb[0] = 0.0;
    if (ADV) wbp[0] = 0.0;
    const double* rdrC = lv + 2 * nr;
    for (int k = 1; k < nr; ++k) {
        double gk = 0.0, gw = 0.0;
        if (in) {
            const double lk = yb[k * P], lk1 = yb[(k - 1) * P];
            const double uk = uc[k * P], uk1 = uc[(k - 1) * P];
            const double mCk = mCp[k * P], mCk1 = mCp[(k - 1) * P];
            const double rhk = rhp[k * P], rhk1 = rhp[(k - 1) * P];
            const double beta = -((((lv[k] * mCk1) * rhk) * lv[nr + k]) * rdrC[k]);
            const double delta = -((((lv[k - 1] * mCk) * rhk1) * lv[nr + k - 1]) * rdrC[k]);
            const double du = uk - uk1;
            gk = -((lk * beta) * (-du) + (lk1 * delta) * du);
            if (ADV) {
                const double dw = (((0.5 * rA) * mCk1) * rrA) * a.rkSign;
                const double rC = ((0.5 * ((wp[k * P] * rA) * mCk1)) * rrA) * a.rkSign;
                const double sg = (double)((rC > 0.0) - (rC < 0.0));
                const double mskM = mCp[(k >= 2 ? k - 2 : 0) * P] * (k <= 1 ? 0.0 : 1.0);
                const double mskP = mCp[(k + 1 < nr ? k + 1 : nr - 1) * P] * (k >= nr - 1 ? 0.0 : 1.0);
                const double drUp = (2.0 * kSixth) * sg;
                const double dkm = (kSixth * (1.0 + sg)) * mskM;
                const double dkp = (kSixth * (1.0 - sg)) * mskP;
                const double uKm2 = k >= 2 ? uc[(k - 2) * P] : 0.0, uKp1 = k + 1 < nr ? uc[(k + 1) * P] : 0.0;
                const double phi = (((-dkm) * uKm2 + ((1.0 + drUp) + dkm) * uk1) + ((1.0 - drUp) + dkp) * uk) - dkp * uKp1;
                const double s1 = (lv[k - 1] * rhk1) * lv[nr + k - 1], s0 = (lv[k] * rhk) * lv[nr + k];
                const double rCb = -((lk1 * s1) - (lk * s0)) * phi;
                gw = rCb * dw;
            }
        }
        kb[k * P] = gk;
        if (ADV) wbp[k * P] = gw;
    }
}

}  // namespace
}  // namespace mj

extern "C" int mj_gad_implicit_r_adjoint(const mj_impr_grid* g, const mj_impr_config* c, const double* kappaRX,
                                         const double* recip_hFac, const double* wFld, const double* u,
                                         const double* ubar, double* ybar, double* kappaRX_bar, double* wFld_bar,
                                         void* stream) {
    MJ_REQUIRE(g && c, "mj_gad_implicit_r_adjoint: null grid/config");
    MJ_REQUIRE(c->batch >= 1 && c->batch <= 65535 && c->ntiles >= 1 && c->nr >= 2 && c->nr <= 150,
               "mj_gad_implicit_r_adjoint: batch in 1..65535, ntiles >= 1, nr in 2..150");
    MJ_REQUIRE(c->sNx >= 1 && c->sNy >= 1 && c->OL >= 0, "mj_gad_implicit_r_adjoint: bad tile layout");
    const int nx = c->sNx + 2 * c->OL, ny = c->sNy + 2 * c->OL;
    const int64_t plane = (int64_t)nx * ny, member = (int64_t)c->ntiles * c->nr * plane;
    MJ_REQUIRE(c->batch == 1 || (c->member_stride3d >= member && (c->mask_member_stride == 0 ||
                                                                  c->mask_member_stride >= member)),
               "mj_gad_implicit_r_adjoint: member strides smaller than one member's arrays");
    MJ_REQUIRE(kappaRX && recip_hFac && u && ubar && ybar && kappaRX_bar && g->maskC && g->dTtracerLev &&
                   g->recip_drF && g->recip_drC,
               "mj_gad_implicit_r_adjoint: null array or grid field");
    MJ_REQUIRE(!c->implicitAdvection || (wFld && wFld_bar && g->rA && g->recip_rA),
               "mj_gad_implicit_r_adjoint: implicit advection needs wFld, wFld_bar, rA, recip_rA");
    const double* ins[] = {kappaRX, recip_hFac, wFld, u, ubar, g->maskC};
    for (const double* p : ins)
        MJ_REQUIRE(p != ybar && p != kappaRX_bar && (!wFld_bar || p != wFld_bar),
                   "mj_gad_implicit_r_adjoint: an output aliases an input");
    ImprArgs a;
    a.T = c->ntiles;
    a.nx = nx;
    a.ny = ny;
    a.sNx = c->sNx;
    a.sNy = c->sNy;
    a.OL = c->OL;
    a.nr = c->nr;
    a.adv = c->implicitAdvection ? 1 : 0;
    a.plane = plane;
    a.ms3 = c->member_stride3d;
    a.mms = c->mask_member_stride;
    a.maskC = g->maskC;
    a.rA = g->rA;
    a.recip_rA = g->recip_rA;
    a.dt = g->dTtracerLev;
    a.rdrF = g->recip_drF;
    a.rdrC = g->recip_drC;
    a.rkSign = g->rkSign;
    a.kap = kappaRX;
    a.rh = recip_hFac;
    a.w = c->implicitAdvection ? wFld : kappaRX;
    a.y = nullptr;
    a.out = nullptr;
    hipStream_t st = as_stream(stream);
    const size_t lds = ((size_t)2 * c->nr * kImprThreads + 3 * (size_t)c->nr) * sizeof(double);
    const void* kern = a.adv ? (const void*)impr_adjoint_kernel<true> : (const void*)impr_adjoint_kernel<false>;
    if (lds > 65536) MJ_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int64_t cols = (int64_t)c->ntiles * plane;
    const dim3 grid((unsigned)((cols + kImprThreads - 1) / kImprThreads), (unsigned)c->batch);
    if (a.adv)
        hipLaunchKernelGGL(impr_adjoint_kernel<true>, grid, dim3(kImprThreads), lds, st, a, u, ubar, ybar,
                           kappaRX_bar, wFld_bar);
    else
        hipLaunchKernelGGL(impr_adjoint_kernel<false>, grid, dim3(kImprThreads), lds, st, a, u, ubar, ybar,
                           kappaRX_bar, (double*)nullptr);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}
