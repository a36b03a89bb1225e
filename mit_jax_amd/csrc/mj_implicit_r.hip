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
// columns of a wave are consecutive points of a row (coalesced).
#include "mj_common.h"

namespace mj {
namespace {

constexpr int kImprThreads = 64;  // one wave per workgroup (LDS: 2 * nr * 64 doubles)
#ifndef MJ_IMPR_PF
#define MJ_IMPR_PF 8
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

__global__ void __launch_bounds__(kImprThreads) impr_kernel(ImprArgs a) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int64_t col = (int64_t)blockIdx.x * kImprThreads + lane;
    const int mem = blockIdx.y;
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
    const double rA = in && a.adv ? a.rA[o2] : 0.0;
    const double rrA = in && a.adv ? a.recip_rA[o2] : 0.0;

    double rmC[kPf], rrh[kPf], rkap[kPf], rw[kPf], ry[kPf];  // slot s: level m with m % kPf == s
    auto load = [&](int s, int lev) {
        const int64_t ok = (int64_t)lev * a.plane;
        ry[s] = y[ok];
        if (in) {
            rmC[s] = a.maskC[om + ok];
            rrh[s] = a.rh[o3 + ok];
            rkap[s] = a.kap[o3 + ok];
            rw[s] = a.adv ? a.w[o3 + ok] : 0.0;
        }
    };
#pragma unroll
    for (int s = 0; s < kPf; ++s) {
        rmC[s] = rrh[s] = rkap[s] = rw[s] = ry[s] = 0.0;
        if (s < nr) load(s, s);
    }
    double mC1 = 0.0;                // maskC of level m-1
    Iface up{0.0, 0.0, 0.0, 0.0};   // interface m (above level m), built by row m-1
    double dp1 = 0.0, ep1 = 0.0, yp1 = 0.0, dp2 = 0.0, ep2 = 0.0, yp2 = 0.0;  // rows m-1, m-2 after elimination
    for (int m0 = 0; m0 < nr; m0 += kPf) {
#pragma unroll
        for (int s = 0; s < kPf; ++s) {
            const int m = m0 + s;
            if (m < nr) {
                const int s1 = (s + 1) % kPf, s2 = (s + 2) % kPf;
                // row m of the matrix (identity on halo columns)
                double ra = 0.0, rb = 0.0, rc = 1.0, rd = 0.0, re = 0.0;
                const double mC0 = rmC[s];
                if (in) {
                    const double dtm = a.dt[m], rdf = a.rdrF[m], rhm = rrh[s];
                    const double mCn = m + 1 < nr ? rmC[s1] : 0.0;
                    // diffusion, reference tracer_implicit.py:54-64
                    if (m >= 1) rb = -(((((dtm * mC1) * rhm) * rdf) * rkap[s]) * a.rdrC[m]);
                    if (m + 1 < nr) rd = -(((((dtm * mCn) * rhm) * rdf) * rkap[s1]) * a.rdrC[m + 1]);
                    rc = 1.0 - (rb + rd);
                    if (a.adv) {
                        // U3 advection, :95-105: level m takes the updates of interface m+1 (its bottom) first,
                        // then those of interface m (its top)
                        Iface lo{0.0, 0.0, 0.0, 0.0};
                        if (m + 1 < nr) {
                            const double mCk = m + 2 < nr ? rmC[s2] : mCn;
                            lo = iface(a, m + 1, rw[s1], rA, rrA, mC0, m >= 1 ? mC1 : mC0, mCk);
                            rb = rb - ((lo.km * dtm) * rhm) * rdf;
                            rc = rc + (((((lo.rC + lo.rUp) + lo.km) * dtm) * rhm) * rdf);
                            rd = rd + (((((lo.rC - lo.rUp) + lo.kp) * dtm) * rhm) * rdf);
                            re = re - ((lo.kp * dtm) * rhm) * rdf;
                        }
                        if (m >= 1) {
                            ra = ra + ((up.km * dtm) * rhm) * rdf;
                            rb = rb - (((((up.rC + up.rUp) + up.km) * dtm) * rhm) * rdf);
                            rc = rc - (((((up.rC - up.rUp) + up.kp) * dtm) * rhm) * rdf);
                            rd = rd + ((up.kp * dtm) * rhm) * rdf;
                        }
                        up = lo;
                    }
                }
                // forward elimination of row m, reference tracer_implicit.py:126-143
                double dq, eq = re, yq;
                if (m == 0) {
                    dq = rd;
                    yq = ry[s];
                    norm3(rc, dq, eq, yq);
                } else if (m == 1) {
                    const double c2 = rc - rb * dp1;
                    dq = rd - rb * ep1;
                    yq = ry[s] - rb * yp1;
                    norm3(c2, dq, eq, yq);
                } else {
                    const double bb = rb - ra * dp2;
                    const double cq = (rc - ra * ep2) - bb * dp1;
                    dq = rd - bb * ep1;
                    yq = (ry[s] - ra * yp2) - bb * yp1;
                    norm3(cq, dq, eq, yq);
                }
                dpL[(int64_t)m * kImprThreads] = dq;
                epL[(int64_t)m * kImprThreads] = eq;
                out[(int64_t)m * a.plane] = yq;
                dp2 = dp1, ep2 = ep1, yp2 = yp1;
                dp1 = dq, ep1 = eq, yp1 = yq;
                mC1 = mC0;
                if (m + kPf < nr) load(s, m + kPf);
            }
        }
    }
    // back substitution, reference tracer_implicit.py:150-160: u(nr-1) = y'(nr-1) is in place already
    if (nr < 2) return;
    double u1 = yp1;                          // u(m+1)
    double u0 = yp2 - u1 * dpL[(int64_t)(nr - 2) * kImprThreads];
    out[(int64_t)(nr - 2) * a.plane] = u0;
    double u2 = u1;
    u1 = u0;
    // y'(m) for m = nr-3 .. 0 are re-read from the output column, kPf levels ahead
    double ryp[kPf];
#pragma unroll
    for (int s = 0; s < kPf; ++s) ryp[s] = nr - 3 - s >= 0 ? out[(int64_t)(nr - 3 - s) * a.plane] : 0.0;
    for (int m0 = nr - 3; m0 >= 0; m0 -= kPf) {
#pragma unroll
        for (int s = 0; s < kPf; ++s) {
            const int m = m0 - s;
            if (m >= 0) {
                const double u = (ryp[s] - u1 * dpL[(int64_t)m * kImprThreads]) - u2 * epL[(int64_t)m * kImprThreads];
                out[(int64_t)m * a.plane] = u;
                u2 = u1;
                u1 = u;
                if (m - kPf >= 0) ryp[s] = out[(int64_t)(m - kPf) * a.plane];
            }
        }
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
    a.w = wFld;
    a.y = gTracer;
    a.out = out;
    const size_t lds = (size_t)2 * c->nr * kImprThreads * sizeof(double);
    if (lds > 65536)
        MJ_HIP(hipFuncSetAttribute((const void*)impr_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const dim3 grid((unsigned)((cols + kImprThreads - 1) / kImprThreads), (unsigned)c->batch);
    hipLaunchKernelGGL(impr_kernel, grid, dim3(kImprThreads), lds, st, a);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}
