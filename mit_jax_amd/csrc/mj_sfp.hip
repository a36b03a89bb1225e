// The CG2D operator and right-hand side on gfx950: UPDATE_CG2D and cg2d_rhs of SOLVE_FOR_PRESSURE
// (reference core/solve_for_pressure.py:29-106).
//
// Both are column sums over the levels of 3-D [T, nr, ny, nx] FP64 fields into 2-D [T, ny, nx] results: HBM-bound
// streams with a few flops per byte. One thread owns one (j, i) column of one member and walks the levels in the
// reference's order (UPDATE_CG2D k = 1..Nr, the divergence of cg2d_b k = Nr..1), so every sum is the literal
// sequential one and the results are bitwise those of the reference's loops (-ffp-contract=off). The level-sum
// kernels give each thread two horizontally adjacent points (flat over the padded plane), so every level is read
// with contiguous 16-byte loads across the lanes; the j+1 row the divergence reads is the neighbouring threads'
// (L1/L2 hits). The 2-D kernels (aC2d, the preconditioner) cover kRows rows x 128 columns per block.
//
// UPDATE_CG2D runs as: sums -> aW2d, aS2d (whole arrays); aC2d on the interior (whole array); the T exchange of
// aC2d; the preconditioner on j, i = 1..sNy+1, 1..sNx+1 (it reads aC2d at i-1 and j-1, i.e. exchanged halo points).
#include "mj_common.h"

namespace mj {

namespace {

constexpr int kCols = 128;        // threads across a row (nx <= 128), 2-D kernels
constexpr int kRows = 8;          // rows per block, 2-D kernels
constexpr int kPairThreads = 256; // level-sum kernels: one thread per pair of adjacent points
#ifndef MJ_SFP_LV
#define MJ_SFP_LV 5
#endif
#ifndef MJ_SFP_RHS_LV
#define MJ_SFP_RHS_LV 2
#endif
constexpr int kLv = MJ_SFP_LV;         // levels per load batch (UPDATE_CG2D sums)
constexpr int kRhsLv = MJ_SFP_RHS_LV;  // levels per load batch (cg2d_rhs)

struct SfpGeo {
    int B, T, nx, ny, sNx, sNy, OL, nr;
    int64_t plane, ms3, ms2;  // ny*nx; member strides
};

__device__ __forceinline__ bool in_range(int v, int lo, int hi) { return v >= lo && v <= hi; }

__device__ __forceinline__ double2 ld2(const double* p) { return *reinterpret_cast<const double2*>(p); }
__device__ __forceinline__ void st2(double* p, double2 v) { *reinterpret_cast<double2*>(p) = v; }

// aW2d, aS2d: ((sum_k ((dyG * drF(k)) * hFacW(k)) * recip_dxC) * cg2dNorm) * implicSurfPress * implicDiv2DFlow on
// Fortran j, i = 1..sNy+1, 1..sNx+1, 0 elsewhere (reference solve_for_pressure.py:38-47 with _ksum :23-26).
// One thread per pair of horizontally adjacent points (flat over the padded plane, nx even: a pair never straddles
// a row), 16-byte loads of each level.
__global__ void __launch_bounds__(kPairThreads)
    sfp_opsum_kernel(SfpGeo g, const double* __restrict__ dyG, const double* __restrict__ dxG,
                     const double* __restrict__ recip_dxC, const double* __restrict__ recip_dyC,
                     const double* __restrict__ drF, const double* __restrict__ hFacW,
                     const double* __restrict__ hFacS, const double* __restrict__ norm, double isp, double idf,
                     double* __restrict__ aW2d, double* __restrict__ aS2d) {
    const int p = blockIdx.x * kPairThreads + threadIdx.x, t = blockIdx.y, m = blockIdx.z;
    const int q0 = 2 * p;
    if (q0 >= g.plane) return;
    const int j = q0 / g.nx, i0 = q0 - j * g.nx;
    const int64_t q = (int64_t)t * g.plane + q0;
    const int64_t o2 = m * g.ms2 + q;
    const bool jin = in_range(j, g.OL, g.OL + g.sNy);
    const bool in0 = jin && in_range(i0, g.OL, g.OL + g.sNx), in1 = jin && in_range(i0 + 1, g.OL, g.OL + g.sNx);
    double w0 = 0.0, w1 = 0.0, s0 = 0.0, s1 = 0.0;
    if (in0 || in1) {
        const double2 cW = ld2(dyG + q), cS = ld2(dxG + q), rW = ld2(recip_dxC + q), rS = ld2(recip_dyC + q);
        const int64_t o3 = m * g.ms3 + (int64_t)t * g.nr * g.plane + q0;
        const double *hW = hFacW + o3, *hS = hFacS + o3;
        // levels in chunks of kLv: the chunk's loads are all issued before its sums (the sums stay in k order)
        int k = 0;
        for (; k + kLv <= g.nr; k += kLv) {
            double2 hw[kLv], hs[kLv];
#pragma unroll
            for (int c = 0; c < kLv; ++c) {
                hw[c] = ld2(hW + (int64_t)(k + c) * g.plane);
                hs[c] = ld2(hS + (int64_t)(k + c) * g.plane);
            }
#pragma unroll
            for (int c = 0; c < kLv; ++c) {
                const double dz = drF[k + c];
                w0 = w0 + ((cW.x * dz) * hw[c].x) * rW.x;
                w1 = w1 + ((cW.y * dz) * hw[c].y) * rW.y;
                s0 = s0 + ((cS.x * dz) * hs[c].x) * rS.x;
                s1 = s1 + ((cS.y * dz) * hs[c].y) * rS.y;
            }
        }
        for (; k < g.nr; ++k) {
            const double dz = drF[k];
            const double2 hw = ld2(hW + (int64_t)k * g.plane), hs = ld2(hS + (int64_t)k * g.plane);
            w0 = w0 + ((cW.x * dz) * hw.x) * rW.x;
            w1 = w1 + ((cW.y * dz) * hw.y) * rW.y;
            s0 = s0 + ((cS.x * dz) * hs.x) * rS.x;
            s1 = s1 + ((cS.y * dz) * hs.y) * rS.y;
        }
        const double nm = norm[m];
        w0 = in0 ? ((w0 * nm) * isp) * idf : 0.0;
        s0 = in0 ? ((s0 * nm) * isp) * idf : 0.0;
        w1 = in1 ? ((w1 * nm) * isp) * idf : 0.0;
        s1 = in1 ? ((s1 * nm) * isp) * idf : 0.0;
    }
    st2(aW2d + o2, make_double2(w0, w1));
    st2(aS2d + o2, make_double2(s0, s1));
}

// aC2d = -(aW(i) + aW(i+1) + aS(j) + aS(j+1) + freeSurfFac cg2dNorm recip_Bo rA / deltaTMom / deltaTFreeSurf) on
// the interior, 0 elsewhere (reference solve_for_pressure.py:48-52).
__global__ void __launch_bounds__(kCols * kRows)
    sfp_aC_kernel(SfpGeo g, const double* __restrict__ aW2d, const double* __restrict__ aS2d,
                  const double* __restrict__ recip_Bo, const double* __restrict__ rA, const double* __restrict__ norm,
                  double fsf, double dtm, double dtfs, double* __restrict__ aC2d) {
    const int i = threadIdx.x, j = blockIdx.x * kRows + threadIdx.y, t = blockIdx.y, m = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int64_t q = (int64_t)t * g.plane + (int64_t)j * g.nx + i;
    const int64_t o2 = m * g.ms2 + q;
    double c = 0.0;
    if (in_range(j, g.OL, g.OL + g.sNy - 1) && in_range(i, g.OL, g.OL + g.sNx - 1)) {
        const double d = ((((fsf * norm[m]) * recip_Bo[q]) * rA[q]) / dtm) / dtfs;
        c = -((((aW2d[o2] + aW2d[o2 + 1]) + aS2d[o2]) + aS2d[o2 + g.nx]) + d);
    }
    aC2d[o2] = c;
}

// The preconditioner on Fortran j, i = 1..sNy+1, 1..sNx+1 from the exchanged aC2d (reference
// solve_for_pressure.py:54-65); other points untouched.
__global__ void __launch_bounds__(kCols * kRows)
    sfp_precond_kernel(SfpGeo g, const double* __restrict__ aW2d, const double* __restrict__ aS2d,
                       const double* __restrict__ aC2d, double off, double* __restrict__ pW, double* __restrict__ pS,
                       double* __restrict__ pC) {
    const int i = threadIdx.x, j = blockIdx.x * kRows + threadIdx.y, t = blockIdx.y, m = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    if (!in_range(j, g.OL, g.OL + g.sNy) || !in_range(i, g.OL, g.OL + g.sNx)) return;
    const int64_t o2 = m * g.ms2 + (int64_t)t * g.plane + (int64_t)j * g.nx + i;
    const double a = aC2d[o2];
    pC[o2] = a == 0.0 ? 1.0 : 1.0 / a;
    const double tW = a + aC2d[o2 - 1];
    const double fW = off * tW;
    pW[o2] = tW == 0.0 ? 0.0 : -aW2d[o2] / (fW * fW);
    const double tS = a + aC2d[o2 - g.nx];
    const double fS = off * tS;
    pS[o2] = tS == 0.0 ? 0.0 : -aS2d[o2] / (fS * fS);
}

struct RhsArgs {
    const double *dyG, *dxG, *rA, *maskInC, *drF, *dfc, *rfc, *d2f;
    const int32_t* kSurfC;
    const double *gU, *gV, *hFacW, *hFacS, *EmPmR, *etaN, *etaH, *Bo_surf;
    double *b, *x;
    double tmpFac, fsf, dtm, dtfs;
    int rfw;
};

// pfu = (((((dyG deepFacC(k)) drF(k)) hFacW) rhoFacC(k)) gU) / deltaTMom (reference solve_for_pressure.py:90-93);
// the factors of a NULL vertical array are 1 and skipped (x * 1 == x).
__device__ __forceinline__ double pf(double c, const double* dfc, const double* rfc, double dz, int k, double h,
                                     double v, double dtm) {
    double xa = c;
    if (dfc) xa = xa * dfc[k];
    xa = (xa * dz) * h;
    if (rfc) xa = xa * rfc[k];
    return (xa * v) / dtm;
}

__device__ __forceinline__ double eta_term(const RhsArgs& a, const SfpGeo& g, int64_t q, double rA, double etaH) {
    double e = a.fsf * rA;
    if (a.d2f) {
        int ks = a.kSurfC[q];
        ks = ks < 1 ? 1 : (ks > g.nr + 1 ? g.nr + 1 : ks);
        e = e * a.d2f[ks - 1];
    }
    return ((e / a.dtm) / a.dtfs) * etaH;
}

// cg2d_x = Bo_surf etaN on the whole array; cg2d_b on the interior (0 elsewhere): the fresh-water term, the
// divergence of the pf fluxes for k = nr..1 in order, the exactConserv etaH term (reference
// solve_for_pressure.py:68-106). One thread per pair of horizontally adjacent points (i0, i0+1) of row j: per level
// it reads hFacW, gU at i0..i0+2 and hFacS, gV at rows j and j+1 (16-byte loads but for i0+2), and the flux at
// i0+1 serves both points.
__global__ void __launch_bounds__(kPairThreads) sfp_rhs_kernel(SfpGeo g, RhsArgs a) {
    const int p = blockIdx.x * kPairThreads + threadIdx.x, t = blockIdx.y, m = blockIdx.z;
    const int q0 = 2 * p;
    if (q0 >= g.plane) return;
    const int j = q0 / g.nx, i0 = q0 - j * g.nx;
    const int64_t q = (int64_t)t * g.plane + q0;
    const int64_t o2 = m * g.ms2 + q;
    {
        const double2 bo = ld2(a.Bo_surf + q), en = ld2(a.etaN + o2);
        st2(a.x + o2, make_double2(bo.x * en.x, bo.y * en.y));
    }
    const bool jin = in_range(j, g.OL, g.OL + g.sNy - 1);
    const bool in0 = jin && in_range(i0, g.OL, g.OL + g.sNx - 1);
    const bool in1 = jin && in_range(i0 + 1, g.OL, g.OL + g.sNx - 1);
    double b0 = 0.0, b1 = 0.0;
    if (in0 || in1) {
        const double2 rA = ld2(a.rA + q);
        if (a.rfw) {
            const double2 em = ld2(a.EmPmR + o2), mk = ld2(a.maskInC + q);
            b0 = (((a.tmpFac * rA.x) * em.x) / a.dtm) * mk.x;
            b1 = (((a.tmpFac * rA.y) * em.y) / a.dtm) * mk.y;
        }
        const double2 cu = ld2(a.dyG + q), cv0 = ld2(a.dxG + q), cv1 = ld2(a.dxG + q + g.nx);
        const double cu2 = a.dyG[q + 2];
        const int64_t o3 = m * g.ms3 + (int64_t)t * g.nr * g.plane + q0;
        const double *hW = a.hFacW + o3, *hS = a.hFacS + o3, *gU = a.gU + o3, *gV = a.gV + o3;
        // levels k = nr-1 .. 0 in chunks of kRhsLv: the chunk's loads are issued before its terms
        for (int k = g.nr - 1; k >= 0; k -= kRhsLv) {
            const int nc = k + 1 < kRhsLv ? k + 1 : kRhsLv;
            double2 h01[kRhsLv], u01[kRhsLv], s0[kRhsLv], s1[kRhsLv], v0[kRhsLv], v1[kRhsLv];
            double h2[kRhsLv], u2[kRhsLv];
#pragma unroll
            for (int c = 0; c < kRhsLv; ++c) {
                if (c < nc) {
                    const int64_t ok = (int64_t)(k - c) * g.plane;
                    h01[c] = ld2(hW + ok);
                    h2[c] = hW[ok + 2];
                    u01[c] = ld2(gU + ok);
                    u2[c] = gU[ok + 2];
                    s0[c] = ld2(hS + ok);
                    s1[c] = ld2(hS + ok + g.nx);
                    v0[c] = ld2(gV + ok);
                    v1[c] = ld2(gV + ok + g.nx);
                }
            }
#pragma unroll
            for (int c = 0; c < kRhsLv; ++c) {
                if (c < nc) {
                    const int kk = k - c;
                    const double dz = a.drF[kk];
                    const double pu0 = pf(cu.x, a.dfc, a.rfc, dz, kk, h01[c].x, u01[c].x, a.dtm);
                    const double pu1 = pf(cu.y, a.dfc, a.rfc, dz, kk, h01[c].y, u01[c].y, a.dtm);
                    const double pu2 = pf(cu2, a.dfc, a.rfc, dz, kk, h2[c], u2[c], a.dtm);
                    const double pv00 = pf(cv0.x, a.dfc, a.rfc, dz, kk, s0[c].x, v0[c].x, a.dtm);
                    const double pv10 = pf(cv1.x, a.dfc, a.rfc, dz, kk, s1[c].x, v1[c].x, a.dtm);
                    const double pv01 = pf(cv0.y, a.dfc, a.rfc, dz, kk, s0[c].y, v0[c].y, a.dtm);
                    const double pv11 = pf(cv1.y, a.dfc, a.rfc, dz, kk, s1[c].y, v1[c].y, a.dtm);
                    b0 = (b0 + pu1) - pu0;
                    b0 = (b0 + pv10) - pv00;
                    b1 = (b1 + pu2) - pu1;
                    b1 = (b1 + pv11) - pv01;
                }
            }
        }
        const double2 eH = ld2(a.etaH + o2);
        b0 = in0 ? b0 - eta_term(a, g, q, rA.x, eH.x) : 0.0;
        b1 = in1 ? b1 - eta_term(a, g, q + 1, rA.y, eH.y) : 0.0;
    }
    st2(a.b + o2, make_double2(b0, b1));
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

struct RhsAdjArgs {
    const double *dyG, *dxG, *rA, *maskInC, *drF, *d2f;
    const int32_t* kSurfC;
    const double *gU, *gV, *hFacW, *hFacS, *Bo_surf, *b_bar, *x_bar;
    double *gU_bar, *gV_bar, *hW_bar, *hS_bar, *E_bar, *eN_bar, *eH_bar;
    double tmpFac, fsf, dtm, dtfs;
    int rfw;
};

// Reverse mode of sfp_rhs_kernel (V4r4 vertical factors): pfu(j, i) enters b(j, i-1) with + and b(j, i) with -
// (pfv likewise in j), so the face cotangent is b_bar(left) - b_bar(right) with b_bar taken as 0 off the interior;
// then gU_bar = (pu ((dyG drF) hFacW)) / deltaTMom, hFacW_bar = (pu ((dyG drF) gU)) / deltaTMom (v alike); the
// 2-D cotangents of EmPmR, etaH (interior) and etaN (everywhere). One thread per pair of adjacent points.
__global__ void __launch_bounds__(kPairThreads) sfp_rhs_adjoint_kernel(SfpGeo g, RhsAdjArgs a) {
    const int p = blockIdx.x * kPairThreads + threadIdx.x, t = blockIdx.y, m = blockIdx.z;
    const int q0 = 2 * p;
    if (q0 >= g.plane) return;
    const int j = q0 / g.nx, i0 = q0 - j * g.nx;
    const int64_t q = (int64_t)t * g.plane + q0;
    const int64_t o2 = m * g.ms2 + q;
    auto inside = [&](int jj, int ii) {
        return jj >= g.OL && jj < g.OL + g.sNy && ii >= g.OL && ii < g.OL + g.sNx;
    };
    const double* bb = a.b_bar + m * g.ms2 + (int64_t)t * g.plane;
    auto bval = [&](int jj, int ii) { return inside(jj, ii) ? bb[jj * g.nx + ii] : 0.0; };
    const double b00 = bval(j, i0), b01 = bval(j, i0 + 1);
    const double pu0 = bval(j, i0 - 1) - b00, pu1 = b00 - b01;
    const double pv0 = bval(j - 1, i0) - b00, pv1 = bval(j - 1, i0 + 1) - b01;
    // 2-D cotangents
    {
        const double2 rA = ld2(a.rA + q);
        double e0 = 0.0, e1 = 0.0, h0 = 0.0, h1 = 0.0;
        const bool in0 = inside(j, i0), in1 = inside(j, i0 + 1);
        if (a.rfw) {
            const double2 mk = ld2(a.maskInC + q);
            if (in0) e0 = b00 * (((a.tmpFac * rA.x) / a.dtm) * mk.x);
            if (in1) e1 = b01 * (((a.tmpFac * rA.y) / a.dtm) * mk.y);
        }
        double f0 = a.fsf * rA.x, f1 = a.fsf * rA.y;
        if (a.d2f) {
            int k0 = a.kSurfC[q], k1 = a.kSurfC[q + 1];
            k0 = k0 < 1 ? 1 : (k0 > g.nr + 1 ? g.nr + 1 : k0);
            k1 = k1 < 1 ? 1 : (k1 > g.nr + 1 ? g.nr + 1 : k1);
            f0 = f0 * a.d2f[k0 - 1];
            f1 = f1 * a.d2f[k1 - 1];
        }
        if (in0) h0 = -(b00 * ((f0 / a.dtm) / a.dtfs));
        if (in1) h1 = -(b01 * ((f1 / a.dtm) / a.dtfs));
        st2(a.E_bar + o2, make_double2(e0, e1));
        st2(a.eH_bar + o2, make_double2(h0, h1));
        double2 xn = make_double2(0.0, 0.0);
        if (a.x_bar) {
            const double2 bo = ld2(a.Bo_surf + q), xb = ld2(a.x_bar + o2);
            xn = make_double2(bo.x * xb.x, bo.y * xb.y);
        }
        st2(a.eN_bar + o2, xn);
    }
    const double2 cu = ld2(a.dyG + q), cv = ld2(a.dxG + q);
    const int64_t o3 = m * g.ms3 + (int64_t)t * g.nr * g.plane + q0;
    for (int k = 0; k < g.nr; ++k) {
        const int64_t ok = o3 + (int64_t)k * g.plane;
        const double dz = a.drF[k];
        const double cu0 = cu.x * dz, cu1 = cu.y * dz, cv0 = cv.x * dz, cv1 = cv.y * dz;
        const double2 hw = ld2(a.hFacW + ok), gu = ld2(a.gU + ok), hs = ld2(a.hFacS + ok), gv = ld2(a.gV + ok);
        st2(a.gU_bar + ok, make_double2((pu0 * (cu0 * hw.x)) / a.dtm, (pu1 * (cu1 * hw.y)) / a.dtm));
        st2(a.hW_bar + ok, make_double2((pu0 * (cu0 * gu.x)) / a.dtm, (pu1 * (cu1 * gu.y)) / a.dtm));
        st2(a.gV_bar + ok, make_double2((pv0 * (cv0 * hs.x)) / a.dtm, (pv1 * (cv1 * hs.y)) / a.dtm));
        st2(a.hS_bar + ok, make_double2((pv0 * (cv0 * gv.x)) / a.dtm, (pv1 * (cv1 * gv.y)) / a.dtm));
    }
}

// Reverse mode of the level sums of UPDATE_CG2D: ct = exch^T(aC2d_bar) (in work), then on j, i = 1..sNy+1,
// 1..sNx+1: aW_tot = (aW_bar - C(j, i)) - C(j, i-1), aS_tot = (aS_bar - C(j, i)) - C(j-1, i) with C = ct on the
// interior and 0 elsewhere; hFacW_bar(k) = (aW_tot ((norm isp) idf)) ((dyG drF(k)) recip_dxC) (hFacS_bar alike),
// 0 off that range. One thread per pair of adjacent points.
__global__ void __launch_bounds__(kPairThreads)
    sfp_opsum_adjoint_kernel(SfpGeo g, const double* __restrict__ dyG, const double* __restrict__ dxG,
                             const double* __restrict__ recip_dxC, const double* __restrict__ recip_dyC,
                             const double* __restrict__ drF, const double* __restrict__ norm, double isp, double idf,
                             const double* __restrict__ aWb, const double* __restrict__ aSb,
                             const double* __restrict__ ct, double* __restrict__ hWb, double* __restrict__ hSb) {
    const int p = blockIdx.x * kPairThreads + threadIdx.x, t = blockIdx.y, m = blockIdx.z;
    const int q0 = 2 * p;
    if (q0 >= g.plane) return;
    const int j = q0 / g.nx, i0 = q0 - j * g.nx;
    const int64_t q = (int64_t)t * g.plane + q0;
    const int64_t o2 = m * g.ms2 + q;
    const double* c2 = ct + m * g.ms2 + (int64_t)t * g.plane;
    auto C = [&](int jj, int ii) {
        const bool in = jj >= g.OL && jj < g.OL + g.sNy && ii >= g.OL && ii < g.OL + g.sNx;
        return in ? c2[jj * g.nx + ii] : 0.0;
    };
    const bool jin = in_range(j, g.OL, g.OL + g.sNy);
    const bool in0 = jin && in_range(i0, g.OL, g.OL + g.sNx), in1 = jin && in_range(i0 + 1, g.OL, g.OL + g.sNx);
    double w0 = 0.0, w1 = 0.0, s0 = 0.0, s1 = 0.0;
    double2 cW = make_double2(0.0, 0.0), cS = cW, rW = cW, rS = cW;
    if (in0 || in1) {
        const double2 aw = ld2(aWb + o2), as = ld2(aSb + o2);
        const double c00 = C(j, i0), c01 = C(j, i0 + 1);
        const double fac = (norm[m] * isp) * idf;
        w0 = ((aw.x - c00) - C(j, i0 - 1)) * fac;
        w1 = ((aw.y - c01) - c00) * fac;
        s0 = ((as.x - c00) - C(j - 1, i0)) * fac;
        s1 = ((as.y - c01) - C(j - 1, i0 + 1)) * fac;
        cW = ld2(dyG + q), cS = ld2(dxG + q), rW = ld2(recip_dxC + q), rS = ld2(recip_dyC + q);
    }
    const int64_t o3 = m * g.ms3 + (int64_t)t * g.nr * g.plane + q0;
    for (int k = 0; k < g.nr; ++k) {
        const double dz = drF[k];
        const int64_t ok = o3 + (int64_t)k * g.plane;
        st2(hWb + ok, make_double2(in0 ? w0 * ((cW.x * dz) * rW.x) : 0.0, in1 ? w1 * ((cW.y * dz) * rW.y) : 0.0));
        st2(hSb + ok, make_double2(in0 ? s0 * ((cS.x * dz) * rS.x) : 0.0, in1 ? s1 * ((cS.y * dz) * rS.y) : 0.0));
    }
}

// INI_CG2D's cg2dNorm (reference core/cg2d.py:108-134, ini_cg2d.F:90-135). Per interior point of a member:
// aW = sum over k = 1..nr in order of (fac ((dyG drF(k)) hFacW)) recip_dxC (aS likewise), fac = implicSurfPress
// implicDiv2DFlow; the member's max of |aW|, |aS| is kept as the bits of a nonnegative double in `bits` (the
// unsigned-integer order of nonnegative doubles is their numeric order, so the atomic maxima are order-free and
// the result is deterministic); ini_norm_finish turns it into 1 / max (1 when the max is 0), in place.
__global__ void __launch_bounds__(256) ini_norm_zero_kernel(unsigned long long* bits, int B) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < B) bits[m] = 0ull;
}

__global__ void __launch_bounds__(256) ini_norm_max_kernel(SfpGeo g, const double* __restrict__ dyG,
                                                           const double* __restrict__ dxG,
                                                           const double* __restrict__ recip_dxC,
                                                           const double* __restrict__ recip_dyC,
                                                           const double* __restrict__ drF,
                                                           const double* __restrict__ hFacW,
                                                           const double* __restrict__ hFacS, double fac,
                                                           unsigned long long* bits) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = blockIdx.y, m = blockIdx.z;
    double mx = 0.0;
    if (e < g.sNx * g.sNy) {
        const int j = e / g.sNx + g.OL, i = e % g.sNx + g.OL;
        const int64_t q = (int64_t)t * g.plane + (int64_t)j * g.nx + i;
        const double cW = dyG[q], cS = dxG[q], rW = recip_dxC[q], rS = recip_dyC[q];
        const int64_t o3 = m * g.ms3 + (int64_t)t * g.nr * g.plane + (int64_t)j * g.nx + i;
        double aW = 0.0, aS = 0.0;
        for (int k = 0; k < g.nr; ++k) {
            const double dz = drF[k];
            const int64_t ok = o3 + (int64_t)k * g.plane;
            aW = aW + (fac * ((cW * dz) * hFacW[ok])) * rW;
            aS = aS + (fac * ((cS * dz) * hFacS[ok])) * rS;
        }
        mx = fmax(fabs(aW), fabs(aS));
    }
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&bits[m], (unsigned long long)__double_as_longlong(mx));
}

__global__ void __launch_bounds__(256) ini_norm_finish_kernel(unsigned long long* bits, int B) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= B) return;
    const double v = __longlong_as_double((long long)bits[m]);
    reinterpret_cast<double*>(bits)[m] = v != 0.0 ? 1.0 / v : 1.0;
}

int check_sfp(const mj_sfp_grid* gr, const mj_fs_params* p, const mj_sfp_config* c, SfpGeo* g, const char* who) {
    MJ_REQUIRE(gr && p && c, "%s: null grid/params/config", who);
    MJ_REQUIRE(c->batch >= 1 && c->batch <= 65535 && c->ntiles >= 1 && c->ntiles <= 65535 && c->nr >= 1,
               "%s: batch and ntiles must be in 1..65535, nr >= 1", who);
    MJ_REQUIRE(c->sNx >= 2 && c->sNx % 2 == 0 && c->sNy >= 1 && c->OL >= 1 && c->sNx + 2 * c->OL <= kCols,
               "%s: need sNx even, sNy, OL >= 1 and sNx + 2 OL <= %d", who, kCols);
    MJ_REQUIRE(c->batch == 1 || (c->member_stride3d % 2 == 0 && c->member_stride2d % 2 == 0),
               "%s: member strides must be even (16-byte aligned members)", who);
    g->B = c->batch;
    g->T = c->ntiles;
    g->sNx = c->sNx;
    g->sNy = c->sNy;
    g->OL = c->OL;
    g->nx = c->sNx + 2 * c->OL;
    g->ny = c->sNy + 2 * c->OL;
    g->nr = c->nr;
    g->plane = (int64_t)g->nx * g->ny;
    g->ms3 = c->member_stride3d;
    g->ms2 = c->member_stride2d;
    MJ_REQUIRE(c->batch == 1 || (g->ms3 >= (int64_t)g->T * g->nr * g->plane && g->ms2 >= (int64_t)g->T * g->plane),
               "%s: member strides smaller than one member's arrays", who);
    MJ_REQUIRE(gr->dyG && gr->dxG && gr->drF, "%s: null grid field", who);
    MJ_REQUIRE(aligned16(gr->dyG) && aligned16(gr->dxG) && aligned16(gr->recip_dxC) && aligned16(gr->recip_dyC) &&
                   aligned16(gr->rA) && aligned16(gr->maskInC),
               "%s: 2-D grid fields must be 16-byte aligned", who);
    MJ_REQUIRE(p->deltaTMom != 0.0 && p->deltaTFreeSurf != 0.0, "%s: deltaTMom and deltaTFreeSurf must be nonzero",
               who);
    return MJ_OK;
}

dim3 sfp_grid(const SfpGeo& g) { return dim3((unsigned)((g.ny + kRows - 1) / kRows), (unsigned)g.T, (unsigned)g.B); }
dim3 pair_grid(const SfpGeo& g) {
    return dim3((unsigned)((g.plane / 2 + kPairThreads - 1) / kPairThreads), (unsigned)g.T, (unsigned)g.B);
}

}  // namespace
}  // namespace mj

using namespace mj;

extern "C" int mj_ini_cg2d_norm(const mj_sfp_grid* grid, const mj_sfp_config* cfg, double implicSurfPress,
                                double implicDiv2DFlow, const double* hFacW, const double* hFacS, double* cg2dNorm,
                                void* stream) {
    SfpGeo g;
    const mj_fs_params unit{1.0, 1.0, 1.0, implicSurfPress, implicDiv2DFlow, 1.0, 0};  // only the checks read it
    int rc = check_sfp(grid, &unit, cfg, &g, "mj_ini_cg2d_norm");
    if (rc) return rc;
    MJ_REQUIRE(grid->recip_dxC && grid->recip_dyC, "mj_ini_cg2d_norm: null grid field");
    MJ_REQUIRE(hFacW && hFacS && cg2dNorm, "mj_ini_cg2d_norm: null array");
    MJ_REQUIRE(cg2dNorm != hFacW && cg2dNorm != hFacS, "mj_ini_cg2d_norm: cg2dNorm aliases an input");
    MJ_REQUIRE(((uintptr_t)cg2dNorm & 7) == 0, "mj_ini_cg2d_norm: cg2dNorm must be 8-byte aligned");
    hipStream_t st = as_stream(stream);
    unsigned long long* bits = reinterpret_cast<unsigned long long*>(cg2dNorm);
    const unsigned nb = (unsigned)((g.B + 255) / 256);
    hipLaunchKernelGGL(ini_norm_zero_kernel, dim3(nb), dim3(256), 0, st, bits, g.B);
    MJ_HIP(hipGetLastError());
    const dim3 grd((unsigned)((g.sNx * g.sNy + 255) / 256), (unsigned)g.T, (unsigned)g.B);
    hipLaunchKernelGGL(ini_norm_max_kernel, grd, dim3(256), 0, st, g, grid->dyG, grid->dxG, grid->recip_dxC,
                       grid->recip_dyC, grid->drF, hFacW, hFacS, implicSurfPress * implicDiv2DFlow, bits);
    MJ_HIP(hipGetLastError());
    hipLaunchKernelGGL(ini_norm_finish_kernel, dim3(nb), dim3(256), 0, st, bits, g.B);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}

extern "C" int mj_update_cg2d(const mj_halo_plan* plan_T, const mj_sfp_grid* grid, const mj_fs_params* p,
                              const mj_sfp_config* cfg, const double* cg2dNorm, double cg2dpcOffDFac,
                              const double* hFacW, const double* hFacS, const double* recip_Bo, double* aW2d,
                              double* aS2d, double* aC2d, double* pW, double* pS, double* pC, void* stream) {
    SfpGeo g;
    int rc = check_sfp(grid, p, cfg, &g, "mj_update_cg2d");
    if (rc) return rc;
    MJ_REQUIRE(plan_T, "mj_update_cg2d: null T plan");
    MJ_REQUIRE(grid->recip_dxC && grid->recip_dyC && grid->rA, "mj_update_cg2d: null grid field");
    MJ_REQUIRE(cg2dNorm && hFacW && hFacS && recip_Bo && aW2d && aS2d && aC2d && pW && pS && pC,
               "mj_update_cg2d: null array");
    MJ_REQUIRE(mj_halo_plan_inplace_ok(plan_T) == 1, "mj_update_cg2d: plan_T must be a scalar kind exchangeable in place");
    const double* ins[] = {hFacW, hFacS, recip_Bo, cg2dNorm};
    const double* outs[] = {aW2d, aS2d, aC2d, pW, pS, pC};
    for (const double* o : outs)
        for (const double* in : ins) MJ_REQUIRE(o != in, "mj_update_cg2d: an output aliases an input");
    for (int x = 0; x < 6; ++x)
        for (int y = x + 1; y < 6; ++y) MJ_REQUIRE(outs[x] != outs[y], "mj_update_cg2d: two outputs alias");
    MJ_REQUIRE(aligned16(hFacW) && aligned16(hFacS) && aligned16(aW2d) && aligned16(aS2d),
               "mj_update_cg2d: hFacW, hFacS, aW2d, aS2d must be 16-byte aligned");
    hipStream_t st = as_stream(stream);
    const dim3 grd = sfp_grid(g), blk(kCols, kRows);
    hipLaunchKernelGGL(sfp_opsum_kernel, pair_grid(g), dim3(kPairThreads), 0, st, g, grid->dyG, grid->dxG, grid->recip_dxC, grid->recip_dyC,
                       grid->drF, hFacW, hFacS, cg2dNorm, p->implicSurfPress, p->implicDiv2DFlow, aW2d, aS2d);
    MJ_HIP(hipGetLastError());
    hipLaunchKernelGGL(sfp_aC_kernel, grd, blk, 0, st, g, (const double*)aW2d, (const double*)aS2d, recip_Bo, grid->rA,
                       cg2dNorm, p->freeSurfFac, p->deltaTMom, p->deltaTFreeSurf, aC2d);
    MJ_HIP(hipGetLastError());
    mj_halo_geom hg{g.B, g.B == 1 ? (int64_t)g.T * g.plane : g.ms2, g.plane, 1, g.plane};
    if ((rc = mj_halo_exchange(plan_T, aC2d, aC2d, &hg, stream))) return rc;
    hipLaunchKernelGGL(sfp_precond_kernel, grd, blk, 0, st, g, (const double*)aW2d, (const double*)aS2d,
                       (const double*)aC2d, cg2dpcOffDFac, pW, pS, pC);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}

extern "C" int mj_cg2d_rhs(const mj_sfp_grid* grid, const mj_fs_params* p, const mj_sfp_config* cfg,
                           const double* gU, const double* gV, const double* hFacW, const double* hFacS,
                           const double* EmPmR, const double* etaN, const double* etaH, const double* Bo_surf,
                           double* cg2d_b, double* cg2d_x, void* stream) {
    SfpGeo g;
    int rc = check_sfp(grid, p, cfg, &g, "mj_cg2d_rhs");
    if (rc) return rc;
    MJ_REQUIRE(grid->rA && grid->maskInC, "mj_cg2d_rhs: null grid field");
    MJ_REQUIRE(!grid->deepFac2F || grid->kSurfC, "mj_cg2d_rhs: deepFac2F needs kSurfC");
    MJ_REQUIRE(gU && gV && hFacW && hFacS && etaN && etaH && Bo_surf && cg2d_b && cg2d_x, "mj_cg2d_rhs: null array");
    MJ_REQUIRE(EmPmR || !p->useRealFreshWaterFlux, "mj_cg2d_rhs: useRealFreshWaterFlux needs EmPmR");
    MJ_REQUIRE(cg2d_b != cg2d_x, "mj_cg2d_rhs: cg2d_b and cg2d_x alias");
    const double* ins[] = {gU, gV, hFacW, hFacS, EmPmR, etaN, etaH, Bo_surf};
    for (const double* in : ins) {
        MJ_REQUIRE(in != cg2d_b && in != cg2d_x, "mj_cg2d_rhs: an output aliases an input");
        MJ_REQUIRE(aligned16(in), "mj_cg2d_rhs: arrays must be 16-byte aligned");
    }
    MJ_REQUIRE(aligned16(cg2d_b) && aligned16(cg2d_x), "mj_cg2d_rhs: arrays must be 16-byte aligned");
    RhsArgs a;
    a.dyG = grid->dyG;
    a.dxG = grid->dxG;
    a.rA = grid->rA;
    a.maskInC = grid->maskInC;
    a.drF = grid->drF;
    a.dfc = grid->deepFacC;
    a.rfc = grid->rhoFacC;
    a.d2f = grid->deepFac2F;
    a.kSurfC = grid->kSurfC;
    a.gU = gU;
    a.gV = gV;
    a.hFacW = hFacW;
    a.hFacS = hFacS;
    a.EmPmR = EmPmR;
    a.etaN = etaN;
    a.etaH = etaH;
    a.Bo_surf = Bo_surf;
    a.b = cg2d_b;
    a.x = cg2d_x;
    a.tmpFac = (p->freeSurfFac * p->mass2rUnit) * p->implicDiv2DFlow;
    a.fsf = p->freeSurfFac;
    a.dtm = p->deltaTMom;
    a.dtfs = p->deltaTFreeSurf;
    a.rfw = p->useRealFreshWaterFlux ? 1 : 0;
    hipLaunchKernelGGL(sfp_rhs_kernel, pair_grid(g), dim3(kPairThreads), 0, as_stream(stream), g, a);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}

extern "C" int mj_cg2d_rhs_adjoint(const mj_sfp_grid* grid, const mj_fs_params* p, const mj_sfp_config* cfg,
                                   const double* gU, const double* gV, const double* hFacW, const double* hFacS,
                                   const double* Bo_surf, const double* cg2d_b_bar, const double* cg2d_x_bar,
                                   double* gU_bar, double* gV_bar, double* hFacW_bar, double* hFacS_bar,
                                   double* EmPmR_bar, double* etaN_bar, double* etaH_bar, void* stream) {
    SfpGeo g;
    int rc = check_sfp(grid, p, cfg, &g, "mj_cg2d_rhs_adjoint");
    if (rc) return rc;
    MJ_REQUIRE(!grid->deepFacC && !grid->rhoFacC, "mj_cg2d_rhs_adjoint: deepFacC / rhoFacC must be NULL (ones)");
    MJ_REQUIRE(grid->rA && grid->maskInC, "mj_cg2d_rhs_adjoint: null grid field");
    MJ_REQUIRE(!grid->deepFac2F || grid->kSurfC, "mj_cg2d_rhs_adjoint: deepFac2F needs kSurfC");
    const double* ins[] = {gU, gV, hFacW, hFacS, Bo_surf, cg2d_b_bar};
    double* outs[] = {gU_bar, gV_bar, hFacW_bar, hFacS_bar, EmPmR_bar, etaN_bar, etaH_bar};
    for (const double* in : ins) MJ_REQUIRE(in && aligned16(in), "mj_cg2d_rhs_adjoint: null or unaligned input");
    MJ_REQUIRE(aligned16(cg2d_x_bar), "mj_cg2d_rhs_adjoint: unaligned cg2d_x_bar");
    for (double* o : outs) {
        MJ_REQUIRE(o && aligned16(o), "mj_cg2d_rhs_adjoint: null or unaligned output");
        for (const double* in : ins) MJ_REQUIRE(o != in, "mj_cg2d_rhs_adjoint: an output aliases an input");
        MJ_REQUIRE(o != cg2d_x_bar, "mj_cg2d_rhs_adjoint: an output aliases an input");
    }
    for (int x = 0; x < 7; ++x)
        for (int y = x + 1; y < 7; ++y) MJ_REQUIRE(outs[x] != outs[y], "mj_cg2d_rhs_adjoint: two outputs alias");
    RhsAdjArgs a;
    a.dyG = grid->dyG;
    a.dxG = grid->dxG;
    a.rA = grid->rA;
    a.maskInC = grid->maskInC;
    a.drF = grid->drF;
    a.d2f = grid->deepFac2F;
    a.kSurfC = grid->kSurfC;
    a.gU = gU;
    a.gV = gV;
    a.hFacW = hFacW;
    a.hFacS = hFacS;
    a.Bo_surf = Bo_surf;
    a.b_bar = cg2d_b_bar;
    a.x_bar = cg2d_x_bar;
    a.gU_bar = gU_bar;
    a.gV_bar = gV_bar;
    a.hW_bar = hFacW_bar;
    a.hS_bar = hFacS_bar;
    a.E_bar = EmPmR_bar;
    a.eN_bar = etaN_bar;
    a.eH_bar = etaH_bar;
    a.tmpFac = (p->freeSurfFac * p->mass2rUnit) * p->implicDiv2DFlow;
    a.fsf = p->freeSurfFac;
    a.dtm = p->deltaTMom;
    a.dtfs = p->deltaTFreeSurf;
    a.rfw = p->useRealFreshWaterFlux ? 1 : 0;
    hipLaunchKernelGGL(sfp_rhs_adjoint_kernel, pair_grid(g), dim3(kPairThreads), 0, as_stream(stream), g, a);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}

extern "C" int mj_update_cg2d_adjoint(const mj_halo_plan* plan_T, const mj_sfp_grid* grid, const mj_fs_params* p,
                                      const mj_sfp_config* cfg, const double* cg2dNorm, const double* aW2d_bar,
                                      const double* aS2d_bar, const double* aC2d_bar, double* hFacW_bar,
                                      double* hFacS_bar, double* work, void* stream) {
    SfpGeo g;
    int rc = check_sfp(grid, p, cfg, &g, "mj_update_cg2d_adjoint");
    if (rc) return rc;
    MJ_REQUIRE(plan_T, "mj_update_cg2d_adjoint: null T plan");
    MJ_REQUIRE(grid->recip_dxC && grid->recip_dyC, "mj_update_cg2d_adjoint: null grid field");
    const double* ins[] = {cg2dNorm, aW2d_bar, aS2d_bar, aC2d_bar};
    double* outs[] = {hFacW_bar, hFacS_bar, work};
    for (const double* in : ins) MJ_REQUIRE(in, "mj_update_cg2d_adjoint: null input");
    for (double* o : outs) {
        MJ_REQUIRE(o && aligned16(o), "mj_update_cg2d_adjoint: null or unaligned output/work");
        for (const double* in : ins) MJ_REQUIRE(o != in, "mj_update_cg2d_adjoint: an output aliases an input");
    }
    MJ_REQUIRE(hFacW_bar != hFacS_bar && work != hFacW_bar && work != hFacS_bar,
               "mj_update_cg2d_adjoint: outputs alias");
    MJ_REQUIRE(aligned16(aW2d_bar) && aligned16(aS2d_bar), "mj_update_cg2d_adjoint: unaligned aW2d_bar/aS2d_bar");
    mj_halo_geom hg{g.B, g.B == 1 ? (int64_t)g.T * g.plane : g.ms2, g.plane, 1, g.plane};
    if ((rc = mj_halo_exchange_adjoint(plan_T, aC2d_bar, nullptr, work, nullptr, &hg, stream))) return rc;
    hipLaunchKernelGGL(sfp_opsum_adjoint_kernel, pair_grid(g), dim3(kPairThreads), 0, as_stream(stream), g, grid->dyG,
                       grid->dxG, grid->recip_dxC, grid->recip_dyC, grid->drF, cg2dNorm, p->implicSurfPress,
                       p->implicDiv2DFlow, aW2d_bar, aS2d_bar, (const double*)work, hFacW_bar, hFacS_bar);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}
