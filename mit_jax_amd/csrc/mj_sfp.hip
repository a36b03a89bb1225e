// The CG2D operator and right-hand side on gfx950: UPDATE_CG2D and cg2d_rhs of SOLVE_FOR_PRESSURE
// (reference core/solve_for_pressure.py:29-106).
//
// Both are column sums over the levels of 3-D [T, nr, ny, nx] FP64 fields into 2-D [T, ny, nx] results: HBM-bound
// streams with a few flops per byte. One thread owns one (j, i) column of one member and walks the levels in the
// reference's order (UPDATE_CG2D k = 1..Nr, the divergence of cg2d_b k = Nr..1), so every sum is the literal
// sequential one and the results are bitwise those of the reference's loops (-ffp-contract=off). A block covers
// kRows rows x 128 columns of one (member, tile): the level loads of a row are contiguous 8-byte words across the
// lanes, and the i+1 / j+1 neighbours the divergence reads are mostly the block's own loads (L1 hits).
//
// UPDATE_CG2D runs as: sums -> aW2d, aS2d (whole arrays); aC2d on the interior (whole array); the T exchange of
// aC2d; the preconditioner on j, i = 1..sNy+1, 1..sNx+1 (it reads aC2d at i-1 and j-1, i.e. exchanged halo points).
#include "mj_common.h"

namespace mj {

namespace {

constexpr int kCols = 128;  // threads across a row (nx <= 128)
constexpr int kRows = 8;    // rows per block

struct SfpGeo {
    int B, T, nx, ny, sNx, sNy, OL, nr;
    int64_t plane, ms3, ms2;  // ny*nx; member strides
};

__device__ __forceinline__ bool in_range(int v, int lo, int hi) { return v >= lo && v <= hi; }

// aW2d, aS2d: ((sum_k ((dyG * drF(k)) * hFacW(k)) * recip_dxC) * cg2dNorm) * implicSurfPress * implicDiv2DFlow on
// Fortran j, i = 1..sNy+1, 1..sNx+1 (reference solve_for_pressure.py:72-80, :101-115 with _ksum :23-26, :119-132),
// 0 elsewhere.
__global__ void __launch_bounds__(kCols * kRows)
    sfp_opsum_kernel(SfpGeo g, const double* __restrict__ dyG, const double* __restrict__ dxG,
                     const double* __restrict__ recip_dxC, const double* __restrict__ recip_dyC,
                     const double* __restrict__ drF, const double* __restrict__ hFacW,
                     const double* __restrict__ hFacS, const double* __restrict__ norm, double isp, double idf,
                     double* __restrict__ aW2d, double* __restrict__ aS2d) {
    const int i = threadIdx.x, j = blockIdx.x * kRows + threadIdx.y, t = blockIdx.y, m = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int64_t q = (int64_t)t * g.plane + (int64_t)j * g.nx + i;
    const int64_t o2 = m * g.ms2 + q;
    double w = 0.0, s = 0.0;
    if (in_range(j, g.OL, g.OL + g.sNy) && in_range(i, g.OL, g.OL + g.sNx)) {
        const double cW = dyG[q], cS = dxG[q], rW = recip_dxC[q], rS = recip_dyC[q];
        const double* hW = hFacW + m * g.ms3 + (int64_t)t * g.nr * g.plane + (int64_t)j * g.nx + i;
        const double* hS = hFacS + m * g.ms3 + (int64_t)t * g.nr * g.plane + (int64_t)j * g.nx + i;
#pragma unroll 10
        for (int k = 0; k < g.nr; ++k) {
            const double dz = drF[k];
            w = w + ((cW * dz) * hW[k * g.plane]) * rW;
            s = s + ((cS * dz) * hS[k * g.plane]) * rS;
        }
        const double nm = norm[m];
        w = ((w * nm) * isp) * idf;
        s = ((s * nm) * isp) * idf;
    }
    aW2d[o2] = w;
    aS2d[o2] = s;
}

// aC2d = -(aW(i) + aW(i+1) + aS(j) + aS(j+1) + freeSurfFac cg2dNorm recip_Bo rA / deltaTMom / deltaTFreeSurf) on
// the interior, 0 elsewhere (reference solve_for_pressure.py:134-138).
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
// solve_for_pressure.py:139-149); other points untouched.
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

// pfu = (((((dyG deepFacC(k)) drF(k)) hFacW) rhoFacC(k)) gU) / deltaTMom (reference solve_for_pressure.py:89-94);
// the factors of a NULL vertical array are 1 and skipped (x * 1 == x).
__device__ __forceinline__ double pf(double c, const double* dfc, const double* rfc, double dz, int k, double h,
                                     double v, double dtm) {
    double xa = c;
    if (dfc) xa = xa * dfc[k];
    xa = (xa * dz) * h;
    if (rfc) xa = xa * rfc[k];
    return (xa * v) / dtm;
}

// cg2d_x = Bo_surf etaN on the whole array; cg2d_b on the interior (0 elsewhere): the fresh-water term, the
// divergence of the pf fluxes for k = nr..1 in order, the exactConserv etaH term (reference
// solve_for_pressure.py:68-106).
__global__ void __launch_bounds__(kCols * kRows) sfp_rhs_kernel(SfpGeo g, RhsArgs a) {
    const int i = threadIdx.x, j = blockIdx.x * kRows + threadIdx.y, t = blockIdx.y, m = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int64_t q = (int64_t)t * g.plane + (int64_t)j * g.nx + i;
    const int64_t o2 = m * g.ms2 + q;
    a.x[o2] = a.Bo_surf[q] * a.etaN[o2];
    double b = 0.0;
    if (in_range(j, g.OL, g.OL + g.sNy - 1) && in_range(i, g.OL, g.OL + g.sNx - 1)) {
        if (a.rfw) b = (((a.tmpFac * a.rA[q]) * a.EmPmR[o2]) / a.dtm) * a.maskInC[q];
        const double cu0 = a.dyG[q], cu1 = a.dyG[q + 1], cv0 = a.dxG[q], cv1 = a.dxG[q + g.nx];
        const int64_t o3 = m * g.ms3 + (int64_t)t * g.nr * g.plane + (int64_t)j * g.nx + i;
        const double *hW = a.hFacW + o3, *hS = a.hFacS + o3, *gU = a.gU + o3, *gV = a.gV + o3;
#pragma unroll 4
        for (int k = g.nr - 1; k >= 0; --k) {
            const int64_t ok = (int64_t)k * g.plane;
            const double dz = a.drF[k];
            const double pu1 = pf(cu1, a.dfc, a.rfc, dz, k, hW[ok + 1], gU[ok + 1], a.dtm);
            const double pu0 = pf(cu0, a.dfc, a.rfc, dz, k, hW[ok], gU[ok], a.dtm);
            const double pv1 = pf(cv1, a.dfc, a.rfc, dz, k, hS[ok + g.nx], gV[ok + g.nx], a.dtm);
            const double pv0 = pf(cv0, a.dfc, a.rfc, dz, k, hS[ok], gV[ok], a.dtm);
            b = (b + pu1) - pu0;
            b = (b + pv1) - pv0;
        }
        double e = a.fsf * a.rA[q];
        if (a.d2f) {
            int ks = a.kSurfC[q];
            ks = ks < 1 ? 1 : (ks > g.nr + 1 ? g.nr + 1 : ks);
            e = e * a.d2f[ks - 1];
        }
        b = b - ((e / a.dtm) / a.dtfs) * a.etaH[o2];
    }
    a.b[o2] = b;
}

int check_sfp(const mj_sfp_grid* gr, const mj_fs_params* p, const mj_sfp_config* c, SfpGeo* g, const char* who) {
    MJ_REQUIRE(gr && p && c, "%s: null grid/params/config", who);
    MJ_REQUIRE(c->batch >= 1 && c->batch <= 65535 && c->ntiles >= 1 && c->ntiles <= 65535 && c->nr >= 1,
               "%s: batch and ntiles must be in 1..65535, nr >= 1", who);
    MJ_REQUIRE(c->sNx >= 1 && c->sNy >= 1 && c->OL >= 1 && c->sNx + 2 * c->OL <= kCols,
               "%s: need sNx, sNy, OL >= 1 and sNx + 2 OL <= %d", who, kCols);
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
    MJ_REQUIRE(p->deltaTMom != 0.0 && p->deltaTFreeSurf != 0.0, "%s: deltaTMom and deltaTFreeSurf must be nonzero",
               who);
    return MJ_OK;
}

dim3 sfp_grid(const SfpGeo& g) { return dim3((unsigned)((g.ny + kRows - 1) / kRows), (unsigned)g.T, (unsigned)g.B); }

}  // namespace
}  // namespace mj

using namespace mj;

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
    hipStream_t st = as_stream(stream);
    const dim3 grd = sfp_grid(g), blk(kCols, kRows);
    hipLaunchKernelGGL(sfp_opsum_kernel, grd, blk, 0, st, g, grid->dyG, grid->dxG, grid->recip_dxC, grid->recip_dyC,
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
    for (const double* in : ins)
        MJ_REQUIRE(in != cg2d_b && in != cg2d_x, "mj_cg2d_rhs: an output aliases an input");
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
    hipLaunchKernelGGL(sfp_rhs_kernel, sfp_grid(g), dim3(kCols, kRows), 0, as_stream(stream), g, a);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}
