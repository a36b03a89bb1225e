// The DST3 interface fluxes alone on gfx950: dst3_adv_x / dst3_adv_y of the reference (pkgs/gad.py:409-453,
// GAD_DST3_ADV_X/Y with calcCFL), for every (member, tile, level) plane at once, for callers that assemble their own
// update from the fluxes. (The fused GAD_ADVECTION kernel evaluates the same interface values inside its sweeps.)
//
// Each thread produces two horizontally adjacent points (even i) of one plane: every operand is one aligned
// 16-byte load per row it spans, so a wavefront's loads cover contiguous 1 KB runs of each row. The arithmetic is
// dst3_flux (mj_common.h), the reference's expressions in their order; points outside the reference's loop range
// are written 0, as the reference's zeros_like output holds them.
#include "mj_common.h"

namespace mj {

struct FluxDev {
    int B, T, K, ny, nx;
    int64_t ms;  // member stride of the 3-D arrays
    double oneSixth;
    const double *dt, *rdeep;                  // [K]
    const double *trans, *vel, *mask, *tr;     // [B][T, K, ny, nx]
    const double* rd;                          // [T, ny, nx]
    double* out;
};

__device__ __forceinline__ double2 ld2d(const double* p) { return *reinterpret_cast<const double2*>(p); }

// blockIdx.x: plane (member, tile, level), tile-major inside a member; blockIdx.y * blockDim.x + threadIdx.x: the
// column pair within the plane.
template <int DIR>
__global__ void __launch_bounds__(256) dst3_adv_kernel(FluxDev a) {
    const int npair = a.ny * a.nx / 2;
    const int pr = blockIdx.y * blockDim.x + threadIdx.x;
    if (pr >= npair) return;
    const int64_t pl = blockIdx.x;
    const int k = (int)(pl % a.K), t = (int)((pl / a.K) % a.T), m = (int)(pl / ((int64_t)a.K * a.T));
    const int q = 2 * pr, j = q / a.nx, i = q - j * a.nx;
    const int64_t plane = (int64_t)a.ny * a.nx;
    const int64_t o = m * a.ms + ((int64_t)t * a.K + k) * plane + q;
    const double *tr = a.tr + o, *msk = a.mask + o;
    const double dt = a.dt[k], rdeep = a.rdeep[k];
    double2 res = make_double2(0.0, 0.0);
    if (DIR == 0) {
        // dst3_adv_x: i = 2 .. nx-2 on every row (Fortran 1-OLx+2 .. sNx+OLx-1)
        // (i even: both points need i >= 2, so the loads at i - 2 stay in the row; T2/M2 only when i + 1 <= nx - 2)
        const bool in0 = i >= 2 && i <= a.nx - 2, in1 = i >= 2 && i + 1 <= a.nx - 2;
        if (in0 || in1) {
            const double2 T0 = ld2d(tr - 2), T1 = ld2d(tr), M0 = ld2d(msk - 2), M1 = ld2d(msk);
            const double2 U = ld2d(a.trans + o), V = ld2d(a.vel + o), R = ld2d(a.rd + (int64_t)t * plane + q);
            if (in0)
                res.x = dst3_flux(U.x, fabs(V.x * dt * R.x * rdeep), a.oneSixth, T0.x, T0.y, T1.x, T1.y, M0.y, M1.x,
                                  M1.y);
            if (in1) {
                const double2 T2 = ld2d(tr + 2), M2 = ld2d(msk + 2);
                res.y = dst3_flux(U.y, fabs(V.y * dt * R.y * rdeep), a.oneSixth, T0.y, T1.x, T1.y, T2.x, M1.x, M1.y,
                                  M2.x);
            }
        }
    } else {
        // dst3_adv_y: j = 2 .. ny-2 (Fortran 1-OLy+2 .. sNy+OLy-1) on every column
        if (j >= 2 && j <= a.ny - 2) {
            const int nx = a.nx;
            const double2 Tm2 = ld2d(tr - 2 * nx), Tm1 = ld2d(tr - nx), T0 = ld2d(tr), Tp1 = ld2d(tr + nx);
            const double2 Mm1 = ld2d(msk - nx), M0 = ld2d(msk), Mp1 = ld2d(msk + nx);
            const double2 U = ld2d(a.trans + o), V = ld2d(a.vel + o), R = ld2d(a.rd + (int64_t)t * plane + q);
            res.x = dst3_flux(U.x, fabs(V.x * dt * R.x * rdeep), a.oneSixth, Tm2.x, Tm1.x, T0.x, Tp1.x, Mm1.x, M0.x,
                              Mp1.x);
            res.y = dst3_flux(U.y, fabs(V.y * dt * R.y * rdeep), a.oneSixth, Tm2.y, Tm1.y, T0.y, Tp1.y, Mm1.y, M0.y,
                              Mp1.y);
        }
    }
    *reinterpret_cast<double2*>(a.out + o) = res;
}

}  // namespace mj

using namespace mj;

namespace {

int dst3_adv(int dir, const mj_dst3_flux_config* c, double oneSixth, const double* deltaTloc, const double* trans,
             const double* vel, const double* mask, const double* tracer, const double* rd, const double* rdeep,
             double* out, void* stream) {
    const char* fn = dir == 0 ? "mj_dst3_adv_x" : "mj_dst3_adv_y";
    MJ_REQUIRE(c, "%s: null config", fn);
    MJ_REQUIRE(deltaTloc && trans && vel && mask && tracer && rd && rdeep && out, "%s: null argument", fn);
    MJ_REQUIRE(c->batch >= 1 && c->ntiles >= 1 && c->nr >= 1, "%s: batch, ntiles, nr must be >= 1", fn);
    MJ_REQUIRE(c->sNx >= 1 && c->sNy >= 1 && c->OLx >= 2 && c->OLy >= 2, "%s: sNx, sNy >= 1 and OLx, OLy >= 2", fn);
    const int nx = c->sNx + 2 * c->OLx, ny = c->sNy + 2 * c->OLy;
    MJ_REQUIRE(nx % 2 == 0, "%s: sNx + 2 OLx must be even", fn);
    const int64_t per_member = (int64_t)c->ntiles * c->nr * ny * nx;
    MJ_REQUIRE(c->member_stride3d >= per_member && c->member_stride3d % 2 == 0,
               "%s: member_stride3d must be even and >= ntiles * nr * ny * nx", fn);
    for (const double* p : {trans, vel, mask, tracer, rd, (const double*)out})
        MJ_REQUIRE(((uintptr_t)p & 15) == 0, "%s: arrays must be 16-byte aligned", fn);
    for (const double* p : {trans, vel, mask, tracer})
        MJ_REQUIRE(p != out, "%s: the output must not alias an input", fn);
    const int64_t planes = (int64_t)c->batch * c->ntiles * c->nr;
    MJ_REQUIRE(planes < ((int64_t)1 << 31), "%s: too many planes", fn);
    FluxDev a{c->batch, c->ntiles, c->nr, ny, nx, c->member_stride3d, oneSixth, deltaTloc, rdeep, trans, vel, mask,
              tracer, rd, out};
    const int threads = 256;
    const dim3 grid((unsigned)planes, (unsigned)((ny * nx / 2 + threads - 1) / threads));
    if (dir == 0)
        hipLaunchKernelGGL(dst3_adv_kernel<0>, grid, dim3(threads), 0, as_stream(stream), a);
    else
        hipLaunchKernelGGL(dst3_adv_kernel<1>, grid, dim3(threads), 0, as_stream(stream), a);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}

}  // namespace

extern "C" int mj_dst3_adv_x(const mj_dst3_flux_config* cfg, double oneSixth, const double* deltaTloc,
                             const double* uTrans, const double* uFld, const double* maskLocW, const double* tracer,
                             const double* recip_dxC, const double* recip_deepFacC, double* uT, void* stream) {
    return dst3_adv(0, cfg, oneSixth, deltaTloc, uTrans, uFld, maskLocW, tracer, recip_dxC, recip_deepFacC, uT,
                    stream);
}

extern "C" int mj_dst3_adv_y(const mj_dst3_flux_config* cfg, double oneSixth, const double* deltaTloc,
                             const double* vTrans, const double* vFld, const double* maskLocS, const double* tracer,
                             const double* recip_dyC, const double* recip_deepFacC, double* vT, void* stream) {
    return dst3_adv(1, cfg, oneSixth, deltaTloc, vTrans, vFld, maskLocS, tracer, recip_dyC, recip_deepFacC, vT,
                    stream);
}
