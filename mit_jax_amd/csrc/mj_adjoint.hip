// Reverse-mode pieces of the CG2D implicit derivative (reference core/cg2d.py:292-348).
//
// The reference differentiates x = A^-1 b through lax.custom_linear_solve with the symmetric linear operator
// _matvec(v) = (A / cg2dNorm) exch(where(inner, v, 0)) on the interior: for a cotangent w of x, the transpose solve
// gives lambda = A^-T w (the same tight CG, A is symmetric), the cotangent of b is lambda, and the cotangent of the
// coefficient arrays is minus the transpose of dA -> _matvec(dA, x) applied to lambda:
//   gaW[q] = -(lambda[q] xe[q-1] + lambda[q-1] xe[q]) / cg2dNorm      (aW multiplies x(i-1) at q and x(i+1) at q-1)
//   gaS[q] = -(lambda[q] xe[q-nx] + lambda[q-nx] xe[q]) / cg2dNorm
//   gaC[q] = -(lambda[q] xe[q]) / cg2dNorm
// with xe = exch(where(inner, x, 0)) and lambda zero off the interior. Entries that no interior stencil reads get 0.
#include "mj_common.h"
#include "mj_internal.h"

namespace mj {

__global__ void cg2d_coeff_grad_kernel(const double* xe, const double* lam, const double* norm, double* gaW,
                                       double* gaS, double* gaC, int64_t ms, int T, int ny, int nx) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    const int plane = ny * nx;
    if (q >= (int64_t)T * plane) return;
    const int r = (int)(q % plane), j = r / nx, i = r - j * nx;
    const double* X = xe + m * ms;
    const double* Lm = lam + m * ms;
    const double nm = norm[m];
    const int64_t o = m * ms + q;
    const double lq = Lm[q], xq = X[q];
    const double w = i > 0 ? lq * X[q - 1] + Lm[q - 1] * xq : 0.0;
    const double s = j > 0 ? lq * X[q - nx] + Lm[q - nx] * xq : 0.0;
    gaW[o] = -(w / nm);
    gaS[o] = -(s / nm);
    gaC[o] = -(lq * xq / nm);
}

__global__ void zero_halo_copy_adj_kernel(const double* v, double* w, int64_t ms, int T, int ny, int nx, int OL,
                                          int sNx, int sNy) {
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

extern "C" int mj_cg2d_coeff_grad(const mj_halo_plan* plan_T, const double* x, const double* lam,
                                  const double* cg2dNorm, double* gaW, double* gaS, double* gaC,
                                  const mj_cg2d_config* cfg, double* work, void* stream) {
    MJ_REQUIRE(plan_T && x && lam && cg2dNorm && gaW && gaS && gaC && cfg && work,
               "mj_cg2d_coeff_grad: null argument");
    MJ_REQUIRE(plan_T->nmaps == 1, "mj_cg2d_coeff_grad: plan_T must be a scalar (T) plan");
    MJ_REQUIRE(cfg->batch >= 1 && cfg->ntiles == plan_T->ntiles, "mj_cg2d_coeff_grad: bad batch / tile count");
    const int nx = cfg->sNx + 2 * cfg->OL, ny = cfg->sNy + 2 * cfg->OL;
    MJ_REQUIRE(nx == plan_T->nx && ny == plan_T->ny, "mj_cg2d_coeff_grad: layout does not match the plan");
    const int64_t n = (int64_t)cfg->ntiles * nx * ny;
    MJ_REQUIRE(cfg->member_stride >= n, "mj_cg2d_coeff_grad: member_stride too small");
    hipStream_t st = as_stream(stream);
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)cfg->batch);
    hipLaunchKernelGGL(zero_halo_copy_adj_kernel, grid, dim3(256), 0, st, x, work, cfg->member_stride, cfg->ntiles,
                       ny, nx, cfg->OL, cfg->sNx, cfg->sNy);
    MJ_HIP(hipGetLastError());
    mj_halo_geom g{cfg->batch, cfg->member_stride, (int64_t)nx * ny, 1, 0};
    int rc = mj_halo_exchange(plan_T, work, work, &g, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(cg2d_coeff_grad_kernel, grid, dim3(256), 0, st, (const double*)work, lam, cg2dNorm, gaW, gaS,
                       gaC, cfg->member_stride, cfg->ntiles, ny, nx);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}
