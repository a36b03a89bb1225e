/* mjamd.h — C ABI of the MI355X-native CG2D / GAD DST3 / exch2 halo library (libmjamd.so, gfx950).
 *
 * Drop-in boundary for the ocean time step's two data-parallel cores of mitgcm_jax (reference file:line citations
 * below are into that repository). Every entry point takes plain device pointers, element strides and sizes; all
 * device buffers belong to the caller; the library owns only the plans it creates (halo tables, DST3 pass tables).
 * Every call is asynchronous on the given HIP stream (NULL = the null stream; the one exception is mj_cg2d_solve in
 * MJ_CG_LAUNCH_PHASES mode, which waits for the stream between chunks) and returns 0 (MJ_OK) or a negative
 * code; mj_last_error() then describes the failure. Solver non-convergence is not an error: it is reported per
 * member in mj_cg2d_diag (reference core/cg2d.py:267-268, diag numIters / lastResidual).
 *
 * Layout (reference mitgcm_jax/layout.py:1-54): FP64, C order, [tile, (k,) j, i] with OLx = OLy halos; batched
 * (ensemble) arguments add a leading member axis addressed by an element stride. All arithmetic is FP64 in the
 * reference's literal operation order with no fused multiply-add, except the explicitly mixed-precision solver.
 *
 * Sum orders of the CG2D global dot products (reference core/cg2d.py:141-169):
 *   MJ_SUM_FORTRAN  per tile s = 0; DO j; DO i; s = s + a(i,j)  (strictly sequential), then tiles in order
 *                   0 + tile1 + ... + tile13 (parallel/global_sum.py:20-26). Bitwise the reference "fortran" order.
 *   MJ_SUM_TREE     per row of sNx points and per tile over its sNy row sums: fold onto 64 lanes
 *                   (w[l] = v[l] + v[l+64] where l+64 < n), then butterfly w[l] += w[l^m], m = 32..1; lane 0 is
 *                   the result. Tiles in the same fixed order. The reference's "tile" order is an XLA-chosen
 *                   tree; this one is fixed and independent of the launch geometry, so results are bitwise
 *                   reproducible for any block size and batch.
 */
#ifndef MJAMD_H
#define MJAMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MJ_OK 0
#define MJ_ERR_ARG (-1)   /* bad shape, null pointer, unsupported option */
#define MJ_ERR_HIP (-2)   /* a HIP runtime call failed */
#define MJ_ERR_SYNC (-3)  /* a device-side grid barrier timed out (the launch was abandoned safely) */

#define MJ_SUM_FORTRAN 0
#define MJ_SUM_TREE 1

/* Launch shapes of mj_cg2d_solve (same arithmetic, same bits):
 *   PERSISTENT    one cooperative launch with grid barriers; asynchronous. Best for small batches.
 *   PHASES        one launch per CG phase, one block per (member, tile); the host enqueues poll_iters iterations
 *                 at a time and reads the number of active members between chunks (the call returns once all
 *                 members have finished). Best for large batches.
 *   PHASES_FIXED  as PHASES, but enqueues exactly max_iters iterations (finished members' blocks exit at once)
 *                 and never reads back: asynchronous and capturable in a HIP graph.
 *   AUTO          PERSISTENT below 256 (member, tile) units, PHASES from there. */
#define MJ_CG_LAUNCH_AUTO 0
#define MJ_CG_LAUNCH_PERSISTENT 1
#define MJ_CG_LAUNCH_PHASES 2
#define MJ_CG_LAUNCH_PHASES_FIXED 3

/* Library version (major*10000 + minor*100 + patch). */
int mj_version(void);

/* Copies the last error message of the calling thread into buf (NUL-terminated); returns its full length. */
int mj_last_error(char* buf, size_t n);

/* ======================================================================================= exch2 halo exchange
 * Replaces Exchanger.scalar / vector / exch_xy / exch_uv_xy (parallel/exchange.py:136-142, :170-186): for each
 * written point d of the output component, out[d] = sign[d] * in_comp[d][src[d]], comp 1 = the u-like input,
 * comp 2 = the v-like input; all other points keep their value. The tables are the written points of the probed
 * exch2 maps (mitgcm_jax/data/exch_maps_13x90x90.npz, parallel/exchange.py:72-92), indices flat over
 * [ntiles, ny, nx]. */
typedef struct mj_halo_plan mj_halo_plan;

/* nmaps = 1 (scalar kind: T, 3D, Z, SMs) or 2 (vector kind: u-map then v-map). For each map m: npts[m] written
 * points with dst[m][p], src[m][p] (flat padded indices), comp[m][p] in {1,2}, sign[m][p] in {-1,+1}. */
int mj_halo_plan_create(int nmaps, const int32_t* npts, const int32_t* const* dst, const int32_t* const* src,
                        const int8_t* const* comp, const int8_t* const* sign, int ntiles, int ny, int nx,
                        mj_halo_plan** out);
int mj_halo_plan_destroy(mj_halo_plan* plan);
/* 1 when no written point is the source of a written point (T, 3D, UVs, UVn, UV3s): in-place fills are exact. */
int mj_halo_plan_inplace_ok(const mj_halo_plan* plan);

/* Geometry of the arrays one halo call covers: element (outer o, tile t, inner n, j, i) lives at
 * o*outer_stride + t*tile_stride + n*inner_stride + j*nx + i. [T, ny, nx]: nouter = ninner = 1;
 * [T, K, ny, nx]: ninner = K, inner_stride = ny*nx, tile_stride = K*ny*nx; [B, T, K, ny, nx]: nouter = B. */
typedef struct {
    int64_t nouter, outer_stride, tile_stride, ninner, inner_stride;
} mj_halo_geom;

/* Scalar kind. out may equal in only when mj_halo_plan_inplace_ok(plan). out must already hold the values of the
 * points the exchange does not write (the reference's "own" values). */
int mj_halo_exchange(const mj_halo_plan* plan, const double* in, double* out, const mj_halo_geom* g, void* stream);
/* Vector kind: (u_out, v_out) from (u_in, v_in). Same aliasing rule. */
int mj_halo_exchange_uv(const mj_halo_plan* plan, const double* u_in, const double* v_in, double* u_out,
                        double* v_out, const mj_halo_geom* g, void* stream);
/* Transpose of mj_halo_exchange / _uv (reverse mode; the reference checks this adjoint identity in
 * test_exchange.py:106-118): from the cotangents of the exchanged arrays, the cotangents of the inputs. A point's
 * cotangent is its own where no map writes it, plus sign * the cotangent of every written point that reads it
 * (summed in a fixed order: deterministic). Scalar plans: vbar_out, vbar_in NULL. Every point of the inputs'
 * cotangent arrays covered by the geometry is written; outputs must not alias inputs. */
int mj_halo_exchange_adjoint(const mj_halo_plan* plan, const double* ubar_out, const double* vbar_out,
                             double* ubar_in, double* vbar_in, const mj_halo_geom* g, void* stream);

/* ======================================================================================= CG2D
 * Replaces cg2d_fortran (core/cg2d.py:205-269) and the solve inside cg2d_solve (:351-375): normalise the rhs by
 * cg2dNorm and 1/max|b|, exchange the first guess, PCG with the 5-point operator aW2d/aS2d/aC2d and the
 * 5-point preconditioner pW/pS/pC in the Fortran update order, stop when err_sq < tol^2 or after max_iters
 * iterations, un-normalise. The halos of x are left as cg2d.F leaves them: the exchanged normalised first guess
 * where the T map writes, the first guess itself elsewhere. Runs all members of a batch in one persistent launch
 * with per-member convergence masks. */
typedef struct {
    double firstResidual;  /* sqrt(err_sq) before the first iteration */
    double lastResidual;   /* sqrt(err_sq) at exit */
    double sumRHS;         /* global sum of the normalised rhs */
    double rhsMax;         /* max |cg2d_b * cg2dNorm| */
    int32_t numIters;      /* iterations done */
    int32_t converged;     /* 1 if err_sq < tol^2 at exit */
    int32_t outerIters;    /* FP64 residual replacements (mixed precision), 0 for the plain solve */
    int32_t launches;      /* kernel launches of the iteration loop (PHASES modes; the same value in every member's
                              record, written when the loop ends); 0 for the persistent launch */
} mj_cg2d_diag;

typedef struct {
    const double *aW2d, *aS2d, *aC2d, *pW, *pS, *pC; /* [T, ny, nx] each (one per operator slot) */
    int64_t member_stride;                           /* elements between operator slots; 0 = one shared slot */
    const int32_t* op_index;                         /* device [B]: slot of each member, or NULL (slot = member) */
} mj_cg2d_ops;

typedef struct {
    int batch;           /* members B >= 1 */
    int ntiles, sNx, sNy, OL; /* 13, 90, 90, 4 (sNx, sNy in 64..128) */
    int64_t member_stride;    /* elements between members of b, x0, x (>= ntiles*ny*nx) */
    double tolerance;    /* cg2dTolerance (normalised target) */
    int max_iters;       /* cg2dMaxIters */
    int sum_order;       /* MJ_SUM_FORTRAN or MJ_SUM_TREE */
    int block_threads;   /* 0 = launch-mode default (1024; the mixed solve: see mj_cg2d_solve_mixed); else a
                            multiple of 64 up to it */
    int grid_blocks;     /* persistent launch: 0 = as many co-resident blocks as useful */
    int launch_mode;     /* MJ_CG_LAUNCH_*; results are bitwise the same in every mode */
    int poll_iters;      /* MJ_CG_LAUNCH_PHASES: iterations enqueued between reads of the active count (0 = 32;
                            the mixed solve: 64) */
} mj_cg2d_config;

/* Device workspace the solve needs (bytes), for the given configuration. */
size_t mj_cg2d_workspace_bytes(const mj_cg2d_config* cfg);

/* b, x0: [B][T, ny, nx] (member_stride apart); x: output, same shape (may not alias b or x0); cg2dNorm: device
 * [B]; diag: device [B]; workspace: device, >= mj_cg2d_workspace_bytes(cfg), 256-byte aligned. plan_T: the T
 * (scalar) halo plan. */
int mj_cg2d_solve(const mj_halo_plan* plan_T, const mj_cg2d_ops* ops, const double* b, const double* x0, double* x,
                  const double* cg2dNorm, const mj_cg2d_config* cfg, void* workspace, size_t workspace_bytes,
                  mj_cg2d_diag* diag, void* stream);

/* Allocation-free variant of mj_cg2d_solve. It has no counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * arguments and results are those of mj_cg2d_solve without the workspace: the solve's scratch is a device array the
 * library holds per device (MJ_CG2D_NOALLOC_SCRATCH_BYTES, kept until the process ends), and the call makes no
 * allocation. A configuration whose mj_cg2d_workspace_bytes exceeds that size is rejected (MJ_ERR_ARG; at 13 tiles
 * of 90 x 90 with OL 4 that is above 11 members in the MJ_SUM_TREE order, above 8 in MJ_SUM_FORTRAN). The T plan
 * must already hold its CG tables, which the plan's first mj_cg2d_solve or mj_cg2d_solve_mixed call builds (with
 * allocations); otherwise MJ_ERR_ARG. Every call on one device uses the same array and leaves its contents there:
 * calls that overlap on one device (other streams, other threads) race and must be serialised by the caller. */
#define MJ_CG2D_NOALLOC_SCRATCH_BYTES (64u << 20)
int mj_cg2d_solve_noalloc(const mj_halo_plan* plan_T, const mj_cg2d_ops* ops, const double* b, const double* x0,
                          double* x, const double* cg2dNorm, const mj_cg2d_config* cfg, mj_cg2d_diag* diag,
                          void* stream);

/* The linear operator of the implicit derivative (core/cg2d.py:292-298 _matvec): y = (A / cg2dNorm) v on the
 * interior, 0 on halos, with v exchanged after zeroing its halos. Batched like mj_cg2d_solve. work: device
 * scratch of B*member_stride doubles. */
int mj_cg2d_matvec(const mj_halo_plan* plan_T, const mj_cg2d_ops* ops, const double* v, double* y,
                   const double* cg2dNorm, const mj_cg2d_config* cfg, double* work, void* stream);

/* Cotangent of the operator coefficients in the implicit derivative of the solve (core/cg2d.py:326-348: the
 * transpose of dA -> -_matvec(dA, x), applied to lambda = A^-T w): gaW[q] = -(lam[q] xe[q-1] + lam[q-1] xe[q]) /
 * cg2dNorm, gaS likewise with q-nx, gaC[q] = -lam[q] xe[q] / cg2dNorm, where xe = exch(where(inner, x, 0)) and
 * lam is zero off the interior. x, lam, gaW, gaS, gaC: [B][T, ny, nx] (cfg->member_stride apart); cg2dNorm: device
 * [B]; work: device scratch of B*member_stride doubles. Only cfg's batch/layout fields are read. */
int mj_cg2d_coeff_grad(const mj_halo_plan* plan_T, const double* x, const double* lam, const double* cg2dNorm,
                       double* gaW, double* gaS, double* gaC, const mj_cg2d_config* cfg, double* work, void* stream);

/* ---- mixed-precision CG2D (no single counterpart in the reference: it solves the system of cg2d_fortran,
 * core/cg2d.py:205-269, to the same normalised stopping rule err < tolerance^2, by FP32 PCG whose recursive
 * residual is replaced by the FP64 true residual from time to time: at each replacement the FP32 solution
 * increment goes into the FP64 solution, the true residual is formed in FP64 and the PCG goes on from the same
 * search direction). True residuals, dot products (MJ_SUM_TREE order), alpha/beta and the solution update are
 * FP64; the operator/preconditioner copies, inner vectors and stencils are FP32. diag: numIters = total inner
 * iterations, outerIters = true residuals that replaced the recursive one (the first included), lastResidual = the
 * last FP64 true residual (normalised). */
typedef struct {
    int op_slots;            /* number of operator slots when ops.op_index is set (else ignored) */
    double inner_tolerance;  /* replace when the recursive residual norm has fallen by this factor since the last
                                replacement (e.g. 1e-4), or below the target */
    int inner_max_iters;     /* or after this many iterations since the last replacement */
    int max_outer;           /* replacements */
} mj_cg2d_mixed_config;

/* Workspace of mj_cg2d_solve_mixed; ops_member_stride as in mj_cg2d_ops (0 = one shared operator). */
size_t mj_cg2d_mixed_workspace_bytes(const mj_cg2d_config* cfg, const mj_cg2d_mixed_config* mcfg,
                                     int64_t ops_member_stride);

/* Arguments as mj_cg2d_solve. cfg: sum_order must be MJ_SUM_TREE; launch_mode AUTO (= PHASES), PHASES or
 * PHASES_FIXED; max_iters caps the total inner iterations; block_threads 0 = 1024 while batch * ntiles blocks fit
 * one per CU, else 512; poll_iters 0 = 64. */
int mj_cg2d_solve_mixed(const mj_halo_plan* plan_T, const mj_cg2d_ops* ops, const double* b, const double* x0,
                        double* x, const double* cg2dNorm, const mj_cg2d_config* cfg,
                        const mj_cg2d_mixed_config* mcfg, void* workspace, size_t workspace_bytes,
                        mj_cg2d_diag* diag, void* stream);

/* ======================================================================================= CG2D operator and rhs
 * The producers either side of the solve in SOLVE_FOR_PRESSURE (core/solve_for_pressure.py): the operator and
 * preconditioner of UPDATE_CG2D (:29-65) and the right-hand side and first guess of cg2d_rhs (:68-106). Both are
 * column sums over k in the reference's order (aW/aS: k = 1..Nr; the divergence of cg2d_b: k = Nr..1), FP64,
 * bitwise equal to the literal loops. */
typedef struct {
    /* FreeSurfParams fields the two producers read (core/free_surface.py:33-52) */
    double deltaTMom, deltaTFreeSurf, freeSurfFac, implicSurfPress, implicDiv2DFlow, mass2rUnit;
    int useRealFreshWaterFlux;
} mj_fs_params;

typedef struct {
    /* 2-D [T, ny, nx], shared by all members */
    const double *dyG, *dxG, *recip_dxC, *recip_dyC, *rA, *maskInC;
    const int32_t* kSurfC; /* 1-based first wet level of each column, nr + 1 when dry (k_surf of maskC); cg2d_rhs */
    /* per level: drF [nr]; deepFacC, rhoFacC [nr] and deepFac2F [nr + 1] (NULL = all ones, the V4r4 values of
       vertical_factors, core/free_surface.py:121-125) */
    const double *drF, *deepFacC, *rhoFacC, *deepFac2F;
} mj_sfp_grid;

typedef struct {
    int batch;               /* members */
    int ntiles, sNx, sNy, OL; /* tile layout of the T plan */
    int nr;                  /* levels of the 3-D arrays */
    int64_t member_stride3d; /* elements between members' 3-D arrays ([T, nr, ny, nx]) */
    int64_t member_stride2d; /* elements between members' 2-D arrays ([T, ny, nx]) */
} mj_sfp_config;

/* INI_CG2D's normalisation (core/cg2d.py:108-134 ini_cg2d_norm / _ini_cg2d_norm; ini_cg2d.F:90-135). Per member m,
 * on the interior points (i, j = 1..sNx, 1..sNy): aW = sum over k = 1..nr in order of
 * (fac ((dyG drF(k)) hFacW)) recip_dxC and aS likewise with dxG, hFacS, recip_dyC, fac = implicSurfPress *
 * implicDiv2DFlow; cg2dNorm[m] = 1 / max(|aW|, |aS|) over the member's tiles (1 when that max is 0). The maximum is
 * order-free, so the result is deterministic. hFacW, hFacS: [B][T, nr, ny, nx] (member_stride3d apart; the
 * reference passes INITIALISE_FIXED's h0FacW/S); grid: dyG, dxG, recip_dxC, recip_dyC, drF; cg2dNorm: device [B],
 * also used as scratch during the call. Only cfg's batch/layout/nr fields are read. */
int mj_ini_cg2d_norm(const mj_sfp_grid* grid, const mj_sfp_config* cfg, double implicSurfPress,
                     double implicDiv2DFlow, const double* hFacW, const double* hFacS, double* cg2dNorm,
                     void* stream);

/* UPDATE_CG2D (core/solve_for_pressure.py:29-65). Per member m: aW2d = ((sum_k dyG drF(k) hFacW recip_dxC) *
 * cg2dNorm[m]) * implicSurfPress * implicDiv2DFlow on j, i = 1..sNy+1, 1..sNx+1 and 0 elsewhere (aS2d likewise);
 * aC2d = -(aW(i) + aW(i+1) + aS(j) + aS(j+1) + freeSurfFac cg2dNorm[m] recip_Bo rA / deltaTMom / deltaTFreeSurf) on
 * the interior, 0 elsewhere, then exchanged with plan_T; pC, pW, pS rewritten on j, i = 1..sNy+1, 1..sNx+1 only
 * (pC = 1/aC, 1 where aC = 0; pW = -aW / (cg2dpcOffDFac (aC(i) + aC(i-1)))^2, 0 where the sum is 0; pS likewise):
 * their other points keep the values they hold. hFacW, hFacS: [B][T, nr, ny, nx]; recip_Bo: [T, ny, nx] shared;
 * cg2dNorm: device [B]; outputs [B][T, ny, nx] (member_stride2d apart), none aliasing an input. */
int mj_update_cg2d(const mj_halo_plan* plan_T, const mj_sfp_grid* grid, const mj_fs_params* p,
                   const mj_sfp_config* cfg, const double* cg2dNorm, double cg2dpcOffDFac, const double* hFacW,
                   const double* hFacS, const double* recip_Bo, double* aW2d, double* aS2d, double* aC2d, double* pW,
                   double* pS, double* pC, void* stream);

/* Reverse mode of mj_update_cg2d in hFacW, hFacS (JAX's transpose of update_cg2d in the reference; the operator is
 * affine in them, and the preconditioner outputs carry no cotangent since CG2D's derivative does not use them,
 * core/cg2d.py:307-348): aC2d_bar goes back through the transposed T exchange to the interior, then to aW (i, i+1) and
 * aS (j, j+1), then to every level of hFacW_bar, hFacS_bar ([B][T, nr, ny, nx], every point written). aW2d_bar,
 * aS2d_bar, aC2d_bar: [B][T, ny, nx]; work: device, B * member_stride2d doubles (T * ny * nx for one member). */
int mj_update_cg2d_adjoint(const mj_halo_plan* plan_T, const mj_sfp_grid* grid, const mj_fs_params* p,
                           const mj_sfp_config* cfg, const double* cg2dNorm, const double* aW2d_bar,
                           const double* aS2d_bar, const double* aC2d_bar, double* hFacW_bar, double* hFacS_bar,
                           double* work, void* stream);

/* cg2d_rhs (core/solve_for_pressure.py:68-106; solve_for_pressure.F:124-255 with CALC_DIV_GHAT): cg2d_x = Bo_surf *
 * etaN on the whole array; cg2d_b = 0 off the interior and on the interior
 *   b = useRealFreshWaterFlux ? freeSurfFac mass2rUnit implicDiv2DFlow rA EmPmR / deltaTMom maskInC : 0
 *   for k = nr..1: b = b + pfu(i+1) - pfu(i);  b = b + pfv(j+1) - pfv(j)
 *     pfu = dyG deepFacC(k) drF(k) hFacW rhoFacC(k) gU / deltaTMom   (pfv with dxG, hFacS, gV)
 *   b = b - freeSurfFac rA deepFac2F(kSurfC) / deltaTMom / deltaTFreeSurf etaH
 * gU, gV, hFacW, hFacS: [B][T, nr, ny, nx]; EmPmR, etaN, etaH, cg2d_b, cg2d_x: [B][T, ny, nx]; Bo_surf: [T, ny,
 * nx] shared; EmPmR may be NULL when useRealFreshWaterFlux is 0. Outputs must not alias inputs. */
int mj_cg2d_rhs(const mj_sfp_grid* grid, const mj_fs_params* p, const mj_sfp_config* cfg, const double* gU,
                const double* gV, const double* hFacW, const double* hFacS, const double* EmPmR, const double* etaN,
                const double* etaH, const double* Bo_surf, double* cg2d_b, double* cg2d_x, void* stream);

/* Reverse mode of mj_cg2d_rhs (JAX's transpose of cg2d_rhs in the reference; no counterpart file): from the
 * cotangents of cg2d_b (its interior) and cg2d_x (NULL: none), the cotangents of gU, gV, hFacW, hFacS ([B][T, nr, ny,
 * nx], every point written), EmPmR, etaH (interior, 0 elsewhere) and etaN ([B][T, ny, nx]). The face cotangent of
 * pfu(j, i) is b_bar(j, i-1) - b_bar(j, i) (b_bar taken as 0 off the interior), pfv alike in j. deepFacC and rhoFacC
 * must be NULL (the V4r4 ones). Outputs must not alias inputs or each other. */
int mj_cg2d_rhs_adjoint(const mj_sfp_grid* grid, const mj_fs_params* p, const mj_sfp_config* cfg, const double* gU,
                        const double* gV, const double* hFacW, const double* hFacS, const double* Bo_surf,
                        const double* cg2d_b_bar, const double* cg2d_x_bar, double* gU_bar, double* gV_bar,
                        double* hFacW_bar, double* hFacS_bar, double* EmPmR_bar, double* etaN_bar, double* etaH_bar,
                        void* stream);

/* ======================================================================================= GAD DST3
 * Replaces gad_advection / advect_tiles (pkgs/gad.py:460-560) with dst3_adv_x/y (:409-453): the 3-pass
 * multi-dimensional DST3 on the LLC cubed sphere with per-facet sweep order, overlap/interior update regions and
 * FILL_CS_CORNER_TR_RL / _UV_RS corner fills; gTracer = (localTij - tracer)/deltaTLev(k) on the whole tile array.
 * With wFld given, a vertical DST3 flux (R) and its divergence are added (not in the reference, which runs the
 * vertical implicitly). */
typedef struct mj_gad_plan mj_gad_plan;

/* Topology per tile (reference TileTopology, gad.py:75-106): face[t] in 1..6, edges[t*4 + {0,1,2,3}] = N, S,
 * E, W facet-edge flags (0/1). */
int mj_gad_plan_create(int ntiles, const int32_t* face, const uint8_t* edges, int sNx, int sNy, int OLx, int OLy,
                       mj_gad_plan** out);
int mj_gad_plan_destroy(mj_gad_plan* plan);
/* Rasterises the plan's update regions: masks[((pass*2 + dir)*ntiles + t)*ny*nx + j*nx + i] = 1 where the
 * X (dir 0) / Y (dir 1) update of pass 0..2 writes (host memory, 6*ntiles*ny*nx bytes); fills[(pass*4+f)*ntiles + t]
 * = fill4dir (0 none) of f = {X before, X after, Y before, Y after}. For checking the plan against the tables. */
int mj_gad_plan_describe(const mj_gad_plan* plan, uint8_t* masks, int32_t* fills);
/* The same rasterisation straight from a topology, on the host only (no device, no plan object): the per-facet pass
 * table of gad_advection.F:333-351 and the update regions / fill calls of :376-774 (reference gad.py:217-383). */
int mj_gad_tables_host(int ntiles, const int32_t* face, const uint8_t* edges, int sNx, int sNy, int OLx, int OLy,
                       uint8_t* masks, int32_t* fills);

typedef struct {
    /* 2-D [T, ny, nx], shared by all members */
    const double *dyG, *dxG, *recip_dxC, *recip_dyC, *recip_rA, *rA, *maskInC;
    /* 3-D [T, K, ny, nx] (member_stride3d apart when per-member) */
    const double *maskW, *maskS, *maskC;
    /* per level [K]: deltaTLev, drF, recip_drF, rhoFacC, recip_rhoFacC, deepFacC, recip_deepFacC, recip_deepFac2C,
       recip_drC (interface k, used by the R flux only) */
    const double *dTlev, *drF, *recip_drF, *rhoFacC, *recip_rhoFacC, *deepFacC, *recip_deepFacC, *recip_deepFac2C,
        *recip_drC;
    double oneSixth;
} mj_gad_grid;

typedef struct {
    int batch;               /* members */
    int nr;                  /* K levels per member array */
    int64_t member_stride3d; /* elements between members' 3-D arrays (u, v, w, tracer, hFac*, gTracer) */
    int64_t mask_member_stride; /* elements between mask slots of maskW/maskS/maskC; 0 = one shared slot */
    const int32_t* nlev;     /* device [batch] active levels per member (levels >= nlev untouched), or NULL */
    const int32_t* mask_index; /* device [batch]: mask slot of each member (members of one level class share
                                  their grid masks), or NULL: slot = member. Ignored when mask_member_stride is 0 */
    int levels_per_block;    /* consecutive levels of one (member, tile) per workgroup of the horizontal passes
                                (0 = 1); a launch shape: the results are the same for every value */
    int mask_slots;          /* number of mask slots when mask_index is set (else ignored) */
} mj_gad_config;

/* Workspace of mj_gad_dst3 (with_w: the R flux runs): tables built once per call (the maskLocW / maskLocS bit planes
 * of every (mask slot, tile, level), the 2-D metrics in the pairs the sweeps load together) and with the R flux the
 * localTij of the batch. */
size_t mj_gad_dst3_workspace_bytes(const mj_gad_config* cfg, int ntiles, int with_w);

/* uFld, vFld, tracer, hFacW, hFacS, recip_hFacC, gTracer: [B][T, K, ny, nx]; wFld: same or NULL (no R flux).
 * work: device scratch (256-byte aligned, not overlapping gTracer) of mj_gad_dst3_workspace_bytes. */
int mj_gad_dst3(const mj_gad_plan* plan, const mj_gad_grid* grid, const mj_gad_config* cfg, const double* uFld,
                const double* vFld, const double* wFld, const double* tracer, const double* hFacW,
                const double* hFacS, const double* recip_hFacC, double* gTracer, void* work, size_t work_bytes,
                void* stream);

/* ---- the DST3 interface fluxes alone, for every (member, tile, level) plane (the reference's flux kernels; the
 * fused mj_gad_dst3 evaluates the same values inside its sweeps). */
typedef struct {
    int batch, ntiles, nr;   /* members, tiles, levels of the 3-D arrays */
    int sNx, sNy, OLx, OLy;  /* tile interior and halos (OLx, OLy >= 2): ny = sNy + 2 OLy, nx = sNx + 2 OLx (even) */
    int64_t member_stride3d; /* elements between members' 3-D arrays (even, >= ntiles * nr * ny * nx) */
} mj_dst3_flux_config;

/* Replaces dst3_adv_x (pkgs/gad.py:409-430; gad_dst3_adv_x.F:74-121 with calcCFL): on every row and at array
 * columns i = 2 .. nx-2, with t = tracer, m = maskLocW, U = uTrans,
 *   Rjp = (t(i+1) - t(i)) m(i+1), Rj = (t(i) - t(i-1)) m(i), Rjm = (t(i-1) - t(i-2)) m(i-1),
 *   c = |uFld deltaTloc(k) recip_dxC recip_deepFacC(k)|, d0 = (2 - c)(1 - c) oneSixth, d1 = (1 - c c) oneSixth,
 *   uT = 0.5 (U + |U|)(t(i-1) + (d0 Rj + d1 Rjm)) + 0.5 (U - |U|)(t(i) - (d0 Rj + d1 Rjp)),
 * and uT = 0 at every other point. uTrans, uFld, maskLocW, tracer, uT: [B][T, nr, ny, nx] (member_stride3d apart);
 * recip_dxC: [T, ny, nx] shared; deltaTloc, recip_deepFacC: [nr]. All 16-byte aligned; uT aliases no input. */
int mj_dst3_adv_x(const mj_dst3_flux_config* cfg, double oneSixth, const double* deltaTloc, const double* uTrans,
                  const double* uFld, const double* maskLocW, const double* tracer, const double* recip_dxC,
                  const double* recip_deepFacC, double* uT, void* stream);
/* Replaces dst3_adv_y (pkgs/gad.py:433-453; gad_dst3_adv_y.F:73-120): the same in j (rows j = 2 .. ny-2, every
 * column) with vTrans, vFld, maskLocS, recip_dyC. */
int mj_dst3_adv_y(const mj_dst3_flux_config* cfg, double oneSixth, const double* deltaTloc, const double* vTrans,
                  const double* vFld, const double* maskLocS, const double* tracer, const double* recip_dyC,
                  const double* recip_deepFacC, double* vT, void* stream);

/* Reverse mode of mj_gad_dst3 (the reference's advect_tiles, pkgs/gad.py:460-539, differentiated as
 * test_gad.py:297-346 does; the R flux, which the reference does not have, differentiated the same way): for the
 * cotangent gTracer_bar of gTracer, writes the cotangents of tracer, uFld and vFld, and with wFld that of wFld
 * (every padded point of [B][T, nr, ny, nx]; levels at or beyond a member's nlev get 0; w at the surface
 * interface, which the R flux does not read, gets 0). The grid, hFac and mask inputs are constants. With wFld the
 * horizontal passes are replayed forward once (localTij into work), the vertical sweep is reversed per column,
 * then the horizontal sweeps per (member, tile, level). work: device, 256-byte aligned, of
 * mj_gad_dst3_adjoint_workspace_bytes(cfg, ntiles, wFld != NULL) bytes (the mask bit planes as in mj_gad_dst3, and the
 * replay scratch). Outputs must not alias inputs or each other. */
size_t mj_gad_dst3_adjoint_workspace_bytes(const mj_gad_config* cfg, int ntiles, int with_w);
int mj_gad_dst3_adjoint(const mj_gad_plan* plan, const mj_gad_grid* grid, const mj_gad_config* cfg,
                        const double* uFld, const double* vFld, const double* wFld, const double* tracer,
                        const double* hFacW, const double* hFacS, const double* recip_hFacC,
                        const double* gTracer_bar, double* tracer_bar, double* uFld_bar, double* vFld_bar,
                        double* wFld_bar, double* work, size_t work_bytes, void* stream);

/* ======================================================================================= GAD_IMPLICIT_R
 * Replaces gad_implicit_r (core/tracer_implicit.py:165-173) with gad_implicit_r_matrix (:34-114) and
 * solve_pentadiagonal (:116-162): per water column, the implicit vertical diffusion (kappaRX) and, with
 * implicitAdvection, the 3rd-order upwind implicit vertical advection (wFld) as a pentadiagonal system on the tile
 * interior (halo columns: identity rows), solved by the elimination of SOLVE_DIAGONAL_KINNER with its zero-pivot
 * guard. Bitwise the reference's expressions. With nr = 1 the tracer is returned unchanged. */
typedef struct {
    const double* maskC;          /* [T, nr, ny, nx] (mask_member_stride apart per member; 0 = shared) */
    const double *rA, *recip_rA;  /* [T, ny, nx], shared (advection only) */
    const double *dTtracerLev, *recip_drF, *recip_drC; /* [nr] */
    double rkSign;                /* -1 for the ocean's z coordinate */
} mj_impr_grid;

typedef struct {
    int batch, ntiles, sNx, sNy, OL, nr;
    int64_t member_stride3d;      /* elements between members' kappaRX, recip_hFac, wFld, gTracer, out */
    int64_t mask_member_stride;   /* elements between members' maskC; 0 = one maskC for all */
    int implicitAdvection;        /* 1: U3 implicit vertical advection (wFld, rA, recip_rA needed) */
} mj_impr_config;

/* kappaRX, recip_hFac, wFld, gTracer (the tracer after the explicit step), out: [B][T, nr, ny, nx]; nr <= 150.
 * out may equal gTracer (in place), no other input. */
int mj_gad_implicit_r(const mj_impr_grid* grid, const mj_impr_config* cfg, const double* kappaRX,
                      const double* recip_hFac, const double* wFld, const double* gTracer, double* out, void* stream);

/* Reverse mode of mj_gad_implicit_r (the reference differentiates the literal solve with JAX): for u the forward
 * result and ubar its cotangent, ybar = A^-T ubar (the same guarded elimination on the transposed diagonals) and the
 * cotangents of kappaRX and (with implicitAdvection) wFld; halo columns: ybar = ubar, no parameter cotangent. nr >= 2.
 * Outputs [B][T, nr, ny, nx], none aliasing an input; wFld_bar may be NULL without advection. */
int mj_gad_implicit_r_adjoint(const mj_impr_grid* grid, const mj_impr_config* cfg, const double* kappaRX,
                              const double* recip_hFac, const double* wFld, const double* u, const double* ubar,
                              double* ybar, double* kappaRX_bar, double* wFld_bar, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MJAMD_H */
