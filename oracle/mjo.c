/* CPU oracle of the CG2D / GAD DST3 / exch2 hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference algorithms (mitgcm_jax, cited file:line per function), compiled with
 * -O2 -ffp-contract=off so every expression is evaluated in its written order without fused multiply-adds, as
 * the reference's XLA build does (conftest.py:25-35: no FMA, no algebraic simplification). Only the tests, the
 * smoke check and the CPU-baseline leg of the benchmark load it; the product never does.
 *
 * Arrays are the reference layout: FP64, C order, [tile, (k,) j, i] with halos (layout.py:1-54). Flat "point"
 * indices p = (t*ny + j)*nx + i address one padded 2-D level of all tiles.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* =================================================================================================== exchange */
/* Exchanger._apply (parallel/exchange.py:136-142) on [T, nlead, ny, nx] arrays:
 * out = where(comp == 1, fa[src], own); out = where(comp == 2, fb[src], out); out *= sign. */
EXPORT void mjo_exchange(const int32_t* src, const int8_t* comp, const int8_t* sign, int T, int plane, int nlead,
                         const double* own, const double* fa, const double* fb, double* out) {
    const int64_t npts = (int64_t)T * plane;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < npts; ++p) {
        const int t = (int)(p / plane), r = (int)(p % plane);
        const int s = src[p], st = s / plane, sr = s % plane;
        for (int k = 0; k < nlead; ++k) {
            const int64_t o = ((int64_t)t * nlead + k) * plane + r;
            const int64_t so = ((int64_t)st * nlead + k) * plane + sr;
            double v = own[o];
            if (comp[p] == 1) v = fa[so];
            if (comp[p] == 2) v = fb[so];
            out[o] = v * (double)sign[p];
        }
    }
}

/* =================================================================================================== sums */
/* The fixed tree of the GPU "tile" order: fold n (64..128) values onto 64 lanes, then a butterfly
 * w[l] = w[l] + w[l ^ m] for m = 32..1; the result is lane 0. */
static double tree_n(const double* v, int n) {
    double w[64], nw[64];
    for (int l = 0; l < 64; ++l) w[l] = (l + 64 < n) ? v[l] + v[l + 64] : v[l];
    for (int m = 32; m >= 1; m >>= 1) {
        for (int l = 0; l < 64; ++l) nw[l] = w[l] + w[l ^ m];
        memcpy(w, nw, sizeof(w));
    }
    return w[0];
}

EXPORT double mjo_tree_n(const double* v, int n) { return tree_n(v, n); }

/* tile_sum (core/cg2d.py:141-161) of an interior field a[T][sNy][sNx] -> per-tile partials.
 * order 0 "fortran": s = 0; DO j; DO i; s = s + a (sequential); order 1: the GPU tree (rows, then tile). */
static void tile_sums(const double* a, int T, int sNy, int sNx, int order, double* part) {
#pragma omp parallel for schedule(static)
    for (int t = 0; t < T; ++t) {
        const double* at = a + (int64_t)t * sNy * sNx;
        if (order == 0) {
            double s = 0.0;
            for (int j = 0; j < sNy; ++j)
                for (int i = 0; i < sNx; ++i) s = s + at[j * sNx + i];
            part[t] = s;
        } else {
            double rows[128];
            for (int j = 0; j < sNy; ++j) rows[j] = tree_n(at + j * sNx, sNx);
            part[t] = tree_n(rows, sNy);
        }
    }
}

/* global_sum_tile (parallel/global_sum.py:20-26): 0 + tile 1 + tile 2 + ... in order. */
static double gsum(const double* a, int T, int sNy, int sNx, int order) {
    double part[256];
    tile_sums(a, T, sNy, sNx, order, part);
    double s = 0.0;
    for (int t = 0; t < T; ++t) s = s + part[t];
    return s;
}

EXPORT double mjo_global_sum(const double* a, int T, int sNy, int sNx, int order) {
    return gsum(a, T, sNy, sNx, order);
}

/* =================================================================================================== CG2D */
typedef struct {
    int T, ny, nx, OL, sNx, sNy;
} Lay;

#define PIX(L, t, j, i) (((int64_t)(t) * (L).ny + (j)) * (L).nx + (i))

/* apply_operator (core/cg2d.py:181-188) at interior point (t, j, i) of full arrays. */
static inline double op_A(const double* aW, const double* aS, const double* aC, const double* x, int64_t p, int nx) {
    return aW[p] * x[p - 1] + aW[p + 1] * x[p + 1] + aS[p] * x[p - nx] + aS[p + nx] * x[p + nx] + aC[p] * x[p];
}
/* apply_preconditioner (core/cg2d.py:191-198). */
static inline double op_M(const double* pW, const double* pS, const double* pC, const double* r, int64_t p, int nx) {
    return pC[p] * r[p] + pW[p] * r[p - 1] + pW[p + 1] * r[p + 1] + pS[p] * r[p - nx] + pS[p + nx] * r[p + nx];
}

static void exch_T(const int32_t* src, const int8_t* comp, const int8_t* sign, Lay L, double* a, double* tmp) {
    const int plane = L.ny * L.nx;
    mjo_exchange(src, comp, sign, L.T, plane, 1, a, a, a, tmp);
    memcpy(a, tmp, sizeof(double) * (size_t)L.T * plane);
}

/* cg2d_fortran (core/cg2d.py:205-269). maps: the full T map (src, comp, sign over T*ny*nx points).
 * diag[0..4] = firstResidual, lastResidual, sumRHS, rhsMax, numIters. Returns 0. */
EXPORT int mjo_cg2d(const int32_t* Tsrc, const int8_t* Tcomp, const int8_t* Tsign, int T, int sNx, int sNy, int OL,
                    const double* aW, const double* aS, const double* aC, const double* pW, const double* pS,
                    const double* pC, const double* cg2d_b, const double* cg2d_x, double* x_out, double cg2dNorm,
                    double tolerance, int max_iters, int sum_order, double* diag) {
    Lay L = {T, sNy + 2 * OL, sNx + 2 * OL, OL, sNx, sNy};
    const int64_t N = (int64_t)T * L.ny * L.nx, Ni = (int64_t)T * sNy * sNx;
    double* x = (double*)malloc(sizeof(double) * N);
    double* r = (double*)calloc(N, sizeof(double));
    double* s = (double*)calloc(N, sizeof(double));
    double* tmp = (double*)malloc(sizeof(double) * N);
    double* b = (double*)malloc(sizeof(double) * Ni);
    double* q = (double*)malloc(sizeof(double) * Ni);
    double* w = (double*)malloc(sizeof(double) * Ni);
    double* ri = (double*)malloc(sizeof(double) * Ni);
    double* si = (double*)malloc(sizeof(double) * Ni);
    const double tol_sq = tolerance * tolerance; /* cg2d.F:111 */
    double eta_qrNM1 = 1.0;                      /* cg2d.F:113 */
#define II(t, j, i) (((int64_t)(t) * sNy + (j)) * sNx + (i))
#define LOOP_I for (int t = 0; t < T; ++t) for (int j = 0; j < sNy; ++j) for (int i = 0; i < sNx; ++i)
    double rhsMax = 0.0;
    LOOP_I {
        b[II(t, j, i)] = cg2d_b[PIX(L, t, j + OL, i + OL)] * cg2dNorm; /* cg2d.F:121 */
        const double a = fabs(b[II(t, j, i)]);
        if (a > rhsMax) rhsMax = a; /* _GLOBAL_MAX_RL: order-free */
    }
    const double rhsNorm = rhsMax != 0.0 ? 1.0 / rhsMax : 1.0; /* cg2d.F:131-132 */
    LOOP_I b[II(t, j, i)] = b[II(t, j, i)] * rhsNorm;          /* cg2d.F:137 */
    memcpy(x, cg2d_x, sizeof(double) * N);
    LOOP_I x[PIX(L, t, j + OL, i + OL)] = cg2d_x[PIX(L, t, j + OL, i + OL)] * rhsNorm; /* cg2d.F:138 */
    exch_T(Tsrc, Tcomp, Tsign, L, x, tmp);                                           /* cg2d.F:147 */
    /* s = 0 (cg2d.F:159-163); r_i = b - A x (:171-177) */
#pragma omp parallel for schedule(static)
    for (int t = 0; t < T; ++t)
        for (int j = 0; j < sNy; ++j)
            for (int i = 0; i < sNx; ++i) {
                const int64_t p = PIX(L, t, j + OL, i + OL);
                ri[II(t, j, i)] = b[II(t, j, i)] - op_A(aW, aS, aC, x, p, L.nx);
            }
    LOOP_I w[II(t, j, i)] = ri[II(t, j, i)] * ri[II(t, j, i)];
    double err_sq = gsum(w, T, sNy, sNx, sum_order); /* cg2d.F:181-194 */
    const double sumRHS = gsum(b, T, sNy, sNx, sum_order);
    memset(r, 0, sizeof(double) * N);
    LOOP_I r[PIX(L, t, j + OL, i + OL)] = ri[II(t, j, i)];
    exch_T(Tsrc, Tcomp, Tsign, L, r, tmp); /* cg2d.F:189 */
    const double firstResidual = sqrt(err_sq);
    int it = 0;
    while (it < max_iters && !(err_sq < tol_sq)) { /* cg2d.F:213, :346 */
#pragma omp parallel for schedule(static)
        for (int t = 0; t < T; ++t)
            for (int j = 0; j < sNy; ++j)
                for (int i = 0; i < sNx; ++i) {
                    const int64_t p = PIX(L, t, j + OL, i + OL);
                    q[II(t, j, i)] = op_M(pW, pS, pC, r, p, L.nx); /* cg2d.F:228-233 */
                    w[II(t, j, i)] = q[II(t, j, i)] * r[p];
                }
        const double eta_qrN = gsum(w, T, sNy, sNx, sum_order); /* cg2d.F:241-252 */
        const double cgBeta = eta_qrN / eta_qrNM1;              /* cg2d.F:254 */
        eta_qrNM1 = eta_qrN;
        LOOP_I {
            const int64_t p = PIX(L, t, j + OL, i + OL);
            si[II(t, j, i)] = q[II(t, j, i)] + cgBeta * s[p]; /* cg2d.F:265-266 */
            s[p] = si[II(t, j, i)];
        }
        exch_T(Tsrc, Tcomp, Tsign, L, s, tmp); /* cg2d.F:273 */
#pragma omp parallel for schedule(static)
        for (int t = 0; t < T; ++t)
            for (int j = 0; j < sNy; ++j)
                for (int i = 0; i < sNx; ++i) {
                    const int64_t p = PIX(L, t, j + OL, i + OL);
                    q[II(t, j, i)] = op_A(aW, aS, aC, s, p, L.nx); /* cg2d.F:285-290 */
                    w[II(t, j, i)] = si[II(t, j, i)] * q[II(t, j, i)];
                }
        double alpha = gsum(w, T, sNy, sNx, sum_order); /* cg2d.F:294-304 */
        alpha = eta_qrN / alpha;                         /* cg2d.F:310 */
        LOOP_I {
            const int64_t p = PIX(L, t, j + OL, i + OL);
            x[p] = x[p] + alpha * si[II(t, j, i)];                 /* cg2d.F:319 */
            ri[II(t, j, i)] = r[p] - alpha * q[II(t, j, i)];       /* cg2d.F:320 */
            w[II(t, j, i)] = ri[II(t, j, i)] * ri[II(t, j, i)];
            r[p] = ri[II(t, j, i)];
        }
        err_sq = gsum(w, T, sNy, sNx, sum_order); /* cg2d.F:324-336 */
        exch_T(Tsrc, Tcomp, Tsign, L, r, tmp);    /* cg2d.F:362 */
        it = it + 1;
    }
    LOOP_I x[PIX(L, t, j + OL, i + OL)] = x[PIX(L, t, j + OL, i + OL)] / rhsNorm; /* cg2d.F:380-391 */
    memcpy(x_out, x, sizeof(double) * N);
    diag[0] = firstResidual;
    diag[1] = sqrt(err_sq);
    diag[2] = sumRHS;
    diag[3] = rhsMax;
    diag[4] = (double)it;
#undef LOOP_I
#undef II
    free(x); free(r); free(s); free(tmp); free(b); free(q); free(w); free(ri); free(si);
    return 0;
}

/* _matvec (core/cg2d.py:292-298): y = (A/cg2dNorm) exch(where(inner, v, 0)) on the interior, 0 elsewhere. */
EXPORT void mjo_cg2d_matvec(const int32_t* Tsrc, const int8_t* Tcomp, const int8_t* Tsign, int T, int sNx, int sNy,
                            int OL, const double* aW, const double* aS, const double* aC, const double* v, double* y,
                            double cg2dNorm) {
    Lay L = {T, sNy + 2 * OL, sNx + 2 * OL, OL, sNx, sNy};
    const int64_t N = (int64_t)T * L.ny * L.nx;
    double* xe = (double*)calloc(N, sizeof(double));
    double* tmp = (double*)malloc(sizeof(double) * N);
    for (int t = 0; t < T; ++t)
        for (int j = OL; j < OL + sNy; ++j)
            for (int i = OL; i < OL + sNx; ++i) xe[PIX(L, t, j, i)] = v[PIX(L, t, j, i)];
    exch_T(Tsrc, Tcomp, Tsign, L, xe, tmp);
    memset(y, 0, sizeof(double) * N);
    for (int t = 0; t < T; ++t)
        for (int j = OL; j < OL + sNy; ++j)
            for (int i = OL; i < OL + sNx; ++i) {
                const int64_t p = PIX(L, t, j, i);
                y[p] = op_A(aW, aS, aC, xe, p, L.nx) / cg2dNorm;
            }
    free(xe);
    free(tmp);
}

/* =================================================================================================== DST3 */
/* dst3_adv_x (pkgs/gad.py:409-430) on one (tile, level) plane: flux at i = 2 .. nx-2 (Fortran 1-OLx+2 ..
 * sNx+OLx-1), all rows; 0 elsewhere. */
static void dst3_x(int ny, int nx, double oneSixth, double dt, double rdeep, const double* uTr, const double* u,
                   const double* m, const double* t, const double* rdxC, double* af) {
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const int q = j * nx + i;
            if (i < 2 || i > nx - 2) {
                af[q] = 0.0;
                continue;
            }
            const double Rjp = (t[q + 1] - t[q]) * m[q + 1];
            const double Rj = (t[q] - t[q - 1]) * m[q];
            const double Rjm = (t[q - 1] - t[q - 2]) * m[q - 1];
            const double uCFL = fabs(u[q] * dt * rdxC[q] * rdeep);
            const double d0 = (2. - uCFL) * (1. - uCFL) * oneSixth;
            const double d1 = (1. - uCFL * uCFL) * oneSixth;
            const double U = uTr[q];
            af[q] = 0.5 * (U + fabs(U)) * (t[q - 1] + (d0 * Rj + d1 * Rjm)) +
                    0.5 * (U - fabs(U)) * (t[q] - (d0 * Rj + d1 * Rjp));
        }
}

/* dst3_adv_y (pkgs/gad.py:433-453): flux at j = 2 .. ny-2, all columns. */
static void dst3_y(int ny, int nx, double oneSixth, double dt, double rdeep, const double* vTr, const double* v,
                   const double* m, const double* t, const double* rdyC, double* af) {
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const int q = j * nx + i;
            if (j < 2 || j > ny - 2) {
                af[q] = 0.0;
                continue;
            }
            const double Rjp = (t[q + nx] - t[q]) * m[q + nx];
            const double Rj = (t[q] - t[q - nx]) * m[q];
            const double Rjm = (t[q - nx] - t[q - 2 * nx]) * m[q - nx];
            const double vCFL = fabs(v[q] * dt * rdyC[q] * rdeep);
            const double d0 = (2. - vCFL) * (1. - vCFL) * oneSixth;
            const double d1 = (1. - vCFL * vCFL) * oneSixth;
            const double V = vTr[q];
            af[q] = 0.5 * (V + fabs(V)) * (t[q - nx] + (d0 * Rj + d1 * Rjm)) +
                    0.5 * (V - fabs(V)) * (t[q] - (d0 * Rj + d1 * Rjp));
        }
}

/* dst3_adv_x / dst3_adv_y (pkgs/gad.py:409-453) on [T, K, ny, nx] arrays: per (tile, level) plane the interface
 * fluxes of dst3_x / dst3_y with deltaTloc(k) and recip_deepFacC(k); rd = recip_dxC / recip_dyC [T, ny, nx]. */
EXPORT void mjo_dst3_adv(int dir, int T, int K, int ny, int nx, double oneSixth, const double* deltaTloc,
                         const double* trans, const double* vel, const double* mask, const double* tracer,
                         const double* rd, const double* rdeep, double* out) {
    const int64_t plane = (int64_t)ny * nx;
#pragma omp parallel for schedule(static)
    for (int tk = 0; tk < T * K; ++tk) {
        const int t = tk / K, k = tk % K;
        const int64_t o = (int64_t)tk * plane;
        if (dir == 0)
            dst3_x(ny, nx, oneSixth, deltaTloc[k], rdeep[k], trans + o, vel + o, mask + o, tracer + o,
                   rd + (int64_t)t * plane, out + o);
        else
            dst3_y(ny, nx, oneSixth, deltaTloc[k], rdeep[k], trans + o, vel + o, mask + o, tracer + o,
                   rd + (int64_t)t * plane, out + o);
    }
}

static void gather_plane(double* a, const int32_t* src, int plane, double* tmp) {
    for (int q = 0; q < plane; ++q) tmp[q] = a[src[q]];
    memcpy(a, tmp, sizeof(double) * plane);
}

/* advect_tiles (pkgs/gad.py:460-539), each (tile, level) plane on its own (the routine is tile- and
 * level-local). vert[8][K]: dTtracerLev, deepFacC, recip_deepFacC, recip_deepFac2C, recip_rhoFacC, drF,
 * recip_drF, rhoFacC. masks[(p*2+d)][T][plane]: mX{p+1} (d=0), mY{p+1} (d=1). fills[(p*4+f)] (f = fXb, fXa,
 * fYb, fYa): a [T][plane] gather table or NULL when no tile calls that fill. uvfill_u/v: [T][plane] indices into
 * concat(maskW plane, maskS plane). out: gTracer, or (want_local) localTij after the passes. */
EXPORT void mjo_advect_tiles(int T, int K, int ny, int nx, const double* vert, double oneSixth, const double* dyG,
                             const double* dxG, const double* recip_dxC, const double* recip_dyC,
                             const double* recip_rA, const double* maskInC, const double* maskW, const double* maskS,
                             const uint8_t* masks, const int32_t* const* fills, const int32_t* uvfill_u,
                             const int32_t* uvfill_v, const double* uFld, const double* vFld, const double* tracer,
                             const double* hFacW, const double* hFacS, const double* recip_hFacC, double* out,
                             int want_local) {
    const int plane = ny * nx;
    const double *dTl = vert, *deepFacC = vert + K, *rdeepC = vert + 2 * K, *rdeep2C = vert + 3 * K,
                 *rrhoFacC = vert + 4 * K, *drF = vert + 5 * K, *rdrF = vert + 6 * K, *rhoFacC = vert + 7 * K;
#pragma omp parallel
    {
        double* uTr = (double*)malloc(sizeof(double) * plane);
        double* vTr = (double*)malloc(sizeof(double) * plane);
        double* mLW = (double*)malloc(sizeof(double) * plane);
        double* mLS = (double*)malloc(sizeof(double) * plane);
        double* coef = (double*)malloc(sizeof(double) * plane);
        double* dUx = (double*)malloc(sizeof(double) * plane);
        double* dVy = (double*)malloc(sizeof(double) * plane);
        double* loc = (double*)malloc(sizeof(double) * plane);
        double* af = (double*)malloc(sizeof(double) * plane);
        double* tmp = (double*)malloc(sizeof(double) * plane);
#pragma omp for schedule(static)
        for (int tk = 0; tk < T * K; ++tk) {
            const int t = tk / K, k = tk % K;
            const int64_t o3 = (int64_t)tk * plane, o2 = (int64_t)t * plane;
            const double dt = dTl[k];
            const double *u = uFld + o3, *v = vFld + o3, *tr = tracer + o3, *hW = hFacW + o3, *hS = hFacS + o3,
                         *rhC = recip_hFacC + o3, *mW = maskW + o3, *mS = maskS + o3;
            for (int q = 0; q < plane; ++q) {
                /* gad.py:482-485: xA = dyG*deepFacC*drF*hFacW; uTrans = uFld*xA*rhoFacC */
                const double xA = dyG[o2 + q] * deepFacC[k] * drF[k] * hW[q];
                const double yA = dxG[o2 + q] * deepFacC[k] * drF[k] * hS[q];
                uTr[q] = u[q] * xA * rhoFacC[k];
                vTr[q] = v[q] * yA * rhoFacC[k];
                /* gad.py:489-490: corner-filled masks */
                const int su = uvfill_u[o2 + q], sv = uvfill_v[o2 + q];
                mLW[q] = su < plane ? mW[su] : mS[su - plane];
                mLS[q] = sv < plane ? mW[sv] : mS[sv - plane];
                /* gad.py:494 */
                coef[q] = dt * rrhoFacC[k] * rhC[q] * rdrF[k] * recip_rA[o2 + q] * rdeep2C[k];
                loc[q] = tr[q];
            }
            /* gad.py:502-509: ddx / ddy with a 0 last column / row */
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const int q = j * nx + i;
                    dUx[q] = i < nx - 1 ? uTr[q + 1] - uTr[q] : 0.0;
                    dVy[q] = j < ny - 1 ? vTr[q + nx] - vTr[q] : 0.0;
                }
            for (int p = 0; p < 3; ++p) {
                for (int d = 0; d < 2; ++d) {
                    const uint8_t* msk = masks + ((int64_t)(p * 2 + d) * T + t) * plane;
                    const int32_t* fb = fills[p * 4 + d * 2];
                    const int32_t* fa = fills[p * 4 + d * 2 + 1];
                    int any = fb != NULL || fa != NULL;
                    for (int q = 0; q < plane && !any; ++q) any = msk[q];
                    if (!any) continue;
                    if (fb) gather_plane(loc, fb + o2, plane, tmp);
                    if (d == 0)
                        dst3_x(ny, nx, oneSixth, dt, rdeepC[k], uTr, u, mLW, loc, recip_dxC + o2, af);
                    else
                        dst3_y(ny, nx, oneSixth, dt, rdeepC[k], vTr, v, mLS, loc, recip_dyC + o2, af);
                    if (fa) gather_plane(loc, fa + o2, plane, tmp);
                    for (int j = 0; j < ny; ++j)
                        for (int i = 0; i < nx; ++i) {
                            const int q = j * nx + i;
                            if (!msk[q]) continue;
                            double daf, dtr;
                            if (d == 0) {
                                daf = i < nx - 1 ? af[q + 1] - af[q] : 0.0;
                                dtr = dUx[q];
                            } else {
                                daf = j < ny - 1 ? af[q + nx] - af[q] : 0.0;
                                dtr = dVy[q];
                            }
                            /* gad.py:511-514 */
                            loc[q] = loc[q] - coef[q] * (daf - tr[q] * dtr) * maskInC[o2 + q];
                        }
                }
            }
            double* o = out + o3;
            if (want_local)
                memcpy(o, loc, sizeof(double) * plane);
            else
                for (int q = 0; q < plane; ++q) o[q] = (loc[q] - tr[q]) / dt; /* gad.py:539 */
        }
        free(uTr); free(vTr); free(mLW); free(mLS); free(coef); free(dUx); free(dVy); free(loc); free(af); free(tmp);
    }
}

/* Vertical DST3 (R) sweep after the horizontal passes — not in the reference (implicitAdvection=T skips
 * gad_advection.F:835-1069). Definition (the same as the library's): interface k (0..K) is the top of cell k,
 * w(k) positive upward, rTrans(k) = w(k) rA mk(k), mk(k) = maskC(k-1) maskC(k) for 1 <= k <= K-1 and 0 at the
 * surface and bottom; DST3 interface value with upstream cell k (then k+1) for rTrans > 0 and downstream k-1
 * (then k-2), CFL |w(k) dt(k) recip_drC(k)|; local(k) -= coef(k) ((wT(k) - wT(k+1)) - tracer(k)(rTrans(k) -
 * rTrans(k+1))) maskC(k); gTracer = (local - tracer)/dt. vert as mjo_advect_tiles, plus recip_drC[K] after. */
EXPORT void mjo_vert_dst3(int T, int K, int ny, int nx, const double* vert, const double* recip_drC, double oneSixth,
                          const double* rA, const double* recip_rA, const double* maskC, const double* wFld,
                          const double* tracer, const double* recip_hFacC, const double* loc_h, double* gT) {
    const int plane = ny * nx;
    const double *dTl = vert, *rdeep2C = vert + 3 * K, *rrhoFacC = vert + 4 * K, *rdrF = vert + 6 * K;
#pragma omp parallel for schedule(static)
    for (int64_t col = 0; col < (int64_t)T * plane; ++col) {
        const int t = (int)(col / plane), q = (int)(col % plane);
#define O3(k) (((int64_t)t * K + (k)) * plane + q)
        double mkv[1024], rT[1024], wT[1024];
        for (int k = 0; k <= K; ++k) {
            mkv[k] = (k >= 1 && k <= K - 1) ? maskC[O3(k - 1)] * maskC[O3(k)] : 0.0;
        }
        for (int k = 0; k <= K; ++k) {
            if (k < 1 || k > K - 1) {
                rT[k] = 0.0;
                wT[k] = 0.0;
                continue;
            }
            rT[k] = wFld[O3(k)] * rA[(int64_t)t * plane + q] * mkv[k];
            const double tm2 = (k + 1 <= K - 1) ? loc_h[O3(k + 1)] : 0.0; /* upstream-upstream (below) */
            const double tm1 = loc_h[O3(k)];
            const double t0 = loc_h[O3(k - 1)];
            const double tp1 = (k - 2 >= 0) ? loc_h[O3(k - 2)] : 0.0;
            const double mm1 = mkv[k + 1], m0 = mkv[k], mp1 = mkv[k - 1];
            const double Rjp = (tp1 - t0) * mp1;
            const double Rj = (t0 - tm1) * m0;
            const double Rjm = (tm1 - tm2) * mm1;
            const double c = fabs(wFld[O3(k)] * dTl[k] * recip_drC[k]);
            const double d0 = (2. - c) * (1. - c) * oneSixth;
            const double d1 = (1. - c * c) * oneSixth;
            const double U = rT[k];
            wT[k] = 0.5 * (U + fabs(U)) * (tm1 + (d0 * Rj + d1 * Rjm)) + 0.5 * (U - fabs(U)) * (t0 - (d0 * Rj + d1 * Rjp));
        }
        for (int k = 0; k < K; ++k) {
            const double tr = tracer[O3(k)];
            const double cf = dTl[k] * rrhoFacC[k] * recip_hFacC[O3(k)] * rdrF[k] * recip_rA[(int64_t)t * plane + q] *
                              rdeep2C[k];
            const double loc = loc_h[O3(k)] - cf * ((wT[k] - wT[k + 1]) - tr * (rT[k] - rT[k + 1])) * maskC[O3(k)];
            gT[O3(k)] = (loc - tr) / dTl[k];
        }
#undef O3
    }
}
