"""CPU oracle of GAD_IMPLICIT_R with the U3 implicit vertical advection and SOLVE_PENTADIAGONAL (reference
core/tracer_implicit.py:34-173) — TEST INFRASTRUCTURE ONLY.

numpy FP64, one elementwise operation at a time in the reference's written order (numpy never fuses a multiply and
an add); the elimination and back substitution as explicit loops over k, vectorised over the columns. Imported only
by tests/ as the checker.

p: dTtracerLev [Nr], implicitDiffusion, rkSign; g: maskC [T, Nr, ny, nx], rA, recip_rA [T, ny, nx], recip_drF,
recip_drC [Nr]; layout sNx, sNy, OL. The deep-atmosphere and anelastic factors are 1 (the reference's DEEPFAC,
RHOFAC, tracer_implicit.py:26-27) and are not multiplied.
"""

import numpy as np

ONE_SIXTH = 1.0 / 6.0


def matrix(p, g, implicitAdvection, kappaRX, recip_hFac, wFld, sNx=90, sNy=90, OL=4):
    """gad_implicit_r_matrix (reference tracer_implicit.py:34-114): (a, b, c, d, e), each [T, Nr, ny, nx]; halo
    columns are identity rows."""
    T, Nr = kappaRX.shape[:2]
    J, I = slice(OL, OL + sNy), slice(OL, OL + sNx)
    dt = np.asarray(p.dTtracerLev, np.float64)[None, :, None, None]
    mC = g.maskC[:, :, J, I]
    rh = recip_hFac[:, :, J, I]
    kap = kappaRX[:, :, J, I]
    rdrF = np.asarray(g.recip_drF, np.float64)[None, :, None, None]
    rdrC = np.asarray(g.recip_drC, np.float64)[None, :, None, None]
    shp = rh.shape
    a = np.zeros(shp)
    b = np.zeros(shp)
    d = np.zeros(shp)
    e = np.zeros(shp)
    if not p.implicitDiffusion:
        raise NotImplementedError("implicitDiffusion = F")
    # :54-56 b(k) = -(dt(k) maskC(k-1) rh(k) rdrF(k) kap(k) rdrC(k)), k = 2..Nr
    b[:, 1:] = -(((((dt[:, 1:] * mC[:, :-1]) * rh[:, 1:]) * rdrF[:, 1:]) * kap[:, 1:]) * rdrC[:, 1:])
    # :59-62 d(k) = -(dt(k) maskC(k+1) rh(k) rdrF(k) kap(k+1) rdrC(k+1)), k = 1..Nr-1
    d[:, :-1] = -(((((dt[:, :-1] * mC[:, 1:]) * rh[:, :-1]) * rdrF[:, :-1]) * kap[:, 1:]) * rdrC[:, 1:])
    c = 1.0 - (b + d)                                                             # :64
    if implicitAdvection:
        # :71-78 interfaces k = 2..Nr (index 0 <-> k = 2): rTrans, rCenter, the U3 masks
        rTrans = (wFld[:, 1:, J, I] * g.rA[:, None, J, I]) * mC[:, :-1]
        ks = np.arange(2, Nr + 1)
        km2 = np.maximum(1, ks - 2) - 1
        kp1 = np.minimum(Nr, ks + 1) - 1
        maskP1 = np.where(ks >= Nr, 0.0, 1.0)[None, :, None, None]
        maskM2 = np.where(ks <= 2, 0.0, 1.0)[None, :, None, None]
        rCenter = ((0.5 * rTrans) * g.recip_rA[:, None, J, I]) * p.rkSign        # :80
        mskM = mC[:, km2] * maskM2                                                # :81
        mskP = mC[:, kp1] * maskP1                                                # :82
        rUpwind = (2.0 * ONE_SIXTH) * np.abs(rCenter)                             # :84
        rC4km = (ONE_SIXTH * (rCenter + np.abs(rCenter))) * mskM                  # :85
        rC4kp = (ONE_SIXTH * (rCenter - np.abs(rCenter))) * mskP                  # :86

        def sc(x, lev):                                                            # :88-89
            return ((x * dt[:, lev]) * rh[:, lev]) * rdrF[:, lev]

        up = slice(1, Nr)
        lo = slice(0, Nr - 1)
        b[:, lo] = b[:, lo] - sc(rC4km, lo)                                       # :97
        c[:, lo] = c[:, lo] + sc((rCenter + rUpwind) + rC4km, lo)                 # :98
        d[:, lo] = d[:, lo] + sc((rCenter - rUpwind) + rC4kp, lo)                 # :99
        e[:, lo] = e[:, lo] - sc(rC4kp, lo)                                       # :100
        a[:, up] = a[:, up] + sc(rC4km, up)                                       # :102
        b[:, up] = b[:, up] - sc((rCenter + rUpwind) + rC4km, up)                 # :103
        c[:, up] = c[:, up] - sc((rCenter - rUpwind) + rC4kp, up)                 # :104
        d[:, up] = d[:, up] + sc(rC4kp, up)                                       # :105

    def full(x, init):                                                            # :107-110
        out = np.full(recip_hFac.shape, init, np.float64)
        out[:, :, J, I] = x
        return out

    return full(a, 0.0), full(b, 0.0), full(c, 1.0), full(d, 0.0), full(e, 0.0)


def _norm(cp, dp, ep, yp):
    """solve_pentadiagonal.F:256-267 (reference tracer_implicit.py:120-124): divide by a nonzero pivot, else 0."""
    nz = cp != 0.0
    rec = 1.0 / np.where(nz, cp, 1.0)
    return np.where(nz, dp * rec, 0.0), np.where(nz, ep * rec, 0.0), np.where(nz, yp * rec, 0.0)


def solve_pentadiagonal(a, b, c, d, e, y):
    """solve_pentadiagonal (reference tracer_implicit.py:116-162) on every column, k = axis 1."""
    Nr = y.shape[1]
    dp = np.empty_like(y)
    ep = np.empty_like(y)
    yp = np.empty_like(y)
    dp[:, 0], ep[:, 0], yp[:, 0] = _norm(c[:, 0], d[:, 0], e[:, 0], y[:, 0])          # :127
    c2 = c[:, 1] - b[:, 1] * dp[:, 0]                                                # :129-131
    d2 = d[:, 1] - b[:, 1] * ep[:, 0]
    y2 = y[:, 1] - b[:, 1] * yp[:, 0]
    dp[:, 1], ep[:, 1], yp[:, 1] = _norm(c2, d2, e[:, 1], y2)                         # :132
    for k in range(2, Nr):                                                           # :134-143
        bb = b[:, k] - a[:, k] * dp[:, k - 2]
        cq = (c[:, k] - a[:, k] * ep[:, k - 2]) - bb * dp[:, k - 1]
        dq = d[:, k] - bb * ep[:, k - 1]
        yq = (y[:, k] - a[:, k] * yp[:, k - 2]) - bb * yp[:, k - 1]
        dp[:, k], ep[:, k], yp[:, k] = _norm(cq, dq, e[:, k], yq)
    u = np.empty_like(y)
    u[:, Nr - 1] = yp[:, Nr - 1]                                                      # :151
    u[:, Nr - 2] = yp[:, Nr - 2] - u[:, Nr - 1] * dp[:, Nr - 2]                       # :152
    for k in range(Nr - 3, -1, -1):                                                  # :154-160
        u[:, k] = (yp[:, k] - u[:, k + 1] * dp[:, k]) - u[:, k + 2] * ep[:, k]
    return u


def gad_implicit_r(p, g, implicitAdvection, kappaRX, recip_hFac, wFld, gTracer, sNx=90, sNy=90, OL=4):
    """GAD_IMPLICIT_R (reference tracer_implicit.py:165-173): the tracer after the implicit vertical advection (U3)
    and diffusion solve, on every padded column (halo columns unchanged through identity rows)."""
    if gTracer.shape[1] <= 1:
        return np.array(gTracer, np.float64)
    return solve_pentadiagonal(*matrix(p, g, implicitAdvection, kappaRX, recip_hFac, wFld, sNx, sNy, OL), gTracer)


def gad_implicit_r_adjoint(p, g, implicitAdvection, kappaRX, recip_hFac, wFld, u, ubar, sNx=90, sNy=90, OL=4):
    """Reverse mode of gad_implicit_r for u = A(kappaRX, wFld)^-1 y (the reference differentiates the literal
    solve with JAX): lam = A^-T ubar (the same elimination on the transposed diagonals), ybar = lam,
    kappaRX_bar(k) = -(lam(k) beta(k) (u(k-1) - u(k)) + lam(k-1) delta(k-1) (u(k) - u(k-1))) with beta, delta the
    derivatives of b(k), d(k-1) in kappaRX(k); wFld_bar(K) from the U3 interface K's coefficients (rows K-1 and K).
    Halo columns: ybar = ubar, no parameter cotangent. Returns (ybar, kappaRX_bar, wFld_bar)."""
    T, Nr = u.shape[:2]
    if Nr <= 1:
        return np.array(ubar, np.float64), np.zeros_like(u), np.zeros_like(u)
    a, b, c, d, e = matrix(p, g, implicitAdvection, kappaRX, recip_hFac, wFld, sNx, sNy, OL)
    aT, bT, dT, eT = (np.zeros_like(a) for _ in range(4))
    aT[:, 2:] = e[:, :-2]
    bT[:, 1:] = d[:, :-1]
    dT[:, :-1] = b[:, 1:]
    eT[:, :-2] = a[:, 2:]
    lam = solve_pentadiagonal(aT, bT, c, dT, eT, ubar)
    J, I = slice(OL, OL + sNy), slice(OL, OL + sNx)
    dt = np.asarray(p.dTtracerLev, np.float64)[None, :, None, None]
    mC = g.maskC[:, :, J, I]
    rh = recip_hFac[:, :, J, I]
    rdrF = np.asarray(g.recip_drF, np.float64)[None, :, None, None]
    rdrC = np.asarray(g.recip_drC, np.float64)[None, :, None, None]
    ui = u[:, :, J, I]
    li = lam[:, :, J, I]
    kb = np.zeros_like(ui)
    beta = -((((dt[:, 1:] * mC[:, :-1]) * rh[:, 1:]) * rdrF[:, 1:]) * rdrC[:, 1:])      # d b(k) / d kap(k)
    delta = -((((dt[:, :-1] * mC[:, 1:]) * rh[:, :-1]) * rdrF[:, :-1]) * rdrC[:, 1:])   # d d(k-1) / d kap(k)
    du = ui[:, 1:] - ui[:, :-1]                                                       # u(k) - u(k-1)
    kb[:, 1:] = -((li[:, 1:] * beta) * (-du) + (li[:, :-1] * delta) * du)
    wb = np.zeros_like(ui)
    if implicitAdvection:
        ks = np.arange(1, Nr)                                                         # interfaces K (0-based)
        rA = g.rA[:, None, J, I]
        rrA = g.recip_rA[:, None, J, I]
        dw = ((0.5 * rA) * mC[:, :-1]) * rrA * p.rkSign                               # d rC / d w(K)
        rC = ((0.5 * ((wFld[:, 1:, J, I] * rA) * mC[:, :-1])) * rrA) * p.rkSign
        sg = np.sign(rC)
        mskM = mC[:, np.maximum(0, ks - 2)] * np.where(ks <= 1, 0.0, 1.0)[None, :, None, None]
        mskP = mC[:, np.minimum(Nr - 1, ks + 1)] * np.where(ks >= Nr - 1, 0.0, 1.0)[None, :, None, None]
        drUp = (2.0 * ONE_SIXTH) * sg
        dkm = (ONE_SIXTH * (1.0 + sg)) * mskM
        dkp = (ONE_SIXTH * (1.0 - sg)) * mskP
        upad = np.concatenate([np.zeros_like(ui[:, :1]), ui, np.zeros_like(ui[:, :1])], axis=1)  # u(-1) .. u(Nr)
        uKm2, uKm1, uK, uKp1 = upad[:, ks - 1], upad[:, ks], upad[:, ks + 1], upad[:, ks + 2]
        phi = ((-dkm * uKm2 + ((1.0 + drUp) + dkm) * uKm1) + ((1.0 - drUp) + dkp) * uK) - dkp * uKp1
        s = (dt * rh) * rdrF                                                          # row scale of level L
        rCb = -((li[:, :-1] * s[:, :-1]) - (li[:, 1:] * s[:, 1:])) * phi
        wb[:, 1:] = rCb * dw
    def full(x):
        out = np.zeros_like(u)
        out[:, :, J, I] = x
        return out
    return lam, full(kb), full(wb)
