"""CPU oracle of UPDATE_CG2D and cg2d_rhs (reference core/solve_for_pressure.py:29-106) — TEST INFRASTRUCTURE ONLY.

numpy FP64, one elementwise operation at a time in the reference's written order (numpy never fuses a multiply
and an add), the level sums as explicit sequential loops. The halo exchange of aC2d is the oracle's exchange
(oracle.scalar with the T map). Imported only by tests/ as the checker.

Arrays: [T, K, ny, nx] 3-D and [T, ny, nx] 2-D, one member; grid: an object with dyG, dxG, recip_dxC, recip_dyC, rA,
maskInC ([T, ny, nx]) and drF ([K]); vertical factors all 1 (reference core/free_surface.py:121-125) unless given.
"""

import numpy as np

from . import scalar


def _sl(lo, hi, OL):
    """Array slice of the Fortran index range lo..hi (inclusive) for halo width OL (reference layout.py)."""
    return slice(lo - 1 + OL, hi + OL)


def ini_cg2d_norm(g, hFacW, hFacS, implicSurfPress, implicDiv2DFlow, sNx=90, sNy=90, OL=4):
    """INI_CG2D's cg2dNorm (reference core/cg2d.py:108-134): on the interior, aW = sum_k (fac ((dyG drF) hFacW))
    recip_dxC, k = 1..Nr in order from 0 (:123-128), aS likewise, fac = implicSurfPress implicDiv2DFlow (:124);
    1 / max(|aW|, |aS|), 1 when the max is 0 (:129-134)."""
    K = hFacW.shape[1]
    J, I = _sl(1, sNy, OL), _sl(1, sNx, OL)
    fac = implicSurfPress * implicDiv2DFlow
    aW = np.zeros_like(hFacW[:, 0, J, I])
    aS = np.zeros_like(aW)
    for k in range(K):
        aW = aW + (fac * ((g.dyG[:, J, I] * g.drF[k]) * hFacW[:, k, J, I])) * g.recip_dxC[:, J, I]
        aS = aS + (fac * ((g.dxG[:, J, I] * g.drF[k]) * hFacS[:, k, J, I])) * g.recip_dyC[:, J, I]
    m = max(float(np.abs(aW).max()), float(np.abs(aS).max()))
    return 1.0 / m if m != 0.0 else 1.0


def update_cg2d(p, cg2dNorm, cg2dpcOffDFac, g, Tmap, hFacW, hFacS, recip_Bo, pW, pS, pC, sNx=90, sNy=90, OL=4):
    """UPDATE_CG2D (reference core/solve_for_pressure.py:29-65). p: deltaTMom, deltaTFreeSurf, freeSurfFac,
    implicSurfPress, implicDiv2DFlow. Returns (aW2d, aS2d, aC2d, pW, pS, pC), new arrays."""
    T, K = hFacW.shape[:2]
    J1, I1 = _sl(1, sNy + 1, OL), _sl(1, sNx + 1, OL)
    z2 = np.zeros_like(hFacW[:, 0])                                              # :38
    # :41-43 terms ((dyG * drF) * hFacW) * recip_dxC, summed k = 1..Nr in order from 0 (_ksum :23-26)
    aW = z2[:, J1, I1].copy()
    aS = z2[:, J1, I1].copy()
    for k in range(K):
        aW = aW + ((g.dyG[:, J1, I1] * g.drF[k]) * hFacW[:, k, J1, I1]) * g.recip_dxC[:, J1, I1]
        aS = aS + ((g.dxG[:, J1, I1] * g.drF[k]) * hFacS[:, k, J1, I1]) * g.recip_dyC[:, J1, I1]
    aW = ((aW * cg2dNorm) * p.implicSurfPress) * p.implicDiv2DFlow                # :44-45
    aS = ((aS * cg2dNorm) * p.implicSurfPress) * p.implicDiv2DFlow
    aW2d = z2.copy()
    aS2d = z2.copy()
    aW2d[:, J1, I1] = aW                                                          # :46-47
    aS2d[:, J1, I1] = aS
    J, I, Jp, Ip = _sl(1, sNy, OL), _sl(1, sNx, OL), _sl(2, sNy + 1, OL), _sl(2, sNx + 1, OL)
    d = ((((p.freeSurfFac * cg2dNorm) * recip_Bo[:, J, I]) * g.rA[:, J, I]) / p.deltaTMom) / p.deltaTFreeSurf
    aC = -((((aW2d[:, J, I] + aW2d[:, J, Ip]) + aS2d[:, J, I]) + aS2d[:, Jp, I]) + d)  # :50-51
    aC2d = z2.copy()
    aC2d[:, J, I] = aC                                                            # :52
    aC2d = scalar(Tmap, aC2d, T, aC2d.shape[1] * aC2d.shape[2])                  # :53 exch_xy
    Im, Jm = _sl(0, sNx, OL), _sl(0, sNy, OL)
    a = aC2d[:, J1, I1]
    pC_n = np.where(a == 0.0, 1.0, 1.0 / np.where(a == 0.0, 1.0, a))              # :57
    tW = a + aC2d[:, J1, Im]                                                      # :58
    fW = cg2dpcOffDFac * np.where(tW == 0.0, 1.0, tW)
    pW_n = np.where(tW == 0.0, 0.0, -aW2d[:, J1, I1] / (fW * fW))                 # :59-60 (x**2 = x*x)
    tS = a + aC2d[:, Jm, I1]                                                      # :61
    fS = cg2dpcOffDFac * np.where(tS == 0.0, 1.0, tS)
    pS_n = np.where(tS == 0.0, 0.0, -aS2d[:, J1, I1] / (fS * fS))                 # :62-63
    pW, pS, pC = np.array(pW, np.float64), np.array(pS, np.float64), np.array(pC, np.float64)
    pW[:, J1, I1] = pW_n
    pS[:, J1, I1] = pS_n
    pC[:, J1, I1] = pC_n
    return aW2d, aS2d, aC2d, pW, pS, pC


def k_surf(maskC):
    """1-based first wet level, Nr+1 for a dry column (reference core/free_surface.py:128-134)."""
    Nr = maskC.shape[1]
    wet = maskC != 0
    return np.where(wet.any(axis=1), np.argmax(wet, axis=1) + 1, Nr + 1).astype(np.int32)


def cg2d_rhs(p, g, gU, gV, hFacW, hFacS, EmPmR, etaN, etaH, Bo_surf, sNx=90, sNy=90, OL=4, deepFacC=None,
             rhoFacC=None, deepFac2F=None, maskC=None):
    """cg2d_b, cg2d_x (reference core/solve_for_pressure.py:68-106). p: deltaTMom, deltaTFreeSurf, freeSurfFac,
    implicDiv2DFlow, mass2rUnit, useRealFreshWaterFlux. The vertical factors default to 1 (deepFac2F then needs no
    kSurf); with deepFac2F given, maskC gives kSurf."""
    T, K = gU.shape[:2]
    J, I = _sl(1, sNy, OL), _sl(1, sNx, OL)
    cg2d_x = Bo_surf * etaN                                                       # :76
    cg2d_b = np.zeros_like(etaN)                                                  # :77
    if p.useRealFreshWaterFlux:                                                   # :78-80
        tmpFac = (p.freeSurfFac * p.mass2rUnit) * p.implicDiv2DFlow
        cg2d_b[:, J, I] = (((tmpFac * g.rA[:, J, I]) * EmPmR[:, J, I]) / p.deltaTMom) * g.maskInC[:, J, I]
    b = cg2d_b[:, J, I].copy()
    J1, I1 = _sl(1, sNy + 1, OL), _sl(1, sNx + 1, OL)
    Ju, Iu = _sl(1, sNy, OL), _sl(1, sNx + 1, OL)
    Jv, Iv = _sl(1, sNy + 1, OL), _sl(1, sNx, OL)
    one = np.ones(K)
    dfc = one if deepFacC is None else deepFacC
    rfc = one if rhoFacC is None else rhoFacC
    for k in range(K - 1, -1, -1):                                                # :101 scan, reverse
        # :90-93 xA = (((dyG * deepFacC) * drF) * hFacW) * rhoFacC; pfu = (xA * gU) / deltaTMom
        xA = (((g.dyG[:, J1, I1] * dfc[k]) * g.drF[k]) * hFacW[:, k, J1, I1]) * rfc[k]
        yA = (((g.dxG[:, J1, I1] * dfc[k]) * g.drF[k]) * hFacS[:, k, J1, I1]) * rfc[k]
        pu = (xA[:, :-1, :] * gU[:, k, Ju, Iu]) / p.deltaTMom                    # [T, sNy, sNx+1]
        pv = (yA[:, :, :-1] * gV[:, k, Jv, Iv]) / p.deltaTMom                    # [T, sNy+1, sNx]
        b = (b + pu[:, :, 1:]) - pu[:, :, :-1]                                    # :97
        b = (b + pv[:, 1:, :]) - pv[:, :-1, :]                                    # :98
    e = p.freeSurfFac * g.rA[:, J, I]
    if deepFac2F is not None:                                                     # :103 deepFac2F(kSurfC)
        e = e * np.asarray(deepFac2F)[k_surf(maskC)[:, J, I] - 1]
    b = b - ((e / p.deltaTMom) / p.deltaTFreeSurf) * etaH[:, J, I]               # :103-105
    cg2d_b[:, J, I] = b
    return cg2d_b, cg2d_x


def cg2d_rhs_adjoint(p, g, gU, gV, hFacW, hFacS, Bo_surf, b_bar, x_bar=None, sNx=90, sNy=90, OL=4, deepFac2F=None,
                     maskC=None):
    """Reverse mode of cg2d_rhs (linear in gU, gV, EmPmR, etaN, etaH; linear in hFacW, hFacS for fixed gU, gV), V4r4
    vertical factors: from the cotangents of cg2d_b (interior) and cg2d_x, the cotangents (gU, gV, hFacW, hFacS,
    EmPmR, etaN, etaH). pfu(j, i) enters b(j, i-1) with + and b(j, i) with -, pfv(j, i) enters b(j-1, i) with + and
    b(j, i) with -."""
    T, K = gU.shape[:2]
    ny, nx = gU.shape[2:]
    J, I = _sl(1, sNy, OL), _sl(1, sNx, OL)
    bb = np.zeros((T, ny, nx))
    bb[:, J, I] = b_bar[:, J, I]
    # face cotangents: u faces (j in J, i = 1..sNx+1), v faces (j = 1..sNy+1, i in I)
    pu = np.zeros((T, ny, nx))
    pu[:, J, OL:OL + sNx + 1] = bb[:, J, OL - 1:OL + sNx] - bb[:, J, OL:OL + sNx + 1]
    pv = np.zeros((T, ny, nx))
    pv[:, OL:OL + sNy + 1, I] = bb[:, OL - 1:OL + sNy, I] - bb[:, OL:OL + sNy + 1, I]
    drF = np.asarray(g.drF, np.float64)[None, :, None, None]
    cu = g.dyG[:, None] * drF                                                       # d xA / d hFacW
    cv = g.dxG[:, None] * drF
    gUb = (pu[:, None] * (cu * hFacW)) / p.deltaTMom
    gVb = (pv[:, None] * (cv * hFacS)) / p.deltaTMom
    hWb = (pu[:, None] * (cu * gU)) / p.deltaTMom
    hSb = (pv[:, None] * (cv * gV)) / p.deltaTMom
    Eb = np.zeros((T, ny, nx))
    if p.useRealFreshWaterFlux:
        tmpFac = (p.freeSurfFac * p.mass2rUnit) * p.implicDiv2DFlow
        Eb[:, J, I] = bb[:, J, I] * (((tmpFac * g.rA[:, J, I]) / p.deltaTMom) * g.maskInC[:, J, I])
    e = p.freeSurfFac * g.rA[:, J, I]
    if deepFac2F is not None:
        e = e * np.asarray(deepFac2F)[k_surf(maskC)[:, J, I] - 1]
    eHb = np.zeros((T, ny, nx))
    eHb[:, J, I] = -(bb[:, J, I] * ((e / p.deltaTMom) / p.deltaTFreeSurf))
    eNb = Bo_surf * x_bar if x_bar is not None else np.zeros((T, ny, nx))
    return gUb, gVb, hWb, hSb, Eb, eNb, eHb


def exchange_transpose(Tmap, abar):
    """Transpose of the scalar exchange on [T, ny, nx]: a written point passes sign * its cotangent to its source and
    keeps none; other points keep their own (summed with np.add.at)."""
    src, comp, sign = Tmap
    a = np.asarray(abar, np.float64).reshape(-1)
    out = a.copy()
    w = comp != 0
    out[w] = 0.0
    np.add.at(out, src[w], sign[w].astype(np.float64) * a[w])
    return out.reshape(np.shape(abar))


def update_cg2d_adjoint(p, cg2dNorm, g, Tmap, aW_bar, aS_bar, aC_bar, K, sNx=90, sNy=90, OL=4):
    """Reverse mode of UPDATE_CG2D in hFacW, hFacS (the operator is affine in them; the preconditioner outputs have no
    cotangent, as CG2D's derivative does not use them): aC2d = exch(aC) gives aC_bar' = exch^T(aC2d_bar) on the
    interior; aW_tot(j, i) = (aW_bar - aC'(j, i)) - aC'(j, i-1), aS_tot(j, i) = (aS_bar - aC'(j, i)) - aC'(j-1, i)
    (aC' taken as 0 off the interior); hFacW_bar(k) = (aW_tot ((norm isp) idf)) ((dyG drF(k)) recip_dxC) on
    j, i = 1..sNy+1, 1..sNx+1, 0 elsewhere (hFacS_bar alike)."""
    T = aW_bar.shape[0]
    J, I = _sl(1, sNy, OL), _sl(1, sNx, OL)
    J1, I1 = _sl(1, sNy + 1, OL), _sl(1, sNx + 1, OL)
    ct = exchange_transpose(Tmap, aC_bar)
    C = np.zeros_like(ct)
    C[:, J, I] = ct[:, J, I]
    Cw = np.zeros_like(C)
    Cw[:, :, 1:] = C[:, :, :-1]                                                     # aC'(j, i-1)
    Cs = np.zeros_like(C)
    Cs[:, 1:, :] = C[:, :-1, :]                                                     # aC'(j-1, i)
    aWt = (aW_bar - C) - Cw
    aSt = (aS_bar - C) - Cs
    fac = (cg2dNorm * p.implicSurfPress) * p.implicDiv2DFlow
    hWb = np.zeros((T, K) + aW_bar.shape[1:])
    hSb = np.zeros_like(hWb)
    drF = np.asarray(g.drF, np.float64)
    for k in range(K):
        hWb[:, k, J1, I1] = (aWt[:, J1, I1] * fac) * ((g.dyG[:, J1, I1] * drF[k]) * g.recip_dxC[:, J1, I1])
        hSb[:, k, J1, I1] = (aSt[:, J1, I1] * fac) * ((g.dxG[:, J1, I1] * drF[k]) * g.recip_dyC[:, J1, I1])
    return hWb, hSb
