"""The CPU oracle of UPDATE_CG2D and cg2d_rhs (oracle/sfp.py) pinned by what the mathematics fixes (CPU only):
closed forms on dyadic data (every operation exact, so any dropped term, wrong index or swapped operand shows),
the telescoping of the cg2d_b divergence over a tile to the net flux through the tile's edge, the sequential level
sum against numpy's sequential accumulate, the preconditioner's defining identities and land cases, and planted
errors that must change the result."""

from types import SimpleNamespace

import numpy as np
import pytest

from oracle import sfp as OS
from tests.conftest import maps

T, NY, NX, OL, N, K = 13, 98, 98, 4, 90, 12
J1, I1 = slice(OL, OL + N + 1), slice(OL, OL + N + 1)
J, I = slice(OL, OL + N), slice(OL, OL + N)

P = SimpleNamespace(deltaTMom=2.0, deltaTFreeSurf=4.0, freeSurfFac=1.0, implicSurfPress=1.0, implicDiv2DFlow=1.0,
                    mass2rUnit=0.125, useRealFreshWaterFlux=True)


def _grid(rng, dyadic=True):
    """Grid fields; dyadic: small integers and powers of two, so that sums and products are exact."""
    def f(lo, hi):
        return rng.integers(lo, hi, size=(T, NY, NX)).astype(np.float64) if dyadic else rng.uniform(lo, hi, (T, NY, NX))
    g = SimpleNamespace(dyG=f(1, 8), dxG=f(1, 8), recip_dxC=2.0 ** -rng.integers(0, 3, size=(T, NY, NX)),
                        recip_dyC=2.0 ** -rng.integers(0, 3, size=(T, NY, NX)), rA=f(1, 16),
                        maskInC=(rng.random((T, NY, NX)) > 0.2).astype(np.float64),
                        drF=rng.integers(1, 9, size=K).astype(np.float64))
    return g


def _hfac(rng, frac=True):
    h = (rng.integers(0, 5, size=(T, K, NY, NX)) / 4.0) if frac else np.ones((T, K, NY, NX))
    return h


def _update_case(seed=0):
    rng = np.random.default_rng(seed)
    g = _grid(rng)
    hW, hS = _hfac(rng), _hfac(rng)
    rBo = 2.0 ** -rng.integers(0, 3, size=(T, NY, NX))
    pW0, pS0, pC0 = (rng.normal(size=(T, NY, NX)) for _ in range(3))
    return g, hW, hS, rBo, pW0, pS0, pC0


def _update_pins(mod, seed=0):
    """The closed-form pins of UPDATE_CG2D on dyadic data; returns the list of failed pins."""
    g, hW, hS, rBo, pW0, pS0, pC0 = _update_case(seed)
    norm = 0.25
    Tm = maps().full("T")
    aW2d, aS2d, aC2d, pW, pS, pC = mod.update_cg2d(P, norm, 0.51, g, Tm, hW, hS, rBo, pW0, pS0, pC0)
    bad = []
    # aW = norm * sum_k dyG drF(k) hFacW(k) recip_dxC (exact here in any order), 0 off j, i = 1..sNy+1, 1..sNx+1
    eW = norm * np.einsum("k,tkji->tji", g.drF, hW[:, :, J1, I1]) * g.dyG[:, J1, I1] * g.recip_dxC[:, J1, I1]
    eS = norm * np.einsum("k,tkji->tji", g.drF, hS[:, :, J1, I1]) * g.dxG[:, J1, I1] * g.recip_dyC[:, J1, I1]
    if not (np.array_equal(aW2d[:, J1, I1], eW) and np.array_equal(aS2d[:, J1, I1], eS)):
        bad.append("aW/aS closed form")
    zero = np.ones((T, NY, NX), bool)
    zero[:, J1, I1] = False
    if aW2d[zero].any() or aS2d[zero].any():
        bad.append("aW/aS zero outside")
    # A applied to a constant (the CG oracle's own stencil, mjo.c op_A): on points whose four neighbours are
    # interior, A 1 = -freeSurfFac norm recip_Bo rA / dtMom / dtFS exactly
    ones = np.ones((T, NY, NX))
    y = mod_matvec(Tm, (aW2d, aS2d, aC2d), ones)
    d = norm * rBo * g.rA / (P.deltaTMom * P.deltaTFreeSurf)
    core = (slice(None), slice(OL + 1, OL + N - 1), slice(OL + 1, OL + N - 1))
    if not np.array_equal(y[core], -d[core]):
        bad.append("A 1 = -d")
    # the halo of aC2d is the T exchange of its interior: written points read their source, others stay 0
    src, comp, sign = Tm
    own = np.zeros((T, NY, NX))
    own[:, J, I] = aC2d[:, J, I]
    flat = aC2d.reshape(-1)
    written = comp != 0
    halo = np.ones((T, NY, NX), bool)
    halo[:, J, I] = False
    if not np.array_equal(flat[written], sign[written] * own.reshape(-1)[src[written]]) or \
            flat[~written & halo.reshape(-1)].any():
        bad.append("aC2d halo = T exchange")
    # preconditioner: defining identities on j, i = 1..sNy+1, 1..sNx+1, untouched elsewhere
    a = aC2d[:, J1, I1]
    nz = a != 0
    if not (np.array_equal(pC[:, J1, I1][~nz], np.ones(int((~nz).sum())))
            and np.allclose(pC[:, J1, I1][nz] * a[nz], 1.0, rtol=0, atol=2e-16)):
        bad.append("pC")
    tW = a + aC2d[:, J1, OL - 1:OL + N]
    okW = tW != 0
    if not (np.allclose((pW[:, J1, I1] * (0.51 * tW) ** 2)[okW], -aW2d[:, J1, I1][okW], rtol=1e-15, atol=0)
            and not pW[:, J1, I1][~okW].any()):
        bad.append("pW")
    tS = a + aC2d[:, OL - 1:OL + N, I1]
    okS = tS != 0
    if not (np.allclose((pS[:, J1, I1] * (0.51 * tS) ** 2)[okS], -aS2d[:, J1, I1][okS], rtol=1e-15, atol=0)
            and not pS[:, J1, I1][~okS].any()):
        bad.append("pS")
    for new, old in ((pW, pW0), (pS, pS0), (pC, pC0)):
        if not np.array_equal(new[zero], old[zero]):
            bad.append("p* untouched outside")
    return bad


def mod_matvec(Tm, A, v):
    import oracle
    return oracle.cg2d_matvec(Tm, A, v, 1.0)


def _planted(old, new):
    """oracle/sfp.py with one line changed (a planted mistake), as a module namespace."""
    from pathlib import Path
    src = Path(OS.__file__).read_text()
    assert src.count(old) == 1, old
    ns = {"__name__": "oracle.sfp_planted", "__package__": "oracle"}
    exec(compile(src.replace(old, new), "sfp_planted.py", "exec"), ns)
    return SimpleNamespace(**ns)


def test_update_cg2d_closed_form_on_dyadic_data():
    assert _update_pins(OS) == []


@pytest.mark.parametrize("old,new", [
    ("hFacW[:, k, J1, I1]) * g.recip_dxC", "hFacS[:, k, J1, I1]) * g.recip_dxC"),
    ("(aW2d[:, J, I] + aW2d[:, J, Ip])", "(aW2d[:, J, I] + aW2d[:, J, I])"),
    ("aS2d[:, Jp, I]) + d)", "aS2d[:, J, I]) + d)"),
    ("/ p.deltaTMom) / p.deltaTFreeSurf\n", "/ p.deltaTMom)\n"),
    ("tW = a + aC2d[:, J1, Im]", "tW = a + aC2d[:, Jm, I1]"),
    ("aW2d = z2.copy()", "aW2d = np.ones_like(z2)"),
])
def test_update_cg2d_planted_errors_fail_the_pins(old, new):
    assert _update_pins(_planted(old, new)) != []


def test_update_cg2d_level_sum_is_sequential():
    """On data where the order matters, the sums equal numpy's sequential accumulate over k, not a reordering."""
    rng = np.random.default_rng(1)
    g = _grid(rng, dyadic=False)
    g.drF = rng.uniform(1.0, 500.0, size=K)
    hW = rng.uniform(0, 1, size=(T, K, NY, NX)) * 10.0 ** rng.uniform(-6, 0, size=(T, K, NY, NX))
    hS = rng.uniform(0, 1, size=(T, K, NY, NX))
    z = np.zeros((T, NY, NX))
    aW2d = OS.update_cg2d(P, 1.0, 0.51, g, maps().full("T"), hW, hS, z, z, z, z)[0]
    terms = ((g.dyG[:, None, J1, I1] * g.drF[None, :, None, None]) * hW[:, :, J1, I1]) * g.recip_dxC[:, None, J1, I1]
    seq = np.add.accumulate(terms, axis=1)[:, -1]
    assert np.array_equal(aW2d[:, J1, I1], seq)
    rev = np.add.accumulate(terms[:, ::-1], axis=1)[:, -1]
    assert not np.array_equal(aW2d[:, J1, I1], rev)  # the data separates the orders


def _rhs_inputs(rng):
    g = _grid(rng)
    gU = rng.integers(-8, 9, size=(T, K, NY, NX)).astype(np.float64)
    gV = rng.integers(-8, 9, size=(T, K, NY, NX)).astype(np.float64)
    hW, hS = _hfac(rng), _hfac(rng)
    EmPmR = rng.integers(-4, 5, size=(T, NY, NX)).astype(np.float64)
    etaN = rng.integers(-4, 5, size=(T, NY, NX)).astype(np.float64)
    etaH = rng.integers(-4, 5, size=(T, NY, NX)).astype(np.float64)
    Bo = rng.integers(1, 4, size=(T, NY, NX)).astype(np.float64)
    return g, gU, gV, hW, hS, EmPmR, etaN, etaH, Bo


def test_cg2d_rhs_divergence_telescopes_to_the_tile_edge_flux():
    rng = np.random.default_rng(2)
    g, gU, gV, hW, hS, EmPmR, etaN, etaH, Bo = _rhs_inputs(rng)
    p = SimpleNamespace(**{**vars(P), "useRealFreshWaterFlux": False})
    z = np.zeros_like(etaH)
    b, x = OS.cg2d_rhs(p, g, gU, gV, hW, hS, None, etaN, z, Bo)
    assert np.array_equal(x, Bo * etaN)
    # face fluxes: u faces i = 1..sNx+1 on rows 1..sNy, v faces j = 1..sNy+1 on columns 1..sNx
    fu = np.einsum("k,tkji->tji", g.drF, hW[:, :, J, I1] * gU[:, :, J, I1]) * g.dyG[:, J, I1] / p.deltaTMom
    fv = np.einsum("k,tkji->tji", g.drF, hS[:, :, J1, I] * gV[:, :, J1, I]) * g.dxG[:, J1, I] / p.deltaTMom
    # per point: the net outflow through the cell's four faces (exact on this data)
    assert np.array_equal(b[:, J, I], (fu[:, :, 1:] - fu[:, :, :-1]) + (fv[:, 1:, :] - fv[:, :-1, :]))
    # per tile: only the edge fluxes remain
    edge = (fu[:, :, -1] - fu[:, :, 0]).sum(axis=1) + (fv[:, -1, :] - fv[:, 0, :]).sum(axis=1)
    assert np.array_equal(b[:, J, I].sum(axis=(1, 2)), edge)
    off = np.ones((T, NY, NX), bool)
    off[:, J, I] = False
    assert not b[off].any()


def test_cg2d_rhs_source_terms_closed_form():
    rng = np.random.default_rng(3)
    g, gU, gV, hW, hS, EmPmR, etaN, etaH, Bo = _rhs_inputs(rng)
    zero3 = np.zeros_like(gU)
    b, _ = OS.cg2d_rhs(P, g, zero3, zero3, hW, hS, EmPmR, etaN, etaH, Bo)
    tmp = P.freeSurfFac * P.mass2rUnit * P.implicDiv2DFlow
    e = tmp * g.rA * EmPmR / P.deltaTMom * g.maskInC - P.freeSurfFac * g.rA / (P.deltaTMom * P.deltaTFreeSurf) * etaH
    assert np.array_equal(b[:, J, I], e[:, J, I])
    # uniform flow through uniform faces: no divergence, only the source terms
    gU1 = np.full_like(gU, 3.0)
    h1 = np.ones_like(hW)
    g.dyG[:] = 2.0
    b2, _ = OS.cg2d_rhs(P, g, gU1, zero3, h1, h1, EmPmR, etaN, etaH, Bo)
    assert np.array_equal(b2[:, J, I], e[:, J, I])


def test_cg2d_rhs_deep_factors_and_ksurf():
    rng = np.random.default_rng(4)
    g, gU, gV, hW, hS, EmPmR, etaN, etaH, Bo = _rhs_inputs(rng)
    maskC = (rng.random((T, K, NY, NX)) > 0.6).astype(np.float64)
    ks = OS.k_surf(maskC)
    wet = maskC != 0
    first = np.where(wet.any(axis=1), wet.argmax(axis=1) + 1, K + 1)
    assert np.array_equal(ks, first) and ks.max() == K + 1 and ks.min() == 1
    d2f = 2.0 ** rng.integers(-2, 3, size=K + 1)
    ones = np.ones(K)
    b1, _ = OS.cg2d_rhs(P, g, gU, gV, hW, hS, EmPmR, etaN, etaH, Bo)
    b2, _ = OS.cg2d_rhs(P, g, gU, gV, hW, hS, EmPmR, etaN, etaH, Bo, deepFacC=ones, rhoFacC=ones,
                        deepFac2F=np.ones(K + 1), maskC=maskC)
    assert np.array_equal(b1, b2)  # unit factors change nothing
    b3, _ = OS.cg2d_rhs(P, g, gU, gV, hW, hS, EmPmR, etaN, etaH, Bo, deepFac2F=d2f, maskC=maskC)
    p = P
    e = -(p.freeSurfFac * g.rA * d2f[ks - 1] / (p.deltaTMom * p.deltaTFreeSurf) * etaH)
    b4, _ = OS.cg2d_rhs(SimpleNamespace(**{**vars(P), "useRealFreshWaterFlux": False}), g, np.zeros_like(gU),
                        np.zeros_like(gV), hW, hS, None, etaN, etaH, Bo, deepFac2F=d2f, maskC=maskC)
    assert np.array_equal(b4[:, J, I], e[:, J, I])
    assert not np.array_equal(b3, b1)


def test_cg2d_rhs_level_order_is_reverse_sequential():
    """On data where the order matters, cg2d_b is the sequential sum of the per-level terms pu(i+1), -pu(i),
    pv(j+1), -pv(j) for k = Nr..1 (numpy's accumulate), not another order."""
    rng = np.random.default_rng(6)
    g = _grid(rng, dyadic=False)
    g.drF = rng.uniform(1.0, 500.0, size=K)
    gU = rng.normal(size=(T, K, NY, NX)) * 10.0 ** rng.uniform(-4, 4, size=(T, K, NY, NX))
    gV = rng.normal(size=(T, K, NY, NX))
    hW, hS = rng.uniform(0, 1, (T, K, NY, NX)), rng.uniform(0, 1, (T, K, NY, NX))
    p = SimpleNamespace(**{**vars(P), "useRealFreshWaterFlux": False})
    z = np.zeros((T, NY, NX))
    b, _ = OS.cg2d_rhs(p, g, gU, gV, hW, hS, None, z, z, z)
    xA = ((g.dyG[:, None, J1, I1] * g.drF[None, :, None, None]) * hW[:, :, J1, I1])
    yA = ((g.dxG[:, None, J1, I1] * g.drF[None, :, None, None]) * hS[:, :, J1, I1])
    pu = (xA[:, :, :-1, :] * gU[:, :, J, I1]) / p.deltaTMom
    pv = (yA[:, :, :, :-1] * gV[:, :, J1, I]) / p.deltaTMom
    terms = np.stack([pu[:, :, :, 1:], -pu[:, :, :, :-1], pv[:, :, 1:, :], -pv[:, :, :-1, :]], axis=2)  # [T,K,4,..]
    seq = np.add.accumulate(terms[:, ::-1].reshape(T, 4 * K, N, N), axis=1)[:, -1]
    assert np.array_equal(b[:, J, I], seq)
    fwd = np.add.accumulate(terms.reshape(T, 4 * K, N, N), axis=1)[:, -1]
    assert not np.array_equal(b[:, J, I], fwd)


def _rhs_pins(mod):
    """The telescoping and source-term pins of cg2d_rhs; returns the list of failed pins."""
    rng = np.random.default_rng(2)
    g, gU, gV, hW, hS, EmPmR, etaN, etaH, Bo = _rhs_inputs(rng)
    bad = []
    p = SimpleNamespace(**{**vars(P), "useRealFreshWaterFlux": False})
    z = np.zeros_like(etaH)
    b, x = mod.cg2d_rhs(p, g, gU, gV, hW, hS, None, etaN, z, Bo)
    if not np.array_equal(x, Bo * etaN):
        bad.append("cg2d_x")
    fu = np.einsum("k,tkji->tji", g.drF, hW[:, :, J, I1] * gU[:, :, J, I1]) * g.dyG[:, J, I1] / p.deltaTMom
    fv = np.einsum("k,tkji->tji", g.drF, hS[:, :, J1, I] * gV[:, :, J1, I]) * g.dxG[:, J1, I] / p.deltaTMom
    if not np.array_equal(b[:, J, I], (fu[:, :, 1:] - fu[:, :, :-1]) + (fv[:, 1:, :] - fv[:, :-1, :])):
        bad.append("divergence")
    zero3 = np.zeros_like(gU)
    b2, _ = mod.cg2d_rhs(P, g, zero3, zero3, hW, hS, EmPmR, etaN, etaH, Bo)
    tmp = P.freeSurfFac * P.mass2rUnit * P.implicDiv2DFlow
    e = tmp * g.rA * EmPmR / P.deltaTMom * g.maskInC - P.freeSurfFac * g.rA / (P.deltaTMom * P.deltaTFreeSurf) * etaH
    if not np.array_equal(b2[:, J, I], e[:, J, I]):
        bad.append("source terms")
    return bad


def test_cg2d_rhs_pins_hold():
    assert _rhs_pins(OS) == []


@pytest.mark.parametrize("old,new", [
    ("b = (b + pu[:, :, 1:]) - pu[:, :, :-1]", "b = (b + pu[:, :, :-1]) - pu[:, :, 1:]"),
    ("gU[:, k, Ju, Iu]", "gV[:, k, Ju, Iu]"),
    ("b = (b + pv[:, 1:, :]) - pv[:, :-1, :]", "b = (b + pv[:, :-1, :]) - pv[:, 1:, :]"),
    ("g.dxG[:, J1, I1] * dfc[k]", "g.dyG[:, J1, I1] * dfc[k]"),
    ("* etaH[:, J, I]", "* etaN[:, J, I]"),
    ("EmPmR[:, J, I]) / p.deltaTMom) * g.maskInC[:, J, I]", "EmPmR[:, J, I]) / p.deltaTMom)"),
    ("cg2d_x = Bo_surf * etaN", "cg2d_x = Bo_surf * etaH"),
])
def test_cg2d_rhs_planted_errors_fail_the_pins(old, new):
    assert _rhs_pins(_planted(old, new)) != []


def test_oracle_matches_the_synthetic_operator_builder():
    """The oracle and the synthetic-config builder (written separately from the same reference lines) agree
    bitwise on the seed grid with the 5-point preconditioner."""
    from mit_jax_amd import synthetic as S
    g = S.make_grid(0)
    norm = S.ini_cg2d_norm(g)
    ref = S.update_cg2d(g, norm, precond="5point")
    p = SimpleNamespace(deltaTMom=S.DT, deltaTFreeSurf=S.DT, freeSurfFac=1.0, implicSurfPress=1.0,
                        implicDiv2DFlow=1.0)
    z = np.zeros((T, NY, NX))
    got = OS.update_cg2d(p, norm, 0.51, g, maps().full("T"), g.hFacW, g.hFacS, np.full((T, NY, NX), 1.0 / S.G), z,
                         z, z)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_cg2d_rhs_adjoint_dot_products():
    """The rhs is linear in (gU, gV, EmPmR, etaH) jointly and, for fixed gU and gV, in (hFacW, hFacS, EmPmR, etaH);
    cg2d_x is linear in etaN: <b_bar, b> equals the sum of <cotangent, input> over each such group (Euler's
    identity), and the same for cg2d_x; a planted sign error in a face cotangent breaks it."""
    rng = np.random.default_rng(9)
    g, gU, gV, hW, hS, EmPmR, etaN, etaH, Bo = _rhs_inputs(rng)
    g.dyG = rng.uniform(1, 2, g.dyG.shape)
    gU, gV = rng.normal(size=gU.shape), rng.normal(size=gV.shape)
    bb, xb = rng.normal(size=etaN.shape), rng.normal(size=etaN.shape)
    b, x = OS.cg2d_rhs(P, g, gU, gV, hW, hS, EmPmR, etaN, etaH, Bo)

    def check(mod):
        gUb, gVb, hWb, hSb, Eb, eNb, eHb = mod.cg2d_rhs_adjoint(P, g, gU, gV, hW, hS, Bo, bb, xb)
        lhs = float((bb * b).sum())
        r1 = float((gUb * gU).sum() + (gVb * gV).sum() + (Eb * EmPmR).sum() + (eHb * etaH).sum())
        r2 = float((hWb * hW).sum() + (hSb * hS).sum() + (Eb * EmPmR).sum() + (eHb * etaH).sum())
        rx = float((eNb * etaN).sum())
        scale = float((np.abs(bb) * np.abs(b)).sum())
        return (abs(lhs - r1) < 1e-12 * scale and abs(lhs - r2) < 1e-12 * scale and
                abs(float((xb * x).sum()) - rx) < 1e-12 * float((np.abs(xb) * np.abs(x)).sum()))

    assert check(OS)
    bad = _planted("pu[:, J, OL:OL + sNx + 1] = bb[:, J, OL - 1:OL + sNx] - bb[:, J, OL:OL + sNx + 1]",
                   "pu[:, J, OL:OL + sNx + 1] = bb[:, J, OL:OL + sNx + 1] - bb[:, J, OL - 1:OL + sNx]")
    assert not check(bad)


def test_update_cg2d_adjoint_dot_product():
    """UPDATE_CG2D is affine in (hFacW, hFacS): <(aW_bar, aS_bar, aC_bar), ops(h) - ops(0)> equals
    <(hFacW_bar, hFacS_bar), h> (the preconditioner outputs carry no cotangent); a planted index error in the
    aC cotangent's neighbour breaks it."""
    rng = np.random.default_rng(11)
    g, hW, hS, rBo, *_ = _update_case(11)
    g.dyG = rng.uniform(1, 2, g.dyG.shape)
    hW, hS = rng.uniform(0, 1, hW.shape), rng.uniform(0, 1, hS.shape)
    Tm = maps().full("T")
    z = np.zeros(rBo.shape)
    ops = OS.update_cg2d(P, 0.25, 0.51, g, Tm, hW, hS, rBo, z, z, z)
    ops0 = OS.update_cg2d(P, 0.25, 0.51, g, Tm, np.zeros_like(hW), np.zeros_like(hS), rBo, z, z, z)
    bars = [rng.normal(size=rBo.shape) for _ in range(3)]

    def check(mod):
        hWb, hSb = mod.update_cg2d_adjoint(P, 0.25, g, Tm, *bars, K)
        lhs = sum(float((b * (o - o0)).sum()) for b, o, o0 in zip(bars, ops[:3], ops0[:3]))
        rhs = float((hWb * hW).sum() + (hSb * hS).sum())
        return abs(lhs - rhs) < 1e-11 * max(abs(lhs), 1.0)

    assert check(OS)
    assert not check(_planted("    Cw[:, :, 1:] = C[:, :, :-1]", "    Cw[:, :, :-1] = C[:, :, 1:]"))
