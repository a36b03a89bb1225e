"""The CPU oracle of GAD_IMPLICIT_R (oracle/implicit_r.py) pinned by what the mathematics fixes (CPU only): the
pentadiagonal elimination against a dense solve, the matrix against a per-interface loop in the Fortran order of
gad_u3c4_impl_r.F (interfaces k = Nr..2 updating levels k and k-1), conservation of the column integral of
hFac drF T under implicit diffusion and advection, identity halo columns, and planted errors that break these."""

from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import implicit_r as OI

T, NY, NX, OL, N = 2, 10, 12, 2, None


def _case(seed, Nr=12, sNx=8, sNy=6, OL=2, dry_bottom=True, mask_w=True):
    rng = np.random.default_rng(seed)
    ny, nx = sNy + 2 * OL, sNx + 2 * OL
    drF = rng.uniform(5.0, 300.0, Nr)
    drC = np.empty(Nr)
    drC[0] = 0.5 * drF[0]
    drC[1:] = 0.5 * (drF[:-1] + drF[1:])
    nwet = rng.integers(2, Nr + 1, size=(T, ny, nx)) if dry_bottom else np.full((T, ny, nx), Nr)
    k = np.arange(Nr)[None, :, None, None]
    maskC = (k < nwet[:, None]).astype(np.float64)
    hFac = np.where(maskC > 0, rng.uniform(0.3, 1.0, (T, Nr, ny, nx)), 0.0)
    rh = np.where(hFac > 0, 1.0 / np.where(hFac > 0, hFac, 1.0), 0.0)
    rA = rng.uniform(0.5e10, 1.5e10, (T, ny, nx))
    g = SimpleNamespace(maskC=maskC, rA=rA, recip_rA=1.0 / rA, recip_drF=1.0 / drF, recip_drC=1.0 / drC)
    p = SimpleNamespace(dTtracerLev=np.full(Nr, 3600.0), implicitDiffusion=True, rkSign=-1.0)
    kap = rng.uniform(1e-5, 1e-1, (T, Nr, ny, nx))
    # w at interfaces k (index k = top of level k), zero at the surface and where either neighbour level is dry
    w = rng.normal(size=(T, Nr, ny, nx)) * 1e-3
    if mask_w:
        w[:, 0] = 0.0
        w[:, 1:] *= maskC[:, 1:] * maskC[:, :-1]
    y = rng.normal(size=(T, Nr, ny, nx)) * 3.0 + 10.0
    return p, g, kap, rh, w, y, hFac, drF, dict(sNx=sNx, sNy=sNy, OL=OL)


def test_pentadiagonal_elimination_matches_a_dense_solve():
    rng = np.random.default_rng(0)
    for Nr in (2, 3, 5, 17):
        shp = (2, Nr, 3, 4)
        a, b, d, e = (rng.normal(size=shp) * 0.2 for _ in range(4))
        a[:, :2] = 0.0
        b[:, :1] = 0.0
        d[:, -1:] = 0.0
        e[:, -2:] = 0.0
        c = 2.0 + rng.random(shp)
        y = rng.normal(size=shp)
        u = OI.solve_pentadiagonal(a, b, c, d, e, y)
        for t in range(2):
            for j in range(3):
                for i in range(4):
                    M = np.zeros((Nr, Nr))
                    for k in range(Nr):
                        M[k, k] = c[t, k, j, i]
                        if k >= 1:
                            M[k, k - 1] = b[t, k, j, i]
                        if k >= 2:
                            M[k, k - 2] = a[t, k, j, i]
                        if k + 1 < Nr:
                            M[k, k + 1] = d[t, k, j, i]
                        if k + 2 < Nr:
                            M[k, k + 2] = e[t, k, j, i]
                    ref = np.linalg.solve(M, y[t, :, j, i])
                    assert np.allclose(u[t, :, j, i], ref, rtol=1e-12, atol=1e-13), (Nr, t, j, i)


def _loop_matrix(p, g, implicitAdvection, kap, rh, w, sNx, sNy, OL):
    """gad_implicit_r.F / gad_u3c4_impl_r.F as scalar loops, 1-based k as in Fortran: diffusion rows, then for
    interfaces k = Nr..2 the updates of level k and of level k-1."""
    Tn, Nr, ny, nx = kap.shape
    A = np.zeros((5,) + kap.shape)
    A[2] = 1.0
    dt, rdrF, rdrC = p.dTtracerLev, g.recip_drF, g.recip_drC
    s6 = 1.0 / 6.0
    for t in range(Tn):
        for j in range(OL, OL + sNy):
            for i in range(OL, OL + sNx):
                mC = g.maskC[t, :, j, i]
                r = rh[t, :, j, i]
                kk = kap[t, :, j, i]
                a, b, c, d, e = (np.zeros(Nr) for _ in range(5))
                for k in range(2, Nr + 1):
                    b[k - 1] = -(((((dt[k - 1] * mC[k - 2]) * r[k - 1]) * rdrF[k - 1]) * kk[k - 1]) * rdrC[k - 1])
                for k in range(1, Nr):
                    d[k - 1] = -(((((dt[k - 1] * mC[k]) * r[k - 1]) * rdrF[k - 1]) * kk[k]) * rdrC[k])
                for k in range(1, Nr + 1):
                    c[k - 1] = 1.0 - (b[k - 1] + d[k - 1])
                if implicitAdvection:
                    for k in range(Nr, 1, -1):
                        rTrans = (w[t, k - 1, j, i] * g.rA[t, j, i]) * mC[k - 2]
                        rC = ((0.5 * rTrans) * g.recip_rA[t, j, i]) * p.rkSign
                        mM = mC[max(1, k - 2) - 1] * (0.0 if k <= 2 else 1.0)
                        mP = mC[min(Nr, k + 1) - 1] * (0.0 if k >= Nr else 1.0)
                        rUp = (2.0 * s6) * abs(rC)
                        km = (s6 * (rC + abs(rC))) * mM
                        kp = (s6 * (rC - abs(rC))) * mP

                        def sc(x, lev):
                            return ((x * dt[lev - 1]) * r[lev - 1]) * rdrF[lev - 1]
                        # level k (interface k is its top)
                        a[k - 1] = a[k - 1] + sc(km, k)
                        b[k - 1] = b[k - 1] - sc((rC + rUp) + km, k)
                        c[k - 1] = c[k - 1] - sc((rC - rUp) + kp, k)
                        d[k - 1] = d[k - 1] + sc(kp, k)
                        # level k-1 (interface k is its bottom)
                        b[k - 2] = b[k - 2] - sc(km, k - 1)
                        c[k - 2] = c[k - 2] + sc((rC + rUp) + km, k - 1)
                        d[k - 2] = d[k - 2] + sc((rC - rUp) + kp, k - 1)
                        e[k - 2] = e[k - 2] - sc(kp, k - 1)
                for n, v in enumerate((a, b, c, d, e)):
                    A[n, t, :, j, i] = v
    return A


@pytest.mark.parametrize("adv", [False, True])
def test_matrix_equals_the_fortran_order_loop(adv):
    p, g, kap, rh, w, y, hFac, drF, lay = _case(1, mask_w=False)
    got = OI.matrix(p, g, adv, kap, rh, w, **lay)
    ref = _loop_matrix(p, g, adv, kap, rh, w, **lay)
    for n in range(5):
        assert np.array_equal(got[n], ref[n]), "abcde"[n]


def _conserves(mod, adv=True, seed=2):
    p, g, kap, rh, w, y, hFac, drF, lay = _case(seed)
    u = mod.gad_implicit_r(p, g, adv, kap, rh, w, y, **lay)
    J, I = slice(lay["OL"], lay["OL"] + lay["sNy"]), slice(lay["OL"], lay["OL"] + lay["sNx"])
    wt = hFac * drF[None, :, None, None]
    before = (wt * y)[:, :, J, I].sum(axis=1)
    after = (wt * u)[:, :, J, I].sum(axis=1)
    return np.allclose(after, before, rtol=1e-13, atol=0)


@pytest.mark.parametrize("adv", [False, True])
def test_column_integral_is_conserved(adv):
    assert _conserves(OI, adv)


def test_halo_columns_are_unchanged_and_single_level_is_identity():
    p, g, kap, rh, w, y, hFac, drF, lay = _case(3)
    u = OI.gad_implicit_r(p, g, True, kap, rh, w, y, **lay)
    halo = np.ones(y.shape[2:], bool)
    halo[lay["OL"]:lay["OL"] + lay["sNy"], lay["OL"]:lay["OL"] + lay["sNx"]] = False
    assert np.array_equal(u[:, :, halo], y[:, :, halo])
    one = OI.gad_implicit_r(p, g, True, kap[:, :1], rh[:, :1], w[:, :1], y[:, :1], **lay)
    assert np.array_equal(one, y[:, :1])


def test_pure_diffusion_of_a_constant_is_the_constant():
    p, g, kap, rh, w, y, hFac, drF, lay = _case(4, dry_bottom=False)
    u = OI.gad_implicit_r(p, g, False, kap, rh, w, np.full_like(y, 7.0), **lay)
    assert np.allclose(u, 7.0, rtol=1e-14, atol=0)


def _planted(old, new):
    src = Path(OI.__file__).read_text()
    assert src.count(old) == 1, old
    ns = {"__name__": "oracle.implicit_r_planted", "__package__": "oracle"}
    exec(compile(src.replace(old, new), "implicit_r_planted.py", "exec"), ns)
    return SimpleNamespace(**ns)


@pytest.mark.parametrize("old,new", [
    ("* kap[:, 1:]) * rdrC[:, 1:])\n    # :59", "* kap[:, :-1]) * rdrC[:, 1:])\n    # :59"),
    ("    c = 1.0 - (b + d)", "    c = 1.0 - (b - d)"),
    ("b[:, lo] = b[:, lo] - sc(rC4km, lo)", "b[:, lo] = b[:, lo] + sc(rC4km, lo)"),
    ("a[:, up] = a[:, up] + sc(rC4km, up)", "a[:, up] = a[:, up] + sc(rC4kp, up)"),
    ("        yq = (y[:, k] - a[:, k] * yp[:, k - 2]) - bb * yp[:, k - 1]",
     "        yq = (y[:, k] - a[:, k] * yp[:, k - 1]) - bb * yp[:, k - 1]"),
])
def test_planted_errors_break_conservation(old, new):
    mod = _planted(old, new)
    assert not _conserves(mod, True)


@pytest.mark.parametrize("old,new", [
    ("rTrans = (wFld[:, 1:, J, I] * g.rA[:, None, J, I]) * mC[:, :-1]",
     "rTrans = (wFld[:, 1:, J, I] * g.rA[:, None, J, I]) * mC[:, 1:]"),
    ("kp1 = np.minimum(Nr, ks + 1) - 1", "kp1 = np.minimum(Nr, ks) - 1"),
    ("maskM2 = np.where(ks <= 2, 0.0, 1.0)", "maskM2 = np.where(ks <= 1, 0.0, 1.0)"),
    ("rUpwind = (2.0 * ONE_SIXTH) * np.abs(rCenter)", "rUpwind = ONE_SIXTH * np.abs(rCenter)"),
])
def test_planted_errors_break_the_fortran_loop_match(old, new):
    p, g, kap, rh, w, y, hFac, drF, lay = _case(1, mask_w=False)
    got = _planted(old, new).matrix(p, g, True, kap, rh, w, **lay)
    ref = _loop_matrix(p, g, True, kap, rh, w, **lay)
    assert not all(np.array_equal(got[n], ref[n]) for n in range(5))


def _adjoint_pins(mod, seed=7):
    """The reverse mode's pins: the dot-product identity for the tracer cotangent (the solve is linear in y) and
    central differences of <ubar, u> in kappaRX and wFld at interior points; returns the failed pins."""
    p, g, kap, rh, w, y, hFac, drF, lay = _case(seed)
    u = mod.gad_implicit_r(p, g, True, kap, rh, w, y, **lay)
    rng = np.random.default_rng(seed + 1)
    ub = rng.normal(size=u.shape)
    yb, kb, wb = mod.gad_implicit_r_adjoint(p, g, True, kap, rh, w, u, ub, **lay)
    bad = []
    L = lambda kap_, w_, y_: float((ub * mod.gad_implicit_r(p, g, True, kap_, rh, w_, y_, **lay)).sum())  # noqa: E731
    y2 = rng.normal(size=y.shape)
    if abs(L(kap, w, y2) - float((yb * y2).sum())) > 1e-12 * abs(L(kap, w, y2)):
        bad.append("dot")
    pts = [(0, 3, 3, 4), (0, 1, 5, 6), (1, 8, 2, 3), (1, 2, 6, 9), (0, 6, 7, 8)]
    for name, arr, gb in (("kap", kap, kb), ("w", w, wb)):
        for pt in pts:
            if g.maskC[pt] == 0:
                continue
            errs = []
            for rel in (1e-4, 1e-5, 1e-6):
                h = rel * max(abs(arr[pt]), 1e-8)
                e = np.zeros_like(arr)
                e[pt] = h
                if name == "kap":
                    fd = (L(kap + e, w, y) - L(kap - e, w, y)) / (2 * h)
                else:
                    fd = (L(kap, w + e, y) - L(kap, w - e, y)) / (2 * h)
                errs.append(abs(fd - gb[pt]) / max(abs(gb[pt]), abs(fd), 1e-30))
            if min(errs) > 1e-5:
                bad.append((name, pt, min(errs)))
    return bad


def test_adjoint_dot_product_and_finite_differences():
    assert _adjoint_pins(OI) == []


@pytest.mark.parametrize("old,new", [
    ("    kb[:, 1:] = -((li[:, 1:] * beta) * (-du) + (li[:, :-1] * delta) * du)",
     "    kb[:, 1:] = -((li[:, 1:] * beta) * (-du) - (li[:, :-1] * delta) * du)"),
    ("        dkm = (ONE_SIXTH * (1.0 + sg)) * mskM", "        dkm = (ONE_SIXTH * (1.0 - sg)) * mskM"),
    ("    bT[:, 1:] = d[:, :-1]", "    bT[:, 1:] = b[:, 1:]"),
    ("        rCb = -((li[:, :-1] * s[:, :-1]) - (li[:, 1:] * s[:, 1:])) * phi",
     "        rCb = -((li[:, :-1] * s[:, 1:]) - (li[:, 1:] * s[:, :-1])) * phi"),
])
def test_adjoint_planted_errors_fail_the_pins(old, new):
    assert _adjoint_pins(_planted(old, new)) != []
