"""The oracle's INI_CG2D normalisation and standalone DST3 interface fluxes, pinned by what the mathematics fixes
(CPU only).

ini_cg2d_norm (reference core/cg2d.py:108-134): on dyadic data every product and sum is exact, so the result must
equal the closed form 1 / max(|fac dyG recip_dxC sum_k drF hFacW|, |... aS ...|) over the interior; large halo
values, a factor fac != 1 and distinct hFacW / hFacS make a dropped factor, a halo point in the max or a swapped
operand change the result.

dst3_adv_x / dst3_adv_y (reference pkgs/gad.py:409-453): DST3 is the exact flux of the quadratic through three cell
averages, so for cell averages of a quadratic and a uniform Courant number C the flux through a face is the
transport times the mean of the quadratic over the swept interval (closed form); a linear tracer with one face
mask cleared pins which mask goes with which difference (the cleared face drops d0 from its own flux and d1 from
the downstream neighbour's); points outside the reference's loop range are 0.
"""

from types import SimpleNamespace

import numpy as np
import pytest

import oracle
from oracle import sfp as OS

T, K, NY, NX, OL, N = 13, 6, 98, 98, 4, 90


# ------------------------------------------------------------------------------------------------ ini_cg2d_norm
def _norm_case(seed=0):
    rng = np.random.default_rng(seed)
    g = SimpleNamespace(dyG=rng.integers(1, 8, (T, NY, NX)).astype(np.float64),
                        dxG=rng.integers(1, 8, (T, NY, NX)).astype(np.float64),
                        recip_dxC=2.0 ** -rng.integers(0, 3, (T, NY, NX)),
                        recip_dyC=2.0 ** -rng.integers(0, 3, (T, NY, NX)),
                        drF=rng.integers(1, 9, K).astype(np.float64))
    hW = rng.integers(0, 5, (T, K, NY, NX)) / 4.0
    hS = rng.integers(0, 5, (T, K, NY, NX)) / 4.0
    inner = np.zeros((NY, NX), bool)
    inner[OL:OL + N, OL:OL + N] = True
    hW[:, :, ~inner] = 64.0  # halo values far above any interior sum: they must not enter the max
    hS[:, :, ~inner] = 64.0
    return g, hW, hS


def _norm_closed_form(g, hW, hS, fac):
    J, I = slice(OL, OL + N), slice(OL, OL + N)
    sW = np.einsum("k,tkji->tji", g.drF, hW[:, :, J, I])  # exact: integers and quarters
    sS = np.einsum("k,tkji->tji", g.drF, hS[:, :, J, I])
    aW = fac * g.dyG[:, J, I] * g.recip_dxC[:, J, I] * sW
    aS = fac * g.dxG[:, J, I] * g.recip_dyC[:, J, I] * sS
    return 1.0 / max(np.abs(aW).max(), np.abs(aS).max())


def test_ini_cg2d_norm_closed_form_and_planted_errors():
    g, hW, hS = _norm_case()
    isp, idf = 0.5, 4.0
    got = OS.ini_cg2d_norm(g, hW, hS, isp, idf)
    cf = _norm_closed_form(g, hW, hS, isp * idf)
    assert got == cf
    assert OS.ini_cg2d_norm(g, hW, hS, 1.0, 1.0) != cf            # factor dropped
    assert OS.ini_cg2d_norm(g, hS, hW, isp, idf) != cf            # hFacW / hFacS swapped
    g2 = SimpleNamespace(**{**vars(g), "recip_dxC": g.recip_dyC, "recip_dyC": g.recip_dxC})
    assert OS.ini_cg2d_norm(g2, hW, hS, isp, idf) != cf           # recip_dxC / recip_dyC swapped
    # a halo point in the max would dominate it (halo hFac 64): the closed form over one more ring differs
    J1 = slice(OL - 1, OL + N + 1)
    g3 = SimpleNamespace(**{k: (v[:, J1, J1] if np.ndim(v) == 3 else v) for k, v in vars(g).items()})
    aw_ring = (isp * idf) * g3.dyG * g3.recip_dxC * np.einsum("k,tkji->tji", g.drF, hW[:, :, J1, J1])
    assert 1.0 / np.abs(aw_ring).max() < cf


def test_ini_cg2d_norm_zero_operator():
    g, hW, hS = _norm_case(1)
    z = np.zeros_like(hW)
    assert OS.ini_cg2d_norm(g, z, z, 1.0, 1.0) == 1.0


def test_synthetic_norm_equals_oracle():
    """The synthetic-input builder's norm (setup code) is the oracle's, factor included."""
    from mit_jax_amd import synthetic as S
    g = S.make_grid(0)
    assert S.ini_cg2d_norm(g) == OS.ini_cg2d_norm(g, g.hFacW, g.hFacS, 1.0, 1.0)
    assert S.ini_cg2d_norm(g, 0.5, 0.25) == OS.ini_cg2d_norm(g, g.hFacW, g.hFacS, 0.5, 0.25)


# ------------------------------------------------------------------------------------------------ DST3 fluxes
def _P(x, a, b, c):
    return a * x + b * x * x / 2.0 + c * x ** 3 / 3.0


def _quad_avg(s, a=3.0, b=0.7, c=-0.02):
    return a + b * s + c * (s * s + 1.0 / 12.0)


@pytest.mark.parametrize("direction", ["x", "y"])
def test_dst3_flux_exact_for_quadratics(direction):
    """Flux through the face at s = i - 1/2: trans * (mean of the quadratic over the interval swept in one step)."""
    a, b, c = 3.0, 0.7, -0.02
    Kf, ny, nx = 2, 14, 16
    Cs = np.array([0.3, -0.25])
    dt = np.array([3600.0, 1800.0])
    rd = np.full((1, ny, nx), 1e-5)
    vel = (Cs / (dt * 1e-5))[None, :, None, None] * np.ones((1, Kf, ny, nx))
    trans = np.sign(Cs)[None, :, None, None] * 5.0 * np.ones((1, Kf, ny, nx))
    n = nx if direction == "x" else ny
    idx = np.arange(n, dtype=np.float64)
    prof = _quad_avg(idx, a, b, c)
    tr = prof[None, None, None, :] if direction == "x" else prof[None, None, :, None]
    tr = np.broadcast_to(tr, (1, Kf, ny, nx)).copy()
    got = oracle.dst3_adv(direction, trans, vel, np.ones_like(tr), tr, rd, dt, np.ones(Kf))
    for k, C in enumerate(Cs):
        s = idx - 0.5
        lo, hi = (s - C, s) if C > 0 else (s, s - C)
        exact = 5.0 * np.sign(C) * (_P(hi, a, b, c) - _P(lo, a, b, c)) / abs(C)
        g = got[0, k] if direction == "x" else got[0, k].T
        rows = slice(None)
        assert np.allclose(g[rows, 2:n - 1], exact[None, 2:n - 1], rtol=1e-13, atol=0), (direction, k)
        assert not g[:, :2].any() and not g[:, n - 1:].any()  # outside i = 2 .. n-2


@pytest.mark.parametrize("direction,C", [("x", 0.3), ("x", -0.3), ("y", 0.2), ("y", -0.2)])
def test_dst3_flux_mask_placement(direction, C):
    """Linear tracer t = i, face mask cleared at i0: for C > 0 the flux U (i - 1 + d0 m(i) + d1 m(i-1)) loses d0
    at i0 and d1 at i0 + 1; for C < 0 the flux U (i - d0 m(i) - d1 m(i+1)) loses d0 at i0 and d1 at i0 - 1."""
    ny, nx = 12, 12
    dt = np.array([3600.0])
    n = nx if direction == "x" else ny
    idx = np.arange(n, dtype=np.float64)
    tr = idx[None, None, None, :] if direction == "x" else idx[None, None, :, None]
    tr = np.broadcast_to(tr, (1, 1, ny, nx)).copy()
    vel = np.full((1, 1, ny, nx), C / (dt[0] * 1e-5))
    U = 2.0 * np.sign(C)
    trans = np.full_like(vel, U)
    i0 = 6
    mask = np.ones_like(tr)
    if direction == "x":
        mask[..., i0] = 0.0
    else:
        mask[..., i0, :] = 0.0
    got = oracle.dst3_adv(direction, trans, vel, mask, tr, np.full((1, ny, nx), 1e-5), dt, np.ones(1))
    g = got[0, 0] if direction == "x" else got[0, 0].T
    cfl = abs(C)
    d0 = (2.0 - cfl) * (1.0 - cfl) / 6.0
    d1 = (1.0 - cfl * cfl) / 6.0
    m = np.ones(n)
    m[i0] = 0.0
    exact = np.zeros(n)
    for i in range(2, n - 1):
        if C > 0:
            exact[i] = U * ((i - 1) + (d0 * m[i] + d1 * m[i - 1]))
        else:
            exact[i] = U * (i - (d0 * m[i] + d1 * m[i + 1]))
    assert np.allclose(g, exact[None, :], rtol=1e-14, atol=1e-14)
    # the two affected faces differ from their unmasked values
    full = U * ((idx - 1) + (d0 + d1)) if C > 0 else U * (idx - (d0 + d1))
    aff = (i0, i0 + 1) if C > 0 else (i0, i0 - 1)
    for i in aff:
        assert abs(g[0, i] - full[i]) > 0.1 * min(d0, d1)
