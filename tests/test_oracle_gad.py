"""The CPU oracle of GAD DST3 pinned by invariants and closed forms (CPU only): tile topology, the library's pass
plan against the oracle's tables, a constant tracer, global conservation, the all-wet facet-corner case and the
planted errors of the reference's tests (test_gad.py:145-294), linearity, and exactness of the DST3 flux for
quadratic profiles (DST3 is the exact semi-Lagrangian flux of the quadratic through three cell averages), in x, y
and — for the vertical flux this library adds — in the vertical."""

import numpy as np
import pytest

import oracle
from oracle import tables as OT
from tests.conftest import dst3_case, maps, oracle_tables

T, K, NY, NX, OL, N = 13, 50, 98, 98, 4, 90
J = slice(OL, OL + N)
I = slice(OL, OL + N)
DT = 3600.0


def test_topology_llc90():
    topo = OT.Topology.from_facets(OT.LLC90_FACETS)
    assert topo.face == (1, 1, 1, 2, 2, 2, 3, 4, 4, 4, 5, 5, 5)
    assert topo.S == (True, False, False, True, False, False, True, True, True, True, True, True, True)
    assert topo.N == (False, False, True, False, False, True, True, True, True, True, True, True, True)
    assert topo.W == (True, True, True, True, True, True, True, True, False, False, True, False, False)
    assert topo.E == (True, True, True, True, True, True, True, False, False, True, False, False, True)
    from mit_jax_amd.gad import TileTopology
    p = TileTopology.from_facets()
    assert (p.face, p.N_edge, p.S_edge, p.E_edge, p.W_edge) == (topo.face, topo.N, topo.S, topo.E, topo.W)


def test_library_plan_equals_oracle_tables():
    """The library's update boxes and fill calls (C++) rasterise to the oracle's masks and fill tables."""
    from mit_jax_amd import gad
    masks, fills = gad.plan_tables_host(gad.TileTopology.from_facets())
    tab = oracle_tables()
    assert np.array_equal(masks, tab["masks"])
    assert np.array_equal(fills, tab["fill_dirs"])


def _run(g, u, v, tr, tab=None, one_sixth=1.0 / 6.0, w=None, vgrid=None):
    from mit_jax_amd import synthetic as S
    g2, vg = S.gad_inputs(g)
    if vgrid is not None:
        vg = vgrid
    tab = tab or oracle_tables()
    if w is None:
        return oracle.advect_tiles(tab, g2, vg, u, v, tr, g.hFacW, g.hFacS, g.recip_hFacC, one_sixth)
    return oracle.gad_dst3(tab, g2, vg, u, v, w, tr, g.hFacW, g.hFacS, g.recip_hFacC, one_sixth)


def _conservation(g, u, v, gT, tr):
    """sum_wet V gT dt - dt sum_wet tracer div_h(U) over tile interiors, and its scale (test_gad.py:178-194)."""
    V = g.rA[:, None] * g.drF[None, :, None, None] * g.hFacC
    xA = g.dyG[:, None] * g.drF[None, :, None, None] * g.hFacW
    yA = g.dxG[:, None] * g.drF[None, :, None, None] * g.hFacS
    uT, vT = u * xA, v * yA
    div = (uT[..., J, OL + 1:OL + N + 1] - uT[..., J, I]) + (vT[..., OL + 1:OL + N + 1, I] - vT[..., J, I])
    wet = g.maskC[..., J, I] > 0
    a = (V[..., J, I] * gT[..., J, I] * DT)[wet]
    b = (tr[..., J, I] * div * DT)[wet]
    return a.sum() - b.sum(), np.abs(a).sum() + np.abs(b).sum()


def _facet3_as_facet1():
    topo = OT.Topology.from_facets(OT.LLC90_FACETS)
    bad = OT.Topology([1 if t == 6 else f for t, f in enumerate(topo.face)], topo.N, topo.S, topo.E, topo.W)
    return OT.tables(bad)


def test_constant_tracer_and_conservation():
    g, u, v, w, tr = dst3_case(0)
    c = np.full_like(tr, 35.0)
    gc = _run(g, u, v, c)
    assert np.abs(gc[..., J, I]).max() * DT / 35.0 < 1e-13
    gT = _run(g, u, v, tr)
    r, s = _conservation(g, u, v, gT, tr)
    assert abs(r) / s < 1e-12
    # planted error: facet 3 swept in facet 1's order (test_gad.py:211-217)
    r_bad, _ = _conservation(g, u, v, _run(g, u, v, tr, tab=_facet3_as_facet1()), tr)
    assert abs(r_bad) / s > 1e-10


def test_planted_errors_change_the_tendency():
    """oneSixth x (1 + 1e-6); facet 5's pass-2 overlap update dropped; one row missing from pass 1's X regions
    (test_gad.py:145-174): each changes gTracer."""
    g, u, v, w, tr = dst3_case(0)
    ref = _run(g, u, v, tr)
    assert not np.array_equal(_run(g, u, v, tr, one_sixth=(1.0 / 6.0) * (1 + 1e-6)), ref)
    tab = dict(oracle_tables())
    m = tab["masks"].copy()
    m[1, 1] = 0
    tab["masks"] = m
    assert not np.array_equal(_run(g, u, v, tr, tab=tab), ref)
    tab = dict(oracle_tables())
    m = tab["masks"].copy().reshape(3, 2, T, NY, NX)
    m[0, 0, :, OL + N - 1] = 0
    tab["masks"] = m.reshape(3, 2, T, NY * NX)
    assert not np.array_equal(_run(g, u, v, tr, tab=tab), ref)


def _allwet():
    """Every cell wet (hFac = 1, masks 1 on every lane), flow uniform in tile coordinates varying with k, halos by
    the signed vector exchange, no flow through the open facet edges; facet-corner halos marked
    (test_gad.py:234-266)."""
    from mit_jax_amd import synthetic as S
    g0 = S.make_grid(0)
    ones = np.ones((T, K, NY, NX))
    g = S.Grid(g0.L, dict(g0.f), dict(g0.v))
    g.f.update(maskW=ones, maskS=ones, maskC=ones, maskInC=np.ones((T, NY, NX)), hFacW=ones, hFacS=ones,
               hFacC=ones, recip_hFacC=ones)
    k = np.arange(K)[None, :, None, None]
    u = np.zeros((T, K, NY, NX))
    v = np.zeros((T, K, NY, NX))
    u[..., J, I] = np.broadcast_to(0.3 * np.cos(0.1 * k), u.shape)[..., J, I]
    v[..., J, I] = np.broadcast_to(-0.2 + 0.01 * k, v.shape)[..., J, I]
    topo = OT.Topology.from_facets(OT.LLC90_FACETS)
    for t in range(T):
        if topo.face[t] in (1, 2) and topo.S[t]:
            v[t, :, OL, :] = 0.0
    u, v = oracle.vector(maps().full("UVs_u"), maps().full("UVs_v"), u, v)
    corners = np.zeros((T, NY, NX), bool)
    for t in range(T):
        sw, se, nw, ne = topo.corners(t)
        for flag, jj, ii in ((sw, slice(0, OL), slice(0, OL)), (se, slice(0, OL), slice(-OL, None)),
                             (nw, slice(-OL, None), slice(0, OL)), (ne, slice(-OL, None), slice(-OL, None))):
            if flag:
                corners[t, jj, ii] = True
    return g, u, v, corners


def test_facet_corners_all_wet():
    """FILL_CS_CORNER_TR_RL / _UV_RS at work (test_gad.py:269-294): a constant with zero facet-corner halos stays
    constant in every tile interior only with the fills; a random exchanged tracer is conserved only with them."""
    g, u, v, corners = _allwet()
    tab = oracle_tables()
    nofill = dict(tab)
    fl = list(tab["fills"])
    for p, f in ((0, 2), (0, 3), (1, 0), (1, 2)):  # fYb1, fYa1, fXb2, fYb2
        fl[p * 4 + f] = None
    nofill["fills"] = fl
    tr0 = np.where(corners[:, None], 0.0, np.full((T, K, NY, NX), 12.0))
    e = np.abs(_run(g, u, v, tr0, tab=tab)[..., J, I]).max() * DT / 12.0
    e_bad = np.abs(_run(g, u, v, tr0, tab=nofill)[..., J, I]).max() * DT / 12.0
    assert e < 1e-13 and e_bad > 1e-6, (e, e_bad)
    rnd = np.random.default_rng(3).normal(size=tr0.shape)
    rnd = np.where(corners[:, None], 0.0, oracle.scalar(maps().full("T"), rnd))
    r, s = _conservation(g, u, v, _run(g, u, v, rnd, tab=tab), rnd)
    r_bad, _ = _conservation(g, u, v, _run(g, u, v, rnd, tab=nofill), rnd)
    assert abs(r) / s < 1e-12 and abs(r_bad) / s > 1e-10, (r / s, r_bad / s)


def test_linear_in_tracer():
    g, u, v, w, tr = dst3_case(0)
    rng = np.random.default_rng(8)
    t2 = np.where(g.maskC > 0, rng.normal(size=tr.shape), 0.0)
    a = _run(g, u, v, 2.0 * tr + 3.0 * t2)
    b = 2.0 * _run(g, u, v, tr) + 3.0 * _run(g, u, v, t2)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def _uniform_grid(dz=10.0):
    """All-wet grid with uniform metrics (dx = dy = 1e5 m, dz) for the closed-form tests."""
    from mit_jax_amd import synthetic as S
    g0 = S.make_grid(0)
    ones = np.ones((T, K, NY, NX))
    o2 = np.ones((T, NY, NX))
    f = dict(g0.f)
    f.update(maskW=ones, maskS=ones, maskC=ones, maskInC=o2, hFacW=ones, hFacS=ones, hFacC=ones, recip_hFacC=ones,
             dxG=o2 * 1e5, dyG=o2 * 1e5, dxC=o2 * 1e5, dyC=o2 * 1e5, recip_dxC=o2 * 1e-5, recip_dyC=o2 * 1e-5,
             rA=o2 * 1e10, recip_rA=o2 * 1e-10)
    v = dict(g0.v)
    v.update(drF=np.full(K, dz), recip_drF=np.full(K, 1.0 / dz), drC=np.full(K, dz), recip_drC=np.full(K, 1.0 / dz))
    return S.Grid(g0.L, f, v)


def _quad_avg(s, a=3.0, b=0.7, c=-0.02):
    """Cell average over [s - 1/2, s + 1/2] of a + b x + c x^2."""
    return a + b * s + c * (s * s + 1.0 / 12.0)


@pytest.mark.parametrize("direction,vel", [("x", 0.3), ("x", -0.25), ("y", 0.2), ("y", -0.3)])
def test_quadratic_profile_advected_exactly(direction, vel):
    """Uniform flow with Courant number C: DST3 moves the cell averages of a quadratic exactly,
    new(i) = avg over [i - 1/2 - C, i + 1/2 - C] (interior points away from tile edges)."""
    g = _uniform_grid()
    C = vel
    U = np.full((T, K, NY, NX), C * 1e5 / DT)  # |U| dt / dx = |C|
    Z = np.zeros_like(U)
    idx = np.arange(NX if direction == "x" else NY, dtype=np.float64)
    prof = _quad_avg(idx)
    if direction == "x":
        tr = np.broadcast_to(prof[None, None, None, :], U.shape).copy()
        gT = _run(g, U, Z, tr)
        exact = (_quad_avg(idx - C) - prof)[None, None, None, :] / DT
        sl = (slice(None), slice(None), J, slice(OL + 3, OL + N - 3))
    else:
        tr = np.broadcast_to(prof[None, None, :, None], U.shape).copy()
        gT = _run(g, Z, U, tr)
        exact = (_quad_avg(idx - C) - prof)[None, None, :, None] / DT
        sl = (slice(None), slice(None), slice(OL + 3, OL + N - 3), I)
    assert abs(C) <= 0.35
    err = np.abs(gT - exact)[sl].max()
    assert err <= 1e-13 * np.abs(prof).max() / DT * 100, err


@pytest.mark.parametrize("wvel", [1e-3, -7e-4])
def test_vertical_quadratic_exact_constant_and_column_budget(wvel):
    """The vertical DST3 flux (not in the reference; defined by this library): a quadratic in the upward
    coordinate is moved exactly by a uniform w (levels away from surface and bottom); a constant stays constant;
    the column integral of the vertical update telescopes to sum_k tracer(k)(rTrans(k) - rTrans(k+1))."""
    g = _uniform_grid(dz=10.0)
    C = wvel * DT / 10.0
    Z = np.zeros((T, K, NY, NX))
    W = np.full_like(Z, wvel)
    s = -np.arange(K, dtype=np.float64)          # upward coordinate of the level centres
    prof = _quad_avg(s)
    tr = np.broadcast_to(prof[None, :, None, None], Z.shape).copy()
    gT = _run(g, Z, Z, tr, w=W, vgrid=None)
    exact = (_quad_avg(s - C) - prof) / DT
    err = np.abs(gT[:, 3:K - 3] - exact[None, 3:K - 3, None, None]).max()
    assert err <= 1e-12 * np.abs(prof).max() / DT, err
    c = np.full_like(Z, 7.5)
    gc = _run(g, Z, Z, c, w=W)
    assert np.abs(gc).max() * DT / 7.5 < 1e-13
    # column budget: sum_k V gT dt = sum_k tracer(k) (rT(k) - rT(k+1)) dt with rT(0) = rT(K) = 0
    rT = np.zeros((K + 1,))
    rT[1:K] = wvel * 1e10
    lhs = (1e10 * 10.0 * gT[0, :, 50, 50] * DT).sum()
    rhs_ = (tr[0, :, 50, 50] * (rT[:-1] - rT[1:]) * DT).sum()
    assert abs(lhs - rhs_) <= 1e-12 * (np.abs(1e10 * 10.0 * gT[0, :, 50, 50] * DT).sum() + 1.0)
