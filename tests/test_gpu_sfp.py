"""GPU parity of the CG2D operator and rhs producers (mj_update_cg2d, mj_cg2d_rhs) against the CPU oracle
(oracle/sfp.py), bitwise: both evaluate the reference's operation order (core/solve_for_pressure.py:29-106) without
fused multiply-adds, the level sums in the same sequential order."""

import ctypes
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import sfp as OS
from tests.conftest import maps, to_dev

pytestmark = pytest.mark.gpu

T, NY, NX, OL, N = 13, 98, 98, 4, 90


@pytest.fixture(scope="module")
def ex():
    from mit_jax_amd.exchange import default_exchanger
    return default_exchanger()


@pytest.fixture(scope="module")
def grid0():
    from mit_jax_amd import synthetic as S
    return S.make_grid(0)


def _fs(**kw):
    from mit_jax_amd.solve_for_pressure import FreeSurfParams
    base = dict(deltaTMom=3600.0, deltaTFreeSurf=3600.0, freeSurfFac=1.0, implicSurfPress=1.0, implicDiv2DFlow=1.0,
                mass2rUnit=1.0 / 1029.0, useRealFreshWaterFlux=True)
    base.update(kw)
    return FreeSurfParams(**base)


def _cp(norm):
    from mit_jax_amd.cg2d import Cg2dParams
    return Cg2dParams(cg2dMaxIters=300, cg2dTolerance=1e-12, cg2dNorm=norm)


def test_update_cg2d_bitwise(dev, ex, grid0):
    from mit_jax_amd import synthetic as S
    from mit_jax_amd import solve_for_pressure as SFP
    g = grid0
    norm = S.ini_cg2d_norm(g)
    rng = np.random.default_rng(10)
    rBo = np.full((T, NY, NX), 1.0 / S.G) * (1.0 + 0.01 * rng.normal(size=(T, NY, NX)))
    pW0, pS0, pC0 = (rng.normal(size=(T, NY, NX)) for _ in range(3))
    print("update_cg2d: seed grid, one member", flush=True)
    got = SFP.update_cg2d(_fs(), _cp(norm), g, ex, to_dev(g.hFacW, dev), to_dev(g.hFacS, dev), to_dev(rBo, dev),
                          to_dev(pW0, dev), to_dev(pS0, dev), to_dev(pC0, dev))
    ref = OS.update_cg2d(_fs(), norm, 0.51, g, maps().full("T"), g.hFacW, g.hFacS, rBo, pW0, pS0, pC0)
    for name, a, b in zip(("aW2d", "aS2d", "aC2d", "pW", "pS", "pC"), got, ref):
        assert np.array_equal(a.cpu().numpy(), b), name


def test_update_cg2d_batched_members(dev, ex):
    """Members with their own level counts (own hFac) and cg2dNorm, one launch: each equals its own oracle."""
    from mit_jax_amd import synthetic as S
    from mit_jax_amd import solve_for_pressure as SFP
    grids = [S.make_grid(0, nlev=n) for n in (15, 30, 50)]
    norms = [S.ini_cg2d_norm(gm) for gm in grids]
    g = grids[2]
    rBo = np.full((T, NY, NX), 1.0 / S.G)
    z = np.zeros((T, NY, NX))
    hW = np.stack([gm.hFacW for gm in grids])
    hS = np.stack([gm.hFacS for gm in grids])
    print("update_cg2d: 3 members, nlev 15/30/50", flush=True)
    import torch
    got = SFP.update_cg2d(_fs(), _cp(torch.tensor(norms, dtype=torch.float64, device=dev)), g, ex, to_dev(hW, dev),
                          to_dev(hS, dev), to_dev(rBo, dev), to_dev(z, dev), to_dev(z, dev), to_dev(z, dev))
    for m, gm in enumerate(grids):
        ref = OS.update_cg2d(_fs(), norms[m], 0.51, gm, maps().full("T"), gm.hFacW, gm.hFacS, rBo, z, z, z)
        for name, a, b in zip(("aW2d", "aS2d", "aC2d", "pW", "pS", "pC"), got, ref):
            assert np.array_equal(a[m].cpu().numpy(), b), (m, name)
    # the synthetic builder of the solve's operators is the same computation
    ops = S.update_cg2d(grids[1], norms[1], precond="5point")
    for name, a, b in zip(("aW2d", "aS2d", "aC2d", "pW", "pS", "pC"), got, ops):
        assert np.array_equal(a[1].cpu().numpy(), b), name


def _rhs_fields(rng, g, K=50):
    gU = rng.normal(size=(T, K, NY, NX)) * g.maskW
    gV = rng.normal(size=(T, K, NY, NX)) * g.maskS
    EmPmR = rng.normal(size=(T, NY, NX)) * 1e-5
    etaN = rng.normal(size=(T, NY, NX))
    etaH = rng.normal(size=(T, NY, NX))
    Bo = np.full((T, NY, NX), 9.81) * (1.0 + 0.01 * rng.normal(size=(T, NY, NX)))
    return gU, gV, EmPmR, etaN, etaH, Bo


@pytest.mark.parametrize("rfw", [True, False])
def test_cg2d_rhs_bitwise(dev, grid0, rfw):
    from mit_jax_amd import solve_for_pressure as SFP
    g = grid0
    rng = np.random.default_rng(11)
    gU, gV, EmPmR, etaN, etaH, Bo = _rhs_fields(rng, g)
    p = _fs(useRealFreshWaterFlux=rfw)
    print(f"cg2d_rhs: seed grid, useRealFreshWaterFlux={rfw}", flush=True)
    b, x = SFP.cg2d_rhs(p, g, to_dev(gU, dev), to_dev(gV, dev), to_dev(g.hFacW, dev), to_dev(g.hFacS, dev),
                        to_dev(EmPmR, dev) if rfw else None, to_dev(etaN, dev), to_dev(etaH, dev), to_dev(Bo, dev))
    rb, rx = OS.cg2d_rhs(p, g, gU, gV, g.hFacW, g.hFacS, EmPmR, etaN, etaH, Bo)
    assert np.array_equal(b.cpu().numpy(), rb)
    assert np.array_equal(x.cpu().numpy(), rx)


def test_cg2d_rhs_batched_and_vertical_factors(dev, grid0):
    """Through the C ABI with non-unit deepFacC, rhoFacC, deepFac2F (and kSurfC): 2 members."""
    import torch
    import mit_jax_amd._lib as L
    g = grid0
    K = 50
    rng = np.random.default_rng(12)
    fields = [_rhs_fields(rng, g) for _ in range(2)]
    dfc = 1.0 + 0.01 * rng.normal(size=K)
    rfc = 1.0 + 0.01 * rng.normal(size=K)
    d2f = 1.0 + 0.01 * rng.normal(size=K + 1)
    ks = OS.k_surf(g.maskC)
    p = _fs()
    keep = dict(dyG=g.dyG, dxG=g.dxG, recip_dxC=g.recip_dxC, recip_dyC=g.recip_dyC, rA=g.rA, maskInC=g.maskInC,
                drF=np.asarray(g.drF), dfc=dfc, rfc=rfc, d2f=d2f)
    kd = {k: to_dev(np.ascontiguousarray(v, np.float64), dev) for k, v in keep.items()}
    kd["ks"] = torch.as_tensor(ks, device=dev)
    gs = L.SfpGrid(*(kd[k].data_ptr() for k in ("dyG", "dxG", "recip_dxC", "recip_dyC", "rA", "maskInC", "ks", "drF",
                                                  "dfc", "rfc", "d2f")))
    cfg = L.SfpConfig(2, T, N, N, OL, K, T * K * NY * NX, T * NY * NX)
    fs = L.FsParams(p.deltaTMom, p.deltaTFreeSurf, p.freeSurfFac, p.implicSurfPress, p.implicDiv2DFlow, p.mass2rUnit, 1)
    st = [to_dev(np.stack([f[i] for f in fields]), dev) for i in range(5)]  # gU, gV, EmPmR, etaN, etaH
    hW = to_dev(np.stack([g.hFacW, g.hFacW]), dev)
    hS = to_dev(np.stack([g.hFacS, g.hFacS]), dev)
    Bo = to_dev(fields[0][5], dev)
    b = torch.empty((2, T, NY, NX), dtype=torch.float64, device=dev)
    x = torch.empty_like(b)
    P = L.ptr
    print("cg2d_rhs: 2 members, vertical factors through the C ABI", flush=True)
    L.check(L.lib().mj_cg2d_rhs(ctypes.byref(gs), ctypes.byref(fs), ctypes.byref(cfg), P(st[0]), P(st[1]), P(hW),
                                P(hS), P(st[2]), P(st[3]), P(st[4]), P(Bo), P(b), P(x), L.stream_ptr()), "rhs")
    for m in range(2):
        gU, gV, EmPmR, etaN, etaH, _ = fields[m]
        rb, rx = OS.cg2d_rhs(p, g, gU, gV, g.hFacW, g.hFacS, EmPmR, etaN, etaH, fields[0][5], deepFacC=dfc,
                             rhoFacC=rfc, deepFac2F=d2f, maskC=g.maskC)
        assert np.array_equal(b[m].cpu().numpy(), rb), m
        assert np.array_equal(x[m].cpu().numpy(), rx), m


def test_producers_feed_the_solve(dev, ex, grid0):
    """GPU operator -> GPU solve: the same solution as the oracle's solve on the oracle's operator (both FP64,
    "fortran" order), with the 5-point preconditioner of UPDATE_CG2D."""
    import oracle
    from mit_jax_amd import synthetic as S
    from mit_jax_amd import cg2d, solve_for_pressure as SFP
    g = grid0
    norm = S.ini_cg2d_norm(g)
    z = np.zeros((T, NY, NX))
    rBo = np.full((T, NY, NX), 1.0 / S.G)
    ops = SFP.update_cg2d(_fs(), _cp(norm), g, ex, to_dev(g.hFacW, dev), to_dev(g.hFacS, dev), to_dev(rBo, dev),
                          to_dev(z, dev), to_dev(z, dev), to_dev(z, dev))
    rhs = S.rhs(g, 3)
    print("producers -> solve", flush=True)
    x, diag = cg2d.cg2d_fortran(ex, ops, to_dev(rhs, dev), to_dev(z, dev), cg2dNorm=norm, tolerance=1e-12,
                                max_iters=2000, sum_order="fortran")
    rops = OS.update_cg2d(_fs(), norm, 0.51, g, maps().full("T"), g.hFacW, g.hFacS, rBo, z, z, z)
    rx, rd = oracle.cg2d(maps().full("T"), rops, rhs, z, cg2dNorm=norm, tolerance=1e-12, max_iters=2000)
    assert np.array_equal(x.cpu().numpy(), rx)
    assert int(diag["numIters"]) == rd["numIters"]


def test_cg2d_rhs_adjoint_bitwise(dev, grid0):
    """mj_cg2d_rhs_adjoint against the oracle's reverse mode (the same face cotangents and expressions), bitwise,
    both cotangent inputs given, useRealFreshWaterFlux on."""
    from mit_jax_amd import solve_for_pressure as SFP
    g = grid0
    rng = np.random.default_rng(13)
    gU, gV, EmPmR, etaN, etaH, Bo = _rhs_fields(rng, g)
    bb, xb = rng.normal(size=etaN.shape), rng.normal(size=etaN.shape)
    p = _fs()
    print("cg2d_rhs_adjoint: seed grid", flush=True)
    got = SFP.cg2d_rhs_adjoint(p, g, to_dev(gU, dev), to_dev(gV, dev), to_dev(g.hFacW, dev), to_dev(g.hFacS, dev),
                               to_dev(Bo, dev), to_dev(bb, dev), to_dev(xb, dev))
    ref = OS.cg2d_rhs_adjoint(p, g, gU, gV, g.hFacW, g.hFacS, Bo, bb, xb)
    for name, a, b in zip(("gU", "gV", "hFacW", "hFacS", "EmPmR", "etaN", "etaH"), got, ref):
        assert np.array_equal(a.cpu().numpy(), b), name


def test_update_cg2d_adjoint_matches_oracle(dev, ex, grid0):
    """mj_update_cg2d_adjoint against the oracle's reverse mode (transposed exchange, aC's neighbours, the level
    factors): equal to 1e-13 relative (the transposed exchange may sum a source's contributions in another order)."""
    from mit_jax_amd import synthetic as S
    from mit_jax_amd import solve_for_pressure as SFP
    g = grid0
    norm = S.ini_cg2d_norm(g)
    rng = np.random.default_rng(14)
    bars = [rng.normal(size=(T, NY, NX)) for _ in range(3)]
    print("update_cg2d_adjoint: seed grid", flush=True)
    got = SFP.update_cg2d_adjoint(_fs(), _cp(norm), g, ex, *(to_dev(b, dev) for b in bars), 50)
    ref = OS.update_cg2d_adjoint(_fs(), norm, g, maps().full("T"), *bars, 50)
    for name, a, b in zip(("hFacW", "hFacS"), got, ref):
        a = a.cpu().numpy()
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max(), name
