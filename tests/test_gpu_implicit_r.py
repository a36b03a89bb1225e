"""GPU parity of GAD_IMPLICIT_R (mj_gad_implicit_r) against the CPU oracle (oracle/implicit_r.py), bitwise: both
evaluate the reference's expressions (core/tracer_implicit.py:34-173) in order without fused multiply-adds; and,
at the ensemble size, conservation of each column's integral of hFac drF T."""

import ctypes
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import implicit_r as OI
from tests.conftest import to_dev

pytestmark = pytest.mark.gpu


def _inputs(nlev=None, seed=20):
    from mit_jax_amd import synthetic as S
    g = S.make_grid(0, nlev=nlev)
    rng = np.random.default_rng(seed)
    u, v, w = S.flow(g, 2)
    tr = S.tracer(g, 3)
    kap = (1e-5 + 1e-2 * rng.random(g.maskC.shape)) * g.maskC
    p = SimpleNamespace(dTtracerLev=np.asarray(g.v["dTlev"]), implicitDiffusion=True, rkSign=-1.0)
    return g, p, kap, w, tr


@pytest.mark.parametrize("adv", [True, False])
def test_gad_implicit_r_bitwise(dev, adv):
    from mit_jax_amd import tracer_implicit as TI
    g, p, kap, w, tr = _inputs()
    print(f"gad_implicit_r: seed grid, implicitAdvection={adv}", flush=True)
    got = TI.gad_implicit_r(p, g, adv, 3, to_dev(kap, dev), to_dev(g.recip_hFacC, dev), to_dev(w, dev),
                            to_dev(tr, dev))
    ref = OI.gad_implicit_r(p, g, adv, kap, g.recip_hFacC, w, tr)
    assert np.array_equal(got.cpu().numpy(), ref)


def test_gad_implicit_r_members_in_place_and_conservation(dev):
    """Three members with their own level counts (own maskC, hFac, w) in one launch through the C ABI, in place;
    each member bitwise equal to its oracle result, and every wet column's hFac drF T integral conserved."""
    import torch
    import mit_jax_amd._lib as L
    cases = [_inputs(n, 30 + n) for n in (15, 30, 50)]
    g0, p = cases[2][0], cases[2][1]
    st = lambda i: np.stack([c[i] for c in cases])  # noqa: E731
    mC = to_dev(np.stack([c[0].maskC for c in cases]), dev)
    rh = to_dev(np.stack([c[0].recip_hFacC for c in cases]), dev)
    kap, w, y = to_dev(st(2), dev), to_dev(st(3), dev), to_dev(st(4), dev)
    keep = [to_dev(np.ascontiguousarray(a, np.float64), dev) for a in
            (g0.rA, g0.recip_rA, p.dTtracerLev, g0.v["recip_drF"], g0.v["recip_drC"])]
    gs = L.ImprGrid(mC.data_ptr(), *(k.data_ptr() for k in keep), -1.0)
    K = 50
    n3 = 13 * K * 98 * 98
    cfg = L.ImprConfig(3, 13, 90, 90, 4, K, n3, n3, 1)
    y0 = y.clone()
    print("gad_implicit_r: 3 members, nlev 15/30/50, in place", flush=True)
    L.check(L.lib().mj_gad_implicit_r(ctypes.byref(gs), ctypes.byref(cfg), L.ptr(kap), L.ptr(rh), L.ptr(w),
                                      L.ptr(y), L.ptr(y), L.stream_ptr()), "mj_gad_implicit_r")
    torch.cuda.synchronize()
    out = y.cpu().numpy()
    for m, (g, pm, km, wm, trm) in enumerate(cases):
        ref = OI.gad_implicit_r(pm, g, True, km, g.recip_hFacC, wm, trm)
        assert np.array_equal(out[m], ref), m
        wt = g.hFacC * np.asarray(g.v["drF"])[None, :, None, None]
        before = (wt * trm)[:, :, 4:94, 4:94].sum(axis=1)
        after = (wt * out[m])[:, :, 4:94, 4:94].sum(axis=1)
        assert np.allclose(after, before, rtol=1e-12, atol=1e-9), m
    assert not torch.equal(y, y0)


@pytest.mark.parametrize("adv", [True, False])
def test_gad_implicit_r_adjoint_matches_oracle(dev, adv):
    """mj_gad_implicit_r_adjoint against the oracle's reverse mode (the same elimination on the transposed
    diagonals, the same cotangent expressions): equal to 1e-12 relative, on the seed grid."""
    from mit_jax_amd import tracer_implicit as TI
    g, p, kap, w, tr = _inputs()
    u = OI.gad_implicit_r(p, g, adv, kap, g.recip_hFacC, w, tr)
    ub = np.random.default_rng(21).normal(size=u.shape)
    print(f"gad_implicit_r_adjoint: seed grid, implicitAdvection={adv}", flush=True)
    got = TI.gad_implicit_r_adjoint(p, g, adv, 3, to_dev(kap, dev), to_dev(g.recip_hFacC, dev), to_dev(w, dev),
                                    to_dev(u, dev), to_dev(ub, dev))
    ref = OI.gad_implicit_r_adjoint(p, g, adv, kap, g.recip_hFacC, w, u, ub)
    for name, a, b in zip(("ybar", "kappaRX_bar", "wFld_bar"), got, ref):
        if a is None:
            assert not adv and name == "wFld_bar"
            continue
        a = a.cpu().numpy()
        scale = np.abs(b).max()
        assert np.abs(a - b).max() <= 1e-12 * scale, (name, np.abs(a - b).max(), scale)


def test_gad_implicit_r_autograd(dev):
    """torch.autograd through gad_implicit_r_ad: the gradient of a weighted sum in gTracer equals the oracle's
    transpose solve; in kappaRX and wFld it matches the oracle's cotangents."""
    import torch
    from mit_jax_amd import tracer_implicit as TI
    g, p, kap, w, tr = _inputs()
    wts = np.random.default_rng(22).normal(size=tr.shape)
    y_t, k_t, w_t = (to_dev(a, dev).requires_grad_(True) for a in (tr, kap, w))
    print("gad_implicit_r_ad backward", flush=True)
    u = TI.gad_implicit_r_ad(p, g, True, 3, k_t, to_dev(g.recip_hFacC, dev), w_t, y_t)
    (u * to_dev(wts, dev)).sum().backward()
    ref = OI.gad_implicit_r_adjoint(p, g, True, kap, g.recip_hFacC, w, u.detach().cpu().numpy(), wts)
    for t, b in zip((y_t, k_t, w_t), ref):
        a = t.grad.cpu().numpy()
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    assert torch.isfinite(k_t.grad).all()


@pytest.mark.parametrize("K", [2, 17, 30])
def test_gad_implicit_r_other_level_counts(dev, K):
    """Level counts other than 50 take the generic sweep (y' through the output column): bitwise the oracle too."""
    from mit_jax_amd import tracer_implicit as TI
    g, p, kap, w, tr = _inputs()
    sl = lambda a: np.ascontiguousarray(a[:, :K])  # noqa: E731
    gk = SimpleNamespace(layout=g.L, maskC=sl(g.maskC), rA=g.rA, recip_rA=g.recip_rA,
                         recip_drF=np.asarray(g.v["recip_drF"])[:K], recip_drC=np.asarray(g.v["recip_drC"])[:K])
    pk = SimpleNamespace(dTtracerLev=np.asarray(p.dTtracerLev)[:K], implicitDiffusion=True, rkSign=-1.0)
    print(f"gad_implicit_r: {K} levels", flush=True)
    got = TI.gad_implicit_r(pk, gk, True, 3, to_dev(sl(kap), dev), to_dev(sl(g.recip_hFacC), dev), to_dev(sl(w), dev),
                            to_dev(sl(tr), dev))
    ref = OI.gad_implicit_r(pk, gk, True, sl(kap), sl(g.recip_hFacC), sl(w), sl(tr))
    assert np.array_equal(got.cpu().numpy(), ref)


def test_gad_implicit_r_refuses_mismatched_shapes(dev):
    """Forward and reverse mode refuse arrays whose shape differs from the tracer's (the kernels index every array
    with the tracer's strides) before any launch."""
    from mit_jax_amd import tracer_implicit as TI
    g, p, kap, w, tr = _inputs()
    kd, rh, wd, y = to_dev(kap, dev), to_dev(g.recip_hFacC, dev), to_dev(w, dev), to_dev(tr, dev)
    short = kd[:, :30].contiguous()
    print("gad_implicit_r: mismatched shapes", flush=True)
    with pytest.raises(ValueError, match="kappaRX"):
        TI.gad_implicit_r(p, g, True, 3, short, rh, wd, y)
    with pytest.raises(ValueError, match="wFld"):
        TI.gad_implicit_r_adjoint(p, g, True, 3, kd, rh, wd[:, :30].contiguous(), y, y)
    with pytest.raises(ValueError, match="ubar"):
        TI.gad_implicit_r_adjoint(p, g, True, 3, kd, rh, wd, y, y[:, :30].contiguous())
    with pytest.raises(ValueError, match="tile layout"):
        TI.gad_implicit_r_adjoint(p, g, True, 3, kd, rh, wd, y[:12].contiguous(), y[:12].contiguous())
