"""GPU: the reference entry points not covered by the solve / advection parity tests, each against the oracle.

* dst3_adv_x / dst3_adv_y (reference pkgs/gad.py:409-453): bitwise the oracle's fluxes on the config-1 inputs,
  batched (two members with different tracers) and with the reference's argument shapes ([1, K, 1, 1] per-level
  factors, [T, 1, ny, nx] metrics).
* _matvec (core/cg2d.py:292-298): bitwise the oracle's, one member and batched over operator slots (op_index).
* gad_advection (pkgs/gad.py:546-560): the public entry, reference branch (implicitAdvection, wFld ignored) bitwise
  the oracle's advect_tiles; the vertical-flux branch bitwise the oracle's horizontal + vertical sweep.
* ini_cg2d_norm (core/cg2d.py:108-134): bitwise the oracle's, on the three level classes in one batched call and
  on dyadic data with a factor != 1.
"""

import numpy as np
import pytest

import oracle
from oracle import sfp as OS
from tests.conftest import dst3_case, maps, oracle_tables, to_dev

pytestmark = pytest.mark.gpu


def _transports(g, u, v):
    """uTrans, vTrans as advect_tiles forms them (pkgs/gad.py:482-485 with deep/rho factors 1)."""
    drF = g.drF[None, :, None, None]
    xA = g.dyG[:, None] * 1.0 * drF * g.hFacW
    yA = g.dxG[:, None] * 1.0 * drF * g.hFacS
    return u * xA * 1.0, v * yA * 1.0


@pytest.mark.parametrize("direction", ["x", "y"])
def test_dst3_adv_bitwise(dev, direction):
    import torch
    from mit_jax_amd import gad
    from mit_jax_amd.layout import LLC90
    g, u, v, _, tr = dst3_case(0)
    uT, vT = _transports(g, u, v)
    trans, vel, mask, rd = (uT, u, g.maskW, g.recip_dxC) if direction == "x" else (vT, v, g.maskS, g.recip_dyC)
    dt = np.asarray(g.dTlev)
    K = tr.shape[1]
    tr2 = tr + 0.5 * np.random.default_rng(7).normal(size=tr.shape) * g.maskC
    fn = gad.dst3_adv_x if direction == "x" else gad.dst3_adv_y
    print(f"dst3_adv_{direction}: one member, reference shapes", flush=True)
    got = fn(LLC90, 1.0 / 6.0, dt.reshape(1, K, 1, 1), to_dev(trans, dev), to_dev(vel, dev), to_dev(mask, dev),
             to_dev(tr, dev), to_dev(rd[:, None], dev), np.ones((1, K, 1, 1)))
    ref = oracle.dst3_adv(direction, trans, vel, mask, tr, rd, dt, np.ones(K))
    assert np.array_equal(got.cpu().numpy(), ref)
    assert np.abs(ref).max() > 0
    print(f"dst3_adv_{direction}: two members", flush=True)
    st = lambda a, b=None: to_dev(np.stack([a, a if b is None else b]), dev)  # noqa: E731
    got2 = fn(LLC90, 1.0 / 6.0, dt, st(trans), st(vel), st(mask), st(tr, tr2), to_dev(rd, dev), np.ones(K))
    ref2 = oracle.dst3_adv(direction, trans, vel, mask, tr2, rd, dt, np.ones(K))
    torch.cuda.synchronize()
    assert np.array_equal(got2[0].cpu().numpy(), ref) and np.array_equal(got2[1].cpu().numpy(), ref2)


def test_matvec_bitwise(dev):
    from mit_jax_amd import cg2d, synthetic as S
    from mit_jax_amd.exchange import default_exchanger
    ex = default_exchanger()
    g, ops5, norm = S.cg2d_case(0, "5point")
    g15, ops15, norm15 = S.cg2d_case(0, "5point", nlev=15)
    rng = np.random.default_rng(11)
    vs = [rng.normal(size=g.L.shape2d) for _ in range(3)]  # halos random too: _matvec zeroes them first
    print("matvec: one member", flush=True)
    y = cg2d.matvec(ex, [to_dev(a, dev) for a in ops5[:3]], to_dev(vs[0], dev), norm).cpu().numpy()
    assert np.array_equal(y, oracle.cg2d_matvec(maps().full("T"), ops5[:3], vs[0], norm))
    print("matvec: batched over operator slots", flush=True)
    slots = [to_dev(np.stack([a, b]), dev) for a, b in zip(ops5[:3], ops15[:3])]
    op_index = [1, 0, 1]
    norms = [norm15, norm, norm15]
    yb = cg2d.matvec(ex, slots, to_dev(np.stack(vs), dev), np.asarray(norms), op_index=op_index).cpu().numpy()
    for m, k in enumerate(op_index):
        ops = (ops5, ops15)[k]
        assert np.array_equal(yb[m], oracle.cg2d_matvec(maps().full("T"), ops[:3], vs[m], norms[m])), m


@pytest.mark.parametrize("implicit", [True, False])
def test_gad_advection_entry(dev, implicit):
    from mit_jax_amd import gad, synthetic as S
    g, u, v, w, tr = dst3_case(0)
    params = gad.GADParams.llc90(implicitAdvection=implicit)
    g2, vg = S.gad_inputs(g)
    gd = {k: to_dev(np.asarray(a), dev) for k, a in {**g2, **vg}.items()}
    print(f"gad_advection implicitAdvection={implicit}", flush=True)
    got = gad.gad_advection(params, gd, to_dev(u, dev), to_dev(v, dev), to_dev(w, dev), to_dev(tr, dev),
                            to_dev(g.hFacW, dev), to_dev(g.hFacS, dev), to_dev(g.recip_hFacC, dev)).cpu().numpy()
    tab = oracle_tables()
    if implicit:
        ref = oracle.advect_tiles(tab, g2, vg, u, v, tr, g.hFacW, g.hFacS, g.recip_hFacC)
    else:
        ref = oracle.gad_dst3(tab, g2, vg, u, v, w, tr, g.hFacW, g.hFacS, g.recip_hFacC)
    assert np.array_equal(got, ref)
    if not implicit:  # the vertical flux changes the tendency
        assert not np.array_equal(ref, oracle.advect_tiles(tab, g2, vg, u, v, tr, g.hFacW, g.hFacS, g.recip_hFacC))


def test_ini_cg2d_norm_bitwise(dev):
    import torch
    from mit_jax_amd import cg2d, synthetic as S
    gs = [S.cg2d_case(0, "5point", nlev=n)[0] for n in (15, 30, 50)]
    hW = to_dev(np.stack([g.hFacW for g in gs]), dev)
    hS = to_dev(np.stack([g.hFacS for g in gs]), dev)
    print("ini_cg2d_norm: three level classes, batched", flush=True)
    got = cg2d.ini_cg2d_norm(gs[2], hW, hS, 1.0, 1.0).cpu().numpy()
    for m, g in enumerate(gs):
        assert got[m] == OS.ini_cg2d_norm(g, g.hFacW, g.hFacS, 1.0, 1.0), m
        assert got[m] == S.cg2d_case(0, "5point", nlev=(15, 30, 50)[m])[2]
    print("ini_cg2d_norm: dyadic data, factor 2", flush=True)
    rng = np.random.default_rng(3)
    g = gs[2]
    gd = S.Grid(g.L, dict(g.f, dyG=rng.integers(1, 8, g.L.shape2d).astype(np.float64),
                          recip_dxC=2.0 ** -rng.integers(0, 3, g.L.shape2d)), dict(g.v))
    h = rng.integers(0, 5, g.hFacW.shape) / 4.0
    one = float(cg2d.ini_cg2d_norm(gd, to_dev(h, dev), to_dev(h * 0.5, dev), 0.5, 4.0))
    torch.cuda.synchronize()
    assert one == OS.ini_cg2d_norm(gd, h, h * 0.5, 0.5, 4.0)
