"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on the same seeded inputs, bitwise.

The library and the oracle are both compiled without fused multiply-adds and evaluate the reference's operation
order, and the CG2D sums use the same order on both sides (the reference's sequential "fortran" order, or the
library's fixed "tile" tree), so every comparison here is exact equality.
"""

import numpy as np
import pytest

import oracle
from tests.conftest import dst3_case, maps, oracle_tables, to_dev

pytestmark = pytest.mark.gpu

KINDS_SCALAR = ("T", "3D", "Z", "SMs")
KINDS_VECTOR = ("UVs", "UVn", "UV3s", "As", "An", "Bs", "Bn", "Ds")


@pytest.fixture(scope="module")
def ex():
    from mit_jax_amd.exchange import default_exchanger
    return default_exchanger()


@pytest.mark.parametrize("kind", KINDS_SCALAR)
@pytest.mark.parametrize("shape", [(13, 98, 98), (13, 7, 98, 98)])
def test_halo_scalar_bitwise(dev, ex, kind, shape):
    rng = np.random.default_rng(hash((kind, len(shape))) % 2**32)
    a = rng.normal(size=shape)
    got = ex.scalar(to_dev(a, dev), kind).cpu().numpy()
    ref = oracle.scalar(maps().full(kind), a)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("kind", KINDS_VECTOR)
@pytest.mark.parametrize("shape", [(13, 98, 98), (13, 5, 98, 98)])
def test_halo_vector_bitwise(dev, ex, kind, shape):
    rng = np.random.default_rng(hash((kind, len(shape), 1)) % 2**32)
    u, v = rng.normal(size=shape), rng.normal(size=shape)
    gu, gv = ex.vector(to_dev(u, dev), to_dev(v, dev), kind)
    ru, rv = oracle.vector(maps().full(kind + "_u"), maps().full(kind + "_v"), u, v)
    assert np.array_equal(gu.cpu().numpy(), ru)
    assert np.array_equal(gv.cpu().numpy(), rv)


def test_halo_inplace_fill_matches_functional(dev, ex):
    rng = np.random.default_rng(5)
    a = to_dev(rng.normal(size=(13, 50, 98, 98)), dev)
    ref = ex.scalar(a, "T")
    b = a.clone()
    ex.fill_(b, "T")
    assert bool((b == ref).all())


def _cg_inputs(precond, seed=1):
    from mit_jax_amd import synthetic as S
    g, ops, norm = S.cg2d_case(0, precond)
    return g, ops, norm, S.rhs(g, seed)


@pytest.mark.parametrize("precond", ["diagonal", "5point"])
@pytest.mark.parametrize("order", ["fortran", "tile"])
def test_cg2d_bitwise(dev, ex, precond, order):
    from mit_jax_amd import cg2d
    g, ops, norm, b = _cg_inputs(precond)
    rng = np.random.default_rng(9)
    x0 = rng.normal(size=b.shape) * 1e-3  # a nonzero first guess: exercises the halo normalisation
    print(f"cg2d {precond} {order}", flush=True)
    x, d = cg2d.cg2d_fortran(ex, [to_dev(o, dev) for o in ops], to_dev(b, dev), to_dev(x0, dev), cg2dNorm=norm,
                             tolerance=1e-12, max_iters=2000, sum_order=order)
    d.check()
    xr, dr = oracle.cg2d(maps().full("T"), ops, b, x0, cg2dNorm=norm, tolerance=1e-12, max_iters=2000,
                         sum_order=order)
    assert int(d["numIters"]) == dr["numIters"]
    assert int(d["converged"]) == 1
    for k in ("firstResidual", "lastResidual", "sumRHS", "rhsMax"):
        assert float(d[k]) == dr[k], (k, float(d[k]), dr[k])
    assert np.array_equal(x.cpu().numpy(), xr)


LAUNCHES = ["persistent", "phases", "phases_fixed"]


@pytest.mark.parametrize("launch", LAUNCHES)
def test_cg2d_batched_members_bitwise(dev, ex, launch):
    """Members with different rhs (one zero rhs: 0 iterations) and different max_iters outcomes in one launch;
    each member equals its own single solve on the oracle, in every launch shape."""
    import torch
    from mit_jax_amd import cg2d
    g, ops, norm, _ = _cg_inputs("5point")
    from mit_jax_amd import synthetic as S
    bs = [S.rhs(g, s) for s in (11, 12)] + [np.zeros(g.L.shape2d), S.rhs(g, 13) * 1e-3]
    B = len(bs)
    b = np.stack(bs)
    x0 = np.zeros_like(b)
    print(f"cg2d batched {launch}", flush=True)
    x, d = cg2d.cg2d_fortran(ex, [to_dev(o, dev) for o in ops], to_dev(b, dev), to_dev(x0, dev), cg2dNorm=norm,
                             tolerance=1e-11, max_iters=2000, sum_order="tile", launch=launch, poll_iters=5)
    d.check()
    xs = x.cpu().numpy()
    for m in range(B):
        xr, dr = oracle.cg2d(maps().full("T"), ops, bs[m], x0[m], cg2dNorm=norm, tolerance=1e-11, max_iters=2000,
                             sum_order="tile")
        assert int(d["numIters"][m]) == dr["numIters"], m
        assert np.array_equal(xs[m], xr), m
    assert int(d["numIters"][2]) == 0
    # max_iters reached: not converged, stops at the cap
    x, d = cg2d.cg2d_fortran(ex, [to_dev(o, dev) for o in ops], to_dev(b, dev), to_dev(x0, dev), cg2dNorm=norm,
                             tolerance=1e-11, max_iters=7, sum_order="tile", launch=launch, poll_iters=3)
    d.check()
    assert d["numIters"].cpu().tolist() == [7, 7, 0, 7]
    assert d["converged"].cpu().tolist() == [0, 0, 1, 0]
    xr, dr = oracle.cg2d(maps().full("T"), ops, bs[1], x0[1], cg2dNorm=norm, tolerance=1e-11, max_iters=7,
                         sum_order="tile")
    assert np.array_equal(x[1].cpu().numpy(), xr)
    torch.cuda.synchronize()


@pytest.mark.parametrize("launch", ["persistent", "phases"])
def test_cg2d_operator_slots_bitwise(dev, ex, launch):
    """Members reading different operator slots (5-point and diagonal preconditioners) in one launch through
    op_index: each equals its own single solve on the oracle, in both sum orders."""
    from mit_jax_amd import cg2d, synthetic as S
    g, ops5, norm = S.cg2d_case(0, "5point")
    _, opsd, _ = S.cg2d_case(0, "diagonal")
    slots = [to_dev(np.stack([a5, ad]), dev) for a5, ad in zip(ops5, opsd)]
    op_index = [1, 0, 0, 1, 1]
    bs = [S.rhs(g, 20 + m) for m in range(len(op_index))]
    b = np.stack(bs)
    for order in ("tile", "fortran"):
        print(f"cg2d slots {launch} {order}", flush=True)
        x, d = cg2d.cg2d_fortran(ex, slots, to_dev(b, dev), to_dev(np.zeros_like(b), dev), cg2dNorm=norm,
                                 tolerance=1e-12, max_iters=2000, sum_order=order, op_index=op_index, launch=launch)
        d.check()
        xs = x.cpu().numpy()
        for m, k in enumerate(op_index):
            xr, dr = oracle.cg2d(maps().full("T"), (ops5, opsd)[k], bs[m], np.zeros_like(bs[m]), cg2dNorm=norm,
                                 tolerance=1e-12, max_iters=2000, sum_order=order)
            assert int(d["numIters"][m]) == dr["numIters"], (order, m)
            assert np.array_equal(xs[m], xr), (order, m)


def test_cg2d_block_size_independent(dev, ex):
    """The reduction trees do not depend on the launch geometry: any launch shape and block size gives the same
    bits."""
    from mit_jax_amd import cg2d
    g, ops, norm, b = _cg_inputs("5point")
    x0 = np.zeros_like(b)
    outs = []
    for launch, bt in (("persistent", 128), ("persistent", 512), ("persistent", 1024), ("phases", 64),
                       ("phases", 256), ("phases", 512), ("phases", 1024)):
        print(f"cg2d geometry {launch} {bt}", flush=True)
        x, d = cg2d.cg2d_fortran(ex, [to_dev(o, dev) for o in ops], to_dev(b, dev), to_dev(x0, dev), cg2dNorm=norm,
                                 tolerance=1e-12, max_iters=2000, sum_order="tile", block_threads=bt, launch=launch)
        d.check()
        outs.append(x.cpu().numpy())
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


def _gad_dev(dev, g, u, v, w, tr):
    from mit_jax_amd import synthetic as S
    g2, vg = S.gad_inputs(g)
    return {k: to_dev(a, dev) for k, a in g2.items()}, vg


def test_dst3_bitwise(dev):
    from mit_jax_amd import gad, synthetic as S
    g, u, v, w, tr = dst3_case(0)
    g2d, vg = _gad_dev(dev, g, u, v, w, tr)
    params = gad.GADParams.llc90()
    gT = gad.advect_tiles(params, None, g2d, vg, to_dev(u, dev), to_dev(v, dev), to_dev(tr, dev),
                          to_dev(g.hFacW, dev), to_dev(g.hFacS, dev), to_dev(g.recip_hFacC, dev))
    g2, _ = S.gad_inputs(g)
    ref = oracle.advect_tiles(oracle_tables(), g2, vg, u, v, tr, g.hFacW, g.hFacS, g.recip_hFacC)
    assert np.array_equal(gT.cpu().numpy(), ref)


def test_dst3_with_vertical_bitwise(dev):
    from mit_jax_amd import gad, synthetic as S
    g, u, v, w, tr = dst3_case(0)
    g2d, vg = _gad_dev(dev, g, u, v, w, tr)
    params = gad.GADParams.llc90(implicitAdvection=False)
    gT = gad.advect_tiles(params, None, g2d, vg, to_dev(u, dev), to_dev(v, dev), to_dev(tr, dev),
                          to_dev(g.hFacW, dev), to_dev(g.hFacS, dev), to_dev(g.recip_hFacC, dev), wFld=to_dev(w, dev))
    g2, _ = S.gad_inputs(g)
    ref = oracle.gad_dst3(oracle_tables(), g2, vg, u, v, w, tr, g.hFacW, g.hFacS, g.recip_hFacC)
    assert np.array_equal(gT.cpu().numpy(), ref)


def test_dst3_batched_members(dev):
    """Two members (perturbed tracer) in one launch, the second with 30 active levels: each equals its own
    oracle run on the active levels; levels beyond a member's count are left untouched."""
    import torch
    from mit_jax_amd import gad, synthetic as S
    g, u, v, w, tr = dst3_case(0)
    g2d, vg = _gad_dev(dev, g, u, v, w, tr)
    tr2 = tr + np.where(g.maskC > 0, np.random.default_rng(4).normal(0, 0.05, tr.shape), 0.0)
    trb = np.stack([tr, tr2])
    B = 2
    params = gad.GADParams.llc90()
    st = lambda a: to_dev(np.stack([a] * B), dev)  # noqa: E731
    out = torch.full((B,) + tr.shape, 7.0, dtype=torch.float64, device=dev)
    gT = gad.advect_tiles(params, None, g2d, vg, st(u), st(v), to_dev(trb, dev), st(g.hFacW), st(g.hFacS),
                          st(g.recip_hFacC), nlev=[50, 30], out=out).cpu().numpy()
    g2, _ = S.gad_inputs(g)
    ref0 = oracle.advect_tiles(oracle_tables(), g2, vg, u, v, tr, g.hFacW, g.hFacS, g.recip_hFacC)
    ref1 = oracle.advect_tiles(oracle_tables(), g2, vg, u, v, tr2, g.hFacW, g.hFacS, g.recip_hFacC)
    assert np.array_equal(gT[0], ref0)
    assert np.array_equal(gT[1][:, :30], ref1[:, :30])
    assert np.all(gT[1][:, 30:] == 7.0)


def test_dst3_levels_per_block_same_bits(dev):
    """The DST3 tile shape (levels per workgroup) changes the launch only: two members with 50 and 30 active
    levels, with and without the R flux, give the same bits for every shape."""
    import torch
    from mit_jax_amd import gad
    g, u, v, w, tr = dst3_case(0)
    g2d, vg = _gad_dev(dev, g, u, v, w, tr)
    tr2 = tr + np.where(g.maskC > 0, np.random.default_rng(5).normal(0, 0.05, tr.shape), 0.0)
    st = lambda a: to_dev(np.stack([a] * 2), dev)  # noqa: E731
    args = (st(u), st(v), to_dev(np.stack([tr, tr2]), dev), st(g.hFacW), st(g.hFacS), st(g.recip_hFacC))
    for implicit in (True, False):
        params = gad.GADParams.llc90(implicitAdvection=implicit)
        wf = None if implicit else st(w)
        ref = gad.advect_tiles(params, None, g2d, vg, *args, wFld=wf, nlev=[50, 30])
        for lpb in (2, 3, 7, 50):
            print(f"levels per block {lpb}, implicitAdvection {implicit}", flush=True)
            got = gad.advect_tiles(params, None, g2d, vg, *args, wFld=wf, nlev=[50, 30], levels_per_block=lpb)
            assert torch.equal(got, ref), (lpb, implicit)


def test_dst3_mask_slots_same_bits(dev):
    """Mask slots picked per member by grid2d["mask_index"] give the same bits as the same masks stored once per
    member, with and without the R flux; all members on slot 0 give the shared-mask result."""
    import torch
    from mit_jax_amd import gad
    g, u, v, w, tr = dst3_case(0)
    g2d, vg = _gad_dev(dev, g, u, v, w, tr)
    keep = (torch.arange(g2d["maskW"].shape[1], device=dev) < 20).to(torch.float64).view(1, -1, 1, 1)
    slots = {k: torch.stack([g2d[k], g2d[k] * keep]) for k in ("maskW", "maskS", "maskC")}
    idx = [1, 0, 1]
    B = len(idx)
    rng = np.random.default_rng(11)
    trs = np.stack([tr + np.where(g.maskC > 0, rng.normal(0, 0.05, tr.shape), 0.0) for _ in range(B)])
    st = lambda a: to_dev(np.stack([a] * B), dev)  # noqa: E731
    args = (st(u), st(v), to_dev(trs, dev), st(g.hFacW), st(g.hFacS), st(g.recip_hFacC))
    per_member = dict(g2d, **{k: s[idx].contiguous() for k, s in slots.items()})
    by_slot = dict(g2d, **slots, mask_index=idx)
    all_zero = dict(g2d, **slots, mask_index=[0] * B)
    for implicit in (True, False):
        print(f"mask slots, implicitAdvection {implicit}", flush=True)
        params = gad.GADParams.llc90(implicitAdvection=implicit)
        wf = None if implicit else st(w)
        ref = gad.advect_tiles(params, None, per_member, vg, *args, wFld=wf)
        got = gad.advect_tiles(params, None, by_slot, vg, *args, wFld=wf)
        assert torch.equal(got, ref), implicit
        shared = gad.advect_tiles(params, None, g2d, vg, *args, wFld=wf)
        assert torch.equal(gad.advect_tiles(params, None, all_zero, vg, *args, wFld=wf), shared), implicit
        assert not torch.equal(got, shared)


def test_runs_are_deterministic(dev, ex):
    from mit_jax_amd import cg2d, gad
    g, ops, norm, b = _cg_inputs("5point")
    opsd = [to_dev(o, dev) for o in ops]
    r = [cg2d.cg2d_fortran(ex, opsd, to_dev(b, dev), to_dev(np.zeros_like(b), dev), cg2dNorm=norm, tolerance=1e-12,
                           max_iters=2000, sum_order="tile")[0].cpu().numpy() for _ in range(2)]
    assert np.array_equal(r[0], r[1])
    g, u, v, w, tr = dst3_case(0)
    g2d, vg = _gad_dev(dev, g, u, v, w, tr)
    params = gad.GADParams.llc90()
    r = [gad.advect_tiles(params, None, g2d, vg, to_dev(u, dev), to_dev(v, dev), to_dev(tr, dev),
                          to_dev(g.hFacW, dev), to_dev(g.hFacS, dev), to_dev(g.recip_hFacC, dev)).cpu().numpy()
         for _ in range(2)]
    assert np.array_equal(r[0], r[1])
