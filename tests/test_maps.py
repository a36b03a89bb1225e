"""exch2 halo maps (CPU): the committed compact tables against the reference's map file, the map invariants of the
reference's tests (test_exchange.py:60-73, :106-118), and the properties the kernels rely on."""

import hashlib
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.conftest import maps

REF_NPZ = Path("/root/reference/mitgcm_jax/data/exch_maps_13x90x90.npz")
T, NY, NX, OL, N = 13, 98, 98, 4, 90
PLANE = NY * NX


def _interior():
    m = np.zeros((T, NY, NX), bool)
    m[:, OL:OL + N, OL:OL + N] = True
    return m.reshape(-1)


@pytest.mark.skipif(not REF_NPZ.exists(), reason="reference map file not mounted")
def test_compact_tables_reproduce_reference_file():
    """Every map rebuilt from the compact written-point tables equals the reference's full map, element for
    element (src, comp, sign)."""
    assert hashlib.sha256(REF_NPZ.read_bytes()).hexdigest() == \
        "a31f0d49754259c3b1304c42f6734886605c0f4f1f72257b0654e8486a9d84fd"
    z = np.load(REF_NPZ, allow_pickle=False)
    names = sorted({k.rsplit(".", 1)[0] for k in z.files if k != "layout"})
    assert names == maps().names()
    assert list(z["layout"]) == [90, 90, 4, 4, 13, 50]
    for n in names:
        src, comp, sign = maps().full(n)
        assert np.array_equal(src, z[n + ".src"]), n
        assert np.array_equal(comp, z[n + ".comp"]), n
        assert np.array_equal(sign, z[n + ".sign"]), n


def test_map_invariants():
    """Sources inside the padded arrays, interior points never written, scalar maps without component swaps or
    sign flips, and exactly 4 x 90 x OL = 1,440 unwritten C-halo points (the open facet edges)."""
    interior = _interior()
    for n in maps().names():
        src, comp, sign = maps().full(n)
        assert src.min() >= 0 and src.max() < T * PLANE, n
        assert np.all(comp[interior] == 0), n
        if n in ("T", "Z", "3D"):
            assert np.all(comp <= 1) and np.all(sign == 1), n
    _, comp, _ = maps().full("T")
    assert ((comp == 0) & ~interior).sum() == 4 * 90 * 4
    assert (comp > 0).sum() == 18112


def test_inplace_safe_kinds():
    """The kinds whose written points are never sources (so a fill in place equals the functional gather) are
    exactly T, 3D, UVs, UVn, UV3s; the library reports the same."""
    safe = set()
    for kind in ("T", "3D", "Z", "SMs"):
        d, s, c, g = maps().compact[kind]
        if not np.isin(s, d).any():
            safe.add(kind)
    for kind in ("UVs", "UVn", "UV3s", "As", "An", "Bs", "Bn", "Ds"):
        du, su, cu, _ = maps().compact[kind + "_u"]
        dv, sv, cv, _ = maps().compact[kind + "_v"]
        hit = False
        for s, c in ((su, cu), (sv, cv)):
            hit |= np.isin(s[c == 1], du).any() or np.isin(s[c == 2], dv).any()
        if not hit:
            safe.add(kind)
    assert safe == {"T", "3D", "UVs", "UVn", "UV3s"}


def test_stencil_ring_sourced_from_interior():
    """The width-1 ring the 5-point stencils read (i = 0, sNx+1 for j = 1..sNy; j = 0, sNy+1 for i = 1..sNx) has
    4,320 written points, all sourced from interior points (the fused CG2D kernel computes them from there);
    360 ring points stay unwritten (open facet edges)."""
    src, comp, _ = maps().full("T")
    ring = np.zeros((NY, NX), bool)
    ring[OL - 1, OL:OL + N] = ring[OL + N, OL:OL + N] = True
    ring[OL:OL + N, OL - 1] = ring[OL:OL + N, OL + N] = True
    ring = np.tile(ring.reshape(-1), T)
    w = ring & (comp > 0)
    assert w.sum() == 4320 and (ring & (comp == 0)).sum() == 360
    assert _interior()[src[w]].all()


def _numpy_apply(m, own, fa, fb):
    """The reference's gather formula (parallel/exchange.py:136-142) in numpy, over the flat point axis."""
    src, comp, sign = m
    out = np.where(comp == 1, fa[..., src], own)
    out = np.where(comp == 2, fb[..., src], out)
    return out * sign


@pytest.mark.parametrize("kind", ["T", "Z", "SMs", "UVs", "As", "Ds"])
def test_oracle_exchange_matches_gather_formula(kind):
    rng = np.random.default_rng(7)
    K = 3
    u = rng.normal(size=(T, K, NY, NX))
    v = rng.normal(size=(T, K, NY, NX))
    flat = lambda a: np.moveaxis(a, 0, 1).reshape(K, T * PLANE)  # noqa: E731
    unflat = lambda f: np.moveaxis(f.reshape(K, T, NY, NX), 0, 1)  # noqa: E731
    if kind in ("T", "Z", "SMs"):
        m = maps().full(kind)
        ref = unflat(_numpy_apply(m, flat(u), flat(u), flat(u)))
        assert np.array_equal(oracle.scalar(m, u), ref)
    else:
        mu, mv = maps().full(kind + "_u"), maps().full(kind + "_v")
        gu, gv = oracle.vector(mu, mv, u, v)
        assert np.array_equal(gu, unflat(_numpy_apply(mu, flat(u), flat(u), flat(v))))
        assert np.array_equal(gv, unflat(_numpy_apply(mv, flat(v), flat(u), flat(v))))


def test_exchange_adjoint_identity():
    """<E x, y> = <x, E^T y>, E^T the scatter-add transpose of the gather (reference test_exchange.py:106-118)."""
    rng = np.random.default_rng(0)
    for kind in ("T", "UVs"):
        if kind == "T":
            src, comp, sign = maps().full("T")
            x, y = rng.normal(size=T * PLANE), rng.normal(size=T * PLANE)
            ex = np.where(comp == 1, x[src], x) * sign
            et = np.where(comp == 0, y, 0.0)
            np.add.at(et, src[comp == 1], (y * sign)[comp == 1])
            assert abs(ex @ y - x @ et) <= 1e-13 * abs(ex @ y)
        else:
            mu, mv = maps().full("UVs_u"), maps().full("UVs_v")
            xu, xv, yu, yv = (rng.normal(size=T * PLANE) for _ in range(4))
            eu = _numpy_apply(mu, xu, xu, xv)
            ev = _numpy_apply(mv, xv, xu, xv)
            bu, bv = np.zeros_like(xu), np.zeros_like(xv)
            for (src, comp, sign), y, own in ((mu, yu, bu), (mv, yv, bv)):
                own += np.where(comp == 0, y, 0.0)
                np.add.at(bu, src[comp == 1], (y * sign)[comp == 1])
                np.add.at(bv, src[comp == 2], (y * sign)[comp == 2])
            lhs = eu @ yu + ev @ yv
            assert abs(lhs - (xu @ bu + xv @ bv)) <= 1e-13 * abs(lhs)
