"""GPU: reverse mode of the DST3 step (gad.advect_tiles_adjoint / advect_tiles_ad), horizontal and with the R flux.

Pinned the way the reference pins jax.grad of advect_tiles (test_gad.py:297-346): gTracer is linear in the tracer,
so the tracer cotangent must pass the dot-product test <gbar, L dT> = <L^T gbar, dT> for random dT over every
padded point (to rounding); in the velocities it is smooth away from u = 0, so the velocity cotangents are checked
against central differences at wet points with |u| > 1e-3 m/s (best relative error of h/|u| in 1e-2..1e-4 below
1e-6, tracer below 1e-7). The R flux (not in the reference) is pinned the same way: the dot-product test of the
whole X/Y/R step in the tracer, and central differences in w at wet interfaces (and in the tracer, u and v through
the composed step; Richardson-extrapolated central differences, relative steps h/|x| in 5e-2..1e-4, errors
relative to max(|ad|, 1e-3 of the field's largest cotangent), as some of these cotangents are small).
"""

import numpy as np
import pytest

from tests.conftest import dst3_case, to_dev

pytestmark = pytest.mark.gpu


def _setup(dev):
    from mit_jax_amd import gad, synthetic as S
    g, u, v, w, tr = dst3_case(0)
    g2, vg = S.gad_inputs(g)
    g2d = {k: to_dev(a, dev) for k, a in g2.items()}
    params = gad.GADParams.llc90()
    return g, g2d, vg, params, u, v, tr


def test_tracer_adjoint_dot_product(dev):
    from mit_jax_amd import gad
    g, g2d, vg, params, u, v, tr = _setup(dev)
    rng = np.random.default_rng(11)
    dT = to_dev(rng.normal(size=tr.shape), dev)
    gbar = to_dev(rng.normal(size=tr.shape), dev)
    fwd = lambda t: gad.advect_tiles(params, None, g2d, vg, to_dev(u, dev), to_dev(v, dev), t,  # noqa: E731
                                     to_dev(g.hFacW, dev), to_dev(g.hFacS, dev), to_dev(g.recip_hFacC, dev))
    LdT = fwd(dT)
    tb, ub, vb = gad.advect_tiles_adjoint(params, None, g2d, vg, to_dev(u, dev), to_dev(v, dev), to_dev(tr, dev),
                                          to_dev(g.hFacW, dev), to_dev(g.hFacS, dev), to_dev(g.recip_hFacC, dev),
                                          gbar)
    lhs = float((gbar * LdT).sum())
    rhs = float((tb * dT).sum())
    print(f"dot test: {lhs:.15e} {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs))
    for a in (tb, ub, vb):
        assert bool(a.isfinite().all())


def test_gradients_vs_finite_differences(dev):
    import torch
    from mit_jax_amd import gad
    g, g2d, vg, params, u, v, tr = _setup(dev)
    L = params.layout
    rng = np.random.default_rng(0)
    inner = np.zeros(tr.shape, bool)
    inner[:, :, 4:94, 4:94] = True
    wi = to_dev(np.where(inner, rng.normal(size=tr.shape) * g.maskC, 0.0), dev)
    H = [to_dev(a, dev) for a in (g.hFacW, g.hFacS, g.recip_hFacC)]

    def J(t, uu, vv):
        return float((wi * gad.advect_tiles(params, None, g2d, vg, uu, vv, t, *H)).sum())

    tr_d, u_d, v_d = (to_dev(a, dev).requires_grad_(True) for a in (tr, u, v))
    out = gad.advect_tiles_ad(params, None, g2d, vg, u_d, v_d, tr_d, *H)
    gtr, gu, gv = torch.autograd.grad((wi * out).sum(), [tr_d, u_d, v_d])
    gtr, gu, gv = (a.cpu().numpy() for a in (gtr, gu, gv))
    assert np.isfinite(gtr).all() and np.isfinite(gu).all() and np.isfinite(gv).all()
    T0, U0, V0 = (to_dev(a, dev) for a in (tr, u, v))
    for t, k, j0, i0 in ((0, 5, 45, 30), (6, 3, 88, 2), (11, 10, 40, 60), (2, 0, 10, 85)):
        cand = [(t, k, L.jj(j), L.ii(i)) for j in range(j0, L.sNy + 1) for i in range(1, L.sNx + 1)
                if (j > j0 or i >= i0)]
        pt = next(q for q in cand if g.maskC[q] > 0 and g.maskW[q] > 0 and g.maskS[q] > 0 and abs(u[q]) > 1e-3
                  and abs(v[q]) > 1e-3)
        h = 1e-2
        e = torch.zeros_like(T0)
        e[pt] = 1.0
        fd = (J(T0 + h * e, U0, V0) - J(T0 - h * e, U0, V0)) / (2 * h)
        rel = abs(fd - gtr[pt]) / abs(gtr[pt])
        print(f"theta {pt}: ad {gtr[pt]:.10e} fd {fd:.10e} rel {rel:.2e}", flush=True)
        assert rel < 1e-7
        for name, base, gg in (("uFld", u, gu), ("vFld", v, gv)):
            best = np.inf
            for rh in (1e-2, 1e-3, 1e-4):
                hh = rh * abs(base[pt])
                if name == "uFld":
                    fd = (J(T0, U0 + hh * e, V0) - J(T0, U0 - hh * e, V0)) / (2 * hh)
                else:
                    fd = (J(T0, U0, V0 + hh * e) - J(T0, U0, V0 - hh * e)) / (2 * hh)
                r = abs(fd - gg[pt]) / abs(gg[pt])
                print(f"{name} {pt} h={hh:.1e}: ad {gg[pt]:.10e} fd {fd:.10e} rel {r:.2e}", flush=True)
                best = min(best, r)
            assert best < 1e-6, (name, pt, best)


def test_batched_members_and_levels(dev):
    """Two members in one reverse launch (the second with 30 active levels): each equals its own single-member
    reverse run on its active levels; inactive levels get zero cotangents."""
    import torch
    from mit_jax_amd import gad
    g, g2d, vg, params, u, v, tr = _setup(dev)
    rng = np.random.default_rng(12)
    gbar = rng.normal(size=(2,) + tr.shape)
    st = lambda a: to_dev(np.stack([a, a]), dev)  # noqa: E731
    tb, ub, vb = gad.advect_tiles_adjoint(params, None, g2d, vg, st(u), st(v), st(tr), st(g.hFacW), st(g.hFacS),
                                          st(g.recip_hFacC), to_dev(gbar, dev), nlev=[50, 30])
    for m in range(2):
        one = gad.advect_tiles_adjoint(params, None, g2d, vg, to_dev(u, dev), to_dev(v, dev), to_dev(tr, dev),
                                       to_dev(g.hFacW, dev), to_dev(g.hFacS, dev), to_dev(g.recip_hFacC, dev),
                                       to_dev(gbar[m], dev))
        K = 50 if m == 0 else 30
        for a, b in zip((tb, ub, vb), one):
            assert torch.equal(a[m][:, :K], b[:, :K])
            if K < 50:
                assert bool((a[m][:, K:] == 0).all())


def _setup_r(dev):
    """Config-1 inputs with the R flux: params with implicitAdvection False, w from continuity."""
    from mit_jax_amd import gad, synthetic as S
    g, u, v, w, tr = dst3_case(0)
    g2, vg = S.gad_inputs(g)
    g2d = {k: to_dev(a, dev) for k, a in g2.items()}
    params = gad.GADParams.llc90(implicitAdvection=False)
    H = [to_dev(a, dev) for a in (g.hFacW, g.hFacS, g.recip_hFacC)]
    return g, g2d, vg, params, u, v, w, tr, H


def test_r_flux_tracer_adjoint_dot_product(dev):
    from mit_jax_amd import gad
    g, g2d, vg, params, u, v, w, tr, H = _setup_r(dev)
    rng = np.random.default_rng(21)
    dT = to_dev(rng.normal(size=tr.shape), dev)
    gbar = to_dev(rng.normal(size=tr.shape), dev)
    U, V, W = (to_dev(a, dev) for a in (u, v, w))
    LdT = gad.advect_tiles(params, None, g2d, vg, U, V, dT, *H, wFld=W)
    bars = gad.advect_tiles_adjoint(params, None, g2d, vg, U, V, to_dev(tr, dev), *H, gbar, wFld=W)
    assert len(bars) == 4
    lhs = float((gbar * LdT).sum())
    rhs = float((bars[0] * dT).sum())
    print(f"R dot test: {lhs:.15e} {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs))
    for a in bars:
        assert bool(a.isfinite().all())
    assert not bool((bars[3][:, 0] != 0).any()), "w at the surface interface is not read: zero cotangent"


def test_r_flux_gradients_vs_finite_differences(dev):
    import torch
    from mit_jax_amd import gad
    g, g2d, vg, params, u, v, w, tr, H = _setup_r(dev)
    L = params.layout
    rng = np.random.default_rng(3)
    inner = np.zeros(tr.shape, bool)
    inner[:, :, 4:94, 4:94] = True
    wi = to_dev(np.where(inner, rng.normal(size=tr.shape) * g.maskC, 0.0), dev)
    T0, U0, V0, W0 = (to_dev(a, dev) for a in (tr, u, v, w))

    def J(t, uu, vv, ww):
        return float((wi * gad.advect_tiles(params, None, g2d, vg, uu, vv, t, *H, wFld=ww)).sum())

    args = [to_dev(a, dev).requires_grad_(True) for a in (tr, u, v, w)]
    out = gad.advect_tiles_ad(params, None, g2d, vg, args[1], args[2], args[0], *H, wFld=args[3])
    grads = torch.autograd.grad((wi * out).sum(), args)
    gtr, gu, gv, gw = (a.cpu().numpy() for a in grads)
    # the explicit reverse mode gives the autograd gradients
    bars = gad.advect_tiles_adjoint(params, None, g2d, vg, U0, V0, T0, *H, wi, wFld=W0)
    for a, b in zip(bars, (gtr, gu, gv, gw)):
        assert np.array_equal(a.cpu().numpy(), b)
    for t, k, j0, i0 in ((0, 5, 45, 30), (6, 12, 60, 2), (11, 20, 40, 60), (2, 1, 10, 85)):
        cand = [(t, k, L.jj(j), L.ii(i)) for j in range(j0, L.sNy + 1) for i in range(1, L.sNx + 1)
                if (j > j0 or i >= i0)]
        pt = next(q for q in cand if g.maskC[q] > 0 and g.maskC[(q[0], q[1] - 1) + q[2:]] > 0 and g.maskW[q] > 0
                  and g.maskS[q] > 0 and abs(u[q]) > 1e-3 and abs(v[q]) > 1e-3 and abs(w[q]) > 1e-7)
        e = torch.zeros_like(T0)
        e[pt] = 1.0
        h = 1e-2
        fd = (J(T0 + h * e, U0, V0, W0) - J(T0 - h * e, U0, V0, W0)) / (2 * h)
        rel = abs(fd - gtr[pt]) / abs(gtr[pt])
        print(f"R theta {pt}: ad {gtr[pt]:.10e} fd {fd:.10e} rel {rel:.2e}", flush=True)
        assert rel < 1e-7
        for name, base, gg in (("uFld", u, gu), ("vFld", v, gv), ("wFld", w, gw)):
            best = np.inf
            i = {"uFld": 1, "vFld": 2, "wFld": 3}[name]

            def D(hh):
                pp, pm = [T0, U0, V0, W0], [T0, U0, V0, W0]
                pp[i] = pp[i] + hh * e
                pm[i] = pm[i] - hh * e
                return (J(*pp) - J(*pm)) / (2 * hh)

            # errors relative to |ad|, or to 1e-3 of the field's largest cotangent where |ad| is smaller than that
            # (there the rounding of J, not the reverse mode, sets the finite-difference error)
            scale = max(abs(gg[pt]), 1e-3 * float(np.abs(gg).max()))
            for rh in (5e-2, 1e-2, 1e-3, 1e-4):
                hh = rh * abs(base[pt])
                fd = (4 * D(hh / 2) - D(hh)) / 3  # Richardson: the h^2 term of the central difference cancels
                r = abs(fd - gg[pt]) / scale
                print(f"R {name} {pt} h={hh:.1e}: ad {gg[pt]:.10e} fd {fd:.10e} rel {r:.2e}", flush=True)
                best = min(best, r)
            assert best < 1e-6, (name, pt, best)
