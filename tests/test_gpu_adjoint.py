"""GPU: reverse mode of the CG2D solve (implicit derivative, mit_jax_amd.adjoint) — the checks of the reference's
test_cg2d.py:214-277 on the synthetic operator, plus the coefficient-cotangent kernel against the oracle bitwise."""

import dataclasses

import numpy as np
import pytest

import oracle
from tests.conftest import maps, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ex():
    from mit_jax_amd.exchange import default_exchanger
    return default_exchanger()


@pytest.fixture(scope="module")
def case():
    from mit_jax_amd import synthetic as S
    from mit_jax_amd.cg2d import Cg2dParams
    g, ops, norm = S.cg2d_case(0, "5point")
    p = Cg2dParams(cg2dMaxIters=1000, cg2dTolerance=1e-12, cg2dNorm=norm, sum_order="tile")
    return g, ops, norm, p


def _inner(g):
    m = np.zeros(g.L.shape2d, bool)
    m[:, 4:94, 4:94] = True
    return m


def test_coeff_grad_kernel_bitwise(dev, ex, case):
    from mit_jax_amd import adjoint
    g, ops, norm, _ = case
    rng = np.random.default_rng(3)
    inner = _inner(g)
    x = rng.normal(size=inner.shape)
    lam = np.where(inner, rng.normal(size=inner.shape), 0.0)
    gk = adjoint.coeff_grad(ex, to_dev(x, dev), to_dev(lam, dev), norm)
    go = oracle.cg2d_coeff_grad(maps().full("T"), x, lam, norm)
    for a, b in zip(gk, go):
        assert np.array_equal(a.cpu().numpy(), b)


def test_transpose_solve(dev, ex, case):
    """z = dJ/db for J = <w, x(b)>: A z = w (relative residual < 1e-11), z = 0 off the interior, and
    <z, b> = <w, x> with x from a forward solve at 1e-13 (rtol 1e-9) (reference test_cg2d.py:214-237)."""
    import torch
    from mit_jax_amd import adjoint, synthetic as S
    g, ops, norm, p = case
    inner = _inner(g)
    rng = np.random.default_rng(1)
    b = to_dev(S.rhs(g, 5), dev).requires_grad_(True)
    x0 = torch.zeros_like(b)
    w = to_dev(rng.normal(size=inner.shape) * inner, dev)
    opsd = [to_dev(o, dev) for o in ops]
    x, d = adjoint.cg2d_solve(p, ex, opsd, b, x0)
    J = (w * x).sum()
    (z,) = torch.autograd.grad(J, [b])
    zn = z.cpu().numpy()
    assert np.all(np.isfinite(zn)) and np.all(zn[~inner] == 0.0)
    Az = oracle.cg2d_matvec(maps().full("T"), ops[:3], zn, norm)
    wn = w.cpu().numpy()
    res = np.abs(Az - wn).max() / np.abs(wn).max()
    assert res < 1e-11, res
    pt = dataclasses.replace(p, cg2dTolerance=1e-13)
    xt, _ = adjoint.cg2d_solve(pt, ex, opsd, b.detach(), x0)
    np.testing.assert_allclose(float((z * b.detach()).sum()), float((w * xt).sum()), rtol=1e-9)


def test_coefficient_and_rhs_gradients_fd(dev, ex, case):
    """d/d(aW2d), d/d(aC2d), d/d(cg2d_b) of J = <w, x> by reverse mode vs central differences along a random
    direction on a wet 3x3 window (reference test_cg2d.py:240-277: tolerances 1e-6, 1e-5, 1e-8 on the best
    relative error of the h sweep; FD forwards at cg2dTolerance 1e-13)."""
    import torch
    from mit_jax_amd import adjoint, synthetic as S
    g, ops, norm, p = case
    inner = _inner(g)
    rng = np.random.default_rng(2)
    w = to_dev(rng.normal(size=inner.shape) * inner, dev)
    opsd = [to_dev(o, dev) for o in ops]
    b0 = to_dev(S.rhs(g, 6), dev)
    x0 = torch.zeros_like(b0)
    tight = dataclasses.replace(p, cg2dTolerance=1e-13)

    def J(pp, aW, aC, b):
        o = [aW, opsd[1], aC] + opsd[3:]
        return (w * adjoint.cg2d_solve(pp, ex, o, b, x0)[0]).sum()

    leaves = [opsd[0].clone().requires_grad_(True), opsd[2].clone().requires_grad_(True),
              b0.clone().requires_grad_(True)]
    grads = torch.autograd.grad(J(p, *leaves), leaves)
    for a in grads:
        assert bool(torch.isfinite(a).all())
    base = [opsd[0], opsd[2], b0]
    t, j0, i0 = 4, 4 + 39, 4 + 39
    assert np.all(ops[2][t, j0:j0 + 3, i0:i0 + 3] != 0)
    for k, name, hs, tol in ((0, "aW2d", (1e-2, 1e-3, 1e-4, 1e-5), 1e-6), (1, "aC2d", (1e-3, 1e-4, 1e-5), 1e-5),
                             (2, "cg2d_b", (1e-1, 1e-2, 1e-3), 1e-8)):
        win = base[k][t, j0:j0 + 3, i0:i0 + 3].abs().max().item()
        dn = np.zeros(inner.shape)
        dn[t, j0:j0 + 3, i0:i0 + 3] = rng.normal(size=(3, 3)) * win
        dd = to_dev(dn, dev)
        gd = float((grads[k] * dd).sum())
        errs = []
        with torch.no_grad():
            for h in hs:
                plus = [a + h * dd if n == k else a for n, a in enumerate(base)]
                minus = [a - h * dd if n == k else a for n, a in enumerate(base)]
                fd = (float(J(tight, *plus)) - float(J(tight, *minus))) / (2 * h)
                errs.append(abs(fd - gd) / abs(gd))
        print(name, ["%.2e" % e for e in errs], flush=True)
        assert min(errs) < tol, (name, errs)


def test_stop_coeff_grad_and_batched_sum(dev, ex, case):
    """stop_coeff_grad: no cotangent reaches the coefficients (reference test_cg2d.py:280-300, ECCO mode). A batch
    of members over one shared operator: the coefficient cotangent is the sum of the members' cotangents."""
    import torch
    from mit_jax_amd import adjoint, synthetic as S
    g, ops, norm, p = case
    inner = _inner(g)
    rng = np.random.default_rng(4)
    opsd = [to_dev(o, dev) for o in ops]
    bs = [to_dev(S.rhs(g, 7 + m), dev) for m in range(2)]
    ws = [to_dev(rng.normal(size=inner.shape) * inner, dev) for _ in range(2)]
    aW = opsd[0].clone().requires_grad_(True)
    pe = dataclasses.replace(p, stop_coeff_grad=True)
    b = bs[0].clone().requires_grad_(True)
    x, _ = adjoint.cg2d_solve(pe, ex, [aW] + opsd[1:], b, torch.zeros_like(b))
    J = (ws[0] * x).sum()
    J.backward()
    assert aW.grad is None and b.grad is not None
    # batched, shared operator
    aW = opsd[0].clone().requires_grad_(True)
    B = torch.stack(bs)
    x, _ = adjoint.cg2d_solve(p, ex, [aW] + opsd[1:], B, torch.zeros_like(B))
    (gB,) = torch.autograd.grad((torch.stack(ws) * x).sum(), [aW])
    gs = []
    for m in range(2):
        a = opsd[0].clone().requires_grad_(True)
        xm, _ = adjoint.cg2d_solve(p, ex, [a] + opsd[1:], bs[m], torch.zeros_like(bs[m]))
        gs.append(torch.autograd.grad((ws[m] * xm).sum(), [a])[0])
    torch.testing.assert_close(gB, gs[0] + gs[1], rtol=1e-12, atol=1e-12 * float(gB.abs().max()))


@pytest.mark.parametrize("kind,shape", [("T", (13, 98, 98)), ("3D", (13, 5, 98, 98)), ("UVs", (13, 98, 98)),
                                        ("UV3s", (13, 5, 98, 98)), ("UVn", (13, 98, 98))])
def test_exchange_transpose_dot_product(dev, ex, kind, shape):
    """<exch(x), y> = <x, exch^T(y)> over every padded point (the adjoint identity of the reference's
    test_exchange.py:106-118), for scalar and vector kinds; the transpose is deterministic (fixed summation order)."""
    rng = np.random.default_rng(7)
    p = ex.plan(kind, dev)
    if p.vector:
        u, v, yu, yv = (to_dev(rng.normal(size=shape), dev) for _ in range(4))
        eu, ev = ex.vector(u, v, kind)
        bu, bv = ex.vector_adjoint(yu, yv, kind)
        lhs = float((eu * yu).sum() + (ev * yv).sum())
        rhs = float((u * bu).sum() + (v * bv).sum())
        bu2, bv2 = ex.vector_adjoint(yu, yv, kind)
        assert np.array_equal(bu.cpu().numpy(), bu2.cpu().numpy()) and np.array_equal(bv.cpu().numpy(),
                                                                                       bv2.cpu().numpy())
    else:
        x, y = (to_dev(rng.normal(size=shape), dev) for _ in range(2))
        lhs = float((ex.scalar(x, kind) * y).sum())
        rhs = float((x * ex.scalar_adjoint(y, kind)).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
