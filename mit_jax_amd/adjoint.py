"""Differentiable CG2D solve: the literal forward with hand-written implicit-derivative reverse mode.

`cg2d_solve(p, ex, ops, cg2d_b, cg2d_x) -> (x, diag)` mirrors the reference's public entry (core/cg2d.py:351-375):
x is the literal cg2d_fortran solution on the interior with 0 halos; it is differentiable (torch.autograd) in
cg2d_b and the operator coefficients aW2d, aS2d, aC2d, and not in the preconditioner, the first guess or cg2dNorm
(core/cg2d.py:326-334). The backward pass is the transpose solve of the implicit derivative: for the cotangent w of
x, lambda = A^-1 where(inner, w, 0) by a tight CG from 0 in the "tile" order at p.adj_tolerance / p.adj_max_iters
(A is symmetric, the same solve serves as its transpose), cotangent of cg2d_b = lambda, cotangent of the
coefficients from mj_cg2d_coeff_grad (the transpose of dA -> -_matvec(dA, x)), skipped with p.stop_coeff_grad.
Batched inputs ([B, T, ny, nx]) are supported; cotangents of shared coefficient arrays are summed over members.
"""

import ctypes

import torch

from . import _lib
from .cg2d import Cg2dParams, _as_batch, _norm_tensor, cg2d_fortran, config, interior


def coeff_grad(ex, x, lam, cg2dNorm):
    """(gaW, gaS, gaC) per member for x, lam [T, ny, nx] or [B, T, ny, nx] (mj_cg2d_coeff_grad)."""
    L = ex.L
    batched = x.dim() == 4
    xb = _as_batch(x, L, "x")
    lb = _as_batch(lam, L, "lam")
    B = xb.shape[0]
    norm = _norm_tensor(cg2dNorm, B, xb.device)
    cfg = config(L, B, 0.0, 0, "tile")
    g = [torch.empty_like(xb) for _ in range(3)]
    work = torch.empty_like(xb)
    _lib.check(_lib.lib().mj_cg2d_coeff_grad(ex.plan("T", xb.device).handle, _lib.ptr(xb), _lib.ptr(lb),
                                             _lib.ptr(norm), *[_lib.ptr(a) for a in g], ctypes.byref(cfg),
                                             _lib.ptr(work), _lib.stream_ptr()), "mj_cg2d_coeff_grad")
    return tuple(g) if batched else tuple(a[0] for a in g)


class _Cg2dImplicit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, aW, aS, aC, b, p, ex, pre, x_first):
        ops = (aW, aS, aC) + tuple(pre)
        x, diag = cg2d_fortran(ex, ops, b, x_first, cg2dNorm=p.cg2dNorm, tolerance=p.cg2dTolerance,
                               max_iters=p.cg2dMaxIters, sum_order=p.sum_order, nIterMin=p.nIterMin,
                               sum_unroll=p.sum_unroll)
        inner = interior(ex.L, b.device)
        xin = torch.where(inner, x, 0.0)
        ctx.save_for_backward(xin, aW, aS, aC, *pre)
        ctx.p, ctx.ex = p, ex
        ctx.mark_non_differentiable(diag.raw)
        ctx.diag = diag
        return xin, diag.raw

    @staticmethod
    def backward(ctx, gx, _graw):
        xin, aW, aS, aC, pW, pS, pC = ctx.saved_tensors
        p, ex = ctx.p, ctx.ex
        inner = interior(ex.L, xin.device)
        rhs = torch.where(inner, gx, 0.0).contiguous()
        y, d = cg2d_fortran(ex, (aW, aS, aC, pW, pS, pC), rhs, torch.zeros_like(rhs), cg2dNorm=p.cg2dNorm,
                            tolerance=p.adj_tolerance, max_iters=p.adj_max_iters, sum_order="tile")
        lam = torch.where(inner, y, 0.0).contiguous()
        gb = lam if ctx.needs_input_grad[3] else None
        ga = [None, None, None]
        if any(ctx.needs_input_grad[:3]) and not p.stop_coeff_grad:
            g = coeff_grad(ex, xin.contiguous(), lam, p.cg2dNorm)
            for k, a in enumerate((aW, aS, aC)):
                if ctx.needs_input_grad[k]:
                    gk = g[k]
                    if gk.dim() > a.dim():  # one operator shared by a batch of members
                        gk = gk.sum(0)
                    ga[k] = gk
        return ga[0], ga[1], ga[2], gb, None, None, None, None


def cg2d_solve(p: Cg2dParams, ex, ops, cg2d_b, cg2d_x):
    """Solve A x = b with the literal CG2D forward and implicit derivatives (reference core/cg2d.py:351-375).
    Returns (x with 0 halos, diag); diag["x_fortran"] is not provided (the halos of the literal solution are those
    of cg2d_fortran, available from cg2d.cg2d_fortran)."""
    from .cg2d import Cg2dDiag
    aW, aS, aC, pW, pS, pC = ops
    if p.stop_coeff_grad:
        aW, aS, aC = (a.detach() for a in (aW, aS, aC))
    x_first = cg2d_x.detach()
    x, raw = _Cg2dImplicit.apply(aW, aS, aC, cg2d_b, p, ex, (pW.detach(), pS.detach(), pC.detach()), x_first)
    return x, Cg2dDiag(raw, cg2d_b.dim() == 4)
