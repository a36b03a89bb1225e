"""GAD_IMPLICIT_R on the GPU with the reference's interface (mitgcm_jax core/tracer_implicit.py).

`gad_implicit_r(p, g, implicitAdvection, advectionScheme, kappaRX, recip_hFac, wFld, gTracer)` (reference
:165-173): the tracer after the implicit vertical diffusion and, with implicitAdvection, the 3rd-order upwind
implicit vertical advection, solved per water column (pentadiagonal elimination with the zero-pivot guard of
solve_pentadiagonal.F), through libmjamd's mj_gad_implicit_r; bitwise the reference's expressions. Arrays are
CUDA float64 [T, Nr, ny, nx] or [B, T, Nr, ny, nx]; maskC [T, Nr, ny, nx] (shared) or per member. p supplies
dTtracerLev (Nr values), implicitDiffusion and rkSign; g supplies layout (or L), maskC, rA, recip_rA, recip_drF and
recip_drC. As the reference, only implicitDiffusion = True and the U3 scheme (advectionScheme 3) are ported.

`gad_implicit_r_adjoint` is the reverse mode (cotangents of gTracer, kappaRX, wFld) and `gad_implicit_r_ad` the
torch.autograd op built on it.
"""

import ctypes

import numpy as np
import torch

from . import _lib

ENUM_UPWIND_3RD = 3


def _f64(a, device, name):
    if isinstance(a, torch.Tensor):
        if a.device.type != "cuda":
            raise TypeError(f"{name}: expected a CUDA tensor")
        return a.to(device=device, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)


def _batch(a, name):
    if not isinstance(a, torch.Tensor) or a.device.type != "cuda" or a.dtype != torch.float64:
        raise TypeError(f"{name}: expected a CUDA float64 tensor")
    if a.dim() == 4:
        a = a.unsqueeze(0)
    if a.dim() != 5:
        raise ValueError(f"{name}: expected [T, Nr, ny, nx] or [B, T, Nr, ny, nx], got {tuple(a.shape)}")
    return a.contiguous()


def _check_shapes(L, y, name, others):
    """The tile layout of y ([B, T, Nr, ny, nx]) and the same shape for every (tensor, name) of others: the kernels
    index all arrays with y's strides."""
    B, T, K, ny, nx = y.shape
    if (T, ny, nx) != (L.nTiles, L.ny, L.nx):
        raise ValueError(f"{name}: tile layout {(T, ny, nx)} is not {(L.nTiles, L.ny, L.nx)}")
    for t, n in others:
        if t is not None and t.shape != y.shape:
            raise ValueError(f"{n}: shape {tuple(t.shape)} differs from {name}'s {tuple(y.shape)}")


def _check_branch(p, implicitAdvection, advectionScheme):
    if not p.implicitDiffusion:
        raise NotImplementedError("GAD_IMPLICIT_R without implicitDiffusion: the V4r4 run always has it")
    if implicitAdvection and advectionScheme != ENUM_UPWIND_3RD:
        raise NotImplementedError(f"GAD_IMPLICIT_R: vertical scheme {advectionScheme} not ported (V4r4: 3)")


def _structs(p, g, L, B, T, K, ny, nx, implicitAdvection, dev):
    mC = _f64(g.maskC, dev, "maskC")
    if mC.dim() == 4:
        mms = 0
        if tuple(mC.shape) != (T, K, ny, nx):
            raise ValueError("maskC: expected [T, Nr, ny, nx] or [B, T, Nr, ny, nx]")
    else:
        if tuple(mC.shape) != (B, T, K, ny, nx):
            raise ValueError("maskC: expected [T, Nr, ny, nx] or [B, T, Nr, ny, nx]")
        mms = T * K * ny * nx
    keep = dict(maskC=mC, rA=_f64(g.rA, dev, "rA"), recip_rA=_f64(g.recip_rA, dev, "recip_rA"),
                dt=_f64(np.asarray(p.dTtracerLev, np.float64).reshape(-1)[:K], dev, "dTtracerLev"),
                rdrF=_f64(np.asarray(g.recip_drF, np.float64).reshape(-1)[:K], dev, "recip_drF"),
                rdrC=_f64(np.asarray(g.recip_drC, np.float64).reshape(-1)[:K], dev, "recip_drC"))
    for n in ("dt", "rdrF", "rdrC"):
        if keep[n].numel() != K:
            raise ValueError(f"{n}: fewer than {K} levels")
    gs = _lib.ImprGrid(keep["maskC"].data_ptr(), keep["rA"].data_ptr(), keep["recip_rA"].data_ptr(),
                       keep["dt"].data_ptr(), keep["rdrF"].data_ptr(), keep["rdrC"].data_ptr(), float(p.rkSign))
    cfg = _lib.ImprConfig(B, T, L.sNx, L.sNy, L.OLx, K, T * K * ny * nx, mms, 1 if implicitAdvection else 0)
    return gs, cfg, keep


def gad_implicit_r_adjoint(p, g, implicitAdvection, advectionScheme, kappaRX, recip_hFac, wFld, u, ubar):
    """Reverse mode of gad_implicit_r at its result u: (gTracer_bar, kappaRX_bar, wFld_bar) for the cotangent ubar
    of u (wFld_bar None without implicit advection); batched like u."""
    _check_branch(p, implicitAdvection, advectionScheme)
    L = getattr(g, "layout", None) or g.L
    batched = u.dim() == 5
    uu, ub = _batch(u, "u"), _batch(ubar, "ubar")
    _check_shapes(L, uu, "u", [(ub, "ubar")])
    B, T, K, ny, nx = uu.shape
    if K <= 1:
        z = torch.zeros_like(uu)
        r = (ub.clone(), z, z.clone() if implicitAdvection else None)
        return r if batched else tuple(x[0] if x is not None else None for x in r)
    kap, rh = _batch(kappaRX, "kappaRX"), _batch(recip_hFac, "recip_hFac")
    w = _batch(wFld, "wFld") if implicitAdvection else None
    _check_shapes(L, uu, "u", [(kap, "kappaRX"), (rh, "recip_hFac"), (w, "wFld")])
    gs, cfg, keep = _structs(p, g, L, B, T, K, ny, nx, implicitAdvection, uu.device)
    yb, kb = torch.empty_like(uu), torch.empty_like(uu)
    wb = torch.empty_like(uu) if implicitAdvection else None
    P = _lib.ptr
    _lib.check(_lib.lib().mj_gad_implicit_r_adjoint(ctypes.byref(gs), ctypes.byref(cfg), P(kap), P(rh), P(w), P(uu),
                                                    P(ub), P(yb), P(kb), P(wb), _lib.stream_ptr()),
               "mj_gad_implicit_r_adjoint")
    r = (yb, kb, wb)
    return r if batched else tuple(x[0] if x is not None else None for x in r)


class _ImplicitR(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gTracer, kappaRX, wFld, p, g, implicitAdvection, advectionScheme, recip_hFac):
        u = gad_implicit_r(p, g, implicitAdvection, advectionScheme, kappaRX, recip_hFac, wFld, gTracer)
        ctx.save_for_backward(kappaRX, wFld, recip_hFac, u)
        ctx.args = (p, g, implicitAdvection, advectionScheme)
        return u

    @staticmethod
    def backward(ctx, ubar):
        kap, w, rh, u = ctx.saved_tensors
        p, g, adv, scheme = ctx.args
        yb, kb, wb = gad_implicit_r_adjoint(p, g, adv, scheme, kap, rh, w, u, ubar.contiguous())
        return yb, kb, wb, None, None, None, None, None


def gad_implicit_r_ad(p, g, implicitAdvection, advectionScheme, kappaRX, recip_hFac, wFld, gTracer):
    """gad_implicit_r differentiable (torch.autograd) in gTracer, kappaRX and wFld (recip_hFac and the grid are
    constants), through mj_gad_implicit_r_adjoint."""
    return _ImplicitR.apply(gTracer, kappaRX, wFld, p, g, implicitAdvection, advectionScheme, recip_hFac)


def gad_implicit_r(p, g, implicitAdvection, advectionScheme, kappaRX, recip_hFac, wFld, gTracer):
    """GAD_IMPLICIT_R (reference core/tracer_implicit.py:165-173). Returns a new array shaped like gTracer."""
    _check_branch(p, implicitAdvection, advectionScheme)
    L = getattr(g, "layout", None) or g.L
    batched = gTracer.dim() == 5
    y = _batch(gTracer, "gTracer")
    _check_shapes(L, y, "gTracer", [])
    B, T, K, ny, nx = y.shape
    out = torch.empty_like(y)
    if K <= 1:
        out.copy_(y)
        return out if batched else out[0]
    kap, rh = _batch(kappaRX, "kappaRX"), _batch(recip_hFac, "recip_hFac")
    w = _batch(wFld, "wFld") if implicitAdvection else None
    _check_shapes(L, y, "gTracer", [(kap, "kappaRX"), (rh, "recip_hFac"), (w, "wFld")])
    gs, cfg, keep = _structs(p, g, L, B, T, K, ny, nx, implicitAdvection, y.device)
    P = _lib.ptr
    _lib.check(_lib.lib().mj_gad_implicit_r(ctypes.byref(gs), ctypes.byref(cfg), P(kap), P(rh), P(w), P(y), P(out),
                                            _lib.stream_ptr()), "mj_gad_implicit_r")
    return out if batched else out[0]
