"""The producers either side of the CG2D solve in SOLVE_FOR_PRESSURE, on the GPU, with the reference's interface
(mitgcm_jax core/solve_for_pressure.py).

`update_cg2d(p, cp, g, ex, hFacW, hFacS, recip_Bo, pW, pS, pC) -> (aW2d, aS2d, aC2d, pW, pS, pC)`
(reference :29-65, update_cg2d.F:59-195): the 5-point operator from the column sums of the r* hFacW/S and the
5-point preconditioner, through libmjamd's mj_update_cg2d. Returns new arrays; pW, pS, pC are copies of the
arguments with j, i = 1..sNy+1, 1..sNx+1 rewritten, as the reference's `.at[].set`. `update_cg2d_adjoint` is its
reverse mode in hFacW, hFacS (mj_update_cg2d_adjoint).

`cg2d_rhs(p, g, gU, gV, hFacW, hFacS, EmPmR, etaN, etaH, Bo_surf) -> (cg2d_b, cg2d_x)` (reference :68-106,
solve_for_pressure.F:124-255 with CALC_DIV_GHAT): through mj_cg2d_rhs; `cg2d_rhs_adjoint` is its reverse mode
(mj_cg2d_rhs_adjoint).

Both are bitwise equal to the literal Fortran-order loops. Field arrays are CUDA float64 [T, K, ny, nx] / [T, ny, nx],
or with a leading member axis [B, ...] for an ensemble (grid fields, recip_Bo and Bo_surf shared). `g` is a grid
object with `layout` (or `L`), `dyG, dxG, recip_dxC, recip_dyC, rA, maskInC, drF` (numpy or CUDA tensors); the
vertical factors are V4r4's (all 1, reference core/free_surface.py:121-125), so the library skips them.
"""

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .cg2d import _norm_tensor


@dataclass(frozen=True)
class FreeSurfParams:
    """The FreeSurfParams fields these producers read (reference core/free_surface.py:33-52); V4r4 defaults.
    A reference FreeSurfParams object with the same attribute names is accepted as well."""
    deltaTMom: float
    deltaTFreeSurf: float
    freeSurfFac: float = 1.0
    implicSurfPress: float = 1.0
    implicDiv2DFlow: float = 1.0
    mass2rUnit: float = 1.0 / 1029.0
    useRealFreshWaterFlux: bool = True


def _layout(g):
    L = getattr(g, "layout", None)
    return L if L is not None else g.L


def _fs(p):
    return _lib.FsParams(float(p.deltaTMom), float(p.deltaTFreeSurf), float(p.freeSurfFac), float(p.implicSurfPress),
                         float(p.implicDiv2DFlow), float(p.mass2rUnit), 1 if p.useRealFreshWaterFlux else 0)


def _dev2(a, device, name):
    if isinstance(a, torch.Tensor):
        if a.device.type != "cuda":
            raise TypeError(f"{name}: expected a CUDA tensor")
        return a.to(dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)


_GRID_FIELDS = ("dyG", "dxG", "recip_dxC", "recip_dyC", "rA", "maskInC")


class _GridDev:
    """Device copies of the grid fields the producers read, and the mj_sfp_grid pointing at them."""

    def __init__(self, g, device):
        L = _layout(g)
        self.keep = {k: _dev2(getattr(g, k), device, k) for k in _GRID_FIELDS}
        for k, t in self.keep.items():
            if tuple(t.shape) != (L.nTiles, L.ny, L.nx):
                raise ValueError(f"grid.{k}: expected [{L.nTiles}, {L.ny}, {L.nx}], got {tuple(t.shape)}")
        self.keep["drF"] = _dev2(np.asarray(getattr(g, "drF")).reshape(-1) if not isinstance(g.drF, torch.Tensor)
                                 else g.drF.reshape(-1), device, "drF")
        k = self.keep
        self.s = _lib.SfpGrid(*(k[n].data_ptr() for n in _GRID_FIELDS), None, k["drF"].data_ptr(), None, None, None)


_grid_cache = {}
_GRID_CACHE_MAX = 4


def _grid(g, device):
    """Device copies of g's fields, kept for the last few grids used (keyed by identity, so a grid object must not
    be modified in place between calls)."""
    key = (id(g), str(device))
    hit = _grid_cache.pop(key, None)
    if hit is None or hit[0] is not g:
        hit = (g, _GridDev(g, device))
    _grid_cache[key] = hit  # most recently used last
    while len(_grid_cache) > _GRID_CACHE_MAX:
        _grid_cache.pop(next(iter(_grid_cache)))
    return hit[1]


def _batch3(a, L, name):
    if not isinstance(a, torch.Tensor) or a.device.type != "cuda" or a.dtype != torch.float64:
        raise TypeError(f"{name}: expected a CUDA float64 tensor")
    if a.dim() == 4:
        a = a.unsqueeze(0)
    if a.dim() != 5 or tuple(a.shape[1:2]) != (L.nTiles,) or tuple(a.shape[3:]) != (L.ny, L.nx):
        raise ValueError(f"{name}: expected [{L.nTiles}, K, {L.ny}, {L.nx}] or [B, ...], got {tuple(a.shape)}")
    return a.contiguous()


def _batch2(a, L, B, name):
    if not isinstance(a, torch.Tensor) or a.device.type != "cuda" or a.dtype != torch.float64:
        raise TypeError(f"{name}: expected a CUDA float64 tensor")
    if a.dim() == 3:
        a = a.unsqueeze(0)
    if a.dim() != 4 or tuple(a.shape[1:]) != L.shape2d:
        raise ValueError(f"{name}: expected [{L.nTiles}, {L.ny}, {L.nx}] or [B, ...], got {tuple(a.shape)}")
    if a.shape[0] == 1 and B > 1:
        a = a.expand(B, *a.shape[1:])
    if a.shape[0] != B:
        raise ValueError(f"{name}: {a.shape[0]} members, expected {B}")
    return a.contiguous()


def _shared2(a, L, device, name):
    t = _dev2(a, device, name)
    if tuple(t.shape) != L.shape2d:
        raise ValueError(f"{name}: expected [{L.nTiles}, {L.ny}, {L.nx}] (shared by all members)")
    return t


def _config(L, B, K):
    return _lib.SfpConfig(B, L.nTiles, L.sNx, L.sNy, L.OLx, K, L.nTiles * K * L.ny * L.nx, L.npad)


def _check_layout(L):
    if L.OLx != L.OLy:
        raise ValueError("the library's tile layout has OLx == OLy")


def ini_cg2d_norm(g, hFacW, hFacS, implicSurfPress, implicDiv2DFlow):
    """cg2dNorm = 1/max|aW2d, aS2d| of the initial operator (reference core/cg2d.py:108-134) on the GPU
    (libmjamd mj_ini_cg2d_norm). hFacW, hFacS: CUDA float64 [T, K, ny, nx] (a float is returned, as the reference
    does) or [B, T, K, ny, nx] (a CUDA tensor of B norms)."""
    L = _layout(g)
    _check_layout(L)
    batched = hFacW.dim() == 5
    hW, hS = _batch3(hFacW, L, "hFacW"), _batch3(hFacS, L, "hFacS")
    if hW.shape != hS.shape:
        raise ValueError("hFacW and hFacS differ in shape")
    B, K, dev = hW.shape[0], hW.shape[2], hW.device
    gd = _grid(g, dev)
    if gd.keep["drF"].numel() < K:
        raise ValueError(f"grid.drF has {gd.keep['drF'].numel()} levels, the hFac arrays {K}")
    norm = torch.empty(B, dtype=torch.float64, device=dev)
    cfg = _config(L, B, K)
    _lib.check(_lib.lib().mj_ini_cg2d_norm(ctypes.byref(gd.s), ctypes.byref(cfg), float(implicSurfPress),
                                           float(implicDiv2DFlow), _lib.ptr(hW), _lib.ptr(hS), _lib.ptr(norm),
                                           _lib.stream_ptr()), "mj_ini_cg2d_norm")
    return norm if batched else float(norm[0])


def update_cg2d(p, cp, g, ex, hFacW, hFacS, recip_Bo, pW, pS, pC):
    """UPDATE_CG2D (reference core/solve_for_pressure.py:29-65). cp: Cg2dParams (cg2dNorm: a scalar or one value
    per member; cg2dpcOffDFac). Returns (aW2d, aS2d, aC2d, pW, pS, pC), batched like hFacW."""
    L = _layout(g)
    _check_layout(L)
    batched = hFacW.dim() == 5
    hW, hS = _batch3(hFacW, L, "hFacW"), _batch3(hFacS, L, "hFacS")
    if hW.shape != hS.shape:
        raise ValueError("hFacW and hFacS differ in shape")
    B, K, dev = hW.shape[0], hW.shape[2], hW.device
    gd = _grid(g, dev)
    if gd.keep["drF"].numel() < K:
        raise ValueError(f"grid.drF has {gd.keep['drF'].numel()} levels, the hFac arrays {K}")
    rBo = _shared2(recip_Bo, L, dev, "recip_Bo")
    norm = _norm_tensor(cp.cg2dNorm, B, dev)
    aW, aS, aC = (torch.empty((B,) + L.shape2d, dtype=torch.float64, device=dev) for _ in range(3))
    pWn, pSn, pCn = (_batch2(a, L, B, n).clone() for a, n in ((pW, "pW"), (pS, "pS"), (pC, "pC")))
    cfg, fs = _config(L, B, K), _fs(p)
    plan = ex.plan("T", dev)
    P = _lib.ptr
    _lib.check(_lib.lib().mj_update_cg2d(plan.handle, ctypes.byref(gd.s), ctypes.byref(fs), ctypes.byref(cfg),
                                         P(norm), float(cp.cg2dpcOffDFac), P(hW), P(hS), P(rBo), P(aW), P(aS), P(aC),
                                         P(pWn), P(pSn), P(pCn), _lib.stream_ptr()), "mj_update_cg2d")
    out = (aW, aS, aC, pWn, pSn, pCn)
    return out if batched else tuple(o[0] for o in out)


def cg2d_rhs(p, g, gU, gV, hFacW, hFacS, EmPmR, etaN, etaH, Bo_surf):
    """cg2d_b and the first guess cg2d_x (reference core/solve_for_pressure.py:68-106). Returns (cg2d_b, cg2d_x),
    batched like gU. EmPmR may be None when p.useRealFreshWaterFlux is false."""
    L = _layout(g)
    _check_layout(L)
    batched = gU.dim() == 5
    arrs = [_batch3(a, L, n) for a, n in ((gU, "gU"), (gV, "gV"), (hFacW, "hFacW"), (hFacS, "hFacS"))]
    if len({tuple(a.shape) for a in arrs}) != 1:
        raise ValueError("gU, gV, hFacW, hFacS differ in shape")
    B, K, dev = arrs[0].shape[0], arrs[0].shape[2], arrs[0].device
    gd = _grid(g, dev)
    if gd.keep["drF"].numel() < K:
        raise ValueError(f"grid.drF has {gd.keep['drF'].numel()} levels, the 3-D arrays {K}")
    if EmPmR is None and p.useRealFreshWaterFlux:
        raise ValueError("useRealFreshWaterFlux needs EmPmR")
    em = _batch2(EmPmR, L, B, "EmPmR") if EmPmR is not None else None
    eN, eH = _batch2(etaN, L, B, "etaN"), _batch2(etaH, L, B, "etaH")
    Bo = _shared2(Bo_surf, L, dev, "Bo_surf")
    b, x = (torch.empty((B,) + L.shape2d, dtype=torch.float64, device=dev) for _ in range(2))
    cfg, fs = _config(L, B, K), _fs(p)
    P = _lib.ptr
    _lib.check(_lib.lib().mj_cg2d_rhs(ctypes.byref(gd.s), ctypes.byref(fs), ctypes.byref(cfg), *(P(a) for a in arrs),
                                      P(em), P(eN), P(eH), P(Bo), P(b), P(x), _lib.stream_ptr()), "mj_cg2d_rhs")
    return (b, x) if batched else (b[0], x[0])


def cg2d_rhs_adjoint(p, g, gU, gV, hFacW, hFacS, Bo_surf, cg2d_b_bar, cg2d_x_bar=None):
    """Reverse mode of cg2d_rhs: from the cotangents of (cg2d_b, cg2d_x), those of (gU, gV, hFacW, hFacS, EmPmR,
    etaN, etaH), batched like gU (the rhs is linear in gU, gV, EmPmR, etaN, etaH, and in hFacW, hFacS for fixed gU,
    gV)."""
    L = _layout(g)
    _check_layout(L)
    batched = gU.dim() == 5
    arrs = [_batch3(a, L, n) for a, n in ((gU, "gU"), (gV, "gV"), (hFacW, "hFacW"), (hFacS, "hFacS"))]
    if len({tuple(a.shape) for a in arrs}) != 1:
        raise ValueError("gU, gV, hFacW, hFacS differ in shape")
    B, K, dev = arrs[0].shape[0], arrs[0].shape[2], arrs[0].device
    gd = _grid(g, dev)
    bb = _batch2(cg2d_b_bar, L, B, "cg2d_b_bar")
    xb = _batch2(cg2d_x_bar, L, B, "cg2d_x_bar") if cg2d_x_bar is not None else None
    Bo = _shared2(Bo_surf, L, dev, "Bo_surf")
    o3 = [torch.empty_like(arrs[0]) for _ in range(4)]
    o2 = [torch.empty((B,) + L.shape2d, dtype=torch.float64, device=dev) for _ in range(3)]
    cfg, fs = _config(L, B, K), _fs(p)
    P = _lib.ptr
    _lib.check(_lib.lib().mj_cg2d_rhs_adjoint(ctypes.byref(gd.s), ctypes.byref(fs), ctypes.byref(cfg),
                                              *(P(a) for a in arrs), P(Bo), P(bb), P(xb), *(P(o) for o in o3),
                                              *(P(o) for o in o2), _lib.stream_ptr()), "mj_cg2d_rhs_adjoint")
    out = tuple(o3) + tuple(o2)
    return out if batched else tuple(o[0] for o in out)


def update_cg2d_adjoint(p, cp, g, ex, aW2d_bar, aS2d_bar, aC2d_bar, nr):
    """Reverse mode of update_cg2d in hFacW, hFacS: from the cotangents of (aW2d, aS2d, aC2d) (the preconditioner's
    carry none), those of (hFacW, hFacS) with nr levels; batched like aW2d_bar."""
    L = _layout(g)
    _check_layout(L)
    batched = aW2d_bar.dim() == 4
    aWb = _batch2(aW2d_bar, L, aW2d_bar.shape[0] if batched else 1, "aW2d_bar")
    B, dev = aWb.shape[0], aWb.device
    aSb, aCb = _batch2(aS2d_bar, L, B, "aS2d_bar"), _batch2(aC2d_bar, L, B, "aC2d_bar")
    gd = _grid(g, dev)
    if gd.keep["drF"].numel() < nr:
        raise ValueError(f"grid.drF has {gd.keep['drF'].numel()} levels, asked for {nr}")
    norm = _norm_tensor(cp.cg2dNorm, B, dev)
    hWb = torch.empty((B, L.nTiles, nr, L.ny, L.nx), dtype=torch.float64, device=dev)
    hSb = torch.empty_like(hWb)
    work = torch.empty((B,) + L.shape2d, dtype=torch.float64, device=dev)
    cfg, fs = _config(L, B, nr), _fs(p)
    plan = ex.plan("T", dev)
    P = _lib.ptr
    _lib.check(_lib.lib().mj_update_cg2d_adjoint(plan.handle, ctypes.byref(gd.s), ctypes.byref(fs), ctypes.byref(cfg),
                                                 P(norm), P(aWb), P(aSb), P(aCb), P(hWb), P(hSb), P(work),
                                                 _lib.stream_ptr()), "mj_update_cg2d_adjoint")
    return (hWb, hSb) if batched else (hWb[0], hSb[0])
