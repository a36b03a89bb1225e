"""GAD DST3 tracer advection on the GPU, with the reference's interface (mitgcm_jax pkgs/gad.py).

`gad_advection(params, g, uFld, vFld, wFld, tracer, hFacW, hFacS, recip_hFacC) -> gTracer` (reference
gad.py:546-560) and `advect_tiles(params, tables, grid2d, vgrid, ...)` (:460-539) run GAD_ADVECTION's three-pass
multi-dimensional DST3 on the LLC cubed sphere in one HIP launch (libmjamd mj_gad_dst3): one workgroup per (tile,
level) keeps the tile in LDS through the per-facet X/Y sweeps, corner fills and update regions, and writes
gTracer = (localTij - tracer)/deltaTLev(k) on the whole tile array. Fields are CUDA float64 [T, K, ny, nx], or
[B, T, K, ny, nx] for an ensemble in the same launch (optional per-member level counts).

`implicitAdvection=True` is the reference branch (vertical advection left to the implicit solver, wFld not read).
`implicitAdvection=False` adds the vertical DST3 flux (R) of wFld after the horizontal passes, a formulation this
library defines (the reference raises NotImplementedError for it, gad.py:164-166); see mj_gad.hip.
`dst3_adv_x` / `dst3_adv_y` (gad.py:409-453) are the interface fluxes alone (libmjamd mj_dst3_adv_x / _y), for
callers that assemble their own update; the fused kernel evaluates the same values inside its sweeps.
"""

import ctypes
import sys
from dataclasses import dataclass, replace

import numpy as np
import torch

from . import _lib
from .layout import LLC90, Layout

ENUM_DST3 = 30
ONE_SIXTH = 1.0 / 6.0
LLC90_FACETS = (90, 270, 90, 270, 90, 90, 270, 90, 270, 90)


@dataclass(frozen=True)
class TileTopology:
    """exch2_myFace and facet-edge flags per tile (reference gad.py:75-106, w2_set_map_tiles.F:161-178,
    w2_set_tile2tiles.F:88-110)."""
    face: tuple
    N_edge: tuple
    S_edge: tuple
    E_edge: tuple
    W_edge: tuple

    @classmethod
    def from_facets(cls, dims_facets=LLC90_FACETS, layout=LLC90):
        L = layout
        face, N, S, E, W = [], [], [], [], []
        for f in range(len(dims_facets) // 2):
            fNx, fNy = dims_facets[2 * f], dims_facets[2 * f + 1]
            for ty in range(fNy // L.sNy):
                for tx in range(fNx // L.sNx):
                    face.append(f + 1)
                    N.append(not (ty + 1) * L.sNy < fNy)
                    S.append(ty == 0)
                    E.append(not (tx + 1) * L.sNx < fNx)
                    W.append(tx == 0)
        if len(face) != L.nTiles:
            raise ValueError(f"facets give {len(face)} tiles, layout has {L.nTiles}")
        return cls(tuple(face), tuple(N), tuple(S), tuple(E), tuple(W))


@dataclass(frozen=True)
class GADParams:
    """One GAD_ADVECTION call (reference gad.py:124-187). Per-level tuples of Nr values."""
    topo: TileTopology
    layout: Layout = LLC90
    dTtracerLev: tuple = (3600.0,) * 50
    deepFacC: tuple = (1.0,) * 50
    recip_deepFacC: tuple = (1.0,) * 50
    recip_deepFac2C: tuple = (1.0,) * 50
    recip_rhoFacC: tuple = (1.0,) * 50
    oneSixth: float = ONE_SIXTH
    advectionScheme: int = ENUM_DST3
    implicitAdvection: bool = True
    useCubedSphereExchange: bool = True
    trIdentity: int = 1

    def __post_init__(self):
        if self.advectionScheme != ENUM_DST3:
            raise NotImplementedError(f"advectionScheme={self.advectionScheme}: only ENUM_DST3=30 is implemented")
        if not self.useCubedSphereExchange:
            raise NotImplementedError("useCubedSphereExchange=F (npass=2 branch) is not implemented")

    @classmethod
    def llc90(cls, dt=3600.0, Nr=50, **kw):
        one = (1.0,) * Nr
        return cls(TileTopology.from_facets(), LLC90, (float(dt),) * Nr, one, one, one, one, **kw)


class GadPlan:
    """Device pass table of a topology (mj_gad_plan) on one device."""

    def __init__(self, topo: TileTopology, L: Layout, device):
        face = np.asarray(topo.face, np.int32)
        edges = np.asarray([[n, s, e, w] for n, s, e, w in zip(topo.N_edge, topo.S_edge, topo.E_edge, topo.W_edge)],
                           np.uint8)
        h = ctypes.c_void_p()
        with torch.cuda.device(device):
            _lib.check(_lib.lib().mj_gad_plan_create(len(face), _lib.ptr(face), _lib.ptr(edges), L.sNx, L.sNy, L.OLx,
                                                     L.OLy, ctypes.byref(h)), "mj_gad_plan_create")
        self.handle = h
        self._destroy = _lib.lib().mj_gad_plan_destroy
        self.ntiles = len(face)
        self.L = L

    def describe(self):
        """(masks uint8 [3, 2, T, ny*nx], fill_dirs int32 [3, 4, T]) rasterised from the plan."""
        T, n = self.ntiles, self.L.ny * self.L.nx
        masks = np.zeros((3, 2, T, n), np.uint8)
        fills = np.zeros((3, 4, T), np.int32)
        _lib.check(_lib.lib().mj_gad_plan_describe(self.handle, _lib.ptr(masks), _lib.ptr(fills)),
                   "mj_gad_plan_describe")
        return masks, fills

    def close(self):
        """Destroys the device plan now (idempotent); the object is unusable afterwards."""
        h, destroy = getattr(self, "handle", None), getattr(self, "_destroy", None)
        self.handle = None
        if h and destroy is not None:
            destroy(h)

    def __del__(self, _finalizing=sys.is_finalizing):
        # the destroy function is bound at creation: module globals may already be gone when this runs late, and at
        # interpreter exit the process end releases the device memory anyway
        if not _finalizing():
            self.close()


def plan_tables_host(topo: TileTopology, L: Layout = LLC90):
    """The library's update regions and fill calls for `topo`, rasterised on the host (no GPU needed):
    (masks uint8 [3, 2, T, ny*nx], fill_dirs int32 [3, 4, T])."""
    face = np.asarray(topo.face, np.int32)
    edges = np.asarray([[n, s, e, w] for n, s, e, w in zip(topo.N_edge, topo.S_edge, topo.E_edge, topo.W_edge)],
                       np.uint8)
    T, n = len(face), L.ny * L.nx
    masks = np.zeros((3, 2, T, n), np.uint8)
    fills = np.zeros((3, 4, T), np.int32)
    _lib.check(_lib.lib().mj_gad_tables_host(T, _lib.ptr(face), _lib.ptr(edges), L.sNx, L.sNy, L.OLx, L.OLy,
                                             _lib.ptr(masks), _lib.ptr(fills)), "mj_gad_tables_host")
    return masks, fills


_plans = {}


def gad_plan(topo, L, device):
    dev = torch.device(device)
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    key = (topo, L, idx)
    if key not in _plans:
        _plans[key] = GadPlan(topo, L, idx)
    return _plans[key]


def gad_tables(params):
    """The device plan standing in for the reference's table dict (gad.py:386-389)."""
    return gad_plan(params.topo, params.layout, torch.cuda.current_device())


GRID2D = ("dyG", "dxG", "recip_dxC", "recip_dyC", "recip_rA", "maskInC", "maskW", "maskS")
VGRID = ("drF", "recip_drF", "rhoFacC")


def _dev(a, device, name):
    if isinstance(a, torch.Tensor):
        if a.device.type != "cuda":
            raise TypeError(f"{name}: expected a CUDA tensor")
        return a.to(dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)


def _prepare(params, tables, grid2d, vgrid, uFld, vFld, tracer, hFacW, hFacS, recip_hFacC, wFld, nlev):
    """Checked device arguments of the DST3 entry points."""
    L = params.layout
    if not isinstance(tracer, torch.Tensor) or tracer.device.type != "cuda":
        raise TypeError("tracer: expected a CUDA tensor (the advection runs on the GPU only)")
    dev = tracer.device
    batched = tracer.dim() == 5
    tr = tracer.to(torch.float64).contiguous()
    tr5 = tr if batched else tr.unsqueeze(0)
    B, T, K, ny, nx = tr5.shape
    if (T, ny, nx) != L.shape2d:
        raise ValueError(f"tracer shape {tuple(tracer.shape)} does not match the layout")
    fields = [_dev(a, dev, n) for a, n in ((uFld, "uFld"), (vFld, "vFld"), (hFacW, "hFacW"), (hFacS, "hFacS"),
                                           (recip_hFacC, "recip_hFacC"))]
    for f, n in zip(fields, ("uFld", "vFld", "hFacW", "hFacS", "recip_hFacC")):
        if tuple(f.shape) != tuple(tr.shape):
            raise ValueError(f"{n}: shape {tuple(f.shape)} != tracer {tuple(tr.shape)}")
    w = None
    if wFld is not None and not params.implicitAdvection:
        w = _dev(wFld, dev, "wFld")
        if tuple(w.shape) != tuple(tr.shape):
            raise ValueError("wFld shape differs from tracer")
    g = {k: _dev(grid2d[k], dev, k) for k in grid2d if grid2d[k] is not None and k != "mask_index"}
    for k in ("dyG", "dxG", "recip_dxC", "recip_dyC", "recip_rA", "maskInC"):
        if tuple(g[k].shape) != L.shape2d:
            raise ValueError(f"{k}: expected {L.shape2d}")
    # masks: one shared slot [T, K, ny, nx], one per member [B, T, K, ny, nx], or C slots [C, T, K, ny, nx] picked
    # per member by grid2d["mask_index"] (members of one level class share their grid masks)
    mshape = tuple(g["maskW"].shape)
    midx = None
    if mshape == (T, K, ny, nx):
        mms = 0
    elif len(mshape) == 5 and mshape[1:] == (T, K, ny, nx) and batched:
        mms = T * K * ny * nx
        mi = grid2d.get("mask_index")
        if mi is not None:
            midx = torch.as_tensor(np.asarray(mi, np.int32).reshape(-1), device=dev) \
                if not isinstance(mi, torch.Tensor) else mi.to(device=dev, dtype=torch.int32).contiguous()
            if midx.numel() != B or int(midx.min()) < 0 or int(midx.max()) >= mshape[0]:
                raise ValueError(f"mask_index: {B} slot numbers in 0..{mshape[0] - 1} expected")
        elif mshape[0] != B:
            raise ValueError(f"maskW: {mshape[0]} mask slots for {B} members need grid2d['mask_index']")
    else:
        raise ValueError(f"maskW: shape {mshape} is neither [T, K, ny, nx], per member nor per slot")
    for k in ("maskS", "maskC"):
        if k in g and tuple(g[k].shape) != mshape:
            raise ValueError(f"{k}: shape {tuple(g[k].shape)} differs from maskW {mshape}")
    vert = {
        "dTlev": params.dTtracerLev, "deepFacC": params.deepFacC, "recip_deepFacC": params.recip_deepFacC,
        "recip_deepFac2C": params.recip_deepFac2C, "recip_rhoFacC": params.recip_rhoFacC,
        "drF": vgrid["drF"], "recip_drF": vgrid["recip_drF"], "rhoFacC": vgrid["rhoFacC"],
        "recip_drC": vgrid.get("recip_drC"),
    }
    vt = {}
    for k, v in vert.items():
        if v is None:
            continue
        t = _dev(np.asarray(v, np.float64) if not isinstance(v, torch.Tensor) else v, dev, k).reshape(-1)
        if t.numel() < K:
            raise ValueError(f"{k}: {t.numel()} levels < {K}")
        vt[k] = t
    if w is not None:
        for k in ("rA", "maskC"):
            if k not in g:
                raise ValueError(f"the vertical flux needs grid2d['{k}']")
        if "recip_drC" not in vt:
            raise ValueError("the vertical flux needs vgrid['recip_drC']")
    P = lambda k, d=g: _lib.ptr(d.get(k))  # noqa: E731
    grid_s = _lib.GadGrid(P("dyG"), P("dxG"), P("recip_dxC"), P("recip_dyC"), P("recip_rA"), P("rA"), P("maskInC"),
                          P("maskW"), P("maskS"), P("maskC"), P("dTlev", vt), P("drF", vt), P("recip_drF", vt),
                          P("rhoFacC", vt), P("recip_rhoFacC", vt), P("deepFacC", vt), P("recip_deepFacC", vt),
                          P("recip_deepFac2C", vt), P("recip_drC", vt), float(params.oneSixth))
    nl = None
    if nlev is not None:
        nl = torch.as_tensor(np.asarray(nlev, np.int32).reshape(-1), device=dev) if not isinstance(nlev, torch.Tensor) \
            else nlev.to(device=dev, dtype=torch.int32).contiguous()
        if nl.numel() != B or int(nl.max()) > K or int(nl.min()) < 0:
            raise ValueError("nlev: one value in 0..K per member")
    cfg = _lib.GadConfig(B, K, T * K * ny * nx, mms, _lib.ptr(nl), _lib.ptr(midx))
    cfg.mask_slots = mshape[0] if len(mshape) == 5 else 1
    plan = tables if isinstance(tables, GadPlan) else gad_plan(params.topo, L, dev)
    keep = (g, vt, nl, midx)
    return plan, grid_s, cfg, tr, fields, w, keep


def _dst3_adv(direction, L, oneSixth, deltaTloc, trans, vel, maskLoc, tracer, rd, recip_deepFacC):
    if not isinstance(tracer, torch.Tensor) or tracer.device.type != "cuda":
        raise TypeError("tracer: expected a CUDA tensor (the flux kernels run on the GPU only)")
    dev = tracer.device
    tr = tracer.to(torch.float64).contiguous()
    tr5 = tr if tr.dim() == 5 else tr.unsqueeze(0)
    if tr5.dim() != 5 or tuple(tr5.shape[1:2]) + tuple(tr5.shape[3:]) != L.shape2d:
        raise ValueError(f"tracer: expected [T, K, ny, nx] or [B, T, K, ny, nx] of the layout, got {tuple(tr.shape)}")
    B, T, K, ny, nx = tr5.shape
    f = [_dev(a, dev, n) for a, n in ((trans, "trans"), (vel, "vel"), (maskLoc, "mask"))]
    for a, n in zip(f, ("trans", "vel", "mask")):
        if tuple(a.shape) != tuple(tr.shape):
            raise ValueError(f"{n}: shape {tuple(a.shape)} != tracer {tuple(tr.shape)}")
    r = _dev(rd, dev, "recip_dxC/recip_dyC").reshape(-1)
    if r.numel() != T * ny * nx:
        raise ValueError("recip_dxC/recip_dyC: expected [T, ny, nx] (or [T, 1, ny, nx])")
    dt = _dev(np.broadcast_to(np.asarray(deltaTloc, np.float64).reshape(-1), (K,)).copy()
              if not isinstance(deltaTloc, torch.Tensor) else deltaTloc.reshape(-1).expand(K), dev, "deltaTloc")
    rdp = _dev(np.broadcast_to(np.asarray(recip_deepFacC, np.float64).reshape(-1), (K,)).copy()
               if not isinstance(recip_deepFacC, torch.Tensor) else recip_deepFacC.reshape(-1).expand(K), dev,
               "recip_deepFacC")
    out = torch.empty_like(tr)
    cfg = _lib.Dst3FluxConfig(B, T, K, L.sNx, L.sNy, L.OLx, L.OLy, T * K * ny * nx)
    fn = _lib.lib().mj_dst3_adv_x if direction == 0 else _lib.lib().mj_dst3_adv_y
    _lib.check(fn(ctypes.byref(cfg), float(oneSixth), _lib.ptr(dt), *[_lib.ptr(a) for a in f], _lib.ptr(tr),
                  _lib.ptr(r), _lib.ptr(rdp), _lib.ptr(out), _lib.stream_ptr()),
               "mj_dst3_adv_x" if direction == 0 else "mj_dst3_adv_y")
    return out


def dst3_adv_x(L, oneSixth, deltaTloc, uTrans, uFld, maskLocW, tracer, recip_dxC, recip_deepFacC):
    """GAD_DST3_ADV_X with calcCFL for all tiles and levels (reference gad.py:409-430) -> the interface flux uT,
    0 outside i = 2 .. nx-2. deltaTloc, recip_deepFacC: K values ([1, K, 1, 1] as the reference passes them, or
    [K]); recip_dxC: [T, ny, nx] or [T, 1, ny, nx]; the rest [T, K, ny, nx] (or [B, T, K, ny, nx]), CUDA float64."""
    return _dst3_adv(0, L, oneSixth, deltaTloc, uTrans, uFld, maskLocW, tracer, recip_dxC, recip_deepFacC)


def dst3_adv_y(L, oneSixth, deltaTloc, vTrans, vFld, maskLocS, tracer, recip_dyC, recip_deepFacC):
    """GAD_DST3_ADV_Y (reference gad.py:433-453) -> vT, 0 outside j = 2 .. ny-2; arguments as dst3_adv_x."""
    return _dst3_adv(1, L, oneSixth, deltaTloc, vTrans, vFld, maskLocS, tracer, recip_dyC, recip_deepFacC)


def _check_out(out, like, name):
    """A caller-supplied output: a contiguous CUDA float64 tensor of the tracer's shape on its device."""
    if not isinstance(out, torch.Tensor) or out.device != like.device or out.dtype != torch.float64:
        raise TypeError(f"{name}: expected a CUDA float64 tensor on {like.device}")
    if tuple(out.shape) != tuple(like.shape) or not out.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous tensor of shape {tuple(like.shape)}")
    return out


def advect_tiles(params, tables, grid2d, vgrid, uFld, vFld, tracer, hFacW, hFacS, recip_hFacC, wFld=None,
                 nlev=None, out=None, levels_per_block=1):
    """GAD_ADVECTION on all tiles (reference gad.py:460-539) -> gTracer, same shape as tracer.

    tables: gad_tables(params) (or None). grid2d: dyG, dxG, recip_dxC, recip_dyC, recip_rA, maskInC [T, ny, nx]
    and maskW, maskS [T, K, ny, nx] (or [B, T, K, ny, nx] per member); with wFld also rA [T, ny, nx] and maskC.
    vgrid: drF, recip_drF, rhoFacC [K] (+ recip_drC with wFld). nlev: per-member level counts (ensemble).
    levels_per_block: the launch's tile shape (consecutive levels per workgroup); the result does not depend on it."""
    plan, grid_s, cfg, tr, fields, w, keep = _prepare(params, tables, grid2d, vgrid, uFld, vFld, tracer, hFacW,
                                                      hFacS, recip_hFacC, wFld, nlev)
    cfg.levels_per_block = int(levels_per_block)
    if out is None:
        # with level counts the levels at or beyond a member's count are not written: they read 0
        gT = torch.zeros_like(tr) if nlev is not None else torch.empty_like(tr)
    else:
        gT = _check_out(out, tr, "out")
    lib = _lib.lib()
    wsb = lib.mj_gad_dst3_workspace_bytes(ctypes.byref(cfg), params.layout.nTiles, int(w is not None))
    work = torch.empty(wsb, dtype=torch.uint8, device=tr.device)
    u, v, hW, hS, rhC = fields
    _lib.check(lib.mj_gad_dst3(plan.handle, ctypes.byref(grid_s), ctypes.byref(cfg), _lib.ptr(u), _lib.ptr(v),
                               _lib.ptr(w), _lib.ptr(tr), _lib.ptr(hW), _lib.ptr(hS), _lib.ptr(rhC), _lib.ptr(gT),
                               _lib.ptr(work), ctypes.c_size_t(wsb), _lib.stream_ptr()), "mj_gad_dst3")
    del keep, work
    return gT


def advect_tiles_adjoint(params, tables, grid2d, vgrid, uFld, vFld, tracer, hFacW, hFacS, recip_hFacC, gT_bar,
                         nlev=None, wFld=None):
    """Reverse mode of advect_tiles (libmjamd mj_gad_dst3_adjoint): cotangents (tracer_bar, uFld_bar, vFld_bar) of
    the inputs for the cotangent gT_bar of gTracer (the reference's jax.grad of advect_tiles in the tracer and the
    velocities, test_gad.py:297-346); with wFld (and params.implicitAdvection False: the R flux) also wFld_bar, as
    a fourth element."""
    plan, grid_s, cfg, tr, fields, w, keep = _prepare(params, tables, grid2d, vgrid, uFld, vFld, tracer, hFacW,
                                                      hFacS, recip_hFacC, wFld, nlev)
    gb = gT_bar.to(torch.float64).contiguous()
    if tuple(gb.shape) != tuple(tr.shape):
        raise ValueError("gT_bar shape differs from tracer")
    bars = [torch.empty_like(tr) for _ in range(4 if w is not None else 3)]
    lib = _lib.lib()
    wsb = lib.mj_gad_dst3_adjoint_workspace_bytes(ctypes.byref(cfg), params.layout.nTiles, int(w is not None))
    work = torch.empty(wsb, dtype=torch.uint8, device=tr.device)
    u, v, hW, hS, rhC = fields
    _lib.check(lib.mj_gad_dst3_adjoint(plan.handle, ctypes.byref(grid_s), ctypes.byref(cfg), _lib.ptr(u), _lib.ptr(v),
                                       _lib.ptr(w), _lib.ptr(tr), _lib.ptr(hW), _lib.ptr(hS), _lib.ptr(rhC),
                                       _lib.ptr(gb), *[_lib.ptr(b) for b in bars[:3]],
                                       _lib.ptr(bars[3]) if w is not None else None, _lib.ptr(work),
                                       ctypes.c_size_t(wsb), _lib.stream_ptr()), "mj_gad_dst3_adjoint")
    del keep, work
    return tuple(bars)


class _AdvectTiles(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tracer, uFld, vFld, wFld, params, tables, grid2d, vgrid, hFacW, hFacS, recip_hFacC, nlev):
        ctx.save_for_backward(tracer, uFld, vFld, wFld)
        ctx.rest = (params, tables, grid2d, vgrid, hFacW, hFacS, recip_hFacC, nlev)
        return advect_tiles(params, tables, grid2d, vgrid, uFld, vFld, tracer, hFacW, hFacS, recip_hFacC, wFld=wFld,
                            nlev=nlev)

    @staticmethod
    def backward(ctx, gT_bar):
        tracer, uFld, vFld, wFld = ctx.saved_tensors
        params, tables, grid2d, vgrid, hFacW, hFacS, recip_hFacC, nlev = ctx.rest
        bars = advect_tiles_adjoint(params, tables, grid2d, vgrid, uFld, vFld, tracer, hFacW, hFacS, recip_hFacC,
                                    gT_bar, nlev=nlev, wFld=wFld)
        wb = bars[3] if len(bars) == 4 else None
        return bars[0], bars[1], bars[2], wb, None, None, None, None, None, None, None, None


def advect_tiles_ad(params, tables, grid2d, vgrid, uFld, vFld, tracer, hFacW, hFacS, recip_hFacC, nlev=None,
                    wFld=None):
    """advect_tiles differentiable in tracer, uFld and vFld (and wFld with the R flux, params.implicitAdvection
    False) through torch.autograd, with the hand-written reverse mode (advect_tiles_adjoint)."""
    if wFld is not None and params.implicitAdvection:
        wFld = None  # the reference branch: w is not read
    return _AdvectTiles.apply(tracer, uFld, vFld, wFld, params, tables, grid2d, vgrid, hFacW, hFacS, recip_hFacC,
                              nlev)


def gad_advection(params, g, uFld, vFld, wFld, tracer, hFacW, hFacS, recip_hFacC):
    """GAD_ADVECTION(advectionScheme=ENUM_DST3) on all tiles -> gTracer (reference gad.py:546-560). g: an object
    or dict with the GRID2D / VGRID fields (plus rA, maskC, recip_drC when implicitAdvection is False)."""
    get = (lambda k: g.get(k)) if isinstance(g, dict) else (lambda k: getattr(g, k, None))
    grid2d = {k: get(k) for k in GRID2D + ("rA", "maskC")}
    vgrid = {k: get(k) for k in VGRID + ("recip_drC",)}
    return advect_tiles(params, None, grid2d, vgrid, uFld, vFld, tracer, hFacW, hFacS, recip_hFacC,
                        wFld=None if params.implicitAdvection else wFld)
