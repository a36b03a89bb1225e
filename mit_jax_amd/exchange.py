"""exch2 halo exchanges on the GPU, with the reference's Exchanger interface (parallel/exchange.py:145-222).

`ExchangeMaps` holds the probed exch2 maps in compact form (the written points of each map: dst, src, comp, sign;
`data/exch2_halo_13x90x90.npz`, made by `data/make_halo_tables.py` from the reference's map file).
`Exchanger` runs them through libmjamd (mj_halo_exchange / mj_halo_exchange_uv) on CUDA float64 tensors
[tile, ..., j, i] (any number of leading levels) and returns new tensors, like the reference's functional gathers.
Kinds: T, 3D, Z, SMs (scalar) and UVs, UVn, UV3s, As, An, Bs, Bn, Ds (vector).
"""

import ctypes
import sys
import threading
from dataclasses import dataclass
from functools import lru_cache
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .layout import Layout

DATA = Path(__file__).resolve().parent / "data" / "exch2_halo_13x90x90.npz"
SOURCE_SHA256 = "a31f0d49754259c3b1304c42f6734886605c0f4f1f72257b0654e8486a9d84fd"
SCALAR_KINDS = ("T", "3D", "Z", "SMs")
VECTOR_KINDS = ("UVs", "UVn", "As", "An", "Bs", "Bn", "Ds", "UV3s")


@dataclass(frozen=True)
class ExchangeMaps:
    layout: Layout
    compact: dict  # map name -> (dst int32, src int32, comp int8, sign int8), written points in dst order

    @classmethod
    def load(cls, path=DATA):
        return _load_maps(str(path))

    def full(self, name):
        """The map over every padded point, as the reference stores it: (src int32, comp int8, sign int8);
        unwritten points keep their own value (src = own index, comp 0, sign +1)."""
        L = self.layout
        n = L.nTiles * L.ny * L.nx
        d, s, c, g = self.compact[name]
        src = np.arange(n, dtype=np.int32)
        comp = np.zeros(n, np.int8)
        sign = np.ones(n, np.int8)
        src[d], comp[d], sign[d] = s, c, g
        return src, comp, sign

    def names(self):
        return sorted(self.compact)

    @classmethod
    def from_full_npz(cls, path):
        """Maps from a file in the reference's format (reference parallel/exchange.py:72-92, ExchangeMaps.save/load:
        per map name `<name>.src`, `.comp`, `.sign` over every padded point, and `layout` = sNx, sNy, OLx, OLy,
        nTiles, Nr), e.g. maps probed from another exch2 configuration. Kept points must be the identity."""
        z = np.load(path, allow_pickle=False)
        sNx, sNy, OLx, OLy, nT, Nr = (int(v) for v in z["layout"])
        L = Layout(sNx, sNy, OLx, OLy, nT, Nr)
        n = nT * L.ny * L.nx
        names = sorted({k.rsplit(".", 1)[0] for k in z.files if k != "layout"})
        comp = {}
        for name in names:
            src = np.asarray(z[name + ".src"]).astype(np.int64)
            cmp_ = np.asarray(z[name + ".comp"]).astype(np.int8)
            sign = np.asarray(z[name + ".sign"]).astype(np.int8)
            if src.size != n or cmp_.size != n or sign.size != n:
                raise ValueError(f"{path}: map {name} does not cover the {n} padded points")
            keep = cmp_ == 0
            if not (np.all(src[keep] == np.arange(n)[keep]) and np.all(sign[keep] == 1)):
                raise ValueError(f"{path}: map {name}: a kept point is not the identity")
            if not (np.all(np.isin(cmp_, (0, 1, 2))) and np.all(np.isin(sign, (-1, 1))) and
                    np.all((src >= 0) & (src < n))):
                raise ValueError(f"{path}: map {name}: comp/sign/src out of range")
            dst = np.nonzero(~keep)[0]
            comp[name] = (dst.astype(np.int32), src[dst].astype(np.int32), cmp_[dst].copy(), sign[dst].copy())
        return cls(L, comp)

    def save_full(self, path):
        """Write the maps in the reference's format (the inverse of from_full_npz)."""
        L = self.layout
        arrs = {"layout": np.array([L.sNx, L.sNy, L.OLx, L.OLy, L.nTiles, L.Nr])}
        for name in self.names():
            arrs[name + ".src"], arrs[name + ".comp"], arrs[name + ".sign"] = self.full(name)
        np.savez_compressed(path, **arrs)


@lru_cache(maxsize=4)
def _load_maps(path):
    z = np.load(path, allow_pickle=False)
    if str(z["source_sha256"]) != SOURCE_SHA256:
        raise ValueError(f"{path}: not made from the known exch2 map file")
    sNx, sNy, OLx, OLy, nT, Nr = (int(v) for v in z["layout"])
    names = sorted({k.rsplit(".", 1)[0] for k in z.files if k not in ("layout", "source_sha256")})
    comp = {n: tuple(np.ascontiguousarray(z[n + s]) for s in (".dst", ".src", ".comp", ".sign")) for n in names}
    return ExchangeMaps(Layout(sNx, sNy, OLx, OLy, nT, Nr), comp)


class HaloPlan:
    """A device halo plan (mj_halo_plan) of one kind on one device."""

    def __init__(self, maps: ExchangeMaps, kind, device):
        L = maps.layout
        self.maps = maps  # held, so that the id keying the plan cache stays this object's
        names = [kind] if kind in SCALAR_KINDS else [kind + "_u", kind + "_v"]
        arrs = [maps.compact[n] for n in names]
        nm = len(arrs)
        npts = (ctypes.c_int32 * nm)(*[len(a[0]) for a in arrs])
        ptrs = [(ctypes.c_void_p * nm)(*[a[i].ctypes.data for a in arrs]) for i in range(4)]
        self._keep = arrs
        self.kind = kind
        self.device = device
        self.vector = nm == 2
        h = ctypes.c_void_p()
        with torch.cuda.device(device):
            _lib.check(_lib.lib().mj_halo_plan_create(nm, npts, *[ctypes.cast(p, ctypes.c_void_p) for p in ptrs],
                                                      L.nTiles, L.ny, L.nx, ctypes.byref(h)), "mj_halo_plan_create")
        self.handle = h
        self._destroy = _lib.lib().mj_halo_plan_destroy
        self.inplace_ok = bool(_lib.lib().mj_halo_plan_inplace_ok(h))

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


_plans = {}
_plans_lock = threading.Lock()


def halo_plan(maps: ExchangeMaps, kind, device):
    """Cached device plan of `kind` for `device` (a torch.device or index)."""
    dev = torch.device(device)
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    key = (id(maps), kind, idx)
    with _plans_lock:
        p = _plans.get(key)
        if p is None or p.maps is not maps:
            p = _plans[key] = HaloPlan(maps, kind, idx)
        return p


def _geom(a, L):
    """mj_halo_geom of a contiguous [T, ..., ny, nx] or [B, T, ..., ny, nx] tensor whose tile axis is `tile_axis`."""
    plane = L.ny * L.nx
    inner = int(np.prod(a.shape[1:-2])) if a.dim() > 3 else 1
    return _lib.HaloGeom(1, a.numel(), inner * plane, inner, plane)


def _require(a, L, what):
    if not isinstance(a, torch.Tensor) or a.device.type != "cuda":
        raise TypeError(f"{what}: expected a CUDA tensor (the exchange runs on the GPU only)")
    if a.dtype != torch.float64:
        raise TypeError(f"{what}: expected float64, got {a.dtype}")
    if a.dim() < 3 or a.shape[0] != L.nTiles or tuple(a.shape[-2:]) != (L.ny, L.nx):
        raise ValueError(f"{what}: expected [{L.nTiles}, ..., {L.ny}, {L.nx}], got {tuple(a.shape)}")
    return a.contiguous()


class Exchanger:
    """Halo exchanges on [tile, ..., j, i] CUDA float64 tensors (reference parallel/exchange.py:145-217)."""

    def __init__(self, maps: ExchangeMaps = None):
        self.maps = maps or ExchangeMaps.load()
        self.L = self.maps.layout

    def plan(self, kind, device):
        return halo_plan(self.maps, kind, device)

    def scalar(self, a, kind="T", out=None):
        a = _require(a, self.L, f"scalar({kind})")
        p = self.plan(kind, a.device)
        if p.vector:
            raise ValueError(f"{kind} is a vector kind")
        if out is None:
            out = a.clone()
        elif out is not a:
            # the exchange writes only the halo points it fills: the rest of out must hold a's values
            if not isinstance(out, torch.Tensor) or out.device != a.device or out.dtype != a.dtype or \
                    tuple(out.shape) != tuple(a.shape) or not out.is_contiguous():
                raise ValueError(f"scalar({kind}): out must be a contiguous tensor like a")
            out.copy_(a)
        elif not p.inplace_ok:
            raise ValueError(f"scalar({kind}): in place (out=a) is not exact for this kind")
        g = _geom(a, self.L)
        _lib.check(_lib.lib().mj_halo_exchange(p.handle, _lib.ptr(a), _lib.ptr(out), ctypes.byref(g),
                                               _lib.stream_ptr()), "mj_halo_exchange")
        return out

    def vector(self, u, v, kind="UVs"):
        u = _require(u, self.L, f"vector({kind}) u")
        v = _require(v, self.L, f"vector({kind}) v")
        if u.shape != v.shape:
            raise ValueError("u and v shapes differ")
        p = self.plan(kind, u.device)
        if not p.vector:
            raise ValueError(f"{kind} is a scalar kind")
        uo, vo = u.clone(), v.clone()
        g = _geom(u, self.L)
        _lib.check(_lib.lib().mj_halo_exchange_uv(p.handle, _lib.ptr(u), _lib.ptr(v), _lib.ptr(uo), _lib.ptr(vo),
                                                  ctypes.byref(g), _lib.stream_ptr()), "mj_halo_exchange_uv")
        return uo, vo

    def scalar_adjoint(self, abar, kind="T"):
        """Transpose of scalar(): the cotangent of the input for the cotangent abar of the exchanged array."""
        abar = _require(abar, self.L, f"scalar_adjoint({kind})")
        p = self.plan(kind, abar.device)
        if p.vector:
            raise ValueError(f"{kind} is a vector kind")
        out = torch.empty_like(abar)
        g = _geom(abar, self.L)
        _lib.check(_lib.lib().mj_halo_exchange_adjoint(p.handle, _lib.ptr(abar), None, _lib.ptr(out), None,
                                                       ctypes.byref(g), _lib.stream_ptr()), "mj_halo_exchange_adjoint")
        return out

    def vector_adjoint(self, ubar, vbar, kind="UVs"):
        """Transpose of vector(): the cotangents of (u, v) for the cotangents of the exchanged pair."""
        ubar = _require(ubar, self.L, f"vector_adjoint({kind}) u")
        vbar = _require(vbar, self.L, f"vector_adjoint({kind}) v")
        p = self.plan(kind, ubar.device)
        if not p.vector:
            raise ValueError(f"{kind} is a scalar kind")
        uo, vo = torch.empty_like(ubar), torch.empty_like(vbar)
        g = _geom(ubar, self.L)
        _lib.check(_lib.lib().mj_halo_exchange_adjoint(p.handle, _lib.ptr(ubar), _lib.ptr(vbar), _lib.ptr(uo),
                                                       _lib.ptr(vo), ctypes.byref(g), _lib.stream_ptr()),
                   "mj_halo_exchange_adjoint")
        return uo, vo

    def scalar_ad(self, a, kind="T"):
        """scalar() differentiable through torch.autograd (backward: scalar_adjoint)."""
        return _ScalarExchange.apply(a, self, kind)

    def vector_ad(self, u, v, kind="UVs"):
        """vector() differentiable through torch.autograd (backward: vector_adjoint)."""
        return _VectorExchange.apply(u, v, self, kind)

    def fill_(self, a, kind="T"):
        """In-place halo fill (only for kinds whose written points are never sources: T, 3D, UVs, UVn, UV3s)."""
        a = _require(a, self.L, f"fill_({kind})")
        p = self.plan(kind, a.device)
        g = _geom(a, self.L)
        _lib.check(_lib.lib().mj_halo_exchange(p.handle, _lib.ptr(a), _lib.ptr(a), ctypes.byref(g),
                                               _lib.stream_ptr()), "mj_halo_exchange")
        return a

    # Fortran names (parallel/exchange.py:181-195)
    def exch_xy(self, a):  # EXCH_XY_RL / EXCH_XYZ_RL / EXCH_3D_RL
        return self.scalar(a, "T")

    def exch_uv_xy(self, u, v, with_signs=True):  # EXCH_UV_XY_RL / EXCH_UV_XYZ_RL
        return self.vector(u, v, "UVs" if with_signs else "UVn")

    def exch_z(self, a):  # EXCH_Z_3D_RL
        return self.scalar(a, "Z")

    def exch_uv_agrid(self, u, v, with_signs=True):
        return self.vector(u, v, "As" if with_signs else "An")

    def exch_uv_bgrid(self, u, v, with_signs=True):
        return self.vector(u, v, "Bs" if with_signs else "Bn")

    # global reductions over tiles (parallel/exchange.py:199-209): one device holds every tile
    def all_tiles(self, per_tile):
        return per_tile

    def global_sum_tile(self, phiTile):
        """GLOBAL_SUM_TILE_RL: 0 + tile 1 + tile 2 + ... in tile order (parallel/global_sum.py:20-26)."""
        s = torch.zeros(phiTile.shape[1:], dtype=phiTile.dtype, device=phiTile.device)
        for t in range(phiTile.shape[0]):
            s = s + phiTile[t]
        return s

    def global_max(self, a):
        return torch.max(a)

    def vary(self, x):
        return x

    def zero_padding(self, a):
        return a


_default = {}


class _ScalarExchange(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, ex, kind):
        ctx.ex, ctx.kind = ex, kind
        return ex.scalar(a, kind)

    @staticmethod
    def backward(ctx, g):
        return ctx.ex.scalar_adjoint(g.contiguous(), ctx.kind), None, None


class _VectorExchange(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, v, ex, kind):
        ctx.ex, ctx.kind = ex, kind
        return ex.vector(u, v, kind)

    @staticmethod
    def backward(ctx, gu, gv):
        if gu is None:
            gu = torch.zeros_like(gv)
        if gv is None:
            gv = torch.zeros_like(gu)
        ub, vb = ctx.ex.vector_adjoint(gu.contiguous(), gv.contiguous(), ctx.kind)
        return ub, vb, None, None


def default_exchanger(layout=None):
    """The LLC90 13-tile exchanger (reference exchange.py:220-222)."""
    key = layout
    if key not in _default:
        ex = Exchanger(ExchangeMaps.load())
        if layout is not None and layout != ex.L:
            raise ValueError(f"no exch2 maps for {layout}")
        _default[key] = ex
    return _default[key]
