"""CPU oracle of the CG2D / GAD DST3 / exch2 halo hot path — TEST INFRASTRUCTURE ONLY.

A restatement of the reference algorithms (mitgcm_jax; citations per function in mjo.c and tables.py): plain C
FP64 compiled without fused multiply-adds (mjo.c, built by the Makefile into _build/libmjo.so) for the arithmetic,
numpy for the per-tile tables. Imported only by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py, always as the checker and never as the thing measured or shipped. The product (mit_jax_amd) does not
import it and has no CPU fallback.
"""

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

from . import tables  # noqa: F401

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libmjo.so"

_lib = None


def build():
    """Compile mjo.c into _build/libmjo.so (gcc -O2 -ffp-contract=off -fopenmp)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "-B", "_build/libmjo.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle library {LIB_PATH} is not built (run oracle.build())")
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        L.mjo_exchange.argtypes = [P, P, P, I, I, I, P, P, P, P]
        L.mjo_tree_n.argtypes = [P, I]
        L.mjo_tree_n.restype = D
        L.mjo_global_sum.argtypes = [P, I, I, I, I]
        L.mjo_global_sum.restype = D
        L.mjo_cg2d.argtypes = [P, P, P, I, I, I, I] + [P] * 9 + [D, D, I, I, P]
        L.mjo_cg2d.restype = I
        L.mjo_cg2d_matvec.argtypes = [P, P, P, I, I, I, I, P, P, P, P, P, D]
        L.mjo_advect_tiles.argtypes = [I, I, I, I, P, D] + [P] * 8 + [P, P, P, P] + [P] * 6 + [P, I]
        L.mjo_dst3_adv.argtypes = [I, I, I, I, I, D] + [P] * 8
        L.mjo_vert_dst3.argtypes = [I, I, I, I, P, P, D, P, P, P, P, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------------------------------------------------ maps
def full_map(src, comp, sign):
    """(src int32, comp int8, sign int8) over all padded points, contiguous."""
    return (np.ascontiguousarray(src, np.int32), np.ascontiguousarray(comp, np.int8),
            np.ascontiguousarray(sign, np.int8))


def exchange(m, own, fa, fb=None, T=13, plane=98 * 98):
    """Exchanger._apply on [T, nlead, ny, nx] (or [T, ny, nx]) arrays; returns a new array."""
    src, comp, sign = full_map(*m)
    own = _f64(own)
    fa = _f64(fa)
    fb = fa if fb is None else _f64(fb)
    nlead = own.size // (T * plane)
    out = np.empty_like(own)
    lib().mjo_exchange(_p(src), _p(comp), _p(sign), T, plane, nlead, _p(own), _p(fa), _p(fb), _p(out))
    return out


def scalar(m, a, T=13, plane=98 * 98):
    return exchange(m, a, a, None, T, plane)


def vector(mu, mv, u, v, T=13, plane=98 * 98):
    return exchange(mu, u, u, v, T, plane), exchange(mv, v, u, v, T, plane)


# ------------------------------------------------------------------------------------------------------ sums
def tree_n(v):
    v = _f64(v)
    return lib().mjo_tree_n(_p(v), v.size)


def global_sum(a_interior, order):
    a = _f64(a_interior)
    T, sNy, sNx = a.shape
    return lib().mjo_global_sum(_p(a), T, sNy, sNx, 0 if order == "fortran" else 1)


# ------------------------------------------------------------------------------------------------------ CG2D
def cg2d(Tmap, ops, b, x0, *, cg2dNorm, tolerance, max_iters, sum_order="fortran", sNx=90, sNy=90, OL=4):
    """cg2d_fortran (core/cg2d.py:205-269). Returns (x, diag dict)."""
    src, comp, sign = full_map(*Tmap)
    ops = [_f64(a) for a in ops]
    b = _f64(b)
    x0 = _f64(x0)
    T = b.shape[0]
    x = np.empty_like(b)
    diag = np.zeros(5)
    lib().mjo_cg2d(_p(src), _p(comp), _p(sign), T, sNx, sNy, OL, *[_p(a) for a in ops], _p(b), _p(x0), _p(x),
                   float(cg2dNorm), float(tolerance), int(max_iters), 0 if sum_order == "fortran" else 1, _p(diag))
    return x, dict(firstResidual=diag[0], lastResidual=diag[1], sumRHS=diag[2], rhsMax=diag[3], numIters=int(diag[4]))


def cg2d_matvec(Tmap, A, v, cg2dNorm, sNx=90, sNy=90, OL=4):
    src, comp, sign = full_map(*Tmap)
    aW, aS, aC = (_f64(a) for a in A)
    v = _f64(v)
    y = np.empty_like(v)
    lib().mjo_cg2d_matvec(_p(src), _p(comp), _p(sign), v.shape[0], sNx, sNy, OL, _p(aW), _p(aS), _p(aC), _p(v),
                          _p(y), float(cg2dNorm))
    return y


def cg2d_coeff_grad(Tmap, x, lam, cg2dNorm, OL=4):
    """Cotangent of (aW2d, aS2d, aC2d) in the implicit derivative of the solve (core/cg2d.py:326-348): the
    transpose of dA -> -_matvec(dA, x) applied to lam (zero off the interior), xe = exch(where(inner, x, 0)):
    gaW[q] = -(lam[q] xe[q-1] + lam[q-1] xe[q]) / cg2dNorm; gaS with q-nx; gaC[q] = -(lam[q] xe[q]) / cg2dNorm.
    numpy, [T, ny, nx]."""
    x = _f64(x)
    T, ny, nx = x.shape
    inner = np.zeros((ny, nx), bool)
    inner[OL:ny - OL, OL:nx - OL] = True
    xe = scalar(Tmap, np.where(inner[None], x, 0.0), T=T, plane=ny * nx)
    lam = _f64(lam)
    gW = np.zeros_like(x)
    gS = np.zeros_like(x)
    gW[:, :, 1:] = lam[:, :, 1:] * xe[:, :, :-1] + lam[:, :, :-1] * xe[:, :, 1:]
    gS[:, 1:, :] = lam[:, 1:, :] * xe[:, :-1, :] + lam[:, :-1, :] * xe[:, 1:, :]
    return -(gW / cg2dNorm), -(gS / cg2dNorm), -(lam * xe / cg2dNorm)


# ------------------------------------------------------------------------------------------------------ DST3
VERT_ORDER = ("dTlev", "deepFacC", "recip_deepFacC", "recip_deepFac2C", "recip_rhoFacC", "drF", "recip_drF",
              "rhoFacC")


def advect_tiles(tab, grid2d, vgrid, uFld, vFld, tracer, hFacW, hFacS, recip_hFacC, oneSixth=1.0 / 6.0,
                 want_local=False):
    """advect_tiles (pkgs/gad.py:460-539) with oracle tables `tab` (tables.tables()). grid2d: dyG, dxG,
    recip_dxC, recip_dyC, recip_rA, maskInC [T, ny, nx]; maskW, maskS [T, K, ny, nx]. vgrid: the VERT_ORDER
    per-level arrays. Returns gTracer (or localTij after the passes)."""
    tracer = _f64(tracer)
    T, K, ny, nx = tracer.shape
    vert = np.concatenate([_f64(vgrid[k]).reshape(K) for k in VERT_ORDER])
    g = {k: _f64(grid2d[k]) for k in ("dyG", "dxG", "recip_dxC", "recip_dyC", "recip_rA", "maskInC", "maskW", "maskS")}
    masks = np.ascontiguousarray(tab["masks"], np.uint8)
    fills = [None if f is None else np.ascontiguousarray(f, np.int32) for f in tab["fills"]]
    farr = (ctypes.c_void_p * 12)(*[None if f is None else f.ctypes.data for f in fills])
    uvu = np.ascontiguousarray(tab["uvfill_u"], np.int32)
    uvv = np.ascontiguousarray(tab["uvfill_v"], np.int32)
    fields = [_f64(a) for a in (uFld, vFld, tracer, hFacW, hFacS, recip_hFacC)]
    out = np.empty_like(tracer)
    lib().mjo_advect_tiles(T, K, ny, nx, _p(vert), float(oneSixth), *[_p(g[k]) for k in
                           ("dyG", "dxG", "recip_dxC", "recip_dyC", "recip_rA", "maskInC", "maskW", "maskS")],
                           _p(masks), ctypes.cast(farr, ctypes.c_void_p), _p(uvu), _p(uvv),
                           *[_p(a) for a in fields], _p(out), int(bool(want_local)))
    return out


def dst3_adv(direction, trans, vel, mask, tracer, rd, deltaTloc, rdeep, oneSixth=1.0 / 6.0):
    """dst3_adv_x (direction "x") / dst3_adv_y ("y") (pkgs/gad.py:409-453) over [T, K, ny, nx] arrays; rd = recip_dxC /
    recip_dyC [T, ny, nx]; deltaTloc, rdeep = recip_deepFacC [K]. Returns the flux array."""
    tracer = _f64(tracer)
    T, K, ny, nx = tracer.shape
    args = [_f64(a) for a in (trans, vel, mask)] + [tracer, _f64(rd)]
    dt, rdp = _f64(np.broadcast_to(deltaTloc, (K,))), _f64(np.broadcast_to(rdeep, (K,)))
    out = np.empty_like(tracer)
    lib().mjo_dst3_adv(0 if direction == "x" else 1, T, K, ny, nx, float(oneSixth), _p(dt), *[_p(a) for a in args],
                       _p(rdp), _p(out))
    return out


def gad_dst3(tab, grid2d, vgrid, uFld, vFld, wFld, tracer, hFacW, hFacS, recip_hFacC, oneSixth=1.0 / 6.0):
    """Horizontal passes, then (wFld given) the vertical DST3 sweep. grid2d additionally needs rA and maskC
    [T, K, ny, nx]; vgrid recip_drC [K]."""
    if wFld is None:
        return advect_tiles(tab, grid2d, vgrid, uFld, vFld, tracer, hFacW, hFacS, recip_hFacC, oneSixth)
    loc = advect_tiles(tab, grid2d, vgrid, uFld, vFld, tracer, hFacW, hFacS, recip_hFacC, oneSixth, want_local=True)
    tracer = _f64(tracer)
    T, K, ny, nx = tracer.shape
    vert = np.concatenate([_f64(vgrid[k]).reshape(K) for k in VERT_ORDER])
    out = np.empty_like(tracer)
    lib().mjo_vert_dst3(T, K, ny, nx, _p(vert), _p(_f64(vgrid["recip_drC"])), float(oneSixth),
                        _p(_f64(grid2d["rA"])), _p(_f64(grid2d["recip_rA"])), _p(_f64(grid2d["maskC"])),
                        _p(_f64(wFld)), _p(tracer), _p(_f64(recip_hFacC)), _p(loc), _p(out))
    return out


def threads():
    """OpenMP threads the oracle uses (OMP_NUM_THREADS, else all cores)."""
    v = os.environ.get("OMP_NUM_THREADS")
    return int(v) if v else os.cpu_count()
