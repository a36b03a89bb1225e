"""CG2D free-surface pressure solve on the GPU, with the reference's interface (mitgcm_jax core/cg2d.py).

`cg2d_fortran(ex, ops, cg2d_b, cg2d_x, *, cg2dNorm, tolerance, max_iters, sum_order, nIterMin=-1,
sum_unroll=1) -> (x, diag)` (reference core/cg2d.py:205-269) runs the literal preconditioned CG of cg2d.F through
libmjamd's batched HIP solve (mj_cg2d_solve): same normalisation, preconditioner, update order and stopping
rule `err_sq < tolerance**2`; x leaves with the Fortran halos (exchanged normalised first guess). Arrays are CUDA
float64 [T, ny, nx], or [B, T, ny, nx] for a batch of members solved in the same launch (per-member convergence;
ops shared [T, ny, nx] or per member [B, T, ny, nx]; cg2dNorm a scalar or B values).

sum_order is required, as in the reference (Cg2dParams defaults it to "fortran"): "fortran" sums each tile
sequentially in j/i order, bitwise the reference's order; "tile" uses the library's fixed wavefront tree (the
reference's "tile" order is XLA's own tree, so the two agree to rounding, not bitwise). sum_unroll only changes the reference's XLA scheduling and is accepted and
ignored.

`cg2d_solve(p, ex, ops, cg2d_b, cg2d_x) -> (x, diag)` (core/cg2d.py:351-375) returns x with 0 halos and is
differentiable (torch.autograd): the backward pass is the transposed solve with the same symmetric operator
(`adjoint.py`).
"""

import ctypes
from dataclasses import dataclass, replace

import numpy as np
import torch

from . import _lib
from .layout import LLC90, interior_mask

SUM_ORDERS = {"fortran": _lib.MJ_SUM_FORTRAN, "tile": _lib.MJ_SUM_TREE}


@dataclass(frozen=True)
class Cg2dParams:
    """CG2D parameters (reference core/cg2d.py:49-93)."""
    cg2dMaxIters: int
    cg2dTolerance: float
    cg2dNorm: float
    cg2dpcOffDFac: float = 0.51
    nIterMin: int = -1
    sum_order: str = "fortran"
    sum_unroll: int = 1
    adj_tolerance: float = 1e-13
    adj_max_iters: int = 2000
    stop_coeff_grad: bool = False


class Cg2dDiag(dict):
    """Per-member diagnostics of a solve (tensors on the device; the reference returns firstResidual,
    lastResidual, numIters, sumRHS, rhsMax, core/cg2d.py:267-268), plus `converged` (1 / 0; -2 when a device
    barrier timed out), `outerIters` (FP64 residual replacements of the mixed-precision solve) and `launches` (kernel launches
    of the solve's iteration loop in the per-phase launch modes, the same in every member's record)."""

    def __init__(self, raw, batched):
        f = raw.view(torch.float64).view(-1, 6)
        i = raw.view(torch.int32).view(-1, 12)
        sel = (lambda t: t) if batched else (lambda t: t[0])
        super().__init__(firstResidual=sel(f[:, 0]), lastResidual=sel(f[:, 1]), sumRHS=sel(f[:, 2]),
                         rhsMax=sel(f[:, 3]), numIters=sel(i[:, 8]), converged=sel(i[:, 9]),
                         outerIters=sel(i[:, 10]), launches=sel(i[:, 11]))
        self.raw = raw

    def check(self):
        """Raise if a device barrier timed out (the launch was abandoned; results are invalid)."""
        if bool((self["converged"] == -2).any()):
            raise _lib.MjError("cg2d: a device grid barrier timed out")
        return self


def _as_batch(a, L, name):
    if not isinstance(a, torch.Tensor) or a.device.type != "cuda" or a.dtype != torch.float64:
        raise TypeError(f"{name}: expected a CUDA float64 tensor")
    if a.dim() == 3:
        a = a.unsqueeze(0)
    if a.dim() != 4 or tuple(a.shape[1:]) != L.shape2d:
        raise ValueError(f"{name}: expected [{L.nTiles}, {L.ny}, {L.nx}] or [B, ...], got {tuple(a.shape)}")
    return a.contiguous()


def _ops_struct(ops, B, L, device, op_index=None):
    """The six operator arrays as one slot [T, ny, nx] shared by every member, one slot per member
    [B, T, ny, nx], or C slots [C, T, ny, nx] with op_index (B slot numbers: members sharing an operator, e.g.
    the same level count, read the same coefficients)."""
    ops = [o if isinstance(o, torch.Tensor) else torch.as_tensor(np.asarray(o), device=device) for o in ops]
    ops = [o.to(device=device, dtype=torch.float64).contiguous() for o in ops]
    shp = {tuple(o.shape) for o in ops}
    if len(shp) != 1:
        raise ValueError("the six operator arrays must have the same shape")
    shp = shp.pop()
    idx = None
    if shp == L.shape2d:
        ms = 0
        if op_index is not None:
            raise ValueError("op_index given with a single shared operator")
    elif len(shp) == 4 and shp[1:] == L.shape2d:
        ms = L.npad
        C = shp[0]
        if op_index is None:
            if C != B:
                raise ValueError(f"{C} operator slots for {B} members need op_index")
        else:
            idx = torch.as_tensor(np.asarray(op_index, np.int32).reshape(-1), device=device) \
                if not isinstance(op_index, torch.Tensor) else op_index.to(device=device, dtype=torch.int32)
            idx = idx.contiguous()
            if idx.numel() != B or int(idx.min()) < 0 or int(idx.max()) >= C:
                raise ValueError(f"op_index: {B} slot numbers in 0..{C - 1} expected")
    else:
        raise ValueError(f"operator shape {shp} is neither {L.shape2d} nor [C, {L.nTiles}, {L.ny}, {L.nx}]")
    s = _lib.Cg2dOps(*[o.data_ptr() for o in ops], ms, idx.data_ptr() if idx is not None else None)
    return s, ops + ([idx] if idx is not None else [])


def _norm_tensor(cg2dNorm, B, device):
    if isinstance(cg2dNorm, torch.Tensor):
        n = cg2dNorm.to(device=device, dtype=torch.float64).reshape(-1)
    else:
        n = torch.as_tensor(np.asarray(cg2dNorm, dtype=np.float64).reshape(-1), device=device)
    if n.numel() == 1 and B > 1:
        n = n.expand(B)
    if n.numel() != B:
        raise ValueError(f"cg2dNorm: {n.numel()} values for {B} members")
    return n.contiguous()


LAUNCH_MODES = {"auto": 0, "persistent": 1, "phases": 2, "phases_fixed": 3}


def config(L, B, tolerance, max_iters, sum_order, block_threads=0, grid_blocks=0, launch="auto", poll_iters=0):
    if sum_order not in SUM_ORDERS:
        raise ValueError(sum_order)
    if launch not in LAUNCH_MODES:
        raise ValueError(f"launch must be one of {sorted(LAUNCH_MODES)}")
    return _lib.Cg2dConfig(B, L.nTiles, L.sNx, L.sNy, L.OLx, L.npad, float(tolerance), int(max_iters),
                           SUM_ORDERS[sum_order], int(block_threads), int(grid_blocks), LAUNCH_MODES[launch],
                           int(poll_iters))


def cg2d_fortran(ex, ops, cg2d_b, cg2d_x, *, cg2dNorm, tolerance, max_iters, sum_order, nIterMin=-1,
                 sum_unroll=1, block_threads=0, grid_blocks=0, out=None, op_index=None, launch="auto", poll_iters=0):
    """CG2D (reference core/cg2d.py:205-269). ops = (aW2d, aS2d, aC2d, pW, pS, pC). Returns (x, diag).
    Batched: cg2d_b, cg2d_x [B, T, ny, nx]; ops shared, per member, or per slot with op_index. launch picks the
    launch shape ("auto", "persistent", "phases", "phases_fixed"); every shape gives the same bits."""
    if nIterMin >= 0:
        raise NotImplementedError("nIterMin >= 0 (min-residual solution) is not ported")
    L = ex.L
    if L.OLx != L.OLy:
        raise ValueError("OLx != OLy")
    batched = isinstance(cg2d_b, torch.Tensor) and cg2d_b.dim() == 4
    b = _as_batch(cg2d_b, L, "cg2d_b")
    x0 = _as_batch(cg2d_x, L, "cg2d_x")
    B = b.shape[0]
    if x0.shape[0] != B:
        raise ValueError("cg2d_b and cg2d_x batch sizes differ")
    dev = b.device
    ops_s, ops_keep = _ops_struct(ops, B, L, dev, op_index)
    norm = _norm_tensor(cg2dNorm, B, dev)
    cfg = config(L, B, tolerance, max_iters, sum_order, block_threads, grid_blocks, launch, poll_iters)
    lib = _lib.lib()
    wsb = lib.mj_cg2d_workspace_bytes(ctypes.byref(cfg))
    if wsb == 0:
        raise _lib.MjError(f"mj_cg2d_workspace_bytes: {_lib.last_error()}")
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    x = torch.empty_like(b) if out is None else out
    diag = torch.empty(B * 48, dtype=torch.uint8, device=dev)
    plan = ex.plan("T", dev)
    _lib.check(lib.mj_cg2d_solve(plan.handle, ctypes.byref(ops_s), _lib.ptr(b), _lib.ptr(x0), _lib.ptr(x),
                                 _lib.ptr(norm), ctypes.byref(cfg), _lib.ptr(ws), ctypes.c_size_t(wsb),
                                 _lib.ptr(diag), _lib.stream_ptr()), "mj_cg2d_solve")
    del ops_keep, ws
    return (x if batched else x[0]), Cg2dDiag(diag, batched)


def matvec(ex, A, v, cg2dNorm, op_index=None):
    """(A / cg2dNorm) v on the interior, 0 on halos, v's halos zeroed then exchanged (reference _matvec,
    core/cg2d.py:292-298). A = (aW2d, aS2d, aC2d). v [T, ny, nx] or [B, T, ny, nx]."""
    L = ex.L
    batched = v.dim() == 4
    vb = _as_batch(v, L, "v")
    B = vb.shape[0]
    dev = vb.device
    A = tuple(A)
    ops_s, keep = _ops_struct(A + A, B, L, dev, op_index)  # the preconditioner slots are not read
    norm = _norm_tensor(cg2dNorm, B, dev)
    cfg = config(L, B, 0.0, 0, "tile")
    y = torch.empty_like(vb)
    work = torch.empty_like(vb)
    _lib.check(_lib.lib().mj_cg2d_matvec(ex.plan("T", dev).handle, ctypes.byref(ops_s), _lib.ptr(vb), _lib.ptr(y),
                                         _lib.ptr(norm), ctypes.byref(cfg), _lib.ptr(work), _lib.stream_ptr()),
               "mj_cg2d_matvec")
    del keep
    return y if batched else y[0]


def ini_cg2d_norm(g, hFacW, hFacS, implicSurfPress, implicDiv2DFlow):
    """cg2dNorm of INI_CG2D (reference core/cg2d.py:108-134) on the GPU; see solve_for_pressure.ini_cg2d_norm."""
    from .solve_for_pressure import ini_cg2d_norm as _norm
    return _norm(g, hFacW, hFacS, implicSurfPress, implicDiv2DFlow)


def cg2d_solve(p: Cg2dParams, ex, ops, cg2d_b, cg2d_x):
    """Solve A x = b with the literal CG2D forward; differentiable in cg2d_b and aW2d/aS2d/aC2d through the
    implicit derivative (reference core/cg2d.py:351-375). Returns (x with 0 halos, diag)."""
    from .adjoint import cg2d_solve as _solve
    return _solve(p, ex, ops, cg2d_b, cg2d_x)


def interior(L=LLC90, device=None):
    return torch.as_tensor(interior_mask(L), device=device)
