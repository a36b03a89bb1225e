"""Mixed-precision CG2D: FP32 PCG with FP64 residual replacement (libmjamd mj_cg2d_solve_mixed).

`cg2d_mixed(ex, ops, cg2d_b, cg2d_x, *, cg2dNorm, tolerance, max_iters, ...) -> (x, diag)` solves the system of
`cg2d.cg2d_fortran` (reference core/cg2d.py:205-269) with the same normalisation, operator, preconditioner and
stopping rule `err < tolerance**2` on the normalised residual, but stops on the FP64 *true* residual
`b - A x` instead of the recursive one. The PCG runs in FP32; whenever its recursive residual has fallen by
`inner_tolerance` since the last replacement (or below the target, or after `inner_max_iters` iterations) the FP32
solution increment goes into the FP64 solution, the true residual is formed in FP64 and replaces the recursive one,
and the PCG goes on from the same search direction (the Krylov space is kept: the iteration count is that of the
FP64 PCG). FP64: true residuals, dot products, alpha/beta, the solution update. FP32: the operator/preconditioner
copies, inner vectors and stencil arithmetic. Arguments, batching, operator slots and the returned x
(un-normalised interior, Fortran halos) are as `cg2d.cg2d_fortran`; diag numIters counts inner iterations and
outerIters the replacements (the first residual included).
"""

import ctypes

import torch

from . import _lib
from .cg2d import Cg2dDiag, _as_batch, _norm_tensor, _ops_struct, config

INNER_TOLERANCE = 1e-4
INNER_MAX_ITERS = 400
MAX_OUTER = 40


def cg2d_mixed(ex, ops, cg2d_b, cg2d_x, *, cg2dNorm, tolerance, max_iters, out=None, op_index=None,
               inner_tolerance=INNER_TOLERANCE, inner_max_iters=INNER_MAX_ITERS, max_outer=MAX_OUTER,
               launch="phases", poll_iters=0, block_threads=0):
    L = ex.L
    batched = isinstance(cg2d_b, torch.Tensor) and cg2d_b.dim() == 4
    b = _as_batch(cg2d_b, L, "cg2d_b")
    x0 = _as_batch(cg2d_x, L, "cg2d_x")
    B = b.shape[0]
    if x0.shape[0] != B:
        raise ValueError("cg2d_b and cg2d_x batch sizes differ")
    if launch not in ("phases", "phases_fixed", "auto"):
        raise ValueError("the mixed solve runs as phase launches: launch must be 'phases' or 'phases_fixed'")
    dev = b.device
    ops_s, ops_keep = _ops_struct(ops, B, L, dev, op_index)
    nslots = ops_keep[0].shape[0] if ops_keep[0].dim() == 4 else 1
    norm = _norm_tensor(cg2dNorm, B, dev)
    cfg = config(L, B, tolerance, max_iters, "tile", block_threads, 0, launch, poll_iters)
    mcfg = _lib.Cg2dMixedConfig(int(nslots), float(inner_tolerance), int(inner_max_iters), int(max_outer))
    lib = _lib.lib()
    wsb = lib.mj_cg2d_mixed_workspace_bytes(ctypes.byref(cfg), ctypes.byref(mcfg), ops_s.member_stride)
    if wsb == 0:
        raise _lib.MjError(f"mj_cg2d_mixed_workspace_bytes: {_lib.last_error()}")
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    x = torch.empty_like(b) if out is None else out
    diag = torch.empty(B * 48, dtype=torch.uint8, device=dev)
    plan = ex.plan("T", dev)
    _lib.check(lib.mj_cg2d_solve_mixed(plan.handle, ctypes.byref(ops_s), _lib.ptr(b), _lib.ptr(x0), _lib.ptr(x),
                                       _lib.ptr(norm), ctypes.byref(cfg), ctypes.byref(mcfg), _lib.ptr(ws),
                                       ctypes.c_size_t(wsb), _lib.ptr(diag), _lib.stream_ptr()),
               "mj_cg2d_solve_mixed")
    del ops_keep, ws
    return (x if batched else x[0]), Cg2dDiag(diag, batched)
