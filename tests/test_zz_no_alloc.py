"""GPU: mj_cg2d_solve_noalloc (the solve with the library's own per-device scratch, no allocation) against
mj_cg2d_solve (the same solve with a caller workspace), through the C ABI.

Two calls in a row, the second with a smaller input (6 members with the persistent launch, then 3 members with
per-phase launches), each bitwise equal to mj_cg2d_solve on the same inputs (solution and diagnostics). A plan
whose CG tables are not built yet, and a batch whose workspace exceeds the scratch, are rejected before anything
is launched.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve_pair(ex, plan, ops, bs, launch):
    """(x, diag bytes) from mj_cg2d_solve and from mj_cg2d_solve_noalloc on the same inputs."""
    import torch
    from mit_jax_amd import _lib, cg2d
    dev = torch.device("cuda:0")
    L = ex.L
    B = bs.shape[0]
    b = torch.as_tensor(bs, device=dev)
    x0 = torch.zeros_like(b)
    ops_s, keep = cg2d._ops_struct(ops[1], B, L, dev)
    norm = torch.full((B,), ops[2], dtype=torch.float64, device=dev)
    cfg = cg2d.config(L, B, 1e-12, 1000, "tile", 0, 0, launch, 0)
    lib = _lib.lib()
    wsb = lib.mj_cg2d_workspace_bytes(ctypes.byref(cfg))
    assert 0 < wsb <= (64 << 20)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    out = []
    for noalloc in (False, True):
        x = torch.full_like(b, np.nan)
        diag = torch.zeros(B * 48, dtype=torch.uint8, device=dev)
        print(f"no-alloc test: B={B} launch={launch} {'mj_cg2d_solve_noalloc' if noalloc else 'mj_cg2d_solve'}",
              flush=True)
        if noalloc:
            rc = lib.mj_cg2d_solve_noalloc(plan.handle, ctypes.byref(ops_s), _lib.ptr(b), _lib.ptr(x0), _lib.ptr(x),
                                           _lib.ptr(norm), ctypes.byref(cfg), _lib.ptr(diag), _lib.stream_ptr())
        else:
            rc = lib.mj_cg2d_solve(plan.handle, ctypes.byref(ops_s), _lib.ptr(b), _lib.ptr(x0), _lib.ptr(x),
                                   _lib.ptr(norm), ctypes.byref(cfg), _lib.ptr(ws), ctypes.c_size_t(wsb),
                                   _lib.ptr(diag), _lib.stream_ptr())
        _lib.check(rc, "solve")
        torch.cuda.synchronize()
        out.append((x.cpu().numpy(), diag.cpu().numpy()))
    del keep, ws
    return out


def test_noalloc_equals_workspace_solve_twice():
    import torch
    from mit_jax_amd import _lib, cg2d, synthetic as S
    from mit_jax_amd.exchange import HaloPlan, default_exchanger
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    dev = torch.device("cuda:0")
    ex = default_exchanger()
    case = S.cg2d_case(0, "5point")
    g = case[0]
    plan = HaloPlan(ex.maps, "T", 0)  # a plan of its own: its CG tables are not built yet
    lib = _lib.lib()

    # rejected before any launch: the plan has no CG tables yet
    bs = np.stack([S.rhs(g, 100 + m) for m in range(6)])
    b = torch.as_tensor(bs, device=dev)
    x = torch.empty_like(b)
    ops_s, keep = cg2d._ops_struct(case[1], 6, ex.L, dev)
    norm = torch.full((6,), case[2], dtype=torch.float64, device=dev)
    diag = torch.zeros(6 * 48, dtype=torch.uint8, device=dev)
    cfg = cg2d.config(ex.L, 6, 1e-12, 1000, "tile")
    rc = lib.mj_cg2d_solve_noalloc(plan.handle, ctypes.byref(ops_s), _lib.ptr(b), _lib.ptr(torch.zeros_like(b)),
                                   _lib.ptr(x), _lib.ptr(norm), ctypes.byref(cfg), _lib.ptr(diag), _lib.stream_ptr())
    assert rc == -1 and "CG tables" in _lib.last_error()
    # rejected: 12 members need more than the scratch holds
    cfg12 = cg2d.config(ex.L, 12, 1e-12, 1000, "tile")
    assert lib.mj_cg2d_workspace_bytes(ctypes.byref(cfg12)) > (64 << 20)
    rc = lib.mj_cg2d_solve_noalloc(plan.handle, ctypes.byref(ops_s), _lib.ptr(b), _lib.ptr(b), _lib.ptr(x),
                                   _lib.ptr(norm), ctypes.byref(cfg12), _lib.ptr(diag), _lib.stream_ptr())
    assert rc == -1 and "scratch" in _lib.last_error()
    del keep

    # first call: 6 members, persistent launch (the workspace solve builds the plan's tables first)
    (xr, dr), (xn, dn) = _solve_pair(ex, plan, case, bs, "persistent")
    assert np.array_equal(xr, xn), "6 members: solutions differ"
    assert np.array_equal(dr, dn), "6 members: diagnostics differ"
    it = dr.view(np.int32).reshape(6, 12)[:, 8]
    assert (dr.view(np.int32).reshape(6, 12)[:, 9] == 1).all() and (it > 50).all()
    # second call: 3 other members, per-phase launches (the scratch holds the first call's contents)
    bs2 = np.stack([S.rhs(g, 200 + m) for m in range(3)])
    (xr, dr), (xn, dn) = _solve_pair(ex, plan, case, bs2, "phases")
    assert np.array_equal(xr, xn), "3 members: solutions differ"
    assert np.array_equal(dr, dn), "3 members: diagnostics differ"
    assert (dr.view(np.int32).reshape(3, 12)[:, 9] == 1).all()
    print(f"no-alloc test: bitwise equal, iterations {it.tolist()}", flush=True)
    plan.close()
