"""Replay of the Fortran oracle's dumped stages through the GPU path: the reference's bitwise gates C01/C02/S07
(reference tests/test_cg2d.py:116-160, tests/test_free_surface.py:78-125) and T10/T20 (tests/test_gad.py:119-130)
run against a jaxdump directory (`dump.DumpIndex`) the moment one is available.

Stage and field names are the reference's: S04_oceanic_phys EmPmR; S05_dynamics gU, gV, etaN, etaH;
S06_update_rstar_T hFacW, hFacS; G00_geometry Bo_surf, recip_Bo; S07_update_cg2d and C01_cg2d_inputs aW2d, aS2d,
aC2d, pW, pS, pC; C01_cg2d_inputs cg2d_b, cg2d_x; C02_cg2d_solution cg2d_x; for GAD_ADVECTION (what it reads inside
THERMODYNAMICS) T01_residual_flow uFld, vFld, wFld, S04_oceanic_phys theta, salt, S11_calc_rstar hFacW, hFacS,
recip_hFacC, the G00_geometry fields and its packed vertical record, and the outputs T10_temp_adv gT_loc,
T20_salt_adv gS_loc. Each function returns the GPU result next to the dumped one (numpy), for the caller to compare
bitwise.
"""

import numpy as np
import torch

OPS = ("aW2d", "aS2d", "aC2d", "pW", "pS", "pC")


def _dev(a, device):
    return torch.as_tensor(np.ascontiguousarray(a, np.float64), device=device)


def replay_cg2d(ds, it, ex, cp, device="cuda:0"):
    """C02: the solve from the dumped C01 inputs. Returns (x, diag, x_dumped)."""
    from . import cg2d
    L = ex.L
    f = lambda st, n: ds.padded(it, st, n, L)  # noqa: E731
    ops = [_dev(f("C01_cg2d_inputs", n), device) for n in OPS]
    x, d = cg2d.cg2d_fortran(ex, ops, _dev(f("C01_cg2d_inputs", "cg2d_b"), device),
                             _dev(f("C01_cg2d_inputs", "cg2d_x"), device), cg2dNorm=cp.cg2dNorm,
                             tolerance=cp.cg2dTolerance, max_iters=cp.cg2dMaxIters, sum_order=cp.sum_order)
    return x.cpu().numpy(), d, f("C02_cg2d_solution", "cg2d_x")


def replay_update_cg2d(ds, it, p, cp, g, ex, pW, pS, pC, device="cuda:0"):
    """S07: UPDATE_CG2D from the dumped r* hFacW/S and recip_Bo; pW, pS, pC: the preconditioner before the call.
    Returns ({name: GPU}, {name: dumped})."""
    from . import solve_for_pressure as sfp
    L = ex.L
    f = lambda st, n: ds.padded(it, st, n, L, squeeze=False)  # noqa: E731
    hW = _dev(f("S06_update_rstar_T", "hFacW"), device)
    hS = _dev(f("S06_update_rstar_T", "hFacS"), device)
    rBo = _dev(ds.padded(it, "G00_geometry", "recip_Bo", L), device)
    out = sfp.update_cg2d(p, cp, g, ex, hW, hS, rBo, _dev(pW, device), _dev(pS, device), _dev(pC, device))
    got = {n: o.cpu().numpy() for n, o in zip(OPS, out)}
    return got, {n: ds.padded(it, "S07_update_cg2d", n, L) for n in OPS}


def replay_cg2d_rhs(ds, it, p, g, ex, device="cuda:0"):
    """C01: cg2d_b and cg2d_x from the dumped dynamics, r* and forcing stages. Returns ((b, x), (b_dumped,
    x_dumped))."""
    from . import solve_for_pressure as sfp
    L = ex.L
    f3 = lambda st, n: _dev(ds.padded(it, st, n, L, squeeze=False), device)  # noqa: E731
    f2 = lambda st, n: _dev(ds.padded(it, st, n, L), device)  # noqa: E731
    b, x = sfp.cg2d_rhs(p, g, f3("S05_dynamics", "gU"), f3("S05_dynamics", "gV"), f3("S06_update_rstar_T", "hFacW"),
                        f3("S06_update_rstar_T", "hFacS"), f2("S04_oceanic_phys", "EmPmR"), f2("S05_dynamics", "etaN"),
                        f2("S05_dynamics", "etaH"), f2("G00_geometry", "Bo_surf"))
    return ((b.cpu().numpy(), x.cpu().numpy()),
            (ds.padded(it, "C01_cg2d_inputs", "cg2d_b", L), ds.padded(it, "C01_cg2d_inputs", "cg2d_x", L)))


GAD_GRID2D = ("dyG", "dxG", "recip_dxC", "recip_dyC", "recip_rA", "rA", "maskInC")
GAD_GRID3D = ("maskW", "maskS", "maskC")
GAD_TRACERS = {"temp": ("theta", "T10_temp_adv", "gT_loc"), "salt": ("salt", "T20_salt_adv", "gS_loc")}


def replay_gad(ds, it, params, device="cuda:0", tracers=("temp", "salt")):
    """T10 / T20: GAD_ADVECTION of theta and salt from the dumped residual flow, r* hFac and geometry (reference
    tests/test_gad.py:119-130 through gad.gad_advection, pkgs/gad.py:546-560), on the whole tile arrays.
    Returns {tracer: (GPU gTracer, dumped)}."""
    from . import gad
    L = params.layout
    f2 = lambda st, n: _dev(ds.padded(it, st, n, L), device)  # noqa: E731
    f3 = lambda st, n: _dev(ds.padded(it, st, n, L, squeeze=False), device)  # noqa: E731
    g = {n: f2("G00_geometry", n) for n in GAD_GRID2D}
    g.update({n: f3("G00_geometry", n) for n in GAD_GRID3D})
    g.update({k: _dev(a, device) for k, a in ds.vertical(it, L).items()})
    u, v, w = (f3("T01_residual_flow", n) for n in ("uFld", "vFld", "wFld"))
    hW, hS, rhC = (f3("S11_calc_rstar", n) for n in ("hFacW", "hFacS", "recip_hFacC"))
    out = {}
    for tr in tracers:
        state, stage, name = GAD_TRACERS[tr]
        got = gad.gad_advection(params, g, u, v, w, f3("S04_oceanic_phys", state), hW, hS, rhC)
        out[tr] = (got.cpu().numpy(), ds.padded(it, stage, name, L, squeeze=False))
    return out
