"""Benchmark: ensemble member-steps/s of the CG2D + DST3 step on LLC90 (13 tiles x 90x90, 50 levels) on one MI355X.

A member-step is one CG2D free-surface solve to the normalised 1e-12 target plus one GAD DST3 tracer step (X/Y flux
passes, corner fills, flux divergence) of one ensemble member. One timed step runs the whole ensemble: one batched
solve (per-member operators through operator slots, per-member convergence) and one batched DST3 step (per-member
level counts). Inputs are seeded synthetic fields resident in HBM before timing starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--members B] [--precision mixed|fp64]
    python bench.py --kernel-profile --block-threads N   (the per-kernel timings alone, the workload for rocprofv3)
    python bench.py --configs   (configs 0-3 one by one, 0-2 beside the FP64 C oracle on the host cores)

The headline (default) is the mixed-precision solve (FP32 PCG + FP64 residual replacement, mj_cg2d_solve_mixed); the
FP64 solve (mj_cg2d_solve) is timed on the same step for `mixed_over_fp64`, and the two solutions are compared
member by member (`error_vs_fp64`). Prints one JSON line.

`roofline` is the solve's step kernel (the dominant kernel), from the solve alone on its stream (eager, not
overlapped), timed with HIP events: achieved = algorithmic bytes per launch / average launch duration (= the solve's
algorithmic bytes / its time; the average includes the few setup kernels and the launch gaps, so it is an upper bound
of the profiler's per-kernel average). Algorithmic bytes are what the recurrence itself must move, per interior
point (N2 = 105,300 per member): per member-iteration the vectors r, s, x read and written (6 streams of s_v bytes);
per operator slot and iteration in which any member of the slot iterates, the six operator and preconditioner arrays
(6 x s_c); per FP64 residual replacement of the mixed solve 64 B (the FP64 true residual and solution update); s_v = s_c = 4 for
the mixed inner solve, 8 for the FP64 solve. `implementation_bytes` counts what this build's kernels are laid out to
stream instead (mixed: 9.5 FP32 vector streams per inner iteration, 68 B per point per replacement, 48 B per point
once; FP64: 11 FP64 streams per iteration, 48 B per point once; coefficients served from cache not counted).
DST3 (`dst3`): per (member, tile, level) the padded 3-D planes the step reads (tracer, u, v, hFacW, hFacS,
recip_hFacC, maskW, maskS; with the R flux also w) and the tendency it writes, 8 B each, plus the 2-D metrics
(6 planes per tile) once. `traffic` = measured HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE
profile of this workload. `cpu_baseline` is the FP64 C oracle (OpenMP, all host cores) on a bounded sample of the
same member-steps. `producers` times the operator builder (UPDATE_CG2D) and the rhs builder (cg2d_rhs) on the same
ensemble, each alone (not part of the headline step, whose operators and rhs are inputs); `implicit_r` likewise
times GAD_IMPLICIT_R (the implicit vertical advection/diffusion solve) on the ensemble's tracer; `dst3_xyr` the DST3
step with the vertical (R) flux. `single_member` gives B = 1, which is latency-bound and so has no roofline
fraction, as microseconds per iteration of the FP64 and the mixed solve for one member of each level class.
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
N2 = 13 * 90 * 90
TOL = 1e-12
LEVEL_CHOICES = (15, 30, 50)  # the member level counts of the ensemble (mit_jax_amd.workload)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--members", type=int, default=128)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="mixed")
    ap.add_argument("--no-autotune", action="store_true",
                    help="run the launch configuration of --block-threads / --inner-tolerance / --overlap, no graph")
    ap.add_argument("--block-threads", type=int, default=0, help="with --no-autotune (0: the library default)")
    ap.add_argument("--inner-tolerance", type=float, default=1e-4, help="with --no-autotune")
    ap.add_argument("--poll-iters", type=int, default=64,
                    help="with --no-autotune: iterations between the solve's active-member polls (the autotuner's "
                         "candidates use 64)")
    ap.add_argument("--overlap", action="store_true", help="with --no-autotune: DST3 concurrent with the solve")
    ap.add_argument("--dst3-levels", type=int, default=1, help="with --no-autotune: DST3 levels per workgroup")
    ap.add_argument("--no-fp64-reference", action="store_true",
                    help="with --precision mixed: skip timing the FP64 solve (the ratio and the error vs FP64)")
    ap.add_argument("--sum-order", choices=("tile", "fortran"), default="tile")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-producers", action="store_true",
                    help="skip the UPDATE_CG2D / cg2d_rhs / GAD_IMPLICIT_R / DST3-with-R measurements")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB device copy (the reference kernel for HBM counter scaling)")
    ap.add_argument("--kernel-profile", action="store_true",
                    help="only the per-kernel timings (solve alone, DST3 alone, DST3 with R) of the launch "
                         "configuration given by --block-threads / --inner-tolerance, three times each (no autotune): "
                         "the workload of the committed kernel-trace profiles")
    ap.add_argument("--configs", action="store_true",
                    help="only the per-configuration table: config 0 (one FP64 solve, diagonal preconditioner), config "
                         "1 (one and 24 DST3 X/Y/R steps), config 2 (64 rhs, batched FP64 and mixed), each beside the "
                         "FP64 C oracle on the host cores, and config 3 (the 24-step rollout and its gradients); "
                         "prints one JSON line")
    ap.add_argument("--profile-solves", action="store_true",
                    help="only run one FP64 and one mixed solve of --members right-hand sides on the shared 5-point "
                         "operator (the workload of the hardware-counter profiles) and exit")
    return ap.parse_args()


def cpu_baseline(E, seconds):
    """FP64 oracle member-steps on the host cores: CG2D solve (fortran order) + DST3 step per member, cycling
    through the three level classes until `seconds` have passed (at least one member-step)."""
    import oracle
    from oracle import tables as OT
    from mit_jax_amd import synthetic as S

    tab = OT.tables(OT.Topology.from_facets(OT.LLC90_FACETS))
    Tm = E["ex"].maps.full("T")
    t0 = time.perf_counter()
    n = 0
    done = []
    while True:
        lv = LEVEL_CHOICES[n % 3]
        g, ops, norm, u, v, tr = E["cls"][lv][:6]
        b = S.rhs(g, 100 + n)
        ta = time.perf_counter()
        oracle.cg2d(Tm, ops, b, np.zeros_like(b), cg2dNorm=norm, tolerance=TOL, max_iters=2000,
                    sum_order="fortran")
        g2, vg = S.gad_inputs(g)
        oracle.advect_tiles(tab, g2, vg, u, v, tr, g.hFacW, g.hFacS, g.recip_hFacC)
        done.append(time.perf_counter() - ta)
        n += 1
        if time.perf_counter() - t0 > seconds and n >= 1:
            break
    el = sum(done)
    return dict(value=n / el, unit="member-steps/s", cores=oracle.threads(), kind="port",
                sample=f"{n} member-steps (levels cycling 15/30/50; FP64 CG2D to 1e-12 + one DST3 step each), "
                       f"{el:.1f} s")


def plain_fp64(E, params, members):
    """The plain FP64 GPU form: one member per call, as a caller of the reference-signature ops would run the
    ensemble: cg2d_fortran (literal FP64 PCG, launch shape picked for one member) then advect_tiles for that
    member's levels. Returns (member-steps/s, the solutions of the first `members` members)."""
    import torch
    from mit_jax_amd import cg2d, gad
    ex = E["ex"]
    plan = gad.gad_tables(params)
    xs = []

    def one(m):
        li = LEVEL_CHOICES.index(int(E["nlev"][m]))
        ops = [o[li] for o in E["ops"]]
        x, _ = cg2d.cg2d_fortran(ex, ops, E["rhs"][m], E["x0"][m], cg2dNorm=E["norm"][m:m + 1], tolerance=TOL,
                                 max_iters=2000, sum_order="tile")
        c = int(E["g2d"]["mask_index"][m])
        g2d = {k: a for k, a in E["g2d"].items() if k != "mask_index"}
        g2d.update(maskW=E["g2d"]["maskW"][c:c + 1], maskS=E["g2d"]["maskS"][c:c + 1])
        gad.advect_tiles(params, plan, g2d, E["vg"], E["u"][m:m + 1], E["v"][m:m + 1], E["tr"][m:m + 1],
                         E["hW"][m:m + 1], E["hS"][m:m + 1], E["rhC"][m:m + 1], nlev=E["nlev_t"][m:m + 1])
        return x

    one(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for m in range(members):
        xs.append(one(m))
    torch.cuda.synchronize()
    return members / (time.perf_counter() - t0), xs


def single_member(E, reps=5):
    """B = 1 (latency-bound, no roofline fraction): for the first member of each level class, the FP64 solve
    (cg2d_fortran, launch shape picked for one member) and the mixed solve alone, microseconds per iteration
    (wall time of `reps` solves / their iterations)."""
    import torch
    from mit_jax_amd import cg2d, mixed
    ex = E["ex"]
    out = {}
    for li, lv in enumerate(LEVEL_CHOICES):
        m = int(np.nonzero(np.asarray(E["nlev"]) == lv)[0][0])
        ops = [o[li] for o in E["ops"]]
        args = (ex, ops, E["rhs"][m], E["x0"][m])
        kw = dict(cg2dNorm=E["norm"][m:m + 1], tolerance=TOL, max_iters=2000)
        row = {"member": m}
        for name, fn, extra in (("fp64", cg2d.cg2d_fortran, {"sum_order": "tile"}), ("mixed", mixed.cg2d_mixed, {})):
            _, d = fn(*args, **kw, **extra)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                _, d = fn(*args, **kw, **extra)
            torch.cuda.synchronize()
            it = int(d["numIters"])
            row[name] = {"iterations": it, "us_per_iteration": (time.perf_counter() - t0) / reps / it * 1e6}
        out[str(lv)] = row
    return out


TRAFFIC_FILE = ROOT / "profiles" / "mx_step_kernel_ensemble128_traffic.md"
TRAFFIC_FILES = (TRAFFIC_FILE, ROOT / "profiles" / "sfp_impr_kernels_ensemble128_traffic.md")


def measured_traffic(kernel, members):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE/WRITE_SIZE profiles of the 128-member
    workload (None for another workload or kernel)."""
    if members != 128:
        return None
    for path in TRAFFIC_FILES:
        if not path.exists():
            continue
        for line in path.read_text().splitlines():
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
                return float(f[2])
    return None


def profile_solves(B):
    """One FP64 (phase launches) and one mixed solve of B right-hand sides on the shared operator."""
    import torch
    from mit_jax_amd import cg2d, mixed, synthetic as S
    from mit_jax_amd.exchange import default_exchanger
    dev = torch.device("cuda:0")
    ex = default_exchanger()
    g, ops, norm = S.cg2d_case(0, "5point")
    opsd = [torch.as_tensor(o, device=dev) for o in ops]
    b = torch.stack([torch.as_tensor(S.rhs(g, 100 + m), device=dev) for m in range(B)])
    x0 = torch.zeros_like(b)
    _, d = cg2d.cg2d_fortran(ex, opsd, b, x0, cg2dNorm=norm, tolerance=TOL, max_iters=2000, sum_order="tile",
                             launch="phases")
    _, d2 = mixed.cg2d_mixed(ex, opsd, b, x0, cg2dNorm=norm, tolerance=TOL, max_iters=4000)
    torch.cuda.synchronize()
    print(json.dumps({"profile_solves": B, "fp64_iterations": int(d["numIters"].max()),
                      "mixed_iterations": int(d2["numIters"].max())}), flush=True)


def producers(E, dev, reps=5):
    """UPDATE_CG2D and cg2d_rhs (the producers either side of the solve) on the whole ensemble through the C ABI,
    outputs preallocated, each timed alone with HIP events on the current stream (median of reps). Algorithmic
    bytes per member: UPDATE_CG2D reads hFacW, hFacS on the (sNy+1) x (sNx+1) points of every level and writes
    aW2d, aS2d, aC2d (whole planes) and pW, pS, pC ((sNy+1) x (sNx+1)); cg2d_rhs reads hFacW, gU on sNy x (sNx+1)
    and hFacS, gV on (sNy+1) x sNx points of every level, etaN (whole plane), EmPmR, etaH (interior), and writes
    cg2d_b, cg2d_x (whole planes). The GPU operators are checked bitwise against the CPU builder's."""
    import ctypes
    import torch
    import mit_jax_amd._lib as L
    from mit_jax_amd import synthetic as S
    B, T, K, ny, nx = E["hW"].shape
    sN = nx - 8
    g0 = E["cls"][50][0]
    f64 = dict(dtype=torch.float64, device=dev)
    keep = {k: torch.as_tensor(np.ascontiguousarray(g0.f[k]), device=dev)
            for k in ("dyG", "dxG", "recip_dxC", "recip_dyC", "rA", "maskInC")}
    keep["drF"] = torch.as_tensor(np.ascontiguousarray(g0.v["drF"]), device=dev)
    grid = L.SfpGrid(*(keep[k].data_ptr() for k in ("dyG", "dxG", "recip_dxC", "recip_dyC", "rA", "maskInC")), None,
                     keep["drF"].data_ptr(), None, None, None)
    cfg = L.SfpConfig(B, T, sN, sN, 4, K, T * K * ny * nx, T * ny * nx)
    fs = L.FsParams(S.DT, S.DT, 1.0, 1.0, 1.0, 1.0 / 1029.0, 1)
    rBo = torch.full((T, ny, nx), 1.0 / S.G, **f64)
    Bo = torch.full((T, ny, nx), S.G, **f64)
    outs = [torch.zeros((B, T, ny, nx), **f64) for _ in range(6)]
    b, x = torch.empty((B, T, ny, nx), **f64), torch.empty((B, T, ny, nx), **f64)
    gen = torch.Generator(device=dev)
    gen.manual_seed(4001)
    eta, etaH = (torch.randn((B, T, ny, nx), generator=gen, **f64) for _ in range(2))
    em = torch.randn((B, T, ny, nx), generator=gen, **f64) * 1e-5
    plan = E["ex"].plan("T", dev)
    P = L.ptr
    lib = L.lib()

    def upd():
        L.check(lib.mj_update_cg2d(plan.handle, ctypes.byref(grid), ctypes.byref(fs), ctypes.byref(cfg), P(E["norm"]),
                                   0.51, P(E["hW"]), P(E["hS"]), P(rBo), *(P(o) for o in outs), L.stream_ptr()),
                "mj_update_cg2d")

    def rhs():
        L.check(lib.mj_cg2d_rhs(ctypes.byref(grid), ctypes.byref(fs), ctypes.byref(cfg), P(E["u"]), P(E["v"]),
                                P(E["hW"]), P(E["hS"]), P(em), P(eta), P(etaH), P(Bo), P(b), P(x), L.stream_ptr()),
                "mj_cg2d_rhs")

    def t(fn):
        fn()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        return float(np.median(ms))

    print("bench: producers (UPDATE_CG2D, cg2d_rhs)", file=sys.stderr, flush=True)
    tu, tr = t(upd), t(rhs)
    same = all(bool(torch.equal(outs[k][m], E["ops"][k][int(E["op_index"][m])])) for k in range(6) for m in range(B))
    p1, pu, pv, plane = (sN + 1) * (sN + 1), sN * (sN + 1), (sN + 1) * sN, ny * nx
    ub = B * T * 8 * (2 * K * p1 + 3 * plane + 3 * p1)
    rb = B * T * 8 * (2 * K * pu + 2 * K * pv + plane + 2 * sN * sN + 2 * plane)
    res = {}
    for name, ms, nb, kern in (("update_cg2d", tu, ub, "sfp_opsum_kernel"), ("cg2d_rhs", tr, rb, "sfp_rhs_kernel")):
        gbs = nb / (ms * 1e-3) / 1e9
        res[name] = {"ms": ms, "member_calls_per_s": B / (ms * 1e-3), "algorithmic_bytes": nb, "achieved_GBs": gbs,
                     "frac": gbs / HBM_PEAK_GBS, "dominant_kernel": kern,
                     "dominant_kernel_traffic": measured_traffic(kern, B)}
    res["update_cg2d"]["operators_equal_cpu_builder"] = same
    return res


def implicit_r(E, dev, reps=5):
    """GAD_IMPLICIT_R (implicit U3 vertical advection + diffusion, pentadiagonal solve per column) on the whole
    ensemble through the C ABI, timed alone with HIP events (median of reps). Inputs: each member's maskC,
    recip_hFacC and w of its level class, the perturbed tracer, a seeded kappaRX. Algorithmic bytes per member:
    maskC, recip_hFac, kappaRX, wFld and the tracer read on every interior column and level, the tracer read on the
    halo columns, the result written everywhere."""
    import ctypes
    import torch
    import mit_jax_amd._lib as L
    B, T, K, ny, nx = E["tr"].shape
    g0 = E["cls"][50][0]
    w = E["per_member"](lambda c: c[6])
    gen = torch.Generator(device=dev)
    gen.manual_seed(4002)
    kap = (1e-5 + 1e-2 * torch.rand(E["tr"].shape, generator=gen, device=dev, dtype=torch.float64)) * E["mC"]
    keep = [torch.as_tensor(np.ascontiguousarray(a, np.float64), device=dev) for a in
            (g0.f["rA"], g0.f["recip_rA"], g0.v["dTlev"], g0.v["recip_drF"], g0.v["recip_drC"])]
    gs = L.ImprGrid(E["mC"].data_ptr(), *(k.data_ptr() for k in keep), -1.0)
    n3 = T * K * ny * nx
    cfg = L.ImprConfig(B, T, nx - 8, ny - 8, 4, K, n3, n3, 1)
    out = torch.empty_like(E["tr"])
    lib = L.lib()

    def run():
        L.check(lib.mj_gad_implicit_r(ctypes.byref(gs), ctypes.byref(cfg), L.ptr(kap), L.ptr(E["rhC"]), L.ptr(w),
                                      L.ptr(E["tr"]), L.ptr(out), L.stream_ptr()), "mj_gad_implicit_r")

    print("bench: GAD_IMPLICIT_R", file=sys.stderr, flush=True)
    run()
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        run()
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    t = float(np.median(ms))
    interior, cols = T * (nx - 8) * (ny - 8), T * ny * nx
    nb = B * K * 8 * (6 * interior + 2 * (cols - interior))
    gbs = nb / (t * 1e-3) / 1e9
    wt = E["tr"].new_tensor(np.asarray(g0.v["drF"]))[None, None, :, None, None] / E["rhC"].clamp(min=1e-300) * E["mC"]
    cons = float(((wt * out).sum(dim=2) - (wt * E["tr"]).sum(dim=2)).abs().max() /
                 (wt * E["tr"]).sum(dim=2).abs().max())
    del kap, w, out
    kname = "impr_reg_kernel" if K == 50 else "impr_kernel"  # the library's sweep with y' in registers at 50 levels
    return {"ms": t, "member_calls_per_s": B / (t * 1e-3), "algorithmic_bytes": nb, "achieved_GBs": gbs,
            "frac": gbs / HBM_PEAK_GBS, "kernel": kname, "traffic": measured_traffic(kname, B),
            "column_integral_max_rel_change": cons}


def copy_probe(dev):
    """A 1 GiB device-to-device copy: the kernel of known traffic that scales the HBM byte counters."""
    import torch
    a = torch.empty(1 << 27, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def solve_bytes(precision, d, op_index):
    """(algorithmic bytes, implementation bytes, operator-slot iterations) of one batched solve from its
    diagnostics (the counts of the module docstring)."""
    iters = d["numIters"].cpu().numpy().astype(np.int64)
    outer = d["outerIters"].cpu().numpy().astype(np.int64)
    B = len(iters)
    slot = op_index.cpu().numpy() if op_index is not None else np.arange(B)
    slot_iters = sum(int(iters[slot == s].max()) for s in np.unique(slot))  # iterations with a member of the slot active
    sv = 8 if precision == "fp64" else 4
    alg = float(iters.sum()) * N2 * 6 * sv + slot_iters * N2 * 6 * sv
    if precision == "fp64":
        impl = float(iters.sum()) * N2 * 11 * 8 + B * N2 * 48
    else:
        alg += float(outer.sum()) * N2 * 64
        impl = float(iters.sum()) * N2 * 9.5 * 4 + float(outer.sum()) * N2 * 68 + B * N2 * 48
    return alg, impl, slot_iters


def dst3_bytes(nlev, with_w):
    """Algorithmic bytes of one batched DST3 launch: per (member, tile, level) 9 padded planes (10 with w) of 8 B,
    plus the 2-D metrics (6 planes per tile) once."""
    planes = 10 if with_w else 9
    return float(sum(planes * 8 * 13 * 98 * 98 * int(n) for n in nlev)) + 6 * 8 * 13 * 98 * 98


def dst3_xyr(E, dev, params, reps=3):
    """The DST3 step with the vertical (R) flux of each member's w (the headline step's DST3 launch plus the
    vertical sweep), timed alone with HIP events on the current stream (median of reps)."""
    import ctypes
    import torch
    from dataclasses import replace as _replace
    import mit_jax_amd._lib as L
    from mit_jax_amd import gad
    w = E["per_member"](lambda c: c[6])
    g2d = dict(E["g2d"], maskC=E["mC_slots"])
    pr = _replace(params, implicitAdvection=False)
    plan, grid_s, gcfg, trc, fields, wd, keep = gad._prepare(pr, None, g2d, dict(E["vg"]), E["u"], E["v"], E["tr"],
                                                             E["hW"], E["hS"], E["rhC"], w, E["nlev_t"])
    u, v, hW, hS, rhC = fields
    out = torch.zeros_like(trc)
    lib = L.lib()
    wsb = lib.mj_gad_dst3_workspace_bytes(ctypes.byref(gcfg), trc.shape[1], 1)
    work = torch.empty(wsb, dtype=torch.uint8, device=dev)

    def run():
        L.check(lib.mj_gad_dst3(plan.handle, ctypes.byref(grid_s), ctypes.byref(gcfg), L.ptr(u), L.ptr(v), L.ptr(wd),
                                L.ptr(trc), L.ptr(hW), L.ptr(hS), L.ptr(rhC), L.ptr(out), L.ptr(work),
                                ctypes.c_size_t(wsb), L.stream_ptr()), "mj_gad_dst3")

    print("bench: DST3 with the R flux", file=sys.stderr, flush=True)
    run()
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        run()
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    t = float(np.median(ms))
    nb = dst3_bytes(E["nlev"], True)
    del w, out, work, keep
    gbs = nb / (t * 1e-3) / 1e9
    return {"ms": t, "member_steps_per_s": len(E["nlev"]) / (t * 1e-3), "algorithmic_bytes": nb,
            "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}


def kernel_times(st, reps=3):
    """The solve alone and the DST3 step alone (eager, polled, not overlapped, no graph), HIP events on the current
    stream (both launch on it): the per-kernel times behind `roofline` and `dst3` (the timed steps may overlap the
    two). Returns (solve ms, DST3 ms, the solve's diagnostics), medians of reps."""
    import torch
    from dataclasses import replace as _replace
    st.drop_graph()
    st.set_config(_replace(st.cfg, graph=False, overlap=False))
    ts, ta = [], []
    for _ in range(reps):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record()
        st._solve()
        e[1].record()
        st._advect()
        e[2].record()
        torch.cuda.synchronize()
        ts.append(e[0].elapsed_time(e[1]))
        ta.append(e[1].elapsed_time(e[2]))
    return float(np.median(ts)), float(np.median(ta)), st.diag().check()


def roofline_of(precision, cg_ms, d, op_index, members, launch):
    """The `roofline` object of a solve timed alone (cg_ms) with diagnostics d."""
    alg, impl, slot_iters = solve_bytes(precision, d, op_index)
    launches = int(d["launches"].max())
    if precision == "fp64":
        kernel = "cg2d_persistent_kernel" if launch == "persistent" else "cg2d_phase_kernel"
    else:
        kernel = "mx_step_kernel"
    achieved = alg / (cg_ms * 1e-3) / 1e9
    tr = measured_traffic(kernel, members)
    return {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": tr, "traffic_source": TRAFFIC_FILE.name if tr else None,
            "launches": launches, "avg_launch_us": cg_ms * 1e3 / launches if launches else None,
            "algorithmic_bytes_per_launch": alg / launches if launches else None, "algorithmic_bytes": alg,
            "operator_slot_iterations": slot_iters, "implementation_bytes": impl,
            "implementation_GBs": impl / (cg_ms * 1e-3) / 1e9, "solve_ms": cg_ms}


def make_step(E, params, precision, args):
    from mit_jax_amd.ensemble import EnsembleStep, StepConfig
    return EnsembleStep(E["ex"], E["ops"], E["op_index"], E["norm"], E["rhs"], E["x0"], params, E["g2d"], E["vg"],
                        E["u"], E["v"], E["tr"], E["hW"], E["hS"], E["rhC"], nlev=E["nlev_t"], tolerance=TOL,
                        max_iters=2000, cfg=StepConfig(precision=precision, sum_order=args.sum_order,
                                                       block_threads=args.block_threads, poll_iters=args.poll_iters,
                                                       inner_tolerance=args.inner_tolerance, overlap=args.overlap,
                                                       dst3_levels=args.dst3_levels))


def kernel_profile(args):
    """The per-kernel timings of one launch configuration alone: the workload of the kernel-trace profiles."""
    import torch
    import mit_jax_amd
    from mit_jax_amd import gad
    from mit_jax_amd.workload import build_ensemble
    dev = torch.device("cuda:0")
    B = args.members
    print(f"bench: building {B} members", file=sys.stderr, flush=True)
    E = build_ensemble(B, dev)
    params = gad.GADParams.llc90()
    st = make_step(E, params, args.precision, args)
    if args.copy_probe:
        copy_probe(dev)
    print("bench: kernel timings (solve alone, DST3 alone)", file=sys.stderr, flush=True)
    cg_ms, gad_ms, d = kernel_times(st)
    gb = dst3_bytes(E["nlev"], False)
    out = {"kernel_profile": B, "precision": args.precision,
           "launch": {k: v for k, v in st.cfg.__dict__.items() if k != "precision"},
           "cg2d_iterations_mean": float(d["numIters"].float().mean()),
           "roofline": roofline_of(args.precision, cg_ms, d, E["op_index"], B, st.cfg.launch),
           "dst3": {"launch_ms": gad_ms, "achieved_GBs": gb / (gad_ms * 1e-3) / 1e9, "algorithmic_bytes": gb,
                    "frac": gb / (gad_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
           "dst3_xyr": dst3_xyr(E, dev, params)}
    print(json.dumps(out), flush=True)
    del st, E
    mit_jax_amd.release()


def _median_ms(fn, reps):
    """Median wall time (ms) of `reps` calls of fn after one warm-up call, each followed by a device sync."""
    import torch
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(ts))


def configs_cpu_baseline(c0, c1, c2):
    """The CPU baseline of the configs table: the FP64 C oracle (OpenMP, all host cores) on the same inputs, timed
    (medians of a few calls; config 2 from 2 of its 64 solves x 32), with its results for the table's equality
    flags. c0 = (ops, b, kw), c1 = (grid2d, vgrid, u, v, w, tracer, hFacW, hFacS, recip_hFacC), c2 = (ops, bs, kw)."""
    import oracle
    from oracle import tables as OT
    from mit_jax_amd.exchange import default_exchanger
    Tm = default_exchanger().maps.full("T")

    def timed(fn, reps):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            r = fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), r

    ops, b, kw = c0
    t0, (x0, d0) = timed(lambda: oracle.cg2d(Tm, ops, b, np.zeros_like(b), sum_order="fortran", **kw), 3)
    tab = OT.tables(OT.Topology.from_facets(OT.LLC90_FACETS))
    t1, g1 = timed(lambda: oracle.gad_dst3(tab, *c1), 3)
    ops2, bs, kw2 = c2
    t2, x2 = timed(lambda: [oracle.cg2d(Tm, ops2, bs[m], np.zeros_like(bs[m]), sum_order="tile", **kw2)[0]
                            for m in range(2)], 1)
    return {"cores": oracle.threads(), "config0": (t0, x0, int(d0["numIters"])), "config1": (t1, g1),
            "config2": (t2 * 32, x2)}


def configs_table():
    """Configs 0-3 of the workload set on the GPU; configs 0-2 beside the FP64 C oracle on the host cores
    (configs_cpu_baseline). Config 0: one solve of the 50-level operator with the diagonal preconditioner to 1e-12
    (FP64, reference sum order, launch picked for one member; and the mixed solve). Config 1: one DST3 step with the
    X/Y/R fluxes (one member, 50 levels) and 24 steps, the tracer updated and re-exchanged after each. Config 2: 64
    rhs on the config-0 operator in one launch, FP64 (tile order) and mixed; the mixed solutions' error against the
    FP64 ones, per rhs (max and mean). The FP64 results of configs 0 and 1, and of config 2's first 2 rhs, are
    compared bitwise with the oracle's. Config 3: the 24-step coupled rollout (CG2D + momentum correction + DST3 X/Y,
    checkpoint every 6 steps) forward alone, and forward with the reverse-mode gradients of its loss w.r.t. T0, u0,
    v0 and b (no CPU counterpart: the oracle has no reverse mode). GB/s are algorithmic bytes (the headline's
    counts) over the wall time."""
    import torch
    from mit_jax_amd import cg2d, gad, mixed, rollout, synthetic as S
    from mit_jax_amd.exchange import default_exchanger
    dev = torch.device("cuda:0")
    ex = default_exchanger()
    Td = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)  # noqa: E731
    inner = (slice(None), slice(4, 94), slice(4, 94))
    out, res = {}, {}

    print("bench: config 0", file=sys.stderr, flush=True)
    g, ops, norm = S.cg2d_case(0, "diagonal")
    b = S.rhs(g, 1)
    opsd = [Td(o) for o in ops]
    bd, x0 = Td(b), torch.zeros((13, 98, 98), dtype=torch.float64, device=dev)
    kw = dict(cg2dNorm=norm, tolerance=TOL, max_iters=4000)
    ms = _median_ms(lambda: res.update(r=cg2d.cg2d_fortran(ex, opsd, bd, x0, sum_order="fortran", **kw)), 5)
    x64, d64 = res["r"]
    it = int(d64["numIters"])
    msm = _median_ms(lambda: res.update(m=mixed.cg2d_mixed(ex, opsd, bd, x0, **kw)), 5)
    xm, dm = res["m"]
    out["config0"] = {"iterations": it, "gpu_fp64_ms": ms, "gpu_fp64_us_per_iteration": ms * 1e3 / it,
                      "gpu_fp64_GBs": it * N2 * 96 / (ms * 1e-3) / 1e9, "gpu_mixed_ms": msm,
                      "mixed_inner_iterations": int(dm["numIters"]), "mixed_replacements": int(dm["outerIters"])}

    print("bench: config 1", file=sys.stderr, flush=True)
    g1 = S.make_grid(0)
    u, v, w = S.flow(g1, 2)
    tr = S.tracer(g1, 3)
    params = gad.GADParams.llc90(implicitAdvection=False)
    g2, vg = S.gad_inputs(g1)
    g2d = {k: Td(a) for k, a in g2.items()}
    g2d["rA"] = Td(g1.rA)
    vgd = {k: Td(np.asarray(a)) for k, a in vg.items()}
    dtd = Td(np.asarray(vg["dTlev"])[None, :, None, None])
    ud, vd, wd, hWd, hSd, rhCd, trd = (Td(a) for a in (u, v, w, g1.hFacW, g1.hFacS, g1.recip_hFacC, tr))
    plan = gad.gad_tables(params)
    step = lambda t: gad.advect_tiles(params, plan, g2d, vgd, ud, vd, t, hWd, hSd, rhCd, wFld=wd)  # noqa: E731
    ms1 = _median_ms(lambda: step(trd), 5)

    def chain():
        t = trd.clone()
        for _ in range(24):
            t = t + step(t) * dtd
            ex.fill_(t, "T")
        return t

    ms24 = _median_ms(chain, 3)
    gT = step(trd).cpu().numpy()
    out["config1"] = {"gpu_step_ms": ms1, "gpu_step_GBs": dst3_bytes([50], True) / (ms1 * 1e-3) / 1e9,
                      "gpu_24_steps_ms": ms24}

    print("bench: config 3", file=sys.stderr, flush=True)
    case = rollout.RolloutCase.synthetic(dev)
    X = [Td(a) for a in (tr, u, v)] + [Td(S.rhs(g1, 1) * 1e-3)]

    def forward():
        with torch.no_grad():
            rollout.rollout(case, *X, nsteps=24, checkpoint_every=6, ex=ex)

    def forward_backward():
        leaves = [x.clone().requires_grad_(True) for x in X]
        T, _, _ = rollout.rollout(case, *leaves, nsteps=24, checkpoint_every=6, ex=ex)
        torch.autograd.grad(rollout.loss(case, T), leaves)

    out["config3"] = {"gpu_forward_24_steps_ms": _median_ms(forward, 3),
                      "gpu_forward_and_gradients_ms": _median_ms(forward_backward, 3)}

    print("bench: config 2", file=sys.stderr, flush=True)
    bs = np.stack([S.rhs(g, s) for s in range(100, 164)])
    bsd, x0s = Td(bs), torch.zeros(bs.shape, dtype=torch.float64, device=dev)
    ms64 = _median_ms(lambda: res.update(b64=cg2d.cg2d_fortran(ex, opsd, bsd, x0s, sum_order="tile", **kw)), 3)
    msmx = _median_ms(lambda: res.update(bmx=mixed.cg2d_mixed(ex, opsd, bsd, x0s, **kw)), 3)
    (xb, db), (xmx, dmx) = res["b64"], res["bmx"]
    one = torch.zeros(64, dtype=torch.int32, device=dev)
    xbn, xmn = xb.cpu().numpy(), xmx.cpu().numpy()
    errs = [float(np.abs(xmn[m] - xbn[m])[inner].max() / np.abs(xbn[m][inner]).max()) for m in range(64)]
    out["config2"] = {"gpu_fp64_ms": ms64, "gpu_fp64_GBs": solve_bytes("fp64", db, one)[0] / (ms64 * 1e-3) / 1e9,
                      "gpu_mixed_ms": msmx, "gpu_mixed_GBs": solve_bytes("mixed", dmx, one)[0] / (msmx * 1e-3) / 1e9,
                      "mixed_over_fp64": ms64 / msmx, "fp64_iterations_mean": float(db["numIters"].float().mean()),
                      "mixed_iterations_mean": float(dmx["numIters"].float().mean()),
                      "all_converged": bool((dmx["converged"] == 1).all() and (db["converged"] == 1).all()),
                      "mixed_error_vs_fp64_max": max(errs), "mixed_error_vs_fp64_mean": float(np.mean(errs))}

    print("bench: configs CPU baseline", file=sys.stderr, flush=True)
    cb = configs_cpu_baseline((ops, b, kw), (dict(g2, rA=g1.rA), vg, u, v, w, tr, g1.hFacW, g1.hFacS,
                                             g1.recip_hFacC), (ops, bs, kw))
    out["cores"] = cb["cores"]
    t0, xr, itr = cb["config0"]
    out["config0"].update(cpu_fp64_s=t0, oracle_iterations=itr,
                          gpu_fp64_equals_oracle=bool(np.array_equal(x64.cpu().numpy(), xr)),
                          mixed_error_vs_oracle=float(np.abs(xm.cpu().numpy() - xr)[inner].max()
                                                      / np.abs(xr[inner]).max()))
    t1, gr = cb["config1"]
    out["config1"].update(cpu_step_s=t1, gpu_step_equals_oracle=bool(np.array_equal(gT, gr)))
    t2, x2 = cb["config2"]
    out["config2"].update(cpu_fp64_s=t2, fp64_equals_oracle_first_2=all(np.array_equal(xbn[m], x2[m])
                                                                        for m in range(2)))
    print(json.dumps({"configs": out}), flush=True)
    import mit_jax_amd
    mit_jax_amd.release()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the ensemble shards over GPUs with no communication, so --gpus {args.gpus}"
              " is not run here (one process per GPU would each run this same single-GPU step)", file=sys.stderr)
        sys.exit(2)
    if args.profile_solves:
        return profile_solves(args.members)
    if args.kernel_profile:
        return kernel_profile(args)
    if args.configs:
        return configs_table()
    import torch
    from mit_jax_amd import cg2d, gad
    from mit_jax_amd.ensemble import autotune

    dev = torch.device("cuda:0")
    B = args.members
    print(f"bench: building {B} members", file=sys.stderr, flush=True)
    from mit_jax_amd.workload import build_ensemble
    E = build_ensemble(B, dev)
    params = gad.GADParams.llc90()
    log = lambda m: print("bench: autotune " + m, file=sys.stderr, flush=True)  # noqa: E731

    def make(precision):
        st = make_step(E, params, precision, args)
        tuned = None
        if not args.no_autotune:
            best, results = autotune(st, precision, steps=2, warmup=1, log=log)
            keys = ("block_threads", "inner_tolerance", "overlap", "dst3_levels", "graph")
            tuned = {"selected": {**best.config.__dict__}, "candidates": len(results),
                     "ms_per_step": [[*(r.config.__dict__[k] for k in keys), round(r.ms, 3)] for r in results],
                     "columns": [*keys, "ms"]}
        return st, tuned

    def timed(st):
        for _ in range(args.warmup):
            st.step()
        torch.cuda.synchronize()
        ev = []
        print(f"bench: timing {st.cfg.precision}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st.step(ev)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        d = st.diag().check()
        return dict(el=el, d=d, cg_ms=float(np.mean([a.elapsed_time(b) for a, b, _ in ev])),
                    gad_ms=float(np.mean([b.elapsed_time(c) for _, b, c in ev])))

    if args.copy_probe:
        copy_probe(dev)
    main_step, tuned = make(args.precision)
    run = timed(main_step)
    sel = main_step.cfg  # the timed configuration (kernel_times then times the parts alone)
    run["cg_ms"], run["gad_ms"], dk = kernel_times(main_step)
    d = run["d"]
    iters = d["numIters"].cpu().numpy().astype(np.int64)
    outer = d["outerIters"].cpu().numpy().astype(np.int64)
    conv = d["converged"].cpu().numpy()
    gad_bytes = dst3_bytes(E["nlev"], False)
    ms = run["el"] / args.steps * 1e3
    cfg = {"workload": f"LLC90 ensemble: {B} members x (CG2D to 1e-12 normalised residual + one DST3 "
                       "step), 13x90x90 tiles, 50 levels, member level counts 15/30/50",
           "members": B, "precision": args.precision, "sum_order": args.sum_order if args.precision == "fp64" else
           "tile", "cg2d_iterations_mean": float(iters.mean()), "cg2d_iterations_max": int(iters.max()),
           "replacements_mean": float(outer.mean()), "all_converged": bool((conv == 1).all()),
           "launch": {k: v for k, v in sel.__dict__.items() if k != "precision"}, "autotune": tuned}
    out = {
        "metric": "member-steps/s, cg2d+DST3 at LLC90; mixed-refined vs FP64 at equal residual",
        "value": B * args.steps / run["el"],
        "unit": "member-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": cfg,
        "roofline": roofline_of(args.precision, run["cg_ms"], dk, E["op_index"], B, sel.launch),
        "dst3": {"launch_ms": run["gad_ms"], "achieved_GBs": gad_bytes / (run["gad_ms"] * 1e-3) / 1e9,
                 "algorithmic_bytes": gad_bytes, "frac": gad_bytes / (run["gad_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                 "kernel": "gad_hz_kernel", "traffic": measured_traffic("gad_hz_kernel", B)},
    }
    if args.precision == "mixed" and not args.no_fp64_reference:
        ref_step, ref_tuned = make("fp64")
        ref = timed(ref_step)
        ref_sel = ref_step.cfg
        ref["cg_ms"], _, dref = kernel_times(ref_step)
        inner = cg2d.interior(E["ex"].L, dev)
        x, xr = main_step.x, ref_step.x
        diff = torch.where(inner, (x - xr).abs(), 0.0).amax(dim=(1, 2, 3))
        scale = torch.where(inner, xr.abs(), 0.0).amax(dim=(1, 2, 3))
        rel = (diff / scale).cpu().numpy()
        fp64_value = B * args.steps / ref["el"]
        out["fp64"] = {"value": fp64_value, "ms_per_step": ref["el"] / args.steps * 1e3, "solve_ms": ref["cg_ms"],
                       "roofline": roofline_of("fp64", ref["cg_ms"], dref, E["op_index"], B, ref_sel.launch),
                       "cg2d_iterations_mean": float(ref["d"]["numIters"].float().mean()),
                       "launch": {k: v for k, v in ref_sel.__dict__.items() if k != "precision"},
                       "autotune": ref_tuned}
        out["mixed_over_fp64"] = out["value"] / fp64_value
        out["mixed_over_fp64_solve"] = ref["cg_ms"] / run["cg_ms"]
        out["error_vs_fp64"] = {"max_rel": float(rel.max()), "mean_rel": float(rel.mean()),
                                "measure": "per member max|x_mixed - x_fp64| / max|x_fp64| over the interior"}
        print("bench: timing the plain FP64 form (one member per call)", file=sys.stderr, flush=True)
        pv, xs = plain_fp64(E, params, B)
        same = all(bool(torch.equal(torch.where(inner, xp, 0.0), torch.where(inner, xr[m], 0.0)))
                   for m, xp in enumerate(xs))
        out["plain_fp64"] = {"value": pv, "form": "one member per call: cg2d_fortran (FP64, launch auto) + "
                                                  "advect_tiles (member's levels)",
                             "solutions_equal_batched_fp64": same}
        out["mixed_over_plain_fp64"] = out["value"] / pv
        out["single_member"] = single_member(E)
    if not args.no_producers:
        out["producers"] = producers(E, dev)
        out["implicit_r"] = implicit_r(E, dev)
        out["dst3_xyr"] = dst3_xyr(E, dev, params)
    if not args.no_cpu_baseline:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(E, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    del main_step, E
    import mit_jax_amd
    mit_jax_amd.release()


if __name__ == "__main__":
    main()
