"""Batched ensemble step (CG2D solve + DST3 tracer step for B members) prepared once, launched eagerly or replayed
from a captured HIP graph, with an autotuner over its launch parameters.

`EnsembleStep` validates and converts every argument once (operator slots, configs, DST3 grid structs), allocates
the solver workspace, diagnostics and outputs once, and then launches the C entry points directly: a step does no
Python-side allocation, validation or host read-back besides the solver's own active-member polls. `capture()`
records one step into a `torch.cuda.CUDAGraph` (a hipGraph on ROCm) with the solve enqueued for a fixed number of
iterations (`MJ_CG_LAUNCH_PHASES_FIXED`: no read-back; members that finish early exit their blocks at once), so a
replay is one graph launch per part (solve, DST3) from the host. The capture cap comes from an eager step
(`cap_from`); after a replay `step()` reads the convergence flags and runs the step again eagerly when a member
stopped at the cap unconverged.

`autotune()` times candidate configurations (launch shape and block size of the solve, inner tolerance of the
mixed solve, polling chunk, eager vs graph) on the actual inputs and returns the fastest; results are deterministic
per configuration (the launch shape never changes the FP64 bits; the mixed solve's bits depend on its inner
tolerance only through the iterates, and every candidate meets the same stopping rule).
"""

import ctypes
import json
import time
from dataclasses import dataclass, field, replace

import torch

from . import _lib, gad
from .cg2d import Cg2dDiag, _as_batch, _norm_tensor, _ops_struct, config


@dataclass(frozen=True)
class StepConfig:
    precision: str = "mixed"      # "mixed" or "fp64"
    launch: str = "auto"          # solve launch shape: auto, persistent, phases
    block_threads: int = 0        # 0 = library default for the shape
    poll_iters: int = 0           # iterations between active-member polls (0 = library default)
    inner_tolerance: float = 1e-4  # mixed: reduction of the recursive residual between FP64 replacements
    inner_max_iters: int = 400
    max_outer: int = 40
    sum_order: str = "tile"       # fp64 only
    graph: bool = False           # replay a captured graph
    overlap: bool = False         # run the DST3 step on a second stream, concurrently with the solve
    dst3_levels: int = 1          # DST3 tile shape: consecutive levels of one (member, tile) per workgroup


@dataclass
class Timing:
    config: StepConfig
    ms: float
    all_converged: bool
    extra: dict = field(default_factory=dict)


class EnsembleStep:
    """One batched step over fixed device buffers. Inputs as `bench.build_ensemble` returns them: ops (six arrays,
    shared [T, ny, nx] or slots [C, T, ny, nx] with op_index), cg2dNorm [B], rhs / x0 [B, T, ny, nx]; DST3 fields
    [B, T, K, ny, nx] with per-member level counts nlev."""

    def __init__(self, ex, ops, op_index, cg2dNorm, rhs, x0, params, g2d, vg, u, v, tr, hFacW, hFacS, recip_hFacC,
                 nlev=None, tolerance=1e-12, max_iters=2000, cfg: StepConfig = StepConfig()):
        self.ex = ex
        self.L = ex.L
        self.rhs = _as_batch(rhs, self.L, "rhs")
        self.x0 = _as_batch(x0, self.L, "x0")
        self.B = self.rhs.shape[0]
        dev = self.rhs.device
        self.dev = dev
        self.ops_s, self._ops_keep = _ops_struct(ops, self.B, self.L, dev, op_index)
        self.nslots = self._ops_keep[0].shape[0] if self._ops_keep[0].dim() == 4 else 1
        self.norm = _norm_tensor(cg2dNorm, self.B, dev)
        self.tolerance = float(tolerance)
        self.max_iters = int(max_iters)
        self.plan_T = ex.plan("T", dev)
        self.x = torch.empty_like(self.rhs)
        self.diag_raw = torch.empty(self.B * 48, dtype=torch.uint8, device=dev)
        plan, grid_s, gcfg, trc, fields, _, keep = gad._prepare(params, None, g2d, vg, u, v, tr, hFacW, hFacS,
                                                                 recip_hFacC, None, nlev)
        self._gad = (plan, grid_s, gcfg, trc, fields, keep)
        self.gwsb = _lib.lib().mj_gad_dst3_workspace_bytes(ctypes.byref(gcfg), ex.L.nTiles, 0)
        if self.gwsb == 0:
            raise _lib.MjError(f"mj_gad_dst3_workspace_bytes: {_lib.last_error()}")
        self.gws = torch.empty(self.gwsb, dtype=torch.uint8, device=dev)
        # levels at or beyond a member's level count are never written by the DST3 launch: they stay 0
        self.gT = torch.zeros_like(trc)
        self.graph = None
        self.fallbacks = 0  # graph replays re-run eagerly because a member hit the captured iteration cap
        self.graph_cfg = None
        self.set_config(cfg)

    # ------------------------------------------------------------------------------------------------ config
    def set_config(self, cfg: StepConfig, max_iters=None, max_outer=None, fixed=False):
        self.cfg = cfg
        self._gad[2].levels_per_block = int(cfg.dst3_levels)
        lib = _lib.lib()
        # the mixed solve caps total inner iterations (about 1.1x the FP64 count to the same residual)
        mi = (self.max_iters if cfg.precision == "fp64" else 2 * self.max_iters) if max_iters is None else int(max_iters)
        launch = "phases_fixed" if fixed else ("phases" if cfg.precision == "mixed" and cfg.launch == "auto"
                                               else cfg.launch)
        if cfg.precision == "mixed":
            self.ccfg = config(self.L, self.B, self.tolerance, mi, "tile", cfg.block_threads, 0, launch,
                               cfg.poll_iters)
            self.mcfg = _lib.Cg2dMixedConfig(int(self.nslots), float(cfg.inner_tolerance), int(cfg.inner_max_iters),
                                             int(cfg.max_outer if max_outer is None else max_outer))
            wsb = lib.mj_cg2d_mixed_workspace_bytes(ctypes.byref(self.ccfg), ctypes.byref(self.mcfg),
                                                    self.ops_s.member_stride)
        else:
            self.ccfg = config(self.L, self.B, self.tolerance, mi, cfg.sum_order, cfg.block_threads, 0, launch,
                               cfg.poll_iters)
            self.mcfg = None
            wsb = lib.mj_cg2d_workspace_bytes(ctypes.byref(self.ccfg))
        if wsb == 0:
            raise _lib.MjError(f"workspace query: {_lib.last_error()}")
        if getattr(self, "ws", None) is None or self.ws.numel() < wsb:
            self.ws = torch.empty(wsb, dtype=torch.uint8, device=self.dev)
        self.wsb = wsb

    # ------------------------------------------------------------------------------------------------ launches
    def _solve(self):
        lib = _lib.lib()
        args = (self.plan_T.handle, ctypes.byref(self.ops_s), _lib.ptr(self.rhs), _lib.ptr(self.x0), _lib.ptr(self.x),
                _lib.ptr(self.norm), ctypes.byref(self.ccfg))
        tail = (_lib.ptr(self.ws), ctypes.c_size_t(self.wsb), _lib.ptr(self.diag_raw), _lib.stream_ptr())
        if self.mcfg is not None:
            _lib.check(lib.mj_cg2d_solve_mixed(*args, ctypes.byref(self.mcfg), *tail), "mj_cg2d_solve_mixed")
        else:
            _lib.check(lib.mj_cg2d_solve(*args, *tail), "mj_cg2d_solve")

    def _advect(self):
        plan, grid_s, gcfg, trc, fields, _ = self._gad
        u, v, hW, hS, rhC = fields
        _lib.check(_lib.lib().mj_gad_dst3(plan.handle, ctypes.byref(grid_s), ctypes.byref(gcfg), _lib.ptr(u),
                                          _lib.ptr(v), None, _lib.ptr(trc), _lib.ptr(hW), _lib.ptr(hS), _lib.ptr(rhC),
                                          _lib.ptr(self.gT), _lib.ptr(self.gws), ctypes.c_size_t(self.gwsb),
                                          _lib.stream_ptr()), "mj_gad_dst3")

    def _side(self):
        if getattr(self, "side", None) is None:
            self.side = torch.cuda.Stream()  # default priority
        return self.side

    def _both(self, e1=None):
        """Solve then DST3, or with overlap the DST3 step forked onto the side stream and joined after the solve."""
        if self.cfg.overlap:
            cur = torch.cuda.current_stream()
            side = self._side()
            side.wait_stream(cur)
            with torch.cuda.stream(side):
                self._advect()
            self._solve()
            if e1 is not None:
                e1.record()
            cur.wait_stream(side)
        else:
            self._solve()
            if e1 is not None:
                e1.record()
            self._advect()

    def step(self, events=None, verify=True):
        """One step: the graph if one is captured and selected, else eager launches. events: optional list that
        receives (start, after-solve, end) HIP events (with overlap, after-solve marks the solve's own stream).

        A captured solve runs for a fixed number of iterations (`capture`). With verify (the default) a replay is
        followed by a read of the members' convergence flags, and when a member stopped at that cap unconverged
        the step runs again eagerly (same inputs, so the eager step's result; counted in `fallbacks`). Without
        verify the caller reads `diag()` itself."""
        e1 = None
        if events is not None:
            e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            e0.record()
        if self.graph is not None and self.cfg.graph:
            if len(self.graph) == 1:
                self.graph[0].replay()
                if e1 is not None:
                    e1.record()
            else:
                self.graph[0].replay()
                if e1 is not None:
                    e1.record()
                self.graph[1].replay()
            if verify and not bool((self.diag()["converged"] == 1).all()):
                self.fallbacks += 1
                self._both()
        else:
            self._both(e1)
        if events is not None:
            e2.record()
            events.append((e0, e1, e2))

    def diag(self):
        return Cg2dDiag(self.diag_raw, True)

    # ------------------------------------------------------------------------------------------------ graphs
    def cap_from(self, d, slack=1.25, extra=16):
        """Capture caps from an eager step's diagnostics: (max_iters, max_outer)."""
        iters = int(d["numIters"].max())
        outer = int(d["outerIters"].max())
        return int(iters * slack) + extra, outer + 2

    def capture(self, max_iters, max_outer=None):
        """Record one step into two graphs, the solve (enqueued for max_iters iterations without read-back) and the
        DST3 step, replayed back to back."""
        cfg = replace(self.cfg, graph=True)
        self.set_config(cfg, max_iters=max_iters, max_outer=max_outer, fixed=True)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):  # warm the fixed-length path once outside the capture
            self._both()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        if cfg.overlap:  # one graph with the DST3 branch forked off the solve
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._both()
            self.graph = (g,)
        else:
            gs, ga = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
            with torch.cuda.graph(gs):
                self._solve()
            with torch.cuda.graph(ga):
                self._advect()
            self.graph = (gs, ga)
        torch.cuda.synchronize()
        self.graph_cfg = cfg
        return self.graph

    def drop_graph(self):
        self.graph = None
        self.graph_cfg = None


# ---------------------------------------------------------------------------------------------------- autotuner
def _time(step: EnsembleStep, steps, warmup):
    for _ in range(warmup):
        step.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step.step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


def candidates(precision, B, T=13):
    """Launch candidates for a batch of B members: persistent shapes only where the units fit one launch."""
    out = []
    for ov in (False, True):
        if precision == "mixed":
            for itol in (1e-3, 1e-4):
                for bt in (256, 448, 512, 1024):
                    out.append(StepConfig("mixed", "phases", bt, 64, itol, overlap=ov))
        else:
            for bt in (128, 256, 512, 1024):
                out.append(StepConfig("fp64", "phases", bt, 64, overlap=ov))
            if B * T <= 512:
                for bt in (512, 1024):
                    out.append(StepConfig("fp64", "persistent", bt, overlap=ov))
    return out


DST3_LEVELS = (1, 2, 5, 10)  # levels per workgroup tried for the DST3 launch


def autotune(step: EnsembleStep, precision="mixed", steps=3, warmup=1, try_graph=True, log=None):
    """Time every candidate on the step's own inputs: the solve's launch shapes (block size, inner tolerance,
    overlap), then the DST3 tile shape (levels per workgroup) with the best of those, then the graph replay of the
    best; leave the fastest selected and return (best Timing, all Timings)."""
    results = []

    def trial(c):
        step.drop_graph()
        step.set_config(c)
        ms = _time(step, steps, warmup)
        d = step.diag()
        conv = bool((d["converged"] == 1).all())
        results.append(Timing(c, ms, conv))
        if log:
            log(json.dumps({"config": c.__dict__, "ms": ms, "converged": conv}))
        return results[-1]

    for c in candidates(precision, step.B):
        trial(c)
    ok = [r for r in results if r.all_converged] or results
    best = min(ok, key=lambda r: r.ms)
    for lv in DST3_LEVELS:
        if lv != best.config.dst3_levels:
            r = trial(replace(best.config, dst3_levels=lv))
            if r.all_converged and r.ms < best.ms:
                best = r
    step.drop_graph()
    step.set_config(best.config)
    if try_graph:
        step.step()
        torch.cuda.synchronize()
        mi, mo = step.cap_from(step.diag())
        step.capture(mi, mo if precision == "mixed" else None)
        ms = _time(step, steps, warmup)
        d = step.diag()
        conv = bool((d["converged"] == 1).all())
        gt = Timing(step.cfg, ms, conv, {"cap_iters": mi})
        results.append(gt)
        if log:
            log(json.dumps({"config": step.cfg.__dict__, "ms": ms, "converged": conv, "cap_iters": mi}))
        if conv and ms < best.ms:
            best = gt
        else:
            step.drop_graph()
            step.set_config(best.config)
    return best, results
