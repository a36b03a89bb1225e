"""A differentiable multi-step rollout of the hot path: CG2D surface-pressure solve + DST3 tracer step, with reverse
mode through the hand-written adjoints and checkpointing.

One step (the coupling is this package's own; the reference couples the same two kernels through its full time
step, core/forward_step.py:300-360 and core/solve_for_pressure.py:109-154):
    eta   = cg2d_solve(p, ex, ops, b, 0)                    (implicit-derivative adjoint, mit_jax_amd.adjoint)
    eta   = exch_xy(eta)                                    (transposed exchange in reverse)
    u    += -dt g (eta(i) - eta(i-1)) recip_dxC maskW       on the interior, every level (SOLVE_FOR_PRESSURE's
    v    += -dt g (eta(j) - eta(j-1)) recip_dyC maskS       momentum correction, solve_for_pressure.py:140-154)
    u, v  = exch_uv_xy(u, v)                                (signed vector exchange)
    gT    = advect_tiles(u, v, T)                           (DST3 X/Y, hand-written reverse mode)
    T     = exch_xy(T + dt gT maskC)
The loss of the 24-step rollout is sum(rA drF hFacC T_final) over the interior; its gradients with respect to T0,
the rhs b and the initial velocities u0, v0 come from torch.autograd. `checkpoint_every` steps form one segment
whose forward is recomputed in the reverse pass (torch.utils.checkpoint), so only segment inputs are stored.
"""

from dataclasses import dataclass

import torch
from torch.utils.checkpoint import checkpoint

from . import adjoint, gad
from .cg2d import Cg2dParams
from .exchange import default_exchanger

G = 9.81


@dataclass
class RolloutCase:
    """Device inputs of a rollout: the operator (six arrays), solver parameters, the DST3 grid and constants."""
    ops: list
    p: Cg2dParams
    params: object
    g2d: dict
    vg: dict
    hFacW: torch.Tensor
    hFacS: torch.Tensor
    recip_hFacC: torch.Tensor
    maskC: torch.Tensor
    maskW: torch.Tensor
    maskS: torch.Tensor
    recip_dxC: torch.Tensor
    recip_dyC: torch.Tensor
    weight: torch.Tensor  # rA drF hFacC on the interior, 0 elsewhere
    dt: float = 3600.0

    @classmethod
    def synthetic(cls, dev, seed=0, tolerance=1e-12):
        """The synthetic LLC90 case: seed grid, 5-point preconditioned operator, DST3 LLC90 parameters."""
        import numpy as np
        from . import synthetic as S
        g, ops, norm = S.cg2d_case(seed, "5point")
        g2, vg = S.gad_inputs(g)
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)  # noqa: E731
        inner = np.zeros(g.L.shape2d, bool)
        inner[:, g.L.OLy:g.L.OLy + g.L.sNy, g.L.OLx:g.L.OLx + g.L.sNx] = True
        w = g.rA[:, None] * g.drF[None, :, None, None] * g.hFacC * inner[:, None]
        return cls(ops=[T(o) for o in ops], p=Cg2dParams(cg2dMaxIters=2000, cg2dTolerance=tolerance, cg2dNorm=norm,
                                                         sum_order="tile"),
                   params=gad.GADParams.llc90(), g2d={k: T(a) for k, a in g2.items()}, vg=vg,
                   hFacW=T(g.hFacW), hFacS=T(g.hFacS), recip_hFacC=T(g.recip_hFacC), maskC=T(g.maskC),
                   maskW=T(g.maskW), maskS=T(g.maskS), recip_dxC=T(g.recip_dxC), recip_dyC=T(g.recip_dyC),
                   weight=T(w))


def step(case: RolloutCase, ex, T, u, v, b):
    """One coupled step; returns (T_new, u_new, v_new). Differentiable in T, u, v, b."""
    L = ex.L
    eta, _ = adjoint.cg2d_solve(case.p, ex, case.ops, b, torch.zeros_like(b))
    eta = ex.scalar_ad(eta, "T")
    j0, j1, i0, i1 = L.OLy, L.OLy + L.sNy, L.OLx, L.OLx + L.sNx
    gx = torch.zeros_like(eta)
    gy = torch.zeros_like(eta)
    gx[:, j0:j1, i0:i1] = (eta[:, j0:j1, i0:i1] - eta[:, j0:j1, i0 - 1:i1 - 1]) * case.recip_dxC[:, j0:j1, i0:i1]
    gy[:, j0:j1, i0:i1] = (eta[:, j0:j1, i0:i1] - eta[:, j0 - 1:j1 - 1, i0:i1]) * case.recip_dyC[:, j0:j1, i0:i1]
    u = u - (case.dt * G) * gx[:, None] * case.maskW
    v = v - (case.dt * G) * gy[:, None] * case.maskS
    u, v = ex.vector_ad(u.contiguous(), v.contiguous(), "UVs")
    gT = gad.advect_tiles_ad(case.params, None, case.g2d, case.vg, u, v, T, case.hFacW, case.hFacS,
                             case.recip_hFacC)
    T = ex.scalar_ad((T + case.dt * gT * case.maskC).contiguous(), "T")
    return T, u, v


def rollout(case: RolloutCase, T0, u0, v0, b, nsteps=24, checkpoint_every=6, ex=None):
    """nsteps coupled steps in segments of checkpoint_every (0: no checkpointing); returns (T, u, v)."""
    ex = ex or default_exchanger()

    def segment(n):
        def run(T, u, v, b):
            for _ in range(n):
                T, u, v = step(case, ex, T, u, v, b)
            return T, u, v
        return run

    T, u, v = T0, u0, v0
    done = 0
    while done < nsteps:
        n = min(checkpoint_every, nsteps - done) if checkpoint_every else nsteps
        if checkpoint_every and torch.is_grad_enabled():
            T, u, v = checkpoint(segment(n), T, u, v, b, use_reentrant=False)
        else:
            T, u, v = segment(n)(T, u, v, b)
        done += n
    return T, u, v


def loss(case: RolloutCase, T):
    """sum(rA drF hFacC T) over the interior."""
    return (case.weight * T).sum()
