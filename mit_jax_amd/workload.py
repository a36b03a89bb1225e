"""The seeded LLC90 ensemble workload of the batched CG2D + DST3 step (13 tiles of 90x90, halo 4, 50 levels).

B members, each with a level count drawn from {15, 30, 50}; members with the same level count share one operator
and one set of grid masks (three operator slots and three mask slots, `op_index` / `mask_index` pick a member's); per
member a rhs (seed 100 + m), the level class's
streamfunction flow and tracer perturbed by tracer noise 0.05 degC on wet cells and velocity x (1 + 0.05 N(0, 1))
(torch generator seeded 4000 on the device), halos re-exchanged. Built on the device once; `bench.py` times steps
over it and the parity tests check every member against the oracle.

`member_host(E, m)` returns one member's inputs as numpy arrays (device copies), sliced to its active levels where
the oracle is to run on them.
"""

import numpy as np
import torch

from . import synthetic as S
from .exchange import default_exchanger

LEVEL_CHOICES = (15, 30, 50)


def member_levels(B, seed=4000):
    """Level count of each of B members (seeded draw from LEVEL_CHOICES)."""
    return np.random.default_rng(seed).choice(LEVEL_CHOICES, size=B)


def level_classes():
    """{levels: (grid, ops (5-point preconditioner), cg2dNorm, u, v, tracer, w)} of the three level classes."""
    cls = {}
    for n in LEVEL_CHOICES:
        g, ops, norm = S.cg2d_case(0, "5point", nlev=int(n))
        u, v, w = S.flow(g, 2)
        tr = S.tracer(g, 3)
        cls[int(n)] = (g, ops, norm, u, v, tr, w)
    return cls


def build_ensemble(B, dev):
    """Per-member inputs on the device: level counts, operator slots, rhs, flow, tracer, hFac and masks."""
    ex = default_exchanger()
    nlev = member_levels(B)
    cls = level_classes()
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)  # noqa: E731
    idx = torch.as_tensor([LEVEL_CHOICES.index(int(n)) for n in nlev], device=dev)

    def per_member(get):
        base = torch.stack([T(get(cls[n])) for n in LEVEL_CHOICES])
        return base.index_select(0, idx).contiguous()

    # members with the same level count share one operator: 3 slots, op_index picks each member's
    ops = [torch.stack([T(cls[n][1][k]) for n in LEVEL_CHOICES]) for k in range(6)]
    op_index = idx.to(torch.int32)
    norm = torch.as_tensor([cls[int(n)][2] for n in nlev], device=dev, dtype=torch.float64)
    rhs = torch.stack([T(S.rhs(cls[int(n)][0], 100 + m)) for m, n in enumerate(nlev)])
    x0 = torch.zeros_like(rhs)
    gen = torch.Generator(device=dev)
    gen.manual_seed(4000)
    def per_class(get):
        return torch.stack([T(get(cls[n])) for n in LEVEL_CHOICES]).contiguous()

    u = per_member(lambda c: c[3])
    v = per_member(lambda c: c[4])
    tr = per_member(lambda c: c[5])
    # grid masks: one slot per level class, member m reads slot idx[m]
    mW = per_class(lambda c: c[0].maskW)
    mS = per_class(lambda c: c[0].maskS)
    mC_slots = per_class(lambda c: c[0].maskC)
    mC = mC_slots.index_select(0, idx).contiguous()  # per member: GAD_IMPLICIT_R's maskC, the tracer noise mask
    hW = per_member(lambda c: c[0].hFacW)
    hS = per_member(lambda c: c[0].hFacS)
    rhC = per_member(lambda c: c[0].recip_hFacC)
    # perturbations: tracer noise 0.05 degC on wet cells, velocity x (1 + 0.05 N(0,1)); halos re-exchanged
    tr += torch.randn(tr.shape, generator=gen, device=dev, dtype=torch.float64) * 0.05 * mC
    u *= 1.0 + 0.05 * torch.randn(u.shape, generator=gen, device=dev, dtype=torch.float64)
    v *= 1.0 + 0.05 * torch.randn(v.shape, generator=gen, device=dev, dtype=torch.float64)
    for m in range(B):
        ex.fill_(tr[m], "T")
        uu, vv = ex.vector(u[m], v[m], "UVs")
        u[m].copy_(uu)
        v[m].copy_(vv)
    g0 = cls[50][0]
    g2d = {k: T(g0.f[k]) for k in ("dyG", "dxG", "recip_dxC", "recip_dyC", "recip_rA", "rA", "maskInC")}
    g2d.update(maskW=mW, maskS=mS, mask_index=idx.to(torch.int32))
    vg = {k: g0.v[k] for k in ("drF", "recip_drF", "rhoFacC", "recip_drC")}
    return dict(ex=ex, nlev=nlev, nlev_t=torch.as_tensor(nlev.astype(np.int32), device=dev), ops=ops,
                op_index=op_index, norm=norm, mC=mC, mC_slots=mC_slots, per_member=per_member,
                rhs=rhs, x0=x0, u=u, v=v, tr=tr, hW=hW, hS=hS, rhC=rhC, g2d=g2d, vg=vg, cls=cls)


def member_host(E, m, active_levels=True):
    """Member m's inputs as numpy arrays: the level class's grid, operator and cg2dNorm, its rhs, and its DST3
    fields (u, v, tracer, hFacW, hFacS, recip_hFacC, maskW, maskS; the first nlev levels when active_levels)."""
    n = int(E["nlev"][m])
    g, ops, norm = E["cls"][n][:3]
    K = n if active_levels else E["tr"].shape[2]
    f = {k: np.ascontiguousarray(E[k][m, :, :K].cpu().numpy())
         for k in ("u", "v", "tr", "hW", "hS", "rhC")}
    c = int(E["g2d"]["mask_index"][m])
    f["maskW"] = np.ascontiguousarray(E["g2d"]["maskW"][c, :, :K].cpu().numpy())
    f["maskS"] = np.ascontiguousarray(E["g2d"]["maskS"][c, :, :K].cpu().numpy())
    return dict(nlev=n, levels=K, grid=g, ops=ops, norm=norm, rhs=E["rhs"][m].cpu().numpy(), **f)
