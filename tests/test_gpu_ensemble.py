"""GPU: the prepared ensemble step — eager step equal to the library's solve/advect calls, graph replays equal to
the eager step on every replay (with eager work and synchronisations between replays), autotuner picks a
converged configuration."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ensemble(dev, B=4):
    import torch
    from mit_jax_amd import gad, synthetic as S
    from mit_jax_amd.exchange import default_exchanger
    ex = default_exchanger()
    g, ops, norm = S.cg2d_case(0, "5point")
    u, v, _ = S.flow(g, 2)
    tr = S.tracer(g, 3)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)  # noqa: E731
    rhs = torch.stack([T(S.rhs(g, 300 + m)) for m in range(B)])
    st = lambda a: T(np.stack([a] * B))  # noqa: E731
    trb = T(np.stack([tr + 0.01 * m for m in range(B)]))
    g2, vg = S.gad_inputs(g)
    return dict(ex=ex, ops=[T(o) for o in ops], norm=torch.full((B,), norm, dtype=torch.float64, device=dev),
                rhs=rhs, x0=torch.zeros_like(rhs), params=gad.GADParams.llc90(), g2d={k: T(a) for k, a in g2.items()},
                vg=vg, u=st(u), v=st(v), tr=trb, hW=st(g.hFacW), hS=st(g.hFacS), rhC=st(g.recip_hFacC))


def _step(E, cfg):
    from mit_jax_amd.ensemble import EnsembleStep
    return EnsembleStep(E["ex"], E["ops"], None, E["norm"], E["rhs"], E["x0"], E["params"], E["g2d"], E["vg"],
                        E["u"], E["v"], E["tr"], E["hW"], E["hS"], E["rhC"], cfg=cfg)


@pytest.mark.parametrize("precision,overlap", [("fp64", False), ("mixed", False), ("mixed", True)])
def test_graph_replays_equal_eager(dev, precision, overlap):
    import torch
    from mit_jax_amd import cg2d, gad, mixed
    from mit_jax_amd.ensemble import StepConfig
    E = _ensemble(dev)
    s = _step(E, StepConfig(precision=precision, overlap=overlap))
    s.step()
    torch.cuda.synchronize()
    x_e, gT_e, d = s.x.clone(), s.gT.clone(), s.diag()
    assert bool((d["converged"] == 1).all())
    # the eager step is the library's own calls
    if precision == "fp64":
        xr, _ = cg2d.cg2d_fortran(E["ex"], E["ops"], E["rhs"], E["x0"], cg2dNorm=E["norm"], tolerance=1e-12,
                                  max_iters=2000, sum_order="tile", launch="phases")
    else:
        xr, _ = mixed.cg2d_mixed(E["ex"], E["ops"], E["rhs"], E["x0"], cg2dNorm=E["norm"], tolerance=1e-12,
                                 max_iters=4000)
    gr = gad.advect_tiles(E["params"], None, E["g2d"], E["vg"], E["u"], E["v"], E["tr"], E["hW"], E["hS"], E["rhC"])
    assert torch.equal(x_e, xr) and torch.equal(gT_e, gr)
    mi, mo = s.cap_from(d)
    s.capture(mi, mo)
    for rep in range(3):
        s.x.zero_()
        s.gT.zero_()
        junk = torch.randn(1 << 20, device=dev).sum()  # eager work between replays
        torch.cuda.synchronize()
        s.step()
        torch.cuda.synchronize()
        print(f"{precision} replay {rep}", flush=True)
        assert torch.equal(s.x, x_e), rep
        assert torch.equal(s.gT, gT_e), rep
        assert bool((s.diag()["converged"] == 1).all())
        del junk


def test_autotune_selects_converged(dev):
    from mit_jax_amd.ensemble import StepConfig, autotune
    E = _ensemble(dev)
    s = _step(E, StepConfig(precision="mixed"))
    best, results = autotune(s, "mixed", steps=1, warmup=1)
    assert best.all_converged and len(results) >= 2
    s.step()
    assert bool((s.diag()["converged"] == 1).all())
