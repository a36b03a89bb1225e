"""GPU: the benchmark workloads at their full sizes against the oracle, member by member.

* config 2: 64 right-hand sides (seeds 100..163) in one mixed-precision launch on the config-0 operator (diagonal
  preconditioner): each solution's true normalised FP64 residual < 1.5e-12 (evaluated with the oracle's operator)
  and its error against the oracle's FP64 solution <= 1e-8 of max |x| (printed per rhs).
* config 4: the B = 32 and B = 128 ensembles of `mit_jax_amd.workload` (level counts 15/30/50, three operator
  slots, perturbed tracer and velocity) through `EnsembleStep`, the object bench.py times: the FP64 step's solution
  and iteration count bitwise the oracle's `cg2d` (tree order) per member; the mixed step's solution as in config 2;
  the DST3 tendency bitwise the oracle's `advect_tiles` on each member's active levels and 0 beyond them; the
  captured-graph replay equal to the eager step.
* config 1: 24 DST3 steps with the vertical (R) flux, T <- T + dt gT then the T exchange, bitwise the oracle's at
  every step.

The per-solve gate follows the reference's (tests/test_cg2d.py:127-145): per solve, the solution and its iteration
count against the oracle's.
"""

import numpy as np
import pytest

import oracle
from tests.conftest import maps, oracle_tables

pytestmark = pytest.mark.gpu

TOL = 1e-12
X_REL = 1e-8
RES_BOUND = 1.5 * TOL


def _true_residual(ops, norm, b, x):
    """Normalised FP64 residual ||(b cg2dNorm - A x) / max|b cg2dNorm||| with the oracle's operator."""
    y = oracle.cg2d_matvec(maps().full("T"), ops[:3], x, norm)
    inner = np.zeros(b.shape, bool)
    inner[:, 4:94, 4:94] = True
    bn = b * norm
    mx = np.abs(bn[inner]).max()
    if mx == 0:
        return 0.0
    r = (bn - y * norm)[inner] / mx
    return float(np.sqrt((r * r).sum()))


def _rel_err(x, xr):
    return float(np.abs(x - xr)[:, 4:94, 4:94].max() / np.abs(xr[:, 4:94, 4:94]).max())


# ------------------------------------------------------------------------------------------------------- config 2
def test_config2_64_rhs_one_mixed_launch(dev):
    import torch
    from mit_jax_amd import mixed, synthetic as S
    from mit_jax_amd.exchange import default_exchanger
    ex = default_exchanger()
    g, ops, norm = S.cg2d_case(0, "diagonal")
    bs = np.stack([S.rhs(g, s) for s in range(100, 164)])
    print("config 2: 64 rhs, one mixed launch", flush=True)
    x, d = mixed.cg2d_mixed(ex, [torch.as_tensor(o, device=dev) for o in ops], torch.as_tensor(bs, device=dev),
                            torch.zeros(bs.shape, dtype=torch.float64, device=dev), cg2dNorm=norm, tolerance=TOL,
                            max_iters=4000)
    xs = x.cpu().numpy()
    assert d["converged"].cpu().tolist() == [1] * 64
    errs = []
    for m in range(64):
        xr, dr = oracle.cg2d(maps().full("T"), ops, bs[m], np.zeros_like(bs[m]), cg2dNorm=norm, tolerance=TOL,
                             max_iters=4000, sum_order="tile")
        res = _true_residual(ops, norm, bs[m], xs[m])
        err = _rel_err(xs[m], xr)
        errs.append(err)
        print(f"rhs {100 + m}: inner {int(d['numIters'][m])} (fp64 {dr['numIters']}), replacements "
              f"{int(d['outerIters'][m])}, true residual {res:.2e}, error vs FP64 oracle {err:.2e}", flush=True)
        assert res < RES_BOUND, (m, res)
        assert err < X_REL, (m, err)
        assert float(d["firstResidual"][m]) == dr["firstResidual"]
    print(f"config 2: max error vs FP64 oracle {max(errs):.2e}, mean {np.mean(errs):.2e}", flush=True)


# ------------------------------------------------------------------------------------------------------- config 4
@pytest.fixture(scope="module", params=[32, 128], ids=["B32", "B128"])
def ens(request, dev):
    import torch
    from mit_jax_amd.workload import build_ensemble
    B = request.param
    print(f"config 4: building {B} members", flush=True)
    E = build_ensemble(B, dev)
    E["B"] = B
    E["oracle_x"] = {}
    yield E
    E.clear()
    torch.cuda.empty_cache()


def _step(E, precision, **kw):
    from mit_jax_amd import gad
    from mit_jax_amd.ensemble import EnsembleStep, StepConfig
    return EnsembleStep(E["ex"], E["ops"], E["op_index"], E["norm"], E["rhs"], E["x0"], gad.GADParams.llc90(),
                        E["g2d"], E["vg"], E["u"], E["v"], E["tr"], E["hW"], E["hS"], E["rhC"], nlev=E["nlev_t"],
                        tolerance=TOL, max_iters=2000, cfg=StepConfig(precision=precision, **kw))


def _oracle_x(E, m):
    """The oracle's FP64 solution of member m (tree order), cached on the ensemble."""
    if m not in E["oracle_x"]:
        li = int(E["nlev"][m])
        g, ops, norm = E["cls"][li][:3]
        b = E["rhs"][m].cpu().numpy()
        E["oracle_x"][m] = oracle.cg2d(maps().full("T"), ops, b, np.zeros_like(b), cg2dNorm=norm, tolerance=TOL,
                                       max_iters=2000, sum_order="tile")
    return E["oracle_x"][m]


def test_config4_fp64_step_bitwise(ens):
    import torch
    E = ens
    st = _step(E, "fp64")
    print(f"config 4 B={E['B']}: FP64 step", flush=True)
    st.step()
    torch.cuda.synchronize()
    d = st.diag().check()
    x = st.x.cpu().numpy()
    it = d["numIters"].cpu().numpy()
    assert bool((d["converged"] == 1).all())
    for m in range(E["B"]):
        xr, dr = _oracle_x(E, m)
        assert int(it[m]) == dr["numIters"], (m, int(it[m]), dr["numIters"])
        assert np.array_equal(x[m], xr), f"member {m} (levels {E['nlev'][m]}): FP64 solution differs"
    print(f"config 4 B={E['B']}: {E['B']} FP64 solutions bitwise the oracle's, iterations "
          f"{int(it.min())}..{int(it.max())}", flush=True)


def test_config4_mixed_step_vs_oracle(ens):
    import torch
    E = ens
    st = _step(E, "mixed")
    print(f"config 4 B={E['B']}: mixed step", flush=True)
    st.step()
    torch.cuda.synchronize()
    d = st.diag().check()
    assert bool((d["converged"] == 1).all())
    x = st.x.cpu().numpy()
    halo = ~np.pad(np.ones((90, 90), bool), 4)[None].repeat(13, 0)
    errs, ress = [], []
    for m in range(E["B"]):
        li = int(E["nlev"][m])
        ops, norm = E["cls"][li][1], E["cls"][li][2]
        xr, _ = _oracle_x(E, m)
        res = _true_residual(ops, norm, E["rhs"][m].cpu().numpy(), x[m])
        err = _rel_err(x[m], xr)
        errs.append(err)
        ress.append(res)
        assert res < RES_BOUND, (m, res)
        assert err < X_REL, (m, err)
        assert np.array_equal(x[m][halo], xr[halo])
    print(f"config 4 B={E['B']}: mixed true residual max {max(ress):.2e}; error vs FP64 oracle max {max(errs):.2e} "
          f"mean {np.mean(errs):.2e}", flush=True)


def test_config4_dst3_bitwise(ens):
    import torch
    from mit_jax_amd import synthetic as S
    from mit_jax_amd.workload import member_host
    E = ens
    st = _step(E, "mixed")
    print(f"config 4 B={E['B']}: DST3 step", flush=True)
    st._advect()
    torch.cuda.synchronize()
    gT = st.gT
    tab = oracle_tables()
    for m in range(E["B"]):
        h = member_host(E, m)
        K = h["levels"]
        g2, vg = S.gad_inputs(h["grid"])
        g2 = dict(g2, maskW=h["maskW"], maskS=h["maskS"])
        vg = {k: np.asarray(a)[:K] for k, a in vg.items()}
        ref = oracle.advect_tiles(tab, g2, vg, h["u"], h["v"], h["tr"], h["hW"], h["hS"], h["rhC"])
        got = gT[m].cpu().numpy()
        assert np.array_equal(got[:, :K], ref), f"member {m} (levels {K}): DST3 tendency differs"
        assert not got[:, K:].any(), f"member {m}: levels beyond {K} not zero"
    print(f"config 4 B={E['B']}: {E['B']} DST3 tendencies bitwise the oracle's", flush=True)


def test_config4_graph_replay_equals_eager(ens):
    import torch
    E = ens
    if E["B"] != 32:
        pytest.skip("graph replay checked at B = 32")
    st = _step(E, "mixed")
    st.step()
    torch.cuda.synchronize()
    x_e, gT_e = st.x.clone(), st.gT.clone()
    mi, mo = st.cap_from(st.diag())
    st.capture(mi, mo)
    for rep in range(3):
        st.x.zero_()
        junk = torch.randn(1 << 20, device=E["rhs"].device).sum()  # eager work between replays
        torch.cuda.synchronize()
        st.step()
        torch.cuda.synchronize()
        print(f"config 4 B=32: replay {rep}", flush=True)
        assert torch.equal(st.x, x_e) and torch.equal(st.gT, gT_e), rep
        del junk


# ------------------------------------------------------------------------------------------------------- config 1
def test_config1_24_steps_with_r_flux_bitwise(dev):
    import torch
    from mit_jax_amd import gad, synthetic as S
    from mit_jax_amd.exchange import default_exchanger
    from tests.conftest import dst3_case
    ex = default_exchanger()
    g, u, v, w, tr = dst3_case(0)
    params = gad.GADParams.llc90(implicitAdvection=False)
    g2, vg = S.gad_inputs(g)
    Td = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)  # noqa: E731
    g2d = {k: Td(a) for k, a in g2.items()}
    g2d["rA"] = Td(g.rA)
    vgd = {k: Td(np.asarray(a)) for k, a in vg.items()}
    dt = np.asarray(vg["dTlev"])[None, :, None, None]
    dtd = Td(dt)
    args = [Td(a) for a in (u, v, w, g.hFacW, g.hFacS, g.recip_hFacC)]
    td = Td(tr)
    tab = oracle_tables()
    Tm = maps().full("T")
    tc = tr.copy()
    for n in range(24):
        gT = gad.advect_tiles(params, None, g2d, vgd, args[0], args[1], td, args[3], args[4], args[5],
                              wFld=args[2])
        td = td + gT * dtd
        ex.fill_(td, "T")
        gr = oracle.gad_dst3(tab, dict(g2, rA=g.rA), vg, u, v, w, tc, g.hFacW, g.hFacS, g.recip_hFacC)
        tc = oracle.scalar(Tm, tc + gr * dt)
        torch.cuda.synchronize()
        assert np.array_equal(gT.cpu().numpy(), gr), f"step {n}: tendency differs"
        assert np.array_equal(td.cpu().numpy(), tc), f"step {n}: tracer differs"
        print(f"config 1: step {n + 1} bitwise (max |gT| {np.abs(gr).max():.3e})", flush=True)
