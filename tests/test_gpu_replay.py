"""The reference's dump-replay gates (tests/test_cg2d.py:116-160, tests/test_free_surface.py:78-125,
tests/test_gad.py:119-130) through the GPU path, on synthetic jaxdump directories written from the CPU oracle's
results: S07 (UPDATE_CG2D), C01 (cg2d_b, cg2d_x), C02 (the solve, halos included, and its iteration count) and
T10 / T20 (GAD_ADVECTION of theta and salt on the whole tile arrays), each bitwise. With dumps of the Fortran model
the same functions (`mit_jax_amd.replay`) run the reference's gates."""

from types import SimpleNamespace

import numpy as np
import pytest

import oracle
from oracle import sfp as OS
from tests.conftest import maps

pytestmark = pytest.mark.gpu

FACE = [1, 1, 1, 2, 2, 2, 3, 4, 4, 4, 5, 5, 5]
OPS = ("aW2d", "aS2d", "aC2d", "pW", "pS", "pC")


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    from mit_jax_amd import dump as D, synthetic as S
    from mit_jax_amd.cg2d import Cg2dParams
    from mit_jax_amd.solve_for_pressure import FreeSurfParams
    g = S.make_grid(0)
    rng = np.random.default_rng(40)
    T, K = 13, 50
    norm = S.ini_cg2d_norm(g)
    p = FreeSurfParams(deltaTMom=S.DT, deltaTFreeSurf=S.DT)
    cp = Cg2dParams(cg2dMaxIters=2000, cg2dTolerance=1e-12, cg2dNorm=norm, sum_order="fortran")
    z = np.zeros((T, 98, 98))
    rBo = np.full((T, 98, 98), 1.0 / S.G)
    Bo = np.full((T, 98, 98), S.G)
    gU = rng.normal(size=(T, K, 98, 98)) * g.maskW * 1e-4
    gV = rng.normal(size=(T, K, 98, 98)) * g.maskS * 1e-4
    EmPmR = rng.normal(size=(T, 98, 98)) * 1e-5
    etaN = rng.normal(size=(T, 98, 98)) * 0.1
    etaH = etaN + rng.normal(size=(T, 98, 98)) * 1e-3
    ops = OS.update_cg2d(p, norm, 0.51, g, maps().full("T"), g.hFacW, g.hFacS, rBo, z, z, z)
    b, x0 = OS.cg2d_rhs(p, g, gU, gV, g.hFacW, g.hFacS, EmPmR, etaN, etaH, Bo)
    x, d = oracle.cg2d(maps().full("T"), ops, b, x0, cg2dNorm=norm, tolerance=1e-12, max_iters=2000,
                       sum_order="fortran")
    fields = {("G00_geometry", "Bo_surf"): Bo, ("G00_geometry", "recip_Bo"): rBo,
              ("S04_oceanic_phys", "EmPmR"): EmPmR, ("S05_dynamics", "gU"): gU, ("S05_dynamics", "gV"): gV,
              ("S05_dynamics", "etaN"): etaN, ("S05_dynamics", "etaH"): etaH,
              ("S06_update_rstar_T", "hFacW"): g.hFacW, ("S06_update_rstar_T", "hFacS"): g.hFacS}
    fields.update({("S07_update_cg2d", n): o for n, o in zip(OPS, ops)})
    fields.update({("C01_cg2d_inputs", n): o for n, o in zip(OPS, ops)})
    fields.update({("C01_cg2d_inputs", "cg2d_b"): b, ("C01_cg2d_inputs", "cg2d_x"): x0,
                   ("C02_cg2d_solution", "cg2d_x"): x})
    d_dir = tmp_path_factory.mktemp("jaxdump")
    it = 8
    for t in range(T):
        with open(D.file_name(d_dir, it, t + 1), "wb") as fh:
            for seq, ((stage, field), a) in enumerate(fields.items()):
                D.append_record(fh, it, seq, stage, field, "RL", t + 1, FACE[t], 90 * (t % 3), 0, a[t])
    return SimpleNamespace(ds=D.DumpIndex(d_dir), it=it, g=g, p=p, cp=cp, iters=d["numIters"])


@pytest.fixture(scope="module")
def ex():
    from mit_jax_amd.exchange import default_exchanger
    return default_exchanger()


def test_replay_update_cg2d(dev, case, ex):
    from mit_jax_amd import replay
    z = np.zeros((13, 98, 98))
    print("replay S07 UPDATE_CG2D", flush=True)
    got, ref = replay.replay_update_cg2d(case.ds, case.it, case.p, case.cp, case.g, ex, z, z, z)
    for n in replay.OPS:
        assert np.array_equal(got[n], ref[n]), n


def test_replay_cg2d_rhs(dev, case, ex):
    from mit_jax_amd import replay
    print("replay C01 cg2d_rhs", flush=True)
    (b, x), (rb, rx) = replay.replay_cg2d_rhs(case.ds, case.it, case.p, case.g, ex)
    assert np.array_equal(b, rb) and np.array_equal(x, rx)


def test_replay_cg2d_solution(dev, case, ex):
    from mit_jax_amd import replay
    print("replay C02 CG2D", flush=True)
    x, d, rx = replay.replay_cg2d(case.ds, case.it, ex, case.cp)
    assert np.array_equal(x, rx)
    assert int(d["numIters"]) == case.iters


def test_replay_gad_t10_t20(dev, tmp_path):
    """T01 residual flow + S04 tracers + S11 hFac + G00 geometry -> T10 gT_loc / T20 gS_loc, bitwise (the
    reference's gate, tests/test_gad.py:119-130); a planted error in the dumped output is caught."""
    from mit_jax_amd import dump as D, gad, replay, synthetic as S
    from tests.conftest import dst3_case, oracle_tables
    g, u, v, w, theta = dst3_case(0)
    salt = 34.0 + 0.3 * np.random.default_rng(8).normal(size=theta.shape) * g.maskC
    salt = oracle.scalar(maps().full("T"), salt)
    g2, vg = S.gad_inputs(g)
    tab = oracle_tables()
    refs = {n: oracle.advect_tiles(tab, g2, vg, u, v, tr, g.hFacW, g.hFacS, g.recip_hFacC)
            for n, tr in (("temp", theta), ("salt", salt))}
    L = g.L
    vert = D.pack_vertical(L, {"drF": g.drF, "recip_drF": g.recip_drF, "rhoFacC": g.rhoFacC,
                               "recip_drC": np.append(g.recip_drC, g.recip_drC[-1])})
    fields = {("G00_geometry", n): g.f[n] for n in replay.GAD_GRID2D + replay.GAD_GRID3D}
    fields[("G00_geometry", "vertical")] = np.broadcast_to(vert, (13,) + vert.shape)
    fields.update({("T01_residual_flow", "uFld"): u, ("T01_residual_flow", "vFld"): v, ("T01_residual_flow", "wFld"): w,
                   ("S04_oceanic_phys", "theta"): theta, ("S04_oceanic_phys", "salt"): salt,
                   ("S11_calc_rstar", "hFacW"): g.hFacW, ("S11_calc_rstar", "hFacS"): g.hFacS,
                   ("S11_calc_rstar", "recip_hFacC"): g.recip_hFacC,
                   ("T10_temp_adv", "gT_loc"): refs["temp"], ("T20_salt_adv", "gS_loc"): refs["salt"]})
    it = 3
    for t in range(13):
        with open(D.file_name(tmp_path, it, t + 1), "wb") as fh:
            for seq, ((stage, field), a) in enumerate(fields.items()):
                D.append_record(fh, it, seq, stage, field, "RL", t + 1, FACE[t], 90 * (t % 3), 0, a[t])
    ds = D.DumpIndex(tmp_path)
    print("replay T10/T20 GAD_ADVECTION", flush=True)
    out = replay.replay_gad(ds, it, gad.GADParams.llc90())
    for n, (got, dumped) in out.items():
        assert np.array_equal(got, dumped), n
        assert np.abs(dumped).max() > 0
    # negative control: oneSixth off by 1e-6 relative fails the gate
    bad = replay.replay_gad(ds, it, gad.GADParams.llc90(oneSixth=(1.0 / 6.0) * (1 + 1e-6)), tracers=("temp",))
    assert not np.array_equal(bad["temp"][0], bad["temp"][1])
