"""The reference's dump-replay gates (tests/test_cg2d.py:116-160, tests/test_free_surface.py:78-125) through the
GPU path, on a synthetic jaxdump directory written from the CPU oracle's results: S07 (UPDATE_CG2D), C01 (cg2d_b,
cg2d_x) and C02 (the solve, halos included, and its iteration count), each bitwise. With dumps of the Fortran model
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
        with open(D.dump_name(d_dir, it, t + 1), "wb") as fh:
            for seq, ((stage, field), a) in enumerate(fields.items()):
                D.write_record(fh, it, seq, stage, field, "RL", t + 1, FACE[t], 90 * (t % 3), 0, a[t])
    return SimpleNamespace(ds=D.DumpSet(d_dir), it=it, g=g, p=p, cp=cp, iters=d["numIters"])


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
