"""The C ABI library and the host-side logic, without a GPU: the library loads, exports every entry point the
header declares, answers pure host queries; the ops refuse to run anywhere but on a GPU (no CPU fallback)."""

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "mjamd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mj_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from mit_jax_amd import _lib
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    nm = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mj_\w+)", nm))
    assert set(names) <= exported
    assert set(names) == set(_lib.SIGNATURES), "the ctypes table must bind exactly the declared entry points"


def test_workspace_query_and_config_validation():
    import ctypes
    from mit_jax_amd import _lib, cg2d
    from mit_jax_amd.layout import LLC90
    cfg = cg2d.config(LLC90, 4, 1e-12, 100, "tile")
    n = _lib.lib().mj_cg2d_workspace_bytes(ctypes.byref(cfg))
    # r, s ping-pong buffers, z = M r and q = A s: 6 x 4 members x 124,852 doubles, plus partials
    assert n >= 6 * 4 * LLC90.npad * 8
    bad = cg2d.config(LLC90, 4, 1e-12, 100, "tile", block_threads=100)
    assert _lib.lib().mj_cg2d_workspace_bytes(ctypes.byref(bad)) == 0
    assert "block_threads" in _lib.last_error()


def test_mixed_workspace_and_config_validation():
    """The mixed solve's workspace query: compact FP32 work vectors (7 x members x tiles x 8,288 values: the 8,100
    interior values and copies of the two edge columns, 128-byte aligned) besides the
    FP64 residual; invalid configurations are refused with a message, before any device work."""
    import ctypes
    from mit_jax_amd import _lib, cg2d
    from mit_jax_amd.layout import LLC90
    L = _lib.lib()
    B = 4
    cfg = cg2d.config(LLC90, B, 1e-12, 100, "tile", launch="phases")
    mcfg = _lib.Cg2dMixedConfig(1, 1e-5, 400, 40)
    n = L.mj_cg2d_mixed_workspace_bytes(ctypes.byref(cfg), ctypes.byref(mcfg), 0)
    assert n >= 7 * B * LLC90.nTiles * 8288 * 4 + B * LLC90.npad * 8
    for field, value, word in (("sNy", 89, "sNy"), ("OL", 0, "OL"), ("sum_order", 0, "MJ_SUM_TREE"),
                               ("tolerance", 0.0, "tolerance"), ("block_threads", 2048, "block_threads")):
        bad = cg2d.config(LLC90, B, 1e-12, 100, "tile", launch="phases")
        setattr(bad, field, value)
        assert L.mj_cg2d_mixed_workspace_bytes(ctypes.byref(bad), ctypes.byref(mcfg), 0) == 0, field
        assert word in _lib.last_error(), (field, _lib.last_error())
    for mbad, word in ((_lib.Cg2dMixedConfig(1, 0.0, 400, 40), "inner_tolerance"),
                       (_lib.Cg2dMixedConfig(1, 1e-5, 0, 40), "inner_max_iters"),
                       (_lib.Cg2dMixedConfig(-1, 1e-5, 400, 40), "op_slots")):
        assert L.mj_cg2d_mixed_workspace_bytes(ctypes.byref(cfg), ctypes.byref(mbad), 0) == 0
        assert word in _lib.last_error(), (word, _lib.last_error())


def test_dst3_workspace_query():
    """mj_gad_dst3's workspace: the per-call tables (corner-filled mask bit planes, 2 x 151 words per (mask slot,
    tile, level); the 2-D metrics in three double pairs per tile point), with the R flux also the batch's localTij;
    mask slots as the configuration gives them (one shared set, one per member, or mask_slots with mask_index)."""
    import ctypes
    from mit_jax_amd import _lib
    L = _lib.lib()
    al = lambda v: (v + 255) // 256 * 256  # noqa: E731
    T, K, B, plane = 13, 50, 4, 98 * 98
    metrics = al(3 * T * plane * 16)
    bits = lambda slots: al(slots * T * K * 2 * 151 * 8)  # noqa: E731
    cfg = _lib.GadConfig(B, K, T * K * plane, 0, None, None)
    assert L.mj_gad_dst3_workspace_bytes(ctypes.byref(cfg), T, 0) == bits(1) + metrics
    assert L.mj_gad_dst3_workspace_bytes(ctypes.byref(cfg), T, 1) == al(B * T * K * plane * 8) + bits(1) + metrics
    per_member = _lib.GadConfig(B, K, T * K * plane, T * K * plane, None, None)
    assert L.mj_gad_dst3_workspace_bytes(ctypes.byref(per_member), T, 0) == bits(B) + metrics
    slots = _lib.GadConfig(B, K, T * K * plane, T * K * plane, None, ctypes.c_void_p(256))
    slots.mask_slots = 3
    assert L.mj_gad_dst3_workspace_bytes(ctypes.byref(slots), T, 0) == bits(3) + metrics
    assert L.mj_gad_dst3_workspace_bytes(ctypes.byref(_lib.GadConfig(0, K, 0, 0, None, None)), T, 0) == 0
    # the reverse mode's workspace holds the same tables after its replay scratch
    n = L.mj_gad_dst3_adjoint_workspace_bytes(ctypes.byref(slots), T, 0)
    assert n == al(1024 * 7 * plane * 8) + bits(3) + metrics


def test_operator_and_rhs_producers_validate_before_device_work():
    """mj_update_cg2d / mj_cg2d_rhs refuse null arguments, bad layouts and strides with a message (MJ_ERR_ARG)
    before touching the device."""
    import ctypes
    from mit_jax_amd import _lib
    L = _lib.lib()
    fs = _lib.FsParams(3600.0, 3600.0, 1.0, 1.0, 1.0, 1.0 / 1029.0, 1)
    dummy = ctypes.c_void_p(16)
    grid = _lib.SfpGrid(*([dummy] * 6), None, dummy, None, None, None)
    good = _lib.SfpConfig(2, 13, 90, 90, 4, 50, 13 * 50 * 98 * 98, 13 * 98 * 98)
    args13 = [dummy] * 8

    def rhs(g, p, c, *arrs):
        return L.mj_cg2d_rhs(g, p, c, *arrs, None)

    assert rhs(None, ctypes.byref(fs), ctypes.byref(good), *args13, dummy, dummy) == -1
    assert "null" in _lib.last_error()
    for field, value, word in (("batch", 0, "batch"), ("nr", 0, "nr"), ("OL", 0, "OL"), ("sNx", 125, "sNx"),
                               ("member_stride3d", 5, "member strides")):
        c = _lib.SfpConfig(2, 13, 90, 90, 4, 50, 13 * 50 * 98 * 98, 13 * 98 * 98)
        setattr(c, field, value)
        assert rhs(ctypes.byref(grid), ctypes.byref(fs), ctypes.byref(c), *args13, dummy, dummy) == -1, field
        assert word in _lib.last_error(), (field, _lib.last_error())
    # useRealFreshWaterFlux without EmPmR; outputs aliasing inputs; deepFac2F without kSurfC
    a = [ctypes.c_void_p(16 * (n + 1)) for n in range(10)]
    assert rhs(ctypes.byref(grid), ctypes.byref(fs), ctypes.byref(good), *a[:4], None, *a[5:8], a[8], a[9]) == -1
    assert "EmPmR" in _lib.last_error()
    assert rhs(ctypes.byref(grid), ctypes.byref(fs), ctypes.byref(good), *a[:8], a[0], a[9]) == -1
    assert "alias" in _lib.last_error()
    g2 = _lib.SfpGrid(*([dummy] * 6), None, dummy, None, None, dummy)
    assert rhs(ctypes.byref(g2), ctypes.byref(fs), ctypes.byref(good), *a[:8], a[8], a[9]) == -1
    assert "kSurfC" in _lib.last_error()
    # update_cg2d: null plan, aliasing outputs
    upd = [ctypes.c_void_p(16 * (n + 1)) for n in range(10)]
    assert L.mj_update_cg2d(None, ctypes.byref(grid), ctypes.byref(fs), ctypes.byref(good), upd[0], 0.51, *upd[1:4],
                            *upd[4:10], None) == -1
    assert "plan" in _lib.last_error()
    bad_fs = _lib.FsParams(0.0, 3600.0, 1.0, 1.0, 1.0, 1.0, 0)
    assert L.mj_update_cg2d(None, ctypes.byref(grid), ctypes.byref(bad_fs), ctypes.byref(good), upd[0], 0.51,
                            *upd[1:4], *upd[4:10], None) == -1
    assert "deltaTMom" in _lib.last_error()


def test_implicit_r_validates_before_device_work():
    """mj_gad_implicit_r refuses null arguments, bad sizes and aliasing with a message (MJ_ERR_ARG) before touching
    the device; the Python op refuses the branches the reference does not port."""
    import ctypes
    from types import SimpleNamespace
    from mit_jax_amd import _lib, tracer_implicit as TI
    L = _lib.lib()
    d = [ctypes.c_void_p(16 * (n + 1)) for n in range(8)]
    grid = _lib.ImprGrid(d[0], d[1], d[2], d[3], d[4], d[5], -1.0)
    n3 = 13 * 50 * 98 * 98

    def call(cfg, *arrs):
        return L.mj_gad_implicit_r(ctypes.byref(grid), ctypes.byref(cfg), *arrs, None)

    good = _lib.ImprConfig(2, 13, 90, 90, 4, 50, n3, 0, 1)
    assert L.mj_gad_implicit_r(None, ctypes.byref(good), d[6], d[7], d[6], d[7], d[6], None) == -1
    for field, value, word in (("nr", 151, "nr"), ("batch", 0, "batch"), ("member_stride3d", 5, "member strides")):
        c = _lib.ImprConfig(2, 13, 90, 90, 4, 50, n3, 0, 1)
        setattr(c, field, value)
        assert call(c, d[6], d[7], d[5], d[4], d[3]) == -1, field
        assert word in _lib.last_error(), (field, _lib.last_error())
    assert call(good, d[6], d[7], None, d[4], d[3]) == -1
    assert "wFld" in _lib.last_error()
    assert call(good, d[6], d[7], d[5], d[4], d[6]) == -1
    assert "alias" in _lib.last_error()
    p = SimpleNamespace(implicitDiffusion=False, dTtracerLev=[3600.0], rkSign=-1.0)
    with pytest.raises(NotImplementedError):
        TI.gad_implicit_r(p, None, True, 3, None, None, None, None)
    with pytest.raises(NotImplementedError):
        TI.gad_implicit_r(SimpleNamespace(implicitDiffusion=True), None, True, 4, None, None, None, None)


def test_unported_branches_raise():
    """As the reference raises at setup for the branches it does not port (gad.py:159-182, cg2d.py:212-213)."""
    from mit_jax_amd import cg2d, gad
    topo = gad.TileTopology.from_facets()
    with pytest.raises(NotImplementedError):
        gad.GADParams(topo, advectionScheme=33)
    with pytest.raises(NotImplementedError):
        gad.GADParams(topo, useCubedSphereExchange=False)
    with pytest.raises(NotImplementedError):
        cg2d.cg2d_fortran(None, None, None, None, cg2dNorm=1.0, tolerance=1e-7, max_iters=10, sum_order="fortran",
                          nIterMin=0)


def test_ops_refuse_cpu_tensors():
    """No CPU fallback: host tensors are rejected before any work."""
    import torch
    from mit_jax_amd import cg2d, gad
    from mit_jax_amd.exchange import Exchanger
    ex = Exchanger()
    a = torch.zeros((13, 98, 98), dtype=torch.float64)
    with pytest.raises(TypeError):
        ex.exch_xy(a)
    with pytest.raises(TypeError):
        cg2d.cg2d_fortran(ex, [a] * 6, a, a, cg2dNorm=1.0, tolerance=1e-7, max_iters=10, sum_order="fortran")
    with pytest.raises(TypeError):
        gad.advect_tiles(gad.GADParams.llc90(), None, {}, {}, a, a, a, a, a, a)


def test_synthetic_inputs_shape_the_configs():
    """The synthetic workload has what the configs ask for: ~30 % land in every facet, an SPD-like operator
    (negative definite diagonal, symmetric couplings checked in test_oracle_cg2d), max CFL 0.3, tracer profile."""
    from mit_jax_amd import synthetic as S
    from tests.conftest import dst3_case
    g, u, v, w, tr = dst3_case(0)
    J = I = slice(4, 94)
    land = 1.0 - g.maskInC[:, J, I]
    for tiles in ((0, 1, 2), (3, 4, 5), (6,), (7, 8, 9), (10, 11, 12)):
        assert 0.2 < land[list(tiles)].mean() < 0.4
    cfl = max(np.abs(u * S.DT * g.recip_dxC[:, None]).max(), np.abs(v * S.DT * g.recip_dyC[:, None]).max(),
              np.abs(w * S.DT * g.recip_drC[None, :, None, None]).max())
    assert abs(cfl - 0.3) < 1e-12
    assert g.depth.max() <= 6000.0 and g.drF.sum() >= 6000.0
    wet = g.maskC[:, 0, J, I] > 0
    assert abs(tr[:, 0, J, I][wet].mean() - (2 + 25 * np.exp(-(-g.rC[0]) / 800.0))) < 0.01
