"""On-disk formats either side of the path (CPU): the exch2 map file in the reference's format (reference
parallel/exchange.py:72-92) and jaxdump dump files (reference io/dump.py:1-40): round trips, the reference's own map
file where mounted, and the errors a damaged file raises."""

from pathlib import Path

import numpy as np
import pytest

from mit_jax_amd import dump as D
from mit_jax_amd.exchange import ExchangeMaps
from mit_jax_amd.layout import LLC90
from tests.conftest import maps

REF_NPZ = Path("/root/reference/mitgcm_jax/data/exch_maps_13x90x90.npz")


def test_map_file_round_trip(tmp_path):
    p = tmp_path / "maps.npz"
    maps().save_full(p)
    back = ExchangeMaps.from_full_npz(p)
    assert back.layout == maps().layout and back.names() == maps().names()
    for n in maps().names():
        for a, b in zip(back.compact[n], maps().compact[n]):
            assert np.array_equal(a, b), n


@pytest.mark.skipif(not REF_NPZ.exists(), reason="reference map file not mounted")
def test_map_file_of_the_reference_loads_to_the_committed_tables():
    back = ExchangeMaps.from_full_npz(REF_NPZ)
    for n in maps().names():
        for a, b in zip(back.compact[n], maps().compact[n]):
            assert np.array_equal(a, b), n


def test_map_file_errors(tmp_path):
    src, comp, sign = maps().full("T")
    src = src.copy()
    keep = np.nonzero(comp == 0)[0][0]
    src[keep] = (src[keep] + 1) % src.size
    L = maps().layout
    np.savez(tmp_path / "bad.npz", layout=np.array([L.sNx, L.sNy, L.OLx, L.OLy, L.nTiles, L.Nr]),
             **{"T.src": src, "T.comp": comp, "T.sign": sign})
    with pytest.raises(ValueError, match="identity"):
        ExchangeMaps.from_full_npz(tmp_path / "bad.npz")
    np.savez(tmp_path / "short.npz", layout=np.array([L.sNx, L.sNy, L.OLx, L.OLy, L.nTiles, L.Nr]),
             **{"T.src": src[:-1], "T.comp": comp[:-1], "T.sign": sign[:-1]})
    with pytest.raises(ValueError, match="cover"):
        ExchangeMaps.from_full_npz(tmp_path / "short.npz")


FACE = [1, 1, 1, 2, 2, 2, 3, 4, 4, 4, 5, 5, 5]


def _write_dump(d, fields, its=(0, 72)):
    """fields: {(stage, field): array [T, (nz,) ny, nx]} written for every iteration (values + it)."""
    for it in its:
        for t in range(LLC90.nTiles):
            with open(D.file_name(d, it, t + 1), "wb") as fh:
                for seq, ((stage, field), a) in enumerate(fields.items()):
                    D.append_record(fh, it, seq, stage, field, "RL", t + 1, FACE[t], 90 * (t % 3), 0, a[t] + it)


def test_dump_round_trip_and_index(tmp_path):
    rng = np.random.default_rng(0)
    T, K, ny, nx = LLC90.nTiles, 5, LLC90.ny, LLC90.nx
    fields = {("S05_dynamics", "gU"): rng.normal(size=(T, K, ny, nx)), ("S05_dynamics", "etaN"):
              rng.normal(size=(T, ny, nx)), ("C01_cg2d_inputs", "cg2d_b"): rng.normal(size=(T, ny, nx))}
    _write_dump(tmp_path, fields)
    ds = D.DumpIndex(tmp_path)
    assert ds.keys() == [(it, st, f) for it in (0, 72) for (st, f) in fields]
    assert len(ds.table) == 2 * T * len(fields)
    for it in (0, 72):
        for (st, f), a in fields.items():
            got = ds.padded(it, st, f, LLC90)
            assert np.array_equal(got, a + it), (it, st, f)
            assert ds.tiles(it, st, f) == list(range(1, T + 1))
            h = ds.header(it, st, f, 8)  # exch2 tile 8 = array tile 7, facet 4
            assert (int(h["face"]), int(h["tbx"]), int(h["olx"]), int(h["oly"]), str(h["kind"])) == (4, 90, 4, 4, "RL")
            v = ds.values(it, st, f, 8)
            assert v.dtype == np.float64 and np.array_equal(v, (a[7] + it).reshape(v.shape))
    # the call order is the header's seq, not the file order of the records
    d2 = tmp_path / "reordered"
    d2.mkdir()
    with open(D.file_name(d2, 5, 1), "wb") as fh:
        D.append_record(fh, 5, 2, "B", "x", "RL", 1, 1, 0, 0, np.zeros((ny, nx)))
        D.append_record(fh, 5, 1, "A", "y", "RL", 1, 1, 0, 0, np.ones((ny, nx)))
    assert D.DumpIndex(d2).keys() == [(5, "A", "y"), (5, "B", "x")]


def test_dump_vertical_record(tmp_path):
    """The packed vertical record of the geometry stage: each array at its row, from column OLx on."""
    rng = np.random.default_rng(1)
    Nr = LLC90.Nr
    arrs = {"drF": rng.random(Nr), "recip_drF": rng.random(Nr), "rhoFacC": rng.random(Nr),
            "drC": rng.random(Nr + 1), "recip_drC": rng.random(Nr + 1), "rC": rng.random(Nr), "rF": rng.random(Nr + 1)}
    v = D.pack_vertical(LLC90, arrs)
    assert v[LLC90.OLy + 10, LLC90.OLx + 3] == arrs["rhoFacC"][3]   # row 11 (1-based)
    assert v[LLC90.OLy, LLC90.OLx + Nr - 1] == arrs["drF"][Nr - 1]   # row 1
    stack = np.broadcast_to(v, (LLC90.nTiles,) + v.shape)
    _write_dump(tmp_path, {("G00_geometry", "vertical"): stack}, its=(0,))
    got = D.DumpIndex(tmp_path).vertical(0, LLC90)
    for k, a in arrs.items():
        assert np.array_equal(got[k], a), k
    with pytest.raises(ValueError, match="holds"):
        D.pack_vertical(LLC90, {"drF": np.ones(Nr + 1)})


def test_dump_errors(tmp_path):
    a = np.zeros((LLC90.nTiles, LLC90.ny, LLC90.nx))
    _write_dump(tmp_path, {("G00_geometry", "rA"): a}, its=(0,))
    f = D.file_name(tmp_path, 0, 1)
    raw = bytearray(f.read_bytes())
    raw[0] ^= 0xFF
    f.write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="magic"):
        D.scan_file(f)
    g = D.file_name(tmp_path, 0, 2)
    g.write_bytes(g.read_bytes()[:-8])
    with pytest.raises(ValueError, match="truncated"):
        D.scan_file(g)
    with pytest.raises(FileNotFoundError):
        D.DumpIndex(tmp_path / "none")
