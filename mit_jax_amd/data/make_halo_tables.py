"""Write `exch2_halo_13x90x90.npz`, the compact exch2 halo tables, from the probed exchange maps.

The source is the reference's data file `mitgcm_jax/data/exch_maps_13x90x90.npz` (read by
`mitgcm_jax/parallel/exchange.py:86-92`): per map name (T, 3D, Z, SMs and u/v maps of the vector kinds) three
arrays over all 13*98*98 padded points, `src` (int32 flat source index), `comp` (int8: 0 keep own value, 1 from the
u-like array, 2 from the v-like array) and `sign` (int8 +-1), plus `layout`.

Every point with comp == 0 keeps its own value (src == own index, sign == +1: checked below), so the maps are fully
described by the written points alone. The compact file keeps, per map name, `dst`, `src` (int32), `comp`, `sign`
(int8) of the written points in increasing `dst` order, the layout, and the sha256 of the source file.

Usage: python mit_jax_amd/data/make_halo_tables.py [path/to/exch_maps_13x90x90.npz]
"""

import hashlib
import sys
from pathlib import Path

import numpy as np

DEFAULT_SRC = "/root/reference/mitgcm_jax/data/exch_maps_13x90x90.npz"
SOURCE_SHA256 = "a31f0d49754259c3b1304c42f6734886605c0f4f1f72257b0654e8486a9d84fd"
OUT = Path(__file__).resolve().parent / "exch2_halo_13x90x90.npz"


def compact(path):
    raw = Path(path).read_bytes()
    sha = hashlib.sha256(raw).hexdigest()
    if sha != SOURCE_SHA256:
        raise ValueError(f"{path}: sha256 {sha} is not the known map file {SOURCE_SHA256}")
    z = np.load(path, allow_pickle=False)
    layout = np.asarray(z["layout"], np.int64)
    names = sorted({k.rsplit(".", 1)[0] for k in z.files if k != "layout"})
    out = {"layout": layout, "source_sha256": np.array(sha)}
    for n in names:
        src = np.asarray(z[n + ".src"], np.int64)
        comp = np.asarray(z[n + ".comp"], np.int8)
        sign = np.asarray(z[n + ".sign"], np.int8)
        own = np.arange(src.size)
        keep = comp == 0
        if not (np.all(src[keep] == own[keep]) and np.all(sign[keep] == 1)):
            raise ValueError(f"{n}: a kept point is not the identity")
        if not np.all(np.isin(sign, (-1, 1))) or not np.all(np.isin(comp, (0, 1, 2))):
            raise ValueError(f"{n}: unexpected comp/sign values")
        dst = np.nonzero(~keep)[0]
        out[n + ".dst"] = dst.astype(np.int32)
        out[n + ".src"] = src[dst].astype(np.int32)
        out[n + ".comp"] = comp[dst]
        out[n + ".sign"] = sign[dst]
    return out


def main(argv):
    path = argv[1] if len(argv) > 1 else DEFAULT_SRC
    arrs = compact(path)
    np.savez_compressed(OUT, **arrs)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main(sys.argv)
