"""Reader (and writer) of jaxdump per-substep dump files, the Fortran oracle's output format that the reference's
bitwise gates replay (reference io/dump.py:1-40, reference/jaxdump/jaxdump.F).

A file `<dir>/jd_<iter %010d>_t<tile %04d>.bin` is a stream of big-endian records: a header of int32 magic
(1245990224), version, iter, seq, char32 stage, char32 field, char4 kind, int32 nz, sNx, sNy, OLx, OLy, tile, face,
tBasex, tBasey, followed by float64 values[nz][sNy + 2 OLy][sNx + 2 OLx] (halos included). The format is the
Fortran tool's, so it is kept as is.

`DumpSet(dir)` indexes every record by (iter, stage, field) -> {tile: Record}; `padded(...)` assembles a field in
this package's device layout [tile, (k,) j, i] (tile t = exch2 tile t + 1, halos as dumped), ready for the GPU ops;
`write_record` writes the same format (synthetic dumps for tests, or dumps of GPU results for comparison with the
Fortran oracle's).
"""

import re
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from pathlib import Path

import numpy as np

MAGIC = 1245990224
VERSION = 1
HEADER = np.dtype([("magic", ">i4"), ("version", ">i4"), ("iter", ">i4"), ("seq", ">i4"),
                   ("stage", "S32"), ("field", "S32"), ("kind", "S4"),
                   ("nz", ">i4"), ("snx", ">i4"), ("sny", ">i4"), ("olx", ">i4"), ("oly", ">i4"),
                   ("tile", ">i4"), ("face", ">i4"), ("tbx", ">i4"), ("tby", ">i4")])
_NAME = re.compile(r"jd_(\d{10})_t(\d{4})\.bin")


@dataclass
class Record:
    """One dumped field of one tile. `data` ([nz, sNy + 2 OLy, sNx + 2 OLx] float64) is read on first use."""
    iter: int
    seq: int
    stage: str
    field: str
    kind: str
    tile: int
    face: int
    tbx: int
    tby: int
    olx: int
    oly: int
    shape: tuple
    path: str
    offset: int
    _data: np.ndarray = None

    @property
    def data(self):
        if self._data is None:
            n = int(np.prod(self.shape))
            with open(self.path, "rb") as fh:
                fh.seek(self.offset)
                buf = fh.read(8 * n)
            if len(buf) != 8 * n:
                raise ValueError(f"{self.path}: record {self.stage}/{self.field} truncated")
            self._data = np.frombuffer(buf, ">f8", count=n).reshape(self.shape).astype(np.float64)
        return self._data

    @property
    def interior(self):
        d = self.data
        return d[:, self.oly:d.shape[1] - self.oly, self.olx:d.shape[2] - self.olx]

    def drop(self):
        """Forget the loaded values (read again on the next use)."""
        self._data = None


def read_file(path, lazy=False):
    """The records of one dump file; lazy: headers only (values read on first use)."""
    path = str(path)
    size = Path(path).stat().st_size
    recs, pos = [], 0
    with open(path, "rb") as fh:
        while pos < size:
            fh.seek(pos)
            raw = fh.read(HEADER.itemsize)
            if len(raw) != HEADER.itemsize:
                raise ValueError(f"{path}: truncated header at byte {pos}")
            h = np.frombuffer(raw, HEADER, count=1)[0]
            if int(h["magic"]) != MAGIC:
                raise ValueError(f"{path}: bad magic at byte {pos}")
            pos += HEADER.itemsize
            nz, sny, snx, oly, olx = (int(h[k]) for k in ("nz", "sny", "snx", "oly", "olx"))
            shape = (nz, sny + 2 * oly, snx + 2 * olx)
            r = Record(int(h["iter"]), int(h["seq"]), h["stage"].decode().strip(), h["field"].decode().strip(),
                       h["kind"].decode().strip(), int(h["tile"]), int(h["face"]), int(h["tbx"]), int(h["tby"]),
                       olx, oly, shape, path, pos)
            if not lazy:
                r.data  # noqa: B018 (read now)
            recs.append(r)
            pos += 8 * int(np.prod(shape))
    if pos != size:
        raise ValueError(f"{path}: last record truncated")
    return recs


def write_record(fh, it, seq, stage, field, kind, tile, face, tbx, tby, values, olx=4, oly=4):
    """Append one record (values [nz, ny, nx] or [ny, nx], halos included) to an open binary file."""
    v = np.asarray(values, np.float64)
    if v.ndim == 2:
        v = v[None]
    nz, ny, nx = v.shape
    h = np.zeros(1, HEADER)
    h["magic"], h["version"], h["iter"], h["seq"] = MAGIC, VERSION, it, seq
    h["stage"], h["field"], h["kind"] = stage.ljust(32).encode(), field.ljust(32).encode(), kind.ljust(4).encode()
    h["nz"], h["snx"], h["sny"], h["olx"], h["oly"] = nz, nx - 2 * olx, ny - 2 * oly, olx, oly
    h["tile"], h["face"], h["tbx"], h["tby"] = tile, face, tbx, tby
    fh.write(h.tobytes())
    fh.write(v.astype(">f8").tobytes())


def dump_name(directory, it, tile):
    return Path(directory) / f"jd_{it:010d}_t{tile:04d}.bin"


class DumpSet:
    """Every record under a dump directory, indexed by (iter, stage, field) -> {tile: Record}; keys in first-seen
    order per iteration (the call order of the dumped model). Headers are read in parallel threads; the index does
    not depend on their number."""

    def __init__(self, directory, threads=16):
        self.dir = Path(directory)
        files = sorted(f for f in self.dir.glob("jd_*_t*.bin") if _NAME.fullmatch(f.name))
        if not files:
            raise FileNotFoundError(f"no jaxdump files in {self.dir}")
        with ThreadPoolExecutor(max(1, min(threads, len(files)))) as pool:
            headers = dict(zip(files, pool.map(lambda f: read_file(f, lazy=True), files)))
        per_iter = {}
        for f in files:
            per_iter.setdefault(int(_NAME.fullmatch(f.name).group(1)), []).append(f)
        self.index, self.order = {}, []
        for it in sorted(per_iter):
            for n, f in enumerate(sorted(per_iter[it])):
                for r in headers[f]:
                    key = (r.iter, r.stage, r.field)
                    if key not in self.index:
                        self.index[key] = {}
                        if n == 0:
                            self.order.append(key)
                    self.index[key][r.tile] = r
        seen = set(self.order)
        self.order += [k for k in self.index if k not in seen]

    def keys(self):
        return list(self.order)

    def tiles(self, it, stage, field):
        return self.index[(it, stage, field)]

    def padded(self, it, stage, field, layout, squeeze=True):
        """The field in the device layout [T, nz, ny, nx] ([T, ny, nx] when nz = 1 and squeeze); tiles that were
        not dumped are NaN."""
        recs = self.index[(it, stage, field)]
        nz = next(iter(recs.values())).shape[0]
        out = np.full((layout.nTiles, nz, layout.ny, layout.nx), np.nan)
        for r in recs.values():
            if r.shape[1:] != (layout.ny, layout.nx):
                raise ValueError(f"{field}: tile {r.tile} has shape {r.shape[1:]}, layout {layout.ny, layout.nx}")
            out[r.tile - 1] = r.data
        return out[:, 0] if squeeze and nz == 1 else out
