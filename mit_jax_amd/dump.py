"""jaxdump per-substep dump files: the Fortran oracle's output format that the reference's bitwise gates replay
(reference io/dump.py:1-40, reference/jaxdump/jaxdump.F).

Format (the Fortran tool's, kept as is): a file `jd_<iter %010d>_t<tile %04d>.bin` is a sequence of big-endian
records, each a fixed header (`HEADER`: int32 magic 1245990224, version, iter, seq; char32 stage; char32 field;
char4 kind; int32 nz, sNx, sNy, OLx, OLy, tile, face, tBasex, tBasey) followed by nz x (sNy + 2 OLy) x (sNx + 2 OLx)
float64 values, halos included.

Design. `DumpIndex(dir)` scans every file once through a read-only memory map and keeps one row per record in a
numpy structured table (`ROW`: iter, seq, tile, face, tBase, nz/ny/nx, halo widths, file number, byte offset of
the values, and the stage/field/kind strings). Nothing else is held: values are a view of the file's memory map,
converted to native float64 when asked for. `keys()` lists (iter, stage, field) in the model's call order, the
header's own (iter, seq); `values()` gives one tile's record and `padded()` assembles [T, (nz,) ny, nx] in this
package's layout (array tile t = exch2 tile t + 1). `vertical()` unpacks the packed 1-D record of the geometry
stage (jaxdump.F JAXDUMP_GEOM group V: row r of a 2-D record holds one vertical array from column OLx on).
`append_record()` writes the same format (synthetic dumps for tests; dumps of GPU results).
"""

import re
from pathlib import Path

import numpy as np

MAGIC = 1245990224
VERSION = 1
HEADER = np.dtype([("magic", ">i4"), ("version", ">i4"), ("iter", ">i4"), ("seq", ">i4"),
                   ("stage", "S32"), ("field", "S32"), ("kind", "S4"),
                   ("nz", ">i4"), ("snx", ">i4"), ("sny", ">i4"), ("olx", ">i4"), ("oly", ">i4"),
                   ("tile", ">i4"), ("face", ">i4"), ("tbx", ">i4"), ("tby", ">i4")])
# packed vertical record of the geometry stage: 1-based row -> (name, length - Nr); only the rows this package
# reads (GAD and the CG2D producers): drF, drC, rC, rF, recip_drF, recip_drC, rhoFacC
VERTICAL_ROWS = {1: ("drF", 0), 2: ("drC", 1), 3: ("rC", 0), 4: ("rF", 1), 5: ("recip_drF", 0),
                 6: ("recip_drC", 1), 11: ("rhoFacC", 0)}

ROW = np.dtype([("iter", np.int32), ("seq", np.int32), ("tile", np.int32), ("face", np.int32), ("tbx", np.int32),
                ("tby", np.int32), ("nz", np.int32), ("ny", np.int32), ("nx", np.int32), ("olx", np.int32),
                ("oly", np.int32), ("file", np.int32), ("offset", np.int64),
                ("stage", "U32"), ("field", "U32"), ("kind", "U4")])
FILE_RE = re.compile(r"jd_(\d{10})_t(\d{4})\.bin")


def file_name(directory, it, tile):
    return Path(directory) / f"jd_{it:010d}_t{tile:04d}.bin"


def scan_file(path, file_no=0):
    """The index rows of one dump file (headers only), in file order. Raises ValueError on a bad magic number or
    a truncated record."""
    path = Path(path)
    size = path.stat().st_size
    rows = []
    if size == 0:
        return np.zeros(0, ROW)
    mm = np.memmap(path, dtype=np.uint8, mode="r")
    pos = 0
    while pos < size:
        if pos + HEADER.itemsize > size:
            raise ValueError(f"{path}: truncated header at byte {pos}")
        h = mm[pos:pos + HEADER.itemsize].view(HEADER)[0]
        if int(h["magic"]) != MAGIC:
            raise ValueError(f"{path}: bad magic number at byte {pos}")
        nz, ny, nx = int(h["nz"]), int(h["sny"]) + 2 * int(h["oly"]), int(h["snx"]) + 2 * int(h["olx"])
        start = pos + HEADER.itemsize
        end = start + 8 * nz * ny * nx
        if end > size:
            raise ValueError(f"{path}: record {h['stage'].decode().strip()}/{h['field'].decode().strip()} truncated")
        rows.append((int(h["iter"]), int(h["seq"]), int(h["tile"]), int(h["face"]), int(h["tbx"]), int(h["tby"]), nz,
                     ny, nx, int(h["olx"]), int(h["oly"]), file_no, start, h["stage"].decode().strip(),
                     h["field"].decode().strip(), h["kind"].decode().strip()))
        pos = end
    del mm
    return np.array(rows, ROW)


class DumpIndex:
    """Every record under a jaxdump directory (see the module docstring)."""

    def __init__(self, directory):
        self.dir = Path(directory)
        self.files = sorted(f for f in self.dir.glob("jd_*_t*.bin") if FILE_RE.fullmatch(f.name))
        if not self.files:
            raise FileNotFoundError(f"no jaxdump files in {self.dir}")
        parts = [scan_file(f, n) for n, f in enumerate(self.files)]
        self.table = np.concatenate(parts) if parts else np.zeros(0, ROW)
        self._where = {}
        for r, row in enumerate(self.table):
            self._where.setdefault((int(row["iter"]), str(row["stage"]), str(row["field"])), {})[int(row["tile"])] = r
        self._maps = {}

    def keys(self):
        """(iter, stage, field) of every dumped field, in the model's call order (iter, then the header's seq)."""
        first = {}
        for k, tiles in self._where.items():
            first[k] = min(int(self.table[r]["seq"]) for r in tiles.values())
        return sorted(first, key=lambda k: (k[0], first[k]))

    def tiles(self, it, stage, field):
        """Sorted exch2 tile numbers (1-based) holding the field."""
        return sorted(self._where[(it, stage, field)])

    def header(self, it, stage, field, tile):
        """The index row of one tile's record (face, tbx, tby, kind, nz, halo widths, ...)."""
        return self.table[self._where[(it, stage, field)][tile]]

    def _map(self, file_no):
        mm = self._maps.get(file_no)
        if mm is None:
            mm = self._maps[file_no] = np.memmap(self.files[file_no], dtype=np.uint8, mode="r")
        return mm

    def values(self, it, stage, field, tile):
        """One tile's record as native float64 [nz, ny, nx] (halos included)."""
        row = self.header(it, stage, field, tile)
        nz, ny, nx, off = int(row["nz"]), int(row["ny"]), int(row["nx"]), int(row["offset"])
        raw = self._map(int(row["file"]))[off:off + 8 * nz * ny * nx]
        return raw.view(">f8").astype(np.float64).reshape(nz, ny, nx)

    def padded(self, it, stage, field, layout, squeeze=True):
        """The field in the device layout [T, nz, ny, nx] ([T, ny, nx] when nz = 1 and squeeze); tiles that were
        not dumped are NaN."""
        where = self._where[(it, stage, field)]
        nz = int(self.table[next(iter(where.values()))]["nz"])
        out = np.full((layout.nTiles, nz, layout.ny, layout.nx), np.nan)
        for tile in where:
            v = self.values(it, stage, field, tile)
            if v.shape[1:] != (layout.ny, layout.nx):
                raise ValueError(f"{field}: tile {tile} has shape {v.shape[1:]}, the layout {layout.ny, layout.nx}")
            out[tile - 1] = v
        return out[:, 0] if squeeze and nz == 1 else out

    def vertical(self, it, layout, stage="G00_geometry", field="vertical"):
        """The packed vertical record (read from exch2 tile 1): {name: 1-D array} for VERTICAL_ROWS."""
        v = self.values(it, stage, field, 1)[0]
        out = {}
        for row, (name, extra) in VERTICAL_ROWS.items():
            out[name] = v[row - 1 + layout.OLy, layout.OLx:layout.OLx + layout.Nr + extra].copy()
        return out

    def close(self):
        self._maps.clear()


def pack_vertical(layout, arrays):
    """The packed vertical record [ny, nx] holding `arrays` ({name: 1-D}) at their VERTICAL_ROWS rows."""
    v = np.zeros((layout.ny, layout.nx))
    rows = {name: (row, extra) for row, (name, extra) in VERTICAL_ROWS.items()}
    for name, a in arrays.items():
        row, extra = rows[name]
        a = np.asarray(a, np.float64).reshape(-1)
        if a.size != layout.Nr + extra:
            raise ValueError(f"{name}: {a.size} values, the record row holds {layout.Nr + extra}")
        v[row - 1 + layout.OLy, layout.OLx:layout.OLx + a.size] = a
    return v


def append_record(fh, it, seq, stage, field, kind, tile, face, tbx, tby, values, olx=4, oly=4):
    """Append one record (values [nz, ny, nx] or [ny, nx], halos included) to a binary file open for writing."""
    v = np.asarray(values, np.float64)
    v = v[None] if v.ndim == 2 else v
    nz, ny, nx = v.shape
    h = np.zeros(1, HEADER)
    for k, val in (("magic", MAGIC), ("version", VERSION), ("iter", it), ("seq", seq), ("nz", nz),
                   ("snx", nx - 2 * olx), ("sny", ny - 2 * oly), ("olx", olx), ("oly", oly), ("tile", tile),
                   ("face", face), ("tbx", tbx), ("tby", tby)):
        h[k] = val
    h["stage"], h["field"], h["kind"] = stage.ljust(32).encode(), field.ljust(32).encode(), kind.ljust(4).encode()
    fh.write(h.tobytes())
    fh.write(v.astype(">f8").tobytes())
