"""Oracle restatement of the GAD_ADVECTION per-tile tables — TEST INFRASTRUCTURE ONLY.

Restates, in numpy, the reference's exch2 tile topology (pkgs/gad.py:75-106, w2_set_map_tiles.F /
w2_set_tile2tiles.F), the facet pass table (gad.py:217-235, gad_advection.F:333-351), the FILL_CS_CORNER_TR_RL and
FILL_CS_CORNER_UV_RS point pairs (gad.py:241-281) and the per-pass update regions and fill tables (gad.py:330-383,
gad_advection.F:376-774), in the reference's table form: boolean update masks and flat per-tile gather tables.
The library builds the same information as update boxes and in-LDS fills in C++ (mj_gad_plan_create); the tests
compare the two.
"""

import numpy as np

LLC90_FACETS = (90, 270, 90, 270, 90, 90, 270, 90, 270, 90)  # dimsFacets of the 13-tile LLC90 exch2 setup


class Topology:
    """exch2_myFace and facet-edge flags per tile (tile t = exch2 tile t+1)."""

    def __init__(self, face, N, S, E, W):
        self.face, self.N, self.S, self.E, self.W = tuple(face), tuple(N), tuple(S), tuple(E), tuple(W)

    @classmethod
    def from_facets(cls, dims, sNx=90, sNy=90):
        # tiles numbered facet by facet, rows of tiles (ty) outer, tx inner; an edge flag is set where the tile
        # edge lies on the facet edge (gad.py:85-106)
        face, N, S, E, W = [], [], [], [], []
        for f in range(len(dims) // 2):
            fNx, fNy = dims[2 * f], dims[2 * f + 1]
            for ty in range(fNy // sNy):
                for tx in range(fNx // sNx):
                    face.append(f + 1)
                    S.append(ty == 0)
                    N.append((ty + 1) * sNy >= fNy)
                    W.append(tx == 0)
                    E.append((tx + 1) * sNx >= fNx)
        return cls(face, N, S, E, W)

    def corners(self, t):
        """(sw, se, nw, ne) of tile t (fill_cs_corner_tr_rl.F:78-85)."""
        return (self.W[t] and self.S[t], self.E[t] and self.S[t], self.W[t] and self.N[t], self.E[t] and self.N[t])

    def edges_u8(self):
        """[T, 4] uint8: N, S, E, W."""
        return np.array([[self.N[t], self.S[t], self.E[t], self.W[t]] for t in range(len(self.face))], np.uint8)


def pass_flags(face, ipass):
    """(overlapOnly, interiorOnly, calc X, calc Y) of pass ipass (1..3) for facet `face` (gad_advection.F:333-351)."""
    if ipass == 1:
        return face % 3 == 0, face % 3 != 0, face in (6, 1, 2), face in (3, 4, 5)
    if ipass == 2:
        return face % 3 == 2, face % 3 == 1, face in (2, 3, 4), face in (5, 6, 1)
    return False, True, face in (5, 6), face in (2, 3)


def fill_tr_pairs(fill4dir, corners, sNx, sNy, OL):
    """FILL_CS_CORNER_TR_RL(fill4dir, withSigns=F) as [(dst (i, j), src (i, j))] in Fortran indices."""
    sw, se, nw, ne = corners
    out = []
    for j in range(1, OL + 1):
        for i in range(1, OL + 1):
            if fill4dir == 1:
                cand = ((sw, (1 - i, 1 - j), (1 - j, i)), (se, (sNx + i, 1 - j), (sNx + j, i)),
                        (nw, (1 - i, sNy + j), (1 - j, sNy + 1 - i)), (ne, (sNx + i, sNy + j), (sNx + j, sNy + 1 - i)))
            else:
                cand = ((sw, (1 - i, 1 - j), (j, 1 - i)), (se, (sNx + i, 1 - j), (sNx + 1 - j, 1 - i)),
                        (nw, (1 - i, sNy + j), (j, sNy + i)), (ne, (sNx + i, sNy + j), (sNx + 1 - j, sNy + i)))
            out += [(d, s) for on, d, s in cand if on]
    return out


def fill_uv_pairs(corners, sNx, sNy, OL):
    """FILL_CS_CORNER_UV_RS(withSigns=F) as [(dst comp, dst (i, j), src comp, src (i, j))], comp 0 = u, 1 = v."""
    sw, se, nw, ne = corners
    R = range(1, OL + 1)
    R2 = range(2, OL + 1)
    out = []
    if sw:
        out += [(0, (1 - i, 1 - j), 1, (1 - j, 1 + i)) for j in R for i in R]
        out += [(1, (1 - i, 1 - j), 0, (1 + j, 1 - i)) for j in R for i in R]
    if se:
        out += [(0, (sNx + i, 1 - j), 1, (sNx + j, i)) for j in R for i in R2]
        out += [(1, (sNx + i, 1 - j), 0, (sNx + 1 - j, 1 - i)) for j in R for i in R]
    if nw:
        out += [(0, (1 - i, sNy + j), 1, (1 - j, sNy + 1 - i)) for j in R for i in R]
        out += [(1, (1 - i, sNy + j), 0, (j, sNy + i)) for j in R2 for i in R]
    if ne:
        out += [(0, (sNx + i, sNy + j), 1, (sNx + j, sNy + 2 - i)) for j in R for i in R2]
        out += [(1, (sNx + i, sNy + j), 0, (sNx + 2 - j, sNy + i)) for j in R2 for i in R]
    return out


def tables(topo, sNx=90, sNy=90, OL=4):
    """Per-tile tables of GAD_ADVECTION. Returns dict with
    masks   uint8 [3, 2, T, ny*nx]: update region of the X (dir 0) / Y (dir 1) block of each pass;
    fills   list of 12 (pass*4 + {fXb, fXa, fYb, fYa}) entries: [T, ny*nx] int32 gather table, or None when no
            tile calls that fill (identity rows for tiles that do not);
    fill_dirs int32 [3, 4, T]: the fill4dir of each call (0 none), for comparison with the library plan;
    uvfill_u, uvfill_v int32 [T, ny*nx]: indices into concat(maskW plane, maskS plane)."""
    T = len(topo.face)
    ny, nx = sNy + 2 * OL, sNx + 2 * OL
    n = ny * nx

    def flat(i, j):
        return (j - 1 + OL) * nx + (i - 1 + OL)

    def box(m, ilo, ihi, jlo, jhi):
        m[(jlo - 1 + OL):(jhi + OL), (ilo - 1 + OL):(ihi + OL)] = 1

    masks = np.zeros((3, 2, T, ny, nx), np.uint8)
    fill_dirs = np.zeros((3, 4, T), np.int32)
    for p in range(1, 4):
        for t in range(T):
            N, S, E, W = topo.N[t], topo.S[t], topo.E[t], topo.W[t]
            ov, io, cX, cY = pass_flags(topo.face[t], p)
            if cX:
                doflux = (not ov) or N or S
                if doflux and ov:
                    fill_dirs[p - 1, 0, t] = 1
                    if p == 1:
                        fill_dirs[p - 1, 1, t] = 2
                mX = masks[p - 1, 0, t]
                if ov:
                    ilo = 1 if W else 2 - OL
                    ihi = sNx if E else sNx + OL - 1
                    if S:
                        box(mX, ilo, ihi, 1 - OL, 0)
                    if N:
                        box(mX, ilo, ihi, sNy + 1, sNy + OL)
                else:
                    jlo = 1 if (io and S) else 1 - OL
                    jhi = sNy if (io and N) else sNy + OL
                    box(mX, 2 - OL, sNx + OL - 1, jlo, jhi)
            if cY:
                doflux = (not ov) or E or W
                if doflux and ov:
                    fill_dirs[p - 1, 2, t] = 2
                    if p == 1:
                        fill_dirs[p - 1, 3, t] = 1
                mY = masks[p - 1, 1, t]
                if ov:
                    jlo = 1 if S else 2 - OL
                    jhi = sNy if N else sNy + OL - 1
                    if W:
                        box(mY, 1 - OL, 0, jlo, jhi)
                    if E:
                        box(mY, sNx + 1, sNx + OL, jlo, jhi)
                else:
                    ilo = 1 if (io and W) else 1 - OL
                    ihi = sNx if (io and E) else sNx + OL
                    box(mY, ilo, ihi, 2 - OL, sNy + OL - 1)
    fills = []
    ident = np.arange(n, dtype=np.int32)
    for p in range(3):
        for f in range(4):
            dirs = fill_dirs[p, f]
            if not dirs.any():
                fills.append(None)
                continue
            tab = np.tile(ident, (T, 1))
            for t in range(T):
                if dirs[t]:
                    pairs = fill_tr_pairs(int(dirs[t]), topo.corners(t), sNx, sNy, OL)
                    dsts = {flat(*d) for d, _ in pairs}
                    assert not dsts & {flat(*s) for _, s in pairs}
                    for d, s in pairs:
                        tab[t, flat(*d)] = flat(*s)
            fills.append(tab)
    uvu = np.tile(ident, (T, 1))
    uvv = np.tile(ident + n, (T, 1))
    for t in range(T):
        for c, d, cs, s in fill_uv_pairs(topo.corners(t), sNx, sNy, OL):
            (uvu if c == 0 else uvv)[t, flat(*d)] = cs * n + flat(*s)
    return dict(masks=masks.reshape(3, 2, T, n), fills=fills, fill_dirs=fill_dirs, uvfill_u=uvu, uvfill_v=uvv)
