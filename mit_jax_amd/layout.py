"""Array layout of the LLC90 tile fields and Fortran-index helpers.

Fields are stored per tile with halos, as MITgcm's `(1-OLx:sNx+OLx, 1-OLy:sNy+OLy, [Nr,] tile)` arrays but with the
tile axis first and i fastest: `[tile, j, i]` (2-D) or `[tile, k, j, i]` (3-D), float64, C-contiguous. Batched
(ensemble) fields put the member axis in front: `[member, tile, (k,) j, i]`. Tile t (0-based) is exch2 tile t+1:
tiles 0-2 facet 1, 3-5 facet 2, 6 facet 3, 7-9 facet 4, 10-12 facet 5.

Fortran index i (1-based, halo points <= 0 or > sNx) sits at array index i - 1 + OLx; level k (1..Nr) at k - 1.
Follows the reference layout (mitgcm_jax/layout.py:1-54) so that arrays pass between the two unchanged.
"""

from dataclasses import dataclass


@dataclass(frozen=True)
class Layout:
    sNx: int = 90
    sNy: int = 90
    OLx: int = 4
    OLy: int = 4
    nTiles: int = 13
    Nr: int = 50

    @property
    def nx(self):
        return self.sNx + 2 * self.OLx

    @property
    def ny(self):
        return self.sNy + 2 * self.OLy

    def is_(self, lo, hi):
        """Array slice of the Fortran i-range lo..hi (inclusive)."""
        return slice(lo - 1 + self.OLx, hi + self.OLx)

    def js(self, lo, hi):
        """Array slice of the Fortran j-range lo..hi (inclusive)."""
        return slice(lo - 1 + self.OLy, hi + self.OLy)

    def ii(self, i):
        """Array index of Fortran i."""
        return i - 1 + self.OLx

    def jj(self, j):
        return j - 1 + self.OLy

    @property
    def shape2d(self):
        return (self.nTiles, self.ny, self.nx)

    @property
    def shape3d(self):
        return (self.nTiles, self.Nr, self.ny, self.nx)

    @property
    def npad(self):
        """Points of one padded 2-D level over all tiles."""
        return self.nTiles * self.ny * self.nx

    def as_array(self):
        return [self.sNx, self.sNy, self.OLx, self.OLy, self.nTiles, self.Nr]


LLC90 = Layout()


def interior_mask(L=LLC90):
    """[ny, nx] bool: Fortran interior points i = 1..sNx, j = 1..sNy."""
    import numpy as np
    m = np.zeros((L.ny, L.nx), bool)
    m[L.js(1, L.sNy), L.is_(1, L.sNx)] = True
    return m
