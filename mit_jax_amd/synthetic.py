"""Seeded synthetic LLC90 inputs for the CG2D and DST3 workloads (13 tiles of 90x90, halo 4, 50 levels).

Setup-time host code (numpy), not part of the hot path: it builds the grid, the CG2D operator and the advection
inputs that the GPU kernels then consume. Halos are filled with the exch2 maps by `_exch_np` (a gather over the
compact tables, used only here while building inputs; the exchange op itself runs on the GPU).

Recipes (reference files the steps follow):
  grid        rA ~ U(0.5e10, 1.5e10) m^2 on tile interiors; face lengths dxG = dyG = dxC = dyC from sqrt(rA)
              averaged to faces; halos by the T / UVn exchanges as INI_CURVILINEAR_GRID does
              (grid/load.py:396-416).
  depth       a smooth random field per LLC facet (facet shapes of io/llc.py:16), 0-6000 m, thresholded per
              facet so that ~30 % of the columns of every facet are land (depth <= 0).
  hFac, masks ini_masks_etc (grid/load.py:485-594): hFacMin = 0.2, hFacMinDr = 5 m, hFacW/S = min of the
              neighbours + UVn exchange, walls where dyG/dxG = 0, masks = hFac != 0.
  operator    UPDATE_CG2D (core/solve_for_pressure.py:29-65) with implicSurfPress = implicDiv2DFlow =
              freeSurfFac = 1, recip_Bo = 1/g, dtMom = dtFS = dt; cg2dNorm from INI_CG2D (core/cg2d.py:108-134).
  flow        per-level streamfunction at cell corners (smooth per facet, amplitude decaying with depth):
              uTrans = psi(j+1) - psi(j), vTrans = -(psi(i+1) - psi(i)), masked; u = uTrans/(dyG drF hFacW);
              UVs exchange; w from continuity by the INTEGRATE_FOR_W recursion (core/free_surface.py:255-279,
              rStarDhDt = 0); all scaled so that the largest |u| dt/dxC, |v| dt/dyC, |w| dt/drC is 0.3.
  tracer      2 + 25 exp(-z/800 m) + N(0, 0.1) on wet cells, T exchange.
"""

from dataclasses import dataclass, field
from functools import lru_cache

import numpy as np

from .layout import LLC90, Layout

G = 9.81
DT = 3600.0
FACET_SHAPES = {1: (270, 90), 2: (270, 90), 3: (90, 90), 4: (90, 270), 5: (90, 270)}


# ------------------------------------------------------------------------------------------------ halo fill
def _load_compact():
    from .exchange import ExchangeMaps
    return ExchangeMaps.load()


def _exch_np(maps, kind, a, b=None):
    """Halo fill of [T, ..., ny, nx] numpy arrays (setup only): scalar kind -> array, vector kind -> (u, v)."""
    def one(mapname, own, fa, fb):
        d, s, c, g = maps.compact[mapname]
        T = own.shape[0]
        plane = own.shape[-1] * own.shape[-2]
        lead = own.shape[1:-2]
        out = own.copy()
        o = out.reshape(T, -1, plane)
        A = fa.reshape(T, -1, plane)
        B = fb.reshape(T, -1, plane) if fb is not None else A
        dt, dr = d // plane, d % plane
        st, sr = s // plane, s % plane
        va = A[st, :, sr]
        vb = B[st, :, sr]
        v = np.where((c == 2)[:, None], vb, va)
        o[dt, :, dr] = v * g[:, None].astype(np.float64)
        return out.reshape((T,) + lead + own.shape[-2:])
    if b is None:
        return one(kind, a, a, None)
    return one(kind + "_u", a, a, b), one(kind + "_v", b, a, b)


# ------------------------------------------------------------------------------------------------ facets
def facets_to_tiles(fa, L=LLC90):
    """{facet: [..., fny, fnx]} -> [..., 13, sNy, sNx] (tiles 0-2 facet 1 rows, 3-5 facet 2, 6 facet 3, 7-9
    facet 4 columns, 10-12 facet 5), as io/llc.py:39-46."""
    n = L.sNx
    t = []
    for f in (1, 2):
        t += [fa[f][..., n * k:n * (k + 1), :] for k in range(3)]
    t.append(fa[3])
    for f in (4, 5):
        t += [fa[f][..., :, n * k:n * (k + 1)] for k in range(3)]
    return np.stack(t, axis=-3)


def _smooth_field(rng, shape, nmodes=24, kmax=6.0):
    """Sum of random Fourier modes on a facet grid: smooth, zero-mean-ish, seeded."""
    ny, nx = shape
    y = (np.arange(ny) + 0.5) / 90.0
    x = (np.arange(nx) + 0.5) / 90.0
    Y, X = np.meshgrid(y, x, indexing="ij")
    f = np.zeros(shape)
    for _ in range(nmodes):
        kx, ky = rng.uniform(-kmax, kmax, 2)
        ph = rng.uniform(0, 2 * np.pi)
        amp = rng.normal() / (1.0 + kx * kx + ky * ky) ** 0.5
        f += amp * np.cos(kx * X + ky * Y + ph)
    return f


def vertical_grid(Nr=50, total=6150.0, dz_top=10.0, dz_bot=250.0):
    """Stretched layer thicknesses (tanh profile scaled to `total` m), interfaces rF (0 at the surface,
    negative downward) and centre-distances drC (drC[0] = half the first layer)."""
    k = np.arange(Nr)
    prof = dz_top + (dz_bot - dz_top) * 0.5 * (1.0 + np.tanh((k - 0.55 * Nr) / (0.16 * Nr)))
    drF = prof * (total / prof.sum())
    rF = np.concatenate([[0.0], -np.cumsum(drF)])
    rC = 0.5 * (rF[:-1] + rF[1:])
    drC = np.empty(Nr)
    drC[0] = 0.5 * drF[0]
    drC[1:] = 0.5 * (drF[:-1] + drF[1:])
    return dict(drF=drF, recip_drF=1.0 / drF, rF=rF, rC=rC, drC=drC, recip_drC=1.0 / drC)


def _recip(a):
    out = np.zeros_like(a)
    nz = a != 0.0
    out[nz] = 1.0 / a[nz]
    return out


# ------------------------------------------------------------------------------------------------ grid
@dataclass
class Grid:
    L: Layout
    f: dict = field(default_factory=dict)   # [T, ny, nx] and [T, K, ny, nx] numpy fields
    v: dict = field(default_factory=dict)   # per-level [K]

    def __getattr__(self, k):
        f = self.__dict__.get("f", {})
        if k in f:
            return f[k]
        v = self.__dict__.get("v", {})
        if k in v:
            return v[k]
        raise AttributeError(k)


def make_grid(seed=0, nlev=None, L=LLC90, land_frac=0.30, hFacMin=0.2, hFacMinDr=5.0):
    """The shared synthetic LLC90 grid. nlev (<= Nr) truncates the ocean to its first nlev levels (the member
    level counts of the ensemble workload)."""
    maps = _load_compact()
    rng = np.random.default_rng(seed)
    T, ny, nx, K = L.nTiles, L.ny, L.nx, L.Nr
    J, I = L.js(1, L.sNy), L.is_(1, L.sNx)
    vg = vertical_grid(K)
    f = {}
    rA = np.zeros(L.shape2d)
    rA[:, J, I] = rng.uniform(0.5e10, 1.5e10, (T, L.sNy, L.sNx))
    rA = _exch_np(maps, "T", rA)
    sq = np.sqrt(rA)
    # faces: west face of cell (j, i) between (j, i-1) and (j, i); south face between (j-1, i) and (j, i)
    dyG = np.zeros(L.shape2d)
    dxG = np.zeros(L.shape2d)
    dyG[:, J, I] = 0.5 * (sq[:, J, L.is_(0, L.sNx - 1)] + sq[:, J, I])
    dxG[:, J, I] = 0.5 * (sq[:, L.js(0, L.sNy - 1), I] + sq[:, J, I])
    dxC, dyC = dyG.copy(), dxG.copy()
    dxC, dyC = _exch_np(maps, "UVn", dxC, dyC)
    dyG, dxG = _exch_np(maps, "UVn", dyG, dxG)
    f.update(rA=rA, recip_rA=_recip(rA), dyG=dyG, dxG=dxG, dxC=dxC, dyC=dyC, recip_dxC=_recip(dxC),
             recip_dyC=_recip(dyC))
    # depth: smooth per facet, land fraction by threshold
    # thresholded per facet, so that every facet (the Arctic cap included) has ~land_frac land
    fac = {k: _smooth_field(rng, s) for k, s in FACET_SHAPES.items()}
    dep = {}
    for k, a in fac.items():
        thr, top = np.quantile(a, land_frac), a.max()
        dep[k] = np.where(a > thr, np.minimum(6000.0 * (a - thr) / (top - thr), 6000.0), 0.0)
    depth = np.zeros(L.shape2d)
    depth[:, J, I] = facets_to_tiles(dep, L)
    depth = _exch_np(maps, "T", depth)
    if nlev is not None:
        depth = np.minimum(depth, -vg["rF"][nlev])
    f["depth"] = depth
    # hFacC (ini_masks_etc.F:117-137 with Ro_surf = 0; the surface pass :160-182 leaves it unchanged)
    R_low = -depth
    hFacC = np.zeros(L.shape3d)
    for k in range(K):
        mn = max(hFacMin, min(hFacMinDr * vg["recip_drF"][k], 1.0))
        tmp = np.clip((vg["rF"][k] - R_low) * vg["recip_drF"][k], 0.0, 1.0)
        hFacC[:, k] = np.where(tmp < mn, np.where(tmp < mn * 0.5, 0.0, mn), tmp)
    hFacC[:, :, :, :] = np.where(depth[:, None] > 0.0, hFacC, 0.0)
    hFacW = np.zeros_like(hFacC)
    hFacS = np.zeros_like(hFacC)
    hFacW[..., 1:] = np.minimum(hFacC[..., 1:], hFacC[..., :-1])
    hFacS[..., 1:, :] = np.minimum(hFacC[..., 1:, :], hFacC[..., :-1, :])
    hFacW, hFacS = _exch_np(maps, "UVn", hFacW, hFacS)
    hFacW = np.where(dyG[:, None] == 0.0, 0.0, hFacW)
    hFacS = np.where(dxG[:, None] == 0.0, 0.0, hFacS)
    maskC = (hFacC != 0.0).astype(np.float64)
    f.update(hFacC=hFacC, hFacW=hFacW, hFacS=hFacS, recip_hFacC=_recip(hFacC), maskC=maskC,
             maskW=(hFacW != 0.0).astype(np.float64), maskS=(hFacS != 0.0).astype(np.float64),
             maskInC=(maskC.max(axis=1) > 0).astype(np.float64))
    one = np.ones(K)
    vg.update(dTlev=np.full(K, DT), rhoFacC=one, recip_rhoFacC=one, deepFacC=one, recip_deepFacC=one,
              recip_deepFac2C=one)
    return Grid(L, f, vg)


# ------------------------------------------------------------------------------------------------ CG2D inputs
def ini_cg2d_norm(g, implicSurfPress=1.0, implicDiv2DFlow=1.0):
    """cg2dNorm = 1/max|aW, aS| of the interior operator (core/cg2d.py:108-134) for the synthetic inputs: the
    per-level terms (fac ((dyG drF) hFacW)) recip_dxC summed over k in order, fac = implicSurfPress implicDiv2DFlow.
    (The GPU op is cg2d.ini_cg2d_norm.)"""
    L = g.L
    J, I = L.js(1, L.sNy), L.is_(1, L.sNx)
    fac = implicSurfPress * implicDiv2DFlow
    aW = np.zeros((L.nTiles, L.sNy, L.sNx))
    aS = np.zeros_like(aW)
    for k in range(L.Nr):
        aW = aW + (fac * ((g.dyG[:, J, I] * g.drF[k]) * g.hFacW[:, k, J, I])) * g.recip_dxC[:, J, I]
        aS = aS + (fac * ((g.dxG[:, J, I] * g.drF[k]) * g.hFacS[:, k, J, I])) * g.recip_dyC[:, J, I]
    m = max(np.abs(aW).max(), np.abs(aS).max())
    return 1.0 / m if m != 0.0 else 1.0


def update_cg2d(g, cg2dNorm, dt=DT, recip_Bo=1.0 / G, precond="5point", pcOffDFac=0.51):
    """UPDATE_CG2D (core/solve_for_pressure.py:29-65): (aW2d, aS2d, aC2d, pW, pS, pC). precond "5point" is the
    reference preconditioner; "diagonal" sets pW = pS = 0 and pC = 1/aC (config 0)."""
    maps = _load_compact()
    L = g.L
    T = L.nTiles
    J1, I1 = L.js(1, L.sNy + 1), L.is_(1, L.sNx + 1)
    J, I, Jp, Ip = L.js(1, L.sNy), L.is_(1, L.sNx), L.js(2, L.sNy + 1), L.is_(2, L.sNx + 1)
    aW = np.zeros((T, L.sNy + 1, L.sNx + 1))
    aS = np.zeros_like(aW)
    for k in range(L.Nr):
        aW = aW + g.dyG[:, J1, I1] * g.drF[k] * g.hFacW[:, k, J1, I1] * g.recip_dxC[:, J1, I1]
        aS = aS + g.dxG[:, J1, I1] * g.drF[k] * g.hFacS[:, k, J1, I1] * g.recip_dyC[:, J1, I1]
    aW = aW * cg2dNorm * 1.0 * 1.0
    aS = aS * cg2dNorm * 1.0 * 1.0
    aW2d = np.zeros(L.shape2d)
    aS2d = np.zeros(L.shape2d)
    aW2d[:, J1, I1] = aW
    aS2d[:, J1, I1] = aS
    aC2d = np.zeros(L.shape2d)
    aC2d[:, J, I] = -(aW2d[:, J, I] + aW2d[:, J, Ip] + aS2d[:, J, I] + aS2d[:, Jp, I]
                      + 1.0 * cg2dNorm * recip_Bo * g.rA[:, J, I] / dt / dt)
    aC2d = _exch_np(maps, "T", aC2d)
    pW = np.zeros(L.shape2d)
    pS = np.zeros(L.shape2d)
    pC = np.zeros(L.shape2d)
    a = aC2d[:, J1, I1]
    pC[:, J1, I1] = np.where(a == 0.0, 1.0, 1.0 / np.where(a == 0.0, 1.0, a))
    if precond == "5point":
        Im, Jm = L.is_(0, L.sNx), L.js(0, L.sNy)
        tW = a + aC2d[:, J1, Im]
        tS = a + aC2d[:, Jm, I1]
        pW[:, J1, I1] = np.where(tW == 0.0, 0.0, -aW2d[:, J1, I1] / (pcOffDFac * np.where(tW == 0.0, 1.0, tW)) ** 2)
        pS[:, J1, I1] = np.where(tS == 0.0, 0.0, -aS2d[:, J1, I1] / (pcOffDFac * np.where(tS == 0.0, 1.0, tS)) ** 2)
    elif precond != "diagonal":
        raise ValueError(precond)
    return aW2d, aS2d, aC2d, pW, pS, pC


def rhs(g, seed=1):
    """N(0, 1) on wet interior points, wet mean removed, 0 on land and halos."""
    L = g.L
    J, I = L.js(1, L.sNy), L.is_(1, L.sNx)
    rng = np.random.default_rng(seed)
    wet = g.maskInC[:, J, I] > 0
    v = rng.normal(size=wet.shape)
    v = np.where(wet, v - v[wet].mean(), 0.0)
    b = np.zeros(L.shape2d)
    b[:, J, I] = v
    return b


@lru_cache(maxsize=4)
def cg2d_case(seed=0, precond="diagonal", nlev=None):
    """Config-0 problem: (grid, ops, cg2dNorm). Cached (the grid takes ~1 s)."""
    g = make_grid(seed, nlev=nlev)
    norm = ini_cg2d_norm(g)
    return g, update_cg2d(g, norm, precond=precond), norm


# ------------------------------------------------------------------------------------------------ DST3 inputs
def flow(g, seed=2, cfl=0.3):
    """(uFld, vFld, wFld) [T, K, ny, nx]: streamfunction flow, w from continuity, max CFL = cfl."""
    maps = _load_compact()
    L = g.L
    T, K = L.nTiles, L.Nr
    J, I = L.js(1, L.sNy), L.is_(1, L.sNx)
    rng = np.random.default_rng(seed)
    amp = np.exp(g.rC / 1000.0)  # rC < 0: decays with depth
    psi_f = {}
    base = {k: _smooth_field(rng, (s[0] + 1, s[1] + 1)) for k, s in FACET_SHAPES.items()}
    lvl = {k: _smooth_field(rng, (s[0] + 1, s[1] + 1), nmodes=8) for k, s in FACET_SHAPES.items()}
    for k in FACET_SHAPES:
        psi_f[k] = (base[k][None] + 0.3 * np.cos(np.arange(K) * 0.2)[:, None, None] * lvl[k][None]) \
            * amp[:, None, None]
    # corner values of each tile: rows/cols 0..90 of the facet block (the +1 corner comes from the facet)
    n = L.sNx
    psi = np.zeros((T, K, n + 1, n + 1))
    tiles = []
    for f in (1, 2):
        tiles += [psi_f[f][:, n * k:n * (k + 1) + 1, :] for k in range(3)]
    tiles.append(psi_f[3])
    for f in (4, 5):
        tiles += [psi_f[f][:, :, n * k:n * (k + 1) + 1] for k in range(3)]
    for t in range(T):
        psi[t] = tiles[t]
    uT = np.zeros(L.shape3d)
    vT = np.zeros(L.shape3d)
    uT[:, :, J, I] = psi[:, :, 1:, :-1] - psi[:, :, :-1, :-1]
    vT[:, :, J, I] = -(psi[:, :, :-1, 1:] - psi[:, :, :-1, :-1])
    uT *= g.maskW
    vT *= g.maskS
    xA = g.dyG[:, None] * g.drF[None, :, None, None] * g.hFacW
    yA = g.dxG[:, None] * g.drF[None, :, None, None] * g.hFacS
    u = np.where(xA != 0.0, uT / np.where(xA != 0.0, xA, 1.0), 0.0)
    v = np.where(yA != 0.0, vT / np.where(yA != 0.0, yA, 1.0), 0.0)
    u, v = _exch_np(maps, "UVs", u, v)
    w = continuity_w(g, u, v)
    c_u = np.abs(u * DT * g.recip_dxC[:, None]).max()
    c_v = np.abs(v * DT * g.recip_dyC[:, None]).max()
    c_w = np.abs(w * DT * g.recip_drC[None, :, None, None]).max()
    s = cfl / max(c_u, c_v, c_w)
    return u * s, v * s, w * s


def continuity_w(g, u, v):
    """wVel from INTEGRATE_FOR_W (core/free_surface.py:255-279, rStarDhDt = 0, deep/rho factors 1): interior
    points, k = Nr..1, then the 3D (T) exchange."""
    maps = _load_compact()
    L = g.L
    K = L.Nr
    J1, I1 = L.js(1, L.sNy + 1), L.is_(1, L.sNx + 1)
    J, I = L.js(1, L.sNy), L.is_(1, L.sNx)
    drF = g.drF[None, :, None, None]
    uT = u[:, :, J1, I1] * g.dyG[:, None, J1, I1] * 1.0 * 1.0 * drF * g.hFacW[:, :, J1, I1]
    vT = v[:, :, J1, I1] * g.dxG[:, None, J1, I1] * 1.0 * 1.0 * drF * g.hFacS[:, :, J1, I1]
    div = uT[:, :, :-1, 1:] - uT[:, :, :-1, :-1] + vT[:, :, 1:, :-1] - vT[:, :, :-1, :-1]
    conv = -div
    rA_i = g.recip_rA[:, J, I]
    mC = g.maskC[:, :, J, I]
    w = np.zeros(L.shape3d)
    wk = conv[:, K - 1] * rA_i * mC[:, K - 1]
    w[:, K - 1, J, I] = wk
    for k in range(K - 2, -1, -1):
        wk = (wk + conv[:, k] * rA_i) * mC[:, k]
        w[:, k, J, I] = wk
    return _exch_np(maps, "3D", w)


def tracer(g, seed=3, noise=0.1):
    """2 + 25 exp(-z/800 m) + N(0, noise) on wet cells, 0 on dry cells, T exchange."""
    maps = _load_compact()
    L = g.L
    rng = np.random.default_rng(seed)
    z = -g.rC[None, :, None, None]
    tr = (2.0 + 25.0 * np.exp(-z / 800.0)) + rng.normal(0.0, noise, L.shape3d)
    tr = np.where(g.maskC > 0, tr, 0.0)
    return _exch_np(maps, "T", tr)


def gad_inputs(g):
    """The grid pieces mj_gad_dst3 reads: (grid2d dict, vgrid dict)."""
    grid2d = {k: g.f[k] for k in ("dyG", "dxG", "recip_dxC", "recip_dyC", "recip_rA", "rA", "maskInC", "maskW",
                                  "maskS", "maskC")}
    vgrid = {k: g.v[k] for k in ("dTlev", "drF", "recip_drF", "rhoFacC", "recip_rhoFacC", "deepFacC",
                                 "recip_deepFacC", "recip_deepFac2C", "recip_drC")}
    return grid2d, vgrid
