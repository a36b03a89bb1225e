// GAD DST3 tracer advection on gfx950.
//
// Horizontal part: the reference's GAD_ADVECTION with implicitAdvection (pkgs/gad.py:460-539, gad_advection.F):
// three passes, each an X and a Y block whose action per tile comes from the facet pass table (which direction
// sweeps, overlap-only / interior-only update regions, FILL_CS_CORNER_TR_RL corner fills before/after the flux).
// Everything is tile-local, so one workgroup owns one (member, tile, level): the 98x98 FP64 working tracer
// (localTij) and the input tracer live in LDS (2 x 76.8 KB, gfx950's 160 KB LDS); each point evaluates the
// interfaces on both its faces, velocities and metrics are read once per sweep that needs them, and only the
// tendency is written. The reference evaluates all six flux arrays on every tile and
// masks; here each tile evaluates only the sweeps its table runs (2-3 of 6).
//
// Vertical part (R, only when wFld is given): a DST3 flux through the level interfaces applied after the
// horizontal passes, with the same non-divergent update form. It is not in the reference (which runs the vertical
// implicitly); its definition is written out at vert_kernel.
#include <string.h>

#include <algorithm>

#include "mj_common.h"
#include "mj_internal.h"

namespace mj {

constexpr int GNY = 98, GNX = 98, GPLANE = GNY * GNX;  // LLC90 tiles with OL = 4

struct GadDev {
    int B, T, K, OLx, OLy, sNx, sNy;
    int64_t ms3, mms;
    const int32_t* mask_index;  // device [B]: mask slot of each member, or NULL (slot = member)
    int lpb;                    // levels per block of the horizontal launch (consecutive levels of one tile)
    const MjTilePlan* tiles;
    const double *dyG, *dxG, *recip_dxC, *recip_dyC, *recip_rA, *rA, *maskInC;
    const double *maskW, *maskS, *maskC;
    const double *dTlev, *drF, *recip_drF, *rhoFacC, *recip_rhoFacC, *deepFacC, *recip_deepFacC, *recip_deepFac2C,
        *recip_drC;
    double oneSixth;
    const double *u, *v, *w, *tr, *hFacW, *hFacS, *recip_hFacC;
    double* out;  // gTracer (no R flux) or localTij after the horizontal passes (R flux follows)
    const int32_t* nlev;
    unsigned long long* mbits;  // maskLocW / maskLocS bit planes per (mask slot, tile, level), gad_maskbits_kernel
    double2* g2;  // per tile, the 2-D metrics a sweep point reads, in pairs (gad_pack2d_kernel): [3][T][GPLANE]
                  // {dyG, recip_dxC} (X faces), {dxG, recip_dyC} (Y faces), {recip_rA, maskInC}
};

// Mask slot of member m: members of one level class share their grid masks.
__device__ __forceinline__ int64_t mask_slot(const GadDev& a, int m) { return a.mask_index ? a.mask_index[m] : m; }
constexpr int kMaskWords = (GPLANE + 63) / 64;
// The bit planes of member m's masks at (tile t, level k): maskLocW's kMaskWords words, then maskLocS's.
__device__ __forceinline__ const unsigned long long* unit_bits(const GadDev& a, int m, int t, int k) {
    const int64_t s = a.mms ? mask_slot(a, m) : 0;
    return a.mbits + ((s * a.T + t) * a.K + k) * 2 * kMaskWords;
}

// maskLocW / maskLocS: maskW / maskS with FILL_CS_CORNER_UV_RS(.FALSE.) applied (gad.py:263-281, :489-490).
// (ia, ja) array indices; returns the corner-filled mask value at that point.
__device__ __forceinline__ double mask_loc(const GadDev& a, const MjTilePlan& tp, const double* mW, const double* mS,
                                           int ja, int ia, bool want_u) {
    const int fi = ia - a.OLx + 1, fj = ja - a.OLy + 1;  // Fortran indices
    const int sNx = a.sNx, sNy = a.sNy;
    auto at = [&](const double* f, int i, int j) { return f[(j - 1 + a.OLy) * GNX + (i - 1 + a.OLx)]; };
    if (fi <= 0 && fj <= 0 && tp.sw) {
        const int i = 1 - fi, j = 1 - fj;
        return want_u ? at(mS, 1 - j, 1 + i) : at(mW, 1 + j, 1 - i);
    }
    if (fi >= sNx + 1 && fj <= 0 && tp.se) {
        const int i = fi - sNx, j = 1 - fj;
        if (want_u) return i >= 2 ? at(mS, sNx + j, i) : at(mW, fi, fj);
        return at(mW, sNx + 1 - j, 1 - i);
    }
    if (fi <= 0 && fj >= sNy + 1 && tp.nw) {
        const int i = 1 - fi, j = fj - sNy;
        if (want_u) return at(mS, 1 - j, sNy + 1 - i);
        return j >= 2 ? at(mW, j, sNy + i) : at(mS, fi, fj);
    }
    if (fi >= sNx + 1 && fj >= sNy + 1 && tp.ne) {
        const int i = fi - sNx, j = fj - sNy;
        if (want_u) return i >= 2 ? at(mS, sNx + j, sNy + 2 - i) : at(mW, fi, fj);
        return j >= 2 ? at(mW, sNx + 2 - j, sNy + i) : at(mS, fi, fj);
    }
    return want_u ? at(mW, fi, fj) : at(mS, fi, fj);
}

// FILL_CS_CORNER_TR_RL(fill4dir, withSigns=.FALSE.) (gad.py:241-260): the (dst, src) pair of thread tid, for each
// corner present and i, j = 1..OL (false: no pair). Destinations are distinct, sources are distinct, and no
// source is a destination.
__device__ __forceinline__ bool corner_pair(const GadDev& a, const MjTilePlan& tp, int dir, int tid, int& dst,
                                            int& src) {
    const int OL = a.OLx;  // OLx == OLy checked on the host
    if (tid >= 4 * OL * OL) return false;
    const int c = tid / (OL * OL), r = tid - c * OL * OL;
    const int j = r / OL + 1, i = r % OL + 1;
    const int sNx = a.sNx, sNy = a.sNy;
    const bool on = c == 0 ? tp.sw : c == 1 ? tp.se : c == 2 ? tp.nw : tp.ne;
    if (!on) return false;
    int di, dj, si, sj;
    if (dir == 1) {
        if (c == 0) { di = 1 - i; dj = 1 - j; si = 1 - j; sj = i; }
        else if (c == 1) { di = sNx + i; dj = 1 - j; si = sNx + j; sj = i; }
        else if (c == 2) { di = 1 - i; dj = sNy + j; si = 1 - j; sj = sNy + 1 - i; }
        else { di = sNx + i; dj = sNy + j; si = sNx + j; sj = sNy + 1 - i; }
    } else {
        if (c == 0) { di = 1 - i; dj = 1 - j; si = j; sj = 1 - i; }
        else if (c == 1) { di = sNx + i; dj = 1 - j; si = sNx + 1 - j; sj = 1 - i; }
        else if (c == 2) { di = 1 - i; dj = sNy + j; si = j; sj = sNy + i; }
        else { di = sNx + i; dj = sNy + j; si = sNx + 1 - j; sj = sNy + i; }
    }
    dst = (dj - 1 + a.OLy) * GNX + (di - 1 + a.OLx);
    src = (sj - 1 + a.OLy) * GNX + (si - 1 + a.OLx);
    return true;
}

// The fill on a tile (LDS or the block's own global plane): all reads precede all writes.
__device__ __forceinline__ void corner_fill(double* loc, const GadDev& a, const MjTilePlan& tp, int dir, int tid) {
    int dst = -1, src = -1;
    double v = 0.0;
    const bool on = corner_pair(a, tp, dir, tid, dst, src);
    if (on) v = loc[src];
    __syncthreads();
    if (on) loc[dst] = v;
    __syncthreads();
}

// Transpose of the fill on a cotangent plane: the destination's cotangent moves to its source.
__device__ __forceinline__ void corner_fill_T(double* bar, const GadDev& a, const MjTilePlan& tp, int dir, int tid) {
    int dst = -1, src = -1;
    double v = 0.0;
    const bool on = corner_pair(a, tp, dir, tid, dst, src);
    if (on) v = bar[dst];
    __syncthreads();
    if (on) {
        bar[src] = bar[src] + v;
        bar[dst] = 0.0;
    }
    __syncthreads();
}

// Unit (tile, member, level block) of a DST3 launch of gridDim.x >= B*T*ceil(K/L) blocks (a multiple of 8, L =
// a.lpb levels per block): units ordered tile-major (tile, member, level block) and cut into 8 contiguous ranges, one
// per XCD (block b runs on XCD b % 8), so each XCD works through about T/8 tiles and their 2-D metrics stay in its L2
// across members and levels. k is the block's first level.
__device__ __forceinline__ bool gad_unit(const GadDev& a, int& t, int& m, int& k) {
    const int nkb = (a.K + a.lpb - 1) / a.lpb;
    const int n = a.B * a.T * nkb;
    const int chunk = (n + 7) / 8;
    const int pos = blockIdx.x / 8;
    const int u = (blockIdx.x % 8) * chunk + pos;
    if (pos >= chunk || u >= n) return false;
    k = (u % nkb) * a.lpb;
    const int tm = u / nkb;
    m = tm % a.B;
    t = tm / a.B;
    return true;
}

// The points of the corner squares of the corners the tile has: mask_loc (corner-filled) into the bit planes.
__device__ __forceinline__ void fix_mask_corners(const GadDev& a, const MjTilePlan& tp, const double* mW,
                                                 const double* mS, unsigned long long* bw, unsigned long long* bs) {
    // corner c (sw, se, nw, ne), point p of its OL x OL square, mask u (W) or v (S)
    const int OL = a.OLx, n = 4 * OL * OL;
    const int tid = threadIdx.x;
    if (tid < 2 * n) {
        const int want_u = tid < n, r = tid % n, c = r / (OL * OL), pp = r % (OL * OL);
        const bool on = c == 0 ? tp.sw : c == 1 ? tp.se : c == 2 ? tp.nw : tp.ne;
        if (on) {
            const int dj = pp / OL, di = pp % OL;
            const int ja = (c < 2) ? dj : a.OLy + a.sNy + dj, ia = (c % 2 == 0) ? di : a.OLx + a.sNx + di;
            const int q = ja * GNX + ia;
            const bool v = mask_loc(a, tp, mW, mS, ja, ia, want_u) != 0.0;
            unsigned long long* word = (want_u ? bw : bs) + (q >> 6);
            const unsigned long long bit = 1ull << (q & 63);
            if (v)
                atomicOr(word, bit);
            else
                atomicAnd(word, ~bit);
        }
    }
    __syncthreads();
}

// maskLocW / maskLocS of one (member, tile, level) as LDS bit planes (the masks are 0/1 by construction, hFac != 0,
// so a bit and its double 0.0/1.0 carry the same value). Off the four corner halo squares mask_loc is the plain
// mask, so every wave ballots the plain maskW / maskS of its words (all its loads in flight at once); then the
// points of the corner squares of the corners the tile has are set from mask_loc.
__device__ __forceinline__ void build_mask_bits(const GadDev& a, const MjTilePlan& tp, const double* mW,
                                                const double* mS, unsigned long long* bw, unsigned long long* bs) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    constexpr int kWordsPerWave = 16;  // words of one wave whose loads are issued together
    for (int w0 = wave; w0 < kMaskWords; w0 += nwaves * kWordsPerWave) {
        double vw[kWordsPerWave], vs[kWordsPerWave];
#pragma unroll
        for (int r = 0; r < kWordsPerWave; ++r) {
            const int q = (w0 + r * nwaves) * 64 + lane;
            const bool in = w0 + r * nwaves < kMaskWords && q < GPLANE;
            vw[r] = in ? mW[q] : 0.0;
            vs[r] = in ? mS[q] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < kWordsPerWave; ++r) {
            const unsigned long long bwv = __ballot(vw[r] != 0.0), bsv = __ballot(vs[r] != 0.0);
            if (lane == 0 && w0 + r * nwaves < kMaskWords) {
                bw[w0 + r * nwaves] = bwv;
                bs[w0 + r * nwaves] = bsv;
            }
        }
    }
    __syncthreads();
    fix_mask_corners(a, tp, mW, mS, bw, bs);
}
// maskLocW / maskLocS of every (mask slot, tile, level) as bit planes in global memory, once per call: the DST3
// blocks of the members that share a slot load its 2 x 151 words instead of the two mask planes (2 x 77 KB).
__global__ void __launch_bounds__(256) gad_maskbits_kernel(GadDev a, int nslots) {
    __shared__ unsigned long long bw[kMaskWords], bs[kMaskWords];
    const int u = blockIdx.x;
    const int k = u % a.K, st = u / a.K, t = st % a.T, sl = st / a.T;
    if (sl >= nslots) return;
    const int64_t bm = sl * a.mms + ((int64_t)t * a.K + k) * GPLANE;
    build_mask_bits(a, a.tiles[t], a.maskW + bm, a.maskS + bm, bw, bs);
    unsigned long long* o = a.mbits + (int64_t)u * 2 * kMaskWords;
    for (int i = threadIdx.x; i < kMaskWords; i += blockDim.x) {
        o[i] = bw[i];
        o[kMaskWords + i] = bs[i];
    }
}

// The 2-D metrics in the pairs the sweeps load together (one 16-byte load instead of two 8-byte ones), per call.
__global__ void __launch_bounds__(256) gad_pack2d_kernel(GadDev a) {
    const int64_t n = (int64_t)a.T * GPLANE;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    a.g2[q] = make_double2(a.dyG[q], a.recip_dxC[q]);
    a.g2[n + q] = make_double2(a.dxG[q], a.recip_dyC[q]);
    a.g2[2 * n + q] = make_double2(a.recip_rA[q], a.maskInC[q]);
}

// The tracer plane into LDS (into loc and loc2) and the unit's mask bit planes, all loads in flight together (one
// memory latency for a 1024-thread block).
__device__ __forceinline__ void load_tile_and_bits(const double* tr, const unsigned long long* gb, double* loc,
                                                   double* loc2, unsigned long long* bw, unsigned long long* bs) {
    const int tid = threadIdx.x, nthr = blockDim.x;
    constexpr int KP = (GPLANE + 1023) / 1024;  // per thread at 1024
    for (int it = 0; it * KP * nthr < GPLANE || it * nthr < 2 * kMaskWords; ++it) {
        double v[KP];
#pragma unroll
        for (int b = 0; b < KP; ++b) {
            const int q = (it * KP + b) * nthr + tid;
            v[b] = q < GPLANE ? tr[q] : 0.0;
        }
        const int wd = it * nthr + tid;
        const unsigned long long word = wd < 2 * kMaskWords ? gb[wd] : 0ull;
#pragma unroll
        for (int b = 0; b < KP; ++b) {
            const int q = (it * KP + b) * nthr + tid;
            if (q < GPLANE) loc[q] = loc2[q] = v[b];
        }
        if (wd < kMaskWords)
            bw[wd] = word;
        else if (wd < 2 * kMaskWords)
            bs[wd - kMaskWords] = word;
    }
    __syncthreads();
}

__device__ __forceinline__ double mbit(const unsigned long long* b, int q) {
    return (double)((b[q >> 6] >> (q & 63)) & 1ull);
}

// The interface fluxes a sweep's update reads: faces q and q + d (d = 1 in X, GNX in Y) of every box point.
template <class F>
__device__ __forceinline__ void flux_boxes(const MjSweep& S, int dir, double* af, F&& flux) {
    for (int bx = 0; bx < S.nbox; ++bx) {
        const MjBox B = S.box[bx];
        const int i1 = B.i1 + (dir == 0 ? 1 : 0), j1 = B.j1 + (dir == 1 ? 1 : 0);
        const int bw = i1 - B.i0, n = bw * (j1 - B.j0);
        for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
            const int q = (B.j0 + idx / bw) * GNX + B.i0 + idx % bw;
            af[q] = flux(q);
        }
    }
}

// Box point idx (row-major, width bw) of box B: array index.
__device__ __forceinline__ int box_point(const MjBox& B, int bw, float inv_bw, int idx) {
    const int r = (int)(((float)idx + 0.5f) * inv_bw);  // idx / bw, exact for idx < 2^14 and bw <= 128
    return (B.j0 + r) * GNX + B.i0 + (idx - r * bw);
}

// One workgroup of 1024 threads per (member, tile, level). Every box point of a sweep evaluates the interface
// values of both its faces (q and q + d) itself, with no flux plane: from the same loads and in the same order of
// operations, so the two points beside a face get the same bits, and the update is the same IEEE expression as
// with one shared evaluation. The point keeps its increment in a register until every thread has read the tracer
// plane, and only then updates it. LDS: the working plane, the input tracer (read by every update and by the
// output) and the mask bits, 156 KB.
constexpr int kHzThreads = 1024;
constexpr int kHzPts = (GPLANE + kHzThreads - 1) / kHzThreads;  // box points per thread (disjoint boxes in the plane)

// One (member, tile, level) plane of the horizontal passes, LDS planes given.
__device__ __forceinline__ void hz_plane(const GadDev& a, int t, int m, int k, double* loc, double* trc,
                                         unsigned long long* mlw, unsigned long long* mls) {
    constexpr int NT = kHzThreads;
    const int tid = threadIdx.x;
    const MjTilePlan& tp = a.tiles[t];
    const int64_t b3 = m * a.ms3 + ((int64_t)t * a.K + k) * GPLANE;
    const int64_t b2 = (int64_t)t * GPLANE;
    const double* __restrict__ tr = a.tr + b3;
    const double* __restrict__ rhC = a.recip_hFacC + b3;
    const double2* __restrict__ am = a.g2 + 2 * (int64_t)a.T * GPLANE + b2;  // {recip_rA, maskInC}
    const double dt = a.dTlev[k];
    const double deepFacC = a.deepFacC[k], rdeepC = a.recip_deepFacC[k], rdeep2C = a.recip_deepFac2C[k];
    const double rhoFacC = a.rhoFacC[k], rrhoFacC = a.recip_rhoFacC[k], drF = a.drF[k], rdrF = a.recip_drF[k];
    const double oneSixth = a.oneSixth;

    load_tile_and_bits(tr, unit_bits(a, m, t, k), loc, trc, mlw, mls);

    for (int pass = 0; pass < 3; ++pass) {
        for (int dir = 0; dir < 2; ++dir) {
            const MjSweep& S = tp.sweep[pass][dir];
            if (!S.run) continue;
            const double* __restrict__ vel = (dir == 0 ? a.u : a.v) + b3;
            const double* __restrict__ hf = (dir == 0 ? a.hFacW : a.hFacS) + b3;
            const double2* __restrict__ gr = a.g2 + dir * (int64_t)a.T * GPLANE + b2;  // {dyG|dxG, recip_dxC|dyC}
            const unsigned long long* mb = dir == 0 ? mlw : mls;
            const int d = dir == 0 ? 1 : GNX, cmax = dir == 0 ? GNX - 2 : GNY - 2;
            auto trans = [&](double vv, double g, double h) { return vv * (g * deepFacC * drF * h) * rhoFacC; };
            // interface value of face q (gad.py:482-485 transport; 0 outside the Fortran flux loops
            // i = 1-OLx+2 .. sNx+OLx-1, gad_dst3_adv_x.F:80, and j = 1-OLy+2 .. sNy+OLy-1, gad_dst3_adv_y.F:78)
            auto in_range = [&](int q) {
                const int j = q / GNX, c = dir == 0 ? q - j * GNX : j;
                return c >= 2 && c <= cmax;
            };
            // the interface values of faces q and q + d of a point: the masked differences Rj, Rjp of face q are
            // Rjm, Rj of face q + d, so they are formed once (the same operations as one face at a time)
            auto faces = [&](int q, double V0, double T0, double R0, double V1, double T1, double R1, double& f0,
                             double& f1) {
                const bool in0 = in_range(q), in1 = in_range(q + d);
                f0 = 0.0;
                f1 = 0.0;
                if (!in0 && !in1) return;
                const double t_m1 = loc[q - d], t_0 = loc[q], t_p1 = loc[q + d];
                const double D0 = (t_0 - t_m1) * mbit(mb, q), Dp1 = (t_p1 - t_0) * mbit(mb, q + d);
                if (in0) {
                    const double Dm1 = (t_m1 - loc[q - 2 * d]) * mbit(mb, q - d);
                    f0 = dst3_flux_r(T0, fabs(V0 * dt * R0 * rdeepC), oneSixth, t_m1, t_0, Dm1, D0, Dp1);
                }
                if (in1) {
                    const double Dp2 = (loc[q + 2 * d] - t_p1) * mbit(mb, q + 2 * d);
                    f1 = dst3_flux_r(T1, fabs(V1 * dt * R1 * rdeepC), oneSixth, t_0, t_p1, D0, Dp1, Dp2);
                }
            };
            if (S.fill_before) corner_fill(loc, a, tp, S.fill_before, tid);
            // the boxes' points as one index range (the boxes are disjoint, so at most GPLANE points)
            const MjBox B0 = S.box[0], B1 = S.box[1];
            const int bw0 = B0.i1 - B0.i0, n0 = S.nbox > 0 ? bw0 * (B0.j1 - B0.j0) : 0;
            const int bw1 = B1.i1 - B1.i0, n1 = S.nbox > 1 ? bw1 * (B1.j1 - B1.j0) : 0;
            const float inv0 = 1.0f / (float)max(bw0, 1), inv1 = 1.0f / (float)max(bw1, 1);
            auto point = [&](int idx) {
                return idx < n0 ? box_point(B0, bw0, inv0, idx) : box_point(B1, bw1, inv1, idx - n0);
            };
            double inc[kHzPts];
#pragma unroll
            for (int r = 0; r < kHzPts; ++r) {
                const int idx = tid + r * NT;
                inc[r] = 0.0;
                if (idx < n0 + n1) {
                    const int qq = point(idx), qd = qq + d;
                    const double2 GR0 = gr[qq], GR1 = gr[qd], AM = am[qq];
                    const double V0 = vel[qq], G0 = GR0.x, H0 = hf[qq], R0 = GR0.y;
                    const double V1 = vel[qd], G1 = GR1.x, H1 = hf[qd], R1 = GR1.y;
                    const double TR = trc[qq], RH = rhC[qq], RA = AM.x, MI = AM.y;
                    const double T0 = trans(V0, G0, H0), T1 = trans(V1, G1, H1);
                    double f0, f1;
                    faces(qq, V0, T0, R0, V1, T1, R1, f0, f1);
                    // local - coef*(d(af) - tracer*d(trans))*maskInC (gad.py:511-514), coefficient gad.py:494
                    const double daf = f1 - f0, dtr = T1 - T0;
                    const double coef = dt * rrhoFacC * RH * rdrF * RA * rdeep2C;
                    inc[r] = coef * (daf - TR * dtr) * MI;
                }
            }
            __syncthreads();  // every read of the plane precedes the updates
            if (S.fill_after) corner_fill(loc, a, tp, S.fill_after, tid);
#pragma unroll
            for (int r = 0; r < kHzPts; ++r) {
                const int idx = tid + r * NT;
                if (idx < n0 + n1) {
                    const int qq = point(idx);
                    loc[qq] = loc[qq] - inc[r];
                }
            }
            __syncthreads();
        }
    }
    double* out = a.out + b3;
    for (int q = tid; q < GPLANE; q += NT) {
        if (a.w)
            out[q] = loc[q];
        else
            out[q] = (loc[q] - trc[q]) / dt;  // gad.py:539
    }
}

// The plane as a call of its own (the multi-level blocks: inlined into the level loop it spills registers).
__device__ __noinline__ void hz_plane_call(const GadDev& a, int t, int m, int k, double* loc, double* trc,
                                           unsigned long long* mlw, unsigned long long* mls) {
    hz_plane(a, t, m, k, loc, trc, mlw, mls);
}

// One workgroup per (member, tile, level block): LEVELS false, one level per block (a.lpb == 1); true, a.lpb
// consecutive levels, one after the other in the same LDS planes.
template <bool LEVELS>
__global__ void __launch_bounds__(kHzThreads, 4) gad_hz_kernel(GadDev a) {
    __shared__ double loc[GPLANE];
    __shared__ double trc[GPLANE];
    __shared__ unsigned long long mlw[kMaskWords], mls[kMaskWords];
    int t, m, k0;
    if (!gad_unit(a, t, m, k0)) return;
    if (!LEVELS) {
        if (a.nlev && k0 >= a.nlev[m]) return;
        hz_plane(a, t, m, k0, loc, trc, mlw, mls);
        return;
    }
    const int kend = min(k0 + a.lpb, a.nlev ? min(a.nlev[m], a.K) : a.K);
    for (int k = k0; k < kend; ++k) {
        if (k > k0) __syncthreads();  // the previous level's output has read the LDS planes
        hz_plane_call(a, t, m, k, loc, trc, mlw, mls);
    }
}

// ------------------------------------------------------------------------------------------------ reverse mode
constexpr int kAdjChunk = 1024;  // blocks per reverse launch (each uses 7 planes of the workspace)
// Partial derivatives of one DST3 interface value f = A (t_m1 + d0 Rj + d1 Rjm) + Bm (t_0 - d0 Rj - d1 Rjp)
// (A = (tr + |tr|)/2, Bm = (tr - |tr|)/2, d0 = (2-c)(1-c)/6, d1 = (1-c^2)/6): w[0..3] = df/dt_m2, df/dt_m1,
// df/dt_0, df/dt_p1 (independent of the tracer), dtr = df/dtr, dc = df/dcfl. The derivative of |x| at 0 is taken
// as 0 (sign(0) = 0), which is what a central difference across x = 0 gives.
struct FluxPartials {
    double w[4], dtr, dc;
};
__device__ __forceinline__ FluxPartials dst3_partials(double tr_, double cfl, double oneSixth, double t_m2,
                                                      double t_m1, double t_0, double t_p1, double m_m1, double m_0,
                                                      double m_p1) {
    const double Rjp = (t_p1 - t_0) * m_p1;
    const double Rj = (t_0 - t_m1) * m_0;
    const double Rjm = (t_m1 - t_m2) * m_m1;
    const double d0 = (2. - cfl) * (1. - cfl) * oneSixth;
    const double d1 = (1. - cfl * cfl) * oneSixth;
    const double A = 0.5 * (tr_ + fabs(tr_)), Bm = 0.5 * (tr_ - fabs(tr_));
    const double sg = tr_ > 0.0 ? 1.0 : (tr_ < 0.0 ? -1.0 : 0.0);
    FluxPartials P;
    P.w[0] = -A * d1 * m_m1;
    P.w[1] = A * (1.0 - d0 * m_0 + d1 * m_m1) + Bm * (d0 * m_0);
    P.w[2] = A * (d0 * m_0) + Bm * (1.0 - d0 * m_0 + d1 * m_p1);
    P.w[3] = -Bm * d1 * m_p1;
    const double Pv = t_m1 + (d0 * Rj + d1 * Rjm), Nv = t_0 - (d0 * Rj + d1 * Rjp);
    P.dtr = 0.5 * (1.0 + sg) * Pv + 0.5 * (1.0 - sg) * Nv;
    const double d0c = (2.0 * cfl - 3.0) * oneSixth, d1c = -2.0 * cfl * oneSixth;
    P.dc = A * (d0c * Rj + d1c * Rjm) - Bm * (d0c * Rj + d1c * Rjp);
    return P;
}

// Reverse mode of gad_hz_kernel (horizontal DST3 without R flux) for one (member, tile, level): given the
// cotangent gbar of gTracer, writes the cotangents of the tracer, uFld and vFld planes (every padded point; the
// grid, hFac and mask inputs are treated as constants). The forward is replayed with a snapshot of the tile
// before each sweep (work: 7 planes per block, the first holds the running cotangent of localTij); the sweeps
// are then reversed: the box update (af, tracer and transport terms), the corner fill after the flux, the flux
// (weights of the four tracer values, transport and CFL terms), the corner fill before it.
//
// With lbar (the R flux's reverse mode) the sweeps start from lbar, the cotangent of localTij after the horizontal
// passes, instead of from gbar, and the tracer cotangent already holds the vertical part (added to, not set).
__global__ void __launch_bounds__(1024) gad_hz_adjoint_kernel(GadDev a, const double* gbar, double* trb, double* ub,
                                                              double* vb, double* work, int blk0, const double* lbar) {
    __shared__ double loc[GPLANE];
    __shared__ double fb[GPLANE];
    __shared__ unsigned long long mlw[kMaskWords], mls[kMaskWords];
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int blk = blk0 + blockIdx.x;
    const int k = blk % a.K;
    const int mt = blk / a.K;
    const int t = mt % a.T, m = mt / a.T;
    const int64_t b3 = m * a.ms3 + ((int64_t)t * a.K + k) * GPLANE;
    double* TB = trb + b3;
    double* UB = ub + b3;
    double* VB = vb + b3;
    if (a.nlev && k >= a.nlev[m]) {  // no dependence: zero cotangents
        for (int q = tid; q < GPLANE; q += nthr) {
            TB[q] = 0.0;
            UB[q] = 0.0;
            VB[q] = 0.0;
        }
        return;
    }
    const MjTilePlan& tp = a.tiles[t];
    const int64_t b2 = (int64_t)t * GPLANE;
    const double* tr = a.tr + b3;
    const double* u = a.u + b3;
    const double* v = a.v + b3;
    const double* hW = a.hFacW + b3;
    const double* hS = a.hFacS + b3;
    const double* rhC = a.recip_hFacC + b3;
    const double *dyG = a.dyG + b2, *dxG = a.dxG + b2, *rdxC = a.recip_dxC + b2, *rdyC = a.recip_dyC + b2;
    const double *rrA = a.recip_rA + b2, *mInC = a.maskInC + b2;
    const double dt = a.dTlev[k];
    const double deepFacC = a.deepFacC[k], rdeepC = a.recip_deepFacC[k], rdeep2C = a.recip_deepFac2C[k];
    const double rhoFacC = a.rhoFacC[k], rrhoFacC = a.recip_rhoFacC[k], drF = a.drF[k], rdrF = a.recip_drF[k];
    const double oneSixth = a.oneSixth;
    double* Lb = work + (size_t)blockIdx.x * 7 * GPLANE;
    double* snap = Lb + GPLANE;

    auto uTrans = [&](int q) { return u[q] * (dyG[q] * deepFacC * drF * hW[q]) * rhoFacC; };
    auto vTrans = [&](int q) { return v[q] * (dxG[q] * deepFacC * drF * hS[q]) * rhoFacC; };
    auto coef = [&](int q) { return dt * rrhoFacC * rhC[q] * rdrF * rrA[q] * rdeep2C; };
    auto in_box = [&](const MjSweep& S, int q) {
        const int j = q / GNX, i = q - j * GNX;
        for (int bx = 0; bx < S.nbox; ++bx)
            if (i >= S.box[bx].i0 && i < S.box[bx].i1 && j >= S.box[bx].j0 && j < S.box[bx].j1) return true;
        return false;
    };
    // flux partials at q in direction dir from the tile state in loc; false outside the flux domain
    auto partials = [&](int dir, int q, FluxPartials& P) {
        const int j = q / GNX, i = q - j * GNX;
        if (dir == 0) {
            if (i < 2 || i > GNX - 2) return false;
            const double cfl = fabs(u[q] * dt * rdxC[q] * rdeepC);
            P = dst3_partials(uTrans(q), cfl, oneSixth, loc[q - 2], loc[q - 1], loc[q], loc[q + 1], mbit(mlw, q - 1),
                              mbit(mlw, q), mbit(mlw, q + 1));
        } else {
            if (j < 2 || j > GNY - 2) return false;
            const double cfl = fabs(v[q] * dt * rdyC[q] * rdeepC);
            P = dst3_partials(vTrans(q), cfl, oneSixth, loc[q - 2 * GNX], loc[q - GNX], loc[q], loc[q + GNX],
                              mbit(mls, q - GNX), mbit(mls, q), mbit(mls, q + GNX));
        }
        return true;
    };
    auto flux = [&](int dir, int q) {
        const int j = q / GNX, i = q - j * GNX;
        if (dir == 0) {
            if (i < 2 || i > GNX - 2) return 0.0;
            const double cfl = fabs(u[q] * dt * rdxC[q] * rdeepC);
            return dst3_flux(uTrans(q), cfl, oneSixth, loc[q - 2], loc[q - 1], loc[q], loc[q + 1], mbit(mlw, q - 1),
                             mbit(mlw, q), mbit(mlw, q + 1));
        }
        if (j < 2 || j > GNY - 2) return 0.0;
        const double cfl = fabs(v[q] * dt * rdyC[q] * rdeepC);
        return dst3_flux(vTrans(q), cfl, oneSixth, loc[q - 2 * GNX], loc[q - GNX], loc[q], loc[q + GNX],
                         mbit(mls, q - GNX), mbit(mls, q), mbit(mls, q + GNX));
    };

    // ---- forward replay, snapshot before each sweep
    int order[6], ns = 0;
    for (int q = tid; q < GPLANE; q += nthr) loc[q] = tr[q];
    {
        const unsigned long long* gb = unit_bits(a, m, t, k);
        for (int wd = tid; wd < 2 * kMaskWords; wd += nthr) {
            if (wd < kMaskWords)
                mlw[wd] = gb[wd];
            else
                mls[wd - kMaskWords] = gb[wd];
        }
    }
    __syncthreads();
    for (int pass = 0; pass < 3; ++pass) {
        for (int dir = 0; dir < 2; ++dir) {
            const MjSweep& S = tp.sweep[pass][dir];
            if (!S.run) continue;
            double* sn = snap + (size_t)ns * GPLANE;
            for (int q = tid; q < GPLANE; q += nthr) sn[q] = loc[q];
            order[ns++] = pass * 2 + dir;
            if (S.fill_before) corner_fill(loc, a, tp, S.fill_before, tid);
            if (S.nbox) {
                flux_boxes(S, dir, fb, [&](int q) { return flux(dir, q); });
                __syncthreads();
            }
            if (S.fill_after) corner_fill(loc, a, tp, S.fill_after, tid);
            for (int bx = 0; bx < S.nbox; ++bx) {
                const MjBox Bx = S.box[bx];
                const int bw = Bx.i1 - Bx.i0, n = bw * (Bx.j1 - Bx.j0);
                for (int idx = tid; idx < n; idx += nthr) {
                    const int q = (Bx.j0 + idx / bw) * GNX + Bx.i0 + idx % bw;
                    double daf, dtr;
                    if (dir == 0) {
                        daf = fb[q + 1] - fb[q];
                        dtr = uTrans(q + 1) - uTrans(q);
                    } else {
                        daf = fb[q + GNX] - fb[q];
                        dtr = vTrans(q + GNX) - vTrans(q);
                    }
                    loc[q] = loc[q] - coef(q) * (daf - tr[q] * dtr) * mInC[q];
                }
            }
            __syncthreads();
        }
    }

    // ---- cotangent of the output (localTij - tracer) / dt, or of localTij itself (lbar)
    for (int q = tid; q < GPLANE; q += nthr) {
        if (lbar) {
            Lb[q] = lbar[b3 + q];
        } else {
            const double g = gbar[b3 + q];
            Lb[q] = g / dt;
            TB[q] = -(g / dt);
        }
        UB[q] = 0.0;
        VB[q] = 0.0;
    }
    __syncthreads();

    // ---- sweeps in reverse
    for (int s = ns - 1; s >= 0; --s) {
        const int pass = order[s] >> 1, dir = order[s] & 1;
        const MjSweep& S = tp.sweep[pass][dir];
        const int d = dir == 0 ? 1 : GNX;
        double* VBd = dir == 0 ? UB : VB;
        const double* vel = dir == 0 ? u : v;
        const double* rdC = dir == 0 ? rdxC : rdyC;
        const double* len = dir == 0 ? dyG : dxG;
        const double* hF = dir == 0 ? hW : hS;
        const double* sn = snap + (size_t)s * GPLANE;
        for (int q = tid; q < GPLANE; q += nthr) loc[q] = sn[q];
        __syncthreads();
        if (S.fill_before) corner_fill(loc, a, tp, S.fill_before, tid);
        auto trans = [&](int q) { return dir == 0 ? uTrans(q) : vTrans(q); };
        // box update: loc[q] -= c_q (af[q+d] - af[q] - tr[q] (T[q+d] - T[q])), c_q = coef(q) maskInC(q)
        for (int p = tid; p < GPLANE; p += nthr) {
            double f = 0.0, tb = 0.0;
            if (S.nbox) {
                const bool inp = in_box(S, p), inm = p - d >= 0 && in_box(S, p - d);
                if (inp) {
                    const double cg = coef(p) * mInC[p] * Lb[p];
                    f = f + cg;
                    tb = tb - cg * tr[p];
                    TB[p] = TB[p] + cg * (trans(p + d) - trans(p));
                }
                if (inm) {
                    const double cg = coef(p - d) * mInC[p - d] * Lb[p - d];
                    f = f - cg;
                    tb = tb + cg * tr[p - d];
                }
            }
            fb[p] = f;
            // transport cotangent -> velocity: T = vel (len deepFacC drF hFac) rhoFacC
            if (tb != 0.0) VBd[p] = VBd[p] + tb * ((len[p] * deepFacC * drF * hF[p]) * rhoFacC);
        }
        __syncthreads();
        if (S.fill_after) corner_fill_T(Lb, a, tp, S.fill_after, tid);
        // flux: Lb[p] += sum_k w_k(q) fb[q] over the fluxes q = p + 2d, p + d, p, p - d that read p
        for (int p = tid; p < GPLANE; p += nthr) {
            double acc = 0.0;
            FluxPartials P;
            for (int kk = 0; kk < 4; ++kk) {
                const int q = p + (2 - kk) * d;  // loc index p is t_m2 (kk=0) .. t_p1 (kk=3) of flux q
                if (q < 0 || q >= GPLANE || fb[q] == 0.0) continue;
                if (partials(dir, q, P)) acc = acc + P.w[kk] * fb[q];
            }
            Lb[p] = Lb[p] + acc;
            if (fb[p] != 0.0 && partials(dir, p, P)) {
                const double g = fb[p];
                const double dT = (len[p] * deepFacC * drF * hF[p]) * rhoFacC;
                const double arg = vel[p] * dt * rdC[p] * rdeepC;
                const double sgn = arg > 0.0 ? 1.0 : (arg < 0.0 ? -1.0 : 0.0);
                VBd[p] = VBd[p] + (P.dtr * g) * dT + (P.dc * g) * (sgn * dt * rdC[p] * rdeepC);
            }
        }
        __syncthreads();
        if (S.fill_before) corner_fill_T(Lb, a, tp, S.fill_before, tid);
    }
    for (int q = tid; q < GPLANE; q += nthr) TB[q] = TB[q] + Lb[q];
}

// Vertical DST3 (R) sweep after the horizontal passes, per column (member, tile, j, i), k = 0..nlev-1 downward.
// Interface k (0..K) is the top of cell k; w(k) = wFld at interface k, positive upward; the transport is
// rTrans(k) = w(k) * rA * mk(k), mk(k) = maskC(k-1) * maskC(k) for 1 <= k <= K-1, and 0 at the surface (k = 0)
// and the bottom (k = K). The interface value is the DST3 formula with the upstream side below for rTrans > 0:
//   Rj = (t(k-1) - t(k)) mk(k), Rjm = (t(k) - t(k+1)) mk(k+1), Rjp = (t(k-2) - t(k-1)) mk(k-1),
//   c = |w(k) dt(k) recip_drC(k)|, d0 = (2-c)(1-c)/6, d1 = (1-c^2)/6,
//   wT(k) = (rT+|rT|)/2 (t(k) + d0 Rj + d1 Rjm) + (rT-|rT|)/2 (t(k-1) - (d0 Rj + d1 Rjp)),
// with t = localTij after the horizontal passes. Update (the horizontal form, net outflow wT(k) - wT(k+1)):
//   local(k) <- local(k) - coef(k) ((wT(k) - wT(k+1)) - tracer(k) (rTrans(k) - rTrans(k+1))) maskC(k),
// gTracer = (local - tracer)/dt(k) on the whole tile array.
__global__ void __launch_bounds__(256) gad_vert_kernel(GadDev a, const double* __restrict__ loc_h) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int mt = blockIdx.y;
    if (q >= GPLANE) return;
    const int t = mt % a.T, m = mt / a.T;
    const int K = a.K;
    const int kmax = a.nlev ? a.nlev[m] : K;
    const int64_t b2 = (int64_t)t * GPLANE + q;
    const double rA = a.rA[b2], rrA = a.recip_rA[b2];
    const double* __restrict__ w = a.w + m * a.ms3 + (int64_t)t * K * GPLANE + q;
    const double* __restrict__ tr = a.tr + m * a.ms3 + (int64_t)t * K * GPLANE + q;
    const double* __restrict__ rhC = a.recip_hFacC + m * a.ms3 + (int64_t)t * K * GPLANE + q;
    const double* __restrict__ lh = loc_h + m * a.ms3 + (int64_t)t * K * GPLANE + q;
    const double* __restrict__ mc = a.maskC + mask_slot(a, m) * a.mms + (int64_t)t * K * GPLANE + q;
    double* __restrict__ out = a.out + m * a.ms3 + (int64_t)t * K * GPLANE + q;
    // the column as a window over k: T(k-1) .. T(k+2), maskC(k-1) .. maskC(k+2) (0 outside 0 .. kmax-1); each level
    // loads one new localTij, maskC and w value (the expressions are those of the formulas above)
    auto Tk = [&](int k) { return (k >= 0 && k < kmax) ? lh[(int64_t)k * GPLANE] : 0.0; };
    auto Mk = [&](int k) { return (k >= 0 && k < kmax) ? mc[(int64_t)k * GPLANE] : 0.0; };
    double Tm1 = 0.0, T0 = Tk(0), Tp1 = Tk(1), Tp2 = Tk(2);
    double Mm1 = 0.0, M0 = Mk(0), Mp1 = Mk(1), Mp2 = Mk(2);
    // interface k + 1 from the window centred on level k: mk(k+1) = M0 Mp1, mk(k+2) = Mp1 Mp2, mk(k) = Mm1 M0
    auto iface = [&](int k, double& wTk, double& rTk) {
        wTk = 0.0;
        rTk = 0.0;
        const int ki = k + 1;
        if (ki < 1 || ki > kmax - 1) return;
        const double wk = w[(int64_t)ki * GPLANE];
        const double mk1 = M0 * Mp1;
        rTk = wk * rA * mk1;
        const double cfl = fabs(wk * a.dTlev[ki] * a.recip_drC[ki]);
        // upstream for positive (upward) transport: cell k+1, then k+2; downstream: k, then k-1
        wTk = dst3_flux(rTk, cfl, a.oneSixth, Tp2, Tp1, T0, Tm1, Mp1 * Mp2, mk1, Mm1 * M0);
    };
    double wTop = 0.0, rTop = 0.0;  // interface 0: the surface
    for (int k = 0; k < kmax; ++k) {
        double wBot, rBot;
        iface(k, wBot, rBot);
        const double trk = tr[(int64_t)k * GPLANE];
        const double cf = a.dTlev[k] * a.recip_rhoFacC[k] * rhC[(int64_t)k * GPLANE] * a.recip_drF[k] * rrA *
                          a.recip_deepFac2C[k];
        const double loc = T0 - cf * ((wTop - wBot) - trk * (rTop - rBot)) * M0;
        out[(int64_t)k * GPLANE] = (loc - trk) / a.dTlev[k];
        wTop = wBot;
        rTop = rBot;
        Tm1 = T0;
        T0 = Tp1;
        Tp1 = Tp2;
        Tp2 = Tk(k + 3);
        Mm1 = M0;
        M0 = Mp1;
        Mp1 = Mp2;
        Mp2 = Mk(k + 3);
    }
}

// Reverse mode of gad_vert_kernel for one column (member, tile, q): from gbar, the cotangent of gTracer, and loc_h,
// the forward's localTij after the horizontal passes, it writes every level of the column of
//   locbar = the cotangent of loc_h (the identity term gbar/dt plus the interface values' tracer weights),
//   trb    = the vertical part of the tracer cotangent: -gbar/dt + e(k) (rTrans(k) - rTrans(k+1)),
//   wb     = the cotangent of w at interfaces 1 .. nlev-1 (0 at the surface and below the member's levels),
// with e(k) = coef(k) maskC(k) gbar(k)/dt(k). Through the update of level k the interface value wT(k) has the
// cotangent e(k-1) - e(k) and rTrans(k) e(k) tracer(k) - e(k-1) tracer(k-1); the interface value's partials
// (dst3_partials: the four tracer weights, d/dtransport, d/dcfl) carry them on to loc_h and to w through
// rTrans = w rA mk and cfl = |w dt recip_drC| (d|x|/dx = sign x, 0 at 0).
__global__ void __launch_bounds__(256) gad_vert_adjoint_kernel(GadDev a, const double* __restrict__ loc_h,
                                                               const double* __restrict__ gbar,
                                                               double* __restrict__ locbar, double* __restrict__ trb,
                                                               double* __restrict__ wb) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int mt = blockIdx.y;
    if (q >= GPLANE) return;
    const int t = mt % a.T, m = mt / a.T;
    const int K = a.K;
    const int kmax = a.nlev ? a.nlev[m] : K;
    const int64_t b2 = (int64_t)t * GPLANE + q;
    const double rA = a.rA[b2], rrA = a.recip_rA[b2];
    const int64_t c3 = m * a.ms3 + (int64_t)t * K * GPLANE + q;
    const double* __restrict__ w = a.w + c3;
    const double* __restrict__ tr = a.tr + c3;
    const double* __restrict__ rhC = a.recip_hFacC + c3;
    const double* __restrict__ lh = loc_h + c3;
    const double* __restrict__ gb = gbar + c3;
    const double* __restrict__ mc = a.maskC + mask_slot(a, m) * a.mms + (int64_t)t * K * GPLANE + q;
    double* __restrict__ lb = locbar + c3;
    double* __restrict__ tb = trb + c3;
    double* __restrict__ wbar = wb + c3;
    auto L = [](int k) { return (int64_t)k * GPLANE; };
    auto Tk = [&](int k) { return (k >= 0 && k < kmax) ? lh[L(k)] : 0.0; };
    auto Mk = [&](int k) { return (k >= 0 && k < kmax) ? mc[L(k)] : 0.0; };
    auto mk = [&](int k) { return (k >= 1 && k <= kmax - 1) ? Mk(k - 1) * Mk(k) : 0.0; };
    auto rT = [&](int k) { return (k >= 1 && k <= kmax - 1) ? w[L(k)] * rA * mk(k) : 0.0; };
    auto e = [&](int k) {
        if (k < 0 || k >= kmax) return 0.0;
        const double cf = a.dTlev[k] * a.recip_rhoFacC[k] * rhC[L(k)] * a.recip_drF[k] * rrA * a.recip_deepFac2C[k];
        return cf * Mk(k) * (gb[L(k)] / a.dTlev[k]);
    };
    for (int k = 0; k < K; ++k) {
        double l = 0.0, tv = 0.0;
        if (k < kmax) {
            const double g = gb[L(k)];
            l = g / a.dTlev[k];
            tv = -(g / a.dTlev[k]) + e(k) * (rT(k) - rT(k + 1));
        }
        lb[L(k)] = l;
        tb[L(k)] = tv;
        wbar[L(k)] = 0.0;
    }
    for (int ki = 1; ki <= kmax - 1; ++ki) {
        const double wk = w[L(ki)];
        const double wTb = e(ki - 1) - e(ki);
        const double arg = wk * a.dTlev[ki] * a.recip_drC[ki];
        const FluxPartials P = dst3_partials(rT(ki), fabs(arg), a.oneSixth, Tk(ki + 1), Tk(ki), Tk(ki - 1),
                                             Tk(ki - 2), mk(ki + 1), mk(ki), mk(ki - 1));
        if (ki + 1 < kmax) lb[L(ki + 1)] = lb[L(ki + 1)] + P.w[0] * wTb;
        lb[L(ki)] = lb[L(ki)] + P.w[1] * wTb;
        lb[L(ki - 1)] = lb[L(ki - 1)] + P.w[2] * wTb;
        if (ki - 2 >= 0) lb[L(ki - 2)] = lb[L(ki - 2)] + P.w[3] * wTb;
        const double rTb = e(ki) * tr[L(ki)] - e(ki - 1) * tr[L(ki - 1)] + P.dtr * wTb;
        const double sgn = arg > 0.0 ? 1.0 : (arg < 0.0 ? -1.0 : 0.0);
        wbar[L(ki)] = rTb * (rA * mk(ki)) + (P.dc * wTb) * (sgn * a.dTlev[ki] * a.recip_drC[ki]);
    }
}

}  // namespace mj

using namespace mj;

mj_gad_plan::~mj_gad_plan() {
    if (d_tiles) (void)hipFree(d_tiles);
}

// ------------------------------------------------------------------------------------------------ plan (host)
// Per-facet pass table of gad_advection.F:333-351 (reference gad.py:217-235) and the update regions / fill calls
// of :376-774 (gad.py:339-375), as boxes. Array index of Fortran i is i - 1 + OLx.
static void pass_table(int face, int pass, bool* overlapOnly, bool* interiorOnly, bool* cX, bool* cY) {
    auto in = [&](int a, int b, int c) { return face == a || face == b || face == c; };
    if (pass == 1) {
        *overlapOnly = face % 3 == 0;
        *interiorOnly = face % 3 != 0;
        *cX = in(6, 1, 2);
        *cY = in(3, 4, 5);
    } else if (pass == 2) {
        *overlapOnly = face % 3 == 2;
        *interiorOnly = face % 3 == 1;
        *cX = in(2, 3, 4);
        *cY = in(5, 6, 1);
    } else {
        *overlapOnly = false;
        *interiorOnly = true;
        *cX = face == 5 || face == 6;
        *cY = face == 2 || face == 3;
    }
}

// Fills P's host tables from the topology (no device work).
static int build_plan_host(mj_gad_plan* P, int ntiles, const int32_t* face, const uint8_t* edges, int sNx, int sNy,
                           int OLx, int OLy) {
    P->ntiles = ntiles;
    P->sNx = sNx;
    P->sNy = sNy;
    P->OLx = OLx;
    P->OLy = OLy;
    P->tiles.resize(ntiles);
    auto ix = [&](int i) { return (int16_t)(i - 1 + OLx); };
    auto jx = [&](int j) { return (int16_t)(j - 1 + OLy); };
    for (int t = 0; t < ntiles; ++t) {
        const bool N = edges[4 * t] != 0, S = edges[4 * t + 1] != 0, E = edges[4 * t + 2] != 0,
                   W = edges[4 * t + 3] != 0;
        MjTilePlan& tp = P->tiles[t];
        memset(&tp, 0, sizeof(tp));
        tp.sw = W && S;  // fill_cs_corner_tr_rl.F:78-85
        tp.se = E && S;
        tp.nw = W && N;
        tp.ne = E && N;
        for (int pass = 1; pass <= 3; ++pass) {
            bool ov, io, cX, cY;
            pass_table(face[t], pass, &ov, &io, &cX, &cY);
            MjSweep& X = tp.sweep[pass - 1][0];
            MjSweep& Y = tp.sweep[pass - 1][1];
            if (cX) {  // gad_advection.F:376-565
                const bool doflux = !ov || N || S;
                X.fill_before = (doflux && ov) ? 1 : 0;
                X.fill_after = (doflux && ov && pass == 1) ? 2 : 0;
                if (ov) {
                    int iMin = 1 - OLx + 1, iMax = sNx + OLx - 1;
                    if (W) iMin = 1;
                    if (E) iMax = sNx;
                    if (S) X.box[X.nbox++] = MjBox{ix(iMin), (int16_t)(ix(iMax) + 1), jx(1 - OLy), (int16_t)(jx(0) + 1)};
                    if (N)
                        X.box[X.nbox++] =
                            MjBox{ix(iMin), (int16_t)(ix(iMax) + 1), jx(sNy + 1), (int16_t)(jx(sNy + OLy) + 1)};
                } else {
                    int jMin = 1 - OLy, jMax = sNy + OLy;
                    if (io && S) jMin = 1;
                    if (io && N) jMax = sNy;
                    X.box[X.nbox++] =
                        MjBox{ix(1 - OLx + 1), (int16_t)(ix(sNx + OLx - 1) + 1), jx(jMin), (int16_t)(jx(jMax) + 1)};
                }
                X.run = X.nbox > 0 || X.fill_before || X.fill_after;
            }
            if (cY) {  // gad_advection.F:585-774
                const bool doflux = !ov || E || W;
                Y.fill_before = (doflux && ov) ? 2 : 0;
                Y.fill_after = (doflux && ov && pass == 1) ? 1 : 0;
                if (ov) {
                    int jMin = 1 - OLy + 1, jMax = sNy + OLy - 1;
                    if (S) jMin = 1;
                    if (N) jMax = sNy;
                    if (W) Y.box[Y.nbox++] = MjBox{ix(1 - OLx), (int16_t)(ix(0) + 1), jx(jMin), (int16_t)(jx(jMax) + 1)};
                    if (E)
                        Y.box[Y.nbox++] =
                            MjBox{ix(sNx + 1), (int16_t)(ix(sNx + OLx) + 1), jx(jMin), (int16_t)(jx(jMax) + 1)};
                } else {
                    int iMin = 1 - OLx, iMax = sNx + OLx;
                    if (io && W) iMin = 1;
                    if (io && E) iMax = sNx;
                    Y.box[Y.nbox++] =
                        MjBox{ix(iMin), (int16_t)(ix(iMax) + 1), jx(1 - OLy + 1), (int16_t)(jx(sNy + OLy - 1) + 1)};
                }
                Y.run = Y.nbox > 0 || Y.fill_before || Y.fill_after;
            }
        }
    }
    return MJ_OK;
}

// What gad_hz_kernel relies on: the update boxes of a sweep are disjoint and lie in the plane together with the
// face q + d of each of their points (so a sweep has at most GPLANE points, kHzPts per thread).
static int check_boxes(const mj_gad_plan& P) {
    for (int t = 0; t < P.ntiles; ++t)
        for (int pass = 0; pass < 3; ++pass)
            for (int dir = 0; dir < 2; ++dir) {
                const MjSweep& S = P.tiles[t].sweep[pass][dir];
                bool ok = S.nbox >= 0 && S.nbox <= 2;
                for (int b = 0; ok && b < S.nbox; ++b) {
                    const MjBox& B = S.box[b];
                    ok = B.i0 >= 0 && B.j0 >= 0 && B.i0 < B.i1 && B.j0 < B.j1 && B.i1 + (dir == 0) <= GNX &&
                         B.j1 + (dir == 1) <= GNY;
                }
                if (ok && S.nbox == 2) {
                    const MjBox &A = S.box[0], &B = S.box[1];
                    ok = A.i1 <= B.i0 || B.i1 <= A.i0 || A.j1 <= B.j0 || B.j1 <= A.j0;
                }
                if (!ok) {
                    set_error("gad plan: tile %d pass %d dir %d: update boxes overlap or leave the plane", t, pass + 1,
                              dir);
                    return MJ_ERR_ARG;
                }
            }
    return MJ_OK;
}

extern "C" int mj_gad_plan_create(int ntiles, const int32_t* face, const uint8_t* edges, int sNx, int sNy, int OLx,
                                  int OLy, mj_gad_plan** out) {
    MJ_REQUIRE(out && face && edges && ntiles > 0, "mj_gad_plan_create: bad argument");
    MJ_REQUIRE(sNx + 2 * OLx == GNX && sNy + 2 * OLy == GNY && OLx == OLy,
               "mj_gad_plan_create: the kernels are built for %dx%d tiles with equal halos (got sNx %d sNy %d OL %d,%d)",
               GNY, GNX, sNx, sNy, OLx, OLy);
    mj_gad_plan* P = new mj_gad_plan();
    build_plan_host(P, ntiles, face, edges, sNx, sNy, OLx, OLy);
    if (check_boxes(*P) != MJ_OK) {
        delete P;
        return MJ_ERR_ARG;
    }
    hipError_t e = hipMalloc((void**)&P->d_tiles, sizeof(MjTilePlan) * ntiles);
    if (e == hipSuccess)
        e = hipMemcpy(P->d_tiles, P->tiles.data(), sizeof(MjTilePlan) * ntiles, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        delete P;
        set_error("mj_gad_plan_create: %s", hipGetErrorString(e));
        return MJ_ERR_HIP;
    }
    *out = P;
    return MJ_OK;
}

extern "C" int mj_gad_plan_destroy(mj_gad_plan* plan) {
    delete plan;
    return MJ_OK;
}

static int describe(const mj_gad_plan* P, uint8_t* masks, int32_t* fills);

extern "C" int mj_gad_plan_describe(const mj_gad_plan* P, uint8_t* masks, int32_t* fills) {
    MJ_REQUIRE(P, "mj_gad_plan_describe: null plan");
    return describe(P, masks, fills);
}

extern "C" int mj_gad_tables_host(int ntiles, const int32_t* face, const uint8_t* edges, int sNx, int sNy, int OLx,
                                  int OLy, uint8_t* masks, int32_t* fills) {
    MJ_REQUIRE(face && edges && ntiles > 0, "mj_gad_tables_host: bad argument");
    MJ_REQUIRE(sNx + 2 * OLx == GNX && sNy + 2 * OLy == GNY && OLx == OLy, "mj_gad_tables_host: unsupported layout");
    mj_gad_plan P;
    P.ntiles = ntiles;
    P.sNx = sNx;
    P.sNy = sNy;
    P.OLx = OLx;
    P.OLy = OLy;
    build_plan_host(&P, ntiles, face, edges, sNx, sNy, OLx, OLy);
    if (check_boxes(P) != MJ_OK) return MJ_ERR_ARG;
    return describe(&P, masks, fills);
}

static int describe(const mj_gad_plan* P, uint8_t* masks, int32_t* fills) {
    const int T = P->ntiles, plane = GPLANE;
    if (masks) memset(masks, 0, (size_t)6 * T * plane);
    for (int t = 0; t < T; ++t)
        for (int pass = 0; pass < 3; ++pass)
            for (int dir = 0; dir < 2; ++dir) {
                const MjSweep& S = P->tiles[t].sweep[pass][dir];
                if (masks)
                    for (int b = 0; b < S.nbox; ++b)
                        for (int j = S.box[b].j0; j < S.box[b].j1; ++j)
                            for (int i = S.box[b].i0; i < S.box[b].i1; ++i)
                                masks[((size_t)(pass * 2 + dir) * T + t) * plane + j * GNX + i] = 1;
                if (fills) {
                    fills[(pass * 4 + dir * 2 + 0) * T + t] = S.fill_before;
                    fills[(pass * 4 + dir * 2 + 1) * T + t] = S.fill_after;
                }
            }
    return MJ_OK;
}

namespace {
size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }
// Mask slots: one shared set (mask_member_stride 0), one per member, or mask_slots sets picked by mask_index.
int mask_slots_of(const mj_gad_config* c) {
    return c->mask_member_stride == 0 ? 1 : (c->mask_index ? c->mask_slots : c->batch);
}
// The bit planes of every (mask slot, tile, level), then the packed 2-D metrics.
size_t mask_bits_bytes(const mj_gad_config* c, int ntiles) {
    return (size_t)mask_slots_of(c) * ntiles * c->nr * 2 * kMaskWords * sizeof(unsigned long long);
}
size_t pre_bytes(const mj_gad_config* c, int ntiles) {
    return al256(mask_bits_bytes(c, ntiles)) + al256((size_t)3 * ntiles * GPLANE * sizeof(double2));
}
// One member-strided FP64 3-D array of the batch (the R flux's localTij, the reverse mode's replay planes).
size_t batch3_bytes(const mj_gad_config* c, int ntiles) {
    return (size_t)c->batch * (size_t)std::max<int64_t>(c->member_stride3d, (int64_t)ntiles * c->nr * GPLANE) *
           sizeof(double);
}
// The reverse mode's part of its workspace (before the bit planes).
size_t adjoint_scratch_bytes(const mj_gad_config* c, int ntiles, int with_w) {
    size_t n = (size_t)kAdjChunk * 7 * GPLANE * sizeof(double);  // the replay scratch of the horizontal reverse
    if (with_w) n += 2 * batch3_bytes(c, ntiles);
    return al256(n);
}
// The per-call tables at `pre` (pre_bytes): the bit planes of maskLocW / maskLocS for every (mask slot, tile,
// level) and the packed 2-D metrics; sets a.mbits and a.g2.
int launch_pre(GadDev& a, const mj_gad_config* c, void* pre, hipStream_t st) {
    const int nslots = mask_slots_of(c);
    a.mbits = (unsigned long long*)pre;
    a.g2 = (double2*)((char*)pre + al256(mask_bits_bytes(c, a.T)));
    hipLaunchKernelGGL(gad_maskbits_kernel, dim3((unsigned)(nslots * a.T * a.K)), dim3(256), 0, st, a, nslots);
    MJ_HIP(hipGetLastError());
    hipLaunchKernelGGL(gad_pack2d_kernel, dim3((unsigned)((a.T * GPLANE + 255) / 256)), dim3(256), 0, st, a);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}
int check_mask_slots(const mj_gad_config* c, int ntiles) {
    MJ_REQUIRE(c->mask_member_stride == 0 || c->mask_index == nullptr || c->mask_slots >= 1,
               "mj_gad: mask_slots (the number of mask slots) is required with mask_index");
    MJ_REQUIRE((int64_t)mask_slots_of(c) * ntiles * c->nr < (1ll << 31), "mj_gad: too many mask slots");
    return MJ_OK;
}

// Device arguments of the DST3 kernels shared by the forward and reverse entry points.
GadDev gad_dev(const mj_gad_plan* P, const mj_gad_grid* g, const mj_gad_config* c, const double* uFld,
               const double* vFld, const double* wFld, const double* tracer, const double* hFacW, const double* hFacS,
               const double* recip_hFacC) {
    GadDev a{};
    a.B = c->batch;
    a.T = P->ntiles;
    a.K = c->nr;
    a.OLx = P->OLx;
    a.OLy = P->OLy;
    a.sNx = P->sNx;
    a.sNy = P->sNy;
    a.ms3 = c->member_stride3d;
    a.mms = c->mask_member_stride;
    a.mask_index = c->mask_member_stride ? c->mask_index : nullptr;
    a.lpb = c->levels_per_block > 0 ? c->levels_per_block : 1;
    a.tiles = P->d_tiles;
    a.dyG = g->dyG;
    a.dxG = g->dxG;
    a.recip_dxC = g->recip_dxC;
    a.recip_dyC = g->recip_dyC;
    a.recip_rA = g->recip_rA;
    a.rA = g->rA;
    a.maskInC = g->maskInC;
    a.maskW = g->maskW;
    a.maskS = g->maskS;
    a.maskC = g->maskC;
    a.dTlev = g->dTlev;
    a.drF = g->drF;
    a.recip_drF = g->recip_drF;
    a.rhoFacC = g->rhoFacC;
    a.recip_rhoFacC = g->recip_rhoFacC;
    a.deepFacC = g->deepFacC;
    a.recip_deepFacC = g->recip_deepFacC;
    a.recip_deepFac2C = g->recip_deepFac2C;
    a.recip_drC = g->recip_drC;
    a.oneSixth = g->oneSixth;
    a.u = uFld;
    a.v = vFld;
    a.w = wFld;
    a.tr = tracer;
    a.hFacW = hFacW;
    a.hFacS = hFacS;
    a.recip_hFacC = recip_hFacC;
    a.nlev = c->nlev;
    return a;
}
}  // namespace

extern "C" size_t mj_gad_dst3_workspace_bytes(const mj_gad_config* c, int ntiles, int with_w) {
    if (!c || c->batch < 1 || c->nr < 1 || ntiles < 1) return 0;
    return (with_w ? al256(batch3_bytes(c, ntiles)) : 0) + pre_bytes(c, ntiles);
}

extern "C" int mj_gad_dst3(const mj_gad_plan* P, const mj_gad_grid* g, const mj_gad_config* c, const double* uFld,
                           const double* vFld, const double* wFld, const double* tracer, const double* hFacW,
                           const double* hFacS, const double* recip_hFacC, double* gTracer, void* work,
                           size_t work_bytes, void* stream) {
    MJ_REQUIRE(P && g && c, "mj_gad_dst3: null plan/grid/config");
    MJ_REQUIRE(uFld && vFld && tracer && hFacW && hFacS && recip_hFacC && gTracer, "mj_gad_dst3: null field");
    MJ_REQUIRE(g->dyG && g->dxG && g->recip_dxC && g->recip_dyC && g->recip_rA && g->maskInC && g->maskW && g->maskS,
               "mj_gad_dst3: null grid field");
    MJ_REQUIRE(g->dTlev && g->drF && g->recip_drF && g->rhoFacC && g->recip_rhoFacC && g->deepFacC &&
                   g->recip_deepFacC && g->recip_deepFac2C,
               "mj_gad_dst3: null vertical grid field");
    MJ_REQUIRE(c->batch >= 1 && c->nr >= 1, "mj_gad_dst3: batch and nr must be >= 1");
    MJ_REQUIRE(c->member_stride3d >= (int64_t)P->ntiles * c->nr * GPLANE || c->batch == 1,
               "mj_gad_dst3: member_stride3d too small");
    MJ_REQUIRE((int64_t)c->batch * P->ntiles * c->nr < (1ll << 31), "mj_gad_dst3: too many (member, tile, level)s");
    int rc = check_mask_slots(c, P->ntiles);
    if (rc) return rc;
    if (wFld) MJ_REQUIRE(g->maskC && g->rA && g->recip_drC, "mj_gad_dst3: the R flux needs maskC, rA, recip_drC");
    const size_t need = mj_gad_dst3_workspace_bytes(c, P->ntiles, wFld != nullptr);
    MJ_REQUIRE(work && work_bytes >= need, "mj_gad_dst3: work %zu < %zu bytes", work ? work_bytes : 0, need);
    MJ_REQUIRE(((uintptr_t)work & 255) == 0, "mj_gad_dst3: work must be 256-byte aligned");
    {
        const char *w0 = (const char*)work, *w1 = w0 + need;
        const char *g0 = (const char*)gTracer, *g1 = g0 + batch3_bytes(c, P->ntiles);
        MJ_REQUIRE(w1 <= g0 || g1 <= w0, "mj_gad_dst3: work must not overlap gTracer");
    }
    double* loc_h = (double*)work;
    GadDev a = gad_dev(P, g, c, uFld, vFld, wFld, tracer, hFacW, hFacS, recip_hFacC);
    a.out = wFld ? loc_h : gTracer;
    hipStream_t st = as_stream(stream);
    if ((rc = launch_pre(a, c, (char*)work + (wFld ? al256(batch3_bytes(c, P->ntiles)) : 0), st))) return rc;
    const int nkb = (c->nr + a.lpb - 1) / a.lpb;
    const unsigned nblk = (unsigned)((c->batch * P->ntiles * nkb + 7) / 8 * 8);
    if (a.lpb > 1)
        hipLaunchKernelGGL(gad_hz_kernel<true>, dim3(nblk), dim3(kHzThreads), 0, st, a);
    else
        hipLaunchKernelGGL(gad_hz_kernel<false>, dim3(nblk), dim3(kHzThreads), 0, st, a);
    MJ_HIP(hipGetLastError());
    if (wFld) {
        a.out = gTracer;
        hipLaunchKernelGGL(gad_vert_kernel, dim3((GPLANE + 255) / 256, (unsigned)(c->batch * P->ntiles)), dim3(256), 0,
                           st, a, (const double*)loc_h);
        MJ_HIP(hipGetLastError());
    }
    return MJ_OK;
}

extern "C" size_t mj_gad_dst3_adjoint_workspace_bytes(const mj_gad_config* c, int ntiles, int with_w) {
    if (!c || c->batch < 1 || c->nr < 1 || ntiles < 1) return 0;
    return adjoint_scratch_bytes(c, ntiles, with_w) + pre_bytes(c, ntiles);
}

extern "C" int mj_gad_dst3_adjoint(const mj_gad_plan* P, const mj_gad_grid* g, const mj_gad_config* c,
                                   const double* uFld, const double* vFld, const double* wFld, const double* tracer,
                                   const double* hFacW, const double* hFacS, const double* recip_hFacC,
                                   const double* gTracer_bar, double* tracer_bar, double* uFld_bar, double* vFld_bar,
                                   double* wFld_bar, double* work, size_t work_bytes, void* stream) {
    MJ_REQUIRE(P && g && c, "mj_gad_dst3_adjoint: null plan/grid/config");
    MJ_REQUIRE(uFld && vFld && tracer && hFacW && hFacS && recip_hFacC && gTracer_bar && tracer_bar && uFld_bar &&
                   vFld_bar && work,
               "mj_gad_dst3_adjoint: null field");
    MJ_REQUIRE(g->dyG && g->dxG && g->recip_dxC && g->recip_dyC && g->recip_rA && g->maskInC && g->maskW && g->maskS,
               "mj_gad_dst3_adjoint: null grid field");
    MJ_REQUIRE(c->batch >= 1 && c->nr >= 1, "mj_gad_dst3_adjoint: batch and nr must be >= 1");
    MJ_REQUIRE(c->member_stride3d >= (int64_t)P->ntiles * c->nr * GPLANE || c->batch == 1,
               "mj_gad_dst3_adjoint: member_stride3d too small");
    MJ_REQUIRE((int64_t)c->batch * P->ntiles * c->nr < (1ll << 31), "mj_gad_dst3_adjoint: too many blocks");
    if (wFld) {
        MJ_REQUIRE(wFld_bar && g->maskC && g->rA && g->recip_drC,
                   "mj_gad_dst3_adjoint: the R flux needs wFld_bar, maskC, rA, recip_drC");
        for (const double* o : {(const double*)tracer_bar, (const double*)uFld_bar, (const double*)vFld_bar})
            MJ_REQUIRE(o != wFld_bar, "mj_gad_dst3_adjoint: wFld_bar aliases another output");
    }
    int rc = check_mask_slots(c, P->ntiles);
    if (rc) return rc;
    const size_t need = mj_gad_dst3_adjoint_workspace_bytes(c, P->ntiles, wFld != nullptr);
    MJ_REQUIRE(work_bytes >= need, "mj_gad_dst3_adjoint: work %zu < %zu bytes", work_bytes, need);
    MJ_REQUIRE(((uintptr_t)work & 255) == 0, "mj_gad_dst3_adjoint: work must be 256-byte aligned");
    hipStream_t st = as_stream(stream);
    const int nblk = c->batch * P->ntiles * c->nr;
    GadDev a = gad_dev(P, g, c, uFld, vFld, nullptr, tracer, hFacW, hFacS, recip_hFacC);
    if ((rc = launch_pre(a, c, (char*)work + adjoint_scratch_bytes(c, P->ntiles, wFld != nullptr), st))) return rc;
    double* lbar = nullptr;
    if (wFld) {
        // forward replay of the horizontal passes (localTij), then the vertical reverse per column
        double* loc_h = work + (size_t)kAdjChunk * 7 * GPLANE;
        lbar = loc_h + batch3_bytes(c, P->ntiles) / sizeof(double);
        GadDev f = gad_dev(P, g, c, uFld, vFld, wFld, tracer, hFacW, hFacS, recip_hFacC);
        f.mbits = a.mbits;
        f.g2 = a.g2;
        f.out = loc_h;
        f.lpb = 1;
        const unsigned nb8 = (unsigned)((nblk + 7) / 8 * 8);
        hipLaunchKernelGGL(gad_hz_kernel<false>, dim3(nb8), dim3(kHzThreads), 0, st, f);
        MJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(gad_vert_adjoint_kernel, dim3((GPLANE + 255) / 256, (unsigned)(c->batch * P->ntiles)),
                           dim3(256), 0, st, f, (const double*)loc_h, gTracer_bar, lbar, tracer_bar, wFld_bar);
        MJ_HIP(hipGetLastError());
    }
    for (int b0 = 0; b0 < nblk; b0 += kAdjChunk) {
        const int n = std::min(kAdjChunk, nblk - b0);
        hipLaunchKernelGGL(gad_hz_adjoint_kernel, dim3((unsigned)n), dim3(1024), 0, st, a, gTracer_bar, tracer_bar,
                           uFld_bar, vFld_bar, work, b0, (const double*)lbar);
        MJ_HIP(hipGetLastError());
    }
    return MJ_OK;
}
