// Device building blocks of the CG2D solves (FP64 and mixed precision): the 5-point stencils in the reference's
// term order, the fixed row/tile reduction trees, and the row sweeps over column pairs.
#pragma once

#include "mj_common.h"

namespace mj {

constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin over them

// ------------------------------------------------------------------------------------------------ stencils
// apply_operator (core/cg2d.py:181-188): aW(i)x(i-1) + aW(i+1)x(i+1) + aS(j)x(j-1) + aS(j+1)x(j+1) + aC x
template <typename T>
__device__ __forceinline__ T op_A(const T* aW, const T* aS, const T* aC, const T* x, int p, int nx) {
    return aW[p] * x[p - 1] + aW[p + 1] * x[p + 1] + aS[p] * x[p - nx] + aS[p + nx] * x[p + nx] + aC[p] * x[p];
}
// apply_preconditioner (core/cg2d.py:191-198): pC r + pW(i)r(i-1) + pW(i+1)r(i+1) + pS(j)r(j-1) + pS(j+1)r(j+1)
template <typename T>
__device__ __forceinline__ T op_M(const T* pW, const T* pS, const T* pC, const T* r, int p, int nx) {
    return pC[p] * r[p] + pW[p] * r[p - 1] + pW[p + 1] * r[p + 1] + pS[p] * r[p - nx] + pS[p + nx] * r[p + nx];
}

// ------------------------------------------------------------------------------------------------ reductions
// Reduction of per-point values over a tile interior, MJ_SUM_TREE: each wave reduces whole rows with wave_tree
// (lane l holds i = l and i = l + 64) into a row-sum array in LDS, then one wave reduces the row sums.
// MJ_SUM_FORTRAN: per-point values are staged in global memory and added in j/i order by lane 0 of one wave,
// which brings them into LDS 64 at a time (tile_of).

// Row loop: wave w of the block takes rows jl = w, w + nwaves, ...; p0 = (jl, il = lane), p1 = (jl, lane + 64).
// Unrolled so that a wave has the loads of several rows in flight at once.
#ifndef MJ_ROW_UNROLL
#define MJ_ROW_UNROLL 2
#endif
#ifndef MJ_SWEEP_UNROLL
#define MJ_SWEEP_UNROLL 1
#endif  // unroll of the row sweeps (more rows' loads in flight per wave)
#define MJ_PRAGMA(x) _Pragma(#x)
#define MJ_UNROLL(n) MJ_PRAGMA(unroll n)
#define MJ_ROWS(...)                                                                                      \
    MJ_UNROLL(MJ_ROW_UNROLL)                                                                              \
    for (int jl = wave; jl < sNy; jl += nwaves) {                                                         \
        const int prow = tbase + (OL + jl) * nx + OL;                                                     \
        const int p0 = prow + lane;                                                                       \
        const bool has1 = lane + 64 < sNx;                                                                \
        const int p1 = prow + lane + 64;                                                                  \
        __VA_ARGS__                                                                                       \
    }

#define MJ_UNIT_IDS                                                                                       \
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;              \
    const int sNx = a.sNx, sNy = a.sNy, OL = a.OL, nx = a.nx, tbase = t * a.plane;                        \
    (void)tid;                                                                                            \
    (void)sNx;                                                                                            \
    (void)OL;                                                                                             \
    (void)nx;                                                                                             \
    (void)tbase;

#ifndef MJ_GROUP
#define MJ_GROUP 1
#endif
// Members that share an operator slot swept together by one block (each coefficient row loaded once for all of
// them). Measured on MI355X at 128 members the per-member blocks (1) are fastest: a group of 4 cuts the bytes the
// sweeps request per member by a third but lengthens each wave's dependent row chain more than that saves.
constexpr int kGroup = MJ_GROUP;

// Per-block LDS: row sums of two reductions per group member (slot r + 2k), values broadcast from thread 0.
static __shared__ double g_rows[2 * kGroup * 128];
static __shared__ double g_bcast[6];
static __shared__ int g_ibcast;

// Store one row's per-point values of reduction r: tree order -> wave_tree into the row sums; fortran order -> the
// unit's global staging area (stage[r][jl][il]).
template <int MODE>
__device__ __forceinline__ void put_row(int r, int jl, int lane, double v0, double v1, bool has1, double* stage,
                                        int sNx, int sNy) {
    if (MODE == MJ_SUM_TREE) {
        const double w = wave_tree(v0, v1, has1);
        if (lane == 0) g_rows[r * 128 + jl] = w;
    } else {
        double* pr = stage + ((size_t)r * sNy + jl) * sNx;
        pr[lane] = v0;
        if (has1) pr[lane + 64] = v1;
    }
}

// After a __syncthreads: tile partial of reduction r of group member k (valid in lane 0 of the calling wave).
template <int MODE>
__device__ __forceinline__ double tile_of(int r, int lane, const double* stage, int sNx, int sNy, int k = 0) {
    if (MODE == MJ_SUM_TREE) {
        const double* rb = g_rows + (r + 2 * k) * 128;
        const bool h = lane + 64 < sNy;
        return wave_tree(rb[lane], h ? rb[lane + 64] : 0.0, h);
    }
    // fortran order: the strict sequence s = ((0 + v[0]) + v[1]) + ... over the staged values in j/i order. The
    // wave loads them 64 at a time (coalesced, the next 64 in flight) into its own slot of the row-sum array (unused
    // in this order: reduction r of group member k takes doubles (2k + r) * 64 ..), and lane 0 adds them from there:
    // only the chain of additions is serial. One wave writes and reads its slot, and a wave's LDS accesses complete
    // in order, so no barrier is needed.
    double* slot = g_rows + (2 * k + r) * 64;
    const double* pr = stage + (size_t)r * sNy * sNx;
    const int n = sNy * sNx;
    double s = 0.0;
    double nxt = lane < n ? pr[lane] : 0.0;
    for (int c = 0; c < n; c += 64) {
        slot[lane] = nxt;
        nxt = c + 64 + lane < n ? pr[c + 64 + lane] : 0.0;
        if (lane == 0) {
            if (c + 64 <= n) {
#pragma unroll
                for (int i = 0; i < 64; ++i) s = s + slot[i];
            } else {
                for (int i = 0; i < n - c; ++i) s = s + slot[i];
            }
        }
    }
    return s;
}

// Reductions 0 and 1 of the unit written to dst0[idx] and dst1[idx] (dst1 NULL: only reduction 0).
template <int MODE>
__device__ __forceinline__ void store_tile_sums(double* dst0, double* dst1, int idx, int lane, int wave, int nwaves,
                                                const double* stage, int sNx, int sNy) {
    __syncthreads();
    const int nr = dst1 ? 2 : 1;
    for (int r = wave; r < nr; r += nwaves) {
        const double v = tile_of<MODE>(r, lane, stage, sNx, sNy);
        if (lane == 0) (r == 0 ? dst0 : dst1)[idx] = v;
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------------------------ row sweeps
// The iteration phases sweep rows: each wave owns a strip of consecutive rows and lane l owns the column pair
// (2l, 2l+1), loaded and stored as one 16-byte access. Down the strip the three rows a 5-point stencil reads are
// kept in registers (one new row loaded per step); the i-1 / i+1 neighbours come from the adjacent lane, and the
// halo columns -1 and sNx from one extra load by the first / last lane. Every value is computed with the same
// expression, in the same term order, as op_A / op_M.
struct Lanes {
    int lane, c, last;  // c: this lane's first column (clamped in range for idle lanes), last: lane of column sNx-1
    bool act;           // lane owns columns
};
__device__ __forceinline__ Lanes lanes_of(int lane, int sNx) {
    Lanes L;
    L.lane = lane;
    L.act = 2 * lane < sNx;
    L.c = L.act ? 2 * lane : sNx - 2;
    L.last = sNx / 2 - 1;
    return L;
}
template <typename T> struct pair_type;
template <> struct pair_type<double> { using type = double2; };
template <> struct pair_type<float> { using type = float2; };
template <typename T> using P2 = typename pair_type<T>::type;

template <typename T>
__device__ __forceinline__ P2<T> ld2(const T* p) { return *reinterpret_cast<const P2<T>*>(p); }
template <typename T>
__device__ __forceinline__ void st2(T* p, P2<T> v) { *reinterpret_cast<P2<T>*>(p) = v; }
// Value at the halo column on this lane's open side (lane 0: column -1, last lane: column sNx), row base prow.
template <typename T>
__device__ __forceinline__ T halo_of(const T* v, int prow, const Lanes& L, int sNx) {
    return (L.lane == 0 || L.lane == L.last) ? v[prow + (L.lane == 0 ? -1 : sNx)] : T(0);
}
// Left neighbour of column c and right neighbour of column c+1.
template <typename T>
__device__ __forceinline__ T left_of(P2<T> v, T h, const Lanes& L) {
    const T w = __shfl_up(v.y, 1, 64);
    return L.lane == 0 ? h : w;
}
template <typename T>
__device__ __forceinline__ T right_of(P2<T> v, T h, const Lanes& L) {
    const T w = __shfl_down(v.x, 1, 64);
    return L.lane == L.last ? h : w;
}
// op_A / op_M at columns c, c+1 of one row: Xm, Xc, Xp rows j-1, j, j+1; xl, xr the outer neighbours.
template <typename T>
__device__ __forceinline__ P2<T> pair_A(P2<T> AW, T awr, P2<T> AS0, P2<T> AS1, P2<T> AC, P2<T> Xm, P2<T> Xc,
                                        P2<T> Xp, T xl, T xr) {
    P2<T> q;
    q.x = AW.x * xl + AW.y * Xc.y + AS0.x * Xm.x + AS1.x * Xp.x + AC.x * Xc.x;
    q.y = AW.y * Xc.x + awr * xr + AS0.y * Xm.y + AS1.y * Xp.y + AC.y * Xc.y;
    return q;
}
template <typename T>
__device__ __forceinline__ P2<T> pair_M(P2<T> PC, P2<T> PW, T pwr, P2<T> PS0, P2<T> PS1, P2<T> Rm, P2<T> Rc,
                                        P2<T> Rp, T rl, T rr) {
    P2<T> q;
    q.x = PC.x * Rc.x + PW.x * rl + PW.y * Rc.y + PS0.x * Rm.x + PS1.x * Rp.x;
    q.y = PC.y * Rc.y + PW.y * Rc.x + pwr * rr + PS0.y * Rm.y + PS1.y * Rp.y;
    return q;
}
// One row's per-point values of reduction r held as column pairs: tree order reduces them with pair_tree (the
// canonical row tree) into the row sums; fortran order stages them.
template <int MODE>
__device__ __forceinline__ void put_row2(int r, int jl, const Lanes& L, double2 v, double* stage, int sNx, int sNy,
                                         int k = 0) {
    if (MODE == MJ_SUM_TREE) {
        const double w = pair_tree(v, sNx, L.lane);
        if (L.lane == 0) g_rows[(r + 2 * k) * 128 + jl] = w;
    } else {
        if (L.act) st2(stage + ((size_t)r * sNy + jl) * sNx + L.c, v);
    }
}

// Strip of rows of this wave: [j0, j1).
#define MJ_STRIP                                                                                         \
    const int rows_ = (sNy + nwaves - 1) / nwaves, j0 = wave * rows_, j1 = min(j0 + rows_, sNy);        \
    const Lanes L = lanes_of(lane, sNx);
#define MJ_ROW(j) (tbase + (OL + (j)) * nx + OL)

// ------------------------------------------------------------------------------------------------ set-up units
// These take any device-argument struct D with the fields b, x0, x, norm, pmax, ms, T, plane, sNx, sNy, OL, nx and
// the full T-map tables full_off/full_dst/full_src/full_sign.
template <class D>
__device__ __forceinline__ double member_rhs_max(const D& a, int m) {
    double mx = 0.0;
    for (int t = 0; t < a.T; ++t) mx = fmax(mx, a.pmax[m * a.T + t]);
    return mx;
}
template <class D>
__device__ __forceinline__ double member_rhs_norm(const D& a, int m) {
    const double mx = member_rhs_max(a, m);
    return mx != 0.0 ? 1.0 / mx : 1.0;
}

// I0: per-tile max |b * cg2dNorm| (cg2d.F:150-160; the maximum is order free).
template <class D>
__device__ __forceinline__ void unit_I0(const D& a, int m, int t) {
    MJ_UNIT_IDS
    const double* b = a.b + m * a.ms;
    const double nm = a.norm[m];
    double mx = 0.0;
    MJ_ROWS({
        mx = fmax(mx, fabs(b[p0] * nm));
        if (has1) mx = fmax(mx, fabs(b[p1] * nm));
    })
    for (int o = 32; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) g_rows[wave] = mx;
    __syncthreads();
    if (tid == 0) {
        double v = 0.0;
        for (int w = 0; w < nwaves; ++w) v = fmax(v, g_rows[w]);
        a.pmax[m * a.T + t] = v;
    }
    __syncthreads();
}

// Halo points of x written by the exchange of the normalised first guess (cg2d.F:165-172): the pre-exchange value
// at the source, normalised where the source is interior.
template <class D>
__device__ __forceinline__ void exchanged_first_guess(const D& a, int m, int t, double rhsNorm) {
    const int plane = a.plane, nx = a.nx, OL = a.OL, sNx = a.sNx, sNy = a.sNy;
    const double* x0 = a.x0 + m * a.ms;
    double* x = a.x + m * a.ms;
    for (int k = a.full_off[t] + threadIdx.x; k < a.full_off[t + 1]; k += blockDim.x) {
        const int s = a.full_src[k];
        const int sq = s % plane, sj = sq / nx, si = sq - sj * nx;
        const bool inner = sj >= OL && sj < OL + sNy && si >= OL && si < OL + sNx;
        const double v = inner ? x0[s] * rhsNorm : x0[s];
        x[a.full_dst[k]] = a.full_sign[k] < 0 ? -v : v;
    }
}

// x = x0 * rhsNorm on the interior, x0 on the halo, then the exchanged halo points; ends with a block barrier.
template <class D>
__device__ __forceinline__ void first_guess(const D& a, int m, int t, double rhsNorm) {
    const int plane = a.plane, nx = a.nx, OL = a.OL, sNx = a.sNx, sNy = a.sNy, tbase = t * plane;
    const double* x0 = a.x0 + m * a.ms;
    double* x = a.x + m * a.ms;
    for (int qq = threadIdx.x; qq < plane; qq += blockDim.x) {
        const int j = qq / nx, i = qq - j * nx, p = tbase + qq;
        const bool inner = j >= OL && j < OL + sNy && i >= OL && i < OL + sNx;
        x[p] = inner ? x0[p] * rhsNorm : x0[p];
    }
    __syncthreads();
    exchanged_first_guess(a, m, t, rhsNorm);
    __syncthreads();
}

// ------------------------------------------------------------------------------------------------ grouped sweeps
// The two CG phases of up to G members that read the same operator slot, swept by one block: each coefficient row
// is loaded once and applied to every member's rows. The recurrence keeps z = M r and q = A s as vectors, as
// cg2d.F does: phase B forms s = z + beta s_old pointwise (ring from the sources), then q = A s by the stencil and
// the partial of sigma = (s, q); phase C forms x += alpha s and r = r_old - alpha q pointwise (ring from the
// sources), then z = M r by the stencil and the partials of err = (r, r) and eta = (z, r). Per member the arithmetic
// is exactly the single-member one (same expressions, same term order, same reduction trees), so results do not
// depend on the grouping.
struct SweepGeo {
    int t, tbase, sNx, sNy, OL, nx;
    const int32_t *ring_off, *ring_dst, *ring_src;
    const int8_t* ring_sign;
};
template <typename T>
struct MemberB {
    const T* z;     // M r_{n-1}
    const T* so;    // s_{n-1} (unused when first)
    T* sn;          // s_n
    T* q;           // out: A s_n
    T beta;
    bool first;     // s_n = z (no s_{n-1} term)
    double* sig;    // where the tile partial of sigma goes
    double* stage;  // MJ_SUM_FORTRAN staging
};
template <typename T>
struct MemberC {
    const T* sv;   // s_n
    const T* q;    // A s_n
    const T* rin;  // r_{n-1}
    T* rout;       // r_n
    T* xs;         // x (FP64) or d (FP32): xs += alpha s
    T* z;          // out: M r_n
    T alpha;
    double* err;
    double* eta;
    double* stage;
};

template <typename T, int G, int MODE>
__device__ __forceinline__ void group_B(const SweepGeo& g, const T* __restrict__ aW, const T* __restrict__ aS,
                                        const T* __restrict__ aC, const MemberB<T> (&mb)[G], int nm) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const int sNx = g.sNx, sNy = g.sNy, OL = g.OL, nx = g.nx, tbase = g.tbase;
    MJ_STRIP
    // s_n = z + beta s_{n-1}: interior rows of the strip, then the ring from the sources
    for (int j = j0; j < j1; ++j) {
        const int p = MJ_ROW(j) + L.c;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            if (k < nm) {
                const P2<T> Z = ld2(mb[k].z + p);
                P2<T> v = Z;
                if (!mb[k].first) {
                    const P2<T> SO = ld2(mb[k].so + p);
                    v.x = Z.x + mb[k].beta * SO.x;
                    v.y = Z.y + mb[k].beta * SO.y;
                }
                if (L.act) st2(mb[k].sn + p, v);
            }
        }
    }
    const int rk0 = g.ring_off[g.t], rk1 = g.ring_off[g.t + 1];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        if (k < nm) {
            for (int kk = rk0 + tid; kk < rk1; kk += blockDim.x) {
                const int s = g.ring_src[kk];
                const T v = s < 0 ? T(0) : mb[k].first ? mb[k].z[s] : mb[k].z[s] + mb[k].beta * mb[k].so[s];
                mb[k].sn[g.ring_dst[kk]] = g.ring_sign[kk] < 0 ? -v : v;
            }
        }
    }
    __syncthreads();
    // q = A s_n, sigma partial (s_n, q)
    if (j0 < j1) {
        int p = MJ_ROW(j0) + L.c;
        P2<T> Xm[G], Xc[G];
#pragma unroll
        for (int k = 0; k < G; ++k)
            if (k < nm) {
                Xm[k] = ld2(mb[k].sn + p - nx);
                Xc[k] = ld2(mb[k].sn + p);
            }
        P2<T> AS0 = ld2(aS + p);
        for (int j = j0; j < j1; ++j, p += nx) {
            const int prow = MJ_ROW(j);
            const P2<T> AS1 = ld2(aS + p + nx), AW = ld2(aW + p), AC = ld2(aC + p);
            const T awr = right_of(AW, halo_of(aW, prow, L, sNx), L);
#pragma unroll
            for (int k = 0; k < G; ++k) {
                if (k < nm) {
                    const P2<T> Xp = ld2(mb[k].sn + p + nx);
                    const T hx = halo_of(mb[k].sn, prow, L, sNx);
                    const P2<T> q = pair_A(AW, awr, AS0, AS1, AC, Xm[k], Xc[k], Xp, left_of(Xc[k], hx, L),
                                           right_of(Xc[k], hx, L));
                    if (L.act) st2(mb[k].q + p, q);
                    double2 v = make_double2((double)Xc[k].x * (double)q.x, (double)Xc[k].y * (double)q.y);
                    if (!L.act) v = make_double2(0.0, 0.0);
                    put_row2<MODE>(0, j, L, v, mb[k].stage, sNx, sNy, k);
                    Xm[k] = Xc[k];
                    Xc[k] = Xp;
                }
            }
            AS0 = AS1;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < G; ++k) {
        if (k < nm && k % nwaves == wave) {
            const double v = tile_of<MODE>(0, lane, mb[k].stage, sNx, sNy, k);
            if (lane == 0) *mb[k].sig = v;
        }
    }
    __syncthreads();
}

template <typename T, int G, int MODE>
__device__ __forceinline__ void group_C(const SweepGeo& g, const T* __restrict__ pW, const T* __restrict__ pS,
                                        const T* __restrict__ pC, const MemberC<T> (&mc)[G], int nm) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const int sNx = g.sNx, sNy = g.sNy, OL = g.OL, nx = g.nx, tbase = g.tbase;
    MJ_STRIP
    // x += alpha s_n, r_n = r_{n-1} - alpha q: interior rows of the strip, then the ring of r_n from the sources
    for (int j = j0; j < j1; ++j) {
        const int p = MJ_ROW(j) + L.c;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            if (k < nm) {
                const P2<T> SV = ld2(mc[k].sv + p), Q = ld2(mc[k].q + p), RI = ld2(mc[k].rin + p),
                            XV = ld2(mc[k].xs + p);
                const T alpha = mc[k].alpha;
                P2<T> xn, rn;
                xn.x = XV.x + alpha * SV.x;
                xn.y = XV.y + alpha * SV.y;
                rn.x = RI.x - alpha * Q.x;
                rn.y = RI.y - alpha * Q.y;
                if (L.act) {
                    st2(mc[k].xs + p, xn);
                    st2(mc[k].rout + p, rn);
                }
            }
        }
    }
    const int rk0 = g.ring_off[g.t], rk1 = g.ring_off[g.t + 1];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        if (k < nm) {
            for (int kk = rk0 + tid; kk < rk1; kk += blockDim.x) {
                const int s = g.ring_src[kk];
                const T v = s < 0 ? T(0) : mc[k].rin[s] - mc[k].alpha * mc[k].q[s];
                mc[k].rout[g.ring_dst[kk]] = g.ring_sign[kk] < 0 ? -v : v;
            }
        }
    }
    __syncthreads();
    // z = M r_n, partials err = (r_n, r_n), eta = (z, r_n)
    if (j0 < j1) {
        int p = MJ_ROW(j0) + L.c;
        P2<T> Rm[G], Rc[G];
#pragma unroll
        for (int k = 0; k < G; ++k)
            if (k < nm) {
                Rm[k] = ld2(mc[k].rout + p - nx);
                Rc[k] = ld2(mc[k].rout + p);
            }
        P2<T> PS0 = ld2(pS + p);
        for (int j = j0; j < j1; ++j, p += nx) {
            const int prow = MJ_ROW(j);
            const P2<T> PS1 = ld2(pS + p + nx), PC = ld2(pC + p), PW = ld2(pW + p);
            const T pwr = right_of(PW, halo_of(pW, prow, L, sNx), L);
#pragma unroll
            for (int k = 0; k < G; ++k) {
                if (k < nm) {
                    const P2<T> Rp = ld2(mc[k].rout + p + nx);
                    const T hr = halo_of(mc[k].rout, prow, L, sNx);
                    const P2<T> zz = pair_M(PC, PW, pwr, PS0, PS1, Rm[k], Rc[k], Rp, left_of(Rc[k], hr, L),
                                            right_of(Rc[k], hr, L));
                    if (L.act) st2(mc[k].z + p, zz);
                    double2 e = make_double2((double)Rc[k].x * (double)Rc[k].x, (double)Rc[k].y * (double)Rc[k].y);
                    double2 h = make_double2((double)zz.x * (double)Rc[k].x, (double)zz.y * (double)Rc[k].y);
                    if (!L.act) {
                        e = make_double2(0.0, 0.0);
                        h = e;
                    }
                    put_row2<MODE>(0, j, L, e, mc[k].stage, sNx, sNy, k);
                    put_row2<MODE>(1, j, L, h, mc[k].stage, sNx, sNy, k);
                    Rm[k] = Rc[k];
                    Rc[k] = Rp;
                }
            }
            PS0 = PS1;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < G; ++k) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            if (k < nm && (2 * k + r) % nwaves == wave) {
                const double v = tile_of<MODE>(r, lane, mc[k].stage, sNx, sNy, k);
                if (lane == 0) *(r == 0 ? mc[k].err : mc[k].eta) = v;
            }
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------------------------ LDS-staged sweeps
// The two CG phases of one member with the stencil operand staged in LDS: the pointwise update writes the new
// vector's interior to global memory (read by the next phase and iteration) and to an LDS tile, the ring is
// computed from its sources into the LDS tile only (no global ring writes: nothing reads them), and the stencil
// reads its neighbours from LDS. The coefficients stream from global memory down each wave's strip of rows (the
// south coefficient row rolls). Every value is the expression group_B / group_C use, in the same term order and
// reduction trees, so the results are bitwise theirs.
//
// LDS tile: rows -1..sNy, columns -1..sNx at lds_at(j, i) = (j + 1) * (sNx + 4) + i + 2 (even offsets for the
// column pairs). Ring entries with no source (open facet edges) carry src < 0 and give 0.
#ifndef MJ_LDS_SWEEPS
#define MJ_LDS_SWEEPS 1
#endif  // 1: the phase launches use the LDS-staged sweeps (0: the global-ring sweeps)
extern __shared__ __align__(16) double mj_lds[];
__host__ __device__ constexpr size_t lds_tile_bytes(int sNx, int sNy, int elem) {
    return (size_t)(sNy + 2) * (sNx + 4) * elem;
}
// Geometry of the mixed solve's FP32 work vectors (interior only) and FP32 coefficient copies. A work vector's
// tile holds sNy rows of CP floats (sNx rounded up to a multiple of 4, so every row starts 16-byte aligned and a
// row is NC float4 chunks; the padding columns always hold 0), then contiguous copies of its west and east columns
// (sNy values each, so that a neighbour's ring read of a column is one contiguous run), rounded up to 128 bytes.
// A coefficient copy's tile holds rows 0..sNy of CPc floats (columns 0..sNx, rounded up to 4): column sNx carries
// the east faces (aW, pW at i = sNx) and row sNy the north faces (aS, pS at j = sNy). The LDS tile of a stencil
// operand holds rows -1..sNy with column i at offset 4 + i of its row (column -1 at 3, the east ring column sNx at
// 4 + sNx) and a row pitch W = 4 + CPc, a multiple of 4.
struct MxGeo {
    int CP, NC, NS, TS, CPc, TSc, W;
    __host__ __device__ int lds(int j, int i) const { return (j + 1) * W + 4 + i; }
};
__host__ __device__ inline MxGeo mx_geo(int sNx, int sNy) {
    MxGeo g;
    g.CP = (sNx + 3) / 4 * 4;
    g.NC = g.CP / 4;
    g.NS = sNy * g.NC;
    g.TS = (sNy * g.CP + 2 * sNy + 31) / 32 * 32;
    g.CPc = (sNx + 1 + 3) / 4 * 4;
    g.TSc = (sNy + 1) * g.CPc;
    g.W = 4 + g.CPc;
    return g;
}
__host__ __device__ inline size_t mx_lds_bytes(int sNx, int sNy) { return (size_t)(sNy + 2) * mx_geo(sNx, sNy).W * 4; }
struct LdsGeo {
    int W;
    __device__ __forceinline__ int at(int j, int i) const { return (j + 1) * W + i + 2; }
};
// Column of this lane's coefficient pair load: the lane after the last one with a column pair loads the east halo
// column sNx (its first value is the last lane's right neighbour), every other lane its own pair.
__device__ __forceinline__ int east_col(const Lanes& L, int sNx) {
    return L.lane == L.last + 1 ? sNx : L.c;
}
// Ring position of a tile-relative halo index r (row-major over ny x nx): LDS offset.
__device__ __forceinline__ int lds_ring(const LdsGeo& G, int r, int nx, int OL) {
    const int jh = r / nx;
    return G.at(jh - OL, r - jh * nx - OL);
}

template <typename T, int MODE>
__device__ __forceinline__ void lds_B(const SweepGeo& g, const T* __restrict__ aW, const T* __restrict__ aS,
                                      const T* __restrict__ aC, const MemberB<T>& mb) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const int sNx = g.sNx, sNy = g.sNy, OL = g.OL, nx = g.nx, tbase = g.tbase;
    T* tile = reinterpret_cast<T*>(mj_lds);
    const LdsGeo G{sNx + 4};
    MJ_STRIP
    const T* __restrict__ z = mb.z;
    const T* __restrict__ so = mb.so;
    T* __restrict__ sn = mb.sn;
    const T beta = mb.beta;
    // s_n = z + beta s_{n-1} on the strip's interior rows, two rows' loads in flight
    for (int j = j0; j < j1; j += 2) {
        const bool two = j + 1 < j1;
        const int p = MJ_ROW(j) + L.c, p2 = two ? p + nx : p;
        const P2<T> Z0 = ld2(z + p), Z1 = ld2(z + p2);
        P2<T> S0 = Z0, S1 = Z1;
        if (!mb.first) {
            const P2<T> O0 = ld2(so + p), O1 = ld2(so + p2);
            S0.x = Z0.x + beta * O0.x;
            S0.y = Z0.y + beta * O0.y;
            S1.x = Z1.x + beta * O1.x;
            S1.y = Z1.y + beta * O1.y;
        }
        if (L.act) {
            st2(sn + p, S0);
            st2(tile + G.at(j, L.c), S0);
            if (two) {
                st2(sn + p2, S1);
                st2(tile + G.at(j + 1, L.c), S1);
            }
        }
    }
    for (int kk = g.ring_off[g.t] + tid; kk < g.ring_off[g.t + 1]; kk += blockDim.x) {
        const int s = g.ring_src[kk];
        T v = T(0);
        if (s >= 0) v = mb.first ? z[s] : z[s] + beta * so[s];
        tile[lds_ring(G, g.ring_dst[kk] - tbase, nx, OL)] = g.ring_sign[kk] < 0 ? -v : v;
    }
    __syncthreads();
    // q = A s_n from the LDS tile; sigma partial (s_n, q)
    if (j0 < j1) {
        const int ce = east_col(L, sNx);
        int pr = MJ_ROW(j0);
        P2<T> AS0 = ld2(aS + pr + L.c);
        for (int j = j0; j < j1; ++j, pr += nx) {
            const int p = pr + L.c;
            const P2<T> AS1 = ld2(aS + p + nx), AC = ld2(aC + p), AW = ld2(aW + pr + ce);
            T awr = __shfl_down(AW.x, 1, 64);
            if (L.last == 63 && lane == 63) awr = aW[pr + sNx];  // sNx = 128: no lane after the last
            const int q0 = G.at(j, L.c);
            const P2<T> Xc = ld2(tile + q0), Xm = ld2(tile + q0 - G.W), Xp = ld2(tile + q0 + G.W);
            const T xl = tile[q0 - 1], xr = tile[q0 + 2];
            const P2<T> q = pair_A(AW, awr, AS0, AS1, AC, Xm, Xc, Xp, xl, xr);
            if (L.act) st2(mb.q + p, q);
            double2 v = make_double2((double)Xc.x * (double)q.x, (double)Xc.y * (double)q.y);
            if (!L.act) v = make_double2(0.0, 0.0);
            put_row2<MODE>(0, j, L, v, mb.stage, sNx, sNy, 0);
            AS0 = AS1;
        }
    }
    __syncthreads();
    if (wave == 0) {
        const double v = tile_of<MODE>(0, lane, mb.stage, sNx, sNy, 0);
        if (lane == 0) *mb.sig = v;
    }
    __syncthreads();
}

template <typename T, int MODE>
__device__ __forceinline__ void lds_C(const SweepGeo& g, const T* __restrict__ pW, const T* __restrict__ pS,
                                      const T* __restrict__ pC, const MemberC<T>& mc) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const int sNx = g.sNx, sNy = g.sNy, OL = g.OL, nx = g.nx, tbase = g.tbase;
    T* tile = reinterpret_cast<T*>(mj_lds);
    const LdsGeo G{sNx + 4};
    MJ_STRIP
    const T* __restrict__ sv = mc.sv;
    const T* __restrict__ qv = mc.q;
    const T* __restrict__ rin = mc.rin;
    T* __restrict__ rout = mc.rout;
    T* __restrict__ xs = mc.xs;
    const T alpha = mc.alpha;
    // x += alpha s_n, r_n = r_{n-1} - alpha q on the strip's interior rows, two rows' loads in flight
    for (int j = j0; j < j1; j += 2) {
        const bool two = j + 1 < j1;
        const int p = MJ_ROW(j) + L.c, p2 = two ? p + nx : p;
        const P2<T> S0 = ld2(sv + p), Q0 = ld2(qv + p), R0 = ld2(rin + p), X0 = ld2(xs + p);
        const P2<T> S1 = ld2(sv + p2), Q1 = ld2(qv + p2), R1 = ld2(rin + p2), X1 = ld2(xs + p2);
        P2<T> x0, r0, x1, r1;
        x0.x = X0.x + alpha * S0.x;
        x0.y = X0.y + alpha * S0.y;
        r0.x = R0.x - alpha * Q0.x;
        r0.y = R0.y - alpha * Q0.y;
        x1.x = X1.x + alpha * S1.x;
        x1.y = X1.y + alpha * S1.y;
        r1.x = R1.x - alpha * Q1.x;
        r1.y = R1.y - alpha * Q1.y;
        if (L.act) {
            st2(xs + p, x0);
            st2(rout + p, r0);
            st2(tile + G.at(j, L.c), r0);
            if (two) {
                st2(xs + p2, x1);
                st2(rout + p2, r1);
                st2(tile + G.at(j + 1, L.c), r1);
            }
        }
    }
    for (int kk = g.ring_off[g.t] + tid; kk < g.ring_off[g.t + 1]; kk += blockDim.x) {
        const int s = g.ring_src[kk];
        T v = T(0);
        if (s >= 0) v = rin[s] - alpha * qv[s];
        tile[lds_ring(G, g.ring_dst[kk] - tbase, nx, OL)] = g.ring_sign[kk] < 0 ? -v : v;
    }
    __syncthreads();
    // z = M r_n from the LDS tile; partials err = (r_n, r_n), eta = (z, r_n)
    if (j0 < j1) {
        const int ce = east_col(L, sNx);
        int pr = MJ_ROW(j0);
        P2<T> PS0 = ld2(pS + pr + L.c);
        for (int j = j0; j < j1; ++j, pr += nx) {
            const int p = pr + L.c;
            const P2<T> PS1 = ld2(pS + p + nx), PC = ld2(pC + p), PW = ld2(pW + pr + ce);
            T pwr = __shfl_down(PW.x, 1, 64);
            if (L.last == 63 && lane == 63) pwr = pW[pr + sNx];
            const int q0 = G.at(j, L.c);
            const P2<T> Rc = ld2(tile + q0), Rm = ld2(tile + q0 - G.W), Rp = ld2(tile + q0 + G.W);
            const T rl = tile[q0 - 1], rr = tile[q0 + 2];
            const P2<T> zz = pair_M(PC, PW, pwr, PS0, PS1, Rm, Rc, Rp, rl, rr);
            if (L.act) st2(mc.z + p, zz);
            double2 e = make_double2((double)Rc.x * (double)Rc.x, (double)Rc.y * (double)Rc.y);
            double2 h = make_double2((double)zz.x * (double)Rc.x, (double)zz.y * (double)Rc.y);
            if (!L.act) {
                e = make_double2(0.0, 0.0);
                h = e;
            }
            put_row2<MODE>(0, j, L, e, mc.stage, sNx, sNy, 0);
            put_row2<MODE>(1, j, L, h, mc.stage, sNx, sNy, 0);
            PS0 = PS1;
        }
    }
    __syncthreads();
    for (int r = wave; r < 2; r += nwaves) {
        const double v = tile_of<MODE>(r, lane, mc.stage, sNx, sNy, 0);
        if (lane == 0) *(r == 0 ? mc.err : mc.eta) = v;
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------------------------ groups
// Member groups: members of one operator slot, in member order, cut into groups of up to kGroup; groups are
// numbered by their first member. grp[g * kGroup + k] = member or -1; *ngroups = number of groups. One block.
// slot(m) = 0 for a shared operator, op_index[m] with slots, m without (every member its own group).
constexpr int kGroupBuildMax = 4096;  // larger batches get one group per member
static __global__ void __launch_bounds__(1024) build_groups_kernel(const int32_t* op_index, int64_t oms, int B,
                                                            int32_t* grp, int* ngroups) {
    __shared__ int s_slot[kGroupBuildMax];
    __shared__ short s_rank[kGroupBuildMax];
    __shared__ int s_n;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int q = tid; q < B * kGroup; q += nt) grp[q] = -1;
    if (B > kGroupBuildMax) {
        __syncthreads();
        for (int m = tid; m < B; m += nt) grp[m * kGroup] = m;
        if (tid == 0) *ngroups = B;
        return;
    }
    for (int m = tid; m < B; m += nt) s_slot[m] = oms == 0 ? 0 : (op_index ? op_index[m] : m);
    if (tid == 0) s_n = 0;
    __syncthreads();
    for (int m = tid; m < B; m += nt) {
        int rank = 0;
        for (int q = 0; q < m; ++q) rank += s_slot[q] == s_slot[m];
        s_rank[m] = (short)rank;
    }
    __syncthreads();
    for (int m = tid; m < B; m += nt) {
        const int want = s_rank[m] - s_rank[m] % kGroup;  // rank of the group's first member
        int lead = m;
        for (int q = m; q >= 0; --q)
            if (s_slot[q] == s_slot[m] && s_rank[q] == want) {
                lead = q;
                break;
            }
        int gid = 0;  // groups whose first member precedes lead
        for (int q = 0; q < lead; ++q) gid += s_rank[q] % kGroup == 0;
        grp[gid * kGroup + s_rank[m] % kGroup] = m;
        if (s_rank[m] % kGroup == 0) atomicAdd(&s_n, 1);
    }
    __syncthreads();
    if (tid == 0) *ngroups = s_n;
}

// Unit (tile t, group gi) of a grouped phase launch of gridDim.x >= ngroups * T blocks (a multiple of kXcds):
// the ngroups * T units, tile-major, are cut into kXcds contiguous ranges, one per XCD (block b runs on XCD
// b % kXcds). Returns false for the blocks beyond.
__device__ __forceinline__ bool group_unit(int ngroups, int T, int& gi, int& t) {
    const int n = ngroups * T;
    const int chunk = (n + kXcds - 1) / kXcds;
    const int pos = blockIdx.x / kXcds;
    const int u = (blockIdx.x % kXcds) * chunk + pos;
    if (pos >= chunk || u >= n) return false;
    t = u / ngroups;
    gi = u - t * ngroups;
    return true;
}

// Unit of a phase launch: gridDim.x is a multiple of kXcds; block b runs on XCD b % kXcds and takes unit
// (b % kXcds) * (gridDim.x / kXcds) + b / kXcds, so each XCD works through one contiguous, tile-major range of
// units and the coefficients of its tiles stay in its L2. Returns false for the padding blocks.
__device__ __forceinline__ bool phase_unit(int B, int T, int& m, int& t) {
    const int chunk = gridDim.x / kXcds;
    const int u = (blockIdx.x % kXcds) * chunk + blockIdx.x / kXcds;
    if (u >= B * T) return false;
    t = u / B;
    m = u - t * B;
    return true;
}

}  // namespace mj
