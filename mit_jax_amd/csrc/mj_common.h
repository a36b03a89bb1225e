// Shared helpers of the MI355X (gfx950) CG2D / DST3 / exch2 halo library.
//
// Everything here is plain FP64 arithmetic in literal order: the library is built with -ffp-contract=off so that
// no multiply-add is fused and every result is the IEEE result of the written operation sequence. The CPU oracle
// is built the same way; that is what makes GPU == oracle bitwise.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/mjamd.h"

namespace mj {

// ---------------------------------------------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int fail_arg(const char* fmt, ...);

#define MJ_HIP(call)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) {                                                                          \
            ::mj::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_));           \
            return MJ_ERR_HIP;                                                                           \
        }                                                                                                \
    } while (0)

#define MJ_REQUIRE(cond, ...)                                                                            \
    do {                                                                                                 \
        if (!(cond)) return ::mj::fail_arg(__VA_ARGS__);                                                 \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------------------------------------------- DST3
// DST3 interface value (gad_dst3_adv_x.F:81-117 via gad.py:420-428): t_m2, t_m1 upstream side for positive
// transport, t_0, t_p1 downstream; m_p1, m_0, m_m1 the face masks of Rjp, Rj, Rjm.
// The same from the masked differences Rjm, Rj, Rjp (the two faces of a point share two of them).
__device__ __forceinline__ double dst3_flux_r(double tr_, double vel_cfl_abs, double oneSixth, double t_m1, double t_0,
                                              double Rjm, double Rj, double Rjp) {
    const double cfl = vel_cfl_abs;
    const double d0 = (2. - cfl) * (1. - cfl) * oneSixth;
    const double d1 = (1. - cfl * cfl) * oneSixth;
    return 0.5 * (tr_ + fabs(tr_)) * (t_m1 + (d0 * Rj + d1 * Rjm)) + 0.5 * (tr_ - fabs(tr_)) * (t_0 - (d0 * Rj + d1 * Rjp));
}
__device__ __forceinline__ double dst3_flux(double tr_, double vel_cfl_abs, double oneSixth, double t_m2, double t_m1,
                                            double t_0, double t_p1, double m_m1, double m_0, double m_p1) {
    const double Rjp = (t_p1 - t_0) * m_p1;
    const double Rj = (t_0 - t_m1) * m_0;
    const double Rjm = (t_m1 - t_m2) * m_m1;
    return dst3_flux_r(tr_, vel_cfl_abs, oneSixth, t_m1, t_0, Rjm, Rj, Rjp);
}

// ---------------------------------------------------------------------------------------------------- reductions
// The fixed reduction tree of the "tile" sum order (documented in mjamd.h): n values (64 <= n <= 128) are folded
// onto the 64 lanes of a wavefront, lane l holding v[l] + v[l+64] (v[l] alone when l+64 >= n), then a 6-step
// butterfly w[l] = w[l] + w[l ^ m], m = 32, 16, ..., 1; lane 0's value is the result. A row of 90 points and a
// tile of 90 row sums are both reduced this way.
//
// Lane 0 of the butterfly only ever adds its partner l + m, so the lower half of each step suffices: lane l < m
// takes the value of lane l + m with the cross-lane moves of the VALU (gfx950 permlane swaps for m = 32, 16, DPP
// row shifts for m = 8 .. 1) instead of LDS permutes. The other lanes end with partial values.
__device__ __forceinline__ double lane_up32(double v) {  // lanes 0..31: the value of lane l + 32
    const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return __hiloint2double((int)b[1], (int)a[1]);
}
__device__ __forceinline__ double lane_up16(double v) {  // lanes 0..15 (and 32..47): the value of lane l + 16
    const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return __hiloint2double((int)b[1], (int)a[1]);
}
template <int K>  // lanes with l % 16 < 16 - K: the value of lane l + K (DPP row_shl:K)
__device__ __forceinline__ double lane_up_row(double v) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int l2 = __builtin_amdgcn_update_dpp(lo, lo, 0x100 + K, 0xf, 0xf, false);
    const int h2 = __builtin_amdgcn_update_dpp(hi, hi, 0x100 + K, 0xf, 0xf, false);
    return __hiloint2double(h2, l2);
}
// Steps m = 16 .. 1 of the butterfly (lane 0's chain).
__device__ __forceinline__ double tree_from16(double w) {
    w = w + lane_up16(w);
    w = w + lane_up_row<8>(w);
    w = w + lane_up_row<4>(w);
    w = w + lane_up_row<2>(w);
    return w + lane_up_row<1>(w);
}
__device__ __forceinline__ double wave_tree(double lo, double hi, bool has_hi) {
    double w = has_hi ? lo + hi : lo;
    w = w + lane_up32(w);
    return tree_from16(w);
}
// The same tree for n values (64 <= n <= 128, n even) held as column pairs: lane L holds v[2L], v[2L+1] (zeros
// beyond the row). In pair terms the fold is v[2L] + v[2L+64], the steps m = 32 .. 2 pair lane L with L + m/2 and
// the last step adds the two halves of lane 0's pair.
__device__ __forceinline__ double pair_tree(double2 v, int n, int lane) {
    const double ux = lane_up32(v.x), uy = lane_up32(v.y);
    const bool hi = 2 * lane + 64 < n;
    double x = hi ? v.x + ux : v.x, y = hi ? v.y + uy : v.y;
    x = x + lane_up16(x);
    y = y + lane_up16(y);
    x = x + lane_up_row<8>(x);
    y = y + lane_up_row<8>(y);
    x = x + lane_up_row<4>(x);
    y = y + lane_up_row<4>(y);
    x = x + lane_up_row<2>(x);
    y = y + lane_up_row<2>(y);
    x = x + lane_up_row<1>(x);
    y = y + lane_up_row<1>(y);
    return x + y;
}

// Sum of the per-tile partials in tile order starting from 0 (GLOBAL_SUM_TILE_RL, reference
// mitgcm_jax/parallel/global_sum.py:20-26).
__device__ __forceinline__ double tile_order_sum(const double* p, int ntiles) {
    double s = 0.0;
    for (int t = 0; t < ntiles; ++t) s = s + p[t];
    return s;
}
// N tile-order sums computed by one whole wave: the lanes load the partials of all N sums at once (one memory
// latency instead of N * ntiles dependent loads), then every lane adds them in tile order from lane reads. Same
// order and bits as tile_order_sum; all 64 lanes of the wave must call it with the same arguments.
template <int N>
__device__ __forceinline__ void wave_tile_sums(const double* const (&p)[N], int ntiles, double (&out)[N]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < N; ++k) out[k] = 0.0;
    for (int t0 = 0; t0 < ntiles; t0 += 64) {
        const int t = t0 + lane, n = ntiles - t0 < 64 ? ntiles - t0 : 64;
        int lo[N], hi[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const double v = t < ntiles ? p[k][t] : 0.0;
            lo[k] = __double2loint(v);
            hi[k] = __double2hiint(v);
        }
        for (int i = 0; i < n; ++i) {
#pragma unroll
            for (int k = 0; k < N; ++k)
                out[k] = out[k] + __hiloint2double(__builtin_amdgcn_readlane(hi[k], i), __builtin_amdgcn_readlane(lo[k], i));
        }
    }
}

// ---------------------------------------------------------------------------------------------------- grid barrier
// A monotonic arrival counter: barrier number k (1, 2, ...) completes when the counter reaches k * gridDim.x.
// Only valid for a cooperative launch (all blocks co-resident). The wait is bounded by wall-clock time: a block
// that waits longer than `limit_ticks` raises the shared error flag and every block leaves its loop at the next
// check, so no wave can spin forever.
struct GridSync {
    unsigned int* counter;  // zeroed before the launch
    int* error;             // zeroed before the launch; 1 = a barrier timed out
    unsigned int nblocks;
    unsigned int target;
    unsigned long long limit_ticks;
};

__device__ __forceinline__ bool grid_sync(GridSync& g) {
    __shared__ int s_ok;
    __syncthreads();
    g.target += g.nblocks;
    if (threadIdx.x == 0) {
        int ok = 1;
        __threadfence();
        atomicAdd(g.counter, 1u);
        const unsigned long long t0 = wall_clock64();
        while (__hip_atomic_load(g.counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < g.target) {
            __builtin_amdgcn_s_sleep(2);
            if (__hip_atomic_load(g.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                ok = 0;
                break;
            }
            if (wall_clock64() - t0 > g.limit_ticks) {
                atomicExch(g.error, 1);
                ok = 0;
                break;
            }
        }
        __threadfence();
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

}  // namespace mj
