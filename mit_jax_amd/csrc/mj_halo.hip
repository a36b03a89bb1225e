// exch2 halo exchange on gfx950: one thread per (written point, level, member), gather with sign.
//
// The tables are the written points of the probed exch2 maps (sorted by destination), so a call moves exactly the
// halo values the Fortran exchange moves (18,112 of the 124,852 padded points per level for the T map) instead of
// the reference's full-array gather (parallel/exchange.py:136-142). Bitwise equal to that gather: each written
// value is sign * source (sign = +-1 exactly), each other point keeps its value.
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "mj_common.h"
#include "mj_internal.h"

namespace mj {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int fail_arg(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return MJ_ERR_ARG;
}

// ------------------------------------------------------------------------------------------------ kernels
__global__ void __launch_bounds__(256) halo_gather_kernel(const int32_t* __restrict__ dst, const int32_t* __restrict__ src,
                                                          const int8_t* __restrict__ comp, const int8_t* __restrict__ sign,
                                                          int npts, const double* in_a, const double* in_b, double* out,
                                                          int ntiles, int plane, int64_t tile_stride, int64_t ninner,
                                                          int64_t inner_stride, int64_t outer_stride) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const int64_t lead = blockIdx.y;  // inner + ninner * outer
    const int64_t n = lead % ninner;
    const int64_t o = lead / ninner;
    const int64_t base = o * outer_stride + n * inner_stride;
    const int d = dst[p];
    const int s = src[p];
    const int dt = d / plane, dr = d - dt * plane;
    const int st = s / plane, sr = s - st * plane;
    const double* from = comp[p] == 2 ? in_b : in_a;
    const double v = from[base + st * tile_stride + sr];
    out[base + dt * tile_stride + dr] = sign[p] < 0 ? -v : v;
}

// Transpose of the exchange (reverse mode): the cotangent of input component c at point q is its own cotangent
// where no map writes q, plus sign * (cotangent of each written point that reads q), summed in table order.
__global__ void __launch_bounds__(256) halo_transpose_kernel(const int32_t* __restrict__ off,
                                                             const int32_t* __restrict__ tdst,
                                                             const int8_t* __restrict__ tmap,
                                                             const int8_t* __restrict__ tsign,
                                                             const int8_t* __restrict__ written, const double* own,
                                                             const double* bar0, const double* bar1, double* out,
                                                             int total, int plane,
                                                             int64_t tile_stride, int64_t ninner, int64_t inner_stride,
                                                             int64_t outer_stride) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= total) return;
    const int64_t lead = blockIdx.y;
    const int64_t n = lead % ninner;
    const int64_t o = lead / ninner;
    const int64_t base = o * outer_stride + n * inner_stride;
    auto at = [&](int p) { return base + (int64_t)(p / plane) * tile_stride + p % plane; };
    double acc = written[q] ? 0.0 : own[at(q)];
    for (int k = off[q]; k < off[q + 1]; ++k) {
        const double* from = tmap[k] == 1 ? bar1 : bar0;  // cotangent of the output of the map the entry is in
        const double v = from[at(tdst[k])];
        acc = acc + (tsign[k] < 0 ? -v : v);
    }
    out[at(q)] = acc;
}

}  // namespace mj

using namespace mj;

// ------------------------------------------------------------------------------------------------ plan
static int upload(const void* h, size_t bytes, void** d) {
    MJ_HIP(hipMalloc(d, bytes > 0 ? bytes : 4));
    if (bytes) MJ_HIP(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
    return MJ_OK;
}

extern "C" int mj_version(void) { return 100; }

extern "C" int mj_last_error(char* buf, size_t n) {
    const std::string& e = g_last_error;
    if (buf && n) {
        size_t k = e.size() < n - 1 ? e.size() : n - 1;
        memcpy(buf, e.data(), k);
        buf[k] = 0;
    }
    return (int)e.size();
}

extern "C" int mj_halo_plan_create(int nmaps, const int32_t* npts, const int32_t* const* dst,
                                   const int32_t* const* src, const int8_t* const* comp,
                                   const int8_t* const* sign, int ntiles, int ny, int nx, mj_halo_plan** out) {
    MJ_REQUIRE(out && npts && dst && src && comp && sign, "mj_halo_plan_create: null argument");
    MJ_REQUIRE(nmaps == 1 || nmaps == 2, "mj_halo_plan_create: nmaps must be 1 or 2 (got %d)", nmaps);
    MJ_REQUIRE(ntiles > 0 && ny > 0 && nx > 0, "mj_halo_plan_create: bad layout");
    const int64_t total = (int64_t)ntiles * ny * nx;
    mj_halo_plan* p = new mj_halo_plan();
    p->nmaps = nmaps;
    p->ntiles = ntiles;
    p->ny = ny;
    p->nx = nx;
    std::vector<char> written[2];
    for (int m = 0; m < nmaps; ++m) {
        const int n = npts[m];
        if (n < 0 || (n > 0 && (!dst[m] || !src[m] || !comp[m] || !sign[m]))) {
            delete p;
            return fail_arg("mj_halo_plan_create: map %d: bad point list", m);
        }
        written[m].assign(total, 0);
        for (int k = 0; k < n; ++k) {
            const int32_t d = dst[m][k], s = src[m][k];
            const int c = comp[m][k], g = sign[m][k];
            if (d < 0 || d >= total || s < 0 || s >= total || (c != 1 && c != 2) || (g != 1 && g != -1) ||
                (nmaps == 1 && c != 1) || (k > 0 && d <= dst[m][k - 1])) {
                delete p;
                return fail_arg("mj_halo_plan_create: map %d point %d invalid (dst %d src %d comp %d sign %d)", m,
                                k, d, s, c, g);
            }
            written[m][d] = 1;
        }
        p->h_dst[m].assign(dst[m], dst[m] + n);
        p->h_src[m].assign(src[m], src[m] + n);
        p->h_comp[m].assign(comp[m], comp[m] + n);
        p->h_sign[m].assign(sign[m], sign[m] + n);
        p->npts[m] = n;
    }
    // in place is exact when no source is a written point of the component it is read from
    int ok = 1;
    for (int m = 0; m < nmaps && ok; ++m)
        for (int k = 0; k < npts[m]; ++k) {
            const int from = comp[m][k] - 1;  // 0 = u-like input, 1 = v-like input
            if (written[from][src[m][k]]) {
                ok = 0;
                break;
            }
        }
    p->inplace_ok = ok;
    for (int m = 0; m < nmaps; ++m) {
        int rc;
        if ((rc = upload(p->h_dst[m].data(), npts[m] * 4, (void**)&p->d_dst[m])) ||
            (rc = upload(p->h_src[m].data(), npts[m] * 4, (void**)&p->d_src[m])) ||
            (rc = upload(p->h_comp[m].data(), npts[m], (void**)&p->d_comp[m])) ||
            (rc = upload(p->h_sign[m].data(), npts[m], (void**)&p->d_sign[m]))) {
            mj_halo_plan_destroy(p);
            return rc;
        }
    }
    *out = p;
    return MJ_OK;
}

mj_halo_plan::~mj_halo_plan() {
    for (int m = 0; m < 2; ++m) {
        if (d_dst[m]) (void)hipFree(d_dst[m]);
        if (d_src[m]) (void)hipFree(d_src[m]);
        if (d_comp[m]) (void)hipFree(d_comp[m]);
        if (d_sign[m]) (void)hipFree(d_sign[m]);
    }
    for (void* q : {(void*)cg_full_off, (void*)cg_full_dst, (void*)cg_full_src, (void*)cg_full_sign,
                    (void*)cg_ring_off, (void*)cg_ring_dst, (void*)cg_ring_src, (void*)cg_ring_sign,
                    (void*)cg_cring})
        if (q) (void)hipFree(q);
    for (int c = 0; c < 2; ++c)
        for (void* q : {(void*)tr_off[c], (void*)tr_dst[c], (void*)tr_map[c], (void*)tr_sign[c], (void*)tr_written[c]})
            if (q) (void)hipFree(q);
}

static int ensure_transpose(mj_halo_plan* P) {
    std::lock_guard<std::mutex> lk(P->tr_mu);
    if (P->tr_off[0]) return MJ_OK;
    const int total = P->ntiles * P->ny * P->nx;
    for (int c = 0; c < P->nmaps; ++c) {
        std::vector<int32_t> cnt(total + 1, 0), dst;
        std::vector<int8_t> map, sg;
        for (int m = 0; m < P->nmaps; ++m)
            for (int k = 0; k < P->npts[m]; ++k)
                if (P->h_comp[m][k] - 1 == c) ++cnt[P->h_src[m][k] + 1];
        for (int q = 0; q < total; ++q) cnt[q + 1] += cnt[q];
        std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
        dst.assign(cnt[total], 0);
        map.assign(cnt[total], 0);
        sg.assign(cnt[total], 0);
        for (int m = 0; m < P->nmaps; ++m)
            for (int k = 0; k < P->npts[m]; ++k)
                if (P->h_comp[m][k] - 1 == c) {
                    const int at = fill[P->h_src[m][k]]++;
                    dst[at] = P->h_dst[m][k];
                    map[at] = (int8_t)m;
                    sg[at] = P->h_sign[m][k];
                }
        std::vector<int8_t> wr(total, 0);
        for (int k = 0; k < P->npts[c]; ++k) wr[P->h_dst[c][k]] = 1;
        int rc;
        if ((rc = upload(cnt.data(), cnt.size() * 4, (void**)&P->tr_off[c])) ||
            (rc = upload(dst.data(), dst.size() * 4, (void**)&P->tr_dst[c])) ||
            (rc = upload(map.data(), map.size(), (void**)&P->tr_map[c])) ||
            (rc = upload(sg.data(), sg.size(), (void**)&P->tr_sign[c])) ||
            (rc = upload(wr.data(), wr.size(), (void**)&P->tr_written[c])))
            return rc;
    }
    return MJ_OK;
}

extern "C" int mj_halo_plan_destroy(mj_halo_plan* plan) {
    delete plan;
    return MJ_OK;
}

extern "C" int mj_halo_plan_inplace_ok(const mj_halo_plan* plan) { return plan ? plan->inplace_ok : 0; }

static int check_geom(const mj_halo_plan* plan, const mj_halo_geom* g) {
    MJ_REQUIRE(plan && g, "halo: null plan or geometry");
    MJ_REQUIRE(g->nouter >= 1 && g->ninner >= 1, "halo: nouter, ninner must be >= 1");
    MJ_REQUIRE(g->nouter * g->ninner <= 65535, "halo: nouter*ninner = %lld exceeds 65535",
               (long long)(g->nouter * g->ninner));
    MJ_REQUIRE(g->tile_stride >= (int64_t)plan->ny * plan->nx, "halo: tile_stride too small");
    return MJ_OK;
}

static int launch_map(const mj_halo_plan* plan, int m, const double* in_a, const double* in_b, double* out,
                      const mj_halo_geom* g, hipStream_t st) {
    const int n = plan->npts[m];
    if (n == 0) return MJ_OK;
    dim3 grid((n + 255) / 256, (unsigned)(g->nouter * g->ninner));
    hipLaunchKernelGGL(halo_gather_kernel, grid, dim3(256), 0, st, plan->d_dst[m], plan->d_src[m], plan->d_comp[m],
                       plan->d_sign[m], n, in_a, in_b, out, plan->ntiles, plan->ny * plan->nx, g->tile_stride,
                       g->ninner, g->inner_stride, g->outer_stride);
    MJ_HIP(hipGetLastError());
    return MJ_OK;
}

extern "C" int mj_halo_exchange(const mj_halo_plan* plan, const double* in, double* out, const mj_halo_geom* g,
                                void* stream) {
    int rc = check_geom(plan, g);
    if (rc) return rc;
    MJ_REQUIRE(plan->nmaps == 1, "mj_halo_exchange: plan is a vector kind (use mj_halo_exchange_uv)");
    MJ_REQUIRE(in && out, "mj_halo_exchange: null array");
    MJ_REQUIRE(in != out || plan->inplace_ok, "mj_halo_exchange: this kind cannot be exchanged in place");
    return launch_map(plan, 0, in, in, out, g, as_stream(stream));
}

extern "C" int mj_halo_exchange_uv(const mj_halo_plan* plan, const double* u_in, const double* v_in, double* u_out,
                                   double* v_out, const mj_halo_geom* g, void* stream) {
    int rc = check_geom(plan, g);
    if (rc) return rc;
    MJ_REQUIRE(plan->nmaps == 2, "mj_halo_exchange_uv: plan is a scalar kind");
    MJ_REQUIRE(u_in && v_in && u_out && v_out, "mj_halo_exchange_uv: null array");
    const bool alias = (const double*)u_out == u_in || (const double*)u_out == v_in ||
                       (const double*)v_out == u_in || (const double*)v_out == v_in;
    MJ_REQUIRE(!alias || plan->inplace_ok, "mj_halo_exchange_uv: this kind cannot be exchanged in place");
    hipStream_t st = as_stream(stream);
    if ((rc = launch_map(plan, 0, u_in, v_in, u_out, g, st))) return rc;
    return launch_map(plan, 1, u_in, v_in, v_out, g, st);
}

extern "C" int mj_halo_exchange_adjoint(const mj_halo_plan* plan, const double* ubar_out, const double* vbar_out,
                                        double* ubar_in, double* vbar_in, const mj_halo_geom* g, void* stream) {
    int rc = check_geom(plan, g);
    if (rc) return rc;
    const bool uv = plan->nmaps == 2;
    MJ_REQUIRE(ubar_out && ubar_in && (!uv || (vbar_out && vbar_in)), "mj_halo_exchange_adjoint: null array");
    MJ_REQUIRE(ubar_in != ubar_out && (!uv || (vbar_in != vbar_out && vbar_in != ubar_out && ubar_in != vbar_out)),
               "mj_halo_exchange_adjoint: outputs must not alias inputs");
    mj_halo_plan* P = const_cast<mj_halo_plan*>(plan);
    if ((rc = ensure_transpose(P))) return rc;
    const int total = P->ntiles * P->ny * P->nx;
    hipStream_t st = as_stream(stream);
    dim3 grid((total + 255) / 256, (unsigned)(g->nouter * g->ninner));
    // map m writes output component m (its own cotangent is that output's); entries of map m read its cotangent
    const double* bar1 = uv ? vbar_out : ubar_out;
    for (int c = 0; c < P->nmaps; ++c) {
        hipLaunchKernelGGL(halo_transpose_kernel, grid, dim3(256), 0, st, P->tr_off[c], P->tr_dst[c], P->tr_map[c],
                           P->tr_sign[c], P->tr_written[c], c == 0 ? ubar_out : vbar_out, ubar_out, bar1,
                           c == 0 ? ubar_in : vbar_in, total, P->ny * P->nx, g->tile_stride, g->ninner,
                           g->inner_stride, g->outer_stride);
        MJ_HIP(hipGetLastError());
    }
    return MJ_OK;
}
