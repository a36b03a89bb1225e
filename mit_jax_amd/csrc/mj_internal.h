// Internal plan structures shared by the library's translation units.
#pragma once

#include <stdint.h>

#include <mutex>
#include <vector>

struct mj_halo_plan {
    int nmaps = 0, ntiles = 0, ny = 0, nx = 0;
    int npts[2] = {0, 0};
    int inplace_ok = 0;
    int32_t* d_dst[2] = {nullptr, nullptr};
    int32_t* d_src[2] = {nullptr, nullptr};
    int8_t* d_comp[2] = {nullptr, nullptr};
    int8_t* d_sign[2] = {nullptr, nullptr};
    std::vector<int32_t> h_dst[2], h_src[2];
    std::vector<int8_t> h_comp[2], h_sign[2];

    // Tables the fused CG2D kernel derives from a scalar plan (built once, on first use, for one (sNx, OL)):
    // per tile t, the written points of tile t ("full", every halo point the map writes) and the subset on the
    // width-1 ring the 5-point stencil reads ("ring"), as offsets [t] .. [t+1] into dst/src/sign lists.
    std::mutex cg_mu;
    int cg_sNx = 0, cg_sNy = 0, cg_OL = 0;
    int32_t *cg_full_off = nullptr, *cg_full_dst = nullptr, *cg_full_src = nullptr;
    int8_t* cg_full_sign = nullptr;
    int32_t *cg_ring_off = nullptr, *cg_ring_dst = nullptr, *cg_ring_src = nullptr;
    int8_t* cg_ring_sign = nullptr;
    int cg_ring_max = 0;  // largest per-tile ring count
    // the same ring entries for the compact work vectors of the mixed solve: the LDS offset of the ring point in
    // the stencil tile and the source's compact index within the member (-1: no source)
    int4* cg_cring = nullptr;  // ring entries of the mixed solve: {compact source, padded source, LDS offset | neg << 30, 0}

    // Transpose tables (built once, on first use): per input component c (0 = u-like, 1 = v-like) and padded
    // point q, the written points that read q (tr_off[c][q] .. tr_off[c][q+1] into tr_dst/tr_map/tr_sign, sorted by
    // map then destination), and per map whether q is written (tr_written[m][q]).
    std::mutex tr_mu;
    int32_t *tr_off[2] = {nullptr, nullptr}, *tr_dst[2] = {nullptr, nullptr};
    int8_t *tr_map[2] = {nullptr, nullptr}, *tr_sign[2] = {nullptr, nullptr};
    int8_t* tr_written[2] = {nullptr, nullptr};

    ~mj_halo_plan();
};

// One update box of a DST3 sweep: array-index ranges [i0, i1) x [j0, j1).
struct MjBox {
    int16_t i0, i1, j0, j1;
};

// What tile t does in one direction of one pass (reference gad.py:339-375 / gad_advection.F:333-774).
struct MjSweep {
    int8_t run;          // block runs (flux needed or a fill happens)
    int8_t fill_before;  // fill4dir of FILL_CS_CORNER_TR_RL before the flux (0 = none)
    int8_t fill_after;   // fill4dir after the flux, before the update (0 = none)
    int8_t nbox;         // update boxes (0..2)
    MjBox box[2];
};

struct MjTilePlan {
    MjSweep sweep[3][2];  // [pass][dir: 0 = X, 1 = Y]
    int8_t sw, se, nw, ne;
};

namespace mj {
// Builds (once per plan, for one (sNx, sNy, OL)) the per-tile full and ring tables of a scalar T plan that the CG2D
// solves use; the ring points must all have interior sources.
int ensure_cg_tables(mj_halo_plan* P, int sNx, int sNy, int OL);
}  // namespace mj

struct mj_gad_plan {
    int ntiles = 0, sNx = 0, sNy = 0, OLx = 0, OLy = 0;
    std::vector<MjTilePlan> tiles;
    MjTilePlan* d_tiles = nullptr;
    ~mj_gad_plan();
};
