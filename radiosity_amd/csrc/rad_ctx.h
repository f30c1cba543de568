// rad_ctx.h - the context behind the opaque rad_ctx handle of the C ABI
// (device buffers of one scene, the pass state, the stream), shared by the
// translation units that implement entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "rad_internal.h"

using rad::BinBuffers;
using rad::Geometry;
using rad::PackBuffers;

struct rad_ctx {
    rad_config cfg{};
    Geometry g{};
    uint32_t K = 1;
    int device = 0;
    hipStream_t st = nullptr;
    // second stream of a split pass (its halves render side by side), with
    // the events that fork it from st and join it back
    hipStream_t st2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::string err;

    // scene
    uint32_t ntri = 0, npatch = 0;
    float *tri = nullptr;          // 9 planes of ntri
    uint32_t *tri_patch = nullptr; // ntri
    float4 *tsph = nullptr;        // ntri bounding spheres (centre, radius)
    float *frames = nullptr;       // npatch*9
    float *refl = nullptr;         // npatch*3
    float *area = nullptr;         // npatch
    float *emis = nullptr;         // npatch*3
    float *unshot = nullptr;       // npatch*3
    float *shot = nullptr;         // npatch*3
    float *ff = nullptr;           // K*npatch
    // one flag per (64 F entries of a row, slot), [row / 64][slot]: set by
    // the tile kernel when it adds to that block of the slot's row, cleared
    // by k_shoot when it has read it; a block whose flag is clear is zero
    uint32_t *fblk = nullptr;      // ceil(rowlen / 64) * K
    uint32_t *hits = nullptr;      // K*npatch
    // F rows [0, ff_dirty) may hold nonzero entries at this point of the
    // stream, rows from ff_dirty on are zero: a clear or a shoot (the shoot
    // kernels zero every entry they read) leaves 0, a kernel that accumulates
    // into k rows raises it to at least k.  Lets a pass skip clearing F.
    uint32_t ff_dirty = 0;

    // per pass
    float *dff = nullptr;          // RES
    uint32_t *item_ids = nullptr;  // K*RES: the item buffer of the last render (patch + 1)
    float *item_depth = nullptr;   // K*RES: its depths (allocated when first asked for)
    unsigned long long *ffq = nullptr; // K*rowlen fixed-point F rows (deterministic mode)
    bool det = false;              // deterministic F rows (rad_set_deterministic)
    // run-form ProcessHemicube (rad_process_runs): per-work-item slot counts
    // and offsets, and the device triplet stream
    uint32_t *runs_slots = nullptr, *runs_offs = nullptr;
    uint32_t runs_items = 0;
    uint32_t *trip_h = nullptr, *trip_i = nullptr;
    float *trip_e = nullptr;
    uint32_t trip_cap = 0;
    float *mvp = nullptr;          // K*5*16
    float *snap = nullptr;         // K*3
    uint32_t *shooters = nullptr;  // K
    uint32_t *nshoot = nullptr;    // 1
    uint32_t *counters = nullptr;  // kCntBlocks blocks of kCntN
    uint32_t *ctl = nullptr;       // kCtlN: stop test, pass and shooter counts
    uint64_t *sel = nullptr;       // 1 (+ pad)
    double *resid = nullptr;       // [0,1] selection residual, [4,5] rad_residual_fp64
    double *parts = nullptr;       // per-block partials of the selection (3 words per block)
    BinBuffers bb{};               // sort-middle rasterizer bins
    uint64_t *cand = nullptr;      // top-K candidates (two halves)
    rad::SelPool *pool = nullptr;  // candidate pool of chained selections
    uint32_t pool_cap = rad::kPoolCap;
    uint32_t small_px = 16;        // in-lane rasterization limit (pixels of a tile)
    int sel_blocks = 1024;
    bool rendered = false;        // item_ids hold the item buffer of the last render
    uint32_t rendered_k = 0;

    // F row length: npatch, or the largest scene of a pack
    uint32_t rowlen = 0;
    uint64_t geom_hash = 0;        // FNV-1a of the uploaded scene arrays (checkpoints)
    bool hash_valid = false;       // geom_hash computed (ingested scenes hash lazily)

    // quad topology (rad_scene_ingest or rad_set_quads; nquad 0 when absent)
    uint32_t nquad = 0;
    uint32_t *patch_quad = nullptr; // npatch
    uint32_t *quad_pbeg = nullptr;  // nquad + 1: first patch of each quad
    uint32_t *quad_nbr = nullptr;   // nquad * 8
    float *qval = nullptr;          // nquad * 3 (vertex colours)
    float *viewfr = nullptr;        // 9: frame of rad_render_view
    uint32_t *view_sh = nullptr;    // 1: shooter slot 0 -> frame 0 of viewfr
    PackBuffers pk{};

    // captured passes (rad_config.graph_passes of them) replayed by rad_step
    hipGraphExec_t gexec = nullptr;
    uint32_t gpasses = 0;

    // dense matrix of the gather: npad x npad bf16 halves (hi, lo)
    uint16_t *fhi = nullptr, *flo = nullptr;
    uint32_t npad = 0;
    bool dense = false;

    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    double t_acc[6] = {0, 0, 0, 0, 0, 0};
    double t_reduce = 0.0;          // device ms of the last standalone ProcessHemicube
    std::string pending_error;
    uint64_t t_passes = 0;
};
