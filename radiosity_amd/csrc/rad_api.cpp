// rad_api.cpp - C ABI of libradhip: context, device memory, pass driver.
//
// The pass driver restates the OnIdle shoot loop (Main.cpp:1137-1309) as a
// sequence of device kernels on the context stream with no host
// synchronisation between passes: select -> prepare -> clear -> raster setup
// -> raster tiles -> reduce -> shoot; the raster stages of a pass of 16 or
// more slots run as two halves on the context's two streams.
#include <cstddef>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/radhip_scene.h"
#include "rad_ctx.h"
#include "rad_internal.h"
#include "rad_patch.h"

using namespace rad;

namespace {

int fail(rad_ctx *c, int code, const std::string &msg)
{
    if (c)
        c->err = msg;
    return code;
}

// 64-bit FNV-1a over a byte range (scene identity of a checkpoint).
uint64_t fnv1a(uint64_t h, const void *p, size_t n)
{
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 0x100000001B3ull;
    }
    return h;
}

// Buffers a captured pass refers to change (scene upload, retiling): the
// next rad_step captures again.
void drop_graph(rad_ctx *c)
{
    if (c->gexec) {
        if (c->st)
            (void)hipStreamSynchronize(c->st);
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
    }
    c->gpasses = 0;
}

#define HIPCHK(c, call)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail((c), RAD_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T> int dalloc(rad_ctx *c, T **p, size_t n)
{
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (n == 0)
        n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, RAD_E_NOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) +
                                        " bytes: " + hipGetErrorString(e));
    }
    return RAD_OK;
}

void dfree(void *p)
{
    if (p)
        (void)hipFree(p);
}

int check_launch(rad_ctx *c, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(c, RAD_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return RAD_OK;
}

void record(rad_ctx *c, size_t idx)
{
    if (c->timing && idx < c->ev.size())
        (void)hipEventRecord(c->ev[idx], c->st);
}

// A pass of 16 or more slots renders as two halves side by side, except a
// timed one (stage events bracket whole stages).
static bool split_render(const rad_ctx *c, uint32_t k) { return k >= kSplitMin && !c->timing; }

// First slot of the second half of a split pass.
static uint32_t split_at(uint32_t k) { return k / 2; }

// Renders the hemicubes of the shooters currently in c->shooters/nshoot
// (k slots): ids != null writes the item buffer (and its depths when depth
// != null), otherwise the tiles are reduced into F rows (and hit counts when
// hits != null); in the deterministic mode through the fixed-point rows and
// a fold into F.  The F rows that
// may be dirty are cleared first (all of them, not only the k this pass
// writes: a render of fewer rows must not leave an earlier writer's rows
// behind for a later pass that skips its clear).
int render_pass(rad_ctx *c, uint32_t k, size_t evbase, uint32_t *ids, float *depth,
                uint32_t *hits, const uint32_t *shooters = nullptr, const float *frames = nullptr,
                bool shoot_half = false, bool prepared = false)
{
    uint32_t rows = c->ff_dirty;
    if (hits && rows < k)
        rows = k;
    // prepared: the previous pass's last shoot already prepared this pass
    // and cleared its bins and counters (no F rows or hits to clear)
    if (!prepared)
        launch_prepare_clear(k, shooters ? shooters : c->shooters, c->nshoot,
                             frames ? frames : c->frames, c->unshot, c->snap, c->mvp, c->bb.count,
                             c->bb.ntiles, c->ff_dirty ? c->ff : nullptr, hits, rows, c->rowlen,
                             c->counters, c->bb.ent_cap, c->ctl, c->st);
    c->ff_dirty = 0;
    record(c, evbase + 2);
    const bool fq = !ids && c->det;
    // stage j (0 setup, 1 bins, 2 tiles) of slots [s0, s1) with counter
    // block h on stream st
    auto stage = [&](int j, uint32_t s0, uint32_t s1, uint32_t h, hipStream_t st) {
        if (j == 0)
            launch_raster_setup(c->g, s0, s1, h, c->tri, c->tsph, c->tri_patch, c->ntri, c->mvp,
                                c->nshoot, c->bb, c->counters, c->pk.slot_tbeg,
                                c->pk.slot_tcnt, c->pk.slot_w, st);
        else if (j == 1)
            launch_bin(s0, s1, h, c->nshoot, c->bb, c->counters, st);
        else
            launch_tile(c->g, s0, s1, h, c->nshoot, c->bb, c->small_px, ids, depth,
                        c->pk.slot_w ? c->pk.dffs : c->dff, c->rowlen, c->pk.slot_pbeg,
                        ids || fq ? nullptr : c->ff, fq ? c->ffq : nullptr, hits, c->counters,
                        c->pk.slot_w, c->pk.slot_dffo, c->fblk, c->K, st);
    };
    // A pass of 16 or more slots renders as two halves side by side: the
    // halves share no bins, records, list entries or counters, so the item
    // buffers and F rows are those of one launch over all slots.  Timed
    // passes stay whole so that the stage events bracket every kernel.
    // With shoot_half, the first half's F rows are shot on the context
    // stream as soon as its tiles are done, while the second half still
    // renders; the caller shoots the second half.
    if (split_render(c, k)) {
        HIPCHK(c, hipEventRecord(c->ev_fork, c->st));
        HIPCHK(c, hipStreamWaitEvent(c->st2, c->ev_fork, 0));
        for (int j = 0; j < 3; ++j) {
            stage(j, 0, split_at(k), 0, c->st);
            stage(j, split_at(k), k, 1, c->st2);
        }
        if (shoot_half)
            launch_shoot(c->cfg.mode, c->K, c->shooters, c->nshoot, c->snap, c->ff, c->fblk,
                         c->refl, c->area, c->npatch, c->unshot, c->shot, c->ctl, 0, split_at(k),
                         false, nullptr, nullptr, c->st);
        HIPCHK(c, hipEventRecord(c->ev_join, c->st2));
        HIPCHK(c, hipStreamWaitEvent(c->st, c->ev_join, 0));
    } else {
        for (int j = 0; j < 3; ++j) {
            stage(j, 0, k, 0, c->st);
            if (j < 2)
                record(c, evbase + 3 + j);
        }
    }
    if (fq)
        launch_fold(c->ffq, c->ff, (size_t)k * c->rowlen, c->st);
    if (!ids)
        c->ff_dirty = k;
    record(c, evbase + 5);
    return check_launch(c, "render kernels");
}

// Reads the pass counters after a synchronisation: a record or bin-entry
// overflow in the last pass, or the sticky flag of an earlier one, means a
// pass lost triangles.
int check_overflow(rad_ctx *c)
{
    uint32_t cnt[kCntN * kCntBlocks];
    hipError_t e = hipMemcpy(cnt, c->counters, sizeof(cnt), hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        return fail(c, RAD_E_HIP, std::string("counters: ") + hipGetErrorString(e));
    for (int h = 0; h < kCntBlocks; ++h) {
        const uint32_t *b = cnt + h * kCntN;
        if (b[kCntOverflow] || b[kCntEntries] > c->bb.ent_cap || cnt[kCntSticky]) {
            uint32_t zero = 0;
            (void)hipMemcpy(c->counters + kCntSticky, &zero, 4, hipMemcpyHostToDevice);
            return fail(c, RAD_E_LIMIT, "rasterizer bins overflowed (records " +
                                            std::to_string(b[kCntSpill]) + " spilled of " +
                                            std::to_string(c->bb.spill_cap) + ", entries " +
                                            std::to_string(b[kCntEntries]) + " of " +
                                            std::to_string(c->bb.ent_cap) + ")");
        }
    }
    return RAD_OK;
}

int select_dev(rad_ctx *c, uint32_t k)
{
    if (k == 1) {
        uint32_t rule = c->cfg.select;
        launch_select_top1(rule, c->unshot, c->area, c->emis, c->npatch, c->shooters, c->nshoot,
                           c->parts, (uint64_t *)(c->parts + 2 * c->sel_blocks), nullptr, c->resid,
                           c->sel_blocks, c->st);
        return check_launch(c, "select");
    }
    if (c->cfg.select == RAD_SELECT_REFERENCE)
        launch_select_reference(c->unshot, c->npatch, k, c->shooters, c->nshoot, c->st);
    else
        launch_select_topk(c->cfg.select, c->unshot, c->area, c->emis, c->npatch, k, c->shooters,
                           c->nshoot, c->cand, nullptr, nullptr, c->resid, c->st);
    return check_launch(c, "select");
}

// Whether a pass's last shoot can leave the next POWER selection's
// per-workgroup lists (top-K of 2 .. 64, one merge level for the lists).
static bool fused_select(const rad_ctx *c)
{
    return !c->pk.nscene && c->cfg.select == RAD_SELECT_POWER && c->K > 1 && c->K <= 64 &&
           shoot_lists(c->npatch) <= 256;
}

// One pass: select, render, shoot.  chain_in: the previous pass already
// selected and prepared this one (no selection or prepare launch);
// chain_out: this pass's last shoot leaves per-workgroup lists of the POWER
// keys of the state it stores, or, in a chain (chain_in too) whose pool is
// armed, pools the keys at or above the pool's bound; k_next_pass then
// selects the next pass's shooters and prepares that pass.  target: the
// keys of the state that the pool's bound must leave above it when
// k_next_pass arms it from lists (0: no arming).  relist: this pass's shoot
// leaves lists whatever the pool's state, and the next selection re-arms.
int one_pass(rad_ctx *c, bool do_select, size_t evbase, bool chain_in = false,
             bool chain_out = false, uint32_t target = 0, bool relist = false)
{
    int rc;
    record(c, evbase + 0);
    if (c->pk.nscene) {
        launch_pack_select(c->pk, c->unshot, c->area, c->st);
        if ((rc = check_launch(c, "pack select")))
            return rc;
    } else if (do_select && !chain_in && (rc = select_dev(c, c->K)) != RAD_OK) {
        return rc;
    }
    record(c, evbase + 1);
    // the first half of a split pass is shot before the second finishes
    // rendering (not for packs, nor with the fixed-point rows, whose fold
    // comes after both halves)
    const bool half = !c->pk.nscene && !c->det && split_render(c, c->K);
    if ((rc = render_pass(c, c->K, evbase, nullptr, nullptr, nullptr, nullptr, nullptr, half,
                          chain_in)) != RAD_OK)
        return rc;
    if (c->pk.nscene)
        launch_pack_shoot(c->pk, c->cfg.mode, c->snap, c->ff, c->rowlen, c->refl, c->area,
                          c->npatch, c->unshot, c->shot, c->st);
    else
        launch_shoot(c->cfg.mode, c->K, c->shooters, c->nshoot, c->snap, c->ff,
                     c->det ? nullptr : c->fblk, c->refl, c->area, c->npatch, c->unshot,
                     c->shot, c->ctl, half ? split_at(c->K) : 0u, c->K, true,
                     chain_out ? c->cand : nullptr,
                     chain_out && chain_in && !relist ? c->pool : nullptr, c->st);
    if (chain_out) {
        const NextPass np{c->K,    c->shooters, c->nshoot,   c->frames,
                          c->snap, c->mvp,      c->bb.count, (size_t)c->K * c->bb.ntiles,
                          c->counters, c->bb.ent_cap};
        launch_next_pass(c->cand, shoot_lists(c->npatch, chain_in && !relist), np, c->unshot,
                         c->area, c->npatch,
                         c->ctl, c->pool, !chain_in || relist, target, c->pool_cap, c->st);
    }
    c->ff_dirty = 0; // the shoot kernels zero the entries they read
    record(c, evbase + 6);
    c->rendered = false;
    return check_launch(c, "shoot");
}

// Screen tile side T (8, 16, 32) and the in-lane triangle limit; (re)sizes
// the per-(shooter, tile) bin arrays.
int set_tiles(rad_ctx *c, uint32_t T, uint32_t small_px)
{
    if (T != 8 && T != 16 && T != 32)
        return fail(c, RAD_E_ARG, "tile_px must be 8, 16 or 32");
    if (small_px == 0 || small_px > 4096)
        return fail(c, RAD_E_ARG, "small_tri_px must be 1..4096");
    uint32_t ntx = (c->g.TW + T - 1) / T, nty = (c->g.TH + T - 1) / T;
    if (ntx * nty > kMaxTilesPerSlab)
        return fail(c, RAD_E_ARG, "too many screen tiles per hemicube (use a larger tile_px)");
    if (c->st)
        (void)hipStreamSynchronize(c->st);
    c->bb.tshift = T == 8 ? 3 : (T == 16 ? 4 : 5);
    if (!c->bb.setup_group)
        c->bb.setup_group = c->K >= 8 ? 8u : c->K;
    c->bb.ntx = ntx;
    c->bb.nty = nty;
    c->bb.ntiles = ntx * nty;
    c->small_px = small_px;
    c->cfg.tile_px = T;
    c->cfg.small_tri_px = small_px;
    drop_graph(c);
    size_t nbins = (size_t)c->K * c->bb.ntiles;
    int rc;
    if ((rc = dalloc(c, &c->bb.count, nbins)) || (rc = dalloc(c, &c->bb.start, nbins)) ||
        (rc = dalloc(c, &c->bb.fill, nbins)) || (rc = dalloc(c, &c->bb.order, nbins)))
        return rc;
    return RAD_OK;
}

int need_scene(rad_ctx *c)
{
    if (c->npatch == 0)
        return fail(c, RAD_E_STATE, "no scene uploaded (rad_scene_upload)");
    return RAD_OK;
}

// Entry points whose F rows span all patches: not with a scene pack.
int need_plain(rad_ctx *c)
{
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    if (c->pk.nscene)
        return fail(c, RAD_E_STATE, "not available while a scene pack is loaded");
    return RAD_OK;
}

} // namespace

extern "C" {

int rad_create(const rad_config *cfg, rad_ctx **out)
{
    if (!cfg || !out)
        return RAD_E_ARG;
    *out = nullptr;
    if (cfg->hemicube < 2 || cfg->hemicube > RAD_MAX_W || (cfg->hemicube & 1))
        return RAD_E_ARG;
    if (cfg->hemicubes < 1 || cfg->hemicubes > RAD_MAX_K)
        return RAD_E_ARG;
    if (cfg->mode > RAD_MODE_PHYSICAL || cfg->select > RAD_SELECT_POWER)
        return RAD_E_ARG;
    rad_ctx *c = new (std::nothrow) rad_ctx();
    if (!c)
        return RAD_E_NOMEM;
    c->cfg = *cfg;
    if (c->cfg.ocl_workitems_x == 0)
        c->cfg.ocl_workitems_x = 4;
    c->K = cfg->hemicubes;
    // Config.cpp:34-45
    c->g.W = cfg->hemicube;
    c->g.TW = c->g.W * 2;
    c->g.TH = (uint32_t)((double)c->g.W * 1.5);
    c->g.RES = c->g.TW * c->g.TH;
    int rc = RAD_OK;
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess)
        e = hipStreamCreate(&c->st);
    if (e == hipSuccess)
        e = hipStreamCreate(&c->st2);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) {
        rad_destroy(c);
        return RAD_E_HIP;
    }
    if ((rc = dalloc(c, &c->dff, c->g.RES)) ||
        (rc = dalloc(c, &c->item_ids, (size_t)c->K * c->g.RES)) ||
        (rc = dalloc(c, &c->mvp, (size_t)c->K * kShooterStride)) || (rc = dalloc(c, &c->snap, (size_t)c->K * 3)) ||
        (rc = dalloc(c, &c->shooters, c->K)) || (rc = dalloc(c, &c->nshoot, 1)) ||
        (rc = dalloc(c, &c->counters, kCntN * kCntBlocks)) || (rc = dalloc(c, &c->sel, 4)) ||
        (rc = dalloc(c, &c->resid, 8)) || (rc = dalloc(c, &c->ctl, kCtlN)) ||
        (rc = dalloc(c, &c->pool, 1))) {
        rad_destroy(c);
        return rc;
    }
    if (hipMemsetAsync(c->counters, 0, kCntN * kCntBlocks * sizeof(uint32_t), c->st) != hipSuccess ||
        hipMemsetAsync(c->ctl, 0, kCtlN * sizeof(uint32_t), c->st) != hipSuccess ||
        hipMemsetAsync(c->pool, 0, sizeof(SelPool), c->st) != hipSuccess ||
        hipMemsetAsync(c->nshoot, 0, sizeof(uint32_t), c->st) != hipSuccess ||
        hipMemsetAsync(c->shooters, 0, c->K * sizeof(uint32_t), c->st) != hipSuccess) {
        rad_destroy(c);
        return RAD_E_HIP;
    }
    launch_dff(c->g, c->dff, c->st);
    if (hipStreamSynchronize(c->st) != hipSuccess) {
        rad_destroy(c);
        return RAD_E_HIP;
    }
    int ncu = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
        ncu = prop.multiProcessorCount;
    c->sel_blocks = ncu * 2;
    // screen tiles: 32x32 when enough hemicubes fill the GPU, else 16x16
    // (32x32 also when a hemicube would have more tiles than the bins hold)
    uint32_t T = cfg->tile_px ? cfg->tile_px : (c->K >= 4 ? 32u : 16u);
    if (!cfg->tile_px && ((c->g.TW + T - 1) / T) * ((c->g.TH + T - 1) / T) > kMaxTilesPerSlab)
        T = 32;
    if ((rc = set_tiles(c, T, cfg->small_tri_px ? cfg->small_tri_px : 64u))) {
        rad_destroy(c);
        return rc;
    }
    if ((rc = dalloc(c, &c->parts, (size_t)c->sel_blocks * 3)) ||
        (rc = rad::static_scratch_init(c))) {
        rad_destroy(c);
        return rc;
    }
    *out = c;
    return RAD_OK;
}

void rad_destroy(rad_ctx *c)
{
    if (!c)
        return;
    if (c->st)
        (void)hipStreamSynchronize(c->st);
    if (c->st2)
        (void)hipStreamSynchronize(c->st2);
    drop_graph(c);
    void *ptrs[] = {c->tri,  c->tri_patch, c->tsph, c->frames, c->refl,     c->area,   c->emis,
                    c->unshot, c->shot,   c->ff,  c->fblk,   c->hits,     c->dff,    c->item_ids,
                    c->item_depth, c->ffq, c->runs_slots, c->runs_offs, c->trip_h, c->trip_i,
                    c->trip_e,
                    c->mvp,  c->snap,     c->shooters, c->nshoot, c->counters, c->ctl, c->sel,
                    c->resid, c->bb.recs, c->bb.count, c->bb.start, c->bb.fill,
                    c->bb.list, c->bb.blk_nrec, c->bb.bbox, c->bb.clipq, c->cand, c->parts, c->bb.order,
                    c->fhi,  c->flo,      c->pk.sc_tbeg, c->pk.sc_pbeg, c->pk.sc_slots,
                    c->pk.sc_slot_beg, c->pk.sc_nsel, c->pk.sc_done, c->pk.sc_den, c->pk.sc_res,
                    c->pk.sc_shot,
                    c->pk.cand, c->pk.patch_scene, c->pk.pass, c->pk.slot_tbeg, c->pk.slot_tcnt,
                    c->pk.slot_pbeg, c->pk.sc_w, c->pk.sc_dffo, c->pk.slot_w, c->pk.slot_dffo,
                    c->pk.dffs, c->patch_quad, c->quad_pbeg, c->quad_nbr, c->qval,
                    c->viewfr, c->view_sh, c->pool};
    for (void *p : ptrs)
        dfree(p);
    for (hipEvent_t e : c->ev)
        (void)hipEventDestroy(e);
    for (hipEvent_t e : {c->ev_fork, c->ev_join})
        if (e)
            (void)hipEventDestroy(e);
    if (c->st2)
        (void)hipStreamDestroy(c->st2);
    if (c->st)
        (void)hipStreamDestroy(c->st);
    delete c;
}

const char *rad_last_error(const rad_ctx *c) { return c ? c->err.c_str() : "null context"; }

int rad_geometry(const rad_ctx *c, uint32_t *tw, uint32_t *th, uint32_t *res)
{
    if (!c)
        return RAD_E_ARG;
    if (tw)
        *tw = c->g.TW;
    if (th)
        *th = c->g.TH;
    if (res)
        *res = c->g.RES;
    return RAD_OK;
}

int rad_precompute_dff(rad_ctx *c, float *out)
{
    if (!c)
        return RAD_E_ARG;
    if (out) {
        HIPCHK(c, hipMemcpyAsync(out, c->dff, c->g.RES * sizeof(float), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    return RAD_OK;
}

} // extern "C"

namespace {

void pack_free(rad_ctx *c)
{
    PackBuffers &p = c->pk;
    void *ptrs[] = {p.sc_tbeg, p.sc_pbeg, p.sc_slots, p.sc_slot_beg, p.sc_nsel, p.sc_done, p.sc_den,
                    p.sc_res,  p.sc_shot, p.cand,    p.patch_scene, p.pass,     p.slot_tbeg,
                    p.slot_tcnt, p.slot_pbeg, p.sc_w, p.sc_dffo, p.slot_w, p.slot_dffo, p.dffs};
    for (void *q : ptrs)
        dfree(q);
    c->pk = PackBuffers{};
}

int check_scene(rad_ctx *c, const float *tri_xyz, const uint32_t *tri_patch, uint32_t ntri,
                uint32_t npatch, const float *frames, const float *refl, const float *area)
{
    if (!tri_xyz || !tri_patch || !frames || !refl || !area || npatch == 0)
        return fail(c, RAD_E_ARG, "scene: null array or zero patches");
    for (uint32_t t = 0; t < ntri; ++t) {
        if (tri_patch[t] >= npatch)
            return fail(c, RAD_E_ARG, "scene: tri_patch[" + std::to_string(t) + "] >= npatch");
        if (t && tri_patch[t] < tri_patch[t - 1])
            return fail(c, RAD_E_ARG, "scene: tri_patch must be non-decreasing (draw order)");
    }
    for (uint64_t i = 0; i < (uint64_t)ntri * 9; ++i)
        if (!std::isfinite(tri_xyz[i]) || std::fabs(tri_xyz[i]) > 1.0e4f)
            return fail(c, RAD_E_ARG, "scene: vertex coordinate not finite or beyond 1e4");
    for (uint32_t i = 0; i < npatch; ++i)
        if (!(area[i] > 0.0f))
            return fail(c, RAD_E_ARG, "scene: patch area must be > 0");
    return RAD_OK;
}

// Device copy of a scene; the rasterizer's record regions are sized for
// setup grids of nblk blocks per shooter and F rows of rowlen entries.
// Device buffers of a scene of ntri triangles and npatch patches (contents
// undefined), the rasterizer's record regions for setup grids of nblk blocks
// per shooter, and F rows of rowlen entries.
int scene_alloc(rad_ctx *c, uint32_t ntri, uint32_t npatch, uint32_t nblk, uint32_t rowlen)
{
    (void)hipStreamSynchronize(c->st);
    drop_graph(c);
    pack_free(c);
    c->nquad = 0;
    c->ntri = c->npatch = 0;
    c->hash_valid = false;
    int rc;
    size_t K = c->K;
    if ((rc = dalloc(c, &c->tri, (size_t)ntri * 9)) || (rc = dalloc(c, &c->tri_patch, ntri)) ||
        (rc = dalloc(c, &c->tsph, ntri)) ||
        (rc = dalloc(c, &c->frames, (size_t)npatch * 9)) || (rc = dalloc(c, &c->refl, (size_t)npatch * 3)) ||
        (rc = dalloc(c, &c->area, npatch)) || (rc = dalloc(c, &c->emis, (size_t)npatch * 3)) ||
        (rc = dalloc(c, &c->unshot, (size_t)npatch * 3)) || (rc = dalloc(c, &c->shot, (size_t)npatch * 3)) ||
        (rc = dalloc(c, &c->ff, K * rowlen)) || (rc = dalloc(c, &c->hits, K * rowlen)) ||
        (rc = dalloc(c, &c->fblk, (size_t)(rowlen + 63) / 64 * K)))
        return rc;
    if (c->det) {
        if ((rc = dalloc(c, &c->ffq, K * rowlen)))
            return rc;
        HIPCHK(c, hipMemset(c->ffq, 0, K * rowlen * 8));
    }
    // records: one region of kRecsPerBlock per (setup block, shooter) plus a
    // spill region; bin entries: three per record slot
    c->bb.nblk = nblk ? nblk : 1;
    c->bb.nz = K <= 4 ? 5u : 1u;
    uint64_t regions = (uint64_t)c->bb.nblk * K * c->bb.nz * kRecsPerBlock;
    uint64_t spill = regions / 8 + 65536;
    if (regions + kCntBlocks * spill > (1ull << 31))
        return fail(c, RAD_E_LIMIT, "scene too large for the rasterizer record regions");
    c->bb.spill_base = (uint32_t)regions;
    c->bb.spill_cap = (uint32_t)spill;
    c->bb.clip_cap = (uint32_t)(spill / 2);
    c->bb.clip_blocks = 512;
    uint64_t went = 2 * (regions + spill);
    c->bb.ent_cap = (uint32_t)(went > 0xF0000000ull ? 0xF0000000ull : went);
    // each half of a split pass has its own spill region, list area and
    // clipping queue
    const uint64_t nrec = regions + kCntBlocks * spill;
    if ((rc = dalloc(c, &c->bb.recs, nrec)) || (rc = dalloc(c, &c->bb.bbox, nrec)) ||
        (rc = dalloc(c, &c->bb.list, (size_t)kCntBlocks * c->bb.ent_cap)) ||
        (rc = dalloc(c, &c->bb.blk_nrec, (size_t)c->bb.nblk * K * c->bb.nz)) ||
        (rc = dalloc(c, &c->bb.clipq, (size_t)kCntBlocks * c->bb.clip_cap)) ||
        (rc = dalloc(c, &c->cand, 2 * topk_cand_half(npatch, c->K))))
        return rc;
    HIPCHK(c, hipMemset(c->ff, 0, K * rowlen * 4));
    HIPCHK(c, hipMemset(c->fblk, 0, (size_t)(rowlen + 63) / 64 * K * 4));
    c->ff_dirty = 0;
    HIPCHK(c, hipMemset(c->hits, 0, K * rowlen * 4));
    c->ntri = ntri;
    c->npatch = npatch;
    c->rowlen = rowlen;
    c->rendered = false;
    c->dense = false;
    return RAD_OK;
}

// 64-bit FNV-1a of a scene's arrays as the host hands them over (triangles
// interleaved, then patch ids, frames, reflectances, areas).
uint64_t scene_hash(uint32_t ntri, uint32_t npatch, const float *tri_xyz,
                    const uint32_t *tri_patch, const float *frames, const float *refl,
                    const float *area)
{
    uint64_t h = 0xCBF29CE484222325ull;
    h = fnv1a(h, &ntri, 4);
    h = fnv1a(h, &npatch, 4);
    h = fnv1a(h, tri_xyz, (size_t)ntri * 36);
    h = fnv1a(h, tri_patch, (size_t)ntri * 4);
    h = fnv1a(h, frames, (size_t)npatch * 36);
    h = fnv1a(h, refl, (size_t)npatch * 12);
    h = fnv1a(h, area, (size_t)npatch * 4);
    return h;
}

int upload_copy(rad_ctx *c, const float *tri_xyz, const uint32_t *tri_patch, uint32_t ntri,
                uint32_t npatch, const float *frames, const float *refl, const float *area)
{
    // triangles -> 9 planes
    std::vector<float> planes((size_t)ntri * 9);
    for (uint32_t t = 0; t < ntri; ++t)
        for (int k = 0; k < 9; ++k)
            planes[(size_t)k * ntri + t] = tri_xyz[(size_t)t * 9 + k];
    HIPCHK(c, hipMemcpy(c->tri, planes.data(), planes.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->tri_patch, tri_patch, (size_t)ntri * 4, hipMemcpyHostToDevice));
    launch_tri_sphere(c->tri, ntri, c->tsph, c->st);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpy(c->frames, frames, (size_t)npatch * 36, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->refl, refl, (size_t)npatch * 12, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->area, area, (size_t)npatch * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->emis, 0, (size_t)npatch * 12));
    HIPCHK(c, hipMemset(c->unshot, 0, (size_t)npatch * 12));
    HIPCHK(c, hipMemset(c->shot, 0, (size_t)npatch * 12));
    c->geom_hash = scene_hash(ntri, npatch, tri_xyz, tri_patch, frames, refl, area);
    c->hash_valid = true;
    return RAD_OK;
}

// A scene whose buffers were allocated but not filled is no scene.
int drop_scene(rad_ctx *c, int rc)
{
    c->ntri = c->npatch = c->nquad = 0;
    c->hash_valid = false;
    return rc;
}

int upload_impl(rad_ctx *c, const float *tri_xyz, const uint32_t *tri_patch, uint32_t ntri,
                uint32_t npatch, const float *frames, const float *refl, const float *area,
                uint32_t nblk, uint32_t rowlen)
{
    int rc;
    if ((rc = scene_alloc(c, ntri, npatch, nblk, rowlen)))
        return drop_scene(c, rc);
    if ((rc = upload_copy(c, tri_xyz, tri_patch, ntri, npatch, frames, refl, area)))
        return drop_scene(c, rc);
    return RAD_OK;
}

// The scene hash of what is on the device (an ingested scene is hashed on
// first use: its arrays are read back in the upload layout).
int ensure_hash(rad_ctx *c)
{
    if (c->hash_valid)
        return RAD_OK;
    const size_t nt = c->ntri, np = c->npatch;
    std::vector<float> planes(nt * 9), tri(nt * 9), fr(np * 9), rf(np * 3), ar(np);
    std::vector<uint32_t> tp(nt);
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(planes.data(), c->tri, nt * 36, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(tp.data(), c->tri_patch, nt * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(fr.data(), c->frames, np * 36, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(rf.data(), c->refl, np * 12, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(ar.data(), c->area, np * 4, hipMemcpyDeviceToHost));
    for (size_t t = 0; t < nt; ++t)
        for (int k = 0; k < 9; ++k)
            tri[t * 9 + k] = planes[(size_t)k * nt + t];
    c->geom_hash = scene_hash(c->ntri, c->npatch, tri.data(), tp.data(), fr.data(), rf.data(),
                              ar.data());
    c->hash_valid = true;
    return RAD_OK;
}

} // namespace

extern "C" {

int rad_scene_upload(rad_ctx *c, const float *tri_xyz, const uint32_t *tri_patch, uint32_t ntri,
                     uint32_t npatch, const float *frames, const float *refl, const float *area)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = check_scene(c, tri_xyz, tri_patch, ntri, npatch, frames, refl, area)))
        return rc;
    return upload_impl(c, tri_xyz, tri_patch, ntri, npatch, frames, refl, area, (ntri + 255) / 256,
                       npatch);
}

int rad_set_state(rad_ctx *c, const float *unshot, const float *shot, const float *emission)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    size_t n = (size_t)c->npatch * 12;
    if (unshot)
        HIPCHK(c, hipMemcpyAsync(c->unshot, unshot, n, hipMemcpyHostToDevice, c->st));
    if (shot)
        HIPCHK(c, hipMemcpyAsync(c->shot, shot, n, hipMemcpyHostToDevice, c->st));
    if (emission)
        HIPCHK(c, hipMemcpyAsync(c->emis, emission, n, hipMemcpyHostToDevice, c->st));
    else if (unshot)
        HIPCHK(c, hipMemcpyAsync(c->emis, unshot, n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return RAD_OK;
}

int rad_get_state(rad_ctx *c, float *unshot, float *shot)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    size_t n = (size_t)c->npatch * 12;
    if (unshot)
        HIPCHK(c, hipMemcpyAsync(unshot, c->unshot, n, hipMemcpyDeviceToHost, c->st));
    if (shot)
        HIPCHK(c, hipMemcpyAsync(shot, c->shot, n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return RAD_OK;
}

int rad_select(rad_ctx *c, uint32_t k, uint32_t *out_ids, uint32_t *out_n)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_plain(c)))
        return rc;
    if (k < 1 || k > c->K)
        return fail(c, RAD_E_ARG, "select: k must be in [1, hemicubes]");
    if ((rc = select_dev(c, k)))
        return rc;
    uint32_t n = 0;
    HIPCHK(c, hipMemcpyAsync(&n, c->nshoot, 4, hipMemcpyDeviceToHost, c->st));
    if (out_ids)
        HIPCHK(c, hipMemcpyAsync(out_ids, c->shooters, (size_t)k * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (out_n)
        *out_n = n;
    return RAD_OK;
}

static int upload_shooters(rad_ctx *c, const uint32_t *shooters, uint32_t k)
{
    if (!shooters || k < 1 || k > c->K)
        return fail(c, RAD_E_ARG, "shooters: k must be in [1, hemicubes]");
    for (uint32_t s = 0; s < k; ++s) {
        if (shooters[s] >= c->npatch)
            return fail(c, RAD_E_ARG, "shooters: patch index out of range");
        for (uint32_t r = 0; r < s; ++r)
            if (shooters[r] == shooters[s])
                return fail(c, RAD_E_ARG, "shooters: duplicate patch index");
    }
    HIPCHK(c, hipMemcpyAsync(c->shooters, shooters, (size_t)k * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->nshoot, &k, 4, hipMemcpyHostToDevice, c->st));
    // the host value k must outlive the async copy
    HIPCHK(c, hipStreamSynchronize(c->st));
    return RAD_OK;
}

int rad_render_itembuf(rad_ctx *c, const uint32_t *shooters, uint32_t k, uint32_t *out_ids,
                       float *out_depth)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_plain(c)) || (rc = upload_shooters(c, shooters, k)))
        return rc;
    if (out_depth && !c->item_depth && (rc = dalloc(c, &c->item_depth, (size_t)c->K * c->g.RES)))
        return rc;
    if ((rc = render_pass(c, k, 0, c->item_ids, out_depth ? c->item_depth : nullptr, nullptr)))
        return rc;
    c->rendered = true;
    c->rendered_k = k;
    size_t n = (size_t)k * c->g.RES;
    if (out_ids)
        HIPCHK(c, hipMemcpyAsync(out_ids, c->item_ids, n * 4, hipMemcpyDeviceToHost, c->st));
    if (out_depth)
        HIPCHK(c, hipMemcpyAsync(out_depth, c->item_depth, n * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return check_overflow(c);
}

static int ensure_events(rad_ctx *c, size_t n);
static int add_stage_times(rad_ctx *c, uint32_t passes);

// Standalone ProcessHemicube of the item buffer in item_ids (k hemicubes),
// timed with HIP events when timing is on (rad_reduce_kernel_ms).
static int reduce_out(rad_ctx *c, uint32_t k, uint32_t *out_hits, float *out_ff)
{
    // F rows [0, max(k, ff_dirty)) and hit rows [0, k) start from zero
    uint32_t rows = c->ff_dirty > k ? c->ff_dirty : k;
    HIPCHK(c, hipMemsetAsync(c->ff, 0, (size_t)rows * c->rowlen * 4, c->st));
    HIPCHK(c, hipMemsetAsync(c->hits, 0, (size_t)k * c->npatch * 4, c->st));
    HIPCHK(c, hipMemsetAsync(c->counters, 0, kCntN * kCntBlocks * 4, c->st));
    int rc;
    if (c->timing && (rc = ensure_events(c, 2)))
        return rc;
    record(c, 0);
    launch_reduce(c->g, k, c->nshoot, c->item_ids, c->dff, c->npatch, c->ff,
                  c->det ? c->ffq : nullptr, out_hits ? c->hits : nullptr, c->counters, c->st);
    record(c, 1);
    if (c->det)
        launch_fold(c->ffq, c->ff, (size_t)k * c->npatch, c->st);
    c->ff_dirty = k;
    if ((rc = check_launch(c, "reduce")))
        return rc;
    size_t nf = (size_t)k * c->npatch;
    if (out_hits)
        HIPCHK(c, hipMemcpyAsync(out_hits, c->hits, nf * 4, hipMemcpyDeviceToHost, c->st));
    if (out_ff)
        HIPCHK(c, hipMemcpyAsync(out_ff, c->ff, nf * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (c->timing) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
        c->t_reduce = ms;
    }
    return RAD_OK;
}

int rad_process_hemicube(rad_ctx *c, uint32_t k, uint32_t *out_hits, float *out_ff)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_plain(c)))
        return rc;
    if (!c->rendered || k != c->rendered_k)
        return fail(c, RAD_E_STATE, "process_hemicube: k must match the last render");
    return reduce_out(c, k, out_hits, out_ff);
}

int rad_process_itembuf(rad_ctx *c, const uint32_t *itembuf, uint32_t k, uint32_t *out_hits,
                        float *out_ff)
{
    if (!c || !itembuf)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_plain(c)))
        return rc;
    if (k < 1 || k > c->K)
        return fail(c, RAD_E_ARG, "process_itembuf: k must be in [1, hemicubes]");
    size_t n = (size_t)k * c->g.RES;
    HIPCHK(c, hipMemcpyAsync(c->item_ids, itembuf, n * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->nshoot, &k, 4, hipMemcpyHostToDevice, c->st));
    // k must outlive the async copy
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->rendered = true;
    c->rendered_k = k;
    return reduce_out(c, k, out_hits, out_ff);
}

int rad_process_runs(rad_ctx *c, uint32_t k, uint32_t cap, uint32_t *out_hemicubes,
                     uint32_t *out_ids, float *out_energies, uint32_t *out_count)
{
    if (!c || !out_count)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_plain(c)))
        return rc;
    if (!c->rendered || k != c->rendered_k)
        return fail(c, RAD_E_STATE, "process_runs: k must match the last render");
    if (cap && (!out_hemicubes || !out_ids || !out_energies))
        return fail(c, RAD_E_ARG, "process_runs: null output arrays");
    const uint32_t wx = c->cfg.ocl_workitems_x;
    const uint32_t nx = wx < c->g.TW ? wx : c->g.TW; // Config.cpp:43-45
    const uint64_t items64 = (uint64_t)k * c->g.TH * nx;
    if (items64 >= (1ull << 31))
        return fail(c, RAD_E_LIMIT, "process_runs: too many work-items");
    const uint32_t items = (uint32_t)items64;
    if (items > c->runs_items) {
        if ((rc = dalloc(c, &c->runs_slots, items)) || (rc = dalloc(c, &c->runs_offs, items + 1))) {
            c->runs_items = 0;
            return rc;
        }
        c->runs_items = items;
    }
    launch_runs_count(c->g, k, nx, c->item_ids, c->runs_slots, c->runs_offs, c->st);
    if ((rc = check_launch(c, "process_runs count")))
        return rc;
    uint32_t total = 0;
    HIPCHK(c, hipMemcpyAsync(&total, c->runs_offs + items, 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    *out_count = total;
    if (total > cap)
        return fail(c, RAD_E_LIMIT, "process_runs: " + std::to_string(total) +
                                        " triplets, output arrays hold " + std::to_string(cap));
    if (total > c->trip_cap) {
        if ((rc = dalloc(c, &c->trip_h, total)) || (rc = dalloc(c, &c->trip_i, total)) ||
            (rc = dalloc(c, &c->trip_e, total))) {
            c->trip_cap = 0;
            return rc;
        }
        c->trip_cap = total;
    }
    launch_runs_write(c->g, k, nx, c->item_ids, c->dff, c->runs_offs, total, c->trip_h, c->trip_i,
                      c->trip_e, c->st);
    if ((rc = check_launch(c, "process_runs write")))
        return rc;
    if (total) {
        HIPCHK(c, hipMemcpyAsync(out_hemicubes, c->trip_h, (size_t)total * 4,
                                 hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipMemcpyAsync(out_ids, c->trip_i, (size_t)total * 4, hipMemcpyDeviceToHost,
                                 c->st));
        HIPCHK(c, hipMemcpyAsync(out_energies, c->trip_e, (size_t)total * 4,
                                 hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    return RAD_OK;
}

int rad_shoot(rad_ctx *c, const uint32_t *shooters, uint32_t k)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_plain(c)) || (rc = upload_shooters(c, shooters, k)))
        return rc;
    if (c->timing && (rc = ensure_events(c, 7)))
        return rc;
    uint32_t saveK = c->K;
    c->K = k;
    rc = one_pass(c, false, 0);
    c->K = saveK;
    if (rc || (c->timing && (rc = add_stage_times(c, 1))))
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return check_overflow(c);
}

static int ensure_events(rad_ctx *c, size_t n)
{
    while (c->ev.size() < n) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        c->ev.push_back(e);
    }
    return RAD_OK;
}

// Captures graph_passes passes (kernels only: every clear in a pass is a
// kernel, no memset or copy node) into one executable graph.
static int capture_graph(rad_ctx *c, uint32_t n)
{
    drop_graph(c);
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    int rc = RAD_OK;
    // inside a replay nothing else touches the state between passes, so
    // every pass but the last selects and prepares its successor.  Every
    // kPoolRearm passes (from the first) a pass's shoot leaves lists and the
    // next selection arms the pool from them.  Armed after pass p's shoot,
    // the pool serves the selections of passes p + 2 .. q (q: the pass
    // before the next such shoot's successor, at most n - 1); each pass
    // shoots K keys at or above the bound while the others only grow (all
    // additions are of non-negative terms), so (q - p) K keys at or above
    // it keep every pool at K or more.  Whatever the bound, a pool is used
    // only when it holds K keys.
    const bool fuse = fused_select(c);
    for (uint32_t p = 0; p < n && rc == RAD_OK; ++p) {
        const uint32_t q = std::min(kPoolRearm * (p / kPoolRearm + 1), n - 1);
        rc = one_pass(c, true, 0, fuse && p > 0, fuse && p + 1 < n,
                      fuse && q >= p + 2 ? (q - p) * c->K : 0u, p % kPoolRearm == 0);
    }
    hipError_t e = hipStreamEndCapture(c->st, &graph);
    if (rc)
        return rc;
    if (e != hipSuccess)
        return fail(c, RAD_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        c->gexec = nullptr;
        return fail(c, RAD_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    c->gpasses = n;
    return RAD_OK;
}

// With timing on: waits for the last `passes` passes and adds their stage
// times (select, prepare + clear, setup, bins, tiles, shoot) to the totals.
static int add_stage_times(rad_ctx *c, uint32_t passes)
{
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (uint32_t p = 0; p < passes; ++p) {
        for (int k = 0; k < 6; ++k) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, c->ev[(size_t)p * 7 + k], c->ev[(size_t)p * 7 + k + 1]);
            c->t_acc[k] += ms;
        }
    }
    c->t_passes += passes;
    return RAD_OK;
}

static int step_impl(rad_ctx *c, uint32_t passes)
{
    int rc;
    if (c->timing && (rc = ensure_events(c, (size_t)passes * 7)))
        return rc;
    uint32_t G = c->cfg.graph_passes;
    if (!c->timing && G && passes >= G) {
        // captured passes assume zero F rows at their start (the first one
        // skips clearing them); every replay ends with a shoot, which leaves
        // them zero again
        if (c->ff_dirty) {
            launch_clear(c->K, c->bb.count, c->bb.ntiles, c->ff, nullptr, c->ff_dirty, c->rowlen,
                         c->counters, c->bb.ent_cap, c->st);
            if ((rc = check_launch(c, "clear")))
                return rc;
            c->ff_dirty = 0;
        }
        if (!c->gexec || c->gpasses != G)
            if ((rc = capture_graph(c, G)))
                return rc;
        for (; passes >= G; passes -= G)
            HIPCHK(c, hipGraphLaunch(c->gexec, c->st));
        c->rendered = false;
    }
    for (uint32_t p = 0; p < passes; ++p)
        if ((rc = one_pass(c, true, (size_t)p * 7)))
            return rc;
    return c->timing ? add_stage_times(c, passes) : RAD_OK;
}

int rad_step(rad_ctx *c, uint32_t passes)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    if ((rc = step_impl(c, passes)))
        return rc;
    return RAD_OK;
}

int rad_prepare(rad_ctx *c)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    const uint32_t G = c->cfg.graph_passes;
    if (c->timing || !G)
        return RAD_OK;
    if (!c->gexec || c->gpasses != G) {
        // captured as rad_step captures, from zero F rows (rad_step clears
        // dirty rows before it replays); the capture runs nothing, so the
        // record of dirty rows and of the last render stays as it was
        const uint32_t dirty = c->ff_dirty;
        const bool rendered = c->rendered;
        c->ff_dirty = 0;
        rc = capture_graph(c, G);
        c->ff_dirty = dirty;
        c->rendered = rendered;
        if (rc)
            return rc;
    }
    HIPCHK(c, hipGraphUpload(c->gexec, c->st));
    return RAD_OK;
}

int rad_finish(rad_ctx *c)
{
    if (!c)
        return RAD_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return c->npatch ? check_overflow(c) : RAD_OK;
}

int rad_residual_fp64(rad_ctx *c, double *out)
{
    if (!c || !out)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    launch_residual(c->unshot, c->area, c->emis, c->npatch, c->parts, nullptr, c->resid + 4,
                    c->sel_blocks, c->st);
    if ((rc = check_launch(c, "residual")))
        return rc;
    double r[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(r, c->resid + 4, 16, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    *out = r[1] > 0 ? r[0] / r[1] : 0.0;
    return RAD_OK;
}

} // extern "C"

namespace {

// The control block as the host sees it after a synchronisation.
struct CtlView {
    uint64_t shooters = 0, passes = 0;
    float last[3] = {0, 0, 0};
    bool stopped = false;
};

int read_ctl(rad_ctx *c, CtlView &v)
{
    uint32_t w[kCtlN];
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(w, c->ctl, sizeof(w), hipMemcpyDeviceToHost));
    v.stopped = w[kCtlStop] != 0;
    memcpy(v.last, &w[kCtlLastE], 12);
    v.shooters = (uint64_t)w[kCtlShotLo] | ((uint64_t)w[kCtlShotLo + 1] << 32);
    v.passes = (uint64_t)w[kCtlPassLo] | ((uint64_t)w[kCtlPassLo + 1] << 32);
    return RAD_OK;
}

// Arms (threshold > 0) or disarms (0) the stop test; clears the stop flag.
int set_stop(rad_ctx *c, double threshold)
{
    uint32_t w[4] = {0u, 0u, 0u, 0u}; // words kCtlStop .. kCtlThresh + 1
    static_assert(kCtlStop == 0 && kCtlThresh == 2, "control block layout");
    memcpy(&w[kCtlThresh], &threshold, 8);
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(c->ctl, w, sizeof(w), hipMemcpyHostToDevice));
    return RAD_OK;
}

float length3(const float *v)
{
    // Vector3f::f_Length (Vector.h:351-354): the sum of squares in float
    return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

} // namespace

extern "C" {

int rad_run(rad_ctx *c, uint64_t max_passes, double target, uint32_t check_every, rad_stats *out)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    if (check_every == 0)
        check_every = c->cfg.shoots ? c->cfg.shoots : 500;
    CtlView v0;
    if ((rc = read_ctl(c, v0)))
        return rc;
    auto t0 = std::chrono::steady_clock::now();
    uint64_t done = 0;
    double r = 0.0;
    if ((rc = rad_residual_fp64(c, &r)))
        return rc;
    while (done < max_passes && !(r <= target)) {
        uint64_t n = max_passes - done;
        if (n > check_every)
            n = check_every;
        if ((rc = step_impl(c, (uint32_t)n)))
            return rc;
        done += n;
        if ((rc = rad_residual_fp64(c, &r)) || (rc = check_overflow(c)))
            return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    auto t1 = std::chrono::steady_clock::now();
    CtlView v1;
    if ((rc = read_ctl(c, v1)))
        return rc;
    if (out) {
        out->passes = done;
        out->shooters = v1.shooters - v0.shooters;
        out->residual = r;
        out->seconds = std::chrono::duration<double>(t1 - t0).count();
        out->last_energy = done ? length3(v1.last) : 0.0f;
        out->reserved = 0;
    }
    return RAD_OK;
}

int rad_run_cycle(rad_ctx *c, uint32_t passes, double stop_energy, rad_cycle_stats *out)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_plain(c)))
        return rc;
    if (!(stop_energy >= 0.0) || !std::isfinite(stop_energy))
        return fail(c, RAD_E_ARG, "run_cycle: stop_energy must be finite and >= 0");
    CtlView v0, v1;
    if ((rc = read_ctl(c, v0)) || (rc = set_stop(c, stop_energy)))
        return rc;
    rc = step_impl(c, passes);
    int rc2 = read_ctl(c, v1);
    int rc3 = set_stop(c, 0.0);
    if (rc || (rc = rc2) || (rc = rc3) || (rc = check_overflow(c)))
        return rc;
    if (out) {
        out->passes = v1.passes - v0.passes;
        out->shooters = v1.shooters - v0.shooters;
        for (int k = 0; k < 3; ++k)
            out->last_energy[k] = out->passes ? v1.last[k] : 0.0f;
        out->stopped = v1.stopped ? 1u : 0u;
    }
    return RAD_OK;
}

int rad_set_deterministic(rad_ctx *c, int on)
{
    if (!c)
        return RAD_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (!on) {
        dfree(c->ffq);
        c->ffq = nullptr;
        c->det = false;
        return RAD_OK;
    }
    if (!c->ffq && c->rowlen) {
        size_t n = (size_t)c->K * c->rowlen;
        int rc;
        if ((rc = dalloc(c, &c->ffq, n)))
            return rc;
        HIPCHK(c, hipMemset(c->ffq, 0, n * 8));
    }
    drop_graph(c); // captured passes name the F accumulation buffers
    c->det = true;
    return RAD_OK;
}

int rad_shot_count(rad_ctx *c, uint64_t *passes, uint64_t *shooters)
{
    if (!c)
        return RAD_E_ARG;
    CtlView v;
    int rc;
    if ((rc = read_ctl(c, v)))
        return rc;
    if (passes)
        *passes = v.passes;
    if (shooters)
        *shooters = v.shooters;
    return RAD_OK;
}

int rad_face_mvps(rad_ctx *c, uint32_t patch, float *out80)
{
    if (!c || !out80)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    if (patch >= c->npatch)
        return fail(c, RAD_E_ARG, "face_mvps: patch out of range");
    uint32_t one = 1;
    HIPCHK(c, hipMemcpyAsync(c->shooters, &patch, 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->nshoot, &one, 4, hipMemcpyHostToDevice, c->st));
    launch_prepare(c->g, 1, c->shooters, c->nshoot, c->frames, c->unshot, c->snap, c->mvp, c->st);
    if ((rc = check_launch(c, "prepare")))
        return rc;
    HIPCHK(c, hipMemcpyAsync(out80, c->mvp, 80 * sizeof(float), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return RAD_OK;
}

int rad_kernel_times(rad_ctx *c, int enable, double *out6, uint64_t *npasses)
{
    if (!c)
        return RAD_E_ARG;
    if (out6)
        for (int k = 0; k < 6; ++k)
            out6[k] = c->t_passes ? c->t_acc[k] / (double)c->t_passes : 0.0;
    if (npasses)
        *npasses = c->t_passes;
    if (enable >= 0) {
        c->timing = enable != 0;
        for (int k = 0; k < 6; ++k)
            c->t_acc[k] = 0;
        c->t_passes = 0;
    }
    return RAD_OK;
}

int rad_reduce_kernel_ms(rad_ctx *c, double *ms)
{
    if (!c || !ms)
        return RAD_E_ARG;
    *ms = c->t_reduce;
    return RAD_OK;
}

int rad_autotune(rad_ctx *c, uint32_t passes, rad_tune_result *out)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_plain(c)))
        return rc;
    if (passes == 0)
        passes = 3;
    rad_tune_result r{};
    const uint32_t tiles[3] = {8, 16, 32}, smalls[6] = {16, 32, 48, 64, 128, 256};
    const uint32_t T0 = c->cfg.tile_px, S0 = c->small_px;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(c, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(c, RAD_E_HIP, "hipEventCreate");
    }
    bool timing = c->timing;
    c->timing = false;
    // K shooters spread evenly over the patch list: early in a solve the
    // selection holds only the few emitters, which is no sample of the work
    std::vector<uint32_t> spread(c->K);
    for (uint32_t s = 0; s < c->K; ++s)
        spread[s] = (uint32_t)(((uint64_t)s * c->npatch + c->npatch / 2) / c->K) % c->npatch;
    uint32_t kk = c->K;
    rc = RAD_OK;
    if (hipMemcpyAsync(c->shooters, spread.data(), (size_t)c->K * 4, hipMemcpyHostToDevice, c->st) !=
            hipSuccess ||
        hipMemcpyAsync(c->nshoot, &kk, 4, hipMemcpyHostToDevice, c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess)
        rc = fail(c, RAD_E_HIP, "autotune: shooter upload");
    if (rc == RAD_OK) {
        for (uint32_t ti = 0; ti < 3 && rc == RAD_OK; ++ti) {
            if (tiles[ti] == 8 && c->K > 2)
                continue;
            for (uint32_t si = 0; si < 6 && rc == RAD_OK && r.ncand < RAD_TUNE_MAX; ++si) {
                if ((rc = set_tiles(c, tiles[ti], smalls[si])))
                    break;
                rc = render_pass(c, c->K, 0, nullptr, nullptr, nullptr); // warm
                (void)hipEventRecord(e0, c->st);
                for (uint32_t p = 0; p < passes && rc == RAD_OK; ++p)
                    rc = render_pass(c, c->K, 0, nullptr, nullptr, nullptr);
                (void)hipEventRecord(e1, c->st);
                if (rc)
                    break;
                float ms = 0.f;
                if (hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    rc = fail(c, RAD_E_HIP, "autotune timing");
                    break;
                }
                uint32_t k = r.ncand++;
                r.cand_tile[k] = tiles[ti];
                r.cand_small[k] = smalls[si];
                r.cand_ms[k] = check_overflow(c) == RAD_OK ? ms / passes : -1.0;
                if (r.cand_ms[k] >= 0 && (r.best_ms == 0.0 || r.cand_ms[k] < r.best_ms)) {
                    r.best_ms = r.cand_ms[k];
                    r.tile_px = tiles[ti];
                    r.small_tri_px = smalls[si];
                }
            }
        }
    }
    c->timing = timing;
    c->rendered = false;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == RAD_OK)
        c->err.clear(); // overflowing candidates are reported in cand_ms only
    if (rc == RAD_OK && r.tile_px == 0)
        rc = fail(c, RAD_E_LIMIT, "autotune: every candidate overflowed the bins");
    int rc2 = set_tiles(c, rc == RAD_OK ? r.tile_px : T0, rc == RAD_OK ? r.small_tri_px : S0);
    if (rc)
        return rc;
    if (rc2)
        return rc2;
    if (out)
        *out = r;
    return RAD_OK;
}

int rad_debug_pool(rad_ctx *c, uint32_t cap, uint32_t *out4)
{
    if (!c || cap > kPoolCap)
        return RAD_E_ARG;
    if (cap) {
        c->pool_cap = cap;
        drop_graph(c); // the cap is an argument of the captured kernels
    }
    if (out4) {
        HIPCHK(c, hipStreamSynchronize(c->st));
        // the header: tau (two words), armed, hits, falls, arms
        static_assert(offsetof(SelPool, armed) == 8 && offsetof(SelPool, arms) == 20,
                      "pool header layout");
        uint32_t head[6];
        HIPCHK(c, hipMemcpy(head, c->pool, sizeof(head), hipMemcpyDeviceToHost));
        out4[0] = head[3];
        out4[1] = head[4];
        out4[2] = head[5];
        out4[3] = head[2];
    }
    return RAD_OK;
}

int rad_debug_counters(rad_ctx *c, uint32_t *out8, uint32_t *caps)
{
    if (!c || !out8)
        return RAD_E_ARG;
    static_assert(kCntN == 8, "eight pass counters");
    HIPCHK(c, hipStreamSynchronize(c->st));
    uint32_t cnt[kCntN * kCntBlocks];
    HIPCHK(c, hipMemcpy(cnt, c->counters, sizeof(cnt), hipMemcpyDeviceToHost));
    // the halves of a split pass: flags OR, unknown ids add up, and the
    // fill levels are the fuller half's (each half has the capacities)
    for (int i = 0; i < kCntN; ++i) {
        out8[i] = 0;
        for (int h = 0; h < kCntBlocks; ++h) {
            const uint32_t v = cnt[h * kCntN + i];
            if (i == kCntOverflow || i == kCntSticky)
                out8[i] |= v;
            else if (i == kCntUnknown || i == kCntRec)
                out8[i] += v;
            else
                out8[i] = out8[i] > v ? out8[i] : v;
        }
    }
    if (caps) {
        caps[0] = c->bb.spill_cap;
        caps[1] = c->bb.ent_cap;
        caps[2] = c->bb.clip_cap;
    }
    return RAD_OK;
}

int rad_debug_bins(rad_ctx *c, uint32_t *out, uint32_t *ntiles, uint32_t *tile_px)
{
    if (!c)
        return RAD_E_ARG;
    if (ntiles)
        *ntiles = c->bb.ntiles;
    if (tile_px)
        *tile_px = 1u << c->bb.tshift;
    if (out)
        HIPCHK(c, hipMemcpy(out, c->bb.count, (size_t)c->K * c->bb.ntiles * 4, hipMemcpyDeviceToHost));
    return RAD_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Dense matrix and Jacobi gather (rad_gather.hip kernels).
namespace {

uint32_t dense_pad(uint32_t n) { return (n + 511u) & ~511u; }

// Split-K factor of the gather product: enough workgroups (128 rows each)
// to give every CU two; with many columns, 8 k slices (one per XCD, so each
// XCD's L2 holds its slice of X).  npad / S stays a multiple of 64.
uint32_t gather_split(uint32_t npad, uint32_t R)
{
    uint32_t rows = npad / 128, S = 1;
    while (S < 8 && (rows * S < 512 || (R > 16 && S < 8)))
        S *= 2;
    return S;
}

int dense_alloc(rad_ctx *c)
{
    uint32_t npad = dense_pad(c->npatch);
    if (c->fhi && c->npad == npad)
        return RAD_OK;
    dfree(c->fhi);
    dfree(c->flo);
    c->fhi = c->flo = nullptr;
    c->dense = false;
    size_t n = (size_t)npad * npad;
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    if (fr && 4 * n + ((size_t)512 << 20) > fr)
        return fail(c, RAD_E_NOMEM, "dense matrix needs " + std::to_string(4 * n) +
                                        " bytes, " + std::to_string(fr) + " free");
    int rc;
    if ((rc = dalloc(c, &c->fhi, n)) || (rc = dalloc(c, &c->flo, n)))
        return rc;
    c->npad = npad;
    HIPCHK(c, hipMemsetAsync(c->fhi, 0, n * 2, c->st));
    HIPCHK(c, hipMemsetAsync(c->flo, 0, n * 2, c->st));
    return RAD_OK;
}

int need_dense(rad_ctx *c)
{
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    if (!c->dense || !c->fhi)
        return fail(c, RAD_E_STATE, "no dense matrix (rad_dense_build / rad_dense_upload)");
    return RAD_OK;
}

// Work buffers of one gather call (R columns, rp = 16 * tiles).
struct GatherBufs {
    uint32_t rt = 1, rp = 16, S = 1, nblk = 1024;
    GemmSplit ks1{}, ks2{}; // k slices of the single-half and two-half products
    float *yp = nullptr, *E = nullptr, *B = nullptr, *rv = nullptr, *out = nullptr;
    uint16_t *xhi = nullptr, *xlo = nullptr;
    double *parts = nullptr;
    ~GatherBufs()
    {
        void *p[] = {yp, E, B, rv, out, xhi, xlo, parts};
        for (void *q : p)
            dfree(q);
    }
};

int gather_alloc(rad_ctx *c, GatherBufs &g, uint32_t R)
{
    g.rt = gemm_tiles(R);
    g.rp = g.rt * 16;
    g.S = gather_split(c->npad, R);
    int ncu = 0;
    HIPCHK(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    g.ks1 = gemm_split(g.rt, false, c->npad, g.S, ncu > 0 ? (uint32_t)ncu : 1u);
    g.ks2 = gemm_split(g.rt, true, c->npad, g.S, ncu > 0 ? (uint32_t)ncu : 1u);
    const uint32_t planes = std::max(g.S, std::max(g.ks1.S2, g.ks2.S2));
    size_t plane = (size_t)g.rp * c->npad;
    int rc;
    if ((rc = dalloc(c, &g.yp, plane * planes)) || (rc = dalloc(c, &g.E, plane)) ||
        (rc = dalloc(c, &g.B, plane)) || (rc = dalloc(c, &g.rv, plane)) ||
        (rc = dalloc(c, &g.out, (size_t)R * c->npatch)) || (rc = dalloc(c, &g.xhi, plane)) ||
        (rc = dalloc(c, &g.xlo, plane)) || (rc = dalloc(c, &g.parts, g.nblk)))
        return rc;
    HIPCHK(c, hipMemsetAsync(g.E, 0, plane * 4, c->st));
    HIPCHK(c, hipMemsetAsync(g.B, 0, plane * 4, c->st));
    HIPCHK(c, hipMemsetAsync(g.rv, 0, plane * 4, c->st));
    HIPCHK(c, hipMemsetAsync(g.xhi, 0, plane * 2, c->st));
    HIPCHK(c, hipMemsetAsync(g.xlo, 0, plane * 2, c->st));
    return RAD_OK;
}

int sum_parts(rad_ctx *c, GatherBufs &g, double *out)
{
    std::vector<double> h(g.nblk);
    HIPCHK(c, hipMemcpyAsync(h.data(), g.parts, g.nblk * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    double s = 0.0;
    for (double v : h)
        s += v;
    *out = s;
    return RAD_OK;
}

} // namespace

extern "C" {

// The shooting-layout build as replays of a captured graph of
// kDenseGraphPasses passes: each pass takes the next K rows from a device
// counter (k_iota_next), renders them and stores them transposed; passes
// beyond the last patch select nobody and store nothing.  The F rows are
// zero at the start of every captured pass (the transposed store zeroes
// what it reads).
static int dense_build_graph(rad_ctx *c, uint32_t npass)
{
    if (c->ff_dirty) {
        launch_clear(c->K, c->bb.count, c->bb.ntiles, c->ff, nullptr, c->ff_dirty, c->rowlen,
                     c->counters, c->bb.ent_cap, c->st);
        c->ff_dirty = 0;
    }
    const uint32_t start = 0u - c->K; // the first pass advances it to row 0
    HIPCHK(c, hipMemcpyAsync(c->ctl + kCtlDenseBase, &start, 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    int rc = RAD_OK;
    for (uint32_t p = 0; p < kDenseGraphPasses && rc == RAD_OK; ++p) {
        launch_iota_next(c->K, c->npatch, c->ctl, c->shooters, c->nshoot, c->st);
        if ((rc = render_pass(c, c->K, 0, nullptr, nullptr, nullptr)))
            break;
        launch_dense_store_t_dev(c->ff, c->K, c->ctl, c->nshoot, c->npatch, c->npad, c->area,
                                 c->fhi, c->flo, c->st);
        c->ff_dirty = 0;
        rc = check_launch(c, "dense store");
    }
    hipError_t e = hipStreamEndCapture(c->st, &graph);
    if (rc) {
        if (graph)
            (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess)
        return fail(c, RAD_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess)
        return fail(c, RAD_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    for (uint32_t r = 0; r < npass && e == hipSuccess; r += kDenseGraphPasses)
        e = hipGraphLaunch(exec, c->st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->st);
    (void)hipGraphExecDestroy(exec);
    if (e != hipSuccess)
        return fail(c, RAD_E_HIP, std::string("dense build graph: ") + hipGetErrorString(e));
    return RAD_OK;
}

int rad_dense_build(rad_ctx *c, uint32_t layout)
{
    if (!c)
        return RAD_E_ARG;
    if (layout > RAD_DENSE_SHOOT)
        return fail(c, RAD_E_ARG, "dense_build: unknown layout");
    int rc;
    if ((rc = need_plain(c)) || (rc = dense_alloc(c)))
        return rc;
    bool timing = c->timing;
    c->timing = false;
    const uint32_t npass = (c->npatch + c->K - 1) / c->K;
    if (layout == RAD_DENSE_SHOOT && npass >= 2 * kDenseGraphPasses) {
        rc = dense_build_graph(c, npass);
    } else {
        for (uint32_t base = 0; base < c->npatch; base += c->K) {
            uint32_t k = c->npatch - base < c->K ? c->npatch - base : c->K;
            launch_iota_shooters(base, k, c->shooters, c->nshoot, c->st);
            if ((rc = render_pass(c, k, 0, nullptr, nullptr, nullptr)))
                break;
            if (layout == RAD_DENSE_ROWS) {
                launch_dense_store(c->ff, k, base, c->npatch, c->npad, c->fhi, c->flo, c->st);
            } else {
                launch_dense_store_t(c->ff, k, base, c->npatch, c->npad, c->area, c->fhi, c->flo,
                                     c->st);
                c->ff_dirty = 0; // the transposed store zeroes the F entries it reads
            }
            if ((rc = check_launch(c, "dense store")))
                break;
        }
    }
    c->timing = timing;
    c->rendered = false;
    if (rc)
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if ((rc = check_overflow(c)))
        return rc;
    c->dense = true;
    return RAD_OK;
}

int rad_dense_upload(rad_ctx *c, const float *m)
{
    if (!c || !m)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_plain(c)) || (rc = dense_alloc(c)))
        return rc;
    size_t n = (size_t)c->npatch * c->npatch;
    float *tmp = nullptr;
    if ((rc = dalloc(c, &tmp, n)))
        return rc;
    hipError_t e = hipMemcpy(tmp, m, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_dense_from_f32(tmp, c->npatch, 0, c->npatch, c->npad, c->fhi, c->flo, c->st);
        e = hipStreamSynchronize(c->st);
    }
    dfree(tmp);
    if (e != hipSuccess)
        return fail(c, RAD_E_HIP, std::string("dense_upload: ") + hipGetErrorString(e));
    c->dense = true;
    return RAD_OK;
}

int rad_dense_download(rad_ctx *c, uint32_t row0, uint32_t nrows, float *out)
{
    if (!c || !out)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_dense(c)))
        return rc;
    if (row0 > c->npatch || nrows > c->npatch - row0)
        return fail(c, RAD_E_ARG, "dense_download: rows out of range");
    size_t n = (size_t)nrows * c->npatch;
    float *tmp = nullptr;
    if ((rc = dalloc(c, &tmp, n)))
        return rc;
    launch_dense_to_f32(c->fhi, c->flo, nrows, row0, c->npatch, c->npad, tmp, c->st);
    hipError_t e = hipMemcpyAsync(out, tmp, n * 4, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->st);
    dfree(tmp);
    if (e != hipSuccess)
        return fail(c, RAD_E_HIP, std::string("dense_download: ") + hipGetErrorString(e));
    return RAD_OK;
}

int rad_dense_release(rad_ctx *c)
{
    if (!c)
        return RAD_E_ARG;
    (void)hipStreamSynchronize(c->st);
    dfree(c->fhi);
    dfree(c->flo);
    c->fhi = c->flo = nullptr;
    c->npad = 0;
    c->dense = false;
    return RAD_OK;
}

int rad_dense_apply(rad_ctx *c, const float *x, uint32_t R, int split, float *y)
{
    if (!c || !x || !y)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_dense(c)))
        return rc;
    if (R == 0 || R > RAD_GATHER_MAX_R)
        return fail(c, RAD_E_ARG, "dense_apply: R must be 1..192");
    GatherBufs g;
    if ((rc = gather_alloc(c, g, R)))
        return rc;
    size_t n = (size_t)R * c->npatch;
    HIPCHK(c, hipMemcpyAsync(g.out, x, n * 4, hipMemcpyHostToDevice, c->st));
    launch_pad_columns(R, c->npad, c->npatch, g.out, nullptr, g.xhi, split ? g.xlo : nullptr, c->st);
    const GemmSplit ks = split ? g.ks2 : g.ks1;
    launch_gemm(g.rt, split != 0, c->fhi, c->flo, g.xhi, g.xlo, c->npad, ks, g.yp, c->st);
    launch_sum_partials(R, c->npad, c->npatch, ks, g.rp, g.yp, g.out, c->st);
    if ((rc = check_launch(c, "dense_apply")))
        return rc;
    HIPCHK(c, hipMemcpyAsync(y, g.out, n * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return RAD_OK;
}

int rad_gather_solve(rad_ctx *c, const float *rhs, uint32_t R, const rad_gather_opts *opts,
                     float *out, rad_gather_stats *stats)
{
    if (!c || !out)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_dense(c)))
        return rc;
    if (R == 0 || R > RAD_GATHER_MAX_R || (!rhs && R != 3))
        return fail(c, RAD_E_ARG, "gather: R must be 1..192 (3 without rhs)");
    uint32_t max_outer = opts && opts->max_outer ? opts->max_outer : 50;
    uint32_t inner = opts && opts->inner ? opts->inner : 8;
    double target = opts ? opts->target : 1e-3;
    auto t0 = std::chrono::steady_clock::now();
    GatherBufs g;
    if ((rc = gather_alloc(c, g, R)))
        return rc;
    const uint32_t N = c->npatch, P = c->npad;
    size_t n = (size_t)R * N;
    if (rhs) {
        HIPCHK(c, hipMemcpyAsync(g.out, rhs, n * 4, hipMemcpyHostToDevice, c->st));
    } else {
        std::vector<float> em((size_t)N * 3), col(n);
        HIPCHK(c, hipMemcpyAsync(em.data(), c->emis, em.size() * 4, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        for (uint32_t i = 0; i < N; ++i)
            for (int ch = 0; ch < 3; ++ch)
                col[(size_t)ch * N + i] = em[(size_t)i * 3 + ch];
        HIPCHK(c, hipMemcpy(g.out, col.data(), n * 4, hipMemcpyHostToDevice));
    }
    launch_pad_columns(R, P, N, g.out, g.E, nullptr, nullptr, c->st);
    // B = 0, r = E; denominator sum_i A_i sum_r |E_ir|
    launch_gather_update(0, g.yp, g.ks1, g.rp, R, P, N, g.E, c->refl, c->area, g.rv, g.B, g.xhi,
                         g.xlo, g.parts, g.nblk, c->st);
    if ((rc = check_launch(c, "gather init")))
        return rc;
    double den = 0.0, num = 0.0;
    if ((rc = sum_parts(c, g, &den)))
        return rc;
    double res = den > 0 ? 1.0 : 0.0;
    uint32_t outer = 0, sweeps = 0;
    rad_gather_stats st{};
    // events around the first sweep product and the residual product of
    // each step (device time of the two product kernels)
    hipEvent_t ev[4] = {};
    bool have_ev = true;
    for (hipEvent_t &e : ev)
        have_ev = have_ev && hipEventCreate(&e) == hipSuccess;
    double t_sweep = 0.0, t_res = 0.0;
    while (res > target && outer < max_outer) {
        // correction d = r + diag(rho) M_bf16 d, `inner` sweeps from d = r,
        // the last one adds d to B and splits B into its two bf16 halves
        for (uint32_t s = 0; s < inner; ++s) {
            if (s == 0 && have_ev)
                (void)hipEventRecord(ev[0], c->st);
            launch_gemm(g.rt, false, c->fhi, nullptr, g.xhi, nullptr, P, g.ks1, g.yp, c->st);
            if (s == 0 && have_ev)
                (void)hipEventRecord(ev[1], c->st);
            launch_gather_update(s + 1 == inner ? 2 : 1, g.yp, g.ks1, g.rp, R, P, N, g.E, c->refl,
                                 c->area, g.rv, g.B, g.xhi, g.xlo, g.parts, g.nblk, c->st);
        }
        sweeps += inner;
        // residual r = E + diag(rho) M B - B from both halves, FP64 sums
        if (have_ev)
            (void)hipEventRecord(ev[2], c->st);
        launch_gemm(g.rt, true, c->fhi, c->flo, g.xhi, g.xlo, P, g.ks2, g.yp, c->st);
        if (have_ev)
            (void)hipEventRecord(ev[3], c->st);
        launch_gather_update(3, g.yp, g.ks2, g.rp, R, P, N, g.E, c->refl, c->area, g.rv, g.B, g.xhi,
                             g.xlo, g.parts, g.nblk, c->st);
        if ((rc = check_launch(c, "gather step")) || (rc = sum_parts(c, g, &num)))
            break;
        if (have_ev) {
            float a = 0.f, b = 0.f;
            (void)hipEventElapsedTime(&a, ev[0], ev[1]);
            (void)hipEventElapsedTime(&b, ev[2], ev[3]);
            t_sweep += a;
            t_res += b;
        }
        res = num / den;
        if (outer < 16)
            st.history[outer] = res;
        ++outer;
    }
    for (hipEvent_t e : ev)
        if (e)
            (void)hipEventDestroy(e);
    if (rc)
        return rc;
    st.sweep_ms = outer ? t_sweep / outer : 0.0;
    st.resid_ms = outer ? t_res / outer : 0.0;
    launch_gather_out(R, P, N, g.B, g.rv, g.out, c->st);
    if ((rc = check_launch(c, "gather out")))
        return rc;
    HIPCHK(c, hipMemcpyAsync(out, g.out, n * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (stats) {
        st.outer = outer;
        st.sweeps = sweeps;
        st.residual = res;
        st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *stats = st;
    }
    return RAD_OK;
}

int rad_gather_check_fp64(rad_ctx *c, const float *B, const float *E, uint32_t R, double *out2)
{
    if (!c || !B || !out2)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_plain(c)))
        return rc;
    if (R == 0 || R > 16 || (!E && R != 3))
        return fail(c, RAD_E_ARG, "gather_check: R must be 1..16 (3 without E)");
    const uint32_t N = c->npatch;
    const size_t n = (size_t)R * N;
    std::vector<float> eh;
    if (!E) {
        std::vector<float> em((size_t)N * 3);
        HIPCHK(c, hipStreamSynchronize(c->st));
        HIPCHK(c, hipMemcpy(em.data(), c->emis, em.size() * 4, hipMemcpyDeviceToHost));
        eh.resize(n);
        for (uint32_t i = 0; i < N; ++i)
            for (uint32_t ch = 0; ch < 3; ++ch)
                eh[(size_t)ch * N + i] = em[(size_t)i * 3 + ch];
        E = eh.data();
    }
    const uint32_t nblk = 1024;
    float *dB = nullptr, *dE = nullptr;
    double *acc = nullptr, *parts = nullptr;
    unsigned long long *q = nullptr;
    auto cleanup = [&]() {
        dfree(dB);
        dfree(dE);
        dfree(acc);
        dfree(parts);
        dfree(q);
    };
    if ((rc = dalloc(c, &dB, n)) || (rc = dalloc(c, &dE, n)) || (rc = dalloc(c, &acc, n)) ||
        (rc = dalloc(c, &parts, 2 * nblk)) || (rc = dalloc(c, &q, 1))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpy(dB, B, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(dE, E, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemset(acc, 0, n * 8);
    if (e == hipSuccess)
        e = hipMemset(q, 0, 8);
    if (e != hipSuccess) {
        cleanup();
        return fail(c, RAD_E_HIP, std::string("gather_check: ") + hipGetErrorString(e));
    }
    bool timing = c->timing;
    c->timing = false;
    // every patch's hemicube once, K per pass, accumulated as it is rendered
    for (uint32_t base = 0; base < N && rc == RAD_OK; base += c->K) {
        uint32_t k = N - base < c->K ? N - base : c->K;
        launch_iota_shooters(base, k, c->shooters, c->nshoot, c->st);
        if ((rc = render_pass(c, k, 0, nullptr, nullptr, nullptr)))
            break;
        launch_check_accum(c->ff, k, base, c->rowlen, N, R, dB, c->area, acc, c->st);
        launch_check_rowsum(c->ff, k, c->rowlen, N, c->refl, q, c->st);
        rc = check_launch(c, "gather_check pass");
    }
    c->timing = timing;
    c->rendered = false;
    if (rc == RAD_OK) {
        launch_check_final(R, N, acc, dE, dB, c->refl, c->area, parts, nblk, c->st);
        rc = check_launch(c, "gather_check final");
    }
    std::vector<double> h(2 * nblk);
    unsigned long long qb = 0;
    if (rc == RAD_OK) {
        e = hipMemcpyAsync(h.data(), parts, h.size() * 8, hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(&qb, q, 8, hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->st);
        if (e != hipSuccess)
            rc = fail(c, RAD_E_HIP, std::string("gather_check: ") + hipGetErrorString(e));
    }
    cleanup();
    if (rc)
        return rc;
    if ((rc = check_overflow(c)))
        return rc;
    double num = 0.0, den = 0.0;
    for (uint32_t b = 0; b < nblk; ++b) {
        num += h[b];
        den += h[nblk + b];
    }
    out2[0] = den > 0 ? num / den : 0.0;
    memcpy(&out2[1], &qb, 8);
    return RAD_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Scene packs (rad_pack.hip kernels).
extern "C" {

int rad_pack_upload(rad_ctx *c, uint32_t nscene, const uint32_t *scene_tri_beg,
                    const uint32_t *scene_patch_beg, const uint32_t *scene_slots,
                    const float *tri_xyz, const uint32_t *tri_patch, uint32_t ntri, uint32_t npatch,
                    const float *frames, const float *refl, const float *area)
{
    if (!c)
        return RAD_E_ARG;
    if (nscene == 0 || !scene_tri_beg || !scene_patch_beg || !scene_slots)
        return fail(c, RAD_E_ARG, "pack: no scenes");
    int rc;
    if ((rc = check_scene(c, tri_xyz, tri_patch, ntri, npatch, frames, refl, area)))
        return rc;
    if (scene_tri_beg[0] != 0 || scene_tri_beg[nscene] != ntri || scene_patch_beg[0] != 0 ||
        scene_patch_beg[nscene] != npatch)
        return fail(c, RAD_E_ARG, "pack: offsets must start at 0 and end at ntri / npatch");
    uint32_t nblk = 1, rowlen = 1, kmax = 1;
    std::vector<uint32_t> patch_scene(npatch);
    for (uint32_t s = 0; s < nscene; ++s) {
        uint32_t t0 = scene_tri_beg[s], t1 = scene_tri_beg[s + 1];
        uint32_t p0 = scene_patch_beg[s], p1 = scene_patch_beg[s + 1];
        if (t1 < t0 || p1 <= p0)
            return fail(c, RAD_E_ARG, "pack: scene " + std::to_string(s) +
                                          " has decreasing offsets or no patch");
        for (uint32_t t = t0; t < t1; ++t)
            if (tri_patch[t] < p0 || tri_patch[t] >= p1)
                return fail(c, RAD_E_ARG, "pack: triangle " + std::to_string(t) +
                                              " names a patch outside its scene");
        if (scene_slots[s] == 0 || scene_slots[s] > c->K)
            return fail(c, RAD_E_ARG, "pack: shooters per scene must be 1..K");
        nblk = std::max(nblk, (t1 - t0 + 255) / 256);
        rowlen = std::max(rowlen, p1 - p0);
        kmax = std::max(kmax, scene_slots[s]);
        for (uint32_t p = p0; p < p1; ++p)
            patch_scene[p] = s;
    }
    if ((rc = upload_impl(c, tri_xyz, tri_patch, ntri, npatch, frames, refl, area, nblk, rowlen)))
        return rc;
    PackBuffers &p = c->pk;
    size_t S = nscene;
    if ((rc = dalloc(c, &p.sc_tbeg, S + 1)) || (rc = dalloc(c, &p.sc_pbeg, S + 1)) ||
        (rc = dalloc(c, &p.sc_slots, S)) || (rc = dalloc(c, &p.sc_slot_beg, S)) ||
        (rc = dalloc(c, &p.sc_nsel, S)) || (rc = dalloc(c, &p.sc_done, S)) ||
        (rc = dalloc(c, &p.sc_den, S)) || (rc = dalloc(c, &p.sc_res, S)) ||
        (rc = dalloc(c, &p.sc_shot, S)) ||
        (rc = dalloc(c, &p.cand, S * kmax)) || (rc = dalloc(c, &p.patch_scene, npatch)) ||
        (rc = dalloc(c, &p.pass, 1)) || (rc = dalloc(c, &p.slot_tbeg, c->K)) ||
        (rc = dalloc(c, &p.slot_tcnt, c->K)) || (rc = dalloc(c, &p.slot_pbeg, c->K))) {
        // the F rows are sized for the largest scene, not for a plain scene
        pack_free(c);
        return drop_scene(c, rc);
    }
    hipError_t e = hipSuccess;
    auto cp = [&](void *d, const void *h, size_t n) {
        if (e == hipSuccess)
            e = hipMemcpy(d, h, n, hipMemcpyHostToDevice);
    };
    auto set = [&](void *d, int v, size_t n) {
        if (e == hipSuccess)
            e = hipMemset(d, v, n);
    };
    cp(p.sc_tbeg, scene_tri_beg, (S + 1) * 4);
    cp(p.sc_pbeg, scene_patch_beg, (S + 1) * 4);
    cp(p.sc_slots, scene_slots, S * 4);
    cp(p.patch_scene, patch_scene.data(), (size_t)npatch * 4);
    set(p.sc_nsel, 0, S * 4);
    set(p.sc_slot_beg, 0, S * 4);
    set(p.sc_done, 0xFF, S * 4);
    set(p.sc_den, 0, S * 8);
    set(p.sc_res, 0, S * 8);
    set(p.sc_shot, 0, S * 4);
    set(p.pass, 0, 4);
    set(p.slot_tbeg, 0, c->K * 4);
    set(p.slot_tcnt, 0, c->K * 4);
    set(p.slot_pbeg, 0, c->K * 4);
    if (e != hipSuccess) {
        pack_free(c);
        return drop_scene(c, fail(c, RAD_E_HIP, std::string("pack upload: ") +
                                                    hipGetErrorString(e)));
    }
    p.nscene = nscene;
    p.kmax = kmax;
    p.K = c->K;
    p.shooters = c->shooters;
    p.nshoot = c->nshoot;
    return RAD_OK;
}

int rad_pack_hemicubes(rad_ctx *c, const uint32_t *scene_w)
{
    if (!c || !scene_w)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    PackBuffers &p = c->pk;
    if (!p.nscene)
        return fail(c, RAD_E_STATE, "no scene pack (rad_pack_upload)");
    const uint32_t S = p.nscene;
    std::vector<uint32_t> ws;
    for (uint32_t s = 0; s < S; ++s) {
        const uint32_t w = scene_w[s];
        if (w < 2 || (w & 1) || w > c->g.W)
            return fail(c, RAD_E_ARG, "pack_hemicubes: scene " + std::to_string(s) +
                                          ": hemicube side must be even, 2 .. the context's");
        ws.push_back(w);
    }
    std::sort(ws.begin(), ws.end());
    ws.erase(std::unique(ws.begin(), ws.end()), ws.end());
    // one ΔFF table per distinct side, concatenated
    std::vector<uint32_t> off(ws.size());
    size_t total = 0;
    for (size_t i = 0; i < ws.size(); ++i) {
        off[i] = (uint32_t)total;
        total += (size_t)3 * ws[i] * ws[i];
    }
    std::vector<uint32_t> dffo(S), slot_fill(c->K, c->g.W);
    for (uint32_t s = 0; s < S; ++s)
        dffo[s] = off[std::lower_bound(ws.begin(), ws.end(), scene_w[s]) - ws.begin()];
    if ((rc = dalloc(c, &p.dffs, total)) || (rc = dalloc(c, &p.sc_w, S)) ||
        (rc = dalloc(c, &p.sc_dffo, S)) || (rc = dalloc(c, &p.slot_w, c->K)) ||
        (rc = dalloc(c, &p.slot_dffo, c->K))) {
        dfree(p.dffs);
        dfree(p.sc_w);
        dfree(p.sc_dffo);
        dfree(p.slot_w);
        dfree(p.slot_dffo);
        p.dffs = nullptr;
        p.sc_w = p.sc_dffo = p.slot_w = p.slot_dffo = nullptr;
        return rc;
    }
    for (size_t i = 0; i < ws.size(); ++i) {
        Geometry gw{ws[i], 2 * ws[i], ws[i] + ws[i] / 2, 3 * ws[i] * ws[i]};
        launch_dff(gw, p.dffs + off[i], c->st);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(p.sc_w, scene_w, S * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(p.sc_dffo, dffo.data(), S * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(p.slot_w, slot_fill.data(), c->K * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemsetAsync(p.slot_dffo, 0, c->K * 4, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    drop_graph(c); // captured passes must read the per-slot sides
    return RAD_OK;
}

int rad_pack_begin(rad_ctx *c, double target)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    PackBuffers &p = c->pk;
    if (!p.nscene)
        return fail(c, RAD_E_STATE, "no scene pack (rad_pack_upload)");
    size_t n = c->npatch, S = p.nscene;
    std::vector<float> em(n * 3), ar(n);
    std::vector<uint32_t> pb(S + 1);
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(em.data(), c->emis, n * 12, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(ar.data(), c->area, n * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(pb.data(), p.sc_pbeg, (S + 1) * 4, hipMemcpyDeviceToHost));
    std::vector<double> den(S, 0.0);
    for (size_t s = 0; s < S; ++s)
        for (uint32_t i = pb[s]; i < pb[s + 1]; ++i)
            den[s] += (double)ar[i] *
                      ((double)em[3 * (size_t)i] + (double)em[3 * (size_t)i + 1] + (double)em[3 * (size_t)i + 2]);
    HIPCHK(c, hipMemcpy(p.sc_den, den.data(), S * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(p.sc_done, 0xFF, S * 4));
    HIPCHK(c, hipMemset(p.sc_shot, 0, S * 4));
    HIPCHK(c, hipMemset(p.pass, 0, 4));
    p.target = target;
    drop_graph(c); // the target is a kernel argument of captured passes
    return RAD_OK;
}

int rad_pack_status(rad_ctx *c, uint32_t *done_pass, double *residual, uint32_t *shooters,
                    uint32_t *passes, uint32_t *remaining)
{
    if (!c)
        return RAD_E_ARG;
    PackBuffers &p = c->pk;
    if (!p.nscene)
        return fail(c, RAD_E_STATE, "no scene pack (rad_pack_upload)");
    HIPCHK(c, hipStreamSynchronize(c->st));
    int rc;
    if ((rc = check_overflow(c)))
        return rc;
    std::vector<uint32_t> done(p.nscene);
    HIPCHK(c, hipMemcpy(done.data(), p.sc_done, p.nscene * 4, hipMemcpyDeviceToHost));
    if (done_pass)
        memcpy(done_pass, done.data(), p.nscene * 4);
    if (residual)
        HIPCHK(c, hipMemcpy(residual, p.sc_res, p.nscene * 8, hipMemcpyDeviceToHost));
    if (shooters)
        HIPCHK(c, hipMemcpy(shooters, p.sc_shot, p.nscene * 4, hipMemcpyDeviceToHost));
    if (passes)
        HIPCHK(c, hipMemcpy(passes, p.pass, 4, hipMemcpyDeviceToHost));
    if (remaining) {
        uint32_t r = 0;
        for (uint32_t d : done)
            r += d == 0xFFFFFFFFu;
        *remaining = r;
    }
    return RAD_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Checkpoints: the solver state in a portable file (the successor of the
// reference's .rr dump of raw Patch structs, Main.cpp:1460-1647).
//   "RADCKPT1" | u32 version 1 | u32 npatch | u32 ntri | u32 mode |
//   u64 scene hash | f32 unshot[3][npatch] | f32 shot[3][npatch] |
//   f32 emission[3][npatch] | u64 FNV-1a of everything before it
// Little-endian, channels as separate planes.
extern "C" {

int rad_checkpoint_save(rad_ctx *c, const char *path)
{
    if (!c || !path)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    if ((rc = ensure_hash(c)))
        return rc;
    const size_t n = c->npatch;
    std::vector<float> u(n * 3), sh(n * 3), em(n * 3), buf(n * 9);
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(u.data(), c->unshot, n * 12, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(sh.data(), c->shot, n * 12, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(em.data(), c->emis, n * 12, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i)
        for (int ch = 0; ch < 3; ++ch) {
            buf[(0 + ch) * n + i] = u[i * 3 + ch];
            buf[(3 + ch) * n + i] = sh[i * 3 + ch];
            buf[(6 + ch) * n + i] = em[i * 3 + ch];
        }
    std::vector<unsigned char> out;
    auto put = [&](const void *p, size_t k) {
        const unsigned char *b = (const unsigned char *)p;
        out.insert(out.end(), b, b + k);
    };
    const uint32_t ver = 1, np = c->npatch, nt = c->ntri, mode = c->cfg.mode;
    put("RADCKPT1", 8);
    put(&ver, 4);
    put(&np, 4);
    put(&nt, 4);
    put(&mode, 4);
    put(&c->geom_hash, 8);
    put(buf.data(), buf.size() * 4);
    uint64_t sum = fnv1a(0xCBF29CE484222325ull, out.data(), out.size());
    put(&sum, 8);
    FILE *f = fopen(path, "wb");
    if (!f)
        return fail(c, RAD_E_ARG, std::string("checkpoint: cannot open ") + path);
    size_t w = fwrite(out.data(), 1, out.size(), f);
    int cl = fclose(f);
    if (w != out.size() || cl != 0)
        return fail(c, RAD_E_ARG, std::string("checkpoint: write failed for ") + path);
    return RAD_OK;
}

int rad_checkpoint_load(rad_ctx *c, const char *path)
{
    if (!c || !path)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    FILE *f = fopen(path, "rb");
    if (!f)
        return fail(c, RAD_E_ARG, std::string("checkpoint: cannot open ") + path);
    std::vector<unsigned char> in;
    unsigned char tmp[65536];
    size_t r;
    while ((r = fread(tmp, 1, sizeof(tmp), f)) > 0)
        in.insert(in.end(), tmp, tmp + r);
    fclose(f);
    const size_t n = c->npatch, head = 8 + 4 * 4 + 8, body = n * 9 * 4;
    if (in.size() < head + 8 || memcmp(in.data(), "RADCKPT1", 8) != 0)
        return fail(c, RAD_E_ARG, "checkpoint: not a checkpoint file");
    uint32_t ver, np, nt, mode;
    uint64_t hash, sum;
    memcpy(&ver, &in[8], 4);
    memcpy(&np, &in[12], 4);
    memcpy(&nt, &in[16], 4);
    memcpy(&mode, &in[20], 4);
    memcpy(&hash, &in[24], 8);
    if (ver != 1)
        return fail(c, RAD_E_ARG, "checkpoint: unknown version " + std::to_string(ver));
    if ((rc = ensure_hash(c)))
        return rc;
    if (np != c->npatch || nt != c->ntri || hash != c->geom_hash)
        return fail(c, RAD_E_ARG, "checkpoint: written for another scene (patch count, triangle "
                                  "count or geometry hash differ)");
    if (in.size() != head + body + 8)
        return fail(c, RAD_E_ARG, "checkpoint: truncated or oversized file");
    memcpy(&sum, &in[head + body], 8);
    if (sum != fnv1a(0xCBF29CE484222325ull, in.data(), head + body))
        return fail(c, RAD_E_ARG, "checkpoint: checksum mismatch (corrupted file)");
    if (mode != c->cfg.mode)
        return fail(c, RAD_E_ARG, "checkpoint: written in another shooting mode");
    std::vector<float> buf(n * 9), u(n * 3), sh(n * 3), em(n * 3);
    memcpy(buf.data(), &in[head], body);
    for (size_t i = 0; i < n; ++i)
        for (int ch = 0; ch < 3; ++ch) {
            u[i * 3 + ch] = buf[(0 + ch) * n + i];
            sh[i * 3 + ch] = buf[(3 + ch) * n + i];
            em[i * 3 + ch] = buf[(6 + ch) * n + i];
        }
    return rad_set_state(c, u.data(), sh.data(), em.data());
}

// ------------------------------------------------------------ scene ingest

int rad_scene_ingest(rad_ctx *c, const float *quads, uint32_t nface, const float *color,
                     const float *emission, const float *shot0, double max_area, int mode,
                     uint32_t *out_ntri, uint32_t *out_npatch, uint32_t *out_nquad)
{
    if (!c)
        return RAD_E_ARG;
    if (!quads || nface == 0 || !(max_area >= 0.0) ||
        (mode != RAD_FACES_QUADS && mode != RAD_FACES_TRIANGLES))
        return fail(c, RAD_E_ARG, "ingest: null faces, zero faces, negative area or bad mode");
    for (uint64_t i = 0; i < (uint64_t)nface * 12; ++i)
        if (!std::isfinite(quads[i]) || std::fabs(quads[i]) > 1.0e4f)
            return fail(c, RAD_E_ARG, "ingest: face coordinate not finite or beyond 1e4");
    std::vector<radpatch::FacePlan> plan;
    uint32_t nq, nt, np;
    if (!radpatch::plan_faces(quads, nface, max_area, mode, plan, nq, nt, np))
        return fail(c, RAD_E_LIMIT, "ingest: more than 2^31 triangles");
    if (np == 0) // Patch::divide yields no children for a quad with a zero-length edge
        return fail(c, RAD_E_ARG, "ingest: the faces yield no patches");
    int rc;
    if ((rc = scene_alloc(c, nt, np, (nt + 255) / 256, np)))
        return drop_scene(c, rc);
    if ((rc = dalloc(c, &c->patch_quad, np)) || (rc = dalloc(c, &c->quad_pbeg, (size_t)nq + 1)) ||
        (rc = dalloc(c, &c->quad_nbr, (size_t)nq * 8)) || (rc = dalloc(c, &c->qval, (size_t)nq * 3)))
        return drop_scene(c, rc);
    // inputs of the kernel: faces, plans, per-face colour / radiosity / illumination
    float *dfaces = nullptr, *dcol = nullptr, *dem = nullptr, *dsh = nullptr;
    FacePlanDev *dplan = nullptr;
    std::vector<FacePlanDev> hp(nface);
    for (uint32_t f = 0; f < nface; ++f)
        hp[f] = {plan[f].native, plan[f].kx, plan[f].ky, plan[f].qbeg, plan[f].tbeg, plan[f].pbeg};
    auto cleanup = [&]() {
        dfree(dfaces);
        dfree(dcol);
        dfree(dem);
        dfree(dsh);
        dfree(dplan);
    };
    if ((rc = dalloc(c, &dfaces, (size_t)nface * 12)) || (rc = dalloc(c, &dplan, nface)) ||
        (color && (rc = dalloc(c, &dcol, (size_t)nface * 3))) ||
        (emission && (rc = dalloc(c, &dem, (size_t)nface * 3))) ||
        (shot0 && (rc = dalloc(c, &dsh, (size_t)nface * 3)))) {
        cleanup();
        return drop_scene(c, rc);
    }
    hipError_t e = hipMemcpyAsync(dfaces, quads, (size_t)nface * 48, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dplan, hp.data(), (size_t)nface * sizeof(FacePlanDev),
                           hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess && color)
        e = hipMemcpyAsync(dcol, color, (size_t)nface * 12, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess && emission)
        e = hipMemcpyAsync(dem, emission, (size_t)nface * 12, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess && shot0)
        e = hipMemcpyAsync(dsh, shot0, (size_t)nface * 12, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->quad_pbeg + nq, &np, 4, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) {
        launch_ingest(nface, nq, dfaces, dplan, dcol, dem, dsh, (uint32_t)mode, nt, c->tri,
                      c->tri_patch, c->frames, c->refl, c->area, c->emis, c->unshot, c->shot,
                      c->patch_quad, c->quad_pbeg, c->quad_nbr, c->st);
        launch_tri_sphere(c->tri, nt, c->tsph, c->st);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->st);
    cleanup();
    if (e != hipSuccess)
        return drop_scene(c, fail(c, RAD_E_HIP, std::string("ingest: ") + hipGetErrorString(e)));
    // a patch of zero area would break the area-normalised shooting
    std::vector<float> ar(np);
    e = hipMemcpy(ar.data(), c->area, (size_t)np * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        return drop_scene(c, fail(c, RAD_E_HIP, std::string("ingest: ") + hipGetErrorString(e)));
    for (uint32_t p = 0; p < np; ++p)
        if (!(ar[p] > 0.0f))
            return drop_scene(c, fail(c, RAD_E_ARG, "ingest: patch " + std::to_string(p) +
                                                        " has zero area (degenerate face)"));
    c->nquad = nq;
    if (out_ntri)
        *out_ntri = nt;
    if (out_npatch)
        *out_npatch = np;
    if (out_nquad)
        *out_nquad = nq;
    return RAD_OK;
}

int rad_scene_download(rad_ctx *c, float *tri_xyz, uint32_t *tri_patch, float *frames,
                       float *refl, float *area)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    const size_t nt = c->ntri, np = c->npatch;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (tri_xyz) {
        std::vector<float> planes(nt * 9);
        HIPCHK(c, hipMemcpy(planes.data(), c->tri, nt * 36, hipMemcpyDeviceToHost));
        for (size_t t = 0; t < nt; ++t)
            for (int k = 0; k < 9; ++k)
                tri_xyz[t * 9 + k] = planes[(size_t)k * nt + t];
    }
    if (tri_patch)
        HIPCHK(c, hipMemcpy(tri_patch, c->tri_patch, nt * 4, hipMemcpyDeviceToHost));
    if (frames)
        HIPCHK(c, hipMemcpy(frames, c->frames, np * 36, hipMemcpyDeviceToHost));
    if (refl)
        HIPCHK(c, hipMemcpy(refl, c->refl, np * 12, hipMemcpyDeviceToHost));
    if (area)
        HIPCHK(c, hipMemcpy(area, c->area, np * 4, hipMemcpyDeviceToHost));
    return RAD_OK;
}

int rad_set_quads(rad_ctx *c, const uint32_t *patch_quad, uint32_t nquad, const uint32_t *quad_nbr)
{
    if (!c)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_scene(c)))
        return rc;
    if (!patch_quad || !quad_nbr || nquad == 0 || nquad > c->npatch)
        return fail(c, RAD_E_ARG, "quads: null arrays or quad count out of range");
    const uint32_t np = c->npatch;
    std::vector<uint32_t> pbeg(nquad + 1, 0);
    if (patch_quad[0] != 0 || patch_quad[np - 1] != nquad - 1)
        return fail(c, RAD_E_ARG, "quads: patch_quad must run from 0 to nquad - 1");
    for (uint32_t p = 1; p < np; ++p) {
        uint32_t d = patch_quad[p] - patch_quad[p - 1];
        if (patch_quad[p] < patch_quad[p - 1] || d > 1)
            return fail(c, RAD_E_ARG, "quads: patch_quad must step by 0 or 1");
        if (d == 1)
            pbeg[patch_quad[p]] = p;
    }
    pbeg[nquad] = np;
    for (uint64_t i = 0; i < (uint64_t)nquad * 8; ++i)
        if (quad_nbr[i] >= nquad)
            return fail(c, RAD_E_ARG, "quads: neighbour index out of range");
    c->nquad = 0; // until the new arrays are complete
    if ((rc = dalloc(c, &c->patch_quad, np)) || (rc = dalloc(c, &c->quad_pbeg, (size_t)nquad + 1)) ||
        (rc = dalloc(c, &c->quad_nbr, (size_t)nquad * 8)) || (rc = dalloc(c, &c->qval, (size_t)nquad * 3)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->patch_quad, patch_quad, (size_t)np * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->quad_pbeg, pbeg.data(), ((size_t)nquad + 1) * 4,
                             hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->quad_nbr, quad_nbr, (size_t)nquad * 32, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->nquad = nquad;
    return RAD_OK;
}

int rad_get_quads(rad_ctx *c, uint32_t *nquad, uint32_t *patch_quad, uint32_t *quad_nbr)
{
    if (!c || !nquad)
        return RAD_E_ARG;
    *nquad = c->nquad;
    if (c->nquad == 0)
        return fail(c, RAD_E_STATE, "no quad topology (rad_scene_ingest or rad_set_quads)");
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (patch_quad)
        HIPCHK(c, hipMemcpy(patch_quad, c->patch_quad, (size_t)c->npatch * 4, hipMemcpyDeviceToHost));
    if (quad_nbr)
        HIPCHK(c, hipMemcpy(quad_nbr, c->quad_nbr, (size_t)c->nquad * 32, hipMemcpyDeviceToHost));
    return RAD_OK;
}

int rad_vertex_colors(rad_ctx *c, float *out)
{
    if (!c || !out)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_plain(c)))
        return rc;
    if (c->nquad == 0)
        return fail(c, RAD_E_STATE, "no quad topology (rad_scene_ingest or rad_set_quads)");
    float *dout = nullptr;
    if ((rc = dalloc(c, &dout, (size_t)c->nquad * 12)))
        return rc;
    launch_vertex_colors(c->nquad, c->quad_pbeg, c->quad_nbr, c->refl, c->area, c->unshot,
                         c->shot, c->qval, dout, c->st);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, (size_t)c->nquad * 48, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->st);
    dfree(dout);
    if (e != hipSuccess)
        return fail(c, RAD_E_HIP, std::string("vertex colours: ") + hipGetErrorString(e));
    return RAD_OK;
}

int rad_render_view(rad_ctx *c, const float *frame, uint32_t *out_ids)
{
    if (!c || !frame || !out_ids)
        return RAD_E_ARG;
    int rc;
    if ((rc = need_plain(c)))
        return rc;
    for (int k = 0; k < 9; ++k)
        if (!std::isfinite(frame[k]))
            return fail(c, RAD_E_ARG, "render_view: frame not finite");
    if ((!c->viewfr && (rc = dalloc(c, &c->viewfr, 9))) ||
        (!c->view_sh && (rc = dalloc(c, &c->view_sh, 1))))
        return rc;
    const uint32_t zero = 0, one = 1;
    HIPCHK(c, hipMemcpyAsync(c->viewfr, frame, 36, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->view_sh, &zero, 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->nshoot, &one, 4, hipMemcpyHostToDevice, c->st));
    if ((rc = render_pass(c, 1, 0, c->item_ids, nullptr, nullptr, c->view_sh, c->viewfr)))
        return rc;
    c->rendered = false; // the item buffer does not hold a shooter's hemicube
    HIPCHK(c, hipMemcpyAsync(out_ids, c->item_ids, (size_t)c->g.RES * 4, hipMemcpyDeviceToHost,
                             c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return check_overflow(c);
}

} // extern "C"
