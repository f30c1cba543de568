// rad_internal.h - shared definitions of libradhip (device records, context).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/radhip.h"

namespace rad {

// Sub-pixel precision of snapped window coordinates (8 bits, as GL
// implementations commonly use) and the guard band of the clipper in NDC.
constexpr int kSub = 256;
constexpr float kGuard = 4096.0f;
// Depth 1.0 (the glClear value, Main.cpp:1148) with id 0 (black background).
constexpr uint64_t kClearKey = (uint64_t)0x3F800000u << 32;
// Big triangles are split into 8x8 pixel tiles, one wavefront per tile.
constexpr int kTile = 8;

// Setup record of a triangle too big for one lane (48 bytes).
struct RasterRec {
    int32_t x[3], y[3];   // snapped face-local window coordinates (1/256 px)
    float dzdx, dzdy, z0; // depth plane anchored at vertex 0
    uint32_t id;          // patch + 1
    int32_t base;         // texel of face-local pixel (0,0) in the key buffer
    int16_t bx0, by0, bx1, by1; // inclusive pixel bbox, clamped to the scissor
};
static_assert(sizeof(RasterRec) == 52, "RasterRec layout");

struct TileItem {
    uint32_t rec;
    uint16_t tx, ty;
};

// Counters used by one pass (reset by the clear kernel).
enum { kCntRec = 0, kCntItem = 1, kCntOverflow = 2, kCntDone = 3, kCntUnknown = 4, kCntN = 8 };

struct Geometry {
    uint32_t W, TW, TH, RES;
};

// Launchers implemented in rad_kernels.hip; every launch goes on `st`.
void launch_dff(const Geometry &g, float *dff, hipStream_t st);
void launch_select_top1(uint32_t mode_rule, const float *unshot, const float *area,
                        const float *emis, uint32_t npatch, uint32_t *shooters,
                        uint32_t *nshoot, double *partials, uint64_t *pkeys,
                        uint32_t *counter, double *resid, int nblocks, hipStream_t st);
void launch_select_topk(uint32_t rule, const float *unshot, const float *area,
                        const float *emis, uint32_t npatch, uint32_t k, uint32_t *shooters,
                        uint32_t *nshoot, uint64_t *cand, double *partials,
                        uint32_t *counter, double *resid, hipStream_t st);
size_t topk_cand_half(uint32_t npatch, uint32_t k);
void launch_select_reference(const float *unshot, uint32_t npatch, uint32_t k,
                             uint32_t *shooters, uint32_t *nshoot, hipStream_t st);
void launch_prepare(const Geometry &g, uint32_t K, const uint32_t *shooters,
                    const uint32_t *nshoot, const float *frames, const float *unshot,
                    float *snap, float *mvp, hipStream_t st);
void launch_clear(const Geometry &g, uint32_t K, uint64_t *keys, float *ff,
                  uint32_t *hits, uint32_t npatch, uint32_t *counters, hipStream_t st);
void launch_raster_setup(const Geometry &g, uint32_t K, const float *tri,
                         const uint32_t *tri_patch, uint32_t ntri, const float *mvp,
                         const uint32_t *nshoot, uint64_t *keys, RasterRec *recs,
                         uint32_t rec_cap, TileItem *items, uint32_t item_cap,
                         uint32_t *counters, uint32_t small_px, hipStream_t st);
void launch_raster_tiles(const RasterRec *recs, const TileItem *items,
                         const uint32_t *counters, uint32_t rec_cap, uint32_t item_cap,
                         int32_t TW, uint64_t *keys, int nblocks, hipStream_t st);
void launch_reduce(const Geometry &g, uint32_t K, const uint32_t *nshoot,
                   const uint64_t *keys, const float *dff, uint32_t npatch, float *ff,
                   uint32_t *hits, uint32_t *counters, hipStream_t st);
void launch_shoot(uint32_t mode, uint32_t K, const uint32_t *shooters,
                  const uint32_t *nshoot, const float *snap, const float *ff,
                  const float *refl, const float *area, uint32_t npatch, float *unshot,
                  float *shot, hipStream_t st);
void launch_ids_to_keys(const uint32_t *ids, uint64_t *keys, size_t n, hipStream_t st);
void launch_split_keys(const uint64_t *keys, uint32_t *ids, float *depth, size_t n,
                       hipStream_t st);
void launch_residual(const float *unshot, const float *area, const float *emis,
                     uint32_t npatch, double *partials, uint32_t *counter, double *resid,
                     int nblocks, hipStream_t st);

} // namespace rad
