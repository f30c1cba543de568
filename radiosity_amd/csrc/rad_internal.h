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
// Largest screen tile of the binned rasterizer (texels per side), the most
// tiles per hemicube slab (LDS histograms), and the records each setup
// workgroup (256 triangles of one shooter) may keep in its own region before
// spilling to the shared spill region.
constexpr int kMaxTile = 32;
constexpr uint32_t kMaxTilesPerSlab = 4096;
constexpr uint32_t kRecsPerBlock = 768;
// Floats per shooter in the MVP buffer: 5 face MVPs + the culling frame.
constexpr int kShooterStride = 96;

// Setup record of one snapped (sub-)triangle of one face of one shooter:
// three 16-byte words (one vector load each).  Its texel bbox lives in the
// parallel bbox array the binning and the tile classification read.
struct alignas(16) RasterRec {
    int32_t x[3], y[3];   // snapped face-local window coordinates (1/256 px)
    float dzdx, dzdy, z0; // depth plane anchored at vertex 0
    uint32_t id;          // patch + 1
    uint32_t sf;          // shooter slot * 8 + face
    uint32_t pad;
};
static_assert(sizeof(RasterRec) == 48, "RasterRec layout");

// Counters used by one pass (reset by the clear kernel).
enum {
    kCntRec = 0,      // records emitted
    kCntOverflow = 2, // a record did not fit
    kCntEntries = 3,  // bin entries (scan total)
    kCntUnknown = 4,  // item-buffer ids beyond npatch
    kCntSpill = 5,    // records in the spill region
    kCntClip = 6,     // triangle-face jobs queued for clipping
    kCntSticky = 7,   // some earlier pass overflowed (cleared by the host only)
    kCntN = 8
};
// A pass may render as two halves on two streams (slots [0, k/2) and
// [k/2, k)), each with its own block of kCntN counters, spill region, clip
// queue and list area; the sticky flag lives in the first block.
constexpr int kCntBlocks = 2;
constexpr uint32_t kSplitMin = 16; // smallest pass rendered as two halves

// Control block of a context (u32 words on the device), written by the
// shoot kernels and read by the pass gate in k_prepare_clear:
enum {
    kCtlStop = 0,      // the stop test fired: later passes are no-ops (nshoot = 0)
    kCtlThresh = 2,    // [2..3] stop threshold on |lastEnergy| (double; 0: off)
    kCtlLastE = 4,     // [4..6] unshot RGB of the pass's last shooter before its
                       // emitter update (lastEnergy, Main.cpp:1290-1292), float bits
    kCtlShotLo = 8,    // [8..9] shooters shot (u64, low word first)
    kCtlPassLo = 10,   // [10..11] passes that shot at least one shooter (u64)
    kCtlDenseBase = 12, // first row of the current pass of a captured dense build
    kCtlN = 16
};

// Bins of the sort-middle rasterizer: one per (shooter slot, tile).
struct BinBuffers {
    RasterRec *recs;       // nblk*K regions of kRecsPerBlock, then kCntBlocks spill regions of spill_cap
    uint64_t *bbox;        // texel bbox per record slot, read by the bin fill
    uint32_t *blk_nrec;    // records per (shooter, setup block)
    uint32_t *count, *start, *fill;
    uint32_t *list;        // kCntBlocks list areas of ent_cap entries
    uint32_t *order;       // bins by descending size class (tile launch order)
    uint32_t nblk;         // setup blocks per shooter and face group (256 triangles each)
    uint32_t nz;           // face groups: 5 (one face per block, small K) or 1
    uint32_t spill_base, spill_cap, ent_cap;
    uint2 *clipq;          // kCntBlocks queues of clip_cap clipping jobs
    uint32_t clip_cap;
    int clip_blocks;
    uint32_t ntx, nty, ntiles, tshift;
    uint32_t setup_group;  // shooter slots one setup workgroup walks
};

struct Geometry {
    uint32_t W, TW, TH, RES;
};

// How one source face splits (planned on the host, rad_patch.h FacePlan):
// native triangle grid of side kx, or a kx x ky quad grid; where its child
// quads, triangles and patches start.
struct FacePlanDev {
    uint32_t native, kx, ky, qbeg, tbeg, pbeg;
};

// Many scenes packed into one context (rad_pack_upload): per-scene ranges
// and per-pass slot arrays (null when no pack is loaded).
struct PackBuffers {
    uint32_t nscene = 0, kmax = 0, K = 0;
    double target = 1e-3;
    uint32_t *sc_tbeg = nullptr, *sc_pbeg = nullptr; // nscene + 1 offsets
    uint32_t *sc_slots = nullptr;                    // shooters per pass per scene
    uint32_t *sc_slot_beg = nullptr, *sc_nsel = nullptr;
    uint32_t *sc_done = nullptr;                     // pass a scene finished in (~0u: not yet)
    double *sc_den = nullptr, *sc_res = nullptr;
    uint32_t *sc_shot = nullptr;                     // shooters shot per scene
    uint32_t *cand = nullptr;                        // nscene x kmax
    uint32_t *patch_scene = nullptr;                 // npatch
    uint32_t *pass = nullptr;                        // passes selected so far
    uint32_t *slot_tbeg = nullptr, *slot_tcnt = nullptr, *slot_pbeg = nullptr; // K
    // scenes at their own hemicube resolution (rad_pack_hemicubes; null: all
    // at the context's): per scene W and offset of its ΔFF table in dffs,
    // the same per slot of the pass
    uint32_t *sc_w = nullptr, *sc_dffo = nullptr, *slot_w = nullptr, *slot_dffo = nullptr;
    float *dffs = nullptr;
    uint32_t *shooters = nullptr, *nshoot = nullptr; // the context's own
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
// Prepare + clear of a pass: ff_rows rows of rowlen entries of F (and hits)
// are zeroed when ff (hits) is non-null; ctl (may be null) gates a stopped
// context.
void launch_prepare_clear(uint32_t K, const uint32_t *shooters, uint32_t *nshoot,
                          const float *frames, const float *unshot, float *snap, float *mvp,
                          uint32_t *bin_count, uint32_t ntiles, float *ff, uint32_t *hits,
                          uint32_t ff_rows, uint32_t rowlen, uint32_t *counters, uint32_t ent_cap,
                          const uint32_t *ctl, hipStream_t st);
void launch_clear(uint32_t K, uint32_t *bin_count, uint32_t ntiles, float *ff, uint32_t *hits,
                  uint32_t ff_rows, uint32_t rowlen, uint32_t *counters, uint32_t ent_cap,
                  hipStream_t st);
// Binned render of the shooter slots [s0, s1) (of which those below nshoot
// are drawn) in three launches: setup (records + bin counts), bin (scan +
// fill), tile (LDS depth test; out_ids != null writes the item buffer, ff or
// ffq != null reduces into F rows and hits).  h selects the counter block,
// spill region, clip queue and list area (0, or 1 for the second half of a
// split pass).
void launch_raster_setup(const Geometry &g, uint32_t s0, uint32_t s1, uint32_t h,
                         const float *tri, const float4 *tsph, const uint32_t *tri_patch,
                         uint32_t ntri, const float *mvp, const uint32_t *nshoot,
                         const BinBuffers &bb, uint32_t *counters, const uint32_t *slot_tbeg,
                         const uint32_t *slot_tcnt, const uint32_t *slot_w, hipStream_t st);
void launch_tri_sphere(const float *tri, uint32_t ntri, float4 *sph, hipStream_t st);
void launch_bin(uint32_t s0, uint32_t s1, uint32_t h, const uint32_t *nshoot,
                const BinBuffers &bb, uint32_t *counters, hipStream_t st);
void launch_tile(const Geometry &g, uint32_t s0, uint32_t s1, uint32_t h, const uint32_t *nshoot,
                 const BinBuffers &bb, uint32_t small_px, uint32_t *out_ids, float *out_depth,
                 const float *dff, uint32_t rowlen, const uint32_t *slot_pbeg, float *ff,
                 unsigned long long *ffq, uint32_t *hits, uint32_t *counters,
                 const uint32_t *slot_w, const uint32_t *slot_dffo, uint32_t *fblk,
                 uint32_t fstride, hipStream_t st);
// Deterministic mode: fixed-point F entries (2^-52 units) into the float rows.
void launch_fold(unsigned long long *ffq, float *ff, size_t n, hipStream_t st);
// ProcessHemicube over an id plane (K x RES u32): ff (or ffq) and hits +=.
void launch_reduce(const Geometry &g, uint32_t K, const uint32_t *nshoot, const uint32_t *ids,
                   const float *dff, uint32_t npatch, float *ff, unsigned long long *ffq,
                   uint32_t *hits, uint32_t *counters, hipStream_t st);
// The run-form triplets of ProcessHemicube: slot counts per work-item and
// their exclusive scan (offs: nitems + 1 entries, offs[nitems] = total), then
// the triplets at those offsets (entries >= cap skipped).
void launch_runs_count(const Geometry &g, uint32_t k, uint32_t nx, const uint32_t *item,
                       uint32_t *slots, uint32_t *offs, hipStream_t st);
void launch_runs_write(const Geometry &g, uint32_t k, uint32_t nx, const uint32_t *item,
                       const float *dff, const uint32_t *offs, uint32_t cap, uint32_t *th,
                       uint32_t *tid, float *te, hipStream_t st);
// What k_next_pass selects into and prepares for the pass after one whose
// last shoot left per-workgroup lists.
struct NextPass {
    uint32_t K;
    uint32_t *shooters, *nshoot;
    const float *frames;
    float *snap, *mvp;
    uint32_t *bin_count;
    size_t nbins;
    uint32_t *counters;
    uint32_t ent_cap;
};
// Candidate pool of a chained POWER selection (device memory).  When armed,
// a pass's last shoot keeps every receiver whose key is at least tau:
// count keys, the first kPoolCap of them in keys (one global atomic per
// shoot workgroup places its keys).  The pool holds every key >= tau, so
// when it has at least K keys its K largest are the state's; the next pass
// then selects from it, else it scans every patch.
constexpr uint32_t kPoolCap = 16384; // most pool keys a selection takes
constexpr uint32_t kPoolRearm = 8;  // passes of a graph between armings
constexpr uint32_t kDenseGraphPasses = 8; // passes per captured dense-build graph
struct SelPool {
    uint64_t tau;
    uint32_t armed;             // the next last shoot pools (else leaves lists)
    uint32_t hits, falls, arms; // selections from the pool, full scans, armings
    uint32_t count;             // keys the last shoot pooled (reset by k_next_pass)
    uint64_t keys[kPoolCap];
};
// fblk (nullable): the F-block flags, [row / 64][slot] with K slots per
// block; null reads every F entry.  Adds the F rows of shooters
// [sbeg, send); the last launch of a pass also moves the shooters' energy,
// runs the stop test and counts the pass, and with lists != null leaves the
// next POWER selection's per-workgroup lists of the new state, or, given
// the pool and when it is armed, adds the keys at or above its bound to it.
void launch_shoot(uint32_t mode, uint32_t K, const uint32_t *shooters,
                  const uint32_t *nshoot, const float *snap, float *ff, uint32_t *fblk,
                  const float *refl, const float *area, uint32_t npatch, float *unshot,
                  float *shot, uint32_t *ctl, uint32_t sbeg, uint32_t send, bool last,
                  uint64_t *lists, SelPool *pool, hipStream_t st);
// The selection and preparation of a pass after one whose last shoot left
// lists or pooled (pool->armed, unless lists_only).  From lists, with
// target != 0, it arms the pool for the shoots that follow: tau is set so
// that at least `target` keys of the state are >= tau.  pool_cap (at most
// kPoolCap): the most pool keys taken before scanning every patch instead.
void launch_next_pass(const uint64_t *lists, uint32_t nlists, const NextPass &np,
                      const float *unshot, const float *area, uint32_t npatch,
                      const uint32_t *ctl, SelPool *pool, bool lists_only, uint32_t target,
                      uint32_t pool_cap, hipStream_t st);
// Per-workgroup selection lists a last shoot leaves (lists != null): their
// count for npatch receivers, when the shoot is given the pool or not.
uint32_t shoot_lists(uint32_t npatch, bool pooled = false);
void launch_residual(const float *unshot, const float *area, const float *emis,
                     uint32_t npatch, double *partials, uint32_t *counter, double *resid,
                     int nblocks, hipStream_t st);

// Launchers implemented in rad_pack.hip.
void launch_pack_select(const PackBuffers &pk, const float *unshot, const float *area,
                        hipStream_t st);
void launch_pack_shoot(const PackBuffers &pk, uint32_t mode, const float *snap, float *ff,
                       uint32_t rowlen, const float *refl, const float *area, uint32_t npatch,
                       float *unshot, float *shot, hipStream_t st);

// Launchers implemented in rad_gather.hip.
void launch_iota_shooters(uint32_t row0, uint32_t k, uint32_t *shooters, uint32_t *nshoot,
                          hipStream_t st);
void launch_dense_store(const float *ff, uint32_t k, uint32_t row0, uint32_t npatch, uint32_t npad,
                        uint16_t *fhi, uint16_t *flo, hipStream_t st);
void launch_dense_store_t(float *ff, uint32_t k, uint32_t row0, uint32_t npatch,
                          uint32_t npad, const float *area, uint16_t *fhi, uint16_t *flo,
                          hipStream_t st);
// A captured dense build: the next pass's rows (ctl[kCtlDenseBase] advanced
// by K, from 0 - K before the first; K rows or the rest of the patches) as
// its shooters, and the transposed store of the pass's rows taken from the
// device (ctl[kCtlDenseBase], nshoot).
void launch_iota_next(uint32_t K, uint32_t npatch, uint32_t *ctl, uint32_t *shooters,
                      uint32_t *nshoot, hipStream_t st);
void launch_dense_store_t_dev(float *ff, uint32_t K, const uint32_t *ctl, const uint32_t *nshoot,
                              uint32_t npatch, uint32_t npad, const float *area, uint16_t *fhi,
                              uint16_t *flo, hipStream_t st);
void launch_dense_from_f32(const float *src, uint32_t rows, uint32_t row0, uint32_t npatch,
                           uint32_t npad, uint16_t *fhi, uint16_t *flo, hipStream_t st);
void launch_dense_to_f32(const uint16_t *fhi, const uint16_t *flo, uint32_t rows, uint32_t row0,
                         uint32_t npatch, uint32_t npad, float *dst, hipStream_t st);
uint32_t gemm_tiles(uint32_t R);
// k slices of a gather product: rows below rsplit are summed over S slices
// of npad / S columns, the rows from rsplit on (the row tiles of the last,
// partial round of workgroups) over S2 slices of about npad / S2 columns,
// so that round also fills the CUs.  The partial products hold max(S, S2)
// planes.
struct GemmSplit {
    uint32_t S, S2, rsplit;
};
GemmSplit gemm_split(uint32_t rt, bool split, uint32_t npad, uint32_t S, uint32_t ncu);
void launch_gemm(uint32_t rt, bool split, const uint16_t *fhi, const uint16_t *flo,
                 const uint16_t *xhi, const uint16_t *xlo, uint32_t npad, GemmSplit ks,
                 float *yp, hipStream_t st);
void launch_gather_update(int kind, const float *yp, GemmSplit ks, uint32_t rp, uint32_t R,
                          uint32_t npad, uint32_t n, const float *E, const float *refl,
                          const float *area, float *rv, float *B, uint16_t *xhi, uint16_t *xlo,
                          double *parts, uint32_t nblocks, hipStream_t st);
// FP64 residual check of a gather solution against rendered FP32 rows
// (shooting layout), one pass of k rows at a time; R <= kCheckMaxR (16).
void launch_check_accum(const float *ff, uint32_t k, uint32_t row0, uint32_t rowlen, uint32_t n,
                        uint32_t R, const float *B, const float *area, double *acc,
                        hipStream_t st);
void launch_check_rowsum(const float *ff, uint32_t k, uint32_t rowlen, uint32_t n,
                         const float *refl, unsigned long long *qbits, hipStream_t st);
void launch_check_final(uint32_t R, uint32_t n, const double *acc, const float *E, const float *B,
                        const float *refl, const float *area, double *parts, uint32_t nblocks,
                        hipStream_t st);
void launch_gather_out(uint32_t R, uint32_t npad, uint32_t n, const float *B, const float *rv,
                       float *out, hipStream_t st);
void launch_pad_columns(uint32_t R, uint32_t npad, uint32_t n, const float *src, float *dst,
                        uint16_t *xhi, uint16_t *xlo, hipStream_t st);
void launch_sum_partials(uint32_t R, uint32_t npad, uint32_t n, GemmSplit ks, uint32_t rp,
                         const float *yp, float *out, hipStream_t st);

// Launchers implemented in rad_ingest.hip.
void launch_ingest(uint32_t nface, uint32_t nquad, const float *faces, const FacePlanDev *plan,
                   const float *color, const float *emission, const float *shot0, uint32_t mode,
                   uint32_t ntri, float *tri, uint32_t *tri_patch, float *frames, float *refl,
                   float *area, float *emis, float *unshot, float *shot, uint32_t *patch_quad,
                   uint32_t *quad_pbeg, uint32_t *quad_nbr, hipStream_t st);
void launch_vertex_colors(uint32_t nquad, const uint32_t *quad_pbeg, const uint32_t *quad_nbr,
                          const float *refl, const float *area, const float *unshot,
                          const float *shot, float *val, float *out, hipStream_t st);

} // namespace rad

struct rad_ctx;
namespace rad {
// rad_static.hip: the device address of the fixed scratch array of
// rad_process_runs_static on the context's (current) device, looked up once
// per device at context creation.
int static_scratch_init(rad_ctx *c);
} // namespace rad
