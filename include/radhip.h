/*
 * radhip.h - C ABI of the MI355X progressive-refinement radiosity library
 * (libradhip.so).  The entry points replace the hot path of
 * david-sabata/Radiosity: Config, precomputeHemicubeFormFactors, the
 * ModelContainer/Patch state, the GL item-buffer render, the OpenCL
 * ProcessHemicube kernel and the host accumulate/shoot loops.  Each
 * declaration cites the reference interface it replaces.
 *
 * Conventions
 *  - Every call returns RAD_OK (0) or a negative RAD_E* code; the message of
 *    the last failure is rad_last_error(ctx).  No C++ exception crosses the ABI.
 *  - Host pointers are caller-owned; arguments named *_or_null may be NULL to
 *    skip a copy.  Device memory is owned by the context.
 *  - One context = one HIP stream on the current device.  A context is not
 *    thread-safe (the reference drives GL/CL from its UI thread); separate
 *    contexts may be used from separate threads.
 *  - Patch p is drawn into the item buffer as p + 1; 0 is background
 *    (Colors.cpp:148, Kernel_ProcessHemicube.h:51).
 *  - An item buffer holds K hemicubes stacked vertically, each TEX_W x TEX_H
 *    texels in the cross layout of FormFactors.cpp:250-272, row-major with
 *    row 0 at the bottom (GL window convention), exactly as the reference's
 *    RGBA8 texture (Main.cpp:277).
 */
#ifndef RADHIP_H
#define RADHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RAD_OK 0
#define RAD_E_ARG -1      /* invalid argument or configuration */
#define RAD_E_STATE -2    /* call out of order (no scene, no selection ...) */
#define RAD_E_HIP -3      /* HIP runtime error (message names the call) */
#define RAD_E_NOMEM -4    /* device allocation failed */
#define RAD_E_LIMIT -5    /* size beyond a documented limit */

/* Shooting update (Main.cpp:1274 vs. physically based) */
#define RAD_MODE_REFERENCE 0 /* unshot_j += dB_i * F_ij * 0.3 * colour_i (Patch.h:13) */
#define RAD_MODE_PHYSICAL 1  /* unshot_j += rho_j * dB_i * F_ij * A_i / A_j */

/* Shooter selection */
#define RAD_SELECT_REFERENCE 0 /* ModelContainer.cpp:259-299 list rule on |dB|^2 */
#define RAD_SELECT_POWER 1     /* top-K by A_i*(r+g+b)(dB_i), ties to the lower index */

#define RAD_MAX_K 256
#define RAD_MAX_W 1024

/* Mirrors Config (Config.cpp:5-10 defaults, CLI Main.cpp:769-785) */
typedef struct rad_config {
    uint32_t hemicube;        /* Config::hemicubeSide: W, even, 2..RAD_MAX_W (default 16) */
    uint32_t hemicubes;       /* Config::hemicubesCount: K shooters per pass, 1..RAD_MAX_K (10) */
    uint32_t shoots;          /* Config::shootsPerCycle: passes per rad_run chunk (500) */
    uint32_t ocl_workitems_x; /* Config::oclWorkitemsX: spans per row for the run form (4) */
    uint32_t mode;            /* RAD_MODE_* */
    uint32_t select;          /* RAD_SELECT_* */
    uint32_t small_tri_px;    /* raster tuning: triangle pixels in a tile drawn by one lane (0: 64) */
    uint32_t tile_px;         /* raster tuning: screen tile side 8, 16 or 32 (0: auto) */
    uint32_t graph_passes;    /* rad_step replays this many passes as one captured
                                 HIP graph (0: plain launches) */
} rad_config;

typedef struct rad_ctx rad_ctx;

/* Config::freeze + InitCLObjects (Config.cpp:30-48, Main.cpp:405-611).
 * Binds the current HIP device, creates the stream and uploads the delta
 * form-factor table. */
int rad_create(const rad_config *cfg, rad_ctx **out);
/* CleanupGLObjects / cleanup of CL objects (Main.cpp:616-...) */
void rad_destroy(rad_ctx *ctx);
const char *rad_last_error(const rad_ctx *ctx);
/* Config::PATCHVIEW_TEX_W/H/RES getters (Config.cpp:115-153) */
int rad_geometry(const rad_ctx *ctx, uint32_t *tex_w, uint32_t *tex_h, uint32_t *tex_res);

/* float* precomputeHemicubeFormFactors() (FormFactors.h:20, FormFactors.cpp:280-339).
 * Writes ONE copy (TEX_RES floats) of the cross-layout table computed on the
 * device; the reference's K replicas (FormFactors.cpp:316-323) are not kept. */
int rad_precompute_dff(rad_ctx *ctx, float *out_or_null);

/* ModelContainer::updateData output (ModelContainer.cpp:81-155) and the
 * per-patch fields the hot path reads (Patch.h:39-55).
 *  tri_xyz     ntri*9 floats, triangle t = (v0,v1,v2); a reference quad
 *              patch is two triangles (0,1,2),(0,2,3) (ModelContainer.cpp:112-117)
 *  tri_patch   ntri patch indices; must be non-decreasing in t (draw order)
 *  patch_frame npatch*9 floats: eye, normal, up (Patch::getCenter/getNormal/getUp)
 *  refl_rgb    npatch*3: REFERENCE: Patch::getColor(); PHYSICAL: rho_j
 *  area        npatch patch areas (PHYSICAL area ratio and residual weights)
 * Resets the state to zero. */
int rad_scene_upload(rad_ctx *ctx, const float *tri_xyz, const uint32_t *tri_patch,
                     uint32_t ntri, uint32_t npatch, const float *patch_frame,
                     const float *refl_rgb, const float *area);

/* Patch::radiosity (unshot) and Patch::illumination (shot), npatch*3 each
 * (Patch.h:46-47).  emission_or_null sets the residual normalisation E
 * (NULL: E = unshot as given). */
int rad_set_state(rad_ctx *ctx, const float *unshot_rgb, const float *shot_rgb,
                  const float *emission_or_null);
int rad_get_state(rad_ctx *ctx, float *unshot_or_null, float *shot_or_null);

/* ModelContainer::getHighestRadiosityPatchesId (ModelContainer.cpp:259-299)
 * with the context's selection rule.  Writes k ids (unfilled slots 0, as the
 * reference) and the number filled.  The selection stays on the device for
 * the next rad_shoot_selected. */
int rad_select(rad_ctx *ctx, uint32_t k, uint32_t *out_ids, uint32_t *out_n);

/* Face loop + DrawPatchLook (Main.cpp:1145-1200, 689-725): renders the
 * hemicubes of the k given shooters into the device item buffer and
 * optionally copies ids (k*TEX_RES) and depths (k*TEX_RES) out. */
int rad_render_itembuf(rad_ctx *ctx, const uint32_t *shooters, uint32_t k,
                       uint32_t *out_ids_or_null, float *out_depth_or_null);

/* ProcessHemicube kernel + host accumulation (Kernel_ProcessHemicube.h:9-70,
 * Main.cpp:1207-1269) over the k hemicubes last rendered: per-patch hit
 * counts and form-factor rows, k*npatch each. */
int rad_process_hemicube(rad_ctx *ctx, uint32_t k, uint32_t *out_hits_or_null,
                         float *out_ff_or_null);

/* Same reduction over a caller-supplied item buffer (k*TEX_RES uint32). */
int rad_process_itembuf(rad_ctx *ctx, const uint32_t *itembuf, uint32_t k,
                        uint32_t *out_hits_or_null, float *out_ff_or_null);

/* The ProcessHemicube kernel's own output (Kernel_ProcessHemicube.h:9-69,
 * launched at Main.cpp:1207-1249) over the k hemicubes last rendered (or
 * given to rad_process_itembuf): per-run triplets (hemicube, patch id,
 * ΔFF sum of the run) for every run of a nonzero id along the spans of
 * TEX_W / ocl_workitems_x texels (Config.cpp:43-45), each work-item reserving
 * 4 slots at a time and padding its last block with (0, 0, 0), exactly as
 * the reference kernel writes p_hemicubes / p_ids / p_energies.  The stream
 * is in work-item order (row, then span), one of the orders the reference's
 * atomics can produce.  *out_count gets the write index (a multiple of 4,
 * the reference's n_last_index, Main.cpp:1231-1233) even when it exceeds cap
 * (then RAD_E_LIMIT and nothing is copied).  Ids are item - 1 whatever their
 * range; the reference's host loop skips those >= npatch (Main.cpp:1259). */
int rad_process_runs(rad_ctx *ctx, uint32_t k, uint32_t cap, uint32_t *out_hemicubes,
                     uint32_t *out_ids, float *out_energies, uint32_t *out_count);

/* No counterpart in the reference: the rad_process_runs triplet stream for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops).  All pointers are device memory: the k item buffers (k*TEX_RES
 * uint32, as rad_process_itembuf takes them), the three triplet arrays of cap
 * entries, and one uint32 that receives the write index (entries at or past
 * cap are not written; the index still counts them).  Its work arrays live in
 * a fixed 4 MiB array per GPU, so k * TEX_H * min(ocl_workitems_x, TEX_W)
 * may be at most 524287 (RAD_E_LIMIT beyond).  The call queues its kernels on
 * the context's stream and returns without waiting or allocating; calls that
 * overlap on one GPU share that array and race. */
int rad_process_runs_static(rad_ctx *ctx, const uint32_t *d_itembuf, uint32_t k, uint32_t cap,
                            uint32_t *d_hemicubes, uint32_t *d_ids, float *d_energies,
                            uint32_t *d_count);

/* One pass of the OnIdle shoot loop (Main.cpp:1137-1295) for the k given
 * shooters: snapshot, render, reduce, shoot, emitter update. */
int rad_shoot(rad_ctx *ctx, const uint32_t *shooters, uint32_t k);

/* passes x (select K with the context rule, then rad_shoot), all on the
 * device without host synchronisation. */
int rad_step(rad_ctx *ctx, uint32_t passes);

/* No counterpart in the reference: captures and uploads the graph of
 * graph_passes passes that rad_step replays, and runs no pass, so that the
 * first rad_step after it pays no capture.  Nothing to do when graph_passes
 * is 0 or kernel timing is on.  A new scene, tile size or pack drops the
 * graph again. */
int rad_prepare(rad_ctx *ctx);

/* Runs passes until the FP64 residual <= residual_target or max_passes;
 * the residual is checked every check_every passes (the residual rule; the
 * reference's own per-pass stop test is rad_run_cycle). */
typedef struct rad_stats {
    uint64_t passes;         /* passes executed */
    uint64_t shooters;       /* shooters shot, counted on the device (<= passes*K) */
    double residual;         /* final FP64 residual */
    double seconds;          /* wall time of the loop */
    float last_energy;       /* lastEnergy.f_Length() of the last pass (Main.cpp:1288-1298):
                                |radiosity| of its last shooter before the emitter update */
    uint32_t reserved;
} rad_stats;
int rad_run(rad_ctx *ctx, uint64_t max_passes, double residual_target,
            uint32_t check_every, rad_stats *out);

/* One OnIdle cycle (Main.cpp:1137-1303): up to `passes` passes (select K
 * with the context rule, render, reduce, shoot) with the reference's stop
 * test after every pass: the cycle ends with the pass whose last shooter's
 * radiosity before its emitter update (lastEnergy, Main.cpp:1288-1292) has
 * f_Length() < stop_energy (0.1 in the reference, Main.cpp:1298; 0 turns
 * the test off).  The test runs on the device, so captured graph replays
 * stop too: passes after the stopping one select nobody.  The host
 * synchronises once per call. */
typedef struct rad_cycle_stats {
    uint64_t passes;         /* passes that shot, the stopping pass included */
    uint64_t shooters;       /* shooters shot */
    float last_energy[3];    /* lastEnergy of the last pass that shot */
    uint32_t stopped;        /* 1: the stop test fired in this cycle */
} rad_cycle_stats;
int rad_run_cycle(rad_ctx *ctx, uint32_t passes, double stop_energy, rad_cycle_stats *out);

/* Deterministic form-factor rows (no counterpart in the reference, whose
 * triplets are appended in the order the OpenCL work-items finish,
 * Kernel_ProcessHemicube.h:52-58, and summed in that order on the host,
 * Main.cpp:1257-1266): on != 0 accumulates every F entry in 64-bit fixed
 * point (units of 2^-52) with integer atomics, whose sum does not depend on
 * their order, and converts the rows to float once per pass; passes, item
 * buffer reductions and states are then identical bit for bit from run to
 * run.  Costs 8 bytes per F entry of device memory and one conversion pass
 * over the rows.  0 returns to float atomics (the default). */
int rad_set_deterministic(rad_ctx *ctx, int on);

/* Passes that shot and shooters shot since rad_create, as counted by the
 * shoot kernels on the device (synchronises).  Either pointer may be NULL. */
int rad_shot_count(rad_ctx *ctx, uint64_t *passes_or_null, uint64_t *shooters_or_null);

/* glFinish/clFinish (Main.cpp:1202, 1228): waits for all work queued on the
 * context's stream; RAD_E_LIMIT if a pass since the last check lost
 * triangles to a full rasterizer bin (capacities are set at upload). */
int rad_finish(rad_ctx *ctx);

/* sum_i A_i (r+g+b)(unshot_i) / sum_i A_i (r+g+b)(E_i), accumulated in FP64 */
int rad_residual_fp64(rad_ctx *ctx, double *out);

/* The five MVP matrices (column-major, 16 floats each, faces UP, DOWN, LEFT,
 * RIGHT, FRONT) the device built for shooter patch p (Camera.cpp:19-52,
 * Main.cpp:1171-1183). */
int rad_face_mvps(rad_ctx *ctx, uint32_t patch, float *out80);

/* Average device time (ms) per launch of each kernel of the last
 * rad_step/rad_run, measured with HIP events on the context stream:
 * out[0] select, [1] prepare+clear, [2] raster setup, [3] binning (scan +
 * fill), [4] tile raster with the fused ProcessHemicube, [5] shoot; also the
 * number of timed passes. */
int rad_kernel_times(rad_ctx *ctx, int enable, double *out6, uint64_t *npasses);

/* Device time (ms, HIP events) of the last standalone ProcessHemicube
 * launch (rad_process_hemicube / rad_process_itembuf) made while timing was
 * enabled with rad_kernel_times; 0 before one. */
int rad_reduce_kernel_ms(rad_ctx *ctx, double *ms);

/* Diagnostics (no counterpart in the reference): copies the per-tile record
 * counts of the last pass (hemicubes * tiles entries, row-major tiles of
 * tile_px) into out and returns the tile count per hemicube in *ntiles. */
int rad_debug_bins(rad_ctx *ctx, uint32_t *out_or_null, uint32_t *ntiles, uint32_t *tile_px);

/* Diagnostics (no counterpart in the reference): the rasterizer's pass
 * counters after the last pass (synchronises): out[0] and [1] unused,
 * [2] nonzero if a record or clipped polygon did not fit,
 * [3] bin entries, [4] item-buffer ids beyond npatch, [5] records in the
 * spill region, [6] triangle-faces queued for clipping, [7] nonzero if an
 * earlier pass overflowed.  Capacities: out_caps (optional, 3 entries) gets
 * the spill capacity, the bin-entry capacity and the clip-queue capacity.
 * A pass of 16 or more slots renders as two halves, each with these
 * capacities: [3], [5] and [6] are then the fuller half's, [4] the sum. */
int rad_debug_counters(rad_ctx *ctx, uint32_t *out8, uint32_t *out_caps_or_null);

/* Diagnostics (no counterpart in the reference): the candidate pool of the
 * POWER selection in captured passes.  Inside a graph, a pass's selection
 * may take the K largest keys from a pool of every receiver whose key is at
 * least a bound, which the previous pass's shoot collected; it does so only
 * when the pool holds K keys (then they are exact) and at most `cap`, else
 * it scans every patch.  cap != 0 sets that most (1 .. 16384, default 16384)
 * and drops the captured graph.  out4 (optional, synchronises) gets the
 * selections taken from the pool, the full scans, the times the bound was
 * set, and whether it is set now, counted since the context was created. */
int rad_debug_pool(rad_ctx *ctx, uint32_t cap, uint32_t *out4_or_null);

/* Autotuner of the rasterizer: times `passes` render passes of K shooters
 * spread evenly over the patches (render + reduce, no shooting, state
 * unchanged; the device selection is overwritten) for each screen tile
 * side {16, 32} (and 8 for K <= 2) and in-lane triangle limit
 * {16, 32, 48, 64, 128, 256}, and keeps the fastest pair.  No counterpart in the
 * reference (its GL driver rasterizes). */
#define RAD_TUNE_MAX 18
typedef struct rad_tune_result {
    uint32_t tile_px, small_tri_px; /* chosen */
    uint32_t ncand, reserved;
    double best_ms;                 /* device ms per render pass, chosen pair */
    uint32_t cand_tile[RAD_TUNE_MAX], cand_small[RAD_TUNE_MAX];
    double cand_ms[RAD_TUNE_MAX];   /* < 0: the candidate overflowed a bin */
} rad_tune_result;
int rad_autotune(rad_ctx *ctx, uint32_t passes, rad_tune_result *out_or_null);

/* Checkpoint / resume of the solver state: the successor of the reference's
 * .rr save and load (Main.cpp:1460-1647, LoadingModel.cpp:4-26), which dumped
 * raw Patch structs.  The file is portable: magic "RADCKPT1", version,
 * patch and triangle counts, shooting mode, a 64-bit FNV-1a hash of the
 * uploaded scene arrays, the unshot / shot / emission planes (f32, one plane
 * per channel) and an FNV-1a checksum.  Loading requires the same scene
 * (hash) and mode and replaces the state as rad_set_state does; a mismatch,
 * truncation or corruption is RAD_E_ARG with the reason in rad_last_error. */
int rad_checkpoint_save(rad_ctx *ctx, const char *path);
int rad_checkpoint_load(rad_ctx *ctx, const char *path);

/* ---------------------------------------------------------------------------
 * Scene packs: many independent scenes solved by one context, every kernel
 * of a pass launched once for all of them.  No counterpart in the reference
 * (one scene per process there); each scene follows the context's shooting
 * mode with the POWER rule inside the scene.
 *
 * The scenes' arrays are concatenated: scene s owns triangles
 * [scene_tri_beg[s], scene_tri_beg[s+1]) and patches [scene_patch_beg[s],
 * scene_patch_beg[s+1]) (nscene + 1 offsets each, first 0, last ntri /
 * npatch); tri_patch holds packed patch indices and each triangle's patch
 * lies in its own scene.  scene_slots[s] (1..K) is the number of shooters
 * the scene asks for per pass while it is active; when the active scenes ask
 * for more than K in all, the K slots are shared out by water-filling (each
 * scene min(quota, L) for the largest level L that fits, the rest one each in
 * an order that rotates every pass).  The pack
 * replaces the context's scene; rad_set_state / rad_get_state / rad_step /
 * rad_finish / rad_residual_fp64 work on the packed arrays, and the calls
 * that take explicit shooters, rad_select, rad_autotune and the dense gather
 * return RAD_E_STATE while a pack is loaded. */
int rad_pack_upload(rad_ctx *ctx, uint32_t nscene, const uint32_t *scene_tri_beg,
                    const uint32_t *scene_patch_beg, const uint32_t *scene_slots,
                    const float *tri_xyz, const uint32_t *tri_patch, uint32_t ntri, uint32_t npatch,
                    const float *patch_frame, const float *refl_rgb, const float *area);
/* Per-scene hemicube resolution of the loaded pack (scene_w[s]: even, 2 up
 * to the context's hemicube side, which sizes the screen tiles): each slot
 * renders and reduces its scene's hemicubes at the scene's own side with
 * that side's ΔFF table, all slots still in one launch per kernel.  Without
 * this call every scene uses the context's side. */
int rad_pack_hemicubes(rad_ctx *ctx, const uint32_t *scene_w);

/* Starts a solve: per-scene normalisation sum_A (r+g+b)(E) from the emission
 * of rad_set_state; a scene is done once its residual <= target. */
int rad_pack_begin(rad_ctx *ctx, double target);
/* Synchronises, then reports per scene the pass it finished in (0xFFFFFFFF:
 * not yet), its residual at the last selection and the shooters it has shot,
 * the passes run and the scenes still active.  Any pointer may be NULL. */
int rad_pack_status(rad_ctx *ctx, uint32_t *done_pass_or_null, double *residual_or_null,
                    uint32_t *shooters_or_null, uint32_t *passes_or_null,
                    uint32_t *remaining_or_null);

/* ---------------------------------------------------------------------------
 * Dense form factors and the Jacobi gather (the precision path).
 *
 * The reference shoots one hemicube row at a time (Main.cpp:1251-1279).  The
 * rows of all N patches form the dense matrix F (F_ij: form factor from
 * patch i to patch j, one hemicube of patch i per row), stored on the device
 * as BF16 plus the BF16 remainder F - bf16(F) (4 B per entry, N rounded up to
 * 512).  The gather solves
 *     B = E + diag(rho) M B          (column r uses the reflectance channel r % 3)
 * with M = F (RAD_DENSE_ROWS, the textbook gathering form) or
 * M_ji = F_ij A_i / A_j (RAD_DENSE_SHOOT: the factor PHYSICAL shooting applies,
 * so the fixed point is the one rad_run converges to).  Jacobi sweeps run on
 * v_mfma_f32_16x16x32_bf16 with FP32 accumulation; each outer step computes the
 * residual r = E + diag(rho) M B - B from both BF16 halves of M and B (about
 * 16 significant bits) in FP64 arithmetic, then solves the correction with
 * BF16 sweeps (iterative refinement).  Residual = sum_i A_i sum_r |r_ir| /
 * sum_i A_i sum_r |E_ir|, the measure rad_run uses for unshot energy.
 * No counterpart in the reference beyond the shooting step it replaces.
 * ------------------------------------------------------------------------- */
#define RAD_DENSE_ROWS 0u
#define RAD_DENSE_SHOOT 1u
#define RAD_GATHER_MAX_R 192u

/* Renders the hemicube of every patch (K per pass, the context's hemicube
 * resolution) and stores the matrix.  Needs 4 * npad^2 bytes of free device
 * memory (RAD_E_NOMEM otherwise).  Replaces rows of Main.cpp:1207-1249. */
int rad_dense_build(rad_ctx *ctx, uint32_t layout);
/* Stores a caller-given matrix M (npatch x npatch f32, row-major) instead. */
int rad_dense_upload(rad_ctx *ctx, const float *m);
/* Rows row0 .. row0+nrows-1 of the stored matrix (hi + lo) as f32. */
int rad_dense_download(rad_ctx *ctx, uint32_t row0, uint32_t nrows, float *out);
/* y = M x for R <= 192 columns (x, y: R x npatch, column r contiguous).
 * split = 0: BF16 M and x (the sweep product); 1: both halves (the residual
 * product). */
int rad_dense_apply(rad_ctx *ctx, const float *x, uint32_t R, int split, float *y);
/* Frees the matrix. */
int rad_dense_release(rad_ctx *ctx);

typedef struct rad_gather_opts {
    uint32_t max_outer;   /* refinement steps (0: 50) */
    uint32_t inner;       /* BF16 Jacobi sweeps per step (0: 8) */
    uint32_t reserved[2];
    double target;        /* stop when the residual <= target (e.g. 1e-3) */
} rad_gather_opts;

typedef struct rad_gather_stats {
    uint32_t outer;       /* refinement steps taken */
    uint32_t sweeps;      /* BF16 sweeps taken */
    uint32_t reserved[2];
    double residual;      /* final residual (FP64 sums) */
    double seconds;       /* wall time of the solve */
    double history[16];   /* residual after each of the first 16 steps */
    double sweep_ms;      /* device time of one BF16 product (HIP events, mean) */
    double resid_ms;      /* device time of one two-half residual product (mean) */
} rad_gather_stats;

/* Solves for R <= 192 right-hand sides.  rhs: R x npatch (column r
 * contiguous), or NULL with R = 3 for the scene's emission.  out (R x npatch)
 * receives B + r, i.e. E + diag(rho) M B, the radiosity including the energy
 * not yet gathered (as shot + unshot after shooting).  Blocks until done. */
int rad_gather_solve(rad_ctx *ctx, const float *rhs, uint32_t R, const rad_gather_opts *opts,
                     float *out, rad_gather_stats *stats);

/* FP64 check of a gather solution (no counterpart in the reference).  B and
 * E: R <= 16 columns of npatch (column r contiguous, reflectance channel
 * r % 3; E NULL: the scene's emission, R = 3).  Every patch's hemicube is
 * rendered again (FP32 rows, K per pass, never stored) and streamed into the
 * FP64 residual r = E + diag(rho) M B - B of the shooting factors
 * M_ji = F_ij A_i / A_j (the matrix RAD_DENSE_SHOOT rounds to BF16):
 * out2[0] = sum_j A_j sum_r |r_rj| / sum_j A_j sum_r |B_rj|, out2[1] =
 * q = max over rows i and channels c of sum_j rho_jc F_ij.  In the
 * area-weighted norm x -> diag(rho) M x contracts by q, so when q < 1 the
 * relative error of B against the exact FP64 solution of the FP32 form
 * factors is at most out2[0] / (1 - q). */
int rad_gather_check_fp64(rad_ctx *ctx, const float *B, const float *E_or_null, uint32_t R,
                          double *out2);

/* ---------------------------------------------------------------------------
 * Scene ingest on the device and vertex-colour export
 * ------------------------------------------------------------------------ */

/* Builds the scene on the device from source faces (nface*12 floats: A, B,
 * C, D each; a triangle is A B C C): replaces Model::subdivide + Patch::divide
 * (Model.cpp:27-60, Patch.cpp:47-222) and ModelContainer::updateData
 * (ModelContainer.cpp:81-155) followed by rad_scene_upload.  The host only
 * plans each face (whether and how it splits); one kernel writes every
 * child's triangles, frame, area, colour, radiosity and illumination
 * (inherited from its face, Patch.cpp:119-122; null arrays are 0) and its
 * neighbour links into the context.  mode is RAD_FACES_QUADS or
 * RAD_FACES_TRIANGLES (radhip_scene.h).  The arrays equal rad_scene_faces'
 * bit for bit; the state is unshot = emission, shot = shot0.  Counts are
 * returned through the optional out pointers.  A face that yields a patch of
 * zero area is RAD_E_ARG (no scene is kept). */
int rad_scene_ingest(rad_ctx *ctx, const float *quads, uint32_t nface, const float *color,
                     const float *emission, const float *shot0, double max_area, int mode,
                     uint32_t *out_ntri, uint32_t *out_npatch, uint32_t *out_nquad);

/* Copies the scene on the device back in rad_scene_upload's layout (any
 * pointer may be null): tri_xyz ntri*9, tri_patch ntri, frames npatch*9,
 * refl npatch*3, area npatch. */
int rad_scene_download(rad_ctx *ctx, float *tri_xyz, uint32_t *tri_patch, float *frames,
                       float *refl, float *area);

/* Quad topology of an uploaded scene (rad_scene_ingest sets it itself):
 * patch_quad (npatch, starting at 0 and stepping by 0 or 1, so the patches of
 * a quad are consecutive) and quad_nbr (nquad*8 neighbour quads in the slot
 * order of Patch::neighbours, as rad_scene.quad_nbr). */
int rad_set_quads(rad_ctx *ctx, const uint32_t *patch_quad, uint32_t nquad,
                  const uint32_t *quad_nbr);
int rad_get_quads(rad_ctx *ctx, uint32_t *nquad, uint32_t *patch_quad, uint32_t *quad_nbr);

/* Smoothed vertex colours of every quad (Colors::smoothShadePatch,
 * Colors.cpp:195-262, the per-frame loop of Main.cpp:1319-1340): out
 * (nquad*12) holds corners A, B, C, D (left-bottom, right-bottom, right-top,
 * left-top), each the mean of colour * (illumination + radiosity) over the
 * quad and its three neighbours at that corner.  A quad of several patches
 * (RAD_FACES_TRIANGLES) uses their area-weighted mean. */
int rad_vertex_colors(rad_ctx *ctx, float *out);

/* Item buffer of a viewpoint that is not a patch: frame = eye, view
 * direction, up (9 floats); out_ids (TEX_RES) as rad_render_itembuf for one
 * shooter.  The FRONT face is the 90-degree perspective image along the
 * view direction. */
int rad_render_view(rad_ctx *ctx, const float *frame, uint32_t *out_ids);

#ifdef __cplusplus
}
#endif
#endif
