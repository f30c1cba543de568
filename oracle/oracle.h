/*
 * oracle.h - CPU restatement of david-sabata/Radiosity's hemicube shooting
 * path.  TEST INFRASTRUCTURE ONLY: the oracle is the checker that tests,
 * the smoke check and the CPU-baseline leg of the benchmark compare the HIP
 * library against.  Product code never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Conventions:
 *   W        hemicube side (Config::hemicubeSide), even
 *   TEX_W    2W, TEX_H 1.5W, TEX_RES = TEX_W*TEX_H  (Config.cpp:37-39)
 *   item     uint32 per texel: 0 = background, patch index + 1 otherwise
 *   frame    9 floats per patch: eye(3), normal(3), up(3)
 *            (Patch.cpp:253-276 getCenter/getNormal/getUp)
 *   tri      9 floats per triangle: v0, v1, v2
 */
#ifndef RADIOSITY_ORACLE_H
#define RADIOSITY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Config.cpp:30-48 */
void orc_geometry(uint32_t W, uint32_t *tex_w, uint32_t *tex_h, uint32_t *tex_res);

/* FormFactors.cpp:23-67: top[W*W], side[W*W/2] */
void orc_dff_top_side(uint32_t W, float *top, float *side);
/* FormFactors.cpp:87-108 (__SUM_TEST) - float sum in the reference's loop order */
float orc_dff_sum_test(uint32_t W);
/* FormFactors.cpp:280-314: one copy of the cross-layout table, TEX_RES floats */
void orc_dff_cross(uint32_t W, float *out);

/* Patch.cpp:253-276 for a quad (v1..v4) */
void orc_quad_frame(const float *v4x3, float *frame9);

/* Camera.cpp:19-52, 97-103; Transform.cpp:26-46, 70-80, 127-156;
 * Vector.h:534-537; Vector.cpp:445-458, 478-...; Main.cpp:1171-1183.
 * face: 0 UP, 1 DOWN, 2 LEFT, 3 RIGHT, 4 FRONT (Main.h:210-211).
 * mvp is column-major m[col*4+row], as Matrix4f::f[col][row]. */
void orc_face_mvp(const float *frame9, int face, float *mvp16);

/* Main.cpp:316-389: viewport (x,y,w,h) and scissor (x,y,w,h) of a face,
 * hemicube slot 0. */
void orc_face_viewport(uint32_t W, int face, int32_t *vp4, int32_t *sc4);

/* Item-buffer render of one hemicube (Main.cpp:1145-1200, DrawPatchLook
 * Main.cpp:689-725, patch shader Shaders.cpp:234-260) under the GL 3.0
 * rules restated in oracle.c: near-plane clipping, guard-band clipping,
 * 8-bit sub-pixel snapping, CCW front faces with back-face culling,
 * pixel-centre sampling with a top-left fill rule, GL_LESS on a float
 * depth buffer cleared to 1.0, triangles drawn in index order.
 * item[TEX_RES] gets patch+1 or 0; depth_or_null[TEX_RES] gets the depth.
 * Returns the number of fragments that passed the scissor and coverage
 * tests (depth-tested or not). */
uint64_t orc_render_hemicube(uint32_t W, const float *tri, const uint32_t *tri_patch,
                             uint32_t ntri, const float *frame9,
                             uint32_t *item, float *depth_or_null);
#define ORC_VAR_DEPTH24 1
#define ORC_VAR_FMA 2
#define ORC_VAR_DIVIDE 4
uint64_t orc_render_hemicube_mt(uint32_t W, const float *tri, const uint32_t *tri_patch,
                                uint32_t ntri, const float *frame, uint32_t *item,
                                int nthreads);
uint64_t orc_render_hemicube_variant(uint32_t W, const float *tri, const uint32_t *tri_patch,
                                     uint32_t ntri, const float *frame, uint32_t *item,
                                     int flags);

/* Kernel_ProcessHemicube.h:23-69 run-length form + host accumulation
 * Main.cpp:1251-1269 for K stacked hemicubes (rows TEX_H*K).
 * workitems_x = Config::oclWorkitemsX (default 4).  ff[K*npatch] (+=).
 * Returns the number of triplets the kernel would have emitted; ids >= npatch
 * are skipped and counted in *unknown (Main.cpp:1259-1262). */
uint64_t orc_process_hemicube_runs(uint32_t W, uint32_t K, uint32_t workitems_x,
                                   const uint32_t *item, const float *dff_cross,
                                   uint32_t npatch, float *ff, uint64_t *unknown);

/* The same kernel's output itself: per-run triplets (hemicube, patch id,
 * float sum of ΔFF along the run) for every run of a nonzero id, each
 * work-item (row y, span gx) reserving 4 slots at a time and padding its
 * last block with (0, 0, 0) triplets (Kernel_ProcessHemicube.h:52-69).
 * Written in work-item order (row, then span), one of the orders the
 * reference's atomics may produce; ids are item - 1 whatever their range.
 * Returns the write index (a multiple of 4); entries at or beyond cap are
 * not written. */
uint64_t orc_process_hemicube_triplets(uint32_t W, uint32_t K, uint32_t workitems_x,
                                       const uint32_t *item, const float *dff_cross,
                                       uint64_t cap, uint32_t *hemis, uint32_t *ids,
                                       float *energies);

/* Direct per-pixel histogram of the same item buffer: hits[K*npatch] and
 * ff64[K*npatch] (FP64 twin).  Both are overwritten. */
void orc_hemicube_histogram(uint32_t W, uint32_t K, const uint32_t *item,
                            const float *dff_cross, uint32_t npatch,
                            uint32_t *hits, double *ff64);

/* ModelContainer.cpp:259-299 selection rule over |unshot|^2.
 * Writes count ids (unfilled = 0) and returns how many were filled. */
uint32_t orc_select_reference(const float *unshot, uint32_t npatch, uint32_t count,
                              uint32_t *ids);
/* Argmax of A_i * (r+g+b) (PHYSICAL) - ties to the lowest index. */
uint32_t orc_select_argmax_power(const float *unshot, const float *area, uint32_t npatch);

/* Shooting update for k shooters whose form-factor rows ff[k*npatch] are
 * given (Main.cpp:1251-1295).
 * mode 0 (REFERENCE): unshot_j += snap_i * F_ij * 0.3 * colour_i (Main.cpp:1274,
 *   Patch.h:13), refl holds the patch colours.
 * mode 1 (PHYSICAL): unshot_j += refl_j * snap_i * F_ij * A_i / A_j.
 * Then shot_i += snap_i, unshot_i -= snap_i for each shooter (Main.cpp:1287-1295).
 * Arithmetic in double on float state (the FP64 twin). */
void orc_shoot(uint32_t mode, uint32_t npatch, const uint32_t *shooters, uint32_t k,
               const double *ff, const float *refl, const float *area,
               double *unshot, double *shot);

/* sum_i A_i (r+g+b)(unshot_i) / sum_i A_i (r+g+b)(emission_i), FP64 */
double orc_residual(const double *unshot, const float *emission, const float *area,
                    uint32_t npatch);

/* One full plain pass on the CPU (the reference CPU path restated):
 * select (argmax power or reference rule, k=1), render, process, shoot.
 * State is float (as the reference's Patch fields).  Returns the shooter. */
uint32_t orc_plain_pass(uint32_t W, uint32_t mode, uint32_t select_rule,
                        const float *tri, const uint32_t *tri_patch, uint32_t ntri,
                        const float *frames, const float *refl, const float *area,
                        uint32_t npatch, const float *dff_cross,
                        float *unshot, float *shot,
                        uint32_t *item_scratch, float *ff_scratch);
uint32_t orc_plain_pass_mt(uint32_t W, uint32_t mode, uint32_t select_rule,
                           const float *tri, const uint32_t *tri_patch, uint32_t ntri,
                           const float *frames, const float *refl, const float *area,
                           uint32_t npatch, const float *dff, float *unshot, float *shot,
                           uint32_t *item, float *ff, int nthreads);

/* POWER top-K: patches with A (r+g+b) > 0, by that key descending, equal
 * keys by index ascending; count ids (unfilled 0), returns the number filled. */
uint32_t orc_select_topk_power(const float *unshot, const float *area, uint32_t npatch,
                               uint32_t count, uint32_t *ids);

/* One OnIdle pass with K hemicubes (Main.cpp:1137-1303) on float state:
 * select, snapshot, per shooter render + run-form ProcessHemicube + receiver
 * update, then the emitter updates.  last_energy3 = lastEnergy, the last
 * shooter's radiosity before its subtraction (Main.cpp:1288-1292).
 * item: TEX_RES scratch, ff: npatch scratch, ids: K (out).  Returns the
 * number of shooters. */
uint32_t orc_pass_k(uint32_t W, uint32_t mode, uint32_t select_rule, uint32_t K,
                    const float *tri, const uint32_t *tri_patch, uint32_t ntri,
                    const float *frames, const float *refl, const float *area, uint32_t npatch,
                    const float *dff, float *unshot, float *shot, uint32_t *item, float *ff,
                    uint32_t *ids, float *last_energy3);

/* FP64 Jacobi for the gather B = E + diag(rho) M B (no reference
 * counterpart: the fixed point the shooting loop of Main.cpp:1251-1295
 * converges to, with M the shooting factors).  M: n x n row-major float;
 * column r of E and B (R x n, column r contiguous) uses reflectance channel
 * r % 3.  Iterates until sum A|r| / sum A|E| <= tol or max_it; returns the
 * iterations taken.  B is the last iterate. */
uint32_t orc_gather_jacobi(uint32_t n, uint32_t R, const float *M, const float *refl,
                           const float *area, const double *E, double tol, uint32_t max_it,
                           double *B);

#ifdef __cplusplus
}
#endif
#endif
