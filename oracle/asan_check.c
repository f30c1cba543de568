/*
 * asan_check.c - drives every function of oracle.c on small synthetic
 * inputs for the sanitizer build (make -C oracle asan): a closed box room
 * with boxes inside and a soup of random triangles around the eye (near-plane
 * and guard-band clipping).  TEST INFRASTRUCTURE ONLY.  Prints one JSON line.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint64_t rs = 88172645463325252ull;
static float urand(float a, float b)
{
    rs ^= rs << 13;
    rs ^= rs >> 7;
    rs ^= rs << 17;
    return a + (b - a) * (float)((rs >> 11) * (1.0 / 9007199254740992.0));
}

int main(void)
{
    const uint32_t W = 16, K = 3;
    uint32_t TW, TH, RES;
    orc_geometry(W, &TW, &TH, &RES);
    const uint32_t ntri = 600, np = 300;
    float *tri = malloc(sizeof(float) * 9 * ntri);
    uint32_t *tp = malloc(sizeof(uint32_t) * ntri);
    float *fr = malloc(sizeof(float) * 9 * np), *refl = malloc(sizeof(float) * 3 * np);
    float *area = malloc(sizeof(float) * np);
    for (uint32_t t = 0; t < ntri; ++t) {
        float cx = urand(-2, 2), cy = urand(-2, 2), cz = urand(-2, 2);
        float s = t % 17 == 0 ? 3.0f : 0.2f;
        for (int v = 0; v < 3; ++v) {
            tri[9 * t + 3 * v + 0] = cx + urand(-s, s);
            tri[9 * t + 3 * v + 1] = cy + urand(-s, s);
            tri[9 * t + 3 * v + 2] = cz + urand(-s, s);
        }
        tp[t] = t / 2;
    }
    for (uint32_t p = 0; p < np; ++p) {
        float q[12];
        memcpy(q, tri + 18 * p, 9 * sizeof(float));
        memcpy(q + 9, tri + 18 * p + 15, 3 * sizeof(float));
        orc_quad_frame(q, fr + 9 * p);
        for (int c = 0; c < 3; ++c)
            refl[3 * p + c] = urand(0.1f, 0.8f);
        area[p] = urand(0.01f, 0.1f);
    }
    float *dff = malloc(sizeof(float) * RES), *top = malloc(sizeof(float) * W * W),
          *side = malloc(sizeof(float) * W * W / 2);
    orc_dff_top_side(W, top, side);
    orc_dff_cross(W, dff);
    float sumt = orc_dff_sum_test(W);
    float mvp[16];
    int32_t vp[4], sc[4];
    for (int f = 0; f < 5; ++f) {
        orc_face_mvp(fr, f, mvp);
        orc_face_viewport(W, f, vp, sc);
    }
    uint32_t *item = malloc(sizeof(uint32_t) * RES * K), *item2 = malloc(sizeof(uint32_t) * RES);
    float *depth = malloc(sizeof(float) * RES);
    uint64_t frags = 0;
    for (uint32_t k = 0; k < K; ++k)
        frags += orc_render_hemicube(W, tri, tp, ntri, fr + 9 * (k * 37), item + k * RES, depth);
    frags += orc_render_hemicube_mt(W, tri, tp, ntri, fr, item2, 3);
    frags += orc_render_hemicube_variant(W, tri, tp, ntri, fr, item2, 7);
    float *ff = calloc((size_t)K * np, sizeof(float));
    uint64_t unk = 0;
    uint64_t trip = orc_process_hemicube_runs(W, K, 4, item, dff, np, ff, &unk);
    uint32_t *hits = malloc(sizeof(uint32_t) * K * np);
    double *ff64 = malloc(sizeof(double) * K * np);
    orc_hemicube_histogram(W, K, item, dff, np, hits, ff64);
    float *u = malloc(sizeof(float) * 3 * np), *shot = calloc(3 * np, sizeof(float));
    for (uint32_t i = 0; i < 3 * np; ++i)
        u[i] = i % 7 == 0 ? urand(0, 5) : 0.0f;
    uint32_t ids[8];
    uint32_t n1 = orc_select_reference(u, np, 8, ids);
    uint32_t a1 = orc_select_argmax_power(u, area, np);
    uint32_t n2 = orc_select_topk_power(u, area, np, 8, ids);
    double *u64 = malloc(sizeof(double) * 3 * np), *s64 = calloc(3 * np, sizeof(double));
    for (uint32_t i = 0; i < 3 * np; ++i)
        u64[i] = u[i];
    uint32_t sh3[3] = {0, 37, 74};
    orc_shoot(1, np, sh3, K, ff64, refl, area, u64, s64);
    double r = orc_residual(u64, u, area, np);
    float *ffs = malloc(sizeof(float) * np), le[3];
    uint32_t last = 0;
    for (int p = 0; p < 3; ++p) {
        last += orc_plain_pass(W, 1, 1, tri, tp, ntri, fr, refl, area, np, dff, u, shot, item2, ffs);
        last += orc_plain_pass_mt(W, 0, 0, tri, tp, ntri, fr, refl, area, np, dff, u, shot, item2,
                                  ffs, 2);
        last += orc_pass_k(W, 1, 1, 4, tri, tp, ntri, fr, refl, area, np, dff, u, shot, item2, ffs,
                           ids, le);
    }
    const uint32_t gn = 40, gR = 3;
    float *M = calloc((size_t)gn * gn, sizeof(float));
    for (uint32_t i = 0; i < gn; ++i)
        for (uint32_t j = 0; j < gn; ++j)
            M[(size_t)i * gn + j] = i == j ? 0.0f : urand(0.0f, 1.0f / gn);
    double *E = calloc((size_t)gR * gn, sizeof(double)), *B = malloc(sizeof(double) * gR * gn);
    E[3] = 1.0;
    E[gn + 7] = 2.0;
    uint32_t it = orc_gather_jacobi(gn, gR, M, refl, area, E, 1e-12, 1000, B);
    printf("{\"ok\": 1, \"sum_test\": %.6f, \"frags\": %llu, \"triplets\": %llu, \"unknown\": %llu, "
           "\"sel\": [%u, %u, %u], \"residual\": %.6g, \"passes\": %u, \"jacobi\": %u}\n",
           (double)sumt, (unsigned long long)frags, (unsigned long long)trip,
           (unsigned long long)unk, n1, a1, n2, r, last, it);
    free(tri); free(tp); free(fr); free(refl); free(area); free(dff); free(top); free(side);
    free(item); free(item2); free(depth); free(ff); free(hits); free(ff64); free(u); free(shot);
    free(u64); free(s64); free(ffs); free(M); free(E); free(B);
    return 0;
}
