/*
 * oracle.c - CPU restatement of david-sabata/Radiosity's hemicube shooting
 * path (Config, delta form factors, hemicube camera, GL item-buffer render,
 * ProcessHemicube, shooting, selection).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests, the smoke check
 * and the benchmark's CPU-baseline leg use it to verify and time the HIP
 * library.  The library never links, loads or calls it.
 *
 * Build with -ffp-contract=off: the float expressions below are evaluated
 * in the written order without fused multiply-adds, except where fmaf() is
 * written explicitly.
 *
 * Parity status per function (how each is pinned, see tests/):
 *   orc_dff_*            closed form, known answer 1.00435 (FormFactors.cpp:106),
 *                        layout invariants
 *   orc_quad_frame       closed form on hand-checked quads
 *   orc_face_mvp         orthonormality, mirror determinant, FRONT axis = normal,
 *                        projection of known points
 *   orc_render_hemicube  brute-force point-in-triangle on tiny inputs, the
 *                        closed-room invariant (no background), analytic
 *                        plate-to-plate form factors.  GL's own vendor
 *                        rasterization (24-bit depth, driver edge rules) cannot be
 *                        run here: parity to the GL driver is unpinned and
 *                        stated as an edge-mismatch bound.
 *   orc_process_*        run-length form == per-pixel histogram (bit-exact hits)
 *   orc_shoot            energy conservation (PHYSICAL), hand-computed 2-patch case
 *   orc_select_*         brute-force restatements on small arrays with ties/zeros
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* FormFactors.cpp:3 */
#define ORC_PI (3.1415926535897932384626433832795028841931f)

/* ------------------------------------------------------------------ Config */

void orc_geometry(uint32_t W, uint32_t *tex_w, uint32_t *tex_h, uint32_t *tex_res)
{
    /* Config.cpp:37-39: TEX_W = W*2, TEX_H = unsigned(W*1.5) in double */
    uint32_t tw = W * 2u;
    uint32_t th = (uint32_t)((double)W * 1.5);
    *tex_w = tw;
    *tex_h = th;
    *tex_res = tw * th;
}

/* ------------------------------------------------------- delta form factors */

void orc_dff_top_side(uint32_t W, float *top, float *side)
{
    /* FormFactors.cpp:34-37 */
    int n = (int)W;
    float hpw = 1.0f / n;
    float pa = 2.0f / n;
    pa *= pa;
    /* FormFactors.cpp:40-52: top face */
    for (int x = 0; x < n; ++x)
        for (int y = 0; y < n; ++y) {
            float dx = ((x - n / 2) / (n / 2.0f)) + hpw;
            float dy = ((y - n / 2) / (n / 2.0f)) + hpw;
            float f = (dx * dx + dy * dy + 1);
            f *= f * ORC_PI;
            top[x + y * n] = pa / f;
        }
    /* FormFactors.cpp:55-67: side face, numerator height offset by one
     * extra half pixel (dy + hpw) */
    for (int x = 0; x < n; ++x)
        for (int y = 0; y < n / 2; ++y) {
            float dx = (x - n / 2) / (n / 2.0f) + hpw;
            float dy = (n / 2 - 1 - y) / (n / 2.0f) + hpw;
            float f = (dx * dx + dy * dy + 1);
            f *= f * ORC_PI;
            side[x + y * n] = (pa * (dy + hpw)) / f;
        }
}

float orc_dff_sum_test(uint32_t W)
{
    int n = (int)W;
    float *top = (float *)malloc(sizeof(float) * n * n);
    float *side = (float *)malloc(sizeof(float) * n * n / 2);
    orc_dff_top_side(W, top, side);
    /* FormFactors.cpp:89-103 */
    float s = 0;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n / 2; ++x)
            s += side[y + x * n];
    s *= 4;
    for (int x = 0; x < n; ++x)
        for (int y = 0; y < n; ++y)
            s += top[x + y * n];
    free(top);
    free(side);
    return s;
}

void orc_dff_cross(uint32_t W, float *out)
{
    uint32_t TW, TH, RES;
    orc_geometry(W, &TW, &TH, &RES);
    uint32_t H = W;
    float *top = (float *)malloc(sizeof(float) * W * W);
    float *side = (float *)malloc(sizeof(float) * W * W / 2);
    orc_dff_top_side(W, top, side);
    /* FormFactors.cpp:295-314 (first copy only; 316-323 replicate it K times) */
    for (uint32_t i = 0; i < RES; ++i) {
        uint32_t x = i % TW, y = i / TW;
        float v = 0.0f;
        if (x >= W / 2 && x < W * 1.5 && y < H)
            v = top[y * W + (x - W / 2)];
        else if (x < W / 2 && y < H)
            v = side[(H / 2 - x) * W - y - 1];
        else if (x >= W * 1.5 && y < H)
            v = side[(x - (uint32_t)(W * 1.5)) * W + y];
        else if (x < W && y >= H)
            v = side[(y - H) * W + x];
        else if (x >= W && y >= H)
            v = side[((H / 2 - 1) - (y - H)) * W + (x - W)];
        out[i] = v;
    }
    free(top);
    free(side);
}

/* ------------------------------------------------------------------ camera */

typedef struct { float x, y, z; } v3;

static v3 v3_sub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static v3 v3_add(v3 a, v3 b) { v3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static v3 v3_neg(v3 a) { v3 r = {-a.x, -a.y, -a.z}; return r; }
/* Vector.h:534-537: a.v_Cross(b) evaluates to b x a */
static v3 v3_vcross(v3 a, v3 b)
{
    v3 r = {b.y * a.z - b.z * a.y, b.z * a.x - b.x * a.z, b.x * a.y - b.y * a.x};
    return r;
}
/* Vector.h:389-399 */
static v3 v3_normalize(v3 a)
{
    float t = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    if (t != 0) {
        t = 1 / t;
        a.x *= t;
        a.y *= t;
        a.z *= t;
    }
    return a;
}

void orc_quad_frame(const float *v, float *frame)
{
    v3 v1 = {v[0], v[1], v[2]}, v2 = {v[3], v[4], v[5]};
    v3 v3_ = {v[6], v[7], v[8]}, v4 = {v[9], v[10], v[11]};
    /* Patch.cpp:261-267 */
    frame[0] = (v1.x + v2.x + v3_.x + v4.x) / 4.0f;
    frame[1] = (v1.y + v2.y + v3_.y + v4.y) / 4.0f;
    frame[2] = (v1.z + v2.z + v3_.z + v4.z) / 4.0f;
    /* Patch.cpp:272-276: (v2-v1).Cross(v4-v1) = (v4-v1) x (v2-v1) */
    v3 n = v3_vcross(v3_sub(v2, v1), v3_sub(v4, v1));
    frame[3] = n.x; frame[4] = n.y; frame[5] = n.z;
    /* Patch.cpp:253-255 */
    v3 u = v3_sub(v4, v1);
    frame[6] = u.x; frame[7] = u.y; frame[8] = u.z;
}

void orc_face_mvp(const float *fr, int face, float *mvp)
{
    v3 eye = {fr[0], fr[1], fr[2]};
    v3 n = {fr[3], fr[4], fr[5]};
    v3 pu = {fr[6], fr[7], fr[8]};
    v3 target, up;
    /* Camera.cpp:26-51 */
    switch (face) {
    case 0: target = pu; up = v3_neg(n); break;               /* UP    */
    case 1: target = v3_neg(pu); up = n; break;               /* DOWN  */
    case 2: target = v3_neg(v3_vcross(n, pu)); up = pu; break; /* LEFT  */
    case 3: target = v3_vcross(n, pu); up = pu; break;        /* RIGHT */
    default: target = n; up = pu; break;                      /* FRONT */
    }
    /* Camera.cpp:97-103 -> Transform.cpp:127-156 LookAt(eye, target + eye, up) */
    v3 dir = v3_normalize(v3_sub(v3_add(target, eye), eye));
    v3 right = v3_normalize(v3_vcross(dir, up));
    v3 up2 = v3_vcross(right, dir);
    float V[4][4]; /* V[col][row] */
    float rr[3] = {right.x, right.y, right.z};
    float uu[3] = {up2.x, up2.y, up2.z};
    float dd[3] = {dir.x, dir.y, dir.z};
    for (int i = 0; i < 3; ++i) {
        V[i][0] = rr[i];
        V[i][1] = uu[i];
        V[i][2] = -dd[i];
        V[i][3] = 0;
        V[3][i] = 0;
    }
    V[3][3] = 1;
    /* Matrix4f::Translate (Vector.cpp:325-330) multiplies by T(-eye) on the
     * right (operator*=, Vector.cpp:478-...).  Columns 0..2 are unchanged;
     * column 3 becomes sum_k V[k][r] * t_k + V[3][r]. */
    float t[4] = {-eye.x, -eye.y, -eye.z, 1.0f};
    float c3[4];
    for (int r = 0; r < 4; ++r)
        c3[r] = V[0][r] * t[0] + V[1][r] * t[1] + V[2][r] * t[2] + V[3][r] * t[3];
    for (int r = 0; r < 4; ++r)
        V[3][r] = c3[r];
    /* Transform.cpp:70-80 Perspective(90, 1, 0.01, 1000): tan(45deg) rounds
     * to 1.0f, so the frustum is +-0.01 at near 0.01 (Transform.cpp:26-46). */
    float fn = 0.01f, ff = 1000.0f;
    float fh = 1.0f * fn, fw = fh * 1.0f;
    float l = -fw, rt = fw, b = -fh, tp = fh;
    float P[4][4];
    memset(P, 0, sizeof(P));
    P[0][0] = 2 * fn / (rt - l);
    P[2][0] = (rt + l) / (rt - l);
    P[1][1] = 2 * fn / (tp - b);
    P[2][1] = (tp + b) / (tp - b);
    P[2][2] = -(ff + fn) / (ff - fn);
    P[3][2] = -2 * ff * fn / (ff - fn);
    P[2][3] = -1;
    P[3][3] = 0;
    /* Main.cpp:1183 t_mvp = P * MV; Matrix4f::ProductOf Vector.cpp:445-458 */
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            mvp[i * 4 + j] = P[0][j] * V[i][0] + P[1][j] * V[i][1] +
                             P[2][j] * V[i][2] + P[3][j] * V[i][3];
}

void orc_face_viewport(uint32_t W, int face, int32_t *vp, int32_t *sc)
{
    int32_t w = (int32_t)W, h = (int32_t)W;
    /* Main.cpp:316-347 (hi = 0) */
    vp[2] = w;
    vp[3] = h;
    switch (face) {
    case 0: vp[0] = 0; vp[1] = h; break;
    case 1: vp[0] = w; vp[1] = h / 2; break;
    case 2: vp[0] = -(w / 2); vp[1] = 0; break;
    case 3: vp[0] = (int32_t)(w * 1.5); vp[1] = 0; break;
    default: vp[0] = w / 2; vp[1] = 0; break;
    }
    /* Main.cpp:354-389 */
    if (face == 0 || face == 1) {
        sc[1] = h;
        sc[3] = h / 2;
    } else {
        sc[1] = 0;
        sc[3] = h;
    }
    sc[2] = (face == 2 || face == 3) ? w / 2 : w;
    switch (face) {
    case 0: case 2: sc[0] = 0; break;
    case 1: sc[0] = w; break;
    case 3: sc[0] = (int32_t)(w * 1.5); break;
    default: sc[0] = w / 2; break;
    }
}

/* --------------------------------------------------------------- rasterizer
 *
 * GL 3.0 rasterization as restated for this path:
 *  - clip coords c = MVP * (x,y,z,1), each row summed left to right
 *  - clip against z >= -w (near) and the guard band |x|,|y| <= G*w with
 *    Sutherland-Hodgman; an intersection is always interpolated from the
 *    inside vertex towards the outside one (so a shared edge clips the same
 *    way from both of its triangles); the far plane is left to the depth
 *    test (depth >= 1 never passes GL_LESS against the 1.0 clear value)
 *  - window coords relative to the face viewport: with iw = 1/w (rounded
 *    once), xl = (x*iw + 1) * W/2, yl likewise, depth z*iw*0.5 + 0.5; x and y
 *    are snapped to 1/256 pixel with round-to-nearest-even
 *  - the clipped polygon is fanned from its first vertex; a sub-triangle
 *    with snapped signed area <= 0 is culled (GL_CULL_FACE, GL_BACK, CCW)
 *  - a pixel (ix,iy) inside the scissor is covered if its centre passes
 *    all three integer edge functions, ties resolved by a top-left rule
 *  - depth z = fmaf(dzdx, px-x0, fmaf(dzdy, py-y0, z0)) with the plane
 *    slopes computed in double and rounded to float; z <= 0 clamps to 0
 *  - GL_LESS: the fragment replaces the stored one if z < stored depth.
 */

#define ORC_SUB 256
#define ORC_GUARD 4096.0f

typedef struct { float c[4]; } cv4; /* x, y, z, w */

static float clip_dist(const cv4 *v, int plane)
{
    switch (plane) {
    case 0: return v->c[2] + v->c[3];
    case 1: return ORC_GUARD * v->c[3] - v->c[0];
    case 2: return ORC_GUARD * v->c[3] + v->c[0];
    case 3: return ORC_GUARD * v->c[3] - v->c[1];
    default: return ORC_GUARD * v->c[3] + v->c[1];
    }
}

static int clip_polygon(cv4 *poly, int n)
{
    cv4 tmp[12];
    for (int plane = 0; plane < 5; ++plane) {
        float d[12];
        int nin = 0;
        for (int i = 0; i < n; ++i) {
            d[i] = clip_dist(&poly[i], plane);
            nin += d[i] >= 0.0f;
        }
        if (nin == n)
            continue;
        if (nin == 0)
            return 0;
        int m = 0;
        for (int i = 0; i < n; ++i) {
            int j = (i + 1 == n) ? 0 : i + 1;
            int ini = d[i] >= 0.0f, inj = d[j] >= 0.0f;
            if (ini)
                tmp[m++] = poly[i];
            if (ini != inj) {
                const cv4 *a = ini ? &poly[i] : &poly[j];
                const cv4 *b = ini ? &poly[j] : &poly[i];
                float da = ini ? d[i] : d[j], db = ini ? d[j] : d[i];
                float t = da / (da - db);
                cv4 r;
                for (int c = 0; c < 4; ++c)
                    r.c[c] = a->c[c] + t * (b->c[c] - a->c[c]);
                tmp[m++] = r;
            }
        }
        n = m;
        memcpy(poly, tmp, sizeof(cv4) * n);
    }
    return n;
}

typedef struct {
    int32_t x, y;
    float z;
} sv;

static inline int32_t floor_shift8(int32_t v) { return v >> 8; } /* arithmetic */

static uint64_t raster_tri(uint32_t W, const sv *a, const sv *b, const sv *c,
                           const int32_t *sc_local, uint32_t id, const int32_t *vp,
                           uint32_t TW, uint32_t *item, float *zbuf, int flags)
{
    int64_t area2 = (int64_t)(b->x - a->x) * (c->y - a->y) -
                    (int64_t)(c->x - a->x) * (b->y - a->y);
    if (area2 <= 0)
        return 0;
    const sv *v[3] = {a, b, c};
    int64_t A[3], B[3], C[3];
    int32_t bias[3];
    for (int e = 0; e < 3; ++e) {
        const sv *p = v[e], *q = v[(e + 1) % 3];
        A[e] = (int64_t)p->y - q->y;
        B[e] = (int64_t)q->x - p->x;
        C[e] = -(A[e] * p->x + B[e] * p->y);
        bias[e] = (A[e] > 0 || (A[e] == 0 && B[e] < 0)) ? 1 : 0;
    }
    int32_t xmin = a->x, xmax = a->x, ymin = a->y, ymax = a->y;
    for (int e = 1; e < 3; ++e) {
        if (v[e]->x < xmin) xmin = v[e]->x;
        if (v[e]->x > xmax) xmax = v[e]->x;
        if (v[e]->y < ymin) ymin = v[e]->y;
        if (v[e]->y > ymax) ymax = v[e]->y;
    }
    int32_t ix0 = floor_shift8(xmin - 128 + 255), ix1 = floor_shift8(xmax - 128);
    int32_t iy0 = floor_shift8(ymin - 128 + 255), iy1 = floor_shift8(ymax - 128);
    if (ix0 < sc_local[0]) ix0 = sc_local[0];
    if (iy0 < sc_local[1]) iy0 = sc_local[1];
    if (ix1 > sc_local[2] - 1) ix1 = sc_local[2] - 1;
    if (iy1 > sc_local[3] - 1) iy1 = sc_local[3] - 1;
    if (ix0 > ix1 || iy0 > iy1)
        return 0;
    /* depth plane */
    double dx1 = (double)(b->x - a->x), dy1 = (double)(b->y - a->y);
    double dx2 = (double)(c->x - a->x), dy2 = (double)(c->y - a->y);
    double dz1 = (double)b->z - (double)a->z, dz2 = (double)c->z - (double)a->z;
    double ar = (double)area2;
    float dzdx = (float)((dz1 * dy2 - dz2 * dy1) / ar);
    float dzdy = (float)((dx1 * dz2 - dx2 * dz1) / ar);
    float z0 = a->z;
    uint64_t frags = 0;
    (void)W;
    for (int32_t iy = iy0; iy <= iy1; ++iy) {
        int32_t py = iy * ORC_SUB + ORC_SUB / 2;
        for (int32_t ix = ix0; ix <= ix1; ++ix) {
            int32_t px = ix * ORC_SUB + ORC_SUB / 2;
            int in = 1;
            for (int e = 0; e < 3; ++e) {
                int64_t E = A[e] * px + B[e] * py + C[e];
                if (E + bias[e] <= 0) { in = 0; break; }
            }
            if (!in)
                continue;
            ++frags;
            float z = fmaf(dzdx, (float)(px - a->x), fmaf(dzdy, (float)(py - a->y), z0));
            if (!(z > 0.0f))
                z = 0.0f;
            if (!(z < 1.0f))
                continue;
            if (flags & ORC_VAR_DEPTH24) /* stored as round(z (2^24 - 1)) */
                z = (float)((double)lrint((double)z * 16777215.0) / 16777215.0);
            uint32_t tx = (uint32_t)(vp[0] + ix), ty = (uint32_t)(vp[1] + iy);
            uint32_t texel = tx + ty * TW;
            if (z < zbuf[texel]) {
                zbuf[texel] = z;
                item[texel] = id;
            }
        }
    }
    return frags;
}

static uint64_t render_impl(uint32_t W, const float *tri, const uint32_t *tri_patch,
                            uint32_t ntri, const float *frame, uint32_t *item,
                            float *depth_or_null, int flags);

uint64_t orc_render_hemicube(uint32_t W, const float *tri, const uint32_t *tri_patch,
                             uint32_t ntri, const float *frame, uint32_t *item,
                             float *depth_or_null)
{
    return render_impl(W, tri, tri_patch, ntri, frame, item, depth_or_null, 0);
}

/* The same render with arithmetic a GL driver may use instead, to measure
 * how many item-buffer texels such differences change (flags may be
 * combined):
 *   ORC_VAR_DEPTH24  depth rounded to 24-bit fixed point before the GL_LESS
 *                    test (the reference's GL_DEPTH_COMPONENT24, Main.cpp:288)
 *   ORC_VAR_FMA      clip coordinates as fused multiply-adds
 *   ORC_VAR_DIVIDE   x / w instead of x * (1 / w) */
uint64_t orc_render_hemicube_variant(uint32_t W, const float *tri, const uint32_t *tri_patch,
                                     uint32_t ntri, const float *frame, uint32_t *item,
                                     int flags)
{
    return render_impl(W, tri, tri_patch, ntri, frame, item, NULL, flags);
}

/* Draws triangles [t0, t1) of the five faces into item / zbuf (already
 * cleared), in draw order. */
static uint64_t render_range(uint32_t W, const float *tri, const uint32_t *tri_patch,
                             uint32_t t0, uint32_t t1, const float *frame, uint32_t *item,
                             float *zbuf, int flags)
{
    uint32_t TW, TH, RES;
    orc_geometry(W, &TW, &TH, &RES);
    float halfW = 0.5f * (float)W;
    uint64_t frags = 0;
    for (int face = 0; face < 5; ++face) {
        float m[16];
        int32_t vp[4], sc[4], scl[4];
        orc_face_mvp(frame, face, m);
        orc_face_viewport(W, face, vp, sc);
        /* scissor in face-local pixels [x0, y0, x1, y1) */
        scl[0] = sc[0] - vp[0];
        scl[1] = sc[1] - vp[1];
        scl[2] = sc[0] + sc[2] - vp[0];
        scl[3] = sc[1] + sc[3] - vp[1];
        for (uint32_t t = t0; t < t1; ++t) {
            cv4 poly[12];
            for (int k = 0; k < 3; ++k) {
                float x = tri[t * 9 + k * 3 + 0], y = tri[t * 9 + k * 3 + 1],
                      z = tri[t * 9 + k * 3 + 2];
                for (int r = 0; r < 4; ++r)
                    poly[k].c[r] = (flags & ORC_VAR_FMA)
                                       ? fmaf(m[0 * 4 + r], x,
                                              fmaf(m[1 * 4 + r], y, fmaf(m[2 * 4 + r], z,
                                                                         m[3 * 4 + r])))
                                       : m[0 * 4 + r] * x + m[1 * 4 + r] * y +
                                             m[2 * 4 + r] * z + m[3 * 4 + r];
            }
            int n = clip_polygon(poly, 3);
            if (n < 3)
                continue;
            sv s[12];
            for (int k = 0; k < n; ++k) {
                float iw = 1.0f / poly[k].c[3];
                float xn = poly[k].c[0] * iw, yn = poly[k].c[1] * iw, zn = poly[k].c[2] * iw;
                if (flags & ORC_VAR_DIVIDE) {
                    xn = poly[k].c[0] / poly[k].c[3];
                    yn = poly[k].c[1] / poly[k].c[3];
                    zn = poly[k].c[2] / poly[k].c[3];
                }
                float xl = (xn + 1.0f) * halfW, yl = (yn + 1.0f) * halfW;
                s[k].x = (int32_t)rintf(xl * (float)ORC_SUB);
                s[k].y = (int32_t)rintf(yl * (float)ORC_SUB);
                s[k].z = zn * 0.5f + 0.5f;
            }
            uint32_t id = tri_patch[t] + 1u;
            for (int k = 1; k + 1 < n; ++k)
                frags += raster_tri(W, &s[0], &s[k], &s[k + 1], scl, id, vp, TW, item, zbuf,
                                    flags);
        }
    }
    return frags;
}

static uint64_t render_impl(uint32_t W, const float *tri, const uint32_t *tri_patch,
                            uint32_t ntri, const float *frame, uint32_t *item,
                            float *depth_or_null, int flags)
{
    uint32_t TW, TH, RES;
    orc_geometry(W, &TW, &TH, &RES);
    float *zbuf = depth_or_null ? depth_or_null : (float *)malloc(sizeof(float) * RES);
    /* Main.cpp:1148 glClear: colour 0, depth 1.0 */
    for (uint32_t i = 0; i < RES; ++i) {
        zbuf[i] = 1.0f;
        item[i] = 0;
    }
    uint64_t frags = render_range(W, tri, tri_patch, 0, ntri, frame, item, zbuf, flags);
    if (!depth_or_null)
        free(zbuf);
    return frags;
}

/* orc_render_hemicube on nthreads OpenMP threads: thread t draws the t-th
 * contiguous slice of the triangles into its own buffers, and the buffers
 * are merged texel by texel in slice order keeping the first strictly
 * nearer fragment - the GL_LESS result of drawing all triangles in order.
 * The CPU baseline's multi-core leg; equal to orc_render_hemicube. */
uint64_t orc_render_hemicube_mt(uint32_t W, const float *tri, const uint32_t *tri_patch,
                                uint32_t ntri, const float *frame, uint32_t *item,
                                int nthreads)
{
    uint32_t TW, TH, RES;
    orc_geometry(W, &TW, &TH, &RES);
    if (nthreads < 1)
        nthreads = 1;
    uint32_t *items = (uint32_t *)malloc(sizeof(uint32_t) * RES * (size_t)nthreads);
    float *zs = (float *)malloc(sizeof(float) * RES * (size_t)nthreads);
    uint64_t frags = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : frags)
    {
        int t = omp_get_thread_num(), nt = omp_get_num_threads();
        uint32_t *it = items + (size_t)t * RES;
        float *z = zs + (size_t)t * RES;
        for (uint32_t i = 0; i < RES; ++i) {
            z[i] = 1.0f;
            it[i] = 0;
        }
        uint32_t a = (uint32_t)((uint64_t)ntri * t / nt), b = (uint32_t)((uint64_t)ntri * (t + 1) / nt);
        frags += render_range(W, tri, tri_patch, a, b, frame, it, z, 0);
#pragma omp barrier
#pragma omp for
        for (uint32_t i = 0; i < RES; ++i) {
            float bz = zs[i];
            uint32_t bi = items[i];
            for (int q = 1; q < nt; ++q)
                if (zs[(size_t)q * RES + i] < bz) {
                    bz = zs[(size_t)q * RES + i];
                    bi = items[(size_t)q * RES + i];
                }
            item[i] = bi;
        }
    }
    free(items);
    free(zs);
    return frags;
}

/* --------------------------------------------------------- ProcessHemicube */

uint64_t orc_process_hemicube_runs(uint32_t W, uint32_t K, uint32_t wix,
                                   const uint32_t *item, const float *dff,
                                   uint32_t npatch, float *ff, uint64_t *unknown)
{
    uint32_t TW, TH, RES;
    orc_geometry(W, &TW, &TH, &RES);
    /* Config.cpp:43-45 */
    uint32_t nx = wix < TW ? wix : TW;
    uint32_t span = TW / nx;
    uint64_t emitted = 0, unk = 0;
    /* Kernel_ProcessHemicube.h:23-69; the ΔFF table is the same for every
     * hemicube (FormFactors.cpp:316-323 replicates one copy). */
    for (uint32_t y = 0; y < TH * K; ++y) {
        const float *frow = dff + (size_t)(y % TH) * TW;
        for (uint32_t gx = 0; gx < nx; ++gx) {
            uint32_t x0 = gx * span < TW - 1 ? gx * span : TW - 1;
            uint32_t x1 = x0 + span < TW ? x0 + span : TW;
            while (x0 < x1) {
                uint32_t id = item[(size_t)y * TW + x0];
                float e = 0;
                while (x0 < x1 && item[(size_t)y * TW + x0] == id) {
                    e += frow[x0];
                    ++x0;
                }
                if (id > 0) {
                    ++emitted;
                    uint32_t hi = y / TH, pid = id - 1;
                    /* Main.cpp:1259-1268 */
                    if (pid >= npatch) {
                        ++unk;
                        continue;
                    }
                    ff[(size_t)hi * npatch + pid] += e;
                }
            }
        }
    }
    if (unknown)
        *unknown = unk;
    return emitted;
}

uint64_t orc_process_hemicube_triplets(uint32_t W, uint32_t K, uint32_t wix,
                                       const uint32_t *item, const float *dff, uint64_t cap,
                                       uint32_t *hemis, uint32_t *ids, float *energies)
{
    uint32_t TW, TH, RES;
    orc_geometry(W, &TW, &TH, &RES);
    uint32_t nx = wix < TW ? wix : TW; /* Config.cpp:43-45 */
    uint32_t span = TW / nx;
    uint64_t count = 0; /* the write index, advanced 4 slots at a time */
    for (uint32_t y = 0; y < TH * K; ++y) {
        const float *frow = dff + (size_t)(y % TH) * TW;
        for (uint32_t gx = 0; gx < nx; ++gx) {
            /* Kernel_ProcessHemicube.h:23-69, one work-item */
            uint32_t x0 = gx * span < TW - 1 ? gx * span : TW - 1;
            uint32_t x1 = x0 + span < TW ? x0 + span : TW;
            uint32_t space = 0;
            uint64_t wid = 0;
            while (x0 < x1) {
                uint32_t id = item[(size_t)y * TW + x0];
                float e = 0;
                while (x0 < x1 && item[(size_t)y * TW + x0] == id) {
                    e += frow[x0];
                    ++x0;
                }
                if (id > 0) {
                    if (!space) {
                        wid = count;
                        count += 4;
                        space = 3;
                    } else {
                        ++wid;
                        --space;
                    }
                    if (wid < cap) {
                        hemis[wid] = y / TH;
                        ids[wid] = id - 1;
                        energies[wid] = e;
                    }
                }
            }
            for (; space > 0; --space) {
                ++wid;
                if (wid < cap) {
                    hemis[wid] = 0;
                    ids[wid] = 0;
                    energies[wid] = 0.0f;
                }
            }
        }
    }
    return count;
}

void orc_hemicube_histogram(uint32_t W, uint32_t K, const uint32_t *item,
                            const float *dff, uint32_t npatch, uint32_t *hits,
                            double *ff64)
{
    uint32_t TW, TH, RES;
    orc_geometry(W, &TW, &TH, &RES);
    memset(hits, 0, sizeof(uint32_t) * (size_t)K * npatch);
    memset(ff64, 0, sizeof(double) * (size_t)K * npatch);
    for (uint32_t hi = 0; hi < K; ++hi)
        for (uint32_t i = 0; i < RES; ++i) {
            uint32_t id = item[(size_t)hi * RES + i];
            if (id == 0 || id - 1 >= npatch)
                continue;
            hits[(size_t)hi * npatch + id - 1] += 1;
            ff64[(size_t)hi * npatch + id - 1] += (double)dff[i];
        }
}

/* ---------------------------------------------------------------- selection */

static float len2(const float *u, uint32_t i)
{
    /* Vector.h:356-359 f_Length2 */
    return u[3 * i] * u[3 * i] + u[3 * i + 1] * u[3 * i + 1] + u[3 * i + 2] * u[3 * i + 2];
}

uint32_t orc_select_reference(const float *unshot, uint32_t npatch, uint32_t count,
                              uint32_t *ids)
{
    /* ModelContainer.cpp:263-298 restated on an array, step by step:
     * push_back, stable ascending sort by |B|^2 (std::list::sort is
     * stable), reverse, truncate to count.  Equal keys therefore swap their
     * relative order on every insertion, exactly as the list does. */
    uint32_t *tops = (uint32_t *)malloc(sizeof(uint32_t) * (count + 2));
    uint32_t n = 0;
    for (uint32_t pi = 0; pi < npatch; ++pi) {
        float e = len2(unshot, pi);
        if (n == 0 || (e > 0 && len2(unshot, tops[n - 1]) <= e)) {
            tops[n++] = pi;
            /* stable insertion sort, ascending */
            for (uint32_t a = 1; a < n; ++a) {
                uint32_t v = tops[a];
                float kv = len2(unshot, v);
                uint32_t b = a;
                while (b > 0 && len2(unshot, tops[b - 1]) > kv) {
                    tops[b] = tops[b - 1];
                    --b;
                }
                tops[b] = v;
            }
            for (uint32_t a = 0; a < n / 2; ++a) {
                uint32_t t = tops[a];
                tops[a] = tops[n - 1 - a];
                tops[n - 1 - a] = t;
            }
            if (n > count)
                n = count;
        }
    }
    for (uint32_t i = 0; i < count; ++i)
        ids[i] = i < n ? tops[i] : 0;
    free(tops);
    return n;
}

uint32_t orc_select_argmax_power(const float *unshot, const float *area, uint32_t npatch)
{
    uint32_t best = 0;
    float bk = -1.0f;
    for (uint32_t i = 0; i < npatch; ++i) {
        float k = area[i] * (unshot[3 * i] + unshot[3 * i + 1] + unshot[3 * i + 2]);
        if (k > bk) {
            bk = k;
            best = i;
        }
    }
    return best;
}

/* ----------------------------------------------------------------- shooting */

void orc_shoot(uint32_t mode, uint32_t npatch, const uint32_t *shooters, uint32_t k,
               const double *ff, const float *refl, const float *area, double *unshot,
               double *shot)
{
    double *snap = (double *)malloc(sizeof(double) * 3 * (k ? k : 1));
    /* Main.cpp:1161 snapshot taken before any update of the pass */
    for (uint32_t s = 0; s < k; ++s)
        for (int c = 0; c < 3; ++c)
            snap[3 * s + c] = unshot[3 * shooters[s] + c];
    for (uint32_t s = 0; s < k; ++s) {
        uint32_t i = shooters[s];
        for (uint32_t j = 0; j < npatch; ++j) {
            double f = ff[(size_t)s * npatch + j];
            if (f == 0.0)
                continue;
            for (int c = 0; c < 3; ++c) {
                double d;
                if (mode == 0) /* Main.cpp:1274 */
                    d = snap[3 * s + c] * f * (double)0.3f * (double)refl[3 * i + c];
                else
                    d = (double)refl[3 * j + c] * snap[3 * s + c] * f *
                        ((double)area[i] / (double)area[j]);
                unshot[3 * j + c] += d;
            }
        }
    }
    /* Main.cpp:1287-1295 */
    for (uint32_t s = 0; s < k; ++s)
        for (int c = 0; c < 3; ++c) {
            shot[3 * shooters[s] + c] += snap[3 * s + c];
            unshot[3 * shooters[s] + c] -= snap[3 * s + c];
        }
    free(snap);
}

double orc_residual(const double *unshot, const float *emission, const float *area,
                    uint32_t npatch)
{
    double num = 0, den = 0;
    for (uint32_t i = 0; i < npatch; ++i) {
        num += (double)area[i] * (unshot[3 * i] + unshot[3 * i + 1] + unshot[3 * i + 2]);
        den += (double)area[i] * ((double)emission[3 * i] + emission[3 * i + 1] +
                                  emission[3 * i + 2]);
    }
    return den > 0 ? num / den : 0.0;
}

static uint32_t plain_pass(uint32_t W, uint32_t mode, uint32_t select_rule, const float *tri,
                           const uint32_t *tri_patch, uint32_t ntri, const float *frames,
                           const float *refl, const float *area, uint32_t npatch,
                           const float *dff, float *unshot, float *shot, uint32_t *item,
                           float *ff, int nthreads);

uint32_t orc_plain_pass(uint32_t W, uint32_t mode, uint32_t select_rule,
                        const float *tri, const uint32_t *tri_patch, uint32_t ntri,
                        const float *frames, const float *refl, const float *area,
                        uint32_t npatch, const float *dff, float *unshot, float *shot,
                        uint32_t *item, float *ff)
{
    return plain_pass(W, mode, select_rule, tri, tri_patch, ntri, frames, refl, area, npatch,
                      dff, unshot, shot, item, ff, 0);
}

/* orc_plain_pass with the hemicube rendered by orc_render_hemicube_mt and
 * the shooting loop split over the same threads; same results. */
uint32_t orc_plain_pass_mt(uint32_t W, uint32_t mode, uint32_t select_rule,
                           const float *tri, const uint32_t *tri_patch, uint32_t ntri,
                           const float *frames, const float *refl, const float *area,
                           uint32_t npatch, const float *dff, float *unshot, float *shot,
                           uint32_t *item, float *ff, int nthreads)
{
    return plain_pass(W, mode, select_rule, tri, tri_patch, ntri, frames, refl, area, npatch,
                      dff, unshot, shot, item, ff, nthreads < 1 ? 1 : nthreads);
}

static uint32_t plain_pass(uint32_t W, uint32_t mode, uint32_t select_rule, const float *tri,
                           const uint32_t *tri_patch, uint32_t ntri, const float *frames,
                           const float *refl, const float *area, uint32_t npatch,
                           const float *dff, float *unshot, float *shot, uint32_t *item,
                           float *ff, int nthreads)
{
    uint32_t i;
    if (select_rule == 0)
        orc_select_reference(unshot, npatch, 1, &i);
    else
        i = orc_select_argmax_power(unshot, area, npatch);
    float snap[3] = {unshot[3 * i], unshot[3 * i + 1], unshot[3 * i + 2]};
    if (nthreads)
        orc_render_hemicube_mt(W, tri, tri_patch, ntri, frames + 9 * (size_t)i, item, nthreads);
    else
        orc_render_hemicube(W, tri, tri_patch, ntri, frames + 9 * (size_t)i, item, NULL);
    memset(ff, 0, sizeof(float) * npatch);
    orc_process_hemicube_runs(W, 1, 4, item, dff, npatch, ff, NULL);
    /* Main.cpp:1272-1275 in float, as the reference's Patch fields */
#pragma omp parallel for num_threads(nthreads ? nthreads : 1) if (nthreads > 1)
    for (uint32_t j = 0; j < npatch; ++j) {
        float f = ff[j];
        if (f == 0.0f)
            continue;
        for (int c = 0; c < 3; ++c) {
            if (mode == 0)
                unshot[3 * j + c] += snap[c] * f * 0.3f * refl[3 * i + c];
            else
                unshot[3 * j + c] += refl[3 * j + c] * snap[c] * f * (area[i] / area[j]);
        }
    }
    for (int c = 0; c < 3; ++c) {
        shot[3 * i + c] += snap[c];
        unshot[3 * i + c] -= snap[c];
    }
    return i;
}

/* POWER top-K: patches with A (r+g+b) > 0 by that key, descending, equal
 * keys by index ascending; returns how many (<= count) were filled. */
uint32_t orc_select_topk_power(const float *unshot, const float *area, uint32_t npatch,
                               uint32_t count, uint32_t *ids)
{
    uint32_t n = 0;
    for (uint32_t i = 0; i < npatch; ++i) {
        float k = area[i] * (unshot[3 * i] + unshot[3 * i + 1] + unshot[3 * i + 2]);
        if (!(k > 0.0f))
            continue;
        /* insertion into the sorted prefix ids[0..n) */
        uint32_t pos = n < count ? n : count;
        while (pos > 0) {
            uint32_t q = ids[pos - 1];
            float kq = area[q] * (unshot[3 * q] + unshot[3 * q + 1] + unshot[3 * q + 2]);
            if (kq >= k)
                break;
            --pos;
        }
        if (pos >= count)
            continue;
        uint32_t last = n < count ? n : count - 1;
        for (uint32_t a = last; a > pos; --a)
            ids[a] = ids[a - 1];
        ids[pos] = i;
        if (n < count)
            ++n;
    }
    for (uint32_t a = n; a < count; ++a)
        ids[a] = 0;
    return n;
}

/* One pass of the OnIdle loop with K hemicubes (Main.cpp:1137-1303), float
 * state as the reference's Patch fields:
 *  - select K shooters (ModelContainer.cpp:259-299 list rule, or the POWER
 *    argmax / top-K); unfilled slots are skipped (Main.cpp:1157,1253,1289);
 *  - snapshot every shooter's radiosity before any update (Main.cpp:1161);
 *  - per shooter in slot order: its hemicube, the run-form ProcessHemicube
 *    sums accumulated in float (Kernel_ProcessHemicube.h:23-69,
 *    Main.cpp:1251-1269), then every receiver updated (Main.cpp:1272-1275);
 *  - emitter update in slot order; lastEnergy is the last shooter's
 *    radiosity before its subtraction (Main.cpp:1286-1295).
 * item: TEX_RES scratch, ff: npatch scratch, ids: K entries (out).
 * last_energy3 gets lastEnergy.  Returns the number of shooters. */
uint32_t orc_pass_k(uint32_t W, uint32_t mode, uint32_t select_rule, uint32_t K,
                    const float *tri, const uint32_t *tri_patch, uint32_t ntri,
                    const float *frames, const float *refl, const float *area, uint32_t npatch,
                    const float *dff, float *unshot, float *shot, uint32_t *item, float *ff,
                    uint32_t *ids, float *last_energy3)
{
    uint32_t n;
    if (select_rule == 0)
        n = orc_select_reference(unshot, npatch, K, ids);
    else if (K == 1) {
        ids[0] = orc_select_argmax_power(unshot, area, npatch);
        n = 1;
    } else
        n = orc_select_topk_power(unshot, area, npatch, K, ids);
    float *snap = (float *)malloc(sizeof(float) * 3 * (n ? n : 1));
    for (uint32_t s = 0; s < n; ++s)
        for (int c = 0; c < 3; ++c)
            snap[3 * s + c] = unshot[3 * ids[s] + c];
    for (uint32_t s = 0; s < n; ++s) {
        const uint32_t i = ids[s];
        orc_render_hemicube(W, tri, tri_patch, ntri, frames + 9 * (size_t)i, item, NULL);
        memset(ff, 0, sizeof(float) * npatch);
        orc_process_hemicube_runs(W, 1, 4, item, dff, npatch, ff, NULL);
        for (uint32_t j = 0; j < npatch; ++j) {
            float f = ff[j];
            if (f == 0.0f)
                continue;
            for (int c = 0; c < 3; ++c) {
                if (mode == 0)
                    unshot[3 * j + c] += snap[3 * s + c] * f * 0.3f * refl[3 * i + c];
                else
                    unshot[3 * j + c] += refl[3 * j + c] * snap[3 * s + c] * f *
                                         (area[i] / area[j]);
            }
        }
    }
    for (int c = 0; c < 3; ++c)
        last_energy3[c] = 0.0f;
    for (uint32_t s = 0; s < n; ++s) {
        const uint32_t i = ids[s];
        for (int c = 0; c < 3; ++c) {
            last_energy3[c] = unshot[3 * i + c];
            shot[3 * i + c] += snap[3 * s + c];
            unshot[3 * i + c] -= snap[3 * s + c];
        }
    }
    free(snap);
    return n;
}

uint32_t orc_gather_jacobi(uint32_t n, uint32_t R, const float *M, const float *refl,
                           const float *area, const double *E, double tol, uint32_t max_it,
                           double *B)
{
    size_t nn = (size_t)R * n;
    double *Y = (double *)malloc(nn * sizeof(double));
    double den = 0.0;
    for (uint32_t r = 0; r < R; ++r)
        for (uint32_t i = 0; i < n; ++i)
            den += (double)area[i] * fabs(E[(size_t)r * n + i]);
    for (size_t e = 0; e < nn; ++e)
        B[e] = E[e];
    uint32_t it = 0;
    for (; it < max_it; ++it) {
        /* Y = E + diag(rho) M B */
        double num = 0.0;
        for (uint32_t i = 0; i < n; ++i) {
            const float *row = M + (size_t)i * n;
            for (uint32_t r = 0; r < R; ++r) {
                const double *b = B + (size_t)r * n;
                double s = 0.0;
                for (uint32_t j = 0; j < n; ++j)
                    s += (double)row[j] * b[j];
                double y = E[(size_t)r * n + i] + (double)refl[(size_t)i * 3 + r % 3] * s;
                Y[(size_t)r * n + i] = y;
                num += (double)area[i] * fabs(y - b[i]);
            }
        }
        for (size_t e = 0; e < nn; ++e)
            B[e] = Y[e];
        if (den <= 0.0 || num / den <= tol) {
            ++it;
            break;
        }
    }
    free(Y);
    return it;
}
