// rad_kernels.hip - CDNA4 (gfx950) kernels of the hemicube shooting path.
//
// Compiled with -ffp-contract=off: the camera, clipping, snapping and depth
// arithmetic below is written operation by operation so the item buffer is
// reproducible bit for bit on any IEEE-754 implementation of the same
// sequence (fmaf only where written).  Division and sqrt are HIP's
// correctly-rounded defaults.
//
// Kernels (one pass over K shooters):
//   k_select_*     shooter selection (+ FP64 residual partial sums)
//   k_prepare      dB snapshot and the five face MVPs per shooter
//   k_clear        item/depth keys to the clear value, F rows to 0, counters
//   k_raster_setup one lane per (triangle, face, shooter): transform, clip,
//                  snap, cull, scissor; small triangles rasterized in-lane,
//                  big ones emitted as 8x8 pixel tiles
//   k_raster_tiles one wavefront per 8x8 tile of a big triangle
//   k_reduce       ProcessHemicube: wave-level run-length reduction of the
//                  item buffer into per-patch form factors and hit counts
//   k_shoot        energy shooting into unshot RGB, emitter update
//
// Depth test: each texel holds a 64-bit key (float depth bits << 32 | id).
// atomicMin of keys gives the nearest fragment, ties to the smallest id,
// which is the fragment GL_LESS keeps when triangles are drawn in id order
// (Main.cpp:1104, DrawPatchLook Main.cpp:720).
#include "rad_internal.h"

#include <hip/hip_runtime.h>

namespace rad {

// ------------------------------------------------------------------ ΔFF table
// FormFactors.cpp:23-67 (values) and 280-314 (cross layout), one thread per
// texel of one TEX_W x TEX_H copy.
__device__ static float dff_top(int n, int x, int y)
{
    float hpw = 1.0f / n;
    float pa = 2.0f / n;
    pa = pa * pa;
    float dx = ((x - n / 2) / (n / 2.0f)) + hpw;
    float dy = ((y - n / 2) / (n / 2.0f)) + hpw;
    float f = (dx * dx + dy * dy + 1);
    f = f * (f * 3.1415926535897932384626433832795028841931f);
    return pa / f;
}

__device__ static float dff_side(int n, int x, int y)
{
    float hpw = 1.0f / n;
    float pa = 2.0f / n;
    pa = pa * pa;
    float dx = (x - n / 2) / (n / 2.0f) + hpw;
    float dy = (n / 2 - 1 - y) / (n / 2.0f) + hpw;
    float f = (dx * dx + dy * dy + 1);
    f = f * (f * 3.1415926535897932384626433832795028841931f);
    return (pa * (dy + hpw)) / f;
}

__global__ void k_dff(uint32_t W, uint32_t TW, uint32_t RES, float *out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= RES)
        return;
    uint32_t x = i % TW, y = i / TW, H = W;
    int n = (int)W;
    float v = 0.0f;
    // side[idx] with idx = xs + ys*W  ->  (xs, ys) = (idx % W, idx / W)
    if (x >= W / 2 && x < W + W / 2 && y < H) {
        v = dff_top(n, (int)(x - W / 2), (int)y);
    } else if (x < W / 2 && y < H) {
        uint32_t idx = (H / 2 - x) * W - y - 1;
        v = dff_side(n, (int)(idx % W), (int)(idx / W));
    } else if (x >= W + W / 2 && y < H) {
        uint32_t idx = (x - (W + W / 2)) * W + y;
        v = dff_side(n, (int)(idx % W), (int)(idx / W));
    } else if (x < W && y >= H) {
        uint32_t idx = (y - H) * W + x;
        v = dff_side(n, (int)(idx % W), (int)(idx / W));
    } else if (x >= W && y >= H) {
        uint32_t idx = ((H / 2 - 1) - (y - H)) * W + (x - W);
        v = dff_side(n, (int)(idx % W), (int)(idx / W));
    }
    out[i] = v;
}

void launch_dff(const Geometry &g, float *dff, hipStream_t st)
{
    k_dff<<<(g.RES + 255) / 256, 256, 0, st>>>(g.W, g.TW, g.RES, dff);
}

// -------------------------------------------------------------------- camera
struct f3 {
    float x, y, z;
};
__device__ static inline f3 sub3(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ static inline f3 add3(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ static inline f3 neg3(f3 a) { return {-a.x, -a.y, -a.z}; }
// Vector.h:534-537: a.v_Cross(b) is b x a
__device__ static inline f3 vcross(f3 a, f3 b)
{
    return {b.y * a.z - b.z * a.y, b.z * a.x - b.x * a.z, b.x * a.y - b.y * a.x};
}
// Vector.h:389-399
__device__ static inline f3 normalize3(f3 a)
{
    float t = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    if (t != 0) {
        t = 1 / t;
        a.x = a.x * t;
        a.y = a.y * t;
        a.z = a.z * t;
    }
    return a;
}

// Camera::lookFromPatch + GetMatrix + Perspective + P*MV (Camera.cpp:19-52,
// 97-103; Transform.cpp:26-46, 70-80, 127-156; Main.cpp:1171-1183).
__device__ static void face_mvp(const float *fr, int face, float *mvp)
{
    f3 eye = {fr[0], fr[1], fr[2]}, n = {fr[3], fr[4], fr[5]}, pu = {fr[6], fr[7], fr[8]};
    f3 target, up;
    switch (face) {
    case 0: target = pu; up = neg3(n); break;
    case 1: target = neg3(pu); up = n; break;
    case 2: target = neg3(vcross(n, pu)); up = pu; break;
    case 3: target = vcross(n, pu); up = pu; break;
    default: target = n; up = pu; break;
    }
    f3 dir = normalize3(sub3(add3(target, eye), eye));
    f3 right = normalize3(vcross(dir, up));
    f3 up2 = vcross(right, dir);
    float V[4][4];
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
    float t[4] = {-eye.x, -eye.y, -eye.z, 1.0f};
    float c3[4];
    for (int r = 0; r < 4; ++r)
        c3[r] = V[0][r] * t[0] + V[1][r] * t[1] + V[2][r] * t[2] + V[3][r] * t[3];
    for (int r = 0; r < 4; ++r)
        V[3][r] = c3[r];
    float fn = 0.01f, ff = 1000.0f;
    float fh = 1.0f * fn, fw = fh * 1.0f;
    float l = -fw, rt = fw, b = -fh, tp = fh;
    float P[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            P[i][j] = 0.0f;
    P[0][0] = 2 * fn / (rt - l);
    P[2][0] = (rt + l) / (rt - l);
    P[1][1] = 2 * fn / (tp - b);
    P[2][1] = (tp + b) / (tp - b);
    P[2][2] = -(ff + fn) / (ff - fn);
    P[3][2] = -2 * ff * fn / (ff - fn);
    P[2][3] = -1;
    P[3][3] = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            mvp[i * 4 + j] = P[0][j] * V[i][0] + P[1][j] * V[i][1] + P[2][j] * V[i][2] +
                             P[3][j] * V[i][3];
}

// One thread per (slot, face): snapshot dB (Main.cpp:1161) and the MVP.
__global__ void k_prepare(uint32_t K, const uint32_t *__restrict__ shooters,
                          const uint32_t *__restrict__ nshoot, const float *__restrict__ frames,
                          const float *__restrict__ unshot, float *__restrict__ snap,
                          float *__restrict__ mvp)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= K * 5)
        return;
    uint32_t s = t / 5, face = t % 5;
    if (s >= *nshoot)
        return;
    uint32_t p = shooters[s];
    if (face == 0) {
        snap[3 * s + 0] = unshot[3 * (size_t)p + 0];
        snap[3 * s + 1] = unshot[3 * (size_t)p + 1];
        snap[3 * s + 2] = unshot[3 * (size_t)p + 2];
    }
    float fr[9];
    for (int k = 0; k < 9; ++k)
        fr[k] = frames[9 * (size_t)p + k];
    float m[16];
    face_mvp(fr, (int)face, m);
    for (int k = 0; k < 16; ++k)
        mvp[(size_t)t * 16 + k] = m[k];
}

void launch_prepare(const Geometry &g, uint32_t K, const uint32_t *shooters,
                    const uint32_t *nshoot, const float *frames, const float *unshot,
                    float *snap, float *mvp, hipStream_t st)
{
    (void)g;
    uint32_t n = K * 5;
    k_prepare<<<(n + 63) / 64, 64, 0, st>>>(K, shooters, nshoot, frames, unshot, snap, mvp);
}

// --------------------------------------------------------------------- clear
__global__ void k_clear(uint64_t *__restrict__ keys, size_t nkeys, float *__restrict__ ff,
                        uint32_t *__restrict__ hits, size_t nff, uint32_t *__restrict__ counters)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    if (i < kCntN && i != kCntDone)
        counters[i] = 0;
    for (size_t k = i; k < nkeys; k += stride)
        keys[k] = kClearKey;
    for (size_t k = i; k < nff; k += stride) {
        ff[k] = 0.0f;
        if (hits)
            hits[k] = 0u;
    }
}

void launch_clear(const Geometry &g, uint32_t K, uint64_t *keys, float *ff, uint32_t *hits,
                  uint32_t npatch, uint32_t *counters, hipStream_t st)
{
    size_t nkeys = (size_t)K * g.RES, nff = (size_t)K * npatch;
    size_t n = nkeys > nff ? nkeys : nff;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    k_clear<<<(unsigned)blocks, 256, 0, st>>>(keys, nkeys, ff, hits, nff, counters);
}

// ---------------------------------------------------------------- rasterizer
struct CV4 {
    float c[4];
};

struct SV {
    int32_t x, y;
    float z;
};

struct FaceView {
    int32_t vpx, vpy;       // viewport origin inside the hemicube slab (Main.cpp:316-347)
    int32_t sx0, sy0, sx1, sy1; // scissor in face-local pixels, [x0,x1) x [y0,y1)
};

__device__ static inline FaceView face_view(int32_t W, int face)
{
    FaceView v;
    int32_t h2 = W / 2;
    switch (face) {
    case 0: v.vpx = 0; v.vpy = W; v.sx0 = 0; v.sy0 = 0; v.sx1 = W; v.sy1 = h2; break;
    case 1: v.vpx = W; v.vpy = h2; v.sx0 = 0; v.sy0 = h2; v.sx1 = W; v.sy1 = W; break;
    case 2: v.vpx = -h2; v.vpy = 0; v.sx0 = h2; v.sy0 = 0; v.sx1 = W; v.sy1 = W; break;
    case 3: v.vpx = W + h2; v.vpy = 0; v.sx0 = 0; v.sy0 = 0; v.sx1 = h2; v.sy1 = W; break;
    default: v.vpx = h2; v.vpy = 0; v.sx0 = 0; v.sy0 = 0; v.sx1 = W; v.sy1 = W; break;
    }
    return v;
}

__device__ static inline float clip_dist(const CV4 &v, int plane)
{
    switch (plane) {
    case 0: return v.c[2] + v.c[3];
    case 1: return kGuard * v.c[3] - v.c[0];
    case 2: return kGuard * v.c[3] + v.c[0];
    case 3: return kGuard * v.c[3] - v.c[1];
    default: return kGuard * v.c[3] + v.c[1];
    }
}

__device__ static inline SV project(const CV4 &v, float halfW)
{
    float iw = 1.0f / v.c[3];
    float xn = v.c[0] * iw, yn = v.c[1] * iw, zn = v.c[2] * iw;
    float xl = (xn + 1.0f) * halfW, yl = (yn + 1.0f) * halfW;
    SV s;
    s.x = (int32_t)rintf(xl * (float)kSub);
    s.y = (int32_t)rintf(yl * (float)kSub);
    s.z = zn * 0.5f + 0.5f;
    return s;
}

struct TriSetup {
    int32_t A[3], B[3];
    int64_t C[3];
    int32_t bias[3];
    float dzdx, dzdy;
    int32_t ix0, iy0, ix1, iy1;
};

// Edge functions, top-left bias, bbox clamped to the scissor and the depth
// plane.  Returns false when culled (area <= 0) or the bbox is empty.
__device__ static inline bool tri_setup(const SV &a, const SV &b, const SV &c,
                                        const FaceView &fv, TriSetup &ts)
{
    int64_t area2 = (int64_t)(b.x - a.x) * (c.y - a.y) - (int64_t)(c.x - a.x) * (b.y - a.y);
    if (area2 <= 0)
        return false;
    int32_t xmin = min(a.x, min(b.x, c.x)), xmax = max(a.x, max(b.x, c.x));
    int32_t ymin = min(a.y, min(b.y, c.y)), ymax = max(a.y, max(b.y, c.y));
    ts.ix0 = max((xmin - 128 + 255) >> 8, fv.sx0);
    ts.ix1 = min((xmax - 128) >> 8, fv.sx1 - 1);
    ts.iy0 = max((ymin - 128 + 255) >> 8, fv.sy0);
    ts.iy1 = min((ymax - 128) >> 8, fv.sy1 - 1);
    if (ts.ix0 > ts.ix1 || ts.iy0 > ts.iy1)
        return false;
    const SV *v[3] = {&a, &b, &c};
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const SV *p = v[e];
        const SV *q = v[e == 2 ? 0 : e + 1];
        int32_t A = p->y - q->y, B = q->x - p->x;
        ts.A[e] = A;
        ts.B[e] = B;
        ts.C[e] = -((int64_t)A * p->x + (int64_t)B * p->y);
        ts.bias[e] = (A > 0 || (A == 0 && B < 0)) ? 1 : 0;
    }
    double dx1 = (double)(b.x - a.x), dy1 = (double)(b.y - a.y);
    double dx2 = (double)(c.x - a.x), dy2 = (double)(c.y - a.y);
    double dz1 = (double)b.z - (double)a.z, dz2 = (double)c.z - (double)a.z;
    double ar = (double)area2;
    ts.dzdx = (float)((dz1 * dy2 - dz2 * dy1) / ar);
    ts.dzdy = (float)((dx1 * dz2 - dx2 * dz1) / ar);
    return true;
}

__device__ static inline void shade_pixel(int32_t ix, int32_t iy, const int32_t *A,
                                          const int32_t *B, const int64_t *C,
                                          const int32_t *bias, float dzdx, float dzdy,
                                          int32_t x0, int32_t y0, float z0, uint32_t id,
                                          int32_t base, int32_t TW, uint64_t *keys)
{
    int32_t px = ix * kSub + kSub / 2, py = iy * kSub + kSub / 2;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        int64_t E = (int64_t)A[e] * px + (int64_t)B[e] * py + C[e];
        if (E + bias[e] <= 0)
            return;
    }
    float z = fmaf(dzdx, (float)(px - x0), fmaf(dzdy, (float)(py - y0), z0));
    if (!(z > 0.0f))
        z = 0.0f;
    if (!(z < 1.0f))
        return;
    uint64_t key = ((uint64_t)__float_as_uint(z) << 32) | id;
    uint64_t *dst = keys + (base + iy * TW + ix);
    if (key < *dst)
        atomicMin((unsigned long long *)dst, (unsigned long long)key);
}

// Rasterize one snapped triangle: in-lane when its bbox is small, otherwise
// as 8x8 tiles for k_raster_tiles.
__device__ static void raster_tri(const SV &a, const SV &b, const SV &c, uint32_t id,
                                  const FaceView &fv, int32_t base, int32_t TW,
                                  uint64_t *keys, RasterRec *recs, uint32_t rec_cap,
                                  TileItem *items, uint32_t item_cap, uint32_t *counters,
                                  uint32_t small_px)
{
    TriSetup ts;
    if (!tri_setup(a, b, c, fv, ts))
        return;
    uint32_t w = (uint32_t)(ts.ix1 - ts.ix0 + 1), h = (uint32_t)(ts.iy1 - ts.iy0 + 1);
    if (w * h > small_px) {
        uint32_t ntx = (w + kTile - 1) / kTile, nty = (h + kTile - 1) / kTile;
        uint32_t nt = ntx * nty;
        // Items are reserved first and a record only when they fit; every
        // item slot below item_cap that is reserved gets written, either
        // pointing at a written record or marked invalid, so the tile kernel
        // never reads an unwritten record.
        uint32_t it = atomicAdd(&counters[kCntItem], nt);
        bool items_ok = it <= item_cap && nt <= item_cap - it;
        uint32_t r = items_ok ? atomicAdd(&counters[kCntRec], 1u) : 0xFFFFFFFFu;
        bool ok = items_ok && r < rec_cap;
        if (ok) {
            RasterRec rr;
            rr.x[0] = a.x; rr.x[1] = b.x; rr.x[2] = c.x;
            rr.y[0] = a.y; rr.y[1] = b.y; rr.y[2] = c.y;
            rr.dzdx = ts.dzdx;
            rr.dzdy = ts.dzdy;
            rr.z0 = a.z;
            rr.id = id;
            rr.base = base;
            rr.bx0 = (int16_t)ts.ix0; rr.by0 = (int16_t)ts.iy0;
            rr.bx1 = (int16_t)ts.ix1; rr.by1 = (int16_t)ts.iy1;
            recs[r] = rr;
        }
        uint32_t lim = it >= item_cap ? 0u : min(nt, item_cap - it);
        for (uint32_t k = 0; k < lim; ++k) {
            TileItem ti;
            ti.rec = ok ? r : 0xFFFFFFFFu;
            ti.tx = (uint16_t)(k % ntx);
            ti.ty = (uint16_t)(k / ntx);
            items[it + k] = ti;
        }
        if (ok)
            return;
        // Queue full: rasterize this triangle in-lane instead.
        atomicOr(&counters[kCntOverflow], 1u);
    }
    for (int32_t iy = ts.iy0; iy <= ts.iy1; ++iy)
        for (int32_t ix = ts.ix0; ix <= ts.ix1; ++ix)
            shade_pixel(ix, iy, ts.A, ts.B, ts.C, ts.bias, ts.dzdx, ts.dzdy, a.x, a.y, a.z,
                        id, base, TW, keys);
}

// Sutherland-Hodgman against the near plane and the guard band; an
// intersection is interpolated from the inside vertex towards the outside.
__device__ static __noinline__ void raster_clipped(CV4 *poly, float halfW, uint32_t id,
                                                   const FaceView &fv, int32_t base,
                                                   int32_t TW, uint64_t *keys,
                                                   RasterRec *recs, uint32_t rec_cap,
                                                   TileItem *items, uint32_t item_cap,
                                                   uint32_t *counters, uint32_t small_px)
{
    CV4 tmp[12];
    int n = 3;
    for (int plane = 0; plane < 5; ++plane) {
        float d[12];
        int nin = 0;
        for (int i = 0; i < n; ++i) {
            d[i] = clip_dist(poly[i], plane);
            nin += d[i] >= 0.0f;
        }
        if (nin == n)
            continue;
        if (nin == 0)
            return;
        int m = 0;
        for (int i = 0; i < n; ++i) {
            int j = (i + 1 == n) ? 0 : i + 1;
            bool ini = d[i] >= 0.0f, inj = d[j] >= 0.0f;
            if (ini)
                tmp[m++] = poly[i];
            if (ini != inj) {
                const CV4 &A = ini ? poly[i] : poly[j];
                const CV4 &B = ini ? poly[j] : poly[i];
                float da = ini ? d[i] : d[j], db = ini ? d[j] : d[i];
                float t = da / (da - db);
                CV4 r;
                for (int c = 0; c < 4; ++c)
                    r.c[c] = A.c[c] + t * (B.c[c] - A.c[c]);
                tmp[m++] = r;
            }
        }
        n = m;
        for (int i = 0; i < n; ++i)
            poly[i] = tmp[i];
    }
    if (n < 3)
        return;
    SV s[12];
    for (int k = 0; k < n; ++k)
        s[k] = project(poly[k], halfW);
    for (int k = 1; k + 1 < n; ++k)
        raster_tri(s[0], s[k], s[k + 1], id, fv, base, TW, keys, recs, rec_cap, items,
                   item_cap, counters, small_px);
}

// Triangles are stored as 9 planes of ntri floats (v0.x, v0.y, ..., v2.z) so
// consecutive lanes load consecutive addresses.
__global__ __launch_bounds__(256) void k_raster_setup(
    uint32_t W, uint32_t TW, uint32_t RES, const float *__restrict__ tri,
    const uint32_t *__restrict__ tri_patch, uint32_t ntri, const float *__restrict__ mvp,
    const uint32_t *__restrict__ nshoot, uint64_t *__restrict__ keys, RasterRec *recs,
    uint32_t rec_cap, TileItem *items, uint32_t item_cap, uint32_t *counters,
    uint32_t small_px)
{
    uint32_t sf = blockIdx.y;
    uint32_t s = sf / 5, face = sf % 5;
    if (s >= *nshoot)
        return;
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri)
        return;
    const float *m = mvp + (size_t)sf * 16;
    float mm[16];
#pragma unroll
    for (int k = 0; k < 16; ++k)
        mm[k] = m[k];
    CV4 poly[12];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float x = tri[(size_t)(3 * k + 0) * ntri + t];
        float y = tri[(size_t)(3 * k + 1) * ntri + t];
        float z = tri[(size_t)(3 * k + 2) * ntri + t];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            poly[k].c[r] = mm[0 * 4 + r] * x + mm[1 * 4 + r] * y + mm[2 * 4 + r] * z +
                           mm[3 * 4 + r];
    }
    int32_t Wi = (int32_t)W;
    FaceView fv = face_view(Wi, (int)face);
    // Trivial rejects; each is conservative by at least one pixel so it never
    // removes a fragment the exact test would keep.
    uint32_t all_out = 0x1f, any_out = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint32_t o = 0;
#pragma unroll
        for (int p = 0; p < 5; ++p)
            o |= (clip_dist(poly[k], p) >= 0.0f ? 0u : 1u) << p;
        all_out &= o;
        any_out |= o;
    }
    if (all_out)
        return;
    float halfW = 0.5f * (float)W;
    int32_t base = (int32_t)(s * RES) + fv.vpy * (int32_t)TW + fv.vpx;
    uint32_t id = tri_patch[t] + 1u;
    if (any_out == 0) {
        // All vertices in front of the near plane (w > 0): reject against the
        // scissor rectangle in clip space with a one-pixel margin.
        float lo_x = 2.0f * (float)(fv.sx0 - 1) / (float)W - 1.0f;
        float hi_x = 2.0f * (float)(fv.sx1 + 1) / (float)W - 1.0f;
        float lo_y = 2.0f * (float)(fv.sy0 - 1) / (float)W - 1.0f;
        float hi_y = 2.0f * (float)(fv.sy1 + 1) / (float)W - 1.0f;
        bool lx = true, hx = true, ly = true, hy = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float w = poly[k].c[3];
            lx = lx && (poly[k].c[0] < lo_x * w);
            hx = hx && (poly[k].c[0] > hi_x * w);
            ly = ly && (poly[k].c[1] < lo_y * w);
            hy = hy && (poly[k].c[1] > hi_y * w);
        }
        if (lx || hx || ly || hy)
            return;
        SV a = project(poly[0], halfW), b = project(poly[1], halfW), c = project(poly[2], halfW);
        raster_tri(a, b, c, id, fv, base, (int32_t)TW, keys, recs, rec_cap, items, item_cap,
                   counters, small_px);
    } else {
        raster_clipped(poly, halfW, id, fv, base, (int32_t)TW, keys, recs, rec_cap, items,
                       item_cap, counters, small_px);
    }
}

void launch_raster_setup(const Geometry &g, uint32_t K, const float *tri,
                         const uint32_t *tri_patch, uint32_t ntri, const float *mvp,
                         const uint32_t *nshoot, uint64_t *keys, RasterRec *recs,
                         uint32_t rec_cap, TileItem *items, uint32_t item_cap,
                         uint32_t *counters, uint32_t small_px, hipStream_t st)
{
    if (ntri == 0)
        return;
    dim3 grid((ntri + 255) / 256, K * 5);
    k_raster_setup<<<grid, 256, 0, st>>>(g.W, g.TW, g.RES, tri, tri_patch, ntri, mvp, nshoot,
                                         keys, recs, rec_cap, items, item_cap, counters,
                                         small_px);
}

// One wavefront per 8x8 tile item; lane l covers pixel (l & 7, l >> 3).
__global__ __launch_bounds__(256) void k_raster_tiles(const RasterRec *__restrict__ recs,
                                                      const TileItem *__restrict__ items,
                                                      const uint32_t *__restrict__ counters,
                                                      uint32_t rec_cap, uint32_t item_cap,
                                                      int32_t TW, uint64_t *__restrict__ keys)
{
    uint32_t nitem = counters[kCntItem];
    if (nitem > item_cap)
        nitem = item_cap;
    uint32_t lane = threadIdx.x & 63;
    uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t it = wave; it < nitem; it += nwaves) {
        TileItem ti = items[it];
        if (ti.rec >= rec_cap)
            continue;
        const RasterRec &r = recs[ti.rec];
        int32_t ix = r.bx0 + ti.tx * kTile + (int32_t)(lane & 7);
        int32_t iy = r.by0 + ti.ty * kTile + (int32_t)(lane >> 3);
        if (ix > r.bx1 || iy > r.by1)
            continue;
        SV a = {r.x[0], r.y[0], r.z0}, b = {r.x[1], r.y[1], 0.f}, c = {r.x[2], r.y[2], 0.f};
        int32_t A[3], B[3], bias[3];
        int64_t C[3];
        const SV *v[3] = {&a, &b, &c};
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const SV *p = v[e];
            const SV *q = v[e == 2 ? 0 : e + 1];
            A[e] = p->y - q->y;
            B[e] = q->x - p->x;
            C[e] = -((int64_t)A[e] * p->x + (int64_t)B[e] * p->y);
            bias[e] = (A[e] > 0 || (A[e] == 0 && B[e] < 0)) ? 1 : 0;
        }
        shade_pixel(ix, iy, A, B, C, bias, r.dzdx, r.dzdy, r.x[0], r.y[0], r.z0, r.id, r.base,
                    TW, keys);
    }
}

void launch_raster_tiles(const RasterRec *recs, const TileItem *items, const uint32_t *counters,
                         uint32_t rec_cap, uint32_t item_cap, int32_t TW, uint64_t *keys,
                         int nblocks, hipStream_t st)
{
    k_raster_tiles<<<nblocks, 256, 0, st>>>(recs, items, counters, rec_cap, item_cap, TW, keys);
}

// ----------------------------------------------------------- ProcessHemicube
// One texel per lane.  A wavefront covers 64 consecutive texels of a row;
// equal (hemicube, id) neighbours form runs (the span walk of
// Kernel_ProcessHemicube.h:36-49), summed with a segmented scan; the last
// lane of each run adds the run's ΔFF sum and pixel count to the patch's
// entry (the host accumulation of Main.cpp:1257-1269 fused in).
__global__ __launch_bounds__(256) void k_reduce(uint32_t RES, const uint32_t *__restrict__ nshoot,
                                                const uint64_t *__restrict__ keys,
                                                const float *__restrict__ dff, uint32_t npatch,
                                                float *__restrict__ ff,
                                                uint32_t *__restrict__ hits,
                                                uint32_t *__restrict__ counters)
{
    size_t total = (size_t)(*nshoot) * RES;
    int lane = threadIdx.x & 63;
    for (size_t base = (size_t)blockIdx.x * blockDim.x; base < total;
         base += (size_t)gridDim.x * blockDim.x) {
        size_t t = base + threadIdx.x;
        uint32_t id = 0, hi = 0, c = 0;
        float f = 0.0f;
        if (t < total) {
            id = (uint32_t)keys[t];
            hi = (uint32_t)(t / RES);
            if (id) {
                f = dff[t - (size_t)hi * RES];
                c = 1;
            }
        }
        uint32_t pid = __shfl_up(id, 1), phi = __shfl_up(hi, 1);
        bool head = lane == 0 || pid != id || phi != hi;
        int ss = head ? lane : 0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            int o = __shfl_up(ss, d);
            if (lane >= d && o > ss)
                ss = o;
        }
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            float fo = __shfl_up(f, d);
            uint32_t co = __shfl_up(c, d);
            if (lane - d >= ss) {
                f += fo;
                c += co;
            }
        }
        uint32_t nid = __shfl_down(id, 1), nhi = __shfl_down(hi, 1);
        bool tail = lane == 63 || nid != id || nhi != hi;
        if (tail && id != 0) {
            uint32_t p = id - 1;
            if (p < npatch) {
                size_t o = (size_t)hi * npatch + p;
                atomicAdd(&ff[o], f);
                if (hits)
                    atomicAdd(&hits[o], c);
            } else {
                // Main.cpp:1259-1262: unknown patch id, counted and skipped
                atomicAdd(&counters[kCntUnknown], 1u);
            }
        }
    }
}

void launch_reduce(const Geometry &g, uint32_t K, const uint32_t *nshoot, const uint64_t *keys,
                   const float *dff, uint32_t npatch, float *ff, uint32_t *hits,
                   uint32_t *counters, hipStream_t st)
{
    size_t total = (size_t)K * g.RES;
    size_t blocks = (total + 255) / 256;
    if (blocks > 16384)
        blocks = 16384;
    k_reduce<<<(unsigned)blocks, 256, 0, st>>>(g.RES, nshoot, keys, dff, npatch, ff, hits,
                                               counters);
}

// ------------------------------------------------------------------ shooting
// Main.cpp:1251-1295 for all receivers at once: one thread per receiver j
// accumulates the contributions of the pass's shooters, then the shooters
// move their snapshot from unshot to shot.
__global__ __launch_bounds__(256) void k_shoot(uint32_t mode, const uint32_t *__restrict__ shooters,
                                               const uint32_t *__restrict__ nshoot,
                                               const float *__restrict__ snap,
                                               const float *__restrict__ ff,
                                               const float *__restrict__ refl,
                                               const float *__restrict__ area, uint32_t npatch,
                                               float *__restrict__ unshot,
                                               float *__restrict__ shot)
{
    __shared__ uint32_t sid[RAD_MAX_K];
    __shared__ float sc[RAD_MAX_K * 4];
    uint32_t n = *nshoot;
    for (uint32_t s = threadIdx.x; s < n; s += blockDim.x) {
        uint32_t i = shooters[s];
        sid[s] = i;
        if (mode == RAD_MODE_REFERENCE) {
            sc[4 * s + 0] = refl[3 * (size_t)i + 0];
            sc[4 * s + 1] = refl[3 * (size_t)i + 1];
            sc[4 * s + 2] = refl[3 * (size_t)i + 2];
        }
        sc[4 * s + 3] = area[i];
    }
    __syncthreads();
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npatch)
        return;
    float u0 = unshot[3 * (size_t)j + 0], u1 = unshot[3 * (size_t)j + 1],
          u2 = unshot[3 * (size_t)j + 2];
    float r0 = 0.f, r1 = 0.f, r2 = 0.f, inv_aj = 0.f;
    if (mode != RAD_MODE_REFERENCE) {
        r0 = refl[3 * (size_t)j + 0];
        r1 = refl[3 * (size_t)j + 1];
        r2 = refl[3 * (size_t)j + 2];
        inv_aj = 1.0f / area[j];
    }
    for (uint32_t s = 0; s < n; ++s) {
        float f = ff[(size_t)s * npatch + j];
        if (f == 0.0f)
            continue;
        float d0 = snap[3 * s + 0], d1 = snap[3 * s + 1], d2 = snap[3 * s + 2];
        if (mode == RAD_MODE_REFERENCE) {
            // Main.cpp:1274: radiosity += dB * F * 0.3 * colour_emitter
            u0 += d0 * f * 0.3f * sc[4 * s + 0];
            u1 += d1 * f * 0.3f * sc[4 * s + 1];
            u2 += d2 * f * 0.3f * sc[4 * s + 2];
        } else {
            float g = f * (sc[4 * s + 3] * inv_aj);
            u0 += r0 * d0 * g;
            u1 += r1 * d1 * g;
            u2 += r2 * d2 * g;
        }
    }
    // Main.cpp:1287-1295
    for (uint32_t s = 0; s < n; ++s) {
        if (sid[s] == j) {
            float d0 = snap[3 * s + 0], d1 = snap[3 * s + 1], d2 = snap[3 * s + 2];
            shot[3 * (size_t)j + 0] += d0;
            shot[3 * (size_t)j + 1] += d1;
            shot[3 * (size_t)j + 2] += d2;
            u0 -= d0;
            u1 -= d1;
            u2 -= d2;
        }
    }
    unshot[3 * (size_t)j + 0] = u0;
    unshot[3 * (size_t)j + 1] = u1;
    unshot[3 * (size_t)j + 2] = u2;
}

void launch_shoot(uint32_t mode, uint32_t K, const uint32_t *shooters, const uint32_t *nshoot,
                  const float *snap, const float *ff, const float *refl, const float *area,
                  uint32_t npatch, float *unshot, float *shot, hipStream_t st)
{
    (void)K;
    k_shoot<<<(npatch + 255) / 256, 256, 0, st>>>(mode, shooters, nshoot, snap, ff, refl, area,
                                                  npatch, unshot, shot);
}

// ----------------------------------------------------------------- selection
// Single-shooter selection, one packed 64-bit key per patch reduced with
// atomicMax into sel[0] (cleared by the previous k_sel_reset):
//  RAD_SELECT_REFERENCE: key |dB|^2 (Vector.h:356-359), ties to the larger
//    index; a zero maximum selects patch 0.  This is what the list rule of
//    ModelContainer.cpp:263-298 yields for count = 1.
//  RAD_SELECT_POWER: key A*(r+g+b), ties to the smaller index.
// The FP64 residual numerator and denominator are accumulated alongside.
__device__ static inline uint64_t sel_key(uint32_t rule, const float *unshot, const float *area,
                                          uint32_t j)
{
    float u0 = unshot[3 * (size_t)j + 0], u1 = unshot[3 * (size_t)j + 1],
          u2 = unshot[3 * (size_t)j + 2];
    if (rule == RAD_SELECT_REFERENCE) {
        float e = u0 * u0 + u1 * u1 + u2 * u2;
        return ((uint64_t)__float_as_uint(e) << 32) | j;
    }
    float e = area[j] * (u0 + u1 + u2);
    if (!(e > 0.0f))
        e = 0.0f;
    return ((uint64_t)__float_as_uint(e) << 32) | (0xFFFFFFFFu - j);
}

__device__ static inline uint32_t sel_decode(uint32_t rule, uint64_t key)
{
    if (rule == RAD_SELECT_REFERENCE)
        return (key >> 32) == 0 ? 0u : (uint32_t)key;
    return 0xFFFFFFFFu - (uint32_t)key;
}

template <typename T> __device__ static inline T wave_sum(T v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1)
        v += __shfl_xor(v, d);
    return v;
}

__device__ static inline uint64_t wave_max_u64(uint64_t v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        uint64_t o = __shfl_xor(v, d);
        v = o > v ? o : v;
    }
    return v;
}

// Block-level reduction of (max key, num, den) for 256 threads; the result
// is valid in thread 0.
__device__ static inline void block_reduce3(uint64_t &best, double &num, double &den)
{
    __shared__ uint64_t sb[4];
    __shared__ double sn[4], sd[4];
    best = wave_max_u64(best);
    num = wave_sum(num);
    den = wave_sum(den);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sb[w] = best;
        sn[w] = num;
        sd[w] = den;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            best = sb[q] > best ? sb[q] : best;
            num += sn[q];
            den += sd[q];
        }
    }
}

// Stage 1: one partial (max key, residual numerator, denominator) per block.
__global__ __launch_bounds__(256) void k_select_top1(uint32_t rule, const float *__restrict__ unshot,
                                                     const float *__restrict__ area,
                                                     const float *__restrict__ emis,
                                                     uint32_t npatch, uint64_t *__restrict__ pkey,
                                                     double *__restrict__ pres)
{
    uint64_t best = 0;
    double num = 0.0, den = 0.0;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < npatch;
         j += gridDim.x * blockDim.x) {
        uint64_t k = sel_key(rule, unshot, area, j);
        best = k > best ? k : best;
        double a = (double)area[j];
        num += a * ((double)unshot[3 * (size_t)j] + (double)unshot[3 * (size_t)j + 1] +
                    (double)unshot[3 * (size_t)j + 2]);
        den += a * ((double)emis[3 * (size_t)j] + (double)emis[3 * (size_t)j + 1] +
                    (double)emis[3 * (size_t)j + 2]);
    }
    block_reduce3(best, num, den);
    if (threadIdx.x == 0) {
        pkey[blockIdx.x] = best;
        pres[2 * blockIdx.x] = num;
        pres[2 * blockIdx.x + 1] = den;
    }
}

// Stage 2 (one block): reduce the partials, decode the shooter.
__global__ __launch_bounds__(256) void k_select_final(uint32_t rule, uint32_t nparts,
                                                      const uint64_t *__restrict__ pkey,
                                                      const double *__restrict__ pres,
                                                      uint32_t *__restrict__ shooters,
                                                      uint32_t *__restrict__ nshoot,
                                                      double *__restrict__ resid)
{
    uint64_t best = 0;
    double num = 0.0, den = 0.0;
    for (uint32_t b = threadIdx.x; b < nparts; b += blockDim.x) {
        best = pkey[b] > best ? pkey[b] : best;
        num += pres[2 * b];
        den += pres[2 * b + 1];
    }
    block_reduce3(best, num, den);
    if (threadIdx.x == 0) {
        if (shooters) {
            shooters[0] = sel_decode(rule, best);
            nshoot[0] = 1;
        }
        resid[0] = num;
        resid[1] = den;
    }
}

void launch_select_top1(uint32_t rule, const float *unshot, const float *area, const float *emis,
                        uint32_t npatch, uint32_t *shooters, uint32_t *nshoot, double *partials,
                        uint64_t *pkeys, uint32_t *counter, double *resid, int nblocks,
                        hipStream_t st)
{
    (void)counter;
    k_select_top1<<<nblocks, 256, 0, st>>>(rule, unshot, area, emis, npatch, pkeys, partials);
    k_select_final<<<1, 256, 0, st>>>(rule, (uint32_t)nblocks, pkeys, partials, shooters, nshoot,
                                      resid);
}

void launch_residual(const float *unshot, const float *area, const float *emis, uint32_t npatch,
                     double *partials, uint32_t *counter, double *resid, int nblocks,
                     hipStream_t st)
{
    (void)counter;
    uint64_t *pkeys = (uint64_t *)(partials + 2 * nblocks);
    k_select_top1<<<nblocks, 256, 0, st>>>(RAD_SELECT_POWER, unshot, area, emis, npatch, pkeys,
                                           partials);
    k_select_final<<<1, 256, 0, st>>>(RAD_SELECT_POWER, (uint32_t)nblocks, pkeys, partials,
                                      nullptr, nullptr, resid);
}

// ------------------------------------------------------------------ helpers
__global__ void k_ids_to_keys(const uint32_t *__restrict__ ids, uint64_t *__restrict__ keys,
                              size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        keys[i] = kClearKey | ids[i];
}

__global__ void k_split_keys(const uint64_t *__restrict__ keys, uint32_t *__restrict__ ids,
                             float *__restrict__ depth, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        uint64_t k = keys[i];
        if (ids)
            ids[i] = (uint32_t)k;
        if (depth)
            depth[i] = __uint_as_float((uint32_t)(k >> 32));
    }
}

void launch_ids_to_keys(const uint32_t *ids, uint64_t *keys, size_t n, hipStream_t st)
{
    size_t b = (n + 255) / 256;
    k_ids_to_keys<<<(unsigned)(b > 4096 ? 4096 : (b ? b : 1)), 256, 0, st>>>(ids, keys, n);
}

void launch_split_keys(const uint64_t *keys, uint32_t *ids, float *depth, size_t n,
                       hipStream_t st)
{
    size_t b = (n + 255) / 256;
    k_split_keys<<<(unsigned)(b > 4096 ? 4096 : (b ? b : 1)), 256, 0, st>>>(keys, ids, depth, n);
}


// ------------------------------------------------------------ top-K (POWER)
// Each block sorts a chunk of kTopkChunk packed keys descending in LDS
// (bitonic) and keeps its first k; passes repeat over the candidates until
// one block remains, which writes the shooters.  Keys are
// (A*(r+g+b) bits << 32 | ~index): larger power first, ties to the smaller
// index; only keys with positive power are selected.
constexpr int kTopkChunk = 2048;
size_t topk_cand_half(uint32_t npatch, uint32_t k);

__device__ static void bitonic_desc(uint64_t *a, int n)
{
    for (int size = 2; size <= n; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < n / 2; i += blockDim.x) {
                int lo = 2 * i - (i & (stride - 1));
                int hi = lo + stride;
                bool desc = ((lo & size) == 0);
                uint64_t x = a[lo], y = a[hi];
                if ((x < y) == desc) {
                    a[lo] = y;
                    a[hi] = x;
                }
            }
            __syncthreads();
        }
}

__global__ __launch_bounds__(256) void k_topk_pass(const float *__restrict__ unshot,
                                                   const float *__restrict__ area,
                                                   const uint64_t *__restrict__ in, uint32_t nin,
                                                   uint32_t k, uint64_t *__restrict__ out,
                                                   uint32_t *__restrict__ shooters,
                                                   uint32_t *__restrict__ nshoot)
{
    __shared__ uint64_t buf[kTopkChunk];
    uint32_t base = blockIdx.x * kTopkChunk;
    for (int i = threadIdx.x; i < kTopkChunk; i += blockDim.x) {
        uint32_t j = base + i;
        uint64_t key = 0;
        if (j < nin)
            key = in ? in[j] : sel_key(RAD_SELECT_POWER, unshot, area, j);
        buf[i] = key;
    }
    __syncthreads();
    bitonic_desc(buf, kTopkChunk);
    if (shooters) {
        // final pass (one block)
        __shared__ uint32_t cnt;
        if (threadIdx.x == 0)
            cnt = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
            uint64_t key = buf[i];
            bool ok = (key >> 32) != 0;
            shooters[i] = ok ? 0xFFFFFFFFu - (uint32_t)key : 0u;
            if (ok)
                atomicAdd(&cnt, 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0)
            *nshoot = cnt;
    } else {
        for (uint32_t i = threadIdx.x; i < k; i += blockDim.x)
            out[(size_t)blockIdx.x * k + i] = buf[i];
    }
}

void launch_select_topk(uint32_t rule, const float *unshot, const float *area, const float *emis,
                        uint32_t npatch, uint32_t k, uint32_t *shooters, uint32_t *nshoot,
                        uint64_t *cand, double *partials, uint32_t *counter, double *resid,
                        hipStream_t st)
{
    (void)rule;
    (void)emis;
    (void)partials;
    (void)counter;
    (void)resid;
    // cand holds two ping-pong halves of topk_cand_size(npatch, k) entries
    uint32_t n = npatch;
    const uint64_t *in = nullptr;
    uint64_t *bufs[2] = {cand, cand + topk_cand_half(npatch, k)};
    int which = 0;
    while (true) {
        uint32_t nb = (n + kTopkChunk - 1) / kTopkChunk;
        if (nb <= 1) {
            k_topk_pass<<<1, 256, 0, st>>>(unshot, area, in, n, k, nullptr, shooters, nshoot);
            return;
        }
        uint64_t *out = bufs[which];
        k_topk_pass<<<nb, 256, 0, st>>>(unshot, area, in, n, k, out, nullptr, nullptr);
        in = out;
        n = nb * k;
        which ^= 1;
    }
}

size_t topk_cand_half(uint32_t npatch, uint32_t k)
{
    size_t nb = ((size_t)npatch + kTopkChunk - 1) / kTopkChunk;
    return nb * k + 1;
}

// ----------------------------------------------- reference rule, K > 1
// ModelContainer.cpp:263-298 simulated exactly by one workgroup.  The
// list's last (smallest) key never decreases during the scan, so a chunk's
// candidates can be prefiltered in parallel against the value at the chunk
// start; the survivors are then inserted one by one in index order by lane 0
// with the list's push_back / stable sort / reverse / truncate semantics:
// every run of equal keys reverses its order and the new element goes
// first among its equals.
__global__ __launch_bounds__(256) void k_select_reference(const float *__restrict__ unshot,
                                                          uint32_t npatch, uint32_t k,
                                                          uint32_t *__restrict__ shooters,
                                                          uint32_t *__restrict__ nshoot)
{
    __shared__ uint32_t tid_[RAD_MAX_K + 1], tmp_id[RAD_MAX_K + 1];
    __shared__ float tkey[RAD_MAX_K + 1], tmp_key[RAD_MAX_K + 1];
    __shared__ uint32_t cand[256];
    __shared__ uint32_t ncand, n;
    if (threadIdx.x == 0)
        n = 0;
    __syncthreads();
    for (uint32_t base = 0; base < npatch; base += blockDim.x) {
        float back = n ? tkey[n - 1] : 0.0f;
        bool empty = n == 0;
        if (threadIdx.x == 0)
            ncand = 0;
        __syncthreads();
        uint32_t j = base + threadIdx.x;
        float e = 0.f;
        bool c = false;
        if (j < npatch) {
            float u0 = unshot[3 * (size_t)j], u1 = unshot[3 * (size_t)j + 1],
                  u2 = unshot[3 * (size_t)j + 2];
            e = u0 * u0 + u1 * u1 + u2 * u2;
            c = (empty && j == base) || (e > 0 && back <= e);
        }
        // compact candidates in index order with a ballot per wave
        uint64_t m = __ballot(c);
        __shared__ uint32_t wcnt[4];
        int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (lane == 0)
            wcnt[w] = __popcll(m);
        __syncthreads();
        uint32_t off = 0;
        for (int q = 0; q < w; ++q)
            off += wcnt[q];
        if (c)
            cand[off + __popcll(m & ((1ull << lane) - 1))] = j;
        if (threadIdx.x == 0)
            ncand = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        __syncthreads();
        if (threadIdx.x == 0) {
            for (uint32_t q = 0; q < ncand; ++q) {
                uint32_t pi = cand[q];
                float u0 = unshot[3 * (size_t)pi], u1 = unshot[3 * (size_t)pi + 1],
                      u2 = unshot[3 * (size_t)pi + 2];
                float ek = u0 * u0 + u1 * u1 + u2 * u2;
                if (!(n == 0 || (ek > 0 && tkey[n - 1] <= ek)))
                    continue;
                // result: groups in descending key order; each existing group
                // reversed; the new element first in its group
                uint32_t m2 = 0;
                bool placed = false;
                uint32_t g0 = 0;
                while (g0 < n || !placed) {
                    if (!placed && (g0 >= n || tkey[g0] <= ek)) {
                        // groups with key > ek are done; the new element's group
                        tmp_id[m2] = pi;
                        tmp_key[m2] = ek;
                        ++m2;
                        placed = true;
                        continue;
                    }
                    uint32_t g1 = g0;
                    while (g1 < n && tkey[g1] == tkey[g0])
                        ++g1;
                    for (uint32_t q2 = g1; q2 > g0; --q2) {
                        tmp_id[m2] = tid_[q2 - 1];
                        tmp_key[m2] = tkey[q2 - 1];
                        ++m2;
                    }
                    g0 = g1;
                }
                if (m2 > k)
                    m2 = k;
                for (uint32_t q2 = 0; q2 < m2; ++q2) {
                    tid_[q2] = tmp_id[q2];
                    tkey[q2] = tmp_key[q2];
                }
                n = m2;
            }
        }
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x)
        shooters[i] = i < n ? tid_[i] : 0u;
    if (threadIdx.x == 0)
        *nshoot = n;
}

void launch_select_reference(const float *unshot, uint32_t npatch, uint32_t k, uint32_t *shooters,
                             uint32_t *nshoot, hipStream_t st)
{
    k_select_reference<<<1, 256, 0, st>>>(unshot, npatch, k, shooters, nshoot);
}

} // namespace rad
