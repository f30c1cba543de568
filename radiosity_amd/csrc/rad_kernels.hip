// rad_kernels.hip - CDNA4 (gfx950) kernels of the hemicube shooting path.
//
// Compiled with -ffp-contract=off: the camera, clipping, snapping and depth
// arithmetic below is written operation by operation so the item buffer is
// reproducible bit for bit on any IEEE-754 implementation of the same
// sequence (fmaf only where written).  Division and sqrt are HIP's
// correctly-rounded defaults.
//
// Kernels (one pass over K shooters):
//   k_select_* / k_topk*   shooter selection (+ FP64 residual partial sums)
//   k_prepare_clear  dB snapshot and the five face MVPs per shooter; bin
//                  counts and pass counters reset, F rows zeroed unless the
//                  previous shoot left them zero
//   k_raster_setup one lane per (triangle, shooter) over the five faces:
//                  cull, transform, clip, snap; records counted into tile bins
//   k_setup_clipped  the few triangle-faces that cross the near plane or the
//                  guard band
//   k_bin_scan / k_bin_fill   bin offsets and record lists
//   k_tile_raster  one workgroup per (shooter, tile): depth test in LDS, then
//                  the tile's item buffer out or ProcessHemicube fused
//   k_reduce       ProcessHemicube over a whole item buffer (wave-level
//                  run-length reduction into form factors and hit counts)
//   k_shoot        energy shooting into unshot RGB, emitter update; zeroes
//                  the F entries it reads
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
// Per shooter the mvp buffer holds kShooterStride floats: five column-major
// MVPs, then the culling frame used by k_raster_setup: eye, unit normal n,
// unit up u (made orthogonal to n) and n x u; then a flag (1 when the frame
// is degenerate or its up is not orthogonal to the normal within 2e-5, which
// turns the conservative culls off for that shooter).
__device__ static void prepare_item(uint32_t t, uint32_t K, const uint32_t *__restrict__ shooters,
                                    const uint32_t *__restrict__ nshoot,
                                    const float *__restrict__ frames,
                                    const float *__restrict__ unshot, float *__restrict__ snap,
                                    float *__restrict__ mvp, const uint32_t *__restrict__ ctl)
{
    if (t >= K * 5 || (ctl && ctl[kCtlStop]))
        return;
    uint32_t s = t / 5, face = t % 5;
    if (s >= *nshoot)
        return;
    uint32_t p = shooters[s];
    float fr[9];
    for (int k = 0; k < 9; ++k)
        fr[k] = frames[9 * (size_t)p + k];
    float *out = mvp + (size_t)s * kShooterStride;
    if (face == 0) {
        snap[3 * s + 0] = unshot[3 * (size_t)p + 0];
        snap[3 * s + 1] = unshot[3 * (size_t)p + 1];
        snap[3 * s + 2] = unshot[3 * (size_t)p + 2];
        f3 n = normalize3({fr[3], fr[4], fr[5]});
        f3 u = {fr[6], fr[7], fr[8]};
        float un = u.x * n.x + u.y * n.y + u.z * n.z;
        u = normalize3({u.x - un * n.x, u.y - un * n.y, u.z - un * n.z});
        f3 v = {n.y * u.z - n.z * u.y, n.z * u.x - n.x * u.z, n.x * u.y - n.y * u.x};
        float cf[12] = {fr[0], fr[1], fr[2], n.x, n.y, n.z, u.x, u.y, u.z, v.x, v.y, v.z};
        for (int k = 0; k < 12; ++k)
            out[80 + k] = cf[k];
        // the culls take the five face axes as n, +-u, +-(n x u); face_mvp
        // aims the side faces along the raw up, so they hold only while up
        // is orthogonal to n
        float ln = sqrtf(fr[3] * fr[3] + fr[4] * fr[4] + fr[5] * fr[5]);
        float lu = sqrtf(fr[6] * fr[6] + fr[7] * fr[7] + fr[8] * fr[8]);
        float cosnu = fabsf(fr[3] * fr[6] + fr[4] * fr[7] + fr[5] * fr[8]) / (ln * lu);
        bool ok = ln > 0.0f && lu > 0.0f && cosnu <= 2e-5f && isfinite(fr[0] + fr[1] + fr[2]);
        out[92] = ok ? 0.0f : 1.0f;
    }
    float m[16];
    face_mvp(fr, (int)face, m);
    for (int k = 0; k < 16; ++k)
        out[face * 16 + k] = m[k];
}

__global__ void k_prepare(uint32_t K, const uint32_t *__restrict__ shooters,
                          const uint32_t *__restrict__ nshoot, const float *__restrict__ frames,
                          const float *__restrict__ unshot, float *__restrict__ snap,
                          float *__restrict__ mvp)
{
    prepare_item(blockIdx.x * blockDim.x + threadIdx.x, K, shooters, nshoot, frames, unshot, snap,
                 mvp, nullptr);
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
// Per pass: bin counts and pass counters reset, F rows (ff non-null) and hit
// counts (hits non-null) to zero.  A kernel, not hipMemsetAsync, so a pass
// can be captured in a graph.
// The previous pass's overflow state is folded into the sticky counter,
// which only the host clears, so a multi-pass call reports any lost pass.
__device__ static void clear_range(size_t i, size_t stride, uint32_t *__restrict__ bin_count,
                                   size_t nbins, float *__restrict__ ff,
                                   uint32_t *__restrict__ hits, size_t nff,
                                   uint32_t *__restrict__ counters, uint32_t ent_cap)
{
    if (i == 0) {
        // every block of pass counters (one per half-pass), the sticky flag
        // in the first
        bool bad = false;
        for (int h = 0; h < kCntBlocks; ++h)
            bad = bad || counters[h * kCntN + kCntOverflow] != 0 ||
                  counters[h * kCntN + kCntEntries] > ent_cap;
        for (int k = 0; k < kCntN * kCntBlocks; ++k)
            if (k != kCntSticky)
                counters[k] = 0;
        if (bad)
            counters[kCntSticky] = 1;
    }
    for (size_t k = i; k < nbins; k += stride)
        bin_count[k] = 0u;
    for (size_t k = i; k < nff; k += stride) {
        if (ff)
            ff[k] = 0.0f;
        if (hits)
            hits[k] = 0u;
    }
}

__global__ void k_clear(uint32_t *__restrict__ bin_count, size_t nbins, float *__restrict__ ff,
                        uint32_t *__restrict__ hits, size_t nff, uint32_t *__restrict__ counters,
                        uint32_t ent_cap)
{
    clear_range((size_t)blockIdx.x * blockDim.x + threadIdx.x, (size_t)gridDim.x * blockDim.x,
                bin_count, nbins, ff, hits, nff, counters, ent_cap);
}

// k_prepare and k_clear of a pass in one launch (the first np workgroups
// prepare, the others clear): both are small, so one launch saves a kernel's
// start-up and drain.
// Pass gate (the loop condition of Main.cpp:1137): once the stop test of a
// previous shoot has fired (ctl[kCtlStop]), the pass selects nobody: nshoot
// is set to 0 for the later kernels of the pass, and the prepare items of
// this launch return at once (they read the flag themselves rather than the
// nshoot this launch writes).
__global__ __launch_bounds__(256) void k_prepare_clear(
    uint32_t np, uint32_t K, const uint32_t *__restrict__ shooters,
    uint32_t *__restrict__ nshoot, const float *__restrict__ frames,
    const float *__restrict__ unshot, float *__restrict__ snap, float *__restrict__ mvp,
    uint32_t *__restrict__ bin_count, size_t nbins, float *__restrict__ ff,
    uint32_t *__restrict__ hits, size_t nff, uint32_t *__restrict__ counters, uint32_t ent_cap,
    const uint32_t *__restrict__ ctl)
{
    if (blockIdx.x < np) {
        prepare_item(blockIdx.x * blockDim.x + threadIdx.x, K, shooters, nshoot, frames, unshot,
                     snap, mvp, ctl);
        return;
    }
    if (ctl && blockIdx.x == np && threadIdx.x == 0 && ctl[kCtlStop])
        *nshoot = 0u;
    clear_range((size_t)(blockIdx.x - np) * blockDim.x + threadIdx.x,
                (size_t)(gridDim.x - np) * blockDim.x, bin_count, nbins, ff, hits, nff, counters,
                ent_cap);
}

void launch_prepare_clear(uint32_t K, const uint32_t *shooters, uint32_t *nshoot,
                          const float *frames, const float *unshot, float *snap, float *mvp,
                          uint32_t *bin_count, uint32_t ntiles, float *ff, uint32_t *hits,
                          uint32_t ff_rows, uint32_t rowlen, uint32_t *counters, uint32_t ent_cap,
                          const uint32_t *ctl, hipStream_t st)
{
    const uint32_t np = (K * 5 + 255) / 256;
    size_t nbins = (size_t)K * ntiles, nff = ff || hits ? (size_t)ff_rows * rowlen : 0;
    size_t n = nbins > nff ? nbins : nff;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    if (blocks == 0)
        blocks = 1;
    k_prepare_clear<<<(unsigned)(np + blocks), 256, 0, st>>>(np, K, shooters, nshoot, frames,
                                                            unshot, snap, mvp, bin_count, nbins,
                                                            ff, hits, nff, counters, ent_cap, ctl);
}

void launch_clear(uint32_t K, uint32_t *bin_count, uint32_t ntiles, float *ff, uint32_t *hits,
                  uint32_t ff_rows, uint32_t rowlen, uint32_t *counters, uint32_t ent_cap,
                  hipStream_t st)
{
    // ff null: the F rows are known to be zero (the last shoot cleared them)
    size_t nbins = (size_t)K * ntiles, nff = ff || hits ? (size_t)ff_rows * rowlen : 0;
    size_t n = nbins > nff ? nbins : nff;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    if (blocks == 0)
        blocks = 1;
    k_clear<<<(unsigned)blocks, 256, 0, st>>>(bin_count, nbins, ff, hits, nff, counters, ent_cap);
}

// ---------------------------------------------------------------- rasterizer
// Sort-middle rasterizer.
//  k_raster_setup  one lane per (triangle, shooter), looping over the five
//                  faces: cull, transform, clip, snap; every surviving
//                  (sub-)triangle becomes a RasterRec and is counted in the
//                  bins of the screen tiles its bbox touches
//  k_bin_scan      exclusive scan of the bin counts
//  k_bin_fill      record indices into their bins
//  k_tile_raster   one workgroup per tile: depth test in LDS with 64-bit
//                  atomicMin, then either writes the tile's ids and depths
//                  (item buffer out) or reduces it into form factors
//                  (ProcessHemicube fused)
struct CV4 {
    float c[4];
};

struct SV {
    int32_t x, y;
    float z;
};

struct FaceView {
    int32_t vpx, vpy;           // viewport origin inside the hemicube slab (Main.cpp:316-347)
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

// Edge functions of a CCW triangle: E_e(p) = A_e p.x + B_e p.y + C_e, inside
// when E_e + bias_e > 0 for all three (top-left rule, y up).
struct Edges {
    int32_t A[3], B[3];
    int64_t C[3];
    int32_t bias[3];
};

__device__ static inline void make_edges(const int32_t *x, const int32_t *y, Edges &e)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int q = k == 2 ? 0 : k + 1;
        int32_t A = y[k] - y[q], B = x[q] - x[k];
        e.A[k] = A;
        e.B[k] = B;
        e.C[k] = -((int64_t)A * x[k] + (int64_t)B * y[k]);
        e.bias[k] = (A > 0 || (A == 0 && B < 0)) ? 1 : 0;
    }
}

// Depth key of face-local pixel (ix, iy) or 0 when the centre is outside
// the triangle or fails the depth range.
__device__ static inline uint64_t frag_key(const Edges &e, const RasterRec &r, int32_t ix,
                                           int32_t iy)
{
    int32_t px = ix * kSub + kSub / 2, py = iy * kSub + kSub / 2;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int64_t E = (int64_t)e.A[k] * px + (int64_t)e.B[k] * py + e.C[k];
        if (E + e.bias[k] <= 0)
            return 0;
    }
    float z = fmaf(r.dzdx, (float)(px - r.x[0]), fmaf(r.dzdy, (float)(py - r.y[0]), r.z0));
    if (!(z > 0.0f))
        z = 0.0f;
    if (!(z < 1.0f))
        return 0;
    return ((uint64_t)__float_as_uint(z) << 32) | r.id;
}

struct BinCtx {
    RasterRec *recs;       // block regions of kRecsPerBlock, then the spill region
    uint64_t *bbox;        // texel bbox of each record slot (x0 | y0 << 16 | x1 << 32 | y1 << 48)
    uint32_t *bin_count;   // K * ntiles
    uint32_t *counters;
    uint32_t *lnrec;       // LDS: records in this block's region
    uint32_t *lcnt;        // LDS: per-tile record counts of this block
    uint32_t region;       // first record of this block's region
    uint32_t spill_base, spill_cap;
    uint2 *clipq;          // (triangle, shooter * 8 + face) jobs that need clipping
    uint32_t clip_cap;
    uint32_t ntx, ntiles, tshift; // tiles per row, per hemicube, log2(tile side)
};

// Emits a snapped triangle (face-local coordinates) as a record into the
// block's region (spill region when full) and counts it in the LDS
// histogram of the tiles its texel bbox touches.
__device__ static void emit_tri(const SV &a, const SV &b, const SV &c, uint32_t id,
                                const FaceView &fv, uint32_t s, int face, const BinCtx &bc)
{
    int64_t area2 = (int64_t)(b.x - a.x) * (c.y - a.y) - (int64_t)(c.x - a.x) * (b.y - a.y);
    if (area2 <= 0)
        return;
    int32_t xmin = min(a.x, min(b.x, c.x)), xmax = max(a.x, max(b.x, c.x));
    int32_t ymin = min(a.y, min(b.y, c.y)), ymax = max(a.y, max(b.y, c.y));
    int32_t ix0 = max((xmin - 128 + 255) >> 8, fv.sx0);
    int32_t ix1 = min((xmax - 128) >> 8, fv.sx1 - 1);
    int32_t iy0 = max((ymin - 128 + 255) >> 8, fv.sy0);
    int32_t iy1 = min((ymax - 128) >> 8, fv.sy1 - 1);
    if (ix0 > ix1 || iy0 > iy1)
        return;
    uint32_t slot = bc.lnrec ? atomicAdd(bc.lnrec, 1u) : kRecsPerBlock;
    uint32_t r;
    bool local = slot < kRecsPerBlock;
    if (local) {
        r = bc.region + slot;
    } else {
        if (bc.lnrec)
            atomicSub(bc.lnrec, 1u);
        uint32_t q = atomicAdd(&bc.counters[kCntSpill], 1u);
        if (q >= bc.spill_cap) {
            atomicOr(&bc.counters[kCntOverflow], 1u);
            return;
        }
        r = bc.spill_base + q;
    }
    double dx1 = (double)(b.x - a.x), dy1 = (double)(b.y - a.y);
    double dx2 = (double)(c.x - a.x), dy2 = (double)(c.y - a.y);
    double dz1 = (double)b.z - (double)a.z, dz2 = (double)c.z - (double)a.z;
    double ar = (double)area2;
    RasterRec rr;
    rr.x[0] = a.x; rr.x[1] = b.x; rr.x[2] = c.x;
    rr.y[0] = a.y; rr.y[1] = b.y; rr.y[2] = c.y;
    rr.dzdx = (float)((dz1 * dy2 - dz2 * dy1) / ar);
    rr.dzdy = (float)((dx1 * dz2 - dx2 * dz1) / ar);
    rr.z0 = a.z;
    rr.id = id;
    rr.sf = s * 8u + (uint32_t)face;
    rr.pad = 0u;
    // bbox in texel coordinates of the hemicube slab
    int32_t tx0 = fv.vpx + ix0, tx1 = fv.vpx + ix1, ty0 = fv.vpy + iy0, ty1 = fv.vpy + iy1;
    bc.recs[r] = rr;
    bc.bbox[r] = (uint64_t)(uint16_t)tx0 | ((uint64_t)(uint16_t)ty0 << 16) |
                 ((uint64_t)(uint16_t)tx1 << 32) | ((uint64_t)(uint16_t)ty1 << 48);
    uint32_t g0 = (uint32_t)tx0 >> bc.tshift, g1 = (uint32_t)tx1 >> bc.tshift;
    uint32_t h0 = (uint32_t)ty0 >> bc.tshift, h1 = (uint32_t)ty1 >> bc.tshift;
    if (local) {
        for (uint32_t gy = h0; gy <= h1; ++gy)
            for (uint32_t gx = g0; gx <= g1; ++gx)
                atomicAdd(&bc.lcnt[gy * bc.ntx + gx], 1u);
    } else {
        uint32_t *cnt = bc.bin_count + (size_t)s * bc.ntiles;
        for (uint32_t gy = h0; gy <= h1; ++gy)
            for (uint32_t gx = g0; gx <= g1; ++gx)
                atomicAdd(&cnt[gy * bc.ntx + gx], 1u);
    }
}

// Clipping against the near plane and the guard band is Sutherland-Hodgman
// (clip_wave, below): planes in clip_dist order, each polygon walked from
// vertex 0, an intersection interpolated from the inside vertex towards the
// outside one.  A convex polygon gains at most one vertex per plane (3 + 5 =
// 8); more sign changes along the outline than two can only come from
// vertices within rounding of a plane, and a polygon that would outgrow
// kClipMax (fewer than two free places before a vertex) is dropped and
// flagged as an overflow.
constexpr int kClipMax = 12;

// Clip-space transform of a triangle: MVP (column-major) times (x, y, z, 1),
// each row summed left to right.
__device__ static inline void transform3(const float *mm, const float (&vx)[3][3], CV4 *poly)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float x = vx[k][0], y = vx[k][1], z = vx[k][2];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            poly[k].c[r] = mm[0 * 4 + r] * x + mm[1 * 4 + r] * y + mm[2 * 4 + r] * z +
                           mm[3 * 4 + r];
    }
}

// Exact per-face pipeline of one triangle up to the record.  Triangles that
// need clipping (near plane or guard band) are queued for k_setup_clipped.
// Scissor rectangle of a face in clip space (x / w, y / w) with a one-pixel
// margin: (lo_x, hi_x, lo_y, hi_y).
__device__ static inline float4 face_scissor_clip(int32_t W, int face)
{
    FaceView fv = face_view(W, face);
    return make_float4(2.0f * (float)(fv.sx0 - 1) / (float)W - 1.0f,
                       2.0f * (float)(fv.sx1 + 1) / (float)W - 1.0f,
                       2.0f * (float)(fv.sy0 - 1) / (float)W - 1.0f,
                       2.0f * (float)(fv.sy1 + 1) / (float)W - 1.0f);
}

// mm: the face's MVP, scis: face_scissor_clip of the face.
__device__ static inline void setup_face(const float *mm, float4 scis, const float (&vx)[3][3],
                                         uint32_t W, uint32_t s, int face, uint32_t id,
                                         uint32_t t, const BinCtx &bc)
{
    CV4 poly[3];
    transform3(mm, vx, poly);
    FaceView fv = face_view((int32_t)W, face);
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
    if (any_out) {
        uint32_t q = atomicAdd(&bc.counters[kCntClip], 1u);
        if (q < bc.clip_cap)
            bc.clipq[q] = make_uint2(t, s * 8u + (uint32_t)face);
        else
            atomicOr(&bc.counters[kCntOverflow], 1u);
        return;
    }
    float halfW = 0.5f * (float)W;
    // All vertices in front of the near plane (w > 0): reject against the
    // scissor rectangle in clip space with a one-pixel margin.
    const float lo_x = scis.x, hi_x = scis.y, lo_y = scis.z, hi_y = scis.w;
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
    emit_tri(a, b, c, id, fv, s, face, bc);
}

// Sutherland-Hodgman of one queued job by one wavefront: lane i holds
// vertex i of the polygon (at most kClipMax).  Per plane every lane finds
// its edge to the next vertex (the last one wraps to vertex 0), emits its
// vertex when inside and the intersection when the edge crosses, at offsets
// from two ballots, and the new polygon is gathered back through the
// wavefront's LDS buffer.  The arithmetic and the vertex order are those of
// the sequential walk (the oracle's clipper), vertex by vertex, and the
// overflow test is its test before the last vertex, where the count is
// largest.  The fan triangles (0, k, k + 1) are emitted by lanes
// k = 1 .. n - 2 at once.
__device__ static void clip_wave(CV4 v, float halfW, uint32_t id, const FaceView &fv, uint32_t s,
                                 int face, const BinCtx &bc, CV4 *lp)
{
    const int lane = (int)(threadIdx.x & 63);
    const uint64_t lt = (1ull << lane) - 1ull;
    int n = 3;
    for (int plane = 0; plane < 5; ++plane) {
        const bool act = lane < n;
        const float d = act ? clip_dist(v, plane) : 0.0f;
        const bool in = act && d >= 0.0f;
        const int nin = __popcll(__ballot(in));
        if (nin == n)
            continue;
        if (nin == 0)
            return;
        const int nx = lane + 1 >= n ? 0 : lane + 1;
        CV4 q;
#pragma unroll
        for (int c = 0; c < 4; ++c)
            q.c[c] = __shfl(v.c[c], nx, 64);
        const float dq = __shfl(d, nx, 64);
        const bool inq = dq >= 0.0f;
        const bool cross = act && in != inq;
        const uint64_t b1 = __ballot(in), b2 = __ballot(cross);
        const int off = __popcll(b1 & lt) + __popcll(b2 & lt);
        const int m = __popcll(b1) + __popcll(b2);
        const int last = (int)((b1 >> (n - 1)) & 1ull) + (int)((b2 >> (n - 1)) & 1ull);
        if (m - last + 2 > kClipMax) {
            if (lane == 0)
                atomicOr(&bc.counters[kCntOverflow], 1u);
            return;
        }
        if (in)
            lp[off] = v;
        if (cross) {
            const CV4 &A = in ? v : q;
            const CV4 &B = in ? q : v;
            float da = in ? d : dq, db = in ? dq : d;
            float t = da / (da - db);
            CV4 r;
            for (int c = 0; c < 4; ++c)
                r.c[c] = A.c[c] + t * (B.c[c] - A.c[c]);
            lp[off + (in ? 1 : 0)] = r;
        }
        __syncthreads();
        n = m;
        if (lane < n)
            v = lp[lane];
        __syncthreads();
    }
    if (n < 3)
        return;
    const SV pv = project(v, halfW);
    SV a, c;
    a.x = __shfl(pv.x, 0, 64);
    a.y = __shfl(pv.y, 0, 64);
    a.z = __shfl(pv.z, 0, 64);
    const int nx = lane + 1 < 64 ? lane + 1 : lane;
    c.x = __shfl(pv.x, nx, 64);
    c.y = __shfl(pv.y, nx, 64);
    c.z = __shfl(pv.z, nx, 64);
    if (lane >= 1 && lane + 1 < n)
        emit_tri(a, pv, c, id, fv, s, face, bc);
}

// Queued (triangle, shooter, face) jobs that cross the near plane or the
// guard band, one per wavefront (a 64-thread workgroup): lanes 0 .. 2
// transform their vertex as transform3 does, then clip_wave; records go to
// the spill region.
__global__ __launch_bounds__(64) void k_setup_clipped(uint32_t W, const float *__restrict__ tri,
                                                       const uint32_t *__restrict__ tri_patch,
                                                       uint32_t ntri, const float *__restrict__ mvp,
                                                       BinCtx bc,
                                                       const uint32_t *__restrict__ slot_w)
{
    __shared__ CV4 lp[kClipMax];
    const uint32_t lane = threadIdx.x;
    const uint32_t n = min(bc.counters[kCntClip], bc.clip_cap);
    for (uint32_t q = blockIdx.x; q < n; q += gridDim.x) {
        const uint2 job = bc.clipq[q];
        const uint32_t t = job.x, s = job.y >> 3;
        const int face = (int)(job.y & 7u);
        CV4 v;
#pragma unroll
        for (int c = 0; c < 4; ++c)
            v.c[c] = 0.0f;
        if (lane < 3) {
            const float x = tri[(size_t)(3 * lane + 0) * ntri + t];
            const float y = tri[(size_t)(3 * lane + 1) * ntri + t];
            const float z = tri[(size_t)(3 * lane + 2) * ntri + t];
            const float *mm = mvp + (size_t)s * kShooterStride + face * 16;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                v.c[r] = mm[0 * 4 + r] * x + mm[1 * 4 + r] * y + mm[2 * 4 + r] * z + mm[3 * 4 + r];
        }
        const uint32_t Ws = slot_w ? slot_w[s] : W;
        FaceView fv = face_view((int32_t)Ws, face);
        clip_wave(v, 0.5f * (float)Ws, tri_patch[t] + 1u, fv, s, face, bc, lp);
        __syncthreads(); // lp is the next job's
    }
}

__device__ static uint32_t cull_faces(const float *__restrict__ sm, const float (&vx)[3][3],
                                      float4 sph, int f0, int f1);

// Bounding sphere of every triangle, once per scene: the centroid as stored
// and the largest vertex distance from it, inflated by 1e-5 relative so it
// bounds the vertices whatever the rounding of the distances.
__global__ __launch_bounds__(256) void k_tri_sphere(const float *__restrict__ tri, uint32_t ntri,
                                                    float4 *__restrict__ sph)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri)
        return;
    float v[3][3];
    for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c)
            v[k][c] = tri[(size_t)(3 * k + c) * ntri + t];
    float cx = (v[0][0] + v[1][0] + v[2][0]) * (1.0f / 3.0f);
    float cy = (v[0][1] + v[1][1] + v[2][1]) * (1.0f / 3.0f);
    float cz = (v[0][2] + v[1][2] + v[2][2]) * (1.0f / 3.0f);
    float r2 = 0.0f;
    for (int k = 0; k < 3; ++k) {
        float dx = v[k][0] - cx, dy = v[k][1] - cy, dz = v[k][2] - cz;
        r2 = fmaxf(r2, dx * dx + dy * dy + dz * dz);
    }
    sph[t] = make_float4(cx, cy, cz, sqrtf(r2) * 1.00001f);
}

void launch_tri_sphere(const float *tri, uint32_t ntri, float4 *sph, hipStream_t st)
{
    if (ntri)
        k_tri_sphere<<<(ntri + 255) / 256, 256, 0, st>>>(tri, ntri, sph);
}

// One workgroup per (256 triangles, shooter[, face]).  Phase 1: one lane
// per triangle runs the conservative culls (cull_faces) and queues the
// surviving (triangle, face) pairs in LDS; phase 2: the lanes take the queue
// so the exact pipeline runs with every lane busy.  Records go to the block's
// own region (LDS slot counter) and are counted in an LDS tile histogram,
// flushed with one global atomic per touched tile.  Triangles are stored as 9
// planes of ntri floats (v0.x, v0.y, ..., v2.z) so consecutive lanes load
// consecutive addresses.
// In a packed launch slot s draws only its scene's triangles
// [slot_tbeg[s], slot_tbeg[s] + slot_tcnt[s]); blockIdx.x walks that range.
__global__ __launch_bounds__(256) void k_raster_setup(uint32_t W, const float *__restrict__ tri,
                                                      const float4 *__restrict__ tsph,
                                                      const uint32_t *__restrict__ tri_patch,
                                                      uint32_t ntri, const float *__restrict__ mvp,
                                                      const uint32_t *__restrict__ nshoot,
                                                      BinCtx bc, uint32_t *__restrict__ blk_nrec,
                                                      const uint32_t *__restrict__ slot_tbeg,
                                                      const uint32_t *__restrict__ slot_tcnt,
                                                      const uint32_t *__restrict__ slot_w,
                                                      uint32_t group, uint32_t slot0,
                                                      uint32_t nslot)
{
    extern __shared__ uint32_t dyn_lds[];
    uint32_t *lcnt = dyn_lds; // 4 x bc.ntiles per-tile counts (dynamic LDS)
    __shared__ float lv[9][256];
    __shared__ uint32_t lid[256];
    __shared__ uint16_t lq[256 * 5 * 4];
    __shared__ uint32_t lnrec2[4], lqn;
    __shared__ float4 lscis[5];
    __shared__ float lmvp[4][80]; // the current shooters' five face MVPs
    const uint32_t ns = *nshoot < nslot ? *nshoot : nslot;
    // hemicube resolution of the slots being drawn (per slot in a pack of
    // scenes at several resolutions; slots drawn together share a scene)
    uint32_t lw = 0u;
    const uint32_t tl0 = blockIdx.x * blockDim.x;
    // gridDim.z == 5 splits the faces over blocks (small K), else one block
    // walks all five faces
    const int f0 = gridDim.z == 1 ? 0 : (int)blockIdx.z, f1 = gridDim.z == 1 ? 5 : f0 + 1;
    // record region of shooter slot sl
    auto rbof = [&](uint32_t sl) { return (sl * gridDim.z + blockIdx.z) * gridDim.x + blockIdx.x; };
    // The block's triangles stay in registers / LDS while it walks `group`
    // consecutive shooter slots (reloaded only when a packed slot changes
    // scene), so vertices are read once per group instead of once per shooter.
    // Up to four slots that draw the same triangles are taken together: their
    // culls share one queue and one exact phase, which divides the barriers
    // and the per-slot set-up by the slot count; each keeps its own record region, tile counts and
    // record count.
    uint32_t ltb = 0xFFFFFFFFu, ltc = 0;
    float vx[3][3];
    float4 sph = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t g = 0; g < group;) {
        const uint32_t s = (slot0 + blockIdx.y * group) + g;
        if (s >= ns)
            break;
        const uint32_t tb = slot_tbeg ? slot_tbeg[s] : 0u, tc = slot_tcnt ? slot_tcnt[s] : ntri;
        uint32_t nsh = 1;
        while (nsh < 4 && g + nsh < group && s + nsh < ns) {
            const uint32_t tb1 = slot_tbeg ? slot_tbeg[s + nsh] : 0u;
            const uint32_t tc1 = slot_tcnt ? slot_tcnt[s + nsh] : ntri;
            if (tb1 != tb || tc1 != tc)
                break;
            ++nsh;
        }
        if (tl0 >= tc) {
            // beyond the slots' scene: empty regions for the bin fill
            if (threadIdx.x < nsh)
                blk_nrec[rbof(s + threadIdx.x)] = 0u;
            g += nsh;
            continue;
        }
        __syncthreads(); // the previous slots' phase 2 and flush are done with LDS
        const uint32_t Ws = slot_w ? slot_w[s] : W;
        if (Ws != lw) {
            lw = Ws;
            if (threadIdx.x < 5)
                lscis[threadIdx.x] = face_scissor_clip((int32_t)Ws, (int)threadIdx.x);
        }
        if (tb != ltb || tc != ltc) {
            ltb = tb;
            ltc = tc;
            if (tl0 + threadIdx.x < tc) {
                const uint32_t t = tb + tl0 + threadIdx.x;
#pragma unroll
                for (int k = 0; k < 3; ++k)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        vx[k][c] = tri[(size_t)(3 * k + c) * ntri + t];
                        lv[3 * k + c][threadIdx.x] = vx[k][c];
                    }
                lid[threadIdx.x] = tri_patch[t] + 1u;
                sph = tsph[t];
            }
        }
        for (uint32_t i = threadIdx.x; i < nsh * bc.ntiles; i += blockDim.x)
            lcnt[i] = 0u;
        if (threadIdx.x < 4)
            lnrec2[threadIdx.x] = 0u;
        if (threadIdx.x == 0)
            lqn = 0u;
        for (uint32_t i = threadIdx.x; i < 80 * nsh; i += blockDim.x)
            lmvp[i / 80][i % 80] = mvp[(size_t)(s + i / 80) * kShooterStride + i % 80];
        __syncthreads();
        // phase 1: cheap culls; surviving (triangle, slot, face) triples go to
        // an LDS queue
        if (tl0 + threadIdx.x < tc) {
            for (uint32_t sh = 0; sh < nsh; ++sh) {
                uint32_t fm = cull_faces(mvp + (size_t)(s + sh) * kShooterStride, vx, sph, f0, f1);
                while (fm) {
                    int f = __ffs(fm) - 1;
                    fm &= fm - 1;
                    lq[atomicAdd(&lqn, 1u)] =
                        (uint16_t)((threadIdx.x << 5) | (sh << 3) | (uint32_t)f);
                }
            }
        }
        __syncthreads();
        // phase 2: the exact per-face pipeline over the compacted triples
        const uint32_t nq = lqn;
        for (uint32_t q = threadIdx.x; q < nq; q += blockDim.x) {
            uint32_t e = lq[q];
            uint32_t tl = e >> 5, sh = (e >> 3) & 3u;
            int face = (int)(e & 7u);
            float wv[3][3];
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    wv[k][c] = lv[3 * k + c][tl];
            BinCtx b2 = bc;
            b2.region = rbof(s + sh) * kRecsPerBlock;
            b2.lnrec = &lnrec2[sh];
            b2.lcnt = lcnt + sh * bc.ntiles;
            setup_face(lmvp[sh] + face * 16, lscis[face], wv, Ws, s + sh, face, lid[tl],
                       tb + tl0 + tl, b2);
        }
        __syncthreads();
        // flush: one global atomic per touched tile, and the regions' record counts
        for (uint32_t sh = 0; sh < nsh; ++sh) {
            uint32_t *cnt = bc.bin_count + (size_t)(s + sh) * bc.ntiles;
            for (uint32_t i = threadIdx.x; i < bc.ntiles; i += blockDim.x)
                if (lcnt[sh * bc.ntiles + i])
                    atomicAdd(&cnt[i], lcnt[sh * bc.ntiles + i]);
        }
        if (threadIdx.x < nsh)
            blk_nrec[rbof(s + threadIdx.x)] = lnrec2[threadIdx.x];
        g += nsh;
    }
}

// Conservative culls of one triangle for one shooter; returns the mask of
// faces in [f0, f1) the exact pipeline must look at.  The triangle's
// bounding sphere (centre, radius; k_tri_sphere) gives the centre's height
// pa over the shooter's plane and its offsets along the face axes, with a
// margin mq for float rounding:
//  - a sphere wholly below the plane (pa + rho + mq < 0) puts every vertex
//    strictly below it: in the side faces the triangle lies in the
//    scissored-away half, in the front face behind the near plane (0.01);
//  - a sphere straddling the band just above the plane falls back to the
//    per-vertex test: no vertex above the plane by more than
//    tol = min(1e-6 |v - eye|_1, 0.0099) (which also bounds the float error
//    of the height), so in the side faces the triangle lies in the
//    scissored-away half or within tol/|v - eye| * W/2 << 0.5 pixel of the
//    horizon, where the nearest kept pixel centre is half a pixel away, and
//    in the front face it is nearer than the near plane.  This removes the
//    shooter's own plane (floor under a floor patch) before any per-face work;
//  - a face whose frustum (90 degree pyramid around its axis, near 0.01)
//    misses the sphere by more than a margin.
__device__ static uint32_t cull_faces(const float *__restrict__ sm, const float (&vx)[3][3],
                                      float4 sph, int f0, int f1)
{
    if (sm[92] != 0.0f)
        return ((1u << f1) - 1u) & ~((1u << f0) - 1u);
    const float ex = sm[80], ey = sm[81], ez = sm[82];
    const float nx = sm[83], ny = sm[84], nz = sm[85];
    const float ux = sm[86], uy = sm[87], uz = sm[88];
    const float wx = sm[89], wy = sm[90], wz = sm[91];
    const float cx = sph.x - ex, cy = sph.y - ey, cz = sph.z - ez, rho = sph.w;
    const float pa = cx * nx + cy * ny + cz * nz;
    const float pb = cx * ux + cy * uy + cz * uz;
    const float pc = cx * wx + cy * wy + cz * wz;
    const float mq = 1e-4f * (fabsf(pa) + fabsf(pb) + fabsf(pc) + rho) + 1e-6f;
    if (pa + rho + mq < 0.0f)
        return 0u;
    if (pa - rho - mq <= 0.0099f) {
        bool below = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float dx = vx[k][0] - ex, dy = vx[k][1] - ey, dz = vx[k][2] - ez;
            float h = dx * nx + dy * ny + dz * nz;
            float tol = fminf(1e-6f * (fabsf(dx) + fabsf(dy) + fabsf(dz)), 0.0099f);
            below = below && (h < tol);
        }
        if (below)
            return 0u;
    }
    const float m = rho * 1.4143f + mq;
    uint32_t mask = 0u;
    for (int face = f0; face < f1; ++face) {
        float q, s1, s2;
        switch (face) {
        case 0: q = pb; s1 = pa; s2 = pc; break;   // UP    (axis  u)
        case 1: q = -pb; s1 = pa; s2 = pc; break;  // DOWN  (axis -u)
        case 2: q = pc; s1 = pa; s2 = pb; break;   // LEFT  (axis n x u)
        case 3: q = -pc; s1 = pa; s2 = pb; break;  // RIGHT (axis -(n x u))
        default: q = pa; s1 = pb; s2 = pc; break;  // FRONT (axis n)
        }
        if (q + m < fabsf(s1) || q + m < fabsf(s2) || q + rho + m < 0.0099f)
            continue;
        mask |= 1u << face;
    }
    return mask;
}

// Exclusive scan of n bin counts by one block of 1024 threads; writes start
// offsets, resets the fill counters and stores the total in
// counters[kCntEntries].  The counts go through LDS in chunks of
// kScanChunk: every thread issues its kScanPer coalesced loads at once, then
// scans kScanPer contiguous counts from LDS, so a chunk costs one memory
// latency instead of one per count.
// Also orders the bins for the tile kernel by descending size class
// (floor(log2(count)) buckets, a counting sort): the largest bins start
// first, so the heavy horizon tiles overlap the light ones instead of
// forming the tail of the launch.  The arrays given start at bin bin0 of the
// pass (a half-pass scans its own bins); the order holds pass bin numbers.
constexpr uint32_t kScanPer = 8, kScanChunk = 1024 * kScanPer;

__device__ static inline uint32_t size_class(uint32_t c)
{
    return c ? 32u - __clz(c) : 0u;
}

__global__ __launch_bounds__(1024) void k_bin_scan(const uint32_t *__restrict__ cnt,
                                                   uint32_t *__restrict__ start,
                                                   uint32_t *__restrict__ fill, uint32_t n,
                                                   uint32_t *__restrict__ counters,
                                                   uint32_t *__restrict__ order, uint32_t bin0)
{
    __shared__ uint32_t lc[kScanChunk];
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t bucket[33], boff[33];
    __shared__ uint32_t carry;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid < 33)
        bucket[tid] = 0u;
    if (tid == 0)
        carry = 0u;
    __syncthreads();
    for (uint32_t c0 = 0; c0 < n; c0 += kScanChunk) {
        uint32_t v[kScanPer];
#pragma unroll
        for (uint32_t k = 0; k < kScanPer; ++k) {
            const uint32_t i = c0 + k * 1024u + tid;
            v[k] = i < n ? cnt[i] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kScanPer; ++k) {
            const uint32_t i = c0 + k * 1024u + tid;
            lc[k * 1024u + tid] = v[k];
            if (i < n) {
                fill[i] = 0u;
                atomicAdd(&bucket[size_class(v[k])], 1u);
            }
        }
        __syncthreads();
        uint32_t x[kScanPer], sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < kScanPer; ++k) {
            x[k] = lc[tid * kScanPer + k];
            sum += x[k];
        }
        // inclusive scan inside the wave, then over the wave totals
        uint32_t inc = sum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t o = __shfl_up(inc, d, 64);
            if (lane >= (uint32_t)d)
                inc += o;
        }
        if (lane == 63)
            wsum[w] = inc;
        const uint32_t base = carry;
        __syncthreads();
        if (w == 0) {
            uint32_t y = lane < 16 ? wsum[lane] : 0u, z = y;
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) {
                uint32_t o = __shfl_up(z, d, 64);
                if (lane >= (uint32_t)d)
                    z += o;
            }
            if (lane < 16)
                wsum[lane] = z - y; // exclusive prefix of the wave totals
            if (lane == 15)
                carry = base + z;
        }
        __syncthreads();
        uint32_t run = base + wsum[w] + inc - sum;
#pragma unroll
        for (uint32_t k = 0; k < kScanPer; ++k) {
            lc[tid * kScanPer + k] = run;
            run += x[k];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kScanPer; ++k) {
            const uint32_t i = c0 + k * 1024u + tid;
            if (i < n)
                start[i] = lc[k * 1024u + tid];
        }
        __syncthreads(); // lc and wsum are reused by the next chunk
    }
    if (tid == 0) {
        counters[kCntEntries] = carry;
        uint32_t acc = 0;
        for (int k = 32; k >= 0; --k) {
            boff[k] = acc;
            acc += bucket[k];
        }
    }
    __syncthreads();
    // Each thread ranks a contiguous run of bins, one at a time: consecutive
    // launches of a class then come from bins about n / 1024 apart (tiles far
    // apart on the hemicube).  That ran measurably faster in the tile kernel
    // than neighbouring tiles launched together, and than batching the loads.
    const uint32_t per = (n + blockDim.x - 1) / blockDim.x;
    const uint32_t b0 = tid * per, b1 = min(n, b0 + per);
    for (uint32_t i = b0; i < b1; ++i)
        order[atomicAdd(&boff[size_class(cnt[i])], 1u)] = bin0 + i;
}

// Same grid as k_raster_setup: block (blk, s) puts the records of its region
// into their bins, reserving each bin's range with one global atomic per
// touched tile.  The extra column blk == nblk handles the spill region.
__device__ static inline void unpack_bbox(uint64_t v, uint32_t tshift, uint32_t &g0,
                                          uint32_t &g1, uint32_t &h0, uint32_t &h1)
{
    g0 = (uint32_t)(v & 0xFFFFu) >> tshift;
    h0 = (uint32_t)((v >> 16) & 0xFFFFu) >> tshift;
    g1 = (uint32_t)((v >> 32) & 0xFFFFu) >> tshift;
    h1 = (uint32_t)(v >> 48) >> tshift;
}

// The spill records of shooter slots [s0, s1) into their tiles' lists: one
// global atomic per entry (records go to the spill region only when a setup
// block's region is full, which is rare).
__device__ static void bin_spill(const RasterRec *__restrict__ recs,
                                 const uint64_t *__restrict__ bbox,
                                 const uint32_t *__restrict__ counters, uint32_t spill_base,
                                 uint32_t spill_cap, const uint32_t *__restrict__ start,
                                 uint32_t *__restrict__ fill, uint32_t *__restrict__ list,
                                 uint32_t ent_cap, uint32_t ntx, uint32_t ntiles, uint32_t tshift,
                                 uint32_t s0, uint32_t s1)
{
    const uint32_t nsp = min(counters[kCntSpill], spill_cap);
    for (uint32_t s = s0; s < s1; ++s) {
        uint32_t *fl = fill + (size_t)s * ntiles;
        const uint32_t *stt = start + (size_t)s * ntiles;
        for (uint32_t q = threadIdx.x; q < nsp; q += blockDim.x) {
            const uint32_t r = spill_base + q;
            if ((recs[r].sf >> 3) != s)
                continue;
            uint32_t g0, g1, h0, h1;
            unpack_bbox(bbox[r], tshift, g0, g1, h0, h1);
            for (uint32_t gy = h0; gy <= h1; ++gy)
                for (uint32_t gx = g0; gx <= g1; ++gx) {
                    const uint32_t b = gy * ntx + gx;
                    const uint32_t pos = stt[b] + atomicAdd(&fl[b], 1u);
                    if (pos < ent_cap)
                        list[pos] = r;
                }
        }
    }
}

__global__ __launch_bounds__(256) void k_bin_fill(const RasterRec *__restrict__ recs,
                                                  const uint64_t *__restrict__ bbox,
                                                  const uint32_t *__restrict__ blk_nrec,
                                                  const uint32_t *__restrict__ counters,
                                                  uint32_t nblk, uint32_t spill_base,
                                                  uint32_t spill_cap,
                                                  const uint32_t *__restrict__ start,
                                                  uint32_t *__restrict__ fill,
                                                  uint32_t *__restrict__ list, uint32_t ent_cap,
                                                  uint32_t ntx, uint32_t ntiles, uint32_t tshift,
                                                  const uint32_t *__restrict__ nshoot,
                                                  uint32_t group, uint32_t slot0,
                                                  uint32_t nslot)
{
    extern __shared__ uint32_t dyn_lds[];
    uint32_t *lcnt = dyn_lds, *lbase = dyn_lds + ntiles; // dynamic LDS, 2 x ntiles
    const uint32_t ns = *nshoot < nslot ? *nshoot : nslot;
    if (blockIdx.x == nblk) {
        if (blockIdx.z == 0)
            bin_spill(recs, bbox, counters, spill_base, spill_cap, start, fill, list, ent_cap, ntx,
                      ntiles, tshift, (slot0 + blockIdx.y * group), min(ns, (slot0 + blockIdx.y * group) + group));
        return;
    }
    // The workgroup walks `group` consecutive shooter slots, like the setup.
    // Each lane holds the bboxes of its kFillPer region slots in registers;
    // the next slot's count and bboxes are loaded while the current slot is
    // binned, so a slot costs no load latency of its own.
    constexpr uint32_t kFillPer = kRecsPerBlock / 256;
    static_assert(kRecsPerBlock % 256 == 0, "k_bin_fill runs 256 threads per region");
    uint64_t bv[kFillPer];
    uint32_t nrec = 0;
    auto prefetch = [&](uint32_t s) {
        const uint32_t blk = (s * gridDim.z + blockIdx.z) * nblk + blockIdx.x;
        nrec = blk_nrec[blk];
#pragma unroll
        for (uint32_t k = 0; k < kFillPer; ++k) {
            const uint32_t q = threadIdx.x + 256u * k;
            bv[k] = q < nrec ? bbox[(size_t)blk * kRecsPerBlock + q] : 0ull;
        }
    };
    if ((slot0 + blockIdx.y * group) < ns)
        prefetch((slot0 + blockIdx.y * group));
    for (uint32_t g = 0; g < group; ++g) {
        const uint32_t s = (slot0 + blockIdx.y * group) + g;
        if (s >= ns)
            break;
        uint32_t *fl = fill + (size_t)s * ntiles;
        const uint32_t *stt = start + (size_t)s * ntiles;
        const uint32_t blk = (s * gridDim.z + blockIdx.z) * nblk + blockIdx.x;
        const uint32_t region = blk * kRecsPerBlock, n = nrec;
        uint64_t cur[kFillPer];
#pragma unroll
        for (uint32_t k = 0; k < kFillPer; ++k)
            cur[k] = bv[k];
        if (g + 1 < group && s + 1 < ns)
            prefetch(s + 1);
        if (n == 0)
            continue;
        __syncthreads(); // the previous slot is done with lcnt / lbase
        for (uint32_t i = threadIdx.x; i < ntiles; i += blockDim.x)
            lcnt[i] = 0u;
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kFillPer; ++k) {
            if (threadIdx.x + 256u * k >= n)
                break;
            uint32_t g0, g1, h0, h1;
            unpack_bbox(cur[k], tshift, g0, g1, h0, h1);
            for (uint32_t gy = h0; gy <= h1; ++gy)
                for (uint32_t gx = g0; gx <= g1; ++gx)
                    atomicAdd(&lcnt[gy * ntx + gx], 1u);
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < ntiles; i += blockDim.x) {
            uint32_t c = lcnt[i];
            lbase[i] = c ? stt[i] + atomicAdd(&fl[i], c) : 0u;
            lcnt[i] = 0u;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kFillPer; ++k) {
            const uint32_t q = threadIdx.x + 256u * k;
            if (q >= n)
                break;
            uint32_t g0, g1, h0, h1;
            unpack_bbox(cur[k], tshift, g0, g1, h0, h1);
            for (uint32_t gy = h0; gy <= h1; ++gy)
                for (uint32_t gx = g0; gx <= g1; ++gx) {
                    uint32_t b = gy * ntx + gx;
                    uint32_t pos = lbase[b] + atomicAdd(&lcnt[b], 1u);
                    if (pos < ent_cap)
                        list[pos] = region + q;
                }
        }
    }
}

// Segmented sums over the runs of equal id in a wavefront, on DPP lane
// moves (no LDS round trip).  A run starts where the id differs from the
// previous lane's; ss = the lane of the last run start at or before this
// lane, read off the wavefront's mask of run starts.  The sums are a
// Hillis-Steele scan inside each row of 16 lanes (row_shr 1, 2, 4, 8; a lane
// adds lane - d only if lane - d >= ss, so only its own run), then the
// row_bcast steps carry the sum at lane 15 / 47 into rows 1 / 3 and at lane 31
// into rows 2 and 3, again only into lanes whose run began at or before that
// lane.  Lanes a DPP move cannot read from give 0.  Returns true in the last
// lane of each run with an id (id != 0), which then holds the run's sums of f
// (and of c when with_c).  Every lane of the wavefront must call it.
template <int CTRL, int ROWS> __device__ static inline float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, true));
}
template <int CTRL, int ROWS> __device__ static inline uint32_t dpp_u(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, true);
}

template <int CTRL, int ROWS>
__device__ static inline void seg_step(float &f, uint32_t &c, bool take, bool with_c)
{
    const float fo = dpp_f<CTRL, ROWS>(f);
    f += take ? fo : 0.0f;
    if (with_c) {
        const uint32_t co = dpp_u<CTRL, ROWS>(c);
        c += take ? co : 0u;
    }
}

// The segmented scan itself, for the wavefront's mask hm of segment starts
// (lane 0 must be one): f (and c) become the sums over the lanes from the
// segment's start up to this lane.
__device__ static inline void seg_scan(uint64_t hm, float &f, uint32_t &c, bool with_c)
{
    const int lane = (int)(threadIdx.x & 63);
    const int ss = 63 - __clzll(hm & ((2ull << lane) - 1ull));
    seg_step<0x111, 0xF>(f, c, lane - 1 >= ss, with_c); // row_shr:1
    seg_step<0x112, 0xF>(f, c, lane - 2 >= ss, with_c); // row_shr:2
    seg_step<0x114, 0xF>(f, c, lane - 4 >= ss, with_c); // row_shr:4
    seg_step<0x118, 0xF>(f, c, lane - 8 >= ss, with_c); // row_shr:8
    seg_step<0x142, 0xA>(f, c, (lane & ~15) - 1 >= ss, with_c); // row_bcast:15
    seg_step<0x143, 0xC>(f, c, 31 >= ss, with_c);               // row_bcast:31
}

__device__ static inline bool wave_runs(uint32_t id, float &f, uint32_t &c, bool with_c)
{
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t pid = dpp_u<0x138, 0xF>(id); // wave_shr:1
    const bool head = lane == 0 || pid != id;
    const uint64_t hm = __ballot(head);
    seg_scan(hm, f, c, with_c);
    const bool tail = lane == 63 || ((hm >> (lane + 1)) & 1ull);
    return tail && id != 0;
}

// F row entries are id - 1 - pbase for ids in [pbase + 1, pbase + rowlen]
// (pbase: first patch of the shooter's scene in a packed launch, else 0).
// Deterministic mode: a run's sum in fixed point, units of 2^-52 (exact for
// sums from 2^-29 up, well below 1e-7 relative under that; F < 4096), added
// with 64-bit integer atomics, so the total does not depend on their order.
__device__ static inline unsigned long long fixed52(float f)
{
    return __float2ull_rn(f * 4503599627370496.0f);
}

// Wave-level ProcessHemicube on 64 consecutive texels of one shooter's
// hemicube hi (any grouping of equal ids gives the same sums; consecutive
// texels make runs long): each run's sum is added by its last lane.  FQ: the
// runs go to the fixed-point rows ffq instead of ff.  Every lane of the
// wavefront must call it.
template <bool FQ>
__device__ static inline void reduce_wave(uint32_t id, float f, uint32_t hi, uint32_t pbase,
                                          uint32_t rowlen, float *__restrict__ ff,
                                          unsigned long long *__restrict__ ffq,
                                          uint32_t *__restrict__ hits,
                                          uint32_t *__restrict__ counters)
{
    uint32_t c = id ? 1u : 0u;
    if (wave_runs(id, f, c, hits != nullptr)) {
        const uint32_t p = id - 1 - pbase;
        if (p < rowlen) {
            const size_t o = (size_t)hi * rowlen + p;
            if constexpr (FQ)
                atomicAdd(&ffq[o], fixed52(f));
            else
                atomicAdd(&ff[o], f);
            if (hits)
                atomicAdd(&hits[o], c);
        } else {
            atomicAdd(&counters[kCntUnknown], 1u);
        }
    }
}

constexpr uint32_t kTileThreads = 256; // k_tile_raster's workgroup size
constexpr uint32_t kTileTable = 256;   // its LDS table of F entries (one slot per thread)
static_assert(kTileTable == kTileThreads, "the table is cleared and flushed one slot per thread");

// The tile kernel's F sums go through a per-workgroup LDS table (hk: F-row
// index + 1, 0 = free; hs, hc: the sums of f and of the texel counts; open
// addressing, 16 probes), which is added to the global row once at the end,
// so that a tile adds each patch it sees to the row once rather than once
// per run.  A run that finds no free slot goes straight to the global row.
__device__ static inline void table_add(uint32_t p, float f, uint32_t c, bool with_c,
                                        uint32_t *hk, float *hs, uint32_t *hc,
                                        float *__restrict__ frow, uint32_t *__restrict__ hrow,
                                        uint32_t *__restrict__ fflag, uint32_t fstride)
{
    const uint32_t key = p + 1;
    uint32_t h = (key * 2654435761u) >> 24; // kTileTable = 2^8
    for (int probe = 0; probe < 16; ++probe, h = (h + 1) & (kTileTable - 1)) {
        uint32_t k = hk[h];
        if (k == 0u)
            k = atomicCAS(&hk[h], 0u, key);
        if (k == 0u || k == key) {
            atomicAdd(&hs[h], f);
            if (with_c)
                atomicAdd(&hc[h], c);
            return;
        }
    }
    atomicAdd(&frow[p], f);
    if (with_c)
        atomicAdd(&hrow[p], c);
    if (fflag)
        fflag[(p >> 6) * fstride] = 1u;
}

// One run (id, sums f and c) of a shooter's hemicube: id 0 (background) adds
// nothing; F entries are id - 1 - pbase (pbase: first patch of the shooter's
// scene in a packed launch, else 0); an id beyond the row is counted (once
// per run, Main.cpp:1259-1262) and skipped.
__device__ static inline void run_add(uint32_t id, float f, uint32_t c, uint32_t pbase,
                                      uint32_t rowlen, bool with_c, uint32_t *hk, float *hs,
                                      uint32_t *hc, float *__restrict__ frow,
                                      uint32_t *__restrict__ hrow, uint32_t *__restrict__ counters,
                                      uint32_t *__restrict__ fflag, uint32_t fstride)
{
    if (!id)
        return;
    const uint32_t p = id - 1 - pbase;
    if (p >= rowlen) {
        atomicAdd(&counters[kCntUnknown], 1u);
        return;
    }
    table_add(p, f, c, with_c, hk, hs, hc, frow, hrow, fflag, fstride);
}

// ProcessHemicube on 256 consecutive texels of one shooter's hemicube, 4
// consecutive texels per lane (ids q, ΔFF f, f = 0 where q = 0).  A run of
// equal ids inside a lane is summed in the lane; a run that reaches a lane's
// last texel (its tail run) may go on into the next lanes, so the tail runs
// are summed across lanes by the segmented scan (a segment starts at a lane
// whose tail run starts inside it, or whose first id differs from the
// previous lane's last).  Every run is added once: a lane adds its middle
// runs, its first run unless that is its tail run (plus the scanned sum the
// previous lane carries into it), and its tail run's scanned sum when the
// next lane does not go on with it.  Any grouping of equal ids gives the
// same F sums.  Every lane of the wavefront must call it.
// emit(id, f, c) adds one run.
template <typename Emit>
__device__ static inline void reduce_quad(const uint32_t (&q)[4], const float (&f)[4], bool with_c,
                                          Emit &&emit)
{
    const int lane = (int)(threadIdx.x & 63);
    const bool b1 = q[1] != q[0], b2 = q[2] != q[1], b3 = q[3] != q[2];
    const bool whole = !(b1 || b2 || b3);
    // first run [0, first boundary), tail run [last boundary, 4)
    float fs = f[0];
    uint32_t fc = 1;
    if (!b1) {
        fs += f[1];
        ++fc;
        if (!b2) {
            fs += f[2];
            ++fc;
            if (!b3) {
                fs += f[3];
                ++fc;
            }
        }
    }
    float ts = f[3];
    uint32_t tc = 1;
    if (!b3) {
        ts += f[2];
        ++tc;
        if (!b2) {
            ts += f[1];
            ++tc;
            if (!b1) {
                ts += f[0];
                ++tc;
            }
        }
    }
    // middle runs: [1, 2) or [1, 3), and [2, 3)
    if (b1 && (b2 || b3))
        emit(q[1], b2 ? f[1] : f[1] + f[2], b2 ? 1u : 2u);
    if (b2 && b3)
        emit(q[2], f[2], 1u);
    const uint32_t prev3 = dpp_u<0x138, 0xF>(q[3]); // wave_shr:1: lane - 1's last id
    const uint32_t next0 = dpp_u<0x130, 0xF>(q[0]); // wave_shl:1: lane + 1's first id
    const bool cont_prev = lane > 0 && prev3 == q[0];
    const bool cont_next = lane < 63 && next0 == q[3];
    seg_scan(__ballot(!whole || !cont_prev), ts, tc, with_c);
    const float sp = dpp_f<0x138, 0xF>(ts); // the previous lane's scanned tail sum
    const uint32_t cp = with_c ? dpp_u<0x138, 0xF>(tc) : 0u;
    if (!whole)
        emit(q[0], cont_prev ? fs + sp : fs, cont_prev ? fc + cp : fc);
    if (!cont_next)
        emit(q[3], ts, tc);
}

__device__ static inline void draw_span(const RasterRec &r, const Edges &e, int32_t vpx,
                                        int32_t vpy, int32_t ax0, int32_t ay0, uint32_t w,
                                        uint32_t area, uint32_t p0, uint32_t step, int32_t ox,
                                        int32_t oy, uint32_t tshift, uint64_t *tk)
{
    for (uint32_t p = p0; p < area; p += step) {
        int32_t tx = ax0 + (int32_t)(p % w), ty = ay0 + (int32_t)(p / w);
        uint64_t key = frag_key(e, r, tx - vpx, ty - vpy);
        uint32_t pp = (uint32_t)((ty - oy) << tshift) + (uint32_t)(tx - ox);
        if (key)
            atomicMin((unsigned long long *)&tk[pp], (unsigned long long)key);
    }
}

// The whole wavefront draws one record's pixels [ax0,ax1] x [ay0,ay1]
// (texels, at most one tile) in 8 x 8 blocks, lane = (lane & 7, lane >> 3)
// inside the block.  The edge values start from the exact 64-bit evaluation
// at the lane's first pixel and step by 8 pixels with 64-bit adds, so every
// inside test equals frag_key's; depth is frag_key's expression.
__device__ static inline void draw_wave(const RasterRec &r, const Edges &e, int32_t vpx,
                                        int32_t vpy, int32_t ax0, int32_t ax1, int32_t ay0,
                                        int32_t ay1, int32_t ox, int32_t oy, uint32_t tshift,
                                        uint64_t *tk)
{
    const int32_t lane = (int32_t)(threadIdx.x & 63), lx = lane & 7, ly = lane >> 3;
    const int32_t px0 = (ax0 - vpx + lx) * kSub + kSub / 2, py0 = (ay0 - vpy + ly) * kSub + kSub / 2;
    int64_t er[3], sx[3], sy[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        er[k] = (int64_t)e.A[k] * px0 + (int64_t)e.B[k] * py0 + e.C[k] + e.bias[k];
        sx[k] = (int64_t)e.A[k] * (8 * kSub);
        sy[k] = (int64_t)e.B[k] * (8 * kSub);
    }
    for (int32_t by = 0; ay0 + by <= ay1; by += 8) {
        const int32_t ty = ay0 + by + ly;
        const int32_t py = py0 + by * kSub;
        const float zr = fmaf(r.dzdy, (float)(py - r.y[0]), r.z0);
        int64_t e0 = er[0], e1 = er[1], e2 = er[2];
        for (int32_t bx = 0; ax0 + bx <= ax1; bx += 8) {
            const int32_t tx = ax0 + bx + lx;
            if (tx <= ax1 && ty <= ay1 && e0 > 0 && e1 > 0 && e2 > 0) {
                const int32_t px = px0 + bx * kSub;
                float z = fmaf(r.dzdx, (float)(px - r.x[0]), zr);
                if (!(z > 0.0f))
                    z = 0.0f;
                if (z < 1.0f) {
                    uint64_t key = ((uint64_t)__float_as_uint(z) << 32) | r.id;
                    uint32_t pp = (uint32_t)((ty - oy) << tshift) + (uint32_t)(tx - ox);
                    atomicMin((unsigned long long *)&tk[pp], (unsigned long long)key);
                }
            }
            e0 += sx[0];
            e1 += sx[1];
            e2 += sx[2];
        }
        er[0] += sy[0];
        er[1] += sy[1];
        er[2] += sy[2];
    }
}

// draw_wave for a record whose snapped vertices span at most 96 pixels in x
// and in y (kWave32Span subpixels): every pixel a lane visits, the 8 x 8
// blocks' overshoot past the bbox included, lies within 16 pixels of the
// vertices' box, so |A (px - x_k)| + |B (py - y_k)| < 2 * 2^15 * 2^15 = 2^31
// and the edge functions step in int32 (E + bias - 1 >= 0, exactly the
// 64-bit test).  Depth is frag_key's expression.
constexpr int32_t kWave32Span = 96 * kSub;

__device__ static inline void draw_wave32(const RasterRec &r, int32_t vpx, int32_t vpy,
                                          int32_t ax0, int32_t ax1, int32_t ay0, int32_t ay1,
                                          int32_t ox, int32_t oy, uint32_t tshift, uint64_t *tk)
{
    const int32_t lane = (int32_t)(threadIdx.x & 63), lx = lane & 7, ly = lane >> 3;
    const int32_t px0 = (ax0 - vpx + lx) * kSub + kSub / 2, py0 = (ay0 - vpy + ly) * kSub + kSub / 2;
    int32_t er[3], sx[3], sy[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int q = k == 2 ? 0 : k + 1;
        const int32_t A = r.y[k] - r.y[q], B = r.x[q] - r.x[k];
        const int32_t bias = (A > 0 || (A == 0 && B < 0)) ? 1 : 0;
        er[k] = A * (px0 - r.x[k]) + B * (py0 - r.y[k]) + bias - 1;
        sx[k] = A * (8 * kSub);
        sy[k] = B * (8 * kSub);
    }
    for (int32_t by = 0; ay0 + by <= ay1; by += 8) {
        const int32_t ty = ay0 + by + ly;
        const int32_t py = py0 + by * kSub;
        const float zr = fmaf(r.dzdy, (float)(py - r.y[0]), r.z0);
        int32_t e0 = er[0], e1 = er[1], e2 = er[2];
        for (int32_t bx = 0; ax0 + bx <= ax1; bx += 8) {
            const int32_t tx = ax0 + bx + lx;
            if (tx <= ax1 && ty <= ay1 && (e0 | e1 | e2) >= 0) {
                const int32_t px = px0 + bx * kSub;
                float z = fmaf(r.dzdx, (float)(px - r.x[0]), zr);
                if (!(z > 0.0f))
                    z = 0.0f;
                if (z < 1.0f) {
                    const uint64_t key = ((uint64_t)__float_as_uint(z) << 32) | r.id;
                    const uint32_t pp = (uint32_t)((ty - oy) << tshift) + (uint32_t)(tx - ox);
                    atomicMin((unsigned long long *)&tk[pp], (unsigned long long)key);
                }
            }
            e0 += sx[0];
            e1 += sx[1];
            e2 += sx[2];
        }
        er[0] += sy[0];
        er[1] += sy[1];
        er[2] += sy[2];
    }
}

// One lane draws a record's pixels [ax0,ax1] x [ay0,ay1] (texels).  When the
// snapped vertices span less than 64 pixels every edge value inside the bbox
// fits in int32, so the edge functions are stepped with 32-bit adds; the
// values equal the 64-bit evaluation exactly.  Depth is the same
// fmaf(dzdx, px-x0, fmaf(dzdy, py-y0, z0)) with the inner term per row;
// (float)(px - x0) is stepped by 256.0f, exact because every value is an
// integer below 2^24.  Every bbox pixel issues its ds_min_u64 (a pixel
// outside the triangle or the depth range with the neutral key ~0), which
// ran faster than branching around the atomic.
// Rows ay0 + ph, ay0 + ph + nph, ... of the bbox (row phase ph of nph; the
// caller's other lanes draw the other phases).
__device__ static inline void draw_lane(const RasterRec &r, int32_t vpx, int32_t vpy,
                                        int32_t ax0, int32_t ax1, int32_t ay0, int32_t ay1,
                                        int32_t ox, int32_t oy, uint32_t tshift, uint64_t *tk,
                                        uint32_t ph, uint32_t nph)
{
    int32_t xmin = min(r.x[0], min(r.x[1], r.x[2])), xmax = max(r.x[0], max(r.x[1], r.x[2]));
    int32_t ymin = min(r.y[0], min(r.y[1], r.y[2])), ymax = max(r.y[0], max(r.y[1], r.y[2]));
    if (xmax - xmin >= (64 << 8) || ymax - ymin >= (64 << 8)) {
        if (ph != 0)
            return; // phase 0 draws these (rare) records whole
        Edges e;
        make_edges(r.x, r.y, e);
        uint32_t w = (uint32_t)(ax1 - ax0 + 1), area = w * (uint32_t)(ay1 - ay0 + 1);
        draw_span(r, e, vpx, vpy, ax0, ay0, w, area, 0, 1, ox, oy, tshift, tk);
        return;
    }
    ay0 += (int32_t)ph;
    if (ay0 > ay1)
        return;
    const int32_t px0 = (ax0 - vpx) * kSub + kSub / 2, py0 = (ay0 - vpy) * kSub + kSub / 2;
    int32_t v[3], dA[3], dB[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int q = k == 2 ? 0 : k + 1;
        int32_t A = r.y[k] - r.y[q], B = r.x[q] - r.x[k];
        int32_t bias = (A > 0 || (A == 0 && B < 0)) ? 1 : 0;
        // E(p) = A (p.x - x_k) + B (p.y - y_k); inside iff E + bias - 1 >= 0
        v[k] = A * (px0 - r.x[k]) + B * (py0 - r.y[k]) + bias - 1;
        dA[k] = A * kSub;
        dB[k] = B * kSub * (int32_t)nph;
    }
    int32_t py = py0;
    const uint32_t w = (uint32_t)(ax1 - ax0 + 1), rstride = nph << tshift;
    const float fx0 = (float)(px0 - r.x[0]);
    uint32_t prow = (uint32_t)((ay0 - oy) << tshift) + (uint32_t)(ax0 - ox);
    for (int32_t ty = ay0; ty <= ay1;
         ty += (int32_t)nph, py += kSub * (int32_t)nph, prow += rstride) {
        float zr = fmaf(r.dzdy, (float)(py - r.y[0]), r.z0);
        int32_t e0 = v[0], e1 = v[1], e2 = v[2];
        float fx = fx0;
        unsigned long long *p = (unsigned long long *)(tk + prow), *pend = p + w;
        for (; p != pend; ++p, fx += 256.0f) {
            float z = fmaf(r.dzdx, fx, zr);
            if (!(z > 0.0f))
                z = 0.0f;
            const bool ok = (e0 | e1 | e2) >= 0 && z < 1.0f;
            atomicMin(p, ok ? (unsigned long long)(((uint64_t)__float_as_uint(z) << 32) | r.id)
                            : ~0ull);
            e0 += dA[0];
            e1 += dA[1];
            e2 += dA[2];
        }
        v[0] += dB[0];
        v[1] += dB[1];
        v[2] += dB[2];
    }
}

// One workgroup per bin = (shooter, tile).  The bin's records are taken 256
// at a time, one per lane.  A record covering at most small_px pixels of the
// tile is drawn by its lane with incremental edge functions; larger ones go
// to an LDS queue that the wavefronts then draw one record per wavefront, one
// pixel per lane.  (Sorting the records by size first, so that a wavefront's
// lanes draw records of about the same size, cost more in loads and barriers
// than the lane divergence it saved.)  Depth is resolved in LDS with 64-bit
// atomicMin, so the drawing order does not matter; the atomic returns nothing
// and is issued without first reading the texel, so no fragment waits on an
// LDS read.  Then the tile either goes
// to the item buffer (out_ids != null; out_depth != null adds the depth
// plane) or is reduced into F rows and hit counts (ff or, in the
// deterministic mode, its fixed-point rows ffq != null).

template <bool FQ>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_tile_raster(
    uint32_t W, uint32_t TW, uint32_t TH, uint32_t RES, const uint32_t *__restrict__ nshoot,
    const RasterRec *__restrict__ recs, const uint64_t *__restrict__ bbox,
    const uint32_t *__restrict__ bin_count, const uint32_t *__restrict__ bin_start,
    const uint32_t *__restrict__ list, uint32_t ent_cap, uint32_t ntx, uint32_t ntiles,
    uint32_t tshift, uint32_t small_px, uint32_t *__restrict__ out_ids,
    float *__restrict__ out_depth, const float *__restrict__ dff,
    uint32_t rowlen, const uint32_t *__restrict__ slot_pbeg, float *__restrict__ ff,
    unsigned long long *__restrict__ ffq, uint32_t *__restrict__ hits,
    uint32_t *__restrict__ counters, const uint32_t *__restrict__ order,
    const uint32_t *__restrict__ slot_w, const uint32_t *__restrict__ slot_dffo,
    uint32_t *__restrict__ fblk, uint32_t fstride)
{
    __shared__ uint64_t tk[kMaxTile * kMaxTile];
    // the chunk queues, then (resolve) the table of F entries
    __shared__ union {
        struct {
            uint32_t perm[kTileThreads];
            uint64_t pbb[kTileThreads];
        } q;
        struct {
            uint32_t key[kTileTable];
            float sum[kTileTable];
            uint32_t cnt[kTileTable];
        } t;
    } sh;
    uint32_t *const perm = sh.q.perm;
    uint64_t *const pbb = sh.q.pbb;
    __shared__ uint32_t nbig;
    const uint32_t bin = order[blockIdx.x];
    const uint32_t s = bin / ntiles;
    if (s >= *nshoot)
        return;
    const uint32_t tile = bin - s * ntiles;
    const uint32_t n = bin_count[bin], st0 = bin_start[bin];
    // an empty bin adds nothing to the F rows (its texels are background)
    if (n == 0 && !out_ids)
        return;
    // a slot of a pack may render at its scene's own resolution; tiles are
    // laid out for the largest, so its slab uses the first ones
    if (slot_w) {
        W = slot_w[s];
        TW = 2 * W;
        TH = W + W / 2;
        dff += slot_dffo[s];
    }
    const uint32_t T = 1u << tshift, TT = T * T;
    const int32_t ox = (int32_t)((tile % ntx) << tshift), oy = (int32_t)((tile / ntx) << tshift);
    const int32_t Wi = (int32_t)W;
    const uint32_t me = threadIdx.x;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(me >> 6);
    constexpr uint32_t kNoRec = ~0u;
    // Lane me's record of the chunk starting at entry cb.  A chunk of m <
    // 256 records would leave waves idle: its records are dealt to
    // ceil(m / 64) wave sets (one per lane), and the nph = 4 / sets waves of a
    // set each draw row phase ph of its in-lane records (every nph-th bbox
    // row); a full chunk keeps one record per lane, nph = 1.
    auto deal = [&](uint32_t cb, uint32_t &ph, uint32_t &nph) -> uint32_t {
        const uint32_t m = min(kTileThreads, n - cb);
        const uint32_t sets = (m + 63u) >> 6;
        nph = sets >= 3u ? 1u : 4u / sets;
        ph = wv / sets;
        const uint32_t k = (wv - ph * sets) * 64u + (me & 63u);
        return ph < nph && k < m && st0 + cb + k < ent_cap ? list[st0 + cb + k] : kNoRec;
    };
    // A lane's entry of the first chunk (list index and bbox) is fetched
    // while the tile clears; each later chunk's list entries are fetched
    // during the chunk before it.
    uint32_t ph = 0, nph = 1;
    uint32_t ri = n ? deal(0u, ph, nph) : kNoRec;
    uint64_t bv = 0;
    if (ri != kNoRec)
        bv = bbox[ri];
    for (uint32_t p = threadIdx.x; p < TT; p += blockDim.x)
        tk[p] = kClearKey;
    for (uint32_t c0 = 0; c0 < n; c0 += blockDim.x) {
        if (me == 0)
            nbig = 0u;
        uint32_t ph_n = 0, nph_n = 1;
        const uint32_t rnext = c0 + blockDim.x < n ? deal(c0 + blockDim.x, ph_n, nph_n) : kNoRec;
        __syncthreads();
        if (ri != kNoRec) {
            const int32_t ax0 = max((int32_t)(bv & 0xFFFFu), ox);
            const int32_t ay0 = max((int32_t)((bv >> 16) & 0xFFFFu), oy);
            const int32_t ax1 = min((int32_t)((bv >> 32) & 0xFFFFu), ox + (int32_t)T - 1);
            const int32_t ay1 = min((int32_t)(bv >> 48), oy + (int32_t)T - 1);
            const uint32_t area = (uint32_t)(ax1 - ax0 + 1) * (uint32_t)(ay1 - ay0 + 1);
            if (area <= small_px) {
                const RasterRec r = recs[ri];
                FaceView fv = face_view(Wi, (int)(r.sf & 7u));
                draw_lane(r, fv.vpx, fv.vpy, ax0, ax1, ay0, ay1, ox, oy, tshift, tk, ph, nph);
            } else if (ph == 0) {
                const uint32_t q = atomicAdd(&nbig, 1u);
                perm[q] = ri;
                pbb[q] = bv;
            }
        }
        __syncthreads();
        // large records: the whole wavefront per record in 8 x 8 pixel
        // blocks; wave w takes queue entries w, w + 4, ...
        const uint32_t nall = nbig;
        for (uint32_t q = me >> 6; q < nall; q += blockDim.x >> 6) {
            const RasterRec rb = recs[perm[q]];
            const uint64_t bb = pbb[q];
            const int32_t ax0 = max((int32_t)(bb & 0xFFFFu), ox);
            const int32_t ax1 = min((int32_t)((bb >> 32) & 0xFFFFu), ox + (int32_t)T - 1);
            const int32_t ay0 = max((int32_t)((bb >> 16) & 0xFFFFu), oy);
            const int32_t ay1 = min((int32_t)(bb >> 48), oy + (int32_t)T - 1);
            FaceView fv = face_view(Wi, (int)(rb.sf & 7u));
            const int32_t xspan = max(rb.x[0], max(rb.x[1], rb.x[2])) - min(rb.x[0], min(rb.x[1], rb.x[2]));
            const int32_t yspan = max(rb.y[0], max(rb.y[1], rb.y[2])) - min(rb.y[0], min(rb.y[1], rb.y[2]));
            if (xspan <= kWave32Span && yspan <= kWave32Span) {
                draw_wave32(rb, fv.vpx, fv.vpy, ax0, ax1, ay0, ay1, ox, oy, tshift, tk);
            } else {
                Edges e;
                make_edges(rb.x, rb.y, e);
                draw_wave(rb, e, fv.vpx, fv.vpy, ax0, ax1, ay0, ay1, ox, oy, tshift, tk);
            }
        }
        ri = rnext;
        ph = ph_n;
        nph = nph_n;
        if (ri != kNoRec)
            bv = bbox[ri];
        __syncthreads();
    }
    // Resolve.  Float F rows (the shooting pass) go through the LDS table in
    // the 4-texels-per-lane layout: wavefront w takes texels [256 w, 256 w +
    // 256) of the tile in row-major order (4 waves cover a 32 x 32 tile); TW
    // and the tile side are multiples of 4, so a lane's texels lie in one row
    // and are all inside the hemicube or all outside, and their ΔFF values
    // (16-byte aligned: RES, and a pack slot's table offset, are multiples of
    // 4) are one load.  The item buffer and the fixed-point rows walk the
    // tile one texel per lane.  Every ΔFF load is issued before the last
    // barrier.
    static_assert(kMaxTile * kMaxTile == 4 * kTileThreads, "one table resolve step per tile");
    constexpr uint32_t kPer = kMaxTile * kMaxTile / kTileThreads;
    const bool table = !FQ && ff != nullptr;
    const bool reduce = ff || ffq;
    const uint32_t rp = (threadIdx.x >> 6) * 256u + (threadIdx.x & 63) * 4u;
    const int32_t rx = ox + (int32_t)(rp & (T - 1)), ry = oy + (int32_t)(rp >> tshift);
    const bool rin = rp < TT && rx < (int32_t)TW && ry < (int32_t)TH;
    float4 d4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float fd[kPer];
    if (table) {
        if (rin)
            d4 = *reinterpret_cast<const float4 *>(dff + (uint32_t)ry * TW + (uint32_t)rx);
        // the queues are done: their LDS becomes the table
        sh.t.key[threadIdx.x] = 0u;
        sh.t.sum[threadIdx.x] = 0.0f;
        sh.t.cnt[threadIdx.x] = 0u;
    } else {
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t p = k * kTileThreads + threadIdx.x;
            const int32_t tx = ox + (int32_t)(p & (T - 1)), ty = oy + (int32_t)(p >> tshift);
            fd[k] = reduce && p < TT && tx < (int32_t)TW && ty < (int32_t)TH
                        ? dff[(uint32_t)ty * TW + (uint32_t)tx]
                        : 0.0f;
        }
    }
    const uint32_t pbase = slot_pbeg ? slot_pbeg[s] : 0u;
    __syncthreads();
    if (table) {
        float *const frow = ff + (size_t)s * rowlen;
        uint32_t *const hrow = hits ? hits + (size_t)s * rowlen : nullptr;
        // this slot's column of the F-block flags
        uint32_t *const fflag = fblk ? fblk + s : nullptr;
        if ((threadIdx.x >> 6) * 256u < TT) { // whole wavefronts
            uint32_t q4[4] = {0u, 0u, 0u, 0u};
            float f4[4] = {0.f, 0.f, 0.f, 0.f};
            if (rin) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    q4[j] = (uint32_t)tk[rp + j];
                f4[0] = q4[0] ? d4.x : 0.f;
                f4[1] = q4[1] ? d4.y : 0.f;
                f4[2] = q4[2] ? d4.z : 0.f;
                f4[3] = q4[3] ? d4.w : 0.f;
            }
            const bool with_c = hrow != nullptr;
            reduce_quad(q4, f4, with_c, [&](uint32_t id, float f, uint32_t c) {
                run_add(id, f, c, pbase, rowlen, with_c, sh.t.key, sh.t.sum, sh.t.cnt, frow, hrow,
                        counters, fflag, fstride);
            });
        }
        __syncthreads();
        const uint32_t key = sh.t.key[threadIdx.x];
        if (key) {
            atomicAdd(&frow[key - 1], sh.t.sum[threadIdx.x]);
            if (hrow)
                atomicAdd(&hrow[key - 1], sh.t.cnt[threadIdx.x]);
            if (fflag)
                fflag[((key - 1) >> 6) * fstride] = 1u;
        }
        return;
    }
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        if (k * kTileThreads >= TT)
            break;
        const uint32_t p = k * kTileThreads + threadIdx.x;
        const int32_t tx = ox + (int32_t)(p & (T - 1)), ty = oy + (int32_t)(p >> tshift);
        const bool inside = p < TT && tx < (int32_t)TW && ty < (int32_t)TH;
        const uint64_t key = inside ? tk[p] : kClearKey;
        const uint32_t texel = (uint32_t)ty * TW + (uint32_t)tx;
        if (out_ids && inside) {
            out_ids[(size_t)s * RES + texel] = (uint32_t)key;
            if (out_depth)
                out_depth[(size_t)s * RES + texel] = __uint_as_float((uint32_t)(key >> 32));
        }
        if (reduce) {
            const uint32_t id = (uint32_t)key;
            reduce_wave<FQ>(id, id ? fd[k] : 0.0f, s, pbase, rowlen, ff, ffq, hits, counters);
        }
    }
}

// k_bin_fill for up to 256 * RPT tiles, software-pipelined over the
// workgroup's shooter slots: the list space of slot g is reserved (one
// returning global atomic per touched tile) and, while those atomics are in
// flight, the records of the previous slot are placed; two LDS sets of
// counts and bases alternate between the slots.  Slots without records are
// skipped.  The placement is k_bin_fill's: a record's entries go to its
// tiles' lists at base + rank, rank from an LDS atomic in lane order.
template <int RPT>
__global__ __launch_bounds__(256) void k_bin_fill_pipe(const RasterRec *__restrict__ recs,
                                                       const uint64_t *__restrict__ bbox,
                                                       const uint32_t *__restrict__ blk_nrec,
                                                       const uint32_t *__restrict__ counters,
                                                       uint32_t nblk, uint32_t spill_base,
                                                       uint32_t spill_cap,
                                                       const uint32_t *__restrict__ start,
                                                       uint32_t *__restrict__ fill,
                                                       uint32_t *__restrict__ list,
                                                       uint32_t ent_cap, uint32_t ntx,
                                                       uint32_t ntiles, uint32_t tshift,
                                                       const uint32_t *__restrict__ nshoot,
                                                       uint32_t group, uint32_t slot0,
                                                       uint32_t nslot)
{
    extern __shared__ uint32_t dyn_lds[]; // 4 x ntiles: counts and bases, twice
    constexpr uint32_t kFillPer = kRecsPerBlock / 256;
    const uint32_t tid = threadIdx.x;
    const uint32_t ns = *nshoot < nslot ? *nshoot : nslot;
    const uint32_t sbeg = slot0 + blockIdx.y * group, send = min(ns, sbeg + group);
    if (blockIdx.x == nblk) { // the last column: the spill records
        if (blockIdx.z == 0)
            bin_spill(recs, bbox, counters, spill_base, spill_cap, start, fill, list, ent_cap, ntx,
                      ntiles, tshift, sbeg, send);
        return;
    }
    auto blk_of = [&](uint32_t sl) { return (sl * gridDim.z + blockIdx.z) * nblk + blockIdx.x; };
    uint64_t bv[kFillPer];
    uint32_t nrec = 0;
    auto prefetch = [&](uint32_t sl) {
        const uint32_t blk = blk_of(sl);
        nrec = blk_nrec[blk];
#pragma unroll
        for (uint32_t k = 0; k < kFillPer; ++k) {
            const uint32_t q = tid + 256u * k;
            bv[k] = q < nrec ? bbox[(size_t)blk * kRecsPerBlock + q] : 0ull;
        }
    };
    // the records of one slot into the lists (bases lb, ranks from lc)
    auto place = [&](const uint64_t (&v)[kFillPer], uint32_t n, uint32_t region, uint32_t *lc,
                     const uint32_t *lb) {
#pragma unroll
        for (uint32_t k = 0; k < kFillPer; ++k) {
            const uint32_t q = tid + 256u * k;
            if (q >= n)
                break;
            uint32_t g0, g1, h0, h1;
            unpack_bbox(v[k], tshift, g0, g1, h0, h1);
            for (uint32_t gy = h0; gy <= h1; ++gy)
                for (uint32_t gx = g0; gx <= g1; ++gx) {
                    const uint32_t b = gy * ntx + gx;
                    const uint32_t pos = lb[b] + atomicAdd(&lc[b], 1u);
                    if (pos < ent_cap)
                        list[pos] = region + q;
                }
        }
    };
    uint64_t pv[kFillPer];
    uint32_t pn = 0, pregion = 0, it = 0;
    if (sbeg < send)
        prefetch(sbeg);
    for (uint32_t sl = sbeg; sl < send; ++sl) {
        const uint32_t n = nrec, region = blk_of(sl) * kRecsPerBlock;
        uint64_t cur[kFillPer];
#pragma unroll
        for (uint32_t k = 0; k < kFillPer; ++k)
            cur[k] = bv[k];
        if (sl + 1 < send)
            prefetch(sl + 1);
        if (n == 0)
            continue;
        uint32_t *lc = dyn_lds + (it & 1u) * 2u * ntiles, *lb = lc + ntiles;
        uint32_t *pc = dyn_lds + ((it & 1u) ^ 1u) * 2u * ntiles, *pb = pc + ntiles;
        __syncthreads(); // the slot two back is placed; the previous slot's bases are written
        for (uint32_t i = tid; i < ntiles; i += 256u)
            lc[i] = 0u;
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kFillPer; ++k) {
            if (tid + 256u * k >= n)
                break;
            uint32_t g0, g1, h0, h1;
            unpack_bbox(cur[k], tshift, g0, g1, h0, h1);
            for (uint32_t gy = h0; gy <= h1; ++gy)
                for (uint32_t gx = g0; gx <= g1; ++gx)
                    atomicAdd(&lc[gy * ntx + gx], 1u);
        }
        __syncthreads();
        // reserve this slot's list space ...
        uint32_t *fl = fill + (size_t)sl * ntiles;
        const uint32_t *stt = start + (size_t)sl * ntiles;
        uint32_t rb[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const uint32_t i = tid + 256u * (uint32_t)r;
            const uint32_t c = i < ntiles ? lc[i] : 0u;
            rb[r] = c ? stt[i] + atomicAdd(&fl[i], c) : 0u;
            if (i < ntiles)
                lc[i] = 0u;
        }
        // ... while the previous slot's records are placed
        if (it)
            place(pv, pn, pregion, pc, pb);
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const uint32_t i = tid + 256u * (uint32_t)r;
            if (i < ntiles)
                lb[i] = rb[r];
        }
#pragma unroll
        for (uint32_t k = 0; k < kFillPer; ++k)
            pv[k] = cur[k];
        pn = n;
        pregion = region;
        ++it;
    }
    if (it) {
        __syncthreads();
        uint32_t *lc = dyn_lds + ((it - 1u) & 1u) * 2u * ntiles;
        place(pv, pn, pregion, lc, lc + ntiles);
    }
}

// k_bin_fill for up to 256 tiles per hemicube, all of a workgroup's shooter
// slots (at most 8) at once: the region records of every slot are counted
// per (slot, tile) in LDS, one returning global atomic per touched (slot,
// tile) reserves the list space (all of a thread's issued before any is
// used), and the records are placed.  Three barriers and one atomic round
// trip per workgroup, where walking the slots one by one takes one round
// trip per slot.  A thread keeps the bboxes of its first kFillKeep records
// in registers for the placement and loads the others again.
constexpr uint32_t kFillSlots = 8, kFillTiles = 256, kFillKeep = 4;

__global__ __launch_bounds__(256) void k_bin_fill_all(const RasterRec *__restrict__ recs,
                                                      const uint64_t *__restrict__ bbox,
                                                      const uint32_t *__restrict__ blk_nrec,
                                                      const uint32_t *__restrict__ counters,
                                                      uint32_t nblk, uint32_t spill_base,
                                                      uint32_t spill_cap,
                                                      const uint32_t *__restrict__ start,
                                                      uint32_t *__restrict__ fill,
                                                      uint32_t *__restrict__ list,
                                                      uint32_t ent_cap, uint32_t ntx,
                                                      uint32_t ntiles, uint32_t tshift,
                                                      const uint32_t *__restrict__ nshoot,
                                                      uint32_t group, uint32_t slot0,
                                                      uint32_t nslot)
{
    __shared__ uint32_t lc[kFillSlots * kFillTiles], lb[kFillSlots * kFillTiles];
    __shared__ uint32_t off[kFillSlots + 1];
    const uint32_t tid = threadIdx.x;
    const uint32_t ns = *nshoot < nslot ? *nshoot : nslot;
    const uint32_t sbeg = slot0 + blockIdx.y * group, send = min(ns, sbeg + group);
    if (blockIdx.x == nblk) { // the last column: the spill records
        if (blockIdx.z == 0)
            bin_spill(recs, bbox, counters, spill_base, spill_cap, start, fill, list, ent_cap, ntx,
                      ntiles, tshift, sbeg, send);
        return;
    }
    if (sbeg >= send)
        return;
    const uint32_t nsl = send - sbeg;
    auto region = [&](uint32_t sl) {
        return ((sbeg + sl) * gridDim.z + blockIdx.z) * nblk + blockIdx.x;
    };
    if (tid < 64) {
        // the slots' record counts and their running offsets
        const uint32_t c = tid < nsl ? blk_nrec[region(tid)] : 0u;
        uint32_t inc = c;
#pragma unroll
        for (int d = 1; d < (int)kFillSlots; d <<= 1) {
            const uint32_t o = __shfl_up(inc, d, 64);
            if ((int)tid >= d)
                inc += o;
        }
        if (tid < kFillSlots)
            off[tid + 1] = inc;
        if (tid == 0)
            off[0] = 0u;
    }
    for (uint32_t i = tid; i < nsl * ntiles; i += 256u)
        lc[i] = 0u;
    __syncthreads();
    const uint32_t total = off[nsl];
    // record f of the workgroup: slot sl (off[sl] <= f < off[sl + 1]) and
    // its index q in the slot's region
    auto locate = [&](uint32_t f, uint32_t &sl, uint32_t &q) {
        sl = 0;
        while (sl + 1 < nsl && f >= off[sl + 1])
            ++sl;
        q = f - off[sl];
    };
    uint64_t keep[kFillKeep];
#pragma unroll
    for (uint32_t k = 0; k < kFillKeep; ++k)
        keep[k] = 0ull;
    uint32_t k = 0;
    for (uint32_t f = tid; f < total; f += 256u, ++k) {
        uint32_t sl, q;
        locate(f, sl, q);
        const uint64_t v = bbox[(size_t)region(sl) * kRecsPerBlock + q];
#pragma unroll
        for (uint32_t j = 0; j < kFillKeep; ++j)
            if (j == k)
                keep[j] = v;
        uint32_t g0, g1, h0, h1;
        unpack_bbox(v, tshift, g0, g1, h0, h1);
        for (uint32_t gy = h0; gy <= h1; ++gy)
            for (uint32_t gx = g0; gx <= g1; ++gx)
                atomicAdd(&lc[sl * ntiles + gy * ntx + gx], 1u);
    }
    __syncthreads();
    // reserve: every count of this thread goes out before any result is used
    constexpr uint32_t kPer = kFillSlots * kFillTiles / 256u;
    uint32_t rb[kPer], st[kPer];
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
        const uint32_t i = tid + 256u * r;
        const uint32_t c = i < nsl * ntiles ? lc[i] : 0u;
        const uint32_t bin = sbeg * ntiles + i;
        rb[r] = c ? atomicAdd(&fill[bin], c) : 0u;
        st[r] = c ? start[bin] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
        const uint32_t i = tid + 256u * r;
        if (i < nsl * ntiles) {
            lb[i] = st[r] + rb[r];
            lc[i] = 0u;
        }
    }
    __syncthreads();
    k = 0;
    for (uint32_t f = tid; f < total; f += 256u, ++k) {
        uint32_t sl, q;
        locate(f, sl, q);
        const uint32_t rr = region(sl) * kRecsPerBlock + q;
        uint64_t v = 0ull;
        if (k < kFillKeep) {
#pragma unroll
            for (uint32_t j = 0; j < kFillKeep; ++j)
                if (j == k)
                    v = keep[j];
        } else {
            v = bbox[rr];
        }
        uint32_t g0, g1, h0, h1;
        unpack_bbox(v, tshift, g0, g1, h0, h1);
        for (uint32_t gy = h0; gy <= h1; ++gy)
            for (uint32_t gx = g0; gx <= g1; ++gx) {
                const uint32_t b = sl * ntiles + gy * ntx + gx;
                const uint32_t pos = lb[b] + atomicAdd(&lc[b], 1u);
                if (pos < ent_cap)
                    list[pos] = rr;
            }
    }
}

void launch_raster_setup(const Geometry &g, uint32_t s0, uint32_t s1, uint32_t h,
                         const float *tri, const float4 *tsph, const uint32_t *tri_patch,
                         uint32_t ntri, const float *mvp, const uint32_t *nshoot,
                         const BinBuffers &bb, uint32_t *counters, const uint32_t *slot_tbeg,
                         const uint32_t *slot_tcnt, const uint32_t *slot_w, hipStream_t st)
{
    BinCtx bc;
    bc.recs = bb.recs;
    bc.bbox = bb.bbox;
    bc.bin_count = bb.count;
    bc.counters = counters + h * kCntN;
    bc.lnrec = nullptr;
    bc.lcnt = nullptr;
    bc.region = 0;
    bc.spill_base = bb.spill_base + h * bb.spill_cap;
    bc.spill_cap = bb.spill_cap;
    bc.clipq = bb.clipq + (size_t)h * bb.clip_cap;
    bc.clip_cap = bb.clip_cap;
    bc.ntx = bb.ntx;
    bc.ntiles = bb.ntiles;
    bc.tshift = bb.tshift;
    const uint32_t n = s1 - s0;
    uint32_t group = bb.setup_group ? bb.setup_group : 1u;
    if (group > n)
        group = n;
    dim3 grid(bb.nblk, (n + group - 1) / group, bb.nz);
    k_raster_setup<<<grid, 256, 4 * bb.ntiles * sizeof(uint32_t), st>>>(
        g.W, tri, tsph, tri_patch, ntri, mvp, nshoot, bc, bb.blk_nrec, slot_tbeg, slot_tcnt,
        slot_w, group, s0, s1);
    bc.lnrec = nullptr;
    k_setup_clipped<<<bb.clip_blocks, 64, 0, st>>>(g.W, tri, tri_patch, ntri, mvp, bc, slot_w);
}

void launch_bin(uint32_t s0, uint32_t s1, uint32_t h, const uint32_t *nshoot,
                const BinBuffers &bb, uint32_t *counters, hipStream_t st)
{
    const uint32_t n = s1 - s0, b0 = s0 * bb.ntiles, nb = n * bb.ntiles;
    uint32_t *ctr = counters + h * kCntN, *list = bb.list + (size_t)h * bb.ent_cap;
    const uint32_t spill_base = bb.spill_base + h * bb.spill_cap;
    k_bin_scan<<<1, 1024, 0, st>>>(bb.count + b0, bb.start + b0, bb.fill + b0, nb, ctr,
                                   bb.order + b0, b0);
    uint32_t group = bb.setup_group ? bb.setup_group : 1u;
    if (group > n)
        group = n;
    dim3 grid(bb.nblk + 1, (n + group - 1) / group, bb.nz);
    if (bb.ntiles > 1024) {
        k_bin_fill<<<grid, 256, 2 * bb.ntiles * sizeof(uint32_t), st>>>(
            bb.recs, bb.bbox, bb.blk_nrec, ctr, bb.nblk, spill_base, bb.spill_cap, bb.start,
            bb.fill, list, bb.ent_cap, bb.ntx, bb.ntiles, bb.tshift, nshoot, group, s0, s1);
        return;
    }
    const size_t lds = 4 * bb.ntiles * sizeof(uint32_t);
    if (bb.ntiles <= kFillTiles && group <= kFillSlots)
        k_bin_fill_all<<<grid, 256, 0, st>>>(
            bb.recs, bb.bbox, bb.blk_nrec, ctr, bb.nblk, spill_base, bb.spill_cap, bb.start,
            bb.fill, list, bb.ent_cap, bb.ntx, bb.ntiles, bb.tshift, nshoot, group, s0, s1);
    else if (bb.ntiles <= 256)
        k_bin_fill_pipe<1><<<grid, 256, lds, st>>>(
            bb.recs, bb.bbox, bb.blk_nrec, ctr, bb.nblk, spill_base, bb.spill_cap, bb.start,
            bb.fill, list, bb.ent_cap, bb.ntx, bb.ntiles, bb.tshift, nshoot, group, s0, s1);
    else
        k_bin_fill_pipe<4><<<grid, 256, lds, st>>>(
            bb.recs, bb.bbox, bb.blk_nrec, ctr, bb.nblk, spill_base, bb.spill_cap, bb.start,
            bb.fill, list, bb.ent_cap, bb.ntx, bb.ntiles, bb.tshift, nshoot, group, s0, s1);
}

void launch_tile(const Geometry &g, uint32_t s0, uint32_t s1, uint32_t h, const uint32_t *nshoot,
                 const BinBuffers &bb, uint32_t small_px, uint32_t *out_ids, float *out_depth,
                 const float *dff, uint32_t rowlen, const uint32_t *slot_pbeg, float *ff,
                 unsigned long long *ffq, uint32_t *hits, uint32_t *counters,
                 const uint32_t *slot_w, const uint32_t *slot_dffo, uint32_t *fblk,
                 uint32_t fstride, hipStream_t st)
{
    const uint32_t nb = (s1 - s0) * bb.ntiles;
    const uint32_t *order = bb.order + (size_t)s0 * bb.ntiles;
    const uint32_t *list = bb.list + (size_t)h * bb.ent_cap;
    uint32_t *ctr = counters + h * kCntN;
    if (ffq)
        k_tile_raster<true><<<nb, kTileThreads, 0, st>>>(
            g.W, g.TW, g.TH, g.RES, nshoot, bb.recs, bb.bbox, bb.count, bb.start, list,
            bb.ent_cap, bb.ntx, bb.ntiles, bb.tshift, small_px, out_ids, out_depth, dff, rowlen,
            slot_pbeg, ff, ffq, hits, ctr, order, slot_w, slot_dffo, nullptr, 0u);
    else
        k_tile_raster<false><<<nb, kTileThreads, 0, st>>>(
            g.W, g.TW, g.TH, g.RES, nshoot, bb.recs, bb.bbox, bb.count, bb.start, list,
            bb.ent_cap, bb.ntx, bb.ntiles, bb.tshift, small_px, out_ids, out_depth, dff, rowlen,
            slot_pbeg, ff, ffq, hits, ctr, order, slot_w, slot_dffo, ff ? fblk : nullptr,
            fstride);
}

// Deterministic mode: n fixed-point entries (units of 2^-52) into the float
// rows, which are zero where the entries are nonzero; the fixed-point
// entries are left zero.
__global__ __launch_bounds__(256) void k_fold(unsigned long long *__restrict__ ffq,
                                              float *__restrict__ ff, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const unsigned long long q = ffq[i];
        if (q) {
            ff[i] = (float)((double)q * 0x1p-52);
            ffq[i] = 0ull;
        }
    }
}

void launch_fold(unsigned long long *ffq, float *ff, size_t n, hipStream_t st)
{
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    if (blocks)
        k_fold<<<(unsigned)blocks, 256, 0, st>>>(ffq, ff, n);
}

// ----------------------------------------------------------- ProcessHemicube
// A workgroup takes kReduceChunk consecutive texels of one hemicube's id
// plane (4 bytes per texel), 4 consecutive texels per lane per step.  Equal
// ids in a wavefront's 256 consecutive texels form runs (the span walk of
// Kernel_ProcessHemicube.h:36-49) summed in the lane and across lanes with
// a segmented scan (reduce_quad); each run's
// ΔFF sum and pixel count go into an LDS hash table keyed by id, which also
// merges the runs of the chunk's other rows; the table is then flushed with
// one global atomic per distinct id (the host accumulation of
// Main.cpp:1257-1269 fused in).  A run that finds no free slot adds to the
// global entry directly.  Acc is float, or the 2^-52 fixed point of the
// deterministic mode (LDS and global sums in 64-bit integers).
constexpr uint32_t kReduceChunk = 4096, kReduceSlotsLog2 = 11, kReduceSlots = 1u << kReduceSlotsLog2;
static_assert(kReduceChunk % 1024 == 0, "k_reduce steps of 1024 texels");
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <typename Acc> __device__ static inline Acc reduce_acc(float f);
template <> __device__ inline float reduce_acc<float>(float f) { return f; }
template <> __device__ inline unsigned long long reduce_acc<unsigned long long>(float f)
{
    return fixed52(f);
}

template <typename Acc>
__global__ __launch_bounds__(256) void k_reduce(uint32_t RES, const uint32_t *__restrict__ nshoot,
                                                const uint32_t *__restrict__ ids_in,
                                                const float *__restrict__ dff, uint32_t npatch,
                                                Acc *__restrict__ ff,
                                                uint32_t *__restrict__ hits,
                                                uint32_t *__restrict__ counters)
{
    __shared__ uint32_t hkey[kReduceSlots], hcnt[kReduceSlots];
    __shared__ Acc hsum[kReduceSlots];
    const uint32_t chunks = (RES + kReduceChunk - 1) / kReduceChunk;
    const uint32_t hi = blockIdx.x / chunks, c0 = (blockIdx.x - hi * chunks) * kReduceChunk;
    if (hi >= *nshoot)
        return;
    for (uint32_t q = threadIdx.x; q < kReduceSlots; q += blockDim.x) {
        hkey[q] = 0u;
        hcnt[q] = 0u;
        hsum[q] = Acc(0);
    }
    __syncthreads();
    const uint32_t *ip = ids_in + (size_t)hi * RES;
    const uint32_t c1 = min(RES, c0 + kReduceChunk);
    const bool with_c = hits != nullptr;
    // one run: into the LDS table (keyed by id), else straight to the row;
    // Main.cpp:1259-1262: an unknown patch id is counted (once per run) and
    // skipped
    auto emit = [&](uint32_t id, float f, uint32_t c) {
        if (!id)
            return;
        if (id - 1 >= npatch) {
            atomicAdd(&counters[kCntUnknown], 1u);
            return;
        }
        const Acc a = reduce_acc<Acc>(f);
        uint32_t h = (id * 2654435761u) >> (32 - kReduceSlotsLog2);
        for (int probe = 0; probe < 16; ++probe, h = (h + 1) & (kReduceSlots - 1)) {
            uint32_t k = hkey[h];
            if (k == 0u)
                k = atomicCAS(&hkey[h], 0u, id);
            if (k == 0u || k == id) {
                atomicAdd(&hsum[h], a);
                if (with_c)
                    atomicAdd(&hcnt[h], c);
                return;
            }
        }
        atomicAdd(&ff[(size_t)hi * npatch + (id - 1)], a);
        if (with_c)
            atomicAdd(&hits[(size_t)hi * npatch + (id - 1)], c);
    };
    // A step covers 1024 texels, 4 consecutive ones per lane (RES and the
    // chunk size are multiples of 4, so the 16-byte loads are aligned and a
    // lane's texels are all in the chunk or all out); the chunk's ids and
    // weights are all loaded before the first step is reduced.
    constexpr uint32_t kSteps = kReduceChunk / (4 * 256);
    u32x4 iv[kSteps];
    float4 fv[kSteps];
#pragma unroll
    for (uint32_t j = 0; j < kSteps; ++j) {
        const uint32_t t = c0 + j * 1024u + threadIdx.x * 4u;
        iv[j] = t < c1 ? __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(ip + t))
                       : u32x4{0u, 0u, 0u, 0u};
        fv[j] = t < c1 ? *reinterpret_cast<const float4 *>(dff + t) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (uint32_t j = 0; j < kSteps; ++j) {
        const uint32_t q[4] = {iv[j].x, iv[j].y, iv[j].z, iv[j].w};
        const float f[4] = {q[0] ? fv[j].x : 0.f, q[1] ? fv[j].y : 0.f, q[2] ? fv[j].z : 0.f,
                            q[3] ? fv[j].w : 0.f};
        reduce_quad(q, f, with_c, emit);
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < kReduceSlots; q += blockDim.x) {
        const uint32_t id = hkey[q];
        if (id == 0u)
            continue;
        atomicAdd(&ff[(size_t)hi * npatch + (id - 1)], hsum[q]);
        if (hits)
            atomicAdd(&hits[(size_t)hi * npatch + (id - 1)], hcnt[q]);
    }
}

void launch_reduce(const Geometry &g, uint32_t K, const uint32_t *nshoot, const uint32_t *ids,
                   const float *dff, uint32_t npatch, float *ff, unsigned long long *ffq,
                   uint32_t *hits, uint32_t *counters, hipStream_t st)
{
    const uint32_t chunks = (g.RES + kReduceChunk - 1) / kReduceChunk;
    if (ffq)
        k_reduce<unsigned long long><<<K * chunks, 256, 0, st>>>(g.RES, nshoot, ids, dff, npatch,
                                                                 ffq, hits, counters);
    else
        k_reduce<float><<<K * chunks, 256, 0, st>>>(g.RES, nshoot, ids, dff, npatch, ff, hits,
                                                    counters);
}

// The ProcessHemicube kernel's own output (Kernel_ProcessHemicube.h:23-69):
// one thread per work-item of the reference (row y of the k stacked
// hemicubes, span gx of TW / nx texels, the last TW - nx * span texels of a
// row left out as there), walking its span texel by texel.  Pass 1 counts the
// slots a work-item reserves (4 per allocation block, runs of id 0 skipped);
// an exclusive scan turns them into offsets; pass 2 writes the triplets
// (hemicube, id - 1, float ΔFF sum of the run in texel order) and the zero
// padding of the last block at those offsets, so the stream comes out in
// work-item order.  Slots at or beyond cap are not written.
__global__ __launch_bounds__(256) void k_runs(uint32_t TW, uint32_t TH, uint32_t nx, uint32_t span,
                                              uint32_t nitems, const uint32_t *__restrict__ item,
                                              const float *__restrict__ dff,
                                              uint32_t *__restrict__ slots,
                                              const uint32_t *__restrict__ offs, uint32_t cap,
                                              uint32_t *__restrict__ th, uint32_t *__restrict__ tid,
                                              float *__restrict__ te)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nitems)
        return;
    const uint32_t y = t / nx, gx = t - y * nx;
    uint32_t x0 = min(gx * span, TW - 1);
    const uint32_t x1 = min(x0 + span, TW);
    const uint32_t *row = item + (size_t)y * TW;
    const float *frow = dff + (size_t)(y % TH) * TW;
    // next: the first slot of the work-item's next allocation block
    uint32_t nruns = 0, next = offs ? offs[t] : 0u, w = 0, space = 0;
    while (x0 < x1) {
        const uint32_t id = row[x0];
        float e = 0.0f;
        while (x0 < x1 && row[x0] == id) {
            if (offs)
                e += frow[x0];
            ++x0;
        }
        if (id == 0u)
            continue;
        ++nruns;
        if (!offs)
            continue;
        if (!space) {
            w = next;
            next += 4u;
            space = 3;
        } else {
            ++w;
            --space;
        }
        if (w < cap) {
            th[w] = y / TH;
            tid[w] = id - 1u;
            te[w] = e;
        }
    }
    if (!offs) {
        slots[t] = (nruns + 3u) & ~3u;
        return;
    }
    for (; space > 0; --space) {
        ++w;
        if (w < cap) {
            th[w] = 0u;
            tid[w] = 0u;
            te[w] = 0.0f;
        }
    }
}

// Exclusive scan of n counts by one workgroup of 1024 threads in chunks of
// 1024 x 8 through LDS: out[i] = sum of in[0..i), out[n] = the total.
__global__ __launch_bounds__(1024) void k_scan_excl(const uint32_t *__restrict__ in,
                                                    uint32_t *__restrict__ out, uint32_t n)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0)
        carry = 0u;
    __syncthreads();
    for (uint32_t c0 = 0; c0 < n; c0 += 1024u * 8u) {
        uint32_t x[8], sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
            const uint32_t i = c0 + tid * 8u + k;
            x[k] = i < n ? in[i] : 0u;
            sum += x[k];
        }
        uint32_t inc = sum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(inc, d, 64);
            if (lane >= (uint32_t)d)
                inc += o;
        }
        if (lane == 63)
            wsum[w] = inc;
        const uint32_t base = carry;
        __syncthreads();
        if (w == 0) {
            const uint32_t y = lane < 16 ? wsum[lane] : 0u;
            uint32_t z = y;
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) {
                const uint32_t o = __shfl_up(z, d, 64);
                if (lane >= (uint32_t)d)
                    z += o;
            }
            if (lane < 16)
                wsum[lane] = z - y;
            if (lane == 15)
                carry = base + z;
        }
        __syncthreads();
        uint32_t run = base + wsum[w] + inc - sum;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
            const uint32_t i = c0 + tid * 8u + k;
            if (i < n)
                out[i] = run;
            run += x[k];
        }
        __syncthreads(); // wsum and carry are reused by the next chunk
    }
    if (tid == 0)
        out[n] = carry;
}

void launch_runs_count(const Geometry &g, uint32_t k, uint32_t nx, const uint32_t *item,
                       uint32_t *slots, uint32_t *offs, hipStream_t st)
{
    const uint32_t nitems = k * g.TH * nx;
    k_runs<<<(nitems + 255) / 256, 256, 0, st>>>(g.TW, g.TH, nx, g.TW / nx, nitems, item, nullptr,
                                                 slots, nullptr, 0u, nullptr, nullptr, nullptr);
    k_scan_excl<<<1, 1024, 0, st>>>(slots, offs, nitems);
}

void launch_runs_write(const Geometry &g, uint32_t k, uint32_t nx, const uint32_t *item,
                       const float *dff, const uint32_t *offs, uint32_t cap, uint32_t *th,
                       uint32_t *tid, float *te, hipStream_t st)
{
    const uint32_t nitems = k * g.TH * nx;
    k_runs<<<(nitems + 255) / 256, 256, 0, st>>>(g.TW, g.TH, nx, g.TW / nx, nitems, item, dff,
                                                 nullptr, offs, cap, th, tid, te);
}

// ------------------------------------------------------------------ shooting
// Stop test of Main.cpp:1288-1300, run by the lane of the pass's last
// shooter j with its unshot value u before the emitter update (that is
// lastEnergy: the shooter's radiosity including what it received during the
// pass, Main.cpp:1290): u goes to ctl[kCtlLastE..+2]; when a threshold is
// set and |u| (Vector3f::f_Length in float, compared with the double
// threshold as `lastEnergy.f_Length() < 0.1` does) is below it,
// ctl[kCtlStop] makes the later passes no-ops.
__device__ static inline void stop_test(uint32_t *__restrict__ ctl, float u0, float u1, float u2)
{
    ctl[kCtlLastE + 0] = __float_as_uint(u0);
    ctl[kCtlLastE + 1] = __float_as_uint(u1);
    ctl[kCtlLastE + 2] = __float_as_uint(u2);
    const double th = *(const double *)(ctl + kCtlThresh);
    const float len = sqrtf(u0 * u0 + u1 * u1 + u2 * u2);
    if (th > 0.0 && (double)len < th)
        ctl[kCtlStop] = 1u;
}

// Passes and shooters counted by the first thread of a shoot launch.
__device__ static inline void count_pass(uint32_t *__restrict__ ctl, uint32_t n)
{
    if (n == 0)
        return;
    *(unsigned long long *)(ctl + kCtlPassLo) += 1ull;
    *(unsigned long long *)(ctl + kCtlShotLo) += n;
}

// Lane exchanges of the top-K networks without ds_bpermute: lane i gets the
// value of lane i ^ M, from DPP lane moves (strides 1, 2, 4, 8 inside a row
// of 16 lanes) and the gfx950 row and half-wave swaps (16, 32).
template <int M> __device__ __forceinline__ uint32_t lane_xor(uint32_t x)
{
    const int lane = threadIdx.x & 63;
    const int xi = (int)x;
    if constexpr (M == 1) {
        return (uint32_t)__builtin_amdgcn_update_dpp(xi, xi, 0xB1, 0xF, 0xF, false); // quad_perm 1,0,3,2
    } else if constexpr (M == 2) {
        return (uint32_t)__builtin_amdgcn_update_dpp(xi, xi, 0x4E, 0xF, 0xF, false); // quad_perm 2,3,0,1
    } else if constexpr (M == 4) {
        const int up = __builtin_amdgcn_update_dpp(xi, xi, 0x104, 0xF, 0xF, false); // row_shl:4, lane i + 4
        const int dn = __builtin_amdgcn_update_dpp(xi, xi, 0x114, 0xF, 0xF, false); // row_shr:4, lane i - 4
        return (uint32_t)((lane & 4) ? dn : up);
    } else if constexpr (M == 8) {
        return (uint32_t)__builtin_amdgcn_update_dpp(xi, xi, 0x128, 0xF, 0xF, false); // row_ror:8
    } else if constexpr (M == 16) {
        // odd rows of the first operand swap with even rows of the second:
        // r[0] holds rows (0, 0, 2, 2) of x, r[1] rows (1, 1, 3, 3)
        const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
        return (lane & 16) ? r[0] : r[1];
    } else {
        static_assert(M == 32, "lane_xor strides 1 .. 32");
        // the upper half of the first operand swaps with the lower half of
        // the second: r[0] holds x's lower half twice, r[1] its upper half
        const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
        return (lane & 32) ? r[0] : r[1];
    }
}

template <int M> __device__ __forceinline__ uint64_t lane_xor_u64(uint64_t v)
{
    return ((uint64_t)lane_xor<M>((uint32_t)(v >> 32)) << 32) | lane_xor<M>((uint32_t)v);
}

// The value of lane 63 - i: each row mirrored (row_mirror), then rows
// r and 3 - r exchanged.
__device__ __forceinline__ uint64_t lane_reverse_u64(uint64_t v)
{
    const int lo = __builtin_amdgcn_update_dpp((int)v, (int)v, 0x140, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(v >> 32), (int)(v >> 32), 0x140, 0xF, 0xF, false);
    const uint64_t m = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
    return lane_xor_u64<32>(lane_xor_u64<16>(m));
}

// A wave-uniform lane's value (lane src, a constant).
__device__ __forceinline__ uint64_t lane_read_u64(uint64_t v, int src)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// One compare-exchange step of a bitonic network at lane stride M.
template <int M> __device__ __forceinline__ uint64_t bitonic_step(uint64_t v, bool take_max)
{
    const uint64_t o = lane_xor_u64<M>(v);
    return take_max ? (v > o ? v : o) : (v < o ? v : o);
}

__device__ __forceinline__ uint64_t wave_sort_desc(uint64_t v)
{
    const int lane = threadIdx.x & 63;
    // stage `size`: lanes in blocks of size sort descending where
    // lane & size == 0; a step at stride m keeps the max in the lower lane
    // of a pair when that block is descending
#define RAD_SORT_STEP(size_, m_) \
    v = bitonic_step<m_>(v, ((lane & (m_)) == 0) == ((lane & (size_)) == 0));
    RAD_SORT_STEP(2, 1)
    RAD_SORT_STEP(4, 2) RAD_SORT_STEP(4, 1)
    RAD_SORT_STEP(8, 4) RAD_SORT_STEP(8, 2) RAD_SORT_STEP(8, 1)
    RAD_SORT_STEP(16, 8) RAD_SORT_STEP(16, 4) RAD_SORT_STEP(16, 2) RAD_SORT_STEP(16, 1)
    RAD_SORT_STEP(32, 16) RAD_SORT_STEP(32, 8) RAD_SORT_STEP(32, 4) RAD_SORT_STEP(32, 2)
    RAD_SORT_STEP(32, 1)
    RAD_SORT_STEP(64, 32) RAD_SORT_STEP(64, 16) RAD_SORT_STEP(64, 8) RAD_SORT_STEP(64, 4)
    RAD_SORT_STEP(64, 2) RAD_SORT_STEP(64, 1)
#undef RAD_SORT_STEP
    return v;
}

// top and b sorted descending over the lanes: the top 64 of both, sorted.
__device__ __forceinline__ uint64_t wave_merge_desc(uint64_t top, uint64_t b)
{
    const int lane = threadIdx.x & 63;
    const uint64_t r = lane_reverse_u64(b);
    uint64_t v = top > r ? top : r;
    v = bitonic_step<32>(v, (lane & 32) == 0);
    v = bitonic_step<16>(v, (lane & 16) == 0);
    v = bitonic_step<8>(v, (lane & 8) == 0);
    v = bitonic_step<4>(v, (lane & 4) == 0);
    v = bitonic_step<2>(v, (lane & 2) == 0);
    v = bitonic_step<1>(v, (lane & 1) == 0);
    return v;
}

// The sorted 64-key lists of the workgroup's waves (top) merged pairwise
// through wl; wave 0 returns the 64 largest keys of all of them.  Every
// thread of the workgroup calls it.
__device__ static uint64_t merge_wave_lists(uint64_t top, uint64_t (*wl)[64])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    wl[w][lane] = top;
    for (int h = nw >> 1; h >= 1; h >>= 1) {
        __syncthreads();
        if (w < h)
            top = wave_merge_desc(wl[w][lane], wl[w + h][lane]);
        __syncthreads();
        if (w < h)
            wl[w][lane] = top;
    }
    return top;
}

// The 64 largest keys of the sorted 64-key lists [l0, l1) by the
// workgroup's waves: each wave merges every nw-th list, loading 16 of them
// at a time (a list whose largest key is not above the 64th of the wave's
// running list cannot contribute),
// then the wave lists are merged pairwise through wl; wave 0 returns the
// result.  Every thread of the workgroup calls it.
__device__ static uint64_t merge_lists_block(const uint64_t *__restrict__ in, uint32_t l0,
                                             uint32_t l1, uint64_t (*wl)[64])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    constexpr int kLoads = 16; // lists a wave loads before merging them
    uint64_t top = 0, thr = 0;
    for (uint32_t lb = l0 + w; lb < l1; lb += (uint32_t)nw * kLoads) {
        uint64_t v[kLoads];
#pragma unroll
        for (int q = 0; q < kLoads; ++q) {
            const uint32_t l = lb + (uint32_t)(q * nw);
            v[q] = l < l1 ? in[(size_t)l * 64 + lane] : 0ull;
        }
#pragma unroll
        for (int q = 0; q < kLoads; ++q) {
            if (lane_read_u64(v[q], 0) <= thr)
                continue;
            top = wave_merge_desc(top, v[q]);
            thr = lane_read_u64(top, 63);
        }
    }
    return merge_wave_lists(top, wl);
}

// Wave 0 of a merge: the patches of the k largest POWER keys (a zero key
// fills no slot) and their count.
__device__ static void write_shooters(uint64_t top, uint32_t k, uint32_t *__restrict__ shooters,
                                      uint32_t *__restrict__ nshoot)
{
    const int lane = threadIdx.x & 63;
    const bool ok = (uint32_t)lane < k && (top >> 32) != 0;
    if ((uint32_t)lane < k)
        shooters[lane] = ok ? 0xFFFFFFFFu - (uint32_t)top : 0u;
    const uint64_t m = __ballot(ok);
    if (lane == 0)
        *nshoot = (uint32_t)__popcll(m);
}

// Main.cpp:1251-1295 for all receivers at once: one thread per receiver j
// accumulates the contributions of the pass's shooters, then the shooters
// move their snapshot from unshot to shot.  Every nonzero F entry it reads
// is set back to zero, so the next pass finds the F rows cleared.
// With lists != null (the last launch of a pass whose next pass selects by
// POWER), each workgroup also leaves the sorted 64 largest selection keys
// of its 1,024 receivers' new state in lists[blockIdx.x * 64 + lane], the
// per-workgroup lists k_topk64_scan would make from the stored state, so
// the next pass's selection only merges them; or, with the pool armed, it
// adds every receiver whose key is at least the pool's tau to the pool
// instead.  NT threads per workgroup: 1,024 with lists (one merge
// level for up to 262,144 receivers); 256 without lists, and with lists
// when the pool is expected (then lists, left only when the pool is not
// armed, cover 256 receivers each).
constexpr uint32_t kShootListThreads = 1024;

template <uint32_t NT>
__global__ __launch_bounds__(NT) void k_shoot(uint32_t mode, const uint32_t *__restrict__ shooters,
                                               const uint32_t *__restrict__ nshoot,
                                               const float *__restrict__ snap,
                                               float *__restrict__ ff,
                                               uint32_t *__restrict__ fblk, uint32_t fstride,
                                               const float *__restrict__ refl,
                                               const float *__restrict__ area, uint32_t npatch,
                                               float *__restrict__ unshot,
                                               float *__restrict__ shot, uint32_t *__restrict__ ctl,
                                               uint32_t sbeg, uint32_t send, uint32_t last,
                                               uint64_t *__restrict__ lists,
                                               SelPool *__restrict__ pool)
{
    __shared__ uint32_t sid[RAD_MAX_K];
    __shared__ float sc[RAD_MAX_K * 4];
    __shared__ uint64_t wl[NT / 64][64];
    __shared__ uint32_t npool, gbase;
    const uint32_t pooled = lists && pool ? pool->armed : 0u;
    if (threadIdx.x == 0)
        npool = 0u;
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
    if (last && blockIdx.x == 0 && threadIdx.x == 0)
        count_pass(ctl, n);
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    // the wavefront's 64 receivers are one block of the F-block flags
    const uint32_t blk = (blockIdx.x * blockDim.x + (threadIdx.x & ~63u)) >> 6;
    const bool live = j < npatch;
    float u0 = 0.f, u1 = 0.f, u2 = 0.f, r0 = 0.f, r1 = 0.f, r2 = 0.f, inv_aj = 0.f;
    if (live) {
        u0 = unshot[3 * (size_t)j + 0];
        u1 = unshot[3 * (size_t)j + 1];
        u2 = unshot[3 * (size_t)j + 2];
        if (mode != RAD_MODE_REFERENCE) {
            r0 = refl[3 * (size_t)j + 0];
            r1 = refl[3 * (size_t)j + 1];
            r2 = refl[3 * (size_t)j + 2];
            inv_aj = 1.0f / area[j];
        }
    }
    // Shooters [sbeg, send) are added here: a pass may be shot in two
    // launches, the first half's as soon as its tiles are done, the second
    // (last) one adding the rest and moving the shooters' energy; the sums
    // are the same float additions in the same order either way.
    // Shooters are taken in groups of 64 and, within a group, in shooter
    // order; only those whose F block of these receivers is flagged are
    // read (an unflagged block is zero), 8 at a time (independent loads in
    // flight).  Entries are added in shooter order, skipping zeros, so the
    // sums are those of reading every entry.  The nonzero entries of a group
    // are set back to zero after the group's last read (a load issued after
    // a store waits for it), and the group's flags are cleared, which leaves
    // the F rows zero for the next pass (its clear skips them).
    const uint32_t se = min(n, send);
    if (blk * 64u < npatch) {
        for (uint32_t g0 = sbeg; g0 < se; g0 += 64) {
            const uint32_t g1 = min(se, g0 + 64);
            uint32_t *const fl = fblk ? fblk + (size_t)blk * fstride + g0 : nullptr;
            const bool mine = g0 + lane < g1;
            uint64_t todo;
            if (fl) {
                const bool set = mine && fl[lane] != 0u;
                todo = __ballot(set);
                if (set)
                    fl[lane] = 0u;
            } else {
                todo = __ballot(mine);
            }
            uint64_t nz = 0;
            while (todo) {
                uint32_t ss[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    ss[q] = todo ? g0 + (uint32_t)__ffsll((long long)todo) - 1u : ~0u;
                    todo &= todo - 1;
                }
                float fb[8];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    fb[q] = live && ss[q] != ~0u ? ff[(size_t)ss[q] * npatch + j] : 0.0f;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const uint32_t s = ss[q];
                    const float f = fb[q];
                    if (s == ~0u || f == 0.0f)
                        continue;
                    nz |= 1ull << (s - g0);
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
            }
            for (uint64_t m = nz; m; m &= m - 1)
                ff[(size_t)(g0 + (uint32_t)__ffsll((long long)m) - 1) * npatch + j] = 0.0f;
        }
    }
    if (live) {
        if (last) {
            // Main.cpp:1287-1295
            if (n && sid[n - 1] == j)
                stop_test(ctl, u0, u1, u2);
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
        }
        unshot[3 * (size_t)j + 0] = u0;
        unshot[3 * (size_t)j + 1] = u1;
        unshot[3 * (size_t)j + 2] = u2;
    }
    if (!lists)
        return;
    // sel_key(RAD_SELECT_POWER) of the state just stored
    uint64_t key = 0ull;
    if (live) {
        float e = area[j] * (u0 + u1 + u2);
        if (!(e > 0.0f))
            e = 0.0f;
        key = ((uint64_t)__float_as_uint(e) << 32) | (0xFFFFFFFFu - j);
    }
    if (pooled) {
        // every receiver whose key is at least tau (a positive power) goes
        // to this workgroup's pool region
        const uint64_t tau = pool->tau;
        const bool c = live && key >= tau;
        const uint64_t m = __ballot(c);
        uint32_t local = 0u;
        if (m) {
            uint32_t base = 0u;
            if (lane == 0)
                base = atomicAdd(&npool, (uint32_t)__popcll(m));
            base = (uint32_t)__shfl((int)base, 0);
            local = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        }
        __syncthreads();
        if (threadIdx.x == 0)
            gbase = npool ? atomicAdd(&pool->count, npool) : 0u;
        __syncthreads();
        const uint32_t idx = gbase + local;
        if (c && idx < kPoolCap)
            pool->keys[idx] = key;
        return;
    }
    uint64_t top = wave_sort_desc(key);
    wl[w][lane] = top;
    for (uint32_t h = NT / 128; h >= 1; h >>= 1) {
        __syncthreads();
        if (w < h)
            top = wave_merge_desc(wl[w][lane], wl[w + h][lane]);
        __syncthreads();
        if (w < h)
            wl[w][lane] = top;
    }
    if (w == 0)
        lists[(size_t)blockIdx.x * 64 + lane] = top;
}

uint32_t shoot_lists(uint32_t npatch, bool pooled)
{
    const uint32_t nt = pooled ? 256u : kShootListThreads;
    return (npatch + nt - 1) / nt;
}

void launch_shoot(uint32_t mode, uint32_t K, const uint32_t *shooters, const uint32_t *nshoot,
                  const float *snap, float *ff, uint32_t *fblk, const float *refl,
                  const float *area, uint32_t npatch, float *unshot, float *shot, uint32_t *ctl,
                  uint32_t sbeg, uint32_t send, bool last, uint64_t *lists, SelPool *pool,
                  hipStream_t st)
{
    if (lists && !pool)
        k_shoot<kShootListThreads><<<shoot_lists(npatch, false), kShootListThreads, 0, st>>>(
            mode, shooters, nshoot, snap, ff, fblk, K, refl, area, npatch, unshot, shot, ctl,
            sbeg, send, last ? 1u : 0u, lists, nullptr);
    else
        k_shoot<256><<<(npatch + 255) / 256, 256, 0, st>>>(mode, shooters, nshoot, snap, ff, fblk,
                                                          K, refl, area, npatch, unshot, shot,
                                                          ctl, sbeg, send, last ? 1u : 0u, lists,
                                                          pool);
}

__device__ static inline uint64_t sel_key(uint32_t rule, const float *unshot, const float *area,
                                          uint32_t j);

// The 64 largest of n unsorted keys: each wave takes every nw-th batch of
// 64, loading 8 batches at a time, and sorts and merges in those that hold
// a key above its running 64th; wave 0 returns the result.  Every thread of
// the workgroup calls it.
__device__ static uint64_t select_keys_block(const uint64_t *__restrict__ pk, uint32_t n,
                                             uint64_t (*wl)[64])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    constexpr int kLoads = 8;
    const uint32_t step = (uint32_t)nw * 64;
    uint64_t top = 0, thr = 0;
    for (uint32_t b0 = (uint32_t)w * 64; b0 < n; b0 += step * kLoads) {
        uint64_t v[kLoads];
#pragma unroll
        for (int q = 0; q < kLoads; ++q) {
            const uint32_t i = b0 + (uint32_t)q * step + lane;
            v[q] = i < n ? pk[i] : 0ull;
        }
#pragma unroll
        for (int q = 0; q < kLoads; ++q) {
            if (__ballot(v[q] > thr) == 0)
                continue;
            top = wave_merge_desc(top, wave_sort_desc(v[q]));
            thr = lane_read_u64(top, 63);
        }
    }
    return merge_wave_lists(top, wl);
}

// The 64 largest POWER keys of all npatch patches, by one workgroup (the
// pool's fallback): as k_topk64_scan, four batches loaded at a time.
__device__ static uint64_t select_all_block(const float *__restrict__ unshot,
                                            const float *__restrict__ area, uint32_t npatch,
                                            uint64_t (*wl)[64])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t step = (uint32_t)nw * 64;
    uint64_t top = 0, thr = 0;
    for (uint32_t base = (uint32_t)w * 64; base < npatch; base += 4 * step) {
        uint64_t kk[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t j = base + (uint32_t)q * step + lane;
            kk[q] = j < npatch ? sel_key(RAD_SELECT_POWER, unshot, area, j) : 0ull;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (__ballot(kk[q] > thr) == 0)
                continue;
            top = wave_merge_desc(top, wave_sort_desc(kk[q]));
            thr = lane_read_u64(top, 63);
        }
    }
    return merge_wave_lists(top, wl);
}

// Arms the pool after the merge of nlists sorted lists whose 64th largest
// key is t64.  The lists hold a subset of the state's keys, so at least as
// many keys of the state as of the lists are >= any tau.  tau steps down
// from t64 by 1/8 octave of power (2^20 in the float bits of the key's
// upper word), at most 8 octaves, until the lists hold `target` keys >= tau
// (t64 itself when target <= 64).  A tau of zero power is never armed.
// hist: 64 words of LDS.  Every thread of the workgroup calls it.
__device__ static void arm_pool(const uint64_t *__restrict__ in, uint32_t nlists, uint64_t t64,
                                uint32_t target, SelPool *__restrict__ pool, uint32_t *hist)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t th = (uint32_t)(t64 >> 32);
    if (threadIdx.x < 64)
        hist[threadIdx.x] = 0u;
    __syncthreads();
    for (uint32_t l = (uint32_t)w; l < nlists; l += (uint32_t)nw) {
        const uint64_t e = in[(size_t)l * 64 + lane];
        const uint32_t eh = (uint32_t)(e >> 32);
        // bucket (th - eh) >> 20 of the keys below t64 within 8 octaves
        // (64: none)
        uint32_t b = 64u;
        if (e < t64 && eh != 0u && th - eh < (64u << 20))
            b = (th - eh) >> 20;
        // a list is sorted, so a bucket's keys are a run of lanes: the run's
        // first lane adds its length
        const uint32_t prev = (uint32_t)__shfl_up((int)b, 1);
        const bool start = lane == 0 || prev != b;
        const uint64_t starts = __ballot(start);
        if (start && b < 64u) {
            const uint64_t after = lane == 63 ? 0ull : starts & (~0ull << (lane + 1));
            const uint32_t next = after ? (uint32_t)__ffsll((long long)after) - 1u : 64u;
            atomicAdd(&hist[b], next - (uint32_t)lane);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t armed = 0u;
        uint64_t tau = 0ull;
        if (th != 0u) {
            uint32_t c = 64u; // the merged 64 are >= t64
            if (c >= target) {
                tau = t64;
                armed = 1u;
            }
            for (uint32_t b = 0; b < 64u && !armed; ++b) {
                c += hist[b];
                if (c >= target) {
                    const uint32_t drop = (b + 1u) << 20;
                    if (th > drop) {
                        tau = (uint64_t)(th - drop + 1u) << 32;
                        armed = 1u;
                    }
                    break;
                }
            }
        }
        if (armed) {
            pool->tau = tau;
            pool->arms += 1u;
        }
        pool->armed = armed;
    }
}

// The start of the pass after one whose last shoot left per-workgroup lists
// or pooled, in one workgroup: the selection and what k_prepare_clear does
// (the stop gate, the shooters' snapshots and MVPs, bin counts and pass
// counters cleared; no F rows or hits, which the shoot leaves zero).
// The selection takes the K largest POWER keys (as k_topk64_merge):
//  - pooled: from the pool when it holds K .. pool_cap keys, which are then
//    exactly the keys >= tau, so their K largest are the state's; otherwise
//    from a scan of every patch, and the pool is disarmed (the next shoot
//    leaves lists);
//  - else: the merge of the lists, then the pool armed for `target` keys
//    (target 0: disarmed).
__global__ __launch_bounds__(1024) void k_next_pass(const uint64_t *__restrict__ lists,
                                                    uint32_t nlists, NextPass np,
                                                    const float *__restrict__ unshot,
                                                    const float *__restrict__ area,
                                                    uint32_t npatch,
                                                    const uint32_t *__restrict__ ctl,
                                                    SelPool *__restrict__ pool,
                                                    uint32_t lists_only, uint32_t target,
                                                    uint32_t pool_cap)
{
    __shared__ uint64_t wl[16][64];
    __shared__ uint32_t hist[64], ssel[64], sn;
    __shared__ uint64_t t64s;
    const int lane = threadIdx.x & 63;
    const uint32_t t = threadIdx.x;
    const uint32_t stop = ctl[kCtlStop];
    const uint32_t pooled = lists_only ? 0u : pool->armed;
    uint64_t best;
    if (pooled) {
        const uint32_t n = pool->count;
        if (n >= np.K && n <= pool_cap) {
            best = select_keys_block(pool->keys, n, wl);
            if (t == 0)
                pool->hits += 1u;
        } else {
            best = select_all_block(unshot, area, npatch, wl);
            if (t == 0) {
                pool->falls += 1u;
                pool->armed = 0u;
            }
        }
    } else {
        best = merge_lists_block(lists, 0u, nlists, wl);
        if (target) {
            if (t == 0)
                t64s = lane_read_u64(best, 63);
            __syncthreads();
            arm_pool(lists, nlists, t64s, target, pool, hist);
        } else if (t == 0) {
            pool->armed = 0u;
        }
    }
    // wave 0: the shooters (as write_shooters) into memory and LDS; the
    // stop gate of k_prepare_clear sets the count to 0
    if (t < 64) {
        const bool ok = (uint32_t)lane < np.K && (best >> 32) != 0;
        const uint32_t id = ok ? 0xFFFFFFFFu - (uint32_t)best : 0u;
        if ((uint32_t)lane < np.K) {
            np.shooters[lane] = id;
            ssel[lane] = id;
        }
        const uint32_t m = stop ? 0u : (uint32_t)__popcll(__ballot(ok));
        if (lane == 0) {
            *np.nshoot = m;
            sn = m;
        }
    }
    __syncthreads();
    // every read of the pool count is done: the next shoot pools from zero
    if (t == 0)
        pool->count = 0u;
    for (uint32_t i = t; i < np.K * 5; i += blockDim.x)
        prepare_item(i, np.K, ssel, &sn, np.frames, unshot, np.snap, np.mvp, nullptr);
    clear_range(t, blockDim.x, np.bin_count, np.nbins, nullptr, nullptr, 0, np.counters,
                np.ent_cap);
}

void launch_next_pass(const uint64_t *lists, uint32_t nlists, const NextPass &np,
                      const float *unshot, const float *area, uint32_t npatch,
                      const uint32_t *ctl, SelPool *pool, bool lists_only, uint32_t target,
                      uint32_t pool_cap, hipStream_t st)
{
    k_next_pass<<<1, 1024, 0, st>>>(lists, nlists, np, unshot, area, npatch, ctl, pool,
                                    lists_only ? 1u : 0u, target,
                                    pool_cap < kPoolCap ? pool_cap : kPoolCap);
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

// ---- K <= 64: one sorted list of 64 keys per wavefront, one key per lane
// (lane 0 largest).  A batch of 64 new keys is sorted with a shuffle bitonic
// network and merged into the list (max against the reversed batch gives a
// bitonic sequence holding the top 64 of both; a half-cleaner pass sorts
// it).  Waves scan the patches (a batch with no key above the wave's 64th
// is skipped), each workgroup merges its wave lists in LDS, then one
// workgroup merges the workgroup lists.  The keys are distinct (they hold
// the patch index), so the result does not depend on the merge order.
constexpr uint32_t kTopkScanWaves = 8;

__global__ __launch_bounds__(kTopkScanWaves * 64) void k_topk64_scan(const float *__restrict__ unshot,
                                                     const float *__restrict__ area,
                                                     uint32_t npatch, uint64_t *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const uint32_t gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * (blockDim.x >> 6);
    __shared__ uint64_t wl[kTopkScanWaves][64];
    uint64_t top = 0, thr = 0; // thr: the list's 64th key
    for (uint32_t base = gw * 64; base < npatch; base += nw * 64) {
        uint32_t j = base + lane;
        uint64_t key = j < npatch ? sel_key(RAD_SELECT_POWER, unshot, area, j) : 0ull;
        if (__ballot(key > thr) == 0)
            continue;
        top = wave_merge_desc(top, wave_sort_desc(key));
        thr = lane_read_u64(top, 63);
    }
    // the workgroup's wave lists: pairwise merge levels in LDS
    const int w = threadIdx.x >> 6;
    wl[w][lane] = top;
    for (int h = (int)kTopkScanWaves / 2; h >= 1; h >>= 1) {
        __syncthreads();
        if (w < h)
            top = wave_merge_desc(wl[w][lane], wl[w + h][lane]);
        __syncthreads();
        if (w < h)
            wl[w][lane] = top;
    }
    if (w == 0)
        out[(size_t)blockIdx.x * 64 + lane] = top;
}

// Merges nlists sorted lists of 64 (in[l * 64 + lane]) per workgroup; the
// last workgroup level (gridDim.x == 1) writes the shooters.
__global__ __launch_bounds__(1024) void k_topk64_merge(const uint64_t *__restrict__ in,
                                                       uint32_t nlists, uint32_t per_block,
                                                       uint32_t k, uint64_t *__restrict__ out,
                                                       uint32_t *__restrict__ shooters,
                                                       uint32_t *__restrict__ nshoot)
{
    __shared__ uint64_t wl[16][64];
    const uint32_t l0 = blockIdx.x * per_block;
    const uint32_t l1 = l0 + per_block < nlists ? l0 + per_block : nlists;
    const uint64_t top = merge_lists_block(in, l0, l1, wl);
    if ((threadIdx.x >> 6) != 0)
        return;
    if (shooters)
        write_shooters(top, k, shooters, nshoot);
    else
        out[(size_t)blockIdx.x * 64 + (threadIdx.x & 63)] = top;
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
    if (k <= 64) {
        // wave lists: about four batches of 64 patches per wave
        uint32_t waves = (npatch + 255) / 256;
        if (waves < kTopkScanWaves)
            waves = kTopkScanWaves;
        uint32_t blocks = (waves + kTopkScanWaves - 1) / kTopkScanWaves;
        if ((size_t)blocks * 64 > topk_cand_half(npatch, k))
            blocks = (uint32_t)(topk_cand_half(npatch, k) / 64);
        uint64_t *a = cand, *b = cand + topk_cand_half(npatch, k);
        k_topk64_scan<<<blocks, kTopkScanWaves * 64, 0, st>>>(unshot, area, npatch, a);
        uint32_t nl = blocks;
        while (nl > 256) {
            uint32_t per = 64, nb = (nl + per - 1) / per;
            k_topk64_merge<<<nb, 1024, 0, st>>>(a, nl, per, k, b, nullptr, nullptr);
            uint64_t *t = a;
            a = b;
            b = t;
            nl = nb;
        }
        k_topk64_merge<<<1, 1024, 0, st>>>(a, nl, nl, k, nullptr, shooters, nshoot);
        return;
    }
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
    size_t bitonic = nb * k + 1;
    // the wave-list path keeps 64 keys per wave, one wave per 256 patches
    size_t waves = ((size_t)npatch + 255) / 256;
    size_t lists = (waves < 4 ? 4 : waves) + 3;
    return bitonic > lists * 64 ? bitonic : lists * 64;
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
