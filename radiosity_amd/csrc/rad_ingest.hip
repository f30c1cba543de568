// rad_ingest.hip - scene ingest and vertex colours on the device.
//
// k_ingest writes a subdivided scene straight into the context's buffers:
// one thread per child quad finds its source face (binary search over the
// faces' first child), computes the child's corners with the arithmetic of
// Patch::divide (Patch.cpp:85-117) or of the native triangle grid, and
// writes its triangles (9 planes), patch ids, frames, areas, inherited
// colour / radiosity / illumination (Patch.cpp:119-122) and neighbour links
// (Patch.cpp:127-222).  The host only plans the faces (how each splits and
// where its children start), which needs the acos / sin of the reference's
// area test; every per-child operation is here, in the same float order as
// the host restatement, so both give the same bits.
//
// k_quad_value / k_vertex_colors restate Colors::smoothShadePatch
// (Colors.cpp:195-262) for every quad at once.
#include "rad_internal.h"

namespace rad {

namespace {

struct P3 {
    float x, y, z;
};
__device__ inline P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline P3 operator*(P3 a, float t) { return {a.x * t, a.y * t, a.z * t}; }
// Vector.h:448-452: division by a scalar multiplies by its reciprocal
__device__ inline P3 divs(P3 a, float t)
{
    t = 1 / t;
    return {a.x * t, a.y * t, a.z * t};
}
// Vector.h:534-537: a.v_Cross(b) is b x a
__device__ inline P3 vcross(P3 a, P3 b)
{
    return {b.y * a.z - b.z * a.y, b.z * a.x - b.x * a.z, b.x * a.y - b.y * a.x};
}

__device__ inline float tri_area(P3 a, P3 b, P3 c)
{
    P3 n = vcross(c - a, b - a); // (b - a) x (c - a)
    return (float)(0.5 * sqrt((double)n.x * n.x + (double)n.y * n.y + (double)n.z * n.z));
}

__device__ inline P3 grid_vertex(P3 A, P3 B, P3 C, uint32_t n, uint32_t i, uint32_t j)
{
    return A + (B - A) * ((float)i / (float)n) + (C - A) * ((float)j / (float)n);
}

} // namespace

__global__ __launch_bounds__(256) void k_ingest(
    uint32_t nface, uint32_t nquad, const float *__restrict__ faces,
    const FacePlanDev *__restrict__ plan, const float *__restrict__ color,
    const float *__restrict__ emission, const float *__restrict__ shot0, uint32_t mode,
    uint32_t ntri, float *__restrict__ tri, uint32_t *__restrict__ tri_patch,
    float *__restrict__ frames, float *__restrict__ refl, float *__restrict__ area,
    float *__restrict__ emis, float *__restrict__ unshot, float *__restrict__ shot,
    uint32_t *__restrict__ patch_quad, uint32_t *__restrict__ quad_pbeg,
    uint32_t *__restrict__ quad_nbr)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nquad)
        return;
    // the face whose children include q: largest f with plan[f].qbeg <= q
    uint32_t lo = 0, hi = nface - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) / 2;
        if (plan[mid].qbeg <= q)
            lo = mid;
        else
            hi = mid - 1;
    }
    const uint32_t f = lo;
    const FacePlanDev pl = plan[f];
    const uint32_t c = q - pl.qbeg;
    const float *fv = faces + 12 * (size_t)f;
    P3 src[4];
    for (int k = 0; k < 4; ++k)
        src[k] = {fv[3 * k], fv[3 * k + 1], fv[3 * k + 2]};
    P3 v[4];
    int ntc = 2;
    uint32_t nb[8];
    if (pl.native) {
        const uint32_t n = pl.kx;
        ntc = 1;
        if (n == 1) {
            v[0] = src[0];
            v[1] = src[1];
            v[2] = src[2];
        } else {
            // row j starts at triangle j (2n - j) = n^2 - (n - j)^2
            uint32_t j = (uint32_t)floor((double)n - sqrt((double)n * n - (double)c));
            while (j > 0 && (uint64_t)j * (2 * n - j) > c)
                --j;
            while ((uint64_t)(j + 1) * (2 * n - (j + 1)) <= c)
                ++j;
            uint32_t k = c - j * (2 * n - j), i = k / 2;
            if (k % 2 == 0) {
                v[0] = grid_vertex(src[0], src[1], src[2], n, i, j);
                v[1] = grid_vertex(src[0], src[1], src[2], n, i + 1, j);
                v[2] = grid_vertex(src[0], src[1], src[2], n, i, j + 1);
            } else {
                v[0] = grid_vertex(src[0], src[1], src[2], n, i + 1, j);
                v[1] = grid_vertex(src[0], src[1], src[2], n, i + 1, j + 1);
                v[2] = grid_vertex(src[0], src[1], src[2], n, i, j + 1);
            }
        }
        v[3] = v[2];
        for (int k = 0; k < 8; ++k)
            nb[k] = c;
    } else {
        const uint32_t kx = pl.kx, ky = pl.ky;
        if (kx * ky == 1) {
            for (int k = 0; k < 4; ++k)
                v[k] = src[k];
        } else {
            P3 A = src[0], B = src[1], C = src[2], D = src[3];
            P3 pCD = divs(C - D, (float)kx);
            P3 pAB = divs(B - A, (float)kx);
            uint32_t col = c % kx, row = c / kx;
            P3 bCD = pCD * (float)col + D;
            P3 bAB = pAB * (float)col + A;
            P3 bCD1 = pCD * (float)(col + 1) + D;
            P3 bAB1 = pAB * (float)(col + 1) + A;
            v[0] = bAB + divs(bCD - bAB, (float)ky) * (float)row;
            v[1] = bAB1 + divs(bCD1 - bAB1, (float)ky) * (float)row;
            v[2] = bAB1 + divs(bCD1 - bAB1, (float)ky) * (float)(row + 1);
            v[3] = bAB + divs(bCD - bAB, (float)ky) * (float)(row + 1);
        }
        uint32_t col = c % kx, row = c / kx;
        bool top = row + 1 >= ky, bottom = row == 0, left = col == 0, right = col + 1 >= kx;
        nb[0] = top ? (left ? c : c - 1) : (left ? c + kx : c + kx - 1);
        nb[1] = top ? c : c + kx;
        nb[2] = top ? (right ? c : c + 1) : (right ? c + kx : c + kx + 1);
        nb[3] = right ? c : c + 1;
        nb[4] = bottom ? (right ? c : c + 1) : (right ? c - kx : c - kx + 1);
        nb[5] = bottom ? c : c - kx;
        nb[6] = bottom ? (left ? c : c - 1) : (left ? c - kx : c - kx - 1);
        nb[7] = left ? c : c - 1;
    }
    for (int k = 0; k < 8; ++k)
        quad_nbr[8 * (size_t)q + k] = pl.qbeg + nb[k];
    const uint32_t t0 = pl.tbeg + (uint32_t)ntc * c;
    const uint32_t p0 = mode == 0 ? pl.pbeg + c : pl.pbeg + (uint32_t)ntc * c;
    quad_pbeg[q] = p0;
    P3 tv[2][3] = {{v[0], v[1], v[2]}, {v[0], v[2], v[3]}};
    for (int h = 0; h < ntc; ++h) {
        const uint32_t t = t0 + h;
        for (int k = 0; k < 3; ++k) {
            tri[(size_t)(3 * k + 0) * ntri + t] = tv[h][k].x;
            tri[(size_t)(3 * k + 1) * ntri + t] = tv[h][k].y;
            tri[(size_t)(3 * k + 2) * ntri + t] = tv[h][k].z;
        }
        const uint32_t p = mode == 0 ? p0 : p0 + h;
        tri_patch[t] = p;
        if (mode == 0 && h == 1)
            continue;
        float fr[9], a;
        if (mode == 0) {
            // Patch.cpp:253-276; area of the two drawn triangles
            P3 eye = {(v[0].x + v[1].x + v[2].x + v[3].x) / 4.0f,
                      (v[0].y + v[1].y + v[2].y + v[3].y) / 4.0f,
                      (v[0].z + v[1].z + v[2].z + v[3].z) / 4.0f};
            P3 n = vcross(v[1] - v[0], v[3] - v[0]);
            P3 up = v[3] - v[0];
            float f9[9] = {eye.x, eye.y, eye.z, n.x, n.y, n.z, up.x, up.y, up.z};
            for (int k = 0; k < 9; ++k)
                fr[k] = f9[k];
            a = tri_area(v[0], v[1], v[2]) + tri_area(v[0], v[2], v[3]);
        } else {
            const P3 *tt = tv[h];
            P3 eye = {(tt[0].x + tt[1].x + tt[2].x) / 3.0f, (tt[0].y + tt[1].y + tt[2].y) / 3.0f,
                      (tt[0].z + tt[1].z + tt[2].z) / 3.0f};
            P3 n = vcross(tt[1] - tt[0], tt[2] - tt[0]);
            P3 up = tt[2] - tt[0];
            float f9[9] = {eye.x, eye.y, eye.z, n.x, n.y, n.z, up.x, up.y, up.z};
            for (int k = 0; k < 9; ++k)
                fr[k] = f9[k];
            a = tri_area(tt[0], tt[1], tt[2]);
        }
        for (int k = 0; k < 9; ++k)
            frames[9 * (size_t)p + k] = fr[k];
        area[p] = a;
        patch_quad[p] = q;
        for (int k = 0; k < 3; ++k) {
            float cc = color ? color[3 * (size_t)f + k] : 0.0f;
            float e = emission ? emission[3 * (size_t)f + k] : 0.0f;
            float s = shot0 ? shot0[3 * (size_t)f + k] : 0.0f;
            refl[3 * (size_t)p + k] = cc;
            emis[3 * (size_t)p + k] = e;
            unshot[3 * (size_t)p + k] = e;
            shot[3 * (size_t)p + k] = s;
        }
    }
}

// Per quad: colour * (illumination + radiosity) (Colors.cpp:198) of its
// patch; a quad of several patches takes their area-weighted mean.
__global__ __launch_bounds__(256) void k_quad_value(uint32_t nquad,
                                                    const uint32_t *__restrict__ quad_pbeg,
                                                    const float *__restrict__ refl,
                                                    const float *__restrict__ area,
                                                    const float *__restrict__ unshot,
                                                    const float *__restrict__ shot,
                                                    float *__restrict__ val)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nquad)
        return;
    const uint32_t pb = quad_pbeg[q], pe = quad_pbeg[q + 1];
    float out[3];
    if (pe - pb == 1) {
        for (int k = 0; k < 3; ++k)
            out[k] = refl[3 * (size_t)pb + k] *
                     (shot[3 * (size_t)pb + k] + unshot[3 * (size_t)pb + k]);
    } else {
        float num[3] = {0.0f, 0.0f, 0.0f}, den = 0.0f;
        for (uint32_t p = pb; p < pe; ++p) {
            float a = area[p];
            for (int k = 0; k < 3; ++k)
                num[k] = num[k] + a * (refl[3 * (size_t)p + k] *
                                       (shot[3 * (size_t)p + k] + unshot[3 * (size_t)p + k]));
            den = den + a;
        }
        for (int k = 0; k < 3; ++k)
            out[k] = num[k] / den;
    }
    for (int k = 0; k < 3; ++k)
        val[3 * (size_t)q + k] = out[k];
}

// Colors::smoothShadePatch: vertex colours lb, rb, rt, lt (corners A, B, C,
// D) each the mean of the quad's value and three neighbours', summed in the
// reference's order and divided by 4 as a multiply by 1/4.
__global__ __launch_bounds__(256) void k_vertex_colors(uint32_t nquad,
                                                       const uint32_t *__restrict__ nbr,
                                                       const float *__restrict__ val,
                                                       float *__restrict__ out)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nquad)
        return;
    uint32_t n[8];
    for (int k = 0; k < 8; ++k)
        n[k] = nbr[8 * (size_t)q + k];
    // neighbour slots per corner: lt 7 0 1, rt 1 2 3, rb 3 4 5, lb 5 6 7
    const int corner_of_out[4] = {5, 3, 1, 7}; // lb, rb, rt, lt: first slot
    for (int o = 0; o < 4; ++o) {
        int s0 = corner_of_out[o];
        uint32_t m[3] = {n[s0], n[(s0 + 1) & 7], n[(s0 + 2) & 7]};
        for (int k = 0; k < 3; ++k) {
            float acc = val[3 * (size_t)q + k];
            acc = acc + val[3 * (size_t)m[0] + k];
            acc = acc + val[3 * (size_t)m[1] + k];
            acc = acc + val[3 * (size_t)m[2] + k];
            out[12 * (size_t)q + 3 * o + k] = acc * (1 / 4.0f);
        }
    }
}

void launch_ingest(uint32_t nface, uint32_t nquad, const float *faces, const FacePlanDev *plan,
                   const float *color, const float *emission, const float *shot0, uint32_t mode,
                   uint32_t ntri, float *tri, uint32_t *tri_patch, float *frames, float *refl,
                   float *area, float *emis, float *unshot, float *shot, uint32_t *patch_quad,
                   uint32_t *quad_pbeg, uint32_t *quad_nbr, hipStream_t st)
{
    k_ingest<<<(nquad + 255) / 256, 256, 0, st>>>(nface, nquad, faces, plan, color, emission,
                                                 shot0, mode, ntri, tri, tri_patch, frames, refl,
                                                 area, emis, unshot, shot, patch_quad, quad_pbeg,
                                                 quad_nbr);
}

void launch_vertex_colors(uint32_t nquad, const uint32_t *quad_pbeg, const uint32_t *quad_nbr,
                          const float *refl, const float *area, const float *unshot,
                          const float *shot, float *val, float *out, hipStream_t st)
{
    k_quad_value<<<(nquad + 255) / 256, 256, 0, st>>>(nquad, quad_pbeg, refl, area, unshot, shot,
                                                     val);
    k_vertex_colors<<<(nquad + 255) / 256, 256, 0, st>>>(nquad, quad_nbr, val, out);
}

} // namespace rad
