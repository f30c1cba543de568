// rad_patch.h - patch geometry shared by the scene sources and the C++
// drop-in layer (host code): vector helpers with the reference's operation
// order, quads, Patch::divide / Model::subdivide, and the PrimitiveModel
// Cornell data.  Compiled with -ffp-contract=off.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace radpatch {

struct V3 {
    float x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, float t) { return {a.x * t, a.y * t, a.z * t}; }
// Vector.h:448-452: division by a scalar multiplies by its reciprocal
inline V3 divs(V3 a, float t)
{
    t = 1 / t;
    return {a.x * t, a.y * t, a.z * t};
}
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
// Vector.h:534-537: a.v_Cross(b) is b x a
inline V3 vcross(V3 a, V3 b)
{
    return {b.y * a.z - b.z * a.y, b.z * a.x - b.x * a.z, b.x * a.y - b.y * a.x};
}
inline V3 cross(V3 a, V3 b) { return vcross(b, a); }

struct Quad {
    V3 v[4];
    V3 color, illum, rad;
    int lightish; // 1 for an emitter source quad
    int group;    // source surface (used to pick ceiling lights)
};

// The decision of Patch::divide (Patch.cpp:47-83): false when the quad is
// already small enough (S <= area * 1.01), else the kx * ky grid it splits
// into.
inline bool divide_grid(const Quad &q, double area, unsigned &kx, unsigned &ky)
{
    V3 A = q.v[0], B = q.v[1], C = q.v[2], D = q.v[3];
    V3 u1 = A - C, u2 = B - D;
    double phi = (double)acosf(dot(u1, u2) / (length(u1) * length(u2)));
    double S = (double)(0.5f * length(u1) * length(u2)) * sin(phi);
    if (S <= (area * 1.01))
        return false;
    double a = sqrt(area);
    kx = (unsigned)ceil((double)length(B - A) / a);
    ky = (unsigned)ceil((double)length(C - B) / a);
    return true;
}

// Corners of child i (column i % kx, row i / kx, [0, 0] at A) of the grid
// (Patch.cpp:85-117).
inline void divide_child(const V3 *v, unsigned kx, unsigned ky, unsigned i, V3 *out)
{
    V3 A = v[0], B = v[1], C = v[2], D = v[3];
    V3 pCD = divs(C - D, (float)kx);
    V3 pAB = divs(B - A, (float)kx);
    unsigned col = i % kx, row = i / kx;
    V3 bCD = pCD * (float)col + D;
    V3 bAB = pAB * (float)col + A;
    V3 bCD1 = pCD * (float)(col + 1) + D;
    V3 bAB1 = pAB * (float)(col + 1) + A;
    out[0] = bAB + divs(bCD - bAB, (float)ky) * (float)row;
    out[1] = bAB1 + divs(bCD1 - bAB1, (float)ky) * (float)row;
    out[2] = bAB1 + divs(bCD1 - bAB1, (float)ky) * (float)(row + 1);
    out[3] = bAB + divs(bCD - bAB, (float)ky) * (float)(row + 1);
}

// Neighbour links of child i in the kx * ky grid (Patch.cpp:127-222), as
// child indices; slots 0..7 run clockwise from the top-left (0 top-left,
// 1 top, 2 top-right, 3 right, 4 bottom-right, 5 bottom, 6 bottom-left,
// 7 left; row + 1 is "up").  At the grid border a diagonal slot falls back to
// the side neighbour along the border, or to the child itself.
inline void divide_neighbours(unsigned kx, unsigned ky, unsigned i, unsigned nb[8])
{
    unsigned col = i % kx, row = i / kx;
    auto at = [kx](unsigned r, unsigned c) { return r * kx + c; };
    bool top = row + 1 >= ky, bottom = row == 0, left = col == 0, right = col + 1 >= kx;
    nb[0] = top ? (left ? i : at(row, col - 1)) : (left ? at(row + 1, col) : at(row + 1, col - 1));
    nb[1] = top ? i : at(row + 1, col);
    nb[2] = top ? (right ? i : at(row, col + 1)) : (right ? at(row + 1, col) : at(row + 1, col + 1));
    nb[3] = right ? i : at(row, col + 1);
    nb[4] = bottom ? (right ? i : at(row, col + 1)) : (right ? at(row - 1, col) : at(row - 1, col + 1));
    nb[5] = bottom ? i : at(row - 1, col);
    nb[6] = bottom ? (left ? i : at(row, col - 1)) : (left ? at(row - 1, col) : at(row - 1, col - 1));
    nb[7] = left ? i : at(row, col - 1);
}

// Patch::divide (Patch.cpp:47-125): returns false when the quad is already
// small enough (S <= area * 1.01).
inline bool divide(const Quad &q, double area, std::vector<Quad> &out)
{
    unsigned kx, ky;
    if (!divide_grid(q, area, kx, ky))
        return false;
    for (unsigned i = 0; i < kx * ky; ++i) {
        Quad c = q;
        divide_child(q.v, kx, ky, i, c.v);
        out.push_back(c);
    }
    return true;
}

// Model::subdivide (Model.cpp:27-60): one pass, children replace parents in
// order.  nbr (optional) receives 8 neighbour links per output quad, as
// indices into the output; an undivided quad links to itself.
inline std::vector<Quad> subdivide(const std::vector<Quad> &in, double area, bool keep_lights,
                                   std::vector<unsigned> *nbr = nullptr)
{
    std::vector<Quad> out;
    for (const Quad &q : in) {
        unsigned base = (unsigned)out.size(), kx = 1, ky = 1;
        if (!(area > 0 && !(keep_lights && q.lightish) && divide_grid(q, area, kx, ky))) {
            kx = ky = 1;
            out.push_back(q);
        } else {
            for (unsigned i = 0; i < kx * ky; ++i) {
                Quad c = q;
                divide_child(q.v, kx, ky, i, c.v);
                out.push_back(c);
            }
        }
        if (nbr)
            for (unsigned i = 0; i < kx * ky; ++i) {
                unsigned nb[8];
                divide_neighbours(kx, ky, i, nb);
                for (int k = 0; k < 8; ++k)
                    nbr->push_back(base + nb[k]);
            }
    }
    return out;
}

// ------------------------------------------------------------ face ingest
// Source faces (OBJ faces or any quads; a triangle face is the quad A B C C,
// as WaveFrontModel.cpp:98-100 pads it) become patches in one of two modes:
//  - kFacesQuads: the reference's path.  Every face is split by
//    Patch::divide and every quad is a patch drawn as the triangles
//    (A, B, C), (A, C, D) (ModelContainer.cpp:103-110);
//  - kFacesTriangles: every triangle is a patch.  Quads are split by
//    Patch::divide into two triangle patches each; a triangle face (C == D
//    bit for bit) is split natively into n * n congruent triangles, with
//    n = ceil(sqrt(S / area)) when S > area * 1.01.
enum { kFacesQuads = 0, kFacesTriangles = 1 };

// Per source face: how it splits and where its children start.
struct FacePlan {
    uint32_t native;    // 1: native triangle grid of side n = kx
    uint32_t kx, ky;    // quad grid (1 x 1 when not divided)
    uint32_t qbeg;      // first child quad
    uint32_t tbeg;      // first triangle
    uint32_t pbeg;      // first patch
};

inline V3 face_vertex(const float *f, int k) { return {f[3 * k], f[3 * k + 1], f[3 * k + 2]}; }

inline bool same_bits(V3 a, V3 b)
{
    return !memcmp(&a, &b, sizeof(V3));
}

inline double tri_area_d(V3 a, V3 b, V3 c)
{
    V3 n = cross(b - a, c - a);
    return 0.5 * sqrt((double)n.x * n.x + (double)n.y * n.y + (double)n.z * n.z);
}

// Plans every face; returns false when the counts exceed 2^31 triangles.
inline bool plan_faces(const float *quads, uint32_t nface, double area, int mode,
                       std::vector<FacePlan> &plan, uint32_t &nquad, uint32_t &ntri,
                       uint32_t &npatch)
{
    plan.resize(nface);
    uint64_t q = 0, t = 0, p = 0;
    for (uint32_t f = 0; f < nface; ++f) {
        const float *fv = quads + 12 * (size_t)f;
        Quad qd{};
        for (int k = 0; k < 4; ++k)
            qd.v[k] = face_vertex(fv, k);
        FacePlan &pl = plan[f];
        pl.native = mode == kFacesTriangles && same_bits(qd.v[2], qd.v[3]);
        pl.kx = pl.ky = 1;
        pl.qbeg = (uint32_t)q;
        pl.tbeg = (uint32_t)t;
        pl.pbeg = (uint32_t)p;
        uint64_t kids;
        if (pl.native) {
            double S = tri_area_d(qd.v[0], qd.v[1], qd.v[2]);
            if (area > 0 && S > area * 1.01)
                pl.kx = pl.ky = (uint32_t)ceil(sqrt(S / area));
            kids = (uint64_t)pl.kx * pl.kx;
            t += kids;
            p += kids;
        } else {
            unsigned kx, ky;
            if (area > 0 && divide_grid(qd, area, kx, ky)) {
                pl.kx = kx;
                pl.ky = ky;
            }
            kids = (uint64_t)pl.kx * pl.ky;
            t += 2 * kids;
            p += mode == kFacesQuads ? kids : 2 * kids;
        }
        q += kids;
        if (t > (1ull << 31))
            return false;
    }
    nquad = (uint32_t)q;
    ntri = (uint32_t)t;
    npatch = (uint32_t)p;
    return true;
}

// Native triangle grid of side n over A, B, C: vertex (i, j) is
// A + (B - A) i/n + (C - A) j/n; row j (towards C) holds 2 (n - j) - 1
// triangles, up (i, j), (i+1, j), (i, j+1) at even positions and down
// (i+1, j), (i+1, j+1), (i, j+1) at odd ones, all wound like A, B, C.
inline V3 tri_grid_vertex(V3 A, V3 B, V3 C, uint32_t n, uint32_t i, uint32_t j)
{
    return A + (B - A) * ((float)i / (float)n) + (C - A) * ((float)j / (float)n);
}

inline void tri_grid_child(V3 A, V3 B, V3 C, uint32_t n, uint32_t t, V3 *out)
{
    if (n == 1) {
        out[0] = A;
        out[1] = B;
        out[2] = C;
        return;
    }
    // row j starts at triangle j (2n - j) = n^2 - (n - j)^2
    uint32_t j = (uint32_t)floor((double)n - sqrt((double)n * n - (double)t));
    while (j > 0 && (uint64_t)j * (2 * n - j) > t)
        --j;
    while ((uint64_t)(j + 1) * (2 * n - (j + 1)) <= t)
        ++j;
    uint32_t k = t - j * (2 * n - j), i = k / 2;
    if (k % 2 == 0) {
        out[0] = tri_grid_vertex(A, B, C, n, i, j);
        out[1] = tri_grid_vertex(A, B, C, n, i + 1, j);
        out[2] = tri_grid_vertex(A, B, C, n, i, j + 1);
    } else {
        out[0] = tri_grid_vertex(A, B, C, n, i + 1, j);
        out[1] = tri_grid_vertex(A, B, C, n, i + 1, j + 1);
        out[2] = tri_grid_vertex(A, B, C, n, i, j + 1);
    }
}

// PrimitiveModel.cpp:92-213 data (vertices x,y,z per corner, 4 corners per quad)
static const float kRoom[60] = {
    0.0f,   0.0f,   5.592f, 5.496f, 0.0f,   5.592f, 5.560f, 5.488f, 5.592f, 0.0f,   5.488f, 5.592f,
    0.0f,   5.488f, 0.0f,   0.0f,   5.488f, 5.592f, 5.560f, 5.488f, 5.592f, 5.560f, 5.488f, 0.0f,
    0.0f,   0.0f,   0.0f,   5.528f, 0.0f,   0.0f,   5.496f, 0.0f,   5.592f, 0.0f,   0.0f,   5.592f,
    0.0f,   0.0f,   0.0f,   0.0f,   0.0f,   5.592f, 0.0f,   5.488f, 5.593f, 0.0f,   5.488f, 0.0f,
    5.496f, 0.0f,   5.592f, 5.528f, 0.0f,   0.0f,   5.560f, 5.488f, 0.0f,   5.560f, 5.488f, 5.592f};
static const float kRoomColors[15] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 1, 0};
static const float kRoomClosure[12] = {5.528f, 0.0f, 0.0f, 0.0f,   0.0f,   0.0f,
                                0.0f,   5.488f, 0.0f, 5.560f, 5.488f, 0.0f};
static const float kCube[60] = {
    1.3f,  1.65f, 0.65f, 2.9f,  1.65f, 1.14f, 2.4f,  1.65f, 2.72f, 0.82f, 1.65f, 2.25f,
    2.9f,  0.0f,  1.14f, 2.4f,  0.0f,  2.72f, 2.4f,  1.65f, 2.72f, 2.9f,  1.65f, 1.14f,
    1.3f,  0.0f,  0.65f, 2.9f,  0.0f,  1.14f, 2.9f,  1.65f, 1.14f, 1.3f,  1.65f, 0.65f,
    0.82f, 0.0f,  2.25f, 1.3f,  0.0f,  0.65f, 1.3f,  1.65f, 0.65f, 0.82f, 1.65f, 2.25f,
    2.4f,  0.0f,  2.72f, 0.82f, 0.0f,  2.25f, 0.82f, 1.65f, 2.25f, 2.4f,  1.65f, 2.72f};
static const float kBlock[60] = {
    4.23f, 3.3f, 2.47f, 4.72f, 3.3f, 4.06f, 3.14f, 3.3f, 4.56f, 2.65f, 3.3f, 2.96f,
    4.23f, 0.0f, 2.47f, 4.72f, 0.0f, 4.06f, 4.72f, 3.3f, 4.06f, 4.23f, 3.3f, 2.47f,
    4.72f, 0.0f, 4.06f, 3.14f, 0.0f, 4.56f, 3.14f, 3.3f, 4.56f, 4.72f, 3.3f, 4.06f,
    3.14f, 0.0f, 4.56f, 2.65f, 0.0f, 2.96f, 2.65f, 3.3f, 2.96f, 3.14f, 3.3f, 4.56f,
    2.65f, 0.0f, 2.96f, 4.23f, 0.0f, 2.47f, 4.23f, 3.3f, 2.47f, 2.65f, 3.3f, 2.96f};

inline Quad quad_from(const float *p, V3 color)
{
    Quad q;
    for (int k = 0; k < 4; ++k)
        q.v[k] = {p[3 * k], p[3 * k + 1], p[3 * k + 2]};
    q.color = color;
    q.illum = {0, 0, 0};
    q.rad = {0, 0, 0};
    q.lightish = 0;
    q.group = 0;
    return q;
}

// Patch::getCenter / getNormal / getUp (Patch.cpp:253-276) of a quad.
inline void quad_frame(const V3 *v, float fr[9])
{
    V3 eye = {(v[0].x + v[1].x + v[2].x + v[3].x) / 4.0f, (v[0].y + v[1].y + v[2].y + v[3].y) / 4.0f,
              (v[0].z + v[1].z + v[2].z + v[3].z) / 4.0f};
    V3 n = vcross(v[1] - v[0], v[3] - v[0]);
    V3 up = v[3] - v[0];
    float f[9] = {eye.x, eye.y, eye.z, n.x, n.y, n.z, up.x, up.y, up.z};
    for (int k = 0; k < 9; ++k)
        fr[k] = f[k];
}

// Frame and area of a quad patch (Patch.cpp:253-276; the area is the sum of
// its two drawn triangles) and of a triangle patch (centroid, the quad rule
// applied to the triangle's edges: normal (B - A) x (C - A), up C - A).
inline float tri_area(V3 a, V3 b, V3 c) { return (float)tri_area_d(a, b, c); }

inline void quad_patch(const V3 *v, float fr[9], float &area)
{
    quad_frame(v, fr);
    area = tri_area(v[0], v[1], v[2]) + tri_area(v[0], v[2], v[3]);
}

inline void tri_patch(const V3 *t, float fr[9], float &area)
{
    V3 eye = {(t[0].x + t[1].x + t[2].x) / 3.0f, (t[0].y + t[1].y + t[2].y) / 3.0f,
              (t[0].z + t[1].z + t[2].z) / 3.0f};
    V3 n = vcross(t[1] - t[0], t[2] - t[0]);
    V3 up = t[2] - t[0];
    float f[9] = {eye.x, eye.y, eye.z, n.x, n.y, n.z, up.x, up.y, up.z};
    for (int k = 0; k < 9; ++k)
        fr[k] = f[k];
    area = tri_area(t[0], t[1], t[2]);
}

// The source models of ModelContainer::load (ModelContainer.cpp:35-48) before
// subdivision: ROOM (+ the light quad), ROOMCLOSURE, CUBE, BLOCK.
inline std::vector<std::vector<Quad>> cornell_models()
{
    std::vector<Quad> room;
    for (int i = 0; i < 5; ++i)
        room.push_back(quad_from(kRoom + 12 * i, {kRoomColors[3 * i], kRoomColors[3 * i + 1],
                                                   kRoomColors[3 * i + 2]}));
    {
        // PrimitiveModel.cpp:22-28: the light, radiosity 100 and illumination 1
        const float l[12] = {3.430f, 5.485f, 2.270f, 2.130f, 5.485f, 2.270f,
                             2.130f, 5.485f, 3.320f, 3.430f, 5.485f, 3.320f};
        Quad q = quad_from(l, {1.0f, 1.0f, 1.0f});
        q.illum = {1.0f, 1.0f, 1.0f};
        q.rad = {100.0f, 100.0f, 100.0f};
        q.lightish = 1;
        room.push_back(q);
    }
    std::vector<Quad> closure = {quad_from(kRoomClosure, {1, 1, 1})};
    std::vector<Quad> cube, block;
    for (int i = 0; i < 5; ++i) {
        cube.push_back(quad_from(kCube + 12 * i, {1, 1, 1}));
        block.push_back(quad_from(kBlock + 12 * i, {1, 1, 1}));
    }
    return {room, closure, cube, block};
}

} // namespace radpatch
