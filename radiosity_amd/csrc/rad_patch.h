// rad_patch.h - patch geometry shared by the scene sources and the C++
// drop-in layer (host code): vector helpers with the reference's operation
// order, quads, Patch::divide / Model::subdivide, and the PrimitiveModel
// Cornell data.  Compiled with -ffp-contract=off.
#pragma once

#include <math.h>

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

// Patch::divide (Patch.cpp:47-125): returns false when the quad is already
// small enough (S <= area * 1.01).
inline bool divide(const Quad &q, double area, std::vector<Quad> &out)
{
    V3 A = q.v[0], B = q.v[1], C = q.v[2], D = q.v[3];
    V3 u1 = A - C, u2 = B - D;
    double phi = (double)acosf(dot(u1, u2) / (length(u1) * length(u2)));
    double S = (double)(0.5f * length(u1) * length(u2)) * sin(phi);
    if (S <= (area * 1.01))
        return false;
    double a = sqrt(area);
    unsigned kx = (unsigned)ceil((double)length(B - A) / a);
    unsigned ky = (unsigned)ceil((double)length(C - B) / a);
    V3 pCD = divs(C - D, (float)kx);
    V3 pAB = divs(B - A, (float)kx);
    for (unsigned i = 0; i < kx * ky; ++i) {
        unsigned col = i % kx, row = i / kx;
        V3 bCD = pCD * (float)col + D;
        V3 bAB = pAB * (float)col + A;
        V3 nA = bAB + divs(bCD - bAB, (float)ky) * (float)row;
        V3 bCD1 = pCD * (float)(col + 1) + D;
        V3 bAB1 = pAB * (float)(col + 1) + A;
        V3 nB = bAB1 + divs(bCD1 - bAB1, (float)ky) * (float)row;
        V3 nC = bAB1 + divs(bCD1 - bAB1, (float)ky) * (float)(row + 1);
        V3 nD = bAB + divs(bCD - bAB, (float)ky) * (float)(row + 1);
        Quad c = q;
        c.v[0] = nA;
        c.v[1] = nB;
        c.v[2] = nC;
        c.v[3] = nD;
        out.push_back(c);
    }
    return true;
}

// Model::subdivide (Model.cpp:27-60): one pass, children replace parents in order.
inline std::vector<Quad> subdivide(const std::vector<Quad> &in, double area, bool keep_lights)
{
    std::vector<Quad> out;
    for (const Quad &q : in) {
        if (area > 0 && !(keep_lights && q.lightish) && divide(q, area, out))
            continue;
        out.push_back(q);
    }
    return out;
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
