// rad_scene.cpp - scene sources (host code): the reference's Cornell data
// with its quad subdivision, and seeded random-box rooms.
//
// Compiled with -ffp-contract=off so the float subdivision arithmetic is the
// written sequence (Patch.cpp:47-125).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "../../include/radhip.h"
#include "../../include/radhip_scene.h"
#include "rad_patch.h"

using namespace radpatch;

namespace {

uint64_t count_tris(const std::vector<Quad> &in, double area)
{
    uint64_t n = 0;
    std::vector<Quad> tmp;
    for (const Quad &q : in) {
        tmp.clear();
        if (divide(q, area, tmp))
            n += 2 * tmp.size();
        else
            n += 2;
    }
    return n;
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float uni(float lo, float hi) { return lo + (hi - lo) * (float)((next() >> 40) * (1.0 / 16777216.0)); }
};

rad_scene *alloc_scene(uint32_t ntri, uint32_t npatch, uint32_t nquad)
{
    rad_scene *s = (rad_scene *)calloc(1, sizeof(rad_scene));
    if (!s)
        return nullptr;
    s->ntri = ntri;
    s->npatch = npatch;
    s->nquad = nquad;
    s->tri = (float *)calloc((size_t)ntri * 9 + 1, sizeof(float));
    s->tri_patch = (uint32_t *)calloc((size_t)ntri + 1, sizeof(uint32_t));
    s->frame = (float *)calloc((size_t)npatch * 9 + 1, sizeof(float));
    s->refl = (float *)calloc((size_t)npatch * 3 + 1, sizeof(float));
    s->area = (float *)calloc((size_t)npatch + 1, sizeof(float));
    s->emission = (float *)calloc((size_t)npatch * 3 + 1, sizeof(float));
    s->shot0 = (float *)calloc((size_t)npatch * 3 + 1, sizeof(float));
    s->patch_quad = (uint32_t *)calloc((size_t)npatch + 1, sizeof(uint32_t));
    s->quad_nbr = (uint32_t *)calloc((size_t)nquad * 8 + 1, sizeof(uint32_t));
    if (!s->tri || !s->tri_patch || !s->frame || !s->refl || !s->area || !s->emission ||
        !s->shot0 || !s->patch_quad || !s->quad_nbr) {
        rad_scene_free(s);
        return nullptr;
    }
    return s;
}

// Flattens quads like ModelContainer::updateData (ModelContainer.cpp:81-155):
// triangles (0,1,2), (0,2,3) per quad.  Per-patch values come from `fill`;
// nbr (8 per quad, or empty) becomes quad_nbr.
template <typename Fill>
rad_scene *flatten(const std::vector<Quad> &quads, bool tri_patches, Fill fill,
                   const std::vector<unsigned> &nbr)
{
    uint32_t nq = (uint32_t)quads.size();
    uint32_t np = tri_patches ? 2 * nq : nq;
    rad_scene *s = alloc_scene(2 * nq, np, nq);
    if (!s)
        return nullptr;
    for (uint32_t q = 0; q < nq; ++q) {
        const Quad &Q = quads[q];
        const V3 *v = Q.v;
        V3 t0[3] = {v[0], v[1], v[2]}, t1[3] = {v[0], v[2], v[3]};
        for (int k = 0; k < 3; ++k) {
            memcpy(&s->tri[(size_t)(2 * q) * 9 + 3 * k], &t0[k], 12);
            memcpy(&s->tri[(size_t)(2 * q + 1) * 9 + 3 * k], &t1[k], 12);
        }
        for (int k = 0; k < 8; ++k)
            s->quad_nbr[8 * (size_t)q + k] = nbr.empty() ? q : nbr[8 * (size_t)q + k];
        if (tri_patches) {
            s->tri_patch[2 * q] = 2 * q;
            s->tri_patch[2 * q + 1] = 2 * q + 1;
            for (int h = 0; h < 2; ++h) {
                uint32_t p = 2 * q + h;
                tri_patch(h ? t1 : t0, &s->frame[(size_t)p * 9], s->area[p]);
                s->patch_quad[p] = q;
                fill(s, p, Q, h);
            }
        } else {
            s->tri_patch[2 * q] = q;
            s->tri_patch[2 * q + 1] = q;
            quad_patch(v, &s->frame[(size_t)q * 9], s->area[q]);
            s->patch_quad[q] = q;
            fill(s, q, Q, 0);
        }
    }
    return s;
}

// Each model's quads subdivided like Model::getPatches(area)
// (PrimitiveModel.cpp:79-84), concatenated in ModelContainer::load order,
// with the neighbour links of every quad.
std::vector<Quad> cornell_quads(double max_area, bool keep_light, std::vector<unsigned> &nbr)
{
    std::vector<Quad> all;
    for (const std::vector<Quad> &m : cornell_models()) {
        std::vector<unsigned> nb;
        std::vector<Quad> d = subdivide(m, max_area, keep_light, &nb);
        for (unsigned x : nb)
            nbr.push_back(x + (unsigned)all.size());
        all.insert(all.end(), d.begin(), d.end());
    }
    return all;
}

// Axis-aligned face as a quad with normal = (v4-v1) x (v2-v1) along n.
Quad face_quad(V3 o, V3 e1, V3 e2, V3 n, int group)
{
    V3 c = cross(e2, e1);
    if (dot(c, n) < 0) {
        V3 t = e1;
        e1 = e2;
        e2 = t;
    }
    Quad q;
    q.v[0] = o;
    q.v[1] = o + e1;
    q.v[2] = o + e1 + e2;
    q.v[3] = o + e2;
    q.color = {1, 1, 1};
    q.illum = {0, 0, 0};
    q.rad = {0, 0, 0};
    q.lightish = 0;
    q.group = group;
    return q;
}

// Child c of face f (plan pl): its corners, four even for a triangle
// (the last one repeated), and its triangles (2, or 1 for a native one).
int face_child(const float *fv, const FacePlan &pl, uint32_t c, V3 *v)
{
    V3 src[4] = {face_vertex(fv, 0), face_vertex(fv, 1), face_vertex(fv, 2), face_vertex(fv, 3)};
    if (pl.native) {
        tri_grid_child(src[0], src[1], src[2], pl.kx, c, v);
        v[3] = v[2];
        return 1;
    }
    if (pl.kx * pl.ky == 1) {
        for (int k = 0; k < 4; ++k)
            v[k] = src[k];
    } else {
        divide_child(src, pl.kx, pl.ky, c, v);
    }
    return 2;
}

} // namespace

extern "C" {

void rad_scene_free(rad_scene *s)
{
    if (!s)
        return;
    free(s->tri);
    free(s->tri_patch);
    free(s->frame);
    free(s->refl);
    free(s->area);
    free(s->emission);
    free(s->shot0);
    free(s->patch_quad);
    free(s->quad_nbr);
    free(s);
}

int rad_scene_cornell(double max_area, int triangle_patches, int physical, uint64_t seed,
                      rad_scene **out)
{
    if (!out || !(max_area >= 0.0))
        return RAD_E_ARG;
    *out = nullptr;
    std::vector<unsigned> nbr;
    std::vector<Quad> quads = cornell_quads(max_area, physical != 0, nbr);
    Rng rng(seed);
    rad_scene *s = flatten(quads, triangle_patches != 0, [&](rad_scene *sc, uint32_t p, const Quad &q, int) {
        if (physical) {
            sc->refl[3 * p + 0] = rng.uni(0.1f, 0.8f);
            sc->refl[3 * p + 1] = rng.uni(0.1f, 0.8f);
            sc->refl[3 * p + 2] = rng.uni(0.1f, 0.8f);
            if (q.lightish) {
                sc->emission[3 * p + 0] = 17.0f;
                sc->emission[3 * p + 1] = 12.0f;
                sc->emission[3 * p + 2] = 4.0f;
            }
        } else {
            sc->refl[3 * p + 0] = q.color.x;
            sc->refl[3 * p + 1] = q.color.y;
            sc->refl[3 * p + 2] = q.color.z;
            sc->emission[3 * p + 0] = q.rad.x;
            sc->emission[3 * p + 1] = q.rad.y;
            sc->emission[3 * p + 2] = q.rad.z;
            sc->shot0[3 * p + 0] = q.illum.x;
            sc->shot0[3 * p + 1] = q.illum.y;
            sc->shot0[3 * p + 2] = q.illum.z;
        }
    }, nbr);
    if (!s)
        return RAD_E_NOMEM;
    *out = s;
    return RAD_OK;
}

int rad_scene_boxes(uint32_t nboxes, float rx, float ry, float rz, uint32_t target_tris,
                    uint32_t nlights, uint64_t seed, rad_scene **out)
{
    if (!out || !(rx > 0 && ry > 0 && rz > 0) || nboxes > 100000)
        return RAD_E_ARG;
    *out = nullptr;
    Rng rng(seed);
    std::vector<Quad> base;
    // room, inward normals; group 1 = ceiling
    base.push_back(face_quad({0, 0, 0}, {rx, 0, 0}, {0, 0, rz}, {0, 1, 0}, 0));   // floor
    base.push_back(face_quad({0, ry, 0}, {rx, 0, 0}, {0, 0, rz}, {0, -1, 0}, 1)); // ceiling
    base.push_back(face_quad({0, 0, 0}, {rx, 0, 0}, {0, ry, 0}, {0, 0, 1}, 0));
    base.push_back(face_quad({0, 0, rz}, {rx, 0, 0}, {0, ry, 0}, {0, 0, -1}, 0));
    base.push_back(face_quad({0, 0, 0}, {0, 0, rz}, {0, ry, 0}, {1, 0, 0}, 0));
    base.push_back(face_quad({rx, 0, 0}, {0, 0, rz}, {0, ry, 0}, {-1, 0, 0}, 0));
    struct Box {
        float x0, y0, z0, x1, y1, z1;
    };
    std::vector<Box> boxes;
    for (uint32_t b = 0; b < nboxes; ++b) {
        Box bx{};
        bool placed = false;
        for (int attempt = 0; attempt < 200 && !placed; ++attempt) {
            float sx = rng.uni(0.2f, 2.0f), sy = rng.uni(0.2f, 2.0f), sz = rng.uni(0.2f, 2.0f);
            sx = std::min(sx, rx * 0.9f);
            sy = std::min(sy, ry * 0.9f);
            sz = std::min(sz, rz * 0.9f);
            float x0 = rng.uni(0.0f, rx - sx), y0 = rng.uni(0.0f, ry - sy), z0 = rng.uni(0.0f, rz - sz);
            bx = {x0, y0, z0, x0 + sx, y0 + sy, z0 + sz};
            placed = true;
            for (const Box &o : boxes)
                if (bx.x0 < o.x1 + 0.01f && o.x0 < bx.x1 + 0.01f && bx.y0 < o.y1 + 0.01f &&
                    o.y0 < bx.y1 + 0.01f && bx.z0 < o.z1 + 0.01f && o.z0 < bx.z1 + 0.01f) {
                    placed = false;
                    break;
                }
            if (attempt == 199)
                placed = true; // crowded room: accept an overlap
        }
        boxes.push_back(bx);
        V3 o = {bx.x0, bx.y0, bx.z0};
        float sx = bx.x1 - bx.x0, sy = bx.y1 - bx.y0, sz = bx.z1 - bx.z0;
        base.push_back(face_quad(o, {sx, 0, 0}, {0, 0, sz}, {0, -1, 0}, 2));
        base.push_back(face_quad(o + V3{0, sy, 0}, {sx, 0, 0}, {0, 0, sz}, {0, 1, 0}, 2));
        base.push_back(face_quad(o, {sx, 0, 0}, {0, sy, 0}, {0, 0, -1}, 2));
        base.push_back(face_quad(o + V3{0, 0, sz}, {sx, 0, 0}, {0, sy, 0}, {0, 0, 1}, 2));
        base.push_back(face_quad(o, {0, 0, sz}, {0, sy, 0}, {-1, 0, 0}, 2));
        base.push_back(face_quad(o + V3{sx, 0, 0}, {0, 0, sz}, {0, sy, 0}, {1, 0, 0}, 2));
    }
    // max patch area giving close to target_tris (count decreases with area)
    double lo = 1e-6, hi = 1e3;
    if (target_tris > 2 * base.size()) {
        for (int it = 0; it < 60; ++it) {
            double mid = sqrt(lo * hi);
            if (count_tris(base, mid) > target_tris)
                lo = mid;
            else
                hi = mid;
        }
    } else {
        hi = 0;
    }
    std::vector<unsigned> nbr;
    std::vector<Quad> quads = subdivide(base, hi, false, &nbr);
    // lights: nlights distinct ceiling quads, both triangles of the first
    // half of picks ... one triangle per pick
    std::vector<uint32_t> ceiling;
    for (uint32_t q = 0; q < quads.size(); ++q)
        if (quads[q].group == 1)
            ceiling.push_back(q);
    std::vector<uint32_t> picks;
    Rng lr(seed ^ 0x5eed1167ull);
    for (uint32_t l = 0; l < nlights && !ceiling.empty(); ++l) {
        uint32_t i = (uint32_t)(lr.next() % ceiling.size());
        picks.push_back(2 * ceiling[i] + (uint32_t)(lr.next() & 1));
        ceiling.erase(ceiling.begin() + i);
    }
    std::sort(picks.begin(), picks.end());
    rad_scene *s = flatten(quads, true, [&](rad_scene *sc, uint32_t p, const Quad &, int) {
        sc->refl[3 * p + 0] = rng.uni(0.1f, 0.8f);
        sc->refl[3 * p + 1] = rng.uni(0.1f, 0.8f);
        sc->refl[3 * p + 2] = rng.uni(0.1f, 0.8f);
        if (std::binary_search(picks.begin(), picks.end(), p)) {
            sc->emission[3 * p + 0] = lr.uni(5.0f, 20.0f);
            sc->emission[3 * p + 1] = lr.uni(5.0f, 20.0f);
            sc->emission[3 * p + 2] = lr.uni(5.0f, 20.0f);
        }
    }, nbr);
    if (!s)
        return RAD_E_NOMEM;
    *out = s;
    return RAD_OK;
}

int rad_cornell_faces(float *quads, float *color, float *emission, float *shot0,
                      uint32_t *nface)
{
    if (!nface)
        return RAD_E_ARG;
    uint32_t n = 0;
    for (const std::vector<Quad> &m : cornell_models())
        for (const Quad &q : m) {
            if (quads)
                for (int k = 0; k < 4; ++k)
                    memcpy(quads + 12 * (size_t)n + 3 * k, &q.v[k], 12);
            if (color)
                memcpy(color + 3 * (size_t)n, &q.color, 12);
            if (emission)
                memcpy(emission + 3 * (size_t)n, &q.rad, 12);
            if (shot0)
                memcpy(shot0 + 3 * (size_t)n, &q.illum, 12);
            ++n;
        }
    *nface = n;
    return RAD_OK;
}

int rad_scene_faces(const float *quads, uint32_t nface, const float *color, const float *emission,
                    const float *shot0, double max_area, int mode, rad_scene **out)
{
    if (!out)
        return RAD_E_ARG;
    *out = nullptr;
    if (!quads || nface == 0 || !(max_area >= 0.0) || (mode != kFacesQuads && mode != kFacesTriangles))
        return RAD_E_ARG;
    for (uint64_t i = 0; i < (uint64_t)nface * 12; ++i)
        if (!std::isfinite(quads[i]) || std::fabs(quads[i]) > 1.0e4f)
            return RAD_E_ARG;
    std::vector<FacePlan> plan;
    uint32_t nq, nt, np;
    if (!plan_faces(quads, nface, max_area, mode, plan, nq, nt, np))
        return RAD_E_LIMIT;
    if (np == 0) // Patch::divide yields no children for a quad with a zero-length edge
        return RAD_E_ARG;
    rad_scene *s = alloc_scene(nt, np, nq);
    if (!s)
        return RAD_E_NOMEM;
    for (uint32_t f = 0; f < nface; ++f) {
        const FacePlan &pl = plan[f];
        const float *fv = quads + 12 * (size_t)f;
        uint32_t kids = pl.native ? pl.kx * pl.kx : pl.kx * pl.ky;
        for (uint32_t c = 0; c < kids; ++c) {
            V3 v[4];
            int ntc = face_child(fv, pl, c, v);
            uint32_t q = pl.qbeg + c;
            unsigned nb[8];
            if (pl.native)
                for (int k = 0; k < 8; ++k)
                    nb[k] = c;
            else
                divide_neighbours(pl.kx, pl.ky, c, nb);
            for (int k = 0; k < 8; ++k)
                s->quad_nbr[8 * (size_t)q + k] = pl.qbeg + nb[k];
            V3 t[2][3] = {{v[0], v[1], v[2]}, {v[0], v[2], v[3]}};
            uint32_t t0 = pl.tbeg + (uint32_t)ntc * c;
            uint32_t p0 = mode == kFacesQuads ? pl.pbeg + c : pl.pbeg + (uint32_t)ntc * c;
            for (int h = 0; h < ntc; ++h) {
                memcpy(&s->tri[(size_t)(t0 + h) * 9], t[h], 36);
                uint32_t p = mode == kFacesQuads ? p0 : p0 + h;
                s->tri_patch[t0 + h] = p;
                if (mode == kFacesQuads && h == 1)
                    continue;
                if (mode == kFacesQuads)
                    quad_patch(v, &s->frame[(size_t)p * 9], s->area[p]);
                else
                    tri_patch(t[h], &s->frame[(size_t)p * 9], s->area[p]);
                s->patch_quad[p] = q;
                // Patch.cpp:119-122: children inherit colour, illumination, radiosity
                for (int k = 0; k < 3; ++k) {
                    s->refl[3 * (size_t)p + k] = color ? color[3 * (size_t)f + k] : 0.0f;
                    s->emission[3 * (size_t)p + k] = emission ? emission[3 * (size_t)f + k] : 0.0f;
                    s->shot0[3 * (size_t)p + k] = shot0 ? shot0[3 * (size_t)f + k] : 0.0f;
                }
            }
        }
    }
    *out = s;
    return RAD_OK;
}

} // extern "C"
