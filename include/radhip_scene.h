/*
 * radhip_scene.h - scene sources for libradhip (host code, no GPU needed).
 *
 * These produce the arrays rad_scene_upload takes: the reference's
 * PrimitiveModel Cornell data subdivided by Patch::divide / Model::subdivide
 * and flattened like ModelContainer::updateData, and seeded random-box rooms
 * for the larger workloads.
 */
#ifndef RADHIP_SCENE_H
#define RADHIP_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rad_scene {
    uint32_t ntri;        /* triangles */
    uint32_t npatch;      /* patches (ids) */
    uint32_t nquad;       /* quads after subdivision */
    uint32_t reserved;
    float *tri;           /* ntri*9: v0, v1, v2 */
    uint32_t *tri_patch;  /* ntri: patch of each triangle, non-decreasing */
    float *frame;         /* npatch*9: eye, normal, up (Patch.cpp:253-276) */
    float *refl;          /* npatch*3: colour (REFERENCE) or rho (PHYSICAL) */
    float *area;          /* npatch */
    float *emission;      /* npatch*3: initial unshot (Patch::radiosity) */
    float *shot0;         /* npatch*3: initial shot (Patch::illumination) */
    uint32_t *patch_quad; /* npatch: quad after subdivision each patch comes from */
    uint32_t *quad_nbr;   /* nquad*8: neighbour quads of every quad (Patch::neighbours,
                             Patch.cpp:127-222): top-left, top, top-right, right,
                             bottom-right, bottom, bottom-left, left; itself where the
                             reference links a patch to itself */
} rad_scene;

/* Cornell box of ModelContainer::load (ModelContainer.cpp:35-48):
 * PrimitiveModel ROOM (+ light), ROOMCLOSURE, CUBE, BLOCK
 * (PrimitiveModel.cpp:4-213), each quad split by Patch::divide(max_area)
 * (Patch.cpp:47-125, Model.cpp:27-60).
 *  triangle_patches 0: patch = quad (two triangles share the id, as the
 *                      reference draws them); 1: every triangle is a patch
 *  physical         0: reference data - colours as reflectance tint, light
 *                      radiosity 100 and illumination 1 copied into every
 *                      child (Patch.cpp:119-122, PrimitiveModel.cpp:20-29)
 *                   1: rho ~ U[0.1,0.8]^3 per patch (seeded), the light quad
 *                      kept whole and emitting (17,12,4) */
int rad_scene_cornell(double max_area, int triangle_patches, int physical, uint64_t seed,
                      rad_scene **out);

/* Room room_x * room_y * room_z (y up) with nboxes axis-aligned boxes of
 * edge U[0.2,2.0] placed uniformly inside (seeded, non-overlapping where the
 * placement allows), subdivided so the triangle count is close to
 * target_tris; every triangle is a patch with rho ~ U[0.1,0.8]^3, and
 * nlights ceiling triangles emit U[5,20]^3. */
int rad_scene_boxes(uint32_t nboxes, float room_x, float room_y, float room_z,
                    uint32_t target_tris, uint32_t nlights, uint64_t seed, rad_scene **out);

void rad_scene_free(rad_scene *s);

/* Face modes of rad_scene_faces and rad_scene_ingest.
 *  RAD_FACES_QUADS     the reference's path: every face is split by
 *                      Patch::divide(max_area) (Patch.cpp:47-125) and every
 *                      quad is one patch drawn as (A,B,C), (A,C,D)
 *                      (ModelContainer.cpp:81-155); a triangle face is the
 *                      quad A B C C (WaveFrontModel.cpp:98-100)
 *  RAD_FACES_TRIANGLES every triangle is a patch: quads are split by
 *                      Patch::divide into two triangle patches each; a
 *                      triangle face (C == D bit for bit) is split natively
 *                      into n*n congruent triangles, n = ceil(sqrt(S/max_area))
 *                      when its area S > 1.01 max_area. */
#define RAD_FACES_QUADS 0
#define RAD_FACES_TRIANGLES 1

/* Wavefront OBJ file as source faces (WaveFrontModel::parse,
 * WaveFrontModel.cpp:15-136): "v x y z [w]" lines give the vertex
 * (x, y, z) / 1000; "f" lines with 3 or 4 vertex references (1-based; text
 * after a reference such as "/2/3" ignored) give a face, a triangle as the
 * quad A B C C; faces with other counts or undefined vertices are skipped;
 * all other lines are ignored.  *quads (nquad*12 floats: A, B, C, D) is
 * released with rad_obj_free.  RAD_E_ARG when the file cannot be opened
 * (nothing returned) or a "v" line has fewer than three numbers (the faces
 * before it are returned, as the reference keeps them). */
int rad_obj_parse(const char *path, float **quads, uint32_t *nquad);
void rad_obj_free(float *quads);

/* The 17 source quads of ModelContainer::load (ModelContainer.cpp:35-48,
 * PrimitiveModel.cpp:4-213, light included) with their colour, radiosity
 * and illumination; any output pointer may be null.  *nface = 17. */
int rad_cornell_faces(float *quads, float *color, float *emission, float *shot0,
                      uint32_t *nface);

/* Scene from source faces (nface*12 floats: A, B, C, D each) on the host:
 * Model::subdivide(max_area) (Model.cpp:27-60) per face, flattened like
 * ModelContainer::updateData, with the neighbour links of
 * Patch.cpp:127-222 (natively split triangles link to themselves).
 * color, emission and shot0 (nface*3 each, may be null = 0; the reference's
 * OBJ patches are black) are inherited by a face's children
 * (Patch.cpp:119-122).  Equals rad_scene_cornell(..., physical 0) on the
 * Cornell faces in RAD_FACES_QUADS mode.  rad_scene_ingest builds the same
 * arrays on the device. */
int rad_scene_faces(const float *quads, uint32_t nface, const float *color, const float *emission,
                    const float *shot0, double max_area, int mode, rad_scene **out);

#ifdef __cplusplus
}
#endif
#endif
