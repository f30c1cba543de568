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

#ifdef __cplusplus
}
#endif
#endif
