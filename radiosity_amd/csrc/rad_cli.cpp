// rad_cli.cpp - headless command line of the drop-in layer.
//
// Restates main + OnIdle (Main.cpp:761-785, 1137-1309) without the window:
// the same `name value` argument pairs (area, hemicube, shoots, hemicubes),
// Config::freeze, the Cornell scene of ModelContainer::load, then shooting
// cycles until the last shooter's energy falls below 0.1 (Main.cpp:1298).
// Extra pairs: `cycles N` caps the cycles (default 1000), `json 1` prints a
// summary line, `obj PATH` loads a WaveFrontModel instead of the Cornell box
// (ModelContainer.cpp:50-54; its patches are black, as the reference makes
// them), `ply PATH` writes the smoothed vertex colours of the result
// (Colors::smoothShadePatch over every patch, on the device) as a coloured
// mesh, and `ppm PATH` writes a flat-shaded render of the result from
// `view ex,ey,ez,dx,dy,dz,ux,uy,uz` (default: inside the Cornell box's open
// side looking in), both tone-mapped with `exposure E` (default: the 99th
// percentile of the vertex colours maps to white).
//
//   radhip_cli area 0.3127 hemicube 128 shoots 10 hemicubes 1
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/radhip_dropin.hpp"

using namespace radhip;

namespace {

unsigned char tone(float v, float exposure)
{
    float x = std::max(0.0f, v * exposure);
    x = std::pow(std::min(x, 1.0f), 1.0f / 2.2f);
    return (unsigned char)std::lround(255.0f * x);
}

// Binary little-endian PLY: four vertices per patch with their smoothed
// colours, one quad face each.
bool write_ply(const char *path, ModelContainer &scene, const std::vector<float> &col,
               float exposure)
{
    FILE *f = fopen(path, "wb");
    if (!f)
        return false;
    unsigned n = scene.getPatchesCount();
    fprintf(f, "ply\nformat binary_little_endian 1.0\nelement vertex %u\n"
               "property float x\nproperty float y\nproperty float z\n"
               "property uchar red\nproperty uchar green\nproperty uchar blue\n"
               "element face %u\nproperty list uchar int vertex_indices\nend_header\n",
            4 * n, n);
    const float *v = scene.getVertices();
    for (unsigned i = 0; i < 4 * n; ++i) {
        fwrite(v + 3 * (size_t)i, 4, 3, f);
        unsigned char rgb[3] = {tone(col[3 * (size_t)i], exposure),
                                tone(col[3 * (size_t)i + 1], exposure),
                                tone(col[3 * (size_t)i + 2], exposure)};
        fwrite(rgb, 1, 3, f);
    }
    for (unsigned i = 0; i < n; ++i) {
        unsigned char four = 4;
        int idx[4] = {(int)(4 * i), (int)(4 * i + 1), (int)(4 * i + 2), (int)(4 * i + 3)};
        fwrite(&four, 1, 1, f);
        fwrite(idx, 4, 4, f);
    }
    return fclose(f) == 0;
}

// The FRONT face of a viewpoint's item buffer, each pixel the colour *
// (illumination + radiosity) of the patch it shows.
bool write_ppm(const char *path, ModelContainer &scene, const float view[9], float exposure)
{
    rad_ctx *c = scene.context();
    unsigned tw, th, res;
    rad_geometry(c, &tw, &th, &res);
    std::vector<uint32_t> ids(res);
    if (rad_render_view(c, view, ids.data()) != RAD_OK) {
        fprintf(stderr, "radhip: %s\n", rad_last_error(c));
        return false;
    }
    unsigned W = Config::HEMICUBE_W();
    FILE *f = fopen(path, "wb");
    if (!f)
        return false;
    fprintf(f, "P6\n%u %u\n255\n", W, W);
    Patch **ps = scene.getPatches();
    for (unsigned r = 0; r < W; ++r) {
        unsigned y = W - 1 - r; // item-buffer rows run bottom-up
        for (unsigned x = 0; x < W; ++x) {
            uint32_t id = ids[(size_t)y * tw + W / 2 + x];
            unsigned char rgb[3] = {0, 0, 0};
            if (id) {
                Patch *p = ps[id - 1];
                Vector3f t = p->getColor() * (p->illumination + p->radiosity);
                rgb[0] = tone(t.x, exposure);
                rgb[1] = tone(t.y, exposure);
                rgb[2] = tone(t.z, exposure);
            }
            fwrite(rgb, 1, 3, f);
        }
    }
    return fclose(f) == 0;
}

} // namespace

int main(int argc, const char **argv)
{
    if ((argc - 1) % 2 != 0) {
        fprintf(stderr, "wrong number of parameters: expected name value pairs "
                        "(area, hemicube, shoots, hemicubes, cycles, json, obj, ply, ppm, "
                        "view, exposure)\n");
        return 2;
    }
    unsigned max_cycles = 1000;
    bool json = false;
    const char *obj = nullptr, *ply = nullptr, *ppm = nullptr;
    float view[9] = {2.78f, 2.73f, 0.2f, 0.0f, 0.0f, 1.0f, 0.0f, 1.0f, 0.0f};
    float exposure = 0.0f;
    for (int i = 1; i < argc; i += 2) {
        const char *k = argv[i], *v = argv[i + 1];
        if (strcmp(k, "area") == 0)
            Config::setMaxPatchArea(atof(v));
        else if (strcmp(k, "hemicube") == 0)
            Config::setHemicubeSide((unsigned)atoi(v));
        else if (strcmp(k, "shoots") == 0)
            Config::setShootsPerCycle((unsigned)atoi(v));
        else if (strcmp(k, "hemicubes") == 0)
            Config::setHemicubesCount((unsigned)atoi(v));
        else if (strcmp(k, "cycles") == 0)
            max_cycles = (unsigned)atoi(v);
        else if (strcmp(k, "json") == 0)
            json = atoi(v) != 0;
        else if (strcmp(k, "obj") == 0)
            obj = v;
        else if (strcmp(k, "ply") == 0)
            ply = v;
        else if (strcmp(k, "ppm") == 0)
            ppm = v;
        else if (strcmp(k, "exposure") == 0)
            exposure = (float)atof(v);
        else if (strcmp(k, "view") == 0) {
            if (sscanf(v, "%f,%f,%f,%f,%f,%f,%f,%f,%f", &view[0], &view[1], &view[2], &view[3],
                       &view[4], &view[5], &view[6], &view[7], &view[8]) != 9) {
                fprintf(stderr, "view takes 9 comma-separated numbers\n");
                return 2;
            }
        }
        else {
            fprintf(stderr, "unknown parameter %s\n", k);
            return 2;
        }
    }
    Config::freeze();
    ModelContainer scene;
    scene.maxPatchArea = Config::MAX_PATCH_AREA();
    if (obj)
        scene.addModel(new WaveFrontModel(obj));
    else
        scene.load();
    if (obj && !lastError().empty()) {
        fprintf(stderr, "radhip: %s\n", lastError().c_str());
        return 1;
    }
    scene.updateData();
    if (!scene.context()) {
        fprintf(stderr, "radhip: %s\n", lastError().c_str());
        return 1;
    }
    printf("patches %u, hemicube %u, %u hemicubes per pass, %u passes per cycle\n",
           scene.getPatchesCount(), Config::HEMICUBE_W(), Config::HEMICUBES_CNT(),
           Config::SHOOTS_PER_CYCLE());
    Shooter shooter(&scene);
    auto t0 = std::chrono::steady_clock::now();
    unsigned cycles = 0;
    float energy = 0.0f;
    while (cycles < max_cycles) {
        energy = shooter.shootCycle();
        if (shooter.failed()) {
            fprintf(stderr, "radhip: %s\n", lastError().c_str());
            return 1;
        }
        ++cycles;
        if (shooter.finished())
            break;
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    double res = shooter.residual();
    printf("Done in %.3f seconds, %llu shooters (%u cycles), last energy %.4g, residual %.4g\n",
           secs, shooter.shooters(), cycles, energy, res);
    if (ply || ppm) {
        unsigned n = scene.getPatchesCount();
        std::vector<float> col((size_t)n * 12);
        if (!scene.smoothShadeAll(col.data())) {
            fprintf(stderr, "radhip: %s\n", lastError().c_str());
            return 1;
        }
        if (!(exposure > 0.0f)) {
            std::vector<float> v(col);
            size_t k = v.size() * 99 / 100;
            std::nth_element(v.begin(), v.begin() + k, v.end());
            exposure = v[k] > 0.0f ? 1.0f / v[k] : 1.0f;
        }
        if (ply && !write_ply(ply, scene, col, exposure)) {
            fprintf(stderr, "radhip: cannot write %s\n", ply);
            return 1;
        }
        if (ppm && !write_ppm(ppm, scene, view, exposure))
            return 1;
    }
    if (json) {
        // sum of the final radiosity (illumination + radiosity) per channel
        double sum[3] = {0, 0, 0};
        Patch **ps = scene.getPatches();
        for (unsigned i = 0; i < scene.getPatchesCount(); ++i) {
            Vector3f t = ps[i]->illumination + ps[i]->radiosity;
            sum[0] += t.x;
            sum[1] += t.y;
            sum[2] += t.z;
        }
        printf("{\"patches\": %u, \"shooters\": %llu, \"cycles\": %u, \"seconds\": %.6f, "
               "\"last_energy\": %.9g, \"residual\": %.9g, \"sum_rgb\": [%.9g, %.9g, %.9g]}\n",
               scene.getPatchesCount(), shooter.shooters(), cycles, secs, energy, res, sum[0],
               sum[1], sum[2]);
    }
    return 0;
}
