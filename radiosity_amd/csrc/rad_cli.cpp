// rad_cli.cpp - headless command line of the drop-in layer.
//
// Restates main + OnIdle (Main.cpp:761-785, 1137-1309) without the window:
// the same `name value` argument pairs (area, hemicube, shoots, hemicubes),
// Config::freeze, the Cornell scene of ModelContainer::load, then shooting
// cycles until the last shooter's energy falls below 0.1 (Main.cpp:1298).
// Extra pairs: `cycles N` caps the cycles (default 1000), `json 1` prints a
// summary line.
//
//   radhip_cli area 0.3127 hemicube 128 shoots 10 hemicubes 1
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>

#include "../../include/radhip_dropin.hpp"

using namespace radhip;

int main(int argc, const char **argv)
{
    if ((argc - 1) % 2 != 0) {
        fprintf(stderr, "wrong number of parameters: expected name value pairs "
                        "(area, hemicube, shoots, hemicubes, cycles, json)\n");
        return 2;
    }
    unsigned max_cycles = 1000;
    bool json = false;
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
        else {
            fprintf(stderr, "unknown parameter %s\n", k);
            return 2;
        }
    }
    Config::freeze();
    ModelContainer scene;
    scene.maxPatchArea = Config::MAX_PATCH_AREA();
    scene.load();
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
        if (!lastError().empty() && energy == 0.0f) {
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
