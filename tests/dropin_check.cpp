// dropin_check.cpp - drives the C++ drop-in layer for the tests; every mode
// prints one JSON line.  Built next to the library by radiosity_amd.build.
//
//   dropin_check config W K WX SHOOTS AREA     Config::freeze arithmetic (CPU)
//   dropin_check patches AREA                  load + updateData arrays (CPU)
//   dropin_check scenes OBJPATH                scene builders and the OBJ reader (CPU)
//   dropin_check dff W K                       precomputeHemicubeFormFactors
//   dropin_check select AREA K SEED            list rule on seeded radiosities
//   dropin_check shoot AREA W K SHOOTS CYCLES  Shooter cycles (until finished()), final fields
//   dropin_check process AREA W                ProcessHemicube triplets
//   dropin_check resume AREA W PATH            shoot, saveToFile, loadFromFile
//   dropin_check obj PATH AREA                 WaveFrontModel patches and links (CPU)
//   dropin_check smooth AREA W CYCLES          smoothShadePatch vs smoothShadeAll
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <string>
#include <vector>

#include "radhip_dropin.hpp"
#include "radhip_scene.h"

using namespace radhip;

static void vec(const char *name, const std::vector<double> &v, bool last = false)
{
    printf("\"%s\": [", name);
    for (size_t i = 0; i < v.size(); ++i)
        printf(i ? ", %.9g" : "%.9g", v[i]);
    printf("]%s", last ? "" : ", ");
}

int main(int argc, char **argv)
{
    if (argc < 2)
        return 2;
    std::string mode = argv[1];
    if (mode == "config" && argc == 7) {
        Config::setHemicubeSide((unsigned)atoi(argv[2]));
        Config::setHemicubesCount((unsigned)atoi(argv[3]));
        Config::setOCLWorkitemsX((unsigned)atoi(argv[4]));
        Config::setShootsPerCycle((unsigned)atoi(argv[5]));
        Config::setMaxPatchArea(atof(argv[6]));
        Config::freeze();
        Config::setHemicubeSide(4); // frozen: ignored
        printf("{\"W\": %u, \"H\": %u, \"TW\": %u, \"TH\": %u, \"RES\": %u, \"WX\": %u, "
               "\"WY\": %u, \"shoots\": %u, \"K\": %u, \"area\": %.9g, \"frozen_err\": %d}\n",
               Config::HEMICUBE_W(), Config::HEMICUBE_H(), Config::PATCHVIEW_TEX_W(),
               Config::PATCHVIEW_TEX_H(), Config::PATCHVIEW_TEX_RES(), Config::OCL_WORKITEMS_X(),
               Config::OCL_WORKITEMS_Y(), Config::SHOOTS_PER_CYCLE(), Config::HEMICUBES_CNT(),
               Config::MAX_PATCH_AREA(), lastError().empty() ? 0 : 1);
        return 0;
    }
    if (mode == "scenes" && argc == 3) {
        // the host scene builders and the OBJ reader (CPU only): seeded box
        // rooms, the Cornell box both ways, the Cornell faces through
        // rad_scene_faces, and an OBJ file
        unsigned long long sum = 0;
        rad_scene *sc = nullptr;
        for (unsigned seed = 1; seed <= 3; ++seed) {
            if (rad_scene_boxes(20 * seed, 6.0f, 3.0f, 6.0f, 3000 * seed, 2, seed, &sc) != RAD_OK)
                return 1;
            sum += sc->ntri + sc->npatch;
            rad_scene_free(sc);
        }
        for (int tp = 0; tp < 2; ++tp) {
            if (rad_scene_cornell(0.3127, tp, tp, 1, &sc) != RAD_OK)
                return 1;
            sum += sc->ntri + sc->npatch + sc->nquad;
            rad_scene_free(sc);
        }
        uint32_t nf = 0;
        rad_cornell_faces(nullptr, nullptr, nullptr, nullptr, &nf);
        std::vector<float> q((size_t)nf * 12), c((size_t)nf * 3), e((size_t)nf * 3),
            sh((size_t)nf * 3);
        rad_cornell_faces(q.data(), c.data(), e.data(), sh.data(), &nf);
        for (int m = 0; m < 2; ++m) {
            if (rad_scene_faces(q.data(), nf, c.data(), e.data(), sh.data(), 0.05, m, &sc) != RAD_OK)
                return 1;
            sum += sc->ntri + sc->npatch + sc->nquad;
            rad_scene_free(sc);
        }
        float *oq = nullptr;
        uint32_t on = 0;
        int orc = rad_obj_parse(argv[2], &oq, &on);
        if (oq) {
            if (on && rad_scene_faces(oq, on, nullptr, nullptr, nullptr, 0.1, 1, &sc) == RAD_OK) {
                sum += sc->ntri;
                rad_scene_free(sc);
            }
            rad_obj_free(oq);
        }
        printf("{\"sum\": %llu, \"obj_faces\": %u, \"obj_rc\": %d}\n", sum, on, orc);
        return 0;
    }
    if (mode == "patches" && argc == 3) {
        ModelContainer mc;
        mc.maxPatchArea = atof(argv[2]);
        mc.load();
        unsigned n = mc.getPatchesCount();
        Patch **ps = mc.getPatches();
        std::vector<double> fr, col, rad, ill;
        for (unsigned i : {0u, 5u, n / 2, n - 1}) {
            Vector3f c = ps[i]->getCenter(), nn = ps[i]->getNormal(), u = ps[i]->getUp();
            for (float f : {c.x, c.y, c.z, nn.x, nn.y, nn.z, u.x, u.y, u.z})
                fr.push_back(f);
        }
        for (unsigned i = 0; i < n; ++i) {
            Vector3f c = ps[i]->getColor();
            col.insert(col.end(), {c.x, c.y, c.z});
            rad.push_back(ps[i]->radiosity.x);
            ill.push_back(ps[i]->illumination.x);
        }
        std::vector<double> v(mc.getVertices(), mc.getVertices() + mc.getVerticesCount());
        std::vector<double> ix(mc.getIndices(), mc.getIndices() + mc.getIndicesCount());
        printf("{\"n\": %u, \"nv\": %u, \"ni\": %u, ", n, mc.getVerticesCount(),
               mc.getIndicesCount());
        vec("frames", fr);
        vec("colors", col);
        vec("radiosity_r", rad);
        vec("illumination_r", ill);
        vec("vertices", v);
        vec("indices", ix, true);
        printf("}\n");
        return 0;
    }
    if (mode == "dff" && argc == 4) {
        Config::setHemicubeSide((unsigned)atoi(argv[2]));
        Config::setHemicubesCount((unsigned)atoi(argv[3]));
        Config::freeze();
        float *f = precomputeHemicubeFormFactors();
        if (!f) {
            fprintf(stderr, "%s\n", lastError().c_str());
            return 1;
        }
        size_t res = Config::PATCHVIEW_TEX_RES(), k = Config::HEMICUBES_CNT();
        bool same = true;
        for (size_t h = 1; h < k; ++h)
            same = same && memcmp(f, f + h * res, res * 4) == 0;
        std::vector<double> first(f, f + res);
        printf("{\"copies_equal\": %d, ", same ? 1 : 0);
        vec("table", first, true);
        printf("}\n");
        delete[] f;
        return 0;
    }
    if (mode == "obj" && argc == 4) {
        WaveFrontModel m(argv[2]);
        std::vector<Patch *> *ps = m.getPatches(atof(argv[3]));
        std::vector<double> v, nb;
        for (size_t i = 0; i < ps->size(); ++i) {
            std::vector<float> c = (*ps)[i]->getVerticesCoords();
            v.insert(v.end(), c.begin(), c.end());
            for (Patch *n : (*ps)[i]->neighbours) {
                size_t j = std::find(ps->begin(), ps->end(), n) - ps->begin();
                nb.push_back(j == ps->size() ? -1.0 : (double)j);
            }
        }
        Vector3f col = ps->empty() ? Vector3f() : (*ps)[0]->getColor();
        printf("{\"n\": %zu, \"error\": %d, \"color0\": [%g, %g, %g], ", ps->size(),
               lastError().empty() ? 0 : 1, col.x, col.y, col.z);
        vec("vertices", v);
        vec("neighbours", nb, true);
        printf("}\n");
        return 0;
    }
    if (mode == "smooth" && argc == 5) {
        Config::setHemicubeSide((unsigned)atoi(argv[3]));
        Config::setHemicubesCount(4);
        Config::setShootsPerCycle(5);
        Config::setMaxPatchArea(atof(argv[2]));
        Config::freeze();
        ModelContainer mc;
        mc.maxPatchArea = Config::MAX_PATCH_AREA();
        mc.load();
        mc.updateData();
        Shooter sh(&mc);
        for (int c = 0; c < atoi(argv[4]); ++c)
            sh.shootCycle();
        unsigned n = mc.getPatchesCount();
        std::vector<float> dev((size_t)n * 12), host((size_t)n * 12);
        bool ok = mc.smoothShadeAll(dev.data());
        for (unsigned i = 0; i < n; ++i)
            Colors::smoothShadePatch(&host[(size_t)i * 12], mc.getPatches()[i]);
        unsigned same = 0;
        double mx = 0;
        for (size_t i = 0; i < host.size(); ++i) {
            same += memcmp(&host[i], &dev[i], 4) == 0;
            mx = std::max(mx, (double)std::fabs(host[i]));
        }
        printf("{\"ok\": %d, \"n\": %u, \"values\": %zu, \"same\": %u, \"max\": %g}\n",
               ok ? 1 : 0, n, host.size(), same, mx);
        return 0;
    }
    if (mode == "resume" && argc == 5) {
        Config::setHemicubeSide((unsigned)atoi(argv[3]));
        Config::setHemicubesCount(4);
        Config::setShootsPerCycle(5);
        Config::setMaxPatchArea(atof(argv[2]));
        Config::freeze();
        ModelContainer a, b;
        a.maxPatchArea = b.maxPatchArea = Config::MAX_PATCH_AREA();
        a.load();
        b.load();
        a.updateData();
        b.updateData();
        Shooter sh(&a);
        sh.shootCycle();
        bool saved = a.saveToFile(argv[4]);
        bool loaded = b.loadFromFile(argv[4]);
        unsigned n = a.getPatchesCount(), same = 0;
        for (unsigned i = 0; i < n; ++i) {
            Patch *p = a.getPatches()[i], *q = b.getPatches()[i];
            same += memcmp(&p->radiosity, &q->radiosity, 12) == 0 &&
                    memcmp(&p->illumination, &q->illumination, 12) == 0;
        }
        printf("{\"saved\": %d, \"loaded\": %d, \"n\": %u, \"same\": %u}\n", saved ? 1 : 0,
               loaded ? 1 : 0, n, same);
        return 0;
    }
    if ((mode == "select" && argc == 5) || (mode == "shoot" && argc == 7) ||
        (mode == "process" && argc == 4)) {
        unsigned W = mode == "select" ? 16 : (unsigned)atoi(argv[3]);
        unsigned K = mode == "select" ? (unsigned)atoi(argv[3])
                                      : (mode == "shoot" ? (unsigned)atoi(argv[4]) : 1);
        Config::setHemicubeSide(W);
        Config::setHemicubesCount(K);
        if (mode == "shoot")
            Config::setShootsPerCycle((unsigned)atoi(argv[5]));
        Config::setMaxPatchArea(atof(argv[2]));
        Config::freeze();
        ModelContainer mc;
        mc.maxPatchArea = Config::MAX_PATCH_AREA();
        mc.load();
        mc.updateData();
        if (!mc.context()) {
            fprintf(stderr, "%s\n", lastError().c_str());
            return 1;
        }
        unsigned n = mc.getPatchesCount();
        Patch **ps = mc.getPatches();
        if (mode == "select") {
            std::mt19937 rng((unsigned)atoi(argv[4]));
            std::vector<double> rr;
            for (unsigned i = 0; i < n; ++i) {
                // a few distinct levels so ties and zeros occur
                float v = (float)(rng() % 7) * 0.5f;
                ps[i]->radiosity = {v, (float)(rng() % 3), 0.25f};
                if (rng() % 5 == 0)
                    ps[i]->radiosity = {0, 0, 0};
                rr.insert(rr.end(), {ps[i]->radiosity.x, ps[i]->radiosity.y, ps[i]->radiosity.z});
            }
            std::vector<Patch *> em(K);
            std::vector<unsigned> ids(K);
            mc.getHighestRadiosityPatchesId(K, em.data(), ids.data());
            std::vector<double> out(ids.begin(), ids.end()), nul;
            for (unsigned k = 0; k < K; ++k)
                nul.push_back(em[k] ? (double)(em[k] == ps[ids[k]]) : -1.0);
            printf("{\"first\": %u, ", mc.getHighestRadiosityPatchId());
            vec("ids", out);
            vec("ptr_ok", nul);
            vec("radiosity", rr, true);
            printf("}\n");
            return 0;
        }
        if (mode == "shoot") {
            Shooter sh(&mc);
            unsigned cycles = (unsigned)atoi(argv[6]), ran = 0;
            float e = 0;
            for (unsigned c = 0; c < cycles && !sh.finished(); ++c, ++ran) {
                e = sh.shootCycle();
                if (sh.failed()) {
                    fprintf(stderr, "%s\n", lastError().c_str());
                    return 1;
                }
            }
            std::vector<double> rad, ill;
            for (unsigned i = 0; i < n; ++i) {
                rad.insert(rad.end(), {ps[i]->radiosity.x, ps[i]->radiosity.y, ps[i]->radiosity.z});
                ill.insert(ill.end(),
                           {ps[i]->illumination.x, ps[i]->illumination.y, ps[i]->illumination.z});
            }
            printf("{\"shooters\": %llu, \"passes\": %llu, \"cycles\": %u, \"finished\": %d, "
                   "\"last_energy\": %.9g, \"residual\": %.9g, ",
                   sh.shooters(), sh.passes(), ran, sh.finished() ? 1 : 0, e, sh.residual());
            vec("radiosity", rad);
            vec("illumination", ill, true);
            printf("}\n");
            return 0;
        }
        // process: the light's hemicube through the C ABI, then the triplets
        rad_ctx *c = mc.context();
        unsigned tw, th, res;
        rad_geometry(c, &tw, &th, &res);
        unsigned light = 0;
        for (unsigned i = 0; i < n; ++i)
            if (ps[i]->radiosity.x > 0) {
                light = i;
                break;
            }
        std::vector<uint32_t> item(res);
        if (rad_render_itembuf(c, &light, 1, item.data(), nullptr) != RAD_OK) {
            fprintf(stderr, "%s\n", rad_last_error(c));
            return 1;
        }
        // output arrays as the reference allocates them (Main.cpp:605-607)
        std::vector<uint32_t> hc(res), idv(res);
        std::vector<float> en(res);
        unsigned m = ProcessHemicube(&mc, item.data(), 1, hc.data(), idv.data(), en.data());
        std::vector<double> ids(idv.begin(), idv.begin() + m), e(en.begin(), en.begin() + m),
            h(hc.begin(), hc.begin() + m), items(item.begin(), item.end());
        printf("{\"light\": %u, \"count\": %u, ", light, m);
        vec("hemicubes", h);
        vec("ids", ids);
        vec("energies", e);
        vec("item", items, true);
        printf("}\n");
        return 0;
    }
    fprintf(stderr, "usage: see the header of dropin_check.cpp\n");
    return 2;
}
