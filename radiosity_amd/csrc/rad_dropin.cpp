// rad_dropin.cpp - the reference's host classes over the C ABI (host code).
//
// Config, Patch, Model, PrimitiveModel, ModelContainer and
// precomputeHemicubeFormFactors keep the reference's names and argument
// meaning; the hot loop of OnIdle (Main.cpp:1137-1309) is Shooter, which
// runs whole passes on the device and writes the patch fields back.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/radhip_dropin.hpp"
#include "../../include/radhip_scene.h"
#include "rad_patch.h"

namespace radhip {

namespace {

std::string g_err;

int setErr(const std::string &m)
{
    g_err = m;
    return -1;
}

int libErr(rad_ctx *c, int rc, const char *what)
{
    if (rc != RAD_OK)
        g_err = std::string(what) + ": " + (c ? rad_last_error(c) : "no context") + " (" +
                std::to_string(rc) + ")";
    return rc;
}

radpatch::V3 toV3(const Vector3f &v) { return {v.x, v.y, v.z}; }
Vector3f fromV3(const radpatch::V3 &v) { return {v.x, v.y, v.z}; }

// Config storage: Config.cpp:5-10 defaults
struct ConfigState {
    bool frozen = false;
    unsigned hemicubeSide = 16, oclWorkitemsX = 4, shootsPerCycle = 500, hemicubesCount = 10;
    double maxPatchArea = 0.5;
    unsigned W = 0, H = 0, TW = 0, TH = 0, RES = 0, WX = 0, WY = 0;
    double area = 0;
};
ConfigState cfg;

bool cfgWritable()
{
    if (cfg.frozen) {
        setErr("Config: trying to modify a frozen configuration");
        return false;
    }
    return true;
}

} // namespace

const std::string &lastError() { return g_err; }

float Vector3f::f_Length() const { return sqrtf(x * x + y * y + z * z); }

// ------------------------------------------------------------------ Config

void Config::setHemicubeSide(unsigned int n)
{
    if (cfgWritable())
        cfg.hemicubeSide = n;
}
void Config::setOCLWorkitemsX(unsigned int n)
{
    if (cfgWritable())
        cfg.oclWorkitemsX = n;
}
void Config::setMaxPatchArea(double n)
{
    if (cfgWritable())
        cfg.maxPatchArea = n;
}
void Config::setShootsPerCycle(unsigned int n)
{
    if (cfgWritable())
        cfg.shootsPerCycle = n;
}
void Config::setHemicubesCount(unsigned int n)
{
    if (cfgWritable())
        cfg.hemicubesCount = n;
}

// Config.cpp:30-48
void Config::freeze()
{
    cfg.frozen = true;
    cfg.W = cfg.H = cfg.hemicubeSide;
    cfg.TW = cfg.W * 2;
    cfg.TH = (unsigned)((double)cfg.H * 1.5);
    cfg.RES = cfg.TW * cfg.TH;
    cfg.area = cfg.maxPatchArea;
    cfg.WX = std::min(cfg.oclWorkitemsX, cfg.TW);
    cfg.WY = cfg.TH * cfg.hemicubesCount;
}

unsigned int Config::HEMICUBE_W() { return cfg.W; }
unsigned int Config::HEMICUBE_H() { return cfg.H; }
unsigned int Config::PATCHVIEW_TEX_W() { return cfg.TW; }
unsigned int Config::PATCHVIEW_TEX_H() { return cfg.TH; }
unsigned int Config::PATCHVIEW_TEX_RES() { return cfg.RES; }
double Config::MAX_PATCH_AREA() { return cfg.area; }
unsigned int Config::OCL_WORKITEMS_X() { return cfg.WX; }
unsigned int Config::OCL_WORKITEMS_Y() { return cfg.WY; }
unsigned int Config::SHOOTS_PER_CYCLE() { return cfg.shootsPerCycle; }
unsigned int Config::HEMICUBES_CNT() { return cfg.hemicubesCount; }
bool Config::isFrozen() { return cfg.frozen; }

// ------------------------------------------------------------------- Patch

Patch::Patch() {}
// Patch.cpp:8-13: a patch made from corners alone is black
Patch::Patch(Vector3f v1, Vector3f v2, Vector3f v3, Vector3f v4)
    : vec1(v1), vec2(v2), vec3(v3), vec4(v4), color(0, 0, 0)
{
}
Patch::Patch(Vector3f v1, Vector3f v2, Vector3f v3, Vector3f v4, Vector3f c)
    : vec1(v1), vec2(v2), vec3(v3), vec4(v4), color(c)
{
}
Patch::Patch(Vector3f v1, Vector3f v2, Vector3f v3, Vector3f v4, Vector3f c, Vector3f ill)
    : illumination(ill), vec1(v1), vec2(v2), vec3(v3), vec4(v4), color(c)
{
}
Patch::Patch(Vector3f v1, Vector3f v2, Vector3f v3, Vector3f v4, Vector3f c, Vector3f ill,
             Vector3f rad)
    : radiosity(rad), illumination(ill), vec1(v1), vec2(v2), vec3(v3), vec4(v4), color(c)
{
}

std::vector<Patch *> *Patch::divide(double area)
{
    radpatch::Quad q{};
    q.v[0] = toV3(vec1);
    q.v[1] = toV3(vec2);
    q.v[2] = toV3(vec3);
    q.v[3] = toV3(vec4);
    unsigned kx, ky;
    if (!radpatch::divide_grid(q, area, kx, ky))
        return nullptr;
    auto *out = new std::vector<Patch *>();
    out->reserve((size_t)kx * ky);
    // Patch.cpp:119-122: children inherit colour, illumination and radiosity
    for (unsigned i = 0; i < kx * ky; ++i) {
        radpatch::V3 v[4];
        radpatch::divide_child(q.v, kx, ky, i, v);
        out->push_back(new Patch(fromV3(v[0]), fromV3(v[1]), fromV3(v[2]), fromV3(v[3]), color,
                                 illumination, radiosity));
    }
    // Patch.cpp:127-222
    for (unsigned i = 0; i < kx * ky; ++i) {
        unsigned nb[8];
        radpatch::divide_neighbours(kx, ky, i, nb);
        for (int k = 0; k < 8; ++k)
            (*out)[i]->neighbours[k] = (*out)[nb[k]];
    }
    return out;
}

std::vector<float> Patch::getVerticesCoords()
{
    return {vec1.x, vec1.y, vec1.z, vec2.x, vec2.y, vec2.z,
            vec3.x, vec3.y, vec3.z, vec4.x, vec4.y, vec4.z};
}

static void frameOf(const Vector3f &a, const Vector3f &b, const Vector3f &c, const Vector3f &d,
                    float fr[9])
{
    radpatch::V3 v[4] = {toV3(a), toV3(b), toV3(c), toV3(d)};
    radpatch::quad_frame(v, fr);
}

Vector3f Patch::getCenter()
{
    float fr[9];
    frameOf(vec1, vec2, vec3, vec4, fr);
    return {fr[0], fr[1], fr[2]};
}
Vector3f Patch::getNormal()
{
    float fr[9];
    frameOf(vec1, vec2, vec3, vec4, fr);
    return {fr[3], fr[4], fr[5]};
}
Vector3f Patch::getUp()
{
    float fr[9];
    frameOf(vec1, vec2, vec3, vec4, fr);
    return {fr[6], fr[7], fr[8]};
}

// ------------------------------------------------------------ Model, Primitive

Model::Model() : patches(new std::vector<Patch *>()) {}

Model::~Model()
{
    if (patches) {
        for (Patch *p : *patches)
            delete p;
        delete patches;
    }
}

// Model.cpp:27-60: one pass, each patch replaced by its children in place.
void Model::subdivide(double area)
{
    if (!(area > 0))
        return;
    auto *out = new std::vector<Patch *>();
    for (Patch *p : *patches) {
        std::vector<Patch *> *kids = p->divide(area);
        if (kids) {
            out->insert(out->end(), kids->begin(), kids->end());
            delete kids;
            delete p;
        } else {
            // Model.cpp:42-48: an undivided patch links to itself
            for (Patch *&n : p->neighbours)
                if (!n)
                    n = p;
            out->push_back(p);
        }
    }
    delete patches;
    patches = out;
}

PrimitiveModel::PrimitiveModel(int t) : type(t) {}

std::vector<Patch *> *PrimitiveModel::getPatches(double area)
{
    if (!built) {
        built = true;
        std::vector<std::vector<radpatch::Quad>> src = radpatch::cornell_models();
        if (type >= 0 && type < (int)src.size())
            for (const radpatch::Quad &q : src[type])
                patches->push_back(new Patch(fromV3(q.v[0]), fromV3(q.v[1]), fromV3(q.v[2]),
                                             fromV3(q.v[3]), fromV3(q.color), fromV3(q.illum),
                                             fromV3(q.rad)));
        subdivide(area);
    }
    return patches;
}

WaveFrontModel::WaveFrontModel(std::string filename)
{
    if (!parse(filename))
        setErr("Unable to load model '" + filename + "'");
}

// WaveFrontModel.cpp:15-136 through rad_obj_parse; the faces read before a
// bad vertex line are kept, as the reference keeps them.
bool WaveFrontModel::parse(std::string filename)
{
    float *q = nullptr;
    uint32_t n = 0;
    int rc = rad_obj_parse(filename.c_str(), &q, &n);
    for (uint32_t f = 0; f < n; ++f) {
        const float *v = q + 12 * (size_t)f;
        patches->push_back(new Patch({v[0], v[1], v[2]}, {v[3], v[4], v[5]}, {v[6], v[7], v[8]},
                                     {v[9], v[10], v[11]}));
    }
    rad_obj_free(q);
    return rc == RAD_OK;
}

// WaveFrontModel.cpp:145-151
std::vector<Patch *> *WaveFrontModel::getPatches(double area)
{
    if (area > 0)
        subdivide(area);
    return patches;
}

// Colors.cpp:195-262
void Colors::smoothShadePatch(float *colors, Patch *p)
{
    auto val = [](Patch *q) { return q->getColor() * (q->illumination + q->radiosity); };
    Patch **n = p->neighbours;
    Vector3f lt = val(p), rt = val(p), rb = val(p), lb = val(p);
    lt += val(n[7]);
    lt += val(n[0]);
    lt += val(n[1]);
    lt = lt / 4;
    rt += val(n[1]);
    rt += val(n[2]);
    rt += val(n[3]);
    rt = rt / 4;
    rb += val(n[3]);
    rb += val(n[4]);
    rb += val(n[5]);
    rb = rb / 4;
    lb += val(n[5]);
    lb += val(n[6]);
    lb += val(n[7]);
    lb = lb / 4;
    const Vector3f out[4] = {lb, rb, rt, lt};
    for (int k = 0; k < 4; ++k) {
        colors[3 * k] = out[k].x;
        colors[3 * k + 1] = out[k].y;
        colors[3 * k + 2] = out[k].z;
    }
}

// ---------------------------------------------------------- ModelContainer

ModelContainer::ModelContainer()
    : maxPatchArea(Config::isFrozen() ? Config::MAX_PATCH_AREA() : 0.5), needRefresh(true),
      ctx(nullptr), ctxK(0)
{
}

ModelContainer::~ModelContainer()
{
    if (ctx)
        rad_destroy(ctx);
    for (Model *m : models)
        delete m;
}

// ModelContainer.cpp:35-48
void ModelContainer::load()
{
    addModel(new PrimitiveModel(PrimitiveModel::ROOM));
    addModel(new PrimitiveModel(PrimitiveModel::ROOMCLOSURE));
    addModel(new PrimitiveModel(PrimitiveModel::CUBE));
    addModel(new PrimitiveModel(PrimitiveModel::BLOCK));
}

int ModelContainer::addModel(Model *m)
{
    models.push_back(m);
    needRefresh = true;
    return (int)models.size() - 1;
}

void ModelContainer::removeModel(int i)
{
    if (i < 0 || i >= (int)models.size())
        return;
    delete models[i];
    models.erase(models.begin() + i);
    needRefresh = true;
}

// ModelContainer.cpp:81-155 (vertices 4 x 3 per patch, indices 0,1,2,0,2,3),
// then the scene goes to the device: two triangles per patch sharing its id.
void ModelContainer::updateData()
{
    patchList.clear();
    vertices.clear();
    indices.clear();
    int offset = 0;
    for (Model *m : models) {
        std::vector<Patch *> *ps = m->getPatches(maxPatchArea);
        for (Patch *p : *ps) {
            std::vector<float> v = p->getVerticesCoords();
            vertices.insert(vertices.end(), v.begin(), v.end());
            for (int k : {0, 1, 2, 0, 2, 3})
                indices.push_back(offset + k);
            // ModelContainer.cpp:112-116
            for (Patch *&n : p->neighbours)
                if (!n)
                    n = p;
            patchList.push_back(p);
            offset += 4;
        }
    }
    needRefresh = false;
    if (!Config::isFrozen()) {
        setErr("ModelContainer::updateData: Config::freeze() first");
        return;
    }
    uint32_t n = (uint32_t)patchList.size();
    if (!ctx || ctxK != Config::HEMICUBES_CNT()) {
        if (ctx)
            rad_destroy(ctx);
        ctx = nullptr;
        rad_config rc{};
        rc.hemicube = Config::HEMICUBE_W();
        rc.hemicubes = Config::HEMICUBES_CNT();
        rc.shoots = Config::SHOOTS_PER_CYCLE();
        rc.ocl_workitems_x = Config::OCL_WORKITEMS_X();
        rc.mode = RAD_MODE_REFERENCE;
        rc.select = RAD_SELECT_REFERENCE;
        if (libErr(nullptr, rad_create(&rc, &ctx), "rad_create") != RAD_OK) {
            ctx = nullptr;
            return;
        }
        ctxK = rc.hemicubes;
    }
    if (n == 0)
        return;
    std::vector<float> tri((size_t)n * 18), frame((size_t)n * 9), refl((size_t)n * 3), area(n);
    std::vector<uint32_t> tp((size_t)n * 2);
    for (uint32_t i = 0; i < n; ++i) {
        const float *v = &vertices[(size_t)i * 12];
        radpatch::V3 q[4];
        for (int k = 0; k < 4; ++k)
            q[k] = {v[3 * k], v[3 * k + 1], v[3 * k + 2]};
        const int order[6] = {0, 1, 2, 0, 2, 3};
        for (int k = 0; k < 6; ++k) {
            tri[(size_t)i * 18 + 3 * k + 0] = q[order[k]].x;
            tri[(size_t)i * 18 + 3 * k + 1] = q[order[k]].y;
            tri[(size_t)i * 18 + 3 * k + 2] = q[order[k]].z;
        }
        tp[2 * i] = tp[2 * i + 1] = i;
        radpatch::quad_frame(q, &frame[(size_t)i * 9]);
        Vector3f c = patchList[i]->getColor();
        refl[3 * i] = c.x;
        refl[3 * i + 1] = c.y;
        refl[3 * i + 2] = c.z;
        radpatch::V3 n0 = radpatch::cross(q[1] - q[0], q[2] - q[0]);
        radpatch::V3 n1 = radpatch::cross(q[2] - q[0], q[3] - q[0]);
        area[i] = 0.5f * (radpatch::length(n0) + radpatch::length(n1));
        if (!(area[i] > 0))
            area[i] = 1e-12f;
    }
    if (libErr(ctx,
               rad_scene_upload(ctx, tri.data(), tp.data(), 2 * n, n, frame.data(), refl.data(),
                                area.data()),
               "rad_scene_upload") != RAD_OK)
        return;
    // the residual is normalised by the radiosity the scene starts with
    std::vector<float> u((size_t)n * 3), s((size_t)n * 3);
    for (uint32_t i = 0; i < n; ++i) {
        Patch *p = patchList[i];
        u[3 * i] = p->radiosity.x;
        u[3 * i + 1] = p->radiosity.y;
        u[3 * i + 2] = p->radiosity.z;
        s[3 * i] = p->illumination.x;
        s[3 * i + 1] = p->illumination.y;
        s[3 * i + 2] = p->illumination.z;
    }
    emission0 = u;
    if (libErr(ctx, rad_set_state(ctx, u.data(), s.data(), u.data()), "rad_set_state") != RAD_OK)
        return;
    // neighbour links as scene indices (a link to a patch outside the scene
    // counts as a link to itself); every patch is its own quad
    std::unordered_map<const Patch *, uint32_t> index;
    index.reserve(n);
    for (uint32_t i = 0; i < n; ++i)
        index[patchList[i]] = i;
    std::vector<uint32_t> pq(n), nbr((size_t)n * 8);
    for (uint32_t i = 0; i < n; ++i) {
        pq[i] = i;
        for (int k = 0; k < 8; ++k) {
            auto it = index.find(patchList[i]->neighbours[k]);
            uint32_t j = it == index.end() ? i : it->second;
            patchList[i]->relativeNeighbours[k] = j;
            nbr[(size_t)i * 8 + k] = j;
        }
    }
    libErr(ctx, rad_set_quads(ctx, pq.data(), n, nbr.data()), "rad_set_quads");
}

bool ModelContainer::smoothShadeAll(float *colors)
{
    rad_ctx *c = context();
    if (!c || patchList.empty() || !colors)
        return false;
    return libErr(c, rad_vertex_colors(c, colors), "rad_vertex_colors") == RAD_OK;
}

float *ModelContainer::getVertices()
{
    if (needRefresh)
        updateData();
    return vertices.data();
}
unsigned int ModelContainer::getVerticesCount()
{
    if (needRefresh)
        updateData();
    return (unsigned)vertices.size();
}
int *ModelContainer::getIndices()
{
    if (needRefresh)
        updateData();
    return indices.data();
}
unsigned int ModelContainer::getIndicesCount()
{
    if (needRefresh)
        updateData();
    return (unsigned)indices.size();
}
Patch **ModelContainer::getPatches()
{
    if (needRefresh)
        updateData();
    return patchList.data();
}
unsigned int ModelContainer::getPatchesCount()
{
    if (needRefresh)
        updateData();
    return (unsigned)patchList.size();
}

rad_ctx *ModelContainer::context()
{
    if (needRefresh)
        updateData();
    return ctx;
}

void ModelContainer::pushState()
{
    rad_ctx *c = context();
    if (!c || patchList.empty())
        return;
    size_t n = patchList.size();
    std::vector<float> u(n * 3), s(n * 3);
    for (size_t i = 0; i < n; ++i) {
        const Patch *p = patchList[i];
        u[3 * i] = p->radiosity.x;
        u[3 * i + 1] = p->radiosity.y;
        u[3 * i + 2] = p->radiosity.z;
        s[3 * i] = p->illumination.x;
        s[3 * i + 1] = p->illumination.y;
        s[3 * i + 2] = p->illumination.z;
    }
    // the emission (residual normalisation) set at upload is kept
    libErr(c, rad_set_state(c, u.data(), s.data(), emission0.data()), "rad_set_state");
}

void ModelContainer::pullState()
{
    rad_ctx *c = context();
    if (!c || patchList.empty())
        return;
    size_t n = patchList.size();
    std::vector<float> u(n * 3), s(n * 3);
    if (libErr(c, rad_get_state(c, u.data(), s.data()), "rad_get_state") != RAD_OK)
        return;
    for (size_t i = 0; i < n; ++i) {
        patchList[i]->radiosity = {u[3 * i], u[3 * i + 1], u[3 * i + 2]};
        patchList[i]->illumination = {s[3 * i], s[3 * i + 1], s[3 * i + 2]};
    }
}

bool ModelContainer::saveToFile(const char *path)
{
    rad_ctx *c = context();
    if (!c)
        return false;
    pushState();
    return libErr(c, rad_checkpoint_save(c, path), "rad_checkpoint_save") == RAD_OK;
}

bool ModelContainer::loadFromFile(const char *path)
{
    rad_ctx *c = context();
    if (!c)
        return false;
    if (libErr(c, rad_checkpoint_load(c, path), "rad_checkpoint_load") != RAD_OK)
        return false;
    pullState();
    return true;
}

unsigned int ModelContainer::getHighestRadiosityPatchId()
{
    if (needRefresh)
        updateData();
    unsigned best = 0;
    float bv = -1.0f;
    for (unsigned i = 0; i < patchList.size(); ++i) {
        float v = patchList[i]->radiosity.f_Length2();
        if (v > bv) {
            bv = v;
            best = i;
        }
    }
    return best;
}

void ModelContainer::getHighestRadiosityPatchesId(unsigned int count, Patch **p_emitters,
                                                  unsigned int *p_emitters_ids)
{
    for (unsigned i = 0; i < count; ++i) {
        p_emitters[i] = nullptr;
        p_emitters_ids[i] = 0;
    }
    rad_ctx *c = context();
    if (!c || count == 0)
        return;
    if (count > ctxK) {
        setErr("getHighestRadiosityPatchesId: count beyond Config::HEMICUBES_CNT()");
        return;
    }
    pushState();
    std::vector<uint32_t> ids(count);
    uint32_t filled = 0;
    if (libErr(c, rad_select(c, count, ids.data(), &filled), "rad_select") != RAD_OK)
        return;
    for (unsigned i = 0; i < filled && i < count; ++i) {
        p_emitters_ids[i] = ids[i];
        p_emitters[i] = patchList[ids[i]];
    }
}

// ------------------------------------------------------ form factors, shooting

float *precomputeHemicubeFormFactors()
{
    if (!Config::isFrozen()) {
        setErr("precomputeHemicubeFormFactors: Config::freeze() first");
        return nullptr;
    }
    rad_config rc{};
    rc.hemicube = Config::HEMICUBE_W();
    rc.hemicubes = 1;
    rad_ctx *c = nullptr;
    if (libErr(nullptr, rad_create(&rc, &c), "rad_create") != RAD_OK)
        return nullptr;
    size_t res = Config::PATCHVIEW_TEX_RES(), k = Config::HEMICUBES_CNT();
    float *out = new float[res * k];
    if (libErr(c, rad_precompute_dff(c, out), "rad_precompute_dff") != RAD_OK) {
        delete[] out;
        rad_destroy(c);
        return nullptr;
    }
    rad_destroy(c);
    // FormFactors.cpp:316-323: one copy per hemicube
    for (size_t h = 1; h < k; ++h)
        memcpy(out + h * res, out, res * sizeof(float));
    return out;
}

Shooter::Shooter(ModelContainer *s) : scene(s) {}

float Shooter::shootCycle()
{
    rad_ctx *c = scene ? scene->context() : nullptr;
    if (!c) {
        error = true;
        setErr("Shooter::shootCycle: the scene has no device context");
        return -1.0f;
    }
    if (stopped)
        return lastEnergy;
    scene->pushState();
    rad_cycle_stats st{};
    // the stop test of Main.cpp:1298 runs on the device after every pass
    if (libErr(c, rad_run_cycle(c, Config::SHOOTS_PER_CYCLE(), 0.1, &st), "rad_run_cycle") !=
        RAD_OK) {
        error = true;
        return -1.0f;
    }
    scene->pullState();
    lastEnergy = Vector3f(st.last_energy[0], st.last_energy[1], st.last_energy[2]).f_Length();
    stopped = st.stopped != 0;
    ++cycles;
    shot += st.shooters;
    npasses += st.passes;
    return lastEnergy;
}

double Shooter::residual()
{
    rad_ctx *c = scene ? scene->context() : nullptr;
    double r = -1.0;
    if (c)
        libErr(c, rad_residual_fp64(c, &r), "rad_residual_fp64");
    return r;
}

unsigned int ProcessHemicube(ModelContainer *scene, const uint32_t *itembuf,
                             unsigned int n_hemicubes, uint32_t *p_hemicubes, uint32_t *p_ids,
                             float *p_energies)
{
    rad_ctx *c = scene ? scene->context() : nullptr;
    if (!c || !itembuf || !p_hemicubes || !p_ids || !p_energies) {
        setErr("ProcessHemicube: no device context or null arrays");
        return 0;
    }
    // the item buffer goes to the device (the reduction into F rows there is
    // not read), then the kernel's triplet stream comes back
    if (libErr(c, rad_process_itembuf(c, itembuf, n_hemicubes, nullptr, nullptr),
               "rad_process_itembuf") != RAD_OK)
        return 0;
    uint32_t n = 0;
    const uint32_t cap = Config::PATCHVIEW_TEX_RES() * n_hemicubes; // Main.cpp:605-607
    if (libErr(c, rad_process_runs(c, n_hemicubes, cap, p_hemicubes, p_ids, p_energies, &n),
               "rad_process_runs") != RAD_OK)
        return 0;
    return n;
}

} // namespace radhip
