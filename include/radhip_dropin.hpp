// radhip_dropin.hpp - C++ drop-in layer over the C ABI of libradhip.
//
// The reference's host API for the shooting path, with the same class names,
// member names, argument meaning and ownership rules, implemented on the
// MI355X library (include/radhip.h).  A caller of the reference switches by
// including this header instead of Config.h, Patch.h, Model.h,
// PrimitiveModel.h, ModelContainer.h and FormFactors.h and using namespace
// radhip; the GL item-buffer render and the OpenCL ProcessHemicube launch
// of OnIdle become one Shooter::shootCycle() call.
//
// Differences by design: patches keep their host fields (radiosity,
// illumination) as the reference's do, and the GPU state is synchronised
// around the calls that need it; the GL display path is not part of this
// layer (its vertex colours are, through smoothShadeAll).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "radhip.h"

namespace radhip {

// Vector3f of Vector.h: the fields and operations the hot path uses.
struct Vector3f {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    Vector3f() = default;
    Vector3f(float a, float b, float c) : x(a), y(b), z(c) {}
    Vector3f operator+(const Vector3f &o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vector3f operator-(const Vector3f &o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vector3f operator*(const Vector3f &o) const { return {x * o.x, y * o.y, z * o.z}; }
    Vector3f operator*(float t) const { return {x * t, y * t, z * t}; }
    // Vector.h:448-452: a multiply by the reciprocal
    Vector3f operator/(float t) const
    {
        t = 1 / t;
        return {x * t, y * t, z * t};
    }
    Vector3f &operator+=(const Vector3f &o)
    {
        x += o.x;
        y += o.y;
        z += o.z;
        return *this;
    }
    Vector3f &operator-=(const Vector3f &o)
    {
        x -= o.x;
        y -= o.y;
        z -= o.z;
        return *this;
    }
    float f_Length2() const { return x * x + y * y + z * z; }
    float f_Length() const;
};

// Config (Config.h:7-39, Config.cpp): static settings, frozen before use.
class Config {
  public:
    static void setHemicubeSide(unsigned int n);
    static void setOCLWorkitemsX(unsigned int n);
    static void setMaxPatchArea(double n);
    static void setShootsPerCycle(unsigned int n);
    static void setHemicubesCount(unsigned int n);
    static void freeze();

    static unsigned int HEMICUBE_W();
    static unsigned int HEMICUBE_H();
    static unsigned int PATCHVIEW_TEX_W();
    static unsigned int PATCHVIEW_TEX_H();
    static unsigned int PATCHVIEW_TEX_RES();
    static double MAX_PATCH_AREA();
    static unsigned int OCL_WORKITEMS_X();
    static unsigned int OCL_WORKITEMS_Y();
    static unsigned int SHOOTS_PER_CYCLE();
    static unsigned int HEMICUBES_CNT();
    static bool isFrozen();
};

// REFLECTIVITY of Patch.h:13.
constexpr float REFLECTIVITY = 0.3f;

// Patch (Patch.h:16-56): a quad with colour, illumination (shot) and
// radiosity (unshot) fields.
class Patch {
  public:
    Patch();
    Patch(Vector3f v1, Vector3f v2, Vector3f v3, Vector3f v4);
    Patch(Vector3f v1, Vector3f v2, Vector3f v3, Vector3f v4, Vector3f color);
    Patch(Vector3f v1, Vector3f v2, Vector3f v3, Vector3f v4, Vector3f color,
          Vector3f illumination);
    Patch(Vector3f v1, Vector3f v2, Vector3f v3, Vector3f v4, Vector3f color,
          Vector3f illumination, Vector3f radiosity);

    // Patch::divide (Patch.cpp:47-125): children of area <= `area` (the
    // caller owns the vector and the patches), or NULL when already small.
    std::vector<Patch *> *divide(double area);
    std::vector<float> getVerticesCoords();

    Vector3f getCenter();
    Vector3f getNormal();
    Vector3f getUp();
    Vector3f getColor() { return color; }
    float getReflectivity() { return REFLECTIVITY; }
    Patch **getNeighbours() { return neighbours; }

    Vector3f radiosity;    // unshot
    Vector3f illumination; // shot

    // Patch.h:48-49: links to the 8 surrounding patches clockwise from the
    // top-left (a patch without a neighbour there links to itself; NULL
    // until Patch::divide, Model::subdivide or updateData sets them), and
    // their indices in the scene (filled by updateData).
    unsigned int relativeNeighbours[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    Patch *neighbours[8] = {nullptr, nullptr, nullptr, nullptr,
                            nullptr, nullptr, nullptr, nullptr};

  protected:
    Vector3f vec1, vec2, vec3, vec4;
    Vector3f color;
};

// Model (Model.h:7-20): a list of patches, subdivided on request.
class Model {
  public:
    Model();
    virtual ~Model();
    virtual std::vector<Patch *> *getPatches(double area = 0) = 0;

  protected:
    void subdivide(double area); // Model.cpp:27-60
    std::vector<Patch *> *patches;
};

// PrimitiveModel (PrimitiveModel.h): the Cornell box pieces.
class PrimitiveModel : public Model {
  public:
    enum { ROOM, ROOMCLOSURE, CUBE, BLOCK };
    explicit PrimitiveModel(int type);
    std::vector<Patch *> *getPatches(double area = 0) override;

  private:
    int type;
    bool built = false;
};

// WaveFrontModel (WaveFrontModel.h): the faces of an OBJ file
// (rad_obj_parse; coordinates / 1000, triangles as quads A B C C), black
// patches as the reference makes them (Patch(v1, v2, v3, v4)).  getPatches
// subdivides on every call with area > 0, as WaveFrontModel.cpp:145-151 does.
class WaveFrontModel : public Model {
  public:
    explicit WaveFrontModel(std::string filename);
    bool parse(std::string filename);
    std::vector<Patch *> *getPatches(double area = 0) override;
};

// Colors::smoothShadePatch (Colors.cpp:195-262): the 12 vertex colour
// components of a patch (corners 1..4 = left-bottom, right-bottom,
// right-top, left-top), each the mean of colour * (illumination +
// radiosity) over the patch and its three neighbours at that corner.
namespace Colors {
void smoothShadePatch(float *colors, Patch *p);
}

// ModelContainer (ModelContainer.h:11-53) whose shooter selection runs on
// the device.  updateData() also uploads the scene to a context made from the
// frozen Config (REFERENCE shooting and selection rules).
class ModelContainer {
  public:
    ModelContainer();
    ~ModelContainer();

    void load(); // ModelContainer.cpp:35-48: the Cornell box

    int addModel(Model *m);
    void removeModel(int i);
    void updateData();

    float *getVertices();
    unsigned int getVerticesCount();
    int *getIndices();
    unsigned int getIndicesCount();
    Patch **getPatches();
    unsigned int getPatchesCount();

    // ModelContainer.cpp:222-237: argmax |radiosity|^2, first index on ties
    // (the reference's unused single variant; a host loop over the fields).
    unsigned int getHighestRadiosityPatchId();
    // ModelContainer.cpp:259-299 on the device: `count` ids by the
    // reference's list rule, unfilled slots NULL / 0.
    void getHighestRadiosityPatchesId(unsigned int count, Patch **p_emitters,
                                      unsigned int *p_emitters_ids);

    double maxPatchArea;

    // SaveToFile / LoadFromFile of Main.cpp:1460-1647 with a path instead
    // of the file dialog: the patch fields in the library's portable
    // checkpoint format (rad_checkpoint_save / rad_checkpoint_load); loading
    // requires the same scene and writes the fields back.
    bool saveToFile(const char *path);
    bool loadFromFile(const char *path);

    // The vertex-colour loop of the display (Main.cpp:1319-1336) for every
    // patch at once on the device, from the device state: colors receives
    // getPatchesCount() * 12 floats in Colors::smoothShadePatch's layout.
    bool smoothShadeAll(float *colors);

    // library context of the uploaded scene (created by updateData)
    rad_ctx *context();
    void pushState(); // patch fields -> device
    void pullState(); // device -> patch fields

  protected:
    bool needRefresh;
    std::vector<Model *> models;
    std::vector<Patch *> patchList;
    std::vector<float> vertices;
    std::vector<int> indices;
    std::vector<float> emission0; // radiosity at upload: the residual's normalisation
    rad_ctx *ctx;
    unsigned int ctxK;
};

// float *precomputeHemicubeFormFactors() (FormFactors.h:20): new[] array of
// PATCHVIEW_TEX_RES * HEMICUBES_CNT floats (HEMICUBES_CNT copies of the
// cross-layout table), computed on the device; the caller delete[]s it.
// Config must be frozen.
float *precomputeHemicubeFormFactors();

// The OnIdle shooting cycle (Main.cpp:1137-1309): up to SHOOTS_PER_CYCLE
// passes of HEMICUBES_CNT shooters (select, render the item buffers,
// ProcessHemicube, accumulate, shoot, emitter update), each followed by the
// reference's stop test, then the patch fields are updated.
class Shooter {
  public:
    explicit Shooter(ModelContainer *scene);
    // One OnIdle call: passes until SHOOTS_PER_CYCLE or until the pass whose
    // lastEnergy.f_Length() < 0.1 (Main.cpp:1137, 1288-1300; computeRadiosity
    // then stays false and later calls shoot nothing).  Returns that length
    // for the cycle's last pass, or -1 when the library failed (failed() is
    // then true and lastError() says why).
    float shootCycle();
    // computeRadiosity == false: the stop test fired (Main.cpp:1298-1300).
    bool finished() const { return stopped; }
    bool failed() const { return error; }
    double residual(); // FP64 area-weighted unshot / emitted
    unsigned long long shooters() const { return shot; }
    unsigned long long passes() const { return npasses; }

  private:
    ModelContainer *scene;
    float lastEnergy = 0.0f;
    bool stopped = false, error = false;
    unsigned cycles = 0;
    unsigned long long shot = 0, npasses = 0;
};

// The ProcessHemicube kernel's host contract (Kernel_ProcessHemicube.h:9-69,
// Main.cpp:1207-1249) over an explicit item buffer of n_hemicubes stacked
// hemicubes (ids: patch + 1, 0 background): writes the kernel's triplets
// (hemicube, patch id, ΔFF sum of one run along a span of
// PATCHVIEW_TEX_W / OCL_WORKITEMS_X texels), 4 slots reserved at a time per
// span with (0, 0, 0) padding, and returns their count (the reference's write
// index n_last_index, a multiple of 4).  The output arrays hold
// PATCHVIEW_TEX_RES * n_hemicubes entries, as the reference allocates them
// (Main.cpp:605-607); a longer stream returns 0 with lastError() set.
unsigned int ProcessHemicube(ModelContainer *scene, const uint32_t *itembuf,
                             unsigned int n_hemicubes, uint32_t *p_hemicubes, uint32_t *p_ids,
                             float *p_energies);

// Last error of this layer (the library's message, or the layer's own).
const std::string &lastError();

} // namespace radhip
