// Host scene container: the data SceneLoader keeps for one scene
// (src/SceneLoader.h:100-160) in the byte layouts of include/rtx_amd.h.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "rtx_amd.h"
#include "rtx_host.h"

namespace rtxh {

struct Vec3 {
    float x = 0, y = 0, z = 0;
    Vec3() = default;
    Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
Vec3 normalize(Vec3 a);

// Column-major 4x4 (GLM convention), m[col*4+row].
struct Mat4 {
    float m[16];
    static Mat4 identity();
    // GLM's mat4 * vec4 summation order: (c0*x + c1*y) + (c2*z + c3*w).
    void apply(const float v[4], float out[4]) const;
};

// PCG32 (O'Neill 2014, pcg32_srandom_r / pcg32_random_r).
struct Pcg32 {
    uint64_t state = 0, inc = 1;
    Pcg32(uint64_t initstate, uint64_t initseq);
    uint32_t next();
    float uniform();                       // [0,1), 24 bits
    float uniform(float lo, float hi) { return lo + (hi - lo) * uniform(); }
};

struct Model {
    std::vector<rtx_vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<int> emissiveFaces; // faces whose material is a light (SceneLoader.cpp:320-323)
    float faceArea(int face, const Mat4 &transform) const; // Model::getFaceArea, src/Model.cpp:34-56
};

// An 8-bit RGBA texture as stb_image hands it to the reference
// (src/SceneLoader.cpp:184-236): RGB sRGB-encoded, alpha linear.
struct Texture {
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> rgba;
};

struct Scene {
    std::vector<Model> models;
    std::vector<rtx_instance> instances;
    std::vector<rtx_material> materials;
    std::vector<rtx_light> lights;
    std::vector<rtx_sphere> spheres;
    std::vector<int32_t> randomLightIndex;
    std::vector<rtx_face_sample> randomTriIndex;
    uint32_t randomTriRows = 0;
    std::vector<float> envMap;
    uint32_t envWidth = 0, envHeight = 0;
    // Texture slots as the reference numbers them: slot 0 is the environment
    // map's (src/SceneLoader.cpp:52-53), material textures from 1 on.
    std::vector<Texture> textures;
    std::vector<std::string> texturePaths;
    rtxh_view view{{0, -10, 4}, {0, 0, 4}, {0, 0, 1}, 28.0f}; // src/SceneLoader.h:147-150

    // scratch for rtx_scene_desc
    std::vector<rtx_model_desc> modelDescs;
    std::vector<rtx_texture_desc> textureDescs;

    // Append a model as its own instance (identity transform) the way the
    // Mitsuba path adds an "obj" shape (src/SceneLoader.cpp:455-495).
    // Returns the instance index; creates an area light if any face is emissive.
    int addModelInstance(Model &&model, const float *transform = nullptr, const float *normalTransform = nullptr);
    // A model without an instance (SceneLoader::addModel, src/SceneLoader.cpp:233-243);
    // returns the model index.
    int addModel(Model &&model);
    // An instance of a model; an area light when the model has emissive
    // faces (src/SceneLoader.cpp:717-728).  Returns the instance index.
    int addInstance(int model, const float *transform, const float *normalTransform);
    // Registers a decoded texture; returns its slot (1, 2, ...).
    int addTexture(Texture &&texture, const std::string &path);
    void addPointLight(Vec3 pos, Vec3 color);
    int addSphere(Vec3 center, float radius, int material);
    int addMaterial(const rtx_material &m);

    // SceneLoader::createLightSamplersBuffer (src/SceneLoader.cpp:847-944):
    // light choice table first, then the per-area-light face tables.
    void buildLightTables();
    void fillDesc(rtx_scene_desc *desc);
};

rtx_material makeDiffuse(Vec3 albedo);
rtx_material makeLight(Vec3 albedo, Vec3 radiance);
rtx_material makeDielectric(float ior);
rtx_material makeRoughConductor(float alpha, float eta, float k, int distribution = RTX_DISTRIBUTION_BECKMANN);
rtx_material makePhong(Vec3 kd, Vec3 ks, float exponent);

// Geometry helpers (flat-shaded faces; one vertex set per face).
void addQuad(Model &m, Vec3 p0, Vec3 p1, Vec3 p2, Vec3 p3, int mat);
void addBox(Model &m, Vec3 lo, Vec3 hi, int mat);
void addColumn(Model &m, Vec3 base, float radius, float height, int sides, int mat);

bool generateCornell(uint64_t seed, Scene &s, std::string &err);
bool generateRoom(uint64_t seed, uint32_t triangles, Scene &s, std::string &err);
bool generatePlates(uint64_t seed, Scene &s, std::string &err);
bool generateSpot(uint64_t seed, Scene &s, std::string &err);
bool loadMitsubaXml(const std::string &path, Scene &s, std::string &err);
bool loadJsonScene(const std::string &path, Scene &s, std::string &err);
// By extension: ".xml" Mitsuba, ".json" the reference's own format.
bool loadSceneFile(const std::string &path, Scene &s, std::string &err);

// 8-bit PNG / baseline JPEG decoding to RGBA (4 channels, alpha 255 when the
// file has none), rows top to bottom as stbi_load returns them.
bool loadImageRgba8(const std::string &path, Texture &out, std::string &err);

bool cameraMatrices(const rtxh_view &view, uint32_t w, uint32_t h, rtx_camera *cam);

bool writeExr(const std::string &path, const float *rgba, uint32_t w, uint32_t h, std::string &err);
bool readExr(const std::string &path, std::vector<float> &rgba, uint32_t &w, uint32_t &h,
             std::string &err);

} // namespace rtxh

struct rtxh_scene {
    rtxh::Scene scene;
};
