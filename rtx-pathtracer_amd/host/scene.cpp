// Host scene layer: light tables, face areas and geometry helpers.
// Restates src/SceneLoader.cpp:847-944 (light sampling tables),
// src/WeightedSampler.cpp:8-33 and src/Model.cpp:34-56 (face area).
#include "scene.h"

#include <cmath>
#include <cstring>
#include <random>

namespace rtxh {

Vec3 normalize(Vec3 a) {
    float l = std::sqrt(dot(a, a));
    return {a.x / l, a.y / l, a.z / l};
}

Mat4 Mat4::identity() {
    Mat4 r;
    std::memset(r.m, 0, sizeof(r.m));
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}

void Mat4::apply(const float v[4], float out[4]) const {
    for (int r = 0; r < 4; ++r) {
        float add0 = m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1];
        float add1 = m[2 * 4 + r] * v[2] + m[3 * 4 + r] * v[3];
        out[r] = add0 + add1;
    }
}

Pcg32::Pcg32(uint64_t initstate, uint64_t initseq) {
    state = 0u;
    inc = (initseq << 1u) | 1u;
    next();
    state += initstate;
    next();
}

uint32_t Pcg32::next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31u));
}

float Pcg32::uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }

float Model::faceArea(int face, const Mat4 &transform) const {
    float p[3][4];
    for (int i = 0; i < 3; ++i) {
        const rtx_vertex &v = vertices[indices[3 * face + i]];
        float in[4] = {v.pos[0], v.pos[1], v.pos[2], 1.0f};
        transform.apply(in, p[i]);
    }
    Vec3 v0(p[0][0], p[0][1], p[0][2]), v1(p[1][0], p[1][1], p[1][2]),
        v2(p[2][0], p[2][1], p[2][2]);
    Vec3 d01 = v1 - v0, d02 = v2 - v0;
    float l01 = std::sqrt(dot(d01, d01));
    float l02 = std::sqrt(dot(d02, d02));
    float cosAlpha = dot(d01, d02) / (l01 * l02);
    float sinAlpha = std::sqrt(1 - cosAlpha * cosAlpha);
    return l01 * l02 * sinAlpha / 2.0f;
}

static void set3(float dst[3], Vec3 v) {
    dst[0] = v.x;
    dst[1] = v.y;
    dst[2] = v.z;
}

static void identityInto(float m[16]) {
    Mat4 id = Mat4::identity();
    std::memcpy(m, id.m, sizeof(id.m));
}

int Scene::addModelInstance(Model &&model, const float *transform, const float *normalTransform) {
    return addInstance(addModel(std::move(model)), transform, normalTransform);
}

int Scene::addModel(Model &&model) {
    // emissive faces: faces whose material is a light (src/SceneLoader.cpp:320-323)
    model.emissiveFaces.clear();
    for (size_t f = 0; f * 3 < model.indices.size(); ++f) {
        int mat = model.vertices[model.indices[3 * f]].materialIndex;
        if (mat >= 0 && mat < (int)materials.size() && materials[mat].type == RTX_MAT_LIGHT)
            model.emissiveFaces.push_back((int)f);
    }
    models.push_back(std::move(model));
    return (int)models.size() - 1;
}

int Scene::addInstance(int model, const float *transform, const float *normalTransform) {
    rtx_instance inst;
    std::memset(&inst, 0, sizeof(inst));
    identityInto(inst.transform);
    identityInto(inst.normalTransform);
    if (transform) std::memcpy(inst.transform, transform, sizeof(inst.transform));
    if (normalTransform) std::memcpy(inst.normalTransform, normalTransform, sizeof(inst.normalTransform));
    inst.modelIndex = model;
    inst.iLight = -1;
    int instanceIndex = (int)instances.size();
    if (!models[model].emissiveFaces.empty()) {
        rtx_light light;
        std::memset(&light, 0, sizeof(light));
        light.type = RTX_LIGHT_AREA;
        light.instanceIndex = (uint32_t)instanceIndex;
        lights.push_back(light);
        inst.iLight = (int)lights.size() - 1;
    }
    instances.push_back(inst);
    return instanceIndex;
}

int Scene::addTexture(Texture &&texture, const std::string &path) {
    if (textures.empty()) { // the environment map's slot
        textures.emplace_back();
        texturePaths.emplace_back();
    }
    textures.push_back(std::move(texture));
    texturePaths.push_back(path);
    return (int)textures.size() - 1;
}

void Scene::addPointLight(Vec3 pos, Vec3 color) { // parsePointLights, src/SceneLoader.cpp:686-702
    rtx_light light;
    std::memset(&light, 0, sizeof(light));
    light.type = RTX_LIGHT_POINT;
    set3(light.color, color);
    set3(light.pos, pos);
    lights.push_back(light);
}

int Scene::addSphere(Vec3 center, float radius, int material) {
    rtx_sphere s;
    s.center[0] = center.x;
    s.center[1] = center.y;
    s.center[2] = center.z;
    s.radius = radius;
    s.materialIndex = material;
    s.iLight = -1;
    if (materials[material].type == RTX_MAT_LIGHT) { // src/SceneLoader.cpp:499-507
        rtx_light light;
        std::memset(&light, 0, sizeof(light));
        light.type = RTX_LIGHT_SPHERE;
        light.instanceIndex = (uint32_t)spheres.size();
        lights.push_back(light);
        s.iLight = (int)lights.size() - 1;
    }
    spheres.push_back(s);
    return (int)spheres.size() - 1;
}

int Scene::addMaterial(const rtx_material &m) {
    materials.push_back(m);
    return (int)materials.size() - 1;
}

namespace {
// WeightedSampler (src/WeightedSampler.h:9-31, src/WeightedSampler.cpp:8-33):
// a default-seeded mt19937 per sampler and a float uniform in [0,1).
struct WeightedSampler {
    std::vector<float> values;
    float total = 0.0f;
    std::mt19937 generator;
    std::uniform_real_distribution<float> distribution{0, 1};
    explicit WeightedSampler(const std::vector<float> &v) : values(v) {
        for (float x : values) total += x;
    }
    int sample() {
        if (values.empty()) return -1;
        float random = distribution(generator);
        random *= total;
        float sum = values[0];
        for (size_t i = 1; i < values.size(); ++i) {
            if (sum > random) return (int)i - 1;
            sum += values[i];
        }
        return (int)values.size() - 1;
    }
    std::vector<float> probabilities() const {
        std::vector<float> p(values.size());
        for (size_t i = 0; i < values.size(); ++i) p[i] = values[i] / total;
        return p;
    }
};
} // namespace

void Scene::buildLightTables() {
    // getLightSamplingVector (src/SceneLoader.cpp:923-944): equal power per light.
    std::vector<float> powers(lights.size(), 1.0f);
    WeightedSampler lightSampler(powers);
    std::vector<float> prob = lightSampler.probabilities();
    for (size_t i = 0; i < lights.size(); ++i) lights[i].sampleProb = prob[i];
    randomLightIndex.assign(RTX_SIZE_LIGHT_RANDOM, 0);
    for (int i = 0; i < RTX_SIZE_LIGHT_RANDOM; ++i) randomLightIndex[i] = lightSampler.sample();

    // getFaceSamplingVector (src/SceneLoader.cpp:861-921)
    randomTriIndex.clear();
    randomTriRows = 0;
    const float pi = 3.14159265358979323846264338327950288f;
    for (rtx_light &light : lights) {
        if (light.type == RTX_LIGHT_POINT || light.type == RTX_LIGHT_ENV_MAP) continue;
        if (light.type == RTX_LIGHT_SPHERE) {
            float radius = spheres[light.instanceIndex].radius;
            light.area = 4 * pi * radius * radius;
            continue;
        }
        const rtx_instance &inst = instances[light.instanceIndex];
        const Model &model = models[inst.modelIndex];
        Mat4 transform;
        std::memcpy(transform.m, inst.transform, sizeof(transform.m));
        std::vector<float> areas(model.emissiveFaces.size());
        for (size_t i = 0; i < areas.size(); ++i)
            areas[i] = model.faceArea(model.emissiveFaces[i], transform);
        WeightedSampler faceSampler(areas);
        light.area = faceSampler.total;
        std::vector<float> probSampleFace = faceSampler.probabilities();
        for (int i = 0; i < RTX_SIZE_TRI_RANDOM; ++i) {
            int s = faceSampler.sample();
            rtx_face_sample fs;
            fs.index = model.emissiveFaces[s];
            fs.sampleProb = probSampleFace[s];
            fs.faceArea = areas[s];
            randomTriIndex.push_back(fs);
        }
        randomTriRows++;
    }
    if (randomTriIndex.empty()) { // dummy row (src/SceneLoader.cpp:909-918)
        randomTriIndex.assign(RTX_SIZE_TRI_RANDOM, rtx_face_sample{0, 0.0f, 0.0f});
        randomTriRows = 1;
    }
}

void Scene::fillDesc(rtx_scene_desc *d) {
    modelDescs.resize(models.size());
    for (size_t i = 0; i < models.size(); ++i) {
        modelDescs[i].vertices = models[i].vertices.data();
        modelDescs[i].vertexCount = (uint32_t)models[i].vertices.size();
        modelDescs[i].indices = models[i].indices.data();
        modelDescs[i].indexCount = (uint32_t)models[i].indices.size();
    }
    std::memset(d, 0, sizeof(*d));
    d->models = modelDescs.data();
    d->modelCount = (uint32_t)modelDescs.size();
    d->instances = instances.data();
    d->instanceCount = (uint32_t)instances.size();
    d->materials = materials.data();
    d->materialCount = (uint32_t)materials.size();
    d->lights = lights.data();
    d->lightCount = (uint32_t)lights.size();
    d->randomLightIndex = randomLightIndex.data();
    d->randomTriIndex = randomTriIndex.data();
    d->randomTriRows = randomTriRows;
    d->spheres = spheres.data();
    d->sphereCount = (uint32_t)spheres.size();
    d->envMap = envMap.empty() ? nullptr : envMap.data();
    d->envWidth = envWidth;
    d->envHeight = envHeight;
    textureDescs.resize(textures.size());
    for (size_t i = 0; i < textures.size(); ++i) {
        textureDescs[i].rgba = textures[i].rgba.empty() ? nullptr : textures[i].rgba.data();
        textureDescs[i].width = textures[i].width;
        textureDescs[i].height = textures[i].height;
    }
    d->textures = textureDescs.empty() ? nullptr : textureDescs.data();
    d->textureCount = (uint32_t)textureDescs.size();
}

static rtx_material blankMaterial() {
    rtx_material m;
    std::memset(&m, 0, sizeof(m));
    m.textureIdDiffuse = -1;
    m.textureIdSpecular = -1;
    return m;
}


// parseXmlBSDF (src/SceneLoader.cpp:558-611) material mappings.
rtx_material makeDiffuse(Vec3 albedo) {
    rtx_material m = blankMaterial();
    m.type = RTX_MAT_DIFFUSE;
    set3(m.diffuse, albedo);
    return m;
}

rtx_material makeLight(Vec3 albedo, Vec3 radiance) {
    rtx_material m = makeDiffuse(albedo);
    set3(m.lightColor, radiance);
    m.type = RTX_MAT_LIGHT;
    return m;
}

rtx_material makeDielectric(float ior) {
    rtx_material m = blankMaterial();
    m.type = RTX_MAT_DIELECTRIC;
    set3(m.specular, Vec3(1, 1, 1));
    m.refractionIndex = ior;
    m.refractionIndexInv = 1.0f / ior;
    return m;
}

rtx_material makeRoughConductor(float alpha, float eta, float k, int distribution) {
    rtx_material m = blankMaterial();
    m.type = RTX_MAT_ROUGH_CONDUCTOR;
    m.distribution = distribution;
    m.roughness = alpha;
    m.eta = eta;
    m.k = k;
    return m;
}

rtx_material makePhong(Vec3 kd, Vec3 ks, float exponent) {
    rtx_material m = blankMaterial();
    m.type = RTX_MAT_PHONG;
    set3(m.diffuse, kd);
    set3(m.specular, ks);
    m.specularHighlight = exponent;
    return m;
}

static rtx_vertex makeVertex(Vec3 p, Vec3 n, int mat) {
    rtx_vertex v;
    std::memset(&v, 0, sizeof(v));
    set3(v.pos, p);
    set3(v.normal, n);
    v.materialIndex = mat;
    return v;
}

void addQuad(Model &m, Vec3 p0, Vec3 p1, Vec3 p2, Vec3 p3, int mat) {
    Vec3 n = normalize(cross(p1 - p0, p2 - p0));
    uint32_t base = (uint32_t)m.vertices.size();
    for (Vec3 p : {p0, p1, p2, p3}) m.vertices.push_back(makeVertex(p, n, mat));
    for (uint32_t i : {0u, 1u, 2u, 0u, 2u, 3u}) m.indices.push_back(base + i);
}

void addBox(Model &m, Vec3 lo, Vec3 hi, int mat) {
    // outward-facing quads
    addQuad(m, {lo.x, lo.y, lo.z}, {lo.x, lo.y, hi.z}, {lo.x, hi.y, hi.z}, {lo.x, hi.y, lo.z}, mat);
    addQuad(m, {hi.x, lo.y, lo.z}, {hi.x, hi.y, lo.z}, {hi.x, hi.y, hi.z}, {hi.x, lo.y, hi.z}, mat);
    addQuad(m, {lo.x, lo.y, lo.z}, {hi.x, lo.y, lo.z}, {hi.x, lo.y, hi.z}, {lo.x, lo.y, hi.z}, mat);
    addQuad(m, {lo.x, hi.y, lo.z}, {lo.x, hi.y, hi.z}, {hi.x, hi.y, hi.z}, {hi.x, hi.y, lo.z}, mat);
    addQuad(m, {lo.x, lo.y, lo.z}, {lo.x, hi.y, lo.z}, {hi.x, hi.y, lo.z}, {hi.x, lo.y, lo.z}, mat);
    addQuad(m, {lo.x, lo.y, hi.z}, {hi.x, lo.y, hi.z}, {hi.x, hi.y, hi.z}, {lo.x, hi.y, hi.z}, mat);
}

// Vertical (+y) column with `sides` smooth-shaded side quads and two
// flat caps, fanned: 2*sides + 2*(sides-2) triangles.
void addColumn(Model &m, Vec3 base, float radius, float height, int sides, int mat) {
    const double twoPi = 6.283185307179586;
    std::vector<Vec3> ring(sides), nrm(sides);
    for (int i = 0; i < sides; ++i) {
        double a = twoPi * i / sides;
        nrm[i] = Vec3((float)std::cos(a), 0.0f, (float)std::sin(a));
        ring[i] = Vec3(base.x + radius * nrm[i].x, base.y, base.z + radius * nrm[i].z);
    }
    uint32_t b = (uint32_t)m.vertices.size();
    for (int i = 0; i < sides; ++i) {
        m.vertices.push_back(makeVertex(ring[i], nrm[i], mat));
        m.vertices.push_back(makeVertex(ring[i] + Vec3(0, height, 0), nrm[i], mat));
    }
    for (int i = 0; i < sides; ++i) {
        uint32_t j = (uint32_t)((i + 1) % sides);
        uint32_t b0 = b + 2 * i, t0 = b0 + 1, b1 = b + 2 * j, t1 = b1 + 1;
        for (uint32_t x : {b0, t0, t1, b0, t1, b1}) m.indices.push_back(x);
    }
    uint32_t top = (uint32_t)m.vertices.size();
    for (int i = 0; i < sides; ++i)
        m.vertices.push_back(makeVertex(ring[i] + Vec3(0, height, 0), Vec3(0, 1, 0), mat));
    for (int i = 1; i + 1 < sides; ++i)
        for (uint32_t x : {top, top + i + 1, top + i}) m.indices.push_back(x);
    uint32_t bot = (uint32_t)m.vertices.size();
    for (int i = 0; i < sides; ++i) m.vertices.push_back(makeVertex(ring[i], Vec3(0, -1, 0), mat));
    for (int i = 1; i + 1 < sides; ++i)
        for (uint32_t x : {bot, bot + i, bot + i + 1}) m.indices.push_back(x);
}

} // namespace rtxh
