// Scene files: the Mitsuba 0.6 XML subset and the JSON format the reference
// loads (SceneLoader, src/SceneLoader.cpp:29-63, 329-348, 405-770), with
// the file layout it assumes: an XML scene finds its OBJ, MTL, texture and
// environment-map files beside it; a JSON scene in models/, materials/ and
// textures/ below its folder (src/SceneLoader.cpp:37-47).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

#include "json.h"
#include "obj.h"
#include "scene.h"
#include "xml.h"

namespace rtxh {

namespace {

[[noreturn]] void fail(const std::string &m) { throw std::runtime_error(m); }

std::string readText(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail("cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

std::string parentDir(const std::string &path) {
    size_t s = path.find_last_of('/');
    return s == std::string::npos ? std::string() : path.substr(0, s);
}

std::string joinPath(const std::string &dir, const std::string &name) {
    if (dir.empty() || (!name.empty() && name[0] == '/')) return name;
    return dir.back() == '/' ? dir + name : dir + "/" + name;
}

// Lexical normalisation of a relative or absolute path: "." and "x/.."
// segments removed (std::filesystem::path::lexically_normal).
std::string lexicallyNormal(const std::string &p) {
    const bool abs = !p.empty() && p[0] == '/';
    std::vector<std::string> parts;
    std::stringstream ss(p);
    std::string seg;
    while (std::getline(ss, seg, '/')) {
        if (seg.empty() || seg == ".") continue;
        if (seg == ".." && !parts.empty() && parts.back() != "..") parts.pop_back();
        else if (!(seg == ".." && abs)) parts.push_back(seg);
    }
    std::string out = abs ? "/" : "";
    for (size_t i = 0; i < parts.size(); ++i) out += (i ? "/" : "") + parts[i];
    return out.empty() ? "." : out;
}

struct Loader {
    Scene &s;
    std::string modelsDir, materialsDir, texturesDir;
    std::map<std::string, int> textureByPath;

    explicit Loader(Scene &scene) : s(scene) {}

    // SceneLoader::addTexture (src/SceneLoader.cpp:184-236): one slot per
    // file (keyed by the joined path before its first backslash becomes a
    // slash), decoded to 8-bit RGBA; a file that cannot be read is an error.
    int addTexture(const std::string &name) {
        std::string key = joinPath(texturesDir, name);
        auto it = textureByPath.find(key);
        if (it != textureByPath.end()) return it->second;
        std::string path = key;
        size_t bs = path.find('\\');
        if (bs != std::string::npos) path[bs] = '/';
        path = lexicallyNormal(path);
        Texture tex;
        std::string err;
        if (!loadImageRgba8(path, tex, err)) fail("could not load texture file " + path + ": " + err);
        int id = s.addTexture(std::move(tex), path);
        textureByPath[key] = id;
        return id;
    }

    ObjData readObjFile(const std::string &name) {
        ObjData obj;
        std::string err;
        if (!readObj(joinPath(modelsDir, name), materialsDir, obj, err)) fail(err);
        return obj;
    }
};

// ------------------------------------------------------------------ XML

// parseCommaSpaceSeparatedVec3 (src/MitsubaXML.h:29-49): numbers read one
// after another, each optionally followed by one ',' or ' '; reading stops
// at the first text that is not a number.  One number fills all three.
Vec3 parseVec3(const char *text) {
    if (!text) fail("missing vector value");
    std::vector<float> vals;
    const char *p = text;
    for (;;) {
        while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\f' || *p == '\v') ++p;
        char *end = nullptr;
        float f = std::strtof(p, &end);
        if (end == p) break;
        vals.push_back(f);
        p = end;
        if (*p == ',' || *p == ' ') ++p;
    }
    if (vals.size() == 1) return {vals[0], vals[0], vals[0]};
    if (vals.size() != 3) fail(std::string("'") + text + "' does not hold 3 values");
    return {vals[0], vals[1], vals[2]};
}

// tinyxml2's FloatAttribute: 0 when absent or not a number.
float floatAttr(const XmlNode *n, const char *key) {
    const char *v = n->attr(key);
    if (!v) return 0.0f;
    char *end = nullptr;
    float f = std::strtof(v, &end);
    return end == v ? 0.0f : f;
}

std::string requiredAttr(const XmlNode *n, const char *key) {
    const char *v = n->attr(key);
    if (!v) fail("<" + n->name + "> has no '" + key + "' attribute");
    return v;
}

// getNamedChild (src/MitsubaXML.h:15-27): the first child element of the
// given tag whose name attribute matches.
const XmlNode *namedChild(const XmlNode *n, const std::string &name, const char *tag) {
    for (const XmlNode *c : n->all(tag)) {
        const char *nm = c->attr("name");
        if (nm && name == nm) return c;
    }
    return nullptr;
}

const XmlNode *requiredChild(const XmlNode *n, const std::string &name, const char *tag) {
    const XmlNode *c = namedChild(n, name, tag);
    if (!c) fail("<" + n->name + "> has no <" + tag + " name=\"" + name + "\">");
    return c;
}

float childFloat(const XmlNode *n, const char *name) { return floatAttr(requiredChild(n, name, "float"), "value"); }
float childSpectrum(const XmlNode *n, const char *name) {
    return floatAttr(requiredChild(n, name, "spectrum"), "value");
}
Vec3 childRgb(const XmlNode *n, const char *name) { return parseVec3(requiredChild(n, name, "rgb")->attr("value")); }
std::string childString(const XmlNode *n, const char *name) {
    return requiredAttr(requiredChild(n, name, "string"), "value");
}
Vec3 childPoint(const XmlNode *n, const char *name) {
    const XmlNode *c = requiredChild(n, name, "point");
    return {floatAttr(c, "x"), floatAttr(c, "y"), floatAttr(c, "z")};
}

void set3(float *d, Vec3 v) {
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
}

// parseCameraSettings (src/SceneLoader.cpp:614-648).  The fov is read only
// when the sensor has a transform; the search for it stops at the first
// <float> without a name attribute, whose value is then taken.
void parseSensor(const XmlNode *scene, Scene &s) {
    const XmlNode *sensor = scene->first("sensor");
    if (!sensor) return;
    const XmlNode *tr = sensor->first("transform");
    if (!tr) return;
    const XmlNode *look = tr->first("lookAt");
    if (!look) look = tr->first("lookat");
    if (look) {
        Vec3 o = parseVec3(look->attr("origin")), t = parseVec3(look->attr("target")), u = parseVec3(look->attr("up"));
        set3(s.view.origin, o);
        set3(s.view.target, t);
        set3(s.view.up, u);
    }
    for (const XmlNode *f : sensor->all("float")) {
        const char *nm = f->attr("name");
        if (!nm || std::strcmp(nm, "fov") == 0) {
            s.view.vfov = floatAttr(f, "value");
            break;
        }
    }
}

rtx_material blank() {
    rtx_material m;
    std::memset(&m, 0, sizeof(m));
    m.textureIdDiffuse = -1;
    m.textureIdSpecular = -1;
    return m;
}

// parseXmlBSDF (src/SceneLoader.cpp:558-611).
rtx_material parseBsdf(const XmlNode *b, const std::map<std::string, int> &textures) {
    const std::string type = requiredAttr(b, "type");
    rtx_material m = blank();
    if (type == "phong") {
        m.type = RTX_MAT_PHONG;
        set3(m.specular, childRgb(b, "specularReflectance"));
        set3(m.diffuse, childRgb(b, "diffuseReflectance"));
        m.specularHighlight = childFloat(b, "exponent");
    } else if (type == "diffuse") {
        m.type = RTX_MAT_DIFFUSE;
        const XmlNode *ref = b->first("ref");
        const char *refName = ref ? ref->attr("name") : nullptr;
        if (refName && std::strcmp(refName, "reflectance") == 0) {
            std::string id = requiredAttr(ref, "id");
            auto it = textures.find(id);
            if (it == textures.end()) fail("unknown texture id '" + id + "'");
            set3(m.diffuse, Vec3(1, 1, 1));
            m.textureIdDiffuse = it->second;
        } else {
            set3(m.diffuse, childRgb(b, "reflectance"));
        }
    } else if (type == "dielectric") {
        m.type = RTX_MAT_DIELECTRIC;
        set3(m.specular, Vec3(1, 1, 1));
        m.refractionIndex = childFloat(b, "intIOR") / childFloat(b, "extIOR");
        m.refractionIndexInv = 1.0f / m.refractionIndex;
    } else if (type == "conductor") {
        const XmlNode *mat = namedChild(b, "material", "string");
        const char *mv = mat ? mat->attr("value") : nullptr;
        if (mat && mv && std::strcmp(mv, "none") == 0) {
            m.type = RTX_MAT_SPECULAR;
            set3(m.specular, Vec3(1, 1, 1));
        } else {
            m.type = RTX_MAT_CONDUCTOR;
            m.eta = childSpectrum(b, "eta");
            m.k = childSpectrum(b, "k");
        }
    } else if (type == "roughconductor") {
        m.type = RTX_MAT_ROUGH_CONDUCTOR;
        // Mitsuba's "distribution" (default beckmann); the reference reads no
        // distribution and always renders Beckmann (src/SceneLoader.cpp:601-605)
        const XmlNode *dist = namedChild(b, "distribution", "string");
        const char *dv = dist ? dist->attr("value") : nullptr;
        if (dv && std::strcmp(dv, "ggx") == 0) m.distribution = RTX_DISTRIBUTION_GGX;
        else if (dv && std::strcmp(dv, "beckmann") != 0) fail(std::string("unsupported microfacet distribution '") + dv + "'");
        m.roughness = childFloat(b, "alpha");
        m.eta = childSpectrum(b, "eta");
        m.k = childSpectrum(b, "k");
    } else {
        fail("unknown material type '" + type + "'");
    }
    return m;
}

// parseXmlShapes (src/SceneLoader.cpp:405-514): the shape's own <bsdf>, else
// its first <ref>; an <emitter> turns a copy of that material into a light;
// an OBJ becomes one model and one instance, transforms are not read.
void parseShapes(const XmlNode *scene, Loader &ld, const std::map<std::string, int> &bsdfs,
                 const std::map<std::string, int> &textures) {
    Scene &s = ld.s;
    for (const XmlNode *shape : scene->all("shape")) {
        const std::string type = requiredAttr(shape, "type");
        int mat = -1;
        if (const XmlNode *b = shape->first("bsdf")) {
            mat = s.addMaterial(parseBsdf(b, textures));
        } else if (const XmlNode *ref = shape->first("ref")) {
            std::string id = requiredAttr(ref, "id");
            auto it = bsdfs.find(id);
            if (it == bsdfs.end()) fail("unknown bsdf id '" + id + "'");
            mat = it->second;
        }
        if (mat < 0) fail("a <shape> has no material");
        if (const XmlNode *em = shape->first("emitter")) {
            rtx_material light = s.materials[mat];
            set3(light.lightColor, childRgb(em, "radiance"));
            light.type = RTX_MAT_LIGHT;
            mat = s.addMaterial(light);
        }
        if (type == "obj") {
            ObjData obj = ld.readObjFile(childString(shape, "filename"));
            Model model;
            std::string err;
            if (!objToModel(obj, -1, mat, s.materials, model, err)) fail(err);
            s.addModelInstance(std::move(model));
        } else if (type == "sphere") {
            float radius = childFloat(shape, "radius");
            Vec3 center = childPoint(shape, "center");
            s.addSphere(center, radius, mat);
        }
    }
}

// parseEnvMap (src/SceneLoader.cpp:370-403): the scene's first <emitter>,
// when it is an envmap, is an RGBA32F lat-long map and the last light.
void parseEnvMap(const XmlNode *scene, Loader &ld) {
    const XmlNode *em = scene->first("emitter");
    if (!em) return;
    const char *type = em->attr("type");
    if (!type || std::strcmp(type, "envmap") != 0) return;
    std::string path = joinPath(ld.texturesDir, childString(em, "filename"));
    std::string err;
    if (!readExr(path, ld.s.envMap, ld.s.envWidth, ld.s.envHeight, err)) fail(err);
    rtx_light light;
    std::memset(&light, 0, sizeof(light));
    light.type = RTX_LIGHT_ENV_MAP;
    ld.s.lights.push_back(light);
}

void loadXml(const std::string &path, Scene &s) {
    XmlNode root;
    std::string err;
    if (!parseXml(readText(path), root, err)) fail(path + ": " + err);
    if (root.name != "scene") fail(path + ": the document element is not <scene>");
    Loader ld(s);
    ld.modelsDir = ld.materialsDir = ld.texturesDir = parentDir(path);
    parseSensor(&root, s);
    std::map<std::string, int> textures; // parseXmlTextures, src/SceneLoader.cpp:516-534
    for (const XmlNode *t : root.all("texture")) {
        std::string id = requiredAttr(t, "id");
        if (textures.count(id)) fail("duplicate texture id '" + id + "'");
        textures[id] = ld.addTexture(childString(t, "filename"));
    }
    std::map<std::string, int> bsdfs; // parseXmlBSDFs, src/SceneLoader.cpp:536-556
    for (const XmlNode *b : root.all("bsdf")) {
        const char *id = b->attr("id");
        std::string key = id ? id : "";
        rtx_material m = parseBsdf(b, textures);
        if (bsdfs.count(key)) fail("duplicate bsdf id '" + key + "'");
        bsdfs[key] = s.addMaterial(m);
    }
    parseShapes(&root, ld, bsdfs, textures);
    parseEnvMap(&root, ld);
}

// ----------------------------------------------------------------- JSON

// glm::translate / rotate / scale, each composed on the right (m * T) as
// parseInstances builds its matrices (src/SceneLoader.cpp:705-754).
// Column-major m[col*4+row].
void translate(Mat4 &m, Vec3 t) {
    // column 3 += t.x*col0 + t.y*col1 + t.z*col2
    for (int row = 0; row < 4; ++row)
        m.m[12 + row] = m.m[0 + row] * t.x + m.m[4 + row] * t.y + m.m[8 + row] * t.z + m.m[12 + row];
}

// Rotation by `deg` degrees about a unit axis (Rodrigues:
// R = cos I + (1 - cos) a a^T + sin [a]x), applied as m * R.
void rotate(Mat4 &m, float deg, Vec3 axis) {
    const float a = deg * 0.01745329251994329576923690768489f;
    const float c = std::cos(a), s = std::sin(a);
    const float ax[3] = {axis.x, axis.y, axis.z};
    float R[3][3]; // R[col][row]
    for (int j = 0; j < 3; ++j) {
        const float t = (1.0f - c) * ax[j];
        for (int k = 0; k < 3; ++k) R[j][k] = t * ax[k];
        R[j][j] = c + R[j][j];
    }
    R[0][1] += s * ax[2];
    R[0][2] -= s * ax[1];
    R[1][0] -= s * ax[2];
    R[1][2] += s * ax[0];
    R[2][0] += s * ax[1];
    R[2][1] -= s * ax[0];
    Mat4 out = m;
    for (int j = 0; j < 3; ++j)
        for (int row = 0; row < 4; ++row)
            out.m[j * 4 + row] = m.m[0 * 4 + row] * R[j][0] + m.m[1 * 4 + row] * R[j][1] + m.m[2 * 4 + row] * R[j][2];
    m = out;
}

void scale(Mat4 &m, Vec3 v) {
    for (int row = 0; row < 4; ++row) {
        m.m[0 + row] *= v.x;
        m.m[4 + row] *= v.y;
        m.m[8 + row] *= v.z;
    }
}

Vec3 jsonVec3(const JsonValue *v, const std::string &what) {
    if (!v || v->type != JsonValue::Array || v->array.size() < 3) fail(what + " must be an array of 3 numbers");
    float f[3];
    for (int i = 0; i < 3; ++i) {
        if (v->array[i].type != JsonValue::Number) fail(what + " must be an array of 3 numbers");
        f[i] = (float)v->array[i].number;
    }
    return {f[0], f[1], f[2]};
}

// The key of a one-entry object ({"name": value}); nlohmann's items().begin()
// gives the smallest key of a larger one.
const std::pair<const std::string, JsonValue> &onlyEntry(const JsonValue &v, const char *what) {
    if (v.type != JsonValue::Object || v.object.empty()) fail(std::string(what) + " entries are {\"name\": ...} objects");
    return *v.object.begin();
}

void loadJson(const std::string &path, Scene &s) {
    JsonValue j;
    std::string err;
    if (!parseJson(readText(path), j, err)) fail(path + ": " + err);
    Loader ld(s);
    std::string dir = parentDir(path);
    ld.modelsDir = joinPath(dir, "models/");
    ld.materialsDir = joinPath(dir, "materials/");
    ld.texturesDir = joinPath(dir, "textures/");

    // parseModels / loadModel (src/SceneLoader.cpp:114-125, 760-770)
    const JsonValue *models = j.get("models");
    if (!models || models->type != JsonValue::Array) fail(path + ": no \"models\" array");
    std::map<std::string, int> modelIndex;
    for (const JsonValue &jm : models->array) {
        const auto &e = onlyEntry(jm, "\"models\"");
        if (e.second.type != JsonValue::String) fail("model '" + e.first + "': the value is not a file name");
        ObjData obj = ld.readObjFile(e.second.string);
        const int offset = (int)s.materials.size();
        for (const ObjMaterial &om : obj.materials) { // addMaterials
            int tex = -1;
            if (!om.diffuseTex.empty()) tex = ld.addTexture(om.diffuseTex);
            if (!om.specularTex.empty()) tex = ld.addTexture(om.specularTex);
            rtx_material m;
            if (!objMaterialToRtx(om, tex, m, err)) fail(err);
            s.addMaterial(m);
        }
        Model model;
        if (!objToModel(obj, offset, -1, s.materials, model, err)) fail(e.second.string + ": " + err);
        modelIndex[e.first] = s.addModel(std::move(model));
    }

    // parseInstances (src/SceneLoader.cpp:705-758)
    const JsonValue *instances = j.get("instances");
    if (!instances || instances->type != JsonValue::Array) fail(path + ": no \"instances\" array");
    for (const JsonValue &ji : instances->array) {
        const auto &e = onlyEntry(ji, "\"instances\"");
        auto it = modelIndex.find(e.first);
        if (it == modelIndex.end()) fail("instance of unknown model '" + e.first + "'");
        const JsonValue &props = e.second;
        Mat4 transform = Mat4::identity(), normalTransform = Mat4::identity();
        if (const JsonValue *t = props.get("translate")) translate(transform, jsonVec3(t, "translate"));
        if (const JsonValue *r = props.get("rotate")) {
            Vec3 deg = jsonVec3(r, "rotate");
            for (Mat4 *m : {&transform, &normalTransform}) {
                rotate(*m, deg.x, Vec3(1, 0, 0));
                rotate(*m, deg.y, Vec3(0, 1, 0));
                rotate(*m, deg.z, Vec3(0, 0, 1));
            }
        }
        if (const JsonValue *sc = props.get("scale")) {
            Vec3 v = jsonVec3(sc, "scale");
            scale(transform, v);
            scale(normalTransform, Vec3(1.0f / v.x, 1.0f / v.y, 1.0f / v.z));
        }
        s.addInstance(it->second, transform.m, normalTransform.m);
    }

    // parsePointLights, after the instances' area lights (src/SceneLoader.cpp:686-702)
    if (const JsonValue *lights = j.get("lights")) {
        if (lights->type != JsonValue::Array) fail("\"lights\" must be an array");
        for (const JsonValue &jl : lights->array)
            s.addPointLight(jsonVec3(jl.get("position"), "light position"), jsonVec3(jl.get("color"), "light color"));
    }

    // parseJsonCamera (src/SceneLoader.cpp:669-681): the field of view stays
    if (const JsonValue *cam = j.get("camera")) {
        set3(s.view.target, jsonVec3(cam->get("target"), "camera target"));
        set3(s.view.origin, jsonVec3(cam->get("origin"), "camera origin"));
        set3(s.view.up, jsonVec3(cam->get("up"), "camera up"));
    }
}

bool endsWith(const std::string &s, const char *suffix) {
    size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

} // namespace

bool loadMitsubaXml(const std::string &path, Scene &s, std::string &err) {
    try {
        loadXml(path, s);
        s.buildLightTables();
    } catch (const std::exception &e) {
        err = e.what();
        return false;
    }
    return true;
}

bool loadJsonScene(const std::string &path, Scene &s, std::string &err) {
    try {
        loadJson(path, s);
        s.buildLightTables();
    } catch (const std::exception &e) {
        err = e.what();
        return false;
    }
    return true;
}

bool loadSceneFile(const std::string &path, Scene &s, std::string &err) {
    if (endsWith(path, ".xml")) return loadMitsubaXml(path, s, err);
    if (endsWith(path, ".json")) return loadJsonScene(path, s, err);
    err = "unknown scene file type (expected .xml or .json): " + path;
    return false;
}

} // namespace rtxh
