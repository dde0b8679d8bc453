// OBJ/MTL reading and OBJ -> model conversion (see obj.h).
#include "obj.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <unordered_map>

namespace rtxh {

namespace {

bool isSpace(char c) { return c == ' ' || c == '\t'; }
bool isLineEnd(char c) { return c == '\0' || c == '\r' || c == '\n'; }
bool isDigit(char c) { return c >= '0' && c <= '9'; }

// Splits a file into lines at "\n", "\r\n" or "\r" (tinyobjloader's
// safeGetline, tiny_obj_loader.h:708-740).
std::vector<std::string> splitLines(const std::string &text) {
    std::vector<std::string> lines;
    std::string cur;
    for (size_t i = 0; i < text.size(); ++i) {
        char c = text[i];
        if (c == '\n') {
            lines.push_back(cur);
            cur.clear();
        } else if (c == '\r') {
            if (i + 1 < text.size() && text[i + 1] == '\n') ++i;
            lines.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (!cur.empty()) lines.push_back(cur);
    return lines;
}

bool readFile(const std::string &path, std::string &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

// parseReal: one whitespace-delimited number, `def` when it does not parse.
float parseReal(const char *&t, double def = 0.0) {
    t += std::strspn(t, " \t");
    const char *end = t + std::strcspn(t, " \t\r");
    double v = def;
    objParseDouble(t, end, &v);
    t = end;
    return (float)v;
}

int parseInt(const char *&t) {
    t += std::strspn(t, " \t");
    int i = std::atoi(t);
    t += std::strcspn(t, " \t\r");
    return i;
}

std::string parseWord(const char *&t) {
    t += std::strspn(t, " \t");
    size_t n = std::strcspn(t, " \t\r");
    std::string s(t, n);
    t += n;
    return s;
}

// 1-based, or negative = relative to the records read so far; 0 is invalid
// (fixIndex, tiny_obj_loader.h:748-769).
bool fixIndex(int idx, int n, int &out) {
    if (idx > 0) {
        out = idx - 1;
        return true;
    }
    if (idx < 0) {
        out = n + idx;
        return true;
    }
    return false;
}

// v, v/vt, v//vn or v/vt/vn (parseTriple, tiny_obj_loader.h:1078-1130).
bool parseCorner(const char *&t, int nv, int nvn, int nvt, ObjCorner &c) {
    c = ObjCorner();
    if (!fixIndex(std::atoi(t), nv, c.v)) return false;
    t += std::strcspn(t, "/ \t\r");
    if (*t != '/') return true;
    ++t;
    if (*t == '/') {
        ++t;
        if (!fixIndex(std::atoi(t), nvn, c.vn)) return false;
        t += std::strcspn(t, "/ \t\r");
        return true;
    }
    if (!fixIndex(std::atoi(t), nvt, c.vt)) return false;
    t += std::strcspn(t, "/ \t\r");
    if (*t != '/') return true;
    ++t;
    if (!fixIndex(std::atoi(t), nvn, c.vn)) return false;
    t += std::strcspn(t, "/ \t\r");
    return true;
}

// Crossing-number point-in-polygon test (W. R. Franklin's pnpoly, the test
// tinyobjloader uses for ear candidates).
bool pointInTriangle(const float *x, const float *y, float tx, float ty) {
    bool inside = false;
    for (int i = 0, j = 2; i < 3; j = i++) {
        if (((y[i] > ty) != (y[j] > ty)) && (tx < (x[j] - x[i]) * (ty - y[i]) / (y[j] - y[i]) + x[i]))
            inside = !inside;
    }
    return inside;
}

// Triangulates one polygon into `out` the way exportGroupsToShape does with
// triangulate = true (tiny_obj_loader.h:1343-1577): project on the two axes
// of the first non-degenerate corner's dominant normal component, then clip
// ears, starting from the first corner and advancing past reflex or occupied
// candidates; a polygon that cannot be clipped any further is dropped.
// Positions are those read up to this point of the file.
void triangulate(const std::vector<ObjCorner> &poly, const std::vector<float> &v, int material,
                 ObjData &out) {
    const size_t n0 = poly.size();
    if (n0 < 3) return;
    auto emit = [&](const ObjCorner &a, const ObjCorner &b, const ObjCorner &c) {
        out.corners.push_back(a);
        out.corners.push_back(b);
        out.corners.push_back(c);
        out.triMaterial.push_back(material);
    };
    auto has = [&](int vi, size_t comp) { return vi >= 0 && (size_t)vi * 3 + comp < v.size(); };

    size_t axes[2] = {1, 2};
    for (size_t k = 0; k < n0; ++k) {
        int a = poly[k].v, b = poly[(k + 1) % n0].v, c = poly[(k + 2) % n0].v;
        if (!has(a, 2) || !has(b, 2) || !has(c, 2)) continue;
        float e0[3], e1[3];
        for (int d = 0; d < 3; ++d) {
            e0[d] = v[3 * b + d] - v[3 * a + d];
            e1[d] = v[3 * c + d] - v[3 * b + d];
        }
        float cx = std::fabs(e0[1] * e1[2] - e0[2] * e1[1]);
        float cy = std::fabs(e0[2] * e1[0] - e0[0] * e1[2]);
        float cz = std::fabs(e0[0] * e1[1] - e0[1] * e1[0]);
        const float eps = std::numeric_limits<float>::epsilon();
        if (cx > eps || cy > eps || cz > eps) {
            if (!(cx > cy && cx > cz)) {
                axes[0] = 0;
                if (cz > cx && cz > cy) axes[1] = 1;
            }
            break;
        }
    }

    float area = 0;
    for (size_t k = 0; k < n0; ++k) {
        int a = poly[k].v, b = poly[(k + 1) % n0].v;
        if (!has(a, axes[0]) || !has(a, axes[1]) || !has(b, axes[0]) || !has(b, axes[1])) continue;
        area += (v[3 * a + axes[0]] * v[3 * b + axes[1]] - v[3 * a + axes[1]] * v[3 * b + axes[0]]) * 0.5f;
    }

    std::vector<ObjCorner> rest = poly;
    size_t guess = 0;
    size_t triesLeft = n0, lastCount = rest.size();
    while (rest.size() > 3 && triesLeft > 0) {
        const size_t n = rest.size();
        if (guess >= n) guess -= n;
        if (lastCount != n) {
            lastCount = n;
            triesLeft = n;
        } else {
            --triesLeft;
        }
        ObjCorner ind[3];
        float x[3], y[3];
        for (size_t k = 0; k < 3; ++k) {
            ind[k] = rest[(guess + k) % n];
            int vi = ind[k].v;
            if (!has(vi, axes[0]) || !has(vi, axes[1])) {
                x[k] = y[k] = 0.0f;
            } else {
                x[k] = v[3 * vi + axes[0]];
                y[k] = v[3 * vi + axes[1]];
            }
        }
        float cross = (x[1] - x[0]) * (y[2] - y[1]) - (y[1] - y[0]) * (x[2] - x[1]);
        if (cross * area < 0.0f) { // reflex corner
            guess += 1;
            continue;
        }
        bool occupied = false;
        for (size_t o = 3; o < n; ++o) {
            int vi = rest[(guess + o) % n].v;
            if (!has(vi, axes[0]) || !has(vi, axes[1])) continue;
            if (pointInTriangle(x, y, v[3 * vi + axes[0]], v[3 * vi + axes[1]])) {
                occupied = true;
                break;
            }
        }
        if (occupied) {
            guess += 1;
            continue;
        }
        emit(ind[0], ind[1], ind[2]);
        rest.erase(rest.begin() + (long)((guess + 1) % n)); // the ear's middle corner
    }
    if (rest.size() == 3) emit(rest[0], rest[1], rest[2]);
}

} // namespace

bool objParseDouble(const char *s, const char *end, double *out) {
    if (s >= end) return false;
    double mantissa = 0.0;
    int exponent = 0;
    bool negative = false, expNegative = false;
    const char *p = s;
    bool leadingDot = false;
    if (*p == '+' || *p == '-') {
        negative = *p == '-';
        ++p;
        if (p != end && *p == '.') leadingDot = true;
    } else if (*p == '.') {
        leadingDot = true;
    } else if (!isDigit(*p)) {
        return false;
    }
    if (!leadingDot) {
        int digits = 0;
        while (p != end && isDigit(*p)) {
            mantissa = mantissa * 10 + (*p - '0');
            ++p;
            ++digits;
        }
        if (digits == 0) return false;
    }
    if (p != end) {
        if (*p == '.') {
            ++p;
            static const double scale[8] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
            int k = 1;
            while (p != end && isDigit(*p)) {
                mantissa += (*p - '0') * (k < 8 ? scale[k] : std::pow(10.0, -k));
                ++k;
                ++p;
            }
        }
        if (p != end && (*p == 'e' || *p == 'E')) {
            ++p;
            if (p != end && (*p == '+' || *p == '-')) {
                expNegative = *p == '-';
                ++p;
            } else if (p == end || !isDigit(*p)) {
                return false;
            }
            int digits = 0;
            while (p != end && isDigit(*p)) {
                // Held below 10^6: any larger exponent already gives inf or 0.
                if (exponent < 100000) exponent = exponent * 10 + (*p - '0');
                ++p;
                ++digits;
            }
            if (expNegative) exponent = -exponent;
            if (digits == 0) return false;
        }
    }
    double mag = exponent ? std::ldexp(mantissa * std::pow(5.0, exponent), exponent) : mantissa;
    *out = (negative ? -1 : 1) * mag;
    return true;
}

bool readMtl(const std::string &path, std::vector<ObjMaterial> &materials,
             std::vector<std::pair<std::string, int>> &names, std::string &err) {
    std::string text;
    if (!readFile(path, text)) {
        err = "cannot open material file " + path;
        return false;
    }
    ObjMaterial m;
    bool hasKd = false;
    auto flush = [&]() {
        bool known = false;
        for (auto &nm : names) known |= nm.first == m.name;
        if (!known) names.emplace_back(m.name, (int)materials.size()); // the first definition wins
        materials.push_back(m);
    };
    for (std::string line : splitLines(text)) {
        size_t last = line.find_last_not_of(" \t");
        line = last == std::string::npos ? std::string() : line.substr(0, last + 1);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        const char *t = line.c_str();
        t += std::strspn(t, " \t");
        if (*t == '\0' || *t == '#') continue;
        if (!std::strncmp(t, "newmtl", 6) && isSpace(t[6])) {
            if (!m.name.empty()) flush();
            m = ObjMaterial();
            m.name = t + 7;
            continue;
        }
        auto rgb = [&](float *dst) {
            t += 2;
            for (int i = 0; i < 3; ++i) dst[i] = parseReal(t);
        };
        if (t[0] == 'K' && t[1] == 'd' && isSpace(t[2])) {
            rgb(m.diffuse);
            hasKd = true;
        } else if (t[0] == 'K' && t[1] == 's' && isSpace(t[2])) {
            rgb(m.specular);
        } else if (t[0] == 'K' && t[1] == 'e' && isSpace(t[2])) {
            rgb(m.emission);
        } else if (t[0] == 'N' && t[1] == 'i' && isSpace(t[2])) {
            t += 2;
            m.ior = parseReal(t);
        } else if (t[0] == 'N' && t[1] == 's' && isSpace(t[2])) {
            t += 2;
            m.shininess = parseReal(t);
        } else if (!std::strncmp(t, "illum", 5) && isSpace(t[5])) {
            t += 6;
            m.illum = parseInt(t);
        } else if ((!std::strncmp(t, "map_Kd", 6) || !std::strncmp(t, "map_Ks", 6)) && isSpace(t[6])) {
            const bool diffuseMap = t[5] == 'd';
            t += 7;
            // ParseTextureNameAndOption (tiny_obj_loader.h:1164-1254): skip
            // the "-option values" groups; the file name is the rest of the line.
            std::string name;
            while (!isLineEnd(*t)) {
                t += std::strspn(t, " \t");
                struct Opt {
                    const char *key;
                    int reals; // numbers that follow; -1 = one word
                };
                static const Opt opts[] = {{"-blendu", -1}, {"-blendv", -1}, {"-clamp", -1}, {"-boost", 1},
                                           {"-bm", 1},      {"-o", 3},       {"-s", 3},      {"-t", 3},
                                           {"-type", -1},   {"-texres", -1}, {"-imfchan", -1}, {"-mm", 2},
                                           {"-colorspace", -1}};
                bool matched = false;
                for (const Opt &o : opts) {
                    size_t L = std::strlen(o.key);
                    if (!std::strncmp(t, o.key, L) && isSpace(t[L])) {
                        t += L;
                        if (o.reals < 0) parseWord(t);
                        else
                            for (int i = 0; i < o.reals; ++i) parseReal(t);
                        matched = true;
                        break;
                    }
                }
                if (!matched) {
                    name = t;
                    t += name.size();
                }
            }
            if (diffuseMap) {
                m.diffuseTex = name;
                if (!hasKd) m.diffuse[0] = m.diffuse[1] = m.diffuse[2] = 0.6f;
            } else {
                m.specularTex = name;
            }
        }
    }
    flush();
    return true;
}

bool readObj(const std::string &path, const std::string &mtlDir, ObjData &out, std::string &err) {
    std::string text;
    if (!readFile(path, text)) {
        err = "cannot open OBJ file " + path;
        return false;
    }
    out = ObjData();
    std::vector<std::pair<std::string, int>> matNames;
    int material = -1;
    // faces of the current group, triangulated at each usemtl change, g, o
    // and at the end of the file, against the positions read so far
    std::vector<std::vector<ObjCorner>> pending;
    auto flushFaces = [&]() {
        for (auto &poly : pending) triangulate(poly, out.v, material, out);
        pending.clear();
    };
    size_t lineNo = 0;
    for (std::string &line : splitLines(text)) {
        ++lineNo;
        const char *t = line.c_str();
        t += std::strspn(t, " \t");
        if (*t == '\0' || *t == '#') continue;
        if (t[0] == 'v' && isSpace(t[1])) {
            t += 2;
            for (int i = 0; i < 3; ++i) out.v.push_back(parseReal(t));
        } else if (t[0] == 'v' && t[1] == 'n' && isSpace(t[2])) {
            t += 3;
            for (int i = 0; i < 3; ++i) out.vn.push_back(parseReal(t));
        } else if (t[0] == 'v' && t[1] == 't' && isSpace(t[2])) {
            t += 3;
            for (int i = 0; i < 2; ++i) out.vt.push_back(parseReal(t));
        } else if ((t[0] == 'f' || t[0] == 'l' || t[0] == 'p') && isSpace(t[1])) {
            const bool face = t[0] == 'f';
            t += 2;
            t += std::strspn(t, " \t");
            std::vector<ObjCorner> poly;
            while (!isLineEnd(*t)) {
                ObjCorner c;
                if (!parseCorner(t, (int)(out.v.size() / 3), (int)(out.vn.size() / 3), (int)(out.vt.size() / 2), c)) {
                    err = path + ":" + std::to_string(lineNo) + ": invalid index (0 is not allowed)";
                    return false;
                }
                poly.push_back(c);
                t += std::strspn(t, " \t\r");
            }
            if (face) pending.push_back(std::move(poly)); // lines and points draw nothing
        } else if (!std::strncmp(t, "usemtl", 6)) {
            t += 6;
            std::string name = parseWord(t);
            int id = -1;
            for (auto &nm : matNames)
                if (nm.first == name) id = nm.second;
            if (id < 0) out.warnings += "material '" + name + "' not found\n";
            if (id != material) {
                flushFaces();
                material = id;
            }
        } else if (!std::strncmp(t, "mtllib", 6) && isSpace(t[6])) {
            // the first of the listed files that can be read
            std::istringstream names(std::string(t + 7));
            std::string file;
            bool found = false, any = false;
            while (!found && std::getline(names, file, ' ')) {
                if (file.empty()) continue;
                any = true;
                std::string merr;
                std::string full = mtlDir.empty() ? file : mtlDir + (mtlDir.back() == '/' ? "" : "/") + file;
                found = readMtl(full, out.materials, matNames, merr);
            }
            if (any && !found) out.warnings += "no material file of '" + std::string(t + 7) + "' could be read\n";
        } else if ((t[0] == 'g' || t[0] == 'o') && isSpace(t[1])) {
            flushFaces();
        }
    }
    flushFaces();
    // indices past the records (the reference would read out of bounds)
    for (const ObjCorner &c : out.corners) {
        if (c.v < 0 || (size_t)c.v * 3 + 2 >= out.v.size() || c.vn >= (int)(out.vn.size() / 3) ||
            c.vt >= (int)(out.vt.size() / 2) || c.vn < -1 || c.vt < -1) {
            err = path + ": a face refers to a vertex, normal or texture coordinate that does not exist";
            return false;
        }
    }
    return true;
}

bool objMaterialToRtx(const ObjMaterial &in, int textureId, rtx_material &m, std::string &err) {
    std::memset(&m, 0, sizeof(m));
    for (int i = 0; i < 3; ++i) {
        m.lightColor[i] = in.emission[i];
        m.diffuse[i] = in.diffuse[i];
        m.specular[i] = in.specular[i];
    }
    m.specularHighlight = in.shininess;
    m.refractionIndex = in.ior;
    m.refractionIndexInv = 1.0f / in.ior;
    m.textureIdDiffuse = textureId;
    m.textureIdSpecular = -1;
    switch (in.illum) {
    case 0:
    case 1: m.type = RTX_MAT_DIFFUSE; break;
    case 2: m.type = RTX_MAT_PHONG; break;
    case 3: m.type = RTX_MAT_SPECULAR; break;
    case 4:
    case 7: m.type = RTX_MAT_DIELECTRIC; break;
    case 11: m.type = RTX_MAT_LIGHT; break;
    default: err = "material '" + in.name + "': unknown illum mode " + std::to_string(in.illum); return false;
    }
    return true;
}

namespace {
struct VertexKey {
    rtx_vertex v;
    bool operator==(const VertexKey &o) const { // float == (so -0 equals +0)
        for (int i = 0; i < 3; ++i)
            if (!(v.pos[i] == o.v.pos[i]) || !(v.normal[i] == o.v.normal[i])) return false;
        return v.texCoord[0] == o.v.texCoord[0] && v.texCoord[1] == o.v.texCoord[1] &&
               v.materialIndex == o.v.materialIndex;
    }
};
struct VertexHash {
    static size_t bits(float f) {
        if (f == 0.0f) return 0; // +0 and -0 compare equal, so they hash alike
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return u;
    }
    size_t operator()(const VertexKey &k) const {
        size_t h = 17;
        for (int i = 0; i < 3; ++i) h = h * 31 + bits(k.v.pos[i]);
        for (int i = 0; i < 3; ++i) h = h * 31 + bits(k.v.normal[i]);
        for (int i = 0; i < 2; ++i) h = h * 31 + bits(k.v.texCoord[i]);
        return h * 31 + (size_t)(uint32_t)k.v.materialIndex;
    }
};
} // namespace

bool objToModel(const ObjData &obj, int materialOffset, int materialOverride,
                const std::vector<rtx_material> &materials, Model &model, std::string &err) {
    model = Model();
    const bool hasNormals = !obj.vn.empty(), hasTexCoords = !obj.vt.empty();
    std::unordered_map<VertexKey, uint32_t, VertexHash> unique;
    const size_t faces = obj.corners.size() / 3;
    for (size_t f = 0; f < faces; ++f) {
        int mat = materialOverride >= 0 ? materialOverride : materialOffset + obj.triMaterial[f];
        if (mat < 0 || mat >= (int)materials.size()) {
            err = "a face has no material (no usemtl before it and no material in the scene)";
            return false;
        }
        rtx_vertex fv[3];
        for (int i = 0; i < 3; ++i) {
            const ObjCorner &c = obj.corners[3 * f + i];
            rtx_vertex &v = fv[i];
            std::memset(&v, 0, sizeof(v));
            for (int d = 0; d < 3; ++d) v.pos[d] = obj.v[3 * c.v + d];
            if (hasNormals) {
                if (c.vn < 0) {
                    err = "a face has no normal index in a file with normals";
                    return false;
                }
                for (int d = 0; d < 3; ++d) v.normal[d] = obj.vn[3 * c.vn + d];
            }
            if (hasTexCoords) {
                if (c.vt < 0) {
                    err = "a face has no texture coordinate index in a file with texture coordinates";
                    return false;
                }
                v.texCoord[0] = obj.vt[2 * c.vt];
                v.texCoord[1] = 1.0f - obj.vt[2 * c.vt + 1];
            }
            v.materialIndex = mat;
        }
        if (!hasNormals) {
            // glm::normalize(glm::cross(b - a, c - a)): x * (1 / sqrt(dot(x, x)))
            Vec3 a(fv[0].pos[0], fv[0].pos[1], fv[0].pos[2]);
            Vec3 b(fv[1].pos[0], fv[1].pos[1], fv[1].pos[2]);
            Vec3 c(fv[2].pos[0], fv[2].pos[1], fv[2].pos[2]);
            Vec3 n = cross(b - a, c - a);
            float inv = 1.0f / std::sqrt(dot(n, n));
            for (int i = 0; i < 3; ++i) {
                fv[i].normal[0] = n.x * inv;
                fv[i].normal[1] = n.y * inv;
                fv[i].normal[2] = n.z * inv;
            }
        }
        for (int i = 0; i < 3; ++i) {
            VertexKey key{fv[i]};
            auto it = unique.find(key);
            if (it != unique.end()) {
                model.indices.push_back(it->second);
                continue;
            }
            // A vertex with a NaN component never equals itself: the
            // reference appends it and then, looking it up again with
            // operator[], gets a fresh entry holding index 0.
            const bool selfEqual = key == key;
            if (selfEqual) unique.emplace(key, (uint32_t)model.vertices.size());
            model.indices.push_back(selfEqual ? (uint32_t)model.vertices.size() : 0u);
            model.vertices.push_back(fv[i]);
        }
    }
    return true;
}

} // namespace rtxh
