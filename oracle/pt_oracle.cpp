// CPU oracle: serial per-path restatement of the reference's path-tracing
// loop.  TEST INFRASTRUCTURE ONLY (see oracle.h).  Every function names the
// reference lines it follows; "ref:" citations are paths in
// FelixFifi/rtx-pathtracer.
#include "oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

// ------------------------------------------------------------------ vectors
template <class R> struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> V3<R> operator*(R s, V3<R> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class R> R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class R> R length(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> V3<R> normalize(V3<R> a) { return a / length(a); }
// GLSL reflect / refract (GLSL 4.60 spec 8.5)
template <class R> V3<R> reflect(V3<R> i, V3<R> n) { return i - R(2) * dot(n, i) * n; }
template <class R> V3<R> refract(V3<R> i, V3<R> n, R eta) {
    R d = dot(n, i);
    R k = R(1) - eta * eta * (R(1) - d * d);
    if (k < R(0)) return V3<R>(0, 0, 0);
    return eta * i - (eta * d + std::sqrt(k)) * n;
}
template <class R> V3<R> f3(const float *p) { return V3<R>((R)p[0], (R)p[1], (R)p[2]); }
template <class R> V3<R> fromD(const double *p) { return V3<R>((R)p[0], (R)p[1], (R)p[2]); }
template <class R> void store3(float *d, V3<R> v) {
    d[0] = (float)v.x;
    d[1] = (float)v.y;
    d[2] = (float)v.z;
}
template <class R> void toD(V3<R> v, double *p) {
    p[0] = (double)v.x;
    p[1] = (double)v.y;
    p[2] = (double)v.z;
}

// column-major mat4 * vec4 (xyz part), GLM order (c0*x + c1*y) + (c2*z + c3*w)
template <class R> V3<R> xform(const float *m, V3<R> v, R w) {
    R o[3];
    for (int r = 0; r < 3; ++r) {
        R a0 = (R)m[0 * 4 + r] * v.x + (R)m[1 * 4 + r] * v.y;
        R a1 = (R)m[2 * 4 + r] * v.z + (R)m[3 * 4 + r] * w;
        o[r] = a0 + a1;
    }
    return {o[0], o[1], o[2]};
}

template <class R> struct Consts {
    static constexpr R PI = R(3.1415926535897932384626433832795);
    static constexpr R INV_PI = R(1) / PI;
    static constexpr R INV_2PI = R(1) / (R(2) * PI);
    static constexpr R HALF_PI = PI / R(2);
    static constexpr R TMIN = R(0.001);    // ref: shaders/raytrace.rgen:52
    static constexpr R TMAX = R(1000000.0); // ref: shaders/raytrace.rgen:53
};

// ------------------------------------------------------------- RNG (ref: random.glsl)
uint32_t tea(uint32_t val0, uint32_t val1) { // ref: shaders/random.glsl:13-27
    uint32_t v0 = val0, v1 = val1, s0 = 0;
    for (uint32_t n = 0; n < 16; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

uint32_t pcg32First(uint64_t initstate, uint64_t initseq) { // O'Neill 2014, pcg32_srandom_r
    uint64_t state = 0, inc = (initseq << 1u) | 1u;
    auto step = [&]() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    };
    step();
    state += initstate;
    step();
    return step();
}

// Bounded rejection loops: the reference's do/while loops have no bound;
// 2^16 rejections in a row has probability < 0.48^65536, so the bound only
// guarantees termination and never changes a result.
constexpr int kMaxRejections = 65536;

// Path signature: FNV-1a-style fold of every discrete traversal outcome in
// path order (extend-ray primitive ids, shadow-ray visibility, probe-ray
// primitive ids).  Two runs of one path with equal signatures and equal
// final LCG states took the same branches everywhere.
constexpr uint32_t kSigVisible = 0x1b873593u;
constexpr uint32_t kSigOccluded = 0x5bd1e995u;
constexpr uint32_t kSigProbe = 0x85ebca6bu;
constexpr uint32_t kSigCacheHit = 0xc2b2ae35u; // an irradiance-cache lookup that ended the path
constexpr uint32_t kSigRRKill = 0x27d4eb2fu;   // a weight-window roulette that ended the path
constexpr uint32_t kSigSplit = 0x165667b1u;    // ^ number of split records a vertex queued
constexpr uint32_t kDirInvalid = 0xFFFFFFFFu;  // DirectionalData.flags INVALID, guiding.glsl:112

// Irradiance cache (shaders/raytrace.rgen:1228-1421, src/IrradianceCache.h:17-30)
constexpr int kMaxNewCacheEntries = 5;  // MAX_NEW_IRRADIANCE_ENTRIES, raytrace.rgen:65
constexpr int kIcStrataPhi = 20;        // N in calculateCacheData, raytrace.rgen:1232
constexpr int kIcStrataTheta = 10;      // M, raytrace.rgen:1233
constexpr int kIcStrata = kIcStrataPhi * kIcStrataTheta;
// Keys of the per-stratum random streams of calculateCacheData: the reference
// draws the 20 phi jitters and 200 sub-paths from one sequential per-pixel
// stream; here each phi column and each stratum owns a TEA-seeded stream so
// the 200 sub-paths of a cache record run in parallel.
constexpr uint32_t kIcPhiKey = 0x9e3779b9u;
constexpr uint32_t kIcStratumKey = 0x7f4a7c15u;

// ADRRS: at most this many queued split records per pixel and frame
// (MAX_SPLITS, raytrace.rgen:82).
constexpr int kMaxSplits = 10;
// Test statistics over all frames rendered since the last orc_adrrs_counters
// call: split records queued, paths ended by the window's roulette, split
// paths traced.
std::atomic<uint64_t> g_adrrsCount[3];

// The cache as a frame reads it: the reference's FP32 buffers.
struct IcSnapshot {
    rtx_cache_header hdr;
    std::vector<rtx_cache_data> e;
    std::vector<rtx_sphere> sph;
    uint32_t count() const { return std::min(hdr.nextCacheSlot, hdr.maxCaches); }
};

// ------------------------------------------------------------------ scene
struct Node {
    double lo[3], hi[3];
    int left, right; // children (internal) or -1
    int start, count; // leaf range into primOrder
};

struct HitRec {
    double t;
    int prim;
    double u, v;
};

} // namespace

struct orc_scene {
    std::vector<std::vector<rtx_vertex>> verts;
    std::vector<std::vector<uint32_t>> idx;
    std::vector<rtx_instance> inst;
    std::vector<rtx_material> mats;
    std::vector<rtx_light> lights;
    std::vector<int32_t> lightIndex;
    std::vector<rtx_face_sample> triTable;
    uint32_t triRows = 0;
    std::vector<rtx_sphere> spheres;
    // world-space triangles: 9 doubles each (instance transform applied in FP64)
    std::vector<double> tri;
    std::vector<float> triF; // the same in FP32 (transform applied in FP32)
    std::vector<int> triInst, triLocal;
    size_t ntri = 0;
    std::vector<Node> nodes;
    std::vector<int> primOrder;
    // path-guiding regions and their mixtures (orc_guiding_set)
    std::vector<rtx_aabb> regions;
    std::vector<rtx_vmm_theta> vmms;
    // ADRRS pixel estimates (estimateImage, raytrace.rgen:120), RGBA per pixel
    std::vector<float> estimate;
    uint32_t estW = 0, estH = 0;
    // textures (binding 7): the environment map (slot 0, RGBA32F) and the
    // 8-bit RGBA material textures (slots 1..), sRGB decoded through srgb[]
    std::vector<float> env;
    uint32_t envW = 0, envH = 0;
    struct Texture {
        uint32_t w = 0, h = 0;
        std::vector<uint8_t> rgba;
    };
    std::vector<Texture> textures;
    float srgb[256];
    std::vector<uint8_t> triAlpha; // per world triangle: the any-hit alpha test applies
};

namespace {

void triBounds(const orc_scene &s, int p, double lo[3], double hi[3]) {
    if (p < (int)s.ntri) {
        const double *t = &s.tri[(size_t)p * 9];
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min({t[a], t[3 + a], t[6 + a]});
            hi[a] = std::max({t[a], t[3 + a], t[6 + a]});
        }
    } else {
        const rtx_sphere &sp = s.spheres[p - s.ntri];
        for (int a = 0; a < 3; ++a) {
            lo[a] = (double)sp.center[a] - std::fabs((double)sp.radius);
            hi[a] = (double)sp.center[a] + std::fabs((double)sp.radius);
        }
    }
    // pad so that FP32 hits computed slightly outside the exact triangle are kept
    for (int a = 0; a < 3; ++a) {
        double pad = 1e-5 * (std::fabs(lo[a]) + std::fabs(hi[a])) + 1e-7;
        lo[a] -= pad;
        hi[a] += pad;
    }
}

// Plain median-split BVH (the oracle does not care about speed beyond
// "finishes in seconds"; closest-hit answers do not depend on the tree).
int buildNode(orc_scene &s, std::vector<std::array<double, 6>> &box, int start, int count) {
    Node n;
    for (int a = 0; a < 3; ++a) {
        n.lo[a] = std::numeric_limits<double>::infinity();
        n.hi[a] = -std::numeric_limits<double>::infinity();
    }
    for (int i = start; i < start + count; ++i) {
        const auto &b = box[s.primOrder[i]];
        for (int a = 0; a < 3; ++a) {
            n.lo[a] = std::min(n.lo[a], b[a]);
            n.hi[a] = std::max(n.hi[a], b[3 + a]);
        }
    }
    int id = (int)s.nodes.size();
    s.nodes.push_back(n);
    if (count <= 4) {
        s.nodes[id].left = s.nodes[id].right = -1;
        s.nodes[id].start = start;
        s.nodes[id].count = count;
        return id;
    }
    double clo[3], chi[3];
    for (int a = 0; a < 3; ++a) {
        clo[a] = std::numeric_limits<double>::infinity();
        chi[a] = -clo[a];
    }
    for (int i = start; i < start + count; ++i) {
        const auto &b = box[s.primOrder[i]];
        for (int a = 0; a < 3; ++a) {
            double c = 0.5 * (b[a] + b[3 + a]);
            clo[a] = std::min(clo[a], c);
            chi[a] = std::max(chi[a], c);
        }
    }
    int axis = 0;
    for (int a = 1; a < 3; ++a)
        if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
    int mid = start + count / 2;
    std::nth_element(s.primOrder.begin() + start, s.primOrder.begin() + mid,
                     s.primOrder.begin() + start + count, [&](int x, int y) {
                         double cx = box[x][axis] + box[x][3 + axis];
                         double cy = box[y][axis] + box[y][3 + axis];
                         return cx < cy || (cx == cy && x < y);
                     });
    int l = buildNode(s, box, start, mid - start);
    int r = buildNode(s, box, mid, start + count - mid);
    s.nodes[id].left = l;
    s.nodes[id].right = r;
    s.nodes[id].start = s.nodes[id].count = 0;
    return id;
}

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), the test
// the GPU kernel uses too.  Barycentrics (u, v) weight vertices 1 and 2,
// as the reference's hit attributes do (raytrace.rchit:28).
template <class R> struct RayPre {
    V3<R> o, d;
    int kx, ky, kz;
    R Sx, Sy, Sz;
    explicit RayPre(V3<R> org, V3<R> dir) : o(org), d(dir) {
        R ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
        kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
        kx = (kz + 1) % 3;
        ky = (kx + 1) % 3;
        if (d[kz] < R(0)) std::swap(kx, ky);
        Sx = d[kx] / d[kz];
        Sy = d[ky] / d[kz];
        Sz = R(1) / d[kz];
    }
};

template <class R>
bool isectTri(const RayPre<R> &rp, V3<R> v0, V3<R> v1, V3<R> v2, R tmin, R tmax, R &t, R &u, R &v) {
    V3<R> A = v0 - rp.o, B = v1 - rp.o, C = v2 - rp.o;
    R Ax = A[rp.kx] - rp.Sx * A[rp.kz], Ay = A[rp.ky] - rp.Sy * A[rp.kz];
    R Bx = B[rp.kx] - rp.Sx * B[rp.kz], By = B[rp.ky] - rp.Sy * B[rp.kz];
    R Cx = C[rp.kx] - rp.Sx * C[rp.kz], Cy = C[rp.ky] - rp.Sy * C[rp.kz];
    R U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (sizeof(R) == 4 && (U == R(0) || V == R(0) || W == R(0))) {
        U = (R)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (R)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (R)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < R(0) || V < R(0) || W < R(0)) && (U > R(0) || V > R(0) || W > R(0))) return false;
    R det = U + V + W;
    if (det == R(0)) return false;
    R Az = rp.Sz * A[rp.kz], Bz = rp.Sz * B[rp.kz], Cz = rp.Sz * C[rp.kz];
    R T = U * Az + V * Bz + W * Cz;
    R tt = T / det;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt;
    u = V / det;
    v = W / det;
    return true;
}

// Analytic sphere (ref: shaders/raytrace.sphere.rint:11-29): both roots are
// reported; the nearest root inside [tmin, tmax] is kept.
template <class R> bool isectSphere(const rtx_sphere &s, V3<R> o, V3<R> d, R tmin, R tmax, R &t) {
    V3<R> oc = o - f3<R>(s.center);
    R dotDOC = dot(d, oc);
    R r = (R)s.radius;
    R rootTerm = dotDOC * dotDOC - dot(oc, oc) + r * r;
    if (rootTerm < R(0)) return false;
    R root = std::sqrt(rootTerm);
    R t1 = -dotDOC + root, t2 = -dotDOC - root;
    bool hit = false;
    R best = tmax;
    if (t1 >= tmin && t1 <= best) {
        best = t1;
        hit = true;
    }
    if (t2 >= tmin && t2 <= best) {
        best = t2;
        hit = true;
    }
    t = best;
    return hit;
}

// ---------------------------------------------------------------- textures
// Vulkan linear filtering with REPEAT addressing (the samplers of
// src/SceneLoader.cpp:209-227): texel centres at (i + 0.5) / size, the four
// nearest texels weighted bilinearly, indices wrapped; one axis.
template <class R> void bilerpAxis(R u, int n, int &i0, int &i1, R &w) {
    R x = u * (R)n - R(0.5);
    if (!(std::fabs(x) < R(1e30))) {
        i0 = i1 = 0;
        w = 0;
        return;
    }
    R xf = std::floor(x);
    w = x - xf;
    int i = (int)std::fmod(xf, (R)n);
    if (i < 0) i += n;
    i0 = i;
    i1 = i + 1 == n ? 0 : i + 1;
}
template <class R> void lerp4(const R *a, const R *b, const R *c, const R *d, R fx, R fy, R *out) {
    R gx = R(1) - fx, gy = R(1) - fy;
    for (int k = 0; k < 4; ++k) {
        R t = a[k] * gx + b[k] * fx, u = c[k] * gx + d[k] * fx;
        out[k] = t * gy + u * fy;
    }
}
// texture(textureSamplers[slot], uv) of an R8G8B8A8_SRGB material texture:
// RGB decoded to linear before filtering, alpha / 255.
template <class R> void texSample(const orc_scene &S, int slot, R u, R v, R out[4]) {
    const orc_scene::Texture &t = S.textures[slot];
    int x0, x1, y0, y1;
    R fx, fy;
    bilerpAxis<R>(u, (int)t.w, x0, x1, fx);
    bilerpAxis<R>(v, (int)t.h, y0, y1, fy);
    R q[4][4];
    const int xs[4] = {x0, x1, x0, x1}, ys[4] = {y0, y0, y1, y1};
    for (int k = 0; k < 4; ++k) {
        const uint8_t *p = &t.rgba[((size_t)ys[k] * t.w + xs[k]) * 4];
        for (int c = 0; c < 3; ++c) q[k][c] = (R)S.srgb[p[c]];
        q[k][3] = (R)((float)p[3] / 255.0f);
    }
    lerp4<R>(q[0], q[1], q[2], q[3], fx, fy, out);
}
// The environment map in direction `dir`, raytrace.rmiss:24-28 and
// raytrace.rgen:410-416: u = atan(x, -z) / 2pi, v = acos(y) / pi; black
// without a map (the reference's default texture, SceneLoader.cpp:80-86).
template <class R> V3<R> envColor(const orc_scene &S, V3<R> dir) {
    if (S.env.empty()) return V3<R>(0, 0, 0);
    R u = std::atan2(dir.x, -dir.z) / (R(2) * Consts<R>::PI);
    R v = std::acos(dir.y) / Consts<R>::PI;
    int x0, x1, y0, y1;
    R fx, fy;
    bilerpAxis<R>(u, (int)S.envW, x0, x1, fx);
    bilerpAxis<R>(v, (int)S.envH, y0, y1, fy);
    R q[4][4];
    const int xs[4] = {x0, x1, x0, x1}, ys[4] = {y0, y0, y1, y1};
    for (int k = 0; k < 4; ++k)
        for (int c = 0; c < 4; ++c) q[k][c] = (R)S.env[((size_t)ys[k] * S.envW + xs[k]) * 4 + c];
    R o[4];
    lerp4<R>(q[0], q[1], q[2], q[3], fx, fy, o);
    return V3<R>(o[0], o[1], o[2]);
}
// GLSL uint(float) on the reference's hardware: saturating, NaN -> 0.
inline uint32_t satU32(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
// raytrace.rahit:21-47: a candidate hit on world triangle p (barycentrics
// bu, bv of vertices 1 and 2, distance t) of a material with a diffuse
// texture is ignored when rnd(seed) > alpha at the candidate's texture
// coordinates; seed = tea(uint(u * 1e8 + origin.x * t), randomUInt), computed
// in FP32 as the shader does.
template <class R> bool alphaAccept(const orc_scene &S, int p, R ox, R t, R bu, R bv, uint32_t randomUInt) {
    const rtx_instance &in = S.inst[S.triInst[p]];
    const auto &vv = S.verts[in.modelIndex];
    const auto &ii = S.idx[in.modelIndex];
    int local = S.triLocal[p];
    const rtx_vertex &v0 = vv[ii[3 * local]], &v1 = vv[ii[3 * local + 1]], &v2 = vv[ii[3 * local + 2]];
    R b0 = R(1) - bu - bv;
    R u = (R)v0.texCoord[0] * b0 + (R)v1.texCoord[0] * bu + (R)v2.texCoord[0] * bv;
    R v = (R)v0.texCoord[1] * b0 + (R)v1.texCoord[1] * bu + (R)v2.texCoord[1] * bv;
    uint32_t st = tea(satU32((float)u * 100000000.0f + (float)ox * (float)t), randomUInt);
    st = 1664525u * st + 1013904223u;
    R r = R(st & 0x00FFFFFFu) / R(16777216.0);
    R px[4];
    texSample<R>(S, S.mats[v0.materialIndex].textureIdDiffuse, u, v, px);
    return !(r > px[3]);
}

template <class R> V3<R> triVert(const orc_scene &s, int p, int k) {
    if (sizeof(R) == 8) {
        const double *q = &s.tri[(size_t)p * 9 + 3 * k];
        return V3<R>((R)q[0], (R)q[1], (R)q[2]);
    }
    const float *q = &s.triF[(size_t)p * 9 + 3 * k];
    return V3<R>((R)q[0], (R)q[1], (R)q[2]);
}

template <class R> bool boxHit(const Node &n, const RayPre<R> &rp, double tmin, double tmax) {
    double t0 = tmin, t1 = tmax;
    for (int a = 0; a < 3; ++a) {
        double o = (double)rp.o[a], d = (double)rp.d[a];
        if (d == 0.0) {
            if (o < n.lo[a] || o > n.hi[a]) return false;
            continue;
        }
        double inv = 1.0 / d;
        double tn = (n.lo[a] - o) * inv, tf = (n.hi[a] - o) * inv;
        if (tn > tf) std::swap(tn, tf);
        t0 = std::max(t0, tn);
        t1 = std::min(t1, tf * (1 + 1e-9));
        if (t0 > t1) return false;
    }
    return true;
}

// Closest hit: smallest t; equal t -> smaller primitive id.  Triangles of
// textured materials pass the any-hit alpha test (randomUInt: the frame's).
template <class R>
bool closest(const orc_scene &s, V3<R> o, V3<R> d, R tmin, R tmax, bool brute, HitRec &h, uint32_t randomUInt = 0) {
    RayPre<R> rp(o, d);
    R bestT = tmax;
    int bestP = -1;
    R bu = 0, bv = 0;
    auto testPrim = [&](int p) {
        R t, u = 0, v = 0;
        bool hit;
        if (p < (int)s.ntri)
            hit = isectTri<R>(rp, triVert<R>(s, p, 0), triVert<R>(s, p, 1), triVert<R>(s, p, 2), tmin,
                              bestT, t, u, v);
        else
            hit = isectSphere<R>(s.spheres[p - s.ntri], o, d, tmin, bestT, t);
        if (!hit) return;                                // hits satisfy t <= bestT
        if (bestP >= 0 && t == bestT && p > bestP) return; // tie: smaller id wins
        if (p < (int)s.ntri && s.triAlpha[p] && !alphaAccept<R>(s, p, o.x, t, u, v, randomUInt)) return;
        bestT = t;
        bestP = p;
        bu = u;
        bv = v;
    };
    size_t nprim = s.ntri + s.spheres.size();
    if (brute || s.nodes.empty()) {
        for (size_t p = 0; p < nprim; ++p) testPrim((int)p);
    } else {
        std::vector<int> stack;
        stack.push_back(0);
        while (!stack.empty()) {
            const Node &n = s.nodes[stack.back()];
            stack.pop_back();
            if (!boxHit<R>(n, rp, (double)tmin, (double)bestT)) continue;
            if (n.left < 0) {
                for (int i = n.start; i < n.start + n.count; ++i) testPrim(s.primOrder[i]);
            } else {
                stack.push_back(n.left);
                stack.push_back(n.right);
            }
        }
    }
    if (bestP < 0) return false;
    h.t = (double)bestT;
    h.prim = bestP;
    h.u = (double)bu;
    h.v = (double)bv;
    return true;
}

template <class R>
bool occluded(const orc_scene &s, V3<R> o, V3<R> d, R tmin, R tmax, bool brute, uint32_t randomUInt = 0) {
    RayPre<R> rp(o, d);
    auto testPrim = [&](int p) {
        R t, u, v;
        if (p < (int)s.ntri)
            return isectTri<R>(rp, triVert<R>(s, p, 0), triVert<R>(s, p, 1), triVert<R>(s, p, 2), tmin,
                               tmax, t, u, v) &&
                   !(s.triAlpha[p] && !alphaAccept<R>(s, p, o.x, t, u, v, randomUInt));
        return isectSphere<R>(s.spheres[p - s.ntri], o, d, tmin, tmax, t);
    };
    size_t nprim = s.ntri + s.spheres.size();
    if (brute || s.nodes.empty()) {
        for (size_t p = 0; p < nprim; ++p)
            if (testPrim((int)p)) return true;
        return false;
    }
    std::vector<int> stack{0};
    while (!stack.empty()) {
        const Node &n = s.nodes[stack.back()];
        stack.pop_back();
        if (!boxHit<R>(n, rp, (double)tmin, (double)tmax)) continue;
        if (n.left < 0) {
            for (int i = n.start; i < n.start + n.count; ++i)
                if (testPrim(s.primOrder[i])) return true;
        } else {
            stack.push_back(n.left);
            stack.push_back(n.right);
        }
    }
    return false;
}

// ------------------------------------------------------------ the tracer
template <class R> struct HitInfo { // ref: shaders/raycommon.glsl:1-13
    V3<R> worldPos, normal;
    R u = 0, v = 0; // textureUV
    int matIndex = -1;
    R t = 0;
    bool isMiss = true, isFrontFace = false, isSphere = false;
    V3<R> missColor;
    int instanceIndex = 0;
    int prim = -1;
};

template <class R> struct Tracer {
    using C = Consts<R>;
    const orc_scene &S;
    const rtx_frame_params &fp;
    const rtx_push_constant &pc;
    uint32_t seed = 0;
    uint32_t primHash = 2166136261u;
    // irradiance cache: the state this frame reads, and what the path asks of it
    const IcSnapshot *ic = nullptr;
    bool icElect = false;
    int numCand = 0;
    V3<R> candO[kMaxNewCacheEntries], candN[kMaxNewCacheEntries];
    // guiding sample record of this path (raytrace.rgen:966-990, 1179-1212):
    // its RTX_MAX_DIRECTIONAL_DATA_PER_PIXEL slots, flags INVALID on entry
    rtx_directional_data *dd = nullptr;
    int curSlot = 0;      // currentSampleOffset of the running raytrace()
    int sampleOffset = 0; // first slot of the running raytrace() (raytrace.rgen:93, 1221)
    int iUpdateDistance = -1;
    R distanceFactor = R(1);
    V3<R> lightSums[RTX_MAX_DIRECTIONAL_DATA_PER_PIXEL], sampleThroughputs[RTX_MAX_DIRECTIONAL_DATA_PER_PIXEL];
    // ADRRS (raytrace.rgen:69-88): the pixel's estimate and its queued splits
    V3<R> estimate = V3<R>(0, 0, 0);
    struct SplitRec { // splitInfo, raytrace.rgen:70-80
        V3<R> origin, normal, wi, throughput;
        int matIndex, depth;
        bool frontFace;
        R u, v;
    };
    SplitRec splits[kMaxSplits];
    int nSplits = 0;
    R lastFirstT = C::TMAX;
    // the frame's randomUInt (any-hit alpha test) and the texture
    // coordinates of the vertex being shaded (diffuse(), phong())
    uint32_t randomUInt = 0;
    R texU = 0, texV = 0;

    Tracer(const orc_scene &s, const rtx_frame_params &f) : S(s), fp(f), pc(f.pc) {}

    // ref: shaders/random.glsl:31-57
    R rnd() {
        seed = 1664525u * seed + 1013904223u;
        return R(seed & 0x00FFFFFFu) / R(16777216.0);
    }
    R negPos() { return rnd() * R(2) - R(1); }
    // int(rnd() * (max + 1)) evaluated in FP32 as the shader does: the
    // product is what decides the integer, so it is kept bit-exact.
    int randomInteger(int max) {
        float u = (float)rnd();
        return (int)(u * (float)(max + 1));
    }

    // ref: shaders/transform.glsl:7-42
    static void coordinateAxis(V3<R> z, V3<R> &x, V3<R> &y) {
        if (std::fabs(z.x) > std::fabs(z.y)) {
            R invLen = R(1) / std::sqrt(z.x * z.x + z.z * z.z);
            y = V3<R>(z.z * invLen, R(0), -z.x * invLen);
        } else {
            R invLen = R(1) / std::sqrt(z.y * z.y + z.z * z.z);
            y = V3<R>(R(0), z.z * invLen, -z.y * invLen);
        }
        x = cross(y, z);
    }
    static V3<R> toWorld(V3<R> v, V3<R> n) {
        V3<R> x, y;
        coordinateAxis(n, x, y);
        return v.x * x + v.y * y + v.z * n;
    }
    static V3<R> sphericalToCartesian(R theta, R phi) {
        return V3<R>(std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta));
    }

    // ref: shaders/random.glsl:60-68
    V3<R> randomOnUnitSphere() {
        V3<R> res;
        for (int it = 0; it < kMaxRejections; ++it) {
            R a = negPos(), b = negPos(), c = negPos();
            res = V3<R>(a, b, c);
            if (!(length(res) > R(1))) break;
        }
        return normalize(res);
    }
    // ref: shaders/random.glsl:86-94
    V3<R> randomInHemisphereCosine(V3<R> n) {
        R u = rnd();
        R sqrt_u = std::sqrt(u);
        R phi = R(2) * C::PI * rnd();
        V3<R> local(sqrt_u * std::cos(phi), sqrt_u * std::sin(phi), std::sqrt(R(1) - u));
        return toWorld(local, n);
    }
    // ref: shaders/random.glsl:97-106
    V3<R> randomInHemisphereCosinePower(V3<R> refl, R p) {
        R u = rnd();
        R cosTheta = std::pow(u, R(1) / (p + R(1)));
        R phi = R(2) * C::PI * rnd();
        R sinTheta = std::sqrt(R(1) - cosTheta * cosTheta);
        V3<R> local(sinTheta * std::cos(phi), sinTheta * std::sin(phi), cosTheta);
        return toWorld(local, refl);
    }
    // ref: shaders/random.glsl:108-115
    V3<R> randomOnSphere(const rtx_sphere &s, V3<R> &normal) {
        for (int it = 0; it < kMaxRejections; ++it) {
            R a = negPos(), b = negPos(), c = negPos();
            normal = V3<R>(a, b, c);
            if (!(length(normal) > R(1))) break;
        }
        normal = normalize(normal);
        return f3<R>(s.center) + normal * (R)s.radius;
    }
    // GGX normal with density D(m) cos(theta_m) (Walter et al. 2007, eq. 35-36:
    // theta_m = atan(a sqrt(u1) / sqrt(1 - u1)), phi_m = 2 pi u2); new, drawn
    // where the reference draws its Beckmann normal (two draws, same order)
    V3<R> randomGGXNormal(const rtx_material &m, V3<R> n) {
        R a = (R)m.roughness;
        R u1 = rnd();
        R thetaM = std::atan(a * std::sqrt(u1) / std::sqrt(R(1) - u1));
        R phiM = R(2) * C::PI * rnd();
        V3<R> localM(std::sin(thetaM) * std::cos(phiM), std::sin(thetaM) * std::sin(phiM), std::cos(thetaM));
        return toWorld(localM, n);
    }
    V3<R> randomMicrofacetNormal(const rtx_material &m, V3<R> n) {
        return m.distribution == RTX_DISTRIBUTION_GGX ? randomGGXNormal(m, n) : randomBeckmannNormal(m, n);
    }
    // ref: shaders/random.glsl:127-136
    V3<R> randomBeckmannNormal(const rtx_material &m, V3<R> n) {
        R a = (R)m.roughness;
        R thetaM = std::atan(std::sqrt(-a * a * std::log(R(1) - rnd())));
        R phiM = R(2) * C::PI * rnd();
        R cosThetaNM = std::cos(thetaM);
        V3<R> localM(std::sin(thetaM) * std::cos(phiM), std::sin(thetaM) * std::sin(phiM), cosThetaNM);
        return toWorld(localM, n);
    }

    // ---------------------------------------------------------- BSDF layer
    // ref: shaders/raytrace.rgen:145-160
    static R fresnel(R eta, R cosThetaI) {
        R sinThetaSqr = eta * eta * (R(1) - cosThetaI * cosThetaI);
        if (sinThetaSqr > R(1)) return R(1);
        R cosThetaT = std::sqrt(R(1) - sinThetaSqr);
        R Rs = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
        R Rp = (cosThetaI - eta * cosThetaT) / (cosThetaI + eta * cosThetaT);
        return (Rs * Rs + Rp * Rp) / R(2);
    }
    // ref: shaders/raytrace.rgen:163-175
    static R fresnelConductor(R cosThetaI, R eta, R k) {
        if (cosThetaI < R(0)) cosThetaI = -cosThetaI;
        R ek = eta * eta + k * k;
        R c2 = cosThetaI * cosThetaI;
        R Rs2 = (ek * c2 - R(2) * eta * cosThetaI + R(1)) / (ek * c2 + R(2) * eta * cosThetaI + R(1));
        R Rp2 = (ek - R(2) * eta * cosThetaI + c2) / (ek + R(2) * eta * cosThetaI + c2);
        return (Rs2 + Rp2) / R(2);
    }
    // GGX D (Walter et al. 2007, eq. 33): a^2 / (pi cos^4 (a^2 + tan^2)^2)
    // = a^2 / (pi ((a^2 - 1) cos^2 + 1)^2) for cos > 0; new (the reference has
    // Beckmann only)
    static R ggxD(const rtx_material &m, V3<R> n, V3<R> h) {
        R c = dot(n, h);
        if (c <= R(0)) return R(0);
        R a2 = (R)m.roughness * (R)m.roughness;
        R t = (a2 - R(1)) * c * c + R(1);
        return a2 / (C::PI * t * t);
    }
    // Smith G1 of GGX (Walter et al. 2007, eq. 34):
    // chi+(v.h / v.n) 2 / (1 + sqrt(1 + a^2 tan^2 theta_v))
    static R ggxG1(const rtx_material &m, V3<R> n, V3<R> h, V3<R> v) {
        R cv = dot(v, n);
        if (dot(v, h) / cv <= R(0)) return R(0);
        R tan2 = (R(1) - cv * cv) / (cv * cv);
        if (tan2 < R(0)) tan2 = R(0);
        R a = (R)m.roughness;
        return R(2) / (R(1) + std::sqrt(R(1) + a * a * tan2));
    }
    // D of the material's microfacet distribution
    static R D(const rtx_material &m, V3<R> n, V3<R> h) {
        return m.distribution == RTX_DISTRIBUTION_GGX ? ggxD(m, n, h) : beckmannD(m, n, h);
    }
    // Beckmann D, ref: shaders/raytrace.rgen:179-202 (pow(E, x) == exp(x))
    static R beckmannD(const rtx_material &m, V3<R> n, V3<R> h) {
        R cosTheta = dot(n, h);
        if (cosTheta <= R(0)) return R(0);
        R theta = std::acos(cosTheta);
        if (std::isnan(theta) || std::isinf(theta)) theta = 0;
        R tanTheta = std::tan(theta);
        if (std::isnan(tanTheta) || std::isinf(tanTheta)) tanTheta = 0;
        R alphaSqr = (R)m.roughness * (R)m.roughness;
        R c4 = cosTheta * cosTheta * cosTheta * cosTheta;
        return std::exp(-tanTheta * tanTheta / alphaSqr) / (C::PI * alphaSqr * c4);
    }
    // ref: shaders/raytrace.rgen:204-222, kept with its two typos: tan() of a
    // cosine, and 2.577*a where 2.577*a^2 is meant.
    static R G1(const rtx_material &m, V3<R> n, V3<R> h, V3<R> v) {
        R thetaV = dot(v, n);
        R c = dot(v, h) / thetaV;
        if (c <= R(0)) return R(0);
        R a = R(1) / ((R)m.roughness * std::tan(thetaV));
        if (a >= R(1.6)) return R(1);
        R a2 = a * a;
        return (R(3.535) * a + R(2.181) * a2) / (R(1) + R(2.276) * a + R(2.577) * a);
    }
    // ref: shaders/raytrace.rgen:224-230 (Beckmann); GGX: the separable Smith
    // form G1(i) G1(o)
    static R G(const rtx_material &m, V3<R> i, V3<R> o, V3<R> n, V3<R> h) {
        if (m.distribution == RTX_DISTRIBUTION_GGX) return ggxG1(m, n, h, i) * ggxG1(m, n, h, o);
        R result = G1(m, n, h, i) * G1(m, n, h, o);
        if (result < R(0)) return R(0);
        return result;
    }
    V3<R> texRgb(int slot) const {
        R px[4];
        texSample<R>(S, slot, texU, texV, px);
        return V3<R>(px[0], px[1], px[2]);
    }
    // ref: shaders/raytrace.rgen:232-239, at the shaded vertex's texture coordinates
    V3<R> diffuse(const rtx_material &m) const {
        if (m.textureIdDiffuse != -1) return f3<R>(m.diffuse) * texRgb(m.textureIdDiffuse) / C::PI;
        return f3<R>(m.diffuse) / C::PI;
    }
    // ref: shaders/raytrace.rgen:242-267
    V3<R> phong(const rtx_material &m, V3<R> n, V3<R> wi, V3<R> wo) const {
        V3<R> res(0, 0, 0);
        R cosThetaWo = dot(wo, n);
        if (cosThetaWo > R(0)) {
            res = res + diffuse(m);
            R dotReflDir = dot(reflect(-wo, n), wi);
            if (dotReflDir > R(0)) {
                R h = (R)m.specularHighlight;
                V3<R> spec = ((h + R(2)) / (R(2) * C::PI)) * f3<R>(m.specular) * std::pow(dotReflDir, h);
                if (m.textureIdSpecular != -1) spec = spec * texRgb(m.textureIdSpecular);
                res = res + spec;
            }
        }
        return cosThetaWo * res;
    }
    // ref: shaders/raytrace.rgen:269-280
    static V3<R> roughConductor(const rtx_material &m, V3<R> n, V3<R> wi, V3<R> wo) {
        V3<R> hr = normalize(wi + wo);
        R cosThetaIHr = dot(wi, hr);
        R cosThetaI = dot(wi, n);
        R cosThetaO = dot(wo, n);
        if (cosThetaI <= R(0)) return V3<R>(0, 0, 0);
        R val = fresnelConductor(cosThetaIHr, (R)m.eta, (R)m.k) * G(m, wi, wo, n, hr) * D(m, n, hr) /
                (R(4) * cosThetaI * cosThetaO);
        return V3<R>(cosThetaO * val, cosThetaO * val, cosThetaO * val);
    }
    // ref: shaders/raytrace.rgen:284-332
    static R pdfBSDF(const rtx_material &m, V3<R> n, V3<R> wi, V3<R> wo) {
        switch (m.type) {
        case RTX_MAT_ROUGH_CONDUCTOR: {
            V3<R> hr = normalize(wi + wo);
            R pm = D(m, n, hr) * std::fabs(dot(hr, n));
            if (pm <= R(0) || dot(wo, hr) <= R(0)) return R(0);
            return pm / (R(4) * std::fabs(dot(wo, hr)));
        }
        case RTX_MAT_PHONG: {
            if (dot(n, wo) < R(0)) return R(0);
            R lD = length(f3<R>(m.diffuse)), lS = length(f3<R>(m.specular));
            R sum = lD + lS;
            if (sum == R(0)) return R(0);
            V3<R> reflected = reflect(-wi, n);
            R h = (R)m.specularHighlight;
            R pdf = 0;
            if (dot(reflected, wo) > R(0)) {
                pdf = (h + R(1)) * std::pow(dot(reflected, wo), h) / (R(2) * C::PI);
                pdf *= lS / sum;
            }
            pdf += dot(wo, n) / C::PI * lD / sum;
            return pdf;
        }
        default:
            return dot(wo, n) / C::PI;
        }
    }
    // ref: shaders/raytrace.rgen:483-552
    R sampleBSDF(const rtx_material &m, V3<R> wi, V3<R> n, bool frontFace, V3<R> &newDir) {
        switch (m.type) {
        case RTX_MAT_ROUGH_CONDUCTOR: {
            V3<R> worldM = randomMicrofacetNormal(m, n);
            newDir = reflect(-wi, worldM);
            return pdfBSDF(m, n, wi, newDir);
        }
        case RTX_MAT_PHONG: {
            R lD = length(f3<R>(m.diffuse)), lS = length(f3<R>(m.specular));
            R sum = lD + lS;
            if (sum == R(0)) return R(0);
            if (rnd() * sum > lD) {
                V3<R> reflected = reflect(-wi, n);
                newDir = randomInHemisphereCosinePower(reflected, (R)m.specularHighlight);
                if (dot(n, newDir) < R(0)) return R(0);
                return pdfBSDF(m, n, wi, newDir);
            }
            newDir = randomInHemisphereCosine(n);
            return pdfBSDF(m, n, wi, newDir);
        }
        case RTX_MAT_SPECULAR:
        case RTX_MAT_CONDUCTOR:
            newDir = reflect(-wi, n);
            return R(1);
        case RTX_MAT_DIELECTRIC: {
            R eta = frontFace ? (R)m.refractionIndexInv : (R)m.refractionIndex;
            R cosThetaI = dot(wi, n);
            R F = fresnel(eta, cosThetaI);
            if (rnd() > F) {
                newDir = refract(-wi, n, eta);
                return R(1) - F;
            }
            newDir = reflect(-wi, n);
            return F;
        }
        default:
            newDir = randomInHemisphereCosine(n);
            return dot(newDir, n) / C::PI;
        }
    }
    // ref: shaders/raytrace.rgen:557-599
    V3<R> evalBsdf(const rtx_material &m, V3<R> n, V3<R> wi, V3<R> wo, bool frontFace) const {
        switch (m.type) {
        case RTX_MAT_DIFFUSE:
        case RTX_MAT_LIGHT:
            return dot(wo, n) * diffuse(m);
        case RTX_MAT_PHONG:
            return phong(m, n, wi, wo);
        case RTX_MAT_ROUGH_CONDUCTOR: {
            V3<R> r = roughConductor(m, n, wi, wo);
            if (std::isnan(r.x)) return V3<R>(0, 0, 0);
            return r;
        }
        case RTX_MAT_DIELECTRIC: {
            R cosThetaI = dot(n, wi);
            R eta = frontFace ? (R)m.refractionIndexInv : (R)m.refractionIndex;
            if (dot(n, wo) < R(0)) return f3<R>(m.specular) * (R(1) - fresnel(eta, cosThetaI));
            return f3<R>(m.specular) * fresnel(eta, cosThetaI);
        }
        case RTX_MAT_CONDUCTOR: {
            R F = fresnelConductor(dot(wi, n), (R)m.eta, (R)m.k);
            return V3<R>(F, F, F);
        }
        case RTX_MAT_SPECULAR:
            return f3<R>(m.specular);
        default:
            return V3<R>(0, 0, 0);
        }
    }
    static bool hasDiscreteDirection(const rtx_material &m) { // ref: raytrace.rgen:733-742
        return m.type == RTX_MAT_DIELECTRIC || m.type == RTX_MAT_SPECULAR || m.type == RTX_MAT_CONDUCTOR;
    }
    static R powerHeuristic(R p1, R p2) { // ref: raytrace.rgen:341-344
        R s = p1 * p1;
        return s / (s + p2 * p2);
    }
    static R balanceHeuristic(R p1, R p2) { return p1 / (p1 + p2); } // ref: raytrace.rgen:346-348
    R heuristic(R a, R b) const { return pc.usePowerHeuristic ? powerHeuristic(a, b) : balanceHeuristic(a, b); }

    // ------------------------------------------------------ traversal + hit shaders
    HitInfo<R> trace(V3<R> o, V3<R> d, R tmin, R tmax) {
        HitInfo<R> info;
        HitRec h;
        if (!closest<R>(S, o, d, tmin, tmax, false, h, randomUInt)) { // ref: raytrace.rmiss:18-31
            info.normal = -d;
            info.t = tmax;
            info.isMiss = true;
            info.missColor = envColor<R>(S, normalize(d)); // black without an environment map
            info.matIndex = -1;
            info.prim = -1;
            return info;
        }
        info.isMiss = false;
        info.prim = h.prim;
        info.t = (R)h.t;
        if (h.prim >= (int)S.ntri) { // ref: raytrace.sphere.rchit:13-35
            int si = h.prim - (int)S.ntri;
            const rtx_sphere &s = S.spheres[si];
            V3<R> pos = o + info.t * d;
            V3<R> normal = normalize(pos - f3<R>(s.center));
            if (dot(d, normal) < R(0)) {
                info.isFrontFace = true;
                info.normal = normal;
            } else {
                info.isFrontFace = false;
                info.normal = -normal;
            }
            info.worldPos = pos;
            info.u = info.v = 0;
            info.matIndex = s.materialIndex;
            info.isSphere = true;
            info.instanceIndex = si;
            return info;
        }
        // ref: raytrace.rchit:16-58
        int inst = S.triInst[h.prim], local = S.triLocal[h.prim];
        const rtx_instance &in = S.inst[inst];
        const auto &vv = S.verts[in.modelIndex];
        const auto &ii = S.idx[in.modelIndex];
        const rtx_vertex &v0 = vv[ii[3 * local + 0]], &v1 = vv[ii[3 * local + 1]], &v2 = vv[ii[3 * local + 2]];
        R b1 = (R)h.u, b2 = (R)h.v, b0 = R(1) - b1 - b2;
        V3<R> normal = f3<R>(v0.normal) * b0 + f3<R>(v1.normal) * b1 + f3<R>(v2.normal) * b2;
        normal = normalize(xform<R>(in.normalTransform, normal, R(0)));
        V3<R> wp = f3<R>(v0.pos) * b0 + f3<R>(v1.pos) * b1 + f3<R>(v2.pos) * b2;
        wp = xform<R>(in.transform, wp, R(1));
        R tu = (R)v0.texCoord[0] * b0 + (R)v1.texCoord[0] * b1 + (R)v2.texCoord[0] * b2;
        R tv = (R)v0.texCoord[1] * b0 + (R)v1.texCoord[1] * b1 + (R)v2.texCoord[1] * b2;
        if (dot(d, normal) < R(0)) {
            info.isFrontFace = true;
            info.normal = normal;
        } else {
            info.isFrontFace = false;
            info.normal = -normal;
        }
        info.worldPos = wp;
        info.u = tu;
        info.v = tv;
        info.matIndex = v0.materialIndex;
        info.isSphere = false;
        info.instanceIndex = inst;
        return info;
    }

    // ref: shaders/raytrace.rgen:358-473
    R sampleLights(V3<R> origin, V3<R> normal, V3<R> &lightDir, V3<R> &lightColor, R &lightDistance) {
        int iRandomLight = randomInteger(RTX_SIZE_LIGHT_RANDOM - 1);
        int iLight = S.lightIndex[iRandomLight];
        const rtx_light &light = S.lights[iLight];
        if (light.type == RTX_LIGHT_POINT) {
            V3<R> toLight = f3<R>(light.pos) - origin;
            lightDistance = length(toLight);
            lightDir = toLight / lightDistance;
            lightColor = f3<R>(light.color) / (lightDistance * lightDistance);
            return (R)light.sampleProb;
        } else if (light.type == RTX_LIGHT_SPHERE) {
            const rtx_sphere &s = S.spheres[light.instanceIndex];
            lightColor = f3<R>(S.mats[s.materialIndex].lightColor);
            V3<R> sphereNormal;
            R area = (R)light.area;
            V3<R> position = randomOnSphere(s, sphereNormal);
            V3<R> toLight = position - origin;
            lightDistance = length(toLight);
            lightDir = normalize(toLight);
            R cosThetaLight = dot(-lightDir, sphereNormal);
            return (R)light.sampleProb * lightDistance * lightDistance / (cosThetaLight * area);
        } else if (light.type == RTX_LIGHT_ENV_MAP) { // uniform hemisphere, pdf 1 / 2pi
            lightDir = randomOnUnitSphere();
            if (dot(normal, lightDir) < R(0)) lightDir = reflect(lightDir, normal);
            lightColor = envColor<R>(S, lightDir);
            lightDistance = C::TMAX;
            return (R)light.sampleProb / (R(2) * C::PI);
        }
        int iModel = S.inst[light.instanceIndex].modelIndex;
        int iRandomTri = randomInteger(RTX_SIZE_TRI_RANDOM - 1);
        const rtx_face_sample &ts = S.triTable[(size_t)iLight * RTX_SIZE_TRI_RANDOM + iRandomTri];
        int iTri = ts.index;
        const auto &vv = S.verts[iModel];
        const auto &ii = S.idx[iModel];
        const rtx_vertex &v0 = vv[ii[3 * iTri + 0]], &v1 = vv[ii[3 * iTri + 1]], &v2 = vv[ii[3 * iTri + 2]];
        R rx = rnd();
        R ry = rnd();
        R sqrtx = std::sqrt(rx);
        R b0 = R(1) - sqrtx, b1 = sqrtx * (R(1) - ry), b2 = ry * sqrtx;
        V3<R> pos = f3<R>(v0.pos) * b0 + f3<R>(v1.pos) * b1 + f3<R>(v2.pos) * b2;
        V3<R> nrm = f3<R>(v0.normal) * b0 + f3<R>(v1.normal) * b1 + f3<R>(v2.normal) * b2;
        const rtx_instance &in = S.inst[light.instanceIndex];
        pos = xform<R>(in.transform, pos, R(1));
        nrm = normalize(xform<R>(in.normalTransform, nrm, R(0)));
        V3<R> toLight = pos - origin;
        lightDistance = length(toLight);
        lightDir = toLight / lightDistance;
        lightColor = f3<R>(S.mats[v0.materialIndex].lightColor);
        R cosThetaLight = dot(-lightDir, nrm);
        if (cosThetaLight < R(0)) cosThetaLight = -cosThetaLight;
        return (R)light.sampleProb * lightDistance * lightDistance / cosThetaLight / (R)light.area;
    }

    static bool neeCapable(const rtx_material &m) { // ref: raytrace.rgen:602-611
        return m.type == RTX_MAT_DIFFUSE || m.type == RTX_MAT_PHONG || m.type == RTX_MAT_LIGHT ||
               m.type == RTX_MAT_ROUGH_CONDUCTOR;
    }

    // ref: shaders/raytrace.rgen:601-731
    V3<R> nextEventEstimation(const rtx_material &m, V3<R> origin, V3<R> wi, V3<R> normal, bool frontFace) {
        if (!neeCapable(m) || S.lights.empty()) return V3<R>(0, 0, 0);
        V3<R> nee(0, 0, 0);
        V3<R> lightDir, lightColor;
        R lightDistance;
        R pdfLights = sampleLights(origin, normal, lightDir, lightColor, lightDistance);
        bool isShadowed = true;
        R cosThetaLight = dot(normal, lightDir);
        if (cosThetaLight > R(0) && pdfLights > R(0)) {
            isShadowed =
                occluded<R>(S, origin, lightDir, C::TMIN, lightDistance * (R(1) - R(0.0001)), false, randomUInt);
            // the outcome of every traced shadow ray joins the path signature
            primHash = (primHash ^ (isShadowed ? kSigOccluded : kSigVisible)) * 16777619u;
        }
        if (!isShadowed) {
            if (pc.enableMIS) {
                R pdfMat = pdfBSDF(m, normal, wi, lightDir);
                R h = heuristic(pdfLights, pdfMat);
                if (std::isnan(h)) return nee;
                nee = nee + evalBsdf(m, normal, wi, lightDir, true) * lightColor * h / pdfLights;
            } else {
                nee = evalBsdf(m, normal, wi, lightDir, true) * lightColor / pdfLights;
            }
        }
        if (pc.enableMIS) {
            V3<R> bsdfDir;
            R pdfMat = sampleBSDF(m, wi, normal, frontFace, bsdfDir);
            if (pdfMat > R(0)) {
                HitInfo<R> pinfo = trace(origin, bsdfDir, C::TMIN, C::TMAX);
                // the signature records the emitter the probe sees, or -1 (a miss or
                // another surface first)
                const bool seesLight = !pinfo.isMiss && S.mats[pinfo.matIndex].type == RTX_MAT_LIGHT;
                primHash = (primHash ^ (kSigProbe ^ (uint32_t)(seesLight ? pinfo.prim : -1))) * 16777619u;
                if (pinfo.isMiss) {
                    lightColor = pinfo.missColor;
                    pdfLights = C::INV_2PI / (R)S.lights.size();
                    R h = heuristic(pdfMat, pdfLights);
                    nee = nee + evalBsdf(m, normal, wi, bsdfDir, true) * lightColor * h / pdfMat;
                } else {
                    const rtx_material &ms = S.mats[pinfo.matIndex];
                    if (ms.type == RTX_MAT_LIGHT) {
                        int iLight = pinfo.isSphere ? S.spheres[pinfo.instanceIndex].iLight
                                                    : S.inst[pinfo.instanceIndex].iLight;
                        const rtx_light &light = S.lights[iLight];
                        lightColor = f3<R>(ms.lightColor);
                        // pdfLight, ref: raytrace.rgen:334-339
                        R cosL = dot(-bsdfDir, pinfo.normal);
                        pdfLights = (R)light.sampleProb * pinfo.t * pinfo.t / cosL / (R)light.area;
                        R h = heuristic(pdfMat, pdfLights);
                        if (std::isnan(h)) return nee;
                        nee = nee + evalBsdf(m, normal, wi, bsdfDir, true) * lightColor * h / pdfMat;
                    }
                }
            }
        }
        return nee;
    }

    // ref: shaders/raytrace.rgen:871-877
    V3<R> multipleNEE(const rtx_material &m, V3<R> origin, V3<R> wi, V3<R> normal, bool ff, int numNEE) {
        V3<R> r(0, 0, 0);
        for (int i = 0; i < numNEE; ++i) r = r + nextEventEstimation(m, origin, wi, normal, ff);
        return r / (R)numNEE;
    }

    // ------------------------------------------------------------ guiding
    // getGuidingRegion (raytrace.rgen:904-921, raytrace.guiding.rint:6-10):
    // the region box holding x (bounds inclusive); of several (a point on a
    // shared face) the smallest index.  -1 outside every region.
    int guidingRegion(V3<R> x) const {
        for (size_t i = 0; i < S.regions.size(); ++i) {
            const rtx_aabb &b = S.regions[i];
            if ((R)b.min[0] <= x.x && x.x <= (R)b.max[0] && (R)b.min[1] <= x.y && x.y <= (R)b.max[1] &&
                (R)b.min[2] <= x.z && x.z <= (R)b.max[2])
                return (int)i;
        }
        return -1;
    }
    // vMF (guiding.glsl:34-47) with parallax compensation
    static R vmf(V3<R> wo, const rtx_vmf_theta &t, V3<R> x, bool parallax) {
        if (t.k == 0.0f) return R(0.07957747155); // M_INV_4PI
        V3<R> mu = f3<R>(t.mu);
        if (parallax && t.distance > 0.0f) mu = normalize(f3<R>(t.target) - x);
        return (R)t.norm * std::exp((R)t.k * (dot(mu, wo) - R(1)));
    }
    static R vmm(V3<R> wo, const rtx_vmm_theta &m, V3<R> x, bool parallax) { // guiding.glsl:57-65
        R res = 0;
        for (int i = 0; i < m.usedDistributions; ++i) res += (R)m.pi[i] * vmf(wo, m.thetas[i], x, parallax);
        return res;
    }
    V3<R> sampleVMF(const rtx_vmf_theta &t, V3<R> x, bool parallax) { // guiding.glsl:67-89
        if (t.k > 0.0f) {
            R r1 = rnd();
            R r2 = rnd();
            R cosTheta = R(1) + std::log(R(1) + (R)t.eMin2K * r1 - r1) / (R)t.k;
            R sinTheta = R(1) - cosTheta * cosTheta <= R(0) ? R(0) : std::sqrt(R(1) - cosTheta * cosTheta);
            R phi = R(2) * C::PI * r2;
            V3<R> mu = f3<R>(t.mu);
            if (parallax && t.distance > 0.0f) mu = normalize(f3<R>(t.target) - x);
            return toWorld(V3<R>(sinTheta * std::cos(phi), sinTheta * std::sin(phi), cosTheta), mu);
        }
        return randomOnUnitSphere();
    }
    V3<R> sampleVMM(const rtx_vmm_theta &m, V3<R> x, bool parallax) { // guiding.glsl:91-104
        R u = rnd();
        int i = 0;
        const int last = m.usedDistributions - 1;
        R piSum = (R)m.pi[0];
        while (piSum < u && i < last) {
            i++;
            piSum += (R)m.pi[i];
        }
        return sampleVMF(m.thetas[i], x, parallax);
    }
    // getNewDirection (raytrace.rgen:923-960): with guiding on, non-discrete
    // materials sample the region's mixture with probability guidingProb
    // and the BSDF otherwise; the pdf is the mix of both.
    R getNewDirection(const rtx_material &m, V3<R> x, V3<R> normal, V3<R> wi, bool frontFace, V3<R> &newDir) {
        if (!(pc.useGuiding && !hasDiscreteDirection(m))) return sampleBSDF(m, wi, normal, frontFace, newDir);
        int region = guidingRegion(x);
        if (region < 0) return R(0);
        const rtx_vmm_theta &vm = S.vmms[region];
        const bool parallax = pc.useParallaxCompensation != 0;
        R pdfMat;
        if (rnd() < (R)pc.guidingProb) {
            newDir = sampleVMM(vm, x, parallax);
            pdfMat = pdfBSDF(m, normal, wi, newDir);
        } else {
            pdfMat = sampleBSDF(m, wi, normal, frontFace, newDir);
        }
        if (dot(newDir, normal) < R(0) || pdfMat <= R(0)) return R(0);
        R pdfGuiding = vmm(newDir, vm, x, parallax);
        if (std::isnan(pdfGuiding)) return sampleBSDF(m, wi, normal, frontFace, newDir);
        return pdfMat * (R(1) - (R)pc.guidingProb) + pdfGuiding * (R)pc.guidingProb; // mix()
    }

    // ------------------------------------------------- guiding sample record
    static bool isMatAlmostDiscrete(const rtx_material &m) { // raytrace.rgen:962-964
        return (m.type == RTX_MAT_ROUGH_CONDUCTOR && m.roughness <= 0.3f) ||
               (m.type == RTX_MAT_PHONG && m.specularHighlight >= 250.0f);
    }
    void updateSamples(V3<R> light) { // raytrace.rgen:973-978
        for (int i = curSlot - 1; i >= sampleOffset; --i) {
            lightSums[i] = lightSums[i] + light;
            light = light * sampleThroughputs[i];
        }
    }
    void commitSamples() { // raytrace.rgen:980-990
        for (int i = sampleOffset; i < curSlot; ++i) {
            R w = length(lightSums[i]) / (R)dd[i].pdf;
            if (w <= R(0)) dd[i].flags = kDirInvalid;
            else dd[i].weight = (float)w;
        }
    }
    void addDistance(R t) { dd[iUpdateDistance].distance = (float)((R)dd[iUpdateDistance].distance + t * distanceFactor); }

    // ---------------------------------------------------- irradiance cache
    // ref: raytrace.rgen:806-813
    bool isICCapable(const rtx_material &m) const {
        if (m.type == RTX_MAT_DIFFUSE || m.type == RTX_MAT_LIGHT) return true;
        return pc.useIrradianceCacheOnGlossy && !hasDiscreteDirection(m);
    }
    // ref: raytrace.rgen:815-828
    V3<R> approxDiffuse(const rtx_material &m, V3<R> normal, V3<R> wi) const {
        switch (m.type) {
        case RTX_MAT_DIFFUSE:
        case RTX_MAT_LIGHT:
        case RTX_MAT_PHONG: return diffuse(m);
        case RTX_MAT_ROUGH_CONDUCTOR: {
            R f = fresnelConductor(dot(wi, normal), (R)m.eta, (R)m.k);
            return V3<R>(f, f, f) * C::INV_PI;
        }
        }
        return V3<R>(-1, -1, -1);
    }
    // queryIrradianceCache (raytrace.rgen:748-773): every cache sphere that
    // contains x (raytrace.irradiance.rint:11-21) goes through the any-hit
    // weighting of raytrace.irradiance.rahit:18-51, here in ascending entry order.
    bool queryIC(V3<R> x, V3<R> n, V3<R> &color) const {
        V3<R> sum(0, 0, 0);
        R total = 0;
        const uint32_t count = ic->count();
        for (uint32_t i = 0; i < count; ++i) {
            const rtx_sphere &sp = ic->sph[i];
            const rtx_cache_data &e = ic->e[i];
            V3<R> c = f3<R>(sp.center);
            if (!(length(x - c) <= (R)sp.radius)) continue;
            V3<R> ni = f3<R>(e.normal);
            R w = R(1) / (length(x - c) / (R)e.harmonicR + std::sqrt(R(1) - dot(n, ni)));
            if (std::isnan(w) || std::isinf(w)) w = R(1000000);
            if (w <= R(1) / (R)pc.irradianceA) continue;
            if (pc.irradianceCachePerformVisibilityCheck && !(R(-0.001) <= dot(x - c, (n + ni) / R(2)))) continue;
            if (pc.useIrradianceGradients) {
                V3<R> col = f3<R>(e.color);
                R E = length(col);
                V3<R> cn = E != R(0) ? normalize(col) : V3<R>(0, 0, 0);
                V3<R> adjusted = cn * (E + dot(cross(ni, n), f3<R>(e.rotGrad)) + dot(x - c, f3<R>(e.transGrad)));
                sum = sum + w * adjusted;
            } else {
                sum = sum + w * f3<R>(e.color);
            }
            total += w;
        }
        if (total > R(0)) {
            color = sum / total;
            return true;
        }
        return false;
    }

    // ------------------------------------------------------------- ADRRS
    // applyWeightWindow (raytrace.rgen:830-869): the window's centre is
    // |estimate / adjoint| (component-wise quotient), lower = 2c/(1+s),
    // upper = s*lower; below it roulette with survival max(v/lower, 0.1),
    // above it split into q = v/upper copies rounded up with probability
    // q - floor(q).  A float-to-int conversion past INT_MAX saturates, as
    // the GPU conversion instructions do.
    R weightWindow(V3<R> throughput, V3<R> adjoint, int &n) {
        R center = length(estimate / adjoint);
        R lower = R(2) * center / (R(1) + (R)pc.adrrsS);
        R upper = (R)pc.adrrsS * lower;
        R v = length(throughput);
        if (std::isnan(lower) || lower <= R(0)) {
            n = 1;
            return R(1);
        }
        if (lower <= v && v <= upper) {
            n = 1;
            return R(1);
        } else if (v <= lower) {
            n = 1;
            return std::max(v / lower, R(0.1));
        }
        R q = v / upper;
        n = q >= R(2147483520.0) ? 2147483646 : (int)q;
        if (rnd() > (R)(n + 1) - q) n++;
        return q;
    }
    // split (raytrace.rgen:883-902): false when MAX_SPLITS records are queued
    bool split(V3<R> origin, V3<R> normal, V3<R> wi, V3<R> throughput, int matIndex, int depth, bool frontFace) {
        if (nSplits >= kMaxSplits) return false;
        splits[nSplits++] = SplitRec{origin, normal, wi, throughput, matIndex, depth, frontFace, texU, texV};
        g_adrrsCount[0]++;
        return true;
    }

    // The bounce loop, ref: shaders/raytrace.rgen:992-1226: the plain method,
    // the irradiance-cache branch and ADRRS (1080-1146), guiding samples and
    // records.  currentDepth is the depth the path starts from (the split
    // loop of main() resumes paths at the depth of their split vertex) and
    // returns the depth reached.  Russian roulette (absent from the
    // reference) is applied to camera paths when pc.enableRR is set.
    V3<R> raytrace(V3<R> origin, V3<R> direction, V3<R> throughput, int &currentDepth, int maxDepth,
                   int maxFollowDiscrete, bool addDirectLights, bool addFirstHitLight, bool useNEE, int numNEE,
                   bool useIC, bool createIC, bool useADRRS, bool allowRR, R &firstT, bool saveSamples = false) {
        bool follow = true;
        int followCount = 0;
        bool addNextDirectLights = addFirstHitLight;
        int depth = currentDepth;
        firstT = C::TMAX;
        curSlot = sampleOffset;
        iUpdateDistance = -1;
        distanceFactor = R(1);
        V3<R> result(0, 0, 0);
        do {
            depth++;
            HitInfo<R> info = trace(origin, direction, C::TMIN, C::TMAX);
            primHash = (primHash ^ (uint32_t)info.prim) * 16777619u;
            if (info.isMiss) {
                if (addDirectLights || addNextDirectLights) {
                    result = result + throughput * info.missColor;
                    if (saveSamples) {
                        updateSamples(info.missColor);
                        if (iUpdateDistance != -1) dd[iUpdateDistance].distance = 0.0f; // infinitely far
                    }
                }
                break;
            }
            const rtx_material &m = S.mats[info.matIndex];
            origin = info.worldPos;
            V3<R> normal = info.normal;
            texU = info.u;
            texV = info.v;
            if (depth == 1) firstT = info.t;
            if (m.type == RTX_MAT_LIGHT && (addDirectLights || addNextDirectLights)) {
                result = result + throughput * f3<R>(m.lightColor);
                if (saveSamples) updateSamples(f3<R>(m.lightColor));
            }
            if (hasDiscreteDirection(m)) {
                addNextDirectLights = true;
                follow = true;
                if (depth >= maxDepth) followCount++;
                if (saveSamples && iUpdateDistance != -1) addDistance(info.t);
            } else {
                addNextDirectLights = false;
                follow = false;
                if ((useIC || (useADRRS && depth > 1)) && isICCapable(m)) {
                    V3<R> irradiance;
                    if (queryIC(origin, normal, irradiance)) {
                        if (useADRRS) { // raytrace.rgen:1087-1116
                            int n;
                            R q = weightWindow(throughput, irradiance, n);
                            if (n == 1) {
                                if (rnd() > q) { // the path dies
                                    primHash = (primHash ^ kSigRRKill) * 16777619u;
                                    g_adrrsCount[1]++;
                                    break;
                                }
                                throughput = throughput / q;
                            } else if (pc.adrrsSplit) {
                                int possible = kMaxSplits - nSplits;
                                if (n - 1 > possible) {
                                    n = possible + 1;
                                    q = (R)n;
                                }
                                throughput = throughput / q;
                                // n copies: this path and n - 1 queued splits
                                for (int k = 0; k < n - 1; ++k)
                                    split(origin, normal, -direction, throughput, info.matIndex, depth, info.isFrontFace);
                                primHash = (primHash ^ (kSigSplit ^ (uint32_t)(n - 1))) * 16777619u;
                            }
                        } else {
                            // cache record found: the path ends here (raytrace.rgen:1114-1122)
                            primHash = (primHash ^ kSigCacheHit) * 16777619u;
                            result = result + throughput * approxDiffuse(m, normal, -direction) * irradiance;
                            result = result + throughput * multipleNEE(m, origin, -direction, normal, info.isFrontFace,
                                                                       numNEE);
                            break;
                        }
                    } else if (createIC && rnd() < (R)pc.irradianceCreateProb) {
                        if (numCand < kMaxNewCacheEntries) { // raytrace.rgen:1125-1137
                            candO[numCand] = origin;
                            candN[numCand] = normal;
                            numCand++;
                        }
                    }
                }
                if (pc.splitOnFirst && depth == 1) { // raytrace.rgen:1142-1146
                    if (split(origin, normal, -direction, throughput * R(0.5), info.matIndex, depth,
                              info.isFrontFace)) {
                        throughput = throughput * R(0.5);
                        primHash = (primHash ^ (kSigSplit ^ 1u)) * 16777619u;
                    }
                }
                if (saveSamples && iUpdateDistance != -1) {
                    addDistance(info.t);
                    if (!isMatAlmostDiscrete(m)) iUpdateDistance = -1; // the record's distance is final
                }
                if (useNEE) {
                    V3<R> nee = multipleNEE(m, origin, -direction, normal, info.isFrontFace, numNEE);
                    result = result + throughput * nee;
                    if (saveSamples) updateSamples(nee);
                }
            }
            V3<R> newDirection;
            R pdf = getNewDirection(m, origin, normal, -direction, info.isFrontFace, newDirection);
            // The reference stops on pdf <= 0 only; a NaN pdf (pdfBSDF of wo == -wi
            // normalises a zero half-vector) would poison the pixel, so it ends the path.
            if (pdf <= R(0) || std::isnan(pdf)) break;
            V3<R> change = evalBsdf(m, normal, -direction, newDirection, info.isFrontFace) / pdf;
            if (std::isnan(change.x) || std::isnan(change.y) || std::isnan(change.z)) break;
            throughput = throughput * change;
            if (saveSamples && curSlot < RTX_MAX_DIRECTIONAL_DATA_PER_PIXEL) {
                rtx_directional_data &d = dd[curSlot];
                if (hasDiscreteDirection(m) || isMatAlmostDiscrete(m)) {
                    d.flags = kDirInvalid; // probabilities, not densities: unusable for fitting
                } else {
                    store3(d.position, origin);
                    store3(d.direction, newDirection);
                    d.pdf = (float)pdf;
                    d.weight = 0.0f;
                    d.distance = 0.0f;
                    d.flags = (uint32_t)guidingRegion(origin);
                    iUpdateDistance = curSlot;
                    distanceFactor = R(1);
                }
                lightSums[curSlot] = V3<R>(0, 0, 0);
                sampleThroughputs[curSlot] = change;
                curSlot++;
            } else if (saveSamples && iUpdateDistance != -1 && m.type == RTX_MAT_DIELECTRIC &&
                       dot(newDirection, normal) < R(0)) {
                // refraction: the next distance is the apparent one (raytrace.rgen:1200-1210)
                R eta = info.isFrontFace ? (R)m.refractionIndex : (R)m.refractionIndexInv;
                distanceFactor = std::fabs(dot(normal, direction) / dot(normal, newDirection)) * eta;
            }
            direction = newDirection;
            if (allowRR && pc.enableRR && depth >= fp.rrStartDepth) {
                R q = std::max(std::fabs(throughput.x), std::max(std::fabs(throughput.y), std::fabs(throughput.z)));
                q = std::min(q, R(0.95));
                if (rnd() >= q) break;
                throughput = throughput / q;
            }
        } while (depth <= maxDepth || (follow && followCount <= maxFollowDiscrete));
        if (saveSamples) {
            commitSamples();
            // the slots of a path that found no light are reused by the next one;
            // otherwise the next starts at sampleOffset + currentSampleOffset (1220-1222)
            if (length(result) > R(0)) sampleOffset += curSlot;
        }
        currentDepth = depth;
        return result;
    }

    // getCameraRay, ref: shaders/raytrace.rgen:1487-1494
    void cameraRay(uint32_t px, uint32_t py, uint32_t w, uint32_t h, V3<R> &o, V3<R> &d) {
        const rtx_camera &cam = fp.cam;
        R jx = negPos();
        R jy = negPos();
        R ux = ((R)px + R(0.5) + jx / R(2)) / (R)w;
        R uy = ((R)py + R(0.5) + jy / R(2)) / (R)h;
        R dx = ux * R(2) - R(1), dy = uy * R(2) - R(1);
        o = xform<R>(cam.viewInverse, V3<R>(0, 0, 0), R(1));
        // target = projInverse * (d.x, d.y, 1, 1); only xyz is used
        V3<R> target = xform<R>(cam.projInverse, V3<R>(dx, dy, R(1)), R(1));
        d = normalize(xform<R>(cam.viewInverse, normalize(target), R(0)));
    }

    // Camera samples of one frame (samplesPerPixel, 1 in prepare frames;
    // raytrace.rgen:1662).
    int samplesThisFrame() const {
        if (pc.useIrradianceCache && pc.isIrradiancePrepareFrame) return 1;
        return pc.samplesPerPixel > 1 ? pc.samplesPerPixel : 1;
    }

    // main(), ref: shaders/raytrace.rgen:1623-1711 for one pixel and frame:
    // the estimate, the cache-update election (the stream's first draw), the
    // camera samples, then the queued splits (each: NEE at its vertex, a new
    // direction, a path resumed at its depth; a split whose direction has
    // pdf <= 0 ends the split loop), the mean over the camera samples.
    // Sample `pathId / (width*height)` of the call is one frame.
    V3<R> renderPath(uint64_t pathId) {
        const uint64_t npix = (uint64_t)fp.width * fp.height;
        uint32_t pixel = (uint32_t)(pathId % npix);
        uint64_t sample = pathId / npix + fp.sampleOffset;
        uint32_t frameSeed = pcg32First(fp.seed, sample);
        seed = tea(pixel, frameSeed);
        randomUInt = frameSeed;
        primHash = 2166136261u;
        const bool useIC = pc.useIrradianceCache != 0;
        const bool useADRRS = pc.useADRRS != 0;
        if (useADRRS && !S.estimate.empty()) estimate = f3<R>(&S.estimate[4 * (size_t)pixel]);
        // raytrace.rgen:1637-1639: the update election is the stream's first draw
        if (useIC && rnd() < (R)pc.irradianceUpdateProb) icElect = true;
        const bool save = pc.updateGuiding != 0 && dd != nullptr;
        // recording frames render nothing when splitting is on (raytrace.rgen:1644-1648)
        if (pc.updateGuiding && ((useADRRS && pc.adrrsSplit) || pc.splitOnFirst)) return V3<R>(0, 0, 0);
        const bool useNEE = pc.enableNEE != 0;
        const int spp = samplesThisFrame();
        // createIC is true for every camera path (raytrace.rgen:1676); candidates
        // only arise where the cache is looked up
        const bool createIC = useIC || useADRRS;
        V3<R> result(0, 0, 0);
        R firstT = C::TMAX;
        for (int k = 0; k < spp; ++k) {
            V3<R> o, d;
            cameraRay(pixel % fp.width, pixel / fp.width, fp.width, fp.height, o, d);
            int depth = 0;
            result = result + raytrace(o, d, V3<R>(1, 1, 1), depth, pc.maxDepth, pc.maxFollowDiscrete, !useNEE, true,
                                       useNEE, pc.numNEE, useIC, createIC, useADRRS, true, firstT, save);
        }
        for (int k = 0; k < nSplits; ++k) { // raytrace.rgen:1684-1708
            const SplitRec sp = splits[k];
            const rtx_material &m = S.mats[sp.matIndex];
            texU = sp.u;
            texV = sp.v;
            if (useNEE) result = result + sp.throughput * multipleNEE(m, sp.origin, sp.wi, sp.normal, sp.frontFace,
                                                                      pc.numNEE);
            V3<R> dir;
            R pdf = getNewDirection(m, sp.origin, sp.normal, sp.wi, sp.frontFace, dir);
            // a NaN pdf or BSDF ratio is treated as pdf <= 0, as in the bounce loop
            if (pdf <= R(0) || std::isnan(pdf)) break;
            V3<R> thr = sp.throughput * evalBsdf(m, sp.normal, sp.wi, dir, sp.frontFace) / pdf;
            if (std::isnan(thr.x) || std::isnan(thr.y) || std::isnan(thr.z)) break;
            int depth = sp.depth;
            g_adrrsCount[2]++;
            result = result + raytrace(sp.origin, dir, thr, depth, pc.maxDepth, pc.maxFollowDiscrete, !useNEE, false,
                                       useNEE, pc.numNEE, useIC, true, useADRRS, true, firstT, save);
        }
        lastFirstT = firstT;
        return result / (R)spp;
    }
};

int checkScene(const orc_scene &s) {
    for (const auto &l : s.lights)
        if (l.type == RTX_LIGHT_ENV_MAP && s.env.empty()) {
            g_err = "environment-map light without an environment map";
            return RTX_E_INVALID;
        }
    for (const auto &m : s.mats)
        for (int32_t tid : {m.textureIdDiffuse, m.textureIdSpecular})
            if (tid != -1 && (tid < 1 || (size_t)tid >= s.textures.size() || s.textures[tid].w == 0)) {
                g_err = "material texture id " + std::to_string(tid) + " names no texture";
                return RTX_E_INVALID;
            }
    return RTX_OK;
}

int checkParams(const orc_scene &s, const rtx_frame_params *fp, bool cacheFrame = false) {
    const rtx_push_constant &pc = fp->pc;
    if (fp->width == 0 || fp->height == 0) {
        g_err = "empty image";
        return RTX_E_INVALID;
    }
    if ((pc.useIrradianceCache || pc.useADRRS) && !cacheFrame) {
        g_err = "oracle: irradiance-cache and ADRRS frames go through orc_ic_frame";
        return RTX_E_UNSUPPORTED;
    }
    if (pc.useIrradianceCache && (pc.numNEE < 1 || pc.irradianceNumNEE < 1)) {
        g_err = "numNEE and irradianceNumNEE must be >= 1 with the irradiance cache";
        return RTX_E_INVALID;
    }
    if (pc.visualizeMode != 0 || pc.showIrradianceCacheOnly) {
        g_err = "oracle: visualisation modes are not restated";
        return RTX_E_UNSUPPORTED;
    }
    if (pc.splitOnFirst && (pc.useIrradianceCache || pc.useADRRS)) {
        // the cache's sub-paths would queue splits on the pixel's list
        g_err = "splitOnFirst together with the irradiance cache or ADRRS is not supported";
        return RTX_E_UNSUPPORTED;
    }
    if (pc.useADRRS) {
        if (s.estimate.empty() || s.estW != fp->width || s.estH != fp->height) {
            g_err = "useADRRS needs an estimate of the frame's size (orc_estimate_set)";
            return RTX_E_STATE;
        }
        if (!(pc.adrrsS >= 1.0f)) { // the reference's range is [1, 5]
            g_err = "adrrsS must be >= 1";
            return RTX_E_INVALID;
        }
    }
    if (pc.samplesPerPixel > 4096) {
        g_err = "samplesPerPixel must be at most 4096";
        return RTX_E_INVALID;
    }
    if (pc.enableNEE && pc.numNEE < 1) {
        g_err = "numNEE must be >= 1";
        return RTX_E_INVALID;
    }
    return RTX_OK;
}

template <class F> void parallelFor(size_t n, int threads, F fn) {
    if (threads <= 0) threads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    std::atomic<size_t> next{0};
    const size_t chunk = 64;
    auto worker = [&]() {
        for (;;) {
            size_t b = next.fetch_add(chunk);
            if (b >= n) break;
            size_t e = std::min(n, b + chunk);
            for (size_t i = b; i < e; ++i) fn(i);
        }
    };
    if (threads == 1) {
        worker();
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(worker);
    for (auto &t : pool) t.join();
}

template <class R> V3<R> vd(const double *p) { return V3<R>((R)p[0], (R)p[1], (R)p[2]); }

// ------------------------------------------------- irradiance-cache frame
// One frame (one 1-spp sample plane) of main() with useIrradianceCache
// (raytrace.rgen:1623-1827): every pixel's path reads the cache as it was at
// the start of the frame; afterwards the frame's update elections (in pixel
// order) and creation candidates (in pixel, then path order) become jobs of
// calculateCacheData, whose results are committed in job order.  The
// reference does the same work with atomics from every pixel at once, so
// its order (and which candidates win the last free slots) is a race; this
// order is the deterministic restatement of it.
struct IcJobRec {
    int type; // 0 = update (updateIrradianceCache, raytrace.rgen:1334-1381), 1 = create (1383-1421)
    uint32_t target;
    double o[3], n[3];
};

template <class R> struct IcResult {
    V3<R> color, rot, trans;
    R harmonicR;
};

// The stratum (k, j) of job `job`: phi from the column stream, theta from
// the stratum stream, whose state then drives the sub-path.
template <class R> void icStratum(uint32_t frameSeed, uint32_t job, int k, int j, R &phi, R &theta, uint32_t &state) {
    using C = Consts<R>;
    uint32_t ps = tea(job * kIcStrataPhi + (uint32_t)k, frameSeed ^ kIcPhiKey);
    ps = 1664525u * ps + 1013904223u;
    R uphi = R(ps & 0x00FFFFFFu) / R(16777216.0);
    phi = R(2) * C::PI * ((R)k + uphi) / (R)kIcStrataPhi; // raytrace.rgen:1247
    uint32_t st = tea((job * kIcStrataPhi + (uint32_t)k) * kIcStrataTheta + (uint32_t)j, frameSeed ^ kIcStratumKey);
    st = 1664525u * st + 1013904223u;
    R utheta = R(st & 0x00FFFFFFu) / R(16777216.0);
    theta = std::asin(std::sqrt(((R)j + utheta) / (R)kIcStrataTheta)); // raytrace.rgen:1257
    state = st;
}

// One sub-path of calculateCacheData (raytrace.rgen:1255-1263): raytrace with
// maxDepth 1, maxFollowDiscrete 10, no direct light on hits, NEE with
// irradianceNumNEE samples, cache lookups on, no new candidates.
template <class R>
void icSubPath(const orc_scene &S, const rtx_frame_params &fp, const IcSnapshot &snap, uint32_t frameSeed,
               uint32_t job, const IcJobRec &jb, int k, int j, double out[4], uint32_t sig[2]) {
    Tracer<R> t(S, fp);
    t.ic = &snap;
    t.randomUInt = frameSeed;
    R phi, theta;
    icStratum<R>(frameSeed, job, k, j, phi, theta, t.seed);
    t.primHash = 2166136261u;
    V3<R> n = vd<R>(jb.n), o = vd<R>(jb.o);
    V3<R> dir = Tracer<R>::toWorld(Tracer<R>::sphericalToCartesian(theta, phi), n);
    R r;
    int depth = 0;
    V3<R> L = t.raytrace(o, dir, V3<R>(1, 1, 1), depth, 1, 10, false, false, true, fp.pc.irradianceNumNEE, true, false,
                         false, false, r);
    out[0] = (double)L.x;
    out[1] = (double)L.y;
    out[2] = (double)L.z;
    out[3] = (double)r;
    sig[0] = t.seed;
    sig[1] = t.primHash;
}

// The accumulation half of calculateCacheData (raytrace.rgen:1264-1315) over
// the 200 sub-path results (L rgb, first-hit distance r) of one job, in the
// reference's loop order.  Column k-1's values feed the k-gradient as the
// reference's previousKLs/previousKRs intend.
template <class R>
IcResult<R> icReduce(uint32_t frameSeed, uint32_t job, V3<R> normal, const double *sub) {
    using C = Consts<R>;
    const int N = kIcStrataPhi, M = kIcStrataTheta;
    R invDistanceSum = 0;
    int numDistances = 0;
    IcResult<R> res;
    res.rot = V3<R>(0, 0, 0);
    res.trans = V3<R>(0, 0, 0);
    V3<R> color(0, 0, 0);
    R prevKL[kIcStrataTheta] = {}, prevKR[kIcStrataTheta] = {};
    for (int k = 0; k < N; ++k) {
        R phi, theta;
        uint32_t st;
        icStratum<R>(frameSeed, job, k, 0, phi, theta, st);
        V3<R> uk = Tracer<R>::toWorld(Tracer<R>::sphericalToCartesian(C::HALF_PI, phi), normal);
        V3<R> vk = Tracer<R>::toWorld(Tracer<R>::sphericalToCartesian(C::HALF_PI, phi + C::HALF_PI), normal);
        V3<R> prevVk = Tracer<R>::toWorld(
            Tracer<R>::sphericalToCartesian(C::HALF_PI, R(2) * C::PI * (R)k / (R)N + C::HALF_PI), normal);
        R prevJR = 0, prevJL = 0;
        for (int j = 0; j < M; ++j) {
            icStratum<R>(frameSeed, job, k, j, phi, theta, st);
            const double *o = sub + 4 * (size_t)(k * M + j);
            V3<R> sample((R)o[0], (R)o[1], (R)o[2]);
            R r = (R)o[3];
            color = color + sample;
            if (r < C::TMAX) {
                invDistanceSum += R(1) / r;
                numDistances++;
            }
            R L = length(sample);
            R prevJTheta = std::asin(std::sqrt((R)j / (R)M));
            R nextJTheta = std::asin(std::sqrt((R)(j + 1) / (R)M));
            R tanTheta = std::tan(theta);
            if (std::isinf(tanTheta) || std::isnan(tanTheta)) tanTheta = 0;
            res.rot = res.rot - tanTheta * L * vk;
            if (j > 0) {
                R cosPrev = std::cos(prevJTheta);
                res.trans = res.trans + uk * R(2) * C::PI / (R)N * std::sin(prevJTheta) * cosPrev * cosPrev /
                                            std::min(r, prevJR) * (L - prevJL);
            }
            if (k > 0)
                res.trans = res.trans + prevVk * (std::sin(nextJTheta) - std::sin(prevJTheta)) /
                                            std::min(r, prevKR[j]) * (L - prevKL[j]);
            prevKL[j] = L;
            prevKR[j] = r;
            prevJL = L;
            prevJR = r;
        }
    }
    R normFactor = C::PI / (R)(M * N);
    res.color = normFactor * color;
    res.rot = res.rot * normFactor;
    if (invDistanceSum == R(0) || numDistances == 0) res.harmonicR = R(-1);
    else res.harmonicR = R(1) / (invDistanceSum / (R)numDistances);
    return res;
}

template <class R> void clampGradients(const rtx_push_constant &pc, V3<R> &rot, V3<R> &trans) { // raytrace.rgen:1318-1329
    R lr = length(rot);
    if (lr > (R)pc.irradianceGradientsMaxLength) rot = rot * ((R)pc.irradianceGradientsMaxLength / lr);
    R lt = length(trans);
    if (lt > (R)pc.irradianceGradientsMaxLength) trans = trans * ((R)pc.irradianceGradientsMaxLength / lt);
}

template <class R> V3<R> mixv(V3<R> x, V3<R> y, R a) { return x * (R(1) - a) + y * a; }
template <class R> R mixs(R x, R y, R a) { return x * (R(1) - a) + y * a; }

// Commit one job's result into the cache buffers (FP32, as the reference's).
template <class R> void icCommit(IcSnapshot &live, const rtx_push_constant &pc, const IcJobRec &jb, IcResult<R> r) {
    if (jb.type == 0) { // updateIrradianceCache, raytrace.rgen:1349-1380
        rtx_cache_data &e = live.e[jb.target];
        R a = std::min((R)e.numUpdates / (R)(e.numUpdates + 1), R(0.95));
        store3(e.color, mixv(r.color, f3<R>(e.color), a));
        V3<R> rot = mixv(r.rot, f3<R>(e.rotGrad), a), trans = mixv(r.trans, f3<R>(e.transGrad), a);
        clampGradients(pc, rot, trans);
        store3(e.rotGrad, rot);
        store3(e.transGrad, trans);
        R h = r.harmonicR > R(0) ? mixs(r.harmonicR, (R)e.harmonicR, a) : (R)e.harmonicR;
        h = std::max(h, (R)pc.irradianceCacheMinRadius);
        e.harmonicR = (float)h;
        live.sph[jb.target].radius = (float)((R)pc.irradianceA * h);
        e.numUpdates += 1;
        return;
    }
    // createIrradianceCache, raytrace.rgen:1383-1421 (slots stay below maxCaches)
    if (r.harmonicR < R(0)) return;
    uint32_t slot = live.hdr.nextCacheSlot;
    if (slot >= live.hdr.maxCaches) return;
    live.hdr.nextCacheSlot = slot + 1;
    R h = std::max(r.harmonicR, (R)pc.irradianceCacheMinRadius);
    clampGradients(pc, r.rot, r.trans);
    rtx_cache_data &e = live.e[slot];
    for (int a = 0; a < 3; ++a) e.normal[a] = (float)jb.n[a];
    store3(e.color, r.color);
    e.harmonicR = (float)h;
    store3(e.rotGrad, r.rot);
    store3(e.transGrad, r.trans);
    e.numUpdates = 1;
    rtx_sphere &sp = live.sph[slot];
    for (int a = 0; a < 3; ++a) sp.center[a] = (float)jb.o[a];
    sp.radius = (float)((R)pc.irradianceA * h);
}

template <class R>
void icFrame(const orc_scene &S, const rtx_frame_params &fp, IcSnapshot &live, int threads, double *rgb,
             uint32_t *sig, std::vector<IcJobRec> &jobs, std::vector<double> &sub, std::vector<uint32_t> &subSig) {
    const size_t npix = (size_t)fp.width * fp.height;
    const IcSnapshot snap = live;
    const uint32_t frameSeed = pcg32First(fp.seed, fp.sampleOffset);
    std::vector<uint8_t> elect(npix);
    std::vector<int> ncand(npix);
    std::vector<std::array<double, 6 * kMaxNewCacheEntries>> cand(npix);
    parallelFor(npix, threads, [&](size_t p) {
        Tracer<R> t(S, fp);
        t.ic = &snap;
        V3<R> v = t.renderPath(p);
        rgb[3 * p + 0] = (double)v.x;
        rgb[3 * p + 1] = (double)v.y;
        rgb[3 * p + 2] = (double)v.z;
        if (sig) {
            sig[2 * p + 0] = t.seed;
            sig[2 * p + 1] = t.primHash;
        }
        elect[p] = t.icElect ? 1 : 0;
        ncand[p] = t.numCand;
        for (int c = 0; c < t.numCand; ++c) {
            toD(t.candO[c], &cand[p][6 * c]);
            toD(t.candN[c], &cand[p][6 * c + 3]);
        }
    });
    jobs.clear();
    const uint32_t count = live.count();
    uint32_t numUpdates = 0; // at most maxCaches update jobs a frame
    for (size_t p = 0; p < npix && numUpdates < live.hdr.maxCaches; ++p) { // raytrace.rgen:1335-1345
        if (!elect[p]) continue;
        uint32_t idx = live.hdr.nextUpdateSlot++;
        if (idx >= count) {
            live.hdr.nextUpdateSlot = 0;
            continue;
        }
        if (live.e[idx].numUpdates < 1) continue;
        IcJobRec jb{0, idx, {}, {}};
        for (int a = 0; a < 3; ++a) {
            jb.o[a] = live.sph[idx].center[a];
            jb.n[a] = live.e[idx].normal[a];
        }
        jobs.push_back(jb);
        numUpdates++;
    }
    if (count < live.hdr.maxCaches) { // createIrradianceCache's early exit, raytrace.rgen:1384-1386
        uint32_t room = live.hdr.maxCaches - count, made = 0;
        for (size_t p = 0; p < npix && made < room; ++p)
            for (int c = 0; c < ncand[p] && made < room; ++c, ++made) {
                IcJobRec jb{1, 0, {}, {}};
                for (int a = 0; a < 3; ++a) {
                    jb.o[a] = cand[p][6 * c + a];
                    jb.n[a] = cand[p][6 * c + 3 + a];
                }
                jobs.push_back(jb);
            }
    }
    const size_t nsub = jobs.size() * kIcStrata;
    sub.assign(nsub * 4, 0.0);
    subSig.assign(nsub * 2, 0u);
    parallelFor(nsub, threads, [&](size_t i) {
        uint32_t job = (uint32_t)(i / kIcStrata), s = (uint32_t)(i % kIcStrata);
        icSubPath<R>(S, fp, snap, frameSeed, job, jobs[job], (int)(s / kIcStrataTheta), (int)(s % kIcStrataTheta),
                     &sub[4 * i], &subSig[2 * i]);
    });
    for (size_t j = 0; j < jobs.size(); ++j) {
        IcResult<R> r = icReduce<R>(frameSeed, (uint32_t)j, vd<R>(jobs[j].n), &sub[4 * j * kIcStrata]);
        icCommit<R>(live, fp.pc, jobs[j], r);
    }
}

} // namespace

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

int orc_scene_create(const rtx_scene_desc *d, orc_scene **out) {
    if (!d || !out) {
        g_err = "NULL argument";
        return RTX_E_INVALID;
    }
    auto s = std::make_unique<orc_scene>();
    for (uint32_t m = 0; m < d->modelCount; ++m) {
        const rtx_model_desc &md = d->models[m];
        s->verts.emplace_back(md.vertices, md.vertices + md.vertexCount);
        s->idx.emplace_back(md.indices, md.indices + md.indexCount);
    }
    s->inst.assign(d->instances, d->instances + d->instanceCount);
    s->mats.assign(d->materials, d->materials + d->materialCount);
    s->lights.assign(d->lights, d->lights + d->lightCount);
    if (d->randomLightIndex) s->lightIndex.assign(d->randomLightIndex, d->randomLightIndex + RTX_SIZE_LIGHT_RANDOM);
    if (d->randomTriIndex)
        s->triTable.assign(d->randomTriIndex, d->randomTriIndex + (size_t)d->randomTriRows * RTX_SIZE_TRI_RANDOM);
    s->triRows = d->randomTriRows;
    s->spheres.assign(d->spheres, d->spheres + d->sphereCount);
    if (d->envMap && d->envWidth && d->envHeight) {
        s->env.assign(d->envMap, d->envMap + (size_t)d->envWidth * d->envHeight * 4);
        s->envW = d->envWidth;
        s->envH = d->envHeight;
    }
    s->textures.resize(d->textures ? d->textureCount : 0);
    for (uint32_t t = 1; t < s->textures.size(); ++t) {
        const rtx_texture_desc &td = d->textures[t];
        if (!td.rgba) continue;
        s->textures[t].w = td.width;
        s->textures[t].h = td.height;
        s->textures[t].rgba.assign(td.rgba, td.rgba + (size_t)td.width * td.height * 4);
    }
    for (int i = 0; i < 256; ++i) { // IEC 61966-2-1 decoding of the SRGB formats
        double c = i / 255.0;
        s->srgb[i] = (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
    }
    int rc = checkScene(*s);
    if (rc != RTX_OK) return rc;
    // flatten instances to world-space triangles (global primitive order:
    // instances in order, each model's triangles in order; spheres after)
    for (uint32_t i = 0; i < d->instanceCount; ++i) {
        const rtx_instance &in = s->inst[i];
        const auto &vv = s->verts[in.modelIndex];
        const auto &ii = s->idx[in.modelIndex];
        for (size_t f = 0; f * 3 + 2 < ii.size(); ++f) {
            for (int k = 0; k < 3; ++k) {
                const rtx_vertex &v = vv[ii[3 * f + k]];
                V3<double> p = xform<double>(in.transform, f3<double>(v.pos), 1.0);
                V3<float> pf = xform<float>(in.transform, f3<float>(v.pos), 1.0f);
                s->tri.push_back(p.x);
                s->tri.push_back(p.y);
                s->tri.push_back(p.z);
                s->triF.push_back(pf.x);
                s->triF.push_back(pf.y);
                s->triF.push_back(pf.z);
            }
            s->triInst.push_back((int)i);
            s->triLocal.push_back((int)f);
        }
    }
    s->ntri = s->triInst.size();
    s->triAlpha.assign(s->ntri, 0);
    for (size_t p = 0; p < s->ntri; ++p) { // raytrace.rahit:37: the first vertex's material
        const rtx_instance &in = s->inst[s->triInst[p]];
        const rtx_vertex &v0 = s->verts[in.modelIndex][s->idx[in.modelIndex][3 * (size_t)s->triLocal[p]]];
        s->triAlpha[p] = s->mats[v0.materialIndex].textureIdDiffuse != -1;
    }
    size_t nprim = s->ntri + s->spheres.size();
    if (nprim > 0) {
        std::vector<std::array<double, 6>> box(nprim);
        for (size_t p = 0; p < nprim; ++p) {
            double lo[3], hi[3];
            triBounds(*s, (int)p, lo, hi);
            box[p] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
        }
        s->primOrder.resize(nprim);
        for (size_t p = 0; p < nprim; ++p) s->primOrder[p] = (int)p;
        buildNode(*s, box, 0, (int)nprim);
    }
    *out = s.release();
    return RTX_OK;
}

void orc_scene_destroy(orc_scene *s) { delete s; }

void orc_adrrs_counters(uint64_t out[3]) {
    for (int k = 0; k < 3; ++k) out[k] = g_adrrsCount[k].exchange(0);
}

int orc_estimate_set(orc_scene *s, const float *rgba, uint32_t width, uint32_t height) {
    if (!s || (!rgba && width && height)) {
        g_err = "NULL argument";
        return RTX_E_INVALID;
    }
    s->estimate.assign(rgba, rgba + 4 * (size_t)width * height);
    s->estW = width;
    s->estH = height;
    return RTX_OK;
}

int orc_guiding_set(orc_scene *s, const rtx_aabb *regions, const rtx_vmm_theta *vmms, uint32_t n) {
    if (!s || (n && (!regions || !vmms))) {
        g_err = "NULL argument";
        return RTX_E_INVALID;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (vmms[i].usedDistributions < 1 || vmms[i].usedDistributions > RTX_MAX_DISTRIBUTIONS) {
            g_err = "usedDistributions out of range";
            return RTX_E_INVALID;
        }
    s->regions.assign(regions, regions + n);
    s->vmms.assign(vmms, vmms + n);
    return RTX_OK;
}



int orc_render_paths(const orc_scene *s, const rtx_frame_params *fp, const uint64_t *ids, size_t n, int prec,
                     int threads, double *rgb, uint32_t *sig) {
    if (!s || !fp || (!ids && n) || (!rgb && n)) {
        g_err = "NULL argument";
        return RTX_E_INVALID;
    }
    int rc = checkParams(*s, fp);
    if (rc != RTX_OK) return rc;
    if (prec != 32 && prec != 64) {
        g_err = "prec must be 32 or 64";
        return RTX_E_INVALID;
    }
    if ((fp->pc.useGuiding || fp->pc.updateGuiding) && s->regions.empty()) {
        g_err = "guiding without guiding regions (orc_guiding_set)";
        return RTX_E_STATE;
    }
    parallelFor(n, threads, [&](size_t i) {
        double r[3];
        uint32_t st, hs;
        if (prec == 64) {
            Tracer<double> t(*s, *fp);
            V3<double> v = t.renderPath(ids[i]);
            toD(v, r);
            st = t.seed;
            hs = t.primHash;
        } else {
            Tracer<float> t(*s, *fp);
            V3<float> v = t.renderPath(ids[i]);
            toD(v, r);
            st = t.seed;
            hs = t.primHash;
        }
        rgb[3 * i + 0] = r[0];
        rgb[3 * i + 1] = r[1];
        rgb[3 * i + 2] = r[2];
        if (sig) {
            sig[2 * i + 0] = st;
            sig[2 * i + 1] = hs;
        }
    });
    return RTX_OK;
}

int orc_render_paths_record(const orc_scene *s, const rtx_frame_params *fp, const uint64_t *ids, size_t n, int prec,
                            int threads, double *rgb, uint32_t *sig, rtx_directional_data *dd) {
    if (!s || !fp || (!ids && n) || (!rgb && n) || (!dd && n)) {
        g_err = "NULL argument";
        return RTX_E_INVALID;
    }
    int rc = checkParams(*s, fp);
    if (rc != RTX_OK) return rc;
    if ((fp->pc.useGuiding || fp->pc.updateGuiding) && s->regions.empty()) {
        g_err = "guiding without guiding regions (orc_guiding_set)";
        return RTX_E_STATE;
    }
    parallelFor(n, threads, [&](size_t i) {
        rtx_directional_data *slots = dd + (size_t)RTX_MAX_DIRECTIONAL_DATA_PER_PIXEL * i;
        for (int k = 0; k < RTX_MAX_DIRECTIONAL_DATA_PER_PIXEL; ++k) {
            slots[k] = rtx_directional_data{};
            slots[k].flags = kDirInvalid; // resetSamplesToInvalid, raytrace.rgen:1615-1621
        }
        double r[3];
        uint32_t st, hs;
        if (prec == 64) {
            Tracer<double> t(*s, *fp);
            t.dd = slots;
            toD(t.renderPath(ids[i]), r);
            st = t.seed;
            hs = t.primHash;
        } else {
            Tracer<float> t(*s, *fp);
            t.dd = slots;
            toD(t.renderPath(ids[i]), r);
            st = t.seed;
            hs = t.primHash;
        }
        for (int a = 0; a < 3; ++a) rgb[3 * i + a] = r[a];
        if (sig) {
            sig[2 * i] = st;
            sig[2 * i + 1] = hs;
        }
    });
    return RTX_OK;
}

int orc_trace_closest(const orc_scene *s, const rtx_ray *rays, size_t n, int brute, double *t, int32_t *prim,
                      double *u, double *v) {
    return orc_trace_closest_prec(s, rays, n, brute, 64, t, prim, u, v);
}

int orc_trace_closest_prec(const orc_scene *s, const rtx_ray *rays, size_t n, int brute, int prec, double *t,
                           int32_t *prim, double *u, double *v) {
    if (!s || (!rays && n)) {
        g_err = "NULL argument";
        return RTX_E_INVALID;
    }
    parallelFor(n, 0, [&](size_t i) {
        const rtx_ray &r = rays[i];
        HitRec h;
        bool hit = prec == 32 ? closest<float>(*s, f3<float>(r.org), f3<float>(r.dir), r.tmin, r.tmax, brute != 0, h)
                              : closest<double>(*s, f3<double>(r.org), f3<double>(r.dir), (double)r.tmin,
                                                (double)r.tmax, brute != 0, h);
        if (t) t[i] = hit ? h.t : (double)r.tmax;
        if (prim) prim[i] = hit ? h.prim : -1;
        if (u) u[i] = hit ? h.u : 0.0;
        if (v) v[i] = hit ? h.v : 0.0;
    });
    return RTX_OK;
}

int orc_trace_occluded(const orc_scene *s, const rtx_ray *rays, size_t n, int brute, uint8_t *occ) {
    if (!s || (!rays && n) || (!occ && n)) {
        g_err = "NULL argument";
        return RTX_E_INVALID;
    }
    parallelFor(n, 0, [&](size_t i) {
        const rtx_ray &r = rays[i];
        occ[i] = occluded<double>(*s, f3<double>(r.org), f3<double>(r.dir), (double)r.tmin, (double)r.tmax,
                                  brute != 0)
                     ? 1
                     : 0;
    });
    return RTX_OK;
}

uint32_t orc_tea(uint32_t v0, uint32_t v1) { return tea(v0, v1); }
uint32_t orc_lcg_step(uint32_t s) { return 1664525u * s + 1013904223u; }
uint32_t orc_pcg32_first(uint64_t a, uint64_t b) { return pcg32First(a, b); }

namespace {
const orc_scene &emptyScene() {
    static orc_scene s;
    return s;
}
rtx_frame_params defaultParams() {
    rtx_frame_params fp;
    std::memset(&fp, 0, sizeof(fp));
    fp.pc.usePowerHeuristic = 1; // src/RayTracingApp.h:92
    return fp;
}
} // namespace

double orc_weight_window(const double estimate[3], const double adjoint[3], const double throughput[3],
                         double adrrsS, uint32_t *state, int *n) {
    rtx_frame_params fp = defaultParams();
    fp.pc.adrrsS = (float)adrrsS;
    Tracer<double> t(emptyScene(), fp);
    t.seed = *state;
    t.estimate = V3<double>(estimate[0], estimate[1], estimate[2]);
    int nn = 1;
    double q = t.weightWindow(V3<double>(throughput[0], throughput[1], throughput[2]),
                              V3<double>(adjoint[0], adjoint[1], adjoint[2]), nn);
    *state = t.seed;
    *n = nn;
    return q;
}

void orc_camera_ray(const rtx_camera *cam, uint32_t w, uint32_t h, uint32_t px, uint32_t py, uint32_t *state,
                    double org[3], double dir[3]) {
    rtx_frame_params fp = defaultParams();
    fp.cam = *cam;
    fp.width = w;
    fp.height = h;
    Tracer<double> t(emptyScene(), fp);
    t.seed = *state;
    V3<double> o, d;
    t.cameraRay(px, py, w, h, o, d);
    *state = t.seed;
    toD(o, org);
    toD(d, dir);
}

int orc_ic_frame(const orc_scene *s, const rtx_frame_params *fp, int prec, int threads, rtx_cache_header *hdr,
                 rtx_cache_data *entries, rtx_sphere *spheres, double *rgb, uint32_t *sig, orc_ic_job *jobs,
                 size_t jobCap, size_t *numJobs, double *sub, uint32_t *subSig) {
    if (!s || !fp || !hdr || !entries || !spheres || !rgb) {
        g_err = "NULL argument";
        return RTX_E_INVALID;
    }
    int rc = checkParams(*s, fp, true);
    if (rc != RTX_OK) return rc;
    if (!fp->pc.useIrradianceCache && !fp->pc.useADRRS) {
        g_err = "orc_ic_frame needs pc.useIrradianceCache or pc.useADRRS";
        return RTX_E_INVALID;
    }
    if (fp->pc.useGuiding && s->regions.empty()) {
        g_err = "useGuiding without guiding regions (orc_guiding_set)";
        return RTX_E_STATE;
    }
    if (prec != 32 && prec != 64) {
        g_err = "prec must be 32 or 64";
        return RTX_E_INVALID;
    }
    IcSnapshot live;
    live.hdr = *hdr;
    live.e.assign(entries, entries + hdr->maxCaches);
    live.sph.assign(spheres, spheres + hdr->maxCaches);
    std::vector<IcJobRec> jr;
    std::vector<double> sb;
    std::vector<uint32_t> ss;
    if (prec == 64) icFrame<double>(*s, *fp, live, threads, rgb, sig, jr, sb, ss);
    else icFrame<float>(*s, *fp, live, threads, rgb, sig, jr, sb, ss);
    *hdr = live.hdr;
    std::copy(live.e.begin(), live.e.end(), entries);
    std::copy(live.sph.begin(), live.sph.end(), spheres);
    if (numJobs) *numJobs = jr.size();
    const size_t keep = std::min(jobCap, jr.size());
    for (size_t j = 0; j < keep; ++j) {
        if (jobs) {
            jobs[j].type = jr[j].type;
            jobs[j].target = jr[j].target;
            for (int a = 0; a < 3; ++a) {
                jobs[j].origin[a] = jr[j].o[a];
                jobs[j].normal[a] = jr[j].n[a];
            }
        }
        if (sub) std::copy(sb.begin() + 4 * j * kIcStrata, sb.begin() + 4 * (j + 1) * kIcStrata, sub + 4 * j * kIcStrata);
        if (subSig)
            std::copy(ss.begin() + 2 * j * kIcStrata, ss.begin() + 2 * (j + 1) * kIcStrata, subSig + 2 * j * kIcStrata);
    }
    return RTX_OK;
}

int orc_ic_calculate(const orc_scene *s, const rtx_frame_params *fp, int prec, const rtx_cache_header *hdr,
                     const rtx_cache_data *entries, const rtx_sphere *spheres, uint32_t job, const double origin[3],
                     const double normal[3], double color[3], double rot[3], double trans[3], double *harmonicR) {
    if (!s || !fp || !hdr || !entries || !spheres || !origin || !normal) {
        g_err = "NULL argument";
        return RTX_E_INVALID;
    }
    int rc = checkParams(*s, fp, true);
    if (rc != RTX_OK) return rc;
    IcSnapshot snap;
    snap.hdr = *hdr;
    snap.e.assign(entries, entries + hdr->maxCaches);
    snap.sph.assign(spheres, spheres + hdr->maxCaches);
    IcJobRec jb{1, 0, {origin[0], origin[1], origin[2]}, {normal[0], normal[1], normal[2]}};
    const uint32_t frameSeed = pcg32First(fp->seed, fp->sampleOffset);
    std::vector<double> sub(4 * kIcStrata);
    uint32_t sig[2];
    for (int st = 0; st < kIcStrata; ++st) {
        if (prec == 64)
            icSubPath<double>(*s, *fp, snap, frameSeed, job, jb, st / kIcStrataTheta, st % kIcStrataTheta, &sub[4 * st], sig);
        else
            icSubPath<float>(*s, *fp, snap, frameSeed, job, jb, st / kIcStrataTheta, st % kIcStrataTheta, &sub[4 * st], sig);
    }
    return orc_ic_reduce(prec, fp->seed, fp->sampleOffset, job, normal, sub.data(), color, rot, trans, harmonicR);
}

int orc_ic_apply(const rtx_frame_params *fp, int prec, rtx_cache_header *hdr, rtx_cache_data *entries,
                 rtx_sphere *spheres, const orc_ic_job *jobs, size_t numJobs, const double *sub) {
    if (!fp || !hdr || !entries || !spheres || (numJobs && (!jobs || !sub))) {
        g_err = "NULL argument";
        return RTX_E_INVALID;
    }
    IcSnapshot live;
    live.hdr = *hdr;
    live.e.assign(entries, entries + hdr->maxCaches);
    live.sph.assign(spheres, spheres + hdr->maxCaches);
    const uint32_t frameSeed = pcg32First(fp->seed, fp->sampleOffset);
    for (size_t j = 0; j < numJobs; ++j) {
        IcJobRec jb{jobs[j].type, jobs[j].target, {}, {}};
        for (int a = 0; a < 3; ++a) {
            jb.o[a] = jobs[j].origin[a];
            jb.n[a] = jobs[j].normal[a];
        }
        const double *sj = sub + 4 * j * kIcStrata;
        if (prec == 64) icCommit<double>(live, fp->pc, jb, icReduce<double>(frameSeed, (uint32_t)j, vd<double>(jb.n), sj));
        else icCommit<float>(live, fp->pc, jb, icReduce<float>(frameSeed, (uint32_t)j, vd<float>(jb.n), sj));
    }
    *hdr = live.hdr;
    std::copy(live.e.begin(), live.e.end(), entries);
    std::copy(live.sph.begin(), live.sph.end(), spheres);
    return RTX_OK;
}

int orc_ic_reduce(int prec, uint64_t seed, uint32_t sample, uint32_t job, const double normal[3], const double *sub,
                  double color[3], double rot[3], double trans[3], double *harmonicR) {
    uint32_t frameSeed = pcg32First(seed, sample);
    if (prec == 64) {
        IcResult<double> r = icReduce<double>(frameSeed, job, vd<double>(normal), sub);
        toD(r.color, color);
        toD(r.rot, rot);
        toD(r.trans, trans);
        *harmonicR = r.harmonicR;
    } else {
        IcResult<float> r = icReduce<float>(frameSeed, job, vd<float>(normal), sub);
        toD(r.color, color);
        toD(r.rot, rot);
        toD(r.trans, trans);
        *harmonicR = r.harmonicR;
    }
    return RTX_OK;
}

void orc_ic_stratum(uint64_t seed, uint32_t sample, uint32_t job, int k, int j, double *phi, double *theta,
                    uint32_t *state) {
    icStratum<double>(pcg32First(seed, sample), job, k, j, *phi, *theta, *state);
}

int orc_ic_query(const orc_scene *s, const rtx_push_constant *pc, const rtx_cache_header *hdr,
                 const rtx_cache_data *entries, const rtx_sphere *spheres, const double x[3], const double n[3],
                 double color[3]) {
    IcSnapshot snap;
    snap.hdr = *hdr;
    snap.e.assign(entries, entries + hdr->maxCaches);
    snap.sph.assign(spheres, spheres + hdr->maxCaches);
    rtx_frame_params fp = defaultParams();
    fp.pc = *pc;
    Tracer<double> t(s ? *s : emptyScene(), fp);
    t.ic = &snap;
    V3<double> c;
    if (!t.queryIC(vd<double>(x), vd<double>(n), c)) return 0;
    toD(c, color);
    return 1;
}

double orc_vmm_pdf(const rtx_vmm_theta *m, const double x[3], const double wo[3], int parallax) {
    return Tracer<double>::vmm(vd<double>(wo), *m, vd<double>(x), parallax != 0);
}

void orc_sample_vmm(const rtx_vmm_theta *m, const double x[3], int parallax, uint32_t *state, double out[3]) {
    Tracer<double> t(emptyScene(), defaultParams());
    t.seed = *state;
    toD(t.sampleVMM(*m, vd<double>(x), parallax != 0), out);
    *state = t.seed;
}

double orc_fresnel(double eta, double c) { return Tracer<double>::fresnel(eta, c); }
double orc_fresnel_conductor(double c, double eta, double k) { return Tracer<double>::fresnelConductor(c, eta, k); }
double orc_beckmann_d(const rtx_material *m, const double n[3], const double h[3]) {
    return Tracer<double>::D(*m, vd<double>(n), vd<double>(h));
}
// D and G1 of the material's microfacet distribution (Beckmann with the
// reference's G1, or GGX)
double orc_microfacet_d(const rtx_material *m, const double n[3], const double h[3]) {
    return Tracer<double>::D(*m, vd<double>(n), vd<double>(h));
}
double orc_microfacet_g1(const rtx_material *m, const double n[3], const double h[3], const double v[3]) {
    if (m->distribution == RTX_DISTRIBUTION_GGX) return Tracer<double>::ggxG1(*m, vd<double>(n), vd<double>(h), vd<double>(v));
    return Tracer<double>::G1(*m, vd<double>(n), vd<double>(h), vd<double>(v));
}
double orc_pdf_bsdf(const rtx_material *m, const double n[3], const double wi[3], const double wo[3]) {
    return Tracer<double>::pdfBSDF(*m, vd<double>(n), vd<double>(wi), vd<double>(wo));
}
void orc_eval_bsdf(const rtx_material *m, const double n[3], const double wi[3], const double wo[3], int ff,
                   double out[3]) {
    rtx_material mm = *m; // without a scene: textures are not applied
    mm.textureIdDiffuse = mm.textureIdSpecular = -1;
    Tracer<double> t(emptyScene(), defaultParams());
    toD(t.evalBsdf(mm, vd<double>(n), vd<double>(wi), vd<double>(wo), ff != 0), out);
}
// Texture look-ups of a scene: texture(textureSamplers[slot], uv) for a
// material texture (slot >= 1) and the environment map's colour in `dir`
// (miss shader form: dir is normalized first).
int orc_tex_sample(const orc_scene *s, int slot, double u, double v, double out[4]) {
    if (!s || slot < 1 || (size_t)slot >= s->textures.size() || s->textures[slot].w == 0) {
        g_err = "no texture in that slot";
        return RTX_E_INVALID;
    }
    texSample<double>(*s, slot, u, v, out);
    return RTX_OK;
}
void orc_env_color(const orc_scene *s, const double dir[3], double out[3]) {
    toD(envColor<double>(*s, normalize(vd<double>(dir))), out);
}
double orc_sample_bsdf(const rtx_material *m, const double wi[3], const double n[3], int ff, uint32_t *state,
                       double wo[3]) {
    rtx_frame_params fp = defaultParams();
    Tracer<double> t(emptyScene(), fp);
    t.seed = *state;
    V3<double> d(0, 0, 0);
    double pdf = t.sampleBSDF(*m, vd<double>(wi), vd<double>(n), ff != 0, d);
    *state = t.seed;
    toD(d, wo);
    return pdf;
}
void orc_sample_cosine(const double n[3], uint32_t *state, double out[3]) {
    rtx_frame_params fp = defaultParams();
    Tracer<double> t(emptyScene(), fp);
    t.seed = *state;
    toD(t.randomInHemisphereCosine(vd<double>(n)), out);
    *state = t.seed;
}
void orc_sample_cosine_power(const double r[3], double p, uint32_t *state, double out[3]) {
    rtx_frame_params fp = defaultParams();
    Tracer<double> t(emptyScene(), fp);
    t.seed = *state;
    toD(t.randomInHemisphereCosinePower(vd<double>(r), p), out);
    *state = t.seed;
}
void orc_sample_beckmann_normal(const rtx_material *m, const double n[3], uint32_t *state, double out[3]) {
    rtx_frame_params fp = defaultParams();
    Tracer<double> t(emptyScene(), fp);
    t.seed = *state;
    toD(t.randomMicrofacetNormal(*m, vd<double>(n)), out);
    *state = t.seed;
}
void orc_sample_unit_sphere(uint32_t *state, double out[3]) {
    rtx_frame_params fp = defaultParams();
    Tracer<double> t(emptyScene(), fp);
    t.seed = *state;
    toD(t.randomOnUnitSphere(), out);
    *state = t.seed;
}
double orc_sample_lights(const orc_scene *s, const double origin[3], const double normal[3], uint32_t *state,
                         double lightDir[3], double lightColor[3], double *lightDistance) {
    rtx_frame_params fp = defaultParams();
    Tracer<double> t(*s, fp);
    t.seed = *state;
    V3<double> ld, lc;
    double dist = 0;
    double pdf = t.sampleLights(vd<double>(origin), vd<double>(normal), ld, lc, dist);
    *state = t.seed;
    toD(ld, lightDir);
    toD(lc, lightColor);
    *lightDistance = dist;
    return pdf;
}

} // extern "C"
