// Device vector math and random numbers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rtx {

#define RTX_DI __device__ __forceinline__

struct V3 {
    float x, y, z;
};
RTX_DI V3 v3(float a, float b, float c) { return V3{a, b, c}; }
RTX_DI V3 v3(float s) { return V3{s, s, s}; }
RTX_DI V3 v3(const float *p) { return V3{p[0], p[1], p[2]}; }
RTX_DI V3 v3(float4 a) { return V3{a.x, a.y, a.z}; }
RTX_DI float4 f4(V3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
RTX_DI V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RTX_DI V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RTX_DI V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
RTX_DI V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
RTX_DI V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
RTX_DI V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
RTX_DI V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
RTX_DI V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
RTX_DI float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RTX_DI V3 cross(V3 a, V3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
RTX_DI float length(V3 a) { return sqrtf(dot(a, a)); }
RTX_DI V3 normalize(V3 a) { return a / length(a); }
// GLSL reflect / refract (GLSL 4.60 section 8.5)
RTX_DI V3 reflect(V3 i, V3 n) { return i - 2.0f * dot(n, i) * n; }
RTX_DI V3 refract(V3 i, V3 n, float eta) {
    float d = dot(n, i);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return v3(0.0f);
    return eta * i - (eta * d + sqrtf(k)) * n;
}
RTX_DI float sel(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// column-major mat4 * (v, w), GLM summation order (c0*x + c1*y) + (c2*z + c3*w)
RTX_DI V3 xform(const float *m, V3 v, float w) {
    float o[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        float a0 = m[0 * 4 + r] * v.x + m[1 * 4 + r] * v.y;
        float a1 = m[2 * 4 + r] * v.z + m[3 * 4 + r] * w;
        o[r] = a0 + a1;
    }
    return {o[0], o[1], o[2]};
}

constexpr float kPi = 3.1415926535897932384626433832795f;
constexpr float kInv2Pi = 1.0f / (2.0f * 3.1415926535897932384626433832795f);
constexpr float kTMin = 0.001f;     // shaders/raytrace.rgen:52
constexpr float kTMax = 1000000.0f; // shaders/raytrace.rgen:53
constexpr int kMaxRejections = 65536; // bound of the rejection loops (never reached in practice)

// Path signature: a fold of every discrete traversal outcome in path order
// (extend-ray primitive ids, shadow-ray visibility, probe-ray primitive
// ids), returned with the final LCG state for per-path parity checks.
constexpr uint32_t kSigPrime = 16777619u;
constexpr uint32_t kSigVisible = 0x1b873593u;
constexpr uint32_t kSigOccluded = 0x5bd1e995u;
constexpr uint32_t kSigProbe = 0x85ebca6bu;
constexpr uint32_t kSigCacheHit = 0xc2b2ae35u; // an irradiance-cache lookup that ended the path
constexpr uint32_t kSigRRKill = 0x27d4eb2fu;   // a weight-window roulette that ended the path
constexpr uint32_t kSigSplit = 0x165667b1u;    // ^ number of split records a vertex queued
RTX_DI uint32_t sigFold(uint32_t h, uint32_t v) { return (h ^ v) * kSigPrime; }

// ------------------------------------------------------------------ RNG
// tea: 16 rounds of the Tiny Encryption Algorithm (shaders/random.glsl:13-27)
RTX_DI uint32_t tea(uint32_t val0, uint32_t val1) {
    uint32_t v0 = val0, v1 = val1, s0 = 0;
#pragma unroll
    for (uint32_t n = 0; n < 16; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

// PCG32 first output after pcg32_srandom(initstate, initseq) (O'Neill 2014):
// the per-sample frame seed that replaces the reference's per-frame
// glm::linearRand (src/RayTracingApp.cpp:1121).
__host__ __device__ inline uint32_t pcg32First(uint64_t initstate, uint64_t initseq) {
    uint64_t inc = (initseq << 1u) | 1u;
    uint64_t state = 0;
    state = state * 6364136223846793005ULL + inc;
    state += initstate;
    state = state * 6364136223846793005ULL + inc;
    uint64_t old = state;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}

// Numerical-Recipes LCG with 24-bit output (shaders/random.glsl:31-57)
struct Rng {
    uint32_t s;
    RTX_DI float next() {
        s = 1664525u * s + 1013904223u;
        return (float)(s & 0x00FFFFFFu) * (1.0f / 16777216.0f);
    }
    RTX_DI float negPos() { return next() * 2.0f - 1.0f; }
    RTX_DI int integer(int max) { return (int)(next() * (float)(max + 1)); }
};

} // namespace rtx
