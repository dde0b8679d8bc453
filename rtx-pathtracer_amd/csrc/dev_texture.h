// Texture reads on the device: the reference's binding 7 (textureSamplers[],
// slot 0 the environment map, slots 1.. material textures) sampled as its
// samplers do, and the any-hit alpha test of raytrace.rahit.
#pragma once

#include "dev_math.h"
#include "internal.h"

namespace rtx {

// Vulkan linear filtering with REPEAT addressing, the samplers of
// src/SceneLoader.cpp:209-227: texel centres at (i + 0.5) / size, the four
// nearest texels weighted bilinearly, indices wrapped.  Axis set-up for
// coordinate u on n texels.
RTX_DI void bilerpAxis(float u, int n, int &i0, int &i1, float &w) {
#pragma clang fp contract(off)
    float x = u * (float)n - 0.5f;
    if (!(fabsf(x) < 1e30f)) { // NaN or infinite coordinates read texel 0
        i0 = i1 = 0;
        w = 0.0f;
        return;
    }
    float xf = floorf(x);
    w = x - xf;
    int i = (int)fmodf(xf, (float)n);
    if (i < 0) i += n;
    i0 = i;
    i1 = i + 1 == n ? 0 : i + 1;
}
RTX_DI float4 lerp4(float4 a, float4 b, float4 c, float4 d, float fx, float fy) {
#pragma clang fp contract(off)
    float gx = 1.0f - fx, gy = 1.0f - fy;
    float4 t = make_float4(a.x * gx + b.x * fx, a.y * gx + b.y * fx, a.z * gx + b.z * fx, a.w * gx + b.w * fx);
    float4 u = make_float4(c.x * gx + d.x * fx, c.y * gx + d.y * fx, c.z * gx + d.z * fx, c.w * gx + d.w * fx);
    return make_float4(t.x * gy + u.x * fy, t.y * gy + u.y * fy, t.z * gy + u.z * fy, t.w * gy + u.w * fy);
}
// One texel of an R8G8B8A8_SRGB texture: RGB decoded to linear, alpha / 255.
RTX_DI float4 texel8(const DevScene &S, int4 info, int x, int y) {
    uint32_t p = S.texels[(size_t)info.x + (size_t)y * (uint32_t)info.y + (uint32_t)x];
    return make_float4(S.srgbToLinear[p & 255u], S.srgbToLinear[(p >> 8) & 255u], S.srgbToLinear[(p >> 16) & 255u],
                       (float)(p >> 24) / 255.0f);
}
// texture(textureSamplers[slot], uv) for a material texture (slot >= 1).
RTX_DI float4 texSample(const DevScene &S, int slot, float u, float v) {
    int4 info = S.texInfo[slot];
    int x0, x1, y0, y1;
    float fx, fy;
    bilerpAxis(u, info.y, x0, x1, fx);
    bilerpAxis(v, info.z, y0, y1, fy);
    return lerp4(texel8(S, info, x0, y0), texel8(S, info, x1, y0), texel8(S, info, x0, y1), texel8(S, info, x1, y1),
                 fx, fy);
}
// The environment map (texture slot 0, RGBA32F) in direction `dir`:
// raytrace.rmiss:24-28 and raytrace.rgen:410-416.  u = atan(x, -z) / 2pi
// (the macro `atan * M_INV_2PI` expands to a division), v = acos(y) / pi.
// Black without a map (the reference's default texture, SceneLoader.cpp:80-86).
RTX_DI V3 envColor(const DevScene &S, V3 dir) {
    if (!S.env) return v3(0.0f);
    float u = atan2f(dir.x, -dir.z) / (2.0f * kPi);
    float v = acosf(dir.y) / kPi;
    int x0, x1, y0, y1;
    float fx, fy;
    bilerpAxis(u, (int)S.envW, x0, x1, fx);
    bilerpAxis(v, (int)S.envH, y0, y1, fy);
    const float4 *e = S.env;
    const size_t w = S.envW;
    float4 c = lerp4(e[y0 * w + x0], e[y0 * w + x1], e[y1 * w + x0], e[y1 * w + x1], fx, fy);
    return v3(c.x, c.y, c.z);
}

// Saturating float -> uint conversion (GLSL uint(float) on the reference's
// hardware: negative and NaN give 0, values past 2^32 - 1 give 2^32 - 1).
RTX_DI uint32_t satU32(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

// raytrace.rahit:21-47: a candidate hit on triangle `prim` (barycentrics bu,
// bv of vertices 1 and 2 at distance t) whose material has a diffuse texture
// is ignored when rnd(seed) > alpha, alpha being the texture's at the
// candidate's texture coordinates and seed = tea(uint(u * 1e8 + origin.x *
// t), randomUInt) (GLSL's uint(vec2) takes the first component).
RTX_DI bool alphaAccept(const DevScene &S, int prim, float ox, float t, float bu, float bv, uint32_t randomUInt) {
#pragma clang fp contract(off)
    const int4 ref = S.triRef[prim];
    const rtx_vertex &v0 = S.vertices[ref.y], &v1 = S.vertices[ref.z], &v2 = S.vertices[ref.w];
    float b0 = 1.0f - bu - bv;
    float u = v0.texCoord[0] * b0 + v1.texCoord[0] * bu + v2.texCoord[0] * bv;
    float v = v0.texCoord[1] * b0 + v1.texCoord[1] * bu + v2.texCoord[1] * bv;
    Rng r{tea(satU32(u * 100000000.0f + ox * t), randomUInt)};
    float alpha = texSample(S, S.materials[v0.materialIndex].textureIdDiffuse, u, v).w;
    return !(r.next() > alpha);
}

} // namespace rtx
