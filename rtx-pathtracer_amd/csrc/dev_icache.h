// Irradiance-cache lookup on the device.  Replaces the reference's cache
// acceleration structure (a refitted AABB BLAS, src/IrradianceCache.cpp:19-44,
// 81-104) and its point query through the intersection and any-hit shaders
// (shaders/raytrace.rgen:748-773, raytrace.irradiance.rint:11-21,
// raytrace.irradiance.rahit:18-51).
//
// Index: a multi-level hashed grid.  A record of radius r lives at the
// level L whose cells (c0 * 2^L) are at least 2r wide, in every cell of that
// level its sphere's box touches (at most 3 per axis with the margin below).
// A lookup visits one cell per occupied level and weighs the records whose
// sphere contains the point, in (level, record id) order.
#pragma once

#include "dev_math.h"
#include "dev_shading.h"
#include "internal.h"

namespace rtx {

constexpr int kIcMaxLevel = 24;
constexpr int kIcAxisBits = 19;
constexpr int32_t kIcAxisHalf = 1 << (kIcAxisBits - 1);

RTX_DI float icCellSize(float c0, int level) { return ldexpf(c0, level); }

RTX_DI int32_t icCellCoord(float v, float cs) {
    float q = floorf(v / cs);
    q = fminf(fmaxf(q, -(float)kIcAxisHalf), (float)(kIcAxisHalf - 1));
    return (int32_t)q;
}

RTX_DI uint64_t icKey(int level, int32_t ix, int32_t iy, int32_t iz) {
    const uint64_t m = (1ull << kIcAxisBits) - 1;
    return ((uint64_t)level << (3 * kIcAxisBits)) | (((uint64_t)(ix + kIcAxisHalf) & m) << (2 * kIcAxisBits)) |
           (((uint64_t)(iy + kIcAxisHalf) & m) << kIcAxisBits) | ((uint64_t)(iz + kIcAxisHalf) & m);
}

RTX_DI uint32_t icBucket(uint64_t key, uint32_t nb) {
    // splitmix64 finaliser
    key ^= key >> 30;
    key *= 0xbf58476d1ce4e5b9ull;
    key ^= key >> 27;
    key *= 0x94d049bb133111ebull;
    key ^= key >> 31;
    return (uint32_t)key & (nb - 1);
}

// The level a sphere of radius r is filed under.
RTX_DI int icLevel(float c0, float r) {
    int level = 0;
    float cs = c0;
    while (cs < 2.0f * r && level < kIcMaxLevel) {
        cs *= 2.0f;
        ++level;
    }
    return level;
}

// The jobs [j0, jEnd) of an IcJobRange other than kIcJobsHost.
RTX_DI void icJobRange(const IcBuffers &ic, int range, uint32_t firstJob, uint32_t &j0, uint32_t &jEnd) {
    switch (range) {
    case kIcJobsUpdates: j0 = 0, jEnd = ic.frame[3]; break;
    case kIcJobsCreations: j0 = ic.frame[3], jEnd = ic.frame[4]; break;
    default: j0 = firstJob, jEnd = ic.frame[4]; break;
    }
}

// raytrace.rgen:806-813
RTX_DI bool isICCapable(int type, bool onGlossy) {
    if (type == RTX_MAT_DIFFUSE || type == RTX_MAT_LIGHT) return true;
    return onGlossy && !hasDiscreteDirection(type);
}

// raytrace.rgen:815-828
RTX_DI V3 approxDiffuse(const Mat &m, V3 normal, V3 wi) {
    switch (m.type) {
    case RTX_MAT_DIFFUSE:
    case RTX_MAT_LIGHT:
    case RTX_MAT_PHONG: return m.kd * (1.0f / kPi);
    case RTX_MAT_ROUGH_CONDUCTOR: return v3(fresnelConductor(dot(wi, normal), m.eta, m.k) * (1.0f / kPi));
    default: return v3(-1.0f);
    }
}

struct IcQueryParams {
    float invA;          // 1 / irradianceA
    bool gradients;      // useIrradianceGradients
    bool visibility;     // irradianceCachePerformVisibilityCheck
};

// The bucket of x's cell at one level of the index.
RTX_DI uint32_t icLevelBucket(const DevIC &ic, float c0, int level, V3 x, uint64_t &key) {
    const float cs = icCellSize(c0, level);
    key = icKey(level, icCellCoord(x.x, cs), icCellCoord(x.y, cs), icCellCoord(x.z, cs));
    return icBucket(key, ic.nb);
}

// One record's term of the weighted mean (raytrace.irradiance.rint:17-20,
// raytrace.irradiance.rahit:18-51), when its sphere contains x.
RTX_DI void icAccept(const DevIC &ic, const IcQueryParams &qp, V3 x, V3 n, const IcItem &item, V3 &sum,
                     float &total) {
    V3 oc = x - v3(item.center);
    if (!(length(oc) <= item.radius)) return;
    const rtx_cache_data &e = ic.entries[item.entry];
    V3 ni = v3(e.normal);
    float w = 1.0f / (length(oc) / e.harmonicR + sqrtf(1.0f - dot(n, ni)));
    if (isnan(w) || isinf(w)) w = 1000000.0f;
    if (w <= qp.invA) return;
    if (qp.visibility && !(-0.001f <= dot(oc, (n + ni) / 2.0f))) return;
    V3 col = v3(e.color);
    if (qp.gradients) {
        float E = length(col);
        V3 cn = E != 0.0f ? normalize(col) : v3(0.0f);
        col = cn * (E + dot(cross(ni, n), v3(e.rotGrad)) + dot(oc, v3(e.transGrad)));
    }
    sum = sum + w * col;
    total += w;
}

// Weighted cache colour at x (normal n); false when no record accepts x.
// Records are weighed in (level, bucket order) order.  The loads are
// arranged so that each level costs about one memory latency: the next
// level's bucket range is fetched while this level's items are read, and the
// next item while this one is weighed.
RTX_DI bool icQuery(const DevIC &ic, const IcQueryParams &qp, V3 x, V3 n, V3 &color) {
    uint32_t lm = ic.meta[0];
    if (lm == 0) return false;
    const float c0 = __uint_as_float(ic.meta[2]);
    V3 sum = v3(0.0f);
    float total = 0.0f;
    uint64_t key;
    uint32_t b = icLevelBucket(ic, c0, __ffs(lm) - 1, x, key);
    uint32_t it = ic.bucketStart[b], end = ic.bucketStart[b + 1];
    for (;;) {
        lm &= lm - 1;
        uint64_t nkey = 0;
        uint32_t nit = 0, nend = 0;
        if (lm != 0) {
            const uint32_t nbk = icLevelBucket(ic, c0, __ffs(lm) - 1, x, nkey);
            nit = ic.bucketStart[nbk];
            nend = ic.bucketStart[nbk + 1];
        }
        if (it < end) {
            IcItem cur = ic.items[it];
            for (;;) {
                const bool more = ++it < end;
                IcItem nxt;
                if (more) nxt = ic.items[it];
                if (cur.key == key) icAccept(ic, qp, x, n, cur, sum, total);
                if (!more) break;
                cur = nxt;
            }
        }
        if (lm == 0) break;
        key = nkey;
        it = nit;
        end = nend;
    }
    if (total > 0.0f) {
        color = sum / total;
        return true;
    }
    return false;
}

} // namespace rtx
