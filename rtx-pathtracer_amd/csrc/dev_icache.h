// Irradiance-cache lookup on the device.  Replaces the reference's cache
// acceleration structure (a refitted AABB BLAS, src/IrradianceCache.cpp:19-44,
// 81-104) and its point query through the intersection and any-hit shaders
// (shaders/raytrace.rgen:748-773, raytrace.irradiance.rint:11-21,
// raytrace.irradiance.rahit:18-51).
//
// Index: a multi-level hashed grid.  A record of radius r lives at the
// level L whose cells (c0 * 2^L) are a little wider than r grown by a
// rounding margin m, in every cell of that level its sphere's box (half-width
// m) touches: at most 3 per axis.  Cells between r and 2r wide keep the
// records a lookup tests but does not accept few (the cell's neighbourhood
// of reach r spans 3r-4r per axis against 4r-6r for cells of 2r-4r).
// Each occupied level has its own range of buckets, and a record files at
// most one item per bucket, so a lookup needs no cell key: it visits one
// bucket per occupied level and weighs the records in it whose sphere
// contains the point, in (level, record id) order.  A record that contains
// the point is filed in the point's cell at its level, so it is in the
// bucket visited there, once; items of other cells that share the bucket
// either do not contain the point or are that same record.
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

// The bucket of a cell key at `level`: the nb buckets are split into
// 2^ceil(log2(occupied levels)) equal ranges, one per occupied level in
// ascending order (levelMask: the occupied levels), and the key hashed
// into its level's range.
RTX_DI uint32_t icBucket(uint64_t key, uint32_t nb, uint32_t levelMask, int level) {
    // splitmix64 finaliser
    key ^= key >> 30;
    key *= 0xbf58476d1ce4e5b9ull;
    key ^= key >> 27;
    key *= 0x94d049bb133111ebull;
    key ^= key >> 31;
    const uint32_t nl = (uint32_t)__popc(levelMask);
    const int partBits = nl <= 1u ? 0 : 32 - __clz((int)(nl - 1u));
    const uint32_t per = nb >> partBits;
    const uint32_t rank = (uint32_t)__popc(levelMask & ((1u << level) - 1u));
    return rank * per + ((uint32_t)key & (per - 1u));
}

// The level a sphere of grown radius m is filed under: the first whose cells
// are at least 1.001 m wide, so that its box (2m wide) spans at most 3 cells
// per axis whatever the rounding of its bounds.
RTX_DI int icLevel(float c0, float m) {
    int level = 0;
    float cs = c0;
    while (cs < 1.001f * m && level < kIcMaxLevel) {
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
RTX_DI uint32_t icLevelBucket(const DevIC &ic, float c0, uint32_t levelMask, int level, V3 x) {
    const float cs = icCellSize(c0, level);
    const uint64_t key = icKey(level, icCellCoord(x.x, cs), icCellCoord(x.y, cs), icCellCoord(x.z, cs));
    return icBucket(key, ic.nb, levelMask, level);
}

// The fields of a record a lookup weighs with (gradients only when used).
struct IcRec {
    V3 normal, color, rotGrad, transGrad;
    float harmonicR;
};

// From the item's second and third 16 bytes, and the entry's gradients.
RTX_DI IcRec icLoadRec(const DevIC &ic, const IcItem *item, bool gradients) {
    const float4 a = reinterpret_cast<const float4 *>(item)[1], b = reinterpret_cast<const float4 *>(item)[2];
    IcRec r;
    r.normal = v3(a);
    r.harmonicR = a.w;
    r.color = v3(b);
    if (gradients) {
        const rtx_cache_data &e = ic.entries[__float_as_uint(b.w)];
        r.rotGrad = v3(e.rotGrad);
        r.transGrad = v3(e.transGrad);
    } else {
        r.rotGrad = r.transGrad = v3(0.0f);
    }
    return r;
}

// One record's term of the weighted mean (raytrace.irradiance.rint:17-20,
// raytrace.irradiance.rahit:18-51) for a point x inside its sphere
// (oc = x - centre).
RTX_DI void icWeigh(const IcQueryParams &qp, V3 n, V3 oc, const IcRec &e, V3 &sum, float &total) {
    const V3 ni = e.normal;
    float w = 1.0f / (length(oc) / e.harmonicR + sqrtf(1.0f - dot(n, ni)));
    if (isnan(w) || isinf(w)) w = 1000000.0f;
    if (w <= qp.invA) return;
    if (qp.visibility && !(-0.001f <= dot(oc, (n + ni) / 2.0f))) return;
    V3 col = e.color;
    if (qp.gradients) {
        float E = length(col);
        V3 cn = E != 0.0f ? normalize(col) : v3(0.0f);
        col = cn * (E + dot(cross(ni, n), e.rotGrad) + dot(oc, e.transGrad));
    }
    sum = sum + w * col;
    total += w;
}

// Items read per step of a lookup: their spheres' loads, and then the loads
// of the records among them that contain the point, go out together.
constexpr int kIcChunk = 4;

RTX_DI void icLoadChunk(const DevIC &ic, uint32_t it, uint32_t end, float4 (&c)[kIcChunk]) {
#pragma unroll
    for (int k = 0; k < kIcChunk; ++k)
        if (it + k < end) c[k] = *reinterpret_cast<const float4 *>(&ic.items[it + k]);
}

// Weighted cache colour at x (normal n); false when no record accepts x.
// Records are weighed in (level, bucket order) order.  The loads are
// arranged so that a step of kIcChunk items costs about one memory latency:
// the next level's bucket range is fetched while this level's items are
// read, a chunk's records while its containment tests run, and the next
// chunk's items while the records are weighed.
RTX_DI bool icQuery(const DevIC &ic, const IcQueryParams &qp, V3 x, V3 n, V3 &color) {
    const uint32_t mask = ic.meta[0];
    uint32_t lm = mask;
    if (lm == 0) return false;
    const float c0 = __uint_as_float(ic.meta[2]);
    V3 sum = v3(0.0f);
    float total = 0.0f;
    uint32_t b = icLevelBucket(ic, c0, mask, __ffs(lm) - 1, x);
    uint32_t it = ic.bucketStart[b], end = ic.bucketStart[b + 1];
    for (;;) {
        lm &= lm - 1;
        uint32_t nit = 0, nend = 0;
        if (lm != 0) {
            const uint32_t nbk = icLevelBucket(ic, c0, mask, __ffs(lm) - 1, x);
            nit = ic.bucketStart[nbk];
            nend = ic.bucketStart[nbk + 1];
        }
        float4 cur[kIcChunk];
        icLoadChunk(ic, it, end, cur);
        while (it < end) {
            bool in[kIcChunk];
            V3 oc[kIcChunk];
            IcRec rec[kIcChunk];
#pragma unroll
            for (int k = 0; k < kIcChunk; ++k) {
                in[k] = it + k < end;
                if (in[k]) {
                    oc[k] = x - v3(cur[k]);
                    in[k] = length(oc[k]) <= cur[k].w;
                }
                if (in[k]) rec[k] = icLoadRec(ic, &ic.items[it + k], qp.gradients);
            }
            it += kIcChunk;
            icLoadChunk(ic, it, end, cur);
#pragma unroll
            for (int k = 0; k < kIcChunk; ++k)
                if (in[k]) icWeigh(qp, n, oc[k], rec[k], sum, total);
        }
        if (lm == 0) break;
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
