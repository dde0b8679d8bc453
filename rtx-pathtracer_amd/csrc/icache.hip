// Irradiance-cache bookkeeping on the device (gfx950).
//
// The reference runs updateIrradianceCache at the start and
// createIrradianceCache at the end of every pixel's invocation
// (shaders/raytrace.rgen:1637-1639, 1819-1825), racing on the header's
// atomics, and refits the cache's acceleration structure on the host after
// the frame (src/RayTracingApp.cpp:116-121).  Here one frame (one 1-spp
// sample plane) reads the cache as it was at its start; the frame's work is
// collected and done afterwards, deterministically:
//   kIcFrameBegin  - zero the frame's counters, note the free slots
//   kIcCand*       - ordered compaction of the creation candidates of a
//                    batch (pixel order, then path order; the first
//                    free-slot-many are kept)
//   kIcPlan        - the update elections served in order (the header's
//                    nextUpdateSlot rule), then the kept candidates: the jobs
//   (the jobs' 200 stratum sub-paths run through the wavefront kernels)
//   kIcReduce      - calculateCacheData's accumulation per job
//   kIcCommit      - updates and creations written in job order
//   kIcIndex*      - rebuild of the lookup index (dev_icache.h)
#include <hip/hip_runtime.h>


#include "dev_icache.h"
#include "dev_math.h"
#include "dev_shading.h"
#include "dev_sort.h"
#include "internal.h"

namespace rtx {

constexpr float kHalfPi = 3.1415926535897932384626433832795f / 2.0f;

// Exclusive scan of one value per thread over a block of BLOCK threads.
template <int BLOCK> RTX_DI uint32_t blockScanExclusive(uint32_t v, uint32_t *sh, uint32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < BLOCK / 64; ++w) {
            uint32_t t = sh[w];
            sh[w] = acc;
            acc += t;
        }
        sh[BLOCK / 64] = acc;
    }
    __syncthreads();
    uint32_t r = sh[wave] + x - v;
    total = sh[BLOCK / 64];
    __syncthreads();
    return r;
}

__global__ void kIcFrameBegin(IcBuffers ic) {
    if (threadIdx.x == 0) {
        uint32_t count = min(ic.hdr->nextCacheSlot, ic.hdr->maxCaches);
        ic.frame[0] = 0;
        ic.frame[1] = 0;
        ic.frame[2] = ic.maxEntries - count;
        ic.frame[3] = 0;
        ic.frame[4] = 0;
    }
}

RTX_DI uint32_t candCount(const PathBuffers &pb, uint32_t pid, uint32_t n) {
    if (pid >= n) return 0;
    return (__float_as_uint(pb.thr[pid].w) & kFlagCandMask) >> kFlagCandShift;
}

__global__ __launch_bounds__(256) void kIcCandCount(PathBuffers pb, uint32_t n, uint32_t *blockSums) {
    __shared__ uint32_t sh[256 / 64 + 1];
    uint32_t total;
    blockScanExclusive<256>(candCount(pb, blockIdx.x * 256 + threadIdx.x, n), sh, total);
    if (threadIdx.x == 0) blockSums[blockIdx.x] = total;
}

// Exclusive scan of the batch's block sums on top of the frame's running
// candidate count; one block.
__global__ __launch_bounds__(1024) void kIcCandScan(uint32_t *blockSums, uint32_t nblk, uint32_t *frame) {
    __shared__ uint32_t sh[1024 / 64 + 1];
    uint32_t carry = frame[1];
    for (uint32_t base = 0; base < nblk; base += 1024) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < nblk ? blockSums[i] : 0u;
        uint32_t total;
        uint32_t ex = blockScanExclusive<1024>(v, sh, total);
        if (i < nblk) blockSums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) frame[1] = carry;
}

__global__ __launch_bounds__(256) void kIcCandScatter(PathBuffers pb, uint32_t n, const uint32_t *blockOffsets,
                                                      IcBuffers ic) {
    __shared__ uint32_t sh[256 / 64 + 1];
    const uint32_t pid = blockIdx.x * 256 + threadIdx.x;
    const uint32_t c = candCount(pb, pid, n);
    uint32_t total;
    uint32_t pos = blockOffsets[blockIdx.x] + blockScanExclusive<256>(c, sh, total);
    const uint32_t room = ic.frame[2];
    for (uint32_t k = 0; k < c; ++k, ++pos) {
        if (pos >= room) break;
        const float4 *src = pb.cand + 2 * ((size_t)pid * kMaxNewCacheEntries + k);
        ic.cand[2 * pos] = src[0];
        ic.cand[2 * pos + 1] = src[1];
    }
}

// The jobs of the frame: update elections served in order by the header's
// nextUpdateSlot rule (updateIrradianceCache, raytrace.rgen:1334-1345; at
// most maxEntries update jobs a frame), then the kept candidates.  `part`
// (IcPlanPart) plans both, or only the updates (frame[3] = frame[4] = their
// count), or only the creations behind the frame[3] updates planned before.
__global__ __launch_bounds__(256) void kIcPlan(IcBuffers ic, int part) {
    __shared__ uint32_t numUpd;
    if (part == kIcPlanCreations) {
        if (threadIdx.x == 0) numUpd = ic.frame[3];
    } else if (threadIdx.x == 0) {
        const uint32_t elections = ic.frame[0];
        const uint32_t count = min(ic.hdr->nextCacheSlot, ic.hdr->maxCaches);
        uint32_t u = 0;
        // with no records every election only resets the slot counter
        if (count == 0 && elections > 0) ic.hdr->nextUpdateSlot = 0;
        for (uint32_t e = 0; count > 0 && e < elections && u < ic.maxEntries; ++e) {
            uint32_t idx = ic.hdr->nextUpdateSlot++;
            if (idx >= count) {
                ic.hdr->nextUpdateSlot = 0;
                continue;
            }
            if (ic.entries[idx].numUpdates < 1) continue;
            ic.jobs[2 * u] = f4(v3(ic.spheres[idx].center), __int_as_float(0));
            ic.jobs[2 * u + 1] = f4(v3(ic.entries[idx].normal), __uint_as_float(idx));
            ++u;
        }
        numUpd = u;
        if (part == kIcPlanUpdates) ic.frame[3] = ic.frame[4] = u;
    }
    __syncthreads();
    if (part == kIcPlanUpdates) return;
    const uint32_t u = numUpd;
    const uint32_t c = min(ic.frame[1], ic.frame[2]);
    for (uint32_t i = threadIdx.x; i < c; i += 256) {
        float4 o = ic.cand[2 * i], nn = ic.cand[2 * i + 1];
        ic.jobs[2 * (u + i)] = make_float4(o.x, o.y, o.z, __int_as_float(1));
        ic.jobs[2 * (u + i) + 1] = make_float4(nn.x, nn.y, nn.z, __uint_as_float(0));
    }
    if (threadIdx.x == 0) {
        ic.frame[3] = u;
        ic.frame[4] = u + c;
    }
}

RTX_DI V3 sphericalToCartesian(float theta, float phi) { // shaders/transform.glsl:41-43
    return v3(sinf(theta) * cosf(phi), sinf(theta) * sinf(phi), cosf(theta));
}

// calculateCacheData's accumulation (raytrace.rgen:1246-1315) over the 200
// sub-path results (radiance, first-hit distance in w) of each job of the
// sub-batch, in the reference's loop order.
__global__ __launch_bounds__(64) void kIcReduce(IcBuffers ic, uint32_t frameSeed, uint32_t firstJob,
                                                uint32_t numJobs, PathBuffers pb, int jobRange, uint32_t outBase) {
    // jobRange (IcJobRange): the jobs, from device memory unless kIcJobsHost;
    // job firstJob + t's sub-path results are at pb.out[outBase + 200 t ..]
    if (jobRange != kIcJobsHost) {
        uint32_t j0, jEnd;
        icJobRange(ic, jobRange, firstJob, j0, jEnd);
        firstJob = j0;
        numJobs = min(numJobs, jEnd > j0 ? jEnd - j0 : 0u);
    }
    // One 64-lane block per job: the lanes compute every stratum's terms
    // (trigonometry, neighbour differences) in parallel, then lane 0 adds
    // them in the loop order of calculateCacheData, so the sums are the
    // sequential loop's.
    __shared__ float4 smp[kIcStrata];                            // sample radiance, w = hit distance
    __shared__ float rotT[3 * kIcStrata], ta[3 * kIcStrata], tb[3 * kIcStrata];
    const uint32_t t = blockIdx.x;
    if (t >= numJobs) return; // uniform per block
    const uint32_t job = firstJob + t;
    const V3 normal = v3(ic.jobs[2 * job + 1]);
    const float4 *sub = pb.out + outBase + (size_t)t * kIcStrata;
    for (int q = (int)threadIdx.x; q < kIcStrata; q += 64) smp[q] = sub[q];
    __syncthreads();
    for (int q = (int)threadIdx.x; q < kIcStrata; q += 64) {
        const int k = q / kIcStrataTheta, j = q % kIcStrataTheta;
        Rng col{tea(job * kIcStrataPhi + (uint32_t)k, frameSeed ^ kIcPhiKey)};
        float phi = 2.0f * kPi * ((float)k + col.next()) / (float)kIcStrataPhi;
        V3 uk = toWorld(sphericalToCartesian(kHalfPi, phi), normal);
        V3 vk = toWorld(sphericalToCartesian(kHalfPi, phi + kHalfPi), normal);
        V3 prevVk = toWorld(sphericalToCartesian(kHalfPi, 2.0f * kPi * (float)k / (float)kIcStrataPhi + kHalfPi),
                            normal);
        Rng rs{tea((job * kIcStrataPhi + (uint32_t)k) * kIcStrataTheta + (uint32_t)j, frameSeed ^ kIcStratumKey)};
        float theta = asinf(sqrtf(((float)j + rs.next()) / (float)kIcStrataTheta));
        const float4 o = smp[q];
        const float r = o.w;
        const float L = length(v3(o));
        float prevJTheta = asinf(sqrtf((float)j / (float)kIcStrataTheta));
        float nextJTheta = asinf(sqrtf((float)(j + 1) / (float)kIcStrataTheta));
        float tanTheta = tanf(theta);
        if (isinf(tanTheta) || isnan(tanTheta)) tanTheta = 0.0f;
        V3 rt = tanTheta * L * vk, a = v3(0.0f), c = v3(0.0f);
        if (j > 0) {
            const float4 pj = smp[q - 1]; // stratum (k, j - 1)
            float cosPrev = cosf(prevJTheta);
            a = uk * 2.0f * kPi / (float)kIcStrataPhi * sinf(prevJTheta) * cosPrev * cosPrev / fminf(r, pj.w) *
                (L - length(v3(pj)));
        }
        if (k > 0) {
            const float4 pk = smp[q - kIcStrataTheta]; // stratum (k - 1, j)
            c = prevVk * (sinf(nextJTheta) - sinf(prevJTheta)) / fminf(r, pk.w) * (L - length(v3(pk)));
        }
        rotT[3 * q] = rt.x, rotT[3 * q + 1] = rt.y, rotT[3 * q + 2] = rt.z;
        ta[3 * q] = a.x, ta[3 * q + 1] = a.y, ta[3 * q + 2] = a.z;
        tb[3 * q] = c.x, tb[3 * q + 1] = c.y, tb[3 * q + 2] = c.z;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    float invDistanceSum = 0.0f;
    int numDistances = 0;
    V3 color = v3(0.0f), rot = v3(0.0f), trans = v3(0.0f);
    for (int q = 0; q < kIcStrata; ++q) {
        const int k = q / kIcStrataTheta, j = q % kIcStrataTheta;
        const float4 o = smp[q];
        color = color + v3(o);
        if (o.w < kTMax) {
            invDistanceSum += 1.0f / o.w;
            numDistances++;
        }
        rot = rot - V3{rotT[3 * q], rotT[3 * q + 1], rotT[3 * q + 2]};
        if (j > 0) trans = trans + V3{ta[3 * q], ta[3 * q + 1], ta[3 * q + 2]};
        if (k > 0) trans = trans + V3{tb[3 * q], tb[3 * q + 1], tb[3 * q + 2]};
    }
    const float normFactor = kPi / (float)(kIcStrataTheta * kIcStrataPhi);
    color = normFactor * color;
    rot = rot * normFactor;
    float harmonicR = (invDistanceSum == 0.0f || numDistances == 0) ? -1.0f
                                                                    : 1.0f / (invDistanceSum / (float)numDistances);
    float *res = ic.jobRes + 10 * (size_t)job;
    res[0] = color.x;
    res[1] = color.y;
    res[2] = color.z;
    res[3] = rot.x;
    res[4] = rot.y;
    res[5] = rot.z;
    res[6] = trans.x;
    res[7] = trans.y;
    res[8] = trans.z;
    res[9] = harmonicR;
}

RTX_DI void clampGradients(float maxLen, V3 &rot, V3 &trans) { // raytrace.rgen:1318-1329
    float lr = length(rot);
    if (lr > maxLen) rot = rot * (maxLen / lr);
    float lt = length(trans);
    if (lt > maxLen) trans = trans * (maxLen / lt);
}
RTX_DI V3 mixv(V3 x, V3 y, float a) { return x * (1.0f - a) + y * a; }
RTX_DI void put3(float *d, V3 v) {
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
}

// Updates (raytrace.rgen:1347-1380) and creations (1383-1421, slots kept
// below maxCaches) in job order.
RTX_DI void commitJobs(IcBuffers ic, const FrameConst *__restrict__ fcp) {
    const rtx_push_constant &pc = fcp->pc;
    const uint32_t numJobs = ic.frame[4];
    for (uint32_t j = 0; j < numJobs; ++j) {
        const float4 o4 = ic.jobs[2 * j], n4 = ic.jobs[2 * j + 1];
        const float *r = ic.jobRes + 10 * (size_t)j;
        V3 color = v3(r[0], r[1], r[2]), rot = v3(r[3], r[4], r[5]), trans = v3(r[6], r[7], r[8]);
        float harmonicR = r[9];
        if (__float_as_int(o4.w) == 0) {
            const uint32_t idx = __float_as_uint(n4.w);
            rtx_cache_data &e = ic.entries[idx];
            float a = fminf((float)e.numUpdates / (float)(e.numUpdates + 1), 0.95f);
            put3(e.color, mixv(color, v3(e.color), a));
            rot = mixv(rot, v3(e.rotGrad), a);
            trans = mixv(trans, v3(e.transGrad), a);
            clampGradients(pc.irradianceGradientsMaxLength, rot, trans);
            put3(e.rotGrad, rot);
            put3(e.transGrad, trans);
            float h = harmonicR > 0.0f ? harmonicR * (1.0f - a) + e.harmonicR * a : e.harmonicR;
            h = fmaxf(h, pc.irradianceCacheMinRadius);
            e.harmonicR = h;
            ic.spheres[idx].radius = pc.irradianceA * h;
            e.numUpdates += 1;
        } else {
            if (harmonicR < 0.0f) continue;
            const uint32_t slot = ic.hdr->nextCacheSlot;
            if (slot >= ic.hdr->maxCaches) continue;
            ic.hdr->nextCacheSlot = slot + 1;
            float h = fmaxf(harmonicR, pc.irradianceCacheMinRadius);
            clampGradients(pc.irradianceGradientsMaxLength, rot, trans);
            rtx_cache_data &e = ic.entries[slot];
            put3(e.normal, v3(n4));
            put3(e.color, color);
            e.harmonicR = h;
            put3(e.rotGrad, rot);
            put3(e.transGrad, trans);
            e.numUpdates = 1;
            rtx_sphere &sp = ic.spheres[slot];
            put3(sp.center, v3(o4));
            sp.radius = pc.irradianceA * h;
        }
    }
}

__global__ void kIcCommit(IcBuffers ic, const FrameConst *__restrict__ fcp) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    commitJobs(ic, fcp);
    if (ic.fullHost) *reinterpret_cast<volatile uint32_t *>(ic.fullHost) = ic.hdr->nextCacheSlot >= ic.hdr->maxCaches;
}

// The freshly created cache (IrradianceCache::createBuffers,
// src/IrradianceCache.cpp:46-75): no records, harmonicR -1, spheres of
// radius 0 at the origin.
__global__ __launch_bounds__(256) void kIcInit(IcBuffers ic) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0) {
        ic.hdr->nextCacheSlot = 0;
        ic.hdr->maxCaches = ic.maxEntries;
        ic.hdr->nextUpdateSlot = 0;
        ic.meta[0] = 0;
        ic.meta[1] = 0;
    }
    if (i >= ic.maxEntries) return;
    rtx_cache_data e;
    for (int a = 0; a < 3; ++a) e.color[a] = e.normal[a] = e.rotGrad[a] = e.transGrad[a] = 0.0f;
    e.harmonicR = -1.0f;
    e.numUpdates = 0;
    ic.entries[i] = e;
    rtx_sphere sp;
    sp.center[0] = sp.center[1] = sp.center[2] = 0.0f;
    sp.radius = 0.0f;
    sp.materialIndex = 0;
    sp.iLight = 0;
    ic.spheres[i] = sp;
}

// ------------------------------------------------------------ lookup index
__global__ __launch_bounds__(256) void kIcIndexClear(IcBuffers ic, float c0) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < ic.nb; i += gridDim.x * 256) {
        ic.bucketCount[i] = 0;
        ic.cursor[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ic.meta[0] = 0;
        ic.meta[1] = 0;
        ic.meta[2] = __float_as_uint(c0);
    }
}

// The cells of record i at its level: a box of the sphere grown by a small
// margin, so that every point the exact containment test accepts lies in one
// of them despite rounding; at most 3 cells per axis.
RTX_DI void icCells(const rtx_sphere &sp, float c0, int &level, int32_t lo[3], int32_t hi[3]) {
    const float r = sp.radius;
    const float m = r * 1.0001f + 1e-6f * (1.0f + fmaxf(fabsf(sp.center[0]), fmaxf(fabsf(sp.center[1]),
                                                                                  fabsf(sp.center[2]))));
    level = icLevel(c0, m);
    const float cs = icCellSize(c0, level);
    for (int a = 0; a < 3; ++a) {
        lo[a] = icCellCoord(sp.center[a] - m, cs);
        hi[a] = icCellCoord(sp.center[a] + m, cs);
        if (hi[a] - lo[a] > 2) hi[a] = lo[a] + 2; // only when r exceeds the top level's reach
    }
}

RTX_DI bool icFiled(const IcBuffers &ic, uint32_t i, uint32_t count, rtx_sphere &sp) {
    if (i >= count) return false;
    sp = ic.spheres[i];
    return sp.radius > 0.0f;
}

// The occupied levels (meta[0]), which the bucket ranges depend on.
__global__ __launch_bounds__(256) void kIcIndexLevels(IcBuffers ic, float c0) {
    const uint32_t count = min(ic.hdr->nextCacheSlot, ic.hdr->maxCaches);
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    rtx_sphere sp;
    if (i >= ic.maxEntries || !icFiled(ic, i, count, sp)) return;
    int level;
    int32_t lo[3], hi[3];
    icCells(sp, c0, level, lo, hi);
    atomicOr(&ic.meta[0], 1u << level);
}

// Pass 1 counts the (record, bucket) items per bucket; pass 2 writes record
// i's bucket keys into its 27 fixed slots in cell order (unused slots keyed
// nb), so that a stable sort by bucket leaves each bucket's items in record
// order and lookups add the same records in the same order every time.  A
// cell whose bucket an earlier cell of the record already has files nothing.
template <bool SCATTER> __global__ __launch_bounds__(256) void kIcIndexFill(IcBuffers ic, float c0) {
    const uint32_t count = min(ic.hdr->nextCacheSlot, ic.hdr->maxCaches);
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ic.maxEntries) return;
    const uint32_t mask = ic.meta[0];
    uint32_t k = 0;
    uint32_t filed[27];
    rtx_sphere sp;
    if (icFiled(ic, i, count, sp)) {
        int level;
        int32_t lo[3], hi[3];
        icCells(sp, c0, level, lo, hi);
        for (int32_t x = lo[0]; x <= hi[0]; ++x)
            for (int32_t y = lo[1]; y <= hi[1]; ++y)
                for (int32_t z = lo[2]; z <= hi[2]; ++z) {
                    const uint32_t b = icBucket(icKey(level, x, y, z), ic.nb, mask, level);
                    bool dup = false;
                    for (uint32_t q = 0; q < k; ++q) dup |= filed[q] == b;
                    if (dup) continue;
                    filed[k] = b;
                    if (!SCATTER) atomicAdd(&ic.bucketCount[b], 1u);
                    else ic.slotKeys[27 * (size_t)i + k] = b;
                    ++k;
                }
    }
    if (SCATTER) {
        for (; k < 27; ++k) ic.slotKeys[27 * (size_t)i + k] = ic.nb;
        for (uint32_t q = 0; q < 27; ++q) ic.slotIdx[27 * (size_t)i + q] = 27u * i + q;
    }
}

// items[j] = the item of the j-th slot in bucket order (record order[j] / 27),
// for the meta[1] items in use.
__global__ __launch_bounds__(256) void kIcIndexGather(IcBuffers ic, const uint32_t *order) {
    const uint32_t n = ic.meta[1];
    for (uint32_t j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) {
        const uint32_t r = order[j] / 27u;
        const rtx_sphere sp = ic.spheres[r];
        const rtx_cache_data &e = ic.entries[r];
        IcItem it;
        for (int a = 0; a < 3; ++a) {
            it.center[a] = sp.center[a];
            it.normal[a] = e.normal[a];
            it.color[a] = e.color[a];
        }
        it.radius = sp.radius;
        it.harmonicR = e.harmonicR;
        it.entry = r;
        it.pad[0] = it.pad[1] = it.pad[2] = it.pad[3] = 0u;
        ic.items[j] = it;
    }
}

// The bucket starts, an exclusive scan of the counts, in three passes over
// blocks of 1024 buckets: each block's total (kIcBucketSums), a scan of the
// totals (kScanInPlace), and each block's own scan plus its offset
// (kIcBucketStarts, whose last block also writes the item count).
__global__ __launch_bounds__(1024) void kIcBucketSums(IcBuffers ic) {
    __shared__ uint32_t sh[1024 / 64 + 1];
    const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
    uint32_t total;
    blockScanExclusive<1024>(i < ic.nb ? ic.bucketCount[i] : 0u, sh, total);
    if (threadIdx.x == 0) ic.blockSums[blockIdx.x] = total;
}
__global__ __launch_bounds__(1024) void kIcBucketStarts(IcBuffers ic) {
    __shared__ uint32_t sh[1024 / 64 + 1];
    const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
    uint32_t total;
    const uint32_t ex = blockScanExclusive<1024>(i < ic.nb ? ic.bucketCount[i] : 0u, sh, total);
    const uint32_t base = ic.blockSums[blockIdx.x];
    if (i < ic.nb) ic.bucketStart[i] = base + ex;
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        ic.bucketStart[ic.nb] = base + total;
        ic.meta[1] = base + total;
        ic.meta[3] = 27u * min(ic.hdr->nextCacheSlot, ic.hdr->maxCaches); // slots of the records in use
    }
}


// ------------------------------------------------------------- launchers
static int blocksFor(size_t n, int block) { return (int)((n + block - 1) / block); }

void launchIcFrameBegin(hipStream_t s, IcBuffers ic) {
    hipLaunchKernelGGL(kIcFrameBegin, dim3(1), dim3(64), 0, s, ic);
}
void launchIcCollect(hipStream_t s, IcBuffers ic, PathBuffers pb, uint32_t numPaths) {
    if (numPaths == 0) return;
    const uint32_t nblk = (uint32_t)blocksFor(numPaths, 256);
    hipLaunchKernelGGL(kIcCandCount, dim3(nblk), dim3(256), 0, s, pb, numPaths, ic.blockSums);
    hipLaunchKernelGGL(kIcCandScan, dim3(1), dim3(1024), 0, s, ic.blockSums, nblk, ic.frame);
    hipLaunchKernelGGL(kIcCandScatter, dim3(nblk), dim3(256), 0, s, pb, numPaths, ic.blockSums, ic);
}
void launchIcPlan(hipStream_t s, IcBuffers ic, int part) {
    hipLaunchKernelGGL(kIcPlan, dim3(1), dim3(256), 0, s, ic, part);
}
void launchIcReduce(hipStream_t s, IcBuffers ic, uint32_t frameSeed, uint32_t firstJob, uint32_t numJobs,
                    PathBuffers pb, int jobRange, uint32_t outBase) {
    if (numJobs == 0) return;
    hipLaunchKernelGGL(kIcReduce, dim3(numJobs), dim3(64), 0, s, ic, frameSeed, firstJob, numJobs, pb, jobRange,
                       outBase);
}
void launchIcCommit(hipStream_t s, IcBuffers ic, const FrameConst *fc) {
    hipLaunchKernelGGL(kIcCommit, dim3(1), dim3(64), 0, s, ic, fc);
}
void launchIcInit(hipStream_t s, IcBuffers ic) {
    hipLaunchKernelGGL(kIcInit, dim3(blocksFor(ic.maxEntries, 256)), dim3(256), 0, s, ic);
}
void launchIcIndex(hipStream_t s, IcBuffers ic, float c0) {
    hipLaunchKernelGGL(kIcIndexClear, dim3(blocksFor(ic.nb, 256)), dim3(256), 0, s, ic, c0);
    hipLaunchKernelGGL(kIcIndexLevels, dim3(blocksFor(ic.maxEntries, 256)), dim3(256), 0, s, ic, c0);
    hipLaunchKernelGGL(kIcIndexFill<false>, dim3(blocksFor(ic.maxEntries, 256)), dim3(256), 0, s, ic, c0);
    const uint32_t nblk = (ic.nb + 1023) / 1024;
    hipLaunchKernelGGL(kIcBucketSums, dim3(nblk), dim3(1024), 0, s, ic);
    hipLaunchKernelGGL(kScanInPlace<0>, dim3(1), dim3(1024), 0, s, ic.blockSums, nblk, (uint32_t *)nullptr);
    hipLaunchKernelGGL(kIcBucketStarts, dim3(nblk), dim3(1024), 0, s, ic);
    hipLaunchKernelGGL(kIcIndexFill<true>, dim3(blocksFor(ic.maxEntries, 256)), dim3(256), 0, s, ic, c0);
    // stable LSD radix sort of (bucket, slot) over the slots of the records in
    // use (meta[3]; keys below 2^(bucketBits + 1): nb itself marks unused
    // slots, which sort last), then the items gathered in that order
    uint32_t *keysA = ic.slotKeys, *valsA = ic.slotIdx, *keysB = ic.slotKeysB, *valsB = ic.slotIdxB;
    launchRadixSortPairs(s, keysA, valsA, keysB, valsB, ic.meta + 3, ic.itemCap, (int)ic.bucketBits + 1, ic.sortHist);
    hipLaunchKernelGGL(kIcIndexGather, dim3(blocksFor(ic.itemCap, 256)), dim3(256), 0, s, ic, valsA);
}

} // namespace rtx
