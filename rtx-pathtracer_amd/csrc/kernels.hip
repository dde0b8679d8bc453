// Wavefront path-tracing kernels for CDNA4 (gfx950).
//
// The reference runs one recursive megakernel per pixel (raygen main(),
// shaders/raytrace.rgen:1623-1827, bounce loop raytrace() 992-1226).  Here
// one bounce of every live path is one pass of five kernels:
//   extend  - closest-hit traversal of the queued rays
//   shade   - hit shader + emission + NEE sampling + continuation sampling,
//             appends shadow / probe / next-extend / done records with
//             wave-ballot + prefix-sum compaction
//   shadow  - any-hit traversal of the NEE shadow rays
//   probe   - closest-hit traversal of the MIS probe rays + their MIS weight
//   finish  - adds the last pending NEE terms of terminated paths
// Every path owns its LCG stream, so results do not depend on queue order.
#include <hip/hip_runtime.h>

#include "dev_guiding.h"
#include "dev_icache.h"
#include "dev_math.h"
#include "dev_shading.h"
#include "dev_traverse.h"
#include "internal.h"

namespace rtx {

constexpr int kTraceBlock = 128;
constexpr int kLdsStack = 20;
constexpr int kShadeBlock = 256;

// --------------------------------------------------------- wave helpers
// Append `pred` lanes to a queue: one atomic per wave, order-preserving
// within the wave (ballot + popcount prefix).  Must be called with the
// whole wave converged.
RTX_DI int waveAppend(bool pred, uint32_t *counter) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return -1;
    int lane = (int)__lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    uint64_t below = lane == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane)));
    return pred ? (int)(base + (uint32_t)__popcll(below)) : -1;
}

// Two appends whose atomics are in flight together: the slots of `a` lanes
// in queue ca and of `b` lanes in queue cb (-1 for the others).
RTX_DI void waveAppend2(bool a, uint32_t *ca, bool b, uint32_t *cb, int &ia, int &ib) {
    const uint64_t ma = __ballot(a), mb = __ballot(b);
    const int lane = (int)__lane_id();
    const int la = ma ? __ffsll((unsigned long long)ma) - 1 : 0, lb = mb ? __ffsll((unsigned long long)mb) - 1 : 0;
    uint32_t basea = 0, baseb = 0;
    if (ma && lane == la) basea = atomicAdd(ca, (uint32_t)__popcll(ma));
    if (mb && lane == lb) baseb = atomicAdd(cb, (uint32_t)__popcll(mb));
    basea = __shfl(basea, la);
    baseb = __shfl(baseb, lb);
    const uint64_t below = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    ia = a ? (int)(basea + (uint32_t)__popcll(ma & below)) : -1;
    ib = b ? (int)(baseb + (uint32_t)__popcll(mb & below)) : -1;
}

// Runs body(k) once for every distinct value k of `key` among the lanes
// with `active`, with exactly the lanes holding that key active and k
// wave-uniform (read with readlane, so per-key data are scalar loads).  The
// lanes of a pass are picked by a ballot mask rather than by comparing key
// with k inside the branch: the compiler would otherwise substitute the
// lane's own key for k there and drop the loop.
template <class F> RTX_DI void forEachKey(bool active, uint32_t key, F &&body) {
    uint64_t todo = __ballot(active);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
        const uint64_t m = __ballot(active && key == k);
        if ((m >> __lane_id()) & 1ull) body(k);
        todo &= ~m;
    }
}

RTX_DI uint32_t waveSum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

template <bool COUNT> RTX_DI void flushVisits(const VisitCount &vc, Counters *cnt, int kind) {
    if (!COUNT) return;
    uint32_t n = waveSum(vc.nodes), p = waveSum(vc.prims);
    if (__lane_id() == 0) {
        atomicAdd(&cnt->nodeVisits[kind], (unsigned long long)n);
        atomicAdd(&cnt->primVisits[kind], (unsigned long long)p);
    }
}

// ------------------------------------------------------ guiding records
// updateSamples (raytrace.rgen:973-978): light reaching a vertex is added to
// the light sums of the running segment's earlier records (slots lo..base-1),
// through their throughputs.
RTX_DI void updateSamples(const PathBuffers &pb, uint32_t pid, int lo, int base, V3 light) {
    float4 *ls = pb.lightSum + (size_t)pid * kMaxRecords;
    const float4 *st = pb.sampleThr + (size_t)pid * kMaxRecords;
    for (int i = base - 1; i >= lo; --i) {
        float4 a = ls[i];
        ls[i] = f4(v3(a) + light, 0.0f);
        light = light * v3(st[i]);
    }
}

RTX_DI rtx_directional_data *recordSlots(const PathBuffers &pb, uint32_t pid) {
    return pb.dirData + (size_t)(pb.dirFirstPixel + pid) * kMaxRecords;
}

RTX_DI bool isMatAlmostDiscrete(const Mat &m) { // raytrace.rgen:962-964
    return (m.type == RTX_MAT_ROUGH_CONDUCTOR && m.roughness <= 0.3f) ||
           (m.type == RTX_MAT_PHONG && m.highlight >= 250.0f);
}

// getCameraRay (raytrace.rgen:1487-1494): a jittered pinhole ray, two draws.
RTX_DI void cameraRay(const FrameConst &fc, uint32_t pixel, Rng &rng, V3 &o, V3 &d) {
    uint32_t px = pixel % fc.width, py = pixel / fc.width;
    float jx = rng.negPos();
    float jy = rng.negPos();
    float ux = ((float)px + 0.5f + jx / 2.0f) / (float)fc.width;
    float uy = ((float)py + 0.5f + jy / 2.0f) / (float)fc.height;
    float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
    o = xform(fc.cam.viewInverse, v3(0.0f), 1.0f);
    V3 target = xform(fc.cam.projInverse, v3(dx, dy, 1.0f), 1.0f);
    d = normalize(xform(fc.cam.viewInverse, normalize(target), 0.0f));
}

// ------------------------------------------------ learned radiance cache
// The training slot of path pid, or -1 for a path that ends into the cache:
// pixels with (pixel + offset) % every == 0 train, the offset moving with
// the sample so that every pixel trains now and then.
RTX_DI int nrcSlot(const FrameConst &fc, BatchDesc b, uint32_t pid) {
    const uint64_t g = b.firstPath + (pid - b.pathBase);
    const uint32_t pixel = (uint32_t)(g % b.npix);
    const uint64_t sample = g / b.npix + fc.sampleOffset;
    const uint32_t every = fc.nrcTrainEvery;
    const uint32_t k = pixel + (uint32_t)((sample * 7919u) % every);
    if (k % every != 0u) return -1;
    const uint32_t slot = k / every;
    return slot < fc.nrcTrainSlots ? (int)slot : -1;
}
// The cache's surface inputs: roughness 1 for Lambert, alpha for rough
// conductors, sqrt(2 / (n + 2)) for Phong lobes (the Beckmann width of a
// Phong exponent), 0 for specular; albedo the reflectance colour.
RTX_DI void nrcSurface(const Mat &m, float &rough, V3 &albedo) {
    switch (m.type) {
    case RTX_MAT_DIFFUSE:
    case RTX_MAT_LIGHT:
        rough = 1.0f;
        albedo = m.kd;
        break;
    case RTX_MAT_PHONG:
        rough = sqrtf(2.0f / (m.highlight + 2.0f));
        albedo = m.kd + m.ks;
        break;
    case RTX_MAT_ROUGH_CONDUCTOR:
        rough = m.roughness;
        albedo = v3(fresnelConductor(1.0f, m.eta, m.k));
        break;
    default:
        rough = 0.0f;
        albedo = m.specular;
    }
}

// State of a path at the start of a raytrace() call from a camera sample:
// depth 0, followCount 0, follow = true, addNextDirectLights = addFirstHitLight = true.
constexpr uint32_t kFlagsCameraStart = kFlagFollow | kFlagAddNext;

// ------------------------------------------------------------ generate
// main() up to the camera ray: seed = tea(pixel, randomUInt) (raytrace.rgen:1625),
// the irradiance-cache update election (1637-1639, the stream's first draw),
// the first camera sample; randomUInt per frame from PCG32.
__global__ __launch_bounds__(256) void kGenerate(const FrameConst *__restrict__ fcp, BatchDesc b, PathBuffers pb,
                                                 Queue q0, Counters *cnt, int qsel, uint32_t *electCounter) {
    const FrameConst &fc = *fcp;
    uint32_t pid = blockIdx.x * 256 + threadIdx.x;
    if (pid == 0) {
        cnt->extend[qsel] = b.numPaths;
        if (fc.nrc) *pb.nrcQueryCount = 0u;
    }
    const bool active = pid < b.numPaths;
    uint64_t g = b.firstPath + pid;
    uint32_t pixel = (uint32_t)(g % b.npix);
    uint64_t sample = g / b.npix + fc.sampleOffset;
    uint32_t frameSeed = pcg32First(fc.seed, sample);
    Rng rng{tea(pixel, frameSeed)};
    bool elect = false;
    if (fc.pc.useIrradianceCache) elect = rng.next() < fc.pc.irradianceUpdateProb;
    if (fc.pc.useIrradianceCache) {
        // only the number of elections matters: they are served in order
        uint64_t m = __ballot(active && elect);
        if (m && __lane_id() == __ffsll((unsigned long long)m) - 1) atomicAdd(electCounter, (uint32_t)__popcll(m));
    }
    if (!active) return;
    V3 o, d;
    cameraRay(fc, pixel, rng, o, d);
    q0.id[pid] = (int)pid;
    q0.org[pid] = f4(o, kTMin);
    q0.dir[pid] = f4(d, kTMax);
    pb.rng[pid] = rng.s;
    pb.hash[pid] = 2166136261u;
    pb.frameSeed[pid] = frameSeed;
    if (fc.record) { // resetSamplesToInvalid, raytrace.rgen:1615-1621
        rtx_directional_data *dd = recordSlots(pb, pid);
        for (int k = 0; k < kMaxRecords; ++k) dd[k].flags = kDirInvalid;
        pb.grec[pid] = make_int4(0, -1, __float_as_int(1.0f), 0);
    }
    if (fc.segments) {
        pb.task[pid] = make_int4((int)fc.samplesPerFrame - 1, 0, 0, 0);
        pb.taskRes[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    pb.thr[pid] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(kFlagsCameraStart));
    pb.result[pid] = make_float4(0.0f, 0.0f, 0.0f, kTMax);
    if (fc.nrc) {
        const int slot = nrcSlot(fc, b, pid);
        if (slot >= 0) pb.nrcTrainTarget[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f); // no target until the path ends
    }
}

// The paths of one serving slot: main() up to the camera ray, as kGenerate
// (one camera sample per path, no records), at pool ids pathBase + local,
// appended to the extend queue qsel (the next bounce's).  With the cache on
// the stream's first draw is the update election, which the serving pool
// does not act on (its caches are read-only) but draws to keep the stream
// the reference's.
__global__ __launch_bounds__(256) void kGenerateM(const SlotDesc *__restrict__ slots, uint32_t slot, PathBuffers pb,
                                                  Queue q0, Counters *cnt, int qsel) {
    const SlotDesc &sd = slots[slot];
    const FrameConst &fc = sd.fc;
    const BatchDesc b = sd.b;
    const uint32_t local = blockIdx.x * 256 + threadIdx.x;
    const bool active = local < b.numPaths;
    const uint32_t pid = b.pathBase + local;
    V3 o = v3(0.0f), d = v3(0.0f, 0.0f, 1.0f);
    if (active) {
        const uint64_t g = b.firstPath + local;
        const uint32_t pixel = (uint32_t)(g % b.npix);
        const uint64_t sample = g / b.npix + fc.sampleOffset;
        const uint32_t frameSeed = pcg32First(fc.seed, sample);
        Rng rng{tea(pixel, frameSeed)};
        if (fc.pc.useIrradianceCache) (void)rng.next();
        cameraRay(fc, pixel, rng, o, d);
        pb.rng[pid] = rng.s;
        pb.hash[pid] = 2166136261u;
        pb.frameSeed[pid] = frameSeed;
        pb.thr[pid] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(kFlagsCameraStart));
        pb.result[pid] = make_float4(0.0f, 0.0f, 0.0f, kTMax);
    }
    const int qi = waveAppend(active, &cnt->extend[qsel]);
    if (active) {
        q0.id[qi] = (int)pid;
        q0.org[qi] = f4(o, kTMin);
        q0.dir[qi] = f4(d, kTMax);
    }
}

// Writes slot `slot`'s descriptor and its live-path count (the batch's
// paths); passed by value so that the host's copy needs no lifetime.
__global__ __launch_bounds__(256) void kInstallSlot(SlotDesc sd, SlotDesc *slots, uint32_t slot, uint32_t *alive) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&sd);
    uint32_t *dst = reinterpret_cast<uint32_t *>(slots + slot);
    for (uint32_t k = threadIdx.x; k < sizeof(SlotDesc) / 4; k += 256) dst[k] = src[k];
    if (threadIdx.x == 0) alive[slot] = sd.b.numPaths;
}

// The sub-paths of calculateCacheData (raytrace.rgen:1246-1263) for jobs
// firstJob.. of the frame's job list: path p is stratum (k, j) = (s / 10,
// s % 10), s = p % 200, of job firstJob + p / 200.  phi comes from the
// column's stream, theta from the stratum's, whose state then drives the
// sub-path; raytrace(..., addFirstHitLight = false, ...) starts with
// addNextDirectLights off.
__global__ __launch_bounds__(256) void kGenerateIC(IcBuffers ic, uint32_t frameSeed, uint32_t firstJob,
                                                   uint32_t numPaths, PathBuffers pb, Queue q0, Counters *cnt,
                                                   int qsel, int jobRange, uint32_t pidBase, uint32_t qBase) {
    // jobRange: which jobs of the frame's list, from device memory when not
    // kIcJobsHost (see IcJobRange); numPaths bounds the paths either way
    uint32_t j0 = firstJob;
    if (jobRange != kIcJobsHost) {
        uint32_t jEnd;
        icJobRange(ic, jobRange, firstJob, j0, jEnd);
        numPaths = min(numPaths, (jEnd > j0 ? jEnd - j0 : 0u) * (uint32_t)kIcStrata);
    }
    const uint32_t local = blockIdx.x * 256 + threadIdx.x;
    if (local == 0) cnt->extend[qsel] = qBase + numPaths; // after the qBase entries already queued
    if (local >= numPaths) return;
    const uint32_t pid = pidBase + local, qi = qBase + local;
    const uint32_t job = j0 + local / kIcStrata, st = local % kIcStrata;
    const uint32_t k = st / kIcStrataTheta, j = st % kIcStrataTheta;
    float4 o4 = ic.jobs[2 * job], n4 = ic.jobs[2 * job + 1];
    V3 n = v3(n4);
    Rng col{tea(job * kIcStrataPhi + k, frameSeed ^ kIcPhiKey)};
    float phi = 2.0f * kPi * ((float)k + col.next()) / (float)kIcStrataPhi;
    Rng rng{tea((job * kIcStrataPhi + k) * kIcStrataTheta + j, frameSeed ^ kIcStratumKey)};
    float theta = asinf(sqrtf(((float)j + rng.next()) / (float)kIcStrataTheta));
    V3 local3 = v3(sinf(theta) * cosf(phi), sinf(theta) * sinf(phi), cosf(theta));
    V3 d = toWorld(local3, n);
    q0.id[qi] = (int)pid;
    q0.org[qi] = f4(v3(o4), kTMin);
    q0.dir[qi] = f4(d, kTMax);
    pb.rng[pid] = rng.s;
    pb.hash[pid] = 2166136261u;
    pb.frameSeed[pid] = frameSeed;
    pb.thr[pid] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(kFlagFollow));
    pb.result[pid] = make_float4(0.0f, 0.0f, 0.0f, kTMax);
}

// The frame's randomUInt for queue entry i, when the any-hit alpha test needs it.
RTX_DI uint32_t raySeed(const DevScene &S, const Queue &q, uint32_t i) {
    return S.alphaTris ? q.pathSeed[((uint32_t)q.id[i] & 0x7fffffffu) / q.seedDiv] : 0u;
}

// The work a bounce starts with (its extend queue, the previous bounce's
// probe, shadow and done lists), written to the host-visible slot `poll`
// (may be NULL) so that the host can end the bounce loop without stopping
// the stream.
RTX_DI void reportBounce(uint32_t *poll, uint32_t ne, uint32_t np, uint32_t ns, uint32_t nd) {
    if (!poll) return;
    poll[0] = ne;
    poll[1] = np;
    poll[2] = ns;
    poll[3] = nd;
}

// MIS probe result (raytrace.rgen:684-726): on a miss the environment colour
// (black without a map) with pdf 1/(2 pi lightCount); on a light the
// light-side pdf (pdfLight, raytrace.rgen:334-339) and the BSDF-side weight.
// base = throughput * f * cos / (numNEE * pdfMat), base.w = pdfMat.
RTX_DI V3 probeContribution(const DevScene &S, bool power, V3 o, V3 d, int prim, float t, float bu, float bv,
                            float4 base, bool &emitter) {
    float pdfMat = base.w;
    emitter = false;
    if (prim < 0) {
        V3 lightColor = envColor(S, normalize(d));
        float pdfL = kInv2Pi / (float)S.lightCount;
        float h = misWeight(power, pdfMat, pdfL);
        return v3(base) * lightColor * h;
    }
    HitInfo hi = hitInfo(S, prim, t, bu, bv, o, d);
    const rtx_material &ms = S.materials[hi.matIndex];
    if (ms.type != RTX_MAT_LIGHT) return v3(0.0f);
    emitter = true;
    int iLight = hi.isSphere ? S.spheres[hi.instanceIndex].iLight : S.instances[hi.instanceIndex].iLight;
    const rtx_light &light = S.lights[iLight];
    float cosL = dot(-d, hi.normal);
    float pdfL = light.sampleProb * t * t / cosL / light.area;
    float h = misWeight(power, pdfMat, pdfL);
    if (isnan(h)) return v3(0.0f);
    return v3(base) * v3(ms.lightColor) * h;
}

// The nearest emitter along (o, d) within [tmin, kTMax]: the emissive
// spheres and S.lightTris, with the closest-hit traversal's rule (least t,
// ties to the smaller primitive id) and its intersection tests on the same
// ray set-up, so the hit is the one the traversal would report if nothing
// else came first.  Returns the primitive id (tHit, bu, bv set) or -1.
RTX_DI int nearestLight(const DevScene &S, V3 o, V3 d, float tmin, float &tHit, float &bu, float &bv) {
    const TravRay r = makeTravRay(o, d);
    int best = -1;
    float tmax = kTMax;
    bu = bv = 0.0f;
    for (uint32_t k = 0; k < S.numSpheres; ++k) {
        const rtx_sphere &sp = S.spheres[k];
        if (S.materials[sp.materialIndex].type != RTX_MAT_LIGHT) continue;
        float t;
        if (isectSphere(sp, r.o, r.d, tmin, tmax, t)) {
            const int id = (int)(S.numTris + k);
            if (best < 0 || t < tmax || id < best) {
                tmax = t;
                best = id;
            }
        }
    }
    for (uint32_t k = 0; k < S.numLightTris; ++k) {
        const float4 *tp = S.lightTris + 3 * (size_t)k;
        float t, u, v;
        if (isectTri(r, tp[0], tp[1], tp[2], tmin, tmax, t, u, v)) {
            const int id = __float_as_int(tp[0].w) & ~kAlphaBit;
            if (best < 0 || t < tmax || id < best) {
                tmax = t;
                best = id;
                bu = u;
                bv = v;
            }
        }
    }
    tHit = tmax;
    return best;
}

// --------------------------------------------------- persistent traversal
// One launch per bounce traces three queues: the bounce's extend rays
// (closest hit into qe.hit), then the previous bounce's MIS probe rays
// (closest hit into qp.hit), then its shadow rays (first hit; visibility into
// pendShadow).  Ray i of the launch is extend ray i for i < ne, probe ray
// i - ne below ne + np, shadow ray i - ne - np after that.  The closest-hit
// rays come first so that the launch's tail is made of short occlusion rays,
// and a bounce pays one launch and one tail.  Work counter: fetch[1].
// COUNT builds count node and triangle visits per queue (visit kinds 0
// extend, 1 shadow, 2 probe).  The kernels come in two builds: without the
// any-hit alpha test (scenes with no textured triangles, the hot path) and
// with it (kTracePA), so that the test's code stays out of the common loop.
template <bool COUNT, bool ALPHA, int LDS>
RTX_DI void traceBody(const DevScene &S, const Queue &qe, int esel, const Queue &qs, const Queue &qp,
                      const PathBuffers &pb, Counters *cnt, uint32_t *poll, int *stackMem) {
    const uint32_t ne = cnt->extend[esel], np = cnt->probe, ns = cnt->shadow;
    const uint32_t nep = ne + np;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&cnt->extendRays, (unsigned long long)ne);
        atomicAdd(&cnt->probeRays, (unsigned long long)np);
        atomicAdd(&cnt->shadowRays, (unsigned long long)ns);
        reportBounce(poll, ne, np, ns, cnt->done[esel ^ 1]);
        // the lists the bounce's kFinish and kShade append to: their previous
        // contents (two bounces back) are consumed
        cnt->extend[esel ^ 1] = 0;
        cnt->done[esel] = 0;
    }
    HybridStack<kTraceBlock, LDS> st;
    int spill[HybridStack<kTraceBlock, LDS>::kSpill];
    st.bind(stackMem + threadIdx.x, spill);
    VisitCount vc, vk[3];
    // the queues' arrays as plain pointers, read in per-queue branches (a
    // select among them becomes a table in scratch and flat loads)
    const float4 *eOrg = qe.org, *eDir = qe.dir, *pOrg = qp.org, *pDir = qp.dir, *sOrg = qs.org, *sDir = qs.dir;
    const int *eId = qe.id, *pId = qp.id, *sId = qs.id;
    float4 *eHit = qe.hit, *pHit = qp.hit;
    float4 *pendShadow = pb.pendShadow, *pendProbe = pb.pendProbe;
    const uint32_t *pathSeed = qe.pathSeed;
    const uint32_t secDiv = qs.seedDiv; // probe and shadow entries: path * numNEE + j
    persistentTrace<kHitLane, COUNT, ALPHA>(
        S, nep + ns, &cnt->fetch[1], st, vc,
        [&](uint32_t i, LaneRay &L) {
            const bool ext = i < ne, shadow = i >= nep;
            const uint32_t j = ext ? i : shadow ? i - nep : i - ne;
            float4 o, d;
            uint32_t id = 0;
            if (ext) {
                o = eOrg[j];
                d = eDir[j];
                if (ALPHA) id = (uint32_t)eId[j];
            } else if (shadow) {
                o = sOrg[j];
                d = sDir[j];
                if (ALPHA) id = (uint32_t)sId[j];
            } else {
                o = pOrg[j];
                d = pDir[j];
                if (ALPHA) id = (uint32_t)pId[j];
            }
            L.r = makeTravRay(v3(o), v3(d));
            L.tmin = o.w;
            L.tmax = d.w;
            L.any = shadow;
            if (ALPHA) {
                id &= ext ? 0x7fffffffu : (uint32_t)~kProbeOccBit;
                L.randomUInt = pathSeed[ext ? id : id / secDiv];
            }
        },
        [&](const LaneRay &L) {
            const bool ext = L.idx < ne;
            if (L.any) {
                if (L.best < 0) { // raytrace.shadow.rmiss:7-10
                    const int sid = sId[L.idx - nep];
                    if (sid & kProbeOccBit) { // a probe's light is visible: its pending record becomes the light's id
                        float4 &p = pendProbe[sid & ~kProbeOccBit];
                        p.w = __int_as_float(-3 - __float_as_int(p.w));
                    } else {
                        pendShadow[sid].w = 1.0f;
                    }
                }
            } else if (ext) {
                eHit[L.idx] = make_float4(L.tmax, __int_as_float(L.best), L.bu, L.bv);
            } else {
                pHit[L.idx - ne] = make_float4(L.tmax, __int_as_float(L.best), L.bu, L.bv);
            }
            if (COUNT) {
                const int k = ext ? 0 : L.any ? 1 : 2;
                for (int m = 0; m < 3; ++m) {
                    if (m == k) {
                        vk[m].nodes += vc.nodes;
                        vk[m].prims += vc.prims;
                    }
                }
                vc = VisitCount{};
            }
        });
    flushVisits<COUNT>(vk[0], cnt, 0);
    flushVisits<COUNT>(vk[1], cnt, 1);
    flushVisits<COUNT>(vk[2], cnt, 2);
}
// LDS: stack entries per thread held in LDS (the rest in scratch), a
// launch-shape choice (rtx_tuning.ldsStackEntries): more entries cost LDS and
// so resident blocks.
template <bool COUNT, int LDS = kLdsStack>
__global__ __launch_bounds__(kTraceBlock) void kTraceP(DevScene S, Queue qe, int esel, Queue qs, Queue qp,
                                                      PathBuffers pb, Counters *cnt, uint32_t *poll) {
    __shared__ int stackMem[LDS * kTraceBlock];
    traceBody<COUNT, false, LDS>(S, qe, esel, qs, qp, pb, cnt, poll, stackMem);
}
template <bool COUNT, int LDS = kLdsStack>
__global__ __launch_bounds__(kTraceBlock) void kTracePA(DevScene S, Queue qe, int esel, Queue qs, Queue qp,
                                                       PathBuffers pb, Counters *cnt, uint32_t *poll) {
    __shared__ int stackMem[LDS * kTraceBlock];
    traceBody<COUNT, true, LDS>(S, qe, esel, qs, qp, pb, cnt, poll, stackMem);
}

// One bounce of the serving pool: the three queues of every slot's paths in
// one persistent launch.  Each lane points its traversal at the scene of the
// slot its ray belongs to when it fetches the ray (pid >> slotShift; probe
// and shadow entries hold NEE slots, pid * neeStride + j).  Block 0 also
// copies every slot's live-path count after the bounce's queue lengths into
// the host-visible slot `poll` (4 + numSlots words).
template <bool ALPHA>
__global__ __launch_bounds__(kTraceBlock) void kTraceM(MultiArgs ma, Queue qe, int esel, Queue qs, Queue qp,
                                                      PathBuffers pb, Counters *cnt, uint32_t *poll) {
    __shared__ int stackMem[kLdsStack * kTraceBlock];
    const uint32_t ne = cnt->extend[esel], np = cnt->probe, ns = cnt->shadow;
    const uint32_t nep = ne + np;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&cnt->extendRays, (unsigned long long)ne);
        atomicAdd(&cnt->probeRays, (unsigned long long)np);
        atomicAdd(&cnt->shadowRays, (unsigned long long)ns);
        reportBounce(poll, ne, np, ns, cnt->done[esel ^ 1]);
        // the lists the bounce's kFinish and kShade append to: their previous
        // contents (two bounces back) are consumed
        cnt->extend[esel ^ 1] = 0;
        cnt->done[esel] = 0;
    }
    if (blockIdx.x == 0 && poll)
        for (uint32_t k = threadIdx.x; k < ma.numSlots; k += kTraceBlock) poll[4 + k] = ma.slotAlive[k];
    HybridStack<kTraceBlock, kLdsStack> st;
    int spill[HybridStack<kTraceBlock, kLdsStack>::kSpill];
    st.bind(stackMem + threadIdx.x, spill);
    VisitCount vc;
    LaneScene ls{};
    const float4 *eOrg = qe.org, *eDir = qe.dir, *pOrg = qp.org, *pDir = qp.dir, *sOrg = qs.org, *sDir = qs.dir;
    const int *eId = qe.id, *pId = qp.id, *sId = qs.id;
    float4 *eHit = qe.hit, *pHit = qp.hit;
    float4 *pendShadow = pb.pendShadow, *pendProbe = pb.pendProbe;
    const uint32_t *frameSeed = pb.frameSeed;
    persistentTrace<kHitLane, false, ALPHA>(
        ls, nep + ns, &cnt->fetch[1], st, vc,
        [&](uint32_t i, LaneRay &L) {
            const bool ext = i < ne, shadow = i >= nep;
            const uint32_t j = ext ? i : shadow ? i - nep : i - ne;
            float4 o, d;
            uint32_t id;
            if (ext) {
                o = eOrg[j];
                d = eDir[j];
                id = (uint32_t)eId[j];
            } else if (shadow) {
                o = sOrg[j];
                d = sDir[j];
                id = (uint32_t)sId[j];
            } else {
                o = pOrg[j];
                d = pDir[j];
                id = (uint32_t)pId[j];
            }
            const uint32_t pid = ext ? (id & ~kSplitVertex) : (id & ~(uint32_t)kProbeOccBit) / ma.neeStride;
            const DevScene &sc = ma.slots[pid >> ma.slotShift].S;
            pointAt(ls, sc, &sc);
            L.r = makeTravRay(v3(o), v3(d));
            L.tmin = o.w;
            L.tmax = d.w;
            L.any = shadow;
            if (ALPHA) L.randomUInt = frameSeed[pid];
        },
        [&](const LaneRay &L) {
            if (L.any) {
                if (L.best < 0) { // raytrace.shadow.rmiss:7-10
                    const int sid = sId[L.idx - nep];
                    if (sid & kProbeOccBit) { // a probe's light is visible
                        float4 &p = pendProbe[sid & ~kProbeOccBit];
                        p.w = __int_as_float(-3 - __float_as_int(p.w));
                    } else {
                        pendShadow[sid].w = 1.0f;
                    }
                }
            } else if (L.idx < ne) {
                eHit[L.idx] = make_float4(L.tmax, __int_as_float(L.best), L.bu, L.bv);
            } else {
                pHit[L.idx - ne] = make_float4(L.tmax, __int_as_float(L.best), L.bu, L.bv);
            }
        });
}

// The BSDF-side half of the MIS probe (raytrace.rgen:684-726) for probe rays
// whose closest hits the trace kernels wrote (probes that found their
// emitter first never enter this queue).
__global__ __launch_bounds__(256) void kProbeResolve(DevScene S, const FrameConst *__restrict__ fcp, Queue q,
                                                     PathBuffers pb, Counters *cnt, int qsel) {
    const uint32_t n = cnt->probe;
    const bool power = fcp->pc.usePowerHeuristic != 0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        float4 o4 = q.org[i], d4 = q.dir[i], base = q.aux[i], hit = q.hit[i];
        int prim = __float_as_int(hit.y);
        bool emitter;
        const V3 c = probeContribution(S, power, v3(o4), v3(d4), prim, hit.x, hit.z, hit.w, base, emitter);
        pb.pendProbe[q.id[i]] = f4(c, __int_as_float(emitter ? prim : -1));
    }
}

// The same over the serving pool: probe entry ids are NEE slots, path * neeStride + j.
__global__ __launch_bounds__(256) void kProbeResolveM(MultiArgs ma, Queue q, PathBuffers pb, Counters *cnt,
                                                      int qsel) {
    const uint32_t n = cnt->probe;
    for (uint32_t base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < n;
        const uint32_t id = active ? (uint32_t)q.id[i] : 0u;
        const uint32_t slot = (id / ma.neeStride) >> ma.slotShift;
        forEachKey(active, slot, [&](uint32_t k) {
            const SlotDesc &sd = ma.slots[k];
            float4 o4 = q.org[i], d4 = q.dir[i], base4 = q.aux[i], hit = q.hit[i];
            int prim = __float_as_int(hit.y);
            bool emitter;
            const V3 c = probeContribution(sd.S, sd.fc.pc.usePowerHeuristic != 0, v3(o4), v3(d4), prim, hit.x, hit.z,
                                           hit.w, base4, emitter);
            pb.pendProbe[id] = f4(c, __int_as_float(emitter ? prim : -1));
        });
    }
}

// Pending NEE terms of the previous vertex, in NEE-sample order.
// pendShadow.w: 0 no shadow ray, -1 traced and occluded, +1 traced and visible.
// pendProbe.w (int bits): -2 no probe ray; the id of the emitter the probe
// sees, or -1 when it sees none (its term is then zero or the environment's);
// -3 - id while the occlusion ray to emitter `id` has not shown it visible
// (the term does not count, and the path signature records -1).
// NEE slot j of path pid is pid * stride + j (stride >= numNEE: FrameConst::neeStride).
// (sh0, pr0): slot 0's records, which the caller may have fetched already.
RTX_DI V3 pendingSum(const PathBuffers &pb, uint32_t pid, uint32_t numNEE, uint32_t stride, uint32_t &hash,
                     float4 sh0, float4 pr0) {
    V3 s = v3(0.0f);
    for (uint32_t j = 0; j < numNEE; ++j) {
        float4 sh = j == 0 ? sh0 : pb.pendShadow[pid * stride + j];
        float4 pr = j == 0 ? pr0 : pb.pendProbe[pid * stride + j];
        if (sh.w != 0.0f) hash = sigFold(hash, sh.w > 0.0f ? kSigVisible : kSigOccluded);
        int probePrim = __float_as_int(pr.w);
        V3 term = v3(pr);
        if (probePrim <= -3) { // its light was occluded
            probePrim = -1;
            term = v3(0.0f);
        }
        if (probePrim != -2) hash = sigFold(hash, kSigProbe ^ (uint32_t)probePrim);
        if (sh.w > 0.0f) term = v3(sh) + term;
        s = s + term;
    }
    return s;
}
RTX_DI V3 pendingSum(const PathBuffers &pb, uint32_t pid, uint32_t numNEE, uint32_t stride, uint32_t &hash) {
    return pendingSum(pb, pid, numNEE, stride, hash, pb.pendShadow[pid * stride], pb.pendProbe[pid * stride]);
}

// ---------------------------------------------------------------- ADRRS
// applyWeightWindow (raytrace.rgen:830-869): window centre |estimate / adjoint|,
// lower = 2c/(1+s), upper = s*lower; roulette below with survival
// max(v/lower, 0.1), split above into q = v/upper copies (rounded up with
// probability q - floor(q), one draw).  Float-to-int saturates past INT_MAX.
RTX_DI float weightWindow(Rng &rng, V3 estimate, V3 adjoint, V3 thr, float s, int &n) {
    const float center = length(estimate / adjoint);
    const float lower = 2.0f * center / (1.0f + s);
    const float upper = s * lower;
    const float v = length(thr);
    n = 1;
    if (isnan(lower) || lower <= 0.0f) return 1.0f;
    if (lower <= v && v <= upper) return 1.0f;
    if (v <= lower) return fmaxf(v / lower, 0.1f);
    const float q = v / upper;
    n = q >= 2147483520.0f ? 2147483646 : (int)q;
    if (rng.next() > (float)(n + 1) - q) n++;
    return q;
}

// split (raytrace.rgen:883-902): queue slot task.y (the caller checks room).
RTX_DI void pushSplit(const PathBuffers &pb, uint32_t pid, int4 &task, const HitInfo &h, V3 wi, V3 thr, int depth) {
    float4 *sr = pb.splitRec + ((size_t)pid * kMaxSplits + (uint32_t)task.y) * 4;
    sr[0] = f4(h.worldPos, __int_as_float(depth));
    sr[1] = f4(h.normal, __uint_as_float((uint32_t)h.matIndex | (h.frontFace ? 0x80000000u : 0u)));
    sr[2] = f4(wi, h.u);
    sr[3] = f4(thr, h.v);
    task.y++;
}

// ----------------------------------------------------------------- shade
// Register budget: two waves per SIMD (at most 256 VGPRs and AGPRs
// together).  Without the bound the compiler reaches 254 VGPRs plus 4 AGPRs
// and the kernel drops to one wave per SIMD, which costs the 1080p cache
// workload about 16% (2.35 s per image instead of 2.74 s).
// One iteration of the bounce loop, raytrace.rgen:1008-1215, for the paths of
// the extend queue; an entry marked kSplitVertex resumes a queued split
// instead (raytrace.rgen:1684-1703: NEE at its vertex, a new direction, then
// the resumed path at the split's depth).
RTX_DI void shadeLane(const DevScene &S, const DevIC &IC, const DevGuiding &G, const FrameConst &fc,
                      const BatchDesc &b, const PathBuffers &pb, const Queue &qin, const Queue &qout, const Queue &qs,
                      const Queue &qp, int32_t *done, Counters *cnt, int qsel, uint32_t i, bool active) {
    const uint32_t numNEE = fc.numNEE;
    const float invNEE = 1.0f / (float)numNEE;
    const bool useNEE = fc.useNEE != 0;
    const bool addDirect = fc.addDirect != 0;
    const bool mis = fc.pc.enableMIS != 0;
    const bool power = fc.pc.usePowerHeuristic != 0;
    const int maxDepth = fc.maxDepth, maxFollow = fc.maxFollow;
    const bool useIC = fc.useIC != 0, createIC = fc.createIC != 0;
    const bool record = fc.record != 0;
    const bool icOnGlossy = fc.pc.useIrradianceCacheOnGlossy != 0;
    const bool segments = fc.segments != 0;
    const bool useADRRS = fc.useADRRS != 0, splitOnFirst = fc.splitOnFirst != 0;
    const bool adrrsSplit = fc.pc.adrrsSplit != 0;
    const float adrrsS = fc.pc.adrrsS;
    IcQueryParams icp;
    icp.invA = 1.0f / fc.pc.irradianceA;
    icp.gradients = fc.pc.useIrradianceGradients != 0;
    icp.visibility = fc.pc.irradianceCachePerformVisibilityCheck != 0;
    GuidingParams gp;
    gp.on = fc.pc.useGuiding != 0;
    gp.parallax = fc.pc.useParallaxCompensation != 0;
    gp.prob = fc.pc.guidingProb;
    uint32_t *nextCount = &cnt->extend[qsel ^ 1];
    bool cont = false, isDone = false, doNee = false;
    uint32_t pid = 0;
    Rng rng{0};
    V3 thr = v3(0.0f), res = v3(0.0f), d = v3(0.0f);
    uint32_t flags = 0, hash = 0;
    HitInfo h;
    Mat m;
    m.type = -1;
    int depth = 0, followCount = 0;
    bool follow = false, addNext = false, icHit = false;
    bool splitV = false, pendTask = false, dead = false, taskDirty = false, nrcQ = false;
    int4 task = make_int4(0, 0, 0, 0); // (camera samples left, splits queued, next split, record sampleOffset)
    uint32_t ncand = 0;
    int4 grec = make_int4(0, -1, 0, 0); // record state: slot, distance slot, distanceFactor bits, NEE base
    float firstT = kTMax;
    if (active) {
        const uint32_t raw = (uint32_t)qin.id[i];
        splitV = (raw & kSplitVertex) != 0;
        pid = raw & ~kSplitVertex;
        float4 hit = qin.hit[i];
        d = v3(qin.dir[i]);
        V3 o = v3(qin.org[i]);
        rng.s = pb.rng[pid];
        float4 t4 = pb.thr[pid];
        thr = v3(t4);
        flags = __float_as_uint(t4.w);
        float4 r4 = pb.result[pid];
        res = v3(r4);
        firstT = r4.w;
        ncand = (flags & kFlagCandMask) >> kFlagCandShift;
        hash = pb.hash[pid];
        if (record) grec = pb.grec[pid];
        if (segments) task = pb.task[pid];
        // slot 0 of the pending NEE records, fetched with the state rather than
        // after the flags that say whether it is needed (the buffers always hold it)
        const float4 sh0 = pb.pendShadow[pid * fc.neeStride], pr0 = pb.pendProbe[pid * fc.neeStride];
        // The hit's geometry and material depend on the queue entry alone:
        // their loads go out before the path state is used, so that the two
        // chains of dependent loads overlap instead of following each other.
        const int prim = __float_as_int(hit.y);
        if (!splitV && prim >= 0) {
            h = hitInfo(S, prim, hit.x, hit.z, hit.w, o, d);
            m = loadMat(&S.materials[h.matIndex]);
            applyTextures(S, m, h.u, h.v);
        }
        if (flags & kFlagPending) {
            V3 nee = pendingSum(pb, pid, numNEE, fc.neeStride, hash, sh0, pr0);
            if (flags & kFlagPendTask) { // a split vertex's NEE joins main()'s result
                float4 tr = pb.taskRes[pid];
                pb.taskRes[pid] = f4(v3(tr) + nee, tr.w);
            } else if (record) { // recorded NEE terms are kept without the throughput
                res = res + v3(pb.grecThr[pid]) * nee;
                updateSamples(pb, pid, task.w, grec.w, nee);
            } else {
                res = res + nee;
            }
        }
        if (splitV) {
            // the next queued split: its vertex, throughput and depth; raytrace(...,
            // addFirstHitLight = false, ...) from there (raytrace.rgen:1684-1703)
            const float4 *sr = pb.splitRec + ((size_t)pid * kMaxSplits + (uint32_t)task.z) * 4;
            const float4 a = sr[0], nb = sr[1], c = sr[2], t = sr[3];
            task.z++;
            taskDirty = true;
            const uint32_t mf = __float_as_uint(nb.w);
            h.worldPos = v3(a);
            h.normal = v3(nb);
            h.matIndex = (int)(mf & 0x7FFFFFFFu);
            h.frontFace = (mf >> 31) != 0;
            h.isSphere = false;
            h.instanceIndex = 0;
            h.u = c.w;
            h.v = t.w;
            d = -v3(c);
            thr = v3(t);
            m = loadMat(&S.materials[h.matIndex]);
            applyTextures(S, m, h.u, h.v);
            depth = __float_as_int(a.w);
            followCount = 0;
            follow = true;
            addNext = false;
            res = v3(0.0f);
            firstT = kTMax;
            doNee = useNEE && neeCapable(m.type) && S.lightCount > 0;
            pendTask = doNee;
        } else {
        hash = sigFold(hash, (uint32_t)prim);
        depth = (int)(flags & kFlagDepthMask) + 1;
        followCount = (int)((flags >> kFlagFollowCountShift) & 0xFFu);
        follow = (flags & kFlagFollow) != 0;
        addNext = (flags & kFlagAddNext) != 0;
        if (prim < 0) {
            // raytrace.rmiss: the environment map (black when the scene has none)
            if (addDirect || addNext) {
                const V3 missColor = envColor(S, normalize(d));
                res = res + thr * missColor;
                if (record) {
                    updateSamples(pb, pid, task.w, grec.x, missColor);
                    if (grec.y != -1) recordSlots(pb, pid)[grec.y].distance = 0.0f; // infinitely far
                }
            }
            isDone = true;
        } else {
            if (depth == 1) firstT = hit.x;
#ifdef RTX_DEVICE_DEBUG
            if (b.firstPath + (pid - b.pathBase) == fc.debugPath)
                printf("[gpu] depth %d prim %d t %.9g mat %d seed %u thr %.6g %.6g %.6g d %.6g %.6g %.6g n %.6g %.6g %.6g ff %d res %.6g\n",
                       depth, prim, hit.x, m.type, rng.s, thr.x, thr.y, thr.z, d.x, d.y, d.z, h.normal.x,
                       h.normal.y, h.normal.z, (int)h.frontFace, res.x);
#endif
            if (m.type == RTX_MAT_LIGHT && (addDirect || addNext)) {
                res = res + thr * m.lightColor;
                if (record) updateSamples(pb, pid, task.w, grec.x, m.lightColor);
            }
            if (hasDiscreteDirection(m.type)) {
                addNext = true;
                follow = true;
                if (depth >= maxDepth) followCount++;
                if (record && grec.y != -1) recordSlots(pb, pid)[grec.y].distance += hit.x * __int_as_float(grec.z);
            } else {
                addNext = false;
                follow = false;
                bool nee = useNEE;
                if (fc.nrc && depth >= (int)fc.nrcDepth && !(flags & kFlagNrcTrain)) {
                    const int slot = nrcSlot(fc, b, pid);
                    if (slot < 0) { // the path ends into the learned cache here, without NEE
                        nrcQ = true;
                        isDone = true;
                        nee = false;
                    } else { // a training path: what it adds from here on is this vertex's target
                        float rough;
                        V3 alb;
                        nrcSurface(m, rough, alb);
                        float4 *ti = pb.nrcTrainIn + 4 * (size_t)slot;
                        ti[0] = f4(h.worldPos, rough);
                        ti[1] = f4(-d, 0.0f);
                        ti[2] = f4(h.normal, 0.0f);
                        ti[3] = f4(alb, 0.0f);
                        pb.nrcPath[2 * (size_t)pid] = f4(thr, __int_as_float(slot));
                        pb.nrcPath[2 * (size_t)pid + 1] = f4(res, __int_as_float(depth));
                        flags |= kFlagNrcTrain;
                    }
                } else if (fc.nrcSuffix && (flags & kFlagNrcTrain) && !(flags & kFlagNrcTrainEnd)) {
                    // self-training: the suffix ends into the cache nrcSuffix vertices past the
                    // training vertex; the cache's prediction there completes the target
                    const int d0 = __float_as_int(pb.nrcPath[2 * (size_t)pid + 1].w);
                    if (depth >= d0 + (int)fc.nrcSuffix) {
                        nrcQ = true;
                        isDone = true;
                        nee = false;
                        flags |= kFlagNrcTrainEnd;
                    }
                }
                if ((useIC || (useADRRS && depth > 1)) && isICCapable(m.type, icOnGlossy)) { // raytrace.rgen:1080-1139
                    V3 irr;
                    if (icQuery(IC, icp, h.worldPos, h.normal, irr)) {
                        if (useADRRS) { // weight window against the cache's irradiance (1087-1116)
                            const uint32_t pixel = (uint32_t)((b.firstPath + (pid - b.pathBase)) % b.npix);
                            int nw;
                            float q = weightWindow(rng, v3(pb.estimate[pixel]), irr, thr, adrrsS, nw);
                            if (nw == 1) {
                                if (rng.next() > q) { // the path dies here, before its NEE
                                    hash = sigFold(hash, kSigRRKill);
                                    dead = true;
                                } else {
                                    thr = thr / q;
                                }
                            } else if (adrrsSplit) {
                                const int room = kMaxSplits - task.y;
                                if (nw - 1 > room) {
                                    nw = room + 1;
                                    q = (float)nw;
                                }
                                thr = thr / q;
                                for (int k = 0; k < nw - 1; ++k) pushSplit(pb, pid, task, h, -d, thr, depth);
                                taskDirty = true;
                                hash = sigFold(hash, kSigSplit ^ (uint32_t)(nw - 1));
                            }
                        } else {
                            // a cache record ends the path after its NEE (1114-1122)
                            hash = sigFold(hash, kSigCacheHit);
                            res = res + thr * approxDiffuse(m, h.normal, -d) * irr;
                            icHit = true;
                            nee = true;
                        }
                    } else if (createIC && rng.next() < fc.pc.irradianceCreateProb) {
                        if (ncand < (uint32_t)kMaxNewCacheEntries) {
                            float4 *c = pb.cand + 2 * ((size_t)pid * kMaxNewCacheEntries + ncand);
                            c[0] = f4(h.worldPos, 0.0f);
                            c[1] = f4(h.normal, 0.0f);
                            ncand++;
                        }
                    }
                }
                if (!dead && splitOnFirst && depth == 1 && task.y < kMaxSplits) { // raytrace.rgen:1142-1146
                    pushSplit(pb, pid, task, h, -d, thr * 0.5f, depth);
                    thr = thr * 0.5f;
                    taskDirty = true;
                    hash = sigFold(hash, kSigSplit ^ 1u);
                }
                if (record && !icHit && !dead && grec.y != -1) { // raytrace.rgen:1153-1160
                    recordSlots(pb, pid)[grec.y].distance += hit.x * __int_as_float(grec.z);
                    if (!isMatAlmostDiscrete(m)) grec.y = -1;
                }
                if (dead) {
                    isDone = true;
                    nee = false;
                }
                doNee = nee && neeCapable(m.type) && S.lightCount > 0;
                if (record) grec.w = grec.x; // the NEE light reaches the records made so far
            }
        }
        } // not a split vertex
    }
    // ---- next event estimation, raytrace.rgen:601-731 (numNEE samples, multipleNEE 871-877)
    // A probe ray (the BSDF-sampled half of MIS) adds only if the first thing
    // it hits is an emitter.  Where the scene has few emitters and no
    // environment, the probe finds its nearest emitter here and traces only
    // an occlusion ray to it, in the shadow queue, instead of a closest-hit
    // ray through the whole scene.
    const bool lightFirst = fc.probeFirstLight && S.lightFirstOk && S.env == nullptr;
    const V3 wi = -d;
    const V3 thrNee = thr; // the throughput NEE terms of this vertex carry
    for (uint32_t j = 0; j < numNEE; ++j) {
        bool wantShadow = false, wantProbe = false, wantOcc = false;
        V3 lightDir = v3(0.0f), bsdfDir = v3(0.0f), probeBase = v3(0.0f);
        float lightDist = 0.0f, pdfMatProbe = 0.0f, occMax = 0.0f;
        const uint32_t slot = pid * fc.neeStride + j;
        if (doNee) {
            V3 lightColor;
            float pdfL = sampleLights(S, rng, h.worldPos, h.normal, lightDir, lightColor, lightDist);
            float cosS = dot(h.normal, lightDir);
            bool traced = cosS > 0.0f && pdfL > 0.0f;
            bool probeAllowed = true;
            float shadowState = traced ? -1.0f : 0.0f; // occluded until the shadow kernel says otherwise
            V3 sh = v3(0.0f);
            if (traced) {
                if (mis) {
                    float pdfMat = pdfBSDF(m, h.normal, wi, lightDir);
                    float hw = misWeight(power, pdfL, pdfMat);
                    if (isnan(hw)) {
                        // The reference returns from NEE here if the shadow ray is
                        // unoccluded (raytrace.rgen:653-655), skipping the probe's draws:
                        // resolve visibility now to keep the random stream in step.
                        int spill[kStackCapacity];
                        PrivateStack ps(spill);
                        VisitCount vc0;
                        TravRay sr = makeTravRay(h.worldPos, lightDir);
                        float tm = lightDist * (1.0f - 0.0001f), bu, bv;
                        bool occl = traverse<true, false>(S, sr, kTMin, tm, bu, bv, ps, vc0, pb.frameSeed[pid]) >= 0;
                        traced = false;
                        shadowState = occl ? -1.0f : 1.0f; // resolved here; contributes nothing
                        if (!occl) probeAllowed = false;
                    } else {
                        sh = evalBsdf(m, h.normal, wi, lightDir, true) * lightColor * hw / pdfL;
                    }
                } else {
                    sh = evalBsdf(m, h.normal, wi, lightDir, true) * lightColor / pdfL;
                }
            }
            const V3 tw = record ? v3(1.0f) : thr;
            pb.pendShadow[slot] = f4(tw * sh * invNEE, shadowState);
            wantShadow = traced;
            bool probeZero = true;
            if (mis && probeAllowed) {
                float pdfMat = sampleBSDF(rng, m, wi, h.normal, h.frontFace, bsdfDir);
                if (pdfMat > 0.0f) {
                    wantProbe = true;
                    probeZero = false;
                    pdfMatProbe = pdfMat;
                    probeBase = (record ? v3(1.0f) : thr) * evalBsdf(m, h.normal, wi, bsdfDir, true) * invNEE /
                                pdfMat;
                }
            }
            if (probeZero) pb.pendProbe[slot] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-2));
            if (wantProbe && lightFirst) {
                float tl, lu, lv;
                const int lp = nearestLight(S, h.worldPos, bsdfDir, kTMin, tl, lu, lv);
                wantProbe = false;
                if (lp >= 0) {
                    bool emitter;
                    const V3 c = probeContribution(S, power, h.worldPos, bsdfDir, lp, tl, lu, lv,
                                                   f4(probeBase, pdfMatProbe), emitter);
                    pb.pendProbe[slot] = f4(c, __int_as_float(-3 - lp)); // counts once shown visible
                    wantOcc = true;
                    occMax = __int_as_float(__float_as_int(tl) - 1); // below the light: t < tl
                } else {
                    pb.pendProbe[slot] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
                }
            }
        }
        int si, pi;
        waveAppend2(wantShadow, &cnt->shadow, wantProbe, &cnt->probe, si, pi);
        if (lightFirst) {
            const int oi = waveAppend(wantOcc, &cnt->shadow);
            if (wantOcc) {
                qs.id[oi] = (int)slot | kProbeOccBit;
                qs.org[oi] = f4(h.worldPos, kTMin);
                qs.dir[oi] = f4(bsdfDir, occMax);
            }
        }
        if (wantShadow) {
            qs.id[si] = (int)slot;
            qs.org[si] = f4(h.worldPos, kTMin);
            qs.dir[si] = f4(lightDir, lightDist * (1.0f - 0.0001f));
        }
        if (wantProbe) {
            qp.id[pi] = (int)slot;
            qp.org[pi] = f4(h.worldPos, kTMin);
            qp.dir[pi] = f4(bsdfDir, kTMax);
            qp.aux[pi] = f4(probeBase, pdfMatProbe);
        }
    }
    if (fc.nrc) { // the learned cache's queries, in the order the waves reach them
        const int qn = waveAppend(nrcQ, pb.nrcQueryCount);
        if (nrcQ) {
            float rough;
            V3 alb;
            nrcSurface(m, rough, alb);
            float4 *q = pb.nrcQuery + 4 * (size_t)qn;
            q[0] = f4(h.worldPos, rough);
            q[1] = f4(-d, 0.0f);
            q[2] = f4(h.normal, 0.0f);
            q[3] = f4(alb, 0.0f);
            pb.nrcQueryAux[qn] = f4(thr, __uint_as_float(pid));
        }
    }
    // ---- continuation, raytrace.rgen:1169-1215
    V3 newDir = v3(0.0f);
    if (icHit) isDone = true;
    if (active && !isDone) {
        float pdf = getNewDirection(rng, G, gp, m, h.worldPos, h.normal, wi, h.frontFace, newDir);
#ifdef RTX_DEVICE_DEBUG
        if (b.firstPath + (pid - b.pathBase) == fc.debugPath) {
            V3 ev = evalBsdf(m, h.normal, wi, newDir, h.frontFace);
            printf("[gpu]   continue pdf %.9g eval %.9g new %.6g %.6g %.6g\n", pdf, ev.x, newDir.x, newDir.y, newDir.z);
        }
#endif
        // NaN guard beyond the reference (which stops on pdf <= 0 only): a NaN
        // pdf or BSDF ratio ends the path instead of poisoning the pixel.
        V3 change = v3(0.0f);
        if (!(pdf <= 0.0f) && !isnan(pdf)) change = evalBsdf(m, h.normal, wi, newDir, h.frontFace) / pdf;
        if (pdf <= 0.0f || isnan(pdf) || isnan(change.x) || isnan(change.y) || isnan(change.z)) {
            isDone = true;
            if (splitV) { // a split without a direction ends the frame's split loop (raytrace.rgen:1698-1700)
                task.z = task.y;
                taskDirty = true;
            }
        } else if (splitV) {
            thr = thr * change;
            cont = true; // raytrace() traces at least once
        } else {
            thr = thr * change;
            if (record && grec.x < kMaxRecords) { // raytrace.rgen:1179-1199
                rtx_directional_data &dd = recordSlots(pb, pid)[grec.x];
                if (hasDiscreteDirection(m.type) || isMatAlmostDiscrete(m)) {
                    dd.flags = kDirInvalid;
                } else {
                    dd.position[0] = h.worldPos.x;
                    dd.position[1] = h.worldPos.y;
                    dd.position[2] = h.worldPos.z;
                    dd.direction[0] = newDir.x;
                    dd.direction[1] = newDir.y;
                    dd.direction[2] = newDir.z;
                    dd.pdf = pdf;
                    dd.weight = 0.0f;
                    dd.distance = 0.0f;
                    dd.flags = (uint32_t)guidingRegion(G, h.worldPos);
                    grec.y = grec.x;
                    grec.z = __float_as_int(1.0f);
                }
                pb.lightSum[(size_t)pid * kMaxRecords + grec.x] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                pb.sampleThr[(size_t)pid * kMaxRecords + grec.x] = f4(change, 0.0f);
                grec.x++;
            } else if (record && grec.y != -1 && m.type == RTX_MAT_DIELECTRIC && dot(newDir, h.normal) < 0.0f) {
                // refraction: the next distance is the apparent one (raytrace.rgen:1200-1210)
                float eta = h.frontFace ? m.ior : m.iorInv;
                grec.z = __float_as_int(fabsf(dot(h.normal, d) / dot(h.normal, newDir)) * eta);
            }
            bool killed = false;
            if (fc.allowRR && fc.pc.enableRR && depth >= fc.rrStartDepth) { // Russian roulette (not in the reference)
                float q = fmaxf(fabsf(thr.x), fmaxf(fabsf(thr.y), fabsf(thr.z)));
                q = fminf(q, 0.95f);
                if (rng.next() >= q) killed = true;
                else thr = thr / q;
            }
            if (killed) {
                isDone = true;
            } else {
                cont = depth <= maxDepth || (follow && followCount <= maxFollow);
                isDone = !cont;
            }
        }
    }
    if (active) {
        uint32_t nf = (uint32_t)depth | ((uint32_t)followCount << kFlagFollowCountShift) |
                      (follow ? kFlagFollow : 0u) | (addNext ? kFlagAddNext : 0u) |
                      (doNee ? kFlagPending : 0u) | (doNee && pendTask ? kFlagPendTask : 0u) |
                      (ncand << kFlagCandShift) | (flags & (kFlagNrcTrain | kFlagNrcTrainEnd));
        if (taskDirty) pb.task[pid] = task;
        pb.rng[pid] = rng.s;
        pb.hash[pid] = hash;
        pb.thr[pid] = f4(thr, __uint_as_float(nf));
        pb.result[pid] = f4(res, firstT);
        if (record) {
            if (doNee) pb.grecThr[pid] = f4(thrNee, 0.0f); // the throughput the pending NEE terms carry
            pb.grec[pid] = grec;
        }
    }
    int qi, di;
    waveAppend2(cont, nextCount, isDone, &cnt->done[qsel], qi, di);
    if (cont) {
        qout.id[qi] = (int)pid;
        qout.org[qi] = f4(h.worldPos, kTMin);
        qout.dir[qi] = f4(newDir, kTMax);
    }
    if (isDone) done[di] = (int)pid;
}

__global__ __launch_bounds__(kShadeBlock, 2) void kShade(DevScene S, DevIC IC, DevGuiding G,
                                                      const FrameConst *__restrict__ fcp, BatchDesc b, PathBuffers pb,
                                                      Queue qin, Queue qout, Queue qs, Queue qp, int32_t *done,
                                                      Counters *cnt, int qsel) {
    const uint32_t n = cnt->extend[qsel];
    for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += gridDim.x * kShadeBlock) {
        const uint32_t i = base + threadIdx.x;
        shadeLane(S, IC, G, *fcp, b, pb, qin, qout, qs, qp, done, cnt, qsel, i, i < n);
    }
}

// Two classes of paths in one wavefront: ids below `split` are the class of
// (fc, b), the others of (fc2, b2) (a frame's camera paths and the cache's
// update sub-paths).  A wave takes one pass per class among its lanes.
__global__ __launch_bounds__(kShadeBlock, 2) void kShade2(DevScene S, DevIC IC, DevGuiding G,
                                                       const FrameConst *__restrict__ fcp, BatchDesc b,
                                                       const FrameConst *__restrict__ fcp2, BatchDesc b2,
                                                       uint32_t split, PathBuffers pb, Queue qin, Queue qout,
                                                       Queue qs, Queue qp, int32_t *done, Counters *cnt, int qsel) {
    const uint32_t n = cnt->extend[qsel];
    for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += gridDim.x * kShadeBlock) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < n;
        const uint32_t cls = active && ((uint32_t)qin.id[i] & ~kSplitVertex) >= split ? 1u : 0u;
        forEachKey(active, cls, [&](uint32_t k) {
            shadeLane(S, IC, G, k ? *fcp2 : *fcp, k ? b2 : b, pb, qin, qout, qs, qp, done, cnt, qsel, i, true);
        });
    }
}

// The same over the serving pool: each queue entry's path belongs to the
// slot pid >> slotShift, whose scene, cache, guide, frame constants and batch
// the lane shades with.  A wave takes one pass per slot among its lanes
// (forEachKey), so the slot's data are read with scalar loads as in the
// single-job kernel.
__global__ __launch_bounds__(kShadeBlock, 2) void kShadeM(MultiArgs ma, PathBuffers pb, Queue qin, Queue qout,
                                                       Queue qs, Queue qp, int32_t *done, Counters *cnt, int qsel) {
    const uint32_t n = cnt->extend[qsel];
    for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += gridDim.x * kShadeBlock) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < n;
        const uint32_t slot = active ? ((uint32_t)qin.id[i] & ~kSplitVertex) >> ma.slotShift : 0u;
        forEachKey(active, slot, [&](uint32_t k) {
            const SlotDesc &sd = ma.slots[k];
            shadeLane(sd.S, sd.ic, sd.g, sd.fc, sd.b, pb, qin, qout, qs, qp, done, cnt, qsel, i, true);
        });
    }
}

// ---------------------------------------------------------------- finish
// The end of a raytrace() call: the last pending NEE terms, commitSamples
// (raytrace.rgen:980-990, 1217-1222).  With segments the segment's result
// joins main()'s, and the path goes on with the frame's next camera sample
// (raytrace.rgen:1666-1681) or its next queued split (1684-1708), appended to
// the next iteration's extend queue; after the last one the pixel's frame
// result is main()'s sum over the camera samples (1710).  Runs in bounce
// qsel, before its kShade, on the paths the previous bounce's kShade ended
// (done list qsel ^ 1, whose NEE rays this bounce traced), appending to the
// next bounce's extend queue (qn, count extend[qsel ^ 1]).
// Returns whether the lane's path ended for good (its result is final).
RTX_DI bool finishLane(const FrameConst &fc, const BatchDesc &b, const PathBuffers &pb, const int32_t *done,
                       Counters *cnt, const Queue &qn, int qsel, uint32_t i, bool active) {
    const uint32_t numNEE = fc.numNEE;
    const bool record = fc.record != 0;
    const bool segments = fc.segments != 0;
    uint32_t *nextCount = &cnt->extend[qsel ^ 1];
    bool again = false, splitNext = false;
    uint32_t pid = 0;
    V3 o = v3(0.0f), dir = v3(0.0f, 0.0f, 1.0f);
    if (active) {
        pid = (uint32_t)done[i];
        float4 r4 = pb.result[pid];
        V3 res = v3(r4);
        uint32_t flags = __float_as_uint(pb.thr[pid].w);
        uint32_t hash = pb.hash[pid];
        int4 grec = record ? pb.grec[pid] : make_int4(0, -1, 0, 0);
        int4 task = segments ? pb.task[pid] : make_int4(0, 0, 0, 0);
        V3 tr = segments ? v3(pb.taskRes[pid]) : v3(0.0f);
        if (flags & kFlagPending) {
            V3 nee = pendingSum(pb, pid, numNEE, fc.neeStride, hash);
            if (flags & kFlagPendTask) {
                tr = tr + nee;
            } else if (record) {
                res = res + v3(pb.grecThr[pid]) * nee;
                updateSamples(pb, pid, task.w, grec.w, nee);
            } else {
                res = res + nee;
            }
        }
        if (flags & kFlagNrcTrain) { // the training vertex's target: what the path added after it
            const float4 tp = pb.nrcPath[2 * (size_t)pid], r0 = pb.nrcPath[2 * (size_t)pid + 1];
            const V3 add = res - v3(r0);
            const V3 t = v3(tp.x != 0.0f ? add.x / tp.x : 0.0f, tp.y != 0.0f ? add.y / tp.y : 0.0f,
                            tp.z != 0.0f ? add.z / tp.z : 0.0f);
            pb.nrcTrainTarget[__float_as_int(tp.w)] = f4(t, 1.0f);
        }
        if (record) { // commitSamples
            rtx_directional_data *dd = recordSlots(pb, pid);
            const float4 *ls = pb.lightSum + (size_t)pid * kMaxRecords;
            for (int k = task.w; k < grec.x; ++k) {
                float w = length(v3(ls[k])) / dd[k].pdf;
                if (w <= 0.0f) dd[k].flags = kDirInvalid;
                else dd[k].weight = w;
            }
            // a path without light leaves its slots to the next one; otherwise
            // the next starts at sampleOffset + currentSampleOffset
            if (length(res) > 0.0f) task.w += grec.x;
        }
        if (!segments) {
            pb.out[pid] = f4(res, r4.w);
            pb.outSig[pid] = make_uint2(pb.rng[pid], hash);
        } else {
            tr = tr + res; // result += raytrace(...)
            const uint32_t keepCand = flags & kFlagCandMask;
            Rng rng{pb.rng[pid]};
            if (task.x > 0) {
                task.x--;
                const uint32_t pixel = (uint32_t)((b.firstPath + (pid - b.pathBase)) % b.npix);
                cameraRay(fc, pixel, rng, o, dir);
                again = true;
                pb.rng[pid] = rng.s;
                pb.thr[pid] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(kFlagsCameraStart | keepCand));
                if (record) pb.grec[pid] = make_int4(task.w, -1, __float_as_int(1.0f), task.w);
            } else if (task.z < task.y) {
                again = splitNext = true;
                pb.thr[pid] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(keepCand));
            } else {
                pb.out[pid] = f4(tr / (float)fc.samplesPerFrame, r4.w);
                pb.outSig[pid] = make_uint2(rng.s, hash);
            }
            if (again) {
                pb.result[pid] = make_float4(0.0f, 0.0f, 0.0f, kTMax);
                pb.hash[pid] = hash;
            }
            pb.task[pid] = task;
            pb.taskRes[pid] = f4(tr, 0.0f);
        }
    }
    int qi = waveAppend(again, nextCount);
    if (again) {
        // a split's entry carries an empty ray (tmax < tmin): traversal finds nothing
        qn.id[qi] = (int)(splitNext ? (pid | kSplitVertex) : pid);
        qn.org[qi] = f4(o, splitNext ? 1.0f : kTMin);
        qn.dir[qi] = f4(dir, splitNext ? -1.0f : kTMax);
    }
    return active && !again;
}

// Clears the secondary queues' counters: they were traced and resolved, and
// kShade refills them.
RTX_DI void finishClears(Counters *cnt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cnt->shadow = 0;
        cnt->probe = 0;
        cnt->fetch[0] = cnt->fetch[1] = cnt->fetch[2] = 0;
    }
}

__global__ __launch_bounds__(256) void kFinish(const FrameConst *__restrict__ fcp, BatchDesc b, PathBuffers pb,
                                               const int32_t *done, Counters *cnt, Queue qn, int qsel) {
    const uint32_t n = cnt->done[qsel ^ 1];
    finishClears(cnt);
    for (uint32_t base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        const uint32_t i = base + threadIdx.x;
        (void)finishLane(*fcp, b, pb, done, cnt, qn, qsel, i, i < n);
    }
}

// kFinish over two classes of paths (kShade2).
__global__ __launch_bounds__(256) void kFinish2(const FrameConst *__restrict__ fcp, BatchDesc b,
                                                const FrameConst *__restrict__ fcp2, BatchDesc b2, uint32_t split,
                                                PathBuffers pb, const int32_t *done, Counters *cnt, Queue qn,
                                                int qsel) {
    const uint32_t n = cnt->done[qsel ^ 1];
    finishClears(cnt);
    for (uint32_t base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < n;
        const uint32_t cls = active && (uint32_t)done[i] >= split ? 1u : 0u;
        forEachKey(active, cls, [&](uint32_t k) { (void)finishLane(k ? *fcp2 : *fcp, k ? b2 : b, pb, done, cnt, qn, qsel, i, true); });
    }
}

// The same over the serving pool, one pass per slot among a wave's lanes;
// the paths that end leave their slot's live count (one atomic per pass).
__global__ __launch_bounds__(256) void kFinishM(MultiArgs ma, PathBuffers pb, const int32_t *done, Counters *cnt,
                                                Queue qn, int qsel) {
    const uint32_t n = cnt->done[qsel ^ 1];
    finishClears(cnt);
    for (uint32_t base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < n;
        const uint32_t slot = active ? (uint32_t)done[i] >> ma.slotShift : 0u;
        forEachKey(active, slot, [&](uint32_t k) {
            const SlotDesc &sd = ma.slots[k];
            const bool ended = finishLane(sd.fc, sd.b, pb, done, cnt, qn, qsel, i, true);
            const uint64_t m = __ballot(ended); // the lanes of slot k only
            if (m && (int)__lane_id() == __ffsll((unsigned long long)m) - 1)
                atomicSub(&ma.slotAlive[k], (uint32_t)__popcll(m));
        });
    }
}

// ------------------------------------------------------------ accumulate
// Adds each pixel's samples of this batch in increasing sample order, so
// the accumulated image does not depend on how samples are batched.
__global__ __launch_bounds__(256) void kAccumulate(BatchDesc b, PathBuffers pb, float4 *accum, float4 *keepRgb,
                                                   uint2 *keepSig, int keep, float4 *estimate) {
    uint32_t j = blockIdx.x * 256 + threadIdx.x;
    uint32_t lanes = b.numPaths < b.npix ? b.numPaths : b.npix;
    if (j >= lanes) return;
    uint32_t pixel = (uint32_t)((b.firstPath + j) % b.npix);
    float4 acc = accum ? accum[pixel] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t pid = j; pid < b.numPaths; pid += b.npix) {
        v = pb.out[pid];
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
        acc.w += 1.0f;
        if (keep) {
            keepRgb[b.firstPath + pid] = v;
            keepSig[b.firstPath + pid] = pb.outSig[pid];
        }
    }
    if (accum) accum[pixel] = acc;
    if (estimate) estimate[pixel] = make_float4(v.x, v.y, v.z, 1.0f); // saveEstimate, raytrace.rgen:1602-1605
}

__global__ void kResetCounters(Counters *cnt, int which) {
    if (threadIdx.x == 0) {
        cnt->extend[which] = 0;
        cnt->shadow = 0;
        cnt->probe = 0;
        cnt->done[0] = cnt->done[1] = 0;
        cnt->fetch[0] = cnt->fetch[1] = cnt->fetch[2] = 0;
    }
}

__global__ __launch_bounds__(256) void kFillU32(uint32_t *p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void kResolveImage(const float4 *accum, uint32_t npix, float inv, float4 *out) {
    uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npix) return;
    float4 a = accum[p];
    out[p] = make_float4(a.x * inv, a.y * inv, a.z * inv, 1.0f);
}

// ---------------------------------------------------------- ray queries
// The two ray queries (closest hit, occlusion) with the persistent lane-refill
// traversal of the wavefront.
template <bool ANY>
__global__ __launch_bounds__(kTraceBlock) void kQueryP(DevScene S, const rtx_ray *rays, uint32_t n, void *out,
                                                      uint32_t *fetch) {
    __shared__ int stackMem[kLdsStack * kTraceBlock];
    HybridStack<kTraceBlock, kLdsStack> st;
    int spill[HybridStack<kTraceBlock, kLdsStack>::kSpill];
    st.bind(stackMem + threadIdx.x, spill);
    VisitCount vc;
    persistentTrace<ANY ? kHitAny : kHitClosest, false, true>( // alpha-tested triangles with randomUInt 0
        S, n, fetch, st, vc,
        [&](uint32_t i, LaneRay &L) {
            rtx_ray ry = rays[i];
            L.r = makeTravRay(v3(ry.org), v3(ry.dir));
            L.tmin = ry.tmin;
            L.tmax = ry.tmax;
            L.randomUInt = 0;
        },
        [&](const LaneRay &L) {
            if (ANY) {
                static_cast<uint8_t *>(out)[L.idx] = L.best >= 0 ? 1 : 0;
            } else {
                rtx_hit h;
                h.t = L.tmax; // the ray's tmax when nothing was hit
                h.prim = L.best;
                h.u = L.bu;
                h.v = L.bv;
                static_cast<rtx_hit *>(out)[L.idx] = h;
            }
        });
}

// ------------------------------------------------------------- launchers
static int gridFor(size_t n, int block) { return (int)((n + block - 1) / block); }

void launchGenerate(hipStream_t s, const FrameConst *fc, BatchDesc b, PathBuffers pb, Queue q0, Counters *cnt,
                    int qsel, uint32_t *electCounter) {
    hipLaunchKernelGGL(kGenerate, dim3(gridFor(b.numPaths, 256)), dim3(256), 0, s, fc, b, pb, q0, cnt, qsel,
                       electCounter);
}
void launchGenerateIC(hipStream_t s, IcBuffers ic, uint32_t frameSeed, uint32_t firstJob, uint32_t numPaths,
                      PathBuffers pb, Queue q0, Counters *cnt, int qsel, int jobRange, uint32_t pidBase,
                      uint32_t qBase) {
    if (numPaths == 0) return;
    hipLaunchKernelGGL(kGenerateIC, dim3(gridFor(numPaths, 256)), dim3(256), 0, s, ic, frameSeed, firstJob, numPaths,
                       pb, q0, cnt, qsel, jobRange, pidBase, qBase);
}
template <int LDS>
static void launchTracePL(hipStream_t s, const DevScene &S, Queue qe, int esel, Queue qs, Queue qp, PathBuffers pb,
                          Counters *cnt, uint32_t *poll, bool countVisits, int blocks) {
    const dim3 g(blocks), t(kTraceBlock);
    if (S.alphaTris != 0) {
        if (countVisits) hipLaunchKernelGGL((kTracePA<true, LDS>), g, t, 0, s, S, qe, esel, qs, qp, pb, cnt, poll);
        else hipLaunchKernelGGL((kTracePA<false, LDS>), g, t, 0, s, S, qe, esel, qs, qp, pb, cnt, poll);
    } else {
        if (countVisits) hipLaunchKernelGGL((kTraceP<true, LDS>), g, t, 0, s, S, qe, esel, qs, qp, pb, cnt, poll);
        else hipLaunchKernelGGL((kTraceP<false, LDS>), g, t, 0, s, S, qe, esel, qs, qp, pb, cnt, poll);
    }
}
void launchTraceP(hipStream_t s, const DevScene &S, Queue qe, int esel, Queue qs, Queue qp, PathBuffers pb,
                  Counters *cnt, uint32_t *poll, bool countVisits, int blocks, int ldsEntries) {
    switch (ldsEntries) {
    case 12: launchTracePL<12>(s, S, qe, esel, qs, qp, pb, cnt, poll, countVisits, blocks); break;
    case 16: launchTracePL<16>(s, S, qe, esel, qs, qp, pb, cnt, poll, countVisits, blocks); break;
    case 24: launchTracePL<24>(s, S, qe, esel, qs, qp, pb, cnt, poll, countVisits, blocks); break;
    default: launchTracePL<kLdsStack>(s, S, qe, esel, qs, qp, pb, cnt, poll, countVisits, blocks); break;
    }
}
void launchProbeResolve(hipStream_t s, const DevScene &S, const FrameConst *fc, Queue q, PathBuffers pb,
                        Counters *cnt, int qsel, int blocks) {
    hipLaunchKernelGGL(kProbeResolve, dim3(blocks), dim3(256), 0, s, S, fc, q, pb, cnt, qsel);
}
void launchShade(hipStream_t s, const DevScene &S, const DevIC &ic, const DevGuiding &g, const FrameConst *fc,
                 BatchDesc b, PathBuffers pb, Queue qin, Queue qout, Queue qs, Queue qp, int32_t *done, Counters *cnt,
                 int qsel, int blocks, const PathClass2 *second) {
    if (second)
        hipLaunchKernelGGL(kShade2, dim3(blocks), dim3(kShadeBlock), 0, s, S, ic, g, fc, b, second->fc, second->b,
                           second->split, pb, qin, qout, qs, qp, done, cnt, qsel);
    else
        hipLaunchKernelGGL(kShade, dim3(blocks), dim3(kShadeBlock), 0, s, S, ic, g, fc, b, pb, qin, qout, qs, qp, done,
                           cnt, qsel);
}
void launchFinish(hipStream_t s, const FrameConst *fc, BatchDesc b, PathBuffers pb, const int32_t *done,
                  Counters *cnt, Queue qnext, int qsel, int blocks, const PathClass2 *second) {
    if (second)
        hipLaunchKernelGGL(kFinish2, dim3(blocks), dim3(256), 0, s, fc, b, second->fc, second->b, second->split, pb,
                           done, cnt, qnext, qsel);
    else
        hipLaunchKernelGGL(kFinish, dim3(blocks), dim3(256), 0, s, fc, b, pb, done, cnt, qnext, qsel);
}
void launchAccumulate(hipStream_t s, BatchDesc b, PathBuffers pb, float4 *accum, float4 *keepRgb, uint2 *keepSig,
                      bool keep, float4 *estimate) {
    uint32_t lanes = b.numPaths < b.npix ? b.numPaths : b.npix;
    hipLaunchKernelGGL(kAccumulate, dim3(gridFor(lanes, 256)), dim3(256), 0, s, b, pb, accum, keepRgb, keepSig,
                       keep ? 1 : 0, estimate);
}
void launchResetCounters(hipStream_t s, Counters *cnt, int which) {
    hipLaunchKernelGGL(kResetCounters, dim3(1), dim3(64), 0, s, cnt, which);
}
void launchQueryP(hipStream_t s, const DevScene &S, const rtx_ray *rays, uint32_t n, void *out, bool closest,
                  uint32_t *fetch, int blocks) {
    if (n == 0) return;
    if (closest) hipLaunchKernelGGL(kQueryP<false>, dim3(blocks), dim3(kTraceBlock), 0, s, S, rays, n, out, fetch);
    else hipLaunchKernelGGL(kQueryP<true>, dim3(blocks), dim3(kTraceBlock), 0, s, S, rays, n, out, fetch);
}
void launchInstallSlot(hipStream_t s, const SlotDesc &sd, SlotDesc *slots, uint32_t slot, uint32_t *alive) {
    hipLaunchKernelGGL(kInstallSlot, dim3(1), dim3(256), 0, s, sd, slots, slot, alive);
}
void launchGenerateM(hipStream_t s, const SlotDesc *slots, uint32_t slot, uint32_t numPaths, PathBuffers pb, Queue q0,
                     Counters *cnt, int qsel) {
    if (numPaths == 0) return;
    hipLaunchKernelGGL(kGenerateM, dim3(gridFor(numPaths, 256)), dim3(256), 0, s, slots, slot, pb, q0, cnt, qsel);
}
void launchTraceM(hipStream_t s, const MultiArgs &ma, bool alpha, Queue qe, int esel, Queue qs, Queue qp,
                  PathBuffers pb, Counters *cnt, uint32_t *poll, int blocks) {
    const dim3 g(blocks), t(kTraceBlock);
    if (alpha) hipLaunchKernelGGL(kTraceM<true>, g, t, 0, s, ma, qe, esel, qs, qp, pb, cnt, poll);
    else hipLaunchKernelGGL(kTraceM<false>, g, t, 0, s, ma, qe, esel, qs, qp, pb, cnt, poll);
}
void launchProbeResolveM(hipStream_t s, const MultiArgs &ma, Queue q, PathBuffers pb, Counters *cnt, int qsel,
                         int blocks) {
    hipLaunchKernelGGL(kProbeResolveM, dim3(blocks), dim3(256), 0, s, ma, q, pb, cnt, qsel);
}
void launchFinishM(hipStream_t s, const MultiArgs &ma, PathBuffers pb, const int32_t *done, Counters *cnt,
                   Queue qnext, int qsel, int blocks) {
    hipLaunchKernelGGL(kFinishM, dim3(blocks), dim3(256), 0, s, ma, pb, done, cnt, qnext, qsel);
}
void launchShadeM(hipStream_t s, const MultiArgs &ma, PathBuffers pb, Queue qin, Queue qout, Queue qs, Queue qp,
                  int32_t *done, Counters *cnt, int qsel, int blocks) {
    hipLaunchKernelGGL(kShadeM, dim3(blocks), dim3(kShadeBlock), 0, s, ma, pb, qin, qout, qs, qp, done, cnt, qsel);
}
void launchFillU32(hipStream_t s, uint32_t *p, uint32_t n, uint32_t v) {
    if (n == 0) return;
    hipLaunchKernelGGL(kFillU32, dim3(gridFor(n, 256)), dim3(256), 0, s, p, n, v);
}
void launchResolveImage(hipStream_t s, const float4 *accum, uint32_t npix, float inv, float4 *out) {
    hipLaunchKernelGGL(kResolveImage, dim3(gridFor(npix, 256)), dim3(256), 0, s, accum, npix, inv, out);
}

} // namespace rtx
