// C ABI of librtxamd.so (include/rtx_amd.h): context, scene upload + BVH
// build, wavefront render orchestration, readback and ray queries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "abi_guard.h"
#include "context.h"
#include "dev_sort.h"
#include "dev_math.h"
#include "internal.h"
#include "nrc.h"

using namespace rtx;

namespace {

#define CHECK_HIP(ctx, expr)                                                                              \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) {                                                                           \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                              \
            return RTX_E_DEVICE;                                                                          \
        }                                                                                                 \
    } while (0)

int fail(rtx_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

template <class T> hipError_t devAlloc(T **p, size_t count, std::vector<void *> &list) {
    void *q = nullptr;
    hipError_t e = devMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) {
        list.push_back(q);
        *p = (T *)q;
    }
    return e;
}

// Replaces one allocation of a context list by a new one of `count` items.
template <class T> hipError_t reallocIn(std::vector<void *> &list, T *&p, size_t count) {
    auto it = std::find(list.begin(), list.end(), (void *)p);
    if (it != list.end()) {
        (void)hipFree(*it);
        list.erase(it);
    }
    p = nullptr;
    return devAlloc(&p, count, list);
}

void freeAll(std::vector<void *> &list) {
    for (void *p : list) (void)hipFree(p);
    list.clear();
}

// mat4 * (v, 1) in FP32 with GLM's summation order
// The sRGB transfer function's inverse (IEC 61966-2-1), as Vulkan applies it
// to each colour component of an SRGB-format texel before filtering.
float srgbDecode(int i) {
    double c = i / 255.0;
    return (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
}

void xformPoint(const float *m, const float *v, float *o) {
    for (int r = 0; r < 3; ++r) {
        float a0 = m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1];
        float a1 = m[2 * 4 + r] * v[2] + m[3 * 4 + r] * 1.0f;
        o[r] = a0 + a1;
    }
}

int validateScene(rtx_ctx *ctx, const rtx_scene_desc *d, uint64_t &numTris) {
    if (!d) return fail(ctx, RTX_E_INVALID, "scene desc is NULL");
    if (d->modelCount && !d->models) return fail(ctx, RTX_E_INVALID, "models is NULL");
    if (d->instanceCount && !d->instances) return fail(ctx, RTX_E_INVALID, "instances is NULL");
    if (d->materialCount == 0 || !d->materials) return fail(ctx, RTX_E_INVALID, "a scene needs materials");
    if (d->lightCount && (!d->lights || !d->randomLightIndex))
        return fail(ctx, RTX_E_INVALID, "lights without light tables");
    if (d->sphereCount && !d->spheres) return fail(ctx, RTX_E_INVALID, "spheres is NULL");
    if (d->envMap && (d->envWidth == 0 || d->envHeight == 0))
        return fail(ctx, RTX_E_INVALID, "environment map without a size");
    if (d->textureCount && !d->textures) return fail(ctx, RTX_E_INVALID, "textures is NULL");
    for (uint32_t t = 1; t < d->textureCount; ++t) {
        const rtx_texture_desc &td = d->textures[t];
        if (!td.rgba || td.width == 0 || td.height == 0)
            return fail(ctx, RTX_E_INVALID, "texture slot " + std::to_string(t) + " has no texels");
        if ((uint64_t)td.width * td.height >= (1ull << 31)) return fail(ctx, RTX_E_UNSUPPORTED, "texture too large");
    }
    for (uint32_t m = 0; m < d->modelCount; ++m) {
        const rtx_model_desc &md = d->models[m];
        if (md.indexCount % 3) return fail(ctx, RTX_E_INVALID, "model index count not a multiple of 3");
        if ((md.indexCount && !md.indices) || (md.vertexCount && !md.vertices))
            return fail(ctx, RTX_E_INVALID, "model buffers are NULL");
        for (uint32_t i = 0; i < md.indexCount; ++i)
            if (md.indices[i] >= md.vertexCount) return fail(ctx, RTX_E_INVALID, "vertex index out of range");
        for (uint32_t v = 0; v < md.vertexCount; ++v)
            if (md.vertices[v].materialIndex < 0 || (uint32_t)md.vertices[v].materialIndex >= d->materialCount)
                return fail(ctx, RTX_E_INVALID, "vertex material index out of range");
    }
    numTris = 0;
    for (uint32_t i = 0; i < d->instanceCount; ++i) {
        const rtx_instance &in = d->instances[i];
        if (in.modelIndex < 0 || (uint32_t)in.modelIndex >= d->modelCount)
            return fail(ctx, RTX_E_INVALID, "instance model index out of range");
        if (in.iLight >= (int)d->lightCount) return fail(ctx, RTX_E_INVALID, "instance light index out of range");
        numTris += d->models[in.modelIndex].indexCount / 3;
    }
    if (numTris >= (1ull << 27)) return fail(ctx, RTX_E_UNSUPPORTED, "more than 2^27 triangles");
    for (uint32_t m = 0; m < d->materialCount; ++m) {
        if (d->materials[m].type < 0 || d->materials[m].type > RTX_MAT_ROUGH_CONDUCTOR)
            return fail(ctx, RTX_E_INVALID, "unknown material type");
        for (int32_t tid : {d->materials[m].textureIdDiffuse, d->materials[m].textureIdSpecular})
            if (tid != -1 && (tid < 1 || (uint32_t)tid >= d->textureCount))
                return fail(ctx, RTX_E_INVALID, "material texture id " + std::to_string(tid) + " names no texture");
    }
    for (uint32_t s = 0; s < d->sphereCount; ++s) {
        const rtx_sphere &sp = d->spheres[s];
        if (sp.materialIndex < 0 || (uint32_t)sp.materialIndex >= d->materialCount)
            return fail(ctx, RTX_E_INVALID, "sphere material index out of range");
        if (sp.iLight >= (int)d->lightCount) return fail(ctx, RTX_E_INVALID, "sphere light index out of range");
    }
    for (uint32_t l = 0; l < d->lightCount; ++l) {
        const rtx_light &li = d->lights[l];
        switch (li.type) {
        case RTX_LIGHT_POINT: break;
        case RTX_LIGHT_SPHERE:
            if (li.instanceIndex >= d->sphereCount) return fail(ctx, RTX_E_INVALID, "sphere light out of range");
            break;
        case RTX_LIGHT_AREA: {
            if (li.instanceIndex >= d->instanceCount) return fail(ctx, RTX_E_INVALID, "area light instance out of range");
            // The reference indexes randomTriIndex by the global light index
            // (shaders/raytrace.rgen:427); rows exist only for area lights.
            if (l >= d->randomTriRows || !d->randomTriIndex)
                return fail(ctx, RTX_E_INVALID, "area light has no face-sampling row (area lights must come first)");
            const rtx_model_desc &md = d->models[d->instances[li.instanceIndex].modelIndex];
            for (int k = 0; k < RTX_SIZE_TRI_RANDOM; ++k) {
                int f = d->randomTriIndex[(size_t)l * RTX_SIZE_TRI_RANDOM + k].index;
                if (f < 0 || (uint32_t)f >= md.indexCount / 3)
                    return fail(ctx, RTX_E_INVALID, "face sample out of range");
            }
            break;
        }
        case RTX_LIGHT_ENV_MAP:
            if (!d->envMap) return fail(ctx, RTX_E_INVALID, "environment-map light without an environment map");
            break;
        default:
            return fail(ctx, RTX_E_INVALID, "unknown light type");
        }
    }
    if (d->lightCount)
        for (int k = 0; k < RTX_SIZE_LIGHT_RANDOM; ++k)
            if (d->randomLightIndex[k] < 0 || (uint32_t)d->randomLightIndex[k] >= d->lightCount)
                return fail(ctx, RTX_E_INVALID, "light table entry out of range");
    return RTX_OK;
}

int ensureBatchBuffers(rtx_ctx *ctx, size_t paths, uint32_t numNEE) {
    if (paths <= ctx->cap && numNEE <= ctx->capNEE) return RTX_OK;
    freeAll(ctx->batchAllocs);
    size_t cap = std::max(paths, ctx->cap);
    uint32_t nee = std::max(numNEE, ctx->capNEE);
    auto &L = ctx->batchAllocs;
    PathBuffers &pb = ctx->pb;
    CHECK_HIP(ctx, devAlloc(&pb.rng, cap, L));
    CHECK_HIP(ctx, devAlloc(&pb.hash, cap, L));
    CHECK_HIP(ctx, devAlloc(&pb.frameSeed, cap, L));
    CHECK_HIP(ctx, devAlloc(&pb.thr, cap, L));
    CHECK_HIP(ctx, devAlloc(&pb.result, cap, L));
    CHECK_HIP(ctx, devAlloc(&pb.pendShadow, cap * nee, L));
    CHECK_HIP(ctx, devAlloc(&pb.pendProbe, cap * nee, L));
    CHECK_HIP(ctx, devAlloc(&pb.out, cap, L));
    CHECK_HIP(ctx, devAlloc(&pb.outSig, cap, L));
    for (int k = 0; k < 2; ++k) {
        CHECK_HIP(ctx, devAlloc(&ctx->q[k].id, cap, L));
        CHECK_HIP(ctx, devAlloc(&ctx->q[k].org, cap, L));
        CHECK_HIP(ctx, devAlloc(&ctx->q[k].dir, cap, L));
        CHECK_HIP(ctx, devAlloc(&ctx->q[k].hit, cap, L));
        ctx->q[k].aux = nullptr;
    }
    // shadow queue: a shadow ray and a probe's occlusion ray per NEE slot
    CHECK_HIP(ctx, devAlloc(&ctx->qs.id, 2 * cap * nee, L));
    CHECK_HIP(ctx, devAlloc(&ctx->qs.org, 2 * cap * nee, L));
    CHECK_HIP(ctx, devAlloc(&ctx->qs.dir, 2 * cap * nee, L));
    ctx->qs.aux = nullptr;
    ctx->qs.hit = nullptr;
    CHECK_HIP(ctx, devAlloc(&ctx->qp.id, cap * nee, L));
    CHECK_HIP(ctx, devAlloc(&ctx->qp.org, cap * nee, L));
    CHECK_HIP(ctx, devAlloc(&ctx->qp.dir, cap * nee, L));
    CHECK_HIP(ctx, devAlloc(&ctx->qp.aux, cap * nee, L));
    CHECK_HIP(ctx, devAlloc(&ctx->qp.hit, cap * nee, L));
    CHECK_HIP(ctx, devAlloc(&ctx->done, cap, L));
    ctx->cap = cap;
    ctx->capNEE = nee;
    return RTX_OK;
}

// Median split on the widest axis of the region centres, depth first.
static int buildGuideRec(std::vector<GuideNode> &out, const rtx_aabb *boxes, std::vector<uint32_t> &ids, size_t lo,
                         size_t hi) {
    const int me = (int)out.size();
    out.push_back(GuideNode{});
    GuideNode nd{};
    for (int a = 0; a < 3; ++a) {
        nd.lo[a] = boxes[ids[lo]].min[a];
        nd.hi[a] = boxes[ids[lo]].max[a];
    }
    float clo[3], chi[3];
    for (int a = 0; a < 3; ++a) clo[a] = chi[a] = 0.5f * (boxes[ids[lo]].min[a] + boxes[ids[lo]].max[a]);
    for (size_t i = lo; i < hi; ++i)
        for (int a = 0; a < 3; ++a) {
            nd.lo[a] = std::min(nd.lo[a], boxes[ids[i]].min[a]);
            nd.hi[a] = std::max(nd.hi[a], boxes[ids[i]].max[a]);
            float c = 0.5f * (boxes[ids[i]].min[a] + boxes[ids[i]].max[a]);
            clo[a] = std::min(clo[a], c);
            chi[a] = std::max(chi[a], c);
        }
    if (hi - lo == 1) {
        nd.region = (int32_t)ids[lo];
        nd.right = -1;
        out[me] = nd;
        return me;
    }
    int axis = 0;
    for (int a = 1; a < 3; ++a)
        if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
    const size_t mid = (lo + hi) / 2;
    std::nth_element(ids.begin() + lo, ids.begin() + mid, ids.begin() + hi, [&](uint32_t x, uint32_t y) {
        float cx = boxes[x].min[axis] + boxes[x].max[axis], cy = boxes[y].min[axis] + boxes[y].max[axis];
        return cx < cy || (cx == cy && x < y);
    });
    nd.region = -1;
    buildGuideRec(out, boxes, ids, lo, mid);
    nd.right = buildGuideRec(out, boxes, ids, mid, hi);
    out[me] = nd;
    return me;
}

// Level-0 cell of the lookup index: the diameter of the smallest record
// (irradianceA * irradianceCacheMinRadius), or 1e-3 when that is 0.
float icBaseCell(const rtx_push_constant &pc) {
    float c0 = 2.0f * pc.irradianceA * pc.irradianceCacheMinRadius;
    return c0 > 1e-6f ? c0 : 1e-3f;
}

// Paths of one wavefront batch of the cache's sub-paths: whole jobs.
constexpr uint32_t kMaxBlockSums = (1u << 26) / 256 + 1;

const char *unsupportedFeature(const rtx_push_constant &pc) {
    if (pc.showIrradianceCacheOnly) return "showIrradianceCacheOnly";
    if (pc.visualizeMode != 0) return "visualizeMode";
    return nullptr;
}

EventPair *nextEvents(rtx_ctx *ctx, int cls) {
    if (ctx->eventsUsed == ctx->events.size()) {
        EventPair p;
        if (hipEventCreate(&p.a) != hipSuccess || hipEventCreate(&p.b) != hipSuccess) return nullptr;
        ctx->events.push_back(p);
    }
    EventPair *p = &ctx->events[ctx->eventsUsed++];
    p->cls = cls;
    return p;
}

} // namespace

static void *noAllocScratch(int device, hipError_t &err);

namespace rtx {
int abiStatus() noexcept {
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return RTX_E_MEMORY;
    } catch (...) {
        return RTX_E_DEVICE;
    }
}
int abiFail(rtx_ctx *ctx) noexcept {
    const int code = abiStatus();
    if (ctx) {
        try {
            throw;
        } catch (const std::exception &e) {
            try {
                ctx->err = std::string("internal error: ") + e.what();
            } catch (...) {
            }
        } catch (...) {
            try {
                ctx->err = "internal error";
            } catch (...) {
            }
        }
    }
    return code;
}
} // namespace rtx

namespace rtx {
std::vector<GuideNode> buildGuideNodes(const rtx_aabb *boxes, uint32_t n) {
    std::vector<GuideNode> out;
    if (n == 0) return out;
    std::vector<uint32_t> ids(n);
    for (uint32_t i = 0; i < n; ++i) ids[i] = i;
    out.reserve(2 * (size_t)n);
    buildGuideRec(out, boxes, ids, 0, n);
    return out;
}
} // namespace rtx

extern "C" {

void rtx_push_constant_defaults(rtx_push_constant *pc) {
    if (!pc) return;
    std::memset(pc, 0, sizeof(*pc));
    // src/RayTracingApp.h:81-130
    pc->previousFrames = (uint32_t)-1;
    pc->maxDepth = 30;
    pc->maxFollowDiscrete = 3;
    pc->samplesPerPixel = 1;
    pc->enableNEE = 1;
    pc->numNEE = 1;
    pc->usePowerHeuristic = 1;
    pc->irradianceA = 0.2f;
    pc->irradianceUpdateProb = 0.00001f;
    pc->irradianceCreateProb = 0.0005f;
    pc->irradianceVisualizationScale = 1.0f;
    pc->irradianceGradientsMaxLength = 5;
    pc->irradianceNumNEE = 1;
    pc->irradianceCacheMinRadius = 0.1f;
    pc->adrrsS = 5;
    pc->adrrsSplit = 1;
    pc->guidingProb = 0.5f;
    pc->guidingVisuScale = 0.5f;
    pc->guidingVisuMax = 1.0f;
    pc->useParallaxCompensation = 1;
    pc->guidingVisuPhiScale = 0.004f;
    pc->guidingVisuThetaScale = 0.01f;
    pc->guidingPiPShowSpheres = 1;
    pc->guidingPiPSize = 0.3f;
}

const char *rtx_last_error(const rtx_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int rtx_ctx_set_tuning(rtx_ctx *ctx, const rtx_tuning *t) try {
    if (!ctx || !t) return RTX_E_INVALID;
    rtx_tuning v = ctx->tuning;
    if (t->traversalKernel) v.traversalKernel = t->traversalKernel;
    if (t->traceBlocksPerCU) v.traceBlocksPerCU = t->traceBlocksPerCU;
    if (t->shadeBlocksPerCU) v.shadeBlocksPerCU = t->shadeBlocksPerCU;
    if (t->batchPaths) v.batchPaths = t->batchPaths;
    if (t->ldsStackEntries) v.ldsStackEntries = t->ldsStackEntries;
    if (t->cacheUpkeep) v.cacheUpkeep = t->cacheUpkeep;
    if (v.traversalKernel != 1)
        return fail(ctx, RTX_E_INVALID, "rtx_tuning.traversalKernel must be 1 (persistent lane-refill traversal)");
    if (v.traceBlocksPerCU < 1 || v.traceBlocksPerCU > 64 || v.shadeBlocksPerCU < 1 || v.shadeBlocksPerCU > 64)
        return fail(ctx, RTX_E_INVALID, "rtx_tuning blocks per CU must lie in [1, 64]");
    if (v.batchPaths < 1024 || v.batchPaths > (1u << 26))
        return fail(ctx, RTX_E_INVALID, "rtx_tuning.batchPaths must lie in [1024, 2^26]");
    if (v.ldsStackEntries != 12 && v.ldsStackEntries != 16 && v.ldsStackEntries != 20 && v.ldsStackEntries != 24)
        return fail(ctx, RTX_E_INVALID, "rtx_tuning.ldsStackEntries must be 12, 16, 20 or 24");
    if (v.cacheUpkeep != 1 && v.cacheUpkeep != 2) return fail(ctx, RTX_E_INVALID, "rtx_tuning.cacheUpkeep must be 1 or 2");
    ctx->tuning = v;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_ctx_get_tuning(const rtx_ctx *ctx, rtx_tuning *t) try {
    if (!ctx || !t) return RTX_E_INVALID;
    *t = ctx->tuning;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(nullptr);
}

int rtx_autotune(rtx_ctx *ctx, const rtx_scene *scene, const rtx_frame_params *probe, uint32_t repeats,
                 rtx_autotune_result *result) try {
    if (!ctx) return RTX_E_INVALID;
    if (!scene || !probe || !result) return fail(ctx, RTX_E_INVALID, "NULL argument");
    if (repeats < 1 || repeats > 16) return fail(ctx, RTX_E_INVALID, "repeats must be in [1, 16]");
    std::memset(result, 0, sizeof(*result));
    const rtx_tuning start = ctx->tuning;
    rtx_tuning cur = start;
    uint32_t trials = 0;
    // fastest of `repeats` renders of the probe under tuning t (device time)
    auto timeWith = [&](const rtx_tuning &t, double &ms) -> int {
        ctx->tuning = t;
        ms = 1e300;
        for (uint32_t k = 0; k < repeats; ++k) {
            rtx_frame_stats st;
            const int rc = rtx_render(ctx, scene, probe, &st);
            if (rc != RTX_OK) return rc;
            ms = std::min(ms, st.msTotal);
            trials++;
        }
        return RTX_OK;
    };
    double best = 0;
    int rc = timeWith(cur, best);
    if (rc != RTX_OK) {
        ctx->tuning = start;
        return rc;
    }
    result->startMs = best;
    const int32_t traceOpts[] = {8, 10, 12, 14, 16, 20};
    const int32_t ldsOpts[] = {12, 16, 20, 24};
    const int32_t shadeOpts[] = {4, 6, 8, 12, 16};
    const uint32_t batchOpts[] = {1u << 20, 1u << 21, 1u << 22, 1u << 23};
    auto descend = [&](auto field, const auto *opts, size_t n) -> int {
        for (size_t i = 0; i < n; ++i) {
            if (*field(cur) == opts[i]) continue;
            rtx_tuning t = cur;
            *field(t) = opts[i];
            double ms = 0;
            const int r = timeWith(t, ms);
            if (r != RTX_OK) return r;
            if (ms < best) {
                best = ms;
                cur = t;
            }
        }
        return RTX_OK;
    };
    rc = descend([](rtx_tuning &t) { return &t.traceBlocksPerCU; }, traceOpts, 6);
    if (rc == RTX_OK) rc = descend([](rtx_tuning &t) { return &t.ldsStackEntries; }, ldsOpts, 4);
    if (rc == RTX_OK) rc = descend([](rtx_tuning &t) { return &t.shadeBlocksPerCU; }, shadeOpts, 5);
    if (rc == RTX_OK) rc = descend([](rtx_tuning &t) { return &t.batchPaths; }, batchOpts, 4);
    if (rc != RTX_OK) {
        ctx->tuning = start;
        return rc;
    }
    ctx->tuning = cur;
    result->best = cur;
    result->bestMs = best;
    result->trials = trials;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_ctx_create(int device, rtx_ctx **out) try {
    if (!out) return RTX_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return RTX_E_DEVICE;
    if (device < 0 || device >= n) return RTX_E_INVALID;
    auto *ctx = new (std::nothrow) rtx_ctx();
    if (!ctx) return RTX_E_MEMORY;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        devMalloc(&ctx->cnt, sizeof(Counters)) != hipSuccess || devMalloc(&ctx->fc, sizeof(FrameConst)) != hipSuccess ||
        hipEventCreate(&ctx->tStart) != hipSuccess || hipEventCreate(&ctx->tStop) != hipSuccess ||
        hipHostMalloc((void **)&ctx->pollRing, kPollSlots * 4 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
        rtx_ctx_destroy(ctx);
        return RTX_E_DEVICE;
    }
    for (hipEvent_t &e : ctx->pollEv) {
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            rtx_ctx_destroy(ctx);
            return RTX_E_DEVICE;
        }
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->cuCount = prop.multiProcessorCount;
    // the allocation-free query's scratch on this device (looked up once per device)
    hipError_t se = hipSuccess;
    ctx->scratch = noAllocScratch(device, se);
    if (!ctx->scratch) ctx->scratchErr = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se);
    *out = ctx;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(nullptr);
}

int rtx_ctx_destroy(rtx_ctx *ctx) try {
    if (!ctx) return RTX_E_INVALID;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    freeAll(ctx->batchAllocs);
    freeAll(ctx->icAllocs);
    freeAll(ctx->guideAllocs);
    freeAll(ctx->fitAllocs);
    freeAll(ctx->fitWorkAllocs);
    freeAll(ctx->recAllocs);
    freeAll(ctx->segAllocs);
    freeAll(ctx->splitAllocs);
    freeAll(ctx->nrcAllocs);
    freeAll(ctx->nrcPathAllocs);
    freeAll(ctx->nrcSlotAllocs);
    if (ctx->estimate) (void)hipFree(ctx->estimate);
    if (ctx->dirData) (void)hipFree(ctx->dirData);
    if (ctx->fcIC) (void)hipFree(ctx->fcIC);
    if (ctx->candBuf) (void)hipFree(ctx->candBuf);
    if (ctx->keepSub) (void)hipFree(ctx->keepSub);
    if (ctx->keepSubSig) (void)hipFree(ctx->keepSubSig);
    if (ctx->accum) (void)hipFree(ctx->accum);
    if (ctx->resolved) (void)hipFree(ctx->resolved);
    if (ctx->keepRgb) (void)hipFree(ctx->keepRgb);
    if (ctx->keepSig) (void)hipFree(ctx->keepSig);
    if (ctx->cnt) (void)hipFree(ctx->cnt);
    if (ctx->fc) (void)hipFree(ctx->fc);
    for (auto &e : ctx->events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (hipEvent_t e : ctx->pollEv)
        if (e) (void)hipEventDestroy(e);
    if (ctx->pollRing) (void)hipHostFree(ctx->pollRing);
    if (ctx->tStart) (void)hipEventDestroy(ctx->tStart);
    if (ctx->tStop) (void)hipEventDestroy(ctx->tStop);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_scene_create(rtx_ctx *ctx, const rtx_scene_desc *d, rtx_scene **out) try {
    if (!ctx || !out) return RTX_E_INVALID;
    *out = nullptr;
    uint64_t numTris = 0;
    int rc = validateScene(ctx, d, numTris);
    if (rc != RTX_OK) return rc;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));

    // flatten instances into world-space triangles (global primitive order)
    std::vector<float> tv((size_t)numTris * 9);
    std::vector<int2> primRef((size_t)numTris);
    size_t t = 0;
    for (uint32_t i = 0; i < d->instanceCount; ++i) {
        const rtx_instance &in = d->instances[i];
        const rtx_model_desc &md = d->models[in.modelIndex];
        for (uint32_t f = 0; f < md.indexCount / 3; ++f, ++t) {
            for (int k = 0; k < 3; ++k)
                xformPoint(in.transform, md.vertices[md.indices[3 * f + k]].pos, &tv[9 * t + 3 * k]);
            primRef[t] = make_int2((int)i, (int)f);
        }
    }
    BvhBuildResult bvh;
    try {
        buildBvh(tv.data(), (uint32_t)numTris, bvh);
    } catch (const std::exception &e) {
        return fail(ctx, RTX_E_INVALID, std::string("BVH build failed: ") + e.what());
    }
    // triangles of materials with a diffuse texture run the any-hit alpha
    // test (raytrace.rahit:37-46 reads the material of the first vertex)
    std::vector<uint8_t> alpha((size_t)numTris, 0);
    uint32_t alphaTris = 0;
    for (size_t p = 0; p < (size_t)numTris; ++p) {
        const rtx_instance &in = d->instances[primRef[p].x];
        const rtx_model_desc &md = d->models[in.modelIndex];
        int mat = md.vertices[md.indices[3 * (size_t)primRef[p].y]].materialIndex;
        if (d->materials[mat].textureIdDiffuse != -1) {
            alpha[p] = 1;
            alphaTris++;
        }
    }
    std::vector<float4> tris((size_t)numTris * 3);
    for (size_t k = 0; k < bvh.primOrder.size(); ++k) {
        uint32_t p = bvh.primOrder[k];
        const float *v = &tv[9 * (size_t)p];
        int pid = (int)p | (alpha[p] ? kAlphaBit : 0);
        float pidBits;
        std::memcpy(&pidBits, &pid, 4);
        tris[3 * k + 0] = make_float4(v[0], v[1], v[2], pidBits);
        tris[3 * k + 1] = make_float4(v[3], v[4], v[5], 0.0f);
        tris[3 * k + 2] = make_float4(v[6], v[7], v[8], 0.0f);
    }
    // the emissive triangles (in BVH order), for probe rays that look for
    // their light first: kept only while few and none is alpha-tested
    std::vector<float4> lightTris;
    bool lightTrisOk = true;
    for (size_t k = 0; k < bvh.primOrder.size() && lightTrisOk; ++k) {
        const uint32_t p = bvh.primOrder[k];
        const rtx_instance &in = d->instances[primRef[p].x];
        const rtx_model_desc &md = d->models[in.modelIndex];
        const int mat = md.vertices[md.indices[3 * (size_t)primRef[p].y]].materialIndex;
        if (d->materials[mat].type != RTX_MAT_LIGHT) continue;
        if (alpha[p] || lightTris.size() >= 3 * (size_t)kProbeLightMax) lightTrisOk = false;
        for (int j = 0; j < 3; ++j) lightTris.push_back(tris[3 * k + j]);
    }
    if (!lightTrisOk) lightTris.clear();
    std::vector<rtx_vertex> verts;
    std::vector<uint32_t> idx, vbase(std::max<uint32_t>(d->modelCount, 1)), ibase(std::max<uint32_t>(d->modelCount, 1));
    for (uint32_t m = 0; m < d->modelCount; ++m) {
        vbase[m] = (uint32_t)verts.size();
        ibase[m] = (uint32_t)idx.size();
        verts.insert(verts.end(), d->models[m].vertices, d->models[m].vertices + d->models[m].vertexCount);
        idx.insert(idx.end(), d->models[m].indices, d->models[m].indices + d->models[m].indexCount);
    }

    auto *sc = new (std::nothrow) rtx_scene();
    if (!sc) return fail(ctx, RTX_E_MEMORY, "out of memory");
    sc->device = ctx->device;
    auto upload = [&](auto **dst, const auto *src, size_t count) -> hipError_t {
        using T = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
        T *p = nullptr;
        hipError_t e = devAlloc(&p, count, sc->allocs);
        if (e != hipSuccess) return e;
        if (count) e = hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice);
        *dst = p;
        return e;
    };
    DevScene &S = sc->view;
    hipError_t e = hipSuccess;
    float4 *nodes = nullptr, *trisD = nullptr;
    int4 *prD = nullptr;
    rtx_vertex *vD = nullptr;
    uint32_t *iD = nullptr, *vbD = nullptr, *ibD = nullptr;
    rtx_instance *instD = nullptr;
    rtx_material *matD = nullptr;
    rtx_light *lightD = nullptr;
    int32_t *liD = nullptr;
    rtx_face_sample *ftD = nullptr;
    rtx_sphere *spD = nullptr;
    if (e == hipSuccess) e = upload(&nodes, reinterpret_cast<const float4 *>(bvh.nodes.data()), bvh.nodes.size() * 8);
    if (e == hipSuccess) e = upload(&trisD, tris.data(), tris.size());
    float4 *lightTrisD = nullptr;
    if (e == hipSuccess && lightTrisOk && !lightTris.empty()) e = upload(&lightTrisD, lightTris.data(), lightTris.size());
    // per world triangle: its instance and the global indices of its vertices
    std::vector<int4> triRef((size_t)numTris);
    for (size_t p = 0; p < (size_t)numTris; ++p) {
        const int inst = primRef[p].x;
        const int model = d->instances[inst].modelIndex;
        const uint32_t *ix = d->models[model].indices + 3 * (size_t)primRef[p].y;
        triRef[p] = make_int4(inst, (int)(vbase[model] + ix[0]), (int)(vbase[model] + ix[1]), (int)(vbase[model] + ix[2]));
    }
    if (e == hipSuccess) e = upload(&prD, triRef.data(), triRef.size());
    if (e == hipSuccess) e = upload(&vD, verts.data(), verts.size());
    if (e == hipSuccess) e = upload(&iD, idx.data(), idx.size());
    if (e == hipSuccess) e = upload(&vbD, vbase.data(), vbase.size());
    if (e == hipSuccess) e = upload(&ibD, ibase.data(), ibase.size());
    if (e == hipSuccess) e = upload(&instD, d->instances, d->instanceCount);
    if (e == hipSuccess) e = upload(&matD, d->materials, d->materialCount);
    if (e == hipSuccess) e = upload(&lightD, d->lights, d->lightCount);
    if (e == hipSuccess) e = upload(&liD, d->randomLightIndex, d->lightCount ? RTX_SIZE_LIGHT_RANDOM : 0);
    if (e == hipSuccess)
        e = upload(&ftD, d->randomTriIndex, d->randomTriIndex ? (size_t)d->randomTriRows * RTX_SIZE_TRI_RANDOM : 0);
    if (e == hipSuccess) e = upload(&spD, d->spheres, d->sphereCount);
    // the area lights' face-table entries with their vertices and colour
    // (validateScene checked rows, instances and face indices)
    float4 *emitD = nullptr;
    rtx_instance *linstD = nullptr;
    const size_t emitRows = d->randomTriIndex ? std::min<size_t>(d->randomTriRows, d->lightCount) : 0;
    if (e == hipSuccess && emitRows && emitRows * RTX_SIZE_TRI_RANDOM * 6 * sizeof(float4) <= kEmitTableBytes) {
        std::vector<float4> emit(emitRows * RTX_SIZE_TRI_RANDOM * 6, make_float4(0, 0, 0, 0));
        std::vector<rtx_instance> linst(emitRows);
        std::memset(linst.data(), 0, linst.size() * sizeof(rtx_instance));
        for (size_t r = 0; r < emitRows; ++r) {
            const rtx_light &L = d->lights[r];
            if (L.type != RTX_LIGHT_AREA) continue;
            const rtx_instance &in = d->instances[L.instanceIndex];
            linst[r] = in;
            const rtx_model_desc &md = d->models[in.modelIndex];
            for (size_t k = 0; k < RTX_SIZE_TRI_RANDOM; ++k) {
                const size_t f = (size_t)d->randomTriIndex[r * RTX_SIZE_TRI_RANDOM + k].index;
                const rtx_vertex &v0 = md.vertices[md.indices[3 * f]], &v1 = md.vertices[md.indices[3 * f + 1]],
                                 &v2 = md.vertices[md.indices[3 * f + 2]];
                const float *lc = d->materials[v0.materialIndex].lightColor;
                float4 *o = &emit[6 * (r * RTX_SIZE_TRI_RANDOM + k)];
                o[0] = make_float4(v0.pos[0], v0.pos[1], v0.pos[2], lc[0]);
                o[1] = make_float4(v1.pos[0], v1.pos[1], v1.pos[2], lc[1]);
                o[2] = make_float4(v2.pos[0], v2.pos[1], v2.pos[2], lc[2]);
                o[3] = make_float4(v0.normal[0], v0.normal[1], v0.normal[2], 0.0f);
                o[4] = make_float4(v1.normal[0], v1.normal[1], v1.normal[2], 0.0f);
                o[5] = make_float4(v2.normal[0], v2.normal[1], v2.normal[2], 0.0f);
            }
        }
        e = upload(&emitD, emit.data(), emit.size());
        if (e == hipSuccess) e = upload(&linstD, linst.data(), linst.size());
    }
    // textures: the environment map as float4 texels, material textures packed
    // one after another, the sRGB decoding table of the R8G8B8A8_SRGB images
    float4 *envD = nullptr;
    uint32_t *texD = nullptr;
    int4 *texInfoD = nullptr;
    float *srgbD = nullptr;
    std::vector<int4> texInfo(std::max<uint32_t>(d->textureCount, 1), make_int4(0, 0, 0, 0));
    std::vector<uint32_t> texels;
    for (uint32_t t = 1; t < d->textureCount; ++t) {
        const rtx_texture_desc &td = d->textures[t];
        texInfo[t] = make_int4((int)texels.size(), (int)td.width, (int)td.height, 0);
        const size_t n = (size_t)td.width * td.height;
        texels.resize(texels.size() + n);
        std::memcpy(texels.data() + texels.size() - n, td.rgba, n * 4);
    }
    std::vector<float> srgb(256);
    for (int i = 0; i < 256; ++i) srgb[i] = srgbDecode(i);
    if (e == hipSuccess && d->envMap)
        e = upload(&envD, reinterpret_cast<const float4 *>(d->envMap), (size_t)d->envWidth * d->envHeight);
    if (e == hipSuccess) e = upload(&texD, texels.data(), texels.size());
    if (e == hipSuccess) e = upload(&texInfoD, texInfo.data(), texInfo.size());
    if (e == hipSuccess) e = upload(&srgbD, srgb.data(), srgb.size());
    if (e != hipSuccess) {
        freeAll(sc->allocs);
        delete sc;
        return fail(ctx, RTX_E_DEVICE, std::string("scene upload: ") + hipGetErrorString(e));
    }
    S.nodes = nodes;
    S.tris = trisD;
    S.triRef = prD;
    S.vertices = vD;
    S.indices = iD;
    S.vertexBase = vbD;
    S.indexBase = ibD;
    S.instances = instD;
    S.materials = matD;
    S.lights = lightD;
    S.lightIndex = liD;
    S.triTable = ftD;
    S.emitTris = emitD;
    S.lightInst = linstD;
    S.lightTris = lightTrisOk ? lightTrisD : nullptr;
    S.numLightTris = lightTrisOk ? (uint32_t)(lightTris.size() / 3) : 0u;
    S.lightFirstOk = lightTrisOk ? 1u : 0u;
    S.spheres = spD;
    S.numTris = (uint32_t)numTris;
    S.numSpheres = d->sphereCount;
    S.lightCount = d->lightCount;
    S.triRows = d->randomTriRows;
    S.numNodes = (uint32_t)bvh.nodes.size();
    S.env = envD;
    S.envW = d->envMap ? d->envWidth : 0;
    S.envH = d->envMap ? d->envHeight : 0;
    S.texels = texD;
    S.texInfo = texInfoD;
    S.numTextures = d->textureCount;
    S.alphaTris = alphaTris;
    S.srgbToLinear = srgbD;
    sc->info.triangles = (uint32_t)numTris;
    sc->info.spheres = d->sphereCount;
    sc->info.nodes = (uint32_t)bvh.nodes.size();
    sc->info.maxDepth = bvh.maxDepth;
    sc->info.deviceBytes = bvh.nodes.size() * sizeof(Bvh4Node) + tris.size() * sizeof(float4) +
                           primRef.size() * sizeof(int4) + verts.size() * sizeof(rtx_vertex) +
                           idx.size() * 4 + (size_t)d->materialCount * sizeof(rtx_material);
    sc->info.buildMs = bvh.buildMs;
    for (int a = 0; a < 3; ++a) {
        sc->info.boundsMin[a] = std::numeric_limits<float>::infinity();
        sc->info.boundsMax[a] = -std::numeric_limits<float>::infinity();
    }
    for (uint32_t i = 0; i < d->instanceCount; ++i) {
        const rtx_instance &in = d->instances[i];
        const rtx_model_desc &md = d->models[in.modelIndex];
        for (uint32_t v = 0; v < md.vertexCount; ++v) {
            float w[3];
            xformPoint(in.transform, md.vertices[v].pos, w);
            for (int a = 0; a < 3; ++a) {
                sc->info.boundsMin[a] = std::min(sc->info.boundsMin[a], w[a]);
                sc->info.boundsMax[a] = std::max(sc->info.boundsMax[a], w[a]);
            }
        }
    }
    for (uint32_t i = 0; i < d->sphereCount; ++i)
        for (int a = 0; a < 3; ++a) {
            sc->info.boundsMin[a] = std::min(sc->info.boundsMin[a], d->spheres[i].center[a] - d->spheres[i].radius);
            sc->info.boundsMax[a] = std::max(sc->info.boundsMax[a], d->spheres[i].center[a] + d->spheres[i].radius);
        }
    *out = sc;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_scene_destroy(rtx_ctx *ctx, rtx_scene *scene) try {
    if (!scene) return RTX_E_INVALID;
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    freeAll(scene->allocs);
    delete scene;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_scene_get_info(const rtx_scene *scene, rtx_scene_info *info) try {
    if (!scene || !info) return RTX_E_INVALID;
    *info = scene->info;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(nullptr);
}

static int guideUpdateAsync(rtx_ctx *ctx);
static int nrcStep(rtx_ctx *ctx, const float4 *in, const float4 *target, uint32_t n);

int rtx_render(rtx_ctx *ctx, const rtx_scene *scene, const rtx_frame_params *p, rtx_frame_stats *stats) try {
    if (!ctx) return RTX_E_INVALID;
    if (!scene || !p) return fail(ctx, RTX_E_INVALID, "NULL scene or params");
    if (p->width == 0 || p->height == 0 || p->spp == 0) return fail(ctx, RTX_E_INVALID, "empty render");
    if (const char *f = unsupportedFeature(p->pc))
        return fail(ctx, RTX_E_UNSUPPORTED, std::string("feature not supported yet: ") + f);
    if (p->pc.maxDepth < 0 || p->pc.maxDepth > 200 || p->pc.maxFollowDiscrete < 0 || p->pc.maxFollowDiscrete > 50)
        return fail(ctx, RTX_E_INVALID, "maxDepth must be in [0, 200] and maxFollowDiscrete in [0, 50]");
    if (p->batchPaths > (1u << 26)) return fail(ctx, RTX_E_INVALID, "batchPaths must be at most 2^26");
    const bool useIC = p->pc.useIrradianceCache != 0;
    const bool useADRRS = p->pc.useADRRS != 0;
    // frames that look the cache up and grow it: useIrradianceCache, or ADRRS
    // (lookups from depth 2 for the weight window, creations where none is found)
    const bool icFrames = useIC || useADRRS;
    const bool splitting = (useADRRS && p->pc.adrrsSplit) || p->pc.splitOnFirst;
    // a cache hit runs multipleNEE even with NEE off (raytrace.rgen:1120-1121)
    const bool needNEE = p->pc.enableNEE || useIC;
    uint32_t numNEE = needNEE ? (uint32_t)p->pc.numNEE : 1u;
    if (needNEE && (p->pc.numNEE < 1 || p->pc.numNEE > 8))
        return fail(ctx, RTX_E_INVALID, "numNEE must be in [1, 8]");
    if (p->pc.samplesPerPixel < 0 || p->pc.samplesPerPixel > 4096)
        return fail(ctx, RTX_E_INVALID, "samplesPerPixel must be in [0, 4096]");
    if (p->pc.splitOnFirst && icFrames)
        // the cache's own sub-paths would queue splits on the pixel's list
        return fail(ctx, RTX_E_UNSUPPORTED, "splitOnFirst together with the irradiance cache or ADRRS");
    if (useADRRS) {
        if (!(p->pc.adrrsS >= 1.0f)) return fail(ctx, RTX_E_INVALID, "adrrsS must be >= 1");
        if (!ctx->estimate || ctx->estW != p->width || ctx->estH != p->height)
            return fail(ctx, RTX_E_STATE,
                        "useADRRS needs a pixel estimate of the frame's size (a storeEstimate render or "
                        "rtx_estimate_write)");
    }
    uint32_t neeCap = numNEE;
    if (icFrames) {
        if (!ctx->icValid)
            return fail(ctx, RTX_E_STATE, "useIrradianceCache and useADRRS need rtx_ic_create first");
        if (p->pc.irradianceNumNEE < 1 || p->pc.irradianceNumNEE > 8)
            return fail(ctx, RTX_E_INVALID, "irradianceNumNEE must be in [1, 8]");
        if (!(p->pc.irradianceA > 0.0f)) return fail(ctx, RTX_E_INVALID, "irradianceA must be > 0");
        neeCap = std::max(neeCap, (uint32_t)p->pc.irradianceNumNEE);
    }
    const bool nrc = (p->flags & RTX_FRAME_NRC) != 0;
    if (nrc) {
        if (!ctx->nrcValid) return fail(ctx, RTX_E_STATE, "RTX_FRAME_NRC needs rtx_nrc_create first");
        // guided sampling (useGuiding) combines with it; recording for the guide's update does not
        if (icFrames || splitting || p->pc.updateGuiding || p->pc.samplesPerPixel > 1)
            return fail(ctx, RTX_E_UNSUPPORTED, "the learned radiance cache with the irradiance cache, ADRRS, "
                                                "splitting, guiding updates or several samples per frame");
    }
    if ((p->pc.useGuiding || p->pc.updateGuiding) && ctx->guide.numRegions == 0)
        return fail(ctx, RTX_E_STATE, "useGuiding/updateGuiding need guiding regions (rtx_guiding_set)");
    if (p->pc.useGuiding && !(p->pc.guidingProb >= 0.0f && p->pc.guidingProb <= 1.0f))
        return fail(ctx, RTX_E_INVALID, "guidingProb must be in [0, 1]");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    const uint64_t npix = (uint64_t)p->width * p->height;
    const uint64_t total = npix * p->spp;
    if (npix > (1ull << 31)) return fail(ctx, RTX_E_INVALID, "image too large");

    // accumulation image
    bool reset = (p->flags & RTX_FRAME_RESET_ACCUM) || p->width != ctx->accW || p->height != ctx->accH;
    if (p->width != ctx->accW || p->height != ctx->accH || !ctx->accum) {
        if (ctx->accum) (void)hipFree(ctx->accum);
        if (ctx->resolved) (void)hipFree(ctx->resolved);
        ctx->accum = ctx->resolved = nullptr;
        CHECK_HIP(ctx, devMalloc(&ctx->accum, npix * sizeof(float4)));
        CHECK_HIP(ctx, devMalloc(&ctx->resolved, npix * sizeof(float4)));
        ctx->accW = p->width;
        ctx->accH = p->height;
    }
    if (reset) {
        CHECK_HIP(ctx, hipMemsetAsync(ctx->accum, 0, npix * sizeof(float4), ctx->stream));
        ctx->accSamples = 0;
    }
    // irradiance-cache prepare frames only build the cache (src/RayTracingApp.cpp:160-161)
    const bool accumulate = !(useIC && p->pc.isIrradiancePrepareFrame);
    // camera samples per pixel and frame (raytrace.rgen:1662)
    const uint32_t samplesPerFrame =
        (useIC && p->pc.isIrradiancePrepareFrame) ? 1u : (uint32_t)std::max(1, p->pc.samplesPerPixel);
    const bool segments = samplesPerFrame > 1 || splitting;

    // batch size: whole sample planes when they fit, else pixel ranges; with
    // the cache one frame (sample plane) at a time, each followed by its upkeep
    uint64_t batch = p->batchPaths ? p->batchPaths : ctx->tuning.batchPaths;
    batch = std::min<uint64_t>(batch, total);
    const bool record = p->pc.updateGuiding != 0;
    if (icFrames || record || nrc) batch = std::min<uint64_t>(batch, npix);
    if (batch >= npix) batch -= batch % npix;
    size_t bufPaths = (size_t)batch;
    // with the cache, room for every job a frame can make (2 x maxCaches, 200
    // sub-paths each) when that stays within 2^26 paths: the frame's upkeep
    // then runs in one batch sized on the device
    if (icFrames)
        bufPaths = std::max<size_t>(bufPaths, std::min<size_t>((size_t)ctx->ic.jobCap * kIcStrata, (size_t)1 << 26));
    // With the cache on, a frame's update jobs are known once its camera
    // paths are generated (the elections are the paths' first draws and the
    // records they update are the frame-start cache's), so their sub-paths
    // can run in the frame's own bounces, beside the camera paths, as a second
    // class of paths with ids from `batch` on; only the creation jobs, which
    // need the frame's candidates, run after it.  The reads are the
    // frame-start cache's either way and every job keeps its index, streams
    // and place in the commit, so the cache and the image are the same as
    // with every job after the frame.
    const bool keepFlag = (p->flags & RTX_FRAME_KEEP_PATHS) != 0;
    const size_t mergedPaths = (size_t)batch + (size_t)ctx->ic.maxEntries * kIcStrata;
    const bool mergeUpkeep = useIC && ctx->tuning.cacheUpkeep == 1 && batch == npix && !keepFlag && !segments &&
                             !record && !nrc && mergedPaths <= ((size_t)1 << 26);
    if (mergeUpkeep) bufPaths = std::max(bufPaths, mergedPaths);
    int rc = ensureBatchBuffers(ctx, bufPaths, neeCap);
    if (rc != RTX_OK) return rc;
    if (segments && ctx->segCap < bufPaths) {
        freeAll(ctx->segAllocs);
        ctx->segCap = 0;
        CHECK_HIP(ctx, devAlloc(&ctx->pb.task, bufPaths, ctx->segAllocs));
        CHECK_HIP(ctx, devAlloc(&ctx->pb.taskRes, bufPaths, ctx->segAllocs));
        ctx->segCap = bufPaths;
    }
    if (splitting && ctx->splitCap < bufPaths) {
        freeAll(ctx->splitAllocs);
        ctx->splitCap = 0;
        CHECK_HIP(ctx, devAlloc(&ctx->pb.splitRec, bufPaths * kMaxSplits * 4, ctx->splitAllocs));
        ctx->splitCap = bufPaths;
    }
    ctx->pb.estimate = useADRRS ? ctx->estimate : nullptr;
    const bool storeEstimate = p->pc.storeEstimate != 0;
    if (storeEstimate && (!ctx->estimate || ctx->estW != p->width || ctx->estH != p->height)) {
        if (ctx->estimate) (void)hipFree(ctx->estimate);
        ctx->estimate = nullptr;
        ctx->estW = ctx->estH = 0;
        CHECK_HIP(ctx, devMalloc(&ctx->estimate, npix * sizeof(float4)));
        CHECK_HIP(ctx, hipMemsetAsync(ctx->estimate, 0, npix * sizeof(float4), ctx->stream));
        ctx->estW = p->width;
        ctx->estH = p->height;
    }
    if (icFrames && ctx->candCap < bufPaths) {
        if (ctx->candBuf) (void)hipFree(ctx->candBuf);
        ctx->candBuf = nullptr;
        ctx->candCap = 0;
        CHECK_HIP(ctx, devMalloc(&ctx->candBuf, bufPaths * kMaxNewCacheEntries * 2 * sizeof(float4)));
        ctx->candCap = bufPaths;
    }
    ctx->pb.cand = ctx->candBuf;
    // learned radiance cache: one query and one training record per path of a
    // batch, training slots for a frame
    uint32_t nrcEvery = 1, nrcSlots = 0;
    if (nrc) {
        if (ctx->nrcPathCap < bufPaths) {
            freeAll(ctx->nrcPathAllocs);
            ctx->nrcPathCap = 0;
            auto &L = ctx->nrcPathAllocs;
            CHECK_HIP(ctx, devAlloc(&ctx->pb.nrcQuery, 4 * bufPaths, L));
            CHECK_HIP(ctx, devAlloc(&ctx->pb.nrcQueryAux, bufPaths, L));
            CHECK_HIP(ctx, devAlloc(&ctx->pb.nrcPath, 2 * bufPaths, L));
            CHECK_HIP(ctx, devAlloc(&ctx->pb.nrcQueryCount, 1, L));
            ctx->nrcPathCap = bufPaths;
        }
        nrcEvery = (uint32_t)std::max<uint64_t>(1, npix / ctx->nrcCfg.trainingPaths);
        nrcSlots = (uint32_t)((npix + nrcEvery - 1) / nrcEvery + 1);
        if (ctx->nrcSlotCap < nrcSlots) {
            freeAll(ctx->nrcSlotAllocs);
            ctx->nrcSlotCap = 0;
            CHECK_HIP(ctx, devAlloc(&ctx->pb.nrcTrainIn, 4 * (size_t)nrcSlots, ctx->nrcSlotAllocs));
            CHECK_HIP(ctx, devAlloc(&ctx->pb.nrcTrainTarget, nrcSlots, ctx->nrcSlotAllocs));
            ctx->nrcSlotCap = nrcSlots;
        }
    }
    if (record) {
        if (ctx->recCap < bufPaths) {
            freeAll(ctx->recAllocs);
            ctx->recCap = 0;
            auto &L = ctx->recAllocs;
            CHECK_HIP(ctx, devAlloc(&ctx->pb.grec, bufPaths, L));
            CHECK_HIP(ctx, devAlloc(&ctx->pb.grecThr, bufPaths, L));
            CHECK_HIP(ctx, devAlloc(&ctx->pb.lightSum, bufPaths * kMaxRecords, L));
            CHECK_HIP(ctx, devAlloc(&ctx->pb.sampleThr, bufPaths * kMaxRecords, L));
            ctx->recCap = bufPaths;
        }
        const size_t slots = (size_t)npix * kMaxRecords;
        if (ctx->dirCap < slots) {
            if (ctx->dirData) (void)hipFree(ctx->dirData);
            ctx->dirData = nullptr;
            ctx->dirCap = 0;
            CHECK_HIP(ctx, devMalloc(&ctx->dirData, slots * sizeof(rtx_directional_data)));
            ctx->dirCap = slots;
        }
        ctx->pb.dirData = ctx->dirData;
        ctx->dirCount = slots;
    }

    const bool keep = (p->flags & RTX_FRAME_KEEP_PATHS) != 0;
    if (keep) {
        if (total > ctx->keepCap) {
            if (ctx->keepRgb) (void)hipFree(ctx->keepRgb);
            if (ctx->keepSig) (void)hipFree(ctx->keepSig);
            ctx->keepRgb = nullptr;
            ctx->keepSig = nullptr;
            ctx->keepCap = 0;
            CHECK_HIP(ctx, devMalloc(&ctx->keepRgb, total * sizeof(float4)));
            CHECK_HIP(ctx, devMalloc(&ctx->keepSig, total * sizeof(uint2)));
            ctx->keepCap = total;
        }
        ctx->keptPaths = total;
    }

    if (record && splitting) {
        // main() returns right after resetting the records when splitting is on
        // (raytrace.rgen:1641-1648): no pixel is traced or written
        CHECK_HIP(ctx, hipMemsetAsync(ctx->dirData, 0xFF, (size_t)npix * kMaxRecords * sizeof(rtx_directional_data),
                                      ctx->stream));
        if (keep) {
            CHECK_HIP(ctx, hipMemsetAsync(ctx->keepRgb, 0, total * sizeof(float4), ctx->stream));
            CHECK_HIP(ctx, hipMemsetAsync(ctx->keepSig, 0, total * sizeof(uint2), ctx->stream));
        }
        CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->paths = total;
        }
        return RTX_OK;
    }

    FrameConst fc;
    std::memset(&fc, 0, sizeof(fc));
    fc.pc = p->pc;
    fc.cam = p->cam;
    fc.width = p->width;
    fc.height = p->height;
    fc.seed = p->seed;
    fc.sampleOffset = p->sampleOffset;
    fc.rrStartDepth = p->rrStartDepth;
    fc.numNEE = numNEE;
    fc.neeStride = numNEE;
    if (mergeUpkeep) fc.neeStride = neeCap; // both classes of paths in one set of NEE slots
    fc.countVisits = (p->flags & RTX_FRAME_COUNT_VISITS) ? 1 : 0;
    // developer aid (a build with -DRTX_DEVICE_DEBUG): RTX_DEBUG_PATH=<path index
    // within this call> prints that path's bounces
    fc.debugPath = ~0ull;
    if (const char *dbg = std::getenv("RTX_DEBUG_PATH")) fc.debugPath = std::strtoull(dbg, nullptr, 10);
    // camera paths: raytrace(..., maxDepth, maxFollowDiscrete, !enableNEE, true, enableNEE, numNEE,
    // useIrradianceCache, createIC = true, ...) (raytrace.rgen:1671)
    fc.maxDepth = p->pc.maxDepth;
    fc.maxFollow = p->pc.maxFollowDiscrete;
    fc.addDirect = p->pc.enableNEE ? 0u : 1u;
    fc.useNEE = p->pc.enableNEE ? 1u : 0u;
    fc.useIC = useIC ? 1u : 0u;
    fc.createIC = icFrames ? 1u : 0u; // candidates arise only where the cache is looked up
    fc.allowRR = 1;
    fc.record = record ? 1u : 0u;
    fc.samplesPerFrame = samplesPerFrame;
    fc.segments = segments ? 1u : 0u;
    fc.useADRRS = useADRRS ? 1u : 0u;
    fc.splitOnFirst = p->pc.splitOnFirst ? 1u : 0u;
    fc.nrc = nrc ? 1u : 0u;
    fc.nrcDepth = nrc ? ctx->nrcCfg.terminateBounces + 1u : 0u;
    fc.nrcSuffix = nrc ? ctx->nrcCfg.trainSuffixBounces : 0u;
    fc.probeFirstLight = 1u; // taken where the scene has few emitters and no environment map
    const bool probesLightFirst = scene->view.lightFirstOk && scene->view.env == nullptr;
    fc.nrcTrainEvery = nrcEvery;
    fc.nrcTrainSlots = nrcSlots;
    CHECK_HIP(ctx, hipMemcpyAsync(ctx->fc, &fc, sizeof(fc), hipMemcpyHostToDevice, ctx->stream));
    // the cache's stratum sub-paths: raytrace(..., 1, 10, false, false, true, irradianceNumNEE,
    // true, false, ...) (raytrace.rgen:1260)
    FrameConst fcIC = fc;
    fcIC.maxDepth = 1;
    fcIC.maxFollow = 10;
    fcIC.addDirect = 0;
    fcIC.useNEE = 1;
    fcIC.useIC = 1;
    fcIC.createIC = 0;
    fcIC.allowRR = 0;
    fcIC.record = 0;
    fcIC.samplesPerFrame = 1;
    fcIC.segments = 0;
    fcIC.useADRRS = 0;
    fcIC.splitOnFirst = 0;
    fcIC.nrc = 0;
    fcIC.numNEE = icFrames ? (uint32_t)p->pc.irradianceNumNEE : 1u;
    fcIC.neeStride = mergeUpkeep ? neeCap : fcIC.numNEE;
    fcIC.debugPath = ~0ull;
    if (icFrames) CHECK_HIP(ctx, hipMemcpyAsync(ctx->fcIC, &fcIC, sizeof(fcIC), hipMemcpyHostToDevice, ctx->stream));
    CHECK_HIP(ctx, hipMemsetAsync(ctx->cnt, 0, sizeof(Counters), ctx->stream));

    const bool countVisits = fc.countVisits != 0;
    const bool timeKernels = (p->flags & RTX_FRAME_TIME_KERNELS) != 0;
    const bool timeTrace = timeKernels || (p->flags & RTX_FRAME_TIME_TRACE) != 0;
    ctx->eventsUsed = 0;
    const int traceBlocks = ctx->cuCount * ctx->tuning.traceBlocksPerCU;
    const int shadeBlocks = ctx->cuCount * ctx->tuning.shadeBlocksPerCU;
    const int maxIter = p->pc.maxDepth + p->pc.maxFollowDiscrete + 3;
    const int maxIterIC = 1 + 10 + 3;
    const DevScene &S = scene->view;
    const DevIC icv = icView(ctx->ic);
    const float c0 = icBaseCell(p->pc);
    hipStream_t s = ctx->stream;

    uint32_t launches[4] = {0, 0, 0, 0};
    auto timed = [&](int cls, auto &&launch) {
        launches[cls]++;
        const bool trace = cls == kClsExtend || cls == kClsShadow;
        EventPair *ev = timeKernels || (timeTrace && trace) ? nextEvents(ctx, cls) : nullptr;
        if (ev) (void)hipEventRecord(ev->a, s);
        launch();
        if (ev) (void)hipEventRecord(ev->b, s);
    };
    // The bounce iterations of one batch whose extend queue 0 is filled.
    // Bounce `it` traces its extend queue together with the previous
    // bounce's shadow and probe rays (one persistent launch), resolves the
    // probes, finishes the paths the previous bounce ended, then shades its
    // own hits, which fills the next extend queue and the secondary queues.
    // A path ends within `iters` iterations of its start, and one more
    // bounce drains its last secondary rays.  With segments a path slot runs
    // a pixel's whole frame (several camera samples, queued splits), so the
    // loop goes on until every queue is empty.  The host learns the queue
    // lengths each bounce started from through the poll ring, kPollLead
    // bounces late, so the stream never waits for the host: once a bounce
    // is known to be followed only by empty ones, the loop ends (the bounces
    // launched meanwhile run over empty queues).
    int itersRun = 0;
    // where the any-hit alpha test finds each ray's frame seed: extend entries
    // are path ids, shadow and probe entries path * numNEE + j
    auto setRaySeeds = [&](uint32_t nee) {
        for (Queue *q : {&ctx->q[0], &ctx->q[1], &ctx->qs, &ctx->qp}) {
            q->pathSeed = ctx->pb.frameSeed;
            q->seedDiv = q == &ctx->qs || q == &ctx->qp ? nee : 1u;
        }
    };
    // developer aid: RTX_LOG_BOUNCES=1 prints the queue lengths of every bounce
    static const bool logBounces = std::getenv("RTX_LOG_BOUNCES") != nullptr;
    auto bounces = [&](const FrameConst *fcd, BatchDesc b, int iters, bool multi,
                       const PathClass2 *second = nullptr) -> int {
        int qsel = 0;
        const int cap = multi ? (int)(samplesPerFrame + kMaxSplits + 1) * (iters + 1) : iters + 1;
        for (int it = 0;; ++it) {
            if (it == cap && !multi) break;
            if (it >= kPollLead) {
                const int j = (it - kPollLead) % kPollSlots;
                CHECK_HIP(ctx, hipEventSynchronize(ctx->pollEv[j]));
                const uint32_t *c = ctx->pollRing + 4 * j;
                if (logBounces) std::fprintf(stderr, "[rtx] bounce %d: extend %u probe %u shadow %u done %u\n",
                                             it - kPollLead, c[0], c[1], c[2], c[3]);
                // c = lengths bounce j started from: extend, probe, shadow, done.
                // With no extend rays, bounce j shades nothing, so bounce j + 1
                // starts empty unless kFinish restarts segments of j's done paths.
                if (c[0] == 0 && (!multi || c[3] == 0)) break;
                if (it >= cap) return fail(ctx, RTX_E_DEVICE, "paths did not finish within the iteration bound");
            }
            uint32_t *poll = ctx->pollRing + 4 * (it % kPollSlots);
            itersRun = std::max(itersRun, it + 1);
            if (it == 0) timed(kClsShade, [&] { launchResetCounters(s, ctx->cnt, qsel ^ 1); });
            timed(kClsExtend, [&] {
                launchTraceP(s, S, ctx->q[qsel], qsel, ctx->qs, ctx->qp, ctx->pb, ctx->cnt, poll, countVisits,
                             traceBlocks, ctx->tuning.ldsStackEntries);
            });
            CHECK_HIP(ctx, hipEventRecord(ctx->pollEv[it % kPollSlots], s));
            if (!probesLightFirst) // otherwise every probe traced an occlusion ray, and the probe queue is empty
                timed(kClsProbe, [&] { launchProbeResolve(s, S, fcd, ctx->qp, ctx->pb, ctx->cnt, qsel, shadeBlocks); });
            timed(kClsShade, [&] {
                launchFinish(s, fcd, b, ctx->pb, ctx->done, ctx->cnt, ctx->q[qsel ^ 1], qsel, shadeBlocks, second);
            });
            timed(kClsShade, [&] {
                launchShade(s, S, icv, ctx->guide, fcd, b, ctx->pb, ctx->q[qsel], ctx->q[qsel ^ 1], ctx->qs, ctx->qp,
                            ctx->done, ctx->cnt, qsel, shadeBlocks, second);
            });
            qsel ^= 1;
        }
        return RTX_OK;
    };

    CHECK_HIP(ctx, hipEventRecord(ctx->tStart, s));
    uint32_t batches = 0;
    if (icFrames) timed(kClsShade, [&] { launchIcIndex(s, ctx->ic, c0); });
    const uint64_t frameLen = (icFrames || record || nrc) ? npix : total; // paths per frame with per-frame state
    for (uint64_t f0 = 0; f0 < total; f0 += frameLen) {
        const uint64_t fEnd = std::min(total, f0 + frameLen);
        if (icFrames) timed(kClsShade, [&] { launchIcFrameBegin(s, ctx->ic); });
        // every training slot starts the frame without a target (w = 0) and
        // with zero inputs: a slot no path of this frame claims still enters
        // the training batch, where its weight gradient is its (zero) delta
        // times its inputs' activations, so those must be finite
        if (nrc)
            timed(kClsShade, [&] {
                launchNrcZero(s, &ctx->pb.nrcTrainTarget[0].x, 4 * (size_t)nrcSlots);
                launchNrcZero(s, &ctx->pb.nrcTrainIn[0].x, 16 * (size_t)nrcSlots);
            });
        for (uint64_t first = f0; first < fEnd; first += batch) {
            BatchDesc b{};
            b.firstPath = first;
            b.numPaths = (uint32_t)std::min<uint64_t>(batch, fEnd - first);
            b.npix = (uint32_t)npix;
            ctx->pb.dirFirstPixel = (uint32_t)(first % npix);
            timed(kClsShade, [&] { launchGenerate(s, ctx->fc, b, ctx->pb, ctx->q[0], ctx->cnt, 0, ctx->ic.frame); });
            int rcb = RTX_OK;
            if (mergeUpkeep) {
                // the update jobs, planned from the elections just drawn; their
                // sub-paths join the camera paths' first extend queue
                const uint32_t frameSeed = pcg32First(p->seed, f0 / npix + p->sampleOffset);
                const uint32_t np = ctx->ic.maxEntries * kIcStrata;
                timed(kClsShade, [&] { launchIcPlan(s, ctx->ic, kIcPlanUpdates); });
                timed(kClsShade, [&] {
                    launchGenerateIC(s, ctx->ic, frameSeed, 0, np, ctx->pb, ctx->q[0], ctx->cnt, 0, kIcJobsUpdates,
                                     (uint32_t)batch, b.numPaths);
                });
                PathClass2 sub{ctx->fcIC, BatchDesc{}, (uint32_t)batch};
                sub.b.numPaths = np;
                sub.b.npix = np;
                sub.b.pathBase = (uint32_t)batch;
                setRaySeeds(neeCap);
                rcb = bounces(ctx->fc, b, std::max(maxIter, maxIterIC), false, &sub);
            } else {
                setRaySeeds(fc.numNEE);
                rcb = bounces(ctx->fc, b, maxIter, segments);
            }
            if (rcb != RTX_OK) return rcb;
            if (nrc) // the paths that ended into the learned cache: throughput x cache
                timed(kClsShade, [&] {
                    launchNrcInfer(s, ctx->nrc, ctx->pb.nrcQuery, b.numPaths, ctx->pb.nrcQueryCount, nullptr,
                                   ctx->pb.nrcQueryAux, ctx->pb.out, 2 * ctx->cuCount,
                                   NrcSelfTrain{ctx->pb.thr, ctx->pb.nrcPath,
                                                fc.nrcSuffix ? ctx->pb.nrcTrainTarget : nullptr, kFlagNrcTrainEnd});
                });
            timed(kClsShade, [&] {
                launchAccumulate(s, b, ctx->pb, accumulate ? ctx->accum : nullptr, ctx->keepRgb, ctx->keepSig, keep,
                                 storeEstimate ? ctx->estimate : nullptr);
            });
            if (icFrames) timed(kClsShade, [&] { launchIcCollect(s, ctx->ic, ctx->pb, b.numPaths); });
            CHECK_HIP(ctx, hipGetLastError());
            batches++;
        }
        if (nrc) { // the learned cache trains on the frame's training vertices
            const uint32_t B = ctx->nrcCfg.trainBatch;
            const uint32_t rows = std::min(nrcSlots, B * ctx->nrcCfg.trainSteps);
            launchNrcZero(s, ctx->nrcLoss, 1);
            ctx->nrcLossSteps = 0;
            for (uint32_t r0 = 0; r0 < rows; r0 += B) {
                const uint32_t nr = std::min(B, rows - r0);
                int rcn = RTX_OK;
                timed(kClsShade, [&] {
                    rcn = nrcStep(ctx, ctx->pb.nrcTrainIn + 4 * (size_t)r0, ctx->pb.nrcTrainTarget + r0, nr);
                });
                if (rcn != RTX_OK) return rcn;
                ctx->nrcLossSteps++;
            }
        }
        // PathGuiding::update after every recording frame when the context owns the mixtures
        // (src/RayTracingApp.cpp:143-151)
        if (record && ctx->fitValid && !(p->flags & RTX_FRAME_NO_GUIDING_UPDATE)) {
            int rcu = 0;
            timed(kClsShade, [&] { rcu = guideUpdateAsync(ctx); });
            if (rcu != RTX_OK) return rcu;
        }
        if (!icFrames) continue;
        // the frame's cache upkeep
        const uint64_t sample = f0 / npix + p->sampleOffset;
        const uint32_t frameSeed = pcg32First(p->seed, sample);
        if (mergeUpkeep) {
            // the update jobs' records from their sub-paths (ids from batch on),
            // then the creation jobs behind them, sized on the device
            const uint32_t np = ctx->ic.maxEntries * kIcStrata;
            timed(kClsShade, [&] {
                launchIcReduce(s, ctx->ic, frameSeed, 0, ctx->ic.maxEntries, ctx->pb, kIcJobsUpdates, (uint32_t)batch);
            });
            // Once an earlier commit left every record slot in use (the flag
            // kIcCommit sets in pinned memory; a value not yet seen reads as 0)
            // the frame has no creation job, and its creation loop would run
            // empty bounces: it is skipped (the plan's update part already set
            // the job count to the updates').
            const bool cacheFull = *reinterpret_cast<volatile uint32_t *>(ctx->ic.fullHost) != 0u;
            if (!cacheFull) {
                timed(kClsShade, [&] { launchIcPlan(s, ctx->ic, kIcPlanCreations); });
                BatchDesc b{};
                b.numPaths = np;
                b.npix = np;
                timed(kClsShade, [&] {
                    launchGenerateIC(s, ctx->ic, frameSeed, 0, np, ctx->pb, ctx->q[0], ctx->cnt, 0, kIcJobsCreations);
                });
                setRaySeeds(neeCap);
                int rcs = bounces(ctx->fcIC, b, maxIterIC, false);
                if (rcs != RTX_OK) return rcs;
                timed(kClsShade, [&] {
                    launchIcReduce(s, ctx->ic, frameSeed, 0, ctx->ic.maxEntries, ctx->pb, kIcJobsCreations);
                });
            }
            timed(kClsShade, [&] { launchIcCommit(s, ctx->ic, ctx->fcIC); });
            timed(kClsShade, [&] { launchIcIndex(s, ctx->ic, c0); });
            CHECK_HIP(ctx, hipGetLastError());
            continue;
        }
        timed(kClsShade, [&] { launchIcPlan(s, ctx->ic); });
        const bool keepSub = keep && fEnd == total;
        const uint32_t jobsPerBatch = (uint32_t)std::max<size_t>(1, ctx->cap / kIcStrata);
        if (!keepSub && ctx->ic.jobCap <= jobsPerBatch) {
            // every job of the frame fits one batch: the launches take the job
            // count from device memory, so no host round trip sizes them
            const uint32_t np = ctx->ic.jobCap * kIcStrata;
            BatchDesc b{};
            b.numPaths = np;
            b.npix = np;
            timed(kClsShade,
                  [&] { launchGenerateIC(s, ctx->ic, frameSeed, 0, np, ctx->pb, ctx->q[0], ctx->cnt, 0, kIcJobsFromFirst); });
            setRaySeeds(fcIC.numNEE);
            int rcs = bounces(ctx->fcIC, b, maxIterIC, false);
            if (rcs != RTX_OK) return rcs;
            timed(kClsShade,
                  [&] { launchIcReduce(s, ctx->ic, frameSeed, 0, ctx->ic.jobCap, ctx->pb, kIcJobsFromFirst); });
            timed(kClsShade, [&] { launchIcCommit(s, ctx->ic, ctx->fcIC); });
            timed(kClsShade, [&] { launchIcIndex(s, ctx->ic, c0); });
            CHECK_HIP(ctx, hipGetLastError());
            continue;
        }
        uint32_t jobInfo[2];
        CHECK_HIP(ctx, hipMemcpyAsync(jobInfo, ctx->ic.frame + 3, sizeof(jobInfo), hipMemcpyDeviceToHost, s));
        CHECK_HIP(ctx, hipStreamSynchronize(s));
        const uint32_t numJobs = jobInfo[1];
        if (keepSub) {
            size_t need = (size_t)numJobs * kIcStrata;
            if (need > ctx->keepSubCap) {
                if (ctx->keepSub) (void)hipFree(ctx->keepSub);
                if (ctx->keepSubSig) (void)hipFree(ctx->keepSubSig);
                ctx->keepSub = nullptr;
                ctx->keepSubSig = nullptr;
                ctx->keepSubCap = 0;
                CHECK_HIP(ctx, devMalloc(&ctx->keepSub, need * sizeof(float4)));
                CHECK_HIP(ctx, devMalloc(&ctx->keepSubSig, need * sizeof(uint2)));
                ctx->keepSubCap = need;
            }
            ctx->keptJobs = numJobs;
            ctx->keptUpdates = jobInfo[0];
        }
        for (uint32_t j0 = 0; j0 < numJobs; j0 += jobsPerBatch) {
            const uint32_t nj = std::min(jobsPerBatch, numJobs - j0);
            const uint32_t np = nj * kIcStrata;
            BatchDesc b{};
            b.firstPath = 0;
            b.numPaths = np;
            b.npix = np;
            timed(kClsShade,
                  [&] { launchGenerateIC(s, ctx->ic, frameSeed, j0, np, ctx->pb, ctx->q[0], ctx->cnt, 0, kIcJobsHost); });
            setRaySeeds(fcIC.numNEE);
            int rcs = bounces(ctx->fcIC, b, maxIterIC, false);
            if (rcs != RTX_OK) return rcs;
            timed(kClsShade, [&] { launchIcReduce(s, ctx->ic, frameSeed, j0, nj, ctx->pb, kIcJobsHost); });
            if (keepSub) {
                CHECK_HIP(ctx, hipMemcpyAsync(ctx->keepSub + (size_t)j0 * kIcStrata, ctx->pb.out, np * sizeof(float4),
                                              hipMemcpyDeviceToDevice, s));
                CHECK_HIP(ctx, hipMemcpyAsync(ctx->keepSubSig + (size_t)j0 * kIcStrata, ctx->pb.outSig,
                                              np * sizeof(uint2), hipMemcpyDeviceToDevice, s));
            }
            CHECK_HIP(ctx, hipGetLastError());
        }
        timed(kClsShade, [&] { launchIcCommit(s, ctx->ic, ctx->fcIC); });
        timed(kClsShade, [&] { launchIcIndex(s, ctx->ic, c0); });
        CHECK_HIP(ctx, hipGetLastError());
    }
    CHECK_HIP(ctx, hipEventRecord(ctx->tStop, s));
    CHECK_HIP(ctx, hipStreamSynchronize(s));
    if (accumulate) ctx->accSamples += p->spp;

    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        Counters c;
        CHECK_HIP(ctx, hipMemcpy(&c, ctx->cnt, sizeof(c), hipMemcpyDeviceToHost));
        stats->paths = total;
        stats->extendRays = c.extendRays;
        stats->shadowRays = c.shadowRays;
        stats->probeRays = c.probeRays;
        for (int k = 0; k < 3; ++k) {
            stats->nodeVisitsByKind[k] = c.nodeVisits[k];
            stats->primVisitsByKind[k] = c.primVisits[k];
            stats->nodeVisits += c.nodeVisits[k];
            stats->primVisits += c.primVisits[k];
        }
        for (int k = 0; k < 4; ++k) stats->launchesByKind[k] = launches[k];
        float ms = 0;
        CHECK_HIP(ctx, hipEventElapsedTime(&ms, ctx->tStart, ctx->tStop));
        stats->msTotal = ms;
        for (size_t k = 0; k < ctx->eventsUsed; ++k) {
            float e = 0;
            CHECK_HIP(ctx, hipEventElapsedTime(&e, ctx->events[k].a, ctx->events[k].b));
            switch (ctx->events[k].cls) {
            case kClsExtend: stats->msExtend += e; break;
            case kClsShadow: stats->msShadow += e; break;
            case kClsProbe: stats->msProbe += e; break;
            default: stats->msShade += e; break;
            }
        }
        stats->batches = batches;
        stats->bounces = (uint32_t)itersRun;
        stats->nodeBytes = kBvh4NodeBytesRead;
        stats->primBytes = 48;
    }
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

// ---------------------------------------------------------- path guiding
// A 16-lobe VMM_Theta (the reference's layout) as a device mixture: the
// lobes past usedDistributions zero, as kToVmm leaves them.
static rtx_vmm_theta_wide widen(const rtx_vmm_theta &v) {
    rtx_vmm_theta_wide w;
    std::memset(&w, 0, sizeof(w));
    for (int k = 0; k < RTX_MAX_LOBES; ++k) {
        w.thetas[k].mu[2] = 1.0f;
        w.thetas[k].distance = -1.0f;
    }
    for (int k = 0; k < RTX_MAX_DISTRIBUTIONS; ++k) {
        w.thetas[k] = v.thetas[k];
        w.pi[k] = v.pi[k];
    }
    for (int c = 0; c < 3; ++c) w.meanPosition[c] = v.meanPosition[c];
    w.usedDistributions = v.usedDistributions;
    return w;
}

static int guideUpload(rtx_ctx *ctx, const rtx_aabb *regions, const rtx_vmm_theta *vmms, uint32_t n) {
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    freeAll(ctx->guideAllocs);
    ctx->guide = DevGuiding{};
    ctx->guideBoxes.assign(regions, regions + n);
    ctx->guideVmms.resize(n);
    for (uint32_t i = 0; i < n; ++i) ctx->guideVmms[i] = widen(vmms[i]);
    ctx->guideVmmsStale = false;
    if (n == 0) return RTX_OK;
    std::vector<GuideNode> nodes = buildGuideNodes(regions, n);
    rtx_aabb *dBoxes = nullptr;
    rtx_vmm_theta_wide *dVmms = nullptr;
    GuideNode *dNodes = nullptr;
    hipError_t e = devAlloc(&dBoxes, n, ctx->guideAllocs);
    if (e == hipSuccess) e = devAlloc(&dVmms, n, ctx->guideAllocs);
    if (e == hipSuccess) e = devAlloc(&dNodes, nodes.size(), ctx->guideAllocs);
    if (e == hipSuccess) e = hipMemcpy(dBoxes, regions, n * sizeof(rtx_aabb), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(dVmms, ctx->guideVmms.data(), n * sizeof(rtx_vmm_theta_wide), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dNodes, nodes.data(), nodes.size() * sizeof(GuideNode), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        freeAll(ctx->guideAllocs);
        return fail(ctx, RTX_E_DEVICE, std::string("guiding upload: ") + hipGetErrorString(e));
    }
    ctx->guide.boxes = dBoxes;
    ctx->guide.vmms = dVmms;
    ctx->guide.nodes = dNodes;
    ctx->guide.numRegions = n;
    ctx->guide.numNodes = (uint32_t)nodes.size();
    return RTX_OK;
}

int rtx_guiding_set(rtx_ctx *ctx, const rtx_aabb *regions, const rtx_vmm_theta *vmms, uint32_t n) try {
    if (!ctx) return RTX_E_INVALID;
    if (n && (!regions || !vmms)) return fail(ctx, RTX_E_INVALID, "NULL argument");
    if (n > (1u << 20)) return fail(ctx, RTX_E_INVALID, "more than 2^20 guiding regions");
    for (uint32_t i = 0; i < n; ++i) {
        if (vmms[i].usedDistributions < 1 || vmms[i].usedDistributions > RTX_MAX_DISTRIBUTIONS)
            return fail(ctx, RTX_E_INVALID, "usedDistributions must be in [1, 16]");
        for (int a = 0; a < 3; ++a)
            if (!(regions[i].min[a] <= regions[i].max[a])) return fail(ctx, RTX_E_INVALID, "empty guiding region box");
    }
    // hand-set mixtures replace any fitted ones
    freeAll(ctx->fitAllocs);
    freeAll(ctx->fitWorkAllocs);
    ctx->fit = GuideFitBuffers{};
    ctx->fitValid = false;
    return guideUpload(ctx, regions, vmms, n);
} catch (...) {
    return rtx::abiFail(ctx);
}

// The host copy of the device mixtures, brought up to date.
static int guideSync(rtx_ctx *ctx) {
    const uint32_t count = (uint32_t)ctx->guideBoxes.size();
    if (ctx->guideVmmsStale && count) {
        CHECK_HIP(ctx, hipSetDevice(ctx->device));
        CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->guideVmms.resize(count);
        CHECK_HIP(ctx, hipMemcpy(ctx->guideVmms.data(), ctx->guide.vmms, count * sizeof(rtx_vmm_theta_wide),
                                 hipMemcpyDeviceToHost));
        ctx->guideVmmsStale = false;
    }
    return RTX_OK;
}

int rtx_guiding_get(rtx_ctx *ctx, rtx_aabb *regions, rtx_vmm_theta *vmms, uint32_t capacity, uint32_t *n) try {
    if (!ctx) return RTX_E_INVALID;
    const uint32_t count = (uint32_t)ctx->guideBoxes.size();
    if (n) *n = count;
    const int rc = guideSync(ctx);
    if (rc != RTX_OK) return rc;
    const uint32_t k = std::min(capacity, count);
    if (regions) std::copy(ctx->guideBoxes.begin(), ctx->guideBoxes.begin() + k, regions);
    if (vmms) {
        for (uint32_t i = 0; i < k; ++i) {
            const rtx_vmm_theta_wide &w = ctx->guideVmms[i];
            if (w.usedDistributions > RTX_MAX_DISTRIBUTIONS)
                return fail(ctx, RTX_E_UNSUPPORTED, "a mixture uses more than 16 lobes: rtx_guiding_get_wide");
            rtx_vmm_theta v;
            for (int c = 0; c < RTX_MAX_DISTRIBUTIONS; ++c) {
                v.thetas[c] = w.thetas[c];
                v.pi[c] = w.pi[c];
            }
            for (int c = 0; c < 3; ++c) v.meanPosition[c] = w.meanPosition[c];
            v.usedDistributions = w.usedDistributions;
            vmms[i] = v;
        }
    }
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_guiding_get_wide(rtx_ctx *ctx, rtx_aabb *regions, rtx_vmm_theta_wide *vmms, uint32_t capacity,
                         uint32_t *n) try {
    if (!ctx) return RTX_E_INVALID;
    const uint32_t count = (uint32_t)ctx->guideBoxes.size();
    if (n) *n = count;
    const int rc = guideSync(ctx);
    if (rc != RTX_OK) return rc;
    const uint32_t k = std::min(capacity, count);
    if (regions) std::copy(ctx->guideBoxes.begin(), ctx->guideBoxes.begin() + k, regions);
    if (vmms) std::copy(ctx->guideVmms.begin(), ctx->guideVmms.begin() + k, vmms);
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

// PathGuiding's constructor (src/PathGuiding.cpp:11-31, createRegions 82-106):
// the scene box grown by Aabb::addEpsilon, split splitCount times along the
// largest axis of every box; each region starts with lightpmm's initial
// mixture (8 lobes, kappa 5, spherical-Fibonacci directions).
int rtx_guiding_init(rtx_ctx *ctx, const rtx_aabb *sceneBox, uint32_t splitCount, int parallax) try {
    if (!ctx) return RTX_E_INVALID;
    if (!sceneBox) return fail(ctx, RTX_E_INVALID, "NULL scene box");
    if (splitCount > 16) return fail(ctx, RTX_E_INVALID, "splitCount must be at most 16");
    rtx_aabb e;
    for (int c = 0; c < 3; ++c) {
        if (!(sceneBox->min[c] <= sceneBox->max[c])) return fail(ctx, RTX_E_INVALID, "empty scene box");
        float ext = sceneBox->max[c] - sceneBox->min[c], ctr = sceneBox->min[c] + 0.5f * ext;
        e.min[c] = ctr - 0.50001f * ext;
        e.max[c] = ctr + 0.50001f * ext;
    }
    std::vector<rtx_aabb> boxes{e};
    for (uint32_t it = 0; it < splitCount; ++it) {
        std::vector<rtx_aabb> next;
        next.reserve(boxes.size() * 2);
        for (const rtx_aabb &x : boxes) {
            float size[3] = {x.max[0] - x.min[0], x.max[1] - x.min[1], x.max[2] - x.min[2]};
            int axis = size[0] > size[1] ? (size[0] > size[2] ? 0 : 2) : (size[1] > size[2] ? 1 : 2);
            rtx_aabb l = x, r = x;
            l.max[axis] -= 0.5f * size[axis];
            r.min[axis] += 0.5f * size[axis];
            next.push_back(l);
            next.push_back(r);
        }
        boxes.swap(next);
    }
    const uint32_t n = (uint32_t)boxes.size();
    std::vector<rtx_vmm_theta> placeholder(n);
    for (auto &v : placeholder) v.usedDistributions = 1;
    int rc = rtx_guiding_set(ctx, boxes.data(), placeholder.data(), n);
    if (rc != RTX_OK) return rc;
    GuideFitBuffers &g = ctx->fit;
    g = GuideFitBuffers{};
    g.numRegions = n;
    g.boxes = ctx->guide.boxes;
    hipError_t err = devAlloc(&g.regions, n, ctx->fitAllocs);
    if (err == hipSuccess) err = devAlloc(&g.regionStart, n + 1, ctx->fitAllocs);
    if (err == hipSuccess) err = devAlloc(&g.counters, 4, ctx->fitAllocs);
    if (err != hipSuccess) {
        freeAll(ctx->fitAllocs);
        return fail(ctx, RTX_E_DEVICE, std::string("guiding state: ") + hipGetErrorString(err));
    }
    ctx->fitParallax = parallax != 0;
    ctx->fitFirst = true;
    ctx->fitValid = true;
    launchGuideInit(ctx->stream, g, 8, 5.0f);
    launchGuideToVmm(ctx->stream, g, ctx->fitParallax, const_cast<rtx_vmm_theta_wide *>(ctx->guide.vmms));
    CHECK_HIP(ctx, hipGetLastError());
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->guideVmmsStale = true;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

static int guideFitWork(rtx_ctx *ctx, size_t slots) {
    GuideFitBuffers &g = ctx->fit;
    if (g.slotCap >= slots) return RTX_OK;
    freeAll(ctx->fitWorkAllocs);
    auto &L = ctx->fitWorkAllocs;
    const uint32_t nblk = (uint32_t)((slots + 1023) / 1024);
    CHECK_HIP(ctx, devAlloc(&g.keysA, slots, L));
    CHECK_HIP(ctx, devAlloc(&g.keysB, slots, L));
    CHECK_HIP(ctx, devAlloc(&g.valsA, slots, L));
    CHECK_HIP(ctx, devAlloc(&g.valsB, slots, L));
    CHECK_HIP(ctx, devAlloc(&g.samples, 2 * slots, L));
    CHECK_HIP(ctx, devAlloc(&g.blockSums, nblk, L));
    CHECK_HIP(ctx, devAlloc(&g.hist, (size_t)nblk * 64, L));
    g.slotCap = (uint32_t)slots;
    g.blockCap = nblk;
    return RTX_OK;
}

static EmProps guideProps() { // src/PathGuiding.cpp:33-40 over VMMFactory.h:45-61
    EmProps P;
    P.minItr = 1;
    P.maxItr = 100;
    P.relLogLikelihoodThreshold = 0.005f;
    P.maxKappa = 50000.0f;
    P.vPrior = 0.01f;
    P.rPrior = 0.0f;
    P.rPriorWeight = 1.0f;
    return P;
}

// Region splitting after an update (PathGuiding::update, src/PathGuiding.cpp:
// 295-301; splitRegion 328-348): a region whose mixture saw more than
// samplesForRegionSplit samples is halved along the largest axis of its box
// (Aabb::splitAabb, src/Shapes.h:30-46); its sample counts are decayed by
// 0.25 and the whole region state is copied to a new last region.  Decided
// on the host from the regions' states, which are small; the region
// hierarchy and mixtures are then rebuilt.
static int guideSplitRegions(rtx_ctx *ctx) {
    GuideFitBuffers &g = ctx->fit;
    const uint32_t n = g.numRegions;
    std::vector<GuideRegion> regs(n);
    CHECK_HIP(ctx, hipMemcpyAsync(regs.data(), g.regions, n * sizeof(GuideRegion), hipMemcpyDeviceToHost,
                                  ctx->stream));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<rtx_aabb> boxes = ctx->guideBoxes;
    bool any = false;
    const size_t cap = ctx->guideOpts.maxRegions ? std::min<size_t>(ctx->guideOpts.maxRegions, 1u << 20) : 1u << 20;
    for (uint32_t r = 0; r < n; ++r) {
        if (!(regs[r].numSamples > (float)ctx->guideOpts.samplesForRegionSplit)) continue;
        if (boxes.size() >= cap) break; // the cap: no further region is split
        const rtx_aabb b = boxes[r];
        const float size[3] = {b.max[0] - b.min[0], b.max[1] - b.min[1], b.max[2] - b.min[2]};
        const int axis = size[0] > size[1] ? (size[0] > size[2] ? 0 : 2) : (size[1] > size[2] ? 1 : 2);
        rtx_aabb left = b, right = b;
        left.max[axis] -= 0.5f * size[axis];
        right.min[axis] += 0.5f * size[axis];
        boxes[r] = left;
        boxes.push_back(right);
        regs[r].numSamples *= 0.25f;
        regs[r].sampleWeight *= 0.25f;
        const GuideRegion copy = regs[r];
        regs.push_back(copy);
        any = true;
    }
    if (!any) return RTX_OK;
    const uint32_t m = (uint32_t)regs.size();
    std::vector<rtx_vmm_theta> placeholder(m);
    for (auto &v : placeholder) v.usedDistributions = 1;
    int rc = guideUpload(ctx, boxes.data(), placeholder.data(), m); // boxes, hierarchy
    if (rc != RTX_OK) return rc;
    CHECK_HIP(ctx, reallocIn(ctx->fitAllocs, g.regions, m));
    CHECK_HIP(ctx, reallocIn(ctx->fitAllocs, g.regionStart, (size_t)m + 1));
    CHECK_HIP(ctx, hipMemcpy(g.regions, regs.data(), m * sizeof(GuideRegion), hipMemcpyHostToDevice));
    g.numRegions = m;
    g.boxes = ctx->guide.boxes;
    launchGuideToVmm(ctx->stream, g, ctx->fitParallax, const_cast<rtx_vmm_theta_wide *>(ctx->guide.vmms));
    CHECK_HIP(ctx, hipGetLastError());
    return RTX_OK;
}

// PathGuiding::update (src/PathGuiding.cpp:276-312) on the samples the last
// recording frame left (stream-ordered, no host wait unless regions split).
static int guideUpdateAsync(rtx_ctx *ctx) {
    if (ctx->dirCount > (size_t)UINT32_MAX) return fail(ctx, RTX_E_INVALID, "too many guiding samples");
    int rc = guideFitWork(ctx, ctx->dirCount);
    if (rc != RTX_OK) return rc;
    launchGuideUpdate(ctx->stream, ctx->fit, ctx->dirData, (uint32_t)ctx->dirCount, guideProps(), ctx->guideOpts,
                      ctx->fitParallax, ctx->fitFirst, const_cast<rtx_vmm_theta_wide *>(ctx->guide.vmms));
    CHECK_HIP(ctx, hipGetLastError());
    ctx->fitFirst = false;
    if (ctx->guideOpts.splitRegions) {
        rc = guideSplitRegions(ctx);
        if (rc != RTX_OK) return rc;
    }
    ctx->guideVmmsStale = true;
    return RTX_OK;
}

int rtx_guiding_update(rtx_ctx *ctx) try {
    if (!ctx) return RTX_E_INVALID;
    if (!ctx->fitValid) return fail(ctx, RTX_E_STATE, "no guiding state to update (rtx_guiding_init)");
    if (ctx->dirCount == 0) return fail(ctx, RTX_E_STATE, "no render with updateGuiding yet");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    int rc = guideUpdateAsync(ctx);
    if (rc != RTX_OK) return rc;
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_guiding_set_options(rtx_ctx *ctx, const rtx_guiding_options *options) try {
    if (!ctx) return RTX_E_INVALID;
    if (!options) return fail(ctx, RTX_E_INVALID, "NULL options");
    if (!(options->splitMinDivergence >= 0.0f) || !(options->mergeMaxDivergence >= 0.0f))
        return fail(ctx, RTX_E_INVALID, "divergence thresholds must be >= 0");
    if (options->maxComponents < 1 || options->maxComponents > RTX_MAX_LOBES)
        return fail(ctx, RTX_E_INVALID, "maxComponents must be in [1, RTX_MAX_LOBES]");
    ctx->guideOpts = *options;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_guiding_get_options(rtx_ctx *ctx, rtx_guiding_options *options) try {
    if (!ctx) return RTX_E_INVALID;
    if (!options) return fail(ctx, RTX_E_INVALID, "NULL options");
    *options = ctx->guideOpts;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_guiding_read_samples(rtx_ctx *ctx, rtx_directional_data *out, size_t count) try {
    if (!ctx) return RTX_E_INVALID;
    if (ctx->dirCount == 0) return fail(ctx, RTX_E_STATE, "no render with updateGuiding yet");
    if (count > ctx->dirCount) return fail(ctx, RTX_E_INVALID, "more samples asked for than recorded");
    if (count == 0) return RTX_OK;
    if (!out) return fail(ctx, RTX_E_INVALID, "NULL argument");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CHECK_HIP(ctx, hipMemcpy(out, ctx->dirData, count * sizeof(rtx_directional_data), hipMemcpyDeviceToHost));
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

// ------------------------------------------------------- irradiance cache
static void icFree(rtx_ctx *ctx) {
    freeAll(ctx->icAllocs);
    if (ctx->ic.fullHost) (void)hipHostFree(ctx->ic.fullHost);
    ctx->ic = IcBuffers{};
    ctx->icValid = false;
}

int rtx_ic_create(rtx_ctx *ctx, uint32_t maxEntries) try {
    if (!ctx) return RTX_E_INVALID;
    if (maxEntries == 0 || maxEntries > (1u << 24))
        return fail(ctx, RTX_E_INVALID, "maxEntries must be in [1, 2^24]");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    icFree(ctx);
    IcBuffers &ic = ctx->ic;
    auto &L = ctx->icAllocs;
    ic.maxEntries = maxEntries;
    ic.nb = 1024;
    // records file up to 27 items (about 12 on surfaces): 16 buckets per record, up to 2^26
    while (ic.nb < 16u * maxEntries && ic.nb < (1u << 26)) ic.nb <<= 1;
    ic.itemCap = 27u * maxEntries;
    ic.bucketBits = 0;
    while ((1u << ic.bucketBits) < ic.nb) ic.bucketBits++;
    ic.jobCap = 2u * maxEntries;
    ic.blockCap = kMaxBlockSums;
    hipError_t e = devAlloc(&ic.hdr, 1, L);
    if (e == hipSuccess) e = devAlloc(&ic.entries, maxEntries, L);
    if (e == hipSuccess) e = devAlloc(&ic.spheres, maxEntries, L);
    if (e == hipSuccess) e = devAlloc(&ic.bucketCount, ic.nb, L);
    if (e == hipSuccess) e = devAlloc(&ic.bucketStart, ic.nb + 1, L);
    if (e == hipSuccess) e = devAlloc(&ic.cursor, ic.nb, L);
    if (e == hipSuccess) e = devAlloc(&ic.meta, 4, L);
    if (e == hipSuccess) e = devAlloc(&ic.items, ic.itemCap, L);
    if (e == hipSuccess) e = devAlloc(&ic.slotKeys, ic.itemCap, L);
    if (e == hipSuccess) e = devAlloc(&ic.slotKeysB, ic.itemCap, L);
    if (e == hipSuccess) e = devAlloc(&ic.slotIdx, ic.itemCap, L);
    if (e == hipSuccess) e = devAlloc(&ic.slotIdxB, ic.itemCap, L);
    if (e == hipSuccess) e = devAlloc(&ic.sortHist, (size_t)((ic.itemCap + kSortBlock - 1) / kSortBlock) * kDigits, L);
    if (e == hipSuccess) e = devAlloc(&ic.frame, 8, L);
    if (e == hipSuccess) e = devAlloc(&ic.cand, 2 * (size_t)maxEntries, L);
    if (e == hipSuccess) e = devAlloc(&ic.jobs, 2 * (size_t)ic.jobCap, L);
    if (e == hipSuccess) e = devAlloc(&ic.jobRes, 10 * (size_t)ic.jobCap, L);
    if (e == hipSuccess) e = devAlloc(&ic.blockSums, ic.blockCap, L);
    if (e == hipSuccess) e = hipHostMalloc((void **)&ic.fullHost, sizeof(uint32_t), hipHostMallocDefault);
    if (e == hipSuccess) *ic.fullHost = 0u;
    if (e == hipSuccess && !ctx->fcIC) e = devMalloc(&ctx->fcIC, sizeof(FrameConst));
    if (e != hipSuccess) {
        icFree(ctx);
        return fail(ctx, RTX_E_DEVICE, std::string("irradiance cache allocation: ") + hipGetErrorString(e));
    }
    return rtx_ic_reset(ctx);
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_ic_reset(rtx_ctx *ctx) try {
    if (!ctx) return RTX_E_INVALID;
    if (ctx->ic.maxEntries == 0) return fail(ctx, RTX_E_STATE, "no irradiance cache (rtx_ic_create)");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    launchIcInit(ctx->stream, ctx->ic);
    CHECK_HIP(ctx, hipGetLastError());
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *ctx->ic.fullHost = 0u;
    ctx->icValid = true;
    ctx->keptJobs = ctx->keptUpdates = 0;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_estimate_write(rtx_ctx *ctx, const float *rgba, uint32_t width, uint32_t height) try {
    if (!ctx) return RTX_E_INVALID;
    if (!rgba || width == 0 || height == 0) return fail(ctx, RTX_E_INVALID, "NULL or empty estimate");
    const size_t npix = (size_t)width * height;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!ctx->estimate || ctx->estW != width || ctx->estH != height) {
        if (ctx->estimate) (void)hipFree(ctx->estimate);
        ctx->estimate = nullptr;
        ctx->estW = ctx->estH = 0;
        CHECK_HIP(ctx, devMalloc(&ctx->estimate, npix * sizeof(float4)));
        ctx->estW = width;
        ctx->estH = height;
    }
    CHECK_HIP(ctx, hipMemcpy(ctx->estimate, rgba, npix * sizeof(float4), hipMemcpyHostToDevice));
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_estimate_read(rtx_ctx *ctx, float *rgba, uint32_t width, uint32_t height) try {
    if (!ctx) return RTX_E_INVALID;
    if (!rgba) return fail(ctx, RTX_E_INVALID, "NULL argument");
    if (!ctx->estimate || ctx->estW != width || ctx->estH != height)
        return fail(ctx, RTX_E_STATE, "no estimate of this size");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CHECK_HIP(ctx, hipMemcpy(rgba, ctx->estimate, (size_t)width * height * sizeof(float4), hipMemcpyDeviceToHost));
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_ic_read(rtx_ctx *ctx, rtx_cache_header *hdr, rtx_cache_data *entries, rtx_sphere *spheres,
                uint32_t capacity) try {
    if (!ctx) return RTX_E_INVALID;
    if (!ctx->icValid) return fail(ctx, RTX_E_STATE, "no irradiance cache (rtx_ic_create)");
    const uint32_t n = std::min(capacity, ctx->ic.maxEntries);
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (hdr) CHECK_HIP(ctx, hipMemcpy(hdr, ctx->ic.hdr, sizeof(rtx_cache_header), hipMemcpyDeviceToHost));
    if (entries && n) CHECK_HIP(ctx, hipMemcpy(entries, ctx->ic.entries, n * sizeof(rtx_cache_data),
                                               hipMemcpyDeviceToHost));
    if (spheres && n) CHECK_HIP(ctx, hipMemcpy(spheres, ctx->ic.spheres, n * sizeof(rtx_sphere),
                                               hipMemcpyDeviceToHost));
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_ic_write(rtx_ctx *ctx, const rtx_cache_header *hdr, const rtx_cache_data *entries,
                 const rtx_sphere *spheres) try {
    if (!ctx) return RTX_E_INVALID;
    if (!ctx->icValid) return fail(ctx, RTX_E_STATE, "no irradiance cache (rtx_ic_create)");
    if (!hdr || !entries || !spheres) return fail(ctx, RTX_E_INVALID, "NULL argument");
    if (hdr->maxCaches != ctx->ic.maxEntries || hdr->nextCacheSlot > hdr->maxCaches)
        return fail(ctx, RTX_E_INVALID, "header does not match the cache (maxCaches, nextCacheSlot)");
    const size_t n = ctx->ic.maxEntries;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CHECK_HIP(ctx, hipMemcpy(ctx->ic.hdr, hdr, sizeof(rtx_cache_header), hipMemcpyHostToDevice));
    CHECK_HIP(ctx, hipMemcpy(ctx->ic.entries, entries, n * sizeof(rtx_cache_data), hipMemcpyHostToDevice));
    CHECK_HIP(ctx, hipMemcpy(ctx->ic.spheres, spheres, n * sizeof(rtx_sphere), hipMemcpyHostToDevice));
    *ctx->ic.fullHost = 0u; // the next commit tells again
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_ic_read_jobs(rtx_ctx *ctx, rtx_ic_job *jobs, float *sub, uint32_t *subSig, uint32_t capacity,
                     uint32_t *numJobs) try {
    if (!ctx) return RTX_E_INVALID;
    if (!ctx->icValid) return fail(ctx, RTX_E_STATE, "no irradiance cache (rtx_ic_create)");
    if (numJobs) *numJobs = ctx->keptJobs;
    const uint32_t n = std::min(capacity, ctx->keptJobs);
    if (n == 0) return RTX_OK;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (jobs) {
        std::vector<float4> j(2 * (size_t)n);
        CHECK_HIP(ctx, hipMemcpy(j.data(), ctx->ic.jobs, j.size() * sizeof(float4), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) {
            float4 o = j[2 * i], nn = j[2 * i + 1];
            int32_t type;
            uint32_t target;
            std::memcpy(&type, &o.w, 4);
            std::memcpy(&target, &nn.w, 4);
            jobs[i].type = type;
            jobs[i].target = target;
            jobs[i].origin[0] = o.x;
            jobs[i].origin[1] = o.y;
            jobs[i].origin[2] = o.z;
            jobs[i].normal[0] = nn.x;
            jobs[i].normal[1] = nn.y;
            jobs[i].normal[2] = nn.z;
        }
    }
    const size_t ns = (size_t)n * kIcStrata;
    if (sub) CHECK_HIP(ctx, hipMemcpy(sub, ctx->keepSub, ns * sizeof(float4), hipMemcpyDeviceToHost));
    if (subSig) CHECK_HIP(ctx, hipMemcpy(subSig, ctx->keepSubSig, ns * sizeof(uint2), hipMemcpyDeviceToHost));
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_read_image(rtx_ctx *ctx, float *rgba, size_t floatCount) try {
    if (!ctx) return RTX_E_INVALID;
    if (!rgba) return fail(ctx, RTX_E_INVALID, "rgba is NULL");
    if (!ctx->accum || ctx->accSamples == 0) return fail(ctx, RTX_E_STATE, "nothing rendered yet");
    size_t npix = (size_t)ctx->accW * ctx->accH;
    if (floatCount < npix * 4) return fail(ctx, RTX_E_INVALID, "output buffer too small");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    launchResolveImage(ctx->stream, ctx->accum, (uint32_t)npix, 1.0f / (float)ctx->accSamples, ctx->resolved);
    CHECK_HIP(ctx, hipGetLastError());
    CHECK_HIP(ctx, hipMemcpyAsync(rgba, ctx->resolved, npix * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_accumulated_samples(const rtx_ctx *ctx, uint32_t *samples) try {
    if (!ctx || !samples) return RTX_E_INVALID;
    *samples = ctx->accSamples;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(nullptr);
}

int rtx_read_paths(rtx_ctx *ctx, float *rgb, uint32_t *sig, size_t n) try {
    if (!ctx) return RTX_E_INVALID;
    if (!ctx->keepRgb || ctx->keptPaths == 0) return fail(ctx, RTX_E_STATE, "no render with RTX_FRAME_KEEP_PATHS");
    if (n > ctx->keptPaths) return fail(ctx, RTX_E_INVALID, "more paths requested than rendered");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    if (rgb) {
        std::vector<float4> tmp(n);
        CHECK_HIP(ctx, hipMemcpy(tmp.data(), ctx->keepRgb, n * sizeof(float4), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            rgb[3 * i + 0] = tmp[i].x;
            rgb[3 * i + 1] = tmp[i].y;
            rgb[3 * i + 2] = tmp[i].z;
        }
    }
    if (sig) CHECK_HIP(ctx, hipMemcpy(sig, ctx->keepSig, n * sizeof(uint2), hipMemcpyDeviceToHost));
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

// ------------------------------------------------ learned radiance cache
// Deterministic parameter initialisation stream (splitmix64).
static uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static float uniform01(uint64_t &x) { return (float)(splitmix64(x) >> 40) * (1.0f / 16777216.0f); }

static void nrcFree(rtx_ctx *ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    freeAll(ctx->nrcAllocs);
    freeAll(ctx->nrcPathAllocs);
    freeAll(ctx->nrcSlotAllocs);
    ctx->nrc = NrcBuffers{};
    ctx->nrcLoss = nullptr;
    ctx->nrcValid = false;
    ctx->nrcPathCap = ctx->nrcSlotCap = 0;
    ctx->pb.nrcQuery = ctx->pb.nrcQueryAux = ctx->pb.nrcPath = nullptr;
    ctx->pb.nrcQueryCount = nullptr;
    ctx->pb.nrcTrainIn = ctx->pb.nrcTrainTarget = nullptr;
}

// Adam restarts (moments and step count), gradients cleared.
static int nrcResetOptimizer(rtx_ctx *ctx) {
    const size_t total = ctx->nrc.gridParams + kNrcMlpParams;
    CHECK_HIP(ctx, hipMemsetAsync(ctx->nrc.grads, 0, total * sizeof(float), ctx->stream));
    CHECK_HIP(ctx, hipMemsetAsync(ctx->nrc.adamM, 0, total * sizeof(float), ctx->stream));
    CHECK_HIP(ctx, hipMemsetAsync(ctx->nrc.adamV, 0, total * sizeof(float), ctx->stream));
    ctx->nrc.step = 0;
    return RTX_OK;
}

// One Adam step on rows [0, n) of device records `in` / `target`; the loss
// adds to ctx->nrcLoss.
static int nrcStep(rtx_ctx *ctx, const float4 *in, const float4 *target, uint32_t n) {
    launchNrcGradients(ctx->stream, ctx->nrc, in, target, n, 1.0f / (3.0f * (float)n), ctx->nrcLoss, ctx->cuCount);
    ctx->nrc.step++;
    launchNrcAdam(ctx->stream, ctx->nrc);
    CHECK_HIP(ctx, hipGetLastError());
    return RTX_OK;
}

// Host rows (rtx_nrc_input, 3 floats of target each) into temporary device
// buffers (4 float4 per input row, one float4 per target).
static int nrcUpload(rtx_ctx *ctx, const rtx_nrc_input *in, const float *target, size_t n, std::vector<void *> &tmp,
                     float4 **inD, float4 **tD) {
    CHECK_HIP(ctx, devAlloc(inD, 4 * n, tmp));
    CHECK_HIP(ctx, hipMemcpyAsync(*inD, in, n * sizeof(rtx_nrc_input), hipMemcpyHostToDevice, ctx->stream));
    if (target && tD) {
        std::vector<float4> t(n);
        for (size_t i = 0; i < n; ++i) t[i] = make_float4(target[3 * i], target[3 * i + 1], target[3 * i + 2], 1.0f);
        CHECK_HIP(ctx, devAlloc(tD, n, tmp));
        CHECK_HIP(ctx, hipMemcpy(*tD, t.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    }
    return RTX_OK;
}

void rtx_nrc_config_defaults(rtx_nrc_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->levels = 16;
    c->log2TableSize = 19;
    c->baseResolution = 16.0f;
    c->maxResolution = 2048.0f;
    c->learningRate = 1e-2f;
    c->beta1 = 0.9f;
    c->beta2 = 0.99f;
    c->epsilon = 1e-15f;
    c->trainBatch = 16384;
    c->trainSteps = 4;
    c->terminateBounces = 2;
    c->trainingPaths = 65536;
    c->trainSuffixBounces = 1;
    c->emaDecay = 0.99f;
    for (int a = 0; a < 3; ++a) {
        c->sceneMin[a] = -1.0f;
        c->sceneMax[a] = 1.0f;
    }
    c->seed = 1;
}

int rtx_nrc_create(rtx_ctx *ctx, const rtx_nrc_config *c) try {
    if (!ctx) return RTX_E_INVALID;
    if (!c) return fail(ctx, RTX_E_INVALID, "NULL config");
    if (c->levels < 1 || c->levels > (uint32_t)kNrcMaxLevels || c->log2TableSize < 8 || c->log2TableSize > 24)
        return fail(ctx, RTX_E_INVALID, "levels must be in [1, 16] and log2TableSize in [8, 24]");
    if (!(c->baseResolution >= 2.0f) || !(c->maxResolution >= c->baseResolution) || !(c->maxResolution <= 1e6f))
        return fail(ctx, RTX_E_INVALID, "need 2 <= baseResolution <= maxResolution <= 1e6");
    if (c->trainBatch < 256 || c->trainBatch > 65536 || c->trainSteps < 1 || c->trainSteps > 64 ||
        c->trainingPaths < 1 || c->terminateBounces > 64 || c->trainSuffixBounces > 64)
        return fail(ctx, RTX_E_INVALID, "trainBatch in [256, 65536], trainSteps in [1, 64], trainingPaths >= 1, "
                                        "terminateBounces and trainSuffixBounces <= 64");
    if (!(c->learningRate > 0.0f) || !(c->beta1 >= 0.0f && c->beta1 < 1.0f) || !(c->beta2 >= 0.0f && c->beta2 < 1.0f) ||
        !(c->epsilon >= 0.0f))
        return fail(ctx, RTX_E_INVALID, "Adam: learningRate > 0, beta1 and beta2 in [0, 1), epsilon >= 0");
    if (!(c->emaDecay >= 0.0f && c->emaDecay < 1.0f)) return fail(ctx, RTX_E_INVALID, "emaDecay must be in [0, 1)");
    for (int a = 0; a < 3; ++a)
        if (!(c->sceneMax[a] > c->sceneMin[a])) return fail(ctx, RTX_E_INVALID, "sceneMax must exceed sceneMin");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    nrcFree(ctx);
    NrcBuffers nb{};
    const uint32_t T = 1u << c->log2TableSize;
    nb.grid.levels = c->levels;
    nb.grid.tableMask = T - 1u;
    const double growth =
        c->levels > 1 ? std::exp((std::log((double)c->maxResolution) - std::log((double)c->baseResolution)) /
                                 (double)(c->levels - 1))
                      : 1.0;
    for (uint32_t l = 0; l < (uint32_t)kNrcMaxLevels; ++l) {
        const double r = std::floor((double)c->baseResolution * std::pow(growth, (double)l));
        nb.grid.res[l] = (float)r;
        nb.grid.dense[l] = (r + 1.0) * (r + 1.0) * (r + 1.0) <= (double)T ? 1u : 0u;
    }
    for (int a = 0; a < 3; ++a) {
        nb.grid.lo[a] = c->sceneMin[a];
        nb.grid.inv[a] = 1.0f / (c->sceneMax[a] - c->sceneMin[a]);
    }
    nb.gridParams = (size_t)c->levels * T * 2;
    const size_t total = nb.gridParams + kNrcMlpParams;
    nb.rowsCap = (c->trainBatch + 255u) & ~255u;
    auto &L = ctx->nrcAllocs;
    CHECK_HIP(ctx, devAlloc(&nb.params, total, L));
    CHECK_HIP(ctx, devAlloc(&nb.grads, total, L));
    CHECK_HIP(ctx, devAlloc(&nb.adamM, total, L));
    CHECK_HIP(ctx, devAlloc(&nb.adamV, total, L));
    CHECK_HIP(ctx, devAlloc(&nb.gridBf16, (size_t)c->levels * T, L));
    CHECK_HIP(ctx, devAlloc(&nb.wFwd, (size_t)kNrcLayers * kNrcWidth * kNrcWidth, L));
    CHECK_HIP(ctx, devAlloc(&nb.wBwd, (size_t)kNrcLayers * kNrcWidth * kNrcWidth, L));
    nb.emaDecay = c->emaDecay;
    if (nb.emaDecay > 0.0f) {
        CHECK_HIP(ctx, devAlloc(&nb.ema, total, L));
        CHECK_HIP(ctx, devAlloc(&nb.gridBf16Ema, (size_t)c->levels * T, L));
        CHECK_HIP(ctx, devAlloc(&nb.wFwdEma, (size_t)kNrcLayers * kNrcWidth * kNrcWidth, L));
    } else {
        nb.ema = nb.params;
        nb.gridBf16Ema = nb.gridBf16;
        nb.wFwdEma = nb.wFwd;
    }
    CHECK_HIP(ctx, devAlloc(&nb.acts, (size_t)kNrcLayers * nb.rowsCap * 64, L));
    CHECK_HIP(ctx, devAlloc(&nb.dz, (size_t)kNrcLayers * nb.rowsCap * 64, L));
    CHECK_HIP(ctx, devAlloc(&nb.pred, (size_t)nb.rowsCap * 4, L));
    CHECK_HIP(ctx, devAlloc(&nb.dFeat, (size_t)nb.rowsCap * 32, L));
    CHECK_HIP(ctx, devAlloc(&nb.partials, (size_t)(nb.rowsCap / 256) * kNrcLayers * kNrcWidth * kNrcWidth, L));
    CHECK_HIP(ctx, devAlloc(&ctx->nrcLoss, 1, L));
    hipStream_t s = ctx->stream;
    CHECK_HIP(ctx, hipMemsetAsync(nb.wFwd, 0, (size_t)kNrcLayers * kNrcWidth * kNrcWidth * 2, s));
    CHECK_HIP(ctx, hipMemsetAsync(nb.wBwd, 0, (size_t)kNrcLayers * kNrcWidth * kNrcWidth * 2, s));
    if (nb.wFwdEma != nb.wFwd)
        CHECK_HIP(ctx, hipMemsetAsync(nb.wFwdEma, 0, (size_t)kNrcLayers * kNrcWidth * kNrcWidth * 2, s));
    CHECK_HIP(ctx, hipMemsetAsync(nb.acts, 0, (size_t)kNrcLayers * nb.rowsCap * 64 * 2, s));
    CHECK_HIP(ctx, hipMemsetAsync(nb.dz, 0, (size_t)kNrcLayers * nb.rowsCap * 64 * 2, s));
    CHECK_HIP(ctx, hipMemsetAsync(ctx->nrcLoss, 0, sizeof(float), s));
    // the table U(-1e-4, 1e-4) (instant-ngp), the MLP Glorot-uniform, the
    // output layer's padding rows zero
    std::vector<float> h(total, 0.0f);
    uint64_t st = c->seed;
    for (size_t i = 0; i < nb.gridParams; ++i) h[i] = (uniform01(st) * 2.0f - 1.0f) * 1e-4f;
    for (int layer = 0; layer < kNrcLayers; ++layer) {
        const int outs = layer < kNrcHidden ? kNrcWidth : 3;
        const float bound = std::sqrt(6.0f / (float)(kNrcWidth + outs));
        float *w = h.data() + nb.gridParams + (size_t)layer * kNrcWidth * kNrcWidth;
        for (int o = 0; o < outs; ++o)
            for (int k = 0; k < kNrcWidth; ++k) w[o * kNrcWidth + k] = (uniform01(st) * 2.0f - 1.0f) * bound;
    }
    CHECK_HIP(ctx, hipMemcpyAsync(nb.params, h.data(), total * sizeof(float), hipMemcpyHostToDevice, s));
    nb.lr = c->learningRate;
    nb.beta1 = c->beta1;
    nb.beta2 = c->beta2;
    nb.eps = c->epsilon;
    ctx->nrc = nb;
    int rc = nrcResetOptimizer(ctx);
    if (rc != RTX_OK) return rc;
    CHECK_HIP(ctx, launchNrcEmaReset(s, ctx->nrc));
    CHECK_HIP(ctx, hipGetLastError());
    CHECK_HIP(ctx, hipStreamSynchronize(s));
    ctx->nrcCfg = *c;
    ctx->nrcValid = true;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_nrc_destroy(rtx_ctx *ctx) try {
    if (!ctx) return RTX_E_INVALID;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    nrcFree(ctx);
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_nrc_get_info(rtx_ctx *ctx, rtx_nrc_info *info) try {
    if (!ctx || !info) return RTX_E_INVALID;
    if (!ctx->nrcValid) return fail(ctx, RTX_E_STATE, "no learned radiance cache (rtx_nrc_create)");
    std::memset(info, 0, sizeof(*info));
    const NrcGrid &g = ctx->nrc.grid;
    info->levels = g.levels;
    info->tableSize = g.tableMask + 1u;
    for (int l = 0; l < kNrcMaxLevels; ++l) {
        info->resolution[l] = g.res[l];
        info->dense[l] = g.dense[l];
    }
    for (int a = 0; a < 3; ++a) {
        info->lo[a] = g.lo[a];
        info->inv[a] = g.inv[a];
    }
    info->gridParams = ctx->nrc.gridParams;
    info->mlpParams = (uint32_t)kNrcMlpParams;
    info->steps = ctx->nrc.step;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    float loss = 0.0f;
    CHECK_HIP(ctx, hipMemcpyAsync(&loss, ctx->nrcLoss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    info->lastLoss = ctx->nrcLossSteps ? loss / (float)ctx->nrcLossSteps : 0.0f;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_nrc_get_params(rtx_ctx *ctx, float *grid, size_t gridCount, float *mlp, size_t mlpCount) try {
    if (!ctx) return RTX_E_INVALID;
    if (!ctx->nrcValid) return fail(ctx, RTX_E_STATE, "no learned radiance cache (rtx_nrc_create)");
    if ((grid && gridCount != ctx->nrc.gridParams) || (mlp && mlpCount != (size_t)kNrcMlpParams))
        return fail(ctx, RTX_E_INVALID, "parameter counts differ from rtx_nrc_info");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (grid) CHECK_HIP(ctx, hipMemcpy(grid, ctx->nrc.params, gridCount * sizeof(float), hipMemcpyDeviceToHost));
    if (mlp)
        CHECK_HIP(ctx, hipMemcpy(mlp, ctx->nrc.params + ctx->nrc.gridParams, mlpCount * sizeof(float),
                                 hipMemcpyDeviceToHost));
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_nrc_set_params(rtx_ctx *ctx, const float *grid, size_t gridCount, const float *mlp, size_t mlpCount) try {
    if (!ctx) return RTX_E_INVALID;
    if (!ctx->nrcValid) return fail(ctx, RTX_E_STATE, "no learned radiance cache (rtx_nrc_create)");
    if ((grid && gridCount != ctx->nrc.gridParams) || (mlp && mlpCount != (size_t)kNrcMlpParams))
        return fail(ctx, RTX_E_INVALID, "parameter counts differ from rtx_nrc_info");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (grid) CHECK_HIP(ctx, hipMemcpy(ctx->nrc.params, grid, gridCount * sizeof(float), hipMemcpyHostToDevice));
    if (mlp) {
        std::vector<float> w(mlp, mlp + mlpCount);
        // the output layer's padding rows stay zero
        for (size_t i = (size_t)kNrcHidden * kNrcWidth * kNrcWidth + 3 * kNrcWidth; i < w.size(); ++i) w[i] = 0.0f;
        CHECK_HIP(ctx, hipMemcpy(ctx->nrc.params + ctx->nrc.gridParams, w.data(), mlpCount * sizeof(float),
                                 hipMemcpyHostToDevice));
    }
    int rc = nrcResetOptimizer(ctx);
    if (rc != RTX_OK) return rc;
    CHECK_HIP(ctx, launchNrcEmaReset(ctx->stream, ctx->nrc));
    CHECK_HIP(ctx, hipGetLastError());
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_nrc_infer(rtx_ctx *ctx, const rtx_nrc_input *in, size_t n, float *rgb) try {
    if (!ctx) return RTX_E_INVALID;
    if (!ctx->nrcValid) return fail(ctx, RTX_E_STATE, "no learned radiance cache (rtx_nrc_create)");
    if (n && (!in || !rgb)) return fail(ctx, RTX_E_INVALID, "NULL argument");
    if (n == 0) return RTX_OK;
    if (n > (size_t)INT32_MAX) return fail(ctx, RTX_E_INVALID, "more than 2^31-1 queries");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<void *> tmp;
    float4 *inD = nullptr, *outD = nullptr;
    int rc = nrcUpload(ctx, in, nullptr, n, tmp, &inD, nullptr);
    if (rc == RTX_OK) {
        hipError_t e = devAlloc(&outD, n, tmp);
        if (e == hipSuccess) {
            launchNrcInfer(ctx->stream, ctx->nrc, inD, (uint32_t)n, nullptr, outD, nullptr, nullptr, 2 * ctx->cuCount);
            e = hipGetLastError();
        }
        std::vector<float4> out(n);
        if (e == hipSuccess) e = hipMemcpyAsync(out.data(), outD, n * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(ctx, RTX_E_DEVICE, std::string("rtx_nrc_infer: ") + hipGetErrorString(e));
        else
            for (size_t i = 0; i < n; ++i) {
                rgb[3 * i] = out[i].x;
                rgb[3 * i + 1] = out[i].y;
                rgb[3 * i + 2] = out[i].z;
            }
    }
    (void)hipStreamSynchronize(ctx->stream);
    freeAll(tmp);
    return rc;
} catch (...) {
    return rtx::abiFail(ctx);
}

static int nrcTrainCall(rtx_ctx *ctx, const rtx_nrc_input *in, const float *target, size_t n, float *gridGrad,
                        float *mlpGrad, float *loss, bool step) {
    if (!ctx) return RTX_E_INVALID;
    if (!ctx->nrcValid) return fail(ctx, RTX_E_STATE, "no learned radiance cache (rtx_nrc_create)");
    if (!in || !target || n == 0) return fail(ctx, RTX_E_INVALID, "no rows");
    if (n > ctx->nrcCfg.trainBatch) return fail(ctx, RTX_E_INVALID, "more rows than trainBatch");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<void *> tmp;
    float4 *inD = nullptr, *tD = nullptr;
    int rc = nrcUpload(ctx, in, target, n, tmp, &inD, &tD);
    hipStream_t s = ctx->stream;
    if (rc == RTX_OK) {
        launchNrcZero(s, ctx->nrcLoss, 1);
        ctx->nrcLossSteps = 1;
        if (step) {
            rc = nrcStep(ctx, inD, tD, (uint32_t)n);
        } else {
            launchNrcGradients(s, ctx->nrc, inD, tD, (uint32_t)n, 1.0f / (3.0f * (float)n), ctx->nrcLoss, ctx->cuCount);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess && gridGrad)
                e = hipMemcpyAsync(gridGrad, ctx->nrc.grads, ctx->nrc.gridParams * sizeof(float), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess && mlpGrad)
                e = hipMemcpyAsync(mlpGrad, ctx->nrc.grads + ctx->nrc.gridParams, kNrcMlpParams * sizeof(float),
                                   hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) launchNrcZeroGrads(s, ctx->nrc); // gradients are zero between calls
            if (e == hipSuccess) e = hipGetLastError();
            if (e != hipSuccess) rc = fail(ctx, RTX_E_DEVICE, std::string("rtx_nrc_gradients: ") + hipGetErrorString(e));
        }
        float l = 0.0f;
        if (rc == RTX_OK) {
            hipError_t e = hipMemcpyAsync(&l, ctx->nrcLoss, sizeof(float), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) rc = fail(ctx, RTX_E_DEVICE, std::string("rtx_nrc_train: ") + hipGetErrorString(e));
        }
        if (loss) *loss = l;
    }
    (void)hipStreamSynchronize(s);
    freeAll(tmp);
    return rc;
}

int rtx_nrc_gradients(rtx_ctx *ctx, const rtx_nrc_input *in, const float *target, size_t n, float *gridGrad,
                      float *mlpGrad, float *loss) try {
    return nrcTrainCall(ctx, in, target, n, gridGrad, mlpGrad, loss, false);
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_nrc_train(rtx_ctx *ctx, const rtx_nrc_input *in, const float *target, size_t n, float *loss) try {
    return nrcTrainCall(ctx, in, target, n, nullptr, nullptr, loss, true);
} catch (...) {
    return rtx::abiFail(ctx);
}

// ------------------------------------------------ allocation-free query
// Device memory of rtx_trace_closest_noalloc: one call's rays, hits and the
// traversal's work counter, in a file-scope array that each device holds
// until the process ends.  Host code never passes `scratch` itself to a
// kernel or a copy: its name is a host-side stand-in, and the device address
// comes from hipGetSymbolAddress, once per device (noAllocScratch).
constexpr size_t kNoAllocRayBytes = (size_t)RTX_NOALLOC_MAX_RAYS * sizeof(rtx_ray);
constexpr size_t kNoAllocHitBytes = (size_t)RTX_NOALLOC_MAX_RAYS * sizeof(rtx_hit);
constexpr size_t kNoAllocBytes = kNoAllocRayBytes + kNoAllocHitBytes + 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kNoAllocDevices = 64;
std::mutex gScratchMutex;
void *gScratch[kNoAllocDevices] = {};
hipError_t gScratchErr[kNoAllocDevices] = {};
bool gScratchDone[kNoAllocDevices] = {};
} // namespace

// The device address of `scratch` on `device` (the current device), looked
// up on the first call for that device.
static void *noAllocScratch(int device, hipError_t &err) {
    if (device < 0 || device >= kNoAllocDevices) {
        err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchDone[device]) {
        void *ptr = nullptr;
        gScratchErr[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        gScratch[device] = gScratchErr[device] == hipSuccess ? ptr : nullptr;
        gScratchDone[device] = true;
    }
    err = gScratchErr[device];
    return gScratch[device];
}

static int queryCommon(rtx_ctx *ctx, const rtx_scene *scene, const rtx_ray *rays, size_t n, void *out,
                       size_t outBytes, bool closest) {
    if (!ctx) return RTX_E_INVALID;
    if (!scene || (n && (!rays || !out))) return fail(ctx, RTX_E_INVALID, "NULL argument");
    if (n == 0) return RTX_OK;
    if (n > (size_t)INT32_MAX) return fail(ctx, RTX_E_INVALID, "ray query: more than 2^31-1 rays");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<void *> tmp;
    rtx_ray *rD = nullptr;
    void *oD = nullptr;
    uint32_t *fetch = nullptr;
    hipError_t e = devAlloc(&rD, n, tmp);
    if (e == hipSuccess) e = devAlloc((char **)&oD, n * outBytes, tmp);
    if (e == hipSuccess) e = devAlloc(&fetch, 1, tmp);
    if (e == hipSuccess) e = hipMemsetAsync(fetch, 0, sizeof(uint32_t), ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rD, rays, n * sizeof(rtx_ray), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        launchQueryP(ctx->stream, scene->view, rD, (uint32_t)n, oD, closest, fetch,
                     ctx->cuCount * ctx->tuning.traceBlocksPerCU);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, oD, n * outBytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    freeAll(tmp);
    if (e != hipSuccess) return fail(ctx, RTX_E_DEVICE, std::string("ray query: ") + hipGetErrorString(e));
    return RTX_OK;
}

int rtx_trace_closest(rtx_ctx *ctx, const rtx_scene *scene, const rtx_ray *rays, size_t n, rtx_hit *hits) try {
    return queryCommon(ctx, scene, rays, n, hits, sizeof(rtx_hit), true);
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_trace_occluded(rtx_ctx *ctx, const rtx_scene *scene, const rtx_ray *rays, size_t n, uint8_t *occluded) try {
    return queryCommon(ctx, scene, rays, n, occluded, 1, false);
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_trace_closest_noalloc(rtx_ctx *ctx, const rtx_scene *scene, const rtx_ray *rays, size_t n,
                              rtx_hit *hits) try {
    if (!ctx) return RTX_E_INVALID;
    if (!scene || (n && (!rays || !hits))) return fail(ctx, RTX_E_INVALID, "NULL argument");
    if (n > RTX_NOALLOC_MAX_RAYS)
        return fail(ctx, RTX_E_INVALID, "rtx_trace_closest_noalloc: at most RTX_NOALLOC_MAX_RAYS (65536) rays per call");
    if (n == 0) return RTX_OK;
    if (!ctx->scratch) return fail(ctx, RTX_E_DEVICE, "rtx_trace_closest_noalloc: " + ctx->scratchErr);
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    char *base = static_cast<char *>(ctx->scratch);
    rtx_ray *rD = reinterpret_cast<rtx_ray *>(base);
    rtx_hit *hD = reinterpret_cast<rtx_hit *>(base + kNoAllocRayBytes);
    uint32_t *fetch = reinterpret_cast<uint32_t *>(base + kNoAllocRayBytes + kNoAllocHitBytes);
    hipStream_t s = ctx->stream;
    launchFillU32(s, fetch, 1, 0u); // a kernel, not a memset node: callers may capture the stream
    CHECK_HIP(ctx, hipMemcpyAsync(rD, rays, n * sizeof(rtx_ray), hipMemcpyHostToDevice, s));
    launchQueryP(s, scene->view, rD, (uint32_t)n, hD, true, fetch, ctx->cuCount * ctx->tuning.traceBlocksPerCU);
    CHECK_HIP(ctx, hipGetLastError());
    CHECK_HIP(ctx, hipMemcpyAsync(hits, hD, n * sizeof(rtx_hit), hipMemcpyDeviceToHost, s));
    CHECK_HIP(ctx, hipStreamSynchronize(s));
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

} // extern "C"
