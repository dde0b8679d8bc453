// Definitions of the C ABI's opaque handles (rtx_ctx, rtx_scene), shared by
// the library's host files: api.cpp (render, caches, queries) and serve.cpp
// (the serving pool).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <string>
#include <vector>

#include "internal.h"
#include "nrc.h"

using namespace rtx;

// hipMalloc for the library's device buffers.  Developer aid: with
// RTX_POISON_ALLOC set in the environment every new buffer is filled with
// 0xFF bytes (NaN floats, -1 integers) before use, so that a read of memory
// no kernel or copy has written shows up in the results instead of reading
// the zeros that fresh pages happen to hold.
inline hipError_t devMalloc(void **p, size_t bytes) {
    static const bool poison = std::getenv("RTX_POISON_ALLOC") != nullptr;
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess && poison) {
        e = hipMemset(*p, 0xFF, bytes);
        if (e == hipSuccess) e = hipDeviceSynchronize(); // before any stream's use of it
    }
    return e;
}
template <class T> hipError_t devMalloc(T **p, size_t bytes) { return devMalloc(reinterpret_cast<void **>(p), bytes); }

struct rtx_scene {
    DevScene view{};
    std::vector<void *> allocs;
    rtx_scene_info info{};
    int device = 0;
};

namespace rtx {

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    int cls = 0;
};

enum { kClsExtend = 0, kClsShadow = 1, kClsProbe = 2, kClsShade = 3 };
// Bounces the device may run ahead of the host's loop-end check, and the
// poll ring's slots (more than kPollLead, so that a slot is read before the
// bounce that reuses it is launched).
constexpr int kPollLead = 3, kPollSlots = 8;

} // namespace rtx

struct rtx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // accumulation image
    float4 *accum = nullptr;
    float4 *resolved = nullptr;
    uint32_t accW = 0, accH = 0, accSamples = 0;
    // wavefront buffers
    size_t cap = 0;
    uint32_t capNEE = 0;
    PathBuffers pb{};
    Queue q[2]{}, qs{}, qp{};
    int32_t *done = nullptr;
    Counters *cnt = nullptr;
    FrameConst *fc = nullptr;
    // bounce-loop end detection: each bounce's trace kernel writes the queue
    // lengths it starts from into a slot of this pinned ring (kPollSlots
    // slots of 4 words), and an event after the trace marks the slot readable
    uint32_t *pollRing = nullptr;
    hipEvent_t pollEv[kPollSlots] = {};
    std::vector<void *> batchAllocs;
    // per-path results of the last render
    float4 *keepRgb = nullptr;
    uint2 *keepSig = nullptr;
    size_t keepCap = 0, keptPaths = 0;
    // timing
    std::vector<EventPair> events;
    size_t eventsUsed = 0;
    hipEvent_t tStart = nullptr, tStop = nullptr;
    int cuCount = 256;
    rtx_tuning tuning{1, 14, 8, 1u << 22, 20, 1, {0, 0}};
    // irradiance cache (rtx_ic_create)
    IcBuffers ic{};
    std::vector<void *> icAllocs;
    bool icValid = false;
    FrameConst *fcIC = nullptr;      // constants of the cache's stratum sub-paths
    float4 *candBuf = nullptr;       // per-path creation candidates of a batch
    size_t candCap = 0;
    // sub-path results of the last cache frame of a render with RTX_FRAME_KEEP_PATHS
    float4 *keepSub = nullptr;
    uint2 *keepSubSig = nullptr;
    size_t keepSubCap = 0;
    uint32_t keptJobs = 0, keptUpdates = 0;
    // guiding sample records (pc.updateGuiding)
    std::vector<void *> recAllocs;    // per-path record state, recCap paths
    size_t recCap = 0;
    rtx_directional_data *dirData = nullptr;
    size_t dirCap = 0, dirCount = 0;  // slots allocated / slots of the last recording render
    // path guiding (rtx_guiding_set / rtx_guiding_init)
    std::vector<void *> guideAllocs;
    DevGuiding guide{};
    std::vector<rtx_aabb> guideBoxes;
    std::vector<rtx_vmm_theta_wide> guideVmms;
    bool guideVmmsStale = false;      // the device mixtures moved on (updates)
    // the update (rtx_guiding_init): per-region mixtures and work buffers
    std::vector<void *> fitAllocs, fitWorkAllocs;
    GuideFitBuffers fit{};
    bool fitValid = false, fitParallax = true, fitFirst = true;
    // src/PathGuiding.h:118-129; 16 lobes at most (shaders/limits.glsl:6), no region or sample cap
    rtx_guiding_options guideOpts{1, 8192, 4096, 4096, 0.5f, 0.025f, 0, 10000, RTX_MAX_DISTRIBUTIONS, 0, 0, 0};
    // frames of several segments: per-path frame state, split records
    std::vector<void *> segAllocs, splitAllocs;
    size_t segCap = 0, splitCap = 0;
    // ADRRS pixel estimate (estimateImage)
    float4 *estimate = nullptr;
    uint32_t estW = 0, estH = 0;
    // learned radiance cache (rtx_nrc_create): parameters and training work,
    // the render's queries and training vertices
    bool nrcValid = false;
    rtx_nrc_config nrcCfg{};
    NrcBuffers nrc{};
    std::vector<void *> nrcAllocs, nrcPathAllocs, nrcSlotAllocs;
    float *nrcLoss = nullptr;   // device: summed loss of the last steps
    uint32_t nrcLossSteps = 0;
    size_t nrcPathCap = 0, nrcSlotCap = 0;
    // device address of the allocation-free query's scratch array (rtx_trace_closest_noalloc)
    void *scratch = nullptr;
    std::string scratchErr;
};

namespace rtx {
// What lookups read of a cache (dev_icache.h).
inline DevIC icView(const IcBuffers &ic) {
    DevIC v;
    v.entries = ic.entries;
    v.spheres = ic.spheres;
    v.bucketStart = ic.bucketStart;
    v.items = ic.items;
    v.meta = ic.meta;
    v.nb = ic.nb;
    return v;
}
} // namespace rtx
