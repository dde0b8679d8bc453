// Frame schedule: the per-frame host logic of RayTracingApp::drawFrame and
// RayTracingApp::raytrace (src/RayTracingApp.cpp:91-215, 1120-1163,
// 1206-1217) around rtx_render.  Plain C++: it plans frames and calls the
// library's entry points.
#include <chrono>
#include <cstring>
#include <new>

#include "abi_guard.h"
#include "rtx_amd.h"

using Clock = std::chrono::steady_clock;

struct rtx_schedule {
    rtx_schedule_config cfg;
    rtx_push_constant pc;           // rtPushConstants
    bool accumulateResults = false;
    bool hasInputChanged = true;    // the first frame starts the accumulation
    uint32_t previousFrames = 0;
    int currentPrepareFrames = 0;
    bool activateADRRSAfterPrepareFrames = false;
    rtx_push_constant backup;       // backupPushConstant
    bool loadBackupNextIteration = false;
    int currentGuidingOptimizations = -1;
    Clock::time_point guidingUpdateStart;
    double guidingUpdateMs = -1.0;
    bool evalXSamples = false;
    int evalCurrentSamples = 0;
    Clock::time_point evalXSamplesStart;
    double evalXSamplesMs = -1.0;
    bool evalXSeconds = false;
    bool evalSecondsStarted = false;
    Clock::time_point evalXSecondsEnd;
    int64_t evalXSecondsTotalSamples = 0;
    int64_t evalXSecondsLastResult = 0;
    uint64_t frames = 0;
    bool planned = false; // begin_frame ran, end_frame not yet
};

namespace {

double msSince(Clock::time_point t) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

// setEstimateRTSettings, src/RayTracingApp.cpp:1206-1217
void setEstimateRTSettings(rtx_push_constant &pc) {
    pc.samplesPerPixel = 16;
    pc.useIrradianceCache = 1;
    pc.useIrradianceCacheOnGlossy = 1;
    pc.enableNEE = 1;
    pc.visualizeMode = 0; // ERayTrace
    pc.maxDepth = 1;
    pc.maxFollowDiscrete = 10;
    pc.showIrradianceCacheOnly = 0;
    pc.numNEE = 5;
    pc.useADRRS = 0;
}

} // namespace

extern "C" {

void rtx_schedule_config_defaults(rtx_schedule_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->prepareFrames = 50;        // src/RayTracingApp.h:176
    c->guidingOptimizations = 6;  // src/RayTracingApp.h:180
    c->accumulate = 1;
    c->evalSamples = 0;
    c->evalSeconds = 0;
    c->sppToOneAfterEval = 0;
}

int rtx_schedule_create(const rtx_schedule_config *cfg, const rtx_push_constant *pc, rtx_schedule **out) try {
    if (!cfg || !pc || !out) return RTX_E_INVALID;
    *out = nullptr;
    if (cfg->prepareFrames < 0 || cfg->guidingOptimizations < -1 || cfg->evalSamples < 0 || cfg->evalSeconds < 0)
        return RTX_E_INVALID;
    auto *s = new (std::nothrow) rtx_schedule();
    if (!s) return RTX_E_MEMORY;
    s->cfg = *cfg;
    s->pc = *pc;
    s->backup = *pc;
    s->accumulateResults = cfg->accumulate != 0;
    s->evalXSamples = cfg->evalSamples > 0;
    s->evalXSeconds = cfg->evalSeconds > 0;
    *out = s;
    return RTX_OK;
} catch (...) {
    return rtx::abiStatus();
}

int rtx_schedule_destroy(rtx_schedule *s) try {
    delete s;
    return RTX_OK;
} catch (...) {
    return rtx::abiStatus();
}

int rtx_schedule_set_push_constants(rtx_schedule *s, const rtx_push_constant *pc) try {
    if (!s || !pc) return RTX_E_INVALID;
    s->pc = *pc;
    s->hasInputChanged = true;
    return RTX_OK;
} catch (...) {
    return rtx::abiStatus();
}

int rtx_schedule_input_changed(rtx_schedule *s) try {
    if (!s) return RTX_E_INVALID;
    s->hasInputChanged = true;
    return RTX_OK;
} catch (...) {
    return rtx::abiStatus();
}

int rtx_schedule_begin_frame(rtx_schedule *s, const rtx_frame_params *base, rtx_frame_params *frame) try {
    if (!s || !base || !frame) return RTX_E_INVALID;
    if (s->planned) return RTX_E_STATE;
    // drawFrame: the eval modes switch accumulation on (src/RayTracingApp.cpp:91-100)
    if (s->evalXSamples && s->evalCurrentSamples == 0) {
        s->evalXSamplesStart = Clock::now();
        s->accumulateResults = true;
        s->hasInputChanged = true;
    }
    if (s->evalXSeconds && s->evalXSecondsTotalSamples == 0) {
        if (!s->evalSecondsStarted) { // evalXSecondsEnd from the start (src/RayTracingApp.cpp:17-20)
            s->evalXSecondsEnd = Clock::now() + std::chrono::seconds(s->cfg.evalSeconds);
            s->evalSecondsStarted = true;
        }
        s->accumulateResults = true;
        s->hasInputChanged = true;
    }
    rtx_push_constant &pc = s->pc;
    // raytrace(): accumulation counter (src/RayTracingApp.cpp:1123-1127)
    if (s->accumulateResults && !s->hasInputChanged) s->previousFrames += 1;
    else s->previousFrames = 0;
    // irradiance-cache prepare frames and the switch to ADRRS (1129-1156)
    if (pc.useIrradianceCache || pc.useADRRS) {
        if (s->currentPrepareFrames < s->cfg.prepareFrames) {
            if (pc.useADRRS) {
                pc.useIrradianceCacheOnGlossy = 1;
                pc.useIrradianceCache = 1;
                pc.useADRRS = 0;
                s->activateADRRSAfterPrepareFrames = true;
            }
            pc.isIrradiancePrepareFrame = 1;
            s->previousFrames = (uint32_t)-1;
            s->currentPrepareFrames++;
        } else if (s->activateADRRSAfterPrepareFrames && s->currentPrepareFrames == s->cfg.prepareFrames) {
            pc.storeEstimate = 1;
            pc.useIrradianceCache = 0;
            pc.isIrradiancePrepareFrame = 0;
            pc.useADRRS = 1;
            s->activateADRRSAfterPrepareFrames = false;
            s->currentPrepareFrames++;
        } else {
            pc.isIrradiancePrepareFrame = 0;
        }
    } else {
        pc.isIrradiancePrepareFrame = 0;
    }
    // the estimate frame renders with setEstimateRTSettings, the next frame
    // restores the settings (1158-1170)
    if (pc.storeEstimate) {
        if (s->loadBackupNextIteration) {
            pc = s->backup;
            pc.storeEstimate = 0;
            s->loadBackupNextIteration = false;
        } else {
            s->backup = pc;
            setEstimateRTSettings(pc);
            s->loadBackupNextIteration = true;
        }
    }
    pc.previousFrames = s->previousFrames;
    *frame = *base;
    frame->pc = pc;
    frame->spp = 1;
    frame->sampleOffset = (uint32_t)s->frames;
    // previousFrames 0 starts the accumulation over; prepare frames (-1) are
    // not accumulated, and the frame after them starts over too
    if (s->previousFrames == 0 || s->previousFrames == (uint32_t)-1) frame->flags |= RTX_FRAME_RESET_ACCUM;
    else frame->flags &= ~RTX_FRAME_RESET_ACCUM;
    // the recording frame after the last optimization gets no update (133-139)
    if (pc.updateGuiding) {
        int next = s->currentGuidingOptimizations == -1 ? 0 : s->currentGuidingOptimizations + 1;
        if (s->cfg.guidingOptimizations != -1 && next >= s->cfg.guidingOptimizations)
            frame->flags |= RTX_FRAME_NO_GUIDING_UPDATE;
    }
    s->hasInputChanged = false;
    s->planned = true;
    return RTX_OK;
} catch (...) {
    return rtx::abiStatus();
}

int rtx_schedule_end_frame(rtx_schedule *s) try {
    if (!s) return RTX_E_INVALID;
    if (!s->planned) return RTX_E_STATE;
    s->planned = false;
    s->frames++;
    rtx_push_constant &pc = s->pc;
    // guiding cadence (src/RayTracingApp.cpp:125-152)
    if (pc.updateGuiding) {
        if (s->currentGuidingOptimizations == -1) {
            s->currentGuidingOptimizations = 0;
            s->guidingUpdateStart = Clock::now();
        } else {
            s->currentGuidingOptimizations++;
        }
        if (s->cfg.guidingOptimizations != -1 && s->currentGuidingOptimizations >= s->cfg.guidingOptimizations) {
            s->guidingUpdateMs = msSince(s->guidingUpdateStart);
            pc.updateGuiding = 0;
            s->currentGuidingOptimizations = -1;
        }
    } else {
        s->currentGuidingOptimizations = -1;
    }
    // eval counters (158-215): prepare frames collect no samples
    const bool prepare = pc.useIrradianceCache && s->currentPrepareFrames < s->cfg.prepareFrames;
    if (s->evalXSamples) {
        if (!prepare) s->evalCurrentSamples += pc.samplesPerPixel;
        if (s->evalCurrentSamples >= s->cfg.evalSamples) {
            s->evalXSamplesMs = msSince(s->evalXSamplesStart);
            s->evalXSamples = false;
            s->evalCurrentSamples = 0;
            if (s->cfg.sppToOneAfterEval) pc.samplesPerPixel = 1;
        }
    }
    if (s->evalXSeconds) {
        if (!prepare) s->evalXSecondsTotalSamples += pc.samplesPerPixel;
        if (Clock::now() > s->evalXSecondsEnd) {
            s->evalXSecondsLastResult = s->evalXSecondsTotalSamples;
            s->evalXSeconds = false;
            s->evalXSecondsTotalSamples = 0;
            if (s->cfg.sppToOneAfterEval) pc.samplesPerPixel = 1;
        }
    }
    return RTX_OK;
} catch (...) {
    return rtx::abiStatus();
}

int rtx_schedule_frame(rtx_ctx *ctx, const rtx_scene *scene, rtx_schedule *s, const rtx_frame_params *base,
                       rtx_frame_stats *stats) try {
    if (!ctx || !scene || !s || !base) return RTX_E_INVALID;
    rtx_frame_params frame;
    const rtx_schedule before = *s;
    int rc = rtx_schedule_begin_frame(s, base, &frame);
    if (rc != RTX_OK) return rc;
    rc = rtx_render(ctx, scene, &frame, stats);
    if (rc != RTX_OK) {
        *s = before; // the frame did not happen
        return rc;
    }
    return rtx_schedule_end_frame(s);
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_schedule_get_state(const rtx_schedule *s, rtx_schedule_state *st) try {
    if (!s || !st) return RTX_E_INVALID;
    std::memset(st, 0, sizeof(*st));
    st->frames = s->frames;
    st->currentPrepareFrames = s->currentPrepareFrames;
    st->currentGuidingOptimizations = s->currentGuidingOptimizations;
    st->previousFrames = s->previousFrames;
    st->evalSamplesRunning = s->evalXSamples;
    st->evalSecondsRunning = s->evalXSeconds;
    st->evalCurrentSamples = s->evalCurrentSamples;
    st->evalSecondsTotalSamples = s->evalXSecondsTotalSamples;
    st->evalSecondsResult = s->evalXSecondsLastResult;
    st->evalSamplesMs = s->evalXSamplesMs;
    st->guidingUpdateMs = s->guidingUpdateMs;
    st->pc = s->pc;
    return RTX_OK;
} catch (...) {
    return rtx::abiStatus();
}

} // extern "C"
