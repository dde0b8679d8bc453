"""The frame schedule (rtx_schedule_*, RayTracingApp::drawFrame/raytrace,
src/RayTracingApp.cpp:91-215, 1120-1217) planned on the host: no device
work happens in begin_frame/end_frame, so these run without a GPU.  Each
test follows one behaviour of the reference's loop frame by frame:
accumulation restarts, irradiance-cache prepare frames, ADRRS's switch to
the estimate frame and back, the guiding cadence and the eval counters."""
import time

import pytest

import rtxamd
from rtxamd import abi


def base():
    fp = abi.FrameParams()
    fp.width, fp.height, fp.seed = 8, 4, 7
    fp.flags = abi.RTX_FRAME_KEEP_PATHS
    return fp


def pc_defaults(**kw):
    pc = rtxamd.push_constant_defaults()
    for k, v in kw.items():
        setattr(pc, k, v)
    return pc


def run(sched, n):
    frames = []
    for _ in range(n):
        f = sched.begin_frame(base())
        frames.append(f)
        sched.end_frame()
    return frames


def reset(f):
    return bool(f.flags & abi.RTX_FRAME_RESET_ACCUM)


def test_accumulation_restarts_on_input_change():
    s = rtxamd.Schedule(pc_defaults(), prepareFrames=0)
    fr = run(s, 3)
    assert [f.pc.previousFrames for f in fr] == [0, 1, 2]
    assert [reset(f) for f in fr] == [True, False, False]
    # every frame is one frame of the reference: spp 1, consecutive sample offsets
    assert [f.spp for f in fr] == [1, 1, 1] and [f.sampleOffset for f in fr] == [0, 1, 2]
    assert all(f.flags & abi.RTX_FRAME_KEEP_PATHS for f in fr) and all(f.seed == 7 for f in fr)
    s.input_changed()
    fr = run(s, 2)
    assert [f.pc.previousFrames for f in fr] == [0, 1] and [reset(f) for f in fr] == [True, False]
    s.set_push_constants(pc_defaults(maxDepth=3))
    f = run(s, 1)[0]
    assert f.pc.maxDepth == 3 and f.pc.previousFrames == 0 and reset(f)
    # without accumulation every frame starts over
    s2 = rtxamd.Schedule(pc_defaults(), accumulate=0)
    assert all(reset(f) and f.pc.previousFrames == 0 for f in run(s2, 3))


def test_irradiance_cache_prepare_frames():
    s = rtxamd.Schedule(pc_defaults(useIrradianceCache=1), prepareFrames=3)
    fr = run(s, 6)
    assert [f.pc.isIrradiancePrepareFrame for f in fr] == [1, 1, 1, 0, 0, 0]
    # prepare frames carry previousFrames = -1; the first real frame starts the accumulation
    assert [f.pc.previousFrames for f in fr] == [0xFFFFFFFF] * 3 + [0, 1, 2]
    assert [reset(f) for f in fr] == [True, True, True, True, False, False]
    assert s.state().currentPrepareFrames == 3


def test_adrrs_prepare_estimate_and_restore():
    pc = pc_defaults(useADRRS=1, samplesPerPixel=2, maxDepth=9, numNEE=2, useIrradianceCacheOnGlossy=0)
    s = rtxamd.Schedule(pc, prepareFrames=2)
    fr = run(s, 5)
    # prepare frames are cache frames with lookups on glossy surfaces
    for f in fr[:2]:
        assert (f.pc.useIrradianceCache, f.pc.useADRRS, f.pc.isIrradiancePrepareFrame,
                f.pc.useIrradianceCacheOnGlossy) == (1, 0, 1, 1)
    # the estimate frame: setEstimateRTSettings (src/RayTracingApp.cpp:1206-1217)
    e = fr[2].pc
    assert (e.storeEstimate, e.samplesPerPixel, e.useIrradianceCache, e.maxDepth, e.maxFollowDiscrete, e.numNEE,
            e.useADRRS, e.enableNEE, e.isIrradiancePrepareFrame) == (1, 16, 1, 1, 10, 5, 0, 1, 0)
    # then the saved settings with ADRRS on, the cache method off
    for f in fr[3:]:
        p = f.pc
        assert (p.useADRRS, p.useIrradianceCache, p.storeEstimate, p.samplesPerPixel, p.maxDepth, p.numNEE,
                p.useIrradianceCacheOnGlossy) == (1, 0, 0, 2, 9, 2, 1)
    # the estimate frame starts the accumulation and the ADRRS frames add to it
    assert [f.pc.previousFrames for f in fr[2:]] == [0, 1, 2]
    assert [reset(f) for f in fr[2:]] == [True, False, False]


def test_guiding_cadence():
    s = rtxamd.Schedule(pc_defaults(updateGuiding=1, useGuiding=1), guidingOptimizations=2)
    fr = run(s, 5)
    # three recording frames: the first two are followed by an update, the
    # third is recorded without one and recording stops
    assert [f.pc.updateGuiding for f in fr] == [1, 1, 1, 0, 0]
    assert [bool(f.flags & abi.RTX_FRAME_NO_GUIDING_UPDATE) for f in fr] == [False, False, True, False, False]
    st = s.state()
    assert st.currentGuidingOptimizations == -1 and st.guidingUpdateMs >= 0 and st.pc.updateGuiding == 0
    # no limit: every recording frame is followed by an update
    s2 = rtxamd.Schedule(pc_defaults(updateGuiding=1), guidingOptimizations=-1)
    fr = run(s2, 4)
    assert all(f.pc.updateGuiding == 1 and not f.flags & abi.RTX_FRAME_NO_GUIDING_UPDATE for f in fr)
    assert s2.state().currentGuidingOptimizations == 3


def test_eval_samples_counts_outside_prepare_frames():
    """The count skips frames while currentPrepareFrames < prepareFrames after
    the frame: the last prepare frame has already raised the counter to
    prepareFrames, so its samples count (src/RayTracingApp.cpp:158-161)."""
    s = rtxamd.Schedule(pc_defaults(useIrradianceCache=1, samplesPerPixel=2), prepareFrames=2, evalSamples=6,
                        sppToOneAfterEval=1, accumulate=0)
    counts = []
    for k in range(4):
        f = s.begin_frame(base())
        assert reset(f) == (k in (0, 1, 2))  # prepare frames, then eval switched accumulation on
        s.end_frame()
        counts.append(s.state().evalCurrentSamples)
    assert counts == [0, 2, 4, 0]  # the fourth frame reaches 6 and the eval ends (counter reset)
    st = s.state()
    assert st.evalSamplesRunning == 0 and st.evalSamplesMs >= 0 and st.pc.samplesPerPixel == 1


def test_eval_seconds():
    s = rtxamd.Schedule(pc_defaults(samplesPerPixel=3), evalSeconds=1)
    run(s, 2)
    assert s.state().evalSecondsRunning == 1 and s.state().evalSecondsTotalSamples == 6
    time.sleep(1.05)
    run(s, 1)
    st = s.state()
    assert st.evalSecondsRunning == 0 and st.evalSecondsResult == 9


def test_schedule_errors():
    s = rtxamd.Schedule(pc_defaults())
    with pytest.raises(rtxamd.RtxError):
        s.end_frame()  # nothing planned
    s.begin_frame(base())
    with pytest.raises(rtxamd.RtxError):
        s.begin_frame(base())  # the planned frame has not ended
    with pytest.raises(rtxamd.RtxError):
        rtxamd.Schedule(pc_defaults(), prepareFrames=-1)
