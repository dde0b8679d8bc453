"""ADRRS and main()'s frame structure on the device (librtxamd.so through
the C ABI) against the oracle: frames of several camera samples, splitOnFirst,
the estimate frame (storeEstimate) and ADRRS frames (weight-window roulette
and splitting at cache hits, the queued splits traced after the samples),
path by path, plus the image mean of ADRRS against the plain method and the
error paths.  Bars as in test_gpu_parity: per pixel-frame the device
diverges no more than the FP32 instantiation of the reference logic (+0.5%)
and agrees within 1e-4 of the FP64 oracle where it does not diverge."""
import numpy as np
import pytest

import rtxamd
from rtxamd import abi
from oracle.binding import OracleScene
import scenes
from parity import check_paths, log

pytestmark = pytest.mark.gpu

W = H = 64
KEEP = abi.RTX_FRAME_KEEP_PATHS | abi.RTX_FRAME_RESET_ACCUM


def params(hs, sample, n_frames=1, **kw):
    kw = dict(dict(maxDepth=8, enableMIS=1), **kw)
    return scenes.params(hs, W, H, n_frames, sample_offset=sample, flags=KEEP, **kw)


@pytest.fixture(scope="module")
def cornell():
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    orc = OracleScene(abi, hs.desc())
    # a cache built by the oracle's prepare frames, then the estimate frame
    # (setEstimateRTSettings, src/RayTracingApp.cpp:1206-1217) on it
    st = rtxamd.IcState.empty(3000)
    for f in range(4):
        orc.ic_frame(params(hs, 100 + f, useIrradianceCache=1, useIrradianceCacheOnGlossy=1,
                            isIrradiancePrepareFrame=1, irradianceCreateProb=0.02, irradianceUpdateProb=0.005), st)
    log(f"oracle-built cache: {st.count} records")
    return hs, orc, st


def estimate_params(hs, sample):
    return params(hs, sample, useIrradianceCache=1, useIrradianceCacheOnGlossy=1, samplesPerPixel=16, maxDepth=1,
                  maxFollowDiscrete=10, numNEE=5, storeEstimate=1, irradianceCreateProb=0.0005,
                  irradianceUpdateProb=0.00001)


def adrrs_params(hs, sample, n_frames=1, **kw):
    kw = dict(dict(useADRRS=1, adrrsS=5.0, adrrsSplit=1, useIrradianceCacheOnGlossy=1, irradianceCreateProb=0.002),
              **kw)
    return params(hs, sample, n_frames, **kw)


@pytest.mark.parametrize("spp", [2, 5])
def test_multi_sample_frames_match_oracle(renderer, cornell, spp):
    hs, orc, _ = cornell
    ds = renderer.upload(hs)
    fp = params(hs, 3, samplesPerPixel=spp)
    renderer.render(ds, fp)
    rgb, sig = renderer.read_paths(W * H)
    ids = np.arange(W * H, dtype=np.uint64)
    orgb, osig = orc.render_paths(fp, ids)
    frgb, fsig = orc.render_paths(fp, ids, prec=32)
    # a frame of spp samples diverges when any of its paths does
    check_paths(f"{spp} camera samples per frame", rgb, sig, orgb, osig, (frgb, fsig), abs_same=1 - 0.02 * spp,
                abs_within=0.99)
    ds.close()


def test_split_on_first_matches_oracle(renderer, cornell):
    hs, orc, _ = cornell
    ds = renderer.upload(hs)
    fp = params(hs, 11, 2, splitOnFirst=1)
    renderer.render(ds, fp)
    rgb, sig = renderer.read_paths(2 * W * H)
    ids = np.arange(2 * W * H, dtype=np.uint64)
    orgb, osig = orc.render_paths(fp, ids)
    frgb, fsig = orc.render_paths(fp, ids, prec=32)
    check_paths("splitOnFirst frames", rgb, sig, orgb, osig, (frgb, fsig), abs_same=0.95, abs_within=0.99)
    ds.close()


def test_estimate_frame_matches_oracle(renderer, cornell):
    """storeEstimate: the estimate is each pixel's frame result."""
    hs, orc, st0 = cornell
    renderer.ic_create(st0.max_entries)
    renderer.ic_write(st0)
    ds = renderer.upload(hs)
    fp = estimate_params(hs, 20)
    renderer.render(ds, fp)
    rgb, sig = renderer.read_paths(W * H)
    est = renderer.estimate_read(W, H)
    assert np.array_equal(est[:, :3], rgb[:, :3].astype(np.float32)) and np.all(est[:, 3] == 1.0)
    orgb, osig, _, _, _ = orc.ic_frame(fp, st0.copy())
    frgb, fsig, _, _, _ = orc.ic_frame(fp, st0.copy(), prec=32)
    check_paths("estimate frame", rgb, sig, orgb, osig, (frgb, fsig), abs_same=0.95, abs_within=0.99)
    ds.close()


def test_adrrs_frame_matches_oracle(renderer, cornell):
    """One ADRRS frame on a given cache and estimate: paths, then the cache's
    creations (ADRRS frames keep creating records, no updates)."""
    hs, orc, st0 = cornell
    est = orc.ic_frame(estimate_params(hs, 20), st0.copy())[0]
    rgba = np.concatenate([est, np.ones((W * H, 1))], axis=1).astype(np.float32)
    orc.estimate_set(rgba, W, H)
    renderer.estimate_write(rgba, W, H)
    renderer.ic_create(st0.max_entries)
    renderer.ic_write(st0)
    ds = renderer.upload(hs)
    fp = adrrs_params(hs, 30)
    st = renderer.render(ds, fp)
    rgb, sig = renderer.read_paths(W * H)
    gstate = renderer.ic_read()
    o64 = st0.copy()
    orgb, osig, ojobs, _, _ = orc.ic_frame(fp, o64)
    frgb, fsig, _, _, _ = orc.ic_frame(fp, st0.copy(), prec=32)
    log(f"ADRRS frame: {st.bounces} iterations; records {gstate.count} device, {o64.count} oracle "
        f"({len(ojobs)} creations)")
    check_paths("ADRRS frame", rgb, sig, orgb, osig, (frgb, fsig), abs_same=0.93, abs_within=0.99)
    assert all(j[0] == 1 for j in ojobs) and len(ojobs) > 0
    assert abs(gstate.count - o64.count) <= max(2, len(ojobs) // 10)
    ds.close()


def test_adrrs_render_is_deterministic(renderer, cornell):
    hs, orc, st0 = cornell
    renderer.estimate_write(np.full((W * H, 4), 0.2, np.float32), W, H)
    ds = renderer.upload(hs)
    out = []
    for _ in range(2):
        renderer.ic_create(st0.max_entries)
        renderer.ic_write(st0)
        renderer.render(ds, adrrs_params(hs, 40, 3))
        out.append(renderer.read_paths(3 * W * H))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    ds.close()


def test_adrrs_image_mean_equals_plain(renderer):
    """The weight window moves variance, not the mean: the reference's
    schedule (prepare frames, the estimate frame, then ADRRS frames) against
    plain frames of the same scene."""
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    ds = renderer.upload(hs)
    w = h = 128
    frames = 32
    def mk(sample, n, **kw):
        return scenes.params(hs, w, h, n, sample_offset=sample, **dict(dict(maxDepth=8, enableMIS=1), **kw))

    renderer.ic_create(5000)
    renderer.render(ds, mk(0, 30, useIrradianceCache=1, useIrradianceCacheOnGlossy=1, isIrradiancePrepareFrame=1,
                           irradianceCreateProb=0.01))
    renderer.render(ds, mk(30, 1, useIrradianceCache=1, useIrradianceCacheOnGlossy=1, samplesPerPixel=16,
                           maxDepth=1, maxFollowDiscrete=10, numNEE=5, storeEstimate=1))
    renderer.render(ds, mk(31, frames, useADRRS=1, adrrsS=5.0, adrrsSplit=1, useIrradianceCacheOnGlossy=1,
                           flags=abi.RTX_FRAME_KEEP_PATHS | abi.RTX_FRAME_RESET_ACCUM))
    a, _ = renderer.read_paths(frames * w * h)
    renderer.render(ds, mk(31, frames, flags=abi.RTX_FRAME_KEEP_PATHS | abi.RTX_FRAME_RESET_ACCUM))
    p, _ = renderer.read_paths(frames * w * h)
    va, vp = a[:, :3].mean(1), p[:, :3].mean(1)
    se = np.hypot(va.std() / np.sqrt(va.size), vp.std() / np.sqrt(vp.size))
    log(f"ADRRS mean {va.mean():.5f}, plain {vp.mean():.5f} (se {se:.1e}); per-pixel-frame variance "
        f"{va.var():.3e} vs {vp.var():.3e}")
    assert abs(va.mean() - vp.mean()) <= 4 * se
    ds.close()


def test_recording_with_splitting_renders_nothing(renderer, cornell):
    hs, orc, _ = cornell
    renderer.guiding_set([[[-1, -1, -1], [1, 1, 1]]],
                         [rtxamd.vmm_theta([rtxamd.vmf_theta((0.0, 0.0, 1.0), 2.0)], [1.0])])
    ds = renderer.upload(hs)
    renderer.render(ds, params(hs, 0, updateGuiding=1, splitOnFirst=1))
    rgb, _ = renderer.read_paths(W * H)
    dd = renderer.guiding_read_samples(W, H)
    assert np.all(rgb[:, :3] == 0) and np.all(dd["flags"] == abi.DIR_INVALID)
    ds.close()


def test_adrrs_errors(renderer, cornell):
    hs, orc, st0 = cornell
    ds = renderer.upload(hs)
    renderer.ic_create(100)
    with pytest.raises(rtxamd.RtxError):  # no estimate of this size
        renderer.estimate_write(np.zeros((16, 4), np.float32), 4, 4)
        renderer.render(ds, adrrs_params(hs, 0))
    renderer.estimate_write(np.zeros((W * H, 4), np.float32), W, H)
    with pytest.raises(rtxamd.RtxError):  # window ratio below the reference's range
        renderer.render(ds, adrrs_params(hs, 0, adrrsS=0.5))
    with pytest.raises(rtxamd.RtxError):  # the cache's sub-paths would queue splits
        renderer.render(ds, params(hs, 0, useIrradianceCache=1, splitOnFirst=1))
    with pytest.raises(rtxamd.RtxError):
        renderer.estimate_read(W + 1, H)
    with pytest.raises(rtxamd.RtxError):
        renderer.render(ds, params(hs, 0, samplesPerPixel=5000))
    ds.close()
