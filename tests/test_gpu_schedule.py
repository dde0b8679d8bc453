"""The frame schedule on the device: rtx_schedule_frame renders exactly the
frames its plan describes (the same images as planning with
begin_frame/end_frame and calling rtx_render by hand, and the oracle's
images for the same plan, frame by frame), the ADRRS sequence
(prepare frames, estimate frame, ADRRS frames) runs end to end with the
estimate frame in the average as in the reference, and the guiding cadence
skips the update after the last recording frame."""
import numpy as np
import pytest

import rtxamd
import scenes
from parity import log
from rtxamd import abi

pytestmark = pytest.mark.gpu

W, H = 48, 32


def base(hs, **kw):
    fp = scenes.params(hs, W, H, 1, flags=0, **kw)
    return fp


def samples(renderer):
    n = rtxamd.C.c_uint32()
    renderer._check(rtxamd.device_lib().rtx_accumulated_samples(renderer._h, rtxamd.C.byref(n)))
    return n.value


@pytest.fixture(scope="module")
def cornell():
    return rtxamd.HostScene.generate(abi.SCENE_CORNELL)


def test_scheduled_frames_equal_planned_renders(renderer, cornell):
    hs = cornell
    ds = renderer.upload(hs)
    fp = base(hs, useIrradianceCache=1, useIrradianceCacheOnGlossy=1, maxDepth=6, enableMIS=1,
              irradianceCreateProb=0.01)
    images = []
    for mode in ("schedule", "manual"):
        renderer.ic_create(2000)
        s = rtxamd.Schedule(fp.pc, prepareFrames=3)
        for _ in range(6):
            if mode == "schedule":
                s.frame(renderer, ds, fp)
            else:
                f = s.begin_frame(fp)
                renderer.render(ds, f)
                s.end_frame()
        images.append(renderer.read_image(W, H))
        assert samples(renderer) == 3  # prepare frames are not accumulated
        assert s.state().frames == 6
    assert np.array_equal(images[0], images[1])
    ds.close()


def test_scheduled_cache_frames_match_oracle(renderer, cornell):
    """The schedule's frames on the device against the oracle following the
    same plan frame by frame from the same empty cache (orc_ic_frame: the
    frame and its upkeep): three prepare frames, then five cache frames.  The
    device accumulates exactly the cache frames, and its image's RelMSE
    against the FP64 oracle's average is at most 3x that of the reference
    logic in FP32 (+1e-6), the bar of the per-path parity tests carried to the
    image (each run grows its own cache, so FP32 branch flips move records)."""
    from oracle.binding import OracleScene
    from parity import relmse
    hs = cornell
    orc = OracleScene(abi, hs.desc())
    ds = renderer.upload(hs)
    fp = base(hs, useIrradianceCache=1, maxDepth=6, enableMIS=1, irradianceCreateProb=0.01,
              irradianceUpdateProb=0.01)
    renderer.ic_create(2000)
    s = rtxamd.Schedule(fp.pc, prepareFrames=3)
    st64, st32 = rtxamd.IcState.empty(2000), rtxamd.IcState.empty(2000)
    acc64, acc32, n = np.zeros((W * H, 3)), np.zeros((W * H, 3)), 0
    for _ in range(8):
        f = s.begin_frame(fp)
        renderer.render(ds, f)
        s.end_frame()
        rgb64 = orc.ic_frame(f, st64, prec=64)[0]
        rgb32 = orc.ic_frame(f, st32, prec=32)[0]
        if not f.pc.isIrradiancePrepareFrame:
            acc64 += rgb64
            acc32 += rgb32
            n += 1
    assert n == 5 and samples(renderer) == n
    img = renderer.read_image(W, H)[..., :3].astype(np.float64)
    ref = (acc64 / n).reshape(H, W, 3)
    e, e32 = relmse(img, ref), relmse((acc32 / n).reshape(H, W, 3), ref)
    dev = renderer.ic_read()
    log(f"scheduled cache frames: records device {dev.count} / oracle {st64.count}; image relmse vs FP64 "
        f"{e:.3e} (reference logic in FP32 {e32:.3e}), mean {img.mean():.5f} vs {ref.mean():.5f}")
    assert abs(dev.count - st64.count) <= max(3, 0.05 * st64.count)
    assert e <= 3 * e32 + 1e-6
    ds.close()


def test_adrrs_schedule_end_to_end(renderer, cornell):
    hs = cornell
    ds = renderer.upload(hs)
    fp = base(hs, useADRRS=1, adrrsS=5.0, adrrsSplit=1, maxDepth=8, enableMIS=1, irradianceCreateProb=0.01)
    renderer.ic_create(3000)
    s = rtxamd.Schedule(fp.pc, prepareFrames=4)
    for k in range(8):
        s.frame(renderer, ds, fp)
    img = renderer.read_image(W, H)
    # the estimate frame and 3 ADRRS frames make the average (previousFrames 0..3)
    assert samples(renderer) == 4
    assert np.isfinite(img).all() and img[..., :3].mean() > 0
    st = s.state()
    assert st.pc.useADRRS == 1 and st.pc.storeEstimate == 0 and st.currentPrepareFrames == 5
    est = renderer.estimate_read(W, H)
    assert np.isfinite(est).all() and est[:, :3].mean() > 0
    log(f"ADRRS schedule: image mean {img[..., :3].mean():.4f}, estimate mean {est[:, :3].mean():.4f}")
    ds.close()


def test_guiding_cadence_skips_the_last_update(renderer, cornell):
    hs = cornell
    ds = renderer.upload(hs)
    box = ds.bounds()
    fp = base(hs, updateGuiding=1, useGuiding=1, guidingProb=0.5, maxDepth=6, enableMIS=1)
    # scheduled: two recording frames, the update after the first only
    renderer.guiding_init(box, 2)
    s = rtxamd.Schedule(fp.pc, guidingOptimizations=1)
    s.frame(renderer, ds, fp)
    s.frame(renderer, ds, fp)
    assert s.state().pc.updateGuiding == 0
    a_boxes, a_vmms = renderer.guiding_get()
    # by hand: one recording frame (the context updates after it)
    renderer.guiding_init(box, 2)
    f = rtxamd.Schedule(fp.pc, guidingOptimizations=1).begin_frame(fp)
    renderer.render(ds, f)
    b_boxes, b_vmms = renderer.guiding_get()
    assert np.array_equal(np.asarray(a_boxes), np.asarray(b_boxes))
    assert len(a_vmms) == len(b_vmms)
    for x, y in zip(a_vmms, b_vmms):
        assert bytes(x) == bytes(y)
    ds.close()


def test_command_line_renders_with_cache_and_guiding(tmp_path):
    """bin/rtx_pathtracer WIDTH HEIGHT IC_SIZE GUIDING_SPLITS scene: the cache
    is created from IC_SIZE and the guiding regions from GUIDING_SPLITS
    (src/RayTracingApp.cpp:44-46, 346-351), the schedule runs the prepare
    frames, the guiding cadence and eval-by-samples, and the EXR holds the
    accumulated image, finite, lit like a plain render of the same scene (the
    cache's interpolation bias is small next to a 4-spp image's noise at this
    scale).  --spp 4 accumulates 3 samples per pixel: the reference's eval
    counter checks currentPrepareFrames after raytrace() has already counted
    the last prepare frame (src/RayTracingApp.cpp:1131-1142, 160-163), so
    that frame's sample is counted though it is not accumulated, and six
    frames run: three prepare frames, three accumulated ones."""
    import subprocess
    from pathlib import Path
    exe = Path(rtxamd.LIB_DIR).parent / "bin" / "rtx_pathtracer"
    args = [str(exe), "64", "48", "400", "3", "gen:cornell", "--ic", "1", "--guiding", "1", "--spp", "4",
            "--prepare", "3", "--guiding-updates", "2", "--depth", "6", "--out", str(tmp_path)]
    log("cli: " + " ".join(args[1:]))
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    log(r.stdout.strip())
    assert r.returncode == 0, r.stderr
    assert ", 6 frames, 3 spp accumulated" in r.stdout
    files = sorted(tmp_path.glob("gen_cornell_3spp_*ms.exr"))
    assert len(files) == 1, list(tmp_path.iterdir())
    img = rtxamd.read_exr(files[0])
    assert img.shape == (48, 64, 4) and np.isfinite(img).all() and (img[..., :3] >= 0).all()
    with rtxamd.Renderer(0) as rr:
        hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL, 1)
        ds = rr.upload(hs)
        fp = scenes.params(hs, 64, 48, 64, maxDepth=6, enableMIS=1)
        rr.render(ds, fp)
        ref = rr.read_image(64, 48)
        ds.close()
    m, mr = img[..., :3].mean(), ref[..., :3].mean()
    log(f"cli image mean {m:.4f}, plain 64 spp {mr:.4f}")
    assert abs(m - mr) < 0.15 * mr
