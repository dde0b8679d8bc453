"""The serving pool (rtx_server_*, librtxamd.so through the C ABI).

Every path draws from its own (pixel, sample) stream and a job's batches are
added to its image in batch order, so:
  * a job's image equals rtx_render of the same frame with batchPaths set to
    the job's batch size, bit for bit (plain method; the multi-job builds of
    the bounce kernels against the single-job ones);
  * a job's image is the same bit for bit alone and in a mix of jobs of other
    scenes, sizes and methods running in the same launches (cache lookups,
    guided sampling, textures, the alpha test, an environment map, NEE
    counts);
and admission follows the framebuffer budget: jobs whose images do not fit
wait until earlier images are read."""
import ctypes as C

import numpy as np
import pytest

import rtxamd
from rtxamd import abi
import scenes
from parity import log

pytestmark = pytest.mark.gpu


def bits(img):
    return np.ascontiguousarray(img).view(np.uint32)


def test_job_equals_render(renderer):
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL, 1)
    ds = renderer.upload(hs)
    slot = 1 << 12
    with rtxamd.Server(renderer, slotPaths=slot, slots=4) as sv:
        # 1920 pixels: two sample planes per batch; 7680 pixels: pixel ranges of one plane
        for w, h, spp in [(48, 40, 6), (96, 80, 2)]:
            fp = scenes.params(hs, w, h, spp, maxDepth=8, enableMIS=1, enableRR=1, seed=5)
            log(f"serve vs render: cornell {w}x{h}x{spp}")
            jid = sv.submit(ds, fp)
            st = sv.wait(jid, timeout=60)
            assert st.state == abi.RTX_JOB_DONE and st.paths == w * h * spp
            img = sv.read_image(jid)
            npix = w * h
            batch = min(slot, npix * spp)
            if batch >= npix:
                batch -= batch % npix
            fp.batchPaths = batch
            renderer.render(ds, fp)
            ref = renderer.read_image(w, h)
            assert st.batches == -(-npix * spp // batch)
            assert np.array_equal(bits(img), bits(ref)), np.abs(img - ref).max()
            assert img[..., :3].mean() > 0.01
    ds.close()


def _ic_state(hs):
    """A context whose cache was built by three prepare frames of its own."""
    st = rtxamd.Renderer(0)
    st.ic_create(2000)
    ds = st.upload(hs)
    for f in range(3):
        fp = scenes.params(hs, 64, 48, 1, maxDepth=8, enableMIS=1, useIrradianceCache=1,
                           isIrradiancePrepareFrame=1, irradianceCreateProb=0.05, sample_offset=50 + f)
        st.render(ds, fp)
    n = st.ic_read().count
    log(f"state cache: {n} records")
    assert n > 10
    return st, ds


def _guide_state(hs, box):
    st = rtxamd.Renderer(0)
    st.guiding_init(box, 3)
    return st


def _jobs(renderer):
    """(scene, params, state) of a mix of jobs: several scenes, sizes, sample
    counts and methods."""
    cornell = rtxamd.HostScene.generate(abi.SCENE_CORNELL, 1)
    room = rtxamd.HostScene.generate(abi.SCENE_ROOM, 1)
    zoo = scenes.material_zoo()
    gallery = scenes.env_gallery()
    y, x = np.mgrid[0:8, 0:8]
    alpha_tex = np.stack([np.full_like(x, 200), np.full_like(x, 100), np.full_like(x, 50),
                          ((x + y) % 4) * 80], -1).astype(np.uint8)
    plate = scenes.textured_plate(alpha_tex, alpha_black=True)
    dev = {name: renderer.upload(h) for name, h in
           [("cornell", cornell), ("room", room), ("zoo", zoo), ("gallery", gallery), ("plate", plate)]}
    ic_state, ic_ds = _ic_state(cornell)
    zb = dev["zoo"].bounds()
    g_state = _guide_state(zoo, zb)
    jobs = [
        ("cornell", dev["cornell"], scenes.params(cornell, 64, 48, 4, maxDepth=8, enableMIS=1, enableRR=1), None),
        ("room", dev["room"], scenes.params(room, 80, 45, 2, maxDepth=12, enableMIS=1), None),
        ("zoo nee2", dev["zoo"], scenes.params(zoo, 64, 64, 3, maxDepth=12, enableMIS=1, numNEE=2, seed=3), None),
        ("gallery", dev["gallery"], scenes.params(gallery, 48, 48, 2, maxDepth=6, enableMIS=1), None),
        ("alpha plate", dev["plate"], scenes.params(plate, 32, 32, 2, maxDepth=3), None),
        ("cornell cache", dev["cornell"],
         scenes.params(cornell, 64, 48, 2, maxDepth=8, enableMIS=1, useIrradianceCache=1, sample_offset=9), ic_state),
        ("zoo guided", dev["zoo"],
         scenes.params(zoo, 48, 48, 2, maxDepth=8, enableMIS=1, useGuiding=1, guidingProb=0.5, seed=7), g_state),
    ]
    keep = (dev, ic_state, ic_ds, g_state)
    return jobs, keep


def test_jobs_same_alone_and_in_mix(renderer):
    jobs, keep = _jobs(renderer)
    with rtxamd.Server(renderer, slotPaths=1 << 12, slots=8, maxNumNEE=2) as sv:
        alone = []
        for name, ds, fp, state in jobs:
            log(f"alone: {name}")
            jid = sv.submit(ds, fp, state)
            sv.wait(jid, timeout=60)
            alone.append(sv.read_image(jid))
        log("mix: all jobs at once")
        ids = [sv.submit(ds, fp, state) for _, ds, fp, state in jobs]
        mixed = []
        for jid in ids:
            st = sv.wait(jid, timeout=60)
            assert st.state == abi.RTX_JOB_DONE and st.latencyMs > 0
            mixed.append(sv.read_image(jid))
        stats = sv.stats()
        log(f"mix: peak slots {stats.peakSlotsBusy}, peak jobs {stats.peakJobsRunning}, bounces {stats.bounces}, "
            f"p50 {stats.latencyP50Ms:.1f} ms, p99 {stats.latencyP99Ms:.1f} ms")
        assert stats.peakJobsRunning >= 3 and stats.peakSlotsBusy >= 4
        for (name, _, fp, _), a, b in zip(jobs, alone, mixed):
            assert np.isfinite(a).all() and a[..., :3].mean() > 0.0, name
            assert np.array_equal(bits(a), bits(b)), (name, np.abs(a - b).max())
    dev, ic_state, ic_ds, g_state = keep
    for ds in dev.values():
        ds.close()
    ic_ds.close()
    ic_state.close()
    g_state.close()


def test_admission_follows_the_framebuffer_budget(renderer):
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL, 1)
    ds = renderer.upload(hs)
    w, h = 64, 48
    image_bytes = w * h * 32
    with rtxamd.Server(renderer, slotPaths=1 << 12, slots=4, framebufferBytes=3 * image_bytes + 1024) as sv:
        fps = [scenes.params(hs, w, h, 2, maxDepth=6, enableMIS=1, seed=11 + k) for k in range(6)]
        ids = [sv.submit(ds, fp) for fp in fps]
        for jid in ids[:3]:
            sv.wait(jid, timeout=60)
        # three images are held and unread: the fourth job cannot be admitted
        st = sv.poll(ids[3])
        assert st.state == abi.RTX_JOB_QUEUED
        images = {}
        for k, jid in enumerate(ids):
            st = sv.wait(jid, timeout=60)
            assert st.state == abi.RTX_JOB_DONE
            images[k] = sv.read_image(jid)  # frees budget for the next queued job
        stats = sv.stats()
        assert stats.jobsDone == 6 and stats.peakJobsRunning <= 3
        assert stats.framebufferPeakBytes <= 3 * image_bytes + 1024
        for k, fp in enumerate(fps):
            fp.batchPaths = (1 << 12) - (1 << 12) % (w * h)
            renderer.render(ds, fp)
            assert np.array_equal(bits(images[k]), bits(renderer.read_image(w, h)))
        # errors
        big = scenes.params(hs, 256, 256, 1)
        with pytest.raises(rtxamd.RtxError) as e:
            sv.submit(ds, big)
        assert e.value.code == abi.RTX_E_MEMORY
        for kw, code in [({"useADRRS": 1}, abi.RTX_E_UNSUPPORTED), ({"samplesPerPixel": 2}, abi.RTX_E_UNSUPPORTED),
                         ({"updateGuiding": 1}, abi.RTX_E_UNSUPPORTED), ({"numNEE": 2}, abi.RTX_E_INVALID),
                         ({"useIrradianceCache": 1}, abi.RTX_E_STATE), ({"useGuiding": 1}, abi.RTX_E_STATE)]:
            with pytest.raises(rtxamd.RtxError) as e:
                sv.submit(ds, scenes.params(hs, 8, 8, 1, **kw))
            assert e.value.code == code, kw
        with pytest.raises(rtxamd.RtxError) as e:
            sv.poll(12345)
        assert e.value.code == abi.RTX_E_INVALID
    ds.close()


def test_latency_statistics_and_shutdown(renderer):
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL, 1)
    ds = renderer.upload(hs)
    sv = rtxamd.Server(renderer, slotPaths=1 << 14, slots=4)
    ids = [sv.submit(ds, scenes.params(hs, 32 * (1 + k % 3), 32, 1 + k % 4, maxDepth=6, seed=k)) for k in range(12)]
    for jid in ids:
        sv.wait(jid, timeout=60)
    st = sv.stats()
    total = sum(32 * (1 + k % 3) * 32 * (1 + k % 4) for k in range(12))
    log(f"12 jobs: p50 {st.latencyP50Ms:.2f} ms, p99 {st.latencyP99Ms:.2f} ms, {st.msamplesPerSec:.2f} Msamples/s, "
        f"{st.bounces} bounces")
    assert st.jobsDone == 12 and st.pathsDone == total
    assert 0 < st.latencyP50Ms <= st.latencyP99Ms <= st.latencyMaxMs
    assert st.msamplesPerSec > 0 and st.extendRays >= total and st.bounces > 0
    # jobs still queued or running when the server is destroyed: no hang
    for k in range(3):
        sv.submit(ds, scenes.params(hs, 64, 64, 4, maxDepth=6, seed=100 + k))
    sv.close()
    ds.close()
