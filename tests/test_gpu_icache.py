"""The irradiance cache on the device (librtxamd.so through the C ABI)
against the oracle: lookups on a given cache state path by path, one
frame's upkeep (its jobs, their 200 stratum sub-paths each, and the records
they produce), determinism, a multi-frame render against the plain method,
and the error paths.  Bars as in test_gpu_parity: per path and sub-path the
device diverges no more than the FP32 instantiation of the reference logic
(+0.5%) and agrees within 1e-4 where it does not diverge; the accumulation
and commit of the device's own sub-path results must match the oracle's
restatement of them to FP32 rounding (rel 1e-4 of the record's scale)."""
import numpy as np
import pytest

import rtxamd
from rtxamd import abi
from oracle.binding import OracleScene, ic_apply
import scenes
from parity import check_paths, log, path_stats, relmse

pytestmark = pytest.mark.gpu

W = H = 64


def cornell_state(orc, hs, frames=4, max_entries=400):
    """A cache built by the oracle's own frames on the Cornell scene."""
    st = rtxamd.IcState.empty(max_entries)
    for f in range(frames):
        fp = scenes.params(hs, W, H, 1, maxDepth=8, enableMIS=1, useIrradianceCache=1,
                           irradianceCreateProb=0.02, irradianceUpdateProb=0.005, sample_offset=100 + f)
        orc.ic_frame(fp, st, prec=64)
    return st


def ic_params(hs, create, update, sample=0, flags=abi.RTX_FRAME_KEEP_PATHS | abi.RTX_FRAME_RESET_ACCUM, **kw):
    return scenes.params(hs, W, H, 1, maxDepth=8, enableMIS=1, useIrradianceCache=1, irradianceCreateProb=create,
                         irradianceUpdateProb=update, sample_offset=sample, flags=flags, **kw)


@pytest.fixture(scope="module")
def cornell():
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    orc = OracleScene(abi, hs.desc())
    st = cornell_state(orc, hs)
    log(f"oracle-built cache: {st.count} records")
    return hs, orc, st


@pytest.mark.parametrize("opts", [{}, {"useIrradianceGradients": 1, "irradianceCachePerformVisibilityCheck": 1}])
def test_lookup_paths_match_oracle(renderer, cornell, opts):
    hs, orc, st0 = cornell
    renderer.ic_create(st0.max_entries)
    renderer.ic_write(st0)
    ds = renderer.upload(hs)
    fp = ic_params(hs, 0.0, 0.0, sample=7, **opts)
    renderer.render(ds, fp)
    rgb, sig = renderer.read_paths(W * H)
    o64, o32 = st0.copy(), st0.copy()
    orgb, osig, jobs, _, _ = orc.ic_frame(fp, o64, prec=64)
    frgb, fsig, _, _, _ = orc.ic_frame(fp, o32, prec=32)
    assert len(jobs) == 0
    check_paths(f"cache lookups {opts}", rgb, sig, orgb, osig, (frgb, fsig), abs_same=0.97, abs_within=0.99)
    # the cache is unchanged by a frame without upkeep
    after = renderer.ic_read()
    assert after.count == st0.count and np.array_equal(after.entries, st0.entries)
    ds.close()


def test_frame_upkeep_matches_oracle(renderer, cornell):
    hs, orc, st0 = cornell
    renderer.ic_create(st0.max_entries)
    renderer.ic_write(st0)
    ds = renderer.upload(hs)
    fp = ic_params(hs, 0.02, 0.01, sample=9)
    renderer.render(ds, fp)
    rgb, sig = renderer.read_paths(W * H)
    gjobs, gsub, gsig = renderer.ic_read_jobs()
    gstate = renderer.ic_read()
    ostate = st0.copy()
    orgb, osig, ojobs, osub, osubsig = orc.ic_frame(fp, ostate, prec=64)
    same, _ = check_paths("cache frame paths", rgb, sig, orgb, osig, abs_same=0.97, abs_within=0.99)
    nu = sum(1 for j in gjobs if j[0] == 0)
    log(f"jobs: device {len(gjobs)} ({nu} updates), oracle {len(ojobs)}; records {gstate.count} vs {ostate.count}")
    assert len(gjobs) > 10
    # the job lists agree where the paths that made them agree
    n = min(len(gjobs), len(ojobs))
    match = np.array([gjobs[i][0] == ojobs[i][0] and gjobs[i][1] == ojobs[i][1] and
                      np.allclose(gjobs[i][2], ojobs[i][2], atol=1e-4) for i in range(n)])
    log(f"  matching jobs {match.mean():.4f}")
    assert abs(len(gjobs) - len(ojobs)) <= max(2, len(ojobs) // 20) and match.mean() >= 0.9
    # sub-paths of matching jobs, stratum by stratum
    idx = np.flatnonzero(match)
    g = gsub[idx].reshape(-1, 4).astype(np.float64)
    o = osub[idx].reshape(-1, 4)
    check_paths("stratum sub-paths", g[:, :3], gsig[idx].reshape(-1, 2), o[:, :3], osubsig[idx].reshape(-1, 2),
                abs_same=0.95, abs_within=0.99)
    hit = (o[:, 3] < 1e6) & (gsig[idx].reshape(-1, 2) == osubsig[idx].reshape(-1, 2)).all(1)
    assert np.allclose(g[hit, 3], o[hit, 3], rtol=2e-5, atol=1e-6)
    # accumulation + commit of the device's own sub-path results
    ref = st0.copy()
    ic_apply(abi, fp, ref, gjobs, gsub, prec=32)
    assert ref.count == gstate.count and ref.hdr.nextUpdateSlot == gstate.hdr.nextUpdateSlot
    c = gstate.count
    assert np.array_equal(ref.entries["numUpdates"][:c], gstate.entries["numUpdates"][:c])
    for f in ("color", "rotGrad", "transGrad"):
        a, b = gstate.entries[f][:c].astype(np.float64), ref.entries[f][:c].astype(np.float64)
        scale = np.maximum(np.abs(b).max(1, keepdims=True), 1e-6)
        err = np.abs(a - b) / scale
        log(f"  {f}: max rel err {err.max():.2e}")
        assert err.max() <= 1e-4, f
    assert np.allclose(gstate.entries["harmonicR"][:c], ref.entries["harmonicR"][:c], rtol=1e-5)
    assert np.allclose(gstate.spheres["center"][:c], ref.spheres["center"][:c], rtol=0, atol=0)
    assert np.allclose(gstate.spheres["radius"][:c], ref.spheres["radius"][:c], rtol=1e-5)
    ds.close()


def test_cache_render_is_deterministic(renderer):
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    ds = renderer.upload(hs)
    out = []
    for _ in range(2):
        renderer.ic_create(500)
        for f in range(3):
            renderer.render(ds, ic_params(hs, 0.01, 0.002, sample=f, flags=abi.RTX_FRAME_RESET_ACCUM,
                                          isIrradiancePrepareFrame=1))
        renderer.render(ds, ic_params(hs, 0.01, 0.002, sample=3, flags=abi.RTX_FRAME_RESET_ACCUM))
        st = renderer.ic_read()
        out.append((renderer.read_image(W, H), st.entries.copy(), st.spheres.copy(), st.count))
    assert out[0][3] == out[1][3] and out[0][3] > 0
    for a, b in zip(out[0][:3], out[1][:3]):
        assert np.array_equal(a, b)
    ds.close()


def test_cache_image_close_to_plain(renderer):
    """Prepare frames, then 64 spp with the cache against 64 spp plain: the
    cache's bias stays small and paths get shorter."""
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    ds = renderer.upload(hs)
    w, h = 128, 128
    base = dict(maxDepth=8, enableMIS=1)
    plain = renderer.render(ds, scenes.params(hs, w, h, 256, seed=5, **base))
    ref = renderer.read_image(w, h)[..., :3]
    renderer.ic_create(10000)
    icp = dict(useIrradianceCache=1, irradianceCreateProb=0.0005, irradianceUpdateProb=0.00001, **base)
    renderer.render(ds, scenes.params(hs, w, h, 30, isIrradiancePrepareFrame=1, **icp))
    n = renderer.ic_read().count
    cst = renderer.render(ds, scenes.params(hs, w, h, 64, seed=9, **icp))
    img = renderer.read_image(w, h)[..., :3]
    bias = img.mean() / ref.mean() - 1.0
    e = relmse(img, ref)
    log(f"cache: {n} records after prepare frames; mean {img.mean():.4f} vs plain {ref.mean():.4f} "
        f"(bias {bias:+.3%}), relmse {e:.3e}; extend rays/path {cst.extendRays / cst.paths:.2f} vs "
        f"{plain.extendRays / plain.paths:.2f}")
    assert n > 100
    assert abs(bias) < 0.05
    assert e < 0.05
    assert cst.extendRays / cst.paths < plain.extendRays / plain.paths
    ds.close()


def test_cache_errors(renderer):
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    ds = renderer.upload(hs)
    r2 = rtxamd.Renderer(0)
    with pytest.raises(rtxamd.RtxError) as e:
        r2.render(ds, ic_params(hs, 0.0, 0.0))
    assert e.value.code == abi.RTX_E_STATE
    r2.close()
    renderer.ic_create(16)
    bad = rtxamd.IcState.empty(8)
    with pytest.raises(rtxamd.RtxError) as e:
        renderer.ic_write(bad)
    assert e.value.code == abi.RTX_E_INVALID
    with pytest.raises(rtxamd.RtxError):
        renderer.render(ds, ic_params(hs, 0.0, 0.0, irradianceA=0.0))
    ds.close()
