"""Guided sampling on the device (getNewDirection with the region mixtures,
shaders/raytrace.rgen:923-960, shaders/guiding.glsl) against the oracle path
by path, and unbiasedness of the guided estimator at image level."""
import numpy as np
import pytest

import rtxamd
from rtxamd import abi
from oracle.binding import OracleScene
import scenes
from parity import check_paths, log

pytestmark = pytest.mark.gpu


def regions_and_mixtures(parallax):
    regions, vmms = [], []
    rng = np.random.default_rng(4)
    for ix in range(2):
        for iy in range(2):
            for iz in range(2):
                lo = [-1.01 + ix * 1.01, -1.01 + iy * 1.01, -0.01 + iz * 1.01]
                hi = [ix * 1.01 - 0.0, iy * 1.01, iz * 1.01 + 1.0]
                regions.append([lo, hi])
                lobes = [rtxamd.vmf_theta((0, 0, 1), 12.0, distance=1.5 if parallax else -1.0,
                                          target=(0.0, 0.0, 1.98)),
                         rtxamd.vmf_theta(rng.normal(size=3), float(rng.uniform(1, 40))),
                         rtxamd.vmf_theta((0, 1, 0), 0.0)]
                vmms.append(rtxamd.vmm_theta(lobes, [0.5, 0.3, 0.2]))
    return regions, vmms


@pytest.mark.parametrize("parallax", [0, 1])
def test_guided_paths_match_oracle(renderer, parallax):
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    regions, vmms = regions_and_mixtures(parallax)
    renderer.guiding_set(regions, vmms)
    orc = OracleScene(abi, hs.desc())
    orc.guiding_set(regions, vmms)
    ds = renderer.upload(hs)
    w = h = 128
    fp = scenes.params(hs, w, h, 2, maxDepth=8, enableMIS=1, useGuiding=1, guidingProb=0.6,
                       useParallaxCompensation=parallax,
                       flags=abi.RTX_FRAME_KEEP_PATHS | abi.RTX_FRAME_RESET_ACCUM)
    renderer.render(ds, fp)
    rgb, sig = renderer.read_paths(w * h * 2)
    ids = np.arange(w * h * 2, dtype=np.uint64)
    orgb, osig = orc.render_paths(fp, ids)
    frgb, fsig = orc.render_paths(fp, ids, prec=32)
    check_paths(f"guided paths (parallax {parallax})", rgb, sig, orgb, osig, (frgb, fsig), abs_same=0.97,
                abs_within=0.99)
    renderer.guiding_set([], [])
    ds.close()


def test_guided_image_mean_equals_plain(renderer):
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    regions, vmms = regions_and_mixtures(1)
    renderer.guiding_set(regions, vmms)
    ds = renderer.upload(hs)
    w = h = 128
    base = dict(maxDepth=8, enableMIS=1, flags=abi.RTX_FRAME_RESET_ACCUM | abi.RTX_FRAME_KEEP_PATHS)
    n = w * h * 64
    renderer.render(ds, scenes.params(hs, w, h, 64, seed=3, **base))
    a, _ = renderer.read_paths(n)
    renderer.render(ds, scenes.params(hs, w, h, 64, seed=4, useGuiding=1, guidingProb=0.5,
                                      useParallaxCompensation=1, **base))
    b, _ = renderer.read_paths(n)
    la, lb = a.mean(1).astype(np.float64), b.mean(1).astype(np.float64)
    se = np.sqrt(la.var() / n + lb.var() / n)
    log(f"plain {la.mean():.5f} guided {lb.mean():.5f} (se {se:.1e})")
    assert abs(la.mean() - lb.mean()) < 4 * se
    renderer.guiding_set([], [])
    ds.close()


def test_guiding_errors(renderer):
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    ds = renderer.upload(hs)
    renderer.guiding_set([], [])
    with pytest.raises(rtxamd.RtxError) as e:
        renderer.render(ds, scenes.params(hs, 8, 8, 1, useGuiding=1))
    assert e.value.code == abi.RTX_E_STATE
    bad = rtxamd.vmm_theta([rtxamd.vmf_theta((0, 0, 1), 1.0)], [1.0])
    bad.usedDistributions = 17
    with pytest.raises(rtxamd.RtxError):
        renderer.guiding_set([[[0, 0, 0], [1, 1, 1]]], [bad])
    ds.close()


@pytest.mark.parametrize("guided", [0, 1])
def test_sample_records_match_oracle(renderer, guided):
    """updateGuiding: the DirectionalData slots of every pixel (position,
    direction, pdf, weight = |light sum| / pdf, distance, region) against the
    oracle's for the paths that take the same branches."""
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    regions, vmms = regions_and_mixtures(1)
    renderer.guiding_set(regions, vmms)
    orc = OracleScene(abi, hs.desc())
    orc.guiding_set(regions, vmms)
    ds = renderer.upload(hs)
    w = h = 96
    fp = scenes.params(hs, w, h, 1, maxDepth=8, enableMIS=1, updateGuiding=1, useGuiding=guided, guidingProb=0.5,
                       sample_offset=5, flags=abi.RTX_FRAME_KEEP_PATHS | abi.RTX_FRAME_RESET_ACCUM)
    renderer.render(ds, fp)
    rgb, sig = renderer.read_paths(w * h)
    dd = renderer.guiding_read_samples(w, h)
    ids = np.arange(w * h, dtype=np.uint64)
    orgb, osig, odd = orc.render_paths_record(fp, ids)
    same, _ = check_paths(f"recording paths (guided {guided})", rgb, sig, orgb, osig, abs_same=0.97,
                          abs_within=0.99)
    g, o = dd[same], odd[same]
    gv, ov = g["flags"] != abi.DIR_INVALID, o["flags"] != abi.DIR_INVALID
    assert np.array_equal(gv, ov)
    assert np.array_equal(g["flags"][gv], o["flags"][ov])
    log(f"  valid records {gv.sum()} in {same.sum()} paths")
    for f in ("position", "direction"):
        assert np.allclose(g[f][gv], o[f][ov], rtol=1e-4, atol=2e-5), f
    for f, tol in (("pdf", 1e-4), ("distance", 1e-4)):
        a, b = g[f][gv].astype(np.float64), o[f][ov].astype(np.float64)
        assert np.all(np.abs(a - b) <= tol * np.abs(b) + 1e-5), f
    a, b = g["weight"][gv].astype(np.float64), o["weight"][ov].astype(np.float64)
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-6)
    log(f"  weight rel err p50 {np.median(rel):.2e} p99 {np.percentile(rel, 99):.2e}")
    assert np.mean(rel <= 1e-4) >= 0.99
    renderer.guiding_set([], [])
    ds.close()
