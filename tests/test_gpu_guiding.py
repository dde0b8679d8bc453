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
