"""Textures and environment maps on the device (librtxamd.so through the C
ABI): path-by-path parity with the oracle on an environment-lit scene with
textured Lambert and Phong surfaces (miss colours, environment NEE, MIS
probe misses, texture look-ups), the closed-form furnace sphere and
textured plate, the any-hit alpha test's pass rate, and a scene file loaded
by the host (the reference's veachMIS) rendered against the oracle.  Bars as
in test_gpu_parity: the device diverges no more than the FP32 instantiation
of the reference logic (+0.5%) and agrees within 1e-4 of the FP64 oracle
where it does not diverge."""
from pathlib import Path

import numpy as np
import pytest

import rtxamd
import scenes
from oracle.binding import OracleScene
from parity import check_paths, log
from rtxamd import abi

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "scenes"
KEEP = abi.RTX_FRAME_KEEP_PATHS | abi.RTX_FRAME_RESET_ACCUM


def paths(renderer, hs, fp, n):
    ds = renderer.upload(hs)
    try:
        renderer.render(ds, fp)
        return renderer.read_paths(n)
    finally:
        ds.close()


def compare(renderer, name, hs, w, h, frames, abs_same, **kw):
    fp = scenes.params(hs, w, h, frames, flags=KEEP, **dict(dict(maxDepth=8, enableMIS=1), **kw))
    n = frames * w * h
    rgb, sig = paths(renderer, hs, fp, n)
    orc = OracleScene(abi, hs.desc())
    ids = np.arange(n, dtype=np.uint64)
    orgb, osig = orc.render_paths(fp, ids)
    frgb, fsig = orc.render_paths(fp, ids, prec=32)
    check_paths(name, rgb, sig, orgb, osig, (frgb, fsig), abs_same=abs_same, abs_within=0.99)
    return rgb


@pytest.mark.parametrize("power", [1, 0])
def test_environment_and_textures_match_oracle(renderer, power):
    compare(renderer, f"environment + textures (power {power})", scenes.env_gallery(), 48, 48, 2, 0.9,
            usePowerHeuristic=power)


def test_environment_without_nee_matches_oracle(renderer):
    compare(renderer, "environment, NEE off", scenes.env_gallery(), 48, 48, 2, 0.9, enableNEE=0)


def test_furnace_on_device(renderer):
    albedo, L = 0.6, 1.5
    hs = scenes.furnace(albedo, L, point_light=True)
    w = h = 48
    frames = 64
    fp = scenes.params(hs, w, h, frames, flags=KEEP, maxDepth=6, enableMIS=1)
    rgb, _ = paths(renderer, hs, fp, frames * w * h)
    per = rgb[:, 0].reshape(frames, w * h)
    background = np.isclose(per, L, atol=1e-5).all(0)
    on = ~np.isclose(per, L, atol=1e-5).any(0)
    m = per[:, on].mean()
    se = per[:, on].mean(1).std() / np.sqrt(frames)
    log(f"furnace: sphere {m:.5f} (expected {albedo * L}, se {se:.1e}), {background.sum()} background pixels")
    assert on.sum() > 500 and background.sum() > 500
    assert abs(m - albedo * L) < 4 * se + 1e-3
    compare(renderer, "furnace", hs, 24, 24, 2, 0.97)


def test_alpha_test_pass_rate_on_device(renderer):
    a8 = 64
    tex = np.full((4, 4, 4), 255, np.uint8)
    tex[..., 3] = a8
    L = 2.0
    hs = scenes.textured_plate(tex, radiance=L, alpha_black=True)
    w = h = 16
    frames = 256
    fp = scenes.params(hs, w, h, frames, flags=KEEP, maxDepth=4, enableMIS=1)
    rgb, _ = paths(renderer, hs, fp, frames * w * h)
    per = rgb[:, 0].reshape(frames, h, w)[:, 6:10, 6:10].reshape(-1)
    assert set(np.unique(np.round(per, 5))) <= {0.0, L}
    frac = (per > 0).mean()
    p = 1 - a8 / 255.0
    log(f"alpha: pass rate {frac:.4f} (expected {p:.4f})")
    assert abs(frac - p) < 4 * np.sqrt(p * (1 - p) / per.size)
    # decisions against the FP32 oracle, which runs the same FP32 hit arithmetic
    orc = OracleScene(abi, hs.desc())
    ids = np.arange(8 * w * h, dtype=np.uint64)
    frgb, _ = orc.render_paths(fp, ids, prec=32)
    agree = ((frgb[:, 0] > 0) == (rgb[:8 * w * h, 0] > 0)).mean()
    log(f"alpha: decisions equal to the FP32 oracle's for {agree:.4f} of paths")
    assert agree > 0.5


def test_loaded_scene_matches_oracle(renderer):
    """The reference's veachMIS scene file, loaded by the host layer."""
    hs = rtxamd.HostScene.load(GOLDEN / "veachMIS" / "veachMIS.xml")
    compare(renderer, "veachMIS.xml", hs, 64, 36, 2, 0.95, maxDepth=6)


def test_texture_errors(renderer):
    s = rtxamd.HostScene.new()
    m = rtxamd.diffuse((0.5,) * 3)
    m.textureIdDiffuse = 2
    s.add_material(m)
    s.add_texture(np.zeros((2, 2, 4), np.uint8))
    hs = s.finalize()
    with pytest.raises(rtxamd.RtxError):
        renderer.upload(hs)
