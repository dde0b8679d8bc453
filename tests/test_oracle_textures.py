"""Textures and environment maps in the oracle (row f3): the sampler rules
(Vulkan bilinear filtering with REPEAT addressing, sRGB decoding before
filtering) at hand-derived texel positions, the environment map's lat-long
mapping (raytrace.rmiss:15-30) at known directions, and closed-form scenes:
a diffuse sphere in a constant environment (a convex object never sees
itself, so it reflects albedo x radiance; this pins the environment light's
pdf, the MIS weights of its NEE and probe rays and the miss colour), a
textured plate, and the any-hit alpha test (raytrace.rahit:21-47) letting
camera rays through with probability 1 - alpha."""
import numpy as np
import pytest

import scenes
from oracle.binding import OracleScene
from rtxamd import abi


def srgb(c):
    c = c / 255.0
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def oracle(hs):
    return OracleScene(abi, hs.desc())


def test_texture_sampler_rules():
    # 2 x 2 texture: texel centres at u, v = 0.25 and 0.75
    tex = np.zeros((2, 2, 4), np.uint8)
    tex[0, 0] = (255, 0, 0, 0)
    tex[0, 1] = (0, 255, 0, 255)
    tex[1, 0] = (0, 0, 255, 51)
    tex[1, 1] = (128, 128, 128, 102)
    o = oracle(scenes.textured_plate(tex))
    lin = np.concatenate([srgb(tex[..., :3].astype(np.float64)), tex[..., 3:] / 255.0], -1)
    # at a texel centre: that texel
    for (u, v), (r, c) in [((0.25, 0.25), (0, 0)), ((0.75, 0.25), (0, 1)), ((0.25, 0.75), (1, 0)),
                           ((0.75, 0.75), (1, 1))]:
        assert np.allclose(o.tex_sample(1, u, v), lin[r, c], atol=1e-7)
    # halfway between centres: the mean (decoded first, then filtered)
    assert np.allclose(o.tex_sample(1, 0.5, 0.25), 0.5 * (lin[0, 0] + lin[0, 1]), atol=1e-7)
    # REPEAT: u = 0 sits halfway between the last and the first column, and
    # v = 1.25 wraps to the first row's centre
    assert np.allclose(o.tex_sample(1, 0.0, 0.25), 0.5 * (lin[0, 1] + lin[0, 0]), atol=1e-7)
    assert np.allclose(o.tex_sample(1, 0.75, 1.25), lin[0, 1], atol=1e-7)
    assert np.allclose(o.tex_sample(1, -0.25, -0.75), lin[0, 1], atol=1e-7)  # (0.75, 0.25)
    assert np.allclose(o.tex_sample(1, -0.25, 2.75), lin[1, 1], atol=1e-7)  # (0.75, 0.75)
    # a quarter of the way: 3:1 weights on both axes
    w = np.array([0.75 * 0.75, 0.25 * 0.75, 0.75 * 0.25, 0.25 * 0.25])
    exp = w[0] * lin[0, 0] + w[1] * lin[0, 1] + w[2] * lin[1, 0] + w[3] * lin[1, 1]
    assert np.allclose(o.tex_sample(1, 0.375, 0.375), exp, atol=1e-7)


def test_environment_lat_long_mapping():
    # 4 x 2 map, every texel distinct; u = atan(x, -z) / 2pi, v = acos(y) / pi
    env = np.zeros((2, 4, 4), np.float32)
    for r in range(2):
        for c in range(4):
            env[r, c, :3] = (c + 1, 10 * (r + 1), 100 * (r * 4 + c))
    hs = scenes.furnace()
    s = scenes.rtxamd.HostScene.new()
    s.add_material(scenes.rtxamd.diffuse((0.5,) * 3))
    s.set_environment(env, add_light=False)
    o = oracle(s.finalize())
    t = env[..., :3].astype(np.float64)
    # +x: u = 0.25 -> x = 0.5 (columns 0 and 1 halfway); y = 0: v = 0.5 -> y = 0.5 (rows 0 and 1)
    assert np.allclose(o.env_color((1, 0, 0)), 0.25 * (t[0, 0] + t[0, 1] + t[1, 0] + t[1, 1]), rtol=1e-6)
    # -z: u = 0 -> x = -0.5: the last column wraps in
    assert np.allclose(o.env_color((0, 0, -1)), 0.25 * (t[0, 3] + t[0, 0] + t[1, 3] + t[1, 0]), rtol=1e-6)
    # +z: u = atan(0, -1) / 2pi = 0.5 -> x = 1.5: columns 1 and 2
    assert np.allclose(o.env_color((0, 0, 1)), 0.25 * (t[0, 1] + t[0, 2] + t[1, 1] + t[1, 2]), rtol=1e-6)
    # straight up: v = 0 -> y = -0.5: the last row wraps in (REPEAT on v too);
    # -z is -0 there, so atan(0, -0) = pi and u = 0.5 (columns 1 and 2)
    up = o.env_color((0, 1, 0))
    assert np.allclose(up, 0.25 * (t[1, 1] + t[1, 2] + t[0, 1] + t[0, 2]), rtol=1e-6)
    # the direction is normalized first
    assert np.allclose(o.env_color((3, 0, 0)), o.env_color((1, 0, 0)))
    # no map: black
    o2 = oracle(hs)
    assert np.allclose(OracleScene(abi, scenes.material_zoo().desc()).env_color((1, 0, 0)), 0)
    assert np.allclose(o2.env_color((0, 1, 0)), 1.5)


def render_mean(hs, w, h, frames, prec=64, **kw):
    fp = scenes.params(hs, w, h, frames, **dict(dict(maxDepth=6, enableMIS=1), **kw))
    o = oracle(hs)
    rgb, _ = o.render_paths(fp, np.arange(frames * w * h, dtype=np.uint64), prec=prec)
    return rgb.reshape(frames, h * w, 3)


@pytest.mark.parametrize("mis,power", [(1, 1), (1, 0), (0, 0)])
@pytest.mark.parametrize("point_light", [False, True])
def test_furnace_sphere_reflects_albedo_times_radiance(mis, power, point_light):
    albedo, L = 0.6, 1.5
    hs = scenes.furnace(albedo, L, point_light=point_light)
    img = render_mean(hs, 24, 24, 48, enableMIS=mis, usePowerHeuristic=power)
    px = img.mean(0)[:, 0]
    per = img[..., 0]
    background = np.isclose(per, L, atol=1e-6).all(0)
    on = ~np.isclose(per, L, atol=1e-6).any(0)  # every jittered camera ray hit the sphere
    assert on.sum() > 150 and background.sum() > 100
    m = per[:, on].mean()
    se = per[:, on].mean(1).std() / np.sqrt(per.shape[0])
    assert abs(m - albedo * L) < 4 * se + 1e-3, (m, albedo * L, se)
    assert np.allclose(px[background], L, atol=1e-6)


def test_furnace_without_environment_light_uses_escaping_paths():
    """No environment light: NEE has nothing to sample, the continuation's
    miss adds the environment (raytrace.rgen:1025-1036 with NEE off)."""
    albedo, L = 0.6, 1.5
    hs = scenes.furnace(albedo, L, env_light=False)
    img = render_mean(hs, 16, 16, 48, enableNEE=0)
    per = img[..., 0]
    on = ~np.isclose(per, L, atol=1e-6).any(0)
    m = per[:, on].mean()
    se = per[:, on].mean(1).std() / np.sqrt(per.shape[0])
    assert abs(m - albedo * L) < 4 * se + 1e-3


def test_textured_plate_reflects_decoded_texture():
    c = np.array([200, 90, 30], np.float64)
    tex = np.zeros((4, 4, 4), np.uint8)
    tex[..., :3] = c
    tex[..., 3] = 255
    L = 2.0
    hs = scenes.textured_plate(tex, diffuse=(0.9, 0.8, 1.0), radiance=L)
    img = render_mean(hs, 16, 16, 32)
    # the plate fills the middle of the image; its expected radiance is kd * L
    centre = img[:, [(8 * 16 + 8), (7 * 16 + 7), (8 * 16 + 7)], :].reshape(-1, 3)
    kd = np.array([0.9, 0.8, 1.0]) * srgb(c)
    m = centre.mean(0)
    se = centre.std(0) / np.sqrt(len(centre))
    assert np.all(np.abs(m - kd * L) < 4 * se + 1e-3), (m, kd * L)


def test_alpha_test_lets_rays_through():
    a8 = 64
    tex = np.full((4, 4, 4), 255, np.uint8)
    tex[..., 3] = a8
    L = 2.0
    hs = scenes.textured_plate(tex, radiance=L, alpha_black=True)
    img = render_mean(hs, 8, 8, 256)
    centre = img[:, 27:29, 0].reshape(-1)
    # a camera ray through the black plate ends at 0 where the candidate is
    # kept (rnd <= alpha) and sees the environment where it is not
    assert set(np.unique(np.round(centre, 6))) <= {0.0, L}
    frac = (centre > 0).mean()
    p = 1 - a8 / 255.0
    assert abs(frac - p) < 4 * np.sqrt(p * (1 - p) / centre.size)
    # The seed takes uint(u * 1e8 + ...): a last-bit change of the hit's
    # barycentrics changes it, so FP32 and FP64 hits decide differently ray by
    # ray (as the reference's hardware barycentrics would); the rate agrees.
    img32 = render_mean(hs, 8, 8, 256, prec=32)
    frac32 = (img32[:, 27:29, 0] > 0).mean()
    assert abs(frac32 - p) < 4 * np.sqrt(p * (1 - p) / centre.size)


def test_texture_and_environment_errors():
    s = scenes.rtxamd.HostScene.new()
    m = scenes.rtxamd.diffuse((0.5,) * 3)
    m.textureIdDiffuse = 3  # no such slot
    s.add_material(m)
    with pytest.raises(RuntimeError, match="texture"):
        oracle(s.finalize())
