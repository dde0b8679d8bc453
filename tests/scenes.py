"""Scenes built through the host scene API for tests (helpers, not tests)."""
import numpy as np

import rtxamd
from rtxamd import abi


def quad(p0, p1, p2, p3):
    return np.array([p0, p1, p2, p3], np.float32), np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def box(lo, hi):
    lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    c = lambda x, y, z: [(lo, hi)[x][0], (lo, hi)[y][1], (lo, hi)[z][2]]
    faces = [
        [c(0, 0, 0), c(0, 0, 1), c(0, 1, 1), c(0, 1, 0)], [c(1, 0, 0), c(1, 1, 0), c(1, 1, 1), c(1, 0, 1)],
        [c(0, 0, 0), c(1, 0, 0), c(1, 0, 1), c(0, 0, 1)], [c(0, 1, 0), c(0, 1, 1), c(1, 1, 1), c(1, 1, 0)],
        [c(0, 0, 0), c(0, 1, 0), c(1, 1, 0), c(1, 0, 0)], [c(0, 0, 1), c(1, 0, 1), c(1, 1, 1), c(0, 1, 1)],
    ]
    pos = np.array([p for f in faces for p in f], np.float32)
    idx = np.array([[4 * i, 4 * i + 1, 4 * i + 2, 4 * i, 4 * i + 2, 4 * i + 3] for i in range(6)], np.uint32)
    return pos, idx.reshape(-1, 3)


def material_zoo(point_light=True, distribution=0):
    """Cornell-like box exercising every material type and light type:
    Lambert, Phong, smooth conductor, mirror, rough conductor, dielectric and
    light materials; area, sphere and point lights (area lights first).
    Point lights are reachable by NEE only, so estimator-equivalence tests
    leave them out (point_light=False).  distribution: the rough conductors'
    microfacet distribution (abi.DISTRIBUTION_BECKMANN or abi.DISTRIBUTION_GGX)."""
    s = rtxamd.HostScene.new()
    white = s.add_material(rtxamd.diffuse((0.7, 0.7, 0.7)))
    red = s.add_material(rtxamd.diffuse((0.6, 0.1, 0.1)))
    ph = s.add_material(rtxamd.phong((0.3, 0.4, 0.5), (0.4, 0.4, 0.3), 30.0))
    ph_sharp = s.add_material(rtxamd.phong((0.1, 0.1, 0.1), (0.8, 0.8, 0.8), 300.0))
    cond = s.add_material(rtxamd.conductor(0.2, 3.0))
    mir = s.add_material(rtxamd.mirror())
    rough = s.add_material(rtxamd.rough_conductor(0.25, 0.37, 2.82, distribution))
    rough_glossy = s.add_material(rtxamd.rough_conductor(0.08, 1.2, 5.0, distribution))
    glass = s.add_material(rtxamd.dielectric(1.5))
    lamp = s.add_material(rtxamd.light((8.0, 8.0, 8.0), (0.5, 0.5, 0.5)))
    ball = s.add_material(rtxamd.light((20.0, 10.0, 5.0)))
    # area light first so that its light index owns face-table row 0
    p, i = quad((-0.3, -0.3, 1.98), (-0.3, 0.3, 1.98), (0.3, 0.3, 1.98), (0.3, -0.3, 1.98))
    s.add_mesh(p, i, lamp)
    walls = [
        (quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)), white),
        (quad((-1, -1, 2), (-1, 1, 2), (1, 1, 2), (1, -1, 2)), white),
        (quad((-1, 1, 0), (1, 1, 0), (1, 1, 2), (-1, 1, 2)), ph),
        (quad((-1, -1, 0), (-1, 1, 0), (-1, 1, 2), (-1, -1, 2)), red),
        (quad((1, -1, 0), (1, -1, 2), (1, 1, 2), (1, 1, 0)), mir),
    ]
    for (p, i), m in walls:
        s.add_mesh(p, i, m)
    for lo, hi, m in [((0.05, -0.6, 0.0), (0.55, -0.1, 0.5), rough), ((-0.7, 0.0, 0.0), (-0.2, 0.5, 1.1), cond),
                      ((-0.2, -0.7, 0.0), (0.0, -0.5, 0.3), ph_sharp), ((0.5, 0.3, 0.0), (0.8, 0.6, 0.4), rough_glossy)]:
        p, i = box(lo, hi)
        s.add_mesh(p, i, m)
    s.add_sphere((0.3, -0.35, 0.75), 0.25, glass)
    s.add_sphere((-0.5, -0.5, 1.5), 0.08, ball)
    if point_light:
        s.add_point_light((0.6, -0.6, 1.6), (1.5, 1.5, 1.5))
    s.set_view((0.0, -3.8, 1.0), (0.0, 0.0, 1.0), (0.0, 0.0, 1.0), 40.0)
    return s.finalize()


def spot_room():
    """The closed room lit only through a spot on its ceiling (a lamp on a
    black box under it; host generator RTXH_SCENE_SPOT): BSDF sampling finds
    the bright spot rarely, a guide that learns its direction lowers the
    variance (the classic case for path guiding)."""
    return rtxamd.HostScene.generate(abi.SCENE_SPOT)


def emitter_over_plane(side=1.0, height=1.0, radiance=5.0, albedo=0.5):
    """A large diffuse floor at z=0 under a square area light of side `side`
    centred at height `height`, facing down; the camera looks straight down
    at the origin with a very narrow field of view."""
    s = rtxamd.HostScene.new()
    lamp = s.add_material(rtxamd.light((radiance,) * 3))
    floor = s.add_material(rtxamd.diffuse((albedo,) * 3))
    h = side / 2
    p, i = quad((-h, -h, height), (-h, h, height), (h, h, height), (h, -h, height))
    s.add_mesh(p, i, lamp)
    p, i = quad((-50, -50, 0), (50, -50, 0), (50, 50, 0), (-50, 50, 0))
    s.add_mesh(p, i, floor)
    s.set_view((0.0, 0.0, 0.5 * height), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.01)
    return s.finalize()


def form_factor_parallel_square(side, height):
    """Differential-area-to-parallel-square form factor, square centred above
    the point: four corner rectangles a x b = side/2 x side/2 (Howell's
    catalogue of configuration factors, case B-3)."""
    X = Y = (side / 2) / height
    f = (X / np.sqrt(1 + X * X) * np.arctan(Y / np.sqrt(1 + X * X)) +
         Y / np.sqrt(1 + Y * Y) * np.arctan(X / np.sqrt(1 + Y * Y))) / (2 * np.pi)
    return 4 * f


def params(hs, w, h, spp, **kw):
    cam = hs.camera(w, h)
    flags = kw.pop("flags", abi.RTX_FRAME_RESET_ACCUM)
    return rtxamd.frame_params(cam, w, h, spp, flags=flags, **kw)


def ceiling_over_point_light(height=1.0, albedo=0.6, intensity=2.0, eps=1e-4):
    """A black floor at z=0, a Lambert ceiling of albedo `albedo` at z=height
    (100 heights wide) and one point light of intensity `intensity` just above
    the floor's origin.  The ceiling's radiance seen from the origin is
    (albedo I / (pi h^2)) cos^3(theta), so the origin's irradiance is
    2 albedo I / (5 h^2) and the harmonic mean distance to the ceiling over
    cosine-weighted directions is 1.5 h."""
    s = rtxamd.HostScene.new()
    black = s.add_material(rtxamd.diffuse((0.0, 0.0, 0.0)))
    ceil = s.add_material(rtxamd.diffuse((albedo,) * 3))
    w = 100.0 * height
    p, i = quad((-w, -w, height), (-w, w, height), (w, w, height), (w, -w, height))
    s.add_mesh(p, i, ceil)
    p, i = quad((-w, -w, 0), (w, -w, 0), (w, w, 0), (-w, w, 0))
    s.add_mesh(p, i, black)
    s.add_point_light((0.0, 0.0, eps), (intensity,) * 3)
    s.set_view((0.0, -2.0, 0.5 * height), (0.0, 0.0, 0.5 * height), (0.0, 0.0, 1.0), 60.0)
    return s.finalize()


def cache_state(records, max_entries, irradiance_a=0.2):
    """An IcState holding `records` = [(center, normal, color, harmonicR), ...]
    with the reference's radius a * harmonicR and numUpdates 1."""
    st = rtxamd.IcState.empty(max_entries)
    for i, (c, n, col, hr) in enumerate(records):
        st.entries[i]["normal"] = n
        st.entries[i]["color"] = col
        st.entries[i]["harmonicR"] = hr
        st.entries[i]["numUpdates"] = 1
        st.spheres[i]["center"] = c
        st.spheres[i]["radius"] = np.float32(irradiance_a) * np.float32(hr)
    st.hdr.nextCacheSlot = len(records)
    return st


def constant_env(radiance, w=8, h=4):
    return np.full((h, w, 4), radiance, np.float32)


def furnace(albedo=0.6, radiance=1.5, env_light=True, point_light=False):
    """A diffuse sphere of albedo `albedo` (radius 1 at the origin) under a
    constant environment of radiance `radiance`.  A convex object never sees
    itself, so every pixel that hits it has expected value albedo *
    radiance and every other pixel exactly `radiance`."""
    s = rtxamd.HostScene.new()
    m = s.add_material(rtxamd.diffuse((albedo,) * 3))
    s.add_sphere((0.0, 0.0, 0.0), 1.0, m)
    if point_light:  # a second light: the environment light's sample probability halves
        s.add_point_light((0.0, 0.0, 50.0), (0.0, 0.0, 0.0))
    s.set_environment(constant_env(radiance), add_light=env_light)
    s.set_view((0.0, -4.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 40.0)
    return s.finalize()


def textured_plate(texture, diffuse=(1.0, 1.0, 1.0), radiance=2.0, alpha_black=False):
    """A 2 x 2 quad at z = 0 facing +z with texture coordinates spanning
    [0,1]^2 and a diffuse texture, under a constant environment; the camera
    looks straight down from z = 3.  With alpha_black the quad's diffuse
    colour is 0, so a pixel is `radiance` where the any-hit test lets the
    camera ray through and 0 where it does not."""
    s = rtxamd.HostScene.new()
    slot = s.add_texture(texture)
    m = rtxamd.diffuse((0.0, 0.0, 0.0) if alpha_black else diffuse)
    m.textureIdDiffuse = slot
    mi = s.add_material(m)
    p, i = quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0))
    s.add_mesh(p, i, mi, texcoords=[(0, 0), (1, 0), (1, 1), (0, 1)])
    s.set_environment(constant_env(radiance), add_light=True)
    s.set_view((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 30.0)
    return s.finalize()


def env_gallery():
    """A non-constant environment (smooth colour bands), a textured Phong
    box, a Lambert floor and an area light: every environment and texture
    path of the shaders (miss colour, environment NEE, MIS probe misses,
    textured diffuse and Phong) in one scene."""
    s = rtxamd.HostScene.new()
    lamp = s.add_material(rtxamd.light((6.0, 6.0, 6.0)))
    floor = s.add_material(rtxamd.diffuse((0.5, 0.5, 0.5)))
    y, x = np.mgrid[0:16, 0:32]
    tex = np.stack([(x * 8) % 256, (y * 16) % 256, ((x + y) * 5) % 256, np.full_like(x, 255)], -1).astype(np.uint8)
    slot = s.add_texture(tex)
    ph = rtxamd.phong((0.6, 0.6, 0.6), (0.3, 0.3, 0.3), 20.0)
    ph.textureIdDiffuse = slot
    phm = s.add_material(ph)
    dm = rtxamd.diffuse((0.9, 0.9, 0.9))
    dm.textureIdDiffuse = slot
    dmi = s.add_material(dm)
    p, i = quad((-0.3, -0.3, 2.5), (-0.3, 0.3, 2.5), (0.3, 0.3, 2.5), (0.3, -0.3, 2.5))
    s.add_mesh(p, i, lamp)
    p, i = quad((-3, -3, 0), (3, -3, 0), (3, 3, 0), (-3, 3, 0))
    s.add_mesh(p, i, floor)
    p, i = box((-0.5, -0.5, 0.0), (0.1, 0.1, 0.8))
    uv = np.tile(np.array([(0, 0), (1, 0), (1, 1), (0, 1)], np.float32), (6, 1))
    s.add_mesh(p, i, phm, texcoords=uv)
    p, i = quad((0.3, -0.8, 0.01), (1.3, -0.8, 0.01), (1.3, 0.2, 0.01), (0.3, 0.2, 0.01))
    s.add_mesh(p, i, dmi, texcoords=[(0, 0), (2, 0), (2, 2), (0, 2)])
    ey, ex = np.mgrid[0:8, 0:16]
    env = np.stack([0.2 + 0.1 * ex, 0.3 + 0.2 * ey, np.full_like(ex, 0.5, dtype=np.float64) + 0.05 * (ex % 3),
                    np.ones_like(ex, dtype=np.float64)], -1).astype(np.float32)
    s.set_environment(env, add_light=True)
    s.set_view((0.0, -4.0, 2.0), (0.0, 0.0, 0.4), (0.0, 0.0, 1.0), 45.0)
    return s.finalize()
