"""The oracle's guided sampling (shaders/guiding.glsl; getNewDirection,
shaders/raytrace.rgen:923-960) pinned by mathematics: every vMF lobe and
mixture integrates to one over the sphere, sampled directions follow the
mixture pdf (chi-square), parallax compensation aims a lobe at its target,
the guided estimator of a Lambert surface's albedo is unbiased, regions
resolve to the smallest containing box, and a guided render has the plain
render's mean."""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

import rtxamd
from rtxamd import abi
from oracle.binding import OracleScene, d3, lib
import scenes


def fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    phi = np.pi * (1 + 5 ** 0.5) * i
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


def three_lobes(parallax=False):
    lobes = [rtxamd.vmf_theta((0, 0, 1), 20.0), rtxamd.vmf_theta((1, 1, 0.2), 5.0),
             rtxamd.vmf_theta((-1, 0.2, -0.3), 0.0)]
    if parallax:
        lobes[0] = rtxamd.vmf_theta((0, 0, 1), 20.0, distance=2.0, target=(0.5, 0.0, 2.0))
    return rtxamd.vmm_theta(lobes, [0.5, 0.3, 0.2])


@pytest.mark.parametrize("kappa", [0.0, 0.5, 5.0, 50.0, 500.0])
def test_vmf_lobe_integrates_to_one(kappa):
    L = lib(abi)
    m = rtxamd.vmm_theta([rtxamd.vmf_theta((0.3, -0.2, 0.9), kappa)], [1.0])
    dirs = fib_sphere(400000)
    x = d3((0, 0, 0))
    vals = np.array([L.orc_vmm_pdf(C.byref(m), x, d3(w), 0) for w in dirs[::10]])
    integral = vals.mean() * 4 * np.pi
    assert abs(integral - 1.0) < (2e-3 if kappa < 100 else 2e-2), integral


def test_mixture_sampling_matches_pdf():
    L = lib(abi)
    m = three_lobes()
    st = C.c_uint32(12345)
    n = 60000
    out = (C.c_double * 3)()
    samples = np.empty((n, 3))
    for i in range(n):
        L.orc_sample_vmm(C.byref(m), d3((0, 0, 0)), 0, C.byref(st), out)
        samples[i] = out[:]
    assert np.allclose(np.linalg.norm(samples, axis=1), 1.0, atol=1e-9)
    # chi-square over 12 x 24 (cos theta, phi) cells of equal solid angle
    nt, nphi = 12, 24
    ct = np.clip(((samples[:, 2] + 1) / 2 * nt).astype(int), 0, nt - 1)
    ph = np.clip(((np.arctan2(samples[:, 1], samples[:, 0]) + np.pi) / (2 * np.pi) * nphi).astype(int), 0, nphi - 1)
    counts = np.bincount(ct * nphi + ph, minlength=nt * nphi)
    expected = np.zeros(nt * nphi)
    sub = 8
    for a in range(nt):
        for b in range(nphi):
            z = -1 + 2 * (a + (np.arange(sub) + 0.5) / sub) / nt
            p = -np.pi + 2 * np.pi * (b + (np.arange(sub) + 0.5) / sub) / nphi
            zz, pp = np.meshgrid(z, p)
            r = np.sqrt(1 - zz ** 2)
            vals = [L.orc_vmm_pdf(C.byref(m), d3((0, 0, 0)), d3(w), 0)
                    for w in np.stack([r * np.cos(pp), r * np.sin(pp), zz], -1).reshape(-1, 3)]
            expected[a * nphi + b] = np.mean(vals) * (4 * np.pi / (nt * nphi)) * n
    keep = expected > 20
    chi2 = ((counts[keep] - expected[keep]) ** 2 / expected[keep]).sum()
    p_value = stats.chi2.sf(chi2, keep.sum() - 1)
    assert p_value > 1e-3, (chi2, keep.sum())


def test_parallax_aims_lobe_at_target():
    L = lib(abi)
    m = rtxamd.vmm_theta([rtxamd.vmf_theta((0, 0, 1), 20.0, distance=2.0, target=(0.5, 0.0, 2.0))], [1.0])
    x = np.array([0.5, 0.0, 0.0])  # the target (0.5, 0, 2) is straight up from here
    up = L.orc_vmm_pdf(C.byref(m), d3(x), d3((0, 0, 1)), 1)
    assert up == pytest.approx(L.orc_vmm_pdf(C.byref(m), d3(x), d3((0, 0, 1)), 0), rel=1e-12)
    # from elsewhere the lobe's peak turns to the target
    p = np.array([-1.0, 0.0, 0.0])
    to_target = (np.array([0.5, 0.0, 2.0]) - p) / np.linalg.norm(np.array([0.5, 0.0, 2.0]) - p)
    assert L.orc_vmm_pdf(C.byref(m), d3(p), d3(to_target), 1) == pytest.approx(up, rel=1e-9)
    assert L.orc_vmm_pdf(C.byref(m), d3(p), d3((0, 0, 1)), 1) < 0.5 * up
    # without a distance (infinitely far) parallax changes nothing
    far = rtxamd.vmm_theta([rtxamd.vmf_theta((0, 0, 1), 20.0)], [1.0])
    assert L.orc_vmm_pdf(C.byref(far), d3(p), d3((0, 0, 1)), 1) == pytest.approx(up, rel=1e-12)


def guided_scene(prob=0.5):
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    orc = OracleScene(abi, hs.desc())
    # 2 x 2 x 1 regions over the box, lobes toward the ceiling light
    regions, vmms = [], []
    for ix in range(2):
        for iy in range(2):
            regions.append([[-1.01 + ix * 1.01, -1.01 + iy * 1.01, -0.01], [ix * 1.01, iy * 1.01, 2.01]])
            vmms.append(rtxamd.vmm_theta([rtxamd.vmf_theta((0, 0, 1), 8.0), rtxamd.vmf_theta((1, 0, 0), 0.0)],
                                         [0.7, 0.3]))
    orc.guiding_set(regions, vmms)
    return hs, orc, regions, vmms


def test_guided_render_is_unbiased():
    hs, orc, _, _ = guided_scene()
    w, h, spp = 24, 24, 48
    base = dict(maxDepth=4, enableMIS=1)
    fp_plain = scenes.params(hs, w, h, spp, **base)
    fp_guided = scenes.params(hs, w, h, spp, useGuiding=1, guidingProb=0.5, **base)
    ids = np.arange(w * h * spp, dtype=np.uint64)
    a, _ = orc.render_paths(fp_plain, ids)
    fp_guided.seed = 77
    b, _ = orc.render_paths(fp_guided, ids)
    la, lb = a.mean(1), b.mean(1)
    se = np.sqrt(la.var() / la.size + lb.var() / lb.size)
    assert abs(la.mean() - lb.mean()) < 4 * se, (la.mean(), lb.mean(), se)


def test_guided_render_uses_the_mixture():
    # with a mixture concentrated downwards every guided sample from the floor is
    # rejected (below the surface): paths end at the first floor hit, unlike plain ones
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    orc = OracleScene(abi, hs.desc())
    down = rtxamd.vmm_theta([rtxamd.vmf_theta((0, 0, -1), 2000.0)], [1.0])
    orc.guiding_set([[[-2, -2, -1], [2, 2, 3]]], [down])
    fp = scenes.params(hs, 16, 16, 4, maxDepth=6, useGuiding=1, guidingProb=1.0)
    rgb, sig = orc.render_paths(fp, np.arange(16 * 16 * 4, dtype=np.uint64))
    plain, psig = orc.render_paths(scenes.params(hs, 16, 16, 4, maxDepth=6), np.arange(16 * 16 * 4, dtype=np.uint64))
    assert rgb.mean() < plain.mean()
    # outside every region the path ends at the first non-discrete vertex
    orc.guiding_set([[[5, 5, 5], [6, 6, 6]]], [down])
    rgb2, _ = orc.render_paths(fp, np.arange(64, dtype=np.uint64))
    assert np.all(np.isfinite(rgb2))


def test_guiding_without_regions_is_an_error():
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    orc = OracleScene(abi, hs.desc())
    with pytest.raises(RuntimeError):
        orc.render_paths(scenes.params(hs, 4, 4, 1, useGuiding=1), np.arange(4, dtype=np.uint64))


# ------------------------------------------------------------- sample record
def record_scene():
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    orc = OracleScene(abi, hs.desc())
    regions = [[[-1.01, -1.01, -0.01], [0.0, 1.01, 2.01]], [[0.0, -1.01, -0.01], [1.01, 1.01, 2.01]]]
    vmms = [rtxamd.vmm_theta([rtxamd.vmf_theta((0, 0, 1), 3.0)], [1.0])] * 2
    orc.guiding_set(regions, vmms)
    return hs, orc, regions


def wall_albedo(p):
    # the Cornell generator's walls: red at x = -1, green at x = +1, white elsewhere
    if abs(p[0] + 1.0) < 1e-4:
        return np.array([0.63, 0.065, 0.05])
    if abs(p[0] - 1.0) < 1e-4:
        return np.array([0.14, 0.45, 0.09])
    return np.array([0.73, 0.73, 0.73])


def test_record_weights_carry_the_path_radiance():
    """NEE off, one bounce: every light reaching the camera path after vertex 0
    passes through slot 0, so lightSums[0] = result / albedo_0 (the Lambert
    throughput change f cos / pdf is the albedo) and weight = |.| / pdf."""
    hs, orc, regions = record_scene()
    n = 64 * 64 * 8
    fp = scenes.params(hs, 64, 64, 8, maxDepth=1, maxFollowDiscrete=0, enableNEE=0, updateGuiding=1)
    rgb, sig, dd = orc.render_paths_record(fp, np.arange(n, dtype=np.uint64))
    checked = 0
    for i in range(n):
        s0 = dd[i, 0]
        if s0["flags"] == abi.DIR_INVALID or rgb[i].max() <= 0:
            continue
        p = s0["position"].astype(np.float64)
        if abs(abs(p[0]) - 1.0) > 1e-4 and abs(p[2]) > 1e-4 and abs(p[1] - 1.0) > 1e-4:
            continue  # a box face or the ceiling: keep to the walls and floor
        want = np.linalg.norm(rgb[i] / wall_albedo(p)) / s0["pdf"]
        assert s0["weight"] == pytest.approx(want, rel=2e-5)
        checked += 1
    assert checked > 100


def test_record_distances_regions_and_invalid_slots():
    hs, orc, regions = record_scene()
    n = 32 * 32 * 2
    fp = scenes.params(hs, 32, 32, 2, maxDepth=6, enableMIS=1, updateGuiding=1)
    rgb, sig, dd = orc.render_paths_record(fp, np.arange(n, dtype=np.uint64))
    valid = dd["flags"] != abi.DIR_INVALID
    assert valid.any(axis=1).mean() > 0.5
    # the region of a record is the region holding its position (smallest index on the shared face)
    pos = dd["position"][valid]
    want = np.where(pos[:, 0] <= 0.0, 0, 1)
    assert np.array_equal(dd["flags"][valid], want)
    # directions are unit vectors in the upper hemisphere of where they start; pdfs positive
    d = dd["direction"][valid]
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    assert np.all(dd["pdf"][valid] > 0) and np.all(dd["weight"][valid] > 0)
    # consecutive Lambert vertices: a record's distance is the segment to the next vertex
    ok = 0
    for i in range(n):
        for k in range(15):
            a, b = dd[i, k], dd[i, k + 1]
            if a["flags"] == abi.DIR_INVALID or b["flags"] == abi.DIR_INVALID or a["distance"] <= 0:
                continue
            seg = np.linalg.norm(b["position"].astype(np.float64) - a["position"])
            assert a["distance"] == pytest.approx(seg, rel=1e-4, abs=1e-5)
            # and the next vertex lies along the recorded direction
            assert np.allclose(a["position"] + a["distance"] * a["direction"], b["position"], atol=1e-4)
            ok += 1
    assert ok > 100
