"""Host scene layer: procedural scenes, light tables, camera, EXR."""
import ctypes as C

import numpy as np
import pytest

import rtxamd
from rtxamd import abi
import scenes


def mesh_arrays(desc):
    """World-space triangles, per-triangle material index, instance index."""
    tris, mats, inst = [], [], []
    for i in range(desc.instanceCount):
        m = desc.models[desc.instances[i].modelIndex]
        v = np.ctypeslib.as_array(C.cast(m.vertices, C.POINTER(C.c_float)), shape=(m.vertexCount, 12))
        idx = np.ctypeslib.as_array(m.indices, shape=(m.indexCount,)).reshape(-1, 3)
        tris.append(v[:, 0:3][idx])
        mats.append(v[:, 10].view(np.int32)[idx[:, 0]])
        inst.append(np.full(len(idx), i))
    return np.concatenate(tris), np.concatenate(mats), np.concatenate(inst)


def materials(desc):
    return [desc.materials[i] for i in range(desc.materialCount)]


def test_cornell_scene_matches_the_workload():
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    d = hs.desc()
    tris, mats, _ = mesh_arrays(d)
    ms = materials(d)
    lightTris = [k for k in range(len(tris)) if ms[mats[k]].type == abi.MAT_LIGHT]
    assert len(tris) - len(lightTris) == 34 and len(lightTris) == 2  # 5 walls + 2 boxes, quad light
    assert d.lightCount == 1 and d.lights[0].type == abi.LIGHT_AREA
    assert list(ms[mats[lightTris[0]]].lightColor) == pytest.approx([17, 12, 4])
    albedos = {tuple(np.round(m.diffuse[:], 3)) for m in ms if m.type == abi.MAT_DIFFUSE}
    assert albedos == {(0.63, 0.065, 0.05), (0.14, 0.45, 0.09), (0.73, 0.73, 0.73)}
    assert d.sphereCount == 1
    sp = d.spheres[0]
    assert sp.radius == pytest.approx(0.3) and ms[sp.materialIndex].type == abi.MAT_DIELECTRIC
    assert ms[sp.materialIndex].refractionIndex == pytest.approx(1.5)
    # light area: the quad is 0.5 x 0.5
    assert d.lights[0].area == pytest.approx(0.25, rel=1e-6)


def test_room_scene_has_exact_triangle_count_and_material_mix():
    hs = rtxamd.HostScene.generate(abi.SCENE_ROOM, 1)
    d = hs.desc()
    tris, mats, inst = mesh_arrays(d)
    assert len(tris) == 262144
    ms = materials(d)
    assert d.lightCount == 4 and all(d.lights[i].type == abi.LIGHT_AREA for i in range(4))
    lo, hi = tris.reshape(-1, 3).min(0), tris.reshape(-1, 3).max(0)
    assert np.all(lo >= [-10 - 1e-4, -1e-4, -10 - 1e-4]) and np.all(hi <= [10 + 1e-4, 10 + 1e-4, 10 + 1e-4])
    # per-object material draws: 50% Lambert, 30% rough conductor, 20% dielectric
    obj = np.array([m.type for m in ms[1:-1]])
    frac = [np.mean(obj == t) for t in (abi.MAT_DIFFUSE, abi.MAT_ROUGH_CONDUCTOR, abi.MAT_DIELECTRIC)]
    assert frac == pytest.approx([0.5, 0.3, 0.2], abs=0.02)
    rough = [m.roughness for m in ms if m.type == abi.MAT_ROUGH_CONDUCTOR]
    assert 0.05 <= min(rough) and max(rough) <= 0.5
    alb = np.array([m.diffuse[:] for m in ms[1:-1] if m.type == abi.MAT_DIFFUSE])
    assert alb.min() >= 0.1 and alb.max() <= 0.8
    # exact counts for other sizes, and a clear error for impossible ones
    assert rtxamd.HostScene.generate(abi.SCENE_ROOM, 3, 10020).triangle_count() == 10020
    with pytest.raises(rtxamd.RtxError):
        rtxamd.HostScene.generate(abi.SCENE_ROOM, 3, 10021)


def generate_canonical_f32(raw):
    """libstdc++ generate_canonical<float, 24> on one mt19937 word:
    float(x) / 2^32, kept below 1."""
    r = np.float32(raw) / np.float32(2.0 ** 32)
    return np.where(r >= 1, np.nextafter(np.float32(1), np.float32(0)), r).astype(np.float32)


def test_light_tables_follow_weighted_sampler_with_default_mt19937():
    """WeightedSampler (src/WeightedSampler.cpp:8-33) draws from a
    default-seeded std::mt19937 (seed 5489); numpy's MT19937 with legacy
    seeding is an independent implementation of the same generator."""
    hs = scenes.material_zoo()
    d = hs.desc()
    nl = d.lightCount
    bg = np.random.MT19937()
    bg._legacy_seeding(5489)
    raw = bg.random_raw(10000).astype(np.uint32)
    u = generate_canonical_f32(raw) * np.float32(nl)
    # sum-walk of equal weights: index = number of prefix sums <= u ... as in sample()
    want = np.minimum((u >= np.float32(1)).astype(int) + (u >= np.float32(2)).astype(int), nl - 1) if nl == 3 else None
    got = np.ctypeslib.as_array(d.randomLightIndex, shape=(10000,))
    assert nl == 3
    assert np.array_equal(got, want)
    for i in range(nl):
        assert d.lights[i].sampleProb == pytest.approx(1 / 3)
    # face table of the (two-triangle, equal-area) area light: fresh generator, same stream
    ft = np.ctypeslib.as_array(C.cast(d.randomTriIndex, C.POINTER(C.c_int32)), shape=(10000, 3))
    u2 = generate_canonical_f32(raw) * np.float32(2 * 0.18)
    face = (u2 >= np.float32(0.18)).astype(int)
    assert np.array_equal(ft[:, 0], face)
    assert np.allclose(ft[:, 1].view(np.float32), 0.5)
    assert np.allclose(ft[:, 2].view(np.float32), 0.18, rtol=1e-6)
    assert d.lights[0].area == pytest.approx(0.36, rel=1e-6)
    # sphere light area 4 pi r^2 (src/SceneLoader.cpp:866-869)
    assert d.lights[1].area == pytest.approx(4 * np.pi * 0.08 ** 2, rel=1e-6)


def test_camera_matrices_are_inverse_pairs_and_look_at_target():
    hs = rtxamd.HostScene.generate(abi.SCENE_ROOM, 1, 10020)
    cam = hs.camera(1920, 1080)
    m = lambda a: np.array(a[:], np.float64).reshape(4, 4).T  # column-major storage
    V, P, Vi, Pi = m(cam.view), m(cam.proj), m(cam.viewInverse), m(cam.projInverse)
    assert np.allclose(V @ Vi, np.eye(4), atol=1e-5) and np.allclose(P @ Pi, np.eye(4), atol=1e-5)
    v = hs.view()
    o = np.array(v.origin[:])
    f = np.array(v.target[:]) - o
    f /= np.linalg.norm(f)
    assert np.allclose(Vi[:3, 3], o, atol=1e-5)
    assert np.allclose(-Vi[:3, 2], f, atol=1e-6)  # camera looks down -z
    assert P[1, 1] < 0  # Y flip of src/CameraController.cpp:104
    assert P[0, 0] * 1920 / 1080 == pytest.approx(-P[1, 1], rel=1e-6)


def test_exr_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.uniform(0, 10, (37, 53, 4)).astype(np.float32)
    p = tmp_path / "x.exr"
    rtxamd.write_exr(p, img)
    back = rtxamd.read_exr(p)
    assert back.shape == img.shape
    assert np.array_equal(back[..., :3], img[..., :3]) and np.all(back[..., 3] == 1)
    with pytest.raises(rtxamd.RtxError):
        rtxamd.read_exr(tmp_path / "missing.exr")


def test_exr_reader_rejects_malformed_files(tmp_path):
    """Truncated files, a negative attribute size, a block outside the data
    window and an uncompressed block of the wrong size are errors, never
    reads past the file (each of these read out of bounds before)."""
    img = np.arange(3 * 2 * 4, dtype=np.float32).reshape(2, 3, 4)
    good = tmp_path / "good.exr"
    rtxamd.write_exr(good, img)
    data = good.read_bytes()
    bad = tmp_path / "bad.exr"

    def rejects(b):
        bad.write_bytes(bytes(b))
        with pytest.raises(rtxamd.RtxError):
            rtxamd.read_exr(bad)

    for n in range(len(data)):
        rejects(data[:n])
    i = data.index(b"channels\0chlist\0") + len(b"channels\0chlist\0")
    rejects(data[:i] + (-1).to_bytes(4, "little", signed=True) + data[i + 4:])
    # two blocks of one line (line number, byte count, 3 pixels of R, G, B
    # floats) after a table of two 8-byte offsets
    table = len(data) - 2 * (8 + 3 * 3 * 4) - 2 * 8
    first = int.from_bytes(data[table:table + 8], "little")
    for y in (2, -1, 1 << 30):
        rejects(data[:first] + y.to_bytes(4, "little", signed=True) + data[first + 4:])
    rejects(data[:first + 4] + (35).to_bytes(4, "little") + data[first + 8:])
    rtxamd.read_exr(good)


def test_reference_exr_reads_if_mounted():
    """Mitsuba's ZIP/HALF ground-truth EXRs decode (CPU only, skipped when the
    reference is not mounted)."""
    from pathlib import Path
    p = Path("/root/reference/scenes/veachMIS/veachMIS.exr")
    if not p.exists():
        pytest.skip("reference not mounted")
    img = rtxamd.read_exr(p)
    assert img.shape == (720, 1280, 4)
    assert np.isfinite(img).all() and img[..., :3].mean() > 0.01


def test_scene_builder_rejects_bad_input():
    s = rtxamd.HostScene.new()
    m = s.add_material(rtxamd.diffuse((0.5, 0.5, 0.5)))
    p, i = scenes.quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0))
    with pytest.raises(rtxamd.RtxError):
        s.add_mesh(p, i, m + 5)
    with pytest.raises(rtxamd.RtxError):
        s.add_sphere((0, 0, 0), 1.0, 7)
