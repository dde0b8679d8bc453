"""Scene files on the host (librtxhost.so): the OBJ/MTL reader, the Mitsuba
XML and JSON scene formats of the reference's SceneLoader
(src/SceneLoader.cpp:29-770) and the tinyobjloader 2.0 behaviour it relies
on.  The checks come from the files themselves (geometry, materials and
lights that the text fixes), from closed forms (polygon areas, rotation
matrices) and from the reference's own veachMIS scene, kept as a fixture."""
import math
from pathlib import Path

import numpy as np
import pytest

import rtxamd
from rtxamd import abi

GOLDEN = Path(__file__).resolve().parent / "golden" / "scenes"


def write(path, text):
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(text)
    return path


def models(hs):
    """Per model: positions (n,3), normals (n,3), uv (n,2), materials (n,), indices (m,3)."""
    d = hs.desc()
    out = []
    for i in range(d.modelCount):
        md = d.models[i]
        vs = [md.vertices[k] for k in range(md.vertexCount)]
        out.append(dict(pos=np.array([list(v.pos) for v in vs], np.float32).reshape(-1, 3),
                        nrm=np.array([list(v.normal) for v in vs], np.float32).reshape(-1, 3),
                        uv=np.array([list(v.texCoord) for v in vs], np.float32).reshape(-1, 2),
                        mat=np.array([v.materialIndex for v in vs], np.int32),
                        idx=np.array([md.indices[k] for k in range(md.indexCount)], np.int64).reshape(-1, 3)))
    return out


def mats(hs):
    d = hs.desc()
    return [d.materials[i] for i in range(d.materialCount)]


def lights(hs):
    d = hs.desc()
    return [d.lights[i] for i in range(d.lightCount)]


def xml_scene(tmp_path, body, name="scene.xml"):
    return write(tmp_path / name, f'<?xml version="1.0"?>\n<scene version="0.6.0">\n{body}\n</scene>\n')


QUAD_DIFFUSE = '<bsdf type="diffuse"><rgb name="reflectance" value="0.5"/></bsdf>'


def obj_shape(fname, inner=QUAD_DIFFUSE):
    return f'<shape type="obj"><string name="filename" value="{fname}"/>{inner}</shape>'


def load_obj(tmp_path, obj_text, name="m.obj"):
    write(tmp_path / name, obj_text)
    return rtxamd.HostScene.load(xml_scene(tmp_path, obj_shape(name)))


def tri_area_xy(p, t):
    a, b, c = p[t[0]], p[t[1]], p[t[2]]
    return 0.5 * ((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]))


# ------------------------------------------------------------------ OBJ

def test_obj_triangle_and_quad(tmp_path):
    hs = load_obj(tmp_path, "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    m = models(hs)[0]
    # a convex quad is clipped at its first corner: (0,1,2) then (0,2,3)
    assert m["idx"].tolist() == [[0, 1, 2], [0, 2, 3]]
    assert np.allclose(m["nrm"], [0, 0, 1])  # per-face normal when the file has none


@pytest.mark.parametrize("poly", [
    # concave "arrow": the first corner is reflex
    [(0, 0), (2, 1), (4, 0), (2, 4)],
    # L shape
    [(0, 0), (3, 0), (3, 1), (1, 1), (1, 3), (0, 3)],
    # star (10 corners, 5 reflex)
    [(math.cos(k * math.pi / 5) * (1 if k % 2 == 0 else 0.4), math.sin(k * math.pi / 5) * (1 if k % 2 == 0 else 0.4))
     for k in range(10)],
])
@pytest.mark.parametrize("plane", ["xy", "xz", "yz"])
def test_obj_ear_clipping_tiles_concave_polygons(tmp_path, poly, plane):
    """n-2 triangles, all of the polygon's orientation, areas summing to the
    shoelace area, each free of the other corners: a valid ear clipping in
    whichever axis pair the dominant normal picks."""
    def embed(x, y):
        return {"xy": (x, y, 0.5), "xz": (x, 0.5, y), "yz": (0.5, x, y)}[plane]
    lines = [f"v {' '.join(repr(c) for c in embed(x, y))}" for x, y in poly]
    lines.append("vn 0 0 1")  # a normal keeps corners from being merged by the face normal
    lines.append("f " + " ".join(f"{i + 1}//1" for i in range(len(poly))))
    hs = load_obj(tmp_path, "\n".join(lines) + "\n")
    m = models(hs)[0]
    p2 = np.array(poly, np.float64)
    pos = m["pos"].astype(np.float64)
    key = {"xy": (0, 1), "xz": (0, 2), "yz": (1, 2)}[plane]
    # map model vertices back to polygon corners
    corner = [int(np.argmin(np.abs(p2 - pos[v, list(key)]).sum(1))) for v in range(len(pos))]
    tris = [[corner[v] for v in t] for t in m["idx"]]
    assert len(tris) == len(poly) - 2
    shoelace = 0.5 * sum(p2[i, 0] * p2[(i + 1) % len(poly), 1] - p2[(i + 1) % len(poly), 0] * p2[i, 1]
                         for i in range(len(poly)))
    areas = [tri_area_xy(p2, t) for t in tris]
    assert all(a * shoelace > 0 for a in areas)
    assert abs(sum(areas) - shoelace) < 1e-5 * abs(shoelace)
    for t in tris:
        a, b, c = p2[t]
        for k in set(range(len(poly))) - set(t):
            q = p2[k]
            d1 = (b[0] - a[0]) * (q[1] - a[1]) - (b[1] - a[1]) * (q[0] - a[0])
            d2 = (c[0] - b[0]) * (q[1] - b[1]) - (c[1] - b[1]) * (q[0] - b[0])
            d3 = (a[0] - c[0]) * (q[1] - c[1]) - (a[1] - c[1]) * (q[0] - c[0])
            inside = (d1 > 1e-9 and d2 > 1e-9 and d3 > 1e-9) or (d1 < -1e-9 and d2 < -1e-9 and d3 < -1e-9)
            assert not inside


def test_obj_index_forms_and_relative_indices(tmp_path):
    obj = """
v 0 0 0
v 1 0 0
v 0 1 0
vt 0.25 0.75
vt 0.5 0.5
vt 1 0
vn 0 0 -1
f -3/1/1 -2/2/1 -1/3/1
f 1//1 2//1 3//1
f 1/3/1 2/2/1 3/1/1
"""
    write(tmp_path / "m.obj", obj)
    with pytest.raises(rtxamd.RtxError):  # texture coordinates exist, the 2nd face has none
        rtxamd.HostScene.load(xml_scene(tmp_path, obj_shape("m.obj")))
    hs = load_obj(tmp_path, obj.replace("f 1//1 2//1 3//1\n", ""), "n.obj")
    m = models(hs)[0]
    # texture v is flipped (1 - v), normals taken from the file
    assert m["idx"].shape == (2, 3)
    uv = m["uv"][m["idx"][0]]
    assert np.allclose(uv, [[0.25, 0.25], [0.5, 0.5], [1.0, 1.0]])
    assert np.allclose(m["nrm"], [0, 0, -1])
    uv2 = m["uv"][m["idx"][1]]
    assert np.allclose(uv2, [[1.0, 1.0], [0.5, 0.5], [0.25, 0.25]])


def test_obj_zero_index_is_an_error(tmp_path):
    write(tmp_path / "m.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n")
    with pytest.raises(rtxamd.RtxError, match="index"):
        rtxamd.HostScene.load(xml_scene(tmp_path, obj_shape("m.obj")))


def test_obj_out_of_range_index_is_an_error(tmp_path):
    write(tmp_path / "m.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n")
    with pytest.raises(rtxamd.RtxError):
        rtxamd.HostScene.load(xml_scene(tmp_path, obj_shape("m.obj")))


def test_obj_numbers(tmp_path):
    """tinyobjloader's number reader: signs, leading dots, exponents; text
    that is not a number reads as 0; an exponent too long for an int
    (undefined behaviour in tinyobjloader) reads as 0 or inf."""
    vals = ["1e-3", ".5", "-.25", "+3", "1.5e+2", "123.456", "nan", "-0", "2.5E-1", "7e-99999999999"]
    expect = [1e-3, 0.5, -0.25, 3.0, 150.0, 123.456, 0.0, 0.0, 0.25, 0.0]
    lines = [f"v {v} 0 0" for v in vals] + ["vn 0 0 1"] + [f"f {i + 1}//1 {i + 1}//1 {i + 1}//1" for i in range(len(vals))]
    m = models(load_obj(tmp_path, "\n".join(lines) + "\n"))[0]
    got = sorted(set(m["pos"][:, 0].tolist()))
    assert got == sorted(set(float(np.float32(e)) for e in expect))


def test_obj_vertex_dedup_and_degenerate_face(tmp_path):
    # two faces share an edge; a degenerate face has a NaN normal, and its
    # corners all index vertex 0 (the reference's map look-up of a NaN key)
    obj = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 2 2 2\nf 1 2 3\nf 1 3 4\nf 5 5 5\n"
    m = models(load_obj(tmp_path, obj))[0]
    assert m["idx"][:2].tolist() == [[0, 1, 2], [0, 2, 3]]
    assert len(m["pos"]) == 4 + 3
    assert m["idx"][2].tolist() == [0, 0, 0]
    assert np.isnan(m["nrm"][4:]).all()


def test_obj_crlf_comments_groups_and_lines(tmp_path):
    obj = "# comment\r\nv 0 0 0\r\nv 1 0 0\r\nv 0 1 0\r\no first\r\ng grp\r\nl 1 2\r\np 3\r\nf 1 2 3\r\n\r\n"
    m = models(load_obj(tmp_path, obj))[0]
    assert m["idx"].tolist() == [[0, 1, 2]]


# ------------------------------------------------------------------ MTL + JSON

MTL = """newmtl white
  Kd 0.3 0.3 0.3
  illum 1
newmtl light
  Kd 0.1 0.2 0.3
  Ke 10 20 30
  illum 11
newmtl glass
  Ks 0.8 0.8 0.8
  Ni 1.5
  illum 4
newmtl shiny
  Kd 0.1 0.1 0.1
  Ks 0.5 0.5 0.5
  Ns 20
  illum 2
newmtl textured
  map_Kd -bm 0.5 -s 1 1 1 tex dir\\check.png
  illum 1
newmtl mirror
  illum 3
"""


def make_png(path, w=4, h=3, mode="RGBA"):
    from PIL import Image
    rng = np.random.default_rng(7)
    a = rng.integers(0, 256, (h, w, 4 if mode == "RGBA" else 3), dtype=np.uint8)
    path.parent.mkdir(parents=True, exist_ok=True)
    Image.fromarray(a, mode).save(path)
    return a


def json_scene(tmp_path, scene, objs, mtl=MTL):
    root = tmp_path / "js"
    for name, text in objs.items():
        write(root / "models" / name, text)
    write(root / "materials" / "m.mtl", mtl)
    if "check.png" in mtl:
        make_png(root / "textures" / "tex dir" / "check.png")
    import json
    return write(root / "scene.json", json.dumps(scene))


def box_obj(mat_lines):
    """Two triangles per usemtl block."""
    s = "mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\nvt 0 0\n"
    for m in mat_lines:
        s += f"usemtl {m}\nf 1/1/1 2/1/1 3/1/1\nf 1/1/1 3/1/1 4/1/1\n"
    return s


def test_json_scene_materials_instances_lights_camera(tmp_path):
    scene = {
        "models": [{"a": "a.obj"}, {"b": "b.obj"}],
        "instances": [
            {"b": {"translate": [1, 2, 3], "rotate": [10, 20, 30], "scale": [2, 3, 4]}},
            {"a": {}},
            {"b": {"rotate": [0, 0, 90]}},
        ],
        "lights": [{"color": [1, 2, 3], "position": [4, 5, 6]}],
        "camera": {"origin": [0, -5, 1], "target": [0, 0, 1], "up": [0, 0, 1]},
    }
    p = json_scene(tmp_path, scene, {"a": "", "a.obj": box_obj(["white", "textured"]),
                                     "b.obj": box_obj(["light", "glass", "shiny", "nosuch", "mirror"])})
    hs = rtxamd.HostScene.load(p)
    d = hs.desc()
    ms = mats(hs)
    # one MTL read per model: 6 materials each, in file order
    assert len(ms) == 12
    types = [m.type for m in ms[:6]]
    assert types == [abi.MAT_DIFFUSE, abi.MAT_LIGHT, abi.MAT_DIELECTRIC, abi.MAT_PHONG, abi.MAT_DIFFUSE,
                     abi.MAT_SPECULAR]
    assert np.allclose(list(ms[1].lightColor), [10, 20, 30]) and np.allclose(list(ms[1].diffuse), [0.1, 0.2, 0.3])
    assert ms[2].refractionIndex == pytest.approx(1.5) and ms[2].refractionIndexInv == pytest.approx(1 / 1.5)
    assert ms[3].specularHighlight == 20 and np.allclose(list(ms[3].specular), [0.5] * 3)
    # map_Kd without its own Kd keeps diffuse 0: LoadMtl's "Kd seen" flag is
    # not reset per material, and "white" set it.  The texture is slot 1
    # (slot 0 is the environment's).
    assert list(ms[4].diffuse) == [0, 0, 0] and ms[4].textureIdDiffuse == 1
    assert d.textureCount == 2
    tex = hs.textures()
    assert tex[0] is None and tex[1].shape == (3, 4, 4)
    # model a uses materials 0 and 4; model b 6 + (1, 2, 3, -1, 5): an
    # unknown usemtl name is material -1, which the offset turns into the
    # material just before this file's (here model a's "mirror")
    mm = models(hs)
    assert sorted(set(mm[0]["mat"].tolist())) == [0, 4]
    per_face = [mm[1]["mat"][t[0]] for t in mm[1]["idx"]]
    assert per_face == [7, 7, 8, 8, 9, 9, 5, 5, 11, 11]
    # instances: b is emissive -> area lights for instances 0 and 2, then the point light
    ls = lights(hs)
    assert [(l.type, l.instanceIndex) for l in ls] == [(abi.LIGHT_AREA, 0), (abi.LIGHT_AREA, 2),
                                                      (abi.LIGHT_POINT, 0)]
    assert np.allclose(list(ls[2].pos), [4, 5, 6]) and np.allclose(list(ls[2].color), [1, 2, 3])
    inst = [d.instances[i] for i in range(d.instanceCount)]
    assert [i.modelIndex for i in inst] == [1, 0, 1]
    assert [i.iLight for i in inst] == [0, -1, 1]
    v = hs.view()
    assert list(v.origin) == [0, -5, 1] and list(v.target) == [0, 0, 1] and v.vfov == 28.0

    def rot(axis, deg):
        c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
        r = np.eye(4)
        i, j = {0: (1, 2), 1: (2, 0), 2: (0, 1)}[axis]
        r[i, i] = r[j, j] = c
        r[i, j], r[j, i] = -s, s
        return r
    T = np.eye(4)
    T[:3, 3] = [1, 2, 3]
    R = rot(0, 10) @ rot(1, 20) @ rot(2, 30)
    S = np.diag([2, 3, 4, 1.0])
    M = np.array(inst[0].transform, np.float64).reshape(4, 4).T  # column-major
    N = np.array(inst[0].normalTransform, np.float64).reshape(4, 4).T
    assert np.allclose(M, T @ R @ S, atol=1e-6)
    assert np.allclose(N, R @ np.diag([1 / 2, 1 / 3, 1 / 4, 1.0]), atol=1e-6)
    M2 = np.array(inst[2].transform, np.float64).reshape(4, 4).T
    assert np.allclose(M2, rot(2, 90), atol=1e-7)
    assert np.array_equal(np.array(inst[1].transform).reshape(4, 4), np.eye(4))


def test_json_scene_errors(tmp_path):
    p = json_scene(tmp_path, {"models": [{"a": "a.obj"}], "instances": [{"zz": {}}]},
                   {"a.obj": box_obj(["white"])})
    with pytest.raises(rtxamd.RtxError, match="unknown model"):
        rtxamd.HostScene.load(p)
    p = json_scene(tmp_path, {"models": [{"a": "missing.obj"}], "instances": []}, {})
    with pytest.raises(rtxamd.RtxError, match="missing.obj"):
        rtxamd.HostScene.load(p)
    p = json_scene(tmp_path, {"models": [{"a": "a.obj"}], "instances": []}, {"a.obj": box_obj(["bad"])},
                   mtl="newmtl bad\nillum 5\n")
    with pytest.raises(rtxamd.RtxError, match="illum"):
        rtxamd.HostScene.load(p)
    # a face before any usemtl in the first model has no material
    p = json_scene(tmp_path, {"models": [{"a": "a.obj"}], "instances": []},
                   {"a.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n"})
    with pytest.raises(rtxamd.RtxError, match="material"):
        rtxamd.HostScene.load(p)
    with pytest.raises(rtxamd.RtxError):
        rtxamd.HostScene.load(write(tmp_path / "x.json", "{\"models\": [}"))
    with pytest.raises(rtxamd.RtxError, match="scene file type"):
        rtxamd.HostScene.load(write(tmp_path / "x.yaml", ""))


def test_mtl_map_kd_before_any_kd_sets_diffuse(tmp_path):
    make_png(tmp_path / "js" / "textures" / "t.png")
    mtl = "newmtl t\nmap_Kd t.png\nillum 1\nnewmtl u\nmap_Ks t.png\nKd 0.2 0.2 0.2\nillum 1\n"
    p = json_scene(tmp_path, {"models": [{"m": "m.obj"}], "instances": [{"m": {}}]},
                   {"m.obj": box_obj(["t", "u"])}, mtl=mtl)
    ms = mats(rtxamd.HostScene.load(p))
    assert list(ms[0].diffuse) == pytest.approx([0.6] * 3) and ms[0].textureIdDiffuse == 1
    # map_Ks lands in textureIdDiffuse too (src/SceneLoader.cpp:176-178); one slot per file
    assert ms[1].textureIdDiffuse == 1 and ms[1].textureIdSpecular == -1


def test_mtl_unnamed_default_material_and_first_definition_wins(tmp_path):
    mtl = "Kd 0.9 0.9 0.9\nillum 1\nnewmtl a\nKd 0.1 0 0\nillum 1\nnewmtl a\nKd 0.2 0 0\nillum 1\n"
    p = json_scene(tmp_path, {"models": [{"m": "m.obj"}], "instances": [{"m": {}}]},
                   {"m.obj": box_obj(["a"])}, mtl=mtl)
    hs = rtxamd.HostScene.load(p)
    ms = mats(hs)
    # properties before the first newmtl are dropped; both "a"s are kept, usemtl picks the first
    assert len(ms) == 2 and ms[0].diffuse[0] == pytest.approx(0.1) and ms[1].diffuse[0] == pytest.approx(0.2)
    assert set(models(hs)[0]["mat"].tolist()) == {0}


# ------------------------------------------------------------------ Mitsuba XML

def test_veach_mis_scene():
    """The reference's veachMIS scene: four sphere lights (radius, power),
    four Beckmann plates and the floor, the sensor's lookAt and fov."""
    hs = rtxamd.HostScene.load(GOLDEN / "veachMIS" / "veachMIS.xml")
    d = hs.desc()
    v = hs.view()
    assert list(v.origin) == [0, 2, 15] and list(v.target) == [0, -2, 2.5] and list(v.up) == [0, 1, 0]
    assert v.vfov == 28
    sp = [d.spheres[i] for i in range(d.sphereCount)]
    assert [s.radius for s in sp] == pytest.approx([0.1, 0.03333, 0.3, 0.9])
    assert [list(s.center) for s in sp] == [[-1.25, 0, 0], [-3.75, 0, 0], [1, 0, 0], [3.75, 0, 0]]
    ms = mats(hs)
    rad = [ms[s.materialIndex].lightColor[0] for s in sp]
    assert rad == pytest.approx([100, 901.803, 11.1111, 1.23457], rel=1e-6)
    ls = lights(hs)
    assert [(l.type, l.instanceIndex) for l in ls] == [(abi.LIGHT_SPHERE, i) for i in range(4)]
    assert [l.area for l in ls] == pytest.approx([4 * math.pi * s.radius ** 2 for s in sp], rel=1e-6)
    assert all(l.sampleProb == 0.25 for l in ls)
    plates = [ms[m["mat"][0]] for m in models(hs)[:4]]
    assert [p.type for p in plates] == [abi.MAT_ROUGH_CONDUCTOR] * 4
    assert [p.roughness for p in plates] == pytest.approx([0.02, 0.06, 0.1, 0.2])
    assert all(p.eta == pytest.approx(0.37) and p.k == pytest.approx(2.82) for p in plates)
    floor = ms[models(hs)[4]["mat"][0]]
    assert floor.type == abi.MAT_DIFFUSE and list(floor.diffuse) == pytest.approx([0.4, 0.04, 0.04])
    assert hs.triangle_count() == 2 * 4 + 4


def test_xml_materials_refs_emitters_and_camera(tmp_path):
    write(tmp_path / "tri.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    make_png(tmp_path / "wood.png", 5, 2, "RGB")
    body = """
  <sensor type="perspective">
    <float name="nearClip" value="1"/>
    <float name="fov" value="45"/>
    <transform name="toWorld"><lookat origin="1 2 3" target="0,0,0" up="0, 0, 1"/></transform>
  </sensor>
  <texture type="bitmap" id="woodTex"><string name="filename" value="wood.png"/></texture>
  <bsdf type="phong" id="ph">
    <rgb name="specularReflectance" value="0.2, 0.3, 0.4"/><rgb name="diffuseReflectance" value="0.5"/>
    <float name="exponent" value="12"/>
  </bsdf>
  <bsdf type="diffuse" id="wood"><ref name="reflectance" id="woodTex"/></bsdf>
  <bsdf type="dielectric" id="glass"><float name="intIOR" value="1.5"/><float name="extIOR" value="1.25"/></bsdf>
  <bsdf type="conductor" id="mirror"><string name="material" value="none"/></bsdf>
  <bsdf type="conductor" id="gold"><spectrum name="eta" value="0.2"/><spectrum name="k" value="3.1"/></bsdf>
  <shape type="obj"><string name="filename" value="tri.obj"/><ref id="ph"/></shape>
  <shape type="obj"><string name="filename" value="tri.obj"/><ref id="ph"/>
    <emitter type="area"><rgb name="radiance" value="5 6 7"/></emitter></shape>
  <shape type="sphere"><point name="center" x="1" y="2"/><float name="radius" value="0.5"/><ref id="wood"/></shape>
  <shape type="obj"><string name="filename" value="tri.obj"/><ref id="glass"/></shape>
  <shape type="obj"><string name="filename" value="tri.obj"/><ref id="mirror"/></shape>
  <shape type="obj"><string name="filename" value="tri.obj"/><ref id="gold"/></shape>
  <shape type="sphere"><point name="center" x="0" y="0" z="9"/><float name="radius" value="1"/>
    <bsdf type="diffuse"><rgb name="reflectance" value="0 0 0"/></bsdf>
    <emitter type="area"><rgb name="radiance" value="2"/></emitter></shape>
"""
    hs = rtxamd.HostScene.load(xml_scene(tmp_path, body))
    v = hs.view()
    # fov 45 (the first <float> has a name that is not fov, so the search goes on)
    assert list(v.origin) == [1, 2, 3] and list(v.target) == [0, 0, 0] and list(v.up) == [0, 0, 1] and v.vfov == 45
    ms = mats(hs)
    # 5 scene-level bsdfs, the emitter copy of ph, the inline sphere bsdf and its emitter copy
    assert len(ms) == 8
    ph, wood, glass, mirror, gold, phl, sd, sdl = ms
    assert ph.type == abi.MAT_PHONG and list(ph.specular) == pytest.approx([0.2, 0.3, 0.4])
    assert list(ph.diffuse) == [0.5] * 3 and ph.specularHighlight == 12
    assert wood.type == abi.MAT_DIFFUSE and list(wood.diffuse) == [1, 1, 1] and wood.textureIdDiffuse == 1
    assert glass.type == abi.MAT_DIELECTRIC and glass.refractionIndex == pytest.approx(1.2)
    assert mirror.type == abi.MAT_SPECULAR and list(mirror.specular) == [1, 1, 1]
    assert gold.type == abi.MAT_CONDUCTOR and gold.eta == pytest.approx(0.2) and gold.k == pytest.approx(3.1)
    assert phl.type == abi.MAT_LIGHT and list(phl.lightColor) == [5, 6, 7] and list(phl.diffuse) == [0.5] * 3
    assert sdl.type == abi.MAT_LIGHT and list(sdl.lightColor) == [2, 2, 2]
    ls = lights(hs)
    assert [(l.type, l.instanceIndex) for l in ls] == [(abi.LIGHT_AREA, 1), (abi.LIGHT_SPHERE, 1)]
    d = hs.desc()
    assert list(d.spheres[0].center) == [1, 2, 0]
    tex = hs.textures()
    assert tex[1].shape == (2, 5, 4) and (tex[1][..., 3] == 255).all()


def test_xml_rough_conductor_distribution(tmp_path):
    """The roughconductor's Mitsuba "distribution" picks the microfacet
    distribution: beckmann (Mitsuba's default, and the only one the reference
    renders) or ggx; others are refused."""
    write(tmp_path / "tri.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")

    def rough(idname, dist):
        d = f'<string name="distribution" value="{dist}"/>' if dist else ""
        return (f'<bsdf type="roughconductor" id="{idname}">{d}<float name="alpha" value="0.1"/>'
                '<spectrum name="eta" value="0.5"/><spectrum name="k" value="2"/></bsdf>')
    body = rough("a", None) + rough("b", "beckmann") + rough("c", "ggx") + "".join(
        obj_shape("tri.obj", f'<ref id="{i}"/>') for i in "abc")
    ms = mats(rtxamd.HostScene.load(xml_scene(tmp_path, body)))
    assert [m.type for m in ms] == [abi.MAT_ROUGH_CONDUCTOR] * 3
    assert [m.distribution for m in ms] == [abi.DISTRIBUTION_BECKMANN, abi.DISTRIBUTION_BECKMANN,
                                            abi.DISTRIBUTION_GGX]
    assert ms[2].roughness == pytest.approx(0.1) and ms[2].eta == pytest.approx(0.5)
    with pytest.raises(rtxamd.RtxError, match="distribution"):
        rtxamd.HostScene.load(xml_scene(tmp_path, rough("d", "phong") + obj_shape("tri.obj", '<ref id="d"/>')))


def test_xml_envmap_is_the_last_light(tmp_path):
    env = np.random.default_rng(3).random((8, 16, 4)).astype(np.float32)
    rtxamd.write_exr(tmp_path / "sky.exr", env)
    body = ('<shape type="sphere"><point name="center" x="0" y="0" z="0"/><float name="radius" value="1"/>'
            + QUAD_DIFFUSE + '<emitter type="area"><rgb name="radiance" value="1"/></emitter></shape>'
            + '<emitter type="envmap"><string name="filename" value="sky.exr"/></emitter>')
    hs = rtxamd.HostScene.load(xml_scene(tmp_path, body))
    ls = lights(hs)
    assert [l.type for l in ls] == [abi.LIGHT_SPHERE, abi.LIGHT_ENV_MAP]
    assert all(l.sampleProb == 0.5 for l in ls)
    e = hs.env_map()
    assert e.shape == (8, 16, 4)
    assert np.array_equal(e[..., :3], env[..., :3])


def test_xml_errors(tmp_path):
    # the reference's cornell-dielectric scene names a shell.obj that is not there
    with pytest.raises(rtxamd.RtxError, match="shell.obj"):
        rtxamd.HostScene.load(xml_scene(tmp_path, obj_shape("shell.obj")))
    with pytest.raises(rtxamd.RtxError, match="no material"):
        rtxamd.HostScene.load(xml_scene(tmp_path, '<shape type="sphere"><float name="radius" value="1"/></shape>'))
    with pytest.raises(rtxamd.RtxError, match="duplicate bsdf"):
        rtxamd.HostScene.load(xml_scene(tmp_path, QUAD_DIFFUSE + QUAD_DIFFUSE))
    with pytest.raises(rtxamd.RtxError, match="unknown material type"):
        rtxamd.HostScene.load(xml_scene(tmp_path, '<bsdf type="plastic" id="p"/>'))
    with pytest.raises(rtxamd.RtxError, match="unknown bsdf"):
        write(tmp_path / "t.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
        rtxamd.HostScene.load(xml_scene(tmp_path, obj_shape("t.obj", '<ref id="nope"/>')))
    with pytest.raises(rtxamd.RtxError, match="reflectance"):
        rtxamd.HostScene.load(xml_scene(tmp_path, '<bsdf type="diffuse" id="d"/>'))
    with pytest.raises(rtxamd.RtxError, match="3 values"):
        rtxamd.HostScene.load(xml_scene(tmp_path, '<bsdf type="diffuse" id="d"><rgb name="reflectance" '
                                                  'value="1 , 2, 3, 4"/></bsdf>'.replace("1 , 2, 3, 4", "1 2")))
    with pytest.raises(rtxamd.RtxError, match="texture"):
        rtxamd.HostScene.load(xml_scene(tmp_path, '<texture id="t"><string name="filename" value="no.png"/></texture>'))
    with pytest.raises(rtxamd.RtxError, match="XML"):
        rtxamd.HostScene.load(write(tmp_path / "bad.xml", "<scene><shape></scene>"))


def test_xml_vector_quirk(tmp_path):
    """'1 , 2, 3' stops after the first number (a space is skipped, then ','
    is not a number) and fills all three components with it."""
    hs = rtxamd.HostScene.load(xml_scene(tmp_path, '<bsdf type="diffuse" id="d"><rgb name="reflectance" '
                                                   'value="0.25 , 2, 3"/></bsdf>'))
    assert list(mats(hs)[0].diffuse) == [0.25] * 3


# ------------------------------------------------------------------ images

@pytest.mark.parametrize("mode,bits", [("RGBA", 8), ("RGB", 8), ("L", 8), ("LA", 8), ("P", 8), ("1", 1),
                                       ("I;16", 16), ("RGB16", 16)])
@pytest.mark.parametrize("interlace", [False, True])
def test_png_decoding_matches_pil(tmp_path, mode, bits, interlace):
    from PIL import Image
    rng = np.random.default_rng(11)
    w, h = 37, 23
    path = tmp_path / "x.png"
    if mode == "RGB16":
        # 16-bit RGB written by hand (PIL does not write it); zlib + filter 0
        import struct
        import zlib
        a16 = rng.integers(0, 65536, (h, w, 3), dtype=np.uint16)
        raw = b"".join(b"\x00" + a16[y].astype(">u2").tobytes() for y in range(h))
        if interlace:
            pytest.skip("hand-written 16-bit file is not interlaced")

        def chunk(t, d):
            return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
        path.write_bytes(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 16, 2, 0, 0, 0))
                         + chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))
        got = rtxamd.decode_image(path)
        assert np.array_equal(got[..., :3], (a16 >> 8).astype(np.uint8)) and (got[..., 3] == 255).all()
        return
    if mode == "P":
        img = Image.fromarray(rng.integers(0, 256, (h, w, 3), dtype=np.uint8), "RGB").quantize(17)
        img.info["transparency"] = 3
    elif mode == "1":
        img = Image.fromarray(rng.integers(0, 2, (h, w), dtype=np.uint8) * 255, "L").convert("1")
    elif mode == "I;16":
        img = Image.fromarray(rng.integers(0, 65536, (h, w), dtype=np.uint16))
    else:
        ch = {"RGBA": 4, "RGB": 3, "L": 1, "LA": 2}[mode]
        a = rng.integers(0, 256, (h, w, ch), dtype=np.uint8)
        img = Image.fromarray(a[..., 0] if ch == 1 else a, mode)
    kw = {"transparency": 3} if mode == "P" else {}
    img.save(path, interlace=interlace, **kw)
    got = rtxamd.decode_image(path)
    ref = Image.open(path)
    if mode == "I;16":
        exp = (np.asarray(ref, np.uint32) >> 8).astype(np.uint8)
        exp = np.dstack([exp, exp, exp, np.full_like(exp, 255)])
    else:
        exp = np.asarray(ref.convert("RGBA"))
    assert got.shape == (h, w, 4)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("subsampling,mode,restart", [(2, "RGB", 0), (0, "RGB", 0), (1, "RGB", 0), (2, "L", 0),
                                                      (2, "RGB", 3)])
def test_jpeg_decoding_close_to_libjpeg(tmp_path, subsampling, mode, restart):
    """Sequential JPEG against PIL's libjpeg: decoders differ only in IDCT
    and upsampling rounding."""
    from PIL import Image
    y, x = np.mgrid[0:61, 0:83]
    a = np.stack([(x * 3) % 256, (y * 4) % 256, ((x + y) * 2) % 256], -1).astype(np.uint8)
    img = Image.fromarray(a, "RGB")
    if mode == "L":
        img = img.convert("L")
    path = tmp_path / "x.jpg"
    kw = dict(quality=92, subsampling=subsampling) if mode == "RGB" else dict(quality=92)
    if restart:
        kw["restart_marker_blocks"] = restart
    img.save(path, **kw)
    got = rtxamd.decode_image(path).astype(np.int32)
    exp = np.asarray(Image.open(path).convert("RGBA")).astype(np.int32)
    diff = np.abs(got - exp)
    assert got.shape == exp.shape
    assert diff.max() <= 6 and diff.mean() < 0.6, (diff.max(), diff.mean())


def test_image_errors(tmp_path):
    with pytest.raises(rtxamd.RtxError):
        rtxamd.decode_image(tmp_path / "none.png")
    with pytest.raises(rtxamd.RtxError, match="format"):
        rtxamd.decode_image(write(tmp_path / "x.bmp", "BM...."))
    from PIL import Image
    Image.fromarray(np.zeros((16, 16, 3), np.uint8)).save(tmp_path / "p.jpg", progressive=True)
    with pytest.raises(rtxamd.RtxError, match="progressive"):
        rtxamd.decode_image(tmp_path / "p.jpg")
