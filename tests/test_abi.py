"""C ABI checks that need no GPU: struct layouts, exported symbols, defaults."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

import rtxamd
from rtxamd import abi

ROOT = Path(__file__).resolve().parent.parent
INC = ROOT / "include"

C_NAMES = {
    abi.Vertex: "rtx_vertex", abi.Material: "rtx_material", abi.Instance: "rtx_instance", abi.Light: "rtx_light",
    abi.FaceSample: "rtx_face_sample", abi.Sphere: "rtx_sphere", abi.PushConstant: "rtx_push_constant",
    abi.Camera: "rtx_camera", abi.ModelDesc: "rtx_model_desc", abi.SceneDesc: "rtx_scene_desc", abi.TextureDesc: "rtx_texture_desc",
    abi.FrameParams: "rtx_frame_params", abi.FrameStats: "rtx_frame_stats", abi.SceneInfo: "rtx_scene_info",
    abi.Ray: "rtx_ray", abi.Hit: "rtx_hit", abi.Tuning: "rtx_tuning", abi.CacheHeader: "rtx_cache_header",
    abi.CacheData: "rtx_cache_data", abi.IcJob: "rtx_ic_job", abi.Aabb: "rtx_aabb", abi.VmfTheta: "rtx_vmf_theta",
    abi.VmmTheta: "rtx_vmm_theta", abi.NrcConfig: "rtx_nrc_config", abi.NrcInput: "rtx_nrc_input",
    abi.NrcInfo: "rtx_nrc_info", abi.ServerConfig: "rtx_server_config", abi.JobDesc: "rtx_job_desc",
    abi.JobStatus: "rtx_job_status", abi.ServerStats: "rtx_server_stats", abi.AutotuneResult: "rtx_autotune_result",
    abi.GuidingOptions: "rtx_guiding_options", abi.VmmThetaWide: "rtx_vmm_theta_wide",
}


def test_ctypes_mirror_matches_c_header(tmp_path):
    """Every ctypes field sits at the offset the C compiler gives it."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rtx_amd.h"', 'int main(void) {']
    for cls, cname in C_NAMES.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", f"-I{INC}", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            cname, field, value = ln.split()
            got[(cname, field)] = int(value)
    for cls, cname in C_NAMES.items():
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


# Device layouts of the reference (std140 / std430 / scalar rules applied to
# shaders/wavefront.glsl:1-45, shaders/raycommon.glsl:19-96).
REFERENCE_LAYOUT = {
    abi.Vertex: (48, {"pos": 0, "normal": 16, "texCoord": 32, "materialIndex": 40}),
    abi.Material: (96, {"lightColor": 0, "diffuse": 16, "specular": 32, "specularHighlight": 44,
                        "transparency": 48, "refractionIndex": 52, "refractionIndexInv": 56, "eta": 60,
                        "k": 64, "roughness": 68, "textureIdDiffuse": 72, "textureIdSpecular": 76,
                        "type": 80}),
    abi.Instance: (144, {"transform": 0, "normalTransform": 64, "modelIndex": 128, "iLight": 132}),
    abi.Light: (40, {"color": 0, "pos": 12, "instanceIndex": 24, "sampleProb": 28, "area": 32, "type": 36}),
    abi.FaceSample: (12, {"index": 0, "sampleProb": 4, "faceArea": 8}),
    abi.Sphere: (24, {"center": 0, "radius": 12, "materialIndex": 16, "iLight": 20}),
    abi.Camera: (256, {"view": 0, "proj": 64, "viewInverse": 128, "projInverse": 192}),
    abi.CacheHeader: (12, {"nextCacheSlot": 0, "maxCaches": 4, "nextUpdateSlot": 8}),
    abi.CacheData: (56, {"color": 0, "normal": 12, "rotGrad": 24, "transGrad": 36, "harmonicR": 48,
                         "numUpdates": 52}),
}


def test_reference_device_layouts():
    for cls, (size, offs) in REFERENCE_LAYOUT.items():
        assert C.sizeof(cls) == size
        for f, o in offs.items():
            assert getattr(cls, f).offset == o, (cls.__name__, f)


def test_push_constant_is_48_fields_in_reference_order():
    ref = (ROOT / "tests" / "golden" / "push_constant_fields.txt").read_text().split()
    assert [f for f, _ in abi.PushConstant._fields_] == ref
    assert C.sizeof(abi.PushConstant) == 192
    for i, (f, _) in enumerate(abi.PushConstant._fields_):
        assert getattr(abi.PushConstant, f).offset == 4 * i


def _declared(header):
    text = (INC / header).read_text()
    return sorted(set(re.findall(r"^[a-z_0-9 \*]+?\b((?:rtx|rtxh)_[a-z0-9_]+)\s*\(", text, re.M)))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], check=True, capture_output=True, text=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_device_library_exports_every_declared_entry_point():
    names = _declared("rtx_amd.h")
    assert "rtx_render" in names and "rtx_scene_create" in names
    exp = _exported(rtxamd.LIB_DIR / "librtxamd.so")
    missing = [n for n in names if n not in exp]
    assert not missing, missing
    lib = C.CDLL(str(rtxamd.LIB_DIR / "librtxamd.so"))
    for n in names:
        assert hasattr(lib, n)


def test_host_library_exports_every_declared_entry_point():
    names = _declared("rtx_host.h")
    assert "rtxh_scene_generate" in names
    exp = _exported(rtxamd.LIB_DIR / "librtxhost.so")
    missing = [n for n in names if n not in exp]
    assert not missing, missing


def test_push_constant_defaults_are_the_references():
    """rtx_push_constant_defaults (host-side function, no device needed)
    fills src/RayTracingApp.h:81-130."""
    pc = abi.PushConstant()
    rtxamd.device_lib().rtx_push_constant_defaults(C.byref(pc))
    ref = abi.push_constant_defaults()
    for f, _ in abi.PushConstant._fields_:
        assert getattr(pc, f) == pytest.approx(getattr(ref, f)), f
    assert pc.maxDepth == 30 and pc.enableNEE == 1 and pc.enableMIS == 0 and pc.usePowerHeuristic == 1
    assert pc.previousFrames == 0xFFFFFFFF


def test_context_creation_without_device_fails_cleanly():
    import os
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = rtxamd.device_lib().rtx_ctx_create(0, C.byref(h))
    assert rc == abi.RTX_E_DEVICE and not h.value


def test_header_constants_match_python_mirror():
    """Every integer #define of the C headers that the Python mirror also
    names has the same value there (frame flags, error codes, limits)."""
    checked = 0
    for h in sorted(INC.glob("*.h")):
        for name, val in re.findall(r"^#define\s+(RTX\w*)\s+\(?(-?(?:0x[0-9A-Fa-f]+|\d+))[uU]?\)?\s*(?:/\*.*)?$",
                                    h.read_text(), re.M):
            if hasattr(abi, name):
                assert getattr(abi, name) == int(val, 0), (h.name, name)
                checked += 1
    assert checked >= 10 and abi.RTX_FRAME_TIME_TRACE == 0x20
