"""Python binding of the MI355X path tracer (ctypes over the C ABI).

librtxamd.so  - HIP kernels behind include/rtx_amd.h (the product)
librtxhost.so - C++ host scene layer behind include/rtx_host.h

Both are built in-tree by __graft_entry__.build() (make in this folder).
Nothing here falls back to a CPU path: a missing library raises.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

from . import abi
from .abi import (FrameParams, FrameStats, Hit, Ray, SceneDesc, SceneInfo, View, Camera,
                  push_constant_defaults)

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_DIR = PKG_DIR / "lib"
BIN_DIR = PKG_DIR / "bin"

_host = None
_dev = None


class RtxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rtx error {code}: {msg}")
        self.code = code


def _load(name):
    path = LIB_DIR / name
    if not path.exists():
        raise RtxError(abi.RTX_E_STATE, f"{path} is not built; run __graft_entry__.build()")
    return C.CDLL(str(path), mode=C.RTLD_GLOBAL)


def host_lib():
    global _host
    if _host is None:
        L = _load("librtxhost.so")
        L.rtxh_last_error.restype = C.c_char_p
        L.rtxh_scene_generate.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]
        L.rtxh_scene_load_xml.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.rtxh_scene_load.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.rtxh_decode_image.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32)]
        L.rtxh_scene_destroy.argtypes = [C.c_void_p]
        L.rtxh_scene_destroy.restype = None
        L.rtxh_scene_desc.argtypes = [C.c_void_p, C.POINTER(SceneDesc)]
        L.rtxh_scene_view.argtypes = [C.c_void_p, C.POINTER(View)]
        L.rtxh_camera.argtypes = [C.POINTER(View), C.c_uint32, C.c_uint32, C.POINTER(Camera)]
        L.rtxh_write_exr.argtypes = [C.c_char_p, C.POINTER(C.c_float), C.c_uint32, C.c_uint32]
        L.rtxh_read_exr.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32)]
        L.rtxh_free.argtypes = [C.c_void_p]
        L.rtxh_free.restype = None
        L.rtxh_scene_new.argtypes = [C.POINTER(C.c_void_p)]
        L.rtxh_scene_add_material.argtypes = [C.c_void_p, C.POINTER(abi.Material), C.POINTER(C.c_int32)]
        L.rtxh_scene_add_mesh.argtypes = [C.c_void_p, C.POINTER(abi.Vertex), C.c_uint32, C.POINTER(C.c_uint32),
                                          C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                          C.POINTER(C.c_int32)]
        L.rtxh_scene_add_sphere.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_float, C.c_int32,
                                            C.POINTER(C.c_int32)]
        L.rtxh_scene_add_point_light.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.rtxh_scene_set_view.argtypes = [C.c_void_p, C.POINTER(View)]
        L.rtxh_scene_add_texture.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), C.c_uint32, C.c_uint32,
                                             C.POINTER(C.c_int32)]
        L.rtxh_scene_set_environment.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_uint32, C.c_uint32, C.c_int]
        L.rtxh_scene_finalize.argtypes = [C.c_void_p]
        L.rtxh_pcg32_first.argtypes = [C.c_uint64, C.c_uint64]
        L.rtxh_pcg32_first.restype = C.c_uint32
        _host = L
    return _host


def device_lib():
    global _dev
    if _dev is None:
        L = _load("librtxamd.so")
        L.rtx_last_error.argtypes = [C.c_void_p]
        L.rtx_last_error.restype = C.c_char_p
        L.rtx_push_constant_defaults.argtypes = [C.POINTER(abi.PushConstant)]
        L.rtx_push_constant_defaults.restype = None
        L.rtx_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.rtx_ctx_destroy.argtypes = [C.c_void_p]
        L.rtx_scene_create.argtypes = [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]
        L.rtx_scene_destroy.argtypes = [C.c_void_p, C.c_void_p]
        L.rtx_scene_get_info.argtypes = [C.c_void_p, C.POINTER(SceneInfo)]
        L.rtx_render.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(FrameParams), C.POINTER(FrameStats)]
        L.rtx_read_image.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_size_t]
        L.rtx_accumulated_samples.argtypes = [C.c_void_p, C.POINTER(C.c_uint32)]
        L.rtx_read_paths.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.c_size_t]
        L.rtx_trace_closest.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(Ray), C.c_size_t, C.POINTER(Hit)]
        L.rtx_trace_occluded.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(Ray), C.c_size_t,
                                         C.POINTER(C.c_uint8)]
        L.rtx_trace_closest_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(Ray), C.c_size_t,
                                                C.POINTER(Hit)]
        L.rtx_ic_create.argtypes = [C.c_void_p, C.c_uint32]
        L.rtx_ic_reset.argtypes = [C.c_void_p]
        L.rtx_ic_read.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]
        L.rtx_ic_write.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.rtx_ic_read_jobs.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_uint32),
                                       C.c_uint32, C.POINTER(C.c_uint32)]
        L.rtx_guiding_set.argtypes = [C.c_void_p, C.POINTER(abi.Aabb), C.POINTER(abi.VmmTheta), C.c_uint32]
        L.rtx_guiding_get.argtypes = [C.c_void_p, C.POINTER(abi.Aabb), C.POINTER(abi.VmmTheta), C.c_uint32,
                                      C.POINTER(C.c_uint32)]
        L.rtx_guiding_read_samples.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.rtx_guiding_get_wide.argtypes = [C.c_void_p, C.POINTER(abi.Aabb), C.POINTER(abi.VmmThetaWide), C.c_uint32,
                                           C.POINTER(C.c_uint32)]
        L.rtx_guiding_init.argtypes = [C.c_void_p, C.POINTER(abi.Aabb), C.c_uint32, C.c_int]
        L.rtx_guiding_update.argtypes = [C.c_void_p]
        L.rtx_guiding_set_options.argtypes = [C.c_void_p, C.POINTER(abi.GuidingOptions)]
        L.rtx_guiding_get_options.argtypes = [C.c_void_p, C.POINTER(abi.GuidingOptions)]
        L.rtx_schedule_config_defaults.argtypes = [C.POINTER(abi.ScheduleConfig)]
        L.rtx_schedule_config_defaults.restype = None
        L.rtx_schedule_create.argtypes = [C.POINTER(abi.ScheduleConfig), C.POINTER(abi.PushConstant),
                                          C.POINTER(C.c_void_p)]
        L.rtx_schedule_destroy.argtypes = [C.c_void_p]
        L.rtx_schedule_set_push_constants.argtypes = [C.c_void_p, C.POINTER(abi.PushConstant)]
        L.rtx_schedule_input_changed.argtypes = [C.c_void_p]
        L.rtx_schedule_begin_frame.argtypes = [C.c_void_p, C.POINTER(abi.FrameParams), C.POINTER(abi.FrameParams)]
        L.rtx_schedule_end_frame.argtypes = [C.c_void_p]
        L.rtx_schedule_frame.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(abi.FrameParams),
                                         C.POINTER(abi.FrameStats)]
        L.rtx_schedule_get_state.argtypes = [C.c_void_p, C.POINTER(abi.ScheduleState)]
        L.rtx_estimate_write.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_uint32, C.c_uint32]
        L.rtx_estimate_read.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_uint32, C.c_uint32]
        fp = C.POINTER(C.c_float)
        L.rtx_nrc_config_defaults.argtypes = [C.POINTER(abi.NrcConfig)]
        L.rtx_nrc_config_defaults.restype = None
        L.rtx_nrc_create.argtypes = [C.c_void_p, C.POINTER(abi.NrcConfig)]
        L.rtx_nrc_destroy.argtypes = [C.c_void_p]
        L.rtx_nrc_get_info.argtypes = [C.c_void_p, C.POINTER(abi.NrcInfo)]
        L.rtx_nrc_get_params.argtypes = [C.c_void_p, fp, C.c_size_t, fp, C.c_size_t]
        L.rtx_nrc_set_params.argtypes = [C.c_void_p, fp, C.c_size_t, fp, C.c_size_t]
        L.rtx_nrc_infer.argtypes = [C.c_void_p, C.POINTER(abi.NrcInput), C.c_size_t, fp]
        L.rtx_nrc_gradients.argtypes = [C.c_void_p, C.POINTER(abi.NrcInput), fp, C.c_size_t, fp, fp, fp]
        L.rtx_nrc_train.argtypes = [C.c_void_p, C.POINTER(abi.NrcInput), fp, C.c_size_t, fp]
        L.rtx_server_config_defaults.argtypes = [C.POINTER(abi.ServerConfig)]
        L.rtx_server_config_defaults.restype = None
        L.rtx_server_create.argtypes = [C.c_void_p, C.POINTER(abi.ServerConfig), C.POINTER(C.c_void_p)]
        L.rtx_server_destroy.argtypes = [C.c_void_p]
        L.rtx_server_submit.argtypes = [C.c_void_p, C.POINTER(abi.JobDesc), C.POINTER(C.c_uint64)]
        L.rtx_server_poll.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(abi.JobStatus)]
        L.rtx_server_wait.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(abi.JobStatus)]
        L.rtx_server_read_image.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_float), C.c_size_t]
        L.rtx_server_get_stats.argtypes = [C.c_void_p, C.POINTER(abi.ServerStats)]
        L.rtx_server_last_error.argtypes = [C.c_void_p]
        L.rtx_server_last_error.restype = C.c_char_p
        L.rtx_autotune.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(FrameParams), C.c_uint32,
                                   C.POINTER(abi.AutotuneResult)]
        L.rtx_ctx_set_tuning.argtypes = [C.c_void_p, C.POINTER(abi.Tuning)]
        L.rtx_ctx_get_tuning.argtypes = [C.c_void_p, C.POINTER(abi.Tuning)]
        _dev = L
    return _dev


def _host_check(rc):
    if rc != abi.RTX_OK:
        raise RtxError(rc, host_lib().rtxh_last_error().decode())


class HostScene:
    """A scene held by the C++ host layer (generated or loaded)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def generate(cls, kind, seed=1, triangles=0):
        h = C.c_void_p()
        _host_check(host_lib().rtxh_scene_generate(int(kind), int(seed), int(triangles), C.byref(h)))
        return cls(h)

    @classmethod
    def new(cls):
        h = C.c_void_p()
        _host_check(host_lib().rtxh_scene_new(C.byref(h)))
        return cls(h)

    def add_material(self, mat):
        i = C.c_int32()
        _host_check(host_lib().rtxh_scene_add_material(self._h, C.byref(mat), C.byref(i)))
        return i.value

    def add_mesh(self, positions, indices, materials, normals=None, transform=None, normal_transform=None,
                 texcoords=None):
        """positions (n,3), indices (m,3), materials: per-vertex or scalar material index,
        texcoords (n,2) or None."""
        positions = np.asarray(positions, np.float32).reshape(-1, 3)
        indices = np.ascontiguousarray(np.asarray(indices, np.uint32).reshape(-1))
        n = positions.shape[0]
        verts = (abi.Vertex * n)()
        if normals is None:
            normals = np.zeros_like(positions)
            tri = indices.reshape(-1, 3)
            fn = np.cross(positions[tri[:, 1]] - positions[tri[:, 0]], positions[tri[:, 2]] - positions[tri[:, 0]])
            fn /= np.linalg.norm(fn, axis=1, keepdims=True)
            for k in range(3):
                normals[tri[:, k]] = fn
        normals = np.asarray(normals, np.float32).reshape(-1, 3)
        mats = np.broadcast_to(np.asarray(materials, np.int32), (n,))
        uv = None if texcoords is None else np.asarray(texcoords, np.float32).reshape(-1, 2)
        for i in range(n):
            verts[i].pos[:] = positions[i].tolist()
            verts[i].normal[:] = normals[i].tolist()
            verts[i].materialIndex = int(mats[i])
            if uv is not None:
                verts[i].texCoord[:] = uv[i].tolist()
        fp = C.POINTER(C.c_float)
        tr = (C.c_float * 16)(*np.asarray(transform, np.float32).reshape(16, order="F")) if transform is not None else None
        ntr = (C.c_float * 16)(*np.asarray(normal_transform, np.float32).reshape(16, order="F")) \
            if normal_transform is not None else None
        inst = C.c_int32()
        _host_check(host_lib().rtxh_scene_add_mesh(
            self._h, verts, n, indices.ctypes.data_as(C.POINTER(C.c_uint32)), indices.size,
            C.cast(tr, fp) if tr is not None else None, C.cast(ntr, fp) if ntr is not None else None, C.byref(inst)))
        return inst.value

    def add_sphere(self, center, radius, material):
        c = (C.c_float * 3)(*center)
        i = C.c_int32()
        _host_check(host_lib().rtxh_scene_add_sphere(self._h, c, float(radius), int(material), C.byref(i)))
        return i.value

    def add_point_light(self, pos, color):
        _host_check(host_lib().rtxh_scene_add_point_light(self._h, (C.c_float * 3)(*pos), (C.c_float * 3)(*color)))

    def set_view(self, origin, target, up, vfov):
        v = View()
        v.origin[:] = list(origin)
        v.target[:] = list(target)
        v.up[:] = list(up)
        v.vfov = vfov
        _host_check(host_lib().rtxh_scene_set_view(self._h, C.byref(v)))

    def add_texture(self, rgba):
        """(h, w, 4) uint8 RGBA (RGB sRGB-encoded) -> texture slot for Material.textureIdDiffuse."""
        a = np.ascontiguousarray(rgba, np.uint8)
        h, w = a.shape[:2]
        slot = C.c_int32()
        _host_check(host_lib().rtxh_scene_add_texture(self._h, a.ctypes.data_as(C.POINTER(C.c_uint8)), w, h,
                                                      C.byref(slot)))
        return slot.value

    def set_environment(self, rgba, add_light=True):
        """(h, w, 4) float32 lat-long environment map; add_light appends the environment light."""
        a = np.ascontiguousarray(rgba, np.float32)
        h, w = a.shape[:2]
        _host_check(host_lib().rtxh_scene_set_environment(self._h, a.ctypes.data_as(C.POINTER(C.c_float)), w, h,
                                                          1 if add_light else 0))

    def finalize(self):
        _host_check(host_lib().rtxh_scene_finalize(self._h))
        return self

    @classmethod
    def load_xml(cls, path):
        h = C.c_void_p()
        _host_check(host_lib().rtxh_scene_load_xml(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, path):
        """A Mitsuba .xml or the reference's .json scene file (rtxh_scene_load)."""
        h = C.c_void_p()
        _host_check(host_lib().rtxh_scene_load(str(path).encode(), C.byref(h)))
        return cls(h)

    def textures(self):
        """Material textures by slot as (height, width, 4) uint8 arrays; slot 0
        (the environment map's) is None."""
        d = self.desc()
        out = []
        for i in range(d.textureCount):
            t = d.textures[i]
            if not t.rgba or t.width == 0:
                out.append(None)
                continue
            a = np.ctypeslib.as_array(t.rgba, shape=(t.height * t.width * 4,))
            out.append(a.reshape(t.height, t.width, 4).copy())
        return out

    def env_map(self):
        """The environment map as a (height, width, 4) float32 array, or None."""
        d = self.desc()
        if not d.envMap:
            return None
        a = np.ctypeslib.as_array(d.envMap, shape=(d.envHeight * d.envWidth * 4,))
        return a.reshape(d.envHeight, d.envWidth, 4).copy()

    def close(self):
        if self._h:
            host_lib().rtxh_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def desc(self):
        d = SceneDesc()
        _host_check(host_lib().rtxh_scene_desc(self._h, C.byref(d)))
        return d

    def view(self):
        v = View()
        _host_check(host_lib().rtxh_scene_view(self._h, C.byref(v)))
        return v

    def camera(self, width, height, view=None):
        cam = Camera()
        v = view if view is not None else self.view()
        _host_check(host_lib().rtxh_camera(C.byref(v), int(width), int(height), C.byref(cam)))
        return cam

    def triangle_count(self):
        d = self.desc()
        total = 0
        for i in range(d.instanceCount):
            total += d.models[d.instances[i].modelIndex].indexCount // 3
        return total


def frame_params(cam, width, height, spp, seed=1, sample_offset=0, flags=0, rr_start_depth=3,
                 batch_paths=0, **pc_overrides):
    fp = FrameParams()
    fp.pc = push_constant_defaults()
    for k, v in pc_overrides.items():
        if not hasattr(fp.pc, k):
            raise KeyError(k)
        setattr(fp.pc, k, v)
    fp.cam = cam
    fp.width = width
    fp.height = height
    fp.spp = spp
    fp.sampleOffset = sample_offset
    fp.seed = seed
    fp.flags = flags
    fp.rrStartDepth = rr_start_depth
    fp.batchPaths = batch_paths
    return fp


def _material(mtype, **kw):
    m = abi.Material()
    m.textureIdDiffuse = -1
    m.textureIdSpecular = -1
    m.type = mtype
    for k, v in kw.items():
        if isinstance(v, (list, tuple)):
            getattr(m, k)[:] = [float(x) for x in v]
        else:
            setattr(m, k, v)
    return m


def diffuse(albedo):
    return _material(abi.MAT_DIFFUSE, diffuse=albedo)


def light(radiance, albedo=(0.0, 0.0, 0.0)):
    return _material(abi.MAT_LIGHT, diffuse=albedo, lightColor=radiance)


def dielectric(ior):
    """parseXmlBSDF 'dielectric', src/SceneLoader.cpp:587-591."""
    return _material(abi.MAT_DIELECTRIC, specular=(1, 1, 1), refractionIndex=ior, refractionIndexInv=1.0 / ior)


def rough_conductor(alpha, eta, k, distribution=abi.DISTRIBUTION_BECKMANN):
    """A rough conductor; distribution abi.DISTRIBUTION_BECKMANN (the
    reference's) or abi.DISTRIBUTION_GGX."""
    return _material(abi.MAT_ROUGH_CONDUCTOR, roughness=alpha, eta=eta, k=k, distribution=distribution)


def conductor(eta, k):
    return _material(abi.MAT_CONDUCTOR, eta=eta, k=k)


def mirror():
    return _material(abi.MAT_SPECULAR, specular=(1, 1, 1))


def phong(kd, ks, exponent):
    return _material(abi.MAT_PHONG, diffuse=kd, specular=ks, specularHighlight=exponent)


def write_exr(path, rgba):
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    h, w = rgba.shape[:2]
    _host_check(host_lib().rtxh_write_exr(str(path).encode(), rgba.ctypes.data_as(C.POINTER(C.c_float)), w, h))


def read_exr(path):
    p = C.POINTER(C.c_float)()
    w, h = C.c_uint32(), C.c_uint32()
    _host_check(host_lib().rtxh_read_exr(str(path).encode(), C.byref(p), C.byref(w), C.byref(h)))
    try:
        arr = np.ctypeslib.as_array(p, shape=(h.value, w.value, 4)).copy()
    finally:
        host_lib().rtxh_free(p)
    return arr


def decode_image(path):
    """PNG / baseline JPEG file -> (height, width, 4) uint8 (rtxh_decode_image)."""
    p = C.POINTER(C.c_uint8)()
    w, h = C.c_uint32(), C.c_uint32()
    _host_check(host_lib().rtxh_decode_image(str(path).encode(), C.byref(p), C.byref(w), C.byref(h)))
    try:
        return np.ctypeslib.as_array(p, shape=(h.value * w.value * 4,)).reshape(h.value, w.value, 4).copy()
    finally:
        host_lib().rtxh_free(p)


def pcg32_first(initstate, initseq):
    return int(host_lib().rtxh_pcg32_first(int(initstate), int(initseq)))


def rays_array(org, dirs, tmin=1e-3, tmax=1e6):
    """Pack (n,3) origins/directions into the rtx_ray layout (n, 8) float32."""
    org = np.asarray(org, np.float32)
    dirs = np.asarray(dirs, np.float32)
    n = org.shape[0]
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = org
    r[:, 3] = tmin
    r[:, 4:7] = dirs
    r[:, 7] = tmax
    return r


def vmf_theta(mu, kappa, distance=-1.0, target=(0.0, 0.0, 0.0)):
    """One lobe as VMF_Theta::setK fills it (src/PathGuiding.h:44-50): kappa
    below 1e-3 becomes 0, norm = k / (2 pi (1 - e^-2k)), eMin2K = e^-2k."""
    t = abi.VmfTheta()
    mu = np.asarray(mu, np.float64)
    mu = mu / np.linalg.norm(mu)
    t.mu[:] = [float(x) for x in mu]
    k = np.float32(0.0 if kappa < 1e-3 else kappa)
    t.k = float(k)
    t.norm = float(np.float32(k / (2 * np.pi * (1 - np.exp(-2.0 * float(k)))))) if k > 0 else 0.0
    t.eMin2K = float(np.float32(np.exp(-2.0 * float(k))))
    t.distance = float(distance)
    t.target[:] = [float(x) for x in target]
    return t


def vmm_theta(lobes, weights, mean_position=(0.0, 0.0, 0.0)):
    """A mixture: lobes = [vmf_theta, ...], weights summing to 1."""
    m = abi.VmmTheta()
    m.usedDistributions = len(lobes)
    for i, (t, w) in enumerate(zip(lobes, weights)):
        m.thetas[i] = t
        m.pi[i] = float(w)
    m.meanPosition[:] = [float(x) for x in mean_position]
    return m


def guiding_arrays(regions, vmms):
    regions = np.asarray(regions, np.float32).reshape(-1, 2, 3)
    boxes = (abi.Aabb * max(1, len(regions)))()
    for i, (lo, hi) in enumerate(regions):
        boxes[i].min[:] = [float(x) for x in lo]
        boxes[i].max[:] = [float(x) for x in hi]
    arr = (abi.VmmTheta * max(1, len(vmms)))()
    for i, v in enumerate(vmms):
        arr[i] = v
    return boxes, arr


class IcState:
    """Host copy of an irradiance cache: the header and the entry and sphere
    buffers of src/IrradianceCache.cpp:46-75, as numpy arrays."""

    def __init__(self, hdr, entries, spheres):
        self.hdr = hdr            # CacheHeader
        self.entries = entries    # CACHE_DTYPE[maxCaches]
        self.spheres = spheres    # SPHERE_DTYPE[maxCaches]

    @classmethod
    def empty(cls, max_entries):
        """The freshly created cache (IrradianceCache::createBuffers): no
        entries, harmonicR -1, spheres of radius 0 at the origin."""
        hdr = abi.CacheHeader(0, max_entries, 0)
        e = np.zeros(max_entries, abi.CACHE_DTYPE)
        e["harmonicR"] = -1.0
        sp = np.zeros(max_entries, abi.SPHERE_DTYPE)
        return cls(hdr, e, sp)

    @property
    def max_entries(self):
        return int(self.hdr.maxCaches)

    @property
    def count(self):
        return int(min(self.hdr.nextCacheSlot, self.hdr.maxCaches))

    def copy(self):
        return IcState(abi.CacheHeader(self.hdr.nextCacheSlot, self.hdr.maxCaches, self.hdr.nextUpdateSlot),
                       self.entries.copy(), self.spheres.copy())

    def ctypes_views(self):
        return (C.cast(C.pointer(self.hdr), C.c_void_p), self.entries.ctypes.data_as(C.c_void_p),
                self.spheres.ctypes.data_as(C.c_void_p))


class DeviceScene:
    def bounds(self):
        i = self.info()
        return np.array([i.boundsMin[:], i.boundsMax[:]], np.float32)

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self._h = handle

    def info(self):
        i = SceneInfo()
        self.ctx._check(device_lib().rtx_scene_get_info(self._h, C.byref(i)))
        return i

    def close(self):
        if self._h and self.ctx._h:
            device_lib().rtx_scene_destroy(self.ctx._h, self._h)
        self._h = None


class Renderer:
    """A librtxamd context on one HIP device."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        rc = device_lib().rtx_ctx_create(int(device), C.byref(self._h))
        if rc != abi.RTX_OK:
            msg = device_lib().rtx_last_error(self._h).decode() if self._h else "context creation failed"
            raise RtxError(rc, msg)

    def _check(self, rc):
        if rc != abi.RTX_OK:
            raise RtxError(rc, device_lib().rtx_last_error(self._h).decode())

    def close(self):
        if self._h:
            device_lib().rtx_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, host_scene):
        d = host_scene.desc()
        h = C.c_void_p()
        self._check(device_lib().rtx_scene_create(self._h, C.byref(d), C.byref(h)))
        return DeviceScene(self, h)

    def tuning(self):
        t = abi.Tuning()
        self._check(device_lib().rtx_ctx_get_tuning(self._h, C.byref(t)))
        return t

    def set_tuning(self, **kw):
        """Launch-shape knobs (traceBlocksPerCU, shadeBlocksPerCU, batchPaths,
        ldsStackEntries, cacheUpkeep); fields left out keep their current values."""
        t = abi.Tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        self._check(device_lib().rtx_ctx_set_tuning(self._h, C.byref(t)))

    def autotune(self, scene, probe, repeats=2):
        """rtx_autotune: times `probe` under candidate launch shapes, keeps
        the fastest on the context and returns the rtx_autotune_result."""
        res = abi.AutotuneResult()
        self._check(device_lib().rtx_autotune(self._h, scene._h, C.byref(probe), int(repeats), C.byref(res)))
        return res

    # --------------------------------------------- learned radiance cache
    def nrc_create(self, scene_min, scene_max, **kw):
        """rtx_nrc_create with the defaults of rtx_nrc_config_defaults, the
        encoded box and any config fields given as keywords."""
        c = abi.NrcConfig()
        device_lib().rtx_nrc_config_defaults(C.byref(c))
        c.sceneMin[:] = [float(x) for x in scene_min]
        c.sceneMax[:] = [float(x) for x in scene_max]
        for k, v in kw.items():
            setattr(c, k, v)
        self._check(device_lib().rtx_nrc_create(self._h, C.byref(c)))
        return c

    def nrc_destroy(self):
        self._check(device_lib().rtx_nrc_destroy(self._h))

    def nrc_info(self):
        info = abi.NrcInfo()
        self._check(device_lib().rtx_nrc_get_info(self._h, C.byref(info)))
        return info

    def nrc_params(self):
        """(table (levels, tableSize, 2), mlp (kNrcMlpParams,)) FP32 masters."""
        info = self.nrc_info()
        grid = np.empty(int(info.gridParams), np.float32)
        mlp = np.empty(int(info.mlpParams), np.float32)
        fp = C.POINTER(C.c_float)
        self._check(device_lib().rtx_nrc_get_params(self._h, grid.ctypes.data_as(fp), grid.size,
                                                     mlp.ctypes.data_as(fp), mlp.size))
        return grid.reshape(info.levels, info.tableSize, 2), mlp

    def nrc_set_params(self, grid=None, mlp=None):
        fp = C.POINTER(C.c_float)
        g = None if grid is None else np.ascontiguousarray(grid, np.float32).reshape(-1)
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32).reshape(-1)
        self._check(device_lib().rtx_nrc_set_params(self._h, None if g is None else g.ctypes.data_as(fp),
                                                     0 if g is None else g.size,
                                                     None if m is None else m.ctypes.data_as(fp),
                                                     0 if m is None else m.size))

    def nrc_infer(self, inputs):
        """inputs: (n, 16) float32 rows laid out as rtx_nrc_input; returns (n, 3)."""
        x = np.ascontiguousarray(inputs, np.float32)
        out = np.empty((len(x), 3), np.float32)
        self._check(device_lib().rtx_nrc_infer(self._h, x.ctypes.data_as(C.POINTER(abi.NrcInput)), len(x),
                                                out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def nrc_gradients(self, inputs, targets):
        """(loss, table gradient, mlp gradient) without a parameter change."""
        x = np.ascontiguousarray(inputs, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        info = self.nrc_info()
        gg = np.empty(int(info.gridParams), np.float32)
        mg = np.empty(int(info.mlpParams), np.float32)
        loss = C.c_float()
        fp = C.POINTER(C.c_float)
        self._check(device_lib().rtx_nrc_gradients(self._h, x.ctypes.data_as(C.POINTER(abi.NrcInput)),
                                                    t.ctypes.data_as(fp), len(x), gg.ctypes.data_as(fp),
                                                    mg.ctypes.data_as(fp), C.byref(loss)))
        return loss.value, gg.reshape(info.levels, info.tableSize, 2), mg

    def nrc_train(self, inputs, targets):
        """One Adam step; returns the loss before it."""
        x = np.ascontiguousarray(inputs, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        loss = C.c_float()
        self._check(device_lib().rtx_nrc_train(self._h, x.ctypes.data_as(C.POINTER(abi.NrcInput)),
                                                t.ctypes.data_as(C.POINTER(C.c_float)), len(x), C.byref(loss)))
        return loss.value

    # ------------------------------------------------------ guiding
    def guiding_set(self, regions, vmms):
        """regions: (n, 2, 3) boxes; vmms: sequence of abi.VmmTheta."""
        boxes, arr = guiding_arrays(regions, vmms)
        self._check(device_lib().rtx_guiding_set(self._h, boxes, arr, len(vmms)))

    def guiding_init(self, scene_box, splits, parallax=True):
        box = abi.Aabb()
        box.min[:] = [float(x) for x in scene_box[0]]
        box.max[:] = [float(x) for x in scene_box[1]]
        self._check(device_lib().rtx_guiding_init(self._h, C.byref(box), int(splits), int(parallax)))

    def guiding_update(self):
        self._check(device_lib().rtx_guiding_update(self._h))

    def guiding_get_wide(self):
        """(boxes (n, 2, 3), [abi.VmmThetaWide]): mixtures of up to 32 lobes."""
        n = C.c_uint32()
        self._check(device_lib().rtx_guiding_get_wide(self._h, None, None, 0, C.byref(n)))
        boxes = (abi.Aabb * max(1, n.value))()
        vmms = (abi.VmmThetaWide * max(1, n.value))()
        self._check(device_lib().rtx_guiding_get_wide(self._h, boxes, vmms, n.value, C.byref(n)))
        arr = np.array([[boxes[i].min[:], boxes[i].max[:]] for i in range(n.value)], np.float32)
        return arr, [vmms[i] for i in range(n.value)]

    def guiding_get(self):
        """(boxes (n, 2, 3), [abi.VmmTheta]) of the context."""
        n = C.c_uint32()
        self._check(device_lib().rtx_guiding_get(self._h, None, None, 0, C.byref(n)))
        boxes = (abi.Aabb * max(1, n.value))()
        vmms = (abi.VmmTheta * max(1, n.value))()
        self._check(device_lib().rtx_guiding_get(self._h, boxes, vmms, n.value, C.byref(n)))
        arr = np.array([[boxes[i].min[:], boxes[i].max[:]] for i in range(n.value)], np.float32)
        return arr, [vmms[i] for i in range(n.value)]

    def guiding_read_samples(self, width, height):
        """The last recording frame's samples: (width*height, 16) DIRECTIONAL_DTYPE."""
        out = np.zeros((width * height, abi.MAX_DIRECTIONAL_DATA_PER_PIXEL), abi.DIRECTIONAL_DTYPE)
        self._check(device_lib().rtx_guiding_read_samples(self._h, out.ctypes.data_as(C.c_void_p), out.size))
        return out

    # ------------------------------------------------ irradiance cache
    def ic_create(self, max_entries):
        self._check(device_lib().rtx_ic_create(self._h, int(max_entries)))
        self._ic_max = int(max_entries)

    def ic_reset(self):
        self._check(device_lib().rtx_ic_reset(self._h))

    def ic_read(self):
        """The device cache as an IcState."""
        st = IcState.empty(self._ic_max)
        hdr, ent, sph = st.ctypes_views()
        self._check(device_lib().rtx_ic_read(self._h, hdr, ent, sph, self._ic_max))
        return st

    def ic_write(self, state):
        hdr, ent, sph = state.ctypes_views()
        self._check(device_lib().rtx_ic_write(self._h, hdr, ent, sph))

    def guiding_options(self):
        o = abi.GuidingOptions()
        self._check(device_lib().rtx_guiding_get_options(self._h, C.byref(o)))
        return o

    def set_guiding_options(self, options=None, **kw):
        """PathGuiding's split/merge switches: an abi.GuidingOptions, or
        keyword overrides of the reference's defaults."""
        o = options if options is not None else abi.guiding_options(**kw)
        self._check(device_lib().rtx_guiding_set_options(self._h, C.byref(o)))

    def estimate_write(self, rgba, width, height):
        """The ADRRS pixel estimate (RGBA float32 per pixel) that frames with
        pc.useADRRS read (the reference's estimateImage)."""
        a = np.ascontiguousarray(rgba, np.float32).reshape(-1)
        if a.size != 4 * width * height:
            raise ValueError("estimate must hold 4 floats per pixel")
        self._check(device_lib().rtx_estimate_write(self._h, a.ctypes.data_as(C.POINTER(C.c_float)), width, height))

    def estimate_read(self, width, height):
        """The pixel estimate as (width*height, 4) float32."""
        a = np.zeros((width * height, 4), np.float32)
        self._check(device_lib().rtx_estimate_read(self._h, a.ctypes.data_as(C.POINTER(C.c_float)), width, height))
        return a

    def ic_read_jobs(self):
        """Jobs of the last kept cache frame: list of (type, target, origin,
        normal), sub-path results (jobs, 200, 4) and signatures (jobs, 200, 2)."""
        n = C.c_uint32()
        self._check(device_lib().rtx_ic_read_jobs(self._h, None, None, None, 0, C.byref(n)))
        nj = n.value
        jobs = (abi.IcJob * max(1, nj))()
        sub = np.zeros((max(1, nj) * 200, 4), np.float32)
        sig = np.zeros((max(1, nj) * 200, 2), np.uint32)
        self._check(device_lib().rtx_ic_read_jobs(self._h, C.cast(jobs, C.c_void_p),
                                                  sub.ctypes.data_as(C.POINTER(C.c_float)),
                                                  sig.ctypes.data_as(C.POINTER(C.c_uint32)), nj, C.byref(n)))
        jl = [(jobs[i].type, jobs[i].target, np.array(jobs[i].origin[:], np.float64),
               np.array(jobs[i].normal[:], np.float64)) for i in range(nj)]
        return jl, sub[:nj * 200].reshape(nj, 200, 4), sig[:nj * 200].reshape(nj, 200, 2)

    def render(self, scene, params):
        st = FrameStats()
        self._check(device_lib().rtx_render(self._h, scene._h, C.byref(params), C.byref(st)))
        return st

    def read_image(self, width, height):
        out = np.empty((height, width, 4), np.float32)
        self._check(device_lib().rtx_read_image(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), out.size))
        return out

    def read_paths(self, n):
        rgb = np.empty((n, 3), np.float32)
        sig = np.empty((n, 2), np.uint32)
        self._check(device_lib().rtx_read_paths(self._h, rgb.ctypes.data_as(C.POINTER(C.c_float)),
                                                sig.ctypes.data_as(C.POINTER(C.c_uint32)), n))
        return rgb, sig

    def trace_closest(self, scene, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hits = np.empty((n, 4), np.float32)
        self._check(device_lib().rtx_trace_closest(self._h, scene._h, rays.ctypes.data_as(C.POINTER(Ray)), n,
                                                   hits.ctypes.data_as(C.POINTER(Hit))))
        t = hits[:, 0].copy()
        prim = hits[:, 1].view(np.int32).copy()
        return t, prim, hits[:, 2].copy(), hits[:, 3].copy()

    def trace_closest_noalloc(self, scene, rays):
        """rtx_trace_closest_noalloc: the closest-hit query through the
        library's fixed device array (at most RTX_NOALLOC_MAX_RAYS rays)."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hits = np.empty((n, 4), np.float32)
        self._check(device_lib().rtx_trace_closest_noalloc(self._h, scene._h, rays.ctypes.data_as(C.POINTER(Ray)),
                                                           n, hits.ctypes.data_as(C.POINTER(Hit))))
        t = hits[:, 0].copy()
        prim = hits[:, 1].view(np.int32).copy()
        return t, prim, hits[:, 2].copy(), hits[:, 3].copy()

    def trace_occluded(self, scene, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        occ = np.empty(n, np.uint8)
        self._check(device_lib().rtx_trace_occluded(self._h, scene._h, rays.ctypes.data_as(C.POINTER(Ray)), n,
                                                    occ.ctypes.data_as(C.POINTER(C.c_uint8))))
        return occ


def job_desc(scene, params, state=None):
    """rtx_job_desc of one frame's worth of work: the frame parameters'
    push constants, camera, size, spp (sample planes), seed, sampleOffset and
    rrStartDepth; state: a Renderer whose cache / guiding regions the job
    reads (pc.useIrradianceCache / pc.useGuiding)."""
    d = abi.JobDesc()
    d.scene = scene._h
    d.state = state._h if state is not None else None
    d.pc = params.pc
    d.cam = params.cam
    d.width, d.height, d.spp = params.width, params.height, params.spp
    d.rrStartDepth = params.rrStartDepth
    d.seed = params.seed
    d.sampleOffset = params.sampleOffset
    return d


class Server:
    """rtx_server: the serving pool on a Renderer's device."""

    def __init__(self, renderer, **cfg):
        c = abi.ServerConfig()
        device_lib().rtx_server_config_defaults(C.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        self.config = c
        self._h = C.c_void_p()
        rc = device_lib().rtx_server_create(renderer._h, C.byref(c), C.byref(self._h))
        if rc != abi.RTX_OK:
            raise RtxError(rc, device_lib().rtx_last_error(renderer._h).decode())
        self._sizes = {}

    def _check(self, rc):
        if rc != abi.RTX_OK:
            raise RtxError(rc, device_lib().rtx_server_last_error(self._h).decode())

    def submit(self, scene, params, state=None):
        d = job_desc(scene, params, state)
        jid = C.c_uint64()
        self._check(device_lib().rtx_server_submit(self._h, C.byref(d), C.byref(jid)))
        self._sizes[jid.value] = (params.width, params.height)
        return jid.value

    def poll(self, jid):
        st = abi.JobStatus()
        self._check(device_lib().rtx_server_poll(self._h, jid, C.byref(st)))
        return st

    def wait(self, jid, timeout=None):
        """rtx_server_wait; with a timeout (seconds) the job is polled instead
        and TimeoutError raised when it is not done in time."""
        st = abi.JobStatus()
        if timeout is None:
            self._check(device_lib().rtx_server_wait(self._h, jid, C.byref(st)))
            return st
        import time
        t_end = time.monotonic() + timeout
        while True:
            self._check(device_lib().rtx_server_poll(self._h, jid, C.byref(st)))
            if st.state >= abi.RTX_JOB_DONE:
                return st
            if time.monotonic() > t_end:
                raise TimeoutError(f"job {jid} not done after {timeout} s (state {st.state}, "
                                   f"{st.batchesDone}/{st.batches} batches)")
            time.sleep(0.0005)

    def read_image(self, jid):
        w, h = self._sizes.pop(jid)
        out = np.empty((h, w, 4), np.float32)
        self._check(device_lib().rtx_server_read_image(self._h, jid, out.ctypes.data_as(C.POINTER(C.c_float)),
                                                        out.size))
        return out

    def stats(self):
        st = abi.ServerStats()
        self._check(device_lib().rtx_server_get_stats(self._h, C.byref(st)))
        return st

    def close(self):
        if self._h:
            device_lib().rtx_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def schedule_config(**kw):
    """rtx_schedule_config with the reference's defaults (50 prepare frames,
    6 guiding optimizations, accumulation on), fields overridden by kw."""
    c = abi.ScheduleConfig()
    device_lib().rtx_schedule_config_defaults(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Schedule:
    """The frame schedule of RayTracingApp (rtx_schedule_*): plans each
    frame's push constants and accumulation, renders it, keeps the counters."""

    def __init__(self, pc, config=None, **kw):
        self._h = C.c_void_p()
        cfg = config if config is not None else schedule_config(**kw)
        rc = device_lib().rtx_schedule_create(C.byref(cfg), C.byref(pc), C.byref(self._h))
        if rc != abi.RTX_OK:
            raise RtxError(rc, "rtx_schedule_create")

    def close(self):
        if self._h:
            device_lib().rtx_schedule_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != abi.RTX_OK:
            raise RtxError(rc, what)

    def set_push_constants(self, pc):
        self._check(device_lib().rtx_schedule_set_push_constants(self._h, C.byref(pc)), "set_push_constants")

    def input_changed(self):
        self._check(device_lib().rtx_schedule_input_changed(self._h), "input_changed")

    def begin_frame(self, base):
        out = abi.FrameParams()
        self._check(device_lib().rtx_schedule_begin_frame(self._h, C.byref(base), C.byref(out)), "begin_frame")
        return out

    def end_frame(self):
        self._check(device_lib().rtx_schedule_end_frame(self._h), "end_frame")

    def frame(self, renderer, scene, base):
        st = abi.FrameStats()
        rc = device_lib().rtx_schedule_frame(renderer._h, scene._h, self._h, C.byref(base), C.byref(st))
        renderer._check(rc)
        return st

    def state(self):
        st = abi.ScheduleState()
        self._check(device_lib().rtx_schedule_get_state(self._h, C.byref(st)), "get_state")
        return st
