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
    def load_xml(cls, path):
        h = C.c_void_p()
        _host_check(host_lib().rtxh_scene_load_xml(str(path).encode(), C.byref(h)))
        return cls(h)

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


class DeviceScene:
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

    def trace_occluded(self, scene, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        occ = np.empty(n, np.uint8)
        self._check(device_lib().rtx_trace_occluded(self._h, scene._h, rays.ctypes.data_as(C.POINTER(Ray)), n,
                                                    occ.ctypes.data_as(C.POINTER(C.c_uint8))))
        return occ
