"""ctypes binding of the CPU oracle (oracle/_build/liborcpt.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB = ORACLE_DIR / "_build" / "liborcpt.so"

_lib = None

# a GPU box grants 16 host cores per GPU; never use more
DEFAULT_THREADS = max(1, min(16, os.cpu_count() or 1))


def build():
    subprocess.run(["make", "-B", "-C", str(ORACLE_DIR)], check=True, stdout=subprocess.DEVNULL)


def lib(abi):
    """Load the oracle; `abi` is rtxamd.abi (for the shared struct layouts)."""
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.orc_last_error.restype = C.c_char_p
        L.orc_scene_create.argtypes = [C.POINTER(abi.SceneDesc), C.POINTER(C.c_void_p)]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_scene_destroy.restype = None
        L.orc_render_paths.argtypes = [C.c_void_p, C.POINTER(abi.FrameParams), C.POINTER(C.c_uint64), C.c_size_t,
                                       C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint32)]
        dp = C.POINTER(C.c_double)
        L.orc_trace_closest.argtypes = [C.c_void_p, C.POINTER(abi.Ray), C.c_size_t, C.c_int, dp,
                                        C.POINTER(C.c_int32), dp, dp]
        L.orc_trace_closest_prec.argtypes = [C.c_void_p, C.POINTER(abi.Ray), C.c_size_t, C.c_int, C.c_int, dp,
                                             C.POINTER(C.c_int32), dp, dp]
        L.orc_trace_occluded.argtypes = [C.c_void_p, C.POINTER(abi.Ray), C.c_size_t, C.c_int, C.POINTER(C.c_uint8)]
        L.orc_tea.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_tea.restype = C.c_uint32
        L.orc_lcg_step.argtypes = [C.c_uint32]
        L.orc_lcg_step.restype = C.c_uint32
        L.orc_pcg32_first.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_pcg32_first.restype = C.c_uint32
        d3 = C.c_double * 3
        L.orc_camera_ray.argtypes = [C.POINTER(abi.Camera), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.POINTER(C.c_uint32), d3, d3]
        L.orc_camera_ray.restype = None
        L.orc_fresnel.argtypes = [C.c_double, C.c_double]
        L.orc_fresnel.restype = C.c_double
        L.orc_fresnel_conductor.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_fresnel_conductor.restype = C.c_double
        L.orc_beckmann_d.argtypes = [C.POINTER(abi.Material), d3, d3]
        L.orc_beckmann_d.restype = C.c_double
        L.orc_pdf_bsdf.argtypes = [C.POINTER(abi.Material), d3, d3, d3]
        L.orc_pdf_bsdf.restype = C.c_double
        L.orc_eval_bsdf.argtypes = [C.POINTER(abi.Material), d3, d3, d3, C.c_int, d3]
        L.orc_eval_bsdf.restype = None
        L.orc_sample_bsdf.argtypes = [C.POINTER(abi.Material), d3, d3, C.c_int, C.POINTER(C.c_uint32), d3]
        L.orc_sample_bsdf.restype = C.c_double
        L.orc_sample_cosine.argtypes = [d3, C.POINTER(C.c_uint32), d3]
        L.orc_sample_cosine.restype = None
        L.orc_sample_cosine_power.argtypes = [d3, C.c_double, C.POINTER(C.c_uint32), d3]
        L.orc_sample_cosine_power.restype = None
        L.orc_sample_beckmann_normal.argtypes = [C.POINTER(abi.Material), d3, C.POINTER(C.c_uint32), d3]
        L.orc_sample_beckmann_normal.restype = None
        L.orc_sample_unit_sphere.argtypes = [C.POINTER(C.c_uint32), d3]
        L.orc_sample_unit_sphere.restype = None
        L.orc_sample_lights.argtypes = [C.c_void_p, d3, d3, C.POINTER(C.c_uint32), d3, d3, C.POINTER(C.c_double)]
        L.orc_sample_lights.restype = C.c_double
        _lib = L
    return _lib


def d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


class OracleScene:
    def __init__(self, abi, scene_desc):
        self.abi = abi
        self.L = lib(abi)
        self._h = C.c_void_p()
        rc = self.L.orc_scene_create(C.byref(scene_desc), C.byref(self._h))
        if rc != 0:
            raise RuntimeError(f"oracle scene: {self.L.orc_last_error().decode()}")

    def __del__(self):
        try:
            if self._h:
                self.L.orc_scene_destroy(self._h)
        except Exception:
            pass

    def render_paths(self, params, ids, prec=64, threads=None):
        threads = DEFAULT_THREADS if threads is None else threads
        ids = np.ascontiguousarray(ids, np.uint64)
        n = ids.size
        rgb = np.empty((n, 3), np.float64)
        sig = np.empty((n, 2), np.uint32)
        rc = self.L.orc_render_paths(self._h, C.byref(params), ids.ctypes.data_as(C.POINTER(C.c_uint64)), n, prec,
                                     threads, rgb.ctypes.data_as(C.POINTER(C.c_double)),
                                     sig.ctypes.data_as(C.POINTER(C.c_uint32)))
        if rc != 0:
            raise RuntimeError(f"oracle render: {self.L.orc_last_error().decode()}")
        return rgb, sig

    def trace_closest(self, rays, brute=False, prec=64):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        t = np.empty(n, np.float64)
        prim = np.empty(n, np.int32)
        u = np.empty(n, np.float64)
        v = np.empty(n, np.float64)
        dp = C.POINTER(C.c_double)
        rc = self.L.orc_trace_closest_prec(self._h, rays.ctypes.data_as(C.POINTER(self.abi.Ray)), n, int(brute),
                                      prec, t.ctypes.data_as(dp), prim.ctypes.data_as(C.POINTER(C.c_int32)),
                                      u.ctypes.data_as(dp), v.ctypes.data_as(dp))
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return t, prim, u, v

    def trace_occluded(self, rays, brute=False):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        occ = np.empty(n, np.uint8)
        rc = self.L.orc_trace_occluded(self._h, rays.ctypes.data_as(C.POINTER(self.abi.Ray)), n, int(brute),
                                       occ.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return occ
