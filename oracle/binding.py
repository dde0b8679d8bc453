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
        L.orc_microfacet_d.argtypes = [C.POINTER(abi.Material), d3, d3]
        L.orc_microfacet_d.restype = C.c_double
        L.orc_microfacet_g1.argtypes = [C.POINTER(abi.Material), d3, d3, d3]
        L.orc_microfacet_g1.restype = C.c_double
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
        L.orc_tex_sample.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double)]
        L.orc_env_color.argtypes = [C.c_void_p, d3, C.POINTER(C.c_double)]
        L.orc_env_color.restype = None
        vp = C.c_void_p
        L.orc_ic_frame.argtypes = [C.c_void_p, C.POINTER(abi.FrameParams), C.c_int, C.c_int, vp, vp, vp,
                                   C.POINTER(C.c_double), C.POINTER(C.c_uint32), vp, C.c_size_t,
                                   C.POINTER(C.c_size_t), C.POINTER(C.c_double), C.POINTER(C.c_uint32)]
        L.orc_ic_reduce.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, d3, C.POINTER(C.c_double), d3, d3,
                                    d3, C.POINTER(C.c_double)]
        L.orc_ic_stratum.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.POINTER(C.c_uint32)]
        L.orc_ic_stratum.restype = None
        L.orc_ic_query.argtypes = [C.c_void_p, C.POINTER(abi.PushConstant), vp, vp, vp, d3, d3, d3]
        L.orc_ic_apply.argtypes = [C.POINTER(abi.FrameParams), C.c_int, vp, vp, vp, vp, C.c_size_t,
                                   C.POINTER(C.c_double)]
        L.orc_render_paths_record.argtypes = [C.c_void_p, C.POINTER(abi.FrameParams), C.POINTER(C.c_uint64),
                                              C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_double),
                                              C.POINTER(C.c_uint32), C.c_void_p]
        L.orc_guiding_set.argtypes = [C.c_void_p, C.POINTER(abi.Aabb), C.POINTER(abi.VmmTheta), C.c_uint32]
        L.orc_vmm_pdf.argtypes = [C.POINTER(abi.VmmTheta), d3, d3, C.c_int]
        L.orc_vmm_pdf.restype = C.c_double
        L.orc_sample_vmm.argtypes = [C.POINTER(abi.VmmTheta), d3, C.c_int, C.POINTER(C.c_uint32), d3]
        L.orc_sample_vmm.restype = None
        L.orc_guide_create.argtypes = [C.POINTER(abi.Aabb), C.c_uint32, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.orc_guide_destroy.argtypes = [C.c_void_p]
        L.orc_guide_destroy.restype = None
        L.orc_guide_regions.argtypes = [C.c_void_p, C.POINTER(abi.Aabb), C.POINTER(abi.VmmTheta), C.c_uint32]
        L.orc_guide_regions.restype = C.c_uint32
        L.orc_guide_regions_wide.argtypes = [C.c_void_p, C.POINTER(abi.Aabb), C.POINTER(abi.VmmThetaWide), C.c_uint32]
        L.orc_guide_regions_wide.restype = C.c_uint32
        L.orc_guide_update.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.orc_guide_last_error.restype = C.c_char_p
        L.orc_guide_set_options.argtypes = [C.c_void_p, C.POINTER(abi.GuidingOptions)]
        d8 = C.c_double * 8
        L.orc_vmf_combine.argtypes = [d8, d8, C.c_int, d8]
        L.orc_vmf_combine.restype = None
        L.orc_vmf_merge.argtypes = [d8, d8, d8]
        L.orc_vmf_merge.restype = None
        L.orc_merge_metric.argtypes = [d8, d8]
        L.orc_merge_metric.restype = C.c_double
        L.orc_eigen2x2.argtypes = [C.c_double * 4, C.c_double * 2, C.c_double * 4]
        L.orc_eigen2x2.restype = None
        L.orc_fastexp.argtypes = [C.c_float]
        L.orc_fastexp.restype = C.c_float
        dp = C.POINTER(C.c_double)
        L.orc_em_mix_doubles.argtypes = []
        L.orc_em_mix_doubles.restype = C.c_int
        L.orc_em_estep.argtypes = [dp, dp, dp, C.c_size_t, dp, dp, dp, dp]
        L.orc_em_estep.restype = C.c_int
        L.orc_em_mstep.argtypes = [dp, dp, dp, dp]
        L.orc_em_mstep.restype = C.c_int
        L.orc_em_fit.argtypes = [dp, dp, dp, dp, C.c_size_t, C.c_int]
        L.orc_em_fit.restype = C.c_int
        L.orc_partition_order.argtypes = [dp, C.c_size_t, C.c_double, C.POINTER(C.c_uint32)]
        L.orc_partition_order.restype = C.c_size_t
        L.orc_ic_calculate.argtypes = [C.c_void_p, C.POINTER(abi.FrameParams), C.c_int, vp, vp, vp, C.c_uint32, d3,
                                       d3, d3, d3, d3, C.POINTER(C.c_double)]
        L.orc_estimate_set.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_uint32, C.c_uint32]
        L.orc_weight_window.argtypes = [d3, d3, d3, C.c_double, C.POINTER(C.c_uint32), C.POINTER(C.c_int)]
        L.orc_weight_window.restype = C.c_double
        L.orc_adrrs_counters.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_adrrs_counters.restype = None
        _lib = L
    return _lib


class IcJob(C.Structure):
    _fields_ = [("type", C.c_int32), ("target", C.c_uint32), ("origin", C.c_double * 3),
                ("normal", C.c_double * 3)]


IC_STRATA = 200  # 20 phi x 10 theta strata of calculateCacheData


def weight_window(abi, estimate, adjoint, throughput, adrrs_s, state):
    """applyWeightWindow in FP64: (q, n, state after)."""
    L = lib(abi)
    st = C.c_uint32(state)
    n = C.c_int(0)
    q = L.orc_weight_window(d3(estimate), d3(adjoint), d3(throughput), adrrs_s, C.byref(st), C.byref(n))
    return q, n.value, st.value


def adrrs_counters(abi):
    """(splits queued, roulette deaths, split paths traced) since the last call."""
    out = (C.c_uint64 * 3)()
    lib(abi).orc_adrrs_counters(out)
    return tuple(int(x) for x in out)


def vmf_lobe(w, mu, kappa):
    """A lobe as the oracle's test hooks take it (8 doubles), normalised as
    lightpmm's calNormalization does (with exact exp here)."""
    mu = np.asarray(mu, np.float64)
    mu = mu / np.linalg.norm(mu)
    e2k = np.exp(-2 * kappa)
    norm = kappa / (2 * np.pi * (1 - e2k)) if kappa > 0 else 1 / (4 * np.pi)
    r = 1 / np.tanh(kappa) - 1 / kappa if kappa > 0 else 0.0
    return (C.c_double * 8)(w, *mu, kappa, r, norm, e2k)


def vmf_combine(abi, a, b, divide):
    out = (C.c_double * 8)()
    lib(abi).orc_vmf_combine(a, b, int(divide), out)
    return np.array(out[:])


def vmf_merge(abi, a, b):
    out = (C.c_double * 8)()
    lib(abi).orc_vmf_merge(a, b, out)
    return np.array(out[:])


def merge_metric(abi, a, b):
    return lib(abi).orc_merge_metric(a, b)


def eigen2x2(abi, m):
    """(eigenvalues, eigenvector matrix with the vectors as columns) in Eigen's order."""
    lam = (C.c_double * 2)()
    vec = (C.c_double * 4)()
    lib(abi).orc_eigen2x2((C.c_double * 4)(*[float(x) for x in np.asarray(m).reshape(-1)]), lam, vec)
    return np.array(lam[:]), np.array(vec[:]).reshape(2, 2)


def ic_reduce(abi, seed, sample, job, normal, sub, prec=64):
    """calculateCacheData's accumulation over one job's sub-path results
    sub: (200, 4) = L rgb, first-hit distance.  Returns color, rot, trans, harmonicR."""
    L = lib(abi)
    sub = np.ascontiguousarray(sub, np.float64)
    c, r, t = (C.c_double * 3)(), (C.c_double * 3)(), (C.c_double * 3)()
    h = C.c_double()
    L.orc_ic_reduce(prec, seed, sample, job, d3(normal), sub.ctypes.data_as(C.POINTER(C.c_double)), c, r, t,
                    C.byref(h))
    return np.array(c[:]), np.array(r[:]), np.array(t[:]), h.value


def ic_apply(abi, params, state, jobs, sub, prec=32):
    """Accumulate and commit `jobs` [(type, target, origin, normal)] with their
    sub-path results sub (J, 200, 4) into `state` (in place)."""
    L = lib(abi)
    arr = (IcJob * max(1, len(jobs)))()
    for i, (t, tg, o, n) in enumerate(jobs):
        arr[i].type, arr[i].target = int(t), int(tg)
        arr[i].origin[:] = [float(x) for x in o]
        arr[i].normal[:] = [float(x) for x in n]
    sub = np.ascontiguousarray(sub, np.float64).reshape(-1, 4)
    hdr, ent, sph = state.ctypes_views()
    rc = L.orc_ic_apply(C.byref(params), prec, hdr, ent, sph, C.cast(arr, C.c_void_p), len(jobs),
                        sub.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise RuntimeError(L.orc_last_error().decode())


def ic_stratum(abi, seed, sample, job, k, j):
    L = lib(abi)
    phi, theta, st = C.c_double(), C.c_double(), C.c_uint32()
    L.orc_ic_stratum(seed, sample, job, k, j, C.byref(phi), C.byref(theta), C.byref(st))
    return phi.value, theta.value, st.value


def ic_query(abi, pc, state, x, n, scene=None):
    """Cache lookup at x with normal n against `state` (rtxamd.IcState); None on a miss."""
    L = lib(abi)
    hdr, ent, sph = state.ctypes_views()
    out = (C.c_double * 3)()
    hit = L.orc_ic_query(scene._h if scene is not None else None, C.byref(pc), hdr, ent, sph, d3(x), d3(n), out)
    return np.array(out[:]) if hit else None


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

    def ic_frame(self, params, state, prec=64, threads=None, job_cap=None):
        """One frame of the irradiance-cache method over every pixel (sample
        params.sampleOffset).  Updates `state` (rtxamd.IcState) in place and
        returns (rgb, sig, jobs, sub, subsig): per pixel, then the frame's
        jobs with their (200, 4) sub-path results and (200, 2) signatures."""
        threads = DEFAULT_THREADS if threads is None else threads
        npix = params.width * params.height
        rgb = np.empty((npix, 3), np.float64)
        sig = np.empty((npix, 2), np.uint32)
        cap = state.max_entries + npix if job_cap is None else job_cap
        jobs = (IcJob * max(1, cap))()
        sub = np.zeros((max(1, cap) * IC_STRATA, 4), np.float64)
        subsig = np.zeros((max(1, cap) * IC_STRATA, 2), np.uint32)
        nj = C.c_size_t()
        hdr, ent, sph = state.ctypes_views()
        rc = self.L.orc_ic_frame(self._h, C.byref(params), prec, threads, hdr, ent, sph,
                                 rgb.ctypes.data_as(C.POINTER(C.c_double)), sig.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 C.cast(jobs, C.c_void_p), cap, C.byref(nj),
                                 sub.ctypes.data_as(C.POINTER(C.c_double)), subsig.ctypes.data_as(C.POINTER(C.c_uint32)))
        if rc != 0:
            raise RuntimeError(f"oracle ic frame: {self.L.orc_last_error().decode()}")
        n = min(nj.value, cap)
        jl = [(jobs[i].type, jobs[i].target, np.array(jobs[i].origin[:]), np.array(jobs[i].normal[:]))
              for i in range(n)]
        return rgb, sig, jl, sub[:n * IC_STRATA].reshape(n, IC_STRATA, 4), subsig[:n * IC_STRATA].reshape(n, IC_STRATA, 2)

    def tex_sample(self, slot, u, v):
        out = (C.c_double * 4)()
        if self.L.orc_tex_sample(self._h, int(slot), float(u), float(v), out) != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return np.array(out[:])

    def env_color(self, direction):
        out = (C.c_double * 3)()
        self.L.orc_env_color(self._h, (C.c_double * 3)(*[float(x) for x in direction]), out)
        return np.array(out[:])

    def estimate_set(self, rgba, width, height):
        """The ADRRS pixel estimate (RGBA float32, width*height) of later frames."""
        self._estimate = np.ascontiguousarray(rgba, np.float32).reshape(-1)
        assert self._estimate.size == 4 * width * height
        rc = self.L.orc_estimate_set(self._h, self._estimate.ctypes.data_as(C.POINTER(C.c_float)), width, height)
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error().decode())

    def guiding_set(self, regions, vmms):
        from rtxamd import guiding_arrays
        boxes, arr = guiding_arrays(regions, vmms)
        rc = self.L.orc_guiding_set(self._h, boxes, arr, len(vmms))
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error().decode())

    def ic_calculate(self, params, state, origin, normal, job=0, prec=64):
        """calculateCacheData at one point: color, rot, trans, harmonicR."""
        hdr, ent, sph = state.ctypes_views()
        c, r, t = (C.c_double * 3)(), (C.c_double * 3)(), (C.c_double * 3)()
        h = C.c_double()
        rc = self.L.orc_ic_calculate(self._h, C.byref(params), prec, hdr, ent, sph, job, d3(origin), d3(normal), c, r,
                                     t, C.byref(h))
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return np.array(c[:]), np.array(r[:]), np.array(t[:]), h.value

    def render_paths_record(self, params, ids, prec=64, threads=None):
        """render_paths with pc.updateGuiding: also each path's 16 DirectionalData slots."""
        threads = DEFAULT_THREADS if threads is None else threads
        ids = np.ascontiguousarray(ids, np.uint64)
        n = ids.size
        rgb = np.empty((n, 3), np.float64)
        sig = np.empty((n, 2), np.uint32)
        dd = np.zeros((n, self.abi.MAX_DIRECTIONAL_DATA_PER_PIXEL), self.abi.DIRECTIONAL_DTYPE)
        rc = self.L.orc_render_paths_record(self._h, C.byref(params), ids.ctypes.data_as(C.POINTER(C.c_uint64)), n,
                                            prec, threads, rgb.ctypes.data_as(C.POINTER(C.c_double)),
                                            sig.ctypes.data_as(C.POINTER(C.c_uint32)), dd.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle render: {self.L.orc_last_error().decode()}")
        return rgb, sig, dd

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


class OracleGuide:
    """PathGuiding's regions and mixtures, refit with the oracle's EM."""

    def __init__(self, abi, scene_box, splits, parallax=True, prec=64, options=None):
        """options: abi.GuidingOptions (default: the reference's, split/merge on)."""
        self.abi = abi
        self.L = lib(abi)
        box = abi.Aabb()
        box.min[:] = [float(x) for x in scene_box[0]]
        box.max[:] = [float(x) for x in scene_box[1]]
        self._h = C.c_void_p()
        rc = self.L.orc_guide_create(C.byref(box), splits, int(parallax), prec, C.byref(self._h))
        if rc != 0:
            raise RuntimeError(self.L.orc_guide_last_error().decode())
        if options is not None:
            self.set_options(options)

    def set_options(self, options):
        """Replace the split/merge and scale options (abi.GuidingOptions)."""
        rc = self.L.orc_guide_set_options(self._h, C.byref(options))
        if rc != 0:
            raise RuntimeError(self.L.orc_guide_last_error().decode())

    def __del__(self):
        try:
            if self._h:
                self.L.orc_guide_destroy(self._h)
        except Exception:
            pass

    def regions(self):
        n = self.L.orc_guide_regions(self._h, None, None, 0)
        boxes = (self.abi.Aabb * n)()
        vmms = (self.abi.VmmTheta * n)()
        self.L.orc_guide_regions(self._h, boxes, vmms, n)
        arr = np.array([[b.min[:], b.max[:]] for b in boxes], np.float32)
        return arr, [vmms[i] for i in range(n)]

    def regions_wide(self):
        """(boxes, [abi.VmmThetaWide]): mixtures of up to 32 lobes."""
        n = self.L.orc_guide_regions_wide(self._h, None, None, 0)
        boxes = (self.abi.Aabb * n)()
        vmms = (self.abi.VmmThetaWide * n)()
        self.L.orc_guide_regions_wide(self._h, boxes, vmms, n)
        arr = np.array([[b.min[:], b.max[:]] for b in boxes], np.float32)
        return arr, [vmms[i] for i in range(n)]

    def update(self, samples):
        samples = np.ascontiguousarray(samples, self.abi.DIRECTIONAL_DTYPE).reshape(-1)
        rc = self.L.orc_guide_update(self._h, samples.ctypes.data_as(C.c_void_p), samples.size)
        if rc != 0:
            raise RuntimeError(self.L.orc_guide_last_error().decode())
