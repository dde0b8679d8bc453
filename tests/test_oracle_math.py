"""Closed forms and distribution tests pinning the oracle's shading layer
(shaders/raytrace.rgen:144-599, shaders/random.glsl:59-136)."""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

import rtxamd
from rtxamd import abi
from oracle.binding import d3, lib

L = lib(abi)


def mat(kind, **kw):
    if kind == "ggx":
        return rtxamd.rough_conductor(distribution=abi.DISTRIBUTION_GGX, **kw)
    return {"diffuse": rtxamd.diffuse, "phong": rtxamd.phong, "rough": rtxamd.rough_conductor}[kind](**kw)


# ------------------------------------------------------------------ Fresnel
def hecht_dielectric(n1, n2, cos_i):
    """Unpolarised Fresnel reflectance from the Fresnel equations (Hecht, Optics)."""
    sin_t = n1 / n2 * np.sqrt(max(0.0, 1 - cos_i * cos_i))
    if sin_t > 1:
        return 1.0
    cos_t = np.sqrt(1 - sin_t * sin_t)
    rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)
    rp = (n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t)
    return 0.5 * (rs * rs + rp * rp)


@pytest.mark.parametrize("n1,n2", [(1.0, 1.5), (1.5, 1.0), (1.0, 2.4)])
def test_fresnel_dielectric_matches_fresnel_equations(n1, n2):
    for c in np.linspace(0.01, 1.0, 37):
        assert L.orc_fresnel(n1 / n2, c) == pytest.approx(hecht_dielectric(n1, n2, c), rel=1e-12, abs=1e-15)
    # normal incidence ((n1-n2)/(n1+n2))^2, total internal reflection beyond the critical angle
    assert L.orc_fresnel(1 / 1.5, 1.0) == pytest.approx(0.04, rel=1e-12)
    if n1 > n2:
        crit = np.sqrt(1 - (n2 / n1) ** 2)
        assert L.orc_fresnel(n1 / n2, crit * 0.99) == 1.0


def test_fresnel_conductor_normal_incidence():
    for eta, k in [(0.2, 3.0), (1.3, 7.5), (0.37, 2.82)]:
        want = ((eta - 1) ** 2 + k * k) / ((eta + 1) ** 2 + k * k)
        assert L.orc_fresnel_conductor(1.0, eta, k) == pytest.approx(want, rel=1e-12)
        assert L.orc_fresnel_conductor(-1.0, eta, k) == pytest.approx(want, rel=1e-12)  # |cos| symmetric
        assert L.orc_fresnel_conductor(0.0, eta, k) == pytest.approx(1.0, rel=1e-12)  # grazing


# ------------------------------------------------------------- Beckmann D
@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5])
def test_beckmann_projected_area_normalisation(alpha):
    """int D(m) (m.n) dw_m = 1 over the hemisphere (microfacet normalisation)."""
    m = mat("rough", alpha=alpha, eta=1.0, k=1.0)
    th = np.linspace(0, np.pi / 2, 20001)[:-1]
    vals = np.array([L.orc_beckmann_d(C.byref(m), d3((0, 0, 1)), d3((np.sin(t), 0, np.cos(t)))) for t in th])
    integral = 2 * np.pi * np.trapezoid(vals * np.cos(th) * np.sin(th), th)
    assert integral == pytest.approx(1.0, rel=2e-3)


def test_beckmann_closed_form():
    m = mat("rough", alpha=0.3, eta=1.0, k=1.0)
    a2 = float(np.float32(0.3)) ** 2  # roughness is stored in FP32
    for t in [0.0, 0.3, 0.9]:
        want = np.exp(-np.tan(t) ** 2 / a2) / (np.pi * a2 * np.cos(t) ** 4)
        got = L.orc_beckmann_d(C.byref(m), d3((0, 0, 1)), d3((np.sin(t), 0, np.cos(t))))
        assert got == pytest.approx(want, rel=1e-10)
    assert L.orc_beckmann_d(C.byref(m), d3((0, 0, 1)), d3((1, 0, -0.1))) == 0.0


# ------------------------------------------------------------------- GGX
def _sph(t, p=0.0):
    return (np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t))


def _hemisphere_integral(f, alpha):
    """int over the upper hemisphere of f(theta, phi) dw, with the theta
    quadrature concentrated where a lobe of width alpha lives."""
    from scipy import integrate
    pts = [min(x, np.pi / 2) for x in (alpha, 3 * alpha, 10 * alpha)]

    def inner(t):
        return integrate.quad(lambda p: f(t, p), 0, 2 * np.pi, limit=200)[0] * np.sin(t)
    return integrate.quad(inner, 0, np.pi / 2, points=pts, limit=400, epsabs=1e-10)[0]


@pytest.mark.parametrize("alpha", [0.005, 0.05, 0.2, 0.5])
def test_ggx_projected_area_normalisation(alpha):
    """int D(m) (m.n) dw_m = 1 (Walter et al. 2007, eq. 4) for the GGX D."""
    from scipy import integrate
    m = mat("ggx", alpha=alpha, eta=1.0, k=1.0)
    pts = [min(x, np.pi / 2) for x in (alpha, 3 * alpha, 10 * alpha)]
    val = integrate.quad(lambda t: 2 * np.pi * L.orc_microfacet_d(C.byref(m), d3((0, 0, 1)), d3(_sph(t)))
                         * np.cos(t) * np.sin(t), 0, np.pi / 2, points=pts, limit=400, epsabs=1e-12)[0]
    assert val == pytest.approx(1.0, rel=1e-6)


def test_ggx_closed_form():
    """D = a^2 / (pi cos^4 (a^2 + tan^2)^2) (Walter et al. 2007, eq. 33)."""
    m = mat("ggx", alpha=0.3, eta=1.0, k=1.0)
    a2 = float(np.float32(0.3)) ** 2  # roughness is stored in FP32
    for t in [0.0, 0.3, 0.9, 1.4]:
        want = a2 / (np.pi * np.cos(t) ** 4 * (a2 + np.tan(t) ** 2) ** 2)
        got = L.orc_microfacet_d(C.byref(m), d3((0, 0, 1)), d3(_sph(t)))
        assert got == pytest.approx(want, rel=1e-10)
    assert L.orc_microfacet_d(C.byref(m), d3((0, 0, 1)), d3((1, 0, -0.1))) == 0.0


@pytest.mark.parametrize("alpha,theta_v", [(0.1, 0.3), (0.3, 1.0), (0.5, 1.35), (0.8, 0.7)])
def test_ggx_smith_g1_projected_area(alpha, theta_v):
    """The visible microsurface projects onto the geometric one:
    int G1(v, m) max(0, v.m) D(m) dw_m = v.n (Heitz 2014, eq. 7), which fixes
    the Smith G1 of a given D.  G1 is constant over the facets that face v,
    so G1(v) = (v.n) / int max(0, v.m) D(m) dw_m."""
    m = mat("ggx", alpha=alpha, eta=1.0, k=1.0)
    v = np.array(_sph(theta_v))
    n = d3((0, 0, 1))

    def f(t, p):
        h = _sph(t, p)
        vm = float(np.dot(v, h))
        if vm <= 0:
            return 0.0
        return L.orc_microfacet_g1(C.byref(m), n, d3(h), d3(v)) * vm * L.orc_microfacet_d(C.byref(m), n, d3(h))
    assert _hemisphere_integral(f, alpha) == pytest.approx(np.cos(theta_v), rel=2e-5)
    # facets facing away from v are invisible
    assert L.orc_microfacet_g1(C.byref(m), n, d3(_sph(0.2, np.pi)), d3(_sph(1.45, 0.0))) == 0.0


def test_ggx_bsdf_is_reciprocal_and_below_horizon_zero():
    """f(wi, wo) = F G D / (4 |wi.n| |wo.n|) is symmetric in wi and wo; the
    evaluated value is f cos(theta_o).  A direction below the horizon gets 0
    (its G1 vanishes), unlike the reference's Beckmann G1."""
    m = mat("ggx", alpha=0.35, eta=0.4, k=3.0)
    n = (0, 0, 1)
    rng = np.random.default_rng(1)
    out1, out2 = (C.c_double * 3)(), (C.c_double * 3)()
    for _ in range(50):
        a = np.array(_sph(rng.uniform(0, 1.4), rng.uniform(0, 2 * np.pi)))
        b = np.array(_sph(rng.uniform(0, 1.4), rng.uniform(0, 2 * np.pi)))
        L.orc_eval_bsdf(C.byref(m), d3(n), d3(a), d3(b), 1, out1)
        L.orc_eval_bsdf(C.byref(m), d3(n), d3(b), d3(a), 1, out2)
        assert out1[0] / b[2] == pytest.approx(out2[0] / a[2], rel=1e-10)
        assert out1[0] > 0
    L.orc_eval_bsdf(C.byref(m), d3(n), d3(_sph(0.3)), d3((0.8, 0.0, -0.6)), 1, out1)
    assert out1[0] == 0.0


def test_ggx_normal_sampler_matches_D_cos():
    """The GGX normal sampler draws m with density D(m) (m.n)."""
    m = mat("ggx", alpha=0.2, eta=1.0, k=1.0)
    st = C.c_uint32(3)
    out = (C.c_double * 3)()
    ms = np.empty((30000, 3))
    for k in range(len(ms)):
        L.orc_sample_beckmann_normal(C.byref(m), d3((0, 0, 1)), C.byref(st), out)
        ms[k] = out[:]
    assert st.value == _lcg_n(3, 2 * len(ms))  # two draws per normal

    def pdf(w):
        return L.orc_microfacet_d(C.byref(m), d3((0, 0, 1)), d3(w)) * max(w[2], 0)

    ct, ph = sphere_grid(24, 8, hemisphere=True)
    ct = 1 - (1 - ct) ** 3  # concentrate bins near the pole
    exp = expected_counts(pdf, ct, ph, len(ms))
    assert chi2_pvalue(ms, exp, ct, ph) > 1e-3


def _lcg_n(state, n):
    for _ in range(n):
        state = L.orc_lcg_step(state)
    return state


# ------------------------------------------------------- sampling vs pdfs
def sphere_grid(nt=24, np_=24, hemisphere=False):
    ct = np.linspace(0 if hemisphere else -1, 1, nt + 1)
    ph = np.linspace(0, 2 * np.pi, np_ + 1)
    return ct, ph


def expected_counts(pdf, ct, ph, n, sub=6):
    exp = np.zeros((len(ct) - 1, len(ph) - 1))
    for i in range(len(ct) - 1):
        cs = np.linspace(ct[i], ct[i + 1], sub + 1)
        cs = 0.5 * (cs[1:] + cs[:-1])
        for j in range(len(ph) - 1):
            ps = np.linspace(ph[j], ph[j + 1], sub + 1)
            ps = 0.5 * (ps[1:] + ps[:-1])
            acc = 0.0
            for c in cs:
                s = np.sqrt(max(0, 1 - c * c))
                for p in ps:
                    acc += pdf(np.array([s * np.cos(p), s * np.sin(p), c]))
            exp[i, j] = acc / (sub * sub) * (ct[i + 1] - ct[i]) * (ph[j + 1] - ph[j]) * n
    return exp


def chi2_pvalue(dirs, exp, ct, ph):
    c = np.clip(dirs[:, 2], -1, 1)
    p = np.mod(np.arctan2(dirs[:, 1], dirs[:, 0]), 2 * np.pi)
    obs, _, _ = np.histogram2d(c, p, bins=[ct, ph])
    mask = exp > 5
    # pool the sparse bins into one
    o = np.append(obs[mask], obs[~mask].sum())
    e = np.append(exp[mask], exp[~mask].sum())
    if e[-1] <= 5:
        o, e = o[:-1], e[:-1]
    e = e * o.sum() / e.sum()
    return stats.chisquare(o, e).pvalue


def draw(m, wi, n, count, seed=7):
    st = C.c_uint32(seed)
    out = np.empty((count, 3))
    pdfs = np.empty(count)
    w = (C.c_double * 3)()
    for k in range(count):
        pdfs[k] = L.orc_sample_bsdf(C.byref(m), d3(wi), d3(n), 1, C.byref(st), w)
        out[k] = w[:]
    return out, pdfs


@pytest.mark.parametrize("kind,kw", [
    ("diffuse", dict(albedo=(0.5, 0.5, 0.5))),
    ("phong", dict(kd=(0.3, 0.3, 0.3), ks=(0.5, 0.5, 0.5), exponent=20.0)),
    ("rough", dict(alpha=0.3, eta=0.4, k=3.0)),
    ("ggx", dict(alpha=0.3, eta=0.4, k=3.0)),
    ("ggx", dict(alpha=0.08, eta=0.4, k=3.0)),
])
def test_sampleBSDF_distribution_matches_pdfBSDF(kind, kw):
    m = mat(kind, **kw)
    n = np.array([0.0, 0.0, 1.0])
    wi = np.array([0.5, 0.0, 0.8660254037844386])
    N = 40000
    dirs, pdfs = draw(m, wi, n, N)
    valid = pdfs > 0
    # the returned pdf equals pdfBSDF of the returned direction
    for k in np.flatnonzero(valid)[:200]:
        assert pdfs[k] == pytest.approx(L.orc_pdf_bsdf(C.byref(m), d3(n), d3(wi), d3(dirs[k])), rel=1e-9)

    def pdf(w):
        return max(0.0, L.orc_pdf_bsdf(C.byref(m), d3(n), d3(wi), d3(w)))

    ct, ph = sphere_grid(16, 16, hemisphere=(kind == "diffuse"))
    exp = expected_counts(pdf, ct, ph, valid.sum())
    assert chi2_pvalue(dirs[valid], exp, ct, ph) > 1e-3


def test_cosine_and_sphere_sampler_moments():
    st = C.c_uint32(99)
    out = (C.c_double * 3)()
    zs, us = [], []
    for _ in range(20000):
        L.orc_sample_cosine(d3((0, 0, 1)), C.byref(st), out)
        zs.append(out[2])
        L.orc_sample_unit_sphere(C.byref(st), out)
        us.append(out[:])
    zs, us = np.array(zs), np.array(us)
    assert zs.min() >= 0 and zs.mean() == pytest.approx(2 / 3, abs=0.01)  # E[cos] under cos/pi
    assert np.allclose(np.linalg.norm(us, axis=1), 1, atol=1e-12)
    assert np.abs(us.mean(0)).max() < 0.02 and np.allclose((us ** 2).mean(0), 1 / 3, atol=0.01)


def test_cosine_power_lobe_moment():
    """E[cos^1] of the (p+1) cos^p / (2 pi) lobe is (p+1)/(p+2)."""
    st = C.c_uint32(5)
    out = (C.c_double * 3)()
    p = 10.0
    cs = []
    for _ in range(20000):
        L.orc_sample_cosine_power(d3((0, 0, 1)), p, C.byref(st), out)
        cs.append(out[2])
    assert np.mean(cs) == pytest.approx((p + 1) / (p + 2), abs=0.005)


def test_beckmann_normal_sampler_matches_D_cos():
    """randomBeckmannNormal samples m with density D(m) (m.n)."""
    m = mat("rough", alpha=0.25, eta=1.0, k=1.0)
    st = C.c_uint32(3)
    out = (C.c_double * 3)()
    ms = np.empty((30000, 3))
    for k in range(len(ms)):
        L.orc_sample_beckmann_normal(C.byref(m), d3((0, 0, 1)), C.byref(st), out)
        ms[k] = out[:]

    def pdf(w):
        return L.orc_beckmann_d(C.byref(m), d3((0, 0, 1)), d3(w)) * max(w[2], 0)

    ct, ph = sphere_grid(24, 8, hemisphere=True)
    ct = 1 - (1 - ct) ** 3  # concentrate bins near the pole
    exp = expected_counts(pdf, ct, ph, len(ms))
    assert chi2_pvalue(ms, exp, ct, ph) > 1e-3


def test_pdf_normalisation():
    n = (0, 0, 1)
    wi = (0.3, 0.1, np.sqrt(1 - 0.1))
    ct = np.linspace(-1, 1, 801)
    ph = np.linspace(0, 2 * np.pi, 401)
    for m, want, hemi in [(mat("diffuse", albedo=(1, 1, 1)), 1.0, True),
                          (mat("rough", alpha=0.3, eta=1.0, k=1.0), 1.0, False),
                          (mat("ggx", alpha=0.3, eta=1.0, k=1.0), 1.0, False)]:
        tot = 0.0
        for c in 0.5 * (ct[1:] + ct[:-1]):
            if hemi and c < 0:
                continue
            s = np.sqrt(1 - c * c)
            for p in 0.5 * (ph[1:] + ph[:-1])[::4]:
                tot += max(0.0, L.orc_pdf_bsdf(C.byref(m), d3(n), d3(wi), d3((s * np.cos(p), s * np.sin(p), c))))
        tot *= (ct[1] - ct[0]) * (ph[1] - ph[0]) * 4
        assert tot == pytest.approx(want, rel=0.02)


# ------------------------------------------------------------------ camera
def test_camera_ray_closed_form():
    """getCameraRay (raytrace.rgen:1487-1494) with the host's matrices equals
    forward + right*x*aspect*tan(fov/2) - up*y*tan(fov/2)."""
    view = rtxamd.View()
    view.origin[:] = [1.0, -3.0, 2.0]
    view.target[:] = [0.2, 0.5, 1.0]
    view.up[:] = [0.0, 0.0, 1.0]
    view.vfov = 35.0
    W, H = 64, 48
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    cam = hs.camera(W, H, view)
    o = np.array(view.origin[:])
    f = np.array(view.target[:]) - o
    f /= np.linalg.norm(f)
    r = np.cross(f, np.array(view.up[:]))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    th = np.tan(np.radians(35.0) / 2)
    org, dirv = (C.c_double * 3)(), (C.c_double * 3)()
    for px, py, seed in [(0, 0, 1), (63, 47, 2), (31, 20, 3), (10, 40, 4)]:
        st = C.c_uint32(seed)
        L.orc_camera_ray(C.byref(cam), W, H, px, py, C.byref(st), org, dirv)
        s1 = L.orc_lcg_step(seed)
        s2 = L.orc_lcg_step(s1)
        jx = (s1 & 0xFFFFFF) / 2**24 * 2 - 1
        jy = (s2 & 0xFFFFFF) / 2**24 * 2 - 1
        dx = (px + 0.5 + jx / 2) / W * 2 - 1
        dy = (py + 0.5 + jy / 2) / H * 2 - 1
        want = f + r * dx * (W / H) * th - u * dy * th
        want /= np.linalg.norm(want)
        assert np.allclose(org[:], o, atol=1e-5)
        assert np.allclose(dirv[:], want, atol=2e-6)
