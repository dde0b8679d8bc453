"""The oracle's MAP-EM steps (oracle/em_oracle.cpp after lightpmm's
VMMFactory.h:157-216, 305-362, 385-486) pinned by closed forms on one FP64
mixture:

* the M-step (parameterUpdate) from given statistics: mixture weights are
  the MAP estimate under the weight prior, (W/S + v) / (1 + K v); mean
  directions the normalised resultants; mean cosines the resultant length
  per weight, blended with the prior as (r N + r0 n0) / (N + n0) where N is
  the component's share of all samples seen, and clamped at the mean cosine
  of maxKappa; concentrations from Banerjee et al.'s (2005) approximation
  (3r - r^3) / (1 - r^2), which is checked against the exact inverse of
  A3(k) = coth(k) - 1/k;
* the E-step's statistics against brute force with the exact exponential
  (the oracle uses lightpmm's fastexp, relative error < 1e-4);
* updateFit's merge of old and new statistics: for one lobe, fitting batch A
  and then updating with batch B gives the fit of A and B together;
* the fit of a single known vMF lobe within its sampling error.

Each closed-form check also shows that the plausible mistakes (a dropped or
mis-blended prior, the wrong sample count in the blend) move the result far
outside its tolerance."""
import ctypes as C

import numpy as np
import pytest

from rtxamd import abi
from oracle.binding import lib

L = lib(abi)
ND = L.orc_em_mix_doubles()
KMAX = (ND - 4) // 6  # lobe slots of the oracle's mixture record
DP = C.POINTER(C.c_double)


def ptr(a):
    return a.ctypes.data_as(DP)


def props(maxKappa=50000.0, vPrior=0.01, rPrior=0.0, rPriorWeight=1.0, minItr=1, maxItr=100, relLL=0.005):
    return np.array([minItr, maxItr, relLL, maxKappa, vPrior, rPrior, rPriorWeight], np.float64)


def A3(k):
    """Mean cosine of a 3-D vMF lobe of concentration k."""
    k = np.asarray(k, float)
    return np.where(k > 1e-8, 1 / np.tanh(np.maximum(k, 1e-8)) - 1 / np.maximum(k, 1e-8), 0.0)


def banerjee(r):
    return (3 * r - r ** 3) / (1 - r * r)


def mix(w, mu, kappa, sampleWeight=0.0, numSamples=0.0, total=0):
    v = np.zeros(ND)
    K = len(w)
    v[0] = K
    for k in range(KMAX):
        c = v[1 + 6 * k: 7 + 6 * k]
        if k < K:
            m = np.asarray(mu[k], float)
            c[:] = [w[k], *(m / np.linalg.norm(m)), kappa[k], float(A3(kappa[k]))]
        else:
            c[:] = [0, 0, 0, 1, 0, 0]
    v[1 + 6 * KMAX:] = [sampleWeight, numSamples, total]
    return v


def unpack(v):
    K = int(v[0])
    c = v[1:1 + 6 * KMAX].reshape(KMAX, 6)[:K]
    return dict(w=c[:, 0].copy(), mu=c[:, 1:4].copy(), kappa=c[:, 4].copy(), r=c[:, 5].copy(),
                sampleWeight=v[1 + 6 * KMAX], numSamples=v[2 + 6 * KMAX], total=v[3 + 6 * KMAX])


def mstep(p, m, rW, W):
    m = m.copy()
    rW = np.ascontiguousarray(rW, np.float64)
    W = np.ascontiguousarray(W, np.float64)
    assert L.orc_em_mstep(ptr(p), ptr(m), ptr(rW), ptr(W)) == 0
    return unpack(m)


def fit(p, m, dirs, weights, update=False):
    m = m.copy()
    dirs = np.ascontiguousarray(dirs, np.float64)
    weights = np.ascontiguousarray(weights, np.float64)
    assert L.orc_em_fit(ptr(p), ptr(m), ptr(dirs), ptr(weights), len(weights), int(update)) == 0
    return m


def sample_vmf(mu, kappa, n, rng):
    """Exact vMF sampler (inverse CDF of the cosine, Wood 1994 for p = 3)."""
    mu = np.asarray(mu, float)
    mu /= np.linalg.norm(mu)
    u, v = rng.random(n), rng.random(n)
    w = 1 + np.log(u + (1 - u) * np.exp(-2 * kappa)) / kappa
    phi = 2 * np.pi * v
    s = np.sqrt(np.maximum(0, 1 - w * w))
    local = np.stack([s * np.cos(phi), s * np.sin(phi), w], 1)
    a = np.array([1.0, 0, 0]) if abs(mu[0]) < 0.9 else np.array([0, 1.0, 0])
    x = np.cross(a, mu)
    x /= np.linalg.norm(x)
    return local @ np.stack([x, np.cross(mu, x), mu])


def test_banerjee_approximation_inverts_the_mean_cosine():
    """(3r - r^3)/(1 - r^2) approximates the inverse of A3 within a few
    per cent over the range the fit uses (Banerjee et al. 2005, Fig. 1)."""
    k = np.array([0.5, 1.0, 2.0, 5.0, 20.0, 100.0, 1000.0])
    assert np.all(np.abs(banerjee(A3(k)) / k - 1) < 0.05)


def test_mstep_closed_form():
    rng = np.random.default_rng(11)
    K = 3
    rW = rng.normal(size=(K, 3))
    W = np.array([2.0, 5.0, 0.5])
    rW *= (W * np.array([0.4, 0.85, 0.95]) / np.linalg.norm(rW, axis=1))[:, None]  # mean cosines < 1
    S, total = W.sum() * 1.3, 24  # some weight went to samples outside the components' statistics
    v0, r0, n0 = 0.05, 0.3, 4.0
    p = props(vPrior=v0, rPrior=r0, rPriorWeight=n0)
    m0 = mix([1 / 3] * 3, [(0, 0, 1)] * 3, [5.0] * 3, sampleWeight=S, numSamples=total, total=total)
    got = mstep(p, m0, rW, W)

    mu = rW / np.linalg.norm(rW, axis=1)[:, None]
    rbar = np.linalg.norm(rW, axis=1) / W
    f32 = lambda x: float(np.float32(x))  # the factory's properties are FP32 fields
    pi = (W / S + f32(v0)) / (1 + K * f32(v0))
    N = total * pi
    r = (rbar * N + f32(r0) * f32(n0)) / (N + f32(n0))
    kappa = banerjee(r)
    assert np.allclose(got["w"], pi, rtol=1e-12)
    assert np.allclose(got["mu"], mu, rtol=1e-12)
    assert np.allclose(got["r"], r, rtol=1e-12)
    assert np.allclose(got["kappa"], kappa, rtol=1e-12)
    # the check has teeth: the usual mistakes are far outside 1e-12
    wrong = {
        "weight prior dropped": (W / S, got["w"]),
        "mean-cosine prior dropped": (banerjee(rbar), got["kappa"]),
        "prior blended with the sample weight": (banerjee((rbar * W + f32(r0) * f32(n0)) / (W + f32(n0))),
                                                 got["kappa"]),
        "prior blended without the weight share": (banerjee((rbar * total + f32(r0) * f32(n0)) / (total + f32(n0))),
                                                   got["kappa"]),
    }
    for name, (alt, have) in wrong.items():
        assert np.max(np.abs(alt / have - 1)) > 1e-3, name


def test_mstep_clamps_at_max_kappa_and_zeroes_tiny_kappa():
    K = 2
    W = np.array([1000.0, 10.0])
    rW = np.array([[0, 0, 1000.0], [0.0, 1e-5, 0.0]])  # all on one direction; nearly isotropic
    m0 = mix([0.5, 0.5], [(0, 0, 1), (1, 0, 0)], [5.0, 5.0], sampleWeight=W.sum(), numSamples=1010, total=1010)
    for maxKappa in (200.0, 50000.0):
        got = mstep(props(maxKappa=maxKappa), m0, rW, W)
        v0 = float(np.float32(0.01))
        pi0 = (W[0] / W.sum() + v0) / (1 + 2 * v0)
        N = 1010 * pi0
        r = N / (N + 1)  # mean cosine 1 blended with the prior r0 = 0, n0 = 1
        r = min(r, 1 - 1 / maxKappa)  # A3 of maxKappa (> 5) as lightpmm's kappaToMeanCosine
        assert got["r"][0] == pytest.approx(r, rel=1e-12)
        assert got["kappa"][0] == pytest.approx(banerjee(r), rel=1e-9)
        if maxKappa == 200.0:
            assert got["r"][0] == pytest.approx(1 - 1 / 200.0, rel=1e-12)
        # a concentration below 1e-3 is zero, with mean cosine zero (VMFKernel::setKappa)
        assert got["kappa"][1] == 0.0 and got["r"][1] == 0.0


def vmf_pdf(mu, kappa, d):
    return kappa / (2 * np.pi * (1 - np.exp(-2 * kappa))) * np.exp(kappa * (d @ mu - 1))


def test_estep_matches_brute_force():
    rng = np.random.default_rng(5)
    w = np.array([0.3, 0.7])
    mu = np.array([[0.0, 0.6, 0.8], [1.0, 0.0, 0.0]])
    kap = np.array([10.0, 3.0])
    m0 = mix(w, mu, kap)
    d = rng.normal(size=(3000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    wt = rng.uniform(0.2, 3.0, len(d))
    rW = np.zeros(3 * KMAX)
    W = np.zeros(KMAX)
    sw, ll = C.c_double(), C.c_double()
    assert L.orc_em_estep(ptr(m0), ptr(np.ascontiguousarray(d)), ptr(wt), len(wt), ptr(rW), ptr(W), C.byref(sw),
                          C.byref(ll)) == 0
    a = np.stack([w[k] * vmf_pdf(mu[k], kap[k], d) for k in range(2)], 1)
    tot = a.sum(1)
    g = a / tot[:, None] * wt[:, None]
    assert np.allclose(W[:2], g.sum(0), rtol=5e-4)
    assert np.allclose(rW[:6].reshape(2, 3), g.T @ d, rtol=5e-4, atol=1e-3 * np.abs(g.T @ d).max())
    assert sw.value == pytest.approx(wt.sum(), rel=1e-12)
    assert ll.value == pytest.approx(float((wt * np.log(tot)).sum()), rel=1e-3)
    # responsibilities must be normalised per sample: without it W would be off by the mixture pdf
    assert abs((a * wt[:, None]).sum(0)[0] / W[0] - 1) > 1e-2


def test_update_fit_of_one_lobe_equals_fit_of_both_batches():
    """updateFit rebuilds the old statistics from the mixture (removing the
    priors) and adds the new batch's: with one lobe every responsibility is
    1, so fit(A) then updateFit(B) is fit(A and B).  Small batches and a
    strong mean-cosine prior make a wrong prior removal visible."""
    rng = np.random.default_rng(7)
    a = sample_vmf((0, 0.3, 1), 6.0, 40, rng)
    b = sample_vmf((0.5, 0, 1), 3.0, 60, rng)
    wa, wb = rng.uniform(0.5, 2, 40), rng.uniform(0.5, 2, 60)
    p = props(rPrior=0.2, rPriorWeight=3.0)
    m0 = mix([1.0], [(0, 0, 1)], [5.0])
    seq = unpack(fit(p, fit(p, m0, a, wa), b, wb, update=True))
    both = unpack(fit(p, m0, np.concatenate([a, b]), np.concatenate([wa, wb])))
    for key in ("w", "mu", "kappa", "r"):
        assert np.allclose(seq[key], both[key], rtol=1e-9), key
    assert seq["sampleWeight"] == pytest.approx(wa.sum() + wb.sum(), rel=1e-12)
    assert seq["numSamples"] == 100 and seq["total"] == 100
    # the closed form of the one-lobe fit
    res = (np.concatenate([wa, wb])[:, None] * np.concatenate([a, b])).sum(0)
    rbar = np.linalg.norm(res) / (wa.sum() + wb.sum())
    r = (rbar * 100 + float(np.float32(0.2)) * 3.0) / (100 + 3.0)
    assert both["r"][0] == pytest.approx(r, rel=1e-12)
    assert np.allclose(both["mu"][0], res / np.linalg.norm(res), rtol=1e-12)
    # without the prior removal the update would see r blended twice
    assert abs(banerjee((rbar * 100) / 103.0) / seq["kappa"][0] - 1) > 1e-3


@pytest.mark.parametrize("kappa", [5.0, 20.0, 80.0])
def test_one_lobe_fit_recovers_kappa_within_sampling_error(kappa):
    rng = np.random.default_rng(int(kappa))
    mu = np.array([0.0, 0.6, 0.8])
    n = 200000
    d = sample_vmf(mu, kappa, n, rng)
    got = unpack(fit(props(), mix([1.0], [(0, 0, 1)], [5.0]), d, np.ones(n)))
    # what the sample determines: r = n |mean| / (n + 1) with the default prior (r0 = 0, n0 = 1)
    rbar = np.linalg.norm(d.mean(0))
    assert got["r"][0] == pytest.approx(rbar * n / (n + 1), rel=1e-9)
    # against the truth: Banerjee's estimate of the true lobe, within 4 standard errors of r
    se_r = d @ mu
    se_r = se_r.std() / np.sqrt(n)
    r_true = float(A3(kappa))
    dk = (banerjee(r_true + 4 * se_r) - banerjee(r_true - 4 * se_r)) / 2
    assert abs(got["kappa"][0] - banerjee(r_true)) < dk + 1e-9, (got["kappa"][0], banerjee(r_true), dk)
    assert np.dot(got["mu"][0], mu) > 1 - 1e-4


def test_partition_order_is_libstdcxx_std_partition(tmp_path):
    """postFit takes its merge and split candidates in the order
    std::partition leaves them (src/PathGuiding.cpp:563-711); the oracle's
    restatement of that order equals the compiler's own std::partition
    (g++ 11, libstdc++) on the same keys, element for element."""
    import subprocess
    src = tmp_path / "part.cpp"
    src.write_text(r'''
#include <algorithm>
#include <cstdio>
#include <utility>
#include <vector>
int main() {
    int n; double pivot;
    while (std::scanf("%d %lf", &n, &pivot) == 2) {
        std::vector<std::pair<double, unsigned>> v(n);
        for (int i = 0; i < n; ++i) { std::scanf("%lf", &v[i].first); v[i].second = i; }
        auto mid = std::partition(v.begin(), v.end(), [&](const std::pair<double, unsigned> &x) { return x.first <= pivot; });
        std::printf("%d", (int)(mid - v.begin()));
        for (auto &x : v) std::printf(" %u", x.second);
        std::printf("\n");
    }
}
''')
    exe = tmp_path / "part"
    subprocess.run(["g++", "-O1", "-o", str(exe), str(src)], check=True)
    rng = np.random.default_rng(9)
    cases = [(int(n), rng.random(int(n)), float(p)) for n, p in zip(rng.integers(0, 40, 60), rng.random(60))]
    cases += [(5, np.array([0.1, 0.1, 0.1, 0.1, 0.1]), 0.5), (5, np.array([0.9] * 5), 0.5)]
    text = "".join(f"{n} {p!r}\n" + " ".join(repr(float(k)) for k in keys) + "\n" for n, keys, p in cases)
    lines = subprocess.run([str(exe)], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
    for (n, keys, p), line in zip(cases, lines):
        want = [int(x) for x in line.split()]
        out = (C.c_uint32 * max(n, 1))()
        keys = np.ascontiguousarray(keys, np.float64)
        mid = L.orc_partition_order(ptr(keys), n, p, out)
        assert [mid] + list(out[:n]) == want
