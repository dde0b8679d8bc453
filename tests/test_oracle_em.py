"""The oracle's guiding update (PathGuiding::update with lightpmm's MAP-EM,
oracle/em_oracle.cpp) pinned by mathematics: fastexp's published error
bound, the region layout of PathGuiding::createRegions, the invariants of a
fitted mixture (weights sum to 1 and respect the weight prior, unit mean
directions, kappa in [0, maxKappa]), the first moment of the samples carried
by the fit, the fit of a single known vMF lobe, incremental refits, and
parallax targets that converge on a point light.  Exact EM trajectories of
lightpmm are parity unpinned (lightpmm cannot be built here: it needs GLM)."""
import numpy as np
import pytest

from rtxamd import abi
from oracle.binding import OracleGuide, lib

# the MAP-EM alone; postFit's split and merge have their own tests
EM_ONLY = abi.guiding_options(splitAndMerge=0)


def sample_vmf(mu, kappa, n, rng):
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


def samples(dirs, region=0, weight=1.0, pos=(0, 0, 0), dist=0.0):
    s = np.zeros(len(dirs), abi.DIRECTIONAL_DTYPE)
    s["direction"] = dirs
    s["position"] = pos
    s["weight"] = weight
    s["pdf"] = 1.0
    s["distance"] = dist
    s["flags"] = region
    return s


def mixture(vmm):
    K = vmm.usedDistributions
    pis = np.array(vmm.pi[:K], np.float64)
    ks = np.array([vmm.thetas[i].k for i in range(K)], np.float64)
    mus = np.array([vmm.thetas[i].mu[:] for i in range(K)], np.float64)
    return pis, ks, mus


def mean_cosine(k):
    return np.where(k > 0, 1 / np.tanh(np.maximum(k, 1e-9)) - 1 / np.maximum(k, 1e-9), 0.0)


def test_fastexp_error_bound():
    L = lib(abi)
    xs = np.linspace(-80, 0, 20001)
    fe = np.array([L.orc_fastexp(float(x)) for x in xs])
    # fastapprox's fastexp: relative error below 1e-4 (Mineiro 2011)
    assert np.max(np.abs(fe / np.exp(xs) - 1)) < 1e-4


def test_regions_split_the_scene_box():
    g = OracleGuide(abi, ([0, 0, 0], [4, 2, 1]), 3, parallax=False)
    boxes, vmms = g.regions()
    assert len(boxes) == 8
    # addEpsilon: 0.50001 of the extent around the centre
    lo, hi = boxes[:, 0].min(0), boxes[:, 1].max(0)
    assert np.allclose(lo, [2 - 0.50001 * 4, 1 - 0.50001 * 2, 0.5 - 0.50001], atol=1e-6)
    assert np.allclose(hi, [2 + 0.50001 * 4, 1 + 0.50001 * 2, 0.5 + 0.50001], atol=1e-6)
    vol = np.prod(boxes[:, 1] - boxes[:, 0], axis=1)
    assert np.allclose(vol, vol[0], rtol=1e-5) and vol.sum() == pytest.approx(np.prod(hi - lo), rel=1e-5)
    # the initial mixtures: 8 lobes of kappa 5 at spherical-Fibonacci directions, equal weights
    pis, ks, mus = mixture(vmms[0])
    assert len(pis) == 8 and np.allclose(pis, 1 / 8) and np.allclose(ks, 5.0)
    assert np.allclose(mus.sum(0), 0, atol=0.15) and np.allclose(np.linalg.norm(mus, axis=1), 1, atol=1e-6)


@pytest.mark.parametrize("prec", [64, 32])
def test_fit_keeps_the_first_moment_and_invariants(prec):
    rng = np.random.default_rng(1)
    d = np.concatenate([sample_vmf((0.3, 0.5, 0.8), 20.0, 12000, rng), sample_vmf((-1, 0, 0.2), 4.0, 8000, rng)])
    w = rng.uniform(0.5, 2.0, len(d))
    g = OracleGuide(abi, ([-1, -1, -1], [1, 1, 1]), 0, parallax=False, prec=prec, options=EM_ONLY)
    g.update(samples(d, weight=w))
    pis, ks, mus = mixture(g.regions()[1][0])
    assert pis.sum() == pytest.approx(1.0, abs=1e-5)
    assert np.all(pis >= 0.01 / (0.01 * 8 + 1) - 1e-7)  # the weight prior's floor
    assert np.allclose(np.linalg.norm(mus, axis=1), 1, atol=1e-5)
    assert np.all(ks >= 0) and np.all(ks <= 50000 * 1.01)
    # the weighted first moment of the samples is what the mixture's mean resultant gives back
    want = (w[:, None] * d).sum(0) / w.sum()
    got = (pis[:, None] * mean_cosine(ks)[:, None] * mus).sum(0)
    assert np.allclose(got, want, atol=0.02), (got, want)


def test_fit_of_one_lobe_concentrates_on_it():
    rng = np.random.default_rng(2)
    mu = np.array([0.0, 0.6, 0.8])
    d = sample_vmf(mu, 50.0, 20000, rng)
    g = OracleGuide(abi, ([-1, -1, -1], [1, 1, 1]), 0, parallax=False, options=EM_ONLY)
    g.update(samples(d))
    pis, ks, mus = mixture(g.regions()[1][0])
    # most of the mass sits in lobes aligned with mu, as concentrated as the source
    near = mus @ mu > 0.98
    assert pis[near].sum() > 0.85
    assert np.average(ks[near], weights=pis[near]) == pytest.approx(50.0, rel=0.35)


def test_update_fit_accumulates():
    """updateFit folds the mixture's own sufficient statistics back in: after a
    second batch the fit carries the first moment of both batches together."""
    rng = np.random.default_rng(3)
    a = sample_vmf((0, 0, 1), 4.0, 6000, rng)
    b = sample_vmf((0.6, 0.0, 0.8), 4.0, 6000, rng)
    g = OracleGuide(abi, ([-1, -1, -1], [1, 1, 1]), 0, parallax=False, options=EM_ONLY)
    g.update(samples(a))
    pis, ks, mus = mixture(g.regions()[1][0])
    first = (pis[:, None] * mean_cosine(ks)[:, None] * mus).sum(0)
    assert np.allclose(first, a.mean(0), atol=0.02)
    g.update(samples(b))
    pis, ks, mus = mixture(g.regions()[1][0])
    both = (pis[:, None] * mean_cosine(ks)[:, None] * mus).sum(0)
    assert np.allclose(both, np.concatenate([a, b]).mean(0), atol=0.03), (both, np.concatenate([a, b]).mean(0))


def test_uncovered_samples_are_skipped():
    """lightpmm skips samples whose mixture pdf is below 1e-8
    (SoftAssignmentWeights::valid): a mixture fitted to one pole ignores a
    batch from the other."""
    rng = np.random.default_rng(5)
    g = OracleGuide(abi, ([-1, -1, -1], [1, 1, 1]), 0, parallax=False, options=EM_ONLY)
    g.update(samples(sample_vmf((0, 0, 1), 200.0, 5000, rng)))
    before = mixture(g.regions()[1][0])
    g.update(samples(sample_vmf((0, 0, -1), 200.0, 5000, rng)))
    after = mixture(g.regions()[1][0])
    assert np.allclose(before[2], after[2], atol=1e-4) and after[2][:, 2].min() > 0.9


def test_parallax_targets_find_a_point_light():
    rng = np.random.default_rng(4)
    light = np.array([0.2, -0.3, 3.0])
    g = OracleGuide(abi, ([-1, -1, -1], [1, 1, 1]), 0, parallax=True, options=EM_ONLY)
    for it in range(3):
        pos = rng.uniform(-0.9, 0.9, (8000, 3))
        v = light - pos
        dist = np.linalg.norm(v, axis=1)
        s = samples(v / dist[:, None], dist=0.0)
        s["position"] = pos
        s["distance"] = dist
        g.update(s)
    vmm = g.regions()[1][0]
    pis, ks, mus = mixture(vmm)
    k = int(np.argmax(pis))
    t = vmm.thetas[k]
    assert t.distance > 0
    # the region's mean position is the centre of its box (0, 0, 0)
    assert np.allclose(vmm.meanPosition[:], 0.0, atol=1e-5)
    assert np.allclose(np.array(t.target[:]), light, atol=0.15 * np.linalg.norm(light)), (t.target[:], light)
