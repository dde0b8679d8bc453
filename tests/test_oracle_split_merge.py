"""postFit's component merge and split in the oracle (oracle/em_oracle.cpp),
pinned by mathematics where the reference's dependencies cannot run here:
lightpmm's vMF product and division against quadrature of the product and
quotient densities (the quotient's exponent is positive, where lightpmm's
fastexp is not Mineiro's: the factor it makes is part of the expectation),
mergeComponent against first-moment conservation, the merge metric against
quadrature of the chi-square divergence
int (A + B - M)^2 / M between a pair and its merge, the 2x2 eigensolver
against numpy's, and whole updates against their intent: merges remove
redundant lobes of a single source, splits add lobes for many sources and
fit them better, and every state stays a valid mixture.  Which eigenvector
sign and row Eigen hands back, and so the exact split directions, are parity
unpinned (Eigen is absent; the restatement follows Eigen 3's RealSchur)."""
import numpy as np
import pytest

from rtxamd import abi
from oracle.binding import OracleGuide, eigen2x2, lib, merge_metric, vmf_combine, vmf_lobe, vmf_merge

from test_oracle_em import EM_ONLY, mixture, sample_vmf, samples


def fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    phi = np.pi * (1 + 5 ** 0.5) * i
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


DIRS = fib_sphere(400_000)
DOMEGA = 4 * np.pi / len(DIRS)


def density(lobe, dirs=DIRS):
    """w * norm * exp(kappa (mu.d - 1)) of a lobe given as the hooks' 8 values."""
    w, mu, k = lobe[0], np.asarray(lobe[1:4]), lobe[4]
    if k == 0:
        return np.full(len(dirs), w / (4 * np.pi))
    norm = k / (2 * np.pi * (1 - np.exp(-2 * k)))
    return w * norm * np.exp(k * (dirs @ mu - 1))


def unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("ka,kb", [(8.0, 20.0), (2.0, 0.5), (40.0, 35.0)])
def test_vmf_product_weight_is_the_product_integral(ka, kb):
    a = vmf_lobe(0.3, unit((0.2, 0.4, 0.9)), ka)
    b = vmf_lobe(0.6, unit((-0.3, 0.5, 0.8)), kb)
    p = vmf_combine(abi, a, b, False)
    quad = np.sum(density(list(a)) * density(list(b))) * DOMEGA
    assert p[0] == pytest.approx(quad, rel=2e-3)
    # the product lobe is the normalised product density
    eta = ka * np.array(a[1:4]) + kb * np.array(b[1:4])
    assert p[4] == pytest.approx(np.linalg.norm(eta), rel=1e-12)
    assert np.allclose(p[1:4], unit(eta), atol=1e-12)


def fastexp(x):
    return lib(abi).orc_fastexp(float(x))


def quotient_exp_factor(num, den):
    """lightpmm's division evaluates exp at t = kappa' - kappa_num + kappa_den
    (>= 0) with its fastexp, whose fastpow2 always adds 1 to the fractional
    part (pmm-vcl.h:158-172): Mineiro's form for negative arguments only.  The
    quotient weight is the exact integral times fastexp(t) / exp(t)."""
    eta = num[4] * np.array(num[1:4]) - den[4] * np.array(den[1:4])
    t = np.linalg.norm(eta) - num[4] + den[4]
    return fastexp(t) / np.exp(t)


@pytest.mark.parametrize("ka,kb", [(30.0, 6.0), (12.0, 11.0)])
def test_vmf_division_weight_is_the_quotient_integral(ka, kb):
    a = vmf_lobe(0.5, unit((0.0, 0.3, 1.0)), ka)
    b = vmf_lobe(0.2, unit((0.2, 0.0, 1.0)), kb)
    q = vmf_combine(abi, a, b, True)
    quad = np.sum(density(list(a)) / density(list(b))) * DOMEGA
    assert q[0] == pytest.approx(quad * quotient_exp_factor(a, b), rel=2e-3)
    eta = ka * np.array(a[1:4]) - kb * np.array(b[1:4])
    assert q[4] == pytest.approx(np.linalg.norm(eta), rel=1e-12)


def test_merge_keeps_weight_and_mean_resultant():
    a = vmf_lobe(0.25, unit((0.1, 0.2, 1.0)), 15.0)
    b = vmf_lobe(0.15, unit((0.4, -0.1, 1.0)), 7.0)
    m = vmf_merge(abi, a, b)
    assert m[0] == pytest.approx(0.4, rel=1e-12)
    want = a[0] * a[5] * np.array(a[1:4]) + b[0] * b[5] * np.array(b[1:4])
    assert np.allclose(m[0] * m[5] * m[1:4], want, atol=1e-12)
    r = m[5]
    assert m[4] == pytest.approx((3 * r - r ** 3) / (1 - r * r), rel=1e-12)  # Banerjee et al. 2005


@pytest.mark.parametrize("sep", [0.05, 0.3, 1.2])
def test_merge_metric_is_the_chi_square_divergence(sep):
    """metric = int A^2/M + 2 AB/M + B^2/M - w_M = int (A + B - M)^2 / M with
    w_M = w_A + w_B; each quotient weight carries lightpmm's fastexp factor."""
    a = vmf_lobe(0.3, unit((0.0, 0.0, 1.0)), 20.0)
    b = vmf_lobe(0.2, unit((np.sin(sep), 0.0, np.cos(sep))), 12.0)
    m = vmf_merge(abi, a, b)
    fa, fb, fm = density(list(a)), density(list(b)), density(list(m))
    aa, bb, ab = vmf_combine(abi, a, a, False), vmf_combine(abi, b, b, False), vmf_combine(abi, a, b, False)
    terms = [(fa * fa, aa), (fa * fb, ab), (fb * fb, bb)]
    quad = [np.sum(f / fm) * DOMEGA * quotient_exp_factor(p, m) for f, p in terms]
    want = quad[0] + 2 * quad[1] + quad[2] - m[0]
    got = merge_metric(abi, a, b)
    exact = np.sum((fa + fb - fm) ** 2 / fm) * DOMEGA
    print(f"sep {sep}: metric {got:.5f}, chi-square {exact:.5f}")
    assert got == pytest.approx(want, rel=5e-3, abs=5e-5), (got, want)
    if sep == 0.05:
        assert exact <= 0.025 and got <= 0.025  # close lobes qualify for merging


def test_eigen2x2_matches_numpy():
    rng = np.random.default_rng(7)
    for _ in range(200):
        b = rng.normal(size=(2, 2)) * rng.uniform(1e-4, 2.0)
        a = b @ b.T  # symmetric positive semidefinite, like a covariance
        lam, V = eigen2x2(abi, a)
        assert np.allclose(np.sort(lam), np.linalg.eigvalsh(a), rtol=1e-10, atol=1e-14 * np.abs(a).max())
        for i in range(2):
            v = V[:, i]
            assert np.linalg.norm(v) == pytest.approx(1.0, abs=1e-12)
            assert np.allclose(a @ v, lam[i] * v, atol=1e-10 * np.abs(a).max())
    # an already diagonal matrix needs no rotation: identity vectors, diagonal order
    lam, V = eigen2x2(abi, np.diag([0.2, 0.5]))
    assert np.array_equal(lam, [0.2, 0.5]) and np.array_equal(V, np.eye(2))
    lam, V = eigen2x2(abi, np.zeros((2, 2)))
    assert np.array_equal(lam, [0, 0]) and np.array_equal(V, np.eye(2))


def valid_mixture(vmm):
    pis, ks, mus = mixture(vmm)
    K = len(pis)
    assert 1 <= K <= 16
    assert pis.sum() == pytest.approx(1.0, abs=1e-5)
    assert np.all(pis >= 0.01 / (0.01 * K + 1) - 1e-6)  # the weight prior's floor
    assert np.allclose(np.linalg.norm(mus, axis=1), 1, atol=1e-5)
    assert np.all(ks >= 0) and np.all(ks <= 50000 * 1.01)
    return pis, ks, mus


def log_likelihood(vmm, dirs):
    pis, ks, mus = mixture(vmm)
    f = np.zeros(len(dirs))
    for p, k, m in zip(pis, ks, mus):
        f += p * density([1.0, *m, k], dirs)
    return np.mean(np.log(f))


@pytest.mark.parametrize("prec", [64, 32])
def test_merges_remove_redundant_lobes(prec):
    """One concentrated source: the 8 initial lobes collapse onto it and,
    with more than minSamplesForMerging samples, are merged."""
    rng = np.random.default_rng(11)
    d = sample_vmf((0.3, -0.2, 0.9), 80.0, 20000, rng)
    g = OracleGuide(abi, ([-1, -1, -1], [1, 1, 1]), 0, parallax=False, prec=prec)
    g.update(samples(d))
    pis, ks, mus = valid_mixture(g.regions()[1][0])
    em = OracleGuide(abi, ([-1, -1, -1], [1, 1, 1]), 0, parallax=False, prec=prec, options=EM_ONLY)
    em.update(samples(d))
    assert em.regions()[1][0].usedDistributions == 8
    print(f"K after split/merge: {len(pis)}")
    assert len(pis) < 8
    near = mus @ unit((0.3, -0.2, 0.9)) > 0.98
    assert pis[near].sum() > 0.9
    held = sample_vmf((0.3, -0.2, 0.9), 80.0, 5000, rng)
    assert log_likelihood(g.regions()[1][0], held) >= log_likelihood(em.regions()[1][0], held) - 0.05


def test_splits_add_lobes_for_many_sources():
    """Twelve concentrated sources: more than the 8 initial lobes are
    needed; splitting grows the mixture and the fit explains held-out
    samples better than the 8-lobe EM."""
    rng = np.random.default_rng(12)
    centres = fib_sphere(12)
    g = OracleGuide(abi, ([-1, -1, -1], [1, 1, 1]), 0, parallax=False)
    em = OracleGuide(abi, ([-1, -1, -1], [1, 1, 1]), 0, parallax=False, options=EM_ONLY)
    for it in range(4):
        d = np.concatenate([sample_vmf(c, 300.0, 1500, rng) for c in centres])
        rng.shuffle(d)
        g.update(samples(d))
        em.update(samples(d))
        valid_mixture(g.regions()[1][0])
    vmm = g.regions()[1][0]
    pis, ks, mus = valid_mixture(vmm)
    held = np.concatenate([sample_vmf(c, 300.0, 500, rng) for c in centres])
    ll, ll_em = log_likelihood(vmm, held), log_likelihood(em.regions()[1][0], held)
    covered = sum(np.any((mus @ c > 0.99) & (pis > 0.02)) for c in centres)
    print(f"K {len(pis)}, sources covered {covered}/12, held-out log-likelihood {ll:.3f} vs EM-only {ll_em:.3f}")
    assert len(pis) > 8
    assert covered >= 10
    assert ll > ll_em


def test_split_merge_with_parallax_keeps_distances():
    """With parallax, split components inherit their parent's distance and
    merged ones take the harmonic mean: every lobe keeps a finite target
    for a point source."""
    rng = np.random.default_rng(13)
    light = np.array([0.4, 0.1, 2.5])
    g = OracleGuide(abi, ([-1, -1, -1], [1, 1, 1]), 0, parallax=True)
    for it in range(3):
        pos = rng.uniform(-0.9, 0.9, (10000, 3))
        v = light - pos
        dist = np.linalg.norm(v, axis=1)
        s = samples(v / dist[:, None])
        s["position"] = pos
        s["distance"] = dist
        g.update(s)
    vmm = g.regions()[1][0]
    pis, ks, mus = valid_mixture(vmm)
    k = int(np.argmax(pis))
    t = vmm.thetas[k]
    assert t.distance > 0
    assert np.allclose(np.array(t.target[:]), light, atol=0.15 * np.linalg.norm(light)), (t.target[:], light)


def test_region_splitting_halves_busy_regions():
    """splitRegions: a region that saw more than samplesForRegionSplit
    samples is halved along its largest axis; both halves carry its mixture
    (with the sample counts decayed) and the new half is appended."""
    rng = np.random.default_rng(14)
    opts = abi.guiding_options(splitRegions=1, samplesForRegionSplit=3000)
    g = OracleGuide(abi, ([0, 0, 0], [4, 2, 1]), 1, parallax=False, options=opts)
    boxes0, _ = g.regions()
    assert len(boxes0) == 2
    d = sample_vmf((0, 0, 1), 20.0, 5000, rng)
    s = samples(d, region=0)
    g.update(s)  # region 0 has 5000 > 3000 samples, region 1 none
    boxes, vmms = g.regions()
    assert len(boxes) == 3
    vol = lambda b: np.prod(b[1] - b[0])
    assert vol(boxes[0]) + vol(boxes[2]) == pytest.approx(vol(boxes0[0]), rel=1e-6)
    assert np.allclose(boxes[1], boxes0[1])
    # halves of the old region 0 along its largest axis (x and y tie at 2.00002: y, the later one)
    size = boxes0[0][1] - boxes0[0][0]
    axis = 0 if size[0] > size[1] else (1 if size[1] > size[2] else 2)
    assert axis == 1
    assert np.allclose(boxes[0][0], boxes0[0][0]) and np.allclose(boxes[2][1], boxes0[0][1])
    assert boxes[0][1][axis] == pytest.approx(boxes[2][0][axis]) == pytest.approx(boxes0[0][0][axis] + size[axis] / 2)
    pa, ka, ma = mixture(vmms[0])
    pb, kb, mb = mixture(vmms[2])
    assert np.array_equal(pa, pb) and np.array_equal(ka, kb) and np.array_equal(ma, mb)
    # the decayed counts: a second batch of the same size has four times the old weight in the refit
    g.update(samples(d, region=2))
    assert len(g.regions()[0]) == 4  # region 2 split again (1250 + 5000 > 3000)
