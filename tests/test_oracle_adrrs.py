"""ADRRS (adjoint-driven Russian roulette and splitting) and the frame
structure around it in the oracle: applyWeightWindow
(shaders/raytrace.rgen:830-869) against the weight-window definition of
Vorba and Krivanek, "Adjoint-Driven Russian Roulette and Splitting in Light
Transport Simulation" (SIGGRAPH 2016): window centre C = estimate / adjoint,
bounds 2C/(1+s) and s * 2C/(1+s), survivors of the roulette lifted to the
lower bound, split copies brought to the upper bound with an expected copy
count equal to the ratio.  The whole estimator is then pinned by
unbiasedness: frames with ADRRS, with splitOnFirst and with several camera
samples per frame have the plain method's image mean (the weight window
only moves variance), while the counters show that roulette and splitting
happened."""
import numpy as np
import pytest

import rtxamd
from rtxamd import abi
from oracle.binding import OracleScene, adrrs_counters, weight_window
import scenes

S = 5.0  # adrrsS default, src/RayTracingApp.h:112


def lcg(s):
    return (1664525 * s + 1013904223) & 0xFFFFFFFF


def ww(est, adj, thr, state=12345):
    return weight_window(abi, est, adj, thr, S, state)


def outcome(v, est, adj, u):
    """(q, n, drew) for throughput v * u."""
    q, n, st = ww(est, adj, v * u)
    return q, n, st != 12345


def bisect(pred, lo, hi, it=80):
    """Smallest v in [lo, hi] with pred(v), pred monotone."""
    for _ in range(it):
        mid = 0.5 * (lo + hi)
        if pred(mid):
            hi = mid
        else:
            lo = mid
    return hi


@pytest.mark.parametrize("est,adj", [((0.3, 0.2, 0.1), (0.5, 0.4, 0.25)), ((2.0, 1.0, 4.0), (0.1, 0.3, 0.2))])
def test_window_bounds_follow_the_definition(est, adj):
    """The thresholds where the window stops rouletting and starts splitting
    have arithmetic mean C = |estimate / adjoint| and ratio s."""
    u = np.array([0.6, 0.0, 0.8])
    c = np.linalg.norm(np.array(est) / np.array(adj))
    lower = bisect(lambda v: outcome(v, est, adj, u)[0] >= 1.0, 0.0, 100 * c)
    upper = bisect(lambda v: outcome(v, est, adj, u)[2], lower, 100 * c)
    assert abs(0.5 * (lower + upper) - c) <= 1e-9 * c
    assert abs(upper / lower - S) <= 1e-9 * S
    # inside: unchanged, no random draw
    for v in np.linspace(lower * (1 + 1e-9), upper * (1 - 1e-9), 7):
        assert outcome(v, est, adj, u) == (1.0, 1, False)


def test_roulette_lifts_survivors_to_the_lower_bound():
    est, adj, u = (0.3, 0.2, 0.1), (0.5, 0.4, 0.25), np.array([0.0, 1.0, 0.0])
    c = np.linalg.norm(np.array(est) / np.array(adj))
    lower = 2 * c / (1 + S)
    for v in (0.95 * lower, 0.5 * lower, 0.1 * lower):
        q, n, drew = outcome(v, est, adj, u)
        assert n == 1 and not drew  # the caller draws the roulette
        assert abs(v / q - lower) <= 1e-12 * lower  # survivor weight v/q = lower
    # survival never below 0.1 (raytrace.rgen:853-855)
    for v in (0.05 * lower, 1e-6 * lower, 0.0):
        assert outcome(v, est, adj, u)[:2] == (0.1, 1)


def test_split_count_has_the_ratio_as_expectation():
    est, adj, u = (0.3, 0.2, 0.1), (0.5, 0.4, 0.25), np.array([1.0, 0.0, 0.0])
    c = np.linalg.norm(np.array(est) / np.array(adj))
    upper = S * 2 * c / (1 + S)
    for ratio in (1.3, 2.75, 7.5):
        v = ratio * upper
        ns, st = [], 1
        for _ in range(20000):
            st = lcg(lcg(st))
            q, n, st2 = ww(est, adj, v * u, st)
            assert st2 == lcg(st)  # one draw
            assert abs(q - ratio) <= 1e-12 * ratio and abs(v / q - upper) <= 1e-12 * upper
            ns.append(n)
        ns = np.array(ns)
        assert set(np.unique(ns)) <= {int(np.floor(ratio)), int(np.floor(ratio)) + 1}
        frac = ratio - np.floor(ratio)
        se = np.sqrt(frac * (1 - frac) / len(ns))
        assert abs(ns.mean() - ratio) <= 4 * se + 1e-12


def test_degenerate_windows():
    # no prediction possible (0 / 0): unchanged
    assert ww((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))[:2] == (1.0, 1)
    # adjoint 0 against a positive estimate: infinite centre, roulette at 0.1
    assert ww((0.5, 0.5, 0.5), (0.0, 1.0, 1.0), (1.0, 1.0, 1.0))[:2] == (0.1, 1)
    # an estimate of 0: every positive throughput lies above the window
    q, n, _ = ww((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (1.0, 0.0, 0.0))
    assert q == 1.0 and n == 1  # lower <= 0: no prediction


# ------------------------------------------------------------ unbiasedness
W = H = 32
FRAMES = 128


@pytest.fixture(scope="module")
def cornell():
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    orc = OracleScene(abi, hs.desc())
    return hs, orc


def plain_params(hs, sample, **kw):
    kw = dict(dict(maxDepth=8, enableMIS=1), **kw)
    return scenes.params(hs, W, H, 1, sample_offset=sample, **kw)


def frames_mean(render, frames):
    """Image mean over frames and its standard error over pixel-frames."""
    vals = np.concatenate([render(f).mean(axis=1) for f in range(frames)])
    return vals.mean(), vals.std() / np.sqrt(len(vals)), vals


def test_adrrs_frames_match_plain_mean(cornell):
    """Prepare frames build the cache, an estimate frame
    (setEstimateRTSettings, src/RayTracingApp.cpp:1206-1217) stores the pixel
    estimates, then ADRRS frames (cache looked up from depth 2, roulette and
    splitting at cache hits) must have the plain image mean."""
    hs, orc = cornell
    st = rtxamd.IcState.empty(2000)
    for f in range(6):
        fp = plain_params(hs, 500 + f, useIrradianceCache=1, useIrradianceCacheOnGlossy=1, isIrradiancePrepareFrame=1,
                          irradianceCreateProb=0.05, irradianceUpdateProb=0.005)
        orc.ic_frame(fp, st)
    assert st.count > 100
    fe = plain_params(hs, 600, useIrradianceCache=1, useIrradianceCacheOnGlossy=1, samplesPerPixel=16, maxDepth=1,
                      maxFollowDiscrete=10, numNEE=5, irradianceCreateProb=0.0005, irradianceUpdateProb=0.00001)
    est, _, _, _, _ = orc.ic_frame(fe, st)
    assert np.all(np.isfinite(est)) and est.mean() > 0
    rgba = np.zeros((W * H, 4), np.float32)
    rgba[:, :3] = est
    rgba[:, 3] = 1.0
    orc.estimate_set(rgba, W, H)
    adrrs_counters(abi)

    def adrrs(f):
        fp = plain_params(hs, 1000 + f, useADRRS=1, adrrsS=S, adrrsSplit=1, useIrradianceCacheOnGlossy=1,
                          irradianceCreateProb=0.0005)
        rgb, _, _, _, _ = orc.ic_frame(fp, st)
        return rgb

    m_a, se_a, va = frames_mean(adrrs, FRAMES)
    queued, killed, traced = adrrs_counters(abi)
    m_p, se_p, vp = frames_mean(lambda f: orc.render_paths(plain_params(hs, 1000 + f),
                                                           np.arange(W * H, dtype=np.uint64))[0], FRAMES)
    print(f"ADRRS mean {m_a:.5f} +- {se_a:.5f}, plain {m_p:.5f} +- {se_p:.5f}; "
          f"splits queued {queued}, traced {traced}, roulette deaths {killed}")
    assert queued > 100 and killed > 100 and 0 < traced <= queued
    assert abs(m_a - m_p) <= 4 * np.hypot(se_a, se_p)


def test_split_on_first_matches_plain_mean(cornell):
    """splitOnFirst (raytrace.rgen:1142-1146) halves the path at its first
    non-discrete vertex and queues the other half: the mean is unchanged and
    every pixel-frame that reached such a vertex traced one split."""
    hs, orc = cornell
    ids = np.arange(W * H, dtype=np.uint64)
    adrrs_counters(abi)
    m_s, se_s, _ = frames_mean(lambda f: orc.render_paths(plain_params(hs, 2000 + f, splitOnFirst=1), ids)[0],
                               FRAMES)
    queued, killed, traced = adrrs_counters(abi)
    m_p, se_p, _ = frames_mean(lambda f: orc.render_paths(plain_params(hs, 2000 + f), ids)[0], FRAMES)
    print(f"splitOnFirst mean {m_s:.5f} +- {se_s:.5f}, plain {m_p:.5f} +- {se_p:.5f}; splits {queued}/{traced}")
    assert killed == 0 and queued > 0.5 * W * H * FRAMES and traced <= queued
    assert abs(m_s - m_p) <= 4 * np.hypot(se_s, se_p)


def test_samples_per_pixel_frames(cornell):
    """samplesPerPixel camera samples on one random stream per frame: the
    same mean as 1-spp frames, a quarter of the per-pixel variance at 4."""
    hs, orc = cornell
    ids = np.arange(W * H, dtype=np.uint64)
    r1 = np.stack([orc.render_paths(plain_params(hs, 3000 + f), ids)[0].mean(axis=1) for f in range(FRAMES)])
    r4 = np.stack([orc.render_paths(plain_params(hs, 3000 + f, samplesPerPixel=4), ids)[0].mean(axis=1)
                   for f in range(FRAMES)])
    se = np.hypot(r1.std() / np.sqrt(r1.size), r4.std() / np.sqrt(r4.size))
    assert abs(r1.mean() - r4.mean()) <= 4 * se
    ratio = np.mean(r4.var(axis=0)) / np.mean(r1.var(axis=0))
    print(f"variance ratio 4 spp / 1 spp: {ratio:.3f}")
    assert 0.15 < ratio < 0.4
    # samplesPerPixel 0 and 1 are the same single-sample frame
    a = orc.render_paths(plain_params(hs, 7, samplesPerPixel=0), ids)
    b = orc.render_paths(plain_params(hs, 7, samplesPerPixel=1), ids)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def diffuse_box():
    """A closed Lambert box lit by a ceiling quad: every vertex records."""
    s = rtxamd.HostScene.new()
    white = s.add_material(rtxamd.diffuse((0.7, 0.7, 0.7)))
    lamp = s.add_material(rtxamd.light((6.0, 6.0, 6.0), (0.5, 0.5, 0.5)))
    p, i = scenes.quad((-0.3, -0.3, 1.98), (-0.3, 0.3, 1.98), (0.3, 0.3, 1.98), (0.3, -0.3, 1.98))
    s.add_mesh(p, i, lamp)
    p, i = scenes.box((-1, -1, 0), (1, 1, 2))
    s.add_mesh(p, i, white)
    s.set_view((0.0, -0.9, 1.0), (0.0, 0.0, 1.0), (0.0, 0.0, 1.0), 60.0)
    return s.finalize()


def test_record_slots_across_samples():
    """With several camera samples per recording frame, the next sample's
    slots start at sampleOffset + currentSampleOffset when the previous one
    found light, and the slots in between stay INVALID
    (raytrace.rgen:1002, 1220-1222; a rule fixed only by the reference's
    source text).  In an all-Lambert scene every vertex writes a slot, so
    the first sample's slot count is its number of written slots."""
    hs = diffuse_box()
    orc = OracleScene(abi, hs.desc())
    orc.guiding_set([[[-1, -1, 0], [1, 1, 2]]], [rtxamd.vmm_theta([rtxamd.vmf_theta((0.0, 0.0, 1.0), 2.0)], [1.0])])
    ids = np.arange(W * H, dtype=np.uint64)
    kw = dict(maxDepth=3, enableMIS=1, updateGuiding=1)
    rgb1, _, d1 = orc.render_paths_record(plain_params(hs, 9, **kw), ids)
    rgb2, _, d2 = orc.render_paths_record(plain_params(hs, 9, samplesPerPixel=2, **kw), ids)
    rgb3, _, d3 = orc.render_paths_record(plain_params(hs, 9, samplesPerPixel=3, **kw), ids)
    checked = 0
    for p in range(W * H):
        n1 = int(np.sum(d1["pdf"][p] != 0))  # slots sample 1 wrote (from 0)
        r2 = 2 * rgb2[p] - rgb1[p]           # sample 2's own result
        if n1 == 0 or not np.any(rgb1[p] > 0) or not np.any(r2 > 1e-12):
            continue
        assert np.array_equal(d1[p][:n1], d2[p][:n1])  # the same first sample
        # sample 1 found light: sample 2 starts at 0 + n1
        n2 = int(np.sum(d2["pdf"][p][n1:] != 0))
        assert np.all(d2["pdf"][p][n1:n1 + n2] != 0)
        start3 = n1 + (n1 + n2)  # sampleOffset + currentSampleOffset
        if n2 == 0 or start3 >= 16:
            continue
        assert np.array_equal(d2[p][:n1 + n2], d3[p][:n1 + n2])
        gap = slice(n1 + n2, start3)
        assert np.all(d3["flags"][p][gap] == abi.DIR_INVALID) and np.all(d3["pdf"][p][gap] == 0)
        assert d3["pdf"][p][start3] != 0 or np.all(d3["pdf"][p][start3:] == 0)
        checked += 1
    assert checked > 100


def test_recording_with_splitting_renders_nothing(cornell):
    """raytrace.rgen:1644-1648: a recording frame with splitting on returns
    before tracing: black pixels and no valid record."""
    hs, orc = cornell
    orc.guiding_set([[[-1, -1, -1], [1, 1, 1]]], [rtxamd.vmm_theta([rtxamd.vmf_theta((0.0, 0.0, 1.0), 2.0)], [1.0])])
    ids = np.arange(W * H, dtype=np.uint64)
    rgb, _, dd = orc.render_paths_record(plain_params(hs, 9, maxDepth=4, updateGuiding=1, splitOnFirst=1), ids)
    assert np.all(rgb == 0) and np.all(dd["flags"] == abi.DIR_INVALID)


def test_adrrs_errors(cornell):
    hs, orc = cornell
    ids = np.arange(4, dtype=np.uint64)
    with pytest.raises(RuntimeError):  # ADRRS frames look the cache up: orc_ic_frame only
        orc.render_paths(plain_params(hs, 0, useADRRS=1), ids)
    st = rtxamd.IcState.empty(10)
    orc2 = OracleScene(abi, hs.desc())
    with pytest.raises(RuntimeError):  # no estimate
        orc2.ic_frame(plain_params(hs, 0, useADRRS=1), st)
    orc2.estimate_set(np.zeros((W * H, 4), np.float32), W, H)
    with pytest.raises(RuntimeError):  # window ratio below the reference's range
        orc2.ic_frame(plain_params(hs, 0, useADRRS=1, adrrsS=0.5), st)
    with pytest.raises(RuntimeError):  # cache sub-paths would queue splits
        orc2.ic_frame(plain_params(hs, 0, useIrradianceCache=1, splitOnFirst=1), st)
