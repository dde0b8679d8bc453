"""The oracle's irradiance cache (shaders/raytrace.rgen:748-828, 1228-1421;
raytrace.irradiance.rint/.rahit) pinned by closed forms and exact rules:
lookup weights and their thresholds, the cosine-weighted stratification,
the irradiance and harmonic-mean distance of a point under a lit ceiling,
the accumulation on constant inputs, and the nextUpdateSlot rule."""
import numpy as np
import pytest

import rtxamd
from rtxamd import abi
from oracle.binding import OracleScene, ic_query, ic_reduce, ic_stratum
import scenes

N_PHI, N_THETA = 20, 10


def pc(**kw):
    p = rtxamd.push_constant_defaults()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


# ------------------------------------------------------------------ lookup
def test_lookup_single_record_returns_its_colour():
    st = scenes.cache_state([((0, 0, 0), (0, 0, 1), (0.3, 0.5, 0.7), 1.0)], 4)
    c = ic_query(abi, pc(irradianceA=0.2), st, (0.05, 0.0, 0.0), (0, 0, 1))
    assert np.allclose(c, (0.3, 0.5, 0.7), rtol=1e-6)


def test_lookup_weights_two_records():
    # same normal: w = 1 / (d / R) = R / d
    recs = [((0, 0, 0), (0, 0, 1), (1.0, 0.0, 0.0), 1.0), ((0.1, 0, 0), (0, 0, 1), (0.0, 1.0, 0.0), 2.0)]
    st = scenes.cache_state(recs, 4, irradiance_a=0.2)
    x = (0.04, 0.0, 0.0)
    w1, w2 = 1.0 / 0.04, 2.0 / 0.06
    c = ic_query(abi, pc(irradianceA=0.2), st, x, (0, 0, 1))
    assert np.allclose(c, (w1 / (w1 + w2), w2 / (w1 + w2), 0.0), rtol=1e-5)


def test_lookup_threshold_and_radius():
    # record radius a*R = 0.2; weight R/d > 1/a holds exactly inside it
    st = scenes.cache_state([((0, 0, 0), (0, 0, 1), (1.0, 1.0, 1.0), 1.0)], 2)
    assert ic_query(abi, pc(irradianceA=0.2), st, (0.199, 0, 0), (0, 0, 1)) is not None
    assert ic_query(abi, pc(irradianceA=0.2), st, (0.201, 0, 0), (0, 0, 1)) is None
    # a tilted normal adds sqrt(1 - n.ni): 0.1 + sqrt(1 - cos 30deg) = 0.466 > a -> rejected
    n30 = (np.sin(np.pi / 6), 0.0, np.cos(np.pi / 6))
    assert ic_query(abi, pc(irradianceA=0.2), st, (0.1, 0, 0), n30) is None
    # ... but accepted with a = 0.5 once the radius allows the point
    st5 = scenes.cache_state([((0, 0, 0), (0, 0, 1), (1.0, 1.0, 1.0), 1.0)], 2, irradiance_a=0.5)
    assert ic_query(abi, pc(irradianceA=0.5), st5, (0.1, 0, 0), n30) is not None


def test_lookup_at_a_record_centre_dominates():
    # d = 0 with the same normal: the weight is infinite and becomes 1e6
    recs = [((0, 0, 0), (0, 0, 1), (1.0, 0.0, 0.0), 1.0), ((0.05, 0, 0), (0, 0, 1), (0.0, 1.0, 0.0), 1.0)]
    st = scenes.cache_state(recs, 4)
    c = ic_query(abi, pc(irradianceA=0.2), st, (0.0, 0.0, 0.0), (0, 0, 1))
    w2 = 1.0 / 0.05
    assert np.allclose(c, (1e6 / (1e6 + w2), w2 / (1e6 + w2), 0.0), rtol=1e-6)


def test_lookup_visibility_half_space():
    # raytrace.irradiance.rahit:25-27: (x - p) . (n + ni) / 2 >= -0.001
    st = scenes.cache_state([((0, 0, 0), (0, 0, 1), (1.0, 1.0, 1.0), 1.0)], 2)
    below = (0.0, 0.0, -0.05)
    assert ic_query(abi, pc(irradianceA=0.2), st, below, (0, 0, 1)) is not None
    assert ic_query(abi, pc(irradianceA=0.2, irradianceCachePerformVisibilityCheck=1), st, below,
                    (0, 0, 1)) is None
    assert ic_query(abi, pc(irradianceA=0.2, irradianceCachePerformVisibilityCheck=1), st, (0, 0, 0.05),
                    (0, 0, 1)) is not None


def test_lookup_gradients_extrapolate():
    # same normal: colour -> normalize(E) * (|E| + (x - p) . transGrad)
    st = scenes.cache_state([((0, 0, 0), (0, 0, 1), (0.6, 0.8, 0.0), 1.0)], 2)
    st.entries[0]["transGrad"] = (2.0, 0.0, 0.0)
    st.entries[0]["rotGrad"] = (0.0, 5.0, 0.0)
    c = ic_query(abi, pc(irradianceA=0.2, useIrradianceGradients=1), st, (0.1, 0, 0), (0, 0, 1))
    assert np.allclose(c, np.array([0.6, 0.8, 0.0]) * (1.0 + 0.2), rtol=1e-6)
    # a rotated normal picks up (ni x n) . rotGrad
    t = 0.05
    n = (np.sin(t), 0.0, np.cos(t))
    st5 = scenes.cache_state([((0, 0, 0), (0, 0, 1), (0.6, 0.8, 0.0), 1.0)], 2, irradiance_a=0.5)
    st5.entries[0]["rotGrad"] = (0.0, 5.0, 0.0)
    c = ic_query(abi, pc(irradianceA=0.5, useIrradianceGradients=1), st5, (0, 0, 0.0), n)
    cross_y = np.sin(t)  # (0,0,1) x (sin t, 0, cos t) = (0, sin t, 0)
    assert np.allclose(c, np.array([0.6, 0.8, 0.0]) * (1.0 + 5.0 * cross_y), rtol=1e-5)


# ----------------------------------------------------------- stratification
def test_strata_are_cosine_weighted_and_stratified():
    sin2 = np.zeros((N_PHI, N_THETA))
    for k in range(N_PHI):
        for j in range(N_THETA):
            phi, theta, _ = ic_stratum(abi, 1, 0, 3, k, j)
            assert 2 * np.pi * k / N_PHI <= phi < 2 * np.pi * (k + 1) / N_PHI
            sin2[k, j] = np.sin(theta) ** 2
            # sin^2(theta) = (j + u) / M: uniform in its stratum -> cos-weighted directions
            assert j / N_THETA <= sin2[k, j] < (j + 1) / N_THETA
    # the phi jitter is shared by the column, the theta jitter is per stratum
    assert np.unique(np.round(sin2, 12)).size == N_PHI * N_THETA


# --------------------------------------------------------- calculateCacheData
@pytest.mark.parametrize("prec", [64, 32])
def test_irradiance_and_harmonic_distance_under_lit_ceiling(prec):
    h, rho, I = 1.5, 0.6, 2.0
    hs = scenes.ceiling_over_point_light(h, rho, I)
    orc = OracleScene(abi, hs.desc())
    fp = scenes.params(hs, 8, 8, 1, enableMIS=0, useIrradianceCache=1, irradianceNumNEE=1)
    st = rtxamd.IcState.empty(8)
    color, rot, trans, hr = orc.ic_calculate(fp, st, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), job=5, prec=prec)
    want_e = 2 * rho * I / (5 * h * h)
    assert np.allclose(color, want_e, rtol=0.01), (color, want_e)
    assert abs(hr - 1.5 * h) < 0.005 * 1.5 * h, hr
    # symmetric set-up: the gradients are small next to E / h
    assert np.linalg.norm(trans) < 0.05 * want_e / h
    assert np.linalg.norm(rot) < 0.05 * want_e


def test_accumulation_on_constant_inputs():
    n = (0.0, 0.0, 1.0)
    sub = np.zeros((200, 4))
    sub[:, :3] = (0.2, 0.4, 0.6)
    sub[:, 3] = 3.0
    color, rot, trans, hr = ic_reduce(abi, 7, 2, 11, n, sub)
    assert np.allclose(color, np.pi * np.array([0.2, 0.4, 0.6]), rtol=1e-12)  # pi/(MN) * MN * L
    assert np.allclose(trans, 0.0, atol=1e-12)  # no L differences
    assert abs(hr - 3.0) < 1e-12
    # nothing hit: harmonicR -1
    sub[:, 3] = 1e6
    assert ic_reduce(abi, 7, 2, 11, n, sub)[3] == -1.0
    # zero radiance: no gradients
    sub[:, :3] = 0.0
    sub[:, 3] = 2.0
    c, r, t, _ = ic_reduce(abi, 7, 2, 11, n, sub)
    assert np.all(c == 0) and np.all(r == 0) and np.all(t == 0)


# ------------------------------------------------------------ frame upkeep
def test_update_elections_follow_the_next_update_slot_rule():
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    orc = OracleScene(abi, hs.desc())
    recs = [((0.0, 0.0, 0.0), (0, 0, 1), (0.1, 0.1, 0.1), 0.5), ((0.5, 0.5, 0.0), (0, 0, 1), (0.2, 0.2, 0.2), 0.5),
            ((-0.5, 0.5, 0.0), (0, 0, 1), (0.3, 0.3, 0.3), 0.5)]
    st = scenes.cache_state(recs, 20)
    fp = scenes.params(hs, 4, 4, 1, maxDepth=2, useIrradianceCache=1, irradianceUpdateProb=1.0,
                       irradianceCreateProb=0.0)
    rgb, sig, jobs, sub, subsig = orc.ic_frame(fp, st, prec=64)
    # 16 elections over 3 records: 0,1,2, reset, 0,1,2, reset, ... (raytrace.rgen:1335-1340)
    assert [j[0] for j in jobs] == [0] * 12
    assert [j[1] for j in jobs] == [0, 1, 2] * 4
    assert st.hdr.nextUpdateSlot == 0 and st.count == 3
    assert list(st.entries["numUpdates"][:3]) == [5, 5, 5]
    # update jobs start at the record's centre and normal
    assert np.allclose(jobs[1][2], (0.5, 0.5, 0.0)) and np.allclose(jobs[1][3], (0, 0, 1))
    # the running mean keeps the radius tied to harmonicR
    assert np.allclose(st.spheres["radius"][:3], 0.2 * st.entries["harmonicR"][:3], rtol=1e-6)
    # at most maxCaches update jobs a frame: with room for 4 the fifth election is not served
    st4 = scenes.cache_state(recs, 4)
    _, _, jobs4, _, _ = orc.ic_frame(fp, st4, prec=64)
    assert [j[1] for j in jobs4] == [0, 1, 2, 0] and st4.hdr.nextUpdateSlot == 1


def test_creations_fill_free_slots_in_order():
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    orc = OracleScene(abi, hs.desc())
    st = rtxamd.IcState.empty(40)
    fp = scenes.params(hs, 8, 8, 1, maxDepth=3, useIrradianceCache=1, irradianceUpdateProb=0.0,
                       irradianceCreateProb=1.0)
    rgb, sig, jobs, sub, subsig = orc.ic_frame(fp, st, prec=64)
    # every IC-capable miss is a candidate: far more than the 40 free slots
    assert len(jobs) == 40 and all(j[0] == 1 for j in jobs)
    assert st.count == 40 and st.hdr.nextCacheSlot == 40
    # record i is job i's point and normal (closed box: every job sees geometry)
    for i in (0, 17, 39):
        assert np.allclose(st.spheres[i]["center"], jobs[i][2].astype(np.float32))
        assert np.allclose(st.entries[i]["normal"], jobs[i][3].astype(np.float32))
        assert st.entries[i]["harmonicR"] >= 0.1 and st.entries[i]["numUpdates"] == 1
    # the next frame has no room: no creation jobs
    fp.sampleOffset = 1
    _, _, jobs2, _, _ = orc.ic_frame(fp, st, prec=64)
    assert len(jobs2) == 0 and st.count == 40
