"""Guided sampling on the device (getNewDirection with the region mixtures,
shaders/raytrace.rgen:923-960, shaders/guiding.glsl) against the oracle path
by path, and unbiasedness of the guided estimator at image level."""
import numpy as np
import pytest

import rtxamd
from rtxamd import abi
from oracle.binding import OracleScene
import scenes
from parity import check_paths, log

pytestmark = pytest.mark.gpu


def regions_and_mixtures(parallax):
    regions, vmms = [], []
    rng = np.random.default_rng(4)
    for ix in range(2):
        for iy in range(2):
            for iz in range(2):
                lo = [-1.01 + ix * 1.01, -1.01 + iy * 1.01, -0.01 + iz * 1.01]
                hi = [ix * 1.01 - 0.0, iy * 1.01, iz * 1.01 + 1.0]
                regions.append([lo, hi])
                lobes = [rtxamd.vmf_theta((0, 0, 1), 12.0, distance=1.5 if parallax else -1.0,
                                          target=(0.0, 0.0, 1.98)),
                         rtxamd.vmf_theta(rng.normal(size=3), float(rng.uniform(1, 40))),
                         rtxamd.vmf_theta((0, 1, 0), 0.0)]
                vmms.append(rtxamd.vmm_theta(lobes, [0.5, 0.3, 0.2]))
    return regions, vmms


@pytest.mark.parametrize("parallax", [0, 1])
def test_guided_paths_match_oracle(renderer, parallax):
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    regions, vmms = regions_and_mixtures(parallax)
    renderer.guiding_set(regions, vmms)
    orc = OracleScene(abi, hs.desc())
    orc.guiding_set(regions, vmms)
    ds = renderer.upload(hs)
    w = h = 128
    fp = scenes.params(hs, w, h, 2, maxDepth=8, enableMIS=1, useGuiding=1, guidingProb=0.6,
                       useParallaxCompensation=parallax,
                       flags=abi.RTX_FRAME_KEEP_PATHS | abi.RTX_FRAME_RESET_ACCUM)
    renderer.render(ds, fp)
    rgb, sig = renderer.read_paths(w * h * 2)
    ids = np.arange(w * h * 2, dtype=np.uint64)
    orgb, osig = orc.render_paths(fp, ids)
    frgb, fsig = orc.render_paths(fp, ids, prec=32)
    check_paths(f"guided paths (parallax {parallax})", rgb, sig, orgb, osig, (frgb, fsig), abs_same=0.97,
                abs_within=0.99)
    renderer.guiding_set([], [])
    ds.close()


def test_guided_image_mean_equals_plain(renderer):
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    regions, vmms = regions_and_mixtures(1)
    renderer.guiding_set(regions, vmms)
    ds = renderer.upload(hs)
    w = h = 128
    base = dict(maxDepth=8, enableMIS=1, flags=abi.RTX_FRAME_RESET_ACCUM | abi.RTX_FRAME_KEEP_PATHS)
    n = w * h * 64
    renderer.render(ds, scenes.params(hs, w, h, 64, seed=3, **base))
    a, _ = renderer.read_paths(n)
    renderer.render(ds, scenes.params(hs, w, h, 64, seed=4, useGuiding=1, guidingProb=0.5,
                                      useParallaxCompensation=1, **base))
    b, _ = renderer.read_paths(n)
    la, lb = a.mean(1).astype(np.float64), b.mean(1).astype(np.float64)
    se = np.sqrt(la.var() / n + lb.var() / n)
    log(f"plain {la.mean():.5f} guided {lb.mean():.5f} (se {se:.1e})")
    assert abs(la.mean() - lb.mean()) < 4 * se
    renderer.guiding_set([], [])
    ds.close()


def test_guiding_errors(renderer):
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    ds = renderer.upload(hs)
    renderer.guiding_set([], [])
    with pytest.raises(rtxamd.RtxError) as e:
        renderer.render(ds, scenes.params(hs, 8, 8, 1, useGuiding=1))
    assert e.value.code == abi.RTX_E_STATE
    bad = rtxamd.vmm_theta([rtxamd.vmf_theta((0, 0, 1), 1.0)], [1.0])
    bad.usedDistributions = 17
    with pytest.raises(rtxamd.RtxError):
        renderer.guiding_set([[[0, 0, 0], [1, 1, 1]]], [bad])
    ds.close()


@pytest.mark.parametrize("guided", [0, 1])
def test_sample_records_match_oracle(renderer, guided):
    """updateGuiding: the DirectionalData slots of every pixel (position,
    direction, pdf, weight = |light sum| / pdf, distance, region) against the
    oracle's for the paths that take the same branches."""
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    regions, vmms = regions_and_mixtures(1)
    renderer.guiding_set(regions, vmms)
    orc = OracleScene(abi, hs.desc())
    orc.guiding_set(regions, vmms)
    ds = renderer.upload(hs)
    w = h = 96
    fp = scenes.params(hs, w, h, 1, maxDepth=8, enableMIS=1, updateGuiding=1, useGuiding=guided, guidingProb=0.5,
                       sample_offset=5, flags=abi.RTX_FRAME_KEEP_PATHS | abi.RTX_FRAME_RESET_ACCUM)
    renderer.render(ds, fp)
    rgb, sig = renderer.read_paths(w * h)
    dd = renderer.guiding_read_samples(w, h)
    ids = np.arange(w * h, dtype=np.uint64)
    orgb, osig, odd = orc.render_paths_record(fp, ids)
    same, _ = check_paths(f"recording paths (guided {guided})", rgb, sig, orgb, osig, abs_same=0.97,
                          abs_within=0.99)
    g, o = dd[same], odd[same]
    gv, ov = g["flags"] != abi.DIR_INVALID, o["flags"] != abi.DIR_INVALID
    assert np.array_equal(gv, ov)
    assert np.array_equal(g["flags"][gv], o["flags"][ov])
    log(f"  valid records {gv.sum()} in {same.sum()} paths")
    # Bars from the spread of the oracle's own FP32 run against its FP64 run on
    # this scene: FP32 drift over several bounces puts about 0.3% of positions
    # and 0.4% of distances/weights outside the per-record bound, never by more
    # than 4e-3 in position (scene size 2) or 4% in weight.
    for f in ("position", "direction"):
        a, b = g[f][gv].astype(np.float64), o[f][ov].astype(np.float64)
        ok = np.all(np.abs(a - b) <= 1e-4 * np.abs(b) + 2e-5, axis=1)
        log(f"  {f}: within 1e-4 {ok.mean():.5f} max abs err {np.abs(a - b).max():.2e}")
        assert ok.mean() >= 0.995 and np.abs(a - b).max() < 1e-2, f
    for f, frac in (("pdf", 0.995), ("distance", 0.99), ("weight", 0.99)):
        a, b = g[f][gv].astype(np.float64), o[f][ov].astype(np.float64)
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-6)
        ok = np.abs(a - b) <= 1e-4 * np.abs(b) + 1e-5
        log(f"  {f}: within 1e-4 {ok.mean():.5f} rel err p50 {np.median(rel):.2e} max {rel.max():.2e}")
        assert ok.mean() >= frac and np.percentile(rel, 99.9) < 0.1, f
    renderer.guiding_set([], [])
    ds.close()


def mixture_pdf_grid(vmm, dirs, x):
    """The mixture pdf (guiding.glsl VMM, with parallax) at each direction."""
    K = vmm.usedDistributions
    out = np.zeros(len(dirs))
    for k in range(K):
        t = vmm.thetas[k]
        if t.k == 0:
            out += vmm.pi[k] / (4 * np.pi)
            continue
        mu = np.array(t.mu[:], np.float64)
        if t.distance > 0:
            mu = np.array(t.target[:]) - x
            mu /= np.linalg.norm(mu)
        out += vmm.pi[k] * t.norm * np.exp(t.k * (dirs @ mu - 1))
    return out


def fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    phi = np.pi * (1 + 5 ** 0.5) * i
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


@pytest.mark.parametrize("split_merge", [0, 1])
def test_guiding_update_matches_oracle_em(renderer, split_merge):
    """Three training frames: the device's region mixtures after each update
    against the oracle's on the same recorded samples (first update: fit,
    later: updateFit), with and without postFit's component merge and split.
    Compared as densities over the sphere from each region's mean position,
    and by component count; the FP32 run of the oracle is the yardstick."""
    from oracle.binding import OracleGuide
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    ds = renderer.upload(hs)
    box = ds.bounds()
    opts = abi.guiding_options(splitAndMerge=split_merge)
    renderer.set_guiding_options(opts)
    renderer.guiding_init(box, 3, parallax=True)
    og = OracleGuide(abi, box, 3, parallax=True, prec=64, options=opts)
    o32 = OracleGuide(abi, box, 3, parallax=True, prec=32, options=opts)
    w = h = 128
    dirs = fib_sphere(4000)
    try:
        for frame in range(3):
            fp = scenes.params(hs, w, h, 1, maxDepth=8, enableMIS=1, updateGuiding=1, useGuiding=int(frame > 0),
                               guidingProb=0.5, sample_offset=frame)
            renderer.render(ds, fp)  # records the frame, then updates the mixtures
            dd = renderer.guiding_read_samples(w, h)
            og.update(dd)
            o32.update(dd)
            boxes, gv = renderer.guiding_get()
            _, ov = og.regions()
            _, fv = o32.regions()
            errs, errs32, sameK, sameK32 = [], [], [], []
            for r in range(len(gv)):
                if (dd["flags"] == r).sum() < 50:
                    continue
                x = np.array(ov[r].meanPosition[:], np.float64)
                pg, po, pf = (mixture_pdf_grid(v, dirs, x) for v in (gv[r], ov[r], fv[r]))
                errs.append(np.abs(pg - po).mean() / po.mean())
                errs32.append(np.abs(pf - po).mean() / po.mean())
                sameK.append(gv[r].usedDistributions == ov[r].usedDistributions)
                sameK32.append(fv[r].usedDistributions == ov[r].usedDistributions)
                assert np.allclose(gv[r].meanPosition[:], ov[r].meanPosition[:], atol=1e-6)
            errs, errs32 = np.array(errs), np.array(errs32)
            ks = [(gv[r].usedDistributions, ov[r].usedDistributions) for r in range(len(gv))]
            log(f"update {frame} (split/merge {split_merge}): {len(errs)} regions, K device/oracle {ks}; density L1 vs "
                f"FP64 oracle: device median {np.median(errs):.2e} max {errs.max():.2e}; FP32 oracle median "
                f"{np.median(errs32):.2e} max {errs32.max():.2e}")
            assert np.mean(sameK) >= np.mean(sameK32) - 0.13
            assert np.median(errs) <= max(1e-3, 2 * np.median(errs32))
            assert errs.max() <= max(0.05, 2 * errs32.max())
            if not split_merge:
                assert all(sameK)
    finally:
        renderer.set_guiding_options()
        renderer.guiding_set([], [])
        ds.close()


def test_region_splitting_matches_oracle(renderer):
    """splitRegions: busy regions are halved after each update; the device's
    boxes equal the oracle's and the split regions' mixtures match."""
    from oracle.binding import OracleGuide
    hs = rtxamd.HostScene.generate(abi.SCENE_CORNELL)
    ds = renderer.upload(hs)
    box = ds.bounds()
    opts = abi.guiding_options(splitRegions=1, samplesForRegionSplit=12000)
    renderer.set_guiding_options(opts)
    renderer.guiding_init(box, 2, parallax=True)
    og = OracleGuide(abi, box, 2, parallax=True, prec=64, options=opts)
    w = h = 96
    dirs = fib_sphere(2000)
    try:
        for frame in range(3):
            fp = scenes.params(hs, w, h, 1, maxDepth=6, enableMIS=1, updateGuiding=1, useGuiding=int(frame > 0),
                               guidingProb=0.5, sample_offset=20 + frame)
            renderer.render(ds, fp)
            dd = renderer.guiding_read_samples(w, h)
            og.update(dd)
            gb, gv = renderer.guiding_get()
            ob, ov = og.regions()
            log(f"frame {frame}: {len(gb)} regions on the device, {len(ob)} in the oracle")
            assert len(gb) == len(ob) and np.array_equal(np.asarray(gb, np.float32), ob)
            errs = []
            for r in range(len(gv)):
                x = np.array(ov[r].meanPosition[:], np.float64)
                pg, po = mixture_pdf_grid(gv[r], dirs, x), mixture_pdf_grid(ov[r], dirs, x)
                errs.append(np.abs(pg - po).mean() / po.mean())
            log(f"  density L1 vs FP64 oracle: median {np.median(errs):.2e} max {np.max(errs):.2e}")
            assert np.median(errs) < 1e-3 and np.max(errs) < 0.05
        assert len(gb) > 4  # regions were split
    finally:
        renderer.set_guiding_options()
        renderer.guiding_set([], [])
        ds.close()


def test_trained_guiding_is_unbiased_and_reduces_variance(renderer):
    """Guiding trained online (four recording frames, each refitting the
    region mixtures) on a room lit only through a bright spot on its ceiling
    (scenes.spot_room): the guided estimator keeps the plain image's mean
    (within 4 standard errors) and lowers the mean per-pixel variance of 32
    samples by at least 30% (measured: 0.50 of plain), a drop of more than 4
    standard errors of the per-pixel differences."""
    hs = scenes.spot_room()
    ds = renderer.upload(hs)
    renderer.guiding_init(ds.bounds(), 6, parallax=True)
    w, h, spp = 160, 120, 32
    base = dict(maxDepth=6, enableMIS=1)
    for f in range(4):
        renderer.render(ds, scenes.params(hs, w, h, 1, updateGuiding=1, useGuiding=int(f > 0), guidingProb=0.5,
                                          sample_offset=1000 + f, **base))
    n = w * h * spp
    keep = abi.RTX_FRAME_KEEP_PATHS | abi.RTX_FRAME_RESET_ACCUM
    renderer.render(ds, scenes.params(hs, w, h, spp, seed=11, flags=keep, **base))
    a, _ = renderer.read_paths(n)
    renderer.render(ds, scenes.params(hs, w, h, spp, seed=12, flags=keep, useGuiding=1, guidingProb=0.5, **base))
    b, _ = renderer.read_paths(n)
    la, lb = a.mean(1).astype(np.float64), b.mean(1).astype(np.float64)
    se = np.sqrt(la.var() / n + lb.var() / n)
    # paths are stored sample plane by sample plane: (spp, pixels)
    va = la.reshape(spp, -1).var(0, ddof=1)
    vb = lb.reshape(spp, -1).var(0, ddof=1)
    d = va - vb
    se_d = d.std() / np.sqrt(d.size)
    log(f"spot room: plain mean {la.mean():.5f} guided {lb.mean():.5f} (se {se:.1e}); mean per-pixel variance plain "
        f"{va.mean():.4f} guided {vb.mean():.4f} (ratio {vb.mean() / va.mean():.3f}, drop {d.mean() / se_d:.1f} se)")
    assert abs(la.mean() - lb.mean()) < 4 * se
    assert vb.mean() <= 0.7 * va.mean() and d.mean() > 4 * se_d
    renderer.guiding_set([], [])
    ds.close()


def mixture_faults(vmm):
    """Broken invariants of a fitted mixture (an empty list when it is valid):
    weights positive, kappa >= 0, unit means.  That the weights sum to 1 is
    not one of them: the refit folds the previous fit back in through the
    weight prior's removal, clamped at zero for lobes below the prior
    (removeWeightPrior), so the sums drift; they are checked against the
    oracle's instead."""
    K = vmm.usedDistributions
    w = np.array(vmm.pi[:K], np.float64)
    mu = np.array([vmm.thetas[k].mu[:] for k in range(K)], np.float64)
    kap = np.array([vmm.thetas[k].k for k in range(K)], np.float64)
    faults = []
    if not (w > 0).all():
        faults.append(f"weights {w[~(w > 0)]}")
    if not (kap >= 0).all():
        faults.append(f"kappa {kap[~(kap >= 0)]}")
    norm = np.linalg.norm(mu, axis=1)
    off = ~(np.abs(norm - 1) <= 1e-4)
    if off.any():
        faults.append(f"|mu| {norm[off]} at kappa {kap[off]}, weight {w[off]}")
    return faults


def compare_guides(dev, orc, dirs):
    """(regions whose lobe counts agree, density L1 of each such region
    relative to the oracle's mean density, weight-sum difference of each)."""
    (db, dv), (ob, ov) = dev, orc
    assert len(db) == len(ob) and np.allclose(db, ob, atol=1e-5)
    same, l1, dw = [], [], []
    for r in range(len(dv)):
        if dv[r].usedDistributions != ov[r].usedDistributions:
            continue
        same.append(r)
        x = np.array(ov[r].meanPosition[:], np.float64)
        pg, po = mixture_pdf_grid(dv[r], dirs, x), mixture_pdf_grid(ov[r], dirs, x)
        l1.append(np.abs(pg - po).mean() / po.mean())
        K = dv[r].usedDistributions
        dw.append(sum(dv[r].pi[:K]) - sum(ov[r].pi[:K]))
    return np.array(same), np.array(l1), np.array(dw)


def test_guiding_at_configuration_3_scale(renderer):
    """Configuration 3's guiding scale on its plates-and-lights scene at
    1280x720: GUIDING_SPLITS = 8 (256 regions), region splitting on with a cap
    of 4096 regions, mixtures of up to 32 lobes, eight recording frames each
    followed by an EM update on at most 2^22 samples (the first valid ones in
    slot order; the first update caps at 2^17, fewer than the frame records,
    so that the cap is exercised).  The FP64 oracle (OracleGuide, same options)
    refits from the same samples alongside:
      * after the first update (whose region splitting already adds regions):
        the same regions, the same lobe counts in every region, densities within 1e-3 (median) / 1e-2 (99th percentile)
        L1;
      * after eight updates: the same regions, grown past 256 and within 4096;
        lobe counts equal in 95% of them, and there densities within 1e-2
        (median) L1 and weight sums within 1e-3 of the oracle's; more than 16
        lobes somewhere and never more than 32; every mixture has positive
        weights, kappa >= 0 and unit means; rtx_guiding_get refuses the wide
        mixtures, rtx_guiding_get_wide returns them; a guided frame renders
        finite and keeps the plain mean."""
    from oracle.binding import OracleGuide
    hs = rtxamd.HostScene.generate(abi.SCENE_PLATES)
    ds = renderer.upload(hs)
    box = ds.bounds()
    w, h = 1280, 720
    first_cap, cap = 1 << 17, 1 << 22

    def options(c):
        return abi.guiding_options(splitRegions=1, maxRegions=4096, maxComponents=32, maxSamplesPerUpdate=c)

    base = dict(maxDepth=12, enableMIS=1)
    renderer.set_guiding_options(options(first_cap))
    renderer.guiding_init(box, 8, parallax=True)
    og = OracleGuide(abi, box, 8, parallax=True, prec=64, options=options(first_cap))
    dirs = fib_sphere(4000)
    try:
        for f in range(8):
            renderer.render(ds, scenes.params(hs, w, h, 1, updateGuiding=1, useGuiding=int(f > 0), guidingProb=0.5,
                                              sample_offset=f, flags=abi.RTX_FRAME_NO_GUIDING_UPDATE, **base))
            dd = renderer.guiding_read_samples(w, h).reshape(-1)
            valid = dd[dd["flags"] != abi.DIR_INVALID]
            renderer.guiding_update()
            og.update(valid)
            if f == 0:
                log(f"configuration 3: {len(valid)} valid samples of {len(dd)} slots; the first update fits "
                    f"{min(len(valid), first_cap)}")
                assert len(valid) > first_cap
                gb, gv = renderer.guiding_get_wide()
                same, l1, dw = compare_guides((gb, gv), og.regions_wide(), dirs)
                log(f"  first update: {len(same)} of {len(gb)} regions with equal lobe counts; density L1 median "
                    f"{np.median(l1):.2e} p99 {np.quantile(l1, 0.99):.2e} max {l1.max():.2e}")
                assert len(same) == len(gb)
                assert np.median(l1) <= 1e-3 and np.quantile(l1, 0.99) <= 1e-2
                renderer.set_guiding_options(options(cap))
                og.set_options(options(cap))
        gb, gv = renderer.guiding_get_wide()
        ob, ov = og.regions_wide()
        ks = np.array([v.usedDistributions for v in gv])
        same, l1, dw = compare_guides((gb, gv), (ob, ov), dirs)
        osum = np.array([sum(v.pi[:v.usedDistributions]) for v in ov])
        log(f"  after 8 updates: {len(gb)} regions, lobes per mixture {ks.min()}..{ks.max()} "
            f"(> 16 in {(ks > 16).sum()} regions); equal lobe counts in {len(same)}, density L1 median "
            f"{np.median(l1):.2e} p95 {np.quantile(l1, 0.95):.2e}; weight sums: oracle's within "
            f"{np.abs(osum - 1).max():.2e} of 1, device's within {np.abs(dw).max():.2e} of the oracle's")
        assert 256 < len(gb) <= 4096
        assert len(same) >= 0.95 * len(gb)
        assert np.median(l1) <= 1e-2 and np.abs(dw).max() <= 1e-3
        assert ks.max() <= 32 and (ks > 16).any()
        faults = [(r, f) for r, v in enumerate(gv) if (f := mixture_faults(v))]
        assert not faults, faults[:4]
        with pytest.raises(rtxamd.RtxError) as e:
            renderer.guiding_get()
        assert e.value.code == abi.RTX_E_UNSUPPORTED
        # a guided frame with the wide mixtures: finite, the plain image's mean
        n = 320 * 180 * 8
        keep = abi.RTX_FRAME_KEEP_PATHS | abi.RTX_FRAME_RESET_ACCUM
        renderer.render(ds, scenes.params(hs, 320, 180, 8, seed=21, flags=keep, **base))
        a, _ = renderer.read_paths(n)
        renderer.render(ds, scenes.params(hs, 320, 180, 8, seed=22, flags=keep, useGuiding=1, guidingProb=0.5,
                                          **base))
        b, _ = renderer.read_paths(n)
        la, lb = a.mean(1).astype(np.float64), b.mean(1).astype(np.float64)
        se = np.sqrt(la.var() / n + lb.var() / n)
        log(f"  guided 32-lobe frame: mean {lb.mean():.5f} vs plain {la.mean():.5f} (se {se:.1e})")
        assert np.isfinite(b).all() and abs(la.mean() - lb.mean()) < 4 * se
    finally:
        renderer.set_guiding_options()
        renderer.guiding_set([], [])
        ds.close()
