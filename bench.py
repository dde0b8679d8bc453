"""Benchmark: Msamples/s at 1920x1080, depth 12, on one MI355X.

Workload: the seeded procedural room of 262,144 triangles (boxes and
24-sided columns in a 20x10x20 m volume, 50% Lambert / 30% rough conductor /
20% dielectric, 4 area lights), 1920x1080 at 64 spp, maxDepth 12, NEE + MIS
(power heuristic), irradiance cache of 10,000 records, FP32.

One step = one full image as the reference produces it with the cache on:
the cache is emptied, 50 prepare frames (1 spp each, cache upkeep only,
src/RayTracingApp.cpp:1131-1145 with irradianceCachePrepareFrames = 50) build
it, then 64 samples per pixel are rendered with the cache (creation and
update probabilities at the reference's defaults).  `value` counts the 64
spp only; the prepare frames are inside the timed step.  --method plain
measures the plain path tracer instead (no cache); its figure is reported
beside the cache's in every run.

Prints one JSON line.  `roofline` prices the dominant kernel (kTraceP, the
persistent traversal that traces one bounce's extend rays together with the
previous bounce's probe and shadow rays) by algorithmic bytes per ray = 32
(ray) + 16 (hit) + nodes x node record + triangles x triangle record (the
library's figures: 112 B read of a 128 B four-wide node, 48 B), with node and
triangle fetch counts from a counting pass of the identical (deterministic)
workload and kernel times from HIP events on the library's stream.  `cpu_baseline` times the FP64 CPU
restatement of the plain method on a pixel subsample of one sample plane.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "rtx-pathtracer_amd"), str(ROOT)]

HBM_PEAK_GBS = 8000.0
RAY_BYTES, HIT_BYTES = 32, 16  # ray record in, hit record out; node and triangle records: the library's figures


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def room_triangles(t):
    """The room generator's nearest valid count: 20 + a multiple of 4, >= 124."""
    t = max(124, int(t))
    return 20 + 4 * ((t - 20) // 4)


def serving_leg(rtxamd, abi, r, a):
    """Serving workload: Poisson arrivals at `serve_rate` jobs/s (seeded) over
    16 resident seeded procedural rooms of 10k to 2M triangles (log-uniform);
    per job a resolution from {256^2, 512^2, 1024^2, 1920x1080}, spp from
    {4, 16, 64, 256}, depth 8-12, irradiance-cache lookups on or off and
    guided sampling on or off (each with probability 1/2).  Each scene's cache
    (10 prepare frames at 256^2) and guiding regions (2^4 regions, 3 training
    frames at 128^2) are built before the stream starts and stay resident
    with the BVHs.  The first `serve_jobs` arrivals of the stream are
    submitted at their arrival times; p50/p99 are over their latencies
    (submission to image complete on the device)."""
    rng = np.random.default_rng(40)
    tris = [room_triangles(np.exp(rng.uniform(np.log(1e4), np.log(2e6)))) for _ in range(16)]
    t0 = time.perf_counter()
    hosts, devs, states = [], [], []
    pc = dict(maxDepth=12, enableMIS=1, enableNEE=1, usePowerHeuristic=1)
    for k, t in enumerate(tris):
        hs = rtxamd.HostScene.generate(abi.SCENE_ROOM, 100 + k, t)
        ds = r.upload(hs)
        st = rtxamd.Renderer(0)
        sds = st.upload(hs)
        st.ic_create(10000)
        cam = hs.camera(256, 256)
        st.render(sds, rtxamd.frame_params(cam, 256, 256, 10, seed=7, flags=abi.RTX_FRAME_RESET_ACCUM,
                                           useIrradianceCache=1, isIrradiancePrepareFrame=1, **pc))
        st.guiding_init(sds.bounds(), 4)
        gcam = hs.camera(128, 128)
        st.render(sds, rtxamd.frame_params(gcam, 128, 128, 3, seed=9, flags=abi.RTX_FRAME_RESET_ACCUM,
                                           updateGuiding=1, useGuiding=1, **pc))
        hosts.append(hs)
        devs.append(ds)
        states.append((st, sds))
    setup_s = time.perf_counter() - t0
    log(f"serving: 16 scenes ({min(tris)}..{max(tris)} triangles, {sum(tris)} total) resident, caches and "
        f"guides built in {setup_s:.1f} s")
    res = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]
    spps = [4, 16, 64, 256]
    jobs = []
    t_arr = 0.0
    for _ in range(a.serve_jobs):
        t_arr += rng.exponential(1.0 / a.serve_rate)
        k = int(rng.integers(16))
        w, h = res[int(rng.integers(4))]
        spp = spps[int(rng.integers(4))]
        depth = int(rng.integers(8, 13))
        ic, guide = bool(rng.integers(2)), bool(rng.integers(2))
        fp = rtxamd.frame_params(hosts[k].camera(w, h), w, h, spp, seed=int(rng.integers(1 << 30)), maxDepth=depth,
                                 enableMIS=1, enableNEE=1, usePowerHeuristic=1, useIrradianceCache=int(ic),
                                 useGuiding=int(guide), guidingProb=0.5)
        jobs.append((t_arr, k, fp, ic or guide))
    with rtxamd.Server(r, slotPaths=1 << 18, slots=16, framebufferBytes=1 << 30) as sv:
        ids = []
        t_start = time.perf_counter()
        for t_arr, k, fp, stateful in jobs:
            dt = t_start + t_arr - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append(sv.submit(devs[k], fp, states[k][0] if stateful else None))
        for jid in ids:
            sv.wait(jid)
            sv.read_image(jid)
        wall = time.perf_counter() - t_start
        st = sv.stats()
    paths = sum(fp.width * fp.height * fp.spp for _, _, fp, _ in jobs)
    out = {
        "workload": (f"{a.serve_jobs} jobs of a Poisson stream at {a.serve_rate:g} jobs/s (seed 40) over 16 resident "
                     f"rooms of {min(tris)}-{max(tris)} triangles; resolutions 256^2..1920x1080, spp 4..256, depth "
                     f"8-12, cache and guiding lookups per job"),
        "jobs": a.serve_jobs, "paths": int(paths),
        "p50_ms": round(st.latencyP50Ms, 1), "p99_ms": round(st.latencyP99Ms, 1),
        "max_ms": round(st.latencyMaxMs, 1), "mean_ms": round(st.latencyMeanMs, 1),
        "msamples_per_s": round(st.msamplesPerSec, 2), "wall_s": round(wall, 2),
        "offered_msamples_per_s": round(paths / (jobs[-1][0] * 1e6), 1),
        "peak_slots_busy": st.peakSlotsBusy, "peak_jobs_running": st.peakJobsRunning,
        "bounces": st.bounces, "mrays_per_s": round((st.extendRays + st.shadowRays + st.probeRays) / (wall * 1e6), 1),
        "pool": "16 slots x 2^18 paths", "setup_s": round(setup_s, 1),
    }
    log(f"serving: {a.serve_jobs} jobs, p50 {out['p50_ms']} ms, p99 {out['p99_ms']} ms, "
        f"{out['msamples_per_s']} Msamples/s")
    for (st_, sds) in states:
        sds.close()
        st_.close()
    for ds in devs:
        ds.close()
    return out


def relmse(img, ref, eps=1e-2, ref_var=None):
    """mean over pixels and channels of (I - R)^2 / (R^2 + eps); with ref_var
    (the per-pixel variance of the reference's own estimate) the reference's
    noise is taken out: E[(I - R)^2] = MSE(I) + Var(R) for independent I, R."""
    d = img[..., :3].astype(np.float64) - ref[..., :3]
    num = d * d if ref_var is None else d * d - ref_var[..., :3]
    return float(np.mean(num / (ref[..., :3] ** 2 + eps)))


def equal_time(render_call, budget_s):
    """Calls render_call(first) until budget_s of wall time have passed;
    returns the number of calls and the seconds spent (the last call may end
    past the budget, and its time is counted)."""
    t0 = time.perf_counter()
    n = 0
    while True:
        render_call(n == 0)
        n += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    return n, time.perf_counter() - t0


def relmse_leg(rtxamd, abi, r, a):
    """Equal-time RelMSE against a plain reference, RelMSE = mean over pixels
    and channels of (I - R)^2 / (R^2 + 1e-2) with the reference's own noise
    taken out (two independent half-references A, B: R = (A + B) / 2,
    Var(R) = (A - B)^2 / 4 per pixel).
    Room (the learned-cache configuration): the benchmark scene at 1920x1080,
    depth 12, NEE + MIS; plain, irradiance cache (10,000 records, 50 prepare
    frames inside the budget, per-frame upkeep) and learned cache (hash grid
    16 x 2 x 2^19, 5 x 64 MLP, 2^16 training suffixes per frame, trained from
    scratch inside the budget); reference `relmse_ref_spp` / 4 plain samples
    per pixel per half (1080p keeps the reference short; its variance is
    subtracted).  Plates (the guiding configuration's plates-and-lights scene)
    at 640x360: plain vs guided sampling (2^8 regions of vMF mixtures, 6
    recording frames with a GPU EM update each inside the budget, then guided
    frames); reference `relmse_ref_spp` / 2 samples per half.  Spot room (a
    closed room lit only through a bright spot on its ceiling, the case
    guiding is for): the same plain vs guided comparison at 640x360, then the
    learned cache alone and with guided sampling (the guide's 6 recording
    frames and the cache's training inside the budget).  Each method gets
    `relmse_seconds` of wall time on the room, a tenth of it on the plates
    and half of it on the spot room."""
    pc = dict(maxDepth=12, enableMIS=1, enableNEE=1, usePowerHeuristic=1)
    RESET = abi.RTX_FRAME_RESET_ACCUM

    def reference(ds, cam, W, H, half_spp):
        halves = []
        for h in range(2):
            acc = np.zeros((H, W, 4))
            chunk = min(64, half_spp)
            for k in range(0, half_spp, chunk):
                r.render(ds, rtxamd.frame_params(cam, W, H, chunk, seed=1001 + h, sample_offset=k, flags=RESET, **pc))
                acc += r.read_image(W, H).astype(np.float64) * chunk
            halves.append(acc / half_spp)
        A, B = halves
        return (A + B) / 2, (A - B) ** 2 / 4

    out = {}
    state = {"spp": 0}
    # ---- room: plain, irradiance cache, learned cache
    T = a.relmse_seconds
    W, H = 1920, 1080
    hs = rtxamd.HostScene.generate(abi.SCENE_ROOM, 1)
    ds = r.upload(hs)
    cam = hs.camera(W, H)
    chunk = 1
    t0 = time.perf_counter()
    half = max(1, a.relmse_ref_spp // 4)
    ref, var = reference(ds, cam, W, H, half)
    ref_s = time.perf_counter() - t0
    room = {"scene": f"room 262144 tris, {W}x{H}, depth 12, NEE+MIS",
            "reference": f"2 x {half} spp plain (independent seeds), FP64 accumulation, its variance subtracted "
                         f"({ref_s:.1f} s)",
            "reference_relmse": relmse(ref, ref, ref_var=-var), "budget_s": T}

    def plain_call(first):
        r.render(ds, rtxamd.frame_params(cam, W, H, chunk, seed=1, sample_offset=state["spp"],
                                         flags=RESET if first else 0, **pc))
        state["spp"] += chunk

    state["spp"] = 0
    n, sec = equal_time(plain_call, T)
    e_plain = relmse(r.read_image(W, H), ref, ref_var=var)
    room["plain"] = {"spp": state["spp"], "seconds": round(sec, 3), "relmse": e_plain}

    r.ic_create(10000)
    icpc = dict(pc, useIrradianceCache=1)

    def ic_call(first):
        if first:
            r.ic_reset()
            r.render(ds, rtxamd.frame_params(cam, W, H, 50, seed=1, flags=RESET, isIrradiancePrepareFrame=1, **icpc))
            state["spp"] = 50
        r.render(ds, rtxamd.frame_params(cam, W, H, chunk, seed=1, sample_offset=state["spp"],
                                         flags=RESET if first else 0, **icpc))
        state["spp"] += chunk

    n, sec = equal_time(ic_call, T)
    e_ic = relmse(r.read_image(W, H), ref, ref_var=var)
    room["irradiance_cache"] = {"spp": state["spp"] - 50, "seconds": round(sec, 3), "relmse": e_ic,
                                "improvement": round(e_plain / e_ic, 3)}

    info = ds.info()

    def nrc_call(first):
        if first:
            r.nrc_create(list(info.boundsMin), list(info.boundsMax), seed=5)
            state["spp"] = 0
        r.render(ds, rtxamd.frame_params(cam, W, H, chunk, seed=1, sample_offset=state["spp"],
                                         flags=abi.RTX_FRAME_NRC | (RESET if first else 0), **pc))
        state["spp"] += chunk

    n, sec = equal_time(nrc_call, T)
    e_nrc = relmse(r.read_image(W, H), ref, ref_var=var)
    ni = r.nrc_info()
    room["learned_cache"] = {"spp": state["spp"], "seconds": round(sec, 3), "relmse": e_nrc,
                             "improvement": round(e_plain / e_nrc, 3), "adam_steps": ni.steps,
                             "last_loss": round(ni.lastLoss, 5)}
    r.nrc_destroy()
    ds.close()
    out["room"] = room
    log(f"relmse room: plain {e_plain:.4g}, cache {e_ic:.4g}, learned {e_nrc:.4g}")

    # ---- plates: plain vs guided
    W, H = 640, 360
    T = a.relmse_seconds / 10
    chunk = 2
    hs = rtxamd.HostScene.generate(abi.SCENE_PLATES, 1)
    ds = r.upload(hs)
    cam = hs.camera(W, H)
    t0 = time.perf_counter()
    half = max(1, a.relmse_ref_spp // 2)
    ref, var = reference(ds, cam, W, H, half)
    ref_s = time.perf_counter() - t0
    plates = {"scene": f"plates and sphere lights, {W}x{H}, depth 12, NEE+MIS",
              "reference": f"2 x {half} spp plain (independent seeds), FP64 accumulation, its variance "
                           f"subtracted ({ref_s:.1f} s)", "budget_s": T}
    state["spp"] = 0
    n, sec = equal_time(plain_call, T)
    e_plain = relmse(r.read_image(W, H), ref, ref_var=var)
    plates["plain"] = {"spp": state["spp"], "seconds": round(sec, 3), "relmse": e_plain}
    bounds = ds.bounds()

    def guided_call(first):
        if first:
            r.guiding_init(bounds, 8)
            r.render(ds, rtxamd.frame_params(cam, W, H, 6, seed=1, flags=RESET, updateGuiding=1, useGuiding=1, **pc))
            state["spp"] = 6
        r.render(ds, rtxamd.frame_params(cam, W, H, chunk, seed=1, sample_offset=state["spp"],
                                         flags=RESET if first else 0, useGuiding=1, **pc))
        state["spp"] += chunk

    n, sec = equal_time(guided_call, T)
    e_g = relmse(r.read_image(W, H), ref, ref_var=var)
    plates["guided"] = {"spp": state["spp"] - 6, "seconds": round(sec, 3), "relmse": e_g,
                        "improvement": round(e_plain / e_g, 3)}
    ds.close()
    out["plates"] = plates
    log(f"relmse plates: plain {e_plain:.4g}, guided {e_g:.4g}")

    # ---- spot room: plain vs guided where the light arrives through a bright spot
    T = a.relmse_seconds / 2
    hs = rtxamd.HostScene.generate(abi.SCENE_SPOT, 1)
    ds = r.upload(hs)
    cam = hs.camera(W, H)
    t0 = time.perf_counter()
    ref, var = reference(ds, cam, W, H, half)
    ref_s = time.perf_counter() - t0
    spot = {"scene": f"closed room lit through a ceiling spot, {W}x{H}, depth 12, NEE+MIS",
            "reference": f"2 x {half} spp plain (independent seeds), FP64 accumulation, its variance "
                         f"subtracted ({ref_s:.1f} s)", "budget_s": T}
    state["spp"] = 0
    n, sec = equal_time(plain_call, T)
    e_plain = relmse(r.read_image(W, H), ref, ref_var=var)
    spot["plain"] = {"spp": state["spp"], "seconds": round(sec, 3), "relmse": e_plain}
    bounds = ds.bounds()
    n, sec = equal_time(guided_call, T)
    e_g = relmse(r.read_image(W, H), ref, ref_var=var)
    spot["guided"] = {"spp": state["spp"] - 6, "seconds": round(sec, 3), "relmse": e_g,
                      "improvement": round(e_plain / e_g, 3)}
    sinfo = ds.info()

    def spot_nrc_call(first, guided):
        if first:
            if guided:
                r.guiding_init(bounds, 8)
                r.render(ds, rtxamd.frame_params(cam, W, H, 6, seed=1, flags=RESET, updateGuiding=1, useGuiding=1,
                                                 **pc))
            r.nrc_create(list(sinfo.boundsMin), list(sinfo.boundsMax), seed=5)
            state["spp"] = 6 if guided else 0
        r.render(ds, rtxamd.frame_params(cam, W, H, chunk, seed=1, sample_offset=state["spp"],
                                         flags=abi.RTX_FRAME_NRC | (RESET if first else 0),
                                         useGuiding=1 if guided else 0, **pc))
        state["spp"] += chunk

    for key, guided in (("learned_cache", False), ("guided_learned_cache", True)):
        n, sec = equal_time(lambda first: spot_nrc_call(first, guided), T)
        e_n = relmse(r.read_image(W, H), ref, ref_var=var)
        spot[key] = {"spp": state["spp"] - (6 if guided else 0), "seconds": round(sec, 3), "relmse": e_n,
                     "improvement": round(e_plain / e_n, 3), "adam_steps": r.nrc_info().steps}
        r.nrc_destroy()
    ds.close()
    out["spot"] = spot
    log(f"relmse spot room: plain {e_plain:.4g}, guided {e_g:.4g}, learned {spot['learned_cache']['relmse']:.4g}, "
        f"guided + learned {spot['guided_learned_cache']['relmse']:.4g}")
    return out


def em_cpu_leg(rtxamd, abi, r):
    """The guiding update (PathGuiding::update: preFit, MAP-EM, split and
    merge, distances) on one recording frame of the plates-and-lights scene
    at 640x360 (2^8 regions), timed on the host the reference's way - the FP32
    oracle, serial over regions on one thread (src/PathGuiding.cpp:281-288) -
    and on the device (rtx_guiding_update on the same samples)."""
    from oracle.binding import OracleGuide
    hs = rtxamd.HostScene.generate(abi.SCENE_PLATES, 1)
    ds = r.upload(hs)
    box = ds.bounds()
    W, H = 640, 360
    r.guiding_init(box, 8)
    r.render(ds, rtxamd.frame_params(hs.camera(W, H), W, H, 1, seed=3, updateGuiding=1, maxDepth=12, enableMIS=1,
                                     flags=abi.RTX_FRAME_RESET_ACCUM | abi.RTX_FRAME_NO_GUIDING_UPDATE))
    dd = r.guiding_read_samples(W, H)
    n = int((dd["flags"] != abi.DIR_INVALID).sum())
    t1 = time.perf_counter()
    r.guiding_update()
    r.guiding_get()  # waits for the update
    gpu_s = time.perf_counter() - t1
    og = OracleGuide(abi, box, 8, parallax=True, prec=32)
    t1 = time.perf_counter()
    og.update(dd)
    cpu_s = time.perf_counter() - t1
    r.guiding_set([], [])
    ds.close()
    return {"samples": n, "regions": 256, "cpu_ms_serial": round(cpu_s * 1e3, 1), "gpu_ms": round(gpu_s * 1e3, 2),
            "cpu_msamples_per_s": round(n / cpu_s / 1e6, 3), "gpu_msamples_per_s": round(n / gpu_s / 1e6, 2),
            "sample": "one recording frame of the plates scene at 640x360, 2^8 regions, FP32, first update"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--triangles", type=int, default=262144)
    ap.add_argument("--method", choices=["ic", "plain"], default="ic")
    ap.add_argument("--ic-size", type=int, default=10000)
    ap.add_argument("--prepare-frames", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--trace-blocks", type=int, default=0, help="traversal blocks per CU (0: default)")
    ap.add_argument("--shade-blocks", type=int, default=0, help="shading blocks per CU (0: default)")
    ap.add_argument("--no-count", action="store_true", help="skip the node/triangle counting pass")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-method comparison step")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no autotune)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-jobs", type=int, default=16, help="jobs of the serving leg's arrival stream")
    ap.add_argument("--serve-rate", type=float, default=40.0, help="Poisson arrival rate of the serving leg, jobs/s")
    ap.add_argument("--no-relmse", action="store_true", help="skip the equal-time RelMSE leg")
    ap.add_argument("--relmse-seconds", type=float, default=2.0, help="equal-time budget per method (room)")
    ap.add_argument("--relmse-ref-spp", type=int, default=1024,
                    help="reference samples per pixel (the room renders a quarter of them at 1080p)")
    ap.add_argument("--nrc-mlp", type=int, default=0, metavar="N",
                    help="only run the learned cache's network on N random queries (rocprofv3 MFMA timing)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates rocprofv3 byte counters)")
    a = ap.parse_args()
    if a.nrc_mlp:
        # the learned cache's network alone (rtx_nrc_infer / rtx_nrc_train on
        # random inputs), for rocprofv3 kernel times of its MFMA kernels
        import rtxamd
        r = rtxamd.Renderer(0)
        r.nrc_create([-1.0, -1.0, -1.0], [1.0, 1.0, 1.0], seed=3)
        rng = np.random.default_rng(1)
        x = rng.uniform(-1, 1, (a.nrc_mlp, 16)).astype(np.float32)
        x[:, 3] = rng.uniform(0.05, 1.0, a.nrc_mlp)  # roughness
        for c in (7, 11, 15):
            x[:, c] = 0.0
        for _ in range(5):
            r.nrc_infer(x)
        B = 16384
        t = rng.uniform(0, 1, (B, 3)).astype(np.float32)
        losses = [r.nrc_train(x[:B], t) for _ in range(8)]
        r.close()
        print(json.dumps({"nrc_queries": a.nrc_mlp, "infer_calls": 5, "train_rows": B, "train_steps": 8,
                          "flop_per_query": 2 * (5 * 64 * 64 + 64 * 3), "loss_first": losses[0],
                          "loss_last": losses[-1]}))
        return
    if a.calibrate_copy:
        import torch
        n = a.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, device="cuda", dtype=torch.float32)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU: every GPU box here has one MI355X (asked for --gpus {a.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import rtxamd
    from rtxamd import abi

    W, H, spp = a.width, a.height, a.spp
    hs = rtxamd.HostScene.generate(abi.SCENE_ROOM, 1, a.triangles)
    r = rtxamd.Renderer(0)
    r.set_tuning(traceBlocksPerCU=a.trace_blocks, shadeBlocksPerCU=a.shade_blocks)
    tun = r.tuning()
    trace_kernel = "kTraceP<false> (persistent BVH traversal per bounce: extend, probe and shadow rays)"
    t0 = time.time()
    ds = r.upload(hs)
    info = ds.info()
    log(f"scene: {info.triangles} triangles, {info.nodes} BVH nodes, depth {info.maxDepth}, "
        f"build {info.buildMs:.0f} ms, upload {time.time() - t0:.2f} s")
    cam = hs.camera(W, H)
    pc = dict(maxDepth=12, enableMIS=1, enableNEE=1, usePowerHeuristic=1)
    icpc = dict(pc, useIrradianceCache=1)  # reference defaults: create 0.0005, update 0.00001
    if a.method == "ic":
        r.ic_create(a.ic_size)

    def params(flags, n_spp=spp, **kw):
        return rtxamd.frame_params(cam, W, H, n_spp, seed=1, flags=flags | abi.RTX_FRAME_RESET_ACCUM, **kw)

    def step(flags, method):
        """One image; returns the frame stats of its renders."""
        if method == "plain":
            return [r.render(ds, params(flags, **pc))]
        r.ic_reset()
        prep = r.render(ds, params(flags, a.prepare_frames, isIrradiancePrepareFrame=1, **icpc))
        img = r.render(ds, rtxamd.frame_params(cam, W, H, spp, seed=1, sample_offset=a.prepare_frames,
                                               flags=flags | abi.RTX_FRAME_RESET_ACCUM, **icpc))
        return [prep, img]

    for _ in range(a.warmup):
        step(0, a.method)
    log("warmup done")
    # launch shapes: rtx_autotune on two frames of the workload (outside the timed region)
    tuning = None
    if not a.no_tune and a.trace_blocks == 0 and a.shade_blocks == 0:
        probe_pc = icpc if a.method == "ic" else pc
        probe = rtxamd.frame_params(cam, W, H, 2, seed=1, sample_offset=a.prepare_frames,
                                    flags=abi.RTX_FRAME_RESET_ACCUM, **probe_pc)
        res = r.autotune(ds, probe, repeats=2)
        b = res.best
        tuning = {"traceBlocksPerCU": b.traceBlocksPerCU, "ldsStackEntries": b.ldsStackEntries,
                  "shadeBlocksPerCU": b.shadeBlocksPerCU, "batchPaths": b.batchPaths,
                  "probe": f"2 frames of the workload ({a.method}), {res.trials} trials",
                  "probe_ms_default": round(res.startMs, 2), "probe_ms_tuned": round(res.bestMs, 2)}
        log(f"autotune: {tuning}")
    else:
        t = r.tuning()
        tuning = {"traceBlocksPerCU": t.traceBlocksPerCU, "ldsStackEntries": t.ldsStackEntries,
                  "shadeBlocksPerCU": t.shadeBlocksPerCU, "batchPaths": t.batchPaths, "probe": None}
    # timed region: K full images; each render ends with a stream synchronize
    keys = ("msTotal", "msExtend", "msShadow", "msProbe", "msShade")
    totals = dict.fromkeys(keys, 0.0)
    launches = [0, 0, 0, 0]
    t_start = time.perf_counter()
    for k in range(a.steps):
        # HIP events around the traversal launches only: an event pair around
        # every launch would cost ~5% of the step
        sts = step(abi.RTX_FRAME_TIME_TRACE, a.method)
        for st in sts:
            for key in keys:
                totals[key] += getattr(st, key)
            for c in range(4):
                launches[c] += st.launchesByKind[c]
        log(f"step {k}: {sum(st.msTotal for st in sts):.1f} ms device")
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1e3 / a.steps
    samples = W * H * spp
    value = samples / (ms_per_step * 1e3)

    # counting pass (same deterministic workload) for the algorithmic bytes
    if a.no_count:
        csts = sts
    else:
        csts = step(abi.RTX_FRAME_COUNT_VISITS, a.method)
    cnt = {f: sum(getattr(st, f) for st in csts) for f in ("extendRays", "shadowRays", "probeRays")}
    nodes = [sum(st.nodeVisitsByKind[k] for st in csts) for k in range(3)]
    prims = [sum(st.primVisitsByKind[k] for st in csts) for k in range(3)]
    NODE_BYTES, TRI_BYTES = csts[0].nodeBytes, csts[0].primBytes  # bytes per node / triangle record read
    ext_rays = cnt["extendRays"]
    ext_bytes = ext_rays * (RAY_BYTES + HIT_BYTES) + nodes[0] * NODE_BYTES + prims[0] * TRI_BYTES
    all_rays = cnt["extendRays"] + cnt["shadowRays"] + cnt["probeRays"]
    all_bytes = all_rays * (RAY_BYTES + HIT_BYTES) + sum(nodes) * NODE_BYTES + sum(prims) * TRI_BYTES
    # the dominant kernel: the persistent per-bounce trace covers every ray;
    # the grid-stride build's extend kernel covers the extend rays
    launches_tr = launches[0] // a.steps
    tr_bytes = all_bytes
    tr_rays = all_rays
    tr_nodes = sum(nodes)
    tr_prims = sum(prims)
    avg_launch_ms = totals["msExtend"] / a.steps / max(1, launches_tr)
    achieved = tr_bytes / max(1, launches_tr) / (avg_launch_ms * 1e-3) / 1e9
    trav_ms = (totals["msExtend"] + totals["msShadow"]) / a.steps
    traffic = None
    tfile = ROOT / "profiles" / f"trace_traffic_room262k_1080p_{a.method}.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    # what the HBM itself carries per launch (rocprofv3 FETCH/WRITE of the same
    # workload) at the launch duration measured here: the traversal's
    # algorithmic bytes are mostly served by L2, so this is far below `achieved`
    hbm_gbs = round(traffic / (avg_launch_ms * 1e-3) / 1e9, 1) if traffic else None
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "kernel": trace_kernel,
        "algorithmic_bytes_per_launch": int(tr_bytes / max(1, launches_tr)),
        "avg_launch_ms": round(avg_launch_ms, 4), "launches_per_step": launches_tr,
        "rays_per_launch": int(tr_rays / max(1, launches_tr)),
        "nodes_per_ray": round(tr_nodes / max(1, tr_rays), 2),
        "tris_per_ray": round(tr_prims / max(1, tr_rays), 2),
        "all_traversal_frac": round(all_bytes / (trav_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "hbm_traffic_gbs": hbm_gbs,
        "note": "achieved = algorithmic bytes (every node and triangle a ray reads, as if from HBM) per launch "
                "over the launch time; the kernel is bound by dependent-load latency and VALU issue, with "
                "the BVH served mostly from L2 (hbm_traffic_gbs is what HBM carries)",
    }
    mrays = all_rays / (ms_per_step * 1e3)

    # the other method, one step, for comparison
    other = "plain" if a.method == "ic" and not a.no_plain else None
    other_value = None
    if other:
        step(0, other)
        t1 = time.perf_counter()
        step(0, other)
        other_value = samples / ((time.perf_counter() - t1) * 1e6)
        log(f"plain method: {other_value:.2f} Msamples/s")

    # FP32 device vs FP64 oracle on one sample plane, pixel subsample
    cpu = None
    parity = None
    if not a.no_cpu:
        from oracle.binding import OracleScene, DEFAULT_THREADS
        orc = OracleScene(abi, hs.desc())
        fp1 = rtxamd.frame_params(cam, W, H, 1, seed=1,
                                  flags=abi.RTX_FRAME_RESET_ACCUM | abi.RTX_FRAME_KEEP_PATHS, **pc)  # plain
        r.render(ds, fp1)
        rgb, sig = r.read_paths(W * H)
        # calibrate: time a small sample, then size the measured sample to the budget
        probe_ids = np.arange(0, W * H, 997, dtype=np.uint64)
        t1 = time.perf_counter()
        orc.render_paths(fp1, probe_ids, threads=DEFAULT_THREADS)
        rate = len(probe_ids) / (time.perf_counter() - t1)
        n = int(min(W * H, max(2000, rate * a.cpu_seconds)))
        stride = max(1, (W * H) // n)
        ids = np.arange(0, W * H, stride, dtype=np.uint64)
        t1 = time.perf_counter()
        orgb, osig = orc.render_paths(fp1, ids, threads=DEFAULT_THREADS)
        cpu_s = time.perf_counter() - t1
        cpu = {"value": round(len(ids) / cpu_s / 1e6, 5), "unit": "Msamples/s", "cores": DEFAULT_THREADS,
               "kind": "port",
               "sample": f"{len(ids)} paths = every {stride}th pixel of sample 0 of the same scene and camera, "
                         f"plain method, FP64 restatement, {DEFAULT_THREADS} threads, {cpu_s:.1f} s"}
        def stats(x, xs, ref, refs):
            same = (xs == refs).all(1)
            rel = np.abs(x - ref).max(1) / np.maximum(np.abs(ref).max(1), 1e-6)
            return {"paths": int(len(x)), "non_diverged": round(float(same.mean()), 5),
                    "within_1e-4": round(float((rel[same] <= 1e-4).mean()), 5),
                    "rel_err_p50": float(np.median(rel[same])), "rel_err_p99": float(np.percentile(rel[same], 99))}

        sub = ids[::8]
        t1 = time.perf_counter()
        r32, s32 = orc.render_paths(fp1, sub, prec=32, threads=DEFAULT_THREADS)
        cpu32_s = time.perf_counter() - t1
        cpu["fp32_value"] = round(len(sub) / cpu32_s / 1e6, 5)
        cpu["fp32_sample"] = f"{len(sub)} paths (every 8th of the FP64 sample), FP32 instantiation, {cpu32_s:.1f} s"
        cpu["host_cpus"] = os.cpu_count()
        cpu["threads_note"] = f"{DEFAULT_THREADS} threads: the host cores one GPU's share of the box grants"
        cpu["em"] = em_cpu_leg(rtxamd, abi, r)
        parity = {"device_vs_fp64": stats(rgb[ids.astype(np.int64)], sig[ids.astype(np.int64)], orgb, osig),
                  "reference_logic_in_fp32_vs_fp64": stats(r32, s32, orgb[::8], osig[::8]),
                  "note": "FP32 shading normals on 2-8 cm smooth columns at ~10 m coordinates make ~20% of "
                          "13-bounce paths branch differently; the CPU FP32 run of the same logic does the same"}

    serving = None if a.no_serve else serving_leg(rtxamd, abi, r, a)
    eq = None if a.no_relmse else relmse_leg(rtxamd, abi, r, a)

    out = {
        "metric": "Msamples/s at 1080p depth 12; equal-time RelMSE vs plain; p99 job latency",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene, PCG32 seed 1)",
        "config": {
            "workload": (f"room {info.triangles} tris, {W}x{H}, {spp} spp, depth 12, NEE+MIS, " +
                         (f"irradiance cache {a.ic_size} records ({a.prepare_frames} prepare frames in the step)"
                          if a.method == "ic" else "plain method")),
            "triangles": info.triangles, "width": W, "height": H, "spp": spp, "max_depth": 12,
            "nee": True, "mis": "power",
            "irradiance_cache": ({"size": a.ic_size, "prepare_frames": a.prepare_frames,
                                  "create_prob": 0.0005, "update_prob": 0.00001} if a.method == "ic" else None),
            "samples_per_step": samples,
        },
        "plain_method_msamples_per_s": round(other_value, 3) if other_value else None,
        "mrays_per_s": round(mrays, 1),
        "rays_per_step": cnt,
        "ms_device_per_step": {"total": round(totals["msTotal"] / a.steps, 1),
                               "trace": round(trav_ms, 1),
                               "other": round(totals["msTotal"] / a.steps - trav_ms, 1)},
        "tuning": tuning,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp32_vs_fp64": parity,
        "serving": serving,
        "equal_time_relmse": eq,
    }
    print(json.dumps(out), flush=True)
    ds.close()
    r.close()


if __name__ == "__main__":
    main()
