"""Per-path comparison helpers shared by the GPU parity tests (not tests)."""
import numpy as np

REL_TOL = 1e-4


def log(*a):
    print(*a, flush=True)


def relerr(gpu, orc):
    return np.abs(gpu - orc).max(1) / np.maximum(np.abs(orc).max(1), 1e-6)


def relmse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))


def path_stats(rgb, sig, orgb, osig):
    same = (sig == osig).all(1)
    rel = relerr(rgb, orgb)
    within = float((rel[same] <= REL_TOL).mean()) if same.any() else 0.0
    return same, rel, within


def check_paths(name, rgb, sig, orgb, osig, f32=None, abs_same=0.0, abs_within=0.0):
    """Device paths against FP64 paths; f32 = (rgb, sig) of the FP32 instantiation."""
    same, rel, within = path_stats(rgb, sig, orgb, osig)
    log(f"{name}: paths {len(rgb)} non-diverged {same.mean():.5f} within 1e-4 {within:.5f} "
        f"p50 {np.median(rel[same]):.2e} p99 {np.percentile(rel[same], 99):.2e} max {rel[same].max():.2e}")
    if f32 is not None:
        s32, r32, w32 = path_stats(f32[0], f32[1], orgb, osig)
        log(f"  reference logic in FP32: non-diverged {s32.mean():.5f} within 1e-4 {w32:.5f}")
        assert same.mean() >= s32.mean() - 0.005
        assert within >= w32 - 0.005
    assert same.mean() >= abs_same
    assert within >= abs_within
    return same, rel


