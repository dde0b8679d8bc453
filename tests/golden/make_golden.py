"""Regenerates the fixtures in tests/golden/.

  push_constant_fields.txt  field names of the reference's push-constant block,
                            read from shaders/raycommon.glsl (run here, where
                            the reference is mounted at /root/reference)
  oracle_c1.npz             per-path FP64 oracle output for the Cornell scene
                            at 64x64 pixels, 2 samples, maxDepth 8, NEE + MIS,
                            Russian roulette from depth 3, seed 1

Usage: python tests/golden/make_golden.py [--reference /root/reference]
"""
import argparse
import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path[:0] = [str(ROOT / "rtx-pathtracer_amd"), str(ROOT)]


def push_constant_fields(reference):
    text = (Path(reference) / "shaders" / "raycommon.glsl").read_text()
    body = re.search(r"struct pushConstant\s*\{(.*?)\};", text, re.S).group(1)
    names = re.findall(r"^\s*\w+\s+(\w+)\s*;", body, re.M)
    (HERE / "push_constant_fields.txt").write_text("\n".join(names) + "\n")
    return names


C1_CONFIG = dict(kind=1, width=64, height=64, spp=2, seed=1, maxDepth=8, enableMIS=1, enableRR=1, rr_start_depth=3)


def c1_params():
    import rtxamd
    from rtxamd import abi
    c = C1_CONFIG
    hs = rtxamd.HostScene.generate(c["kind"], c["seed"])
    cam = hs.camera(c["width"], c["height"])
    fp = rtxamd.frame_params(cam, c["width"], c["height"], c["spp"], seed=c["seed"], rr_start_depth=c["rr_start_depth"],
                             maxDepth=c["maxDepth"], enableMIS=c["enableMIS"], enableRR=c["enableRR"],
                             flags=abi.RTX_FRAME_RESET_ACCUM | abi.RTX_FRAME_KEEP_PATHS)
    return hs, fp


def oracle_c1():
    from rtxamd import abi
    from oracle.binding import OracleScene
    hs, fp = c1_params()
    n = fp.width * fp.height * fp.spp
    rgb, sig = OracleScene(abi, hs.desc()).render_paths(fp, np.arange(n, dtype=np.uint64))
    np.savez_compressed(HERE / "oracle_c1.npz", rgb=rgb, sig=sig)
    return rgb, sig


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    a = ap.parse_args()
    if Path(a.reference).exists():
        print(len(push_constant_fields(a.reference)), "push-constant fields")
    oracle_c1()
    print("oracle_c1.npz written")
