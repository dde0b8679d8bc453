"""The command line (bin/rtx_pathtracer, after src/main.cpp:8-35): the
reference's positional arguments WIDTH HEIGHT IC_SIZE GUIDING_SPLITS
scenes... and this build's options.  Argument errors are caught before any
device is touched; the GPU run lives in test_gpu_schedule.py."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent.parent / "rtx-pathtracer_amd" / "bin" / "rtx_pathtracer"


def run(*args):
    return subprocess.run([str(BIN), *map(str, args)], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("args", [
    (),                                        # nothing
    (64, 64, 0, 0),                            # no scene (src/main.cpp:9-12 needs more than 5 args)
    ("x", 64, 0, 0, "gen:cornell"),            # WIDTH not a number
    (64, 64, -1, 0, "gen:cornell"),            # IC_SIZE not a count
    (64, 64, 0, 13, "gen:cornell"),            # more than 2^12 guiding regions
    (64, 64, 0, 0, "gen:cornell", "--ic", 1),  # the cache needs IC_SIZE > 0
    (64, 64, 0, 0, "gen:cornell", "--bogus"),  # unknown option
    (64, 64, 0, 0, "gen:cornell", "--spp", 0),
])
def test_bad_arguments_print_usage(args):
    r = run(*args)
    assert r.returncode == 2
    assert "usage: rtx_pathtracer WIDTH HEIGHT IC_SIZE GUIDING_SPLITS" in r.stderr


def test_valid_arguments_reach_the_device():
    """Well-formed arguments get past parsing; without a HIP device the tool
    says so (on a GPU box it renders, see test_gpu_schedule.py)."""
    import rtxamd  # noqa: F401
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present: the GPU test covers the render")
    r = run(64, 64, 100, 2, "gen:cornell", "--ic", 1, "--guiding", 1, "--spp", 2)
    assert r.returncode == 1 and "no HIP device" in r.stderr
