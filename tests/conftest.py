import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "rtx-pathtracer_amd"
for p in (str(PKG), str(ROOT), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    # the libraries are built by __graft_entry__.build(); build them here only if absent
    need = [PKG / "lib" / "librtxhost.so", PKG / "lib" / "librtxamd.so"]
    if not all(p.exists() for p in need):
        subprocess.run(["make", "-C", str(PKG), "-j8"], check=True, stdout=subprocess.DEVNULL)
    if not (ROOT / "oracle" / "_build" / "liborcpt.so").exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def renderer():
    import rtxamd
    r = rtxamd.Renderer(0)
    yield r
    r.close()
