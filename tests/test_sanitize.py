"""The oracle's C entry points and the library's host-only ordering code
(mclccd_amd/csrc/hostorder.hpp) under AddressSanitizer and
UndefinedBehaviorSanitizer: `make -C oracle sanitize` builds
oracle/sanitize_check.cpp with -fsanitize=address,undefined and runs it over
generated 2D and 3D scenes (empty to three primitives, shared and coincident
vertices, reptri off, CTCD).  Any sanitizer report or failed cross-check is a
non-zero exit."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_oracle_and_host_order_under_sanitizers():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True,
                       timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    last = [l for l in r.stdout.splitlines() if l.startswith("sanitize_check:")]
    assert last and last[-1].endswith(" 0 failed"), out[-4000:]
