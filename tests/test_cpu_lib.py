"""CPU-side checks of the native library: it loads, exports every entry point
that include/mclg.h declares, and fails cleanly (no crash) without a GPU."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    out = set()
    for name in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not name.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", name)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        out |= set(re.findall(r"\b(mclg_[a-z0-9_]+)\s*\(", text))
    return out


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("mclg_create", "mclg_update", "mclg_update_device", "mclg_traverse", "mclg_traverse_device",
              "mclg_copy_hits", "mclg_copy_discrete", "mclg_set_rank"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from mclccd_amd import _lib

    L = C.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    from mclccd_amd import _lib

    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_default_options_match_reference():
    # BVHTree.hpp:53-63
    from mclccd_amd import _lib

    o = _lib.Options()
    _lib.lib.mclg_default_options(C.byref(o))
    assert (o.box_eta, o.vf_ccd_eta, o.ee_ccd_eta) == (1e-6, 1e-6, 1e-6)
    assert (o.reptri, o.threaded, o.discrete, o.continuous, o.vf_one_sided, o.ee_robust) == (1, 1, 1, 1, 0, 1)


def test_python_options_defaults():
    from mclccd_amd.tree import Options

    o = Options()
    assert o.box_eta == 1e-6 and o.reptri and o.continuous and o.discrete and not o.vf_one_sided


def test_create_without_gpu_fails_cleanly():
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    from mclccd_amd import _lib

    h = C.c_void_p()
    rc = _lib.lib.mclg_create(3, C.byref(h))
    assert rc != 0 and not h.value
    assert _lib.lib.mclg_create(4, C.byref(h)) == _lib.ERR_ARG


def test_one_hip_runtime_per_process():
    """Loading the library must not bring a second HIP runtime next to torch's
    (the first runtime to initialise would hold the GPU and the other would
    see no device).  Checked in a fresh interpreter that loads the library
    before anything imports torch."""
    import subprocess
    import sys

    code = ("from mclccd_amd import _lib\n"
            "import torch\n"
            "libs = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}\n"
            "print(len(libs))\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().splitlines()[-1] == "1"
