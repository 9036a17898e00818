"""The ACCD kernels run each stencil for at most three iterations in lock-step
and suspend the ones still running, to be resumed later from (t1, t2, t3, g)
(mclccd_amd/csrc/geom.hpp, Accd).  This CPU test compiles that device code for
the host (tests/cpp/accd_resume_check.cpp, hipcc host-only) and checks, on
every candidate stencil of generated scenes, that both the uninterrupted run
and suspend + resume on a fresh state give the oracle's additive_ccd hit and
TOI bit for bit (NarrowPhase.cpp:634-711), with the same number of distance
evaluations, in double and in float (the float tree's ACCD on the
float-rounded coordinates)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mclccd_amd import scenes
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"
SRC = os.path.join(ROOT, "tests", "cpp", "accd_resume_check.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("needs hipcc for the host build of the device code")
    exe = str(tmp_path_factory.mktemp("accd") / "accd_resume_check")
    cmd = [HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
           f"-I{os.path.join(ROOT, 'mclccd_amd', 'csrc')}", f"-I{os.path.join(ROOT, 'include')}", SRC, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _run(exe, path, dim, V0, V1, sten, is_vf, eta):
    with open(path, "wb") as f:
        np.array([dim], np.int32).tofile(f)
        np.array([len(V0)], np.int64).tofile(f)
        np.ascontiguousarray(V0, np.float64).tofile(f)
        np.ascontiguousarray(V1, np.float64).tofile(f)
        np.array([len(sten)], np.int64).tofile(f)
        s4 = np.zeros((len(sten), 4), np.int32)
        s4[:, : sten.shape[1]] = sten
        s4.tofile(f)
        np.ascontiguousarray(is_vf, np.uint8).tofile(f)
        np.array([eta], np.float64).tofile(f)
    r = subprocess.run([exe, path], capture_output=True, text=True, timeout=600)
    line = [l for l in r.stdout.splitlines() if l.startswith("accd_resume_check:")]
    assert r.returncode == 0 and line, r.stdout[-2000:] + r.stderr[-2000:]
    parts = line[-1].split()
    return int(parts[1]), int(parts[3]), int(parts[7])


SCENES = {
    "cloth": lambda: scenes.cloth_fold(n=48),
    "layers": lambda: scenes.stacked_layers(layers=6, n=40),
    "soup": lambda: scenes.random_soup(3000, seed=5),
    "disk": lambda: scenes.disk2d(npts=4000, seed=2),
}


@pytest.mark.parametrize("name", list(SCENES))
def test_suspended_stencils_resume_to_the_oracle_toi(checker, tmp_path, name):
    V0, V1, P, th = SCENES[name]()
    o = orc.run_scene(V0, V1, P, th)
    sten, vf = o.candidates()
    dim = V0.shape[1]
    if dim == 2:
        vf = np.ones(len(sten), np.uint8)
    n, suspended, suspended_f32 = _run(checker, str(tmp_path / "in.bin"), dim, V0, V1, sten, vf, th)
    assert n == len(sten) > 1000
    assert suspended > 0 and suspended_f32 > 0  # the resume path is exercised on every scene, in both precisions


def test_long_iterations_resume(checker, tmp_path):
    # Stencils built to iterate long: slow approach (tangential motion with a
    # small normal component), near-parallel edges, and the 1000-iteration
    # receding case (test_oracle_pins.test_accd_max_iter_returns_true).
    rng = np.random.default_rng(11)
    V0s, V1s, st, vf = [], [], [], []
    base = 0
    for k in range(4000):
        is_vf = k % 2 == 0
        p = rng.uniform(-1, 1, (4, 3)) * np.array([1.0, 1.0, 0.02])
        if not is_vf:
            p[2:] = p[:2] + np.array([0.0, 0.0, 0.01]) + rng.normal(0, 0.02, (2, 3))
        tang = rng.normal(0, 1, 3) * np.array([1.0, 1.0, 0.0])
        disp = np.zeros((4, 3))
        disp[0] = tang * rng.uniform(0.1, 3.0) + np.array([0, 0, -rng.uniform(0.0, 0.05)])
        disp[1:] = rng.normal(0, 0.05, (3, 3))
        V0s.append(p)
        V1s.append(p + disp)
        st.append(base + np.arange(4))
        vf.append(1 if is_vf else 0)
        base += 4
    n, suspended, suspended_f32 = _run(checker, str(tmp_path / "long.bin"), 3, np.concatenate(V0s),
                                       np.concatenate(V1s), np.array(st, np.int32), np.array(vf, np.uint8), 1e-3)
    assert n == 4000 and suspended > 100 and suspended_f32 > 100
    R = 5000.0
    ang = np.array([0, 2 * np.pi / 3, 4 * np.pi / 3])
    face = np.stack([R * np.cos(ang), R * np.sin(ang), np.zeros(3)], 1)
    face -= face.mean(0)
    v0 = np.vstack([[0, 0, 1.0], face])
    v1 = np.vstack([[0, 0, 2.0], -face])
    n, suspended, _ = _run(checker, str(tmp_path / "max.bin"), 3, v0, v1, np.array([[0, 1, 2, 3]], np.int32),
                           np.array([1], np.uint8), 1e-12)
    assert n == 1 and suspended == 1
