"""GPU vs oracle on the reference's example meshes (tests/golden, MIT):
hand.obj (7,234 vertices, 9,556 triangles) static and deformed, and the
2D UV layout of cathead.obj (397 edges) folded over itself."""
import os

import numpy as np
import pytest

from mclccd_amd.meshes import read_obj, unique_edges
from tests.test_gpu_parity import check_parity, run_both

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def hand():
    V, _, F, _ = read_obj(os.path.join(GOLD, "hand.obj.gz"))
    return V, F


@pytest.mark.parametrize("strict", [True, False])
def test_hand_static(gpu, strict):
    # no motion: l_p = 0, so ACCD reports nothing; the discrete hits are the
    # mesh's own self-intersections
    V, F = hand()
    t, o = run_both(V, V.copy(), F, 1e-6, strict)
    assert t.counts()["hits"] == 0 and o.count(2) == 0
    assert t.counts()["discrete"] == o.count(3) > 0
    check_parity(t, o, strict)


@pytest.mark.parametrize("strict", [True, False])
def test_hand_deformed(gpu, strict):
    V, F = hand()
    rng = np.random.default_rng(0)
    scale = (V.max(0) - V.min(0)).max()
    V1 = V + 0.002 * scale * np.sin(6 * V[:, [1, 2, 0]]) + rng.normal(0, 1e-3, V.shape)
    t, o = run_both(V, np.ascontiguousarray(V1), F, 1e-3, strict)
    assert o.count(2) > 1000
    check_parity(t, o, strict)


def test_cathead_uv_fold(gpu):
    _, VT, _, FT = read_obj(os.path.join(GOLD, "cathead.obj.gz"))
    E = unique_edges(FT)
    U0 = np.ascontiguousarray(VT)
    c = U0.mean(0)
    w = U0[:, 0].max() - U0[:, 0].min()
    U1 = U0.copy()
    m = U0[:, 0] > c[0]
    U1[m, 0] -= 0.3 * w * (U0[m, 0] - c[0]) / (U0[:, 0].max() - c[0])
    U1[m, 1] += 0.05 * w
    t, o = run_both(U0, U1, E, 1e-3, True)
    assert o.count(2) > 100
    check_parity(t, o, True)


def test_cli_hand(gpu, tmp_path):
    """The headless example (python -m mclccd_amd.cli) on hand.obj: the
    reference example's report (dillo.cpp:32-80) against the oracle."""
    import io

    from mclccd_amd.cli import run
    from oracle.oracle import run_scene

    buf = io.StringIO()
    r = run(os.path.join(GOLD, "hand.obj.gz"), out=str(tmp_path / "hand"), stream=buf)
    V, F = hand()
    eps = float(np.finfo(np.float32).eps)
    o = run_scene(V, V.copy(), F, eps)
    faces = np.unique(o.discrete().reshape(-1))
    assert r["faces"] == len(F) and r["verts"] == len(V)
    assert r["ccd_hits"] == 0 and r["discrete_pairs"] == o.count(3) > 0
    assert r["discrete_faces"] == len(faces)
    assert np.array_equal(np.loadtxt(tmp_path / "hand_discrete.txt", dtype=np.int64), faces)
    # one box per internal volume below the root, 12 edges each
    assert r["box_edges"] == 12 * (len(F) - 2)
    text = buf.getvalue()
    assert f"Num discrete isect: {len(faces)}" in text and "Num ccd isect: 0" in text
