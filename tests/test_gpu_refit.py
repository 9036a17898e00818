"""Refit-only update (mclg_update_refit; no reference counterpart): after a
full update, later frames keep the hierarchy and recompute the boxes.  The
results must equal the oracle's (a rebuild) on every frame, the kept
hierarchy must be the first frame's, and the boxes the new frame's."""
import numpy as np
import pytest

from mclccd_amd import _lib, scenes
from oracle.oracle import run_scene
from tests.test_gpu_parity import check_parity

pytestmark = pytest.mark.gpu


def _frames(V0, V1, count, seed=0):
    """A sequence of (start, end) positions: the scene's motion, then drifts."""
    rng = np.random.default_rng(seed)
    out = [(V0, V1)]
    a, b = V0, V1
    scale = float((V0.max(0) - V0.min(0)).max())
    for _ in range(count - 1):
        d = 0.02 * scale * np.sin(3 * a[:, [1, 2, 0]] + rng.uniform(0, 6)) + rng.normal(0, 1e-3 * scale, a.shape)
        a, b = b, np.ascontiguousarray(b + d)
        out.append((np.ascontiguousarray(a), b))
    return out


@pytest.mark.parametrize("scene", ["cloth_fold", "soup"])
def test_refit_matches_rebuild(gpu, scene):
    from mclccd_amd.tree import BVHTree

    if scene == "cloth_fold":
        V0, V1, F, th = scenes.cloth_fold(n=32)
    else:
        V0, V1, F, th = scenes.random_soup(300, seed=3, share=True, motion=0.2)
        th = 1e-2
    t = BVHTree(3)
    t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = th
    topo = None
    for k, (A, B) in enumerate(_frames(V0, V1, 4)):
        t.update(A, B, F, refit=k > 0)
        if k == 0:
            topo = t.tree()
        else:  # the first frame's hierarchy and leaf order are kept
            for x, y in zip(t.tree(), topo):
                assert np.array_equal(np.asarray(x), np.asarray(y))
        t.run(A, B, F, keep_candidates=True, keep_leaf_pairs=True)
        o = run_scene(A, B, F, th)
        c = check_parity(t, o, strict=False)
        print(f"[refit] {scene} frame {k}: {c['leaf_pairs']} leaf pairs, {c['hits']} hits")


def test_refit_needs_same_mesh(gpu):
    from mclccd_amd.tree import BVHTree

    V0, V1, F, th = scenes.random_soup(50, seed=1, share=True)
    t = BVHTree(3)
    with pytest.raises(_lib.MclgError) as e:
        t.update(V0, V1, F, refit=True)  # no previous update
    assert e.value.code == _lib.ERR_STATE
    t.update(V0, V1, F)
    with pytest.raises(_lib.MclgError) as e:
        t.update(V0, V1, F[:-1], refit=True)  # another primitive count
    assert e.value.code == _lib.ERR_STATE
    t.update(V0, V1, F, refit=True)
    assert t.np == len(F)
