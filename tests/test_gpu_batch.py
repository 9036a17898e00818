"""Batched scenes (one pass over many independent scenes) vs the oracle run on
each scene alone, and vs the single-scene GPU path."""
import numpy as np
import pytest

from mclccd_amd import scenes
from oracle.oracle import run_scene
from tests.parity import canon_pairs, hit_map

pytestmark = pytest.mark.gpu


def small_batch(count=24, seed=5):
    out = scenes.batch_sheets(count=count, nmin=4, nmax=24, seed=seed)
    # vary the gap per scene, add a one-triangle scene and a soup
    out = [(V0, V1, F, th * (1 + 0.5 * (i % 3))) for i, (V0, V1, F, th) in enumerate(out)]
    V0, V1, F, _ = scenes.random_soup(60, seed=seed, share=True, motion=0.2)
    out.append((V0, V1, F, 5e-3))
    V0, V1, F, _ = scenes.random_soup(1, seed=seed, share=False)
    out.append((V0, V1, F, 1e-3))
    return out


def check_batch_vs_oracle(b, sc_list, masks=None):
    sc_h, s_h, v_h, t_h = b.hits()
    sc_d, p_d = b.discrete()
    per_h, per_d = b.scene_counts()
    for k, (V0, V1, F, th) in enumerate(sc_list):
        if masks is not None and masks[k] is not None:
            from oracle.oracle import Oracle

            o = Oracle(V0.shape[1])
            o.set_options(box_eta=th, vf_ccd_eta=th, ee_ccd_eta=th)
            o.update(V0, V1, F, active=masks[k])
            o.traverse()
        else:
            o = run_scene(V0, V1, F, th)
        rs, rv, rt = o.hits()
        m = sc_h == k
        gm, rm = hit_map(s_h[m], v_h[m], t_h[m]), hit_map(rs, rv, rt)
        assert gm.keys() == rm.keys(), f"scene {k}"
        for key in gm:  # VF bit-identical; EE within the orientation tolerance
            if key[0] == 1:
                assert gm[key] == rm[key], f"scene {k}"
            else:
                assert np.allclose(gm[key], rm[key], rtol=1e-9, atol=0), f"scene {k}"
        assert per_h[k] == len(rs) and per_d[k] == o.count(3)
        assert canon_pairs(p_d[sc_d == k]) == canon_pairs(o.discrete()), f"scene {k}"


@pytest.mark.parametrize("graph", [1, 0])
def test_batch_matches_oracle_per_scene(gpu, graph):
    from mclccd_amd.batch import from_scenes

    sc_list = small_batch()
    b, V0, V1 = from_scenes(sc_list)
    b.set_params(use_graph=graph)
    b.run(V0, V1)
    check_batch_vs_oracle(b, sc_list)
    # a second frame on the same inputs (graph replay) gives the same results
    c1 = b.counts()
    b.run(V0, V1)
    assert b.counts() == c1
    check_batch_vs_oracle(b, sc_list)


@pytest.mark.parametrize("graph", [1, 0])
def test_batch_active_masks(gpu, graph):
    # BVHTree::update's active mask per scene (BVHTree.cpp:27,88,114-115):
    # every third scene unmasked (all ones), the others with random vertex
    # masks; each scene equals the oracle updated with its own mask.  The mask
    # then changes between graph replays, and is finally removed.
    from mclccd_amd.batch import from_scenes

    sc_list = small_batch(count=12, seed=8)
    b, V0, V1 = from_scenes(sc_list)
    b.set_params(use_graph=graph)
    rng = np.random.default_rng(3)
    for frame in range(2):
        masks = [None if k % 3 == 0 else (rng.uniform(size=len(s[0])) < 0.3).astype(np.int32)
                 for k, s in enumerate(sc_list)]
        full = np.concatenate([np.ones(len(s[0]), np.int32) if m is None else m for m, s in zip(masks, sc_list)])
        b.set_active(full)
        b.run(V0, V1)
        check_batch_vs_oracle(b, sc_list, masks)
    b.set_active(None)
    b.run(V0, V1)
    check_batch_vs_oracle(b, sc_list)


def test_batch_matches_single_scene_path(gpu):
    from mclccd_amd.batch import from_scenes
    from mclccd_amd.tree import BVHTree

    sc_list = scenes.batch_sheets(count=6, nmin=16, nmax=40, seed=9)
    b, V0, V1 = from_scenes(sc_list)
    b.run(V0, V1)
    sc_h, s_h, v_h, t_h = b.hits()
    for k, (a0, a1, F, th) in enumerate(sc_list):
        t = BVHTree(3)
        t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = th
        t.update(a0, a1, F)
        t.run(a0, a1, F)
        gs, gv, gt = t.hits()
        m = sc_h == k
        # both orient pairs by the lower primitive index: bit-identical records
        assert sorted(zip(map(tuple, s_h[m].tolist()), v_h[m].tolist(), t_h[m].tolist())) == \
            sorted(zip(map(tuple, gs.tolist()), gv.tolist(), gt.tolist()))


def test_batch_autotune_and_regrowth(gpu, tmp_path):
    from mclccd_amd.batch import from_scenes

    sc_list = scenes.batch_sheets(count=16, nmin=8, nmax=48, seed=2)
    b, V0, V1 = from_scenes(sc_list)
    b.run(V0, V1)
    ref = b.counts()
    cache = str(tmp_path / "tune.txt")
    b.autotune(cache, reps=1)
    p = b.params()
    b.run(V0, V1)
    assert b.counts() == ref
    # the cache returns the same choice without timing again
    b2, _, _ = from_scenes(sc_list)
    b2.autotune(cache, reps=1)
    assert b2.params() == p
    assert open(cache).read().count("|") == 1
