"""The ACCD kernels append hits to the hit list themselves (staged per wave,
one counter atomic per flush); when the list overflows its capacity the
traverse falls back to scanning the per-chunk hit counts and compacting the
per-candidate TOIs.  The first traverse of a tree sizes the list for half of
the candidates, so a scene where more than half of the candidates hit takes
the fallback on its first frame and the direct path on the next.  Both must
give the oracle's hits (BVHTree.cpp:428-461), VF hits before EE hits."""
import numpy as np
import pytest

from mclccd_amd import scenes
from oracle.oracle import run_scene
from tests.parity import hit_map
from tests.test_gpu_parity import gpu_tree

pytestmark = pytest.mark.gpu


def _check(t, o, label):
    # the reference's pair orientation (set_rank): every record bit-identical.
    # Failures are reduced to counts: pytest's verbose diff of two maps with
    # hundreds of thousands of entries runs for minutes.
    hs, hv, ht = t.hits()
    rs, rv, rt = o.hits()
    assert t.counts()["hits"] == len(hs) == len(rs), label
    gm, rm = hit_map(hs, hv, ht, canonical=False), hit_map(rs, rv, rt, canonical=False)
    if gm != rm:
        differ = sum(1 for k in gm.keys() | rm.keys() if gm.get(k) != rm.get(k))
        pytest.fail(f"{label}: {differ} of {len(rm)} hit keys differ from the oracle")
    nvf = int(np.count_nonzero(hv))
    assert bool((hv[:nvf] == 1).all() and (hv[nvf:] == 0).all()), f"{label}: VF hits do not precede EE hits"
    _, _, gt = t.candidates()
    assert int((gt >= 0).sum()) == len(hs), f"{label}: hit count differs from the candidates with toi >= 0"


@pytest.mark.parametrize("scene", ["layers", "soup"])
def test_fallback_then_direct(gpu, scene):
    if scene == "layers":  # most candidates hit: the first frame overflows
        V0, V1, F, th = scenes.stacked_layers(layers=4, n=40)
    else:  # few candidates hit: every frame direct
        V0, V1, F, _ = scenes.random_soup(400, seed=3, share=True, motion=0.05)
        th = 1e-2
    o = run_scene(V0, V1, F, th)
    t = gpu_tree(3)
    t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = th
    t.update(V0, V1, F)
    t.set_rank(o.rank())
    for frame in range(3):
        t.run(V0, V1, F, keep_candidates=True)
        c = t.counts()
        if frame == 0:
            overflow = c["hits"] > max(1024, c["candidates"] // 2)
            assert overflow == (scene == "layers"), (c["hits"], c["candidates"])
        # after an overflow the list holds every candidate: long-running stencils are suspended
        assert (c["suspended"] > 0) == (overflow and frame > 0), (frame, c)
        _check(t, o, f"{scene} frame {frame}")
    print(f"[hits] {scene}: {c['candidates']} candidates, {c['hits']} hits")


def test_direct_hits_2d(gpu):
    V0, V1, E, th = scenes.disk2d(npts=3000, edges="all")
    o = run_scene(V0, V1, E, th, dim=2)
    t = gpu_tree(2)
    t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = th
    t.update(V0, V1, E)
    t.set_rank(o.rank())
    for frame in range(2):
        t.run(V0, V1, E, keep_candidates=True)
        _check(t, o, f"disk2d frame {frame}")
