"""The 512-scene batched mix at its benchmark configuration (batch_sheets
defaults: n log-uniform in [16, 180], seed 3, theta 1e-3; about 7.1 M
triangles and 68 M CCD hits) through mclg_batch_run with the autotuned
parameters and graph replay, every scene checked against the oracle run on
that scene alone (the same per-scene problem as the reference's
BVHTree::update + traverse, include/MCL/BVHTree.cpp:27-159):

  - per-scene hit and discrete counts equal the oracle's;
  - VF hit records (stencil, toi) are bit-identical as multisets;
  - EE hits: the same stencils (as unordered pairs of edges: the batch orients
    pairs by primitive index, the reference by its KdBVH), TOIs within 1e-9
    relative and never above the oracle's by more than that;
  - discrete pairs equal as unordered pairs.

Records are compared per scene through sorted packed keys (numpy), not
Python containers."""
import numpy as np
import pytest

from mclccd_amd import scenes
from oracle.oracle import run_scene

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def _keys(sten, is_vf):
    """One uint64 per hit: kind bit, then four 15-bit scene-local vertex ids,
    EE stencils as the ordered pair of their two (owner-ordered) edges."""
    s = np.asarray(sten, np.int64).copy()
    ee = np.asarray(is_vf) == 0
    a, b = s[:, 0] * 32768 + s[:, 1], s[:, 2] * 32768 + s[:, 3]
    swap = ee & (b < a)
    s[swap] = s[swap][:, [2, 3, 0, 1]]
    assert s.min() >= 0 and s.max() < 32768
    k = (s[:, 0] << 45) | (s[:, 1] << 30) | (s[:, 2] << 15) | s[:, 3]
    return (k | (np.asarray(is_vf, np.int64) << 60)).astype(np.uint64)


def _canon_disc(p):
    p = np.sort(np.asarray(p, np.int64).reshape(-1, 2), axis=1)
    return np.sort(p[:, 0] * (1 << 20) + p[:, 1])


def _by_scene(sc, n):
    order = np.argsort(sc, kind="stable")
    bounds = np.searchsorted(sc[order], np.arange(n + 1))
    return order, bounds


@pytest.mark.timeout(900)
def test_c4_mix_matches_oracle_per_scene(gpu):
    import torch

    from mclccd_amd.batch import from_scenes

    sc_list = scenes.batch_sheets()
    assert len(sc_list) == 512
    b, V0, V1 = from_scenes(sc_list)
    dev = torch.device("cuda:0")
    dV0, dV1 = torch.from_numpy(V0).to(dev), torch.from_numpy(V1).to(dev)
    torch.cuda.synchronize()
    b.run(dV0, dV1)  # sizes every buffer
    b.autotune(None, reps=1)
    assert b.params()["use_graph"] == 1
    b.run(dV0, dV1)  # captures the graph with the tuned parameters
    first = b.counts()
    first_per = b.scene_counts()
    b.run(dV0, dV1)  # replay
    c = b.counts()
    assert c == first
    per_h, per_d = b.scene_counts()
    assert np.array_equal(per_h, first_per[0]) and np.array_equal(per_d, first_per[1])
    assert c["hits"] > 5e7, c

    sc_h, s_h, v_h, t_h = b.hits()
    assert len(sc_h) == c["hits"]
    order, bounds = _by_scene(sc_h, 512)
    sc_d, p_d = b.discrete()
    dorder, dbounds = _by_scene(sc_d, 512)
    print(f"[c4] {c['hits']} hits, {c['discrete']} discrete, params {b.params()}", flush=True)
    worst = 0.0
    for k, (a0, a1, F, th) in enumerate(sc_list):
        o = run_scene(a0, a1, F, th)
        rs, rv, rt = o.hits()
        assert per_h[k] == len(rs) == bounds[k + 1] - bounds[k], f"scene {k}: hit count"
        assert per_d[k] == o.count(3), f"scene {k}: discrete count"
        idx = order[bounds[k]:bounds[k + 1]]
        gk, rk = _keys(s_h[idx], v_h[idx]), _keys(rs, rv)
        gt, rtt = t_h[idx], np.asarray(rt)
        # sort by (key, toi): equal keys (no reptri duplicates here) pair up in order
        go, ro = np.lexsort((gt, gk)), np.lexsort((rtt, rk))
        gk, gt, rk, rtt = gk[go], gt[go], rk[ro], rtt[ro]
        assert np.array_equal(gk, rk), f"scene {k}: hit stencils differ"
        vf = (gk >> np.uint64(60)) == 1
        assert np.array_equal(gt[vf].view(np.int64), rtt[vf].view(np.int64)), f"scene {k}: VF toi not bit-identical"
        d = np.abs(gt[~vf] - rtt[~vf])
        assert np.all(d <= RTOL * np.abs(rtt[~vf])), f"scene {k}: EE toi beyond 1e-9 relative"
        assert np.all(gt[~vf] <= rtt[~vf] * (1 + RTOL)), f"scene {k}: EE toi above the oracle's"
        if len(d):
            worst = max(worst, float((d / np.maximum(np.abs(rtt[~vf]), 1e-300)).max()))
        di = dorder[dbounds[k]:dbounds[k + 1]]
        assert np.array_equal(_canon_disc(p_d[di]), _canon_disc(o.discrete())), f"scene {k}: discrete pairs"
    print(f"[c4] every scene matches; worst EE toi relative difference {worst:.3g}", flush=True)
