"""GPU (gfx950, through the C ABI) vs the CPU oracle on the same seeded inputs.

Bar: identical leaf-pair, candidate, hit and discrete sets; TOIs bit-identical
when the GPU orients pairs like the reference's KdBVH (strict mode: the oracle's
rank is handed to the GPU), and within 1e-9 relative for EE stencils in the
default orientation (the EE TOI's rounding depends on which edge comes first).
"""
import numpy as np
import pytest

from mclccd_amd import scenes
from oracle.oracle import run_scene
from tests.parity import (boundary_pairs, canon_pairs, canon_stencils, compare_hits, hit_map, oriented_pairs,
                          oriented_stencils)

pytestmark = pytest.mark.gpu

TOI_RTOL = 1e-9  # north-star tolerance for the default (unoriented) comparison


def gpu_tree(dim):
    from mclccd_amd.tree import BVHTree

    return BVHTree(dim)


def run_both(V0, V1, P, theta, strict, active=None, **opts):
    dim = V0.shape[1]
    o = run_scene(V0, V1, P, theta, dim=dim, **opts) if active is None else None
    if active is not None:
        from oracle.oracle import Oracle

        o = Oracle(dim)
        o.set_options(box_eta=theta, vf_ccd_eta=theta, ee_ccd_eta=theta, **opts)
        o.update(V0, V1, P, active)
        o.traverse()
    t = gpu_tree(dim)
    t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = theta
    for k, v in opts.items():
        setattr(t.options, k, v)
    t.update(V0, V1, P, active)
    if strict:
        t.set_rank(o.rank())
    t.run(V0, V1, P, keep_candidates=True, keep_leaf_pairs=True)
    # the same problem without a kept candidate list: no per-candidate TOIs,
    # the ACCD kernels write the hit list directly (the benchmark's path)
    f = gpu_tree(dim)
    f.options.box_eta = f.options.vf_ccd_eta = f.options.ee_ccd_eta = theta
    for k, v in opts.items():
        setattr(f.options, k, v)
    f.update(V0, V1, P, active)
    if strict:
        f.set_rank(o.rank())
    f.run(V0, V1, P)
    t.fused = f
    return t, o


def check_parity(t, o, strict, order_dependent=None):
    """order_dependent: canonical leaf pairs whose discrete result depends on
    the argument order (degenerate triangles); without the reference's
    orientation (strict) they are left out of the discrete comparison."""
    # leaves: FP64 boxes, active flags and representative-triangle flags, bit-identical
    gb, ga, gf = t.leaves()
    ob, oa = o.boxes()
    assert np.array_equal(gb, ob)
    assert np.array_equal(ga, oa)
    assert np.array_equal(gf, o.flags())
    # leaf pairs
    gl, ol = t.leaf_pairs(), o.leaf_pairs()
    assert len(gl) == len(ol)
    if strict:
        assert oriented_pairs(gl) == oriented_pairs(ol)
    else:
        assert canon_pairs(gl) == canon_pairs(ol)
    # candidates
    gs, gv, gt = t.candidates()
    os_, ov = o.candidates()
    assert len(gs) == len(os_)
    c = t.counts()
    assert c["vf_candidates"] == int(np.count_nonzero(ov))
    assert 0 <= c["ee_dist_evals"] <= c["dist_evals"]
    if int(np.count_nonzero(ov)) == len(ov):
        assert c["ee_dist_evals"] == 0
    if strict:
        assert oriented_stencils(gs, gv) == oriented_stencils(os_, ov)
    else:
        assert canon_stencils(gs, gv) == canon_stencils(os_, ov)
    # hits
    hs, hv, ht = t.hits()
    rs, rv, rt = o.hits()
    assert c["hits"] == len(rs)
    gm, rm = hit_map(hs, hv, ht, canonical=not strict), hit_map(rs, rv, rt, canonical=not strict)
    if strict:
        assert gm == rm  # bit-identical TOIs
    else:
        # VF TOIs do not depend on orientation: bit-identical
        gvf = {k: v for k, v in gm.items() if k[0] == 1}
        rvf = {k: v for k, v in rm.items() if k[0] == 1}
        assert gvf == rvf
        bad, maxrel, over = compare_hits(gm, rm, canonical=True, rtol=TOI_RTOL)
        assert bad == 0 and over == 0, (bad, maxrel, over)
    # candidate toi agrees with the hit list
    assert int((gt >= 0).sum()) == len(hs)
    # discrete
    gd, od = t.discrete(), o.discrete()
    if strict:
        assert len(gd) == len(od)
        assert oriented_pairs(gd) == oriented_pairs(od)
    else:
        skip = order_dependent or set()
        assert canon_pairs(gd) - skip == canon_pairs(od) - skip
        if not skip:
            assert len(gd) == len(od)
    f = getattr(t, "fused", None)
    if f is not None:  # direct hit list: the records of the run above, bit for bit (same orientation)
        fc = f.counts()
        for k in ("leaf_pairs", "candidates", "vf_candidates", "hits", "discrete", "dist_evals", "ee_dist_evals"):
            assert fc[k] == c[k], (k, fc[k], c[k])
        fs, fv, ft = f.hits()
        rec = lambda s, v, x: sorted(zip(map(tuple, s.tolist()), v.tolist(), x.tolist()))
        assert rec(fs, fv, ft) == rec(hs, hv, ht)
        assert oriented_pairs(f.discrete()) == oriented_pairs(gd)
    return c


@pytest.mark.parametrize("strict", [True, False])
def test_cloth_fold(gpu, strict):
    V0, V1, F, th = scenes.cloth_fold()
    t, o = run_both(V0, V1, F, th, strict)
    c = check_parity(t, o, strict)
    assert c["leaf_pairs"] == 68104
    assert c["hits"] > 1000


@pytest.mark.parametrize("strict", [True, False])
def test_stacked_layers_small(gpu, strict):
    # the 16-layer workload at 24x24 quads per layer (18,432 tris)
    V0, V1, F, th = scenes.stacked_layers(n=24)
    t, o = run_both(V0, V1, F, th, strict)
    c = check_parity(t, o, strict)
    assert c["leaf_pairs"] > 16 * len(F)  # more than the initial queue capacity: exercises regrowth


@pytest.mark.parametrize("edges", ["all", "boundary"])
def test_disk2d_small(gpu, edges):
    V0, V1, E, th = scenes.disk2d(npts=4000, edges=edges)
    t, o = run_both(V0, V1, E, th, strict=True)
    check_parity(t, o, strict=True)
    t, o = run_both(V0, V1, E, th, strict=False)
    check_parity(t, o, strict=False)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_random_soup(gpu, seed):
    V0, V1, F, th = scenes.random_soup(400, seed=seed, share=seed % 2 == 0, motion=0.2)
    for strict in (True, False):
        t, o = run_both(V0, V1, F, 1e-2, strict)
        c = check_parity(t, o, strict)
    assert c["discrete"] > 0


def test_static_no_ccd_hits(gpu):
    # V0 == V1: l_p == 0 for every stencil, so ACCD never hits (NarrowPhase.cpp:671-673)
    V0, _, F, th = scenes.cloth_fold()
    t, o = run_both(V0, V0.copy(), F, th, strict=False)
    c = check_parity(t, o, strict=False)
    assert c["hits"] == 0 and c["leaf_pairs"] > 0


@pytest.mark.parametrize("opt", [dict(reptri=False), dict(continuous=False), dict(discrete=False)])
def test_options(gpu, opt):
    V0, V1, F, th = scenes.random_soup(300, seed=7, share=True, motion=0.15)
    t, o = run_both(V0, V1, F, 1e-2, strict=True, **opt)
    c = check_parity(t, o, strict=True)
    if opt.get("continuous") is False:
        assert c["hits"] == 0 and c["candidates"] == 0
    if opt.get("discrete") is False:
        assert c["discrete"] == 0


def test_active_mask(gpu):
    V0, V1, F, th = scenes.cloth_fold(n=32)
    rng = np.random.default_rng(5)
    active = (rng.uniform(size=len(V0)) < 0.3).astype(np.int32)
    t, o = run_both(V0, V1, F, th, strict=True, active=active)
    check_parity(t, o, strict=True)


@pytest.mark.parametrize("nprim", [0, 1, 2, 3])
def test_tiny(gpu, nprim):
    V0, V1, F, th = scenes.random_soup(3, seed=11, share=False, spread=0.05, size=0.2, motion=0.1)
    F = F[:nprim]
    t = gpu_tree(3)
    t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = 1e-2
    t.update(V0, V1, F)
    t.run(V0, V1, F, keep_candidates=True, keep_leaf_pairs=True)
    c = t.counts()
    if nprim <= 1:
        assert c == {"leaf_pairs": 0, "candidates": 0, "hits": 0, "discrete": 0, "dist_evals": 0,
                     "vf_candidates": 0, "ee_dist_evals": 0, "order_dependent": 0, "traverse_steps": 0,
                     "suspended": 0}
    else:
        o = run_scene(V0, V1, F, 1e-2)
        check_parity(t, o, strict=False)


def test_eta_clamp(gpu):
    # update() clamps the narrow-phase gaps to box_eta (BVHTree.cpp:39-40)
    V0, V1, F, _ = scenes.cloth_fold(n=8)
    t = gpu_tree(3)
    t.options.box_eta = 1e-4
    t.options.vf_ccd_eta = 1e-2
    t.options.ee_ccd_eta = 1e-5
    t.update(V0, V1, F)
    assert t.options.vf_ccd_eta == 1e-4 and t.options.ee_ccd_eta == 1e-5


def test_bad_indices_rejected(gpu):
    V0, V1, F, th = scenes.cloth_fold(n=4)
    t = gpu_tree(3)
    bad = F.copy()
    bad[3, 1] = len(V0)  # out of range
    from mclccd_amd._lib import MclgError

    with pytest.raises(MclgError):
        t.update(V0, V1, bad)


@pytest.mark.parametrize("eta,exact", [(1e-3, False), (1e-6, False), (0.0, True)])
def test_leaf_test_at_fp32_boundaries(gpu, eta, exact):
    # leaf boxes meeting within a few FP32 / FP64 ulps of the leaf test's
    # threshold: the traversal decides such pairs on FP32 bounds only when the
    # bracket of the FP64 value settles them, otherwise on the FP64 boxes
    V0, V1, F = boundary_pairs(eta, exact)
    t, o = run_both(V0, V1, F, eta, strict=False)
    c = check_parity(t, o, strict=False)
    assert 0 < c["leaf_pairs"] < len(F) // 2
