"""Degenerate and extreme inputs, GPU vs the oracle (strict and canonical):
zero-area and repeated-vertex triangles, zero-length edges, duplicated
coincident geometry, coordinates far from the origin (coarse FP32 node boxes),
an all-inactive mask, and the 2D path with zero-length edges."""
import numpy as np
import pytest

from mclccd_amd import scenes
from oracle import oracle as O
from tests.test_gpu_parity import check_parity, run_both

pytestmark = pytest.mark.gpu


def _degenerate_soup(seed):
    V0, V1, F, _ = scenes.random_soup(300, seed=seed, share=True, motion=0.2)
    rng = np.random.default_rng(seed)
    F = F.copy()
    V0, V1 = V0.copy(), V1.copy()
    k = len(F)
    # a vertex repeated within a face (BVHTree.cpp:53-79 sees a degenerate edge)
    for i in rng.choice(k, 20, replace=False):
        F[i, 2] = F[i, 0]
    # collinear (zero-area) faces: third vertex moved onto the first edge, at both times
    for i in rng.choice(k, 20, replace=False):
        a, b, c = F[i]
        if len({a, b, c}) == 3:
            V0[c] = 0.5 * (V0[a] + V0[b])
            V1[c] = 0.5 * (V1[a] + V1[b])
    # a vertex that does not move while its neighbours do
    V1[: len(V1) // 10] = V0[: len(V0) // 10]
    return np.ascontiguousarray(V0), np.ascontiguousarray(V1), F


def _order_dependent_pairs(V1, F, pairs):
    """Leaf pairs whose tri-tri result (tt_isect.hpp, as the reference calls it
    with the pair's orientation) changes when the two triangles are swapped.
    This happens in degenerate configurations: a zero-area triangle, or
    coplanar triangles whose orientation tests meet exact zeros (coincident
    copies).  The reference's orientation comes from its KdBVH, so without it
    (the GPU's own leaf order) such a pair may legitimately come out the other
    way; with it (strict mode) every result is identical."""
    out = set()
    for a, b in np.asarray(pairs).tolist():
        if set(F[a]) & set(F[b]):
            continue
        if O.tri_tri(V1[F[a]], V1[F[b]]) != O.tri_tri(V1[F[b]], V1[F[a]]):
            out.add((min(a, b), max(a, b)))
    return out


@pytest.mark.parametrize("seed", [21, 22])
def test_degenerate_triangles(gpu, seed):
    V0, V1, F = _degenerate_soup(seed)
    t, o = run_both(V0, V1, F, 1e-2, strict=True)  # the reference's orientation: everything identical
    check_parity(t, o, strict=True)
    dep = _order_dependent_pairs(V1, F, o.leaf_pairs())
    assert len(dep) > 0  # the scene exercises the order-dependent case
    t, o = run_both(V0, V1, F, 1e-2, strict=False)
    c = check_parity(t, o, strict=False, order_dependent=dep)
    assert c["hits"] > 0


def test_coincident_copies(gpu):
    # every triangle twice, on separate vertices at the same positions: contact
    # at distance 0 (toi 0 for approaching pairs) and coplanar discrete tests
    V0, V1, F, _ = scenes.random_soup(150, seed=5, share=True, motion=0.1)
    n = len(V0)
    W0, W1 = np.concatenate([V0, V0]), np.concatenate([V1, V1])
    G = np.concatenate([F, F + n]).astype(np.int32)
    t, o = run_both(W0, W1, G, 1e-2, strict=True)
    c = check_parity(t, o, strict=True)
    assert c["discrete"] >= len(F)
    dep = _order_dependent_pairs(W1, G, o.leaf_pairs())
    t, o = run_both(W0, W1, G, 1e-2, strict=False)
    check_parity(t, o, strict=False, order_dependent=dep)


@pytest.mark.parametrize("offset", [1e3, 1e6])
def test_far_from_origin(gpu, offset):
    # FP32 node boxes lose precision far from the origin; the traversal's
    # FP64 fallback must still decide every leaf pair exactly
    V0, V1, F, _ = scenes.random_soup(300, seed=8, share=True, motion=0.2)
    shift = np.array([offset, -0.7 * offset, 0.3 * offset])
    t, o = run_both(np.ascontiguousarray(V0 + shift), np.ascontiguousarray(V1 + shift), F, 1e-2, strict=True)
    c = check_parity(t, o, strict=True)
    assert c["leaf_pairs"] > 0


def test_all_inactive(gpu):
    V0, V1, F, th = scenes.cloth_fold(n=16)
    t, o = run_both(V0, V1, F, th, strict=True, active=np.zeros(len(V0), np.int32))
    c = check_parity(t, o, strict=True)
    assert c["leaf_pairs"] == 0 and c["hits"] == 0


def test_2d_zero_length_edges(gpu):
    V0, V1, E, th = scenes.disk2d(npts=2000, edges="all")
    E = E.copy()
    rng = np.random.default_rng(3)
    idx = rng.choice(len(E), 30, replace=False)
    E[idx, 1] = E[idx, 0]  # zero-length edges
    for strict in (True, False):
        t, o = run_both(V0, V1, E, th, strict)
        check_parity(t, o, strict)
