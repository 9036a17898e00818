"""The reference's KdBVH order (KdBVH.hpp:103-134,192-222) on the CPU: the
library's host build (mclg_kdbvh_order, used for MCLG_ORIENT_KDBVH and
threaded = false) against the oracle's, and the oracle's own build pinned by
what the median split fixes independently of any implementation:

  - every node's left part holds the floor(m/2) smallest centres (2 of 3 for a
    node of three) along axis depth mod dim, for distinct centres;
  - the node numbering (the order the reference pushes volumes, which is its
    frontlist order, BVHTree.cpp:168-185) depends on the leaf count only;
  - the serial traversal (threaded = false, BVHTree.cpp:152-157) reaches leaf
    pairs in increasing (frontlist index of the node separating the two leaves,
    position of the first, position of the second) — the order the library's
    callback replay uses."""
import numpy as np
import pytest

from mclccd_amd import scenes
from mclccd_amd.tree import kdbvh_order
from oracle.oracle import run_scene


def _centres(o, n, dim):
    b, _ = o.boxes()
    b = b.reshape(n, 6, dim)
    return (b[:, 0] + b[:, 1]) / 2.0


def _split(frm, to):
    return frm + 2 if to - frm == 3 else frm + (to - frm) // 2


def _gapnode_from_count(n):
    """Node index of each gap, numbered in push order (children first)."""
    gap = np.full(max(n - 1, 0), -1, np.int64)
    nxt = 0
    # iterative post-order over [from, to) ranges
    st = [(0, n, False)]
    while st:
        frm, to, done = st.pop()
        if to - frm < 2:
            continue
        mid = frm + 1 if to - frm == 2 else _split(frm, to)
        if done:
            gap[mid - 1] = nxt
            nxt += 1
            continue
        st.append((frm, to, True))
        if to - frm > 2:
            st.append((mid, to, False))
            st.append((frm, mid, False))
    return gap


SCENES = {
    "cloth": lambda: scenes.cloth_fold(n=24),
    "layers": lambda: scenes.stacked_layers(n=12),
    "disk": lambda: scenes.disk2d(npts=2000),
    "soup": lambda: scenes.random_soup(400, seed=3, share=True, motion=0.2),
    # a static grid: centres on a lattice, with many equal coordinates (ties)
    "grid": lambda: (lambda V, F: (V, V.copy(), F, 1e-3))(*scenes.grid_sheet(20)),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_library_order_equals_oracle(name):
    V0, V1, F, th = SCENES[name]()
    o = run_scene(V0, V1, F, th, threaded=False, candidates=False)
    n, dim = len(F), V0.shape[1]
    rank, gap = kdbvh_order(_centres(o, n, dim))
    assert np.array_equal(rank, o.rank())
    assert np.array_equal(gap, _gapnode_from_count(n))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_partition_is_the_median_split(seed):
    # distinct centres: the median split fixes every node's leaf set
    V0, V1, F, th = scenes.random_soup(700, seed=seed, share=False, motion=0.1)
    o = run_scene(V0, V1, F, th, threaded=False, candidates=False)
    n, dim = len(F), V0.shape[1]
    cen = _centres(o, n, dim)
    assert all(len(np.unique(cen[:, k])) == n for k in range(dim))
    by_pos = np.empty((n, dim))
    by_pos[o.rank()] = cen
    checked = 0
    st = [(0, n, 0)]
    while st:
        frm, to, depth = st.pop()
        if to - frm < 3:
            continue
        mid, ax = _split(frm, to), depth % dim
        assert by_pos[frm:mid, ax].max() < by_pos[mid:to, ax].min(), (frm, to, depth)
        checked += 1
        st.append((frm, mid, depth + 1))
        st.append((mid, to, depth + 1))
    assert checked > n // 4
    # the library's build makes the same choices on the same centres
    rank, _ = kdbvh_order(cen)
    assert np.array_equal(rank, o.rank())


def _range_max(a, lo, hi):
    return int(a[lo:hi + 1].max())


@pytest.mark.parametrize("name", ["cloth", "layers", "soup", "disk"])
def test_serial_traversal_order(name):
    V0, V1, F, th = SCENES[name]()
    o = run_scene(V0, V1, F, th, threaded=False, candidates=False)
    n = len(F)
    gap = _gapnode_from_count(n)
    rank = o.rank()
    lp = o.leaf_pairs()  # in the serial visiting order, left-subtree leaf first
    assert len(lp) > 100
    ra, rb = rank[lp[:, 0]], rank[lp[:, 1]]
    assert np.all(ra < rb)  # collide orients the left-subtree leaf first
    node = np.array([_range_max(gap, a, b - 1) for a, b in zip(ra, rb)])
    key = np.stack([node, ra, rb], axis=1)
    # strictly increasing keys, i.e. the serial order is exactly this sort
    d = np.diff(key, axis=0)
    first = np.where(d[:, 0] != 0, d[:, 0], np.where(d[:, 1] != 0, d[:, 1], d[:, 2]))
    assert np.all(first > 0)
    # events come pair by pair, hits before the pair's discrete hit
    pair, kind, _ = o.events()
    assert np.all(np.diff(pair) >= 0)
    same = pair[1:] == pair[:-1]
    assert not np.any(same & (kind[:-1] == 2))
