"""traverse(BVHTraverse*) (BVHTree.cpp:162-166) over the host copy of the tree:
the reference's point-in-triangle example (examples/traverser.cpp) on the UV
layout of its cathead.obj, against brute force over every triangle; walk rules
(every volume but the root offered, every object once, early stop); merged
volumes contain their children."""
import os

import numpy as np
import pytest

from mclccd_amd.meshes import read_obj

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def barys(p, a, b, c):
    # examples/traverser.cpp:8-26 (point_triangle_barys), in FP64
    v0, v1, v2 = b - a, c - a, p - a
    d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
    d20, d21 = v2 @ v0, v2 @ v1
    inv = 1.0 / (d00 * d11 - d01 * d01)
    r1 = (d11 * d20 - d01 * d21) * inv
    r2 = (d00 * d21 - d01 * d20) * inv
    return np.array([1.0 - r1 - r2, r1, r2])


class PointInTriangle:
    def __init__(self, pt, V, F):
        self.pt, self.V, self.F, self.in_tri = pt, V, F, []

    def intersect_volume(self, box):
        return box.contains(self.pt)

    def intersect_object(self, leaf):
        f = self.F[leaf.idx]
        if barys(self.pt, *self.V[f]).min() >= 0:
            self.in_tri.append(leaf.idx)
        return False


class Collect:
    def __init__(self, stop_after=None):
        self.volumes, self.objects, self.stop_after = [], [], stop_after

    def intersect_volume(self, box):
        self.volumes.append(box)
        return True

    def intersect_object(self, leaf):
        self.objects.append(leaf.idx)
        return self.stop_after is not None and len(self.objects) >= self.stop_after


def test_point_in_triangle_cathead(gpu):
    from mclccd_amd.tree import BVHTree

    _, VT, _, FT = read_obj(os.path.join(GOLDEN, "cathead.obj.gz"))
    t = BVHTree(2)
    t.update(VT, VT, FT)  # as the example: a 2D tree over the UV triangles
    lo, hi = VT.min(0), VT.max(0)
    rng = np.random.default_rng(3)
    pts = lo + rng.uniform(0, 1, (200, 2)) * (hi - lo)
    hits = 0
    for p in pts:
        q = PointInTriangle(p, VT, FT)
        t.traverse(q)
        brute = [i for i in range(len(FT)) if barys(p, *VT[FT[i]]).min() >= 0]
        assert sorted(q.in_tri) == brute
        hits += len(brute)
    assert hits > 50


def test_walk_rules(gpu):
    from mclccd_amd import scenes
    from mclccd_amd.tree import BVHTree

    V0, V1, F, th = scenes.cloth_fold(n=16)
    t = BVHTree(3)
    t.options.box_eta = th
    t.update(V0, V1, F)
    c = Collect()
    t.traverse(c)
    assert sorted(c.objects) == list(range(len(F)))
    assert len(c.volumes) == len(F) - 2  # every internal volume but the root's
    s = Collect(stop_after=1)
    t.traverse(s)
    assert len(s.objects) == 1
    # merged volumes: each contains its children's swept, t0 and t1 boxes
    lc, rc, order, root = t.tree()
    b, a, _ = t.leaves()
    n = len(F)
    m = t._mirror
    for p in range(n - 1):
        for ch in (lc[p], rc[p]):
            assert np.all(m.lo[p] <= m.lo[ch]) and np.all(m.hi[ch] <= m.hi[p])
            assert m.act[p] >= m.act[ch]
    # the root's volume is the union of every leaf's
    lb = b.reshape(n, 6, 3)
    assert np.array_equal(m.lo[root][0], lb[:, 0].min(0)) and np.array_equal(m.hi[root][0], lb[:, 1].max(0))
    assert np.array_equal(np.sort(order), np.arange(n))
