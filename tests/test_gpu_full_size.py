"""The plain path at full size (16 stacked 180x180 cloth layers, 1,036,800
triangles; every edge of the 2D disk of 300k triangles), checked through properties that
do not need a full CPU run:
  - sampled leaf pairs satisfy the reference's leaf test (BVHTree.cpp:400-426)
    on the leaf boxes, and for sampled leaves the pairs found are exactly the
    brute-force set over all leaves (nothing missed, nothing extra, no duplicate);
  - sampled candidate stencils get the oracle's ACCD TOI bit for bit (the same
    stencil in the same order gives the same arithmetic);
  - the hit list holds exactly the candidates with toi >= 0 (as multisets of
    (stencil, toi) records, through a 64-bit hash sum per kind), VF hits first,
    in a frame whose hit list outgrew its capacity, in the next (appended
    directly) and in one that keeps no per-candidate TOIs;
  - sampled discrete pairs intersect at V1 and sampled non-discrete leaf pairs
    do not (the oracle's tri-tri / segment test).
The leaf boxes themselves are pinned bit for bit against the oracle at smaller
sizes (tests/test_gpu_parity.py)."""
import numpy as np
import pytest

from mclccd_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _leaf_test(B, A, i, j, d):
    """BVHTree.cpp:400-426 vectorised over index arrays i, j (boxes as the
    library returns them: swept, t0, t1, each lo then hi)."""
    lo = lambda k, b: B[:, k * 2 * d:(k * 2 + 1) * d][b]
    hi = lambda k, b: B[:, (k * 2 + 1) * d:(k * 2 + 2) * d][b]
    ok = A[i].astype(bool) | A[j].astype(bool)
    ok &= np.all(lo(0, i) <= hi(0, j), axis=1) & np.all(lo(0, j) <= hi(0, i), axis=1)
    empty = lambda k, b: np.any(lo(k, b) > hi(k, b), axis=1)
    any_empty = empty(1, i) | empty(1, j) | empty(2, i) | empty(2, j)
    sep = np.any(((lo(1, i) > hi(1, j)) & (lo(2, i) > hi(2, j))) | ((hi(1, i) < lo(1, j)) & (hi(2, i) < lo(2, j))),
                 axis=1)
    return ok & (any_empty | ~sep)


def _run(V0, V1, P, th):
    from mclccd_amd.tree import BVHTree

    t = BVHTree(V0.shape[1])
    t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = th
    t.update(V0, V1, P)
    t.run(V0, V1, P, keep_candidates=True, keep_leaf_pairs=True)
    return t


def _check_pairs(t, n, d, rng):
    B, A, _ = t.leaves()
    pairs = t.leaf_pairs()
    assert len(pairs) == t.counts()["leaf_pairs"] > 0
    s = pairs[rng.choice(len(pairs), min(len(pairs), 200_000), replace=False)]
    assert np.all(s[:, 0] != s[:, 1])
    assert np.all(_leaf_test(B, A, s[:, 0], s[:, 1], d)), "a reported leaf pair fails the leaf test"
    # completeness for sampled leaves: brute force against every leaf
    q = rng.choice(n, 48, replace=False)
    mine = {int(x): set() for x in q}
    for col, other in ((0, 1), (1, 0)):
        m = np.isin(pairs[:, col], q)
        for a, b in pairs[m][:, [col, other]].tolist():
            assert b not in mine[a], "duplicate leaf pair"
            mine[a].add(b)
    allj = np.arange(n)
    for a in q.tolist():
        hit = _leaf_test(B, A, np.full(n, a), allj, d)
        hit[a] = False
        assert set(np.nonzero(hit)[0].tolist()) == mine[a], f"leaf {a}: pair set differs from brute force"
    return pairs


def _check_candidates_and_hits(t, V0, V1, th, rng):
    import torch

    vs, vt, es, et = t.device_candidates()
    for sten, toi, vf in ((vs, vt, True), (es, et, False)):
        if len(sten) == 0:
            continue
        idx = torch.as_tensor(rng.choice(len(sten), min(len(sten), 100_000), replace=False), device=sten.device)
        S = sten[idx].cpu().numpy()
        T = toi[idx].cpu().numpy()
        want = O.toi_list(V0, V1, S, np.full(len(S), vf, np.uint8), th, th)
        assert np.array_equal(T, want), f"{'VF' if vf else 'EE'} TOIs differ from the oracle on the same stencils"
    hs, hv, ht = t.device_hits()
    # The ACCD kernels append hits in the order their waves finish, so the
    # hit list is compared with the candidates as a multiset: per kind, the
    # record count and the wrapping int64 sum of a 64-bit mix of each record
    # (stencil, toi bits).
    want = _candidate_hashes(vs, vt, es, et)
    assert _hit_hashes(hs, hv, ht) == want, "hit records differ from the candidates with toi >= 0"
    nvf = want[1][0]
    assert len(hs) == t.counts()["hits"]
    assert bool((hv[:nvf] == 1).all()) and bool((hv[nvf:] == 0).all()), "VF hits do not all precede EE hits"
    return want


# Worked in slices of 8M rows: torch's boolean-mask indexing of a whole
# 297M-row queue returned zero rows past ~67M input rows on this software stack.
_STEP = 1 << 23


def _candidate_hashes(vs, vt, es, et):
    """{1: (count, hash) of VF candidates with toi >= 0, 0: the same for EE}."""
    import torch

    out = {}
    for sten, toi, flag in ((vs, vt, 1), (es, et, 0)):
        n, h = 0, torch.zeros((), dtype=torch.int64, device=sten.device)
        for a in range(0, len(sten), _STEP):
            m = toi[a:a + _STEP] >= 0
            n += int(m.sum())
            h += torch.where(m, _record_hash(sten[a:a + _STEP], toi[a:a + _STEP]), 0).sum()
        out[flag] = (n, int(h))
    return out


def _hit_hashes(hs, hv, ht):
    import torch

    out = {}
    for flag in (1, 0):
        n, h = 0, torch.zeros((), dtype=torch.int64, device=hs.device)
        for a in range(0, len(hs), _STEP):
            m = hv[a:a + _STEP] == flag
            n += int(m.sum())
            h += torch.where(m, _record_hash(hs[a:a + _STEP], ht[a:a + _STEP]), 0).sum()
        out[flag] = (n, int(h))
    return out


def _record_hash(sten, toi):
    """A 64-bit mix of each (stencil, toi) record, as wrapping int64 arithmetic."""
    import torch

    s = sten.to(torch.int64)
    h = (s[:, 0] * -7046029254386353131 + s[:, 1] * -4658895280553007687 + s[:, 2] * 6620516959819538809 +
         s[:, 3] * -3372029247567499371 + toi.view(torch.int64) * 7809847782465536322)
    h = h ^ (h >> 31)
    h = h * -4265267296055464877
    return h ^ (h >> 29)


def _check_discrete(t, V1, P, pairs, rng):
    d = V1.shape[1]
    disc = t.discrete()
    test = (lambda a, b: O.tri_tri(V1[P[a]], V1[P[b]])) if d == 3 else (lambda a, b: O.edge_edge2(V1[P[a]], V1[P[b]]))
    for a, b in disc[rng.choice(len(disc), min(len(disc), 20_000), replace=False)].tolist():
        assert test(a, b), (a, b)
    dset = set(map(tuple, np.sort(disc, axis=1).tolist()))
    checked = 0
    for a, b in pairs[rng.choice(len(pairs), 40_000, replace=False)].tolist():
        if (min(a, b), max(a, b)) in dset or set(P[a]) & set(P[b]):
            continue
        assert not test(a, b) or not test(b, a), (a, b)  # order-dependent (degenerate) pairs aside
        checked += 1
    assert checked > 1000


def test_full_size_3d(gpu):
    V0, V1, F, th = scenes.stacked_layers()
    # frame 1 outgrows the first hit-list capacity (the compaction pass); the
    # grown list holds every candidate, so frame 2 appends directly, suspending
    # the ACCD stencils still running after three iterations, and frame 3,
    # without kept candidates, writes no per-candidate TOIs at all
    t = _run(V0, V1, F, th)
    first = t.counts()
    assert first["suspended"] == 0
    first = {k: v for k, v in first.items() if k != "suspended"}
    t.run(V0, V1, F, keep_candidates=True, keep_leaf_pairs=True)
    c2 = t.counts()
    assert c2.pop("suspended") > 0 and c2 == first
    rng = np.random.default_rng(0)
    pairs = _check_pairs(t, len(F), 3, rng)
    want = _check_candidates_and_hits(t, V0, V1, th, rng)
    _check_discrete(t, V1, F, pairs, rng)
    t.run(V0, V1, F)
    c3 = t.counts()
    assert c3.pop("suspended") > 0 and c3 == first
    assert _hit_hashes(*t.device_hits()) == want, "frame 3 hits differ"
    print(f"[full] 3D {len(F)} tris: {first}")


def test_full_size_2d(gpu):
    # every edge of the 300k-triangle disk (the hull alone has under 200 edges)
    V0, V1, E, th = scenes.disk2d(edges="all")
    t = _run(V0, V1, E, th)
    rng = np.random.default_rng(1)
    pairs = _check_pairs(t, len(E), 2, rng)
    _check_candidates_and_hits(t, V0, V1, th, rng)
    _check_discrete(t, V1, E, pairs, rng)
    print(f"[full] 2D {len(E)} edges: {t.counts()}")
