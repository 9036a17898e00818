"""Callback replay order and the reference's orientation, GPU vs the oracle.

- threaded = False: the Python tree's callbacks arrive in the oracle's
  single-threaded collide order (BVHTree.cpp:152-157,230-302), record for
  record, TOIs bit-identical; with filter_pair and narrow_phase, each leaf pair
  filters all of its candidates before its narrow phase and appends
  (get_candidates, BVHTree.cpp:327-395, then :256-265); append_discrete
  returning True stops everything after it (:236-239).
- The default (threaded) order is deterministic per leaf pair, and a stop
  delivers a prefix of it.
- MCLG_ORIENT_KDBVH without any help from the oracle: oriented leaf pairs,
  stencils, hits (bit-identical) and discrete pairs equal the oracle's, on the
  degenerate inputs whose discrete result depends on the orientation.
- MCLG_CHECK_ORDER lists exactly the pairs whose discrete result changes when
  the two primitives are swapped.
"""
import numpy as np
import pytest

from mclccd_amd import scenes
from oracle import oracle as O
from oracle.oracle import run_scene
from tests.parity import oriented_pairs
from tests.test_gpu_edge_cases import _degenerate_soup, _order_dependent_pairs

pytestmark = pytest.mark.gpu

CASES = {
    "soup": lambda: scenes.random_soup(400, seed=2, share=True, motion=0.2)[:3] + (1e-2,),
    "layers": lambda: scenes.stacked_layers(n=12),
    "degenerate": lambda: _degenerate_soup(21) + (1e-2,),
    "disk": lambda: scenes.disk2d(npts=3000),
}


def _tree(dim, th, threaded=True, orient=None):
    from mclccd_amd.tree import BVHTree

    t = BVHTree(dim)
    t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = th
    t.options.threaded = threaded
    if orient:
        t.set_orientation(orient)
    return t


def _record(t, log, stop_at=None):
    count = [0]

    def ap(s, v, toi):
        log.append(("h",) + tuple(s) + (int(v), float(toi)))

    def ad(p0, p1):
        log.append(("d", p0, p1))
        count[0] += 1
        return stop_at is not None and count[0] == stop_at

    t.append_pair, t.append_discrete = ap, ad


def _oracle_sequence(o):
    ss, sv, st = o.hits()
    sd = o.discrete()
    out = []
    for _, kind, idx in zip(*o.events()):
        if kind == 1:
            out.append(("h",) + tuple(int(x) for x in ss[idx]) + (int(sv[idx]), float(st[idx])))
        elif kind == 2:
            out.append(("d", int(sd[idx, 0]), int(sd[idx, 1])))
    return out


@pytest.mark.parametrize("case", sorted(CASES))
def test_serial_sequence_equals_oracle(gpu, case):
    V0, V1, F, th = CASES[case]()
    o = run_scene(V0, V1, F, th, threaded=False, candidates=False)
    expect = _oracle_sequence(o)
    assert len(expect) > 100
    t = _tree(V0.shape[1], th, threaded=False)
    t.update(V0, V1, F)
    log = []
    _record(t, log)
    t.traverse(V0, V1, F)
    assert log == expect
    nd = sum(1 for e in expect if e[0] == "d")
    if nd >= 2:  # stop at the middle discrete hit: exactly the prefix up to it
        k = (nd + 1) // 2
        cut = [i for i, e in enumerate(expect) if e[0] == "d"][k - 1]
        log2 = []
        _record(t, log2, stop_at=k)
        t.traverse(V0, V1, F)
        assert log2 == expect[:cut + 1]


def test_serial_filter_and_narrow_phase_order(gpu):
    V0, V1, F, th = CASES["soup"]()
    o = run_scene(V0, V1, F, th, threaded=False, candidates=True)
    cs, cv = o.candidates()
    sd = o.discrete()
    pair, kind, idx = o.events()

    def nphase(s, v):  # a deterministic stand-in narrow phase: some misses, some hits
        return ((s[0] * 7 + s[1] * 3 + s[2]) % 11) / 10.0 - 0.3

    def keep(s):
        return (s[0] + s[3]) % 3 != 0

    expect = []
    for p in np.unique(pair):
        m = pair == p
        cands = [tuple(int(x) for x in cs[i]) + (int(cv[i]),) for i in idx[m & (kind == 0)]]
        expect += [("f",) + c for c in cands]  # get_candidates filters all of the pair's stencils first
        for c in cands:
            if keep(c[:4]):
                toi = nphase(c[:4], c[4])
                expect.append(("n",) + c)
                if toi >= 0:
                    expect.append(("h",) + c + (toi,))
        for i in idx[m & (kind == 2)]:
            expect.append(("d", int(sd[i, 0]), int(sd[i, 1])))
    t = _tree(3, th, threaded=False)
    t.update(V0, V1, F)
    log = []
    t.filter_pair = lambda s, v: (log.append(("f",) + tuple(s) + (int(v),)), not keep(s))[1]
    t.narrow_phase = lambda s, v: (log.append(("n",) + tuple(s) + (int(v),)), nphase(s, v))[1]
    t.append_pair = lambda s, v, toi: log.append(("h",) + tuple(s) + (int(v), toi))
    t.append_discrete = lambda p0, p1: log.append(("d", p0, p1)) and False
    t.traverse(V0, V1, F)
    assert log == expect


def test_threaded_order_is_deterministic_and_stops(gpu):
    V0, V1, F, th = CASES["soup"]()
    t = _tree(3, th)
    t.update(V0, V1, F)
    full = []
    _record(t, full)
    t.traverse(V0, V1, F)
    again = []
    _record(t, again)
    t.traverse(V0, V1, F)
    assert again == full
    d_at = [i for i, e in enumerate(full) if e[0] == "d"]
    assert len(d_at) >= 3
    part = []
    _record(t, part, stop_at=3)
    t.traverse(V0, V1, F)
    assert part == full[:d_at[2] + 1]


@pytest.mark.parametrize("case", ["degenerate", "coincident", "disk"])
def test_kdbvh_orientation_without_the_oracle(gpu, case):
    if case == "coincident":
        V0, V1, F, _ = scenes.random_soup(150, seed=5, share=True, motion=0.1)
        n = len(V0)
        V0, V1 = np.concatenate([V0, V0]), np.concatenate([V1, V1])
        F, th = np.concatenate([F, F + n]).astype(np.int32), 1e-2
    else:
        V0, V1, F, th = CASES[case]()
    o = run_scene(V0, V1, F, th)
    t = _tree(V0.shape[1], th, orient="kdbvh")
    t.update(V0, V1, F)
    t.run(V0, V1, F, keep_candidates=True, keep_leaf_pairs=True)
    assert oriented_pairs(t.leaf_pairs()) == oriented_pairs(o.leaf_pairs())
    gs, gv, _ = t.candidates()
    os_, ov = o.candidates()
    key = lambda s, v: sorted(zip(map(tuple, np.asarray(s).tolist()), np.asarray(v).tolist()))
    assert key(gs, gv) == key(os_, ov)
    hs, hv, ht = t.hits()
    rs, rv, rt = o.hits()
    rec = lambda s, v, x: sorted(zip(map(tuple, s.tolist()), v.tolist(), x.tolist()))
    assert rec(hs, hv, ht) == rec(rs, rv, rt)  # EE included: bit-identical TOIs
    gd, od = t.discrete(), o.discrete()
    assert len(gd) == len(od) and oriented_pairs(gd) == oriented_pairs(od)
    # the default traverse (no candidate list kept, hits written directly) gives the same records
    t.run(V0, V1, F)
    hs, hv, ht = t.hits()
    assert rec(hs, hv, ht) == rec(rs, rv, rt)
    assert oriented_pairs(t.discrete()) == oriented_pairs(od)


@pytest.mark.parametrize("case", ["degenerate", "coincident", "disk"])
def test_check_order_lists_the_order_dependent_pairs(gpu, case):
    if case == "coincident":
        V0, V1, F, _ = scenes.random_soup(150, seed=5, share=True, motion=0.1)
        n = len(V0)
        V0, V1 = np.concatenate([V0, V0]), np.concatenate([V1, V1])
        F, th = np.concatenate([F, F + n]).astype(np.int32), 1e-2
    else:
        V0, V1, F, th = CASES[case]()
    o = run_scene(V0, V1, F, th)
    t = _tree(V0.shape[1], th)
    t.update(V0, V1, F)
    t.run(V0, V1, F, check_order=True)
    got = t.order_dependent()
    if V0.shape[1] == 3:
        expect = _order_dependent_pairs(V1, F, o.leaf_pairs())
    else:
        expect = set()
        for a, b in o.leaf_pairs().tolist():
            if set(F[a]) & set(F[b]):
                continue
            if O.edge_edge2(V1[F[a]], V1[F[b]]) != O.edge_edge2(V1[F[b]], V1[F[a]]):
                expect.add((min(a, b), max(a, b)))
    assert {(min(a, b), max(a, b)) for a, b in got.tolist()} == expect
    if case != "disk":
        assert len(expect) > 0
    # a discrete pair outside that list does not depend on the orientation
    disc = {(min(a, b), max(a, b)) for a, b in t.discrete().tolist()}
    ref = {(min(a, b), max(a, b)) for a, b in o.discrete().tolist()}
    assert disc - expect == ref - expect


def test_mismatched_rows_return_silently(gpu):
    # BVHTree.hpp:160,179: V0 and V1 with different row counts -> no update, no traverse
    V0, V1, F, th = CASES["soup"]()
    t = _tree(3, th)
    t.update(V0, V1, F)
    log = []
    _record(t, log)
    t.traverse(V0, V1, F)
    assert len(log) > 0
    t.update(V0, V1[:-1], F)  # returns silently; the last update's tree stays
    again = []
    _record(t, again)
    t.traverse(V0, V1, F)
    assert again == log
    none = []
    _record(t, none)
    t.traverse(V0, V1[:-1], F)
    assert none == []
