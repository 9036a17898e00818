"""The C++ drop-in facade (include/MCL/BVHTree.hpp) driven by a compiled C++
program (tests/cpp/facade_check.cpp): callback replay, filter_pair,
narrow_phase and the discrete stop, checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from mclccd_amd import scenes
from oracle.oracle import run_scene
from tests.parity import canon_pairs, hit_map

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "facade_check")


def write_mesh(path, V0, V1, F, eta):
    with open(path, "wb") as f:
        np.array([len(V0), len(F)], np.int32).tofile(f)
        np.array([eta], np.float64).tofile(f)
        np.ascontiguousarray(V0, np.float64).tofile(f)
        np.ascontiguousarray(V1, np.float64).tofile(f)
        np.ascontiguousarray(F, np.int32).tofile(f)


def test_facade_callbacks(gpu, tmp_path):
    assert os.path.exists(BIN), "facade_check is built by __graft_entry__.build()"
    V0, V1, F, _ = scenes.random_soup(400, seed=2, share=True, motion=0.2)
    eta = 1e-2
    write_mesh(tmp_path / "in.bin", V0, V1, F, eta)
    r = subprocess.run([BIN, str(tmp_path / "in.bin"), str(tmp_path / "out.txt")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = open(tmp_path / "out.txt").read().splitlines()
    rec = {l.split()[0]: list(map(float, l.split()[1:])) for l in lines
           if not l.startswith(("hit ", "disc ", "seqfull ", "seqstop ", "fseq "))}
    hits = [l.split()[1:] for l in lines if l.startswith("hit ")]
    disc = [tuple(map(int, l.split()[1:])) for l in lines if l.startswith("disc ")]

    o = run_scene(V0, V1, F, eta)
    rs, rv, rt = o.hits()
    # 1. append_pair received exactly the oracle's hits (TOI within the EE orientation tolerance)
    gs = np.array([[int(x) for x in h[:4]] for h in hits], np.int32).reshape(-1, 4)
    gv = np.array([int(h[4]) for h in hits], np.uint8)
    gt = np.array([float(h[5]) for h in hits])
    gm, rm = hit_map(gs, gv, gt), hit_map(rs, rv, rt)
    assert gm.keys() == rm.keys()
    for k in gm:
        assert np.allclose(gm[k], rm[k], rtol=1e-9, atol=0)
    n_pair, n_vf, n_disc, c_hits, c_disc, c_cand = rec["plain"]
    assert n_pair == len(rs) == c_hits and n_vf == int(rv.sum())
    assert canon_pairs(disc) == canon_pairs(o.discrete()) and n_disc == c_disc == o.count(3)
    # 2. filter_pair sees every candidate; EE filtered out -> only VF hits
    n_filter_calls, n_after, n_after_vf = rec["filter"]
    assert n_filter_calls == o.count(1) and n_after == n_after_vf == int(rv.sum())
    # 3. narrow_phase replaces ACCD for every candidate
    n_np, n_np_pairs, bad = rec["narrow"]
    assert n_np == o.count(1) and n_np_pairs == o.count(1) and bad == 0
    # 4. stop after the first discrete hit
    assert rec["stop"][0] == (1 if o.count(3) > 0 else 0)
    # 5. leaves mirror: every vertex owned once
    nleaves, owned = rec["leaves"]
    assert nleaves == len(F) and owned == len(np.unique(F))
    # visitor traversal: every object once, point queries equal brute force over
    # the leaf boxes, a true from intersectObject stops the walk
    nseen, once, bad_idx, nvol = rec["visit_all"]
    assert nseen == once == len(F) and bad_idx == 0 and nvol == len(F) - 2
    npts, mism, found = rec["visit_points"]
    assert npts > 0 and mism == 0 and found >= npts  # each point lies in its own triangles' boxes
    assert rec["visit_stop"][0] == 1
    # the float tree (threaded = false): the oracle's float tree's serial
    # callback sequence, float TOIs bit for bit; leaf boxes equal to float arithmetic
    n_float_hits, box_bad, eta_f = rec["float"]
    assert n_float_hits > 0 and box_bad == 0 and eta_f == np.float32(eta)
    of = run_scene(V0, V1, F, eta, threaded=False, candidates=False, f32=True)
    fs, fv, ft = of.hits()
    fd = of.discrete()
    fexp = []
    for _, kind, idx in zip(*of.events()):
        if kind == 1:
            fexp.append(("h",) + tuple(int(x) for x in fs[idx]) + (int(fv[idx]), float(np.float32(ft[idx]))))
        elif kind == 2:
            fexp.append(("d", int(fd[idx, 0]), int(fd[idx, 1])))
    fgot = []
    for l in lines:
        if l.startswith("fseq h "):
            x = l.split()[2:]
            fgot.append(("h",) + tuple(int(v) for v in x[:5]) + (float(np.float32(float(x[5]))),))
        elif l.startswith("fseq d "):
            fgot.append(("d",) + tuple(int(v) for v in l.split()[2:]))
    assert fgot == fexp and sum(e[0] == "h" for e in fexp) == n_float_hits
    # refit on the same frame: the rebuilt tree's hits
    n_a, n_b = rec["refit"]
    assert n_a == n_b == len(rs)
    # serial traversal (threaded = false): the oracle's single-threaded collide
    # sequence record for record (TOIs bit-identical: the KdBVH orientation), and
    # a stop at the middle discrete hit delivers exactly the prefix up to it
    total_d, stop_at = (int(x) for x in rec["serial"])
    assert total_d == o.count(3) and total_d >= 2
    full = [l.split()[1:] for l in lines if l.startswith("seqfull ")]
    stop = [l.split()[1:] for l in lines if l.startswith("seqstop ")]
    os_ = run_scene(V0, V1, F, eta, threaded=False, candidates=False)
    ss, sv, st = os_.hits()
    sd = os_.discrete()
    expect = []
    for _, kind, idx in zip(*os_.events()):
        if kind == 1:
            expect.append(["h"] + [str(int(x)) for x in ss[idx]] + [str(int(sv[idx])), float(st[idx])])
        elif kind == 2:
            expect.append(["d", str(int(sd[idx, 0])), str(int(sd[idx, 1]))])
    got = [e[:6] + [float(e[6])] if e[0] == "h" else e for e in full]
    assert got == expect
    cut = [i for i, e in enumerate(expect) if e[0] == "d"][stop_at - 1]
    got_stop = [e[:6] + [float(e[6])] if e[0] == "h" else e for e in stop]
    assert got_stop == expect[:cut + 1]
    # matrix overloads: the raw path's hits; mismatched V0 / V1 rows return silently
    n_mat, n_after_bad_update, n_bad_traverse, c_hits = rec["matrix"]
    assert n_mat == c_hits == len(rs) and n_after_bad_update == len(rs) and n_bad_traverse == 0
