"""Pins of the CPU oracle (oracle/mclccd_oracle.hpp) by independent means.

The reference ships no tests or golden vectors, so each oracle function is
checked against something the mathematics fixes: brute force, closed forms,
exact rational predicates, dense sampling.  A plausible slip in the oracle (a
dropped term, a swapped index, a wrong comparison) fails one of these.
"""
import itertools
from fractions import Fraction

import numpy as np
import pytest

from mclccd_amd import scenes
from oracle import oracle as orc
from tests.parity import quirk_region, quirk_stencils

# ---------------------------------------------------------------------------
# leaf-pair set: tree traversal == O(n^2) brute force (the set is tree-independent)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("case", ["cloth", "soup0", "soup1", "layers", "disk"])
def test_leaf_pairs_equal_brute_force(case):
    if case == "cloth":
        V0, V1, P, th = scenes.cloth_fold()
    elif case == "layers":
        V0, V1, P, th = scenes.stacked_layers(layers=4, n=16)
    elif case == "disk":
        V0, V1, P, th = scenes.disk2d(npts=2000)
    else:
        V0, V1, P, th = scenes.random_soup(300, seed=int(case[-1]), share=case == "soup0")
    o = orc.run_scene(V0, V1, P, th)
    lp = o.leaf_pairs()
    bf = o.brute_force_pairs()
    assert len(lp) == len(bf)
    assert set(map(tuple, np.sort(lp, axis=1).tolist())) == set(map(tuple, bf.tolist()))
    if case == "cloth":
        assert len(lp) == 68104


def test_leaf_boxes_closed_form():
    V0, V1, F, th = scenes.random_soup(200, seed=3, share=True)
    o = orc.run_scene(V0, V1, F, 0.01)
    boxes, act = o.boxes()
    X0, X1 = V0[F], V1[F]  # (np, 3, 3)
    t0lo, t0hi = X0.min(1) - 0.01, X0.max(1) + 0.01
    t1lo, t1hi = X1.min(1) - 0.01, X1.max(1) + 0.01
    swlo = np.minimum(X0.min(1), X1.min(1)) - 0.01
    swhi = np.maximum(X0.max(1), X1.max(1)) + 0.01
    exp = np.concatenate([swlo, swhi, t0lo, t0hi, t1lo, t1hi], axis=1)
    assert np.array_equal(boxes, exp)
    assert act.all()


def test_active_flag():
    V0, V1, F, th = scenes.random_soup(100, seed=4, share=True)
    active = np.zeros(len(V0), np.int32)
    active[F[:10, 0]] = 1
    o = orc.Oracle(3)
    o.set_options(box_eta=th, vf_ccd_eta=th, ee_ccd_eta=th)
    o.update(V0, V1, F, active)
    _, act = o.boxes()
    exp = active[F].any(axis=1)
    assert np.array_equal(act.astype(bool), exp)
    # inactive-inactive pairs never appear
    o.traverse()
    lp = o.leaf_pairs()
    assert (exp[lp[:, 0]] | exp[lp[:, 1]]).all()


def _reptri_expected(F):
    """Owner of a vertex / edge = its first (face, slot) occurrence, computed with
    numpy's unique(return_index) over the slots in face order."""
    npr, pdim = F.shape
    flat = F.reshape(-1)
    _, first_v = np.unique(flat, return_index=True)
    vflag = np.zeros(npr * pdim, bool)
    vflag[first_v] = True
    out = vflag.reshape(npr, pdim).astype(np.uint8)
    if pdim == 3:
        a = F.reshape(-1)
        b = F[:, [1, 2, 0]].reshape(-1)
        key = np.minimum(a, b).astype(np.int64) * (1 << 32) + np.maximum(a, b)
        _, first_e = np.unique(key, return_index=True)
        eflag = np.zeros(npr * 3, bool)
        eflag[first_e] = True
        e = eflag.reshape(npr, 3).astype(np.uint8)
    else:
        e = np.zeros((npr, 3), np.uint8)
    bits = np.zeros(npr, np.uint8)
    for j in range(pdim):
        bits |= out[:, j] << j
    for j in range(3):
        bits |= e[:, j] << (3 + j)
    return bits


@pytest.mark.parametrize("case", ["cloth", "soup", "disk"])
def test_reptri_closed_form(case):
    if case == "cloth":
        V0, V1, P, th = scenes.cloth_fold(n=16)
    elif case == "soup":
        V0, V1, P, th = scenes.random_soup(300, seed=5, share=True)
    else:
        V0, V1, P, th = scenes.disk2d(npts=1000)
    o = orc.run_scene(V0, V1, P, th)
    assert np.array_equal(o.flags(), _reptri_expected(P))


def test_candidates_invariants():
    V0, V1, F, th = scenes.cloth_fold(n=24)
    o = orc.run_scene(V0, V1, F, th)
    s, v = o.candidates()
    flags = o.flags()
    keys = set()
    for st, vf in zip(s.tolist(), v.tolist()):
        if vf:
            assert st[0] not in st[1:]  # no shared vertex
            k = (1, st[0], tuple(sorted(st[1:])))
        else:
            assert len(set(st)) == 4
            e0, e1 = tuple(sorted(st[:2])), tuple(sorted(st[2:]))
            k = (0,) + tuple(sorted([e0, e1]))
        assert k not in keys  # reptri: every stencil once
        keys.add(k)
    # every VF vertex is owned by some face, every EE edge too
    nv = len(V0)
    owner_faces = {}
    for f in range(len(F)):
        for j in range(3):
            if flags[f] >> j & 1:
                owner_faces[F[f, j]] = f
    assert len(owner_faces) == nv
    # expected counts from the leaf pairs, computed with numpy
    lp = o.leaf_pairs()
    f0, f1 = F[lp[:, 0]], F[lp[:, 1]]
    fa, fb = flags[lp[:, 0]], flags[lp[:, 1]]
    nvf = 0
    for x in range(3):
        sh0 = (f0[:, [x]] == f1).any(1)
        sh1 = (f1[:, [x]] == f0).any(1)
        nvf += int(((fa >> x & 1).astype(bool) & ~sh0).sum() + ((fb >> x & 1).astype(bool) & ~sh1).sum())
    nee = 0
    for x in range(3):
        for y in range(3):
            a0, a1 = f0[:, x], f0[:, (x + 1) % 3]
            b0, b1 = f1[:, y], f1[:, (y + 1) % 3]
            ok = (fa >> (3 + x) & 1).astype(bool) & (fb >> (3 + y) & 1).astype(bool)
            ok &= (a0 != b0) & (a0 != b1) & (a1 != b0) & (a1 != b1)
            nee += int(ok.sum())
    assert int(v.sum()) == nvf and int((v == 0).sum()) == nee


# ---------------------------------------------------------------------------
# ACCD closed forms (NarrowPhase.cpp:634-711 with s = 1)
# ---------------------------------------------------------------------------
def tau1(d0, xi, lp):
    """ACCD's first step, 0.9 (d0^2 - xi^2) / ((d0 + xi) l_p): for an approaching
    pair whose distance falls linearly the algorithm stops after it (fact of
    the s = 1 variant), so it is the returned toi."""
    return 0.9 * (d0 * d0 - xi * xi) / ((d0 + xi) * lp)


@pytest.mark.parametrize("h,xi", [(0.5, 1e-3), (0.1, 1e-4), (1e-2, 1e-3), (1.0, 0.0), (0.25, 1e-6)])
def test_accd_vertex_falls_on_face(h, xi):
    # vertex above the face interior falls straight down with speed 1; face static
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    v0 = np.vstack([[0.25, 0.25, h], tri])
    v1 = np.vstack([[0.25, 0.25, h - 1.0], tri])
    # centred displacements: vertex (0,0,-3/4), face (0,0,1/4): l_p = 1 exactly; d0 = h
    assert orc.pair_distance(3, v0, True) == h
    hit, toi, ev = orc.accd(3, v0, v1, xi, True)
    assert hit and toi == tau1(h, xi, 1.0)
    assert ev == 4  # d0, the t_l = 0 re-evaluation, d1, d2
    # conservative: the gap at toi is still above xi
    assert h - toi > xi


def test_accd_edges_cross():
    # edge along x at height h falls onto an edge along y; l_p = 1/2 + 1/2
    h, xi = 0.3, 1e-3
    v0 = np.array([[-1, 0, h], [1, 0, h], [0, -1, 0], [0, 1, 0]], float)
    v1 = v0.copy()
    v1[:2, 2] -= 1.0
    assert orc.pair_distance(3, v0, False) == h
    hit, toi, _ = orc.accd(3, v0, v1, xi, False)
    assert hit and toi == tau1(h, xi, 1.0)


def test_accd_2d_point_edge():
    h, xi = 0.4, 1e-3
    v0 = np.array([[0.0, h], [-1.0, 0.0], [1.0, 0.0]])
    v1 = v0.copy()
    v1[0, 1] -= 1.5  # centred: point (0,-1), edge ends (0, 0.5): l_p = 1.5
    assert orc.pair_distance(2, v0, True) == h
    hit, toi, _ = orc.accd(2, v0, v1, xi, True)
    assert hit and toi == tau1(h, xi, 1.5)


def test_accd_miss_and_special_cases():
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    # too far to reach within t <= 1: tau1 > 1 -> miss
    v0 = np.vstack([[0.25, 0.25, 2.0], tri])
    v1 = np.vstack([[0.25, 0.25, 1.0], tri])
    assert not orc.accd(3, v0, v1, 1e-3, True)[0]
    # static: l_p = 0 -> miss even when touching (NarrowPhase.cpp:671-673 before :682)
    v0 = np.vstack([[0.25, 0.25, 0.0], tri])
    assert not orc.accd(3, v0, v0, 1e-3, True)[0]
    # already within the gap: hit at toi = 0
    v0 = np.vstack([[0.25, 0.25, 5e-4], tri])
    v1 = np.vstack([[0.25, 0.25, 1.0], tri])
    hit, toi, ev = orc.accd(3, v0, v1, 1e-3, True)
    assert hit and toi == 0.0 and ev == 1
    # receding: the distance only grows -> the steps add up past t = 1 -> miss
    v0 = np.vstack([[0.25, 0.25, 0.1], tri])
    v1 = np.vstack([[0.25, 0.25, 1.1], tri])
    assert not orc.accd(3, v0, v1, 1e-3, True)[0]


def test_accd_max_iter_returns_true():
    # A receding vertex above a face that swings through its centroid: l_p is
    # huge, so the steps are tiny and the receding distance never triggers the
    # break; after 1000 iterations the reference returns true (its assert is
    # compiled out in release builds).
    R = 5000.0
    ang = np.array([0, 2 * np.pi / 3, 4 * np.pi / 3])
    face = np.stack([R * np.cos(ang), R * np.sin(ang), np.zeros(3)], 1)
    face -= face.mean(0)
    v0 = np.vstack([[0, 0, 1.0], face])
    v1 = np.vstack([[0, 0, 2.0], -face])
    hit, toi, ev = orc.accd(3, v0, v1, 1e-12, True)
    assert hit and ev == 1001 and 0 < toi < 1


def test_accd_conservative_on_cloth():
    # invariant of conservative advancement: along the linear path the distance
    # stays above the gap up to the returned toi
    V0, V1, F, th = scenes.cloth_fold(n=16)
    o = orc.run_scene(V0, V1, F, th)
    s, v, t = o.hits()
    rng = np.random.default_rng(0)
    idx = rng.choice(len(s), size=min(300, len(s)), replace=False)
    for i in idx:
        st = s[i][: 4 if v[i] == 0 else 4]
        a, b = V0[st], V1[st]
        for u in np.linspace(0, t[i], 12):
            x = a + u * (b - a)
            assert orc.pair_distance(3, x, bool(v[i])) > th * (1 - 1e-9)


# ---------------------------------------------------------------------------
# projections vs dense sampling
# ---------------------------------------------------------------------------
def test_point_triangle_dense():
    rng = np.random.default_rng(1)
    g = np.linspace(0, 1, 241)
    S, T = np.meshgrid(g, g)
    m = S + T <= 1
    S, T = S[m], T[m]
    n_quirk = 0
    for _ in range(200):
        a, b, c = rng.normal(size=(3, 3))
        q = rng.normal(size=3) * 1.5
        pt = orc.point_on_triangle(q, a, b, c)
        quirk = quirk_region(q, a, b, c)
        if quirk is not None:
            n_quirk += 1
            assert np.allclose(pt, quirk, atol=1e-12)
            continue
        d = np.linalg.norm(q - pt)
        pts = a + S[:, None] * (b - a) + T[:, None] * (c - a)
        dmin = np.linalg.norm(pts - q, axis=1).min()
        assert d <= dmin + 1e-12
        assert d >= dmin - 2e-2 * max(1.0, np.linalg.norm(b - a) + np.linalg.norm(c - a))
    assert n_quirk > 0
    # interior projection equals the plane distance
    tri = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0]], float)
    assert np.array_equal(orc.point_on_triangle([0.5, 0.5, 3.0], *tri), [0.5, 0.5, 0.0])
    # vertex regions return the vertex
    assert np.array_equal(orc.point_on_triangle([-1.0, -1.0, 1.0], *tri), tri[0])
    assert np.array_equal(orc.point_on_triangle([0.0, 5.0, 1.0], *tri), tri[2])


def test_edge_edge_dense():
    rng = np.random.default_rng(2)
    g = np.linspace(0, 1, 401)
    for _ in range(40):
        p0, p1, q0, q1 = rng.normal(size=(4, 3))
        v, bary = orc.edge_to_edge(p0, p1, q0, q1)
        d = np.linalg.norm(v)
        A = p0 + g[:, None] * (p1 - p0)
        B = q0 + g[:, None] * (q1 - q0)
        dmin = np.linalg.norm(A[:, None, :] - B[None, :, :], axis=2).min()
        assert d <= dmin + 1e-12
        assert d >= dmin - 1e-2 * (np.linalg.norm(p1 - p0) + np.linalg.norm(q1 - q0))
        # the returned vector joins the closest points given by bary
        c1 = bary[0] * p0 + bary[1] * p1
        c2 = bary[2] * q0 + bary[3] * q1
        assert np.allclose(c2 - c1, v, atol=1e-12)


def test_point_edge_2d_dense():
    rng = np.random.default_rng(3)
    g = np.linspace(0, 1, 2001)
    for _ in range(50):
        p, e0, e1 = rng.normal(size=(3, 2))
        pt = orc.point_on_edge2(p, e0, e1)
        d = np.linalg.norm(p - pt)
        dmin = np.linalg.norm(e0 + g[:, None] * (e1 - e0) - p, axis=1).min()
        assert dmin - 1e-3 <= d <= dmin + 1e-12
    # degenerate edge returns e0
    assert np.array_equal(orc.point_on_edge2([1.0, 2.0], [0.5, 0.5], [0.5, 0.5]), [0.5, 0.5])


# ---------------------------------------------------------------------------
# discrete predicates vs exact rational geometry
# ---------------------------------------------------------------------------
def _sat_tri_tri(T1, T2):
    """Exact separating-axis test of two closed triangles (rational coordinates):
    they intersect iff no axis among the face normals, the edge-edge cross
    products and the in-plane edge normals separates their projections."""
    T1 = [[Fraction(int(x)) for x in p] for p in T1]
    T2 = [[Fraction(int(x)) for x in p] for p in T2]

    def sub(a, b):
        return [a[i] - b[i] for i in range(3)]

    def cross(a, b):
        return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]

    def dot(a, b):
        return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]

    e1 = [sub(T1[(i + 1) % 3], T1[i]) for i in range(3)]
    e2 = [sub(T2[(i + 1) % 3], T2[i]) for i in range(3)]
    n1, n2 = cross(e1[0], e1[1]), cross(e2[0], e2[1])
    axes = [n1, n2] + [cross(a, b) for a in e1 for b in e2]
    axes += [cross(n1, e) for e in e1] + [cross(n2, e) for e in e2]
    for ax in axes:
        if ax == [0, 0, 0]:
            continue
        p1 = [dot(ax, p) for p in T1]
        p2 = [dot(ax, p) for p in T2]
        if max(p1) < min(p2) or max(p2) < min(p1):
            return False
    return True


def test_tri_tri_exact_integer_grid():
    # small integer coordinates: every product in the predicate is exact in
    # double, so the floating-point predicate must equal the exact answer,
    # including touching, coplanar and degenerate-looking configurations
    rng = np.random.default_rng(7)
    n_hit = n_copl = 0
    for it in range(3000):
        span = 3 if it % 3 == 0 else 6
        T1 = rng.integers(-span, span + 1, (3, 3)).astype(float)
        T2 = rng.integers(-span, span + 1, (3, 3)).astype(float)
        if it % 4 == 0:
            T2[:, 2] = 0
            T1[:, 2] = 0  # coplanar in z = 0
            n_copl += 1
        if np.linalg.norm(np.cross(T1[1] - T1[0], T1[2] - T1[0])) == 0:
            continue
        if np.linalg.norm(np.cross(T2[1] - T2[0], T2[2] - T2[0])) == 0:
            continue
        exp = _sat_tri_tri(T1, T2)
        got = orc.tri_tri(T1, T2)
        assert got == exp, (T1.tolist(), T2.tolist(), got, exp)
        n_hit += exp
    assert n_hit > 300 and n_copl > 300


def test_edge_edge_2d_exact():
    rng = np.random.default_rng(8)
    seen = {"hit": 0, "miss": 0, "parallel": 0}
    for _ in range(4000):
        p0, p1, q0, q1 = rng.integers(-5, 6, (4, 2))
        r, s, n = p1 - p0, q1 - q0, q0 - p0
        rxs = int(r[0] * s[1] - r[1] * s[0])
        got = orc.edge_edge2(np.array([p0, p1], float), np.array([q0, q1], float))
        if rxs == 0:  # parallel (collinear included): reported as no hit
            assert not got
            seen["parallel"] += 1
            continue
        nxr = int(n[0] * r[1] - n[1] * r[0])
        nxs = int(n[0] * s[1] - n[1] * s[0])
        if nxr == 0:  # q0 on the line of p: the reference's sign-change rule
            exp = ((q0[0] - p0[0] < 0) != (q0[0] - p1[0] < 0)) or ((q0[1] - p0[1] < 0) != (q0[1] - p1[1] < 0))
            assert got == exp
            continue
        t, u = Fraction(nxs, rxs), Fraction(nxr, rxs)
        if t in (0, 1) or u in (0, 1):
            continue  # the reciprocal-multiply may round an endpoint either way
        exp = 0 <= t <= 1 and 0 <= u <= 1
        assert got == exp
        seen["hit" if exp else "miss"] += 1
    assert min(seen.values()) > 100


# ---------------------------------------------------------------------------
# reference TOI gradient (dual numbers) vs central differences, on the CPU
# ---------------------------------------------------------------------------
def _general_position_cloth(n=16, seed=5):
    # the rest pose of cloth_fold lies exactly on a grid (parallel edges, tied
    # displacement norms), where the TOI has kinks; jitter both poses
    V0, V1, F, th = scenes.cloth_fold(n=n)
    rg = np.random.default_rng(seed)
    return V0 + rg.normal(0, 1e-4, V0.shape), V1 + rg.normal(0, 1e-3, V1.shape), F, th


def test_oracle_toi_gradient_vs_finite_differences():
    V0, V1, F, th = _general_position_cloth()
    o = orc.run_scene(V0, V1, F, th)
    s, v, toi = o.hits()
    assert v.sum() > 50 and (1 - v).sum() > 50
    # the oracle appends hits concurrently (as the reference's callbacks run):
    # sample them in a canonical order so every run checks the same ones
    order = np.lexsort((s[:, 3], s[:, 2], s[:, 1], s[:, 0], v))
    s, v, toi = s[order], v[order], toi[order]
    rng = np.random.default_rng(0)
    h = 1e-7
    sten1 = np.arange(4, dtype=np.int32)[None, :]
    checked = 0
    for i in rng.choice(len(s), 200, replace=False):
        st = s[i][:4]
        a, b = V0[st], V1[st]
        vf = v[i:i + 1]
        hit, t0, ga, gb = orc.accd_grad(3, a, b, th, bool(v[i]))
        assert hit and t0 == toi[i]
        _, s0 = orc.toi_list(a, b, sten1, vf, th, th, signature=True)
        for arr, g in ((0, ga), (1, gb)):
            j, k = int(rng.integers(0, 4)), int(rng.integers(0, 3))
            ap, bp = a.copy(), b.copy()
            (ap if arr == 0 else bp)[j, k] += h
            am, bm = a.copy(), b.copy()
            (am if arr == 0 else bm)[j, k] -= h
            tp, sp = orc.toi_list(ap, bp, sten1, vf, th, th, signature=True)
            tm, sm = orc.toi_list(am, bm, sten1, vf, th, th, signature=True)
            # a step that changes the forward pass's branches crosses a kink
            # of the TOI, where central differences do not apply
            if tp[0] < 0 or tm[0] < 0 or not (sp[0] == sm[0] == s0[0]):
                continue
            assert abs((tp[0] - tm[0]) / (2 * h) - g[j, k]) <= 1e-6 * max(1.0, abs(g[j, k]))
            checked += 1
    assert checked > 350


def test_oracle_toi_gradient_at_a_kink_is_one_sided():
    # two parallel edges at t0 (distance 0.0625): moving q0 away keeps the
    # distance (attained at q1), moving it closer shrinks it at unit rate, so
    # the TOI has a kink; the reference gradient follows the branch the
    # forward pass takes, i.e. equals one of the one-sided derivatives
    a = np.array([[0.125, 0.375, 0.0], [0.125, 0.4375, 0.0], [0.1875, 0.375, 0.0], [0.1875, 0.4375, 0.0]])
    b = np.array([[0.12736789, 0.37431296, 0.19628921], [0.12380183, 0.43635634, 0.10455922],
                  [0.18655589, 0.37648112, 0.20963031], [0.18549369, 0.43685593, 0.11128993]])
    h = 1e-8
    hit, t0, ga, gb = orc.accd_grad(3, a, b, 1e-3, False)
    assert hit
    ap, am = a.copy(), a.copy()
    ap[2, 0] += h
    am[2, 0] -= h
    fwd = (orc.accd(3, ap, b, 1e-3, False)[1] - t0) / h
    bwd = (t0 - orc.accd(3, am, b, 1e-3, False)[1]) / h
    assert abs(fwd - bwd) > 1.0
    assert min(abs(ga[2, 0] - fwd), abs(ga[2, 0] - bwd)) < 1e-5


def test_oracle_toi_gradient_in_quirk_region():
    # there s = clamp(-e/c) is not a minimiser, so the gradient must carry
    # ds/dx; holding s (as in the other branches) fails this check
    h, eta, checked = 1e-7, 1e-2, 0
    for v0, v1 in quirk_stencils(30):
        hit, toi, g0, g1 = orc.accd_grad(3, v0, v1, eta, True)
        if not hit:
            continue
        for arr, g in ((0, g0), (1, g1)):
            for j in range(4):
                for k in range(3):
                    A0, A1 = v0.copy(), v1.copy()
                    (A0 if arr == 0 else A1)[j, k] += h
                    B0, B1 = v0.copy(), v1.copy()
                    (B0 if arr == 0 else B1)[j, k] -= h
                    hp, tp, _ = orc.accd(3, A0, A1, eta, True)
                    hm, tm, _ = orc.accd(3, B0, B1, eta, True)
                    assert hp and hm
                    assert abs((tp - tm) / (2 * h) - g[j, k]) <= 1e-6 * max(1.0, abs(g[j, k]))
                    checked += 1
    assert checked >= 600
