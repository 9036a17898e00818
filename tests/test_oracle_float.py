"""Pins of the oracle's float tree (BVHTree<float,DIM>, BVHTree.cpp:504-507).

- ACCD at T = float (oracle additive_ccd<D, float>, NarrowPhase.cpp:634-711)
  against a second, independent restatement below in numpy float32 scalars:
  every '+', '-', '*', '/', sqrt a float operation, except the step
  t_l = 0.9 * (...) / (...), whose double literal makes C++ compute the
  product and quotient in double before the float store.  Hit flag, TOI bits
  and distance-evaluation count must agree on every stencil.  A restatement
  that took 1/3 in double or evaluated the step in float disagrees on some of
  these stencils (checked below by planting each change in the Python copy),
  and the float TOIs differ from the double path's on many hits.
- Float leaf boxes: float min/max of the float vertices padded by the float
  eta in float arithmetic (BVHTree.cpp:83-124 at T = float), numpy float32.
- Float KdBVH centres: (min + max) / 2 in float — the rank order equals
  mclg_kdbvh_order over float32 centres.
- The float tree's hits are exactly the float ACCD hits of its candidates.
"""
import numpy as np
import pytest

from mclccd_amd import scenes
from oracle import oracle as O
from oracle.oracle import run_scene

f32, f64 = np.float32, np.float64


def _dot(a, b):
    if len(a) == 3:
        return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]
    return a[0] * b[0] + a[1] * b[1]


def _sub(a, b):
    return [x - y for x, y in zip(a, b)]


def _norm(a):
    return np.sqrt(_dot(a, a))


def _clamp_zero_one(v):  # Projection.hpp:63
    return f32(0) if v < 0 else (f32(1) if v > 1 else v)


def _point_on_triangle(q, p1, p2, p3):  # Projection.hpp:60-145 at T = float
    e0, e1, v0 = _sub(p2, p1), _sub(p3, p1), _sub(p1, q)
    a, b, c, d, e = _dot(e0, e0), _dot(e0, e1), _dot(e1, e1), _dot(e0, v0), _dot(e1, v0)
    det = a * c - b * b
    s = b * e - c * d
    t = b * d - a * e
    zero, one, two = f32(0), f32(1), f32(2)
    if s + t < det:
        if s < zero:
            if t < zero:
                if d < zero:
                    s, t = _clamp_zero_one(-d / a), zero
                else:
                    s, t = zero, _clamp_zero_one(-e / c)
            else:
                s, t = zero, _clamp_zero_one(-e / c)
        elif t < zero:
            s, t = _clamp_zero_one(-d / a), zero
        else:
            inv = one / det
            s, t = s * inv, t * inv
    else:
        if s < zero:
            t0, t1 = b + d, c + e
            if t1 > t0:
                s = _clamp_zero_one((t1 - t0) / (a - two * b + c))
                t = one - s
            else:
                t, s = _clamp_zero_one(-e / c), zero
        elif t < zero:
            if a + d > b + e:
                s = _clamp_zero_one((c + e - b - d) / (a - two * b + c))
                t = one - s
            else:
                s, t = _clamp_zero_one(-e / c), zero
        else:
            s = _clamp_zero_one((c + e - b - d) / (a - two * b + c))
            t = one - s
    return [(p1[k] + e0[k] * s) + e1[k] * t for k in range(3)]


def _clamp01(x):  # Projection.hpp:250-253: std::min(T(1), std::max(x, T(0)))
    m = f32(0) if x < f32(0) else x
    return m if m < f32(1) else f32(1)


def _edge_to_edge(p0, p1, q0, q1):  # Projection.hpp:241-302 at T = float
    d1, d2, r = _sub(p1, p0), _sub(q1, q0), _sub(p0, q0)
    a, e, f = _dot(d1, d1), _dot(d2, d2), _dot(d2, r)
    c, b = _dot(d1, r), _dot(d1, d2)
    denom = a * e - b * b
    s = _clamp01((b * f - c * e) / denom) if denom != f32(0) else f32(0)
    tnom = b * s + f
    if tnom < 0 or e == 0:
        t = f32(0)
        s = f32(0) if a == 0 else _clamp01(-c / a)
    elif tnom > e:
        t = f32(1)
        s = f32(0) if a == 0 else _clamp01((b - c) / a)
    else:
        t = tnom / e
    return [(q0[k] + t * d2[k]) - (p0[k] + s * d1[k]) for k in range(3)]


def _point_on_edge(p, e0, e1):  # Projection.hpp:193-204 at T = float
    e = _sub(e1, e0)
    l2 = _dot(e, e)
    if l2 <= 0.0:
        return e0
    t = _dot(_sub(p, e0), e) / l2
    if t < 0.0:
        return e0
    if t > 1.0:
        return e1
    return [e0[0] + t * e[0], e0[1] + t * e[1]]


def _pair_distance(x, is_vf):
    if len(x[0]) == 2:
        return _norm(_sub(x[0], _point_on_edge(x[0], x[1], x[2])))
    if is_vf:
        return _norm(_sub(x[0], _point_on_triangle(x[0], x[1], x[2], x[3])))
    return _norm(_edge_to_edge(x[0], x[1], x[2], x[3]))


def accd_float(v0, v1, eta, is_vf, plant=None):
    """NarrowPhase.cpp:634-711 with T = float, in numpy float32 scalars.
    plant: a deliberate deviation (test of the test)."""
    ns, dim = len(v0), len(v0[0])
    v0 = [[f32(c) for c in row] for row in v0]
    v1 = [[f32(c) for c in row] for row in v1]
    s, xsi, t_c = f32(1), f32(eta), f32(1)
    p = [_sub(v1[i], v0[i]) for i in range(ns)]
    x = [list(row) for row in v0]
    pbar = [f32(0)] * dim
    w = f32(1) / f32(ns) if plant != "third_in_double" else f64(1) / f64(ns)
    for i in range(ns):
        pbar = [f32(pbar[k] + w * p[i][k]) for k in range(dim)]
    p = [_sub(p[i], pbar) for i in range(ns)]
    lf = ls = f32(0)
    for i in range(ns):
        n = _norm(p[i])
        if (i == 0) if is_vf else (i <= 1):
            lf = n if lf < n else lf
        else:
            ls = n if ls < n else ls
    l_p = lf + ls
    if l_p <= 0:
        return False, -1.0, 0
    d = _pair_distance(x, is_vf)
    evals = 1
    g = s * (d * d - xsi * xsi) / (d - xsi)
    toi = f32(0)
    t_l = (f32(1) - s) * (d * d - xsi * xsi) / ((d + xsi) * l_p)
    if d <= xsi:
        return True, 0.0, evals
    for it in range(1000):
        x = [[x[i][k] + t_l * p[i][k] for k in range(dim)] for i in range(ns)]
        d = _pair_distance(x, is_vf)
        evals += 1
        eps = (d * d - xsi * xsi) / (d + xsi)
        if toi > 0 and it > 0 and eps < g:
            break
        toi = toi + t_l
        if toi > t_c:
            return False, -1.0, evals
        if plant == "step_in_float":
            t_l = f32(0.9) * (d * d - xsi * xsi) / ((d + xsi) * l_p)
        else:
            t_l = f32(f64(0.9) * f64(d * d - xsi * xsi) / f64((d + xsi) * l_p))
    return True, float(toi), evals


def _stencils(dim, n, seed):
    """Candidate stencils of a float tree on a random scene (VF and EE)."""
    if dim == 3:
        V0, V1, F, _ = scenes.random_soup(120, seed=seed, share=True, motion=0.25)
    else:
        V0, V1, F, _ = scenes.disk2d(npts=400, seed=seed)
    V0, V1 = V0.astype(np.float32).astype(np.float64), V1.astype(np.float32).astype(np.float64)
    o = run_scene(V0, V1, F, 1e-2, f32=True)
    sten, vf = o.candidates()
    rng = np.random.default_rng(seed)
    pick = rng.choice(len(sten), size=min(n, len(sten)), replace=False)
    out = []
    for i in pick:
        ids = [int(v) for v in sten[i][:dim + 1]]
        out.append((V0[ids], V1[ids], bool(vf[i]) or dim == 2))
    return out


@pytest.mark.parametrize("dim,seed", [(3, 1), (3, 2), (2, 3)])
def test_float_accd_equals_float32_restatement(dim, seed):
    cases = _stencils(dim, 700, seed)
    hits = differ_from_double = 0
    for a, b, vf in cases:
        want = accd_float(a, b, 1e-2, vf)
        got = O.accd_f32(dim, a, b, 1e-2, vf)
        assert got[0] == want[0] and got[2] == want[2], (a, b, vf, got, want)
        assert got[1] == want[1], (a, b, vf, got, want)  # TOI bit for bit
        hits += got[0]
        dh = O.accd(dim, a, b, float(np.float32(1e-2)), vf)
        differ_from_double += got[0] and dh[0] and dh[1] != got[1]
    assert hits > 50
    assert differ_from_double > 10  # the float path is not the double one rounded


@pytest.mark.parametrize("plant", ["third_in_double", "step_in_float"])
def test_planted_float_mistakes_are_caught(plant):
    dim = 2 if plant == "third_in_double" else 3
    cases = _stencils(dim, 400, 5)
    bad = sum(accd_float(a, b, 1e-2, vf, plant=plant)[1:] != O.accd_f32(dim, a, b, 1e-2, vf)[1:]
              for a, b, vf in cases)
    assert bad > 0


@pytest.mark.parametrize("dim", [3, 2])
def test_float_tree_boxes_centres_and_hits(dim):
    if dim == 3:
        V0, V1, F, _ = scenes.random_soup(150, seed=4, share=True, motion=0.2)
    else:
        V0, V1, F, _ = scenes.disk2d(npts=300, seed=4)
    eta = 1e-2
    o = run_scene(V0, V1, F, eta, f32=True)
    # boxes: float min/max of float vertices, padded in float
    a0, a1 = V0.astype(np.float32), V1.astype(np.float32)
    ef = np.float32(eta)
    b, _ = o.boxes()
    for i, f in enumerate(F):
        lo0, hi0 = a0[f].min(0), a0[f].max(0)
        lo1, hi1 = a1[f].min(0), a1[f].max(0)
        want = np.concatenate([np.minimum(lo0, lo1) - ef, np.maximum(hi0, hi1) + ef, lo0 - ef, hi0 + ef,
                               lo1 - ef, hi1 + ef]).astype(np.float64)
        assert np.array_equal(b[i], want), i
    # KdBVH order over float centres
    d = dim
    cen = ((b[:, :d].astype(np.float32) + b[:, d:2 * d].astype(np.float32)) / np.float32(2)).astype(np.float64)
    from mclccd_amd.tree import kdbvh_order

    rank, _ = kdbvh_order(cen)
    assert np.array_equal(rank, o.rank())
    # hits = float ACCD over the candidates
    sten, vf = o.candidates()
    hs, hv, ht = o.hits()
    got = sorted(zip(map(tuple, hs.tolist()), hv.tolist(), ht.tolist()))
    V0f, V1f = a0.astype(np.float64), a1.astype(np.float64)
    want = []
    for s, v in zip(sten.tolist(), vf.tolist()):
        ids = s[:d + 1]
        h, t, _ = O.accd_f32(d, V0f[ids], V1f[ids], eta, bool(v) or d == 2)
        if h:
            want.append((tuple(s), v, t))
    assert got == sorted(want) and len(got) > 20
    assert all(np.float32(t) == t for t in ht)
