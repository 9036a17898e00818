"""Pins of the CPU oracle's CTCD narrow phase (the vf_one_sided path).

The reference ships no tests or vectors for CTCD, so each part is checked
against what the mathematics fixes:
  - the Jenkins-Traub root finder against numpy.roots (companion-matrix
    eigenvalues), including complex pairs, zeros at the origin and a zero
    leading coefficient;
  - the sign intervals of a polynomial over [0, 1] against its known roots;
  - vertex-face times of impact on a static face against dense sampling of the
    geometric conditions (distance to the plane <= eta and the projection
    inside the triangle), whose first time is what CTCD reports;
  - closed-form times for the vertex-edge and vertex-vertex sub-tests and the
    2D vertex-edge test;
  - the wrong-side filter's orientation (approach from the normal's side hits,
    from behind is rejected) in 2D and 3D;
  - the ray-box prefilter against sampling of the segment.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _match_roots(got_r, got_i, want, tol):
    want = list(want)
    for r, i in zip(got_r, got_i):
        z = complex(r, i)
        d = [abs(z - w) for w in want]
        k = int(np.argmin(d))
        assert d[k] <= tol * max(1.0, abs(want[k])), (z, want)
        want.pop(k)
    assert not want


@pytest.mark.parametrize("deg", [3, 4, 5, 6])
def test_rpoly_real_roots(deg):
    rng = np.random.default_rng(deg)
    for _ in range(200):
        roots = rng.uniform(-2, 2, deg)
        roots[np.argsort(roots)] += np.arange(deg) * 0.05  # keep them apart
        c = np.poly(roots) * rng.uniform(0.5, 2.0)
        n, zr, zi = O.rpoly(c)
        assert n == deg
        _match_roots(zr[:n], zi[:n], roots, 1e-8)


@pytest.mark.parametrize("deg", [3, 4, 6])
def test_rpoly_complex_pairs(deg):
    rng = np.random.default_rng(10 + deg)
    for _ in range(100):
        z = complex(rng.uniform(-1, 1), rng.uniform(0.2, 1))
        rest = rng.uniform(-2, 2, deg - 2)
        roots = np.concatenate([[z, z.conjugate()], rest])
        c = np.real(np.poly(roots))
        n, zr, zi = O.rpoly(c)
        assert n == deg
        _match_roots(zr[:n], zi[:n], roots, 1e-7)
        np.testing.assert_allclose(np.sort_complex(zr[:n] + 1j * zi[:n]), np.sort_complex(np.roots(c)), atol=1e-7)


def test_rpoly_special_cases():
    # zeros at the origin are reported first, exactly
    n, zr, zi = O.rpoly([1.0, -3.0, 2.0, 0.0, 0.0])
    assert n == 4 and zr[0] == 0.0 and zr[1] == 0.0
    assert sorted(zr[2:4]) == pytest.approx([1.0, 2.0])
    # zero leading coefficient: failure code
    assert O.rpoly([0.0, 1.0, 2.0, 3.0])[0] == -1
    # x^3 (only zeros at the origin): the reference returns -1
    assert O.rpoly([1.0, 0.0, 0.0, 0.0])[0] == -1


def test_find_intervals_signs():
    # (t - 0.3)(t - 0.7): <= 0 on [0.3, 0.7], >= 0 outside
    c = np.poly([0.3, 0.7])
    iv = O.find_intervals(c, pos=False)
    assert len(iv) == 1 and iv[0] == pytest.approx((0.3, 0.7), abs=1e-15)
    iv = O.find_intervals(c, pos=True)
    assert [tuple(x) for x in iv] == [pytest.approx((0.0, 0.3)), pytest.approx((0.7, 1.0))]
    # degree 4 and 6 with roots inside and outside [0, 1] (the root finder path)
    # (t-.2)(t-.5)(t-.8)(t+3) is negative on [0, .2] and [.5, .8]
    c4 = np.poly([0.2, 0.5, 0.8, -3.0])
    iv = O.find_intervals(c4, pos=False)
    assert [tuple(x) for x in iv] == [pytest.approx((0.0, 0.2), abs=1e-12), pytest.approx((0.5, 0.8), abs=1e-12)]
    iv = O.find_intervals(c4, pos=True)
    assert [tuple(x) for x in iv] == [pytest.approx((0.2, 0.5), abs=1e-12), pytest.approx((0.8, 1.0), abs=1e-12)]
    # five of the six factors negative on [0, .1]: p <= 0 there, then alternating
    c6 = np.poly([0.1, 0.25, 0.4, 0.6, 1.5, -0.5])
    want_neg = [(0.0, 0.1), (0.25, 0.4), (0.6, 1.0)]
    got = O.find_intervals(c6 * -1.0, pos=True)
    assert [tuple(x) for x in got] == [pytest.approx(w, abs=1e-12) for w in want_neg]
    assert [tuple(x) for x in O.find_intervals(c6, pos=False)] == [pytest.approx(w, abs=1e-12) for w in want_neg]
    # constant polynomial: whole interval or nothing
    assert O.find_intervals([0.0, 0.0, -1.0], pos=False) == [(0.0, 1.0)]
    assert O.find_intervals([0.0, 0.0, 1.0], pos=False) == []


def _tri():
    return np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])


def test_vertex_face_closed_form_and_wrong_side():
    eta = 1e-3
    f = _tri()
    p0, p1 = np.array([0.2, 0.2, 1.0]), np.array([0.2, 0.2, -1.0])
    # face normal (q2-q1)x(q3-q1) = +z: approaching from +z is the right side
    v0, v1 = np.vstack([p0, f]), np.vstack([p1, f])
    code, toi = O.ctcd_vf(3, v0, v1, eta)
    assert code == 1 and toi == pytest.approx((1 - eta) / 2, abs=1e-12)
    # from behind: only a wrong-side hit
    code, toi = O.ctcd_vf(3, np.vstack([p1, f]), np.vstack([p0, f]), eta)
    assert code == -1 and toi == -1.0
    # passing beside the face: no hit
    q0, q1 = np.array([2.0, 2.0, 1.0]), np.array([2.0, 2.0, -1.0])
    assert O.ctcd_vf(3, np.vstack([q0, f]), np.vstack([q1, f]), eta)[0] == 0


def _brute_vf(p0, p1, f, eta, m=200001):
    t = np.linspace(0.0, 1.0, m)
    p = p0[None] + t[:, None] * (p1 - p0)[None]
    n = np.cross(f[1] - f[0], f[2] - f[0])
    n = n / np.linalg.norm(n)
    dist = (p - f[0]) @ n
    # barycentric coordinates of the projection
    e0, e1 = f[1] - f[0], f[2] - f[0]
    w = p - f[0]
    d00, d01, d11 = e0 @ e0, e0 @ e1, e1 @ e1
    d20, d21 = w @ e0, w @ e1
    den = d00 * d11 - d01 * d01
    b1 = (d11 * d20 - d01 * d21) / den
    b2 = (d00 * d21 - d01 * d20) / den
    ok = (np.abs(dist) <= eta) & (b1 >= 0) & (b2 >= 0) & (b1 + b2 <= 1)
    idx = np.nonzero(ok)[0]
    return (t[idx[0]] if len(idx) else None), 1.0 / (m - 1)


def test_vertex_face_dense_sampling():
    """Static face, moving vertex: the first time both conditions hold."""
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(300):
        f = rng.uniform(-1, 1, (3, 3))
        p0 = rng.uniform(-1, 1, 3)
        p1 = rng.uniform(-1, 1, 3)
        eta = 10 ** rng.uniform(-3, -1.5)
        hit, ts = O.ctcd_subtest(0, np.vstack([p0, f]), np.vstack([p1, f]), eta)
        want, dt = _brute_vf(p0, p1, f, eta)
        if want is None:
            continue  # no sample satisfies both conditions (a graze between samples stays unchecked)
        assert hit, (p0, p1, f, eta)
        # the sampled first time lies within one step after the exact one
        assert min(ts) <= want + 1e-9 and want - min(ts) <= dt + 1e-9, (min(ts), want)
        checked += 1
    assert checked > 20


def test_vertex_edge_and_vertex_vertex_closed_forms():
    eta = 1e-3
    # edge (0,0,0)-(1,0,0) static; vertex crossing it at x = 0.5 along -y
    x0 = np.array([[0.5, 1.0, 0.0], [0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    x1 = np.array([[0.5, -1.0, 0.0], [0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    hit, ts = O.ctcd_subtest(1, x0, x1, eta)
    assert hit and min(ts) == pytest.approx((1 - eta) / 2, abs=1e-12)
    # vertex passing beyond the edge's end: no vertex-edge hit
    x0b, x1b = x0.copy(), x1.copy()
    x0b[0, 0] = x1b[0, 0] = 1.5
    assert not O.ctcd_subtest(1, x0b, x1b, eta)[0]
    # vertex-vertex: q1 moves through q2 at t = 0.5; |1 - 2t| <= eta from (1-eta)/2
    y0 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    y1 = np.array([[2.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    hit, ts = O.ctcd_subtest(2, y0, y1, eta)
    assert hit and min(ts) == pytest.approx((1 - eta) / 2, abs=1e-12)
    # touching at the start: roots -eta and eta.  The reference pushes 0 for
    # [0, 0], then min(curr_t, t1) = -eta for the two later (accumulated)
    # checks; the query skips negative times, so the hit time is 0.
    hit, ts = O.ctcd_subtest(2, np.array([[1.0, 0.0, 0.0], [1.0, 0.0, 0.0]]), y1, eta)
    assert hit and list(ts) == [0.0, -eta, -eta]


def test_ctcd_2d_closed_form_and_wrong_side():
    e = np.array([[0.0, 0.0], [1.0, 0.0]])
    # n = (dy, -dx) = (0, -1): the right side is y < 0
    good0, good1 = np.vstack([[0.5, -1.0], e]), np.vstack([[0.5, 1.0], e])
    code, toi = O.ctcd_vf(2, good0, good1, 1e-3)
    assert code == 1 and toi == 0.5
    code, toi = O.ctcd_vf(2, good1, good0, 1e-3)
    assert code == -1
    # passing beside the segment
    code, _ = O.ctcd_vf(2, np.vstack([[1.5, -1.0], e]), np.vstack([[1.5, 1.0], e]), 1e-3)
    assert code == 0


def test_wrong_side_orientation_3d():
    f = _tri()
    v0 = np.vstack([[0.2, 0.2, 1.0], f])
    v1 = np.vstack([[0.2, 0.2, -1.0], f])
    assert not O.wrong_side(3, v0, v1, 0.5)
    assert O.wrong_side(3, v1, v0, 0.5)
    assert not O.wrong_side(3, v1, v0, 0.0)  # resting contact: no way to tell


def test_ray_box_against_sampling():
    rng = np.random.default_rng(7)
    lo, hi = np.array([-0.5, -0.5, -0.5]), np.array([0.5, 0.5, 0.5])
    for _ in range(2000):
        p0, p1 = rng.uniform(-2, 2, 3), rng.uniform(-2, 2, 3)
        t = np.linspace(0, 1, 4001)[:, None]
        pts = p0 + t * (p1 - p0)
        inside = np.all((pts >= lo) & (pts <= hi), axis=1)
        got = O.ray_box(p0, p1, lo, hi)
        if inside.any():
            assert got
        # a segment that misses the box by a clear margin is rejected
        d = np.max(np.maximum(lo - pts, pts - hi), axis=1).min()
        if d > 1e-3:
            assert not got


def test_tree_uses_ctcd_for_vf_when_one_sided():
    """A falling vertex over a static face: the one-sided tree reports the
    CTCD time, the default tree the ACCD time."""
    V0 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.2, 0.2, 0.5], [1.2, 0.2, 0.5], [0.2, 1.2, 0.5]])
    V1 = V0.copy()
    V1[3:, 2] = -0.5
    P = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    eta = 1e-3
    one = O.Oracle(3)
    one.set_options(box_eta=eta, vf_ccd_eta=eta, ee_ccd_eta=eta, vf_one_sided=True)
    one.update(V0, V1, P)
    one.traverse()
    sten, is_vf, toi = one.hits()
    vf = {tuple(s): t for s, v, t in zip(sten, is_vf, toi) if v}
    assert vf[(3, 0, 1, 2)] == pytest.approx((0.5 - eta) / 1.0, abs=1e-12)
    for s, t in vf.items():
        code, t2 = O.ctcd_vf(3, V0[list(s)], V1[list(s)], eta)
        assert code == 1 and t2 == t
