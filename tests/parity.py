"""Comparison helpers for GPU-vs-oracle parity (test infrastructure).

Orientation: the reference orients each leaf pair by its KdBVH in-order
position (include/MCL/BVHTree.cpp:288-301), which fixes the order of the two
edges of an EE stencil and the argument order of a discrete hit.  Canonical
forms compare results independently of orientation; in strict mode (the GPU
given the oracle's rank) results are compared as oriented records.
"""
from __future__ import annotations

import numpy as np


def canon_pairs(pairs) -> set:
    p = np.sort(np.asarray(pairs, dtype=np.int64).reshape(-1, 2), axis=1)
    return set(map(tuple, p.tolist()))


def oriented_pairs(pairs) -> list:
    return sorted(map(tuple, np.asarray(pairs).reshape(-1, 2).tolist()))


def _canon_key(s, vf):
    s = tuple(int(x) for x in s)
    if vf:
        return (1,) + s
    e0, e1 = s[:2], s[2:]
    return (0,) + (e0 + e1 if e0 <= e1 else e1 + e0)


def canon_stencils(sten, is_vf) -> list:
    return sorted(_canon_key(s, v) for s, v in zip(np.asarray(sten).tolist(), np.asarray(is_vf).tolist()))


def oriented_stencils(sten, is_vf) -> list:
    return sorted((int(v),) + tuple(s) for s, v in zip(np.asarray(sten).tolist(), np.asarray(is_vf).tolist()))


def hit_map(sten, is_vf, toi, canonical=True) -> dict:
    """key -> sorted list of toi (lists, because without reptri a stencil may repeat)."""
    out: dict = {}
    for s, v, t in zip(np.asarray(sten).tolist(), np.asarray(is_vf).tolist(), np.asarray(toi).tolist()):
        k = _canon_key(s, v) if canonical else (int(v),) + tuple(s)
        out.setdefault(k, []).append(t)
    for k in out:
        out[k].sort()
    return out


def compare_hits(gpu, ref, canonical: bool, rtol: float = 0.0):
    """Returns (n_mismatch_keys, max_rel_diff, n_over) between two hit maps.
    n_over counts GPU TOIs above the reference's by more than rtol (relative)."""
    keys = set(gpu) | set(ref)
    bad = 0
    maxrel = 0.0
    over = 0
    for k in keys:
        a, b = gpu.get(k), ref.get(k)
        if a is None or b is None or len(a) != len(b):
            bad += 1
            continue
        for x, y in zip(a, b):
            d = abs(x - y)
            rel = d / max(abs(y), 1e-300) if d > 0 else 0.0
            maxrel = max(maxrel, rel)
            if x > y * (1 + rtol) + (0 if y != 0 else rtol):
                over += 1
            if rel > rtol:
                bad += 1
    return bad, maxrel, over


def grad_close(got, ref, rtol: float = 1e-6, floor: float = 1e-12) -> float:
    """Worst per-entry error of a gradient against a reference gradient, in
    units of the allowed error rtol*|ref| + floor*max|ref|: <= 1 passes.  The
    floor only matters for entries near zero, where the summation order of
    many contributions decides the last bits."""
    got, ref = np.asarray(got, np.float64), np.asarray(ref, np.float64)
    tol = rtol * np.abs(ref) + floor * max(float(np.abs(ref).max()), 1e-300)
    return float((np.abs(got - ref) / tol).max())


def quirk_region(q, a, b, c):
    """Projection.hpp:123-133: for s >= 0, t < 0, s + t >= det and a+d <= b+e the
    reference sets s = clamp(-e/c), t = 0, i.e. a point of edge (p1,p2) chosen
    with edge (p1,p3)'s parameter, which is not the closest point.  Returns
    that point when q falls in the region, else None."""
    e0, e1, w = b - a, c - a, a - q
    A, B, C = e0 @ e0, e0 @ e1, e1 @ e1
    D, E = e0 @ w, e1 @ w
    det = A * C - B * B
    s, t = B * E - C * D, B * D - A * E
    if s + t >= det and s >= 0 and t < 0 and not (A + D > B + E):
        r = -E / C
        return a + e0 * min(max(r, 0.0), 1.0)
    return None


def quirk_stencils(count, seed=11):
    """VF stencils (vertex, triangle) whose vertex starts in the reference's
    point_on_triangle quirk region, moving toward the quirk point and
    stopping 0.005 short of it; the triangle drifts slightly (no l_p ties)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        a, b, c = rng.normal(size=(3, 3))
        q = rng.normal(size=3) * 1.5
        pt = quirk_region(q, a, b, c)
        if pt is None:
            continue
        e1, w = c - a, a - q
        if not 0.05 < -(e1 @ w) / (e1 @ e1) < 0.95:
            continue
        v0 = np.array([q, a, b, c])
        v1 = v0.copy()
        v1[0] = q + (pt - q) * (1 - 0.5e-2 / np.linalg.norm(q - pt))
        v1[1:] += rng.normal(0, 1e-3, (3, 3))
        out.append((v0, v1))
    return out


def boundary_pairs(eta: float, exact_floats: bool = False):
    """Isolated triangle pairs whose leaf boxes meet within a few FP32 / FP64
    ulps of the leaf test's threshold (BVHTree.cpp:401-426), along each axis,
    from above and below, on the swept boxes (static pairs) and on the t1 boxes
    (t0 separated, so only t1 decides).  Pair j sits near 20 j along x, so the
    FP32 spacing ranges from 1e-7 to ~1e-3 across the scene.  With
    exact_floats the coordinates are FP32 values (FP32 bounds then equal the
    FP64 ones for eta = 0, ties included).  Returns V0, V1, F."""
    V0, V1, F = [], [], []
    j = 0
    for kind in ("swept", "t1"):
        for axis in range(3):
            for side in (1, -1):
                for d32 in range(-3, 4):
                    for d64 in (range(-2, 3) if not exact_floats else (0,)):
                        base = np.zeros(3)
                        base[0] = 20.0 * j + 1.0
                        c = base[axis] if axis == 0 else 1.0 + 0.25 * (j % 7)
                        if exact_floats:
                            c = float(np.float32(c))
                        u32 = float(np.spacing(np.float32(c)))
                        u64 = float(np.spacing(c))
                        gap = 2.0 * eta + d32 * u32 + d64 * u64
                        e = np.eye(3)[axis]
                        o1, o2 = np.eye(3)[(axis + 1) % 3], np.eye(3)[(axis + 2) % 3]
                        org = base.copy()
                        org[axis] = 0.0
                        # A: extreme coordinate c along `axis` (max if side=1, min if side=-1)
                        A = [org + side * c * e, org + side * (c - 0.3) * e + 0.1 * o1,
                             org + side * (c - 0.2) * e + 0.1 * o2]
                        b = c + gap  # B's extreme coordinate facing A
                        B = [org + side * b * e + 0.05 * o1, org + side * (b + 0.3) * e + 0.02 * o2,
                             org + side * (b + 0.2) * e + 0.1 * o1 + 0.1 * o2]
                        A, B = np.array(A), np.array(B)
                        if exact_floats:
                            A, B = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
                        B0 = B.copy()
                        if kind == "t1":
                            B0 = B + side * 0.4 * e  # separated at t0 along the same axis
                        base_i = len(V0)
                        V0 += list(A) + list(B0)
                        V1 += list(A) + list(B)
                        F += [[base_i, base_i + 1, base_i + 2], [base_i + 3, base_i + 4, base_i + 5]]
                        j += 1
    return np.array(V0), np.array(V1), np.array(F, np.int32)
