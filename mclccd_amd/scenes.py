"""Deterministic synthetic scenes for the collision path (numpy, FP64).

Every generator returns ``(V0, V1, P, theta)`` with ``V0``/``V1`` row-major
``nv x dim`` float64, ``P`` row-major ``np x pdim`` int32 and ``theta`` the ACCD
thickness to use as ``box_eta = vf_ccd_eta = ee_ccd_eta`` (the reference clamps
the narrow-phase gaps to ``box_eta``, include/MCL/BVHTree.cpp:39-40).

Workloads:
  cloth_fold        64x64-quad sheet folded onto itself (8,192 tris), seed 0
  stacked_layers    16 layers of 180x180 quads (1,036,800 tris), rigid tilt +
                    z-twist + noise per layer, seed 2
  disk2d            150k random points in the unit disk, Delaunay triangulated,
                    swirl motion; the 2D tree's primitives are edges, seed 1
  batch_sheets      512 independent folded sheets, n in [16,180] log-uniform, seed 3
"""
from __future__ import annotations

import math

import numpy as np


def grid_sheet(n: int, z: float = 0.0):
    """(n+1)^2 vertices on the unit square, vertex (i,j) = (i/n, j/n, z) at
    index i*(n+1)+j; quad (i,j) with corners a=(i,j), b=(i+1,j), c=(i+1,j+1),
    d=(i,j+1) is split into triangles (a,b,c) and (a,c,d)."""
    m = n + 1
    ii, jj = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    V = np.stack([ii.ravel() / n, jj.ravel() / n, np.full(m * m, z)], axis=1).astype(np.float64)
    qi, qj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = (qi * m + qj).ravel()
    b = ((qi + 1) * m + qj).ravel()
    c = ((qi + 1) * m + qj + 1).ravel()
    d = (qi * m + qj + 1).ravel()
    F = np.empty((2 * n * n, 3), dtype=np.int32)
    F[0::2] = np.stack([a, b, c], axis=1)
    F[1::2] = np.stack([a, c, d], axis=1)
    return V, F


def cloth_fold(n: int = 64, amp: float = 0.3, fx: float = 3.0, fy: float = 2.0, noise_rel: float = 0.01,
               seed: int = 0, theta: float = 1e-3):
    """Sheet on the unit square at z=0; V1 = V0 + amp*sin(fx*pi*x)*sin(fy*pi*y) in z
    plus i.i.d. Gaussian noise of sigma = noise_rel * edge length."""
    rng = np.random.default_rng(seed)
    V0, F = grid_sheet(n)
    V1 = V0.copy()
    V1[:, 2] += amp * np.sin(fx * np.pi * V0[:, 0]) * np.sin(fy * np.pi * V0[:, 1])
    V1 += rng.standard_normal(V0.shape) * (noise_rel / n)
    return V0, np.ascontiguousarray(V1), F, theta


def _rot(axis, ang):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    x, y, z = axis
    c, s = math.cos(ang), math.sin(ang)
    C = 1.0 - c
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


def stacked_layers(layers: int = 16, n: int = 180, spacing: float = 0.02, max_tilt: float = 0.3,
                   max_twist: float = 0.5, sigma: float = 0.005, seed: int = 2, theta: float = 1e-4):
    """Stacked cloth layers at z = spacing*k.  Per layer (in layer order) the
    generator draws a rotation axis uniform on the sphere, a tilt angle
    U(0,max_tilt) and a twist angle U(0,max_twist); V1 is the layer rotated
    about its centroid by the tilt, then about the z axis through the centroid
    by the twist; finally Gaussian noise sigma is added to every V1 coordinate."""
    rng = np.random.default_rng(seed)
    V0s, V1s, Fs = [], [], []
    off = 0
    for k in range(layers):
        V, F = grid_sheet(n, z=spacing * k)
        axis = rng.standard_normal(3)
        tilt = rng.uniform(0.0, max_tilt)
        twist = rng.uniform(0.0, max_twist)
        R = _rot([0.0, 0.0, 1.0], twist) @ _rot(axis, tilt)
        c = V.mean(axis=0)
        V1 = (V - c) @ R.T + c
        V0s.append(V)
        V1s.append(V1)
        Fs.append(F + off)
        off += V.shape[0]
    V0 = np.ascontiguousarray(np.concatenate(V0s))
    V1 = np.concatenate(V1s)
    V1 += rng.standard_normal(V1.shape) * sigma
    return V0, np.ascontiguousarray(V1), np.ascontiguousarray(np.concatenate(Fs).astype(np.int32)), theta


def disk2d(npts: int = 150_000, sigma: float = 1e-3, seed: int = 1, theta: float = 1e-3, edges: str = "all"):
    """2D disk: npts uniform points in the unit disk, Delaunay triangulated.
    V1 = V0 + 0.4*r*exp(-r^2) * t_hat + N(0, sigma^2), t_hat the unit tangent.
    The 2D tree takes edges: ``edges='all'`` gives every triangulation edge,
    ``edges='boundary'`` the hull edges only.  Edges are listed (min,max)
    sorted lexicographically."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.uniform(0.0, 1.0, npts))
    ang = rng.uniform(0.0, 2 * np.pi, npts)
    V0 = np.stack([r * np.cos(ang), r * np.sin(ang)], axis=1)
    tri = Delaunay(V0).simplices.astype(np.int64)
    rr = np.linalg.norm(V0, axis=1)
    # 0.4*r*exp(-r^2) * (-y, x)/r
    vel = 0.4 * np.exp(-rr * rr)[:, None] * np.stack([-V0[:, 1], V0[:, 0]], axis=1)
    V1 = V0 + vel + rng.standard_normal(V0.shape) * sigma
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    e.sort(axis=1)
    if edges == "boundary":
        key = e[:, 0] * npts + e[:, 1]
        uk, cnt = np.unique(key, return_counts=True)
        uk = uk[cnt == 1]
    else:
        uk = np.unique(e[:, 0] * npts + e[:, 1])
    E = np.stack([uk // npts, uk % npts], axis=1).astype(np.int32)
    return np.ascontiguousarray(V0), np.ascontiguousarray(V1), np.ascontiguousarray(E), theta


def batch_sheets(count: int = 512, nmin: int = 16, nmax: int = 180, seed: int = 3, theta: float = 1e-3):
    """``count`` independent folded sheets.  Per scene (in order) draws
    n = round(exp(U(ln nmin, ln nmax))), amplitude U(0.1,0.5), integer fold
    frequencies fx, fy in {1..4}; noise sigma = 0.01/n from the same stream.
    Returns a list of (V0, V1, F, theta)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        amp = rng.uniform(0.1, 0.5)
        fx = int(rng.integers(1, 5))
        fy = int(rng.integers(1, 5))
        V0, F = grid_sheet(n)
        V1 = V0.copy()
        V1[:, 2] += amp * np.sin(fx * np.pi * V0[:, 0]) * np.sin(fy * np.pi * V0[:, 1])
        V1 += rng.standard_normal(V0.shape) * (0.01 / n)
        out.append((V0, np.ascontiguousarray(V1), F, theta))
    return out


def random_soup(ntri: int, seed: int, spread: float = 1.0, size: float = 0.15, motion: float = 0.1,
                share: bool = True):
    """Small random triangle soup (with some shared vertices when ``share``)
    for edge-case tests; V1 = V0 + uniform motion."""
    rng = np.random.default_rng(seed)
    if share:
        nv = max(3, int(ntri * 1.2))
        V0 = rng.uniform(0, spread, (nv, 3))
        F = np.empty((ntri, 3), dtype=np.int32)
        for i in range(ntri):
            a = rng.integers(0, nv)
            # pick two other vertices near a so that triangles stay small
            d = np.linalg.norm(V0 - V0[a], axis=1)
            near = np.argsort(d)[1:8]
            b, c = rng.choice(near, 2, replace=False)
            F[i] = (a, b, c)
    else:
        c0 = rng.uniform(0, spread, (ntri, 1, 3))
        V0 = (c0 + rng.uniform(-size, size, (ntri, 3, 3))).reshape(-1, 3)
        F = np.arange(3 * ntri, dtype=np.int32).reshape(ntri, 3)
    V1 = V0 + rng.uniform(-motion, motion, V0.shape)
    return np.ascontiguousarray(V0), np.ascontiguousarray(V1), F, 1e-3
