"""Minimal OBJ reading for the fixture meshes (positions, UVs, triangles)."""
from __future__ import annotations

import gzip

import numpy as np


def read_obj(path: str):
    """(V [n,3], VT [m,2], F [f,3] position indices, FT [f,3] uv indices or None); 0-based."""
    opener = gzip.open if path.endswith(".gz") else open
    V, VT, F, FT = [], [], [], []
    with opener(path, "rt") as fh:
        for line in fh:
            t = line.split()
            if not t:
                continue
            if t[0] == "v":
                V.append([float(x) for x in t[1:4]])
            elif t[0] == "vt":
                VT.append([float(x) for x in t[1:3]])
            elif t[0] == "f":
                idx = [c.split("/") for c in t[1:4]]
                F.append([int(c[0]) - 1 for c in idx])
                if all(len(c) > 1 and c[1] for c in idx):
                    FT.append([int(c[1]) - 1 for c in idx])
    V = np.asarray(V, np.float64)
    VT = np.asarray(VT, np.float64).reshape(-1, 2)
    F = np.asarray(F, np.int32)
    FT = np.asarray(FT, np.int32) if len(FT) == len(F) and len(FT) else None
    return V, VT, F, FT


def unique_edges(F: np.ndarray) -> np.ndarray:
    """Undirected edges of a triangle list, lower index first, sorted."""
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    e.sort(axis=1)
    return np.unique(e, axis=0).astype(np.int32)
