"""Headless mesh check on the GPU path: the steps of the reference's example
viewer (examples/dillo.cpp:32-80), without the viewer.

    python -m mclccd_amd.cli MESH.obj[.gz] [--eta E] [--out PREFIX] [--dim 3]

Loads a triangle mesh, builds the tree with box_eta = float epsilon (as
dillo.cpp:58), runs update(V, V, F) and traverse(V, V, F) (static: no CCD hits,
the discrete self-intersections), gathers the internal volumes with a visitor
(dillo.cpp:112-146 collects 12 edges per box) and prints the same report:
face and vertex counts, build time, CCD time, discrete and CCD counts, box
gathering time.  With --out, writes PREFIX_boxes.obj (the box edges as OBJ
lines) and PREFIX_discrete.txt (the faces in a discrete intersection), which
is what the viewer draws.  --dim 2 reads the mesh's UV layout instead.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from .meshes import read_obj
from .tree import BVHTree


class BoxEdges:
    """Visitor collecting the 12 edges of every internal volume (dillo.cpp:112-146)."""

    def __init__(self, dim):
        self.dim = dim
        self.e0, self.e1 = [], []

    def intersect_volume(self, box) -> bool:
        lo, hi = np.asarray(box.lo, np.float64), np.asarray(box.hi, np.float64)
        if self.dim == 2:
            a, b, c, d = lo, np.array([hi[0], lo[1]]), hi, np.array([lo[0], hi[1]])
            for p, q in ((a, b), (b, c), (c, d), (d, a)):
                self.e0.append(p)
                self.e1.append(q)
            return True
        corner = lambda x, y, z: np.array([(lo, hi)[x][0], (lo, hi)[y][1], (lo, hi)[z][2]])
        for (x0, y0, z0), (x1, y1, z1) in (
                ((0, 0, 0), (1, 0, 0)), ((0, 0, 0), (0, 0, 1)), ((1, 0, 1), (1, 0, 0)), ((1, 0, 1), (0, 0, 1)),
                ((0, 1, 0), (1, 1, 0)), ((0, 1, 0), (0, 1, 1)), ((1, 1, 1), (1, 1, 0)), ((1, 1, 1), (0, 1, 1)),
                ((0, 0, 1), (0, 1, 1)), ((0, 0, 0), (0, 1, 0)), ((1, 0, 0), (1, 1, 0)), ((1, 0, 1), (1, 1, 1))):
            self.e0.append(corner(x0, y0, z0))
            self.e1.append(corner(x1, y1, z1))
        return True

    def intersect_object(self, leaf) -> bool:
        return False


def run(path: str, eta: float | None = None, out: str | None = None, dim: int = 3, stream=sys.stdout) -> dict:
    def say(*a):
        print(*a, file=stream, flush=True)

    V, VT, F, FT = read_obj(path)
    if dim == 2:
        if FT is None:
            raise ValueError(f"{path} has no UV layout")
        V, F = VT, FT
        P = np.ascontiguousarray(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]))
        P = np.unique(np.sort(P, axis=1), axis=0).astype(np.int32)  # the layout's edges
    else:
        P = F
    say(f"Num faces: {len(P)}, num verts: {len(V)}")
    tree = BVHTree(dim)
    tree.options.box_eta = float(np.finfo(np.float32).eps) if eta is None else float(eta)
    t0 = time.perf_counter()
    tree.update(V, V, P)
    t1 = time.perf_counter()
    say(f"Building the tree: {1e3 * (t1 - t0):.3f} ms")
    tree.traverse(V, V, P)
    t2 = time.perf_counter()
    say(f"Performing CCD: {1e3 * (t2 - t1):.3f} ms")
    c = tree.counts()
    disc = tree.discrete()
    faces = np.unique(disc.reshape(-1)) if len(disc) else np.zeros(0, np.int32)
    say(f"Num discrete isect: {len(faces)}")
    say(f"Num ccd isect: {c['hits']}")
    col = BoxEdges(dim)
    tree.visit(col)
    t3 = time.perf_counter()
    say(f"Gathering boxes to render: {1e3 * (t3 - t2):.3f} ms ({len(col.e0)} edges)")
    if out:
        with open(out + "_boxes.obj", "w") as fh:
            for p, q in zip(col.e0, col.e1):
                fh.write("v " + " ".join(f"{x:.17g}" for x in p) + (" 0" if dim == 2 else "") + "\n")
                fh.write("v " + " ".join(f"{x:.17g}" for x in q) + (" 0" if dim == 2 else "") + "\n")
            for i in range(len(col.e0)):
                fh.write(f"l {2 * i + 1} {2 * i + 2}\n")
        np.savetxt(out + "_discrete.txt", faces, fmt="%d")
    return dict(faces=len(P), verts=len(V), discrete_faces=len(faces), discrete_pairs=int(c["discrete"]),
                ccd_hits=int(c["hits"]), box_edges=len(col.e0))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m mclccd_amd.cli", description=__doc__.split("\n\n")[0])
    ap.add_argument("mesh")
    ap.add_argument("--eta", type=float, default=None, help="box_eta (default: float epsilon)")
    ap.add_argument("--out", default=None, help="prefix of the box-edge OBJ and discrete-face list")
    ap.add_argument("--dim", type=int, default=3, choices=(2, 3))
    a = ap.parse_args(argv)
    run(a.mesh, a.eta, a.out, a.dim)
    return 0


if __name__ == "__main__":
    sys.exit(main())
