"""Batched scenes on one GPU: many independent ``BVHTree`` problems per frame.

Each scene is what one ``mcl::BVHTree<double,dim>`` update + traverse solves
(reference include/MCL/BVHTree.cpp:27-500), with its own gaps.  The library
solves all of them in one pass (one LBVH keyed by scene and Morton code, the
frame replayed as a hipGraph); results come back tagged with their scene and
in scene-local indices.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import lib
from .tree import Options, _is_device, _ptr


class Batch:
    def __init__(self, dim: int = 3):
        h = C.c_void_p()
        rc = lib.mclg_batch_create(dim, C.byref(h))
        if rc != _lib.OK:
            raise _lib.MclgError(rc, "mclg_batch_create failed (no HIP device?)")
        self._h = h
        self.dim = dim
        self.nscenes = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.mclg_batch_destroy(h)
            self._h = None

    def _check(self, rc):
        if rc != _lib.OK:
            raise _lib.MclgError(rc, lib.mclg_batch_last_error(self._h).decode())

    def set_scenes(self, prims, nverts, options):
        """prims: list of (np_s x dim) int arrays of scene-local indices;
        nverts: vertex count of each scene; options: one Options or a list."""
        n = len(prims)
        voff = np.zeros(n + 1, np.int32)
        voff[1:] = np.cumsum(nverts)
        poff = np.zeros(n + 1, np.int32)
        poff[1:] = np.cumsum([len(p) for p in prims])
        P = np.ascontiguousarray(np.concatenate([np.asarray(p, np.int32).reshape(-1, self.dim) for p in prims]))
        opts = options if isinstance(options, (list, tuple)) else [options]
        arr = (_lib.Options * len(opts))(*[o.to_c() for o in opts])
        self._check(lib.mclg_batch_set_scenes(self._h, n, _ptr(voff), _ptr(poff), _ptr(P), arr, len(opts)))
        self.nscenes, self.voff, self.poff = n, voff, poff

    def set_active(self, active):
        """BVHTree::update's active mask for every scene: one int per
        concatenated vertex (scenes without a mask: all ones), or None for no
        mask.  Kept for later frames (mclg_batch_set_active)."""
        if active is None:
            self._check(lib.mclg_batch_set_active(self._h, None, 0))
            return
        dev = _is_device(active)
        if not dev:
            active = np.ascontiguousarray(active, np.int32)
        if len(active) != int(self.voff[-1]):
            raise ValueError(f"active needs {int(self.voff[-1])} entries, got {len(active)}")
        self._check(lib.mclg_batch_set_active(self._h, _ptr(active), int(dev)))
        self._active = active  # a device mask must outlive the copy

    def run(self, V0, V1):
        """One frame on concatenated vertices (numpy, or CUDA float64 tensors)."""
        dev = _is_device(V0)
        if not dev:
            V0 = np.ascontiguousarray(V0, np.float64)
            V1 = np.ascontiguousarray(V1, np.float64)
        self._check(lib.mclg_batch_run(self._h, _ptr(V0), _ptr(V1), int(dev)))

    def last_ms(self) -> float:
        return float(lib.mclg_batch_last_ms(self._h))

    def counts(self) -> dict:
        c = _lib.Counts()
        self._check(lib.mclg_batch_get_counts(self._h, C.byref(c)))
        return {f: int(getattr(c, f)) for f, _ in c._fields_}

    def scene_counts(self):
        h = np.zeros(self.nscenes, np.int64)
        d = np.zeros(self.nscenes, np.int64)
        self._check(lib.mclg_batch_scene_counts(self._h, _ptr(h), _ptr(d)))
        return h, d

    def hits(self):
        n = self.counts()["hits"]
        sc = np.empty(n, np.int32)
        s = np.empty((n, 4), np.int32)
        v = np.empty(n, np.uint8)
        t = np.empty(n, np.float64)
        self._check(lib.mclg_batch_copy_hits(self._h, _ptr(sc), _ptr(s), _ptr(v), _ptr(t), n))
        return sc, s, v, t

    def discrete(self):
        n = self.counts()["discrete"]
        sc = np.empty(n, np.int32)
        p = np.empty((n, 2), np.int32)
        self._check(lib.mclg_batch_copy_discrete(self._h, _ptr(sc), _ptr(p), n))
        return sc, p

    def set_params(self, use_graph=-1, tr_waves=0, blocks_per_cu=0, chunk=0):
        self._check(lib.mclg_batch_set_params(self._h, int(use_graph), int(tr_waves), int(blocks_per_cu), int(chunk)))

    def params(self) -> dict:
        out = np.zeros(4, np.int32)
        self._check(lib.mclg_batch_get_params(self._h, _ptr(out)))
        return dict(zip(("use_graph", "tr_waves", "blocks_per_cu", "chunk"), out.tolist()))

    def autotune(self, cache_path: str | None = None, reps: int = 3):
        self._check(lib.mclg_batch_autotune(self._h, cache_path.encode() if cache_path else None, int(reps)))


def from_scenes(scenes, dim=3, **options):
    """Builds a Batch from a list of (V0, V1, P, theta) scenes; returns
    (batch, V0 concatenated, V1 concatenated).  Keyword options (e.g.
    vf_one_sided=True) apply to every scene."""
    b = Batch(dim)
    opts = []
    for _, _, _, th in scenes:
        o = Options()
        o.box_eta = o.vf_ccd_eta = o.ee_ccd_eta = th
        for k, v in options.items():
            setattr(o, k, v)
        opts.append(o)
    b.set_scenes([s[2] for s in scenes], [len(s[0]) for s in scenes], opts)
    V0 = np.ascontiguousarray(np.concatenate([s[0] for s in scenes]))
    V1 = np.ascontiguousarray(np.concatenate([s[1] for s in scenes]))
    return b, V0, V1
