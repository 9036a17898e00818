"""Python mirror of ``mcl::BVHTree<double,DIM>`` over the gfx950 library.

Same names and meaning as the reference (include/MCL/BVHTree.hpp:30-148):
``options`` (box_eta, vf_ccd_eta, ee_ccd_eta, reptri, threaded, discrete,
continuous, vf_one_sided, ee_robust), ``update(V0, V1, P, active)``,
``traverse(V0, V1, P)`` and the four callbacks ``append_pair(sten, is_vf, toi)``,
``append_discrete(p0, p1) -> bool``, ``filter_pair(sten, is_vf) -> bool`` and
``narrow_phase(sten, is_vf) -> toi``.

The GPU computes every result of a traverse; the callbacks are then replayed
on the calling thread, serially, leaf pair by leaf pair in the order of
mclg_copy_events (with options.threaded False, or the KdBVH orientation, the
reference's own serial order, BVHTree.cpp:152-157; its threaded run calls them
concurrently from TBB workers, so any serial order is one it can produce).
Per leaf pair, as collide does (BVHTree.cpp:252-283):
  1. continuous: filter_pair (if set) on each of the pair's candidates in
     get_candidates order, then for each one kept the toi —
     narrow_phase(sten, is_vf) when that callback is set, else the GPU's ACCD
     result — and append_pair when toi >= 0.  Without filter_pair and
     narrow_phase only the hits are read back.
  2. discrete: append_discrete(p0, p1) for the pair's discrete hit; a True
     return stops the replay, so no later pair delivers anything (the
     reference's stop flag, BVHTree.cpp:236-239,280-281).

``traverse(visitor)`` (BVHTree.hpp:96) walks a host copy of the tree with a
``BVHTraverse``-like object: ``intersect_volume(box) -> bool`` (True: go down)
and ``intersect_object(leaf) -> bool`` (True: stop), see ``visit``.

Inputs may be numpy arrays (copied to the GPU) or CUDA torch tensors (used in
place: float64 vertices, int32 indices, contiguous).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import lib


class Options:
    """BVHTree::Options (BVHTree.hpp:42-64) with the reference defaults."""

    __slots__ = ("box_eta", "vf_ccd_eta", "ee_ccd_eta", "reptri", "threaded", "discrete", "continuous",
                 "vf_one_sided", "ee_robust")

    def __init__(self):
        self.box_eta = 1e-6
        self.vf_ccd_eta = 1e-6
        self.ee_ccd_eta = 1e-6
        self.reptri = True
        self.threaded = True
        self.discrete = True
        self.continuous = True
        self.vf_one_sided = False
        self.ee_robust = True

    def to_c(self) -> _lib.Options:
        return _lib.Options(self.box_eta, self.vf_ccd_eta, self.ee_ccd_eta, int(self.reptri), int(self.threaded),
                            int(self.discrete), int(self.continuous), int(self.vf_one_sided), int(self.ee_robust))

    def from_c(self, o: _lib.Options):
        for f in self.__slots__:
            v = getattr(o, f)
            setattr(self, f, bool(v) if isinstance(getattr(self, f), bool) else v)


def _is_device(x) -> bool:
    return hasattr(x, "data_ptr") and getattr(x, "is_cuda", False)


def _ptr(a):
    if a is None:
        return None
    if _is_device(a):
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


def _host_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _host_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _dev_check(t, dtype_name):
    if not t.is_contiguous():
        raise ValueError("device inputs must be contiguous")
    if str(t.dtype) != dtype_name:
        raise ValueError(f"device input must be {dtype_name}, got {t.dtype}")
    return t


class BVHTree:
    """``mcl::BVHTree<T, dim>`` on one GPU; T = double (default) or, with
    ``dtype=numpy.float32``, float: coordinates and etas rounded to float,
    boxes padded in float and ACCD in float arithmetic (mclg_set_precision).
    Inputs are float64 either way (float32 host arrays are widened exactly)."""

    def __init__(self, dim: int = 3, dtype=np.float64):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        dt = np.dtype(dtype)
        if dt not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise ValueError("dtype must be float64 or float32")
        self.dim = dim
        self.dtype = dt
        self.options = Options()
        self.append_pair = None
        self.append_discrete = None
        self.filter_pair = None
        self.narrow_phase = None
        h = C.c_void_p()
        rc = lib.mclg_create(dim, C.byref(h))
        if rc != _lib.OK:
            raise _lib.MclgError(rc, lib.mclg_last_error(None).decode())
        self._h = h
        if dt == np.dtype(np.float32):
            self._check(lib.mclg_set_precision(h, _lib.PRECISION_F32))
        self.nv = 0
        self.np = 0
        self.serial = 0  # counts updates and traversals (the adjoint refers to the last traverse)
        self._mirror = None  # host copy of the tree for visitor walks (per update)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.mclg_destroy(h)
            self._h = None

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc):
        if rc != _lib.OK:
            raise _lib.MclgError(rc, lib.mclg_last_error(self._h).decode())

    def _push_options(self):
        o = self.options.to_c()
        self._check(lib.mclg_set_options(self._h, C.byref(o)))

    def _pull_options(self):
        o = _lib.Options()
        self._check(lib.mclg_get_options(self._h, C.byref(o)))
        self.options.from_c(o)

    # -- reference API ------------------------------------------------------
    def update(self, V0, V1, P, active=None, refit=False):
        """BVHTree::update (BVHTree.hpp:73-82): builds the BVH of P's primitives
        over the swept boxes of V0 -> V1.  Mismatched row counts return silently
        as in the reference (BVHTree.hpp:160).

        refit=True (no reference counterpart; mclg_update_refit): keeps the last
        full update's leaf order and hierarchy and recomputes only the boxes;
        same P, vertex and primitive counts required.  Results equal a rebuild's;
        the kept hierarchy only changes how fast traverse runs."""
        if len(V0) != len(V1):
            return
        self._push_options()
        dev = _is_device(V0)
        if dev:
            V0, V1 = _dev_check(V0, "torch.float64"), _dev_check(V1, "torch.float64")
            P = _dev_check(P, "torch.int32")
            if active is not None:
                active = _dev_check(active, "torch.int32")
                if len(active) != len(V0):
                    active = None
            nv, npr, pdim = int(V0.shape[0]), int(P.shape[0]), int(P.shape[1]) if P.ndim == 2 else 1
            fn = lib.mclg_update_refit_device if refit else lib.mclg_update_device
            rc = fn(self._h, _ptr(V0), _ptr(V1), nv, _ptr(P), npr, pdim, _ptr(active))
        else:
            V0, V1, P = _host_f64(V0), _host_f64(V1), _host_i32(P)
            if P.ndim == 1:
                P = P.reshape(-1, 1)
            if active is not None:
                active = _host_i32(active)
                if len(active) != len(V0):
                    active = None
            nv, npr, pdim = V0.shape[0], P.shape[0], P.shape[1]
            fn = lib.mclg_update_refit if refit else lib.mclg_update
            rc = fn(self._h, _ptr(V0), _ptr(V1), nv, _ptr(P), npr, pdim, _ptr(active))
        self.serial += 1  # the last traverse's hits no longer match the vertex records
        self._check(rc)
        self._pull_options()
        self.nv, self.np = nv, npr
        self._mirror = None

    def traverse(self, V0, V1=None, P=None):
        """BVHTree::traverse (BVHTree.hpp:86-93) with serial callback replay;
        traverse(visitor) is BVHTree::traverse(BVHTraverse*), see ``visit``."""
        if V1 is None and P is None and hasattr(V0, "intersect_volume"):
            return self.visit(V0)
        if len(V0) != len(V1):
            return
        need_cand = self.options.continuous and (self.narrow_phase is not None or self.filter_pair is not None)
        cont = self.options.continuous and (self.append_pair is not None or need_cand)
        disc = self.options.discrete and self.append_discrete is not None
        self.run(V0, V1, P, keep_candidates=need_cand, keep_leaf_pairs=cont or disc)
        if cont or disc:
            self._replay(need_cand, cont, disc)

    def _replay(self, need_cand, cont, disc):
        ev, grp = self.events(over_candidates=need_cand)
        if cont:
            sten, is_vf, toi = self.candidates() if need_cand else self.hits()
        dp = self.discrete() if disc else None
        k, n = 0, len(ev)
        while k < n:  # one leaf pair per iteration: its records, then its discrete hit
            j = k
            while j < n and grp[j] == grp[k]:
                j += 1
            recs = [int(e) for e in ev[k:j] if e >= 0]
            dis = [int(-1 - e) for e in ev[k:j] if e < 0]
            k = j
            if cont:
                items = [(tuple(int(x) for x in sten[i]), bool(is_vf[i]), i) for i in recs]
                if self.filter_pair is not None:  # get_candidates filters all of the pair's stencils first
                    items = [it for it in items if not self.filter_pair(it[0], it[1])]
                for s, v, i in items:
                    t = self.narrow_phase(s, v) if self.narrow_phase is not None else float(toi[i])
                    if t >= 0 and self.append_pair is not None:
                        self.append_pair(s, v, t)
            if disc:
                for d in dis:
                    if self.append_discrete(int(dp[d, 0]), int(dp[d, 1])):
                        return  # stop: every later collide returns at entry

    def events(self, over_candidates=False):
        """(events, group) of the last traverse in replay order (mclg_copy_events):
        events >= 0 index hits() (or candidates()), events < 0 are discrete
        hit -1-e; group numbers the leaf pair of each event."""
        n = C.c_int64(0)
        self._check(lib.mclg_copy_events(self._h, int(over_candidates), None, None, 0, C.byref(n)))
        ev = np.empty(n.value, np.int64)
        grp = np.empty(n.value, np.int64)
        self._check(lib.mclg_copy_events(self._h, int(over_candidates), _ptr(ev), _ptr(grp), n.value, C.byref(n)))
        return ev, grp

    # -- bulk access (no callbacks) --------------------------------------------
    def run(self, V0, V1, P, keep_candidates=False, keep_leaf_pairs=False, check_order=False):
        """GPU traverse only; results stay on the device (see hits(), discrete()).
        check_order: also list the pairs whose discrete result depends on the
        argument order (order_dependent())."""
        self._push_options()
        self.serial += 1
        flags = (_lib.KEEP_CANDIDATES if keep_candidates else 0) | (_lib.KEEP_LEAF_PAIRS if keep_leaf_pairs else 0)
        flags |= _lib.CHECK_ORDER if check_order else 0
        if _is_device(V0):
            V0, V1 = _dev_check(V0, "torch.float64"), _dev_check(V1, "torch.float64")
            P = _dev_check(P, "torch.int32")
            rc = lib.mclg_traverse_device(self._h, _ptr(V0), _ptr(V1), _ptr(P), flags)
        else:
            V0, V1, P = _host_f64(V0), _host_f64(V1), _host_i32(P)
            rc = lib.mclg_traverse(self._h, _ptr(V0), _ptr(V1), _ptr(P), flags)
        self._check(rc)

    def counts(self) -> dict:
        c = _lib.Counts()
        self._check(lib.mclg_get_counts(self._h, C.byref(c)))
        return {f: int(getattr(c, f)) for f, _ in c._fields_}

    def hits(self):
        n = self.counts()["hits"]
        s = np.empty((n, 4), np.int32)
        v = np.empty(n, np.uint8)
        t = np.empty(n, np.float64)
        self._check(lib.mclg_copy_hits(self._h, _ptr(s), _ptr(v), _ptr(t), n))
        return s, v, t

    def device_hits(self):
        """(sten [n,4] int32, is_vf [n] uint8, toi [n] f64) as torch views of the
        library's device buffers (no copy); valid until the next traverse."""
        import torch

        n = self.counts()["hits"]
        r = _lib.DeviceResults()
        self._check(lib.mclg_get_device_results(self._h, C.byref(r)))
        dev = torch.device("cuda", torch.cuda.current_device())
        if n == 0:
            return (torch.zeros((0, 4), dtype=torch.int32, device=dev), torch.zeros(0, dtype=torch.uint8, device=dev),
                    torch.zeros(0, dtype=torch.float64, device=dev))

        def view(ptr, shape, typestr):
            class _A:
                __cuda_array_interface__ = {"shape": shape, "typestr": typestr, "data": (int(ptr), False),
                                            "version": 3}
            return torch.as_tensor(_A(), device=dev)

        return view(r.hit_sten4, (n, 4), "<i4"), view(r.hit_is_vf, (n,), "|u1"), view(r.hit_toi, (n,), "<f8")

    def device_candidates(self):
        """The candidate queues of the last traverse as torch views of the
        library's device buffers (no copy; valid until the next traverse):
        (vf_sten [n_vf,4] int32, vf_toi [n_vf] f64, ee_sten [n_ee,4] int32,
        ee_toi [n_ee] f64); toi -1 where the narrow phase found no contact.
        Needs a traverse run with keep_candidates."""
        import torch

        r = _lib.DeviceResults()
        self._check(lib.mclg_get_device_results(self._h, C.byref(r)))
        if (r.n_vf and not r.vf_toi) or (r.n_ee and not r.ee_toi):
            raise RuntimeError("device_candidates: the last traverse did not keep candidates (keep_candidates=True)")
        dev = torch.device("cuda", torch.cuda.current_device())

        def view(ptr, shape, typestr, dtype):
            if shape[0] == 0 or not ptr:
                return torch.zeros(shape, dtype=dtype, device=dev)

            class _A:
                __cuda_array_interface__ = {"shape": shape, "typestr": typestr, "data": (int(ptr), False),
                                            "version": 3}
            return torch.as_tensor(_A(), device=dev)

        nvf, nee = int(r.n_vf), int(r.n_ee)
        return (view(r.vf_sten4, (nvf, 4), "<i4", torch.int32), view(r.vf_toi, (nvf,), "<f8", torch.float64),
                view(r.ee_sten4, (nee, 4), "<i4", torch.int32), view(r.ee_toi, (nee,), "<f8", torch.float64))

    def discrete(self):
        n = self.counts()["discrete"]
        out = np.empty((n, 2), np.int32)
        self._check(lib.mclg_copy_discrete(self._h, _ptr(out), n))
        return out

    def candidates(self):
        n = self.counts()["candidates"]
        s = np.empty((n, 4), np.int32)
        v = np.empty(n, np.uint8)
        t = np.empty(n, np.float64)
        self._check(lib.mclg_copy_candidates(self._h, _ptr(s), _ptr(v), _ptr(t), n))
        return s, v, t

    def order_dependent(self):
        """Leaf pairs of the last traverse (run with check_order) whose discrete
        result changes when the two primitives are swapped."""
        n = self.counts()["order_dependent"]
        out = np.empty((n, 2), np.int32)
        self._check(lib.mclg_copy_order_dependent(self._h, _ptr(out), n))
        return out

    def leaf_pairs(self):
        n = self.counts()["leaf_pairs"]
        out = np.empty((n, 2), np.int32)
        self._check(lib.mclg_copy_leaf_pairs(self._h, _ptr(out), n))
        return out

    def leaves(self):
        """(boxes [np, 6*dim], active [np], reptri flags [np]) of the last update."""
        n = self.np
        b = np.empty((n, 6 * self.dim), np.float64)
        a = np.empty(n, np.uint8)
        f = np.empty(n, np.uint8)
        self._check(lib.mclg_copy_leaves(self._h, _ptr(b), _ptr(a), _ptr(f), n))
        return b, a, f

    def tree(self):
        """The last update's tree: (lchild, rchild, order, root).  Internal node k
        (< np-1) has children lchild[k], rchild[k]; node np-1+s is the leaf at
        sorted position s, primitive order[s]."""
        n = self.np
        lc = np.empty(max(n - 1, 0), np.int32)
        rc = np.empty(max(n - 1, 0), np.int32)
        order = np.empty(n, np.int32)
        root = C.c_int(-1)
        self._check(lib.mclg_copy_tree(self._h, _ptr(lc), _ptr(rc), _ptr(order), C.byref(root), n))
        return lc, rc, order, root.value

    def visit(self, visitor):
        """BVHTree::traverse(BVHTraverse*) (BVHTree.cpp:162-166, the KdBVH walk of
        Eigen::BVIntersect): from the root, each visited node's child volumes go
        to ``visitor.intersect_volume(box)`` (True queues the child), then its
        child objects to ``visitor.intersect_object(leaf)`` (True ends the walk);
        queued nodes are taken last in, first out.  ``box`` is a ``Box``: swept
        ``lo``/``hi``, ``t0``/``t1`` as (lo, hi) and ``active``; an internal
        volume merges its children's boxes (BVHLeaf.cpp:9-23).  ``leaf`` is a
        ``Leaf`` (idx, box).  The tree is the GPU's LBVH, so the volumes and the
        visiting order differ from the reference's KdBVH; the objects reached by
        a query that never stops early do not."""
        if self.np == 0:
            return
        if self._mirror is None:
            self._mirror = _HostTree(self)
        self._mirror.walk(visitor)

    def set_orientation(self, mode: str):
        """'index' (default: lower primitive index first) or 'kdbvh' (the
        reference's KdBVH orientation, rebuilt on the host at every update;
        mclg_set_orientation).  Takes effect at the next update."""
        m = {"index": _lib.ORIENT_INDEX, "kdbvh": _lib.ORIENT_KDBVH}[mode]
        self._check(lib.mclg_set_orientation(self._h, m))

    def set_rank(self, rank):
        """Orient leaf pairs by rank (the reference's KdBVH in-order position)."""
        if rank is None:
            self._check(lib.mclg_set_rank(self._h, None, 0))
            return
        r = _host_i32(rank)
        self._check(lib.mclg_set_rank(self._h, _ptr(r), len(r)))

    def timings(self) -> dict:
        ms = np.zeros(_lib.NUM_TIMINGS)
        self._check(lib.mclg_get_timings(self._h, _ptr(ms), _lib.NUM_TIMINGS))
        return dict(zip(_lib.TIMING_NAMES, ms.tolist()))

    def set_stream(self, stream_ptr: int | None):
        self._check(lib.mclg_set_stream(self._h, C.c_void_p(stream_ptr) if stream_ptr else None))


class Box:
    """BoundingBox (BVHLeaf.hpp:12-29): swept box lo/hi, t0 and t1 as (lo, hi), active."""

    __slots__ = ("lo", "hi", "t0", "t1", "active")

    def __init__(self, lo, hi, t0, t1, active):
        self.lo, self.hi, self.t0, self.t1, self.active = lo, hi, t0, t1, active

    def contains(self, p) -> bool:
        p = np.asarray(p, np.float64)
        return bool(np.all(self.lo <= p) and np.all(p <= self.hi))


class Leaf:
    """BVHLeaf (BVHLeaf.hpp:31-52): primitive index and its box."""

    __slots__ = ("idx", "box")

    def __init__(self, idx, box):
        self.idx, self.box = idx, box


class _HostTree:
    """Host copy of a BVHTree's last update for visitor walks: leaf boxes and the
    internal volumes, merged children first (BVHLeaf.cpp:9-23: union of the swept,
    t0 and t1 boxes, active if either is)."""

    def __init__(self, t):
        b, a, _ = t.leaves()
        lc, rc, order, root = t.tree()
        d, n = t.dim, t.np
        self.n, self.root, self.lc, self.rc, self.order = n, root, lc, rc, order
        # boxes of nodes in node-id order: internal 0..n-2, then leaves by sorted position
        lo = np.empty((2 * n - 1, 3, d))  # [swept, t0, t1] lower bounds
        hi = np.empty((2 * n - 1, 3, d))
        act = np.empty(2 * n - 1, bool)
        lb = b[order].reshape(n, 6, d)
        lo[n - 1:], hi[n - 1:] = lb[:, 0::2], lb[:, 1::2]
        act[n - 1:] = a[order] != 0
        if n > 1:
            post = []  # post-order of the internal nodes
            st = [(root, False)]
            while st:
                p, done = st.pop()
                if done:
                    post.append(p)
                    continue
                st.append((p, True))
                for c in (lc[p], rc[p]):
                    if c < n - 1:
                        st.append((c, False))
            for p in post:  # the reference's extend: lo = (b < a) ? b : a, hi = (a < b) ? b : a
                l, r = lc[p], rc[p]
                lo[p] = np.where(lo[r] < lo[l], lo[r], lo[l])
                hi[p] = np.where(hi[l] < hi[r], hi[r], hi[l])
                act[p] = act[l] or act[r]
        self.lo, self.hi, self.act = lo, hi, act

    def _box(self, i):
        lo, hi = self.lo[i], self.hi[i]
        return Box(lo[0], hi[0], (lo[1], hi[1]), (lo[2], hi[2]), bool(self.act[i]))

    def walk(self, visitor):
        n = self.n
        if n == 1:
            visitor.intersect_object(Leaf(int(self.order[0]), self._box(0)))
            return
        todo = [self.root]
        while todo:
            p = todo.pop()
            ch = (int(self.lc[p]), int(self.rc[p]))
            for c in ch:
                if c < n - 1 and visitor.intersect_volume(self._box(c)):
                    todo.append(c)
            for c in ch:
                if c >= n - 1 and visitor.intersect_object(Leaf(int(self.order[c - (n - 1)]), self._box(c))):
                    return


def kdbvh_order(centres):
    """The reference KdBVH's leaf order for swept-box centres (n x dim), computed
    on the host by the library (mclg_kdbvh_order; no GPU needed): (rank, gapnode)
    — rank[i] the in-order position of leaf i, gapnode[g] the frontlist index of
    the node that splits positions g and g+1."""
    c = np.ascontiguousarray(centres, np.float64)
    n, dim = c.shape
    rank = np.empty(n, np.int32)
    gap = np.empty(max(n - 1, 0), np.int32)
    rc = lib.mclg_kdbvh_order(dim, _ptr(c), n, _ptr(rank), _ptr(gap) if n > 1 else None)
    if rc != _lib.OK:
        raise _lib.MclgError(rc, "mclg_kdbvh_order: bad arguments")
    return rank, gap
