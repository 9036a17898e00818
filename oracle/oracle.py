"""ctypes wrapper over the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Used by tests/, ``__graft_entry__.smoke()`` (as the checker) and the CPU
baseline leg of bench.py.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


class OracleOptions(C.Structure):
    _fields_ = [("box_eta", C.c_double), ("vf_ccd_eta", C.c_double), ("ee_ccd_eta", C.c_double),
                ("reptri", C.c_int), ("threaded", C.c_int), ("discrete", C.c_int), ("continuous", C.c_int),
                ("vf_one_sided", C.c_int), ("f32", C.c_int)]


def build() -> str:
    """Compile the oracle with its Makefile (g++ -O3, no -march)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P, I, D, LL = C.c_void_p, C.c_int, C.c_double, C.c_longlong
        L.oracle_create.restype = P
        L.oracle_create.argtypes = [I]
        L.oracle_destroy.argtypes = [P]
        L.oracle_set_options.argtypes = [P, C.POINTER(OracleOptions), I]
        L.oracle_get_etas.argtypes = [P, P]
        L.oracle_update.argtypes = [P, P, P, P, I, I, P]
        L.oracle_traverse.argtypes = [P, P, P, P, I, I]
        L.oracle_last_ms.restype = D
        L.oracle_last_ms.argtypes = [P, I]
        L.oracle_count.restype = LL
        L.oracle_count.argtypes = [P, I]
        for f in ("oracle_copy_leaf_pairs", "oracle_copy_discrete", "oracle_copy_rank", "oracle_copy_flags"):
            getattr(L, f).argtypes = [P, P]
        L.oracle_copy_candidates.argtypes = [P, P, P]
        L.oracle_events.restype = LL
        L.oracle_events.argtypes = [P, P, P, P]
        L.oracle_copy_hits.argtypes = [P, P, P, P]
        L.oracle_copy_boxes.argtypes = [P, P, P]
        L.oracle_brute_force_pairs.restype = LL
        L.oracle_brute_force_pairs.argtypes = [P, P, LL]
        L.oracle_accd.argtypes = [I, P, P, D, I, P, P]
        L.oracle_accd_f32.argtypes = [I, P, P, D, I, P, P]
        L.oracle_pair_distance.restype = D
        L.oracle_pair_distance.argtypes = [I, P, I]
        L.oracle_point_on_triangle.argtypes = [P, P, P, P, P]
        L.oracle_edge_to_edge.argtypes = [P, P, P, P, P, P]
        L.oracle_point_on_edge2.argtypes = [P, P, P, P]
        L.oracle_tri_tri.argtypes = [P, P]
        L.oracle_edge_edge2.argtypes = [P, P]
        L.oracle_max_threads.restype = I
        L.oracle_accd_grad.argtypes = [I, P, P, D, I, P, P, P]
        L.oracle_toi_list.argtypes = [I, P, P, P, P, LL, D, D, P, P]
        L.oracle_toi_grad_list.argtypes = [I, P, P, I, P, P, P, LL, D, D, P, P]
        L.oracle_ctcd_vf.argtypes = [I, P, P, D, P]
        L.oracle_rpoly.argtypes = [P, I, P, P]
        L.oracle_find_intervals.argtypes = [P, I, I, P, P, I]
        L.oracle_ctcd_subtest.argtypes = [I, P, P, D, P, I, P]
        L.oracle_wrong_side.argtypes = [I, P, P, D]
        L.oracle_ray_box.argtypes = [P, P, P, P]
        L.oracle_ctcd_list.argtypes = [I, P, P, P, LL, D, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Oracle:
    """CPU restatement of ``mcl::BVHTree<double,dim>`` (update + traverse)."""

    def __init__(self, dim: int):
        self.dim = dim
        self._h = lib().oracle_create(dim)
        if not self._h:
            raise ValueError("dim must be 2 or 3")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.oracle_destroy(h)
            self._h = None

    def set_options(self, box_eta=1e-6, vf_ccd_eta=1e-6, ee_ccd_eta=1e-6, reptri=True, threaded=True,
                    discrete=True, continuous=True, vf_one_sided=False, nthreads=0, f32=False):
        o = OracleOptions(box_eta, vf_ccd_eta, ee_ccd_eta, int(reptri), int(threaded), int(discrete),
                          int(continuous), int(vf_one_sided), int(f32))
        lib().oracle_set_options(self._h, C.byref(o), int(nthreads))

    def etas(self):
        out = np.zeros(3)
        lib().oracle_get_etas(self._h, _p(out))
        return out

    def update(self, V0, V1, P, active=None):
        self._V0, self._V1, self._P = _f64(V0), _f64(V1), _i32(P)
        self._act = _i32(active) if active is not None else None
        npr, pdim = self._P.shape
        lib().oracle_update(self._h, _p(self._V0), _p(self._V1), _p(self._P), npr, pdim, _p(self._act))

    def traverse(self, V0=None, V1=None, P=None, leaf_pairs=True, candidates=True):
        V0 = self._V0 if V0 is None else _f64(V0)
        V1 = self._V1 if V1 is None else _f64(V1)
        P = self._P if P is None else _i32(P)
        lib().oracle_traverse(self._h, _p(V0), _p(V1), _p(P), int(leaf_pairs), int(candidates))

    def last_ms(self):
        return lib().oracle_last_ms(self._h, 0), lib().oracle_last_ms(self._h, 1)

    def count(self, which: int) -> int:
        return int(lib().oracle_count(self._h, which))

    def leaf_pairs(self):
        out = np.zeros((self.count(0), 2), np.int32)
        lib().oracle_copy_leaf_pairs(self._h, _p(out))
        return out

    def candidates(self):
        n = self.count(1)
        s = np.zeros((n, 4), np.int32)
        v = np.zeros(n, np.uint8)
        lib().oracle_copy_candidates(self._h, _p(s), _p(v))
        return s, v

    def hits(self):
        n = self.count(2)
        s = np.zeros((n, 4), np.int32)
        v = np.zeros(n, np.uint8)
        t = np.zeros(n, np.float64)
        lib().oracle_copy_hits(self._h, _p(s), _p(v), _p(t))
        return s, v, t

    def discrete(self):
        out = np.zeros((self.count(3), 2), np.int32)
        lib().oracle_copy_discrete(self._h, _p(out))
        return out

    def events(self):
        """(pair, kind, idx) of a single-threaded traverse (threaded=False): the
        records in the order collide produced them — kind 0 candidate, 1 hit, 2
        discrete, idx into candidates() / hits() / discrete(), pair the ordinal of
        the leaf pair reached."""
        n = int(lib().oracle_events(self._h, None, None, None))
        if n < 0:
            raise RuntimeError("oracle events need a single-threaded traverse (threaded=False)")
        pair = np.zeros(n, np.int64)
        kind = np.zeros(n, np.int32)
        idx = np.zeros(n, np.int64)
        lib().oracle_events(self._h, _p(pair), _p(kind), _p(idx))
        return pair, kind, idx

    def rank(self):
        out = np.zeros(self._P.shape[0], np.int32)
        lib().oracle_copy_rank(self._h, _p(out))
        return out

    def flags(self):
        out = np.zeros(self._P.shape[0], np.uint8)
        lib().oracle_copy_flags(self._h, _p(out))
        return out

    def boxes(self):
        n = self._P.shape[0]
        out = np.zeros((n, 6 * self.dim), np.float64)
        act = np.zeros(n, np.uint8)
        lib().oracle_copy_boxes(self._h, _p(out), _p(act))
        return out, act

    def brute_force_pairs(self):
        n = int(lib().oracle_brute_force_pairs(self._h, None, 0))
        out = np.zeros((n, 2), np.int32)
        lib().oracle_brute_force_pairs(self._h, _p(out), n)
        return out


def accd(dim, verts0, verts1, eta, is_vf=True):
    """Returns (hit, toi, distance evaluations) of the literal ACCD restatement."""
    v0, v1 = _f64(verts0), _f64(verts1)
    toi = C.c_double(-1.0)
    ev = C.c_longlong(0)
    h = lib().oracle_accd(dim, _p(v0), _p(v1), float(eta), int(is_vf), C.byref(toi), C.byref(ev))
    return bool(h), toi.value, ev.value


def accd_f32(dim, verts0, verts1, eta, is_vf=True):
    """(hit, toi, distance evaluations) of ACCD at T = float on float(verts)
    (NarrowPhaseACCD<float,DIM>)."""
    v0, v1 = _f64(verts0), _f64(verts1)
    toi = C.c_double(-1.0)
    ev = C.c_longlong(0)
    h = lib().oracle_accd_f32(dim, _p(v0), _p(v1), float(eta), int(is_vf), C.byref(toi), C.byref(ev))
    return bool(h), toi.value, ev.value


def pair_distance(dim, verts, is_vf=True):
    v = _f64(verts)
    return lib().oracle_pair_distance(dim, _p(v), int(is_vf))


def point_on_triangle(q, a, b, c):
    q, a, b, c = (_f64(x) for x in (q, a, b, c))
    out = np.zeros(3)
    lib().oracle_point_on_triangle(_p(q), _p(a), _p(b), _p(c), _p(out))
    return out


def edge_to_edge(p0, p1, q0, q1):
    p0, p1, q0, q1 = (_f64(x) for x in (p0, p1, q0, q1))
    out = np.zeros(3)
    bary = np.zeros(4)
    lib().oracle_edge_to_edge(_p(p0), _p(p1), _p(q0), _p(q1), _p(out), _p(bary))
    return out, bary


def point_on_edge2(p, e0, e1):
    p, e0, e1 = (_f64(x) for x in (p, e0, e1))
    out = np.zeros(2)
    lib().oracle_point_on_edge2(_p(p), _p(e0), _p(e1), _p(out))
    return out


def tri_tri(f0, f1):
    a, b = _f64(f0).reshape(9), _f64(f1).reshape(9)
    return bool(lib().oracle_tri_tri(_p(a), _p(b)))


def edge_edge2(e0, e1):
    a, b = _f64(e0).reshape(4), _f64(e1).reshape(4)
    return bool(lib().oracle_edge_edge2(_p(a), _p(b)))


def accd_grad(dim, verts0, verts1, eta, is_vf=True):
    """(hit, toi, d toi/d verts0, d toi/d verts1) by dual numbers."""
    v0, v1 = _f64(verts0), _f64(verts1)
    g0 = np.zeros_like(v0)
    g1 = np.zeros_like(v1)
    toi = C.c_double(-1.0)
    h = lib().oracle_accd_grad(dim, _p(v0), _p(v1), float(eta), int(is_vf), _p(g0), _p(g1), C.byref(toi))
    return bool(h), toi.value, g0, g1


def toi_list(V0, V1, sten, is_vf, eta_vf, eta_ee, signature=False):
    """ACCD TOI per stencil (-1: no hit); with ``signature`` also a uint64 per
    stencil hashing every branch taken (equal signatures: same smooth formula)."""
    V0, V1 = _f64(V0), _f64(V1)
    s = _i32(sten).reshape(-1, 4)
    v = np.ascontiguousarray(is_vf, np.uint8)
    out = np.zeros(len(s))
    sig = np.zeros(len(s), np.uint64) if signature else None
    lib().oracle_toi_list(V0.shape[1], _p(V0), _p(V1), _p(s), _p(v), len(s), float(eta_vf), float(eta_ee), _p(out),
                          _p(sig))
    return (out, sig) if signature else out


def toi_grad_list(V0, V1, sten, is_vf, w, eta_vf, eta_ee):
    """sum_i w_i d toi_i / d(V0, V1) over a stencil list (dual numbers)."""
    V0, V1 = _f64(V0), _f64(V1)
    s = _i32(sten).reshape(-1, 4)
    v = np.ascontiguousarray(is_vf, np.uint8)
    w = _f64(w)
    g0 = np.zeros_like(V0)
    g1 = np.zeros_like(V1)
    lib().oracle_toi_grad_list(V0.shape[1], _p(V0), _p(V1), len(V0), _p(s), _p(v), _p(w), len(s), float(eta_vf),
                               float(eta_ee), _p(g0), _p(g1))
    return g0, g1


def ctcd_vf(dim, verts0, verts1, eta):
    """CTCD vertex-face query with the wrong-side test: (code, toi); code 1 hit,
    0 no hit, -1 only wrong-side hits (toi -1 unless code is 1)."""
    v0, v1 = _f64(verts0), _f64(verts1)
    toi = C.c_double(-1.0)
    r = lib().oracle_ctcd_vf(dim, _p(v0), _p(v1), float(eta), C.byref(toi))
    return r, toi.value


def rpoly(coeffs):
    """Jenkins-Traub zeros of a polynomial (decreasing powers): (count, real, imag)."""
    op = _f64(coeffs)
    deg = len(op) - 1
    zr, zi = np.zeros(max(deg, 1)), np.zeros(max(deg, 1))
    n = lib().oracle_rpoly(_p(op), deg, _p(zr), _p(zi))
    return n, zr, zi


def find_intervals(coeffs, pos):
    """CTCD's sign intervals of a polynomial over [0, 1] as a list of (l, u)."""
    op = _f64(coeffs)
    lo, hi = np.zeros(8), np.zeros(8)
    n = lib().oracle_find_intervals(_p(op), len(op) - 1, int(pos), _p(lo), _p(hi), 8)
    return [(lo[i], hi[i]) for i in range(n)]


def ctcd_subtest(kind, x0, x1, eta):
    """One 3D sub-test: kind 0 vertex-face, 1 vertex-edge, 2 vertex-vertex.  Returns (hit, times)."""
    a, b = _f64(x0).reshape(-1), _f64(x1).reshape(-1)
    out = np.zeros(512)
    nt = C.c_int(0)
    h = lib().oracle_ctcd_subtest(int(kind), _p(a), _p(b), float(eta), _p(out), 512, C.byref(nt))
    return bool(h), out[:min(nt.value, 512)].copy()


def wrong_side(dim, verts0, verts1, t):
    v0, v1 = _f64(verts0), _f64(verts1)
    return bool(lib().oracle_wrong_side(dim, _p(v0), _p(v1), float(t)))


def ray_box(p0, p1, bmin, bmax):
    p0, p1, bmin, bmax = (_f64(x) for x in (p0, p1, bmin, bmax))
    return bool(lib().oracle_ray_box(_p(p0), _p(p1), _p(bmin), _p(bmax)))


def ctcd_list(V0, V1, sten, eta):
    """CTCD TOI per vertex-face stencil (-1: no hit)."""
    V0, V1 = _f64(V0), _f64(V1)
    s = _i32(sten).reshape(-1, 4)
    out = np.zeros(len(s))
    lib().oracle_ctcd_list(V0.shape[1], _p(V0), _p(V1), _p(s), len(s), float(eta), _p(out))
    return out


def run_scene(V0, V1, P, theta, dim=None, leaf_pairs=True, candidates=True, nthreads=0, **opts):
    """update + traverse with box_eta = vf = ee = theta; returns the Oracle."""
    dim = dim or V0.shape[1]
    o = Oracle(dim)
    o.set_options(box_eta=theta, vf_ccd_eta=theta, ee_ccd_eta=theta, nthreads=nthreads, **opts)
    o.update(V0, V1, P)
    o.traverse(leaf_pairs=leaf_pairs, candidates=candidates)
    return o
