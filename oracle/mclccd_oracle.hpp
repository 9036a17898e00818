// CPU oracle for the mclccd collision path — TEST INFRASTRUCTURE ONLY.
//
// This header restates, in plain C++17 with no Eigen and no TBB, the algorithm
// of the reference's `mcl::BVHTree<double,DIM>` update/traverse path so that
// the HIP implementation can be checked against it.  Only tests, the smoke
// check and the CPU-baseline leg of the benchmark may link it; the product
// library never does.
//
// What it restates (reference paths, MIT licence, (c) Matt Overby 2021; the
// triangle-triangle predicate is Guigue & Devillers 2003 via Eric Haines'
// jgt-code, MIT):
//   update            include/MCL/BVHTree.cpp:27-126
//   traverse/collide  include/MCL/BVHTree.cpp:129-159,168-302
//   get_candidates    include/MCL/BVHTree.cpp:304-398
//   boxes_intersect   include/MCL/BVHTree.cpp:400-426
//   narrow phase      include/MCL/BVHTree.cpp:428-500
//   box union         include/MCL/BVHLeaf.cpp:9-23
//   KdBVH build       include/MCL/ccd_internal/KdBVH.hpp:103-222
//   ACCD              include/MCL/NarrowPhase.cpp:608-750
//   projections       include/MCL/ccd_internal/Projection.hpp:60-145,193-204,241-302
//   tri-tri           include/MCL/ccd_internal/tt_isect.hpp:71-270,499-585
//   2D edge-edge      include/MCL/NarrowPhase.cpp:229-257
//   CTCD (vf_one_sided)  ctcd_oracle.hpp
//
// Arithmetic follows the reference's evaluation order as Eigen 3.4 evaluates
// it for fixed-size vectors on x86-64 without FMA: dot(a,b) = (a0*b0 + a1*b1)
// + a2*b2, norm = sqrt(dot), element-wise expressions left to right.  Build
// with -O3 and without -march / -ffast-math so that no contraction happens.
//
// Parity pins (what checks each part, since the reference ships no tests):
//   leaf boxes, reptri          closed forms / brute force   (tests/test_oracle_*.py)
//   leaf-pair set               O(n^2) brute force with the exact leaf test
//   candidates                  invariants on the brute-force pair set
//   ACCD                        closed forms (vertex falling on a face, crossing
//                               edges), static => no hit, receding => no hit,
//                               conservativeness along the path
//   projections                 dense-sampling minimum distance
//   tri-tri, 2D edge-edge       exact rational predicates (Python fractions)
//   KdBVH partition             median-split property per node (tests/test_kdbvh_order.py)
//   float tree (Options::f32)   ACCD vs a numpy float32 restatement, boxes vs
//                               numpy float32 (tests/test_oracle_float.py)
//   Eigen 3.4 reduction order   parity unpinned: no reference test fixes it
#pragma once

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <set>
#include <string>
#include <type_traits>
#include <unordered_set>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "ctcd_oracle.hpp"

namespace mclo {

// ---------------------------------------------------------------------------
// Small vector helpers in the Eigen 3.4 fixed-size evaluation order.
// ---------------------------------------------------------------------------
// The ACCD path below is written over a scalar type T: double for the checker
// itself, a forward-mode dual number (oracle/dual.hpp) for the reference TOI
// gradient.  With T = double every expression is the plain FP64 one.
inline double osqrt(double x) { return std::sqrt(x); }
inline float osqrt(float x) { return std::sqrt(x); }

template <int D, typename T = double> inline T vdot(const T* a, const T* b) {
  if constexpr (D == 3) return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
  else return a[0] * b[0] + a[1] * b[1];
}
template <int D, typename T = double> inline T vnorm(const T* a) { return osqrt(vdot<D, T>(a, a)); }

// std::max / std::min as libstdc++ defines them: max(a,b) = a<b ? b : a,
// min(a,b) = b<a ? b : a.  The NaN and signed-zero behaviour matters.
template <typename T> inline T smax(const T& a, const T& b) { return (a < b) ? b : a; }
// Branch signature (test instrumentation for finite-difference checks): while
// branch_sig points to a word on this thread, the distance and ACCD routines
// fold every branch they take (closest-point region, clamp state of each
// ratio, l_p argmax, step count) into it.  Two inputs with equal signatures
// evaluate the same smooth formula.
inline thread_local uint64_t* branch_sig = nullptr;
inline void note_branch(uint64_t code) {
  if (branch_sig) *branch_sig = (*branch_sig ^ (code + 0x9e3779b97f4a7c15ULL)) * 0x100000001b3ULL;
}
template <typename T> inline int clamp_state(const T& x) { return x < 0.0 ? 0 : (x > 1.0 ? 2 : 1); }

// Held<T>::of(x) is x for plain arithmetic; a derivative-carrying scalar type
// specialises it to drop x's derivatives (x treated as a constant).
template <typename T> struct Held {
  static T of(const T& x) { return x; }
};
template <typename T> inline T smin(const T& a, const T& b) { return (b < a) ? b : a; }

// ---------------------------------------------------------------------------
// Eigen::AlignedBox semantics (Eigen 3.4): empty = [+max, lowest]; extend is
// cwise min/max; intersects is inclusive.
// ---------------------------------------------------------------------------
template <int D> struct ABox {
  double lo[D], hi[D];
  void set_empty() {
    for (int k = 0; k < D; ++k) { lo[k] = DBL_MAX; hi[k] = -DBL_MAX; }
  }
  bool is_empty() const {
    for (int k = 0; k < D; ++k) if (lo[k] > hi[k]) return true;
    return false;
  }
  void extend(const double* p) {
    for (int k = 0; k < D; ++k) { lo[k] = smin(lo[k], p[k]); hi[k] = smax(hi[k], p[k]); }
  }
  void extend(const ABox& b) {
    for (int k = 0; k < D; ++k) { lo[k] = smin(lo[k], b.lo[k]); hi[k] = smax(hi[k], b.hi[k]); }
  }
  bool intersects(const ABox& b) const {
    for (int k = 0; k < D; ++k) if (!(lo[k] <= b.hi[k])) return false;
    for (int k = 0; k < D; ++k) if (!(b.lo[k] <= hi[k])) return false;
    return true;
  }
};

// mcl::BoundingBox: swept box + t0/t1 boxes + active flag (BVHLeaf.hpp:12-29).
template <int D> struct BBox {
  ABox<D> sw, t0, t1;
  bool active = true;
  BBox() { sw.set_empty(); t0.set_empty(); t1.set_empty(); }
  // BVHLeaf.cpp:9-23
  BBox merged(const BBox& b) const {
    BBox r;
    r.active = active || b.active;
    r.sw.extend(sw.lo); r.sw.extend(sw.hi); r.sw.extend(b.sw.lo); r.sw.extend(b.sw.hi);
    r.t0.extend(t0); r.t0.extend(b.t0);
    r.t1.extend(t1); r.t1.extend(b.t1);
    return r;
  }
};

// BVHTree.cpp:400-426 on two boxes.
template <int D> inline bool boxes_intersect(const BBox<D>& l, const BBox<D>& r) {
  if (!l.active && !r.active) return false;
  if (!l.sw.intersects(r.sw)) return false;
  if (l.t0.is_empty() || r.t0.is_empty()) return true;
  if (l.t1.is_empty() || r.t1.is_empty()) return true;
  for (int k = 0; k < D; ++k) {
    if ((l.t0.lo[k] > r.t0.hi[k] && l.t1.lo[k] > r.t1.hi[k]) ||
        (l.t0.hi[k] < r.t0.lo[k] && l.t1.hi[k] < r.t1.lo[k]))
      return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Projections (Projection.hpp).
// ---------------------------------------------------------------------------
// Projection.hpp:60-145 — closest point of triangle (a,b,c) to q.
template <typename T = double>
inline void point_on_triangle(const T* q, const T* a, const T* b, const T* c, T* out) {
  auto clamp01 = [](const T& v) {
    note_branch(30 + clamp_state(v));
    return v < 0.0 ? T(0.0) : (v > 1.0 ? T(1.0) : v);
  };
  T e0[3], e1[3], w[3];
  for (int k = 0; k < 3; ++k) { e0[k] = b[k] - a[k]; e1[k] = c[k] - a[k]; w[k] = a[k] - q[k]; }
  const T A = vdot<3, T>(e0, e0), B = vdot<3, T>(e0, e1), C = vdot<3, T>(e1, e1);
  const T Dd = vdot<3, T>(e0, w), E = vdot<3, T>(e1, w);
  const T det = A * C - B * B;
  T s = B * E - C * Dd;
  T t = B * Dd - A * E;
  bool quirk = false;  // Projection.hpp:131 (s >= 0, t < 0, s + t >= det, a + d <= b + e)
  int region;
  if (s + t < det) {
    if (s < 0.0) {
      if (t < 0.0) {
        if (Dd < 0.0) { region = 1; s = clamp01(-Dd / A); t = 0.0; }
        else { region = 2; s = 0.0; t = clamp01(-E / C); }
      } else { region = 3; s = 0.0; t = clamp01(-E / C); }
    } else if (t < 0.0) {
      region = 4; s = clamp01(-Dd / A); t = 0.0;
    } else {
      region = 5;
      const T inv = T(1.0) / det;
      s = s * inv; t = t * inv;
    }
  } else {
    if (s < 0.0) {
      const T t0 = B + Dd, t1 = C + E;
      if (t1 > t0) {
        region = 6;
        const T num = t1 - t0, den = A - T(2.0) * B + C;
        s = clamp01(num / den); t = T(1.0) - s;
      } else { region = 7; t = clamp01(-E / C); s = 0.0; }
    } else if (t < 0.0) {
      if (A + Dd > B + E) {
        region = 8;
        const T num = C + E - B - Dd, den = A - T(2.0) * B + C;
        s = clamp01(num / den); t = T(1.0) - s;
      } else { region = 9; s = clamp01(-E / C); t = 0.0; quirk = true; }
    } else {
      region = 10;
      const T num = C + E - B - Dd, den = A - T(2.0) * B + C;
      s = clamp01(num / den); t = T(1.0) - s;
    }
  }
  note_branch(10 + region);
  // Every branch but the quirk one yields the minimiser (s,t) of |q - x(s,t)|
  // over the triangle (or over the edge it clamps to), where the distance's
  // derivative is the partial at fixed (s,t) (envelope theorem).  Holding
  // them keeps rounding noise out of the derivative; the quirk branch's s is
  // not a minimiser and keeps its derivative.
  if (!quirk) { s = Held<T>::of(s); t = Held<T>::of(t); }
  for (int k = 0; k < 3; ++k) out[k] = (a[k] + e0[k] * s) + e1[k] * t;
}

// Projection.hpp:193-204 — closest point of segment (e0,e1) to p (2D).
template <typename T = double> inline void point_on_edge2(const T* p, const T* e0, const T* e1, T* out) {
  T e[2] = {e1[0] - e0[0], e1[1] - e0[1]};
  const T len2 = vdot<2, T>(e, e);
  if (len2 <= 0.0) { note_branch(40); out[0] = e0[0]; out[1] = e0[1]; return; }
  T pe[2] = {p[0] - e0[0], p[1] - e0[1]};
  const T t = Held<T>::of(vdot<2, T>(pe, e) / len2);  // minimiser: derivative at fixed t
  note_branch(41 + clamp_state(t));
  if (t < 0.0) { out[0] = e0[0]; out[1] = e0[1]; return; }
  if (t > 1.0) { out[0] = e1[0]; out[1] = e1[1]; return; }
  out[0] = e0[0] + t * e[0];
  out[1] = e0[1] + t * e[1];
}

// Projection.hpp:241-302 — shortest vector from segment p to segment q.
template <typename T = double>
inline void edge_to_edge(const T* p0, const T* p1, const T* q0, const T* q1, T* out, T* bary = nullptr) {
  auto clamp01 = [](const T& x) {
    note_branch(50 + clamp_state(x));
    return smin<T>(T(1.0), smax<T>(x, T(0.0)));
  };
  T d1[3], d2[3], r[3];
  for (int k = 0; k < 3; ++k) { d1[k] = p1[k] - p0[k]; d2[k] = q1[k] - q0[k]; r[k] = p0[k] - q0[k]; }
  const T a = vdot<3, T>(d1, d1), e = vdot<3, T>(d2, d2), f = vdot<3, T>(d2, r);
  T s = 0.0, t = 0.0;
  const T c = vdot<3, T>(d1, r), b = vdot<3, T>(d1, d2);
  const T denom = a * e - b * b;
  if (denom != 0.0) s = clamp01((b * f - c * e) / denom);
  else s = 0.0;
  const T tnom = b * s + f;
  note_branch(60 + (denom != 0.0) + 2 * (tnom < 0.0 || e == 0.0 ? 0 : (tnom > e ? 1 : 2)) + 8 * (a == 0.0));
  if (tnom < 0.0 || e == 0.0) {
    t = 0.0;
    s = (a == 0.0) ? T(0.0) : clamp01(-c / a);
  } else if (tnom > e) {
    t = 1.0;
    s = (a == 0.0) ? T(0.0) : clamp01((b - c) / a);
  } else {
    t = tnom / e;
  }
  // (s,t) minimises the segment distance: derivative at fixed (s,t)
  s = Held<T>::of(s);
  t = Held<T>::of(t);
  for (int k = 0; k < 3; ++k) {
    const T c1 = p0[k] + s * d1[k];
    const T c2 = q0[k] + t * d2[k];
    out[k] = c2 - c1;
  }
  if (bary) { bary[0] = T(1.0) - s; bary[1] = s; bary[2] = T(1.0) - t; bary[3] = t; }
}

// NarrowPhase.cpp:714-750
template <int D, typename T = double> inline T pair_distance(const T (*x)[D], bool is_vf) {
  if constexpr (D == 2) {
    T pt[2];
    point_on_edge2<T>(x[0], x[1], x[2], pt);
    T d[2] = {x[0][0] - pt[0], x[0][1] - pt[1]};
    return vnorm<2, T>(d);
  } else {
    if (is_vf) {
      T pt[3];
      point_on_triangle<T>(x[0], x[1], x[2], x[3], pt);
      T d[3] = {x[0][0] - pt[0], x[0][1] - pt[1], x[0][2] - pt[2]};
      return vnorm<3, T>(d);
    }
    T v[3];
    edge_to_edge<T>(x[0], x[1], x[2], x[3], v);
    return vnorm<3, T>(v);
  }
}

struct AccdStats {
  int64_t dist_evals = 0;  // pair_distance calls
  int64_t iters = 0;       // loop iterations entered
};

// NarrowPhase.cpp:634-711 — additive CCD, evaluated literally (including the
// redundant first-iteration distance evaluation) so it is the checker.
template <int D, typename T = double>
inline bool additive_ccd(const T (*v0)[D], const T (*v1)[D], double eta, bool is_vf, T& toi, AccdStats* st = nullptr) {
  constexpr int ns = D + 1;
  // s, t_c and 1/ns are T in the reference: float for the float tree
  using S = std::conditional_t<std::is_same_v<T, float>, float, double>;
  const S s = 1, t_c = 1;
  const T xsi = eta;
  T p[ns][D], x[ns][D], pbar[D];
  for (int k = 0; k < D; ++k) pbar[k] = 0.0;
  const S inv_ns = S(1) / S(ns);
  for (int i = 0; i < ns; ++i)
    for (int k = 0; k < D; ++k) {
      p[i][k] = v1[i][k] - v0[i][k];
      x[i][k] = v0[i][k];
      pbar[k] = pbar[k] + inv_ns * p[i][k];
    }
  for (int i = 0; i < ns; ++i)
    for (int k = 0; k < D; ++k) p[i][k] = p[i][k] - pbar[k];
  T lf = 0.0, ls = 0.0;
  int af = -1, as = -1;  // argmax (signature only)
  for (int i = 0; i < ns; ++i) {
    const bool first = is_vf ? i == 0 : i <= 1;
    const T n = vnorm<D, T>(p[i]);
    if (first) { if (lf < n) af = i; lf = smax<T>(lf, n); }
    else { if (ls < n) as = i; ls = smax<T>(ls, n); }
  }
  note_branch(70 + af + 8 * as);
  const T l_p = lf + ls;
  if (l_p <= 0.0) return false;
  T d = pair_distance<D, T>(x, is_vf);
  if (st) st->dist_evals++;
  const T g = s * (d * d - xsi * xsi) / (d - xsi);
  toi = 0.0;
  T t_l = (1 - s) * (d * d - xsi * xsi) / ((d + xsi) * l_p);
  if (d <= xsi) return true;
  const int max_iter = 1000;
  for (int iter = 0; iter < max_iter; ++iter) {
    if (st) st->iters++;
    for (int i = 0; i < ns; ++i)
      for (int k = 0; k < D; ++k) x[i][k] = x[i][k] + t_l * p[i][k];
    d = pair_distance<D, T>(x, is_vf);
    if (st) st->dist_evals++;
    const T eps = (d * d - xsi * xsi) / (d + xsi);
    if (toi > 0.0 && iter > 0 && eps < g) { note_branch(200 + iter); break; }
    toi = toi + t_l;
    if (toi > t_c) { note_branch(1300 + iter); return false; }
    t_l = 0.9 * (d * d - xsi * xsi) / ((d + xsi) * l_p);
  }
  return true;  // also reached at max_iter (the release-build assert is compiled out)
}

// ---------------------------------------------------------------------------
// Guigue–Devillers triangle-triangle overlap (tt_isect.hpp), same arithmetic
// and branch structure; macros restated as functions.
// ---------------------------------------------------------------------------
namespace tt {
inline void sub3(double* d, const double* a, const double* b) { d[0] = a[0] - b[0]; d[1] = a[1] - b[1]; d[2] = a[2] - b[2]; }
inline void cross3(double* d, const double* a, const double* b) {
  d[0] = a[1] * b[2] - a[2] * b[1];
  d[1] = a[2] * b[0] - a[0] * b[2];
  d[2] = a[0] * b[1] - a[1] * b[0];
}
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double orient2(const double* a, const double* b, const double* c) {
  return (a[0] - c[0]) * (b[1] - c[1]) - (a[1] - c[1]) * (b[0] - c[0]);
}

// tt_isect.hpp:502-534 (INTERSECTION_TEST_VERTEX)
inline int isect_vertex(const double* P1, const double* Q1, const double* R1, const double* P2, const double* Q2,
                        const double* R2) {
  if (orient2(R2, P2, Q1) >= 0.0) {
    if (orient2(R2, Q2, Q1) <= 0.0) {
      if (orient2(P1, P2, Q1) > 0.0) return orient2(P1, Q2, Q1) <= 0.0 ? 1 : 0;
      if (orient2(P1, P2, R1) >= 0.0) return orient2(Q1, R1, P2) >= 0.0 ? 1 : 0;
      return 0;
    }
    if (orient2(P1, Q2, Q1) <= 0.0) {
      if (orient2(R2, Q2, R1) <= 0.0) return orient2(Q1, R1, Q2) >= 0.0 ? 1 : 0;
      return 0;
    }
    return 0;
  }
  if (orient2(R2, P2, R1) >= 0.0) {
    if (orient2(Q1, R1, R2) >= 0.0) return orient2(P1, P2, R1) >= 0.0 ? 1 : 0;
    if (orient2(Q1, R1, Q2) >= 0.0) return orient2(R2, R1, Q2) >= 0.0 ? 1 : 0;
    return 0;
  }
  return 0;
}

// tt_isect.hpp:538-549 (INTERSECTION_TEST_EDGE)
inline int isect_edge(const double* P1, const double* Q1, const double* R1, const double* P2, const double* Q2,
                      const double* R2) {
  (void)Q2;
  if (orient2(R2, P2, Q1) >= 0.0) {
    if (orient2(P1, P2, Q1) >= 0.0) return orient2(P1, Q1, R2) >= 0.0 ? 1 : 0;
    if (orient2(Q1, R1, P2) >= 0.0) return orient2(R1, P1, P2) >= 0.0 ? 1 : 0;
    return 0;
  }
  if (orient2(R2, P2, R1) >= 0.0) {
    if (orient2(P1, P2, R1) >= 0.0) {
      if (orient2(P1, R1, R2) >= 0.0) return 1;
      return orient2(Q1, R1, R2) >= 0.0 ? 1 : 0;
    }
    return 0;
  }
  return 0;
}

// tt_isect.hpp:553-569
inline int ccw_tri_tri_2d(const double* p1, const double* q1, const double* r1, const double* p2, const double* q2,
                          const double* r2) {
  if (orient2(p2, q2, p1) >= 0.0) {
    if (orient2(q2, r2, p1) >= 0.0) {
      if (orient2(r2, p2, p1) >= 0.0) return 1;
      return isect_edge(p1, q1, r1, p2, q2, r2);
    }
    if (orient2(r2, p2, p1) >= 0.0) return isect_edge(p1, q1, r1, r2, p2, q2);
    return isect_vertex(p1, q1, r1, p2, q2, r2);
  }
  if (orient2(q2, r2, p1) >= 0.0) {
    if (orient2(r2, p2, p1) >= 0.0) return isect_edge(p1, q1, r1, q2, r2, p2);
    return isect_vertex(p1, q1, r1, q2, r2, p2);
  }
  return isect_vertex(p1, q1, r1, r2, p2, q2);
}

// tt_isect.hpp:572-585
inline int tri_tri_2d(const double* p1, const double* q1, const double* r1, const double* p2, const double* q2,
                      const double* r2) {
  if (orient2(p1, q1, r1) < 0.0) {
    if (orient2(p2, q2, r2) < 0.0) return ccw_tri_tri_2d(p1, r1, q1, p2, r2, q2);
    return ccw_tri_tri_2d(p1, r1, q1, p2, q2, r2);
  }
  if (orient2(p2, q2, r2) < 0.0) return ccw_tri_tri_2d(p1, q1, r1, p2, r2, q2);
  return ccw_tri_tri_2d(p1, q1, r1, p2, q2, r2);
}

// tt_isect.hpp:215-270
inline int coplanar_3d(const double* p1, const double* q1, const double* r1, const double* p2, const double* q2,
                       const double* r2, const double* N1) {
  double P1[2], Q1[2], R1[2], P2[2], Q2[2], R2[2];
  const double nx = N1[0] < 0 ? -N1[0] : N1[0];
  const double ny = N1[1] < 0 ? -N1[1] : N1[1];
  const double nz = N1[2] < 0 ? -N1[2] : N1[2];
  auto put = [](double* o, const double* v, int i, int j) { o[0] = v[i]; o[1] = v[j]; };
  if (nx > nz && nx >= ny) {  // project onto YZ
    put(P1, q1, 2, 1); put(Q1, p1, 2, 1); put(R1, r1, 2, 1);
    put(P2, q2, 2, 1); put(Q2, p2, 2, 1); put(R2, r2, 2, 1);
  } else if (ny > nz && ny >= nx) {  // project onto XZ
    put(P1, q1, 0, 2); put(Q1, p1, 0, 2); put(R1, r1, 0, 2);
    put(P2, q2, 0, 2); put(Q2, p2, 0, 2); put(R2, r2, 0, 2);
  } else {  // project onto XY
    put(P1, p1, 0, 1); put(Q1, q1, 0, 1); put(R1, r1, 0, 1);
    put(P2, p2, 0, 1); put(Q2, q2, 0, 1); put(R2, r2, 0, 1);
  }
  return tri_tri_2d(P1, Q1, R1, P2, Q2, R2);
}

// tt_isect.hpp:91-102 (CHECK_MIN_MAX)
inline int check_min_max(const double* p1, const double* q1, const double* r1, const double* p2, const double* q2,
                         const double* r2) {
  double v1[3], v2[3], N[3];
  sub3(v1, p2, q1); sub3(v2, p1, q1); cross3(N, v1, v2);
  sub3(v1, q2, q1);
  if (dot3(v1, N) > 0.0) return 0;
  sub3(v1, p2, p1); sub3(v2, r1, p1); cross3(N, v1, v2);
  sub3(v1, r2, p1);
  if (dot3(v1, N) > 0.0) return 0;
  return 1;
}

// tt_isect.hpp:108-130 (TRI_TRI_3D)
inline int tri_tri_3d_canon(const double* p1, const double* q1, const double* r1, const double* p2, const double* q2,
                            const double* r2, double dp2, double dq2, double dr2, const double* N1) {
  if (dp2 > 0.0) {
    if (dq2 > 0.0) return check_min_max(p1, r1, q1, r2, p2, q2);
    if (dr2 > 0.0) return check_min_max(p1, r1, q1, q2, r2, p2);
    return check_min_max(p1, q1, r1, p2, q2, r2);
  }
  if (dp2 < 0.0) {
    if (dq2 < 0.0) return check_min_max(p1, q1, r1, r2, p2, q2);
    if (dr2 < 0.0) return check_min_max(p1, q1, r1, q2, r2, p2);
    return check_min_max(p1, r1, q1, p2, q2, r2);
  }
  if (dq2 < 0.0) {
    if (dr2 >= 0.0) return check_min_max(p1, r1, q1, q2, r2, p2);
    return check_min_max(p1, q1, r1, p2, q2, r2);
  }
  if (dq2 > 0.0) {
    if (dr2 > 0.0) return check_min_max(p1, r1, q1, p2, q2, r2);
    return check_min_max(p1, q1, r1, q2, r2, p2);
  }
  if (dr2 > 0.0) return check_min_max(p1, q1, r1, r2, p2, q2);
  if (dr2 < 0.0) return check_min_max(p1, r1, q1, r2, p2, q2);
  return coplanar_3d(p1, q1, r1, p2, q2, r2, N1);
}

// tt_isect.hpp:141-211
inline int tri_tri_overlap_3d(const double* p1, const double* q1, const double* r1, const double* p2,
                              const double* q2, const double* r2) {
  double v1[3], v2[3], N1[3], N2[3];
  sub3(v1, p2, r2); sub3(v2, q2, r2); cross3(N2, v1, v2);
  sub3(v1, p1, r2); const double dp1 = dot3(v1, N2);
  sub3(v1, q1, r2); const double dq1 = dot3(v1, N2);
  sub3(v1, r1, r2); const double dr1 = dot3(v1, N2);
  if (((dp1 * dq1) > 0.0) && ((dp1 * dr1) > 0.0)) return 0;
  sub3(v1, q1, p1); sub3(v2, r1, p1); cross3(N1, v1, v2);
  sub3(v1, p2, r1); const double dp2 = dot3(v1, N1);
  sub3(v1, q2, r1); const double dq2 = dot3(v1, N1);
  sub3(v1, r2, r1); const double dr2 = dot3(v1, N1);
  if (((dp2 * dq2) > 0.0) && ((dp2 * dr2) > 0.0)) return 0;
  if (dp1 > 0.0) {
    if (dq1 > 0.0) return tri_tri_3d_canon(r1, p1, q1, p2, r2, q2, dp2, dr2, dq2, N1);
    if (dr1 > 0.0) return tri_tri_3d_canon(q1, r1, p1, p2, r2, q2, dp2, dr2, dq2, N1);
    return tri_tri_3d_canon(p1, q1, r1, p2, q2, r2, dp2, dq2, dr2, N1);
  }
  if (dp1 < 0.0) {
    if (dq1 < 0.0) return tri_tri_3d_canon(r1, p1, q1, p2, q2, r2, dp2, dq2, dr2, N1);
    if (dr1 < 0.0) return tri_tri_3d_canon(q1, r1, p1, p2, q2, r2, dp2, dq2, dr2, N1);
    return tri_tri_3d_canon(p1, q1, r1, p2, r2, q2, dp2, dr2, dq2, N1);
  }
  if (dq1 < 0.0) {
    if (dr1 >= 0.0) return tri_tri_3d_canon(q1, r1, p1, p2, r2, q2, dp2, dr2, dq2, N1);
    return tri_tri_3d_canon(p1, q1, r1, p2, q2, r2, dp2, dq2, dr2, N1);
  }
  if (dq1 > 0.0) {
    if (dr1 > 0.0) return tri_tri_3d_canon(p1, q1, r1, p2, r2, q2, dp2, dr2, dq2, N1);
    return tri_tri_3d_canon(q1, r1, p1, p2, q2, r2, dp2, dq2, dr2, N1);
  }
  if (dr1 > 0.0) return tri_tri_3d_canon(r1, p1, q1, p2, q2, r2, dp2, dq2, dr2, N1);
  if (dr1 < 0.0) return tri_tri_3d_canon(r1, p1, q1, p2, r2, q2, dp2, dr2, dq2, N1);
  return coplanar_3d(p1, q1, r1, p2, q2, r2, N1);
}
}  // namespace tt

// NarrowPhase.cpp:229-257 — 2D segment-segment test.
inline bool discrete_edge_edge2(const double* p0, const double* p1, const double* q0, const double* q1) {
  const double eps = DBL_EPSILON;
  const double n[2] = {q0[0] - p0[0], q0[1] - p0[1]};
  const double r[2] = {p1[0] - p0[0], p1[1] - p0[1]};
  const double s[2] = {q1[0] - q0[0], q1[1] - q0[1]};
  const double rxs = r[0] * s[1] - r[1] * s[0];
  if (std::abs(rxs) < eps) return false;
  const double nxr = n[0] * r[1] - n[1] * r[0];
  if (std::abs(nxr) < eps)
    return ((q0[0] - p0[0] < 0) != (q0[0] - p1[0] < 0)) || ((q0[1] - p0[1] < 0) != (q0[1] - p1[1] < 0));
  const double nxs = n[0] * s[1] - n[1] * s[0];
  const double inv = 1.0 / rxs;
  const double t = nxs * inv, u = nxr * inv;
  return (t >= 0) && (t <= 1) && (u >= 0) && (u <= 1);
}

// ---------------------------------------------------------------------------
// Options (BVHTree.hpp:42-64)
// ---------------------------------------------------------------------------
struct Options {
  double box_eta = 1e-6, vf_ccd_eta = 1e-6, ee_ccd_eta = 1e-6;
  bool reptri = true, threaded = true, discrete = true, continuous = true;
  bool vf_one_sided = false;  // CTCD for vertex-face stencils (ctcd_oracle.hpp)
  // BVHTree<float,DIM> (BVHTree.cpp:504-507): float coordinates and etas, boxes
  // padded in float, float KdBVH centres, NarrowPhaseACCD<float,DIM>; CTCD and
  // the discrete tests widen to double (NarrowPhase.cpp:208-219,259-269,425-444)
  bool f32 = false;
};

struct Stencil {
  int v[4];
  bool is_vf;
};

// ---------------------------------------------------------------------------
// The tree: leaves + a median-split KdBVH (KdBVH.hpp) + the traversal.
// ---------------------------------------------------------------------------
template <int D> class Tree {
 public:
  Options options;

  // per leaf (BVHLeaf.hpp:31-52); flags start at (1,1,1) as the constructor sets them
  struct Leaf {
    int idx = -1;
    int v[3] = {1, 1, 1}, e[3] = {1, 1, 1};
    BBox<D> box;
  };
  std::vector<Leaf> leaves;

  // KdBVH storage (KdBVH.hpp:224-226): internal volumes, child indices and the
  // reordered object list (as leaf indices).
  std::vector<BBox<D>> vols;
  std::vector<int> children;
  std::vector<int> objects;  // objects[pos] = leaf index
  std::vector<int> rank;     // rank[leaf] = pos (in-order position; left subtree first)

  // outputs of traverse
  std::vector<std::pair<int, int>> leaf_pairs;  // (left leaf, right leaf) as collide orients them
  std::vector<Stencil> candidates;
  struct Hit { Stencil s; double toi; };
  std::vector<Hit> hits;
  std::vector<std::pair<int, int>> discrete;
  AccdStats stats;
  bool want_leaf_pairs = false, want_candidates = false;
  int nthreads = 0;  // 0 = OpenMP default

  // BVHTree.cpp:27-126
  void update(const double* V0_, const double* V1_, const int* P, int np, int pdim, const int* active = nullptr) {
    if (np == 0) {
      leaves.clear(); vols.clear(); children.clear(); objects.clear(); rank.clear();
      return;
    }
    std::vector<double> r0, r1;
    const double* V0 = as_t(V0_, P, np, pdim, r0);
    const double* V1 = as_t(V1_, P, np, pdim, r1);
    if (options.f32) {  // the float tree's options are floats
      options.box_eta = (float)options.box_eta;
      options.vf_ccd_eta = (float)options.vf_ccd_eta;
      options.ee_ccd_eta = (float)options.ee_ccd_eta;
    }
    options.vf_ccd_eta = std::min(options.vf_ccd_eta, options.box_eta);
    options.ee_ccd_eta = std::min(options.ee_ccd_eta, options.box_eta);
    bool redo_reptri = false;
    if ((int)leaves.size() != np) {
      leaves.resize(np);
      redo_reptri = pdim == 2 || pdim == 3;
    }
    if (options.reptri && redo_reptri) compute_reptri(P, np, pdim);
    const double eta = options.box_eta;
#pragma omp parallel for schedule(static) num_threads(nthreads_eff())
    for (int i = 0; i < np; ++i) {
      Leaf& L = leaves[i];
      L.box.active = active == nullptr;
      L.idx = i;
      L.box.sw.set_empty(); L.box.t0.set_empty(); L.box.t1.set_empty();
      for (int j = 0; j < pdim; ++j) {
        const int vi = P[i * pdim + j];
        double a[D], b[D];
        for (int k = 0; k < D; ++k) { a[k] = V0[vi * D + k]; b[k] = V1[vi * D + k]; }
        L.box.t0.extend(a); L.box.t1.extend(b);
        L.box.sw.extend(a); L.box.sw.extend(b);
        if (active != nullptr && active[vi]) L.box.active = true;
      }
      if (options.f32) {  // float - float, float + float
        const float e = (float)eta;
        for (int k = 0; k < D; ++k) {
          L.box.t0.lo[k] = (float)L.box.t0.lo[k] - e; L.box.t1.lo[k] = (float)L.box.t1.lo[k] - e;
          L.box.sw.lo[k] = (float)L.box.sw.lo[k] - e;
          L.box.t0.hi[k] = (float)L.box.t0.hi[k] + e; L.box.t1.hi[k] = (float)L.box.t1.hi[k] + e;
          L.box.sw.hi[k] = (float)L.box.sw.hi[k] + e;
        }
        continue;
      }
      for (int k = 0; k < D; ++k) {
        L.box.t0.lo[k] -= eta; L.box.t1.lo[k] -= eta; L.box.sw.lo[k] -= eta;
        L.box.t0.hi[k] += eta; L.box.t1.hi[k] += eta; L.box.sw.hi[k] += eta;
      }
    }
    build_kdbvh();
  }

  // The coordinates the tree computes on: V itself, or for the float tree
  // float(V) (a copy of the vertices P uses; others are never read).
  const double* as_t(const double* V, const int* P, int np, int pdim, std::vector<double>& buf) const {
    if (!options.f32 || !V) return V;
    int nv = 0;
    for (long long i = 0; i < (long long)np * pdim; ++i) nv = std::max(nv, P[i] + 1);
    buf.resize((size_t)nv * D);
    for (size_t i = 0; i < buf.size(); ++i) buf[i] = (float)V[i];
    return buf.data();
  }

  // BVHTree.cpp:53-79 — first-seen vertex / edge ownership, in face order.
  void compute_reptri(const int* P, int np, int pdim) {
    std::unordered_set<int> seen_v;
    std::set<std::pair<int, int>> seen_e;  // same identity as the reference's "a b" string keys
    for (int i = 0; i < np; ++i) {
      Leaf& L = leaves[i];
      for (int j = 0; j < 3; ++j) { L.v[j] = 0; L.e[j] = 0; }
      for (int j = 0; j < pdim; ++j) {
        const int vi = P[i * pdim + j];
        if (seen_v.insert(vi).second) L.v[j] = 1;
        if (pdim != 3) continue;
        int a = vi, b = P[i * pdim + ((j + 1) % 3)];
        if (b < a) std::swap(a, b);
        if (seen_e.insert({a, b}).second) L.e[j] = 1;
      }
    }
  }

  // KdBVH.hpp:103-134,192-222
  struct VIPair { double c[D]; int second; };
  void build_kdbvh() {
    vols.clear(); children.clear(); objects.clear(); rank.clear();
    const int n = (int)leaves.size();
    objects.resize(n);
    for (int i = 0; i < n; ++i) objects[i] = i;
    if (n < 2) { rank.assign(n, 0); return; }
    std::vector<VIPair> cent(n);
    for (int i = 0; i < n; ++i) {
      for (int k = 0; k < D; ++k)  // AlignedBox::center() in T
        cent[i].c[k] = options.f32 ? (double)(((float)leaves[i].box.sw.lo[k] + (float)leaves[i].box.sw.hi[k]) / 2.0f)
                                   : (leaves[i].box.sw.lo[k] + leaves[i].box.sw.hi[k]) / 2.0;
      cent[i].second = i;
    }
    vols.reserve(n - 1);
    children.reserve(2 * n - 2);
    build_rec(cent, 0, n, 0, n);
    for (int i = 0; i < n; ++i) objects[i] = cent[i].second;
    rank.assign(n, 0);
    for (int i = 0; i < n; ++i) rank[objects[i]] = i;
  }
  void build_rec(std::vector<VIPair>& cent, int from, int to, int dim, int n) {
    auto cmp = [dim](const VIPair& a, const VIPair& b) { return a.c[dim] < b.c[dim]; };
    if (to - from == 2) {
      vols.push_back(leaves[cent[from].second].box.merged(leaves[cent[from + 1].second].box));
      children.push_back(from + n - 1);
      children.push_back(from + n);
    } else if (to - from == 3) {
      const int mid = from + 2;
      std::nth_element(cent.begin() + from, cent.begin() + mid, cent.begin() + to, cmp);
      build_rec(cent, from, mid, (dim + 1) % D, n);
      const int i1 = (int)vols.size() - 1;
      vols.push_back(vols[i1].merged(leaves[cent[mid].second].box));
      children.push_back(i1);
      children.push_back(mid + n - 1);
    } else {
      const int mid = from + (to - from) / 2;
      std::nth_element(cent.begin() + from, cent.begin() + mid, cent.begin() + to, cmp);
      build_rec(cent, from, mid, (dim + 1) % D, n);
      const int i1 = (int)vols.size() - 1;
      build_rec(cent, mid, to, (dim + 1) % D, n);
      const int i2 = (int)vols.size() - 1;
      vols.push_back(vols[i1].merged(vols[i2]));
      children.push_back(i1);
      children.push_back(i2);
    }
  }

  // node = (index, is_leaf); KdBVH.hpp:141-170 + BVHTree.cpp:187-220
  using Node = std::pair<int, bool>;
  void get_children(const Node& nd, Node& l, Node& r) const {
    const int nb = (int)vols.size();
    const int c0 = children[2 * nd.first], c1 = children[2 * nd.first + 1];
    if (c1 < nb) { l = {c0, false}; r = {c1, false}; }
    else if (c0 >= nb) { l = {objects[c0 - nb], true}; r = {objects[c0 - nb + 1], true}; }
    else { l = {c0, false}; r = {objects[c1 - nb], true}; }
  }
  const BBox<D>& get_box(const Node& nd) const { return nd.second ? leaves[nd.first].box : vols[nd.first]; }

  // BVHTree.cpp:168-185 — post-order list of (left child, right child)
  void make_frontlist(const Node& nd, std::vector<std::pair<Node, Node>>& q) const {
    if (nd.second) return;
    Node l, r;
    get_children(nd, l, r);
    make_frontlist(l, q);
    make_frontlist(r, q);
    q.emplace_back(l, r);
  }

  // The order in which collide produces its records: per leaf pair reached
  // (ordinal pair), its candidates (kind 0, when kept), hits (kind 1) and
  // discrete hit (kind 2), each an index into that list.  With one thread
  // (threaded = false) this is the reference's serial callback sequence.
  struct Event {
    long long pair;
    int kind;
    long long idx;
  };
  std::vector<Event> events;

  struct Local {
    std::vector<std::pair<int, int>> lp;
    std::vector<Stencil> cand;
    std::vector<Hit> hits;
    std::vector<std::pair<int, int>> disc;
    std::vector<Event> ev;
    long long npairs = 0;
    AccdStats st;
  };

  // BVHTree.cpp:129-159
  void traverse(const double* V0_, const double* V1_, const int* P) {
    leaf_pairs.clear(); candidates.clear(); hits.clear(); discrete.clear(); events.clear(); stats = AccdStats();
    if (leaves.size() <= 1) return;
    std::vector<double> r0, r1;
    const double* V0 = as_t(V0_, P, (int)leaves.size(), D, r0);
    const double* V1 = as_t(V1_, P, (int)leaves.size(), D, r1);
    std::vector<std::pair<Node, Node>> q;
    q.reserve(leaves.size());
    make_frontlist(Node((int)vols.size() - 1, false), q);
    const int nq = (int)q.size();
    int nt = options.threaded ? nthreads_eff() : 1;
    std::vector<Local> loc(nt);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int i = 0; i < nq; ++i) {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      collide(V0, V1, P, q[i].first, q[i].second, loc[tid]);
    }
    // merge the per-thread lists; each is released once copied, so the peak is
    // about one copy of the results beside the merged lists
    size_t nlp = 0, nc = 0, nh = 0, nd = 0;
    for (auto& L : loc) { nlp += L.lp.size(); nc += L.cand.size(); nh += L.hits.size(); nd += L.disc.size(); }
    leaf_pairs.reserve(nlp); candidates.reserve(nc); hits.reserve(nh); discrete.reserve(nd);
    for (auto& L : loc) {
      leaf_pairs.insert(leaf_pairs.end(), L.lp.begin(), L.lp.end());
      candidates.insert(candidates.end(), L.cand.begin(), L.cand.end());
      hits.insert(hits.end(), L.hits.begin(), L.hits.end());
      discrete.insert(discrete.end(), L.disc.begin(), L.disc.end());
      if (nt == 1) events = std::move(L.ev);
      stats.dist_evals += L.st.dist_evals;
      stats.iters += L.st.iters;
      L = Local();
    }
  }

  // BVHTree.cpp:230-302 (the stop flag only matters with a user callback that
  // asks to stop; this checker never stops).
  void collide(const double* V0, const double* V1, const int* P, const Node& l, const Node& r, Local& out) const {
    if (!boxes_intersect<D>(get_box(l), get_box(r))) return;
    if (l.second && r.second) {
      if (want_leaf_pairs) out.lp.emplace_back(l.first, r.first);
      const long long pi = out.npairs++;
      if (options.continuous) {
        Stencil cs[15];
        const int nc = get_candidates(l.first, r.first, P, cs);
        for (int c = 0; c < nc; ++c) {
          if (want_candidates) {
            out.ev.push_back({pi, 0, (long long)out.cand.size()});
            out.cand.push_back(cs[c]);
          }
          double toi = narrow_phase(V0, V1, cs[c], &out.st);
          if (toi >= 0) {
            out.ev.push_back({pi, 1, (long long)out.hits.size()});
            out.hits.push_back({cs[c], toi});
          }
        }
      }
      if (options.discrete) {
        if (discrete_test(V1, P + l.first * D, P + r.first * D)) {
          out.ev.push_back({pi, 2, (long long)out.disc.size()});
          out.disc.emplace_back(l.first, r.first);
        }
      }
      return;
    }
    Node a, b;
    if (l.second) {
      get_children(r, a, b);
      collide(V0, V1, P, l, a, out);
      collide(V0, V1, P, l, b, out);
    } else {
      get_children(l, a, b);
      collide(V0, V1, P, a, r, out);
      collide(V0, V1, P, b, r, out);
    }
  }

  // BVHTree.cpp:304-398
  int get_candidates(int p0, int p1, const int* P, Stencil* out) const {
    const Leaf& L0 = leaves[p0];
    const Leaf& L1 = leaves[p1];
    int f0[D], f1[D];
    for (int i = 0; i < D; ++i) { f0[i] = P[p0 * D + i]; f1[i] = P[p1 * D + i]; }
    auto shares = [](int v, const int* f) {
      for (int i = 0; i < D; ++i) if (v == f[i]) return true;
      return false;
    };
    int n = 0;
    for (int i = 0; i < D; ++i) {
      if (!L0.v[i] || shares(f0[i], f1)) continue;
      Stencil& s = out[n++];
      s.is_vf = true;
      s.v[0] = f0[i]; s.v[1] = s.v[2] = s.v[3] = -1;
      for (int j = 0; j < D; ++j) s.v[j + 1] = f1[j];
    }
    for (int i = 0; i < D; ++i) {
      if (!L1.v[i] || shares(f1[i], f0)) continue;
      Stencil& s = out[n++];
      s.is_vf = true;
      s.v[0] = f1[i]; s.v[1] = s.v[2] = s.v[3] = -1;
      for (int j = 0; j < D; ++j) s.v[j + 1] = f0[j];
    }
    if (D != 3) return n;
    for (int i = 0; i < 3; ++i) {
      if (!L0.e[i]) continue;
      for (int j = 0; j < 3; ++j) {
        if (!L1.e[j]) continue;
        const int a = f0[i], b = f0[(i + 1) % 3], c = f1[j], d = f1[(j + 1) % 3];
        if (a == c || a == d || b == c || b == d) continue;
        Stencil& s = out[n++];
        s.is_vf = false;
        s.v[0] = a; s.v[1] = b; s.v[2] = c; s.v[3] = d;
      }
    }
    return n;
  }

  // BVHTree.cpp:428-461
  double narrow_phase(const double* V0, const double* V1, const Stencil& s, AccdStats* st) const {
    double a[D + 1][D], b[D + 1][D];
    for (int i = 0; i < D + 1; ++i)
      for (int k = 0; k < D; ++k) { a[i][k] = V0[s.v[i] * D + k]; b[i][k] = V1[s.v[i] * D + k]; }
    double toi = -1;
    const double eta = s.is_vf ? options.vf_ccd_eta : options.ee_ccd_eta;
    if (s.is_vf && options.vf_one_sided) {  // the float specialisation widens to double, stores a float
      if (ctcd::query_vf<D>(a, b, eta, toi) != 1) return -1;
      return options.f32 ? (double)(float)toi : toi;
    }
    if (options.f32) {  // NarrowPhaseACCD<float,DIM>
      float fa[D + 1][D], fb[D + 1][D], ft = -1;
      for (int i = 0; i < D + 1; ++i)
        for (int k = 0; k < D; ++k) { fa[i][k] = (float)a[i][k]; fb[i][k] = (float)b[i][k]; }
      if (additive_ccd<D, float>(fa, fb, eta, s.is_vf, ft, st)) return ft;
      return -1;
    }
    if (additive_ccd<D>(a, b, eta, s.is_vf, toi, st)) return toi;
    return -1;
  }

  // BVHTree.cpp:463-500
  bool discrete_test(const double* V, const int* p0, const int* p1) const {
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j)
        if (p0[i] == p1[j]) return false;
    if constexpr (D == 3) {
      return tt::tri_tri_overlap_3d(V + p0[0] * 3, V + p0[1] * 3, V + p0[2] * 3, V + p1[0] * 3, V + p1[1] * 3,
                                    V + p1[2] * 3) != 0;
    } else {
      return discrete_edge_edge2(V + p0[0] * 2, V + p0[1] * 2, V + p1[0] * 2, V + p1[1] * 2);
    }
  }

  int nthreads_eff() const {
#ifdef _OPENMP
    return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    return 1;
#endif
  }
};

// O(n^2) brute force over all unordered leaf pairs with the exact leaf test.
// Independent of any tree: pins the traversal's leaf-pair set.
template <int D> inline std::vector<std::pair<int, int>> brute_force_pairs(const Tree<D>& t) {
  const int n = (int)t.leaves.size();
  std::vector<std::vector<std::pair<int, int>>> per(n);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      if (boxes_intersect<D>(t.leaves[i].box, t.leaves[j].box)) per[i].emplace_back(i, j);
  std::vector<std::pair<int, int>> out;
  for (auto& v : per) out.insert(out.end(), v.begin(), v.end());
  return out;
}

}  // namespace mclo
