// FP64 narrow-phase geometry for gfx950: VF/EE distances, ACCD, discrete
// triangle-triangle (3D) and segment-segment (2D) tests.
//
// Every expression keeps the reference's evaluation order (x86-64 Release, no
// FMA): dot(a,b) = (a0*b0 + a1*b1) + a2*b2, norm = sqrt(dot), element-wise
// left to right.  The translation unit is compiled with -ffp-contract=off and
// the pragma below; f64 '/' and sqrt lower to the correctly rounded sequences.
// Restructured for SIMD execution: each projection region selects a
// (numerator, denominator) pair so one division serves every branch; the
// triangle-triangle canonical permutation is an index selection followed by a
// single min-max check.
#pragma once
#include "common.hpp"

#pragma clang fp contract(off)

namespace mclg {

// Scalar type T: double for BVHTree<double,DIM>, float for BVHTree<float,DIM>
// (NarrowPhase.cpp:757-770).  Float '+', '-', '*' are the hardware's IEEE
// single-precision operations; float '/' and sqrt go through double and are
// rounded once to float, which is the correctly rounded float result (binary64
// carries more than 2*24+2 bits, so the second rounding cannot differ).
template <typename T> struct vec3 { T x, y, z; };
template <typename T> struct vec2 { T x, y; };
using d3 = vec3<double>;
using d2 = vec2<double>;

template <typename T> __device__ __forceinline__ vec3<T> sub(const vec3<T>& a, const vec3<T>& b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <typename T> __device__ __forceinline__ vec2<T> sub(const vec2<T>& a, const vec2<T>& b) {
  return {a.x - b.x, a.y - b.y};
}
template <typename T> __device__ __forceinline__ T dot(const vec3<T>& a, const vec3<T>& b) {
  return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
template <typename T> __device__ __forceinline__ T dot(const vec2<T>& a, const vec2<T>& b) { return a.x * b.x + a.y * b.y; }
// libstdc++ std::max / std::min (NaN and signed zero as the reference sees them)
template <typename T> __device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }
template <typename T> __device__ __forceinline__ T smin(T a, T b) { return (b < a) ? b : a; }
__device__ __forceinline__ double cr_div(double a, double b) { return a / b; }
__device__ __forceinline__ float cr_div(float a, float b) { return (float)((double)a / (double)b); }
// Correctly rounded FP64 square root.  On gfx950 sqrt lowers to rsq, one
// Goldschmidt step for (g, h) ~ (sqrt x, 1/(2 sqrt x)) and two Markstein
// corrections g += (x - g*g)*h, wrapped in a scaling of x by 2^256 (result by
// 2^-128) when x < 2^-767 and a select that returns x for +-0 and +inf.  For x
// in [2^-767, DBL_MAX] the scalings are by 2^0 and the select keeps the
// iteration's value, so the bare iteration below gives the same bits with
// eight fewer vector instructions; other inputs take the full sequence.  The
// host build of this code (the CPU checks) uses the C library's sqrt.
__device__ __forceinline__ double cr_sqrt(double a) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (a >= 0x1p-767 && a <= 0x1.fffffffffffffp+1023) {
    const double y = __builtin_amdgcn_rsq(a);
    double g = a * y, h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, a);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, a);
    return __builtin_fma(d, h, g);
  }
#endif
  return __builtin_sqrt(a);
}
__device__ __forceinline__ float cr_sqrt(float a) { return (float)cr_sqrt((double)a); }

// ---------------------------------------------------------------------------
// VF distance: |q - closest point of triangle (a,b,c)|
// (Projection.hpp:60-145 then NarrowPhase.cpp:740-741)
// ---------------------------------------------------------------------------
enum : int { R_S = 0, R_T = 1, R_IN = 2, R_EDGE = 3 };

template <typename T>
__device__ __forceinline__ T dist_vf(const vec3<T>& q, const vec3<T>& a, const vec3<T>& b, const vec3<T>& c) {
  const vec3<T> e0 = sub(b, a), e1 = sub(c, a), w = sub(a, q);
  const T A = dot(e0, e0), B = dot(e0, e1), C = dot(e1, e1);
  const T Dd = dot(e0, w), E = dot(e1, w);
  const T det = A * C - B * B;
  T s = B * E - C * Dd;
  T t = B * Dd - A * E;
  T num, den;
  int kind;
  if (s + t < det) {
    if (s < 0) {
      if (t < 0 && Dd < 0) { kind = R_S; num = -Dd; den = A; }
      else { kind = R_T; num = -E; den = C; }
    } else if (t < 0) { kind = R_S; num = -Dd; den = A; }
    else { kind = R_IN; num = T(1); den = det; }
  } else {
    const T ee = A - T(2) * B + C;
    if (s < 0) {
      const T t0 = B + Dd, t1 = C + E;
      if (t1 > t0) { kind = R_EDGE; num = t1 - t0; den = ee; }
      else { kind = R_T; num = -E; den = C; }
    } else if (t < 0) {
      if (A + Dd > B + E) { kind = R_EDGE; num = C + E - B - Dd; den = ee; }
      else { kind = R_S; num = -E; den = C; }
    } else { kind = R_EDGE; num = C + E - B - Dd; den = ee; }
  }
  const T r = cr_div(num, den);
  const T rc = r < 0 ? T(0) : (r > 1 ? T(1) : r);  // clamp_zero_one
  if (kind == R_S) { s = rc; t = T(0); }
  else if (kind == R_T) { s = T(0); t = rc; }
  else if (kind == R_IN) { s *= r; t *= r; }
  else { s = rc; t = T(1) - s; }
  const vec3<T> pt = {(a.x + e0.x * s) + e1.x * t, (a.y + e0.y * s) + e1.y * t, (a.z + e0.z * s) + e1.z * t};
  const vec3<T> dv = sub(q, pt);
  return cr_sqrt(dot(dv, dv));
}

// ---------------------------------------------------------------------------
// EE distance: |shortest vector between segments p and q|
// (Projection.hpp:241-302 then NarrowPhase.cpp:745-746)
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ T clamp01(T x) { return smin(T(1), smax(x, T(0))); }

template <typename T>
__device__ __forceinline__ T dist_ee(const vec3<T>& p0, const vec3<T>& p1, const vec3<T>& q0, const vec3<T>& q1) {
  const vec3<T> d1 = sub(p1, p0), d2 = sub(q1, q0), r = sub(p0, q0);
  const T a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  const T c = dot(d1, r), b = dot(d1, d2);
  const T denom = a * e - b * b;
  T s = T(0), t;
  if (denom != T(0)) s = clamp01(cr_div(b * f - c * e, denom));
  const T tnom = b * s + f;
  T num, den;
  int kind;  // 0: t=0, s=clamp(-c/a); 1: t=1, s=clamp((b-c)/a); 2: t=tnom/e
  if (tnom < 0 || e == 0) { kind = 0; num = -c; den = a; t = T(0); }
  else if (tnom > e) { kind = 1; num = b - c; den = a; t = T(1); }
  else { kind = 2; num = tnom; den = e; }
  if (kind == 2 || a != T(0)) {
    const T q = cr_div(num, den);
    if (kind == 2) t = q;
    else s = clamp01(q);
  } else {
    s = T(0);
  }
  const vec3<T> v = {(q0.x + t * d2.x) - (p0.x + s * d1.x), (q0.y + t * d2.y) - (p0.y + s * d1.y),
                     (q0.z + t * d2.z) - (p0.z + s * d1.z)};
  return cr_sqrt(dot(v, v));
}

// ---------------------------------------------------------------------------
// 2D VF distance: |p - closest point of segment (e0,e1)| (Projection.hpp:193-204)
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ T dist_pe2(const vec2<T>& p, const vec2<T>& e0, const vec2<T>& e1) {
  const vec2<T> e = sub(e1, e0);
  const T len2 = dot(e, e);
  vec2<T> pt = e0;
  if (len2 > T(0) || !(len2 <= T(0))) {  // i.e. not (len2 <= 0)
    const vec2<T> pe = sub(p, e0);
    const T t = cr_div(dot(pe, e), len2);
    if (t < T(0)) pt = e0;
    else if (t > T(1)) pt = e1;
    else pt = {e0.x + t * e.x, e0.y + t * e.y};
  }
  const vec2<T> dv = sub(p, pt);
  return cr_sqrt(dot(dv, dv));
}

// ---------------------------------------------------------------------------
// ACCD (NarrowPhase.cpp:634-711), s = 1, t_c = 1, max_iter = 1000.
//
// Work the reference does whose result provably cannot change the outcome is
// skipped: (1) iteration 0 moves x by t_l0 = 0*(...)/(...) = +0, so when t_l0
// is exactly zero and every displacement is finite x is unchanged up to the
// sign of zero coordinates, which no comparison and no squared norm can see —
// its distance equals d0 and is not recomputed; (2) while toi is not > 0 the
// break test cannot fire, so when toi + t_l > 1 the reference's next distance
// is unused and the miss is returned without evaluating it.  g and the first
// t_l are only needed once the pair is not already within the gap.  Returns 1
// on hit (toi set), 0 otherwise; evals counts the distance evaluations made.
// With T = float every quantity is a float as in NarrowPhaseACCD<float,DIM>,
// except the step t_l = 0.9 * (...) / (...), whose double literal makes the
// reference compute the product and quotient in double before storing a float.
// ---------------------------------------------------------------------------
template <int D, bool VF, typename T> struct Stencil;

template <bool VF, typename T> struct Stencil<3, VF, T> {
  vec3<T> x[4], p[4];
  __device__ __forceinline__ T dist() const {
    if (VF) return dist_vf(x[0], x[1], x[2], x[3]);
    return dist_ee(x[0], x[1], x[2], x[3]);
  }
  __device__ __forceinline__ void step(T tl) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      x[i].x = x[i].x + tl * p[i].x;
      x[i].y = x[i].y + tl * p[i].y;
      x[i].z = x[i].z + tl * p[i].z;
    }
  }
  // v0[i], v1[i] are the stencil's vertices at t0 and t1; returns l_p, sets fin
  __device__ __forceinline__ T setup(const vec3<T>* v0, const vec3<T>* v1, bool& fin) {
    const T w = T(1) / T(4);
    vec3<T> pb = {T(0), T(0), T(0)};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      p[i] = sub(v1[i], v0[i]);
      x[i] = v0[i];
      pb.x = pb.x + w * p[i].x;
      pb.y = pb.y + w * p[i].y;
      pb.z = pb.z + w * p[i].z;
    }
    // l_p = max |p| over the first group + max |p| over the second.  The
    // maxima are taken over the squared norms and rooted once per group: a
    // correctly rounded square root is monotone, so sqrt(max) = max(sqrt)
    // bit for bit (NaNs are skipped by smax either way).
    T lf = T(0), ls = T(0), sum = T(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      p[i] = sub(p[i], pb);
      sum += (p[i].x + p[i].y) + p[i].z;
      const T n2 = dot(p[i], p[i]);
      const bool first = VF ? (i == 0) : (i <= 1);
      if (first) lf = smax(lf, n2);
      else ls = smax(ls, n2);
    }
    fin = __builtin_isfinite(sum);
    return cr_sqrt(lf) + cr_sqrt(ls);
  }
};

template <typename T> struct Stencil<2, true, T> {
  vec2<T> x[3], p[3];
  __device__ __forceinline__ T dist() const { return dist_pe2(x[0], x[1], x[2]); }
  __device__ __forceinline__ void step(T tl) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      x[i].x = x[i].x + tl * p[i].x;
      x[i].y = x[i].y + tl * p[i].y;
    }
  }
  __device__ __forceinline__ T setup(const vec2<T>* v0, const vec2<T>* v1, bool& fin) {
    const T w = T(1) / T(3);
    vec2<T> pb = {T(0), T(0)};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      p[i] = sub(v1[i], v0[i]);
      x[i] = v0[i];
      pb.x = pb.x + w * p[i].x;
      pb.y = pb.y + w * p[i].y;
    }
    T lf = T(0), ls = T(0), sum = T(0);  // squared norms, rooted once per group (as in 3D)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      p[i] = sub(p[i], pb);
      sum += p[i].x + p[i].y;
      const T n2 = dot(p[i], p[i]);
      if (i == 0) lf = smax(lf, n2);
      else ls = smax(ls, n2);
    }
    fin = __builtin_isfinite(sum);
    return cr_sqrt(lf) + cr_sqrt(ls);
  }
};

// The iteration is split so that lock-step waves need not wait for their
// longest stencil: start<STOP>() runs the setup, d0 and iterations [0, STOP);
// a stencil still running at the top of iteration STOP = 3 can be suspended
// with (t1, t2, t3, g) — the steps of iterations 0..2 after the skipped
// iteration 0, and the break threshold — and resumed later: resume() rebuilds
// p and l_p (the same arithmetic), replays the two position updates
// x0 + t1*p, then + t2*p (the same roundings), and toi = ((0 + 0) + t1) + t2,
// so run() continues with exactly the state the uninterrupted loop has.  Three
// evaluations are the common case (d0, d1, and d2 where the break test fires).
enum : int { ACCD_MISS = 0, ACCD_HIT = 1, ACCD_RUN = 2, ACCD_SUSPENDED = 3 };
constexpr int kAccdSuspendIter = 3;

template <int D, bool VF, typename T> struct Accd {
  Stencil<D, VF, T> st;
  T xsi, l_p, g, toi, tl, t1, t2;
  int iter, evals;
  bool skip, fin;
  bool replayable;  // iteration 0 was the skipped no-op step: (t1, t2) rebuild x

  // ACCD_HIT (toi set), ACCD_MISS, or ACCD_RUN at the top of iteration STOP.
  template <int STOP, typename V> __device__ __forceinline__ int start(const V* v0, const V* v1, T xsi_) {
    xsi = xsi_;
    l_p = st.setup(v0, v1, fin);
    evals = 0;
    replayable = false;
    if (l_p <= 0) return ACCD_MISS;
    T d = st.dist();
    evals = 1;
    toi = T(0);
    if (d <= xsi) return ACCD_HIT;
    g = cr_div(T(1) * (d * d - xsi * xsi), d - xsi);
    tl = cr_div((T(1) - T(1)) * (d * d - xsi * xsi), (d + xsi) * l_p);
    skip = tl == T(0) && fin;  // iteration 0 leaves x unchanged: its distance is d0
    replayable = skip;
    iter = 0;
    return loop<STOP>(d);
  }

  template <typename V>
  __device__ __forceinline__ void resume(const V* v0, const V* v1, T xsi_, T t1_, T t2_, T t3_, T g_) {
    xsi = xsi_;
    l_p = st.setup(v0, v1, fin);
    st.step(t1_);
    st.step(t2_);
    toi = T(0);
    toi += T(0);
    toi += t1_;
    toi += t2_;
    tl = t3_;
    g = g_;
    iter = kAccdSuspendIter;
    skip = false;
    evals = 0;
  }

  // iterations [iter, 1000): ACCD_HIT or ACCD_MISS
  __device__ __forceinline__ int run() {
    T d = T(0);  // assigned before use: skip is false from iteration 1 on
    return loop<1000>(d);
  }

  // One distance evaluation site per iteration (reference loop NarrowPhase.cpp:684-707).
  template <int STOP> __device__ __forceinline__ int loop(T& d) {
    for (; iter < STOP; ++iter) {
      const bool can_break = toi > T(0) && iter > 0;
      if (!can_break && toi + tl > T(1)) return ACCD_MISS;  // the reference's distance here is unused
      if (!skip) {
        st.step(tl);
        d = st.dist();
        ++evals;
      }
      skip = false;
      if (can_break) {
        const T eps = cr_div(d * d - xsi * xsi, d + xsi);
        if (eps < g) return ACCD_HIT;
      }
      toi += tl;
      if (toi > T(1)) return ACCD_MISS;
      tl = (T)(0.9 * (double)(d * d - xsi * xsi) / (double)((d + xsi) * l_p));
      if (STOP <= kAccdSuspendIter) {
        if (iter == 0) t1 = tl;
        else if (iter == 1) t2 = tl;
      }
    }
    // max_iter reached: the reference returns true (its assert is compiled out)
    return STOP >= 1000 ? ACCD_HIT : ACCD_RUN;
  }
};

template <int D, bool VF, typename T, typename V>
__device__ __forceinline__ int accd(const V* v0, const V* v1, T xsi, T& toi, int& evals) {
  Accd<D, VF, T> a;
  int s = a.template start<kAccdSuspendIter>(v0, v1, xsi);
  if (s == ACCD_RUN) s = a.run();
  toi = a.toi;
  evals = a.evals;
  return s == ACCD_HIT ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Discrete triangle-triangle overlap at t1 (Guigue & Devillers 2003, as in
// tt_isect.hpp:71-270,499-585): orientation predicates, canonical permutation,
// min-max interval check; coplanar pairs are projected to 2D.
// ---------------------------------------------------------------------------
__device__ __forceinline__ d3 cross(const d3& a, const d3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// plain left-to-right sum (the predicate's own DOT)
__device__ __forceinline__ double dot_lr(const d3& a, const d3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double orient2(const d2& a, const d2& b, const d2& c) {
  return (a.x - c.x) * (b.y - c.y) - (a.y - c.y) * (b.x - c.x);
}

// tt_isect.hpp:502-534
__device__ inline int isect_vertex2(const d2& P1, const d2& Q1, const d2& R1, const d2& P2, const d2& Q2,
                                    const d2& R2) {
  if (orient2(R2, P2, Q1) >= 0.0) {
    if (orient2(R2, Q2, Q1) <= 0.0) {
      if (orient2(P1, P2, Q1) > 0.0) return orient2(P1, Q2, Q1) <= 0.0;
      if (orient2(P1, P2, R1) >= 0.0) return orient2(Q1, R1, P2) >= 0.0;
      return 0;
    }
    if (orient2(P1, Q2, Q1) <= 0.0 && orient2(R2, Q2, R1) <= 0.0) return orient2(Q1, R1, Q2) >= 0.0;
    return 0;
  }
  if (orient2(R2, P2, R1) >= 0.0) {
    if (orient2(Q1, R1, R2) >= 0.0) return orient2(P1, P2, R1) >= 0.0;
    if (orient2(Q1, R1, Q2) >= 0.0) return orient2(R2, R1, Q2) >= 0.0;
  }
  return 0;
}

// tt_isect.hpp:538-549
__device__ inline int isect_edge2(const d2& P1, const d2& Q1, const d2& R1, const d2& P2, const d2& R2) {
  if (orient2(R2, P2, Q1) >= 0.0) {
    if (orient2(P1, P2, Q1) >= 0.0) return orient2(P1, Q1, R2) >= 0.0;
    if (orient2(Q1, R1, P2) >= 0.0) return orient2(R1, P1, P2) >= 0.0;
    return 0;
  }
  if (orient2(R2, P2, R1) >= 0.0 && orient2(P1, P2, R1) >= 0.0) {
    if (orient2(P1, R1, R2) >= 0.0) return 1;
    return orient2(Q1, R1, R2) >= 0.0;
  }
  return 0;
}

// tt_isect.hpp:553-569 (counter-clockwise T1)
__device__ inline int ccw_tri_tri2(const d2& p1, const d2& q1, const d2& r1, const d2& p2, const d2& q2,
                                   const d2& r2) {
  if (orient2(p2, q2, p1) >= 0.0) {
    if (orient2(q2, r2, p1) >= 0.0) {
      if (orient2(r2, p2, p1) >= 0.0) return 1;
      return isect_edge2(p1, q1, r1, p2, r2);
    }
    if (orient2(r2, p2, p1) >= 0.0) return isect_edge2(p1, q1, r1, r2, q2);
    return isect_vertex2(p1, q1, r1, p2, q2, r2);
  }
  if (orient2(q2, r2, p1) >= 0.0) {
    if (orient2(r2, p2, p1) >= 0.0) return isect_edge2(p1, q1, r1, q2, p2);
    return isect_vertex2(p1, q1, r1, q2, r2, p2);
  }
  return isect_vertex2(p1, q1, r1, r2, p2, q2);
}

// tt_isect.hpp:572-585
__device__ inline int tri_tri2(const d2& p1, const d2& q1, const d2& r1, const d2& p2, const d2& q2,
                               const d2& r2) {
  const bool cw1 = orient2(p1, q1, r1) < 0.0;
  const bool cw2 = orient2(p2, q2, r2) < 0.0;
  return ccw_tri_tri2(p1, cw1 ? r1 : q1, cw1 ? q1 : r1, p2, cw2 ? r2 : q2, cw2 ? q2 : r2);
}

// tt_isect.hpp:215-270 — project on the plane maximising the projected area
__device__ inline int coplanar_tri_tri3(const d3& p1, const d3& q1, const d3& r1, const d3& p2, const d3& q2,
                                        const d3& r2, const d3& N1) {
  const double nx = N1.x < 0 ? -N1.x : N1.x;
  const double ny = N1.y < 0 ? -N1.y : N1.y;
  const double nz = N1.z < 0 ? -N1.z : N1.z;
  if (nx > nz && nx >= ny)  // YZ, with p and q exchanged
    return tri_tri2({q1.z, q1.y}, {p1.z, p1.y}, {r1.z, r1.y}, {q2.z, q2.y}, {p2.z, p2.y}, {r2.z, r2.y});
  if (ny > nz && ny >= nx)  // XZ, with p and q exchanged
    return tri_tri2({q1.x, q1.z}, {p1.x, p1.z}, {r1.x, r1.z}, {q2.x, q2.z}, {p2.x, p2.z}, {r2.x, r2.z});
  return tri_tri2({p1.x, p1.y}, {q1.x, q1.y}, {r1.x, r1.y}, {p2.x, p2.y}, {q2.x, q2.y}, {r2.x, r2.y});
}

// tt_isect.hpp:91-102
__device__ __forceinline__ int check_min_max(const d3& p1, const d3& q1, const d3& r1, const d3& p2, const d3& q2,
                                             const d3& r2) {
  d3 N = cross(sub(p2, q1), sub(p1, q1));
  if (dot_lr(sub(q2, q1), N) > 0.0) return 0;
  N = cross(sub(p2, p1), sub(r1, p1));
  if (dot_lr(sub(r2, p1), N) > 0.0) return 0;
  return 1;
}

__device__ __forceinline__ d3 pick3(int i, const d3& a, const d3& b, const d3& c) { return i == 0 ? a : (i == 1 ? b : c); }
__device__ __forceinline__ double pickd(int i, double a, double b, double c) { return i == 0 ? a : (i == 1 ? b : c); }

// tt_isect.hpp:141-211 + 108-130.  T1 = (p1,q1,r1), T2 = (p2,q2,r2).
__device__ inline int tri_tri3(const d3& p1, const d3& q1, const d3& r1, const d3& p2, const d3& q2, const d3& r2) {
  const d3 N2 = cross(sub(p2, r2), sub(q2, r2));
  const double dp1 = dot_lr(sub(p1, r2), N2), dq1 = dot_lr(sub(q1, r2), N2), dr1 = dot_lr(sub(r1, r2), N2);
  if (((dp1 * dq1) > 0.0) && ((dp1 * dr1) > 0.0)) return 0;
  const d3 N1 = cross(sub(q1, p1), sub(r1, p1));
  const double dp2 = dot_lr(sub(p2, r1), N1), dq2 = dot_lr(sub(q2, r1), N1), dr2 = dot_lr(sub(r2, r1), N1);
  if (((dp2 * dq2) > 0.0) && ((dp2 * dr2) > 0.0)) return 0;

  // T1 rotation: 0 = (p1,q1,r1), 1 = (q1,r1,p1), 2 = (r1,p1,q1); sw = T2 becomes (p2,r2,q2)
  int rot1, sw;
  if (dp1 > 0.0) {
    if (dq1 > 0.0) { rot1 = 2; sw = 1; }
    else if (dr1 > 0.0) { rot1 = 1; sw = 1; }
    else { rot1 = 0; sw = 0; }
  } else if (dp1 < 0.0) {
    if (dq1 < 0.0) { rot1 = 2; sw = 0; }
    else if (dr1 < 0.0) { rot1 = 1; sw = 0; }
    else { rot1 = 0; sw = 1; }
  } else if (dq1 < 0.0) {
    if (dr1 >= 0.0) { rot1 = 1; sw = 1; }
    else { rot1 = 0; sw = 0; }
  } else if (dq1 > 0.0) {
    if (dr1 > 0.0) { rot1 = 0; sw = 1; }
    else { rot1 = 1; sw = 0; }
  } else {
    if (dr1 > 0.0) { rot1 = 2; sw = 0; }
    else if (dr1 < 0.0) { rot1 = 2; sw = 1; }
    else return coplanar_tri_tri3(p1, q1, r1, p2, q2, r2, N1);
  }
  const d3 a1 = pick3(rot1, p1, q1, r1), b1 = pick3(rot1, q1, r1, p1), c1 = pick3(rot1, r1, p1, q1);
  const d3 a2 = p2, b2 = sw ? r2 : q2, c2 = sw ? q2 : r2;
  const double da = dp2, db = sw ? dr2 : dq2, dc = sw ? dq2 : dr2;

  // TRI_TRI_3D: swap T1's last two (s1) and rotate T2 (rot2: 0 = (a,b,c), 1 = (c,a,b), 2 = (b,c,a))
  int s1, rot2;
  if (da > 0.0) {
    if (db > 0.0) { s1 = 1; rot2 = 1; }
    else if (dc > 0.0) { s1 = 1; rot2 = 2; }
    else { s1 = 0; rot2 = 0; }
  } else if (da < 0.0) {
    if (db < 0.0) { s1 = 0; rot2 = 1; }
    else if (dc < 0.0) { s1 = 0; rot2 = 2; }
    else { s1 = 1; rot2 = 0; }
  } else if (db < 0.0) {
    if (dc >= 0.0) { s1 = 1; rot2 = 2; }
    else { s1 = 0; rot2 = 0; }
  } else if (db > 0.0) {
    if (dc > 0.0) { s1 = 1; rot2 = 0; }
    else { s1 = 0; rot2 = 2; }
  } else {
    if (dc > 0.0) { s1 = 0; rot2 = 1; }
    else if (dc < 0.0) { s1 = 1; rot2 = 1; }
    else return coplanar_tri_tri3(a1, b1, c1, a2, b2, c2, N1);
  }
  const d3 u2 = pick3(rot2, a2, c2, b2), v2 = pick3(rot2, b2, a2, c2), w2 = pick3(rot2, c2, b2, a2);
  return check_min_max(a1, s1 ? c1 : b1, s1 ? b1 : c1, u2, v2, w2);
}

// NarrowPhase.cpp:229-257 — 2D segment-segment test
__device__ __forceinline__ int edge_edge2(const d2& p0, const d2& p1, const d2& q0, const d2& q1) {
  const double eps = 2.220446049250313080847263336181640625e-16;  // DBL_EPSILON
  const d2 n = sub(q0, p0), r = sub(p1, p0), s = sub(q1, q0);
  const double rxs = r.x * s.y - r.y * s.x;
  if (__builtin_fabs(rxs) < eps) return 0;
  const double nxr = n.x * r.y - n.y * r.x;
  if (__builtin_fabs(nxr) < eps)
    return ((q0.x - p0.x < 0) != (q0.x - p1.x < 0)) || ((q0.y - p0.y < 0) != (q0.y - p1.y < 0));
  const double nxs = n.x * s.y - n.y * s.x;
  const double inv = 1.0 / rxs;
  const double t = nxs * inv, u = nxr * inv;
  return (t >= 0) && (t <= 1) && (u >= 0) && (u <= 1);
}

// Gathers of the packed vertex records (both time steps, or t1 only).
template <int D> struct VLoad;
template <> struct VLoad<3> {
  using V = d3;
  // record {x0, x1, pad}: the six values in three 16 B loads
  __device__ __forceinline__ static void load(const double* rec, int v, d3& a, d3& b) {
    const double2* r = reinterpret_cast<const double2*>(rec + (size_t)v * 8);
    const double2 r0 = r[0], r1 = r[1], r2 = r[2];
    a = {r0.x, r0.y, r1.x};
    b = {r1.y, r2.x, r2.y};
  }
  __device__ __forceinline__ static d3 load_t1(const double* rec, int v) {
    const double* r = rec + (size_t)v * 8;
    const double2 r2 = reinterpret_cast<const double2*>(r)[2];
    return {r[3], r2.x, r2.y};
  }
};
template <> struct VLoad<2> {
  using V = d2;
  __device__ __forceinline__ static void load(const double* rec, int v, d2& a, d2& b) {
    const double2* r = reinterpret_cast<const double2*>(rec + (size_t)v * 4);
    const double2 r0 = r[0], r1 = r[1];
    a = {r0.x, r0.y};
    b = {r1.x, r1.y};
  }
  __device__ __forceinline__ static d2 load_t1(const double* rec, int v) {
    const double2 r1 = reinterpret_cast<const double2*>(rec + (size_t)v * 4)[1];
    return {r1.x, r1.y};
  }
};

}  // namespace mclg
