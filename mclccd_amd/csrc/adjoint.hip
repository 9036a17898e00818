// TOI adjoint (no reference counterpart): gradient of L = sum_i w_i toi_i over
// the CCD hits of the last traverse with respect to V0 and V1.
//
// Forward model (NarrowPhase.cpp:634-711 with s = 1): with centred
// displacements p, c = 0.9 / l_p and phi(d) = (d^2 - xi^2)/(d + xi) (phi' = 1),
//   toi = sum_{k=1..K} tau_k,  tau_k = c phi(d_{k-1}),
//   d_j = D(x0 + S_j p),  S_j = tau_1 + ... + tau_j  (S_0 = 0),
// K being the number of steps added before the loop stopped (K = 1 for the
// common hit that breaks at iteration 2).  Reverse sweep: the adjoint of
// tau_K is w and G_j = G_{j+1} (1 + c grad D(x_j) . p); each d_j receives
// G_{j+1} c, which flows to x0 (directly) and to p (through S_j); l_p receives
// -sum_k G_k tau_k / l_p and passes it to the two maximising displacements.
// D's gradient is taken at the closest points the forward distance computes
// (the reference's branches, including its point_on_triangle quirk region),
// so the result is the derivative of the TOI wherever it is smooth.
#include <algorithm>
#include <cstdint>

#include "../../include/mclg.h"
#include "context.hpp"
#include "geom.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace mclg {

__device__ __forceinline__ d3 add(const d3& a, const d3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 scl(const d3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ d2 add(const d2& a, const d2& b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ d2 scl(const d2& a, double s) { return {a.x * s, a.y * s}; }

// Distance gradients.  Every closest-point branch of the reference yields the
// minimiser (s,t) of the distance over the primitive (or over the edge it
// clamps to), so the derivative is the partial at fixed (s,t) (envelope
// theorem): no chain through the projection's ratios, whose rounding noise
// would be amplified near degenerate (parallel, coplanar) configurations.
// The exception is the reference's point_on_triangle quirk branch
// (Projection.hpp:131: s = clamp(-e/c), t = 0 for s >= 0, t < 0,
// s + t >= det, a + d <= b + e), whose s is not a minimiser and is
// differentiated through.

// VF distance and its gradient (scaled by seed) w.r.t. q, a, b, c.
__device__ double dist_vf_grad(const d3& q, const d3& a, const d3& b, const d3& c, double seed, d3& gq, d3& ga,
                               d3& gb, d3& gc) {
  const d3 e0 = sub(b, a), e1 = sub(c, a), w = sub(a, q);
  const double A = dot(e0, e0), B = dot(e0, e1), C = dot(e1, e1);
  const double Dd = dot(e0, w), E = dot(e1, w);
  const double det = A * C - B * B;
  const double s0 = B * E - C * Dd, t0 = B * Dd - A * E;
  // st: 0 s=clamp(r),t=0   1 s=0,t=clamp(r)   2 s=s0 r,t=t0 r   3 s=clamp(r),t=1-s
  int st;
  bool quirk = false;
  double num, den;
  if (s0 + t0 < det) {
    if (s0 < 0) {
      if (t0 < 0 && Dd < 0) { st = 0; num = -Dd; den = A; }
      else { st = 1; num = -E; den = C; }
    } else if (t0 < 0) { st = 0; num = -Dd; den = A; }
    else { st = 2; num = 1.0; den = det; }
  } else {
    const double ee = A - 2.0 * B + C;
    if (s0 < 0) {
      if (C + E > B + Dd) { st = 3; num = C + E - B - Dd; den = ee; }
      else { st = 1; num = -E; den = C; }
    } else if (t0 < 0) {
      if (A + Dd > B + E) { st = 3; num = C + E - B - Dd; den = ee; }
      else { st = 0; num = -E; den = C; quirk = true; }
    } else { st = 3; num = C + E - B - Dd; den = ee; }
  }
  const double r = num / den;
  const double rc = r < 0 ? 0.0 : (r > 1 ? 1.0 : r);
  double s, t;
  if (st == 0) { s = rc; t = 0.0; }
  else if (st == 1) { s = 0.0; t = rc; }
  else if (st == 2) { s = s0 * r; t = t0 * r; }
  else { s = rc; t = 1.0 - s; }
  const d3 pt = add(add(a, scl(e0, s)), scl(e1, t));
  const d3 v = sub(q, pt);
  const double d = __builtin_sqrt(dot(v, v));
  if (!(d > 0.0)) return d;
  const d3 vb = scl(v, seed / d);  // d|v|/dv
  // pt = (1-s-t) a + s b + t c at fixed (s,t)
  gq = add(gq, vb);
  ga = sub(ga, scl(vb, 1.0 - s - t));
  gb = sub(gb, scl(vb, s));
  gc = sub(gc, scl(vb, t));
  if (quirk && r >= 0.0 && r <= 1.0) {
    // s = -E/C with E = e1.(a-q), C = e1.e1: add (dd/ds) ds/dx
    const double sb = -dot(e0, vb);  // dd/ds
    const double Eb = -sb / C, Cb = sb * E / (C * C);
    const d3 e1b = add(scl(e1, 2.0 * Cb), scl(w, Eb));
    const d3 wb = scl(e1, Eb);
    gc = add(gc, e1b);
    ga = add(ga, sub(wb, e1b));
    gq = sub(gq, wb);
  }
  return d;
}

// EE distance and its gradient w.r.t. p0, p1, q0, q1.
__device__ double dist_ee_grad(const d3& p0, const d3& p1, const d3& q0, const d3& q1, double seed, d3& gp0, d3& gp1,
                               d3& gq0, d3& gq1) {
  const d3 d1 = sub(p1, p0), d2 = sub(q1, q0), r = sub(p0, q0);
  const double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  const double c = dot(d1, r), b = dot(d1, d2);
  const double denom = a * e - b * b;
  double s = 0.0, t;
  if (denom != 0.0) s = clamp01((b * f - c * e) / denom);
  const double tnom = b * s + f;
  if (tnom < 0 || e == 0) { t = 0.0; s = a == 0 ? 0.0 : clamp01(-c / a); }
  else if (tnom > e) { t = 1.0; s = a == 0 ? 0.0 : clamp01((b - c) / a); }
  else { t = tnom / e; }
  const d3 v = sub(add(q0, scl(d2, t)), add(p0, scl(d1, s)));
  const double dist = __builtin_sqrt(dot(v, v));
  if (!(dist > 0.0)) return dist;
  const d3 vb = scl(v, seed / dist);
  gq0 = add(gq0, scl(vb, 1.0 - t));
  gq1 = add(gq1, scl(vb, t));
  gp0 = sub(gp0, scl(vb, 1.0 - s));
  gp1 = sub(gp1, scl(vb, s));
  return dist;
}

// 2D point-edge distance and its gradient w.r.t. p, e0, e1.
__device__ double dist_pe2_grad(const d2& p, const d2& e0, const d2& e1, double seed, d2& gp, d2& ge0, d2& ge1) {
  const d2 e = sub(e1, e0);
  const double len2 = dot(e, e);
  double t = 0.0;  // closest point e0 + t e
  d2 pt = e0;
  if (!(len2 <= 0.0)) {
    const d2 pe = sub(p, e0);
    const double u = dot(pe, e) / len2;
    if (u > 1.0) { t = 1.0; pt = e1; }
    else if (!(u < 0.0)) { t = u; pt = {e0.x + u * e.x, e0.y + u * e.y}; }
  }
  const d2 v = sub(p, pt);
  const double d = __builtin_sqrt(dot(v, v));
  if (!(d > 0.0)) return d;
  const d2 vb = scl(v, seed / d);
  gp = add(gp, vb);
  ge0 = sub(ge0, scl(vb, 1.0 - t));
  ge1 = sub(ge1, scl(vb, t));
  return d;
}

template <int D> struct AdjV;
template <> struct AdjV<3> { using V = d3; };
template <> struct AdjV<2> { using V = d2; };

template <int D, bool VF>
__device__ __forceinline__ double dist_g(const typename AdjV<D>::V* x, double seed, typename AdjV<D>::V* g) {
  if constexpr (D == 3) {
    if (VF) return dist_vf_grad(x[0], x[1], x[2], x[3], seed, g[0], g[1], g[2], g[3]);
    return dist_ee_grad(x[0], x[1], x[2], x[3], seed, g[0], g[1], g[2], g[3]);
  } else {
    return dist_pe2_grad(x[0], x[1], x[2], seed, g[0], g[1], g[2]);
  }
}

constexpr int ADJ_T = 256;     // threads per block
constexpr int ADJ_STORE = 8;   // step sums S_j kept in LDS (j < ADJ_STORE); later ones are replayed

// One hit: accumulates w * d toi / d(V0, V1) into gV0 / gV1 (atomics).
// S points to this thread's LDS column (stride ADJ_T).  Every per-vertex array
// is indexed by unrolled constants only, so nothing lives in scratch memory.
template <int D, bool VF>
__device__ void toi_adjoint(const int* sv, const double* rec, double xsi, double w, double* S, bool skip_k1,
                            double* gV0, double* gV1) {
  using V = typename AdjV<D>::V;
  constexpr int ns = D + 1;
  V v0[ns];
  Stencil<D, VF, double> st;
  double l_p;
  {
    V v1[ns];
#pragma unroll
    for (int i = 0; i < ns; ++i) VLoad<D>::load(rec, sv[i], v0[i], v1[i]);
    bool fin;
    l_p = st.setup(v0, v1, fin);
  }
  if (!(l_p > 0)) return;
  // maximisers of l_p = max_first |p_i| + max_second |p_i| (first maximum wins)
  int i_first = -1, i_second = -1;
  {
    double lf = 0.0, ls = 0.0;
#pragma unroll
    for (int i = 0; i < ns; ++i) {
      const double n = __builtin_sqrt(dot(st.p[i], st.p[i]));
      const bool first = VF ? (i == 0) : (i <= 1);
      if (first) { if (lf < n) { lf = n; i_first = i; } }
      else if (ls < n) { ls = n; i_second = i; }
    }
  }
  const double d0 = st.dist();
  if (d0 <= xsi) return;  // toi = 0 whatever the inputs nearby
  const double c = 0.9 / l_p;
  // forward: count the steps K summed into toi, keep S_j (j < ADJ_STORE)
  const double g = (d0 * d0 - xsi * xsi) / (d0 - xsi);
  const double tl1 = 0.9 * (d0 * d0 - xsi * xsi) / ((d0 + xsi) * l_p);
  double toi = 0.0, tl = tl1;
  S[0] = 0.0;
  int K = 0;  // a loop that reaches max_iter sums steps 1..999 (K = 999)
  for (int iter = 1; iter < 1000; ++iter) {
    const bool can_break = toi > 0.0;
    if (!can_break && toi + tl > 1.0) return;
    st.step(tl);
    const double d = st.dist();
    if (can_break && (d * d - xsi * xsi) / (d + xsi) < g) break;
    toi += tl;
    ++K;
    if (K < ADJ_STORE) S[K * ADJ_T] = S[(K - 1) * ADJ_T] + tl;
    if (toi > 1.0) return;
    tl = 0.9 * (d * d - xsi * xsi) / ((d + xsi) * l_p);
  }
  if (skip_k1 && K == 1) return;  // done by the one-step kernel
  // reverse sweep
  V gx0[ns], gp[ns];
#pragma unroll
  for (int i = 0; i < ns; ++i) { gx0[i] = {}; gp[i] = {}; }
  double G = w, glp = 0.0;
  for (int j = K - 1; j >= 0; --j) {
    double Sj;
    if (j < ADJ_STORE) {
      Sj = S[j * ADJ_T];
    } else {  // long tails: replay the steps up to j from x0
#pragma unroll
      for (int i = 0; i < ns; ++i) st.x[i] = v0[i];
      double tt = tl1;
      Sj = 0.0;
      for (int m = 1; m <= j; ++m) {
        Sj += tt;
        st.step(tt);
        const double dm = st.dist();
        tt = 0.9 * (dm * dm - xsi * xsi) / ((dm + xsi) * l_p);
      }
    }
    V xj[ns], gd[ns];
#pragma unroll
    for (int i = 0; i < ns; ++i) {
      xj[i] = add(v0[i], scl(st.p[i], Sj));
      gd[i] = {};
    }
    const double dj = dist_g<D, VF>(xj, 1.0, gd);
    const double tau = c * ((dj * dj - xsi * xsi) / (dj + xsi));
    const double dbar = G * c;
    double gdotp = 0.0;
#pragma unroll
    for (int i = 0; i < ns; ++i) {
      gx0[i] = add(gx0[i], scl(gd[i], dbar));
      gp[i] = add(gp[i], scl(gd[i], dbar * Sj));
      gdotp += dot(gd[i], st.p[i]);
    }
    glp -= G * tau / l_p;
    if (j >= 1) G = G * (1.0 + c * gdotp);
  }
  // l_p = |p_first| + |p_second| at the maximisers; then the centring
  // p_i = P_i - mean(P): gP_i = gp_i - mean(gp), and V1 = V0 + P
  V mean = {};
#pragma unroll
  for (int i = 0; i < ns; ++i) {
    if (i == i_first || i == i_second) {
      const double n = __builtin_sqrt(dot(st.p[i], st.p[i]));
      if (n > 0) gp[i] = add(gp[i], scl(st.p[i], glp / n));
    }
    mean = add(mean, gp[i]);
  }
  mean = scl(mean, 1.0 / ns);
#pragma unroll
  for (int i = 0; i < ns; ++i) {
    const V gP = sub(gp[i], mean);
    const V g0 = sub(gx0[i], gP);
    double* o0 = gV0 + (size_t)sv[i] * D;
    double* o1 = gV1 + (size_t)sv[i] * D;
    atomicAdd(o0 + 0, g0.x);
    atomicAdd(o0 + 1, g0.y);
    atomicAdd(o1 + 0, gP.x);
    atomicAdd(o1 + 1, gP.y);
    if constexpr (D == 3) {
      atomicAdd(o0 + 2, g0.z);
      atomicAdd(o1 + 2, gP.z);
    }
  }
}

// Per-block LDS accumulator: open-addressing table vertex -> (dL/dV0, dL/dV1)
// summed with LDS atomics; flushed with one global atomic per component and
// distinct vertex.  A crowded table sends the contribution to global memory.
constexpr int ADJ_TAB_BITS = 10, ADJ_TAB = 1 << ADJ_TAB_BITS, ADJ_PROBES = 16;
constexpr int ADJ_R = 16;  // hits per thread per block of the one-step kernel

struct AdjTable {
  int* key;     // ADJ_TAB entries, -1 = empty
  double* acc;  // 2*D components, component-major (acc[c*ADJ_TAB + slot])
};

template <int D>
__device__ __forceinline__ void adj_accum(const AdjTable& tab, int v, const double* val, double* gV0, double* gV1) {
  unsigned h = ((unsigned)v * 2654435761u) >> (32 - ADJ_TAB_BITS);
  for (int probe = 0; probe < ADJ_PROBES; ++probe) {
    // a slot's key never changes once set: read it first (a broadcast when
    // many lanes look up one vertex), claim an empty slot with a CAS
    int k = __hip_atomic_load(&tab.key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == -1) k = atomicCAS(&tab.key[h], -1, v);
    if (k == -1 || k == v) {
#pragma unroll
      for (int c = 0; c < 2 * D; ++c) atomicAdd(&tab.acc[c * ADJ_TAB + h], val[c]);
      return;
    }
    h = (h + 1) & (ADJ_TAB - 1);
  }
#pragma unroll
  for (int c = 0; c < D; ++c) {
    atomicAdd(gV0 + (size_t)v * D + c, val[c]);
    atomicAdd(gV1 + (size_t)v * D + c, val[D + c]);
  }
}

// One-step hits (the loop breaks at its second iteration: K = 1, nearly every
// hit): toi = tau_1 = c phi(d0), so the gradient needs D's gradient at x0
// only and no step sums.  Returns false, adding nothing, when the hit takes
// more steps; hits that are not CCD hits add nothing and return true.
template <int D, bool VF>
__device__ bool toi_adjoint_k1(const int* sv, const double* rec, double xsi, double w, const AdjTable& tab,
                               double* gV0, double* gV1) {
  using V = typename AdjV<D>::V;
  constexpr int ns = D + 1;
  Stencil<D, VF, double> st;
  double l_p;
  {
    V v0[ns], v1[ns];
#pragma unroll
    for (int i = 0; i < ns; ++i) VLoad<D>::load(rec, sv[i], v0[i], v1[i]);
    bool fin;
    l_p = st.setup(v0, v1, fin);
  }
  if (!(l_p > 0)) return true;
  const double d0 = st.dist();
  if (d0 <= xsi) return true;
  const double g = (d0 * d0 - xsi * xsi) / (d0 - xsi);
  const double tl1 = 0.9 * (d0 * d0 - xsi * xsi) / ((d0 + xsi) * l_p);
  if (tl1 > 1.0) return true;  // iteration 1: toi + t_l > 1 with toi = 0
  st.step(tl1);
  const double d1 = st.dist();
  if (!(tl1 > 0.0)) return false;  // the break test cannot fire at iteration 2
  const double tl2 = 0.9 * (d1 * d1 - xsi * xsi) / ((d1 + xsi) * l_p);
  st.step(tl2);
  const double d2 = st.dist();
  if (!((d2 * d2 - xsi * xsi) / (d2 + xsi) < g)) return false;
  // reverse: dtoi/dx0_i = c grad_i D(x0); dtoi/dl_p = -toi / l_p to the maximisers
  const double c = 0.9 / l_p;
  V v0[ns], gd[ns];  // x0 is reloaded (cached) rather than kept live through the forward steps
#pragma unroll
  for (int i = 0; i < ns; ++i) {
    V unused;
    VLoad<D>::load(rec, sv[i], v0[i], unused);
    gd[i] = {};
  }
  const double dd = dist_g<D, VF>(v0, w * c, gd);
  const double glp = -w * (c * ((dd * dd - xsi * xsi) / (dd + xsi))) / l_p;
  V gp[ns], mean = {};
  {
    double lf = 0.0, ls = 0.0;
    int i_first = -1, i_second = -1;
#pragma unroll
    for (int i = 0; i < ns; ++i) {
      const double n = __builtin_sqrt(dot(st.p[i], st.p[i]));
      const bool first = VF ? (i == 0) : (i <= 1);
      if (first) { if (lf < n) { lf = n; i_first = i; } }
      else if (ls < n) { ls = n; i_second = i; }
    }
#pragma unroll
    for (int i = 0; i < ns; ++i) {
      gp[i] = {};
      if (i == i_first || i == i_second) {
        const double n = __builtin_sqrt(dot(st.p[i], st.p[i]));
        if (n > 0) gp[i] = scl(st.p[i], glp / n);
      }
      mean = add(mean, gp[i]);
    }
  }
  mean = scl(mean, 1.0 / ns);
#pragma unroll
  for (int i = 0; i < ns; ++i) {
    const V gP = sub(gp[i], mean);
    const V g0 = sub(gd[i], gP);
    double val[2 * D];
    val[0] = g0.x;
    val[1] = g0.y;
    val[D] = gP.x;
    val[D + 1] = gP.y;
    if constexpr (D == 3) {
      val[2] = g0.z;
      val[5] = gP.z;
    }
    adj_accum<D>(tab, sv[i], val, gV0, gV1);
  }
  return true;
}

// One-step kernel, one launch per stencil kind.  Entries k in [0, nh) are hit
// order[k] (or hit k when order is null); skey[k] (sorted (kind, first
// vertex) keys, kind at bit kind_shift) lets blocks of the other kind leave at
// once.  Longer hits are appended to queue (at most qcap; *qn counts all).
// Three waves per SIMD: the vertex-face kernel needs 172 VGPRs unconstrained
// (two waves); capped at 168 it spills 8 bytes and the adjoint drops from ~78
// to ~75 ms at 8.3e8 hits.  The LDS table allows three blocks per CU either way.
template <int D, bool VF>
__global__ __launch_bounds__(ADJ_T) __attribute__((amdgpu_waves_per_eu(3))) void k_toi_adjoint_k1(const int4* __restrict__ sten,
                                                          const uint8_t* __restrict__ is_vf,
                                                          const double* __restrict__ w,
                                                          const uint32_t* __restrict__ order,
                                                          const uint32_t* __restrict__ skey, int kind_shift,
                                                          long long nh, const double* __restrict__ rec, double eta,
                                                          double* gV0, double* gV1, long long* queue, long long qcap,
                                                          unsigned long long* qn) {
  __shared__ int s_key[ADJ_TAB];
  __shared__ double s_acc[2 * D * ADJ_TAB];
  const uint32_t my_kind = (D == 3 && !VF) ? 1u : 0u;
  const long long base = blockIdx.x * (long long)(ADJ_T * ADJ_R);
  if (skey) {  // block-uniform: keys are sorted by kind first
    const long long last = (base + ADJ_T * ADJ_R < nh ? base + ADJ_T * ADJ_R : nh) - 1;
    const uint32_t k0 = skey[base] >> kind_shift, k1 = skey[last] >> kind_shift;
    if (k0 > my_kind || k1 < my_kind) return;
  }
  for (int j = threadIdx.x; j < ADJ_TAB; j += ADJ_T) {
    s_key[j] = -1;
#pragma unroll
    for (int c = 0; c < 2 * D; ++c) s_acc[c * ADJ_TAB + j] = 0.0;
  }
  __syncthreads();
  const AdjTable tab{s_key, s_acc};
  for (int r = 0; r < ADJ_R; ++r) {
    const long long k = base + r * ADJ_T + threadIdx.x;
    bool more = false;
    long long i = k;
    if (k < nh) {
      bool mine;
      if (skey) {
        i = order[k];
        mine = (skey[k] >> kind_shift) == my_kind;
      } else {
        mine = D == 2 || (is_vf[i] != 0) == VF;
      }
      if (mine) {
        const double wi = w[i];
        if (wi != 0.0) {
          const int4 s4 = sten[i];
          const int sv[4] = {s4.x, s4.y, s4.z, s4.w};
          more = !toi_adjoint_k1<D, VF>(sv, rec, eta, wi, tab, gV0, gV1);
        }
      }
    }
    const long long slot = wave_append(more, qn);
    if (more && slot < qcap) queue[slot] = i;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < ADJ_TAB; j += ADJ_T) {
    const int v = s_key[j];
    if (v < 0) continue;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const double a0 = s_acc[c * ADJ_TAB + j], a1 = s_acc[(D + c) * ADJ_TAB + j];
      if (a0 != 0.0) atomicAdd(gV0 + (size_t)v * D + c, a0);
      if (a1 != 0.0) atomicAdd(gV1 + (size_t)v * D + c, a1);
    }
  }
}

// Sort keys of the hits: (kind, cell of the first vertex) with kind = 1 for
// EE, 0 else, at bit D * ADJ_CELL_BITS<D>; the cell is the Morton code of the
// vertex's t0 position on a 2^ADJ_CELL_BITS grid per axis over the bounds of
// the leaf-box centres (cbounds, ordered encoding), clamped.
template <int D> constexpr int adj_cell_bits() { return D == 3 ? 5 : 7; }

template <int D>
__global__ void k_adj_keys(const int4* __restrict__ sten, const uint8_t* __restrict__ is_vf, long long nh,
                           const double* __restrict__ rec, const unsigned long long* __restrict__ cbounds,
                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  constexpr int B = adj_cell_bits<D>();
  constexpr double cells = (double)(1 << B);
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nh) return;
  const uint32_t kind = (D == 3 && !is_vf[i]) ? 1u : 0u;
  const double* r = rec + (size_t)sten[i].x * VRec<D>::stride;
  uint32_t q[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double lo = ord_to_dbl(cbounds[k]), hi = ord_to_dbl(cbounds[D + k]);
    const double ext = hi - lo;
    double u = ext > 0 ? (r[k] - lo) / ext * cells : 0.0;
    u = u < 0 ? 0.0 : (u > cells - 1 ? cells - 1 : u);  // NaN -> cells - 1
    q[k] = (uint32_t)u;
  }
  uint32_t code = 0;
#pragma unroll
  for (int b = B - 1; b >= 0; --b)
#pragma unroll
    for (int k = 0; k < D; ++k) code = (code << 1) | ((q[k] >> b) & 1u);
  keys[i] = (kind << (D * B)) | code;
  vals[i] = (uint32_t)i;
}

// Multi-step hits: over the queue (queue != nullptr, n entries), or over all
// hits skipping the one-step ones (queue overflowed).
template <int D, bool VF>
__global__ __launch_bounds__(ADJ_T) void k_toi_adjoint(const int4* __restrict__ sten,
                                                       const uint8_t* __restrict__ is_vf,
                                                       const double* __restrict__ w, const long long* queue,
                                                       long long n, const double* __restrict__ rec, double eta,
                                                       double* gV0, double* gV1) {
  __shared__ double s_S[ADJ_STORE * ADJ_T];
  const long long k = blockIdx.x * (long long)ADJ_T + threadIdx.x;
  if (k >= n) return;
  const long long i = queue ? queue[k] : k;
  if (D == 3 && (is_vf[i] != 0) != VF) return;
  const double wi = w[i];
  if (wi == 0.0) return;
  const int4 s4 = sten[i];
  const int sv[4] = {s4.x, s4.y, s4.z, s4.w};
  toi_adjoint<D, VF>(sv, rec, eta, wi, s_S + threadIdx.x, queue == nullptr, gV0, gV1);
}

// Device scratch of one adjoint call: the visiting order (radix sort of the
// hits by (kind, first vertex); keys null: hits in list order) and the queue
// of multi-step hits with its counter.
struct AdjScratch {
  uint32_t *keys = nullptr, *vals = nullptr, *ktmp = nullptr, *vtmp = nullptr, *hist = nullptr;
  long long* queue = nullptr;
  long long qcap = 0;
  unsigned long long* qn = nullptr;
};

template <int D, bool VF>
static void launch_adjoint(Ctx* c, const double* gw, const uint32_t* order, const uint32_t* skey, int kind_shift,
                           long long nh, double eta, double* g0, double* g1, const AdjScratch& S) {
  const unsigned g = (unsigned)((nh + ADJ_T * ADJ_R - 1) / (ADJ_T * ADJ_R));
  k_toi_adjoint_k1<D, VF><<<g, ADJ_T, 0, c->stream>>>(c->d_hit_sten, c->d_hit_vf, gw, order, skey, kind_shift, nh,
                                                      c->d_rec, eta, g0, g1, S.queue, S.qcap, S.qn);
}

// The adjoint of the last traverse's hits into device gradients g0, g1
// (dim*nv each, zeroed here) from device weights gw; S.keys null: no sort.
// Synchronises the context's stream (reads the multi-step count).
static int toi_grad_run(Ctx* c, const double* gw, double* g0, double* g1, const AdjScratch& S) {
  const long long nh = c->counts.hits;
  launch_fill_u64(reinterpret_cast<unsigned long long*>(g0), (long long)c->dim * c->nv, 0ull, c->stream);
  launch_fill_u64(reinterpret_cast<unsigned long long*>(g1), (long long)c->dim * c->nv, 0ull, c->stream);
  c->adj_multi = 0;
  if (nh == 0) return MCLG_OK;
  launch_fill_u64(S.qn, 1, 0ull, c->stream);
  // Visit hits grouped by (kind, spatial cell of the first vertex): a 16-bit
  // key (two radix passes), 32^3 cells in 3D.  The sort is stable, so within
  // a cell the hits keep the hit list's order, in which the hits of one leaf
  // pair sit together, and a block's LDS table sees the same few vertices
  // again and again.  Measured at 1M triangles: 81 ms, against 92 ms for
  // groups of 32 consecutive first-vertex indices and 152 ms for the full
  // (kind, first vertex) order, whose blocks see one vertex with all of its
  // partners.  The caller's hit order is untouched.
  const int kind_shift = c->dim * (c->dim == 3 ? adj_cell_bits<3>() : adj_cell_bits<2>());
  const uint32_t *order = nullptr, *skey = nullptr;
  if (S.keys) {
    const unsigned gk = (unsigned)((nh + 255) / 256);
    if (c->dim == 3)
      k_adj_keys<3><<<gk, 256, 0, c->stream>>>(c->d_hit_sten, c->d_hit_vf, nh, c->d_rec, c->d_cbounds, S.keys, S.vals);
    else
      k_adj_keys<2><<<gk, 256, 0, c->stream>>>(c->d_hit_sten, c->d_hit_vf, nh, c->d_rec, c->d_cbounds, S.keys, S.vals);
    launch_radix_sort(S.keys, S.vals, S.ktmp, S.vtmp, S.hist, (int)nh, kind_shift + 1, c->stream);
    order = S.vals;
    skey = S.keys;
  }
  if (c->dim == 3) {
    launch_adjoint<3, true>(c, gw, order, skey, kind_shift, nh, c->opt.vf_ccd_eta, g0, g1, S);
    launch_adjoint<3, false>(c, gw, order, skey, kind_shift, nh, c->opt.ee_ccd_eta, g0, g1, S);
  } else {
    launch_adjoint<2, true>(c, gw, order, skey, kind_shift, nh, c->opt.vf_ccd_eta, g0, g1, S);
  }
  if (hipGetLastError() != hipSuccess) return c->fail(MCLG_ERR_HIP, "toi_grad: kernel launch failed");
  // the context's pinned counter mirror receives the multi-step count
  if (hipMemcpyAsync(c->h_counters, S.qn, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return c->fail(MCLG_ERR_HIP, "toi_grad: reading the multi-step count failed");
  const unsigned long long nq = c->h_counters[0];
  if (nq > 0) {
    const bool all = nq > (unsigned long long)S.qcap;  // queue overflow: rescan every hit
    const long long n = all ? nh : (long long)nq;
    const long long* q = all ? nullptr : S.queue;
    const unsigned g = (unsigned)((n + ADJ_T - 1) / ADJ_T);
    if (c->dim == 3) {
      k_toi_adjoint<3, true><<<g, ADJ_T, 0, c->stream>>>(c->d_hit_sten, c->d_hit_vf, gw, q, n, c->d_rec,
                                                         c->opt.vf_ccd_eta, g0, g1);
      k_toi_adjoint<3, false><<<g, ADJ_T, 0, c->stream>>>(c->d_hit_sten, c->d_hit_vf, gw, q, n, c->d_rec,
                                                          c->opt.ee_ccd_eta, g0, g1);
    } else {
      k_toi_adjoint<2, true><<<g, ADJ_T, 0, c->stream>>>(c->d_hit_sten, c->d_hit_vf, gw, q, n, c->d_rec,
                                                         c->opt.vf_ccd_eta, g0, g1);
    }
    if (hipGetLastError() != hipSuccess) return c->fail(MCLG_ERR_HIP, "toi_grad: kernel launch failed");
  }
  c->adj_multi = (long long)nq;
  return MCLG_OK;
}

// Preconditions shared by both entry points: hits of an ACCD traverse on the
// current tree.
static int toi_grad_check(Ctx* c) {
  if (c->hits_one_sided)
    return c->fail(MCLG_ERR_UNSUPPORTED, "toi_grad: the TOI adjoint covers ACCD only, not the vf_one_sided CTCD path");
  if (c->hits_f32)
    return c->fail(MCLG_ERR_UNSUPPORTED, "toi_grad: the TOI adjoint is FP64; the last traverse ran in float precision");
  if (c->hits_gen != c->tree_gen)
    return c->fail(MCLG_ERR_STATE,
                   "toi_grad: no traverse since the last update or refit (the hits would be paired with the new "
                   "vertex positions); traverse again first");
  return MCLG_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free scratch: one static array per GPU (each device holds its own
// copy for the life of the process).  Host code must use the address from
// hipGetSymbolAddress, never the array's name (a host-side stand-in).
// ---------------------------------------------------------------------------
constexpr long long kNoAllocHits = 1ll << 20;  // MCLG_NOALLOC_MAX_HITS
static_assert(kNoAllocHits == MCLG_NOALLOC_MAX_HITS, "header and scratch agree");
constexpr size_t kNaKeys = sizeof(uint32_t) * kNoAllocHits;  // keys, vals, ktmp, vtmp: one each
constexpr size_t kNaHist = sizeof(uint32_t) * ((256 * ((kNoAllocHits + 4095) / 4096) * 2 + 255) / 256 * 256);
constexpr size_t kNaQueue = sizeof(long long) * kNoAllocHits;  // multi-step queue: room for every hit
constexpr size_t kNaBytes = 4 * kNaKeys + kNaHist + kNaQueue + 256;

__device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {};

// The scratch's device address on the current device (looked up once per device).
static int noalloc_scratch(Ctx* c, AdjScratch& S) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != c->device)
    return c->fail(MCLG_ERR_STATE, "toi_grad_noalloc: the context's device %d is not current", c->device);
  if (dev < 0 || dev >= kMaxDevices) return c->fail(MCLG_ERR_UNSUPPORTED, "toi_grad_noalloc: device %d", dev);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return c->fail(MCLG_ERR_HIP, "toi_grad_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = p;
  }
  unsigned char* base = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  if (radix_hist_words((int)kNoAllocHits) * sizeof(uint32_t) > kNaHist)
    return c->fail(MCLG_ERR_STATE, "toi_grad_noalloc: histogram scratch too small");
  S.keys = reinterpret_cast<uint32_t*>(base);
  S.vals = reinterpret_cast<uint32_t*>(base + kNaKeys);
  S.ktmp = reinterpret_cast<uint32_t*>(base + 2 * kNaKeys);
  S.vtmp = reinterpret_cast<uint32_t*>(base + 3 * kNaKeys);
  S.hist = reinterpret_cast<uint32_t*>(base + 4 * kNaKeys);
  S.queue = reinterpret_cast<long long*>(base + 4 * kNaKeys + kNaHist);
  S.qcap = kNoAllocHits;
  S.qn = reinterpret_cast<unsigned long long*>(base + 4 * kNaKeys + kNaHist + kNaQueue);
  return MCLG_OK;
}

}  // namespace mclg

using namespace mclg;

#define CHECK(x)                     \
  do {                               \
    int _rc = (x);                   \
    if (_rc != MCLG_OK) return _rc;  \
  } while (0)
#define HIPCHECK(ctx, call) CHECK((ctx)->hip((call), #call))

extern "C" int mclg_toi_grad(mclg_ctx* h, const double* w, double* dV0, double* dV1, int device_ptrs) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !w || !dV0 || !dV1) return MCLG_ERR_ARG;
  CHECK(toi_grad_check(c));
  const long long nh = c->counts.hits;
  const size_t gbytes = sizeof(double) * c->dim * (size_t)c->nv;
  double *gw = nullptr, *g0 = nullptr, *g1 = nullptr;
  if (device_ptrs) {
    gw = const_cast<double*>(w);
    g0 = dV0;
    g1 = dV1;
  } else {
    CHECK(c->grow((void**)&c->d_adj_w, &c->cap_adj_w, sizeof(double) * (size_t)std::max(nh, 1LL), "adjoint"));
    CHECK(c->grow((void**)&c->d_adj_g, &c->cap_adj_g, 2 * gbytes, "adjoint"));
    gw = c->d_adj_w;
    g0 = c->d_adj_g;
    g1 = c->d_adj_g + (size_t)c->dim * c->nv;
    if (nh > 0) HIPCHECK(c, hipMemcpyAsync(gw, w, sizeof(double) * nh, hipMemcpyHostToDevice, c->stream));
  }
  AdjScratch S;
  if (nh > 0) {
    CHECK(c->grow((void**)&c->d_adj_n, &c->cap_adj_n, sizeof(unsigned long long), "adjoint"));
    if (c->cap_adj_q == 0) {
      // queue of multi-step hits (rare); an overflow rescans all hits instead
      CHECK(c->grow((void**)&c->d_adj_q, &c->cap_adj_q_bytes, sizeof(long long) * kAdjQueue, "adjoint"));
      c->cap_adj_q = kAdjQueue;
    }
    S.queue = c->d_adj_q;
    S.qcap = (long long)c->cap_adj_q;
    S.qn = c->d_adj_n;
    if (nh <= (long long)INT32_MAX) {  // the visiting order holds 32-bit hit indices
      const size_t b4 = sizeof(uint32_t) * (size_t)nh;
      CHECK(c->grow((void**)&c->d_adj_keys, &c->cap_adj_keys, b4, "adjoint"));
      CHECK(c->grow((void**)&c->d_adj_vals, &c->cap_adj_vals, b4, "adjoint"));
      CHECK(c->grow((void**)&c->d_adj_ktmp, &c->cap_adj_ktmp, b4, "adjoint"));
      CHECK(c->grow((void**)&c->d_adj_vtmp, &c->cap_adj_vtmp, b4, "adjoint"));
      CHECK(c->grow((void**)&c->d_adj_hist, &c->cap_adj_hist, sizeof(uint32_t) * radix_hist_words((int)nh),
                    "adjoint"));
      S.keys = c->d_adj_keys; S.vals = c->d_adj_vals; S.ktmp = c->d_adj_ktmp; S.vtmp = c->d_adj_vtmp;
      S.hist = c->d_adj_hist;
    }
  }
  c->timing = true;
  c->mark(14);
  CHECK(toi_grad_run(c, gw, g0, g1, S));
  c->mark(15);
  if (!device_ptrs) {
    HIPCHECK(c, hipMemcpyAsync(dV0, g0, gbytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(dV1, g1, gbytes, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c->ev[14], c->ev[15]) == hipSuccess) c->times[10] = ms;
  return MCLG_OK;
}

// The same adjoint with every scratch buffer in the static array above: no
// device or host allocation.  Device pointers only.
extern "C" int mclg_toi_grad_noalloc(mclg_ctx* h, const double* w, double* dV0, double* dV1) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !w || !dV0 || !dV1) return MCLG_ERR_ARG;
  CHECK(toi_grad_check(c));
  const long long nh = c->counts.hits;
  if (nh > kNoAllocHits)
    return c->fail(MCLG_ERR_CAPACITY,
                   "toi_grad_noalloc: %lld hits exceed the static scratch's maximum of %lld; use mclg_toi_grad",
                   nh, kNoAllocHits);
  AdjScratch S;
  CHECK(noalloc_scratch(c, S));
  c->timing = true;
  c->mark(14);
  CHECK(toi_grad_run(c, w, dV0, dV1, S));
  c->mark(15);
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c->ev[14], c->ev[15]) == hipSuccess) c->times[10] = ms;
  return MCLG_OK;
}
