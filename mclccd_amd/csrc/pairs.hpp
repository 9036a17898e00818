// Device helpers shared by the kernels that read the leaf-pair list: the two
// faces of a pair, the reference's representative-triangle candidate rule,
// V1 gathers for the discrete test, and per-wave LDS staging of outputs.
#pragma once
#include "geom.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace mclg {

// int4 as a clang vector: the operand type of the non-temporal builtins
typedef int nt_int4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Per-wave LDS staging of output records: lanes deposit records at
// ballot-computed slots; a full buffer is written out with one global atomic.
// Called with the whole wave converged (pred carries per-lane validity).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void stage_flush(T* buf, int cnt, T* __restrict__ dst, unsigned long long cap,
                                            unsigned long long* counter) {
  wave_lds_sync();
  unsigned long long base = 0;
  const int lane = lane_id();
  if (lane == 0) base = atomicAdd(counter, (unsigned long long)cnt);
  base = __shfl(base, 0);
  for (int i = lane; i < cnt; i += kWave)
    if (base + i < cap) dst[base + i] = buf[i];
  wave_lds_sync();
}

template <typename T, int CAP>
__device__ __forceinline__ void stage_push(T* buf, int& cnt, bool pred, const T& v, T* __restrict__ dst,
                                           unsigned long long cap, unsigned long long* counter) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return;
  if (pred) buf[cnt + popc_below(m)] = v;
  cnt += __popcll(m);
  if (cnt > CAP - kWave) {
    stage_flush(buf, cnt, dst, cap, counter);
    cnt = 0;
  }
}

// A face's vertex ids and reptri flags (bits v0..v2, e0..e2) from its head.
template <int D> struct FaceIds {
  int p[D];
  uint32_t flags;
};
template <int D> __device__ __forceinline__ FaceIds<D> face_ids(int4 h) {
  FaceIds<D> f;
  f.p[0] = h.x;
  f.p[1] = h.y;
  if constexpr (D == 3) {
    f.p[2] = h.z;
    f.flags = (uint32_t)h.w;
  } else {
    f.flags = (uint32_t)h.z;
  }
  return f;
}

// The pair's two faces, lower rank first (BVHTree.cpp:252-283 visits each leaf
// pair once in the KdBVH's order; the rank reproduces that orientation); pa,
// pb receive their sorted positions.
template <int D>
__device__ __forceinline__ void pair_faces(const FaceArrays& F, const int2 sp, FaceIds<D>& fa, FaceIds<D>& fb, int& pa,
                                           int& pb) {
  const int4 x = F.head[sp.x], y = F.head[sp.y];
  const int rx = F.rank[sp.x], ry = F.rank[sp.y];
  const bool keep = rx < ry;
  fa = face_ids<D>(keep ? x : y);
  fb = face_ids<D>(keep ? y : x);
  pa = keep ? sp.x : sp.y;
  pb = keep ? sp.y : sp.x;
}
template <int D>
__device__ __forceinline__ void load_pair_faces(const int2* __restrict__ pairs, const FaceArrays& F, long long i,
                                                FaceIds<D>& fa, FaceIds<D>& fb, int& pa, int& pb) {
  pair_faces<D>(F, pairs[i], fa, fb, pa, pb);
}

// BVHTree.cpp:327-397: VF f0->f1 (codes 0..2), f1->f0 (3..5), EE (3i + j)
template <int D>
__device__ __forceinline__ void candidate_masks(const FaceIds<D>& ha, const FaceIds<D>& hb, uint32_t& vfm,
                                                uint32_t& eem) {
  vfm = 0;
  eem = 0;
#pragma unroll
  for (int x = 0; x < D; ++x) {
    bool sh0 = false, sh1 = false;
#pragma unroll
    for (int y = 0; y < D; ++y) { sh0 |= ha.p[x] == hb.p[y]; sh1 |= hb.p[x] == ha.p[y]; }
    if (((ha.flags >> x) & 1u) && !sh0) vfm |= 1u << x;
    if (((hb.flags >> x) & 1u) && !sh1) vfm |= 1u << (3 + x);
  }
  if constexpr (D == 3) {
#pragma unroll
    for (int x = 0; x < 3; ++x) {
#pragma unroll
      for (int y = 0; y < 3; ++y) {
        const int a0 = ha.p[x], a1 = ha.p[(x + 1) % 3], b0 = hb.p[y], b1 = hb.p[(y + 1) % 3];
        const bool ok = ((ha.flags >> (3 + x)) & 1u) && ((hb.flags >> (3 + y)) & 1u) && a0 != b0 && a0 != b1 &&
                        a1 != b0 && a1 != b1;
        if (ok) eem |= 1u << (3 * x + y);
      }
    }
  }
}

// The stencil of candidate code c of a pair (faces a = (a0,a1,a2), b = (b0,b1,b2)):
// VF c < 3: vertex c of a against face b; 3 <= c < 6: vertex c-3 of b against a
// (BVHTree.cpp:328-369, -1 padding in 2D); EE c = 3x + y: edge x of a against
// edge y of b (BVHTree.cpp:372-395).  Plain values: indexing an array, or
// selecting between two, would place them in scratch memory.
template <int D, bool VF>
__device__ __forceinline__ int4 candidate_stencil(int c, int a0, int a1, int a2, int b0, int b1, int b2) {
  int4 st;
  if constexpr (VF) {
    const bool fwd = c < 3;
    const int k = fwd ? c : c - 3;
    st.x = fwd ? (k == 0 ? a0 : (k == 1 ? a1 : a2)) : (k == 0 ? b0 : (k == 1 ? b1 : b2));
    st.y = fwd ? b0 : a0;
    st.z = fwd ? b1 : a1;
    st.w = D == 3 ? (fwd ? b2 : a2) : -1;
  } else {
    const int x = c / 3, y = c - 3 * (c / 3);
    st.x = x == 0 ? a0 : (x == 1 ? a1 : a2);
    st.y = x == 0 ? a1 : (x == 1 ? a2 : a0);
    st.z = y == 0 ? b0 : (y == 1 ? b1 : b2);
    st.w = y == 0 ? b1 : (y == 1 ? b2 : b0);
  }
  return st;
}

// V1 of a face's vertices (from the per-vertex V1 array).
template <int D> __device__ __forceinline__ void face_v1(const double* v1, const FaceIds<D>& f, double* x) {
#pragma unroll
  for (int j = 0; j < D; ++j) {
    if constexpr (D == 3) {
      const double2* r = reinterpret_cast<const double2*>(v1 + 4 * (size_t)f.p[j]);
      const double2 r0 = r[0], r1 = r[1];
      x[3 * j] = r0.x; x[3 * j + 1] = r0.y; x[3 * j + 2] = r1.x;
    } else {
      const double2 r0 = reinterpret_cast<const double2*>(v1 + 2 * (size_t)f.p[j])[0];
      x[2 * j] = r0.x; x[2 * j + 1] = r0.y;
    }
  }
}

// The discrete test at V1 (BVHTree.cpp:464-500) of a pair, and with swap the
// same test with the two primitives exchanged (0 when not asked for).
template <int D, bool SWAP>
__device__ __forceinline__ void discrete_test(const double* v1, const FaceIds<D>& ha, const FaceIds<D>& hb, bool valid,
                                              int& dh, int& dh_swap) {
  bool shared = false;
#pragma unroll
  for (int x = 0; x < D; ++x)
#pragma unroll
    for (int y = 0; y < D; ++y) shared |= ha.p[x] == hb.p[y];
  dh = 0;
  dh_swap = 0;
  if (valid && !shared) {
    double a[3 * D], b[3 * D];
    face_v1<D>(v1, ha, a);
    face_v1<D>(v1, hb, b);
    if constexpr (D == 3) {
      const d3 p1{a[0], a[1], a[2]}, q1{a[3], a[4], a[5]}, r1{a[6], a[7], a[8]};
      const d3 p2{b[0], b[1], b[2]}, q2{b[3], b[4], b[5]}, r2{b[6], b[7], b[8]};
      dh = tri_tri3(p1, q1, r1, p2, q2, r2);
      if (SWAP) dh_swap = tri_tri3(p2, q2, r2, p1, q1, r1);
    } else {
      const d2 p0{a[0], a[1]}, p1{a[2], a[3]}, q0{b[0], b[1]}, q1{b[2], b[3]};
      dh = edge_edge2(p0, p1, q0, q1);
      if (SWAP) dh_swap = edge_edge2(q0, q1, p0, p1);
    }
  }
}

}  // namespace mclg
