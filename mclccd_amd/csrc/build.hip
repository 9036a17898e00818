// BVH build kernels for gfx950: vertex packing, representative-triangle flags,
// leaf boxes, Morton codes, LSD radix sort, bottom-up LBVH (Apetrei 2014) with
// FP32 outward-rounded node boxes, and the child-pair records the traversal reads.
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace mclg {

// ---------------------------------------------------------------------------
// Vertex records (both time steps of a vertex in one 64 B / 32 B record)
// ---------------------------------------------------------------------------
// f32: a float-precision tree computes on float(V) (BVHTree<float,DIM> takes
// float vertices); the records hold those values exactly.
template <int D>
__global__ void k_pack(const double* __restrict__ V0, const double* __restrict__ V1, int nv, double* __restrict__ rec,
                       double* __restrict__ v1, bool f32) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  auto rd = [f32](double x) { return f32 ? (double)(float)x : x; };
  double* r = rec + (size_t)v * VRec<D>::stride;
  if constexpr (D == 3) {
    const double x1 = rd(V1[3 * (size_t)v]), y1 = rd(V1[3 * (size_t)v + 1]), z1 = rd(V1[3 * (size_t)v + 2]);
    r[0] = rd(V0[3 * (size_t)v]); r[1] = rd(V0[3 * (size_t)v + 1]); r[2] = rd(V0[3 * (size_t)v + 2]);
    r[3] = x1; r[4] = y1; r[5] = z1; r[6] = 0.0; r[7] = 0.0;
    if (v1) {
      double* o = v1 + (size_t)v * 4;
      o[0] = x1; o[1] = y1; o[2] = z1; o[3] = 0.0;
    }
  } else {
    const double x1 = rd(V1[2 * (size_t)v]), y1 = rd(V1[2 * (size_t)v + 1]);
    r[0] = rd(V0[2 * (size_t)v]); r[1] = rd(V0[2 * (size_t)v + 1]);
    r[2] = x1; r[3] = y1;
    if (v1) {
      v1[2 * (size_t)v] = x1;
      v1[2 * (size_t)v + 1] = y1;
    }
  }
}

void launch_pack(int dim, const double* V0, const double* V1, int nv, double* rec, hipStream_t s, double* v1, bool f32) {
  if (nv <= 0) return;
  const int g = (nv + 255) / 256;
  if (dim == 3) k_pack<3><<<g, 256, 0, s>>>(V0, V1, nv, rec, v1, f32);
  else k_pack<2><<<g, 256, 0, s>>>(V0, V1, nv, rec, v1, f32);
}

// ---------------------------------------------------------------------------
// Index validation: min and max of an int array.  Each block reduces in LDS
// and issues one atomic pair (same-address atomics serialise, so there are at
// most MM_BLOCKS of them).
// ---------------------------------------------------------------------------
constexpr int MM_BLOCKS = 256, MM_T = 256;

__global__ __launch_bounds__(MM_T) void k_minmax_i32(const int* __restrict__ a, long long n, int* out) {
  __shared__ int s_lo[MM_T / kWave], s_hi[MM_T / kWave];
  int lo = 0x7fffffff, hi = -0x7fffffff - 1;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int v = a[i];
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const int l2 = __shfl_xor(lo, o), h2 = __shfl_xor(hi, o);
    lo = l2 < lo ? l2 : lo;
    hi = h2 > hi ? h2 : hi;
  }
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) { s_lo[w] = lo; s_hi[w] = hi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < MM_T / kWave; ++k) {
      lo = s_lo[k] < lo ? s_lo[k] : lo;
      hi = s_hi[k] > hi ? s_hi[k] : hi;
    }
    atomicMin(out, lo);
    atomicMax(out + 1, hi);
  }
}

void launch_minmax_i32(const int* a, long long n, int* out2, hipStream_t s) {
  if (n <= 0) return;
  long long g = (n + MM_T - 1) / MM_T;
  if (g > MM_BLOCKS) g = MM_BLOCKS;
  k_minmax_i32<<<(int)g, MM_T, 0, s>>>(a, n, out2);
}

// ---------------------------------------------------------------------------
// Fills
// ---------------------------------------------------------------------------
template <typename T> __global__ void k_fill(T* p, long long n, T v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}
template <typename T> static void fill(T* p, long long n, T v, hipStream_t s) {
  if (n <= 0) return;
  long long g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  k_fill<T><<<(int)g, 256, 0, s>>>(p, n, v);
}
void launch_fill_u8(uint8_t* p, long long n, uint8_t v, hipStream_t s) { fill<uint8_t>(p, n, v, s); }
void launch_fill_i32(int* p, long long n, int v, hipStream_t s) { fill<int>(p, n, v, s); }
void launch_fill_u64(unsigned long long* p, long long n, unsigned long long v, hipStream_t s) { fill<unsigned long long>(p, n, v, s); }

// ---------------------------------------------------------------------------
// Representative-triangle flags (BVHTree.cpp:53-79): a vertex belongs to its
// first (face, slot) occurrence, an edge {a,b} to the first (face, slot) whose
// slot edge (P[j], P[(j+1)%3]) it is.  "First" is the minimum of face*pdim+slot,
// found with atomicMin (vertices) and an open-addressing hash of edge keys.
// ---------------------------------------------------------------------------
__global__ void k_vertex_owner(const int* __restrict__ P, long long nslots, int* owner) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  atomicMin(owner + P[i], (int)i);
}

__device__ __forceinline__ uint32_t hash64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return (uint32_t)k;
}

constexpr unsigned long long kEmptyKey = ~0ull;

__device__ __forceinline__ uint64_t edge_key(int a, int b) {
  if (b < a) { const int t = a; a = b; b = t; }
  return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

__global__ void k_edge_insert(const int* __restrict__ P, long long nslots, unsigned long long* keys, int* vals,
                              uint32_t mask) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  const long long f = i / 3;
  const int j = (int)(i - f * 3);
  const uint64_t key = edge_key(P[i], P[f * 3 + (j + 1) % 3]);
  uint32_t h = hash64(key) & mask;
  for (;;) {
    const unsigned long long prev = atomicCAS(keys + h, kEmptyKey, (unsigned long long)key);
    if (prev == kEmptyKey || prev == key) { atomicMin(vals + h, (int)i); return; }
    h = (h + 1) & mask;
  }
}

__device__ __forceinline__ int edge_owner(const unsigned long long* keys, const int* vals, uint32_t mask, uint64_t key) {
  uint32_t h = hash64(key) & mask;
  for (;;) {
    const unsigned long long k = keys[h];
    if (k == key) return vals[h];
    if (k == kEmptyKey) return -1;  // unreachable: every slot edge was inserted
    h = (h + 1) & mask;
  }
}

// flags: bits 0..2 = v[j], bits 3..5 = e[j]
__global__ void k_reptri_flags(const int* __restrict__ P, int np, int pdim, const int* __restrict__ owner,
                               const unsigned long long* __restrict__ keys, const int* __restrict__ vals, uint32_t mask,
                               uint8_t* __restrict__ flags) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= np) return;
  uint32_t fl = 0;
  for (int j = 0; j < pdim && j < 3; ++j) {
    const long long slot = (long long)f * pdim + j;
    if (owner[P[slot]] == (int)slot) fl |= 1u << j;
    if (pdim == 3) {
      const uint64_t key = edge_key(P[slot], P[(long long)f * 3 + (j + 1) % 3]);
      if (edge_owner(keys, vals, mask, key) == (int)slot) fl |= 1u << (3 + j);
    }
  }
  flags[f] = (uint8_t)fl;
}

void launch_reptri(const int* P, int np, int pdim, int nv, int* owner, unsigned long long* keys, int* vals,
                   uint32_t table_size, uint8_t* flags, hipStream_t s) {
  const long long nslots = (long long)np * pdim;
  fill<int>(owner, nv, 0x7fffffff, s);
  const int g = (int)((nslots + 255) / 256);
  k_vertex_owner<<<g, 256, 0, s>>>(P, nslots, owner);
  if (pdim == 3) {
    fill<unsigned long long>(keys, table_size, kEmptyKey, s);
    fill<int>(vals, table_size, 0x7fffffff, s);
    k_edge_insert<<<g, 256, 0, s>>>(P, nslots, keys, vals, table_size - 1);
  }
  k_reptri_flags<<<(np + 255) / 256, 256, 0, s>>>(P, np, pdim, owner, keys, vals, table_size - 1, flags);
}

// ---------------------------------------------------------------------------
// Leaf boxes (BVHTree.cpp:82-121): t0 box, t1 box and swept box of the prim's
// vertices, each padded by exactly box_eta; active if no mask or any vertex
// active.  Also reduces the bounds of the swept-box centres for Morton codes.
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(256) void k_leaf_boxes(const double* __restrict__ rec, const int* __restrict__ P, int np,
                                                    int pdim, const int* __restrict__ active, double eta,
                                                    LeafBox<D>* __restrict__ boxes, uint8_t* __restrict__ act,
                                                    double* __restrict__ partial, SceneArgs S, bool f32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (S.prim_scene && i < np) eta = S.box_eta[S.prim_scene[i]];
  double clo[D], chi[D];
#pragma unroll
  for (int k = 0; k < D; ++k) { clo[k] = __builtin_huge_val(); chi[k] = -__builtin_huge_val(); }
  if (i < np) {
    LeafBox<D> b;
    const double big = 1.7976931348623157e308;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      b.sw_lo[k] = b.t0_lo[k] = b.t1_lo[k] = big;
      b.sw_hi[k] = b.t0_hi[k] = b.t1_hi[k] = -big;
    }
    bool a = active == nullptr;
    for (int j = 0; j < pdim; ++j) {
      const int v = P[(long long)i * pdim + j];
      const double* r = rec + (size_t)v * VRec<D>::stride;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const double x0 = r[k], x1 = r[VRec<D>::t1 + k];
        b.t0_lo[k] = (x0 < b.t0_lo[k]) ? x0 : b.t0_lo[k];
        b.t0_hi[k] = (b.t0_hi[k] < x0) ? x0 : b.t0_hi[k];
        b.t1_lo[k] = (x1 < b.t1_lo[k]) ? x1 : b.t1_lo[k];
        b.t1_hi[k] = (b.t1_hi[k] < x1) ? x1 : b.t1_hi[k];
        b.sw_lo[k] = (x0 < b.sw_lo[k]) ? x0 : b.sw_lo[k];
        b.sw_hi[k] = (b.sw_hi[k] < x0) ? x0 : b.sw_hi[k];
        b.sw_lo[k] = (x1 < b.sw_lo[k]) ? x1 : b.sw_lo[k];
        b.sw_hi[k] = (b.sw_hi[k] < x1) ? x1 : b.sw_hi[k];
      }
      if (active != nullptr && active[v]) a = true;
    }
    if (f32) {
      // float tree (float coordinates, float eta): the float sum or
      // difference, i.e. the exact one rounded once to float (one binary64
      // operation then a rounding to binary32 gives exactly that)
#pragma unroll
      for (int k = 0; k < D; ++k) {
        b.t0_lo[k] = (double)(float)(b.t0_lo[k] - eta); b.t1_lo[k] = (double)(float)(b.t1_lo[k] - eta);
        b.sw_lo[k] = (double)(float)(b.sw_lo[k] - eta);
        b.t0_hi[k] = (double)(float)(b.t0_hi[k] + eta); b.t1_hi[k] = (double)(float)(b.t1_hi[k] + eta);
        b.sw_hi[k] = (double)(float)(b.sw_hi[k] + eta);
      }
    } else {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        b.t0_lo[k] -= eta; b.t1_lo[k] -= eta; b.sw_lo[k] -= eta;
        b.t0_hi[k] += eta; b.t1_hi[k] += eta; b.sw_hi[k] += eta;
      }
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double c = (b.sw_lo[k] + b.sw_hi[k]) / 2.0;
      clo[k] = c; chi[k] = c;
    }
    boxes[i] = b;
    act[i] = a ? 1 : 0;
  }
  if (S.prim_scene) {
    // several scenes: per-scene bounds; a wave whose primitives all belong to
    // one scene reduces first and issues one atomic per bound
    const int sc = i < np ? S.prim_scene[i] : -1;
    const int sc0 = __shfl(sc, 0);
    if (__ballot(i < np && sc != sc0) == 0) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        for (int o = 32; o > 0; o >>= 1) {
          const double l2 = __shfl_xor(clo[k], o), h2 = __shfl_xor(chi[k], o);
          clo[k] = l2 < clo[k] ? l2 : clo[k];
          chi[k] = h2 > chi[k] ? h2 : chi[k];
        }
      }
      if (lane_id() == 0 && sc0 >= 0) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
          atomicMin(S.bounds + (size_t)sc0 * 2 * D + k, (unsigned long long)dbl_to_ord(clo[k]));
          atomicMax(S.bounds + (size_t)sc0 * 2 * D + D + k, (unsigned long long)dbl_to_ord(chi[k]));
        }
      }
    } else if (i < np) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        atomicMin(S.bounds + (size_t)sc * 2 * D + k, (unsigned long long)dbl_to_ord(clo[k]));
        atomicMax(S.bounds + (size_t)sc * 2 * D + D + k, (unsigned long long)dbl_to_ord(chi[k]));
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    for (int o = 32; o > 0; o >>= 1) {
      const double l2 = __shfl_xor(clo[k], o), h2 = __shfl_xor(chi[k], o);
      clo[k] = l2 < clo[k] ? l2 : clo[k];
      chi[k] = h2 > chi[k] ? h2 : chi[k];
    }
  }
  __shared__ double s_lo[4][D], s_hi[4][D];
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < D; ++k) { s_lo[w][k] = clo[k]; s_hi[w][k] = chi[k]; }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      double lo = s_lo[0][k], hi = s_hi[0][k];
      for (int j = 1; j < 4; ++j) {
        lo = s_lo[j][k] < lo ? s_lo[j][k] : lo;
        hi = s_hi[j][k] > hi ? s_hi[j][k] : hi;
      }
      partial[(size_t)blockIdx.x * 2 * D + k] = lo;
      partial[(size_t)blockIdx.x * 2 * D + D + k] = hi;
    }
  }
}

// Reduces the per-block bounds to cbounds (ordered encoding, lo then hi).
template <int D>
__global__ __launch_bounds__(256) void k_reduce_bounds(const double* __restrict__ partial, int nb,
                                                       unsigned long long* cbounds) {
  double lo[D], hi[D];
#pragma unroll
  for (int k = 0; k < D; ++k) { lo[k] = __builtin_huge_val(); hi[k] = -__builtin_huge_val(); }
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double l = partial[(size_t)b * 2 * D + k], h = partial[(size_t)b * 2 * D + D + k];
      lo[k] = l < lo[k] ? l : lo[k];
      hi[k] = h > hi[k] ? h : hi[k];
    }
  __shared__ double s_lo[4][D], s_hi[4][D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    for (int o = 32; o > 0; o >>= 1) {
      const double l2 = __shfl_xor(lo[k], o), h2 = __shfl_xor(hi[k], o);
      lo[k] = l2 < lo[k] ? l2 : lo[k];
      hi[k] = h2 > hi[k] ? h2 : hi[k];
    }
  }
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0)
#pragma unroll
    for (int k = 0; k < D; ++k) { s_lo[w][k] = lo[k]; s_hi[w][k] = hi[k]; }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < D; ++k) {
      double l = s_lo[0][k], h = s_hi[0][k];
      for (int j = 1; j < 4; ++j) {
        l = s_lo[j][k] < l ? s_lo[j][k] : l;
        h = s_hi[j][k] > h ? s_hi[j][k] : h;
      }
      cbounds[k] = dbl_to_ord(l);
      cbounds[D + k] = dbl_to_ord(h);
    }
}

int leaf_box_blocks(int np) { return (np + 255) / 256; }

void launch_leaf_boxes(int dim, const double* rec, const int* P, int np, int pdim, const int* active, double eta,
                       void* boxes, uint8_t* act, double* partial, unsigned long long* cbounds, const SceneArgs& S,
                       hipStream_t s, bool f32) {
  const int g = leaf_box_blocks(np);
  if (S.prim_scene) {
    const unsigned long long lo0 = dbl_to_ord(__builtin_huge_val()), hi0 = dbl_to_ord(-__builtin_huge_val());
    launch_bounds_init(S.bounds, S.n, dim, lo0, hi0, s);
  }
  if (dim == 3) {
    k_leaf_boxes<3><<<g, 256, 0, s>>>(rec, P, np, pdim, active, eta, (LeafBox<3>*)boxes, act, partial, S, f32);
    if (!S.prim_scene) k_reduce_bounds<3><<<1, 256, 0, s>>>(partial, g, cbounds);
  } else {
    k_leaf_boxes<2><<<g, 256, 0, s>>>(rec, P, np, pdim, active, eta, (LeafBox<2>*)boxes, act, partial, S, f32);
    if (!S.prim_scene) k_reduce_bounds<2><<<1, 256, 0, s>>>(partial, g, cbounds);
  }
}

__global__ void k_bounds_init(unsigned long long* b, int nscenes, int dim, unsigned long long lo0, unsigned long long hi0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nscenes * 2 * dim) return;
  b[i] = (i % (2 * dim)) < dim ? lo0 : hi0;
}
void launch_bounds_init(unsigned long long* b, int nscenes, int dim, unsigned long long lo0, unsigned long long hi0,
                        hipStream_t s) {
  const int n = nscenes * 2 * dim;
  k_bounds_init<<<(n + 255) / 256, 256, 0, s>>>(b, nscenes, dim, lo0, hi0);
}

// ---------------------------------------------------------------------------
// Morton codes of swept-box centres: 10 bits per axis (3D), 16 (2D).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // 10 bits -> every 3rd bit
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ uint32_t spread2(uint32_t v) {  // 16 bits -> every 2nd bit
  v &= 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

template <int D>
__global__ void k_morton(const LeafBox<D>* __restrict__ boxes, int np, const unsigned long long* __restrict__ cbounds,
                         SceneArgs S, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                         uint32_t* __restrict__ mort) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const LeafBox<D>& b = boxes[i];
  const unsigned long long* bd = S.prim_scene ? S.bounds + (size_t)S.prim_scene[i] * 2 * D : cbounds;
  constexpr double cells = D == 3 ? 1024.0 : 65536.0;
  uint32_t q[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double lo = ord_to_dbl(bd[k]), hi = ord_to_dbl(bd[D + k]);
    const double c = (b.sw_lo[k] + b.sw_hi[k]) / 2.0;
    const double ext = hi - lo;
    double u = ext > 0 ? (c - lo) / ext * cells : 0.0;
    u = u < 0 ? 0.0 : (u > cells - 1 ? cells - 1 : u);
    q[k] = (uint32_t)u;
  }
  uint32_t code;
  if constexpr (D == 3) code = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
  else code = (spread2(q[0]) << 1) | spread2(q[1]);
  keys[i] = code;
  vals[i] = (uint32_t)i;
  if (mort) mort[i] = code;
}

void launch_morton(int dim, const void* boxes, int np, const unsigned long long* cbounds, const SceneArgs& S,
                   uint32_t* keys, uint32_t* vals, uint32_t* mort, hipStream_t s) {
  const int g = (np + 255) / 256;
  if (dim == 3) k_morton<3><<<g, 256, 0, s>>>((const LeafBox<3>*)boxes, np, cbounds, S, keys, vals, mort);
  else k_morton<2><<<g, 256, 0, s>>>((const LeafBox<2>*)boxes, np, cbounds, S, keys, vals, mort);
}

// Scene keys of the Morton-sorted order, and the Morton codes regathered after
// the stable sort by scene: together the (scene, Morton) order.
__global__ void k_scene_keys(const uint32_t* __restrict__ order, int n, const int* __restrict__ prim_scene,
                             uint32_t* __restrict__ hi) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) hi[j] = (uint32_t)prim_scene[order[j]];
}
__global__ void k_gather_u32(const uint32_t* __restrict__ order, int n, const uint32_t* __restrict__ src,
                             uint32_t* __restrict__ dst) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) dst[j] = src[order[j]];
}
void launch_scene_keys(const uint32_t* order, int n, const int* prim_scene, uint32_t* hi, hipStream_t s) {
  k_scene_keys<<<(n + 255) / 256, 256, 0, s>>>(order, n, prim_scene, hi);
}
void launch_gather_u32(const uint32_t* order, int n, const uint32_t* src, uint32_t* dst, hipStream_t s) {
  k_gather_u32<<<(n + 255) / 256, 256, 0, s>>>(order, n, src, dst);
}

// ---------------------------------------------------------------------------
// LSD radix sort of (uint32 key, uint32 value), 8-bit digits.
//   hist:    per-tile digit counts (LDS atomics)      -> hist[digit*ntiles+tile]
//   scan:    exclusive scan of hist in digit-major order (three phases)
//   scatter: stable per-tile ranking by wave match (8 ballots per key)
// ---------------------------------------------------------------------------
constexpr int RS_T = 256, RS_ROUNDS = 16, RS_TILE = RS_T * RS_ROUNDS;

__global__ __launch_bounds__(RS_T) void k_radix_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                     uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const long long i = base + r * RS_T + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan in three phases: per-chunk sums, a one-block scan of those,
// then each chunk scanned in place from its offset.  SC_T threads x SC_PER
// entries per chunk; T is uint32_t (radix histograms) or unsigned long long.
constexpr int SC_T = 1024, SC_PER = 8, SC_CHUNK = SC_T * SC_PER;

template <typename T> __device__ __forceinline__ T wave_incl_scan_t(T v) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const T u = __shfl_up(v, o);
    if (lane_id() >= o) v += u;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread (SC_T threads); also returns the total.
template <typename T> __device__ __forceinline__ T block_excl_scan(T v, T* s_w, T& total) {
  const int w = threadIdx.x / kWave, nw = SC_T / kWave;
  const T incl = wave_incl_scan_t(v);
  if (lane_id() == kWave - 1) s_w[w] = incl;
  __syncthreads();
  if (w == 0) {
    const T x = lane_id() < nw ? s_w[lane_id()] : T(0);
    const T xi = wave_incl_scan_t(x);
    if (lane_id() < nw) s_w[lane_id()] = xi - x;
    if (lane_id() == nw - 1) s_w[nw] = xi;
  }
  __syncthreads();
  total = s_w[nw];
  return s_w[w] + incl - v;
}

template <typename T>
__global__ __launch_bounds__(SC_T) void k_scan_sums(const T* __restrict__ in, long long n, T* __restrict__ part) {
  __shared__ T s_w[SC_T / kWave + 1];
  const long long base = (long long)blockIdx.x * SC_CHUNK;
  T v = 0;
#pragma unroll
  for (int r = 0; r < SC_PER; ++r) {
    const long long i = base + r * SC_T + threadIdx.x;
    if (i < n) v += in[i];
  }
  T total;
  (void)block_excl_scan(v, s_w, total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

template <typename T> __global__ __launch_bounds__(SC_T) void k_scan_top(T* part, long long nb) {
  __shared__ T s_w[SC_T / kWave + 1];
  T carry = 0;
  for (long long b0 = 0; b0 < nb; b0 += SC_T) {
    const long long i = b0 + threadIdx.x;
    const T v = i < nb ? part[i] : T(0);
    T total;
    const T ex = block_excl_scan(v, s_w, total);
    if (i < nb) part[i] = carry + ex;
    carry += total;
    __syncthreads();
  }
}

template <typename T>
__global__ __launch_bounds__(SC_T) void k_scan_apply(T* __restrict__ data, long long n, const T* __restrict__ part) {
  __shared__ T s_w[SC_T / kWave + 1];
  const long long base = (long long)blockIdx.x * SC_CHUNK + (long long)threadIdx.x * SC_PER;
  T x[SC_PER], sum = 0;
#pragma unroll
  for (int r = 0; r < SC_PER; ++r) {
    x[r] = base + r < n ? data[base + r] : T(0);
    sum += x[r];
  }
  T total;
  T run = part[blockIdx.x] + block_excl_scan(sum, s_w, total);
#pragma unroll
  for (int r = 0; r < SC_PER; ++r) {
    if (base + r < n) data[base + r] = run;
    run += x[r];
  }
}

template <typename T> static void launch_scan_t(T* data, long long n, T* part, hipStream_t s) {
  if (n <= 0) return;
  const unsigned nb = (unsigned)((n + SC_CHUNK - 1) / SC_CHUNK);
  k_scan_sums<T><<<nb, SC_T, 0, s>>>(data, n, part);
  k_scan_top<T><<<1, SC_T, 0, s>>>(part, (long long)nb);
  k_scan_apply<T><<<nb, SC_T, 0, s>>>(data, n, part);
}

static void launch_scan(uint32_t* hist, int total, uint32_t* part, hipStream_t s) {
  launch_scan_t<uint32_t>(hist, total, part, s);
}

long long scan_part_words(long long n) { return (n + SC_CHUNK - 1) / SC_CHUNK; }

void launch_scan_u64(unsigned long long* data, long long n, unsigned long long* part, hipStream_t s) {
  launch_scan_t<unsigned long long>(data, n, part, s);
}

__global__ __launch_bounds__(RS_T) void k_radix_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                        uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int n,
                                                        int shift, const uint32_t* __restrict__ hist) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t wcnt[RS_T / kWave][256];
  const int t = threadIdx.x, w = t / kWave;
  base[t] = hist[(size_t)t * gridDim.x + blockIdx.x];
#pragma unroll
  for (int j = 0; j < RS_T / kWave; ++j) wcnt[j][t] = 0;
  __syncthreads();
  const long long tile = (long long)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const long long i = tile + r * RS_T + t;
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u;
    const uint32_t val = valid ? vin[i] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const uint64_t m = __ballot((d >> bit) & 1u);
      peers &= ((d >> bit) & 1u) ? m : ~m;
    }
    const int rank = popc_below(peers);
    if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t off = base[d] + rank;
      for (int j = 0; j < w; ++j) off += wcnt[j][d];
      kout[off] = key;
      vout[off] = val;
    }
    __syncthreads();
    uint32_t tot = 0;
#pragma unroll
    for (int j = 0; j < RS_T / kWave; ++j) { tot += wcnt[j][t]; wcnt[j][t] = 0; }
    base[t] += tot;
    __syncthreads();
  }
}

int radix_tiles(int n) { return (n + RS_TILE - 1) / RS_TILE; }
size_t radix_hist_words(int n) {
  const size_t h = 256 * (size_t)radix_tiles(n);
  return h + (h + SC_CHUNK - 1) / SC_CHUNK;  // digit counts, then the scan's chunk sums
}

// Sorts (keys, vals) in place (ping-pong through tmp buffers); nbits rounded up to 8.
void launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* ktmp, uint32_t* vtmp, uint32_t* hist, int n, int nbits,
                       hipStream_t s) {
  if (n <= 1) return;
  const int tiles = radix_tiles(n);
  uint32_t *ki = keys, *vi = vals, *ko = ktmp, *vo = vtmp;
  int passes = (nbits + 7) / 8;
  if (passes & 1) ++passes;  // even number of passes: result lands back in keys/vals
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    k_radix_hist<<<tiles, RS_T, 0, s>>>(ki, n, shift, hist);
    launch_scan(hist, 256 * tiles, hist + 256 * (size_t)tiles, s);
    k_radix_scatter<<<tiles, RS_T, 0, s>>>(ki, vi, ko, vo, n, shift, hist);
    uint32_t* t;
    t = ki; ki = ko; ko = t;
    t = vi; vi = vo; vo = t;
  }
}

// ---------------------------------------------------------------------------
// LBVH: bottom-up agglomerative build (Apetrei 2014).  Internal node k splits
// sorted leaves k | k+1; the split key orders splits by (Morton XOR, trailing
// ones of k, k) so that runs of equal codes form balanced subtrees and every
// comparison is strict.  Node ids: internal k -> k, leaf s -> (n-1)+s.
// Boxes are FP32 rounded outward from the FP64 leaf boxes: every culling test
// on them is conservative; the exact FP64 test runs at leaf pairs.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t split_key(const uint32_t* lo, const uint32_t* hi, int k) {
  const uint64_t a = ((uint64_t)(hi ? hi[k] : 0u) << 32) | lo[k];
  const uint64_t b = ((uint64_t)(hi ? hi[k + 1] : 0u) << 32) | lo[k + 1];
  const uint64_t x = a ^ b;
  const uint32_t hb = x ? 64u - (uint32_t)__clzll((long long)x) : 0u;  // highest differing bit + 1
  const uint32_t ones = (uint32_t)__builtin_ctz(~(uint32_t)k);
  return ((uint64_t)hb << 32) | ((uint64_t)ones << 26) | (uint32_t)k;
}

template <int D> __device__ __forceinline__ void merge_node(NodeF<D>& o, const NodeF<D>& a, const NodeF<D>& b) {
#pragma unroll
  for (int k = 0; k < D; ++k) {
    o.t0_lo[k] = fminf(a.t0_lo[k], b.t0_lo[k]); o.t0_hi[k] = fmaxf(a.t0_hi[k], b.t0_hi[k]);
    o.t1_lo[k] = fminf(a.t1_lo[k], b.t1_lo[k]); o.t1_hi[k] = fmaxf(a.t1_hi[k], b.t1_hi[k]);
  }
  o.meta = (a.meta | b.meta) & 0x80000000u;
}

// Parent of the node spanning sorted leaves [l, r] (Apetrei 2014): internal
// node r (this node is its left child) or l-1 (right child), whichever split
// orders lower.
__device__ __forceinline__ bool parent_is_right(const uint32_t* keys, const uint32_t* hi, int l, int r, int n) {
  return l == 0 || (r != n - 1 && split_key(keys, hi, r) < split_key(keys, hi, l - 1));
}

// A leaf whose FP64 t0 or t1 box is empty (a coordinate axis of NaNs) passes
// the exact test whenever the swept boxes overlap; its FP32 t0 and t1 boxes
// are set to the swept box so that no culling test (here or at an ancestor,
// whose boxes are unions) separates it at both times.
template <int D> __device__ __forceinline__ NodeF<D> leaf_node(const LeafBox<D>& b, bool active) {
  NodeF<D> L;
  bool emp = false;
#pragma unroll
  for (int k = 0; k < D; ++k) emp |= (b.t0_lo[k] > b.t0_hi[k]) | (b.t1_lo[k] > b.t1_hi[k]);
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const float sl = f32_down(b.sw_lo[k]), sh = f32_up(b.sw_hi[k]);
    L.t0_lo[k] = emp ? sl : f32_down(b.t0_lo[k]); L.t0_hi[k] = emp ? sh : f32_up(b.t0_hi[k]);
    L.t1_lo[k] = emp ? sl : f32_down(b.t1_lo[k]); L.t1_hi[k] = emp ? sh : f32_up(b.t1_hi[k]);
  }
  L.first = L.last = 0;
  L.child = emp ? 1 : 0;
  L.meta = active ? 0x80000000u : 0u;
  return L;
}

// Phase A: each block of LB_T threads owns sorted leaves [B, E) and builds every
// node whose split lies in [B, E-2] and whose children it completes, with LDS
// visit counters (workgroup scope only).  A completed node whose parent split
// is a block boundary (B-1 or E-1) is handed to phase B.
constexpr int LB_T = 256;

template <int D>
__global__ __launch_bounds__(LB_T) void k_lbvh_local(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ hi,
                                                     const uint32_t* __restrict__ order, int n,
                                                     const LeafBox<D>* __restrict__ boxes,
                                                     const uint8_t* __restrict__ act, NodeF<D>* nodes,
                                                     SortedLeaf<D>* __restrict__ sboxes, int* lchild,
                                                     int* rchild, int* rng_l, int* rng_r, int* visits, int4* frontier,
                                                     unsigned long long* nfrontier, SceneArgs S) {
  __shared__ NodeF<D> s_node[2 * LB_T];  // [0, LB_T): leaves, [LB_T, 2 LB_T): internal nodes
  __shared__ int s_vis[LB_T], s_lc[LB_T], s_rc[LB_T], s_rl[LB_T], s_rr[LB_T];
  const int B = blockIdx.x * LB_T, E = min(B + LB_T, n);
  const int t = threadIdx.x, s = B + t;
  s_vis[t] = 0;
  if (t == LB_T - 1 && s < n - 1) visits[s] = 0;  // boundary split: phase B's counter
  if (s < n) {
    const uint32_t prim = order[s];
    const LeafBox<D> lb = boxes[prim];
    NodeF<D> L = leaf_node<D>(lb, act[prim] != 0);
    SortedLeaf<D> sl;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      sl.t0_lo[k] = lb.t0_lo[k]; sl.t0_hi[k] = lb.t0_hi[k];
      sl.t1_lo[k] = lb.t1_lo[k]; sl.t1_hi[k] = lb.t1_hi[k];
      sl.sw_lo[k] = lb.sw_lo[k]; sl.sw_hi[k] = lb.sw_hi[k];
    }
    sl.child = L.child;
    sl.meta = L.meta;
    sboxes[s] = sl;
    L.first = L.last = s;
    nodes[(n - 1) + s] = L;
    s_node[t] = L;
  }
  __syncthreads();
  if (s < n && n > 1) {
    int l = s, r = s, node = (n - 1) + s;
    NodeF<D> cur = s_node[t];
    for (;;) {
      const bool right = parent_is_right(keys, hi, l, r, n);
      const int p = right ? r : l - 1;
      if (p < B || p > E - 2) {  // parent split on a block boundary: phase B
        const unsigned long long k = atomicAdd(nfrontier, 1ull);
        frontier[k] = make_int4(node, l, r, 0);
        break;
      }
      const int q = p - B;
      if (right) { s_lc[q] = node; s_rl[q] = l; }
      else { s_rc[q] = node; s_rr[q] = r; }
      __threadfence_block();
      if (atomicAdd(&s_vis[q], 1) == 0) break;  // the sibling carries on
      __threadfence_block();
      l = s_rl[q];
      r = s_rr[q];
      const int lc = s_lc[q], rc = s_rc[q];
      const NodeF<D>& a = lc < n - 1 ? s_node[LB_T + (lc - B)] : s_node[lc - (n - 1) - B];
      const NodeF<D>& b = rc < n - 1 ? s_node[LB_T + (rc - B)] : s_node[rc - (n - 1) - B];
      merge_node<D>(cur, a, b);
      cur.child = lc;
      cur.first = l;
      cur.last = r;
      s_node[LB_T + q] = cur;
      nodes[p] = cur;
      node = p;
      if (l == 0 && r == n - 1) break;  // root of the whole tree
    }
  }
  __syncthreads();
  // publish this block's partial nodes for phase B
  if (s < E - 1) {
    visits[s] = s_vis[t];
    lchild[s] = s_lc[t]; rng_l[s] = s_rl[t];
    rchild[s] = s_rc[t]; rng_r[s] = s_rr[t];
  }
}

// Phase B: the frontier nodes climb the block-crossing part of the tree with
// device-scope release/acquire (one pair per crossing node arrival).
template <int D>
__global__ __launch_bounds__(256) void k_lbvh_cross(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ hi,
                                                    int n, NodeF<D>* nodes, int* lchild, int* rchild, int* rng_l,
                                                    int* rng_r, int* visits, const int4* __restrict__ frontier,
                                                    const unsigned long long* nfrontier, SceneArgs S) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= (long long)*nfrontier) return;
  const int4 f = frontier[k];
  int node = f.x, l = f.y, r = f.z;
  for (;;) {
    const bool right = parent_is_right(keys, hi, l, r, n);
    const int p = right ? r : l - 1;
    if (right) { lchild[p] = node; rng_l[p] = l; }
    else { rchild[p] = node; rng_r[p] = r; }
    __threadfence();
    if (__hip_atomic_fetch_add(visits + p, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
    __threadfence();
    l = __hip_atomic_load(rng_l + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    r = __hip_atomic_load(rng_r + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int lc = __hip_atomic_load(lchild + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int rc = __hip_atomic_load(rchild + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const NodeF<D> a = nodes[lc], b = nodes[rc];
    NodeF<D> o;
    merge_node<D>(o, a, b);
    o.child = lc;
    o.first = l;
    o.last = r;
    nodes[p] = o;
    node = p;
    if (l == 0 && r == n - 1) return;
  }
}

// Child-pair records of the internal nodes, and the index of the root (the
// internal node covering every leaf).
template <int D>
__global__ void k_child_pairs(const NodeF<D>* __restrict__ nodes, int n, const int* __restrict__ lchild,
                              const int* __restrict__ rchild, ChildPair<D>* __restrict__ cp, int* root) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n - 1) return;
  const int lc = lchild[p], rc = rchild[p];
  NodeF<D> a = nodes[lc], b = nodes[rc];
  a.meta = (a.meta & 0x80000000u) | (uint32_t)(lc + 1);
  b.meta = (b.meta & 0x80000000u) | (uint32_t)(rc + 1);
  cp[p].c[0] = a;
  cp[p].c[1] = b;
  if (a.first == 0 && b.last == n - 1) *root = p;
}

int lbvh_frontier_cap(int n) { return 2 * ((n + LB_T - 1) / LB_T) + 2; }

void launch_lbvh(int dim, const uint32_t* keys, const uint32_t* hi, const uint32_t* order, int n, const void* boxes,
                 const uint8_t* act, void* nodes, void* sboxes, int* lchild, int* rchild, int* rng_l, int* rng_r,
                 int* visits, int4* frontier, unsigned long long* nfrontier, const SceneArgs& S, hipStream_t s,
                 void* cp, int* root) {
  if (n < 1) return;
  fill<unsigned long long>(nfrontier, 1, 0ull, s);
  const int g = (n + LB_T - 1) / LB_T;
  const int gf = (lbvh_frontier_cap(n) + 255) / 256;
  if (dim == 3) {
    k_lbvh_local<3><<<g, LB_T, 0, s>>>(keys, hi, order, n, (const LeafBox<3>*)boxes, act, (NodeF<3>*)nodes,
                                       (SortedLeaf<3>*)sboxes, lchild, rchild, rng_l, rng_r, visits, frontier,
                                       nfrontier, S);
    if (n > 1) k_lbvh_cross<3><<<gf, 256, 0, s>>>(keys, hi, n, (NodeF<3>*)nodes, lchild, rchild, rng_l, rng_r, visits,
                                                  frontier, nfrontier, S);
    if (cp && n > 1)
      k_child_pairs<3><<<(n - 1 + 255) / 256, 256, 0, s>>>((const NodeF<3>*)nodes, n, lchild, rchild,
                                                          (ChildPair<3>*)cp, root);
  } else {
    k_lbvh_local<2><<<g, LB_T, 0, s>>>(keys, hi, order, n, (const LeafBox<2>*)boxes, act, (NodeF<2>*)nodes,
                                       (SortedLeaf<2>*)sboxes, lchild, rchild, rng_l, rng_r, visits, frontier,
                                       nfrontier, S);
    if (n > 1) k_lbvh_cross<2><<<gf, 256, 0, s>>>(keys, hi, n, (NodeF<2>*)nodes, lchild, rchild, rng_l, rng_r, visits,
                                                  frontier, nfrontier, S);
    if (cp && n > 1)
      k_child_pairs<2><<<(n - 1 + 255) / 256, 256, 0, s>>>((const NodeF<2>*)nodes, n, lchild, rchild,
                                                          (ChildPair<2>*)cp, root);
  }
}

}  // namespace mclg
