// Shared device/host definitions for the gfx950 collision kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace mclg {

constexpr int kWave = 64;  // CDNA wavefront width

// Packed vertex record: one record per vertex holding both time steps, so a
// stencil gathers one contiguous 64 B (3D) or 32 B (2D) record per vertex.
//   3D: x0 y0 z0 x1 y1 z1 pad pad      2D: x0 y0 x1 y1
// The six 3D values are contiguous so that three 16 B loads fetch them (four
// with a pad after z0): ACCD 28.4 -> 28.1 ms at 1M triangles.
template <int D> struct VRec;
template <> struct VRec<3> { static constexpr int stride = 8; static constexpr int t1 = 3; };
template <> struct VRec<2> { static constexpr int stride = 4; static constexpr int t1 = 2; };

// FP64 leaf box: swept, t0, t1 (lo then hi) — 6*D doubles, reference layout of
// mcl::BoundingBox (BVHLeaf.hpp:12-29) flattened.
template <int D> struct LeafBox { double sw_lo[D], sw_hi[D], t0_lo[D], t0_hi[D], t1_lo[D], t1_hi[D]; };

// A leaf in sorted order as the traversal reads it: the FP64 t0 and t1 boxes
// and the leaf's flags (child: 1 if its t0 or t1 box is empty; meta:
// active<<31) in the first two 64 B lines, the swept box behind them for the
// empty-box fallback.  192 B (3D) / 128 B (2D), 64 B aligned.
template <int D> struct alignas(64) SortedLeaf {
  double t0_lo[D], t0_hi[D], t1_lo[D], t1_hi[D];
  int child;
  uint32_t meta;
  double sw_lo[D], sw_hi[D];
};

// Tree node (64 B, one aligned half line): the t0 and t1 boxes rounded outward
// to FP32 (conservative for the culling test), meta = active<<31 (the child-pair
// copies below put the node's index in the low bits), and the range of sorted
// leaves the node covers.  The swept box is their
// union: the reference builds it from the same points with the same padding
// (BVHTree.cpp:94-117), and min/max and monotone rounding commute, so the union
// equals the FP32-rounded FP64 swept box exactly.  child is the first child of
// an internal node; of a leaf it is 1 if its FP64 t0 or t1 box is empty (then
// both FP32 boxes hold the swept box, so that no culling test separates it at
// t0 and t1), else 0.
template <int D> struct alignas(16) NodeF;
template <> struct alignas(16) NodeF<3> {
  float t0_lo[3], t0_hi[3], t1_lo[3], t1_hi[3];
  int child;
  uint32_t meta;
  int first, last;  // sorted leaves covered: [first, last]
};
template <> struct alignas(16) NodeF<2> {
  float t0_lo[2], t0_hi[2], t1_lo[2], t1_hi[2];
  int child;
  uint32_t meta;
  int first, last;
};

static_assert(sizeof(NodeF<3>) == 64 && sizeof(NodeF<2>) == 48, "node layout");
static_assert(sizeof(SortedLeaf<3>) == 192 && sizeof(SortedLeaf<2>) == 128, "sorted leaf layout");

// Both children of internal node p side by side (one 128 B line in 3D), as the
// traversal reads them: copies of their NodeF with meta = (index+1) |
// active<<31, the child's own node index in place of its escape link.
template <int D> struct ChildPair {
  NodeF<D> c[2];
};
inline size_t child_pair_bytes(int dim) { return dim == 3 ? sizeof(ChildPair<3>) : sizeof(ChildPair<2>); }

// Face data of the candidate kernels, which gather two faces per leaf pair at
// random: compact arrays in sorted (Morton) order, rebuilt by every traverse,
// and V1 per vertex.  A 16 B head puts eight leaves in one 128 B line, so the
// gathers touch far fewer lines than one whole-face record per leaf would.
//   head[s]  {p0, p1, p2, reptri flags} (3D) / {p0, p1, flags, 0} (2D)
//   rank[s]  orients pairs (lower rank first); the primitive index unless a rank was set
//   prim[s]  primitive index (discrete hits, kept leaf pairs)
//   v1[v]    V1 of vertex v: {x, y, z, 0} (3D) / {x, y} (2D)
struct FaceArrays {
  int4* head = nullptr;
  int* rank = nullptr;
  int* prim = nullptr;
  double* v1 = nullptr;
};
inline size_t face_bytes(size_t np) { return (sizeof(int4) + 2 * sizeof(int)) * np; }
inline size_t v1_stride(int dim) { return dim == 3 ? 4 : 2; }
// The leaf arrays inside one allocation of face_bytes(np) bytes.
inline FaceArrays face_arrays(void* base, size_t np, double* v1) {
  FaceArrays f;
  f.head = static_cast<int4*>(base);
  f.rank = reinterpret_cast<int*>(f.head + np);
  f.prim = f.rank + np;
  f.v1 = v1;
  return f;
}

// Device counters of one traversal (all 64-bit).
enum Counter : int {
  C_LEAF_PAIRS = 0,
  C_CAND_VF,
  C_CAND_EE,
  C_HITS,
  C_DISCRETE,
  C_DIST_EVALS,
  C_ERRORS,     // kernels that detected an inconsistent state
  C_EE_EVALS,   // distance evaluations of edge-edge stencils (also counted in C_DIST_EVALS)
  C_ORDER_DEP,  // leaf pairs whose discrete result depends on the argument order (when checked)
  C_TR_STEPS,   // traversal wave-steps (child-pair records visited, summed over waves)
  C_SUSPENDED,  // ACCD stencils suspended after three iterations and resumed
  C_CTCD_SEL,   // VF stencils that pass the CTCD prefilter (the selection pass's list length)
  C_COUNT = 12
};

// ---- wave helpers ---------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0)); }
__device__ __forceinline__ int popc_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

// Wave-aggregated append: every active lane with pred gets a slot in a global
// list whose length lives in *counter.  Returns the slot (or -1).
__device__ __forceinline__ long long wave_append(bool pred, unsigned long long* counter) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return -1;
  const int leader = __ffsll((long long)m) - 1;
  unsigned long long base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
  base = __shfl(base, leader);
  return pred ? (long long)(base + popc_below(m)) : -1;
}

// Exclusive prefix sum of v over the wave's lanes (inactive lanes count 0).
__device__ __forceinline__ uint32_t wave_prefix_u32(uint32_t v) {
  uint32_t incl = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t u = __shfl_up(incl, o);
    if (lane_id() >= o) incl += u;
  }
  return incl - v;
}

// Order-preserving map of a double onto uint64 (for atomic min/max).
__host__ __device__ __forceinline__ uint64_t dbl_to_ord(double x) {
  uint64_t b;
  __builtin_memcpy(&b, &x, 8);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double ord_to_dbl(uint64_t o) {
  uint64_t b = (o >> 63) ? (o & 0x7fffffffffffffffull) : ~o;
  double x;
  __builtin_memcpy(&x, &b, 8);
  return x;
}

// One-ulp steps of a float toward -inf / +inf (finite or infinite input).
__device__ __forceinline__ float step_down(float f) {
  if (f == 0.0f) return -__int_as_float(1);
  const int b = __float_as_int(f);
  return __int_as_float(f > 0.0f ? b - 1 : b + 1);
}
__device__ __forceinline__ float step_up(float f) {
  if (f == 0.0f) return __int_as_float(1);
  const int b = __float_as_int(f);
  return __int_as_float(f > 0.0f ? b + 1 : b - 1);
}
// FP32 rounding toward -inf / +inf of a double (outward box rounding).
__device__ __forceinline__ float f32_down(double x) {
  float f = (float)x;
  if ((double)f > x) f = step_down(f);
  return f;
}
__device__ __forceinline__ float f32_up(double x) {
  float f = (float)x;
  if ((double)f < x) f = step_up(f);
  return f;
}

// blockIdx remapped so that each XCD (blocks are dealt round-robin to the 8
// XCDs) receives a contiguous range of logical blocks: neighbouring work (in
// Morton order) then shares one L2.  The grid must be a multiple of 8 blocks
// (launchers round up with xcd_grid); the map is then a bijection on
// [0, gridDim.x) and logical blocks past the work simply find no items.
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x;
  const int per = gridDim.x / 8;
  return (b % 8) * per + b / 8;
}
inline int xcd_grid(long long items, int per_block) {
  long long b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  return (int)((b + 7) / 8 * 8);
}

// A wave-uniform 64-bit index as the compiler can see it: the same value read
// from the first active lane.
__device__ __forceinline__ long long uniform_index(long long i) {
  return ((long long)__builtin_amdgcn_readfirstlane((int)(i >> 32)) << 32) |
         (unsigned)__builtin_amdgcn_readfirstlane((int)i);
}

// A load at a wave-uniform address through the scalar cache (constant address
// space).  Only for data that no thread of the running kernel writes: the
// scalar cache does not see this kernel's vector stores (it is invalidated at
// every kernel launch).  Unlike a vector load it does not wait behind the
// wave's earlier vector stores, which share the vector loads' in-order counter.
template <typename T> __device__ __forceinline__ T scalar_load(const T* p) {
  return *(const __attribute__((address_space(4))) T*)p;
}

}  // namespace mclg
