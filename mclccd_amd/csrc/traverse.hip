// Self-traversal and narrow phase for gfx950.
//
// k_traverse: one lane per sorted leaf s (the query), 64 consecutive leaves per
//   wave walking the LBVH together (child-pair records, wave-uniform stack);
//   lane s pairs only with leaves after it, so every unordered leaf pair is
//   found once.  Internal nodes are culled with the FP32 outward-rounded boxes;
//   leaves are decided by the reference's exact test (BVHTree.cpp:401-426) on
//   the FP64 boxes (kept in sorted order).  Pairs (sorted positions) are staged
//   per wave in LDS and appended in batches.
// k_face_records: per sorted leaf, the compact head (P, reptri flags), rank
//   and primitive index.
// k_expand: one lane per leaf pair: orientation, the discrete test and the
//   representative-triangle candidates (BVHTree.cpp:252-398) into the VF / EE
//   stencil queues.
// k_accd: one lane per stencil, FP64 ACCD (NarrowPhase.cpp:634-711).
#include <cmath>
#include <type_traits>

#include "ctcd.hpp"
#include "pairs.hpp"

#pragma clang fp contract(off)

namespace mclg {

// ---------------------------------------------------------------------------
// Leaf-pair tests
// ---------------------------------------------------------------------------
// Node cull on the FP32 outward-rounded boxes: active, and no axis
// separated at both t0 and t1.  It never rejects a pair the exact test accepts:
// FP32 boxes that separate contain FP64 boxes that separate, and a leaf with an
// empty t0 or t1 box holds its swept box in both FP32 boxes (the exact test then
// asks only for overlapping swept boxes, the unions of the t0 and t1 boxes).
// Branch-free, so the node's loads are all issued at once.
template <int D> __device__ __forceinline__ bool cull_f32(const NodeF<D>& q, const NodeF<D>& b) {
  bool ok = ((q.meta | b.meta) & 0x80000000u) != 0;
#pragma unroll
  for (int k = 0; k < D; ++k)
    ok &= !(((q.t0_lo[k] > b.t0_hi[k]) & (q.t1_lo[k] > b.t1_hi[k])) |
            ((q.t0_hi[k] < b.t0_lo[k]) & (q.t1_hi[k] < b.t1_lo[k])));
  return ok;
}

// BVHTree.cpp:401-426 on two leaves (symmetric in l and r), reading the FP64
// boxes in place one axis at a time: false if both inactive or the swept boxes
// are disjoint; true if a t0 or t1 box is empty; else false iff some axis
// separates both at t0 and at t1.  The traversal's fallback for pairs with an
// empty box, kept out of line and small in registers.
template <int D>
__device__ __attribute__((noinline)) bool leaf_test_fp64(const SortedLeaf<D>* l, bool la, const SortedLeaf<D>* r,
                                                         bool ra) {
  bool ov = la | ra, emp = false, sep = false;
#pragma nounroll
  for (int k = 0; k < D; ++k) {
    const double lsl = l->sw_lo[k], lsh = l->sw_hi[k], rsl = r->sw_lo[k], rsh = r->sw_hi[k];
    ov &= (lsl <= rsh) & (rsl <= lsh);
    const double l0l = l->t0_lo[k], l0h = l->t0_hi[k], l1l = l->t1_lo[k], l1h = l->t1_hi[k];
    const double r0l = r->t0_lo[k], r0h = r->t0_hi[k], r1l = r->t1_lo[k], r1h = r->t1_hi[k];
    emp |= (l0l > l0h) | (r0l > r0h) | (l1l > l1h) | (r1l > r1h);
    sep |= ((l0l > r0h) & (l1l > r1h)) | ((l0h < r0l) & (l1h < r1l));
  }
  return ov & (emp | !sep);
}

// The t0 and t1 FP64 boxes of a sorted leaf (the head of its record).  With
// both boxes non-empty the exact test is (la | ra) & !sep: disjoint swept boxes
// on an axis separate that axis at t0 and at t1, so the swept overlap is implied.
// FP64 boxes hold no NaN (box extension skips NaN coordinates; an axis of NaNs
// leaves the box empty, which takes the fallback).
template <int D> struct Box01 {
  double t0_lo[D], t0_hi[D], t1_lo[D], t1_hi[D];
};
template <int D> __device__ __forceinline__ const Box01<D>* box01(const SortedLeaf<D>* b) {
  return reinterpret_cast<const Box01<D>*>(b->t0_lo);
}
template <int D> __device__ __forceinline__ bool separated(const Box01<D>& q, const Box01<D>& b) {
  bool sep = false;
#pragma unroll
  for (int a = 0; a < D; ++a)
    sep |= ((q.t0_lo[a] > b.t0_hi[a]) & (q.t1_lo[a] > b.t1_hi[a])) | ((q.t0_hi[a] < b.t0_lo[a]) & (q.t1_hi[a] < b.t1_lo[a]));
  return sep;
}

// The same cull with the query's exact FP64 t0 / t1 boxes against a node's
// outward-rounded FP32 boxes (converted exactly): the FP64 box lies inside the
// query's FP32 box, so this culls at least as much and never a pair the exact
// test accepts.  A query with an empty FP64 t0 or t1 box (qemp) never culls.
template <int D> __device__ __forceinline__ bool cull_q64(const Box01<D>& q, bool qa, bool qemp, const NodeF<D>& b) {
  bool sep = false;
#pragma unroll
  for (int k = 0; k < D; ++k)
    sep |= ((q.t0_lo[k] > (double)b.t0_hi[k]) & (q.t1_lo[k] > (double)b.t1_hi[k])) |
           ((q.t0_hi[k] < (double)b.t0_lo[k]) & (q.t1_hi[k] < (double)b.t1_lo[k]));
  return (qa | ((b.meta & 0x80000000u) != 0)) & (qemp | !sep);
}

// Whole-record loads as 16-byte vectors (sizes are multiples of 16).  With
// PIN the loaded registers are consumed by an empty asm statement (no memory
// access in it), so the compiler cannot split off and delay any part.
template <typename T, bool PIN = false> __device__ __forceinline__ T load16(const T* p) {
  static_assert(sizeof(T) % 16 == 0, "16-byte multiple");
  int4 raw[sizeof(T) / 16];
  const int4* s = reinterpret_cast<const int4*>(p);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 16); ++i) raw[i] = s[i];
  if (PIN) {
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 16); ++i)
      asm volatile("" : : "v"(raw[i].x), "v"(raw[i].y), "v"(raw[i].z), "v"(raw[i].w));
  }
  T v;
  __builtin_memcpy(&v, raw, sizeof(T));
  return v;
}

// Pairs found by a wave are staged in LDS and written out in batches: one
// global atomic per flush instead of one per emitting step.  768 pairs keep a
// block at 26 KB of LDS, so five blocks (the kernel's register limit) share a CU.
constexpr int TR_MAX_WAVES = kTraverseMaxWaves, TR_BUF = 768;

// Called by the whole wave (converged).
__device__ __forceinline__ void flush_pairs(const int2* buf, int cnt, int2* __restrict__ pairs, unsigned long long cap,
                                            unsigned long long* counter) {
  const int lane = lane_id();
  wave_lds_sync();
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(counter, (unsigned long long)cnt);
  base = __shfl(base, 0);
  for (int i = lane; i < cnt; i += kWave)  // streamed: the walk's nodes stay in cache
    if (base + i < cap) {
      const int2 v = buf[i];
      __builtin_nontemporal_store(((long long)(unsigned)v.y << 32) | (unsigned)v.x,
                                  reinterpret_cast<long long*>(pairs + base + i));
    }
  wave_lds_sync();
}

// Packet traversal: the 64 lanes of a wave hold 64 consecutive sorted leaves
// (queries) and walk the tree together, depth first from the root with a
// wave-uniform stack.  Each step reads the child-pair record of one internal
// node (both children, one 128 B line, at a wave-uniform address: scalar
// loads) and culls each lane's query against both children; a child is
// entered if any lane wants it.  Lane j takes part only where a child covers
// a sorted leaf in (s_j, last_j] (last_j: the last leaf of its scene, n-1 for
// one scene), so each unordered leaf pair is found once, by its lower leaf.
constexpr int TR_STACK = 96;  // deeper than any LBVH over 32-bit keys and int indices

// Leaf children and internal children covering at most TR_BUCKET sorted leaves
// are not walked: their leaves (contiguous in sorted order) are tested in the
// same step, one scalar-loaded FP64 box each.  A step's fixed work (the
// child-pair load, ballots, stack) costs more than a few leaf tests: at 1M
// triangles the traversal (then testing leaves on FP32 bounds) took 11.7 ms
// without buckets, 10.8 / 10.2 / 10.1 / 10.1 ms with buckets of 4 / 8 / 16 /
// 32 leaves (12.1 ms with 2); with the FP64 leaf records, 10.5 / 9.6 / 9.3 /
// 9.0 / 9.2 / 9.3 ms with 4 / 8 / 16 / 32 / 64 / 128.
constexpr int TR_BUCKET = 32;

// Stages the lanes' pairs (s, k) of one leaf k; flushes a full buffer.
__device__ __forceinline__ void stage_pairs(bool emit, int s, int k, int2* buf, int& cnt, int2* __restrict__ pairs,
                                            unsigned long long cap, unsigned long long* counter) {
  const uint64_t m = __ballot(emit);
  if (m == 0) return;
  if (emit) buf[cnt + popc_below(m)] = make_int2(s, k);
  cnt += __popcll(m);
  if (cnt > TR_BUF - kWave) {
    flush_pairs(buf, cnt, pairs, cap, counter);
    cnt = 0;
  }
}

template <int D>
// Six waves per SIMD (80 VGPRs): the register target keeps the two-leaf bucket
// loop below from costing a wave (81 VGPRs otherwise; 4 SGPRs then spill to a
// VGPR's lanes).
__global__ __launch_bounds__(TR_MAX_WAVES * kWave) __attribute__((amdgpu_waves_per_eu(6)))
void k_traverse(const NodeF<D>* __restrict__ nodes, const ChildPair<D>* __restrict__ cp, const int* __restrict__ root,
                const SortedLeaf<D>* __restrict__ sboxes, int n, int2* __restrict__ pairs, unsigned long long cap,
                unsigned long long* counters, const uint32_t* __restrict__ scene_of_sorted,
                const int* __restrict__ scene_off) {
  __shared__ int2 s_buf[TR_MAX_WAVES][TR_BUF];
  __shared__ int s_stack[TR_MAX_WAVES][TR_STACK];
  const int w = threadIdx.x / kWave, lane = lane_id();
  const int s = xcd_block() * blockDim.x + threadIdx.x;
  const bool valid = s < n - 1;
  if (__ballot(valid) == 0) return;  // wave-uniform: the wave starts at or after the last leaf
  Box01<D> qd;
  bool qa = false, qemp = false;
  int last = -1;
  if (valid) {
    const int2 cm = *reinterpret_cast<const int2*>(&sboxes[s].child);
    qd = load16(box01<D>(sboxes + s));
    qa = ((uint32_t)cm.y >> 31) != 0;
    qemp = cm.x != 0;
    last = scene_off ? scene_off[scene_of_sorted[s] + 1] - 1 : n - 1;
  }
  // The query is in registers before the walk, so that no wait inside the walk
  // is for a vector load other than the bucket prefetches below.
  __builtin_amdgcn_s_waitcnt(0);
  int wave_last = last;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int v = __shfl_xor(wave_last, o);
    wave_last = v > wave_last ? v : wave_last;
  }
  wave_last = __builtin_amdgcn_readfirstlane(wave_last);
  int p = *root;  // wave-uniform
  int sp = 0;     // stack depth (uniform)
  // each internal node is entered at most once: more steps means a malformed tree
  int budget = n;
  int cnt = 0;  // staged pairs (wave-uniform)
  unsigned steps = 0;
  uint32_t pf = 0;  // bucket prefetches: their values are only consumed after the walk
  for (;;) {
    ++steps;
    if (--budget < 0) {
      if (lane == 0) atomicAdd(counters + C_ERRORS, 1ull);
      break;
    }
    const ChildPair<D> pr = cp[p];
    bool g[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const NodeF<D>& c = pr.c[j];
      const bool want = valid && c.first <= wave_last && c.last > s && c.first <= last && cull_q64<D>(qd, qa, qemp, c);
      g[j] = __ballot(want) != 0;
    }
    // leaves and small subtrees: their sorted leaves are tested here, on the
    // FP64 t0 / t1 boxes (exact; the swept overlap is implied unless a box is empty)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const NodeF<D>& c = pr.c[j];
      if (!g[j] || c.last - c.first >= TR_BUCKET) continue;
      // The bucket's records (the two 64 B lines the test reads, per leaf) are
      // pulled into L2 by one vector load before the scalar loads below ask for
      // them one leaf at a time: vector loads wait on their own counter, so
      // the scalar loads' waits do not include them, and the scalar cache's
      // misses then hit L2 (1M triangles: 9.0 -> 8.6 ms).
      if (lane < 2 * (c.last - c.first + 1))
        pf = reinterpret_cast<const uint32_t*>(sboxes + c.first + (lane >> 1))[(lane & 1) * 16];
      // two leaves per iteration: both records' scalar loads are issued before
      // one wait, so a bucket waits on half as many scalar-cache round trips
      // (8.5 -> 8.1 ms; three or more would not fit the scalar registers)
      auto test = [&](int k, const Box01<D>& B, uint32_t bm, int bemp) {
        const bool in = valid & (k > s) & (k <= last);
        const bool ba = (bm >> 31) != 0;
        bool emit = in & (qa | ba) & !separated<D>(qd, B);
        const bool slow = in & (qemp | (bemp != 0));
        if (__ballot(slow) != 0) {
          if (slow) emit = leaf_test_fp64<D>(sboxes + s, qa, sboxes + k, ba);
        }
        stage_pairs(emit, s, k, s_buf[w], cnt, pairs, cap, counters + C_LEAF_PAIRS);
      };
      for (int k = c.first; k <= c.last; k += 2) {
        const int k1 = k < c.last ? k + 1 : k;
        const SortedLeaf<D>& L0 = sboxes[k];  // two 64 B lines each
        const SortedLeaf<D>& L1 = sboxes[k1];
        const Box01<D> B0 = *box01<D>(&L0), B1 = *box01<D>(&L1);
        const uint32_t bm0 = L0.meta, bm1 = L1.meta;
        const int be0 = L0.child, be1 = L1.child;
        test(k, B0, bm0, be0);
        if (k < c.last) test(k1, B1, bm1, be1);
      }
      g[j] = false;
    }
    const int ia = (int)(pr.c[0].meta & 0x7fffffffu) - 1, ib = (int)(pr.c[1].meta & 0x7fffffffu) - 1;
    if (g[0]) {
      if (g[1]) {  // enter the first child, the second waits on the stack
        if (sp == TR_STACK) {
          if (lane == 0) atomicAdd(counters + C_ERRORS, 1ull);
          break;
        }
        if (lane == 0) s_stack[w][sp] = ib;
        ++sp;
      }
      p = ia;
    } else if (g[1]) {
      p = ib;
    } else {
      if (sp == 0) break;
      --sp;
      wave_lds_sync();
      p = __builtin_amdgcn_readfirstlane(s_stack[w][sp]);
    }
  }
  if (cnt > 0) flush_pairs(s_buf[w], cnt, pairs, cap, counters + C_LEAF_PAIRS);
  asm volatile("" ::"v"(pf));  // keeps the prefetches (a use the compiler cannot remove)
  if (lane == 0) atomicAdd(counters + C_TR_STEPS, (unsigned long long)steps);
}

void launch_traverse(int dim, const void* nodes, const void* cp, const int* root, const void* sboxes, int n,
                     int2* pairs, unsigned long long cap, unsigned long long* counters,
                     const uint32_t* scene_of_sorted, const int* scene_off, int waves, hipStream_t s) {
  if (n < 2) return;
  waves = waves < 1 ? 4 : (waves > TR_MAX_WAVES ? TR_MAX_WAVES : waves);
  const int g = xcd_grid(n - 1, waves * kWave);
  if (dim == 3)
    k_traverse<3><<<g, waves * kWave, 0, s>>>((const NodeF<3>*)nodes, (const ChildPair<3>*)cp, root,
                                                 (const SortedLeaf<3>*)sboxes, n, pairs, cap, counters, scene_of_sorted,
                                                 scene_off);
  else
    k_traverse<2><<<g, waves * kWave, 0, s>>>((const NodeF<2>*)nodes, (const ChildPair<2>*)cp, root,
                                                 (const SortedLeaf<2>*)sboxes, n, pairs, cap, counters, scene_of_sorted,
                                                 scene_off);
}

// ---------------------------------------------------------------------------
// Face arrays and pair conversion
// ---------------------------------------------------------------------------
template <int D>
__global__ void k_face_records(const uint32_t* __restrict__ order, int n, const int* __restrict__ P,
                               const uint8_t* __restrict__ flags, const int* __restrict__ rank, FaceArrays F) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int prim = (int)order[s];
  const int* p = P + (size_t)prim * D;
  const int fl = (int)flags[prim];
  F.head[s] = D == 3 ? make_int4(p[0], p[1], p[D - 1], fl) : make_int4(p[0], p[1], fl, 0);
  F.rank[s] = rank ? rank[prim] : prim;
  F.prim[s] = prim;
}

void launch_face_records(int dim, const uint32_t* order, int n, const int* P, const uint8_t* flags, const int* rank,
                         const FaceArrays& faces, hipStream_t s) {
  if (n < 1) return;
  const int g = (n + 255) / 256;
  if (dim == 3) k_face_records<3><<<g, 256, 0, s>>>(order, n, P, flags, rank, faces);
  else k_face_records<2><<<g, 256, 0, s>>>(order, n, P, flags, rank, faces);
}

__global__ void k_pairs_to_prims(int2* pairs, long long n, FaceArrays F) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int2 pr = pairs[i];
  const int ra = F.rank[pr.x], rb = F.rank[pr.y];
  const int a = F.prim[pr.x], b = F.prim[pr.y];
  pairs[i] = ra < rb ? make_int2(a, b) : make_int2(b, a);
}

void launch_pairs_to_prims(int2* pairs, long long n, const FaceArrays& faces, hipStream_t s) {
  if (n < 1) return;
  const unsigned g = (unsigned)((n + 255) / 256);
  k_pairs_to_prims<<<g, 256, 0, s>>>(pairs, n, faces);
}

// ---------------------------------------------------------------------------
// Narrow phase
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// k_expand: per leaf pair (BVHTree.cpp:252-283) the discrete test at V1 and the
// representative-triangle candidates, appended to a VF and an EE stencil queue.
// ---------------------------------------------------------------------------
struct ExpandArgs {
  const int2* pairs;                     // sorted positions
  const unsigned long long* npairs_dev;  // pairs found by the traversal (device counter)
  unsigned long long pair_cap;
  FaceArrays faces;                      // leaf arrays by sorted position, V1 by vertex
  int discrete, continuous;
  int4* vfq;
  unsigned long long vf_cap;
  int4* eeq;
  unsigned long long ee_cap;
  int2* disc;
  unsigned long long disc_cap;
  int2* odep;  // non-null: pairs whose discrete result depends on the argument order
  unsigned long long odep_cap;
  unsigned long long* counters;
  // per 64-pair chunk: VF / EE candidate counts, then (scanned) output offsets;
  // entry nchunks_cap holds the totals after the scan
  unsigned long long* off_vf;
  unsigned long long* off_ee;
  long long nchunks_cap;
  int chunk;  // consecutive 64-item chunks a wave takes per grid-stride step
};

constexpr int EX_WAVES = 4, EX_DSTAGE = 128;

// Pass 1: candidate counts per 64-pair chunk (zero past the last pair).  A
// wave takes EC_U consecutive chunks per step and issues all their pair loads,
// then all their head loads, before using any: the pass is bound by the
// latency of those two dependent loads, not by bandwidth.
constexpr int EC_U = 4;
template <int D> __global__ __launch_bounds__(EX_WAVES * kWave) void k_expand_count(ExpandArgs A) {
  const int w = threadIdx.x / kWave, lane = lane_id();
  const long long npairs = (long long)(*A.npairs_dev < A.pair_cap ? *A.npairs_dev : A.pair_cap);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    A.off_vf[A.nchunks_cap] = 0;
    A.off_ee[A.nchunks_cap] = 0;
  }
  const long long stride = (long long)gridDim.x * EX_WAVES * EC_U;
  for (long long c0 = ((long long)xcd_block() * EX_WAVES + w) * EC_U; c0 < A.nchunks_cap; c0 += stride) {
    // unconditional loads (indices clamped into the list; the selects come after
    // every load was issued), so that all of them are in flight together
    int2 sp[EC_U];
    bool in[EC_U];
#pragma unroll
    for (int u = 0; u < EC_U; ++u) {
      const long long i = (c0 + u) * kWave + lane;
      in[u] = A.continuous && i < npairs;
      sp[u] = A.pairs[i & -(long long)(i < npairs)];  // masks, not selects: no branch per load
    }
    int4 ha[EC_U], hb[EC_U];
#pragma unroll
    for (int u = 0; u < EC_U; ++u) {  // counts do not depend on orientation: the heads suffice
      const int m = -(int)in[u];
      ha[u] = A.faces.head[sp[u].x & m];
      hb[u] = A.faces.head[sp[u].y & m];
    }
#pragma unroll
    for (int u = 0; u < EC_U; ++u) {
      uint32_t c = 0;  // VF count in the low half, EE count in the high half (at most 6 and 9 per lane)
      if (in[u]) {
        uint32_t vfm, eem;
        candidate_masks<D>(face_ids<D>(ha[u]), face_ids<D>(hb[u]), vfm, eem);
        c = (uint32_t)__popc(vfm) | ((uint32_t)__popc(eem) << 16);
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
      const long long ch = c0 + u;
      if (lane == 0 && ch < A.nchunks_cap) {
        A.off_vf[ch] = c & 0xffffu;
        A.off_ee[ch] = c >> 16;
      }
    }
  }
}

__global__ void k_expand_totals(const unsigned long long* off_vf, const unsigned long long* off_ee, long long nch,
                                unsigned long long* counters) {
  counters[C_CAND_VF] = off_vf[nch];
  counters[C_CAND_EE] = off_ee[nch];
}

// A wave's stencils of one chunk, staged in LDS in lane order and written out
// by consecutive lanes (coalesced) at the chunk's scanned offset.  Vertex ids
// are plain values (a0..a2 of the first face, b0..b2 of the second): indexing
// an array, or selecting between two, would place them in scratch memory.
template <int D, bool VF>
__device__ __forceinline__ void write_stencils(int4* stage, uint32_t mask, int a0, int a1, int a2, int b0, int b1,
                                               int b2, int4* __restrict__ dst, unsigned long long base,
                                               unsigned long long cap) {
  const uint32_t count = (uint32_t)__popc(mask);
  uint32_t p = wave_prefix_u32(count);
  const uint32_t total = __shfl(p + count, kWave - 1);
  while (mask) {
    const int c = __ffs(mask) - 1;
    mask &= mask - 1;
    stage[p++] = candidate_stencil<D, VF>(c, a0, a1, a2, b0, b1, b2);
  }
  wave_lds_sync();
  for (uint32_t j = lane_id(); j < total; j += kWave)  // streamed: read back by the ACCD kernels from HBM
    if (base + j < cap) {
      const int4 v = stage[j];
      __builtin_nontemporal_store(nt_int4{v.x, v.y, v.z, v.w}, reinterpret_cast<nt_int4*>(dst + base + j));
    }
  wave_lds_sync();
}

// The t-th chunk a wave takes: A.chunk consecutive chunks per grid-stride step.
__device__ __forceinline__ long long expand_chunk(const ExpandArgs& A, int w, long long t) {
  const long long stride = (long long)gridDim.x * EX_WAVES * A.chunk;
  return ((long long)xcd_block() * EX_WAVES + w) * A.chunk + (t / A.chunk) * stride + t % A.chunk;
}

// Pass 2: the discrete test and the candidates, written at the scanned offsets
// (each chunk's stencils in lane order: the queues follow the pair list).
// ODEP: also evaluate the discrete test with the two primitives swapped and
// list the pairs whose result changes (its staging buffer only exists then,
// so the default kernel keeps the LDS size that allows four blocks per CU).
// Each chunk's pairs are loaded one chunk ahead, before the stores they would
// otherwise queue behind.
template <int D, bool ODEP> __global__ __launch_bounds__(EX_WAVES * kWave) void k_expand(ExpandArgs A) {
  __shared__ int2 s_d[EX_WAVES][EX_DSTAGE];
  __shared__ int2 s_o[EX_WAVES][ODEP ? EX_DSTAGE : 1];
  __shared__ int4 s_st[EX_WAVES][kWave * 9];  // at most 9 EE (6 VF) stencils per pair
  const int w = threadIdx.x / kWave, lane = lane_id();
  int nd = 0, no = 0;  // staged discrete / order-dependent records (wave-uniform)
  const long long npairs = (long long)(*A.npairs_dev < A.pair_cap ? *A.npairs_dev : A.pair_cap);
  const long long nchunks = (npairs + kWave - 1) / kWave;
  long long ch = expand_chunk(A, w, 0);
  int2 next = make_int2(0, 0);
  if (ch < nchunks && ch * kWave + lane < npairs) next = A.pairs[ch * kWave + lane];
  for (long long t = 0; ch < nchunks; ++t) {
    const long long i = ch * kWave + lane;
    const bool valid = i < npairs;
    const int2 sp = next;
    const long long cn = expand_chunk(A, w, t + 1);
    if (cn < nchunks && cn * kWave + lane < npairs) next = A.pairs[cn * kWave + lane];
    FaceIds<D> ha = {}, hb = {};
    int pa = 0, pb = 0;
    if (valid) pair_faces<D>(A.faces, sp, ha, hb, pa, pb);
    if (A.discrete) {  // BVHTree.cpp:464-500
      int dh, dh_swap;
      discrete_test<D, ODEP>(A.faces.v1, ha, hb, valid, dh, dh_swap);
      int2 rec = make_int2(0, 0);
      if (dh || dh_swap) rec = make_int2(A.faces.prim[pa], A.faces.prim[pb]);
      stage_push<int2, EX_DSTAGE>(s_d[w], nd, dh != 0, rec, A.disc, A.disc_cap, A.counters + C_DISCRETE);
      if constexpr (ODEP)
        stage_push<int2, EX_DSTAGE>(s_o[w], no, dh != dh_swap, rec, A.odep, A.odep_cap, A.counters + C_ORDER_DEP);
    }
    if (!A.continuous) {
      ch = cn;
      continue;
    }
    uint32_t vfm = 0, eem = 0;
    if (valid) candidate_masks<D>(ha, hb, vfm, eem);
    const int x0 = ha.p[0], x1 = ha.p[1], x2 = ha.p[D - 1], y0 = hb.p[0], y1 = hb.p[1], y2 = hb.p[D - 1];
    // the chunk's offsets (written by the scan) through the scalar cache: the EE
    // offset then does not wait for the VF stencils' stores (1M triangles:
    // 8.5 -> 8.1 ms)
    const long long chu = uniform_index(ch);
    const unsigned long long ov = scalar_load(A.off_vf + chu);
    const unsigned long long oe = D == 3 ? scalar_load(A.off_ee + chu) : 0ull;
    write_stencils<D, true>(s_st[w], vfm, x0, x1, x2, y0, y1, y2, A.vfq, ov, A.vf_cap);
    if constexpr (D == 3)
      write_stencils<D, false>(s_st[w], eem, x0, x1, x2, y0, y1, y2, A.eeq, oe, A.ee_cap);
    ch = cn;
  }
  if (nd) stage_flush(s_d[w], nd, A.disc, A.disc_cap, A.counters + C_DISCRETE);
  if (ODEP && no) stage_flush(s_o[w], no, A.odep, A.odep_cap, A.counters + C_ORDER_DEP);
}

// ---------------------------------------------------------------------------
// k_accd: FP64 ACCD (NarrowPhase.cpp:634-711) over a packed stencil queue, one
// stencil per lane; hits staged in LDS.  VF and EE queues run as separate
// launches so every SIMD step executes one distance kernel.
// ---------------------------------------------------------------------------
struct AccdArgs {
  const int4* q;
  const unsigned long long* n_dev;  // queue length (device counter)
  unsigned long long cap;
  const double* rec;
  double eta;
  const int* vert_scene;    // several scenes: per-scene gaps
  const double* scene_eta;
  double* cand_toi;         // per-stencil toi (-1 = no hit); may be null with hit_sten
  unsigned long long* chunk_hits;  // hits per 64-stencil chunk; may be null with hit_sten
  unsigned long long* counters;
  // direct hit list (null: the compaction pass builds it from cand_toi)
  int4* hit_sten;
  uint8_t* hit_vf;
  double* hit_toi;
  unsigned long long hit_cap;
  int f32;                  // float-precision tree (CTCD: TOI rounded to float)
  int suspend;               // suspend stencils still running after three iterations (see k_accd)
  uint32_t* sel;             // 3D CTCD: queue indices of the stencils that pass the prefilter (null: none)
};

constexpr int AC_WAVES = 4;

// Direct hit list: each wave stages its hits (queue index, toi) in LDS and
// appends them with one counter atomic per flush of up to N hits.  3D:
// 4 x 448 x 12 B = 21 KB per block, which with the suspended-stencil stacks
// (18 KB) keeps the four blocks per CU that the kernels' registers allow; 2D
// (fewer registers, five blocks) stages 256.  Queue indices are 32-bit: the
// host only asks for the direct list when both queues are shorter than 2^32.
template <int D> constexpr int ac_stage() { return D == 3 ? 448 : 256; }  // multiples of 64

// Per-wave stack of suspended stencils (Accd::resume state), SoA in LDS.
// A wave resumes 64 of them as soon as it holds 64, so it never holds more
// than 63 + 64.
constexpr int kSuspendCap = 2 * kWave;
struct SuspendStack {
  uint32_t idx[AC_WAVES][kSuspendCap];
  double t1[AC_WAVES][kSuspendCap], t2[AC_WAVES][kSuspendCap], t3[AC_WAVES][kSuspendCap], g[AC_WAVES][kSuspendCap];
};

template <int N> struct HitStage {
  uint32_t idx[AC_WAVES][N];
  double toi[AC_WAVES][N];
};

template <bool VF, int N>
__device__ __forceinline__ void flush_hits(HitStage<N>& S, int w, int& cnt, const AccdArgs& A);

// Called by the whole wave (converged) after each 64-stencil chunk.
template <bool VF, int N>
__device__ __forceinline__ void stage_hits(HitStage<N>& S, int w, int& cnt, bool hit, long long i, double toi,
                                           const AccdArgs& A, bool last) {
  const uint64_t m = __ballot(hit);
  if (hit) {
    const int p = cnt + popc_below(m);
    S.idx[w][p] = (uint32_t)i;
    S.toi[w][p] = toi;
  }
  cnt += __popcll(m);
  if (cnt == 0 || (cnt <= N - kWave && !last)) return;
  flush_hits<VF>(S, w, cnt, A);
}

template <bool VF, int N>
__device__ __forceinline__ void flush_hits(HitStage<N>& S, int w, int& cnt, const AccdArgs& A) {
  const int lane = lane_id();
  wave_lds_sync();
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(A.counters + C_HITS, (unsigned long long)cnt);
  base = __shfl(base, 0);
  for (int k = lane; k < cnt; k += kWave)
    if (base + k < A.hit_cap) {
      const unsigned long long pos = base + k;
      const int4 sv = A.q[S.idx[w][k]];  // read by this wave a few chunks ago
      __builtin_nontemporal_store(nt_int4{sv.x, sv.y, sv.z, sv.w}, reinterpret_cast<nt_int4*>(A.hit_sten + pos));
      __builtin_nontemporal_store(S.toi[w][k], A.hit_toi + pos);
      __builtin_nontemporal_store((uint8_t)(VF ? 1 : 0), A.hit_vf + pos);
    }
  wave_lds_sync();
  cnt = 0;
}

// One lane per stencil, 64 consecutive stencils per wave step; the TOI (or
// -1) goes to cand_toi and the chunk's hit count to chunk_hits.  With
// A.hit_sten the hits are also appended to the hit list (staged per wave,
// one counter atomic per flush), so the compaction pass only runs when the
// list overflowed hit_cap.
template <typename T> struct VecOf;
template <> struct VecOf<double> {
  template <int D> using V = typename VLoad<D>::V;
};
template <> struct VecOf<float> {
  template <int D> using V = typename std::conditional<D == 3, vec3<float>, vec2<float>>::type;
};
__device__ __forceinline__ void narrow(const d3& a, vec3<float>& o) { o = {(float)a.x, (float)a.y, (float)a.z}; }
__device__ __forceinline__ void narrow(const d2& a, vec2<float>& o) { o = {(float)a.x, (float)a.y}; }
__device__ __forceinline__ void narrow(const d3& a, d3& o) { o = a; }
__device__ __forceinline__ void narrow(const d2& a, d2& o) { o = a; }

// A stencil's vertices at t0 and t1 (narrowed to T) and its gap.
template <int D, typename T, typename VT>
__device__ __forceinline__ double load_stencil(const AccdArgs& A, const int4 sv, VT* v0, VT* v1) {
  using V = typename VLoad<D>::V;
  const int id[4] = {sv.x, sv.y, sv.z, sv.w};
#pragma unroll
  for (int k = 0; k < D + 1; ++k) {
    V a, b;
    VLoad<D>::load(A.rec, id[k], a, b);
    narrow(a, v0[k]);
    narrow(b, v1[k]);
  }
  return A.vert_scene ? A.scene_eta[A.vert_scene[sv.x]] : A.eta;
}

// T = double: the FP64 path; T = float: BVHTree<float,DIM>'s float ACCD on the
// float-valued vertex records of a float-precision tree (exact conversions).
//
// Lock-step waves run every stencil of a chunk for at most three distance
// evaluations (most stencils need exactly three); the stencils still running
// then are pushed on the wave's LDS stack of suspended stencils, and as soon as
// it holds 64 the wave resumes them as one chunk (their vertex records were
// read a few chunks earlier and are still cached).  A chunk therefore no
// longer waits for its longest stencil.  Stencils whose iteration 0 was not
// the no-op step run to completion in place; the results are those of the
// uninterrupted loop (Accd::resume).  Without suspension (A.suspend = 0, as
// whenever per-chunk hit counts are written) every stencil runs in place.
template <int D, bool VF, typename T>
__global__ __launch_bounds__(AC_WAVES * kWave) __attribute__((amdgpu_waves_per_eu(4))) void k_accd(AccdArgs A) {
  using VT = typename VecOf<T>::template V<D>;
  __shared__ HitStage<ac_stage<D>()> S;
  __shared__ SuspendStack Q;
  const int w = threadIdx.x / kWave, lane = lane_id();
  long long evals = 0;
  int staged = 0;
  int held = 0;  // suspended stencils on the wave's stack (wave-uniform)
  unsigned pushed = 0;  // stencils this wave suspended (wave-uniform)
  const bool sus = A.suspend != 0;
  const long long n = (long long)(*A.n_dev < A.cap ? *A.n_dev : A.cap);
  const long long nchunks = (n + kWave - 1) / kWave;
  const long long stride = (long long)gridDim.x * AC_WAVES;
  long long ch = (long long)xcd_block() * AC_WAVES + w;
  while (ch < nchunks || held > 0) {
    // the next chunk: 64 suspended stencils once the stack holds them (or the
    // rest of the stack after the last fresh chunk), else 64 fresh stencils
    const bool resume = held >= kWave || ch >= nchunks;
    long long i = 0;
    bool valid;
    int st = ACCD_MISS;
    Accd<D, VF, T> a;
    if (resume) {
      const int cnt = held < kWave ? held : kWave;
      held -= cnt;
      valid = lane < cnt;
      if (valid) {
        const int k = held + lane;  // the top cnt entries
        i = (long long)Q.idx[w][k];
        VT v0[D + 1], v1[D + 1];
        const double eta = load_stencil<D, T>(A, A.q[i], v0, v1);
        a.resume(v0, v1, (T)eta, (T)Q.t1[w][k], (T)Q.t2[w][k], (T)Q.t3[w][k], (T)Q.g[w][k]);
        st = ACCD_RUN;
      }
      wave_lds_sync();  // the entries are read before later chunks push over them
    } else {
      i = ch * kWave + lane;
      valid = i < n;
      if (valid) {
        VT v0[D + 1], v1[D + 1];
        const double eta = load_stencil<D, T>(A, A.q[i], v0, v1);
        st = a.template start<kAccdSuspendIter>(v0, v1, (T)eta);
      }
      if (sus) {
        const bool push = st == ACCD_RUN && a.replayable;
        const uint64_t m = __ballot(push);
        if (push) {
          const int k = held + popc_below(m);
          Q.idx[w][k] = (uint32_t)i;
          Q.t1[w][k] = (double)a.t1;
          Q.t2[w][k] = (double)a.t2;
          Q.t3[w][k] = (double)a.tl;
          Q.g[w][k] = (double)a.g;
          st = ACCD_SUSPENDED;
        }
        held += __popcll(m);
        pushed += (unsigned)__popcll(m);
        wave_lds_sync();
      }
    }
    if (st == ACCD_RUN) st = a.run();
    const bool hit = st == ACCD_HIT;
    const double toi = hit ? (double)a.toi : -1.0;
    if (valid) {
      evals += a.evals;
      if (A.cand_toi && st != ACCD_SUSPENDED) __builtin_nontemporal_store(toi, A.cand_toi + i);
    }
    if (!resume) {
      if (A.chunk_hits) {  // only without suspension (the launcher's rule)
        const uint64_t m = __ballot(hit);
        if (lane == 0) A.chunk_hits[ch] = (unsigned long long)__popcll(m);
      }
      ch += stride;
    }
    if (A.hit_sten) stage_hits<VF>(S, w, staged, hit, i, toi, A, false);
  }
  if (A.hit_sten && staged) flush_hits<VF>(S, w, staged, A);
  // one atomic per block and counter (same-address atomics serialise)
  __shared__ long long s_ev[AC_WAVES];
  __shared__ unsigned s_sus[AC_WAVES];
  for (int o = 32; o > 0; o >>= 1) evals += __shfl_xor(evals, o);
  if (lane == 0) {
    s_ev[w] = evals;
    s_sus[w] = pushed;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sus = pushed;
    for (int k = 1; k < AC_WAVES; ++k) {
      evals += s_ev[k];
      sus += s_sus[k];
    }
    if (evals) {
      atomicAdd(A.counters + C_DIST_EVALS, (unsigned long long)evals);
      if (!VF) atomicAdd(A.counters + C_EE_EVALS, (unsigned long long)evals);
    }
    if (sus) atomicAdd(A.counters + C_SUSPENDED, sus);
  }
}

// CTCD vertex-face narrow phase (options.vf_one_sided; BVHTree.cpp:446-447):
// the VF queue, outputs and per-chunk hit counts of the VF ACCD kernel.  cosr,
// sinr: the root finder's shift rotation, cos/sin(94 * 0.017453293).  SEL: only
// the stencils listed in A.sel (counters[C_CTCD_SEL] of them) by the selection
// pass; lock-step waves then hold no lane that the prefilter would retire at
// once while the others run the root finder.
template <int D, bool SEL>
__global__ __launch_bounds__(AC_WAVES * kWave) void k_ctcd(AccdArgs A, double cosr, double sinr) {
  using V = typename VLoad<D>::V;
  __shared__ HitStage<ac_stage<D>()> S;
  const int w = threadIdx.x / kWave, lane = lane_id();
  int staged = 0;
  const unsigned long long cnt = SEL ? A.counters[C_CTCD_SEL] : *A.n_dev;
  const long long n = (long long)(cnt < A.cap ? cnt : A.cap);
  const long long nchunks = (n + kWave - 1) / kWave;
  const long long stride = (long long)gridDim.x * AC_WAVES;
  for (long long ch = (long long)xcd_block() * AC_WAVES + w; ch < nchunks; ch += stride) {
    const long long j = ch * kWave + lane;
    long long i = j;
    int hit = 0;
    double toi = -1.0;
    if (j < n) {
      if (SEL) i = (long long)A.sel[j];
      const int4 sv = A.q[i];
      const int id[4] = {sv.x, sv.y, sv.z, sv.w};
      V v0[D + 1], v1[D + 1];
#pragma unroll
      for (int k = 0; k < D + 1; ++k) VLoad<D>::load(A.rec, id[k], v0[k], v1[k]);
      const double eta = A.vert_scene ? A.scene_eta[A.vert_scene[sv.x]] : A.eta;
      hit = ctcd::Query<D>::run(v0, v1, eta, toi, cosr, sinr) == 1;
      if (A.f32 && hit) toi = (double)(float)toi;  // NarrowPhase.cpp:425-444: t_impact = t (a float)
      if (A.cand_toi) __builtin_nontemporal_store(hit ? toi : -1.0, A.cand_toi + i);
    }
    if (!SEL) {
      const uint64_t m = __ballot(hit != 0);
      if (lane == 0 && A.chunk_hits) A.chunk_hits[ch] = (unsigned long long)__popcll(m);
    }
    if (A.hit_sten) stage_hits<true>(S, w, staged, hit != 0, i, toi, A, ch + stride >= nchunks);
  }
}

// The selection pass of the 3D CTCD: the prefilter per VF stencil; the ones
// that pass are appended to A.sel (per-wave LDS staging, one counter atomic per
// flush), the others are misses.
constexpr int CT_SEL_STAGE = 512;
__global__ __launch_bounds__(AC_WAVES * kWave) void k_ctcd_select(AccdArgs A) {
  using V = typename VLoad<3>::V;
  __shared__ uint32_t s_sel[AC_WAVES][CT_SEL_STAGE];
  const int w = threadIdx.x / kWave, lane = lane_id();
  int staged = 0;
  const long long n = (long long)(*A.n_dev < A.cap ? *A.n_dev : A.cap);
  const long long nchunks = (n + kWave - 1) / kWave;
  const long long stride = (long long)gridDim.x * AC_WAVES;
  for (long long ch = (long long)xcd_block() * AC_WAVES + w; ch < nchunks; ch += stride) {
    const long long i = ch * kWave + lane;
    bool pass = false;
    if (i < n) {
      const int4 sv = A.q[i];
      const int id[4] = {sv.x, sv.y, sv.z, sv.w};
      V v0[4], v1[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) VLoad<3>::load(A.rec, id[k], v0[k], v1[k]);
      const double eta = A.vert_scene ? A.scene_eta[A.vert_scene[sv.x]] : A.eta;
      pass = ctcd::Query<3>::prefilter(v0, v1, eta);
      if (!pass && A.cand_toi) __builtin_nontemporal_store(-1.0, A.cand_toi + i);
    }
    stage_push<uint32_t, CT_SEL_STAGE>(s_sel[w], staged, pass, (uint32_t)i, A.sel, A.cap, A.counters + C_CTCD_SEL);
  }
  if (staged) stage_flush(s_sel[w], staged, A.sel, A.cap, A.counters + C_CTCD_SEL);
}

__global__ void k_hits_total(const unsigned long long* off, long long last, unsigned long long* counters) {
  counters[C_HITS] = off[last];
}

// Each chunk's hits in lane order at off[ch].  A wave takes HC_U consecutive
// 64-stencil chunks per step and issues all their loads before any store, so
// each wave keeps several independent loads in flight.  (Plain accesses:
// non-temporal loads or stores made this kernel slower.)
constexpr int HC_U = 8;

template <bool VF>
__global__ __launch_bounds__(256) void k_hits_compact(const int4* __restrict__ q, const unsigned long long* n_dev,
                                                      unsigned long long cap, const double* __restrict__ toi,
                                                      const unsigned long long* __restrict__ off, int4* hit_sten,
                                                      uint8_t* hit_vf, double* hit_toi, unsigned long long hit_cap) {
  const int lane = lane_id();
  const long long n = (long long)(*n_dev < cap ? *n_dev : cap);
  const long long nchunks = (n + kWave - 1) / kWave;
  const long long stride = (long long)gridDim.x * (blockDim.x / kWave) * HC_U;
  for (long long c0 = ((long long)xcd_block() * (blockDim.x / kWave) + threadIdx.x / kWave) * HC_U; c0 < nchunks;
       c0 += stride) {
    double t[HC_U];
    int4 st[HC_U];
    unsigned long long base[HC_U];
#pragma unroll
    for (int u = 0; u < HC_U; ++u) {
      const long long i = (c0 + u) * kWave + lane;
      const bool in = i < n;
      t[u] = in ? toi[i] : -1.0;
      st[u] = in ? q[i] : make_int4(0, 0, 0, 0);
      base[u] = c0 + u < nchunks ? off[c0 + u] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < HC_U; ++u) {
      const bool hit = t[u] >= 0.0;
      const uint64_t m = __ballot(hit);
      const unsigned long long pos = base[u] + (unsigned long long)popc_below(m);
      if (hit && pos < hit_cap) {
        hit_sten[pos] = st[u];
        hit_toi[pos] = t[u];
        hit_vf[pos] = VF ? 1 : 0;
      }
    }
  }
}

static int grid_for(long long chunks, int waves, int num_cus, int per_cu) {
  long long b = (chunks + waves - 1) / waves;
  const long long maxb = (long long)num_cus * per_cu;
  if (b > maxb) b = maxb;
  if (b < 1) b = 1;
  return (int)((b + 7) / 8 * 8);
}

long long expand_chunks(unsigned long long pair_cap) { return (long long)((pair_cap + kWave - 1) / kWave); }

void launch_expand(int dim, const ExpandLaunch& L, hipStream_t s) {
  ExpandArgs A;
  A.pairs = L.pairs; A.npairs_dev = L.npairs_dev; A.pair_cap = L.pair_cap; A.faces = L.faces;
  A.discrete = L.discrete; A.continuous = L.continuous;
  A.vfq = L.vfq; A.vf_cap = L.vf_cap; A.eeq = L.eeq; A.ee_cap = L.ee_cap;
  A.disc = L.disc; A.disc_cap = L.disc_cap; A.counters = L.counters;
  A.odep = L.odep; A.odep_cap = L.odep_cap;
  A.off_vf = L.off_vf; A.off_ee = L.off_ee; A.nchunks_cap = expand_chunks(L.pair_cap);
  A.chunk = L.chunk < 1 ? 1 : L.chunk;
  const long long items = L.npairs_hint >= 0 ? L.npairs_hint : (long long)L.pair_cap;
  if (items <= 0) return;
  const int g = grid_for((items + kWave - 1) / kWave / A.chunk + 1, EX_WAVES, L.num_cus, L.blocks_per_cu);
  if (L.continuous) {  // pass 1 + scans: exact output offsets, no shared queue counter
    const int gc = grid_for((A.nchunks_cap + EC_U - 1) / EC_U, EX_WAVES, L.num_cus, L.blocks_per_cu);
    if (dim == 3) k_expand_count<3><<<gc, EX_WAVES * kWave, 0, s>>>(A);
    else k_expand_count<2><<<gc, EX_WAVES * kWave, 0, s>>>(A);
    launch_scan_u64(A.off_vf, A.nchunks_cap + 1, L.scan_part, s);
    launch_scan_u64(A.off_ee, A.nchunks_cap + 1, L.scan_part, s);
    k_expand_totals<<<1, 1, 0, s>>>(A.off_vf, A.off_ee, A.nchunks_cap, A.counters);
  }
  const bool od = A.odep != nullptr && L.discrete;
  if (dim == 3) {
    if (od) k_expand<3, true><<<g, EX_WAVES * kWave, 0, s>>>(A);
    else k_expand<3, false><<<g, EX_WAVES * kWave, 0, s>>>(A);
  } else {
    if (od) k_expand<2, true><<<g, EX_WAVES * kWave, 0, s>>>(A);
    else k_expand<2, false><<<g, EX_WAVES * kWave, 0, s>>>(A);
  }
}

long long accd_chunks(unsigned long long cap) { return (long long)((cap + kWave - 1) / kWave); }

void launch_accd(int dim, bool vf, const AccdLaunch& L, hipStream_t s) {
  AccdArgs A;
  A.q = L.q; A.n_dev = L.n_dev; A.cap = L.cap; A.rec = L.rec; A.eta = L.eta;
  A.vert_scene = L.vert_scene; A.scene_eta = L.scene_eta; A.cand_toi = L.cand_toi;
  A.chunk_hits = L.chunk_hits; A.counters = L.counters; A.f32 = L.f32;
  const bool direct = L.hit_sten && accd_direct_hits(L.cap);
  A.hit_sten = direct ? L.hit_sten : nullptr; A.hit_vf = L.hit_vf; A.hit_toi = L.hit_toi; A.hit_cap = L.hit_cap;
  if (!direct && (!A.cand_toi || !A.chunk_hits)) return;  // caller error: no output at all
  const long long items = L.n_hint >= 0 ? L.n_hint : (long long)L.cap;
  if (items <= 0) return;
  const int g = grid_for((items + kWave - 1) / kWave, AC_WAVES, L.num_cus, L.blocks_per_cu);
  A.suspend = direct && !A.chunk_hits && !L.no_suspend;  // suspended hits would miss their chunk's count
  // 3D CTCD on the prefilter's survivors only (the same condition: hits of a
  // listed stencil would miss their chunk's count)
  A.sel = (dim == 3 && direct && !A.chunk_hits && L.ctcd_sel) ? L.ctcd_sel : nullptr;
  if (vf && L.one_sided) {
    // the host libm's values, as the reference computes them (rpoly.h:74-77)
    static const double rot = 94.0 * 0.017453293, cosr = std::cos(rot), sinr = std::sin(rot);
    if (A.sel) {  // counters[C_CTCD_SEL] was cleared by the caller
      k_ctcd_select<<<g, AC_WAVES * kWave, 0, s>>>(A);
      k_ctcd<3, true><<<g, AC_WAVES * kWave, 0, s>>>(A, cosr, sinr);
    } else if (dim == 3) {
      k_ctcd<3, false><<<g, AC_WAVES * kWave, 0, s>>>(A, cosr, sinr);
    } else {
      k_ctcd<2, false><<<g, AC_WAVES * kWave, 0, s>>>(A, cosr, sinr);
    }
    return;
  }
  if (L.f32) {
    if (dim == 3) {
      if (vf) k_accd<3, true, float><<<g, AC_WAVES * kWave, 0, s>>>(A);
      else k_accd<3, false, float><<<g, AC_WAVES * kWave, 0, s>>>(A);
    } else {
      k_accd<2, true, float><<<g, AC_WAVES * kWave, 0, s>>>(A);
    }
    return;
  }
  if (dim == 3) {
    if (vf) k_accd<3, true, double><<<g, AC_WAVES * kWave, 0, s>>>(A);
    else k_accd<3, false, double><<<g, AC_WAVES * kWave, 0, s>>>(A);
  } else {
    k_accd<2, true, double><<<g, AC_WAVES * kWave, 0, s>>>(A);
  }
}

void launch_hits(int dim, const HitsLaunch& L, bool scan, hipStream_t s) {
  const long long nv = accd_chunks(L.vf_cap), ne = dim == 3 ? accd_chunks(L.ee_cap) : 0;
  if (scan) {
    launch_scan_u64(L.off, nv + ne + 1, L.scan_part, s);
    k_hits_total<<<1, 1, 0, s>>>(L.off, nv + ne, L.counters);
  }
  const int gv = grid_for((nv + HC_U - 1) / HC_U, 4, L.num_cus, L.blocks_per_cu);
  k_hits_compact<true><<<gv, 256, 0, s>>>(L.vfq, L.n_vf_dev, L.vf_cap, L.vf_toi, L.off, L.hit_sten, L.hit_vf,
                                         L.hit_toi, L.hit_cap);
  if (dim == 3 && ne > 0) {
    const int ge = grid_for((ne + HC_U - 1) / HC_U, 4, L.num_cus, L.blocks_per_cu);
    k_hits_compact<false><<<ge, 256, 0, s>>>(L.eeq, L.n_ee_dev, L.ee_cap, L.ee_toi, L.off + nv, L.hit_sten,
                                            L.hit_vf, L.hit_toi, L.hit_cap);
  }
}

}  // namespace mclg
