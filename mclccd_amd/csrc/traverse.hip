// Self-traversal and narrow phase for gfx950.
//
// k_traverse: one lane per sorted leaf s; the lane walks the LBVH stacklessly
//   from escape(leaf s), i.e. over exactly the subtrees covering leaves s+1..n-1,
//   so every unordered leaf pair is visited once.  Internal nodes are culled with
//   the FP32 outward-rounded boxes; at a leaf the reference's exact FP64 test
//   (BVHTree.cpp:401-426) decides.  Pairs are appended to a device queue with
//   wave-aggregated atomics.
// k_narrow: one wave takes 64 leaf pairs, runs the discrete test per pair (one
//   lane each), expands the representative-triangle candidates
//   (BVHTree.cpp:305-398) into per-wave LDS lists split by kind (VF / EE), then
//   runs FP64 ACCD on 64 stencils at a time so that each SIMD step executes a
//   single distance kernel.
#include "geom.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace mclg {

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Leaf-pair tests
// ---------------------------------------------------------------------------
template <int D> __device__ __forceinline__ bool cull_f32(const NodeF<D>& q, const NodeF<D>& b) {
  if (((q.meta | b.meta) & 0x80000000u) == 0) return false;
#pragma unroll
  for (int k = 0; k < D; ++k)
    if (!(q.sw_lo[k] <= b.sw_hi[k]) || !(b.sw_lo[k] <= q.sw_hi[k])) return false;
#pragma unroll
  for (int k = 0; k < D; ++k)
    if ((q.t0_lo[k] > b.t0_hi[k] && q.t1_lo[k] > b.t1_hi[k]) || (q.t0_hi[k] < b.t0_lo[k] && q.t1_hi[k] < b.t1_lo[k]))
      return false;
  return true;
}

template <int D> __device__ __forceinline__ bool empty_box(const double* lo, const double* hi) {
#pragma unroll
  for (int k = 0; k < D; ++k)
    if (lo[k] > hi[k]) return true;
  return false;
}

// BVHTree.cpp:401-426 on two leaves (symmetric in l and r)
template <int D>
__device__ __forceinline__ bool leaf_test(const LeafBox<D>& l, bool la, const LeafBox<D>& r, bool ra) {
  if (!la && !ra) return false;
#pragma unroll
  for (int k = 0; k < D; ++k)
    if (!(l.sw_lo[k] <= r.sw_hi[k])) return false;
#pragma unroll
  for (int k = 0; k < D; ++k)
    if (!(r.sw_lo[k] <= l.sw_hi[k])) return false;
  if (empty_box<D>(l.t0_lo, l.t0_hi) || empty_box<D>(r.t0_lo, r.t0_hi)) return true;
  if (empty_box<D>(l.t1_lo, l.t1_hi) || empty_box<D>(r.t1_lo, r.t1_hi)) return true;
#pragma unroll
  for (int k = 0; k < D; ++k)
    if ((l.t0_lo[k] > r.t0_hi[k] && l.t1_lo[k] > r.t1_hi[k]) || (l.t0_hi[k] < r.t0_lo[k] && l.t1_hi[k] < r.t1_lo[k]))
      return false;
  return true;
}

// Pairs found by a wave are staged in LDS and written out in batches: one
// global atomic per flush instead of one per emitting step.
constexpr int TR_WAVES = 4, TR_BUF = 512;

__device__ __forceinline__ void flush_pairs(const int2* buf, int cnt, int2* __restrict__ pairs, unsigned long long cap,
                                            unsigned long long* counter) {
  // executed by the wave's active lanes (inside or after the divergent walk)
  const uint64_t act = __ballot(1);
  const int nact = __popcll(act), r = popc_below(act);
  const int leader = __ffsll((long long)act) - 1;
  unsigned long long base = 0;
  if (r == 0) base = atomicAdd(counter, (unsigned long long)cnt);
  base = __shfl(base, leader);
  for (int i = r; i < cnt; i += nact)
    if (base + i < cap) pairs[base + i] = buf[i];
}

template <int D>
__global__ __launch_bounds__(TR_WAVES * kWave) void k_traverse(const NodeF<D>* __restrict__ nodes,
                                                               const LeafBox<D>* __restrict__ boxes,
                                                               const uint8_t* __restrict__ act,
                                                               const uint32_t* __restrict__ order, int n,
                                                               const int* __restrict__ rank, int2* __restrict__ pairs,
                                                               unsigned long long cap, unsigned long long* counters) {
  __shared__ int2 s_buf[TR_WAVES][TR_BUF];
  __shared__ int s_cnt[TR_WAVES];
  const int w = threadIdx.x / kWave;
  const int s = xcd_block() * blockDim.x + threadIdx.x;
  const bool valid = s < n - 1;
  if (lane_id() == 0) s_cnt[w] = 0;
  wave_lds_sync();
  NodeF<D> q;
  LeafBox<D> qb;
  bool qa = false;
  int a = 0, ra = 0;
  int node = -1;
  if (valid) {
    q = nodes[(n - 1) + s];
    a = (int)order[s];
    qb = boxes[a];
    qa = act[a] != 0;
    ra = rank ? rank[a] : a;
    node = (int)(q.meta & 0x7fffffffu) - 1;
  }
  // a pre-order walk visits each of the 2n-1 nodes at most once: more steps
  // than that means a malformed tree, which is reported instead of looping
  int budget = 2 * n;
  while (node >= 0) {
    if (--budget < 0) {
      atomicAdd(counters + C_ERRORS, 1ull);
      break;
    }
    const NodeF<D> nd = nodes[node];
    bool emit = false;
    int b = 0;
    int next = (int)(nd.meta & 0x7fffffffu) - 1;
    if (cull_f32<D>(q, nd)) {
      if (node >= n - 1) {
        b = nd.child;
        const LeafBox<D> bb = boxes[b];
        emit = leaf_test<D>(qb, qa, bb, act[b] != 0);
      } else {
        next = nd.child;
      }
    }
    const uint64_t m = __ballot(emit);
    if (m) {
      const int cnt = s_cnt[w];
      if (emit) {
        const int rb = rank ? rank[b] : b;
        s_buf[w][cnt + popc_below(m)] = ra < rb ? make_int2(a, b) : make_int2(b, a);
      }
      const int nc = cnt + __popcll(m);
      wave_lds_sync();
      if (nc > TR_BUF - kWave) {
        flush_pairs(s_buf[w], nc, pairs, cap, counters + C_LEAF_PAIRS);
        wave_lds_sync();
        if (popc_below(__ballot(1)) == 0) s_cnt[w] = 0;
      } else if (popc_below(__ballot(1)) == 0) {
        s_cnt[w] = nc;
      }
      wave_lds_sync();
    }
    node = next;
  }
  wave_lds_sync();
  const int cnt = s_cnt[w];
  if (cnt > 0) flush_pairs(s_buf[w], cnt, pairs, cap, counters + C_LEAF_PAIRS);
}

void launch_traverse(int dim, const void* nodes, const void* boxes, const uint8_t* act, const uint32_t* order, int n,
                     const int* rank, int2* pairs, unsigned long long cap, unsigned long long* counters, hipStream_t s) {
  if (n < 2) return;
  const int g = xcd_grid(n - 1, TR_WAVES * kWave);
  if (dim == 3)
    k_traverse<3><<<g, TR_WAVES * kWave, 0, s>>>((const NodeF<3>*)nodes, (const LeafBox<3>*)boxes, act, order, n, rank, pairs, cap,
                                    counters);
  else
    k_traverse<2><<<g, TR_WAVES * kWave, 0, s>>>((const NodeF<2>*)nodes, (const LeafBox<2>*)boxes, act, order, n, rank, pairs, cap,
                                    counters);
}

// ---------------------------------------------------------------------------
// Narrow phase
// ---------------------------------------------------------------------------
template <int D> struct VLoad;
template <> struct VLoad<3> {
  using V = d3;
  __device__ __forceinline__ static void load(const double* rec, int v, d3& a, d3& b) {
    const double2* r = reinterpret_cast<const double2*>(rec + (size_t)v * 8);
    const double2 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
    a = {r0.x, r0.y, r1.x};
    b = {r2.x, r2.y, r3.x};
  }
  __device__ __forceinline__ static d3 load_t1(const double* rec, int v) {
    const double2* r = reinterpret_cast<const double2*>(rec + (size_t)v * 8);
    const double2 r2 = r[2], r3 = r[3];
    return {r2.x, r2.y, r3.x};
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

// exclusive prefix over the wave of a small count (< 2^bits) and the total
template <int BITS> __device__ __forceinline__ int wave_prefix(int c, int& total) {
  int pre = 0, tot = 0;
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    const uint64_t m = __ballot((c >> b) & 1);
    pre += popc_below(m) << b;
    tot += __popcll(m) << b;
  }
  total = tot;
  return pre;
}

struct NarrowArgs {
  const int2* pairs;
  long long npairs;
  const int* P;
  const uint8_t* flags;
  const double* rec;
  double vf_eta, ee_eta;
  int discrete, continuous, want_candidates;
  int4* hit_sten;
  uint8_t* hit_vf;
  double* hit_toi;
  unsigned long long hit_cap;
  int2* disc;
  unsigned long long disc_cap;
  int4* cand_sten;
  uint8_t* cand_vf;
  double* cand_toi;
  unsigned long long cand_cap;
  unsigned long long* counters;
};

template <int D, bool VF>
__device__ __forceinline__ void run_stencils(const NarrowArgs& A, const uint16_t* list, int count, const int* fl,
                                             int lane, long long& evals) {
  using V = typename VLoad<D>::V;
  for (int base = 0; base < count; base += kWave) {
    const int e = base + lane;
    const bool live = e < count;
    int sv[4] = {-1, -1, -1, -1};
    if (live) {
      const uint32_t ent = list[e];
      const int src = (int)(ent >> 4), code = (int)(ent & 15u);
      const int* f = fl + src * 6;  // f0 = f[0..2], f1 = f[3..5]
      if (VF) {
        const int from = code < 3 ? 0 : 3, to = code < 3 ? 3 : 0, i = code < 3 ? code : code - 3;
        sv[0] = f[from + i];
#pragma unroll
        for (int j = 0; j < D; ++j) sv[1 + j] = f[to + j];
      } else {
        const int c = code - 6, i = c / 3, j = c - 3 * (c / 3);
        sv[0] = f[i]; sv[1] = f[(i + 1) % 3]; sv[2] = f[3 + j]; sv[3] = f[3 + (j + 1) % 3];
      }
    }
    long long cslot = -1;
    if (A.want_candidates) {
      cslot = wave_append(live, A.counters + C_CANDIDATES);
      if (live && cslot < (long long)A.cand_cap) {
        A.cand_sten[cslot] = make_int4(sv[0], sv[1], sv[2], sv[3]);
        A.cand_vf[cslot] = VF ? 1 : 0;
      }
    }
    int hit = 0, ev = 0;
    double toi = -1.0;
    if (live) {
      V v0[D + 1], v1[D + 1];
#pragma unroll
      for (int i = 0; i < D + 1; ++i) VLoad<D>::load(A.rec, sv[i], v0[i], v1[i]);
      hit = accd<D, VF>(v0, v1, VF ? A.vf_eta : A.ee_eta, toi, ev);
    }
    evals += ev;
    if (live && cslot >= 0 && cslot < (long long)A.cand_cap) A.cand_toi[cslot] = hit ? toi : -1.0;
    const long long slot = wave_append(hit != 0, A.counters + C_HITS);
    if (hit && slot < (long long)A.hit_cap) {
      A.hit_sten[slot] = make_int4(sv[0], sv[1], sv[2], sv[3]);
      A.hit_vf[slot] = VF ? 1 : 0;
      A.hit_toi[slot] = toi;
    }
  }
}

constexpr int NW = 4;  // waves per block

template <int D> __global__ __launch_bounds__(NW * kWave) void k_narrow(NarrowArgs A) {
  __shared__ uint16_t s_vf[NW][kWave * 2 * D];
  __shared__ uint16_t s_ee[NW][D == 3 ? kWave * 9 : 1];
  __shared__ int s_f[NW][kWave * 6];
  const int w = threadIdx.x / kWave, lane = lane_id();
  const long long nchunks = (A.npairs + kWave - 1) / kWave;
  const long long gw = (long long)xcd_block() * NW + w;
  const long long stride = (long long)gridDim.x * NW;
  long long evals = 0;
  unsigned long long ncand = 0;
  for (long long ch = gw; ch < nchunks; ch += stride) {
    const long long i = ch * kWave + lane;
    const bool valid = i < A.npairs;
    int f0[3] = {-1, -1, -1}, f1[3] = {-1, -1, -1};
    int2 pr = make_int2(0, 0);
    uint32_t fa = 0, fb = 0;
    if (valid) {
      pr = A.pairs[i];
#pragma unroll
      for (int j = 0; j < D; ++j) { f0[j] = A.P[(size_t)pr.x * D + j]; f1[j] = A.P[(size_t)pr.y * D + j]; }
      fa = A.flags[pr.x];
      fb = A.flags[pr.y];
    }
    // ---- discrete test at V1 (BVHTree.cpp:464-500) ----
    if (A.discrete) {
      bool shared = false;
#pragma unroll
      for (int x = 0; x < D; ++x)
#pragma unroll
        for (int y = 0; y < D; ++y) shared |= f0[x] == f1[y];
      int dh = 0;
      if (valid && !shared) {
        if constexpr (D == 3) {
          const d3 p1 = VLoad<3>::load_t1(A.rec, f0[0]), q1 = VLoad<3>::load_t1(A.rec, f0[1]),
                   r1 = VLoad<3>::load_t1(A.rec, f0[2]);
          const d3 p2 = VLoad<3>::load_t1(A.rec, f1[0]), q2 = VLoad<3>::load_t1(A.rec, f1[1]),
                   r2 = VLoad<3>::load_t1(A.rec, f1[2]);
          dh = tri_tri3(p1, q1, r1, p2, q2, r2);
        } else {
          const d2 p0 = VLoad<2>::load_t1(A.rec, f0[0]), p1 = VLoad<2>::load_t1(A.rec, f0[1]);
          const d2 q0 = VLoad<2>::load_t1(A.rec, f1[0]), q1 = VLoad<2>::load_t1(A.rec, f1[1]);
          dh = edge_edge2(p0, p1, q0, q1);
        }
      }
      const long long slot = wave_append(dh != 0, A.counters + C_DISCRETE);
      if (dh && slot < (long long)A.disc_cap) A.disc[slot] = pr;
    }
    if (!A.continuous) continue;
    // ---- candidate codes (BVHTree.cpp:305-398) ----
    uint32_t vfm = 0, eem = 0;
    if (valid) {
#pragma unroll
      for (int x = 0; x < D; ++x) {
        bool sh0 = false, sh1 = false;
#pragma unroll
        for (int y = 0; y < D; ++y) { sh0 |= f0[x] == f1[y]; sh1 |= f1[x] == f0[y]; }
        if (((fa >> x) & 1u) && !sh0) vfm |= 1u << x;
        if (((fb >> x) & 1u) && !sh1) vfm |= 1u << (3 + x);
      }
      if constexpr (D == 3) {
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          if (!((fa >> (3 + x)) & 1u)) continue;
#pragma unroll
          for (int y = 0; y < 3; ++y) {
            if (!((fb >> (3 + y)) & 1u)) continue;
            const int a0 = f0[x], a1 = f0[(x + 1) % 3], b0 = f1[y], b1 = f1[(y + 1) % 3];
            if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) continue;
            eem |= 1u << (3 * x + y);
          }
        }
      }
    }
    int nvf, nee = 0;
    int off = wave_prefix<3>(__popc(vfm), nvf);
    for (uint32_t m = vfm; m; m &= m - 1) s_vf[w][off++] = (uint16_t)((lane << 4) | (__ffs(m) - 1));
    if constexpr (D == 3) {
      off = wave_prefix<4>(__popc(eem), nee);
      for (uint32_t m = eem; m; m &= m - 1) s_ee[w][off++] = (uint16_t)((lane << 4) | (6 + __ffs(m) - 1));
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) { s_f[w][lane * 6 + j] = f0[j]; s_f[w][lane * 6 + 3 + j] = f1[j]; }
    wave_lds_sync();
    ncand += (unsigned long long)(nvf + nee);
    run_stencils<D, true>(A, s_vf[w], nvf, s_f[w], lane, evals);
    if constexpr (D == 3) run_stencils<D, false>(A, s_ee[w], nee, s_f[w], lane, evals);
    wave_lds_sync();
  }
  // per-wave totals
  for (int o = 32; o > 0; o >>= 1) evals += __shfl_xor(evals, o);
  if (lane == 0) {
    atomicAdd(A.counters + C_DIST_EVALS, (unsigned long long)evals);
    if (!A.want_candidates) atomicAdd(A.counters + C_CANDIDATES, ncand);
  }
}

void launch_narrow(int dim, const NarrowLaunch& L, hipStream_t s) {
  if (L.npairs <= 0) return;
  NarrowArgs A;
  A.pairs = L.pairs; A.npairs = L.npairs; A.P = L.P; A.flags = L.flags; A.rec = L.rec;
  A.vf_eta = L.vf_eta; A.ee_eta = L.ee_eta;
  A.discrete = L.discrete; A.continuous = L.continuous; A.want_candidates = L.want_candidates;
  A.hit_sten = L.hit_sten; A.hit_vf = L.hit_vf; A.hit_toi = L.hit_toi; A.hit_cap = L.hit_cap;
  A.disc = L.disc; A.disc_cap = L.disc_cap;
  A.cand_sten = L.cand_sten; A.cand_vf = L.cand_vf; A.cand_toi = L.cand_toi; A.cand_cap = L.cand_cap;
  A.counters = L.counters;
  const long long chunks = (L.npairs + kWave - 1) / kWave;
  long long blocks = (chunks + NW - 1) / NW;
  const long long maxb = (long long)L.num_cus * L.blocks_per_cu;
  if (blocks > maxb) blocks = maxb;
  blocks = (blocks + 7) / 8 * 8;
  if (dim == 3) k_narrow<3><<<(int)blocks, NW * kWave, 0, s>>>(A);
  else k_narrow<2><<<(int)blocks, NW * kWave, 0, s>>>(A);
}

}  // namespace mclg
