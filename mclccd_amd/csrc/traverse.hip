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
constexpr int TR_MAX_WAVES = 8, TR_BUF = 512;

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
__global__ __launch_bounds__(TR_MAX_WAVES * kWave) void k_traverse(const NodeF<D>* __restrict__ nodes,
                                                               const LeafBox<D>* __restrict__ boxes,
                                                               const uint8_t* __restrict__ act,
                                                               const uint32_t* __restrict__ order, int n,
                                                               const int* __restrict__ rank, int2* __restrict__ pairs,
                                                               unsigned long long cap, unsigned long long* counters,
                                                               const uint32_t* __restrict__ scene_of_sorted,
                                                               const int* __restrict__ scene_root) {
  __shared__ int2 s_buf[TR_MAX_WAVES][TR_BUF];
  __shared__ int s_cnt[TR_MAX_WAVES];
  const int w = threadIdx.x / kWave;
  const int s = xcd_block() * blockDim.x + threadIdx.x;
  const bool valid = s < n - 1;
  if (lane_id() == 0) s_cnt[w] = 0;
  wave_lds_sync();
  NodeF<D> q;
  LeafBox<D> qb;
  bool qa = false;
  int a = 0, ra = 0;
  int node = -1, end = -1;
  if (valid) {
    q = nodes[(n - 1) + s];
    a = (int)order[s];
    qb = boxes[a];
    qa = act[a] != 0;
    ra = rank ? rank[a] : a;
    node = (int)(q.meta & 0x7fffffffu) - 1;
    // several scenes: the walk ends where the pre-order leaves this leaf's scene subtree
    if (scene_root) end = (int)(nodes[scene_root[scene_of_sorted[s]]].meta & 0x7fffffffu) - 1;
    if (node == end) node = -1;
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
    if (next == end) next = -1;
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
                     const int* rank, int2* pairs, unsigned long long cap, unsigned long long* counters,
                     const uint32_t* scene_of_sorted, const int* scene_root, int waves, hipStream_t s) {
  if (n < 2) return;
  waves = waves < 1 ? 4 : (waves > TR_MAX_WAVES ? TR_MAX_WAVES : waves);
  const int g = xcd_grid(n - 1, waves * kWave);
  if (dim == 3)
    k_traverse<3><<<g, waves * kWave, 0, s>>>((const NodeF<3>*)nodes, (const LeafBox<3>*)boxes, act, order, n, rank,
                                                 pairs, cap, counters, scene_of_sorted, scene_root);
  else
    k_traverse<2><<<g, waves * kWave, 0, s>>>((const NodeF<2>*)nodes, (const LeafBox<2>*)boxes, act, order, n, rank,
                                                 pairs, cap, counters, scene_of_sorted, scene_root);
}

// ---------------------------------------------------------------------------
// Narrow phase
// ---------------------------------------------------------------------------
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

// ---------------------------------------------------------------------------
// k_expand: per leaf pair (BVHTree.cpp:252-283) the discrete test at V1 and the
// representative-triangle candidates, appended to a VF and an EE stencil queue.
// ---------------------------------------------------------------------------
struct ExpandArgs {
  const int2* pairs;
  const unsigned long long* npairs_dev;  // pairs found by the traversal (device counter)
  unsigned long long pair_cap;
  const int* P;
  const uint8_t* flags;
  const double* rec;
  int discrete, continuous;
  int4* vfq;
  unsigned long long vf_cap;
  int4* eeq;
  unsigned long long ee_cap;
  int2* disc;
  unsigned long long disc_cap;
  unsigned long long* counters;
  int chunk;  // consecutive 64-item chunks a wave takes per grid-stride step
};

constexpr int EX_WAVES = 4, EX_STAGE = 256, EX_DSTAGE = 128;

template <int D> __global__ __launch_bounds__(EX_WAVES * kWave) void k_expand(ExpandArgs A) {
  __shared__ int4 s_vf[EX_WAVES][EX_STAGE];
  __shared__ int4 s_ee[EX_WAVES][D == 3 ? EX_STAGE : 1];
  __shared__ int2 s_d[EX_WAVES][EX_DSTAGE];
  const int w = threadIdx.x / kWave, lane = lane_id();
  int nvf = 0, nee = 0, nd = 0;  // staged records (wave-uniform)
  const long long npairs = (long long)(*A.npairs_dev < A.pair_cap ? *A.npairs_dev : A.pair_cap);
  const long long nchunks = (npairs + kWave - 1) / kWave;
  const long long stride = (long long)gridDim.x * EX_WAVES * A.chunk;
  for (long long c0 = ((long long)xcd_block() * EX_WAVES + w) * A.chunk; c0 < nchunks; c0 += stride)
  for (long long ch = c0; ch < c0 + A.chunk && ch < nchunks; ++ch) {
    const long long i = ch * kWave + lane;
    const bool valid = i < npairs;
    int f0[3] = {-1, -1, -1}, f1[3] = {-2, -2, -2};
    int2 pr = make_int2(0, 0);
    uint32_t fa = 0, fb = 0;
    if (valid) {
      pr = A.pairs[i];
#pragma unroll
      for (int j = 0; j < D; ++j) { f0[j] = A.P[(size_t)pr.x * D + j]; f1[j] = A.P[(size_t)pr.y * D + j]; }
      fa = A.flags[pr.x];
      fb = A.flags[pr.y];
    }
    if (A.discrete) {  // BVHTree.cpp:464-500
      bool shared = false;
#pragma unroll
      for (int x = 0; x < D; ++x)
#pragma unroll
        for (int y = 0; y < D; ++y) shared |= f0[x] == f1[y];
      int dh = 0;
      if (valid && !shared) {
        if constexpr (D == 3) {
          dh = tri_tri3(VLoad<3>::load_t1(A.rec, f0[0]), VLoad<3>::load_t1(A.rec, f0[1]), VLoad<3>::load_t1(A.rec, f0[2]),
                        VLoad<3>::load_t1(A.rec, f1[0]), VLoad<3>::load_t1(A.rec, f1[1]), VLoad<3>::load_t1(A.rec, f1[2]));
        } else {
          dh = edge_edge2(VLoad<2>::load_t1(A.rec, f0[0]), VLoad<2>::load_t1(A.rec, f0[1]), VLoad<2>::load_t1(A.rec, f1[0]),
                          VLoad<2>::load_t1(A.rec, f1[1]));
        }
      }
      stage_push<int2, EX_DSTAGE>(s_d[w], nd, dh != 0, pr, A.disc, A.disc_cap, A.counters + C_DISCRETE);
    }
    if (!A.continuous) continue;
    // BVHTree.cpp:327-397: VF f0->f1 (codes 0..2), f1->f0 (3..5), EE (6 + 3i + j)
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
#pragma unroll
          for (int y = 0; y < 3; ++y) {
            const int a0 = f0[x], a1 = f0[(x + 1) % 3], b0 = f1[y], b1 = f1[(y + 1) % 3];
            const bool ok = ((fa >> (3 + x)) & 1u) && ((fb >> (3 + y)) & 1u) && a0 != b0 && a0 != b1 && a1 != b0 && a1 != b1;
            if (ok) eem |= 1u << (3 * x + y);
          }
        }
      }
    }
    // push one code per lane per round; rounds = the wave's largest count
    for (;;) {
      const bool has = vfm != 0;
      if (__ballot(has) == 0) break;
      int4 st = make_int4(-1, -1, -1, -1);
      if (has) {
        const int code = __ffs(vfm) - 1;
        vfm &= vfm - 1;
        const bool fwd = code < 3;  // vertex of f0 against face f1, else the reverse
        const int k = fwd ? code : code - 3;
        const int a0 = fwd ? f0[0] : f1[0], a1 = fwd ? f0[1] : f1[1], a2 = fwd ? f0[2] : f1[2];
        st.x = k == 0 ? a0 : (k == 1 ? a1 : a2);
        st.y = fwd ? f1[0] : f0[0];
        st.z = fwd ? f1[1] : f0[1];
        st.w = D == 3 ? (fwd ? f1[2] : f0[2]) : -1;
      }
      stage_push<int4, EX_STAGE>(s_vf[w], nvf, has, st, A.vfq, A.vf_cap, A.counters + C_CAND_VF);
    }
    if constexpr (D == 3) {
      for (;;) {
        const bool has = eem != 0;
        if (__ballot(has) == 0) break;
        int4 st = make_int4(-1, -1, -1, -1);
        if (has) {
          const int c = __ffs(eem) - 1;
          eem &= eem - 1;
          const int x = c / 3, y = c - 3 * (c / 3);
          st.x = x == 0 ? f0[0] : (x == 1 ? f0[1] : f0[2]);
          st.y = x == 0 ? f0[1] : (x == 1 ? f0[2] : f0[0]);
          st.z = y == 0 ? f1[0] : (y == 1 ? f1[1] : f1[2]);
          st.w = y == 0 ? f1[1] : (y == 1 ? f1[2] : f1[0]);
        }
        stage_push<int4, EX_STAGE>(s_ee[w], nee, has, st, A.eeq, A.ee_cap, A.counters + C_CAND_EE);
      }
    }
  }
  if (nd) stage_flush(s_d[w], nd, A.disc, A.disc_cap, A.counters + C_DISCRETE);
  if (nvf) stage_flush(s_vf[w], nvf, A.vfq, A.vf_cap, A.counters + C_CAND_VF);
  if constexpr (D == 3)
    if (nee) stage_flush(s_ee[w], nee, A.eeq, A.ee_cap, A.counters + C_CAND_EE);
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
  double* cand_toi;  // per-stencil toi (-1 = no hit) or null
  int4* hit_sten;
  uint8_t* hit_vf;
  double* hit_toi;
  unsigned long long hit_cap;
  unsigned long long* counters;
  int chunk;
};

struct HitRec {
  int4 s;
  double t;
};

constexpr int AC_WAVES = 4, AC_STAGE = 256;

template <int D, bool VF> __global__ __launch_bounds__(AC_WAVES * kWave) void k_accd(AccdArgs A) {
  using V = typename VLoad<D>::V;
  __shared__ HitRec s_h[AC_WAVES][AC_STAGE];
  const int w = threadIdx.x / kWave, lane = lane_id();
  int nh = 0;
  long long evals = 0;
  const long long n = (long long)(*A.n_dev < A.cap ? *A.n_dev : A.cap);
  const long long nchunks = (n + kWave - 1) / kWave;
  const long long stride = (long long)gridDim.x * AC_WAVES * A.chunk;
  for (long long c0 = ((long long)xcd_block() * AC_WAVES + w) * A.chunk; c0 < nchunks; c0 += stride)
  for (long long ch = c0; ch < c0 + A.chunk && ch < nchunks; ++ch) {
    const long long i = ch * kWave + lane;
    const bool live = i < n;
    int4 sv = make_int4(0, 0, 0, 0);
    int hit = 0, ev = 0;
    double toi = -1.0;
    if (live) {
      sv = A.q[i];
      const int id[4] = {sv.x, sv.y, sv.z, sv.w};
      V v0[D + 1], v1[D + 1];
#pragma unroll
      for (int k = 0; k < D + 1; ++k) VLoad<D>::load(A.rec, id[k], v0[k], v1[k]);
      const double eta = A.vert_scene ? A.scene_eta[A.vert_scene[sv.x]] : A.eta;
      hit = accd<D, VF>(v0, v1, eta, toi, ev);
      if (A.cand_toi) A.cand_toi[i] = hit ? toi : -1.0;
    }
    evals += ev;
    // hits: staged, then written to the three SoA arrays
    const uint64_t m = __ballot(hit != 0);
    if (m) {
      if (hit) s_h[w][nh + popc_below(m)] = HitRec{sv, toi};
      nh += __popcll(m);
      if (nh > AC_STAGE - kWave) {
        wave_lds_sync();
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(A.counters + C_HITS, (unsigned long long)nh);
        base = __shfl(base, 0);
        for (int k = lane; k < nh; k += kWave)
          if (base + k < A.hit_cap) {
            A.hit_sten[base + k] = s_h[w][k].s;
            A.hit_toi[base + k] = s_h[w][k].t;
            A.hit_vf[base + k] = VF ? 1 : 0;
          }
        nh = 0;
        wave_lds_sync();
      }
    }
  }
  if (nh) {
    wave_lds_sync();
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(A.counters + C_HITS, (unsigned long long)nh);
    base = __shfl(base, 0);
    for (int k = lane; k < nh; k += kWave)
      if (base + k < A.hit_cap) {
        A.hit_sten[base + k] = s_h[w][k].s;
        A.hit_toi[base + k] = s_h[w][k].t;
        A.hit_vf[base + k] = VF ? 1 : 0;
      }
  }
  for (int o = 32; o > 0; o >>= 1) evals += __shfl_xor(evals, o);
  if (lane == 0 && evals) atomicAdd(A.counters + C_DIST_EVALS, (unsigned long long)evals);
}

static int grid_for(long long chunks, int waves, int num_cus, int per_cu) {
  long long b = (chunks + waves - 1) / waves;
  const long long maxb = (long long)num_cus * per_cu;
  if (b > maxb) b = maxb;
  if (b < 1) b = 1;
  return (int)((b + 7) / 8 * 8);
}

void launch_expand(int dim, const ExpandLaunch& L, hipStream_t s) {
  ExpandArgs A;
  A.pairs = L.pairs; A.npairs_dev = L.npairs_dev; A.pair_cap = L.pair_cap; A.P = L.P; A.flags = L.flags; A.rec = L.rec;
  A.discrete = L.discrete; A.continuous = L.continuous;
  A.vfq = L.vfq; A.vf_cap = L.vf_cap; A.eeq = L.eeq; A.ee_cap = L.ee_cap;
  A.disc = L.disc; A.disc_cap = L.disc_cap; A.counters = L.counters;
  A.chunk = L.chunk < 1 ? 1 : L.chunk;
  const long long items = L.npairs_hint >= 0 ? L.npairs_hint : (long long)L.pair_cap;
  if (items <= 0) return;
  const int g = grid_for((items + kWave - 1) / kWave / A.chunk + 1, EX_WAVES, L.num_cus, L.blocks_per_cu);
  if (dim == 3) k_expand<3><<<g, EX_WAVES * kWave, 0, s>>>(A);
  else k_expand<2><<<g, EX_WAVES * kWave, 0, s>>>(A);
}

void launch_accd(int dim, bool vf, const AccdLaunch& L, hipStream_t s) {
  AccdArgs A;
  A.q = L.q; A.n_dev = L.n_dev; A.cap = L.cap; A.rec = L.rec; A.eta = L.eta;
  A.vert_scene = L.vert_scene; A.scene_eta = L.scene_eta; A.cand_toi = L.cand_toi;
  A.hit_sten = L.hit_sten; A.hit_vf = L.hit_vf; A.hit_toi = L.hit_toi; A.hit_cap = L.hit_cap;
  A.counters = L.counters;
  A.chunk = L.chunk < 1 ? 1 : L.chunk;
  const long long items = L.n_hint >= 0 ? L.n_hint : (long long)L.cap;
  if (items <= 0) return;
  const int g = grid_for((items + kWave - 1) / kWave / A.chunk + 1, AC_WAVES, L.num_cus, L.blocks_per_cu);
  if (dim == 3) {
    if (vf) k_accd<3, true><<<g, AC_WAVES * kWave, 0, s>>>(A);
    else k_accd<3, false><<<g, AC_WAVES * kWave, 0, s>>>(A);
  } else {
    k_accd<2, true><<<g, AC_WAVES * kWave, 0, s>>>(A);
  }
}

}  // namespace mclg
