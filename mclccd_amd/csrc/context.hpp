// Host state of one collision context (one mcl::BVHTree on one GPU).
#pragma once
#include <cstdarg>
#include <string>
#include <vector>

#include "../../include/mclg.h"
#include "common.hpp"

namespace mclg {

constexpr int kNumEvents = 19;
constexpr size_t kAdjQueue = size_t(1) << 22;  // adjoint multi-step queue entries

struct BatchState;

struct Ctx {
  int dim = 3;
  int device = 0;
  int num_cus = 256;
  // tunables (the batch autotuner searches them)
  // blocks per CU in a queue kernel's grid: more than fit at once, so the
  // hardware hands out the rest as blocks finish, which evens out chunks of
  // uneven cost (ACCD per 1M-triangle frame: 32 measured ~2 ms faster than 8,
  // 64 a further ~0.35 ms; 16 slower than 32)
  int blocks_per_cu = 64;
  int chunk = 1;           // consecutive 64-item chunks per wave step
  int tr_waves = 4;        // waves per traversal block
  mclg_options opt{};
  hipStream_t own_stream = nullptr, stream = nullptr;
  hipEvent_t ev[kNumEvents] = {};
  bool timing = false;
  double times[MCLG_NUM_TIMINGS] = {};
  std::string err;

  // tree state
  int nv = 0, np = 0, pdim = 0;
  bool have_tree = false;
  bool rank_set = false;
  // orientation: MCLG_ORIENT_INDEX (primitive index, or a rank from
  // mclg_set_rank) or MCLG_ORIENT_KDBVH (the reference's KdBVH order, rebuilt
  // on the host at every update; also used while opt.threaded == 0)
  int orient = MCLG_ORIENT_INDEX;
  // MCLG_PRECISION_F32: BVHTree<float,DIM> arithmetic (float coordinates,
  // boxes and ACCD); hits_f32: the last traverse ran in it
  int precision = MCLG_PRECISION_F64;
  bool hits_f32 = false;
  bool kd_valid = false;           // kd_rank / kd_gapnode describe the current leaves
  std::vector<int> kd_rank;        // in-order position of each primitive
  std::vector<int> kd_gapnode;     // frontlist index of the node splitting positions g, g+1
  bool keep_cand = false, keep_pairs = false;
  bool hits_one_sided = false;  // the last traverse used CTCD for VF (no TOI adjoint for it)
  // generations: tree_gen counts updates and refits; hits_gen is the tree_gen
  // the last complete traverse ran on (kNoGen: none since the last update),
  // so the adjoint can refuse hits whose vertex records were replaced
  static constexpr unsigned long long kNoGen = ~0ull;
  unsigned long long tree_gen = 0, hits_gen = kNoGen;
  mclg_counts counts{0, 0, 0, 0, 0};

  // device buffers (grown on demand; cap in bytes)
  double* d_rec = nullptr; size_t cap_rec = 0;
  uint8_t* d_flags = nullptr; size_t cap_flags = 0;
  int* d_owner = nullptr; size_t cap_owner = 0;
  unsigned long long* d_ekeys = nullptr; size_t cap_ekeys = 0;
  int* d_evals = nullptr; size_t cap_evals = 0;
  void* d_boxes = nullptr; size_t cap_boxes = 0;
  void* d_sboxes = nullptr; size_t cap_sboxes = 0;  // FP64 leaf boxes in sorted order
  void* d_face = nullptr; size_t cap_face = 0;      // face arrays: head, rank, prim per sorted leaf (per traverse)
  double* d_v1 = nullptr; size_t cap_v1 = 0;        // V1 per vertex (per traverse)
  bool pairs_are_prims = false;                     // d_pairs converted to primitive pairs
  unsigned long long* d_ex_off = nullptr; size_t cap_ex_off = 0;    // candidate offsets per 64-pair chunk
  unsigned long long* d_ex_part = nullptr; size_t cap_ex_part = 0;  // their scan's chunk sums
  unsigned long long* d_hit_off = nullptr; size_t cap_hit_off = 0;  // hits per 64-stencil chunk, then offsets
  unsigned long long* d_hit_part = nullptr; size_t cap_hit_part = 0;
  uint8_t* d_act = nullptr; size_t cap_act = 0;
  unsigned long long* d_cbounds = nullptr; size_t cap_cbounds = 0;
  double* d_partial = nullptr; size_t cap_partial = 0;
  uint32_t* d_keys = nullptr; size_t cap_keys = 0;
  uint32_t* d_order = nullptr; size_t cap_order = 0;
  uint32_t* d_ktmp = nullptr; size_t cap_ktmp = 0;
  uint32_t* d_vtmp = nullptr; size_t cap_vtmp = 0;
  uint32_t* d_hist = nullptr; size_t cap_hist = 0;
  void* d_nodes = nullptr; size_t cap_nodes = 0;
  void* d_cp = nullptr; size_t cap_cp = 0;          // child pairs of the internal nodes
  int* d_root = nullptr; size_t cap_root = 0;       // root node index
  int* d_lchild = nullptr; size_t cap_lchild = 0;
  int* d_rchild = nullptr; size_t cap_rchild = 0;
  int* d_rngl = nullptr; size_t cap_rngl = 0;
  int* d_rngr = nullptr; size_t cap_rngr = 0;
  int* d_visits = nullptr; size_t cap_visits = 0;
  int4* d_frontier = nullptr; size_t cap_frontier = 0;
  unsigned long long* d_nfrontier = nullptr; size_t cap_nfrontier = 0;
  unsigned long long* d_counters = nullptr; size_t cap_counters = 0;
  int2* d_pairs = nullptr; size_t cap_pairs_bytes = 0;
  unsigned long long pair_cap = 0;
  int4* d_hit_sten = nullptr; size_t cap_hit_sten = 0;
  uint8_t* d_hit_vf = nullptr; size_t cap_hit_vf = 0;
  double* d_hit_toi = nullptr; size_t cap_hit_toi = 0;
  unsigned long long hit_cap = 0;
  int2* d_disc = nullptr; size_t cap_disc = 0;
  unsigned long long disc_cap = 0;
  int2* d_odep = nullptr; size_t cap_odep = 0;  // order-dependent discrete pairs (MCLG_CHECK_ORDER)
  unsigned long long odep_cap = 0;
  bool check_order = false;
  int4* d_vfq = nullptr; size_t cap_vfq = 0;
  int4* d_eeq = nullptr; size_t cap_eeq = 0;
  unsigned long long vf_cap = 0, ee_cap = 0;
  double* d_vf_toi = nullptr; size_t cap_vf_toi = 0;
  double* d_ee_toi = nullptr; size_t cap_ee_toi = 0;
  uint32_t* d_ctcd_sel = nullptr; size_t cap_ctcd_sel = 0;  // 3D CTCD: stencils that pass the prefilter
  long long n_vf = 0, n_ee = 0;
  int* d_rank = nullptr; size_t cap_rank = 0;
  int* d_minmax = nullptr; size_t cap_minmax = 0;
  double* d_adj_w = nullptr; size_t cap_adj_w = 0;
  double* d_adj_g = nullptr; size_t cap_adj_g = 0;
  long long* d_adj_q = nullptr; size_t cap_adj_q_bytes = 0; size_t cap_adj_q = 0;  // multi-step hit queue
  unsigned long long* d_adj_n = nullptr; size_t cap_adj_n = 0;
  long long adj_multi = 0;  // multi-step hits in the last adjoint
  uint32_t *d_adj_keys = nullptr, *d_adj_vals = nullptr, *d_adj_ktmp = nullptr, *d_adj_vtmp = nullptr,
           *d_adj_hist = nullptr;  // adjoint visiting order (radix sort of the hits)
  size_t cap_adj_keys = 0, cap_adj_vals = 0, cap_adj_ktmp = 0, cap_adj_vtmp = 0, cap_adj_hist = 0;
  double* d_stage_v0 = nullptr; size_t cap_stage_v0 = 0;
  double* d_stage_v1 = nullptr; size_t cap_stage_v1 = 0;
  int* d_stage_p = nullptr; size_t cap_stage_p = 0;
  int* d_stage_act = nullptr; size_t cap_stage_act = 0;

  BatchState* batch = nullptr;  // set for batched-scene contexts

  // pinned host mirrors
  unsigned long long* h_counters = nullptr;
  int* h_minmax = nullptr;

  ~Ctx();
  int fail(int code, const char* fmt, ...);
  int hip(hipError_t e, const char* what);
  int grow(void** p, size_t* cap, size_t need, const char* what);
  void mark(int i);
  int check_indices(const int* P, long long n, int nv, const char* what);
  int read_counters();
  int stage(const double* V0, const double* V1, int nv, const int* P, long long nidx, const int* active);
  // refit: keep the last build's leaf order and Morton keys (so its hierarchy)
  // and recompute only the boxes; needs the same nv, np and pdim.
  int update(const double* V0, const double* V1, int nv, const int* P, int np, int pdim, const int* active,
             bool refit = false);
  int traverse(const double* V0, const double* V1, const int* P, unsigned flags);
  int kdbvh_from_leaves();  // the reference's KdBVH order of the current leaves -> rank
  int events(bool cand, int64_t* out, int64_t* group, int64_t cap, int64_t* nout);
  FaceArrays faces() { return face_arrays(d_face, (size_t)np, d_v1); }
};

}  // namespace mclg
