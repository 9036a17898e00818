// Batched scenes: many independent BVHTree problems solved in one pass.
//
// All scenes' vertices and primitives are concatenated; the LBVH is keyed by
// (scene, Morton code), so every scene's leaves are one contiguous sorted range
// and a query leaf only pairs with leaves of its own range (no cross-scene
// pairs).  One frame is
// a fixed sequence of kernels with no host synchronisation inside — queue
// lengths are read from device counters — so it is captured once as a hipGraph
// and replayed.  Per-scene gaps (box/vf/ee eta) are honoured; the boolean
// options must agree across scenes.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mclg.h"
#include "context.hpp"
#include "kernels.hpp"

using namespace mclg;

#define CHECK(x)                     \
  do {                               \
    int _rc = (x);                   \
    if (_rc != MCLG_OK) return _rc;  \
  } while (0)
#define HIPCHECK(ctx, call) CHECK((ctx)->hip((call), #call))

namespace mclg {

struct BatchState {
  int nscenes = 0;
  int nv = 0, np = 0;
  std::vector<int> voff, poff;
  std::vector<mclg_options> opts;
  int scene_bits = 1;
  // device scene data
  int* d_P = nullptr; size_t cap_P = 0;
  int* d_prim_scene = nullptr; size_t cap_prim_scene = 0;
  int* d_vert_scene = nullptr; size_t cap_vert_scene = 0;
  int* d_prim_off = nullptr; size_t cap_prim_off = 0;
  double* d_eta = nullptr; size_t cap_eta = 0;  // [box | vf | ee] x nscenes
  unsigned long long* d_sbounds = nullptr; size_t cap_sbounds = 0;
  uint32_t* d_khi = nullptr; size_t cap_khi = 0;
  uint32_t* d_mort = nullptr; size_t cap_mort = 0;
  double* d_V0 = nullptr; size_t cap_V0 = 0;
  double* d_V1 = nullptr; size_t cap_V1 = 0;
  unsigned long long* d_scene_counts = nullptr; size_t cap_scene_counts = 0;
  // per-vertex active mask of the concatenated vertices (BVHTree::update's
  // active, every scene at once); has_active false: every leaf active
  int* d_active = nullptr; size_t cap_active = 0;
  bool has_active = false;
  // graph
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_valid = false;
  double last_ms = 0;
};

static void drop_graph(BatchState* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_valid = false;
}

static SceneArgs scene_args(BatchState* b) {
  SceneArgs S;
  S.n = b->nscenes;
  S.prim_scene = b->d_prim_scene;
  S.vert_scene = b->d_vert_scene;
  S.prim_off = b->d_prim_off;
  S.box_eta = b->d_eta;
  S.vf_eta = b->d_eta + b->nscenes;
  S.ee_eta = b->d_eta + 2 * b->nscenes;
  S.bounds = b->d_sbounds;
  return S;
}

// Allocates every buffer a frame needs at the current capacities.
static int batch_buffers(Ctx* c) {
  BatchState* b = c->batch;
  const int np = b->np, nv = b->nv, dim = c->dim;
  const size_t box = dim == 3 ? sizeof(LeafBox<3>) : sizeof(LeafBox<2>);
  const size_t node = dim == 3 ? sizeof(NodeF<3>) : sizeof(NodeF<2>);
  CHECK(c->grow((void**)&c->d_rec, &c->cap_rec, sizeof(double) * (dim == 3 ? 8 : 4) * (size_t)nv, "vertex records"));
  CHECK(c->grow((void**)&c->d_boxes, &c->cap_boxes, box * (size_t)np, "leaf boxes"));
  CHECK(c->grow((void**)&c->d_act, &c->cap_act, (size_t)np, "active"));
  CHECK(c->grow((void**)&c->d_cbounds, &c->cap_cbounds, sizeof(unsigned long long) * 6, "bounds"));
  CHECK(c->grow((void**)&c->d_partial, &c->cap_partial, sizeof(double) * 6 * (size_t)leaf_box_blocks(np), "bounds"));
  CHECK(c->grow((void**)&c->d_keys, &c->cap_keys, sizeof(uint32_t) * (size_t)np, "keys"));
  CHECK(c->grow((void**)&c->d_order, &c->cap_order, sizeof(uint32_t) * (size_t)np, "order"));
  CHECK(c->grow((void**)&c->d_ktmp, &c->cap_ktmp, sizeof(uint32_t) * (size_t)np, "keys tmp"));
  CHECK(c->grow((void**)&c->d_vtmp, &c->cap_vtmp, sizeof(uint32_t) * (size_t)np, "vals tmp"));
  CHECK(c->grow((void**)&c->d_hist, &c->cap_hist, sizeof(uint32_t) * radix_hist_words(np), "histogram"));
  CHECK(c->grow((void**)&b->d_khi, &b->cap_khi, sizeof(uint32_t) * (size_t)np, "scene keys"));
  CHECK(c->grow((void**)&b->d_mort, &b->cap_mort, sizeof(uint32_t) * (size_t)np, "morton"));
  CHECK(c->grow((void**)&c->d_nodes, &c->cap_nodes, node * (size_t)(2 * np - 1), "nodes"));
  CHECK(c->grow(&c->d_cp, &c->cap_cp, child_pair_bytes(dim) * (size_t)std::max(np - 1, 1), "child pairs"));
  CHECK(c->grow((void**)&c->d_root, &c->cap_root, sizeof(int), "root"));
  CHECK(c->grow(&c->d_sboxes, &c->cap_sboxes, sorted_leaf_bytes(dim) * (size_t)np, "sorted leaf boxes"));
  CHECK(c->grow(&c->d_face, &c->cap_face, face_bytes((size_t)np), "face arrays"));
  CHECK(c->grow((void**)&c->d_v1, &c->cap_v1, sizeof(double) * v1_stride(dim) * (size_t)nv, "V1"));
  const size_t ni = (size_t)std::max(np - 1, 1);
  CHECK(c->grow((void**)&c->d_lchild, &c->cap_lchild, sizeof(int) * ni, "lchild"));
  CHECK(c->grow((void**)&c->d_rchild, &c->cap_rchild, sizeof(int) * ni, "rchild"));
  CHECK(c->grow((void**)&c->d_rngl, &c->cap_rngl, sizeof(int) * ni, "rng_l"));
  CHECK(c->grow((void**)&c->d_rngr, &c->cap_rngr, sizeof(int) * ni, "rng_r"));
  CHECK(c->grow((void**)&c->d_visits, &c->cap_visits, sizeof(int) * ni, "visits"));
  CHECK(c->grow((void**)&c->d_frontier, &c->cap_frontier, sizeof(int4) * (size_t)lbvh_frontier_cap(np), "frontier"));
  CHECK(c->grow((void**)&c->d_nfrontier, &c->cap_nfrontier, sizeof(unsigned long long), "frontier"));
  CHECK(c->grow((void**)&c->d_counters, &c->cap_counters, sizeof(unsigned long long) * C_COUNT, "counters"));
  if (c->pair_cap == 0) c->pair_cap = std::max<unsigned long long>(1024, 16ull * (unsigned long long)np);
  if (c->vf_cap == 0) c->vf_cap = std::max<unsigned long long>(1024, 16ull * (unsigned long long)np);
  if (c->ee_cap == 0 && dim == 3) c->ee_cap = std::max<unsigned long long>(1024, 32ull * (unsigned long long)np);
  if (c->disc_cap == 0) c->disc_cap = std::max<unsigned long long>(1024, (unsigned long long)np);
  if (c->hit_cap == 0) c->hit_cap = std::max<unsigned long long>(1024, 16ull * (unsigned long long)np);
  CHECK(c->grow((void**)&c->d_pairs, &c->cap_pairs_bytes, sizeof(int2) * c->pair_cap, "leaf pairs"));
  const long long nch = expand_chunks(c->pair_cap) + 1;
  CHECK(c->grow((void**)&c->d_ex_off, &c->cap_ex_off, sizeof(unsigned long long) * 2 * (size_t)nch, "candidate offsets"));
  CHECK(c->grow((void**)&c->d_ex_part, &c->cap_ex_part, sizeof(unsigned long long) * (size_t)scan_part_words(nch),
                "candidate offsets"));
  CHECK(c->grow((void**)&c->d_vfq, &c->cap_vfq, sizeof(int4) * c->vf_cap, "VF stencils"));
  if (dim == 3) CHECK(c->grow((void**)&c->d_eeq, &c->cap_eeq, sizeof(int4) * c->ee_cap, "EE stencils"));
  CHECK(c->grow((void**)&c->d_disc, &c->cap_disc, sizeof(int2) * c->disc_cap, "discrete"));
  CHECK(c->grow((void**)&c->d_hit_sten, &c->cap_hit_sten, sizeof(int4) * c->hit_cap, "hits"));
  CHECK(c->grow((void**)&c->d_hit_vf, &c->cap_hit_vf, c->hit_cap, "hits"));
  CHECK(c->grow((void**)&c->d_hit_toi, &c->cap_hit_toi, sizeof(double) * c->hit_cap, "hits"));
  CHECK(c->grow((void**)&c->d_vf_toi, &c->cap_vf_toi, sizeof(double) * c->vf_cap, "candidate toi"));
  if (dim == 3) CHECK(c->grow((void**)&c->d_ee_toi, &c->cap_ee_toi, sizeof(double) * c->ee_cap, "candidate toi"));
  {
    const long long nh = accd_chunks(c->vf_cap) + (dim == 3 ? accd_chunks(c->ee_cap) : 0) + 1;
    CHECK(c->grow((void**)&c->d_hit_off, &c->cap_hit_off, sizeof(unsigned long long) * (size_t)nh, "hit offsets"));
    CHECK(c->grow((void**)&c->d_hit_part, &c->cap_hit_part, sizeof(unsigned long long) * (size_t)scan_part_words(nh),
                  "hit offsets"));
  }
  return MCLG_OK;
}

// One frame: update + traverse of every scene.  Enqueues only (no sync, no
// allocation, no memset): safe to capture into a graph.
static void batch_enqueue(Ctx* c) {
  BatchState* b = c->batch;
  const int dim = c->dim, np = b->np;
  hipStream_t s = c->stream;
  const SceneArgs S = scene_args(b);
  const mclg_options& o = b->opts[0];
  launch_pack(dim, b->d_V0, b->d_V1, b->nv, c->d_rec, s, c->d_v1);
  launch_fill_u64(c->d_counters, C_COUNT, 0ull, s);
  launch_leaf_boxes(dim, c->d_rec, b->d_P, np, dim, b->has_active ? b->d_active : nullptr, 0.0, c->d_boxes, c->d_act,
                    c->d_partial, c->d_cbounds, S, s);
  launch_morton(dim, c->d_boxes, np, c->d_cbounds, S, c->d_keys, c->d_order, b->d_mort, s);
  launch_radix_sort(c->d_keys, c->d_order, c->d_ktmp, c->d_vtmp, c->d_hist, np, 32, s);
  launch_scene_keys(c->d_order, np, b->d_prim_scene, b->d_khi, s);
  launch_radix_sort(b->d_khi, c->d_order, c->d_ktmp, c->d_vtmp, c->d_hist, np, b->scene_bits, s);
  launch_gather_u32(c->d_order, np, b->d_mort, c->d_keys, s);
  launch_lbvh(dim, c->d_keys, b->d_khi, c->d_order, np, c->d_boxes, c->d_act, c->d_nodes, c->d_sboxes, c->d_lchild,
              c->d_rchild, c->d_rngl, c->d_rngr, c->d_visits, c->d_frontier, c->d_nfrontier, S, s, c->d_cp,
              c->d_root);
  const FaceArrays F = face_arrays(c->d_face, (size_t)np, c->d_v1);
  launch_face_records(dim, c->d_order, np, b->d_P, c->d_flags, nullptr, F, s);
  launch_traverse(dim, c->d_nodes, c->d_cp, c->d_root, c->d_sboxes, np, c->d_pairs, c->pair_cap, c->d_counters,
                  b->d_khi, b->d_prim_off, c->tr_waves, s);
  if (!(o.continuous || o.discrete)) return;
  ExpandLaunch E;
  E.pairs = c->d_pairs; E.npairs_dev = c->d_counters + C_LEAF_PAIRS; E.pair_cap = c->pair_cap; E.npairs_hint = -1;
  E.faces = F; E.discrete = o.discrete; E.continuous = o.continuous;
  {
    const long long nch = expand_chunks(c->pair_cap) + 1;
    E.off_vf = c->d_ex_off; E.off_ee = c->d_ex_off + nch; E.scan_part = c->d_ex_part;
  }
  E.vfq = c->d_vfq; E.vf_cap = c->vf_cap; E.eeq = c->d_eeq; E.ee_cap = dim == 3 ? c->ee_cap : 0;
  E.disc = c->d_disc; E.disc_cap = c->disc_cap; E.counters = c->d_counters;
  E.num_cus = c->num_cus; E.blocks_per_cu = c->blocks_per_cu; E.chunk = c->chunk;
  launch_expand(dim, E, s);
  if (!o.continuous) return;
  const long long ncv = accd_chunks(c->vf_cap), nce = dim == 3 ? accd_chunks(c->ee_cap) : 0;
  launch_fill_u64(c->d_hit_off, ncv + nce + 1, 0ull, s);
  AccdLaunch L;
  L.rec = c->d_rec; L.vert_scene = b->d_vert_scene;
  L.counters = c->d_counters; L.num_cus = c->num_cus; L.blocks_per_cu = c->blocks_per_cu; 
  L.q = c->d_vfq; L.n_dev = c->d_counters + C_CAND_VF; L.cap = c->vf_cap; L.n_hint = -1; L.eta = 0;
  L.scene_eta = S.vf_eta; L.cand_toi = c->d_vf_toi; L.chunk_hits = c->d_hit_off;
  L.one_sided = o.vf_one_sided != 0;
  // the ACCD kernels append the hits (counters[C_HITS] was cleared with the
  // others); an overflow reruns the frame with a larger list
  const bool direct = accd_direct_hits(c->vf_cap) && accd_direct_hits(c->ee_cap);
  if (direct) {  // per-candidate TOIs and per-chunk counts only serve the compaction pass
    L.hit_sten = c->d_hit_sten; L.hit_vf = c->d_hit_vf; L.hit_toi = c->d_hit_toi; L.hit_cap = c->hit_cap;
    L.cand_toi = nullptr; L.chunk_hits = nullptr;
  }
  launch_accd(dim, true, L, s);
  L.one_sided = 0;
  if (dim == 3) {
    L.q = c->d_eeq; L.n_dev = c->d_counters + C_CAND_EE; L.cap = c->ee_cap; L.scene_eta = S.ee_eta;
    if (!direct) { L.cand_toi = c->d_ee_toi; L.chunk_hits = c->d_hit_off + ncv; }
    launch_accd(dim, false, L, s);
  }
  if (direct) return;
  HitsLaunch H;
  H.vfq = c->d_vfq; H.n_vf_dev = c->d_counters + C_CAND_VF; H.vf_cap = c->vf_cap; H.vf_toi = c->d_vf_toi;
  H.eeq = c->d_eeq; H.n_ee_dev = c->d_counters + C_CAND_EE; H.ee_cap = c->ee_cap; H.ee_toi = c->d_ee_toi;
  H.off = c->d_hit_off; H.scan_part = c->d_hit_part; H.counters = c->d_counters;
  H.hit_sten = c->d_hit_sten; H.hit_vf = c->d_hit_vf; H.hit_toi = c->d_hit_toi; H.hit_cap = c->hit_cap;
  H.num_cus = c->num_cus; H.blocks_per_cu = c->blocks_per_cu;
  launch_hits(dim, H, true, s);
}

// Runs one frame on the staged inputs; grows and recaptures on overflow.
static int batch_frame(Ctx* c) {
  BatchState* b = c->batch;
  for (int attempt = 0;; ++attempt) {
    CHECK(batch_buffers(c));
    const auto t0 = std::chrono::steady_clock::now();
    if (b->use_graph) {
      if (!b->graph_valid) {
        drop_graph(b);
        HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        batch_enqueue(c);
        HIPCHECK(c, hipStreamEndCapture(c->stream, &b->graph));
        HIPCHECK(c, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_valid = true;
      }
      HIPCHECK(c, hipGraphLaunch(b->exec, c->stream));
    } else {
      batch_enqueue(c);
    }
    HIPCHECK(c, hipGetLastError());
    CHECK(c->read_counters());
    b->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    const unsigned long long* h = c->h_counters;
    if (h[C_ERRORS]) return c->fail(MCLG_ERR_STATE, "batch traversal: malformed BVH");
    bool again = false;
    auto bump = [&](unsigned long long need, unsigned long long& cap) {
      if (need > cap) { cap = need + need / 4; again = true; }
    };
    bump(h[C_LEAF_PAIRS], c->pair_cap);
    bump(h[C_CAND_VF], c->vf_cap);
    bump(h[C_CAND_EE], c->ee_cap);
    bump(h[C_DISCRETE], c->disc_cap);
    bump(h[C_HITS], c->hit_cap);
    if (!again) break;
    b->graph_valid = false;  // capacities are baked into the graph
    if (attempt > 6) return c->fail(MCLG_ERR_STATE, "batch: result sizes did not converge");
  }
  const unsigned long long* h = c->h_counters;
  c->counts.leaf_pairs = (int64_t)h[C_LEAF_PAIRS];
  c->n_vf = (long long)h[C_CAND_VF];
  c->n_ee = (long long)h[C_CAND_EE];
  c->counts.candidates = (int64_t)(h[C_CAND_VF] + h[C_CAND_EE]);
  c->counts.hits = (int64_t)h[C_HITS];
  c->counts.discrete = (int64_t)h[C_DISCRETE];
  c->counts.dist_evals = (int64_t)h[C_DIST_EVALS];
  c->counts.suspended = (int64_t)h[C_SUSPENDED];
  c->counts.vf_candidates = (int64_t)h[C_CAND_VF];
  c->counts.ee_dist_evals = (int64_t)h[C_EE_EVALS];
  c->counts.traverse_steps = (int64_t)h[C_TR_STEPS];
  c->keep_cand = false;
  return MCLG_OK;
}

__global__ void k_scene_hist(const int4* __restrict__ hits, long long nh, const int* __restrict__ vert_scene,
                             const int2* __restrict__ disc, long long nd, const int* __restrict__ prim_scene,
                             unsigned long long* counts, int nscenes) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < nh) atomicAdd(counts + vert_scene[hits[i].x], 1ull);
  if (i < nd) atomicAdd(counts + nscenes + prim_scene[disc[i].x], 1ull);
}

}  // namespace mclg

static Ctx* B(mclg_batch* h) { return reinterpret_cast<Ctx*>(h); }

extern "C" {

int mclg_batch_create(int dim, mclg_batch** out) {
  if (!out) return MCLG_ERR_ARG;
  mclg_ctx* c = nullptr;
  const int rc = mclg_create(dim, &c);
  if (rc != MCLG_OK) return rc;
  Ctx* x = reinterpret_cast<Ctx*>(c);
  x->batch = new BatchState();
  *out = reinterpret_cast<mclg_batch*>(x);
  return MCLG_OK;
}

void mclg_batch_destroy(mclg_batch* h) {
  Ctx* c = B(h);
  if (!c) return;
  BatchState* b = c->batch;
  if (b) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(b);
    void* bufs[] = {b->d_P, b->d_prim_scene, b->d_vert_scene, b->d_prim_off, b->d_eta, b->d_sbounds,
                    b->d_khi, b->d_mort, b->d_V0, b->d_V1, b->d_scene_counts};
    for (void* p : bufs)
      if (p) (void)hipFree(p);
    delete b;
    c->batch = nullptr;
  }
  mclg_destroy(reinterpret_cast<mclg_ctx*>(c));
}

const char* mclg_batch_last_error(const mclg_batch* h) {
  return mclg_last_error(reinterpret_cast<const mclg_ctx*>(h));
}

int mclg_batch_set_scenes(mclg_batch* h, int nscenes, const int* voff, const int* poff, const int* P,
                          const mclg_options* opts, int nopts) {
  Ctx* c = B(h);
  if (!c || !c->batch || nscenes < 1 || !voff || !poff || !P || !opts || (nopts != 1 && nopts != nscenes))
    return MCLG_ERR_ARG;
  BatchState* b = c->batch;
  const int dim = c->dim;
  if (voff[0] != 0 || poff[0] != 0) return c->fail(MCLG_ERR_ARG, "batch: offsets must start at 0");
  for (int sc = 0; sc < nscenes; ++sc)
    if (voff[sc + 1] < voff[sc] || poff[sc + 1] < poff[sc]) return c->fail(MCLG_ERR_ARG, "batch: offsets must not decrease");
  const int nv = voff[nscenes], np = poff[nscenes];
  if (np > (1 << 26)) return c->fail(MCLG_ERR_ARG, "batch: %d primitives exceeds the 2^26 limit", np);
  b->opts.assign(nscenes, opts[0]);
  for (int sc = 0; sc < nscenes; ++sc) {
    mclg_options o = opts[nopts == 1 ? 0 : sc];
    if (o.reptri != opts[0].reptri || o.discrete != opts[0].discrete || o.continuous != opts[0].continuous ||
        o.vf_one_sided != opts[0].vf_one_sided)
      return c->fail(MCLG_ERR_UNSUPPORTED, "batch: reptri/discrete/continuous/vf_one_sided must agree across scenes");
    o.vf_ccd_eta = std::min(o.vf_ccd_eta, o.box_eta);  // BVHTree.cpp:39-40
    o.ee_ccd_eta = std::min(o.ee_ccd_eta, o.box_eta);
    b->opts[sc] = o;
  }
  // global vertex indices, checked against each scene's own vertex range
  std::vector<int> Pg((size_t)np * dim), pscene(np), vscene(nv);
  for (int sc = 0; sc < nscenes; ++sc) {
    const int nvs = voff[sc + 1] - voff[sc];
    for (int f = poff[sc]; f < poff[sc + 1]; ++f) {
      pscene[f] = sc;
      for (int j = 0; j < dim; ++j) {
        const int v = P[(size_t)f * dim + j];
        if (v < 0 || v >= nvs) return c->fail(MCLG_ERR_ARG, "batch: scene %d primitive %d index %d outside [0,%d)", sc, f, v, nvs);
        Pg[(size_t)f * dim + j] = v + voff[sc];
      }
    }
    for (int v = voff[sc]; v < voff[sc + 1]; ++v) vscene[v] = sc;
  }
  b->nscenes = nscenes;
  b->nv = nv;
  b->np = np;
  b->voff.assign(voff, voff + nscenes + 1);
  b->poff.assign(poff, poff + nscenes + 1);
  b->scene_bits = 1;
  while ((1 << b->scene_bits) < nscenes) ++b->scene_bits;
  std::vector<double> eta(3 * (size_t)nscenes);
  for (int sc = 0; sc < nscenes; ++sc) {
    eta[sc] = b->opts[sc].box_eta;
    eta[nscenes + sc] = b->opts[sc].vf_ccd_eta;
    eta[2 * nscenes + sc] = b->opts[sc].ee_ccd_eta;
  }
  CHECK(c->grow((void**)&b->d_P, &b->cap_P, sizeof(int) * std::max<size_t>(Pg.size(), 1), "batch P"));
  CHECK(c->grow((void**)&b->d_prim_scene, &b->cap_prim_scene, sizeof(int) * std::max(np, 1), "batch"));
  CHECK(c->grow((void**)&b->d_vert_scene, &b->cap_vert_scene, sizeof(int) * std::max(nv, 1), "batch"));
  CHECK(c->grow((void**)&b->d_prim_off, &b->cap_prim_off, sizeof(int) * (nscenes + 1), "batch"));
  CHECK(c->grow((void**)&b->d_eta, &b->cap_eta, sizeof(double) * eta.size(), "batch"));
  CHECK(c->grow((void**)&b->d_sbounds, &b->cap_sbounds, sizeof(unsigned long long) * 2 * dim * nscenes, "batch"));
  CHECK(c->grow((void**)&b->d_V0, &b->cap_V0, sizeof(double) * dim * std::max<size_t>(nv, 1), "batch V0"));
  CHECK(c->grow((void**)&b->d_V1, &b->cap_V1, sizeof(double) * dim * std::max<size_t>(nv, 1), "batch V1"));
  HIPCHECK(c, hipMemcpy(b->d_P, Pg.data(), sizeof(int) * Pg.size(), hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(b->d_prim_scene, pscene.data(), sizeof(int) * np, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(b->d_vert_scene, vscene.data(), sizeof(int) * nv, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(b->d_prim_off, poff, sizeof(int) * (nscenes + 1), hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(b->d_eta, eta.data(), sizeof(double) * eta.size(), hipMemcpyHostToDevice));
  // representative triangles of the concatenated mesh equal each scene's own:
  // scenes share no vertex and keep their face order (BVHTree.cpp:53-79)
  CHECK(c->grow((void**)&c->d_flags, &c->cap_flags, (size_t)std::max(np, 1), "flags"));
  launch_fill_u8(c->d_flags, np, 0x3f, c->stream);
  if (b->opts[0].reptri && (dim == 2 || dim == 3) && np > 0) {
    CHECK(c->grow((void**)&c->d_owner, &c->cap_owner, sizeof(int) * (size_t)std::max(nv, 1), "vertex owners"));
    uint32_t ts = 1024;
    while (ts < 2u * 3u * (uint32_t)np) ts <<= 1;
    if (dim == 3) {
      CHECK(c->grow((void**)&c->d_ekeys, &c->cap_ekeys, sizeof(unsigned long long) * ts, "edge keys"));
      CHECK(c->grow((void**)&c->d_evals, &c->cap_evals, sizeof(int) * ts, "edge owners"));
    }
    launch_reptri(b->d_P, np, dim, nv, c->d_owner, c->d_ekeys, c->d_evals, ts, c->d_flags, c->stream);
  }
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->opt = b->opts[0];
  c->np = np;
  c->nv = nv;
  c->pdim = dim;
  c->have_tree = false;
  drop_graph(b);
  return MCLG_OK;
}

int mclg_batch_set_params(mclg_batch* h, int use_graph, int tr_waves, int blocks_per_cu, int chunk) {
  Ctx* c = B(h);
  if (!c || !c->batch) return MCLG_ERR_ARG;
  if (use_graph >= 0) c->batch->use_graph = use_graph != 0;
  if (tr_waves > 0) c->tr_waves = std::min(tr_waves, kTraverseMaxWaves);
  if (blocks_per_cu > 0) c->blocks_per_cu = blocks_per_cu;
  if (chunk > 0) c->chunk = chunk;
  drop_graph(c->batch);
  return MCLG_OK;
}

int mclg_batch_get_params(const mclg_batch* h, int* out4) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c || !c->batch || !out4) return MCLG_ERR_ARG;
  out4[0] = c->batch->use_graph;
  out4[1] = c->tr_waves;
  out4[2] = c->blocks_per_cu;
  out4[3] = c->chunk;
  return MCLG_OK;
}

int mclg_batch_set_active(mclg_batch* h, const int* active, int device_ptrs) {
  Ctx* c = B(h);
  if (!c || !c->batch) return MCLG_ERR_ARG;
  BatchState* b = c->batch;
  if (b->nscenes == 0) return c->fail(MCLG_ERR_STATE, "batch: no scenes set");
  const bool on = active != nullptr;
  if (on) {
    CHECK(c->grow((void**)&b->d_active, &b->cap_active, sizeof(int) * std::max<size_t>((size_t)b->nv, 1), "active"));
    if (b->nv > 0)
      HIPCHECK(c, hipMemcpyAsync(b->d_active, active, sizeof(int) * (size_t)b->nv,
                                 device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
  }
  if (on != b->has_active) drop_graph(b);  // the graph holds the mask's pointer (or none), not its contents
  b->has_active = on;
  return MCLG_OK;
}

int mclg_batch_run(mclg_batch* h, const double* V0, const double* V1, int device_ptrs) {
  Ctx* c = B(h);
  if (!c || !c->batch) return MCLG_ERR_ARG;
  BatchState* b = c->batch;
  if (b->nscenes == 0) return c->fail(MCLG_ERR_STATE, "batch: no scenes set");
  if (!V0 || !V1) return c->fail(MCLG_ERR_ARG, "batch: null vertices");
  const size_t bytes = sizeof(double) * c->dim * (size_t)b->nv;
  const hipMemcpyKind k = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  HIPCHECK(c, hipMemcpyAsync(b->d_V0, V0, bytes, k, c->stream));
  HIPCHECK(c, hipMemcpyAsync(b->d_V1, V1, bytes, k, c->stream));
  return batch_frame(c);
}

double mclg_batch_last_ms(const mclg_batch* h) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  return c && c->batch ? c->batch->last_ms : -1.0;
}

int mclg_batch_get_counts(const mclg_batch* h, mclg_counts* out) {
  return mclg_get_counts(reinterpret_cast<const mclg_ctx*>(h), out);
}

// Per-scene numbers of CCD hits and discrete hits (int64[nscenes] each).
int mclg_batch_scene_counts(mclg_batch* h, int64_t* hits, int64_t* discrete) {
  Ctx* c = B(h);
  if (!c || !c->batch) return MCLG_ERR_ARG;
  BatchState* b = c->batch;
  const int n = b->nscenes;
  CHECK(c->grow((void**)&b->d_scene_counts, &b->cap_scene_counts, sizeof(unsigned long long) * 2 * n, "scene counts"));
  launch_fill_u64(b->d_scene_counts, 2 * n, 0ull, c->stream);
  const long long nh = c->counts.hits, nd = c->counts.discrete, m = std::max(nh, nd);
  if (m > 0)
    k_scene_hist<<<(int)((m + 255) / 256), 256, 0, c->stream>>>(c->d_hit_sten, nh, b->d_vert_scene, c->d_disc, nd,
                                                                 b->d_prim_scene, b->d_scene_counts, n);
  std::vector<unsigned long long> hc(2 * (size_t)n);
  HIPCHECK(c, hipMemcpyAsync(hc.data(), b->d_scene_counts, sizeof(unsigned long long) * 2 * n, hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; ++i) {
    if (hits) hits[i] = (int64_t)hc[i];
    if (discrete) discrete[i] = (int64_t)hc[n + i];
  }
  return MCLG_OK;
}

// Hits with their scene and scene-local vertex indices (-1 padding kept).
int mclg_batch_copy_hits(mclg_batch* h, int* scene, int* sten4, unsigned char* is_vf, double* toi, int64_t cap) {
  Ctx* c = B(h);
  if (!c || !c->batch) return MCLG_ERR_ARG;
  BatchState* b = c->batch;
  const int64_t n = c->counts.hits;
  std::vector<int> g((size_t)n * 4);
  CHECK(mclg_copy_hits(reinterpret_cast<mclg_ctx*>(c), g.data(), is_vf, toi, cap));
  for (int64_t i = 0; i < n; ++i) {
    const int v0 = g[4 * i];
    const int sc = (int)(std::upper_bound(b->voff.begin(), b->voff.end(), v0) - b->voff.begin()) - 1;
    if (scene) scene[i] = sc;
    if (sten4)
      for (int k = 0; k < 4; ++k) sten4[4 * i + k] = g[4 * i + k] < 0 ? -1 : g[4 * i + k] - b->voff[sc];
  }
  return MCLG_OK;
}

// Discrete hits with their scene and scene-local primitive indices.
int mclg_batch_copy_discrete(mclg_batch* h, int* scene, int* pairs2, int64_t cap) {
  Ctx* c = B(h);
  if (!c || !c->batch) return MCLG_ERR_ARG;
  BatchState* b = c->batch;
  const int64_t n = c->counts.discrete;
  std::vector<int> g((size_t)n * 2);
  CHECK(mclg_copy_discrete(reinterpret_cast<mclg_ctx*>(c), g.data(), cap));
  for (int64_t i = 0; i < n; ++i) {
    const int sc = (int)(std::upper_bound(b->poff.begin(), b->poff.end(), g[2 * i]) - b->poff.begin()) - 1;
    if (scene) scene[i] = sc;
    if (pairs2) {
      pairs2[2 * i] = g[2 * i] - b->poff[sc];
      pairs2[2 * i + 1] = g[2 * i + 1] - b->poff[sc];
    }
  }
  return MCLG_OK;
}

// Picks waves per traversal block, resident blocks per CU and the queue chunk
// size by timing frames on the staged inputs; the choice is cached per
// (architecture, dim, scene count, primitive count) in cache_path (NULL: no cache).
int mclg_batch_autotune(mclg_batch* h, const char* cache_path, int reps) {
  Ctx* c = B(h);
  if (!c || !c->batch) return MCLG_ERR_ARG;
  BatchState* b = c->batch;
  if (b->nscenes == 0) return c->fail(MCLG_ERR_STATE, "autotune: no scenes");
  hipDeviceProp_t prop;
  HIPCHECK(c, hipGetDeviceProperties(&prop, c->device));
  std::string arch = prop.gcnArchName;
  arch = arch.substr(0, arch.find(':'));
  std::ostringstream key;
  // the format tag changes whenever the parameters' meaning or ranges do, so
  // entries written by another version are never reused
  key << "v2 " << arch << " dim=" << c->dim << " scenes=" << b->nscenes << " prims=" << b->np
      << " graph=" << b->use_graph;
  if (cache_path) {
    std::ifstream in(cache_path);
    std::string line;
    while (std::getline(in, line)) {
      const size_t bar = line.find('|');
      if (bar == std::string::npos || line.substr(0, bar) != key.str()) continue;
      int tw, bpc, ch;
      // a hand-edited or damaged entry is ignored (and re-tuned), never clamped silently
      if (std::sscanf(line.c_str() + bar + 1, "%d %d %d", &tw, &bpc, &ch) == 3 && tw >= 1 &&
          tw <= kTraverseMaxWaves && bpc >= 1 && ch >= 1) {
        c->tr_waves = tw;
        c->blocks_per_cu = bpc;
        c->chunk = ch;
        drop_graph(b);
        return MCLG_OK;
      }
    }
  }
  reps = reps < 1 ? 3 : reps;
  double best = 1e30;
  int bt = c->tr_waves, bb = c->blocks_per_cu, bc = c->chunk;
  const int tws[] = {2, kTraverseMaxWaves}, bpcs[] = {4, 8, 16, 32, 64}, chs[] = {1, 4};
  for (int tw : tws)
    for (int bpc : bpcs)
      for (int ch : chs) {
        c->tr_waves = tw;
        c->blocks_per_cu = bpc;
        c->chunk = ch;
        drop_graph(b);
        CHECK(batch_frame(c));  // capture (and size) outside the timing
        double t = 1e30;
        for (int r = 0; r < reps; ++r) {
          CHECK(batch_frame(c));
          t = std::min(t, b->last_ms);
        }
        if (t < best) { best = t; bt = tw; bb = bpc; bc = ch; }
      }
  c->tr_waves = bt;
  c->blocks_per_cu = bb;
  c->chunk = bc;
  drop_graph(b);
  if (cache_path) {
    std::ofstream out(cache_path, std::ios::app);
    out << key.str() << "|" << bt << " " << bb << " " << bc << " " << best << "\n";
  }
  return MCLG_OK;
}

}  // extern "C"
