// C ABI (include/mclg.h) over the gfx950 kernels: context, device buffers,
// update / traverse orchestration, result copies.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mclg.h"
#include "context.hpp"
#include "kernels.hpp"

using namespace mclg;

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
namespace mclg {

int Ctx::fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  err = buf;
  return code;
}

int Ctx::hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return MCLG_OK;
  (void)hipGetLastError();
  return fail(MCLG_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

int Ctx::grow(void** p, size_t* cap, size_t need, const char* what) {
  if (need <= *cap && *p) return MCLG_OK;
  size_t n = std::max(need + need / 4, (size_t)256);
  if (*p) {
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip(e, "hipStreamSynchronize");
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
  }
  if (hipMalloc(p, n) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(MCLG_ERR_NOMEM, "hipMalloc of %zu bytes for %s failed", n, what);
  }
  *cap = n;
  return MCLG_OK;
}

void Ctx::mark(int i) {
  if (timing) (void)hipEventRecord(ev[i], stream);
}

}  // namespace mclg

#define CHECK(x)                     \
  do {                               \
    int _rc = (x);                   \
    if (_rc != MCLG_OK) return _rc;  \
  } while (0)
#define HIPCHECK(ctx, call) CHECK((ctx)->hip((call), #call))

static size_t ldbox_bytes(int dim) { return dim == 3 ? sizeof(LeafBox<3>) : sizeof(LeafBox<2>); }
static size_t node_bytes(int dim) { return dim == 3 ? sizeof(NodeF<3>) : sizeof(NodeF<2>); }

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

void mclg_default_options(mclg_options* o) {
  if (!o) return;
  o->box_eta = 1e-6;
  o->vf_ccd_eta = 1e-6;
  o->ee_ccd_eta = 1e-6;
  o->reptri = 1;
  o->threaded = 1;
  o->discrete = 1;
  o->continuous = 1;
  o->vf_one_sided = 0;
  o->ee_robust = 1;
}

// Why the last mclg_create failed (mclg_last_error(NULL)).
static thread_local std::string g_create_error = "no error";

int mclg_create(int dim, mclg_ctx** out) {
  if (!out) return MCLG_ERR_ARG;
  *out = nullptr;
  if (dim != 2 && dim != 3) return MCLG_ERR_ARG;
  Ctx* c = new Ctx();
  c->dim = dim;
  mclg_default_options(&c->opt);
  auto failed = [&](hipError_t e, const char* what) {
    g_create_error = std::string("mclg_create: ") + what + ": " + hipGetErrorString(e);
    delete c;
    return MCLG_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipGetDevice(&c->device)) != hipSuccess) return failed(e, "hipGetDevice");
  if ((e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking)) != hipSuccess)
    return failed(e, "hipStreamCreateWithFlags");
  c->stream = c->own_stream;
  for (int i = 0; i < kNumEvents; ++i)
    if ((e = hipEventCreate(&c->ev[i])) != hipSuccess) return failed(e, "hipEventCreate");
  int cus = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
  c->num_cus = cus > 0 ? cus : 256;
  if ((e = hipHostMalloc((void**)&c->h_counters, sizeof(unsigned long long) * C_COUNT, 0)) != hipSuccess)
    return failed(e, "hipHostMalloc");
  if ((e = hipHostMalloc((void**)&c->h_minmax, sizeof(int) * 2, 0)) != hipSuccess) return failed(e, "hipHostMalloc");
  *out = reinterpret_cast<mclg_ctx*>(c);
  return MCLG_OK;
}

void mclg_destroy(mclg_ctx* h) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  delete c;
}

int mclg_set_options(mclg_ctx* h, const mclg_options* o) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !o) return MCLG_ERR_ARG;
  c->opt = *o;
  return MCLG_OK;
}

int mclg_get_options(const mclg_ctx* h, mclg_options* o) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c || !o) return MCLG_ERR_ARG;
  *o = c->opt;
  return MCLG_OK;
}

const char* mclg_last_error(const mclg_ctx* h) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  return c ? c->err.c_str() : g_create_error.c_str();
}

int mclg_set_stream(mclg_ctx* h, void* s) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  c->stream = s ? (hipStream_t)s : c->own_stream;
  return MCLG_OK;
}

}  // extern "C"

namespace mclg {

// Checks that every index of P (n ints) lies in [0, nv).
int Ctx::check_indices(const int* P, long long n, int nv, const char* what) {
  if (n <= 0) return MCLG_OK;
  CHECK(grow((void**)&d_minmax, &cap_minmax, sizeof(int) * 2, "minmax"));
  launch_fill_i32(d_minmax, 1, 0x7fffffff, stream);
  launch_fill_i32(d_minmax + 1, 1, -0x7fffffff - 1, stream);
  launch_minmax_i32(P, n, d_minmax, stream);
  HIPCHECK(this, hipMemcpyAsync(h_minmax, d_minmax, sizeof(int) * 2, hipMemcpyDeviceToHost, stream));
  HIPCHECK(this, hipStreamSynchronize(stream));
  if (h_minmax[0] < 0 || h_minmax[1] >= nv)
    return fail(MCLG_ERR_ARG, "%s: index range [%d, %d] outside [0, %d)", what, h_minmax[0], h_minmax[1], nv);
  return MCLG_OK;
}

// BVHTree.cpp:27-126
int Ctx::update(const double* V0, const double* V1, int nv_, const int* P, int np_, int pdim_, const int* active,
                bool refit) {
  if (np_ < 0 || pdim_ < 1 || nv_ < 0) return fail(MCLG_ERR_ARG, "update: bad sizes nv=%d np=%d pdim=%d", nv_, np_, pdim_);
  if (refit && !(have_tree && np_ == np && nv_ == nv && pdim_ == pdim))
    return fail(MCLG_ERR_STATE, "refit: needs a previous update with the same nv, np and pdim (have %d/%d/%d, got %d/%d/%d)",
                have_tree ? nv : 0, have_tree ? np : 0, have_tree ? pdim : 0, nv_, np_, pdim_);
  ++tree_gen;  // from here on the last traverse's hits no longer match the vertex records
  if (np_ == 0) {  // BVHTree.cpp:29-34
    np = 0;
    have_tree = false;
    rank_set = false;
    kd_valid = false;
    return MCLG_OK;
  }
  if (nv_ == 0 || !V0 || !V1 || !P) return fail(MCLG_ERR_ARG, "update: null input");
  if (np_ > (1 << 26)) return fail(MCLG_ERR_ARG, "update: %d primitives exceeds the 2^26 limit", np_);
  if (precision == MCLG_PRECISION_F32) {  // BVHTree<float,DIM>: the options are floats
    opt.box_eta = (double)(float)opt.box_eta;
    opt.vf_ccd_eta = (double)(float)opt.vf_ccd_eta;
    opt.ee_ccd_eta = (double)(float)opt.ee_ccd_eta;
  }
  const bool f32 = precision == MCLG_PRECISION_F32;
  // BVHTree.cpp:39-40
  opt.vf_ccd_eta = std::min(opt.vf_ccd_eta, opt.box_eta);
  opt.ee_ccd_eta = std::min(opt.ee_ccd_eta, opt.box_eta);
  timing = true;
  mark(0);
  CHECK(check_indices(P, (long long)np_ * pdim_, nv_, "update: P"));
  const int stride = dim == 3 ? 8 : 4;
  CHECK(grow((void**)&d_rec, &cap_rec, sizeof(double) * stride * (size_t)nv_, "vertex records"));
  launch_pack(dim, V0, V1, nv_, d_rec, stream, nullptr, f32);
  mark(1);
  // leaf resize + reptri (BVHTree.cpp:44-80): flags survive a resize for the
  // surviving leaves; new leaves start at (1,1,1)/(1,1,1) as BVHLeaf() sets them
  bool redo = false;
  if (np_ != np) {
    if ((size_t)np_ > cap_flags / 1 || !d_flags) {
      uint8_t* old = d_flags;
      size_t oldcap = cap_flags;
      d_flags = nullptr;
      cap_flags = 0;
      CHECK(grow((void**)&d_flags, &cap_flags, (size_t)np_, "flags"));
      if (old) {
        if (np > 0) HIPCHECK(this, hipMemcpyAsync(d_flags, old, (size_t)std::min(np, np_), hipMemcpyDeviceToDevice, stream));
        HIPCHECK(this, hipStreamSynchronize(stream));
        (void)hipFree(old);
      }
      (void)oldcap;
    }
    if (np_ > np) launch_fill_u8(d_flags + np, np_ - np, 0x3f, stream);
    redo = pdim_ == 2 || pdim_ == 3;
    rank_set = false;
  }
  if (opt.reptri && redo) {
    CHECK(grow((void**)&d_owner, &cap_owner, sizeof(int) * (size_t)nv_, "vertex owners"));
    uint32_t ts = 1024;
    while (ts < 2u * 3u * (uint32_t)np_) ts <<= 1;
    if (pdim_ == 3) {
      CHECK(grow((void**)&d_ekeys, &cap_ekeys, sizeof(unsigned long long) * ts, "edge keys"));
      CHECK(grow((void**)&d_evals, &cap_evals, sizeof(int) * ts, "edge owners"));
    }
    launch_reptri(P, np_, pdim_, nv_, d_owner, d_ekeys, d_evals, ts, d_flags, stream);
  }
  np = np_;
  nv = nv_;
  pdim = pdim_;
  mark(2);
  // leaf boxes
  CHECK(grow((void**)&d_boxes, &cap_boxes, ldbox_bytes(dim) * (size_t)np, "leaf boxes"));
  CHECK(grow((void**)&d_act, &cap_act, (size_t)np, "active"));
  CHECK(grow((void**)&d_cbounds, &cap_cbounds, sizeof(unsigned long long) * 6, "bounds"));
  CHECK(grow((void**)&d_partial, &cap_partial, sizeof(double) * 6 * (size_t)leaf_box_blocks(np), "bounds"));
  launch_leaf_boxes(dim, d_rec, P, np, pdim, active, opt.box_eta, d_boxes, d_act, d_partial, d_cbounds, SceneArgs(), stream,
                    f32);
  // the reference's KdBVH order (KdBVH.hpp:192-222) when asked for, or for a
  // serial traversal whose callback order must be the reference's
  if (orient == MCLG_ORIENT_KDBVH || !opt.threaded) {
    CHECK(kdbvh_from_leaves());
  } else if (kd_valid) {  // ranks of an earlier KdBVH update no longer apply
    rank_set = false;
    kd_valid = false;
  }
  mark(3);
  // Morton + sort
  CHECK(grow((void**)&d_keys, &cap_keys, sizeof(uint32_t) * (size_t)np, "keys"));
  CHECK(grow((void**)&d_order, &cap_order, sizeof(uint32_t) * (size_t)np, "order"));
  CHECK(grow((void**)&d_ktmp, &cap_ktmp, sizeof(uint32_t) * (size_t)np, "keys tmp"));
  CHECK(grow((void**)&d_vtmp, &cap_vtmp, sizeof(uint32_t) * (size_t)np, "vals tmp"));
  CHECK(grow((void**)&d_hist, &cap_hist, sizeof(uint32_t) * radix_hist_words(np), "histogram"));
  if (!refit) launch_morton(dim, d_boxes, np, d_cbounds, SceneArgs(), d_keys, d_order, nullptr, stream);
  mark(4);
  if (!refit) launch_radix_sort(d_keys, d_order, d_ktmp, d_vtmp, d_hist, np, 32, stream);
  mark(5);
  // LBVH
  CHECK(grow((void**)&d_nodes, &cap_nodes, node_bytes(dim) * (size_t)(2 * np - 1), "nodes"));
  const size_t ni = (size_t)std::max(np - 1, 1);
  CHECK(grow((void**)&d_lchild, &cap_lchild, sizeof(int) * ni, "lchild"));
  CHECK(grow((void**)&d_rchild, &cap_rchild, sizeof(int) * ni, "rchild"));
  CHECK(grow((void**)&d_rngl, &cap_rngl, sizeof(int) * ni, "rng_l"));
  CHECK(grow((void**)&d_rngr, &cap_rngr, sizeof(int) * ni, "rng_r"));
  CHECK(grow((void**)&d_visits, &cap_visits, sizeof(int) * ni, "visits"));
  CHECK(grow((void**)&d_frontier, &cap_frontier, sizeof(int4) * (size_t)lbvh_frontier_cap(np), "frontier"));
  CHECK(grow((void**)&d_nfrontier, &cap_nfrontier, sizeof(unsigned long long), "frontier"));
  CHECK(grow(&d_sboxes, &cap_sboxes, sorted_leaf_bytes(dim) * (size_t)np, "sorted leaf boxes"));
  CHECK(grow(&d_cp, &cap_cp, child_pair_bytes(dim) * ni, "child pairs"));
  CHECK(grow((void**)&d_root, &cap_root, sizeof(int), "root"));
  launch_lbvh(dim, d_keys, nullptr, d_order, np, d_boxes, d_act, d_nodes, d_sboxes, d_lchild, d_rchild, d_rngl, d_rngr,
              d_visits, d_frontier, d_nfrontier, SceneArgs(), stream, d_cp, d_root);
  mark(6);
  HIPCHECK(this, hipGetLastError());
  HIPCHECK(this, hipStreamSynchronize(stream));
  for (int i = 0; i < 6; ++i) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
    times[i] = ms;
  }
  have_tree = true;
  return MCLG_OK;
}

int Ctx::read_counters() {
  HIPCHECK(this, hipMemcpyAsync(h_counters, d_counters, sizeof(unsigned long long) * C_COUNT, hipMemcpyDeviceToHost, stream));
  HIPCHECK(this, hipStreamSynchronize(stream));
  return MCLG_OK;
}

// BVHTree.cpp:129-159 (+ collide / get_candidates / narrow phase / discrete)
int Ctx::traverse(const double* V0, const double* V1, const int* P, unsigned flags) {
  counts = mclg_counts{};
  hits_gen = kNoGen;
  hits_one_sided = opt.vf_one_sided != 0;
  hits_f32 = precision == MCLG_PRECISION_F32;
  keep_cand = (flags & MCLG_KEEP_CANDIDATES) != 0;
  keep_pairs = (flags & MCLG_KEEP_LEAF_PAIRS) != 0;
  check_order = (flags & MCLG_CHECK_ORDER) != 0;
  if (!have_tree || np <= 1) {  // BVHTree.cpp:131-132
    hits_gen = tree_gen;
    return MCLG_OK;
  }
  if (pdim != dim) return fail(MCLG_ERR_ARG, "traverse: primitives must have pdim == dim (%d != %d)", pdim, dim);
  if (!V0 || !V1 || !P) return fail(MCLG_ERR_ARG, "traverse: null input");
  timing = true;
  mark(7);
  CHECK(check_indices(P, (long long)np * dim, nv, "traverse: P"));
  CHECK(grow((void**)&d_v1, &cap_v1, sizeof(double) * v1_stride(dim) * (size_t)nv, "V1"));
  launch_pack(dim, V0, V1, nv, d_rec, stream, d_v1, hits_f32);
  CHECK(grow(&d_face, &cap_face, face_bytes((size_t)np), "face arrays"));
  launch_face_records(dim, d_order, np, P, d_flags, rank_set ? d_rank : nullptr, faces(), stream);
  pairs_are_prims = false;
  CHECK(grow((void**)&d_counters, &cap_counters, sizeof(unsigned long long) * C_COUNT, "counters"));
  launch_fill_u64(d_counters, C_COUNT, 0ull, stream);
  if (pair_cap == 0) pair_cap = std::max<unsigned long long>(1024, 16ull * (unsigned long long)np);
  CHECK(grow((void**)&d_pairs, &cap_pairs_bytes, sizeof(int2) * pair_cap, "leaf pairs"));
  for (;;) {
    mark(8);
    launch_traverse(dim, d_nodes, d_cp, d_root, d_sboxes, np, d_pairs, pair_cap, d_counters, nullptr, nullptr, tr_waves,
                    stream);
    mark(9);
    HIPCHECK(this, hipGetLastError());
    CHECK(read_counters());
    if (h_counters[C_ERRORS]) return fail(MCLG_ERR_STATE, "traversal: malformed BVH (walk exceeded 2n steps)");
    if (h_counters[C_LEAF_PAIRS] <= pair_cap) break;
    pair_cap = h_counters[C_LEAF_PAIRS] + h_counters[C_LEAF_PAIRS] / 4;
    CHECK(grow((void**)&d_pairs, &cap_pairs_bytes, sizeof(int2) * pair_cap, "leaf pairs"));
    launch_fill_u64(d_counters, C_COUNT, 0ull, stream);
  }
  const long long npairs = (long long)h_counters[C_LEAF_PAIRS];
  counts.leaf_pairs = npairs;
  counts.traverse_steps = (int64_t)h_counters[C_TR_STEPS];
  n_vf = n_ee = 0;
  if (npairs > 0 && (opt.continuous || opt.discrete)) {
    // candidates + discrete (BVHTree.cpp:252-283)
    if (vf_cap == 0) vf_cap = std::max<unsigned long long>(1024, (unsigned long long)(1.5 * npairs));
    if (ee_cap == 0 && dim == 3) ee_cap = std::max<unsigned long long>(1024, (unsigned long long)(2.5 * npairs));
    if (disc_cap == 0) disc_cap = 1024;
    if (odep_cap == 0) odep_cap = 256;
    for (int attempt = 0;; ++attempt) {
      CHECK(grow((void**)&d_vfq, &cap_vfq, sizeof(int4) * vf_cap, "VF stencils"));
      if (dim == 3) CHECK(grow((void**)&d_eeq, &cap_eeq, sizeof(int4) * ee_cap, "EE stencils"));
      CHECK(grow((void**)&d_disc, &cap_disc, sizeof(int2) * disc_cap, "discrete"));
      if (check_order) CHECK(grow((void**)&d_odep, &cap_odep, sizeof(int2) * odep_cap, "order-dependent pairs"));
      launch_fill_u64(d_counters + C_CAND_VF, C_COUNT - C_CAND_VF, 0ull, stream);
      ExpandLaunch E;
      E.pairs = d_pairs; E.npairs_dev = d_counters + C_LEAF_PAIRS; E.pair_cap = pair_cap; E.npairs_hint = npairs;
      E.faces = faces(); E.blocks_per_cu = blocks_per_cu; E.chunk = chunk;
      const long long nch = expand_chunks(pair_cap) + 1;
      CHECK(grow((void**)&d_ex_off, &cap_ex_off, sizeof(unsigned long long) * 2 * (size_t)nch, "candidate offsets"));
      CHECK(grow((void**)&d_ex_part, &cap_ex_part, sizeof(unsigned long long) * (size_t)scan_part_words(nch),
                 "candidate offsets"));
      E.off_vf = d_ex_off; E.off_ee = d_ex_off + nch; E.scan_part = d_ex_part;
      E.discrete = opt.discrete; E.continuous = opt.continuous;
      E.vfq = d_vfq; E.vf_cap = vf_cap; E.eeq = d_eeq; E.ee_cap = dim == 3 ? ee_cap : 0;
      E.disc = d_disc; E.disc_cap = disc_cap; E.counters = d_counters; E.num_cus = num_cus;
      E.odep = check_order ? d_odep : nullptr; E.odep_cap = odep_cap;
      mark(10);
      launch_expand(dim, E, stream);
      mark(11);
      HIPCHECK(this, hipGetLastError());
      CHECK(read_counters());
      bool again = false;
      if (h_counters[C_CAND_VF] > vf_cap) { vf_cap = h_counters[C_CAND_VF] + h_counters[C_CAND_VF] / 4; again = true; }
      if (h_counters[C_CAND_EE] > ee_cap) { ee_cap = h_counters[C_CAND_EE] + h_counters[C_CAND_EE] / 4; again = true; }
      if (h_counters[C_DISCRETE] > disc_cap) { disc_cap = h_counters[C_DISCRETE] + h_counters[C_DISCRETE] / 4; again = true; }
      if (check_order && h_counters[C_ORDER_DEP] > odep_cap) {
        odep_cap = h_counters[C_ORDER_DEP] + h_counters[C_ORDER_DEP] / 4;
        again = true;
      }
      if (!again) break;
      if (attempt > 4) return fail(MCLG_ERR_STATE, "candidate generation: sizes did not converge");
    }
    n_vf = (long long)h_counters[C_CAND_VF];
    n_ee = (long long)h_counters[C_CAND_EE];
    counts.candidates = n_vf + n_ee;
    counts.vf_candidates = n_vf;
    counts.discrete = (int64_t)h_counters[C_DISCRETE];
    counts.order_dependent = check_order ? (int64_t)h_counters[C_ORDER_DEP] : 0;
    // FP64 ACCD over the two queues (BVHTree.cpp:428-461)
    if (opt.continuous && n_vf + n_ee > 0) {
      if (hit_cap == 0) hit_cap = std::max<unsigned long long>(1024, (unsigned long long)(n_vf + n_ee) / 2);
      // the ACCD kernels append the hit list themselves; the scan + compaction
      // pass below only runs when it overflowed hit_cap.  A list with room for
      // every candidate cannot overflow: then the per-candidate TOIs are only
      // written for MCLG_KEEP_CANDIDATES and the per-chunk hit counts not at all.
      const bool direct = accd_direct_hits(vf_cap) && accd_direct_hits(ee_cap);
      const bool no_overflow = direct && hit_cap >= (unsigned long long)(n_vf + n_ee);
      const bool cand_toi = keep_cand || !no_overflow;
      if (cand_toi) {
        CHECK(grow((void**)&d_vf_toi, &cap_vf_toi, sizeof(double) * (size_t)std::max(n_vf, 1LL), "candidate toi"));
        CHECK(grow((void**)&d_ee_toi, &cap_ee_toi, sizeof(double) * (size_t)std::max(n_ee, 1LL), "candidate toi"));
      }
      const long long ncv = accd_chunks(vf_cap), nce = dim == 3 ? accd_chunks(ee_cap) : 0;
      if (!no_overflow) {
        CHECK(grow((void**)&d_hit_off, &cap_hit_off, sizeof(unsigned long long) * (size_t)(ncv + nce + 1),
                   "hit offsets"));
        CHECK(grow((void**)&d_hit_part, &cap_hit_part,
                   sizeof(unsigned long long) * (size_t)scan_part_words(ncv + nce + 1), "hit offsets"));
        launch_fill_u64(d_hit_off, ncv + nce + 1, 0ull, stream);
      }
      launch_fill_u64(d_counters + C_HITS, 1, 0ull, stream);
      launch_fill_u64(d_counters + C_DIST_EVALS, 1, 0ull, stream);
      launch_fill_u64(d_counters + C_SUSPENDED, 1, 0ull, stream);
      AccdLaunch L;
      L.rec = d_rec; L.counters = d_counters; L.num_cus = num_cus; L.blocks_per_cu = blocks_per_cu; 
      L.vert_scene = nullptr; L.scene_eta = nullptr; L.f32 = hits_f32;
      if (direct) {
        CHECK(grow((void**)&d_hit_sten, &cap_hit_sten, sizeof(int4) * hit_cap, "hits"));
        CHECK(grow((void**)&d_hit_vf, &cap_hit_vf, hit_cap, "hits"));
        CHECK(grow((void**)&d_hit_toi, &cap_hit_toi, sizeof(double) * hit_cap, "hits"));
        L.hit_sten = d_hit_sten; L.hit_vf = d_hit_vf; L.hit_toi = d_hit_toi; L.hit_cap = hit_cap;
        if (no_overflow && opt.vf_one_sided && dim == 3) {  // CTCD on the prefilter's survivors only
          CHECK(grow((void**)&d_ctcd_sel, &cap_ctcd_sel, sizeof(uint32_t) * (size_t)vf_cap, "CTCD selection"));
          launch_fill_u64(d_counters + C_CTCD_SEL, 1, 0ull, stream);
          L.ctcd_sel = d_ctcd_sel;
        }
      }
      mark(12);
      L.q = d_vfq; L.n_dev = d_counters + C_CAND_VF; L.cap = vf_cap; L.n_hint = n_vf;
      L.eta = opt.vf_ccd_eta; L.cand_toi = cand_toi ? d_vf_toi : nullptr;
      L.chunk_hits = no_overflow ? nullptr : d_hit_off;
      L.one_sided = opt.vf_one_sided != 0;
      launch_accd(dim, true, L, stream);
      L.one_sided = 0;
      mark(16);
      L.q = d_eeq; L.n_dev = d_counters + C_CAND_EE; L.cap = ee_cap; L.n_hint = n_ee;
      L.eta = opt.ee_ccd_eta; L.cand_toi = cand_toi ? d_ee_toi : nullptr;
      L.chunk_hits = no_overflow ? nullptr : d_hit_off + ncv;
      if (dim == 3) launch_accd(dim, false, L, stream);
      mark(17);
      HitsLaunch H;
      H.vfq = d_vfq; H.n_vf_dev = d_counters + C_CAND_VF; H.vf_cap = vf_cap; H.vf_toi = d_vf_toi;
      H.eeq = d_eeq; H.n_ee_dev = d_counters + C_CAND_EE; H.ee_cap = ee_cap; H.ee_toi = d_ee_toi;
      H.off = d_hit_off; H.scan_part = d_hit_part; H.counters = d_counters;
      H.num_cus = num_cus; H.blocks_per_cu = blocks_per_cu;
      bool scanned = false;
      for (int attempt = 0;; ++attempt) {
        if (!(direct && attempt == 0)) {
          CHECK(grow((void**)&d_hit_sten, &cap_hit_sten, sizeof(int4) * hit_cap, "hits"));
          CHECK(grow((void**)&d_hit_vf, &cap_hit_vf, hit_cap, "hits"));
          CHECK(grow((void**)&d_hit_toi, &cap_hit_toi, sizeof(double) * hit_cap, "hits"));
          H.hit_sten = d_hit_sten; H.hit_vf = d_hit_vf; H.hit_toi = d_hit_toi; H.hit_cap = hit_cap;
          launch_hits(dim, H, !scanned, stream);  // a regrowth redoes the compaction only
          scanned = true;
        }
        mark(13);
        HIPCHECK(this, hipGetLastError());
        CHECK(read_counters());
        if (h_counters[C_HITS] <= hit_cap) break;
        if (no_overflow) return fail(MCLG_ERR_STATE, "narrow phase: more hits than candidates");
        // hits never outnumber candidates: a list that size never needs this pass
        // again, and lets the ACCD kernels suspend long-running stencils
        hit_cap = (unsigned long long)(n_vf + n_ee);
        if (attempt > 4) return fail(MCLG_ERR_STATE, "narrow phase: sizes did not converge");
      }
      counts.hits = (int64_t)h_counters[C_HITS];
      counts.dist_evals = (int64_t)h_counters[C_DIST_EVALS];
      counts.suspended = (int64_t)h_counters[C_SUSPENDED];
      counts.ee_dist_evals = (int64_t)h_counters[C_EE_EVALS];
    }
  }
  if (keep_pairs && npairs > 0) {  // sorted positions -> primitive pairs, lower rank first
    launch_pairs_to_prims(d_pairs, npairs, faces(), stream);
    HIPCHECK(this, hipGetLastError());
    pairs_are_prims = true;
  }
  HIPCHECK(this, hipStreamSynchronize(stream));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ev[7], ev[8]);
  times[6] = ms;
  (void)hipEventElapsedTime(&ms, ev[8], ev[9]);
  times[7] = ms;
  times[8] = times[9] = times[11] = times[12] = 0;
  if (npairs > 0 && (opt.continuous || opt.discrete)) {
    (void)hipEventElapsedTime(&ms, ev[10], ev[11]);
    times[8] = ms;
    if (opt.continuous && n_vf + n_ee > 0) {
      (void)hipEventElapsedTime(&ms, ev[12], ev[13]);
      times[9] = ms;
      (void)hipEventElapsedTime(&ms, ev[12], ev[16]);
      times[11] = ms;
      (void)hipEventElapsedTime(&ms, ev[16], ev[17]);
      times[12] = ms;
    }
  }
  hits_gen = tree_gen;
  return MCLG_OK;
}

Ctx::~Ctx() {
  if (stream) (void)hipStreamSynchronize(stream);
  void* bufs[] = {d_rec, d_flags, d_owner, d_ekeys, d_evals, d_boxes, d_act, d_cbounds, d_partial, d_keys, d_order, d_ktmp,
                  d_vtmp, d_hist, d_sboxes, d_face, d_v1, d_cp, d_root, d_ex_off, d_ex_part, d_hit_off, d_hit_part, d_nodes, d_lchild, d_rchild, d_rngl, d_rngr, d_visits, d_frontier, d_nfrontier, d_counters, d_pairs,
                  d_hit_sten, d_hit_vf, d_hit_toi, d_disc, d_odep, d_vfq, d_eeq, d_vf_toi, d_ee_toi, d_ctcd_sel, d_rank, d_minmax, d_adj_w, d_adj_g, d_adj_q, d_adj_n, d_adj_keys, d_adj_vals, d_adj_ktmp, d_adj_vtmp, d_adj_hist,
                  d_stage_v0, d_stage_v1, d_stage_p, d_stage_act};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  for (int i = 0; i < kNumEvents; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (h_counters) (void)hipHostFree(h_counters);
  if (h_minmax) (void)hipHostFree(h_minmax);
  if (own_stream) (void)hipStreamDestroy(own_stream);
}

// Host inputs are staged into context-owned device buffers.
int Ctx::stage(const double* V0, const double* V1, int nv_, const int* P, long long nidx, const int* active) {
  CHECK(grow((void**)&d_stage_v0, &cap_stage_v0, sizeof(double) * dim * (size_t)nv_, "staging"));
  CHECK(grow((void**)&d_stage_v1, &cap_stage_v1, sizeof(double) * dim * (size_t)nv_, "staging"));
  CHECK(grow((void**)&d_stage_p, &cap_stage_p, sizeof(int) * (size_t)std::max(nidx, 1LL), "staging"));
  HIPCHECK(this, hipMemcpyAsync(d_stage_v0, V0, sizeof(double) * dim * (size_t)nv_, hipMemcpyHostToDevice, stream));
  HIPCHECK(this, hipMemcpyAsync(d_stage_v1, V1, sizeof(double) * dim * (size_t)nv_, hipMemcpyHostToDevice, stream));
  if (nidx > 0) HIPCHECK(this, hipMemcpyAsync(d_stage_p, P, sizeof(int) * (size_t)nidx, hipMemcpyHostToDevice, stream));
  if (active) {
    CHECK(grow((void**)&d_stage_act, &cap_stage_act, sizeof(int) * (size_t)nv_, "staging"));
    HIPCHECK(this, hipMemcpyAsync(d_stage_act, active, sizeof(int) * (size_t)nv_, hipMemcpyHostToDevice, stream));
  }
  return MCLG_OK;
}

}  // namespace mclg

extern "C" {

int mclg_update_device(mclg_ctx* h, const double* V0, const double* V1, int nv, const int* P, int np, int pdim,
                       const int* active) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  return c->update(V0, V1, nv, P, np, pdim, active);
}

int mclg_update(mclg_ctx* h, const double* V0, const double* V1, int nv, const int* P, int np, int pdim,
                const int* active) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  if (np < 0 || pdim < 1 || nv < 0) return c->fail(MCLG_ERR_ARG, "update: bad sizes");
  if (np == 0) return c->update(nullptr, nullptr, nv, nullptr, 0, pdim, nullptr);
  if (!V0 || !V1 || !P || nv == 0) return c->fail(MCLG_ERR_ARG, "update: null input");
  CHECK(c->stage(V0, V1, nv, P, (long long)np * pdim, active));
  return c->update(c->d_stage_v0, c->d_stage_v1, nv, c->d_stage_p, np, pdim, active ? c->d_stage_act : nullptr);
}

int mclg_update_refit(mclg_ctx* h, const double* V0, const double* V1, int nv, const int* P, int np, int pdim,
                      const int* active) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  if (np <= 0 || pdim < 1 || nv <= 0 || !V0 || !V1 || !P) return c->fail(MCLG_ERR_ARG, "refit: bad input");
  CHECK(c->stage(V0, V1, nv, P, (long long)np * pdim, active));
  return c->update(c->d_stage_v0, c->d_stage_v1, nv, c->d_stage_p, np, pdim, active ? c->d_stage_act : nullptr, true);
}

int mclg_update_refit_device(mclg_ctx* h, const double* V0, const double* V1, int nv, const int* P, int np, int pdim,
                             const int* active) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  return c->update(V0, V1, nv, P, np, pdim, active, true);
}

int mclg_traverse_device(mclg_ctx* h, const double* V0, const double* V1, const int* P, unsigned flags) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  return c->traverse(V0, V1, P, flags);
}

int mclg_traverse(mclg_ctx* h, const double* V0, const double* V1, const int* P, unsigned flags) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  if (!c->have_tree || c->np <= 1) return c->traverse(V0, V1, P, flags);
  if (c->pdim != c->dim) return c->fail(MCLG_ERR_ARG, "traverse: primitives must have pdim == dim");
  if (!V0 || !V1 || !P) return c->fail(MCLG_ERR_ARG, "traverse: null input");
  CHECK(c->stage(V0, V1, c->nv, P, (long long)c->np * c->dim, nullptr));
  return c->traverse(c->d_stage_v0, c->d_stage_v1, c->d_stage_p, flags);
}

int mclg_get_counts(const mclg_ctx* h, mclg_counts* out) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c || !out) return MCLG_ERR_ARG;
  *out = c->counts;
  return MCLG_OK;
}

static int copy_out(Ctx* c, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return MCLG_OK;
  HIPCHECK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return MCLG_OK;
}

int mclg_copy_hits(mclg_ctx* h, int* sten4, unsigned char* is_vf, double* toi, int64_t cap) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  const int64_t n = c->counts.hits;
  if (cap < n) return c->fail(MCLG_ERR_CAPACITY, "copy_hits: cap %lld < %lld", (long long)cap, (long long)n);
  if (sten4) CHECK(copy_out(c, sten4, c->d_hit_sten, sizeof(int4) * (size_t)n));
  if (is_vf) CHECK(copy_out(c, is_vf, c->d_hit_vf, (size_t)n));
  if (toi) CHECK(copy_out(c, toi, c->d_hit_toi, sizeof(double) * (size_t)n));
  return MCLG_OK;
}

int mclg_copy_discrete(mclg_ctx* h, int* pairs2, int64_t cap) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  const int64_t n = c->counts.discrete;
  if (cap < n) return c->fail(MCLG_ERR_CAPACITY, "copy_discrete: cap < count");
  if (pairs2) CHECK(copy_out(c, pairs2, c->d_disc, sizeof(int2) * (size_t)n));
  return MCLG_OK;
}

int mclg_copy_order_dependent(mclg_ctx* h, int* pairs2, int64_t cap) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  if (!c->check_order) return c->fail(MCLG_ERR_STATE, "copy_order_dependent: the traverse had no MCLG_CHECK_ORDER");
  const int64_t n = c->counts.order_dependent;
  if (cap < n) return c->fail(MCLG_ERR_CAPACITY, "copy_order_dependent: cap < count");
  if (pairs2) CHECK(copy_out(c, pairs2, c->d_odep, sizeof(int2) * (size_t)n));
  return MCLG_OK;
}

int mclg_copy_candidates(mclg_ctx* h, int* sten4, unsigned char* is_vf, double* toi, int64_t cap) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  const int64_t n = c->counts.candidates;
  if (cap < n) return c->fail(MCLG_ERR_CAPACITY, "copy_candidates: cap < count");
  if (toi && !c->keep_cand && n > 0)
    return c->fail(MCLG_ERR_STATE, "copy_candidates: traverse was not asked to keep candidate tois");
  if (sten4) {
    CHECK(copy_out(c, sten4, c->d_vfq, sizeof(int4) * (size_t)c->n_vf));
    CHECK(copy_out(c, sten4 + 4 * c->n_vf, c->d_eeq, sizeof(int4) * (size_t)c->n_ee));
  }
  if (is_vf) {
    std::memset(is_vf, 1, (size_t)c->n_vf);
    std::memset(is_vf + c->n_vf, 0, (size_t)c->n_ee);
  }
  if (toi) {
    CHECK(copy_out(c, toi, c->d_vf_toi, sizeof(double) * (size_t)c->n_vf));
    CHECK(copy_out(c, toi + c->n_vf, c->d_ee_toi, sizeof(double) * (size_t)c->n_ee));
  }
  return MCLG_OK;
}

int mclg_copy_leaf_pairs(mclg_ctx* h, int* pairs2, int64_t cap) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  const int64_t n = c->counts.leaf_pairs;
  if (cap < n) return c->fail(MCLG_ERR_CAPACITY, "copy_leaf_pairs: cap < count");
  if (n > 0 && !c->pairs_are_prims)
    return c->fail(MCLG_ERR_STATE, "copy_leaf_pairs: the last traverse did not keep them (MCLG_KEEP_LEAF_PAIRS)");
  if (pairs2) CHECK(copy_out(c, pairs2, c->d_pairs, sizeof(int2) * (size_t)n));
  return MCLG_OK;
}

int mclg_copy_leaves(mclg_ctx* h, double* boxes, unsigned char* active, unsigned char* flags, int64_t cap) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  const int64_t n = c->have_tree ? c->np : 0;
  if (cap < n) return c->fail(MCLG_ERR_CAPACITY, "copy_leaves: cap < np");
  if (boxes) CHECK(copy_out(c, boxes, c->d_boxes, sizeof(double) * 6 * c->dim * (size_t)n));
  if (active) CHECK(copy_out(c, active, c->d_act, (size_t)n));
  if (flags) CHECK(copy_out(c, flags, c->d_flags, (size_t)n));
  return MCLG_OK;
}

int mclg_copy_tree(mclg_ctx* h, int* lchild, int* rchild, int* order, int* root, int64_t cap) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  const int64_t n = c->have_tree ? c->np : 0;
  if (cap < n) return c->fail(MCLG_ERR_CAPACITY, "copy_tree: cap < np");
  if (root) {
    if (n <= 1) *root = n == 1 ? 0 : -1;
    else CHECK(copy_out(c, root, c->d_root, sizeof(int)));
  }
  if (n > 1 && lchild) CHECK(copy_out(c, lchild, c->d_lchild, sizeof(int) * (size_t)(n - 1)));
  if (n > 1 && rchild) CHECK(copy_out(c, rchild, c->d_rchild, sizeof(int) * (size_t)(n - 1)));
  if (n > 0 && order) CHECK(copy_out(c, order, c->d_order, sizeof(int) * (size_t)n));
  return MCLG_OK;
}

int mclg_get_device_results(const mclg_ctx* h, mclg_device_results* out) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c || !out) return MCLG_ERR_ARG;
  out->hit_sten4 = (const int*)c->d_hit_sten;
  out->hit_is_vf = c->d_hit_vf;
  out->hit_toi = c->d_hit_toi;
  out->discrete2 = (const int*)c->d_disc;
  out->vf_sten4 = (const int*)c->d_vfq;
  out->vf_toi = c->keep_cand ? c->d_vf_toi : nullptr;
  out->n_vf = c->n_vf;
  out->ee_sten4 = (const int*)c->d_eeq;
  out->ee_toi = c->keep_cand ? c->d_ee_toi : nullptr;
  out->n_ee = c->n_ee;
  out->leaf_pairs2 = c->pairs_are_prims ? (const int*)c->d_pairs : nullptr;
  return MCLG_OK;
}

int mclg_set_rank(mclg_ctx* h, const int* rank, int np) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  c->kd_valid = false;  // an explicit rank replaces the library's KdBVH order until the next update
  if (!rank) { c->rank_set = false; return MCLG_OK; }
  if (np != c->np || np <= 0) return c->fail(MCLG_ERR_ARG, "set_rank: np %d != tree size %d", np, c->np);
  CHECK(c->grow((void**)&c->d_rank, &c->cap_rank, sizeof(int) * (size_t)np, "rank"));
  HIPCHECK(c, hipMemcpyAsync(c->d_rank, rank, sizeof(int) * (size_t)np, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->rank_set = true;
  return MCLG_OK;
}

int mclg_get_timings(const mclg_ctx* h, double* ms, int n) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c || !ms) return MCLG_ERR_ARG;
  for (int i = 0; i < n && i < MCLG_NUM_TIMINGS; ++i) ms[i] = c->times[i];
  return MCLG_OK;
}

}  // extern "C"
