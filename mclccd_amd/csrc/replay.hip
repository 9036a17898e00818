// Host side of the reference's orientation and callback order.
//
// (1) The reference's KdBVH leaf order (KdBVH.hpp:103-134,192-222), computed
//     on the host from the swept-box centres: the in-order position ("rank") of
//     each leaf orients every leaf pair (collide puts the left-subtree leaf
//     first, BVHTree.cpp:288-301), which fixes the EE stencil order and the
//     argument order of the discrete test.  The post-order index of the node
//     that splits each pair of consecutive positions is the index of that
//     node's entry in the frontlist (BVHTree.cpp:168-185), which fixes the
//     order in which a serial traversal (threaded = false, BVHTree.cpp:152-157)
//     reaches the pair.
// (2) The replay order of the callbacks: leaf pairs in that order (or in rank
//     order without a KdBVH), and per pair its candidates in get_candidates
//     order (BVHTree.cpp:327-395) followed by its discrete hit
//     (BVHTree.cpp:252-283).  Events name records of the hit (or candidate)
//     and discrete lists, which the GPU wrote in no particular order.
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/mclg.h"
#include "context.hpp"
#include "hostorder.hpp"

#define CHECK_RC(x)                  \
  do {                               \
    int _rc = (x);                   \
    if (_rc != MCLG_OK) return _rc;  \
  } while (0)
#define HIPCHECK_RC(call) CHECK_RC(hip((call), #call))

namespace mclg {

using host::kd_order;

// Recomputes the reference's KdBVH leaf order of the current leaves (host).
int Ctx::kdbvh_from_leaves() {
  kd_rank.assign((size_t)np, 0);
  kd_gapnode.assign((size_t)std::max(np - 1, 0), 0);
  kd_valid = false;
  if (np <= 0) return MCLG_OK;
  // the swept box is the first 2*dim doubles of each leaf box
  const size_t pitch = dim == 3 ? sizeof(LeafBox<3>) : sizeof(LeafBox<2>);
  std::vector<double> sw((size_t)np * 2 * dim);
  HIPCHECK_RC(hipMemcpy2DAsync(sw.data(), sizeof(double) * 2 * dim, d_boxes, pitch, sizeof(double) * 2 * dim,
                               (size_t)np, hipMemcpyDeviceToHost, stream));
  HIPCHECK_RC(hipStreamSynchronize(stream));
  // AlignedBox::center(): (min + max) / 2, in T (a float tree's boxes hold floats)
  std::vector<double> cen((size_t)np * dim);
  const bool f32 = precision == MCLG_PRECISION_F32;
  for (size_t i = 0; i < (size_t)np; ++i)
    for (int k = 0; k < dim; ++k) {
      const double lo = sw[i * 2 * dim + k], hi = sw[i * 2 * dim + dim + k];
      cen[i * dim + k] = f32 ? (double)(((float)lo + (float)hi) / 2.0f) : (lo + hi) / 2.0;
    }
  if (dim == 3) kd_order<3>(cen.data(), np, kd_rank.data(), np > 1 ? kd_gapnode.data() : nullptr);
  else kd_order<2>(cen.data(), np, kd_rank.data(), np > 1 ? kd_gapnode.data() : nullptr);
  CHECK_RC(grow((void**)&d_rank, &cap_rank, sizeof(int) * (size_t)np, "rank"));
  HIPCHECK_RC(hipMemcpyAsync(d_rank, kd_rank.data(), sizeof(int) * (size_t)np, hipMemcpyHostToDevice, stream));
  HIPCHECK_RC(hipStreamSynchronize(stream));
  rank_set = true;
  kd_valid = true;
  return MCLG_OK;
}

// Callback replay order of the last traverse (see mclg_copy_events).
int Ctx::events(bool cand, int64_t* out, int64_t* group, int64_t cap, int64_t* nout) {
  const int64_t nrec = cand ? counts.candidates : counts.hits;
  const int64_t ndisc = counts.discrete;
  const int64_t total = (opt.continuous ? nrec : 0) + (opt.discrete ? ndisc : 0);
  if (nout) *nout = total;
  if (!out && !group) return MCLG_OK;
  if (cap < total) return fail(MCLG_ERR_CAPACITY, "copy_events: cap %lld < %lld", (long long)cap, (long long)total);
  if (total == 0) return MCLG_OK;
  if (hits_gen != tree_gen) return fail(MCLG_ERR_STATE, "copy_events: no traverse since the last update");
  if (cand && !keep_cand) return fail(MCLG_ERR_STATE, "copy_events: the traverse did not keep candidates");
  const int64_t npairs = counts.leaf_pairs;
  if (npairs > 0 && !pairs_are_prims)
    return fail(MCLG_ERR_STATE, "copy_events: the traverse did not keep leaf pairs (MCLG_KEEP_LEAF_PAIRS)");
  // leaf pairs (oriented, lower rank first) and the primitives' vertices and flags
  std::vector<int> pairs((size_t)npairs * 2);
  if (npairs > 0)
    HIPCHECK_RC(hipMemcpyAsync(pairs.data(), d_pairs, sizeof(int) * 2 * (size_t)npairs, hipMemcpyDeviceToHost, stream));
  std::vector<int4> head((size_t)np);
  std::vector<int> prim((size_t)np), rank((size_t)np);
  const FaceArrays F = faces();
  HIPCHECK_RC(hipMemcpyAsync(head.data(), F.head, sizeof(int4) * (size_t)np, hipMemcpyDeviceToHost, stream));
  HIPCHECK_RC(hipMemcpyAsync(prim.data(), F.prim, sizeof(int) * (size_t)np, hipMemcpyDeviceToHost, stream));
  HIPCHECK_RC(hipMemcpyAsync(rank.data(), F.rank, sizeof(int) * (size_t)np, hipMemcpyDeviceToHost, stream));
  // records and discrete hits
  std::vector<int> sten((size_t)nrec * 4);
  std::vector<unsigned char> vf((size_t)nrec);
  std::vector<int> disc((size_t)ndisc * 2);
  if (opt.continuous && nrec > 0) {
    if (cand) {
      HIPCHECK_RC(hipMemcpyAsync(sten.data(), d_vfq, sizeof(int4) * (size_t)n_vf, hipMemcpyDeviceToHost, stream));
      if (n_ee > 0)
        HIPCHECK_RC(hipMemcpyAsync(sten.data() + 4 * (size_t)n_vf, d_eeq, sizeof(int4) * (size_t)n_ee,
                                   hipMemcpyDeviceToHost, stream));
      std::memset(vf.data(), 1, (size_t)n_vf);
      std::memset(vf.data() + n_vf, 0, (size_t)n_ee);
    } else {
      HIPCHECK_RC(hipMemcpyAsync(sten.data(), d_hit_sten, sizeof(int4) * (size_t)nrec, hipMemcpyDeviceToHost, stream));
      HIPCHECK_RC(hipMemcpyAsync(vf.data(), d_hit_vf, (size_t)nrec, hipMemcpyDeviceToHost, stream));
    }
  }
  if (opt.discrete && ndisc > 0)
    HIPCHECK_RC(hipMemcpyAsync(disc.data(), d_disc, sizeof(int) * 2 * (size_t)ndisc, hipMemcpyDeviceToHost, stream));
  HIPCHECK_RC(hipStreamSynchronize(stream));

  const int D = dim;
  std::vector<int> P((size_t)np * D);
  std::vector<unsigned char> flags((size_t)np);
  std::vector<int> rk((size_t)np);
  for (int s = 0; s < np; ++s) {
    const int p = prim[(size_t)s];
    const int4 h = head[(size_t)s];
    P[(size_t)p * D] = h.x;
    P[(size_t)p * D + 1] = h.y;
    if (D == 3) P[(size_t)p * D + 2] = h.z;
    flags[(size_t)p] = (unsigned char)(D == 3 ? h.w : h.z);
    rk[(size_t)p] = rank[(size_t)s];
  }
  const std::vector<int>* gap = kd_valid && np > 1 ? &kd_gapnode : nullptr;
  const int64_t n = host::build_events(D, P.data(), flags.data(), rk.data(), pairs.data(), npairs, gap,
                                       opt.continuous != 0, opt.discrete != 0, cand, sten.data(), vf.data(), nrec,
                                       disc.data(), ndisc, out, group);
  if (n < 0) return fail(MCLG_ERR_STATE, "copy_events: candidate missing from the candidate list");
  if (n != total) return fail(MCLG_ERR_STATE, "copy_events: %lld of %lld records placed", (long long)n, (long long)total);
  return MCLG_OK;
}

}  // namespace mclg

using namespace mclg;

extern "C" {

int mclg_kdbvh_order(int dim, const double* centres, int n, int* rank, int* gapnode) {
  if ((dim != 2 && dim != 3) || n < 0 || (n > 0 && (!centres || !rank))) return MCLG_ERR_ARG;
  if (dim == 3) kd_order<3>(centres, n, rank, n > 1 ? gapnode : nullptr);
  else kd_order<2>(centres, n, rank, n > 1 ? gapnode : nullptr);
  return MCLG_OK;
}

int mclg_set_orientation(mclg_ctx* h, int mode) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  if (mode != MCLG_ORIENT_INDEX && mode != MCLG_ORIENT_KDBVH)
    return c->fail(MCLG_ERR_ARG, "set_orientation: unknown mode %d", mode);
  c->orient = mode;
  if (mode == MCLG_ORIENT_INDEX && c->kd_valid) {  // the library's KdBVH ranks no longer apply
    c->rank_set = false;
    c->kd_valid = false;
  }
  return MCLG_OK;
}

int mclg_get_orientation(const mclg_ctx* h, int* mode) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c || !mode) return MCLG_ERR_ARG;
  *mode = c->orient;
  return MCLG_OK;
}

int mclg_set_precision(mclg_ctx* h, int precision) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  if (precision != MCLG_PRECISION_F64 && precision != MCLG_PRECISION_F32)
    return c->fail(MCLG_ERR_ARG, "set_precision: unknown precision %d", precision);
  // T is fixed for the life of a tree (a template parameter in the reference):
  // the precision is chosen before the first update
  if (precision != c->precision && c->tree_gen != 0)
    return c->fail(MCLG_ERR_STATE, "set_precision: the context has already been updated; choose the precision first");
  c->precision = precision;
  return MCLG_OK;
}

int mclg_get_precision(const mclg_ctx* h, int* precision) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c || !precision) return MCLG_ERR_ARG;
  *precision = c->precision;
  return MCLG_OK;
}

int mclg_copy_events(mclg_ctx* h, int over_candidates, int64_t* events, int64_t* group, int64_t cap,
                     int64_t* n_events) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  return c->events(over_candidates != 0, events, group, cap, n_events);
}

}  // extern "C"
