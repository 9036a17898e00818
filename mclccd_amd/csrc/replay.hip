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

#define CHECK_RC(x)                  \
  do {                               \
    int _rc = (x);                   \
    if (_rc != MCLG_OK) return _rc;  \
  } while (0)
#define HIPCHECK_RC(call) CHECK_RC(hip((call), #call))

namespace mclg {

namespace {

template <int D> struct KdCentre {
  double c[D];
  int second;
};

// KdBVH::build (KdBVH.hpp:192-222): the same partitions by std::nth_element
// with the same comparator on the same initial order, so the same leaf
// arrangement; internal nodes are numbered in the order the reference pushes
// their volumes (children first).  gapnode[g]: node splitting positions g, g+1.
template <int D>
void kd_build(std::vector<KdCentre<D>>& cent, int from, int to, int dim, int& next, int* gapnode) {
  auto cmp = [dim](const KdCentre<D>& a, const KdCentre<D>& b) { return a.c[dim] < b.c[dim]; };
  if (to - from == 2) {
    gapnode[from] = next++;
    return;
  }
  if (to - from == 3) {
    const int mid = from + 2;
    std::nth_element(cent.begin() + from, cent.begin() + mid, cent.begin() + to, cmp);
    kd_build<D>(cent, from, mid, (dim + 1) % D, next, gapnode);
    gapnode[mid - 1] = next++;
    return;
  }
  const int mid = from + (to - from) / 2;
  std::nth_element(cent.begin() + from, cent.begin() + mid, cent.begin() + to, cmp);
  kd_build<D>(cent, from, mid, (dim + 1) % D, next, gapnode);
  kd_build<D>(cent, mid, to, (dim + 1) % D, next, gapnode);
  gapnode[mid - 1] = next++;
}

template <int D> void kd_order(const double* centres, int n, int* rank, int* gapnode) {
  std::vector<KdCentre<D>> cent((size_t)n);
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < D; ++k) cent[(size_t)i].c[k] = centres[(size_t)i * D + k];
    cent[(size_t)i].second = i;
  }
  std::vector<int> gn;
  int* gp = gapnode;
  if (!gp) {
    gn.assign((size_t)std::max(n - 1, 1), 0);
    gp = gn.data();
  }
  int next = 0;
  if (n >= 2) kd_build<D>(cent, 0, n, 0, next, gp);
  for (int i = 0; i < n; ++i) rank[cent[(size_t)i].second] = i;
}

// Range maximum over gapnode (sparse table): the frontlist index of the lowest
// common ancestor of the leaves at positions a < b is the largest node index
// among the gaps a .. b-1 (ancestors are pushed after their descendants).
struct RangeMax {
  std::vector<std::vector<int>> t;
  explicit RangeMax(const std::vector<int>& a) {
    t.push_back(a);
    for (size_t w = 1; 2 * w <= a.size(); w *= 2) {
      const std::vector<int>& p = t.back();
      std::vector<int> q(a.size() - 2 * w + 1);
      for (size_t i = 0; i < q.size(); ++i) q[i] = std::max(p[i], p[i + w]);
      t.push_back(std::move(q));
    }
  }
  int query(int lo, int hi) const {  // inclusive, lo <= hi
    int k = 0;
    while ((2u << k) <= (unsigned)(hi - lo + 1)) ++k;
    return std::max(t[(size_t)k][(size_t)lo], t[(size_t)k][(size_t)(hi - (1 << k) + 1)]);
  }
};

// Records (hits or candidates) looked up by stencil.  Without representative
// triangles the same stencil can come from several leaf pairs (with equal
// TOIs); each occurrence is handed out once.
struct StencilIndex {
  struct Rec {
    uint64_t key;
    int64_t idx;
  };
  std::vector<Rec> recs;
  std::vector<uint8_t> used;
  const std::vector<int>* sten = nullptr;
  const std::vector<unsigned char>* vf = nullptr;
  static uint64_t hash(const int* s, bool vf) {
    uint64_t h = vf ? 0x9e3779b97f4a7c15ull : 0x7f4a7c159e3779b9ull;
    for (int i = 0; i < 4; ++i) {
      h ^= (uint64_t)(uint32_t)s[i] + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
      h *= 0xff51afd7ed558ccdull;
    }
    return h ^ (h >> 29);
  }
  void build(const std::vector<int>& s, const std::vector<unsigned char>& v) {
    sten = &s;
    vf = &v;
    const size_t n = v.size();
    recs.resize(n);
    for (size_t i = 0; i < n; ++i) recs[i] = {hash(&s[4 * i], v[i] != 0), (int64_t)i};
    std::sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) { return a.key < b.key || (a.key == b.key && a.idx < b.idx); });
    used.assign(n, 0);
  }
  int64_t take(const int* s, bool is_vf) {
    const uint64_t k = hash(s, is_vf);
    auto it = std::lower_bound(recs.begin(), recs.end(), k, [](const Rec& r, uint64_t key) { return r.key < key; });
    for (; it != recs.end() && it->key == k; ++it) {
      const int64_t i = it->idx;
      if (used[(size_t)i] || ((*vf)[(size_t)i] != 0) != is_vf) continue;
      if (std::memcmp(&(*sten)[4 * (size_t)i], s, 4 * sizeof(int)) != 0) continue;
      used[(size_t)i] = 1;
      return i;
    }
    return -1;
  }
};

}  // namespace

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
  // AlignedBox::center(): (min + max) / 2
  std::vector<double> cen((size_t)np * dim);
  for (size_t i = 0; i < (size_t)np; ++i)
    for (int k = 0; k < dim; ++k) cen[i * dim + k] = (sw[i * 2 * dim + k] + sw[i * 2 * dim + dim + k]) / 2.0;
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
  // visiting order of the leaf pairs: the serial frontlist order of the
  // reference's KdBVH when it is known, else by rank
  std::vector<int64_t> ord((size_t)npairs);
  for (int64_t i = 0; i < npairs; ++i) ord[(size_t)i] = i;
  if (kd_valid && np > 1) {
    RangeMax rm(kd_gapnode);
    std::vector<int> node((size_t)npairs);
    for (int64_t i = 0; i < npairs; ++i) {
      const int ra = rk[(size_t)pairs[2 * i]], rb = rk[(size_t)pairs[2 * i + 1]];
      node[(size_t)i] = rm.query(std::min(ra, rb), std::max(ra, rb) - 1);
    }
    std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
      if (node[(size_t)a] != node[(size_t)b]) return node[(size_t)a] < node[(size_t)b];
      const int a0 = rk[(size_t)pairs[2 * a]], b0 = rk[(size_t)pairs[2 * b]];
      if (a0 != b0) return a0 < b0;
      return rk[(size_t)pairs[2 * a + 1]] < rk[(size_t)pairs[2 * b + 1]];
    });
  } else {
    std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
      const int a0 = rk[(size_t)pairs[2 * a]], b0 = rk[(size_t)pairs[2 * b]];
      if (a0 != b0) return a0 < b0;
      return rk[(size_t)pairs[2 * a + 1]] < rk[(size_t)pairs[2 * b + 1]];
    });
  }
  StencilIndex idx;
  if (opt.continuous) idx.build(sten, vf);
  StencilIndex dix;  // discrete hits keyed like stencils (p0, p1, -1, -1)
  std::vector<int> dsten((size_t)ndisc * 4, -1);
  std::vector<unsigned char> dvf((size_t)ndisc, 0);
  for (int64_t i = 0; i < ndisc; ++i) {
    dsten[4 * (size_t)i] = disc[2 * (size_t)i];
    dsten[4 * (size_t)i + 1] = disc[2 * (size_t)i + 1];
  }
  if (opt.discrete) dix.build(dsten, dvf);

  int64_t n = 0, g = 0;
  auto emit = [&](int64_t e) {
    if (out) out[n] = e;
    if (group) group[n] = g;
    ++n;
  };
  for (int64_t q = 0; q < npairs; ++q) {
    const int64_t i = ord[(size_t)q];
    const int p0 = pairs[2 * (size_t)i], p1 = pairs[2 * (size_t)i + 1];
    const int* f0 = &P[(size_t)p0 * D];
    const int* f1 = &P[(size_t)p1 * D];
    const unsigned fl0 = flags[(size_t)p0], fl1 = flags[(size_t)p1];
    const int64_t n_before = n;
    if (opt.continuous) {
      auto shares = [D](int v, const int* f) {
        for (int k = 0; k < D; ++k)
          if (v == f[k]) return true;
        return false;
      };
      int s[4];
      for (int side = 0; side < 2; ++side) {  // VF f0 -> f1, then f1 -> f0
        const int* a = side ? f1 : f0;
        const int* b = side ? f0 : f1;
        const unsigned fl = side ? fl1 : fl0;
        for (int k = 0; k < D; ++k) {
          if (!((fl >> k) & 1u) || shares(a[k], b)) continue;
          s[0] = a[k];
          s[1] = b[0];
          s[2] = b[1];
          s[3] = D == 3 ? b[2] : -1;
          const int64_t r = idx.take(s, true);
          if (r >= 0) emit(r);
          else if (cand) return fail(MCLG_ERR_STATE, "copy_events: candidate missing from the candidate list");
        }
      }
      if (D == 3) {
        for (int x = 0; x < 3; ++x) {
          if (!((fl0 >> (3 + x)) & 1u)) continue;
          for (int y = 0; y < 3; ++y) {
            if (!((fl1 >> (3 + y)) & 1u)) continue;
            s[0] = f0[x];
            s[1] = f0[(x + 1) % 3];
            s[2] = f1[y];
            s[3] = f1[(y + 1) % 3];
            if (s[0] == s[2] || s[0] == s[3] || s[1] == s[2] || s[1] == s[3]) continue;
            const int64_t r = idx.take(s, false);
            if (r >= 0) emit(r);
            else if (cand) return fail(MCLG_ERR_STATE, "copy_events: candidate missing from the candidate list");
          }
        }
      }
    }
    if (opt.discrete && ndisc > 0) {
      const int d[4] = {p0, p1, -1, -1};
      const int64_t r = dix.take(d, false);
      if (r >= 0) emit(-1 - r);
    }
    if (n > n_before) ++g;
  }
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

int mclg_copy_events(mclg_ctx* h, int over_candidates, int64_t* events, int64_t* group, int64_t cap,
                     int64_t* n_events) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return MCLG_ERR_ARG;
  return c->events(over_candidates != 0, events, group, cap, n_events);
}

}  // extern "C"
