// Host-only parts of the reference's orientation and callback order (plain
// C++, no HIP): the KdBVH leaf order and the per-leaf-pair event sequence of
// callback replay.  Used by replay.hip; also compiled on its own (with the
// sanitizers) by the oracle's checks.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace mclg {
namespace host {

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

// rank[i]: in-order position of leaf i; gapnode (n-1 entries, may be null).
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
  int query(int lo, int hi) const {  // inclusive, 0 <= lo <= hi < size
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
  const int* sten = nullptr;
  const unsigned char* vf = nullptr;
  static uint64_t hash(const int* s, bool vf) {
    uint64_t h = vf ? 0x9e3779b97f4a7c15ull : 0x7f4a7c159e3779b9ull;
    for (int i = 0; i < 4; ++i) {
      h ^= (uint64_t)(uint32_t)s[i] + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
      h *= 0xff51afd7ed558ccdull;
    }
    return h ^ (h >> 29);
  }
  void build(const int* s, const unsigned char* v, size_t n) {
    sten = s;
    vf = v;
    recs.resize(n);
    for (size_t i = 0; i < n; ++i) recs[i] = {hash(s + 4 * i, v[i] != 0), (int64_t)i};
    std::sort(recs.begin(), recs.end(),
              [](const Rec& a, const Rec& b) { return a.key < b.key || (a.key == b.key && a.idx < b.idx); });
    used.assign(n, 0);
  }
  int64_t take(const int* s, bool is_vf) {
    const uint64_t k = hash(s, is_vf);
    auto it = std::lower_bound(recs.begin(), recs.end(), k, [](const Rec& r, uint64_t key) { return r.key < key; });
    for (; it != recs.end() && it->key == k; ++it) {
      const int64_t i = it->idx;
      if (used[(size_t)i] || (vf[(size_t)i] != 0) != is_vf) continue;
      if (std::memcmp(sten + 4 * (size_t)i, s, 4 * sizeof(int)) != 0) continue;
      used[(size_t)i] = 1;
      return i;
    }
    return -1;
  }
};

// The replay sequence of one traverse.  P (np x D), flags (reptri bits v0..v2,
// e0..e2) and rank per primitive; pairs: npairs oriented primitive pairs
// (lower rank first); gapnode: the KdBVH's gap nodes for the serial order, or
// null for rank order.  Records (nrec stencils + is_vf) are the hits or, with
// cand, every candidate; disc: ndisc discrete pairs.  Writes events (record
// index, or -1-d for discrete hit d) and their leaf-pair group numbers; either
// output may be null.  Returns the event count, or -1 when a candidate is
// missing from a candidate list (inconsistent inputs).
inline int64_t build_events(int D, const int* P, const unsigned char* flags, const int* rank, const int* pairs,
                            int64_t npairs, const std::vector<int>* gapnode, bool continuous, bool discrete,
                            bool cand, const int* sten, const unsigned char* vf, int64_t nrec, const int* disc,
                            int64_t ndisc, int64_t* out, int64_t* group) {
  std::vector<int64_t> ord((size_t)npairs);
  for (int64_t i = 0; i < npairs; ++i) ord[(size_t)i] = i;
  auto by_rank = [&](int64_t a, int64_t b) {
    const int a0 = rank[pairs[2 * a]], b0 = rank[pairs[2 * b]];
    if (a0 != b0) return a0 < b0;
    return rank[pairs[2 * a + 1]] < rank[pairs[2 * b + 1]];
  };
  if (gapnode && !gapnode->empty()) {
    RangeMax rm(*gapnode);
    std::vector<int> node((size_t)npairs);
    for (int64_t i = 0; i < npairs; ++i) {
      const int ra = rank[pairs[2 * i]], rb = rank[pairs[2 * i + 1]];
      node[(size_t)i] = rm.query(std::min(ra, rb), std::max(ra, rb) - 1);
    }
    std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
      if (node[(size_t)a] != node[(size_t)b]) return node[(size_t)a] < node[(size_t)b];
      return by_rank(a, b);
    });
  } else {
    std::sort(ord.begin(), ord.end(), by_rank);
  }
  StencilIndex idx;
  if (continuous) idx.build(sten, vf, (size_t)nrec);
  StencilIndex dix;  // discrete hits keyed like stencils (p0, p1, -1, -1)
  std::vector<int> dsten((size_t)ndisc * 4, -1);
  std::vector<unsigned char> dvf((size_t)ndisc, 0);
  for (int64_t i = 0; i < ndisc; ++i) {
    dsten[4 * (size_t)i] = disc[2 * i];
    dsten[4 * (size_t)i + 1] = disc[2 * i + 1];
  }
  if (discrete) dix.build(dsten.data(), dvf.data(), (size_t)ndisc);

  int64_t n = 0, g = 0;
  auto emit = [&](int64_t e) {
    if (out) out[n] = e;
    if (group) group[n] = g;
    ++n;
  };
  auto shares = [D](int v, const int* f) {
    for (int k = 0; k < D; ++k)
      if (v == f[k]) return true;
    return false;
  };
  for (int64_t q = 0; q < npairs; ++q) {
    const int64_t i = ord[(size_t)q];
    const int p0 = pairs[2 * i], p1 = pairs[2 * i + 1];
    const int* f0 = P + (size_t)p0 * D;
    const int* f1 = P + (size_t)p1 * D;
    const unsigned fl0 = flags[p0], fl1 = flags[p1];
    const int64_t n_before = n;
    if (continuous) {
      int s[4];
      for (int side = 0; side < 2; ++side) {  // VF f0 -> f1, then f1 -> f0 (BVHTree.cpp:327-369)
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
          else if (cand) return -1;
        }
      }
      if (D == 3) {  // EE (BVHTree.cpp:372-395)
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
            else if (cand) return -1;
          }
        }
      }
    }
    if (discrete && ndisc > 0) {
      const int d[4] = {p0, p1, -1, -1};
      const int64_t r = dix.take(d, false);
      if (r >= 0) emit(-1 - r);
    }
    if (n > n_before) ++g;
  }
  return n;
}

}  // namespace host
}  // namespace mclg
