// Exercises the drop-in facade include/MCL/BVHTree.hpp the way the reference's
// README uses mcl::BVHTree (README.md:19-50): update, callbacks, traverse.
// Reads a mesh written by tests/test_gpu_facade.py, writes what the callbacks
// received.  Usage: facade_check <in.bin> <out.txt>
//   in.bin: int32 nv, int32 nf, double eta, V0[nv*3], V1[nv*3], int32 F[nf*3]
#include <MCL/BVHTree.hpp>

#include <algorithm>
#include <cstdio>
#include <type_traits>
#include <utility>
#include <map>
#include <set>
#include <vector>

// Visitors for traverse(BVHTraverse*) (BVHTree.hpp:17-27, examples/traverser.cpp):
// every volume and object, points in leaf boxes, and a stop at the first object.
struct VisitAll : mcl::BVHTraverse<double, 3> {
  std::vector<int> seen;
  long volumes = 0;
  bool intersectVolume(const VolumeType&) override {
    ++volumes;
    return true;
  }
  bool intersectObject(const ObjectType& o) override {
    seen.push_back(o.idx);
    return false;
  }
};
struct PointInBox : mcl::BVHTraverse<double, 3> {
  double pt[3];
  std::set<int> in;
  bool intersectVolume(const VolumeType& v) override { return v.contains(pt); }
  bool intersectObject(const ObjectType& o) override {
    if (o.box.contains(pt)) in.insert(o.idx);
    return false;
  }
};
struct StopFirst : mcl::BVHTraverse<double, 3> {
  long objects = 0;
  bool intersectVolume(const VolumeType&) override { return true; }
  bool intersectObject(const ObjectType&) override {
    ++objects;
    return true;
  }
};

// A minimal row-major matrix with the interface the matrix overloads read.
template <typename E> struct MatT {
  long r, c;
  std::vector<E> v;
  MatT(long r_, long c_, std::vector<E> v_) : r(r_), c(c_), v(std::move(v_)) {}
  long rows() const { return r; }
  long cols() const { return c; }
  E operator()(long i, long j) const { return v[(size_t)(i * c + j)]; }
};
typedef MatT<double> Mat;
typedef MatT<int> IMat;

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int nv = 0, nf = 0;
  double eta = 0;
  if (std::fread(&nv, 4, 1, f) != 1 || std::fread(&nf, 4, 1, f) != 1 || std::fread(&eta, 8, 1, f) != 1) return 2;
  std::vector<double> V0((size_t)nv * 3), V1((size_t)nv * 3);
  std::vector<int> F((size_t)nf * 3);
  if (std::fread(V0.data(), 8, V0.size(), f) != V0.size() || std::fread(V1.data(), 8, V1.size(), f) != V1.size() ||
      std::fread(F.data(), 4, F.size(), f) != F.size())
    return 2;
  std::fclose(f);
  FILE* out = std::fopen(argv[2], "w");
  if (!out) return 2;

  mcl::BVHTree<double, 3> tree;
  // the reference's public typedefs (BVHTree.hpp:35-39)
  static_assert(std::is_same<mcl::BVHTree<double, 3>::PairType, std::pair<mcl::Vector4i, bool>>::value, "PairType");
  static_assert(std::is_same<mcl::BVHTree<double, 3>::NodeIndex, std::pair<int, bool>>::value, "NodeIndex");
  static_assert(std::is_same<mcl::BVHTree<double, 3>::LeafType, mcl::BVHLeaf<double, 3>>::value, "LeafType");
  tree.options.box_eta = eta;
  tree.options.vf_ccd_eta = eta;
  tree.options.ee_ccd_eta = eta;
  tree.update(V0.data(), V1.data(), F.data(), nf, 3);

  // 1. plain: append_pair / append_discrete collect everything
  long n_pair = 0, n_vf = 0, n_disc = 0;
  tree.append_pair = [&](const mcl::Vector4i& s, bool is_vf, const double& toi) {
    ++n_pair;
    n_vf += is_vf;
    std::fprintf(out, "hit %d %d %d %d %d %.17g\n", s[0], s[1], s[2], s[3], (int)is_vf, toi);
  };
  tree.append_discrete = [&](int p0, int p1) {
    ++n_disc;
    std::fprintf(out, "disc %d %d\n", p0, p1);
    return false;
  };
  tree.traverse(V0.data(), V1.data(), F.data());
  const mclg_counts c = tree.counts();
  std::fprintf(out, "plain %ld %ld %ld %lld %lld %lld\n", n_pair, n_vf, n_disc, (long long)c.hits,
               (long long)c.discrete, (long long)c.candidates);

  // 2. filter_pair drops every EE candidate: only VF hits reach append_pair
  long n_filter_calls = 0, n_after = 0, n_after_vf = 0;
  tree.append_pair = [&](const mcl::Vector4i&, bool is_vf, const double&) {
    ++n_after;
    n_after_vf += is_vf;
  };
  tree.append_discrete = nullptr;
  tree.filter_pair = [&](const mcl::Vector4i&, bool is_vf) {
    ++n_filter_calls;
    return !is_vf;
  };
  tree.traverse(V0.data(), V1.data(), F.data());
  std::fprintf(out, "filter %ld %ld %ld\n", n_filter_calls, n_after, n_after_vf);

  // 3. custom narrow_phase: called for every unfiltered candidate, its toi is used
  long n_np = 0, n_np_pairs = 0, bad_toi = 0;
  tree.filter_pair = nullptr;
  tree.narrow_phase = [&](const mcl::Vector4i&, bool) {
    ++n_np;
    return 0.25;
  };
  tree.append_pair = [&](const mcl::Vector4i&, bool, const double& toi) {
    ++n_np_pairs;
    bad_toi += toi != 0.25;
  };
  tree.traverse(V0.data(), V1.data(), F.data());
  std::fprintf(out, "narrow %ld %ld %ld\n", n_np, n_np_pairs, bad_toi);

  // 4. stop: append_discrete returning true ends the discrete replay
  long n_stop = 0;
  tree.narrow_phase = nullptr;
  tree.append_pair = nullptr;
  tree.append_discrete = [&](int, int) {
    ++n_stop;
    return true;
  };
  tree.traverse(V0.data(), V1.data(), F.data());
  std::fprintf(out, "stop %ld\n", n_stop);

  // 5. leaves mirror
  tree.sync_leaves();
  long owned_v = 0;
  for (const auto& L : tree.leaves) owned_v += L.v[0] + L.v[1] + L.v[2];
  std::fprintf(out, "leaves %zu %ld\n", tree.leaves.size(), owned_v);

  // 6. visitor traversal over the host copy of the tree
  VisitAll all;
  tree.traverse(&all);
  std::vector<int> count((size_t)nf, 0);
  long bad_idx = 0;
  for (int i : all.seen) {
    if (i < 0 || i >= nf) ++bad_idx;
    else ++count[(size_t)i];
  }
  long once = 0;
  for (int c : count) once += c == 1;
  std::fprintf(out, "visit_all %zu %ld %ld %ld\n", all.seen.size(), once, bad_idx, all.volumes);
  long points = 0, mismatches = 0, found = 0;
  for (int v = 0; v < nv; v += 7) {
    PointInBox q;
    for (int k = 0; k < 3; ++k) q.pt[k] = V0[(size_t)v * 3 + k];
    tree.traverse(&q);
    std::set<int> brute;
    for (const auto& L : tree.leaves)
      if (L.box.contains(q.pt)) brute.insert(L.idx);
    ++points;
    found += (long)q.in.size();
    mismatches += q.in != brute;
  }
  std::fprintf(out, "visit_points %ld %ld %ld\n", points, mismatches, found);
  StopFirst sf;
  tree.traverse(&sf);
  std::fprintf(out, "visit_stop %ld\n", sf.objects);
  // 7. the float instantiation (MCLG_PRECISION_F32): float vertices, the
  // serial order (threaded = false), every callback logged for comparison
  // with the oracle's float tree; leaf boxes checked against float arithmetic
  {
    std::vector<float> F0(V0.begin(), V0.end()), F1(V1.begin(), V1.end());
    mcl::BVHTree<float, 3> tf;
    tf.options.box_eta = tf.options.vf_ccd_eta = tf.options.ee_ccd_eta = (float)eta;
    tf.options.threaded = false;
    long nf_hits = 0;
    tf.append_pair = [&](const mcl::Vector4i& s, bool is_vf, const float& toi) {
      std::fprintf(out, "fseq h %d %d %d %d %d %.9g\n", s[0], s[1], s[2], s[3], (int)is_vf, (double)toi);
      ++nf_hits;
    };
    tf.append_discrete = [&](int p0, int p1) {
      std::fprintf(out, "fseq d %d %d\n", p0, p1);
      return false;
    };
    tf.update(F0.data(), F1.data(), F.data(), nf, 3);
    tf.traverse(F0.data(), F1.data(), F.data());
    // leaf boxes equal the float arithmetic of the reference's float tree
    // (BVHTree.cpp:83-124 with T = float): min/max of float vertices, padded in float
    tf.sync_leaves();
    long box_bad = 0;
    const float ef = (float)eta;
    for (size_t i = 0; i < tf.leaves.size(); ++i)
      for (int k = 0; k < 3; ++k) {
        float lo = 3.0e38f, hi = -3.0e38f, hi1 = -3.0e38f;
        for (int j = 0; j < 3; ++j) {
          const int v = F[i * 3 + j];
          lo = std::min(lo, std::min(F0[(size_t)v * 3 + k], F1[(size_t)v * 3 + k]));
          hi = std::max(hi, std::max(F0[(size_t)v * 3 + k], F1[(size_t)v * 3 + k]));
          hi1 = std::max(hi1, F1[(size_t)v * 3 + k]);
        }
        box_bad += tf.leaves[i].box.min()[k] != lo - ef;
        box_bad += tf.leaves[i].box.max()[k] != hi + ef;
        box_bad += tf.leaves[i].box.t1.max()[k] != hi1 + ef;
      }
    std::fprintf(out, "float %ld %ld %g\n", nf_hits, box_bad, (double)tf.options.vf_ccd_eta);
  }
  // 8. refit: the same frame through refit gives the rebuilt tree's hits
  {
    long n_a = 0, n_b = 0;
    tree.narrow_phase = nullptr;
    tree.filter_pair = nullptr;
    tree.append_discrete = nullptr;
    tree.append_pair = [&](const mcl::Vector4i&, bool, const double&) { ++n_a; };
    tree.update(V0.data(), V1.data(), F.data(), nf, 3);
    tree.traverse(V0.data(), V1.data(), F.data());
    tree.append_pair = [&](const mcl::Vector4i&, bool, const double&) { ++n_b; };
    tree.refit(V0.data(), V1.data(), F.data(), nf, 3);
    tree.traverse(V0.data(), V1.data(), F.data());
    std::fprintf(out, "refit %ld %ld\n", n_a, n_b);
  }
  // 9. the serial traversal (threaded = false): the reference's own callback
  // sequence, once in full and once with append_discrete stopping at the
  // middle discrete hit (BVHTree.cpp:236-239: nothing after the stop)
  {
    mcl::BVHTree<double, 3> ts;
    ts.options.box_eta = ts.options.vf_ccd_eta = ts.options.ee_ccd_eta = eta;
    ts.options.threaded = false;
    const char* tag = "seqfull";
    long n_d = 0, stop_at = -1;
    ts.append_pair = [&](const mcl::Vector4i& s, bool is_vf, const double& toi) {
      std::fprintf(out, "%s h %d %d %d %d %d %.17g\n", tag, s[0], s[1], s[2], s[3], (int)is_vf, toi);
    };
    ts.append_discrete = [&](int p0, int p1) {
      std::fprintf(out, "%s d %d %d\n", tag, p0, p1);
      return ++n_d == stop_at;
    };
    ts.update(V0.data(), V1.data(), F.data(), nf, 3);
    ts.traverse(V0.data(), V1.data(), F.data());
    const long total_d = n_d;
    tag = "seqstop";
    n_d = 0;
    stop_at = (total_d + 1) / 2;
    ts.traverse(V0.data(), V1.data(), F.data());
    std::fprintf(out, "serial %ld %ld\n", total_d, stop_at);
  }
  // 10. the matrix overloads (BVHTree.hpp:151-186) with a plain row-major
  // matrix type: the raw path's hits; V0 and V1 with different row counts
  // return silently (update leaves the tree as it was, traverse calls nothing)
  {
    Mat M0(nv, 3, V0), M1(nv, 3, V1);
    Mat Mshort(nv - 1, 3, std::vector<double>(V1.begin(), V1.end() - 3));
    IMat MF(nf, 3, F);
    mcl::BVHTree<double, 3> tm;
    tm.options.box_eta = tm.options.vf_ccd_eta = tm.options.ee_ccd_eta = eta;
    long n_mat = 0, n_after_bad_update = 0, n_bad_traverse = 0;
    tm.append_pair = [&](const mcl::Vector4i&, bool, const double&) { ++n_mat; };
    std::vector<int> active((size_t)nv, 1);
    tm.update(M0, M1, MF, active);
    tm.traverse(M0, M1, MF);
    const mclg_counts c1 = tm.counts();
    tm.update(M0, Mshort, MF);  // silent return: the tree stays
    tm.append_pair = [&](const mcl::Vector4i&, bool, const double&) { ++n_after_bad_update; };
    tm.traverse(V0.data(), V1.data(), F.data());
    tm.append_pair = [&](const mcl::Vector4i&, bool, const double&) { ++n_bad_traverse; };
    tm.traverse(M0, Mshort, MF);
    std::fprintf(out, "matrix %ld %ld %ld %lld\n", n_mat, n_after_bad_update, n_bad_traverse, (long long)c1.hits);
  }
  std::fclose(out);
  return 0;
}
