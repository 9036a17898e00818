// Drop-in replacement of mcl::BVHTree<double,DIM> (mattoverby/mclccd,
// include/MCL/BVHTree.hpp:30-148) running on an MI355X through the C ABI of
// mclg.h.  Header-only; link against libmclg.so.
//
// Same class name, the same Options fields and defaults, the same four
// std::function callbacks and the same update/traverse overloads.  Differences
// a user can see:
//   - Callbacks are replayed on the calling thread, serially, after the GPU has
//     produced every result (the reference calls them concurrently from TBB
//     workers).  Order: for each candidate stencil filter_pair (if set), then the
//     toi (narrow_phase if set, else the GPU's ACCD) and append_pair when
//     toi >= 0; then append_discrete for each discrete hit until one returns true.
//   - The tree is the GPU's LBVH: the `tree` member (an Eigen KdBVH) does not
//     exist.  `leaves` is filled on demand by sync_leaves().
//   - vf_one_sided (the CTCD narrow phase) has no GPU path: update/traverse
//     throw std::runtime_error for it.  GPU failures throw std::runtime_error.
//   - Only T = double is provided (the FP64 path).
// When Eigen is available the Eigen overloads of update/traverse are provided
// and the callbacks take Eigen::Vector4i, as in the reference.
#ifndef MCL_BVHTREE_HPP
#define MCL_BVHTREE_HPP 1

#include <algorithm>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../mclg.h"

#if defined(__has_include)
#if __has_include(<Eigen/Core>)
#include <Eigen/Core>
#define MCLG_HAVE_EIGEN 1
#endif
#endif

namespace mcl {

#ifdef MCLG_HAVE_EIGEN
using Vector4i = Eigen::Vector4i;
#else
// Minimal stand-in for Eigen::Vector4i (indexing and comparison only).
struct Vector4i {
  int v[4];
  Vector4i() : v{-1, -1, -1, -1} {}
  Vector4i(int a, int b, int c, int d) : v{a, b, c, d} {}
  int& operator[](int i) { return v[i]; }
  const int& operator[](int i) const { return v[i]; }
  bool operator==(const Vector4i& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2] && v[3] == o.v[3]; }
};
#endif

// BVHLeaf.hpp:12-52 without Eigen: the three boxes flattened, active, idx and
// the representative-triangle flags.
template <typename T, int DIM> struct BVHLeaf {
  int idx = -1;
  int v[3] = {1, 1, 1}, e[3] = {1, 1, 1};
  T sw_min[DIM], sw_max[DIM], t0_min[DIM], t0_max[DIM], t1_min[DIM], t1_max[DIM];
  bool active = true;
};

template <typename T, int DIM> class BVHTree;

template <int DIM> class BVHTree<double, DIM> {
  static_assert(DIM == 2 || DIM == 3, "DIM must be 2 or 3");

 public:
  typedef double T;
  typedef BVHLeaf<T, DIM> LeafType;

  // BVHTree.hpp:42-64
  struct Options {
    T box_eta;
    T vf_ccd_eta;
    T ee_ccd_eta;
    bool reptri;
    bool threaded;
    bool discrete;
    bool continuous;
    bool vf_one_sided;
    bool ee_robust;
    Options()
        : box_eta(1e-6), vf_ccd_eta(1e-6), ee_ccd_eta(1e-6), reptri(true), threaded(true), discrete(true),
          continuous(true), vf_one_sided(false), ee_robust(true) {}
  } options;

  // BVHTree.hpp:98-119
  std::function<void(const Vector4i& sten, bool is_vf, const T& toi)> append_pair;
  std::function<bool(int p0, int p1)> append_discrete;
  std::function<bool(const Vector4i& sten, bool is_vf)> filter_pair;
  std::function<T(const Vector4i& sten, bool is_vf)> narrow_phase;

  // Host mirror of the leaves, filled by sync_leaves().
  std::vector<LeafType> leaves;

  BVHTree() {
    mclg_ctx* c = nullptr;
    check(mclg_create(DIM, &c), c);
    ctx_.reset(c, [](mclg_ctx* p) { mclg_destroy(p); });
  }
  virtual ~BVHTree() {}

  // BVHTree.hpp:81-82.  nv is implied by P (1 + the largest index).
  void update(const T* V0, const T* V1, const int* P, int np, int pdim, const int* active = nullptr) {
    push_options();
    int nv = 0;
    for (long long i = 0; i < (long long)np * pdim; ++i) nv = std::max(nv, P[i] + 1);
    check(mclg_update(ctx_.get(), V0, V1, nv, P, np, pdim, active));
    np_ = np;
    pull_options();
  }

  // BVHTree.hpp:86
  void traverse(const T* V0, const T* V1, const int* P) const {
    const bool need_cand = (bool)filter_pair || (bool)narrow_phase;
    const_cast<BVHTree*>(this)->push_options();
    if (np_ <= 1) return;  // BVHTree.cpp:131-132
    check(mclg_traverse(ctx_.get(), V0, V1, P, need_cand ? MCLG_KEEP_CANDIDATES : 0u));
    mclg_counts c;
    check(mclg_get_counts(ctx_.get(), &c));
    if (options.continuous && (append_pair || need_cand)) {
      const int64_t n = need_cand ? c.candidates : c.hits;
      std::vector<int> sten((size_t)n * 4);
      std::vector<unsigned char> vf((size_t)n);
      std::vector<double> toi((size_t)n);
      if (need_cand) check(mclg_copy_candidates(ctx_.get(), sten.data(), vf.data(), toi.data(), n));
      else check(mclg_copy_hits(ctx_.get(), sten.data(), vf.data(), toi.data(), n));
      for (int64_t i = 0; i < n; ++i) {
        const Vector4i s(sten[4 * i], sten[4 * i + 1], sten[4 * i + 2], sten[4 * i + 3]);
        const bool is_vf = vf[i] != 0;
        if (filter_pair && filter_pair(s, is_vf)) continue;
        const T t = narrow_phase ? narrow_phase(s, is_vf) : toi[i];
        if (t >= 0 && append_pair) append_pair(s, is_vf, t);
      }
    }
    if (options.discrete && append_discrete && c.discrete > 0) {
      std::vector<int> d((size_t)c.discrete * 2);
      check(mclg_copy_discrete(ctx_.get(), d.data(), c.discrete));
      for (int64_t i = 0; i < c.discrete; ++i)
        if (append_discrete(d[2 * i], d[2 * i + 1])) break;
    }
  }

#ifdef MCLG_HAVE_EIGEN
  // BVHTree.hpp:151-186: row-major copies, silent return on mismatched rows
  template <typename DerivedV, typename DerivedP>
  void update(const Eigen::MatrixBase<DerivedV>& V0, const Eigen::MatrixBase<DerivedV>& V1,
              const Eigen::MatrixBase<DerivedP>& P, const Eigen::VectorXi& active = Eigen::VectorXi()) {
    if (V0.rows() != V1.rows()) return;
    const int* act = active.size() == V0.rows() ? active.data() : nullptr;
    typedef Eigen::Matrix<T, Eigen::Dynamic, DIM, Eigen::RowMajor> RV;
    typedef Eigen::Matrix<int, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor> RI;
    RV a = V0.template cast<T>(), b = V1.template cast<T>();
    RI p = P.template cast<int>();
    push_options();
    check(mclg_update(ctx_.get(), a.data(), b.data(), (int)a.rows(), p.data(), (int)p.rows(), (int)p.cols(), act));
    np_ = (int)p.rows();
    pull_options();
  }
  template <typename DerivedV, typename DerivedP>
  void traverse(const Eigen::MatrixBase<DerivedV>& V0, const Eigen::MatrixBase<DerivedV>& V1,
                const Eigen::MatrixBase<DerivedP>& P) {
    if (V0.rows() != V1.rows()) return;
    typedef Eigen::Matrix<T, Eigen::Dynamic, DIM, Eigen::RowMajor> RV;
    typedef Eigen::Matrix<int, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor> RI;
    RV a = V0.template cast<T>(), b = V1.template cast<T>();
    RI p = P.template cast<int>();
    traverse(a.data(), b.data(), p.data());
  }
#endif

  // Fills `leaves` from the device (boxes, active, reptri flags, idx).
  void sync_leaves() {
    const int64_t n = (int64_t)np_;
    std::vector<double> b((size_t)n * 6 * DIM);
    std::vector<unsigned char> a((size_t)n), f((size_t)n);
    check(mclg_copy_leaves(ctx_.get(), b.data(), a.data(), f.data(), n));
    leaves.assign((size_t)n, LeafType());
    for (int64_t i = 0; i < n; ++i) {
      LeafType& L = leaves[(size_t)i];
      L.idx = (int)i;
      const double* x = b.data() + i * 6 * DIM;
      for (int k = 0; k < DIM; ++k) {
        L.sw_min[k] = x[k]; L.sw_max[k] = x[DIM + k];
        L.t0_min[k] = x[2 * DIM + k]; L.t0_max[k] = x[3 * DIM + k];
        L.t1_min[k] = x[4 * DIM + k]; L.t1_max[k] = x[5 * DIM + k];
      }
      L.active = a[(size_t)i] != 0;
      for (int j = 0; j < 3; ++j) { L.v[j] = (f[(size_t)i] >> j) & 1; L.e[j] = (f[(size_t)i] >> (3 + j)) & 1; }
    }
  }

  // Results of the last traverse without callbacks (bulk access).
  mclg_counts counts() const {
    mclg_counts c;
    check(mclg_get_counts(ctx_.get(), &c));
    return c;
  }
  mclg_ctx* context() const { return ctx_.get(); }

 private:
  std::shared_ptr<mclg_ctx> ctx_;
  int np_ = 0;

  static void check(int rc, mclg_ctx* c) {
    if (rc != MCLG_OK)
      throw std::runtime_error(std::string("mclg: ") + (c ? mclg_last_error(c) : "no HIP device / bad argument"));
  }
  void check(int rc) const { check(rc, ctx_.get()); }

  void push_options() {
    mclg_options o;
    o.box_eta = options.box_eta;
    o.vf_ccd_eta = options.vf_ccd_eta;
    o.ee_ccd_eta = options.ee_ccd_eta;
    o.reptri = options.reptri;
    o.threaded = options.threaded;
    o.discrete = options.discrete;
    o.continuous = options.continuous;
    o.vf_one_sided = options.vf_one_sided;
    o.ee_robust = options.ee_robust;
    check(mclg_set_options(ctx_.get(), &o));
  }
  void pull_options() {
    mclg_options o;
    check(mclg_get_options(ctx_.get(), &o));
    options.vf_ccd_eta = o.vf_ccd_eta;  // clamped by update (BVHTree.cpp:39-40)
    options.ee_ccd_eta = o.ee_ccd_eta;
  }
};

}  // namespace mcl

#endif
