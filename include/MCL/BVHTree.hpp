// Drop-in replacement of mcl::BVHTree<T,DIM> (mattoverby/mclccd,
// include/MCL/BVHTree.hpp:30-148) running on an MI355X through the C ABI of
// mclg.h.  Header-only; link against libmclg.so.
//
// Same class name, the same Options fields and defaults, the same four
// std::function callbacks and the same update/traverse overloads.  Differences
// a user can see:
//   - Callbacks are replayed on the calling thread, serially, after the GPU has
//     produced every result (the reference's threaded run calls them
//     concurrently from TBB workers).  Order: leaf pair by leaf pair, as collide
//     does — filter_pair on the pair's candidates, the toi (narrow_phase if set,
//     else the GPU's ACCD) and append_pair when toi >= 0, then append_discrete,
//     whose true return ends the replay.  With options.threaded false the pairs
//     come in the reference's serial order and are oriented by its KdBVH, so the
//     callback sequence is the reference's own (mclg_copy_events).
//   - The tree is the GPU's LBVH: the `tree` member (an Eigen KdBVH) does not
//     exist.  `leaves` is filled on demand by sync_leaves().  traverse(BVHTraverse*)
//     walks a host copy of the LBVH with the KdBVH walk's rules, so a visitor
//     sees other internal volumes, in another order, than on the reference.
//   - vf_one_sided runs the CTCD narrow phase with its wrong-side test on the
//     GPU for vertex-face stencils (EE stays ACCD), as BVHTree.cpp:446-447 does.
//     GPU failures throw std::runtime_error.
//   - One thread at a time per tree: update and traverse write the tree's
//     device buffers (separate trees may run on separate threads).
//   - T = double and T = float (BVHTree.cpp:504-507).  A float tree runs the
//     library in MCLG_PRECISION_F32: its float vertices are widened exactly for
//     the C ABI, and the GPU pads boxes in float, orders the KdBVH by float
//     centres and runs ACCD in float arithmetic as NarrowPhaseACCD<float,DIM>
//     does; CTCD and the discrete tests widen to double as the reference's
//     float specialisations do.  TOIs are float values.
// The matrix overloads of update/traverse take Eigen matrices as in the
// reference, or any matrix type with rows(), cols() and operator()(i, j); when
// Eigen is available the callbacks take Eigen::Vector4i, as in the reference.
#ifndef MCL_BVHTREE_HPP
#define MCL_BVHTREE_HPP 1

#include <algorithm>
#include <cstdint>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../mclg.h"

#if defined(__has_include)
#if __has_include(<Eigen/Core>)
#include <Eigen/Core>
#include <Eigen/Geometry>
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

#ifdef MCLG_HAVE_EIGEN
template <typename T, int DIM> using AlignedBox = Eigen::AlignedBox<T, DIM>;
#else
// The part of Eigen::AlignedBox the leaves and visitors use: an empty box
// (min = +max, max = lowest) grown by extend with Eigen's min/max forms.
template <typename T, int DIM> class AlignedBox {
 public:
  AlignedBox() { setEmpty(); }
  void setEmpty() {
    for (int k = 0; k < DIM; ++k) { lo_[k] = std::numeric_limits<T>::max(); hi_[k] = std::numeric_limits<T>::lowest(); }
  }
  const T* min() const { return lo_; }
  const T* max() const { return hi_; }
  T* min() { return lo_; }
  T* max() { return hi_; }
  bool isEmpty() const {
    for (int k = 0; k < DIM; ++k)
      if (lo_[k] > hi_[k]) return true;
    return false;
  }
  bool contains(const T* p) const {
    for (int k = 0; k < DIM; ++k)
      if (!(lo_[k] <= p[k] && p[k] <= hi_[k])) return false;
    return true;
  }
  AlignedBox& extend(const T* p) {
    for (int k = 0; k < DIM; ++k) { lo_[k] = p[k] < lo_[k] ? p[k] : lo_[k]; hi_[k] = hi_[k] < p[k] ? p[k] : hi_[k]; }
    return *this;
  }
  AlignedBox& extend(const AlignedBox& b) {
    for (int k = 0; k < DIM; ++k) {
      lo_[k] = b.lo_[k] < lo_[k] ? b.lo_[k] : lo_[k];
      hi_[k] = hi_[k] < b.hi_[k] ? b.hi_[k] : hi_[k];
    }
    return *this;
  }

 private:
  T lo_[DIM], hi_[DIM];
};
#endif

// BVHLeaf.hpp:12-29: the swept box with its t0 and t1 boxes and the active flag.
template <typename T, int DIM> class BoundingBox : public AlignedBox<T, DIM> {
 public:
  bool active = true;
  AlignedBox<T, DIM> t0, t1;
  // BVHLeaf.cpp:9-23
  BoundingBox merged(const BoundingBox& b) const {
    BoundingBox m;
    m.active = this->active || b.active;
    m.extend(this->min());
    m.extend(this->max());
    m.extend(b.min());
    m.extend(b.max());
    m.t0.extend(this->t0);
    m.t0.extend(b.t0);
    m.t1.extend(this->t1);
    m.t1.extend(b.t1);
    return m;
  }
};

// BVHLeaf.hpp:31-52: primitive index, representative-triangle flags, boxes.
template <typename T, int DIM> struct BVHLeaf {
  typedef BoundingBox<T, DIM> BoxType;
  int idx = -1;
  int v[3] = {1, 1, 1}, e[3] = {1, 1, 1};
  BoxType box;
};

// BVHTree.hpp:17-27: a user traversal (point-in-element queries and the like).
template <typename T, int DIM> class BVHTraverse {
 public:
  typedef BVHLeaf<T, DIM> ObjectType;
  typedef typename BVHLeaf<T, DIM>::BoxType VolumeType;
  virtual ~BVHTraverse() {}
  virtual bool intersectVolume(const VolumeType&) = 0;  // true: go down this branch
  virtual bool intersectObject(const ObjectType&) = 0;  // true: stop the traversal
};

template <typename T, int DIM> class BVHTree;

template <typename T, int DIM> class BVHTree {
  static_assert(DIM == 2 || DIM == 3, "DIM must be 2 or 3");
  static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value, "T must be double or float");

 public:
  typedef BVHLeaf<T, DIM> LeafType;
  // BVHTree.hpp:36-39 (VecType needs Eigen, as in the reference)
#ifdef MCLG_HAVE_EIGEN
  typedef Eigen::Matrix<T, DIM, 1> VecType;
#endif
  typedef std::pair<Vector4i, bool> PairType;  // [sten, is_vf]
  typedef std::pair<int, bool> NodeIndex;      // [index, isleaf]

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
    if (std::is_same<T, float>::value) check(mclg_set_precision(c, MCLG_PRECISION_F32), c);
  }
  virtual ~BVHTree() {}

  // BVHTree.hpp:81-82.  nv is implied by P (1 + the largest index).
  void update(const T* V0, const T* V1, const int* P, int np, int pdim, const int* active = nullptr) {
    push_options();
    int nv = 0;
    for (long long i = 0; i < (long long)np * pdim; ++i) nv = std::max(nv, P[i] + 1);
    check(mclg_update(ctx_.get(), widen(V0, nv, w0_), widen(V1, nv, w1_), nv, P, np, pdim, active));
    np_ = np;
    nv_ = nv;
    mirror_ = false;
    pull_options();
  }

  // Refit-only update (no reference counterpart): same arguments as update for
  // the same mesh; keeps the last update's hierarchy and recomputes the boxes
  // (mclg_update_refit).  Results equal a rebuild's.
  void refit(const T* V0, const T* V1, const int* P, int np, int pdim, const int* active = nullptr) {
    push_options();
    check(mclg_update_refit(ctx_.get(), widen(V0, nv_, w0_), widen(V1, nv_, w1_), nv_, P, np, pdim, active));
    mirror_ = false;
    pull_options();
  }

  // BVHTree.hpp:86.  Callbacks are replayed leaf pair by leaf pair in the
  // order of mclg_copy_events (the reference's serial order when
  // options.threaded is false), as collide does for each pair
  // (BVHTree.cpp:252-283): filter_pair on each of its candidates, then the toi
  // (narrow_phase, or the GPU's ACCD) and append_pair of those kept, then
  // append_discrete, whose true return stops the replay.
  void traverse(const T* V0, const T* V1, const int* P) const {
    const bool need_cand = options.continuous && ((bool)filter_pair || (bool)narrow_phase);
    const bool cont = options.continuous && ((bool)append_pair || need_cand);
    const bool disc = options.discrete && (bool)append_discrete;
    const_cast<BVHTree*>(this)->push_options();
    if (np_ <= 1) return;  // BVHTree.cpp:131-132
    BVHTree* self = const_cast<BVHTree*>(this);
    const unsigned flags = (need_cand ? MCLG_KEEP_CANDIDATES : 0u) | (cont || disc ? MCLG_KEEP_LEAF_PAIRS : 0u);
    check(mclg_traverse(ctx_.get(), widen(V0, nv_, self->w0_), widen(V1, nv_, self->w1_), P, flags));
    if (!cont && !disc) return;
    mclg_counts c;
    check(mclg_get_counts(ctx_.get(), &c));
    int64_t ne = 0;
    check(mclg_copy_events(ctx_.get(), need_cand ? 1 : 0, nullptr, nullptr, 0, &ne));
    std::vector<int64_t> ev((size_t)ne), grp((size_t)ne);
    check(mclg_copy_events(ctx_.get(), need_cand ? 1 : 0, ev.data(), grp.data(), ne, &ne));
    const int64_t n = cont ? (need_cand ? c.candidates : c.hits) : 0;
    std::vector<int> sten((size_t)n * 4);
    std::vector<unsigned char> vf((size_t)n);
    std::vector<double> toi((size_t)n);
    if (cont && need_cand) check(mclg_copy_candidates(ctx_.get(), sten.data(), vf.data(), toi.data(), n));
    else if (cont) check(mclg_copy_hits(ctx_.get(), sten.data(), vf.data(), toi.data(), n));
    std::vector<int> d(disc ? (size_t)c.discrete * 2 : 0);
    if (disc && c.discrete > 0) check(mclg_copy_discrete(ctx_.get(), d.data(), c.discrete));
    std::vector<int64_t> kept;
    for (int64_t k = 0; k < ne;) {  // one leaf pair: its records, then its discrete hit
      int64_t j = k;
      while (j < ne && grp[(size_t)j] == grp[(size_t)k]) ++j;
      if (cont) {
        kept.clear();
        for (int64_t q = k; q < j; ++q) {  // get_candidates filters every stencil of the pair first
          const int64_t i = ev[(size_t)q];
          if (i < 0) continue;
          if (filter_pair && filter_pair(stencil(sten, i), vf[(size_t)i] != 0)) continue;
          kept.push_back(i);
        }
        for (int64_t i : kept) {
          const Vector4i s = stencil(sten, i);
          const bool is_vf = vf[(size_t)i] != 0;
          const T t = narrow_phase ? narrow_phase(s, is_vf) : (T)toi[(size_t)i];
          if (t >= 0 && append_pair) append_pair(s, is_vf, t);
        }
      }
      if (disc)
        for (int64_t q = k; q < j; ++q) {
          const int64_t e = ev[(size_t)q];
          if (e >= 0) continue;
          const int64_t di = -1 - e;
          if (append_discrete(d[2 * (size_t)di], d[2 * (size_t)di + 1])) return;  // stop (BVHTree.cpp:236-239)
        }
      k = j;
    }
  }

  // BVHTree.hpp:151-186: the matrix overloads.  Any matrix type with rows(),
  // cols() and operator()(i, j) — Eigen::MatrixBase among them — is copied to
  // row-major arrays (V as T, P as int); V0 and V1 with different row counts
  // return silently, as in the reference.  active: any vector of int with
  // size() and data() (Eigen::VectorXi, std::vector<int>), used only when its
  // size is the vertex count.  V0 and V1 need dim columns (else a silent return
  // too: the reference's fixed-width copy has no other outcome).
  template <typename MV, typename MP, typename = decltype(std::declval<const MV&>()(0, 0) + std::declval<const MV&>().cols())>
  void update(const MV& V0, const MV& V1, const MP& P) {
    update_rows(V0, V1, P, nullptr, 0);
  }
  template <typename MV, typename MP, typename VA,
            typename = decltype(std::declval<const MV&>()(0, 0) + std::declval<const MV&>().cols() +
                                *std::declval<const VA&>().data() + std::declval<const VA&>().size())>
  void update(const MV& V0, const MV& V1, const MP& P, const VA& active) {
    update_rows(V0, V1, P, active.data(), (long long)active.size());
  }
  template <typename MV, typename MP, typename = decltype(std::declval<const MV&>()(0, 0) + std::declval<const MV&>().cols())>
  void traverse(const MV& V0, const MV& V1, const MP& P) const {
    if (V0.rows() != V1.rows() || V0.cols() != DIM || V1.cols() != DIM) return;
    std::vector<T> a, b;
    std::vector<int> p;
    row_major(V0, a);
    row_major(V1, b);
    row_major(P, p);
    traverse(a.data(), b.data(), p.data());
  }

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
        L.box.min()[k] = (T)x[k]; L.box.max()[k] = (T)x[DIM + k];
        L.box.t0.min()[k] = (T)x[2 * DIM + k]; L.box.t0.max()[k] = (T)x[3 * DIM + k];
        L.box.t1.min()[k] = (T)x[4 * DIM + k]; L.box.t1.max()[k] = (T)x[5 * DIM + k];
      }
      L.box.active = a[(size_t)i] != 0;
      for (int j = 0; j < 3; ++j) { L.v[j] = (f[(size_t)i] >> j) & 1; L.e[j] = (f[(size_t)i] >> (3 + j)) & 1; }
    }
  }

  // BVHTree.hpp:96, BVHTree.cpp:162-166 (Eigen::BVIntersect over the KdBVH):
  // from the root, the child volumes of each visited node are offered to
  // intersectVolume (true queues the child), then its child objects to
  // intersectObject (true ends the walk); queued nodes are taken last in,
  // first out.  Walks a host copy of the last update's tree (LBVH), whose
  // internal volumes are merges of their children's boxes as in the reference.
  void traverse(BVHTraverse<T, DIM>* trav) const {
    const_cast<BVHTree*>(this)->sync_mirror();
    if (root_ < 0 || !trav) return;
    const int ni = np_ - 1;  // internal nodes; leaf s is node ni + s
    if (np_ == 1) {
      trav->intersectObject(leaves[(size_t)order_[0]]);
      return;
    }
    std::vector<int> todo(1, root_);
    while (!todo.empty()) {
      const int p = todo.back();
      todo.pop_back();
      const int ch[2] = {lchild_[(size_t)p], rchild_[(size_t)p]};
      for (int c : ch)
        if (c < ni && trav->intersectVolume(volumes_[(size_t)c])) todo.push_back(c);
      for (int c : ch)
        if (c >= ni && trav->intersectObject(leaves[(size_t)order_[(size_t)(c - ni)]])) return;
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
  int np_ = 0, nv_ = 0;
  std::vector<double> w0_, w1_;  // FP64 copies of float vertices

  // The device path reads FP64 vertices: double input is passed through,
  // float input widened (exactly) into buf.
  static const double* widen(const double* V, int, std::vector<double>&) { return V; }
  static const double* widen(const float* V, int nv, std::vector<double>& buf) {
    buf.assign(V, V + (size_t)nv * DIM);
    return buf.data();
  }
  // host copy of the tree for visitor traversals (valid until the next update)
  bool mirror_ = false;
  int root_ = -1;
  std::vector<int> lchild_, rchild_, order_;
  std::vector<typename LeafType::BoxType> volumes_;  // internal nodes

  void sync_mirror() {
    if (mirror_) return;
    sync_leaves();
    const int n = np_, ni = std::max(n - 1, 0);
    lchild_.assign((size_t)ni, -1);
    rchild_.assign((size_t)ni, -1);
    order_.assign((size_t)n, 0);
    check(mclg_copy_tree(ctx_.get(), lchild_.data(), rchild_.data(), order_.data(), &root_, n));
    // internal volumes, children first (an explicit post-order)
    volumes_.assign((size_t)ni, typename LeafType::BoxType());
    auto box = [&](int c) -> const typename LeafType::BoxType& {
      return c < ni ? volumes_[(size_t)c] : leaves[(size_t)order_[(size_t)(c - ni)]].box;
    };
    if (n > 1) {
      std::vector<std::pair<int, bool>> st(1, std::make_pair(root_, false));
      while (!st.empty()) {
        const std::pair<int, bool> t = st.back();
        st.pop_back();
        const int p = t.first, l = lchild_[(size_t)p], r = rchild_[(size_t)p];
        if (t.second) {
          volumes_[(size_t)p] = box(l).merged(box(r));
        } else {
          st.emplace_back(p, true);
          if (l < ni) st.emplace_back(l, false);
          if (r < ni) st.emplace_back(r, false);
        }
      }
    }
    mirror_ = true;
  }

  // Row-major copy of a matrix (rows(), cols(), operator()(i, j)).
  template <typename M, typename E> static void row_major(const M& m, std::vector<E>& out) {
    const long long r = (long long)m.rows(), c = (long long)m.cols();
    out.resize((size_t)(r * c));
    for (long long i = 0; i < r; ++i)
      for (long long j = 0; j < c; ++j) out[(size_t)(i * c + j)] = static_cast<E>(m(i, j));
  }
  template <typename MV, typename MP>
  void update_rows(const MV& V0, const MV& V1, const MP& P, const int* active, long long nactive) {
    if (V0.rows() != V1.rows() || V0.cols() != DIM || V1.cols() != DIM) return;  // BVHTree.hpp:160
    std::vector<T> a, b;
    std::vector<int> p;
    row_major(V0, a);
    row_major(V1, b);
    row_major(P, p);
    const int nv = (int)V0.rows(), np = (int)P.rows(), pdim = (int)P.cols();
    push_options();
    check(mclg_update(ctx_.get(), widen(a.data(), nv, w0_), widen(b.data(), nv, w1_), nv, p.data(), np, pdim,
                      nactive == nv ? active : nullptr));
    np_ = np;
    nv_ = nv;
    mirror_ = false;
    pull_options();
  }
  static Vector4i stencil(const std::vector<int>& s, int64_t i) {
    return Vector4i(s[4 * (size_t)i], s[4 * (size_t)i + 1], s[4 * (size_t)i + 2], s[4 * (size_t)i + 3]);
  }
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
    options.vf_ccd_eta = (T)o.vf_ccd_eta;  // clamped by update (BVHTree.cpp:39-40)
    options.ee_ccd_eta = (T)o.ee_ccd_eta;
  }
};

}  // namespace mcl

#endif
