// C entry points over the CPU oracle (mclccd_oracle.hpp) for ctypes.
// TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the
// CPU-baseline leg of bench.py, never by the product library.
#include "dual.hpp"
#include "mclccd_oracle.hpp"

#include <chrono>
#include <memory>

namespace {

struct OracleCtx {
  int dim;
  std::unique_ptr<mclo::Tree<2>> t2;
  std::unique_ptr<mclo::Tree<3>> t3;
  double last_update_ms = 0, last_traverse_ms = 0;
};

template <int D> mclo::Tree<D>& tree(OracleCtx* c);
template <> mclo::Tree<2>& tree<2>(OracleCtx* c) { return *c->t2; }
template <> mclo::Tree<3>& tree<3>(OracleCtx* c) { return *c->t3; }

using clk = std::chrono::steady_clock;
double ms_since(clk::time_point t0) { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }

}  // namespace

struct oracle_options {
  double box_eta, vf_ccd_eta, ee_ccd_eta;
  int reptri, threaded, discrete, continuous;
  int vf_one_sided;
  int f32;  // BVHTree<float,DIM> arithmetic
};

template <int D> static void set_opts(OracleCtx* c, const oracle_options* o, int nthreads) {
  auto& t = tree<D>(c);
  t.options.box_eta = o->box_eta;
  t.options.vf_ccd_eta = o->vf_ccd_eta;
  t.options.ee_ccd_eta = o->ee_ccd_eta;
  t.options.reptri = o->reptri != 0;
  t.options.threaded = o->threaded != 0;
  t.options.discrete = o->discrete != 0;
  t.options.continuous = o->continuous != 0;
  t.options.vf_one_sided = o->vf_one_sided != 0;
  t.options.f32 = o->f32 != 0;
  t.nthreads = nthreads;
}

extern "C" {

void* oracle_create(int dim) {
  if (dim != 2 && dim != 3) return nullptr;
  auto* c = new OracleCtx();
  c->dim = dim;
  if (dim == 2) c->t2.reset(new mclo::Tree<2>());
  else c->t3.reset(new mclo::Tree<3>());
  return c;
}

void oracle_destroy(void* h) { delete static_cast<OracleCtx*>(h); }

int oracle_set_options(void* h, const oracle_options* o, int nthreads) {
  auto* c = static_cast<OracleCtx*>(h);
  if (c->dim == 2) set_opts<2>(c, o, nthreads);
  else set_opts<3>(c, o, nthreads);
  return 0;
}

int oracle_get_etas(void* h, double* out3) {
  auto* c = static_cast<OracleCtx*>(h);
  const mclo::Options& o = c->dim == 2 ? c->t2->options : c->t3->options;
  out3[0] = o.box_eta; out3[1] = o.vf_ccd_eta; out3[2] = o.ee_ccd_eta;
  return 0;
}

int oracle_update(void* h, const double* V0, const double* V1, const int* P, int np, int pdim, const int* active) {
  auto* c = static_cast<OracleCtx*>(h);
  auto t0 = clk::now();
  if (c->dim == 2) c->t2->update(V0, V1, P, np, pdim, active);
  else c->t3->update(V0, V1, P, np, pdim, active);
  c->last_update_ms = ms_since(t0);
  return 0;
}

int oracle_traverse(void* h, const double* V0, const double* V1, const int* P, int want_leaf_pairs,
                    int want_candidates) {
  auto* c = static_cast<OracleCtx*>(h);
  auto t0 = clk::now();
  if (c->dim == 2) {
    c->t2->want_leaf_pairs = want_leaf_pairs; c->t2->want_candidates = want_candidates;
    c->t2->traverse(V0, V1, P);
  } else {
    c->t3->want_leaf_pairs = want_leaf_pairs; c->t3->want_candidates = want_candidates;
    c->t3->traverse(V0, V1, P);
  }
  c->last_traverse_ms = ms_since(t0);
  return 0;
}

double oracle_last_ms(void* h, int which) {
  auto* c = static_cast<OracleCtx*>(h);
  return which == 0 ? c->last_update_ms : c->last_traverse_ms;
}

// which: 0 leaf pairs, 1 candidates, 2 hits, 3 discrete, 4 distance evals, 5 loop iterations
long long oracle_count(void* h, int which) {
  auto* c = static_cast<OracleCtx*>(h);
  auto get = [&](auto& t) -> long long {
    switch (which) {
      case 0: return (long long)t.leaf_pairs.size();
      case 1: return (long long)t.candidates.size();
      case 2: return (long long)t.hits.size();
      case 3: return (long long)t.discrete.size();
      case 4: return t.stats.dist_evals;
      case 5: return t.stats.iters;
    }
    return -1;
  };
  return c->dim == 2 ? get(*c->t2) : get(*c->t3);
}

// pairs: int[2n]
int oracle_copy_leaf_pairs(void* h, int* out) {
  auto* c = static_cast<OracleCtx*>(h);
  auto cp = [&](auto& t) { for (size_t i = 0; i < t.leaf_pairs.size(); ++i) { out[2 * i] = t.leaf_pairs[i].first; out[2 * i + 1] = t.leaf_pairs[i].second; } };
  if (c->dim == 2) cp(*c->t2); else cp(*c->t3);
  return 0;
}

int oracle_copy_discrete(void* h, int* out) {
  auto* c = static_cast<OracleCtx*>(h);
  auto cp = [&](auto& t) { for (size_t i = 0; i < t.discrete.size(); ++i) { out[2 * i] = t.discrete[i].first; out[2 * i + 1] = t.discrete[i].second; } };
  if (c->dim == 2) cp(*c->t2); else cp(*c->t3);
  return 0;
}

// sten: int[4n], is_vf: uint8[n]
int oracle_copy_candidates(void* h, int* sten, unsigned char* is_vf) {
  auto* c = static_cast<OracleCtx*>(h);
  auto cp = [&](auto& t) {
    for (size_t i = 0; i < t.candidates.size(); ++i) {
      for (int k = 0; k < 4; ++k) sten[4 * i + k] = t.candidates[i].v[k];
      is_vf[i] = t.candidates[i].is_vf;
    }
  };
  if (c->dim == 2) cp(*c->t2); else cp(*c->t3);
  return 0;
}

int oracle_copy_hits(void* h, int* sten, unsigned char* is_vf, double* toi) {
  auto* c = static_cast<OracleCtx*>(h);
  auto cp = [&](auto& t) {
    for (size_t i = 0; i < t.hits.size(); ++i) {
      for (int k = 0; k < 4; ++k) sten[4 * i + k] = t.hits[i].s.v[k];
      is_vf[i] = t.hits[i].s.is_vf;
      toi[i] = t.hits[i].toi;
    }
  };
  if (c->dim == 2) cp(*c->t2); else cp(*c->t3);
  return 0;
}

// The serial record sequence of a single-threaded traverse (threaded = false):
// per event the leaf-pair ordinal, kind (0 candidate, 1 hit, 2 discrete) and
// index into that list.  Returns the count (outputs may be null), or -1 when
// the traverse ran on several threads.
long long oracle_events(void* h, long long* pair, int* kind, long long* idx) {
  auto* c = static_cast<OracleCtx*>(h);
  auto cp = [&](auto& t) -> long long {
    if (t.options.threaded && t.nthreads_eff() > 1) return -1;
    for (size_t i = 0; i < t.events.size(); ++i) {
      if (pair) pair[i] = t.events[i].pair;
      if (kind) kind[i] = t.events[i].kind;
      if (idx) idx[i] = t.events[i].idx;
    }
    return (long long)t.events.size();
  };
  return c->dim == 2 ? cp(*c->t2) : cp(*c->t3);
}

// rank[leaf] = in-order position in the KdBVH (left subtree first)
int oracle_copy_rank(void* h, int* rank) {
  auto* c = static_cast<OracleCtx*>(h);
  auto cp = [&](auto& t) { for (size_t i = 0; i < t.rank.size(); ++i) rank[i] = t.rank[i]; };
  if (c->dim == 2) cp(*c->t2); else cp(*c->t3);
  return 0;
}

// reptri flags: bits 0..2 = v[0..2], bits 3..5 = e[0..2]
int oracle_copy_flags(void* h, unsigned char* out) {
  auto* c = static_cast<OracleCtx*>(h);
  auto cp = [&](auto& t) {
    for (size_t i = 0; i < t.leaves.size(); ++i) {
      unsigned f = 0;
      for (int j = 0; j < 3; ++j) { f |= (t.leaves[i].v[j] ? 1u : 0u) << j; f |= (t.leaves[i].e[j] ? 1u : 0u) << (3 + j); }
      out[i] = (unsigned char)f;
    }
  };
  if (c->dim == 2) cp(*c->t2); else cp(*c->t3);
  return 0;
}

// leaf boxes: per leaf [sw.lo, sw.hi, t0.lo, t0.hi, t1.lo, t1.hi] (6*D doubles) and active
int oracle_copy_boxes(void* h, double* out, unsigned char* active) {
  auto* c = static_cast<OracleCtx*>(h);
  auto cp = [&](auto& t, auto dimtag) {
    constexpr int D = decltype(dimtag)::value;
    for (size_t i = 0; i < t.leaves.size(); ++i) {
      const auto& b = t.leaves[i].box;
      double* o = out + i * 6 * D;
      for (int k = 0; k < D; ++k) {
        o[k] = b.sw.lo[k]; o[D + k] = b.sw.hi[k];
        o[2 * D + k] = b.t0.lo[k]; o[3 * D + k] = b.t0.hi[k];
        o[4 * D + k] = b.t1.lo[k]; o[5 * D + k] = b.t1.hi[k];
      }
      active[i] = b.active;
    }
  };
  if (c->dim == 2) cp(*c->t2, std::integral_constant<int, 2>());
  else cp(*c->t3, std::integral_constant<int, 3>());
  return 0;
}

// Brute-force leaf pairs (i<j); returns count, fills out if cap suffices.
long long oracle_brute_force_pairs(void* h, int* out, long long cap) {
  auto* c = static_cast<OracleCtx*>(h);
  std::vector<std::pair<int, int>> v = c->dim == 2 ? mclo::brute_force_pairs<2>(*c->t2) : mclo::brute_force_pairs<3>(*c->t3);
  if ((long long)v.size() <= cap)
    for (size_t i = 0; i < v.size(); ++i) { out[2 * i] = v[i].first; out[2 * i + 1] = v[i].second; }
  return (long long)v.size();
}

// ---- per-primitive kernels -------------------------------------------------
// verts0/verts1: (dim+1)*dim doubles each. Returns hit (0/1); toi written on hit.
int oracle_accd(int dim, const double* verts0, const double* verts1, double eta, int is_vf, double* toi,
                long long* dist_evals) {
  mclo::AccdStats st;
  double t = -1;
  bool hit;
  if (dim == 3) hit = mclo::additive_ccd<3>(reinterpret_cast<const double(*)[3]>(verts0), reinterpret_cast<const double(*)[3]>(verts1), eta, is_vf != 0, t, &st);
  else hit = mclo::additive_ccd<2>(reinterpret_cast<const double(*)[2]>(verts0), reinterpret_cast<const double(*)[2]>(verts1), eta, true, t, &st);
  if (toi) *toi = hit ? t : -1.0;
  if (dist_evals) *dist_evals = st.dist_evals;
  return hit ? 1 : 0;
}

// additive_ccd at T = float (NarrowPhaseACCD<float,DIM>) on float(verts);
// returns hit, toi written as the float value.
int oracle_accd_f32(int dim, const double* verts0, const double* verts1, double eta, int is_vf, double* toi,
                    long long* dist_evals) {
  mclo::AccdStats st;
  float a[12], b[12], t = -1;
  for (int i = 0; i < (dim + 1) * dim; ++i) { a[i] = (float)verts0[i]; b[i] = (float)verts1[i]; }
  bool hit;
  if (dim == 3) hit = mclo::additive_ccd<3, float>(reinterpret_cast<const float(*)[3]>(a), reinterpret_cast<const float(*)[3]>(b), (float)eta, is_vf != 0, t, &st);
  else hit = mclo::additive_ccd<2, float>(reinterpret_cast<const float(*)[2]>(a), reinterpret_cast<const float(*)[2]>(b), (float)eta, true, t, &st);
  if (toi) *toi = hit ? (double)t : -1.0;
  if (dist_evals) *dist_evals = st.dist_evals;
  return hit ? 1 : 0;
}

double oracle_pair_distance(int dim, const double* verts, int is_vf) {
  if (dim == 3) return mclo::pair_distance<3>(reinterpret_cast<const double(*)[3]>(verts), is_vf != 0);
  return mclo::pair_distance<2>(reinterpret_cast<const double(*)[2]>(verts), true);
}

void oracle_point_on_triangle(const double* q, const double* a, const double* b, const double* c, double* out) {
  mclo::point_on_triangle(q, a, b, c, out);
}

void oracle_edge_to_edge(const double* p0, const double* p1, const double* q0, const double* q1, double* out, double* bary) {
  mclo::edge_to_edge(p0, p1, q0, q1, out, bary);
}

void oracle_point_on_edge2(const double* p, const double* e0, const double* e1, double* out) {
  mclo::point_on_edge2(p, e0, e1, out);
}

int oracle_tri_tri(const double* f0, const double* f1) {
  return mclo::tt::tri_tri_overlap_3d(f0, f0 + 3, f0 + 6, f1, f1 + 3, f1 + 6);
}

int oracle_edge_edge2(const double* e0, const double* e1) {
  return mclo::discrete_edge_edge2(e0, e0 + 2, e1, e1 + 2) ? 1 : 0;
}

// ---- CTCD (vf_one_sided) ----------------------------------------------------
// query_ccd_vf with the wrong-side test: 1 hit (toi written), 0 none, -1 only wrong-side hits.
int oracle_ctcd_vf(int dim, const double* verts0, const double* verts1, double eta, double* toi) {
  double t = -1;
  int r;
  if (dim == 3) r = mclo::ctcd::query_vf<3>(reinterpret_cast<const double(*)[3]>(verts0), reinterpret_cast<const double(*)[3]>(verts1), eta, t);
  else r = mclo::ctcd::query_vf<2>(reinterpret_cast<const double(*)[2]>(verts0), reinterpret_cast<const double(*)[2]>(verts1), eta, t);
  if (toi) *toi = r == 1 ? t : -1.0;
  return r;
}

// Jenkins-Traub: number of zeros (or -1); real and imaginary parts in zr, zi.
int oracle_rpoly(const double* op, int degree, double* zr, double* zi) {
  mclo::ctcd::Rpoly rf;
  return rf.solve(op, degree, zr, zi);
}

// CTCD intervals of a polynomial (op is copied; the reference normalises in place).
int oracle_find_intervals(const double* op, int n, int pos, double* lo, double* hi, int cap) {
  double c[7];
  for (int i = 0; i <= n && i < 7; ++i) c[i] = op[i];
  std::vector<mclo::ctcd::Interval> v;
  mclo::ctcd::find_intervals(c, n, v, pos != 0);
  for (int i = 0; i < (int)v.size() && i < cap; ++i) { lo[i] = v[i].l; hi[i] = v[i].u; }
  return (int)v.size();
}

// Sub-tests of the 3D query (kind 0 vertex-face, 1 vertex-edge, 2 vertex-vertex),
// verts = the sub-test's points at t0 then t1 (4, 3 or 2 points of 3 doubles).
// Returns the hit flag; the candidate times go to out (count in *nt).
int oracle_ctcd_subtest(int kind, const double* x0, const double* x1, double eta, double* out, int cap, int* nt) {
  std::vector<double> all_t;
  bool hit = false;
  if (kind == 0) hit = mclo::ctcd::vertex_face(x0, x0 + 3, x0 + 6, x0 + 9, x1, x1 + 3, x1 + 6, x1 + 9, eta, all_t);
  else if (kind == 1) hit = mclo::ctcd::vertex_edge3(x0, x0 + 3, x0 + 6, x1, x1 + 3, x1 + 6, eta, all_t);
  else hit = mclo::ctcd::vertex_vertex3(x0, x0 + 3, x1, x1 + 3, eta, all_t);
  for (int i = 0; i < (int)all_t.size() && i < cap; ++i) out[i] = all_t[i];
  *nt = (int)all_t.size();
  return hit ? 1 : 0;
}

int oracle_wrong_side(int dim, const double* verts0, const double* verts1, double t) {
  if (dim == 3) return mclo::ctcd::wrong_side3(reinterpret_cast<const double(*)[3]>(verts0), reinterpret_cast<const double(*)[3]>(verts1), t);
  return mclo::ctcd::wrong_side2(reinterpret_cast<const double(*)[2]>(verts0), reinterpret_cast<const double(*)[2]>(verts1), t);
}

int oracle_ray_box(const double* p0, const double* p1, const double* bmin, const double* bmax) {
  return mclo::ctcd::ray_box(p0, p1, bmin, bmax) ? 1 : 0;
}

}  // extern "C"

// ---- TOI and its reference gradient ---------------------------------------
namespace {
template <int D>
bool accd_grad_one(const double* verts0, const double* verts1, double eta, bool is_vf, double* g0, double* g1,
                   double* toi_out) {
  constexpr int ns = D + 1, N = 2 * ns * D;
  using T = mclo::Dual<N>;
  T a[ns][D], b[ns][D];
  for (int i = 0; i < ns; ++i)
    for (int k = 0; k < D; ++k) {
      a[i][k] = T(verts0[i * D + k]);
      a[i][k].g[i * D + k] = 1.0;
      b[i][k] = T(verts1[i * D + k]);
      b[i][k].g[ns * D + i * D + k] = 1.0;
    }
  T toi(-1.0);
  const bool hit = mclo::additive_ccd<D, T>(a, b, eta, is_vf, toi);
  if (toi_out) *toi_out = hit ? toi.v : -1.0;
  for (int j = 0; j < ns * D; ++j) {
    g0[j] = hit ? toi.g[j] : 0.0;
    g1[j] = hit ? toi.g[ns * D + j] : 0.0;
  }
  return hit;
}
}  // namespace

extern "C" {

// d toi / d(verts0, verts1) by forward-mode dual numbers through the oracle's
// additive_ccd; g0, g1: (dim+1)*dim each.  Returns hit.
int oracle_accd_grad(int dim, const double* verts0, const double* verts1, double eta, int is_vf, double* g0,
                     double* g1, double* toi) {
  if (dim == 3) return accd_grad_one<3>(verts0, verts1, eta, is_vf != 0, g0, g1, toi) ? 1 : 0;
  return accd_grad_one<2>(verts0, verts1, eta, true, g0, g1, toi) ? 1 : 0;
}

// toi (or -1) of every stencil of a list.
// toi[i] = ACCD TOI of stencil i (-1: no hit); sig (optional) = its branch signature.
void oracle_toi_list(int dim, const double* V0, const double* V1, const int* sten, const unsigned char* is_vf,
                     long long n, double eta_vf, double eta_ee, double* toi, unsigned long long* sig) {
#pragma omp parallel for schedule(dynamic, 256)
  for (long long i = 0; i < n; ++i) {
    uint64_t word = 0;
    mclo::branch_sig = sig ? &word : nullptr;
    double a[12], b[12], t = -1;
    const int ns = dim + 1;
    for (int j = 0; j < ns; ++j)
      for (int k = 0; k < dim; ++k) { a[j * dim + k] = V0[(size_t)sten[4 * i + j] * dim + k]; b[j * dim + k] = V1[(size_t)sten[4 * i + j] * dim + k]; }
    const bool vf = dim == 2 || is_vf[i];
    bool hit;
    if (dim == 3) hit = mclo::additive_ccd<3>(reinterpret_cast<const double(*)[3]>(a), reinterpret_cast<const double(*)[3]>(b), vf ? eta_vf : eta_ee, vf, t);
    else hit = mclo::additive_ccd<2>(reinterpret_cast<const double(*)[2]>(a), reinterpret_cast<const double(*)[2]>(b), eta_vf, true, t);
    toi[i] = hit ? t : -1.0;
    if (sig) sig[i] = word;
    mclo::branch_sig = nullptr;
  }
}

// CTCD TOI per vertex-face stencil (-1: no hit, or only wrong-side hits).
void oracle_ctcd_list(int dim, const double* V0, const double* V1, const int* sten, long long n, double eta,
                      double* toi) {
#pragma omp parallel for schedule(dynamic, 256)
  for (long long i = 0; i < n; ++i) {
    double a[12], b[12];
    for (int j = 0; j <= dim; ++j)
      for (int k = 0; k < dim; ++k) { a[j * dim + k] = V0[(size_t)sten[4 * i + j] * dim + k]; b[j * dim + k] = V1[(size_t)sten[4 * i + j] * dim + k]; }
    oracle_ctcd_vf(dim, a, b, eta, toi + i);
  }
}

// gV0, gV1 (nv*dim, overwritten) = sum_i w_i d toi_i / d(V0, V1) over a stencil list.
void oracle_toi_grad_list(int dim, const double* V0, const double* V1, int nv, const int* sten,
                          const unsigned char* is_vf, const double* w, long long n, double eta_vf, double eta_ee,
                          double* gV0, double* gV1) {
  const size_t m = (size_t)nv * dim;
  std::vector<double> acc0(m, 0.0), acc1(m, 0.0);
#pragma omp parallel
  {
    std::vector<double> l0(m, 0.0), l1(m, 0.0);
#pragma omp for schedule(dynamic, 256)
    for (long long i = 0; i < n; ++i) {
      double a[12], b[12], g0[12], g1[12];
      const int ns = dim + 1;
      for (int j = 0; j < ns; ++j)
        for (int k = 0; k < dim; ++k) { a[j * dim + k] = V0[(size_t)sten[4 * i + j] * dim + k]; b[j * dim + k] = V1[(size_t)sten[4 * i + j] * dim + k]; }
      const bool vf = dim == 2 || is_vf[i];
      if (dim == 3) accd_grad_one<3>(a, b, vf ? eta_vf : eta_ee, vf, g0, g1, nullptr);
      else accd_grad_one<2>(a, b, eta_vf, true, g0, g1, nullptr);
      for (int j = 0; j < ns; ++j)
        for (int k = 0; k < dim; ++k) {
          l0[(size_t)sten[4 * i + j] * dim + k] += w[i] * g0[j * dim + k];
          l1[(size_t)sten[4 * i + j] * dim + k] += w[i] * g1[j * dim + k];
        }
    }
#pragma omp critical
    for (size_t j = 0; j < m; ++j) { acc0[j] += l0[j]; acc1[j] += l1[j]; }
  }
  std::memcpy(gV0, acc0.data(), sizeof(double) * m);
  std::memcpy(gV1, acc1.data(), sizeof(double) * m);
}

int oracle_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
