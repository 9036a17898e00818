// TEST INFRASTRUCTURE ONLY: a CPU driver built with AddressSanitizer and
// UndefinedBehaviorSanitizer (make -C oracle sanitize) over the oracle's C
// entry points (oracle_capi.cpp) and the library's host-only ordering code
// (mclccd_amd/csrc/hostorder.hpp: the KdBVH leaf order and the callback
// replay sequence).  Scenes are generated here; every check is a property or a
// cross-check between the two, so a sanitizer report or a failed check ends
// the run with a non-zero status.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <numeric>
#include <set>
#include <tuple>
#include <vector>

#include "../mclccd_amd/csrc/hostorder.hpp"

struct oracle_options {
  double box_eta, vf_ccd_eta, ee_ccd_eta;
  int reptri, threaded, discrete, continuous;
  int vf_one_sided;
  int f32;
};

extern "C" {
void* oracle_create(int dim);
void oracle_destroy(void* h);
int oracle_set_options(void* h, const oracle_options* o, int nthreads);
int oracle_update(void* h, const double* V0, const double* V1, const int* P, int np, int pdim, const int* active);
int oracle_traverse(void* h, const double* V0, const double* V1, const int* P, int want_leaf_pairs, int want_candidates);
long long oracle_count(void* h, int which);
int oracle_copy_leaf_pairs(void* h, int* out);
int oracle_copy_discrete(void* h, int* out);
int oracle_copy_candidates(void* h, int* sten, unsigned char* is_vf);
int oracle_copy_hits(void* h, int* sten, unsigned char* is_vf, double* toi);
long long oracle_events(void* h, long long* pair, int* kind, long long* idx);
int oracle_copy_rank(void* h, int* rank);
int oracle_copy_flags(void* h, unsigned char* out);
int oracle_copy_boxes(void* h, double* out, unsigned char* active);
long long oracle_brute_force_pairs(void* h, int* out, long long cap);
int oracle_accd(int dim, const double* verts0, const double* verts1, double eta, int is_vf, double* toi,
                long long* dist_evals);
int oracle_accd_f32(int dim, const double* verts0, const double* verts1, double eta, int is_vf, double* toi,
                    long long* dist_evals);
int oracle_accd_grad(int dim, const double* verts0, const double* verts1, double eta, int is_vf, double* g0,
                     double* g1, double* toi);
void oracle_toi_list(int dim, const double* V0, const double* V1, const int* sten, const unsigned char* is_vf,
                     long long n, double eta_vf, double eta_ee, double* toi, unsigned long long* sig);
void oracle_ctcd_list(int dim, const double* V0, const double* V1, const int* sten, long long n, double eta,
                      double* toi);
}

namespace {

long long g_checks = 0, g_failed = 0;

void expect(bool ok, const char* scene, const char* what) {
  ++g_checks;
  if (!ok) {
    ++g_failed;
    std::fprintf(stderr, "FAILED [%s] %s\n", scene, what);
  }
}

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 0x9e3779b97f4a7c15ull + 1) {}
  uint64_t next() {  // splitmix64
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  int below(int n) { return (int)(next() % (uint64_t)n); }
};

struct Scene {
  const char* name;
  int dim;
  std::vector<double> V0, V1;
  std::vector<int> F;
  double eta;
  int reptri = 1, one_sided = 0, f32 = 0;
  int nprims() const { return (int)F.size() / dim; }
  int nverts() const { return (int)V0.size() / dim; }
};

// Small primitives (3D triangles or 2D edges) scattered at constant density,
// about a third of the corners shared with a recent primitive, moving by up to
// 'motion' per axis.
Scene soup(const char* name, int dim, int n, uint64_t seed, double motion) {
  Scene s{name, dim, {}, {}, {}, 1e-2};
  Rng r(seed);
  const double L = 0.2 * std::pow((double)n, 1.0 / dim);
  for (int i = 0; i < n; ++i) {
    double c[3];
    for (int k = 0; k < dim; ++k) c[k] = L * r.uni();
    for (int j = 0; j < dim; ++j) {
      if (i > 0 && r.uni() < 0.33) {
        const int from = std::max(0, i - 12) + r.below(std::min(i, 12));
        const int v = s.F[(size_t)from * dim + r.below(dim)];
        bool dup = false;
        for (int q = 0; q < j; ++q) dup |= s.F[(size_t)i * dim + q] == v;
        if (!dup) {
          s.F.push_back(v);
          continue;
        }
      }
      s.F.push_back(s.nverts());
      for (int k = 0; k < dim; ++k) {
        const double x = c[k] + 0.1 * (r.uni() - 0.5);
        s.V0.push_back(x);
        s.V1.push_back(x + motion * (r.uni() - 0.5));
      }
    }
  }
  return s;
}

// Two k x k triangulated sheets: one resting, one 0.05 above it that tilts
// through it (crossing it along x = 0.5 at the end of the step).
Scene sheets(int k) {
  Scene s{"sheets", 3, {}, {}, {}, 1e-3};
  for (int layer = 0; layer < 2; ++layer) {
    const int base = s.nverts();
    for (int y = 0; y <= k; ++y)
      for (int x = 0; x <= k; ++x) {
        const double px = (x + 0.37 * layer) / k, py = (y + 0.21 * layer) / k;
        s.V0.insert(s.V0.end(), {px, py, layer ? 0.05 : 0.0});
        s.V1.insert(s.V1.end(), {px, py, layer ? 0.3 * (px - 0.5) : 0.0});
      }
    for (int y = 0; y < k; ++y)
      for (int x = 0; x < k; ++x) {
        const int a = base + y * (k + 1) + x, b = a + 1, c = a + k + 1, d = c + 1;
        s.F.insert(s.F.end(), {a, b, d, a, d, c});
      }
  }
  return s;
}

// A soup and a coincident copy of it with its own vertex ids.
Scene coincident(int n) {
  Scene s = soup("coincident", 3, n, 5, 0.1);
  const int nv = s.nverts();
  const size_t nf = s.F.size();
  s.V0.insert(s.V0.end(), s.V0.begin(), s.V0.end());
  s.V1.insert(s.V1.end(), s.V1.begin(), s.V1.end());
  for (size_t i = 0; i < nf; ++i) s.F.push_back(s.F[i] + nv);
  return s;
}

Scene prefix(const Scene& s, int np, const char* name) {
  Scene t = s;
  t.name = name;
  t.F.resize((size_t)np * s.dim);
  return t;
}

struct Run {
  std::vector<int> pairs, disc, rank, csten, hsten;
  std::vector<unsigned char> flags, cvf, hvf;
  std::vector<double> htoi, boxes;
  std::vector<long long> epair, eidx;
  std::vector<int> ekind;
};

Run run_oracle(const Scene& s, int threaded, int nthreads, bool cand) {
  void* h = oracle_create(s.dim);
  oracle_options o{s.eta, s.eta, s.eta, s.reptri, threaded, 1, 1, s.one_sided, s.f32};
  oracle_set_options(h, &o, nthreads);
  const int np = s.nprims();
  oracle_update(h, s.V0.data(), s.V1.data(), s.F.data(), np, s.dim, nullptr);
  oracle_traverse(h, s.V0.data(), s.V1.data(), s.F.data(), 1, cand ? 1 : 0);
  Run r;
  r.pairs.resize((size_t)oracle_count(h, 0) * 2);
  oracle_copy_leaf_pairs(h, r.pairs.data());
  r.disc.resize((size_t)oracle_count(h, 3) * 2);
  oracle_copy_discrete(h, r.disc.data());
  const size_t nc = (size_t)oracle_count(h, 1), nh = (size_t)oracle_count(h, 2);
  r.csten.resize(4 * nc);
  r.cvf.resize(nc);
  oracle_copy_candidates(h, r.csten.data(), r.cvf.data());
  r.hsten.resize(4 * nh);
  r.hvf.resize(nh);
  r.htoi.resize(nh);
  oracle_copy_hits(h, r.hsten.data(), r.hvf.data(), r.htoi.data());
  r.rank.resize((size_t)np);
  r.flags.resize((size_t)np);
  r.boxes.resize((size_t)np * 6 * s.dim);
  std::vector<unsigned char> act((size_t)np);
  if (np > 0) {
    oracle_copy_rank(h, r.rank.data());
    oracle_copy_flags(h, r.flags.data());
    oracle_copy_boxes(h, r.boxes.data(), act.data());
  }
  const long long ne = oracle_events(h, nullptr, nullptr, nullptr);
  if (ne >= 0) {
    r.epair.resize((size_t)ne);
    r.ekind.resize((size_t)ne);
    r.eidx.resize((size_t)ne);
    oracle_events(h, r.epair.data(), r.ekind.data(), r.eidx.data());
  }
  if (threaded == 0) {  // the tree's leaf pairs against brute force over its leaf boxes
    const long long nb = oracle_brute_force_pairs(h, nullptr, 0);
    std::vector<int> bf((size_t)nb * 2);
    oracle_brute_force_pairs(h, bf.data(), nb);
    std::set<std::pair<int, int>> a, b;
    for (size_t i = 0; i < r.pairs.size(); i += 2)
      a.insert({std::min(r.pairs[i], r.pairs[i + 1]), std::max(r.pairs[i], r.pairs[i + 1])});
    for (size_t i = 0; i < bf.size(); i += 2) b.insert({bf[i], bf[i + 1]});
    expect(a == b && a.size() * 2 == r.pairs.size(), s.name, "leaf pairs == brute force");
  }
  oracle_destroy(h);
  return r;
}

using HitKey = std::tuple<int, int, int, int, int, double>;
std::multiset<HitKey> hit_set(const Run& r) {
  std::multiset<HitKey> m;
  for (size_t i = 0; i < r.hvf.size(); ++i)
    m.insert({r.hsten[4 * i], r.hsten[4 * i + 1], r.hsten[4 * i + 2], r.hsten[4 * i + 3], r.hvf[i], r.htoi[i]});
  return m;
}

// The replay sequence rebuilt by the host code from the records in shuffled
// order equals the oracle's serial sequence (kinds 1/2, or 0/2 over candidates).
void check_replay(const Scene& s, const Run& r, bool cand, Rng& rng) {
  const int np = s.nprims();
  std::vector<int> rank((size_t)np), gap((size_t)std::max(np - 1, 0));
  std::vector<double> cen((size_t)np * s.dim);
  const int D = s.dim;
  for (int i = 0; i < np; ++i)
    for (int k = 0; k < D; ++k)
      cen[(size_t)i * D + k] = s.f32 ? (double)(((float)r.boxes[(size_t)i * 6 * D + k] +
                                                 (float)r.boxes[(size_t)i * 6 * D + D + k]) / 2.0f)
                                     : (r.boxes[(size_t)i * 6 * D + k] + r.boxes[(size_t)i * 6 * D + D + k]) / 2.0;
  if (D == 3) mclg::host::kd_order<3>(cen.data(), np, rank.data(), np > 1 ? gap.data() : nullptr);
  else mclg::host::kd_order<2>(cen.data(), np, rank.data(), np > 1 ? gap.data() : nullptr);
  expect(rank == r.rank, s.name, "host KdBVH rank == oracle rank");
  bool left_first = true;
  for (size_t i = 0; i < r.pairs.size(); i += 2) left_first &= rank[r.pairs[i]] < rank[r.pairs[i + 1]];
  expect(left_first, s.name, "oracle leaf pairs put the lower rank first");

  const size_t n = cand ? r.cvf.size() : r.hvf.size();
  const int* sten = cand ? r.csten.data() : r.hsten.data();
  const unsigned char* vf = cand ? r.cvf.data() : r.hvf.data();
  std::vector<size_t> perm(n), dperm(r.disc.size() / 2);
  std::iota(perm.begin(), perm.end(), 0);
  std::iota(dperm.begin(), dperm.end(), 0);
  for (size_t i = n; i > 1; --i) std::swap(perm[i - 1], perm[(size_t)rng.below((int)i)]);
  for (size_t i = dperm.size(); i > 1; --i) std::swap(dperm[i - 1], dperm[(size_t)rng.below((int)i)]);
  std::vector<int> ss(4 * n), sd(r.disc.size());
  std::vector<unsigned char> sv(n);
  for (size_t j = 0; j < n; ++j) {
    for (int k = 0; k < 4; ++k) ss[4 * j + k] = sten[4 * perm[j] + k];
    sv[j] = vf[perm[j]];
  }
  for (size_t j = 0; j < dperm.size(); ++j) {
    sd[2 * j] = r.disc[2 * dperm[j]];
    sd[2 * j + 1] = r.disc[2 * dperm[j] + 1];
  }
  std::vector<int> P(s.F.begin(), s.F.begin() + (size_t)np * D);
  const long long total = (long long)(n + dperm.size());
  std::vector<int64_t> ev((size_t)total + 1), grp((size_t)total + 1);
  const int64_t got = mclg::host::build_events(D, P.data(), r.flags.data(), rank.data(), r.pairs.data(),
                                               (int64_t)(r.pairs.size() / 2), &gap, true, true, cand, ss.data(),
                                               sv.data(), (int64_t)n, sd.data(), (int64_t)dperm.size(), ev.data(),
                                               grp.data());
  expect(got == total, s.name, cand ? "every candidate placed" : "every hit placed");
  if (got != total) return;
  // the oracle's sequence, by content (equal stencils of different leaf pairs are interchangeable)
  std::vector<std::vector<int>> want, have;
  std::vector<long long> want_pair;
  for (size_t e = 0; e < r.ekind.size(); ++e) {
    const int kd = r.ekind[e];
    const size_t i = (size_t)r.eidx[e];
    if (kd == 2) want.push_back({2, r.disc[2 * i], r.disc[2 * i + 1]});
    else if (kd == (cand ? 0 : 1)) {
      const int* st = cand ? &r.csten[4 * i] : &r.hsten[4 * i];
      want.push_back({kd, st[0], st[1], st[2], st[3], cand ? r.cvf[i] : r.hvf[i]});
    } else continue;
    want_pair.push_back(r.epair[e]);
  }
  for (int64_t e = 0; e < got; ++e) {
    const int64_t x = ev[(size_t)e];
    if (x < 0) {
      const size_t d = (size_t)(-1 - x);
      have.push_back({2, sd[2 * d], sd[2 * d + 1]});
    } else {
      const int* st = &ss[4 * (size_t)x];
      have.push_back({cand ? 0 : 1, st[0], st[1], st[2], st[3], sv[(size_t)x]});
    }
  }
  expect(have == want, s.name, cand ? "candidate replay == oracle serial order" : "hit replay == oracle serial order");
  // group numbers: one per leaf pair that has an event, in visiting order
  bool groups_ok = want_pair.size() == (size_t)got;
  for (size_t e = 1; groups_ok && e < want_pair.size(); ++e) {
    const bool same = want_pair[e] == want_pair[e - 1];
    groups_ok = grp[e] == grp[e - 1] + (same ? 0 : 1);
  }
  if (groups_ok && got > 0) groups_ok = grp[0] == 0;
  expect(groups_ok, s.name, "event groups follow the leaf pairs");
}

// Every hit's TOI again from the single-stencil entry, the list entry and the
// dual-number gradient; no candidate outside the hit list reports a hit.
void check_narrow(const Scene& s, const Run& serial) {
  const int D = s.dim;
  std::vector<double> toi(serial.cvf.size());
  if (s.one_sided) {
    std::vector<int> vsten;
    std::vector<size_t> at;
    for (size_t i = 0; i < serial.cvf.size(); ++i)
      if (serial.cvf[i] || D == 2) {
        vsten.insert(vsten.end(), serial.csten.begin() + 4 * i, serial.csten.begin() + 4 * i + 4);
        at.push_back(i);
      }
    std::vector<double> t(at.size());
    oracle_ctcd_list(D, s.V0.data(), s.V1.data(), vsten.data(), (long long)at.size(), s.eta, t.data());
    long long nvf_hits = 0;
    for (double x : t) nvf_hits += x >= 0;
    long long want = 0;
    for (size_t i = 0; i < serial.hvf.size(); ++i) want += serial.hvf[i] || D == 2;
    expect(nvf_hits == want, s.name, "CTCD list hits == traverse's VF hits");
    return;
  }
  if (s.f32) {  // the float tree: float ACCD per candidate on float(V)
    for (size_t i = 0; i < toi.size(); ++i) {
      double a[12], b[12];
      for (int j = 0; j <= D; ++j)
        for (int k = 0; k < D; ++k) {
          a[j * D + k] = s.V0[(size_t)serial.csten[4 * i + j] * D + k];
          b[j * D + k] = s.V1[(size_t)serial.csten[4 * i + j] * D + k];
        }
      if (!oracle_accd_f32(D, a, b, (float)s.eta, serial.cvf[i] || D == 2, &toi[i], nullptr)) toi[i] = -1;
    }
  } else {
    oracle_toi_list(D, s.V0.data(), s.V1.data(), serial.csten.data(), serial.cvf.data(), (long long)serial.cvf.size(),
                    s.eta, s.eta, toi.data(), nullptr);
  }
  std::multiset<HitKey> from_list;
  for (size_t i = 0; i < toi.size(); ++i)
    if (toi[i] >= 0)
      from_list.insert({serial.csten[4 * i], serial.csten[4 * i + 1], serial.csten[4 * i + 2], serial.csten[4 * i + 3],
                        serial.cvf[i], toi[i]});
  expect(from_list == hit_set(serial), s.name, "toi list over candidates == traverse hits");
  if (s.f32) return;  // the single-stencil and gradient entries are FP64
  bool same = true, grad_ok = true;
  const size_t step = std::max<size_t>(1, serial.hvf.size() / 64);
  for (size_t i = 0; i < serial.hvf.size(); i += step) {
    double a[12], b[12], t = -1, tg = -1, g0[12], g1[12];
    long long ev = 0;
    for (int j = 0; j <= D; ++j)
      for (int k = 0; k < D; ++k) {
        a[j * D + k] = s.V0[(size_t)serial.hsten[4 * i + j] * D + k];
        b[j * D + k] = s.V1[(size_t)serial.hsten[4 * i + j] * D + k];
      }
    const int hit = oracle_accd(D, a, b, s.eta, serial.hvf[i], &t, &ev);
    same &= hit == 1 && t == serial.htoi[i] && ev > 0;
    const int hg = oracle_accd_grad(D, a, b, s.eta, serial.hvf[i], g0, g1, &tg);
    bool finite = true;
    for (int j = 0; j < (D + 1) * D; ++j) finite &= std::isfinite(g0[j]) && std::isfinite(g1[j]);
    grad_ok &= hg == 1 && tg == t && finite;
  }
  expect(same, s.name, "single-stencil ACCD == traverse TOI");
  expect(grad_ok, s.name, "dual-number ACCD: same TOI, finite gradient");
}

void check_scene(const Scene& s, uint64_t seed) {
  Rng rng(seed);
  const Run serial = run_oracle(s, 0, 1, true);
  const Run hits_only = run_oracle(s, 0, 1, false);
  const Run threaded = run_oracle(s, 1, 4, true);
  expect(hit_set(serial) == hit_set(threaded) && hit_set(serial) == hit_set(hits_only), s.name,
         "threaded hits == serial hits");
  auto dset = [](const Run& r) {
    std::multiset<std::pair<int, int>> m;
    for (size_t i = 0; i < r.disc.size(); i += 2) m.insert({r.disc[i], r.disc[i + 1]});
    return m;
  };
  expect(dset(serial) == dset(threaded), s.name, "threaded discrete == serial discrete");
  expect(serial.pairs.size() == threaded.pairs.size(), s.name, "threaded leaf pairs == serial leaf pairs");
  check_replay(s, serial, true, rng);
  check_replay(s, hits_only, false, rng);
  check_narrow(s, serial);
  std::printf("%-14s dim %d  prims %6d  leaf pairs %7zu  candidates %7zu  hits %6zu  discrete %5zu\n", s.name, s.dim,
              s.nprims(), serial.pairs.size() / 2, serial.cvf.size(), serial.hvf.size(), serial.disc.size() / 2);
}

}  // namespace

int main() {
  std::vector<Scene> scenes;
  scenes.push_back(soup("soup3", 3, 600, 2, 0.2));
  scenes.push_back(sheets(14));
  scenes.push_back(coincident(150));
  scenes.push_back(soup("edges2", 2, 800, 7, 0.2));
  Scene noreptri = soup("soup3-noreptri", 3, 300, 11, 0.2);
  noreptri.reptri = 0;
  scenes.push_back(noreptri);
  Scene ctcd = soup("soup3-ctcd", 3, 300, 13, 0.2);
  ctcd.one_sided = 1;
  scenes.push_back(ctcd);
  Scene f3 = soup("soup3-f32", 3, 300, 19, 0.2);
  f3.f32 = 1;
  scenes.push_back(f3);
  Scene f2 = soup("edges2-f32", 2, 400, 23, 0.2);
  f2.f32 = 1;
  scenes.push_back(f2);
  Scene ctcd2 = soup("edges2-ctcd", 2, 300, 17, 0.2);
  ctcd2.one_sided = 1;
  scenes.push_back(ctcd2);
  for (int np = 0; np <= 3; ++np) {
    static const char* names3[] = {"soup3-0", "soup3-1", "soup3-2", "soup3-3"};
    static const char* names2[] = {"edges2-0", "edges2-1", "edges2-2", "edges2-3"};
    scenes.push_back(prefix(scenes[0], np, names3[np]));
    scenes.push_back(prefix(scenes[3], np, names2[np]));
  }
  uint64_t seed = 1;
  for (const Scene& s : scenes) check_scene(s, seed++);
  std::printf("sanitize_check: %zu scenes, %lld checks, %lld failed\n", scenes.size(), g_checks, g_failed);
  return g_failed == 0 ? 0 : 1;
}
