// Host build of the library's ACCD iteration (mclccd_amd/csrc/geom.hpp, the
// code the GPU kernels run) checked against the oracle's additive_ccd
// (NarrowPhase.cpp:634-711): the uninterrupted run, and every stencil that the
// kernels would suspend after three iterations, resumed from its stored state
// (t1, t2, t3, g as doubles) on a fresh Accd — hit flag and TOI bits equal.
//
// Compiled host-only by hipcc (-x hip --cuda-host-only) with __device__ also
// marking host functions; x86-64 SSE2 arithmetic without contraction is the
// same IEEE arithmetic as the GPU's correctly rounded operations.
//
// Input (binary, little endian): int32 dim, int64 nv, double V0[nv*dim],
// double V1[nv*dim], int64 n, int32 sten[n*4], uint8 is_vf[n], double eta.
// Output: one line "accd_resume_check: <n> stencils, <s> suspended, <f> failed".
#include <hip/hip_runtime.h>
#undef __device__
#define __device__ __attribute__((host)) __attribute__((device))
#include "geom.hpp"
#include "../../oracle/mclccd_oracle.hpp"

#include <cstdio>
#include <cstring>
#include <vector>

namespace {

bool same(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

template <int D, bool VF>
bool check(const double (*a)[D], const double (*b)[D], double eta, long long& suspended) {
  using V = typename std::conditional<D == 3, mclg::d3, mclg::d2>::type;
  V v0[D + 1], v1[D + 1];
  for (int j = 0; j < D + 1; ++j) {
    double* p0 = reinterpret_cast<double*>(&v0[j]);
    double* p1 = reinterpret_cast<double*>(&v1[j]);
    for (int k = 0; k < D; ++k) { p0[k] = a[j][k]; p1[k] = b[j][k]; }
  }
  double rt = -1;
  const bool rh = mclo::additive_ccd<D>(a, b, eta, VF, rt);
  const double want = rh ? rt : -1.0;
  // uninterrupted
  double t = -1;
  int ev = 0;
  const int h = mclg::accd<D, VF, double>(v0, v1, eta, t, ev);
  if ((h != 0) != rh || (h && !same(t, want))) return false;
  // suspended after three iterations, resumed on a fresh state
  mclg::Accd<D, VF, double> s;
  int st = s.template start<mclg::kAccdSuspendIter>(v0, v1, eta);
  if (st == mclg::ACCD_RUN && s.replayable) {
    ++suspended;
    const double t1 = s.t1, t2 = s.t2, t3 = s.tl, g = s.g;
    mclg::Accd<D, VF, double> r;
    r.resume(v0, v1, eta, t1, t2, t3, g);
    // the resumed state is the loop's state at the top of iteration 3, bit for bit
    if (std::memcmp(&r.st, &s.st, sizeof(s.st)) != 0 || !same(r.toi, s.toi) || !same(r.tl, s.tl) ||
        !same(r.g, s.g) || !same(r.l_p, s.l_p) || r.iter != s.iter)
      return false;
    st = r.run();
    const double got = st == mclg::ACCD_HIT ? r.toi : -1.0;
    if ((st == mclg::ACCD_HIT) != rh || !same(got, want)) return false;
    if (ev != s.evals + r.evals) return false;  // no evaluation repeated or lost
  } else {
    if (st == mclg::ACCD_RUN) st = s.run();
    const double got = st == mclg::ACCD_HIT ? s.toi : -1.0;
    if ((st == mclg::ACCD_HIT) != rh || !same(got, want)) return false;
  }
  return true;
}

template <typename X> bool rd(FILE* f, X* p, size_t n) { return std::fread(p, sizeof(X), n, f) == n; }

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s input.bin\n", argv[0]);
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int dim = 0;
  long long nv = 0, n = 0;
  if (!rd(f, &dim, 1) || !rd(f, &nv, 1)) return 2;
  std::vector<double> V0((size_t)nv * dim), V1((size_t)nv * dim);
  if (!rd(f, V0.data(), V0.size()) || !rd(f, V1.data(), V1.size()) || !rd(f, &n, 1)) return 2;
  std::vector<int> S((size_t)n * 4);
  std::vector<unsigned char> vf((size_t)n);
  double eta = 0;
  if (!rd(f, S.data(), S.size()) || !rd(f, vf.data(), vf.size()) || !rd(f, &eta, 1)) return 2;
  std::fclose(f);
  long long failed = 0, suspended = 0;
  for (long long i = 0; i < n; ++i) {
    bool ok;
    if (dim == 3) {
      double a[4][3], b[4][3];
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 3; ++k) {
          a[j][k] = V0[(size_t)S[4 * i + j] * 3 + k];
          b[j][k] = V1[(size_t)S[4 * i + j] * 3 + k];
        }
      ok = vf[i] ? check<3, true>(a, b, eta, suspended) : check<3, false>(a, b, eta, suspended);
    } else {
      double a[3][2], b[3][2];
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 2; ++k) {
          a[j][k] = V0[(size_t)S[4 * i + j] * 2 + k];
          b[j][k] = V1[(size_t)S[4 * i + j] * 2 + k];
        }
      ok = check<2, true>(a, b, eta, suspended);
    }
    if (!ok && failed++ < 5) std::printf("mismatch at stencil %lld\n", i);
  }
  std::printf("accd_resume_check: %lld stencils, %lld suspended, %lld failed\n", n, suspended, failed);
  return failed ? 1 : 0;
}
