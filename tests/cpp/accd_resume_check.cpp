// Host build of the library's ACCD iteration (mclccd_amd/csrc/geom.hpp, the
// code the GPU kernels run) checked against the oracle's additive_ccd
// (NarrowPhase.cpp:634-711): the uninterrupted run, and every stencil that the
// kernels would suspend after three iterations, resumed from its stored state
// (t1, t2, t3, g as doubles) on a fresh Accd — hit flag and TOI bits equal; in
// double (BVHTree<double,DIM>) and in float (BVHTree<float,DIM>, on the
// float-rounded input, against the oracle's float ACCD).
//
// Compiled host-only by hipcc (-x hip --cuda-host-only) with __device__ also
// marking host functions; x86-64 SSE2 arithmetic without contraction is the
// same IEEE arithmetic as the GPU's correctly rounded operations.
//
// Input (binary, little endian): int32 dim, int64 nv, double V0[nv*dim],
// double V1[nv*dim], int64 n, int32 sten[n*4], uint8 is_vf[n], double eta.
// Output: one line "accd_resume_check: <n> stencils, <s> suspended, <f> failed,
// <sf> suspended in float".
#include <hip/hip_runtime.h>
#undef __device__
#define __device__ __attribute__((host)) __attribute__((device))
#include "geom.hpp"
#include "../../oracle/mclccd_oracle.hpp"

#include <cstdio>
#include <cstring>
#include <vector>

namespace {

template <typename T> bool same_t(T a, T b) { return std::memcmp(&a, &b, sizeof(T)) == 0; }

// T = double: BVHTree<double,DIM>; T = float: BVHTree<float,DIM> on the
// float-rounded coordinates and gap (NarrowPhaseACCD<float,DIM>).  The
// suspended state travels as doubles, as in the kernels.
template <int D, bool VF, typename T>
bool check(const double (*a)[D], const double (*b)[D], double eta, long long& suspended) {
  using V = typename std::conditional<D == 3, mclg::vec3<T>, mclg::vec2<T>>::type;
  V v0[D + 1], v1[D + 1];
  T fa[D + 1][D], fb[D + 1][D];
  for (int j = 0; j < D + 1; ++j) {
    T* p0 = reinterpret_cast<T*>(&v0[j]);
    T* p1 = reinterpret_cast<T*>(&v1[j]);
    for (int k = 0; k < D; ++k) {
      p0[k] = fa[j][k] = (T)a[j][k];
      p1[k] = fb[j][k] = (T)b[j][k];
    }
  }
  const T xsi = (T)eta;
  T rt = -1;
  const bool rh = mclo::additive_ccd<D, T>(fa, fb, (double)xsi, VF, rt);
  const T want = rh ? rt : T(-1);
  // uninterrupted
  T t = -1;
  int ev = 0;
  const int h = mclg::accd<D, VF, T>(v0, v1, xsi, t, ev);
  if ((h != 0) != rh || (h && !same_t(t, want))) return false;
  // suspended after three iterations, resumed on a fresh state
  mclg::Accd<D, VF, T> s;
  int st = s.template start<mclg::kAccdSuspendIter>(v0, v1, xsi);
  if (st == mclg::ACCD_RUN && s.replayable) {
    ++suspended;
    const double t1 = (double)s.t1, t2 = (double)s.t2, t3 = (double)s.tl, g = (double)s.g;
    mclg::Accd<D, VF, T> r;
    r.resume(v0, v1, xsi, (T)t1, (T)t2, (T)t3, (T)g);
    // the resumed state is the loop's state at the top of iteration 3, bit for bit
    if (std::memcmp(&r.st, &s.st, sizeof(s.st)) != 0 || !same_t(r.toi, s.toi) || !same_t(r.tl, s.tl) ||
        !same_t(r.g, s.g) || !same_t(r.l_p, s.l_p) || r.iter != s.iter)
      return false;
    st = r.run();
    const T got = st == mclg::ACCD_HIT ? r.toi : T(-1);
    if ((st == mclg::ACCD_HIT) != rh || !same_t(got, want)) return false;
    if (ev != s.evals + r.evals) return false;  // no evaluation repeated or lost
  } else {
    if (st == mclg::ACCD_RUN) st = s.run();
    const T got = st == mclg::ACCD_HIT ? s.toi : T(-1);
    if ((st == mclg::ACCD_HIT) != rh || !same_t(got, want)) return false;
  }
  return true;
}

template <int D, bool VF>
bool check_both(const double (*a)[D], const double (*b)[D], double eta, long long& suspended, long long& sus_f) {
  return check<D, VF, double>(a, b, eta, suspended) && check<D, VF, float>(a, b, eta, sus_f);
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
  long long failed = 0, suspended = 0, sus_f = 0;
  for (long long i = 0; i < n; ++i) {
    bool ok;
    if (dim == 3) {
      double a[4][3], b[4][3];
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 3; ++k) {
          a[j][k] = V0[(size_t)S[4 * i + j] * 3 + k];
          b[j][k] = V1[(size_t)S[4 * i + j] * 3 + k];
        }
      ok = vf[i] ? check_both<3, true>(a, b, eta, suspended, sus_f) : check_both<3, false>(a, b, eta, suspended, sus_f);
    } else {
      double a[3][2], b[3][2];
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 2; ++k) {
          a[j][k] = V0[(size_t)S[4 * i + j] * 2 + k];
          b[j][k] = V1[(size_t)S[4 * i + j] * 2 + k];
        }
      ok = check_both<2, true>(a, b, eta, suspended, sus_f);
    }
    if (!ok && failed++ < 5) std::printf("mismatch at stencil %lld\n", i);
  }
  std::printf("accd_resume_check: %lld stencils, %lld suspended, %lld failed, %lld suspended in float\n", n,
              suspended, failed, sus_f);
  return failed ? 1 : 0;
}
