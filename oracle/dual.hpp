// Forward-mode dual numbers for the oracle's reference TOI gradient
// (TEST INFRASTRUCTURE ONLY).  A Dual<N> carries a value and its derivatives
// with respect to N inputs; comparisons look at the value only, so the
// branch structure of the function is the one the plain FP64 code takes.
#pragma once
#include <cmath>

namespace mclo {

template <int N> struct Dual {
  double v = 0.0;
  double g[N] = {};
  Dual() = default;
  Dual(double x) : v(x) {}  // NOLINT: constants convert implicitly
};

template <typename T> struct Held;
template <int N> struct Held<Dual<N>> {
  static Dual<N> of(const Dual<N>& x) { return Dual<N>(x.v); }
};

template <int N> inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v + b.v);
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] + b.g[i];
  return r;
}
template <int N> inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v - b.v);
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] - b.g[i];
  return r;
}
template <int N> inline Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r(-a.v);
  for (int i = 0; i < N; ++i) r.g[i] = -a.g[i];
  return r;
}
template <int N> inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v * b.v);
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
  return r;
}
template <int N> inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v / b.v);
  for (int i = 0; i < N; ++i) r.g[i] = (a.g[i] - r.v * b.g[i]) / b.v;
  return r;
}
template <int N> inline Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> inline Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> inline Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> inline Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> inline Dual<N> operator*(const Dual<N>& a, double b) { return a * Dual<N>(b); }
template <int N> inline Dual<N> operator*(double a, const Dual<N>& b) { return Dual<N>(a) * b; }
template <int N> inline Dual<N> operator/(const Dual<N>& a, double b) { return a / Dual<N>(b); }
template <int N> inline Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }

#define MCLO_DUAL_CMP(op)                                                                                   \
  template <int N> inline bool operator op(const Dual<N>& a, const Dual<N>& b) { return a.v op b.v; }   \
  template <int N> inline bool operator op(const Dual<N>& a, double b) { return a.v op b; }             \
  template <int N> inline bool operator op(double a, const Dual<N>& b) { return a op b.v; }
MCLO_DUAL_CMP(<)
MCLO_DUAL_CMP(>)
MCLO_DUAL_CMP(<=)
MCLO_DUAL_CMP(>=)
MCLO_DUAL_CMP(==)
MCLO_DUAL_CMP(!=)
#undef MCLO_DUAL_CMP

template <int N> inline Dual<N> osqrt(const Dual<N>& a) {
  Dual<N> r(std::sqrt(a.v));
  const double k = r.v > 0 ? 0.5 / r.v : 0.0;
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] * k;
  return r;
}

}  // namespace mclo
