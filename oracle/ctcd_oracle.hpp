// CPU oracle, CTCD part — TEST INFRASTRUCTURE ONLY (included by mclccd_oracle.hpp).
//
// The narrow phase that BVHTree selects for vertex-face stencils when
// options.vf_one_sided is set (include/MCL/BVHTree.cpp:446-447): continuous
// time collision detection by polynomial intervals, then the wrong-side filter.
// Restated with no Eigen; reference paths (MIT, (c) Matt Overby 2021; the CTCD
// and Distance sources are Etienne Vouga's public-domain collisiondetection
// code; the root finder is C. Bond's C translation of Jenkins & Traub,
// ACM TOMS Algorithm 493):
//   query_ccd_vf (CTCD)    include/MCL/NarrowPhase.cpp:283-330 (2D), 332-419 (3D)
//   hit_wrong_side_vf      include/MCL/NarrowPhase.cpp:21-89
//   query_ray_box          include/MCL/NarrowPhase.cpp:122-171
//   TimeInterval           include/MCL/ccd_internal/CTCD.hpp:10-28, CTCD.cpp:17-43
//   getQuadRoots .. findIntervals  include/MCL/ccd_internal/CTCD.cpp:45-184
//   planePoly3D, distancePoly3D    include/MCL/ccd_internal/CTCD.cpp:222-264
//   vertexFaceCTCD         include/MCL/ccd_internal/CTCD.cpp:435-534
//   vertexEdgeCTCD (3D)    include/MCL/ccd_internal/CTCD.cpp:536-634
//   vertexEdgeCTCD (2D)    include/MCL/ccd_internal/CTCD.cpp:636-713
//   vertexVertexCTCD (3D)  include/MCL/ccd_internal/CTCD.cpp:715-817
//   vertexFaceDistance     include/MCL/ccd_internal/Distance.hpp:34-122
//   RootFinder::rpoly      include/MCL/ccd_internal/rpoly.h:56-762
//
// Arithmetic order follows the reference expressions as written (left to
// right; Eigen 3.4 fixed-size dot = (x0*y0 + x1*y1) + x2*y2; std::min/max as
// libstdc++ defines them; std::sort on at most 16 values is libstdc++'s
// insertion sort, which fixes where a NaN ends up).
//
// One deliberate difference: two loops of the root finder's bound estimate
// (rpoly.h:158-179) never end when a coefficient is NaN; here they stop after
// a cap that finite coefficients cannot reach (the first loop divides x by 10
// until it underflows, at most ~650 steps; Newton converges in a handful).
//
// Parity pins: tests/test_oracle_ctcd.py (roots against numpy, closed-form
// times of impact for vertex-face / vertex-edge / vertex-vertex approaches,
// wrong-side rejection, dense sampling of the vertex-face conditions).
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

namespace mclo {
namespace ctcd {

inline double mx(double a, double b) { return (a < b) ? b : a; }  // std::max
inline double mn(double a, double b) { return (b < a) ? b : a; }  // std::min

struct V3 { double x[3]; };
inline V3 sub(const double* a, const double* b) { return {{a[0] - b[0], a[1] - b[1], a[2] - b[2]}}; }
inline V3 sub(const V3& a, const V3& b) { return sub(a.x, b.x); }
inline V3 add(const V3& a, const V3& b) { return {{a.x[0] + b.x[0], a.x[1] + b.x[1], a.x[2] + b.x[2]}}; }
inline double dot(const V3& a, const V3& b) { return (a.x[0] * b.x[0] + a.x[1] * b.x[1]) + a.x[2] * b.x[2]; }
inline V3 cross(const V3& a, const V3& b) {
  return {{a.x[1] * b.x[2] - a.x[2] * b.x[1], a.x[2] * b.x[0] - a.x[0] * b.x[2], a.x[0] * b.x[1] - a.x[1] * b.x[0]}};
}
inline V3 load3(const double* p) { return {{p[0], p[1], p[2]}}; }

// CTCD.hpp:12-20 / CTCD.cpp:17-43
struct Interval {
  double l = 0, u = 0;
  Interval() = default;
  Interval(double tl, double tu) : l(tl), u(tu) {
    if (l > u) { double t = l; l = u; u = t; }
    l = mx(l, 0.0);
    u = mn(u, 1.0);
  }
};
inline bool overlap(const Interval& a, const Interval& b) { return !(a.l > b.u || b.l > a.u); }
inline bool overlap(const std::vector<Interval>& v) {
  for (size_t i = 0; i < v.size(); ++i)
    for (size_t j = i + 1; j < v.size(); ++j)
      if (!overlap(v[i], v[j])) return false;
  return true;
}
inline Interval intersect(const std::vector<Interval>& v) {
  Interval r(0.0, 1.0);
  for (const Interval& it : v) {
    r.l = mx(it.l, r.l);
    r.u = mn(it.u, r.u);
  }
  return r;
}

// libstdc++ std::sort for n <= 16: plain insertion sort with the unguarded inner loop.
inline void small_sort(double* a, int n) {
  for (int i = 1; i < n; ++i) {
    const double val = a[i];
    if (val < a[0]) {
      for (int j = i; j > 0; --j) a[j] = a[j - 1];
      a[0] = val;
    } else {
      int j = i;
      while (val < a[j - 1]) { a[j] = a[j - 1]; --j; }
      a[j] = val;
    }
  }
}

// ---------------------------------------------------------------------------
// Jenkins–Traub real-coefficient root finder (rpoly.h:56-762; TOMS 493).
// Coefficients in decreasing powers; returns the number of zeros found (real
// parts in zr, imaginary in zi) or -1.
// ---------------------------------------------------------------------------
class Rpoly {
 public:
  int solve(const double* op, int degree, double* zr, double* zi) {
    const double base = 2.0;
    eta = 2.22e-16;
    const double infin = 3.4e38, smalno = 1.2e-38;
    are = eta;
    mre = eta;
    const double lo = smalno / eta;
    double xx = std::sqrt(0.5), yy = -xx;
    double rot = 94.0;
    rot *= 0.017453293;
    const double cosr = std::cos(rot), sinr = std::sin(rot);
    n = degree;
    if (op[0] == 0.0) return -1;
    while (op[n] == 0.0) {
      const int j = degree - n;
      zr[j] = 0.0;
      zi[j] = 0.0;
      n--;
    }
    if (n < 1) return -1;
    double temp[7], pt[7];
    for (int i = 0; i <= n; i++) p[i] = op[i];
    for (;;) {  // rpoly.h:98 (label _40): one zero or pair of zeros per pass
      if (n == 1) {
        zr[degree - 1] = -p[1] / p[0];
        zi[degree - 1] = 0.0;
        n -= 1;
        break;
      }
      if (n == 2) {
        quad(p[0], p[1], p[2], &zr[degree - 2], &zi[degree - 2], &zr[degree - 1], &zi[degree - 1]);
        n -= 2;
        break;
      }
      double max = 0.0, min = infin;
      for (int i = 0; i <= n; i++) {
        const double x = std::fabs(p[i]);
        if (x > max) max = x;
        if (x != 0.0 && x < min) min = x;
      }
      double sc = lo / min;
      bool scale = true;
      if (sc > 1.0 && infin / sc < max) scale = false;
      if (scale && sc <= 1.0) {
        if (max < 10.0) scale = false;
        else if (sc == 0.0) sc = smalno;
      }
      if (scale) {
        const int l = (int)(std::log(sc) / std::log(base) + 0.5);
        const double factor = std::pow(base * 1.0, l);
        if (factor != 1.0)
          for (int i = 0; i <= n; i++) p[i] = factor * p[i];
      }
      // rpoly.h:135 (label _110): lower bound on the moduli of the zeros
      for (int i = 0; i <= n; i++) pt[i] = std::fabs(p[i]);
      pt[n] = -pt[n];
      double x = std::exp((std::log(-pt[n]) - std::log(pt[0])) / (double)n);
      if (pt[n - 1] != 0.0) {
        const double xm = -pt[n] / pt[n - 1];
        if (xm < x) x = xm;
      }
      for (int it = 0; it < kChopCap; ++it) {
        const double xm = x * 0.1;
        double ff = pt[0];
        for (int i = 1; i <= n; i++) ff = ff * xm + pt[i];
        if (ff <= 0.0) break;
        x = xm;
      }
      double dx = x;
      for (int it = 0; it < kNewtonCap && std::fabs(dx / x) > 0.005; ++it) {
        double ff = pt[0], df = ff;
        for (int i = 1; i < n; i++) {
          ff = ff * x + pt[i];
          df = df * x + ff;
        }
        ff = ff * x + pt[n];
        dx = ff / df;
        x -= dx;
      }
      const double bnd = x;
      // derivative as the initial K polynomial, five unshifted steps
      const int nm1 = n - 1;
      for (int i = 1; i < n; i++) k[i] = (double)(n - i) * p[i] / (double)n;
      k[0] = p[0];
      const double aa = p[n], bb = p[n - 1];
      bool zerok = (k[n - 1] == 0);
      for (int jj = 0; jj < 5; jj++) {
        const double cc = k[n - 1];
        if (!zerok) {
          const double t = -aa / cc;
          for (int i = 0; i < nm1; i++) {
            const int j = n - i - 1;
            k[j] = t * k[j - 1] + p[j];
          }
          k[0] = p[0];
          zerok = (std::fabs(k[n - 1]) <= std::fabs(bb) * eta * 10.0);
        } else {
          for (int i = 0; i < nm1; i++) {
            const int j = n - i - 1;
            k[j] = k[j - 1];
          }
          k[0] = 0.0;
          zerok = (k[n - 1] == 0.0);
        }
      }
      for (int i = 0; i < n; i++) temp[i] = k[i];
      bool found = false;
      for (int cnt = 0; cnt < 20; cnt++) {
        const double xxx = cosr * xx - sinr * yy;
        yy = sinr * xx + cosr * yy;
        xx = xxx;
        sr = bnd * xx;
        si = bnd * yy;
        u = -2.0 * sr;
        v = bnd;
        int nz;
        fxshfr(20 * (cnt + 1), &nz);
        if (nz != 0) {
          const int j = degree - n;
          zr[j] = szr;
          zi[j] = szi;
          n -= nz;
          for (int i = 0; i <= n; i++) p[i] = qp[i];
          if (nz != 1) {
            zr[j + 1] = lzr;
            zi[j + 1] = lzi;
          }
          found = true;
          break;
        }
        for (int i = 0; i < n; i++) k[i] = temp[i];
      }
      if (!found) break;  // no convergence with 20 shifts
    }
    return degree - n;
  }

  static constexpr int kChopCap = 2000, kNewtonCap = 1000;

 private:
  double p[7], qp[7], k[7], qk[7], svk[7];
  double sr, si, u, v, a, b, c, d, a1, a3, a7, e, f, g, h, szr, szi, lzr, lzi;
  double eta, are, mre;
  int n;

  // rpoly.h:284-390 — fixed-shift stage; the goto labels become a small state machine.
  void fxshfr(int l2, int* nz) {
    double ui, vi, s = 0, svu = 0, svv = 0;
    double ots = 0, otv = 0;
    int type;
    *nz = 0;
    double betav = 0.25, betas = 0.25;
    double oss = sr, ovv = v;
    quadsd(n, &u, &v, p, qp, &a, &b);
    calcsc(&type);
    for (int j = 0; j < l2; j++) {
      nextk(&type);
      calcsc(&type);
      newest(type, &ui, &vi);
      const double vv = vi;
      double ss = 0.0;
      if (k[n - 1] != 0.0) ss = -p[n] / k[n - 1];
      double tv = 1.0, ts = 1.0;
      if (!(j == 0 || type == 3)) {
        if (vv != 0.0) tv = std::fabs((vv - ovv) / vv);
        if (ss != 0.0) ts = std::fabs((ss - oss) / ss);
        double tvv = 1.0;
        if (tv < otv) tvv = tv * otv;
        double tss = 1.0;
        if (ts < ots) tss = ts * ots;
        const bool vpass = (tvv < betav), spass = (tss < betas);
        if (spass || vpass) {
          svu = u;
          svv = v;
          for (int i = 0; i < n; i++) svk[i] = k[i];
          s = ss;
          bool vtry = false, stry = false;
          enum { QUAD, LIN, RESTORE } st = ((spass && (!vpass)) || tss < tvv) ? LIN : QUAD;
          for (;;) {
            if (st == QUAD) {  // label _20
              quadit(&ui, &vi, nz);
              if (*nz > 0) return;
              vtry = true;
              betav *= 0.25;
              if (stry || !spass) { st = RESTORE; continue; }
              for (int i = 0; i < n; i++) k[i] = svk[i];
              st = LIN;
              continue;
            }
            if (st == LIN) {  // label _40
              int iflag;
              realit(&s, nz, &iflag);
              if (*nz > 0) return;
              stry = true;
              betas *= 0.25;
              if (iflag == 0) { st = RESTORE; continue; }
              ui = -(s + s);
              vi = s * s;
              st = QUAD;
              continue;
            }
            // label _50
            u = svu;
            v = svv;
            for (int i = 0; i < n; i++) k[i] = svk[i];
            if (vpass && !vtry) { st = QUAD; continue; }
            quadsd(n, &u, &v, p, qp, &a, &b);
            calcsc(&type);
            break;
          }
        }
      }
      // label _70
      ovv = vv;
      oss = ss;
      otv = tv;
      ots = ts;
    }
  }

  // rpoly.h:397-474 — variable-shift iteration for a quadratic factor.
  void quadit(double* uu, double* vv, int* nz) {
    double ui, vi, mp, omp = 0, ee, relstp = 0, t, zm;
    int type;
    bool tried = false;
    *nz = 0;
    u = *uu;
    v = *vv;
    int j = 0;
    for (;;) {
      quad(1.0, u, v, &szr, &szi, &lzr, &lzi);
      if (std::fabs(std::fabs(szr) - std::fabs(lzr)) > 0.01 * std::fabs(lzr)) return;
      quadsd(n, &u, &v, p, qp, &a, &b);
      mp = std::fabs(a - szr * b) + std::fabs(szi * b);
      zm = std::sqrt(std::fabs(v));
      ee = 2.0 * std::fabs(qp[0]);
      t = -szr * b;
      for (int i = 1; i < n; i++) ee = ee * zm + std::fabs(qp[i]);
      ee = ee * zm + std::fabs(a + t);
      ee *= (5.0 * mre + 4.0 * are);
      ee = ee - (5.0 * mre + 2.0 * are) * (std::fabs(a + t) + std::fabs(b) * zm) + 2.0 * are * std::fabs(t);
      if (mp <= 20.0 * ee) {
        *nz = 2;
        return;
      }
      j++;
      if (j > 20) return;
      if (!(j < 2) && !(relstp > 0.01 || mp < omp || tried)) {
        // a cluster stalls convergence: five fixed-shift steps near it
        if (relstp < eta) relstp = eta;
        relstp = std::sqrt(relstp);
        u = u - u * relstp;
        v = v + v * relstp;
        quadsd(n, &u, &v, p, qp, &a, &b);
        for (int i = 0; i < 5; i++) {
          calcsc(&type);
          nextk(&type);
        }
        tried = true;
        j = 0;
      }
      // label _50
      omp = mp;
      calcsc(&type);
      nextk(&type);
      calcsc(&type);
      newest(type, &ui, &vi);
      if (vi == 0.0) return;
      relstp = std::fabs((vi - v) / vi);
      u = ui;
      v = vi;
    }
  }

  // rpoly.h:480-566 — variable-shift iteration for a real zero.
  void realit(double* sss, int* nz, int* iflag) {
    double pv, kv, t = 0, s, ms, mp, omp = 0, ee;
    *nz = 0;
    s = *sss;
    *iflag = 0;
    int j = 0;
    for (;;) {
      pv = p[0];
      qp[0] = pv;
      for (int i = 1; i <= n; i++) {
        pv = pv * s + p[i];
        qp[i] = pv;
      }
      mp = std::fabs(pv);
      ms = std::fabs(s);
      ee = (mre / (are + mre)) * std::fabs(qp[0]);
      for (int i = 1; i <= n; i++) ee = ee * ms + std::fabs(qp[i]);
      if (mp <= 20.0 * ((are + mre) * ee - mre * mp)) {
        *nz = 1;
        szr = s;
        szi = 0.0;
        return;
      }
      j++;
      if (j > 10) return;
      if (!(j < 2) && !(std::fabs(t) > 0.001 * std::fabs(s - t) || mp < omp)) {
        *iflag = 1;  // a cluster near the real axis: try a quadratic iteration
        *sss = s;
        return;
      }
      omp = mp;
      kv = k[0];
      qk[0] = kv;
      for (int i = 1; i < n; i++) {
        kv = kv * s + k[i];
        qk[i] = kv;
      }
      if (std::fabs(kv) <= std::fabs(k[n - 1]) * 10.0 * eta) {
        k[0] = 0.0;
        for (int i = 1; i < n; i++) k[i] = qk[i - 1];
      } else {
        t = -pv / kv;
        k[0] = qp[0];
        for (int i = 1; i < n; i++) k[i] = t * qk[i - 1] + qp[i];
      }
      kv = k[0];
      for (int i = 1; i < n; i++) kv = kv * s + k[i];
      t = 0.0;
      if (std::fabs(kv) > std::fabs(k[n - 1] * 10.0 * eta)) t = -pv / kv;
      s += t;
    }
  }

  // rpoly.h:573-603
  void calcsc(int* type) {
    quadsd(n - 1, &u, &v, k, qk, &c, &d);
    if (!(std::fabs(c) > std::fabs(k[n - 1] * 100.0 * eta)) && !(std::fabs(d) > std::fabs(k[n - 2] * 100.0 * eta))) {
      *type = 3;  // the quadratic is almost a factor of K
      return;
    }
    if (std::fabs(d) < std::fabs(c)) {
      *type = 1;
      e = a / c;
      f = d / c;
      g = u * e;
      h = v * b;
      a3 = a * e + (h / c + g) * b;
      a1 = b - a * (d / c);
      a7 = a + g * d + h * f;
      return;
    }
    *type = 2;
    e = a / d;
    f = c / d;
    g = u * b;
    h = v * b;
    a3 = (a + g) * e + h * (b / d);
    a1 = b * f - a;
    a7 = (f + u) * a + h;
  }

  // rpoly.h:608-641
  void nextk(int* type) {
    if (*type == 3) {
      k[0] = 0.0;
      k[1] = 0.0;
      for (int i = 2; i < n; i++) k[i] = qk[i - 2];
      return;
    }
    double temp = a;
    if (*type == 1) temp = b;
    if (std::fabs(a1) <= std::fabs(temp) * eta * 10.0) {
      k[0] = 0.0;
      k[1] = -a7 * qp[0];
      for (int i = 2; i < n; i++) k[i] = a3 * qk[i - 2] - a7 * qp[i - 1];
      return;
    }
    a7 /= a1;
    a3 /= a1;
    k[0] = qp[0];
    k[1] = qp[1] - a7 * qp[0];
    for (int i = 2; i < n; i++) k[i] = a3 * qk[i - 2] - a7 * qp[i - 1] + qp[i];
  }

  // rpoly.h:647-682
  void newest(int type, double* uu, double* vv) {
    double a4, a5;
    if (type == 3) {
      *uu = 0.0;
      *vv = 0.0;
      return;
    }
    if (type == 2) {
      a4 = (a + g) * f + h;
      a5 = (f + u) * c + v * d;
    } else {
      a4 = a + u * b + h * f;
      a5 = c + (u + v * f) * d;
    }
    const double b1 = -k[n - 1] / p[n];
    const double b2 = -(k[n - 2] + b1 * p[n - 1]) / p[n];
    const double c1 = v * b2 * a1;
    const double c2 = b1 * a7;
    const double c3 = b1 * b1 * a3;
    const double c4 = c1 - c2 - c3;
    const double temp = a5 + b1 * a4 - c4;
    if (temp == 0.0) {
      *uu = 0.0;
      *vv = 0.0;
      return;
    }
    *uu = u - (u * (c3 + c2) + v * (b1 * a1 + b2 * a7)) / temp;
    *vv = v * (1.0 + c4 / temp);
  }

  // rpoly.h:687-700 — divide p by 1 + u z + v z^2 (quotient q, remainder a, b)
  static void quadsd(int nn, const double* u_, const double* v_, const double* pp, double* q, double* a_, double* b_) {
    *b_ = pp[0];
    q[0] = *b_;
    *a_ = pp[1] - (*b_) * (*u_);
    q[1] = *a_;
    for (int i = 2; i <= nn; i++) {
      const double c_ = pp[i] - (*a_) * (*u_) - (*b_) * (*v_);
      q[i] = c_;
      *b_ = *a_;
      *a_ = c_;
    }
  }

 public:
  // rpoly.h:706-760 — zeros of a z^2 + b1 z + c
  static void quad(double a_, double b1, double c_, double* sr_, double* si_, double* lr, double* li) {
    if (a_ == 0.0) {
      *sr_ = (b1 != 0.0) ? -c_ / b1 : 0.0;
      *lr = 0.0;
      *si_ = 0.0;
      *li = 0.0;
      return;
    }
    if (c_ == 0.0) {
      *sr_ = 0.0;
      *lr = -b1 / a_;
      *si_ = 0.0;
      *li = 0.0;
      return;
    }
    double b_ = b1 / 2.0, d_, e_;
    if (std::fabs(b_) < std::fabs(c_)) {
      e_ = (c_ < 0.0) ? -a_ : a_;
      e_ = b_ * (b_ / std::fabs(c_)) - e_;
      d_ = std::sqrt(std::fabs(e_)) * std::sqrt(std::fabs(c_));
    } else {
      e_ = 1.0 - (a_ / b_) * (c_ / b_);
      d_ = std::sqrt(std::fabs(e_)) * std::fabs(b_);
    }
    if (e_ < 0.0) {
      *sr_ = -b_ / a_;
      *lr = *sr_;
      *si_ = std::fabs(d_ / a_);
      *li = -(*si_);
    } else {
      if (b_ >= 0.0) d_ = -d_;
      *lr = (-b_ + d_) / a_;
      *sr_ = 0.0;
      if (*lr != 0.0) *sr_ = (c_ / *lr) / a_;
      *si_ = 0.0;
      *li = 0.0;
    }
  }
};

// ---------------------------------------------------------------------------
// CTCD.cpp:45-184
// ---------------------------------------------------------------------------
inline int quad_roots(double a, double b, double c, double& t0, double& t1) {
  int roots = 0;
  const int sign = (b < 0) ? -1 : 1;
  const double D = b * b - 4 * a * c;
  if (D >= 0) {
    roots = 2;
    const double q = -0.5 * (b + sign * std::sqrt(D));
    t0 = q / a;
    t1 = c / q;
    if (t0 > t1) { const double t = t0; t0 = t1; t1 = t; }
  }
  return roots;
}

inline void check_interval(double t1, double t2, const double* op, int degree, std::vector<Interval>& out, bool pos) {
  t1 = mx(0.0, t1);
  t2 = mx(0.0, t2);
  t1 = mn(1.0, t1);
  t2 = mn(1.0, t2);
  const double tmid = (t2 + t1) / 2;
  double f = op[0];
  for (int i = 1; i <= degree; i++) {
    f *= tmid;
    f += op[i];
  }
  if (pos && f >= 0) out.emplace_back(t1, t2);
  else if (!pos && f <= 0) out.emplace_back(t1, t2);
}

inline bool could_have_roots(const double* op, int degree, bool pos) {
  double result = 0;
  if ((pos && op[0] > 0) || (!pos && op[0] < 0)) result = op[0];
  for (int i = 1; i < degree; i++) {
    result *= 1.0;
    if ((pos && op[i] > 0) || (!pos && op[i] < 0)) result += op[i];
  }
  result *= 1.0;
  result += op[degree];
  return !((pos && result < 0) || (!pos && result > 0));
}

// op: degree n+1 coefficients, decreasing powers; modified in place
// (normalised, leading zeros dropped) as the reference does.
inline void find_intervals(double* op, int n, std::vector<Interval>& out, bool pos) {
  if (n > 6) return;
  int roots = 0, rd = n;
  double time[6], zeroi[6];
  double maxval = 0;
  for (int i = 0; i <= n; i++) maxval = mx(maxval, std::fabs(op[i]));
  if (maxval != 0)
    for (int i = 0; i <= n; i++) op[i] /= maxval;
  for (int i = 0; i < n; i++) {
    if (op[i] == 0) rd--;
    else break;
  }
  if (rd < n)
    for (int i = 0; i <= rd; i++) op[i] = op[i + n - rd];
  if (rd > 2) {
    if (!could_have_roots(op, rd, pos)) return;
    Rpoly rf;
    roots = rf.solve(op, rd, time, zeroi);
  } else if (rd == 2) {
    roots = quad_roots(op[0], op[1], op[2], time[0], time[1]);
  } else if (rd == 1) {
    time[0] = -op[1] / op[0];
    roots = 1;
  } else {
    if ((!pos && op[0] <= 0) || (pos && op[0] >= 0)) out.emplace_back(0, 1.0);
    return;
  }
  if (roots > 0) {
    small_sort(time, roots);
    if (time[0] >= 0) check_interval(0, time[0], op, rd, out, pos);
    for (int i = 0; i < roots - 1; i++)
      if (!((time[i] < 0 && time[i + 1] < 0) || (time[i] > 1.0 && time[i + 1] > 1.0)))
        check_interval(time[i], time[i + 1], op, rd, out, pos);
    if (time[roots - 1] <= 1.0) check_interval(time[roots - 1], 1.0, op, rd, out, pos);
  } else {
    check_interval(0.0, 1.0, op, rd, out, pos);
  }
}

// CTCD.cpp:222-236
inline void plane_poly(const V3& x10, const V3& x20, const V3& x30, const V3& v10, const V3& v20, const V3& v30,
                       std::vector<Interval>& out) {
  double op[4];
  op[0] = dot(v10, cross(v20, v30));
  op[1] = dot(x10, cross(v20, v30)) + dot(v10, cross(x20, v30)) + dot(v10, cross(v20, x30));
  op[2] = dot(x10, cross(x20, v30)) + dot(x10, cross(v20, x30)) + dot(v10, cross(x20, x30));
  op[3] = dot(x10, cross(x20, x30));
  find_intervals(op, 3, out, true);
}

// CTCD.cpp:238-264
inline void distance_poly(const V3& x10, const V3& x20, const V3& x30, const V3& v10, const V3& v20, const V3& v30,
                          double minD2, std::vector<Interval>& out) {
  const double A = dot(v10, cross(v20, v30));
  const double B = dot(x10, cross(v20, v30)) + dot(v10, cross(x20, v30)) + dot(v10, cross(v20, x30));
  const double C = dot(x10, cross(x20, v30)) + dot(x10, cross(v20, x30)) + dot(v10, cross(x20, x30));
  const double D = dot(x10, cross(x20, x30));
  const V3 E = cross(x20, x30);
  const V3 F = add(cross(x20, v30), cross(v20, x30));
  const V3 G = cross(v20, v30);
  double op[7];
  op[0] = A * A;
  op[1] = 2 * A * B;
  op[2] = B * B + 2 * A * C - dot(G, G) * minD2;
  op[3] = 2 * A * D + 2 * B * C - 2 * dot(G, F) * minD2;
  op[4] = 2 * B * D + C * C - (2 * dot(G, E) + dot(F, F)) * minD2;
  op[5] = 2 * C * D - 2 * dot(F, E) * minD2;
  op[6] = D * D - dot(E, E) * minD2;
  find_intervals(op, 6, out, false);
}

// Each sub-test appends its candidate times to all_t only when it reports a
// hit (so the caller's list holds the times of the first sub-test that hits).

// CTCD.cpp:435-534
inline bool vertex_face(const double* q0s, const double* q1s, const double* q2s, const double* q3s, const double* q0e,
                        const double* q1e, const double* q2e, const double* q3e, double eta, std::vector<double>& all_t) {
  const double minD = eta * eta;
  const V3 v0 = sub(q0e, q0s), v1 = sub(q1e, q1s), v2 = sub(q2e, q2s), v3 = sub(q3e, q3s);
  std::vector<Interval> coplane, e1, e2, e3;
  plane_poly(sub(q0s, q1s), cross(sub(q3s, q1s), sub(q2s, q1s)), sub(q3s, q1s), sub(v0, v1),
             cross(sub(v3, v1), sub(v2, v1)), sub(v3, v1), e1);
  if (e1.empty()) return false;
  plane_poly(sub(q0s, q2s), cross(sub(q1s, q2s), sub(q3s, q2s)), sub(q1s, q2s), sub(v0, v2),
             cross(sub(v1, v2), sub(v3, v2)), sub(v1, v2), e2);
  if (e2.empty()) return false;
  plane_poly(sub(q0s, q3s), cross(sub(q2s, q3s), sub(q1s, q3s)), sub(q2s, q3s), sub(v0, v3),
             cross(sub(v2, v3), sub(v1, v3)), sub(v2, v3), e3);
  if (e3.empty()) return false;
  distance_poly(sub(q0s, q1s), sub(q2s, q1s), sub(q3s, q1s), sub(v0, v1), sub(v2, v1), sub(v3, v1), minD, coplane);
  if (coplane.empty()) return false;
  bool col = false;
  for (const Interval& i0 : coplane)
    for (const Interval& i1 : e1)
      for (const Interval& i2 : e2)
        for (const Interval& i3 : e3) {
          const std::vector<Interval> iv = {i0, i1, i2, i3};
          if (overlap(iv)) {
            all_t.push_back(intersect(iv).l);
            col = true;
          }
        }
  return col;
}

// CTCD.cpp:536-634
inline bool vertex_edge3(const double* q0s, const double* q1s, const double* q2s, const double* q0e, const double* q1e,
                         const double* q2e, double eta, std::vector<double>& all_t) {
  double op[5];
  const double minD = eta * eta;
  const V3 v0 = sub(q0e, q0s), v1 = sub(q1e, q1s), v2 = sub(q2e, q2s);
  std::vector<Interval> colin, e1, e2;
  const V3 ab = sub(q2s, q1s), ac = sub(q0s, q1s), cb = sub(q2s, q0s);
  const V3 vab = sub(v2, v1), vac = sub(v0, v1), vcb = sub(v2, v0);
  double c = dot(ab, ac), b = dot(ac, vab) + dot(ab, vac), a = dot(vab, vac);
  op[0] = a; op[1] = b; op[2] = c;
  find_intervals(op, 2, e1, true);
  if (e1.empty()) return false;
  c = dot(ab, cb);
  b = dot(cb, vab) + dot(ab, vcb);
  a = dot(vab, vcb);
  op[0] = a; op[1] = b; op[2] = c;
  find_intervals(op, 2, e2, true);
  if (e2.empty()) return false;
  const double A = dot(ab, ab), B = 2 * dot(ab, vab), C = dot(vab, vab);
  const double D = dot(ac, ac), E = 2 * dot(ac, vac), F = dot(vac, vac);
  const double G = dot(ac, ab), H = dot(vab, ac) + dot(vac, ab), I = dot(vab, vac);
  op[4] = A * D - G * G - minD * A;
  op[3] = B * D + A * E - 2 * G * H - minD * B;
  op[2] = B * E + A * F + C * D - H * H - 2 * G * I - minD * C;
  op[1] = B * F + C * E - 2 * H * I;
  op[0] = C * F - I * I;
  find_intervals(op, 4, colin, false);
  if (colin.empty()) return false;
  bool col = false;
  for (const Interval& i0 : colin)
    for (const Interval& i1 : e1)
      for (const Interval& i2 : e2) {
        const std::vector<Interval> iv = {i0, i1, i2};
        if (overlap(iv)) {
          all_t.push_back(intersect(iv).l);
          col = true;
        }
      }
  return col;
}

// CTCD.cpp:715-817 (the all_t != nullptr branches)
inline bool vertex_vertex3(const double* q1s, const double* q2s, const double* q1e, const double* q2e, double eta,
                           std::vector<double>& all_t) {
  int roots = 0;
  const double min_d = eta * eta;
  double t1 = 0, t2 = 0;
  const V3 v1 = sub(q1e, q1s), v2 = sub(q2e, q2s);
  const V3 p1 = load3(q1s), p2 = load3(q2s);
  const double a = dot(v1, v1) + dot(v2, v2) - 2 * dot(v1, v2);
  const double b = 2 * (dot(v1, p1) - dot(v2, p1) - dot(v1, p2) + dot(v2, p2));
  const double c = dot(p1, p1) + dot(p2, p2) - 2 * dot(p1, p2) - min_d;
  if (a != 0) {
    roots = quad_roots(a, b, c, t1, t2);
  } else if (b != 0) {
    t1 = -c / b;
    roots = 1;
  } else {
    if (c <= 0) {
      all_t.push_back(0.0);
      return true;
    }
    return false;
  }
  const double op[3] = {a, b, c};
  std::vector<Interval> iv;  // accumulates across the checks, as in the reference
  if (roots == 2) {
    double curr_t = 1.0;
    check_interval(0, t1, op, 2, iv, false);
    if (!iv.empty()) {
      curr_t = 0;
      all_t.push_back(curr_t);
    }
    check_interval(t1, t2, op, 2, iv, false);
    if (!iv.empty()) {
      curr_t = mn(curr_t, t1);
      all_t.push_back(curr_t);
    }
    check_interval(t2, 1.0, op, 2, iv, false);
    if (!iv.empty()) {
      curr_t = mn(t2, curr_t);
      all_t.push_back(curr_t);
    }
    return !all_t.empty();
  } else if (roots == 1) {
    check_interval(0, t1, op, 2, iv, false);
    if (!iv.empty()) {
      all_t.push_back(0.0);
      return true;
    }
    check_interval(t1, 1.0, op, 2, iv, false);
    if (!iv.empty()) {
      all_t.push_back(t1);
      return true;
    }
    return false;
  }
  check_interval(0, 1.0, op, 2, iv, false);
  if (!iv.empty()) {
    all_t.push_back(0.0);
    return true;
  }
  return false;
}

// CTCD.cpp:636-713
inline bool vertex_edge2(const double* q0s, const double* q1s, const double* q2s, const double* q0e, const double* q1e,
                         const double* q2e, std::vector<double>& all_t) {
  auto norm2d = [](double x, double y) { return std::sqrt(x * x + y * y); };
  auto on_line = [&](const double* p, const double* a, const double* b) {
    const double len = norm2d(a[0] - b[0], a[1] - b[1]);
    if (norm2d(p[0] - a[0], p[1] - a[1]) > len) return false;
    if (norm2d(p[0] - b[0], p[1] - b[1]) > len) return false;
    return true;
  };
  const double u0[2] = {q1s[0] - q0s[0], q1s[1] - q0s[1]}, v0[2] = {q2s[0] - q0s[0], q2s[1] - q0s[1]};
  const double u1[2] = {q1e[0] - q0e[0], q1e[1] - q0e[1]}, v1[2] = {q2e[0] - q0e[0], q2e[1] - q0e[1]};
  const double up[2] = {u1[0] - u0[0], u1[1] - u0[1]}, vp[2] = {v1[0] - v0[0], v1[1] - v0[1]};
  // 2x2 determinants with columns (x, y): x0*y1 - x1*y0
  const double c = u0[0] * v0[1] - u0[1] * v0[0];
  const double b = (u0[0] * vp[1] - u0[1] * vp[0]) + (up[0] * v0[1] - up[1] * v0[0]);
  const double a = up[0] * vp[1] - up[1] * vp[0];
  double t0 = 2, t1 = 2;
  if (quad_roots(a, b, c, t0, t1) == 0) return false;
  auto at = [](const double* s, const double* e, double t, double* o) {
    o[0] = s[0] * (1.0 - t) + e[0] * t;
    o[1] = s[1] * (1.0 - t) + e[1] * t;
  };
  double p[2], a0[2], b0[2];
  at(q0s, q0e, t0, p); at(q1s, q1e, t0, a0); at(q2s, q2e, t0, b0);
  const bool on0 = on_line(p, a0, b0);
  at(q0s, q0e, t1, p); at(q1s, q1e, t1, a0); at(q2s, q2e, t1, b0);
  const bool on1 = on_line(p, a0, b0);
  bool hit = false;
  if (t0 >= 0 && t0 <= 1 && on0) { all_t.push_back(t0); hit = true; }
  if (t1 >= 0 && t1 <= 1 && on1) { all_t.push_back(t1); hit = true; }
  return hit;
}

// Distance.hpp:34-122 — barycentric coordinates of the closest point (clamped).
inline void vertex_face_barys(const V3& p, const V3& q0, const V3& q1, const V3& q2, double* bary) {
  const V3 ab = sub(q1, q0), ac = sub(q2, q0), ap = sub(p, q0);
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0 && d2 <= 0) { bary[0] = 1.0; bary[1] = 0.0; bary[2] = 0.0; return; }
  const V3 bp = sub(p, q1);
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0 && d4 <= d3) { bary[0] = 0.0; bary[1] = 1.0; bary[2] = 0.0; return; }
  const double vc = d1 * d4 - d3 * d2;
  if ((vc <= 0) && (d1 >= 0) && (d3 <= 0)) {
    const double v = d1 / (d1 - d3);
    bary[0] = 1.0 - v; bary[1] = v; bary[2] = 0;
    return;
  }
  const V3 cp = sub(p, q2);
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0 && d5 <= d6) { bary[0] = 0; bary[1] = 0; bary[2] = 1.0; return; }
  const double vb = d5 * d2 - d1 * d6;
  if ((vb <= 0) && (d2 >= 0) && (d6 <= 0)) {
    const double w = d2 / (d2 - d6);
    bary[0] = 1 - w; bary[1] = 0; bary[2] = w;
    return;
  }
  const double va = d3 * d6 - d5 * d4;
  if ((va <= 0) && (d4 - d3 >= 0) && (d5 - d6 >= 0)) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    bary[0] = 0; bary[1] = 1.0 - w; bary[2] = w;
    return;
  }
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom, w = vc * denom;
  const double u = 1.0 - v - w;
  auto clamp01 = [](double x) { return (x < 0.0) ? 0.0 : (1.0 < x) ? 1.0 : x; };  // std::clamp
  bary[0] = clamp01(u); bary[1] = clamp01(v); bary[2] = clamp01(w);
}

// Eigen 3.4 stableNormalize: w = max |x_i|, z = |x/w|^2, x /= sqrt(z)*w when z > 0.
template <int D> inline void stable_normalize(double* x) {
  double w = std::fabs(x[0]);
  for (int i = 1; i < D; ++i) w = mx(w, std::fabs(x[i]));
  double y[D];
  for (int i = 0; i < D; ++i) y[i] = x[i] / w;
  const double z = (D == 3) ? (y[0] * y[0] + y[1] * y[1]) + y[2] * y[2] : y[0] * y[0] + y[1] * y[1];
  if (z > 0) {
    const double s = std::sqrt(z) * w;
    for (int i = 0; i < D; ++i) x[i] = x[i] / s;
  }
}

// NarrowPhase.cpp:58-89
inline bool wrong_side3(const double (*v0)[3], const double (*v1)[3], double t) {
  if (std::fabs(t) <= 0.0) return false;
  V3 xt[4];
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) xt[i].x[k] = v0[i][k] * (1.0 - t) + v1[i][k] * t;
  V3 n = cross(sub(xt[2], xt[1]), sub(xt[3], xt[1]));
  if (dot(n, n) <= 0) return false;
  stable_normalize<3>(n.x);
  double bary[3];
  vertex_face_barys(xt[0], xt[1], xt[2], xt[3], bary);
  V3 vel;
  for (int k = 0; k < 3; ++k) {
    const double apex = v1[0][k] - v0[0][k];
    const double f0 = (bary[0] * v0[1][k] + bary[1] * v0[2][k]) + bary[2] * v0[3][k];
    const double f1 = (bary[0] * v1[1][k] + bary[1] * v1[2][k]) + bary[2] * v1[3][k];
    vel.x[k] = apex - (f1 - f0);
  }
  return dot(n, vel) > 0.0;
}

// NarrowPhase.cpp:21-55
inline bool wrong_side2(const double (*v0)[2], const double (*v1)[2], double t) {
  if (std::fabs(t) <= 0.0) return false;
  double p[2], e0[2], e1[2];
  for (int k = 0; k < 2; ++k) {
    p[k] = v0[0][k] * (1.0 - t) + v1[0][k] * t;
    e0[k] = v0[1][k] * (1.0 - t) + v1[1][k] * t;
    e1[k] = v0[2][k] * (1.0 - t) + v1[2][k] * t;
  }
  double n[2] = {e1[1] - e0[1], -(e1[0] - e0[0])};
  stable_normalize<2>(n);
  auto norm2d = [](double x, double y) { return std::sqrt(x * x + y * y); };
  const double denom = norm2d(e0[0] - e1[0], e0[1] - e1[1]);
  if (denom < DBL_EPSILON) return false;
  const double b0 = norm2d(e1[0] - p[0], e1[1] - p[1]) / denom;
  const double b1 = norm2d(e0[0] - p[0], e0[1] - p[1]) / denom;
  double vel[2];
  for (int k = 0; k < 2; ++k) {
    const double e0b = b0 * v0[1][k] + b1 * v0[2][k];
    const double e1b = b0 * v1[1][k] + b1 * v1[2][k];
    vel[k] = (v1[0][k] - v0[0][k]) - (e1b - e0b);
  }
  return n[0] * vel[0] + n[1] * vel[1] > 0.0;
}

// NarrowPhase.cpp:122-161 — ray (p0 -> p1) against the box [bmin, bmax].
inline bool ray_box(const double* p0, const double* p1, const double* bmin, const double* bmax) {
  V3 dir = sub(p1, p0);
  const double t0 = 0;
  const double t1 = std::sqrt(dot(dir, dir));
  auto contains = [&](const double* q) {
    for (int k = 0; k < 3; ++k)
      if (!(bmin[k] <= q[k])) return false;
    for (int k = 0; k < 3; ++k)
      if (!(q[k] <= bmax[k])) return false;
    return true;
  };
  if (contains(p0) || contains(p1)) return true;
  const double z = dot(dir, dir);
  if (z > 0) {
    const double s = std::sqrt(z);
    for (int k = 0; k < 3; ++k) dir.x[k] = dir.x[k] / s;
  }
  double inv[3];
  bool sign[3];
  for (int k = 0; k < 3; ++k) inv[k] = 1. / dir.x[k];
  for (int k = 0; k < 3; ++k) sign[k] = inv[k] < 0;
  const double* bounds[2] = {bmin, bmax};
  double tmin = (bounds[sign[0]][0] - p0[0]) * inv[0];
  double tmax = (bounds[1 - sign[0]][0] - p0[0]) * inv[0];
  const double tymin = (bounds[sign[1]][1] - p0[1]) * inv[1];
  const double tymax = (bounds[1 - sign[1]][1] - p0[1]) * inv[1];
  if ((tmin > tymax) || (tymin > tmax)) return false;
  if (tymin > tmin) tmin = tymin;
  if (tymax < tmax) tmax = tymax;
  const double tzmin = (bounds[sign[2]][2] - p0[2]) * inv[2];
  const double tzmax = (bounds[1 - sign[2]][2] - p0[2]) * inv[2];
  if ((tmin > tzmax) || (tzmin > tmax)) return false;
  if (tzmin > tmin) tmin = tzmin;
  if (tzmax < tmax) tmax = tzmax;
  if (!((tmin < t1) && (tmax > t0))) return false;
  return true;
}

// NarrowPhase.cpp:283-419 with test_wrong_side = true (BVHTree passes
// vf_one_sided for it): 1 = hit at toi, 0 = no hit, -1 = only wrong-side hits.
template <int D> inline int query_vf(const double (*v0)[D], const double (*v1)[D], double eta, double& toi) {
  std::vector<double> all_t;
  if constexpr (D == 2) {
    if (!vertex_edge2(v0[0], v0[1], v0[2], v1[0], v1[1], v1[2], all_t)) return 0;
    // all_t holds at most two values: the insertion sort fixes NaN placement
    small_sort(all_t.data(), (int)all_t.size());
    for (double t : all_t)
      if (!wrong_side2(v0, v1, t)) { toi = t; return 1; }
    return -1;
  } else {
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) { lo[k] = DBL_MAX; hi[k] = -DBL_MAX; }
    for (int i = 1; i <= 3; ++i)
      for (const double* q : {v0[i], v1[i]})
        for (int k = 0; k < 3; ++k) { lo[k] = mn(lo[k], q[k]); hi[k] = mx(hi[k], q[k]); }
    for (int k = 0; k < 3; ++k) { lo[k] -= eta; hi[k] += eta; }
    if (!ray_box(v0[0], v1[0], lo, hi)) return 0;
    bool hit = vertex_face(v0[0], v0[1], v0[2], v0[3], v1[0], v1[1], v1[2], v1[3], eta, all_t);
    for (int e = 0; e < 3 && !hit; ++e) {
      const int a = 1 + e, b = 1 + (e + 1) % 3;
      hit = vertex_edge3(v0[0], v0[a], v0[b], v1[0], v1[a], v1[b], eta, all_t);
    }
    for (int w = 0; w < 3 && !hit; ++w) hit = vertex_vertex3(v0[0], v0[1 + w], v1[0], v1[1 + w], eta, all_t);
    if (!hit) return 0;
    if (all_t.size() <= 16) small_sort(all_t.data(), (int)all_t.size());
    else std::sort(all_t.begin(), all_t.end());
    for (double t : all_t) {
      if (t < 0) continue;
      if (!wrong_side3(v0, v1, t)) { toi = t; return 1; }
    }
    return -1;
  }
}

}  // namespace ctcd
}  // namespace mclo
