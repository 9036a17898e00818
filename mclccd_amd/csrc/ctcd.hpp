// FP64 CTCD vertex-face narrow phase for gfx950: the path BVHTree takes for
// vertex-face stencils when options.vf_one_sided is set
// (include/MCL/BVHTree.cpp:446-447 -> NarrowPhaseCTCD::query_ccd_vf,
// include/MCL/NarrowPhase.cpp:283-419, with its wrong-side test).
//
// One stencil per lane.  Each sub-test builds the polynomials whose sign on
// [0, 1] encodes a contact condition (include/MCL/ccd_internal/CTCD.cpp:222-264,
// 435-817), turns each into at most degree+1 sign intervals (CTCD.cpp:45-184,
// Jenkins-Traub for degree > 2: include/MCL/ccd_internal/rpoly.h:56-762) and
// intersects them.  The reference collects every combination's start time in
// a vector and sorts it; here a sub-test keeps only the distinct start times,
// in order, in a fixed per-lane list: an overlap's start is the largest of its
// intervals' starts, so a vertex-face test has at most 1 + 7 + 3*4 distinct
// values and the first right-side time of the sorted list is unchanged.
//
// Arithmetic order is the reference's (contraction off, libstdc++ min/max,
// correctly rounded '/' and sqrt).  The root finder's exp and log come from
// the device math library; the rotation of its shifts, cos/sin of a constant,
// is computed on the host (kernel argument) so it is the host libm's value.
// Two loops of the root-finder's bound estimate stop after a cap that only
// non-finite coefficients reach (the reference would not terminate there).
//
// GPU structure: a lane walks its stencil's sub-tests one polynomial per pass
// of a single loop, so lanes of a wave in different sub-tests still share the
// root-finder call; the root finder's arrays are indexed only by compile-time
// constants (guarded loops over the maximum degree, select chains for a
// runtime index), so its state is kept in registers rather than scratch.
#pragma once
#include "geom.hpp"

#pragma clang fp contract(off)

namespace mclg {
namespace ctcd {

__device__ __forceinline__ d3 cross3(const d3& a, const d3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ d3 add3(const d3& a, const d3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }

// Small per-lane arrays stay in registers only if every access has a
// compile-time index: loops run over the fixed maximum size with a guard, and
// a runtime index goes through a select chain.
template <int N> __device__ __forceinline__ double at(const double (&x)[N], int i) {
  double r = x[0];
#pragma unroll
  for (int j = 1; j < N; ++j) r = (i == j) ? x[j] : r;
  return r;
}
template <int N> __device__ __forceinline__ void put(double (&x)[N], int i, double v) {
#pragma unroll
  for (int j = 0; j < N; ++j)
    if (i == j) x[j] = v;
}

// Sign intervals of one polynomial over [0, 1] (TimeInterval, CTCD.hpp:10-28):
// at most degree+1 <= 7.
struct Ivals {
  double l[7], u[7];
  int n;
  __device__ __forceinline__ void push(double t1, double t2) {
    if (n >= 7) return;  // at most degree+1 <= 7 intervals: never taken
    double a = t1, b = t2;
    if (a > b) { const double s = a; a = b; b = s; }
    put(l, n, smax(a, 0.0));
    put(u, n, smin(b, 1.0));
    ++n;
  }
};

// Candidate contact times of the first sub-test that hits.
struct Times {
  double t[20];
  int n;
  // interval sub-tests: distinct values, ascending (the values are finite)
  __device__ void insert_sorted(double x) {
    if (n >= 20) return;  // at most 1 + 7 + 3*4 distinct starts: never taken
    int i = 0;
    while (i < n && t[i] < x) ++i;
    if (i < n && t[i] == x) return;
    for (int j = n; j > i; --j) t[j] = t[j - 1];
    t[i] = x;
    ++n;
  }
};

// libstdc++ insertion sort (std::sort on <= 16 values), which fixes where a NaN lands.
__device__ __forceinline__ void insertion_sort(double* a, int n) {
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

// The same sort on the first n (<= 6) of six register values.
__device__ __forceinline__ void insertion_sort6(double (&a)[6], int n) {
#pragma unroll
  for (int i = 1; i < 6; ++i) {
    if (i >= n) break;
    const double val = a[i];
    if (val < a[0]) {
#pragma unroll
      for (int j = i; j > 0; --j) a[j] = a[j - 1];
      a[0] = val;
    } else {
      // unguarded: stops at j = 1 at the latest, since !(val < a[0])
      bool moving = true;
#pragma unroll
      for (int j = i; j >= 1; --j) {
        if (!moving) break;
        if (val < a[j - 1]) {
          a[j] = a[j - 1];
        } else {
          a[j] = val;
          moving = false;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Jenkins-Traub (TOMS 493) over degree <= 6, same steps and arithmetic as the
// reference's C translation (rpoly.h); the goto structure of the fixed-shift
// stage is a three-state loop, and every array loop runs to the maximum size
// under a guard so the state stays in registers.
// ---------------------------------------------------------------------------
constexpr double kEta = 2.22e-16, kAre = kEta, kMre = kEta;

struct JT {
  double p[7], qp[7], k[7], qk[7], svk[7];
  double u, v, a, b, c, d, a1, a3, a7, e, f, g, h, szr, szi, lzr, lzi;
  int n;

  // rpoly.h:687-700 — p / (1 + u z + v z^2): quotient q, remainder (ra, rb)
  __device__ __forceinline__ static void quadsd(int nn, double uu, double vv, const double (&pp)[7], double (&q)[7],
                                                double& ra, double& rb) {
    rb = pp[0];
    q[0] = rb;
    ra = pp[1] - rb * uu;
    q[1] = ra;
#pragma unroll
    for (int i = 2; i < 7; i++) {
      if (i <= nn) {
        const double cc = pp[i] - ra * uu - rb * vv;
        q[i] = cc;
        rb = ra;
        ra = cc;
      }
    }
  }

  // rpoly.h:706-760
  __device__ __forceinline__ static void quad(double a_, double b1, double c_, double& sr, double& si, double& lr,
                                              double& li) {
    if (a_ == 0.0) {
      sr = (b1 != 0.0) ? -c_ / b1 : 0.0;
      lr = 0.0;
      si = 0.0;
      li = 0.0;
      return;
    }
    if (c_ == 0.0) {
      sr = 0.0;
      lr = -b1 / a_;
      si = 0.0;
      li = 0.0;
      return;
    }
    const double hb = b1 / 2.0;
    double dd, ee;
    if (fabs(hb) < fabs(c_)) {
      ee = (c_ < 0.0) ? -a_ : a_;
      ee = hb * (hb / fabs(c_)) - ee;
      dd = sqrt(fabs(ee)) * sqrt(fabs(c_));
    } else {
      ee = 1.0 - (a_ / hb) * (c_ / hb);
      dd = sqrt(fabs(ee)) * fabs(hb);
    }
    if (ee < 0.0) {
      sr = -hb / a_;
      lr = sr;
      si = fabs(dd / a_);
      li = -si;
    } else {
      if (hb >= 0.0) dd = -dd;
      lr = (-hb + dd) / a_;
      sr = 0.0;
      if (lr != 0.0) sr = (c_ / lr) / a_;
      si = 0.0;
      li = 0.0;
    }
  }

  // rpoly.h:573-603; returns the type
  __device__ __forceinline__ int calcsc() {
    quadsd(n - 1, u, v, k, qk, c, d);
    if (!(fabs(c) > fabs(at(k, n - 1) * 100.0 * kEta)) && !(fabs(d) > fabs(at(k, n - 2) * 100.0 * kEta))) return 3;
    if (fabs(d) < fabs(c)) {
      e = a / c;
      f = d / c;
      g = u * e;
      h = v * b;
      a3 = a * e + (h / c + g) * b;
      a1 = b - a * (d / c);
      a7 = a + g * d + h * f;
      return 1;
    }
    e = a / d;
    f = c / d;
    g = u * b;
    h = v * b;
    a3 = (a + g) * e + h * (b / d);
    a1 = b * f - a;
    a7 = (f + u) * a + h;
    return 2;
  }

  // rpoly.h:608-641
  __device__ __forceinline__ void nextk(int type) {
    if (type == 3) {
      k[0] = 0.0;
      k[1] = 0.0;
#pragma unroll
      for (int i = 2; i < 7; i++)
        if (i < n) k[i] = qk[i - 2];
      return;
    }
    const double temp = (type == 1) ? b : a;
    if (fabs(a1) <= fabs(temp) * kEta * 10.0) {
      k[0] = 0.0;
      k[1] = -a7 * qp[0];
#pragma unroll
      for (int i = 2; i < 7; i++)
        if (i < n) k[i] = a3 * qk[i - 2] - a7 * qp[i - 1];
      return;
    }
    a7 /= a1;
    a3 /= a1;
    k[0] = qp[0];
    k[1] = qp[1] - a7 * qp[0];
#pragma unroll
    for (int i = 2; i < 7; i++)
      if (i < n) k[i] = a3 * qk[i - 2] - a7 * qp[i - 1] + qp[i];
  }

  // rpoly.h:647-682
  __device__ __forceinline__ void newest(int type, double& uu, double& vv) const {
    if (type == 3) {
      uu = 0.0;
      vv = 0.0;
      return;
    }
    double a4, a5;
    if (type == 2) {
      a4 = (a + g) * f + h;
      a5 = (f + u) * c + v * d;
    } else {
      a4 = a + u * b + h * f;
      a5 = c + (u + v * f) * d;
    }
    const double pn = at(p, n);
    const double b1 = -at(k, n - 1) / pn;
    const double b2 = -(at(k, n - 2) + b1 * at(p, n - 1)) / pn;
    const double c1 = v * b2 * a1;
    const double c2 = b1 * a7;
    const double c3 = b1 * b1 * a3;
    const double c4 = c1 - c2 - c3;
    const double temp = a5 + b1 * a4 - c4;
    if (temp == 0.0) {
      uu = 0.0;
      vv = 0.0;
      return;
    }
    uu = u - (u * (c3 + c2) + v * (b1 * a1 + b2 * a7)) / temp;
    vv = v * (1.0 + c4 / temp);
  }

  // rpoly.h:397-474; returns the zeros found (0 or 2)
  __device__ __forceinline__ int quadit(double uu, double vv) {
    double omp = 0, relstp = 0;
    bool tried = false;
    u = uu;
    v = vv;
    int j = 0;
    for (;;) {
      quad(1.0, u, v, szr, szi, lzr, lzi);
      if (fabs(fabs(szr) - fabs(lzr)) > 0.01 * fabs(lzr)) return 0;
      quadsd(n, u, v, p, qp, a, b);
      const double mp = fabs(a - szr * b) + fabs(szi * b);
      const double zm = sqrt(fabs(v));
      double ee = 2.0 * fabs(qp[0]);
      const double t = -szr * b;
#pragma unroll
      for (int i = 1; i < 7; i++)
        if (i < n) ee = ee * zm + fabs(qp[i]);
      ee = ee * zm + fabs(a + t);
      ee *= (5.0 * kMre + 4.0 * kAre);
      ee = ee - (5.0 * kMre + 2.0 * kAre) * (fabs(a + t) + fabs(b) * zm) + 2.0 * kAre * fabs(t);
      if (mp <= 20.0 * ee) return 2;
      j++;
      if (j > 20) return 0;
      if (!(j < 2) && !(relstp > 0.01 || mp < omp || tried)) {
        if (relstp < kEta) relstp = kEta;
        relstp = sqrt(relstp);
        u = u - u * relstp;
        v = v + v * relstp;
        quadsd(n, u, v, p, qp, a, b);
        for (int i = 0; i < 5; i++) nextk(calcsc());
        tried = true;
        j = 0;
      }
      omp = mp;
      nextk(calcsc());
      const int type = calcsc();
      double ui, vi;
      newest(type, ui, vi);
      if (vi == 0.0) return 0;
      relstp = fabs((vi - v) / vi);
      u = ui;
      v = vi;
    }
  }

  // rpoly.h:480-566; returns the zeros found (0 or 1); iflag: a near-double real zero
  __device__ __forceinline__ int realit(double& sss, int& iflag) {
    double t = 0, omp = 0;
    double s = sss;
    iflag = 0;
    int j = 0;
    for (;;) {
      double pv = p[0];
      qp[0] = pv;
#pragma unroll
      for (int i = 1; i < 7; i++) {
        if (i <= n) {
          pv = pv * s + p[i];
          qp[i] = pv;
        }
      }
      const double mp = fabs(pv);
      const double ms = fabs(s);
      double ee = (kMre / (kAre + kMre)) * fabs(qp[0]);
#pragma unroll
      for (int i = 1; i < 7; i++)
        if (i <= n) ee = ee * ms + fabs(qp[i]);
      if (mp <= 20.0 * ((kAre + kMre) * ee - kMre * mp)) {
        szr = s;
        szi = 0.0;
        return 1;
      }
      j++;
      if (j > 10) return 0;
      if (!(j < 2) && !(fabs(t) > 0.001 * fabs(s - t) || mp < omp)) {
        iflag = 1;
        sss = s;
        return 0;
      }
      omp = mp;
      double kv = k[0];
      qk[0] = kv;
#pragma unroll
      for (int i = 1; i < 7; i++) {
        if (i < n) {
          kv = kv * s + k[i];
          qk[i] = kv;
        }
      }
      if (fabs(kv) <= fabs(at(k, n - 1)) * 10.0 * kEta) {
        k[0] = 0.0;
#pragma unroll
        for (int i = 1; i < 7; i++)
          if (i < n) k[i] = qk[i - 1];
      } else {
        t = -pv / kv;
        k[0] = qp[0];
#pragma unroll
        for (int i = 1; i < 7; i++)
          if (i < n) k[i] = t * qk[i - 1] + qp[i];
      }
      kv = k[0];
#pragma unroll
      for (int i = 1; i < 7; i++)
        if (i < n) kv = kv * s + k[i];
      t = 0.0;
      if (fabs(kv) > fabs(at(k, n - 1) * 10.0 * kEta)) t = -pv / kv;
      s += t;
    }
  }

  // rpoly.h:284-390 with the shift (sr, v); returns the zeros found
  __device__ __forceinline__ int fxshfr(int l2, double sr) {
    double betav = 0.25, betas = 0.25;
    double oss = sr, ovv = v, ots = 0, otv = 0;
    quadsd(n, u, v, p, qp, a, b);
    int type = calcsc();
    for (int j = 0; j < l2; j++) {
      nextk(type);
      type = calcsc();
      double ui, vi;
      newest(type, ui, vi);
      const double vv = vi;
      double ss = 0.0;
      const double kn1 = at(k, n - 1);
      if (kn1 != 0.0) ss = -at(p, n) / kn1;
      double tv = 1.0, ts = 1.0;
      if (!(j == 0 || type == 3)) {
        if (vv != 0.0) tv = fabs((vv - ovv) / vv);
        if (ss != 0.0) ts = fabs((ss - oss) / ss);
        const double tvv = (tv < otv) ? tv * otv : 1.0;
        const double tss = (ts < ots) ? ts * ots : 1.0;
        const bool vpass = tvv < betav, spass = tss < betas;
        if (spass || vpass) {
          const double svu = u, svv = v;
#pragma unroll
          for (int i = 0; i < 7; i++)
            if (i < n) svk[i] = k[i];
          double s = ss;
          bool vtry = false, stry = false;
          // 0: quadratic iteration, 1: linear iteration, 2: restore and go on
          int st = ((spass && !vpass) || tss < tvv) ? 1 : 0;
          for (;;) {
            if (st == 0) {
              const int nz = quadit(ui, vi);
              if (nz > 0) return nz;
              vtry = true;
              betav *= 0.25;
              if (stry || !spass) { st = 2; continue; }
#pragma unroll
              for (int i = 0; i < 7; i++)
                if (i < n) k[i] = svk[i];
              st = 1;
              continue;
            }
            if (st == 1) {
              int iflag;
              const int nz = realit(s, iflag);
              if (nz > 0) return nz;
              stry = true;
              betas *= 0.25;
              if (iflag == 0) { st = 2; continue; }
              ui = -(s + s);
              vi = s * s;
              st = 0;
              continue;
            }
            u = svu;
            v = svv;
#pragma unroll
            for (int i = 0; i < 7; i++)
              if (i < n) k[i] = svk[i];
            if (vpass && !vtry) { st = 0; continue; }
            quadsd(n, u, v, p, qp, a, b);
            type = calcsc();
            break;
          }
        }
      }
      ovv = vv;
      oss = ss;
      otv = tv;
      ots = ts;
    }
    return 0;
  }

  // rpoly.h:56-268: zeros (real parts in zr) of op[0..degree], or -1
  __device__ __forceinline__ int solve(const double (&op)[7], int degree, double (&zr)[6], double cosr, double sinr) {
    const double infin = 3.4e38, smalno = 1.2e-38, lo = smalno / kEta;
    double xx = sqrt(0.5), yy = -xx;
    n = degree;
    if (op[0] == 0.0) return -1;
    while (at(op, n) == 0.0) {
      put(zr, degree - n, 0.0);
      n--;
    }
    if (n < 1) return -1;
    double temp[7], pt[7];
#pragma unroll
    for (int i = 0; i < 7; i++)
      if (i <= n) p[i] = op[i];
    for (;;) {
      if (n == 1) {
        put(zr, degree - 1, -p[1] / p[0]);
        n -= 1;
        break;
      }
      if (n == 2) {
        double sr, si, lr, li;
        quad(p[0], p[1], p[2], sr, si, lr, li);
        put(zr, degree - 2, sr);
        put(zr, degree - 1, lr);
        n -= 2;
        break;
      }
      double mx = 0.0, mnv = infin;
#pragma unroll
      for (int i = 0; i < 7; i++) {
        if (i <= n) {
          const double x = fabs(p[i]);
          if (x > mx) mx = x;
          if (x != 0.0 && x < mnv) mnv = x;
        }
      }
      double sc = lo / mnv;
      bool scale = !(sc > 1.0 && infin / sc < mx);
      if (scale && sc <= 1.0) {
        if (mx < 10.0) scale = false;
        else if (sc == 0.0) sc = smalno;
      }
      if (scale) {
        const int l = (int)(log(sc) / log(2.0) + 0.5);
        const double factor = pow(2.0, (double)l);
        if (factor != 1.0) {
#pragma unroll
          for (int i = 0; i < 7; i++)
            if (i <= n) p[i] = factor * p[i];
        }
      }
#pragma unroll
      for (int i = 0; i < 7; i++)
        if (i <= n) pt[i] = fabs(p[i]);
      const double ptn = -at(pt, n);
      put(pt, n, ptn);
      double x = exp((log(-ptn) - log(pt[0])) / (double)n);
      const double ptn1 = at(pt, n - 1);
      if (ptn1 != 0.0) {
        const double xm = -ptn / ptn1;
        if (xm < x) x = xm;
      }
      for (int it = 0; it < 2000; ++it) {
        const double xm = x * 0.1;
        double ff = pt[0];
#pragma unroll
        for (int i = 1; i < 7; i++)
          if (i <= n) ff = ff * xm + pt[i];
        if (ff <= 0.0) break;
        x = xm;
      }
      double dx = x;
      for (int it = 0; it < 1000 && fabs(dx / x) > 0.005; ++it) {
        double ff = pt[0], df = ff;
#pragma unroll
        for (int i = 1; i < 7; i++) {
          if (i < n) {
            ff = ff * x + pt[i];
            df = df * x + ff;
          }
        }
        ff = ff * x + ptn;
        dx = ff / df;
        x -= dx;
      }
      const double bnd = x;
#pragma unroll
      for (int i = 1; i < 7; i++)
        if (i < n) k[i] = (double)(n - i) * p[i] / (double)n;
      k[0] = p[0];
      const double aa = at(p, n), bb = at(p, n - 1);
      bool zerok = (at(k, n - 1) == 0);
      for (int jj = 0; jj < 5; jj++) {
        const double cc = at(k, n - 1);
        if (!zerok) {
          const double t = -aa / cc;
#pragma unroll
          for (int j = 6; j >= 1; --j)
            if (j <= n - 1) k[j] = t * k[j - 1] + p[j];
          k[0] = p[0];
          zerok = (fabs(at(k, n - 1)) <= fabs(bb) * kEta * 10.0);
        } else {
#pragma unroll
          for (int j = 6; j >= 1; --j)
            if (j <= n - 1) k[j] = k[j - 1];
          k[0] = 0.0;
          zerok = (at(k, n - 1) == 0.0);
        }
      }
#pragma unroll
      for (int i = 0; i < 7; i++)
        if (i < n) temp[i] = k[i];
      bool found = false;
      for (int cnt = 0; cnt < 20; cnt++) {
        const double xxx = cosr * xx - sinr * yy;
        yy = sinr * xx + cosr * yy;
        xx = xxx;
        const double sr = bnd * xx;
        u = -2.0 * sr;
        v = bnd;
        const int nz = fxshfr(20 * (cnt + 1), sr);
        if (nz != 0) {
          const int j = degree - n;
          put(zr, j, szr);
          n -= nz;
#pragma unroll
          for (int i = 0; i < 7; i++)
            if (i <= n) p[i] = qp[i];
          if (nz != 1) put(zr, j + 1, lzr);
          found = true;
          break;
        }
#pragma unroll
        for (int i = 0; i < 7; i++)
          if (i < n) k[i] = temp[i];
      }
      if (!found) break;
    }
    return degree - n;
  }
};

// CTCD.cpp:45-63
__device__ __forceinline__ int quad_roots(double a, double b, double c, double& t0, double& t1) {
  const int sign = (b < 0) ? -1 : 1;
  const double D = b * b - 4 * a * c;
  if (!(D >= 0)) return 0;
  const double q = -0.5 * (b + sign * sqrt(D));
  t0 = q / a;
  t1 = c / q;
  if (t0 > t1) { const double t = t0; t0 = t1; t1 = t; }
  return 2;
}

// CTCD.cpp:66-86 (op: degree+1 coefficients, decreasing powers)
template <int N>
__device__ __forceinline__ void check_interval(double t1, double t2, const double (&op)[N], int degree, Ivals& out,
                                               bool pos) {
  t1 = smax(0.0, t1);
  t2 = smax(0.0, t2);
  t1 = smin(1.0, t1);
  t2 = smin(1.0, t2);
  const double tmid = (t2 + t1) / 2;
  double f = op[0];
#pragma unroll
  for (int i = 1; i < N; i++) {
    if (i <= degree) {
      f *= tmid;
      f += op[i];
    }
  }
  if (pos ? (f >= 0) : (f <= 0)) out.push(t1, t2);
}

// CTCD.cpp:88-103
__device__ __forceinline__ bool could_have_roots(const double (&op)[7], int degree, bool pos) {
  double result = 0;
  if (pos ? op[0] > 0 : op[0] < 0) result = op[0];
#pragma unroll
  for (int i = 1; i < 7; i++) {
    if (i < degree) {
      result *= 1.0;
      if (pos ? op[i] > 0 : op[i] < 0) result += op[i];
    }
  }
  result *= 1.0;
  result += at(op, degree);
  return !(pos ? result < 0 : result > 0);
}

// CTCD.cpp:105-184; op (n+1 coefficients, decreasing powers) is normalised in place.
__device__ __forceinline__ void find_intervals(double (&op)[7], int n, Ivals& out, bool pos, double cosr,
                                               double sinr) {
  out.n = 0;
  double time[6];
  double maxval = 0;
#pragma unroll
  for (int i = 0; i < 7; i++)
    if (i <= n) maxval = smax(maxval, fabs(op[i]));
  if (maxval != 0) {
#pragma unroll
    for (int i = 0; i < 7; i++)
      if (i <= n) op[i] /= maxval;
  }
  int rd = n;
  bool lead = true;
#pragma unroll
  for (int i = 0; i < 6; i++) {
    if (i < n && lead) {
      if (op[i] == 0) rd--;
      else lead = false;
    }
  }
  if (rd < n) {
    const int sh = n - rd;
#pragma unroll
    for (int i = 0; i < 7; i++)
      if (i <= rd) op[i] = at(op, i + sh);
  }
  int roots;
  if (rd > 2) {
    if (!could_have_roots(op, rd, pos)) return;
    JT jt;
    roots = jt.solve(op, rd, time, cosr, sinr);
  } else if (rd == 2) {
    roots = quad_roots(op[0], op[1], op[2], time[0], time[1]);
  } else if (rd == 1) {
    time[0] = -op[1] / op[0];
    roots = 1;
  } else {
    if (pos ? op[0] >= 0 : op[0] <= 0) out.push(0.0, 1.0);
    return;
  }
  if (roots > 0) {
    insertion_sort6(time, roots);
    if (time[0] >= 0) check_interval(0, time[0], op, rd, out, pos);
#pragma unroll
    for (int i = 0; i < 5; i++)
      if (i < roots - 1 && !((time[i] < 0 && time[i + 1] < 0) || (time[i] > 1.0 && time[i + 1] > 1.0)))
        check_interval(time[i], time[i + 1], op, rd, out, pos);
    const double last = at(time, roots - 1);
    if (last <= 1.0) check_interval(last, 1.0, op, rd, out, pos);
  } else {
    check_interval(0.0, 1.0, op, rd, out, pos);
  }
}

// CTCD.cpp:222-236 (op[0..3])
__device__ __forceinline__ void plane_poly(const d3& x10, const d3& x20, const d3& x30, const d3& v10, const d3& v20,
                                           const d3& v30, double (&op)[7]) {
  op[0] = dot(v10, cross3(v20, v30));
  op[1] = dot(x10, cross3(v20, v30)) + dot(v10, cross3(x20, v30)) + dot(v10, cross3(v20, x30));
  op[2] = dot(x10, cross3(x20, v30)) + dot(x10, cross3(v20, x30)) + dot(v10, cross3(x20, x30));
  op[3] = dot(x10, cross3(x20, x30));
}

// CTCD.cpp:238-264 (op[0..6])
__device__ __forceinline__ void distance_poly(const d3& x10, const d3& x20, const d3& x30, const d3& v10,
                                              const d3& v20, const d3& v30, double minD2, double (&op)[7]) {
  const double A = dot(v10, cross3(v20, v30));
  const double B = dot(x10, cross3(v20, v30)) + dot(v10, cross3(x20, v30)) + dot(v10, cross3(v20, x30));
  const double C = dot(x10, cross3(x20, v30)) + dot(x10, cross3(v20, x30)) + dot(v10, cross3(x20, x30));
  const double D = dot(x10, cross3(x20, x30));
  const d3 E = cross3(x20, x30);
  const d3 F = add3(cross3(x20, v30), cross3(v20, x30));
  const d3 G = cross3(v20, v30);
  op[0] = A * A;
  op[1] = 2 * A * B;
  op[2] = B * B + 2 * A * C - dot(G, G) * minD2;
  op[3] = 2 * A * D + 2 * B * C - 2 * dot(G, F) * minD2;
  op[4] = 2 * B * D + C * C - (2 * dot(G, E) + dot(F, F)) * minD2;
  op[5] = 2 * C * D - 2 * dot(F, E) * minD2;
  op[6] = D * D - dot(E, E) * minD2;
}

__device__ __forceinline__ bool overlap(double al, double au, double bl, double bu) { return !(al > bu || bl > au); }

// Polynomial `part` of vertex-edge sub-test (vertex q0, edge q1-q2), CTCD.cpp:536-607:
// 0 and 1 the two projection quadratics (>= 0), 2 the colinearity quartic (<= 0).
__device__ __forceinline__ int vertex_edge_poly(int part, const d3& q0s, const d3& q1s, const d3& q2s, const d3& q0e,
                                                const d3& q1e, const d3& q2e, double minD, double (&op)[7], bool& pos) {
  const d3 v0 = sub(q0e, q0s), v1 = sub(q1e, q1s), v2 = sub(q2e, q2s);
  const d3 ab = sub(q2s, q1s), ac = sub(q0s, q1s), cb = sub(q2s, q0s);
  const d3 vab = sub(v2, v1), vac = sub(v0, v1), vcb = sub(v2, v0);
  if (part == 0) {
    op[0] = dot(vab, vac);
    op[1] = dot(ac, vab) + dot(ab, vac);
    op[2] = dot(ab, ac);
    pos = true;
    return 2;
  }
  if (part == 1) {
    op[0] = dot(vab, vcb);
    op[1] = dot(cb, vab) + dot(ab, vcb);
    op[2] = dot(ab, cb);
    pos = true;
    return 2;
  }
  const double A = dot(ab, ab), B = 2 * dot(ab, vab), C = dot(vab, vab);
  const double D = dot(ac, ac), E = 2 * dot(ac, vac), F = dot(vac, vac);
  const double G = dot(ac, ab), H = dot(vab, ac) + dot(vac, ab), I = dot(vab, vac);
  op[4] = A * D - G * G - minD * A;
  op[3] = B * D + A * E - 2 * G * H - minD * B;
  op[2] = B * E + A * F + C * D - H * H - 2 * G * I - minD * C;
  op[1] = B * F + C * E - 2 * H * I;
  op[0] = C * F - I * I;
  pos = false;
  return 4;
}

// CTCD.cpp:715-817 with the candidate list (its values appended in order;
// the check intervals accumulate, as in the reference)
__device__ __noinline__ bool vertex_vertex3(const d3& q1s, const d3& q2s, const d3& q1e, const d3& q2e, double eta,
                                            Times& T) {
  const double min_d = eta * eta;
  const d3 v1 = sub(q1e, q1s), v2 = sub(q2e, q2s);
  const double a = dot(v1, v1) + dot(v2, v2) - 2 * dot(v1, v2);
  const double b = 2 * (dot(v1, q1s) - dot(v2, q1s) - dot(v1, q2s) + dot(v2, q2s));
  const double c = dot(q1s, q1s) + dot(q2s, q2s) - 2 * dot(q1s, q2s) - min_d;
  double t1 = 0, t2 = 0;
  int roots;
  if (a != 0) {
    roots = quad_roots(a, b, c, t1, t2);
  } else if (b != 0) {
    t1 = -c / b;
    roots = 1;
  } else {
    if (c <= 0) {
      if (T.n < 20) T.t[T.n++] = 0.0;
      return true;
    }
    return false;
  }
  const double op[3] = {a, b, c};
  Ivals iv;
  iv.n = 0;
  if (roots == 2) {
    double curr = 1.0;
    check_interval(0, t1, op, 2, iv, false);
    if (iv.n) {
      curr = 0;
      if (T.n < 20) T.t[T.n++] = curr;
    }
    check_interval(t1, t2, op, 2, iv, false);
    if (iv.n) {
      curr = smin(curr, t1);
      if (T.n < 20) T.t[T.n++] = curr;
    }
    check_interval(t2, 1.0, op, 2, iv, false);
    if (iv.n) {
      curr = smin(t2, curr);
      if (T.n < 20) T.t[T.n++] = curr;
    }
    return T.n > 0;
  }
  if (roots == 1) {
    check_interval(0, t1, op, 2, iv, false);
    if (iv.n) {
      if (T.n < 20) T.t[T.n++] = 0.0;
      return true;
    }
    check_interval(t1, 1.0, op, 2, iv, false);
    if (iv.n) {
      if (T.n < 20) T.t[T.n++] = t1;
      return true;
    }
    return false;
  }
  check_interval(0, 1.0, op, 2, iv, false);
  if (iv.n) {
    if (T.n < 20) T.t[T.n++] = 0.0;
    return true;
  }
  return false;
}

// Distance.hpp:34-122: clamped barycentric coordinates of p's closest point on (q0, q1, q2)
__device__ __forceinline__ void vf_barys(const d3& p, const d3& q0, const d3& q1, const d3& q2, double* w3) {
  const d3 ab = sub(q1, q0), ac = sub(q2, q0), ap = sub(p, q0);
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0 && d2 <= 0) { w3[0] = 1.0; w3[1] = 0.0; w3[2] = 0.0; return; }
  const d3 bp = sub(p, q1);
  const double d3_ = dot(ab, bp), d4 = dot(ac, bp);
  if (d3_ >= 0 && d4 <= d3_) { w3[0] = 0.0; w3[1] = 1.0; w3[2] = 0.0; return; }
  const double vc = d1 * d4 - d3_ * d2;
  if ((vc <= 0) && (d1 >= 0) && (d3_ <= 0)) {
    const double v = d1 / (d1 - d3_);
    w3[0] = 1.0 - v; w3[1] = v; w3[2] = 0;
    return;
  }
  const d3 cp = sub(p, q2);
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0 && d5 <= d6) { w3[0] = 0; w3[1] = 0; w3[2] = 1.0; return; }
  const double vb = d5 * d2 - d1 * d6;
  if ((vb <= 0) && (d2 >= 0) && (d6 <= 0)) {
    const double w = d2 / (d2 - d6);
    w3[0] = 1 - w; w3[1] = 0; w3[2] = w;
    return;
  }
  const double va = d3_ * d6 - d5 * d4;
  if ((va <= 0) && (d4 - d3_ >= 0) && (d5 - d6 >= 0)) {
    const double w = (d4 - d3_) / ((d4 - d3_) + (d5 - d6));
    w3[0] = 0; w3[1] = 1.0 - w; w3[2] = w;
    return;
  }
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom, w = vc * denom;
  const double u = 1.0 - v - w;
  w3[0] = (u < 0.0) ? 0.0 : (1.0 < u) ? 1.0 : u;
  w3[1] = (v < 0.0) ? 0.0 : (1.0 < v) ? 1.0 : v;
  w3[2] = (w < 0.0) ? 0.0 : (1.0 < w) ? 1.0 : w;
}

__device__ __forceinline__ d3 lerp3(const d3& a, const d3& b, double t) {
  return {a.x * (1.0 - t) + b.x * t, a.y * (1.0 - t) + b.y * t, a.z * (1.0 - t) + b.z * t};
}
__device__ __forceinline__ d2 lerp2(const d2& a, const d2& b, double t) {
  return {a.x * (1.0 - t) + b.x * t, a.y * (1.0 - t) + b.y * t};
}

// NarrowPhase.cpp:58-89 (the normal scaled as Eigen's stableNormalize does)
__device__ __noinline__ bool wrong_side3(const d3* q0, const d3* q1, double t) {
  if (fabs(t) <= 0.0) return false;
  const d3 x0 = lerp3(q0[0], q1[0], t), x1 = lerp3(q0[1], q1[1], t), x2 = lerp3(q0[2], q1[2], t),
           x3 = lerp3(q0[3], q1[3], t);
  d3 n = cross3(sub(x2, x1), sub(x3, x1));
  if (dot(n, n) <= 0) return false;
  {
    const double w = smax(smax(fabs(n.x), fabs(n.y)), fabs(n.z));
    const d3 y = {n.x / w, n.y / w, n.z / w};
    const double z = dot(y, y);
    if (z > 0) {
      const double s = sqrt(z) * w;
      n = {n.x / s, n.y / s, n.z / s};
    }
  }
  double w3[3];
  vf_barys(x0, x1, x2, x3, w3);
  const d3 apex = sub(q1[0], q0[0]);
  const d3 f0 = {(w3[0] * q0[1].x + w3[1] * q0[2].x) + w3[2] * q0[3].x, (w3[0] * q0[1].y + w3[1] * q0[2].y) + w3[2] * q0[3].y,
                 (w3[0] * q0[1].z + w3[1] * q0[2].z) + w3[2] * q0[3].z};
  const d3 f1 = {(w3[0] * q1[1].x + w3[1] * q1[2].x) + w3[2] * q1[3].x, (w3[0] * q1[1].y + w3[1] * q1[2].y) + w3[2] * q1[3].y,
                 (w3[0] * q1[1].z + w3[1] * q1[2].z) + w3[2] * q1[3].z};
  const d3 vel = sub(apex, sub(f1, f0));
  return dot(n, vel) > 0.0;
}

// NarrowPhase.cpp:21-55
__device__ __forceinline__ bool wrong_side2(const d2* q0, const d2* q1, double t) {
  if (fabs(t) <= 0.0) return false;
  const d2 p = lerp2(q0[0], q1[0], t), e0 = lerp2(q0[1], q1[1], t), e1 = lerp2(q0[2], q1[2], t);
  d2 n = {e1.y - e0.y, -(e1.x - e0.x)};
  {
    const double w = smax(fabs(n.x), fabs(n.y));
    const d2 y = {n.x / w, n.y / w};
    const double z = dot(y, y);
    if (z > 0) {
      const double s = sqrt(z) * w;
      n = {n.x / s, n.y / s};
    }
  }
  const d2 de = sub(e0, e1);
  const double denom = sqrt(dot(de, de));
  if (denom < 2.220446049250313080847e-16) return false;
  const d2 a = sub(e1, p), b = sub(e0, p);
  const double b0 = sqrt(dot(a, a)) / denom, b1 = sqrt(dot(b, b)) / denom;
  const d2 e0b = {b0 * q0[1].x + b1 * q0[2].x, b0 * q0[1].y + b1 * q0[2].y};
  const d2 e1b = {b0 * q1[1].x + b1 * q1[2].x, b0 * q1[1].y + b1 * q1[2].y};
  const d2 vel = sub(sub(q1[0], q0[0]), sub(e1b, e0b));
  return dot(n, vel) > 0.0;
}

// NarrowPhase.cpp:122-161
__device__ __forceinline__ bool ray_box(const d3& p0, const d3& p1, const d3& bmin, const d3& bmax) {
  d3 dir = sub(p1, p0);
  const double t1 = sqrt(dot(dir, dir));
  const bool in0 = bmin.x <= p0.x && bmin.y <= p0.y && bmin.z <= p0.z && p0.x <= bmax.x && p0.y <= bmax.y && p0.z <= bmax.z;
  const bool in1 = bmin.x <= p1.x && bmin.y <= p1.y && bmin.z <= p1.z && p1.x <= bmax.x && p1.y <= bmax.y && p1.z <= bmax.z;
  if (in0 || in1) return true;
  const double z = dot(dir, dir);
  if (z > 0) {
    const double s = sqrt(z);
    dir = {dir.x / s, dir.y / s, dir.z / s};
  }
  const double ix = 1. / dir.x, iy = 1. / dir.y, iz = 1. / dir.z;
  double tmin = ((ix < 0 ? bmax.x : bmin.x) - p0.x) * ix;
  double tmax = ((ix < 0 ? bmin.x : bmax.x) - p0.x) * ix;
  const double tymin = ((iy < 0 ? bmax.y : bmin.y) - p0.y) * iy;
  const double tymax = ((iy < 0 ? bmin.y : bmax.y) - p0.y) * iy;
  if ((tmin > tymax) || (tymin > tmax)) return false;
  if (tymin > tmin) tmin = tymin;
  if (tymax < tmax) tmax = tymax;
  const double tzmin = ((iz < 0 ? bmax.z : bmin.z) - p0.z) * iz;
  const double tzmax = ((iz < 0 ? bmin.z : bmax.z) - p0.z) * iz;
  if ((tmin > tzmax) || (tzmin > tmax)) return false;
  if (tzmin > tmin) tmin = tzmin;
  if (tzmax < tmax) tmax = tzmax;
  return (tmin < t1) && (tmax > 0.0);
}

// CTCD.cpp:636-713 (2D: eta unused by the reference)
__device__ __forceinline__ bool vertex_edge2(const d2* q0, const d2* q1, Times& T) {
  const d2 u0 = sub(q0[1], q0[0]), v0 = sub(q0[2], q0[0]);
  const d2 u1 = sub(q1[1], q1[0]), v1 = sub(q1[2], q1[0]);
  const d2 up = sub(u1, u0), vp = sub(v1, v0);
  const double c = u0.x * v0.y - u0.y * v0.x;
  const double b = (u0.x * vp.y - u0.y * vp.x) + (up.x * v0.y - up.y * v0.x);
  const double a = up.x * vp.y - up.y * vp.x;
  double t0 = 2, t1 = 2;
  if (quad_roots(a, b, c, t0, t1) == 0) return false;
  auto on_line = [&](double t) {
    const d2 p = lerp2(q0[0], q1[0], t), e0 = lerp2(q0[1], q1[1], t), e1 = lerp2(q0[2], q1[2], t);
    const d2 d = sub(e0, e1), a0 = sub(p, e0), a1 = sub(p, e1);
    const double len = sqrt(dot(d, d));
    if (sqrt(dot(a0, a0)) > len) return false;
    if (sqrt(dot(a1, a1)) > len) return false;
    return true;
  };
  const bool on0 = on_line(t0), on1 = on_line(t1);
  bool hit = false;
  if (t0 >= 0 && t0 <= 1 && on0) { if (T.n < 20) T.t[T.n++] = t0; hit = true; }
  if (t1 >= 0 && t1 <= 1 && on1) { if (T.n < 20) T.t[T.n++] = t1; hit = true; }
  return hit;
}

// NarrowPhase.cpp:283-419 with test_wrong_side: 1 hit at toi, 0 none, -1 wrong side only.
template <int D> struct Query;

template <> struct Query<3> {
  // Combines the sign intervals of a sub-test (CTCD.cpp:506-523 / 608-625): the
  // start of every overlapping combination, TimeInterval::intersect from (0, 1)
  // in list order (the distance/colinearity list first).
  __device__ static bool combine(const Ivals* iv, bool vf, Times& T) {
    const Ivals& co = iv[vf ? 3 : 2];
    const Ivals& e1 = iv[0];
    const Ivals& e2 = iv[1];
    bool col = false;
    for (int i = 0; i < co.n; i++)
      for (int j = 0; j < e1.n; j++) {
        if (!overlap(co.l[i], co.u[i], e1.l[j], e1.u[j])) continue;
        for (int k = 0; k < e2.n; k++) {
          if (!overlap(co.l[i], co.u[i], e2.l[k], e2.u[k]) || !overlap(e1.l[j], e1.u[j], e2.l[k], e2.u[k])) continue;
          double lo = smax(co.l[i], 0.0);
          lo = smax(e1.l[j], lo);
          lo = smax(e2.l[k], lo);
          if (!vf) {
            T.insert_sorted(lo);
            col = true;
            continue;
          }
          const Ivals& e3 = iv[2];
          for (int l = 0; l < e3.n; l++) {
            if (!overlap(co.l[i], co.u[i], e3.l[l], e3.u[l]) || !overlap(e1.l[j], e1.u[j], e3.l[l], e3.u[l]) ||
                !overlap(e2.l[k], e2.u[k], e3.l[l], e3.u[l]))
              continue;
            T.insert_sorted(smax(e3.l[l], lo));
            col = true;
          }
        }
      }
    return col;
  }

  // The query's prefilter: the vertex's path against the face's box over both
  // time steps, padded by eta; false means no hit (run returns 0).
  __device__ static bool prefilter(const d3* q0, const d3* q1, double eta) {
    d3 lo = {1.79769313486231570815e+308, 1.79769313486231570815e+308, 1.79769313486231570815e+308};
    d3 hi = {-lo.x, -lo.y, -lo.z};
#pragma unroll
    for (int i = 1; i <= 3; ++i) {
      const d3 ps[2] = {q0[i], q1[i]};
      for (const d3& q : ps) {
        lo = {smin(lo.x, q.x), smin(lo.y, q.y), smin(lo.z, q.z)};
        hi = {smax(hi.x, q.x), smax(hi.y, q.y), smax(hi.z, q.z)};
      }
    }
    lo = {lo.x - eta, lo.y - eta, lo.z - eta};
    hi = {hi.x + eta, hi.y + eta, hi.z + eta};
    return ray_box(q0[0], q1[0], lo, hi);
  }

  __device__ static int run(const d3* q0, const d3* q1, double eta, double& toi, double cosr, double sinr) {
    if (!prefilter(q0, q1, eta)) return 0;
    const double minD = eta * eta;
    Times T;
    T.n = 0;
    Ivals iv[4];  // the current sub-test's sign intervals
    bool hit = false;
    // One polynomial per pass, so every lane reaches the root finder from the
    // same place whatever sub-test it is in.  Phase 0-2: the vertex-face edge
    // planes, 3: its distance polynomial (CTCD.cpp:435-534); 4 + 3e + part: the
    // vertex-edge sub-test of face edge e (CTCD.cpp:536-634); 13: vertex-vertex.
    int ph = 0;
    while (ph < 13) {
      double op[7];
      int deg, slot;
      bool pos;
      if (ph < 4) {
        slot = ph;
        const d3 v0 = sub(q1[0], q0[0]), v1 = sub(q1[1], q0[1]), v2 = sub(q1[2], q0[2]), v3 = sub(q1[3], q0[3]);
        if (ph == 0) {
          plane_poly(sub(q0[0], q0[1]), cross3(sub(q0[3], q0[1]), sub(q0[2], q0[1])), sub(q0[3], q0[1]), sub(v0, v1),
                     cross3(sub(v3, v1), sub(v2, v1)), sub(v3, v1), op);
        } else if (ph == 1) {
          plane_poly(sub(q0[0], q0[2]), cross3(sub(q0[1], q0[2]), sub(q0[3], q0[2])), sub(q0[1], q0[2]), sub(v0, v2),
                     cross3(sub(v1, v2), sub(v3, v2)), sub(v1, v2), op);
        } else if (ph == 2) {
          plane_poly(sub(q0[0], q0[3]), cross3(sub(q0[2], q0[3]), sub(q0[1], q0[3])), sub(q0[2], q0[3]), sub(v0, v3),
                     cross3(sub(v2, v3), sub(v1, v3)), sub(v2, v3), op);
        } else {
          distance_poly(sub(q0[0], q0[1]), sub(q0[2], q0[1]), sub(q0[3], q0[1]), sub(v0, v1), sub(v2, v1),
                        sub(v3, v1), minD, op);
        }
        deg = ph < 3 ? 3 : 6;
        pos = ph < 3;
      } else {
        const int e = (ph - 4) / 3, part = (ph - 4) % 3;
        slot = part;
        // face edge e runs from face vertex e to face vertex e+1 (mod 3)
        const d3 a0 = pick3(e, q0[1], q0[2], q0[3]), b0 = pick3(e, q0[2], q0[3], q0[1]);
        const d3 a1 = pick3(e, q1[1], q1[2], q1[3]), b1 = pick3(e, q1[2], q1[3], q1[1]);
        deg = vertex_edge_poly(part, q0[0], a0, b0, q1[0], a1, b1, minD, op, pos);
      }
      Ivals cur;
      find_intervals(op, deg, cur, pos, cosr, sinr);
      iv[slot] = cur;
      if (cur.n == 0) {  // the sub-test fails: on to the next one
        ph = ph < 4 ? 4 : 4 + 3 * ((ph - 4) / 3 + 1);
        continue;
      }
      if (ph == 3 || (ph >= 4 && (ph - 4) % 3 == 2)) {
        if (combine(iv, ph == 3, T)) {
          hit = true;
          break;
        }
      }
      ph++;
    }
    bool raw = false;  // vertex-vertex lists are appended unsorted
    for (int w = 0; w < 3 && !hit; ++w) {
      hit = vertex_vertex3(q0[0], q0[1 + w], q1[0], q1[1 + w], eta, T);
      raw = true;
    }
    if (!hit) return 0;
    if (raw) insertion_sort(T.t, T.n);
    for (int i = 0; i < T.n; ++i) {
      const double t = T.t[i];
      if (t < 0) continue;
      if (!wrong_side3(q0, q1, t)) {
        toi = t;
        return 1;
      }
    }
    return -1;
  }
};

template <> struct Query<2> {
  __device__ static int run(const d2* q0, const d2* q1, double eta, double& toi, double, double) {
    Times T;
    T.n = 0;
    if (!vertex_edge2(q0, q1, T)) return 0;
    insertion_sort(T.t, T.n);
    for (int i = 0; i < T.n; ++i)
      if (!wrong_side2(q0, q1, T.t[i])) {
        toi = T.t[i];
        return 1;
      }
    return -1;
  }
};

}  // namespace ctcd
}  // namespace mclg
