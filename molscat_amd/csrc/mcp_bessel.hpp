// Riccati-Bessel functions of real order for the Y -> K matching, evaluated on
// the device, one channel per thread.
//
// The algorithm is Temme's published method (N.M. Temme, J. Comput. Phys. 19
// (1975) 324 and 21 (1976) 343) in the combination used by the reference's
// RBESJY/BESSJY/RGAMMA (source_code/rbesjy.f) and RBESSK/MBESSK
// (source_code/rbessk.f): series for small x, the PQ continued fraction for
// intermediate x, Hankel's asymptotic series for large x, upward recurrence in
// the order with exponential scaling, and CF1 for J in the classically
// forbidden region.  The region boundaries, iteration limits and the
// Chebyshev coefficients of 1/Gamma(1-x) are Temme's and have to match the
// reference to reproduce its K matrices to rounding.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace mcp {

struct BesselStatus {
  int code;  // 0 ok, >0 the reference's STOP code ('BESSJY n' / 'MBESSK n')
};

__device__ inline double temme_rgamma(double x, double& odd, double& even) {
  constexpr double b[12] = {-0.283876542276024, -0.076852840844786, +0.001706305071096,
                            +0.001271927136655, +0.000076309597586, -0.000004971736704,
                            -0.000000865920800, -0.000000033126120, +0.000000001745136,
                            +0.000000000242310, +0.000000000009161, -0.000000000000170};
  const double x2 = x * x * 8.0;
  double alfa = -0.000000000000001, beta = 0.0;
  for (int i = 12; i >= 2; i -= 2) {
    beta = -(2 * alfa + beta);
    alfa = -beta * x2 - alfa + b[i - 1];
  }
  even = (beta / 2.0 + alfa) * x2 - alfa + 0.921870293650453;
  alfa = -0.000000000000034;
  beta = 0.0;
  for (int i = 11; i >= 1; i -= 2) {
    beta = -(2 * alfa + beta);
    alfa = -beta * x2 - alfa + b[i - 1];
  }
  odd = 2 * (alfa + beta);
  return odd * x + even;
}

// Ordinary Bessel J_v, Y_v and derivatives with exponential scaling:
// J = cj*exp(ej), J' = dj*exp(ej), Y = cy*exp(ey), Y' = dy*exp(ey).
__device__ inline int bessel_jy_scaled(double v, double x, double& cj, double& dj, double& ej, double& cy,
                                       double& dy, double& ey) {
  const double eps = 1.0e-15;
  const int maxit = 1000;
  const double pi = acos(-1.0);
  const double xmin = 3.0;
  const double xmax = 5.0 - log10(eps);
  const int na = (int)(v + 0.5);
  double a = v - na;
  double f, g, pa = 0.0, qa = 0.0, pa1 = 0.0, qa1 = 0.0, cc = 0.0, ss = 0.0, dd = 0.0;
  if (x < xmin) {
    const double b = x / 2.0;
    double d = -log(b);
    double e = a * d;
    double c = (fabs(a) < eps) ? 1.0 / pi : a / sin(a * pi);
    const double s = (fabs(e) < eps) ? 1.0 : sinh(e) / e;
    e = exp(e);
    double p, q;
    double gg = e * temme_rgamma(a, p, q);
    e = (e + 1.0 / e) / 2.0;
    f = 2 * c * (p * e + q * s * d);
    e = a * a;
    p = gg * c;
    q = 1.0 / gg / pi;
    c = a * pi / 2.0;
    double r = (fabs(c) < eps) ? 1.0 : sin(c) / c;
    r = pi * c * r * r;
    c = 1.0;
    d = -b * b;
    double ya = f + r * q;
    double ya1 = p;
    bool ok = false;
    for (int n = 1; n <= maxit; ++n) {
      f = (f * n + p + q) / (n * n - e);
      c = c * d / n;
      p = p / (n - a);
      q = q / (n + a);
      const double g1 = c * (f + r * q);
      const double h = c * p - n * g1;
      ya = ya + g1;
      ya1 = ya1 + h;
      if (fabs(g1) / (1.0 + fabs(ya)) + fabs(h) / (1.0 + fabs(ya1)) < eps) {
        ok = true;
        break;
      }
    }
    if (!ok) return 1;
    f = -ya;
    g = -ya1 / b;
  } else if (x < xmax) {
    const double c = 0.25 - a * a;
    const double b = x + x;
    const double e = (x * cos(a * pi) / pi / eps) * (x * cos(a * pi) / pi / eps);
    double p = 1.0, q = -x, r = 1.0 + x * x, s = r;
    int n = 2;
    bool ok = false;
    for (; n <= maxit; ++n) {
      const double d = (n - 1 + c / n) / s;
      p = (2 * n - p * d) / (n + 1);
      q = (-b + q * d) / (n + 1);
      s = p * p + q * q;
      r = r * s;
      if (r * n * n > e) {
        ok = true;
        break;
      }
    }
    if (!ok) return 2;
    p = p / s;
    f = p;
    q = -q / s;
    g = q;
    for (int m = n; m >= 1; --m) {
      const double rr = (m + 1) * (2.0 - p) - 2.0;
      const double s2 = b + (m + 1) * q;
      const double d = (m - 1 + c / m) / (rr * rr + s2 * s2);
      p = d * rr;
      q = d * s2;
      const double e2 = f + 1.0;
      f = p * e2 - g * q;
      g = q * e2 + p * g;
    }
    f = 1.0 + f;
    double d = f * f + g * g;
    pa = f / d;
    qa = -g / d;
    d = a + 0.5 - p;
    q = q + x;
    pa1 = (pa * q - qa * d) / x;
    qa1 = (qa * q + pa * d) / x;
    const double bb = x - pi * (a + 0.5) / 2.0;
    cc = cos(bb);
    ss = sin(bb);
    dd = sqrt(2.0 / x / pi);
    f = dd * (pa * ss + qa * cc);
    g = dd * (qa1 * ss - pa1 * cc);
  } else {
    double p = 0.0, q = 0.0;
    for (int ia = 0; ia <= 1; ++ia) {
      pa = p;
      qa = q;
      const double y = 4.0 * (a + ia) * (a + ia);
      const double z = 8.0 * x;
      double d = 0.0, w = -1.0, tp = 1.0, tq = 0.0;
      p = 1.0;
      q = 0.0;
      bool ok = false;
      for (int k = 1; k <= maxit; ++k) {
        d = d + z;
        w = w + 2.0;
        tq = +tp * (y - w * w) / d;
        q = q + tq;
        d = d + z;
        w = w + 2.0;
        tp = -tq * (y - w * w) / d;
        p = p + tp;
        if (fabs(tp) + fabs(tq) < eps) {
          ok = true;
          break;
        }
      }
      if (!ok) return 3;
      p = p - 0.5 * tp;
      q = q - 0.5 * tq;
    }
    pa1 = p;
    qa1 = q;
    const double bb = x - pi * (a + 0.5) / 2.0;
    cc = cos(bb);
    ss = sin(bb);
    dd = sqrt(2.0 / x / pi);
    f = dd * (pa * ss + qa * cc);
    g = dd * (qa1 * ss - pa1 * cc);
  }
  // upward recurrence Y(a) -> Y(v) with scaling by 1/16
  double p16 = 0.0;
  if (na > 0) {
    const double y = 2.0 / x;
    for (int n = 1; n <= na; ++n) {
      const double h = y * (a + n) * g - f;
      f = g;
      g = h;
      while (fabs(f) > 4.0) {
        p16 += 1.0;
        f *= 0.0625;
        g *= 0.0625;
      }
    }
  }
  cy = f;
  dy = (v / x) * f - g;
  double sy = sqrt(cy * cy + dy * dy);
  cy /= sy;
  dy /= sy;
  ey = log(sy) + p16 * log(16.0);
  const double vv = fmax(xmin, v);
  if (x >= vv) {
    f = dd * (pa * cc - qa * ss);
    g = dd * (qa1 * cc + pa1 * ss);
    if (na > 0) {
      const double y = 2.0 / x;
      for (int n = 1; n <= na; ++n) {
        const double h = y * (a + n) * g - f;
        f = g;
        g = h;
      }
    }
    cj = f;
    dj = (v / x) * f - g;
    const double sj = sqrt(cj * cj + dj * dj);
    cj /= sj;
    dj /= sj;
    ej = log(sj);
  } else {
    double ap = 1.0, aa = v / x, bp = 0.0, bb = 1.0;
    f = 0.0;
    g = 0.0;
    const double y = 2.0 / x;
    const double w = y / pi;
    bool ok = false;
    for (int n = 1; n <= maxit; ++n) {
      const double an = y * (v + n) * aa - ap;
      ap = aa;
      aa = an;
      const double bn = y * (v + n) * bb - bp;
      bp = bb;
      bb = bn;
      if (fabs(bb) > fabs(aa)) {
        ap /= bb;
        aa /= bb;
        bp /= bb;
        bb = 1.0;
        if (fabs(aa - f) < eps * fabs(f)) {
          cj = w / (dy - cy * aa);
          dj = aa * cj;
          ok = true;
          break;
        }
        f = aa;
      } else {
        bp /= aa;
        bb /= aa;
        ap /= aa;
        aa = 1.0;
        if (fabs(bb - g) < eps * fabs(g)) {
          dj = w / (dy * bb - cy);
          cj = bb * dj;
          ok = true;
          break;
        }
        g = bb;
      }
    }
    if (!ok) return 4;
    const double sj = sqrt(cj * cj + dj * dj);
    cj /= sj;
    dj /= sj;
    ej = log(sj) - ey;
  }
  return 0;
}

// Riccati-Bessel j^, j^', n^, n^' of order ell (>= -1/2) at x > 0.
__device__ inline int riccati_jy(double ell, double x, double& uj, double& ujp, double& un, double& unp) {
  const double v = ell + 0.5;
  double cj, dj, ej, cy, dy, ey;
  const int rc = bessel_jy_scaled(v, x, cj, dj, ej, cy, dy, ey);
  if (rc) return rc;
  const double pi = acos(-1.0);
  const double ex = 0.5 * log(pi * x / 2.0);
  dj = dj + cj / (2.0 * x);
  const double sj = sqrt(cj * cj + dj * dj);
  cj /= sj;
  dj /= sj;
  ej = ej + log(sj) + ex;
  dy = dy + cy / (2.0 * x);
  const double sy = sqrt(cy * cy + dy * dy);
  cy /= sy;
  dy /= sy;
  ey = ey + log(sy) + ex;
  uj = cj * exp(ej);
  ujp = dj * exp(ej);
  un = cy * exp(ey);
  unp = dy * exp(ey);
  return 0;
}

// K'/K for the modified Riccati-Bessel function of the third kind.
__device__ inline int riccati_k_ratio(double ell, double x, double& ratio) {
  const double v = ell + 0.5;
  const double eps = 1.0e-15;
  const int maxit = 1000;
  const double pi = acos(-1.0);
  const double xmin = 1.0;
  const int na = (int)(v + 0.5);
  const double a = v - na;
  double f, g;
  if (x < xmin) {
    const double b = x / 2.0;
    const double d = -log(b);
    double e = a * d;
    double c = a * pi;
    c = (fabs(c) < eps) ? 1.0 : c / sin(c);
    const double s = (fabs(e) < eps) ? 1.0 : sinh(e) / e;
    e = exp(e);
    double p, q;
    const double gg = e * temme_rgamma(a, p, q);
    e = (e + 1.0 / e) / 2.0;
    f = c * (p * e + q * s * d);
    e = a * a;
    p = 0.5 * gg * c;
    q = 0.5 / gg;
    c = 1.0;
    const double dd = b * b;
    double ak = f, ak1 = p;
    bool ok = false;
    for (int n = 1; n <= maxit; ++n) {
      f = (f * n + p + q) / (n * n - e);
      c = c * dd / n;
      p = p / (n - a);
      q = q / (n + a);
      const double g1 = c * (p - n * f);
      const double h = c * f;
      ak = ak + h;
      ak1 = ak1 + g1;
      if (h / ak + fabs(g1) / ak1 < eps) {
        ok = true;
        break;
      }
    }
    if (!ok) return 11;
    f = ak;
    g = ak1 / b;
  } else {
    const double c = 0.25 - a * a;
    g = 1.0;
    f = 0.0;
    const double e = x * cos(a * pi) / pi / eps;
    int n = 1;
    bool ok = false;
    for (; n <= maxit; ++n) {
      const double h = (2 * (n + x) * g - (n - 1 + c / n) * f) / (n + 1);
      f = g;
      g = h;
      if (h * n > e) {
        ok = true;
        break;
      }
    }
    if (!ok) return 12;
    double p = f / g, q = p;
    const double b = x + x;
    const double e2 = b - 2.0;
    for (int m = n; m >= 1; --m) {
      p = (m - 1 + c / m) / (e2 + (m + 1) * (2.0 - p));
      q = p * (q + 1.0);
    }
    f = sqrt(pi / b) / (1.0 + q);
    g = f * (a + x + 0.5 - p) / x;
  }
  if (na > 0) {
    const double y = 2.0 / x;
    for (int n = 1; n <= na; ++n) {
      const double h = y * (a + n) * g + f;
      f = g;
      g = h;
      while (fabs(f) > 4.0) {
        f *= 0.0625;
        g *= 0.0625;
      }
    }
  }
  double ck = f, dk = (v / x) * f - g;
  const double sk = sqrt(ck * ck + dk * dk);
  ck /= sk;
  dk /= sk;
  dk = dk + ck / (2.0 * x);
  ratio = dk / ck;
  return 0;
}

}  // namespace mcp
