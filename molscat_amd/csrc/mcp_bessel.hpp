// Riccati-Bessel functions of real order for the Y -> K matching (YTOK,
// source_code/ytok.f:24-45), one channel per thread on the device; the same
// source builds for the host for the CPU tests.
//
// Orders ell >= -1/2 enter as nu = ell + 1/2, split as nu = mu + m with the
// fractional part mu in (-1/2, 1/2] and m a non-negative integer, so that an
// integer ell gives mu = +1/2 (no cancellation in Y_mu at small x).
//
// Published methods used (the code is organised around them, not around any
// other implementation):
//   * 1/Gamma(1+z) by its Taylor series, coefficients of Abramowitz & Stegun
//     6.1.34; it gives Temme's auxiliary functions gamma_1, gamma_2 by
//     splitting the series into odd and even powers.
//   * Temme's series (N.M. Temme, J. Comput. Phys. 19 (1975) 324) for
//     Y_mu, Y_mu+1 and K_mu, K_mu+1 when x < 2.
//   * Steed's complex continued fraction for (J' + iY')/(J + iY) at order mu
//     (A.R. Barnett et al., Comput. Phys. Commun. 8 (1974) 377) and the
//     continued fraction for the Tricomi-function ratio that gives
//     K_mu+1/K_mu (I.J. Thompson & A.R. Barnett, J. Comput. Phys. 64 (1986)
//     490) when x >= 2.
//   * The continued fraction for J'_nu/J_nu, with the sign of J_nu from the
//     signs of the denominators, and the Wronskian J Y' - J' Y = 2/(pi x) to
//     normalise J; continued fractions evaluated by the modified Lentz method
//     (W.J. Lentz, Appl. Opt. 15 (1976) 668; Thompson & Barnett 1986).
//   * Order recurrences: J downwards (stable), Y upwards, K's ratio upwards.
// Values are carried as mantissa x 2^exponent through the recurrences, so
// nothing overflows before the final result.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace mcp {
namespace rb {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTiny = 1e-300;
constexpr double kEps = 1e-16;
constexpr double kStepTol = 4.5e-16;  // two ulps of 1: a continued fraction's step has converged
constexpr int kRescale = 600;  // binary exponent at which recurrences rescale

// status codes (returned as the matching status of the instance)
constexpr int kBadArgument = 1;
constexpr int kNoConvergeRatioJ = 2;
constexpr int kNoConvergeSteed = 3;
constexpr int kNoConvergeTemme = 4;
constexpr int kNoConvergeK = 11;

// gm = (1/Gamma(1-mu) - 1/Gamma(1+mu)) / (2 mu), gp = (1/Gamma(1-mu) + 1/Gamma(1+mu)) / 2,
// from 1/Gamma(1+z) = sum_k a_k z^k: gm = -sum_j a_{2j+1} mu^{2j}, gp = sum_j a_{2j} mu^{2j}.
__host__ __device__ inline void inv_gamma_parts(double mu, double& gm, double& gp) {
  constexpr double a[26] = {1.0,
                            0.5772156649015329,
                            -0.6558780715202538,
                            -0.0420026350340952,
                            0.1665386113822915,
                            -0.0421977345555443,
                            -0.0096219715278770,
                            0.0072189432466630,
                            -0.0011651675918591,
                            -0.0002152416741149,
                            0.0001280502823882,
                            -0.0000201348547807,
                            -0.0000012504934821,
                            0.0000011330272320,
                            -0.0000002056338417,
                            0.0000000061160950,
                            0.0000000050020075,
                            -0.0000000011812746,
                            0.0000000001043427,
                            0.0000000000077823,
                            -0.0000000000036968,
                            0.0000000000005100,
                            -0.0000000000000206,
                            -0.0000000000000054,
                            0.0000000000000014,
                            0.0000000000000001};
  const double m2 = mu * mu;
  double odd = 0.0, even = 0.0;
  for (int j = 12; j >= 0; --j) {
    odd = odd * m2 + a[2 * j + 1];
    even = even * m2 + a[2 * j];
  }
  gm = -odd;
  gp = even;
}

// Temme's series at |mu| <= 1/2, 0 < x < 2.  Ordinary kind (modified =
// false): c0 = Y_mu, c1 = Y_mu+1.  Modified kind: c0 = K_mu, c1 = K_mu+1.
__host__ __device__ inline int temme_pair(double mu, double x, bool modified, double& c0, double& c1) {
  double gm, gp;
  inv_gamma_parts(mu, gm, gp);
  const double lg = log(2.0 / x);
  const double sigma = mu * lg;
  const double mpi = mu * kPi;
  const double ratio_sin = (mu == 0.0) ? 1.0 : mpi / sin(mpi);
  const double sinhc = (sigma == 0.0) ? 1.0 : sinh(sigma) / sigma;
  const double big_f = ratio_sin * (cosh(sigma) * gm + sinhc * lg * gp);
  const double big_p = exp(sigma) / (gp - mu * gm);   // (2/x)^mu Gamma(1+mu)
  const double big_q = exp(-sigma) / (gp + mu * gm);  // (x/2)^mu Gamma(1-mu)
  double fk, pk, qk, extra = 0.0, quarter;
  if (modified) {
    fk = big_f;
    pk = 0.5 * big_p;
    qk = 0.5 * big_q;
    quarter = 0.25 * x * x;
  } else {
    fk = (2.0 / kPi) * big_f;
    pk = big_p / kPi;
    qk = big_q / kPi;
    quarter = -0.25 * x * x;
    const double half = 0.5 * mpi;  // (2/mu) sin^2(mu pi/2) = pi * half * sinc(half)^2
    const double sc = (half == 0.0) ? 1.0 : sin(half) / half;
    extra = kPi * half * sc * sc;
  }
  double ck = 1.0;
  double g = fk + extra * qk;
  double sum0 = g, sum1 = pk;
  for (int k = 1; k <= 400; ++k) {
    const double kd = (double)k;
    fk = (kd * fk + pk + qk) / (kd * kd - mu * mu);
    pk /= (kd - mu);
    qk /= (kd + mu);
    ck *= quarter / kd;
    g = fk + extra * qk;
    const double t0 = ck * g;
    const double t1 = ck * (pk - kd * g);
    sum0 += t0;
    sum1 += t1;
    if (fabs(t0) <= kEps * fabs(sum0) && fabs(t1) <= kEps * fabs(sum1)) {
      if (modified) {
        c0 = sum0;
        c1 = sum1 * (2.0 / x);
      } else {
        c0 = -sum0;
        c1 = -sum1 * (2.0 / x);
      }
      return 0;
    }
  }
  return kNoConvergeTemme;
}

// Modified Lentz evaluation of b0 + a1/(b1 + a2/(b2 + ...)) with the terms
// supplied by the functor, which returns false when out of terms.
template <typename Terms>
__host__ __device__ inline int lentz(double b0, Terms terms, int maxit, double& value, int* flips = nullptr) {
  double h = (b0 == 0.0) ? kTiny : b0;
  double cc = h, dd = 0.0;
  int neg = 0;
  for (int k = 1; k <= maxit; ++k) {
    double ak, bk;
    terms(k, ak, bk);
    dd = bk + ak * dd;
    if (dd == 0.0) dd = kTiny;
    cc = bk + ak / cc;
    if (cc == 0.0) cc = kTiny;
    dd = 1.0 / dd;
    if (dd < 0.0) ++neg;
    const double step = cc * dd;
    h *= step;
    if (fabs(step - 1.0) < kStepTol) {
      value = h;  // with b0 = 0 the starting kTiny cancels in the first step
      if (flips) *flips = neg;
      return 0;
    }
  }
  return 1;
}

// J'_nu / J_nu, with the sign of J_nu relative to J_nu+K for K -> infinity
// (positive) from the signs of the Lentz denominators.
__host__ __device__ inline int ratio_j(double nu, double x, double& f, double& sign) {
  const double two_over_x = 2.0 / x;
  // J_nu+1/J_nu = 1/(b1 - 1/(b2 - ...)), b_k = 2(nu+k)/x
  auto terms = [&](int k, double& ak, double& bk) {
    ak = (k == 1) ? 1.0 : -1.0;
    bk = two_over_x * (nu + k);
  };
  const int maxit = 2000 + 4 * (int)fmin(x, 2.0e6);
  double h;
  int flips = 0;
  if (lentz(0.0, terms, maxit, h, &flips)) return kNoConvergeRatioJ;
  f = nu / x - h;
  sign = (flips & 1) ? -1.0 : 1.0;
  return 0;
}

// (ur + i ui) = (nr + i ni) / (dr + i di) by Smith's scaling, safe for any
// representable denominator.
__host__ __device__ inline void cdiv(double nr, double ni, double dr, double di, double& ur, double& ui) {
  if (fabs(dr) >= fabs(di)) {
    const double t = di / dr, den = dr + di * t;
    ur = (nr + ni * t) / den;
    ui = (ni - nr * t) / den;
  } else {
    const double t = dr / di, den = dr * t + di;
    ur = (nr * t + ni) / den;
    ui = (ni * t - nr) / den;
  }
}

// Steed's continued fraction: p + i q = (J'_mu + i Y'_mu) / (J_mu + i Y_mu) for x >= 2:
// p + iq = -1/(2x) + i + (i/x) a1 / G,  G = b1 + a2/(b2 + a3/(b3 + ...)),
// a_k = (k - 1/2)^2 - mu^2, b_k = 2(x + k i).  G (|b1| >= 4) by modified Lentz on
// (re, im) pairs.
__host__ __device__ inline int steed_pq(double mu, double x, double& p, double& q) {
  const double a1 = 0.25 - mu * mu;
  double gr = 2.0 * x, gi = 2.0;  // G_1 = b1
  double cr = gr, ci = gi, dr = 0.0, di = 0.0;
  bool done = (a1 == 0.0);  // mu = +-1/2: the fraction ends at its first term
  for (int k = 2; k <= 20000 && !done; ++k) {
    const double ak = (k - 0.5) * (k - 0.5) - mu * mu;
    const double br = 2.0 * x, bi = 2.0 * k;
    // d = 1 / (b + a d)
    double tr = br + ak * dr, ti = bi + ak * di;
    if (tr == 0.0 && ti == 0.0) tr = kTiny;
    cdiv(1.0, 0.0, tr, ti, dr, di);
    // c = b + a / c
    double qr, qi;
    cdiv(ak, 0.0, cr, ci, qr, qi);
    cr = br + qr;
    ci = bi + qi;
    if (cr == 0.0 && ci == 0.0) cr = kTiny;
    const double sr = cr * dr - ci * di, si = cr * di + ci * dr;
    const double nr = gr * sr - gi * si, ni = gr * si + gi * sr;
    gr = nr;
    gi = ni;
    done = fabs(sr - 1.0) + fabs(si) < kStepTol;
  }
  if (!done) return kNoConvergeSteed;
  double hr, hi;  // a1 / G
  cdiv(a1, 0.0, gr, gi, hr, hi);
  // (i/x) (hr + i hi) = (-hi + i hr)/x
  p = -0.5 / x - hi / x;
  q = 1.0 + hr / x;
  return 0;
}

// K_mu+1 / K_mu for x >= 2: (mu + 1/2 + x + (mu^2 - 1/4) r) / x with
// r = 1/(2(1+x) - c1/(2(2+x) - c2/(...))), c_n = (n + 1/2)^2 - mu^2.
__host__ __device__ inline int ratio_k_cf(double mu, double x, double& rk) {
  auto terms = [&](int k, double& ak, double& bk) {
    ak = (k == 1) ? 1.0 : -((k - 0.5) * (k - 0.5) - mu * mu);
    bk = 2.0 * (k + x);
  };
  double r;
  if (lentz(0.0, terms, 20000, r)) return kNoConvergeK;
  rk = (mu + 0.5 + x + (mu * mu - 0.25) * r) / x;
  return 0;
}

__host__ __device__ inline void split_order(double nu, double& mu, int& m) {
  m = (int)ceil(nu - 0.5);
  if (m < 0) m = 0;
  mu = nu - m;
}

// Riccati-Bessel functions j^ = sqrt(pi x/2) J_nu, n^ = sqrt(pi x/2) Y_nu
// (nu = ell + 1/2) and their x-derivatives, at order ell >= -1/2, x > 0.
__host__ __device__ inline int riccati_jy(double ell, double x, double& uj, double& ujp, double& un, double& unp) {
  if (!(x > 0.0) || ell < -0.5) return kBadArgument;
  const double nu = ell + 0.5;
  double mu;
  int m;
  split_order(nu, mu, m);
  double fnu, sgn;
  int rc = ratio_j(nu, x, fnu, sgn);
  if (rc) return rc;
  // J downwards from nu to mu on an unnormalised solution started at 1
  double j = 1.0, jd = fnu;
  int jexp = 0;
  for (int k = m; k >= 1; --k) {
    const double order = mu + k;
    const double jlow = (order / x) * j + jd;
    jd = ((order - 1.0) / x) * jlow - j;
    j = jlow;
    int e;
    frexp(j, &e);
    if (e > kRescale) {
      j = ldexp(j, -kRescale);
      jd = ldexp(jd, -kRescale);
      jexp += kRescale;
    }
  }
  const double fmu = jd / j;
  const double wron = 2.0 / (kPi * x);
  double jmu, ymu, ymud;
  if (x < 2.0) {
    double y0, y1;
    rc = temme_pair(mu, x, false, y0, y1);
    if (rc) return rc;
    ymu = y0;
    ymud = (mu / x) * y0 - y1;
    jmu = wron / (ymud - fmu * ymu);
  } else {
    double p, q;
    rc = steed_pq(mu, x, p, q);
    if (rc) return rc;
    const double gam = (p - fmu) / q;
    // the unnormalised solution has the sign of J_nu times sgn at order mu
    jmu = copysign(sqrt(wron / (q * (1.0 + gam * gam))), sgn * j);
    ymu = gam * jmu;
    ymud = p * ymu + q * jmu;
  }
  // J_nu = J_mu * (j(nu) / j(mu)) with j(nu) = 1, j(mu) = j 2^jexp
  const double jnu_m = jmu / j;  // times 2^-jexp
  const double jnud_m = fnu * jnu_m;
  // Y upwards from mu to nu
  double ylo = ymu, yhi = (mu / x) * ymu - ymud;  // Y_mu, Y_mu+1
  int yexp = 0;
  for (int k = 1; k <= m; ++k) {
    const double order = mu + k;
    const double ynext = (2.0 * order / x) * yhi - ylo;
    ylo = yhi;
    yhi = ynext;
    int e;
    frexp(yhi, &e);
    if (e > kRescale) {
      ylo = ldexp(ylo, -kRescale);
      yhi = ldexp(yhi, -kRescale);
      yexp += kRescale;
    }
  }
  const double ynu_m = ylo, ynud_m = (nu / x) * ylo - yhi;  // times 2^yexp
  const double s = sqrt(0.5 * kPi * x);
  const double half_x = 0.5 / x;
  uj = ldexp(s * jnu_m, -jexp);
  ujp = ldexp(s * (jnud_m + half_x * jnu_m), -jexp);
  un = ldexp(s * ynu_m, yexp);
  unp = ldexp(s * (ynud_m + half_x * ynu_m), yexp);
  return 0;
}

// d/dx log(sqrt(x) K_nu(x)) at nu = ell + 1/2: the log-derivative of the
// exponentially decaying modified Riccati-Bessel function.
__host__ __device__ inline int riccati_k_ratio(double ell, double x, double& ratio) {
  if (!(x > 0.0) || ell < -0.5) return kBadArgument;
  const double nu = ell + 0.5;
  double mu;
  int m;
  split_order(nu, mu, m);
  double r;  // K_order+1 / K_order
  if (x < 2.0) {
    double k0, k1;
    const int rc = temme_pair(mu, x, true, k0, k1);
    if (rc) return rc + (kNoConvergeK - kNoConvergeTemme);
    r = k1 / k0;
  } else {
    const int rc = ratio_k_cf(mu, x, r);
    if (rc) return rc;
  }
  // K_o+2 = K_o + (2(o+1)/x) K_o+1, so r_o+1 = 1/r_o + 2(o+1)/x
  for (int k = 1; k <= m; ++k) r = 1.0 / r + 2.0 * (mu + k) / x;
  ratio = nu / x - r + 0.5 / x;
  return 0;
}

}  // namespace rb

using rb::riccati_jy;
using rb::riccati_k_ratio;

}  // namespace mcp
