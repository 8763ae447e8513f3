// First-order step integrals of the VIVS propagator (the quantities PERT2
// supplies, source_code/pert2.f:36-222, for a constant perturbation): over a
// step of length h, for two channels whose reference solutions solve
// y'' = -k^2 y with k^2 = DIF (k^2 < 0: closed),
//   c(r) = cos(k r)  (cosh),   s(r) = sin(k r)/k  (sinh(|k| r)/|k|),
//   Icc = int_0^h c_i c_j dr,  Ics = int_0^h c_i s_j dr,
//   Isc = int_0^h s_i c_j dr,  Iss = int_0^h s_i s_j dr.
// Written in terms of z = DIF h^2 and the entire functions
//   C(z) = cos(sqrt z),  S(z) = sin(sqrt z)/sqrt z,  T(z) = (1 - cos(sqrt z))/z
// (their hyperbolic continuations for z < 0), so one set of formulas covers
// open, closed and mixed pairs.  Three evaluations, chosen so that none
// loses digits:
//   * both |z| <= 2: the double power series of the integrals, term by term
//     int_0^1 t^p dt = 1/(p+1) over the Taylor series of C and S;
//   * z_i, z_j apart (opposite signs, or |z_i - z_j| >= max|z| / 2): the
//     divided differences that integrating the product of two solutions
//     gives, e.g. Iss = h^3 (S_i C_j - C_i S_j) / (z_i - z_j);
//   * same sign and close: the products turned into sums, whose sum and
//     difference arguments x_i +- x_j (x = |k| h) enter only through S and T,
//     e.g. Icc = h/2 (S(sigma d^2) + S(sigma s^2)) with d = x_i - x_j.
// The same source builds for the host for the CPU tests.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace mcp {

struct StumpffCST {
  double c, s, t;
};

__host__ __device__ inline StumpffCST stumpff_cst(double z) {
  StumpffCST f;
  if (fabs(z) <= 1.0) {
    // C = sum (-z)^k/(2k)!, S = sum (-z)^k/(2k+1)!, T = sum (-z)^k/(2k+2)!
    double term = 1.0, c = 0.0, s = 0.0, t = 0.0;
    for (int k = 0; k < 12; ++k) {
      c += term;
      const double u = term / (double)(2 * k + 1);
      s += u;
      const double v = u / (double)(2 * k + 2);
      t += v;
      term = -z * v;
    }
    f.c = c;
    f.s = s;
    f.t = t;
  } else if (z > 0.0) {
    const double w = sqrt(z);
    const double q = sin(0.5 * w) / (0.5 * w);
    f.c = cos(w);
    f.s = sin(w) / w;
    f.t = 0.5 * q * q;
  } else {
    const double w = sqrt(-z);
    const double q = sinh(0.5 * w) / (0.5 * w);
    f.c = cosh(w);
    f.s = sinh(w) / w;
    f.t = 0.5 * q * q;
  }
  return f;
}

struct StepIntegrals {
  double cc, cs, sc, ss;
};

__host__ __device__ inline StepIntegrals step_integrals(double zi, double zj, double h) {
  StepIntegrals r;
  const double ai = fabs(zi), aj = fabs(zj);
  const double zmax = ai > aj ? ai : aj;
  if (zmax <= 2.0) {
    // powers (-z)^m/(2m)! and (-z)^m/(2m+1)! of both channels
    constexpr int K = 14;
    double ci[K], si[K], cj[K], sj[K];
    double pi = 1.0, pj = 1.0;
    for (int m = 0; m < K; ++m) {
      ci[m] = pi;
      si[m] = pi / (double)(2 * m + 1);
      cj[m] = pj;
      sj[m] = pj / (double)(2 * m + 1);
      pi = -zi * si[m] / (double)(2 * m + 2);
      pj = -zj * sj[m] / (double)(2 * m + 2);
    }
    double cc = 0.0, cs = 0.0, sc = 0.0, ss = 0.0;
    for (int m = K - 1; m >= 0; --m)
      for (int n = K - 1 - m; n >= 0; --n) {
        const double p = (double)(2 * (m + n));
        cc += ci[m] * cj[n] / (p + 1.0);
        cs += ci[m] * sj[n] / (p + 2.0);
        sc += si[m] * cj[n] / (p + 2.0);
        ss += si[m] * sj[n] / (p + 3.0);
      }
    r.cc = h * cc;
    r.cs = h * h * cs;
    r.sc = h * h * sc;
    r.ss = h * h * h * ss;
    return r;
  }
  const bool same_sign = (zi > 0.0) == (zj > 0.0);
  const double gap = fabs(zi - zj);
  if (!same_sign || gap >= 0.5 * zmax) {
    const StumpffCST fi = stumpff_cst(zi), fj = stumpff_cst(zj);
    const double inv = 1.0 / (zi - zj);
    r.cc = h * (zi * fi.s * fj.c - zj * fj.s * fi.c) * inv;
    r.ss = h * h * h * (fi.s * fj.c - fi.c * fj.s) * inv;
    r.cs = -h * h * (1.0 - fi.c * fj.c - zi * fi.s * fj.s) * inv;
    r.sc = h * h * (1.0 - fi.c * fj.c - zj * fi.s * fj.s) * inv;
    return r;
  }
  // same sign, close: both |x| >= 1 here
  const double sg = zi > 0.0 ? 1.0 : -1.0;
  const double xi = sqrt(ai), xj = sqrt(aj);
  const double d = xi - xj, s = xi + xj;
  const StumpffCST fd = stumpff_cst(sg * d * d), fs = stumpff_cst(sg * s * s);
  r.cc = 0.5 * h * (fd.s + fs.s);
  r.ss = sg * 0.5 * h * h * h * (fd.s - fs.s) / (xi * xj);
  r.cs = 0.5 * h * h * (s * fs.t - d * fd.t) / xj;
  r.sc = 0.5 * h * h * (s * fs.t + d * fd.t) / xi;
  return r;
}

}  // namespace mcp
