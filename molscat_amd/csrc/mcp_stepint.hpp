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
//
// With the derivative perturbations (IVP, IVPP) the perturbation over the
// step is V0 + V1 (r - h/2) + V2 (r - h/2)^2 / 2, so the same four products
// are also integrated against the weights w1 = r - h/2 and w2 = (r - h/2)^2/2
// (step_integrals_w).  In the scaled variable t = r/h, c_i(t) = cos(x_i t),
// s_i(t) = sin(x_i t)/x_i (x^2 = z), the three evaluations become:
//   * both |z| <= 2: the double series with int_0^1 t^P w(t) dt in closed
//     form, P/(2(P+1)(P+2)) for w1 and (P^2+P+2)/(8(P+1)(P+2)(P+3)) for w2;
//   * apart: integration by parts on the Wronskian form: for solutions f
//     (of z_i) and g (of z_j), f g = (f'g - f g')'/(z_j - z_i), so
//     int w f g = ([w (f'g - f g')]_0^1 - int w' (f'g - f g'))/(z_j - z_i),
//     where w' lowers the weight and f' stays in the span of c and s
//     (c' = -z s, s' = c): each weight's four integrals from the previous
//     weight's;
//   * same sign and close: the products turned into sums as above; about the
//     step centre u = t - 1/2, cos(a t) = cos(a u) cos(a/2) - sin(a u) sin(a/2),
//     so every weighted integral is a product of cos(a/2) or sin(a/2) with
//     one of two symmetric moments, O1(a) = int u sin(a u) du and
//     E2(a) = int u^2 cos(a u) du over [-1/2, 1/2] (their hyperbolic forms for
//     closed pairs), each a power series in (a/2)^2 for |a/2| <= 2 and its
//     closed form beyond.
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

// Weighted step integrals: w1 = (r - h/2) and w2 = (r - h/2)^2 / 2 over
// the step (the derivative perturbations of VIVS, the quantities PERT1 and
// PERT2 form from V1 and V2, source_code/pert2.f:111-132,146-180).
struct StepIntegralsW {
  StepIntegrals w1, w2;
};

// symmetric moments over u in [-1/2, 1/2] at y = a/2: O1 = int u sin(a u)
// (sinh), E2 = int u^2 cos(a u) (cosh); sg = +1 trigonometric, -1 hyperbolic
__host__ __device__ inline void stepint_moments(double y, double sg, double& o1, double& e2) {
  const double q = y * y;
  if (fabs(y) <= 2.0) {
    // O1 = (y/2) sum (-sg q)^k / ((2k+1)! (2k+3)), E2 = (1/4) sum (-sg q)^k / ((2k)! (2k+3))
    double pe = 1.0;  // (-sg q)^k / (2k)!
    double so = 0.0, se = 0.0;
    for (int k = 0; k < 15; ++k) {
      const double po = pe / (double)(2 * k + 1);
      se += pe / (double)(2 * k + 3);
      so += po / (double)(2 * k + 3);
      pe = -sg * q * po / (double)(2 * k + 2);
    }
    o1 = 0.5 * y * so;
    e2 = 0.25 * se;
    return;
  }
  if (sg > 0.0) {
    const double sy = sin(y), cy = cos(y);
    o1 = (sy - y * cy) / (2.0 * q);
    e2 = sy / (4.0 * y) + cy / (2.0 * q) - sy / (2.0 * q * y);
  } else {
    const double sy = sinh(y), cy = cosh(y);
    o1 = (y * cy - sy) / (2.0 * q);
    e2 = sy / (4.0 * y) - cy / (2.0 * q) + sy / (2.0 * q * y);
  }
}

// int_0^1 cos(a t) w(t) dt and int_0^1 sin(a t) w(t) dt (cosh, sinh for
// sg = -1) for w1 = t - 1/2 and w2 = (t - 1/2)^2 / 2
struct StepMoments {
  double c1, s1, c2, s2;
};

__host__ __device__ inline StepMoments stepint_trig_moments(double a, double sg) {
  const double y = 0.5 * a;
  double o1, e2;
  stepint_moments(y, sg, o1, e2);
  StepMoments m;
  if (sg > 0.0) {
    const double sy = sin(y), cy = cos(y);
    m.c1 = -sy * o1;
    m.s1 = cy * o1;
    m.c2 = 0.5 * cy * e2;
    m.s2 = 0.5 * sy * e2;
  } else {
    const double sy = sinh(y), cy = cosh(y);
    m.c1 = sy * o1;
    m.s1 = cy * o1;
    m.c2 = 0.5 * cy * e2;
    m.s2 = 0.5 * sy * e2;
  }
  return m;
}

__host__ __device__ inline StepIntegralsW step_integrals_w(double zi, double zj, double h) {
  StepIntegralsW r;
  StepIntegrals& a = r.w1;
  StepIntegrals& b = r.w2;
  const double ai = fabs(zi), aj = fabs(zj);
  const double zmax = ai > aj ? ai : aj;
  const double h2 = h * h, h3 = h2 * h, h4 = h3 * h, h5 = h4 * h;
  if (zmax <= 2.0) {
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
    // weights of t^P: P/(2(P+1)(P+2)) and (P^2+P+2)/(8(P+1)(P+2)(P+3))
    auto w1 = [](double p) { return p / (2.0 * (p + 1.0) * (p + 2.0)); };
    auto w2 = [](double p) { return (p * p + p + 2.0) / (8.0 * (p + 1.0) * (p + 2.0) * (p + 3.0)); };
    double cc1 = 0.0, cs1 = 0.0, sc1 = 0.0, ss1 = 0.0, cc2 = 0.0, cs2 = 0.0, sc2 = 0.0, ss2 = 0.0;
    for (int m = K - 1; m >= 0; --m)
      for (int n = K - 1 - m; n >= 0; --n) {
        const double p = (double)(2 * (m + n));
        const double tcc = ci[m] * cj[n], tcs = ci[m] * sj[n], tsc = si[m] * cj[n], tss = si[m] * sj[n];
        cc1 += tcc * w1(p);
        cs1 += tcs * w1(p + 1.0);
        sc1 += tsc * w1(p + 1.0);
        ss1 += tss * w1(p + 2.0);
        cc2 += tcc * w2(p);
        cs2 += tcs * w2(p + 1.0);
        sc2 += tsc * w2(p + 1.0);
        ss2 += tss * w2(p + 2.0);
      }
    a.cc = h2 * cc1;
    a.cs = h3 * cs1;
    a.sc = h3 * sc1;
    a.ss = h4 * ss1;
    b.cc = h3 * cc2;
    b.cs = h4 * cs2;
    b.sc = h4 * sc2;
    b.ss = h5 * ss2;
    return r;
  }
  const bool same_sign = (zi > 0.0) == (zj > 0.0);
  const double gap = fabs(zi - zj);
  if (!same_sign || gap >= 0.5 * zmax) {
    // scaled solutions' values at t = 1 and the Wronskian forms
    // Om = f'g - f g' of the four products at t = 1 and t = 0
    const StumpffCST fi = stumpff_cst(zi), fj = stumpff_cst(zj);
    const double inv = 1.0 / (zj - zi);
    const double om1cc = -zi * fi.s * fj.c + zj * fi.c * fj.s, om0cc = 0.0;
    const double om1cs = -zi * fi.s * fj.s - fi.c * fj.c, om0cs = -1.0;
    const double om1sc = fi.c * fj.c + zj * fi.s * fj.s, om0sc = 1.0;
    const double om1ss = fi.c * fj.s - fi.s * fj.c, om0ss = 0.0;
    // weight 1
    const double j0cc = (om1cc - om0cc) * inv, j0cs = (om1cs - om0cs) * inv;
    const double j0sc = (om1sc - om0sc) * inv, j0ss = (om1ss - om0ss) * inv;
    const double j1cc = (0.5 * (om1cc + om0cc) - (-zi * j0sc + zj * j0cs)) * inv;
    const double j1cs = (0.5 * (om1cs + om0cs) - (-zi * j0ss - j0cc)) * inv;
    const double j1sc = (0.5 * (om1sc + om0sc) - (j0cc + zj * j0ss)) * inv;
    const double j1ss = (0.5 * (om1ss + om0ss) - (j0cs - j0sc)) * inv;
    // weight 2: w2(1) = w2(0) = 1/8, w2' = w1
    const double j2cc = (0.125 * (om1cc - om0cc) - (-zi * j1sc + zj * j1cs)) * inv;
    const double j2cs = (0.125 * (om1cs - om0cs) - (-zi * j1ss - j1cc)) * inv;
    const double j2sc = (0.125 * (om1sc - om0sc) - (j1cc + zj * j1ss)) * inv;
    const double j2ss = (0.125 * (om1ss - om0ss) - (j1cs - j1sc)) * inv;
    a.cc = h2 * j1cc;
    a.cs = h3 * j1cs;
    a.sc = h3 * j1sc;
    a.ss = h4 * j1ss;
    b.cc = h3 * j2cc;
    b.cs = h4 * j2cs;
    b.sc = h4 * j2sc;
    b.ss = h5 * j2ss;
    return r;
  }
  // same sign, close: products to sums in d = x_i - x_j and s = x_i + x_j
  const double sg = zi > 0.0 ? 1.0 : -1.0;
  const double xi = sqrt(ai), xj = sqrt(aj);
  const StepMoments md = stepint_trig_moments(xi - xj, sg), ms = stepint_trig_moments(xi + xj, sg);
  const double rij = sg * 0.5 / (xi * xj);
  a.cc = h2 * 0.5 * (md.c1 + ms.c1);
  a.ss = h4 * rij * (md.c1 - ms.c1);
  a.cs = h3 * 0.5 * (ms.s1 - md.s1) / xj;
  a.sc = h3 * 0.5 * (ms.s1 + md.s1) / xi;
  b.cc = h3 * 0.5 * (md.c2 + ms.c2);
  b.ss = h5 * rij * (md.c2 - ms.c2);
  b.cs = h4 * 0.5 * (ms.s2 - md.s2) / xj;
  b.sc = h4 * 0.5 * (ms.s2 + md.s2) / xi;
  return r;
}

}  // namespace mcp
