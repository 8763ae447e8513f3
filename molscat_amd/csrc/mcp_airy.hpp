// Airy functions and the single-channel Airy sector propagator of the AIRY
// propagator (AIPROP/SPROPN, source_code/spropn.f:82-258), host and device.
//
// The sector propagator.  In a sector of width w whose reference W is
// -eig + slope*r about the midpoint, the solutions are Ai(x), Bi(x) with
// x = alpha (r + beta), alpha^3 = slope, beta = -eig/slope; with x1, x2 the
// ends and D = Ai(x1)Bi(x2) - Ai(x2)Bi(x1) the sector's log-derivative
// imbedding elements are
//   y1 = alpha (Bi'(x1)Ai(x2) - Ai'(x1)Bi(x2)) / D,   y2 = alpha / (pi D),
//   y4 = alpha (Ai(x1)Bi'(x2) - Bi(x1)Ai'(x2)) / D.
// Where |x| is large (SPROPN's test log|eig| - 2/3 log|slope| > 3) and the
// sector is long against the local wavelength, the three products are
// formed from the asymptotic expansions (Abramowitz & Stegun
// 10.4.59-10.4.68) with the phase difference zeta2 - zeta1 computed from
// delta/x1 by log1p/expm1, so no cancellation; exponential factors are
// carried separately (D, y1, y4 share e^|zeta2 - zeta1|, y2 gets its
// inverse), so nothing overflows.  Elsewhere no Airy function is needed: with
// Phi the transfer matrix of y'' = x y from x1 to x2 (det Phi = 1) and the
// Wronskian 1/pi, D = Phi12/pi and the two numerators are Phi11/pi and
// Phi22/pi, so y1 = alpha Phi11/Phi12, y2 = alpha/Phi12, y4 = alpha
// Phi22/Phi12; Phi comes from Taylor re-expansion of the equation, which
// gives Phi12 = delta + ... without cancellation however narrow the sector.
#pragma once
#include <hip/hip_runtime.h>

#include <math.h>

namespace mcp {

constexpr double kPi = 3.14159265358979323846;

// u_k, v_k of the asymptotic expansions: u_k = u_{k-1} (6k-5)(6k-1)/(72k),
// v_k = -(6k+1)/(6k-1) u_k.  sum_k c_k z^{-k} over |z| >= 18.
__host__ __device__ inline void airy_asym_sums(double z, double& lsum, double& psum) {
  double u = 1.0, l = 1.0, p = 1.0, zk = 1.0;
  const double zi = 1.0 / z;
  for (int k = 1; k < 60; ++k) {
    u *= (6.0 * k - 5.0) * (6.0 * k - 1.0) / (72.0 * k);
    zk *= zi;
    const double tl = u * zk;
    const double tp = -(6.0 * k + 1.0) / (6.0 * k - 1.0) * tl;
    l += tl;
    p += tp;
    if (fabs(tl) < 1e-18 * fabs(l) && fabs(tp) < 1e-18 * fabs(p)) break;
  }
  lsum = l;
  psum = p;
}

// even and odd parts with alternating signs (the oscillatory expansions):
// q1 = sum (-1)^k u_2k z^-2k, q2 = sum (-1)^k u_2k+1 z^-(2k+1), r1, r2 the same with v
__host__ __device__ inline void airy_asym_osc(double z, double& q1, double& q2, double& r1, double& r2) {
  double u = 1.0, zk = 1.0;
  const double zi = 1.0 / z;
  q1 = 1.0;
  q2 = 0.0;
  r1 = 1.0;
  r2 = 0.0;
  for (int k = 1; k < 60; ++k) {
    u *= (6.0 * k - 5.0) * (6.0 * k - 1.0) / (72.0 * k);
    zk *= zi;
    const double v = -(6.0 * k + 1.0) / (6.0 * k - 1.0) * u;
    const int m = k >> 1;
    const double sg = (m & 1) ? -1.0 : 1.0;
    const double tu = sg * u * zk, tv = sg * v * zk;
    if (k & 1) {
      q2 += tu;
      r2 += tv;
    } else {
      q1 += tu;
      r1 += tv;
    }
    if (fabs(u * zk) < 1e-18 && k > 2) break;
  }
}

// Advance (y, y') of y'' = x y from x0 to x0 + h by its Taylor series:
// a_{n+2} = (x0 a_n + a_{n-1}) / ((n+1)(n+2)).
__host__ __device__ inline void airy_taylor(double x0, double h, double& y, double& yp) {
  double am1 = 0.0, a0 = y, a1 = yp;
  double sy = a0 + a1 * h, syp = a1;
  double hn = h;  // h^(n+1)
  int small = 0;
  for (int n = 0; n < 120; ++n) {
    const double a2 = (x0 * a0 + am1) / ((n + 1.0) * (n + 2.0));
    const double ty = a2 * hn * h;           // a_{n+2} h^{n+2}
    const double typ = (n + 2.0) * a2 * hn;  // (n+2) a_{n+2} h^{n+1}
    sy += ty;
    syp += typ;
    hn *= h;
    am1 = a0;
    a0 = a1;
    a1 = a2;
    if (fabs(ty) <= 1e-18 * fabs(sy) && fabs(typ) <= 1e-18 * fabs(syp)) {
      if (++small >= 3) break;
    } else {
      small = 0;
    }
  }
  y = sy;
  yp = syp;
}

// Transfer matrix of y'' = x y from x1 to x1 + delta (columns: the solutions
// with (y, y')(x1) = (1, 0) and (0, 1)), in steps short against the local
// scale.  The width delta is taken as given, not as the difference of two
// rounded end points.
__host__ __device__ inline void airy_transfer(double x1, double delta, double& p11, double& p12, double& p21,
                                              double& p22) {
  const double xm = fmax(fabs(x1), fabs(x1 + delta));
  const double scale = xm > 1.0 ? sqrt(xm) : 1.0;
  const int ns = (int)ceil(fabs(delta) * scale / 0.5);
  double ya = 1.0, ypa = 0.0, yb = 0.0, ypb = 1.0;
  if (ns > 0) {
    const double h = delta / ns;
    for (int s = 0; s < ns; ++s) {
      const double x = x1 + s * h;
      airy_taylor(x, h, ya, ypa);
      airy_taylor(x, h, yb, ypb);
    }
  }
  p11 = ya;
  p21 = ypa;
  p12 = yb;
  p22 = ypb;
}

// y1, y2, y4 of one channel (see the file comment).  slope == 0 is treated
// as slope = 1e-30, as SPROPN does.
__host__ __device__ inline void airy_sector(double width, double eig, double slope, double& y1, double& y2,
                                            double& y4) {
  double darg = 1.0e10;
  if (slope != 0.0) darg = log(fabs(eig)) - 2.0 / 3.0 * log(fabs(slope));
  if (slope == 0.0) slope = 1.0e-30;
  const double alpha = cbrt(slope);
  const double beta = -eig / slope;
  const double x1 = alpha * (beta - 0.5 * width);
  const double x2 = alpha * (beta + 0.5 * width);
  const double delta = alpha * width;
  double d = 1.0, n1 = 0.0, n4 = 0.0, dz = 0.0;  // products over e^|dz|, dz = zeta2 - zeta1
  // the transfer matrix needs |delta| sqrt|x| / (1/2) Taylor steps: short
  // sectors take it at any |x|, long ones at large |x| the asymptotic forms
  const double xm = fmax(fabs(x1), fabs(x2));
  const double nsteps = fabs(delta) * (xm > 1.0 ? sqrt(xm) : 1.0) / 0.5;
  if (darg > 3.0 && x1 != 0.0 && fabs(delta / x1) < 0.5 && nsteps > 32.0) {
    if (x1 > 0.0) {
      const double t = delta / x1;
      const double z1 = 2.0 / 3.0 * x1 * sqrt(x1);
      dz = z1 * expm1(1.5 * log1p(t));
      const double z2 = z1 + dz;
      const double rho = exp(0.25 * log1p(t));
      double la, pa, lb, pb, lc, pc, ld, pd;
      airy_asym_sums(-z1, la, pa);
      airy_asym_sums(z1, lb, pb);
      airy_asym_sums(z2, lc, pc);
      airy_asym_sums(-z2, ld, pd);
      const double e = exp(-2.0 * fabs(dz));
      const double ep = dz >= 0.0 ? 1.0 : e, em = dz >= 0.0 ? e : 1.0;  // e^{dz}, e^{-dz} over e^|dz|
      d = (ep * la * lc - em * lb * ld) / (2.0 * kPi * sqrt(x1) * rho);
      n1 = (em * pb * ld + ep * pa * lc) / (2.0 * kPi * rho);
      n4 = rho * (ep * la * pc + em * lb * pd) / (2.0 * kPi);
    } else {
      const double zz1 = -x1;
      const double t = -delta / zz1;
      const double z1 = 2.0 / 3.0 * zz1 * sqrt(zz1);
      dz = z1 * expm1(1.5 * log1p(t));
      const double z2 = z1 + dz;
      const double rho = exp(0.25 * log1p(t));
      double q1a, q2a, r1a, r2a, q1b, q2b, r1b, r2b;
      airy_asym_osc(z1, q1a, q2a, r1a, r2a);
      airy_asym_osc(z2, q1b, q2b, r1b, r2b);
      const double s = -sin(dz), c = cos(dz);  // sin, cos of (xi1 - xi2)
      d = (s * (q1a * q1b + q2a * q2b) + c * (q1a * q2b - q2a * q1b)) / (kPi * sqrt(zz1) * rho);
      n1 = (c * (r1a * q1b + r2a * q2b) - s * (r1a * q2b - r2a * q1b)) / (kPi * rho);
      n4 = rho * (c * (q1a * r1b + q2a * r2b) - s * (q1a * r2b - q2a * r1b)) / kPi;
      dz = 0.0;  // oscillatory: no exponential factor
    }
  } else {
    double p11, p12, p21, p22;
    airy_transfer(x1, delta, p11, p12, p21, p22);
    y1 = alpha * p11 / p12;
    y2 = alpha / p12;
    y4 = alpha * p22 / p12;
    return;
  }
  y1 = alpha * n1 / d;
  y4 = alpha * n4 / d;
  y2 = alpha * exp(-fabs(dz)) / (kPi * d);
}

}  // namespace mcp
