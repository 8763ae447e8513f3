// COUPLE for ITYPE = 1 (atom + linear rigid rotor, source_code/couple.f:
// 51-143): the Percival-Seaton coefficients
//   VL(lambda, packed(irow, icol)) = (-1)^(jc+jr+J) sqrt((2jc+1)(2jr+1)(2lc+1)(2lr+1))
//       (jc lambda jr; 0 0 0) (lc lambda lr; 0 0 0) {jr lr J; lc jc lambda}
// for every pair of channels (irow <= icol, the packed upper triangle column
// by column, lambda fastest), one thread per coefficient on the device; the
// same source builds for the host for the CPU tests.
//
// Method (written from the closed forms, not from any routine): the 3j
// symbols with zero projections by their factorial closed form, the 6j
// symbol by Racah's single sum.  Every factorial enters through ln(k!),
// tabulated as an unevaluated sum hi + lo of two doubles, and each term's
// logarithm is accumulated with exact two-sums, so the cancellation of the
// large logarithms (ln(320!) ~ 1.5e3) costs nothing: a term is exp(s) (1 + e)
// with s its rounded logarithm and e the accumulated error, good to a few
// ulp.  The alternating Racah sum has at most lambda + 1 terms.  The
// integer product under the square root is exact in a double.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace mcp {

// ln(k!) table: hi[k] + lo[k], k = 0..kmax
struct LnFact {
  const double* hi;
  const double* lo;
  int kmax;
};

struct DDSum {
  double s = 0.0, e = 0.0;
  __host__ __device__ inline void add(double hi, double lo) {
    const double t = s + hi;
    const double bb = t - s;
    e += ((s - (t - bb)) + (hi - bb)) + lo;
    s = t;
  }
};

__host__ __device__ inline void add_lnf(DDSum& a, const LnFact& f, int k, double w) {
  // w = +-1 or +-1/2: exact scalings of the pair
  a.add(w * f.hi[k], w * f.lo[k]);
}

__host__ __device__ inline bool triangle(int a, int b, int c) {
  return c >= (a > b ? a - b : b - a) && c <= a + b;
}

// ln|(a b c; 0 0 0)| into acc, its sign into *sign; false when it vanishes
__host__ __device__ inline bool threej000_log(const LnFact& f, int a, int b, int c, DDSum& acc, int* sign) {
  const int jsum = a + b + c;
  if ((jsum & 1) || !triangle(a, b, c)) return false;
  const int g = jsum / 2;
  add_lnf(acc, f, jsum - 2 * a, 0.5);
  add_lnf(acc, f, jsum - 2 * b, 0.5);
  add_lnf(acc, f, jsum - 2 * c, 0.5);
  add_lnf(acc, f, jsum + 1, -0.5);
  add_lnf(acc, f, g, 1.0);
  add_lnf(acc, f, g - a, -1.0);
  add_lnf(acc, f, g - b, -1.0);
  add_lnf(acc, f, g - c, -1.0);
  if (g & 1) *sign = -*sign;
  return true;
}

// 0.5 ln Delta(a b c) = 0.5 ln[(a+b-c)! (a-b+c)! (-a+b+c)! / (a+b+c+1)!]
__host__ __device__ inline void delta_log(const LnFact& f, int a, int b, int c, DDSum& acc) {
  add_lnf(acc, f, a + b - c, 0.5);
  add_lnf(acc, f, a - b + c, 0.5);
  add_lnf(acc, f, -a + b + c, 0.5);
  add_lnf(acc, f, a + b + c + 1, -0.5);
}

// {a b c; d e f} times exp(base) and the sign: Racah's sum
//   sum_t (-1)^t (t+1)! / [(t-a-b-c)! (t-a-e-f)! (t-d-b-f)! (t-d-e-c)!
//                          (a+b+d+e-t)! (a+c+d+f-t)! (b+c+e+f-t)!]
// times sqrt(Delta(abc) Delta(aef) Delta(dbf) Delta(dec)).
__host__ __device__ inline double sixj_scaled(const LnFact& f, int a, int b, int c, int d, int e, int ff,
                                              DDSum base, int sign) {
  if (!triangle(a, b, c) || !triangle(a, e, ff) || !triangle(d, b, ff) || !triangle(d, e, c)) return 0.0;
  delta_log(f, a, b, c, base);
  delta_log(f, a, e, ff, base);
  delta_log(f, d, b, ff, base);
  delta_log(f, d, e, c, base);
  const int a1 = a + b + c, a2 = a + e + ff, a3 = d + b + ff, a4 = d + e + c;
  const int b1 = a + b + d + e, b2 = a + c + d + ff, b3 = b + c + e + ff;
  int tmin = a1 > a2 ? a1 : a2;
  tmin = tmin > a3 ? tmin : a3;
  tmin = tmin > a4 ? tmin : a4;
  int tmax = b1 < b2 ? b1 : b2;
  tmax = tmax < b3 ? tmax : b3;
  double sum = 0.0;
  for (int t = tmin; t <= tmax; ++t) {
    DDSum x = base;
    add_lnf(x, f, t + 1, 1.0);
    add_lnf(x, f, t - a1, -1.0);
    add_lnf(x, f, t - a2, -1.0);
    add_lnf(x, f, t - a3, -1.0);
    add_lnf(x, f, t - a4, -1.0);
    add_lnf(x, f, b1 - t, -1.0);
    add_lnf(x, f, b2 - t, -1.0);
    add_lnf(x, f, b3 - t, -1.0);
    const double term = exp(x.s) * (1.0 + x.e);
    sum += (t & 1) ? -term : term;
  }
  return sign < 0 ? -sum : sum;
}

// ln(k!) for k = 0..kmax on the host: a compensated sum of logl(k) in the
// x87 extended format, split into hi + lo
inline void build_lnfact(int kmax, double* hi, double* lo) {
  long double s = 0.0L, comp = 0.0L;
  hi[0] = lo[0] = 0.0;
  for (int k = 1; k <= kmax; ++k) {
    const long double y = logl((long double)k) - comp;
    const long double t = s + y;
    comp = (t - s) - y;
    s = t;
    const double h = (double)s;
    hi[k] = h;
    lo[k] = (double)(s - (long double)h);
  }
}

// The largest k of ln(k!) that percival_seaton reads for channel quantum
// numbers up to jmax, lmax, total J and lambda up to lmx.
__host__ __device__ inline int couple_lnf_need(int jmax, int lmax, int jtot, int lmx) {
  // 6j: (t+1)! with t <= jr + lr + lc + jc; 3j: (a+b+c+1)!
  const int a = 2 * (jmax + lmax) + 2;
  const int b = 2 * (jmax > lmax ? jmax : lmax) + lmx + 2;
  const int c = jtot + jmax + lmax + 2;
  int m = a > b ? a : b;
  return m > c ? m : c;
}

__host__ __device__ inline double percival_seaton(const LnFact& f, int jc, int lc, int jr, int lr, int jtot, int lam) {
  if (lam < (jc > jr ? jc - jr : jr - jc) || lam < (lc > lr ? lc - lr : lr - lc)) return 0.0;
  DDSum acc;
  int sign = 1;
  if (!threej000_log(f, jc, lam, jr, acc, &sign)) return 0.0;
  if (!threej000_log(f, lc, lam, lr, acc, &sign)) return 0.0;
  const double v = sixj_scaled(f, jr, lr, jtot, lc, jc, lam, acc, sign);
  if (v == 0.0) return 0.0;
  const double p = (double)(2 * jc + 1) * (double)(2 * jr + 1) * (double)(2 * lc + 1) * (double)(2 * lr + 1);
  const double out = sqrt(p) * v;
  return ((jc + jr + jtot) & 1) ? -out : out;
}

// Column of the packed upper triangle that holds packed index k
// (k = icol (icol + 1) / 2 + irow, irow <= icol)
__host__ __device__ inline int packed_col(long long k) {
  int c = (int)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
  while ((long long)(c + 1) * (c + 2) / 2 <= k) ++c;
  while ((long long)c * (c + 1) / 2 > k) --c;
  return c;
}

#ifdef __HIPCC__
// one thread per (packed pair, lambda); vl[k * nlam + il], lambda fastest.
// jch/lch: the channels' rotor j and orbital L.
__global__ void __launch_bounds__(256) k_couple_rotor(const int* __restrict__ jch, const int* __restrict__ lch,
                                                      long long ntri, int jtot, const int* __restrict__ lam,
                                                      int nlam, LnFact f, double* __restrict__ vl) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ntri * nlam) return;
  const long long k = idx / nlam;
  const int il = (int)(idx - k * nlam);
  const int icol = packed_col(k);
  const int irow = (int)(k - (long long)icol * (icol + 1) / 2);
  vl[idx] = percival_seaton(f, jch[icol], lch[icol], jch[irow], lch[irow], jtot, lam[il]);
}
#endif

}  // namespace mcp
