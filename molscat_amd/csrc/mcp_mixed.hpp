// Mixed-precision batched inversions for the large-N tree (the configs with
// N >= 200): each matrix A (FP64) is copied to FP32, inverted there by the
// 32-pivot kernel (k_syminv_wide<float>, which also gives the inertia from
// its FP32 pivots), and the inverse is won back to FP64 by Newton-Schulz
// steps X <- X (2I - A X) on the FP64 batched products.  The first step's
// residual max|I - A X0| decides whether the FP32 result is trusted: at
// most kMixedTrust it is (the condition number is then below ~1e-2/eps_32, so
// no eigenvalue can have changed sign under the FP32 rounding and the
// inertia is exact); above it the matrix is inverted again in FP64.
// Matrices flagged beforehand (the sectors below a given radius: the
// short-range start, where closed channels grow exponentially) are FP64
// throughout.
#pragma once
#include <hip/hip_runtime.h>

namespace mcp {

constexpr double kMixedTrust = 1e-2;

// The batch's matrices arrive with their lower triangles only: a gets its
// upper triangle mirrored from the lower (the Newton products read A in
// full), a32 = (float) a; flag[m] = 1 (FP64) for (m % per) < cut, else 0.
__global__ void k_mixed_down(int n, int nb, int per, int cut, double* __restrict__ a, float* __restrict__ a32,
                             int* __restrict__ flag) {
  const size_t nn = (size_t)n * n;
  const size_t tot = nn * nb;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const size_t m = e / nn, r = e - m * nn;
    const int j = (int)(r / n), i = (int)(r - (size_t)j * n);
    const double v = i >= j ? a[e] : a[m * nn + j + (size_t)i * n];
    if (i < j) a[e] = v;
    a32[e] = (float)v;
    if (r == 0) {
      flag[m] = ((int)m % per) < cut ? 1 : 0;
    }
  }
}

// x = (double) x32 for the matrices with flag 0
__global__ void k_mixed_up(int n, int nb, const float* __restrict__ x32, double* __restrict__ x,
                           const int* __restrict__ flag) {
  const size_t nn = (size_t)n * n;
  const size_t tot = nn * nb;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x)
    if (flag[e / nn] == 0) x[e] = (double)x32[e];
}

// one workgroup per matrix: flag 2 (FP64 again) where max|I - T| > tol
__global__ void __launch_bounds__(256) k_mixed_resid(int n, const double* __restrict__ t, int* __restrict__ flag,
                                                     double tol) {
  __shared__ double sm[4];
  const int m = blockIdx.x;
  if (flag[m] != 0) return;
  const double* T = t + (size_t)m * n * n;
  double r = 0.0;
  for (int e = threadIdx.x; e < n * n; e += 256) {
    const int j = e / n, i = e - j * n;
    r = fmax(r, fabs((i == j ? 1.0 : 0.0) - T[e]));
  }
  for (int off = 32; off > 0; off >>= 1) r = fmax(r, __shfl_xor(r, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double mx = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
    if (!(mx <= tol)) flag[m] = 2;
  }
}

// a = (x + x^T)/2 for the matrices with flag 0 (the others hold their FP64
// inverse already); inv64 counts the matrices inverted in FP64
__global__ void k_mixed_finish(int n, int nb, const double* __restrict__ x, double* __restrict__ a,
                               const int* __restrict__ flag, int* __restrict__ inv64) {
  const size_t nn = (size_t)n * n;
  const size_t tot = nn * nb;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const size_t m = e / nn, r = e - m * nn;
    if (flag[m] != 0) {
      if (r == 0 && inv64) atomicAdd(inv64, 1);
      continue;
    }
    const int j = (int)(r / n), i = (int)(r - (size_t)j * n);
    const double* X = x + m * nn;
    a[e] = 0.5 * (X[i + (size_t)j * n] + X[j + (size_t)i * n]);
  }
}

}  // namespace mcp
