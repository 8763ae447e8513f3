// YTRANS (source_code/ytrans.f90:1-485) for blocks with NCONST > 0 and
// NRSQ = 0 on the MOLSCAT path (one operator to diagonalise, channels
// grouped by their orbital L).
//
// Per instance: the asymptotic operator C = sum_c VL(MXLAM+c) P_c with the
// instance's constants P_c (ytrans.f90:116-140, the P of HAMMAT's constant
// terms), diagonalised within each set of channels of equal CENT (elements
// between sets are not used, ytrans.f90:147-215); every eigenvector's
// largest component (the first of equals) made positive (:262-275); the
// icol-th eigenvector of a set, in ascending eigenvalue order, is the
// asymptotic channel at the set's icol-th channel position, with EINT its
// eigenvalue (:330-350); then Y <- EVEC^T Y EVEC (TRNSFM with neither
// early return nor symmetrisation, source_code/trnsfm.f:20-32).
//
// The whole operator is diagonalised at once by the LDS Jacobi solver with
// the elements between sets set to zero: a rotation of two channels of one
// set leaves every zero between sets exactly zero, so each eigenvector stays
// inside its set, and the sets' eigenpairs are read off in the global
// ascending order.  One workgroup per instance, everything in LDS.
#pragma once
#include <hip/hip_runtime.h>

#include "mcp_airyprop.hpp"
#include "mcp_device.hpp"
#include "mcp_kernels.hpp"

namespace mcp {

__global__ void __launch_bounds__(kNT) k_ytrans(BlockArgs B, const double* __restrict__ pconst, double* __restrict__ y,
                                                double* __restrict__ eint_out) {
  extern __shared__ double smem[];
  const int n = B.n, lda = n;
  const size_t nn = (size_t)n * n;
  const int inst = blockIdx.x;
  LdsCarve c(smem);
  double* A = c.take<double>(nn);  // C, then scratch of the transformation
  double* V = c.take<double>(nn);  // eigenvectors, then Y
  double* T = c.take<double>(nn);  // EVEC
  const int half = (n + 1) / 2 + 1;
  double* rc = c.take<double>(half);
  double* rs = c.take<double>(half);
  int* rp = c.take<int>(half);
  int* rq = c.take<int>(half);
  double* sv = c.take<double>(kNT / 64);
  double* ev = c.take<double>(n);
  int* order = c.take<int>(n);
  double* tmp = c.take<double>(n);
  int* pos = c.take<int>(n);
  const double* pc = pconst + (size_t)inst * B.nconst;
  // C within the sets of equal CENT, zero between them
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    double s = 0.0;
    if (B.cent[i] == B.cent[j]) {
      const double* v = B.vl + tri_index(i, j) * (size_t)B.nham + B.mxlam;
      for (int k = 0; k < B.nconst; ++k) s += v[k] * pc[k];
    }
    A[i + j * lda] = s;
  }
  __syncthreads();
  jacobi_eigh<kNT>(A, V, n, lda, rc, rs, rp, rq, sv);
  sort_eig<kNT>(A, V, n, lda, ev, order, tmp);
  __syncthreads();
  // eigenvector m (ascending) belongs to the set of its largest component;
  // it takes the set's next channel position in channel order
  if (threadIdx.x == 0) {
    for (int i = 0; i < n; ++i) order[i] = 0;  // order[]: position taken
    for (int m = 0; m < n; ++m) {
      int imax = 0;
      double big = -1.0;
      for (int i = 0; i < n; ++i) {
        const double a = fabs(V[i + m * lda]);
        if (a > big) {
          big = a;
          imax = i;
        }
      }
      const double cs = B.cent[imax];
      int p = 0;
      while (p < n && (order[p] || B.cent[p] != cs)) ++p;
      if (p == n) p = imax;  // unreachable: each set has as many vectors as channels
      order[p] = 1;
      pos[m] = p;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int m = e / n, i = e - m * n;
    T[i + pos[m] * lda] = V[i + m * lda];
  }
  for (int m = threadIdx.x; m < n; m += kNT) eint_out[(size_t)inst * n + pos[m]] = ev[m];
  __syncthreads();
  if (y == nullptr) return;
  double* yg = y + (size_t)inst * nn;
  for (size_t e = threadIdx.x; e < nn; e += kNT) V[e] = yg[e];
  __syncthreads();
  // A = Y^T T, then Y <- A^T T
  lds_gemm(V, true, T, false, A, n);
  __syncthreads();
  lds_gemm(A, true, T, false, V, n);
  __syncthreads();
  for (size_t e = threadIdx.x; e < nn; e += kNT) yg[e] = V[e];
}

}  // namespace mcp
