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
//
// With the centrifugal operator among the terms (NRSQ > 0, k_ytrans_rsq)
// the channels are not grouped by L: the operators are the constant one C
// (when NCONST > 0) and L^2 = sum_r VL(MXLAM+NCONST+r) (ytrans.f90:95-140,
// P = 1 for the NRSQ terms), and MULTOP's recursion (source_code/
// multop.f90:55-180) resolves them: C diagonalised over all channels, L^2
// turned into C's eigenbasis and diagonalised within each set of C
// eigenvalues closer than DEGTOL to their neighbour (elements between the
// sets zero, so each eigenvector stays in its set), the sets in ascending C
// order and ascending L^2 within a set; the total eigenvectors V_C V_L with
// the largest component positive; EINT from C's eigenvalues and CENT from
// L^2's (ytrans.f90:264-275,353-361,411-430).  Without constant terms L^2
// alone is diagonalised and EINT stays as it is.
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

__global__ void __launch_bounds__(kNT) k_ytrans_rsq(BlockArgs B, const double* __restrict__ pconst,
                                                    double* __restrict__ y, double* __restrict__ eint_out,
                                                    double* __restrict__ cent_out, double degtol) {
  extern __shared__ double smem[];
  const int n = B.n, lda = n;
  const size_t nn = (size_t)n * n;
  const int inst = blockIdx.x;
  LdsCarve c(smem);
  double* A = c.take<double>(nn);  // operators, then scratch
  double* V = c.take<double>(nn);  // eigenvectors of the first operator, then Y
  double* T = c.take<double>(nn);  // L^2 in V's basis, then EVEC
  double* W = c.take<double>(nn);  // eigenvectors of L^2 within the sets
  const int half = (n + 1) / 2 + 1;
  double* rc = c.take<double>(half);
  double* rs = c.take<double>(half);
  int* rp = c.take<int>(half);
  int* rq = c.take<int>(half);
  double* sv = c.take<double>(kNT / 64);
  double* ev = c.take<double>(n);
  double* e1 = c.take<double>(n);
  double* e2 = c.take<double>(n);
  int* order = c.take<int>(n);
  double* tmp = c.take<double>(n);
  int* setid = c.take<int>(n);
  int* pos = c.take<int>(n);
  const bool two = B.nconst > 0;
  const double* pc = two ? pconst + (size_t)inst * B.nconst : nullptr;
  const int rsq0 = B.mxlam + B.nconst;
  // the first operator: C, or L^2 when there are no constant terms
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    const double* v = B.vl + tri_index(i, j) * (size_t)B.nham;
    double s = 0.0;
    if (two)
      for (int k = 0; k < B.nconst; ++k) s += v[B.mxlam + k] * pc[k];
    else
      for (int k = 0; k < B.nrsq; ++k) s += v[rsq0 + k];
    A[i + j * lda] = s;
  }
  __syncthreads();
  jacobi_eigh<kNT>(A, V, n, lda, rc, rs, rp, rq, sv);
  __syncthreads();
  sort_eig<kNT>(A, V, n, lda, ev, order, tmp);
  __syncthreads();
  if (!two) {
    for (int e = threadIdx.x; e < n * n; e += kNT) T[e] = V[e];
    for (int i = threadIdx.x; i < n; i += kNT) {
      cent_out[(size_t)inst * n + i] = ev[i];
      eint_out[(size_t)inst * n + i] = B.eint[i];
    }
  } else {
    for (int i = threadIdx.x; i < n; i += kNT) e1[i] = ev[i];
    // L^2 in C's eigenbasis: W = V^T L^2 V
    for (int e = threadIdx.x; e < n * n; e += kNT) {
      const int j = e / n, i = e - j * n;
      const double* v = B.vl + tri_index(i, j) * (size_t)B.nham;
      double s = 0.0;
      for (int k = 0; k < B.nrsq; ++k) s += v[rsq0 + k];
      T[i + j * lda] = s;
    }
    __syncthreads();
    lds_gemm(T, false, V, false, A, n);
    __syncthreads();
    lds_gemm(V, true, A, false, W, n);
    __syncthreads();
    // the sets of C eigenvalues within DEGTOL of their neighbour
    if (threadIdx.x == 0) {
      setid[0] = 0;
      for (int p = 1; p < n; ++p) setid[p] = setid[p - 1] + (fabs(e1[p] - e1[p - 1]) > degtol ? 1 : 0);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += kNT) {
      const int j = e / n, i = e - j * n;
      if (setid[i] != setid[j]) {
        W[i + j * lda] = 0.0;
      } else if (i < j) {
        const double s = 0.5 * (W[i + j * lda] + W[j + i * lda]);
        W[i + j * lda] = s;
        W[j + i * lda] = s;
      }
    }
    __syncthreads();
    jacobi_eigh<kNT>(W, T, n, lda, rc, rs, rp, rq, sv);
    __syncthreads();
    // eigenvector m belongs to the set of its largest component; within a
    // set, ascending L^2 takes the set's positions in order
    if (threadIdx.x == 0) {
      for (int m = 0; m < n; ++m) {
        int imax = 0;
        double big = -1.0;
        for (int i = 0; i < n; ++i) {
          const double a = fabs(T[i + m * lda]);
          if (a > big) {
            big = a;
            imax = i;
          }
        }
        order[m] = setid[imax];
      }
      for (int m = 0; m < n; ++m) {
        const double wm = W[m + m * lda];
        int rank = 0, first = -1;
        for (int p = 0; p < n; ++p)
          if (setid[p] == order[m]) {
            if (first < 0) first = p;
          }
        for (int k = 0; k < n; ++k) {
          if (order[k] != order[m]) continue;
          const double wk = W[k + k * lda];
          if (wk < wm || (wk == wm && k < m)) ++rank;
        }
        pos[m] = first + rank;
        e2[first + rank] = wm;
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += kNT) {
      const int m = e / n, i = e - m * n;
      A[i + pos[m] * lda] = T[i + m * lda];
    }
    __syncthreads();
    // the total eigenvectors V_C V_L
    lds_gemm(V, false, A, false, T, n);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kNT) {
      cent_out[(size_t)inst * n + i] = e2[i];
      eint_out[(size_t)inst * n + i] = e1[i];
    }
  }
  __syncthreads();
  // the largest component (the first of equals) of every vector positive
  for (int j = threadIdx.x; j < n; j += kNT) {
    double big = 0.0;
    for (int i = 0; i < n; ++i)
      if (fabs(T[i + j * lda]) > fabs(big)) big = T[i + j * lda];
    tmp[j] = big < 0.0 ? -1.0 : 1.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += kNT) T[e] *= tmp[e / n];
  __syncthreads();
  if (y == nullptr) return;
  double* yg = y + (size_t)inst * nn;
  for (size_t e = threadIdx.x; e < nn; e += kNT) V[e] = yg[e];
  __syncthreads();
  lds_gemm(V, true, T, false, A, n);
  __syncthreads();
  lds_gemm(A, true, T, false, V, n);
  __syncthreads();
  for (size_t e = threadIdx.x; e < nn; e += kNT) yg[e] = V[e];
}

}  // namespace mcp
