// AIRY propagator (AIPROP, source_code/aiprop.f:86-366, TOLHI = 0) on the
// device, for N within the LDS-resident eigen-solver.
//
// Grid: point 0 is RSTART; step k (0-based) has its Gauss points RA_k, RB_k
// (RNOW -+ DRNOW/(2*1.732050807), source_code/potent.f:66-67) at points
// 1+2k, 2+2k, and its width DRNOW_k in a separate array.
//
// The local basis of a step does not depend on the energy: -W = ERED - W0,
// so POTENT's eigenvectors are those of -W0_avg and its eigenvalues shift by
// ERED.  k_airy_sector computes, once per grid and block, each step's basis
// V_k, eigenvalues of -W0_avg (ascending, DIAGVC) and the diagonal HP_k of
// sqrt(3)(W(RB)-W(RA))/DRNOW in that basis; k_airy_tmat the basis changes
// T_0 = V_0, T_k = V_{k-1}^T V_k (aiprop.f:154-161).  k_airy_prop then runs
// one workgroup per instance: Z <- sym(T^T Z T) (TRNSFM's DSYMM + DSYR2K,
// source_code/trnsfm.f:37-40), the sector propagator per channel (SPROPN,
// molscat_amd/csrc/mcp_airy.hpp), Z <- Y4 - Y2 (Z + Y1)^-1 Y2 with NODES +=
// KOUNT outward, N - KOUNT inward (aiprop.f:196-244), and finally back to
// the primitive basis with V_last (aiprop.f:359-366).
#pragma once
#include <hip/hip_runtime.h>

#include "mcp_airy.hpp"
#include "mcp_device.hpp"
#include "mcp_kernels.hpp"

namespace mcp {

constexpr double kPotentSq3 = 1.732050807;  // the constant of source_code/potent.f:49, as written

// one workgroup per step
__global__ void __launch_bounds__(kNT) k_airy_sector(BlockArgs B, GridArgs G, const double* __restrict__ drnow,
                                                     double* __restrict__ vout, double* __restrict__ eout,
                                                     double* __restrict__ hout) {
  extern __shared__ double smem[];
  const int n = B.n, lda = n;
  const int k = blockIdx.x;
  const int ia = 1 + 2 * k, ib = 2 + 2 * k;
  LdsCarve c(smem);
  double* A = c.take<double>((size_t)n * lda);
  double* V = c.take<double>((size_t)n * lda);
  const int half = (n + 1) / 2 + 1;
  double* rc = c.take<double>(half);
  double* rs = c.take<double>(half);
  int* rp = c.take<int>(half);
  int* rq = c.take<int>(half);
  double* sv = c.take<double>(kNT / 64);
  double* ev = c.take<double>(n);
  int* order = c.take<int>(n);
  double* tmp = c.take<double>(n);
  double* pa = c.take<double>(B.nham);
  double* pb = c.take<double>(B.nham);
  load_pvec(B, G, ia, nullptr, pa);
  load_pvec(B, G, ib, nullptr, pb);
  __syncthreads();
  const double ra = G.r[ia], rb = G.r[ib];
  const double rsa = rsq_of(ra), rsb = rsq_of(rb);
  // local wavevector matrix -(W(RA) + W(RB))/2 at zero energy
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    double wa = sumlam(B, i, j, pa), wb = sumlam(B, i, j, pb);
    if (i == j) {
      wa = hammat_diag(B, i, wa, ra, rsa, 0.0);
      wb = hammat_diag(B, i, wb, rb, rsb, 0.0);
    }
    A[i + j * lda] = -0.5 * (wa + wb);
  }
  __syncthreads();
  jacobi_eigh<kNT>(A, V, n, lda, rc, rs, rp, rq, sv);
  sort_eig<kNT>(A, V, n, lda, ev, order, tmp);
  __syncthreads();
  // dW/dR = sqrt(3)(W(RB) - W(RA))/DRNOW; its diagonal in the local basis
  const double fact = kPotentSq3 / drnow[k];
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    double wa = sumlam(B, i, j, pa), wb = sumlam(B, i, j, pb);
    if (i == j) {
      wa = hammat_diag(B, i, wa, ra, rsa, 0.0);
      wb = hammat_diag(B, i, wb, rb, rsb, 0.0);
    }
    A[i + j * lda] = fact * (wb - wa);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += kNT) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      double u = 0.0;
      for (int j = 0; j < n; ++j) u += A[i + j * lda] * V[j + q * lda];
      s += V[i + q * lda] * u;
    }
    hout[(size_t)k * n + q] = s;
    eout[(size_t)k * n + q] = ev[q];
  }
  double* vo = vout + (size_t)k * n * n;
  for (int e = threadIdx.x; e < n * n; e += kNT) vo[e] = V[e % n + (e / n) * lda];
}

// T_0 = V_0, T_k = V_{k-1}^T V_k; one workgroup per step
__global__ void __launch_bounds__(kNT) k_airy_tmat(int n, const double* __restrict__ v, double* __restrict__ t) {
  const int k = blockIdx.x;
  const size_t nn = (size_t)n * n;
  const double* vk = v + (size_t)k * nn;
  double* tk = t + (size_t)k * nn;
  if (k == 0) {
    for (size_t e = threadIdx.x; e < nn; e += kNT) tk[e] = vk[e];
    return;
  }
  const double* vp = v + (size_t)(k - 1) * nn;
  for (size_t e = threadIdx.x; e < nn; e += kNT) {
    const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
    double s = 0.0;
    for (int m = 0; m < n; ++m) s += vp[m + (size_t)i * n] * vk[m + (size_t)j * n];
    tk[e] = s;
  }
}

// C = op(A) op(B) for n x n matrices in LDS (leading dimension n), all threads
__device__ __forceinline__ void lds_gemm(const double* A, bool ta, const double* Bm, bool tb, double* C, int n) {
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    double s = 0.0;
    for (int m = 0; m < n; ++m) {
      const double a = ta ? A[m + i * n] : A[i + m * n];
      const double b = tb ? Bm[j + m * n] : Bm[m + j * n];
      s += a * b;
    }
    C[i + j * n] = s;
  }
}

// one workgroup per instance; outward = RSTOP > RSTART
__global__ void __launch_bounds__(kNT) k_airy_prop(int n, int nstep, const double* __restrict__ tmat,
                                                   const double* __restrict__ eig0, const double* __restrict__ hp,
                                                   const double* __restrict__ drnow, const double* __restrict__ vlast,
                                                   InstArgs I, int outward) {
  extern __shared__ double smem[];
  const int inst = blockIdx.x;
  const size_t nn = (size_t)n * n;
  LdsCarve c(smem);
  double* Z = c.take<double>(nn);
  double* A = c.take<double>(nn);
  double* T = c.take<double>(nn);
  double* y1 = c.take<double>(n);
  double* y2 = c.take<double>(n);
  double* y4 = c.take<double>(n);
  SweepScratch s = carve_sweep(c, n);
  const double ered = I.ered[inst];
  double* zg = I.y + (size_t)inst * nn;
  for (size_t e = threadIdx.x; e < nn; e += kNT) Z[e] = zg[e];
  int nodes = 0, status = 0;
  const bool bk = n > 30;
  for (int k = 0; k < nstep; ++k) {
    const double* tk = tmat + (size_t)k * nn;
    for (size_t e = threadIdx.x; e < nn; e += kNT) T[e] = tk[e];
    __syncthreads();
    // Z <- (X + X^T)/2, X = T^T (Z T)
    lds_gemm(Z, false, T, false, A, n);
    __syncthreads();
    lds_gemm(T, true, A, false, Z, n);
    __syncthreads();
    for (size_t e = threadIdx.x; e < nn; e += kNT) {
      const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
      if (i > j) {
        const double v = 0.5 * (Z[i + j * n] + Z[j + i * n]);
        Z[i + j * n] = v;
        Z[j + i * n] = v;
      }
    }
    const double w = drnow[k];
    for (int i = threadIdx.x; i < n; i += kNT) {
      double a, b, d;
      airy_sector(w, eig0[(size_t)k * n + i] + ered, hp[(size_t)k * n + i], a, b, d);
      y1[i] = a;
      y2[i] = b;
      y4[i] = d;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kNT) Z[i + i * n] += y1[i];
    __syncthreads();
    const int kc = sweep_inv<kNT>(Z, n, n, bk, s);
    if (kc > n || kc < 0) {
      status = k + 1;
      break;
    }
    nodes += outward ? kc : n - kc;
    for (size_t e = threadIdx.x; e < nn; e += kNT) {
      const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
      double v = -(y2[i] * Z[e] * y2[j]);
      if (i == j) v += y4[i];
      Z[e] = v;
    }
    __syncthreads();
  }
  if (status == 0) {
    // back to the primitive basis: Z <- V Z V^T
    for (size_t e = threadIdx.x; e < nn; e += kNT) T[e] = vlast[e];
    __syncthreads();
    lds_gemm(T, false, Z, false, A, n);
    __syncthreads();
    lds_gemm(A, false, T, true, Z, n);
    __syncthreads();
    for (size_t e = threadIdx.x; e < nn; e += kNT) {
      const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
      zg[e] = 0.5 * (Z[i + j * n] + Z[j + i * n]);
    }
  }
  if (threadIdx.x == 0) {
    I.nodes[inst] = nodes;
    I.status[inst] = status;
  }
}

}  // namespace mcp
