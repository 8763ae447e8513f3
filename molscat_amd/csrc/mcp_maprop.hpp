// MAPROP (IPROP = 7, the quasi-adiabatic modified log-derivative propagator
// of source_code/maprop.f:24-186) for N within LDS, on the LD grid (RSTART,
// then the midpoint and end of each sector).
//
// Everything but the reference potential's eigenvalues is independent of
// the energy: HAMMAT is called with ERED = 0 (maprop.f:48, 62, 171), so the
// adiabatic basis T_k of W at each sector midpoint, the basis change P_k =
// T_{k-1}^T T_k and the correction C_k = T_{k-1}^T A_k T_{k-1} with A_k =
// (W_start - W_mid) DR/6 + (W_start - W_prevmid) DR/6 (the second term 0 in
// the first sector) are computed once per grid and block, every sector in
// parallel: k_maprop_basis (Jacobi DIAGVC, ascending) and k_maprop_mats.
// k_maprop_prop then runs one workgroup per instance through the sectors:
// Y += C_k, Y <- sym(P_k^T Y P_k), the half-sector reference terms of each
// adiabat (tan / tanh of sqrt|EIVAL - ERED| DR/2), the inversion with its
// inertia (not counted when some adiabat lies above 24/DR^2, NCHECK), the
// Y update, and at the end Y <- T Y T^T + (W_end - W_mid) DR/6.
// Constant terms (NCONST > 0) enter W with the instance's constants
// (pconst), so with them the bases belong to one instance and the caller
// runs the instances one at a time.
#pragma once
#include <hip/hip_runtime.h>

#include "mcp_airyprop.hpp"
#include "mcp_device.hpp"
#include "mcp_eig.hpp"
#include "mcp_kernels.hpp"

namespace mcp {

// W at grid point k with ERED = 0 from P values pv; v = column-major index
__device__ __forceinline__ double ma_w(const BlockArgs& B, const double* pv, double r, double rsq, int i, int j) {
  double w = sumlam(B, i, j, pv);
  if (i == j) w = hammat_diag(B, i, w, r, rsq, 0.0);
  return w;
}

// one workgroup per sector: T_k and the eigenvalues of W at its midpoint
__global__ void __launch_bounds__(kNT) k_maprop_basis(BlockArgs B, GridArgs G, double* __restrict__ tout,
                                                      double* __restrict__ eout, const double* __restrict__ pconst) {
  extern __shared__ double smem[];
  const int n = B.n;
  const size_t nn = (size_t)n * n;
  const int k = blockIdx.x;
  LdsCarve c(smem);
  double* M = c.take<double>(nn);
  double* V = c.take<double>(nn);
  const int half = (n + 1) / 2 + 1;
  double* rc = c.take<double>(half);
  double* rs = c.take<double>(half);
  int* rp = c.take<int>(half);
  int* rq = c.take<int>(half);
  double* sv = c.take<double>(kNT / 64);
  double* ev = c.take<double>(n);
  int* order = c.take<int>(n);
  double* tmp = c.take<double>(n);
  double* pv = c.take<double>(B.nham > 0 ? B.nham : 1);
  const int km = 2 * k + 1;
  load_pvec(B, G, km, pconst, pv);
  __syncthreads();
  const double r = G.r[km], rsq = rsq_of(r);
  for (size_t e = threadIdx.x; e < nn; e += kNT) {
    const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
    M[e] = ma_w(B, pv, r, rsq, i, j);
  }
  __syncthreads();
  jacobi_eigh<kNT>(M, V, n, n, rc, rs, rp, rq, sv);
  __syncthreads();
  sort_eig<kNT>(M, V, n, n, ev, order, tmp);
  __syncthreads();
  double* to = tout + (size_t)k * nn;
  for (size_t e = threadIdx.x; e < nn; e += kNT) to[e] = V[e];
  for (int i = threadIdx.x; i < n; i += kNT) eout[(size_t)k * n + i] = ev[i];
}

// one workgroup per sector: C_k, P_k (and, in the last sector, the end
// correction qend); cr = DR/6
__global__ void __launch_bounds__(kNT) k_maprop_mats(BlockArgs B, GridArgs G, int nstep, double cr,
                                                     const double* __restrict__ tout, double* __restrict__ cout,
                                                     double* __restrict__ pout, double* __restrict__ qend,
                                                     const double* __restrict__ pconst) {
  extern __shared__ double smem[];
  const int n = B.n;
  const size_t nn = (size_t)n * n;
  const int k = blockIdx.x;
  LdsCarve c(smem);
  double* A = c.take<double>(nn);
  double* T = c.take<double>(nn);
  double* X = c.take<double>(nn);
  double* ps = c.take<double>(B.nham > 0 ? B.nham : 1);
  double* pm = c.take<double>(B.nham > 0 ? B.nham : 1);
  double* pp = c.take<double>(B.nham > 0 ? B.nham : 1);
  const int ks = 2 * k, km = 2 * k + 1;
  load_pvec(B, G, ks, pconst, ps);
  load_pvec(B, G, km, pconst, pm);
  if (k > 0) load_pvec(B, G, ks - 1, pconst, pp);
  __syncthreads();
  const double rs = G.r[ks], rm = G.r[km];
  const double rss = rsq_of(rs), rsm = rsq_of(rm);
  const double rp = k > 0 ? G.r[ks - 1] : 0.0, rsp = k > 0 ? rsq_of(rp) : 0.0;
  for (size_t e = threadIdx.x; e < nn; e += kNT) {
    const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
    const double ws = ma_w(B, ps, rs, rss, i, j);
    double v = cr * (ws - ma_w(B, pm, rm, rsm, i, j));
    if (k > 0) v = v + cr * (ws - ma_w(B, pp, rp, rsp, i, j));
    A[e] = v;
    T[e] = k > 0 ? tout[(size_t)(k - 1) * nn + e] : (i == j ? 1.0 : 0.0);
  }
  __syncthreads();
  // C_k = T^T A T, symmetrised (TRNSFM, maprop.f:72)
  lds_gemm(A, false, T, false, X, n);
  __syncthreads();
  lds_gemm(T, true, X, false, A, n);
  __syncthreads();
  double* co = cout + (size_t)k * nn;
  for (size_t e = threadIdx.x; e < nn; e += kNT) {
    const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
    co[e] = 0.5 * (A[e] + A[(size_t)j + (size_t)i * n]);
  }
  // P_k = T_{k-1}^T T_k (maprop.f:120)
  lds_gemm(T, true, tout + (size_t)k * nn, false, X, n);
  __syncthreads();
  double* po = pout + (size_t)k * nn;
  for (size_t e = threadIdx.x; e < nn; e += kNT) po[e] = X[e];
  if (k == nstep - 1) {
    // (W_end - W_mid) DR/6 of the last sector (maprop.f:171-177)
    __syncthreads();
    load_pvec(B, G, ks + 2, pconst, pp);
    __syncthreads();
    const double re = G.r[ks + 2], rse = rsq_of(re);
    for (size_t e = threadIdx.x; e < nn; e += kNT) {
      const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
      qend[e] = cr * (ma_w(B, pp, re, rse, i, j) - ma_w(B, pm, rm, rsm, i, j));
    }
  }
}

// one workgroup per instance
__global__ void __launch_bounds__(kNT) k_maprop_prop(int n, int nstep, double dr, const double* __restrict__ eout,
                                                     const double* __restrict__ cout,
                                                     const double* __restrict__ pout,
                                                     const double* __restrict__ tlast,
                                                     const double* __restrict__ qend, InstArgs I, int outward,
                                                     int bk) {
  extern __shared__ double smem[];
  const size_t nn = (size_t)n * n;
  const int inst = blockIdx.x;
  LdsCarve c(smem);
  double* Y = c.take<double>(nn);
  double* P = c.take<double>(nn);
  double* X = c.take<double>(nn);
  double* y1 = c.take<double>(n);
  double* y2 = c.take<double>(n);
  double* sv = c.take<double>(kNT / 64);
  SweepScratch s = carve_sweep(c, n);
  const double ered = I.ered[inst];
  double* yg = I.y + (size_t)inst * nn;
  for (size_t e = threadIdx.x; e < nn; e += kNT) Y[e] = yg[e];
  const double wmax = 24.0 / (dr * dr);
  int nodes = 0, status = 0;
  for (int k = 0; k < nstep; ++k) {
    const double* ck = cout + (size_t)k * nn;
    const double* pk = pout + (size_t)k * nn;
    for (size_t e = threadIdx.x; e < nn; e += kNT) {
      Y[e] = Y[e] + ck[e];
      P[e] = pk[e];
    }
    __syncthreads();
    lds_gemm(Y, false, P, false, X, n);
    __syncthreads();
    lds_gemm(P, true, X, false, Y, n);
    __syncthreads();
    double nchk = 0.0;
    for (size_t e = threadIdx.x; e < nn; e += kNT) {
      const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
      if (i < j) {
        const double v = 0.5 * (Y[e] + Y[(size_t)j + (size_t)i * n]);
        Y[e] = v;
        Y[(size_t)j + (size_t)i * n] = v;
      }
    }
    for (int i = threadIdx.x; i < n; i += kNT) {
      const double wref = eout[(size_t)k * n + i] - ered;
      const double flam = 0.5 * sqrt(fabs(wref));
      double a, b;
      if (wref < 0.0) {
        const double tn = tan(flam * dr);
        a = flam / tn - flam * tn;
        b = flam / tn + flam * tn;
      } else {
        if (wref > wmax) nchk = 1.0;
        const double th = tanh(flam * dr);
        a = flam / th + flam * th;
        b = flam / th - flam * th;
      }
      y1[i] = a;
      y2[i] = b;
    }
    const double ncheck = block_max<kNT>(nchk, sv);
    for (int i = threadIdx.x; i < n; i += kNT) Y[(size_t)i * n + i] += y1[i];
    __syncthreads();
    const int kc = sweep_inv<kNT>(Y, n, n, bk != 0, s);
    if (kc > n || kc < 0) {
      status = k + 1;
      break;
    }
    if (ncheck == 0.0) nodes += outward ? kc : n - kc;
    for (size_t e = threadIdx.x; e < nn; e += kNT) {
      const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
      double v = -y2[i] * Y[e] * y2[j];
      if (i == j) v += y1[i];
      Y[e] = v;
    }
    __syncthreads();
  }
  if (status == 0) {
    // back to the primitive basis and the end correction (maprop.f:180-184)
    for (size_t e = threadIdx.x; e < nn; e += kNT) P[e] = tlast[e];
    __syncthreads();
    lds_gemm(P, false, Y, false, X, n);
    __syncthreads();
    lds_gemm(X, false, P, true, Y, n);
    __syncthreads();
    for (size_t e = threadIdx.x; e < nn; e += kNT) {
      const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
      yg[e] = 0.5 * (Y[e] + Y[(size_t)j + (size_t)i * n]) + qend[e];
    }
  }
  if (threadIdx.x == 0) {
    I.nodes[inst] = nodes;
    I.status[inst] = status;
  }
}

}  // namespace mcp
