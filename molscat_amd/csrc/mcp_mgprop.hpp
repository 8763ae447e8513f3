// MGPROP (IPROP = 8, the symplectic log-derivative propagators of
// source_code/mgprop.f:20-78) for N within LDS: one workgroup per instance
// runs every sector.  Per substep k of a sector: Y += B_k H W(R), then, for
// A_k != 0, Y <- (I - (I + A_k H Y)^-1) / (A_k H) with the inertia of
// I + A_k H Y added to NODES (subtracted for A_k < 0), and R += A_k H.
// The grid holds the radius of every W evaluation in the reference's order
// (RSTART + (ISTEP-1) H, then R += A_k H), one P-table row per evaluation;
// with the scheme's last A zero the first substep of a sector reuses the
// last W (mgprop.f:26-36), which is the same row read again.
#pragma once
#include <hip/hip_runtime.h>

#include "mcp_device.hpp"
#include "mcp_kernels.hpp"

namespace mcp {

constexpr int kMgMaxSub = 6;

struct MgScheme {
  int m;                  // substeps per sector
  double a[kMgMaxSub];    // position advances (SYMSEL's A)
  double b[kMgMaxSub];    // potential weights (SYMSEL's B)
};

__global__ void __launch_bounds__(kNT) k_mgprop(BlockArgs B, GridArgs G, InstArgs I, MgScheme S, int nstep,
                                                double h, int bk) {
  extern __shared__ double smem[];
  const int n = B.n, lda = n;
  const size_t nn = (size_t)n * n;
  const int inst = blockIdx.x;
  LdsCarve c(smem);
  double* Y = c.take<double>(nn);
  double* pv = c.take<double>(B.nham > 0 ? B.nham : 1);
  SweepScratch s = carve_sweep(c, n);
  const double ered = I.ered[inst];
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  double* yg = I.y + (size_t)inst * nn;
  for (size_t e = threadIdx.x; e < nn; e += kNT) Y[e] = yg[e];
  int nodes = 0, status = 0;
  int kr = -1;  // P-table row of the last W evaluation
  const bool reuse = S.a[S.m - 1] == 0.0;
  for (int ks = 1; ks <= nstep && status == 0; ++ks) {
    for (int k = 0; k < S.m; ++k) {
      if (k != 0 || ks == 1 || !reuse) ++kr;
      __syncthreads();
      load_pvec(B, G, kr, pconst, pv);
      __syncthreads();
      const double r = G.r[kr], rsq = rsq_of(r);
      const double bh = S.b[k] * h;
      const double ak = S.a[k];
      const double ah = ak * h;
      for (size_t e = threadIdx.x; e < nn; e += kNT) {
        const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
        double w = sumlam(B, i, j, pv);
        if (i == j) w = hammat_diag(B, i, w, r, rsq, ered);
        double v = Y[e] + bh * w;
        if (ak != 0.0) {
          v = v * ah;
          if (i == j) v = v + 1.0;
        }
        Y[e] = v;
      }
      __syncthreads();
      if (ak != 0.0) {
        const int kc = sweep_inv<kNT>(Y, n, lda, bk != 0, s);
        if (kc > n || kc < 0) {
          status = ks;
          break;
        }
        nodes += ak >= 0.0 ? kc : -kc;
        const double d1 = 1.0 / ah;
        for (size_t e = threadIdx.x; e < nn; e += kNT) {
          const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
          double v = -Y[e];
          if (i == j) v = v + 1.0;
          Y[e] = v * d1;
        }
        __syncthreads();
      }
    }
  }
  if (status == 0) {
    for (size_t e = threadIdx.x; e < nn; e += kNT) {
      const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
      yg[e] = i >= j ? Y[e] : Y[(size_t)j + (size_t)i * n];
    }
  }
  if (threadIdx.x == 0) {
    I.nodes[inst] = nodes;
    I.status[inst] = status;
  }
}

}  // namespace mcp
