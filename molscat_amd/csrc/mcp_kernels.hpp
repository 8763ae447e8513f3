// mcprop kernels for gfx950: the W(R) build, batched symmetric inversion with
// inertia, the log-derivative propagators with the matrix resident in LDS
// (one workgroup per (block, energy/field) instance, the whole radial loop in
// one launch), the YINIT start, a batched eigen-solver and the Y -> K -> S
// matching.
#pragma once
#include <hip/hip_runtime.h>

#include "mcp_bessel.hpp"
#include "mcp_device.hpp"

namespace mcp {

constexpr int kNT = 256;  // threads per workgroup for the LDS kernels

// ---------------------------------------------------------------------------
// HAMMAT / SUMLAM (source_code/hammat.f:28-113, source_code/sumlam.f:29-36)
//
// VL is packed as VL(NHAM, N(N+1)/2): lambda fastest, then the upper
// triangle (row <= col) column by column.  The P vector of one radius is
// [ MXLAM radial terms | NCONST constants | NRSQ copies of min(1/R^2,1e16) ],
// already multiplied by EP2RU where HAMMAT does so.
// ---------------------------------------------------------------------------
struct BlockArgs {
  int n, nham, mxlam, nconst, nrsq;
  const double* vl;    // [ntri][nham]
  const double* eint;  // [n]
  const double* cent;  // [n]
  double conlen4inv;   // 1/CONLEN^4, 0 when CONLEN <= 0
};

struct GridArgs {
  int npts;            // number of tabulated radii
  const double* r;     // [npts]
  const double* ptab;  // [npts][mxlam]
};

__device__ __forceinline__ size_t tri_index(int i, int j) {
  const int a = i < j ? i : j;
  const int b = i < j ? j : i;
  return (size_t)b * (size_t)(b + 1) / 2 + (size_t)a;
}

__device__ __forceinline__ double sumlam(const BlockArgs& B, int i, int j, const double* pv) {
  const double* v = B.vl + tri_index(i, j) * (size_t)B.nham;
  double s = 0.0;
  for (int l = 0; l < B.nham; ++l) s += v[l] * pv[l];
  return s;
}

// Diagonal additions in HAMMAT's order (hammat.f:89-113).
__device__ __forceinline__ double hammat_diag(const BlockArgs& B, int i, double s, double r, double rsq,
                                              double ered) {
  if (B.nrsq == 0) s = s + rsq * B.cent[i];
  if (B.nconst == 0) s = s + B.eint[i];
  if (B.conlen4inv > 0.0) s = s + r * r * B.conlen4inv;
  s = s - ered;
  return s;
}

__device__ __forceinline__ double rsq_of(double r) {
  const double x = 1.0 / (r * r);
  return x < 1.0e16 ? x : 1.0e16;
}

// Fill pv[0..nham) for tabulated radius index k (all threads; caller syncs).
__device__ __forceinline__ void load_pvec(const BlockArgs& B, const GridArgs& G, int k, const double* pconst,
                                          double* pv) {
  const double rsq = rsq_of(G.r[k]);
  for (int l = threadIdx.x; l < B.nham; l += blockDim.x) {
    double v;
    if (l < B.mxlam)
      v = G.ptab[(size_t)k * B.mxlam + l];
    else if (l < B.mxlam + B.nconst)
      v = pconst[l - B.mxlam];
    else
      v = rsq;
    pv[l] = v;
  }
}

// Standalone W build: W[m](i,j) for nr radii (grid indices idx[m]) at one
// energy; one thread per packed element, both triangles written.
__global__ void k_hammat(BlockArgs B, GridArgs G, const int* __restrict__ idx, int nr, const double* __restrict__ pconst,
                         double ered, double* __restrict__ w) {
  const int n = B.n;
  const size_t ntri = (size_t)n * (n + 1) / 2;
  const int m = blockIdx.y;
  if (m >= nr) return;
  const int k = idx[m];
  const double r = G.r[k];
  const double rsq = rsq_of(r);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < ntri; e += (size_t)gridDim.x * blockDim.x) {
    // invert the packed index: column b, row a <= b
    int b = (int)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
    while ((size_t)(b + 1) * (b + 2) / 2 <= e) ++b;
    while ((size_t)b * (b + 1) / 2 > e) --b;
    const int a = (int)(e - (size_t)b * (b + 1) / 2);
    const double* v = B.vl + e * (size_t)B.nham;
    double s = 0.0;
    for (int l = 0; l < B.nham; ++l) {
      double pl;
      if (l < B.mxlam)
        pl = G.ptab[(size_t)k * B.mxlam + l];
      else if (l < B.mxlam + B.nconst)
        pl = pconst[l - B.mxlam];
      else
        pl = rsq;
      s += v[l] * pl;
    }
    if (a == b) s = hammat_diag(B, a, s, r, rsq, ered);
    double* W = w + (size_t)m * n * n;
    W[a + (size_t)b * n] = s;
    W[b + (size_t)a * n] = s;
  }
}

// ---------------------------------------------------------------------------
// LDS layout helper
// ---------------------------------------------------------------------------
struct LdsCarve {
  char* p;
  __device__ explicit LdsCarve(void* base) : p((char*)base) {}
  template <class T>
  __device__ T* take(size_t count) {
    T* r = (T*)p;
    size_t bytes = (count * sizeof(T) + 15) & ~(size_t)15;
    p += bytes;
    return r;
  }
};

__device__ __forceinline__ SweepScratch carve_sweep(LdsCarve& c, int n) {
  SweepScratch s;
  s.cu = c.take<double>(n);
  s.cv = c.take<double>(n);
  s.cx = c.take<double>(n);
  s.cy = c.take<double>(n);
  s.perm = c.take<int>(n);
  s.sv = c.take<double>(kNT / 64);
  s.si = c.take<int>(kNT / 64);
  return s;
}

// ---------------------------------------------------------------------------
// SYMINV batch: lower triangle in, full symmetric inverse out
// (source_code/syminv.f:1-97 / syminl-alloc.f:1-43).
// status: KOUNT, N+1 (zero pivot, natural order), or -INFO (BK).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kNT) k_syminv(int n, const double* __restrict__ a, double* __restrict__ out,
                                                int* __restrict__ kount, int bk) {
  extern __shared__ double smem[];
  LdsCarve c(smem);
  const int lda = n;
  double* A = c.take<double>((size_t)n * lda);
  SweepScratch s = carve_sweep(c, n);
  const double* src = a + (size_t)blockIdx.x * n * n;
  const Tile2D<kNT> t2;
  for (int j = t2.tj; j < n; j += Tile2D<kNT>::kWaves)
    for (int i = t2.ti; i < n; i += 64) {
      const int aa = i > j ? i : j, bb = i > j ? j : i;
      A[i + j * lda] = src[aa + (size_t)bb * n];
    }
  __syncthreads();
  const int k = sweep_inv<kNT>(A, n, lda, bk != 0, s);
  double* dst = out + (size_t)blockIdx.x * n * n;
  for (int j = t2.tj; j < n; j += Tile2D<kNT>::kWaves)
    for (int i = t2.ti; i < n; i += 64) dst[i + (size_t)j * n] = A[i + j * lda];
  if (threadIdx.x == 0) kount[blockIdx.x] = k;
}

// ---------------------------------------------------------------------------
// MDPROP energy-independent sector matrices (source_code/mdprop.f:114-135):
// for step ks the midpoint C = grid index 2ks-1,
//   UC = [ I/8 + D4 * W_off(C) ]^-1 - 8 I   (diagonal overwritten by 1/8)
// with the inversion's KOUNT (summed into NSAVE).  One workgroup per step.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kNT) k_mdprop_uc(BlockArgs B, GridArgs G, const double* __restrict__ pconst,
                                                   int step0, int nsteps_here, double h, int bk,
                                                   double* __restrict__ uc, int* __restrict__ kount) {
  extern __shared__ double smem[];
  const int ks = step0 + blockIdx.x;  // 1-based step
  if ((int)blockIdx.x >= nsteps_here) return;
  const int n = B.n, lda = n;
  LdsCarve c(smem);
  double* A = c.take<double>((size_t)n * lda);
  double* pv = c.take<double>(B.nham);
  SweepScratch s = carve_sweep(c, n);
  const double d1 = h * h / 3.0;
  const double d4 = -d1 / 16.0;
  load_pvec(B, G, 2 * ks - 1, pconst, pv);
  __syncthreads();
  const Tile2D<kNT> t2;
  for (int j = t2.tj; j < n; j += Tile2D<kNT>::kWaves)
    for (int i = t2.ti; i < n; i += 64) A[i + j * lda] = (i == j) ? 0.125 : d4 * sumlam(B, i, j, pv);
  __syncthreads();
  const int k = sweep_inv<kNT>(A, n, lda, bk != 0, s);
  double* dst = uc + (size_t)blockIdx.x * n * n;
  for (int j = t2.tj; j < n; j += Tile2D<kNT>::kWaves)
    for (int i = t2.ti; i < n; i += 64) dst[i + (size_t)j * n] = A[i + j * lda] - ((i == j) ? 8.0 : 0.0);
  if (threadIdx.x == 0) kount[blockIdx.x] = k;
}

// ---------------------------------------------------------------------------
// MDPROP main loop with Y resident in LDS (source_code/mdprop.f:64-279).
// One workgroup per instance (energy or field).  Grid index 0 is RSTART,
// 2ks-1 the midpoint C and 2ks the end B of step ks.  UC comes from
// k_mdprop_uc (shared by all instances).  On exit Y holds Y(RSTOP) (both
// triangles), nodes = sum KOUNT(Y) - NSAVE, status = 0 or the failing step.
// ---------------------------------------------------------------------------
struct InstArgs {
  const double* ered;    // [ninst]
  const double* pconst;  // [ninst][nconst] (NCONST values x EP2RU) or null
  double* y;             // [ninst][n*n] in/out
  int* nodes;            // [ninst]
  int* status;           // [ninst]
};

// NT threads per instance: 256 (four waves) or 64 (one wave, no cross-wave
// barriers, more instances resident per CU); the arithmetic per element and
// the pivot decisions do not depend on NT.  Four waves per SIMD for NT =
// 256 (at most 128 VGPRs, no spills): the field scans run many instances
// per CU.
// One instance's MDPROP with Y in LDS (the body of k_mdprop_lds and of the
// persistent multi-block launch k_mdprop_lds_queue): yg is Y(RSTART) on
// entry and Y(RSTOP) on exit, *nodes_out and *status_out the results.
// With wrec (BOUND's WAVE option, source_code/mdprop.f:170-176,212-245) step
// ks also writes W3 = W W2 to wrec + (ks-1) N^2 (column-major): W the first
// inverse [H(Y(A) + Y1(A,C))]^-1 (kept in wtmp, N^2), W2 = Y23 [H(Y(C) +
// Y1(C,B))]^-1 Y23 the second; W3 carries the wavefunction from the step's
// end back to its start.
template <int NT>
__device__ __forceinline__ void mdprop_lds_run(const BlockArgs& B, const GridArgs& G, double ered,
                                               const double* pconst, double* yg, int* nodes_out, int* status_out,
                                               int nstep, double h, const double* __restrict__ uc,
                                               const int* __restrict__ uc_kount, int bk, double* smem,
                                               double* wrec = nullptr, double* wtmp = nullptr) {
  const int n = B.n, lda = n;
  LdsCarve c(smem);
  double* Y = c.take<double>((size_t)n * lda);
  double* pv = c.take<double>(B.nham);
  double* y14 = c.take<double>(n);
  double* y23 = c.take<double>(n);
  double* u2 = c.take<double>(n);
  SweepScratch s = carve_sweep(c, n);
  const Tile2D<NT> t2;
  const double d1 = h * h / 3.0;
  const double d2 = 2.0 * d1;
  const double half = 0.5 * h;

  // NSAVE = sum of the sector-matrix inertias (mdprop.f:133,277-279)
  int nsave = 0;
  {
    double ps = 0.0, pbad = 0.0;
    for (int k = threadIdx.x; k < nstep; k += NT) {
      const int kc = uc_kount[k];
      ps += (double)kc;
      if (kc > n || kc < 0) pbad += 1.0;
    }
    ps = block_sum<NT>(ps, s.sv);
    pbad = block_sum<NT>(pbad, s.sv);
    if (pbad > 0.0) {
      if (threadIdx.x == 0) {
        *nodes_out = 0;
        *status_out = -1;
      }
      return;
    }
    nsave = (int)ps;
  }

  // Y <- H*Y + D1*W(RSTART)   (mdprop.f:89-93)
  load_pvec(B, G, 0, pconst, pv);
  __syncthreads();
  {
    const double r0 = G.r[0], rsq0 = rsq_of(r0);
    for (int j = t2.tj; j < n; j += Tile2D<NT>::kWaves)
      for (int i = t2.ti; i < n; i += 64) {
        if (i < j) continue;
        double w = sumlam(B, i, j, pv);
        if (i == j) w = hammat_diag(B, i, w, r0, rsq0, ered);
        const double v = h * yg[i + (size_t)j * n] + d1 * w;
        Y[i + j * lda] = v;
        Y[j + i * lda] = v;
      }
  }
  __syncthreads();
  int nodes = 0, status = 0;
  for (int ks = 1; ks <= nstep; ++ks) {
    // midpoint C: reference-potential terms from DIAG(C) (mdprop.f:143-164)
    load_pvec(B, G, 2 * ks - 1, pconst, pv);
    __syncthreads();
    {
      const double rc = G.r[2 * ks - 1], rsqc = rsq_of(rc);
      for (int i = threadIdx.x; i < n; i += NT) {
        const double wref = hammat_diag(B, i, sumlam(B, i, i, pv), rc, rsqc, ered);
        const double arg = half * sqrt(fabs(wref));
        double a14, a23;
        if (wref < 0.0) {
          const double tn = tan(arg);
          a14 = arg / tn - arg * tn;
          a23 = arg / tn + arg * tn;
        } else {
          const double th = tanh(arg);
          a14 = arg / th + arg * th;
          a23 = arg / th - arg * th;
        }
        u2[i] = 2.0 * a14;
        const double y14c = a14 - d1 * wref;
        y14[i] = y14c;
        y23[i] = a23;
        Y[i + i * lda] = Y[i + i * lda] + y14c;
      }
    }
    __syncthreads();
    int kc = sweep_inv<NT>(Y, n, lda, bk != 0, s);
    if (kc > n || kc < 0) {
      status = ks;
      break;
    }
    nodes += kc;
    if (wrec) {
      for (int j = t2.tj; j < n; j += Tile2D<NT>::kWaves)
        for (int i = t2.ti; i < n; i += 64) {
          if (i < j) continue;
          const double v = Y[i + j * lda];
          wtmp[i + (size_t)j * n] = v;
          wtmp[j + (size_t)i * n] = v;
        }
      __syncthreads();
    }
    // Y <- U(C) + 2 Y14 - Y23 Y Y23   (mdprop.f:181-185)
    {
      const double* U = uc + (size_t)(ks - 1) * n * n;
      for (int j = t2.tj; j < n; j += Tile2D<NT>::kWaves)
        for (int i = t2.ti; i < n; i += 64) {
          if (i < j) continue;
          double uij = U[i + (size_t)j * n];
          if (i == j) uij = uij + u2[i];
          const double v = uij - y23[i] * Y[i + j * lda] * y23[j];
          Y[i + j * lda] = v;
          Y[j + i * lda] = v;
        }
    }
    __syncthreads();
    kc = sweep_inv<NT>(Y, n, lda, bk != 0, s);
    if (kc > n || kc < 0) {
      status = ks;
      break;
    }
    nodes += kc;
    if (wrec) {
      // W3(i,j) = sum_m W(i,m) y23[m] Y(m,j) y23[j], Y's lower triangle
      double* dst = wrec + (size_t)(ks - 1) * n * n;
      for (int e = threadIdx.x; e < n * n; e += NT) {
        const int j = e / n, i = e - j * n;
        double acc = 0.0;
        for (int m = 0; m < n; ++m) {
          const double ymj = m >= j ? Y[m + j * lda] : Y[j + m * lda];
          acc += wtmp[i + (size_t)m * n] * (y23[m] * ymj);
        }
        dst[e] = acc * y23[j];
      }
      __syncthreads();
    }
    // Y <- D2*W(B) - Y23 Y Y23 + Y14   (mdprop.f:197-230)
    const double dd2 = (ks == nstep) ? d1 : d2;
    load_pvec(B, G, 2 * ks, pconst, pv);
    __syncthreads();
    {
      const double rb = G.r[2 * ks], rsqb = rsq_of(rb);
      for (int j = t2.tj; j < n; j += Tile2D<NT>::kWaves)
        for (int i = t2.ti; i < n; i += 64) {
          if (i < j) continue;
          double w = sumlam(B, i, j, pv);
          if (i == j) w = hammat_diag(B, i, w, rb, rsqb, ered);
          double v = dd2 * w - y23[i] * Y[i + j * lda] * y23[j];
          if (i == j) v = v + y14[i];
          Y[i + j * lda] = v;
          Y[j + i * lda] = v;
        }
    }
    __syncthreads();
  }
  const double hi = 1.0 / h;
  for (int j = t2.tj; j < n; j += Tile2D<NT>::kWaves)
    for (int i = t2.ti; i < n; i += 64) yg[i + (size_t)j * n] = hi * Y[i + j * lda];
  if (threadIdx.x == 0) {
    *nodes_out = nodes - nsave;
    *status_out = status;
  }
}

template <int NT>
__global__ void __launch_bounds__(NT, NT == 256 ? 4 : 1) k_mdprop_lds(BlockArgs B, GridArgs G, InstArgs I, int nstep, double h,
                                                    const double* __restrict__ uc,
                                                    const int* __restrict__ uc_kount, int bk) {
  extern __shared__ double smem[];
  const int inst = blockIdx.x;
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  mdprop_lds_run<NT>(B, G, I.ered[inst], pconst, I.y + (size_t)inst * B.n * B.n, I.nodes + inst, I.status + inst,
                     nstep, h, uc, uc_kount, bk, smem);
}

// MDPROP of one instance writing BOUND's WAVE records (wrec: nstep N x N
// matrices; wtmp: N^2 scratch).
__global__ void __launch_bounds__(kNT) k_mdprop_lds_wave(BlockArgs B, GridArgs G, InstArgs I, int nstep, double h,
                                                         const double* __restrict__ uc,
                                                         const int* __restrict__ uc_kount, int bk, double* wrec,
                                                         double* wtmp) {
  extern __shared__ double smem[];
  mdprop_lds_run<kNT>(B, G, I.ered[0], I.pconst, I.y, I.nodes, I.status, nstep, h, uc, uc_kount, bk, smem, wrec,
                      wtmp);
}

// WVBKSB's back-substitution (source_code/wvbksb.f:64-114) of one segment
// per workgroup: from psi0 at the matching point, psi(A) = W3 psi(B) with
// the segment's records taken last to first; out[k] (N values) is psi after
// k steps back from the matching point, k = 0 .. nstep.  Segment s has
// nstep[s] records at recs[s] and writes out[s].
struct WaveSeg {
  const double* recs;
  double* out;
  int nstep;
};

__global__ void __launch_bounds__(kNT) k_wvbksb(int n, const double* __restrict__ psi0, WaveSeg s0, WaveSeg s1) {
  extern __shared__ double smem[];
  const WaveSeg S = blockIdx.x == 0 ? s0 : s1;
  double* pa = smem;
  double* pb = smem + n;
  for (int i = threadIdx.x; i < n; i += kNT) {
    pa[i] = psi0[i];
    S.out[i] = psi0[i];
  }
  __syncthreads();
  for (int k = 0; k < S.nstep; ++k) {
    const double* w3 = S.recs + (size_t)(S.nstep - 1 - k) * n * n;
    for (int i = threadIdx.x; i < n; i += kNT) {
      double acc = 0.0;
      for (int m = 0; m < n; ++m) acc += w3[i + (size_t)m * n] * pa[m];
      pb[i] = acc;
      S.out[(size_t)(k + 1) * n + i] = acc;
    }
    __syncthreads();
    double* t = pa;
    pa = pb;
    pb = t;
  }
}

// One block of a multi-block MDPROP launch: its arguments, its
// energy-independent sector matrices and its instances' arrays.
struct QueueBlock {
  BlockArgs B;
  const double* uc;    // [nstep][n*n] U(C)
  const int* uc_kount; // [nstep]
  const double* ered;  // [ninst]
  double* y;           // [ninst][n*n]
  int* nodes;          // [ninst]
  int* status;         // [ninst]
  int bk;
};

// MDPROP for the instances of several blocks of uneven N on one radial grid
// in one persistent launch: every workgroup takes the next (block,
// instance) item from the device counter *head (zero at launch) until the
// list is exhausted.  Items are listed longest first (by N^3), so the long
// ones start first and the short ones fill in behind them.
template <int NT>
__global__ void __launch_bounds__(NT, NT == 256 ? 4 : 1) k_mdprop_lds_queue(const QueueBlock* __restrict__ blocks,
                                                                          const int2* __restrict__ items, int nitems,
                                                                          unsigned* __restrict__ head, GridArgs G,
                                                                          int nstep, double h) {
  extern __shared__ double smem[];
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(head, 1u);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= nitems) break;
    const int2 w = items[it];
    const QueueBlock qb = blocks[w.x];
    const int nn = qb.B.n * qb.B.n;
    mdprop_lds_run<NT>(qb.B, G, qb.ered[w.y], nullptr, qb.y + (size_t)w.y * nn, qb.nodes + w.y, qb.status + w.y,
                       nstep, h, qb.uc, qb.uc_kount, qb.bk, smem);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// LDPROP (Johnson, source_code/ldprop.f:24-147) with Z and U in LDS.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kNT) k_ldprop_lds(BlockArgs B, GridArgs G, InstArgs I, int nstep, double h,
                                                    int bk) {
  extern __shared__ double smem[];
  const int inst = blockIdx.x;
  const int n = B.n, lda = n;
  LdsCarve c(smem);
  double* Z = c.take<double>((size_t)n * lda);
  double* U = c.take<double>((size_t)n * lda);
  double* pv = c.take<double>(B.nham);
  SweepScratch s = carve_sweep(c, n);
  const Tile2D<kNT> t2;
  const double ered = I.ered[inst];
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  const double d1 = h * h / 3.0;
  const double d2 = 2.0 * d1;
  const double d4 = -d1 / 16.0;
  double* yg = I.y + (size_t)inst * n * n;

  load_pvec(B, G, 0, pconst, pv);
  __syncthreads();
  {
    const double r0 = G.r[0], rsq0 = rsq_of(r0);
    for (int j = t2.tj; j < n; j += Tile2D<kNT>::kWaves)
      for (int i = t2.ti; i < n; i += 64) {
        if (i < j) continue;
        double w = sumlam(B, i, j, pv);
        if (i == j) w = hammat_diag(B, i, w, r0, rsq0, ered);
        double v = h * yg[i + (size_t)j * n] + d1 * w;
        if (i == j) v = 1.0 + v;
        Z[i + j * lda] = v;
        Z[j + i * lda] = v;
      }
  }
  __syncthreads();
  int nodes = 0, status = 0;
  for (int ks = 1; ks <= nstep; ++ks) {
    load_pvec(B, G, 2 * ks - 1, pconst, pv);
    __syncthreads();
    {
      const double rc = G.r[2 * ks - 1], rsqc = rsq_of(rc);
      for (int j = t2.tj; j < n; j += Tile2D<kNT>::kWaves)
        for (int i = t2.ti; i < n; i += 64) {
          double w = sumlam(B, i, j, pv);
          if (i == j) w = hammat_diag(B, i, w, rc, rsqc, ered);
          double v = d4 * w;
          if (i == j) v = 0.125 + v;
          U[i + j * lda] = v;
        }
    }
    __syncthreads();
    const int ncu = sweep_inv<kNT>(U, n, lda, bk != 0, s);
    if (ncu > n || ncu < 0) {
      status = ks;
      break;
    }
    int ncz = sweep_inv<kNT>(Z, n, lda, bk != 0, s);
    if (ncz > n || ncz < 0) {
      status = ks;
      break;
    }
    nodes += ncz;
    for (int j = t2.tj; j < n; j += Tile2D<kNT>::kWaves)
      for (int i = t2.ti; i < n; i += 64) {
        double v = U[i + j * lda] - Z[i + j * lda];
        if (i == j) v = v - 6.0;
        Z[i + j * lda] = v;
      }
    __syncthreads();
    ncz = sweep_inv<kNT>(Z, n, lda, bk != 0, s);
    if (ncz > n || ncz < 0) {
      status = ks;
      break;
    }
    nodes += ncz - ncu;
    const double dd2 = (ks == nstep) ? d1 : d2;
    load_pvec(B, G, 2 * ks, pconst, pv);
    __syncthreads();
    {
      const double rb = G.r[2 * ks], rsqb = rsq_of(rb);
      for (int j = t2.tj; j < n; j += Tile2D<kNT>::kWaves)
        for (int i = t2.ti; i < n; i += 64) {
          double w = sumlam(B, i, j, pv);
          if (i == j) w = hammat_diag(B, i, w, rb, rsqb, ered);
          double u = dd2 * w;
          if (i == j) u = u + 2.0;
          Z[i + j * lda] = u - Z[i + j * lda];
        }
    }
    __syncthreads();
  }
  const double hi = 1.0 / h;
  for (int j = t2.tj; j < n; j += Tile2D<kNT>::kWaves)
    for (int i = t2.ti; i < n; i += 64) {
      double v = hi * Z[i + j * lda];
      if (i == j) v = v - hi;
      yg[i + (size_t)j * n] = v;
    }
  if (threadIdx.x == 0) {
    I.nodes[inst] = nodes;
    I.status[inst] = status;
  }
}

// ---------------------------------------------------------------------------
// Batched symmetric eigen-decomposition (DIAGVC/DIAGVL, source_code/diagvc.f):
// eigenvalues ascending, eigenvectors as columns (sign fixed so that the
// largest component of each vector is positive).
// ---------------------------------------------------------------------------
template <int NT>
__device__ void sort_eig(double* A, double* V, int n, int lda, double* ev, int* order, double* tmp) {
  // A diagonal holds eigenvalues; rank-sort (stable by index) into ev/order
  for (int i = threadIdx.x; i < n; i += NT) {
    const double di = A[i + i * lda];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = A[j + j * lda];
      if (dj < di || (dj == di && j < i)) ++rank;
    }
    order[rank] = i;
    ev[rank] = di;
  }
  __syncthreads();
  // permute columns of V into A (A no longer needed), fix signs
  for (int e = threadIdx.x; e < n * n; e += NT) {
    const int j = e / n, i = e - j * n;
    A[i + j * lda] = V[i + order[j] * lda];
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += NT) {
    double big = 0.0;
    for (int i = 0; i < n; ++i)
      if (fabs(A[i + j * lda]) > fabs(big)) big = A[i + j * lda];
    tmp[j] = big < 0.0 ? -1.0 : 1.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += NT) {
    const int j = e / n, i = e - j * n;
    V[i + j * lda] = A[i + j * lda] * tmp[j];
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kNT) k_eigh(int n, const double* __restrict__ a, double* __restrict__ evals,
                                              double* __restrict__ evecs, int* __restrict__ sweeps) {
  extern __shared__ double smem[];
  const int lda = n;
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
  const double* src = a + (size_t)blockIdx.x * n * n;
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    const int aa = i > j ? i : j, bb = i > j ? j : i;
    A[i + j * lda] = src[aa + (size_t)bb * n];
  }
  __syncthreads();
  const int ns = jacobi_eigh<kNT>(A, V, n, lda, rc, rs, rp, rq, sv);
  sort_eig<kNT>(A, V, n, lda, ev, order, tmp);
  for (int i = threadIdx.x; i < n; i += kNT) evals[(size_t)blockIdx.x * n + i] = ev[i];
  if (evecs) {
    double* dst = evecs + (size_t)blockIdx.x * n * n;
    for (int e = threadIdx.x; e < n * n; e += kNT) dst[e] = V[e % n + (e / n) * lda];
  }
  if (threadIdx.x == 0 && sweeps) sweeps[blockIdx.x] = ns;
}

// ---------------------------------------------------------------------------
// ODPROP (source_code/odprop.f): the single-channel LDMD propagator SCCTRL
// runs for N = 1 (source_code/scctrl.f:183-196).  One thread per instance.
// U(R) = W(R) + ERED at the sector ends r[2i], W at the midpoints r[2i+1] of
// the LD grid; Q_i = U_i - W_{i-1}; per sector the reference-potential
// terms of the midpoint and Y <- Y1 - Y2^2 / (Y + Y1 + (U_i - W_i + Q_i) DR/6);
// Y(RSTOP) = Y + Q_NPT DR/6.  NODES = 0, as in the reference.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double od_u(const BlockArgs& B, const GridArgs& G, int k, const double* pconst) {
  const double r = G.r[k];
  const double rsq = rsq_of(r);
  const double* v = B.vl;  // N = 1: one packed element
  double s = 0.0;
  for (int l = 0; l < B.nham; ++l) {
    const double pl = l < B.mxlam ? G.ptab[(size_t)k * B.mxlam + l] : (l < B.mxlam + B.nconst ? pconst[l - B.mxlam] : rsq);
    s += v[l] * pl;
  }
  return hammat_diag(B, 0, s, r, rsq, 0.0);
}

__global__ void k_odprop(BlockArgs B, GridArgs G, InstArgs I, int ninst, int nstep, double dr) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= ninst) return;
  const double* pconst = I.pconst ? I.pconst + (size_t)m * B.nconst : nullptr;
  const double ered = I.ered[m];
  const double dr6 = dr / 6.0;
  double y = I.y[m];
  double wprev = 0.0;
  for (int i = 0; i < nstep; ++i) {
    const double ui = od_u(B, G, 2 * i, pconst);
    const double wi = od_u(B, G, 2 * i + 1, pconst);
    const double qi = i == 0 ? 0.0 : ui - wprev;
    const double uu = (ui - wi + qi) * dr6;
    const double wref = wi - ered;
    const double flam = 0.5 * sqrt(fabs(wref));
    double y1, y2;
    if (wref < 0.0) {
      const double tn = tan(flam * dr);
      y1 = flam / tn - flam * tn;
      y2 = flam / tn + flam * tn;
    } else {
      const double tn = tanh(flam * dr);
      y1 = flam / tn + flam * tn;
      y2 = flam / tn - flam * tn;
    }
    y2 = y2 * y2;
    y = y1 - y2 / (y + (y1 + uu));
    wprev = wi;
  }
  const double qlast = (od_u(B, G, 2 * nstep, pconst) - wprev) * dr6;
  I.y[m] = y + qlast;
  I.nodes[m] = 0;
  I.status[m] = 0;
}

// ---------------------------------------------------------------------------
// BDCTRL matching (source_code/bdctrl.f:221-257): M = Y_out - Y_in (the
// outward and inward log-derivative matrices at RMATCH), DIAGVC of M
// (ascending), NEGCNT = #{eig <= 0}, EIGMIN = the eigenvalue of smallest
// modulus (the last such in ascending order, as the reference's scan keeps
// ties), PSIMCH = its eigenvector with the largest component (the first such)
// made positive, NODEC = NODES_out + NODES_in + NEGCNT.  One workgroup per
// instance; nodes_out/nodes_in may be the same array offset by ninst.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kNT) k_bound_match(int n, const double* __restrict__ yout,
                                                     const double* __restrict__ yin, const int* __restrict__ nodes_out,
                                                     const int* __restrict__ nodes_in, int* __restrict__ nodec,
                                                     double* __restrict__ eigmin, int* __restrict__ imin,
                                                     double* __restrict__ psimch, double* __restrict__ evals) {
  extern __shared__ double smem[];
  const int lda = n;
  const int m = blockIdx.x;
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
  __shared__ int s_imin, s_neg, s_big;
  const double* yo = yout + (size_t)m * n * n;
  const double* yi = yin + (size_t)m * n * n;
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    const int aa = i > j ? i : j, bb = i > j ? j : i;
    const size_t src = aa + (size_t)bb * n;
    A[i + j * lda] = yo[src] - yi[src];
  }
  __syncthreads();
  jacobi_eigh<kNT>(A, V, n, lda, rc, rs, rp, rq, sv);
  sort_eig<kNT>(A, V, n, lda, ev, order, tmp);
  if (threadIdx.x == 0) {
    int neg = 0, im = 0;
    double em = 1.0e30;
    for (int i = 0; i < n; ++i) {
      if (ev[i] <= 0.0) ++neg;
      if (fabs(ev[i]) > fabs(em)) continue;
      em = ev[i];
      im = i;
    }
    int big = 0;
    double bv = 0.0;
    for (int i = 0; i < n; ++i) {
      const double v = V[i + im * lda];
      if (fabs(v) > fabs(bv)) {
        bv = v;
        big = i;
      }
    }
    s_imin = im;
    s_neg = neg;
    s_big = big;
    eigmin[m] = em;
    imin[m] = im;
    nodec[m] = nodes_out[m] + nodes_in[m] + neg;
  }
  __syncthreads();
  const double sgn = V[s_big + s_imin * lda] < 0.0 ? -1.0 : 1.0;
  for (int i = threadIdx.x; i < n; i += kNT) {
    if (psimch) psimch[(size_t)m * n + i] = sgn * V[i + s_imin * lda];
    if (evals) evals[(size_t)m * n + i] = ev[i];
  }
}

// ---------------------------------------------------------------------------
// YINIT (source_code/yinit.f:70-150) with ADIAMN/WKBMN: W(R) diagonalised,
// Y = V diag(y) V^T with y = +sqrt(w) (closed, outward) / -sqrt(w) (closed,
// inward) / BCYO (open).  One workgroup per instance; grid index k0.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kNT) k_yinit(BlockArgs B, GridArgs G, InstArgs I, int k0, int outward,
                                               double bcyo, int adiab) {
  extern __shared__ double smem[];
  const int inst = blockIdx.x;
  const int n = B.n, lda = n;
  LdsCarve c(smem);
  double* A = c.take<double>((size_t)n * lda);
  double* V = c.take<double>((size_t)n * lda);
  const int half = (n + 1) / 2 + 1;
  double* rc = c.take<double>(half);
  double* rs = c.take<double>(half);
  int* rp = c.take<int>(half);
  int* rq = c.take<int>(half);
  double* sv = c.take<double>(kNT / 64);
  double* pv = c.take<double>(B.nham);
  double* yd = c.take<double>(n);
  const double ered = I.ered[inst];
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  load_pvec(B, G, k0, pconst, pv);
  __syncthreads();
  const double r0 = G.r[k0], rsq0 = rsq_of(r0);
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    double w = sumlam(B, i, j, pv);
    if (i == j) w = hammat_diag(B, i, w, r0, rsq0, ered);
    A[i + j * lda] = w;
  }
  __syncthreads();
  if (adiab) {
    jacobi_eigh<kNT>(A, V, n, lda, rc, rs, rp, rq, sv);
  } else {
    for (int e = threadIdx.x; e < n * n; e += kNT) {
      const int j = e / n, i = e - j * n;
      V[i + j * lda] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += kNT) {
    const double w = A[i + i * lda];
    double y;
    if (w > 0.0)
      y = outward ? sqrt(fabs(w)) : -sqrt(fabs(w));
    else
      y = bcyo;
    yd[i] = y;
  }
  __syncthreads();
  double* yg = I.y + (size_t)inst * n * n;
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += V[i + k * lda] * yd[k] * V[j + k * lda];
    yg[i + (size_t)j * n] = s;
  }
}

// ---------------------------------------------------------------------------
// Y -> K (YTOK/YTOKG/KSYM, source_code/ytok.f, ytokg.f, ksym.f) and K -> S
// (KTOS, source_code/ktos.f).  One workgroup per instance; matrices in global
// scratch (the solve runs once per S-matrix).  nb is the CHNSRT order; the
// first nopen entries are the open channels.
// Output SR/SI are nopen x nopen (ld = nopen) in NB order.
// ---------------------------------------------------------------------------
struct MatchArgs {
  int n;
  const double* y;     // [ninst][n*n]
  const double* wvec;  // [ninst][n]  sign(E-EINT) sqrt|E-EINT|
  const int* nopen;    // [ninst]
  const int* nb;       // [ninst][n]
  const double* cent;  // [n] shared, or [ninst][n] with cent_stride = n
  int cent_stride;
  double rval;
  double* work;        // [ninst][n*n + n*n + 4n]
  double* kmat;        // [ninst][n*n] (nopen x nopen used)
  double* sr;          // [ninst][n*n]
  double* si;          // [ninst][n*n]
  int* status;         // [ninst]
};

// k_match runs in two phases around k_match_lu (the forward elimination of
// YTOKG's DGESV): kMatchSetup builds the Riccati-Bessel functions and the two
// matching matrices, kMatchFinish does the back substitution, KSYM and KTOS.
// The diagonal of U is kept in the (no longer needed) SJ slot of the work area.
constexpr int kMatchSetup = 0;
constexpr int kMatchFinish = 1;

constexpr int kMatchBadArg = 99;  // NOPEN or an NB entry out of range

__global__ void __launch_bounds__(kNT) k_match(MatchArgs M, int phase) {
  __shared__ int bad;
  const int inst = blockIdx.x;
  const int n = M.n;
  const int nop = M.nopen[inst];
  const int* nb = M.nb + (size_t)inst * n;
  if (phase != kMatchFinish) {
    // the arguments of callers that pass device arrays are checked here:
    // an instance with NOPEN outside [0, N] or an NB entry outside [0, N)
    // gets status kMatchBadArg and is skipped by every later pass
    if (threadIdx.x == 0) bad = (nop < 0 || nop > n) ? 1 : 0;
    __syncthreads();
    for (int ii = threadIdx.x; ii < n && !bad; ii += kNT)
      if (nb[ii] < 0 || nb[ii] >= n) atomicMax(&bad, 1);
    __syncthreads();
    if (bad) {
      if (threadIdx.x == 0) M.status[inst] = kMatchBadArg;
      return;
    }
    __syncthreads();
  }
  const double* y = M.y + (size_t)inst * n * n;
  const double* wv = M.wvec + (size_t)inst * n;
  double* km = M.work + (size_t)inst * (2 * (size_t)n * n + 4 * (size_t)n);
  double* t = km + (size_t)n * n;
  double* sj = t + (size_t)n * n;
  double* sjp = sj + n;
  double* sn = sjp + n;
  double* snp = sn + n;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  if (nop <= 0) {
    if (threadIdx.x == 0) M.status[inst] = 0;
    return;
  }
  if (phase == kMatchFinish) {
    if (M.status[inst] != 0) return;
  } else {
  // Riccati-Bessel functions per channel (ytok.f:24-45)
  for (int ii = threadIdx.x; ii < n; ii += kNT) {
    const int nx = nb[ii];
    const double ell = sqrt(M.cent[(size_t)inst * M.cent_stride + nx] + 0.25) - 0.5;
    if (ii < nop) {
      const double dw = wv[nx];
      double uj, ujp, un, unp;
      const int rc = riccati_jy(ell, dw * M.rval, uj, ujp, un, unp);
      if (rc) atomicMax(&bad, rc);
      const double rd = sqrt(dw);
      sj[nx] = uj / rd;
      sjp[nx] = ujp * rd;
      sn[nx] = un / rd;
      snp[nx] = unp * rd;
    } else {
      const double dw = fabs(wv[nx]);
      sj[nx] = 0.0;
      sjp[nx] = 0.0;
      sn[nx] = 1.0;
      if (dw == 0.0) {
        // a channel exactly at its threshold: the k -> 0 limit of
        // k K'/K, the log-derivative of the decaying zero-energy solution
        // R^-ell (RBESSK itself has no value at zero argument)
        snp[nx] = -ell / M.rval;
      } else {
        double ratio;
        const int rc = riccati_k_ratio(ell, dw * M.rval, ratio);
        if (rc) atomicMax(&bad, rc);
        snp[nx] = ratio * dw;
      }
    }
  }
  __syncthreads();
  // T(i,j) = Y(nb_i,nb_j) SJ(nb_j) - delta_ij SJP;  KM(i,j) = Y(nb_i,nb_j) SN(nb_j) - delta_ij SNP
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    const int a = nb[i], b = nb[j];
    const int aa = a > b ? a : b, bb = a > b ? b : a;
    const double yv = y[aa + (size_t)bb * n];  // lower triangle (DSYFIL('U') mirrors it)
    double kv = yv * sn[b];
    if (i == j) kv -= snp[b];
    km[i + (size_t)j * n] = kv;
    if (j < nop) {
      double tv = yv * sj[b];
      if (i == j) tv -= sjp[b];
      t[i + (size_t)j * n] = tv;
    }
  }
  __syncthreads();
  __syncthreads();
  if (threadIdx.x == 0) M.status[inst] = bad;
  return;
  }
  const double* ud = sj;  // diagonal of U (k_match_lu)
  // back substitution U x = t by columns of U: x_i = t_i / U_ii, then
  // t_r -= U_ri x_i for r < i, every thread on its rows (column i of U is
  // contiguous); the right-hand sides share the pass
  for (int i = n - 1; i >= 0; --i) {
    const double uii = ud[i];
    for (int e = threadIdx.x; e < i * nop; e += kNT) {
      const int j = e / i, r = e - j * i;
      t[r + (size_t)j * n] -= km[r + (size_t)i * n] * (t[i + (size_t)j * n] / uii);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nop; j += kNT) t[i + (size_t)j * n] /= uii;
    __syncthreads();
  }
  // K = sym(X[0:nop, 0:nop])  (ksym.f)
  double* K = M.kmat + (size_t)inst * n * n;
  for (int e = threadIdx.x; e < nop * nop; e += kNT) {
    const int j = e / nop, i = e - j * nop;
    K[i + (size_t)j * nop] = 0.5 * (t[i + (size_t)j * n] + t[j + (size_t)i * n]);
  }
  __syncthreads();
  // KTOS: A = 0.5*K*K + 0.5 I (SPD), A^-1 by the natural-order sweep in scratch
  double* A = km;  // reuse (nop x nop, ld nop)
  for (int e = threadIdx.x; e < nop * nop; e += kNT) {
    const int j = e / nop, i = e - j * nop;
    double s = 0.0;
    for (int k = 0; k < nop; ++k) s += K[i + (size_t)k * nop] * K[k + (size_t)j * nop];
    s = 0.5 * s;
    if (i == j) s += 0.5;
    A[i + (size_t)j * nop] = s;
  }
  __syncthreads();
  // Gauss-Jordan on the SPD matrix (no pivoting needed): natural-order sweep
  double* colk = t;  // reuse
  for (int k = 0; k < nop; ++k) {
    const double d = A[k + (size_t)k * nop];
    for (int i = threadIdx.x; i < nop; i += kNT) colk[i] = A[i + (size_t)k * nop];
    __syncthreads();
    const double dinv = 1.0 / d;
    for (int e = threadIdx.x; e < nop * nop; e += kNT) {
      const int j = e / nop, i = e - j * nop;
      double v;
      if (i == k)
        v = (j == k) ? -dinv : colk[j] * dinv;
      else if (j == k)
        v = colk[i] * dinv;
      else {
        const int a = i > j ? i : j, b = i > j ? j : i;
        v = A[i + (size_t)j * nop] - (colk[a] * dinv) * colk[b];
      }
      A[i + (size_t)j * nop] = v;
    }
    __syncthreads();
  }
  double* SR = M.sr + (size_t)inst * n * n;
  double* SI = M.si + (size_t)inst * n * n;
  for (int e = threadIdx.x; e < nop * nop; e += kNT) {
    const int j = e / nop, i = e - j * nop;
    double s = 0.0;
    for (int k = 0; k < nop; ++k) s += (-A[i + (size_t)k * nop]) * K[k + (size_t)j * nop];
    SI[i + (size_t)j * nop] = s;
    SR[i + (size_t)j * nop] = -A[i + (size_t)j * nop] - ((i == j) ? 1.0 : 0.0);
  }
  if (threadIdx.x == 0) M.status[inst] = 0;
}

}  // namespace mcp
