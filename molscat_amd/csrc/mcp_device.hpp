// Device building blocks shared by the mcprop kernels (gfx950, wave64).
//
//  * block_argmax      first-index arg-max of a per-thread candidate over a
//                      workgroup (DPP/shuffle inside the wave, LDS across waves)
//  * sweep_inv         in-place inverse of a symmetric matrix held in LDS
//                      (both triangles) by the symmetric sweep operator, with
//                      the inertia count SYMINV returns (source_code/syminv.f).
//  * jacobi_eigh       parallel (round-robin) cyclic Jacobi eigen-decomposition
//                      of a symmetric matrix in LDS (stands in for DSYEVR as
//                      called by DIAGVC, source_code/diagvc.f:46).
//
// All matrices are column-major with leading dimension `lda`.
#pragma once
#include <hip/hip_runtime.h>
#include <climits>
#include <cmath>

namespace mcp {

constexpr double kBkAlpha = 0.6403882032022076;  // (1 + sqrt(17)) / 8, Bunch-Kaufman

// 2-D thread tiling used for N x N element sweeps: lanes run down a column
// (consecutive LDS addresses), waves step across columns.
template <int NT>
struct Tile2D {
  static constexpr int kWaves = NT / 64;
  int ti, tj;
  __device__ Tile2D() : ti(threadIdx.x & 63), tj(threadIdx.x >> 6) {}
};

template <int NT>
__device__ __forceinline__ void block_argmax(double& v, int& idx, double* sv, int* si) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sv[w] = v;
    si[w] = idx;
  }
  __syncthreads();
  v = sv[0];
  idx = si[0];
#pragma unroll
  for (int k = 1; k < NT / 64; ++k) {
    const double ov = sv[k];
    const int oi = si[k];
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
  __syncthreads();
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sv) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sv[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < NT / 64; ++k) s += sv[k];
  __syncthreads();
  return s;
}

// LDS scratch needed by sweep_inv besides the matrix: 4 n-vectors of double,
// n ints, NT/64 doubles + NT/64 ints.
struct SweepScratch {
  double* cu;
  double* cv;
  double* cx;
  double* cy;
  int* perm;
  double* sv;
  int* si;
};

// In-place symmetric inverse by the sweep operator.
//
// Sweeping index p of a symmetric matrix (pivot d = A[p,p]) maps
//   A[p,p] -> -1/d,  A[i,p] = A[p,i] -> A[i,p]/d,  A[i,j] -> A[i,j] - A[i,p]A[p,j]/d,
// and sweeping every index yields -inv(A).  The unswept part of the matrix is
// always the Schur complement, so the pivots are exactly the D of an L.D.L^T
// factorisation taken in the same order, and the inertia is their sign count.
//
//   bk == false : natural order 0..n-1, the pivot sequence of the in-situ
//                 Wilkinson-Reinsch L.D.L^T of SYMINV for N <= 30
//                 (source_code/syminv.f:21-49); returns KOUNT, or n+1 on an
//                 exactly zero pivot (syminv.f:95).
//   bk == true  : Bunch-Kaufman choice of 1x1 / 2x2 pivots, the decisions of
//                 DSYTF2('L') (called through DSYTRF at
//                 source_code/syminl-alloc.f:21), tracked through a position
//                 permutation instead of moving rows; inertia counted as
//                 DSYNEG does (source_code/dsyneg.f:47-73).  Returns the
//                 inertia, or -(k+1) when the candidate column is exactly zero
//                 (DSYTRF INFO = k+1).
//
// Each element update is computed from the unordered index pair (max, min), so
// the matrix stays bitwise symmetric.  Must be called by all NT threads.
template <int NT>
__device__ int sweep_inv(double* A, int n, int lda, bool bk, const SweepScratch& s) {
  const int tid = threadIdx.x;
  const Tile2D<NT> t2;
  int kount = 0;
  if (bk) {
    for (int i = tid; i < n; i += NT) s.perm[i] = i;
  }
  __syncthreads();
  int k = 0;
  while (k < n) {
    int p = k, q = -1;
    if (bk) {
      const int ck = s.perm[k];
      const double absakk = fabs(A[ck + ck * lda]);
      double v = -1.0;
      int idx = INT_MAX;
      for (int i = k + 1 + tid; i < n; i += NT) {
        const double tv = fabs(A[s.perm[i] + ck * lda]);
        if (tv > v) {
          v = tv;
          idx = i;
        }
      }
      block_argmax<NT>(v, idx, s.sv, s.si);
      const double colmax = v < 0.0 ? 0.0 : v;
      const int imaxpos = idx;
      if (fmax(absakk, colmax) == 0.0) return -(k + 1);
      int kp = k, kstep = 1;
      if (!(absakk >= kBkAlpha * colmax)) {
        const int ci = s.perm[imaxpos];
        double rv = -1.0;
        int ri = INT_MAX;
        for (int j = k + tid; j < n; j += NT) {
          if (j == imaxpos) continue;
          const double tv = fabs(A[ci + s.perm[j] * lda]);
          if (tv > rv) {
            rv = tv;
            ri = j;
          }
        }
        block_argmax<NT>(rv, ri, s.sv, s.si);
        const double rowmax = rv;
        if (absakk >= kBkAlpha * colmax * (colmax / rowmax)) {
          kp = k;
        } else if (fabs(A[ci + ci * lda]) >= kBkAlpha * rowmax) {
          kp = imaxpos;
        } else {
          kp = imaxpos;
          kstep = 2;
        }
      }
      const int kk = k + kstep - 1;
      const int pkk = s.perm[kk];
      const int pkp = s.perm[kp];
      __syncthreads();
      if (kp != kk && tid == 0) {
        s.perm[kk] = pkp;
        s.perm[kp] = pkk;
      }
      if (kstep == 1) {
        p = (kp != kk) ? pkp : pkk;
      } else {
        p = ck;
        q = (kp != kk) ? pkp : pkk;
      }
    }
    if (q < 0) {
      const double d = A[p + p * lda];
      if (d == 0.0) return n + 1;
      if (d < 0.0) ++kount;
      const double dinv = 1.0 / d;
      for (int i = tid; i < n; i += NT) {
        const double c = A[i + p * lda];
        s.cu[i] = c;
        s.cx[i] = c * dinv;
      }
      __syncthreads();
      for (int j = t2.tj; j < n; j += Tile2D<NT>::kWaves) {
        const bool jp = (j == p);
        for (int i = t2.ti; i < n; i += 64) {
          double val;
          if (i == p) {
            val = jp ? -dinv : s.cx[j];
          } else if (jp) {
            val = s.cx[i];
          } else {
            const int a = i > j ? i : j;
            const int b = i > j ? j : i;
            val = A[i + j * lda] - s.cx[a] * s.cu[b];
          }
          A[i + j * lda] = val;
        }
      }
      __syncthreads();
      k += 1;
    } else {
      const double a = A[p + p * lda];
      const double b = A[q + p * lda];
      const double c = A[q + q * lda];
      {
        const double tt = fabs(b);
        const double dd = (c / tt) * a - tt;
        if (dd < 0.0) ++kount;
      }
      const double det = a * c - b * b;
      const double d00 = c / det, d01 = -b / det, d11 = a / det;
      for (int i = tid; i < n; i += NT) {
        const double u = A[i + p * lda];
        const double w = A[i + q * lda];
        s.cu[i] = u;
        s.cv[i] = w;
        s.cx[i] = u * d00 + w * d01;
        s.cy[i] = u * d01 + w * d11;
      }
      __syncthreads();
      for (int j = t2.tj; j < n; j += Tile2D<NT>::kWaves) {
        const bool jp = (j == p), jq = (j == q);
        for (int i = t2.ti; i < n; i += 64) {
          const bool ip = (i == p), iq = (i == q);
          double val;
          if ((ip || iq) && (jp || jq)) {
            val = -((ip && jp) ? d00 : ((iq && jq) ? d11 : d01));
          } else if (ip || iq) {
            val = ip ? s.cx[j] : s.cy[j];
          } else if (jp || jq) {
            val = jp ? s.cx[i] : s.cy[i];
          } else {
            const int aa = i > j ? i : j;
            const int bb = i > j ? j : i;
            val = A[i + j * lda] - (s.cx[aa] * s.cu[bb] + s.cy[aa] * s.cv[bb]);
          }
          A[i + j * lda] = val;
        }
      }
      __syncthreads();
      k += 2;
    }
  }
  for (int j = t2.tj; j < n; j += Tile2D<NT>::kWaves)
    for (int i = t2.ti; i < n; i += 64) A[i + j * lda] = -A[i + j * lda];
  __syncthreads();
  return kount;
}

// Round-robin pairing for parallel Jacobi: m (even) indices, round r in
// [0, m-1), slot t in [0, m/2).  Index m-1 stays fixed, the others rotate.
__device__ __forceinline__ void rr_pair(int m, int r, int t, int& p, int& q) {
  const int mm = m - 1;
  int a, b;
  if (t == 0) {
    a = mm;
    b = r;
  } else {
    a = (r + t) % mm;
    b = (r - t + mm) % mm;
  }
  p = a < b ? a : b;
  q = a < b ? b : a;
}

// Parallel cyclic Jacobi: on exit A is (numerically) diagonal and V holds the
// eigenvectors as columns (A_in = V diag(A_out) V^T).  rot holds 4*(m/2)
// doubles/ints of LDS scratch: cos, sin, p, q.  Returns the number of sweeps.
template <int NT>
__device__ int jacobi_eigh(double* A, double* V, int n, int lda, double* rc, double* rs, int* rp, int* rq,
                           double* sv, int max_sweeps = 30) {
  const int tid = threadIdx.x;
  const Tile2D<NT> t2;
  for (int j = t2.tj; j < n; j += Tile2D<NT>::kWaves)
    for (int i = t2.ti; i < n; i += 64) V[i + j * lda] = (i == j) ? 1.0 : 0.0;
  __syncthreads();
  if (n == 1) return 0;
  const int m = (n + 1) & ~1;
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int j = t2.tj; j < n; j += Tile2D<NT>::kWaves)
      for (int i = t2.ti; i < n; i += 64) {
        const double a = A[i + j * lda];
        if (i == j)
          dia += a * a;
        else
          off += a * a;
      }
    off = block_sum<NT>(off, sv);
    dia = block_sum<NT>(dia, sv);
    if (off <= 1.0e-30 * dia || off == 0.0) break;
    for (int r = 0; r < m - 1; ++r) {
      for (int t = tid; t < m / 2; t += NT) {
        int p, q;
        rr_pair(m, r, t, p, q);
        double c = 1.0, sn = 0.0;
        if (q < n) {
          const double apq = A[p + q * lda];
          if (apq != 0.0) {
            const double app = A[p + p * lda], aqq = A[q + q * lda];
            const double theta = (aqq - app) / (2.0 * apq);
            const double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(tt * tt + 1.0);
            sn = tt * c;
          }
        }
        rc[t] = c;
        rs[t] = sn;
        rp[t] = p;
        rq[t] = q;
      }
      __syncthreads();
      // rows: A <- J^T A  (rows p and q of each pair)
      for (int e = tid; e < (m / 2) * n; e += NT) {
        const int t = e / n, j = e - t * n;
        const int p = rp[t], q = rq[t];
        if (q >= n) continue;
        const double c = rc[t], sn = rs[t];
        const double ap = A[p + j * lda], aq = A[q + j * lda];
        A[p + j * lda] = c * ap - sn * aq;
        A[q + j * lda] = sn * ap + c * aq;
      }
      __syncthreads();
      // columns: A <- A J, V <- V J
      for (int e = tid; e < (m / 2) * n; e += NT) {
        const int t = e / n, i = e - t * n;
        const int p = rp[t], q = rq[t];
        if (q >= n) continue;
        const double c = rc[t], sn = rs[t];
        const double ap = A[i + p * lda], aq = A[i + q * lda];
        A[i + p * lda] = c * ap - sn * aq;
        A[i + q * lda] = sn * ap + c * aq;
        const double vp = V[i + p * lda], vq = V[i + q * lda];
        V[i + p * lda] = c * vp - sn * vq;
        V[i + q * lda] = sn * vp + c * vq;
      }
      __syncthreads();
    }
  }
  return sweep;
}

}  // namespace mcp
