// Symmetric eigen-decomposition beyond the LDS-resident Jacobi solver, and
// the large-N AIRY propagator built on it.
//
// DIAGVC (source_code/diagvc.f:29-60) diagonalises with LAPACK DSYEVR: every
// eigenvalue in ascending order, the eigenvectors in the columns of EVEC.
// Here, for a batch of n x n matrices in global memory (lower triangle read),
// in four kernels:
//   k_sytrd  Householder reduction A = Q T Q^T, one workgroup per matrix.
//            Reflector j is H_j = I - tau_j v_j v_j^T with v_j(j+1) = 1 and
//            v_j(i), i >= j+2, left in A(i, j).  The two-sided update of
//            reflector j-1, A <- A - v w^T - w v^T with w = p - tau/2 (p.v) v
//            and p = tau A v, is applied in the same pass over the trailing
//            lower triangle that forms p for reflector j, so the trailing
//            matrix is read and written once per reflector.
//   k_stebz  every eigenvalue of T by bisection on the Sturm count, one
//            thread per eigenvalue, which yields them in ascending order.
//   k_stein  the eigenvectors of T by inverse iteration with partial
//            pivoting, all eigenvalues of a matrix at once (one thread
//            each); clusters of eigenvalues closer than 1e-3 |T|_1 to their
//            neighbour orthogonalised in eigenvalue order after every solve
//            (the thresholds of LAPACK's DSTEIN); the largest component
//            made positive.
//   k_ormtr  V = Q Z: the reflectors, last first, applied to columns of Z
//            held in registers (16 lanes per column).
//
// The AIRY steps of a grid (source_code/aiprop.f:86-366, TOLHI = 0) do not
// depend on the energy: the matrices -(W(RA)+W(RB))/2 and
// sqrt(3)(W(RB)-W(RA))/DRNOW of every step are built at once (k_airy_wd),
// diagonalised as one batch, and the slopes HP = diag(V^T D V) and basis
// changes T_k = V_{k-1}^T V_k formed with the batched matrix-core products.
// The steps of all energies then run as a tree of merged step maps (the
// kernels at the end of this file), or in sequence: Z <- sym(T^T Z T) (two
// products), the inversion of Z + Y1 on a team of workgroups, and
// Z <- Y4 - Y2 (Z + Y1)^-1 Y2 with the node count (k_airy_m, k_airy_z).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>

#include "mcp_airy.hpp"
#include "mcp_device.hpp"
#include "mcp_large.hpp"

namespace mcp {

constexpr int kET = 512;             // threads of k_sytrd and k_stebz
constexpr int kERows = kET / 64;     // rows per lane in k_sytrd
constexpr int kEMaxN = 64 * kERows;  // largest n of this solver
constexpr int kOGroups = 16;         // lanes per column in k_ormtr
constexpr int kORows = kEMaxN / kOGroups;

// every memory counter drained, then the barrier: global stores of one wave
// are complete when another wave of the workgroup reads them
__device__ __forceinline__ void sync_mem() {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* sv) {
  v = wave_max(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sv[w] = v;
  __syncthreads();
  double s = sv[0];
#pragma unroll
  for (int k = 1; k < NT / 64; ++k) s = fmax(s, sv[k]);
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------------------
// Householder tridiagonalisation; d, e, tau: n per matrix (e(n-1), tau(n-2),
// tau(n-1) set to 0)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kET) k_sytrd(int n, double* __restrict__ a, double* __restrict__ dg,
                                               double* __restrict__ eg, double* __restrict__ taug) {
  constexpr int NW = kET / 64;
  __shared__ double pv[kEMaxN], pw[kEMaxN], vn[kEMaxN], scol[kEMaxN];
  __shared__ double part[NW][kEMaxN];
  __shared__ double sv[NW];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const size_t nn = (size_t)n * n;
  double* A = a + blockIdx.x * nn;
  double* d = dg + (size_t)blockIdx.x * n;
  double* e = eg + (size_t)blockIdx.x * n;
  double* tau = taug + (size_t)blockIdx.x * n;
  bool pending = false;  // (pv, pw) of reflector j-1 not yet applied to columns >= j
  for (int j = 0; j + 2 < n; ++j) {
    double* col = A + (size_t)j * n;
    // column j brought up to date
    if (pending && tid < n - j) {
      const int i = j + tid;
      col[i] -= pv[i] * pw[j] + pw[i] * pv[j];
    }
    sync_mem();
    // reflector j from x = A(j+1:, j): H x = beta e_1
    double s2 = 0.0;
    if (tid < n - j - 2) {
      const double x = col[j + 2 + tid];
      s2 = x * x;
    }
    const double sigma = block_sum<kET>(s2, sv);
    const double alpha = col[j + 1];
    double beta = alpha, t = 0.0, scale = 0.0;
    if (sigma > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
      t = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    if (tid == 0) {
      d[j] = col[j];
      e[j] = beta;
      tau[j] = t;
    }
    if (tid < n - j - 1) {
      const int i = j + 1 + tid;
      double v = 1.0;
      if (i > j + 1) {
        v = col[i] * scale;
        col[i] = v;
      }
      vn[i] = v;
    }
    __syncthreads();
    // trailing columns k >= j+1, rows i >= k: the pending update, then
    // p = A22 v (lower triangle: row i of column k adds A(i,k) v(k) to p(i)
    // and, below the diagonal, A(i,k) v(i) to p(k)); column k on wave k % NW
    double rv[kERows], rpv[kERows], rpw[kERows], acc[kERows];
#pragma unroll
    for (int r = 0; r < kERows; ++r) {
      const int i = lane + 64 * r;
      const bool in = i > j && i < n;
      rv[r] = in ? vn[i] : 0.0;
      rpv[r] = (in && pending) ? pv[i] : 0.0;
      rpw[r] = (in && pending) ? pw[i] : 0.0;
      acc[r] = 0.0;
    }
    int k = j + 1 + ((wv - (j + 1)) % NW + NW) % NW;
    double x[kERows];
#pragma unroll
    for (int r = 0; r < kERows; ++r) {
      const int i = lane + 64 * r;
      x[r] = (k < n && i >= k && i < n) ? A[i + (size_t)k * n] : 0.0;
    }
    for (; k < n; k += NW) {
      const int kn = k + NW;
      double xn[kERows];
#pragma unroll
      for (int r = 0; r < kERows; ++r) {
        const int i = lane + 64 * r;
        xn[r] = (kn < n && i >= kn && i < n) ? A[i + (size_t)kn * n] : 0.0;
      }
      double* ck = A + (size_t)k * n;
      const double vk = vn[k];
      const double pvk = pending ? pv[k] : 0.0, pwk = pending ? pw[k] : 0.0;
      double sk = 0.0;
#pragma unroll
      for (int r = 0; r < kERows; ++r) {
        const int i = lane + 64 * r;
        if (i >= k && i < n) {
          double xv = x[r];
          if (pending) {
            xv -= rpv[r] * pwk + rpw[r] * pvk;
            ck[i] = xv;
          }
          acc[r] += xv * vk;
          if (i > k) sk += xv * rv[r];
        }
      }
      sk = wave_sum(sk);
      if (lane == 0) scol[k] = sk;
#pragma unroll
      for (int r = 0; r < kERows; ++r) x[r] = xn[r];
    }
#pragma unroll
    for (int r = 0; r < kERows; ++r) part[wv][lane + 64 * r] = acc[r];
    sync_mem();
    // p = tau A22 v; w = p - tau/2 (p.v) v becomes the pending update
    const int i4 = j + 1 + tid;
    double p = 0.0;
    if (i4 < n) {
      double s = scol[i4];
#pragma unroll
      for (int w = 0; w < NW; ++w) s += part[w][i4];
      p = t * s;
    }
    const double pdv = block_sum<kET>(i4 < n ? p * vn[i4] : 0.0, sv);
    if (i4 < n) {
      pw[i4] = p - 0.5 * t * pdv * vn[i4];
      pv[i4] = vn[i4];
    }
    pending = true;
    __syncthreads();
  }
  if (tid == 0) {
    if (n == 1) {
      d[0] = A[0];
      e[0] = 0.0;
      tau[0] = 0.0;
    } else {
      const int j = n - 2;
      double a00 = A[j + (size_t)j * n], a10 = A[j + 1 + (size_t)j * n], a11 = A[j + 1 + (size_t)(j + 1) * n];
      if (pending) {
        a00 -= pv[j] * pw[j] + pw[j] * pv[j];
        a10 -= pv[j + 1] * pw[j] + pw[j + 1] * pv[j];
        a11 -= pv[j + 1] * pw[j + 1] + pw[j + 1] * pv[j + 1];
      }
      d[j] = a00;
      e[j] = a10;
      d[j + 1] = a11;
      e[j + 1] = 0.0;
      tau[j] = 0.0;
      tau[j + 1] = 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// eigenvalues of the tridiagonal (d, e) by bisection
// ---------------------------------------------------------------------------
// number of eigenvalues below x (negative pivots of T - x I), pivots smaller
// than pivmin taken as -pivmin
__device__ __forceinline__ int sturm_below(const double* d, const double* e2, int n, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < 0.0 ? 1 : 0;
  for (int k = 1; k < n; ++k) {
    q = (d[k] - x) - e2[k - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0 ? 1 : 0;
  }
  return c;
}

__global__ void __launch_bounds__(kET) k_stebz(int n, const double* __restrict__ dg, const double* __restrict__ eg,
                                               double* __restrict__ evg) {
  __shared__ double d[kEMaxN], e2[kEMaxN];
  __shared__ double sv[kET / 64];
  const int tid = threadIdx.x;
  const double* dm = dg + (size_t)blockIdx.x * n;
  const double* em = eg + (size_t)blockIdx.x * n;
  double lo = DBL_MAX, hi = -DBL_MAX, emax = 0.0;
  for (int i = tid; i < n; i += kET) {
    const double di = dm[i];
    const double el = i > 0 ? fabs(em[i - 1]) : 0.0, er = i + 1 < n ? fabs(em[i]) : 0.0;
    d[i] = di;
    const double q = i + 1 < n ? em[i] * em[i] : 0.0;
    e2[i] = q;
    lo = fmin(lo, di - el - er);
    hi = fmax(hi, di + el + er);
    emax = fmax(emax, q);
  }
  lo = -block_max<kET>(-lo, sv);
  hi = block_max<kET>(hi, sv);
  emax = block_max<kET>(emax, sv);
  const double eps = DBL_EPSILON;
  const double pivmin = DBL_MIN * fmax(1.0, emax);
  const double tnorm = fmax(fabs(lo), fabs(hi));
  const double widen = 2.0 * eps * tnorm * n + 4.0 * pivmin;
  lo -= widen;
  hi += widen;
  const double atol = fmax(eps * tnorm, pivmin);
  for (int i = tid; i < n; i += kET) {
    double a0 = lo, b0 = hi;
    for (int it = 0; it < 256; ++it) {
      const double m = 0.5 * (a0 + b0);
      if (b0 - a0 <= fmax(2.0 * eps * fmax(fabs(a0), fabs(b0)), atol) || m <= a0 || m >= b0) break;
      if (sturm_below(d, e2, n, m, pivmin) > i)
        b0 = m;
      else
        a0 = m;
    }
    evg[(size_t)blockIdx.x * n + i] = 0.5 * (a0 + b0);
  }
}

// ---------------------------------------------------------------------------
// eigenvectors of the tridiagonal by inverse iteration; one workgroup per
// matrix, thread j for eigenvalue j.  Every iteration solves all n systems
// at once, then orthogonalises each cluster's iterates by modified
// Gram-Schmidt in eigenvalue order (step s: every later member of a cluster
// removes its component along the cluster's member s), so the last solve of
// each vector is followed by its orthogonalisation against the cluster's
// earlier vectors as in DSTEIN.  Interleaved storage: element k of
// eigenvalue j at [k n + j] (the threads step through k together); the
// iterates in x, the factors in work (5 n^2 per matrix).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double stein_start(int j, int k) {
  unsigned h = (unsigned)(j + 1) * 2654435761u ^ (unsigned)(k + 1) * 2246822519u;
  h ^= h >> 15;
  h *= 2654435761u;
  h ^= h >> 13;
  return (double)(h & 0xFFFFFFu) / 8388608.0 - 1.0;  // (-1, 1)
}

constexpr int kSU = 8;  // elements per batch of loads in k_stein's sweeps

__global__ void __launch_bounds__(kET) k_stein(int n, const double* __restrict__ dg, const double* __restrict__ eg,
                                              const double* __restrict__ evg, double* __restrict__ xg,
                                              double* __restrict__ work) {
  __shared__ double sd[kEMaxN], se[kEMaxN];  // the tridiagonal
  __shared__ int cst[kEMaxN];       // first eigenvalue of each eigenvalue's cluster
  __shared__ double lamv[kEMaxN];   // the shifts (DSTEIN's perturbation of close eigenvalues)
  __shared__ double snorm[kEMaxN];  // squared norms of the iterates
  __shared__ double sv[kET / 64];
  __shared__ int maxcl;
  const int j = threadIdx.x;
  const size_t mat = blockIdx.x, nn = (size_t)n * n;
  const double* ev = evg + mat * n;
  double* X = xg + mat * nn;
  double* F = work + mat * 5 * nn;  // multipliers
  double* UI = F + nn;              // reciprocal pivots of U
  double* U1 = UI + nn;             // first superdiagonal of U
  double* U2 = U1 + nn;             // second superdiagonal of U
  double* PV = U2 + nn;             // 1 where rows k, k+1 were interchanged
  double tn = 0.0;
  if (j < n) {
    const double* d = dg + mat * n;
    const double* e = eg + mat * n;
    sd[j] = d[j];
    se[j] = j + 1 < n ? e[j] : 0.0;
    tn = fabs(d[j]) + (j > 0 ? fabs(e[j - 1]) : 0.0) + (j + 1 < n ? fabs(e[j]) : 0.0);
  }
  tn = block_max<kET>(tn, sv);
  const double ortol = 1e-3 * tn;
  const double eps = DBL_EPSILON;
  const double tiny = fmax(eps * tn, DBL_MIN);
  if (j == 0) {
    int c0 = 0, mx = 1;
    double lprev = 0.0;
    for (int i = 0; i < n; ++i) {
      double lam = ev[i];
      if (i > 0 && ev[i] - ev[i - 1] > ortol) c0 = i;
      if (i > c0) {
        const double pert = 10.0 * fabs(eps * lam);
        if (lam - lprev < pert) lam = lprev + pert;
      }
      lprev = lam;
      cst[i] = c0;
      lamv[i] = lam;
      mx = max(mx, i - c0 + 1);
    }
    maxcl = mx;
  }
  __syncthreads();
  double ss = 0.0;
  if (j < n) {
    const double lam = lamv[j];
    // T - lam I = P L U by partial pivoting; pivots below tiny raised to tiny
    double dk = sd[0] - lam, duk = n > 1 ? se[0] : 0.0;
#pragma unroll 4
    for (int k = 0; k + 1 < n; ++k) {
      const size_t o = (size_t)k * n + j;
      const double sub = se[k], dn = sd[k + 1] - lam, cn = se[k + 1];
      double u;
      if (fabs(dk) >= fabs(sub)) {
        const double f = dk != 0.0 ? sub / dk : 0.0;
        F[o] = f;
        u = dk;
        U1[o] = duk;
        U2[o] = 0.0;
        PV[o] = 0.0;
        dk = dn - f * duk;
        duk = cn;
      } else {
        const double f = dk / sub;
        F[o] = f;
        u = sub;
        U1[o] = dn;
        U2[o] = cn;
        PV[o] = 1.0;
        dk = duk - f * dn;
        duk = -f * cn;
      }
      if (fabs(u) < tiny) u = u < 0.0 ? -tiny : tiny;
      UI[o] = 1.0 / u;
    }
    if (fabs(dk) < tiny) dk = dk < 0.0 ? -tiny : tiny;
    const size_t o = (size_t)(n - 1) * n + j;
    F[o] = 0.0;
    UI[o] = 1.0 / dk;
    U1[o] = 0.0;
    U2[o] = 0.0;
    PV[o] = 0.0;
    for (int k = 0; k < n; ++k) {
      const double v = stein_start(j, k);
      X[(size_t)k * n + j] = v;
      ss += v * v;
    }
  }
  for (int it = 0; it < 3; ++it) {
    if (j < n) {
      const double sc = 1.0 / sqrt(ss);
      // b <- L^-1 P^T (sc b); every batch's loads issued before its
      // recurrence (the loads of b read ahead of the stores)
      double cur = sc * X[j];
      int k = 0;
      for (; k + kSU < n; k += kSU) {
        double nx[kSU], f[kSU], pv[kSU];
#pragma unroll
        for (int u = 0; u < kSU; ++u) {
          const size_t o = (size_t)(k + u) * n + j;
          nx[u] = sc * X[o + n];
          f[u] = F[o];
          pv[u] = PV[o];
        }
#pragma unroll
        for (int u = 0; u < kSU; ++u) {
          const size_t o = (size_t)(k + u) * n + j;
          if (pv[u] != 0.0) {
            X[o] = nx[u];
            cur = cur - f[u] * nx[u];
          } else {
            X[o] = cur;
            cur = nx[u] - f[u] * cur;
          }
        }
      }
      for (; k + 1 < n; ++k) {
        const size_t o = (size_t)k * n + j;
        const double nxt = sc * X[o + n], f = F[o];
        if (PV[o] != 0.0) {
          X[o] = nxt;
          cur = cur - f * nxt;
        } else {
          X[o] = cur;
          cur = nxt - f * cur;
        }
      }
      X[(size_t)(n - 1) * n + j] = cur;
      // x <- U^-1 b, from the last row up
      double x1 = 0.0, x2 = 0.0;
      ss = 0.0;
      k = n - 1;
      for (; k + 1 >= kSU; k -= kSU) {
        double b[kSU], a1[kSU], a2[kSU], ri[kSU];
#pragma unroll
        for (int u = 0; u < kSU; ++u) {
          const size_t o = (size_t)(k - u) * n + j;
          b[u] = X[o];
          a1[u] = U1[o];
          a2[u] = U2[o];
          ri[u] = UI[o];
        }
#pragma unroll
        for (int u = 0; u < kSU; ++u) {
          const double xk = (b[u] - a1[u] * x1 - a2[u] * x2) * ri[u];
          X[(size_t)(k - u) * n + j] = xk;
          ss += xk * xk;
          x2 = x1;
          x1 = xk;
        }
      }
      for (; k >= 0; --k) {
        const size_t o = (size_t)k * n + j;
        const double xk = (X[o] - U1[o] * x1 - U2[o] * x2) * UI[o];
        X[o] = xk;
        ss += xk * xk;
        x2 = x1;
        x1 = xk;
      }
      snorm[j] = ss;
    }
    sync_mem();
    // modified Gram-Schmidt inside the clusters, all clusters together
    for (int st = 0; st + 1 < maxcl; ++st) {
      // the pivot of this thread's cluster; none for threads beyond n
      const int m = j < n ? cst[j] + st : j;
      if (m < j) {
        double acc[kSU];
#pragma unroll
        for (int u = 0; u < kSU; ++u) acc[u] = 0.0;
        int k = 0;
        for (; k + kSU <= n; k += kSU) {
          double a[kSU], b[kSU];
#pragma unroll
          for (int u = 0; u < kSU; ++u) {
            a[u] = X[(size_t)(k + u) * n + j];
            b[u] = X[(size_t)(k + u) * n + m];
          }
#pragma unroll
          for (int u = 0; u < kSU; ++u) acc[u] += a[u] * b[u];
        }
        for (; k < n; ++k) acc[0] += X[(size_t)k * n + j] * X[(size_t)k * n + m];
        double dt = 0.0;
#pragma unroll
        for (int u = 0; u < kSU; ++u) dt += acc[u];
        const double f = dt / snorm[m];
#pragma unroll
        for (int u = 0; u < kSU; ++u) acc[u] = 0.0;
        k = 0;
        for (; k + kSU <= n; k += kSU) {
          double a[kSU], b[kSU];
#pragma unroll
          for (int u = 0; u < kSU; ++u) {
            a[u] = X[(size_t)(k + u) * n + j];
            b[u] = X[(size_t)(k + u) * n + m];
          }
#pragma unroll
          for (int u = 0; u < kSU; ++u) {
            const double v = a[u] - f * b[u];
            X[(size_t)(k + u) * n + j] = v;
            acc[u] += v * v;
          }
        }
        for (; k < n; ++k) {
          const size_t o = (size_t)k * n + j;
          const double v = X[o] - f * X[(size_t)k * n + m];
          X[o] = v;
          acc[0] += v * v;
        }
        double s2 = 0.0;
#pragma unroll
        for (int u = 0; u < kSU; ++u) s2 += acc[u];
        ss = s2;
        snorm[j] = s2;
      }
      sync_mem();
    }
  }
  if (j < n) {
    // unit 2-norm, largest component (first of equals) positive
    double vmax = -1.0, vk = 0.0;
    for (int k = 0; k < n; ++k) {
      const double v = X[(size_t)k * n + j];
      if (fabs(v) > vmax) {
        vmax = fabs(v);
        vk = v;
      }
    }
    double sc = 1.0 / sqrt(ss);
    if (vk < 0.0) sc = -sc;
    for (int k = 0; k < n; ++k) X[(size_t)k * n + j] *= sc;
  }
}

// ---------------------------------------------------------------------------
// V = Q Z; grid (ceil(n/16), nmat), 256 threads: lane l of wave w holds rows
// l%16 + 16 r of column 16 bx + 4 w + l/16.  Z as k_stein leaves it
// (element i of vector c at [i n + c]); V column-major.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_ormtr(int n, const double* __restrict__ a, const double* __restrict__ taug,
                                               const double* __restrict__ zg, double* __restrict__ vg) {
  const int mat = blockIdx.y;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane & (kOGroups - 1), cw = lane / kOGroups;
  const int col = blockIdx.x * 16 + wv * 4 + cw;
  const size_t nn = (size_t)n * n;
  const double* A = a + mat * nn;
  const double* tau = taug + (size_t)mat * n;
  const bool live = col < n;
  double z[kORows];
#pragma unroll
  for (int r = 0; r < kORows; ++r) {
    const int i = g + kOGroups * r;
    z[r] = (live && i < n) ? zg[mat * nn + (size_t)i * n + col] : 0.0;
  }
  for (int j = n - 3; j >= 0; --j) {
    const double t = tau[j];
    if (t == 0.0) continue;
    const double* vj = A + (size_t)j * n;
    double v[kORows];
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < kORows; ++r) {
      const int i = g + kOGroups * r;
      v[r] = i == j + 1 ? 1.0 : ((i > j + 1 && i < n) ? vj[i] : 0.0);
      s += v[r] * z[r];
    }
#pragma unroll
    for (int off = kOGroups / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    s *= t;
#pragma unroll
    for (int r = 0; r < kORows; ++r) z[r] -= s * v[r];
  }
  if (live) {
#pragma unroll
    for (int r = 0; r < kORows; ++r) {
      const int i = g + kOGroups * r;
      if (i < n) vg[mat * nn + (size_t)col * n + i] = z[r];
    }
  }
}

// ---------------------------------------------------------------------------
// AIRY for large N
// ---------------------------------------------------------------------------
// step matrices of every step k: A_k = -(W(RA_k) + W(RB_k))/2 at zero
// energy, D_k = sqrt(3)(W(RB_k) - W(RA_k))/DRNOW_k (POTENT,
// source_code/potent.f:56-108); one thread per element, VL in registers
constexpr int kAiryMaxHam = 16;
__global__ void __launch_bounds__(256) k_airy_wd(BlockArgs B, GridArgs G, int nstep, const double* __restrict__ drnow,
                                                 double sq3, double* __restrict__ aout, double* __restrict__ dout) {
  const int n = B.n;
  const size_t nn = (size_t)n * n;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn) return;
  const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
  const double* vp = B.vl + tri_index(i, j) * (size_t)B.nham;
  double vl[kAiryMaxHam];
#pragma unroll
  for (int l = 0; l < kAiryMaxHam; ++l) vl[l] = l < B.nham ? vp[l] : 0.0;
  for (int k = 0; k < nstep; ++k) {
    double w2[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int ip = 1 + 2 * k + s;
      const double r = G.r[ip], rsq = rsq_of(r);
      double acc = 0.0;
#pragma unroll
      for (int l = 0; l < kAiryMaxHam; ++l) {
        if (l < B.nham) {
          const double pl = l < B.mxlam ? G.ptab[(size_t)ip * B.mxlam + l] : rsq;
          acc += vl[l] * pl;
        }
      }
      if (i == j) acc = hammat_diag(B, i, acc, r, rsq, 0.0);
      w2[s] = acc;
    }
    aout[k * nn + e] = -0.5 * (w2[0] + w2[1]);
    dout[k * nn + e] = (sq3 / drnow[k]) * (w2[1] - w2[0]);
  }
}

// hp(k, q) = V_k(:, q) . P_k(:, q) with P_k = D_k V_k
__global__ void k_airy_hp(int n, int nstep, const double* __restrict__ v, const double* __restrict__ p,
                          double* __restrict__ hout) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nstep * n) return;
  const size_t k = idx / n, q = idx - k * n;
  const double* vc = v + k * n * n + q * n;
  const double* pc = p + k * n * n + q * n;
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += vc[i] * pc[i];
  hout[idx] = s;
}

// Y1, Y2, Y4 of every step for every instance (SPROPN per channel), laid out
// [inst][k][3][n]
__global__ void k_airy_y(int n, int nstep, int ninst, const double* __restrict__ eig0, const double* __restrict__ hp,
                         const double* __restrict__ drnow, const double* __restrict__ ered, double* __restrict__ yout) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)ninst * nstep * n) return;
  const int i = (int)(idx % n);
  const size_t ks = idx / n;
  const int k = (int)(ks % nstep), inst = (int)(ks / nstep);
  double y1, y2, y4;
  airy_sector(drnow[k], eig0[(size_t)k * n + i] + ered[inst], hp[(size_t)k * n + i], y1, y2, y4);
  double* y = yout + ks * 3 * n;
  y[i] = y1;
  y[n + i] = y2;
  y[2 * n + i] = y4;
}

// M = (X + X^T)/2 + diag(Y1) of step k, X = T_k^T Z T_k of each instance
__global__ void k_airy_m(int n, int ninst, int k, int nstep, const double* __restrict__ x, const double* __restrict__ y,
                         double* __restrict__ m) {
  const size_t nn = (size_t)n * n;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nn * ninst) return;
  const size_t inst = idx / nn, e = idx - inst * nn;
  const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
  const double* xi = x + inst * nn;
  double v = 0.5 * (xi[i + (size_t)j * n] + xi[j + (size_t)i * n]);
  if (i == j) v += y[(inst * nstep + k) * 3 * n + i];
  m[idx] = v;
}

// Z = -Y2 M^-1 Y2 + diag(Y4) of step k; NODES += KOUNT outward, N - KOUNT
// inward (aiprop.f:205-209); an inversion failure marks the instance with
// its first failing step (status k+1)
__global__ void k_airy_z(int n, int ninst, int k, int nstep, const double* __restrict__ minv,
                         const int* __restrict__ kount, const double* __restrict__ y, double* __restrict__ z,
                         int* __restrict__ nodes, int* __restrict__ status, int outward) {
  const size_t nn = (size_t)n * n;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nn * ninst) return;
  const size_t inst = idx / nn, e = idx - inst * nn;
  const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
  const double* yy = y + (inst * nstep + k) * 3 * n;
  double v = -(yy[n + i] * minv[idx] * yy[n + j]);
  if (i == j) v += yy[2 * n + i];
  z[idx] = v;
  if (e == 0 && status[inst] == 0) {
    const int kc = kount[inst];
    if (kc < 0 || kc > n)
      status[inst] = k + 1;
    else
      nodes[inst] += outward ? kc : n - kc;
  }
}

// Z_out = (X + X^T)/2, X = V_last Z V_last^T
__global__ void k_airy_out(int n, int ninst, const double* __restrict__ x, double* __restrict__ zout) {
  const size_t nn = (size_t)n * n;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nn * ninst) return;
  const size_t inst = idx / nn, e = idx - inst * nn;
  const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
  const double* xi = x + inst * nn;
  zout[idx] = 0.5 * (xi[i + (size_t)j * n] + xi[j + (size_t)i * n]);
}

// ---------------------------------------------------------------------------
// AIRY solved in parallel over the radial grid.  Step k maps Z of basis k-1
// to Z of basis k as Z -> y4 - y2 (y1 + T^T Z T)^-1 y2 (diagonal y's,
// T = T_k orthogonal), which is Z -> Y4 - Y2^T (Z + Y1)^-1 Y2 with
// Y1 = T y1 T^T, Y2 = T y2, Y4 = y4: the map form the merges of mcp_tree.hpp
// compose.  The pivots of the merges are another elimination order of the
// block-tridiagonal matrix whose sequential pivots are y1 + T^T Z T, so
// their negative-pivot counts have the same sum (KOUNT outward; N - KOUNT
// per step inward).
// ---------------------------------------------------------------------------
// leaf segment (Y1 pending, Y2, Y4) of step k for each instance, and
// B = T_k diag(y1) in zb for the product Y1 = B T_k^T
__global__ void k_airy_leaf(int n, int ninst, int nstep, const double* __restrict__ tmat,
                            const double* __restrict__ y, double* __restrict__ zb, double* __restrict__ seg) {
  const size_t nn = (size_t)n * n;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nn * nstep * ninst) return;
  const size_t ik = idx / nn, e = idx - ik * nn;  // ik = inst * nstep + k
  const int k = (int)(ik % nstep);
  const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
  const double t = tmat[(size_t)k * nn + e];
  const double* yy = y + ik * 3 * n;
  zb[idx] = t * yy[j];
  double* sg = seg + ik * 3 * nn;
  sg[nn + e] = t * yy[n + j];
  sg[2 * nn + e] = i == j ? yy[2 * n + i] : 0.0;
}

// A0 + Y1 of the whole map, A0 the starting Z
__global__ void k_airy_final_w(int n, int ninst, const double* __restrict__ z0, const double* __restrict__ seg,
                               double* __restrict__ zb) {
  const size_t nn = (size_t)n * n;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nn * ninst) return;
  const size_t inst = idx / nn, e = idx - inst * nn;
  const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
  zb[idx] = z0[idx] + sym_tile(seg + inst * 3 * nn, n, i, j);
}

__global__ void k_airy_tree_nodes(int ninst, int nstep, int n, int outward, const int* __restrict__ acc,
                                  const int* __restrict__ bad, int* __restrict__ nodes, int* __restrict__ status) {
  const int inst = blockIdx.x * blockDim.x + threadIdx.x;
  if (inst >= ninst) return;
  nodes[inst] = outward ? acc[inst] : nstep * n - acc[inst];
  status[inst] = bad[inst] ? 1 : 0;
}

}  // namespace mcp
