// Large-N path (matrices that do not fit in LDS, e.g. N = 441): blocked
// symmetric sweep with the Bunch-Kaufman pivot choice of dsytf2.f on a matrix
// in global memory, stored in its lower triangle like the reference's
// propagators (mdprop.f:89-93), with one panel of up to kPB pivots in LDS.
//
// Panel step.  The kPB candidate columns are loaded into LDS.  Wavefront 0
// factors the candidate block S = L D L^T in natural order; every thread then
// forms W = A(:,P) L^-T, whose column k is the Schur column of pivot k at the
// moment it would be taken, so the Bunch-Kaufman acceptance test
// |d_k| >= alpha * colmax_k is evaluated for all candidates at once.  The
// leading run of candidates that passes is swept as one block pivot
// (X = A(:,P) S^-1, A <- A - X A(P,:), A(:,P) <- X, A(P,P) <- -S^-1).  When the
// first candidate fails, the panel falls back to pivot-by-pivot selection
// with row interchanges (wave_panel).
//
// One instance (one matrix, or one propagation) runs on a team of g
// workgroups; the lower triangle is cut into kTile x kTile tiles and tile t
// belongs to rank t % g.  Every workgroup repeats the panel factorisation on
// identical data, applies the update to its own tiles, and a per-instance
// barrier separates panels (g = 1 needs none).
#pragma once
#include <hip/hip_runtime.h>

#include "mcp_device.hpp"
#include "mcp_kernels.hpp"

namespace mcp {

typedef double dvec4 __attribute__((ext_vector_type(4)));

// D = A B + C on the FP64 matrix cores (v_mfma_f64_16x16x4f64): A operand of
// lane l = element (l%16, l/16) of the 16x4 factor, B operand = element
// (l/16, l%16) of the 4x16 factor, register v of the accumulator = element
// (l/16 + 4v, l%16) of the 16x16 product.
__device__ __forceinline__ dvec4 mma(double a, double b, dvec4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kPB = 16;    // pivots per panel
constexpr int kMaxW = 8;   // wavefronts per workgroup the scratch is laid out for (512 threads)
constexpr int kTile = 64;  // ownership / update tile edge

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void wave_argmax(double& v, int& idx) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// element (i, c) of a symmetric matrix held in its lower triangle
__device__ __forceinline__ double lget(const double* A, int n, int i, int c) {
  return i >= c ? A[i + (size_t)c * n] : A[c + (size_t)i * n];
}

// --------------------------------------------------------------------------
// per-instance workgroup team and barrier
// --------------------------------------------------------------------------
struct Team {
  int g;          // workgroups per instance
  int rank;       // this workgroup's rank in [0, g)
  unsigned* ctr;  // barrier counter of the instance (zero at launch)
  unsigned epoch; // barriers passed so far
  int* abort;     // set when a barrier wait exceeds its bound
};

constexpr unsigned kSpinLimit = 1u << 26;  // bounded wait (~seconds), never a hang

__device__ inline bool team_sync(Team& t) {
  // every wave's stores have reached the L2 before thread 0's release writes
  // it back (a workgroup barrier alone does not wait for them)
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (t.g == 1) return true;
  __shared__ int s_ok;
  if (threadIdx.x == 0) {
    t.epoch += 1;
    const unsigned target = t.epoch * (unsigned)t.g;
    __hip_atomic_fetch_add(t.ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    int ok = 1;
    // relaxed polls and one acquire after: an acquire per poll would
    // invalidate the caches on every iteration
    while (__hip_atomic_load(t.ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > kSpinLimit ||
          __hip_atomic_load(t.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        __hip_atomic_store(t.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    s_ok = ok;
  } else {
    t.epoch += 1;
  }
  __syncthreads();
  return s_ok != 0;
}

// tile row of lower-triangle tile number tau (tiles numbered row by row)
__device__ __forceinline__ int tile_row(int tau) {
  int bi = 0;
  while ((bi + 1) * (bi + 2) / 2 <= tau) ++bi;
  return bi;
}

// owner rank of lower element (r, c), r >= c
__device__ __forceinline__ bool owns(const Team& t, int r, int c) {
  if (t.g == 1) return true;
  const int bi = r / kTile, bj = c / kTile;
  return (bi * (bi + 1) / 2 + bj) % t.g == t.rank;
}

// f(i, j) for every lower element (i >= j) of this workgroup's tiles; within
// a tile consecutive threads take consecutive rows of one column
template <int NT = kNT, class F>
__device__ __forceinline__ void for_owned_lower(const Team& t, int n, F&& f) {
  const int nt = (n + kTile - 1) / kTile;
  const int ntt = nt * (nt + 1) / 2;
  for (int tau = t.rank; tau < ntt; tau += t.g) {
    const int bi = tile_row(tau);
    const int bj = tau - bi * (bi + 1) / 2;
    for (int e = threadIdx.x; e < kTile * kTile; e += NT) {
      const int i = bi * kTile + (e & (kTile - 1));
      const int j = bj * kTile + e / kTile;
      if (i < n && j < n && i >= j) f(i, j);
    }
  }
}

// leading dimension of the X and G panels in LDS: n rounded up to whole tiles
__host__ __device__ constexpr int panel_ld(int n) { return (n + kTile - 1) / kTile * kTile; }

struct LargeScratch {
  int ld;        // panel_ld(n)
  double* X;     // [kPB][ld] A(:,P) S^-1 (block path) / pivot-by-pivot factors
  double* G;     // [kPB][ld] candidate columns of the panel (slots)
  double* Dinv;  // [kPB*kPB] inverse of the panel pivot block
  double* Lb;    // [kPB*kPB] LDL^T of the candidate block (D on the diagonal)
  double* Li;    // [kPB*kPB] L^-1
  double* cmax;  // [kMaxW][kPB] per-wave partial column maxima
  double* mc;    // [n] current column of the candidate
  double* mci;   // [n] current column of the row candidate
  double* wv;    // [2*kPB]
  double* gc;    // [kPB] G row of the candidate at the panel pivots
  int* perm;     // [n] pivot order (positions -> index)
  int* inP;      // [n] panel pivot number + 1, or 0
  int* swept;    // [n] 1 once the index has been pivoted
  int* pidx;     // [kPB] pivot indices of the panel
  int* pslot;    // [kPB] G slot of each pivot
  int* slotpos;  // [kPB] G slot of panel position kstart+s
  int* ctl;      // [8] panel control shared with the other waves
  double* sv;
  int* si;
};

__device__ inline LargeScratch carve_large(LdsCarve& c, int n) {
  LargeScratch s;
  s.ld = panel_ld(n);
  s.X = c.take<double>((size_t)kPB * s.ld);
  s.G = c.take<double>((size_t)kPB * s.ld);
  s.Dinv = c.take<double>(kPB * kPB);
  s.Lb = c.take<double>(kPB * kPB);
  s.Li = c.take<double>(kPB * kPB);
  s.cmax = c.take<double>(kMaxW * kPB);
  s.mc = c.take<double>(n);
  s.mci = c.take<double>(n);
  s.wv = c.take<double>(2 * kPB);
  s.gc = c.take<double>(kPB);
  s.perm = c.take<int>(n);
  s.inP = c.take<int>(n);
  s.swept = c.take<int>(n);
  s.pidx = c.take<int>(kPB);
  s.pslot = c.take<int>(kPB);
  s.slotpos = c.take<int>(kPB);
  s.ctl = c.take<int>(8);
  s.sv = c.take<double>(kMaxW);
  s.si = c.take<int>(kMaxW);
  return s;
}

// --------------------------------------------------------------------------
// block step on the candidate window
// --------------------------------------------------------------------------
// broadcast of a double from lane src of the wavefront
__device__ __forceinline__ double bcast(double x, int src) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, src);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Natural-order block pivot over the candidate window (positions kstart ..
// kstart+npre-1, loaded in G slots 0..npre-1).  All threads; returns the
// number np of accepted pivots (0: use the pivot-by-pivot path).  On np > 0:
// X holds A(:,P) S^-1 (rows outside P), Dinv holds S^-1, pidx/pslot/inP the
// pivots, ctl[1] the inertia of the accepted pivots.
template <int NT = kNT>
__device__ __attribute__((always_inline)) int block_panel(int n, int kstart, int npre, bool bk,
                                                          const LargeScratch& s) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int ld = s.ld;
  if (tid < 64) {
    // S = L D L^T in natural order: lane j holds row j of the Schur
    // complement; the pivot row is broadcast from its lane (dsytf2.f 1x1 step)
    if (lane < npre) s.inP[s.perm[kstart + lane]] = lane + 1;
    const int pj = lane < npre ? s.perm[kstart + lane] : 0;
    double r[kPB];
#pragma unroll
    for (int c = 0; c < kPB; ++c) r[c] = (lane < npre && c < npre) ? s.G[c * ld + pj] : 0.0;
    int f = npre;
#pragma unroll
    for (int k = 0; k < kPB; ++k) {
      if (k < f) {
        const double d = bcast(r[k], k);
        if (d == 0.0) {
          f = k;
        } else {
          const double dinv = 1.0 / d;
          const double l = r[k] * dinv;
#pragma unroll
          for (int c = k + 1; c < kPB; ++c) {
            const double rkc = bcast(r[c], k);
            if (lane > k) r[c] -= l * rkc;
          }
          if (lane > k) r[k] = l;
        }
      }
    }
    if (lane < kPB) {
#pragma unroll
      for (int t = 0; t < kPB; ++t) s.Lb[lane + t * kPB] = (t <= lane && t < f && lane < f) ? r[t] : 0.0;
    }
    wave_sync();
    // L^-1 by columns: lane c forms L^-1 e_c (rows of the factor beyond f are 0)
    if (lane < kPB) {
      double x[kPB];
#pragma unroll
      for (int j = 0; j < kPB; ++j) {
        double v = (j == lane) ? 1.0 : 0.0;
#pragma unroll
        for (int m = 0; m < j; ++m) v -= s.Lb[j + m * kPB] * x[m];
        x[j] = (j < f && lane < f) ? v : 0.0;
      }
#pragma unroll
      for (int j = 0; j < kPB; ++j) s.Li[j + lane * kPB] = x[j];
    }
    if (lane == 0) s.ctl[4] = f;
  }
  __syncthreads();
  const int f = s.ctl[4];
  // W = A(:,P) L^-T: column k is the Schur column of candidate k at its pivot
  // time; maxima over the rows still unpivoted then (remaining rows, and
  // window rows after k)
  double cm[kPB];
#pragma unroll
  for (int k = 0; k < kPB; ++k) cm[k] = 0.0;
  for (int i = tid; i < n; i += NT) {
    const int rk = s.swept[i] ? 0 : (s.inP[i] ? s.inP[i] - 1 : kPB);
    double g[kPB];
#pragma unroll
    for (int t = 0; t < kPB; ++t) g[t] = s.G[t * ld + i];
#pragma unroll
    for (int k = 0; k < kPB; ++k) {
      double w = 0.0;
#pragma unroll
      for (int t = 0; t <= k; ++t) w += g[t] * s.Li[k + t * kPB];
      if (k < rk) cm[k] = fmax(cm[k], fabs(w));
    }
  }
#pragma unroll
  for (int k = 0; k < kPB; ++k) {
    const double v = wave_max(cm[k]);
    if (lane == 0) s.cmax[wid * kPB + k] = v;
  }
  __syncthreads();
  // accepted run: the candidates before the first one that fails the test;
  // then S^-1 = L^-T D^-1 L^-1 of the accepted block
  if (tid < 64) {
    bool pass = false, neg = false;
    double d = 1.0;
    if (lane < f) {
      double colmax = 0.0;
#pragma unroll
      for (int wq = 0; wq < NT / 64; ++wq) colmax = fmax(colmax, s.cmax[wq * kPB + lane]);
      d = s.Lb[lane + lane * kPB];
      pass = !bk || (fabs(d) >= kBkAlpha * colmax);
      neg = d < 0.0;
    }
    const int np = __ffsll((unsigned long long)__ballot(!pass)) - 1;  // lanes >= f fail
    const int kount = __popcll((unsigned long long)__ballot(neg) & ((1ull << np) - 1ull));
    if (lane >= np && lane < npre) s.inP[s.perm[kstart + lane]] = 0;
    if (lane < kPB) {
      s.pidx[lane] = lane < np ? s.perm[kstart + lane] : 0;
      s.pslot[lane] = lane;
      s.wv[lane] = lane < np ? 1.0 / d : 0.0;  // D^-1 of the accepted pivots
    }
    if (lane == 0) {
      s.ctl[0] = np;
      s.ctl[1] = kount;
      s.ctl[2] = 0;
      s.ctl[3] = kstart + np;
    }
    wave_sync();
    for (int e = lane; e < kPB * kPB; e += 64) {
      const int a = e & (kPB - 1), b = e / kPB;
      double v = 0.0;
#pragma unroll
      for (int k = 0; k < kPB; ++k)
        if (k >= a && k >= b && k < np) v += s.Li[k + a * kPB] * s.wv[k] * s.Li[k + b * kPB];
      s.Dinv[e] = v;
    }
  }
  __syncthreads();
  const int np = s.ctl[0];
  if (np == 0) return 0;
  // X = A(:,P) S^-1 for the rows outside P (S^-1 is zero beyond np)
  for (int i = tid; i < n; i += NT) {
    if (s.inP[i]) continue;
    double g[kPB];
#pragma unroll
    for (int t = 0; t < kPB; ++t) g[t] = s.G[t * ld + i];
#pragma unroll
    for (int q = 0; q < kPB; ++q) {
      double x = 0.0;
#pragma unroll
      for (int t = 0; t < kPB; ++t) x += g[t] * s.Dinv[t + q * kPB];
      s.X[q * ld + i] = x;
    }
  }
  __syncthreads();
  return np;
}

// --------------------------------------------------------------------------
// pivot-by-pivot panel with Bunch-Kaufman interchanges (wavefront 0)
// --------------------------------------------------------------------------
// v = a_c - X * G[c, pslot[.]] for the original column a_c of index c (a_c
// from a G slot, or gathered from the lower-stored matrix when slot < 0).
// X rows at and beyond np are zero.
__device__ __forceinline__ void wave_cur_col(const double* A, int n, int c, int slot, int np, const LargeScratch& s,
                                             double* v) {
  const int lane = threadIdx.x & 63;
  if (lane < kPB) s.gc[lane] = (lane < np) ? s.G[(size_t)s.pslot[lane] * s.ld + c] : 0.0;
  wave_sync();
  const double* ac = slot >= 0 ? s.G + (size_t)slot * s.ld : nullptr;
  for (int i = lane; i < n; i += 64) {
    double x = ac ? ac[i] : lget(A, n, i, c);
#pragma unroll
    for (int t = 0; t < kPB; ++t) x -= s.X[(size_t)t * s.ld + i] * s.gc[t];
    v[i] = x;
  }
  wave_sync();
}

// Leaves in s.ctl: [0] pivots in the panel, [1] inertia of the panel, [2]
// error code (0, -(k+1) for a zero candidate column, n+1 for a zero pivot),
// [3] next position.  X must be zero on entry.  The caller makes the results
// visible with a workgroup barrier.
__device__ __attribute__((always_inline)) void wave_panel(const double* A, int n, int kstart, bool bk, const LargeScratch& s) {
  const int lane = threadIdx.x & 63;
  int np = 0, kount = 0, err = 0;
  int k = kstart;
  const int npre = min(kPB, n - kstart);
  while (np <= kPB - 2 && k < n) {
    const int ck = s.perm[k];
    wave_cur_col(A, n, ck, s.slotpos[np], np, s, s.mc);
    int p = ck, q = -1;
    const double* colp = s.mc;
    if (bk) {
      const double absakk = fabs(s.mc[ck]);
      double v = -1.0;
      int idx = INT_MAX;
      for (int i = k + 1 + lane; i < n; i += 64) {
        const double tv = fabs(s.mc[s.perm[i]]);
        if (tv > v) {
          v = tv;
          idx = i;
        }
      }
      wave_argmax(v, idx);
      const double colmax = v < 0.0 ? 0.0 : v;
      if (fmax(absakk, colmax) == 0.0) {
        err = -(k + 1);
        break;
      }
      if (!(absakk >= kBkAlpha * colmax)) {
        const int imaxpos = idx;
        const int ci = s.perm[imaxpos];
        const bool pre = (imaxpos - kstart) < npre;
        wave_cur_col(A, n, ci, pre ? s.slotpos[imaxpos - kstart] : -1, np, s, s.mci);
        double rv = -1.0;
        int ri = INT_MAX;
        for (int j = k + lane; j < n; j += 64) {
          if (j == imaxpos) continue;
          const double tv = fabs(s.mci[s.perm[j]]);
          if (tv > rv) {
            rv = tv;
            ri = j;
          }
        }
        wave_argmax(rv, ri);
        const double rowmax = rv;
        int kp = k, kstep = 1;
        if (absakk >= kBkAlpha * colmax * (colmax / rowmax)) {
          kp = k;
        } else if (fabs(s.mci[ci]) >= kBkAlpha * rowmax) {
          kp = imaxpos;
          p = ci;
          colp = s.mci;
        } else {
          kp = imaxpos;
          kstep = 2;
          p = ck;
          q = ci;
        }
        const int kk = k + kstep - 1;
        if (kp != kk) {
          const int pkk = s.perm[kk], pkp = s.perm[kp];
          const int skk = s.slotpos[kk - kstart];
          const int skp = pre ? s.slotpos[kp - kstart] : -1;
          wave_sync();
          if (lane == 0) {
            s.perm[kk] = pkp;
            s.perm[kp] = pkk;
            if (pre) {
              s.slotpos[kk - kstart] = skp;
              s.slotpos[kp - kstart] = skk;
            }
          }
          if (!pre) {
            double* dst = s.G + (size_t)skk * s.ld;
            for (int i = lane; i < n; i += 64) dst[i] = lget(A, n, i, pkp);
          }
          wave_sync();
        }
      }
    }
    if (q < 0) {
      const double d = colp[p];
      if (d == 0.0) {
        err = n + 1;
        break;
      }
      if (d < 0.0) ++kount;
      const double dinv = 1.0 / d;
      if (lane < kPB) s.wv[lane] = (lane < np) ? s.X[(size_t)lane * s.ld + p] : 0.0;
      wave_sync();
      for (int i = lane; i < n; i += 64) {
        const double md = colp[i] * dinv;
#pragma unroll
        for (int t = 0; t < kPB; ++t) {
          const double x = s.X[(size_t)t * s.ld + i];
          s.X[(size_t)t * s.ld + i] = (t < np) ? x - md * s.wv[t] : ((t == np) ? md : x);
        }
      }
      for (int e = lane; e < (np + 1) * (np + 1); e += 64) {
        const int a = e % (np + 1), b = e / (np + 1);
        double v;
        if (a < np && b < np)
          v = s.Dinv[a + b * kPB] + (s.wv[a > b ? a : b] * dinv) * s.wv[a > b ? b : a];
        else if (a < np)
          v = -s.wv[a] * dinv;
        else if (b < np)
          v = -s.wv[b] * dinv;
        else
          v = dinv;
        s.Dinv[a + b * kPB] = v;
      }
      if (lane == 0) {
        s.pidx[np] = p;
        s.pslot[np] = s.slotpos[np];
        s.inP[p] = np + 1;
      }
      wave_sync();
      np += 1;
      k += 1;
    } else {
      const double a = s.mc[p], b = s.mc[q], c = s.mci[q];
      {
        const double tt = fabs(b);
        const double dd = (c / tt) * a - tt;
        if (dd < 0.0) ++kount;
      }
      const double det = a * c - b * b;
      const double s00 = c / det, s01 = -b / det, s11 = a / det;
      if (lane < kPB) {
        s.wv[lane] = (lane < np) ? s.X[(size_t)lane * s.ld + p] : 0.0;
        s.wv[kPB + lane] = (lane < np) ? s.X[(size_t)lane * s.ld + q] : 0.0;
      }
      wave_sync();
      for (int i = lane; i < n; i += 64) {
        const double m0 = s.mc[i], m1 = s.mci[i];
        const double z0 = m0 * s00 + m1 * s01;
        const double z1 = m0 * s01 + m1 * s11;
#pragma unroll
        for (int t = 0; t < kPB; ++t) {
          const double x = s.X[(size_t)t * s.ld + i];
          s.X[(size_t)t * s.ld + i] =
              (t < np) ? x - (z0 * s.wv[t] + z1 * s.wv[kPB + t]) : ((t == np) ? z0 : ((t == np + 1) ? z1 : x));
        }
      }
      for (int e = lane; e < (np + 2) * (np + 2); e += 64) {
        const int aa = e % (np + 2), bb = e / (np + 2);
        double v;
        if (aa < np && bb < np) {
          const int hi = aa > bb ? aa : bb, lo = aa > bb ? bb : aa;
          const double wph = s.wv[hi], wqh = s.wv[kPB + hi];
          const double wpl = s.wv[lo], wql = s.wv[kPB + lo];
          v = s.Dinv[aa + bb * kPB] + (wph * (s00 * wpl + s01 * wql) + wqh * (s01 * wpl + s11 * wql));
        } else if (aa < np || bb < np) {
          const int r = aa < np ? aa : bb;
          const int col = aa < np ? bb - np : aa - np;
          const double wp = s.wv[r], wq = s.wv[kPB + r];
          v = (col == 0) ? -(wp * s00 + wq * s01) : -(wp * s01 + wq * s11);
        } else {
          const int a2 = aa - np, b2 = bb - np;
          v = (a2 == 0 && b2 == 0) ? s00 : ((a2 == 1 && b2 == 1) ? s11 : s01);
        }
        s.Dinv[aa + bb * kPB] = v;
      }
      if (lane == 0) {
        s.pidx[np] = p;
        s.pidx[np + 1] = q;
        s.pslot[np] = s.slotpos[np];
        s.pslot[np + 1] = s.slotpos[np + 1];
        s.inP[p] = np + 1;
        s.inP[q] = np + 2;
      }
      wave_sync();
      np += 2;
      k += 2;
    }
  }
  if (lane == 0) {
    s.ctl[0] = np;
    s.ctl[1] = kount;
    s.ctl[2] = err;
    s.ctl[3] = k;
  }
}

// --------------------------------------------------------------------------
// panel update of this workgroup's tiles
// --------------------------------------------------------------------------
// A(i,j) -= sum_q X(i,q) A(p_q,j) for i, j outside the panel (64 x 64 tiles on
// the matrix cores), then the panel rows/columns
// A(i,p_q) = X(i,q) and the pivot block -Dinv.  X rows beyond np are zero.
template <int NT = kNT>
__device__ __attribute__((always_inline)) void apply_panel(double* A, int n, int np, const LargeScratch& s,
                                                           const Team& t) {
  if (np == 0) return;
  const int ld = s.ld;
  // 64 x 64 tiles of the lower triangle on the matrix cores: wave w takes the
  // 32 x 32 quadrant (w & 1, w >> 1), four 16 x 16 products over the 16
  // pivots each, computed transposed (G X^T) so that register v of lane l
  // is A(row0 + l%16, col0 + l/16 + 4v) and loads and stores run down
  // columns.  Tile values are fetched two tiles ahead of the products (one
  // workgroup per CU: the load latency has no other wave to hide behind).
  // NT = 512 (one workgroup per matrix only): the two 4-wave halves take
  // alternate tiles
  static_assert(NT == 256 || NT == 512, "apply_panel: 4 or 8 wavefronts");
  constexpr int kSub = NT / 256;
  const int lane = threadIdx.x & 63, w = (threadIdx.x >> 6) & 3, sub = threadIdx.x >> 8;
  const int q4 = lane >> 4, r16 = lane & 15;
  const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
  const int nt = (n + kTile - 1) / kTile;
  const int ntt = nt * (nt + 1) / 2;
  // operand rows of the pivots 4c + l/16 (X rows beyond np are zero)
  const double* gq[4];
  const double* xq[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    gq[c] = s.G + s.pslot[4 * c + q4] * ld;
    xq[c] = s.X + (4 * c + q4) * ld;
  }
  auto tile_at = [&](int tau, int& r0, int& c0) {
    const int bi = tile_row(tau);
    const int bj = tau - bi * (bi + 1) / 2;
    r0 = bi * kTile;
    c0 = bj * kTile;
  };
  auto fetch = [&](int r0, int c0, dvec4 (&av)[2][2]) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = min(r0 + wm + 16 * a + r16, n - 1);
          const int j = min(c0 + wn + 16 * b + q4 + 4 * v, n - 1);
          av[a][b][v] = A[i + (size_t)j * n];
        }
  };
  // two tiles ahead: av the current tile, an the next, the one after that
  // fetched at the top of each iteration.  With 8 wavefronts (two per SIMD)
  // one tile ahead: the other wavefront covers the latency, and the
  // registers of a third tile would not fit the halved register budget.
  constexpr bool kDeep = kSub == 1;
  const int step = t.g * kSub;
  int tau = t.rank * kSub + sub;
  dvec4 av[2][2], an[2][2];
  int r0 = 0, c0 = 0, rn = 0, cn = 0;
  if (tau < ntt) {
    tile_at(tau, r0, c0);
    fetch(r0, c0, av);
  }
  if (kDeep && tau + step < ntt) {
    tile_at(tau + step, rn, cn);
    fetch(rn, cn, an);
  }
  while (tau < ntt) {
    const int tn2 = tau + (kDeep ? 2 : 1) * step;
    int r2 = 0, c2 = 0;
    dvec4 a2[2][2];
    if (tn2 < ntt) {
      tile_at(tn2, r2, c2);
      fetch(r2, c2, a2);
    }
    dvec4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = dvec4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double xo[2], go[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) xo[a] = xq[c][r0 + wm + 16 * a + r16];
#pragma unroll
      for (int b = 0; b < 2; ++b) go[b] = gq[c][c0 + wn + 16 * b + r16];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = mma(go[b], xo[a], acc[a][b]);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = r0 + wm + 16 * a + r16;
          const int j = c0 + wn + 16 * b + q4 + 4 * v;
          if (i < n && j < n && i >= j && !s.inP[i] && !s.inP[j]) A[i + (size_t)j * n] = av[a][b][v] - acc[a][b][v];
        }
    tau += step;
    if constexpr (kDeep) {
      r0 = rn;
      c0 = cn;
      rn = r2;
      cn = c2;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          av[a][b] = an[a][b];
          an[a][b] = a2[a][b];
        }
    } else {
      r0 = r2;
      c0 = c2;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) av[a][b] = a2[a][b];
    }
  }
  for (int q = 0; q < np; ++q) {
    const int p = s.pidx[q];
    for (int i = threadIdx.x; i < n; i += NT) {
      if (s.inP[i]) continue;
      const int lo = i < p ? i : p, hi = i < p ? p : i;
      if (owns(t, hi, lo)) A[hi + (size_t)lo * n] = s.X[q * ld + i];
    }
  }
  for (int e = threadIdx.x; e < np * np; e += NT) {
    const int a = e % np, b = e / np;
    const int pa = s.pidx[a], pb = s.pidx[b];
    if (pa >= pb && owns(t, pa, pb)) A[pa + (size_t)pb * n] = -s.Dinv[a + b * kPB];
  }
}

// In-place -inv(A) of the symmetric n x n matrix A held in its lower
// triangle (global, ld = n), by the blocked sweep on a team of workgroups:
// on exit the lower triangle holds -inv(A) (the caller folds the sign into
// its next pass).  Returns the inertia (as sweep_inv), -(k+1) for an exactly
// zero candidate column, n+1 for an exactly zero pivot, or n+2 when a team
// barrier gave up.  Must be entered after a team barrier.
template <int NT = kNT>
__device__ __attribute__((always_inline)) int sweep_neginv_lower(double* A, int n, bool bk, const LargeScratch& s, Team& tm) {
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += NT) {
    s.perm[i] = i;
    s.inP[i] = 0;
    s.swept[i] = 0;
  }
  __syncthreads();
  int kount = 0;
  int k = 0;
  while (k < n) {
    const int kstart = k;
    const int npre = min(kPB, n - k);
    // candidate columns: the lower part of column p (rows i >= p, coalesced
    // in i) and the row segment A(p, i < p) (coalesced in p for a natural run)
    constexpr int kU = 8;  // loads in flight per thread
    const int ld = s.ld;
    for (int base = tid; base < kPB * ld; base += kU * NT) {
      double v[kU];
      bool st[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int e = base + u * NT;
        const int sl = e / ld, i = e - sl * ld;
        const int p = (e < kPB * ld && sl < npre) ? s.perm[kstart + sl] : 0;
        st[u] = e < kPB * ld && (sl >= npre || i >= n || i >= p);
        v[u] = (st[u] && sl < npre && i < n) ? A[i + (size_t)p * n] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (st[u]) s.G[base + u * NT] = v[u];
    }
    for (int base = tid; base < npre * n; base += kU * NT) {
      double v[kU];
      int dst[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int e = base + u * NT;
        const int i = e / npre, sl = e - i * npre;
        const int p = e < npre * n ? s.perm[kstart + sl] : 0;
        dst[u] = (e < npre * n && i < p) ? sl * ld + i : -1;
        v[u] = dst[u] >= 0 ? A[p + (size_t)i * n] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (dst[u] >= 0) s.G[dst[u]] = v[u];
    }
    if (tid < kPB) s.slotpos[tid] = tid;
    __syncthreads();
    int np = block_panel<NT>(n, kstart, npre, bk, s);
    if (np == 0) {
      for (int e = tid; e < kPB * ld; e += NT) s.X[e] = 0.0;
      __syncthreads();
      if (tid < 64) wave_panel(A, n, kstart, bk, s);
      __syncthreads();
      np = s.ctl[0];
    }
    kount += s.ctl[1];
    const int err = s.ctl[2];
    if (err != 0) return err;
    k = s.ctl[3];
    apply_panel<NT>(A, n, np, s, tm);
    __syncthreads();
    if (tid < np) {
      s.inP[s.pidx[tid]] = 0;
      s.swept[s.pidx[tid]] = 1;
    }
    if (!team_sync(tm)) return n + 2;
  }
  return kount;
}

__device__ inline Team make_team(int g, unsigned* ctrs, int inst, int* abort_flag) {
  Team t;
  t.g = g;
  t.rank = (int)blockIdx.x % g;
  t.ctr = ctrs ? ctrs + inst : nullptr;
  t.epoch = 0;
  t.abort = abort_flag;
  return t;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// zero the per-instance barrier counters (a fill kernel, graph-capture safe)
__global__ void k_zero_u32(unsigned* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// one workgroup of NT threads per matrix; NT = 512 puts two wavefronts on
// each SIMD of the CU the panels' LDS confines it to
template <int NT>
__global__ void __launch_bounds__(NT) k_syminv_large(int n, double* __restrict__ a, int* __restrict__ kount, int bk) {
  extern __shared__ double smem[];
  LdsCarve c(smem);
  LargeScratch s = carve_large(c, n);
  Team tm = make_team(1, nullptr, 0, nullptr);
  double* A = a + (size_t)blockIdx.x * n * n;
  const int k = sweep_neginv_lower<NT>(A, n, bk != 0, s, tm);
  // full symmetric inverse out (SYMINL hands back the lower triangle)
  for_owned_lower<NT>(tm, n, [&](int i, int j) { A[i + (size_t)j * n] = -A[i + (size_t)j * n]; });
  __syncthreads();
  for_owned_lower<NT>(tm, n, [&](int i, int j) {
    if (i != j) A[j + (size_t)i * n] = A[i + (size_t)j * n];
  });
  if (threadIdx.x == 0) kount[blockIdx.x] = k;
}

// MDPROP sector matrices UC(k) = [I/8 + D4 W_off]^-1 - 8I, both triangles.
__global__ void __launch_bounds__(kNT) k_mdprop_uc_large(BlockArgs B, GridArgs G, const double* __restrict__ pconst,
                                                         double h, int bk, double* __restrict__ uc,
                                                         int* __restrict__ kount) {
  extern __shared__ double smem[];
  const int ks = 1 + blockIdx.x;
  const int n = B.n;
  LdsCarve c(smem);
  LargeScratch s = carve_large(c, n);
  double* pv = c.take<double>(B.nham);
  Team tm = make_team(1, nullptr, 0, nullptr);
  const double d1 = h * h / 3.0;
  const double d4 = -d1 / 16.0;
  load_pvec(B, G, 2 * ks - 1, pconst, pv);
  __syncthreads();
  double* A = uc + (size_t)blockIdx.x * n * n;
  for_owned_lower(tm, n, [&](int i, int j) {
    A[i + (size_t)j * n] = (i == j) ? 0.125 : d4 * sumlam(B, i, j, pv);
  });
  __syncthreads();
  const int k = sweep_neginv_lower(A, n, bk != 0, s, tm);
  for_owned_lower(tm, n, [&](int i, int j) {
    const size_t e = i + (size_t)j * n;
    A[e] = (i == j) ? -A[e] - 8.0 : -A[e];
  });
  __syncthreads();
  // full storage: the register-resident propagator reads whole columns
  for_owned_lower(tm, n, [&](int i, int j) {
    if (i != j) A[j + (size_t)i * n] = A[i + (size_t)j * n];
  });
  if (threadIdx.x == 0) kount[blockIdx.x] = k;
}

// MDPROP on a team of g workgroups per instance (grid = ninst * g); Y held in
// its lower triangle during the propagation, written in full at the end.
__global__ void __launch_bounds__(kNT) k_mdprop_large(BlockArgs B, GridArgs G, InstArgs I, int nstep, double h,
                                                      const double* __restrict__ uc,
                                                      const int* __restrict__ uc_kount, int bk, int g,
                                                      unsigned* __restrict__ ctrs, int* __restrict__ abort_flag) {
  extern __shared__ double smem[];
  const int inst = blockIdx.x / g;
  const int n = B.n;
  LdsCarve c(smem);
  LargeScratch s = carve_large(c, n);
  double* pv = c.take<double>(B.nham);
  double* y14 = c.take<double>(n);
  double* y23 = c.take<double>(n);
  double* u2 = c.take<double>(n);
  Team tm = make_team(g, ctrs, inst, abort_flag);
  const double ered = I.ered[inst];
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  const double d1 = h * h / 3.0;
  const double d2 = 2.0 * d1;
  const double half = 0.5 * h;
  double* Y = I.y + (size_t)inst * n * n;

  int nsave = 0;
  {
    double ps = 0.0, pbad = 0.0;
    for (int k = threadIdx.x; k < nstep; k += kNT) {
      const int kc = uc_kount[k];
      ps += (double)kc;
      if (kc > n || kc < 0) pbad += 1.0;
    }
    ps = block_sum<kNT>(ps, s.sv);
    pbad = block_sum<kNT>(pbad, s.sv);
    if (pbad > 0.0) {
      if (threadIdx.x == 0 && tm.rank == 0) {
        I.nodes[inst] = 0;
        I.status[inst] = -1;
      }
      return;
    }
    nsave = (int)ps;
  }
  load_pvec(B, G, 0, pconst, pv);
  __syncthreads();
  {
    const double r0 = G.r[0], rsq0 = rsq_of(r0);
    for_owned_lower(tm, n, [&](int i, int j) {
      double w = sumlam(B, i, j, pv);
      if (i == j) w = hammat_diag(B, i, w, r0, rsq0, ered);
      const size_t e = i + (size_t)j * n;
      Y[e] = h * Y[e] + d1 * w;
    });
  }
  int nodes = 0, status = 0;
  bool ok = true;
  for (int ks = 1; ks <= nstep; ++ks) {
    // DIAG(C) terms (mdprop.f:143-164): every workgroup, redundantly
    load_pvec(B, G, 2 * ks - 1, pconst, pv);
    __syncthreads();
    {
      const double rc = G.r[2 * ks - 1], rsqc = rsq_of(rc);
      for (int i = threadIdx.x; i < n; i += kNT) {
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
        if (owns(tm, i, i)) Y[i + (size_t)i * n] = Y[i + (size_t)i * n] + y14c;
      }
    }
    if (!(ok = team_sync(tm))) break;
    int kc = sweep_neginv_lower(Y, n, bk != 0, s, tm);
    if (kc > n || kc < 0) {
      status = ks;
      break;
    }
    nodes += kc;
    // Y <- UC + 2 Y14 - Y23 Y Y23 with Y holding -inv (mdprop.f:181-185)
    {
      const double* U = uc + (size_t)(ks - 1) * n * n;
      for_owned_lower(tm, n, [&](int i, int j) {
        const size_t e = i + (size_t)j * n;
        double uij = U[e];
        if (i == j) uij = uij + u2[i];
        Y[e] = uij + y23[i] * Y[e] * y23[j];
      });
    }
    if (!(ok = team_sync(tm))) break;
    kc = sweep_neginv_lower(Y, n, bk != 0, s, tm);
    if (kc > n || kc < 0) {
      status = ks;
      break;
    }
    nodes += kc;
    // Y <- D2 W(B) - Y23 Y Y23 + Y14 with Y holding -inv (mdprop.f:197-230)
    const double dd2 = (ks == nstep) ? d1 : d2;
    load_pvec(B, G, 2 * ks, pconst, pv);
    __syncthreads();
    {
      const double rb = G.r[2 * ks], rsqb = rsq_of(rb);
      for_owned_lower(tm, n, [&](int i, int j) {
        const size_t e = i + (size_t)j * n;
        double w = sumlam(B, i, j, pv);
        if (i == j) w = hammat_diag(B, i, w, rb, rsqb, ered);
        double v = dd2 * w + y23[i] * Y[e] * y23[j];
        if (i == j) v = v + y14[i];
        Y[e] = v;
      });
    }
    if (!(ok = team_sync(tm))) break;
  }
  if (ok && status == 0) {
    // Y/H, written in full
    const double hi = 1.0 / h;
    for_owned_lower(tm, n, [&](int i, int j) {
      const double v = hi * Y[i + (size_t)j * n];
      Y[i + (size_t)j * n] = v;
      Y[j + (size_t)i * n] = v;
    });
  }
  if (threadIdx.x == 0 && tm.rank == 0) {
    I.nodes[inst] = nodes - nsave;
    I.status[inst] = ok ? status : -2;
  }
}

// LDPROP on a team of g workgroups per instance; Z and U (per-instance
// workspace) held in their lower triangles.
__global__ void __launch_bounds__(kNT) k_ldprop_large(BlockArgs B, GridArgs G, InstArgs I, int nstep, double h,
                                                      double* __restrict__ uwork, int bk, int g,
                                                      unsigned* __restrict__ ctrs, int* __restrict__ abort_flag) {
  extern __shared__ double smem[];
  const int inst = blockIdx.x / g;
  const int n = B.n;
  LdsCarve c(smem);
  LargeScratch s = carve_large(c, n);
  double* pv = c.take<double>(B.nham);
  Team tm = make_team(g, ctrs, inst, abort_flag);
  const double ered = I.ered[inst];
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  const double d1 = h * h / 3.0;
  const double d2 = 2.0 * d1;
  const double d4 = -d1 / 16.0;
  double* Z = I.y + (size_t)inst * n * n;
  double* U = uwork + (size_t)inst * n * n;
  load_pvec(B, G, 0, pconst, pv);
  __syncthreads();
  {
    const double r0 = G.r[0], rsq0 = rsq_of(r0);
    for_owned_lower(tm, n, [&](int i, int j) {
      double w = sumlam(B, i, j, pv);
      if (i == j) w = hammat_diag(B, i, w, r0, rsq0, ered);
      const size_t e = i + (size_t)j * n;
      double v = h * Z[e] + d1 * w;
      if (i == j) v = 1.0 + v;
      Z[e] = v;
    });
  }
  int nodes = 0, status = 0;
  bool ok = true;
  for (int ks = 1; ks <= nstep; ++ks) {
    load_pvec(B, G, 2 * ks - 1, pconst, pv);
    __syncthreads();
    {
      const double rc = G.r[2 * ks - 1], rsqc = rsq_of(rc);
      for_owned_lower(tm, n, [&](int i, int j) {
        double w = sumlam(B, i, j, pv);
        if (i == j) w = hammat_diag(B, i, w, rc, rsqc, ered);
        double v = d4 * w;
        if (i == j) v = 0.125 + v;
        U[i + (size_t)j * n] = v;
      });
    }
    if (!(ok = team_sync(tm))) break;
    const int ncu = sweep_neginv_lower(U, n, bk != 0, s, tm);
    if (ncu > n || ncu < 0) {
      status = ks;
      break;
    }
    int ncz = sweep_neginv_lower(Z, n, bk != 0, s, tm);
    if (ncz > n || ncz < 0) {
      status = ks;
      break;
    }
    nodes += ncz;
    // Z <- U - Z - 6I with both holding -inv
    for_owned_lower(tm, n, [&](int i, int j) {
      const size_t e = i + (size_t)j * n;
      double v = -U[e] + Z[e];
      if (i == j) v = v - 6.0;
      Z[e] = v;
    });
    if (!(ok = team_sync(tm))) break;
    ncz = sweep_neginv_lower(Z, n, bk != 0, s, tm);
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
      for_owned_lower(tm, n, [&](int i, int j) {
        const size_t e = i + (size_t)j * n;
        double w = sumlam(B, i, j, pv);
        if (i == j) w = hammat_diag(B, i, w, rb, rsqb, ered);
        double u = dd2 * w;
        if (i == j) u = u + 2.0;
        Z[e] = u + Z[e];
      });
    }
    if (!(ok = team_sync(tm))) break;
  }
  if (ok && status == 0) {
    const double hi = 1.0 / h;
    for_owned_lower(tm, n, [&](int i, int j) {
      double v = hi * Z[i + (size_t)j * n];
      if (i == j) v = v - hi;
      Z[i + (size_t)j * n] = v;
      Z[j + (size_t)i * n] = v;
    });
  }
  if (threadIdx.x == 0 && tm.rank == 0) {
    I.nodes[inst] = nodes;
    I.status[inst] = ok ? status : -2;
  }
}

// YINIT for N beyond the LDS eigen-solver with every channel locally closed:
// the adiabatic WKB start V diag(sqrt(w)) V^T of yinit.f:76-150 is the
// principal square root of W, computed by the Denman-Beavers iteration
// (Y <- (Y + Z^-1)/2, Z <- (Z + Y^-1)/2 from Y = W/c, Z = I) with the
// blocked sweep (lower triangles).  status = 0, or 1 when W is not positive
// definite (an open channel at the start radius needs BCYOMN,
// yinit.f:124-131).
__global__ void __launch_bounds__(kNT) k_yinit_large(BlockArgs B, GridArgs G, InstArgs I, int k0, int outward,
                                                     double* __restrict__ work) {
  extern __shared__ double smem[];
  const int inst = blockIdx.x;
  const int n = B.n;
  const size_t nn = (size_t)n * n;
  LdsCarve c(smem);
  LargeScratch s = carve_large(c, n);
  double* pv = c.take<double>(B.nham);
  Team tm = make_team(1, nullptr, 0, nullptr);
  const double ered = I.ered[inst];
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  double* Y = I.y + (size_t)inst * nn;
  double* Z = work + (size_t)inst * 3 * nn;
  double* T = Z + nn;
  double* U = T + nn;
  load_pvec(B, G, k0, pconst, pv);
  __syncthreads();
  const double r0 = G.r[k0], rsq0 = rsq_of(r0);
  double tr = 0.0;
  for_owned_lower(tm, n, [&](int i, int j) {
    double w = sumlam(B, i, j, pv);
    if (i == j) {
      w = hammat_diag(B, i, w, r0, rsq0, ered);
      tr += w;
    }
    Y[i + (size_t)j * n] = w;
  });
  tr = block_sum<kNT>(tr, s.sv);
  const double scale = tr / n;
  bool bad = !(scale > 0.0);
  if (!bad) {
    for_owned_lower(tm, n, [&](int i, int j) {
      const size_t e = i + (size_t)j * n;
      Y[e] /= scale;
      Z[e] = (i == j) ? 1.0 : 0.0;
    });
    __syncthreads();
    for (int it = 0; it < 40; ++it) {
      for_owned_lower(tm, n, [&](int i, int j) {
        const size_t e = i + (size_t)j * n;
        T[e] = Y[e];
        U[e] = Z[e];
      });
      __syncthreads();
      const int ky = sweep_neginv_lower(T, n, true, s, tm);
      const int kz = sweep_neginv_lower(U, n, true, s, tm);
      if (ky != 0 || kz != 0) {
        bad = true;
        break;
      }
      double dmax = 0.0, ymax = 0.0;
      for_owned_lower(tm, n, [&](int i, int j) {
        const size_t e = i + (size_t)j * n;
        const double yn = 0.5 * (Y[e] - U[e]);
        const double zn = 0.5 * (Z[e] - T[e]);
        dmax = fmax(dmax, fabs(yn - Y[e]));
        ymax = fmax(ymax, fabs(yn));
        Y[e] = yn;
        Z[e] = zn;
      });
      int dummy = 0;
      block_argmax<kNT>(dmax, dummy, s.sv, s.si);
      int dummy2 = 0;
      block_argmax<kNT>(ymax, dummy2, s.sv, s.si);
      if (it >= 2 && dmax <= 1.0e-15 * ymax) break;
    }
  }
  const double f = (outward ? 1.0 : -1.0) * sqrt(scale > 0.0 ? scale : 0.0);
  __syncthreads();
  for_owned_lower(tm, n, [&](int i, int j) {
    const double v = f * Y[i + (size_t)j * n];
    Y[i + (size_t)j * n] = v;
    Y[j + (size_t)i * n] = v;
  });
  if (threadIdx.x == 0 && I.status) I.status[inst] = bad ? 1 : 0;
}

// ---------------------------------------------------------------------------
// large-N YINIT as separate stages around team inversions (the iteration of
// k_yinit_large, for one or a few instances): one workgroup per instance.
// work: [ninst][2][n*n] the two matrices of an iteration (Y, Z copies),
// then their inverses; zs: [ninst][n*n] Z.
// ---------------------------------------------------------------------------
// Y = W(R0) / scale, scale = trace / n, Z = I; st[inst] = 1 when scale <= 0
__global__ void __launch_bounds__(kNT) k_yl_start(BlockArgs B, GridArgs G, InstArgs I, int k0,
                                                  double* __restrict__ zs, double* __restrict__ scale,
                                                  int* __restrict__ st) {
  __shared__ double sv[kNT / 64];
  __shared__ double pv[64];
  const int inst = blockIdx.x;
  const int n = B.n;
  const size_t nn = (size_t)n * n;
  const double ered = I.ered[inst];
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  double* Y = I.y + (size_t)inst * nn;
  double* Z = zs + (size_t)inst * nn;
  const bool lds_pv = B.nham <= 64;
  if (lds_pv) load_pvec(B, G, k0, pconst, pv);
  __syncthreads();
  const double r0 = G.r[k0], rsq0 = rsq_of(r0);
  double tr = 0.0;
  for (size_t e = threadIdx.x; e < nn; e += kNT) {
    const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
    double w;
    if (lds_pv) {
      w = sumlam(B, i, j, pv);
    } else {
      w = 0.0;
      const double* v = B.vl + tri_index(i, j) * (size_t)B.nham;
      for (int l = 0; l < B.nham; ++l) {
        const double pl = l < B.mxlam ? G.ptab[(size_t)k0 * B.mxlam + l]
                                      : (l < B.mxlam + B.nconst ? pconst[l - B.mxlam] : rsq0);
        w += v[l] * pl;
      }
    }
    if (i == j) {
      w = hammat_diag(B, i, w, r0, rsq0, ered);
      tr += w;
    }
    Y[e] = w;
    Z[e] = i == j ? 1.0 : 0.0;
  }
  tr = block_sum<kNT>(tr, sv);
  const double sc = tr / n;
  __syncthreads();
  if (sc > 0.0)
    for (size_t e = threadIdx.x; e < nn; e += kNT) Y[e] /= sc;
  if (threadIdx.x == 0) {
    scale[inst] = sc;
    st[inst] = sc > 0.0 ? 0 : 1;
  }
}

// the iteration's matrices: work[inst] = (Y, Z)
__global__ void k_yl_copy(int n, int ninst, const double* __restrict__ y, const double* __restrict__ zs,
                          double* __restrict__ work) {
  const size_t nn = (size_t)n * n;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn * ninst) return;
  const int inst = (int)(e / nn);
  const size_t r = e - (size_t)inst * nn;
  work[(size_t)inst * 2 * nn + r] = y[e];
  work[(size_t)inst * 2 * nn + nn + r] = zs[e];
}

// Y <- (Y + Z^-1)/2, Z <- (Z + Y^-1)/2 from the inverses in work; done[inst]
// = 1 once converged (the test of k_yinit_large), st = 1 on a non-positive
// inverse
__global__ void __launch_bounds__(kNT) k_yl_update(int n, int it, double* __restrict__ y, double* __restrict__ zs,
                                                   const double* __restrict__ work, const int* __restrict__ kc,
                                                   int* __restrict__ done, int* __restrict__ st) {
  __shared__ double sv[kNT / 64];
  __shared__ int si[kNT / 64];
  const int inst = blockIdx.x;
  if (done[inst] || st[inst]) return;
  if (kc[2 * inst] != 0 || kc[2 * inst + 1] != 0) {
    if (threadIdx.x == 0) st[inst] = 1;
    return;
  }
  const size_t nn = (size_t)n * n;
  double* Y = y + (size_t)inst * nn;
  double* Z = zs + (size_t)inst * nn;
  const double* Yi = work + (size_t)inst * 2 * nn;
  const double* Zi = Yi + nn;
  // from the lower triangles, mirrored (the iterates stay exactly symmetric)
  double dmax = 0.0, ymax = 0.0;
  for (size_t e = threadIdx.x; e < nn; e += kNT) {
    const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
    if (i < j) continue;
    const double yn = 0.5 * (Y[e] + Zi[e]);
    const double zn = 0.5 * (Z[e] + Yi[e]);
    dmax = fmax(dmax, fabs(yn - Y[e]));
    ymax = fmax(ymax, fabs(yn));
    const size_t et = j + (size_t)i * n;
    Y[e] = yn;
    Y[et] = yn;
    Z[e] = zn;
    Z[et] = zn;
  }
  int d0 = 0;
  block_argmax<kNT>(dmax, d0, sv, si);
  int d1 = 0;
  block_argmax<kNT>(ymax, d1, sv, si);
  if (threadIdx.x == 0 && it >= 2 && dmax <= 1.0e-15 * ymax) done[inst] = 1;
}

// Y <- +-sqrt(scale) Y, symmetrised from the lower triangle
__global__ void k_yl_finish(int n, int ninst, int outward, const double* __restrict__ scale, double* __restrict__ y) {
  const size_t nn = (size_t)n * n;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn * ninst) return;
  const int inst = (int)(e / nn);
  const size_t r = e - (size_t)inst * nn;
  const int j = (int)(r / n), i = (int)(r - (size_t)j * n);
  if (i < j) return;
  const double sc = scale[inst];
  const double f = (outward ? 1.0 : -1.0) * sqrt(sc > 0.0 ? sc : 0.0);
  double* Y = y + (size_t)inst * nn;
  const double v = f * Y[r];
  Y[r] = v;
  Y[j + (size_t)i * n] = v;
}

// DGESV's forward elimination (LU with partial pivoting, the right-hand
// sides carried along; ytokg.f:56) of KM (n x n) and T (n x nop) in the
// work area, on a team of g workgroups per instance: column j (and right-hand
// side c, as column n + c) belongs to rank j % g.  Every rank searches the
// pivot column itself (same data, same choice: the first of the largest
// |values|, as IDAMAX), takes the multipliers with the interchange applied
// virtually, and swaps and updates only its own columns; the lower part of
// the pivot column is left as it was and U's diagonal goes to the SJ slot.
// One team barrier per column.  LDS: 5n doubles (dynamic).
__global__ void __launch_bounds__(kNT) k_match_lu(MatchArgs M, int g, unsigned* __restrict__ ctrs,
                                                  int* __restrict__ abort_flag) {
  __shared__ double sv[kNT / 64];
  __shared__ int si[kNT / 64];
  extern __shared__ double smem[];
  const int inst = blockIdx.x / g;
  Team tm = make_team(g, ctrs, inst, abort_flag);
  const int n = M.n;
  const int nop = M.nopen[inst];
  if (nop <= 0 || M.status[inst] != 0) return;
  double* km = M.work + (size_t)inst * (2 * (size_t)n * n + 4 * (size_t)n);
  double* t = km + (size_t)n * n;
  double* ud = t + (size_t)n * n;
  double* lcol = smem;      // [n] multipliers
  double* urow = smem + n;  // [2 (n + nop)] rows piv and k at the owned columns
  const int ncol = n + nop;
  for (int k = 0; k < n; ++k) {
    double v = -1.0;
    int idx = INT_MAX;
    for (int i = k + threadIdx.x; i < n; i += kNT) {
      const double a = fabs(km[i + (size_t)k * n]);
      if (a > v) {
        v = a;
        idx = i;
      }
    }
    block_argmax<kNT>(v, idx, sv, si);
    if (v == 0.0) {
      if (threadIdx.x == 0 && tm.rank == 0) M.status[inst] = 100 + k;
      break;
    }
    const int piv = idx;
    const double pv = km[piv + (size_t)k * n], akk = km[k + (size_t)k * n];
    if (threadIdx.x == 0 && tm.rank == k % g) ud[k] = pv;
    const double dinv = 1.0 / pv;
    for (int i = k + 1 + threadIdx.x; i < n; i += kNT) lcol[i] = (i == piv ? akk : km[i + (size_t)k * n]) * dinv;
    // the owned columns j > k: their rows k and piv first
    const int j0 = k + 1 + ((tm.rank - (k + 1) % g) + g) % g;  // first owned column > k
    for (int j = j0 + threadIdx.x * g; j < ncol; j += kNT * g) {
      double* col = j < n ? km + (size_t)j * n : t + (size_t)(j - n) * n;
      urow[j] = col[piv];
      urow[n + nop + j] = col[k];
    }
    __syncthreads();
    const int own = j0 < ncol ? (ncol - j0 + g - 1) / g : 0;
    const int rows = n - k - 1;
    for (int e = threadIdx.x; e < own * (rows + 1); e += kNT) {
      const int jc = e / (rows + 1), r = e - jc * (rows + 1);
      const int j = j0 + jc * g;
      double* col = j < n ? km + (size_t)j * n : t + (size_t)(j - n) * n;
      const double ukj = urow[j];
      if (r == rows) {
        col[k] = ukj;
      } else {
        const int i = k + 1 + r;
        const double a = (i == piv) ? urow[n + nop + j] : col[i];
        col[i] = a - lcol[i] * ukj;
      }
    }
    if (!team_sync(tm)) {
      if (threadIdx.x == 0) M.status[inst] = 200;
      break;
    }
  }
}

}  // namespace mcp
