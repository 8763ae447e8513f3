// Batched symmetric inversion with inertia for matrices beyond the
// LDS-resident sizes (N up to kWideMaxN = 448), one workgroup of 512 threads
// per matrix: SYMINV/SYMINL's job (source_code/syminv.f:11-14,
// syminl-alloc.f:17-41: the full inverse of the lower triangle and DSYNEG's
// count of negative pivots) for the batched inversions of the large-N paths.
//
// The method is a blocked symmetric sweep with the Bunch-Kaufman pivot
// choice of dsytf2: each pass takes the next kWP = 32 positions as
// candidates, eliminates the candidate panel G = A(:,P) in natural order (so
// every candidate's Bunch-Kaufman column is known at its pivot time), and
// sweeps the leading run of candidates that pass the test as one block
// pivot: X = G S^-1, A <- A - X G^T off P, A(:,P) <- X, A(P,P) <- -S^-1.
// The matrix crosses HBM once per 32 pivots.
//
// Layout of a pass.  The original panel sits in LDS (32 x 448 doubles) as
// the operand of the update.  Each wave eliminates the candidate block and
// its own 64 rows of the panel in registers, with no barrier between the
// steps; S^-1 comes from a register sweep of the accepted block on one
// wave.  The update runs on the matrix cores over 32 x 32 blocks of
// the lower triangle (wave w a contiguous run of them, numbered down the
// columns): X_j^T = S^-1 G_j^T is formed per column block and kept in
// registers in the A-operand layout, and per 16 x 16 sub-tile eight
// v_mfma_f64_16x16x4f64 over K = 32 give (X_j G_i^T)^T, whose register v in
// lane l is the update of A(i0 + l%16, j0 + l/16 + 4v), so loads and stores
// run down columns; blocks are fetched two ahead.
//
// When the first candidate fails the test, the pass takes one dsytf2 step
// instead (1x1 at the candidate, 1x1 at the row of the column maximum after
// an interchange, or the 2x2 block of both, counted by the DSYNEG rule) and
// applies it as a block pivot of one or two.
#pragma once
#include <hip/hip_runtime.h>

#include "mcp_large.hpp"

namespace mcp {

constexpr int kWP = 32;          // candidate positions per pass
constexpr int kWNT = 512;        // threads per workgroup (one panel row each)
constexpr int kWW = kWNT / 64;   // wavefronts
constexpr int kWideMaxN = 448;   // 32 x 448 doubles of candidates fill 112 KiB of LDS

// Element type of the inversion: double (the FP64 path) or float (the FP32
// factorisation of the mixed-precision path).  The 16x16x4 matrix-core
// products of both types take A operand lane l = (row l%16, k l/16) and B
// operand lane l = (k l/16, column l%16); their accumulators differ:
// register v of lane l is row l/16 + 4v for FP64 and 4(l/16) + v for FP32
// (both measured on the device), which wrow gives.
template <typename T>
struct WideVec;
template <>
struct WideVec<double> {
  typedef double type __attribute__((ext_vector_type(4)));
};
template <>
struct WideVec<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <typename T>
using wvec4 = typename WideVec<T>::type;

__device__ __forceinline__ wvec4<double> wmma(double a, double b, wvec4<double> c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ wvec4<float> wmma(float a, float b, wvec4<float> c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T>
__device__ __forceinline__ int wrow(int q, int v) {
  return sizeof(T) == 8 ? q + 4 * v : 4 * q + v;
}
__device__ __forceinline__ float bcast(float x, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), src));
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ void wave_argmax(float& v, int& idx) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
}

template <typename T>
struct WideScratch {
  int ld;        // panel_ld(n): candidate slot stride
  T* G;     // [kWP][ld] candidate columns (slot t = column of pivot t)
  T* Sinv;  // [kWP][kWP] S^-1 of the accepted block, zero beyond np
  T* dd;    // [kWP] D of the candidate block
  T* cmax;  // [kWW][kWP] per-wave column maxima of the eliminated panel
  int* perm;     // [n] position -> index
  int* ipos;     // [n] index -> position
  int* inP;      // [n] window / pivot number + 1, or 0
  int* swept;    // [n] 1 once pivoted
  int* pidx;     // [kWP] pivot index of slot t
  int* ctl;      // [8] 0 np, 1 inertia of the pass, 2 error, 3 next position, 4 f
  T* sv;    // [kWW]
  int* si;       // [kWW]
};

template <typename T>
__device__ inline WideScratch<T> carve_wide(LdsCarve& c, int n) {
  WideScratch<T> s;
  s.ld = panel_ld(n);
  s.G = c.take<T>((size_t)kWP * s.ld);
  s.Sinv = c.take<T>(kWP * kWP);
  s.dd = c.take<T>(kWP);
  s.cmax = c.take<T>(kWW * kWP);
  s.perm = c.take<int>(n);
  s.ipos = c.take<int>(n);
  s.inP = c.take<int>(n);
  s.swept = c.take<int>(n);
  s.pidx = c.take<int>(kWP);
  s.ctl = c.take<int>(8);
  s.sv = c.take<T>(kWW);
  s.si = c.take<int>(kWW);
  return s;
}

// column p of the lower-stored symmetric A into dst (rows 0..n-1; rows from
// n to ld-1 zero): the lower part is contiguous, the row segment strided
template <typename T>
__device__ __forceinline__ void wide_load_col(const T* A, int n, int ld, int p, T* dst) {
  for (int i = threadIdx.x; i < ld; i += kWNT) dst[i] = i < n ? (i >= p ? A[i + (size_t)p * n] : A[p + (size_t)i * n]) : 0.0;
}

// Candidate panel of the pass at position k.  Thread i loads row i of
// G = A(:, perm[k..k+npre-1]) (all loads in flight together) into LDS.
// Then every wave, independently, eliminates the panel in natural order:
// lanes 0..31 hold the candidate rows (cr, the pivot block S, eliminated
// redundantly in each wave so that no barrier separates the steps), every
// lane holds its own row i = 64 w + lane (orow); at step kk the pivot row is
// broadcast from lane kk of cr.  After step kk, orow[kk] is W(i, kk), the
// Bunch-Kaufman column of candidate kk at its pivot time.  Leaves the
// pivots in dd, the per-wave column maxima in cmax and f (leading nonzero
// pivots) in ctl[4].
template <typename T>
__device__ __forceinline__ void wide_panel(const T* A, int n, int k, int npre, const WideScratch<T>& s) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int ld = s.ld;
  {
    T r[kWP];
#pragma unroll
    for (int c = 0; c < kWP; ++c) {
      T v = 0.0;
      if (tid < n && c < npre) {
        const int p = s.perm[k + c];
        v = tid >= p ? A[tid + (size_t)p * n] : A[p + (size_t)tid * n];
      }
      r[c] = v;
    }
    if (tid < ld) {
#pragma unroll
      for (int c = 0; c < kWP; ++c) s.G[(size_t)c * ld + tid] = r[c];
    }
  }
  if (tid < npre) s.inP[s.perm[k + tid]] = tid + 1;
  __syncthreads();
  const int i = tid;  // this lane's own row
  const int rk = i < n ? (s.swept[i] ? 0 : (s.inP[i] ? s.inP[i] - 1 : kWP)) : 0;
  const int prow = lane < npre ? s.perm[k + lane] : 0;
  T cr[kWP], orow[kWP];
#pragma unroll
  for (int c = 0; c < kWP; ++c) {
    cr[c] = (lane < npre && c < npre) ? s.G[(size_t)c * ld + prow] : 0.0;
    orow[c] = i < ld ? s.G[(size_t)c * ld + i] : 0.0;
  }
  int f = npre;
  T dl = 0.0;  // lane kk: pivot kk
#pragma unroll
  for (int kk = 0; kk < kWP; ++kk) {
    const T d = bcast(cr[kk], kk);
    if (kk < f && d == 0.0) f = kk;
    if (kk < f) {
      if (lane == kk) dl = d;
      const T dinv = 1.0 / d;
      const T lc = lane > kk ? cr[kk] * dinv : 0.0;
      const T lo = orow[kk] * dinv;
#pragma unroll
      for (int c = 0; c < kWP; ++c) {
        if (c > kk) {
          const T u = bcast(cr[c], kk);
          cr[c] -= lc * u;
          orow[c] -= lo * u;
        }
      }
      if (lane > kk) cr[kk] = lc;
    }
  }
  // column maxima of W over this wave's rows still unpivoted at each step
#pragma unroll
  for (int kk = 0; kk < kWP; ++kk) {
    const T cv = wave_max((kk < f && kk < rk) ? fabs(orow[kk]) : 0.0);
    if (lane == 0) s.cmax[w * kWP + kk] = cv;
  }
  if (w == 0) {
    if (lane < kWP) s.dd[lane] = lane < f ? dl : 0.0;
    if (lane == 0) s.ctl[4] = f;
  }
}

// Wave 0: the accepted run (the leading candidates whose pivot passes
// |d_k| >= alpha * colmax_k) and its inertia.
template <typename T>
__device__ __forceinline__ void wide_accept(int kstart, int npre, int f, bool bk, const WideScratch<T>& s) {
  const int lane = threadIdx.x & 63;
  bool pass = false, neg = false;
  if (lane < f) {
    T colmax = 0.0;
#pragma unroll
    for (int w = 0; w < kWW; ++w) colmax = fmax(colmax, s.cmax[w * kWP + lane]);
    const T d = s.dd[lane];
    pass = !bk || (fabs(d) >= kBkAlpha * colmax);
    neg = d < 0.0;
  }
  const int np = __ffsll((unsigned long long)__ballot(!pass)) - 1;  // lanes >= f fail
  const unsigned long long below = np >= 64 ? ~0ull : ((1ull << np) - 1ull);
  const int kc = __popcll((unsigned long long)__ballot(neg) & below);
  if (lane >= np && lane < npre) s.inP[s.perm[kstart + lane]] = 0;
  if (lane < kWP) s.pidx[lane] = lane < np ? s.perm[kstart + lane] : 0;
  if (lane == 0) {
    s.ctl[0] = np;
    s.ctl[1] = kc;
    s.ctl[2] = 0;
    s.ctl[3] = kstart + np;
  }
}

// S^-1 of the accepted np x np block on wave 0 by the symmetric sweep in
// natural order, lane a holding row a of the block in registers and each
// pivot row broadcast from its lane (sweeping every index leaves -S^-1).
// Sinv is zero beyond np.
template <typename T>
__device__ __forceinline__ void wide_sinv(int np, const WideScratch<T>& s) {
  const int lane = threadIdx.x & 63;
  const int pa = lane < np ? s.pidx[lane] : 0;
  T t[kWP];
#pragma unroll
  for (int b = 0; b < kWP; ++b) t[b] = (lane < np && b < np) ? s.G[(size_t)b * s.ld + pa] : 0.0;
#pragma unroll
  for (int kk = 0; kk < kWP; ++kk) {
    if (kk < np) {
      const T dinv = 1.0 / bcast(t[kk], kk);
      const T tk = t[kk];  // this row's entry in the pivot column
#pragma unroll
      for (int b = 0; b < kWP; ++b) {
        if (b != kk) {
          const T u = bcast(t[b], kk) * dinv;  // pivot row, scaled
          t[b] = lane == kk ? u : t[b] - tk * u;
        }
      }
      t[kk] = lane == kk ? -dinv : tk * dinv;
    }
  }
  if (lane < kWP) {
#pragma unroll
    for (int b = 0; b < kWP; ++b) s.Sinv[lane * kWP + b] = (lane < np && b < np) ? -t[b] : 0.0;
  }
}

// Products G(rows of tile ti, :) B for the kWP x kWP matrix B given as
// bsrc[col * kWP + t] = B(t, col), on the matrix cores: block r covers
// columns 16r..16r+15; register v of lane l is row i0 + l/16 + 4v, column
// 16r + l%16.  f(row, col, value) for every element.
template <typename T, class F>
__device__ __forceinline__ void wide_rowtile(int ti, const WideScratch<T>& s, const T* bsrc, F&& f) {
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, r16 = lane & 15;
  const int i0 = 16 * ti;
  T ga[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) ga[c] = s.G[(size_t)(4 * c + q) * s.ld + i0 + r16];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    wvec4<T> acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 8; ++c) acc = wmma(ga[c], bsrc[(16 * r + r16) * kWP + 4 * c + q], acc);
#pragma unroll
    for (int v = 0; v < 4; ++v) f(i0 + wrow<T>(q, v), 16 * r + r16, acc[v]);
  }
}

// argmax of v over the workgroup (ties to the smaller idx)
template <typename T>
__device__ __forceinline__ void wide_argmax(T& v, int& idx, const WideScratch<T>& s) {
  wave_argmax(v, idx);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    s.sv[w] = v;
    s.si[w] = idx;
  }
  __syncthreads();
  v = s.sv[0];
  idx = s.si[0];
  for (int q = 1; q < kWW; ++q) {
    const T ov = s.sv[q];
    const int oi = s.si[q];
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
  __syncthreads();
}

// One dsytf2 step at position k when the first candidate fails the block
// test (all threads).  Slot 0 holds column perm[k]; the step's pivot
// columns end in slots 0 (and 1), ctl as after wide_accept.
template <typename T>
__device__ __attribute__((always_inline)) void wide_bk_step(const T* A, int n, int k, const WideScratch<T>& s) {
  const int tid = threadIdx.x;
  const int ld = s.ld;
  const int ck = s.perm[k];
  T v = -1.0;
  int idx = INT_MAX;
  for (int i = tid; i < n; i += kWNT)
    if (i != ck && !s.swept[i]) {
      const T a = fabs(s.G[i]);
      if (a > v || (a == v && s.ipos[i] < idx)) {
        v = a;
        idx = s.ipos[i];
      }
    }
  wide_argmax<T>(v, idx, s);
  const T colmax = v < 0.0 ? 0.0 : v;
  const T absakk = fabs(s.G[ck]);
  if (fmax(absakk, colmax) == 0.0) {
    if (tid == 0) s.ctl[2] = -(k + 1);
    __syncthreads();
    return;
  }
  int p = ck, pq = -1;
  if (!(absakk >= kBkAlpha * colmax)) {
    const int imax = s.perm[idx];
    wide_load_col<T>(A, n, ld, imax, s.G + ld);
    __syncthreads();
    T rv = -1.0;
    int ri = INT_MAX;
    for (int i = tid; i < n; i += kWNT)
      if (i != imax && !s.swept[i]) {
        const T a = fabs(s.G[ld + i]);
        if (a > rv || (a == rv && s.ipos[i] < ri)) {
          rv = a;
          ri = s.ipos[i];
        }
      }
    wide_argmax<T>(rv, ri, s);
    const T rowmax = rv;
    if (absakk >= kBkAlpha * colmax * (colmax / rowmax)) {
      p = ck;
    } else if (fabs(s.G[ld + imax]) >= kBkAlpha * rowmax) {
      p = imax;
    } else {
      p = ck;
      pq = imax;
    }
  }
  // a 1x1 pivot at imax takes slot 0
  if (p != ck)
    for (int i = tid; i < ld; i += kWNT) s.G[i] = s.G[(size_t)ld + i];
  __syncthreads();
  if (tid == 0) {
    const int ps = pq >= 0 ? pq : p;  // the index that moves to position kk
    const int kk = pq >= 0 ? k + 1 : k;
    if (ps != s.perm[kk]) {
      const int pos = s.ipos[ps], other = s.perm[kk];
      s.perm[kk] = ps;
      s.perm[pos] = other;
      s.ipos[ps] = kk;
      s.ipos[other] = pos;
    }
    for (int e = 0; e < kWP * kWP; ++e) s.Sinv[e] = 0.0;
    int kc = 0, err = 0, np;
    if (pq < 0) {
      const T d = s.G[p];
      if (d == 0.0) {
        err = n + 1;
      } else {
        if (d < 0.0) ++kc;
        s.Sinv[0] = 1.0 / d;
      }
      s.pidx[0] = p;
      np = 1;
    } else {
      const T a = s.G[p], bb = s.G[pq], cc = s.G[(size_t)ld + pq];
      const T tb = fabs(bb);
      if ((cc / tb) * a - tb < 0.0) ++kc;  // DSYNEG's 2x2 rule
      const T det = a * cc - bb * bb;
      s.Sinv[0] = cc / det;
      s.Sinv[1] = -bb / det;
      s.Sinv[kWP] = -bb / det;
      s.Sinv[kWP + 1] = a / det;
      s.pidx[0] = p;
      s.pidx[1] = pq;
      np = 2;
    }
    s.ctl[0] = err ? 0 : np;
    s.ctl[1] = kc;
    s.ctl[2] = err;
    s.ctl[3] = k + np;
  }
  __syncthreads();
  if (s.ctl[2] == 0 && tid < s.ctl[0]) s.inP[s.pidx[tid]] = tid + 1;
  __syncthreads();
}

// The block pivot of the pass on the lower triangle of A (see the header).
template <typename T>
__device__ __forceinline__ void wide_apply(T* A, int n, int np, const WideScratch<T>& s) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int q = lane >> 4, r16 = lane & 15;
  const int ld = s.ld;
  const int nt = (n + 15) >> 4;   // 16-row tiles
  const int nb = (nt + 1) >> 1;   // 32-row blocks
  const int nbb = nb * (nb + 1) / 2;
  const int per = (nbb + kWW - 1) / kWW;
  int blk = w * per;
  const int blk1 = min(nbb, blk + per);
  if (blk < blk1) {
    int bj = 0, base = 0;
    while (base + (nb - bj) <= blk) {
      base += nb - bj;
      ++bj;
    }
    int ci = bj + (blk - base), cj = bj;  // compute cursor (block row, block column)
    int li = ci, lj = cj;                 // load cursor
    auto advance = [&](int& bi_, int& bj_) {
      if (++bi_ >= nb) {
        ++bj_;
        bi_ = bj_;
      }
    };
    // values of block (bi, bj): sub-tile (a, b), register v
    auto fetch = [&](int bi_, int bj_, T (&av)[2][2][4]) {
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int i = min(32 * bi_ + 16 * a + r16, n - 1);
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int j = min(32 * bj_ + 16 * b + wrow<T>(q, v), n - 1);
            av[a][b][v] = A[i + (size_t)j * n];
          }
      }
    };
    constexpr int kPD = 2;  // blocks in flight per wave
    T ring[kPD][2][2][4];
#pragma unroll
    for (int u = 0; u < kPD; ++u)
      if (blk + u < blk1) {
        fetch(li, lj, ring[u]);
        advance(li, lj);
      }
    int cur = -1;
    T xa[2][8];  // X_j of the block's two column tiles, A operands
    for (int t0 = blk; t0 < blk1; t0 += kPD) {
#pragma unroll
      for (int u = 0; u < kPD; ++u) {
        const int t = t0 + u;
        if (t < blk1) {
          if (cj != cur) {
            // X_j^T = S^-1 G_j^T per column tile: block r of rows 16r..; register v
            // holds pivot index 16r + wrow(q, v), chunk 4r+v of the tile products
#pragma unroll
            for (int b = 0; b < 2; ++b) {
              const int j0 = 32 * cj + 16 * b;
#pragma unroll
              for (int r = 0; r < 2; ++r) {
                wvec4<T> acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int c = 0; c < 8; ++c)
                  acc = wmma(s.Sinv[(16 * r + r16) * kWP + 4 * c + q], s.G[(size_t)(4 * c + q) * ld + j0 + r16], acc);
#pragma unroll
                for (int v = 0; v < 4; ++v) xa[b][4 * r + v] = acc[v];
              }
            }
            cur = cj;
          }
          const bool diag = ci == cj;
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            const int i0 = 32 * ci + 16 * a;
            const int i = i0 + r16;
            const bool rowok = i < n && !s.inP[i];
            T gb[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) gb[c] = s.G[(size_t)(16 * (c >> 2) + wrow<T>(q, c & 3)) * ld + i0 + r16];
#pragma unroll
            for (int b = 0; b < 2; ++b) {
              if (diag && b > a) continue;  // strictly above the diagonal
              wvec4<T> acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
              for (int c = 0; c < 8; ++c) acc = wmma(xa[b][c], gb[c], acc);
              const int j0 = 32 * cj + 16 * b;
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const int j = j0 + wrow<T>(q, v);
                if (rowok && j < n && i >= j && !s.inP[j]) A[i + (size_t)j * n] = ring[u][a][b][v] - acc[v];
              }
            }
          }
          advance(ci, cj);
          if (t + kPD < blk1) {
            fetch(li, lj, ring[u]);
            advance(li, lj);
          }
        }
      }
    }
  }
  // the pivot columns, A(i, p_q) = X(i, q) = (G S^-1)(i, q) off P (S^-1 is
  // symmetric, so it serves as its own column-major source), and the block -S^-1
  for (int ti = w; ti < nt; ti += kWW) {
    wide_rowtile<T>(ti, s, s.Sinv, [&](int i, int qq, T x) {
      if (i < n && qq < np && !s.inP[i]) {
        const int p = s.pidx[qq];
        const int hi = i > p ? i : p, lo = i > p ? p : i;
        A[hi + (size_t)lo * n] = x;
      }
    });
  }
  for (int e = threadIdx.x; e < np * np; e += kWNT) {
    const int a = e % np, b = e / np;
    const int pa = s.pidx[a], pb = s.pidx[b];
    if (pa >= pb) A[pa + (size_t)pb * n] = -s.Sinv[a * kWP + b];
  }
}

// -inv(A) into the lower triangle of A (n <= kWideMaxN), by passes of up to
// kWP pivots.  Returns the inertia, -(k+1) for an exactly zero column at
// position k, or n+1 for an exactly zero pivot.
template <typename T>
__device__ __attribute__((always_inline)) int wide_neginv_lower(T* A, int n, bool bk, const WideScratch<T>& s) {
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += kWNT) {
    s.perm[i] = i;
    s.ipos[i] = i;
    s.inP[i] = 0;
    s.swept[i] = 0;
  }
  __syncthreads();
  int kount = 0, k = 0;
  while (k < n) {
    const int npre = min(kWP, n - k);
    wide_panel<T>(A, n, k, npre, s);
    __syncthreads();
    const int f = s.ctl[4];
    if (tid < 64) wide_accept<T>(k, npre, f, bk, s);
    __syncthreads();
    int np = s.ctl[0];
    if (np > 0) {
      if (tid < 64) wide_sinv<T>(np, s);
      __syncthreads();
    } else {
      // the window's marks go; the step marks its own pivots
      if (tid < npre) s.inP[s.perm[k + tid]] = 0;
      __syncthreads();
      if (!bk) return n + 1;
      wide_bk_step<T>(A, n, k, s);
      if (s.ctl[2] != 0) return s.ctl[2];
      np = s.ctl[0];
    }
    kount += s.ctl[1];
    wide_apply<T>(A, n, np, s);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid < np) {
      s.inP[s.pidx[tid]] = 0;
      s.swept[s.pidx[tid]] = 1;
    }
    k = s.ctl[3];
    __syncthreads();
  }
  return kount;
}

// One workgroup per matrix (n x n, column-major, stride n^2 between
// matrices), the lower triangle read; out: the full symmetric inverse and
// kount = number of negative pivots (or the error code, the matrix then
// undefined).  With sel, only the matrices m with sel[m] == selval.
template <typename T>
__global__ void __launch_bounds__(kWNT) k_syminv_wide(int n, T* __restrict__ a, int* __restrict__ kount, int bk,
                                                     const int* __restrict__ sel, int selval) {
  extern __shared__ double smem[];
  if (sel && sel[blockIdx.x] != selval) return;
  LdsCarve c(smem);
  WideScratch<T> s = carve_wide<T>(c, n);
  T* A = a + (size_t)blockIdx.x * n * n;
  const int k = wide_neginv_lower<T>(A, n, bk != 0, s);
  // -(-inv) into both triangles: 16 x 16 tiles of the lower triangle, each
  // negated in place and written transposed above the diagonal through a
  // per-wave LDS tile (the candidate area is free by now)
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int q = lane >> 4, r16 = lane & 15;
    const int nt = (n + 15) >> 4;
    T* tb = s.G + w * 16 * 17;
    for (int tj = 0; tj < nt; ++tj)
      for (int ti = tj + w; ti < nt; ti += kWW) {
        T v4[4];
        const int i = 16 * ti + r16;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int j = 16 * tj + q + 4 * v;
          v4[v] = (i < n && j < n && i >= j) ? -A[i + (size_t)j * n] : 0.0;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int j = 16 * tj + q + 4 * v;
          if (i < n && j < n && i >= j) A[i + (size_t)j * n] = v4[v];
          tb[(q + 4 * v) * 17 + r16] = v4[v];  // tb[col][row]
        }
        wave_sync();
        // transposed: element (row 16 tj + r16, column 16 ti + q + 4v) = tile(q + 4v, r16)
        const int jr = 16 * tj + r16;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ic = 16 * ti + q + 4 * v;
          if (jr < n && ic < n && ic > jr) A[jr + (size_t)ic * n] = tb[r16 * 17 + q + 4 * v];
        }
        wave_sync();
      }
  }
  if (threadIdx.x == 0) kount[blockIdx.x] = k;
}

}  // namespace mcp
