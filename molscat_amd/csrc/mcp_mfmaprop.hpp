// MDPROP for large N (mdprop.f:118-233) on the FP64 matrix cores, with Y
// resident in the registers of a team of g workgroups (one instance per
// team, kMT threads per workgroup).
//
// Matrix-core conventions (v_mfma_f64_16x16x4f64 on gfx950): the A operand
// of lane l is element (l%16, l/16) of the 16x4 factor, the B operand element
// (l/16, l%16) of the 4x16 factor, and accumulator register v of lane l is
// element (l/16 + 4v, l%16) of the 16x16 product.  In a chain of four MFMAs
// over K = 16 whose chunk c covers K = 4c..4c+3, register c of a product
// whose rows run over K is exactly the B operand of chunk c: one product
// feeds the next without lanes exchanging data.
//
// Ownership: row tiles of 16 rows, tile t held by wave t % kMW of every
// workgroup (at most kMTW tiles per wave); column blocks of 16 columns, block
// B by rank B % g (at most RB blocks per rank).  The accumulator of (t, B)
// holds in lane l, register v, the element Y(16t + l%16, 16B + l/16 + 4v):
// the product's rows run over Y's columns and its columns over Y's rows.
//
// The symmetric inversions are the blocked sweep to -inv of the
// global-memory kernel (natural-order pivot blocks of up to 16 with the
// Bunch-Kaufman test of dsytf2.f applied to every candidate at once, and a
// single Bunch-Kaufman step with interchange when the first candidate
// fails).  For the candidate columns G = A(:,P) (P = positions k..k+15) and
// S = A(P,P):
//   S = L D L^T in the registers of lanes 0..15 of wave 0;
//   W = G L^-T, whose column kk over the rows not pivoted before kk is the
//   column dsytf2 tests at step kk; np = the leading candidates that pass;
//   S^-1 = L^-T D^-1 L^-1 over the first np; X = G S^-1;
//   A <- A - X A(P,:) off P, A(P,:) <- S^-1 A(P,:), A(:,P) <- X,
//   A(P,P) <- -S^-1.
// W, S^-1, X and the update are MFMA chains.  After each block the owners of
// the next 16 candidate columns publish them to a double-buffered exchange
// area in global memory, and the team meets at one barrier.
#pragma once
#include <hip/hip_runtime.h>

#include "mcp_large.hpp"

namespace mcp {

constexpr int kMW = 8;                   // waves per workgroup
constexpr int kMT = kMW * 64;            // threads per workgroup
constexpr int kMTW = 4;                  // row tiles per wave
constexpr int kMMaxN = kMW * kMTW * 16;  // largest N
constexpr int kMSlots = kPB + 1;         // exchange slots: the window and one interchanged column

struct MfmaScratch {
  int ld;        // exchange leading dimension (whole 64-row tiles)
  double* Rw;    // [16*16] L^-1, row-major
  double* dw;    // [32] D and its reciprocal
  double* cmax;  // [kMW][16] column maxima of W per wave
  double* Dinv;  // [16*16] S^-1 of a Bunch-Kaufman step
  double* pv;    // [nham]
  double* y14;   // [n]
  double* y23;   // [n]
  double* u2;    // [n]
  int* perm;     // [n] position -> index
  int* ipos;     // [n] index -> position
  int* sl;       // [kPB] exchange slot of each pivot of a Bunch-Kaufman step
  int* ctl;      // [8]
  double* sv;    // [kMW]
  int* si;       // [kMW]
};

__device__ __forceinline__ MfmaScratch carve_mfma(LdsCarve& c, int n, int nham) {
  MfmaScratch s;
  s.ld = panel_ld(n);
  s.Rw = c.take<double>(kPB * kPB);
  s.dw = c.take<double>(2 * kPB);
  s.cmax = c.take<double>(kMW * kPB);
  s.Dinv = c.take<double>(kPB * kPB);
  s.pv = c.take<double>(nham);
  s.y14 = c.take<double>(n);
  s.y23 = c.take<double>(n);
  s.u2 = c.take<double>(n);
  s.perm = c.take<int>(n);
  s.ipos = c.take<int>(n);
  s.sl = c.take<int>(kPB);
  s.ctl = c.take<int>(8);
  s.sv = c.take<double>(kMW);
  s.si = c.take<int>(kMW);
  return s;
}

// f(i, j, a) -> new a over the owned elements with i, j < n
template <int RB, class F>
__device__ __forceinline__ void mf_for_owned(dvec4 (&T)[kMTW][RB], int n, const Team& tm, F&& f) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int q = lane >> 4, r16 = lane & 15;
#pragma unroll
  for (int tt = 0; tt < kMTW; ++tt) {
    const int i = 16 * (w + kMW * tt) + r16;
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const int jb = 16 * (tm.rank + b * tm.g) + q;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = jb + 4 * v;
        if (i < n && j < n) T[tt][b][v] = f(i, j, T[tt][b][v]);
      }
    }
  }
}

// owners write the columns j with slot(j) = u >= 0 to E + u*ld, rows of
// every tile (zero beyond n)
template <int RB, class F>
__device__ __forceinline__ void mf_publish(const dvec4 (&T)[kMTW][RB], int n, int ld, const Team& tm, double* E,
                                           F&& slot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int q = lane >> 4, r16 = lane & 15;
  const int ntile = (n + 15) >> 4;
#pragma unroll
  for (int b = 0; b < RB; ++b) {
    const int jb = 16 * (tm.rank + b * tm.g) + q;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int j = jb + 4 * v;
      const int u = j < n ? slot(j) : -1;
      if (u >= 0) {
#pragma unroll
        for (int tt = 0; tt < kMTW; ++tt) {
          const int t = w + kMW * tt;
          const int i = 16 * t + r16;
          if (t < ntile) E[(size_t)u * ld + i] = i < n ? T[tt][b][v] : 0.0;
        }
      }
    }
  }
}

__device__ __forceinline__ double shfl_d(double x, int src) { return __shfl(x, src, 64); }

// -inv(A) of the symmetric matrix in the team's accumulators, by the blocked
// sweep.  Returns the inertia, -(k+1) for an exactly zero candidate column,
// n+1 for an exactly zero pivot, n+2 when a team barrier gave up.  E: this
// instance's exchange area, 2 x kMSlots x ld doubles.  Entered after a team
// barrier; left after one.
template <int RB>
__device__ __attribute__((always_inline)) int mf_sweep(dvec4 (&T)[kMTW][RB], int n, bool bk, const MfmaScratch& s,
                                                       Team& tm, double* E) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int q = lane >> 4, r16 = lane & 15;
  const int ld = s.ld;
  const int ntile = (n + 15) >> 4;
  for (int i = tid; i < n; i += kMT) {
    s.perm[i] = i;
    s.ipos[i] = i;
  }
  __syncthreads();
  mf_publish<RB>(T, n, ld, tm, E, [&](int j) { return j < kPB ? j : -1; });
  if (!team_sync(tm)) return n + 2;
  const double* Rw = s.Rw;
  const double* dw = s.dw;
  int kount = 0, k = 0, parity = 0;
  while (k < n) {
    double* Ec = E + (size_t)parity * kMSlots * ld;
    const int npre = min(kPB, n - k);
    // candidate columns at this wave's rows: B operands, chunk c = candidates 4c..4c+3
    double Gb[kMTW][4];
#pragma unroll
    for (int tt = 0; tt < kMTW; ++tt) {
      const int t = w + kMW * tt;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int u = 4 * c + q;
        Gb[tt][c] = (t < ntile && u < npre) ? Ec[(size_t)u * ld + 16 * t + r16] : 0.0;
      }
    }
    // candidate rows A(P,j) at the owned columns: A operands, chunk c =
    // candidates 4c..4c+3 (beyond np they meet zero B operands)
    double Ga[RB][4];
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const int blk = tm.rank + b * tm.g;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int u = 4 * c + q;
        Ga[b][c] = (blk < ntile && u < npre) ? Ec[(size_t)u * ld + 16 * blk + r16] : 0.0;
      }
    }
    // S = L D L^T in wave 0, row a in lane a (dsytf2.f 1x1 steps in natural
    // order); R accumulates L^-1 by the same eliminations
    if (w == 0) {
      double Sr[kPB], Rr[kPB];
      const int prow = r16 < npre ? s.perm[k + r16] : 0;
#pragma unroll
      for (int c = 0; c < kPB; ++c) Sr[c] = (lane < npre && c < npre) ? Ec[(size_t)c * ld + prow] : 0.0;
      // straight-line steps: once a pivot is exactly zero (or the window
      // ends) every later multiplier is zero.  At step kk only S(a, kk+1..)
      // and R(a, ..kk) are live: R's column kk starts as e_kk there.
      bool live = true;
      int f = 0;
      double dk = 0.0, dki = 0.0;
#pragma unroll
      for (int kk = 0; kk < kPB; ++kk) {
        Rr[kk] = lane == kk ? 1.0 : 0.0;
        const double d = bcast(Sr[kk], kk);
        live = live && kk < npre && d != 0.0;
        // 1/d: hardware estimate refined by two Newton steps
        double r = __builtin_amdgcn_rcp(d);
        r = fma(r, fma(-d, r, 1.0), r);
        r = fma(r, fma(-d, r, 1.0), r);
        if (live) f = kk + 1;
        if (live && lane == kk) {
          dk = d;
          dki = r;
        }
        // rows at and above the pivot take a zero multiplier: x - 0*y
        // leaves a finite x unchanged, with no predication
        const double la = (live && lane > kk) ? Sr[kk] * r : 0.0;
#pragma unroll
        for (int c = kk + 1; c < kPB; ++c) Sr[c] -= la * bcast(Sr[c], kk);
#pragma unroll
        for (int c = 0; c <= kk; ++c) Rr[c] -= la * bcast(Rr[c], kk);
      }
      if (lane < kPB) {
#pragma unroll
        for (int c = 0; c < kPB; ++c) s.Rw[lane * kPB + c] = Rr[c];
        s.dw[lane] = dk;
        s.dw[kPB + lane] = dki;
      }
      if (lane == 0) s.ctl[4] = f;
    }
    __syncthreads();
    const int f = s.ctl[4];
    const double dk = lane < kPB ? s.dw[lane] : 0.0;
    // W = G L^-T and its masked column maxima
    {
      double la4[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) la4[c] = Rw[r16 * kPB + 4 * c + q];
      double cm[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int tt = 0; tt < kMTW; ++tt) {
        const int t = w + kMW * tt;
        if (t < ntile) {
          const int i = 16 * t + r16;
          const int rk = (i < n && s.ipos[i] >= k) ? min(s.ipos[i] - k, kPB) : 0;
          if (__ballot(rk > 0) != 0ull) {
            dvec4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int c = 0; c < 4; ++c) acc = mma(la4[c], Gb[tt][c], acc);
#pragma unroll
            for (int v = 0; v < 4; ++v)
              if (q + 4 * v < rk) cm[v] = fmax(cm[v], fabs(acc[v]));
          }
        }
      }
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) cm[v] = fmax(cm[v], shfl_d(cm[v], lane ^ off));
      if (r16 == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) s.cmax[w * kPB + q + 4 * v] = cm[v];
      }
    }
    __syncthreads();
    // the leading candidates that pass the test, and their inertia
    int np, kpan;
    {
      bool pass = false, neg = false;
      if (lane < f) {
        double colmax = 0.0;
#pragma unroll
        for (int wq = 0; wq < kMW; ++wq) colmax = fmax(colmax, s.cmax[wq * kPB + lane]);
        pass = !bk || (fabs(dk) >= kBkAlpha * colmax);
        neg = dk < 0.0;
      }
      np = __ffsll((unsigned long long)__ballot(!pass)) - 1;
      kpan = __popcll((unsigned long long)__ballot(neg) & ((1ull << np) - 1ull));
    }
    dvec4 Sv = {0.0, 0.0, 0.0, 0.0};  // lane l, register v: S^-1(l/16 + 4v, l%16)
    if (np > 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int kk = 4 * c + q;
        double rv = 0.0, rd = 0.0;
        if (kk < np) {
          rv = Rw[kk * kPB + r16];
          rd = rv * dw[kPB + kk];
        }
        Sv = mma(rv, rd, Sv);
      }
    } else {
      // one Bunch-Kaufman step at position k (dsytf2.f): candidate ck, the
      // maximum of its column over the unpivoted rows at imax
      const int ck = s.perm[k];
      double v = -1.0;
      int idx = INT_MAX;
      if (tid < n && tid != ck && s.ipos[tid] >= k) {
        v = fabs(Ec[tid]);
        idx = s.ipos[tid];
      }
      block_argmax<kMT>(v, idx, s.sv, s.si);
      const double colmax = v < 0.0 ? 0.0 : v;
      const double absakk = fabs(Ec[ck]);
      if (fmax(absakk, colmax) == 0.0) return -(k + 1);
      int p = ck, pq = -1;
      if (!(absakk >= kBkAlpha * colmax)) {
        const int imax = s.perm[idx];
        // the interchange candidate's column, published by its owner
        mf_publish<RB>(T, n, ld, tm, Ec, [&](int j) { return j == imax ? kPB : -1; });
        if (!team_sync(tm)) return n + 2;
        const double* e16 = Ec + (size_t)kPB * ld;
        double rv = -1.0;
        int ri = INT_MAX;
        if (tid < n && tid != imax && s.ipos[tid] >= k) {
          rv = fabs(e16[tid]);
          ri = s.ipos[tid];
        }
        block_argmax<kMT>(rv, ri, s.sv, s.si);
        const double rowmax = rv;
        if (absakk >= kBkAlpha * colmax * (colmax / rowmax)) {
          p = ck;
        } else if (fabs(e16[imax]) >= kBkAlpha * rowmax) {
          p = imax;
        } else {
          p = ck;
          pq = imax;
        }
      }
      if (tid == 0) {
        // the pivots go to position k (and k+1)
        const int ps = pq >= 0 ? pq : p;
        const int kk = pq >= 0 ? k + 1 : k;
        if (ps != s.perm[kk]) {
          const int pos = s.ipos[ps], other = s.perm[kk];
          s.perm[kk] = ps;
          s.perm[pos] = other;
          s.ipos[ps] = kk;
          s.ipos[other] = pos;
        }
        const int s0 = p == ck ? 0 : kPB;
        for (int u = 0; u < kPB; ++u) s.sl[u] = -1;
        s.sl[0] = s0;
        for (int e = 0; e < kPB * kPB; ++e) s.Dinv[e] = 0.0;
        int kc = 0;
        if (pq < 0) {
          const double d = Ec[(size_t)s0 * ld + p];
          if (d == 0.0) {
            s.ctl[2] = n + 1;
          } else {
            if (d < 0.0) ++kc;
            s.Dinv[0] = 1.0 / d;
            s.ctl[2] = 0;
          }
          s.ctl[0] = 1;
        } else {
          const double a = Ec[p], bb = Ec[pq], cc = Ec[(size_t)kPB * ld + pq];
          const double tb = fabs(bb);
          if ((cc / tb) * a - tb < 0.0) ++kc;
          const double det = a * cc - bb * bb;
          s.Dinv[0] = cc / det;
          s.Dinv[1] = -bb / det;
          s.Dinv[kPB] = -bb / det;
          s.Dinv[kPB + 1] = a / det;
          s.sl[1] = kPB;
          s.ctl[0] = 2;
          s.ctl[2] = 0;
        }
        s.ctl[1] = kc;
      }
      __syncthreads();
      if (s.ctl[2] != 0) return s.ctl[2];
      np = s.ctl[0];
      kpan = s.ctl[1];
      // operands of the step's pivots, pivot u from exchange slot sl[u]
      int slot4[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        Sv[c] = s.Dinv[(q + 4 * c) * kPB + r16];
        slot4[c] = s.sl[4 * c + q];
      }
#pragma unroll
      for (int tt = 0; tt < kMTW; ++tt) {
        const int t = w + kMW * tt;
#pragma unroll
        for (int c = 0; c < 4; ++c)
          Gb[tt][c] = (t < ntile && slot4[c] >= 0) ? Ec[(size_t)slot4[c] * ld + 16 * t + r16] : 0.0;
      }
#pragma unroll
      for (int b = 0; b < RB; ++b) {
        const int blk = tm.rank + b * tm.g;
#pragma unroll
        for (int c = 0; c < 4; ++c)
          Ga[b][c] = (blk < ntile && slot4[c] >= 0) ? Ec[(size_t)slot4[c] * ld + 16 * blk + r16] : 0.0;
      }
    }
    // X = G S^-1 at the owned columns: register c of Xa[b] in lane l is
    // X(16B + l%16, 4c + l/16), the A operand of chunk c of the update
    dvec4 Xa[RB];
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      dvec4 acc = {0.0, 0.0, 0.0, 0.0};
      if (tm.rank + b * tm.g < ntile) {
#pragma unroll
        for (int c = 0; c < 4; ++c) acc = mma(Sv[c], Ga[b][c], acc);
      }
      Xa[b] = acc;
    }
    // A(i,j) <- A(i,j) - G(i,:) X(j,:)^T off P; on pivot row a the B operand
    // is e_a and the entry starts from 0, giving X(j,a) = (S^-1 A(P,j))_a;
    // then the pivot columns: X(i,b) off P, -S^-1(a,b) on P
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const int blk = tm.rank + b * tm.g;
      if (blk < ntile) {
        int bj[4];
        bool pc[4];
        bool lane_aligned = true, any = false;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int j = 16 * blk + q + 4 * v;
          bj[v] = j < n ? s.ipos[j] - k : -1;
          pc[v] = bj[v] >= 0 && bj[v] < np;
          any |= pc[v];
          lane_aligned &= pc[v] && bj[v] == q + 4 * v;
        }
        const bool has_pc = __ballot(any) != 0ull;
        const bool aligned = __ballot(!lane_aligned) == 0ull;
#pragma unroll
        for (int tt = 0; tt < kMTW; ++tt) {
          const int t = w + kMW * tt;
          if (t < ntile) {
            const int i = 16 * t + r16;
            const int ai = i < n ? s.ipos[i] - k : -1;
            const bool prow = ai >= 0 && ai < np;
            dvec4 acc = T[tt][b];
            double gn[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) gn[c] = prow ? (4 * c + q == ai ? 1.0 : 0.0) : -Gb[tt][c];
            if (prow) {
#pragma unroll
              for (int v = 0; v < 4; ++v)
                if (!pc[v]) acc[v] = 0.0;
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) acc = mma(Xa[b][c], gn[c], acc);
            if (has_pc) {
              // X at this wave's rows (lane l, register v: X(16t + l%16, l/16 + 4v))
              dvec4 xt = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
              for (int c = 0; c < 4; ++c) xt = mma(Sv[c], Gb[tt][c], xt);
              if (aligned) {
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[v] = prow ? -Sv[v] : xt[v];
              } else {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                  const int bv = pc[v] ? bj[v] : 0;
                  const int rr = bv >> 2;
                  const int sx = r16 + 16 * (bv & 3), ss = (prow ? ai : 0) + 16 * (bv & 3);
                  const double x0 = shfl_d(xt[0], sx), x1 = shfl_d(xt[1], sx);
                  const double x2 = shfl_d(xt[2], sx), x3 = shfl_d(xt[3], sx);
                  const double s0 = shfl_d(Sv[0], ss), s1 = shfl_d(Sv[1], ss);
                  const double s2 = shfl_d(Sv[2], ss), s3 = shfl_d(Sv[3], ss);
                  const double xv = rr == 0 ? x0 : (rr == 1 ? x1 : (rr == 2 ? x2 : x3));
                  const double sw = rr == 0 ? s0 : (rr == 1 ? s1 : (rr == 2 ? s2 : s3));
                  if (pc[v]) acc[v] = prow ? -sw : xv;
                }
              }
            }
            T[tt][b] = acc;
          }
        }
      }
    }
    kount += kpan;
    k += np;
    parity ^= 1;
    if (k < n) {
      const int kn = k;
      mf_publish<RB>(T, n, ld, tm, E + (size_t)parity * kMSlots * ld, [&](int j) {
        const int u = s.ipos[j] - kn;
        return (u >= 0 && u < kPB) ? u : -1;
      });
    }
    if (!team_sync(tm)) return n + 2;
  }
  return kount;
}

// Symmetric inversion with inertia on the matrix cores for a few matrices:
// matrix b (n x n, full, column-major at a + b n^2) on a team of g
// workgroups of kMT threads (latency-bound: for one or a handful of
// matrices, where a single workgroup per matrix would leave the device
// idle).  a is overwritten by the full inverse, kount[b] by the number of
// negative pivots (or the sweep's error code, the matrix then left as it
// was).  E: exchange areas, 2 x kMSlots x ld doubles per matrix.
template <int RB>
__global__ void __launch_bounds__(kMT) k_syminv_mfma(int n, double* __restrict__ a, int* __restrict__ kount, int bk,
                                                     int g, unsigned* __restrict__ ctrs, int* __restrict__ abort_flag,
                                                     double* __restrict__ xchg) {
  extern __shared__ double smem[];
  const int inst = blockIdx.x / g;
  LdsCarve c(smem);
  MfmaScratch s = carve_mfma(c, n, 0);
  Team tm = make_team(g, ctrs, inst, abort_flag);
  double* E = xchg + (size_t)inst * 2 * kMSlots * s.ld;
  double* A = a + (size_t)inst * n * n;
  dvec4 T[kMTW][RB];
#pragma unroll
  for (int tt = 0; tt < kMTW; ++tt)
#pragma unroll
    for (int b = 0; b < RB; ++b) T[tt][b] = dvec4{0.0, 0.0, 0.0, 0.0};
  mf_for_owned<RB>(T, n, tm, [&](int i, int j, double) { return A[i + (size_t)j * n]; });
  const int k = mf_sweep<RB>(T, n, bk != 0, s, tm, E);
  // the sweep leaves -inv
  if (k >= 0 && k <= n)
    mf_for_owned<RB>(T, n, tm, [&](int i, int j, double v) {
      A[i + (size_t)j * n] = -v;
      return v;
    });
  if (threadIdx.x == 0 && tm.rank == 0) kount[inst] = k;
}

// DIAG(C) reference-potential terms of every sector for every instance
// (mdprop.f:143-164): y14 = A14 - D1 W_ii(C), y23 = A23, u2 = 2 A14, laid
// out [inst][3][nstep][n].  One thread per (instance, sector, channel).
__global__ void k_mdprop_diag(BlockArgs B, GridArgs G, InstArgs I, int ninst, int nstep, double h,
                              double* __restrict__ out) {
  const int n = B.n;
  const size_t per = (size_t)nstep * n;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= per * ninst) return;
  const int inst = (int)(e / per);
  const size_t r = e - (size_t)inst * per;
  const int ks = (int)(r / n) + 1, ii = (int)(r - (size_t)(ks - 1) * n);
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  const int kc = 2 * ks - 1;
  const double rc = G.r[kc], rsqc = rsq_of(rc);
  const double* v = B.vl + tri_index(ii, ii) * (size_t)B.nham;
  double sw = 0.0;
  for (int l = 0; l < B.nham; ++l) {
    const double pl = l < B.mxlam ? G.ptab[(size_t)kc * B.mxlam + l]
                                  : (l < B.mxlam + B.nconst ? pconst[l - B.mxlam] : rsqc);
    sw += v[l] * pl;
  }
  const double wref = hammat_diag(B, ii, sw, rc, rsqc, I.ered[inst]);
  const double d1 = h * h / 3.0;
  const double arg = 0.5 * h * sqrt(fabs(wref));
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
  double* o = out + (size_t)inst * 3 * per + r;
  o[0] = a14 - d1 * wref;
  o[per] = a23;
  o[2 * per] = 2.0 * a14;
}

// MDPROP with Y in the team's accumulators (grid = ninst * g workgroups of
// kMT threads); uc holds the sector matrices U(C) in full, diag the output
// of k_mdprop_diag.
template <int RB>
__global__ void __launch_bounds__(kMT) k_mdprop_mfma(BlockArgs B, GridArgs G, InstArgs I, int nstep, double h,
                                                     const double* __restrict__ uc,
                                                     const int* __restrict__ uc_kount, int bk, int g,
                                                     unsigned* __restrict__ ctrs, int* __restrict__ abort_flag,
                                                     double* __restrict__ xchg, const double* __restrict__ diag) {
  extern __shared__ double smem[];
  const int inst = blockIdx.x / g;
  const int n = B.n;
  LdsCarve c(smem);
  MfmaScratch s = carve_mfma(c, n, B.nham);
  Team tm = make_team(g, ctrs, inst, abort_flag);
  double* E = xchg + (size_t)inst * 2 * kMSlots * s.ld;
  const double ered = I.ered[inst];
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  const double d1 = h * h / 3.0;
  const double d2 = 2.0 * d1;
  double* Yg = I.y + (size_t)inst * n * n;

  int nsave = 0;
  {
    double ps = 0.0, pbad = 0.0;
    for (int k = threadIdx.x; k < nstep; k += kMT) {
      const int kc = uc_kount[k];
      ps += (double)kc;
      if (kc > n || kc < 0) pbad += 1.0;
    }
    ps = block_sum<kMT>(ps, s.sv);
    pbad = block_sum<kMT>(pbad, s.sv);
    if (pbad > 0.0) {
      if (threadIdx.x == 0 && tm.rank == 0) {
        I.nodes[inst] = 0;
        I.status[inst] = -1;
      }
      return;
    }
    nsave = (int)ps;
  }
  dvec4 T[kMTW][RB];
#pragma unroll
  for (int tt = 0; tt < kMTW; ++tt)
#pragma unroll
    for (int b = 0; b < RB; ++b) T[tt][b] = dvec4{0.0, 0.0, 0.0, 0.0};
  load_pvec(B, G, 0, pconst, s.pv);
  __syncthreads();
  {
    const double r0 = G.r[0], rsq0 = rsq_of(r0);
    mf_for_owned<RB>(T, n, tm, [&](int i, int j, double) {
      double wv = sumlam(B, i, j, s.pv);
      if (i == j) wv = hammat_diag(B, i, wv, r0, rsq0, ered);
      return h * Yg[i + (size_t)j * n] + d1 * wv;
    });
  }
  int nodes = 0, status = 0;
  bool ok = true;
  for (int ks = 1; ks <= nstep && ok && status == 0; ++ks) {
    // DIAG(C) terms of this sector
    __syncthreads();
    {
      const double* dg = diag + (size_t)inst * 3 * nstep * n + (size_t)(ks - 1) * n;
      for (int ii = threadIdx.x; ii < n; ii += kMT) {
        s.y14[ii] = dg[ii];
        s.y23[ii] = dg[(size_t)nstep * n + ii];
        s.u2[ii] = dg[(size_t)2 * nstep * n + ii];
      }
    }
    __syncthreads();
    mf_for_owned<RB>(T, n, tm, [&](int i, int j, double a) { return i == j ? a + s.y14[i] : a; });
    // every read of the exchange area ends before the barrier that closes
    // its panel, so a sweep may publish its first window at once
    for (int half = 0; half < 2; ++half) {
      const int kc = mf_sweep<RB>(T, n, bk != 0, s, tm, E);
      if (kc == n + 2) {
        ok = false;
        break;
      }
      if (kc > n || kc < 0) {
        status = ks;
        break;
      }
      nodes += kc;
      if (half == 0) {
        // Y <- UC + 2 Y14 - Y23 Y Y23 with Y holding -inv (mdprop.f:181-185)
        const double* U = uc + (size_t)(ks - 1) * n * n;
        mf_for_owned<RB>(T, n, tm, [&](int i, int j, double a) {
          double uij = U[i + (size_t)j * n];
          if (i == j) uij = uij + s.u2[i];
          return uij + s.y23[i] * a * s.y23[j];
        });
      } else {
        // Y <- D2 W(B) - Y23 Y Y23 + Y14 with Y holding -inv (mdprop.f:197-230)
        const double dd2 = (ks == nstep) ? d1 : d2;
        __syncthreads();
        load_pvec(B, G, 2 * ks, pconst, s.pv);
        __syncthreads();
        const double rb = G.r[2 * ks], rsqb = rsq_of(rb);
        mf_for_owned<RB>(T, n, tm, [&](int i, int j, double a) {
          double wv = sumlam(B, i, j, s.pv);
          if (i == j) wv = hammat_diag(B, i, wv, rb, rsqb, ered);
          double v = dd2 * wv + s.y23[i] * a * s.y23[j];
          if (i == j) v = v + s.y14[i];
          return v;
        });
      }
    }
  }
  if (ok && status == 0) {
    const double hi = 1.0 / h;
    mf_for_owned<RB>(T, n, tm, [&](int i, int j, double a) {
      Yg[i + (size_t)j * n] = hi * a;
      return a;
    });
  }
  if (threadIdx.x == 0 && tm.rank == 0) {
    I.nodes[inst] = nodes - nsave;
    I.status[inst] = ok ? status : -2;
  }
}

}  // namespace mcp
