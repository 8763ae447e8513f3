// MDPROP for large N solved in parallel over the radial grid.
//
// The recurrence of mdprop.f:118-233 is a chain of maps of the form
//   X -> y4 - y3 (X + y1)^-1 y2,
// two per sector: H1 with y1 = diag(Y14), y2 = y3 = diag(Y23), y4 = UC +
// diag(2 A14) (the first inversion and mdprop.f:181-185), and H2 with y1 = 0,
// y2 = y3 = diag(Y23), y4 = D2 W(B) + diag(Y14) (the second inversion and
// mdprop.f:197-230).  Two such maps compose into one of the same form
// (R after L):
//   Z = (l4 + r1)^-1,
//   Y1 = l1 - l2 Z l3,  Y2 = l2 Z r2,  Y3 = r3 Z l3,  Y4 = r4 - r3 Z r2,
// with Y1, Y4 symmetric and Y3 = Y2^T when the parts are.  The whole
// propagation is therefore reduced pairwise in a tree of log2(2 nstep)
// levels, every level one batch of independent merges (an inversion and five
// matrix products each), and applied to the starting matrix at the end.
//
// Node count.  The sequential recurrence inverts the pivots of the block
// L D L^T of one block-tridiagonal matrix; the tree inverts the pivots of
// another elimination order of the same matrix (each merge pivot l4 + r1 is
// the Schur complement of an interface block).  By the inertia additivity of
// Schur complements both sums of negative-pivot counts equal the matrix's
// number of negative eigenvalues, so sum(KOUNT) - NSAVE is the same integer.
//
// Storage: a segment is (Y1, Y2, Y4), three n x n column-major matrices.
#pragma once
#include <hip/hip_runtime.h>

#include "mcp_mfmaprop.hpp"

namespace mcp {

// matrix of batch item (inst, m): base + inst*si + m*sm
struct SegRef {
  const double* base;
  size_t si, sm;
  __device__ __forceinline__ const double* at(int inst, int m) const { return base + inst * si + m * sm; }
};

struct SegOut {
  double* base;
  size_t si, sm;
  __device__ __forceinline__ double* at(int inst, int m) const { return base + inst * si + m * sm; }
};

// element (i, j) of a symmetric matrix of which only the 64 x 64 tiles on
// and below the diagonal are current (the products with LOW)
__device__ __forceinline__ double sym_tile(const double* a, int n, int i, int j) {
  return (i / 64 >= j / 64) ? a[i + (size_t)j * n] : a[j + (size_t)i * n];
}

// W(i, j) at grid point k (HAMMAT: sumlam plus the diagonal terms)
__device__ __forceinline__ double w_elem(const BlockArgs& B, const GridArgs& G, int k, const double* pconst, int i,
                                         int j, double ered) {
  const double r = G.r[k], rsq = rsq_of(r);
  const double* v = B.vl + tri_index(i, j) * (size_t)B.nham;
  double s = 0.0;
  for (int l = 0; l < B.nham; ++l) {
    const double pl = l < B.mxlam ? G.ptab[(size_t)k * B.mxlam + l]
                                  : (l < B.mxlam + B.nconst ? pconst[l - B.mxlam] : rsq);
    s += v[l] * pl;
  }
  if (i == j) s = hammat_diag(B, i, s, r, rsq, ered);
  return s;
}

// ---------------------------------------------------------------------------
// Batched n x n products on the FP64 matrix cores:
//   Co = alpha op(A) op(B) + beta Ci,
// batch item b = (b / per, b % per) on grid x, one 64 x 64 tile of Co per
// workgroup on grid y
// (four waves of 32 x 32, each 2 x 2 MFMA tiles), k in steps of 32 through
// LDS, the next step's operands fetched into registers during the products.
// The MFMA computes the tile transposed (op(B)^T op(A)^T), so that register v
// of lane l is Co(m0 + l%16, n0 + l/16 + 4v) and stores run down columns.
// ---------------------------------------------------------------------------
constexpr int kGT = 64;   // tile edge
constexpr int kGK = 32;   // k step
constexpr int kGPad = 4;  // LDS row padding (doubles)
constexpr int kGPer = kGT * kGK / 256;  // elements of each operand tile per thread

// one k step of op(X) (rows = the tile's m or n side) into registers:
// element e = tid + 256 s; along the contiguous direction of X
template <bool T>
__device__ __forceinline__ void gemm_fetch(const double* x, int n, int r0, int k0, double (&v)[kGPer]) {
#pragma unroll
  for (int s = 0; s < kGPer; ++s) {
    const int e = (int)threadIdx.x + 256 * s;
    // not transposed: rows contiguous (e % 64 = row); transposed: k contiguous (e % 32 = k)
    const int rr = T ? e / kGK : e % kGT, kk = T ? e % kGK : e / kGT;
    const int gr = r0 + rr, gk = k0 + kk;
    double u = 0.0;
    if (gr < n && gk < n) u = T ? x[gk + (size_t)gr * n] : x[gr + (size_t)gk * n];
    v[s] = u;
  }
}

template <bool T>
__device__ __forceinline__ void gemm_stash(double (*S)[kGT + kGPad], const double (&v)[kGPer]) {
#pragma unroll
  for (int s = 0; s < kGPer; ++s) {
    const int e = (int)threadIdx.x + 256 * s;
    const int rr = T ? e / kGK : e % kGT, kk = T ? e % kGK : e / kGT;
    S[kk][rr] = v[s];
  }
}

// LOW: only the tiles on and below the diagonal of Co (a symmetric result),
// ny = nt (nt + 1) / 2 tiles per product, else nt^2.
//
// One-dimensional grid of kXcd * ceil(nbatch / kXcd) * ny workgroups, mapped
// for the XCDs: workgroups are dispatched round-robin over the eight XCDs
// in order of their index, so workgroup L runs on XCD L % 8; the L / 8-th
// workgroup of an XCD is tile (L / 8) % ny of product XCD + 8 (L / 8 / ny).
// Each product's tiles therefore run on one XCD one after another, and the
// panels of its operands are re-read from that XCD's L2 instead of memory.
constexpr int kXcd = 8;
template <bool TA, bool TB, bool LOW = false>
__global__ void __launch_bounds__(256) k_gemm_seg(int n, int per, int nbatch, int ny, SegRef A, SegRef Bm,
                                                  SegRef Ci, SegOut Co, double alpha, double beta) {
  __shared__ double As[kGK][kGT + kGPad];
  __shared__ double Bs[kGK][kGT + kGPad];
  const int lx = (int)(blockIdx.x % kXcd), kx = (int)(blockIdx.x / kXcd);
  const int pi = kx / ny, tile = kx - pi * ny;
  const int bt = lx + kXcd * pi;
  if (bt >= nbatch) return;
  const int inst = bt / per, m = bt - inst * per;
  const int nt = (n + kGT - 1) / kGT;
  int ti = tile % nt, tj = tile / nt;
  if (LOW) {
    // tile t = ti (ti + 1) / 2 + tj, tj <= ti
    const int t = tile;
    ti = 0;
    while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
    tj = t - ti * (ti + 1) / 2;
  }
  const int m0 = ti * kGT, n0 = tj * kGT;
  const double* a = A.at(inst, m);
  const double* bb = Bm.at(inst, m);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int q = lane >> 4, r16 = lane & 15;
  const int mb = (w & 1) * 32, nb = (w >> 1) * 32;
  dvec4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = dvec4{0.0, 0.0, 0.0, 0.0};
  // op(A) rows are the tile's m, op(B)^T rows its n; B transposed is stored
  // like A not transposed and vice versa
  double va[kGPer], vb[kGPer];
  gemm_fetch<TA>(a, n, m0, 0, va);
  gemm_fetch<!TB>(bb, n, n0, 0, vb);
  for (int k0 = 0; k0 < n; k0 += kGK) {
    __syncthreads();
    gemm_stash<TA>(As, va);
    gemm_stash<!TB>(Bs, vb);
    __syncthreads();
    // the next k step's operands are in flight during this step's products
    if (k0 + kGK < n) {
      gemm_fetch<TA>(a, n, m0, k0 + kGK, va);
      gemm_fetch<!TB>(bb, n, n0, k0 + kGK, vb);
    }
#pragma unroll
    for (int c = 0; c < kGK / 4; ++c) {
      const int kk = 4 * c + q;
      double av[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) av[i] = As[kk][mb + 16 * i + r16];
#pragma unroll
      for (int j = 0; j < 2; ++j) bv[j] = Bs[kk][nb + 16 * j + r16];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mma(bv[j], av[i], acc[i][j]);
    }
  }
  const double* ci = beta != 0.0 ? Ci.at(inst, m) : nullptr;
  double* co = Co.at(inst, m);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gm = m0 + mb + 16 * i + r16, gn = n0 + nb + 16 * j + q + 4 * v;
        if (gm < n && gn < n) {
          const size_t idx = gm + (size_t)gn * n;
          double r = alpha * acc[i][j][v];
          if (ci) r += beta * ci[idx];
          co[idx] = r;
        }
      }
}

// ---------------------------------------------------------------------------
// elementwise stages; one thread per element of the batch
// ---------------------------------------------------------------------------
// c1 = UC(k) + diag(2 A14) of every (instance, sector), into Zb
__global__ void k_tree_c1(int n, int ninst, int nstep, const double* __restrict__ uc, const double* __restrict__ diag,
                          double* __restrict__ zb) {
  const size_t nn = (size_t)n * n;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn * nstep * ninst) return;
  const size_t mat = e / nn, r = e - mat * nn;
  const int inst = (int)(mat / nstep), k = (int)(mat - (size_t)inst * nstep);
  const int j = (int)(r / n), i = (int)(r - (size_t)j * n);
  double v = uc[(size_t)k * nn + r];
  if (i == j) v += diag[((size_t)inst * 3 + 2) * nstep * n + (size_t)k * n + i];
  zb[e] = v;
}

// The leaves without U(C).  c1 = UC + diag(2 A14) = V^-1 + E with V = I/8 +
// D4 W_off(C) (mdprop.f:127-135) and E = diag(2 A14 - 8), so by Woodbury
// c1^-1 = E^-1 - E^-1 (V + E^-1)^-1 E^-1: one inversion of the sector's
// M = V + E^-1 replaces the energy-independent inversion of V and the leaf's
// inversion of c1.  The node count needs neg(c1) - neg(V) (the leaf's KOUNT
// less its share of NSAVE); the inertias of [[V, I], [I, -E]] eliminated in
// either order give neg(c1) - neg(V) = n - pos(E) - neg(M).  Used when every
// |E_ii| is well away from zero (k_tree_guard).

// flag = 1 when some |2 A14 - 8| < 0.5 in the batch
__global__ void k_tree_guard(int n, int ninst, int nstep, const double* __restrict__ diag, int* __restrict__ flag) {
  const size_t per = (size_t)nstep * n;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= per * ninst) return;
  const int inst = (int)(e / per);
  const size_t r = e - (size_t)inst * per;
  const double u2 = diag[(size_t)inst * 3 * per + 2 * per + r];
  if (!(fabs(u2 - 8.0) >= 0.5)) atomicMax(flag, 1);
}

// The leaf builders: one element (i, j) per thread over kLeafChunk sectors
// of one instance, its VL entries in registers (VL is then read once per
// chunk instead of once per sector); grid (ceil(n^2/256), ninst * nchunk).
constexpr int kLeafChunk = 32;
constexpr int kLeafMaxHam = 16;  // NHAM of the tree path

// sum_l VL(l, (i,j)) P_l at grid point kp
__device__ __forceinline__ double leaf_sum(const BlockArgs& B, const GridArgs& G, const double* pconst,
                                           const double (&vl)[kLeafMaxHam], int kp, double rsq) {
  double s = 0.0;
#pragma unroll
  for (int l = 0; l < kLeafMaxHam; ++l) {
    if (l < B.nham) {
      const double pl = l < B.mxlam ? G.ptab[(size_t)kp * B.mxlam + l]
                                    : (l < B.mxlam + B.nconst ? pconst[l - B.mxlam] : rsq);
      s += vl[l] * pl;
    }
  }
  return s;
}

// M = V + E^-1 of every (instance, sector) into zb; epos[inst*nstep + k] += pos(E)
__global__ void __launch_bounds__(256) k_tree_m(BlockArgs B, GridArgs G, InstArgs I, int ninst, int nstep, double h,
                                                const double* __restrict__ diag, double* __restrict__ zb,
                                                int* __restrict__ epos, int full) {
  const int n = B.n;
  const size_t nn = (size_t)n * n;
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nn) return;
  const int nch = (nstep + kLeafChunk - 1) / kLeafChunk;
  const int inst = (int)blockIdx.y / nch, k0 = ((int)blockIdx.y % nch) * kLeafChunk;
  const int k1 = min(nstep, k0 + kLeafChunk);
  const int j = (int)(r / n), i = (int)(r - (size_t)j * n);
  // the one-workgroup inversion reads the lower triangle only
  if (!full && i / 64 < j / 64) return;
  const size_t mat0 = (size_t)inst * nstep;
  if (i == j) {
    const double* u2 = diag + ((size_t)inst * 3 + 2) * nstep * n;
    for (int k = k0; k < k1; ++k) {
      const double u = u2[(size_t)k * n + i];
      if (u - 8.0 > 0.0) atomicAdd(epos + mat0 + k, 1);
      zb[(mat0 + k) * nn + r] = 0.125 + 1.0 / (u - 8.0);
    }
    return;
  }
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  const double d4 = -(h * h / 3.0) / 16.0;
  const double* vp = B.vl + tri_index(i, j) * (size_t)B.nham;
  double vl[kLeafMaxHam];
#pragma unroll
  for (int l = 0; l < kLeafMaxHam; ++l) vl[l] = l < B.nham ? vp[l] : 0.0;
  for (int k = k0; k < k1; ++k) {
    const int kc = 2 * (k + 1) - 1;
    zb[(mat0 + k) * nn + r] = d4 * leaf_sum(B, G, pconst, vl, kc, rsq_of(G.r[kc]));
  }
}

// leaf counts n - pos(E) - neg(M) into the instances' totals
__global__ void k_tree_count_leaf(int n, int nb, int per, const int* __restrict__ kount, const int* __restrict__ epos,
                                  int* __restrict__ acc, int* __restrict__ bad) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int inst = b / per;
  const int kc = kount[b];
  if (kc < 0 || kc > n)
    atomicMax(bad + inst, 1);
  else
    atomicAdd(acc + inst, n - epos[b] - kc);
}

// the leaf segments: Z of the first half-step from zb (fused: zb holds
// (V + E^-1)^-1 and Z = E^-1 - E^-1 zb E^-1; else zb holds Z = c1^-1),
// M = D Z D with D = diag(Y23), Y1 = diag(Y14) - M, Y2 = M,
// Y4 = D2 W(B) + diag(Y14) - M
__global__ void __launch_bounds__(256) k_tree_leaf(BlockArgs B, GridArgs G, InstArgs I, int ninst, int nstep,
                                                   double h, const double* __restrict__ diag,
                                                   const double* __restrict__ zb, int fused, double* __restrict__ seg) {
  const int n = B.n;
  const size_t nn = (size_t)n * n;
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nn) return;
  const int nch = (nstep + kLeafChunk - 1) / kLeafChunk;
  const int inst = (int)blockIdx.y / nch, k0 = ((int)blockIdx.y % nch) * kLeafChunk;
  const int k1 = min(nstep, k0 + kLeafChunk);
  const int j = (int)(r / n), i = (int)(r - (size_t)j * n);
  const size_t mat0 = (size_t)inst * nstep;
  const size_t per = (size_t)nstep * n;
  const double* y14 = diag + (size_t)inst * 3 * per;
  const double* y23 = y14 + per;
  const double* u2 = y23 + per;
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  const double d1 = h * h / 3.0;
  const double ered = I.ered[inst];
  const double* vp = B.vl + tri_index(i, j) * (size_t)B.nham;
  double vl[kLeafMaxHam];
#pragma unroll
  for (int l = 0; l < kLeafMaxHam; ++l) vl[l] = l < B.nham ? vp[l] : 0.0;
  for (int k = k0; k < k1; ++k) {
    const size_t o = (size_t)k * n;
    double z = zb[(mat0 + k) * nn + r];
    if (fused) {
      const double ei = 1.0 / (u2[o + i] - 8.0), ej = 1.0 / (u2[o + j] - 8.0);
      const double zz = -ei * z * ej;
      z = i == j ? zz + ei : zz;
    }
    const double mij = y23[o + i] * z * y23[o + j];
    const double dd2 = (k == nstep - 1) ? d1 : 2.0 * d1;
    const int kb = 2 * (k + 1);
    const double rb = G.r[kb], rsq = rsq_of(rb);
    double w = leaf_sum(B, G, pconst, vl, kb, rsq);
    if (i == j) w = hammat_diag(B, i, w, rb, rsq, ered);
    const double wb = dd2 * w;
    const double yd = i == j ? y14[o + i] : 0.0;
    double* sg = seg + (mat0 + k) * 3 * nn;
    sg[r] = yd - mij;
    sg[nn + r] = mij;
    sg[2 * nn + r] = wb + yd - mij;
  }
}

// merge pivots l4 + r1 of segments (2m, 2m+1) of level kl into Zb[inst*per + m]
__global__ void k_tree_w(int n, int ninst, int kl, int per, const double* __restrict__ seg,
                         double* __restrict__ zb, int full) {
  const size_t nn = (size_t)n * n;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn * per * ninst) return;
  const size_t mat = e / nn, r = e - mat * nn;
  const int inst = (int)(mat / per), m = (int)(mat - (size_t)inst * per);
  const int j = (int)(r / n), i = (int)(r - (size_t)j * n);
  // the one-workgroup inversion reads the lower triangle only
  if (!full && i / 64 < j / 64) return;
  const double* L = seg + ((size_t)inst * kl + 2 * m) * 3 * nn;
  const double* R = L + 3 * nn;
  zb[e] = sym_tile(L + 2 * nn, n, i, j) + sym_tile(R, n, i, j);
}

// the odd last segment of a level passes to the next unchanged
__global__ void k_tree_pass(int n, int ninst, int kl, int kn, const double* __restrict__ in, double* __restrict__ out) {
  const size_t sz = 3 * (size_t)n * n;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= sz * ninst) return;
  const int inst = (int)(e / sz);
  const size_t r = e - (size_t)inst * sz;
  out[((size_t)inst * kn + kn - 1) * sz + r] = in[((size_t)inst * kl + kl - 1) * sz + r];
}

// final pivot A0 + Y1, A0 = h Y(RSTART) + D1 W(RSTART) (mdprop.f:118-126)
__global__ void k_tree_final_w(BlockArgs B, GridArgs G, InstArgs I, int ninst, double h, const double* __restrict__ seg,
                               double* __restrict__ zb) {
  const int n = B.n;
  const size_t nn = (size_t)n * n;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn * ninst) return;
  const int inst = (int)(e / nn);
  const size_t r = e - (size_t)inst * nn;
  const int j = (int)(r / n), i = (int)(r - (size_t)j * n);
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  const double d1 = h * h / 3.0;
  const double a0 = h * I.y[e] + d1 * w_elem(B, G, 0, pconst, i, j, I.ered[inst]);
  zb[e] = a0 + sym_tile(seg + (size_t)inst * 3 * nn, n, i, j);
}

// Y(RSTOP) = A / h
__global__ void k_tree_out(int n, int ninst, double h, const double* __restrict__ a, double* __restrict__ y) {
  const size_t nn = (size_t)n * n;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn * ninst) return;
  const size_t inst = e / nn, r = e - inst * nn;
  const int j = (int)(r / n), i = (int)(r - (size_t)j * n);
  y[e] = sym_tile(a + inst * nn, n, i, j) / h;
}

// the tiles above the diagonal of each instance's matrix from those below
__global__ void k_tree_mirror(int n, int ninst, double* __restrict__ a) {
  const size_t nn = (size_t)n * n;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn * ninst) return;
  const size_t inst = e / nn, r = e - inst * nn;
  const int j = (int)(r / n), i = (int)(r - (size_t)j * n);
  if (i / 64 < j / 64) a[e] = a[inst * nn + j + (size_t)i * n];
}

// negative-pivot counts of a batch of inversions (item b of instance b / per)
// into the instances' totals; a failed inversion marks the instance
__global__ void k_tree_count(int n, int nb, int per, const int* __restrict__ kount, int* __restrict__ acc,
                             int* __restrict__ bad) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int inst = b / per;
  const int kc = kount[b];
  if (kc < 0 || kc > n)
    atomicMax(bad + inst, 1);
  else
    atomicAdd(acc + inst, kc);
}

__global__ void k_tree_zero(int* __restrict__ p, int cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) p[i] = 0;
}

// NODES = sum(KOUNT) - NSAVE and STATUS (uc_kount null: NSAVE already in the
// leaf counts)
__global__ void k_tree_nodes(int ninst, int nstep, const int* __restrict__ uc_kount, const int* __restrict__ acc,
                             const int* __restrict__ bad, int* __restrict__ nodes, int* __restrict__ status) {
  const int inst = blockIdx.x * blockDim.x + threadIdx.x;
  if (inst >= ninst) return;
  int nsave = 0;
  if (uc_kount)
    for (int k = 0; k < nstep; ++k) nsave += uc_kount[k];
  nodes[inst] = acc[inst] - nsave;
  status[inst] = bad[inst] ? 1 : 0;
}

}  // namespace mcp
