// MDPROP with Y resident in the registers of one wavefront (N <= 64).
//
// The LDS-resident kernel (k_mdprop_lds) shares each instance's matrix
// between four waves and pays several workgroup barriers per pivot of its
// two inversions per sector.  For batches that fill the chip (field scans,
// many energies) one wavefront per instance does better: lane i keeps row i
// of Y in VGPRs, a pivot's row is broadcast with v_readlane into SGPRs and
// fed straight into the FMAs, and nothing in the sector loop needs LDS or a
// barrier.  The arithmetic follows the same recurrence as k_mdprop_lds
// (source_code/mdprop.f:64-279) and the same pivot decisions (natural order
// for N <= 30, Bunch-Kaufman as DSYTF2 takes them beyond); results agree with
// it to rounding, not bitwise, because each lane updates its whole row from
// its own values (the LDS kernel rebuilds the upper triangle from the lower
// one at every step).
#pragma once
#include <hip/hip_runtime.h>

#include "mcp_kernels.hpp"

namespace mcp {

constexpr int kRegMaxN = 64;    // one row per lane
constexpr int kRegMaxHam = 4;   // P values held in scalar registers
constexpr int kRegWaves = 4;    // instances (wavefronts) per workgroup
constexpr int kRegChunk = 4;    // row elements per broadcast group (bounds the live SGPRs)

// dynamic LDS of k_mdprop_reg<NP>
// (one 64 x (NP + 1) tile per wave: a row per lane, padded against bank
// conflicts)
constexpr size_t reg_lds_bytes(int np) { return (size_t)kRegWaves * 64 * (np + 1) * sizeof(double); }

__device__ __forceinline__ double rl_d(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double rfl_d(double v) {
  const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
  const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

// A lane's row: NP doubles in registers.  Constant indices (after
// unrolling) are registers; the one runtime index the sweep needs, a
// wave-uniform pivot column, is read through a switch on the scalar index
// (a branch tree on the SALU) rather than by addressing the row, which
// would move it to scratch memory.  Nothing writes the row at a runtime
// index.
#define MCP_REG_CASES(X)                                                                                  \
  X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17)      \
  X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32) X(33)       \
  X(34) X(35) X(36) X(37) X(38) X(39) X(40) X(41) X(42) X(43) X(44) X(45) X(46) X(47) X(48) X(49)       \
  X(50) X(51) X(52) X(53) X(54) X(55) X(56) X(57) X(58) X(59) X(60) X(61) X(62) X(63)

template <int NP>
struct RegRow {
  static_assert(NP % 16 == 0 && NP <= 64, "rows of 16, 32, 48 or 64");
  double v[NP];
  __device__ __forceinline__ double get(int j) const { return v[j]; }
  __device__ __forceinline__ void set(int j, double x) { v[j] = x; }
  __device__ __forceinline__ double pick(int p) const {
    double u = 0.0;
    switch (p) {
#define MCP_PICK(j)                       \
  case j:                                 \
    if constexpr (j < NP) u = v[j < NP ? j : 0]; \
    break;
      MCP_REG_CASES(MCP_PICK)
#undef MCP_PICK
      default:
        break;
    }
    return u;
  }
};

// Largest v over the wave, smallest idx among equal v; the result is made
// wave-uniform.
__device__ __forceinline__ void wave_argmax_u(double& v, int& idx) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
  v = rfl_d(v);
  idx = __builtin_amdgcn_readfirstlane(idx);
}

__device__ __forceinline__ double wave_sum_u(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return rfl_d(v);
}

__device__ __forceinline__ void reg_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rebuild the upper triangle from the lower one (DSYFIL, source_code/
// dsyfil.f), as the reference's propagators keep only the lower triangle
// valid between inversions: every lane writes its row to this wave's LDS
// tile and takes A[j, lane] for j > lane back.  The updates between the
// inversions round (i, j) and (j, i) differently (each lane computes its own
// row); without this the triangles drift apart, and the drift is amplified
// where Y passes a pole.  T holds 64 x (NP + 1) doubles (a row per lane).
template <int NP>
__device__ __forceinline__ void reg_symmetrize(RegRow<NP>& r, int lane, double* T) {
  constexpr int S = NP + 1;
#pragma unroll
  for (int j = 0; j < NP; ++j) T[lane * S + j] = r.get(j);
  reg_wave_sync();
  // every element comes back from the tile: A[j, lane] above the diagonal,
  // the lane's own value on and below it (selecting between the tile and
  // the register instead would make the compiler address the row in
  // scratch memory)
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int at = (j > lane) ? j * S + lane : lane * S + j;
    r.set(j, T[at]);
  }
  reg_wave_sync();
}

// In-register symmetric inverse by the sweep operator (the same operator and
// pivot choices as sweep_inv in mcp_device.hpp).  Lane i < n holds row i in
// r[0..n); rows and columns n..NP-1 are zero and stay zero.
//
// During the sweep the diagonal is held apart (dg, one element per lane; the
// diagonal slot of r is scratch), and the off-diagonal elements carry a row
// scale and a column scale: the true A[i,j] = sig_i * r_i[j] * tau_j, with
// lane j holding tau_j.  Sweeping index p then writes no register of r at a
// runtime index: rows other than p take one FMA per element with the
// broadcast row p (its own column masked out of the broadcast, so column p
// keeps its stored values; the product of the two stored factors is formed
// first so that the stored matrix stays bitwise symmetric), row p and
// column p change only their scales
// (A[p,j] -> A[p,j]/d, A[i,p] -> A[i,p]/d), and the new diagonal -1/d goes
// to dg.  A 2x2 pivot (rare: Bunch-Kaufman takes one only when both
// diagonals are small against their coupling) mixes two rows and two
// columns; it goes through the wave's LDS tile T (64 x (NP + 1) doubles),
// where the true values are written out, updated and read back with the
// scales reset.  The scales are folded back at the end.
//
// Returns the inertia; fail receives n+1 for an exactly zero pivot, or
// -(k+1) when the Bunch-Kaufman candidate column k is exactly zero.
template <int NP>
__device__ __forceinline__ int reg_sweep_inv(RegRow<NP>& r, int n, bool bk, int lane, double* T, int& fail) {
  constexpr int S = NP + 1;
  int kount = 0;
  fail = 0;
  const bool live = lane < n;
  double dg = 0.0;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const double x = r.get(j);
    dg = (j == lane) ? x : dg;
    r.set(j, (j == lane) ? 0.0 : x);
  }
  double sig = 1.0;  // scale of this lane's row
  double tau = 1.0;  // scale of column `lane`
  int permv = lane;  // lane t holds perm[t]: the index at position t
  int mypos = lane;  // the position of this lane's index
  int k = 0;
  while (k < n) {
    int p = k, q = -1;
    if (bk) {
      const int ck = __builtin_amdgcn_readlane(permv, k);
      const double tck = rl_d(tau, ck);
      const double colv = (lane == ck) ? 0.0 : fabs(sig * r.pick(ck) * tck);  // |A[lane, ck]|
      const double absakk = fabs(rl_d(dg, ck));
      double v = (live && mypos > k) ? colv : -1.0;
      int idx = (live && mypos > k) ? mypos : INT_MAX;
      wave_argmax_u(v, idx);
      const double colmax = v < 0.0 ? 0.0 : v;
      const int imaxpos = idx;
      if (fmax(absakk, colmax) == 0.0) {
        fail = -(k + 1);
        return kount;
      }
      int kp = k, kstep = 1;
      if (!(absakk >= kBkAlpha * colmax)) {
        const int ci = __builtin_amdgcn_readlane(permv, imaxpos);
        const double tci = rl_d(tau, ci);
        const double rowv = (lane == ci) ? 0.0 : fabs(sig * r.pick(ci) * tci);  // |A[lane, ci]|
        double rv = (live && mypos >= k && mypos != imaxpos) ? rowv : -1.0;
        int ri = (live && mypos >= k && mypos != imaxpos) ? mypos : INT_MAX;
        wave_argmax_u(rv, ri);
        const double rowmax = rv;
        const double aii = fabs(rl_d(dg, ci));
        if (absakk >= kBkAlpha * colmax * (colmax / rowmax)) {
          kp = k;
        } else if (aii >= kBkAlpha * rowmax) {
          kp = imaxpos;
        } else {
          kp = imaxpos;
          kstep = 2;
        }
      }
      const int kk = k + kstep - 1;
      const int pkk = __builtin_amdgcn_readlane(permv, kk);
      const int pkp = __builtin_amdgcn_readlane(permv, kp);
      if (kp != kk) {
        if (lane == kk) permv = pkp;
        else if (lane == kp) permv = pkk;
        if (lane == pkk) mypos = kp;
        else if (lane == pkp) mypos = kk;
      }
      if (kstep == 1) {
        p = (kp != kk) ? pkp : pkk;
      } else {
        p = ck;
        q = (kp != kk) ? pkp : pkk;
      }
    }
    if (q < 0) {
      const double rp = r.pick(p);
      const double tp = rl_d(tau, p), sp = rl_d(sig, p);
      const double u = (lane == p) ? dg : sig * rp * tp;  // A[lane, p]
      const double d = rl_d(dg, p);
      if (d == 0.0) {
        fail = n + 1;
        return kount;
      }
      if (d < 0.0) ++kount;
      const double dinv = 1.0 / d;
      // rows i != p: r_i <- r_i - (A[i,p]/d) row_p, in stored units.  The
      // row and column scales are equal (both change only at a pivot, by
      // the same factor), so the stored matrix is symmetric; forming the
      // product r_i[p] * r_p[j] first keeps it bitwise symmetric, which
      // the propagation needs: a drift between the triangles is amplified
      // where Y passes a pole.
      const double f = tp * sp * dinv;
      const double rq = (lane == p || !live) ? 0.0 : rp;
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        double sj = rl_d(r.get(j), p);
        sj = (j == p) ? 0.0 : sj;
        r.set(j, fma(-(rq * sj), f, r.get(j)));
      }
      if (lane == p) {
        dg = -dinv;
        sig = sp * dinv;
        tau = tp * dinv;
      } else {
        dg = dg - (u * dinv) * u;
      }
      k += 1;
    } else {
      // the true values of this lane's row to the tile (the diagonal from dg)
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const double tj = rl_d(tau, j);
        const double t = (sig * r.get(j)) * tj;
        T[lane * S + j] = (j == lane) ? dg : t;
      }
      reg_wave_sync();
      const double a = T[p * S + p], b = T[q * S + p], cc = T[q * S + q];
      {
        const double tt = fabs(b);
        const double dd = (cc / tt) * a - tt;
        if (dd < 0.0) ++kount;
      }
      const double det = a * cc - b * b;
      const double d00 = cc / det, d01 = -b / det, d11 = a / det;
      const double u = T[lane * S + p], w = T[lane * S + q];  // A[lane, p], A[lane, q]
      const double x = u * d00 + w * d01;
      const double y = u * d01 + w * d11;
      const bool ip = lane == p, iq = lane == q;
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const double s1 = T[p * S + j], s2 = T[q * S + j];  // A[p, j], A[q, j]
        const double mine = T[lane * S + j];
        const bool jp = j == p, jq = j == q;
        double val;
        if ((ip || iq) && (jp || jq)) {
          val = -((ip && jp) ? d00 : ((iq && jq) ? d11 : d01));
        } else if (ip || iq) {
          // rows p, q: D^-1 [row p; row q]
          val = ip ? (s1 * d00 + s2 * d01) : (s1 * d01 + s2 * d11);
        } else if (jp || jq) {
          val = jp ? x : y;
        } else {
          val = mine - (x * s1 + y * s2);
        }
        r.set(j, val);
      }
      reg_wave_sync();
      // the rows computed here differ from their mirror images by rounding:
      // back to one symmetric matrix from the lower triangle
      reg_symmetrize<NP>(r, lane, T);
      dg = r.get(0);
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        dg = (j == lane) ? r.get(j) : dg;
        r.set(j, (j == lane) ? 0.0 : r.get(j));
      }
      sig = 1.0;
      tau = 1.0;
      k += 2;
    }
  }
  // the sweep leaves -A^-1: fold the scales back and restore the diagonal
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const double tj = rl_d(tau, j);
    double t = -(sig * r.get(j)) * tj;
    t = (j == lane) ? -dg : t;
    r.set(j, t);
  }
  return kount;
}

// SUMLAM for element (i, j) with all kRegMaxHam coefficients read: pv is
// zero beyond NHAM and the block's VL is padded with kRegMaxHam zeros, so
// the sum is SUMLAM's without a branch per coefficient.
__device__ __forceinline__ double reg_sumlam(const double* vl, const BlockArgs& B, int i, int j, const double* pv) {
  const double* v = vl + tri_index(i, j) * (size_t)B.nham;
  double s = 0.0;
#pragma unroll
  for (int l = 0; l < kRegMaxHam; ++l) s += v[l] * pv[l];
  return s;
}

__device__ __forceinline__ void reg_load_pvec(const BlockArgs& B, const GridArgs& G, int k, const double* pconst,
                                              double* pv) {
  const double rsq = rsq_of(G.r[k]);
#pragma unroll
  for (int l = 0; l < kRegMaxHam; ++l) {
    double v = 0.0;
    if (l < B.nham) {
      if (l < B.mxlam)
        v = G.ptab[(size_t)k * B.mxlam + l];
      else if (l < B.mxlam + B.nconst)
        v = pconst[l - B.mxlam];
      else
        v = rsq;
    }
    pv[l] = v;
  }
}

// W(B) without its diagonal for every sector ks (grid index 2 ks), the
// energy-independent half of mdprop.f:204 when the constant terms are
// diagonal: wb[ks-1][i + j n] = SUMLAM(i, j) over the radial and 1/R^2
// terms (the constants' VL is zero off the diagonal), the diagonal left to
// the propagation.  One workgroup per sector.
__global__ void __launch_bounds__(256) k_mdprop_wb(BlockArgs B, GridArgs G, int nstep, double* __restrict__ wb) {
  const int ks = blockIdx.x + 1;
  if (ks > nstep) return;
  const int n = B.n;
  const int k = 2 * ks;
  const double rsq = rsq_of(G.r[k]);
  double* W = wb + (size_t)(ks - 1) * n * n;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e % n, j = e / n;
    const double* v = B.vl + tri_index(i, j) * (size_t)B.nham;
    double s = 0.0;
    for (int l = 0; l < B.nham; ++l) {
      double pl;
      if (l < B.mxlam)
        pl = G.ptab[(size_t)k * B.mxlam + l];
      else if (l < B.mxlam + B.nconst)
        pl = 0.0;
      else
        pl = rsq;
      s += v[l] * pl;
    }
    W[e] = (i == j) ? 0.0 : s;
  }
}

// The reference-potential terms of every (instance, sector, channel)
// (mdprop.f:143-164): W_ii at the sector midpoint C, then with arg =
// (H/2) sqrt|W_ii| the values y14 = a14 - D1 W_ii, y23 = a23 and 2 a14
// (a14, a23 = arg/tan(arg) -+ arg tan(arg), or with tanh where W_ii > 0).
// mid[((inst * nstep + ks - 1) * 3 + t) * n + i], t = 0, 1, 2 for y14, y23,
// 2 a14.  Computed ahead of the propagation because tan and tanh, inlined
// into it, would crowd the matrix out of the register file.
__global__ void __launch_bounds__(256) k_mdprop_mid(BlockArgs B, GridArgs G, InstArgs I, int ninst, int nstep,
                                                    double h, double* __restrict__ mid) {
  const int n = B.n;
  const size_t total = (size_t)ninst * nstep * n;
  const double d1 = h * h / 3.0;
  const double half = 0.5 * h;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % n);
    const size_t rest = e / n;
    const int ks = (int)(rest % nstep) + 1;
    const int inst = (int)(rest / nstep);
    const int k = 2 * ks - 1;
    const double rc = G.r[k], rsqc = rsq_of(rc);
    const double* v = B.vl + tri_index(i, i) * (size_t)B.nham;
    const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
    double s = 0.0;
    for (int l = 0; l < B.nham; ++l) {
      double pl;
      if (l < B.mxlam)
        pl = G.ptab[(size_t)k * B.mxlam + l];
      else if (l < B.mxlam + B.nconst)
        pl = pconst[l - B.mxlam];
      else
        pl = rsqc;
      s += v[l] * pl;
    }
    const double wref = hammat_diag(B, i, s, rc, rsqc, I.ered[inst]);
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
    double* out = mid + rest * 3 * n;
    out[i] = a14 - d1 * wref;
    out[n + i] = a23;
    out[2 * n + i] = 2.0 * a14;
  }
}

// One instance per wavefront, kRegWaves instances per workgroup, each wave
// with a 64 x (NP + 1) LDS tile (reg_symmetrize and the 2x2 pivots).  Same
// arguments and results as k_mdprop_lds, plus wb from k_mdprop_wb and mid
// from k_mdprop_mid.
//
// Both updates between the inversions have one form, Y <- a M - Y23 Y Y23
// plus diagonal terms, with M = U(C) and a = 1 after the first inversion of
// a sector (mdprop.f:181-185) and M = W(B) and a = D2 after the second
// (mdprop.f:197-230), so the sector loop holds one pass over the row.  The
// passes run over all NP registers without branches: indices beyond N (and
// lanes beyond N) read an in-range element and keep zero.
template <int NP>
__global__ void __launch_bounds__(64 * kRegWaves) k_mdprop_reg(BlockArgs B, GridArgs G, InstArgs I, int ninst,
                                                               int nstep, double h, const double* __restrict__ uc,
                                                               const int* __restrict__ uc_kount,
                                                               const double* __restrict__ wb,
                                                               const double* __restrict__ mid, int bk) {
  extern __shared__ double smem[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int inst = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kRegWaves + wave));
  if (inst >= ninst) return;
  double* T = smem + (size_t)wave * 64 * (NP + 1);
  const int n = B.n;
  const bool live = lane < n;
  const int li = live ? lane : 0;
  const double ered = I.ered[inst];
  const double* pconst = I.pconst ? I.pconst + (size_t)inst * B.nconst : nullptr;
  double* yg = I.y + (size_t)inst * n * n;
  const double d1 = h * h / 3.0;
  const double d2 = 2.0 * d1;
  const double half = 0.5 * h;

  // NSAVE = sum of the sector-matrix inertias (mdprop.f:133,277-279)
  int nsave;
  {
    double ps = 0.0, pbad = 0.0;
    for (int k = lane; k < nstep; k += 64) {
      const int kc = uc_kount[k];
      ps += (double)kc;
      if (kc > n || kc < 0) pbad += 1.0;
    }
    ps = wave_sum_u(ps);
    pbad = wave_sum_u(pbad);
    if (pbad > 0.0) {
      if (lane == 0) {
        I.nodes[inst] = 0;
        I.status[inst] = -1;
      }
      return;
    }
    nsave = (int)ps;
  }

  RegRow<NP> r;
  double pv[kRegMaxHam];
  // Y <- H*Y + D1*W(RSTART)   (mdprop.f:89-93)
  reg_load_pvec(B, G, 0, pconst, pv);
  {
    const double r0 = G.r[0], rsq0 = rsq_of(r0);
    const double wd = hammat_diag(B, li, reg_sumlam(B.vl, B, li, li, pv), r0, rsq0, ered);
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const bool ok = live && j < n;
      const int jj = j < n ? j : 0;
      double w = reg_sumlam(B.vl, B, li, jj, pv);
      w = (j == lane) ? wd : w;
      const double v = h * yg[li + (size_t)jj * n] + d1 * w;
      r.set(j, ok ? v : 0.0);
    }
  }
  // midpoint terms of sector ks from k_mdprop_mid: y14, y23, 2*a14
  double y14 = 0.0, y23 = 0.0, u2 = 0.0;
  const double* midi = mid + (size_t)inst * nstep * 3 * n;
  auto midpoint = [&](int ks) {
    const double* m = midi + (size_t)(ks - 1) * 3 * n;
    y14 = m[li];
    y23 = m[n + li];
    u2 = m[2 * n + li];
  };
  midpoint(1);
#pragma unroll
  for (int j = 0; j < NP; ++j) r.set(j, (j == lane && live) ? r.get(j) + y14 : r.get(j));

  int nodes = 0, status = 0;
  for (int hs = 0; hs < 2 * nstep; ++hs) {
    const int ks = hs / 2 + 1;
    const bool first = (hs & 1) == 0;
    int fail;
    reg_symmetrize<NP>(r, lane, T);
    const int kc = reg_sweep_inv<NP>(r, n, bk != 0, lane, T, fail);
    if (fail != 0 || kc > n) {
      status = ks;
      break;
    }
    nodes += kc;
    // first: Y <- U(C) + 2 Y14 - Y23 Y Y23   (mdprop.f:181-185)
    // second: Y <- D2*W(B) - Y23 Y Y23 + Y14, then the next sector's Y14 on
    // the diagonal (mdprop.f:197-230 and 158-164)
    const double y23_here = y23;
    double a = 1.0, dgn = 0.0, extra = 0.0;
    const double* M = first ? uc + (size_t)(ks - 1) * n * n : wb + (size_t)(ks - 1) * n * n;
    if (!first) {
      a = (ks == nstep) ? d1 : d2;
      reg_load_pvec(B, G, 2 * ks, pconst, pv);
      const double rb = G.r[2 * ks], rsqb = rsq_of(rb);
      dgn = a * hammat_diag(B, li, reg_sumlam(B.vl, B, li, li, pv), rb, rsqb, ered);
      extra = y14;
      if (ks < nstep) midpoint(ks + 1);
    }
    const double y14_next = (!first && ks < nstep) ? y14 : 0.0;
    // the lane's row base and N pass through empty asm statements so that
    // the addresses and the per-element tests j < N are formed here (a few
    // scalar operations) instead of being hoisted out of the sector loop and
    // held in (spilled) registers
    int lo = li, nn = n;
    asm volatile("" : "+v"(lo));
    asm volatile("" : "+s"(nn));
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const bool ok = live && j < nn;
      const int jj = j < nn ? j : 0;
      const double y23j = rl_d(y23_here, jj);
      const double m = M[lo + (size_t)jj * nn];
      double am = a * m;
      am = (j == lane) ? (first ? m + u2 : dgn) : am;
      double v = am - y23_here * r.get(j) * y23j;
      v = (j == lane && !first) ? (v + extra) + y14_next : v;
      r.set(j, ok ? v : 0.0);
    }
  }
  // Y/H from the lower triangle, mirrored (mdprop.f:269-275)
  const double hi = 1.0 / h;
  if (live) {
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      if (j < n && j <= lane) {
        const double v = hi * r.get(j);
        yg[lane + (size_t)j * n] = v;
        yg[j + (size_t)lane * n] = v;
      }
    }
  }
  if (lane == 0) {
    I.nodes[inst] = nodes - nsave;
    I.status[inst] = status;
  }
}

}  // namespace mcp
