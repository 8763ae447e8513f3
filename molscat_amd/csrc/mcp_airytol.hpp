// AIRY with TOLHI > 0 (AIPROP, source_code/aiprop.f:86-327, the reference's
// default long-range propagator): the steps are chosen as the propagation
// goes, from CORR's estimates of the step just taken (source_code/corr.f:
// 43-60), so the potential is evaluated in the kernel (POTENL's terms, as
// for VIVS) at each step's two Gauss points.
//
// One workgroup per instance runs the whole loop: POTENT (W at RA and RB
// with the energy, the local basis of -(W(RA)+W(RB))/2 by the Jacobi
// solver, ascending; the derivative matrix sqrt(3)(W(RB)-W(RA))/DRNOW in
// that basis, its diagonal HP and XLARGE = its largest off-diagonal
// magnitude, source_code/potent.f:56-116), CORR (CDIAG from the mean |W''|
// implied by the eigenvalue change and slopes, COFF = mean sqrt|eig| x
// XLARGE x |DRNOW|^3 / 12), Z <- sym(T^T Z T) with T = V_prev^T V, the
// Airy sector propagator per channel, Z <- Y4 - Y2 (Z + Y1)^-1 Y2 with the
// inertia into NODES, and the next step DRNOW x min(2, (TOLHI /
// max(CDIAG, COFF))^(1/POWR)), the last one cut at RSTOP (aiprop.f:
// 288-316).  Steps are recorded as (RNOW, DRNOW) pairs; a replaying
// instance follows a record instead, as the energies after the first do
// from the reference's scratch file.  The six N x N matrices sit in LDS
// (N up to about 56 on gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include "mcp_airy.hpp"
#include "mcp_airyprop.hpp"
#include "mcp_vivs.hpp"

namespace mcp {

constexpr int kAiryTolMats = 6;  // Z, A, V, VP, X, TS

constexpr int kAiryTolOk = 0;
constexpr int kAiryTolTooManySteps = -1;  // "AIRY PROPAGATION NOT FINISHED" (aiprop.f:331-333)
constexpr int kAiryTolReplayShort = -2;   // a replay past the record's end

struct AiryTolArgs {
  double rstart, rstop, dr0, powr, tolhi;
  int replay;          // 0: own step control; 1: steps from the record
  int record;          // own control: write the steps to the record (one instance)
  int max_steps;       // record capacity in steps (also the step limit)
  int nrec;            // steps in the record (replay)
  double* steps;       // [max_steps][2] RNOW, DRNOW
  const double* ered;  // [ninst]
  double* z;           // [ninst][n*n] Z(RSTART) in, Z(RSTOP) out (primitive basis)
  int inst0;
  double* work;  // [ninst][kAiryTolMats*n*n] when the matrices are not in LDS
  int lds_mats;
  int* nodes;   // [ninst]
  int* nstep;   // [ninst]
  int* status;  // [ninst]; positive: 1-based step of a SYMINV failure (aiprop.f:211-215)
};

struct AiryTolState {
  double rnow, drnow, drmid, rlast, rnext;
  double cdiag, coff;
  int kstep, last, k, stop;
};

// -(W(RA) + W(RB))/2 into A and sqrt(3)(W(RB) - W(RA))/DRNOW into X, W with
// the energy (POTENT, potent.f:66-101); pa, pb hold P at RA and RB
__device__ __forceinline__ void airytol_potent_mats(const BlockArgs& B, const double* pa, const double* pb,
                                                    double ra, double rb, double drnow, double ered, double* A,
                                                    double* X) {
  const int n = B.n;
  const double rsa = rsq_of(ra), rsb = rsq_of(rb);
  const double fact = kPotentSq3 / drnow;
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    double wa = sumlam(B, i, j, pa), wb = sumlam(B, i, j, pb);
    if (i == j) {
      wa = hammat_diag(B, i, wa, ra, rsa, ered);
      wb = hammat_diag(B, i, wb, rb, rsb, ered);
    }
    A[e] = -0.5 * (wa + wb);
    X[e] = fact * (wb - wa);
  }
}

__global__ void __launch_bounds__(kNT) k_airy_tolhi(BlockArgs B, PotlArgs P, AiryTolArgs A) {
  extern __shared__ double smem[];
  const int n = B.n;
  const size_t nn = (size_t)n * n;
  const int tid = threadIdx.x;
  const int inst = A.inst0 + blockIdx.x;
  LdsCarve c(smem);
  __shared__ AiryTolState st;
  double* eig = c.take<double>(n);
  double* eigold = c.take<double>(n);
  double* hp = c.take<double>(n);
  double* y1 = c.take<double>(n);
  double* y2 = c.take<double>(n);
  double* y4 = c.take<double>(n);
  const int half = (n + 1) / 2 + 1;
  double* rc = c.take<double>(half);
  double* rs = c.take<double>(half);
  int* rp = c.take<int>(half);
  int* rq = c.take<int>(half);
  double* sv = c.take<double>(kNT / 64);
  int* order = c.take<int>(n);
  double* tmp = c.take<double>(n);
  double* pa = c.take<double>(B.nham > 0 ? B.nham : 1);
  double* pb = c.take<double>(B.nham > 0 ? B.nham : 1);
  SweepScratch sw = carve_sweep(c, n);
  double* mats = A.lds_mats ? c.take<double>(kAiryTolMats * nn) : A.work + (size_t)inst * kAiryTolMats * nn;
  double* Z = mats;
  double* M = Z + nn;   // the matrix to diagonalise, then T
  double* V = M + nn;   // this step's basis
  double* VP = V + nn;  // the previous step's basis
  double* X = VP + nn;  // the derivative matrix, then in the local basis
  double* TS = X + nn;  // product scratch
  const double ered = A.ered[inst];
  const bool own = A.replay == 0;
  const bool outward = A.rstop > A.rstart;
  const bool bk = n > 30;
  double* zg = A.z + (size_t)inst * nn;
  for (size_t e = tid; e < nn; e += kNT) Z[e] = zg[e];
  // EIGOLD: the eigenvalues of -W(RSTART) (WAVEIG, source_code/waveig.f:38-46)
  vivs_pvec(B, P, A.rstart, pa);
  __syncthreads();
  vivs_hammat(B, pa, A.rstart, ered, M);
  __syncthreads();
  for (size_t e = tid; e < nn; e += kNT) M[e] = -M[e];
  __syncthreads();
  jacobi_eigh<kNT>(M, V, n, n, rc, rs, rp, rq, sv);
  __syncthreads();
  sort_eig<kNT>(M, V, n, n, eigold, order, tmp);
  __syncthreads();
  if (tid == 0) {
    st.drnow = A.dr0;
    st.drmid = A.dr0 * 0.5;
    st.rlast = A.rstart;
    st.rnow = st.rlast + st.drmid;
    st.rnext = st.rlast + A.dr0;
    st.last = 0;
    st.k = 0;
    st.kstep = 0;
    st.stop = 0;
  }
  int nodes = 0, status = kAiryTolOk;
  __syncthreads();
  while (true) {
    if (tid == 0) {
      if (!own) {
        if (st.k >= A.nrec || st.k >= A.max_steps) {
          st.stop = 1;
        } else {
          st.rnow = A.steps[(size_t)st.k * 2];
          st.drnow = A.steps[(size_t)st.k * 2 + 1];
          ++st.k;
          st.last = st.k == A.nrec;
        }
      }
      if (!st.stop) ++st.kstep;
    }
    __syncthreads();
    if (st.stop) {
      status = kAiryTolReplayShort;
      break;
    }
    const double rnow = st.rnow, drnow = st.drnow, drmid = st.drmid;
    const int kstep = st.kstep;
    const double ra = rnow - 0.5 * drnow / kPotentSq3;
    const double rb = rnow + 0.5 * drnow / kPotentSq3;
    vivs_pvec(B, P, ra, pa);
    vivs_pvec(B, P, rb, pb);
    __syncthreads();
    airytol_potent_mats(B, pa, pb, ra, rb, drnow, ered, M, X);
    __syncthreads();
    jacobi_eigh<kNT>(M, V, n, n, rc, rs, rp, rq, sv);
    __syncthreads();
    sort_eig<kNT>(M, V, n, n, eig, order, tmp);
    __syncthreads();
    // the derivative matrix in the local basis: HP its diagonal, XLARGE its
    // largest off-diagonal magnitude
    lds_gemm(X, false, V, false, TS, n);
    __syncthreads();
    lds_gemm(V, true, TS, false, X, n);
    __syncthreads();
    double xl = 0.0;
    for (size_t e = tid; e < nn; e += kNT) {
      const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
      if (i > j) xl = fmax(xl, fabs(X[e]));
    }
    for (int i = tid; i < n; i += kNT) hp[i] = X[(size_t)i * n + i];
    const double xlarge = block_max<kNT>(xl, sv);
    // CORR (corr.f:43-60)
    double sd = 0.0, sk = 0.0;
    const double fc = 2.0 / (drmid * drmid);
    for (int i = tid; i < n; i += kNT) {
      const double w2p = -fc * (eig[i] - eigold[i] + drmid * hp[i]);
      sd += fabs(w2p);
      sk += sqrt(fabs(eig[i]));
    }
    sd = block_sum<kNT>(sd, sv);
    sk = block_sum<kNT>(sk, sv);
    const double dr3 = fabs(drnow * drnow * drnow);
    const double cdiag = dr3 * (sd / (double)n) / 12.0;
    const double coff = (sk / (double)n) * xlarge * dr3 / 12.0;
    for (int i = tid; i < n; i += kNT) eigold[i] = eig[i];
    // T = V (first step) or V_prev^T V, then Z <- sym(T^T Z T)
    if (kstep == 1) {
      for (size_t e = tid; e < nn; e += kNT) M[e] = V[e];
    } else {
      lds_gemm(VP, true, V, false, M, n);
    }
    __syncthreads();
    lds_gemm(Z, false, M, false, TS, n);
    __syncthreads();
    lds_gemm(M, true, TS, false, Z, n);
    __syncthreads();
    vivs_symmetrise(Z, n);
    for (int i = tid; i < n; i += kNT) {
      double a, b, d;
      airy_sector(drnow, eig[i], hp[i], a, b, d);
      y1[i] = a;
      y2[i] = b;
      y4[i] = d;
    }
    __syncthreads();
    for (int i = tid; i < n; i += kNT) Z[(size_t)i * n + i] += y1[i];
    __syncthreads();
    const int kc = sweep_inv<kNT>(Z, n, n, bk, sw);
    if (kc > n || kc < 0) {
      status = kstep;
      break;
    }
    nodes += outward ? kc : n - kc;
    for (size_t e = tid; e < nn; e += kNT) {
      const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
      double v = -(y2[i] * Z[e] * y2[j]);
      if (i == j) v += y4[i];
      Z[e] = v;
      VP[e] = V[e];
    }
    if (tid == 0) {
      if (A.record && kstep <= A.max_steps) {
        A.steps[(size_t)(kstep - 1) * 2] = rnow;
        A.steps[(size_t)(kstep - 1) * 2 + 1] = drnow;
      }
      if (st.last) {
        st.stop = 1;
      } else if (own) {
        if (kstep >= A.max_steps) {
          st.stop = 2;
        } else {
          // the next step (aiprop.f:288-316)
          const double cmax = fmax(fabs(cdiag), fabs(coff));
          double factor = pow(A.tolhi / cmax, 1.0 / A.powr);
          double dr = st.drnow;
          if (dr < 0.0 && st.rnext + dr > 0.0) {
            const double q = 1.0 + dr / st.rnext;
            factor = factor * (q * q * q);
          }
          if (factor > 2.0) factor = 2.0;
          dr = dr * factor;
          st.rlast = st.rnext;
          st.rnext = st.rnext + dr;
          const bool last = (A.rstop - A.rstart) * (A.rstop - st.rnext) < 0.0;
          if (last) {
            st.rnext = A.rstop;
            dr = st.rnext - st.rlast;
          }
          st.last = last ? 1 : 0;
          const double rnew = st.rlast + 0.5 * dr;
          st.drmid = rnew - st.rnow;
          st.rnow = rnew;
          st.drnow = dr;
        }
      }
    }
    __syncthreads();
    if (st.stop) {
      if (st.stop == 2) status = kAiryTolTooManySteps;
      break;
    }
  }
  __syncthreads();
  if (status == kAiryTolOk || status == kAiryTolTooManySteps) {
    // back to the primitive basis with the last step's basis (aiprop.f:359-366)
    lds_gemm(VP, false, Z, false, TS, n);
    __syncthreads();
    lds_gemm(TS, false, VP, true, Z, n);
    __syncthreads();
    for (size_t e = tid; e < nn; e += kNT) {
      const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
      zg[e] = 0.5 * (Z[(size_t)i + (size_t)j * n] + Z[(size_t)j + (size_t)i * n]);
    }
  }
  if (tid == 0) {
    A.nodes[inst] = nodes;
    A.nstep[inst] = st.kstep;
    A.status[inst] = status;
  }
}

}  // namespace mcp
