// VIVS (IPROP = 4): the R-matrix propagator of source_code/vvprop.f with
// first-order perturbation integrals over each step (the quantities of
// source_code/pert2.f and pert1.f, evaluated as in mcp_stepint.hpp) and the
// interval control of source_code/delrd.f, under the /LDVVCM/ controls
// (VivsOpts; the reference's defaults, source_code/mol.driver.f:277-349, are
// IV, IPERT, ISYM, IALPHA = 6 steps per interval in geometric progression
// with ALPHA from ALPHA1 = 1 to ALPHA2 = 1.5 across the range, IALFP off,
// XSQMAX = 1e4).  IVP / IVPP (or IDIAG) add DERMAT's first / second
// derivatives of W at the step centre (source_code/dermat.f) as the
// perturbations V1 (r - h/2) and V2 (r - h/2)^2 / 2; ISHIFT moves the
// reference potential by h^2/24 times the second derivative's diagonal;
// IALPHA <= 0 takes each step from DELRD and ends an interval when the
// corrections' estimates grow (vvprop.f:480-500); IALFP predicts ALPHA from
// the diagonal corrections' trend over the interval (vvprop.f:630-660).
//
// One workgroup per instance runs the whole propagation: the radii are
// chosen as it goes (the step control reads the perturbation estimates of
// the step just taken), so the potential is evaluated on the device from
// POTENL's inverse-power / exponential terms at each step centre and each
// interval's diagonalisation point.  Per step: W at the step centre turned
// into the interval's local basis (two N^3 products), the reference
// solutions per channel, the step integrals per channel pair accumulated into
// EYE11/EYE12/EYE22, and the estimates of the corrections (block maxima).
// Per interval: the Jacobi eigen-solver on W at RMIDI, the R-matrix turned
// into the new basis, and at the interval's end the new R-matrix from
// (G1' + D1' + R (G2' + D2')) X = G1 + D1 + R (G2 + D2) by LU with partial
// pivoting (DGESV), symmetrised.  The N x N work matrices live in LDS when
// they fit, else in a per-instance global workspace behind the same
// pointers.
//
// Steps are recorded as (RNOW, DRNOW, RMIDI, flags) per step: flags bit 0 =
// the step opens an interval, bit 1 = the last step, bit 2 = the step ends
// an interval.  A replaying instance
// takes its steps from such a record instead of its own control, as the
// reference's energies after the first do from the scratch file
// (vvprop.f:155-160, 350-352, 410, 590).
#pragma once
#include <hip/hip_runtime.h>

#include "mcp_airyprop.hpp"
#include "mcp_eig.hpp"
#include "mcp_kernels.hpp"
#include "mcp_stepint.hpp"

namespace mcp {

constexpr int kVivsIalpha = 6;
constexpr double kVivsAlpha1 = 1.0;
constexpr double kVivsAlpha2 = 1.5;
constexpr double kVivsXsqmax = 1.0e4;
constexpr int kVivsMats = 11;  // N x N work matrices per instance

// status codes of one instance
constexpr int kVivsOk = 0;
constexpr int kVivsStepTooSmall = 1;  // DELRD's "STEP SIZE TOO SMALL" (delrd.f:57-60)
constexpr int kVivsTooManySteps = 2;  // more steps than the record holds
constexpr int kVivsSingular = 3;      // an exactly zero pivot in the interval-end solve
constexpr int kVivsReplayShort = 4;   // the record ends before its last step

// POTENL case 1 (source_code/potenl.f:160-200): radial term lam[t] gets
// a[t] * x^npow[t] (npow != 0) or a[t] * exp(e[t] * x) (npow == 0) at
// x = R * RSCALE; P = EP2RU * the sum, clamped at 1e60 first (HAMMAT).
struct PotlArgs {
  int nterm;
  const int* lam;
  const int* npow;
  const double* a;
  const double* e;
  double rscale, ep2ru;
};

// the /LDVVCM/ controls (source_code/vvprop.f:57-82)
struct VivsOpts {
  int iv, ivp, ivpp, ishift, idiag, ipert, isym, ialpha, ialfp;
  double alpha1, alpha2, xsqmax;
};

__host__ __device__ inline VivsOpts vivs_default_opts() {
  VivsOpts o;
  o.iv = 1;
  o.ivp = 0;
  o.ivpp = 0;
  o.ishift = 0;
  o.idiag = 0;
  o.ipert = 1;
  o.isym = 1;
  o.ialpha = kVivsIalpha;
  o.ialfp = 0;
  o.alpha1 = kVivsAlpha1;
  o.alpha2 = kVivsAlpha2;
  o.xsqmax = kVivsXsqmax;
  return o;
}

struct VivsArgs {
  VivsOpts opt;
  double rstart, rstop, drnow, drmax, tldiag, toff;
  int replay;            // 0: own step control; 1: steps from the record
  int record;            // own control: write the steps to the record (one instance)
  int max_steps;         // record capacity in steps
  int nrec;              // replay: steps the record holds (0 when the recording instance failed)
  double* steps;         // [max_steps][4]
  const double* ered;    // [ninst]
  double* rmat;          // [ninst][n*n] R(RSTART) in, R(RSTOP) out
  double* work;          // [ninst][kVivsMats*n*n] when the matrices are not in LDS
  int lds_mats;
  int inst0;             // instance index of blockIdx.x == 0
  int* nstep;            // [ninst]
  int* status;           // [ninst]
};

struct VivsState {
  double rnow, drnow, drint, diagi, rmid, rmidi, alpha, balpha;
  double coff, cdiag, soff, sdiag, soff1, sdiag1, coffl;
  double xbar, xsbar, ebar, exbar, alfnew;
  int istep, icrmat, last, k, stop, status, piv, endint;
};

__device__ __forceinline__ void vivs_pvec(const BlockArgs& B, const PotlArgs& P, double r, double* pv) {
  for (int l = threadIdx.x; l < B.mxlam; l += blockDim.x) {
    const double x = r * P.rscale;
    double s = 0.0;
    for (int t = 0; t < P.nterm; ++t) {
      if (P.lam[t] != l) continue;
      if (P.npow[t] != 0) {
        const double rr = (P.npow[t] < 0 && fabs(x) < 1.0e-8) ? 1.0e-8 : x;
        s += P.a[t] * pow(rr, (double)P.npow[t]);
      } else {
        s += P.a[t] * exp(P.e[t] * x);
      }
    }
    if (fabs(s) > 1.0e60) s = copysign(1.0e60, s);
    pv[l] = P.ep2ru * s;
  }
}

// POTENL with ICNTRL = ider (source_code/potenl.f:177-185): the radial terms'
// derivatives with respect to x = R * RSCALE, times EP2RU, unclamped, as
// DERMAT takes them (source_code/dermat.f:69-80)
__device__ __forceinline__ void vivs_pvec_der(const BlockArgs& B, const PotlArgs& P, double r, int ider, double* pv) {
  for (int l = threadIdx.x; l < B.mxlam; l += blockDim.x) {
    const double x = r * P.rscale;
    double s = 0.0;
    for (int t = 0; t < P.nterm; ++t) {
      if (P.lam[t] != l) continue;
      const int np = P.npow[t];
      double term;
      if (np != 0) {
        const double rr = (np < 0 && fabs(x) < 1.0e-8) ? 1.0e-8 : x;
        term = pow(rr, (double)np);
        term = ider == 1 ? (double)np * term / rr : (double)(np * (np - 1)) * term / (rr * rr);
      } else {
        const double e = P.e[t];
        term = exp(e * x) * (ider == 1 ? e : e * e);
      }
      s += P.a[t] * term;
    }
    pv[l] = P.ep2ru * s;
  }
}

// DERMAT (source_code/dermat.f:38-126; NCONST = NRSQ = 0): the ider'th
// derivative of W at R into the full column-major M.  The diagonal takes the
// centrifugal term's derivative, -2/R^3 or 6/R^4 (clamped at 1e16) times
// L(L+1) (the reference sets RSQ in its NRSQ > 0 branch only, dermat.f:
// 99-110, and adds RSQ*CENT for NRSQ = 0 with the variable unset; its
// meaning is this derivative), and CONLEN's confinement 2R/CONLEN^4 or
// 2/CONLEN^4.
__device__ __forceinline__ void vivs_dermat(const BlockArgs& B, const double* pv, double r, int ider, double* M) {
  const int n = B.n;
  double rsq = ider == 1 ? -2.0 / (r * r * r) : 6.0 / (r * r * r * r);
  rsq = fmin(rsq, 1.0e16);
  const double conf = ider == 1 ? 2.0 * r * B.conlen4inv : 2.0 * B.conlen4inv;
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    double s = sumlam(B, i, j, pv);
    if (i == j) s = s + rsq * B.cent[i] + conf;
    M[e] = s;
  }
}

// W(R) (HAMMAT, NCONST = NRSQ = 0) into the full column-major M; all threads
__device__ __forceinline__ void vivs_hammat(const BlockArgs& B, const double* pv, double r, double ered, double* M) {
  const int n = B.n;
  const double rsq = rsq_of(r);
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    double s = sumlam(B, i, j, pv);
    if (i == j) s = hammat_diag(B, i, s, r, rsq, ered);
    M[e] = s;
  }
}

__device__ __forceinline__ void vivs_symmetrise(double* M, int n) {
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int j = e / n, i = e - j * n;
    if (i < j) {
      const double s = 0.5 * (M[i + j * n] + M[j + i * n]);
      M[i + j * n] = s;
      M[j + i * n] = s;
    }
  }
}

// DELRD (source_code/delrd.f:1-60); returns false for a step below 1e-6 DRMAX
__device__ __forceinline__ bool vivs_delrd(double& dr, double cdiag, double coff, double tol, double drmax, double e1,
                                           double en, double rnow, double rmax) {
  const double rtol = 0.80 * tol;
  const double dfact = cdiag != 0.0 ? pow(rtol / cdiag, 0.333) : 2.5;
  const double ofact = coff != 0.0 ? pow(rtol / coff, 0.333) : 2.5;
  double factor = fmin(dfact, ofact);
  if (en > 0.0)
    factor = fmin(factor, rnow > 0.10 * rmax ? 2.5 : 1.6);
  else if (e1 > 0.0)
    factor = fmin(factor, 1.20);
  else
    factor = fmin(factor, 1.15);
  dr = dr * factor;
  if (en < 0.0) {
    const double drexp = 4.0 / sqrt(-en);
    if (dr > drexp) dr = drexp;
  }
  if (dr > drmax) dr = drmax;
  return !(dr < 1.0e-06 * drmax);
}

// One channel pair i <= j of a step (the role of PERT2, source_code/
// pert2.f:36-222): the pair's perturbation V0_ij + V1_ij (r - h/2) + V2_ij
// (r - h/2)^2 / 2 (use1 / use2: V1 / V2 present; pert1: PERT1, V2 on the
// diagonal only, source_code/pert1.f:160-187) integrated against the
// products of the reference solutions (mcp_stepint.hpp) into
// V0 (c_i c_j), V1 (c_i s_j; its transposed entry c_j s_i) and V2 (s_i s_j),
// and the integrals of the step-start solutions g = a' s + a c against the
// perturbation accumulated into E11 (g1 g1), E12 (g1 g2) and E22 (g2 g2).
// z = DIF h^2 per channel; a1/a1p regular, a2/a2p irregular step-start
// values.
__device__ __forceinline__ void vivs_pair_integrals(int n, int i, int j, const double* z, double h, double* V0,
                                                    double* V1, double* V2, double* E11, double* E12, double* E22,
                                                    const double* a1, const double* a2, const double* a1p,
                                                    const double* a2p, bool use1, bool use2, bool pert1) {
  const size_t ij = (size_t)i + (size_t)j * n, ji = (size_t)j + (size_t)i * n;
  const double v = V0[ij];
  const StepIntegrals q = step_integrals(z[i], z[j], h);
  double cc = v * q.cc, cs = v * q.cs, sc = v * q.sc, ss = v * q.ss;
  if (use1 || use2) {
    const double v1 = use1 ? V1[ij] : 0.0;
    const double v2 = (use2 && !(pert1 && i != j)) ? V2[ij] : 0.0;
    const StepIntegralsW qw = step_integrals_w(z[i], z[j], h);
    cc += v1 * qw.w1.cc + v2 * qw.w2.cc;
    cs += v1 * qw.w1.cs + v2 * qw.w2.cs;
    sc += v1 * qw.w1.sc + v2 * qw.w2.sc;
    ss += v1 * qw.w1.ss + v2 * qw.w2.ss;
  }
  V0[ij] = cc;
  V0[ji] = cc;
  V1[ij] = cs;
  V1[ji] = sc;
  V2[ij] = ss;
  V2[ji] = ss;
  // int (p_i s_i + u_i c_i) V (q_j s_j + w_j c_j) dr
  auto gvg = [&](double p_i, double u_i, double q_j, double w_j) {
    return p_i * q_j * ss + p_i * w_j * sc + u_i * q_j * cs + u_i * w_j * cc;
  };
  const double e11 = E11[ij] + gvg(a1p[i], a1[i], a1p[j], a1[j]);
  const double e22 = E22[ij] + gvg(a2p[i], a2[i], a2p[j], a2[j]);
  const double e12 = E12[ij] + gvg(a1p[i], a1[i], a2p[j], a2[j]);
  if (i != j) E12[ji] += a1p[j] * a2p[i] * ss + a1p[j] * a2[i] * cs + a1[j] * a2p[i] * sc + a1[j] * a2[i] * cc;
  E11[ij] = e11;
  E11[ji] = e11;
  E22[ij] = e22;
  E22[ji] = e22;
  E12[ij] = e12;
}

// DGESV on column-major n x n A (overwritten by its LU factors) and n
// right-hand sides X (overwritten by the solution): partial pivoting on the
// first largest |A(i,k)| (IDAMAX), multipliers by the pivot's reciprocal,
// then the two triangular solves.  Returns false on an exactly zero pivot.
__device__ bool vivs_gesv(double* A, double* X, int n, VivsState& st) {
  const int tid = threadIdx.x;
  bool ok = true;
  for (int k = 0; k < n; ++k) {
    if (tid < 64) {
      double best = -1.0;
      int bi = k;
      for (int i = k + tid; i < n; i += 64) {
        const double v = fabs(A[i + k * n]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (tid == 0) st.piv = bi;
    }
    __syncthreads();
    const int p = st.piv;
    if (A[p + k * n] == 0.0) ok = false;
    if (p != k) {
      for (int c = tid; c < 2 * n; c += kNT) {
        double* M = c < n ? A : X;
        const int cc = c < n ? c : c - n;
        const double t = M[k + cc * n];
        M[k + cc * n] = M[p + cc * n];
        M[p + cc * n] = t;
      }
    }
    __syncthreads();
    const double akk = A[k + k * n];
    const double rinv = akk != 0.0 ? 1.0 / akk : 0.0;
    for (int i = k + 1 + tid; i < n; i += kNT) A[i + k * n] *= rinv;
    __syncthreads();
    const int m = n - k - 1;
    for (int e = tid; e < m * (m + n); e += kNT) {
      const int c = e / m, i = k + 1 + (e - c * m);
      const double l = A[i + k * n];
      if (c < m) {
        const int j = k + 1 + c;
        A[i + j * n] -= l * A[k + j * n];
      } else {
        const int cc = c - m;
        X[i + cc * n] -= l * X[k + cc * n];
      }
    }
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {
    const double ukk = A[k + k * n];
    for (int c = tid; c < n; c += kNT) X[k + c * n] = X[k + c * n] / ukk;
    __syncthreads();
    for (int e = tid; e < k * n; e += kNT) {
      const int c = e / k, i = e - c * k;
      X[i + c * n] -= A[i + k * n] * X[k + c * n];
    }
    __syncthreads();
  }
  return ok;
}

__global__ void __launch_bounds__(kNT) k_vivs(BlockArgs B, PotlArgs P, VivsArgs A) {
  extern __shared__ double smem[];
  const int n = B.n;
  const size_t nn = (size_t)n * n;
  const int tid = threadIdx.x;
  const int inst = A.inst0 + blockIdx.x;
  LdsCarve c(smem);
  __shared__ VivsState st;
  double* g1 = c.take<double>(n);
  double* g1p = c.take<double>(n);
  double* g2 = c.take<double>(n);
  double* g2p = c.take<double>(n);
  double* dg = c.take<double>(n);
  double* a1 = c.take<double>(n);
  double* a1p = c.take<double>(n);
  double* b1 = c.take<double>(n);
  double* b1p = c.take<double>(n);
  double* xsq = c.take<double>(n);
  double* xk = c.take<double>(n);
  double* cosx = c.take<double>(n);
  double* sinx = c.take<double>(n);
  double* sine = c.take<double>(n);
  const int half = (n + 1) / 2 + 1;
  double* rc = c.take<double>(half);
  double* rs = c.take<double>(half);
  int* rp = c.take<int>(half);
  int* rq = c.take<int>(half);
  double* sv = c.take<double>(kNT / 64);
  double* ev = c.take<double>(n);
  double* tmp = c.take<double>(n);
  int* order = c.take<int>(n);
  double* pv = c.take<double>(B.nham > 0 ? B.nham : 1);
  double* mats = A.lds_mats ? c.take<double>(kVivsMats * nn) : A.work + (size_t)inst * kVivsMats * nn;
  double* R = mats;
  double* W0 = R + nn;
  double* W1 = W0 + nn;
  double* W2 = W1 + nn;
  double* WM = W2 + nn;
  double* E11 = WM + nn;
  double* E12 = E11 + nn;
  double* E22 = E12 + nn;
  double* VN = E22 + nn;
  double* VO = VN + nn;
  double* TS = VO + nn;
  const double ered = A.ered[inst];
  const bool own = A.replay == 0;
  const VivsOpts& o = A.opt;
  const bool ivd = o.iv || o.idiag, ivpd = o.ivp || o.idiag, ivppd = o.ivpp || o.idiag;
  const bool need_w2 = ivppd || o.ishift;
  const bool do_pert = o.iv || o.ivp || o.ivpp;
  const bool pert1 = o.ivp && !o.ivpp;
  double* rg = A.rmat + (size_t)inst * nn;
  for (size_t e = tid; e < nn; e += kNT) {
    const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
    R[e] = rg[e];
    VN[e] = i == j ? 1.0 : 0.0;
    E11[e] = 0.0;
    E12[e] = 0.0;
    E22[e] = 0.0;
  }
  for (int i = tid; i < n; i += kNT) {
    g1[i] = 0.0;
    g1p[i] = 1.0;
    g2[i] = 1.0;
    g2p[i] = 0.0;
  }
  if (tid == 0) {
    st.istep = 1;
    st.drint = A.drnow;
    st.diagi = A.rstart + 0.5 * st.drint;
    st.rmid = A.rstart;
    st.rmidi = st.diagi;
    st.drnow = A.drnow;
    st.alpha = 1.0;
    st.balpha = 0.0;
    st.alfnew = 1.0;
    if (o.ialpha > 0) {
      st.balpha = (o.alpha2 - o.alpha1) / (A.rstop - A.rstart);
      st.alpha = o.alpha1 + st.balpha * (st.diagi - A.rstart);
      if (o.ialfp) st.alpha = o.alpha1;
      st.drnow = st.alpha == 1.0 ? st.drint / o.ialpha
                                 : st.drint * (st.alpha - 1.0) / (pow(st.alpha, (double)o.ialpha) - 1.0);
    }
    st.rnow = A.rstart + st.drnow;
    st.coffl = 0.0;
    st.xbar = st.xsbar = st.ebar = st.exbar = 0.0;
    st.endint = 0;
    st.icrmat = 1;
    st.last = 0;
    st.k = 0;
    st.stop = 0;
    st.status = kVivsOk;
    st.soff1 = st.sdiag1 = 0.0;
  }
  __syncthreads();
  while (true) {
    if (!own && tid == 0) {
      if (st.k >= A.nrec) {
        st.status = kVivsReplayShort;
        st.stop = 1;
      } else {
        const double* rec = A.steps + (size_t)st.k * 4;
        st.rnow = rec[0];
        st.drnow = rec[1];
        if (st.icrmat) st.rmidi = rec[2];
        st.last = (((int)rec[3]) & 2) != 0;
        st.endint = (((int)rec[3]) & 4) != 0;
        ++st.k;
      }
    }
    __syncthreads();
    if (st.stop) break;
    const double drnow = st.drnow;
    const double rcent = st.rnow - 0.5 * drnow;
    const bool icrmat = st.icrmat != 0;
    vivs_pvec(B, P, rcent, pv);
    __syncthreads();
    vivs_hammat(B, pv, rcent, ered, W0);
    if (icrmat && own && tid == 0) st.rmidi = st.diagi;
    __syncthreads();
    // DERMAT's derivatives at the step centre (vvprop.f:236-247)
    if (need_w2) {
      vivs_pvec_der(B, P, rcent, 2, pv);
      __syncthreads();
      vivs_dermat(B, pv, rcent, 2, W2);
      __syncthreads();
    }
    if (ivpd) {
      vivs_pvec_der(B, P, rcent, 1, pv);
      __syncthreads();
      vivs_dermat(B, pv, rcent, 1, W1);
      __syncthreads();
    }
    if (icrmat) {
      const double rmidi = st.rmidi;
      if (rmidi == rcent) {
        for (size_t e = tid; e < nn; e += kNT) TS[e] = W0[e];
      } else {
        vivs_pvec(B, P, rmidi, pv);
        __syncthreads();
        vivs_hammat(B, pv, rmidi, ered, TS);
      }
      for (size_t e = tid; e < nn; e += kNT) VO[e] = VN[e];
      __syncthreads();
      jacobi_eigh<kNT>(TS, VN, n, n, rc, rs, rp, rq, sv);
      __syncthreads();
      sort_eig<kNT>(TS, VN, n, n, ev, order, tmp);
      __syncthreads();
    }
    // the potential in the local basis (TRNSFM, vvprop.f:270)
    lds_gemm(W0, false, VN, false, TS, n);
    __syncthreads();
    lds_gemm(VN, true, TS, false, W0, n);
    __syncthreads();
    vivs_symmetrise(W0, n);
    __syncthreads();
    if (ivpd) {
      lds_gemm(W1, false, VN, false, TS, n);
      __syncthreads();
      lds_gemm(VN, true, TS, false, W1, n);
      __syncthreads();
      vivs_symmetrise(W1, n);
      __syncthreads();
    }
    if (need_w2) {
      lds_gemm(W2, false, VN, false, TS, n);
      __syncthreads();
      lds_gemm(VN, true, TS, false, W2, n);
      __syncthreads();
      vivs_symmetrise(W2, n);
      __syncthreads();
    }
    if (icrmat) {
      // R into the new basis (vvprop.f:292-300)
      lds_gemm(VO, true, VN, false, TS, n);
      __syncthreads();
      lds_gemm(R, false, TS, false, WM, n);
      __syncthreads();
      lds_gemm(TS, true, WM, false, R, n);
      __syncthreads();
      if (o.isym) vivs_symmetrise(R, n);
      if (tid == 0) st.icrmat = 0;
    }
    // reference potential and zeroth-order solutions (vvprop.f:305-395)
    const double factor = o.ishift ? drnow * drnow / 24.0 : 0.0;
    for (int i = tid; i < n; i += kNT) {
      const size_t ii = i + (size_t)i * n;
      double dif, w0d = 0.0;
      if (o.ishift) {
        // ISHIFT: the reference potential moved by h^2/24 W2_ii (vvprop.f:305-312)
        const double d2 = W2[ii];
        dif = -W0[ii] - factor * d2;
        w0d = -factor * d2;
      } else {
        dif = -W0[ii];
      }
      W0[ii] = w0d;
      dg[i] = dif;
      xsq[i] = dif * drnow * drnow;
      const double xl = sqrt(fabs(dif));
      const double x = xl * drnow;
      double sx, cx;
      if (dif < 0.0) {
        sx = sinh(x) / xl;
        cx = cosh(x);
      } else {
        sx = sin(x) / xl;
        cx = cos(x);
      }
      sinx[i] = sx;
      cosx[i] = cx;
      xk[i] = x;
      const double a = g1p[i], b = g1[i], cc = g2p[i], d = g2[i];
      a1[i] = b;
      a1p[i] = a;
      b1[i] = d;
      b1p[i] = cc;
      g1[i] = a * sx + b * cx;
      g1p[i] = a * cx - dif * b * sx;
      g2[i] = cc * sx + d * cx;
      g2p[i] = cc * cx - dif * d * sx;
    }
    __syncthreads();
    // contributions not wanted (vvprop.f:318-332): without IV, IVP or IVPP
    // no step integrals are taken, and the estimates read the potential
    // matrices kept (all of them under IDIAG, none otherwise)
    if (!ivd) {
      for (size_t e = tid; e < nn; e += kNT) W0[e] = 0.0;
    }
    if (!do_pert) {
      for (size_t e = tid; e < nn; e += kNT) {
        if (!ivpd) W1[e] = 0.0;
        if (!ivppd) W2[e] = 0.0;
      }
    }
    __syncthreads();
    // the step integrals of every pair i <= j
    const int ntri = n * (n + 1) / 2;
    if (do_pert) {
      for (int e = tid; e < ntri; e += kNT) {
        int jj = (int)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
        while ((jj + 1) * (jj + 2) / 2 <= e) ++jj;
        while (jj * (jj + 1) / 2 > e) --jj;
        const int ii = e - jj * (jj + 1) / 2;
        vivs_pair_integrals(n, ii, jj, xsq, drnow, W0, W1, W2, E11, E12, E22, a1, b1, a1p, b1p, ivpd, ivppd,
                            pert1);
      }
    }
    __syncthreads();
    if (own) {
      // normalised sizes of the step-start solutions (vvprop.f:416-432)
      const double drint = st.drint;
      for (int i = tid; i < n; i += kNT) {
        const double ad = fabs(dg[i]);
        double an = 1.0 / sqrt(a1p[i] * a1p[i] / ad + a1[i] * a1[i]);
        double bn = 1.0 / sqrt(b1p[i] * b1p[i] / ad + b1[i] * b1[i]);
        double anp = drnow * an / xk[i];
        double bnp = drnow * bn / xk[i];
        double w = 1.0;
        if (!(dg[i] > 0.0)) {
          w = exp(-xk[i] * drint / drnow);
          if (dg[i] < -o.xsqmax) w = 0.0;
          an *= w;
          bn *= w;
          anp *= w;
          bnp *= w;
        }
        a1[i] = an;
        b1[i] = bn;
        a1p[i] = anp;
        b1p[i] = bnp;
        sine[i] = w;
      }
      __syncthreads();
      // largest first-order corrections (vvprop.f:440-478)
      double coff = 0.0, cdiag = 0.0, soff = 0.0, sdiag = 0.0;
      for (size_t e = tid; e < nn; e += kNT) {
        const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
        const size_t ji = (size_t)j + (size_t)i * n;
        const double p1 = g1[j] * E12[e] - g2[j] * E11[e];
        const double p2 = g1[i] * E22[e] - g2[i] * E12[e];
        const double p1p = g1p[j] * E12[e] - g2p[j] * E11[e];
        const double p2p = g1p[i] * E22[e] - g2p[i] * E12[e];
        const double m4 = fmax(fmax(fabs(p1) * a1[j], fabs(p2) * b1[i]), fmax(fabs(p1p) * a1p[j], fabs(p2p) * b1p[i]));
        if (i != j)
          coff = fmax(coff, m4);
        else
          cdiag = fmax(cdiag, m4);
        if (j <= i) {
          const double ccij = W0[e], ccji = W0[ji], csij = W1[e], csji = W1[ji], ssij = W2[e], ssji = W2[ji];
          const double e1 = fabs(sinx[j] * csji - cosx[j] * ssij) * sine[j] * xk[j] / drnow;
          const double e2 = fabs(sinx[i] * ccij - cosx[i] * csji) * sine[i];
          const double e3 = fabs(cosx[j] * csji + dg[j] * sinx[j] * ssij) * sine[j];
          const double e4 = fabs(cosx[i] * ccij + dg[i] * sinx[i] * csji) * sine[i] * drnow / xk[i];
          const double e5 = fabs(sinx[i] * csij - cosx[i] * ssji) * sine[i] * xk[i] / drnow;
          const double e6 = fabs(sinx[j] * ccji - cosx[j] * csij) * sine[j];
          const double e7 = fabs(cosx[i] * csij + dg[i] * sinx[i] * ssji) * sine[i];
          const double e8 = fabs(cosx[j] * ccji + dg[j] * sinx[j] * csij) * sine[j] * drnow / xk[j];
          const double m8 = fmax(fmax(fmax(e1, e2), fmax(e3, e4)), fmax(fmax(e5, e6), fmax(e7, e8)));
          if (i != j)
            soff = fmax(soff, m8);
          else
            sdiag = fmax(sdiag, m8);
        }
      }
      coff = block_max<kNT>(coff, sv);
      cdiag = block_max<kNT>(cdiag, sv);
      soff = block_max<kNT>(soff, sv);
      sdiag = block_max<kNT>(sdiag, sv);
      if (tid == 0) {
        st.coff = coff;
        st.cdiag = cdiag;
        st.soff = soff == 0.0 ? 1.0e-30 : soff;
        st.sdiag = sdiag;
        // does the interval end here (vvprop.f:480-500)
        int endint;
        if (st.last) {
          endint = 1;
        } else if (o.ialpha > 0) {
          endint = st.istep % o.ialpha == 0;
        } else {
          // IALPHA <= 0: when the irregular solution's slope would grow past
          // 1e4, or the corrections' estimates grow towards their tolerances
          endint = -1;
          const double g2pmax = g2p[n - 1] * cosx[n - 1];
          if (fabs(g2pmax) > 1.0e4) {
            endint = 1;
          } else if (st.coffl != 0.0) {
            const double facc = fmin(st.coff / st.coffl, 2.0);
            const double facs = fmin(st.soff / st.soff1, 2.0);
            if (facc * st.coff > 0.8 * A.toff || facs * st.soff > 0.8 * A.tldiag) endint = 1;
          }
          if (endint < 0) {
            st.coffl = st.coff;
            st.soff1 = st.soff;
            st.sdiag1 = st.sdiag;
            endint = !(st.coff < 0.8 * A.toff && st.cdiag < 0.8 * A.toff);
          }
        }
        st.endint = endint;
        // every run ends: at its last step, or here after max_steps
        if (st.istep > A.max_steps) {
          st.status = kVivsTooManySteps;
          st.stop = 1;
        } else if (A.record) {
          double* rec = A.steps + (size_t)(st.istep - 1) * 4;
          rec[0] = st.rnow;
          rec[1] = st.drnow;
          rec[2] = st.rmidi;
          rec[3] = (double)((icrmat ? 1 : 0) + (st.last ? 2 : 0) + (endint ? 4 : 0));
        }
      }
    }
    __syncthreads();
    if (st.stop) break;
    const bool endint = st.endint != 0;
    if (endint) {
      // G + D for both solutions and their derivatives (vvprop.f:440-452,
      // 510-519); D = 0 without IPERT (vvprop.f:460-466)
      const bool pert = o.ipert != 0;
      for (size_t e = tid; e < nn; e += kNT) {
        const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
        const size_t ji = (size_t)j + (size_t)i * n;
        const double p1 = pert ? g1[j] * E12[e] - g2[j] * E11[e] : 0.0;
        const double p2 = pert ? g1[i] * E22[e] - g2[i] * E12[e] : 0.0;
        const double p1p = pert ? g1p[j] * E12[e] - g2p[j] * E11[e] : 0.0;
        const double p2p = pert ? g1p[i] * E22[e] - g2p[i] * E12[e] : 0.0;
        W0[e] = i == j ? p1p + g1p[i] : p1p;
        W2[e] = i == j ? p1 + g1[i] : p1;
        WM[ji] = i == j ? p2 + g2[i] : p2;
        W1[ji] = i == j ? p2p + g2p[i] : p2p;
      }
      if (own && tid == 0) {
        st.coffl = 0.0;
        st.soff1 = st.soff;
        st.sdiag1 = st.sdiag;
      }
      __syncthreads();
      // TS = W0 + R W1, VO = W2 + R WM (DSYMM, vvprop.f:520-523)
      for (size_t e = tid; e < nn; e += kNT) {
        const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
        double s1 = 0.0, s2 = 0.0;
        for (int m = 0; m < n; ++m) {
          const double r = R[i + (size_t)m * n];
          s1 += r * W1[m + (size_t)j * n];
          s2 += r * WM[m + (size_t)j * n];
        }
        TS[e] = s1 + W0[e];
        VO[e] = s2 + W2[e];
      }
      __syncthreads();
      const bool ok = vivs_gesv(TS, VO, n, st);
      if (!ok && tid == 0 && st.status == kVivsOk) st.status = kVivsSingular;
      for (size_t e = tid; e < nn; e += kNT) {
        const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
        R[e] = o.isym ? 0.5 * (VO[e] + VO[(size_t)j + (size_t)i * n]) : VO[(size_t)j + (size_t)i * n];
        E11[e] = 0.0;
        E12[e] = 0.0;
        E22[e] = 0.0;
      }
      for (int i = tid; i < n; i += kNT) {
        g1[i] = 0.0;
        g2[i] = 1.0;
        g1p[i] = 1.0;
        g2p[i] = 0.0;
      }
      if (tid == 0) st.icrmat = 1;
    }
    __syncthreads();
    if (tid == 0) {
      if (!st.last) ++st.istep;
      if (st.last) {
        st.stop = 1;
      } else if (own) {
        // next step and interval (vvprop.f:600-665)
        const double rstop = A.rstop;
        double rnow = st.rnow;
        st.xbar += rnow;
        st.xsbar += rnow * rnow;
        st.ebar += st.sdiag;
        st.exbar += rnow * st.sdiag;
        double dr;
        if (o.ialpha > 0) {
          dr = st.drnow * st.alpha;
        } else {
          // IALPHA <= 0: the step from DELRD on the step corrections
          double soff = st.soff;
          const double tmxx0 = fmax(0.5 * A.toff, A.tldiag);
          if (st.drint != st.drnow) soff = 0.8 * A.tldiag * pow(soff / (0.8 * tmxx0), 1.5);
          dr = st.drnow;
          if (!vivs_delrd(dr, st.sdiag, soff, A.tldiag, A.drmax, dg[0], dg[n - 1], rnow, rstop)) {
            st.status = kVivsStepTooSmall;
            st.stop = 1;
          }
        }
        if (st.icrmat) {
          double drint = rnow - st.rmid;
          double drint1 = drint;
          if (o.ialpha > 0) {
            // IALFP's prediction of ALPHA from the trend of the diagonal
            // corrections over the interval (vvprop.f:630-660)
            const double ia = (double)o.ialpha;
            const double xbar = st.xbar / ia, xsbar = st.xsbar / ia, ebar = st.ebar / ia, exbar = st.exbar / ia;
            const double slope = o.ialpha == 1 ? 0.0 : (exbar - xbar * ebar) / (xsbar - xbar * xbar);
            const double bint = ebar - xbar * slope;
            double emax = bint + slope * rnow;
            const double emin = bint + slope * (rnow - drint);
            st.alfnew = st.alpha;
            if (o.ialpha > 1) {
              if (emax == 0.0) emax = 1.0e-30;
              double fac = emin / emax;
              if (fac <= 0.0) {
                fac = emin <= 0.0 ? 0.9 : 1.1;
              } else {
                fac = pow(fac, 1.0 / (double)(3 * o.ialpha - 3));
                fac = fmax(fmin(fac, 1.1), 0.9);
              }
              st.alfnew = st.alpha * fac;
            }
          }
          st.xbar = st.xsbar = st.ebar = st.exbar = 0.0;
          const double tmxx = fmax(A.toff, A.tldiag);
          double coff = st.coff;
          if (drint != dr) coff = 0.8 * tmxx * pow(coff / (0.8 * tmxx), 1.5);
          bool okd = vivs_delrd(drint, st.cdiag, coff, tmxx, A.drmax, dg[0], dg[n - 1], rnow, rstop);
          double soff1 = st.soff1;
          if (drint1 != dr) soff1 = A.tldiag * pow(2.0 * soff1 / A.tldiag, 1.5);
          okd = vivs_delrd(drint1, st.sdiag1, soff1, A.tldiag, A.drmax, dg[0], dg[n - 1], rnow, rstop) && okd;
          if (!okd) {
            st.status = kVivsStepTooSmall;
            st.stop = 1;
          }
          if (fabs(drint1) < fabs(drint)) drint = drint1;
          if (drint < dr) drint = dr;
          if (fabs(rstop - rnow - drint) < fabs(0.01 * drint)) drint = rstop - rnow;
          if ((rstop - rnow - drint) * drint < 0.0) drint = rstop - rnow;
          st.drint = drint;
          st.rmid = rnow;
          st.diagi = rnow + 0.5 * drint;
          if (o.ialpha > 0) {
            st.alpha = o.alpha1 + st.balpha * (st.diagi - A.rstart);
            if (o.ialfp) st.alpha = st.alfnew;
            dr = st.alpha == 1.0 ? drint / o.ialpha
                                 : drint * (st.alpha - 1.0) / (pow(st.alpha, (double)o.ialpha) - 1.0);
          }
        }
        if (fabs(rstop - rnow - dr) < fabs(0.01 * dr)) dr = rstop - rnow;
        if ((rstop - rnow - dr) * dr < 0.0) dr = rstop - rnow;
        rnow = rnow + dr;
        st.drnow = dr;
        st.rnow = rnow;
        if (fabs((rnow - rstop) / dr) < 0.005) st.last = 1;
      }
    }
    __syncthreads();
    if (st.stop) break;
  }
  __syncthreads();
  // back to the primitive basis (vvprop.f:670-682)
  lds_gemm(VN, false, R, false, WM, n);
  __syncthreads();
  lds_gemm(WM, false, VN, true, R, n);
  __syncthreads();
  if (o.isym) vivs_symmetrise(R, n);
  __syncthreads();
  for (size_t e = tid; e < nn; e += kNT) rg[e] = R[e];
  if (tid == 0) {
    A.nstep[inst] = st.istep;
    A.status[inst] = st.status;
  }
}

}  // namespace mcp
