/*
 * mcprop: MI355X (gfx950) engine for the coupled-channel log-derivative
 * propagation and matching path of MOLSCAT / BOUND / FIELD.
 *
 * C ABI.  All matrices are column-major FP64 with leading dimension N unless
 * stated otherwise; integers are 32-bit; every entry point returns MCP_OK (0)
 * or a negative MCP_ERR_* code, and mcp_last_error() gives the message.
 * Per-instance failures (a singular inversion at one energy) are reported in
 * the per-instance status array and never affect the other instances of the
 * batch.  A context is used from one host thread at a time.
 *
 * Units are the reference's reduced units: energies in 1/RUNIT^2 (ERED,
 * EINT = CM2RU * E[cm-1]), radii in RUNIT, P_lambda(R) already multiplied by
 * EP2RU (HAMMAT steps 1-6, source_code/hammat.f:38-73, done by the caller).
 */
#ifndef MCPROP_H
#define MCPROP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCP_OK 0
#define MCP_ERR_ARG -1
#define MCP_ERR_HIP -2
#define MCP_ERR_UNSUPPORTED -3
#define MCP_ERR_NOMEM -4

/* Propagator code for ODPROP (source_code/odprop.f), the single-channel
 * LDMD propagator, which has no IPROP value of its own in the reference:
 * SCCTRL selects it for N = 1 under IPROP = 6 (source_code/scctrl.f:183-196),
 * as mcp_smatrix_ld does. */
#define MCP_PROP_ODPROP 106

typedef struct mcp_ctx mcp_ctx;

/* Context on HIP device `device`.  No reference counterpart (the reference
 * keeps its state in COMMON blocks, e.g. /PRPSCR/ at source_code/mdprop.f:56). */
int mcp_ctx_create(int device, mcp_ctx **ctx);
void mcp_ctx_destroy(mcp_ctx *ctx);
const char *mcp_last_error(const mcp_ctx *ctx);
/* Largest N the LDS-resident kernels take for the MDPROP / LDPROP / eigen
 * paths on this device, the LDS bytes per workgroup and the CU count. */
int mcp_ctx_info(const mcp_ctx *ctx, int32_t *max_n_md, int32_t *max_n_ld, int32_t *max_n_eig, int32_t *lds_bytes,
                 int32_t *num_cu);

/* SYMINV for a batch: replaces SYMINV(A,IA,N,IFAIL) (source_code/syminv.f:1,
 * with SYMINL source_code/syminl-alloc.f:1 for N > 30).  Reads the lower
 * triangle of each a[m], writes the full symmetric inverse to ainv[m] and
 * kount[m] = number of negative eigenvalues (IFAIL), N+1 for an exactly zero
 * pivot when N <= 30, -INFO (< 0) for an exactly singular column when N > 30. */
int mcp_syminv_batched(mcp_ctx *ctx, int32_t n, int32_t nmat, const double *a, double *ainv, int32_t *kount);

/* DIAGVC for a batch (source_code/diagvc.f:29-60, DSYEVR): eigenvalues
 * ascending, eigenvectors as columns (evecs may be NULL).  LDS-resident
 * Jacobi up to mcp_ctx_info's max_n_eig; beyond it, up to N = 512, a
 * global-memory solver (Householder tridiagonalisation, bisection, inverse
 * iteration) that reads the lower triangle. */
int mcp_eigh_batched(mcp_ctx *ctx, int32_t n, int32_t nmat, const double *a, double *evals, double *evecs);

/* Upload one (JTOT, IB) block: the coupling VL(NHAM, N(N+1)/2) with NHAM =
 * mxlam + nconst + nrsq (lambda fastest, upper triangle column by column, as
 * COUPLE/CPL9 build it: source_code/couple.f:120-127, sumlam.f:31-36), EINT
 * and CENT (source_code/base.f:386-409).  conlen <= 0 disables the harmonic
 * confinement term (hammat.f:101-106). */
int mcp_block_create(mcp_ctx *ctx, int32_t n, int32_t mxlam, int32_t nconst, int32_t nrsq, const double *vl,
                     const double *eint, const double *cent, double conlen, int32_t *block_id);
int mcp_block_destroy(mcp_ctx *ctx, int32_t block_id);

/* COUPLE for ITYPE = 1 (atom + linear rigid rotor, source_code/couple.f:
 * 51-143) on the device: the Percival-Seaton coefficients of the n channels
 * (rotor j = jch[i], orbital L = lch[i]) of total J = jtot for the nlam
 * Legendre terms lambdas[], written to vl in the reference's layout
 * VL(NLAM, N(N+1)/2) (lambda fastest, the upper triangle packed column by
 * column, source_code/couple.f:120-127).  Accurate to a few ulp of the
 * exact coefficients. */
int mcp_couple_rotor(mcp_ctx *ctx, int32_t n, const int32_t *jch, const int32_t *lch, int32_t jtot, int32_t nlam,
                     const int32_t *lambdas, double *vl);
/* mcp_block_create with VL made on the device by the same coefficients
 * (MXLAM = nlam, NCONST = NRSQ = 0): the (JTOT, IB) block of an ITYPE = 1
 * run without VL crossing the host bus (source_code/mol.driver.f:1235-1244
 * allocates VL, couple.f:51-143 fills it). */
int mcp_block_create_rotor(mcp_ctx *ctx, int32_t n, const int32_t *jch, const int32_t *lch, int32_t jtot,
                           int32_t nlam, const int32_t *lambdas, const double *eint, const double *cent,
                           double conlen, int32_t *block_id);

/* Upload a radial grid: npts radii and P(lambda, R) = EP2RU * POTENL(R*RSCALE)
 * (clamped at 1e60, after SCAPOT/PERTRB), ptab[k*mxlam + l].  For the LD
 * propagators the grid is R0 = RSTART, R_{2k-1} = midpoint, R_{2k} = end of
 * step k, accumulated as R = R + H (source_code/mdprop.f:113,203), so npts =
 * 2*NSTEP + 1, and dr is the step DR (H = DR/2 enters the propagator
 * coefficients, mdprop.f:64-68); dr = 0 for a grid not used by LDPROP/MDPROP. */
int mcp_grid_create(mcp_ctx *ctx, int32_t mxlam, int32_t npts, const double *r, const double *ptab, double dr,
                    int32_t *grid_id);
int mcp_grid_destroy(mcp_ctx *ctx, int32_t grid_id);

/* Upload an AIRY grid (AIPROP with TOLHI = 0, source_code/aiprop.f:86-327):
 * nstep steps of widths drnow[k] (one sign: positive outward, negative
 * inward), radii r = {RSTART, RA_0, RB_0, RA_1, RB_1, ...} (2*nstep+1) where
 * RA_k, RB_k = RNOW_k -+ DRNOW_k/(2*1.732050807) are POTENT's Gauss points
 * (source_code/potent.f:66-67) and RNOW_k the midpoints accumulated as
 * AIPROP does; ptab holds P(lambda, R) at those radii as for
 * mcp_grid_create. */
int mcp_grid_create_airy(mcp_ctx *ctx, int32_t mxlam, int32_t nstep, const double *r, const double *drnow,
                         const double *ptab, int32_t *grid_id);

/* HAMMAT (source_code/hammat.f:1-116) at nr grid points idx[] and energy
 * ered; pconst = the NCONST constants (x EP2RU) or NULL.  w: nr full N x N. */
int mcp_hammat(mcp_ctx *ctx, int32_t block_id, int32_t grid_id, int32_t nr, const int32_t *idx,
               const double *pconst, double ered, double *w);

/* YINIT (source_code/yinit.f:1-158, ADIAMN=WKBMN=.TRUE.) at grid point k0 for
 * ninst instances; y: ninst N x N. */
int mcp_yinit(mcp_ctx *ctx, int32_t block_id, int32_t grid_id, int32_t k0, int32_t outward, double bcyo,
              int32_t ninst, const double *ered, const double *pconst, double *y);

/* Log-derivative propagation over the whole grid for ninst instances of one
 * block: iprop = 6 replaces MDPROP (source_code/mdprop.f:1-4), iprop = 5
 * replaces LDPROP (source_code/ldprop.f:1-4), iprop = 7 replaces MAPROP
 * (source_code/maprop.f:1-4; NCONST = 0, N up to the LDS-resident limit),
 * iprop = MCP_PROP_ODPROP
 * replaces ODPROP (source_code/odprop.f:1-4; N = 1, NRSQ = 0; it takes the
 * sector ends and midpoints from the LD grid).  y holds the starting log-
 * derivative matrices on entry and Y(RSTOP) on exit; nodes = NODES; status =
 * 0 or the 1-based step at which an inversion failed ("MATRIX INVERSION
 * ERROR AT STEP K", mdprop.f:282).  pconst: ninst x nconst or NULL. */
int mcp_prop_ld(mcp_ctx *ctx, int32_t block_id, int32_t grid_id, int32_t iprop, int32_t ninst, const double *ered,
                const double *pconst, double *y, int32_t *nodes, int32_t *status);

/* MDPROP (iprop 6, source_code/mdprop.f:1-4) for the instances of nblk
 * blocks of uneven N on one LD grid, as one batched call: the reference's
 * partial-wave loop (source_code/mol.driver.f:1181-1856) calls MDPROP once
 * per (JTOT, IB, energy); here every block's energy-independent sector
 * matrices U(C) are built once and all (block, energy) instances go through
 * one persistent launch that takes them longest (N^3) first from a device
 * queue.  ninst[b] instances of block block_ids[b]; ered, nodes and status
 * hold the instances block after block, y their N_b x N_b matrices block
 * after block (Y(RSTART) in, Y(RSTOP) out).  Every block needs NCONST = 0,
 * N >= 2 and N within the LDS-resident MDPROP limit (mcp_ctx_info).  With
 * mcp_set_graph(ctx, 1) the launch sequence is captured as a HIP graph per
 * (grid, block set, instance counts) and replayed by later calls with the
 * same key. */
int mcp_prop_ld_multi(mcp_ctx *ctx, int32_t nblk, const int32_t *block_ids, int32_t grid_id, const int32_t *ninst,
                      const double *ered, double *y, int32_t *nodes, int32_t *status);

/* 1: mcp_prop_ld_multi captures its launches as a HIP graph and replays it
 * for later calls with the same grid, blocks and instance counts; 0 (the
 * default): ordinary launches.  No reference counterpart. */
int mcp_set_graph(mcp_ctx *ctx, int32_t on);

/* MGPROP's symplectic schemes (SYMSEL, source_code/mgprop.f:103-150):
 * imgsel = 4 (CS4, 5 substeps) or 5 (MA5, 6 substeps); a, b receive the
 * scheme's A (position advances) and B (potential weights), nsub entries. */
int mcp_mg_scheme(int32_t imgsel, int32_t *nsub, double *a, double *b);

/* MGPROP (source_code/mgprop.f:1-4; IPROP = 8, the symplectic log-derivative
 * propagators) for ninst instances of one block over nstep sectors of width
 * dr.  The grid (mcp_grid_create, its dr unused) holds the radius of every W
 * evaluation in the reference's order: per sector R = RSTART + (ISTEP-1)*DR,
 * then R += A_k*DR after substep k, one evaluation per substep except that
 * the first substep of sectors 2.. reuses the last W when the scheme's last
 * A is zero (CS4): m + (nstep-1)(m-1) radii for CS4, nstep*m for MA5.  y
 * holds Y(RSTART) on entry and Y(RSTOP) on exit; nodes = NODES (the inertia
 * of I + A_k DR Y, subtracted when A_k < 0); status = 0 or the 1-based
 * sector of an inversion failure.  N up to the LDS-resident limit. */
int mcp_prop_mg(mcp_ctx *ctx, int32_t block_id, int32_t grid_id, int32_t imgsel, int32_t nstep, double dr,
                int32_t ninst, const double *ered, const double *pconst, double *y, int32_t *nodes,
                int32_t *status);

/* AIPROP (source_code/aiprop.f:1-4, TOLHI = 0, NCONST = 0) over an AIRY grid
 * for ninst energies of one block: y holds Z (the log-derivative matrix) at
 * the grid's RSTART on entry and at its end on exit; nodes = NODES (KOUNT
 * summed outward, N - KOUNT inward); status = 0 or the 1-based step of an
 * inversion failure ("INSTABILITY IN SYMINV IN AIPROP", aiprop.f:212). */
int mcp_prop_airy(mcp_ctx *ctx, int32_t block_id, int32_t grid_id, int32_t ninst, const double *ered, double *y,
                  int32_t *nodes, int32_t *status);

/* POTENL's radial terms (source_code/potenl.f:160-200, the &POTL LAMBDA /
 * NTERM / NPOWER / A / E form) for the propagators that choose their radii as
 * they go: term t adds a[t] * x^npower[t] (npower != 0; x below 1e-8 raised
 * to 1e-8 for negative powers) or a[t] * exp(e[t] * x) (npower == 0) to
 * radial term lam[t] (0-based, < mxlam) at x = R * rscale; P(lam, R) = ep2ru
 * * the sum, clamped at 1e60 first (HAMMAT, source_code/hammat.f:38-56).  A
 * caller with SCAPOT's SCALAM != 1 (source_code/scapot-default.f:19, applied
 * to the clamped terms before EP2RU) passes ep2ru * SCALAM. */
int mcp_block_set_potl(mcp_ctx *ctx, int32_t block_id, int32_t nterm, const int32_t *lam, const int32_t *npower,
                       const double *a, const double *e, double rscale, double ep2ru);

/* VVPROP (source_code/vvprop.f:1-4; VIVS, IPROP = 4) for ninst energies of
 * one block (NCONST = NRSQ = 0, potential from mcp_block_set_potl) under the
 * options of mcp_set_vivs_options (by default the reference's: IV, IPERT,
 * ISYM, IALPHA = 6, IALFP off, source_code/mol.driver.f:277-349): rmat (ninst N x N) holds the R-matrix
 * (the inverse of Y, source_code/scctrl.f:150-166) at rstart on entry and at
 * rstop on exit; drnow = the first step, drmax = DRMAX, tldiag / toff = TLDIAG
 * / TOFF (scctrl.f:157-158: 0.064 sqrt(TOLHI / 0.001) for both).  The step
 * record steps (max_steps x 4: RNOW, DRNOW, RMIDI, flags with bit 0 = the
 * step opens an interval, bit 1 = the last step, bit 2 = the step ends an
 * interval) stands for the reference's
 * scratch file: replay 0 = every instance its own step control (steps, if
 * not NULL, receives instance 0's); replay 1 = every instance follows steps;
 * replay 2 = instance 0 chooses the steps, written to steps when not NULL,
 * and the others follow them, as the energies after the first do with
 * ISCRU (vvprop.f:155-160).  nstep = NSTEP; status = 0, 1 (DELRD: step size
 * too small, delrd.f:57), 2 (more than max_steps steps), 3 (an exactly zero
 * pivot in the interval-end DGESV, vvprop.f:600) or 4 (the record ended
 * before its last step). */
int mcp_prop_vivs(mcp_ctx *ctx, int32_t block_id, int32_t ninst, const double *ered, double *rmat, double rstart,
                  double rstop, double drnow, double drmax, double tldiag, double toff, int32_t replay,
                  int32_t max_steps, double *steps, int32_t *nstep, int32_t *status);

/* The /LDVVCM/ controls of VVPROP (source_code/vvprop.f:57-82, set in
 * source_code/mol.driver.f:277-349 and mol.prpini.f:75-83) for this
 * context's later mcp_prop_vivs calls; nonzero = .TRUE.  iv, ivp, ivpp: the
 * perturbation corrections from the constant part of W, its first and its
 * second derivative (DERMAT, source_code/dermat.f; PERT1 when ivp without
 * ivpp, else PERT2); ishift: the reference potential moved by h^2/24 times
 * the second derivative's diagonal; idiag: every diagonal correction (the
 * IVD/IVPD/IVPPD of vvprop.f:137-139; the driver's IDIAG default of also
 * setting iv, ivp, ivpp, ishift and ipert is the caller's, as MOL.PRPINI
 * does it); ipert: apply the corrections (else the steps only); isym:
 * symmetrise R; ialpha > 0: steps per interval in geometric progression with
 * ALPHA from alpha1 to alpha2 across the range (ialfp: predicted from the
 * diagonal corrections' trend), ialpha <= 0: each step from DELRD, an
 * interval ended when the corrections' estimates grow; xsqmax: closed
 * channels deeper than this (reduced units) drop out of the estimates.  The
 * defaults are 1, 0, 0, 0, 0, 1, 1, 6, 0, 1.0, 1.5, 1e4.  Derivatives are
 * taken with respect to R * RSCALE, as DERMAT takes them. */
int mcp_set_vivs_options(mcp_ctx *ctx, int32_t iv, int32_t ivp, int32_t ivpp, int32_t ishift, int32_t idiag,
                         int32_t ipert, int32_t isym, int32_t ialpha, int32_t ialfp, double alpha1, double alpha2,
                         double xsqmax);

/* AIPROP with TOLHI > 0 (source_code/aiprop.f:1-4, the steps chosen as the
 * propagation goes, aiprop.f:288-316) for ninst energies of one block
 * (NCONST = NRSQ = 0, potential from mcp_block_set_potl): z (ninst N x N)
 * holds Z, the log-derivative matrix, at rstart on entry and at rstop on
 * exit; dr0 = the first step (DRSET's DRFIT, source_code/drset.f:41-84),
 * powr = POWR, tolhi = TOLHI (0 < TOLHI < 1).  The step record steps
 * (max_steps x 2: RNOW, DRNOW of each step) stands for the reference's
 * scratch file: replay 0 = every instance its own steps (instance 0's
 * written to steps when not NULL, their count to *nrec); replay 1 = every
 * instance follows the *nrec steps of steps; replay 2 = instance 0 chooses
 * the steps (written back as for 0) and the others follow them, as the
 * energies after the first do with ISCRU.  nodes = NODES (KOUNT outward,
 * N - KOUNT inward); nstep = NSTEP; status = 0, the 1-based step of a
 * SYMINV instability ("INSTABILITY IN SYMINV IN AIPROP", aiprop.f:211-215),
 * -1 (not finished in max_steps steps: Z at the last point reached, as
 * aiprop.f:331-333) or -2 (a replay past the end of its record).  N up to
 * 448: the six N x N work matrices in LDS up to about N = 56 on gfx950, in
 * a per-instance global workspace beyond (slower: the eigen-solver and the
 * products then run on global memory). */
int mcp_prop_airy_tolhi(mcp_ctx *ctx, int32_t block_id, int32_t ninst, const double *ered, double *z, double rstart,
                        double rstop, double dr0, double powr, double tolhi, int32_t replay, int32_t max_steps,
                        double *steps, int32_t *nrec, int32_t *nodes, int32_t *nstep, int32_t *status);

/* YTRANS (source_code/ytrans.f90:1) for a block with NCONST > 0 and NRSQ = 0
 * on the MOLSCAT path: for each of ninst instances (pconst: ninst x nconst,
 * the constants as for mcp_prop_ld) the asymptotic operator, the sum of the
 * constant VL terms times their constants, is diagonalised within each set
 * of channels of equal CENT; eint (ninst x N) receives the thresholds EINT in
 * channel order (ytrans.f90:330-350) and, when y is not NULL, y (ninst N x N)
 * is turned into the asymptotic basis, Y <- EVEC^T Y EVEC (TRNSFM,
 * ytrans.f90:418-422).  WVEC, NOPEN and the channel order NB then follow
 * from eint by GTWVEC and CHNSRT, as the reference's driver does before the
 * propagation (source_code/mol.driver.f:1336-1342).  mcp_smatrix_ld and
 * mcp_smatrix_segments apply the same transformation to Y themselves when
 * the block has NCONST > 0.  N up to the LDS-resident limit (≈80). */
int mcp_ytrans(mcp_ctx *ctx, int32_t block_id, int32_t ninst, const double *pconst, double *y, double *eint);

/* YTRANS (source_code/ytrans.f90:1-485) as mcp_ytrans, for blocks with
 * NCONST > 0 and/or the centrifugal operator among the terms (NRSQ > 0),
 * with cent (ninst x N, may be NULL) receiving the CENT that YTOK takes in
 * the asymptotic basis.  NRSQ > 0: no grouping by L; MULTOP's resolution
 * (source_code/multop.f90:55-180) of the constant operator C and L^2 (the
 * NRSQ VL terms with P = 1): C diagonalised, L^2 diagonalised within each
 * set of C eigenvalues closer than DEGTOL (mcp_set_degtol) to their
 * neighbour, channels in ascending C then ascending L^2, eigenvectors with
 * the largest component positive, EINT = C's and CENT = L^2's eigenvalues
 * (ytrans.f90:264-275,353-361,411-430); without constant terms L^2 alone
 * (pconst may then be NULL) and EINT unchanged.  NRSQ = 0: CENT is the
 * block's.  mcp_smatrix_ld and mcp_smatrix_segments apply the same
 * transformation and match with these CENT.  Extra operators (NEXTRA) are
 * not taken. */
int mcp_ytrans_cent(mcp_ctx *ctx, int32_t block_id, int32_t ninst, const double *pconst, double *y, double *eint,
                    double *cent);

/* DEGTOL (source_code/mol.driver.f:286, 1e-10 by default): eigenvalues of
 * YTRANS's constant operator closer than this are one degenerate set. */
int mcp_set_degtol(mcp_ctx *ctx, double degtol);

/* YTOK + KTOS (source_code/ytok.f:1, ytokg.f:1, ktos.f:1) for ninst
 * instances: y (N x N), wvec (N, GTWVEC's signed wave vectors), nopen, nb
 * (N, 0-based CHNSRT order), cent (N), rval = RSTOP.  Writes kmat, sr, si as
 * nopen x nopen matrices with leading dimension nopen at stride N*N.
 * status: 0, a Bessel failure code, or 100+k for a singular DGESV pivot. */
int mcp_ytok_ktos(mcp_ctx *ctx, int32_t n, int32_t ninst, const double *y, const double *wvec,
                  const int32_t *nopen, const int32_t *nb, const double *cent, double rval, double *kmat,
                  double *sr, double *si, int32_t *status);

/* mcp_ytok_ktos without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Every array is a device pointer with mcp_ytok_ktos's layout:
 * inputs d_y [ninst][N*N], d_wvec [ninst][N], d_nopen [ninst], d_nb
 * [ninst][N], d_cent [N]; outputs d_kmat, d_sr, d_si [ninst][N*N] (nopen x
 * nopen, leading dimension nopen) and d_status [ninst] (as mcp_ytok_ktos,
 * or 99 for an instance whose NOPEN or NB is out of range).  The work
 * matrices and team barrier counters live in a fixed 32 MiB device array of
 * the library, one per GPU, held until the process ends: calls needing more
 * (mcp_ytok_ktos_noalloc_bytes(N, ninst) bytes; e.g. 10 instances at N=441)
 * fail with MCP_ERR_UNSUPPORTED.  Every call on a GPU shares that array, so
 * calls that overlap on one GPU (two contexts, two host threads) race.  The
 * inputs must be complete on the device before the call; it runs on the
 * context's stream and returns after that stream has finished. */
int mcp_ytok_ktos_noalloc(mcp_ctx *ctx, int32_t n, int32_t ninst, const double *d_y, const double *d_wvec,
                          const int32_t *d_nopen, const int32_t *d_nb, const double *d_cent, double rval,
                          double *d_kmat, double *d_sr, double *d_si, int32_t *d_status);

/* Scratch bytes mcp_ytok_ktos_noalloc needs for ninst instances of N
 * channels (the limit is 32 MiB).  No counterpart in the reference. */
size_t mcp_ytok_ktos_noalloc_bytes(int32_t n, int32_t ninst);

/* Whole MOLSCAT solve for one block on one LD grid (SCCTRL with one segment,
 * source_code/scctrl.f:80-289): YINIT at R0, propagation, YTOK, KTOS, with
 * everything kept on the device; per-instance nopen/nb/wvec as for
 * mcp_ytok_ktos; rval is the segment's nominal RSTOP, which YTOK matches at
 * (scctrl.f:287), not the radius accumulated step by step.  sr/si/kmat as
 * there; nodes/status per instance (status < 0: propagation failed at step
 * -status). */
int mcp_smatrix_ld(mcp_ctx *ctx, int32_t block_id, int32_t grid_id, int32_t iprop, int32_t ninst,
                   const double *ered, const double *pconst, const double *wvec, const int32_t *nopen,
                   const int32_t *nb, double rval, double *kmat, double *sr, double *si, int32_t *nodes,
                   int32_t *status);

/* SCCTRL with up to 8 segments (source_code/scctrl.f:80-289): YINIT at the
 * first grid's start, then each segment's propagator (iprops[k] = 5, 6 or
 * 9; 6 with N = 1 runs ODPROP, scctrl.f:183-196) on grid_ids[k], and YTOK +
 * KTOS at rval (the last segment's nominal RSTOP).  Outputs as for
 * mcp_smatrix_ld; nodes = NODES summed over the segments; status < 0:
 * -(100000*k + step) for a failure in segment k (0-based), -1000000 for an
 * open channel at the start radius. */
int mcp_smatrix_segments(mcp_ctx *ctx, int32_t block_id, int32_t nseg, const int32_t *grid_ids,
                         const int32_t *iprops, int32_t ninst, const double *ered, const double *pconst,
                         const double *wvec, const int32_t *nopen, const int32_t *nb, double rval, double *kmat,
                         double *sr, double *si, int32_t *nodes, int32_t *status);

/* BDCTRL matching (source_code/bdctrl.f:76-257) for ninst energies (or
 * fields, through pconst) of one block, two log-derivative segments as
 * BDPSET lays them out for RMID = RMATCH (source_code/bdpset.f:26-42):
 * YINIT at the first point of grid_out and propagation outward (DR > 0,
 * iprop_out 6 = MDPROP / 5 = LDPROP / 7 = MAPROP, as in mcp_prop_ld) to
 * RMATCH; YINIT with RSTART > RSTOP at the first point of grid_in and
 * propagation inward (DR < 0, iprop_in likewise; MAPROP counts N - KOUNT
 * nodes per sector inward, source_code/maprop.f:145) to RMATCH;
 * matching matrix M = Y_out - Y_in and DIAGVC.  Per instance: nodec = NODEC
 * (the nodes of both propagations plus NEGCNT = #{eig(M) <= 0}), eigmin =
 * EIGMIN (the eigenvalue of smallest modulus), imin = its 0-based index in
 * ascending order, psimch (N, or NULL) = PSIMCH with its largest component
 * positive, evals (N, or NULL) = eig(M) ascending.  status = 0, -k for an
 * inversion failure at outward step k, -(100000+k) inward, -1000000 for an
 * open channel at a start radius. */
int mcp_bound_match(mcp_ctx *ctx, int32_t block_id, int32_t grid_out, int32_t grid_in, int32_t iprop_out,
                    int32_t iprop_in, int32_t ninst, const double *ered, const double *pconst, int32_t *nodec,
                    double *eigmin, int32_t *imin, double *psimch, double *evals, int32_t *status);

/* BDCTRL's wavefunction (source_code/bdctrl.f:266-360, for MDPROP in both
 * directions, BOUND's IWAVE with IPROPS = IPROPL = 6) at one energy ered
 * (reduced units) of block_id on the outward grid gid_out (RMIN -> RMATCH,
 * DR > 0) and the inward grid gid_in (RMAX -> RMATCH, DR < 0): the matching
 * of mcp_bound_match (nodec = NODEC, eigmin = EIGMIN, psimch = PSIMCH, the
 * eigenvector of the matching matrix for EIGMIN, largest component
 * positive), with both propagations writing MDPROP's WAVE records
 * (source_code/mdprop.f:170-176,212-245: per step W3 = W W2, kept on the
 * device), then WVBKSB's back-substitution (source_code/wvbksb.f:64-114)
 * of PSIMCH through each segment's records, last to first: psi_out
 * ((NSTEP_out + 1) x N, row k = psi after k steps from RMATCH towards RMIN)
 * and psi_in ((NSTEP_in + 1) x N, towards RMAX), both unnormalised (row 0 =
 * PSIMCH).  The normalisation (WVBKSB's extended Simpson rule) and the
 * file are the host's.  status as for mcp_bound_match.  N within the
 * LDS-resident MDPROP limit. */
int mcp_bound_wave(mcp_ctx *ctx, int32_t block_id, int32_t gid_out, int32_t gid_in, double ered,
                   const double *pconst, int32_t *nodec, double *eigmin, double *psimch, double *psi_out,
                   double *psi_in, int32_t *status);

/* Workgroups that cooperate on one instance on the sequential large-N
 * kernels (N beyond the LDS-resident ones): g >= 1 asks for g (the
 * register-resident MDPROP takes ceil(N/16) by default, and between
 * ceil(N/32) and ceil(N/16) when a smaller g is asked for; the matching's
 * LU takes at most |g|), 0 restores the automatic choice, g < 0 keeps the
 * sequential MDPROP on the global-memory kernel with -g workgroups and the
 * large-N YINIT and AIRY inversions on one workgroup per instance (with
 * g = -1 no cooperative launch outside that sequential kernel).  No
 * reference counterpart. */
int mcp_set_team_size(mcp_ctx *ctx, int32_t g);

/* How the workgroup-team kernels (the inversions on teams, the sequential
 * large-N propagators, the matching's LU) are launched: cooperative != 0
 * (default) through hipLaunchCooperativeKernel; 0 as ordinary launches of
 * the same grids after the same residency check, for tools that cannot
 * trace cooperative launches.  No reference counterpart. */
int mcp_set_team_launch(mcp_ctx *ctx, int32_t cooperative);

/* How MDPROP runs for large N: 0 (default) solves the recurrence in parallel
 * over the radial grid, as a tree of merged sector maps, where the device
 * memory holds it, else sequentially; 1 always sequentially on the
 * register-resident matrix-core kernel; 2 sequentially on the global-memory
 * kernel.  Adding 4 puts DIAGVC (mcp_eigh_batched) and AIRY on the
 * global-memory eigen-solver for every N (by default only N beyond the
 * LDS-resident one, up to 512).  Same S-matrix and node count to rounding.
 * No reference counterpart. */
int mcp_set_large_mode(mcp_ctx *ctx, int32_t mode);

/* Batched inversions of the large-N paths (N beyond the LDS-resident
 * sizes): 1 (default) passes of 32 pivots for N <= 448, one workgroup per
 * matrix; 0 passes of 16 pivots.  Same inverse to rounding, same inertia.
 * No reference counterpart (SYMINV's blocking is LAPACK's). */
int mcp_set_inverse_kernel(mcp_ctx *ctx, int32_t mode);

/* Mixed precision for the large-N tree's batched inversions (the MDPROP and
 * AIRY trees, N <= 448): each inverse computed in FP32 on the matrix cores
 * (with its inertia from the FP32 pivots) and won back to FP64 by
 * newton_steps Newton-Schulz steps X <- X (2I - A X) in FP64; a matrix whose
 * FP32 inverse leaves max|I - A X| > 1e-2 is inverted again in FP64, and the
 * MDPROP sectors starting below r_fp64 (RUNIT; the short-range start where
 * closed channels grow exponentially) stay FP64, as do YINIT, the Y
 * recurrence's elementwise steps and the matching.  newton_steps = 0 (the
 * default) is FP64 throughout.  SYMINV's job (source_code/syminv.f) in
 * another precision; no reference counterpart. */
int mcp_set_precision(mcp_ctx *ctx, int32_t newton_steps, double r_fp64);

/* Since the last mcp_set_precision: matrices that took the mixed path, and
 * how many of them were inverted in FP64 after all (short-range or
 * untrusted FP32 inverse).  No reference counterpart. */
int mcp_precision_stats(const mcp_ctx *ctx, int64_t *mixed_matrices, int64_t *fp64_again);

/* Threads per instance of the LDS-resident MDPROP kernel: 64 (one wave) or
 * 256 (four waves, also what 0, the default, gives).  The results do not
 * depend on it (same arithmetic per element, same pivots).  No reference
 * counterpart: a tuning knob for the host program. */
int mcp_set_lds_threads(mcp_ctx *ctx, int32_t nt);

/* Which MDPROP kernel runs for N <= 64 (NHAM <= 4): 0 (default) chooses
 * the register-resident kernel (one wavefront per instance, each lane one
 * row of Y in registers) for batches of 512 instances or more and the
 * LDS-resident one below; 1 always the LDS-resident kernel; 2 always the
 * register-resident one.  Same recurrence and pivot decisions, results equal
 * to rounding.  No reference counterpart: a tuning knob for the host
 * program. */
int mcp_set_md_kernel(mcp_ctx *ctx, int32_t mode);

/* Wall time in milliseconds of the last mcp_prop_ld / mcp_smatrix_ld call's
 * device work, measured with HIP events on the library's stream, split into
 * the energy-independent sector matrices and the propagation loop. */
int mcp_last_timing(const mcp_ctx *ctx, double *ms_sector, double *ms_prop, double *ms_total);

/* The same two times for segment k (0-based) of the last
 * mcp_smatrix_segments call: for LD, the sector matrices and the
 * propagation; for AIRY, the grid's energy-independent step bases, slopes
 * and basis changes, and the propagation.  MCP_ERR_ARG for k beyond the
 * segments of that call.  No reference counterpart. */
int mcp_last_segment_timing(const mcp_ctx *ctx, int32_t k, double *ms_sector, double *ms_prop);

#ifdef __cplusplus
}
#endif
#endif
