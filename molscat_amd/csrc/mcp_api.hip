// C-ABI entry points of libmcprop (see include/mcprop.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mcprop.h"
#include "mcp_kernels.hpp"
#include "mcp_large.hpp"
#include "mcp_wide.hpp"
#include "mcp_mixed.hpp"
#include "mcp_airyprop.hpp"
#include "mcp_mfmaprop.hpp"
#include "mcp_tree.hpp"
#include "mcp_eig.hpp"
#include "mcp_ytrans.hpp"
#include "mcp_vivs.hpp"
#include "mcp_airytol.hpp"
#include "mcp_mgprop.hpp"
#include "mcp_maprop.hpp"
#include "mcp_regprop.hpp"
#include "mcp_couple.hpp"

namespace {

using mcp::kNT;
// threads per workgroup of the batched one-workgroup-per-matrix inversions
constexpr int kInvNT = 512;
// Instances from which the register-resident MDPROP (one wavefront per
// instance) is chosen over the LDS-resident one for N <= 64: enough to put
// a few waves on every CU.
constexpr int kRegAutoMinInst = 512;
// Largest buffer of the register-resident MDPROP's midpoint terms (3 N
// doubles per instance and sector); more instances go in chunks.
constexpr size_t kRegMidBytes = size_t(8) << 30;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    if (p) {
      (void)hipDeviceSynchronize();  // kernels queued earlier may still read it
      (void)hipFree(p);
      p = nullptr;
      bytes = 0;
    }
    hipError_t e = hipMalloc(&p, b ? b : 16);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  void release() {
    if (p) {
      (void)hipDeviceSynchronize();
      (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct Block {
  int n, mxlam, nconst, nrsq, nham;
  bool const_offdiag = false;  // some constant term couples two channels
  double conlen4inv;
  DevBuf vl, eint, cent;
  std::vector<double> h_cent;
  // POTENL's terms for the propagators that choose their radii as they go
  int potl_nterm = -1;  // -1: not set
  double potl_rscale = 1.0, potl_ep2ru = 1.0;
  DevBuf potl_lam, potl_npow, potl_a, potl_e;
};

struct Grid {
  int mxlam, npts;
  double dr = 0.0;
  DevBuf r, ptab;
  std::vector<double> h_r;
  bool airy = false;  // AIRY grid: RSTART, then RA_k, RB_k per step; widths in w
  int nstep = 0;
  DevBuf w;
  std::vector<double> h_w;
};

// LDS carve sizes, mirroring LdsCarve::take (16-byte rounded pieces).
struct LdsSize {
  size_t b = 0;
  template <class T>
  LdsSize& add(size_t count) {
    b += (count * sizeof(T) + 15) & ~(size_t)15;
    return *this;
  }
  LdsSize& sweep(int n) {
    add<double>(n).add<double>(n).add<double>(n).add<double>(n).add<int>(n);
    add<double>(kNT / 64).add<int>(kNT / 64);
    return *this;
  }
};

size_t lds_syminv(int n) { return LdsSize().add<double>((size_t)n * n).sweep(n).b; }
size_t lds_uc(int n, int nham) { return LdsSize().add<double>((size_t)n * n).add<double>(nham).sweep(n).b; }
size_t lds_md(int n, int nham) {
  return LdsSize().add<double>((size_t)n * n).add<double>(nham).add<double>(n).add<double>(n).add<double>(n).sweep(n).b;
}
size_t lds_ld(int n, int nham) {
  return LdsSize().add<double>((size_t)n * n).add<double>((size_t)n * n).add<double>(nham).sweep(n).b;
}
size_t lds_eig(int n) {
  const int half = (n + 1) / 2 + 1;
  return LdsSize()
      .add<double>((size_t)n * n)
      .add<double>((size_t)n * n)
      .add<double>(half)
      .add<double>(half)
      .add<int>(half)
      .add<int>(half)
      .add<double>(kNT / 64)
      .add<double>(n)
      .add<int>(n)
      .add<double>(n)
      .b;
}
size_t lds_yinit(int n, int nham) {
  const int half = (n + 1) / 2 + 1;
  return LdsSize()
      .add<double>((size_t)n * n)
      .add<double>((size_t)n * n)
      .add<double>(half)
      .add<double>(half)
      .add<int>(half)
      .add<int>(half)
      .add<double>(kNT / 64)
      .add<double>(nham)
      .add<double>(n)
      .b;
}

size_t lds_ytrans_rsq(int n) {
  // same pieces, in the same order, as k_ytrans_rsq carves them
  const size_t nn = (size_t)n * n;
  const size_t half = (size_t)(n + 1) / 2 + 1;
  LdsSize z;
  z.add<double>(nn).add<double>(nn).add<double>(nn).add<double>(nn);
  z.add<double>(half).add<double>(half).add<int>(half).add<int>(half);
  z.add<double>(kNT / 64).add<double>(n).add<double>(n).add<double>(n).add<int>(n).add<double>(n).add<int>(n).add<int>(n);
  return z.b;
}
size_t lds_ytrans(int n) {
  // same pieces, in the same order, as k_ytrans carves them
  const size_t nn = (size_t)n * n;
  const size_t half = (size_t)(n + 1) / 2 + 1;
  LdsSize z;
  z.add<double>(nn).add<double>(nn).add<double>(nn).add<double>(half).add<double>(half).add<int>(half).add<int>(half);
  z.add<double>(kNT / 64).add<double>(n).add<int>(n).add<double>(n).add<int>(n);
  return z.b;
}
// same pieces, in the same order, as mcp::k_vivs carves them
size_t lds_vivs(int n, int nham, bool mats) {
  const size_t half = (size_t)(n + 1) / 2 + 1;
  LdsSize z;
  for (int v = 0; v < 14; ++v) z.add<double>(n);
  z.add<double>(half).add<double>(half).add<int>(half).add<int>(half).add<double>(kNT / 64);
  z.add<double>(n).add<double>(n).add<int>(n).add<double>(nham > 0 ? nham : 1);
  if (mats) z.add<double>((size_t)mcp::kVivsMats * n * n);
  return z.b;
}
// same pieces, in the same order, as mcp::k_airy_tolhi carves them
size_t lds_airytol(int n, int nham, bool mats) {
  const size_t half = (size_t)(n + 1) / 2 + 1;
  LdsSize z;
  for (int v = 0; v < 6; ++v) z.add<double>(n);
  z.add<double>(half).add<double>(half).add<int>(half).add<int>(half).add<double>(kNT / 64);
  z.add<int>(n).add<double>(n).add<double>(nham > 0 ? nham : 1).add<double>(nham > 0 ? nham : 1).sweep(n);
  if (mats) z.add<double>((size_t)mcp::kAiryTolMats * n * n);
  return z.b;
}
// same pieces, in the same order, as the MAPROP kernels carve them
size_t lds_ma_basis(int n, int nham) {
  const size_t half = (size_t)(n + 1) / 2 + 1;
  LdsSize z;
  z.add<double>((size_t)n * n).add<double>((size_t)n * n);
  z.add<double>(half).add<double>(half).add<int>(half).add<int>(half).add<double>(kNT / 64);
  z.add<double>(n).add<int>(n).add<double>(n).add<double>(nham > 0 ? nham : 1);
  return z.b;
}
size_t lds_ma_mats(int n, int nham) {
  const size_t p = nham > 0 ? nham : 1;
  return LdsSize().add<double>((size_t)n * n).add<double>((size_t)n * n).add<double>((size_t)n * n)
      .add<double>(p).add<double>(p).add<double>(p).b;
}
size_t lds_ma_prop(int n) {
  return LdsSize().add<double>((size_t)n * n).add<double>((size_t)n * n).add<double>((size_t)n * n)
      .add<double>(n).add<double>(n).add<double>(kNT / 64).sweep(n).b;
}
size_t lds_airy_sector(int n, int nham) { return lds_eig(n) + LdsSize().add<double>(nham).add<double>(nham).b; }
size_t lds_airy_prop(int n) {
  return LdsSize().add<double>((size_t)n * n).add<double>((size_t)n * n).add<double>((size_t)n * n)
      .add<double>(n).add<double>(n).add<double>(n).sweep(n).b;
}

// same pieces, in the same order, as mcp::carve_reg
size_t lds_mfma(int n, int nham) {
  LdsSize z;
  z.add<double>(mcp::kPB * mcp::kPB).add<double>(2 * mcp::kPB).add<double>(mcp::kMW * mcp::kPB);
  z.add<double>(mcp::kPB * mcp::kPB).add<double>(nham).add<double>(n).add<double>(n).add<double>(n);
  z.add<int>(n).add<int>(n).add<int>(mcp::kPB).add<int>(8).add<double>(mcp::kMW).add<int>(mcp::kMW);
  return z.b;
}

size_t lds_large(int n, int nham, int nvec) {
  LdsSize z;
  // same pieces, in the same order, as mcp::carve_large
  const size_t ld = mcp::panel_ld(n);
  z.add<double>(mcp::kPB * ld).add<double>(mcp::kPB * ld).add<double>(mcp::kPB * mcp::kPB);
  z.add<double>(mcp::kPB * mcp::kPB).add<double>(mcp::kPB * mcp::kPB).add<double>(mcp::kMaxW * mcp::kPB);
  z.add<double>(n).add<double>(n).add<double>(2 * mcp::kPB).add<double>(mcp::kPB);
  z.add<int>(n).add<int>(n).add<int>(n).add<int>(mcp::kPB);
  z.add<int>(mcp::kPB).add<int>(mcp::kPB).add<int>(8);
  z.add<double>(mcp::kMaxW).add<int>(mcp::kMaxW);
  if (nham > 0) z.add<double>(nham);
  for (int v = 0; v < nvec; ++v) z.add<double>(n);
  return z.b;
}

// same pieces, in the same order, as mcp::carve_wide<T>
template <typename T = double>
size_t lds_wide(int n) {
  LdsSize z;
  const size_t ld = mcp::panel_ld(n);
  z.add<T>(mcp::kWP * ld);
  z.add<T>(mcp::kWP * mcp::kWP);
  z.add<T>(mcp::kWP);
  z.add<T>(mcp::kWW * mcp::kWP);
  z.add<int>(n).add<int>(n).add<int>(n).add<int>(n).add<int>(mcp::kWP).add<int>(8);
  z.add<T>(mcp::kWW);
  z.add<int>(mcp::kWW);
  return z.b;
}

}  // namespace

// Device memory of the allocation-free matching entry point
// (mcp_ytok_ktos_noalloc): the team barrier counters and YTOKG's work
// matrices.  One copy per GPU, held until the process ends; host code uses
// the address hipGetSymbolAddress gives on each device, never this name.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct mcp_ctx {
  mcp::VivsOpts vivs_opt = mcp::vivs_default_opts();
  int64_t coop_refused = 0;  // cooperative launches refused and taken as ordinary ones
  double degtol = 1.0e-10;   // DEGTOL of YTRANS's degenerate sets (source_code/mol.driver.f:286)
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t lds_max = 65536;
  int num_cu = 0;
  std::string err;
  std::map<int, std::unique_ptr<Block>> blocks;
  std::map<int, std::unique_ptr<Grid>> grids;
  int next_id = 1;
  DevBuf s_a, s_b, s_c, s_d, s_e, s_f, s_g, s_h, s_i, s_j, s_k;
  DevBuf s_y;  // YINIT square-root workspace (large N)
  DevBuf s_bar;  // per-instance team barrier counters + abort flag
  int team = 0;  // workgroups per instance on the large-N path (0: automatic)
  bool no_reg = false;  // keep the large-N MDPROP on the global-memory kernel
  int large_mode = 0;   // large-N MDPROP: 0 tree where memory allows, 1 sequential matrix-core, 2 global memory
  bool eig_global = false;  // DIAGVC and AIRY on the global-memory eigen-solver for every N
  bool coop = true;         // team kernels through the cooperative launch API
  int lds_threads = 0;      // threads per instance of the LDS-resident MDPROP (0: 256)
  int md_kernel = 0;        // small-N MDPROP: 0 automatic, 1 LDS-resident, 2 register-resident
  int inv_kernel = 1;       // batched large-N inversions: 1 32-pivot passes (N <= 448), 0 16-pivot passes
  unsigned char* scratch_dev = nullptr;  // this device's address of scratch (mcp_ytok_ktos_noalloc)
  // multi-block MDPROP (mcp_prop_ld_multi): sector matrices, inertias, block
  // table, item list, ERED, Y, nodes/status, queue head; and its graph
  DevBuf s_qu, s_qk, s_qb, s_qi, s_qe, s_qy, s_qn, s_qh;
  bool use_graph = false;
  // mixed precision on the large-N tree's batched inversions (0: FP64)
  int mixed_steps = 0;       // Newton steps after the FP32 inverse
  double mixed_r64 = 0.0;    // sectors starting below this radius stay FP64
  long long mixed_fp64 = 0;  // matrices the last calls inverted in FP64 again (guard)
  long long mixed_total = 0; // matrices the last calls took the mixed path for
  DevBuf s_m32, s_mx, s_mtt, s_mx2, s_mflag, s_mcnt;
  struct MultiGraph {
    std::vector<int> key;            // grid id, nblk, then (block id, ninst) pairs
    std::vector<void*> bufs;         // device buffers the graph was captured on
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  } mgraph;
  DevBuf s_u;  // LDPROP sector-matrix workspace (large N)
  DevBuf s_av, s_at, s_ae, s_ah;  // AIRY step bases, basis changes, eigenvalues, slopes
  DevBuf s_x;                     // column exchange of the register-resident MDPROP
  DevBuf s_dg;                    // its per-sector reference-potential terms
  DevBuf s_rw, s_rm;              // register-resident MDPROP: W(B) off the diagonal, midpoint terms
  DevBuf s_seg0, s_seg1;          // segments of the parallel-in-R MDPROP, two tree levels
  DevBuf s_zb, s_pq;              // its merge pivots / inverses and products
  DevBuf s_kc, s_acc;             // its negative-pivot counts
  DevBuf s_yv;                    // large-N YINIT: scales, flags, inertias
  DevBuf lnf_hi, lnf_lo;          // ln(k!) as hi + lo for the coupling coefficients
  int lnf_kmax = -1;
  DevBuf s_flag;                  // a device-side decision flag
  DevBuf s_aa, s_ad, s_az;        // large-N AIRY: step matrices / reflectors, slopes' matrices, Z of T / D V
  DevBuf s_ed, s_ew;              // global-memory eigen-solver: d, e, tau; inverse-iteration factors
  DevBuf s_ay, s_ax, s_am, s_akc; // large-N AIRY per instance: Y1 Y2 Y4, products, pivots, inertias
  DevBuf s_ye;                    // asymptotic thresholds of YTRANS
  DevBuf s_yc;                    // YTRANS with NRSQ > 0: each instance's CENT
  DevBuf s_vr, s_vw, s_vs, s_ve, s_vn, s_vst;  // VIVS: R, work matrices, step record, ERED, NSTEP, status
  DevBuf s_vnd;                                // AIRY with TOLHI > 0: NODES
  DevBuf s_mt, s_mc, s_mp, s_me, s_mq;         // MAPROP: bases, corrections, basis changes, eigenvalues, end term
  double t_sector = 0.0, t_prop = 0.0, t_total = 0.0;
  int seg_timed = 0;                       // segments of the last mcp_smatrix_segments call
  double seg_ms[8][2] = {};                // their (energy-independent, propagation) times

  int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }
  int hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return MCP_OK;
    if (e == hipErrorOutOfMemory) {
      size_t freeb = 0, totalb = 0;
      (void)hipGetLastError();
      (void)hipMemGetInfo(&freeb, &totalb);
      return fail(MCP_ERR_HIP, "%s: %s (device memory free %.2f of %.2f GiB)", what, hipGetErrorString(e),
                  freeb / 1073741824.0, totalb / 1073741824.0);
    }
    return fail(MCP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  }
};

#define HIPCHK(ctx, call)                               \
  do {                                                  \
    int _rc = (ctx)->hip((call), #call);                \
    if (_rc != MCP_OK) return _rc;                      \
  } while (0)

namespace {

// Raise every LDS kernel's dynamic-LDS limit to what the device allows after
// its static LDS; lds_max becomes the smallest such limit.
int set_lds_limits(mcp_ctx* c) {
  const void* fns[] = {(const void*)mcp::k_syminv,         (const void*)mcp::k_mdprop_uc,
                       (const void*)mcp::k_mdprop_lds<kNT>, (const void*)mcp::k_mdprop_lds<64>,     (const void*)mcp::k_ldprop_lds,
                       (const void*)mcp::k_eigh,           (const void*)mcp::k_bound_match,           (const void*)mcp::k_yinit,
                       (const void*)mcp::k_syminv_large<kInvNT>,   (const void*)mcp::k_mdprop_uc_large, (const void*)mcp::k_syminv_wide<double>, (const void*)mcp::k_syminv_wide<float>,
                       (const void*)mcp::k_mdprop_lds_queue<kNT>, (const void*)mcp::k_mdprop_lds_queue<64>,
                       (const void*)mcp::k_mdprop_large,   (const void*)mcp::k_ldprop_large,
                       (const void*)mcp::k_yinit_large,    (const void*)mcp::k_airy_sector,
                       (const void*)mcp::k_airy_prop,      (const void*)mcp::k_mdprop_mfma<1>, (const void*)mcp::k_mdprop_mfma<2>,
                       (const void*)mcp::k_syminv_mfma<1>, (const void*)mcp::k_match_lu,
                       (const void*)mcp::k_ytrans,         (const void*)mcp::k_vivs,           (const void*)mcp::k_airy_tolhi,
                       (const void*)mcp::k_mgprop,         (const void*)mcp::k_maprop_basis,
                       (const void*)mcp::k_maprop_mats,    (const void*)mcp::k_maprop_prop,
                       (const void*)mcp::k_mdprop_reg<16>, (const void*)mcp::k_mdprop_reg<32>,
                       (const void*)mcp::k_mdprop_reg<48>, (const void*)mcp::k_mdprop_reg<64>};
  const size_t total = c->lds_max;
  size_t lim = total;
  for (const void* f : fns) {
    hipFuncAttributes at;
    size_t stat = 0;
    if (hipFuncGetAttributes(&at, f) == hipSuccess) stat = at.sharedSizeBytes;
    (void)hipGetLastError();
    const size_t dyn = total > stat + 64 ? total - stat - 64 : 0;
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess) {
      (void)hipGetLastError();
      lim = std::min<size_t>(lim, 65536 - stat - 64);
    } else {
      lim = std::min(lim, dyn);
    }
  }
  c->lds_max = lim;
  return MCP_OK;
}

int max_n_for(size_t lds_max, size_t (*f)(int, int), int nham) {
  int n = 1;
  while (n < 4096 && f(n + 1, nham) <= lds_max) ++n;
  return n;
}

size_t w_uc(int n, int nham) { return lds_uc(n, nham); }
size_t w_md(int n, int nham) { return std::max(lds_md(n, nham), lds_uc(n, nham)); }
size_t w_ld(int n, int nham) { return lds_ld(n, nham); }
size_t w_eig(int n, int) { return lds_eig(n); }

template <class T>
int upload(mcp_ctx* c, DevBuf& b, const T* h, size_t count) {
  HIPCHK(c, b.ensure(count * sizeof(T)));
  if (count) HIPCHK(c, hipMemcpyAsync(b.p, h, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return MCP_OK;
}

template <class T>
int download(mcp_ctx* c, T* h, const DevBuf& b, size_t count) {
  if (count) HIPCHK(c, hipMemcpyAsync(h, b.p, count * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  return MCP_OK;
}

mcp::BlockArgs block_args(const Block& b) {
  mcp::BlockArgs a;
  a.n = b.n;
  a.nham = b.nham;
  a.mxlam = b.mxlam;
  a.nconst = b.nconst;
  a.nrsq = b.nrsq;
  a.vl = b.vl.as<double>();
  a.eint = b.eint.as<double>();
  a.cent = b.cent.as<double>();
  a.conlen4inv = b.conlen4inv;
  return a;
}

mcp::GridArgs grid_args(const Grid& g) {
  mcp::GridArgs a;
  a.npts = g.npts;
  a.r = g.r.as<double>();
  a.ptab = g.ptab.as<double>();
  return a;
}

int lookup(mcp_ctx* c, int32_t bid, int32_t gid, Block** b, Grid** g) {
  auto ib = c->blocks.find(bid);
  if (ib == c->blocks.end()) return c->fail(MCP_ERR_ARG, "unknown block id %d", bid);
  *b = ib->second.get();
  if (g) {
    auto ig = c->grids.find(gid);
    if (ig == c->grids.end()) return c->fail(MCP_ERR_ARG, "unknown grid id %d", gid);
    *g = ig->second.get();
    if ((*g)->mxlam != (*b)->mxlam)
      return c->fail(MCP_ERR_ARG, "grid has %d radial terms, block expects %d", (*g)->mxlam, (*b)->mxlam);
  }
  return MCP_OK;
}

// Workgroups per instance for the large-N propagators: all instances' teams
// must be co-resident (one workgroup per CU at this LDS size).
int team_size(mcp_ctx* c, int ninst) {
  int cap = std::max(1, c->num_cu / std::max(1, ninst));
  int g = c->team > 0 ? c->team : 32;
  return std::max(1, std::min(g, cap));
}

int prep_team(mcp_ctx* c, int ninst) {
  HIPCHK(c, c->s_bar.ensure((size_t)(ninst + 1) * sizeof(unsigned)));
  hipLaunchKernelGGL(mcp::k_zero_u32, dim3((ninst + 1 + 255) / 256), dim3(256), 0, c->stream, c->s_bar.as<unsigned>(),
                     ninst + 1);
  HIPCHK(c, hipGetLastError());
  return MCP_OK;
}

unsigned nblk256(size_t cnt) { return (unsigned)((cnt + 255) / 256); }

// Team kernels need all their workgroups resident at once.  By default they
// go through hipLaunchCooperativeKernel; with c->coop off they are ordinary
// launches of the same grids after the same residency check (every team is
// sized to fit the device, and a team barrier that waits too long sets the
// abort flag instead of hanging).  A cooperative launch the runtime refuses
// in its present state (hipErrorIllegalState, seen while another process
// held the device) takes the ordinary launch instead: same kernel, same
// results.
int launch_team(mcp_ctx* c, const void* kern, dim3 grid, dim3 block, void** args, size_t lds) {
  if (c->coop) {
    const hipError_t e = hipLaunchCooperativeKernel(kern, grid, block, args, lds, c->stream);
    if (e == hipSuccess) return MCP_OK;
    if (e != hipErrorIllegalState) HIPCHK(c, e);
    (void)hipGetLastError();
    ++c->coop_refused;
  }
  int per_cu = 0;
  HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, (int)(block.x * block.y * block.z), lds));
  if ((size_t)grid.x * grid.y * grid.z > (size_t)per_cu * c->num_cu)
    return c->fail(MCP_ERR_UNSUPPORTED, "team of %u workgroups does not fit the device", grid.x * grid.y * grid.z);
  HIPCHK(c, hipLaunchKernel(kern, grid, block, args, lds, c->stream));
  return MCP_OK;
}

// Full symmetric inverses and inertias of nb n x n matrices at a (stride
// n^2) in place, each on a team of ceil(n/16) workgroups on the matrix cores
// (for a few matrices: latency-bound, the whole device on them).  False when
// the device cannot hold one team.
bool team_inverse_fits(mcp_ctx* c, int n) {
  const int g = (n + mcp::kPB - 1) / mcp::kPB;
  return n <= mcp::kMMaxN && g <= c->num_cu && lds_mfma(n, 0) <= c->lds_max;
}

int team_inverse(mcp_ctx* c, int n, int nb, double* a, int* kount, int bk) {
  const int g = (n + mcp::kPB - 1) / mcp::kPB;
  const size_t lr = lds_mfma(n, 0);
  const int chunk = std::max(1, std::min(nb, c->num_cu / g));
  HIPCHK(c, c->s_x.ensure((size_t)chunk * 2 * mcp::kMSlots * mcp::panel_ld(n) * sizeof(double)));
  HIPCHK(c, c->s_bar.ensure((size_t)(chunk + 1) * sizeof(unsigned)));
  double* xg = c->s_x.as<double>();
  for (int b0 = 0; b0 < nb; b0 += chunk) {
    const int m = std::min(chunk, nb - b0);
    unsigned* ctrs = c->s_bar.as<unsigned>();
    int* abort_flag = (int*)(ctrs + m);
    hipLaunchKernelGGL(mcp::k_zero_u32, dim3((m + 1 + 255) / 256), dim3(256), 0, c->stream, ctrs, m + 1);
    HIPCHK(c, hipGetLastError());
    double* ab = a + (size_t)b0 * n * n;
    int* kb = kount + b0;
    int n_ = n, bk_ = bk, g_ = g;
    void* args[] = {&n_, &ab, &kb, &bk_, &g_, &ctrs, &abort_flag, &xg};
    const int rt = launch_team(c, (const void*)mcp::k_syminv_mfma<1>, dim3(m * g), dim3(mcp::kMT), args, lr);
    if (rt) return rt;
  }
  return MCP_OK;
}

// LOW: a symmetric result, only its tiles on and below the diagonal
template <bool TA, bool TB, bool LOW = false>
int tree_gemm(mcp_ctx* c, int n, int nbatch, int per, mcp::SegRef A, mcp::SegRef Bm, mcp::SegRef Ci, mcp::SegOut Co,
              double alpha, double beta) {
  const int nt = (n + mcp::kGT - 1) / mcp::kGT;
  const int ny = LOW ? nt * (nt + 1) / 2 : nt * nt;
  const long long grid = (long long)mcp::kXcd * ((nbatch + mcp::kXcd - 1) / mcp::kXcd) * ny;
  if (grid > 0x7fffffffLL) return c->fail(MCP_ERR_UNSUPPORTED, "tree_gemm: %d products of N=%d", nbatch, n);
  hipLaunchKernelGGL((mcp::k_gemm_seg<TA, TB, LOW>), dim3((unsigned)grid), dim3(256), 0, c->stream, n, per, nbatch, ny,
                     A, Bm, Ci, Co, alpha, beta);
  HIPCHK(c, hipGetLastError());
  return MCP_OK;
}

// Device bytes the parallel-in-R MDPROP needs for ninst instances.

// Symmetric inversions of one tree level: nb matrices at zb, inertias to kc.
// Small batches, where one workgroup per matrix leaves the device idle, go
// to workgroup teams (unless a negative team size was set).  At N=441 a
// team launch inverts 9 matrices in ~0.5 ms and a batched launch of up to
// 256 matrices takes one inversion's ~2.5 ms: teams win up to 45.
constexpr int kTreeTeamMax = 45;
bool tree_on_teams(mcp_ctx* c, int n, int nb) { return nb <= kTreeTeamMax && !c->no_reg && team_inverse_fits(c, n); }
// nb full symmetric inverses with inertias, one workgroup per matrix: 32
// pivots per pass where the candidate window fits the LDS, else 16
bool wide_fits(mcp_ctx* c, int n) {
  return c->inv_kernel == 1 && n <= mcp::kWideMaxN && lds_wide(n) <= c->lds_max;
}
int batched_inverse(mcp_ctx* c, int n, int nb, double* a, int* kc, int bk) {
  if (nb <= 0) return MCP_OK;
  if (wide_fits(c, n)) {
    hipLaunchKernelGGL(mcp::k_syminv_wide<double>, dim3(nb), dim3(mcp::kWNT), lds_wide(n), c->stream, n, a, kc, bk,
                       (const int*)nullptr, 0);
  } else {
    const size_t ll = lds_large(n, 0, 0);
    if (ll > c->lds_max) return c->fail(MCP_ERR_UNSUPPORTED, "inverse: N=%d exceeds the large-N panel limit", n);
    hipLaunchKernelGGL(mcp::k_syminv_large<kInvNT>, dim3(nb), dim3(kInvNT), ll, c->stream, n, a, kc, bk);
  }
  HIPCHK(c, hipGetLastError());
  return MCP_OK;
}
// nb inverses in FP32 won back to FP64 by Newton-Schulz steps (mcp_mixed.hpp);
// matrix m is FP64 throughout when (m % per) < cut, and again when its FP32
// inverse is not trusted.
int mixed_inverse(mcp_ctx* c, int n, int nb, int per, int cut, double* a, int* kc, int bk) {
  const size_t nn = (size_t)n * n;
  HIPCHK(c, c->s_m32.ensure((size_t)nb * nn * sizeof(float)));
  HIPCHK(c, c->s_mx.ensure((size_t)nb * nn * sizeof(double)));
  HIPCHK(c, c->s_mtt.ensure((size_t)nb * nn * sizeof(double)));
  HIPCHK(c, c->s_mx2.ensure((size_t)nb * nn * sizeof(double)));
  HIPCHK(c, c->s_mflag.ensure((size_t)nb * sizeof(int)));
  HIPCHK(c, c->s_mcnt.ensure(sizeof(int)));
  float* a32 = c->s_m32.as<float>();
  double* x = c->s_mx.as<double>();
  double* t = c->s_mtt.as<double>();
  double* x2 = c->s_mx2.as<double>();
  int* flag = c->s_mflag.as<int>();
  const unsigned ge = (unsigned)std::min<size_t>(nblk256((size_t)nb * nn), 65536);
  hipLaunchKernelGGL(mcp::k_mixed_down, dim3(ge), dim3(256), 0, c->stream, n, nb, per, cut, a, a32, flag);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(mcp::k_syminv_wide<float>, dim3(nb), dim3(mcp::kWNT), lds_wide<float>(n), c->stream, n, a32, kc,
                     bk, (const int*)flag, 0);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(mcp::k_mixed_up, dim3(ge), dim3(256), 0, c->stream, n, nb, (const float*)a32, x, (const int*)flag);
  HIPCHK(c, hipGetLastError());
  int rc;
  const mcp::SegRef A{a, nn, 0};
  for (int step = 0; step < c->mixed_steps; ++step) {
    const mcp::SegRef X{x, nn, 0}, T{t, nn, 0};
    // T = A X; X' = 2X - X T
    if ((rc = tree_gemm<false, false>(c, n, nb, 1, A, X, A, mcp::SegOut{t, nn, 0}, 1.0, 0.0))) return rc;
    if (step == 0) {
      hipLaunchKernelGGL(mcp::k_mixed_resid, dim3(nb), dim3(256), 0, c->stream, n, (const double*)t, flag,
                         mcp::kMixedTrust);
      HIPCHK(c, hipGetLastError());
    }
    if ((rc = tree_gemm<false, false>(c, n, nb, 1, X, T, X, mcp::SegOut{x2, nn, 0}, -1.0, 2.0))) return rc;
    std::swap(x, x2);
  }
  // the FP64 inversions: flagged beforehand (1) or by the guard (2)
  hipLaunchKernelGGL(mcp::k_syminv_wide<double>, dim3(nb), dim3(mcp::kWNT), lds_wide(n), c->stream, n, a, kc, bk,
                     (const int*)flag, 1);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(mcp::k_syminv_wide<double>, dim3(nb), dim3(mcp::kWNT), lds_wide(n), c->stream, n, a, kc, bk,
                     (const int*)flag, 2);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(mcp::k_tree_zero, dim3(1), dim3(64), 0, c->stream, c->s_mcnt.as<int>(), 1);
  hipLaunchKernelGGL(mcp::k_mixed_finish, dim3(ge), dim3(256), 0, c->stream, n, nb, (const double*)x, a,
                     (const int*)flag, c->s_mcnt.as<int>());
  HIPCHK(c, hipGetLastError());
  int cnt = 0;
  HIPCHK(c, hipMemcpyAsync(&cnt, c->s_mcnt.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->mixed_fp64 += cnt;
  c->mixed_total += nb;
  return MCP_OK;
}

// The tree's batched inversions: per matrices per instance, the first cut
// of each instance's (short-range) FP64 when mixed precision is on.
int tree_invert(mcp_ctx* c, int n, int nb, double* zb, int* kc, int bk, int per = 1, int cut = 0) {
  if (tree_on_teams(c, n, nb)) return team_inverse(c, n, nb, zb, kc, bk);
  if (c->mixed_steps > 0 && n <= mcp::kWideMaxN && lds_wide<float>(n) <= c->lds_max && cut < per)
    return mixed_inverse(c, n, nb, per, cut, zb, kc, bk);
  return batched_inverse(c, n, nb, zb, kc, bk);
}

// Buffers of a tree over nleaf segments per instance: two levels of
// segments, the merge pivots and inverses, the products, the inertias.
struct TreeBufs {
  double *in, *out, *zb, *pq;
  int *kc, *acc, *bad;
};

int tree_alloc(mcp_ctx* c, int n, int ninst, int nleaf, TreeBufs& t) {
  const size_t nn = (size_t)n * n, sz = 3 * nn;
  const size_t k1 = (size_t)std::max((nleaf + 1) / 2, 1);
  // the segment levels, and with mixed precision the FP32 copies and Newton
  // iterates of the widest batch (the leaves).  When one of them must grow,
  // all are given back first: the device then holds the new total, not the
  // new sizes on top of larger ones kept from an earlier, wider call (which
  // the caller's chunking counts as free).
  const size_t nleafs = (size_t)ninst * nleaf;
  const bool mixed = c->mixed_steps > 0;
  DevBuf* bufs[] = {&c->s_seg0, &c->s_seg1, &c->s_zb, &c->s_pq, &c->s_m32, &c->s_mx, &c->s_mtt, &c->s_mx2};
  const size_t need[] = {nleafs * sz * sizeof(double),
                         (size_t)ninst * k1 * sz * sizeof(double),
                         nleafs * nn * sizeof(double),
                         (size_t)ninst * k1 * 2 * nn * sizeof(double),
                         mixed ? nleafs * nn * sizeof(float) : 0,
                         mixed ? nleafs * nn * sizeof(double) : 0,
                         mixed ? nleafs * nn * sizeof(double) : 0,
                         mixed ? nleafs * nn * sizeof(double) : 0};
  bool grow = false;
  for (int i = 0; i < 8; ++i) grow |= need[i] > 0 && (!bufs[i]->p || bufs[i]->bytes < need[i]);
  if (grow)
    for (DevBuf* b : bufs) b->release();
  for (int i = 0; i < 8; ++i)
    if (need[i] > 0) HIPCHK(c, bufs[i]->ensure(need[i]));
  HIPCHK(c, c->s_kc.ensure((size_t)2 * ninst * nleaf * sizeof(int)));
  HIPCHK(c, c->s_acc.ensure((size_t)2 * ninst * sizeof(int)));
  t.in = c->s_seg0.as<double>();
  t.out = c->s_seg1.as<double>();
  t.zb = c->s_zb.as<double>();
  t.pq = c->s_pq.as<double>();
  t.kc = c->s_kc.as<int>();
  t.acc = c->s_acc.as<int>();
  t.bad = t.acc + ninst;
  hipLaunchKernelGGL(mcp::k_tree_zero, dim3(nblk256(2 * ninst)), dim3(256), 0, c->stream, t.acc, 2 * ninst);
  HIPCHK(c, hipGetLastError());
  return MCP_OK;
}

// The levels of a tree: the nleaf segments per instance at t.in merged
// pairwise, (2m, 2m+1) -> m, until one is left (at t.in); the merge pivots'
// negative-pivot counts summed into t.acc.
int tree_levels(mcp_ctx* c, int n, int ninst, int nleaf, TreeBufs& t, int bk, int cut_leaf = 0) {
  const size_t nn = (size_t)n * n, sz = 3 * nn;
  int rc;
  int span = 1;  // leaves per segment at this level
  for (int kl = nleaf; kl > 1;) {
    const int per = kl / 2, kn = (kl + 1) / 2;
    const int nb = ninst * per;
    // merge m covers leaves 2 m span ..; those starting below cut_leaf are short-range
    const int cut = (cut_leaf + 2 * span - 1) / (2 * span);
    hipLaunchKernelGGL(mcp::k_tree_w, dim3(nblk256((size_t)nb * nn)), dim3(256), 0, c->stream, n, ninst, kl, per,
                       t.in, t.zb, tree_on_teams(c, n, nb) ? 1 : 0);
    HIPCHK(c, hipGetLastError());
    if ((rc = tree_invert(c, n, nb, t.zb, t.kc, bk, per, cut))) return rc;
    hipLaunchKernelGGL(mcp::k_tree_count, dim3(nblk256(nb)), dim3(256), 0, c->stream, n, nb, per, t.kc, t.acc, t.bad);
    HIPCHK(c, hipGetLastError());
    const size_t si = (size_t)kl * sz, sm = 2 * sz;
    const mcp::SegRef Z{t.zb, (size_t)per * nn, nn};
    const mcp::SegRef L1{t.in, si, sm}, L2{t.in + nn, si, sm};
    const mcp::SegRef R2{t.in + sz + nn, si, sm}, R4{t.in + sz + 2 * nn, si, sm};
    const mcp::SegRef P{t.pq, (size_t)per * 2 * nn, 2 * nn}, Q{t.pq + nn, (size_t)per * 2 * nn, 2 * nn};
    const mcp::SegOut Po{t.pq, (size_t)per * 2 * nn, 2 * nn}, Qo{t.pq + nn, (size_t)per * 2 * nn, 2 * nn};
    const mcp::SegOut O1{t.out, (size_t)kn * sz, sz}, O2{t.out + nn, (size_t)kn * sz, sz};
    const mcp::SegOut O4{t.out + 2 * nn, (size_t)kn * sz, sz};
    // P = Z l2^T; Y1 = l1 - l2 P; Y2 = P^T r2; Q = Z r2; Y4 = r4 - r2^T Q
    // (Y1, Y4 symmetric: their lower tiles only)
    if ((rc = tree_gemm<false, true>(c, n, nb, per, Z, L2, Z, Po, 1.0, 0.0))) return rc;
    if ((rc = tree_gemm<false, false, true>(c, n, nb, per, L2, P, L1, O1, -1.0, 1.0))) return rc;
    if ((rc = tree_gemm<true, false>(c, n, nb, per, P, R2, P, O2, 1.0, 0.0))) return rc;
    if ((rc = tree_gemm<false, false>(c, n, nb, per, Z, R2, Z, Qo, 1.0, 0.0))) return rc;
    if ((rc = tree_gemm<true, false, true>(c, n, nb, per, R2, Q, R4, O4, -1.0, 1.0))) return rc;
    if (kl & 1) {
      hipLaunchKernelGGL(mcp::k_tree_pass, dim3(nblk256((size_t)ninst * sz)), dim3(256), 0, c->stream, n, ninst, kl,
                         kn, t.in, t.out);
      HIPCHK(c, hipGetLastError());
    }
    std::swap(t.in, t.out);
    kl = kn;
    span *= 2;
  }
  return MCP_OK;
}

// The whole map applied to the starting matrices: t.zb holds A0 + Y1 of each
// instance's last segment (at t.in); res = Y4 - Y2^T (A0 + Y1)^-1 Y2, its
// tiles on and below the diagonal.
int tree_final(mcp_ctx* c, int n, int ninst, TreeBufs& t, int bk, double* res, bool fp64 = false) {
  const size_t nn = (size_t)n * n, sz = 3 * nn;
  int rc;
  if ((rc = tree_invert(c, n, ninst, t.zb, t.kc, bk, 1, fp64 ? 1 : 0))) return rc;
  hipLaunchKernelGGL(mcp::k_tree_count, dim3(nblk256(ninst)), dim3(256), 0, c->stream, n, ninst, 1, t.kc, t.acc,
                     t.bad);
  HIPCHK(c, hipGetLastError());
  const mcp::SegRef Z{t.zb, nn, 0}, Y2{t.in + nn, sz, 0}, Y4{t.in + 2 * nn, sz, 0}, Q{t.pq, nn, 0};
  const mcp::SegOut Qo{t.pq, nn, 0}, Ao{res, nn, 0};
  if ((rc = tree_gemm<false, false>(c, n, ninst, 1, Z, Y2, Z, Qo, 1.0, 0.0))) return rc;
  if ((rc = tree_gemm<true, false, true>(c, n, ninst, 1, Y2, Q, Y4, Ao, -1.0, 1.0))) return rc;
  return MCP_OK;
}

size_t tree_bytes(int n, int ninst, int nstep) {
  const size_t nn = (size_t)n * n, k1 = (size_t)(nstep + 1) / 2;
  return sizeof(double) * nn * ninst * (3 * (size_t)nstep + 3 * k1 + (size_t)nstep + 2 * k1) +
         sizeof(int) * ((size_t)ninst * nstep + 2 * (size_t)ninst);
}

// MDPROP for large N as a tree of imbedding-map merges (mcp_tree.hpp): diag
// the output of k_mdprop_diag; fused: the leaves from M = V + E^-1 (no U(C)),
// else uc, uck the energy-independent sector matrices and their inertias.
// Y in I.y is replaced by Y(RSTOP).
int tree_mdprop(mcp_ctx* c, const mcp::BlockArgs& B, const mcp::GridArgs& G, const mcp::InstArgs& I, int ninst,
                int nstep, double h, bool fused, const double* uc, const int* uck, const double* diag, int bk,
                int cut_leaf = 0) {
  const int n = B.n;
  const size_t nn = (size_t)n * n;
  TreeBufs t;
  int rc;
  if ((rc = tree_alloc(c, n, ninst, nstep, t))) return rc;
  // leaves: each sector's two half-steps merged, Z = (UC + 2 A14)^-1
  const size_t nleaf = (size_t)ninst * nstep;
  const dim3 leaf_grid(nblk256(nn), (unsigned)(ninst * ((nstep + mcp::kLeafChunk - 1) / mcp::kLeafChunk)));
  if (fused) {
    int* epos = t.kc + nleaf;
    hipLaunchKernelGGL(mcp::k_tree_zero, dim3(nblk256(nleaf)), dim3(256), 0, c->stream, epos, (int)nleaf);
    hipLaunchKernelGGL(mcp::k_tree_m, leaf_grid, dim3(256), 0, c->stream, B, G, I, ninst, nstep, h, diag, t.zb, epos,
                       tree_on_teams(c, n, (int)nleaf) ? 1 : 0);
    if ((rc = tree_invert(c, n, (int)nleaf, t.zb, t.kc, bk, nstep, cut_leaf))) return rc;
    hipLaunchKernelGGL(mcp::k_tree_count_leaf, dim3(nblk256(nleaf)), dim3(256), 0, c->stream, n, (int)nleaf, nstep,
                       t.kc, epos, t.acc, t.bad);
  } else {
    hipLaunchKernelGGL(mcp::k_tree_c1, dim3(nblk256(nleaf * nn)), dim3(256), 0, c->stream, n, ninst, nstep, uc, diag,
                       t.zb);
    if ((rc = tree_invert(c, n, (int)nleaf, t.zb, t.kc, bk, nstep, cut_leaf))) return rc;
    hipLaunchKernelGGL(mcp::k_tree_count, dim3(nblk256(nleaf)), dim3(256), 0, c->stream, n, (int)nleaf, nstep, t.kc,
                       t.acc, t.bad);
  }
  hipLaunchKernelGGL(mcp::k_tree_leaf, leaf_grid, dim3(256), 0, c->stream, B, G, I, ninst, nstep, h, diag, t.zb,
                     fused ? 1 : 0, t.in);
  HIPCHK(c, hipGetLastError());
  if ((rc = tree_levels(c, n, ninst, nstep, t, bk, cut_leaf))) return rc;
  // the starting matrix: Y = Y4 - Y2^T (A0 + Y1)^-1 Y2
  hipLaunchKernelGGL(mcp::k_tree_final_w, dim3(nblk256((size_t)ninst * nn)), dim3(256), 0, c->stream, B, G, I, ninst, h,
                     t.in, t.zb);
  HIPCHK(c, hipGetLastError());
  if ((rc = tree_final(c, n, ninst, t, bk, t.out, cut_leaf > 0))) return rc;
  hipLaunchKernelGGL(mcp::k_tree_out, dim3(nblk256((size_t)ninst * nn)), dim3(256), 0, c->stream, n, ninst, h, t.out,
                     I.y);
  hipLaunchKernelGGL(mcp::k_tree_nodes, dim3(nblk256(ninst)), dim3(256), 0, c->stream, ninst, nstep,
                     fused ? nullptr : uck, t.acc, t.bad, I.nodes, I.status);
  HIPCHK(c, hipGetLastError());
  return MCP_OK;
}

// Device-side propagation of ninst instances whose Y already sits in dy.
int prop_ld_device(mcp_ctx* c, Block& b, Grid& g, int iprop, int ninst, const double* d_ered, const double* d_pconst,
                   double* dy, int* d_nodes, int* d_status) {
  if (iprop != 5 && iprop != 6 && iprop != 7 && iprop != MCP_PROP_ODPROP)
    return c->fail(MCP_ERR_ARG, "iprop %d: only 5 (LDPROP), 6 (MDPROP), 7 (MAPROP) and %d (ODPROP)", iprop,
                   MCP_PROP_ODPROP);
  if ((g.npts - 1) % 2 != 0 || g.npts < 3) return c->fail(MCP_ERR_ARG, "LD grid needs 2*NSTEP+1 points");
  const int nstep = (g.npts - 1) / 2;
  const int n = b.n;
  if (g.dr == 0.0) return c->fail(MCP_ERR_ARG, "grid was not created with an LD step DR");
  const double h = g.dr / 2.0;  // H = DR/2 (mdprop.f:64)
  const int bk = n > 30 ? 1 : 0;
  mcp::BlockArgs B = block_args(b);
  mcp::GridArgs G = grid_args(g);
  mcp::InstArgs I;
  I.ered = d_ered;
  I.pconst = d_pconst;
  I.y = dy;
  I.nodes = d_nodes;
  I.status = d_status;
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  if (iprop == MCP_PROP_ODPROP) {
    if (n != 1) return c->fail(MCP_ERR_ARG, "ODPROP needs N = 1 (N = %d)", n);
    if (b.nrsq != 0) return c->fail(MCP_ERR_UNSUPPORTED, "ODPROP with NRSQ > 0 (the reference leaves those P unset)");
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    hipLaunchKernelGGL(mcp::k_odprop, dim3((ninst + 63) / 64), dim3(64), 0, c->stream, B, G, I, ninst, nstep, g.dr);
    HIPCHK(c, hipGetLastError());
  } else if (iprop == 6 && b.const_offdiag && ninst > 1 && d_pconst) {
    // constant terms off the diagonal enter W_off(C), so U(C) is particular
    // to each instance's constants: the instances one at a time
    const size_t nn = (size_t)n * n;
    for (int i = 0; i < ninst; ++i) {
      int rc = prop_ld_device(c, b, g, iprop, 1, d_ered + i, d_pconst + (size_t)i * b.nconst, dy + i * nn,
                              d_nodes + i, d_status + i);
      if (rc) return rc;
    }
    return MCP_OK;
  } else if (iprop == 6) {
    // energy-independent sector matrices (shared by all instances)
    HIPCHK(c, c->s_e.ensure((size_t)nstep * n * n * sizeof(double)));
    HIPCHK(c, c->s_f.ensure((size_t)nstep * sizeof(int)));
    const size_t need = lds_md(n, b.nham);
    if (need <= c->lds_max && lds_uc(n, b.nham) <= c->lds_max) {
      hipLaunchKernelGGL(mcp::k_mdprop_uc, dim3(nstep), dim3(kNT), lds_uc(n, b.nham), c->stream, B, G, d_pconst, 1,
                         nstep, h, bk, c->s_e.as<double>(), c->s_f.as<int>());
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
      // four waves per instance by default; one wave (mcp_set_lds_threads)
      // measured slower at N = 12 and 48 even with 16 instances per CU
      const int nt = c->lds_threads ? c->lds_threads : kNT;
      const bool reg_ok = n <= mcp::kRegMaxN && b.nham <= mcp::kRegMaxHam && !b.const_offdiag &&
                          mcp::reg_lds_bytes((n + 15) / 16 * 16) <= c->lds_max;
      const bool use_reg = reg_ok && (c->md_kernel == 2 || (c->md_kernel == 0 && ninst >= kRegAutoMinInst));
      if (use_reg) {
        // W(B) off the diagonal for every sector, shared by the instances
        HIPCHK(c, c->s_rw.ensure((size_t)nstep * n * n * sizeof(double)));
        hipLaunchKernelGGL(mcp::k_mdprop_wb, dim3(nstep), dim3(256), 0, c->stream, B, G, nstep, c->s_rw.as<double>());
        HIPCHK(c, hipGetLastError());
        // the midpoint terms of every (instance, sector, channel), in chunks
        // of instances that keep them within kRegMidBytes
        const size_t per_inst = (size_t)nstep * 3 * n * sizeof(double);
        int chunk = (int)std::min<size_t>((size_t)ninst, std::max<size_t>(1, kRegMidBytes / per_inst));
        if (chunk < ninst) chunk = std::max(mcp::kRegWaves, chunk / mcp::kRegWaves * mcp::kRegWaves);
        HIPCHK(c, c->s_rm.ensure((size_t)chunk * per_inst));
        const int np = (n + 15) / 16 * 16;
        for (int i0 = 0; i0 < ninst; i0 += chunk) {
          const int m = std::min(chunk, ninst - i0);
          mcp::InstArgs Ic = I;
          Ic.ered = I.ered + i0;
          Ic.pconst = I.pconst ? I.pconst + (size_t)i0 * b.nconst : nullptr;
          Ic.y = I.y + (size_t)i0 * n * n;
          Ic.nodes = I.nodes + i0;
          Ic.status = I.status + i0;
          const size_t tot = (size_t)m * nstep * n;
          const int mblocks = (int)std::min<size_t>((tot + 255) / 256, (size_t)c->num_cu * 64);
          hipLaunchKernelGGL(mcp::k_mdprop_mid, dim3(mblocks), dim3(256), 0, c->stream, B, G, Ic, m, nstep, h,
                             c->s_rm.as<double>());
          HIPCHK(c, hipGetLastError());
          const dim3 grid((m + mcp::kRegWaves - 1) / mcp::kRegWaves), blk(64 * mcp::kRegWaves);
          const double* ucp = c->s_e.as<double>();
          const int* uck = c->s_f.as<int>();
          const double* wbp = c->s_rw.as<double>();
          const double* midp = c->s_rm.as<double>();
          if (np == 16)
            hipLaunchKernelGGL(mcp::k_mdprop_reg<16>, grid, blk, mcp::reg_lds_bytes(16), c->stream, B, G, Ic, m, nstep,
                               h, ucp, uck, wbp, midp, bk);
          else if (np == 32)
            hipLaunchKernelGGL(mcp::k_mdprop_reg<32>, grid, blk, mcp::reg_lds_bytes(32), c->stream, B, G, Ic, m, nstep,
                               h, ucp, uck, wbp, midp, bk);
          else if (np == 48)
            hipLaunchKernelGGL(mcp::k_mdprop_reg<48>, grid, blk, mcp::reg_lds_bytes(48), c->stream, B, G, Ic, m, nstep,
                               h, ucp, uck, wbp, midp, bk);
          else
            hipLaunchKernelGGL(mcp::k_mdprop_reg<64>, grid, blk, mcp::reg_lds_bytes(64), c->stream, B, G, Ic, m, nstep,
                               h, ucp, uck, wbp, midp, bk);
          HIPCHK(c, hipGetLastError());
        }
      } else if (nt == 64)
        hipLaunchKernelGGL(mcp::k_mdprop_lds<64>, dim3(ninst), dim3(64), need, c->stream, B, G, I, nstep, h,
                           c->s_e.as<double>(), c->s_f.as<int>(), bk);
      else
        hipLaunchKernelGGL(mcp::k_mdprop_lds<kNT>, dim3(ninst), dim3(kNT), need, c->stream, B, G, I, nstep, h,
                           c->s_e.as<double>(), c->s_f.as<int>(), bk);
      HIPCHK(c, hipGetLastError());
    } else {
      const size_t lu = lds_large(n, b.nham, 0), lm = lds_large(n, b.nham, 3);
      if (lm > c->lds_max) return c->fail(MCP_ERR_UNSUPPORTED, "MDPROP: N=%d exceeds the large-N panel limit", n);
      int rc;
      const double* ucp = c->s_e.as<double>();
      const int* uck = c->s_f.as<int>();
      // the energy-independent U(C) sector matrices, when a path needs them
      bool uc_done = false;
      auto need_uc = [&]() -> int {
        if (!uc_done) {
          hipLaunchKernelGGL(mcp::k_mdprop_uc_large, dim3(nstep), dim3(kNT), lu, c->stream, B, G, d_pconst, h, bk,
                             c->s_e.as<double>(), c->s_f.as<int>());
          HIPCHK(c, hipGetLastError());
          uc_done = true;
        }
        return MCP_OK;
      };
      // parallel-in-R tree where the device memory holds two of its levels
      if (c->large_mode == 0 && b.nham <= mcp::kLeafMaxHam) {
        size_t freeb = 0, totalb = 0;
        (void)hipMemGetInfo(&freeb, &totalb);
        const size_t held = c->s_seg0.bytes + c->s_seg1.bytes + c->s_zb.bytes + c->s_pq.bytes + c->s_dg.bytes +
                            c->s_m32.bytes + c->s_mx.bytes + c->s_mtt.bytes + c->s_mx2.bytes;
        size_t per = tree_bytes(n, 1, nstep) + (size_t)3 * nstep * n * sizeof(double);
        // mixed precision: the FP32 copies, the Newton iterates and products
        // of the widest level (the leaves)
        if (c->mixed_steps > 0) per += (size_t)nstep * n * n * (3 * sizeof(double) + sizeof(float)) + nstep * sizeof(int);
        // the short-range sectors (starting below mixed_r64) stay FP64
        int cut_leaf = 0;
        if (c->mixed_steps > 0)
          while (cut_leaf < nstep && g.h_r[2 * cut_leaf] < c->mixed_r64) ++cut_leaf;
        const size_t room = freeb + held > (size_t(2) << 30) ? freeb + held - (size_t(2) << 30) : 0;
        // as many instances per tree as the device memory holds, in turn
        const int chunk = (int)std::min<size_t>((size_t)ninst, room / per);
        if (chunk >= 1) {
          HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
          HIPCHK(c, c->s_flag.ensure(sizeof(int)));
          for (int i0 = 0; i0 < ninst; i0 += chunk) {
            const int m = std::min(chunk, ninst - i0);
            mcp::InstArgs Ic = I;
            Ic.ered = I.ered + i0;
            Ic.pconst = I.pconst ? I.pconst + (size_t)i0 * b.nconst : nullptr;
            Ic.y = I.y + (size_t)i0 * n * n;
            Ic.nodes = I.nodes + i0;
            Ic.status = I.status + i0;
            HIPCHK(c, c->s_dg.ensure((size_t)m * 3 * nstep * n * sizeof(double)));
            const size_t ndiag = (size_t)m * nstep * n;
            hipLaunchKernelGGL(mcp::k_mdprop_diag, dim3((unsigned)((ndiag + 255) / 256)), dim3(256), 0, c->stream, B,
                               G, Ic, m, nstep, h, c->s_dg.as<double>());
            HIPCHK(c, hipGetLastError());
            // leaves without U(C) unless some |2 A14 - 8| is small
            int flag = 0;
            HIPCHK(c, hipMemsetAsync(c->s_flag.p, 0, sizeof(int), c->stream));
            hipLaunchKernelGGL(mcp::k_tree_guard, dim3((unsigned)((ndiag + 255) / 256)), dim3(256), 0, c->stream, n, m,
                               nstep, c->s_dg.as<double>(), c->s_flag.as<int>());
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, hipMemcpyAsync(&flag, c->s_flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            const bool fused = flag == 0;
            if (!fused && (rc = need_uc())) return rc;
            if ((rc = tree_mdprop(c, B, G, Ic, m, nstep, h, fused, ucp, uck, c->s_dg.as<double>(), bk, cut_leaf)))
              return rc;
          }
          HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
          return MCP_OK;
        }
      }
      if ((rc = need_uc())) return rc;
      HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
      if ((rc = prep_team(c, ninst))) return rc;
      unsigned* ctrs = c->s_bar.as<unsigned>();
      int* abort_flag = (int*)(ctrs + ninst);
      // register-resident matrix-core kernel when the instances' teams fit
      // the device: one column block per workgroup, or two where there are
      // not CUs enough (or a smaller team was asked for)
      const int nblk = (n + mcp::kPB - 1) / mcp::kPB;
      int rb = 1, g = nblk;
      if ((c->team > 0 && c->team < nblk) || (size_t)nblk * ninst > (size_t)c->num_cu) {
        rb = 2;
        g = std::min(nblk, std::max(c->team, (nblk + 1) / 2));
      }
      const size_t lr = lds_mfma(n, b.nham);
      if (n <= mcp::kMMaxN && lr <= c->lds_max && !c->no_reg && c->large_mode <= 1 &&
          (size_t)g * ninst <= (size_t)c->num_cu) {
        HIPCHK(c, c->s_x.ensure((size_t)ninst * 2 * mcp::kMSlots * mcp::panel_ld(n) * sizeof(double)));
        HIPCHK(c, c->s_dg.ensure((size_t)ninst * 3 * nstep * n * sizeof(double)));
        double* xg = c->s_x.as<double>();
        const double* dgp = c->s_dg.as<double>();
        const size_t ndiag = (size_t)ninst * nstep * n;
        hipLaunchKernelGGL(mcp::k_mdprop_diag, dim3((unsigned)((ndiag + 255) / 256)), dim3(256), 0, c->stream, B, G, I,
                           ninst, nstep, h, c->s_dg.as<double>());
        HIPCHK(c, hipGetLastError());
        void* args[] = {&B,  &G, &I, (void*)&nstep, (void*)&h, &ucp, &uck, (void*)&bk, (void*)&g, &ctrs, &abort_flag,
                        &xg, &dgp};
        const void* kern = rb == 1 ? (const void*)mcp::k_mdprop_mfma<1> : (const void*)mcp::k_mdprop_mfma<2>;
        if ((rc = launch_team(c, kern, dim3(ninst * g), dim3(mcp::kMT), args, lr))) return rc;
      } else {
        const int g = team_size(c, ninst);
        void* args[] = {&B, &G, &I, (void*)&nstep, (void*)&h, &ucp, &uck, (void*)&bk, (void*)&g, &ctrs, &abort_flag};
        if ((rc = launch_team(c, (const void*)mcp::k_mdprop_large, dim3(ninst * g), dim3(kNT), args, lm))) return rc;
      }
    }
  } else if (iprop == 7) {
    // MAPROP: the energy-independent bases, corrections and basis changes of
    // every sector, then one workgroup per instance.  Constant terms enter
    // W, so with NCONST > 0 the bases are particular to each instance's
    // constants: the instances one at a time.
    if (b.nconst > 0 && ninst > 1) {
      const size_t nn1 = (size_t)n * n;
      for (int i = 0; i < ninst; ++i) {
        int r2 = prop_ld_device(c, b, g, iprop, 1, d_ered + i, d_pconst + (size_t)i * b.nconst, dy + i * nn1,
                                d_nodes + i, d_status + i);
        if (r2) return r2;
      }
      return MCP_OK;
    }
    const size_t lb = lds_ma_basis(n, b.nham), lm = lds_ma_mats(n, b.nham), lp = lds_ma_prop(n);
    if (std::max(lb, std::max(lm, lp)) > c->lds_max)
      return c->fail(MCP_ERR_UNSUPPORTED, "MAPROP: N=%d exceeds the LDS-resident limit", n);
    const size_t nn = (size_t)n * n;
    HIPCHK(c, c->s_mt.ensure((size_t)nstep * nn * sizeof(double)));
    HIPCHK(c, c->s_mc.ensure((size_t)nstep * nn * sizeof(double)));
    HIPCHK(c, c->s_mp.ensure((size_t)nstep * nn * sizeof(double)));
    HIPCHK(c, c->s_me.ensure((size_t)nstep * n * sizeof(double)));
    HIPCHK(c, c->s_mq.ensure(nn * sizeof(double)));
    double* mt = c->s_mt.as<double>();
    hipLaunchKernelGGL(mcp::k_maprop_basis, dim3(nstep), dim3(kNT), lb, c->stream, B, G, mt, c->s_me.as<double>(),
                       d_pconst);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(mcp::k_maprop_mats, dim3(nstep), dim3(kNT), lm, c->stream, B, G, nstep, g.dr / 6.0, mt,
                       c->s_mc.as<double>(), c->s_mp.as<double>(), c->s_mq.as<double>(), d_pconst);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    hipLaunchKernelGGL(mcp::k_maprop_prop, dim3(ninst), dim3(kNT), lp, c->stream, n, nstep, g.dr,
                       c->s_me.as<double>(), c->s_mc.as<double>(), c->s_mp.as<double>(), mt + (size_t)(nstep - 1) * nn,
                       c->s_mq.as<double>(), I, g.dr > 0.0 ? 1 : 0, bk);
    HIPCHK(c, hipGetLastError());
  } else {
    const size_t need = lds_ld(n, b.nham);
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    if (need <= c->lds_max) {
      hipLaunchKernelGGL(mcp::k_ldprop_lds, dim3(ninst), dim3(kNT), need, c->stream, B, G, I, nstep, h, bk);
      HIPCHK(c, hipGetLastError());
    } else {
      const size_t ll = lds_large(n, b.nham, 0);
      if (ll > c->lds_max) return c->fail(MCP_ERR_UNSUPPORTED, "LDPROP: N=%d exceeds the large-N panel limit", n);
      HIPCHK(c, c->s_u.ensure((size_t)ninst * n * n * sizeof(double)));
      const int g = team_size(c, ninst);
      int rc;
      if ((rc = prep_team(c, ninst))) return rc;
      double* uw = c->s_u.as<double>();
      unsigned* ctrs = c->s_bar.as<unsigned>();
      int* abort_flag = (int*)(ctrs + ninst);
      void* args[] = {&B, &G, &I, (void*)&nstep, (void*)&h, &uw, (void*)&bk, (void*)&g, &ctrs, &abort_flag};
      if ((rc = launch_team(c, (const void*)mcp::k_ldprop_large, dim3(ninst * g), dim3(kNT), args, ll))) return rc;
    }
  }
  HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  return MCP_OK;
}

// DIAGVC on the global-memory solver (mcp_eig.hpp): nmat matrices at a
// (device; lower triangle read, overwritten by the reflectors) to evals
// (nmat x n, ascending) and evecs (nmat x n x n, columns); zt scratch of
// nmat n^2 doubles.
int eig_global(mcp_ctx* c, int n, int nmat, double* a, double* evals, double* evecs, double* zt) {
  if (n > mcp::kEMaxN) return c->fail(MCP_ERR_UNSUPPORTED, "eigen-solver: N=%d above %d", n, mcp::kEMaxN);
  // the inverse iteration's factors take 5 n^2 doubles a matrix: batches of
  // at most 256 matrices
  const int chunk = std::min(nmat, 256);
  HIPCHK(c, c->s_ed.ensure((size_t)3 * nmat * n * sizeof(double)));
  HIPCHK(c, c->s_ew.ensure((size_t)5 * chunk * n * n * sizeof(double)));
  double* d = c->s_ed.as<double>();
  double* e = d + (size_t)nmat * n;
  double* tau = e + (size_t)nmat * n;
  for (int m0 = 0; m0 < nmat; m0 += chunk) {
    const int m = std::min(chunk, nmat - m0);
    const size_t o = (size_t)m0 * n, onn = o * n;
    hipLaunchKernelGGL(mcp::k_sytrd, dim3(m), dim3(mcp::kET), 0, c->stream, n, a + onn, d + o, e + o, tau + o);
    hipLaunchKernelGGL(mcp::k_stebz, dim3(m), dim3(mcp::kET), 0, c->stream, n, d + o, e + o, evals + o);
    hipLaunchKernelGGL(mcp::k_stein, dim3(m), dim3(mcp::kET), 0, c->stream, n, d + o, e + o, evals + o, zt + onn,
                       c->s_ew.as<double>());
    hipLaunchKernelGGL(mcp::k_ormtr, dim3((n + 15) / 16, m), dim3(256), 0, c->stream, n, a + onn, tau + o, zt + onn,
                       evecs + onn);
    HIPCHK(c, hipGetLastError());
  }
  return MCP_OK;
}

// YTRANS on the device (mcp_ytrans.hpp): the asymptotic thresholds of ninst
// instances (d_pconst: ninst x nconst) into d_eint, and their Y (dy, may be
// null) turned into the asymptotic basis.
int ytrans_device(mcp_ctx* c, Block& b, int ninst, const double* d_pconst, double* dy, double* d_eint) {
  if (b.nrsq > 0) {
    // the centrifugal operator among the terms: CENT per instance into s_yc
    const size_t lr = lds_ytrans_rsq(b.n);
    if (lr > c->lds_max) return c->fail(MCP_ERR_UNSUPPORTED, "ytrans: N=%d exceeds the LDS-resident limit", b.n);
    HIPCHK(c, c->s_yc.ensure((size_t)ninst * b.n * sizeof(double)));
    hipLaunchKernelGGL(mcp::k_ytrans_rsq, dim3(ninst), dim3(kNT), lr, c->stream, block_args(b), d_pconst, dy, d_eint,
                       c->s_yc.as<double>(), c->degtol);
    HIPCHK(c, hipGetLastError());
    return MCP_OK;
  }
  if (b.nconst <= 0) return c->fail(MCP_ERR_ARG, "ytrans: the block has no constant terms (NCONST = 0)");
  const size_t lds = lds_ytrans(b.n);
  if (lds > c->lds_max) return c->fail(MCP_ERR_UNSUPPORTED, "ytrans: N=%d exceeds the LDS-resident limit", b.n);
  hipLaunchKernelGGL(mcp::k_ytrans, dim3(ninst), dim3(kNT), lds, c->stream, block_args(b), d_pconst, dy, d_eint);
  HIPCHK(c, hipGetLastError());
  return MCP_OK;
}

// AIRY for N beyond the LDS-resident kernels: the energy-independent step
// data of the grid on the global-memory eigen-solver and the batched
// matrix-core products, then per step the two products, the team inversion
// and the elementwise updates for all instances together.
int airy_prop_global(mcp_ctx* c, Block& b, Grid& g, int ninst, const double* d_ered, double* dy, int* d_nodes,
                     int* d_status) {
  const int n = b.n, nstep = g.nstep;
  if (n > mcp::kEMaxN) return c->fail(MCP_ERR_UNSUPPORTED, "AIRY: N=%d above %d", n, mcp::kEMaxN);
  if (b.nham > mcp::kAiryMaxHam) return c->fail(MCP_ERR_UNSUPPORTED, "AIRY: more than %d VL terms", mcp::kAiryMaxHam);
  const size_t nn = (size_t)n * n, sn = (size_t)nstep * nn;
  HIPCHK(c, c->s_aa.ensure(sn * sizeof(double)));
  HIPCHK(c, c->s_ad.ensure(sn * sizeof(double)));
  HIPCHK(c, c->s_az.ensure(sn * sizeof(double)));
  HIPCHK(c, c->s_av.ensure(sn * sizeof(double)));
  HIPCHK(c, c->s_at.ensure(sn * sizeof(double)));
  HIPCHK(c, c->s_ae.ensure((size_t)nstep * n * sizeof(double)));
  HIPCHK(c, c->s_ah.ensure((size_t)nstep * n * sizeof(double)));
  HIPCHK(c, c->s_ay.ensure((size_t)ninst * nstep * 3 * n * sizeof(double)));
  HIPCHK(c, c->s_ax.ensure((size_t)ninst * nn * sizeof(double)));
  HIPCHK(c, c->s_am.ensure((size_t)ninst * nn * sizeof(double)));
  HIPCHK(c, c->s_akc.ensure((size_t)ninst * sizeof(int)));
  double* aa = c->s_aa.as<double>();
  double* ad = c->s_ad.as<double>();
  double* az = c->s_az.as<double>();
  double* av = c->s_av.as<double>();
  double* at = c->s_at.as<double>();
  double* ae = c->s_ae.as<double>();
  double* ah = c->s_ah.as<double>();
  double* ay = c->s_ay.as<double>();
  double* ax = c->s_ax.as<double>();
  double* am = c->s_am.as<double>();
  int* akc = c->s_akc.as<int>();
  const mcp::BlockArgs B = block_args(b);
  const mcp::GridArgs G = grid_args(g);
  const double* w = g.w.as<double>();
  int rc;
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  // every step's basis V_k, eigenvalues, slopes HP = diag(V^T D V), and the
  // basis changes T_0 = V_0, T_k = V_{k-1}^T V_k
  hipLaunchKernelGGL(mcp::k_airy_wd, dim3(nblk256(nn)), dim3(256), 0, c->stream, B, G, nstep, w, mcp::kPotentSq3, aa,
                     ad);
  HIPCHK(c, hipGetLastError());
  if ((rc = eig_global(c, n, nstep, aa, ae, av, az))) return rc;
  const mcp::SegRef Dk{ad, nn, 0}, Vk{av, nn, 0}, Vn{av + nn, nn, 0};
  if ((rc = tree_gemm<false, false>(c, n, nstep, 1, Dk, Vk, Dk, mcp::SegOut{az, nn, 0}, 1.0, 0.0))) return rc;
  hipLaunchKernelGGL(mcp::k_airy_hp, dim3(nblk256((size_t)nstep * n)), dim3(256), 0, c->stream, n, nstep, av, az, ah);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(at, av, nn * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  if (nstep > 1 &&
      (rc = tree_gemm<true, false>(c, n, nstep - 1, 1, Vk, Vn, Vk, mcp::SegOut{at + nn, nn, 0}, 1.0, 0.0)))
    return rc;
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  // the steps, all instances together
  hipLaunchKernelGGL(mcp::k_airy_y, dim3(nblk256((size_t)ninst * nstep * n)), dim3(256), 0, c->stream, n, nstep, ninst,
                     ae, ah, w, d_ered, ay);
  HIPCHK(c, hipGetLastError());
  const int outward = g.h_w[0] > 0.0 ? 1 : 0;
  const int bk = n > 30 ? 1 : 0;
  const mcp::SegRef Vl{av + (size_t)(nstep - 1) * nn, 0, 0};
  if (c->large_mode == 0) {
    // in parallel over the steps where the device memory holds the tree
    size_t freeb = 0, totalb = 0;
    (void)hipMemGetInfo(&freeb, &totalb);
    const size_t held = c->s_seg0.bytes + c->s_seg1.bytes + c->s_zb.bytes + c->s_pq.bytes + c->s_m32.bytes +
                        c->s_mx.bytes + c->s_mtt.bytes + c->s_mx2.bytes;
    // mixed precision adds the FP32 copies and Newton buffers of the widest level
    const size_t mixed_extra =
        c->mixed_steps > 0 ? (size_t)ninst * nstep * nn * (3 * sizeof(double) + sizeof(float)) : 0;
    if (freeb + held > tree_bytes(n, ninst, nstep) + mixed_extra + (size_t(2) << 30)) {
      TreeBufs t;
      if ((rc = tree_alloc(c, n, ninst, nstep, t))) return rc;
      const size_t nl = (size_t)ninst * nstep;
      hipLaunchKernelGGL(mcp::k_airy_leaf, dim3(nblk256(nl * nn)), dim3(256), 0, c->stream, n, ninst, nstep, at, ay,
                         t.zb, t.in);
      HIPCHK(c, hipGetLastError());
      // Y1 = (T diag(y1)) T^T
      if ((rc = tree_gemm<false, true>(c, n, (int)nl, nstep, mcp::SegRef{t.zb, (size_t)nstep * nn, nn},
                                       mcp::SegRef{at, 0, nn}, mcp::SegRef{t.zb, 0, 0},
                                       mcp::SegOut{t.in, (size_t)nstep * 3 * nn, 3 * nn}, 1.0, 0.0)))
        return rc;
      if ((rc = tree_levels(c, n, ninst, nstep, t, bk))) return rc;
      hipLaunchKernelGGL(mcp::k_airy_final_w, dim3(nblk256((size_t)ninst * nn)), dim3(256), 0, c->stream, n, ninst,
                         dy, t.in, t.zb);
      HIPCHK(c, hipGetLastError());
      if ((rc = tree_final(c, n, ninst, t, bk, dy))) return rc;
      hipLaunchKernelGGL(mcp::k_tree_mirror, dim3(nblk256((size_t)ninst * nn)), dim3(256), 0, c->stream, n, ninst, dy);
      hipLaunchKernelGGL(mcp::k_airy_tree_nodes, dim3(nblk256(ninst)), dim3(256), 0, c->stream, ninst, nstep, n, outward,
                         t.acc, t.bad, d_nodes, d_status);
      HIPCHK(c, hipGetLastError());
      // back to the primitive basis: Z <- sym(V Z V^T)
      const mcp::SegRef Z{dy, nn, 0}, X{ax, nn, 0};
      if ((rc = tree_gemm<false, false>(c, n, ninst, 1, Vl, Z, Vl, mcp::SegOut{ax, nn, 0}, 1.0, 0.0))) return rc;
      if ((rc = tree_gemm<false, true>(c, n, ninst, 1, X, Vl, X, mcp::SegOut{am, nn, 0}, 1.0, 0.0))) return rc;
      hipLaunchKernelGGL(mcp::k_airy_out, dim3(nblk256((size_t)ninst * nn)), dim3(256), 0, c->stream, n, ninst, am, dy);
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
      return MCP_OK;
    }
  }
  HIPCHK(c, hipMemsetAsync(d_nodes, 0, ninst * sizeof(int), c->stream));
  HIPCHK(c, hipMemsetAsync(d_status, 0, ninst * sizeof(int), c->stream));
  // a negative team size keeps the inversions on one workgroup per instance
  const bool team = !c->no_reg && team_inverse_fits(c, n);
  const size_t ll = lds_large(n, 0, 0);
  if (!team && ll > c->lds_max) return c->fail(MCP_ERR_UNSUPPORTED, "AIRY: N=%d exceeds the inversion limit", n);
  const unsigned ne = nblk256((size_t)ninst * nn);
  const mcp::SegRef Z{dy, nn, 0}, X{ax, nn, 0};
  for (int k = 0; k < nstep; ++k) {
    const mcp::SegRef Tk{at + (size_t)k * nn, 0, 0};
    // Z <- T^T (Z T), then M = sym(Z) + diag(Y1)
    if ((rc = tree_gemm<false, false>(c, n, ninst, 1, Z, Tk, Z, mcp::SegOut{ax, nn, 0}, 1.0, 0.0))) return rc;
    if ((rc = tree_gemm<true, false>(c, n, ninst, 1, Tk, X, X, mcp::SegOut{dy, nn, 0}, 1.0, 0.0))) return rc;
    hipLaunchKernelGGL(mcp::k_airy_m, dim3(ne), dim3(256), 0, c->stream, n, ninst, k, nstep, dy, ay, am);
    HIPCHK(c, hipGetLastError());
    if (team) {
      if ((rc = team_inverse(c, n, ninst, am, akc, bk))) return rc;
    } else {
      if ((rc = batched_inverse(c, n, ninst, am, akc, bk))) return rc;
    }
    hipLaunchKernelGGL(mcp::k_airy_z, dim3(ne), dim3(256), 0, c->stream, n, ninst, k, nstep, am, akc, ay, dy, d_nodes,
                       d_status, outward);
    HIPCHK(c, hipGetLastError());
  }
  // back to the primitive basis: Z <- sym(V Z V^T), V of the last step
  if ((rc = tree_gemm<false, false>(c, n, ninst, 1, Vl, Z, Vl, mcp::SegOut{ax, nn, 0}, 1.0, 0.0))) return rc;
  if ((rc = tree_gemm<false, true>(c, n, ninst, 1, X, Vl, X, mcp::SegOut{am, nn, 0}, 1.0, 0.0))) return rc;
  hipLaunchKernelGGL(mcp::k_airy_out, dim3(ne), dim3(256), 0, c->stream, n, ninst, am, dy);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  return MCP_OK;
}

// AIRY propagation of ninst instances whose Z already sits in dy (TOLHI = 0).
int airy_prop_device(mcp_ctx* c, Block& b, Grid& g, int ninst, const double* d_ered, double* dy, int* d_nodes,
                     int* d_status) {
  if (!g.airy) return c->fail(MCP_ERR_ARG, "AIRY propagation needs a grid from mcp_grid_create_airy");
  if (b.nconst != 0)
    return c->fail(MCP_ERR_UNSUPPORTED, "AIRY with NCONST > 0 (the local basis would depend on the instance)");
  const int n = b.n, nstep = g.nstep;
  const size_t ls = lds_airy_sector(n, b.nham), lp = lds_airy_prop(n);
  if (ls > c->lds_max || lp > c->lds_max || c->eig_global)
    return airy_prop_global(c, b, g, ninst, d_ered, dy, d_nodes, d_status);
  const size_t nn = (size_t)n * n;
  HIPCHK(c, c->s_av.ensure((size_t)nstep * nn * sizeof(double)));
  HIPCHK(c, c->s_at.ensure((size_t)nstep * nn * sizeof(double)));
  HIPCHK(c, c->s_ae.ensure((size_t)nstep * n * sizeof(double)));
  HIPCHK(c, c->s_ah.ensure((size_t)nstep * n * sizeof(double)));
  mcp::BlockArgs B = block_args(b);
  mcp::GridArgs G = grid_args(g);
  const double* w = g.w.as<double>();
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  hipLaunchKernelGGL(mcp::k_airy_sector, dim3(nstep), dim3(kNT), ls, c->stream, B, G, w, c->s_av.as<double>(),
                     c->s_ae.as<double>(), c->s_ah.as<double>());
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(mcp::k_airy_tmat, dim3(nstep), dim3(kNT), 0, c->stream, n, c->s_av.as<double>(),
                     c->s_at.as<double>());
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  mcp::InstArgs I;
  I.ered = d_ered;
  I.pconst = nullptr;
  I.y = dy;
  I.nodes = d_nodes;
  I.status = d_status;
  const int outward = g.h_w[0] > 0.0 ? 1 : 0;
  const double* vlast = c->s_av.as<double>() + (size_t)(nstep - 1) * nn;
  hipLaunchKernelGGL(mcp::k_airy_prop, dim3(ninst), dim3(kNT), lp, c->stream, n, nstep, c->s_at.as<double>(),
                     c->s_ae.as<double>(), c->s_ah.as<double>(), w, vlast, I, outward);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  return MCP_OK;
}

int record_timing(mcp_ctx* c) {
  float a = 0.f, b = 0.f;
  HIPCHK(c, hipEventSynchronize(c->ev[2]));
  HIPCHK(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
  HIPCHK(c, hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
  c->t_sector = a;
  c->t_prop = b;
  c->t_total = (double)a + (double)b;
  return MCP_OK;
}

}  // namespace

extern "C" {

int mcp_ctx_create(int device, mcp_ctx** out) {
  if (!out) return MCP_ERR_ARG;
  *out = nullptr;
  auto c = std::make_unique<mcp_ctx>();
  c->device = device;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return MCP_ERR_HIP;
  if (hipSetDevice(device) != hipSuccess) return MCP_ERR_HIP;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MCP_ERR_HIP;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    fprintf(stderr, "mcprop: built for gfx950, device %d is %s\n", device, prop.gcnArchName);
    return MCP_ERR_UNSUPPORTED;
  }
  c->num_cu = prop.multiProcessorCount;
  c->lds_max = prop.sharedMemPerBlock;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return MCP_ERR_HIP;
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) return MCP_ERR_HIP;
  set_lds_limits(c.get());
  // the allocation-free entry point's scratch, at its address on this device
  if (hipGetSymbolAddress((void**)&c->scratch_dev, HIP_SYMBOL(scratch)) != hipSuccess || !c->scratch_dev) {
    (void)hipGetLastError();
    fprintf(stderr, "mcprop: no address for the scratch array on device %d\n", device);
    return MCP_ERR_HIP;
  }
  *out = c.release();
  return MCP_OK;
}

void mcp_ctx_destroy(mcp_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->mgraph.exec) (void)hipGraphExecDestroy(c->mgraph.exec);
  if (c->mgraph.graph) (void)hipGraphDestroy(c->mgraph.graph);
  c->blocks.clear();
  c->grids.clear();
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* mcp_last_error(const mcp_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mcp_ctx_info(const mcp_ctx* c, int32_t* max_n_md, int32_t* max_n_ld, int32_t* max_n_eig, int32_t* lds_bytes,
                 int32_t* num_cu) {
  if (!c) return MCP_ERR_ARG;
  if (max_n_md) *max_n_md = max_n_for(c->lds_max, w_md, 8);
  if (max_n_ld) *max_n_ld = max_n_for(c->lds_max, w_ld, 8);
  if (max_n_eig) *max_n_eig = max_n_for(c->lds_max, w_eig, 8);
  if (lds_bytes) *lds_bytes = (int32_t)c->lds_max;
  if (num_cu) *num_cu = c->num_cu;
  return MCP_OK;
}

int mcp_syminv_batched(mcp_ctx* c, int32_t n, int32_t nmat, const double* a, double* ainv, int32_t* kount) {
  if (!c || n <= 0 || nmat < 0 || !a || !ainv || !kount) return c ? c->fail(MCP_ERR_ARG, "syminv: bad args") : MCP_ERR_ARG;
  if (nmat == 0) return MCP_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t lds = lds_syminv(n);
  const size_t cnt = (size_t)nmat * n * n;
  int rc;
  if (lds > c->lds_max) {
    const size_t ll = lds_large(n, 0, 0);
    if (ll > c->lds_max) return c->fail(MCP_ERR_UNSUPPORTED, "syminv: N=%d exceeds the large-N panel limit", n);
    if ((rc = upload(c, c->s_b, a, cnt))) return rc;
    HIPCHK(c, c->s_c.ensure((size_t)nmat * sizeof(int)));
    if ((rc = batched_inverse(c, n, nmat, c->s_b.as<double>(), c->s_c.as<int>(), n > 30 ? 1 : 0))) return rc;
    if ((rc = download(c, ainv, c->s_b, cnt))) return rc;
    if ((rc = download(c, kount, c->s_c, (size_t)nmat))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MCP_OK;
  }
  if ((rc = upload(c, c->s_a, a, cnt))) return rc;
  HIPCHK(c, c->s_b.ensure(cnt * sizeof(double)));
  HIPCHK(c, c->s_c.ensure((size_t)nmat * sizeof(int)));
  hipLaunchKernelGGL(mcp::k_syminv, dim3(nmat), dim3(kNT), lds, c->stream, n, c->s_a.as<double>(),
                     c->s_b.as<double>(), c->s_c.as<int>(), n > 30 ? 1 : 0);
  HIPCHK(c, hipGetLastError());
  if ((rc = download(c, ainv, c->s_b, cnt))) return rc;
  if ((rc = download(c, kount, c->s_c, (size_t)nmat))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MCP_OK;
}

int mcp_eigh_batched(mcp_ctx* c, int32_t n, int32_t nmat, const double* a, double* evals, double* evecs) {
  if (!c || n <= 0 || nmat < 0 || !a || !evals) return c ? c->fail(MCP_ERR_ARG, "eigh: bad args") : MCP_ERR_ARG;
  if (nmat == 0) return MCP_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t lds = lds_eig(n);
  const size_t cnt = (size_t)nmat * n * n;
  int rc;
  if (lds > c->lds_max || c->eig_global) {
    if ((rc = upload(c, c->s_a, a, cnt))) return rc;
    HIPCHK(c, c->s_b.ensure(cnt * sizeof(double)));
    HIPCHK(c, c->s_c.ensure((size_t)nmat * n * sizeof(double)));
    HIPCHK(c, c->s_d.ensure(cnt * sizeof(double)));
    if ((rc = eig_global(c, n, nmat, c->s_a.as<double>(), c->s_c.as<double>(), c->s_b.as<double>(),
                         c->s_d.as<double>())))
      return rc;
    if ((rc = download(c, evals, c->s_c, (size_t)nmat * n))) return rc;
    if (evecs && (rc = download(c, evecs, c->s_b, cnt))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MCP_OK;
  }
  if ((rc = upload(c, c->s_a, a, cnt))) return rc;
  HIPCHK(c, c->s_b.ensure(cnt * sizeof(double)));
  HIPCHK(c, c->s_c.ensure((size_t)nmat * n * sizeof(double)));
  hipLaunchKernelGGL(mcp::k_eigh, dim3(nmat), dim3(kNT), lds, c->stream, n, c->s_a.as<double>(),
                     c->s_c.as<double>(), evecs ? c->s_b.as<double>() : nullptr, (int*)nullptr);
  HIPCHK(c, hipGetLastError());
  if ((rc = download(c, evals, c->s_c, (size_t)nmat * n))) return rc;
  if (evecs && (rc = download(c, evecs, c->s_b, cnt))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MCP_OK;
}

int mcp_block_create(mcp_ctx* c, int32_t n, int32_t mxlam, int32_t nconst, int32_t nrsq, const double* vl,
                     const double* eint, const double* cent, double conlen, int32_t* block_id) {
  if (!c || n <= 0 || mxlam < 0 || nconst < 0 || nrsq < 0 || !vl || !eint || !cent || !block_id)
    return c ? c->fail(MCP_ERR_ARG, "block_create: bad args") : MCP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  auto b = std::make_unique<Block>();
  b->n = n;
  b->mxlam = mxlam;
  b->nconst = nconst;
  b->nrsq = nrsq;
  b->nham = mxlam + nconst + nrsq;
  b->conlen4inv = conlen > 0.0 ? 1.0 / (conlen * conlen * conlen * conlen) : 0.0;
  const size_t ntri = (size_t)n * (n + 1) / 2;
  int rc;
  // kRegMaxHam zeros past the end: the register-resident MDPROP reads
  // kRegMaxHam coefficients per element whatever NHAM (times P = 0 beyond it)
  HIPCHK(c, b->vl.ensure((ntri * b->nham + mcp::kRegMaxHam) * sizeof(double)));
  HIPCHK(c, hipMemsetAsync(b->vl.p, 0, (ntri * b->nham + mcp::kRegMaxHam) * sizeof(double), c->stream));
  if ((rc = upload(c, b->vl, vl, ntri * b->nham))) return rc;
  if ((rc = upload(c, b->eint, eint, (size_t)n))) return rc;
  if ((rc = upload(c, b->cent, cent, (size_t)n))) return rc;
  b->h_cent.assign(cent, cent + n);
  for (int j = 0; j < n && !b->const_offdiag; ++j)
    for (int i = 0; i < j && !b->const_offdiag; ++i) {
      const double* v = vl + ((size_t)j * (j + 1) / 2 + i) * b->nham + mxlam;
      for (int k = 0; k < nconst; ++k)
        if (v[k] != 0.0) b->const_offdiag = true;
    }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int id = c->next_id++;
  c->blocks[id] = std::move(b);
  *block_id = id;
  return MCP_OK;
}

int mcp_block_destroy(mcp_ctx* c, int32_t id) {
  if (!c) return MCP_ERR_ARG;
  (void)hipStreamSynchronize(c->stream);
  return c->blocks.erase(id) ? MCP_OK : c->fail(MCP_ERR_ARG, "unknown block id %d", id);
}

// ---------------------------------------------------------------------------
// COUPLE for ITYPE = 1 on the device (source_code/couple.f:51-143)
// ---------------------------------------------------------------------------
static int couple_rotor_device(mcp_ctx* c, int32_t n, const int32_t* jch, const int32_t* lch, int32_t jtot,
                               int32_t nlam, const int32_t* lambdas, double* d_vl) {
  int jmax = 0, lmax = 0, lmx = 0;
  for (int i = 0; i < n; ++i) {
    if (jch[i] < 0 || lch[i] < 0) return c->fail(MCP_ERR_ARG, "couple_rotor: negative j or L at channel %d", i);
    jmax = std::max(jmax, (int)jch[i]);
    lmax = std::max(lmax, (int)lch[i]);
  }
  for (int i = 0; i < nlam; ++i) {
    if (lambdas[i] < 0) return c->fail(MCP_ERR_ARG, "couple_rotor: negative lambda");
    lmx = std::max(lmx, (int)lambdas[i]);
  }
  const int need = mcp::couple_lnf_need(jmax, lmax, jtot, lmx);
  if (need > (1 << 20)) return c->fail(MCP_ERR_ARG, "couple_rotor: quantum numbers too large (ln k! up to %d)", need);
  if (need > c->lnf_kmax) {
    const int kmax = std::max(need, 2047);
    std::vector<double> hi(kmax + 1), lo(kmax + 1);
    mcp::build_lnfact(kmax, hi.data(), lo.data());
    int rc;
    if ((rc = upload(c, c->lnf_hi, hi.data(), hi.size()))) return rc;
    if ((rc = upload(c, c->lnf_lo, lo.data(), lo.size()))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->lnf_kmax = kmax;
  }
  DevBuf dj, dl, dlam;
  int rc;
  if ((rc = upload(c, dj, jch, (size_t)n))) return rc;
  if ((rc = upload(c, dl, lch, (size_t)n))) return rc;
  if ((rc = upload(c, dlam, lambdas, (size_t)nlam))) return rc;
  const long long ntri = (long long)n * (n + 1) / 2;
  const long long tot = ntri * nlam;
  mcp::LnFact f{(const double*)c->lnf_hi.p, (const double*)c->lnf_lo.p, c->lnf_kmax};
  if (tot > 0) {
    hipLaunchKernelGGL(mcp::k_couple_rotor, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream,
                       (const int*)dj.p, (const int*)dl.p, ntri, (int)jtot, (const int*)dlam.p, (int)nlam, f, d_vl);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));  // the small uploads above are freed on return
  return MCP_OK;
}

int mcp_couple_rotor(mcp_ctx* c, int32_t n, const int32_t* jch, const int32_t* lch, int32_t jtot, int32_t nlam,
                     const int32_t* lambdas, double* vl) {
  if (!c || n <= 0 || nlam <= 0 || jtot < 0 || !jch || !lch || !lambdas || !vl)
    return c ? c->fail(MCP_ERR_ARG, "couple_rotor: bad args") : MCP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  DevBuf d;
  const size_t cnt = (size_t)n * (n + 1) / 2 * nlam;
  HIPCHK(c, d.ensure(cnt * sizeof(double)));
  int rc;
  if ((rc = couple_rotor_device(c, n, jch, lch, jtot, nlam, lambdas, (double*)d.p))) return rc;
  if ((rc = download(c, vl, d, cnt))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MCP_OK;
}

int mcp_block_create_rotor(mcp_ctx* c, int32_t n, const int32_t* jch, const int32_t* lch, int32_t jtot,
                           int32_t nlam, const int32_t* lambdas, const double* eint, const double* cent, double conlen,
                           int32_t* block_id) {
  if (!c || n <= 0 || nlam <= 0 || jtot < 0 || !jch || !lch || !lambdas || !eint || !cent || !block_id)
    return c ? c->fail(MCP_ERR_ARG, "block_create_rotor: bad args") : MCP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  auto b = std::make_unique<Block>();
  b->n = n;
  b->mxlam = nlam;
  b->nconst = 0;
  b->nrsq = 0;
  b->nham = nlam;
  b->conlen4inv = conlen > 0.0 ? 1.0 / (conlen * conlen * conlen * conlen) : 0.0;
  const size_t ntri = (size_t)n * (n + 1) / 2;
  HIPCHK(c, b->vl.ensure((ntri * b->nham + mcp::kRegMaxHam) * sizeof(double)));
  HIPCHK(c, hipMemsetAsync(b->vl.p, 0, (ntri * b->nham + mcp::kRegMaxHam) * sizeof(double), c->stream));
  int rc;
  if ((rc = couple_rotor_device(c, n, jch, lch, jtot, nlam, lambdas, (double*)b->vl.p))) return rc;
  if ((rc = upload(c, b->eint, eint, (size_t)n))) return rc;
  if ((rc = upload(c, b->cent, cent, (size_t)n))) return rc;
  b->h_cent.assign(cent, cent + n);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int id = c->next_id++;
  c->blocks[id] = std::move(b);
  *block_id = id;
  return MCP_OK;
}

int mcp_grid_create(mcp_ctx* c, int32_t mxlam, int32_t npts, const double* r, const double* ptab, double dr,
                    int32_t* grid_id) {
  if (!c || mxlam < 0 || npts <= 0 || !r || (mxlam > 0 && !ptab) || !grid_id)
    return c ? c->fail(MCP_ERR_ARG, "grid_create: bad args") : MCP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  auto g = std::make_unique<Grid>();
  g->mxlam = mxlam;
  g->npts = npts;
  g->dr = dr;
  int rc;
  if ((rc = upload(c, g->r, r, (size_t)npts))) return rc;
  if ((rc = upload(c, g->ptab, ptab, (size_t)npts * mxlam))) return rc;
  g->h_r.assign(r, r + npts);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int id = c->next_id++;
  c->grids[id] = std::move(g);
  *grid_id = id;
  return MCP_OK;
}

int mcp_grid_create_airy(mcp_ctx* c, int32_t mxlam, int32_t nstep, const double* r, const double* drnow,
                         const double* ptab, int32_t* grid_id) {
  if (!c || mxlam < 0 || nstep <= 0 || !r || !drnow || (mxlam > 0 && !ptab) || !grid_id)
    return c ? c->fail(MCP_ERR_ARG, "grid_create_airy: bad args") : MCP_ERR_ARG;
  for (int k = 0; k < nstep; ++k)
    if (!(drnow[k] != 0.0) || (drnow[k] > 0.0) != (drnow[0] > 0.0))
      return c->fail(MCP_ERR_ARG, "grid_create_airy: step widths must be nonzero and of one sign");
  int32_t id = 0;
  int rc;
  if ((rc = mcp_grid_create(c, mxlam, 2 * nstep + 1, r, ptab, 0.0, &id))) return rc;
  Grid& g = *c->grids[id];
  g.airy = true;
  g.nstep = nstep;
  if ((rc = upload(c, g.w, drnow, (size_t)nstep))) return rc;
  g.h_w.assign(drnow, drnow + nstep);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *grid_id = id;
  return MCP_OK;
}

int mcp_set_graph(mcp_ctx* c, int32_t on) {
  if (!c || on < 0 || on > 1) return c ? c->fail(MCP_ERR_ARG, "graph mode %d: 0 or 1", on) : MCP_ERR_ARG;
  c->use_graph = on != 0;
  return MCP_OK;
}

int mcp_prop_ld_multi(mcp_ctx* c, int32_t nblk, const int32_t* block_ids, int32_t gid, const int32_t* ninst,
                      const double* ered, double* y, int32_t* nodes, int32_t* status) {
  if (!c || nblk < 0 || (nblk > 0 && (!block_ids || !ninst || !ered || !y || !nodes || !status)))
    return c ? c->fail(MCP_ERR_ARG, "prop_ld_multi: bad args") : MCP_ERR_ARG;
  if (nblk == 0) return MCP_OK;
  auto ig = c->grids.find(gid);
  if (ig == c->grids.end()) return c->fail(MCP_ERR_ARG, "prop_ld_multi: unknown grid id %d", gid);
  Grid& g = *ig->second;
  if (g.airy || (g.npts - 1) % 2 != 0 || g.npts < 3 || g.dr == 0.0)
    return c->fail(MCP_ERR_ARG, "prop_ld_multi: needs an LD grid (2*NSTEP+1 points and its DR)");
  const int nstep = (g.npts - 1) / 2;
  const double h = g.dr / 2.0;
  std::vector<Block*> bl(nblk);
  size_t nitems = 0, ucsum = 0, ysum = 0, ldsmax = 0;
  for (int b = 0; b < nblk; ++b) {
    auto ib = c->blocks.find(block_ids[b]);
    if (ib == c->blocks.end()) return c->fail(MCP_ERR_ARG, "prop_ld_multi: unknown block id %d", block_ids[b]);
    Block& B = *ib->second;
    if (B.mxlam != g.mxlam) return c->fail(MCP_ERR_ARG, "prop_ld_multi: block %d has %d radial terms, grid %d", block_ids[b], B.mxlam, g.mxlam);
    if (B.nconst != 0) return c->fail(MCP_ERR_UNSUPPORTED, "prop_ld_multi: blocks with NCONST terms go through mcp_prop_ld");
    if (B.n < 2) return c->fail(MCP_ERR_UNSUPPORTED, "prop_ld_multi: N=1 blocks take ODPROP (mcp_prop_ld)");
    const size_t need = std::max(lds_md(B.n, B.nham), lds_uc(B.n, B.nham));
    if (need > c->lds_max) return c->fail(MCP_ERR_UNSUPPORTED, "prop_ld_multi: N=%d beyond the LDS-resident MDPROP", B.n);
    if (ninst[b] < 0) return c->fail(MCP_ERR_ARG, "prop_ld_multi: ninst[%d] < 0", b);
    ldsmax = std::max(ldsmax, lds_md(B.n, B.nham));
    bl[b] = &B;
    nitems += ninst[b];
    ucsum += (size_t)nstep * B.n * B.n;
    ysum += (size_t)ninst[b] * B.n * B.n;
  }
  if (nitems == 0) return MCP_OK;
  HIPCHK(c, hipSetDevice(c->device));
  // device buffers (allocated before any capture; a graph is only replayed
  // while the buffers it was captured on are still these)
  HIPCHK(c, c->s_qu.ensure(ucsum * sizeof(double)));
  HIPCHK(c, c->s_qk.ensure((size_t)nblk * nstep * sizeof(int)));
  HIPCHK(c, c->s_qb.ensure((size_t)nblk * sizeof(mcp::QueueBlock)));
  HIPCHK(c, c->s_qi.ensure(nitems * sizeof(int2)));
  HIPCHK(c, c->s_qe.ensure(nitems * sizeof(double)));
  HIPCHK(c, c->s_qy.ensure(ysum * sizeof(double)));
  HIPCHK(c, c->s_qn.ensure(2 * nitems * sizeof(int)));
  HIPCHK(c, c->s_qh.ensure(sizeof(unsigned)));
  // block table and the item list, longest (N^3) first
  std::vector<mcp::QueueBlock> qb(nblk);
  std::vector<std::pair<long long, int2>> its;
  its.reserve(nitems);
  {
    size_t uo = 0, yo = 0, io = 0;
    for (int b = 0; b < nblk; ++b) {
      const int n = bl[b]->n;
      qb[b].B = block_args(*bl[b]);
      qb[b].uc = c->s_qu.as<double>() + uo;
      qb[b].uc_kount = c->s_qk.as<int>() + (size_t)b * nstep;
      qb[b].ered = c->s_qe.as<double>() + io;
      qb[b].y = c->s_qy.as<double>() + yo;
      qb[b].nodes = c->s_qn.as<int>() + io;
      qb[b].status = c->s_qn.as<int>() + nitems + io;
      qb[b].bk = n > 30 ? 1 : 0;
      for (int i = 0; i < ninst[b]; ++i) its.push_back({-(long long)n * n * n, make_int2(b, i)});
      uo += (size_t)nstep * n * n;
      yo += (size_t)ninst[b] * n * n;
      io += ninst[b];
    }
  }
  std::stable_sort(its.begin(), its.end(), [](const auto& a, const auto& b2) { return a.first < b2.first; });
  std::vector<int2> items(nitems);
  for (size_t i = 0; i < nitems; ++i) items[i] = its[i].second;
  int rc;
  if ((rc = upload(c, c->s_qe, ered, nitems))) return rc;
  if ((rc = upload(c, c->s_qy, y, ysum))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->s_qb.p, qb.data(), nblk * sizeof(mcp::QueueBlock), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->s_qi.p, items.data(), nitems * sizeof(int2), hipMemcpyHostToDevice, c->stream));
  // resident workgroups of the persistent launch
  const int nt = c->lds_threads == 64 ? 64 : kNT;
  const void* qk = nt == 64 ? (const void*)mcp::k_mdprop_lds_queue<64> : (const void*)mcp::k_mdprop_lds_queue<kNT>;
  int per_cu = 0;
  HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, qk, nt, ldsmax));
  const int nwg = (int)std::min<size_t>(nitems, (size_t)std::max(1, per_cu) * c->num_cu);
  mcp::GridArgs G = grid_args(g);
  const int nitems_i = (int)nitems;
  auto issue = [&]() -> int {
    for (int b = 0; b < nblk; ++b) {
      mcp::BlockArgs B = qb[b].B;
      hipLaunchKernelGGL(mcp::k_mdprop_uc, dim3(nstep), dim3(kNT), lds_uc(B.n, B.nham), c->stream, B, G,
                         (const double*)nullptr, 1, nstep, h, qb[b].bk, const_cast<double*>(qb[b].uc),
                         const_cast<int*>(qb[b].uc_kount));
      HIPCHK(c, hipGetLastError());
    }
    hipLaunchKernelGGL(mcp::k_zero_u32, dim3(1), dim3(64), 0, c->stream, c->s_qh.as<unsigned>(), 1);
    HIPCHK(c, hipGetLastError());
    if (nt == 64)
      hipLaunchKernelGGL(mcp::k_mdprop_lds_queue<64>, dim3(nwg), dim3(64), ldsmax, c->stream,
                         c->s_qb.as<const mcp::QueueBlock>(), c->s_qi.as<const int2>(), nitems_i,
                         c->s_qh.as<unsigned>(), G, nstep, h);
    else
      hipLaunchKernelGGL(mcp::k_mdprop_lds_queue<kNT>, dim3(nwg), dim3(kNT), ldsmax, c->stream,
                         c->s_qb.as<const mcp::QueueBlock>(), c->s_qi.as<const int2>(), nitems_i,
                         c->s_qh.as<unsigned>(), G, nstep, h);
    HIPCHK(c, hipGetLastError());
    return MCP_OK;
  };
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  if (c->use_graph) {
    // one graph per radial grid and block set: the sector matrices of every
    // block and the persistent launch, replayed while the key and the
    // buffers stay the same
    std::vector<int> key = {gid, nblk, nt};
    for (int b = 0; b < nblk; ++b) {
      key.push_back(block_ids[b]);
      key.push_back(ninst[b]);
    }
    std::vector<void*> bufs = {c->s_qu.p, c->s_qk.p, c->s_qb.p, c->s_qi.p, c->s_qe.p, c->s_qy.p, c->s_qn.p, c->s_qh.p,
                               g.r.p, g.ptab.p};
    for (int b = 0; b < nblk; ++b) {
      bufs.push_back(bl[b]->vl.p);
      bufs.push_back(bl[b]->eint.p);
      bufs.push_back(bl[b]->cent.p);
    }
    auto& mg = c->mgraph;
    if (!mg.exec || mg.key != key || mg.bufs != bufs) {
      if (mg.exec) (void)hipGraphExecDestroy(mg.exec);
      if (mg.graph) (void)hipGraphDestroy(mg.graph);
      mg.exec = nullptr;
      mg.graph = nullptr;
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
      const int rci = issue();
      hipGraph_t gr = nullptr;
      const hipError_t ec = hipStreamEndCapture(c->stream, &gr);
      if (rci) return rci;
      HIPCHK(c, ec);
      mg.graph = gr;
      HIPCHK(c, hipGraphInstantiate(&mg.exec, mg.graph, nullptr, nullptr, 0));
      mg.key = key;
      mg.bufs = bufs;
    }
    HIPCHK(c, hipGraphLaunch(mg.exec, c->stream));
  } else {
    if ((rc = issue())) return rc;
  }
  HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  if ((rc = download(c, y, c->s_qy, ysum))) return rc;
  std::vector<int> ns(2 * nitems);
  HIPCHK(c, hipMemcpyAsync(ns.data(), c->s_qn.p, 2 * nitems * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < nitems; ++i) {
    nodes[i] = ns[i];
    status[i] = ns[nitems + i];
  }
  return record_timing(c);
}

int mcp_prop_airy(mcp_ctx* c, int32_t bid, int32_t gid, int32_t ninst, const double* ered, double* y, int32_t* nodes,
                  int32_t* status) {
  if (!c || ninst < 0 || !ered || !y || !nodes || !status)
    return c ? c->fail(MCP_ERR_ARG, "prop_airy: bad args") : MCP_ERR_ARG;
  Block* b;
  Grid* g;
  int rc;
  if ((rc = lookup(c, bid, gid, &b, &g))) return rc;
  if (ninst == 0) return MCP_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t cnt = (size_t)ninst * b->n * b->n;
  if ((rc = upload(c, c->s_g, ered, (size_t)ninst))) return rc;
  if ((rc = upload(c, c->s_i, y, cnt))) return rc;
  HIPCHK(c, c->s_j.ensure((size_t)ninst * 2 * sizeof(int)));
  int* dn = c->s_j.as<int>();
  if ((rc = airy_prop_device(c, *b, *g, ninst, c->s_g.as<double>(), c->s_i.as<double>(), dn, dn + ninst))) return rc;
  if ((rc = download(c, y, c->s_i, cnt))) return rc;
  HIPCHK(c, hipMemcpyAsync(nodes, dn, ninst * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(status, dn + ninst, ninst * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return record_timing(c);
}

int mcp_grid_destroy(mcp_ctx* c, int32_t id) {
  if (!c) return MCP_ERR_ARG;
  (void)hipStreamSynchronize(c->stream);
  return c->grids.erase(id) ? MCP_OK : c->fail(MCP_ERR_ARG, "unknown grid id %d", id);
}

int mcp_hammat(mcp_ctx* c, int32_t bid, int32_t gid, int32_t nr, const int32_t* idx, const double* pconst,
               double ered, double* w) {
  if (!c || nr < 0 || !idx || !w) return c ? c->fail(MCP_ERR_ARG, "hammat: bad args") : MCP_ERR_ARG;
  Block* b;
  Grid* g;
  int rc;
  if ((rc = lookup(c, bid, gid, &b, &g))) return rc;
  if (nr == 0) return MCP_OK;
  if (b->nconst > 0 && !pconst) return c->fail(MCP_ERR_ARG, "hammat: block has NCONST=%d but pconst is NULL", b->nconst);
  for (int m = 0; m < nr; ++m)
    if (idx[m] < 0 || idx[m] >= g->npts) return c->fail(MCP_ERR_ARG, "hammat: grid index %d out of range", idx[m]);
  HIPCHK(c, hipSetDevice(c->device));
  const int n = b->n;
  if ((rc = upload(c, c->s_a, idx, (size_t)nr))) return rc;
  if (b->nconst > 0 && (rc = upload(c, c->s_d, pconst, (size_t)b->nconst))) return rc;
  HIPCHK(c, c->s_b.ensure((size_t)nr * n * n * sizeof(double)));
  const size_t ntri = (size_t)n * (n + 1) / 2;
  const int gx = (int)std::min<size_t>((ntri + 255) / 256, 1024);
  hipLaunchKernelGGL(mcp::k_hammat, dim3(gx, nr), dim3(256), 0, c->stream, block_args(*b), grid_args(*g),
                     c->s_a.as<int>(), nr, b->nconst > 0 ? c->s_d.as<double>() : nullptr, ered, c->s_b.as<double>());
  HIPCHK(c, hipGetLastError());
  if ((rc = download(c, w, c->s_b, (size_t)nr * n * n))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MCP_OK;
}

// YINIT on the device for ninst instances into I.y (LDS eigen path, or the
// large-N square root).  d_status (ninst ints) receives 1 for an instance
// whose start matrix is not positive definite on the large path.
static int yinit_device(mcp_ctx* c, Block& b, Grid& g, mcp::InstArgs I, int ninst, int k0, int outward, double bcyo,
                        int* d_status) {
  const int n = b.n;
  const size_t lds = lds_yinit(n, b.nham);
  if (lds <= c->lds_max) {
    hipLaunchKernelGGL(mcp::k_yinit, dim3(ninst), dim3(kNT), lds, c->stream, block_args(b), grid_args(g), I, k0,
                       outward, bcyo, 1);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemsetAsync(d_status, 0, ninst * sizeof(int), c->stream));
    return MCP_OK;
  }
  const size_t ll = lds_large(n, b.nham, 0);
  if (ll > c->lds_max) return c->fail(MCP_ERR_UNSUPPORTED, "yinit: N=%d exceeds the large-N panel limit", n);
  HIPCHK(c, c->s_y.ensure((size_t)ninst * 3 * n * n * sizeof(double)));
  I.status = d_status;
  const int nblk = (n + mcp::kPB - 1) / mcp::kPB;
  if (!c->no_reg && team_inverse_fits(c, n) && (size_t)2 * ninst * nblk <= (size_t)c->num_cu) {
    // the square-root iteration of k_yinit_large in stages, both inverses
    // of an iteration on teams of workgroups
    const size_t nn = (size_t)n * n;
    double* work = c->s_y.as<double>();
    double* zs = work + 2 * nn * ninst;
    HIPCHK(c, c->s_yv.ensure((size_t)ninst * (sizeof(double) + 4 * sizeof(int))));
    double* scale = c->s_yv.as<double>();
    int* done = (int*)(scale + ninst);
    int* kc = done + ninst;
    hipLaunchKernelGGL(mcp::k_yl_start, dim3(ninst), dim3(kNT), 0, c->stream, block_args(b), grid_args(g), I, k0, zs,
                       scale, d_status);
    HIPCHK(c, hipMemsetAsync(done, 0, ninst * sizeof(int), c->stream));
    std::vector<int> hd(2 * (size_t)ninst);
    int rc;
    for (int it = 0; it < 40; ++it) {
      hipLaunchKernelGGL(mcp::k_yl_copy, dim3(nblk256(ninst * nn)), dim3(256), 0, c->stream, n, ninst, I.y, zs, work);
      HIPCHK(c, hipGetLastError());
      if ((rc = team_inverse(c, n, 2 * ninst, work, kc, 1))) return rc;
      hipLaunchKernelGGL(mcp::k_yl_update, dim3(ninst), dim3(kNT), 0, c->stream, n, it, I.y, zs, work, kc, done,
                         d_status);
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipMemcpyAsync(hd.data(), done, ninst * sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipMemcpyAsync(hd.data() + ninst, d_status, ninst * sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      bool all = true;
      for (int i = 0; i < ninst; ++i) all = all && (hd[i] || hd[ninst + i]);
      if (all) break;
    }
    hipLaunchKernelGGL(mcp::k_yl_finish, dim3(nblk256(ninst * nn)), dim3(256), 0, c->stream, n, ninst, outward, scale,
                       I.y);
    HIPCHK(c, hipGetLastError());
    return MCP_OK;
  }
  hipLaunchKernelGGL(mcp::k_yinit_large, dim3(ninst), dim3(kNT), ll, c->stream, block_args(b), grid_args(g), I, k0,
                     outward, c->s_y.as<double>());
  HIPCHK(c, hipGetLastError());
  return MCP_OK;
}

static int upload_inst(mcp_ctx* c, const Block& b, int ninst, const double* ered, const double* pconst) {
  int rc;
  if ((rc = upload(c, c->s_g, ered, (size_t)ninst))) return rc;
  if (b.nconst > 0) {
    if (!pconst) return c->fail(MCP_ERR_ARG, "block has NCONST=%d but pconst is NULL", b.nconst);
    if ((rc = upload(c, c->s_h, pconst, (size_t)ninst * b.nconst))) return rc;
  }
  return MCP_OK;
}

int mcp_yinit(mcp_ctx* c, int32_t bid, int32_t gid, int32_t k0, int32_t outward, double bcyo, int32_t ninst,
              const double* ered, const double* pconst, double* y) {
  if (!c || ninst < 0 || !ered || !y) return c ? c->fail(MCP_ERR_ARG, "yinit: bad args") : MCP_ERR_ARG;
  Block* b;
  Grid* g;
  int rc;
  if ((rc = lookup(c, bid, gid, &b, &g))) return rc;
  if (k0 < 0 || k0 >= g->npts) return c->fail(MCP_ERR_ARG, "yinit: grid index %d out of range", k0);
  if (ninst == 0) return MCP_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const int n = b->n;
  if ((rc = upload_inst(c, *b, ninst, ered, pconst))) return rc;
  HIPCHK(c, c->s_i.ensure((size_t)ninst * n * n * sizeof(double)));
  HIPCHK(c, c->s_j.ensure((size_t)ninst * 2 * sizeof(int)));
  mcp::InstArgs I;
  I.ered = c->s_g.as<double>();
  I.pconst = b->nconst > 0 ? c->s_h.as<double>() : nullptr;
  I.y = c->s_i.as<double>();
  I.nodes = nullptr;
  I.status = nullptr;
  if ((rc = yinit_device(c, *b, *g, I, ninst, k0, outward, bcyo, c->s_j.as<int>()))) return rc;
  if ((rc = download(c, y, c->s_i, (size_t)ninst * n * n))) return rc;
  std::vector<int> st(ninst);
  HIPCHK(c, hipMemcpyAsync(st.data(), c->s_j.p, ninst * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < ninst; ++i)
    if (st[i]) return c->fail(MCP_ERR_UNSUPPORTED, "yinit: locally open channel at the start radius (instance %d)", i);
  return MCP_OK;
}

int mcp_prop_ld(mcp_ctx* c, int32_t bid, int32_t gid, int32_t iprop, int32_t ninst, const double* ered,
                const double* pconst, double* y, int32_t* nodes, int32_t* status) {
  if (!c || ninst < 0 || !ered || !y || !nodes || !status)
    return c ? c->fail(MCP_ERR_ARG, "prop_ld: bad args") : MCP_ERR_ARG;
  Block* b;
  Grid* g;
  int rc;
  if ((rc = lookup(c, bid, gid, &b, &g))) return rc;
  if (ninst == 0) return MCP_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const int n = b->n;
  if ((rc = upload_inst(c, *b, ninst, ered, pconst))) return rc;
  if ((rc = upload(c, c->s_i, y, (size_t)ninst * n * n))) return rc;
  HIPCHK(c, c->s_j.ensure((size_t)ninst * 2 * sizeof(int)));
  int* dn = c->s_j.as<int>();
  if ((rc = prop_ld_device(c, *b, *g, iprop, ninst, c->s_g.as<double>(), b->nconst > 0 ? c->s_h.as<double>() : nullptr,
                           c->s_i.as<double>(), dn, dn + ninst)))
    return rc;
  if ((rc = download(c, y, c->s_i, (size_t)ninst * n * n))) return rc;
  std::vector<int> tmp(2 * (size_t)ninst);
  HIPCHK(c, hipMemcpyAsync(tmp.data(), dn, 2 * ninst * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < ninst; ++i) {
    nodes[i] = tmp[i];
    status[i] = tmp[ninst + i];
  }
  return record_timing(c);
}

// YTOK + KTOS for ninst instances whose MatchArgs point at device memory:
// the setup pass, the forward elimination on workgroup teams (counters:
// ninst + 1 zeroed words, the last the abort flag), the finishing pass.
// Allocates nothing.
static int match_run(mcp_ctx* c, int n, int ninst, const mcp::MatchArgs& M, unsigned* ctrs) {
  hipLaunchKernelGGL(mcp::k_match, dim3(ninst), dim3(kNT), 0, c->stream, M, mcp::kMatchSetup);
  HIPCHK(c, hipGetLastError());
  {
    // the forward elimination on teams of workgroups: g ranks per instance
    // where the CUs hold them (a single workgroup for many instances); a
    // team size set on the context caps it
    const int gmax = c->team > 0 ? c->team : 32;
    const int g = std::max(1, std::min({gmax, (n + 7) / 8, c->num_cu / std::max(1, ninst)}));
    const size_t lu_lds = (size_t)5 * n * sizeof(double);
    if (lu_lds > c->lds_max) return c->fail(MCP_ERR_UNSUPPORTED, "matching: N=%d exceeds the LU's LDS", n);
    int rc;
    hipLaunchKernelGGL(mcp::k_zero_u32, dim3((ninst + 1 + 255) / 256), dim3(256), 0, c->stream, ctrs, ninst + 1);
    HIPCHK(c, hipGetLastError());
    int* abort_flag = (int*)(ctrs + ninst);
    mcp::MatchArgs Mc = M;
    if (g == 1) {
      hipLaunchKernelGGL(mcp::k_match_lu, dim3(ninst), dim3(kNT), lu_lds, c->stream, Mc, 1, ctrs, abort_flag);
      HIPCHK(c, hipGetLastError());
    } else {
      int g_ = g;
      void* args[] = {&Mc, &g_, &ctrs, &abort_flag};
      if ((rc = launch_team(c, (const void*)mcp::k_match_lu, dim3(ninst * g), dim3(kNT), args, lu_lds))) return rc;
    }
  }
  hipLaunchKernelGGL(mcp::k_match, dim3(ninst), dim3(kNT), 0, c->stream, M, mcp::kMatchFinish);
  HIPCHK(c, hipGetLastError());
  return MCP_OK;
}

static int match_device(mcp_ctx* c, int n, int ninst, const double* d_y, const double* wvec, const int32_t* nopen,
                        const int32_t* nb, const double* d_cent, double rval, int cent_stride = 0) {
  int rc;
  if ((rc = upload(c, c->s_a, wvec, (size_t)ninst * n))) return rc;
  if ((rc = upload(c, c->s_c, nopen, (size_t)ninst))) return rc;
  if ((rc = upload(c, c->s_d, nb, (size_t)ninst * n))) return rc;
  HIPCHK(c, c->s_b.ensure((size_t)ninst * (2 * (size_t)n * n + 4 * (size_t)n) * sizeof(double)));
  HIPCHK(c, c->s_k.ensure((size_t)ninst * 3 * n * n * sizeof(double) + (size_t)ninst * sizeof(int)));
  HIPCHK(c, c->s_bar.ensure((size_t)(ninst + 1) * sizeof(unsigned)));
  mcp::MatchArgs M;
  M.n = n;
  M.y = d_y;
  M.wvec = c->s_a.as<double>();
  M.nopen = c->s_c.as<int>();
  M.nb = c->s_d.as<int>();
  M.cent = d_cent;
  M.cent_stride = cent_stride;
  M.rval = rval;
  M.work = c->s_b.as<double>();
  M.kmat = c->s_k.as<double>();
  M.sr = M.kmat + (size_t)ninst * n * n;
  M.si = M.sr + (size_t)ninst * n * n;
  M.status = (int*)(M.si + (size_t)ninst * n * n);
  return match_run(c, n, ninst, M, c->s_bar.as<unsigned>());
}

static int fetch_match(mcp_ctx* c, int n, int ninst, const int32_t* nopen, double* kmat, double* sr, double* si,
                       int32_t* status) {
  std::vector<double> all((size_t)ninst * 3 * n * n);
  std::vector<int> st(ninst);
  HIPCHK(c, hipMemcpyAsync(all.data(), c->s_k.p, all.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(st.data(), (char*)c->s_k.p + all.size() * sizeof(double), ninst * sizeof(int),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t nn = (size_t)n * n;
  for (int i = 0; i < ninst; ++i) {
    const int nop = nopen[i];
    for (size_t e = 0; e < (size_t)nop * nop; ++e) {
      if (kmat) kmat[i * nn + e] = all[i * nn + e];
      sr[i * nn + e] = all[(size_t)ninst * nn + i * nn + e];
      si[i * nn + e] = all[2 * (size_t)ninst * nn + i * nn + e];
    }
    status[i] = st[i];
  }
  return MCP_OK;
}

int mcp_ytok_ktos(mcp_ctx* c, int32_t n, int32_t ninst, const double* y, const double* wvec, const int32_t* nopen,
                  const int32_t* nb, const double* cent, double rval, double* kmat, double* sr, double* si,
                  int32_t* status) {
  if (!c || n <= 0 || ninst < 0 || !y || !wvec || !nopen || !nb || !cent || !sr || !si || !status)
    return c ? c->fail(MCP_ERR_ARG, "ytok_ktos: bad args") : MCP_ERR_ARG;
  if (ninst == 0) return MCP_OK;
  for (int i = 0; i < ninst; ++i)
    if (nopen[i] < 0 || nopen[i] > n) return c->fail(MCP_ERR_ARG, "ytok_ktos: nopen[%d]=%d out of range", i, nopen[i]);
  for (size_t e = 0; e < (size_t)ninst * n; ++e)
    if (nb[e] < 0 || nb[e] >= n) return c->fail(MCP_ERR_ARG, "ytok_ktos: nb out of range");
  HIPCHK(c, hipSetDevice(c->device));
  int rc;
  if ((rc = upload(c, c->s_i, y, (size_t)ninst * n * n))) return rc;
  if ((rc = upload(c, c->s_e, cent, (size_t)n))) return rc;
  if ((rc = match_device(c, n, ninst, c->s_i.as<double>(), wvec, nopen, nb, c->s_e.as<double>(), rval))) return rc;
  return fetch_match(c, n, ninst, nopen, kmat, sr, si, status);
}

size_t mcp_ytok_ktos_noalloc_bytes(int32_t n, int32_t ninst) {
  if (n <= 0 || ninst <= 0) return 0;
  const size_t ctr = ((size_t)(ninst + 1) * sizeof(unsigned) + 255) / 256 * 256;
  return ctr + (size_t)ninst * (2 * (size_t)n * n + 4 * (size_t)n) * sizeof(double);
}

int mcp_ytok_ktos_noalloc(mcp_ctx* c, int32_t n, int32_t ninst, const double* d_y, const double* d_wvec,
                          const int32_t* d_nopen, const int32_t* d_nb, const double* d_cent, double rval,
                          double* d_kmat, double* d_sr, double* d_si, int32_t* d_status) {
  if (!c || n <= 0 || ninst < 0 || !d_y || !d_wvec || !d_nopen || !d_nb || !d_cent || !d_kmat || !d_sr || !d_si ||
      !d_status)
    return c ? c->fail(MCP_ERR_ARG, "ytok_ktos_noalloc: bad args") : MCP_ERR_ARG;
  if (!(rval > 0.0)) return c->fail(MCP_ERR_ARG, "ytok_ktos_noalloc: matching radius %g must be positive", rval);
  if (ninst == 0) return MCP_OK;
  const size_t need = mcp_ytok_ktos_noalloc_bytes(n, ninst);
  if (need > kNoAllocScratchBytes)
    return c->fail(MCP_ERR_UNSUPPORTED,
                   "ytok_ktos_noalloc: N=%d with %d instances needs %zu bytes of scratch, the maximum is %zu", n, ninst,
                   need, kNoAllocScratchBytes);
  if (!c->scratch_dev) return c->fail(MCP_ERR_HIP, "ytok_ktos_noalloc: no scratch address on this device");
  HIPCHK(c, hipSetDevice(c->device));
  unsigned* ctrs = (unsigned*)c->scratch_dev;
  const size_t ctr = ((size_t)(ninst + 1) * sizeof(unsigned) + 255) / 256 * 256;
  mcp::MatchArgs M;
  M.n = n;
  M.y = d_y;
  M.wvec = d_wvec;
  M.nopen = d_nopen;
  M.nb = d_nb;
  M.cent = d_cent;
  M.cent_stride = 0;
  M.rval = rval;
  M.work = (double*)(c->scratch_dev + ctr);
  M.kmat = d_kmat;
  M.sr = d_sr;
  M.si = d_si;
  M.status = d_status;
  int rc;
  if ((rc = match_run(c, n, ninst, M, ctrs))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MCP_OK;
}

int mcp_smatrix_ld(mcp_ctx* c, int32_t bid, int32_t gid, int32_t iprop, int32_t ninst, const double* ered,
                   const double* pconst, const double* wvec, const int32_t* nopen, const int32_t* nb, double rval,
                   double* kmat, double* sr, double* si, int32_t* nodes, int32_t* status) {
  if (!c || ninst < 0 || !ered || !wvec || !nopen || !nb || !sr || !si || !nodes || !status)
    return c ? c->fail(MCP_ERR_ARG, "smatrix_ld: bad args") : MCP_ERR_ARG;
  Block* b;
  Grid* g;
  int rc;
  if ((rc = lookup(c, bid, gid, &b, &g))) return rc;
  if (ninst == 0) return MCP_OK;
  const int n = b->n;
  for (int i = 0; i < ninst; ++i)
    if (nopen[i] < 0 || nopen[i] > n) return c->fail(MCP_ERR_ARG, "smatrix_ld: nopen[%d]=%d out of range", i, nopen[i]);
  for (size_t e = 0; e < (size_t)ninst * n; ++e)
    if (nb[e] < 0 || nb[e] >= n) return c->fail(MCP_ERR_ARG, "smatrix_ld: nb out of range");
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = upload_inst(c, *b, ninst, ered, pconst))) return rc;
  HIPCHK(c, c->s_i.ensure((size_t)ninst * n * n * sizeof(double)));
  HIPCHK(c, c->s_j.ensure((size_t)ninst * 3 * sizeof(int)));
  mcp::InstArgs I;
  I.ered = c->s_g.as<double>();
  I.pconst = b->nconst > 0 ? c->s_h.as<double>() : nullptr;
  I.y = c->s_i.as<double>();
  I.nodes = nullptr;
  I.status = nullptr;
  int* dn = c->s_j.as<int>();
  if ((rc = yinit_device(c, *b, *g, I, ninst, 0, 1, 0.0, dn + 2 * ninst))) return rc;
  // SCCTRL runs ODPROP for a single channel (source_code/scctrl.f:183-196)
  const int ip = (iprop == 6 && n == 1) ? MCP_PROP_ODPROP : iprop;
  if ((rc = prop_ld_device(c, *b, *g, ip, ninst, I.ered, I.pconst, I.y, dn, dn + ninst))) return rc;
  if (!(rval > 0.0)) return c->fail(MCP_ERR_ARG, "smatrix_ld: matching radius %g must be positive", rval);
  if (b->nconst > 0 || b->nrsq > 0) {
    // YTRANS before YTOK (source_code/scctrl.f:274-287)
    HIPCHK(c, c->s_ye.ensure((size_t)ninst * n * sizeof(double)));
    if ((rc = ytrans_device(c, *b, ninst, I.pconst, I.y, c->s_ye.as<double>()))) return rc;
  }
  // with the centrifugal operator, YTOK's CENT are YTRANS's, per instance
  const double* dcent = b->nrsq > 0 ? c->s_yc.as<double>() : b->cent.as<double>();
  if ((rc = match_device(c, n, ninst, I.y, wvec, nopen, nb, dcent, rval, b->nrsq > 0 ? n : 0))) return rc;
  std::vector<int> tmp(3 * (size_t)ninst);
  HIPCHK(c, hipMemcpyAsync(tmp.data(), dn, 3 * ninst * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if ((rc = fetch_match(c, n, ninst, nopen, kmat, sr, si, status))) return rc;
  for (int i = 0; i < ninst; ++i) {
    nodes[i] = tmp[i];
    if (tmp[ninst + i] != 0) status[i] = -tmp[ninst + i];  // propagation failure: -(step)
    if (tmp[2 * ninst + i] != 0) status[i] = -1000000;       // open channel at the start radius
  }
  return record_timing(c);
}

int mcp_smatrix_segments(mcp_ctx* c, int32_t bid, int32_t nseg, const int32_t* gids, const int32_t* iprops,
                         int32_t ninst, const double* ered, const double* pconst, const double* wvec,
                         const int32_t* nopen, const int32_t* nb, double rval, double* kmat, double* sr, double* si,
                         int32_t* nodes, int32_t* status) {
  if (!c || nseg <= 0 || nseg > 8 || !gids || !iprops || ninst < 0 || !ered || !wvec || !nopen || !nb || !sr || !si ||
      !nodes || !status)
    return c ? c->fail(MCP_ERR_ARG, "smatrix_segments: bad args") : MCP_ERR_ARG;
  Block* b;
  Grid* g[8];
  int rc;
  for (int k = 0; k < nseg; ++k)
    if ((rc = lookup(c, bid, gids[k], &b, &g[k]))) return rc;
  if (ninst == 0) return MCP_OK;
  const int n = b->n;
  for (int i = 0; i < ninst; ++i)
    if (nopen[i] < 0 || nopen[i] > n) return c->fail(MCP_ERR_ARG, "smatrix_segments: nopen[%d] out of range", i);
  for (size_t e = 0; e < (size_t)ninst * n; ++e)
    if (nb[e] < 0 || nb[e] >= n) return c->fail(MCP_ERR_ARG, "smatrix_segments: nb out of range");
  if (!(rval > 0.0)) return c->fail(MCP_ERR_ARG, "smatrix_segments: matching radius %g must be positive", rval);
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = upload_inst(c, *b, ninst, ered, pconst))) return rc;
  HIPCHK(c, c->s_i.ensure((size_t)ninst * n * n * sizeof(double)));
  HIPCHK(c, c->s_j.ensure((size_t)ninst * (2 * nseg + 1) * sizeof(int)));
  mcp::InstArgs I;
  I.ered = c->s_g.as<double>();
  I.pconst = b->nconst > 0 ? c->s_h.as<double>() : nullptr;
  I.y = c->s_i.as<double>();
  I.nodes = nullptr;
  I.status = nullptr;
  int* dn = c->s_j.as<int>();  // [yinit | nodes, status per segment]
  // YINIT on the first segment only (source_code/scctrl.f:116-121)
  if ((rc = yinit_device(c, *b, *g[0], I, ninst, 0, 1, 0.0, dn))) return rc;
  c->seg_timed = 0;
  for (int k = 0; k < nseg; ++k) {
    int* sn = dn + ninst * (1 + 2 * k);
    if (iprops[k] == 9) {
      if ((rc = airy_prop_device(c, *b, *g[k], ninst, I.ered, I.y, sn, sn + ninst))) return rc;
    } else {
      const int ip = (iprops[k] == 6 && n == 1) ? MCP_PROP_ODPROP : iprops[k];
      if ((rc = prop_ld_device(c, *b, *g[k], ip, ninst, I.ered, I.pconst, I.y, sn, sn + ninst))) return rc;
    }
    if ((rc = record_timing(c))) return rc;
    c->seg_ms[k][0] = c->t_sector;
    c->seg_ms[k][1] = c->t_prop;
    c->seg_timed = k + 1;
  }
  if (b->nconst > 0 || b->nrsq > 0) {
    HIPCHK(c, c->s_ye.ensure((size_t)ninst * n * sizeof(double)));
    if ((rc = ytrans_device(c, *b, ninst, I.pconst, I.y, c->s_ye.as<double>()))) return rc;
  }
  const double* dcent = b->nrsq > 0 ? c->s_yc.as<double>() : b->cent.as<double>();
  if ((rc = match_device(c, n, ninst, I.y, wvec, nopen, nb, dcent, rval, b->nrsq > 0 ? n : 0))) return rc;
  std::vector<int> tmp((size_t)ninst * (2 * nseg + 1));
  HIPCHK(c, hipMemcpyAsync(tmp.data(), dn, tmp.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if ((rc = fetch_match(c, n, ninst, nopen, kmat, sr, si, status))) return rc;
  for (int i = 0; i < ninst; ++i) {
    int nd = 0;
    for (int k = nseg - 1; k >= 0; --k) {
      nd += tmp[(size_t)ninst * (1 + 2 * k) + i];
      const int st = tmp[(size_t)ninst * (2 + 2 * k) + i];
      if (st != 0) status[i] = -(100000 * k + st);  // segment k (0-based) failed at step st
    }
    nodes[i] = nd;
    if (tmp[i] != 0) status[i] = -1000000;
  }
  return MCP_OK;
}

int mcp_bound_match(mcp_ctx* c, int32_t bid, int32_t gid_out, int32_t gid_in, int32_t iprop_out, int32_t iprop_in,
                    int32_t ninst, const double* ered, const double* pconst, int32_t* nodec, double* eigmin,
                    int32_t* imin, double* psimch, double* evals, int32_t* status) {
  if (!c || ninst < 0 || !ered || !nodec || !eigmin || !imin || !status)
    return c ? c->fail(MCP_ERR_ARG, "bound_match: bad args") : MCP_ERR_ARG;
  Block* b;
  Grid* go;
  Grid* gi;
  int rc;
  if ((rc = lookup(c, bid, gid_out, &b, &go))) return rc;
  if ((rc = lookup(c, bid, gid_in, &b, &gi))) return rc;
  if (ninst == 0) return MCP_OK;
  const int n = b->n;
  if (lds_eig(n) > c->lds_max)
    return c->fail(MCP_ERR_UNSUPPORTED, "bound_match: N=%d exceeds the LDS eigen-solver limit", n);
  if (!(go->dr > 0.0) || !(gi->dr < 0.0))
    return c->fail(MCP_ERR_ARG, "bound_match: the outward grid needs DR > 0 and the inward grid DR < 0");
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = upload_inst(c, *b, ninst, ered, pconst))) return rc;
  const size_t nn = (size_t)ninst * n * n;
  HIPCHK(c, c->s_i.ensure(nn * sizeof(double)));
  HIPCHK(c, c->s_k.ensure(nn * sizeof(double)));
  HIPCHK(c, c->s_j.ensure((size_t)ninst * 9 * sizeof(int)));
  HIPCHK(c, c->s_d.ensure((size_t)ninst * (2 + 2 * (size_t)n) * sizeof(double)));
  // sector-matrix workspace for the longer grid first: no reallocation between the two propagations
  const int nstep_max = std::max(go->npts, gi->npts) / 2;
  HIPCHK(c, c->s_e.ensure((size_t)nstep_max * n * n * sizeof(double)));
  HIPCHK(c, c->s_f.ensure((size_t)nstep_max * sizeof(int)));
  mcp::InstArgs I;
  I.ered = c->s_g.as<double>();
  I.pconst = b->nconst > 0 ? c->s_h.as<double>() : nullptr;
  I.nodes = nullptr;
  I.status = nullptr;
  int* dn = c->s_j.as<int>();  // [nodes_out | status_out | yinit_out | nodes_in | status_in | yinit_in | nodec | imin]
  double* yo = c->s_i.as<double>();
  double* yi = c->s_k.as<double>();
  I.y = yo;
  if ((rc = yinit_device(c, *b, *go, I, ninst, 0, 1, 0.0, dn + 2 * ninst))) return rc;
  if ((rc = prop_ld_device(c, *b, *go, iprop_out, ninst, I.ered, I.pconst, yo, dn, dn + ninst))) return rc;
  I.y = yi;
  if ((rc = yinit_device(c, *b, *gi, I, ninst, 0, 0, 0.0, dn + 5 * ninst))) return rc;
  if ((rc = prop_ld_device(c, *b, *gi, iprop_in, ninst, I.ered, I.pconst, yi, dn + 3 * ninst, dn + 4 * ninst)))
    return rc;
  double* de = c->s_d.as<double>();  // [eigmin | psimch | evals]
  hipLaunchKernelGGL(mcp::k_bound_match, dim3(ninst), dim3(kNT), lds_eig(n), c->stream, n, yo, yi, dn, dn + 3 * ninst,
                     dn + 6 * ninst, de, dn + 7 * ninst, de + ninst, de + ninst + (size_t)ninst * n);
  HIPCHK(c, hipGetLastError());
  std::vector<int> tmp(8 * (size_t)ninst);
  HIPCHK(c, hipMemcpyAsync(tmp.data(), dn, 8 * ninst * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(eigmin, de, ninst * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (psimch)
    HIPCHK(c, hipMemcpyAsync(psimch, de + ninst, (size_t)ninst * n * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
  if (evals)
    HIPCHK(c, hipMemcpyAsync(evals, de + ninst + (size_t)ninst * n, (size_t)ninst * n * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < ninst; ++i) {
    nodec[i] = tmp[6 * ninst + i];
    imin[i] = tmp[7 * ninst + i];
    status[i] = 0;
    if (tmp[ninst + i] != 0) status[i] = -tmp[ninst + i];                // outward inversion failure at step
    if (tmp[4 * ninst + i] != 0) status[i] = -(100000 + tmp[4 * ninst + i]);  // inward
    if (tmp[2 * ninst + i] != 0 || tmp[5 * ninst + i] != 0) status[i] = -1000000;
  }
  return MCP_OK;
}

int mcp_bound_wave(mcp_ctx* c, int32_t bid, int32_t gid_out, int32_t gid_in, double ered, const double* pconst,
                   int32_t* nodec, double* eigmin, double* psimch, double* psi_out, double* psi_in,
                   int32_t* status) {
  if (!c || !nodec || !eigmin || !psimch || !psi_out || !psi_in || !status)
    return c ? c->fail(MCP_ERR_ARG, "bound_wave: bad args") : MCP_ERR_ARG;
  Block* b;
  Grid* go;
  Grid* gi;
  int rc;
  if ((rc = lookup(c, bid, gid_out, &b, &go))) return rc;
  if ((rc = lookup(c, bid, gid_in, &b, &gi))) return rc;
  const int n = b->n;
  const size_t nn = (size_t)n * n;
  if (lds_eig(n) > c->lds_max || lds_md(n, b->nham) > c->lds_max || lds_uc(n, b->nham) > c->lds_max)
    return c->fail(MCP_ERR_UNSUPPORTED, "bound_wave: N=%d exceeds the LDS-resident MDPROP limit", n);
  if (!(go->dr > 0.0) || !(gi->dr < 0.0))
    return c->fail(MCP_ERR_ARG, "bound_wave: the outward grid needs DR > 0 and the inward grid DR < 0");
  if ((go->npts - 1) % 2 != 0 || (gi->npts - 1) % 2 != 0 || go->npts < 3 || gi->npts < 3)
    return c->fail(MCP_ERR_ARG, "bound_wave: LD grids need 2*NSTEP+1 points");
  const int nout = (go->npts - 1) / 2, nin = (gi->npts - 1) / 2;
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = upload_inst(c, *b, 1, &ered, pconst))) return rc;
  HIPCHK(c, c->s_i.ensure(nn * sizeof(double)));
  HIPCHK(c, c->s_k.ensure(nn * sizeof(double)));
  HIPCHK(c, c->s_j.ensure(9 * sizeof(int)));
  HIPCHK(c, c->s_d.ensure((2 + 2 * (size_t)n) * sizeof(double)));
  HIPCHK(c, c->s_e.ensure((size_t)std::max(nout, nin) * nn * sizeof(double)));
  HIPCHK(c, c->s_f.ensure((size_t)std::max(nout, nin) * sizeof(int)));
  // the records of both propagations, the first inverse of a step, psi
  HIPCHK(c, c->s_a.ensure(((size_t)nout + nin + 1) * nn * sizeof(double)));
  HIPCHK(c, c->s_b.ensure(((size_t)nout + nin + 2) * n * sizeof(double)));
  double* rec_out = c->s_a.as<double>();
  double* rec_in = rec_out + (size_t)nout * nn;
  double* wtmp = rec_in + (size_t)nin * nn;
  double* d_pout = c->s_b.as<double>();
  double* d_pin = d_pout + ((size_t)nout + 1) * n;
  mcp::InstArgs I;
  I.ered = c->s_g.as<double>();
  I.pconst = b->nconst > 0 ? c->s_h.as<double>() : nullptr;
  int* dn = c->s_j.as<int>();  // [nodes_out, status_out, yinit_out, nodes_in, status_in, yinit_in, nodec, imin]
  double* yo = c->s_i.as<double>();
  double* yi = c->s_k.as<double>();
  const mcp::BlockArgs B = block_args(*b);
  // one direction: YINIT, U(C) of every step, MDPROP writing its records
  auto run = [&](Grid& g, int outward, double* y, int* d_nodes, int* d_status, int* d_yst, double* recs) -> int {
    mcp::InstArgs Ii = I;
    Ii.y = y;
    Ii.nodes = d_nodes;
    Ii.status = d_status;
    int r2 = yinit_device(c, *b, g, Ii, 1, 0, outward, 0.0, d_yst);
    if (r2) return r2;
    const int nstep = (g.npts - 1) / 2;
    const double h = g.dr / 2.0;
    const int bk = n > 30 ? 1 : 0;
    const mcp::GridArgs G = grid_args(g);
    hipLaunchKernelGGL(mcp::k_mdprop_uc, dim3(nstep), dim3(kNT), lds_uc(n, b->nham), c->stream, B, G, Ii.pconst, 1,
                       nstep, h, bk, c->s_e.as<double>(), c->s_f.as<int>());
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(mcp::k_mdprop_lds_wave, dim3(1), dim3(kNT), lds_md(n, b->nham), c->stream, B, G, Ii, nstep, h,
                       c->s_e.as<double>(), c->s_f.as<int>(), bk, recs, wtmp);
    HIPCHK(c, hipGetLastError());
    return MCP_OK;
  };
  if ((rc = run(*go, 1, yo, dn, dn + 1, dn + 2, rec_out))) return rc;
  if ((rc = run(*gi, 0, yi, dn + 3, dn + 4, dn + 5, rec_in))) return rc;
  double* de = c->s_d.as<double>();  // [eigmin | psimch | evals]
  hipLaunchKernelGGL(mcp::k_bound_match, dim3(1), dim3(kNT), lds_eig(n), c->stream, n, yo, yi, dn, dn + 3, dn + 6, de,
                     dn + 7, de + 1, de + 1 + n);
  HIPCHK(c, hipGetLastError());
  mcp::WaveSeg s0, s1;
  s0.recs = rec_out;
  s0.out = d_pout;
  s0.nstep = nout;
  s1.recs = rec_in;
  s1.out = d_pin;
  s1.nstep = nin;
  hipLaunchKernelGGL(mcp::k_wvbksb, dim3(2), dim3(kNT), 2 * (size_t)n * sizeof(double), c->stream, n, de + 1, s0, s1);
  HIPCHK(c, hipGetLastError());
  int tmp[8];
  HIPCHK(c, hipMemcpyAsync(tmp, dn, 8 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(eigmin, de, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(psimch, de + 1, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(psi_out, d_pout, ((size_t)nout + 1) * n * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(psi_in, d_pin, ((size_t)nin + 1) * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *nodec = tmp[6];
  *status = 0;
  if (tmp[1] != 0) *status = -tmp[1];
  if (tmp[4] != 0) *status = -(100000 + tmp[4]);
  if (tmp[2] != 0 || tmp[5] != 0) *status = -1000000;
  return MCP_OK;
}

int mcp_set_large_mode(mcp_ctx* c, int32_t mode) {
  if (!c || mode < 0 || mode > 7 || (mode & 3) > 2)
    return c ? c->fail(MCP_ERR_ARG, "large mode %d: 0, 1 or 2, plus 4", mode) : MCP_ERR_ARG;
  c->large_mode = mode & 3;
  c->eig_global = (mode & 4) != 0;
  return MCP_OK;
}

int mcp_set_vivs_options(mcp_ctx* c, int32_t iv, int32_t ivp, int32_t ivpp, int32_t ishift, int32_t idiag,
                         int32_t ipert, int32_t isym, int32_t ialpha, int32_t ialfp, double alpha1, double alpha2,
                         double xsqmax) {
  if (!c) return MCP_ERR_ARG;
  if (ialpha > 0 && !(alpha1 > 0.0 && alpha2 > 0.0))
    return c->fail(MCP_ERR_ARG, "vivs_options: IALPHA > 0 needs ALPHA1, ALPHA2 > 0");
  if (!(xsqmax > 0.0)) return c->fail(MCP_ERR_ARG, "vivs_options: XSQMAX must be positive");
  mcp::VivsOpts o;
  o.iv = iv != 0;
  o.ivp = ivp != 0;
  o.ivpp = ivpp != 0;
  o.ishift = ishift != 0;
  o.idiag = idiag != 0;
  o.ipert = ipert != 0;
  o.isym = isym != 0;
  o.ialpha = ialpha;
  o.ialfp = ialfp != 0;
  o.alpha1 = alpha1;
  o.alpha2 = alpha2;
  o.xsqmax = xsqmax;
  c->vivs_opt = o;
  return MCP_OK;
}

int mcp_set_precision(mcp_ctx* c, int32_t newton_steps, double r_fp64) {
  if (!c || newton_steps < 0 || newton_steps > 4)
    return c ? c->fail(MCP_ERR_ARG, "precision: 0 (FP64) to 4 Newton steps, not %d", newton_steps) : MCP_ERR_ARG;
  c->mixed_steps = newton_steps;
  c->mixed_r64 = r_fp64;
  c->mixed_fp64 = 0;
  c->mixed_total = 0;
  return MCP_OK;
}

int mcp_precision_stats(const mcp_ctx* c, int64_t* mixed_matrices, int64_t* fp64_again) {
  if (!c || !mixed_matrices || !fp64_again) return MCP_ERR_ARG;
  *mixed_matrices = c->mixed_total;
  *fp64_again = c->mixed_fp64;
  return MCP_OK;
}

int mcp_set_inverse_kernel(mcp_ctx* c, int32_t mode) {
  if (!c || mode < 0 || mode > 1) return c ? c->fail(MCP_ERR_ARG, "inverse kernel %d: 0 or 1", mode) : MCP_ERR_ARG;
  c->inv_kernel = mode;
  return MCP_OK;
}

int mcp_ytrans(mcp_ctx* c, int32_t bid, int32_t ninst, const double* pconst, double* y, double* eint) {
  return mcp_ytrans_cent(c, bid, ninst, pconst, y, eint, nullptr);
}

int mcp_set_degtol(mcp_ctx* c, double degtol) {
  if (!c || !(degtol >= 0.0)) return c ? c->fail(MCP_ERR_ARG, "degtol: must be >= 0") : MCP_ERR_ARG;
  c->degtol = degtol;
  return MCP_OK;
}

int mcp_ytrans_cent(mcp_ctx* c, int32_t bid, int32_t ninst, const double* pconst, double* y, double* eint,
                    double* cent) {
  Block* b;
  int rc;
  if (!c || ninst < 0 || !eint) return c ? c->fail(MCP_ERR_ARG, "ytrans: bad args") : MCP_ERR_ARG;
  if ((rc = lookup(c, bid, 0, &b, nullptr))) return rc;
  if (ninst == 0) return MCP_OK;
  const int n = b->n;
  if (b->nconst <= 0 && b->nrsq <= 0)
    return c->fail(MCP_ERR_ARG, "ytrans: the block has neither constant nor centrifugal terms");
  if (b->nconst > 0 && !pconst) return c->fail(MCP_ERR_ARG, "ytrans: NCONST > 0 needs pconst");
  HIPCHK(c, hipSetDevice(c->device));
  if (b->nconst > 0 && (rc = upload(c, c->s_h, pconst, (size_t)ninst * b->nconst))) return rc;
  HIPCHK(c, c->s_ye.ensure((size_t)ninst * n * sizeof(double)));
  double* dy = nullptr;
  if (y) {
    if ((rc = upload(c, c->s_i, y, (size_t)ninst * n * n))) return rc;
    dy = c->s_i.as<double>();
  }
  if ((rc = ytrans_device(c, *b, ninst, b->nconst > 0 ? c->s_h.as<double>() : nullptr, dy, c->s_ye.as<double>())))
    return rc;
  if ((rc = download(c, eint, c->s_ye, (size_t)ninst * n))) return rc;
  if (y && (rc = download(c, y, c->s_i, (size_t)ninst * n * n))) return rc;
  if (cent) {
    if (b->nrsq > 0) {
      if ((rc = download(c, cent, c->s_yc, (size_t)ninst * n))) return rc;
    } else {
      std::vector<double> c0(n);
      HIPCHK(c, hipMemcpyAsync(c0.data(), b->cent.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (int i = 0; i < ninst; ++i) std::copy(c0.begin(), c0.end(), cent + (size_t)i * n);
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MCP_OK;
}

int mcp_block_set_potl(mcp_ctx* c, int32_t bid, int32_t nterm, const int32_t* lam, const int32_t* npower,
                       const double* a, const double* e, double rscale, double ep2ru) {
  if (!c || nterm < 0 || (nterm > 0 && (!lam || !npower || !a || !e)))
    return c ? c->fail(MCP_ERR_ARG, "block_set_potl: bad args") : MCP_ERR_ARG;
  Block* b;
  int rc;
  if ((rc = lookup(c, bid, 0, &b, nullptr))) return rc;
  for (int t = 0; t < nterm; ++t)
    if (lam[t] < 0 || lam[t] >= b->mxlam)
      return c->fail(MCP_ERR_ARG, "block_set_potl: term %d belongs to radial term %d of %d", t, lam[t], b->mxlam);
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = upload(c, b->potl_lam, lam, (size_t)nterm))) return rc;
  if ((rc = upload(c, b->potl_npow, npower, (size_t)nterm))) return rc;
  if ((rc = upload(c, b->potl_a, a, (size_t)nterm))) return rc;
  if ((rc = upload(c, b->potl_e, e, (size_t)nterm))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  b->potl_nterm = nterm;
  b->potl_rscale = rscale;
  b->potl_ep2ru = ep2ru;
  return MCP_OK;
}

int mcp_prop_vivs(mcp_ctx* c, int32_t bid, int32_t ninst, const double* ered, double* rmat, double rstart,
                  double rstop, double drnow, double drmax, double tldiag, double toff, int32_t replay,
                  int32_t max_steps, double* steps, int32_t* nstep, int32_t* status) {
  if (!c || ninst < 0 || !ered || !rmat || !nstep || !status || replay < 0 || replay > 2 || max_steps <= 0)
    return c ? c->fail(MCP_ERR_ARG, "prop_vivs: bad args") : MCP_ERR_ARG;
  if (replay == 1 && !steps) return c->fail(MCP_ERR_ARG, "prop_vivs: replay without a step record");
  if (!(rstop > rstart) || !(drnow > 0.0) || !(drmax > 0.0) || !(tldiag > 0.0) || !(toff > 0.0))
    return c->fail(MCP_ERR_ARG, "prop_vivs: needs RSTOP > RSTART and positive DRNOW, DRMAX, TLDIAG, TOFF");
  Block* b;
  int rc;
  if ((rc = lookup(c, bid, 0, &b, nullptr))) return rc;
  if (ninst == 0) return MCP_OK;
  if (b->nconst != 0 || b->nrsq != 0)
    return c->fail(MCP_ERR_UNSUPPORTED, "prop_vivs: blocks with NCONST or NRSQ terms are not supported");
  if (b->potl_nterm < 0) return c->fail(MCP_ERR_ARG, "prop_vivs: the block has no potential (mcp_block_set_potl)");
  const int n = b->n;
  const size_t nn = (size_t)n * n;
  HIPCHK(c, hipSetDevice(c->device));
  const bool lds_mats = lds_vivs(n, b->nham, true) <= c->lds_max;
  const size_t lds = lds_vivs(n, b->nham, lds_mats);
  if (lds > c->lds_max) return c->fail(MCP_ERR_UNSUPPORTED, "prop_vivs: N=%d exceeds the LDS limit", n);
  if ((rc = upload(c, c->s_ve, ered, (size_t)ninst))) return rc;
  if ((rc = upload(c, c->s_vr, rmat, (size_t)ninst * nn))) return rc;
  HIPCHK(c, c->s_vs.ensure((size_t)max_steps * 4 * sizeof(double)));
  HIPCHK(c, c->s_vn.ensure((size_t)ninst * sizeof(int)));
  HIPCHK(c, c->s_vst.ensure((size_t)ninst * sizeof(int)));
  if (!lds_mats) HIPCHK(c, c->s_vw.ensure((size_t)ninst * mcp::kVivsMats * nn * sizeof(double)));
  if (replay == 1 && (rc = upload(c, c->s_vs, steps, (size_t)max_steps * 4))) return rc;
  mcp::PotlArgs P;
  P.nterm = b->potl_nterm;
  P.lam = b->potl_lam.as<int>();
  P.npow = b->potl_npow.as<int>();
  P.a = b->potl_a.as<double>();
  P.e = b->potl_e.as<double>();
  P.rscale = b->potl_rscale;
  P.ep2ru = b->potl_ep2ru;
  mcp::VivsArgs A;
  A.opt = c->vivs_opt;
  A.rstart = rstart;
  A.rstop = rstop;
  A.drnow = drnow;
  A.drmax = drmax;
  A.tldiag = tldiag;
  A.toff = toff;
  A.max_steps = max_steps;
  A.nrec = replay == 1 ? max_steps : 0;
  A.steps = c->s_vs.as<double>();
  A.ered = c->s_ve.as<double>();
  A.rmat = c->s_vr.as<double>();
  A.work = lds_mats ? nullptr : c->s_vw.as<double>();
  A.lds_mats = lds_mats ? 1 : 0;
  A.nstep = c->s_vn.as<int>();
  A.status = c->s_vst.as<int>();
  const mcp::BlockArgs B = block_args(*b);
  // replay 2: instance 0 chooses and records the steps, the others follow
  // them (the reference's scratch-file sequence over energies)
  const bool record = replay == 2 || (replay == 0 && steps);
  A.inst0 = 0;
  A.replay = replay == 1 ? 1 : 0;
  A.record = record ? 1 : 0;
  const int first = replay == 2 ? 1 : (record ? 1 : ninst);
  hipLaunchKernelGGL(mcp::k_vivs, dim3(first), dim3(kNT), lds, c->stream, B, P, A);
  HIPCHK(c, hipGetLastError());
  if (first < ninst) {
    if (replay == 2) {
      // the record holds instance 0's steps only if it finished; a failed
      // recording leaves nrec = 0, so the followers stop with kVivsReplayShort
      // instead of reading steps left in the buffer by an earlier call
      int n0 = 0, st0 = 0;
      HIPCHK(c, hipMemcpyAsync(&n0, c->s_vn.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipMemcpyAsync(&st0, c->s_vst.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      A.nrec = st0 == 0 ? std::min(n0, (int)max_steps) : 0;
    }
    A.inst0 = first;
    A.replay = replay == 0 ? 0 : 1;
    A.record = 0;
    hipLaunchKernelGGL(mcp::k_vivs, dim3(ninst - first), dim3(kNT), lds, c->stream, B, P, A);
    HIPCHK(c, hipGetLastError());
  }
  if ((rc = download(c, rmat, c->s_vr, (size_t)ninst * nn))) return rc;
  if ((rc = download(c, nstep, c->s_vn, (size_t)ninst))) return rc;
  if ((rc = download(c, status, c->s_vst, (size_t)ninst))) return rc;
  if (steps && replay != 1 && (rc = download(c, steps, c->s_vs, (size_t)max_steps * 4))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MCP_OK;
}

int mcp_prop_airy_tolhi(mcp_ctx* c, int32_t bid, int32_t ninst, const double* ered, double* z, double rstart,
                        double rstop, double dr0, double powr, double tolhi, int32_t replay, int32_t max_steps,
                        double* steps, int32_t* nrec, int32_t* nodes, int32_t* nstep, int32_t* status) {
  if (!c || ninst < 0 || !ered || !z || !nodes || !nstep || !status || replay < 0 || replay > 2 || max_steps <= 0)
    return c ? c->fail(MCP_ERR_ARG, "prop_airy_tolhi: bad args") : MCP_ERR_ARG;
  if (replay == 1 && (!steps || !nrec || *nrec <= 0 || *nrec > max_steps))
    return c->fail(MCP_ERR_ARG, "prop_airy_tolhi: replay needs a record of 1..max_steps steps");
  if (!(tolhi > 0.0 && tolhi < 1.0) || !(powr > 0.0) || dr0 == 0.0 || rstop == rstart ||
      (rstop > rstart) != (dr0 > 0.0))
    return c->fail(MCP_ERR_ARG, "prop_airy_tolhi: needs 0 < TOLHI < 1, POWR > 0 and DR0 pointing to RSTOP");
  Block* b;
  int rc;
  if ((rc = lookup(c, bid, 0, &b, nullptr))) return rc;
  if (ninst == 0) return MCP_OK;
  if (b->nconst != 0 || b->nrsq != 0)
    return c->fail(MCP_ERR_UNSUPPORTED, "prop_airy_tolhi: blocks with NCONST or NRSQ terms are not supported");
  if (b->potl_nterm < 0)
    return c->fail(MCP_ERR_ARG, "prop_airy_tolhi: the block has no potential (mcp_block_set_potl)");
  const int n = b->n;
  const size_t nn = (size_t)n * n;
  HIPCHK(c, hipSetDevice(c->device));
  // the six N x N matrices in LDS when they fit (N up to about 56 on
  // gfx950), else in a per-instance global workspace behind the same pointers
  if (n > mcp::kEMaxN) return c->fail(MCP_ERR_UNSUPPORTED, "prop_airy_tolhi: N=%d above %d", n, mcp::kEMaxN);
  const bool lds_mats = lds_airytol(n, b->nham, true) <= c->lds_max;
  const size_t lds = lds_airytol(n, b->nham, lds_mats);
  if (lds > c->lds_max)
    return c->fail(MCP_ERR_UNSUPPORTED, "prop_airy_tolhi: N=%d exceeds the LDS limit", n);
  if (!lds_mats) HIPCHK(c, c->s_vw.ensure((size_t)ninst * mcp::kAiryTolMats * nn * sizeof(double)));
  if ((rc = upload(c, c->s_ve, ered, (size_t)ninst))) return rc;
  if ((rc = upload(c, c->s_vr, z, (size_t)ninst * nn))) return rc;
  HIPCHK(c, c->s_vs.ensure((size_t)max_steps * 2 * sizeof(double)));
  HIPCHK(c, c->s_vn.ensure((size_t)ninst * sizeof(int)));
  HIPCHK(c, c->s_vnd.ensure((size_t)ninst * sizeof(int)));
  HIPCHK(c, c->s_vst.ensure((size_t)ninst * sizeof(int)));
  if (replay == 1 && (rc = upload(c, c->s_vs, steps, (size_t)(*nrec) * 2))) return rc;
  mcp::PotlArgs P;
  P.nterm = b->potl_nterm;
  P.lam = b->potl_lam.as<int>();
  P.npow = b->potl_npow.as<int>();
  P.a = b->potl_a.as<double>();
  P.e = b->potl_e.as<double>();
  P.rscale = b->potl_rscale;
  P.ep2ru = b->potl_ep2ru;
  mcp::AiryTolArgs A;
  A.rstart = rstart;
  A.rstop = rstop;
  A.dr0 = dr0;
  A.powr = powr;
  A.tolhi = tolhi;
  A.max_steps = max_steps;
  A.nrec = replay == 1 ? *nrec : 0;
  A.steps = c->s_vs.as<double>();
  A.ered = c->s_ve.as<double>();
  A.z = c->s_vr.as<double>();
  A.nodes = c->s_vnd.as<int>();
  A.nstep = c->s_vn.as<int>();
  A.status = c->s_vst.as<int>();
  A.work = lds_mats ? nullptr : c->s_vw.as<double>();
  A.lds_mats = lds_mats ? 1 : 0;
  const mcp::BlockArgs B = block_args(*b);
  // replay 2: instance 0 chooses and records the steps, the others follow
  const bool record = replay == 2 || (replay == 0 && steps);
  A.inst0 = 0;
  A.replay = replay == 1 ? 1 : 0;
  A.record = record ? 1 : 0;
  const int first = record ? 1 : ninst;
  hipLaunchKernelGGL(mcp::k_airy_tolhi, dim3(first), dim3(kNT), lds, c->stream, B, P, A);
  HIPCHK(c, hipGetLastError());
  if (first < ninst) {
    int n0 = 0, st0 = 0;
    if (replay == 2) {
      // the record's length is instance 0's step count, if it finished
      HIPCHK(c, hipMemcpyAsync(&n0, c->s_vn.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipMemcpyAsync(&st0, c->s_vst.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      A.nrec = st0 == 0 ? n0 : 0;
    }
    A.inst0 = first;
    A.replay = replay == 0 ? 0 : 1;
    A.record = 0;
    hipLaunchKernelGGL(mcp::k_airy_tolhi, dim3(ninst - first), dim3(kNT), lds, c->stream, B, P, A);
    HIPCHK(c, hipGetLastError());
  }
  if ((rc = download(c, z, c->s_vr, (size_t)ninst * nn))) return rc;
  if ((rc = download(c, nodes, c->s_vnd, (size_t)ninst))) return rc;
  if ((rc = download(c, nstep, c->s_vn, (size_t)ninst))) return rc;
  if ((rc = download(c, status, c->s_vst, (size_t)ninst))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (steps && replay != 1) {
    const int k = std::min(nstep[0], (int)max_steps);
    if ((rc = download(c, steps, c->s_vs, (size_t)k * 2))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (nrec) *nrec = k;
  }
  return MCP_OK;
}

// SYMSEL's schemes (source_code/mgprop.f:103-150): IMGSEL = 4 -> CS4, 5 -> MA5
static bool mg_scheme(int imgsel, mcp::MgScheme& S) {
  static const double a4[5] = {0.20517766154229, 0.40302128160421, -0.12092087633891, 0.51272193319241, 0.0};
  static const double b4[5] = {0.061758858135626, 0.33897802655364, 0.61479130717558, -0.14054801465937,
                               0.12501982279453};
  static const double a5[6] = {0.33983962583911, -0.088601336903027, 0.58585647682596, -0.60303935653649,
                               0.32358079655470, 0.44236379421975};
  static const double b5[6] = {0.11939002928757, 0.69892737038248, -0.17131235827160, 0.40126950225135,
                               0.010705081848236, -0.058979625498031};
  const double* a = imgsel == 4 ? a4 : imgsel == 5 ? a5 : nullptr;
  const double* b = imgsel == 4 ? b4 : b5;
  if (!a) return false;
  S.m = imgsel == 4 ? 5 : 6;
  for (int k = 0; k < mcp::kMgMaxSub; ++k) {
    S.a[k] = k < S.m ? a[k] : 0.0;
    S.b[k] = k < S.m ? b[k] : 0.0;
  }
  return true;
}

int mcp_mg_scheme(int32_t imgsel, int32_t* nsub, double* a, double* b) {
  mcp::MgScheme S;
  if (!nsub || !a || !b || !mg_scheme(imgsel, S)) return MCP_ERR_ARG;
  *nsub = S.m;
  for (int k = 0; k < S.m; ++k) {
    a[k] = S.a[k];
    b[k] = S.b[k];
  }
  return MCP_OK;
}

int mcp_prop_mg(mcp_ctx* c, int32_t bid, int32_t gid, int32_t imgsel, int32_t nstep, double dr, int32_t ninst,
                const double* ered, const double* pconst, double* y, int32_t* nodes, int32_t* status) {
  if (!c || nstep <= 0 || ninst < 0 || !ered || !y || !nodes || !status)
    return c ? c->fail(MCP_ERR_ARG, "prop_mg: bad args") : MCP_ERR_ARG;
  mcp::MgScheme S;
  if (!mg_scheme(imgsel, S)) return c->fail(MCP_ERR_ARG, "prop_mg: IMGSEL must be 4 (CS4) or 5 (MA5)");
  Block* b;
  Grid* g;
  int rc;
  if ((rc = lookup(c, bid, gid, &b, &g))) return rc;
  if (ninst == 0) return MCP_OK;
  const bool reuse = S.a[S.m - 1] == 0.0;
  const long long need = reuse ? (long long)S.m + (long long)(nstep - 1) * (S.m - 1) : (long long)nstep * S.m;
  if ((long long)g->npts != need)
    return c->fail(MCP_ERR_ARG, "prop_mg: the grid has %d radii, the scheme needs %lld for %d sectors", g->npts,
                   need, nstep);
  if (b->nconst > 0 && !pconst) return c->fail(MCP_ERR_ARG, "prop_mg: the block needs its constants (pconst)");
  const int n = b->n;
  const size_t lds = LdsSize().add<double>((size_t)n * n).add<double>(b->nham > 0 ? b->nham : 1).sweep(n).b;
  if (lds > c->lds_max) return c->fail(MCP_ERR_UNSUPPORTED, "prop_mg: N=%d exceeds the LDS-resident limit", n);
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = upload_inst(c, *b, ninst, ered, pconst))) return rc;
  if ((rc = upload(c, c->s_i, y, (size_t)ninst * n * n))) return rc;
  HIPCHK(c, c->s_vnd.ensure((size_t)ninst * sizeof(int)));
  HIPCHK(c, c->s_vst.ensure((size_t)ninst * sizeof(int)));
  mcp::InstArgs I;
  I.ered = c->s_g.as<double>();
  I.pconst = b->nconst > 0 ? c->s_h.as<double>() : nullptr;
  I.y = c->s_i.as<double>();
  I.nodes = c->s_vnd.as<int>();
  I.status = c->s_vst.as<int>();
  hipLaunchKernelGGL(mcp::k_mgprop, dim3(ninst), dim3(kNT), lds, c->stream, block_args(*b), grid_args(*g), I, S,
                     nstep, dr, n > 30 ? 1 : 0);
  HIPCHK(c, hipGetLastError());
  if ((rc = download(c, y, c->s_i, (size_t)ninst * n * n))) return rc;
  if ((rc = download(c, nodes, c->s_vnd, (size_t)ninst))) return rc;
  if ((rc = download(c, status, c->s_vst, (size_t)ninst))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MCP_OK;
}

int mcp_set_md_kernel(mcp_ctx* c, int32_t mode) {
  if (!c || mode < 0 || mode > 2) return c ? c->fail(MCP_ERR_ARG, "md_kernel: 0, 1 or 2") : MCP_ERR_ARG;
  c->md_kernel = mode;
  return MCP_OK;
}

int mcp_set_lds_threads(mcp_ctx* c, int32_t nt) {
  if (!c || (nt != 0 && nt != 64 && nt != kNT)) return c ? c->fail(MCP_ERR_ARG, "lds_threads: 0, 64 or 256") : MCP_ERR_ARG;
  c->lds_threads = nt;
  return MCP_OK;
}

int mcp_set_team_launch(mcp_ctx* c, int32_t cooperative) {
  if (!c) return MCP_ERR_ARG;
  c->coop = cooperative != 0;
  return MCP_OK;
}

int mcp_set_team_size(mcp_ctx* c, int32_t g) {
  if (!c) return MCP_ERR_ARG;
  c->no_reg = g < 0;
  c->team = g < 0 ? -g : g;
  return MCP_OK;
}

int mcp_last_segment_timing(const mcp_ctx* c, int32_t k, double* ms_sector, double* ms_prop) {
  if (!c || k < 0 || k >= c->seg_timed) return MCP_ERR_ARG;
  if (ms_sector) *ms_sector = c->seg_ms[k][0];
  if (ms_prop) *ms_prop = c->seg_ms[k][1];
  return MCP_OK;
}

int mcp_last_timing(const mcp_ctx* c, double* ms_sector, double* ms_prop, double* ms_total) {
  if (!c) return MCP_ERR_ARG;
  if (ms_sector) *ms_sector = c->t_sector;
  if (ms_prop) *ms_prop = c->t_prop;
  if (ms_total) *ms_total = c->t_total;
  return MCP_OK;
}

}  // extern "C"
