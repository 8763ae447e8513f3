// Host build of the device Percival-Seaton coefficients
// (molscat_amd/csrc/mcp_couple.hpp) for the CPU tests: the same source,
// compiled for the host, checked against exact rational arithmetic.
#include <vector>

#include "../../molscat_amd/csrc/mcp_couple.hpp"

namespace {
std::vector<double> g_hi, g_lo;
}

extern "C" {
int couple_table_h(int kmax) {
  g_hi.assign(kmax + 1, 0.0);
  g_lo.assign(kmax + 1, 0.0);
  mcp::build_lnfact(kmax, g_hi.data(), g_lo.data());
  return kmax;
}
int couple_need_h(int jmax, int lmax, int jtot, int lmx) { return mcp::couple_lnf_need(jmax, lmax, jtot, lmx); }
double percival_seaton_h(int jc, int lc, int jr, int lr, int jtot, int lam) {
  mcp::LnFact f{g_hi.data(), g_lo.data(), (int)g_hi.size() - 1};
  return mcp::percival_seaton(f, jc, lc, jr, lr, jtot, lam);
}
void lnfact_h(int k, double* out) {
  out[0] = g_hi[k];
  out[1] = g_lo[k];
}
int packed_col_h(long long k) { return mcp::packed_col(k); }
}
