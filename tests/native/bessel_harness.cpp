// Host build of the device Riccati-Bessel routines
// (molscat_amd/csrc/mcp_bessel.hpp) for the CPU tests: the same source,
// compiled for the host, checked against mpmath.
#include "../../molscat_amd/csrc/mcp_bessel.hpp"

extern "C" {
int riccati_jy_h(double ell, double x, double* out) {
  return mcp::riccati_jy(ell, x, out[0], out[1], out[2], out[3]);
}
int riccati_k_ratio_h(double ell, double x, double* out) { return mcp::riccati_k_ratio(ell, x, out[0]); }
}
