// Host build of the device Airy routines (molscat_amd/csrc/mcp_airy.hpp) for
// the CPU tests: the same source, compiled for the host, checked against
// mpmath.
#include "../../molscat_amd/csrc/mcp_airy.hpp"

extern "C" {
void airy_transfer_h(double x1, double delta, double* out) {
  mcp::airy_transfer(x1, delta, out[0], out[1], out[2], out[3]);
}
void airy_sector_h(double width, double eig, double slope, double* out) {
  mcp::airy_sector(width, eig, slope, out[0], out[1], out[2]);
}
}
