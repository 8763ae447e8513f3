// Host build of the VIVS step integrals (molscat_amd/csrc/mcp_stepint.hpp)
// for the CPU tests.
#include "../../molscat_amd/csrc/mcp_stepint.hpp"

extern "C" {
void step_integrals_h(double zi, double zj, double h, double* out) {
  const mcp::StepIntegrals r = mcp::step_integrals(zi, zj, h);
  out[0] = r.cc;
  out[1] = r.cs;
  out[2] = r.sc;
  out[3] = r.ss;
}
}
