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

extern "C" {
// out[0..3]: cc, cs, sc, ss against w1 = r - h/2; out[4..7] against (r - h/2)^2 / 2
void step_integrals_w_h(double zi, double zj, double h, double* out) {
  const mcp::StepIntegralsW r = mcp::step_integrals_w(zi, zj, h);
  out[0] = r.w1.cc;
  out[1] = r.w1.cs;
  out[2] = r.w1.sc;
  out[3] = r.w1.ss;
  out[4] = r.w2.cc;
  out[5] = r.w2.cs;
  out[6] = r.w2.sc;
  out[7] = r.w2.ss;
}
}
