"""molscat_amd: MI355X (gfx950) engine for the MOLSCAT/BOUND/FIELD
log-derivative propagation and matching path.

  * ``libmcprop.so`` (csrc/, C ABI in include/mcprop.h): HIP kernels.
  * ``engine.Engine``: ctypes handle on the C ABI.
  * ``system``: host-side set-up (units, rotor basis and coupling, POTENL
    terms, grids, channel bookkeeping).
  * ``scctrl``: SCCTRL-style control of one block (segments, matching).
"""
__all__ = ["engine", "system", "build"]
