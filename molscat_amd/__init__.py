"""molscat_amd: MI355X (gfx950) engine for the MOLSCAT/BOUND/FIELD
log-derivative propagation and matching path.

  * ``libmcprop.so`` (csrc/, C ABI in include/mcprop.h): HIP kernels.
  * ``engine.Engine``: ctypes handle on the C ABI.
  * ``system``: host-side set-up (units, rotor basis and coupling, POTENL
    terms, grids, channel bookkeeping).
  * ``driver``: the MOLSCAT partial-wave / energy loop (SCCTRL's segments
    and matching per block), the batched sweep and the field scan.
  * ``bound``: BOUND's level search, FIELD's field search and the
    wavefunction.
  * ``smatfile``: the saved S-matrix file (ISAVEU) and restart.
"""
__all__ = ["engine", "system", "build"]
