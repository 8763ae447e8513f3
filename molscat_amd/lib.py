"""ctypes binding of libmcprop.so (the C ABI declared in include/mcprop.h).

The library is built in-tree by ``molscat_amd.build.build_lib()``.  Loading
fails loudly when it is missing or when no gfx950 device is present: the
package has no CPU fallback for the propagation path.
"""
from __future__ import annotations

import ctypes
import os
from functools import lru_cache

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcprop.so")

MCP_OK = 0
MCP_ERR_ARG = -1
MCP_ERR_HIP = -2
MCP_ERR_UNSUPPORTED = -3

_i32 = ctypes.c_int32
_dbl = ctypes.c_double
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "mcp_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "mcp_ctx_destroy": (None, [_vp]),
    "mcp_last_error": (ctypes.c_char_p, [_vp]),
    "mcp_ctx_info": (ctypes.c_int, [_vp, _pi, _pi, _pi, _pi, _pi]),
    "mcp_syminv_batched": (ctypes.c_int, [_vp, _i32, _i32, _pd, _pd, _pi]),
    "mcp_eigh_batched": (ctypes.c_int, [_vp, _i32, _i32, _pd, _pd, _pd]),
    "mcp_block_create": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32, _pd, _pd, _pd, _dbl, _pi]),
    "mcp_block_destroy": (ctypes.c_int, [_vp, _i32]),
    "mcp_couple_rotor": (ctypes.c_int, [_vp, _i32, _pi, _pi, _i32, _i32, _pi, _pd]),
    "mcp_block_create_rotor": (ctypes.c_int, [_vp, _i32, _pi, _pi, _i32, _i32, _pi, _pd, _pd, _dbl, _pi]),
    "mcp_grid_create": (ctypes.c_int, [_vp, _i32, _i32, _pd, _pd, _dbl, _pi]),
    "mcp_grid_destroy": (ctypes.c_int, [_vp, _i32]),
    "mcp_hammat": (ctypes.c_int, [_vp, _i32, _i32, _i32, _pi, _pd, _dbl, _pd]),
    "mcp_yinit": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32, _dbl, _i32, _pd, _pd, _pd]),
    "mcp_prop_ld": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32, _pd, _pd, _pd, _pi, _pi]),
    "mcp_ytok_ktos": (ctypes.c_int, [_vp, _i32, _i32, _pd, _pd, _pi, _pi, _pd, _dbl, _pd, _pd, _pd, _pi]),
    "mcp_ytok_ktos_noalloc": (ctypes.c_int, [_vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _dbl, _vp, _vp, _vp, _vp]),
    "mcp_ytok_ktos_noalloc_bytes": (ctypes.c_size_t, [_i32, _i32]),
    "mcp_smatrix_ld": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32, _pd, _pd, _pd, _pi, _pi, _dbl, _pd, _pd, _pd,
                                       _pi, _pi]),
    "mcp_grid_create_airy": (ctypes.c_int, [_vp, _i32, _i32, _pd, _pd, _pd, _pi]),
    "mcp_prop_airy": (ctypes.c_int, [_vp, _i32, _i32, _i32, _pd, _pd, _pi, _pi]),
    "mcp_smatrix_segments": (ctypes.c_int, [_vp, _i32, _i32, _pi, _pi, _i32, _pd, _pd, _pd, _pi, _pi, _dbl, _pd, _pd,
                                             _pd, _pi, _pi]),
    "mcp_bound_match": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _pd, _pd, _pi, _pd, _pi, _pd, _pd,
                                        _pi]),
    "mcp_bound_wave": (ctypes.c_int, [_vp, _i32, _i32, _i32, _dbl, _pd, _pi, _pd, _pd, _pd, _pd, _pi]),
    "mcp_set_team_size": (ctypes.c_int, [_vp, _i32]),
    "mcp_set_large_mode": (ctypes.c_int, [_vp, _i32]),
    "mcp_set_inverse_kernel": (ctypes.c_int, [_vp, _i32]),
    "mcp_set_precision": (ctypes.c_int, [_vp, _i32, _dbl]),
    "mcp_set_vivs_options": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _dbl, _dbl,
                                             _dbl]),
    "mcp_set_md_kernel": (ctypes.c_int, [_vp, _i32]),
    "mcp_precision_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "mcp_prop_ld_multi": (ctypes.c_int, [_vp, _i32, _pi, _i32, _pi, _pd, _pd, _pi, _pi]),
    "mcp_set_graph": (ctypes.c_int, [_vp, _i32]),
    "mcp_set_team_launch": (ctypes.c_int, [_vp, _i32]),
    "mcp_set_lds_threads": (ctypes.c_int, [_vp, _i32]),
    "mcp_ytrans": (ctypes.c_int, [_vp, _i32, _i32, _pd, _pd, _pd]),
    "mcp_ytrans_cent": (ctypes.c_int, [_vp, _i32, _i32, _pd, _pd, _pd, _pd]),
    "mcp_set_degtol": (ctypes.c_int, [_vp, _dbl]),
    "mcp_last_timing": (ctypes.c_int, [_vp, _pd, _pd, _pd]),
    "mcp_last_segment_timing": (ctypes.c_int, [_vp, _i32, _pd, _pd]),
    "mcp_block_set_potl": (ctypes.c_int, [_vp, _i32, _i32, _pi, _pi, _pd, _pd, _dbl, _dbl]),
    "mcp_prop_vivs": (ctypes.c_int, [_vp, _i32, _i32, _pd, _pd, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _i32, _i32,
                                      _pd, _pi, _pi]),
    "mcp_mg_scheme": (ctypes.c_int, [_i32, _pi, _pd, _pd]),
    "mcp_prop_mg": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32, _dbl, _i32, _pd, _pd, _pd, _pi, _pi]),
    "mcp_prop_airy_tolhi": (ctypes.c_int, [_vp, _i32, _i32, _pd, _pd, _dbl, _dbl, _dbl, _dbl, _dbl, _i32, _i32,
                                            _pd, _pi, _pi, _pi, _pi]),
}


class McpError(RuntimeError):
    pass


@lru_cache(maxsize=1)
def load() -> ctypes.CDLL:
    """Load libmcprop.so and bind every exported entry point."""
    if not os.path.exists(LIB_PATH):
        raise McpError(f"{LIB_PATH} is not built; run molscat_amd.build.build_lib() (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def dptr(a: np.ndarray):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_pd)


def iptr(a: np.ndarray):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_pi)
