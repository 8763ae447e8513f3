"""CPU checks of the drop-in boundary: the built library exports every
symbol include/mcprop.h declares, and the ctypes binding covers them.  No
compute call is made (no GPU here)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mcprop.h")
LIB = os.path.join(ROOT, "molscat_amd", "libmcprop.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|size_t|const char \*)\s*\*?\s*(mcp_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("mcp_ctx_create", "mcp_syminv_batched", "mcp_prop_ld", "mcp_ytok_ktos", "mcp_smatrix_ld"):
        assert s in syms


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmcprop.so not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (mcp_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


def test_ctypes_binding_covers_header():
    from molscat_amd import lib

    assert set(declared_symbols()) == set(lib.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmcprop.so not built")
def test_library_loads_without_gpu():
    import ctypes

    ctypes.CDLL(LIB)  # dlopen only: no HIP call is made


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmcprop.so not built")
def test_mg_scheme_table_is_the_reference_table():
    """MGPROP's CS4 and MA5 coefficients as the library holds them (a host
    table, no GPU call) equal the oracle's restatement of SYMSEL bit for
    bit, and other IMGSEL values are refused."""
    import ctypes

    import numpy as np

    from molscat_amd import lib
    from oracle import molscat_oracle as O

    L = lib.load()
    for sel, (a_ref, b_ref) in O.MG_SCHEMES.items():
        m = ctypes.c_int32()
        a = np.zeros(6)
        b = np.zeros(6)
        assert L.mcp_mg_scheme(sel, ctypes.byref(m), lib.dptr(a), lib.dptr(b)) == 0
        assert m.value == len(a_ref)
        assert np.array_equal(a[: m.value], np.array(a_ref)) and np.array_equal(b[: m.value], np.array(b_ref))
    m = ctypes.c_int32()
    assert L.mcp_mg_scheme(3, ctypes.byref(m), lib.dptr(np.zeros(6)), lib.dptr(np.zeros(6))) == -1


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmcprop.so not built")
def test_entry_points_refuse_a_null_context():
    """Argument checks come before any device call: a NULL context gives
    MCP_ERR_ARG for the propagators added beside the LD ones."""
    import ctypes

    import numpy as np

    from molscat_amd import lib

    L = lib.load()
    d = lib.dptr(np.zeros(4))
    i = lib.iptr(np.zeros(4, dtype=np.int32))
    assert L.mcp_prop_vivs(None, 1, 1, d, d, 1.0, 2.0, 0.01, 5.0, 0.02, 0.02, 0, 10, d, i, i) == -1
    assert L.mcp_prop_airy_tolhi(None, 1, 1, d, d, 1.0, 2.0, 0.01, 3.0, 1e-4, 0, 10, d, i, i, i, i) == -1
    assert L.mcp_prop_mg(None, 1, 1, 5, 10, 0.01, 1, d, None, d, i, i) == -1
    assert L.mcp_block_set_potl(None, 1, 0, None, None, None, None, 1.0, 1.0) == -1
    assert L.mcp_set_lds_threads(None, 64) == -1
