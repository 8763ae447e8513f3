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
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(mcp_\w+)\s*\(", text, re.M)))


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
