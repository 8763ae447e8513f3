"""mcp_ytok_ktos_noalloc, the matching entry point that allocates no device
memory, against mcp_ytok_ktos (same kernels, host arrays and library
buffers) on the same inputs: two calls in a row, the second smaller, each
bitwise equal to the allocating entry point's K, S and status.  Also: the
size limit refused with MCP_ERR_UNSUPPORTED, and an out-of-range NB entry
reported per instance (status 99) with the other instances unaffected.
Device arrays come from hipMalloc of the HIP runtime the library itself is
linked against (loaded by its soname once the library is loaded), so the
test's buffers and the library live in one runtime."""
import ctypes

import numpy as np
import pytest

from molscat_amd import system as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def _case(jtot, energies, seed):
    pot = S.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    blk = S.rotor_block(jtot, 1, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)
    n = blk.n
    rng = np.random.default_rng(seed)
    ered = np.asarray(energies) * blk.cm2ru
    wv, nop = zip(*[S.gtwvec(e, blk.eint) for e in ered])
    nb = np.tile(S.chnsrt(blk.eint), (len(ered), 1)).astype(np.int32)
    a = rng.standard_normal((len(ered), n, n))
    y = 3.0 * (a + np.transpose(a, (0, 2, 1)))
    return blk, y, np.array(wv), np.array(nop, dtype=np.int32), nb


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the library's own HIP runtime."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.live = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(arr.nbytes, 8)) == 0
        self.live.append(p)
        assert self.rt.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        assert self.rt.hipDeviceSynchronize() == 0
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _noalloc(eng, n, y, wv, nop, nb, cent, rval):
    m = y.shape[0]
    hip = _Hip()
    try:
        # column-major N x N per instance, as the C ABI takes them
        d_y = hip.put(np.transpose(y, (0, 2, 1)).reshape(m, n * n).astype(np.float64))
        d_w = hip.put(np.asarray(wv, dtype=np.float64))
        d_o = hip.put(np.asarray(nop, dtype=np.int32))
        d_b = hip.put(np.asarray(nb, dtype=np.int32))
        d_c = hip.put(np.asarray(cent, dtype=np.float64))
        d_k = hip.put(np.zeros((m, n * n)))
        d_sr = hip.put(np.zeros((m, n * n)))
        d_si = hip.put(np.zeros((m, n * n)))
        d_st = hip.put(np.full(m, -7, dtype=np.int32))
        rc = eng.lib.mcp_ytok_ktos_noalloc(eng.ctx, n, m, d_y, d_w, d_o, d_b, d_c, float(rval), d_k, d_sr, d_si,
                                           d_st)
        out = dict(rc=rc, K=[], SR=[], SI=[], STATUS=hip.get(d_st, (m,), np.int32))
        if rc == 0:
            k = hip.get(d_k, (m, n * n), np.float64)
            sr = hip.get(d_sr, (m, n * n), np.float64)
            si = hip.get(d_si, (m, n * n), np.float64)
            for i in range(m):
                q = int(nop[i])
                out["K"].append(k[i, : q * q].reshape(q, q, order="F"))
                out["SR"].append(sr[i, : q * q].reshape(q, q, order="F"))
                out["SI"].append(si[i, : q * q].reshape(q, q, order="F"))
        return out
    finally:
        hip.free()


def test_noalloc_matches_allocating_entry_point_twice(eng):
    rval = 20.0
    for call, (jtot, energies, seed) in enumerate([(10, [900.0, 950.0, 1000.0, 1050.0, 1100.0, 1200.0], 1),
                                                   (4, [700.0, 1000.0], 2)]):
        blk, y, wv, nop, nb = _case(jtot, energies, seed)
        n = blk.n
        print(f"call {call}: N={n} ninst={len(energies)} nopen={list(nop)}", flush=True)
        k_ref, sr_ref, si_ref, st_ref = eng.ytok_ktos(y, wv, nop, nb, blk.cent, rval)
        out = _noalloc(eng, n, y, wv, nop, nb, blk.cent, rval)
        assert out["rc"] == 0
        assert np.array_equal(out["STATUS"], st_ref)
        assert np.all(st_ref == 0)
        for i in range(len(energies)):
            assert np.array_equal(out["K"][i], k_ref[i])
            assert np.array_equal(out["SR"][i], sr_ref[i])
            assert np.array_equal(out["SI"][i], si_ref[i])
            s = out["SR"][i] + 1j * out["SI"][i]
            assert np.abs(s @ s.conj().T - np.eye(len(s))).max() < 1e-12


def test_noalloc_refuses_beyond_its_scratch(eng):
    n, m = 441, 11
    assert eng.lib.mcp_ytok_ktos_noalloc_bytes(n, m) > 32 * 2 ** 20
    assert eng.lib.mcp_ytok_ktos_noalloc_bytes(n, 10) <= 32 * 2 ** 20
    # full-size, valid device arrays (NOPEN = 0 everywhere), so nothing could
    # be read out of range even if the size check did not stop the call
    y = np.zeros((m, n, n))
    out = _noalloc(eng, n, y, np.ones((m, n)), np.zeros(m, dtype=np.int32), np.zeros((m, n), dtype=np.int32),
                   np.zeros(n), 10.0)
    assert out["rc"] == -3
    assert b"scratch" in eng.lib.mcp_last_error(eng.ctx)


def test_noalloc_bad_channel_index_is_per_instance(eng):
    blk, y, wv, nop, nb = _case(10, [900.0, 1000.0, 1100.0], 3)
    nb = nb.copy()
    nb[1, 3] = blk.n + 5
    out = _noalloc(eng, blk.n, y, wv, nop, nb, blk.cent, 20.0)
    assert out["rc"] == 0
    assert list(out["STATUS"]) == [0, 99, 0]
    good = np.array([True, False, True])
    k_ref, sr_ref, si_ref, st_ref = eng.ytok_ktos(y[good], wv[good], nop[good], nb[good], blk.cent, 20.0)
    for j, i in enumerate([0, 2]):
        assert np.array_equal(out["SR"][i], sr_ref[j])
