"""Pins the oracle's symmetric inversion (SYMINV inline path and the
Bunch-Kaufman DSYTF2/DSYNEG/DSYTRI restatement) against what the
mathematics fixes: A*inv(A) = I, the inertia equals the number of negative
eigenvalues (Sylvester's law), exact zero pivots give N+1 on the inline
path, and 2x2 pivots are exercised by matrices with a zero diagonal."""
import itertools

import numpy as np
import pytest

from oracle import molscat_oracle as O


def _rand_sym(n, rng, shift=0.0):
    a = rng.standard_normal((n, n))
    return 0.5 * (a + a.T) + shift * np.eye(n)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 12, 30, 31, 48, 64, 100])
def test_inverse_and_inertia(n):
    rng = np.random.default_rng(1000 + n)
    for _ in range(4):
        a = _rand_sym(n, rng)
        inv, k = O.syminv(a)
        assert k == int(np.sum(np.linalg.eigvalsh(a) < 0.0))
        cond = np.linalg.cond(a)
        assert np.abs(a @ inv - np.eye(n)).max() < 1e-13 * cond * n


def test_inline_zero_pivot_gives_n_plus_1():
    # leading pivot exactly zero: source_code/syminv.f:24,95
    a = np.array([[0.0, 1.0], [1.0, 2.0]])
    _, k = O.syminv(a)
    assert k == 3
    # zero second pivot after elimination
    a = np.array([[1.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 5.0]])
    _, k = O.syminv(a)
    assert k == 4


def test_bunch_kaufman_two_by_two_pivots():
    # zero diagonal forces 2x2 pivots: inertia of [[0,B],[B^T,0]] is n/2 negative
    n = 40
    rng = np.random.default_rng(7)
    b = rng.standard_normal((n // 2, n // 2)) + 3 * np.eye(n // 2)
    a = np.zeros((n, n))
    a[: n // 2, n // 2:] = b
    a[n // 2:, : n // 2] = b.T
    f = np.array(a, order="F")
    import ctypes

    ipiv = np.zeros(n, dtype=np.int32)
    lib = O._lib()
    info = lib.orc_dsytf2_l(f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, n,
                            ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert info == 0
    assert np.any(ipiv < 0), "expected at least one 2x2 pivot"
    inv, k = O.syminv(a)
    assert k == n // 2
    assert np.abs(a @ inv - np.eye(n)).max() < 1e-11


def test_brute_force_small_inertia():
    # all 3x3 symmetric matrices with entries in {-1, 0, 1} and nonzero det
    vals = [-1.0, 0.0, 1.0]
    count = 0
    for d0, d1, d2, o01, o02, o12 in itertools.product(vals, repeat=6):
        a = np.array([[d0, o01, o02], [o01, d1, o12], [o02, o12, d2]])
        if abs(np.linalg.det(a)) < 1e-12:
            continue
        inv, k = O.syminv(a)
        if k == 4:  # inline path met an exact zero pivot: no inverse returned
            continue
        assert k == int(np.sum(np.linalg.eigvalsh(a) < 0))
        assert np.allclose(a @ inv, np.eye(3), atol=1e-12)
        count += 1
    assert count > 100


def test_bk_path_matches_inline_path_on_spd():
    # for N <= 30 both algorithms apply; on SPD matrices they must agree
    rng = np.random.default_rng(3)
    for n in (10, 25):
        a = _rand_sym(n, rng, shift=2 * n)
        inv1, k1 = O.syminv(a)
        import ctypes

        f = np.array(a, order="F")
        ipiv = np.zeros(n, dtype=np.int32)
        lib = O._lib()
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        assert lib.orc_dsytf2_l(f.ctypes.data_as(dp), n, n, ipiv.ctypes.data_as(ip)) == 0
        k2 = lib.orc_dsyneg_l(f.ctypes.data_as(dp), n, n, ipiv.ctypes.data_as(ip))
        assert lib.orc_dsytri_l(f.ctypes.data_as(dp), n, n, ipiv.ctypes.data_as(ip)) == 0
        inv2 = np.tril(f) + np.tril(f, -1).T
        assert k1 == k2 == 0
        assert np.abs(inv1 - inv2).max() < 1e-14 * np.abs(inv1).max() * n


@pytest.mark.parametrize("n", [31, 48, 100, 441])
def test_lapack_inversion_matches_restatement(n):
    """The CPU baseline's inversion (OpenBLAS DSYTRF + DSYNEG + DSYTRI) gives
    the restated SYMINL's inverse and inertia, including matrices with a zero
    diagonal (2x2 pivots)."""
    rng = np.random.default_rng(77 + n)
    for t in range(3):
        a = _rand_sym(n, rng)
        if t == 2:
            np.fill_diagonal(a, 0.0)
        x, k = O.syminv_lapack(a)
        y, k2 = O.syminv(a)
        assert k == k2 == int(np.sum(np.linalg.eigvalsh(a) < 0.0))
        assert np.abs(x - y).max() < 1e-10 * np.abs(y).max()
