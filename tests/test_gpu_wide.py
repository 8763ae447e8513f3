"""The 32-pivot batched inversion (k_syminv_wide, the default for the
large-N paths up to N=448) against the 16-pivot one (mcp_set_inverse_kernel
0) and against numpy on the same matrices: SYMINV's job, the full inverse of
the lower triangle and the count of negative pivots (source_code/syminv.f,
syminl-alloc.f, dsyneg.f).

Bars: inertia equal to the number of negative eigenvalues (numpy eigvalsh)
and to the 16-pivot kernel's, exactly; residual |A A^-1 - I| below
1e-14 * cond * N (backward stability); the two kernels' inverses within
1e-13 * cond of each other; output bitwise symmetric; an exactly zero
column reported as -(k+1) like the 16-pivot kernel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.set_inverse_kernel(1)
    e.close()


def _both(eng, a):
    eng.set_inverse_kernel(0)
    i16, k16 = eng.syminv(a)
    eng.set_inverse_kernel(1)
    i32, k32 = eng.syminv(a)
    return i16, k16, i32, k32


@pytest.mark.parametrize("n", [141, 200, 257, 333, 441, 448])
def test_wide_vs_narrow_and_numpy(eng, n):
    rng = np.random.default_rng(100 + n)
    m = 3
    a = rng.standard_normal((m, n, n))
    a = 0.5 * (a + np.transpose(a, (0, 2, 1)))
    a[1, : n // 4, : n // 4] = 0.0          # zero block: 2x2 pivots and interchanges
    a[2] += np.diag(np.linspace(5.0, 40.0, n) * np.sign(rng.standard_normal(n)))  # mostly natural order
    print(f"wide inverse n={n}", flush=True)
    i16, k16, i32, k32 = _both(eng, a)
    for q in range(m):
        neg = int(np.sum(np.linalg.eigvalsh(a[q]) < 0))
        cond = np.linalg.cond(a[q])
        res = np.abs(a[q] @ i32[q] - np.eye(n)).max()
        d = np.abs(i32[q] - i16[q]).max() / np.abs(i16[q]).max()
        print(f"  m={q} inertia {k32[q]} (numpy {neg}) res {res:.1e} cond {cond:.1e} |d| {d:.1e}", flush=True)
        assert k32[q] == k16[q] == neg
        assert res < 1e-14 * cond * n
        assert d < 1e-13 * cond
        assert np.array_equal(i32[q], i32[q].T)


def test_wide_batch_of_mdprop_like_matrices(eng):
    # the N=441 rotor's leaf matrices V + E^-1 are diagonally dominant with
    # small couplings; a batch of 300 such with a few near-singular ones
    n = 441
    rng = np.random.default_rng(7)
    m = 300
    c = 0.02 * rng.standard_normal((m, n, n))
    a = 0.5 * (c + np.transpose(c, (0, 2, 1)))
    diag = rng.uniform(-3.0, 3.0, size=(m, n))
    diag[::37, 5] = 1e-9
    a[:, np.arange(n), np.arange(n)] += diag
    print("wide inverse batch of 300 at n=441", flush=True)
    i16, k16, i32, k32 = _both(eng, a)
    assert np.array_equal(k16, k32)
    for q in range(0, m, 29):
        neg = int(np.sum(np.linalg.eigvalsh(a[q]) < 0))
        assert k32[q] == neg
        cond = np.linalg.cond(a[q])
        assert np.abs(a[q] @ i32[q] - np.eye(n)).max() < 1e-14 * cond * n


def test_wide_zero_column_code(eng):
    n = 300
    rng = np.random.default_rng(3)
    a = rng.standard_normal((1, n, n))
    a = 0.5 * (a + np.transpose(a, (0, 2, 1)))
    a[0, 0, :] = 0.0
    a[0, :, 0] = 0.0
    _, k16, _, k32 = _both(eng, a)
    assert k16[0] == k32[0] == -1
