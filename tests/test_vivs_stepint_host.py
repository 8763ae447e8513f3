"""The VIVS step integrals of the device code (molscat_amd/csrc/
mcp_stepint.hpp; the quantities PERT2 supplies, source_code/pert2.f:36-222)
built for the host and checked against the integrals they stand for,
evaluated at 40 digits by mpmath: over a step h, with c = cos(k r) or
cosh(|k| r) and s = sin(k r)/k or sinh(|k| r)/|k|, the integrals of c_i c_j,
c_i s_j, s_i c_j and s_i s_j, for open, closed and mixed pairs in each of
the three evaluations (both arguments small; apart; same sign and close),
at their boundaries, and for equal and nearly equal k.  Tolerance 4e-15
relative to the largest of the four integrals' scales (h, h^2, h^2, h^3
times the product of the two solutions' sizes).  The same for the integrals
against the weights r - h/2 and (r - h/2)^2/2 of the derivative
perturbations (IVP, IVPP; scales one and two powers of h higher)."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def slib(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("stepint") / "libstepint_h.so")
    subprocess.check_call([hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared", "-Xclang",
                           "-target-feature", "-Xclang", "-packed-fp32-ops", "-o", out,
                           os.path.join(HERE, "native", "stepint_harness.cpp")], stderr=subprocess.DEVNULL)
    lib = ctypes.CDLL(out)
    lib.step_integrals_h.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.POINTER(ctypes.c_double)]
    lib.step_integrals_w_h.argtypes = lib.step_integrals_h.argtypes
    return lib


def _exact(zi, zj, h, p=0):
    import mpmath as mp

    mp.mp.dps = 40
    H = mp.mpf(h)
    w = [lambda r: 1, lambda r: r - H / 2, lambda r: (r - H / 2) ** 2 / 2][p]

    def fns(z):
        k2 = mp.mpf(z) / (H * H)
        if k2 >= 0:
            k = mp.sqrt(k2)
            return (lambda r: mp.cos(k * r)), (lambda r: mp.sin(k * r) / k if k != 0 else r)
        k = mp.sqrt(-k2)
        return (lambda r: mp.cosh(k * r)), (lambda r: mp.sinh(k * r) / k)

    ci, si = fns(zi)
    cj, sj = fns(zj)
    q = lambda f: mp.quad(lambda r: f(r) * w(r), [0, H / 2, H])
    return [float(q(lambda r: ci(r) * cj(r))), float(q(lambda r: ci(r) * sj(r))),
            float(q(lambda r: si(r) * cj(r))), float(q(lambda r: si(r) * sj(r)))]


CASES = [
    # both small: series
    (0.3, 1.7), (-1.2, 0.8), (-2.0, -2.0), (1e-9, -1e-9), (0.0, 0.0), (2.0, 1.99),
    # apart
    (30.0, 12.0), (-25.0, -9.0), (20.0, -15.0), (2.5, -2.5), (400.0, 0.01), (-400.0, 0.5), (3.0, 1.4),
    (-900.0, -100.0), (5000.0, 20.0),
    # same sign and close
    (30.0, 30.0), (30.0, 30.002), (-30.0, -30.002), (-30.0, -30.0), (2.01, 2.0), (2.5, 1.3),
    (900.0, 880.0), (-600.0, -500.0), (400.0, 200.1),
]


@pytest.mark.parametrize("zi,zj", CASES)
def test_step_integrals_against_quadrature(slib, zi, zj):
    h = 0.37
    out = (ctypes.c_double * 4)()
    slib.step_integrals_h(zi, zj, h, out)
    want = _exact(zi, zj, h)

    def size(z):
        x = np.sqrt(abs(z))
        return np.cosh(x) if z < 0 else 1.0

    scale = size(zi) * size(zj)
    scales = [h * scale, h * h * scale, h * h * scale, h ** 3 * scale]
    for name, got, ref, sc in zip(("cc", "cs", "sc", "ss"), out, want, scales):
        assert abs(got - ref) <= 4e-15 * sc, (zi, zj, name, got, ref, abs(got - ref) / sc)


def test_symmetry_under_exchange(slib):
    rng = np.random.default_rng(9)
    a = (ctypes.c_double * 4)()
    b = (ctypes.c_double * 4)()
    for _ in range(300):
        zi, zj = rng.uniform(-50, 50, 2) * rng.choice([1e-3, 1.0, 10.0], 2)
        slib.step_integrals_h(zi, zj, 0.5, a)
        slib.step_integrals_h(zj, zi, 0.5, b)
        sc = max(1.0, abs(a[0]))
        assert abs(a[0] - b[0]) <= 1e-14 * sc and abs(a[3] - b[3]) <= 1e-14 * max(1.0, abs(a[3]))
        assert abs(a[1] - b[2]) <= 1e-14 * max(1.0, abs(a[1])) and abs(a[2] - b[1]) <= 1e-14 * max(1.0, abs(a[2]))


@pytest.mark.parametrize("zi,zj", CASES + [(1e-3, 1.5e-3), (0.0135, 0.0135), (-2.0, 1e-4), (60.0, 58.0)])
def test_weighted_step_integrals_against_quadrature(slib, zi, zj):
    h = 0.37
    out = (ctypes.c_double * 8)()
    slib.step_integrals_w_h(zi, zj, h, out)

    def size(z):
        x = np.sqrt(abs(z))
        return np.cosh(x) if z < 0 else 1.0

    scale = size(zi) * size(zj)
    for p in (1, 2):
        want = _exact(zi, zj, h, p)
        got = out[4 * (p - 1):4 * p]
        scales = [h ** (1 + p) * scale, h ** (2 + p) * scale, h ** (2 + p) * scale, h ** (3 + p) * scale]
        for name, g, ref, sc in zip(("cc", "cs", "sc", "ss"), got, want, scales):
            assert abs(g - ref) <= 4e-15 * sc, (zi, zj, p, name, g, ref, abs(g - ref) / sc)
