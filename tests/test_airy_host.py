"""The device Airy sector propagator (molscat_amd/csrc/mcp_airy.hpp) built
for the host and checked against its definition evaluated with mpmath at
40 digits: y1 = a (Bi'1 Ai2 - Ai'1 Bi2)/D, y2 = a/(pi D), y4 = a (Ai1 Bi'2 -
Bi1 Ai'2)/D over seeded sector widths, eigenvalues and slopes spanning the
oscillatory, exponential, turning-point and asymptotic regimes.  Tolerance:
1e-11 relative to the largest of |y1|, |y2|, |y4| (the worst case measured is
~5e-13)."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def airy_lib(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("airy") / "libairy_h.so")
    subprocess.check_call([hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared", "-Xclang",
                           "-target-feature", "-Xclang", "-packed-fp32-ops", "-o", out,
                           os.path.join(HERE, "native", "airy_harness.cpp")],
                          stderr=subprocess.DEVNULL)
    lib = ctypes.CDLL(out)
    p = ctypes.POINTER(ctypes.c_double)
    lib.airy_sector_h.argtypes = [ctypes.c_double] * 3 + [p]
    lib.airy_transfer_h.argtypes = [ctypes.c_double] * 2 + [p]
    return lib


def _exact(w, eig, slope):
    import mpmath as mp

    mp.mp.dps = 40
    al = mp.cbrt(mp.mpf(slope)) if slope > 0 else -mp.cbrt(-mp.mpf(slope))
    be = -mp.mpf(eig) / slope
    x1 = al * (be - mp.mpf(w) / 2)
    x2 = al * (be + mp.mpf(w) / 2)
    a1, b1, a2, b2 = mp.airyai(x1), mp.airybi(x1), mp.airyai(x2), mp.airybi(x2)
    ap1, bp1, ap2, bp2 = mp.airyai(x1, 1), mp.airybi(x1, 1), mp.airyai(x2, 1), mp.airybi(x2, 1)
    d = a1 * b2 - a2 * b1
    return [float(al * (bp1 * a2 - ap1 * b2) / d), float(al / (mp.pi * d)), float(al * (a1 * bp2 - b1 * ap2) / d)]


def test_sector_propagator_against_mpmath(airy_lib):
    rng = np.random.default_rng(3)
    cases = [(0.01, 5.0, 0.3), (0.2, -3.0, 2.0), (0.05, 0.0, 1.0), (1e-3, 1e-4, 1e-8), (0.5, 7700.0, 5e-4),
             (0.1, -50.0, -3.0), (0.3, 1e-6, 10.0)]
    for _ in range(300):
        cases.append((10 ** rng.uniform(-3, 0), rng.choice([-1, 1]) * 10 ** rng.uniform(-4, 4),
                      rng.choice([-1, 1]) * 10 ** rng.uniform(-8, 4)))
    worst = 0.0
    for w, eig, slope in cases:
        o = (ctypes.c_double * 3)()
        airy_lib.airy_sector_h(w, eig, slope, o)
        ex = _exact(w, eig, slope)
        sc = max(abs(x) for x in ex)
        err = max(abs(o[k] - ex[k]) for k in range(3)) / sc
        worst = max(worst, err)
        assert err < 1e-11, (w, eig, slope, list(o), ex)
    print("worst relative error", worst)


def test_transfer_matrix_is_unimodular_and_matches_airy(airy_lib):
    import mpmath as mp

    mp.mp.dps = 30
    for x1, dl in [(-12.0, 0.7), (0.3, 2.5), (15.0, -4.0), (-1.0, 1e-7)]:
        o = (ctypes.c_double * 4)()
        airy_lib.airy_transfer_h(x1, dl, o)
        p11, p12, p21, p22 = o
        assert abs(p11 * p22 - p12 * p21 - 1.0) < 1e-12 * max(1.0, abs(p11 * p22))
        # Phi maps (Ai, Ai')(x1) to (Ai, Ai')(x1 + delta)
        a = mp.airyai(x1), mp.airyai(x1, 1)
        b = mp.airyai(x1 + dl), mp.airyai(x1 + dl, 1)
        got = p11 * float(a[0]) + p12 * float(a[1])
        assert abs(got - float(b[0])) < 1e-12 * max(abs(float(b[0])), abs(p11 * float(a[0])))


def test_potential_table_is_the_scalar_evaluation():
    """PowerPotential.table (the host P tables of every grid) gives the
    per-radius evaluation, powers, exponentials and the 1e-8 floor for
    negative powers included, to 2 ulp (numpy's vectorised pow rounds the
    last bit differently from the scalar libm pow on some arguments)."""
    from molscat_amd import decks
    from molscat_amd import system as S

    pots = [decks.rotor441().potential(), decks.field48().potl,
            S.PowerPotential([[(2.0, -6), (0.5, 0, -1.3)], [(1.5, 1), (-0.25, -12)]], rm=1.0, epsil=1.0)]
    x = np.concatenate([[0.0, 1e-9, 2e-8], np.linspace(0.3, 40.0, 997)])
    for pot in pots:
        tab = pot.table(x)
        ref = np.stack([pot(v) for v in x])
        assert np.all(np.abs(tab - ref) <= 4.5e-16 * np.abs(ref))
