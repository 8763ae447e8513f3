"""The device Percival-Seaton coefficients (COUPLE for ITYPE = 1,
source_code/couple.f:51-143; molscat_amd/csrc/mcp_couple.hpp) built for the
host and checked against exact rational arithmetic: the ln(k!) table to
2e-16 relative of ln(k!) (the pair hi + lo against the exact factorial's
logarithm at 40 digits), single coefficients over j <= 40, L <= 160,
JTOT <= 120 and lambda <= 8 to 4e-15 absolute (the coefficients are O(1)),
whole blocks against system.rotor_coupling, and the packed-index decode."""
import ctypes
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from molscat_amd import system as S

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def clib(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("couple") / "libcouple_h.so")
    subprocess.check_call([hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared", "-Xclang",
                           "-target-feature", "-Xclang", "-packed-fp32-ops", "-o", out,
                           os.path.join(HERE, "native", "couple_harness.cpp")], stderr=subprocess.DEVNULL)
    lib = ctypes.CDLL(out)
    lib.couple_table_h.argtypes = [ctypes.c_int]
    lib.couple_need_h.argtypes = [ctypes.c_int] * 4
    lib.percival_seaton_h.argtypes = [ctypes.c_int] * 6
    lib.percival_seaton_h.restype = ctypes.c_double
    lib.lnfact_h.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.packed_col_h.argtypes = [ctypes.c_longlong]
    lib.couple_table_h(2047)
    return lib


def test_lnfact_table(clib):
    from decimal import Decimal, getcontext

    getcontext().prec = 50
    out = (ctypes.c_double * 2)()
    fact = 1
    for k in range(0, 2048):
        if k > 0:
            fact *= k
        if k in (0, 1, 2, 3, 10, 57, 170, 171, 500, 1000, 1500, 2047):
            exact = Decimal(fact).ln() if k > 1 else Decimal(0)
            clib.lnfact_h(k, out)
            got = Decimal(out[0]) + Decimal(out[1])
            assert abs(got - exact) <= Decimal("2e-16") * max(Decimal(1), exact), (k, got, exact)


def test_need_covers_the_arguments(clib):
    # every ln(k!) index the coefficient reads stays within the bound
    assert clib.couple_need_h(40, 160, 120, 8) >= 40 + 160 + 40 + 160 + 1
    assert clib.couple_need_h(40, 160, 120, 8) >= 120 + 40 + 160 + 1


def test_single_coefficients_against_exact(clib):
    rng = np.random.default_rng(12)
    worst = 0.0
    cases = 0
    for _ in range(3000):
        jtot = int(rng.integers(0, 121))
        lam = int(rng.integers(0, 9))
        jc = int(rng.integers(0, 41))
        # mostly pairs the 3j parity allows (j + lambda + j' even), some not
        steps = list(range(-lam, lam + 1, 2)) if rng.random() < 0.9 else list(range(-lam, lam + 1))
        jr = jc + int(rng.choice(steps))
        lc = int(rng.integers(abs(jtot - jc), jtot + jc + 1))
        lr = lc + int(rng.choice(steps))
        if jr < 0 or jr > 40 or lr < 0:
            continue
        want = 0.0
        if abs(jc - jr) <= lam <= jc + jr and abs(lc - lr) <= lam <= lc + lr:
            t = S.threej000(jc, lam, jr)
            u = S.threej000(lc, lam, lr)
            if t and u:
                s6 = S.sixj(jr, lr, jtot, lc, jc, lam)
                want = math.sqrt((2 * jc + 1) * (2 * jr + 1) * (2 * lc + 1) * (2 * lr + 1)) * t * u * s6
                if (jc + jr + jtot) % 2:
                    want = -want
        got = clib.percival_seaton_h(jc, lc, jr, lr, jtot, lam)
        worst = max(worst, abs(got - want))
        cases += want != 0.0
    assert cases > 1000
    assert worst < 4e-15, worst


@pytest.mark.parametrize("jtot,ib,jlev,lams", [(10, 1, [0, 2, 4, 6], [0, 2]), (60, 2, list(range(0, 41, 2)), [0, 2, 4, 6]),
                                               (3, 2, [0, 1, 2, 3, 4, 5], [0, 1, 2, 3])])
def test_block_against_host_coupling(clib, jtot, ib, jlev, lams):
    chans = S.rotor_channels(jtot, ib, jlev)
    if len(chans) > 200:
        chans = chans[:200]       # the exact host coupling is slow; a leading sub-block
    want = S.rotor_coupling(jtot, chans, jlev, lams)
    n = len(chans)
    got = np.zeros_like(want)
    k = 0
    for icol in range(n):
        for irow in range(icol + 1):
            jc, lc = jlev[chans[icol][0]], chans[icol][1]
            jr, lr = jlev[chans[irow][0]], chans[irow][1]
            for il, lam in enumerate(lams):
                got[il, k] = clib.percival_seaton_h(jc, lc, jr, lr, jtot, lam)
            k += 1
    assert np.abs(got - want).max() < 4e-15


def test_packed_column_decode(clib):
    for n in (1, 2, 3, 100, 441, 5000):
        for icol in {0, 1, n // 2, n - 1}:
            for irow in {0, icol // 2, icol}:
                k = icol * (icol + 1) // 2 + irow
                assert clib.packed_col_h(k) == icol
