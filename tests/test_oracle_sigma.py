"""The oracle's SPROC cross sections (source_code/sproc.f:380-430, the sum
over (JTOT, IB) blocks accumulated into ACCSIG and printed times JSTEP,
sproc.f:524-530, 1190-1239) pinned by the reference's own example output:
examples/output/molscat-basic1.output (deck examples/input/molscat-
basic1.input: ITYPE=1 rotor BE=30, j=0,2,4,6, URED=20, LJ + P2 term,
MDPROP with DR=0.001 from FINDRM's RSTART to RMAX=20 RM, E=1250 cm-1, JTOT =
10 and 20 with JSTEP = 10, both parities).  The printed values carry 7
significant digits; detailed balance between the printed up and down
transitions holds to that precision too."""
import json
import os

import numpy as np
import pytest

from oracle import molscat_oracle as O


def _basic1_blocks():
    pot = O.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    for jtot in (10, 20):
        for ib in (1, 2):
            blk, runit = O.make_itype1_block(jtot, ib, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)
            if blk.n > 0:
                yield jtot, ib, blk, runit


@pytest.fixture(scope="module")
def printed(golden_dir):
    with open(os.path.join(golden_dir, "ref_basic1_sigma.json")) as f:
        d = json.load(f)
    sig = np.zeros((3, 3))
    for e in d["sigma"]:
        sig[e["f"] - 1, e["i"] - 1] = e["sigma"]
    return sig


@pytest.mark.slow
def test_partial_wave_sum_matches_reference(printed):
    energy = 1250.0
    jlevels = [0, 2, 4, 6]
    acc = np.zeros((3, 3))
    for jtot, ib, blk, runit in _basic1_blocks():
        ered = energy * blk.cm2ru
        grid, _ = O.ld_grid(blk, 0.5, 20.0, ered, drs=0.001, irmset=9)
        out = O.smatrix_ld(blk, ered, grid, iprop=6)
        acc += O.sigma_block(out["SR"], out["SI"], out["NB"], out["NOPEN"], blk.levidx, jlevels, out["WVEC"],
                             jtot, runit, 3, wght=10.0)
    # 5e-6: the printed values carry 6-7 significant digits (measured 7.5e-7)
    rel = np.abs(acc - printed) / np.abs(printed)
    assert rel.max() < 5e-6, rel


def test_printed_cross_sections_obey_detailed_balance(printed):
    # (2j_i + 1) k_i^2 sigma(i -> f) = (2j_f + 1) k_f^2 sigma(f -> i)
    e = 1250.0
    elev = [30.0 * j * (j + 1) for j in (0, 2, 4)]
    g = [1.0, 5.0, 9.0]
    for i in range(3):
        for f in range(3):
            lhs = g[i] * (e - elev[i]) * printed[f, i]
            rhs = g[f] * (e - elev[f]) * printed[i, f]
            assert abs(lhs - rhs) <= 2e-6 * max(abs(lhs), abs(rhs))


def test_host_findrm_matches_oracle():
    """The driver's FINDRM (molscat_amd.system.findrm, fed the lowest
    adiabat) against the oracle's on the basic1 blocks, the lowest adiabat
    from numpy here (the driver takes it from the device)."""
    from molscat_amd import system as S

    pot_s = S.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    for jtot, ib, blk, runit in _basic1_blocks():
        hb = S.rotor_block(jtot, ib, [0, 2, 4, 6], 30.0, [0, 2], pot_s, 20.0)
        ered = 1250.0 * hb.cm2ru
        assert abs(ered - 1250.0 * blk.cm2ru) < 1e-12 * ered
        w = S.hammat_host(hb, 0.9, ered)
        assert np.abs(w - blk.hammat(0.9, ered)[0]).max() < 1e-12 * np.abs(w).max()
        lowest = lambda r: float(np.linalg.eigvalsh(S.hammat_host(hb, r, ered))[0])
        rs, rt = S.findrm(lowest, 0.5, hb.eint, hb.cent, ered, 9)
        ors, ort = O.findrm(blk, 0.5, ered, 9)
        assert abs(rs - ors) < 1e-12 and abs(rt - ort) < 1e-12
