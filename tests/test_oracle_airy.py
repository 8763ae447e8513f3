"""Pins the oracle's AIRY propagator (AIPROP with TOLHI = 0; its Airy
sector propagator evaluated from the Airy functions with mpmath) against
the reference's own example output: the 'IPROPS=9, IPROPL=9' run of
examples/output/molscat-all_propagators.output (ITYPE=1 rotor, JTOT=10,
IB=1; AIRY with constant steps from RSTART to RMID=1.0 and steps growing as
R^(4/3) to RMAX=20).  The step counts are the reference's (149, 791); the
S-matrix at 1000 cm-1 agrees to 2.3e-10 (tolerance 5e-10: the printed
values carry 14 digits, the remaining difference is the reference's own
Airy function accuracy, SCAIRY tables and asymptotic series).  Also pins
the sector propagator on closed forms: a constant W (slope -> 0) gives the
free-propagation imbedding elements."""
import json
import math
import os

import numpy as np
import pytest

from oracle import molscat_oracle as O


def test_airy_sector_constant_potential_limit():
    # W = -k^2 constant: y1 = y4 = k cot(k w), y2 = k / sin(k w)
    k, w = 1.7, 0.3
    y1, y2, y4 = O.airy_sector(w, np.array([k * k]), np.array([1e-12]))
    assert abs(y1[0] - k / math.tan(k * w)) < 1e-8
    assert abs(y4[0] - k / math.tan(k * w)) < 1e-8
    assert abs(abs(y2[0]) - k / math.sin(k * w)) < 1e-8
    # closed channel W = +kappa^2: kappa coth, kappa / sinh
    kap = 2.2
    y1, y2, y4 = O.airy_sector(w, np.array([-kap * kap]), np.array([1e-12]))
    assert abs(y1[0] - kap / math.tanh(kap * w)) < 1e-8
    assert abs(abs(y2[0]) - kap / math.sinh(kap * w)) < 1e-8


@pytest.mark.slow
def test_airy_smatrix_matches_reference(golden_dir):
    with open(os.path.join(golden_dir, "ref_all_propagators_smat.json")) as f:
        sec = json.load(f)["sections"]["model system: ITYPE=1; IPROPS=9, IPROPL=9"]
    pot = O.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    blk, _ = O.make_itype1_block(10, 1, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)
    eredmx = 1000.0 * blk.cm2ru
    rstart, _ = O.findrm(blk, 0.5, eredmx, 9)
    cays = O.calck(0.0, eredmx, blk.eint)
    n1, dr1 = O.drset(rstart, 1.0, 12.0, 0.0, cays, O.UNSET, 9, 0.0)
    n2, dr2 = O.drset(1.0, 20.0, 12.0, 0.0, cays, O.UNSET, 9, 4.0 / 3.0)
    assert [p[3] for p in sec["props"][:2]] == [n1, n2]
    ered = 1000.0 * blk.cm2ru
    y = O.yinit(blk, ered, rstart)
    y, _ = O.aiprop(y, blk, ered, rstart, 1.0, n1, dr1, 0.0)
    y, _ = O.aiprop(y, blk, ered, 1.0, 20.0, n2, dr2, 4.0 / 3.0)
    wvec, nopen = O.gtwvec(ered, blk.eint)
    k = O.ytok(O.chnsrt(blk.eint), wvec, blk.n, nopen, y, 20.0, blk.cent)
    sr, si = O.ktos(k)
    ref = sec["smat"][0]
    assert np.abs(sr - np.array(ref["sr"])).max() < 5e-10
    assert np.abs(si - np.array(ref["si"])).max() < 5e-10


@pytest.fixture(scope="module")
def tolhi_run(golden_dir):
    """The 'IPROPS=6, IPROPL=9' run: MDPROP from RSTART to RMID = 1.4 RTURN
    (177 steps), then AIRY with the default TOLHIL = 1e-4 and POWRL = POWRX =
    3 to RMAX = 20, the first step DRSET's DRFIT = pi / (CAY * STEPS); 900
    cm-1 replays the steps of 1000 cm-1 (ISCRU = 30)."""
    with open(os.path.join(golden_dir, "ref_all_propagators_smat.json")) as f:
        sec = json.load(f)["sections"]["model system: ITYPE=1; IPROPS=6, IPROPL=9"]
    pot = O.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    blk, _ = O.make_itype1_block(10, 1, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)
    eredmx = 1000.0 * blk.cm2ru
    rstart, rturn = O.findrm(blk, 0.5, eredmx, 9)
    rmid = 1.4 * rturn
    cays = O.calck(0.0, eredmx, blk.eint)
    n1, dr1 = O.drset(rstart, rmid, 12.0, 0.0, cays, O.UNSET, 6, 0.0)
    _, dr2 = O.drset(rmid, 20.0, 12.0, 1e-4, cays, O.UNSET, 9, 3.0)
    ld = O.Grid(rstart, rmid, n1, dr1)
    first = O.smatrix_md_airy(blk, 1000.0 * blk.cm2ru, ld, 20.0, dr2)
    second = O.smatrix_md_airy(blk, 900.0 * blk.cm2ru, ld, 20.0, dr2, replay=first["STEPS"])
    return sec, blk, ld, dr2, first, second


@pytest.mark.slow
def test_airy_tolhi_smatrix_matches_reference(tolhi_run):
    sec, blk, ld, dr2, first, second = tolhi_run
    props = sec["props"]
    assert props[0][3] == ld.nstep == 177
    assert abs(dr2 - math.pi / (O.calck(0.0, 1000.0 * blk.cm2ru, blk.eint) * 12.0)) < 1e-18
    assert props[1][0] == "AIPROP" and props[1][3] == len(first["STEPS"]) == len(second["STEPS"]) == 226
    # 5e-10 as for TOLHI = 0: 14 printed digits and the reference's own Airy
    # function accuracy (measured 2.0e-10 and 1.8e-10)
    for out, sm in ((first, sec["smat"][0]), (second, sec["smat"][1])):
        assert np.abs(out["SR"] - np.array(sm["sr"])).max() < 5e-10
        assert np.abs(out["SI"] - np.array(sm["si"])).max() < 5e-10


@pytest.mark.slow
def test_airy_tolhi_second_energy_is_the_replay(tolhi_run):
    """900 cm-1 on its own step control lands 1.8e-6 from the printed
    S-matrix: the printed one is the replay of the first energy's steps."""
    sec, blk, ld, dr2, first, second = tolhi_run
    own = O.smatrix_md_airy(blk, 900.0 * blk.cm2ru, ld, 20.0, dr2)
    assert np.abs(own["SR"] - np.array(sec["smat"][1]["sr"])).max() > 1e-7


def test_airy_tolhi_linear_potential_doubles_the_step():
    """A potential linear in R: CORR's diagonal estimate (the W'' implied by
    the eigenvalue change and slope) vanishes and, uncoupled, so does COFF;
    every step then grows by the cap, 2, until the last one is cut at
    RSTOP."""
    n = 2
    vl = np.array([[1.0, 0.0, 0.5]])  # diag(1, 0.5) x the radial term
    pot = O.PowerPotential([[(3.0, 1)]], rm=1.0, epsil=1.0)
    blk = O.Block(n=n, vl=vl, eint=np.zeros(n), cent=np.zeros(n), lvals=np.zeros(n, dtype=int),
                  levidx=np.arange(n), potl=pot, rscale=1.0, ep2ru=1.0, cm2ru=1.0)
    z0 = np.diag([1.0, 2.0])
    _, _, steps = O.aiprop_tolhi(z0, blk, 50.0, 1.0, 9.0, 0.01, 3.0, 1e-4)
    dr = [s[1] for s in steps]
    for a, b in zip(dr[:-2], dr[1:-1]):
        assert abs(b - 2.0 * a) < 1e-12 * b
    assert abs(steps[-1][0] + 0.5 * steps[-1][1] - 9.0) < 1e-12
