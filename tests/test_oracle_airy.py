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
