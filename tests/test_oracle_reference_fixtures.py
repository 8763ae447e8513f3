"""Pins the oracle against the reference's own example outputs
(examples/output/molscat-all_propagators.output, written as JSON by
tests/golden/extract_reference_fixtures.py).

System: examples/input/molscat-all_propagators.input, ITYPE=1 rigid rotor
(BE=30, j=0,2,4,6), LJ 12-6 isotropic + P2 term, URED=20, JTOT=10, IB=1,
RMIN from FINDRM (IRMSET=9), RMAX=20 RM, STEPS=12, energies 1000 and
900 cm-1 (the second replays the first energy's grid)."""
import json
import os

import numpy as np
import pytest

from oracle import molscat_oracle as O


@pytest.fixture(scope="module")
def system():
    pot = O.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    blk, runit = O.make_itype1_block(10, 1, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)
    grid, rturn = O.ld_grid(blk, 0.5, 20.0, 1000.0 * blk.cm2ru, steps=12.0, irmset=9)
    return blk, runit, grid, rturn


@pytest.fixture(scope="module")
def ref(golden_dir):
    with open(os.path.join(golden_dir, "ref_all_propagators_smat.json")) as f:
        return json.load(f)["sections"]


def test_grid_matches_printed(system, ref):
    blk, runit, grid, rturn = system
    props = ref["model system: ITYPE=1; IPROPS=6"]["props"]
    assert props[0][0] == "MDPROP"
    assert abs(grid.rstart - props[0][1]) < 5e-5  # printed F12.4
    assert grid.nstep == props[0][3] == 8072
    assert abs(rturn - 0.761) < 5e-4  # 'INNER CLASSICAL TURNING POINT AT R = 0.761'


def test_open_channel_order_and_wavevectors(system, ref):
    blk, runit, grid, _ = system
    sm = ref["model system: ITYPE=1; IPROPS=6"]["smat"][0]
    ered = 1000.0 * blk.cm2ru
    wvec, nopen = O.gtwvec(ered, blk.eint)
    nb = O.chnsrt(blk.eint)
    assert nopen == sm["nopen"]
    for i, ch in enumerate(sm["channels"]):
        assert nb[i] + 1 == ch["channel"]
        assert blk.lvals[nb[i]] == ch["L"]
        assert abs(wvec[nb[i]] / runit - ch["wvec_per_ang"]) < 1e-7 * ch["wvec_per_ang"]


@pytest.mark.parametrize("iprop,label,tol", [
    # measured 1.2e-12 and 2.7e-12
    (6, "model system: ITYPE=1; IPROPS=6", 5e-12),
    # MGPROP with IMGSEL = 5 (MA5); measured 3.1e-11 and 3.5e-12, while CS4
    # lands 1e-3 away
    (8, "model system: ITYPE=1; IPROPS=8", 5e-11),
    # MAPROP (LDMA): measured 4.5e-12 and 4.9e-12
    (7, "model system: ITYPE=1; IPROPS=7", 1e-11),
])
@pytest.mark.parametrize("ie,energy", [(0, 1000.0), (1, 900.0)])
def test_smatrix_matches_reference(system, ref, iprop, label, tol, ie, energy):
    blk, runit, grid, _ = system
    out = O.smatrix_ld(blk, energy * blk.cm2ru, grid, iprop=iprop, imgsel=5)
    sm = ref[label]["smat"][ie]
    assert np.abs(out["SR"] - np.array(sm["sr"])).max() < tol
    assert np.abs(out["SI"] - np.array(sm["si"])).max() < tol
    s = out["SR"] + 1j * out["SI"]
    assert np.abs(s @ s.conj().T - np.eye(s.shape[0])).max() < 1e-12
    assert np.abs(s - s.T).max() < 1e-13


def _sdiff(out, sm):
    return max(np.abs(out["SR"] - np.array(sm["sr"])).max(), np.abs(out["SI"] - np.array(sm["si"])).max())


def test_ldprop_matches_reference(system, ref):
    """LDPROP's S is a thousand times more sensitive to the grid's start
    than MDPROP's (1e-5 A of RSTART moves S by 1e-8 against 3e-12), so it
    pins FINDRM: the phase integral starts at the secant iterate before the
    converged turning point (findrm.f:109,179); started at the turning
    point itself, RSTART moves by 1.4e-5 A and S by 1.5e-8.  The second
    energy replays the first energy's matrices with the diagonal shifted
    (ISCRU, ldprop.f:34-38,58-63,107-112); evaluated afresh it lands 6.5e-11
    away, the replay's order of rounding amplified by the recurrence."""
    blk, runit, grid, _ = system
    label = "model system: ITYPE=1; IPROPS=5"
    e1, e2 = 1000.0 * blk.cm2ru, 900.0 * blk.cm2ru
    first = O.smatrix_ld(blk, e1, grid, iprop=5)
    assert _sdiff(first, ref[label]["smat"][0]) < 5e-12  # measured 1.0e-12
    replay = O.smatrix_ld(blk, e2, grid, iprop=5, efirst=e1)
    assert _sdiff(replay, ref[label]["smat"][1]) < 5e-12  # measured 1.1e-12
    afresh = O.smatrix_ld(blk, e2, grid, iprop=5)
    assert _sdiff(afresh, ref[label]["smat"][1]) < 1e-10  # measured 6.5e-11
    assert int(replay["NODES"]) == int(afresh["NODES"])
    # the turning-point start of the phase integral is what the printed
    # values rule out
    g0 = O.Grid(grid.rstart + 1.4129e-5, grid.rstop, grid.nstep, (grid.rstop - grid.rstart - 1.4129e-5) / grid.nstep)
    assert _sdiff(O.smatrix_ld(blk, e1, g0, iprop=5), ref[label]["smat"][0]) > 1e-8


@pytest.mark.parametrize("imgsel,order,nsteps", [(4, 4, (20, 40, 80)), (5, 5, (20, 40))])
def test_mgprop_order_on_a_constant_potential(imgsel, order, nsteps):
    """A single open channel with W = -k^2 constant: Y(R) = -k tan(k(R-R0) -
    atan(Y0/k)) exactly.  CS4 is of fourth order and MA5 of fifth (halving
    the step divides the error by about 16 and 32 until rounding takes
    over); a wrong coefficient breaks the order."""
    k, r0, r1, y0 = 1.3, 0.0, 2.0, 0.4
    hm = lambda r: (np.array([[-k * k]]), None)
    exact = -k * np.tan(k * (r1 - r0) - np.arctan(y0 / k))
    errs = []
    for nstep in nsteps:
        y, _ = O.mgprop(np.array([[y0]]), hm, r0, nstep, (r1 - r0) / nstep, imgsel)
        errs.append(abs(y[0, 0] - exact))
    want = 2.0 ** order
    for e1, e2 in zip(errs[:-1], errs[1:]):
        assert 0.75 * want < e1 / e2 < 1.25 * want, errs
