"""The oracle's VIVS propagator (source_code/vvprop.f, PERT2 source_code/
pert2.f, DELRD source_code/delrd.f; the default options IV, IPERT, ISYM,
IALPHA = 6) pinned by:
- the reference's own example output: the 'IPROPS=5, IPROPL=4' run of
  examples/output/molscat-all_propagators.output (ITYPE=1 rotor, JTOT=10,
  IB=1; LDPROP from RSTART to RMID = 1.4 RTURN in 177 steps, then VVPROP to
  RMAX = 20 in 1830 steps of its own step control; 900 cm-1 replays the
  steps of 1000 cm-1 from the scratch file, ISCRU = 30);
- PERT2's step integrals against quadrature of the integrals they stand
  for, V0_ij times the integrals over the step of c_i c_j, c_i s_j and
  s_i s_j (c = cos or cosh, s = sin/lambda or sinh/lambda), and the EYE
  accumulations as the integrals of g_i V0_ij g_j, in every branch of PERT2
  (closed form, the series about XI = XJ for open and for closed pairs, the
  series about XSQ = 0, the diagonal);
- a constant coupled W, where the interval's local basis diagonalises W at
  every step, the perturbation vanishes, and the R-matrix must follow the
  exact solution in W's eigenchannels;
- DELRD's growth limits."""
import json
import math
import os

import numpy as np
import pytest
from scipy.integrate import quad

from oracle import molscat_oracle as O

EPS = np.finfo(np.float64).eps


def _ref_fns(dif, h):
    lam = math.sqrt(abs(dif))
    if dif >= 0.0:
        return (lambda r: math.cos(lam * r)), (lambda r: math.sin(lam * r) / lam)
    return (lambda r: math.cosh(lam * r)), (lambda r: math.sinh(lam * r) / lam)


def _pert2_inputs(difs, h):
    difs = np.asarray(difs, dtype=np.float64)
    x = np.sqrt(np.abs(difs)) * h
    op = difs >= 0.0
    cx = np.where(op, np.cos(x), np.cosh(x))
    sine = np.where(op, np.sin(x), -np.sinh(x))
    return cx, sine, difs * h * h, x


H = 0.37
PERT2_CASES = {
    "closed form, open pair": [30.0, 12.0],
    "closed form, closed pair": [-25.0, -9.0],
    "closed form, mixed pair": [20.0, -15.0],
    "closed form, mixed pair with XI ~ XJ": [(2.0 / H) ** 2, -(2.0008 / H) ** 2],
    "series XI ~ XJ, open": [30.0, 30.0 + 2e-3],
    "series XI ~ XJ, closed": [-30.0, -30.0 - 2e-3],
    "series XSQ ~ 0, same sign": [1e-3, 1e-3 + 1e-6],
    "series XSQ ~ 0, opposite sign": [1e-4, -1.1e-4],
}


def _branch(x, xsq):
    if abs(x[0] - x[1]) > 1e-3:
        return 130
    if abs(xsq[0] + xsq[1]) < 1e-3:
        return 150
    if (xsq[0] < 0 and xsq[1] < 0) or xsq[0] * xsq[1] >= 0:
        return 140
    return 130


@pytest.mark.parametrize("case", list(PERT2_CASES))
def test_pert2_step_integrals_match_quadrature(case):
    difs = PERT2_CASES[case]
    cx, sx, xsq, x = _pert2_inputs(difs, H)
    want = {"closed": 130, "series XI": 140, "series XSQ": 150}
    assert _branch(x, xsq) == next(v for k, v in want.items() if case.startswith(k))
    v0 = np.array([[0.7, -0.4], [-0.4, 0.3]])
    a1, a1p = np.array([0.2, -0.5]), np.array([1.1, 0.6])
    a2, a2p = np.array([0.9, 1.3]), np.array([-0.3, 0.4])
    e11, e12, e22 = (np.full((2, 2), 0.25) for _ in range(3))
    cc, cs, ss = O.pert2_v0(cx, sx, xsq, x, H, v0, a1, a2, a1p, a2p, e11, e12, e22)
    # PERT2's closed forms cancel terms of size H^3/X^2 (CS) and H^3/X^4 (SS)
    xm = float(x.min())
    tol_cc = 1e-13
    tol_cs = 1e-13 + 100.0 * EPS / xm ** 2
    tol_ss = 1e-13 + 10.0 * EPS / xm ** 4
    fns = [_ref_fns(d, H) for d in difs]

    def q(f):
        return quad(f, 0.0, H, epsabs=0.0, epsrel=1e-13, limit=200)[0]

    for i in range(2):
        for j in range(2):
            ci, si = fns[i]
            cj, sj = fns[j]
            g1i = lambda r: a1p[i] * si(r) + a1[i] * ci(r)
            g1j = lambda r: a1p[j] * sj(r) + a1[j] * cj(r)
            g2i = lambda r: a2p[i] * si(r) + a2[i] * ci(r)
            g2j = lambda r: a2p[j] * sj(r) + a2[j] * cj(r)
            v = v0[i, j]
            checks = [
                (cc[i, j], v * q(lambda r: ci(r) * cj(r)), tol_cc),
                (cs[i, j], v * q(lambda r: ci(r) * sj(r)), tol_cs),
                (ss[i, j], v * q(lambda r: si(r) * sj(r)), tol_ss),
                (e11[i, j] - 0.25, v * q(lambda r: g1i(r) * g1j(r)), tol_ss),
                (e12[i, j] - 0.25, v * q(lambda r: g1i(r) * g2j(r)), tol_ss),
                (e22[i, j] - 0.25, v * q(lambda r: g2i(r) * g2j(r)), tol_ss),
            ]
            for got, ref, tol in checks:
                assert abs(got - ref) <= tol * abs(ref), (case, i, j, got, ref)


def test_delrd_growth_limits():
    # no corrections: the growth factor is the cap for the channel situation
    assert O.delrd(1.0, 0.0, 0.0, 0.02, 5.0, 1.0, 2.0, 1.0, 20.0) == 1.6  # all open, R <= RMAX/10
    assert O.delrd(1.0, 0.0, 0.0, 0.02, 5.0, 1.0, 2.0, 3.0, 20.0) == 2.5  # all open, beyond
    assert abs(O.delrd(1.0, 0.0, 0.0, 0.02, 50.0, 1.0, -1.0, 3.0, 20.0) - 1.2) < 1e-15  # some closed
    # all closed: 1.15, capped at 4 / sqrt(-EN)
    assert abs(O.delrd(1.0, 0.0, 0.0, 0.02, 50.0, -1.0, -0.01, 3.0, 20.0) - 1.15) < 1e-15
    assert O.delrd(1.0, 0.0, 0.0, 0.02, 50.0, -1.0, -100.0, 3.0, 20.0) == 0.4
    # a correction of the tolerance's size scales by (0.8 TOL / C)^0.333
    f = O.delrd(1.0, 0.016, 0.0, 0.02, 5.0, 1.0, 2.0, 3.0, 20.0)
    assert abs(f - 1.0) < 1e-15
    f = O.delrd(1.0, 0.0, 0.128, 0.02, 5.0, 1.0, 2.0, 3.0, 20.0)
    assert abs(f - 0.125 ** 0.333) < 1e-15
    assert O.delrd(4.0, 0.0, 0.0, 0.02, 5.0, 1.0, 2.0, 3.0, 20.0) == 5.0  # DRMAX
    with pytest.raises(RuntimeError):
        O.delrd(1e-9, 1.0, 1.0, 0.02, 5.0, 1.0, 2.0, 3.0, 20.0)


def _constant_block(w, eint):
    """A block whose W is R-independent: three L = 0 channels coupled by a
    constant matrix (POTENL's exponential form with exponent 0)."""
    n = w.shape[0]
    vl = np.array([[w[i, j] for j in range(n) for i in range(j + 1)]])
    pot = O.PowerPotential([[(1.0, 0, 0.0)]], rm=1.0, epsil=1.0)
    return O.Block(n=n, vl=vl, eint=eint, cent=np.zeros(n), lvals=np.zeros(n, dtype=int),
                   levidx=np.arange(n), potl=pot, rscale=1.0, ep2ru=1.0, cm2ru=1.0)


def test_constant_coupling_follows_exact_solution():
    rng = np.random.default_rng(11)
    n = 3
    m = rng.standard_normal((n, n))
    wc = 0.5 * (m + m.T)
    eint = np.array([0.0, 2.0, 9.0])
    blk = _constant_block(wc, eint)
    ered = 4.0
    w, _ = blk.hammat(1.0, ered)
    lam, u = np.linalg.eigh(w)
    r0, r1 = 1.0, 7.3
    y0c = np.array([0.7, -1.3, 2.1])
    rmat0 = u @ np.diag(1.0 / y0c) @ u.T
    rmat, steps = O.vvprop(rmat0, blk, ered, r0, r1, 0.05, *O.vivs_tolerances(1e-6))
    assert abs(steps[-1].rnow - r1) < 1e-12 and steps[-1].last
    assert len(steps) % O.VIVS_IALPHA == 0 or steps[-1].last
    d = r1 - r0
    yexact = np.empty(n)
    for c in range(n):
        if lam[c] < 0.0:
            k = math.sqrt(-lam[c])
            yexact[c] = (-k * math.sin(k * d) + y0c[c] * math.cos(k * d)) / (math.cos(k * d) + y0c[c] * math.sin(k * d) / k)
        else:
            k = math.sqrt(lam[c])
            yexact[c] = (k * math.sinh(k * d) + y0c[c] * math.cosh(k * d)) / (math.cosh(k * d) + y0c[c] * math.sinh(k * d) / k)
    rex = u @ np.diag(1.0 / yexact) @ u.T
    assert np.abs(rmat - rex).max() < 1e-10 * np.abs(rex).max()


@pytest.fixture(scope="module")
def fixture_run(golden_dir):
    with open(os.path.join(golden_dir, "ref_all_propagators_smat.json")) as f:
        sec = json.load(f)["sections"]["model system: ITYPE=1; IPROPS=5, IPROPL=4"]
    pot = O.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    blk, _ = O.make_itype1_block(10, 1, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)
    eredmx = 1000.0 * blk.cm2ru
    rstart, rturn = O.findrm(blk, 0.5, eredmx, 9)
    rmid = 1.4 * rturn  # RVFAC
    cays = O.calck(0.0, eredmx, blk.eint)
    nstep, dr = O.drset(rstart, rmid, 12.0, 0.0, cays, O.UNSET, 5, 0.0)
    grid = O.Grid(rstart, rmid, nstep, dr)
    first = O.smatrix_ld_vivs(blk, 1000.0 * blk.cm2ru, grid, 20.0)
    second = O.smatrix_ld_vivs(blk, 900.0 * blk.cm2ru, grid, 20.0, replay=first["STEPS"])
    return sec, blk, grid, first, second


@pytest.mark.slow
def test_vivs_smatrices_match_reference(fixture_run):
    sec, blk, grid, first, second = fixture_run
    props = sec["props"]
    assert props[0][0] == "LDPROP" and props[0][3] == grid.nstep
    assert abs(props[0][2] - grid.rstop) < 5e-4
    assert props[1][0] == "VVPROP" and props[1][3] == len(first["STEPS"]) == len(second["STEPS"])
    # measured 1.0e-12 and 1.5e-12
    for out, sm in ((first, sec["smat"][0]), (second, sec["smat"][1])):
        assert out["NOPEN"] == sm["nopen"]
        assert np.abs(out["SR"] - np.array(sm["sr"])).max() < 5e-12
        assert np.abs(out["SI"] - np.array(sm["si"])).max() < 5e-12
        s = out["SR"] + 1j * out["SI"]
        assert np.abs(s @ s.conj().T - np.eye(s.shape[0])).max() < 1e-12


@pytest.mark.slow
def test_second_energy_needs_the_replayed_steps(fixture_run):
    """900 cm-1 on its own step control takes other steps and lands far
    from the printed S-matrix: the printed one is the replay."""
    sec, blk, grid, first, second = fixture_run
    own = O.smatrix_ld_vivs(blk, 900.0 * blk.cm2ru, grid, 20.0)
    assert len(own["STEPS"]) != len(second["STEPS"])
    assert np.abs(own["SR"] - np.array(sec["smat"][1]["sr"])).max() > 1e-6
