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


# ---------------------------------------------------------------------------
# the derivative perturbations (IVP, IVPP, IDIAG, ISHIFT), DERMAT and the
# step-control options (IALPHA <= 0, IALFP, ISYM, IPERT)
# ---------------------------------------------------------------------------
PERT2_W_CASES = dict(PERT2_CASES)
# both channels near threshold with close x: the series about XSQ = 0 at
# an XSQ where its double-precision evaluation keeps ~1e-10 of scale
PERT2_W_CASES["series XSQ ~ 0, moderate"] = [0.01, -0.0105]


@pytest.mark.parametrize("case", list(PERT2_W_CASES))
def test_pert2_derivative_terms_match_quadrature(case):
    """The V1 (r - h/2) and V2 (r - h/2)^2/2 parts of PERT2's integrals
    (V1, V2 the first and second derivatives of W at the step centre)
    against quadrature of the integrals they stand for, in every branch.
    Error measured against each integral's scale (|V| h^(1+p), h^(2+p),
    h^(3+p) for cc, cs, ss with p = 1, 2): PERT2's expressions divide by
    F6 ~ X^4 and, for the diagonal V2 terms, by XSQ once more, so they keep
    EPS/X^4 and EPS/X^6 of scale near threshold; the series about XSQ = 0
    stops at second order in XSQ (truncation ~XSQ^2 of scale)."""
    difs = PERT2_W_CASES[case]
    cx, sx, xsq, x = _pert2_inputs(difs, H)
    if case.endswith("moderate"):
        assert _branch(x, xsq) == 150
    v1 = np.array([[0.5, -0.8], [-0.8, 1.2]])
    v2 = np.array([[-0.6, 0.9], [0.9, 0.4]])
    z = np.zeros(2)
    xm = float(x.min())
    trunc = 10.0 * float(np.abs(xsq).max()) ** 2 if _branch(x, xsq) == 150 else 0.0
    fns = [_ref_fns(d, H) for d in difs]
    for p, kw, w, tol in ((1, dict(v1=v1), lambda r: r - H / 2, 1e-13 + 10.0 * EPS / xm ** 4 + trunc),
                          (2, dict(v2=v2), lambda r: 0.5 * (r - H / 2) ** 2, 1e-13 + 10.0 * EPS / xm ** 6 + trunc)):
        e = [np.zeros((2, 2)) for _ in range(3)]
        cc, cs, ss = O.pert2(cx, sx, xsq, x, H, np.zeros((2, 2)), z, z, z, z, *e, **kw)
        vv = v1 if p == 1 else v2

        def q(f):
            return quad(lambda r: f(r) * w(r), 0.0, H, epsabs=0.0, epsrel=1e-13, limit=200)[0]

        for i in range(2):
            for j in range(2):
                ci, si = fns[i]
                cj, sj = fns[j]
                for got, f, sc in ((cc[i, j], lambda r: ci(r) * cj(r), H ** (1 + p)),
                                   (cs[i, j], lambda r: ci(r) * sj(r), H ** (2 + p)),
                                   (ss[i, j], lambda r: si(r) * sj(r), H ** (3 + p))):
                    ref = vv[i, j] * q(f)
                    assert abs(got - ref) <= tol * abs(vv[i, j]) * sc, (case, p, i, j, got, ref)


def test_pert2_series_branch_sign_of_the_reference():
    """The reference's series about XSQ = 0 (pert2.f:175) writes R2V1 with
    (XSQI - XSQJ): its CS and CS' for the V1 term then carry the wrong sign
    (about -h^3/12 where the integral is +h^3/12); the oracle's sign matches
    the integrals."""
    difs = [0.01, -0.0105]
    cx, sx, xsq, x = _pert2_inputs(difs, H)
    assert _branch(x, xsq) == 150
    v1 = np.array([[0.0, 1.0], [1.0, 0.0]])
    z = np.zeros(2)
    fi, fj = _ref_fns(difs[0], H), _ref_fns(difs[1], H)
    ref = quad(lambda r: fi[0](r) * fj[1](r) * (r - H / 2), 0.0, H, epsabs=0.0, epsrel=1e-13)[0]
    out = {}
    for lit in (False, True):
        e = [np.zeros((2, 2)) for _ in range(3)]
        _, cs, _ = O.pert2(cx, sx, xsq, x, H, np.zeros((2, 2)), z, z, z, z, *e, v1=v1, ref_sign_150=lit)
        out[lit] = cs[0, 1]
    # corrected: within the series' truncation (XSQ^2 ~ 2e-6 of scale)
    assert abs(out[False] - ref) < 1e-5 * H ** 3
    # as written: near -h^3/12 instead of +h^3/12
    assert out[True] < 0.0 < ref and abs(out[True] - ref) > 0.1 * H ** 3


def test_pert1_keeps_the_second_derivative_on_the_diagonal_only():
    """PERT1 (IVP without IVPP, source_code/pert1.f): the V2 terms only on
    the diagonal (pert1.f:160-187), i.e. PERT2 with V2's off-diagonal part
    zero."""
    difs = [30.0, 12.0, -9.0]
    cx, sx, xsq, x = _pert2_inputs(difs, H)
    rng = np.random.default_rng(4)
    m1, m2 = rng.standard_normal((2, 3, 3))
    v0, v1, v2 = 0.5 * (m1 + m1.T), 0.5 * (m2 + m2.T), np.diag([0.3, -0.7, 1.1]) + 0.2 * np.ones((3, 3))
    a = [rng.standard_normal(3) for _ in range(4)]
    e1 = [np.zeros((3, 3)) for _ in range(3)]
    e2 = [np.zeros((3, 3)) for _ in range(3)]
    r1 = O.pert2(cx, sx, xsq, x, H, v0, *a, *e1, v1=v1, v2=v2, pert1=True)
    r2 = O.pert2(cx, sx, xsq, x, H, v0, *a, *e2, v1=v1, v2=np.diag(np.diag(v2)))
    for p, q in zip(list(r1) + e1, list(r2) + e2):
        assert np.array_equal(p, q)


def _rotor_block(rscale=1.0):
    pot = O.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    blk, _ = O.make_itype1_block(10, 1, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)
    if rscale != 1.0:
        blk = O.Block(n=blk.n, vl=blk.vl, eint=blk.eint, cent=blk.cent, lvals=blk.lvals, levidx=blk.levidx,
                      potl=blk.potl, rscale=rscale, ep2ru=blk.ep2ru, cm2ru=blk.cm2ru)
    return blk


@pytest.mark.parametrize("r", [0.9, 1.3, 3.0])
def test_dermat_is_the_derivative_of_w(r):
    """DERMAT (source_code/dermat.f) against fourth-order central
    differences of HAMMAT's W(R) (step 1e-3; truncation ~1e-12, rounding
    ~EPS/h^2 of W's scale)."""
    blk = _rotor_block()
    h = 1e-3
    f = {k: blk.hammat(r + k * h, 0.0)[0] for k in (-2, -1, 0, 1, 2)}
    d1 = (8.0 * (f[1] - f[-1]) - (f[2] - f[-2])) / (12.0 * h)
    d2 = (-f[2] + 16.0 * f[1] - 30.0 * f[0] + 16.0 * f[-1] - f[-2]) / (12.0 * h * h)
    w1, w2 = blk.dermat(r, 1), blk.dermat(r, 2)
    assert np.abs(w1 - d1).max() < 1e-8 * np.abs(d1).max()
    assert np.abs(w2 - d2).max() < 1e-6 * np.abs(d2).max()
    assert np.allclose(w1, w1.T, rtol=0, atol=0) and np.allclose(w2, w2.T, rtol=0, atol=0)


def test_dermat_takes_the_potential_derivative_in_the_scaled_distance():
    """With RSCALE != 1 DERMAT differentiates the radial terms with respect
    to R*RSCALE (its POTENL call, dermat.f:69-80) and the centrifugal term
    with respect to R: the potential part is d/dR / RSCALE."""
    blk = _rotor_block(rscale=2.0)
    r, h = 0.7, 1e-4
    cent = lambda rr: np.diag(blk.cent / (rr * rr))
    fp = lambda rr: blk.hammat(rr, 0.0)[0] - cent(rr) - np.diag(blk.eint)
    d1 = (fp(r + h) - fp(r - h)) / (2.0 * h)
    w1 = blk.dermat(r, 1) - np.diag(-2.0 * blk.cent / r ** 3)
    assert np.abs(w1 - d1 / 2.0).max() < 1e-6 * np.abs(d1).max()


@pytest.fixture(scope="module")
def options_run():
    """R from 1.1 to 8 RM on the rotor deck (N = 12, 1000 cm-1) under
    several option sets, from one R(1.1); the converged S at 8 RM from MDPROP
    in 12000 sectors (6000 agree to 8e-10)."""
    blk = _rotor_block()
    ered = 1000.0 * blk.cm2ru
    r0, r1 = 1.1, 8.0
    y0 = O.yinit(blk, ered, 0.75)
    ymid, _ = O.mdprop(y0, lambda r: blk.hammat(r, ered), 0.75, 1000, (r0 - 0.75) / 1000)
    yend, _ = O.mdprop(ymid, lambda r: blk.hammat(r, ered), r0, 12000, (r1 - r0) / 12000)
    wvec, nopen = O.gtwvec(ered, blk.eint)
    nb = O.chnsrt(blk.eint)

    def smat(y):
        sr, si = O.ktos(O.ytok(nb, wvec, blk.n, nopen, y, r1, blk.cent))
        return sr + 1j * si

    rm0, _ = O.syminv(ymid)
    V = O.VivsOptions
    tl, to = O.vivs_tolerances(1e-4)
    sets = {"default": V(), "nopert": V(ipert=False), "idiag": O.vivs_idiag_defaults(V(idiag=True)),
            "ialpha0+idiag": O.vivs_idiag_defaults(V(idiag=True, ialpha=0)),
            "ialfp+idiag": O.vivs_idiag_defaults(V(idiag=True, ialfp=True)), "nosym": V(isym=False)}
    out = {}
    for k, o in sets.items():
        r, steps = O.vvprop(rm0, blk, ered, r0, r1, 0.002, tl, to, opts=o)
        out[k] = (np.abs(smat(O.syminv(r)[0]) - smat(yend)).max(), steps, r)
    return out


@pytest.mark.slow
def test_vivs_corrections_improve_accuracy(options_run):
    """First-order perturbation theory: the corrections help (default better
    than IPERT off, measured 5.6e-5 vs 1.3e-4), the diagonal and derivative
    corrections of IDIAG much more (5.0e-7, same steps), and DELRD steps
    (IALPHA = 0) with IDIAG stay close (1.8e-5 in 121 steps instead of
    726)."""
    e = {k: v[0] for k, v in options_run.items()}
    assert e["default"] < 0.6 * e["nopert"]
    assert e["idiag"] < 0.02 * e["default"]
    assert e["ialpha0+idiag"] < 5e-5
    assert len(options_run["ialpha0+idiag"][1]) < len(options_run["idiag"][1]) / 3
    # IALFP moves the steps inside the intervals
    ra = [s.rnow for s in options_run["ialfp+idiag"][1]]
    rb = [s.rnow for s in options_run["idiag"][1]]
    assert ra != rb and e["ialfp+idiag"] < 2.0 * e["idiag"]
    # ISYM off: R need not be symmetric, but the physics is the same
    rn = options_run["nosym"][2]
    assert np.abs(rn - options_run["default"][2]).max() < 1e-8 * np.abs(rn).max()


@pytest.mark.parametrize("name", ["idiag", "ialpha0+idiag", "ialfp+idiag", "nosym"])
def test_constant_coupling_exact_under_options(name):
    """A constant coupled W (no derivatives): every option set follows the
    exact solution in W's eigenchannels."""
    rng = np.random.default_rng(11)
    n = 3
    m = rng.standard_normal((n, n))
    wc = 0.5 * (m + m.T)
    blk = _constant_block(wc, np.array([0.0, 2.0, 9.0]))
    ered = 4.0
    w, _ = blk.hammat(1.0, ered)
    lam, u = np.linalg.eigh(w)
    r0, r1 = 1.0, 7.3
    y0c = np.array([0.7, -1.3, 2.1])
    rmat0 = u @ np.diag(1.0 / y0c) @ u.T
    V = O.VivsOptions
    o = {"idiag": O.vivs_idiag_defaults(V(idiag=True)),
         "ialpha0+idiag": O.vivs_idiag_defaults(V(idiag=True, ialpha=0)),
         "ialfp+idiag": O.vivs_idiag_defaults(V(idiag=True, ialfp=True)), "nosym": V(isym=False)}[name]
    rmat, steps = O.vvprop(rmat0, blk, ered, r0, r1, 0.05, *O.vivs_tolerances(1e-6), opts=o)
    assert abs(steps[-1].rnow - r1) < 1e-12 and steps[-1].last
    d = r1 - r0
    yexact = np.empty(n)
    for c in range(n):
        k = math.sqrt(abs(lam[c]))
        if lam[c] < 0.0:
            yexact[c] = (-k * math.sin(k * d) + y0c[c] * math.cos(k * d)) / (math.cos(k * d) + y0c[c] * math.sin(k * d) / k)
        else:
            yexact[c] = (k * math.sinh(k * d) + y0c[c] * math.cosh(k * d)) / (math.cosh(k * d) + y0c[c] * math.sinh(k * d) / k)
    rex = u @ np.diag(1.0 / yexact) @ u.T
    assert np.abs(rmat - rex).max() < 1e-10 * np.abs(rex).max()
