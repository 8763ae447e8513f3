"""Pins oracle functions that no reference fixture covers against closed
forms: Riccati-Bessel functions against scipy, the log-derivative
propagators against exact single-channel solutions, K->S for one open
channel, CALCA for a hard sphere, and HAMMAT against a dense restatement
of its definition."""
import math

import numpy as np
import pytest
from scipy import special

from oracle import molscat_oracle as O


@pytest.mark.parametrize("ell", [0.0, 1.0, 2.0, 7.0, 30.0, 0.3, 5.5])
@pytest.mark.parametrize("x", [0.2, 1.5, 4.0, 12.0, 40.0, 90.0])
def test_rbesjy_against_scipy(ell, x):
    uj, ujp, un, unp = O.rbesjy(ell, x)
    v = ell + 0.5
    # Riccati-Bessel: j^(x) = sqrt(pi x/2) J_{l+1/2}(x), n^ likewise with Y
    f = math.sqrt(math.pi * x / 2.0)
    jv, yv = special.jv(v, x), special.yv(v, x)
    jvp, yvp = special.jvp(v, x), special.yvp(v, x)
    ej = f * jv
    ejp = f * jvp + jv * math.sqrt(math.pi / (8.0 * x))
    en = f * yv
    enp = f * yvp + yv * math.sqrt(math.pi / (8.0 * x))
    for a, b in ((uj, ej), (ujp, ejp), (un, en), (unp, enp)):
        assert abs(a - b) <= 1e-11 * max(1.0, abs(b))


@pytest.mark.parametrize("ell", [0.0, 2.0, 9.0, 0.4])
@pytest.mark.parametrize("x", [0.3, 2.0, 25.0])
def test_rbessk_against_scipy(ell, x):
    v = ell + 0.5
    # modified Riccati-Bessel of the third kind ~ sqrt(x) K_v(x)
    k = special.kve(v, x)
    kp = special.kvp(v, x) * math.exp(x)
    ratio = (kp + k / (2.0 * x)) / k
    assert abs(O.rbessk(ell, x) - ratio) < 1e-12 * max(1.0, abs(ratio))


@pytest.mark.parametrize("prop", ["md", "ld"])
def test_constant_potential_closed_form(prop):
    # y'' = W y with W = -k^2: log-derivative of sin(k(r - r0)) is k cot(k(r - r0))
    k = 2.3
    r0 = 0.0
    rs, re = 0.4, 3.1
    nstep = 400
    dr = (re - rs) / nstep
    f = O.mdprop if prop == "md" else O.ldprop
    # MDPROP uses N>1 code; pad with an uncoupled closed channel
    w = np.diag([-k * k, 5.0])
    hm = lambda r: (w.copy(), np.diag(w).copy())
    y0 = np.diag([k / math.tan(k * (rs - r0)), math.sqrt(5.0)])
    y, nodes = f(y0, hm, rs, nstep, dr)
    exact = k / math.tan(k * (re - r0))
    tol = 1e-10 if prop == "md" else 2e-6  # LDMD is exact (up to rounding) for a constant diagonal W
    assert abs(y[0, 0] - exact) < tol * max(1.0, abs(exact))
    assert abs(y[1, 1] - math.sqrt(5.0) / math.tanh(math.sqrt(5.0) * 1e9)) < 1e-6
    # nodes of sin(k r) between rs and re: zeros at pi/k = 1.366, 2pi/k = 2.732
    assert nodes == 2


def test_free_particle_converges_to_riccati_bessel():
    # W = L(L+1)/r^2 - k^2: Y -> k j^'/j^ with O(h^4) error (MDPROP)
    L, k = 3, 1.7
    cent = L * (L + 1)
    rs, re = 0.05, 6.0

    def hm(r):
        w = np.diag([cent / r ** 2 - k * k, 50.0])
        return w, np.diag(w).copy()

    errs = []
    for nstep in (200, 400):
        y0 = np.diag([math.sqrt(cent / rs ** 2 - k * k), math.sqrt(50.0)])
        y, _ = O.mdprop(y0, hm, rs, nstep, (re - rs) / nstep)
        uj, ujp, _, _ = O.rbesjy(float(L), k * re)
        errs.append(abs(y[0, 0] - k * ujp / uj))
    assert errs[1] < errs[0] / 10.0  # ~16x for O(h^4)
    assert errs[1] < 1e-4


def test_ktos_single_channel():
    # one open channel: K = tan(delta), S = exp(2 i delta) (code convention)
    for delta in (0.3, -1.1, 1.4):
        sr, si = O.ktos(np.array([[math.tan(delta)]]))
        s = complex(sr[0, 0], si[0, 0])
        assert abs(s - complex(math.cos(2 * delta), math.sin(2 * delta))) < 1e-14


def test_calca_hard_sphere():
    # s-wave hard sphere of radius a: delta = -k a, scattering length a
    a = 1.7
    for k in (1e-4, 1e-3):
        d = -k * a
        s = complex(math.cos(2 * d), math.sin(2 * d))
        sc = O.calca(s.real, s.imag, k, 0)
        assert abs(sc.real - a) < 1e-6 * a * 10
        assert abs(sc.imag) < 1e-9


def test_hammat_dense_definition():
    pot = O.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    blk, runit = O.make_itype1_block(6, 2, [0, 2, 4], 20.0, [0, 2], pot, 2.0, runit=1.0)
    r, ered = 4.2, 123.0 * blk.cm2ru
    w, d = blk.hammat(r, ered)
    p = blk.ep2ru * pot(r / 3.5)
    dense = sum(p[k] * blk.vl_full[k] for k in range(2))
    dense += np.diag(blk.cent / r ** 2 + blk.eint - ered)
    assert np.abs(w - dense).max() < 1e-12 * np.abs(dense).max()
    assert np.array_equal(d, np.diag(w))
    # isotropic term is diagonal with coefficient 1 (Percival-Seaton, lambda=0)
    assert np.allclose(blk.vl_full[0], np.eye(blk.n), atol=1e-14)


def _single_channel(c0, cent):
    pot = O.PowerPotential([[(1.0, 0, 0.0)]])  # P(R) = 1
    return O.Block(n=1, vl=np.array([[c0]]), eint=np.zeros(1), cent=np.array([cent]), lvals=np.zeros(1, dtype=int),
                   levidx=np.zeros(1, dtype=int), potl=pot, rscale=1.0, ep2ru=1.0, cm2ru=1.0)


def test_odprop_constant_potential_is_exact():
    # ODPROP's reference potential is the midpoint W: exact for a constant W
    k, rs, re, nstep = 2.3, 0.4, 3.1, 400
    blk = _single_channel(-k * k, 0.0)
    y, nodes = O.odprop(k / math.tan(k * rs), blk, 0.0, rs, nstep, (re - rs) / nstep)
    exact = k / math.tan(k * re)
    assert abs(y - exact) < 1e-10 * max(1.0, abs(exact))
    assert nodes == 0  # ODPROP does not count nodes (source_code/odprop.f)


def test_odprop_free_particle_converges_to_riccati_bessel():
    L, k, rs, re = 3, 1.7, 0.05, 6.0
    blk = _single_channel(-k * k, float(L * (L + 1)))
    uj, ujp, _, _ = O.rbesjy(float(L), k * re)
    errs = []
    for nstep in (200, 400):
        y, _ = O.odprop(math.sqrt(L * (L + 1) / rs ** 2 - k * k), blk, 0.0, rs, nstep, (re - rs) / nstep)
        errs.append(abs(y - k * ujp / uj))
    assert errs[1] < errs[0] / 10.0
    assert errs[1] < 1e-4


def test_maprop_constant_coupled_potential_is_exact():
    """MAPROP (LDMA) for an R-independent coupled W: the adiabatic basis is
    W's eigenbasis at every sector, the corrections vanish, and the half-
    sector reference terms are exact, so Y must follow the exact solution
    in each eigenchannel (cot for open, coth-like for closed), and NODES
    counts the zeros of the open eigenchannels' solutions."""
    rng = np.random.default_rng(3)
    n = 3
    m = rng.standard_normal((n, n))
    wc = 0.5 * (m + m.T)
    vl = np.array([[wc[i, j] for j in range(n) for i in range(j + 1)]])
    pot = O.PowerPotential([[(1.0, 0, 0.0)]], rm=1.0, epsil=1.0)
    blk = O.Block(n=n, vl=vl, eint=np.array([0.0, 1.5, 12.0]), cent=np.zeros(n), lvals=np.zeros(n, dtype=int),
                  levidx=np.arange(n), potl=pot, rscale=1.0, ep2ru=1.0, cm2ru=1.0)
    ered = 5.0
    lam, u = np.linalg.eigh(blk.hammat(1.0, ered)[0])
    assert np.sum(lam < 0.0) == 2
    rs, re, r0 = 0.5, 6.0, 0.1
    yc0 = np.empty(n)
    for c in range(n):
        k = math.sqrt(abs(lam[c]))
        yc0[c] = k / math.tan(k * (rs - r0)) if lam[c] < 0.0 else k / math.tanh(k * (rs - r0))
    y0 = u @ np.diag(yc0) @ u.T
    nstep = 300
    y, nodes = O.maprop(y0, blk, ered, rs, nstep, (re - rs) / nstep)
    yc = np.empty(n)
    zeros = 0
    for c in range(n):
        k = math.sqrt(abs(lam[c]))
        if lam[c] < 0.0:
            yc[c] = k / math.tan(k * (re - r0))
            zeros += int(math.floor(k * (re - r0) / math.pi)) - int(math.floor(k * (rs - r0) / math.pi))
        else:
            yc[c] = k / math.tanh(k * (re - r0))
    yex = u @ np.diag(yc) @ u.T
    assert np.abs(y - yex).max() < 1e-9 * np.abs(yex).max()
    assert nodes == zeros
