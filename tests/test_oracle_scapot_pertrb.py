"""SCAPOT and PERTRB in HAMMAT's P array (source_code/hammat.f:55-67,
scapot-default.f, pertrb.f): the oracle's restatement pinned by identities
that the mathematics fixes, and the host P table (molscat_amd.system.hammat_p,
what the device's grids are built from) against the oracle.

Pins:
  * SCALAM scales the radial terms only: W(SCALAM) - W(1) is
    (SCALAM - 1) times the radial part of W, and constant terms and the
    diagonal (EINT, L(L+1)/R^2, ERED) are untouched;
  * PERTRB with IPOW = 2 on the lambda = 0 term of an ITYPE=1 rotor (whose
    coupling matrix is the identity) is a centrifugal shift: the same W as
    L(L+1) + DELTAN * EP2RU / RSCALE^2 on every channel; IPOW = 0 is a
    constant shift DELTAN * EP2RU on the diagonal (a wrong index, power,
    length unit or energy unit fails one of these);
  * PERTRB on a constant term (IPOW = 0) shifts that constant by
    DELTAN * EP2RU; an index outside 1..NHAM changes nothing (pertrb.f:19).
"""
import numpy as np
import pytest

from molscat_amd import decks
from molscat_amd import system as S
from oracle import molscat_oracle as O

POT_TERMS = [[(1.0, -12), (-2.0, -6)], [(-0.2, -6)], [(0.05, -6)]]
RADII = [0.62, 0.9, 1.3, 2.7, 6.0]


def _rotor(scalam=1.0, expval=None):
    pot = S.PowerPotential(POT_TERMS, rm=3.5, epsil=50.0)
    blk = S.rotor_block(4, 2, [0, 2, 4, 6], 30.0, [0, 2, 4], pot, 20.0, runit=1.0)
    blk.scalam, blk.expval = scalam, expval
    ob = O.Block(n=blk.n, vl=blk.vl, eint=blk.eint, cent=blk.cent, lvals=blk.lvals, levidx=blk.levidx,
                 potl=O.PowerPotential(pot.terms, rm=pot.rm, epsil=pot.epsil), rscale=blk.rscale, ep2ru=blk.ep2ru,
                 cm2ru=blk.cm2ru, scalam=scalam, expval=expval)
    return blk, ob


def _radial_part(ob, r):
    p = np.array(ob.potl(r * ob.rscale)) * ob.ep2ru
    return np.tensordot(p, ob.vl_full[: len(p)], axes=(0, 0))


def test_lambda0_coupling_is_identity():
    _, ob = _rotor()
    # the 3j/6j products give 1 to rounding
    assert np.abs(ob.vl_full[0] - np.eye(ob.n)).max() < 1e-14


@pytest.mark.parametrize("scalam", [0.0, 0.93, 1.7])
def test_scalam_scales_the_radial_terms_only(scalam):
    _, o1 = _rotor()
    _, os_ = _rotor(scalam=scalam)
    ered = 0.4 * o1.cm2ru
    for r in RADII:
        w1, _ = o1.hammat(r, ered)
        ws, _ = os_.hammat(r, ered)
        rad = _radial_part(o1, r)
        scale = np.abs(rad).max()
        assert np.abs((ws - w1) - (scalam - 1.0) * rad).max() <= 1e-13 * scale, r


@pytest.mark.parametrize("ipow", [0, 2])
def test_pertrb_on_lambda0_is_a_diagonal_shift(ipow):
    dn = 0.37
    _, o1 = _rotor()
    _, op = _rotor(expval=(1, ipow, dn))
    ered = 0.4 * o1.cm2ru
    for r in RADII:
        w1, _ = o1.hammat(r, ered)
        wp, _ = op.hammat(r, ered)
        if ipow == 2:
            # DELTAN / (R RSCALE)^2 in EPSIL units = L(L+1) shift over R^2
            shifted = O.Block(n=o1.n, vl=o1.vl, eint=o1.eint, cent=o1.cent + dn * o1.ep2ru / o1.rscale ** 2,
                              lvals=o1.lvals, levidx=o1.levidx, potl=o1.potl, rscale=o1.rscale, ep2ru=o1.ep2ru,
                              cm2ru=o1.cm2ru)
            want, _ = shifted.hammat(r, ered)
        else:
            want = w1 + dn * o1.ep2ru * np.eye(o1.n)
        scale = np.abs(w1).max()
        assert np.abs(wp - want).max() <= 1e-13 * scale, (r, np.abs(wp - want).max())
        # off-diagonal untouched
        off = ~np.eye(o1.n, dtype=bool)
        assert np.abs((wp - w1)[off]).max() <= 1e-13 * scale


@pytest.mark.parametrize("i", [0, 4, 99])
def test_pertrb_outside_the_table_changes_nothing(i):
    _, o1 = _rotor()
    _, op = _rotor(expval=(i, 3, 0.5))
    for r in RADII:
        assert np.array_equal(o1.hammat(r, 0.0)[0], op.hammat(r, 0.0)[0])


def test_pertrb_on_a_constant_term():
    fb = decks.field48()
    dn = 2.5e-4
    from test_oracle_ytrans import _oracle_block

    ob = _oracle_block(fb, 300.0)
    ob.expval = (fb.mxlam + 1, 0, dn)
    p = ob.pvec(1.3)
    assert p[fb.mxlam] == pytest.approx(300.0 * fb.cm2ru + dn * fb.ep2ru, rel=1e-14)
    # the host's constants alike, and an R-dependent shift refused
    fb.expval = (fb.mxlam + 1, 0, dn)
    assert fb.pconst(300.0)[0] == pytest.approx(p[fb.mxlam], rel=1e-15)
    fb.expval = (fb.mxlam + 1, 3, dn)
    with pytest.raises(ValueError):
        fb.pconst(300.0)


@pytest.mark.parametrize("setting", [dict(scalam=0.9), dict(expval=(2, 6, 0.3)), dict(expval=(1, 2, -0.1)),
                                     dict(scalam=1.1, expval=(3, 0, 0.02))])
def test_host_table_matches_oracle(setting):
    blk, ob = _rotor(**setting)
    r = np.array(RADII)
    tab = blk.ptab(r)
    for k, x in enumerate(RADII):
        want = ob.pvec(x)
        assert np.abs(tab[k] - want).max() <= 1e-15 * np.abs(want).max(), (setting, x)


def test_host_clamp_before_scaling():
    # |P| > 1e60 is clamped before SCALAM (hammat.f:42-49 precede the SCAPOT call)
    p = np.array([[3e70, -2e65, 1.0]])
    out = S.hammat_p(p, np.array([1.0]), 2.0, scalam=0.25)
    assert np.array_equal(out, np.array([[0.5e60, -0.5e60, 0.5]]))
