"""The oracle's YTRANS (source_code/ytrans.f90, NCONST > 0, NRSQ = 0) pinned by
what the mathematics fixes:
- a diagonal asymptotic operator gives a permutation within each set of
  equal L (entries +1) with the set's thresholds in ascending order;
- a two-channel set gives the closed-form eigenvalues and rotation angle;
- elements between different L are not used;
- free particles: with no potential and a non-diagonal constant operator,
  the exact log-derivative matrix (Riccati-Bessel functions of each
  asymptotic channel, turned into the primitive basis) must come out of
  YTRANS + YTOK + KTOS as S = I.  A transposed transformation or thresholds
  assigned to the wrong channels break this."""
import numpy as np
import pytest
from scipy import special

from molscat_amd import decks
from oracle import molscat_oracle as O


def _oracle_block(fb, field_g):
    return O.Block(n=fb.n, vl=fb.vl, eint=fb.eint, cent=fb.cent, lvals=fb.lvals, levidx=fb.levidx, potl=fb.potl,
                   rscale=fb.rscale, ep2ru=fb.ep2ru, cm2ru=fb.cm2ru, nconst=1, pconst=fb.pconst(field_g))


def _random_zeeman(seed):
    rng = np.random.default_rng(seed)
    z = rng.uniform(-2.0, 2.0, (12, 12))
    return 0.5 * (z + z.T)


def test_diagonal_operator_is_a_sorted_permutation():
    fb = decks.field48()
    ob = _oracle_block(fb, 300.0)
    _, eint, evec = O.ytrans(None, ob)
    assert np.array_equal(np.abs(evec), (np.abs(evec) > 0.5).astype(float))
    assert np.all(evec >= 0.0)
    c = O.ytrans(None, ob)[2].T @ np.diag(np.diag(
        np.tensordot(ob.pconst, ob.vl_full[3:], axes=(0, 0)))) @ evec
    assert np.allclose(np.diag(c), eint, rtol=0, atol=1e-15)
    for ll in (0, 2, 4, 6):
        idx = np.where(fb.lvals == ll)[0]
        assert np.all(np.diff(eint[idx]) >= 0.0)
        # the set's channels only
        assert np.all(evec[np.ix_(idx, np.setdiff1d(np.arange(fb.n), idx))] == 0.0)


def test_two_channel_set_closed_form():
    fb = decks.field48()
    ob = _oracle_block(fb, 1.0)
    n = fb.n
    # constant operator [[a, b], [b, c]] on channels 0 and 4 (both L = 0), zero elsewhere
    a, b, c = 0.7, -0.3, -0.2
    m = np.zeros((n, n))
    m[0, 0], m[0, 4], m[4, 0], m[4, 4] = a, b, b, c
    ob.vl_full = ob.vl_full.copy()
    ob.vl_full[3] = m / ob.pconst[0]
    _, eint, evec = O.ytrans(None, ob)
    lo = 0.5 * (a + c) - np.hypot(0.5 * (a - c), b)
    hi = 0.5 * (a + c) + np.hypot(0.5 * (a - c), b)
    # the set L = 0 is channels 0, 4, ..., 44: lo, ten zeros, hi in channel order
    assert abs(eint[0] - lo) < 1e-15 and abs(eint[44] - hi) < 1e-15
    assert np.all(eint[4:44:4] == 0.0)
    # eigenvector of lo: (cos t, sin t) up to sign, tan 2t = 2b/(a - c)
    t = 0.5 * np.arctan2(2 * b, a - c) + 0.5 * np.pi
    v = np.array([np.cos(t), np.sin(t)])
    v = v if v[np.argmax(np.abs(v))] > 0 else -v
    assert np.allclose(evec[[0, 4], 0], v, atol=1e-15)


def test_elements_between_sets_are_not_used():
    fb = decks.field48(zeeman=_random_zeeman(5))
    ob = _oracle_block(fb, 200.0)
    _, e0, v0 = O.ytrans(None, ob)
    ob.vl_full = ob.vl_full.copy()
    bump = np.zeros((fb.n, fb.n))
    bump[0, 1] = bump[1, 0] = 0.9  # L = 0 with L = 2
    ob.vl_full[3] = ob.vl_full[3] + bump
    _, e1, v1 = O.ytrans(None, ob)
    assert np.array_equal(e0, e1) and np.array_equal(v0, v1)


def _free_y(eint, ered, cent, r):
    """Log-derivatives of the regular free solutions of each asymptotic channel."""
    y = np.empty(len(eint))
    for i, (ei, cc) in enumerate(zip(eint, cent)):
        ll = int(round(np.sqrt(cc + 0.25) - 0.5))
        if ered > ei:
            k = np.sqrt(ered - ei)
            x = k * r
            f = x * special.spherical_jn(ll, x)
            fp = special.spherical_jn(ll, x) + x * special.spherical_jn(ll, x, derivative=True)
        else:
            k = np.sqrt(ei - ered)
            x = k * r
            f = x * special.spherical_in(ll, x)
            fp = special.spherical_in(ll, x) + x * special.spherical_in(ll, x, derivative=True)
        y[i] = k * fp / f
    return y


def free_particle_case(field_g=400.0, r=30.0):
    fb = decks.field48(zeeman=_random_zeeman(7))
    ob = _oracle_block(fb, field_g)
    _, eint, evec = O.ytrans(None, ob)
    # an energy between the thresholds: some asymptotic channels open, some closed
    ered = float(np.median(eint))
    yprim = evec @ np.diag(_free_y(eint, ered, fb.cent, r)) @ evec.T
    return fb, ob, ered, yprim, r


def test_free_particles_give_unit_s_matrix():
    fb, ob, ered, yprim, r = free_particle_case()
    yt, eint, _ = O.ytrans(yprim, ob)
    wvec, nopen = O.gtwvec(ered, eint)
    assert 0 < nopen < fb.n
    nb = O.chnsrt(eint)
    k = O.ytok(nb, wvec, fb.n, nopen, yt, r, fb.cent)
    sr, si = O.ktos(k)
    assert np.abs(k).max() < 1e-9
    assert np.abs(sr - np.eye(nopen)).max() < 1e-12 and np.abs(si).max() < 1e-9
    # without YTRANS the primitive basis is not the asymptotic one: S != I
    kbad = O.ytok(nb, wvec, fb.n, nopen, yprim, r, fb.cent)
    assert np.abs(kbad).max() > 1e-3
