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


# ---------------------------------------------------------------------------
# NRSQ > 0: the centrifugal operator among the terms, resolved with the
# thresholds by MULTOP
# ---------------------------------------------------------------------------
def turned_rotor_blocks(seed=5):
    """The all-propagators rotor deck (JTOT = 10, IB = 1, j = 0..6, N = 12)
    and the same physics in a basis turned by a random orthogonal Q: every
    radial term Q^T VL Q, the thresholds as a constant term Q^T EINT Q (P = 1,
    EINT = 0) and the centrifugal term as the operator Q^T L(L+1) Q (NRSQ =
    1).  YTRANS must undo the turn: channels by ascending threshold, then
    ascending L within a level."""
    pot = O.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    blk, _ = O.make_itype1_block(10, 1, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)
    n = blk.n
    q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, n)))
    pack = lambda m: np.array([m[i, j] for j in range(n) for i in range(j + 1)])
    vls = [pack(q.T @ blk.vl_full[k] @ q) for k in range(blk.vl.shape[0])]
    vls.append(pack(q.T @ np.diag(blk.eint) @ q))
    vls.append(pack(q.T @ np.diag(blk.cent) @ q))
    tb = O.Block(n=n, vl=np.array(vls), eint=np.zeros(n), cent=np.zeros(n), lvals=np.zeros(n, dtype=int),
                 levidx=np.zeros(n, dtype=int), potl=blk.potl, rscale=blk.rscale, ep2ru=blk.ep2ru, cm2ru=blk.cm2ru,
                 nconst=1, pconst=np.array([1.0]), nrsq=1)
    return blk, tb, q


def test_ytrans_rsq_recovers_thresholds_and_l():
    blk, tb, q = turned_rotor_blocks()
    _, eint, evec, cent = O.ytrans_rsq(None, tb)
    order = sorted(range(blk.n), key=lambda i: (blk.eint[i], blk.lvals[i]))
    assert np.allclose(eint, blk.eint[order], rtol=0, atol=1e-12 * np.abs(blk.eint).max())
    assert np.allclose(cent, blk.cent[order], rtol=0, atol=1e-9)
    # the columns are the turned unit vectors, largest component positive
    for m, i in enumerate(order):
        t = q.T[:, i] * np.sign(q.T[np.argmax(np.abs(q.T[:, i])), i])
        assert np.abs(evec[:, m] - t).max() < 1e-9


def _label_map(blk, eint_t, cent_t, nb_o, nb_t, nopen):
    """Open channel a of the turned problem (NB order) -> the position of
    the original open channel with the same threshold and L(L+1)."""
    out = []
    for a in range(nopen):
        k = nb_t[a]
        i = next(i for i in range(blk.n) if abs(blk.eint[i] - eint_t[k]) < 1e-9 * abs(blk.eint).max()
                 and abs(blk.cent[i] - cent_t[k]) < 1e-6)
        out.append(int(np.where(nb_o[:nopen] == i)[0][0]))
    return out


def test_turned_ytrans_then_ytok_reproduces_k():
    """Y at 20 RM of the plain deck (MDPROP), turned by Q: YTRANS (C and
    L^2 resolved by MULTOP) then YTOK with YTRANS's CENT gives the plain
    deck's K-matrix, channel for channel (matched by threshold and L) up to
    the sign of each asymptotic channel: K''_ab = d_a d_b K_pi(a)pi(b), to
    1e-11 of K's scale."""
    blk, tb, q = turned_rotor_blocks()
    ered = 1000.0 * blk.cm2ru
    rstart, _ = O.findrm(blk, 0.5, ered, 9)
    g = O.Grid(rstart, 20.0, 2000, (20.0 - rstart) / 2000)
    a = O.smatrix_ld(blk, ered, g)
    y = a["Y"]
    y2, eint2, evec2, cent2 = O.ytrans_rsq(q.T @ y @ q, tb)
    wvec2, nopen2 = O.gtwvec(ered, eint2)
    nb2 = O.chnsrt(eint2)
    assert nopen2 == a["NOPEN"]
    k2 = O.ytok(nb2, wvec2, blk.n, nopen2, y2, 20.0, cent2)
    pi = _label_map(blk, eint2, cent2, a["NB"], nb2, nopen2)
    # the sign of each asymptotic channel: its EVEC column against the turned unit vector
    d = np.array([np.sign(evec2[:, nb2[m]] @ q.T[:, a["NB"][pi[m]]]) for m in range(nopen2)])
    kp = a["K"][np.ix_(pi, pi)] * d[:, None] * d[None, :]
    assert np.abs(k2 - kp).max() < 1e-11 * max(1.0, np.abs(kp).max())


def test_turned_basis_smatrix_converges_to_the_same():
    """Through the whole path (YINIT, MDPROP, YTRANS, YTOK, KTOS) the turned
    deck's S-matrix approaches the plain deck's as the step shrinks: MDPROP
    takes W's diagonal as its reference potential, so its discretisation
    error depends on the basis, and the difference falls at fourth order
    (4000 -> 8000 sectors: at least 8 times smaller)."""
    blk, tb, _ = turned_rotor_blocks()
    ered = 1000.0 * blk.cm2ru
    rstart, _ = O.findrm(blk, 0.5, ered, 9)
    diffs = []
    for ns in (4000, 8000):
        g = O.Grid(rstart, 20.0, ns, (20.0 - rstart) / ns)
        a = O.smatrix_ld(blk, ered, g)
        b = O.smatrix_ld(tb, ered, g)
        assert a["NOPEN"] == b["NOPEN"] and a["NODES"] == b["NODES"]
        _, eint2, _, cent2 = O.ytrans_rsq(None, tb)
        pi = _label_map(blk, eint2, cent2, a["NB"], b["NB"], a["NOPEN"])
        sa = (a["SR"] + 1j * a["SI"])[np.ix_(pi, pi)]
        sb = b["SR"] + 1j * b["SI"]
        diffs.append(np.abs(np.diag(sa) - np.diag(sb)).max())
    assert diffs[0] > 8.0 * diffs[1]
