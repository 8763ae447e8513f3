"""COUPLE for ITYPE = 1 on the device (mcp_couple_rotor,
mcp_block_create_rotor; source_code/couple.f:51-143) through the C ABI:
the coefficients against the exact rational ones of system.rotor_coupling
for blocks up to the N = 441 block of the rotor deck (JTOT = 60, IB = 2,
j = 0..40, lambda = 0..6), and a block made on the device propagating as
the block uploaded with the exact VL."""
import numpy as np
import pytest

from molscat_amd import decks
from molscat_amd import system as S

pytestmark = pytest.mark.gpu

POT = S.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("jtot,ib,jlev,lams", [(10, 1, [0, 2, 4, 6], [0, 2]), (0, 2, [0, 2, 4], [0, 2, 4]),
                                               (7, 2, [0, 1, 2, 3, 4, 5, 6], [0, 1, 2, 3, 4]),
                                               (120, 1, list(range(0, 41, 2)), [0, 2, 4, 6])])
def test_coefficients_against_exact(eng, jtot, ib, jlev, lams):
    blk = S.rotor_block(jtot, ib, jlev, 30.0, lams, POT, 20.0)
    got = eng.couple_rotor(blk.jch, blk.lvals, jtot, lams)
    d = np.abs(got - blk.vl).max()
    print(f"JTOT={jtot} IB={ib} N={blk.n} NLAM={len(lams)} max|dVL| = {d:.1e}", flush=True)
    assert got.shape == blk.vl.shape and d < 4e-15


def test_rotor441_block(eng):
    deck = decks.rotor441()
    blk = deck.block()
    got = eng.couple_rotor(blk.jch, blk.lvals, blk.jtot, blk.lambdas)
    d = np.abs(got - blk.vl).max()
    print(f"rotor441 N={blk.n}: max|dVL| = {d:.1e}, nonzero {np.count_nonzero(blk.vl)}", flush=True)
    assert d < 4e-15
    assert np.array_equal(got == 0.0, blk.vl == 0.0)


def test_device_block_propagates_as_uploaded(eng):
    blk = S.rotor_block(10, 2, [0, 2, 4, 6, 8], 30.0, [0, 2], POT, 20.0)
    rstart, rstop, nstep = 0.65, 20.0, 3000
    dr = (rstop - rstart) / nstep
    r = S.ld_radii(rstart, dr, nstep)
    g = eng.grid(r, blk.ptab(r), dr)
    ered = np.array([900.0, 1100.0, 1300.0]) * blk.cm2ru
    wv, nops = zip(*[S.gtwvec(e, blk.eint) for e in ered])
    nb = S.chnsrt(blk.eint)
    outs = []
    for b in (eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam),
              eng.block_rotor(blk.jch, blk.lvals, blk.jtot, blk.lambdas, blk.eint, blk.cent)):
        outs.append(eng.smatrix_ld(b, g, 6, ered, np.array(wv), np.array(nops), nb, rstop))
        eng.destroy_block(b)
    a, b = outs
    assert np.all(a["STATUS"] == 0) and np.all(b["STATUS"] == 0)
    assert np.array_equal(a["NODES"], b["NODES"])
    d = max(max(np.abs(x - y).max() for x, y in zip(a[k], b[k])) for k in ("SR", "SI"))
    print(f"N={blk.n}: S from device VL vs uploaded exact VL max|dS| = {d:.1e}", flush=True)
    # VL differs in the last bits (4e-15 at most); the propagation carries
    # such differences to S at the level of its own rounding (measured 4.7e-12)
    assert d < 2e-11


def test_refuses_bad_arguments(eng):
    from molscat_amd.lib import McpError

    with pytest.raises(McpError):
        eng.couple_rotor([0, -1], [1, 1], 1, [0])
    with pytest.raises(McpError):
        eng.couple_rotor([0, 1], [1, 1], 1, [-2])
