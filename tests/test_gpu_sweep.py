"""The batched partial-wave sweep's device path: MDPROP for several blocks
of uneven N in one persistent launch (mcp_prop_ld_multi), with and without
the captured HIP graph.

Bars: every instance of every block bitwise equal to the single-block call
(mcp_prop_ld, the same per-instance arithmetic), node counts and status
equal; graph replays (three, with eager calls between them) each bitwise
equal to the eager result; one block's Y(RSTOP) and node count against the
oracle's MDPROP (source_code/mdprop.f) to 1e-10 relative."""
import numpy as np
import pytest

from molscat_amd import decks
from molscat_amd import system as S
from oracle import molscat_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.set_graph(False)
    e.close()


@pytest.fixture(scope="module")
def setup(eng):
    deck = decks.rotor441()
    pot = deck.potential()
    jl = list(range(0, 13, 2))
    blks = [S.rotor_block(jt, ib, jl, decks.BE, deck.lambdas, pot, decks.URED, runit=1.0)
            for jt, ib in ((0, 2), (3, 2), (9, 1), (14, 2))]
    rstart, dr, nstep = 2.0, 0.01, 600
    r = S.ld_radii(rstart, dr, nstep)
    hb = [eng.block(b.vl, b.eint, b.cent, mxlam=b.mxlam) for b in blks]
    g = eng.grid(r, blks[0].ptab(r), dr)
    ereds = [np.array([50.0, 200.0, 380.0, 500.0][: 2 + (k % 3)]) * b.cm2ru for k, b in enumerate(blks)]
    y0s = [eng.yinit(h, g, e) for h, e in zip(hb, ereds)]
    return blks, hb, g, ereds, y0s, (rstart, dr, nstep)


def test_multi_block_equals_single_block(eng, setup):
    blks, hb, g, ereds, y0s, _ = setup
    print("multi-block MDPROP N =", [b.n for b in blks], "instances", [len(e) for e in ereds], flush=True)
    outs = eng.prop_ld_multi(hb, g, ereds, y0s)
    for b, h, e, y0, (ym, nm, sm) in zip(blks, hb, ereds, y0s, outs):
        ys, ns, ss = eng.prop_ld(h, g, 6, e, y0)
        assert np.array_equal(sm, ss) and np.all(ss == 0)
        assert np.array_equal(nm, ns)
        assert np.array_equal(ym, ys), (b.n, np.abs(ym - ys).max())


def test_graph_replays_equal_eager(eng, setup):
    blks, hb, g, ereds, y0s, _ = setup
    eager = eng.prop_ld_multi(hb, g, ereds, y0s)
    eng.set_graph(True)
    try:
        for rep in range(3):
            print(f"graph replay {rep}", flush=True)
            got = eng.prop_ld_multi(hb, g, ereds, y0s)
            for (ye, ne, se), (yg, ng, sg) in zip(eager, got):
                assert np.array_equal(ye, yg) and np.array_equal(ne, ng) and np.array_equal(se, sg)
            # eager work between replays
            eng.set_graph(False)
            eng.prop_ld(hb[0], g, 6, ereds[0], y0s[0])
            eng.set_graph(True)
    finally:
        eng.set_graph(False)


def test_multi_block_against_oracle(eng, setup):
    blks, hb, g, ereds, y0s, (rstart, dr, nstep) = setup
    outs = eng.prop_ld_multi(hb, g, ereds, y0s)
    k = 2
    b = blks[k]
    ob = O.Block(n=b.n, vl=b.vl, eint=b.eint, cent=b.cent, lvals=b.lvals, levidx=b.levidx,
                 potl=O.PowerPotential(b.potl.terms, rm=b.potl.rm, epsil=b.potl.epsil), rscale=b.rscale,
                 ep2ru=b.ep2ru, cm2ru=b.cm2ru)
    for m, e in enumerate(ereds[k]):
        yr, nr = O.mdprop(y0s[k][m], lambda r: ob.hammat(r, e), rstart, nstep, dr)
        rel = np.abs(outs[k][0][m] - yr).max() / np.abs(yr).max()
        print(f"  N={b.n} E={e / b.cm2ru:.0f} nodes {int(outs[k][1][m])} (oracle {nr}) rel|dY| {rel:.1e}", flush=True)
        assert int(outs[k][1][m]) == nr
        assert rel < 1e-10
