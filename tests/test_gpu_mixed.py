"""Mixed precision on the large-N tree (mcp_set_precision): batched
inversions in FP32 on the matrix cores, won back to FP64 by Newton-Schulz
steps, FP64 below 3 A, YINIT / recurrence / matching in FP64.

Bars: the N=441 rotor solve (3600 MDPROP sectors + 162 AIRY steps) against
the oracle fixture with max|dS| <= 1e-8 and both node counts exact; a
smaller N >= 200 block at four energies against the FP64 device path with
max|dS| <= 1e-8 and node counts equal; the guard statistics reported."""
import json
import os

import numpy as np
import pytest

from molscat_amd import decks
from molscat_amd import system as S

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.set_precision(0)
    e.close()


def test_mixed_rotor441_against_fixture(eng):
    with open(os.path.join(HERE, "golden", "rotor441_oracle.json")) as f:
        fx = json.load(f)
    deck = decks.rotor441()
    blk = deck.block()
    seg = deck.ld
    r = seg.radii()
    ra, wa = deck.airy_grid()
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), seg.dr)
    ga = eng.airy_grid(ra, wa, blk.ptab(ra))
    ered = deck.energies[0] * blk.cm2ru
    wv, nopen = S.gtwvec(ered, blk.eint)
    nb = S.chnsrt(blk.eint)
    eng.set_precision(2, 3.0)
    try:
        print("mixed precision N=441 solve", flush=True)
        y0 = eng.yinit(b, g, [ered])
        y, nodes, st = eng.prop_ld(b, g, 6, [ered], y0)
        assert st[0] == 0 and int(nodes[0]) == fx["nodes_ld"]
        scale = np.abs(np.asarray(fx["y_rmid_diag"])).max()
        dy = np.abs(np.diag(y[0]) - np.asarray(fx["y_rmid_diag"])).max() / scale
        out = eng.smatrix_segments(b, [g, ga], [6, 9], [ered], wv, [nopen], nb, deck.airy_stop)
        stats = eng.precision_stats()
    finally:
        eng.set_precision(0)
    assert out["STATUS"][0] == 0
    ds = max(np.abs(out["SR"][0] - np.asarray(fx["sr"])).max(), np.abs(out["SI"][0] - np.asarray(fx["si"])).max())
    print(f"  rel|dY(20 A)| {dy:.1e} max|dS| {ds:.1e} nodes {int(out['NODES'][0])} stats {stats}", flush=True)
    assert int(out["NODES"][0]) == fx["nodes_ld"] + fx["nodes_airy"]
    assert ds <= 1e-8
    assert stats["mixed_matrices"] > 0


def test_mixed_vs_fp64_at_n_over_200(eng):
    deck = decks.rotor441()
    pot = deck.potential()
    blk = S.rotor_block(20, 1, deck.jlevels, decks.BE, deck.lambdas, pot, decks.URED, runit=1.0)
    assert blk.n >= 200
    seg = deck.ld
    r = seg.radii()
    ra, wa = deck.airy_grid()
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), seg.dr)
    ga = eng.airy_grid(ra, wa, blk.ptab(ra))
    ered = np.array([365.0, 400.0, 450.0, 490.0]) * blk.cm2ru
    wv, nops = zip(*[S.gtwvec(e, blk.eint) for e in ered])
    assert min(nops) > 0
    nb = S.chnsrt(blk.eint)
    print(f"mixed vs FP64 N={blk.n}, 4 energies, nopen {list(nops)}", flush=True)
    ref = eng.smatrix_segments(b, [g, ga], [6, 9], ered, np.array(wv), np.array(nops), nb, deck.airy_stop)
    eng.set_precision(2, 3.0)
    try:
        mix = eng.smatrix_segments(b, [g, ga], [6, 9], ered, np.array(wv), np.array(nops), nb, deck.airy_stop)
        stats = eng.precision_stats()
    finally:
        eng.set_precision(0)
    worst = 0.0
    for m in range(len(ered)):
        assert ref["STATUS"][m] == 0 and mix["STATUS"][m] == 0
        assert int(ref["NODES"][m]) == int(mix["NODES"][m])
        worst = max(worst, np.abs(ref["SR"][m] - mix["SR"][m]).max(), np.abs(ref["SI"][m] - mix["SI"][m]).max())
    print(f"  max|dS| mixed vs FP64 {worst:.1e}, nodes {list(ref['NODES'])}, stats {stats}", flush=True)
    assert worst <= 1e-8
