"""The N=441 rotor workload at its own size (the S-matrix bench.py times):
YINIT at 2 A, MDPROP over 3600 sectors to 20 A, AIRY (TOLHI = 0) over the
R^(4/3) grid to 200 A, YTOK + KTOS, all through the C ABI, against the
oracle's FP64 solve committed in tests/golden/rotor441_oracle.json
(tests/golden/make_rotor441_fixture.py).

Bars: S to 1e-10 absolute; node counts bit-exact after each segment; Y(20 A)
and Z(200 A) to 1e-10 relative to their largest element on the stored
diagonal and columns, and their Frobenius norms to 1e-12 relative.  The
default (tree over the radial grid) and the sequential recurrence
(mcp_set_large_mode(1)) must give the same node count over all 3600
sectors: the Haynsworth inertia argument for the tree's merge pivots at
full size.  The workgroup-team kernels issued as ordinary launches (the
mode the kernel profiles are taken in) give the cooperative launches'
S-matrix bit for bit at this size."""
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
    e.close()


@pytest.fixture(scope="module")
def fx():
    with open(os.path.join(HERE, "golden", "rotor441_oracle.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def setup(eng):
    deck = decks.rotor441()
    blk = deck.block()
    seg = deck.ld
    r = seg.radii()
    ra, wa = deck.airy_grid()
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), seg.dr)
    ga = eng.airy_grid(ra, wa, blk.ptab(ra))
    ered = deck.energies[0] * blk.cm2ru
    return deck, blk, b, g, ga, ered


def _check_matrix(m, diag, cols, fro, what):
    scale = max(np.abs(diag).max(), max(np.abs(c).max() for c in cols.values()))
    err = np.abs(np.diag(m) - np.asarray(diag)).max()
    for c, v in cols.items():
        err = max(err, np.abs(m[:, int(c)] - np.asarray(v)).max())
    rel = err / scale
    frel = abs(np.linalg.norm(m) - fro) / fro
    print(f"  {what}: max rel err {rel:.2e}, |fro| rel {frel:.2e}", flush=True)
    assert rel < 1e-10, (what, rel)
    assert frel < 1e-12, (what, frel)


def test_smatrix_segments_full_size(eng, fx, setup):
    deck, blk, b, g, ga, ered = setup
    assert blk.n == fx["n"] == 441
    wv, nopen = S.gtwvec(ered, blk.eint)
    nb = S.chnsrt(blk.eint)
    assert nopen == fx["nopen"]
    print(f"rotor441 smatrix_segments N={blk.n} nopen={nopen}", flush=True)
    out = eng.smatrix_segments(b, [g, ga], [6, 9], [ered], wv, [nopen], nb, deck.airy_stop)
    assert out["STATUS"][0] == 0
    dsr = np.abs(out["SR"][0] - np.asarray(fx["sr"])).max()
    dsi = np.abs(out["SI"][0] - np.asarray(fx["si"])).max()
    dk = np.abs(out["K"][0] - np.asarray(fx["k"])).max() / max(1.0, np.abs(fx["k"]).max())
    print(f"  max|dSR|={dsr:.2e} max|dSI|={dsi:.2e} rel|dK|={dk:.2e} nodes={int(out['NODES'][0])}", flush=True)
    assert int(out["NODES"][0]) == fx["nodes_ld"] + fx["nodes_airy"]
    assert dsr < 1e-10 and dsi < 1e-10 and dk < 1e-10


def test_segments_against_oracle(eng, fx, setup):
    deck, blk, b, g, ga, ered = setup
    print("rotor441 YINIT + MDPROP (tree) + AIRY, segment by segment", flush=True)
    y0 = eng.yinit(b, g, [ered])
    y, nodes, st = eng.prop_ld(b, g, 6, [ered], y0)
    assert st[0] == 0
    assert int(nodes[0]) == fx["nodes_ld"], (int(nodes[0]), fx["nodes_ld"])
    _check_matrix(y[0], fx["y_rmid_diag"], fx["y_rmid_cols"], fx["y_rmid_fro"], "Y(20 A)")
    z, nz, st = eng.prop_airy(b, ga, [ered], y)
    assert st[0] == 0
    assert int(nz[0]) == fx["nodes_airy"], (int(nz[0]), fx["nodes_airy"])
    _check_matrix(z[0], fx["z_rmax_diag"], fx["z_rmax_cols"], fx["z_rmax_fro"], "Z(200 A)")


def test_tree_and_sequential_node_counts_agree(eng, fx, setup):
    deck, blk, b, g, ga, ered = setup
    y0 = eng.yinit(b, g, [ered])
    print("rotor441 sequential recurrence over 3600 sectors", flush=True)
    eng.set_large_mode(1)
    try:
        ys, ns, st = eng.prop_ld(b, g, 6, [ered], y0)
    finally:
        eng.set_large_mode(0)
    assert st[0] == 0
    yt, nt, st = eng.prop_ld(b, g, 6, [ered], y0)
    assert st[0] == 0
    print(f"  nodes sequential {int(ns[0])} tree {int(nt[0])} oracle {fx['nodes_ld']}", flush=True)
    assert int(ns[0]) == int(nt[0]) == fx["nodes_ld"]
    rel = np.abs(ys[0] - yt[0]).max() / np.abs(yt[0]).max()
    assert rel < 1e-10, rel


def test_plain_team_launch_bitwise_full_size(eng, setup):
    deck, blk, b, g, ga, ered = setup
    wv, nopen = S.gtwvec(ered, blk.eint)
    nb = S.chnsrt(blk.eint)
    print("rotor441: cooperative and ordinary team launches", flush=True)
    a = eng.smatrix_segments(b, [g, ga], [6, 9], [ered], wv, [nopen], nb, deck.airy_stop)
    eng.set_team_launch(False)
    try:
        p = eng.smatrix_segments(b, [g, ga], [6, 9], [ered], wv, [nopen], nb, deck.airy_stop)
    finally:
        eng.set_team_launch(True)
    assert a["STATUS"][0] == 0 and p["STATUS"][0] == 0
    assert np.array_equal(a["SR"][0], p["SR"][0]) and np.array_equal(a["SI"][0], p["SI"][0])
    assert np.array_equal(a["NODES"], p["NODES"])
