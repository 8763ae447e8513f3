"""GPU FIELD search (molscat_amd.bound.field_levels) on the field deck with
a diagonal Zeeman operator of slopes -100..100 MHz/G (so that levels of
other internal states cross an energy fixed 5 cm-1 below the lowest
threshold within 0..1000 G): the batched scan's node counts against the
oracle's BOUND matching at sample fields (bit-exact), and every crossing
found checked with the oracle: EIGMIN vanishes there (to 1e-6 of the
largest eigenvalue of the matching matrix) and NODEC steps by one across
it."""
import numpy as np
import pytest

from molscat_amd import bound as BD
from molscat_amd import decks
from molscat_amd import system as S
from oracle import molscat_oracle as O

from test_oracle_ytrans import _oracle_block

pytestmark = pytest.mark.gpu

RMIN, RMATCH, RMAX, DR = 3.5, 8.0, 40.0, 0.02
DELTA = 5.0


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def _deck():
    return decks.field48(zeeman=np.diag(np.linspace(-100.0, 100.0, 12)))


def _oracle_at(fb, field_g):
    ob = _oracle_block(fb, field_g)
    _, eint, _ = O.ytrans(None, ob)
    ered = eint.min() - DELTA * fb.cm2ru
    (r0, dro, no), (r1, dri, ni) = S.bound_segments(RMIN, RMATCH, RMAX, DR)
    return O.bound_match(ob, ered, O.BoundGrid(RMIN, RMATCH, RMAX, no, dro, ni, dri))


def test_field_crossings(eng):
    fb = _deck()
    scan = np.linspace(0.0, 1000.0, 257)
    print(f"FIELD search N={fb.n}, {len(scan)} fields, {DELTA} cm-1 below threshold", flush=True)
    found = BD.field_levels(eng, fb, DELTA, scan, RMIN, RMATCH, RMAX, DR)
    print(f"  {len(found)} crossings: " + ", ".join(f"{x['FIELD']:.6f} G ({x['NODEC_LO']}->{x['NODEC_HI']})"
                                                   for x in found), flush=True)
    assert len(found) >= 1
    # every crossing of the scan is found, each a single step of NODEC
    b = eng.block(fb.vl, fb.eint, fb.cent, mxlam=fb.mxlam, nconst=fb.nconst)
    (r0, dro, no), (r1, dri, ni) = S.bound_segments(RMIN, RMATCH, RMAX, DR)
    ro, ri = S.ld_radii(r0, dro, no), S.ld_radii(r1, dri, ni)
    go, gi = eng.grid(ro, fb.ptab(ro), dro), eng.grid(ri, fb.ptab(ri), dri)
    pc = np.stack([fb.pconst(f) for f in scan])
    eint, _ = eng.ytrans(b, pc)
    nod = eng.bound_match(b, go, gi, eint.min(axis=1) - DELTA * fb.cm2ru, pconst=pc)["NODEC"].astype(int)
    assert len(found) == int(np.abs(np.diff(nod)).sum())
    assert all(abs(x["NODEC_HI"] - x["NODEC_LO"]) == 1 for x in found)
    assert all(a["FIELD"] < c["FIELD"] for a, c in zip(found[:-1], found[1:]))
    for x in found[:3]:
        ref = _oracle_at(fb, x["FIELD"])
        scale = np.abs(ref["EVALS"]).max()
        print(f"  B={x['FIELD']:.8f} G oracle EIGMIN {ref['EIGMIN']:.2e} (scale {scale:.2e})", flush=True)
        assert abs(ref["EIGMIN"]) < 1e-6 * scale
        lo = _oracle_at(fb, x["FIELD"] - 1e-3)
        hi = _oracle_at(fb, x["FIELD"] + 1e-3)
        assert int(lo["NODEC"]) == x["NODEC_LO"] and int(hi["NODEC"]) == x["NODEC_HI"]
        assert abs(x["NODEC_HI"] - x["NODEC_LO"]) == 1


def test_scan_node_counts_vs_oracle(eng):
    fb = _deck()
    b = eng.block(fb.vl, fb.eint, fb.cent, mxlam=fb.mxlam, nconst=fb.nconst)
    (r0, dro, no), (r1, dri, ni) = S.bound_segments(RMIN, RMATCH, RMAX, DR)
    ro, ri = S.ld_radii(r0, dro, no), S.ld_radii(r1, dri, ni)
    go, gi = eng.grid(ro, fb.ptab(ro), dro), eng.grid(ri, fb.ptab(ri), dri)
    fields = np.array([0.0, 250.0, 500.0, 1000.0])
    pc = np.stack([fb.pconst(f) for f in fields])
    eint, _ = eng.ytrans(b, pc)
    out = eng.bound_match(b, go, gi, eint.min(axis=1) - DELTA * fb.cm2ru, pconst=pc)
    for m, f in enumerate(fields):
        ref = _oracle_at(fb, f)
        assert int(out["NODEC"][m]) == int(ref["NODEC"])
