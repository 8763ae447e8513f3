"""The LDS-resident MDPROP with one wave per instance (64 threads) and with
four (256): the arithmetic per element and the pivot decisions do not depend
on the thread count, so S-matrices and node counts must be bitwise the same
(N = 36, the Bunch-Kaufman path, and the N = 48 field deck with the fields
as the instances' constants)."""
import numpy as np
import pytest

from molscat_amd import decks
from molscat_amd import system as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def _both(eng, run):
    outs = []
    for nt in (64, 256):
        eng.set_lds_threads(nt)
        print(f"  lds threads {nt}", flush=True)
        outs.append(run())
    eng.set_lds_threads(0)
    return outs


def test_rotor36_bitwise(eng):
    deck = decks.rotor36()
    blk = deck.block()
    seg = deck.ld
    r = seg.radii()
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), seg.dr)
    ered = np.array([100.0, 250.0, 400.0]) * blk.cm2ru
    wv, nops = zip(*[S.gtwvec(e, blk.eint) for e in ered])
    nb = S.chnsrt(blk.eint)
    print(f"MDPROP N={blk.n} {seg.nstep} sectors", flush=True)
    a, c = _both(eng, lambda: eng.smatrix_ld(b, g, 6, ered, np.array(wv), np.array(nops), nb, seg.rstop))
    assert np.array_equal(a["NODES"], c["NODES"]) and np.all(a["STATUS"] == 0)
    for m in range(len(ered)):
        assert np.array_equal(a["SR"][m], c["SR"][m]) and np.array_equal(a["SI"][m], c["SI"][m])


def test_field48_bitwise(eng):
    fb = decks.field48()
    b = eng.block(fb.vl, fb.eint, fb.cent, mxlam=fb.mxlam, nconst=fb.nconst)
    nstep, h = S.drset_constant(3.5, 25.0, 0.02)
    r = S.ld_radii(3.5, h, nstep)
    g = eng.grid(r, fb.ptab(r), h)
    pc = np.stack([fb.pconst(f) for f in (10.0, 400.0, 800.0)])
    eint, _ = eng.ytrans(b, pc)
    ered = eint.min(axis=1) + 1e-3 * fb.cm2ru
    wv, nops = zip(*[S.gtwvec(ered[m], eint[m]) for m in range(3)])
    nb = np.stack([S.chnsrt(eint[m]) for m in range(3)])
    print(f"MDPROP field deck N={fb.n}", flush=True)
    a, c = _both(eng, lambda: eng.smatrix_ld(b, g, 6, ered, np.array(wv), np.array(nops), nb, 25.0, pconst=pc))
    assert np.array_equal(a["NODES"], c["NODES"]) and np.all(a["STATUS"] == 0)
    for m in range(3):
        assert np.array_equal(a["SR"][m], c["SR"][m]) and np.array_equal(a["SI"][m], c["SI"][m])
