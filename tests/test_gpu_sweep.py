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


def test_sweep_driver_sigma_vs_per_block_and_oracle(eng):
    """The sweep driver (molscat_amd/driver.py:sweep) on a small rotor
    (even j = 0..12, JTOT 0, 3, 7, both parities, three energies, LDMD 2-10 A
    then AIRY 10-30 A with TOLHI = 0): sigma summed over the blocks from the
    persistent multi-block launch, from one batched call per block, and from
    the oracle (YINIT, MDPROP, AIPROP, YTOK, KTOS and SPROC per block,
    source_code/sproc.f:380-430) agree to 1e-10 relative; node counts equal."""
    from molscat_amd import driver

    deck = decks.rotor441()
    pot = deck.potential()
    jl = list(range(0, 13, 2))
    energies = [100.0, 300.0, 450.0]
    rstart, rmid, dr = 2.0, 10.0, 0.01
    nstep, h = S.drset_constant(rstart, rmid, dr)
    r = S.ld_radii(rstart, h, nstep)
    nair, dr0 = S.drset_airy(rmid, 30.0, 0.2, 4.0 / 3.0)
    ra, wa = S.airy_grid(rmid, 30.0, nair, dr0, 4.0 / 3.0)
    args = (jl, decks.BE, deck.lambdas, pot, decks.URED, energies, [0, 3, 7], [1, 2], (r, h), (ra, wa), 30.0)
    print(f"sweep driver: LD {nstep} sectors, AIRY {nair} steps", flush=True)
    multi = driver.sweep(eng, *args)
    per_block = driver.sweep(eng, *args, small_max=0)
    scale = np.abs(multi.sigma).max()
    d = np.abs(multi.sigma - per_block.sigma).max() / scale
    print(f"  sigma multi vs per-block rel {d:.1e}", flush=True)
    assert d < 1e-10
    for a, b in zip(multi.blocks, per_block.blocks):
        assert a.nodes == b.nodes
    # the oracle, block by block
    acc = np.zeros_like(multi.sigma)
    nsig = multi.sigma.shape[1]
    for run in multi.blocks:
        blk = S.rotor_block(run.jtot, run.ib, jl, decks.BE, deck.lambdas, pot, decks.URED, runit=1.0)
        ob = O.Block(n=blk.n, vl=blk.vl, eint=blk.eint, cent=blk.cent, lvals=blk.lvals, levidx=blk.levidx,
                     potl=O.PowerPotential(blk.potl.terms, rm=blk.potl.rm, epsil=blk.potl.epsil),
                     rscale=blk.rscale, ep2ru=blk.ep2ru, cm2ru=blk.cm2ru)
        nb = O.chnsrt(ob.eint)
        for m, e in enumerate(energies):
            ered = e * blk.cm2ru
            wv, nop = O.gtwvec(ered, ob.eint)
            y = O.yinit(ob, ered, rstart)
            y, n1 = O.mdprop(y, lambda rr: ob.hammat(rr, ered), rstart, nstep, h)
            z, n2 = O.aiprop(y, ob, ered, rmid, 30.0, nair, dr0, 4.0 / 3.0)
            assert run.nodes[m] == n1 + n2, (run.jtot, run.ib, m)
            if nop == 0:
                continue
            k = O.ytok(nb, wv, blk.n, nop, z, 30.0, ob.cent)
            sr, si = O.ktos(k)
            acc[m] += O.sigma_block(sr, si, nb, nop, blk.levidx, jl, wv, run.jtot, 1.0, nsig)
    d = np.abs(multi.sigma - acc).max() / scale
    print(f"  sigma vs oracle rel {d:.1e}", flush=True)
    assert d < 1e-10


def test_channel_at_threshold_is_the_zero_energy_limit(eng):
    """An energy exactly at a threshold (360 cm-1 = B j(j+1) for j = 2, one of
    the 64 sweep energies) gives those channels k = 0; the matching uses the
    k -> 0 limit of their closed-channel log-derivative, -ell/R.  S at the
    threshold must be the limit of S just below it: |S(E_thr) - S(E_thr -
    d)| falls as sqrt(d), the closed channels' kappa (an s-wave one has
    log-derivative -kappa), from 1e-7 to 1e-9 cm-1 below by a factor 10."""
    deck = decks.rotor441()
    pot = deck.potential()
    blk = S.rotor_block(2, 2, list(range(0, 13, 2)), decks.BE, deck.lambdas, pot, decks.URED, runit=1.0)
    rstart, rstop, dr = 2.0, 12.0, 0.01
    nstep, h = S.drset_constant(rstart, rstop, dr)
    r = S.ld_radii(rstart, h, nstep)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), h)
    ecm = [360.0, 360.0 - 1e-7, 360.0 - 1e-9]
    ered = np.array(ecm) * blk.cm2ru
    wv, nops = zip(*[S.gtwvec(e, blk.eint) for e in ered])
    assert np.any(np.array(wv[0]) == 0.0) and len(set(nops)) == 1
    nb = S.chnsrt(blk.eint)
    out = eng.smatrix_ld(b, g, 6, ered, np.array(wv), np.array(nops), nb, rstop)
    assert np.all(out["STATUS"] == 0)
    s = [out["SR"][m] + 1j * out["SI"][m] for m in range(3)]
    d7 = np.abs(s[0] - s[1]).max()
    d9 = np.abs(s[0] - s[2]).max()
    print(f"  |S(thr) - S(thr - 1e-7)| {d7:.1e}, |S(thr) - S(thr - 1e-9)| {d9:.1e}", flush=True)
    assert d9 < 1e-6 and 0.07 * d7 < d9 < 0.13 * d7
