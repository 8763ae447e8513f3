"""GPU parity of the large-N path (matrices beyond the LDS-resident limit:
blocked Bunch-Kaufman sweep in global memory) against the oracle."""
import numpy as np
import pytest

from molscat_amd import system as S
from oracle import molscat_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def _oracle_block(blk):
    pot = O.PowerPotential(blk.potl.terms, rm=blk.potl.rm, epsil=blk.potl.epsil)
    return O.Block(n=blk.n, vl=blk.vl, eint=blk.eint, cent=blk.cent, lvals=blk.lvals, levidx=blk.levidx,
                   potl=pot, rscale=blk.rscale, ep2ru=blk.ep2ru, cm2ru=blk.cm2ru)


@pytest.mark.parametrize("n", [140, 200, 333, 441])
def test_syminv_large_vs_oracle(eng, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((2, n, n))
    a = 0.5 * (a + np.transpose(a, (0, 2, 1)))
    # one indefinite with a zero-diagonal block, forcing 2x2 pivots
    a[1, : n // 4, : n // 4] = 0.0
    print(f"syminv large n={n}", flush=True)
    inv, k = eng.syminv(a)
    for m in range(2):
        ok = int(np.sum(np.linalg.eigvalsh(a[m]) < 0))
        assert k[m] == ok
        cond = np.linalg.cond(a[m])
        res = np.abs(a[m] @ inv[m] - np.eye(n)).max()
        # backward-stable inverse: residual ~ eps * cond * n
        assert res < 1e-14 * cond * n, (res, cond)
        assert np.array_equal(inv[m], inv[m].T)


def _big_block():
    # atom + rotor, even j = 0..24, JTOT=60, block 2: N = 169 (beyond the LDS limit)
    c = np.random.default_rng(1).uniform(0.05, 0.4, size=3)
    pot = S.lj_rotor_potential([1.0, *c], 3.5, 50.0)
    return S.rotor_block(60, 2, list(range(0, 25, 2)), 60.0, [0, 2, 4, 6], pot, 1.5, runit=1.0)


@pytest.mark.parametrize("iprop", [6, 5])
def test_smatrix_large_vs_oracle(eng, iprop):
    blk = _big_block()
    assert blk.n == 169
    rstart, rstop, nstep = 2.0, 8.0, 300
    dr = (rstop - rstart) / nstep
    r = S.ld_radii(rstart, dr, nstep)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), dr)
    energies = [300.0, 2000.0]
    ered = np.array(energies) * blk.cm2ru
    nb = S.chnsrt(blk.eint)
    wv, nops = zip(*[S.gtwvec(e, blk.eint) for e in ered])
    print(f"large iprop={iprop} N={blk.n} nstep={nstep} nopen={nops}", flush=True)
    out = eng.smatrix_ld(b, g, iprop, ered, np.array(wv), np.array(nops), nb, rstop)
    ob = _oracle_block(blk)
    for m, e in enumerate(ered):
        ref = O.smatrix_ld(ob, e, O.Grid(rstart, rstop, nstep, dr), iprop=iprop)
        assert out["STATUS"][m] == 0
        assert int(out["NODES"][m]) == int(ref["NODES"])
        tol = 1e-10 if iprop == 6 else 1e-9
        d = max(np.abs(out["SR"][m] - ref["SR"]).max(), np.abs(out["SI"][m] - ref["SI"]).max())
        print(f"  E={energies[m]} nopen={nops[m]} max|dS|={d:.2e} nodes={int(out['NODES'][m])}")
        assert d < tol
        s = out["SR"][m] + 1j * out["SI"][m]
        assert np.abs(s @ s.conj().T - np.eye(len(s))).max() < 1e-12


def test_yinit_large_vs_oracle(eng):
    blk = _big_block()
    r = S.ld_radii(2.0, 0.01, 4)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), 0.01)
    ered = 300.0 * blk.cm2ru
    y0 = eng.yinit(b, g, [ered])[0]
    oy0 = O.yinit(_oracle_block(blk), ered, r[0])
    assert np.abs(y0 - oy0).max() < 1e-11 * np.abs(oy0).max()


def test_team_size_bitwise_identical(eng):
    """Teams of 1, 4 and 32 workgroups per instance run the same arithmetic
    per element: Y(RSTOP) and the node count must be bitwise identical."""
    blk = _big_block()
    rstart, rstop, nstep = 2.0, 5.0, 60
    dr = (rstop - rstart) / nstep
    r = S.ld_radii(rstart, dr, nstep)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), dr)
    ered = 2000.0 * blk.cm2ru
    y0 = eng.yinit(b, g, [ered])
    outs = []
    eng.set_large_mode(1)  # the sequential matrix-core kernel
    for team in (1, 4, 32):
        eng.set_team_size(team)
        print(f"team={team}", flush=True)
        y, nodes, st = eng.prop_ld(b, g, 6, [ered], y0)
        assert st[0] == 0
        outs.append((y[0], int(nodes[0])))
    # the global-memory kernel (negative team size) sums in another order
    eng.set_team_size(-8)
    print("team=8 global-memory kernel", flush=True)
    yg, nodes_g, st = eng.prop_ld(b, g, 6, [ered], y0)
    eng.set_team_size(0)
    eng.set_large_mode(0)
    assert st[0] == 0
    for y, nodes in outs[1:]:
        assert nodes == outs[0][1]
        assert np.array_equal(y, outs[0][0])
    assert int(nodes_g[0]) == outs[0][1]
    assert np.abs(yg[0] - outs[0][0]).max() < 1e-12 * np.abs(yg[0]).max()


@pytest.mark.parametrize("nstep", [300, 301, 40])
def test_tree_matches_sequential(eng, nstep):
    """The parallel-in-R MDPROP (tree of merged sector maps) against the
    sequential recurrence on the same grid: node counts identical (the
    inertia of one block-tridiagonal matrix in two elimination orders), Y to
    1e-10 of its largest element.  An odd sector count passes a segment
    through a level unmerged.  At 40 sectors (DR = 0.15) some channel has
    |2 A14 - 8| < 0.5, so the leaves are built from U(C) instead of from
    V + E^-1."""
    blk = _big_block()
    rstart, rstop = 2.0, 8.0
    dr = (rstop - rstart) / nstep
    r = S.ld_radii(rstart, dr, nstep)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), dr)
    ered = np.array([300.0, 2000.0]) * blk.cm2ru
    y0 = eng.yinit(b, g, ered)
    print(f"tree vs sequential N={blk.n} nstep={nstep}", flush=True)
    eng.set_large_mode(1)
    ys, ns, ss = eng.prop_ld(b, g, 6, ered, y0)
    eng.set_large_mode(0)
    yt, nt, st = eng.prop_ld(b, g, 6, ered, y0)
    for m in range(len(ered)):
        assert ss[m] == 0 and st[m] == 0
        assert int(nt[m]) == int(ns[m])
        err = np.abs(yt[m] - ys[m]).max() / np.abs(ys[m]).max()
        print(f"  E={ered[m] / blk.cm2ru:.0f} nodes={int(nt[m])} rel|dY|={err:.2e}", flush=True)
        assert err < 1e-10


def test_plain_team_launch_bitwise(eng):
    """Ordinary launches of the team kernels (for tracing tools) run the same
    grids as the cooperative launches: Y(RSTOP) and the node counts of a
    tree whose small levels invert on teams are bitwise the same."""
    blk = _big_block()
    rstart, rstop, nstep = 2.0, 8.0, 40
    dr = (rstop - rstart) / nstep
    r = S.ld_radii(rstart, dr, nstep)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), dr)
    ered = np.array([300.0, 2000.0]) * blk.cm2ru
    y0 = eng.yinit(b, g, ered)
    print("team launches: cooperative", flush=True)
    y1, n1, s1 = eng.prop_ld(b, g, 6, ered, y0)
    eng.set_team_launch(False)
    print("team launches: ordinary", flush=True)
    y0b = eng.yinit(b, g, ered)
    y2, n2, s2 = eng.prop_ld(b, g, 6, ered, y0b)
    eng.set_team_launch(True)
    assert np.array_equal(y0, y0b)
    assert np.array_equal(s1, s2) and np.array_equal(n1, n2)
    assert np.array_equal(y1, y2)
