"""GPU parity of the LDS-resident path against the oracle, through the C ABI.

Inputs are seeded; tolerances are written in each test.  The oracle and the
GPU get identical block data (VL, EINT, CENT, radii, P table), so the
comparison is of the kernels alone."""
import numpy as np
import pytest

from molscat_amd import system as S
from oracle import molscat_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    print("engine:", e.info(), flush=True)
    yield e
    e.close()


def _oracle_block(blk: S.RotorBlock) -> O.Block:
    pot = O.PowerPotential(blk.potl.terms, rm=blk.potl.rm, epsil=blk.potl.epsil)
    return O.Block(n=blk.n, vl=blk.vl, eint=blk.eint, cent=blk.cent, lvals=blk.lvals, levidx=blk.levidx,
                   potl=pot, rscale=blk.rscale, ep2ru=blk.ep2ru, cm2ru=blk.cm2ru)


def _rand_sym(n, rng):
    a = rng.standard_normal((n, n))
    return 0.5 * (a + a.T)


@pytest.mark.parametrize("n", [1, 2, 7, 12, 30, 31, 36, 48, 64])
def test_syminv_batched_vs_oracle(eng, n):
    rng = np.random.default_rng(100 + n)
    a = np.stack([_rand_sym(n, rng) for _ in range(6)])
    inv, k = eng.syminv(a)
    for m in range(len(a)):
        oinv, ok = O.syminv(a[m])
        assert k[m] == ok == int(np.sum(np.linalg.eigvalsh(a[m]) < 0))
        cond = np.linalg.cond(a[m])
        # tolerance: rounding of two different elimination orders, scaled by cond
        assert np.abs(inv[m] - oinv).max() <= 1e-13 * cond * max(1.0, np.abs(oinv).max())
        assert np.array_equal(inv[m], inv[m].T)  # bitwise symmetric


def test_syminv_zero_pivot_and_singular(eng):
    a = np.array([[0.0, 1.0], [1.0, 2.0]])
    _, k = eng.syminv(a)
    assert k == 3  # N+1 (source_code/syminv.f:95)
    z = np.zeros((40, 40))
    _, k = eng.syminv(z)
    assert k == -1  # DSYTRF INFO=1


@pytest.mark.parametrize("n", [1, 5, 12, 36])
def test_eigh_vs_numpy(eng, n):
    rng = np.random.default_rng(7 + n)
    a = np.stack([_rand_sym(n, rng) for _ in range(3)])
    ev, vec = eng.eigh(a)
    for m in range(3):
        ref = np.linalg.eigvalsh(a[m])
        assert np.abs(ev[m] - ref).max() < 1e-13 * max(1.0, np.abs(ref).max())
        rec = vec[m] @ np.diag(ev[m]) @ vec[m].T
        assert np.abs(rec - a[m]).max() < 1e-12 * max(1.0, np.abs(a[m]).max())


def _allprop_block():
    pot = S.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    return S.rotor_block(10, 1, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)


def _config0_block():
    # rotor B=60 cm-1, even j=0..10, J=10 block 2, URED=1.5, LJ eps=50 Rm=3.5, seeded c_lambda
    c = np.random.default_rng(0).uniform(0.1, 0.4, size=2)
    pot = S.lj_rotor_potential([1.0, c[0], c[1]], 3.5, 50.0)
    return S.rotor_block(10, 2, [0, 2, 4, 6, 8, 10], 60.0, [0, 2, 4], pot, 1.5, runit=1.0)


def test_hammat_vs_oracle(eng):
    blk = _config0_block()
    r = S.ld_radii(2.0, 0.01, 50)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), 0.01)
    ered = 100.0 * blk.cm2ru
    idx = [0, 1, 37, 100]
    w = eng.hammat(b, g, idx, ered)
    ob = _oracle_block(blk)
    for m, k in enumerate(idx):
        ow, _ = ob.hammat(r[k], ered)
        assert np.abs(w[m] - ow).max() <= 1e-13 * np.abs(ow).max()


@pytest.mark.parametrize("case", ["allprop", "config0"])
@pytest.mark.parametrize("iprop", [6, 5])
def test_smatrix_vs_oracle(eng, case, iprop):
    if case == "allprop":
        blk = _allprop_block()
        rstart, rstop, nstep, energies = 0.641657747984697, 20.0, 8072, [1000.0, 900.0]
    else:
        blk = _config0_block()
        rstart, rstop, nstep, energies = 2.0, 30.0, 2800, [100.0, 250.0]
    dr = (rstop - rstart) / nstep
    r = S.ld_radii(rstart, dr, nstep)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), dr)
    ered = np.array(energies) * blk.cm2ru
    nb = S.chnsrt(blk.eint)
    wv, nops = zip(*[S.gtwvec(e, blk.eint) for e in ered])
    print(f"case={case} iprop={iprop} N={blk.n} nstep={nstep} energies={energies}", flush=True)
    out = eng.smatrix_ld(b, g, iprop, ered, np.array(wv), np.array(nops), nb, rstop)
    ob = _oracle_block(blk)
    for m, e in enumerate(ered):
        ref = O.smatrix_ld(ob, e, O.Grid(rstart, rstop, nstep, dr), iprop=iprop)
        assert out["STATUS"][m] == 0
        assert int(out["NODES"][m]) == int(ref["NODES"])  # bit-exact integer
        # FP64 GPU vs FP64 oracle, absolute on S (|S| <= 1): measured at most
        # 1.9e-12 (MDPROP) and 2.2e-12 (LDPROP) on these decks
        tol = 1e-11
        dsr = np.abs(out["SR"][m] - ref["SR"]).max()
        dsi = np.abs(out["SI"][m] - ref["SI"]).max()
        print(f"  E={energies[m]} max|dSR|={dsr:.2e} max|dSI|={dsi:.2e} nodes={int(out['NODES'][m])}")
        assert dsr < tol and dsi < tol
        s = out["SR"][m] + 1j * out["SI"][m]
        assert np.abs(s @ s.conj().T - np.eye(len(s))).max() < 1e-12


@pytest.mark.parametrize("iprop", [6, 5])
def test_propagation_only_vs_oracle(eng, iprop):
    """Y(RSTOP) from the oracle's own start matrix: isolates the propagator."""
    blk = _allprop_block()
    rstart, rstop, nstep = 0.641657747984697, 20.0, 8072
    dr = (rstop - rstart) / nstep
    r = S.ld_radii(rstart, dr, nstep)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), dr)
    ered = 1000.0 * blk.cm2ru
    ob = _oracle_block(blk)
    y0 = O.yinit(ob, ered, rstart)
    y, nodes, st = eng.prop_ld(b, g, iprop, [ered], y0[None])
    f = O.mdprop if iprop == 6 else O.ldprop
    oy, onodes = f(y0.copy(), lambda rr: ob.hammat(rr, ered), rstart, nstep, dr)
    assert st[0] == 0 and int(nodes[0]) == onodes
    # compare through the oracle's matching: the open-channel block decides S
    wv, nop = S.gtwvec(ered, blk.eint)
    nb = S.chnsrt(blk.eint)
    sg = O.ktos(O.ytok(nb, wv, blk.n, nop, y[0], rstop, blk.cent))
    so = O.ktos(O.ytok(nb, wv, blk.n, nop, oy, rstop, blk.cent))
    d = max(np.abs(sg[0] - so[0]).max(), np.abs(sg[1] - so[1]).max())
    print(f"iprop={iprop} propagation-only max|dS| = {d:.2e} nodes {int(nodes[0])} vs {onodes}")
    assert d < 1e-11  # measured 2.9e-13 (MDPROP), 4.7e-13 (LDPROP)


def test_match_on_oracle_y(eng):
    """YTOK/KTOS on the oracle's Y(RSTOP): isolates the matching kernel."""
    blk = _allprop_block()
    rstart, rstop, nstep = 0.641657747984697, 20.0, 8072
    dr = (rstop - rstart) / nstep
    ered = 1000.0 * blk.cm2ru
    ob = _oracle_block(blk)
    y0 = O.yinit(ob, ered, rstart)
    oy, _ = O.mdprop(y0.copy(), lambda rr: ob.hammat(rr, ered), rstart, nstep, dr)
    wv, nop = S.gtwvec(ered, blk.eint)
    nb = S.chnsrt(blk.eint)
    k, sr, si, st = eng.ytok_ktos(oy[None], wv, [nop], nb, blk.cent, rstop)
    ok = O.ytok(nb, wv, blk.n, nop, oy, rstop, blk.cent)
    osr, osi = O.ktos(ok)
    d = max(np.abs(sr[0] - osr).max(), np.abs(si[0] - osi).max())
    print(f"match-only max|dS| = {d:.2e}")
    assert st[0] == 0 and d < 1e-12


def test_yinit_and_match_vs_oracle(eng):
    blk = _config0_block()
    r = S.ld_radii(2.0, 0.01, 10)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), 0.01)
    ered = 100.0 * blk.cm2ru
    y0 = eng.yinit(b, g, [ered])[0]
    oy0 = O.yinit(_oracle_block(blk), ered, r[0])
    assert np.abs(y0 - oy0).max() < 1e-12 * np.abs(oy0).max()
    # matching on a synthetic symmetric Y at R=30 (N=36, 1 open channel)
    rng = np.random.default_rng(5)
    y = _rand_sym(blk.n, rng) + np.diag(np.abs(blk.eint - ered) ** 0.5)
    wv, nop = S.gtwvec(ered, blk.eint)
    nb = S.chnsrt(blk.eint)
    k, sr, si, st = eng.ytok_ktos(y[None], wv, [nop], nb, blk.cent, 30.0)
    ok = O.ytok(nb, wv, blk.n, nop, y, 30.0, blk.cent)
    osr, osi = O.ktos(ok)
    assert st[0] == 0
    assert np.abs(k[0] - ok).max() < 1e-12 * max(1.0, np.abs(ok).max())
    assert np.abs(sr[0] - osr).max() < 1e-12 and np.abs(si[0] - osi).max() < 1e-12


def _single_channel_block():
    # JTOT = 0, j = 0 only: the single channel (j=0, L=0) of parity block 2
    pot = S.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    return S.rotor_block(0, 2, [0], 30.0, [0, 2], pot, 20.0)


def test_odprop_single_channel_vs_oracle(eng):
    """N = 1: mcp_smatrix_ld follows SCCTRL and runs ODPROP; mcp_prop_ld with
    iprop 6 runs MDPROP.  Both against the oracle; tolerance 1e-10 on S."""
    blk = _single_channel_block()
    assert blk.n == 1
    rstart, rstop, nstep = 0.64, 20.0, 4000
    dr = (rstop - rstart) / nstep
    r = S.ld_radii(rstart, dr, nstep)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), dr)
    energies = [50.0, 400.0, 900.0]  # k*DR/2 <= 0.57
    ered = np.array(energies) * blk.cm2ru
    nb = S.chnsrt(blk.eint)
    wv, nops = zip(*[S.gtwvec(e, blk.eint) for e in ered])
    print(f"ODPROP N=1 nstep={nstep} energies={energies}", flush=True)
    out = eng.smatrix_ld(b, g, 6, ered, np.array(wv), np.array(nops), nb, rstop)
    ob = _oracle_block(blk)
    grid = O.Grid(rstart, rstop, nstep, dr)
    for m, e in enumerate(ered):
        ref = O.smatrix_ld(ob, e, grid, iprop=6)  # the oracle also dispatches N=1 to ODPROP
        assert out["STATUS"][m] == 0 and int(out["NODES"][m]) == 0
        assert abs(out["SR"][m][0, 0] - ref["SR"][0, 0]) < 1e-10
        assert abs(out["SI"][m][0, 0] - ref["SI"][0, 0]) < 1e-10
    # MDPROP at N = 1 through mcp_prop_ld
    y0 = np.stack([O.yinit(ob, e, rstart) for e in ered])
    y, nodes, st = eng.prop_ld(b, g, 6, ered, y0)
    for m, e in enumerate(ered):
        yr, nr = O.mdprop(y0[m], lambda rr: ob.hammat(rr, e), rstart, nstep, dr)
        print(f"  MDPROP N=1 E={energies[m]}: nodes GPU {int(nodes[m])} oracle {nr}", flush=True)
        assert st[m] == 0 and int(nodes[m]) == nr
        assert abs(y[m][0, 0] - yr[0, 0]) < 1e-10 * max(1.0, abs(yr[0, 0]))
