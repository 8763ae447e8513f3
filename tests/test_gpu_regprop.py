"""The register-resident MDPROP (one wavefront per instance, each lane one
row of Y, mcp_set_md_kernel(ctx, 2)) against the LDS-resident kernel
(mcp_set_md_kernel(ctx, 1)) and the oracle: the same recurrence and pivot
decisions with a different order of rounding.  Node counts are equal; the
S-matrices of the two kernels agree to 1e-9 over every energy of a sweep
that crosses resonances (the rotor decks: energies where Y is near a pole
amplify rounding differences; the LDS kernel sits 1e-9 from the oracle's
Y there too) and to 1e-10 on the field deck; against the oracle S to 1e-10
with nodes exact at fixed instances.  Cases: N = 12 (natural-order
pivots, NP = 16), N = 36 (Bunch-Kaufman, NP = 48), the N = 48 hyperfine
field deck with the fields as the instances' constants (NP = 48), and a
rotor block of N = 56 (NP = 64); the oracle's MDPROP on a few instances of
each."""
import numpy as np
import pytest

from molscat_amd import decks
from molscat_amd import system as S
from oracle import molscat_oracle as O

from test_oracle_ytrans import _oracle_block

pytestmark = pytest.mark.gpu

POT_TERMS = [[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]]


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.set_md_kernel(0)
    e.close()


def _both(eng, run):
    out = {}
    for mode in (1, 2):
        eng.set_md_kernel(mode)
        print(f"  md_kernel {mode}", flush=True)
        out[mode] = run()
    eng.set_md_kernel(0)
    return out[1], out[2]


def _compare_y(lds, reg):
    (y1, n1, s1), (y2, n2, s2) = lds, reg
    assert np.all(s1 == 0) and np.all(s2 == 0), (s1, s2)
    assert np.array_equal(n1, n2), (n1, n2)
    scale = np.abs(y1).max(axis=(1, 2))
    err = np.abs(y1 - y2).max(axis=(1, 2)) / scale
    print(f"  Y register vs LDS: rel|dY| median {np.median(err):.1e} max {err.max():.1e} (instance {err.argmax()}), "
          f"nodes {int(n1.min())}..{int(n1.max())}", flush=True)
    # the output is symmetric (the lower triangle mirrored)
    assert np.array_equal(y2, np.transpose(y2, (0, 2, 1)))
    return err


def _s_diff(a, b):
    return np.array([max(np.abs(x - y).max(), np.abs(u - v).max())
                     for x, y, u, v in zip(a["SR"], b["SR"], a["SI"], b["SI"])])


def _compare_s(lds, reg, tol):
    assert np.all(lds["STATUS"] == 0) and np.all(reg["STATUS"] == 0)
    assert np.array_equal(lds["NODES"], reg["NODES"])
    d = _s_diff(lds, reg)
    print(f"  S register vs LDS: max|dS| = {d.max():.2e} (instance {d.argmax()}), median {np.median(d):.1e}",
          flush=True)
    assert d.max() < tol
    return int(d.argmax())


def _rotor_case(eng, jlev, jtot, ib, lambdas, rstart, rstop, nstep, energies):
    pot = S.PowerPotential(POT_TERMS if len(lambdas) == 2 else [[(1.0, -12), (-2.0, -6)]] * len(lambdas),
                           rm=3.5, epsil=50.0)
    blk = S.rotor_block(jtot, ib, jlev, 30.0, lambdas, pot, 20.0)
    dr = (rstop - rstart) / nstep
    r = S.ld_radii(rstart, dr, nstep)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), dr)
    ob = O.Block(n=blk.n, vl=blk.vl, eint=blk.eint, cent=blk.cent, lvals=blk.lvals, levidx=blk.levidx,
                 potl=O.PowerPotential(pot.terms, rm=pot.rm, epsil=pot.epsil), rscale=blk.rscale, ep2ru=blk.ep2ru,
                 cm2ru=blk.cm2ru)
    ered = np.asarray(energies) * blk.cm2ru
    y0 = np.stack([O.yinit(ob, e, rstart) for e in ered])
    return blk, ob, b, g, ered, y0, dr


def _oracle_s_err(out, m, ref):
    return max(np.abs(out["SR"][m] - ref["SR"]).max(), np.abs(out["SI"][m] - ref["SI"]).max())


def _rotor_check(eng, blk, ob, b, g, ered, y0, dr, rstart, rstop, nstep, picks):
    _compare_y(*_both(eng, lambda: eng.prop_ld(b, g, 6, ered, y0)))
    wv, nops = zip(*[S.gtwvec(e, blk.eint) for e in ered])
    nb = S.chnsrt(blk.eint)
    lds, reg = _both(eng, lambda: eng.smatrix_ld(b, g, 6, ered, np.array(wv), np.array(nops), nb, rstop))
    worst = _compare_s(lds, reg, 1e-8)
    grid = O.Grid(rstart, rstop, nstep, dr)
    for m in picks:
        ref = O.smatrix_ld(ob, ered[m], grid, iprop=6)
        d = _oracle_s_err(reg, m, ref)
        print(f"  [{m}] oracle nodes={int(ref['NODES'])} max|dS|={d:.2e}", flush=True)
        assert int(reg["NODES"][m]) == int(ref["NODES"]) and d < 1e-10
    # where the kernels differ most (an energy near a pole of Y), the
    # register kernel is as close to the oracle as the LDS kernel
    ref = O.smatrix_ld(ob, ered[worst], grid, iprop=6)
    dl, dr_ = _oracle_s_err(lds, worst, ref), _oracle_s_err(reg, worst, ref)
    print(f"  worst instance {worst}: |dS| vs oracle LDS {dl:.2e} register {dr_:.2e}", flush=True)
    assert dr_ < max(1e-10, 3.0 * dl)


@pytest.mark.parametrize("jlev,jtot,ib,nexp", [([0, 2, 4, 6], 10, 1, 12), ([0, 2, 4, 6, 8, 10], 10, 2, 36)])
def test_rotor_register_vs_lds_and_oracle(eng, jlev, jtot, ib, nexp):
    energies = np.linspace(900.0, 1100.0, 600)
    rstart, rstop, nstep = 0.7, 12.0, 400
    blk, ob, b, g, ered, y0, dr = _rotor_case(eng, jlev, jtot, ib, [0, 2], rstart, rstop, nstep, energies)
    assert blk.n == nexp
    print(f"rotor N={blk.n}, {len(ered)} energies, {nstep} sectors", flush=True)
    _rotor_check(eng, blk, ob, b, g, ered, y0, dr, rstart, rstop, nstep, (0, 299, 599))


def test_rotor56_register_vs_lds(eng):
    energies = np.linspace(1500.0, 1700.0, 520)
    rstart, rstop, nstep = 0.7, 10.0, 300
    blk, ob, b, g, ered, y0, dr = _rotor_case(eng, [0, 2, 4, 6, 8, 10, 12, 14], 14, 1, [0, 2], rstart, rstop,
                                              nstep, energies)
    assert blk.n == 56, blk.n
    print(f"rotor N={blk.n} (rows of 64), {len(ered)} energies", flush=True)
    _rotor_check(eng, blk, ob, b, g, ered, y0, dr, rstart, rstop, nstep, (257,))


def test_hyperfine48_fields_register_vs_lds_and_oracle(eng):
    fb = decks.hyperfine48()
    b = eng.block(fb.vl, fb.eint, fb.cent, mxlam=fb.mxlam, nconst=fb.nconst)
    nstep, h = S.drset_constant(3.5, 25.0, 0.02)
    r = S.ld_radii(3.5, h, nstep)
    g = eng.grid(r, fb.ptab(r), h)
    fields = np.linspace(0.0, 1000.0, 1024)
    pc = np.stack([fb.pconst(f) for f in fields])
    eint, _ = eng.ytrans(b, pc)
    ered = eint.min(axis=1) + 1.0e-4 * fb.cm2ru
    y0 = eng.yinit(b, g, ered, pconst=pc)
    print(f"hyperfine N={fb.n}, {len(fields)} fields, {nstep} sectors", flush=True)
    _compare_y(*_both(eng, lambda: eng.prop_ld(b, g, 6, ered, y0, pconst=pc)))
    wv = np.empty_like(eint)
    nops = np.empty(len(fields), dtype=np.int32)
    nb = np.empty(eint.shape, dtype=np.int32)
    for m in range(len(fields)):
        wv[m], nops[m] = S.gtwvec(ered[m], eint[m])
        nb[m] = S.chnsrt(eint[m])
    lds, reg = _both(eng, lambda: eng.smatrix_ld(b, g, 6, ered, wv, nops, nb, 25.0, pconst=pc))
    worst = _compare_s(lds, reg, 1e-7)
    grid = O.Grid(3.5, 25.0, nstep, h)
    for m in (3, 700):
        ob = _oracle_block(fb, fields[m])
        ref = O.smatrix_ld(ob, ered[m], grid, iprop=6)
        d = _oracle_s_err(reg, m, ref)
        print(f"  B={fields[m]:.3f} G oracle nodes={int(ref['NODES'])} max|dS|={d:.2e}", flush=True)
        assert int(reg["NODES"][m]) == int(ref["NODES"]) and d < 1e-10
    ob = _oracle_block(fb, fields[worst])
    ref = O.smatrix_ld(ob, ered[worst], grid, iprop=6)
    dl, dr_ = _oracle_s_err(lds, worst, ref), _oracle_s_err(reg, worst, ref)
    print(f"  worst instance {worst} (B={fields[worst]:.3f} G): |dS| vs oracle LDS {dl:.2e} register {dr_:.2e}",
          flush=True)
    assert dr_ < max(1e-10, 3.0 * dl)
