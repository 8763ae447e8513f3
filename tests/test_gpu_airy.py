"""GPU parity of the AIRY propagator (mcp_prop_airy, mcp_smatrix_segments
with IPROP = 9), through the C ABI.

System: the reference's molscat-all_propagators deck (ITYPE=1 rotor, BE=30,
j=0,2,4,6, URED=20, JTOT=10, IB=1), the 'IPROPS=9, IPROPL=9' run: RSTART
from FINDRM, AIRY with constant steps to RMID=1.0 and steps growing as
R^(4/3) to RMAX=20 (TOLHIS = TOLHIL = 0, STEPS=12).  Checks: the S-matrices
at 1000 and 900 cm-1 against the reference's printed values (the oracle,
whose Airy propagator is evaluated with mpmath, sits 2.3e-10 from them, so
the tolerance is 5e-10); the propagation alone against the oracle on the
first segment (outward) and on an inward segment, |dZ| <= 1e-10 relative,
node counts exact."""
import json
import os

import numpy as np
import pytest

from molscat_amd import system as S
from oracle import molscat_oracle as O

pytestmark = pytest.mark.gpu

POT_TERMS = [[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]]


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def system():
    pot = S.PowerPotential(POT_TERMS, rm=3.5, epsil=50.0)
    blk = S.rotor_block(10, 1, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)
    opot = O.PowerPotential(pot.terms, rm=pot.rm, epsil=pot.epsil)
    ob = O.Block(n=blk.n, vl=blk.vl, eint=blk.eint, cent=blk.cent, lvals=blk.lvals, levidx=blk.levidx, potl=opot,
                 rscale=blk.rscale, ep2ru=blk.ep2ru, cm2ru=blk.cm2ru)
    eredmx = 1000.0 * blk.cm2ru
    rstart, _ = O.findrm(ob, 0.5, eredmx, 9)
    cays = O.calck(0.0, eredmx, blk.eint)
    drfit = np.pi / (cays * 12.0)
    n1, dr1 = S.drset_airy(rstart, 1.0, drfit, 0.0)
    n2, dr2 = S.drset_airy(1.0, 20.0, drfit, 4.0 / 3.0)
    return blk, ob, rstart, (n1, dr1), (n2, dr2)


def _airy_grid(eng, blk, r0, r1, nstep, dr0, powr):
    r, w = S.airy_grid(r0, r1, nstep, dr0, powr)
    return eng.airy_grid(r, w, blk.ptab(r)), r, w


def test_airy_smatrix_matches_reference(eng, system, golden_dir):
    blk, ob, rstart, (n1, dr1), (n2, dr2) = system
    assert (n1, n2) == (149, 791)  # the step counts the reference prints
    with open(os.path.join(golden_dir, "ref_all_propagators_smat.json")) as f:
        sec = json.load(f)["sections"]["model system: ITYPE=1; IPROPS=9, IPROPL=9"]
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g1, _, _ = _airy_grid(eng, blk, rstart, 1.0, n1, dr1, 0.0)
    g2, _, _ = _airy_grid(eng, blk, 1.0, 20.0, n2, dr2, 4.0 / 3.0)
    energies = [1000.0, 900.0]
    ered = np.array(energies) * blk.cm2ru
    nb = S.chnsrt(blk.eint)
    wv, nops = zip(*[S.gtwvec(e, blk.eint) for e in ered])
    print(f"AIRY N={blk.n} steps {n1}+{n2} energies={energies}", flush=True)
    out = eng.smatrix_segments(b, [g1, g2], [9, 9], ered, np.array(wv), np.array(nops), nb, 20.0)
    for m in range(len(energies)):
        assert out["STATUS"][m] == 0
        ref = sec["smat"][m]
        dsr = np.abs(out["SR"][m] - np.array(ref["sr"])).max()
        dsi = np.abs(out["SI"][m] - np.array(ref["si"])).max()
        print(f"  E={energies[m]} max|dSR|={dsr:.2e} max|dSI|={dsi:.2e} nodes={int(out['NODES'][m])}", flush=True)
        assert dsr < 5e-10 and dsi < 5e-10
        s = out["SR"][m] + 1j * out["SI"][m]
        assert np.abs(s @ s.conj().T - np.eye(len(s))).max() < 1e-12


@pytest.mark.parametrize("direction", ["outward", "inward"])
def test_airy_propagation_vs_oracle(eng, system, direction):
    blk, ob, rstart, (n1, dr1), _ = system
    if direction == "outward":
        r0, r1, nstep, dr0, powr = rstart, 1.0, n1, dr1, 0.0
    else:
        r0, r1, nstep, powr = 5.0, 3.0, 40, 0.0
        dr0 = -(r0 - r1) / nstep
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g, _, _ = _airy_grid(eng, blk, r0, r1, nstep, dr0, powr)
    ered = np.array([1000.0, 700.0]) * blk.cm2ru
    z0 = np.stack([O.yinit(ob, e, r0, zout=(direction == "outward"), bcyomn=0.0) for e in ered])
    z, nodes, st = eng.prop_airy(b, g, ered, z0)
    for m, e in enumerate(ered):
        zr, nr = O.aiprop(z0[m], ob, e, r0, r1, nstep, dr0, powr)
        assert st[m] == 0
        assert int(nodes[m]) == nr, (direction, int(nodes[m]), nr)
        err = np.abs(z[m] - zr).max() / max(1.0, np.abs(zr).max())
        print(f"  {direction} E={e / blk.cm2ru:.0f} nodes={nr} rel|dZ|={err:.2e}", flush=True)
        assert err < 1e-10
