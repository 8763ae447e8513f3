"""GPU parity of MAPROP (IPROP = 7 of mcp_prop_ld / mcp_smatrix_ld, the
quasi-adiabatic modified log-derivative propagator) through the C ABI:
- the reference's 'IPROPS=7' run (examples/output/molscat-all_
  propagators.output; ITYPE=1 rotor, N = 12, 8072 sectors from FINDRM's
  RSTART to 20 RM): YINIT + MAPROP + YTOK/KTOS on the device against the
  printed S-matrices at both energies, node counts against the oracle's;
- a constant coupled W, where MAPROP is exact: Y against the exact solution
  in W's eigenchannels and NODES against their zeros."""
import json
import math
import os

import numpy as np
import pytest

from molscat_amd import system as S
from oracle import molscat_oracle as O

pytestmark = pytest.mark.gpu

LABEL = "model system: ITYPE=1; IPROPS=7"


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def _grid(eng, blk, rstart, dr, nstep):
    r = S.ld_radii(rstart, dr, nstep)
    ptab = np.stack([blk.ep2ru * np.clip(blk.potl(x * blk.rscale), -1e60, 1e60) for x in r])
    return eng.grid(r, ptab, dr)


def test_maprop_fixture_smatrices(eng, golden_dir):
    with open(os.path.join(golden_dir, "ref_all_propagators_smat.json")) as f:
        sec = json.load(f)["sections"][LABEL]
    pot = O.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    blk, _ = O.make_itype1_block(10, 1, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)
    grid, _ = O.ld_grid(blk, 0.5, 20.0, 1000.0 * blk.cm2ru, steps=12.0, irmset=9)
    assert grid.nstep == sec["props"][0][3] == 8072
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.vl.shape[0])
    g = _grid(eng, blk, grid.rstart, grid.dr, grid.nstep)
    ered = np.array([1000.0, 900.0]) * blk.cm2ru
    wv, nops = zip(*[S.gtwvec(e, blk.eint) for e in ered])
    nb = S.chnsrt(blk.eint)
    print(f"MAPROP N={blk.n} {grid.nstep} sectors", flush=True)
    out = eng.smatrix_ld(b, g, 7, ered, np.array(wv), np.array(nops), nb, grid.rstop)
    for m in range(2):
        assert out["STATUS"][m] == 0
        sm = sec["smat"][m]
        d = max(np.abs(out["SR"][m] - np.array(sm["sr"])).max(), np.abs(out["SI"][m] - np.array(sm["si"])).max())
        ref = O.smatrix_ld(blk, ered[m], grid, iprop=7)
        print(f"  E={[1000, 900][m]} cm-1 max|dS| vs printed {d:.2e} nodes={int(out['NODES'][m])} "
              f"(oracle {ref['NODES']})", flush=True)
        # the oracle sits 3.8e-12 and 5.4e-12 from the printed values
        assert d < 2e-11
        assert int(out["NODES"][m]) == int(ref["NODES"])


def test_maprop_constant_coupling_exact(eng):
    rng = np.random.default_rng(3)
    n = 3
    m = rng.standard_normal((n, n))
    wc = 0.5 * (m + m.T)
    vl = np.array([[wc[i, j] for j in range(n) for i in range(j + 1)]])
    eint = np.array([0.0, 1.5, 12.0])
    b = eng.block(vl, eint, np.zeros(n), mxlam=1)
    ered = 5.0
    w = wc + np.diag(eint - ered)
    lam, u = np.linalg.eigh(w)
    rs, re, r0 = 0.5, 6.0, 0.1
    nstep = 300
    dr = (re - rs) / nstep
    r = S.ld_radii(rs, dr, nstep)
    g = eng.grid(r, np.ones((len(r), 1)), dr)
    yc0 = np.array([math.sqrt(abs(x)) / (math.tan if x < 0 else math.tanh)(math.sqrt(abs(x)) * (rs - r0))
                    for x in lam])
    y0 = u @ np.diag(yc0) @ u.T
    print("MAPROP constant coupling N=3", flush=True)
    y, nodes, st = eng.prop_ld(b, g, 7, [ered], y0[None])
    assert st[0] == 0
    yc = np.array([math.sqrt(abs(x)) / (math.tan if x < 0 else math.tanh)(math.sqrt(abs(x)) * (re - r0))
                   for x in lam])
    zeros = sum(int(math.floor(math.sqrt(-x) * (re - r0) / math.pi)) -
                int(math.floor(math.sqrt(-x) * (rs - r0) / math.pi)) for x in lam if x < 0)
    yex = u @ np.diag(yc) @ u.T
    assert np.abs(y[0] - yex).max() < 1e-9 * np.abs(yex).max()
    assert int(nodes[0]) == zeros


def test_maprop_with_constant_terms_vs_oracle(eng):
    """MAPROP for a block with NCONST > 0 (the synthetic field deck, N = 48,
    a rotated Zeeman operator): the adiabatic bases include each instance's
    constants, so the instances run one at a time; three fields in one call
    against the oracle's MAPROP from the same Y, node counts exact."""
    from molscat_amd import decks
    from test_oracle_ytrans import _oracle_block, _random_zeeman

    fb = decks.field48(zeeman=_random_zeeman(5))
    b = eng.block(fb.vl, fb.eint, fb.cent, mxlam=fb.mxlam, nconst=fb.nconst)
    fields = [25.0, 300.0, 900.0]
    pc = np.stack([fb.pconst(f) for f in fields])
    ns, r0, r1 = 240, 4.0, 10.0
    dr = (r1 - r0) / ns
    r = S.ld_radii(r0, dr, ns)
    g = eng.grid(r, fb.ptab(r), dr)
    rng = np.random.default_rng(4)
    a = rng.standard_normal((len(fields), fb.n, fb.n))
    y0 = a + np.transpose(a, (0, 2, 1))
    ered = np.full(len(fields), -2.0 * fb.cm2ru)
    print(f"MAPROP with NCONST=1 N={fb.n} {ns} sectors, fields {fields}", flush=True)
    y, nodes, st = eng.prop_ld(b, g, 7, ered, y0, pconst=pc)
    assert np.all(st == 0)
    for m, f in enumerate(fields):
        ob = _oracle_block(fb, f)
        yo, no = O.maprop(y0[m].copy(), ob, ered[m], r0, ns, dr)
        d = np.abs(y[m] - yo).max() / np.abs(yo).max()
        print(f"  B={f} G nodes {int(nodes[m])} (oracle {no}) rel|dY| {d:.2e}", flush=True)
        assert int(nodes[m]) == no
        # measured up to 2.0e-11
        assert d < 2e-10
