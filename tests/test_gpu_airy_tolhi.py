"""GPU parity of AIRY with TOLHI > 0 (mcp_prop_airy_tolhi, AIPROP choosing
its steps, the reference's default long-range propagator) through the C
ABI:
- the reference's 'IPROPS=6, IPROPL=9' run (examples/output/molscat-all_
  propagators.output; ITYPE=1 rotor, N = 12): YINIT + MDPROP on the device
  to RMID = 1.4 RTURN, AIRY with TOLHIL = 1e-4, POWRL = 3 to RMAX = 20 for
  1000 cm-1 choosing its steps and 900 cm-1 replaying them, YTOK + KTOS:
  226 steps and both S-matrices against the printed ones; Z(RMAX) and the
  steps against the oracle's AIPROP from the same Z(RMID);
- the replay modes: the first instance's run and record are the same alone
  and in a batch, and a replay of the record is bitwise the batch's
  replaying instance;
- inward propagation from RMAX (BOUND's outer segment) against the oracle."""
import json
import math
import os

import numpy as np
import pytest

from molscat_amd import system as S
from oracle import molscat_oracle as O

pytestmark = pytest.mark.gpu

LABEL = "model system: ITYPE=1; IPROPS=6, IPROPL=9"


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def deck(golden_dir):
    with open(os.path.join(golden_dir, "ref_all_propagators_smat.json")) as f:
        sec = json.load(f)["sections"][LABEL]
    pot = O.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    blk, _ = O.make_itype1_block(10, 1, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)
    eredmx = 1000.0 * blk.cm2ru
    rstart, rturn = O.findrm(blk, 0.5, eredmx, 9)
    rmid = 1.4 * rturn
    cays = O.calck(0.0, eredmx, blk.eint)
    nstep, dr = O.drset(rstart, rmid, 12.0, 0.0, cays, O.UNSET, 6, 0.0)
    _, dr2 = O.drset(rmid, 20.0, 12.0, 1e-4, cays, O.UNSET, 9, 3.0)
    return sec, blk, O.Grid(rstart, rmid, nstep, dr), dr2


def _upload(eng, blk):
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.vl.shape[0])
    eng.set_potl(b, blk.potl, blk.rscale, blk.ep2ru)
    return b


def _z_at_rmid(eng, b, blk, ld, ered):
    r = S.ld_radii(ld.rstart, ld.dr, ld.nstep)
    ptab = np.stack([blk.ep2ru * np.clip(blk.potl(x * blk.rscale), -1e60, 1e60) for x in r])
    g = eng.grid(r, ptab, ld.dr)
    y0 = eng.yinit(b, g, ered)
    y, nodes, st = eng.prop_ld(b, g, 6, ered, y0)
    assert np.all(st == 0)
    return y


def test_airy_tolhi_fixture_smatrices(eng, deck):
    sec, blk, ld, dr2 = deck
    b = _upload(eng, blk)
    ered = np.array([1000.0, 900.0]) * blk.cm2ru
    print(f"AIRY TOLHI=1e-4 fixture N={blk.n}: MDPROP {ld.nstep} steps to {ld.rstop:.4f}", flush=True)
    z = _z_at_rmid(eng, b, blk, ld, ered)
    zout, nodes, nstep, st, rec = eng.prop_airy_tolhi(b, ered, z, ld.rstop, 20.0, dr2, 3.0, 1e-4, replay=2)
    print(f"  steps {list(nstep)} status {list(st)}", flush=True)
    assert np.all(st == 0)
    assert list(nstep) == [sec["props"][1][3]] * 2 == [226, 226]
    for m in range(2):
        wvec, nopen = S.gtwvec(ered[m], blk.eint)
        nb = S.chnsrt(blk.eint)
        _, sr, si, sst = eng.ytok_ktos(zout[m:m + 1], wvec[None], [nopen], nb, blk.cent, 20.0)
        sm = sec["smat"][m]
        d = max(np.abs(sr[0] - np.array(sm["sr"])).max(), np.abs(si[0] - np.array(sm["si"])).max())
        print(f"  E={[1000, 900][m]} cm-1 max|dS| vs printed {d:.2e}", flush=True)
        assert sst[0] == 0
        assert d < 5e-10
    # Z(RMAX) and the steps against the oracle's AIPROP from the same Z(RMID)
    zo, no, steps = O.aiprop_tolhi(z[0], blk, ered[0], ld.rstop, 20.0, dr2, 3.0, 1e-4)
    assert len(steps) == 226 and no == nodes[0]
    # 1e-8 relative: CORR's CDIAG is a second difference of the eigenvalues
    # (EIGNOW - EIGOLD + DRMID HP, times 2/DRMID^2), so the eigen-solvers'
    # rounding reaches the step sizes amplified (measured 3.3e-9 on DRNOW)
    ref_steps = np.array(steps)
    drel = (np.abs(rec - ref_steps) / np.abs(ref_steps)).max()
    print(f"  steps vs oracle rel {drel:.2e}", flush=True)
    assert drel < 1e-8
    dz = np.abs(zout[0] - zo).max() / np.abs(zo).max()
    print(f"  Z(RMAX) vs oracle rel {dz:.2e}", flush=True)
    assert dz < 1e-9


def test_airy_tolhi_replay_modes_bitwise(eng, deck):
    sec, blk, ld, dr2 = deck
    b = _upload(eng, blk)
    ered = np.array([1000.0, 900.0, 950.0]) * blk.cm2ru
    z = _z_at_rmid(eng, b, blk, ld, ered)
    print("AIRY TOLHI replay modes", flush=True)
    z2, n2, k2, s2, rec2 = eng.prop_airy_tolhi(b, ered, z, ld.rstop, 20.0, dr2, 3.0, 1e-4, replay=2)
    z0, n0, k0, s0, rec0 = eng.prop_airy_tolhi(b, ered[:1], z[:1], ld.rstop, 20.0, dr2, 3.0, 1e-4, replay=0)
    z1, n1, k1, s1, _ = eng.prop_airy_tolhi(b, ered[1:], z[1:], ld.rstop, 20.0, dr2, 3.0, 1e-4, replay=1,
                                             steps=rec2)
    assert np.all(s2 == 0) and np.all(s0 == 0) and np.all(s1 == 0)
    assert k0[0] == k2[0] == k2[1] == k2[2] == k1[0] == k1[1]
    assert np.array_equal(rec0, rec2)
    assert np.array_equal(z0[0], z2[0]) and n0[0] == n2[0]
    assert np.array_equal(z1, z2[1:]) and np.array_equal(n1, n2[1:])
    # a replay follows its record to the record's last step
    _, _, kshort, sshort, _ = eng.prop_airy_tolhi(b, ered[1:], z[1:], ld.rstop, 20.0, dr2, 3.0, 1e-4, replay=1,
                                                  steps=rec2[:10])
    assert np.all(sshort == 0) and np.all(kshort == 10)


def test_airy_tolhi_inward_vs_oracle(eng, deck):
    """Inward from RMAX, as BOUND propagates its outer segment: negative
    steps (the growth factor's (1 + DRNOW/RNEXT)^3 term) and NODES = N -
    KOUNT per step, at an energy below every threshold, against the
    oracle's AIPROP on the same start: steps, node count and Z."""
    sec, blk, ld, dr2 = deck
    b = _upload(eng, blk)
    ered = -3.0 * blk.cm2ru
    z0 = O.yinit(blk, ered, 20.0, zout=False)
    print(f"AIRY TOLHI=1e-4 inward 20 -> 1.2 RM, N={blk.n}", flush=True)
    z, nodes, nstep, st, rec = eng.prop_airy_tolhi(b, [ered], z0[None], 20.0, 1.2, -0.05, 3.0, 1e-4, replay=0)
    zo, no, steps = O.aiprop_tolhi(z0, blk, ered, 20.0, 1.2, -0.05, 3.0, 1e-4)
    print(f"  steps {nstep[0]} (oracle {len(steps)}) nodes {nodes[0]} (oracle {no})", flush=True)
    assert st[0] == 0 and nstep[0] == len(steps) and nodes[0] == no
    dz = np.abs(z[0] - zo).max() / np.abs(zo).max()
    print(f"  Z vs oracle rel {dz:.2e}", flush=True)
    assert dz < 1e-9


def test_airy_tolhi_beyond_lds_vs_oracle(eng):
    """N = 72 (JTOT = 30, even j = 0..16): the six N x N matrices no longer
    fit in LDS and live in a per-instance global workspace behind the same
    pointers.  Inward from 20 to 8 RM below every threshold against the
    oracle's AIPROP on the same start: steps, node count and Z."""
    pot = O.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    blk, _ = O.make_itype1_block(30, 1, list(range(0, 17, 2)), 30.0, [0, 2], pot, 20.0)
    assert blk.n == 72
    b = _upload(eng, blk)
    ered = -3.0 * blk.cm2ru
    z0 = O.yinit(blk, ered, 20.0, zout=False)
    print(f"AIRY TOLHI=1e-4 inward 20 -> 8 RM, N={blk.n} (matrices in global memory)", flush=True)
    z, nodes, nstep, st, rec = eng.prop_airy_tolhi(b, [ered], z0[None], 20.0, 8.0, -0.05, 3.0, 1e-4, replay=0)
    zo, no, steps = O.aiprop_tolhi(z0, blk, ered, 20.0, 8.0, -0.05, 3.0, 1e-4)
    print(f"  steps {nstep[0]} (oracle {len(steps)}) nodes {nodes[0]} (oracle {no})", flush=True)
    assert st[0] == 0 and nstep[0] == len(steps) and nodes[0] == no
    dz = np.abs(z[0] - zo).max() / np.abs(zo).max()
    print(f"  Z vs oracle rel {dz:.2e}", flush=True)
    # measured 2.3e-15
    assert dz < 1e-11


def test_airy_tolhi_scalam_vs_oracle(eng, deck):
    """SCAPOT's SCALAM (source_code/scapot-default.f:19) on the potential
    AIRY evaluates where its step control lands: the inward propagation
    above with SCALAM = 0.93 (folded into the library's EP2RU factor by
    Engine.set_potl) against the oracle's block with the same SCALAM; steps,
    node count and Z, and the unscaled oracle run differs."""
    import copy

    sec, blk0, ld, dr2 = deck
    blk = copy.copy(blk0)
    blk.scalam = 0.93
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.vl.shape[0])
    eng.set_potl(b, blk.potl, blk.rscale, blk.ep2ru, blk.scalam)
    ered = -3.0 * blk.cm2ru
    z0 = O.yinit(blk, ered, 20.0, zout=False)
    print(f"AIRY TOLHI=1e-4 inward with SCALAM=0.93, N={blk.n}", flush=True)
    z, nodes, nstep, st, rec = eng.prop_airy_tolhi(b, [ered], z0[None], 20.0, 1.2, -0.05, 3.0, 1e-4, replay=0)
    zo, no, steps = O.aiprop_tolhi(z0, blk, ered, 20.0, 1.2, -0.05, 3.0, 1e-4)
    z1, _, _ = O.aiprop_tolhi(z0, blk0, ered, 20.0, 1.2, -0.05, 3.0, 1e-4)
    print(f"  steps {nstep[0]} (oracle {len(steps)}) nodes {nodes[0]} (oracle {no})", flush=True)
    assert st[0] == 0 and nstep[0] == len(steps) and nodes[0] == no
    dz = np.abs(z[0] - zo).max() / np.abs(zo).max()
    d1 = np.abs(z1 - zo).max() / np.abs(zo).max()
    print(f"  Z vs oracle rel {dz:.2e} (unscaled differs by {d1:.2e})", flush=True)
    assert d1 > 1e-3
    assert dz < 1e-9
