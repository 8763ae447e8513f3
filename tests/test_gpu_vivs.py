"""GPU parity of the VIVS propagator (mcp_prop_vivs, VVPROP) through the
C ABI:
- the reference's 'IPROPS=5, IPROPL=4' run (examples/output/molscat-all_
  propagators.output; ITYPE=1 rotor, N = 12): YINIT + LDPROP on the device
  to RMID, R = Y^-1, VIVS to RMAX = 20 for 1000 cm-1 choosing its steps and
  900 cm-1 replaying them, Y = R^-1, YTOK + KTOS: 1830 steps and both
  S-matrices against the printed ones; R(RMAX) against the oracle's VVPROP
  from the same R(RMID);
- a constant coupled W against the exact solution;
- the replay modes: the first instance's run and the record it writes are
  the same alone and in a batch, and a replay of the record is bitwise the
  batch's replaying instance;
- N = 56 (work matrices beyond LDS) against the oracle;
- the /LDVVCM/ options (derivative perturbations, ISHIFT, IDIAG, IALPHA,
  IALFP, ISYM, IPERT) against the oracle with the same options, and the
  replay of interval ends decided by the corrections' estimates."""
import json
import math
import os

import numpy as np
import pytest

from molscat_amd import system as S
from oracle import molscat_oracle as O

from test_oracle_vivs import _constant_block

pytestmark = pytest.mark.gpu

LABEL = "model system: ITYPE=1; IPROPS=5, IPROPL=4"


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def _upload(eng, blk):
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.vl.shape[0])
    eng.set_potl(b, blk.potl, blk.rscale, blk.ep2ru)
    return b


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
    nstep, dr = O.drset(rstart, rmid, 12.0, 0.0, cays, O.UNSET, 5, 0.0)
    return sec, blk, O.Grid(rstart, rmid, nstep, dr)


def test_vivs_fixture_smatrices(eng, deck):
    sec, blk, ld = deck
    b = _upload(eng, blk)
    r = S.ld_radii(ld.rstart, ld.dr, ld.nstep)
    ptab = np.stack([blk.ep2ru * np.clip(blk.potl(x * blk.rscale), -1e60, 1e60) for x in r])
    g = eng.grid(r, ptab, ld.dr)
    ered = np.array([1000.0, 900.0]) * blk.cm2ru
    print(f"VIVS fixture N={blk.n}: LDPROP {ld.nstep} steps to {ld.rstop:.4f}", flush=True)
    y0 = eng.yinit(b, g, ered)
    y, _, st = eng.prop_ld(b, g, 5, ered, y0)
    assert np.all(st == 0)
    rmat, _ = eng.syminv(y)
    tldiag, toff = O.vivs_tolerances(1e-4)
    print("VIVS 1.065 -> 20: 1000 cm-1 own steps, 900 cm-1 replayed", flush=True)
    rout, nstep, vst, rec = eng.prop_vivs(b, ered, rmat, ld.rstop, 20.0, ld.dr, tldiag, toff, replay=2)
    assert np.all(vst == 0)
    assert list(nstep) == [sec["props"][1][3]] * 2 == [1830, 1830]
    yv, _ = eng.syminv(rout)
    for m in range(2):
        wvec, nopen = S.gtwvec(ered[m], blk.eint)
        nb = S.chnsrt(blk.eint)
        _, sr, si, sst = eng.ytok_ktos(yv[m:m + 1], wvec[None], [nopen], nb, blk.cent, 20.0)
        sm = sec["smat"][m]
        d = max(np.abs(sr[0] - np.array(sm["sr"])).max(), np.abs(si[0] - np.array(sm["si"])).max())
        print(f"  E={[1000, 900][m]} cm-1 nopen={nopen} max|dS| vs printed {d:.2e}", flush=True)
        assert sst[0] == 0 and nopen == sm["nopen"]
        assert d < 1e-11  # measured 9.9e-13 and 1.6e-12
    # R(RMAX) against the oracle's VVPROP from the same R(RMID)
    ro, steps = O.vvprop(rmat[0], blk, ered[0], ld.rstop, 20.0, ld.dr, tldiag, toff)
    ro2, _ = O.vvprop(rmat[1], blk, ered[1], ld.rstop, 20.0, ld.dr, tldiag, toff, replay=steps)
    assert len(steps) == 1830
    got = np.array([[s.rnow, s.drnow] for s in steps])
    assert np.abs(rec[:, :2] - got).max() < 1e-10
    for m, ref in enumerate((ro, ro2)):
        dr_ = np.abs(rout[m] - ref).max() / np.abs(ref).max()
        print(f"  R(RMAX) vs oracle rel {dr_:.2e}", flush=True)
        assert dr_ < 1e-9


def test_vivs_constant_coupling_exact(eng):
    rng = np.random.default_rng(11)
    n = 3
    m = rng.standard_normal((n, n))
    blk = _constant_block(0.5 * (m + m.T), np.array([0.0, 2.0, 9.0]))
    b = _upload(eng, blk)
    ered = 4.0
    lam, u = np.linalg.eigh(blk.hammat(1.0, ered)[0])
    r0, r1 = 1.0, 7.3
    y0c = np.array([0.7, -1.3, 2.1])
    rmat0 = u @ np.diag(1.0 / y0c) @ u.T
    print("VIVS constant coupling N=3", flush=True)
    rout, nstep, st, _ = eng.prop_vivs(b, [ered], rmat0[None], r0, r1, 0.05, *O.vivs_tolerances(1e-6), replay=0)
    assert st[0] == 0
    d = r1 - r0
    ye = np.empty(n)
    for c in range(n):
        k = math.sqrt(abs(lam[c]))
        if lam[c] < 0.0:
            ye[c] = (-k * math.sin(k * d) + y0c[c] * math.cos(k * d)) / (math.cos(k * d) + y0c[c] * math.sin(k * d) / k)
        else:
            ye[c] = (k * math.sinh(k * d) + y0c[c] * math.cosh(k * d)) / (math.cosh(k * d) + y0c[c] * math.sinh(k * d) / k)
    rex = u @ np.diag(1.0 / ye) @ u.T
    assert np.abs(rout[0] - rex).max() < 1e-10 * np.abs(rex).max()


def test_vivs_replay_modes_bitwise(eng, deck):
    sec, blk, ld = deck
    b = _upload(eng, blk)
    rng = np.random.default_rng(2)
    a = rng.standard_normal((3, blk.n, blk.n))
    rmat = 0.05 * (a + np.transpose(a, (0, 2, 1)))
    ered = np.array([1000.0, 900.0, 950.0]) * blk.cm2ru
    tl, to = O.vivs_tolerances(1e-4)
    print("VIVS replay modes", flush=True)
    r2, n2, s2, rec2 = eng.prop_vivs(b, ered, rmat, 1.1, 12.0, 0.003, tl, to, replay=2)
    r0, n0, s0, rec0 = eng.prop_vivs(b, ered[:1], rmat[:1], 1.1, 12.0, 0.003, tl, to, replay=0)
    r1, n1, s1, _ = eng.prop_vivs(b, ered[1:], rmat[1:], 1.1, 12.0, 0.003, tl, to, replay=1, steps=rec2)
    assert np.all(s2 == 0) and np.all(s0 == 0) and np.all(s1 == 0)
    assert n0[0] == n2[0] == n2[1] == n2[2] == n1[0] == n1[1]
    assert np.array_equal(rec0, rec2)
    assert np.array_equal(r0[0], r2[0])
    assert np.array_equal(r1, r2[1:])
    # the instances' own step control differs from the replay
    ro, no, so, _ = eng.prop_vivs(b, ered[1:2], rmat[1:2], 1.1, 12.0, 0.003, tl, to, replay=0)
    assert so[0] == 0 and no[0] != n2[1]


def test_vivs_beyond_lds_vs_oracle(eng):
    """N = 56 (JTOT = 20, even j = 0..14, lambda = 0, 2, 4): the work
    matrices in the global workspace; 0.7 -> 3 RM, 966 steps."""
    pot = O.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)], [(0.1, -6)]], rm=3.5, epsil=50.0)
    blk, _ = O.make_itype1_block(20, 1, list(range(0, 15, 2)), 30.0, [0, 2, 4], pot, 20.0)
    assert blk.n == 56
    b = _upload(eng, blk)
    ered = 1500.0 * blk.cm2ru
    y0 = O.yinit(blk, ered, 0.7)
    rmat0, _ = O.syminv(y0)
    tl, to = O.vivs_tolerances(1e-4)
    print(f"VIVS N={blk.n} 0.7 -> 3 RM", flush=True)
    rout, nstep, st, _ = eng.prop_vivs(b, [ered], rmat0[None], 0.7, 3.0, 0.002, tl, to, replay=0)
    ro, steps = O.vvprop(rmat0, blk, ered, 0.7, 3.0, 0.002, tl, to)
    assert st[0] == 0 and nstep[0] == len(steps)
    d = np.abs(rout[0] - ro).max() / np.abs(ro).max()
    print(f"  {nstep[0]} steps, rel |dR| {d:.2e}", flush=True)
    # 5e-9: VIVS's R depends on the interval bases at its truncation level,
    # and near-degenerate adiabats leave those bases ill-determined; the
    # oracle alone moves by 4.6e-10 when its DIAGVC switches from LAPACK's
    # dsyevd to dsyevr (same steps); the device's Jacobi bases measured 1.7e-9
    assert d < 5e-9


def test_vivs_replay_after_failed_recording(eng, deck):
    """replay = 2 when the recording instance stops early (here: more steps
    than the record holds, status 2): the followers must not follow a record
    that was never finished, nor one an earlier call left in the buffer.
    They stop at once with status 4 (the record ends before its last step)
    with their step counter still at the first step (following the
    recording's 10 steps would leave it at 11)."""
    sec, blk, ld = deck
    b = _upload(eng, blk)
    rng = np.random.default_rng(5)
    a = rng.standard_normal((3, blk.n, blk.n))
    rmat = 0.05 * (a + np.transpose(a, (0, 2, 1)))
    ered = np.array([1000.0, 900.0, 950.0]) * blk.cm2ru
    tl, to = O.vivs_tolerances(1e-4)
    print("VIVS replay after a failed recording", flush=True)
    # a complete run first, so the device's step buffer holds a valid record
    _, n_ok, s_ok, _ = eng.prop_vivs(b, ered, rmat, 1.1, 12.0, 0.003, tl, to, replay=2)
    assert np.all(s_ok == 0) and n_ok[0] > 10
    _, n_bad, s_bad, _ = eng.prop_vivs(b, ered, rmat, 1.1, 12.0, 0.003, tl, to, replay=2, max_steps=10)
    print(f"  statuses {list(s_bad)} steps {list(n_bad)}", flush=True)
    assert s_bad[0] == 2
    assert list(s_bad[1:]) == [4, 4]
    assert list(n_bad[1:]) == [1, 1]


VIVS_OPTION_SETS = {
    "ivp": dict(ivp=True),
    "ivpp": dict(ivpp=True),
    "ishift": dict(ishift=True),
    "idiag": dict(idiag=True, iv=True, ivp=True, ivpp=True, ishift=True, ipert=True),
    "idiag-only": dict(idiag=True),
    "ialpha0+idiag": dict(idiag=True, iv=True, ivp=True, ivpp=True, ishift=True, ipert=True, ialpha=0),
    "ialfp+idiag": dict(idiag=True, iv=True, ivp=True, ivpp=True, ishift=True, ipert=True, ialfp=True),
    "ialpha3": dict(ialpha=3, alpha1=1.1, alpha2=1.3),
    "nosym": dict(isym=False),
    "nopert": dict(ipert=False),
}


@pytest.mark.parametrize("name", list(VIVS_OPTION_SETS))
def test_vivs_options_vs_oracle(eng, deck, name):
    """The /LDVVCM/ options (mcp_set_vivs_options) against the oracle's
    VVPROP with the same options, each instance choosing its own steps: the
    rotor deck (N = 12) from 1.1 to 12 RM at 1000 cm-1, from a seeded R.
    The device takes the step integrals of the derivative terms from its own
    closed forms (mcp_stepint.hpp), the oracle from PERT2's."""
    sec, blk, ld = deck
    b = _upload(eng, blk)
    kw = VIVS_OPTION_SETS[name]
    rng = np.random.default_rng(7)
    a = rng.standard_normal((blk.n, blk.n))
    rmat = 0.05 * (a + a.T)
    ered = 1000.0 * blk.cm2ru
    tl, to = O.vivs_tolerances(1e-4)
    print(f"VIVS options {name}", flush=True)
    eng.set_vivs_options(**kw)
    try:
        rout, nstep, st, rec = eng.prop_vivs(b, [ered], rmat[None], 1.1, 12.0, 0.003, tl, to, replay=0)
    finally:
        eng.set_vivs_options()
    ro, steps = O.vvprop(rmat, blk, ered, 1.1, 12.0, 0.003, tl, to, opts=O.VivsOptions(**kw))
    d = np.abs(rout[0] - ro).max() / np.abs(ro).max()
    print(f"  {nstep[0]} steps (oracle {len(steps)}), rel |dR| {d:.2e}", flush=True)
    assert st[0] == 0 and nstep[0] == len(steps)
    # the record's interval ends are the oracle's
    assert [bool(int(f) & 4) for f in rec[:, 3]] == [s.endint for s in steps]
    # measured 5.9e-13 to 4.3e-12 over the option sets
    assert d < 2e-11


def test_vivs_options_replay(eng, deck):
    """IALPHA = 0 with IDIAG: the interval ends come from the corrections'
    estimates, so a replaying instance must take them from the record (bit
    2): the batch's second energy equals the oracle replaying the first
    energy's steps, and a replay=1 run of the record is bitwise the same."""
    sec, blk, ld = deck
    b = _upload(eng, blk)
    kw = VIVS_OPTION_SETS["ialpha0+idiag"]
    rng = np.random.default_rng(8)
    a = rng.standard_normal((2, blk.n, blk.n))
    rmat = 0.05 * (a + np.transpose(a, (0, 2, 1)))
    ered = np.array([1000.0, 900.0]) * blk.cm2ru
    tl, to = O.vivs_tolerances(1e-4)
    print("VIVS options replay", flush=True)
    eng.set_vivs_options(**kw)
    try:
        r2, n2, s2, rec = eng.prop_vivs(b, ered, rmat, 1.1, 12.0, 0.003, tl, to, replay=2)
        r1, n1, s1, _ = eng.prop_vivs(b, ered[1:], rmat[1:], 1.1, 12.0, 0.003, tl, to, replay=1, steps=rec)
    finally:
        eng.set_vivs_options()
    assert np.all(s2 == 0) and s1[0] == 0 and n1[0] == n2[1] == n2[0]
    assert np.array_equal(r1[0], r2[1])
    opts = O.VivsOptions(**kw)
    _, steps = O.vvprop(rmat[0], blk, ered[0], 1.1, 12.0, 0.003, tl, to, opts=opts)
    ro, _ = O.vvprop(rmat[1], blk, ered[1], 1.1, 12.0, 0.003, tl, to, replay=steps, opts=opts)
    d = np.abs(r2[1] - ro).max() / np.abs(ro).max()
    print(f"  {n2[1]} steps, rel |dR| {d:.2e}", flush=True)
    # measured 4.5e-11
    assert d < 2e-10


def test_vivs_scalam_vs_oracle(eng, deck):
    """SCAPOT's SCALAM (source_code/scapot-default.f:19) on the potential
    VIVS evaluates as it goes, and on DERMAT's derivatives with IVP and
    IVPP: the device (SCALAM folded into the library's EP2RU factor by
    Engine.set_potl) against the oracle's block with SCALAM = 0.93, whose
    HAMMAT and DERMAT scale the clamped radial terms; the unscaled run
    differs, so the factor is not lost on either side."""
    import copy

    sec, blk0, ld = deck
    blk = copy.copy(blk0)
    blk.scalam = 0.93
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.vl.shape[0])
    eng.set_potl(b, blk.potl, blk.rscale, blk.ep2ru, blk.scalam)
    rng = np.random.default_rng(11)
    a = rng.standard_normal((blk.n, blk.n))
    rmat = 0.05 * (a + a.T)
    ered = 1000.0 * blk.cm2ru
    tl, to = O.vivs_tolerances(1e-4)
    kw = dict(ivp=True, ivpp=True)
    print("VIVS SCALAM=0.93 with IVP, IVPP", flush=True)
    eng.set_vivs_options(**kw)
    try:
        rout, nstep, st, _ = eng.prop_vivs(b, [ered], rmat[None], 1.1, 12.0, 0.003, tl, to, replay=0)
    finally:
        eng.set_vivs_options()
    ro, steps = O.vvprop(rmat, blk, ered, 1.1, 12.0, 0.003, tl, to, opts=O.VivsOptions(**kw))
    r1, _ = O.vvprop(rmat, blk0, ered, 1.1, 12.0, 0.003, tl, to, opts=O.VivsOptions(**kw))
    d = np.abs(rout[0] - ro).max() / np.abs(ro).max()
    d1 = np.abs(r1 - ro).max() / np.abs(ro).max()
    print(f"  {nstep[0]} steps (oracle {len(steps)}), rel |dR| {d:.2e} (unscaled differs by {d1:.2e})", flush=True)
    assert st[0] == 0 and nstep[0] == len(steps)
    assert d1 > 1e-3
    assert d < 2e-11
