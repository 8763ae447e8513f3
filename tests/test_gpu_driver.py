"""GPU run of the MOLSCAT scattering loop (molscat_amd.driver.cross_sections)
through the C ABI:
- the reference's molscat-basic1 example (ITYPE=1 rotor, JTOT = 10, 20
  with JSTEP = 10, both parities, MDPROP DR = 0.001 from FINDRM's RSTART to
  20 RM, E = 1250 cm-1): the printed integral cross sections, which carry
  6-7 significant digits;
- the energies of a block batched into one call: 1250 cm-1 alone and with
  1100 and 800 cm-1 beside it (same grid, set at the largest energy) give
  the same cross sections, and the batched ones obey detailed balance."""
import json
import os
import time

import numpy as np
import pytest

from molscat_amd import driver as D
from molscat_amd import system as S

pytestmark = pytest.mark.gpu

POT = S.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
JLEV = [0, 2, 4, 6]


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def printed(golden_dir):
    with open(os.path.join(golden_dir, "ref_basic1_sigma.json")) as f:
        d = json.load(f)
    sig = np.zeros((3, 3))
    for e in d["sigma"]:
        sig[e["f"] - 1, e["i"] - 1] = e["sigma"]
    return sig


def _run(eng, energies):
    return D.cross_sections(eng, JLEV, 30.0, [0, 2], POT, 20.0, energies, 10, 20, 10, 0.5, 20.0, 0.001)


def test_basic1_cross_sections(eng, printed):
    print("driver: basic1 JTOT=10,20 both parities, E=1250 cm-1", flush=True)
    t0 = time.perf_counter()
    cs = _run(eng, [1250.0])
    dt = time.perf_counter() - t0
    for br in cs.blocks:
        print(f"  JTOT={br.jtot} IB={br.ib} N={br.n} RSTART={br.rstart:.4f} NSTEP={br.nstep} "
              f"NOPEN={br.nopen} status={br.status}", flush=True)
        assert br.status == [0]
    assert [br.nstep for br in cs.blocks] == [19364, 19367, 19363, 19367]
    rel = np.abs(cs.sigma[0] - printed) / np.abs(printed)
    print(f"  max rel dSIG vs printed {rel.max():.2e} ({dt:.2f} s for 4 blocks)", flush=True)
    # 5e-6: the printed values carry 6-7 significant digits
    assert rel.max() < 5e-6


def test_batched_energies(eng):
    one = _run(eng, [1250.0])
    many = _run(eng, [1250.0, 1100.0, 800.0])
    assert np.abs(many.sigma[0] - one.sigma[0]).max() <= 1e-13 * np.abs(one.sigma[0]).max()
    elev = [30.0 * j * (j + 1) for j in (0, 2, 4)]
    g = [1.0, 5.0, 9.0]
    for m, e in enumerate(many.energies):
        sig = many.sigma[m]
        for i in range(3):
            for f in range(3):
                if e <= elev[i] or e <= elev[f]:
                    continue
                lhs = g[i] * (e - elev[i]) * sig[f, i]
                rhs = g[f] * (e - elev[f]) * sig[i, f]
                assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), abs(rhs))


@pytest.mark.parametrize("iprops,ipropl,label,tol", [
    (6, 9, "model system: ITYPE=1; IPROPS=6, IPROPL=9", 2e-11),  # measured 6.4e-12
    (5, 4, "model system: ITYPE=1; IPROPS=5, IPROPL=4", 1e-11),  # measured 1.7e-12
])
def test_two_segment_runs_match_reference(eng, golden_dir, iprops, ipropl, label, tol):
    """The all-propagators deck's two-segment runs through the driver: the
    LD segment with STEPS = 12 to RMID = 1.4 x FINDRM's turning point, then
    AIRY with the default TOLHIL = 1e-4 (POWRL = 3) or VIVS to 20 RM, at 1000
    and 900 cm-1 (the second energy replaying the first's long-range steps);
    the printed S-matrices."""
    with open(os.path.join(golden_dir, "ref_all_propagators_smat.json")) as f:
        sec = json.load(f)["sections"][label]
    print(f"driver IPROPS={iprops} IPROPL={ipropl}", flush=True)
    cs = D.cross_sections(eng, JLEV, 30.0, [0, 2], POT, 20.0, [1000.0, 900.0], 10, 10, 1, 0.5, 20.0,
                          iprop=iprops, ibs=(1,), steps=12.0, iprop_long=ipropl, rvfac=1.4)
    br = cs.blocks[0]
    assert br.status == [0, 0] and br.nstep == sec["props"][0][3]
    for m in range(2):
        sm = sec["smat"][m]
        d = max(np.abs(br.sr[m] - np.array(sm["sr"])).max(), np.abs(br.si[m] - np.array(sm["si"])).max())
        print(f"  E={[1000, 900][m]} cm-1 max|dS| vs printed {d:.2e}", flush=True)
        assert d < tol


def test_saved_smatrices_and_restart(eng, tmp_path):
    """cross_sections writes every S-matrix to the saved S-matrix file as
    SPROC does; a run stopped inside its third symmetry block and restarted
    with IRSTRT = 2 recomputes only the blocks after the last complete one,
    recreates the earlier cross sections from the file and leaves the same
    file behind."""
    from molscat_amd import smatfile as F

    energies = [1250.0, 1100.0, 800.0]
    args = (eng, JLEV, 30.0, [0, 2], POT, 20.0, energies, 4, 6, 1, 0.5, 20.0, 0.004)
    full = str(tmp_path / "full.sav")
    print("driver: JTOT 4..6 both parities, 3 energies, saved S-matrices", flush=True)
    a = D.cross_sections(*args, isaveu=full, label="basic1 deck, JTOT 4-6")
    h, ents = F.read(full)
    assert h.label == "basic1 deck, JTOT 4-6" and np.array_equal(h.energies, energies)
    want = [(br, m) for br in a.blocks for m in range(3) if br.nopen[m] > 0]
    assert len(ents) == len(want) == 18
    low = lambda s: np.tril(s) + np.tril(s, -1).T   # the lower triangle is what SPROC writes
    for e, (br, m) in zip(ents, want):
        assert (e.jtot, e.ib, e.inrg, e.nopen) == (br.jtot, br.ib, m + 1, br.nopen[m])
        assert np.array_equal(e.sr, low(br.sr[m])) and np.array_equal(e.si, low(br.si[m]))
    # stopped after the first energy of the third block (JTOT 5, IB 1)
    _, _, ends = F.scan(full)
    cut = str(tmp_path / "cut.sav")
    with open(full, "rb") as f:
        data = f.read()
    with open(cut, "wb") as f:
        f.write(data[:ends[7] + 20])
    print("driver: restart after the second block", flush=True)
    b = D.cross_sections(*args, isaveu=cut, irstrt=2, label="basic1 deck, JTOT 4-6")
    assert [(br.jtot, br.ib) for br in b.blocks] == [(5, 1), (5, 2), (6, 1), (6, 2)]
    assert np.abs(b.sigma - a.sigma).max() <= 1e-13 * np.abs(a.sigma).max()
    _, ents2 = F.read(cut)
    assert len(ents2) == len(ents)
    for e, f2 in zip(ents, ents2):
        assert (e.jtot, e.ib, e.inrg) == (f2.jtot, f2.ib, f2.inrg)
        assert np.abs(e.sr - f2.sr).max() <= 1e-13 and np.abs(e.si - f2.si).max() <= 1e-13
