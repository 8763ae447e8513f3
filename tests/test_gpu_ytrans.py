"""GPU parity of YTRANS (mcp_ytrans and its use inside mcp_smatrix_ld for
blocks with NCONST > 0) on the field-dependent deck (N = 48, the sets of
equal L interleaved, a non-diagonal Zeeman operator so that the asymptotic
basis is a real rotation), through the C ABI:
- thresholds and the transformed Y against the oracle's YTRANS (1e-12);
- free particles: the exact primitive-basis Y of the oracle test turned
  into S = I by mcp_ytrans + mcp_ytok_ktos;
- whole S-matrix solves (YINIT, MDPROP with the field as the instance
  constant, YTRANS, YTOK, KTOS) at three fields in one call against the
  oracle's, with the node count, for a diagonal and a rotated Zeeman
  operator."""
import numpy as np
import pytest

from molscat_amd import decks
from molscat_amd import system as S
from oracle import molscat_oracle as O

from test_oracle_ytrans import _oracle_block, _random_zeeman, free_particle_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def _block(eng, fb):
    return eng.block(fb.vl, fb.eint, fb.cent, mxlam=fb.mxlam, nconst=fb.nconst)


def test_ytrans_vs_oracle(eng):
    fb = decks.field48(zeeman=_random_zeeman(5))
    b = _block(eng, fb)
    fields = [25.0, 300.0, 1000.0]
    pc = np.stack([fb.pconst(f) for f in fields])
    rng = np.random.default_rng(3)
    y = rng.standard_normal((len(fields), fb.n, fb.n))
    y = y + np.transpose(y, (0, 2, 1))
    print(f"ytrans N={fb.n} fields={fields}", flush=True)
    eint, yt = eng.ytrans(b, pc, y)
    eint0, none = eng.ytrans(b, pc)
    assert none is None and np.array_equal(eint, eint0)
    for m, f in enumerate(fields):
        yr, er, _ = O.ytrans(y[m], _oracle_block(fb, f))
        de = np.abs(eint[m] - er).max() / np.abs(er).max()
        dy = np.abs(yt[m] - yr).max() / np.abs(yr).max()
        print(f"  B={f} G |dEINT|={de:.1e} |dY|={dy:.1e}", flush=True)
        assert de < 1e-12 and dy < 1e-12


def test_free_particles_unit_s_matrix(eng):
    fb, ob, ered, yprim, r = free_particle_case()
    b = _block(eng, fb)
    eint, yt = eng.ytrans(b, ob.pconst[None, :], yprim[None])
    wvec, nopen = S.gtwvec(ered, eint[0])
    nb = S.chnsrt(eint[0])
    print(f"free particles N={fb.n} nopen={nopen}", flush=True)
    _, sr, si, st = eng.ytok_ktos(yt, wvec, [nopen], nb, fb.cent, r)
    sr, si = sr[0], si[0]
    assert st[0] == 0
    assert np.abs(sr - np.eye(nopen)).max() < 1e-12 and np.abs(si).max() < 1e-9


@pytest.mark.parametrize("zeeman", ["diagonal", "rotated"])
def test_field_smatrix_vs_oracle(eng, zeeman):
    """A diagonal Zeeman operator leaves U(C) the same for every field (one
    shared set of sector matrices); a rotated one couples channels, so each
    field propagates with its own."""
    fb = decks.field48(zeeman=None if zeeman == "diagonal" else _random_zeeman(5))
    b = _block(eng, fb)
    rstart, rstop, dr = 3.5, 25.0, 0.02
    nstep, dr = S.drset_constant(rstart, rstop, dr)
    r = S.ld_radii(rstart, dr, nstep)
    g = eng.grid(r, fb.ptab(r), dr)
    fields = [100.0, 450.0, 900.0]
    pc = np.stack([fb.pconst(f) for f in fields])
    eint, _ = eng.ytrans(b, pc)
    # collision energy 0.05 cm-1 above each field's lowest threshold
    ered = np.array([e.min() + 0.05 * fb.cm2ru for e in eint])
    wv, nops = zip(*[S.gtwvec(ered[m], eint[m]) for m in range(len(fields))])
    nb = np.stack([S.chnsrt(eint[m]) for m in range(len(fields))])
    print(f"field S-matrices N={fb.n} {nstep} sectors nopen={nops}", flush=True)
    out = eng.smatrix_ld(b, g, 6, ered, np.array(wv), np.array(nops), nb, rstop, pconst=pc)
    for m, f in enumerate(fields):
        ref = O.smatrix_ld(_oracle_block(fb, f), ered[m], O.Grid(rstart, rstop, nstep, dr), iprop=6)
        assert out["STATUS"][m] == 0
        assert int(ref["NOPEN"]) == nops[m]
        d = max(np.abs(out["SR"][m] - ref["SR"]).max(), np.abs(out["SI"][m] - ref["SI"]).max())
        print(f"  B={f} G nopen={nops[m]} max|dS|={d:.2e} nodes={int(out['NODES'][m])} (oracle {ref['NODES']})",
              flush=True)
        assert int(out["NODES"][m]) == int(ref["NODES"])
        # 5e-10: the device sweep and the oracle's DSYTF2/DSYTRI round
        # differently over the 1075 inversions (measured up to 1.5e-10; the
        # YTRANS of a diagonal Zeeman operator is an exact permutation on both)
        assert d < 5e-10
        s = out["SR"][m] + 1j * out["SI"][m]
        assert np.abs(s @ s.conj().T - np.eye(len(s))).max() < 1e-11


def test_field_bound_match_vs_oracle(eng):
    """BOUND matching on the field deck (rotated Zeeman operator, so each
    instance propagates with its own U(C)): two fields x two energies below
    the lowest threshold in one call; NODEC exact, EIGMIN and the matching
    matrix's eigenvalues against the oracle's."""
    fb = decks.field48(zeeman=_random_zeeman(5))
    b = _block(eng, fb)
    (r0, dro, no), (r1, dri, ni) = S.bound_segments(3.5, 8.0, 40.0, 0.02)
    ro = S.ld_radii(r0, dro, no)
    ri = S.ld_radii(r1, dri, ni)
    go = eng.grid(ro, fb.ptab(ro), dro)
    gi = eng.grid(ri, fb.ptab(ri), dri)
    og = O.BoundGrid(3.5, 8.0, 40.0, no, dro, ni, dri)
    fields = [200.0, 200.0, 700.0, 700.0]
    below = [3.0, 40.0, 3.0, 40.0]
    pc = np.stack([fb.pconst(f) for f in fields])
    eint, _ = eng.ytrans(b, pc)
    ered = np.array([eint[m].min() - below[m] * fb.cm2ru for m in range(len(fields))])
    print(f"field BOUND N={fb.n} {no}+{ni} sectors", flush=True)
    out = eng.bound_match(b, go, gi, ered, pconst=pc)
    for m in range(len(fields)):
        ref = O.bound_match(_oracle_block(fb, fields[m]), ered[m], og)
        assert out["STATUS"][m] == 0
        de = abs(out["EIGMIN"][m] - ref["EIGMIN"]) / np.abs(ref["EVALS"]).max()
        dv = np.abs(out["EVALS"][m] - ref["EVALS"]).max() / np.abs(ref["EVALS"]).max()
        print(f"  B={fields[m]} G E=thr-{below[m]} cm-1 NODEC={int(out['NODEC'][m])} (oracle {ref['NODEC']}) "
              f"|dEIGMIN|={de:.1e} |dEVALS|={dv:.1e}", flush=True)
        assert int(out["NODEC"][m]) == int(ref["NODEC"])
        assert de < 1e-9 and dv < 1e-9
