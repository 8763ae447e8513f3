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
  operator;
- NRSQ > 0 (the centrifugal operator among the terms, MULTOP's resolution
  of C and L^2): the rotor deck turned by a random rotation, YTRANS against
  the oracle and the plain deck's K-matrix recovered; the whole solve
  against the oracle's with node counts exact."""
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


# ---------------------------------------------------------------------------
# NRSQ > 0: the centrifugal operator among the terms (mcp_ytrans_cent,
# k_ytrans_rsq; MULTOP's resolution of C and L^2)
# ---------------------------------------------------------------------------
def _turned(eng):
    from test_oracle_ytrans import turned_rotor_blocks

    blk, tb, q = turned_rotor_blocks()
    b = eng.block(tb.vl, tb.eint, tb.cent, mxlam=tb.vl.shape[0] - 2, nconst=1, nrsq=1)
    return blk, tb, q, b


def test_ytrans_rsq_vs_oracle(eng):
    """The turned rotor deck (N = 12): thresholds, CENT and Y in the
    asymptotic basis against the oracle's MULTOP restatement, for a Y turned
    from the plain deck's Y at 20 RM; the K-matrix from the device's YTOK
    with those CENT equals the plain deck's channel for channel."""
    from test_oracle_ytrans import _label_map

    blk, tb, q, b = _turned(eng)
    ered = 1000.0 * blk.cm2ru
    rstart, _ = O.findrm(blk, 0.5, ered, 9)
    g = O.Grid(rstart, 20.0, 2000, (20.0 - rstart) / 2000)
    a = O.smatrix_ld(blk, ered, g)
    yt = q.T @ a["Y"] @ q
    print(f"ytrans NRSQ=1 N={blk.n}", flush=True)
    eint, cent, y2 = eng.ytrans_cent(b, np.array([[1.0]]), yt[None])
    yr, er, _, cr = O.ytrans_rsq(yt, tb)
    de = np.abs(eint[0] - er).max() / np.abs(er).max()
    dc = np.abs(cent[0] - cr).max()
    dy = np.abs(y2[0] - yr).max() / np.abs(yr).max()
    print(f"  EINT {de:.2e} CENT {dc:.2e} Y {dy:.2e}", flush=True)
    # measured 3.8e-15, 4.8e-13 (of L(L+1) up to 240) and 1.2e-15
    assert de < 1e-13 and dc < 1e-10 and dy < 1e-12
    wvec, nopen = O.gtwvec(ered, eint[0])
    nb = O.chnsrt(eint[0])
    k, sr, si, st = eng.ytok_ktos(y2, wvec[None], [nopen], nb, cent[0], 20.0)
    assert st[0] == 0
    pi = _label_map(blk, eint[0], cent[0], a["NB"], nb, nopen)
    kd = k[0]
    kp = a["K"][np.ix_(pi, pi)]
    assert np.abs(np.abs(kd) - np.abs(kp)).max() < 1e-10 * max(1.0, np.abs(kp).max())
    assert np.abs(np.diag(kd) - np.diag(kp)).max() < 1e-10 * max(1.0, np.abs(kp).max())


def test_smatrix_rsq_vs_oracle(eng):
    """The whole solve on the turned deck (YINIT, MDPROP, YTRANS with C and
    L^2, YTOK with YTRANS's CENT per instance, KTOS) at two energies in one
    call against the oracle's on the same grid, node counts exact."""
    blk, tb, q, b = _turned(eng)
    energies = np.array([1000.0, 900.0]) * blk.cm2ru
    rstart, _ = O.findrm(blk, 0.5, energies[0], 9)
    ns = 3000
    dr = (20.0 - rstart) / ns
    r = S.ld_radii(rstart, dr, ns)
    ptab = np.stack([tb.ep2ru * np.clip(tb.potl(x * tb.rscale), -1e60, 1e60) for x in r])
    gr = eng.grid(r, ptab, dr)
    eint, _, _ = eng.ytrans_cent(b, np.array([[1.0]]))
    print(f"smatrix NRSQ=1 N={blk.n} {ns} sectors", flush=True)
    for m, e in enumerate(energies):
        wvec, nopen = O.gtwvec(e, eint[0])
        nb = O.chnsrt(eint[0])
        out = eng.smatrix_ld(b, gr, 6, [e], wvec[None], [nopen], nb, 20.0, pconst=np.array([[1.0]]))
        ref = O.smatrix_ld(tb, e, O.Grid(rstart, 20.0, ns, dr))
        assert out["STATUS"][0] == 0
        nodes = out["NODES"]
        d = max(np.abs(out["SR"][0] - ref["SR"]).max(), np.abs(out["SI"][0] - ref["SI"]).max())
        print(f"  E={e / blk.cm2ru:.0f} cm-1 nodes {int(nodes[0])} (oracle {ref['NODES']}) max|dS| {d:.2e}", flush=True)
        assert int(nodes[0]) == ref["NODES"]
        # measured 1.2e-11 and 5.0e-12
        assert d < 1e-10
