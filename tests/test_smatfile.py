"""The saved S-matrix file (molscat_amd/smatfile.py; SPROC's ISAVEU
records, source_code/sproc.f:254-281,797-807, read as source_code/sread.f
does): round trips, the Fortran runtime reading what the writer wrote and
the reader reading what the Fortran runtime wrote (tests/native/
smatfile_check.f90, built with amdflang), the K form with EPSUM's
eigenphase sum, and RESTRT's choice of where a stopped run resumes."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from molscat_amd import smatfile as F

HERE = os.path.dirname(os.path.abspath(__file__))


def _sym(rng, n):
    a = rng.standard_normal((n, n))
    return 0.5 * (a + a.T)


def _header(**kw):
    base = dict(label="ITYPE=1 rotor test", itype=1, jstate=np.array([[0, 1], [2, 2], [4, 3]]), ured=20.0,
                elevel=np.array([0.0, 180.0, 600.0]), energies=np.array([1000.0, 900.0]))
    base.update(kw)
    return F.Header(**base)


def _entries(rng, hdr, blocks):
    out = []
    for jtot, ib in blocks:
        for inrg in range(1, len(hdr.energies) + 1):
            nopen = 1 + (jtot + ib + inrg) % 3
            out.append(dict(jtot=jtot, inrg=inrg, ib=ib, energy=float(hdr.energies[inrg - 1]),
                            level=rng.integers(1, 4, nopen), lval=rng.integers(0, 20, nopen),
                            wvec=rng.uniform(1.0, 9.0, nopen), sr=_sym(rng, nopen), si=_sym(rng, nopen),
                            efv=rng.standard_normal(hdr.nefvp + 1), eref=float(rng.standard_normal()),
                            iexch=int(rng.integers(0, 3)), wght=float(rng.uniform(0.5, 2.0))))
    return out


def _write(path, hdr, ents, kmat=False):
    with F.SMatrixWriter(path, hdr, kmat=kmat) as w:
        for e in ents:
            w.write(e["jtot"], e["inrg"], e["ib"], e["energy"], e["level"], e["lval"], e["wvec"], e["sr"], e["si"],
                    k=e.get("k"), efv=e["efv"], eref=e["eref"], iexch=e["iexch"], wght=e["wght"])


def _same_entry(got, e):
    assert (got.jtot, got.inrg, got.ib, got.jfield, got.iexch) == (e["jtot"], e["inrg"], e["ib"], 1, e["iexch"])
    assert got.energy == e["energy"] and got.eref == e["eref"] and got.wght == e["wght"]
    assert np.array_equal(got.efv, e["efv"])
    assert np.array_equal(got.level, e["level"]) and np.array_equal(got.lval, e["lval"])
    assert np.array_equal(got.wvec, e["wvec"])


def test_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    hdr = _header(efv_names=[("MAGNETIC Z FIELD", "GAUSS"), ("ELECTRIC FIELD", "KV/CM")], isvefv=1, nfield=3,
                  nconst=1, ndgvl=2)
    ents = _entries(rng, hdr, [(0, 1), (0, 2), (1, 1)])
    path = str(tmp_path / "s.sav")
    _write(path, hdr, ents)
    h, got = F.read(path)
    assert (h.label, h.itype, h.ured, h.iprogm, h.nfield, h.nconst, h.ndgvl) == (hdr.label, 1, 20.0, 19, 3, 1, 2)
    assert np.array_equal(h.jstate, hdr.jstate) and np.array_equal(h.elevel, hdr.elevel)
    assert np.array_equal(h.energies, hdr.energies) and h.efv_names == hdr.efv_names and h.nefvp == 2
    assert len(got) == len(ents)
    for g, e in zip(got, ents):
        _same_entry(g, e)
        # SREAD mirrors the lower triangle written row by row
        assert np.array_equal(g.sr, np.tril(e["sr"]) + np.tril(e["sr"], -1).T)
        assert np.array_equal(g.si, np.tril(e["si"]) + np.tril(e["si"], -1).T)
    # record sizes: header record 80 + 3*4 + 8 + 4 bytes between 4-byte markers
    with open(path, "rb") as f:
        assert int.from_bytes(f.read(4), "little") == 104


def test_ibound_writes_cent_as_double(tmp_path):
    rng = np.random.default_rng(4)
    hdr = _header(ibound=1)
    ents = _entries(rng, hdr, [(2, 1)])
    for e in ents:
        e["lval"] = rng.uniform(0.0, 30.0, len(e["level"]))
    path = str(tmp_path / "c.sav")
    _write(path, hdr, ents)
    _, got = F.read(path)
    for g, e in zip(got, ents):
        assert g.lval.dtype == np.float64 and np.array_equal(g.lval, e["lval"])


def test_k_form_and_eigenphase_sum(tmp_path):
    rng = np.random.default_rng(5)
    hdr = _header()
    ents = _entries(rng, hdr, [(0, 1), (3, 2)])
    for e in ents:
        e["k"] = _sym(rng, len(e["level"]))
    path = str(tmp_path / "k.sav")
    _write(path, hdr, ents, kmat=True)
    _, got = F.read(path, kmat=True)
    last = 10.0
    for g, e in zip(got, ents):
        _same_entry(g, e)
        assert np.array_equal(g.sdiag, np.diag(e["sr"]) + 1j * np.diag(e["si"]))
        assert np.array_equal(g.k, np.tril(e["k"]) + np.tril(e["k"], -1).T)
        # EPSUM: sum of atan(eigenvalues of K) / pi, within 1/2 of the previous value
        raw = float(np.sum(np.arctan(np.linalg.eigvalsh(e["k"])))) / math.pi
        assert abs(g.esum - last) <= 0.5 + 1e-12
        assert abs((g.esum - raw) - round(g.esum - raw)) < 1e-12
        last = g.esum


def test_eigenphase_sum_single_channel():
    assert F.eigenphase_sum(np.array([[1.0]]), previous=0.0) == pytest.approx(0.25)
    # moved by a whole unit towards the previous value
    assert F.eigenphase_sum(np.array([[1.0]]), previous=3.0) == pytest.approx(3.25)


@pytest.fixture(scope="module")
def fortran_check(tmp_path_factory):
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if fc is None:
        pytest.skip("amdflang not available")
    d = tmp_path_factory.mktemp("fortran")
    exe = str(d / "smatfile_check")
    subprocess.check_call([fc, "-O1", os.path.join(HERE, "native", "smatfile_check.f90"), "-o", exe], cwd=str(d))
    return exe


def _parse(text):
    out = {"JSTATE": [], "ELEVEL": [], "ENERGY": [], "EFVNAME": [], "ENTRY": []}
    cur = None
    for ln in text.splitlines():
        key, _, rest = ln.partition(" ")
        if key == "ENTRY":
            cur = {"head": [int(x) for x in rest.split()], "efv": [], "chan": [], "s": []}
            out["ENTRY"].append(cur)
        elif key == "REAL":
            cur["real"] = [float(x) for x in rest.split()]
        elif key == "EFVVAL":
            cur["efv"].append(float(rest))
        elif key == "CHAN":
            a, b, c = rest.split()
            cur["chan"].append((int(a), int(b), float(c)))
        elif key == "S":
            a, b, c, d = rest.split()
            cur["s"].append((int(a), int(b), float(c), float(d)))
        elif key in ("ELEVEL", "ENERGY"):
            out[key].append(float(rest))
        elif key == "EFVNAME":
            n, u = rest.split("|")
            out[key].append((n.rstrip(), u.rstrip()))
        elif key == "JSTATE":
            out[key] = [int(x) for x in rest.split()]
        else:
            out[key] = rest
    return out


def test_fortran_reads_the_writer(tmp_path, fortran_check):
    rng = np.random.default_rng(6)
    hdr = _header(efv_names=[("MAGNETIC Z FIELD", "GAUSS")], isvefv=2)
    ents = _entries(rng, hdr, [(4, 1), (4, 2)])
    path = str(tmp_path / "py.sav")
    _write(path, hdr, ents)
    txt = subprocess.run([fortran_check, "read", path], check=True, capture_output=True, text=True).stdout
    p = _parse(txt)
    assert p["LABEL"] == hdr.label
    assert [int(x) for x in p["HEAD"].split()] == [1, 3, 2, 19]
    assert float(p["URED"]) == 20.0
    assert p["JSTATE"] == [0, 2, 4, 1, 2, 3]            # column-major JSTATE(NSTATE, NQN)
    assert p["ELEVEL"] == list(hdr.elevel) and p["ENERGY"] == list(hdr.energies)
    assert p["EFVNAME"] == hdr.efv_names
    assert len(p["ENTRY"]) == len(ents)
    for fe, e in zip(p["ENTRY"], ents):
        n = len(e["level"])
        assert fe["head"] == [e["jtot"], e["inrg"], e["ib"], 1, e["iexch"], n]
        assert fe["real"] == [e["energy"], e["eref"], e["wght"]]
        assert fe["efv"] == list(e["efv"])             # EFV(0:NEFVP), NEFVP = max(1, 2)
        assert fe["chan"] == [(int(e["level"][i]), int(e["lval"][i]), float(e["wvec"][i])) for i in range(n)]
        assert fe["s"] == [(i + 1, j + 1, float(e["sr"][i, j]), float(e["si"][i, j]))
                           for i in range(n) for j in range(i + 1)]


def test_reader_reads_the_fortran_runtime(tmp_path, fortran_check):
    path = str(tmp_path / "f.sav")
    subprocess.run([fortran_check, "write", path], check=True)
    h, got = F.read(path)
    assert h.label == "written by the Fortran runtime" and h.iprogm == 19 and h.ured == 20.0
    assert np.array_equal(h.jstate, [[0, 1], [2, 2], [4, 3]])
    assert h.efv_names == [("MAGNETIC Z FIELD", "GAUSS"), ("ELECTRIC FIELD", "KV/CM")] and h.isvefv == 1
    assert np.array_equal(h.energies, [1000.0, 900.0])
    assert len(got) == 2
    for k, g in enumerate(got, start=1):
        n = k + 1
        assert (g.jtot, g.inrg, g.ib, g.jfield, g.iexch, g.nopen) == (10, k, 1, 1, 0, n)
        assert g.energy == [1000.0, 900.0][k - 1] and g.eref == -3.25 and g.wght == 1.0
        assert np.array_equal(g.efv, [0.0, 250.0 * k, 1.5])
        assert np.array_equal(g.level, np.arange(1, n + 1)) and np.array_equal(g.lval, 2 * np.arange(1, n + 1) - 1)
        assert np.array_equal(g.wvec, 0.5 * np.arange(1, n + 1) + 0.25 * k)
        i, j = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), indexing="ij")
        s = 0.125 * k + 0.01 * (i * j) + 1e-3 * (i + j)
        assert np.array_equal(g.sr, s) and np.array_equal(g.si, -s)


def test_restart_point(tmp_path):
    """A run stopped inside an entry: RESTRT keeps up to the last block
    whose last energy was written (irstrt = 2) or the last JTOT whose last
    block was (irstrt = 1); the writer continues there and the file reads
    as if the run had not stopped."""
    rng = np.random.default_rng(7)
    hdr = _header()
    blocks = [(0, 1), (0, 2), (1, 1), (1, 2), (2, 1)]
    ents = _entries(rng, hdr, blocks)
    full = str(tmp_path / "full.sav")
    _write(full, hdr, ents)
    _, _, ends = F.scan(full)
    # stop after the first energy of (1, 2) and a few bytes of the next entry
    cut = str(tmp_path / "cut.sav")
    with open(full, "rb") as f:
        data = f.read()
    with open(cut, "wb") as f:
        f.write(data[:ends[7] + 13])
    h, keep, off, done = F.restart_point(cut, mxpar=2, irstrt=2)
    assert done == {(0, 1), (0, 2), (1, 1)} and len(keep) == 6 and off == ends[6]
    h, keep, off, done = F.restart_point(cut, mxpar=2, irstrt=1)
    assert done == {(0, 1), (0, 2)} and len(keep) == 4 and off == ends[4]
    # resume with irstrt = 2 and write the rest
    _, keep, off, done = F.restart_point(cut, mxpar=2, irstrt=2)
    with F.SMatrixWriter(cut, hdr, append_at=off) as w:
        for e in ents[len(keep):]:
            w.write(e["jtot"], e["inrg"], e["ib"], e["energy"], e["level"], e["lval"], e["wvec"], e["sr"], e["si"],
                    efv=e["efv"], eref=e["eref"], iexch=e["iexch"], wght=e["wght"])
    with open(cut, "rb") as f1, open(full, "rb") as f2:
        assert f1.read() == f2.read()


def test_restart_of_an_empty_run(tmp_path):
    hdr = _header()
    path = str(tmp_path / "e.sav")
    with F.SMatrixWriter(path, hdr):
        pass
    h, keep, off, done = F.restart_point(path, mxpar=2)
    assert keep == [] and done == set() and off == os.path.getsize(path)


def test_rejects_old_formats(tmp_path):
    hdr = _header(iprogm=16)
    path = str(tmp_path / "old.sav")
    with F.SMatrixWriter(path, hdr):
        pass
    with pytest.raises(F.FormatError):
        F.read(path)
