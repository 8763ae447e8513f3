"""MOLSCAT's saved S-matrix file (unit ISAVEU): the writer SPROC uses, the
reader SREAD/HDREAD/SLPRD implement, and RESTRT's scan for where a run
stopped.

The file is Fortran unformatted sequential: every record is its payload
between two 4-byte little-endian byte counts; INTEGER is 4 bytes, DOUBLE
PRECISION 8, CHARACTER(n) n bytes blank-padded.  Output format IPROGM = 19
(source_code/mol.driver.f:227), the records in this order:

  header (once, source_code/sproc.f:797-807)
    LABEL(80) ITYPE NSTATE NQN URED IPROGM
    JSTATE(1:NSTATE*NQN)                      (column-major JSTATE(NSTATE, NQN))
    NLEVEL ELEVEL(1:NLEVEL)
    NDGVL NCONST NRSQ IBOUND ITPSUB
    NEFV ISVEFV (EFVNAM(20) EFVUNT(6), 1:NEFV)
    NFIELD NNRG ENERGY(1:NNRG)
  per (JTOT, IB, field, energy) with NOPEN > 0 (source_code/sproc.f:254-281)
    JTOT INRG IB JFIELD ENERGN EFV(0:NEFVP) EREF IEXCH WGHT NOPEN
    (LEVEL, L, WVEC, 1:NOPEN)                 L as INTEGER, or CENT as DOUBLE when IBOUND > 0
    ((SREAL(I,J), J=1,I), I=1,NOPEN)          or the K form: (SREAL(I,I), SIMAG(I,I)), K's lower rows
    ((SIMAG(I,J), J=1,I), I=1,NOPEN)
    [ESUM]                                    K form only (source_code/sproc.f:360)

NEFVP = max(NEFV, ISVEFV) when NEFV > 0, else 0 (source_code/sread.f:
HDREAD).  LEVEL is the level index (1-based) of the channel's pair state,
WVEC the channel wavevector in 1/Angstrom, channels in the order of the
S-matrix's rows (CHNSRT's)."""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field
from typing import BinaryIO, Iterator, List, Optional, Sequence, Tuple

import numpy as np

IPROGM = 19
LEFVN, LEFVU = 20, 6


class FormatError(RuntimeError):
    """The file is not an S-matrix file of output format 17 or later, or a
    record is cut short."""


def _chars(s: str, n: int) -> bytes:
    b = s.encode("ascii")
    if len(b) > n:
        raise ValueError(f"'{s}' is longer than CHARACTER({n})")
    return b.ljust(n, b" ")


class _Rec:
    """Payload of one record, built item by item."""

    def __init__(self):
        self.parts: List[bytes] = []

    def i4(self, *v):
        self.parts.append(np.asarray(v, dtype="<i4").tobytes())
        return self

    def f8(self, *v):
        self.parts.append(np.asarray(v, dtype="<f8").tobytes())
        return self

    def raw(self, b: bytes):
        self.parts.append(b)
        return self

    def bytes(self) -> bytes:
        return b"".join(self.parts)


def _write_record(f: BinaryIO, payload: bytes) -> None:
    n = len(payload)
    if n > 0x7FFFFFFF:
        raise ValueError("record longer than 2 GiB")
    mark = struct.pack("<i", n)
    f.write(mark)
    f.write(payload)
    f.write(mark)


def _read_record(f: BinaryIO) -> Optional[bytes]:
    head = f.read(4)
    if len(head) == 0:
        return None
    if len(head) < 4:
        raise FormatError("record marker cut short")
    (n,) = struct.unpack("<i", head)
    if n < 0:
        raise FormatError("continued (sub)records are not used by S-matrix files")
    payload = f.read(n)
    tail = f.read(4)
    if len(payload) < n or len(tail) < 4 or struct.unpack("<i", tail)[0] != n:
        raise FormatError("record cut short")
    return payload


class _Cursor:
    """Reads the items of one record in order."""

    def __init__(self, payload: bytes):
        self.b = payload
        self.p = 0

    def _take(self, n):
        if self.p + n > len(self.b):
            raise FormatError("record shorter than its items")
        v = self.b[self.p:self.p + n]
        self.p += n
        return v

    def i4(self, k=None):
        v = np.frombuffer(self._take(4 * (1 if k is None else k)), dtype="<i4")
        return int(v[0]) if k is None else v.astype(np.int64)

    def f8(self, k=None):
        v = np.frombuffer(self._take(8 * (1 if k is None else k)), dtype="<f8")
        return float(v[0]) if k is None else v.copy()

    def chars(self, n):
        return self._take(n).decode("ascii").rstrip(" ")


@dataclass
class Header:
    label: str
    itype: int
    jstate: np.ndarray          # (NSTATE, NQN), integer
    ured: float
    elevel: np.ndarray          # cm-1
    energies: np.ndarray        # cm-1
    nfield: int = 1
    ndgvl: int = 0
    nconst: int = 0
    nrsq: int = 0
    ibound: int = 0
    itpsub: int = 0
    efv_names: List[Tuple[str, str]] = field(default_factory=list)   # (EFVNAM, EFVUNT) per external variable
    isvefv: int = 0
    iprogm: int = IPROGM

    @property
    def nefvp(self) -> int:
        nefv = len(self.efv_names)
        return max(nefv, self.isvefv) if nefv > 0 else 0


@dataclass
class Entry:
    """One saved S-matrix (or K form): the first three records of
    source_code/sproc.f:254-281 and, for the K form, ESUM."""
    jtot: int
    inrg: int                   # 1-based energy index
    ib: int                     # symmetry block (M)
    jfield: int
    energy: float               # cm-1
    efv: np.ndarray             # EFV(0:NEFVP)
    eref: float
    iexch: int
    wght: float
    level: np.ndarray           # (NOPEN,) level index of each open channel, 1-based
    lval: np.ndarray            # (NOPEN,) L (IBOUND = 0) or CENT (IBOUND > 0)
    wvec: np.ndarray            # (NOPEN,) 1/Angstrom
    sr: Optional[np.ndarray] = None     # (NOPEN, NOPEN) symmetric
    si: Optional[np.ndarray] = None
    sdiag: Optional[np.ndarray] = None  # K form: complex diagonal of S
    k: Optional[np.ndarray] = None      # K form: K matrix, symmetric
    esum: Optional[float] = None

    @property
    def nopen(self) -> int:
        return len(self.level)


def _lower_rows(a: np.ndarray) -> np.ndarray:
    n = a.shape[0]
    return np.concatenate([a[i, :i + 1] for i in range(n)]) if n else np.zeros(0)


def _from_lower_rows(v: np.ndarray, n: int) -> np.ndarray:
    """SREAD: ((S(I,J), J=1,I), I=1,N), then S(J,I) = S(I,J)."""
    s = np.zeros((n, n))
    p = 0
    for i in range(n):
        s[i, :i + 1] = v[p:p + i + 1]
        p += i + 1
    return s + np.tril(s, -1).T


def eigenphase_sum(k: np.ndarray, previous: float = 10.0) -> float:
    """EPSUM (source_code/epsum.f): sum of atan of K's eigenvalues over pi,
    moved by whole units to within 1/2 of the previous value (10 on the
    first call)."""
    k = np.asarray(k, dtype=np.float64)
    if k.shape[0] == 1:
        s = math.atan(k[0, 0])
    else:
        s = float(np.sum(np.arctan(np.linalg.eigvalsh(k))))
    s /= math.pi
    delta = previous - s + 0.5
    if delta <= 0.0:
        delta -= 1.0
    return s + float(int(delta))


class SMatrixWriter:
    """Writes the header on open and one entry per write().  kmat=True
    writes the K form (IPHSUM > 0 in the reference): diag(S), K's lower
    rows and the eigenphase sum."""

    def __init__(self, path: str, header: Header, kmat: bool = False, append_at: Optional[int] = None):
        self.header = header
        self.kmat = kmat
        self._esum_last = 10.0
        if append_at is None:
            self.f = open(path, "wb")
            self._write_header()
        else:
            # a restart: keep the file up to the end of its last complete
            # entry and continue from there, as the reference's unit is
            # positioned after RESTRT's last read
            self.f = open(path, "r+b")
            self.f.truncate(append_at)
            self.f.seek(append_at)

    def _write_header(self):
        h = self.header
        js = np.asarray(h.jstate, dtype=np.int64)
        nstate, nqn = js.shape
        _write_record(self.f, _Rec().raw(_chars(h.label, 80)).i4(h.itype, nstate, nqn).f8(h.ured)
                      .i4(h.iprogm).bytes())
        _write_record(self.f, _Rec().i4(*js.T.reshape(-1)).bytes())
        _write_record(self.f, _Rec().i4(len(h.elevel)).f8(*h.elevel).bytes())
        _write_record(self.f, _Rec().i4(h.ndgvl, h.nconst, h.nrsq, h.ibound, h.itpsub).bytes())
        r = _Rec().i4(len(h.efv_names), h.isvefv)
        for name, unit in h.efv_names:
            r.raw(_chars(name, LEFVN)).raw(_chars(unit, LEFVU))
        _write_record(self.f, r.bytes())
        _write_record(self.f, _Rec().i4(h.nfield, len(h.energies)).f8(*h.energies).bytes())

    def write(self, jtot: int, inrg: int, ib: int, energy: float, level: Sequence[int], lval, wvec: Sequence[float],
              sr: np.ndarray, si: np.ndarray, k: Optional[np.ndarray] = None, jfield: int = 1,
              efv: Optional[Sequence[float]] = None, eref: float = 0.0, iexch: int = 0, wght: float = 1.0) -> None:
        h = self.header
        nopen = len(level)
        if nopen == 0:
            return
        efv = np.zeros(h.nefvp + 1) if efv is None else np.asarray(efv, dtype=np.float64)
        if len(efv) != h.nefvp + 1:
            raise ValueError(f"EFV needs {h.nefvp + 1} values (EFV(0:NEFVP))")
        _write_record(self.f, _Rec().i4(jtot, inrg, ib, jfield).f8(energy).f8(*efv).f8(eref).i4(iexch).f8(wght)
                      .i4(nopen).bytes())
        r = _Rec()
        for i in range(nopen):
            r.i4(int(level[i]))
            if h.ibound == 0:
                r.i4(int(lval[i]))
            else:
                r.f8(float(lval[i]))
            r.f8(float(wvec[i]))
        _write_record(self.f, r.bytes())
        sr = np.asarray(sr, dtype=np.float64)
        si = np.asarray(si, dtype=np.float64)
        if not self.kmat:
            _write_record(self.f, _Rec().f8(*_lower_rows(sr)).bytes())
            _write_record(self.f, _Rec().f8(*_lower_rows(si)).bytes())
        else:
            if k is None:
                raise ValueError("the K form needs the K matrix")
            d = np.empty(2 * nopen)
            d[0::2], d[1::2] = np.diag(sr), np.diag(si)
            _write_record(self.f, _Rec().f8(*d).bytes())
            _write_record(self.f, _Rec().f8(*_lower_rows(np.asarray(k, dtype=np.float64))).bytes())
            self._esum_last = eigenphase_sum(k, self._esum_last)
            _write_record(self.f, _Rec().f8(self._esum_last).bytes())
        self.f.flush()

    def close(self):
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _read_header(f: BinaryIO) -> Header:
    rec = _read_record(f)
    if rec is None:
        raise FormatError("empty file")
    c = _Cursor(rec)
    label = c.chars(80)
    itype, nstate, nqn = c.i4(), c.i4(), c.i4()
    ured = c.f8()
    iprogm = c.i4()
    if iprogm < 17:
        raise FormatError(f"output format IPROGM = {iprogm}; only 17 and later are read")
    js = _Cursor(_read_record(f) or b"").i4(nstate * nqn).reshape(nqn, nstate).T
    c = _Cursor(_read_record(f) or b"")
    nlev = c.i4()
    elevel = c.f8(nlev)
    c = _Cursor(_read_record(f) or b"")
    ndgvl, nconst, nrsq, ibound, itpsub = (c.i4() for _ in range(5))
    c = _Cursor(_read_record(f) or b"")
    nefv, isvefv = c.i4(), c.i4()
    names = [(c.chars(LEFVN), c.chars(LEFVU)) for _ in range(nefv)]
    c = _Cursor(_read_record(f) or b"")
    nfield, nnrg = c.i4(), c.i4()
    energies = c.f8(nnrg)
    return Header(label=label, itype=itype, jstate=js, ured=ured, elevel=elevel, energies=energies, nfield=nfield,
                  ndgvl=ndgvl, nconst=nconst, nrsq=nrsq, ibound=ibound, itpsub=itpsub, efv_names=names,
                  isvefv=isvefv, iprogm=iprogm)


def _read_entry(f: BinaryIO, h: Header, kmat: bool) -> Optional[Entry]:
    """SLPRD / KLPRD with IREAD = 0 (source_code/sread.f); None at a clean
    end of file, FormatError when the entry is cut short."""
    rec = _read_record(f)
    if rec is None:
        return None
    c = _Cursor(rec)
    jtot, inrg, ib, jfield = c.i4(), c.i4(), c.i4(), c.i4()
    energy = c.f8()
    efv = c.f8(h.nefvp + 1)
    eref = c.f8()
    iexch = c.i4()
    wght = c.f8()
    nopen = c.i4()
    rec = _read_record(f)
    if rec is None:
        raise FormatError("entry cut short after its first record")
    c = _Cursor(rec)
    level = np.zeros(nopen, dtype=np.int64)
    lval = np.zeros(nopen, dtype=np.int64 if h.ibound == 0 else np.float64)
    wvec = np.zeros(nopen)
    for i in range(nopen):
        level[i] = c.i4()
        lval[i] = c.i4() if h.ibound == 0 else c.f8()
        wvec[i] = c.f8()
    e = Entry(jtot=jtot, inrg=inrg, ib=ib, jfield=jfield, energy=energy, efv=efv, eref=eref, iexch=iexch, wght=wght,
              level=level, lval=lval, wvec=wvec)
    r1, r2 = _read_record(f), _read_record(f)
    if r1 is None or r2 is None:
        raise FormatError("entry cut short before its matrices")
    ntri = nopen * (nopen + 1) // 2
    if not kmat:
        e.sr = _from_lower_rows(_Cursor(r1).f8(ntri), nopen)
        e.si = _from_lower_rows(_Cursor(r2).f8(ntri), nopen)
    else:
        d = _Cursor(r1).f8(2 * nopen)
        e.sdiag = d[0::2] + 1j * d[1::2]
        e.k = _from_lower_rows(_Cursor(r2).f8(ntri), nopen)
        r3 = _read_record(f)
        if r3 is None:
            raise FormatError("K-form entry without its eigenphase sum")
        e.esum = _Cursor(r3).f8()
    return e


def read(path: str, kmat: bool = False) -> Tuple[Header, List[Entry]]:
    """The header and every complete entry; an entry cut short at the end
    (a run that stopped while writing) is dropped."""
    h, entries, _ = scan(path, kmat)
    return h, entries


def scan(path: str, kmat: bool = False) -> Tuple[Header, List[Entry], List[int]]:
    """As read(), plus the byte offsets: the end of the header, then the
    end of each complete entry."""
    with open(path, "rb") as f:
        h = _read_header(f)
        entries: List[Entry] = []
        ends = [f.tell()]
        while True:
            try:
                e = _read_entry(f, h, kmat)
            except FormatError:
                break
            if e is None:
                break
            entries.append(e)
            ends.append(f.tell())
    return h, entries, ends


def iterate(path: str, kmat: bool = False) -> Iterator[Entry]:
    with open(path, "rb") as f:
        h = _read_header(f)
        while True:
            e = _read_entry(f, h, kmat)
            if e is None:
                return
            yield e


def restart_point(path: str, mxpar: int, irstrt: int = 2, kmat: bool = False):
    """RESTRT's choice of where to resume (source_code/restrt.f:19-27,
    250-296), reading the entries in file order: with irstrt = 2 a
    symmetry block is complete at its entry for the last energy and field
    (INRG = NNRG, JFIELD = NFIELD); with irstrt = 1 a JTOT set is complete
    at that entry of its last block (IB = mxpar).  Entries after the last
    complete unit are dropped.  Returns (header, kept entries, the byte
    offset to continue writing at, the set of complete (JTOT, IB))."""
    if irstrt not in (1, 2):
        raise ValueError("restart_point: IRSTRT 1 (after a complete JTOT set) or 2 (after a complete block)")
    h, entries, ends = scan(path, kmat)
    nnrg = len(h.energies)
    last = -1
    for i, e in enumerate(entries):
        if e.inrg == nnrg and e.jfield == h.nfield and (irstrt == 2 or e.ib == mxpar):
            last = i
    keep = entries[:last + 1]
    return h, keep, ends[last + 1], {(e.jtot, e.ib) for e in keep}
