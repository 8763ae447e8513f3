"""ORACLE — test infrastructure only.

FP64 CPU restatement of the MOLSCAT scattering path (units, ITYPE=1 rotor
basis and coupling, POTENL powers, HAMMAT, YINIT, FINDRM, the
log-derivative propagators MDPROP/LDPROP/ODPROP, Y->K->S matching and the
cross-section/scattering-length formulas), written from the reference's
Fortran as a checker for the HIP engine.  Dense inversions go through the
plain-C restatement in ``oracle/csrc/oracle_la.c`` (SYMINV inline path,
Bunch-Kaufman DSYTF2 + DSYNEG + DSYTRI); eigen-decompositions and the one
general solve use numpy's LAPACK (``eigh`` for DIAGVC/DIAGVL, ``solve``
for DGESV), the same third-party dependency the reference links
(source_code/Makefile:12, version unpinned).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.

Every function cites the reference routine (file:line under source_code/)
it restates.  Matrices are numpy arrays; index conventions are 0-based.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field
from functools import lru_cache
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_la.so")

# ---------------------------------------------------------------------------
# physical constants: source_code/physical_constants_module.f:7-88
# ---------------------------------------------------------------------------
PI_VALUE = 3.141592653589793
PLANCK_SI = 6.62607015e-34
HBAR_SI = PLANCK_SI / 2.0 / PI_VALUE
C_SI = 2.99792458e8
C_CM = C_SI * 1.0e2
AMU_SI = 1.66053906892e-27
ANGSTROM_SI = 1.0e-10
BFCT = HBAR_SI / (4.0 * PI_VALUE * C_CM) / (ANGSTROM_SI) ** 2 / AMU_SI
HZ_IN_INV_CM = 3.335640951e-11
MHZ_IN_INV_CM = HZ_IN_INV_CM * 1.0e6
BOHR_MAGNETON_INV_CM_PER_G = 0.46686447719 / 1.0e4


# ---------------------------------------------------------------------------
# C restatement of the inversion routines
# ---------------------------------------------------------------------------
def build_lib(force: bool = False) -> str:
    """Compile oracle/csrc/oracle_la.c with gcc (oracle/Makefile)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle_la.so"])
    return _LIB_PATH


@lru_cache(maxsize=1)
def _lib():
    build_lib()
    lib = ctypes.CDLL(_LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.orc_syminv.argtypes = [dp, ctypes.c_int, ctypes.c_int]
    lib.orc_syminv.restype = ctypes.c_int
    lib.orc_syminv_inline.argtypes = [dp, ctypes.c_int, ctypes.c_int]
    lib.orc_syminv_inline.restype = ctypes.c_int
    lib.orc_dsytf2_l.argtypes = [dp, ctypes.c_int, ctypes.c_int, ip]
    lib.orc_dsytf2_l.restype = ctypes.c_int
    lib.orc_dsyneg_l.argtypes = [dp, ctypes.c_int, ctypes.c_int, ip]
    lib.orc_dsyneg_l.restype = ctypes.c_int
    lib.orc_dsytri_l.argtypes = [dp, ctypes.c_int, ctypes.c_int, ip]
    lib.orc_dsytri_l.restype = ctypes.c_int
    return lib


def syminv(a: np.ndarray) -> Tuple[np.ndarray, int]:
    """SYMINV (source_code/syminv.f:1-97 with SYMINL source_code/syminl-alloc.f).

    Takes a symmetric matrix (only the lower triangle is read), returns the
    full symmetric inverse (lower triangle mirrored, as every caller on the
    path uses only the lower triangle or mirrors it) and KOUNT, the number of
    negative eigenvalues.  KOUNT = N+1 flags an exact zero pivot on the
    inline path; a negative value is -INFO from DSYTRF/DSYTRI.
    """
    n = a.shape[0]
    f = np.array(a, dtype=np.float64, order="F", copy=True)
    k = _lib().orc_syminv(f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, n)
    lo = np.tril(f)
    full = lo + np.tril(f, -1).T
    return full, int(k)


@lru_cache(maxsize=1)
def _openblas():
    """numpy's bundled OpenBLAS (ILP64; the LAPACK names carry a scipy_
    prefix and a 64_ suffix): the library numpy's own matrix products use, so
    the process has one BLAS thread pool.  Used only by syminv_lapack, the
    CPU baseline's inversion."""
    import glob

    d = os.path.join(os.path.dirname(os.path.dirname(np.__file__)), "numpy.libs")
    cands = sorted(glob.glob(os.path.join(d, "libscipy_openblas64_*.so")))
    if not cands:
        raise RuntimeError("numpy's OpenBLAS not found")
    lib = ctypes.CDLL(cands[0])
    ip = ctypes.POINTER(ctypes.c_int64)
    dp = ctypes.POINTER(ctypes.c_double)
    cp = ctypes.c_char_p
    lib.scipy_dsytrf_64_.restype = None
    lib.scipy_dsytri_64_.restype = None
    lib.scipy_dsytrf_64_.argtypes = [cp, ip, dp, ip, ip, dp, ip, ip, ctypes.c_size_t]
    lib.scipy_dsytri_64_.argtypes = [cp, ip, dp, ip, ip, dp, ip, ctypes.c_size_t]
    return lib


def syminv_lapack(a: np.ndarray) -> Tuple[np.ndarray, int]:
    """SYMINL as the reference runs it for N > 30 (source_code/syminl-alloc.f:
    17-41): blocked DSYTRF('L') and DSYTRI('L') from an optimised LAPACK (here
    numpy's OpenBLAS, threaded as its thread pool allows) with DSYNEG's
    inertia count from the block-diagonal factor (source_code/dsyneg.f:47-73).
    Same interface as syminv; the CPU baseline's inversion."""
    lib = _openblas()
    n = a.shape[0]
    f = np.array(a, dtype=np.float64, order="F", copy=True)
    ipiv = np.zeros(n, dtype=np.int64)
    i64 = ctypes.c_int64
    nn = i64(n)
    info = i64(0)
    dp = ctypes.POINTER(ctypes.c_double)
    ipp = ctypes.POINTER(i64)
    lwork = i64(-1)
    wq = np.zeros(1)
    lib.scipy_dsytrf_64_(b"L", ctypes.byref(nn), f.ctypes.data_as(dp), ctypes.byref(nn), ipiv.ctypes.data_as(ipp),
                         wq.ctypes.data_as(dp), ctypes.byref(lwork), ctypes.byref(info), 1)
    lw = max(1, int(wq[0]))
    work = np.zeros(max(lw, n))
    lwork = i64(len(work))
    lib.scipy_dsytrf_64_(b"L", ctypes.byref(nn), f.ctypes.data_as(dp), ctypes.byref(nn), ipiv.ctypes.data_as(ipp),
                         work.ctypes.data_as(dp), ctypes.byref(lwork), ctypes.byref(info), 1)
    if info.value != 0:
        return f, -abs(int(info.value))
    # DSYNEG: 1x1 blocks by sign, 2x2 blocks (ipiv(k) = ipiv(k+1) < 0 in the
    # lower factor) by the scaled determinant
    kount = 0
    k = 0
    while k < n:
        if ipiv[k] > 0:
            if f[k, k] < 0.0:
                kount += 1
            k += 1
        else:
            aa, bb, cc = f[k, k], f[k + 1, k], f[k + 1, k + 1]
            tt = abs(bb)
            if (cc / tt) * aa - tt < 0.0:
                kount += 1
            k += 2
    lib.scipy_dsytri_64_(b"L", ctypes.byref(nn), f.ctypes.data_as(dp), ctypes.byref(nn), ipiv.ctypes.data_as(ipp),
                         work.ctypes.data_as(dp), ctypes.byref(info), 1)
    lo = np.tril(f)
    return lo + np.tril(f, -1).T, kount


def syminv_raw(a: np.ndarray) -> Tuple[np.ndarray, int]:
    """SYMINV returning the raw column-major output (upper stale for N>30)."""
    n = a.shape[0]
    f = np.array(a, dtype=np.float64, order="F", copy=True)
    k = _lib().orc_syminv(f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, n)
    return f, int(k)


# ---------------------------------------------------------------------------
# ITYPE=1 atom + linear rigid rotor basis and coupling
# ---------------------------------------------------------------------------
def rotor_levels(jlevels: Sequence[int], be: float) -> np.ndarray:
    """Level energies (cm-1): source_code/setbas.f:77-81 with ALPHAE=DE=0."""
    return np.array([(be - 0.0 / 2.0) * float(j * (j + 1)) - 0.0 for j in jlevels])


def itype1_channels(jtot: int, ib: int, jlevels: Sequence[int]) -> List[Tuple[int, int]]:
    """Channel list (state index, L) of one symmetry block.

    source_code/base.f:276-352 (unified ITYPE code, ITP=1): L runs from
    LMIN to LMAX in the outer loop, pair states in input order inside;
    block IB keeps channels with (L + j + JTOT) mod 2 == IB mod 2 and
    |JTOT-L| <= j <= JTOT+L.  JMIN starts at 0 (source_code/base.f:512).
    """
    jmax = max(max(jlevels), 0)
    jmin = min(min(jlevels), 0)
    lmax = jtot + jmax
    lmin = jtot - jmax
    if lmin < 0:
        lmin = max(jmin - jtot, 0)
    ipty = ib - 2 * (ib // 2)
    chans = []
    for li in range(lmin, lmax + 1):
        jk = abs(jtot - li)
        jtop = jtot + li
        for i, ji in enumerate(jlevels):
            lpj = li + ji + jtot
            if (lpj - 2 * (lpj // 2)) != ipty:
                continue
            if ji < jk or ji > jtop:
                continue
            chans.append((i, li))
    return chans


@lru_cache(maxsize=None)
def _w3j000(a: int, b: int, c: int) -> float:
    from sympy.physics.wigner import wigner_3j

    return float(wigner_3j(a, b, c, 0, 0, 0))


@lru_cache(maxsize=None)
def _w6j(a: int, b: int, c: int, d: int, e: int, f: int) -> float:
    from sympy.physics.wigner import wigner_6j

    return float(wigner_6j(a, b, c, d, e, f))


def itype1_coupling(jtot: int, chans, jlevels, lambdas: Sequence[int]) -> np.ndarray:
    """Percival-Seaton coupling VL(NHAM, N(N+1)/2), lambda fastest.

    source_code/couple.f:51-143: packed index runs over the upper triangle
    column by column (row <= col); value
    sqrt((2jc+1)(2jr+1)(2Lc+1)(2Lr+1)) (jc lam jr;000)(Lc lam Lr;000)
    {jr Lr J; Lc jc lam} (-1)^(jc+jr+J) (J6J's {J1 J2 J3; L1 L2 L3} with
    J1=jr, J2=Lr, J3=JTOT, L1=Lc, L2=jc, L3=lam, source_code/j6j.f:10-13).
    The 3j/6j symbols are exact
    (sympy) here, independent of the reference's recursions.
    """
    n = len(chans)
    nham = len(lambdas)
    vl = np.zeros((nham, n * (n + 1) // 2))
    idx = 0
    for icol in range(n):
        jc = jlevels[chans[icol][0]]
        lc = chans[icol][1]
        for irow in range(icol + 1):
            jr = jlevels[chans[irow][0]]
            lr = chans[irow][1]
            for il, lam in enumerate(lambdas):
                t = _w3j000(jc, lam, jr)
                if t == 0.0:
                    continue
                u = _w3j000(lc, lam, lr)
                if u == 0.0:
                    continue
                s = _w6j(jr, lr, jtot, lc, jc, lam)
                if s == 0.0:
                    continue
                v = math.sqrt((2 * jc + 1) * (2 * jr + 1) * (2 * lc + 1) * (2 * lr + 1)) * t * u * s
                if (jc + jr + jtot) % 2 == 1:
                    v = -v
                vl[il, idx] = v
            idx += 1
    return vl


def packed_to_full(vl_col: np.ndarray, n: int) -> np.ndarray:
    """Unpack one VL block (upper triangle, column by column) into a full matrix."""
    m = np.zeros((n, n))
    iu = np.triu_indices(n)
    # triu_indices is row-major over (row, col) pairs; build column order
    rows, cols = [], []
    for c in range(n):
        for r in range(c + 1):
            rows.append(r)
            cols.append(c)
    m[rows, cols] = vl_col
    m[cols, rows] = vl_col
    del iu
    return m


# ---------------------------------------------------------------------------
# POTENL case 1 (powers and exponentials)
# ---------------------------------------------------------------------------
@dataclass
class PowerPotential:
    """POTENL case 1, source_code/potenl.f:160-200.

    terms[lam_index] = list of (A, n) with n != 0 meaning A*R**n, or
    (A, 0, E) meaning A*exp(E*R).  R is in the potential's length unit
    (R_prog * RSCALE); the result is in EPSIL units.
    """

    terms: List[List[tuple]]
    rm: float = 1.0
    epsil: float = 1.0

    def __call__(self, r: float) -> np.ndarray:
        p = np.zeros(len(self.terms))
        for i, tl in enumerate(self.terms):
            tot = 0.0
            for t in tl:
                if t[1] != 0:
                    rr = r
                    if t[1] < 0 and abs(r) < 1.0e-8:
                        rr = 1.0e-8
                    tot += t[0] * rr ** t[1]
                else:
                    tot += t[0] * math.exp(t[2] * r)
            p[i] = tot
        return p

    def deriv(self, r: float, ider: int) -> np.ndarray:
        """POTENL with ICNTRL = IDER (source_code/potenl.f:177-185): the
        IDER'th derivative of each radial term with respect to its argument
        (R_prog * RSCALE), N R^(N-1) and N (N-1) R^(N-2) for powers, E^IDER
        exp(E R) for exponentials; no clamp."""
        p = np.zeros(len(self.terms))
        for i, tl in enumerate(self.terms):
            tot = 0.0
            for t in tl:
                if t[1] != 0:
                    rr = r
                    if t[1] < 0 and abs(r) < 1.0e-8:
                        rr = 1.0e-8
                    term = rr ** t[1]
                    if ider == 1:
                        term = float(t[1]) * term / rr
                    elif ider == 2:
                        term = float(t[1] * (t[1] - 1)) * term / (rr * rr)
                else:
                    term = math.exp(t[2] * r) * t[2] ** ider
                tot += t[0] * term
            p[i] = tot
        return p


def lj_rotor_potential(cvals: Sequence[float], rm: float, epsil: float) -> PowerPotential:
    """V_lam = c_lam * eps * [(Rm/R)^12 - 2 (Rm/R)^6] as &POTL NTERM=2,
    NPOWER=-12,-6, A=c,-2c (the form of examples/input/molscat-basic1.input)."""
    return PowerPotential([[(c, -12), (-2.0 * c, -6)] for c in cvals], rm=rm, epsil=epsil)


# ---------------------------------------------------------------------------
# one (JTOT, IB) block and its W(R)
# ---------------------------------------------------------------------------
@dataclass
class Block:
    n: int
    vl: np.ndarray  # (NHAM, N(N+1)/2)
    eint: np.ndarray  # reduced units
    cent: np.ndarray  # L(L+1)
    lvals: np.ndarray
    levidx: np.ndarray  # 0-based level index of each channel
    potl: Callable[[float], np.ndarray]
    rscale: float
    ep2ru: float
    cm2ru: float
    vl_full: np.ndarray = field(default=None)  # (NHAM, N, N)
    nconst: int = 0  # constant terms: the VL blocks after the radial ones
    nrsq: int = 0  # terms multiplying 1/R^2 (the centrifugal operator): the last nrsq VL blocks
    pconst: Optional[np.ndarray] = None  # their P values, VCONST * CM2RU (hammat.f:57-73)
    scalam: float = 1.0  # SCAPOT's factor (source_code/scapot-default.f)
    expval: Optional[Tuple[int, int, float]] = None  # PERTRB's /EXPVAL/ (I, IPOW, DELTAN), I 1-based

    def __post_init__(self):
        if self.vl_full is None:
            self.vl_full = np.stack([packed_to_full(self.vl[k], self.n) for k in range(self.vl.shape[0])])

    def pvec(self, r: float) -> np.ndarray:
        """HAMMAT's P array (source_code/hammat.f:38-73), in its order: the
        radial terms at R*RSCALE (EPSIL units) clamped to 1e60; the NCONST
        constants appended as VCONST/EP2CM; SCAPOT multiplies the radial
        terms by SCALAM; PERTRB adds DELTAN/(R*RSCALE)**IPOW to P(I) for
        1 <= I <= MXLAM+NCONST (source_code/pertrb.f:19); then everything
        times EP2RU."""
        x = r * self.rscale
        p = np.array(self.potl(x), dtype=np.float64)
        p = np.where(np.abs(p) > 1.0e60, np.sign(p) * 1.0e60, p)
        mxlam = len(p)
        p = p * self.scalam
        i, ipow, dn = (0, 0, 0.0) if self.expval is None else (int(self.expval[0]), int(self.expval[1]),
                                                                float(self.expval[2]))
        if 0 < i <= mxlam:
            p[i - 1] = p[i - 1] + dn / x ** ipow
        p = self.ep2ru * p
        if self.nconst > 0:
            pc = np.array(self.pconst, dtype=np.float64)
            if mxlam < i <= mxlam + self.nconst:
                # the constant in EPSIL units (VCONST / EP2CM) takes the shift
                pc[i - mxlam - 1] = self.ep2ru * (pc[i - mxlam - 1] / self.ep2ru + dn / x ** ipow)
            p = np.concatenate([p, pc])
        return p

    def dermat(self, r: float, ider: int) -> np.ndarray:
        """DERMAT (source_code/dermat.f:38-126, NUMDER off, NRSQ = 0): the
        IDER'th derivative of W at R: POTENL's derivatives at R*RSCALE
        (derivatives with respect to that argument, as DERMAT takes them)
        times EP2RU, zero for the NCONST constant terms, SCAPOT's SCALAM, no
        PERTRB (dermat.f:93-95); on the diagonal the centrifugal term's
        derivative, -2/R^3 or 6/R^4 (clamped at 1e16) times L(L+1).  For
        NRSQ = 0 the reference adds RSQ*CENT with RSQ set only in its NRSQ > 0
        branch (dermat.f:99-110), so its value there is whatever the variable
        held; the derivative of the centrifugal term is what the code means."""
        x = r * self.rscale
        p = self.potl.deriv(x, ider)
        p = self.ep2ru * p
        p = p * self.scalam
        if self.nconst > 0:
            p = np.concatenate([p, np.zeros(self.nconst)])
        w = np.tensordot(p, self.vl_full, axes=(0, 0))
        rsq = -2.0 / r ** 3 if ider == 1 else 6.0 / r ** 4
        rsq = min(rsq, 1.0e16)
        w[np.diag_indices(self.n)] += rsq * self.cent
        return w

    def hammat(self, r: float, ered: float) -> Tuple[np.ndarray, np.ndarray]:
        """HAMMAT: source_code/hammat.f:1-116 with SUMLAM source_code/sumlam.f:29-36
        (NRSQ = 0; EINT added only when NCONST = 0, hammat.f:95-99).

        Returns the full W(N,N) and DIAG(N) = diag(W).
        """
        rsq = min(1.0 / (r * r), 1.0e16)
        p = self.pvec(r)
        if self.nrsq > 0:
            # the centrifugal operator's terms take P = 1/R^2 (hammat.f:74-83)
            p = np.concatenate([p, np.full(self.nrsq, rsq)])
        w = np.tensordot(p, self.vl_full, axes=(0, 0))
        d = np.diag(w).copy()
        if self.nrsq == 0:
            d = d + rsq * self.cent
        if self.nconst == 0:
            d = d + self.eint
        d = d - ered
        np.fill_diagonal(w, d)
        return w, d.copy()


def make_itype1_block(jtot, ib, jlevels, be, lambdas, potl: PowerPotential, ured, runit=None):
    """BASE + COUPLE for ITYPE=1, plus the unit conversions of
    source_code/mol.driver.f:734-735 and source_code/chckru.f (RUNIT unset
    means RUNIT = RPUNIT = RM)."""
    if runit is None:
        runit = potl.rm
    rscale = runit / potl.rm
    cm2ru = ured * 1.0 * runit * runit / BFCT
    ep2ru = potl.epsil * cm2ru
    chans = itype1_channels(jtot, ib, jlevels)
    n = len(chans)
    elev = rotor_levels(jlevels, be)
    eint = np.array([cm2ru * elev[c[0]] for c in chans])
    lv = np.array([c[1] for c in chans])
    cent = np.array([float(l * (l + 1)) for l in lv])
    vl = itype1_coupling(jtot, chans, jlevels, lambdas)
    return Block(n=n, vl=vl, eint=eint, cent=cent, lvals=lv, levidx=np.array([c[0] for c in chans]),
                 potl=potl, rscale=rscale, ep2ru=ep2ru, cm2ru=cm2ru), runit


# ---------------------------------------------------------------------------
# start-up, grid and step control
# ---------------------------------------------------------------------------
def calck(ep: float, ered: float, eint: np.ndarray) -> float:
    """CALCK: source_code/calck.f."""
    c = 0.0
    for e in eint:
        if ep + ered - e <= 0.0:
            continue
        c = max(c, math.sqrt(ep + ered - e))
    return c


def drcalc(rstart: float, rstop: float, kstep: int, nstep: int, pow_: float) -> float:
    """DRCALC: source_code/drcalc.f:23-38."""
    powint = 1.0 - pow_
    if powint != 0.0:
        powinv = 1.0 / powint
        zstart = rstart ** powint
        zstop = rstop ** powint
        dz = (zstop - zstart) / float(nstep)
        rnow = (zstart + float(kstep + 1) * dz) ** powinv
        rlast = (zstart + float(kstep) * dz) ** powinv
        return rnow - rlast
    dz = rstart * ((rstop / rstart) ** (1.0 / float(nstep)) - 1.0)
    return dz * (1.0 + dz / rstart) ** kstep


UNSET = -1.0e30  # stands for the reference's 'unset' sentinel


def drset(rstart, rstop, step, tolhi, cay, drseg, iprop, pow_, dr_prev=None):
    """DRSET: source_code/drset.f:30-86.  Returns (NSTEP, DR)."""
    rsmall = min(rstart, rstop)
    rlarge = max(rstart, rstop)
    powint = 1.0 - pow_
    prange = rlarge - rsmall
    nstep = 0
    dr = dr_prev
    if iprop == 4:
        if drseg != UNSET:
            dr = drseg
        return nstep, dr
    elif step > 0.0:
        drfit = math.acos(-1.0) / (cay * step)
    elif drseg != UNSET:
        drfit = abs(drseg)
    else:
        drfit = abs(dr_prev)
    if tolhi == 0.0:
        rratio = drfit / rsmall
        if pow_ == 0.0:
            nstep = max(int(round_half_away(prange / drfit)), 1)
            dr = prange / float(nstep)
        elif pow_ == 1.0:
            nstep = int(round_half_away(math.log(rlarge / rsmall) / math.log(rratio + 1.0)))
            dr = drcalc(rstart, rstop, 0, nstep, pow_)
        else:
            zratio = (rlarge / rsmall) ** powint
            zrm1 = zratio - 1.0
            nstep = int(round_half_away(zrm1 / ((rratio + 1.0) ** powint - 1.0)))
            dr = drcalc(rstart, rstop, 0, nstep, pow_)
    elif tolhi < 1.0:
        dr = drfit
    if rstop < rstart:
        dr = -abs(dr)
    return nstep, dr


def round_half_away(x: float) -> float:
    """Fortran NINT."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def findrm(blk: Block, rmin: float, ered: float, irmset: int, cdrive: str = "M") -> Tuple[float, float]:
    """FINDRM: source_code/findrm.f (returns RSTART, RTURN).

    DIAGVL (lowest adiabat) is numpy eigvalsh, ascending like DSYEVR.
    """
    n = blk.n

    def lowest(r):
        w, _ = blk.hammat(r, ered)
        return np.linalg.eigvalsh(w)[0]

    if cdrive != "M":
        rturn = rmin
        itry = 1
    else:
        rturn = 1.0e30
        for i in range(n):
            dif = ered - blk.eint[i]
            if dif < 0.0:
                continue
            rcent = math.sqrt(blk.cent[i] / dif)
            rcent = max(rcent, rmin)
            rturn = min(rturn, rcent)
        itry = 0
        if rturn == rmin:
            itry = 1
    while True:  # label 10
        itry += 1
        rstart = rturn
        if itry > 25:
            return rmin, rturn
        diag1 = lowest(rstart)
        rturn = 0.999 * rstart
        converged = False
        retry = False
        dv1 = 0.0
        for _ in range(100):
            v2 = lowest(rturn)
            dv1 = (v2 - diag1) / (rturn - rstart)
            diag1 = v2
            if dv1 >= 0.0:
                if itry == 1:
                    rturn = rmin
                    retry = True
                    break
                return rmin, 2.0 * rmin
            rstart = rturn
            v1 = v2
            dr = -v1 / dv1
            if dr < -0.3 * rturn:
                dr = -0.3 * rturn
            rturn = rturn + dr
            if dr > 1.0e3:
                break
            if rturn <= 0.0:
                break
            if abs(dr / rturn) <= 1.0e-3:
                converged = True
                break
        if retry:
            continue
        if not converged:
            return rmin, rturn
        break
    # label 160
    if irmset <= 0:
        return rmin, rturn
    nstep = 3 + irmset // 3
    target = 2.303 * float(irmset)
    dr = (1.5 * target / math.sqrt(abs(dv1)) / float(nstep)) ** (2.0 / 3.0)
    # RSTART still holds the secant iterate before the converged RTURN
    # (findrm.f:109 RSTART=RTURN, then RTURN=RTURN+DR and the test at :118);
    # the phase integral starts there (:179 RNEXT=RSTART-DR)
    phase = 0.0
    while True:  # label 220
        restart = False
        xk2 = 0.0
        for istep in range(1, nstep + 1):
            rnext = rstart - dr
            if rnext < 0.0:
                return 0.0, rturn
            d1 = lowest(rnext)
            drnext = 0.0
            if d1 <= 0.0:
                raise NotImplementedError("FINDRM inner allowed region branch")
            xk1 = math.sqrt(d1)
            if istep == 1:
                phase = (2.0 / 3.0) * xk1 * dr
            else:
                xkavg = 0.5 * (xk1 + xk2)
                phase = phase + dr * xkavg
            drnext = max(drnext, (target - phase) / xk1)
            xk2 = xk1
            rstart = rnext
            if istep < nstep:
                dr = drnext / float(nstep - istep)
            if drnext <= 0.0:
                break
            if istep < nstep and dr > 0.5 * rstart and dr > 0.5 * rmin:
                dr = 0.02 * rstart
                restart = True
                break
        if not restart:
            break
    return rstart, rturn


def yinit(blk: Block, ered: float, r: float, zout: bool = True,
          bcyomn: Optional[float] = None, bcycmx: float = -1.0, wkb: bool = True,
          adiab: bool = True) -> np.ndarray:
    """YINIT: source_code/yinit.f:70-150 (ADIAMN/ADIAMX default .TRUE.,
    WKBMN/WKBMX default .TRUE.)."""
    w, _ = blk.hammat(r, ered)
    n = blk.n
    if adiab:
        ev, evec = np.linalg.eigh(w)
    else:
        ev = np.diag(w).copy()
    yd = np.zeros(n)
    nopen = 0
    for i in range(n):
        wv = ev[i]
        if wv > 0.0:
            if zout:
                yd[i] = math.sqrt(abs(wv)) if wkb else -1.0
            else:
                yd[i] = -math.sqrt(abs(wv)) if wkb else bcycmx
        else:
            if zout:
                if bcyomn is None:
                    raise RuntimeError("YINIT: locally open channel at RMIN and BCYOMN unset")
                yd[i] = bcyomn
            else:
                yd[i] = bcyomn if bcyomn is not None else 0.0
            if abs(wv) > 1.0e-10:
                nopen += 1
    if adiab:
        return (evec * yd[None, :]) @ evec.T
    return np.diag(yd)


# ---------------------------------------------------------------------------
# propagators
# ---------------------------------------------------------------------------
def _sym_lower(m: np.ndarray) -> np.ndarray:
    lo = np.tril(m)
    return lo + np.tril(m, -1).T


def mdprop(y: np.ndarray, hammat: Callable[[float], Tuple[np.ndarray, np.ndarray]],
           rstart: float, nstep: int, dr: float, inv: Callable = None,
           wave: Optional[list] = None) -> Tuple[np.ndarray, int]:
    """MDPROP (diabatic modified log-derivative, IPROP=6): source_code/mdprop.f:64-279.

    ``hammat(r)`` returns (W, DIAG) at the propagation energy.  Returns the
    final log-derivative matrix and NODES.  ``inv`` replaces syminv (the CPU
    baseline passes syminv_lapack).  With ``wave`` a list, BOUND's WAVE
    records are appended to it, one per step (mdprop.f:170-176, 212-245):
    W3 = W W2 with W = [H(Y(A) + Y1(A,C))]^-1 after the first inversion and
    W2 = Y23 [H(Y(C) + Y1(C,B))]^-1 Y23 after the second, the matrix that
    carries the wavefunction back from the step's end B to its start A.
    """
    syminv = inv if inv is not None else globals()["syminv"]
    n = y.shape[0]
    h = dr / 2.0
    d1 = h * h / 3.0
    d2 = 2.0 * d1
    d4 = -d1 / 16.0
    half = 0.5 * h
    r = rstart
    nsave = 0
    nodes = 0
    u, diag = hammat(r)
    y = h * y + d1 * u
    idx = np.arange(n)
    for kstep in range(1, nstep + 1):
        r = r + h
        u, diag = hammat(r)
        u = d4 * u
        u[idx, idx] = 0.125
        u, kount = syminv(u)
        if kount > n or kount < 0:
            raise RuntimeError(f"MDPROP: MATRIX INVERSION ERROR AT STEP K = {kstep}")
        nsave += kount
        u[idx, idx] -= 8.0
        wref = diag
        arg = half * np.sqrt(np.abs(wref))
        y14 = np.empty(n)
        y23 = np.empty(n)
        neg = wref < 0.0
        tn = np.tan(arg[neg])
        y14[neg] = arg[neg] / tn - arg[neg] * tn
        y23[neg] = arg[neg] / tn + arg[neg] * tn
        pos = ~neg
        th = np.tanh(arg[pos])
        y14[pos] = arg[pos] / th + arg[pos] * th
        y23[pos] = arg[pos] / th - arg[pos] * th
        u[idx, idx] += 2.0 * y14
        y14 = y14 - d1 * diag
        y[idx, idx] += y14
        y, kount = syminv(y)
        if kount > n or kount < 0:
            raise RuntimeError(f"MDPROP: MATRIX INVERSION ERROR AT STEP K = {kstep}")
        nodes += kount
        w_first = _sym_lower(y) if wave is not None else None
        y = u - y23[:, None] * y * y23[None, :]
        y, kount = syminv(y)
        if kount > n or kount < 0:
            raise RuntimeError(f"MDPROP: MATRIX INVERSION ERROR AT STEP K = {kstep}")
        nodes += kount
        if wave is not None:
            wave.append(w_first @ _sym_lower(y23[:, None] * y * y23[None, :]))
        if kstep == nstep:
            d2 = d1
        r = r + h
        u, _ = hammat(r)
        u = d2 * u
        y = u - y23[:, None] * y * y23[None, :]
        y[idx, idx] += y14
    y = _sym_lower(y / h)
    return y, nodes - nsave


def ldprop(z: np.ndarray, hammat, rstart: float, nstep: int, dr: float,
           eshift: Optional[float] = None) -> Tuple[np.ndarray, int]:
    """LDPROP (Johnson log-derivative, IPROP=5): source_code/ldprop.f:24-147.

    With ``eshift`` the matrices are replayed as the reference does for the
    second and later energies of a run with ISCRU (IREAD): ``hammat``
    evaluates W at the first energy, the stored scaled matrix is formed and
    its diagonal shifted afterwards by -ESHIFT, -D4*ESHIFT or -D2*ESHIFT
    (ldprop.f:34-38, 58-63, 107-112), ESHIFT = ERED - EFIRST."""
    n = z.shape[0]
    h = dr / 2.0
    d1 = h * h / 3.0
    d2 = 2.0 * d1
    d4 = -d1 / 16.0
    r = rstart
    nodes = 0
    idx = np.arange(n)
    u, _ = hammat(r)
    if eshift is not None:
        u = u.copy()
        u[idx, idx] -= eshift
    z = h * z + d1 * u
    z[idx, idx] += 1.0
    for istep in range(1, nstep + 1):
        r = r + h
        u, _ = hammat(r)
        u = d4 * u
        u[idx, idx] += 0.125
        if eshift is not None:
            u[idx, idx] += -d4 * eshift
        u, ncu = syminv(u)
        if ncu > n or ncu < 0:
            raise RuntimeError("LDPROP: ERROR IN SYMINV")
        z, ncz = syminv(z)
        if ncz > n or ncz < 0:
            raise RuntimeError("LDPROP: ERROR IN SYMINV")
        nodes += ncz
        z = u - z
        z[idx, idx] -= 6.0
        z, ncz = syminv(z)
        if ncz > n or ncz < 0:
            raise RuntimeError("LDPROP: ERROR IN SYMINV")
        nodes += ncz - ncu
        r = r + h
        if istep == nstep:
            d2 = d1
        u, _ = hammat(r)
        u = d2 * u
        u[idx, idx] += 2.0
        if eshift is not None:
            u[idx, idx] += -d2 * eshift
        z = u - z
    hi = 1.0 / h
    z = _sym_lower(hi * z)
    z[idx, idx] -= hi
    return z, nodes


def maprop(y: np.ndarray, blk: Block, ered: float, rstart: float, nstep: int, dr: float) -> Tuple[np.ndarray, int]:
    """MAPROP (quasi-adiabatic modified log-derivative, IPROP=7):
    source_code/maprop.f:24-186.  Per sector the adiabatic basis T_k of W at
    the sector midpoint (ERED = 0, so energy-independent; DIAGVC ascending),
    the correction C_k = T_{k-1}^T [(W_start - W_mid) DR/6 + Q] T_{k-1} with
    Q = (W_end - W_mid) DR/6 of the previous sector (0 at the first), Y +=
    C_k, Y <- P^T Y P with P = T_{k-1}^T T_k, the half-sector reference
    terms of each adiabat (tan/tanh of sqrt|EIVAL - ERED| DR/2), the
    inversion with its inertia added to NODES unless some adiabat exceeds
    24/DR^2 (NCHECK), and the Y update; at the end Y <- T Y T^T + Q."""
    n = blk.n
    nodes = 0
    r = rstart
    t = np.eye(n)
    q = np.zeros((n, n))
    u, _ = blk.hammat(r, 0.0)
    y = np.array(y, dtype=np.float64)
    idx = np.arange(n)
    for kstep in range(1, nstep + 1):
        r = r + 0.5 * dr
        w, _ = blk.hammat(r, 0.0)
        cr = dr / 6.0
        u = cr * (u - w) + q
        u = t.T @ u @ t
        u = 0.5 * (u + u.T)
        y = y + u
        eival, v = np.linalg.eigh(w)
        q = t.T @ v
        t = v
        y = q.T @ y @ q
        y = 0.5 * (y + y.T)
        wref = eival - ered
        flam = 0.5 * np.sqrt(np.abs(wref))
        wmax = 24.0 / (dr * dr)
        op = wref < 0.0
        tn = np.tan(flam * dr)
        th = np.tanh(flam * dr)
        y1 = np.where(op, flam / tn - flam * tn, flam / th + flam * th)
        y2 = np.where(op, flam / tn + flam * tn, flam / th - flam * th)
        ncheck = int(np.sum(~op & (wref > wmax)))
        y[idx, idx] += y1
        y, ncount = syminv(y)
        if ncount > n:
            raise RuntimeError(f"MATRIX INVERSION ERROR IN MAPROP AT STEP K = {kstep}")
        if dr < 0.0:  # RSTART > RSTOP (maprop.f:145)
            ncount = n - ncount
        if ncheck == 0:
            nodes += ncount
        y = -(y2[:, None] * y * y2[None, :])
        y[idx, idx] += y1
        r = r + 0.5 * dr
        u, _ = blk.hammat(r, 0.0)
        q = cr * (u - w)
    y = t @ y @ t.T
    y = 0.5 * (y + y.T) + q
    return y, nodes


# MGPROP's symplectic schemes (source_code/mgprop.f:103-150, SYMSEL): IMGSEL
# = 4 selects CS4 (5 substeps, the last with A = 0), 5 selects MA5 (6 substeps)
MG_SCHEMES = {
    4: ((0.20517766154229, 0.40302128160421, -0.12092087633891, 0.51272193319241, 0.0),
        (0.061758858135626, 0.33897802655364, 0.61479130717558, -0.14054801465937, 0.12501982279453)),
    5: ((0.33983962583911, -0.088601336903027, 0.58585647682596, -0.60303935653649, 0.32358079655470,
         0.44236379421975),
        (0.11939002928757, 0.69892737038248, -0.17131235827160, 0.40126950225135, 0.010705081848236,
         -0.058979625498031)),
}


def mgprop(y: np.ndarray, hammat, rstart: float, nstep: int, dr: float, imgsel: int = 4) -> Tuple[np.ndarray, int]:
    """MGPROP (symplectic log-derivative, IPROP=8): source_code/mgprop.f:
    20-78.  Per sector of width H = DR and each substep k: Y += B_k H W(R),
    then, for A_k != 0, Y <- (I - (I + A_k H Y)^-1) / (A_k H) with NODES
    +/-= the inertia of I + A_k H Y (the sign of A_k), and R += A_k H; W is
    not re-evaluated at the first substep of a sector when the scheme's last
    A is zero (R has not moved)."""
    a, b = MG_SCHEMES[imgsel]
    m = len(a)
    n = y.shape[0]
    h = dr
    nodes = 0
    idx = np.arange(n)
    u = None
    y = np.array(y, dtype=np.float64)
    for istep in range(1, nstep + 1):
        r = rstart + (istep - 1) * h
        for k in range(m):
            if k != 0 or istep == 1 or a[m - 1] != 0.0:
                u, _ = hammat(r)
            y = y + (b[k] * h) * u
            if a[k] != 0.0:
                y = y * (a[k] * h)
                y[idx, idx] += 1.0
                yi, ncu = syminv(y)
                if ncu > n or ncu < 0:
                    raise RuntimeError("MGPROP: ERROR IN SYMINV")
                nodes += ncu if a[k] >= 0.0 else -ncu
                y = -yi
                y[idx, idx] += 1.0
                y = y * (1.0 / (a[k] * h))
            r = r + a[k] * h
    return _sym_lower(y), nodes


def odprop(y0: float, blk: Block, ered: float, rstart: float, nstep: int, dr: float) -> Tuple[float, int]:
    """ODPROP (single-channel LDMD): source_code/odprop.f (NCONST=NRSQ=0,
    CONLEN=0, no scratch replay).  NODES is always 0 there."""
    npt = nstep + 1
    dr6 = dr / 6.0
    h = dr / 2.0
    u = np.zeros(npt)
    w = np.zeros(npt)
    eintep = blk.eint[0]
    r = rstart
    pts = []
    for i in range(npt):
        pts.append(blk.pvec(r))
        r += dr
    u[:] = eintep
    for i in range(npt):
        u[i] += float(np.dot(blk.vl[:, 0], pts[i]))
    r = rstart
    for i in range(npt):
        rsq = min(1.0 / (r * r), 1.0e16)
        u[i] += blk.cent[0] * rsq
        r += dr
    r = rstart + h
    mids = []
    for i in range(nstep):
        mids.append(blk.pvec(r))
        r += dr
    w[:] = eintep
    for i in range(nstep):
        w[i] += float(np.dot(blk.vl[:, 0], mids[i]))
    # P(:,NPT) keeps the end-point value in the reference (odprop.f loops to NPT)
    w[npt - 1] += float(np.dot(blk.vl[:, 0], pts[npt - 1]))
    r = rstart + h
    for i in range(nstep):
        rsq = min(1.0 / (r * r), 1.0e16)
        w[i] += blk.cent[0] * rsq
        r += dr
    q = np.zeros(npt)
    for i in range(1, npt):
        q[i] = u[i] - w[i - 1]
    qlast = q[npt - 1] * dr6
    for i in range(nstep):
        u[i] = (u[i] - w[i] + q[i]) * dr6
    y = y0
    for i in range(nstep):
        wref = w[i] - ered
        flam = 0.5 * math.sqrt(abs(wref))
        if wref < 0.0:
            tn = math.tan(flam * dr)
            y1 = flam / tn - flam * tn
            y2 = flam / tn + flam * tn
        else:
            tn = math.tanh(flam * dr)
            y1 = flam / tn + flam * tn
            y2 = flam / tn - flam * tn
        y2 = y2 * y2
        y = y1 - y2 / (y + (y1 + u[i]))
    return y + qlast, 0


# ---------------------------------------------------------------------------
# Riccati-Bessel functions (Temme's methods), source_code/rbesjy.f, rbessk.f
# ---------------------------------------------------------------------------
_TEMME_B = (-0.283876542276024, -0.076852840844786, +0.001706305071096, +0.001271927136655,
            +0.000076309597586, -0.000004971736704, -0.000000865920800, -0.000000033126120,
            +0.000000001745136, +0.000000000242310, +0.000000000009161, -0.000000000000170)


def _rgamma(x: float) -> Tuple[float, float, float]:
    """1/Gamma(1-x) with odd/even parts (Temme 1975), source_code/rbesjy.f RGAMMA."""
    x2 = x * x * 8.0
    alfa = -0.000000000000001
    beta = 0.0
    for i in range(12, 1, -2):
        beta = -(2 * alfa + beta)
        alfa = -beta * x2 - alfa + _TEMME_B[i - 1]
    even = (beta / 2.0 + alfa) * x2 - alfa + 0.921870293650453
    alfa = -0.000000000000034
    beta = 0.0
    for i in range(11, 0, -2):
        beta = -(2 * alfa + beta)
        alfa = -beta * x2 - alfa + _TEMME_B[i - 1]
    odd = 2 * (alfa + beta)
    return odd * x + even, odd, even


def _bessjy(v: float, x: float):
    """BESSJY: source_code/rbesjy.f:52-283.  Returns (CJ,DJ,EJ,CY,DY,EY)."""
    eps = 1.0e-15
    maxit = 1000
    pi = math.acos(-1.0)
    xmin = 3.0
    xmax = 5.0 - math.log10(eps)
    na = int(v + 0.5)
    a = v - na
    if x < xmin:
        b = x / 2.0
        d = -math.log(b)
        e = a * d
        c = 1.0 / pi if abs(a) < eps else a / math.sin(a * pi)
        s = 1.0 if abs(e) < eps else math.sinh(e) / e
        e = math.exp(e)
        rg, p, q = _rgamma(a)
        g = e * rg
        e = (e + 1.0 / e) / 2.0
        f = 2 * c * (p * e + q * s * d)
        e = a * a
        p = g * c
        q = 1.0 / g / pi
        c = a * pi / 2.0
        r = 1.0 if abs(c) < eps else math.sin(c) / c
        r = pi * c * r * r
        c = 1.0
        d = -b * b
        ya = f + r * q
        ya1 = p
        ok = False
        for n in range(1, maxit + 1):
            f = (f * n + p + q) / (n * n - e)
            c = c * d / n
            p = p / (n - a)
            q = q / (n + a)
            g = c * (f + r * q)
            h = c * p - n * g
            ya = ya + g
            ya1 = ya1 + h
            dl = abs(g) / (1.0 + abs(ya))
            dl1 = abs(h) / (1.0 + abs(ya1))
            if dl + dl1 < eps:
                ok = True
                break
        if not ok:
            raise RuntimeError("BESSJY 1")
        f = -ya
        g = -ya1 / b
        # d, pa, qa etc. are not set on this branch; the J part below uses
        # CF1 unless X >= max(XMIN,V), which cannot hold here.
        dd = pa = qa = pa1 = qa1 = cc = ss = 0.0
    elif x < xmax:
        c = 0.25 - a * a
        b = x + x
        e = (x * math.cos(a * pi) / pi / eps) ** 2
        p = 1.0
        q = -x
        r = 1.0 + x * x
        s = r
        ok = False
        for n in range(2, maxit + 1):
            d = (n - 1 + c / n) / s
            p = (2 * n - p * d) / (n + 1)
            q = (-b + q * d) / (n + 1)
            s = p * p + q * q
            r = r * s
            if r * n * n > e:
                ok = True
                break
        if not ok:
            raise RuntimeError("BESSJY 2")
        p = p / s
        f = p
        q = -q / s
        g = q
        for m in range(n, 0, -1):
            r = (m + 1) * (2.0 - p) - 2.0
            s = b + (m + 1) * q
            d = (m - 1 + c / m) / (r * r + s * s)
            p = d * r
            q = d * s
            e = f + 1.0
            f = p * e - g * q
            g = q * e + p * g
        f = 1.0 + f
        d = f * f + g * g
        pa = f / d
        qa = -g / d
        d = a + 0.5 - p
        q = q + x
        pa1 = (pa * q - qa * d) / x
        qa1 = (qa * q + pa * d) / x
        b = x - pi * (a + 0.5) / 2.0
        cc = math.cos(b)
        ss = math.sin(b)
        dd = math.sqrt(2.0 / x / pi)
        f = dd * (pa * ss + qa * cc)
        g = dd * (qa1 * ss - pa1 * cc)
    else:
        p = 0.0
        q = 0.0
        pa = qa = 0.0
        for ia in range(2):
            pa = p
            qa = q
            yy = 4.0 * (a + ia) ** 2
            z = 8.0 * x
            d = 0.0
            w = -1.0
            p = 1.0
            q = 0.0
            tp = 1.0
            ok = False
            for _k in range(maxit):
                d = d + z
                w = w + 2.0
                tq = +tp * (yy - w * w) / d
                q = q + tq
                d = d + z
                w = w + 2.0
                tp = -tq * (yy - w * w) / d
                p = p + tp
                if abs(tp) + abs(tq) < eps:
                    ok = True
                    break
            if not ok:
                raise RuntimeError("BESSJY 3")
            p = p - 0.5 * tp
            q = q - 0.5 * tq
        pa1 = p
        qa1 = q
        b = x - pi * (a + 0.5) / 2.0
        cc = math.cos(b)
        ss = math.sin(b)
        dd = math.sqrt(2.0 / x / pi)
        f = dd * (pa * ss + qa * cc)
        g = dd * (qa1 * ss - pa1 * cc)
    # upward recurrence for Y
    p = 0.0
    if na > 0:
        yv = 2.0 / x
        for n in range(1, na + 1):
            h = yv * (a + n) * g - f
            f = g
            g = h
            while abs(f) > 4.0:
                p = p + 1.0
                f = 0.0625 * f
                g = 0.0625 * g
    cy = f
    dy = (v / x) * f - g
    sy = math.sqrt(cy * cy + dy * dy)
    cy = cy / sy
    dy = dy / sy
    ey = math.log(sy) + p * math.log(16.0)
    vv = max(xmin, v)
    if x >= vv:
        f = dd * (pa * cc - qa * ss)
        g = dd * (qa1 * cc + pa1 * ss)
        if na > 0:
            yv = 2.0 / x
            for n in range(1, na + 1):
                h = yv * (a + n) * g - f
                f = g
                g = h
        cj = f
        dj = (v / x) * f - g
        sj = math.sqrt(cj * cj + dj * dj)
        cj = cj / sj
        dj = dj / sj
        ej = math.log(sj)
    else:
        ap = 1.0
        aa = v / x
        bp = 0.0
        bb = 1.0
        f = 0.0
        g = 0.0
        yv = 2.0 / x
        w = yv / pi
        ok = False
        for n in range(1, maxit + 1):
            an = yv * (v + n) * aa - ap
            ap = aa
            aa = an
            bn = yv * (v + n) * bb - bp
            bp = bb
            bb = bn
            if abs(bb) > abs(aa):
                ap = ap / bb
                aa = aa / bb
                bp = bp / bb
                bb = 1.0
                if abs(aa - f) < eps * abs(f):
                    cj = w / (dy - cy * aa)
                    dj = aa * cj
                    ok = True
                    break
                f = aa
            else:
                bp = bp / aa
                bb = bb / aa
                ap = ap / aa
                aa = 1.0
                if abs(bb - g) < eps * abs(g):
                    dj = w / (dy * bb - cy)
                    cj = bb * dj
                    ok = True
                    break
                g = bb
        if not ok:
            raise RuntimeError("BESSJY 4")
        sj = math.sqrt(cj * cj + dj * dj)
        cj = cj / sj
        dj = dj / sj
        ej = math.log(sj) - ey
    return cj, dj, ej, cy, dy, ey


def rbesjy(ell: float, x: float) -> Tuple[float, float, float, float]:
    """Riccati-Bessel j, j', y, y' of fractional order: source_code/rbesjy.f:1-50."""
    if x <= 0.0 or ell < -0.5:
        raise RuntimeError("RBESJY 0")
    v = ell + 0.5
    cj, dj, ej, cy, dy, ey = _bessjy(v, x)
    pi = math.acos(-1.0)
    ex = 0.5 * math.log(pi * x / 2.0)
    dj = dj + cj / (2.0 * x)
    sj = math.sqrt(cj * cj + dj * dj)
    cj = cj / sj
    dj = dj / sj
    ej = ej + math.log(sj) + ex
    dy = dy + cy / (2.0 * x)
    sy = math.sqrt(cy * cy + dy * dy)
    cy = cy / sy
    dy = dy / sy
    ey = ey + math.log(sy) + ex
    return cj * math.exp(ej), dj * math.exp(ej), cy * math.exp(ey), dy * math.exp(ey)


def _mbessk(v: float, x: float):
    """MBESSK: source_code/rbessk.f:41-148.  Returns (CK,DK,EK)."""
    eps = 1.0e-15
    maxit = 1000
    pi = math.acos(-1.0)
    xmin = 1.0
    na = int(v + 0.5)
    a = v - na
    if x < xmin:
        b = x / 2.0
        d = -math.log(b)
        e = a * d
        c = a * pi
        c = 1.0 if abs(c) < eps else c / math.sin(c)
        s = 1.0 if abs(e) < eps else math.sinh(e) / e
        e = math.exp(e)
        rg, p, q = _rgamma(a)
        g = e * rg
        e = (e + 1.0 / e) / 2.0
        f = c * (p * e + q * s * d)
        e = a * a
        p = 0.5 * g * c
        q = 0.5 / g
        c = 1.0
        d = b * b
        ak = f
        ak1 = p
        ok = False
        for n in range(1, maxit + 1):
            f = (f * n + p + q) / (n * n - e)
            c = c * d / n
            p = p / (n - a)
            q = q / (n + a)
            g = c * (p - n * f)
            h = c * f
            ak = ak + h
            ak1 = ak1 + g
            if h / ak + abs(g) / ak1 < eps:
                ok = True
                break
        if not ok:
            raise RuntimeError("MBESSK 1")
        f = ak
        g = ak1 / b
        ex = 0.0
    else:
        c = 0.25 - a * a
        g = 1.0
        f = 0.0
        e = x * math.cos(a * pi) / pi / eps
        ok = False
        for n in range(1, maxit + 1):
            h = (2 * (n + x) * g - (n - 1 + c / n) * f) / (n + 1)
            f = g
            g = h
            if h * n > e:
                ok = True
                break
        if not ok:
            raise RuntimeError("MBESSK 2")
        p = f / g
        q = p
        b = x + x
        e = b - 2.0
        for m in range(n, 0, -1):
            p = (m - 1 + c / m) / (e + (m + 1) * (2.0 - p))
            q = p * (q + 1.0)
        f = math.sqrt(pi / b) / (1.0 + q)
        g = f * (a + x + 0.5 - p) / x
        ex = x
    p = 0.0
    if na > 0:
        yv = 2.0 / x
        for n in range(1, na + 1):
            h = yv * (a + n) * g + f
            f = g
            g = h
            while abs(f) > 4.0:
                p = p + 1.0
                f = 0.0625 * f
                g = 0.0625 * g
    ck = f
    dk = (v / x) * f - g
    sk = math.sqrt(ck * ck + dk * dk)
    ck = ck / sk
    dk = dk / sk
    ek = math.log(sk) + p * math.log(16.0) - ex
    return ck, dk, ek


def rbessk(ell: float, x: float) -> float:
    """Ratio K'/K of the modified Riccati-Bessel function: source_code/rbessk.f:1-39."""
    if x <= 0.0 or ell < -0.5:
        raise RuntimeError("RBESSK 0")
    v = ell + 0.5
    ck, dk, ek = _mbessk(v, x)
    dk = dk + ck / (2.0 * x)
    return dk / ck


# ---------------------------------------------------------------------------
# matching
# ---------------------------------------------------------------------------
def chnsrt(eint: np.ndarray) -> np.ndarray:
    """CHNSRT: source_code/chnsrt.f (exchange sort, 0-based NB)."""
    n = len(eint)
    nb = list(range(n))
    for i in range(n - 1):
        for j in range(i + 1, n):
            if eint[nb[i]] <= eint[nb[j]]:
                continue
            if abs(eint[nb[i]] - eint[nb[j]]) < 1.0e-11:
                continue
            nb[i], nb[j] = nb[j], nb[i]
    return np.array(nb)


def gtwvec(ered: float, eint: np.ndarray) -> Tuple[np.ndarray, int]:
    """GTWVEC: source_code/gtwvec.f:18-30."""
    dif = ered - eint
    wvec = np.sign(dif) * np.sqrt(np.abs(dif))
    wvec = np.where(dif == 0.0, 0.0, wvec)
    nopen = int(np.sum(dif > 0.0))
    return wvec, nopen


def ytok(nb, wvec, n, nopen, y, rval, cent) -> np.ndarray:
    """YTOK + YTOKG + KSYM: source_code/ytok.f:21-48, ytokg.f:21-76, ksym.f."""
    sj = np.zeros(n)
    sjp = np.zeros(n)
    sn = np.zeros(n)
    snp = np.zeros(n)
    for i in range(nopen):
        nx = nb[i]
        dw = wvec[nx]
        darg = dw * rval
        ell = math.sqrt(cent[nx] + 0.25) - 0.5
        uj, ujp, un, unp = rbesjy(ell, darg)
        rdw = math.sqrt(dw)
        sj[nx] = uj / rdw
        sjp[nx] = ujp * rdw
        sn[nx] = un / rdw
        snp[nx] = unp * rdw
    for i in range(nopen, n):
        nx = nb[i]
        dw = abs(wvec[nx])
        darg = dw * rval
        ell = math.sqrt(cent[nx] + 0.25) - 0.5
        ratio = rbessk(ell, darg)
        sn[nx] = 1.0
        snp[nx] = ratio * dw
    ys = _sym_lower(y)  # DSYFIL('U') fills the upper from the lower triangle
    yp = ys[np.ix_(nb, nb)]
    t = yp[:, :nopen] * sj[nb[:nopen]][None, :]
    for i in range(nopen):
        t[i, i] -= sjp[nb[i]]
    km = yp * sn[nb][None, :]
    km[np.arange(n), np.arange(n)] -= snp[nb]
    x = np.linalg.solve(km, t)
    k = x[:nopen, :nopen].copy()
    return 0.5 * (k + k.T)


def ktos(k: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """KTOS: source_code/ktos.f:20-40 (Im S = +2 (I+K^2)^-1 K, as coded)."""
    nop = k.shape[0]
    sr = 0.5 * (k @ k)
    sr[np.arange(nop), np.arange(nop)] += 0.5
    sr, _ = syminv(sr)
    si = sr @ k
    sr = sr.copy()
    sr[np.arange(nop), np.arange(nop)] -= 1.0
    return sr, si


def calca(sr_ii: float, si_ii: float, wvec: float, ll: int, lrpow: int = 6) -> complex:
    """CALCA: source_code/calca.f:42-59."""
    smag = math.sqrt(sr_ii ** 2 + si_ii ** 2)
    if smag > 1.0:
        sr_ii /= smag
        si_ii /= smag
    if sr_ii <= -1.0:
        return complex(1.0e30, 0.0)
    m = min(lrpow - 2, ll + ll + 1)
    return complex(1.0 - sr_ii, -si_ii) / (complex(1.0 + sr_ii, si_ii) * complex(0.0, wvec ** m))


def sigma_block(sr, si, nb, nopen, levidx, jlevels, wvec, jtot, runit, nsig, wght=1.0):
    """SPROC partial cross sections: source_code/sproc.f:380-430 with
    DEGENF (ITYPE=1: 2j_i+1, source_code/base.f:1126-1129).  Returns
    SIGMA(F, I) (0-based levels)."""
    xj = float(2 * jtot + 1) * math.pi
    if wght > 0.0:
        xj *= wght
    sig = np.zeros((nsig, nsig))
    for icol in range(nopen):
        lcol = levidx[nb[icol]]
        for irow in range(nopen):
            lrow = levidx[nb[irow]]
            s_r = sr[irow, icol]
            s_i = si[irow, icol]
            smag = s_r ** 2 + s_i ** 2
            if lrow < nsig and lcol < nsig:
                dd = wvec[nb[irow]] ** 2 / runit ** 2
                if irow == icol:
                    smag = (1.0 - s_r) ** 2 + s_i ** 2
                degen = float(2 * jlevels[lrow] + 1)
                sig[lcol, lrow] += smag * xj / (degen * dd)
    return sig


# ---------------------------------------------------------------------------
# SCCTRL-like single-segment control for one (JTOT, IB) block
# ---------------------------------------------------------------------------
@dataclass
class Grid:
    rstart: float
    rstop: float
    nstep: int
    dr: float


def ld_grid(blk: Block, rmin: float, rmax: float, eredmx: float, steps: float = -10.0,
            drs: float = UNSET, irmset: int = 9) -> Tuple[Grid, float]:
    """Start point (FINDRM) and constant-step grid (DRSET) for an LD segment,
    as set up at the first energy: source_code/mol.driver.f:1490-1550 and
    source_code/scctrl.f:99-108."""
    if irmset > 0:
        rstart, rturn = findrm(blk, rmin, eredmx, irmset)
    else:
        rstart, rturn = rmin, rmin
    cays = calck(0.0, eredmx, blk.eint)
    nstep, dr = drset(rstart, rmax, steps, 0.0, cays, drs, 6, 0.0)
    return Grid(rstart, rmax, nstep, dr), rturn


def ytrans(y: Optional[np.ndarray], blk: Block) -> Tuple[Optional[np.ndarray], np.ndarray, np.ndarray]:
    """YTRANS (source_code/ytrans.f90:79-422) for NCONST > 0, NRSQ = 0 on the
    MOLSCAT path (one operator, channels grouped by L): the operator
    sum_c VL(MXLAM+c) P_c (ytrans.f90:116-140) diagonalised in each set of
    channels of equal L (:147-215, DIAGVC through MULTOP), every eigenvector
    with its largest component (first of equals, IDAMAX) positive
    (:262-275), eigenvector icol of a set at the set's icol-th channel with
    EINT its eigenvalue (:330-350), Y <- EVEC^T Y EVEC (TRNSFM,
    source_code/trnsfm.f:20-32).  Returns (Y or None, EINT, EVEC)."""
    n = blk.n
    if blk.nrsq > 0:
        return ytrans_rsq(y, blk)[:3]
    mx = blk.vl.shape[0] - blk.nconst
    c = np.tensordot(np.asarray(blk.pconst, dtype=np.float64), blk.vl_full[mx:], axes=(0, 0))
    evec = np.zeros((n, n))
    eint = np.zeros(n)
    for ll in sorted(set(int(x) for x in blk.lvals)):
        idx = np.array([i for i in range(n) if int(blk.lvals[i]) == ll])
        w, v = np.linalg.eigh(c[np.ix_(idx, idx)])
        for icol in range(len(idx)):
            col = v[:, icol]
            if col[int(np.argmax(np.abs(col)))] < 0.0:
                col = -col
            evec[idx, idx[icol]] = col
            eint[idx[icol]] = w[icol]
    yt = None if y is None else evec.T @ y @ evec
    return yt, eint, evec


def ytrans_rsq(y: Optional[np.ndarray], blk: Block, degtol: float = 1.0e-10):
    """YTRANS with the centrifugal operator among the terms (NRSQ > 0,
    source_code/ytrans.f90:79-430, MULTOP source_code/multop.f90:55-230):
    no grouping by L; the operators are C = sum_c VL(MXLAM+c) P_c (when
    NCONST > 0) and L^2 = sum_r VL(MXLAM+NCONST+r) (P = 1).  With both: C
    diagonalised (ascending), L^2 turned into its eigenbasis, within each
    set of C eigenvalues differing by at most DEGTOL from their neighbour
    L^2's block diagonalised (ascending; a set of one keeps its diagonal
    element), the total eigenvectors V_C V_L (EVEORD with the identity
    order), EINT = C's eigenvalues; with L^2 alone its eigenvectors and EINT
    unchanged.  Every eigenvector with its largest component (IDAMAX)
    positive; CENT = L^2's eigenvalues (ytrans.f90:353-361,411-430);
    Y <- EVEC^T Y EVEC.  Returns (Y or None, EINT, EVEC, CENT)."""
    n = blk.n
    mx = blk.vl.shape[0] - blk.nconst - blk.nrsq
    l2 = blk.vl_full[mx + blk.nconst:].sum(axis=0)
    if blk.nconst > 0:
        c = np.tensordot(np.asarray(blk.pconst, dtype=np.float64), blk.vl_full[mx:mx + blk.nconst], axes=(0, 0))
        e1, v1 = np.linalg.eigh(c)
        w = v1.T @ l2 @ v1
        e2 = np.empty(n)
        v2 = np.zeros((n, n))
        beg = 0
        while beg < n:
            end = beg + 1
            while end < n and abs(e1[end] - e1[end - 1]) <= degtol:
                end += 1
            ev, vv = np.linalg.eigh(w[beg:end, beg:end])
            e2[beg:end] = ev
            v2[beg:end, beg:end] = vv
            beg = end
        evec = v1 @ v2
        eint = e1
        cent = e2
    else:
        cent, evec = np.linalg.eigh(l2)
        eint = np.array(blk.eint, dtype=np.float64)
    for j in range(n):
        col = evec[:, j]
        if col[int(np.argmax(np.abs(col)))] < 0.0:
            evec[:, j] = -col
    yt = None if y is None else evec.T @ y @ evec
    return yt, eint, evec, cent


def smatrix_ld(blk: Block, ered: float, grid: Grid, iprop: int = 6, imgsel: int = 4, inv: Callable = None,
               efirst: Optional[float] = None):
    """YINIT + MDPROP/LDPROP/ODPROP + (YTRANS when NCONST > 0 or NRSQ > 0, whose
    CENT then go to YTOK) + YTOK + KTOS
    for one energy.  ``inv`` replaces MDPROP's syminv; ``efirst`` (LDPROP)
    replays the first energy's matrices as ldprop.f does with ISCRU.

    Returns dict with Y (final), K, SR, SI (NB-ordered open channels), NB,
    NOPEN, WVEC, NODES."""
    tr = blk.nconst > 0 or blk.nrsq > 0
    eint = ytrans(None, blk)[1] if tr else blk.eint
    cent = ytrans_rsq(None, blk)[3] if blk.nrsq > 0 else blk.cent
    wvec, nopen = gtwvec(ered, eint)
    nb = chnsrt(eint)
    y = yinit(blk, ered, grid.rstart)
    hm = lambda r: blk.hammat(r, ered)
    if iprop == 6 and blk.n > 1:
        y, nodes = mdprop(y, hm, grid.rstart, grid.nstep, grid.dr, inv=inv)
    elif iprop == 6:
        yy, nodes = odprop(float(y[0, 0]), blk, ered, grid.rstart, grid.nstep, grid.dr)
        y = np.array([[yy]])
    elif iprop == 5 and efirst is not None:
        y, nodes = ldprop(y, lambda r: blk.hammat(r, efirst), grid.rstart, grid.nstep, grid.dr,
                          eshift=ered - efirst)
    elif iprop == 5:
        y, nodes = ldprop(y, hm, grid.rstart, grid.nstep, grid.dr)
    elif iprop == 7:
        y, nodes = maprop(y, blk, ered, grid.rstart, grid.nstep, grid.dr)
    elif iprop == 8:
        y, nodes = mgprop(y, hm, grid.rstart, grid.nstep, grid.dr, imgsel)
    else:
        raise ValueError(iprop)
    if tr:
        y = ytrans(y, blk)[0]
    out = dict(Y=y, NB=nb, NOPEN=nopen, WVEC=wvec, NODES=nodes)
    if nopen > 0:
        k = ytok(nb, wvec, blk.n, nopen, y, grid.rstop, cent)
        sr, si = ktos(k)
        out.update(K=k, SR=sr, SI=si)
    return out


# ---------------------------------------------------------------------------
# BOUND matching: source_code/bdctrl.f:76-257 with the two-segment layout of
# source_code/bd.prpini.f (RMID = RMATCH) and source_code/bdpset.f:26-42
# ---------------------------------------------------------------------------
@dataclass
class BoundGrid:
    rmin: float
    rmatch: float
    rmax: float
    nout: int
    drout: float
    nin: int
    drin: float


def bound_grid(rmin: float, rmatch: float, rmax: float, drs: float, drl: float = UNSET) -> BoundGrid:
    """Segments of BDCTRL: outward RMIN -> RMATCH with DRS, inward RMAX ->
    RMATCH with DRL (DRS when unset), each through DRSET with STEP = -10
    (source_code/bdctrl.f:110-111, bd.prpini.f:130-150)."""
    if drl == UNSET:
        drl = drs
    nout, drout = drset(rmin, rmatch, -10.0, 0.0, 0.0, drs, 6, 0.0)
    nin, drin = drset(rmax, rmatch, -10.0, 0.0, 0.0, drl, 6, 0.0)
    return BoundGrid(rmin, rmatch, rmax, nout, drout, nin, drin)


def bound_match(blk: Block, ered: float, g: BoundGrid, iprop_out: int = 6, iprop_in: int = 6):
    """BDCTRL for one energy: YINIT + propagation outward to RMATCH and
    inward to RMATCH, matching matrix M = Y_out - Y_in, its eigenvalues
    (DIAGVC, ascending), NEGCNT = #{eig <= 0}, EIGMIN = the eigenvalue of
    smallest modulus (the last such in ascending order), PSIMCH = its
    eigenvector with the largest component made positive, NODEC = nodes of
    both propagations + NEGCNT (source_code/bdctrl.f:221-257).  iprop 6
    (MDPROP), 5 (LDPROP) or 7 (MAPROP) in either direction.  Returns a
    dict with NODEC, EIGMIN, IMIN (0-based), PSIMCH, EVALS, M."""
    hm = lambda r: blk.hammat(r, ered)
    props = {6: mdprop, 5: ldprop,
             7: lambda y, hm_, r0, ns, dr: maprop(y, blk, ered, r0, ns, dr)}
    y_out = yinit(blk, ered, g.rmin, zout=True)
    y_out, n_out = props[iprop_out](y_out, hm, g.rmin, g.nout, g.drout)
    y_in = yinit(blk, ered, g.rmax, zout=False)
    y_in, n_in = props[iprop_in](y_in, hm, g.rmax, g.nin, g.drin)
    m = y_out - y_in
    ev, vec = np.linalg.eigh(m)
    negcnt = int(np.count_nonzero(ev <= 0.0))
    eigmin, imin = 1.0e30, 0
    for i in range(len(ev)):
        if abs(ev[i]) > abs(eigmin):
            continue
        eigmin, imin = float(ev[i]), i
    psi = vec[:, imin].copy()
    big = 0.0
    for v in psi:
        if abs(v) > abs(big):
            big = v
    psi *= math.copysign(1.0, big)
    return dict(NODEC=n_out + n_in + negcnt, NODES_OUT=n_out, NODES_IN=n_in, NEGCNT=negcnt, EIGMIN=eigmin,
                IMIN=imin, PSIMCH=psi, EVALS=ev, M=m)


def wvbksb(psi_match: np.ndarray, recs: List[np.ndarray]) -> np.ndarray:
    """WVBKSB's back-substitution (source_code/wvbksb.f:64-114): from the
    wavefunction at the matching point, psi at the start of each step from
    psi at its end, psi(A) = W3 psi(B), the records read last to first.
    Returns psi at the segment's NSTEP+1 grid points, [0] at the matching
    point and [NSTEP] at the segment's far end."""
    out = np.empty((len(recs) + 1, len(psi_match)))
    p = np.array(psi_match, dtype=np.float64)
    out[0] = p
    for k, w3 in enumerate(reversed(recs)):
        p = w3 @ p
        out[k + 1] = p
    return out


def simpson_alt_weights(nstep: int) -> np.ndarray:
    """WVBKSB's quadrature weights over NSTEP+1 equally spaced points (in
    units of the step): the alternative extended Simpson rule (Numerical
    Recipes eq. 4.1.14, source_code/wvbksb.f:76-82,139-170): 17/48, 59/48,
    43/48, 49/48 at both ends, 1 inside, filled from both ends towards the
    middle with the central point counted once."""
    w = np.zeros(nstep + 1)
    ends = [17.0 / 48.0, 59.0 / 48.0, 43.0 / 48.0, 49.0 / 48.0]
    for istep in range(1, (nstep + 2) // 2 + 1):
        right, left = istep - 1, nstep + 1 - istep
        a = ends[istep - 1] if istep <= 4 else 1.0
        w[right] += a
        if left != right:
            w[left] += a
    return w


def bound_wavefunction(blk: Block, ered: float, g: BoundGrid) -> dict:
    """BDCTRL's wavefunction (source_code/bdctrl.f:266-360) for MDPROP in
    both directions: the WAVE records of the outward and inward
    propagations, PSIMCH (bound_match) carried back to RMIN through the
    outward records and out to RMAX through the inward ones (wvbksb), the
    normalisation integral SUMPSI of each channel from both segments
    (simpson_alt_weights times |DR|), and the wavefunction from RMIN to RMAX
    divided by sqrt(sum SUMPSI).  Returns R, PSI (normalised), SUMPSI and
    the bound_match dict."""
    hm = lambda r: blk.hammat(r, ered)
    rec_out: list = []
    rec_in: list = []
    y_out = yinit(blk, ered, g.rmin, zout=True)
    y_out, n_out = mdprop(y_out, hm, g.rmin, g.nout, g.drout, wave=rec_out)
    y_in = yinit(blk, ered, g.rmax, zout=False)
    y_in, n_in = mdprop(y_in, hm, g.rmax, g.nin, g.drin, wave=rec_in)
    m = bound_match(blk, ered, g)
    psi0 = m["PSIMCH"]
    p_out = wvbksb(psi0, rec_out)  # [0] at RMATCH ... [nout] at RMIN
    p_in = wvbksb(psi0, rec_in)  # [0] at RMATCH ... [nin] at RMAX
    sumpsi = (simpson_alt_weights(g.nout)[:, None] * p_out ** 2).sum(0) * abs(g.drout) + \
        (simpson_alt_weights(g.nin)[:, None] * p_in ** 2).sum(0) * abs(g.drin)
    r_out = g.rmatch - np.arange(g.nout + 1) * g.drout
    r_in = g.rmatch - np.arange(g.nin + 1) * g.drin
    r = np.concatenate([r_out[::-1], r_in[1:]])
    psi = np.concatenate([p_out[::-1], p_in[1:]]) / math.sqrt(float(sumpsi.sum()))
    return dict(R=r, PSI=psi, SUMPSI=sumpsi, MATCH=m)


# ---------------------------------------------------------------------------
# AIRY propagator (AIPROP, IPROP=9): source_code/aiprop.f:1-392 with POTENT
# (source_code/potent.f:56-119), SPROPN (source_code/spropn.f:82-258) and the
# step sequence of DRSET/DRCALC for TOLHI = 0.  The single-channel Airy
# imbedding propagator is evaluated here from its definition with mpmath at
# 40 digits (independent of the reference's SCAIRY tables and asymptotic
# series): for the linear reference potential x = alpha*(beta + r) of a
# sector with ends x1, x2 and D = Ai(x1)Bi(x2) - Ai(x2)Bi(x1),
#   y1 = alpha (Bi'(x1)Ai(x2) - Ai'(x1)Bi(x2)) / D,
#   y2 = alpha / (pi D),
#   y4 = alpha (Ai(x1)Bi'(x2) - Bi(x1)Ai'(x2)) / D,
# which is what SPROPN's modulus-phase and asymptotic-product forms compute.
# ---------------------------------------------------------------------------
AIRY_SQ3 = 1.732050807  # the sqrt(3) of source_code/potent.f:49 (as written there)


def airy_steps(rstart: float, rstop: float, nstep: int, dr0: float, powr: float) -> List[Tuple[float, float]]:
    """(RNOW, DRNOW) of each AIPROP step for TOLHI = 0, accumulated exactly as
    source_code/aiprop.f:86-91,288-327 does (RNEXT += DRNOW, last step ends
    at RSTOP)."""
    out = []
    drnow = dr0
    rlast = rstart
    rnow = rlast + drnow * 0.5
    rnext = rlast + drnow
    last = False
    kstep = 0
    while True:
        kstep += 1
        out.append((rnow, drnow))
        if last:
            break
        drnow = drcalc(rstart, rstop, kstep, nstep, powr)
        rlast = rnext
        rnext = rnext + drnow
        last = (nstep - kstep) <= 1
        if last:
            rnext = rstop
            drnow = rnext - rlast
        rnew = rlast + 0.5 * drnow
        rnow = rnew
    return out


def airy_sector(width: float, eig: np.ndarray, hp: np.ndarray):
    """y1, y2, y4 per channel for a sector of the given width whose local
    wavevector-squared eigenvalues are eig with slopes hp (SPROPN's inputs:
    the reference W is -eig + hp*r about the sector midpoint, so x =
    alpha*(r + beta) with alpha^3 = hp and beta = -eig/hp)."""
    import mpmath as mp

    mp.mp.dps = 40
    n = len(eig)
    y1 = np.empty(n)
    y2 = np.empty(n)
    y4 = np.empty(n)
    for i in range(n):
        slope = hp[i] if hp[i] != 0.0 else 1.0e-30
        alpha = math.copysign(abs(slope) ** (1.0 / 3.0), slope)
        beta = -eig[i] / slope
        x1 = mp.mpf(alpha) * (mp.mpf(beta) - mp.mpf(width) / 2)
        x2 = mp.mpf(alpha) * (mp.mpf(beta) + mp.mpf(width) / 2)
        a1, b1 = mp.airyai(x1), mp.airybi(x1)
        a2, b2 = mp.airyai(x2), mp.airybi(x2)
        ap1, bp1 = mp.airyai(x1, 1), mp.airybi(x1, 1)
        ap2, bp2 = mp.airyai(x2, 1), mp.airybi(x2, 1)
        d = a1 * b2 - a2 * b1
        y1[i] = float(alpha * (bp1 * a2 - ap1 * b2) / d)
        y2[i] = float(alpha / (mp.pi * d))
        y4[i] = float(alpha * (a1 * bp2 - b1 * ap2) / d)
    return y1, y2, y4


def airy_potent(blk: Block, ered: float, rnow: float, drnow: float):
    """POTENT (source_code/potent.f:56-119): W at the two Gauss points, the
    local wavevector matrix -(W(RA)+W(RB))/2 diagonalised (DIAGVC,
    ascending), and the diagonal of sqrt(3)(W(RB)-W(RA))/DRNOW in the local
    basis.  Returns (eig, V, hp) with V's columns the local basis."""
    ra = rnow - 0.5 * drnow / AIRY_SQ3
    rb = rnow + 0.5 * drnow / AIRY_SQ3
    wa, _ = blk.hammat(ra, ered)
    wb, _ = blk.hammat(rb, ered)
    wavg = -0.5 * (wa + wb)
    wd = (AIRY_SQ3 / drnow) * (wb - wa)
    eig, v = np.linalg.eigh(wavg)
    hp = np.einsum("ji,jk,ki->i", v, wd, v)
    return eig, v, hp


def aiprop(z: np.ndarray, blk: Block, ered: float, rstart: float, rstop: float, nstep: int, dr0: float,
           powr: float, inv: Callable = None, sector: Callable = None,
           limit: Optional[int] = None) -> Tuple[np.ndarray, int]:
    """AIPROP for TOLHI = 0 (source_code/aiprop.f:86-327): per step the local
    basis of POTENT, Z <- T^T Z T with T = V_prev^T V_now, the Airy sector
    propagator, Z <- Y4 - Y2 (Z + Y1)^-1 Y2 with NODES += KOUNT (outward) or
    NCH - KOUNT (inward); finally back to the primitive basis.  ``inv``
    replaces syminv and ``sector`` airy_sector; ``limit`` stops after that
    many steps (the CPU baseline's sample)."""
    inv = inv if inv is not None else syminv
    sector = sector if sector is not None else airy_sector
    n = blk.n
    nodes = 0
    vprev = None
    for kstep, (rnow, drnow) in enumerate(airy_steps(rstart, rstop, nstep, dr0, powr)):
        if limit is not None and kstep >= limit:
            break
        eig, v, hp = airy_potent(blk, ered, rnow, drnow)
        t = v if vprev is None else vprev.T @ v
        z = t.T @ z @ t
        y1, y2, y4 = sector(drnow, eig, hp)
        zz = z + np.diag(y1)
        zi, kount = inv(zz)
        if kount > n:
            raise RuntimeError("AIPROP: instability in SYMINV")
        nodes += kount if rstop > rstart else n - kount
        z = -(y2[:, None] * zi * y2[None, :]) + np.diag(y4)
        vprev = v
    z = vprev @ z @ vprev.T
    return 0.5 * (z + z.T), nodes


def aiprop_tolhi(z: np.ndarray, blk: Block, ered: float, rstart: float, rstop: float, dr0: float, powr: float,
                 tolhi: float, replay: Optional[List[Tuple[float, float]]] = None):
    """AIPROP with TOLHI > 0 (source_code/aiprop.f:86-327): the steps chosen
    as the propagation goes.  EIGOLD starts as the eigenvalues of -W(RSTART)
    (WAVEIG, source_code/waveig.f:38-46); per step POTENT, CORR's estimates
    (source_code/corr.f:43-60: CDIAG from the mean |W''| implied by the
    eigenvalue change and slope, COFF from the mean local wavevector times
    XLARGE, the largest off-diagonal element of the derivative matrix in the
    local basis, potent.f:108-116), the step as for TOLHI = 0, then DRNOW <-
    DRNOW * min(2, (TOLHI / max(CDIAG, COFF))^(1/POWR)) with the last step
    cut at RSTOP (aiprop.f:288-316).  With replay (the (RNOW, DRNOW) of an
    earlier energy) the steps are those, as with the reference's scratch
    file for the energies after the first.  Returns (Z at RSTOP in the
    primitive basis, NODES, the steps taken)."""
    n = blk.n
    nodes = 0
    eigold = np.linalg.eigvalsh(-blk.hammat(rstart, ered)[0])
    drnow = dr0
    drmid = drnow * 0.5
    rlast = rstart
    rnow = rlast + drmid
    rnext = rlast + drnow
    last = False
    vprev = None
    steps: List[Tuple[float, float]] = []
    k = 0
    while True:
        if replay is not None:
            rnow, drnow = replay[k]
            k += 1
            last = k == len(replay)
        ra = rnow - 0.5 * drnow / AIRY_SQ3
        rb = rnow + 0.5 * drnow / AIRY_SQ3
        wa, _ = blk.hammat(ra, ered)
        wb, _ = blk.hammat(rb, ered)
        eig, v = np.linalg.eigh(-0.5 * (wa + wb))
        wdl = v.T @ ((AIRY_SQ3 / drnow) * (wb - wa)) @ v
        hp = np.diag(wdl).copy()
        xlarge = float(np.abs(np.tril(wdl, -1)).max()) if n > 1 else 0.0
        fac = 2.0 / (drmid ** 2)
        w2p = -fac * (eig - eigold + drmid * hp)
        cdiag = abs(drnow ** 3) * (np.sum(np.abs(w2p)) / n) / 12.0
        cay = np.sum(np.sqrt(np.abs(eig))) / n
        coff = cay * xlarge * abs(drnow ** 3) / 12.0
        eigold = eig
        t = v if vprev is None else vprev.T @ v
        z = t.T @ z @ t
        z = 0.5 * (z + z.T)
        y1, y2, y4 = airy_sector(drnow, eig, hp)
        zi, kount = syminv(z + np.diag(y1))
        if kount > n:
            raise RuntimeError("AIPROP: instability in SYMINV")
        nodes += kount if rstop > rstart else n - kount
        z = -(y2[:, None] * zi * y2[None, :]) + np.diag(y4)
        vprev = v
        steps.append((rnow, drnow))
        if last:
            break
        if replay is not None:
            continue
        cmax = max(abs(cdiag), abs(coff))
        factor = (tolhi / cmax) ** (1.0 / powr) if cmax > 0.0 else math.inf
        if drnow < 0.0 and rnext + drnow > 0.0:
            factor = factor * (1.0 + drnow / rnext) ** 3
        factor = min(factor, 2.0)
        drnow = drnow * factor
        rlast = rnext
        rnext = rnext + drnow
        last = (rstop - rstart) * (rstop - rnext) < 0.0
        if last:
            rnext = rstop
            drnow = rnext - rlast
        rnew = rlast + 0.5 * drnow
        drmid = rnew - rnow
        rnow = rnew
        if len(steps) >= 500000:
            raise RuntimeError("AIRY PROPAGATION NOT FINISHED IN 500000 STEPS")
    z = vprev @ z @ vprev.T
    return 0.5 * (z + z.T), nodes, steps


def smatrix_md_airy(blk: Block, ered: float, ld: Grid, rmax: float, dr0: float, tolhi: float = 1.0e-4,
                    powr: float = 3.0, replay: Optional[List[Tuple[float, float]]] = None):
    """The two-segment run IPROPS = 6, IPROPL = 9 with TOLHIL > 0
    (source_code/scctrl.f:99-208): YINIT and MDPROP on the constant-step
    grid RSTART -> RMID, AIPROP with its own step control RMID -> RMAX (the
    first step DRSET's DRFIT, source_code/drset.f:41-84; POWRL = POWRX = 3,
    source_code/mol.prpini.f:108-116), YTOK + KTOS.  Returns the smatrix_ld
    dict plus STEPS (the AIRY steps; pass them as replay for a later
    energy)."""
    wvec, nopen = gtwvec(ered, blk.eint)
    nb = chnsrt(blk.eint)
    y = yinit(blk, ered, ld.rstart)
    y, nodes = mdprop(y, lambda r: blk.hammat(r, ered), ld.rstart, ld.nstep, ld.dr)
    y, n2, steps = aiprop_tolhi(y, blk, ered, ld.rstop, rmax, dr0, powr, tolhi, replay=replay)
    out = dict(Y=y, NB=nb, NOPEN=nopen, WVEC=wvec, NODES=nodes + n2, STEPS=steps)
    if nopen > 0:
        k = ytok(nb, wvec, blk.n, nopen, y, rmax, blk.cent)
        sr, si = ktos(k)
        out.update(K=k, SR=sr, SI=si)
    return out


# ---------------------------------------------------------------------------
# VIVS (IPROP = 4): source_code/vvprop.f with PERT2 (source_code/pert2.f) and
# DELRD (source_code/delrd.f), the options of the reference's defaults
# (source_code/mol.driver.f:277-349): IV = .TRUE. (constant perturbation
# only: no DERMAT, IVP = IVPP = ISHIFT = IDIAG = .FALSE.), IPERT = .TRUE.,
# ISYM = .TRUE., IALPHA = 6 steps per interval in geometric progression
# with ALPHA running from ALPHA1 = 1 to ALPHA2 = 1.5 (IALFP = .FALSE.),
# XSQMAX = 1e4.
# ---------------------------------------------------------------------------
VIVS_IALPHA = 6
VIVS_ALPHA1 = 1.0
VIVS_ALPHA2 = 1.5
VIVS_XSQMAX = 1.0e4
VIVS_DRMAX = 5.0


@dataclass
class VivsOptions:
    """The /LDVVCM/ controls of VVPROP (source_code/vvprop.f:57-82), with
    the reference's defaults (source_code/mol.driver.f:277-349).  IDIAG as
    MOL.PRPINI leaves it: IDIAG itself only adds the diagonal terms (IVD,
    IVPD, IVPPD); the driver's IDIAG default also turns on IV, IVP, IVPP,
    ISHIFT and IPERT (source_code/mol.prpini.f:75-83, vivs_idiag_defaults)."""
    iv: bool = True
    ivp: bool = False
    ivpp: bool = False
    ishift: bool = False
    idiag: bool = False
    ipert: bool = True
    isym: bool = True
    ialpha: int = VIVS_IALPHA
    ialfp: bool = False
    alpha1: float = VIVS_ALPHA1
    alpha2: float = VIVS_ALPHA2
    xsqmax: float = VIVS_XSQMAX


def vivs_idiag_defaults(o: VivsOptions) -> VivsOptions:
    """MOL.PRPINI for IPROPL = 4 (source_code/mol.prpini.f:75-83): IDIAG
    switches on IV, IVP, IVPP, ISHIFT and IPERT."""
    if o.idiag:
        o = VivsOptions(**{**o.__dict__, "iv": True, "ivp": True, "ivpp": True, "ishift": True, "ipert": True})
    return o


def vivs_tolerances(tolhi: float) -> Tuple[float, float]:
    """TLDIAG and TOFF from the segment's TOLHI (source_code/scctrl.f:157-158)."""
    tldiag = 0.064 * math.sqrt(tolhi / 0.001)
    return tldiag, tldiag


def delrd(dr: float, cdiag: float, coff: float, tol: float, drmax: float, e1: float, en: float, rnow: float,
          rmax: float) -> float:
    """DELRD: source_code/delrd.f:1-60 (new interval size from the largest
    diagonal and off-diagonal corrections)."""
    rtol = 0.8 * tol
    dfact = 2.5 if cdiag == 0.0 else (rtol / cdiag) ** 0.333
    ofact = 2.5 if coff == 0.0 else (rtol / coff) ** 0.333
    factor = min(dfact, ofact)
    if en > 0.0:
        factor = min(factor, 1.6 if rnow <= 0.1 * rmax else 2.5)
    elif e1 > 0.0:
        factor = min(factor, 1.20)
    else:
        factor = min(factor, 1.15)
    dr = dr * factor
    if en < 0.0:
        dr = min(dr, 4.0 / math.sqrt(-en))
    dr = min(dr, drmax)
    if dr < 1.0e-6 * drmax:
        raise RuntimeError("DELRD: STEP SIZE TOO SMALL")
    return dr


def pert2(cx, sx, xsq, x, h, v0, a1, a2, a1p, a2p, eye11, eye12, eye22, v1=None, v2=None, pert1=False,
          ref_sign_150=False):
    """PERT2 (source_code/pert2.f:1-230), or PERT1 (source_code/pert1.f:1-224)
    with pert1=True: the step integrals CC, CS, SS of the perturbation
    V0 + V1 (r - h/2) + V2 (r - h/2)^2 / 2 (r measured from the step start;
    V1, V2 the first and second derivatives of W at the step centre, DERMAT)
    between the step's reference solutions, and their accumulation into
    EYE11, EYE12, EYE22 with the step-start values A1, A1P (regular) and A2,
    A2P (irregular).  PERT1 keeps V2 on the diagonal only (pert1.f:160-187).
    v1 = v2 = None: the V1 and V2 terms are exact zeros, as PERT2 gives them
    for V1 = V2 = 0.  Returns (CC, CS, SS); the EYE are updated in place.
    Vectorised over the pairs I <= J; each pair takes PERT2's branch:
    |XI - XJ| > TOL (1e-3) or one open and one closed channel: the closed
    form; else |XSQI + XSQJ| < TOL: the series about XSQ = 0; else the
    series about XI = XJ (both open or both closed).

    The series about XSQ = 0 gives R2V1 as HC (1/12 + ...) (XSQI - XSQJ)
    (pert2.f:175, and pert1.f alike); its leading term makes the V1 parts
    of CS and CS' come out as -h^3/12 where the integrals of c_i s_j (r -
    h/2) tend to +h^3/12 (both solutions near threshold), i.e. the
    difference enters with the wrong sign.  This restatement takes the sign
    that matches the integrals (checked against quadrature); ref_sign_150 =
    True keeps the reference's expression as written."""
    n = len(x)
    hsq = h * h
    hinv = 1.0 / h
    hsqinv = 1.0 / hsq
    hf = h ** 3
    iu, ju = np.triu_indices(n)
    xi, xj = x[iu], x[ju]
    sxi, sxj = sx[iu], sx[ju]
    cxi, cxj = cx[iu], cx[ju]
    xsqi, xsqj = xsq[iu], xsq[ju]
    xsxi = xi * sxi
    xsxj = xsxj0 = xj * sxj
    diag = iu == ju
    f5 = xsqi * xsqj
    f6 = xsxi * xsxj
    f7 = cxi * cxj
    sd = xsqi - xsqj
    total = xsqi + xsqj
    coef0 = hf / (2.0 * f6)
    coef = coef0 * v0[iu, ju]
    use1 = v1 is not None
    use2 = v2 is not None
    hc = h * coef0 * v1[iu, ju] if use1 else None
    hhc = hsq * coef0 * v2[iu, ju] if use2 else None
    zero = np.zeros(len(iu))
    d = xi - xj
    tol = 1.0e-3
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        near = np.abs(d) <= tol
        b150 = near & (np.abs(total) < tol)
        bothc = (xsqi < 0.0) & (xsqj < 0.0)
        b140 = near & ~b150 & (bothc | (f5 >= 0.0))
        b130 = ~(b150 | b140)
        # closed form (label 130)
        f25 = 1.0 / sd
        f8 = xsxj * cxi
        f9 = xsxi * cxj
        f12 = 2.0 * coef * f25
        r1_130 = f12 * (f9 - f8)
        r2_130 = f12 * (xsxi - xsxj)
        f10 = sd * sd
        f18 = 2.0 * total
        if use1:
            f13 = hc * f25
            r1v1_130 = f13 * ((f18 * (f7 - 1.0) + 4.0 * f6) * f25 + f9 - f8)
            r2v1_130 = f13 * (f18 * (cxi - cxj) * f25 + xsxi + xsxj)
        if use2:
            f19 = 0.25 * f25
            f24 = -8.0 * (total + 2.0 * xsqj)
            f26 = -8.0 * (total + 2.0 * xsqi)
            f31 = hhc / f10
            r1v2_130 = f31 * (((f10 + f24) * f9 - (f10 + f26) * f8) * f19 + total * (f7 + 1.0) + 2.0 * f6)
            r2v2_130 = f31 * (((f10 + f24) * xsxi - (f10 + f26) * xsxj) * f19 + total * (cxi + cxj))
        # series about XI = XJ (labels 110/120/140)
        s = xi + xj
        ssq = s * s
        dsq = d * d
        shalf = s * 0.5
        c2 = np.where(bothc, np.cosh(shalf), np.cos(shalf))
        s2 = np.where(bothc, -np.sinh(shalf), np.sin(shalf))
        ssq = np.where(bothc, -ssq, ssq)
        dsq = np.where(bothc, -dsq, dsq)
        ss2 = s * s2
        f14 = ss2 / ssq
        f16 = dsq * (1.0 / 24.0)
        f17 = dsq / 80.0
        f15 = ss2 * c2
        f36 = dsq * (1.0 / 168.0)
        f37 = dsq * (1.0 / 112.0)
        f38 = dsq * (1.0 / 40.0)
        f39 = 1.0 - dsq * (1.0 / 8.0) * (1.0 - dsq * (1.0 / 48.0) * (1.0 - dsq * (1.0 / 120.0)))
        f40 = 1.0 / ssq
        r1_140 = coef * (1.0 - dsq * (1.0 / 6.0) * (1.0 - dsq / 20.0 * (1.0 - dsq * (1.0 / 42.0))) + 2.0 * f15 * f40)
        r2_140 = coef * (f14 * f39 + 0.5 * c2 * (1.0 - f16 * (1.0 - f17 * (1.0 - f36)))) * 2.0
        if use1:
            r1v1_140 = hc * ((-2.0 * ss2 * f14 + f15) * f40 - f16 * (1.0 - dsq * (1.0 / 15.0) * (1.0 - 3.0 * f37)))
            r2v1_140 = hc * (-s2 * (1.0 / 12.0) * (1.0 - f38 * (1.0 - f37))
                             + (0.5 * c2 - f14) / s * (1.0 - f16 * (1.0 - f17))) * d
        if use2:
            f41 = (ssq - 8.0) * 0.5 * f14
            r1v2_140 = hhc * (c2 * (2.0 * c2 + f41) * f40
                              + (1.0 - dsq * 0.2 * (1.0 - 11.0 * f36 * (1.0 - dsq * (1.0 / 36.0)))) * (1.0 / 12.0)) * 0.5
            r2v2_140 = hhc * (f39 * (c2 + f41 * 0.5) * f40
                              + (1.0 - f38 * (3.0 - f37 * (5.0 - 7.0 * dsq * (1.0 / 216.0)))) * c2 * (1.0 / 24.0))
        # series about XSQ = 0 (label 150)
        g3 = xsqi * xsqi
        f20 = total * total
        f22 = 4.0 * xsqi * xsqj
        f23 = xsqj * xsqj
        f29 = total * (1.0 / 6.0)
        f28 = 14.0 * f5
        f30 = coef * 2.0
        r1_150 = f30 * (1.0 - f29 + (f20 + f22) * (1.0 / 120.0) - total * (g3 + f28 + f23) * (1.0 / 5040.0))
        r2_150 = f30 * (1.0 - f29 + (xsqi * total + f23) * (1.0 / 120.0) - (g3 + f23) * total * (1.0 / 5040.0))
        if use1:
            r1v1_150 = -hc * (1.0 / 12.0) * (total - (f20 + f22) * (1.0 / 15.0)
                                             + total * (g3 + f28 + f23) * (1.0 / 560.0))
            r2v1_150 = hc * (1.0 / 12.0 + total * (1.0 / 180.0)
                             - (3.0 * f20 - 2.0 * xsqi * xsqj) * (1.0 / 20160.0)) * (sd if ref_sign_150 else -sd)
        if use2:
            f21 = total * 0.2
            g7 = g3 * xsqi
            r1v2_150 = hhc * (1.0 - f21 + 11.0 * (f20 + f22) * (1.0 / 840.0)
                              - 11.0 * total * (g3 + f28 + f23) * (1.0 / 30240.0)) * (1.0 / 12.0)
            r2v2_150 = hhc * (1.0 - f21 + (11.0 * f20 - 19.0 * f5) * (1.0 / 840.0)
                              - (11.0 * (g7 + f23 * xsqj) + 3.0 * f5 * total) * (1.0 / 30240.0)) * (1.0 / 12.0)
        # diagonal (label 160)
        g1 = xsxi / xsqi
        r1_d = coef * (1.0 + cxi * g1)
        r2_d = coef * (g1 + cxi)
        if use1:
            r1v1_d = hc * (cxi - g1) * (0.5 * g1)
        if use2:
            r1v2_d = hhc * (cxi * (cxi + (xsqi - 2.0) * (0.5 * g1)) / xsqi * 0.25 + 1.0 / 24.0)
            r2v2_d = hhc * ((6.0 + xsqi) * cxi * (1.0 / 3.0) + (xsqi - 2.0) * g1) / (8.0 * xsqi)

    def pick(dv, v130, v140, v150):
        return np.where(diag, dv, np.where(b130, v130, np.where(b140, v140, v150)))

    r1 = pick(r1_d, r1_130, r1_140, r1_150)
    r2 = pick(r2_d, r2_130, r2_140, r2_150)
    r1v1 = pick(r1v1_d, r1v1_130, r1v1_140, r1v1_150) if use1 else zero
    r2v1 = pick(zero, r2v1_130, r2v1_140, r2v1_150) if use1 else zero
    if use2:
        r1v2 = pick(r1v2_d, r1v2_130, r1v2_140, r1v2_150)
        r2v2 = pick(r2v2_d, r2v2_130, r2v2_140, r2v2_150)
        if pert1:
            # PERT1: the second derivative on the diagonal only
            r1v2 = np.where(diag, r1v2, 0.0)
            r2v2 = np.where(diag, r2v2, 0.0)
    else:
        r1v2 = r2v2 = zero
    # label 170 (pert2.f:199-211; pert1.f:180-200 with R1V2 = R2V2 = 0 off the diagonal)
    r1s = r1 + r1v1 + r1v2
    r2s = r2 + r2v1 + r2v2
    term1 = -cxj * r1s
    ccij = f6 * r1s * hsqinv
    csij = xsxi * (r2s + term1) * hinv
    ssij = r1 - r1v1 + r1v2 - cxj * (r2 - r2v1 + r2v2) - cxi * r2s + f7 * r1s
    csji = xsxj0 * (r2 - r2v1 + r2v2 - cxi * r1s) * hinv
    cc = np.zeros((n, n))
    cs = np.zeros((n, n))
    ss = np.zeros((n, n))
    cc[iu, ju] = ccij
    cc[ju, iu] = ccij
    cs[iu, ju] = csij
    cs[ju, iu] = csji
    ss[iu, ju] = ssij
    ss[ju, iu] = ssij
    t1 = a1p[ju] * ssij + a1[ju] * csji
    t2 = a1p[ju] * csij + a1[ju] * ccij
    t3 = a2p[ju] * ssij + a2[ju] * csji
    t4 = a2p[ju] * csij + a2[ju] * ccij
    e11 = a1p[iu] * t1 + a1[iu] * t2 + eye11[iu, ju]
    e22 = a2p[iu] * t3 + a2[iu] * t4 + eye22[iu, ju]
    e12 = a1p[iu] * t3 + a1[iu] * t4 + eye12[iu, ju]
    off = ~diag
    e21 = a2p[iu[off]] * t1[off] + a2[iu[off]] * t2[off] + eye12[ju[off], iu[off]]
    eye11[iu, ju] = e11
    eye11[ju, iu] = e11
    eye22[iu, ju] = e22
    eye22[ju, iu] = e22
    eye12[iu, ju] = e12
    eye12[ju[off], iu[off]] = e21
    return cc, cs, ss


def pert2_v0(cx, sx, xsq, x, h, v0, a1, a2, a1p, a2p, eye11, eye12, eye22):
    """PERT2 with V1 = V2 = 0 (the default IV option): pert2 without its V1
    and V2 terms."""
    return pert2(cx, sx, xsq, x, h, v0, a1, a2, a1p, a2p, eye11, eye12, eye22)


@dataclass
class VivsStep:
    """One VVPROP step as the reference's scratch file records it for the
    energies after the first (source_code/vvprop.f:350-352): its end RNOW,
    width DRNOW, whether it opens an interval (ICRMAT, with the interval's
    diagonalisation point RMIDI) and whether it is the last."""
    rnow: float
    drnow: float
    newint: bool
    rmidi: float
    last: bool
    endint: bool = False  # the step ends an interval (NEWINT on the scratch file, vvprop.f:590-593)


def vvprop(rmat: np.ndarray, blk: Block, ered: float, rstart: float, rstop: float, drnow: float,
           tldiag: float, toff: float, drmax: float = VIVS_DRMAX,
           replay: Optional[List[VivsStep]] = None,
           opts: Optional[VivsOptions] = None) -> Tuple[np.ndarray, List[VivsStep]]:
    """VVPROP (source_code/vvprop.f:1-750) with the /LDVVCM/ controls of
    opts (the reference's defaults when None: IV, IPERT, ISYM, IALPHA = 6,
    IALFP = .FALSE.): propagates the R-matrix (the inverse of Y) from RSTART
    to RSTOP.  Per step: W at the step centre in the interval's local basis
    (the eigenvectors of W at the interval's centre RMIDI, ascending,
    DIAGVC), its diagonal the reference potential and its off-diagonal part
    the perturbation; with IVP / IVPP (or IDIAG) DERMAT's first / second
    derivatives of W at the step centre, in the same basis, add the
    perturbations V1 (r - h/2) and V2 (r - h/2)^2/2; ISHIFT moves the
    reference potential by h^2/24 times the second derivative's diagonal
    (vvprop.f:238-330).  Sine/cosine (sinh/cosh) reference solutions G1
    (regular) and G2 (irregular) of each channel; PERT2's (or PERT1's) step
    integrals accumulated over the interval.  At each interval end the
    R-matrix R <- (G1' + D1' + R (G2' + D2'))^-1 (G1 + D1 + R (G2 + D2)),
    symmetrised under ISYM, D the first-order corrections (zero without
    IPERT; DSYMM + DGESV, vvprop.f:510-618).  Intervals of IALPHA steps in
    geometric progression (ALPHA predicted from the diagonal corrections'
    trend under IALFP), or for IALPHA <= 0 steps from DELRD with an interval
    ended when the corrections' estimates grow (vvprop.f:480-500); the next
    interval size from DELRD.  With replay (the steps of an earlier energy),
    the steps and interval ends are the recorded ones, as with the
    reference's scratch file (the energies after the first, vvprop.f:155-160,
    410, 590).  Returns (R at RSTOP in the primitive basis, the steps
    taken)."""
    o = VivsOptions() if opts is None else opts
    n = blk.n
    rmat = np.array(rmat, dtype=np.float64)
    ialpha = o.ialpha
    ivd = o.iv or o.idiag
    ivpd = o.ivp or o.idiag
    ivppd = o.ivpp or o.idiag
    need_w2 = ivppd or o.ishift
    do_pert = o.ivpp or o.ivp or o.iv
    pert1 = o.ivp and not o.ivpp
    g1 = np.zeros(n)
    g1p = np.ones(n)
    g2 = np.ones(n)
    g2p = np.zeros(n)
    vecnew = np.eye(n)
    eye11 = np.zeros((n, n))
    eye12 = np.zeros((n, n))
    eye22 = np.zeros((n, n))
    steps: List[VivsStep] = []
    icrmat = True
    last = False
    coffl = 0.0
    soff1 = sdiag1 = 0.0
    istep = 1
    drint = drnow
    diagi = rstart + 0.5 * drint
    rmid = rstart
    xbar = xsbar = ebar = exbar = 0.0
    alpha = alfnew = 1.0
    balpha = 0.0
    if ialpha > 0:
        balpha = (o.alpha2 - o.alpha1) / (rstop - rstart)
        alpha = o.alpha1 + balpha * (diagi - rstart)
        if o.ialfp:
            alpha = o.alpha1
        drnow = drint / ialpha if alpha == 1.0 else drint * (alpha - 1.0) / (alpha ** ialpha - 1.0)
    rnow = rstart + drnow
    rmidi = diagi
    k = 0
    while True:
        rec_end = False
        if replay is not None:
            st = replay[k]
            rnow, drnow, last = st.rnow, st.drnow, st.last
            if st.newint:
                rmidi = st.rmidi
            rec_end = st.endint
            k += 1
        rcent = rnow - 0.5 * drnow
        w0, _ = blk.hammat(rcent, ered)
        w1 = blk.dermat(rcent, 1) if ivpd else None
        w2 = blk.dermat(rcent, 2) if need_w2 else None
        factor = drnow * drnow / 24.0 if o.ishift else 0.0
        newint = icrmat
        if icrmat:
            if replay is None:
                rmidi = diagi
            wm = w0 if rmidi == rcent else blk.hammat(rmidi, ered)[0]
            vecold = vecnew
            _, vecnew = np.linalg.eigh(wm)
        w0 = vecnew.T @ w0 @ vecnew
        w0 = 0.5 * (w0 + w0.T)
        if w1 is not None:
            w1 = vecnew.T @ w1 @ vecnew
            w1 = 0.5 * (w1 + w1.T)
        if w2 is not None:
            w2 = vecnew.T @ w2 @ vecnew
            w2 = 0.5 * (w2 + w2.T)
        if icrmat:
            tstore = vecold.T @ vecnew
            rmat = tstore.T @ rmat @ tstore
            if o.isym:
                rmat = 0.5 * (rmat + rmat.T)
        icrmat = False
        if o.ishift:
            d2 = np.diag(w2).copy()
            diag = -np.diag(w0) - factor * d2
            np.fill_diagonal(w0, -factor * d2)
        else:
            diag = -np.diag(w0).copy()
            np.fill_diagonal(w0, 0.0)
        # contributions not wanted (vvprop.f:318-332)
        if not ivd:
            w0 = w0 * 0.0
        if w1 is not None and not ivpd:
            w1 = None
        if w2 is not None and not ivppd:
            w2 = None
        # zeroth-order solutions over the step (vvprop.f:360-395)
        dif = diag
        xsq = dif * drnow * drnow
        xl = np.sqrt(np.abs(dif))
        xk = xl * drnow
        op = dif >= 0.0
        with np.errstate(over="ignore"):
            sx = np.where(op, np.sin(xk), np.sinh(xk)) / xl
            cx = np.where(op, np.cos(xk), np.cosh(xk))
        sinx = sx
        sine = np.where(op, sx * xl, -(sx * xl))
        cosx = cx
        a1, a1p, b1, b1p = g1.copy(), g1p.copy(), g2.copy(), g2p.copy()
        g1 = a1p * sx + a1 * cx
        g1p = a1p * cx - dif * a1 * sx
        g2 = b1p * sx + b1 * cx
        g2p = b1p * cx - dif * b1 * sx
        g2pmax = g2p[n - 1] * cx[n - 1]
        if do_pert:
            cc, cs, ss = pert2(cosx, sine, xsq, xk, drnow, w0, a1, b1, a1p, b1p, eye11, eye12, eye22,
                               v1=w1, v2=w2, pert1=pert1)
        else:
            # no PERT call: the estimates read the potential matrices themselves
            cc = w0
            cs = w1 if w1 is not None else np.zeros((n, n))
            ss = w2 if w2 is not None else np.zeros((n, n))
        # first-order corrections and their size estimates (vvprop.f:410-470)
        prt1 = g1[None, :] * eye12 - g2[None, :] * eye11
        prt2 = g1[:, None] * eye22 - g2[:, None] * eye12
        prt1p = g1p[None, :] * eye12 - g2p[None, :] * eye11
        prt2p = g1p[:, None] * eye22 - g2p[:, None] * eye12
        if replay is None:
            an = 1.0 / np.sqrt(a1p * a1p / np.abs(diag) + a1 * a1)
            bn = 1.0 / np.sqrt(b1p * b1p / np.abs(diag) + b1 * b1)
            anp = drnow * an / xk
            bnp = drnow * bn / xk
            expx = np.where(diag > 0.0, 1.0, np.exp(-xk * drint / drnow))
            expx = np.where((diag <= 0.0) & (diag < -o.xsqmax), 0.0, expx)
            wgt = expx
            an, bn, anp, bnp = an * expx, bn * expx, anp * expx, bnp * expx
            e = np.maximum.reduce([np.abs(prt1) * an[None, :], np.abs(prt2) * bn[:, None],
                                   np.abs(prt1p) * anp[None, :], np.abs(prt2p) * bnp[:, None]])
            eye_n = np.eye(n, dtype=bool)
            coff = float(e[~eye_n].max()) if n > 1 else 0.0
            cdiag = float(np.diag(e).max())
            coff = max(coff, 0.0)
            sj = sinx[None, :]
            si = sinx[:, None]
            cj = cosx[None, :]
            ci = cosx[:, None]
            dj = diag[None, :]
            di = diag[:, None]
            wj = wgt[None, :]
            wi = wgt[:, None]
            kj = xk[None, :]
            ki = xk[:, None]
            ccij, ccji = cc, cc.T
            csij, csji = cs, cs.T
            ssij, ssji = ss, ss.T
            es = np.maximum.reduce([
                np.abs(sj * csji - cj * ssij) * wj * kj / drnow,
                np.abs(si * ccij - ci * csji) * wi,
                np.abs(cj * csji + dj * sj * ssij) * wj,
                np.abs(ci * ccij + di * si * csji) * wi * drnow / ki,
                np.abs(si * csij - ci * ssji) * wi * ki / drnow,
                np.abs(sj * ccji - cj * csij) * wj,
                np.abs(ci * csij + di * si * ssji) * wi,
                np.abs(cj * ccji + dj * sj * csij) * wj * drnow / kj])
            lowr = np.tril(np.ones((n, n), dtype=bool), -1)
            soff = float(es[lowr].max()) if n > 1 else 0.0
            sdiag = float(np.diag(es).max())
            if soff == 0.0:
                soff = 1.0e-30
        if not o.ipert:
            prt1 = prt2 = prt1p = prt2p = np.zeros((n, n))
        # is this the interval's last step (vvprop.f:480-500)
        if last:
            endint = True
        elif replay is not None:
            endint = rec_end
        elif ialpha > 0:
            endint = istep % ialpha == 0
        else:
            endint = None
            if abs(g2pmax) > 1.0e4:
                endint = True
            elif coffl != 0.0:
                facc = min(coff / coffl, 2.0)
                facs = min(soff / soff1, 2.0)
                if facc * coff > 0.8 * toff or facs * soff > 0.8 * tldiag:
                    endint = True
            if endint is None:
                coffl = coff
                soff1 = soff
                sdiag1 = sdiag
                endint = not (coff < 0.8 * toff and cdiag < 0.8 * toff)
        steps.append(VivsStep(rnow, drnow, newint, rmidi, last, endint))
        if endint:
            if replay is None:
                coffl = 0.0
                soff1 = soff
                sdiag1 = sdiag
            icrmat = True
            # vvprop.f:510-618
            w0m = prt1p + np.diag(g1p)
            w1m = prt2p.T + np.diag(g2p)
            w2m = prt1 + np.diag(g1)
            wm = prt2.T + np.diag(g2)
            ts = w0m + rmat @ w1m
            vo = w2m + rmat @ wm
            vo = np.linalg.solve(ts, vo)
            rmat = 0.5 * (vo + vo.T) if o.isym else vo.T.copy()
            g1 = np.zeros(n)
            g2 = np.ones(n)
            g1p = np.ones(n)
            g2p = np.zeros(n)
            eye11 = np.zeros((n, n))
            eye12 = np.zeros((n, n))
            eye22 = np.zeros((n, n))
        if not last:
            istep += 1
        if last:
            break
        if replay is not None:
            continue
        # next step and interval (vvprop.f:600-665)
        xbar += rnow
        xsbar += rnow * rnow
        ebar += sdiag
        exbar += rnow * sdiag
        tmxx = max(0.5 * toff, tldiag)
        if drint != drnow:
            soff = 0.8 * tldiag * (soff / (0.8 * tmxx)) ** 1.5
        if ialpha > 0:
            drnow = drnow * alpha
        else:
            drnow = delrd(drnow, sdiag, soff, tldiag, drmax, diag[0], diag[n - 1], rnow, rstop)
        if icrmat:
            drint = rnow - rmid
            drint1 = drint
            if ialpha > 0:
                xbar /= ialpha
                xsbar /= ialpha
                ebar /= ialpha
                exbar /= ialpha
                slope = 0.0 if ialpha == 1 else (exbar - xbar * ebar) / (xsbar - xbar * xbar)
                bint = ebar - xbar * slope
                emax = bint + slope * rnow
                emin = bint + slope * (rnow - drint)
                alfnew = alpha
                if ialpha > 1:
                    if emax == 0.0:
                        emax = 1.0e-30
                    fac = emin / emax
                    if fac <= 0.0:
                        fac = 0.9 if emin <= 0.0 else 1.1
                    else:
                        fac = min(max(fac ** (1.0 / float(3 * ialpha - 3)), 0.9), 1.1)
                    alfnew = alpha * fac
            xbar = xsbar = ebar = exbar = 0.0
            tmxx = max(toff, tldiag)
            if drint != drnow:
                coff = 0.8 * tmxx * (coff / (0.8 * tmxx)) ** 1.5
            drint = delrd(drint, cdiag, coff, tmxx, drmax, diag[0], diag[n - 1], rnow, rstop)
            if drint1 != drnow:
                soff1 = tldiag * (2.0 * soff1 / tldiag) ** 1.5
            drint1 = delrd(drint1, sdiag1, soff1, tldiag, drmax, diag[0], diag[n - 1], rnow, rstop)
            if abs(drint1) < abs(drint):
                drint = drint1
            if drint < drnow:
                drint = drnow
            if abs(rstop - rnow - drint) < abs(0.01 * drint):
                drint = rstop - rnow
            if (rstop - rnow - drint) * drint < 0.0:
                drint = rstop - rnow
            rmid = rnow
            diagi = rnow + 0.5 * drint
            if ialpha > 0:
                alpha = o.alpha1 + balpha * (diagi - rstart)
                if o.ialfp:
                    alpha = alfnew
                drnow = drint / ialpha if alpha == 1.0 else drint * (alpha - 1.0) / (alpha ** ialpha - 1.0)
        if abs(rstop - rnow - drnow) < abs(0.01 * drnow):
            drnow = rstop - rnow
        if (rstop - rnow - drnow) * drnow < 0.0:
            drnow = rstop - rnow
        rnow = rnow + drnow
        if abs((rnow - rstop) / drnow) < 0.005:
            last = True
    # back to the primitive basis (vvprop.f:670-682)
    rmat = vecnew @ rmat @ vecnew.T
    return (0.5 * (rmat + rmat.T) if o.isym else rmat), steps


def smatrix_ld_vivs(blk: Block, ered: float, ld: Grid, rmax: float, tolhi: float = 1.0e-4,
                    replay: Optional[List[VivsStep]] = None, opts: Optional[VivsOptions] = None):
    """The two-segment run IPROPS = 5, IPROPL = 4 (source_code/scctrl.f:
    99-208): YINIT and LDPROP on the constant-step grid RSTART -> RMID, Y
    inverted to R, VVPROP RMID -> RMAX with the LD segment's step as its
    first step (DRSET returns at once for IPROP = 4, source_code/drset.f:
    37-39) and TLDIAG = TOFF from TOLHI, R inverted to Y, YTOK + KTOS;
    opts = the /LDVVCM/ controls (VivsOptions).
    Returns the smatrix_ld dict plus STEPS (the VVPROP steps; pass them as
    replay for a later energy, as the reference does with its scratch
    file)."""
    wvec, nopen = gtwvec(ered, blk.eint)
    nb = chnsrt(blk.eint)
    y = yinit(blk, ered, ld.rstart)
    y, nodes = ldprop(y, lambda r: blk.hammat(r, ered), ld.rstart, ld.nstep, ld.dr)
    rm, _ = syminv(y)
    tldiag, toff = vivs_tolerances(tolhi)
    rm, steps = vvprop(rm, blk, ered, ld.rstop, rmax, ld.dr, tldiag, toff, replay=replay, opts=opts)
    y, _ = syminv(rm)
    out = dict(Y=y, NB=nb, NOPEN=nopen, WVEC=wvec, NODES=nodes, STEPS=steps)
    if nopen > 0:
        k = ytok(nb, wvec, blk.n, nopen, y, rmax, blk.cent)
        sr, si = ktos(k)
        out.update(K=k, SR=sr, SI=si)
    return out
