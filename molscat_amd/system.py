"""Host-side problem set-up for the propagation path: units, the atom +
rigid-rotor (ITYPE=1) channel basis and Percival-Seaton coupling, POTENL
power/exponential radial terms, LD radial grids, and the O(N) channel
bookkeeping done before and after propagation (GTWVEC, CHNSRT, cross
sections, scattering length).  None of this is on the device: it runs once
per (JTOT, IB) block or once per S-matrix, as in the reference driver.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from fractions import Fraction
from functools import lru_cache
from typing import List, Optional, Sequence, Tuple

import numpy as np

# Physical constants, CODATA 2022 as in source_code/physical_constants_module.f:7-88.
_PI = 3.141592653589793
_HBAR = 6.62607015e-34 / 2.0 / _PI
_C_CM = 2.99792458e8 * 1.0e2
_AMU = 1.66053906892e-27
BFCT = _HBAR / (4.0 * _PI * _C_CM) / (1.0e-10) ** 2 / _AMU  # cm-1 amu A^2 (hbar^2/2)
MHZ_IN_INV_CM = 3.335640951e-11 * 1.0e6


def cm2ru(ured: float, runit: float = 1.0) -> float:
    """CM2RU = URED*MUNIT*RUNIT^2/BFCT (source_code/mol.driver.f:734)."""
    return ured * 1.0 * runit * runit / BFCT


# ---------------------------------------------------------------------------
# angular momentum algebra (exact rational arithmetic)
# ---------------------------------------------------------------------------
@lru_cache(maxsize=None)
def _fact(n: int) -> int:
    return math.factorial(n)


@lru_cache(maxsize=None)
def threej000(a: int, b: int, c: int) -> float:
    """(a b c; 0 0 0) by its closed form."""
    j = a + b + c
    if j % 2 or c < abs(a - b) or c > a + b:
        return 0.0
    g = j // 2
    sq = Fraction(_fact(j - 2 * a) * _fact(j - 2 * b) * _fact(j - 2 * c), _fact(j + 1))
    rat = Fraction(_fact(g), _fact(g - a) * _fact(g - b) * _fact(g - c))
    v = math.sqrt(float(sq * rat * rat))
    return -v if g % 2 else v


def _tri(a, b, c) -> bool:
    return abs(a - b) <= c <= a + b


@lru_cache(maxsize=None)
def sixj(a: int, b: int, c: int, d: int, e: int, f: int) -> float:
    """{a b c; d e f} by Racah's formula in exact rational arithmetic."""
    if not (_tri(a, b, c) and _tri(a, e, f) and _tri(d, b, f) and _tri(d, e, c)):
        return 0.0

    def delta2(x, y, z):
        return Fraction(_fact(x + y - z) * _fact(x - y + z) * _fact(-x + y + z), _fact(x + y + z + 1))

    d2 = delta2(a, b, c) * delta2(a, e, f) * delta2(d, b, f) * delta2(d, e, c)
    tmin = max(a + b + c, a + e + f, d + b + f, d + e + c)
    tmax = min(a + b + d + e, a + c + d + f, b + c + e + f)
    s = Fraction(0)
    for t in range(tmin, tmax + 1):
        den = (_fact(t - a - b - c) * _fact(t - a - e - f) * _fact(t - d - b - f) * _fact(t - d - e - c)
               * _fact(a + b + d + e - t) * _fact(a + c + d + f - t) * _fact(b + c + e + f - t))
        term = Fraction(_fact(t + 1), den)
        s += -term if t % 2 else term
    if s == 0:
        return 0.0
    v = math.sqrt(float(s * s * d2))
    return v if s > 0 else -v


# ---------------------------------------------------------------------------
# ITYPE=1 atom + linear rigid rotor
# ---------------------------------------------------------------------------
def rotor_channels(jtot: int, ib: int, jlevels: Sequence[int]) -> List[Tuple[int, int]]:
    """(state index, L) list of block IB, in BASE's order (source_code/base.f:276-352)."""
    jmax = max(max(jlevels), 0)
    jmin = min(min(jlevels), 0)
    lmax = jtot + jmax
    lmin = jtot - jmax
    if lmin < 0:
        lmin = max(jmin - jtot, 0)
    ipty = ib % 2
    out = []
    for li in range(lmin, lmax + 1):
        for i, ji in enumerate(jlevels):
            if (li + ji + jtot) % 2 != ipty:
                continue
            if ji < abs(jtot - li) or ji > jtot + li:
                continue
            out.append((i, li))
    return out


def rotor_coupling(jtot: int, chans, jlevels, lambdas: Sequence[int]) -> np.ndarray:
    """VL(NHAM, N(N+1)/2), Percival-Seaton coefficients (source_code/couple.f:51-143)."""
    n = len(chans)
    vl = np.zeros((len(lambdas), n * (n + 1) // 2))
    idx = 0
    for icol in range(n):
        jc, lc = jlevels[chans[icol][0]], chans[icol][1]
        for irow in range(icol + 1):
            jr, lr = jlevels[chans[irow][0]], chans[irow][1]
            for il, lam in enumerate(lambdas):
                if abs(jc - jr) > lam or abs(lc - lr) > lam:
                    continue
                t = threej000(jc, lam, jr)
                if t == 0.0:
                    continue
                u = threej000(lc, lam, lr)
                if u == 0.0:
                    continue
                s = sixj(jr, lr, jtot, lc, jc, lam)
                if s == 0.0:
                    continue
                v = math.sqrt((2 * jc + 1) * (2 * jr + 1) * (2 * lc + 1) * (2 * lr + 1)) * t * u * s
                vl[il, idx] = -v if (jc + jr + jtot) % 2 else v
            idx += 1
    return vl


@dataclass
class PowerPotential:
    """&POTL case-1 radial terms (source_code/potenl.f:160-200): per lambda a
    list of (A, n) for A*R**n or (A, 0, E) for A*exp(E*R); R in RM units,
    result in EPSIL units."""

    terms: List[List[tuple]]
    rm: float = 1.0
    epsil: float = 1.0

    def __call__(self, r: float) -> np.ndarray:
        out = np.zeros(len(self.terms))
        for i, tl in enumerate(self.terms):
            s = 0.0
            for t in tl:
                if t[1] != 0:
                    rr = 1.0e-8 if (t[1] < 0 and abs(r) < 1.0e-8) else r
                    s += t[0] * rr ** t[1]
                else:
                    s += t[0] * math.exp(t[2] * r)
            out[i] = s
        return out

    def table(self, x: np.ndarray) -> np.ndarray:
        """The same terms at many radii at once: (len(x), number of terms),
        each term evaluated as __call__ does (pow, exp), summed in the same
        order."""
        x = np.asarray(x, dtype=np.float64)
        out = np.zeros((len(x), len(self.terms)))
        for i, tl in enumerate(self.terms):
            s = np.zeros(len(x))
            for t in tl:
                if t[1] != 0:
                    xx = np.where(np.abs(x) < 1.0e-8, 1.0e-8, x) if t[1] < 0 else x
                    s = s + t[0] * np.power(xx, float(t[1]))
                else:
                    s = s + t[0] * np.exp(t[2] * x)
            out[:, i] = s
        return out


def lj_rotor_potential(cvals: Sequence[float], rm: float, epsil: float) -> PowerPotential:
    return PowerPotential([[(c, -12), (-2.0 * c, -6)] for c in cvals], rm=rm, epsil=epsil)


def hammat_p(p: np.ndarray, x: np.ndarray, ep2ru: float, scalam: float = 1.0,
             expval: Optional[Tuple[int, int, float]] = None) -> np.ndarray:
    """HAMMAT's P array (source_code/hammat.f:38-73) from POTENL's radial
    coefficients p (npts, MXLAM) in EPSIL units at x = R*RSCALE: each |P|
    clamped to 1e60, SCAPOT's scaling of the radial terms by SCALAM
    (source_code/scapot-default.f), PERTRB's shift P(I) += DELTAN / x**IPOW
    for /EXPVAL/ = (I, IPOW, DELTAN) with 1 <= I <= MXLAM (source_code/
    pertrb.f; I outside the table changes nothing here, a constant term is
    the block's pconst), then EP2RU (EPSIL units to reduced units)."""
    p = np.where(np.abs(p) > 1.0e60, np.sign(p) * 1.0e60, p)
    p = p * scalam
    if expval is not None:
        i, ipow, dn = int(expval[0]), int(expval[1]), float(expval[2])
        if 1 <= i <= p.shape[1]:
            p = p.copy()
            p[:, i - 1] = p[:, i - 1] + dn / np.asarray(x, dtype=np.float64) ** ipow
    return ep2ru * p


def pertrb_const(pconst: np.ndarray, mxlam: int, ep2ru: float,
                 expval: Optional[Tuple[int, int, float]] = None) -> np.ndarray:
    """PERTRB on the NCONST constant terms (HAMMAT's P(MXLAM+1..), source_code/
    hammat.f:55-67): P = VCONST / EP2CM in EPSIL units, shifted by DELTAN when
    /EXPVAL/ names one of them, then EP2RU; pconst arrives as VCONST * CM2RU.
    The device takes the constants once per instance, so a shift that depends
    on R (IPOW != 0) is refused."""
    pconst = np.asarray(pconst, dtype=np.float64)
    if expval is None:
        return pconst
    i, ipow, dn = int(expval[0]), int(expval[1]), float(expval[2])
    k = i - mxlam
    if not 1 <= k <= len(pconst):
        return pconst
    if ipow != 0:
        raise ValueError("PERTRB on a constant term with IPOW != 0 makes it depend on R; "
                         "constants are taken once per instance")
    out = pconst.copy()
    out[k - 1] = ep2ru * (pconst[k - 1] / ep2ru + dn)
    return out


@dataclass
class RotorBlock:
    """One (JTOT, IB) block of an atom + rigid rotor system in reduced units."""

    jtot: int
    ib: int
    chans: list
    eint: np.ndarray
    cent: np.ndarray
    lvals: np.ndarray
    levidx: np.ndarray
    cm2ru: float
    ep2ru: float
    rscale: float
    potl: PowerPotential
    runit: float
    jch: np.ndarray                                       # rotor j of each channel
    lambdas: Tuple[int, ...]                              # the Legendre terms of VL
    jlevels: Tuple[int, ...]
    scalam: float = 1.0                                   # SCAPOT's factor (efvs SCALAM)
    expval: Optional[Tuple[int, int, float]] = None       # PERTRB's /EXPVAL/ (I, IPOW, DELTAN)
    _vl: Optional[np.ndarray] = field(default=None, repr=False)

    @property
    def vl(self) -> np.ndarray:
        """VL(MXLAM, N(N+1)/2) by COUPLE's exact coefficients, made on first
        use (the device makes the same coefficients itself through
        mcp_block_create_rotor without it)."""
        if self._vl is None:
            self._vl = rotor_coupling(self.jtot, self.chans, list(self.jlevels), self.lambdas)
        return self._vl

    def set_vl(self, vl: np.ndarray) -> None:
        """Use vl (e.g. the device's coefficients, mcp_couple_rotor) as VL."""
        vl = np.asarray(vl, dtype=np.float64)
        if vl.shape != (self.mxlam, self.n * (self.n + 1) // 2):
            raise ValueError(f"VL shape {vl.shape}")
        self._vl = vl

    @property
    def mxlam(self) -> int:
        return len(self.lambdas)

    @property
    def n(self) -> int:
        return len(self.eint)

    def ptab(self, r: np.ndarray) -> np.ndarray:
        """HAMMAT's P(lambda, R) in reduced units at the radii r (hammat_p)."""
        x = np.asarray(r, dtype=np.float64) * self.rscale
        return hammat_p(self.potl.table(x), x, self.ep2ru, self.scalam, self.expval)


def rotor_block(jtot, ib, jlevels, be, lambdas, potl: PowerPotential, ured, runit=None) -> RotorBlock:
    """BASE + COUPLE for ITYPE=1 (units: source_code/mol.driver.f:734-735,
    source_code/chckru.f: RUNIT defaults to RM)."""
    if runit is None:
        runit = potl.rm
    c2r = cm2ru(ured, runit)
    chans = rotor_channels(jtot, ib, jlevels)
    elev = [be * float(j * (j + 1)) for j in jlevels]
    eint = np.array([c2r * elev[c[0]] for c in chans])
    lv = np.array([c[1] for c in chans], dtype=np.int64)
    cent = np.array([float(l * (l + 1)) for l in lv])
    return RotorBlock(jtot=jtot, ib=ib, chans=chans, eint=eint, cent=cent, lvals=lv,
                      levidx=np.array([c[0] for c in chans], dtype=np.int64), cm2ru=c2r,
                      ep2ru=potl.epsil * c2r, rscale=runit / potl.rm, potl=potl, runit=runit,
                      jch=np.array([jlevels[c[0]] for c in chans], dtype=np.int64), lambdas=tuple(lambdas),
                      jlevels=tuple(jlevels))


# ---------------------------------------------------------------------------
# grids and channel bookkeeping
# ---------------------------------------------------------------------------
def ld_radii(rstart: float, dr: float, nstep: int) -> np.ndarray:
    """Radii visited by MDPROP/LDPROP: RSTART, then R = R + H twice per step
    (source_code/mdprop.f:72,113,203), accumulated in the same order."""
    h = dr / 2.0
    r = np.empty(2 * nstep + 1)
    x = rstart
    r[0] = x
    for k in range(1, 2 * nstep + 1):
        x = x + h
        r[k] = x
    return r


def drcalc(rstart: float, rstop: float, kstep: int, nstep: int, powr: float) -> float:
    """DRCALC: width of step kstep+1 of a grid whose steps grow as R^POWR
    (source_code/drcalc.f:23-38)."""
    powint = 1.0 - powr
    if powint != 0.0:
        powinv = 1.0 / powint
        zstart = rstart ** powint
        dz = (rstop ** powint - zstart) / float(nstep)
        return (zstart + float(kstep + 1) * dz) ** powinv - (zstart + float(kstep) * dz) ** powinv
    dz = rstart * ((rstop / rstart) ** (1.0 / float(nstep)) - 1.0)
    return dz * (1.0 + dz / rstart) ** kstep


def drset_airy(rstart: float, rstop: float, drfit: float, powr: float) -> Tuple[int, float]:
    """DRSET for TOLHI = 0 (source_code/drset.f:62-80): number of steps and
    the first step's width, constant steps for POWR = 0, steps growing as
    R^POWR otherwise."""
    rsmall, rlarge = min(rstart, rstop), max(rstart, rstop)
    prange = rlarge - rsmall
    if powr == 0.0:
        nstep = max(int(math.floor(prange / drfit + 0.5)), 1)
        dr = prange / float(nstep)
    else:
        rratio = drfit / rsmall
        if powr == 1.0:
            nstep = int(math.floor(math.log(rlarge / rsmall) / math.log(rratio + 1.0) + 0.5))
        else:
            powint = 1.0 - powr
            zrm1 = (rlarge / rsmall) ** powint - 1.0
            nstep = int(math.floor(zrm1 / ((rratio + 1.0) ** powint - 1.0) + 0.5))
        dr = drcalc(rstart, rstop, 0, nstep, powr)
    if rstop < rstart:
        dr = -abs(dr)
    return nstep, dr


AIRY_SQ3 = 1.732050807  # the sqrt(3) of POTENT's Gauss points, as written in source_code/potent.f:49


def airy_grid(rstart: float, rstop: float, nstep: int, dr0: float, powr: float):
    """AIPROP's steps for TOLHI = 0 (source_code/aiprop.f:86-91,288-327):
    (r, drnow) with r = (RSTART, RA_0, RB_0, ...) the Gauss points of each
    step (RNOW -+ 0.5*DRNOW/1.732050807, source_code/potent.f:66-67) and
    drnow the widths, accumulated in the reference's order."""
    rnows, widths = [], []
    drnow = dr0
    rlast = rstart
    rnow = rlast + drnow * 0.5
    rnext = rlast + drnow
    last = False
    kstep = 0
    while True:
        kstep += 1
        rnows.append(rnow)
        widths.append(drnow)
        if last:
            break
        drnow = drcalc(rstart, rstop, kstep, nstep, powr)
        rlast = rnext
        rnext = rnext + drnow
        last = (nstep - kstep) <= 1
        if last:
            rnext = rstop
            drnow = rnext - rlast
        rnow = rlast + 0.5 * drnow
    r = [rstart]
    for rn, w in zip(rnows, widths):
        r.append(rn - 0.5 * w / AIRY_SQ3)
        r.append(rn + 0.5 * w / AIRY_SQ3)
    return np.array(r), np.array(widths)


def drset_constant(rstart: float, rstop: float, drfit: float) -> Tuple[int, float]:
    """Constant-step branch of DRSET (source_code/drset.f:62-64)."""
    prange = abs(rstop - rstart)
    nstep = max(int(math.floor(prange / drfit + 0.5)), 1)
    return nstep, prange / nstep


def bound_segments(rmin: float, rmatch: float, rmax: float, drs: float, drl: Optional[float] = None):
    """The two log-derivative segments of BDCTRL when RMID = RMATCH
    (source_code/bd.prpini.f:24-62, bdpset.f:26-42): outward RMIN -> RMATCH
    with DRS and inward RMAX -> RMATCH with DRL (DRS when unset), each
    through DRSET with STEP = -10 (source_code/bdctrl.f:110-111).  Returns
    ((RSTART, DR, NSTEP) outward, (RSTART, DR < 0, NSTEP) inward)."""
    nout, dro = drset_constant(rmin, rmatch, drs)
    nin, dri = drset_constant(rmax, rmatch, drl if drl is not None else drs)
    return (rmin, dro, nout), (rmax, -dri, nin)


def hammat_host(blk: "RotorBlock", r: float, ered: float) -> np.ndarray:
    """W(R) of a block on the host (HAMMAT, source_code/hammat.f:28-113 with
    NCONST = NRSQ = 0): the driver's set-up needs it at a few trial radii
    (FINDRM); the propagators build W on the device."""
    n = blk.n
    p = blk.ptab(np.array([r]))[0]
    packed = np.tensordot(p, blk.vl, axes=(0, 0))
    w = np.empty((n, n))
    k = 0
    for j in range(n):
        for i in range(j + 1):
            w[i, j] = w[j, i] = packed[k]
            k += 1
    rsq = min(1.0 / (r * r), 1.0e16)
    w[np.diag_indices(n)] += rsq * blk.cent + blk.eint - ered
    return w


def findrm(lowest, rmin: float, eint: np.ndarray, cent: np.ndarray, ered: float, irmset: int):
    """FINDRM (source_code/findrm.f, the MOLSCAT driver, not ITYPE=8): the
    inner classical turning point of the lowest adiabat by secant steps from
    the innermost centrifugal turning point of the open channels, then, for
    IRMSET > 0, inwards in 3 + IRMSET/3 steps, from the secant iterate
    before the converged turning point, until the WKB phase integral under
    the barrier reaches 2.303 IRMSET.  lowest(R) returns the lowest
    eigenvalue of W(R) (DIAGVL).  Returns (RSTART, RTURN)."""
    rturn = 1.0e30
    for e, c in zip(eint, cent):
        dif = ered - e
        if dif < 0.0:
            continue
        rturn = min(rturn, max(math.sqrt(c / dif), rmin))
    itry = 1 if rturn == rmin else 0
    dv1 = 0.0
    while True:
        itry += 1
        rstart = rturn
        if itry > 25:
            return rmin, rturn
        d_prev = lowest(rstart)
        rturn = 0.999 * rstart
        found = retry = False
        for _ in range(100):
            v2 = lowest(rturn)
            dv1 = (v2 - d_prev) / (rturn - rstart)
            d_prev = v2
            if dv1 >= 0.0:
                if itry != 1:
                    return rmin, 2.0 * rmin
                rturn, retry = rmin, True
                break
            rstart = rturn
            dr = -v2 / dv1
            if dr < -0.3 * rturn:
                dr = -0.3 * rturn
            rturn = rturn + dr
            if dr > 1.0e3 or rturn <= 0.0:
                break
            if abs(dr / rturn) <= 1.0e-3:
                found = True
                break
        if retry:
            continue
        if not found:
            return rmin, rturn
        break
    if irmset <= 0:
        return rmin, rturn
    nstep = 3 + irmset // 3
    target = 2.303 * float(irmset)
    dr = (1.5 * target / math.sqrt(abs(dv1)) / float(nstep)) ** (2.0 / 3.0)
    # the phase integral starts at rstart, the secant iterate before the
    # converged rturn (findrm.f:109 sets RSTART=RTURN before the last step
    # and :179 starts from RSTART), not at rturn itself
    while True:
        phase = 0.0
        xk_prev = 0.0
        again = False
        for istep in range(1, nstep + 1):
            rnext = rstart - dr
            if rnext < 0.0:
                return 0.0, rturn
            d1 = lowest(rnext)
            if d1 <= 0.0:
                raise NotImplementedError("FINDRM: classically allowed region inside the turning point")
            xk = math.sqrt(d1)
            phase = (2.0 / 3.0) * xk * dr if istep == 1 else phase + dr * 0.5 * (xk + xk_prev)
            drnext = max(0.0, (target - phase) / xk)
            xk_prev = xk
            rstart = rnext
            if istep < nstep:
                dr = drnext / float(nstep - istep)
            if drnext <= 0.0:
                break
            if istep < nstep and dr > 0.5 * rstart and dr > 0.5 * rmin:
                dr = 0.02 * rstart
                again = True
                break
        if not again:
            return rstart, rturn


def calck(ep: float, ered: float, eint: np.ndarray) -> float:
    """CALCK (source_code/calck.f): the largest open-channel wavevector."""
    dif = ep + ered - np.asarray(eint)
    dif = dif[dif > 0.0]
    return float(np.sqrt(dif).max()) if len(dif) else 0.0


def gtwvec(ered: float, eint: np.ndarray) -> Tuple[np.ndarray, int]:
    """GTWVEC: source_code/gtwvec.f:18-30."""
    dif = ered - eint
    w = np.copysign(np.sqrt(np.abs(dif)), dif)
    return w, int(np.sum(dif > 0.0))


def chnsrt(eint: np.ndarray) -> np.ndarray:
    """CHNSRT: source_code/chnsrt.f (exchange sort; 0-based)."""
    nb = list(range(len(eint)))
    for i in range(len(eint) - 1):
        for j in range(i + 1, len(eint)):
            if eint[nb[i]] <= eint[nb[j]] or abs(eint[nb[i]] - eint[nb[j]]) < 1.0e-11:
                continue
            nb[i], nb[j] = nb[j], nb[i]
    return np.array(nb, dtype=np.int32)


def partial_sigma(sr, si, nb, nopen, levidx, jlevels, wvec, jtot, runit, nsig, wght=1.0) -> np.ndarray:
    """SPROC partial cross sections (source_code/sproc.f:380-430), ITYPE=1
    degeneracy 2j_i+1 (source_code/base.f:1126-1129).  SIGMA[F, I]."""
    xj = float(2 * jtot + 1) * math.pi * (wght if wght > 0.0 else 1.0)
    sig = np.zeros((nsig, nsig))
    for icol in range(nopen):
        lcol = levidx[nb[icol]]
        for irow in range(nopen):
            lrow = levidx[nb[irow]]
            if lrow >= nsig or lcol >= nsig:
                continue
            s_r, s_i = sr[irow, icol], si[irow, icol]
            smag = (1.0 - s_r) ** 2 + s_i ** 2 if irow == icol else s_r ** 2 + s_i ** 2
            dd = wvec[nb[irow]] ** 2 / runit ** 2
            sig[lcol, lrow] += smag * xj / (float(2 * jlevels[lrow] + 1) * dd)
    return sig


def scattering_length(sr_ii: float, si_ii: float, k: float, ll: int, lrpow: int = 6) -> complex:
    """CALCA: source_code/calca.f:42-59."""
    smag = math.hypot(sr_ii, si_ii)
    if smag > 1.0:
        sr_ii, si_ii = sr_ii / smag, si_ii / smag
    if sr_ii <= -1.0:
        return complex(1.0e30, 0.0)
    m = min(lrpow - 2, 2 * ll + 1)
    return complex(1.0 - sr_ii, -si_ii) / (complex(1.0 + sr_ii, si_ii) * complex(0.0, k ** m))
