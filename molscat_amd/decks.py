"""Synthetic decks for the propagation-path workloads (atom + homonuclear
rigid rotor, LJ 12-6 Legendre terms with seeded anisotropy, RUNIT = 1 A,
IRMSET = 0).  Each deck is the &INPUT/&BASIS/&POTL content the reference
would read, reduced to what the propagation path consumes.

  rotor36   B=60 cm-1, even j=0..10, JTOT=10, block IB=2 (N=36), URED=1.5,
            lambda=0,2,4 with c ~ U[0.1,0.4] (seed 0), LDMD 2->30 A at
            DR=0.01 (2800 sectors), E=100 cm-1.
  rotor441  even j=0..40, JTOT=60, IB=2 (N=441), lambda=0..6 with
            c ~ U[0.05,0.4] (seed 1), LDMD 2->20 A at DR=0.005 (3600
            sectors), then AIRY 20->200 A, E=300 cm-1.
  sweep     the rotor441 system for JTOT=0..120, IB=1,2 (blocks of N from 20
            to 441) at 64 energies uniform in [10, 500] cm-1.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import system as S

RM = 3.5       # Angstrom
EPSIL = 50.0   # cm-1
URED = 1.5     # amu
BE = 60.0      # cm-1


def anisotropy(seed: int, lo: float, hi: float, nlam: int) -> np.ndarray:
    """c_lambda for lambda > 0, drawn U[lo, hi] from numpy's PCG64(seed)."""
    return np.random.default_rng(seed).uniform(lo, hi, size=nlam)


@dataclass
class LDSegment:
    rstart: float
    rstop: float
    dr_in: float

    @property
    def nstep(self) -> int:
        return S.drset_constant(self.rstart, self.rstop, self.dr_in)[0]

    @property
    def dr(self) -> float:
        return S.drset_constant(self.rstart, self.rstop, self.dr_in)[1]

    def radii(self) -> np.ndarray:
        return S.ld_radii(self.rstart, self.dr, self.nstep)


@dataclass
class Deck:
    name: str
    jlevels: List[int]
    lambdas: List[int]
    cvals: List[float]
    jtot: int
    ib: int
    energies: List[float]
    ld: LDSegment
    airy_stop: float = 0.0      # 0: no AIRY segment
    airy_dr: float = 0.0        # first AIRY step (TOLHIL=0, POWRL=4/3)
    extra: dict = field(default_factory=dict)

    def potential(self) -> S.PowerPotential:
        return S.lj_rotor_potential([1.0, *self.cvals], RM, EPSIL)

    def airy_grid(self):
        """(r, drnow) of the AIRY segment from the LD segment's end to
        airy_stop: TOLHIL = 0, steps growing as R^POWRL with POWRL = 4/3
        (the AIRY default, source_code/mol.prpini.f:111-116), the first
        airy_dr wide (DRSET, source_code/drset.f:62-80)."""
        powr = 4.0 / 3.0
        nstep, dr0 = S.drset_airy(self.ld.rstop, self.airy_stop, self.airy_dr, powr)
        return S.airy_grid(self.ld.rstop, self.airy_stop, nstep, dr0, powr)

    def block(self, jtot=None, ib=None) -> S.RotorBlock:
        return S.rotor_block(self.jtot if jtot is None else jtot, self.ib if ib is None else ib, self.jlevels, BE,
                             self.lambdas, self.potential(), URED, runit=1.0)


def rotor36() -> Deck:
    c = anisotropy(0, 0.1, 0.4, 2)
    return Deck("rotor36", list(range(0, 11, 2)), [0, 2, 4], list(c), 10, 2, [100.0], LDSegment(2.0, 30.0, 0.01))


def rotor441() -> Deck:
    c = anisotropy(1, 0.05, 0.4, 3)
    return Deck("rotor441", list(range(0, 41, 2)), [0, 2, 4, 6], list(c), 60, 2, [300.0],
                LDSegment(2.0, 20.0, 0.005), airy_stop=200.0, airy_dr=0.2)


def sweep_energies() -> np.ndarray:
    return np.linspace(10.0, 500.0, 64)


# ---------------------------------------------------------------------------
# field-dependent deck (a synthetic BASE9 system with NCONST = 1)
# ---------------------------------------------------------------------------
MHZ_IN_INV_CM = 3.335640951e-11 * 1.0e6  # source_code/physical_constants_module.f


@dataclass
class FieldBlock:
    """A synthetic BASE9-style block: 12 internal states x L = 0, 2, 4, 6
    (N = 48, channel 4*alpha + l, so the sets of equal L interleave), three
    radial terms (POTENL case 1, EPSIL = 1 cm-1, RM = 5 A): a singlet
    4000 cm-1 and a triplet 200 cm-1 LJ 12-6 well and a dipolar d(R) =
    0.5 (RM/R)^3, with coupling matrices P_s x I_L, (I - P_s) x I_L and
    D x C_L (P_s a projector on four seeded internal states, D a seeded
    symmetric internal operator, C_L coupling L and L+-2); one constant term,
    the Zeeman operator diag(s_alpha) x I_L with slopes s ~ U[-2, 2] MHz/G,
    whose constant is the field B in G (HAMMAT's P = VCONST * CM2RU).
    EINT is unused (NCONST > 0: hammat.f:95-99); the thresholds are YTRANS's.
    URED = 40 amu, RUNIT = 1 A."""

    vl: np.ndarray
    eint: np.ndarray
    cent: np.ndarray
    lvals: np.ndarray
    levidx: np.ndarray
    cm2ru: float
    ep2ru: float
    rscale: float
    potl: S.PowerPotential
    mxlam: int = 3
    nconst: int = 1
    scalam: float = 1.0                                    # SCAPOT's factor on the radial terms
    expval: Optional[Tuple[int, int, float]] = None        # PERTRB's /EXPVAL/ (I, IPOW, DELTAN)

    @property
    def n(self) -> int:
        return len(self.eint)

    def ptab(self, r: np.ndarray) -> np.ndarray:
        x = np.asarray(r, dtype=np.float64) * self.rscale
        return S.hammat_p(self.potl.table(x), x, self.ep2ru, self.scalam, self.expval)

    def pconst(self, field_g: float) -> np.ndarray:
        return S.pertrb_const(np.array([field_g * self.cm2ru]), self.mxlam, self.ep2ru, self.expval)


def _pack_upper(m: np.ndarray) -> np.ndarray:
    n = m.shape[0]
    return np.array([m[i, j] for j in range(n) for i in range(j + 1)])


def field48(seed: int = 4, zeeman=None) -> FieldBlock:
    """The field-dependent deck; zeeman replaces the internal Zeeman operator
    (12 x 12 symmetric, in MHz/G) when given."""
    rng = np.random.default_rng(seed)
    na, lv = 12, np.array([0, 2, 4, 6])
    q, _ = np.linalg.qr(rng.standard_normal((na, na)))
    ps = q[:, :4] @ q[:, :4].T
    d = rng.standard_normal((na, na))
    d = 0.3 * (d + d.T) / 2.0
    cl = np.zeros((4, 4))
    for a in range(4):
        for b in range(4):
            if abs(int(lv[a]) - int(lv[b])) == 2:
                cl[a, b] = 1.0 / np.sqrt(3.0)
            elif a == b and lv[a] > 0:
                cl[a, b] = -0.5
    slopes = rng.uniform(-2.0, 2.0, na)
    z = np.diag(slopes) if zeeman is None else np.asarray(zeeman, dtype=np.float64)
    il = np.eye(4)
    mats = [np.kron(ps, il), np.kron(np.eye(na) - ps, il), np.kron(d, cl), np.kron(z * MHZ_IN_INV_CM, il)]
    vl = np.stack([_pack_upper(m) for m in mats])
    ured, runit, rm = 40.0, 1.0, 5.0
    c2r = S.cm2ru(ured, runit)
    potl = S.PowerPotential([[(4000.0, -12), (-8000.0, -6)], [(200.0, -12), (-400.0, -6)], [(0.5, -3)]],
                            rm=rm, epsil=1.0)
    lvals = np.tile(lv, na)
    return FieldBlock(vl=vl, eint=np.zeros(na * 4), cent=(lvals * (lvals + 1)).astype(np.float64), lvals=lvals,
                      levidx=np.repeat(np.arange(na), 4), cm2ru=c2r, ep2ru=1.0 * c2r, rscale=runit / rm, potl=potl)


def hyperfine48(seed: int = 11, nsinglet: int = 12, dnorm: float = 1.0e-3) -> FieldBlock:
    """The field-scan system of the batched BOUND/FIELD workload: N = 48
    s-wave hyperfine channels (L = 0 throughout, so YTRANS has one set) with
    H(R, B) = T + V_s(R) P_s + V_t(R) P_t + B Z + d(R) D:
      * P_s = Q_s Q_s^T over the first nsinglet columns of a seeded random
        orthogonal 48 x 48 matrix Q (QR of a Gaussian matrix, signs fixed by
        R's diagonal), P_t = I - P_s;
      * V_s, V_t LJ 12-6 curves of depth 4000 and 200 cm-1 at RM = 5 A
        (POTENL case 1, EPSIL = 1 cm-1);
      * Z = diag(s_i), Zeeman slopes s_i ~ U[-2, 2] MHz/G, one per channel,
        a constant term whose value is the field B in G;
      * D a seeded symmetric matrix of spectral norm dnorm cm-1 with
        d(R) = (RM/R)^3.
    URED = 40 amu, RUNIT = 1 A; EINT unused (NCONST = 1), the thresholds
    are YTRANS's: s_i B."""
    rng = np.random.default_rng(seed)
    n = 48
    g = rng.standard_normal((n, n))
    q, r = np.linalg.qr(g)
    q = q * np.sign(np.diag(r))
    ps = q[:, :nsinglet] @ q[:, :nsinglet].T
    pt = np.eye(n) - ps
    slopes = rng.uniform(-2.0, 2.0, n)
    d = rng.standard_normal((n, n))
    d = 0.5 * (d + d.T)
    d *= dnorm / np.abs(np.linalg.eigvalsh(d)).max()
    mats = [ps, pt, d, np.diag(slopes * MHZ_IN_INV_CM)]
    vl = np.stack([_pack_upper(m) for m in mats])
    ured, runit, rm = 40.0, 1.0, 5.0
    c2r = S.cm2ru(ured, runit)
    potl = S.PowerPotential([[(4000.0, -12), (-8000.0, -6)], [(200.0, -12), (-400.0, -6)], [(1.0, -3)]],
                            rm=rm, epsil=1.0)
    return FieldBlock(vl=vl, eint=np.zeros(n), cent=np.zeros(n), lvals=np.zeros(n, dtype=np.int64),
                      levidx=np.arange(n), cm2ru=c2r, ep2ru=1.0 * c2r, rscale=runit / rm, potl=potl)
