"""MOLSCAT's scattering loop over the engine for the atom + rigid-rotor
(ITYPE=1) basis: the partial-wave and energy loop of the driver
(source_code/mol.driver.f:1181-1856): one log-derivative segment (IPROPS =
5, 6 or 7), or that segment to RMID followed by AIRY with TOLHI > 0
(IPROPL = 9) or VIVS (IPROPL = 4) to RMAX.

For each JTOT = JTOTL, JTOTL + JSTEP, ..., JTOTU and each parity block IB:
the block's channels and couplings (BASE + COUPLE), FINDRM's start radius
and DRSET's constant-step grid, both at the largest energy as the driver
sets them up (mol.driver.f:1490-1550), then every energy of the block in one
batched S-matrix call on the device (YINIT, MDPROP / LDPROP, YTOK, KTOS),
and SPROC's state-to-state cross sections (source_code/sproc.f:380-430)
accumulated over the blocks (ACCSIG, sproc.f:524-530).  The totals are
returned multiplied by JSTEP, as SPROC prints them (sproc.f:1190-1239).

The lowest adiabat that FINDRM follows is taken from the device eigen-
solver; W at its trial radii is built on the host (a handful of N x N
matrices per block).

field_scan is the same loop over an external field for a block with one
constant term: the fields are the instances of one batched call, and the
output is CALCA's scattering length of the s-wave entrance channel.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

from . import system as S


@dataclass
class BlockRun:
    """What one (JTOT, IB) block did: its size, grid and per-energy NOPEN,
    NODES and status."""
    jtot: int
    ib: int
    n: int
    rstart: float
    nstep: int
    dr: float
    nopen: List[int]
    nodes: List[int]
    status: List[int]
    sr: list = field(default_factory=list)    # per energy, NOPEN x NOPEN in CHNSRT order
    si: list = field(default_factory=list)


@dataclass
class CrossSections:
    energies: np.ndarray          # cm-1
    sigma: np.ndarray             # [energy][F][I], Angstrom^2, x JSTEP
    blocks: List[BlockRun] = field(default_factory=list)


def device_block(eng, blk):
    """An ITYPE = 1 block on the device with its VL made there by COUPLE's
    coefficients (mcp_block_create_rotor), so the host never forms it."""
    return eng.block_rotor(blk.jch, blk.lvals, blk.jtot, blk.lambdas, blk.eint, blk.cent)


def _lowest_on_device(eng, blk, ered):
    def lowest(r: float) -> float:
        w = S.hammat_host(blk, r, ered)
        ev, _ = eng.eigh(w[None], vectors=False)
        return float(np.asarray(ev).reshape(-1)[0])
    return lowest


def _long_range(eng, b, blk, ered, y, iprop_long, rmid, rmax, dr_ld, drfit, tolhi, powr):
    """The long-range segment RMID -> RMAX for every energy of a block, the
    first energy choosing the steps and the others replaying them (the
    scratch-file sequence): AIRY (IPROPL = 9, TOLHI > 0, first step DRFIT,
    source_code/drset.f:41-84) or VIVS (IPROPL = 4, R = Y^-1 and back,
    source_code/scctrl.f:150-166, the LD segment's step first)."""
    if iprop_long == 9:
        z, _, _, st, _ = eng.prop_airy_tolhi(b, ered, y, rmid, rmax, drfit, powr, tolhi, replay=2)
        return z, st
    tl = 0.064 * math.sqrt(tolhi / 0.001)
    n = y.shape[-1]
    rm, k1 = eng.syminv(y)
    rr, _, st, _ = eng.prop_vivs(b, ered, rm, rmid, rmax, dr_ld, tl, tl, replay=2)
    yv, k2 = eng.syminv(rr)
    # SYMINV's failure codes (IFAIL = N+1 for a zero pivot, negative for
    # LAPACK's INFO): the instance fails like a failed propagation
    bad = (np.asarray(k1) > n) | (np.asarray(k1) < 0) | (np.asarray(k2) > n) | (np.asarray(k2) < 0)
    st = np.where((st == 0) & bad, -1, st)
    return yv, st


def cross_sections(eng, jlevels: Sequence[int], be: float, lambdas: Sequence[int], potl: S.PowerPotential,
                   ured: float, energies: Sequence[float], jtotl: int, jtotu: int, jstep: int, rmin: float,
                   rmax: float, dr: float = None, iprop: int = 6, irmset: int = 9, ibs: Sequence[int] = (1, 2),
                   runit=None, steps: float = 0.0, iprop_long: int = 0, rvfac: float = 0.0, rmid: float = 0.0,
                   tolhi: float = 1.0e-4, powr: float = 3.0, isaveu: str = None, irstrt: int = 0,
                   label: str = "") -> CrossSections:
    """Integral cross sections SIG(F, I) over the open levels at each
    energy, summed over JTOT (step JSTEP) and the parity blocks ibs; rmin,
    rmax and dr in RUNIT (RM by default), energies in cm-1.  The LD step is
    dr, or pi / (CAY * steps) when steps > 0 (DRSET, source_code/drset.f:
    41-64; CAY the largest open wavevector at the largest energy).  With
    iprop_long = 9 (AIRY, TOLHI > 0) or 4 (VIVS) the LD segment ends at RMID
    (rmid, or rvfac x FINDRM's turning point, source_code/mol.driver.f:
    1504-1506) and the long-range propagator carries Y on to RMAX.

    isaveu names an S-matrix file (molscat_amd/smatfile.py) that every
    S-matrix is written to as SPROC does (source_code/sproc.f:254-281).
    irstrt = 1 or 2 restarts from that file as RESTRT does
    (source_code/restrt.f:19-27): the cross sections of the blocks it
    holds complete (every JTOT set, or every symmetry block, up to the last
    complete one) are recreated from the saved S-matrices, those blocks are
    skipped, and the file continues after them."""
    if iprop not in (5, 6, 7):
        raise ValueError("cross_sections: the LD segment's IPROPS must be 5, 6 or 7")
    if iprop_long not in (0, 4, 9):
        raise ValueError("cross_sections: IPROPL must be 0 (none), 4 (VIVS) or 9 (AIRY, TOLHI > 0)")
    energies = np.asarray(energies, dtype=np.float64)
    elev = np.array([be * float(j * (j + 1)) for j in jlevels])
    nsig = int(np.sum(elev < energies.max()))
    acc = np.zeros((len(energies), nsig, nsig))
    runs: List[BlockRun] = []
    save, done = _open_save(isaveu, irstrt, label, jlevels, be, ured, energies, max(ibs), acc)
    for jtot in range(jtotl, jtotu + 1, jstep):
        for ib in ibs:
            if (jtot, ib) in done:
                continue
            blk = S.rotor_block(jtot, ib, jlevels, be, lambdas, potl, ured, runit=runit)
            if blk.n == 0:
                continue
            ered = energies * blk.cm2ru
            eredmx = float(ered.max())
            if irmset > 0:
                # FINDRM's W on the host from the device's coupling coefficients
                blk.set_vl(eng.couple_rotor(blk.jch, blk.lvals, blk.jtot, blk.lambdas))
                rstart, rturn = S.findrm(_lowest_on_device(eng, blk, eredmx), rmin, blk.eint, blk.cent, eredmx,
                                         irmset)
            else:
                rstart, rturn = rmin, rmin
            drfit = math.pi / (S.calck(0.0, eredmx, blk.eint) * steps) if steps > 0.0 else dr
            rend = rmax
            if iprop_long:
                rend = rmid if rmid > 0.0 else rvfac * rturn
            nstep, h = S.drset_constant(rstart, rend, drfit)
            r = S.ld_radii(rstart, h, nstep)
            b = device_block(eng, blk)
            g = eng.grid(r, blk.ptab(r), h)
            wv, nops = zip(*[S.gtwvec(e, blk.eint) for e in ered])
            nb = S.chnsrt(blk.eint)
            if not iprop_long:
                out = eng.smatrix_ld(b, g, iprop, ered, np.array(wv), np.array(nops), nb, rmax)
            else:
                if getattr(blk, "expval", None) is not None:
                    raise NotImplementedError("PERTRB (/EXPVAL/) on the potential that VIVS and AIRY with "
                                              "TOLHI > 0 evaluate as they go is not supported")
                eng.set_potl(b, blk.potl, blk.rscale, blk.ep2ru, getattr(blk, "scalam", 1.0))
                y0 = eng.yinit(b, g, ered)
                y, nodes_ld, st = eng.prop_ld(b, g, iprop, ered, y0)
                y, st2 = _long_range(eng, b, blk, ered, y, iprop_long, rend, rmax, h, drfit, tolhi, powr)
                _, sr, si, st3 = eng.ytok_ktos(y, np.array(wv), np.array(nops), nb, blk.cent, rmax)
                status = np.where(st != 0, st, np.where(st2 != 0, st2, st3))
                out = dict(SR=sr, SI=si, NODES=nodes_ld, STATUS=status)
            eng.destroy_block(b)
            eng.destroy_grid(g)
            failed = [m for m in range(len(energies)) if out["STATUS"][m] != 0 and nops[m] > 0]
            if failed:
                # the reference stops the run on a failed propagation
                # (source_code/mdprop.f:282-287); so does the driver
                raise RuntimeError(f"cross_sections: JTOT={jtot} IB={ib}: propagation failed at energies "
                                   f"{[float(energies[m]) for m in failed]} (status "
                                   f"{[int(out['STATUS'][m]) for m in failed]})")
            for m in range(len(energies)):
                if nops[m] == 0:
                    continue
                acc[m] += S.partial_sigma(out["SR"][m], out["SI"][m], nb, nops[m], blk.levidx, jlevels, wv[m],
                                          jtot, blk.runit, nsig, wght=1.0)
                if save is not None:
                    op = nb[:nops[m]]
                    save.write(jtot, m + 1, ib, float(energies[m]), blk.levidx[op] + 1, blk.lvals[op],
                               np.asarray(wv[m])[op] / blk.runit, out["SR"][m], out["SI"][m])
            runs.append(BlockRun(jtot, ib, blk.n, rstart, nstep, h, list(nops), [int(x) for x in out["NODES"]],
                                 [int(x) for x in out["STATUS"]], sr=list(out["SR"]), si=list(out["SI"])))
    if save is not None:
        save.close()
    return CrossSections(energies=energies, sigma=acc * float(jstep), blocks=runs)


def _open_save(path, irstrt, label, jlevels, be, ured, energies, mxpar, acc):
    """The S-matrix file of cross_sections: a new one with SPROC's header
    (ITYPE = 1: JSTATE = (j, level index), NQN = 2; source_code/sproc.f:
    797-807), or, for irstrt = 1 or 2, the existing one cut after its last
    complete unit with that unit's cross sections added to acc.  Returns
    (writer or None, set of (JTOT, IB) blocks already done)."""
    from . import smatfile as F

    if path is None:
        if irstrt:
            raise ValueError("cross_sections: a restart needs the S-matrix file (isaveu)")
        return None, set()
    nlev = len(jlevels)
    hdr = F.Header(label=label, itype=1, jstate=np.stack([np.asarray(jlevels), np.arange(1, nlev + 1)], axis=1),
                   ured=float(ured), elevel=np.array([be * float(j * (j + 1)) for j in jlevels]),
                   energies=np.asarray(energies, dtype=np.float64))
    if not irstrt:
        return F.SMatrixWriter(path, hdr), set()
    old, keep, offset, done = F.restart_point(path, mxpar, irstrt)
    if len(old.energies) != len(energies) or np.abs(old.energies - energies).max() > 1e-8:
        raise ValueError("cross_sections: the S-matrix file was written for other energies")
    nsig = acc.shape[1]
    for e in keep:
        acc[e.inrg - 1] += S.partial_sigma(e.sr, e.si, np.arange(e.nopen), e.nopen, e.level - 1, jlevels, e.wvec,
                                           e.jtot, 1.0, nsig, wght=e.wght)
    return F.SMatrixWriter(path, hdr, append_at=offset), done


@dataclass
class FieldScan:
    fields: np.ndarray        # G
    ered: np.ndarray          # reduced collision energies used
    a: np.ndarray             # complex scattering length of the entrance channel, RUNIT
    nopen: np.ndarray
    status: np.ndarray


def field_scan(eng, fb, fields: Sequence[float], rstart: float, rstop: float, dr: float,
               ecoll_cm: float = 1.0e-4, iprop: int = 6) -> FieldScan:
    """Scattering length a(B) over many fields for a block with one constant
    term (the field), every field one instance of one batched call: YTRANS's
    thresholds for all fields (mcp_ytrans), the collision energy ecoll_cm
    above each field's lowest threshold, YINIT + MDPROP + YTRANS + YTOK/KTOS
    on the device (mcp_smatrix_ld with the fields as the instances'
    constants), and CALCA (source_code/calca.f:42-59) on the s-wave
    entrance channel, the lowest threshold's channel of the L = 0 set."""
    fields = np.asarray(fields, dtype=np.float64)
    b = eng.block(fb.vl, fb.eint, fb.cent, mxlam=fb.mxlam, nconst=fb.nconst)
    nstep, h = S.drset_constant(rstart, rstop, dr)
    r = S.ld_radii(rstart, h, nstep)
    g = eng.grid(r, fb.ptab(r), h)
    pc = np.stack([fb.pconst(f) for f in fields])
    eint, _ = eng.ytrans(b, pc)
    ered = eint.min(axis=1) + ecoll_cm * fb.cm2ru
    wv = np.empty_like(eint)
    nops = np.empty(len(fields), dtype=np.int32)
    nb = np.empty(eint.shape, dtype=np.int32)
    for m in range(len(fields)):
        wv[m], nops[m] = S.gtwvec(ered[m], eint[m])
        nb[m] = S.chnsrt(eint[m])
    out = eng.smatrix_ld(b, g, iprop, ered, wv, nops, nb, rstop, pconst=pc)
    a = np.full(len(fields), np.nan + 0j)
    for m in range(len(fields)):
        if out["STATUS"][m] != 0 or nops[m] == 0:
            continue
        e0 = int(nb[m][0])
        if int(fb.lvals[e0]) != 0:
            continue
        a[m] = S.scattering_length(out["SR"][m][0, 0], out["SI"][m][0, 0], wv[m][e0], 0)
    eng.destroy_block(b)
    eng.destroy_grid(g)
    return FieldScan(fields=fields, ered=ered, a=a, nopen=nops, status=np.asarray(out["STATUS"]))


@dataclass
class SweepResult:
    energies: np.ndarray                 # cm-1
    sigma: np.ndarray                    # [energy][NSIG][NSIG], summed over the blocks
    blocks: List[BlockRun]
    seconds: dict                        # device-call wall time by part
    tuned: dict                          # N -> the autotuner's choice used


def sweep(eng, jlevels: Sequence[int], be: float, lambdas: Sequence[int], potl: S.PowerPotential, ured: float,
          energies: Sequence[float], jtots: Sequence[int], ibs: Sequence[int], ld_grid, airy_grid, rmax: float,
          tuner=None, runit=1.0, progress=None, small_max: int = -1, mixed=None) -> SweepResult:
    """The partial-wave and energy sweep on one radial grid for every block:
    the reference's JTOT / IB / energy loops (source_code/mol.driver.f:
    1181-1856) for a fixed LDMD grid (ld_grid = (r, dr) of mcp_grid_create)
    and a fixed AIRY grid (airy_grid = (r, drnow), TOLHI = 0; or None), with
    SPROC's cross sections (source_code/sproc.f:380-430, 524-530) summed over
    the blocks.

    Blocks within the LDS-resident MDPROP limit go through one persistent
    launch together (mcp_prop_ld_multi: every block's energy-independent
    sector matrices, then all (block, energy) instances longest first),
    followed by AIRY and YTOK/KTOS per block; larger blocks each take one
    batched S-matrix call (mcp_smatrix_segments) with the setting the
    autotuner chose for their N (small_max >= 0 lowers the first group's
    N limit, e.g. 0 sends every block the second way).  Every energy of a
    block shares its U(C)
    sector matrices and AIRY step bases.  mixed = (newton_steps, r_fp64,
    n_min) runs the large blocks with N >= n_min in mixed precision
    (mcp_set_precision).  A failed instance raises, as the reference stops
    (source_code/mdprop.f:282-287)."""
    import time as _time

    energies = np.asarray(energies, dtype=np.float64)
    elev = np.array([be * float(j * (j + 1)) for j in jlevels])
    nsig = int(np.sum(elev < energies.max()))
    acc = np.zeros((len(energies), nsig, nsig))
    info = eng.info()
    blocks = []
    for jtot in jtots:
        for ib in ibs:
            blk = S.rotor_block(jtot, ib, jlevels, be, lambdas, potl, ured, runit=runit)
            if blk.n > 0:
                blocks.append((jtot, ib, blk))
    r_ld, dr = ld_grid
    proto = blocks[0][2]
    g = eng.grid(r_ld, proto.ptab(r_ld), dr)
    ga = None
    if airy_grid is not None:
        ra, wa = airy_grid
        ga = eng.airy_grid(ra, wa, proto.ptab(ra))
    secs = {"small_ld": 0.0, "small_airy_match": 0.0, "large": 0.0, "tune": 0.0}
    tuned = {}
    runs: List[BlockRun] = []

    def accumulate(jtot, ib, blk, sr, si, nodes, status, wv, nops, nb):
        failed = [m for m in range(len(energies)) if status[m] != 0 and nops[m] > 0]
        if failed:
            raise RuntimeError(f"sweep: JTOT={jtot} IB={ib}: propagation failed at energies "
                               f"{[float(energies[m]) for m in failed]} (status {[int(status[m]) for m in failed]})")
        for m in range(len(energies)):
            if nops[m] > 0:
                acc[m] += S.partial_sigma(sr[m], si[m], nb, nops[m], blk.levidx, jlevels, wv[m], jtot, blk.runit,
                                          nsig, wght=1.0)
        runs.append(BlockRun(jtot, ib, blk.n, float(r_ld[0]), (len(r_ld) - 1) // 2, dr, list(nops),
                             [int(x) for x in nodes], [int(x) for x in status], sr=list(sr), si=list(si)))

    def channels(blk):
        ered = energies * blk.cm2ru
        wv, nops = zip(*[S.gtwvec(e, blk.eint) for e in ered])
        return ered, np.array(wv), np.array(nops, dtype=np.int32), S.chnsrt(blk.eint)

    nsmall = info["max_n_md"] if small_max < 0 else min(small_max, info["max_n_md"])
    small = [t for t in blocks if t[2].n >= 2 and t[2].n <= nsmall]
    large = [t for t in blocks if t not in small]
    # the small blocks: one persistent MDPROP launch for all of them
    if small:
        t0 = _time.perf_counter()
        hs = [device_block(eng, b) for _, _, b in small]
        chans = [channels(b) for _, _, b in small]
        y0s = [eng.yinit(h, g, c[0]) for h, c in zip(hs, chans)]
        outs = eng.prop_ld_multi(hs, g, [c[0] for c in chans], y0s)
        secs["small_ld"] += _time.perf_counter() - t0
        t0 = _time.perf_counter()
        for (jtot, ib, blk), h, (ered, wv, nops, nb), (y, nodes, st) in zip(small, hs, chans, outs):
            nodes = nodes.copy()
            if ga is not None:
                y, n2, st2 = eng.prop_airy(h, ga, ered, y)
                nodes = nodes + n2
                st = np.where(st != 0, st, st2)
            _, sr, si, st3 = eng.ytok_ktos(y, wv, nops, nb, blk.cent, rmax)
            st = np.where(st != 0, st, st3)
            accumulate(jtot, ib, blk, sr, si, nodes, st, wv, nops, nb)
            eng.destroy_block(h)
            if progress:
                progress(jtot, ib, blk.n)
        secs["small_airy_match"] += _time.perf_counter() - t0
    # the large blocks: one batched call each, with the tuned setting
    for jtot, ib, blk in large:
        ered, wv, nops, nb = channels(blk)
        if tuner is not None:
            t0 = _time.perf_counter()
            ch = tuner.tune_block(blk, ered)
            secs["tune"] += _time.perf_counter() - t0
            tuned[blk.n] = ch.__dict__.copy()
            ch.apply(eng)
        t0 = _time.perf_counter()
        h = device_block(eng, blk)
        grids, iprops = ([g, ga], [6, 9]) if ga is not None else ([g], [6])
        use_mixed = mixed is not None and blk.n >= mixed[2]
        if use_mixed:
            eng.set_precision(mixed[0], mixed[1])
        try:
            out = eng.smatrix_segments(h, grids, iprops, ered, wv, nops, nb, rmax)
        finally:
            if use_mixed:
                eng.set_precision(0)
        eng.destroy_block(h)
        secs["large"] += _time.perf_counter() - t0
        if tuner is not None:
            from .autotune import reset

            reset(eng)
        accumulate(jtot, ib, blk, out["SR"], out["SI"], out["NODES"], out["STATUS"], wv, nops, nb)
        if progress:
            progress(jtot, ib, blk.n)
    eng.destroy_grid(g)
    if ga is not None:
        eng.destroy_grid(ga)
    if tuner is not None:
        tuner.save()
    runs.sort(key=lambda r: (r.jtot, r.ib))
    return SweepResult(energies=energies, sigma=acc, blocks=runs, seconds=secs, tuned=tuned)
