"""Bound-state search on the device's batched BOUND matching.

BOUND locates a level by the node count NODEC(E), the number of states
below E, and converges on the energy where the smallest eigenvalue of the
matching matrix, EIGMIN(E), passes through zero (source_code/bd.driver.f:
879-1215).  Here every iteration evaluates many trial energies of every
level being sought in one ``mcp_bound_match`` call: multisection on the
node count until the bracket is narrow, then Illinois regula falsi on EIGMIN
inside the bracket.
"""
from __future__ import annotations

from typing import Dict, List

import numpy as np


def find_levels(eng, b, g_out, g_in, cm2ru: float, emin: float, emax: float, iprop_out: int = 6,
                iprop_in: int = 6, nsect: int = 15, width: float = 1e-4, etol: float = 1e-12,
                maxit: int = 60) -> List[Dict]:
    """Levels between emin and emax (cm-1) of one block.

    Returns one dict per level: STATE (1-based count of states below it),
    ENERGY (cm-1), EIGMIN at that energy, and the node counts at the
    ends of the range (NODES_EMIN, NODES_EMAX)."""
    ends = eng.bound_match(b, g_out, g_in, [emin * cm2ru, emax * cm2ru], iprop_out, iprop_in)
    if np.any(ends["STATUS"] != 0):
        raise RuntimeError(f"bound_match failed at the range ends: {ends['STATUS']}")
    nlo, nhi = int(ends["NODEC"][0]), int(ends["NODEC"][1])
    states = list(range(nlo + 1, nhi + 1))
    lo = {n: emin for n in states}
    hi = {n: emax for n in states}

    def evaluate(energies):
        r = eng.bound_match(b, g_out, g_in, np.asarray(energies) * cm2ru, iprop_out, iprop_in)
        if np.any(r["STATUS"] != 0):
            raise RuntimeError(f"bound_match failed: {r['STATUS']}")
        return r["NODEC"], r["EIGMIN"]

    # multisection on the node count: state n lies where NODEC steps to n
    for _ in range(maxit):
        active = [n for n in states if hi[n] - lo[n] > width]
        if not active:
            break
        pts = {n: np.linspace(lo[n], hi[n], nsect + 2)[1:-1] for n in active}
        nod, _ = evaluate(np.concatenate([pts[n] for n in active]))
        k = 0
        for n in active:
            for e in pts[n]:
                if nod[k] >= n:
                    hi[n] = min(hi[n], e)
                else:
                    lo[n] = max(lo[n], e)
                k += 1
    # Illinois regula falsi on EIGMIN inside each bracket
    _, f = evaluate(np.concatenate([[lo[n], hi[n]] for n in states]) if states else np.zeros(0))
    fa = {n: f[2 * i] for i, n in enumerate(states)}
    fb = {n: f[2 * i + 1] for i, n in enumerate(states)}
    side = {n: 0 for n in states}
    best = {n: (lo[n] if abs(fa[n]) < abs(fb[n]) else hi[n], min(fa[n], fb[n], key=abs)) for n in states}
    for _ in range(maxit):
        active = [n for n in states if hi[n] - lo[n] > etol * max(1.0, abs(hi[n])) and fa[n] * fb[n] < 0.0]
        if not active:
            break
        trial = [(fb[n] * lo[n] - fa[n] * hi[n]) / (fb[n] - fa[n]) for n in active]
        _, f = evaluate(trial)
        for n, e, fe in zip(active, trial, f):
            if abs(fe) <= abs(best[n][1]):
                best[n] = (e, fe)
            if fe == 0.0:
                lo[n] = hi[n] = e
                continue
            if fe * fb[n] > 0.0:
                hi[n], fb[n] = e, fe
                if side[n] == -1:
                    fa[n] *= 0.5
                side[n] = -1
            else:
                lo[n], fa[n] = e, fe
                if side[n] == 1:
                    fb[n] *= 0.5
                side[n] = 1
    return [dict(STATE=n, ENERGY=best[n][0], EIGMIN=best[n][1], NODES_EMIN=nlo, NODES_EMAX=nhi) for n in states]


def field_levels(eng, fb, delta_cm: float, fields, rmin: float, rmatch: float, rmax: float, dr: float,
                 width: float = 1e-6, btol: float = 1e-10, maxit: int = 60) -> List[Dict]:
    """The fields at which a bound state lies delta_cm below the lowest
    threshold, for a block with one constant term (the field): FIELD's
    search (source_code/fld.driver.f:657-1318) over B at an energy fixed
    relative to each field's threshold.  The whole scan of `fields` is one
    batched BOUND matching (YTRANS thresholds, YINIT, MDPROP out to RMATCH
    and in from RMAX, NODEC and EIGMIN per field); wherever NODEC steps
    between neighbouring fields a level crosses (a step of k is split into
    k brackets by halving), and each bracket is narrowed by bisection on
    NODEC and then converged by Illinois regula falsi on EIGMIN(B), all
    brackets together in each batched call.
    Returns one dict per crossing: FIELD (G), EIGMIN there, NODEC on both
    sides of the crossing (NODEC_LO at the lower field)."""
    from . import system as S

    b = eng.block(fb.vl, fb.eint, fb.cent, mxlam=fb.mxlam, nconst=fb.nconst)
    (r0, dro, no), (r1, dri, ni) = S.bound_segments(rmin, rmatch, rmax, dr)
    ro = S.ld_radii(r0, dro, no)
    ri = S.ld_radii(r1, dri, ni)
    go = eng.grid(ro, fb.ptab(ro), dro)
    gi = eng.grid(ri, fb.ptab(ri), dri)

    def evaluate(bs):
        bs = np.asarray(bs, dtype=np.float64)
        pc = np.stack([fb.pconst(x) for x in bs])
        eint, _ = eng.ytrans(b, pc)
        ered = eint.min(axis=1) - delta_cm * fb.cm2ru
        r = eng.bound_match(b, go, gi, ered, pconst=pc)
        if np.any(r["STATUS"] != 0):
            raise RuntimeError(f"bound_match failed: {r['STATUS']}")
        return r["NODEC"].astype(int), r["EIGMIN"]

    try:
        fields = np.asarray(fields, dtype=np.float64)
        nod, _ = evaluate(fields)
        brackets = [[fields[k], fields[k + 1], int(nod[k]), int(nod[k + 1])] for k in range(len(fields) - 1)
                    if nod[k] != nod[k + 1]]
        # brackets over which NODEC steps by more than one hold several
        # crossings: halve them until each holds one step
        for _ in range(maxit):
            multi = [br for br in brackets if abs(br[3] - br[2]) > 1 and br[1] - br[0] > width]
            if not multi:
                break
            mids = [0.5 * (br[0] + br[1]) for br in multi]
            nm, _ = evaluate(mids)
            keep = [br for br in brackets if not (abs(br[3] - br[2]) > 1 and br[1] - br[0] > width)]
            for br, m, n in zip(multi, mids, nm):
                for sub in ([br[0], m, br[2], int(n)], [m, br[1], int(n), br[3]]):
                    if sub[2] != sub[3]:
                        keep.append(sub)
            brackets = sorted(keep, key=lambda x: x[0])
        # bisection on NODEC: keep the sub-interval that still has the step
        for _ in range(maxit):
            active = [br for br in brackets if br[1] - br[0] > width]
            if not active:
                break
            mids = [0.5 * (br[0] + br[1]) for br in active]
            nm, _ = evaluate(mids)
            for br, m, n in zip(active, mids, nm):
                if n != br[2]:
                    br[1], br[3] = m, int(n)
                else:
                    br[0] = m
        out = []
        if not brackets:
            return out
        _, f = evaluate(np.concatenate([[br[0], br[1]] for br in brackets]))
        fa = [f[2 * i] for i in range(len(brackets))]
        fbv = [f[2 * i + 1] for i in range(len(brackets))]
        best = [(br[0], fa[i]) if abs(fa[i]) < abs(fbv[i]) else (br[1], fbv[i]) for i, br in enumerate(brackets)]
        side = [0] * len(brackets)
        for _ in range(maxit):
            act = [i for i, br in enumerate(brackets) if br[1] - br[0] > btol * max(1.0, abs(br[1]))
                   and fa[i] * fbv[i] < 0.0]
            if not act:
                break
            trial = [(fbv[i] * brackets[i][0] - fa[i] * brackets[i][1]) / (fbv[i] - fa[i]) for i in act]
            _, ft = evaluate(trial)
            for i, x, fx in zip(act, trial, ft):
                if abs(fx) <= abs(best[i][1]):
                    best[i] = (x, fx)
                if fx == 0.0:
                    brackets[i][0] = brackets[i][1] = x
                    continue
                if fx * fbv[i] > 0.0:
                    brackets[i][1], fbv[i] = x, fx
                    if side[i] == -1:
                        fa[i] *= 0.5
                    side[i] = -1
                else:
                    brackets[i][0], fa[i] = x, fx
                    if side[i] == 1:
                        fbv[i] *= 0.5
                    side[i] = 1
        for i, br in enumerate(brackets):
            out.append(dict(FIELD=best[i][0], EIGMIN=best[i][1], NODEC_LO=br[2], NODEC_HI=br[3]))
    finally:
        eng.destroy_block(b)
        eng.destroy_grid(go)
        eng.destroy_grid(gi)
    return out


def levels_over_fields(eng, fb, fields, e_lo: float, e_hi: float, rmin: float, rmatch: float, rmax: float,
                       dr: float, nsect: int = 3, width: float = 1e-5, etol: float = 1e-12,
                       maxit: int = 60) -> Dict:
    """BOUND at every field of `fields` for a block with one constant term
    (the field): the levels between e_lo and e_hi cm-1 relative to each
    field's lowest threshold (YTRANS's), located as BOUND does (source_code/
    bd.driver.f:879-1215: NODEC(E) counts the states below E, EIGMIN(E)
    vanishes at a level), every field's trial energies in each batched
    ``mcp_bound_match`` call: the node counts at both ends of the window for
    all fields, multisection on NODEC for every (field, state) pair until
    its bracket is narrower than `width`, then Illinois regula falsi on
    EIGMIN inside the brackets.

    Returns arrays over the pairs: FIELD_INDEX, STATE (the 1-based count of
    states below the level at that field), ENERGY (cm-1 relative to the
    field's lowest threshold), EIGMIN there; and per field NODEC_LO,
    NODEC_HI (the node counts at e_lo and e_hi), THRESHOLD (cm-1)."""
    from . import system as S

    fields = np.asarray(fields, dtype=np.float64)
    nf = len(fields)
    b = eng.block(fb.vl, fb.eint, fb.cent, mxlam=fb.mxlam, nconst=fb.nconst)
    (r0, dro, no), (r1, dri, ni) = S.bound_segments(rmin, rmatch, rmax, dr)
    ro = S.ld_radii(r0, dro, no)
    ri = S.ld_radii(r1, dri, ni)
    go = eng.grid(ro, fb.ptab(ro), dro)
    gi = eng.grid(ri, fb.ptab(ri), dri)
    try:
        pc = np.stack([fb.pconst(x) for x in fields])
        eint, _ = eng.ytrans(b, pc)
        thr = eint.min(axis=1)

        def evaluate(idx, erel):
            r = eng.bound_match(b, go, gi, thr[idx] + np.asarray(erel) * fb.cm2ru, pconst=pc[idx])
            if np.any(r["STATUS"] != 0):
                raise RuntimeError(f"bound_match failed: {r['STATUS']}")
            return r["NODEC"].astype(np.int64), r["EIGMIN"]

        ends, _ = evaluate(np.concatenate([np.arange(nf), np.arange(nf)]),
                           np.concatenate([np.full(nf, e_lo), np.full(nf, e_hi)]))
        nlo, nhi = ends[:nf], ends[nf:]
        cnt = np.maximum(nhi - nlo, 0)
        pf = np.repeat(np.arange(nf), cnt)
        pn = np.concatenate([np.arange(nlo[f] + 1, nhi[f] + 1) for f in range(nf)]) if pf.size else \
            np.zeros(0, dtype=np.int64)
        lo = np.full(pf.size, float(e_lo))
        hi = np.full(pf.size, float(e_hi))
        # multisection on the node count: state n lies where NODEC steps to n
        frac = np.arange(1, nsect + 1) / (nsect + 1.0)
        for _ in range(maxit):
            act = np.where(hi - lo > width)[0]
            if act.size == 0:
                break
            pts = lo[act, None] + (hi - lo)[act, None] * frac[None, :]
            nod, _ = evaluate(np.repeat(pf[act], nsect), pts.ravel())
            above = nod.reshape(act.size, nsect) >= pn[act, None]
            # the first trial point at or above the state, and the last below it
            hi[act] = np.where(above.any(axis=1), np.where(above, pts, np.inf).min(axis=1), hi[act])
            lo[act] = np.where((~above).any(axis=1), np.where(~above, pts, -np.inf).max(axis=1), lo[act])
        # Illinois regula falsi on EIGMIN inside each bracket
        if pf.size:
            _, f = evaluate(np.concatenate([pf, pf]), np.concatenate([lo, hi]))
            fa, fbv = f[:pf.size].copy(), f[pf.size:].copy()
        else:
            fa = fbv = np.zeros(0)
        side = np.zeros(pf.size, dtype=np.int64)
        best_e = np.where(np.abs(fa) < np.abs(fbv), lo, hi)
        best_f = np.where(np.abs(fa) < np.abs(fbv), fa, fbv)
        for _ in range(maxit):
            act = np.where((hi - lo > etol * np.maximum(1.0, np.abs(hi))) & (fa * fbv < 0.0))[0]
            if act.size == 0:
                break
            trial = (fbv[act] * lo[act] - fa[act] * hi[act]) / (fbv[act] - fa[act])
            _, ft = evaluate(pf[act], trial)
            better = np.abs(ft) <= np.abs(best_f[act])
            best_e[act] = np.where(better, trial, best_e[act])
            best_f[act] = np.where(better, ft, best_f[act])
            zero = ft == 0.0
            right = (~zero) & (ft * fbv[act] > 0.0)
            left = (~zero) & ~right
            a = act[zero]
            lo[a] = hi[a] = trial[zero]
            a = act[right]
            hi[a], fbv[a] = trial[right], ft[right]
            fa[a] = np.where(side[a] == -1, 0.5 * fa[a], fa[a])
            side[a] = -1
            a = act[left]
            lo[a], fa[a] = trial[left], ft[left]
            fbv[a] = np.where(side[a] == 1, 0.5 * fbv[a], fbv[a])
            side[a] = 1
        return dict(FIELD_INDEX=pf, STATE=pn, ENERGY=best_e, EIGMIN=best_f, NODEC_LO=nlo, NODEC_HI=nhi,
                    THRESHOLD=thr / fb.cm2ru)
    finally:
        eng.destroy_block(b)
        eng.destroy_grid(go)
        eng.destroy_grid(gi)


def simpson_weights(nstep: int) -> np.ndarray:
    """Quadrature weights (in units of the step) over nstep + 1 equally
    spaced points as WVBKSB applies them (source_code/wvbksb.f:139-170):
    the alternative extended Simpson rule, 17/48, 59/48, 43/48, 49/48 from
    each end and 1 inside, each end's weights laid from both ends towards
    the middle so the central point is counted once."""
    ends = (17.0 / 48.0, 59.0 / 48.0, 43.0 / 48.0, 49.0 / 48.0)
    w = np.zeros(nstep + 1)
    for k in range((nstep + 2) // 2):
        a = ends[k] if k < 4 else 1.0
        w[k] += a
        if nstep - k != k:
            w[nstep - k] += a
    return w


def wavefunction(eng, b, g_out, g_in, ered: float, pconst=None) -> Dict:
    """BOUND's wavefunction at one energy (reduced units) of a level found
    on the grids g_out (RMIN -> RMATCH) and g_in (RMAX -> RMATCH), MDPROP in
    both directions (source_code/bdctrl.f:266-360): the device's matching
    and back-substitution (``mcp_bound_wave``), then each channel's
    normalisation integral SUMPSI from both segments and the wavefunction
    from RMIN to RMAX divided by sqrt(sum SUMPSI).  Returns R, PSI (npts,
    N), SUMPSI (N), NODEC, EIGMIN, PSIMCH."""
    w = eng.bound_wave(b, g_out, g_in, ered, pconst)
    if w["STATUS"] != 0:
        raise RuntimeError(f"bound_wave failed: status {w['STATUS']}")
    po, pi = w["PSI_OUT"], w["PSI_IN"]
    nout, nin = po.shape[0] - 1, pi.shape[0] - 1
    dro = abs(float(g_out.r[2] - g_out.r[0]))
    dri = abs(float(g_in.r[2] - g_in.r[0]))
    sumpsi = (simpson_weights(nout)[:, None] * po * po).sum(0) * dro + \
        (simpson_weights(nin)[:, None] * pi * pi).sum(0) * dri
    r = np.concatenate([np.asarray(g_out.r)[::2], np.asarray(g_in.r)[::2][::-1][1:]])
    psi = np.concatenate([po[::-1], pi[1:]]) / np.sqrt(sumpsi.sum())
    return dict(R=r, PSI=psi, SUMPSI=sumpsi, NODEC=w["NODEC"], EIGMIN=w["EIGMIN"], PSIMCH=w["PSIMCH"])
