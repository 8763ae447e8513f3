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
