"""Autotuner for the batched propagation: per block size N, which device
variant runs a batch of energies fastest, measured on the device and kept in
a JSON table.

The choices are the library's own switches (no reference counterpart; the
reference has one CPU code path):

  * N within the LDS-resident MDPROP limit: threads per instance
    (mcp_set_lds_threads: 256, four waves, or 64, one wave);
  * N beyond it: the parallel-in-R tree with 32- or 16-pivot batched
    inversions (mcp_set_large_mode 0 with mcp_set_inverse_kernel 1 or 0),
    or the sequential recurrence on the matrix cores (mode 1).

Each candidate propagates a batch of `ninst` instances over a short grid of
`nsec` sectors of the workload's own block (timed after a warm-up call); the
fastest per-sector-instance wins.  Results are kept per device architecture
and N in a JSON file (by default ~/.cache/molscat_amd/autotune.json, or the
path in MOLSCAT_AMD_TUNE), so a later run reads them instead of measuring.
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import system as S

DEFAULT_PATH = os.path.join(os.path.expanduser("~"), ".cache", "molscat_amd", "autotune.json")


@dataclass
class Choice:
    lds_threads: int = 0      # 0: the library default (256)
    large_mode: int = 0       # 0 tree, 1 sequential matrix cores
    inverse_kernel: int = 1   # 1 32-pivot passes, 0 16-pivot passes
    ms_per_sector_instance: float = 0.0

    def apply(self, eng) -> None:
        eng.set_lds_threads(self.lds_threads)
        eng.set_large_mode(self.large_mode)
        eng.set_inverse_kernel(self.inverse_kernel)


def reset(eng) -> None:
    Choice().apply(eng)


@dataclass
class Autotuner:
    eng: object
    path: str = field(default_factory=lambda: os.environ.get("MOLSCAT_AMD_TUNE", DEFAULT_PATH))
    table: Dict[str, Dict[str, dict]] = field(default_factory=dict)
    arch: str = "gfx950"

    def __post_init__(self):
        if os.path.exists(self.path):
            try:
                with open(self.path) as f:
                    self.table = json.load(f)
            except (OSError, ValueError):
                self.table = {}

    def _entries(self) -> Dict[str, dict]:
        return self.table.setdefault(self.arch, {})

    def save(self) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.table, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def lookup(self, n: int) -> Optional[Choice]:
        e = self._entries().get(str(n))
        return Choice(**e) if e else None

    def candidates(self, n: int, ninst: int) -> List[Choice]:
        info = self.eng.info()
        if n <= info["max_n_md"]:
            return [Choice(lds_threads=256), Choice(lds_threads=64)]
        out = [Choice(large_mode=0, inverse_kernel=1), Choice(large_mode=0, inverse_kernel=0)]
        # the sequential matrix-core kernel needs a team of about N/32 CUs per
        # instance resident at once; beyond that it falls back to a kernel
        # known to be an order of magnitude slower, so it is not a candidate
        if ninst * ((n + 31) // 32) <= info["num_cu"]:
            out.append(Choice(large_mode=1, inverse_kernel=1))
        return out

    def tune_block(self, blk, ered: np.ndarray, nsec: int = 48, dr: float = 0.005, rstart: float = 2.0,
                   force: bool = False) -> Choice:
        """The fastest choice for blk's N with the batch of energies ered;
        measured unless the table already has N (force re-measures)."""
        if not force:
            got = self.lookup(blk.n)
            if got is not None:
                return got
        eng = self.eng
        r = S.ld_radii(rstart, dr, nsec)
        b = eng.block_rotor(blk.jch, blk.lvals, blk.jtot, blk.lambdas, blk.eint, blk.cent)
        g = eng.grid(r, blk.ptab(r), dr)
        best = None
        try:
            y0 = eng.yinit(b, g, ered)
            for ch in self.candidates(blk.n, len(ered)):
                ch.apply(eng)
                y, nodes, st = eng.prop_ld(b, g, 6, ered, y0)   # warm-up (and allocations)
                if np.any(st != 0):
                    continue
                t0 = time.perf_counter()
                eng.prop_ld(b, g, 6, ered, y0)
                ch.ms_per_sector_instance = 1e3 * (time.perf_counter() - t0) / (nsec * len(ered))
                if best is None or ch.ms_per_sector_instance < best.ms_per_sector_instance:
                    best = ch
        finally:
            reset(eng)
            eng.destroy_block(b)
            eng.destroy_grid(g)
        if best is None:
            best = Choice()
        self._entries()[str(blk.n)] = best.__dict__.copy()
        return best
