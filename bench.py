"""Benchmark of the propagation path on one MI355X.

Workload: atom + rigid rotor, even j = 0..40, JTOT = 60, parity block 2
(N = 441 channels), LJ 12-6 Legendre terms lambda = 0..6, E = 300 cm-1;
YINIT at 2 A, diabatic modified log-derivative propagation (MDPROP) from 2
to 20 A at DR = 0.005 A (3600 sectors), AIRY propagation from 20 to 200 A
(TOLHIL = 0, 162 steps growing as R^(4/3) from 0.2 A), Y->K->S matching at
200 A.  One "step" is one complete S-matrix solve with the block's VL and
the P(R) tables already resident in HBM.

Prints one JSON line: value = S-matrices/s, plus sector-steps/s; the
roofline of the LDMD propagation stage (the energy-independent sector
matrices and the recurrence, solved as a tree of merged sector maps),
priced at the reference algorithm's work per sector-step and timed with HIP
events on the library's stream; the AIRY segment's times beside it; and a
CPU baseline: the oracle's FP64 solve in the reference's call structure
with LAPACK inversions, on a bounded sample of LD sectors and AIRY steps,
on this host's cores (one process with P BLAS threads for one instance; P
single-threaded processes over instances for the sweep and field scans),
timed before the GPU is touched.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "S-matrices/s and sector-steps/s at N=441, mixed vs FP64 at max|ΔS|≤1e-8"
PEAK_FP64 = 78.6e12  # MI355X dense FP64 (vector and matrix), FLOP/s
# HBM bytes the LD stage of one solve moves (FETCH_SIZE + WRITE_SIZE over the
# stage's dispatches, each scaled by the counter's reading for a plain 256 MiB
# copy), measured by the counter passes of profiles/rotor_N441_ld_stage_traffic.txt;
# it changes only when the stage's kernels do
LD_STAGE_TRAFFIC_BYTES = 364.447e9


def _host_cpu() -> dict:
    """The host's CPU model, logical and physical core counts, and P, the
    worker count of the CPU baselines: this process's CPU affinity, at most
    16 (the host share of one GPU)."""
    model, cores, phys = "unknown", set(), "0"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                key, _, val = ln.partition(":")
                key, val = key.strip(), val.strip()
                if key == "model name" and model == "unknown":
                    model = val
                elif key == "physical id":
                    phys = val
                elif key == "core id":
                    cores.add((phys, val))
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0))
    return {"model": model, "logical_cpus": os.cpu_count(), "physical_cores": len(cores) or None,
            "affinity_cpus": aff, "p": max(1, min(aff, 16))}


def _oracle_block(O, blk, **kw):
    return O.Block(n=blk.n, vl=blk.vl, eint=blk.eint, cent=blk.cent, lvals=blk.lvals, levidx=blk.levidx,
                   potl=O.PowerPotential(blk.potl.terms, rm=blk.potl.rm, epsil=blk.potl.epsil), rscale=blk.rscale,
                   ep2ru=blk.ep2ru, cm2ru=blk.cm2ru, **kw)


def _cpu_solve_sample(ob, ered, ld, airy, rmatch, nld, nai, reps) -> dict:
    """The oracle's FP64 solve of one (block, energy) instance in the
    reference's call structure, its symmetric inversions through LAPACK
    DSYTRF/DSYTRI for N > 30 as SYMINV hands them to SYMINL
    (source_code/syminl-alloc.f:17-41): YINIT, the first nld sectors of the
    LD grid ld = (rstart, nstep, dr), the first nai AIRY steps of airy =
    (rstart, rstop, nstep, dr0, powr) (or none), YTOK + KTOS at rmatch.  Per
    part the median of reps runs; est = the time of the whole solve, each
    part scaled to its full step count.  AIRY's per-channel sector
    coefficients (O(N) Airy-function values per step, which the oracle
    evaluates with mpmath at 40 digits) are computed before the timed runs
    and not counted."""
    from oracle import molscat_oracle as O

    inv = O.syminv_lapack if ob.n > 30 else O.syminv
    rstart, nstep, dr = ld
    hm = lambda r: ob.hammat(r, ered)
    wvec, nopen = O.gtwvec(ered, ob.eint)
    nb = O.chnsrt(ob.eint)
    cache = []
    if airy is not None and nai > 0:
        ra0, ra1, nair, dr0, powr = airy
        for rn, dn in O.airy_steps(ra0, ra1, nair, dr0, powr)[:nai]:
            eig, _, hp = O.airy_potent(ob, ered, rn, dn)
            cache.append(O.airy_sector(dn, eig, hp))
    parts = {"yinit": [], "ld": [], "airy": [], "match": []}
    for _ in range(reps):
        t0 = time.perf_counter()
        y = O.yinit(ob, ered, rstart)
        t1 = time.perf_counter()
        y, _ = O.mdprop(y, hm, rstart, nld, dr, inv=inv)
        t2 = time.perf_counter()
        if cache:
            it = iter(cache)
            y, _ = O.aiprop(y, ob, ered, ra0, ra1, nair, dr0, powr, inv=inv, sector=lambda *a: next(it),
                            limit=len(cache))
        t3 = time.perf_counter()
        if nopen > 0:
            O.ktos(O.ytok(nb, wvec, ob.n, nopen, y, rmatch, ob.cent))
        t4 = time.perf_counter()
        parts["yinit"].append(t1 - t0)
        parts["ld"].append((t2 - t1) / nld)
        parts["airy"].append((t3 - t2) / max(1, len(cache)))
        parts["match"].append(t4 - t3)
    med = {k: float(np.median(v)) for k, v in parts.items()}
    nair_full = airy[2] if airy is not None else 0
    med["est"] = med["yinit"] + nstep * med["ld"] + (nair_full * med["airy"] if cache else 0.0) + med["match"]
    med["n"] = ob.n
    return med


_POOL_JOBS: list = []


def _pool_run(i):
    """One job of a CPU-baseline pool, single-threaded: ("sample", args of
    _cpu_solve_sample), ("smat", (block, energy, grid)) for the oracle's
    whole solve, ("bound", (block, energy, bound grid)) for BOUND's
    matching.  The job list is set before the pool forks."""
    from threadpoolctl import threadpool_limits

    from oracle import molscat_oracle as O

    kind, job = _POOL_JOBS[i]
    with threadpool_limits(1):
        if kind == "sample":
            return _cpu_solve_sample(*job)
        ob, ered, grid = job
        if kind == "smat":
            O.smatrix_ld(ob, ered, grid, iprop=6, inv=O.syminv_lapack if ob.n > 30 else None)
        else:
            O.bound_match(ob, ered, grid)
    return None


def _cpu_pool_samples(jobs, p):
    """_cpu_solve_sample over jobs in p single-threaded processes at once
    (forked before this process touches the GPU); each job's times are its
    own, taken while all p workers run."""
    import multiprocessing as mp

    _POOL_JOBS[:] = [("sample", j) for j in jobs]
    with mp.get_context("fork").Pool(p) as pool:
        return pool.map(_pool_run, range(len(jobs)), chunksize=1)


def _cpu_field(fb, fields, nstep, h, bound, reps=3) -> dict:
    """P single-threaded processes over fields (forked before this process
    touches the GPU): the oracle's whole solve of each field of an even
    sample of the scan (YTRANS's thresholds, YINIT, MDPROP, YTRANS, YTOK +
    KTOS), eight fields per worker, and BOUND's matching (both propagations,
    NODEC and EIGMIN) at four fields per worker; wall clock of each pass,
    the median of reps passes."""
    import multiprocessing as mp

    from oracle import molscat_oracle as O

    host = _host_cpu()
    p = host["p"]
    sub = fields[np.linspace(0, len(fields) - 1, 8 * p).astype(int)]
    obs = [_oracle_block(O, fb, nconst=1, pconst=fb.pconst(f)) for f in sub]
    thr = [O.ytrans(None, ob)[1].min() for ob in obs]
    grid = O.Grid(3.5, 25.0, nstep, h)
    scat = [("smat", (ob, t + 1.0e-4 * fb.cm2ru, grid)) for ob, t in zip(obs, thr)]
    bnd = []
    if bound is not None:
        rmin, rmatch, rmax, drb = bound
        bg = O.bound_grid(rmin, rmatch, rmax, drb)
        bnd = [("bound", (ob, t - 1.0 * fb.cm2ru, bg)) for ob, t in zip(obs[::2], thr[::2])]  # 4 per worker
    _POOL_JOBS[:] = scat + bnd
    out = {"host": host, "cores": p}
    with mp.get_context("fork").Pool(p) as pool:
        for key, idx in (("smat", range(len(scat))), ("bound", range(len(scat), len(scat) + len(bnd)))):
            if len(idx) == 0:
                continue
            walls = []
            for _ in range(reps):
                t0 = time.perf_counter()
                pool.map(_pool_run, idx, chunksize=1)
                walls.append(time.perf_counter() - t0)
            w = float(np.median(walls))
            out[key] = (len(idx) / w, w, len(idx))
    return out


def _cpu_sweep(deck, blks, ne, nld, nai, reps=3) -> dict:
    """P single-threaded processes over (block, energy) instances (forked
    before this process touches the GPU): P instances whose block sizes
    are even quantiles of the blocks' distinct N (both ends included), each
    timed by _cpu_solve_sample while all P run; the per-instance time at
    each block's N by log-log interpolation between the sampled sizes; the
    blocks' total over ne energies each, divided by P."""
    from molscat_amd import decks
    from molscat_amd import system as S
    from oracle import molscat_oracle as O

    host = _host_cpu()
    p = host["p"]
    seg = deck.ld
    powr = 4.0 / 3.0
    nair, dr0 = S.drset_airy(seg.rstop, deck.airy_stop, deck.airy_dr, powr)
    by_n = {b.n: b for b in blks}
    ns = sorted(by_n)
    pick = [ns[int(round(q))] for q in np.linspace(0, len(ns) - 1, p)]
    ev = decks.sweep_energies()
    jobs = []
    for k, n in enumerate(pick):
        b = by_n[n]
        e = float(ev[(k * ne) // p]) * b.cm2ru
        jobs.append((_oracle_block(O, b), e, (seg.rstart, seg.nstep, seg.dr),
                     (seg.rstop, deck.airy_stop, nair, dr0, powr), deck.airy_stop, nld, nai, reps))
    res = _cpu_pool_samples(jobs, p)
    tn = {}
    for r in res:
        tn.setdefault(r["n"], []).append(r["est"])
    nn = np.array(sorted(tn), dtype=float)
    tt = np.array([np.median(tn[n]) for n in sorted(tn)])
    est = {b.n: float(np.exp(np.interp(np.log(b.n), np.log(nn), np.log(tt)))) for b in blks}
    t_one = sum(est[b.n] for b in blks) * ne
    return {"host": host, "cores": p, "s_one_process": t_one, "wall": t_one / p,
            "per_instance_s": {str(int(n)): float(t) for n, t in zip(nn, tt)}}


def _cpu_rotor441(blk, deck, ered, nld, nai, reps=3) -> dict:
    """One process, BLAS threads = P.  The oracle's solve of the
    N=441 deck: YINIT, nld of the 3600 LD sectors, nai of the AIRY steps,
    YTOK + KTOS; the median of reps runs of each part, scaled to the whole
    solve."""
    from threadpoolctl import threadpool_limits

    from molscat_amd import system as S
    from oracle import molscat_oracle as O

    host = _host_cpu()
    seg = deck.ld
    ob = _oracle_block(O, blk)
    powr = 4.0 / 3.0
    nair, dr0 = S.drset_airy(seg.rstop, deck.airy_stop, deck.airy_dr, powr)
    with threadpool_limits(host["p"]):
        m = _cpu_solve_sample(ob, ered, (seg.rstart, seg.nstep, seg.dr), (seg.rstop, deck.airy_stop, nair, dr0, powr),
                              deck.airy_stop, nld, nai, reps)
    return {"value": 1.0 / m["est"], "unit": "S-matrices/s", "cores": host["p"], "kind": "port",
            "host": host, "s_per_solve": m["est"], "sector_steps_per_s": 1.0 / m["ld"],
            "airy_steps_per_s": 1.0 / m["airy"], "ms_yinit": m["yinit"] * 1e3, "ms_match": m["match"] * 1e3,
            "sample": f"the oracle (numpy + LAPACK DSYTRF/DSYTRI from numpy's OpenBLAS, {host['p']} BLAS threads, "
                      f"one process) at N={blk.n}: YINIT, the first {nld} of {seg.nstep} LD sectors, the first {nai} "
                      f"of {nair} AIRY steps, YTOK + KTOS; median of {reps} runs per part, scaled to the whole solve "
                      f"({m['est']:.1f} s); not counted: AIRY's per-channel Airy-function values (O(N) per step)"}


def _copy(args) -> int:
    """A plain device copy of known size (256 MiB, float64) for calibrating
    the HBM counters FETCH_SIZE / WRITE_SIZE against bytes moved."""
    import torch

    n = 32 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "copy", "bytes_read": n * 8, "bytes_written": n * 8}))
    return 0


def _field(args) -> int:
    """The field scan of the N = 48 hyperfine deck (decks.hyperfine48: s-wave
    channels, singlet/triplet projectors from a seeded random orthogonal
    48 x 48 matrix, 48 Zeeman slopes ~ U[-2, 2] MHz/G, a dipolar coupling
    of norm 1e-3 cm-1 times (RM/R)^3, one constant term, the field B):
      * scattering: 4096 fields uniform in [0, 1000] G in one batched call,
        the collision energy 1e-4 cm-1 above each field's lowest threshold;
        YTRANS's thresholds for all fields (mcp_ytrans), then YINIT at
        3.5 A, MDPROP to 25 A at DR = 0.02 A (1075 sectors), YTRANS and
        YTOK/KTOS (mcp_smatrix_ld), timed from the first of the two device
        calls to the last; CALCA's a(B) on the host afterwards;
      * bound states: BOUND's node count and matching determinant at every
        field in one batched call, 1 cm-1 below each field's lowest
        threshold (outward 3.5 -> 6 A, inward 40 -> 6 A at DR = 0.02 A),
        then the levels between 0.9 and 0.8 cm-1 below the lowest threshold
        (a cluster of 12 split by the field) located at --bound-fields of
        the fields >= 50 G (bound.levels_over_fields: every field's trial
        energies in each batched call).
    Beside them the oracle's rates on the host's cores (_cpu_field)."""
    from molscat_amd import bound as BD
    from molscat_amd import decks
    from molscat_amd import system as S
    from molscat_amd.engine import Engine

    fb = decks.hyperfine48()
    nf = args.fields
    fields = np.linspace(0.0, 1000.0, nf)
    nstep, h = S.drset_constant(3.5, 25.0, 0.02)
    rmin, rmatch, rmax, drb = 3.5, 6.0, 40.0, 0.02
    # the CPU baseline first: its worker processes fork before the GPU is touched
    cpu_raw = (_cpu_field(fb, fields, nstep, h, (rmin, rmatch, rmax, drb) if args.bound_fields > 0 else None,
                          args.cpu_reps) if args.cpu_sectors > 0 else None)
    eng = Engine(0)
    eng.set_md_kernel(args.md_kernel)
    b = eng.block(fb.vl, fb.eint, fb.cent, mxlam=fb.mxlam, nconst=fb.nconst)
    r = S.ld_radii(3.5, h, nstep)
    g = eng.grid(r, fb.ptab(r), h)
    pc = np.stack([fb.pconst(f) for f in fields])
    eint, _ = eng.ytrans(b, pc)
    ered = eint.min(axis=1) + 1.0e-4 * fb.cm2ru
    wv = np.empty_like(eint)
    nops = np.empty(nf, dtype=np.int32)
    nb = np.empty(eint.shape, dtype=np.int32)
    for m in range(nf):
        wv[m], nops[m] = S.gtwvec(ered[m], eint[m])
        nb[m] = S.chnsrt(eint[m])

    def solve():
        eng.ytrans(b, pc)
        out = eng.smatrix_ld(b, g, 6, ered, wv, nops, nb, 25.0, pconst=pc)
        if np.any(out["STATUS"] != 0):
            raise RuntimeError("field scan failed")
        return out

    for _ in range(args.warmup):
        solve()
    t0 = time.perf_counter()
    ms_prop = []
    for _ in range(args.steps):
        out = solve()
        ms_prop.append(eng.last_timing()["ms_prop"])
    dt = (time.perf_counter() - t0) / args.steps
    a = [S.scattering_length(out["SR"][m][0, 0], out["SI"][m][0, 0], wv[m][int(nb[m][0])], 0) for m in range(nf)]
    # the MDPROP kernel's work: per sector-step of one instance two
    # inversions, two W builds and the elementwise passes
    n, nham = fb.n, fb.mxlam + fb.nconst
    flops = nf * nstep * (2.0 * n ** 3 + nham * n * (n + 1) + 12.0 * n * n)
    kern_s = float(np.mean(ms_prop)) * 1e-3

    # BOUND: node counts and EIGMIN at every field in one call, then levels
    # (skipped with --bound-fields 0)
    bound_line = None
    if args.bound_fields > 0:
        (r0, dro, no), (r1, dri, ni) = S.bound_segments(rmin, rmatch, rmax, drb)
        ro, ri = S.ld_radii(r0, dro, no), S.ld_radii(r1, dri, ni)
        go, gi = eng.grid(ro, fb.ptab(ro), dro), eng.grid(ri, fb.ptab(ri), dri)
        e_lo, e_hi = -0.9, -0.8
        ebd = eint.min(axis=1) - 1.0 * fb.cm2ru
        eng.bound_match(b, go, gi, ebd, pconst=pc)
        t1 = time.perf_counter()
        bm = eng.bound_match(b, go, gi, ebd, pconst=pc)
        t_nodec = time.perf_counter() - t1
        if np.any(bm["STATUS"] != 0):
            raise RuntimeError("bound matching failed")
        # a subset of the scan's fields away from B = 0 (where the clusters are degenerate)
        sub = fields[fields >= 50.0][:: max(1, nf // max(1, args.bound_fields))]
        t1 = time.perf_counter()
        lev = BD.levels_over_fields(eng, fb, sub, e_lo, e_hi, rmin, rmatch, rmax, drb, width=1e-7)
        t_lev = time.perf_counter() - t1
        bound_line = {"fields_one_call": nf, "ms_nodec_all_fields": t_nodec * 1e3,
                      "bound_evaluations_per_s": nf / t_nodec,
                      "nodec_range": [int(bm["NODEC"].min()), int(bm["NODEC"].max())],
                      "sectors_per_field": no + ni, "nodec_at_cm": -1.0, "level_window_cm": [e_lo, e_hi],
                      "level_fields": len(sub), "levels_found": int(len(lev["STATE"])), "level_search_s": t_lev,
                      "levels_per_s": len(lev["STATE"]) / t_lev if t_lev > 0 else None}
    cpu = None
    if cpu_raw is not None:
        v, w, k = cpu_raw["smat"]
        cpu = {"value": v, "unit": "S-matrices/s", "cores": cpu_raw["cores"], "kind": "port",
               "host": cpu_raw["host"],
               "sample": f"the oracle's whole solve (YTRANS thresholds, YINIT, MDPROP {nstep} sectors, YTRANS, "
                         f"YTOK + KTOS; LAPACK DSYTRF/DSYTRI inversions) of {k} fields spread over the scan, "
                         f"{cpu_raw['cores']} single-threaded processes; median of {args.cpu_reps} passes "
                         f"({w:.2f} s)"}
        if "bound" in cpu_raw:
            vb, wb_, kb = cpu_raw["bound"]
            cpu["bound_evaluations_per_s"] = vb
            cpu["bound_sample"] = (f"BOUND's matching (outward {no} + inward {ni} sectors, NODEC + EIGMIN) at {kb} "
                                   f"fields, 1 cm-1 below threshold, {cpu_raw['cores']} processes; median of "
                                   f"{args.cpu_reps} passes ({wb_:.2f} s)")
    line = {
        "metric": "S-matrices/s over a field scan (N=48, 4096 fields, scattering length a(B))",
        "value": nf / dt, "unit": "S-matrices/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"hyperfine deck N=48, {nf} fields in [0, 1000] G, E = threshold + 1e-4 cm-1, "
                               f"MDPROP 3.5-25 A ({nstep} sectors), one batched call", "fields": nf},
        "a_range": [float(np.nanmin(np.real(a))), float(np.nanmax(np.real(a)))],
        "roofline": {"bound": "mfma", "kernel": ("MDPROP stage of the scan (k_mdprop_reg<48>, one wavefront per field)"
                                                if args.md_kernel == 2 or (args.md_kernel == 0 and nf >= 512) else
                                                "MDPROP stage of the scan (k_mdprop_lds, one workgroup per field)"),
                     "achieved": flops / kern_s / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": flops / kern_s / PEAK_FP64, "traffic": None, "flops_per_launch": flops,
                     "ms_per_launch": kern_s * 1e3},
        "bound": bound_line,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


def _sweep(args) -> int:
    """The batched partial-wave and energy sweep: the N=441 workload's rotor
    and potential for JTOT = 0, jstep, ..., 120 (a stratified subset of the
    121 JTOT values), both parities, 64 energies uniform in [10, 500] cm-1,
    on one radial grid for every block (LDMD 2-20 A at DR = 0.005 A, 3600
    sectors, then AIRY 20-200 A with TOLHIL = 0, 162 steps), YTOK/KTOS at
    200 A and SPROC's cross sections summed over the blocks
    (molscat_amd/driver.py:sweep).  Blocks within the LDS-resident MDPROP
    limit share one persistent launch; the others take one batched call
    each with the autotuner's setting for their N, measured first (not
    timed).  One timed pass over the subset; the full sweep's time is
    extrapolated from it by the sum of N^3 over its blocks."""
    from molscat_amd import decks, driver
    from molscat_amd import system as S
    from molscat_amd.autotune import Autotuner
    from molscat_amd.engine import Engine

    deck = decks.rotor441()
    seg = deck.ld
    r = seg.radii()
    ra, wa = deck.airy_grid()
    energies = decks.sweep_energies()[: args.energies_sweep]
    jtots = list(range(0, 121, args.jstep))
    pot = deck.potential()

    def blk_of(jt, ib):
        return S.rotor_block(jt, ib, deck.jlevels, decks.BE, deck.lambdas, pot, decks.URED, runit=1.0)

    sizes_all = [blk_of(jt, ib).n for jt in range(0, 121) for ib in (1, 2)]
    sizes_all = [n for n in sizes_all if n > 0]
    # the CPU baseline first: its worker processes fork before the GPU is touched
    cpu_raw = None
    if args.cpu_sectors > 0:
        subset = [b for b in (blk_of(jt, ib) for jt in jtots for ib in (1, 2)) if b.n > 0]
        cpu_raw = _cpu_sweep(deck, subset, len(energies), args.cpu_sectors, args.cpu_airy, args.cpu_reps)
    eng = Engine(0)
    tuner = Autotuner(eng, path=args.tune_file) if args.tune_file else Autotuner(eng)
    t0 = time.perf_counter()
    for jt in jtots:
        for ib in (1, 2):
            b = blk_of(jt, ib)
            if b.n > eng.info()["max_n_md"]:
                tuner.tune_block(b, np.asarray(energies) * b.cm2ru)
    tune_s = time.perf_counter() - t0
    print(f"sweep: tuned in {tune_s:.1f} s: " + ", ".join(f"N={k}: {v}" for k, v in sorted(
        tuner._entries().items(), key=lambda kv: int(kv[0]))), file=sys.stderr, flush=True)

    def progress(jt, ib, n):
        print(f"sweep: JTOT={jt} IB={ib} N={n} done ({time.perf_counter() - t1:.1f} s)", file=sys.stderr, flush=True)

    t1 = time.perf_counter()
    res = driver.sweep(eng, deck.jlevels, decks.BE, deck.lambdas, pot, decks.URED, energies, jtots, (1, 2),
                       (r, seg.dr), (ra, wa), deck.airy_stop, tuner=tuner, progress=progress)
    wall = time.perf_counter() - t1
    sizes = [run.n for run in res.blocks]
    nsmat = sum(sum(1 for q in run.nopen if q > 0) for run in res.blocks)
    n3 = float(sum(n ** 3 for n in sizes))
    n3_all = float(sum(n ** 3 for n in sizes_all))
    ne = len(energies)
    nstep = seg.nstep
    nham = len(deck.lambdas)
    # reference work of the LDMD stage per sector-step with B = ne energies
    # sharing U(C) and W's off-diagonal: (2 + 1/B) N^3 + (2/B) NHAM N(N+1) + 12 N^2
    flops_ld = sum(nstep * ne * ((2.0 + 1.0 / ne) * n ** 3 + (2.0 / ne) * nham * n * (n + 1) + 12.0 * n * n)
                   for n in sizes)
    cpu = None
    if cpu_raw is not None:
        cpu = {"value": nsmat / cpu_raw["wall"], "unit": "S-matrices/s", "cores": cpu_raw["cores"], "kind": "port",
               "host": cpu_raw["host"], "per_instance_s": cpu_raw["per_instance_s"],
               "sample": f"the oracle's solve (YINIT, MDPROP, AIRY, YTOK + KTOS; LAPACK DSYTRF/DSYTRI inversions "
                         f"for N > 30) of {cpu_raw['cores']} (block, energy) instances at even quantiles of the "
                         f"subset's block sizes, {cpu_raw['cores']} single-threaded processes at once, "
                         f"{args.cpu_sectors} LD sectors + {args.cpu_airy} AIRY steps each (median of "
                         f"{args.cpu_reps}), scaled to whole solves, interpolated in N over the subset's blocks "
                         f"x {ne} energies and divided by the process count ({cpu_raw['wall']:.0f} s); not "
                         f"counted: AIRY's per-channel Airy-function values"}
    line = {
        "metric": "S-matrices/s over the partial-wave x energy sweep (JTOT 0..120, both parities, 64 energies)",
        "value": nsmat / wall, "unit": "S-matrices/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"rotor sweep, JTOT {jtots[0]}..{jtots[-1]} step {args.jstep} x IB 1,2 "
                               f"({len(sizes)} blocks, N {min(sizes)}..{max(sizes)}) x {ne} energies in [10, 500] "
                               f"cm-1, LDMD 2-20 A ({nstep} sectors) + AIRY 20-200 A (TOLHIL=0), one grid",
                   "blocks": len(sizes), "energies": ne},
        "s_matrices": nsmat,
        "sector_steps_per_s": nstep * ne * len(sizes) / wall,
        "full_sweep_seconds_extrapolated": wall * n3_all / n3,
        "seconds": {k: round(v, 3) for k, v in res.seconds.items()},
        "tuned": {str(k): v for k, v in res.tuned.items()},
        "tune_seconds": tune_s,
        "roofline": {"bound": "mfma", "kernel": "whole sweep, priced at the LDMD stage's reference work",
                     "achieved": flops_ld / wall / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": flops_ld / wall / PEAK_FP64, "traffic": None},
        "cpu_baseline": cpu,
    }
    eng.close()
    print(json.dumps(line))
    return 0


def _mixed_sweep(args) -> int:
    """The mixed-precision sweep: the blocks of the sweep's stratified
    subset with N >= 200 (JTOT = 0, jstep, ..., 120, both parities, 64
    energies, the same grid as --workload sweep), solved twice: FP64, and
    with the large-N tree's batched inversions in FP32 + Newton steps (FP64
    below 3 A; YINIT, the recurrence's elementwise steps, AIRY's steps and
    the matching FP64).  Reports both rates, their ratio, max|dS| over every
    S-matrix and the largest relative change of the summed cross sections."""
    from molscat_amd import decks, driver
    from molscat_amd import system as S
    from molscat_amd.autotune import Autotuner
    from molscat_amd.engine import Engine

    deck = decks.rotor441()
    seg = deck.ld
    r = seg.radii()
    ra, wa = deck.airy_grid()
    energies = decks.sweep_energies()[: args.energies_sweep]
    pot = deck.potential()
    jtots = [jt for jt in range(0, 121, args.jstep)]
    big = [b for b in (S.rotor_block(jt, ib, deck.jlevels, decks.BE, deck.lambdas, pot, decks.URED, runit=1.0)
                       for jt in jtots for ib in (1, 2)) if b.n >= 200]
    blocks = [(b.jtot, b.ib) for b in big]
    jt_use = sorted({jt for jt, _ in blocks})
    # the CPU baseline first: its worker processes fork before the GPU is touched
    cpu_raw = (_cpu_sweep(deck, big, len(energies), args.cpu_sectors, args.cpu_airy, args.cpu_reps)
               if args.cpu_sectors > 0 else None)
    eng = Engine(0)
    tuner = Autotuner(eng, path=args.tune_file) if args.tune_file else Autotuner(eng)

    def progress(jt, ib, n):
        print(f"mixed-sweep: JTOT={jt} IB={ib} N={n} done", file=sys.stderr, flush=True)

    runs = {}
    for label, mixed in (("fp64", None), ("mixed", (args.mixed_steps, 3.0, 200))):
        t1 = time.perf_counter()
        res = driver.sweep(eng, deck.jlevels, decks.BE, deck.lambdas, pot, decks.URED, energies, jt_use, (1, 2),
                           (r, seg.dr), (ra, wa), deck.airy_stop, tuner=tuner, progress=progress, small_max=0,
                           mixed=mixed)
        runs[label] = (res, time.perf_counter() - t1)
        print(f"mixed-sweep: {label} {runs[label][1]:.1f} s", file=sys.stderr, flush=True)
    r64, w64 = runs["fp64"]
    rmx, wmx = runs["mixed"]
    nsmat = sum(sum(1 for q in run.nopen if q > 0) for run in r64.blocks if run.n >= 200)
    ds = 0.0
    nodes_equal = True
    for a, b in zip(r64.blocks, rmx.blocks):
        nodes_equal &= a.nodes == b.nodes
        for sa, sb, ia, ib_ in zip(a.sr, b.sr, a.si, b.si):
            if np.size(sa):
                ds = max(ds, float(np.abs(np.asarray(sa) - np.asarray(sb)).max()),
                         float(np.abs(np.asarray(ia) - np.asarray(ib_)).max()))
    sig_scale = np.abs(r64.sigma).max()
    dsig = float(np.abs(r64.sigma - rmx.sigma).max() / sig_scale)
    line = {
        "metric": "S-matrices/s of the N>=200 blocks of the sweep, mixed precision vs FP64 at max|dS|<=1e-8",
        "value": nsmat / wmx, "unit": "S-matrices/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": wmx * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"rotor sweep blocks with N>=200: JTOT {jt_use} x IB 1,2 ({len(r64.blocks)} blocks) x "
                               f"{len(energies)} energies, LDMD 2-20 A (3600 sectors) + AIRY 20-200 A",
                   "newton_steps": args.mixed_steps, "fp64_below_A": 3.0},
        "fp64_value": nsmat / w64, "ratio_mixed_to_fp64": w64 / wmx, "max_abs_dS": ds,
        "max_rel_dsigma": dsig, "nodes_equal": bool(nodes_equal), "s_matrices": nsmat,
        "cpu_baseline": None if cpu_raw is None else {
            "value": nsmat / cpu_raw["wall"], "unit": "S-matrices/s", "cores": cpu_raw["cores"], "kind": "port",
            "host": cpu_raw["host"], "per_instance_s": cpu_raw["per_instance_s"],
            "sample": f"the oracle's FP64 solve (YINIT, MDPROP, AIRY, YTOK + KTOS; LAPACK DSYTRF/DSYTRI) of "
                      f"{cpu_raw['cores']} (block, energy) instances at even quantiles of the N >= 200 blocks' sizes, "
                      f"{cpu_raw['cores']} single-threaded processes at once, {args.cpu_sectors} LD sectors + "
                      f"{args.cpu_airy} AIRY steps each (median of {args.cpu_reps}), scaled to whole solves over "
                      f"every block x {len(energies)} energies and divided by the process count "
                      f"({cpu_raw['wall']:.0f} s); not counted: AIRY's per-channel Airy-function values"},
    }
    eng.close()
    print(json.dumps(line))
    return 0


def _allprop(args) -> int:
    """Every propagator of the reference's all-propagators deck (ITYPE=1
    rotor BE=30, j=0,2,4,6, JTOT=10, IB=1, N=12, URED=20; FINDRM's RSTART,
    STEPS=12 at 1000 cm-1, RMAX=20 RM) for a batch of energies in [900, 1100]
    cm-1, one workgroup per energy: MDPROP, LDPROP, MAPROP and MGPROP (MA5)
    over 8072 sectors; MDPROP to RMID = 1.4 RTURN then AIRY with TOLHI = 1e-4
    (the first energy choosing the steps, the others replaying them); LDPROP
    to RMID then VIVS (likewise).  Wall time of the whole batch through the C
    ABI (host buffers in and out), after one warm-up; beside it the oracle's
    rate for one energy on one host thread."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    import math

    from molscat_amd import system as S
    from molscat_amd.engine import Engine

    pot = S.PowerPotential([[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]], rm=3.5, epsil=50.0)
    blk = S.rotor_block(10, 1, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)
    eng = Engine(0)
    eredmx = 1000.0 * blk.cm2ru
    lowest = lambda r: float(np.asarray(eng.eigh(S.hammat_host(blk, r, eredmx)[None], vectors=False)[0]).reshape(-1)[0])
    rstart, rturn = S.findrm(lowest, 0.5, blk.eint, blk.cent, eredmx, 9)
    rmid = 1.4 * rturn
    cays = S.calck(0.0, eredmx, blk.eint)
    drfit = math.pi / (cays * 12.0)
    nfull, drf = S.drset_constant(rstart, 20.0, drfit)
    nmid, drm = S.drset_constant(rstart, rmid, drfit)
    k = args.energies
    ered = np.linspace(900.0, 1100.0, k) * blk.cm2ru
    wv, nops = zip(*[S.gtwvec(e, blk.eint) for e in ered])
    wv, nops = np.array(wv), np.array(nops)
    nb = S.chnsrt(blk.eint)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    eng.set_potl(b, blk.potl, blk.rscale, blk.ep2ru)

    def grid_of(r, dr):
        return eng.grid(r, blk.ptab(r), dr)

    gfull = grid_of(S.ld_radii(rstart, drf, nfull), drf)
    gmid = grid_of(S.ld_radii(rstart, drm, nmid), drm)
    rmg = eng.mg_radii(rstart, drf, nfull, 5)
    gmg = grid_of(rmg, drf)
    tl = 0.064 * math.sqrt(1e-4 / 0.001)

    def match(y, rval):
        _, _, _, st = eng.ytok_ktos(y, wv, nops, nb, blk.cent, rval)
        if np.any(st != 0):
            raise RuntimeError("matching failed")

    def run_ld(iprop):
        out = eng.smatrix_ld(b, gfull, iprop, ered, wv, nops, nb, 20.0)
        if np.any(out["STATUS"] != 0):
            raise RuntimeError(f"iprop {iprop} failed")

    def run_mg():
        y0 = eng.yinit(b, gmg, ered)
        y, _, st = eng.prop_mg(b, gmg, 5, nfull, drf, ered, y0)
        if np.any(st != 0):
            raise RuntimeError("MGPROP failed")
        match(y, 20.0)

    def run_airy():
        y0 = eng.yinit(b, gmid, ered)
        y, _, st = eng.prop_ld(b, gmid, 6, ered, y0)
        z, _, _, st2, _ = eng.prop_airy_tolhi(b, ered, y, rmid, 20.0, drfit, 3.0, 1e-4, replay=2)
        if np.any(st != 0) or np.any(st2 != 0):
            raise RuntimeError("MDPROP + AIRY failed")
        match(z, 20.0)

    def run_vivs():
        y0 = eng.yinit(b, gmid, ered)
        y, _, st = eng.prop_ld(b, gmid, 5, ered, y0)
        rm, _ = eng.syminv(y)
        rr, _, st2, _ = eng.prop_vivs(b, ered, rm, rmid, 20.0, drm, tl, tl, replay=2)
        if np.any(st != 0) or np.any(st2 != 0):
            raise RuntimeError("LDPROP + VIVS failed")
        yv, _ = eng.syminv(rr)
        match(yv, 20.0)

    runs = {"MDPROP": lambda: run_ld(6), "LDPROP": lambda: run_ld(5), "MAPROP": lambda: run_ld(7),
            "MGPROP_MA5": run_mg, "MDPROP+AIRY_TOLHI": run_airy, "LDPROP+VIVS": run_vivs}
    res = {}
    for name, fn in runs.items():
        fn()
        t0 = time.perf_counter()
        fn()
        dt = time.perf_counter() - t0
        res[name] = {"ms": dt * 1e3, "s_matrices_per_s": k / dt}
        print(f"# {name}: {dt * 1e3:.1f} ms for {k} energies", file=sys.stderr, flush=True)
    cpu = {}
    if args.cpu_sectors > 0:
        from oracle import molscat_oracle as O

        ob, _ = O.make_itype1_block(10, 1, [0, 2, 4, 6], 30.0, [0, 2],
                                    O.PowerPotential(pot.terms, rm=pot.rm, epsil=pot.epsil), 20.0)
        e0 = 1000.0 * ob.cm2ru
        full = O.Grid(rstart, 20.0, nfull, drf)
        mid = O.Grid(rstart, rmid, nmid, drm)
        cpu_runs = {"MDPROP": lambda: O.smatrix_ld(ob, e0, full, iprop=6),
                    "LDPROP": lambda: O.smatrix_ld(ob, e0, full, iprop=5),
                    "MAPROP": lambda: O.smatrix_ld(ob, e0, full, iprop=7),
                    "MGPROP_MA5": lambda: O.smatrix_ld(ob, e0, full, iprop=8, imgsel=5),
                    "MDPROP+AIRY_TOLHI": lambda: O.smatrix_md_airy(ob, e0, mid, 20.0, drfit),
                    "LDPROP+VIVS": lambda: O.smatrix_ld_vivs(ob, e0, mid, 20.0)}
        for name, fn in cpu_runs.items():
            t0 = time.perf_counter()
            fn()
            cpu[name] = 1.0 / (time.perf_counter() - t0)
            res[name]["cpu_s_matrices_per_s"] = cpu[name]
    line = {
        "metric": "S-matrices/s per propagator, all-propagators deck (N=12)", "value": res["MDPROP"]["s_matrices_per_s"],
        "unit": "S-matrices/s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": res["MDPROP"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"ITYPE=1 rotor N=12 JTOT=10 IB=1, {k} energies in [900, 1100] cm-1, "
                               f"{nfull} LD sectors (RSTART {rstart:.4f} to 20 RM) or {nmid} to RMID {rmid:.4f} "
                               f"then AIRY TOLHI=1e-4 / VIVS to 20 RM", "energies": k},
        "propagators": res,
        "cpu_baseline": None if not cpu else {
            "value": cpu["MDPROP"], "unit": "S-matrices/s", "cores": 1, "kind": "port",
            "sample": "the oracle, one energy per propagator, single thread (its AIRY sector propagator "
                      "evaluates the Airy functions with mpmath, which dominates that line)"},
    }
    print(json.dumps(line))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sectors", type=int, default=150,
                    help="LD sectors per instance in the CPU baseline sample (0: skip the baseline)")
    ap.add_argument("--cpu-airy", type=int, default=6, help="AIRY steps per instance in the CPU baseline sample")
    ap.add_argument("--cpu-reps", type=int, default=3, help="runs of the CPU baseline sample (the median is kept)")
    ap.add_argument("--team", type=int, default=0, help="workgroups per instance on the large-N path (0: auto)")
    ap.add_argument("--plain-launch", action="store_true",
                    help="ordinary launches for the team kernels (for profilers that cannot trace cooperative ones)")
    ap.add_argument("--inverse-kernel", type=int, default=1,
                    help="batched large-N inversions: 1 32-pivot passes (default), 0 16-pivot passes")
    ap.add_argument("--jstep", type=int, default=30, help="JTOT step of the sweep's stratified subset")
    ap.add_argument("--mixed-steps", type=int, default=2,
                    help="Newton steps of the mixed-precision leg (0: no mixed leg)")
    ap.add_argument("--energies-sweep", type=int, default=64, help="energies per block for --workload sweep")
    ap.add_argument("--tune-file", default="", help="autotuner table (default ~/.cache/molscat_amd/autotune.json)")
    ap.add_argument("--workload", choices=["rotor441", "allprop", "field", "copy", "sweep", "mixed-sweep"], default="rotor441",
                    help="rotor441: the N=441 deck (default); allprop: every propagator on the N=12 deck; "
                         "field: the N=48 field scan")
    ap.add_argument("--bound-fields", type=int, default=128,
                    help="--workload field: fields (an even subset of the scan) whose levels are located")
    ap.add_argument("--fields", type=int, default=4096, help="fields per call for --workload field")
    ap.add_argument("--md-kernel", type=int, default=0, choices=[0, 1, 2],
                    help="MDPROP kernel for N <= 64 (--workload field): 0 automatic, 1 LDS-resident, "
                         "2 register-resident")
    ap.add_argument("--energies", type=int, default=256, help="energies batched per call for --workload allprop")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(independent (J, parity, energy) instances would shard across GPUs with no exchange "
              f"but the final cross-section sum)", file=sys.stderr)
        return 2
    if args.workload == "allprop":
        return _allprop(args)
    if args.workload == "field":
        return _field(args)
    if args.workload == "copy":
        return _copy(args)
    if args.workload == "sweep":
        return _sweep(args)
    if args.workload == "mixed-sweep":
        return _mixed_sweep(args)

    from molscat_amd import decks
    from molscat_amd import system as S
    from molscat_amd.engine import Engine

    deck = decks.rotor441()
    blk = deck.block()
    seg = deck.ld
    n, nstep = blk.n, seg.nstep
    r = seg.radii()
    ered = deck.energies[0] * blk.cm2ru
    wv, nopen = S.gtwvec(ered, blk.eint)
    nb = S.chnsrt(blk.eint)

    ra, wa = deck.airy_grid()
    nairy = len(wa)

    # the CPU baseline first, before this process touches the GPU
    cpu = _cpu_rotor441(blk, deck, ered, args.cpu_sectors, args.cpu_airy, args.cpu_reps) if args.cpu_sectors > 0 else None
    eng = Engine(0)
    eng.set_inverse_kernel(args.inverse_kernel)
    eng.set_team_size(args.team)
    if args.plain_launch:
        eng.set_team_launch(False)
    b = eng.block(blk.vl, blk.eint, blk.cent, mxlam=blk.mxlam)
    g = eng.grid(r, blk.ptab(r), seg.dr)
    ga = eng.airy_grid(ra, wa, blk.ptab(ra))

    def solve():
        out = eng.smatrix_segments(b, [g, ga], [6, 9], [ered], wv, [nopen], nb, deck.airy_stop)
        if out["STATUS"][0] != 0:
            raise RuntimeError(f"solve failed: status {out['STATUS'][0]}")
        return out

    for _ in range(args.warmup):
        solve()
    t_prop, t_sec, t_abas, t_aprop = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = solve()
        tl, ta = eng.last_segment_timing(0), eng.last_segment_timing(1)
        t_prop.append(tl["ms_prop"])
        t_sec.append(tl["ms_sector"])
        t_abas.append(ta["ms_sector"])
        t_aprop.append(ta["ms_prop"])
    wall = time.perf_counter() - t0
    s = out["SR"][0] + 1j * out["SI"][0]
    unit_err = float(np.abs(s @ s.conj().T - np.eye(len(s))).max())

    ms_step = 1e3 * wall / args.steps
    smat_per_s = args.steps / wall
    # the reference algorithm's work per LD sector-step (one instance, B = 1):
    # (2 + 1/B) N^3 + (2/B) NHAM N(N+1) + 12 N^2 (two inversions and the
    # energy-independent U(C) inversion, two W builds, the updates)
    nham = blk.mxlam
    f_ss = 3.0 * n ** 3 + 2.0 * nham * n * (n + 1) + 12.0 * n * n
    flops_alg = nstep * f_ss
    ms_sec = float(np.mean(t_sec))
    ms_prop = float(np.mean(t_prop))
    ms_stage = ms_sec + ms_prop
    achieved = flops_alg / (ms_stage * 1e-3)
    # what the tree executes: one leaf inversion per sector (V + E^-1, no
    # U(C) at this step size), (nstep - 1) merges of one inversion and five
    # n x n products, the final application
    flops_exec = (1.0 * nstep * n ** 3 + (nstep - 1) * 11.0 * n ** 3 + 5.0 * n ** 3)
    line = {
        "metric": METRIC,
        "value": smat_per_s,
        "unit": "S-matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"rotor N={n} JTOT=60 E=300 cm-1, YINIT at 2 A, LDMD 2-20 A DR={seg.dr:.4f} A "
                        f"({nstep} sectors), AIRY 20-{deck.airy_stop:.0f} A ({nairy} steps, TOLHIL=0, R^(4/3)), "
                        f"YTOK/KTOS matching at {deck.airy_stop:.0f} A, one instance per solve",
            "n_channels": n, "sectors": nstep, "airy_steps": nairy, "precision": "fp64", "batch_instances": 1,
        },
        "sector_steps_per_s": nstep * smat_per_s,
        "kernel_ms": {"mdprop_sector_matrices": ms_sec, "mdprop_recurrence": ms_prop,
                      "airy_step_bases": float(np.mean(t_abas)), "airy_propagation": float(np.mean(t_aprop))},
        "unitarity_err": unit_err,
        "roofline": {"bound": "mfma",
                     "kernel": "LDMD propagation stage: k_mdprop_diag + tree (k_tree_m, k_syminv_wide, "
                               "k_gemm_seg, k_tree_*)",
                     "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": LD_STAGE_TRAFFIC_BYTES,
                     "flops_per_launch": flops_alg, "ms_per_launch": ms_stage,
                     "executed_flops": flops_exec,
                     "executed_tflops": flops_exec / (ms_stage * 1e-3) / 1e12},
    }
    if args.mixed_steps > 0:
        # the same solve with the tree's batched inversions in FP32 + Newton
        # steps (FP64 below 3 A, YINIT / recurrence / AIRY steps / matching FP64)
        s64 = out["SR"][0] + 1j * out["SI"][0]
        eng.set_precision(args.mixed_steps, 3.0)
        for _ in range(args.warmup):
            solve()
        eng.set_precision(args.mixed_steps, 3.0)  # statistics of the timed solves only
        t0 = time.perf_counter()
        for _ in range(args.steps):
            outm = solve()
        wall_m = time.perf_counter() - t0
        stats = eng.precision_stats()
        eng.set_precision(0)
        sm = outm["SR"][0] + 1j * outm["SI"][0]
        line["mixed"] = {
            "value": args.steps / wall_m, "unit": "S-matrices/s", "ms_per_step": 1e3 * wall_m / args.steps,
            "ratio_to_fp64": (args.steps / wall_m) / smat_per_s,
            "max_abs_dS_vs_fp64": float(np.abs(sm - s64).max()),
            "nodes_equal": int(outm["NODES"][0]) == int(out["NODES"][0]),
            "newton_steps": args.mixed_steps, "fp64_below_A": 3.0,
            "inversions_mixed_per_solve": stats["mixed_matrices"] / args.steps,
            "inversions_fp64_again_per_solve": stats["fp64_again"] / args.steps,
            "precision": "tree inversions FP32 (matrix cores) + FP64 Newton-Schulz; everything else FP64",
        }
    eng.close()
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
