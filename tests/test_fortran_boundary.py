"""The Fortran side of the boundary: the ISO_C_BINDING module and the drop-in
MDPROP / LDPROP / MAPROP / AIPROP / SYMINV / YTOK / VVPROP / BDCTRL (molscat_amd/fortran),
compiled with amdflang and linked into a test host program
(tests/native/fortran_host.f90) that calls them with the reference's
argument lists (source_code/mdprop.f:1-4, ldprop.f:1-4, maprop.f:1-4,
aiprop.f:1-4, syminv.f:1, ytok.f:1).

CPU: the module's interfaces reach the library (MGPROP's scheme table, as
the ctypes binding sees it) and the propagator entry points refuse a NULL
context.  GPU: one deck (the reference's all-propagators rotor, N=12,
E=1000 cm-1) through every drop-in against the oracle; the second MDPROP
call at the same energy reuses the uploaded block and grid."""
import os
import subprocess

import numpy as np
import pytest

from molscat_amd import build as B
from molscat_amd import system as S

HOST = B.FORTRAN_HOST
POT_TERMS = [[(1.0, -12), (-2.0, -6)], [(-0.2, -6)]]


def _host():
    if not os.path.exists(HOST):
        pytest.fail(f"{HOST} is not built (molscat_amd.build.build_fortran)")
    return HOST


def test_fortran_module_cpu_checks():
    out = subprocess.run([_host(), "cpu"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    lines = {ln.split()[0] + ("" if not ln.startswith("NULLCTX") else " " + ln.split()[1]): ln.split()
             for ln in out.stdout.splitlines() if ln.strip()}
    assert "CPU_CHECKS_DONE" in lines
    from molscat_amd.lib import load

    lib = load()
    import ctypes

    for sel, key in ((4, "MG4"), (5, "MG5")):
        ns = ctypes.c_int32()
        a = np.zeros(8)
        b = np.zeros(8)
        rc = lib.mcp_mg_scheme(sel, ctypes.byref(ns), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               b.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        f = lines[key]
        assert int(f[1]) == rc == 0
        assert int(f[2]) == ns.value
        vals = np.array([float(x) for x in f[3:]])
        assert np.array_equal(vals, np.concatenate([a[:ns.value], b[:ns.value]]))
    assert int(lines["MG3"][1]) == -1  # no scheme 3
    for k in ("NULLCTX syminv", "NULLCTX block_create", "NULLCTX prop_ld", "NULLCTX bound_match"):
        assert int(lines[k][2]) == -1, k  # MCP_ERR_ARG


def _write_deck(path, blk, terms, ered, rstart, rmid, nstep, dr, nstepa, drnow, powr, rmax, y0, nopen, nb, wvec):
    n = blk.n
    with open(path, "w") as f:
        f.write(f"{n} {blk.mxlam} {blk.mxlam}\n")
        f.write(" ".join(repr(float(x)) for x in np.asarray(blk.vl).T.ravel()) + "\n")
        f.write(" ".join(repr(float(x)) for x in blk.eint) + "\n")
        f.write(" ".join(repr(float(x)) for x in blk.cent) + "\n")
        tl = [(i + 1, t[1], t[0], t[2] if len(t) > 2 else 0.0) for i, lt in enumerate(terms) for t in lt]
        f.write(f"{len(tl)}\n")
        for lam, pw, a, e in tl:
            f.write(f"{lam} {pw} {a!r} {e!r}\n")
        f.write("1.0\n")
        f.write(f"{blk.ep2ru!r} {blk.cm2ru!r} {blk.rscale!r}\n")
        f.write(f"{ered!r}\n")
        f.write(f"{rstart!r} {rmid!r} {nstep} {dr!r}\n")
        f.write(f"{nstepa} {drnow!r} {powr!r} {rmax!r}\n")
        f.write(" ".join(repr(float(x)) for x in np.asarray(y0).ravel(order="F")) + "\n")
        f.write(f"{nopen}\n")
        f.write(" ".join(str(int(i) + 1) for i in nb) + "\n")
        f.write(" ".join(repr(float(x)) for x in wvec) + "\n")


def _read_out(path, n):
    secs = {}
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    i = 0
    while i < len(lines):
        head = lines[i]
        key = head[0]
        i += 1
        vals = []
        while i < len(lines) and not lines[i][0].isalpha():
            vals.append(float(lines[i][0].replace("D", "E")))
            i += 1
        secs[key] = (head[1:], np.array(vals))
    return secs


@pytest.mark.gpu
def test_fortran_dropins_vs_oracle(tmp_path):
    from oracle import molscat_oracle as O

    pot = S.PowerPotential(POT_TERMS, rm=3.5, epsil=50.0)
    blk = S.rotor_block(10, 1, [0, 2, 4, 6], 30.0, [0, 2], pot, 20.0)
    opot = O.PowerPotential(pot.terms, rm=pot.rm, epsil=pot.epsil)
    ob = O.Block(n=blk.n, vl=blk.vl, eint=blk.eint, cent=blk.cent, lvals=blk.lvals, levidx=blk.levidx, potl=opot,
                 rscale=blk.rscale, ep2ru=blk.ep2ru, cm2ru=blk.cm2ru)
    n = blk.n
    ered = 1000.0 * blk.cm2ru
    rstart, rmid, nstep = 0.641671863671356, 8.0, 300
    dr = (rmid - rstart) / nstep
    rmax, powr = 12.0, 4.0 / 3.0
    nstepa, drnow = S.drset_airy(rmid, rmax, 0.1, powr)
    wvec, nopen = S.gtwvec(ered, blk.eint)
    nb = S.chnsrt(blk.eint)
    y0 = O.yinit(ob, ered, rstart)
    deck, res = str(tmp_path / "deck.txt"), str(tmp_path / "out.txt")
    _write_deck(deck, blk, POT_TERMS, ered, rstart, rmid, nstep, dr, nstepa, drnow, powr, rmax, y0, nopen, nb, wvec)
    print(f"fortran drop-ins: N={n} LD {rstart:.3f}-{rmid} ({nstep}) AIRY {rmid}-{rmax} ({nstepa})", flush=True)
    run = subprocess.run([_host(), "run", deck, res], capture_output=True, text=True, timeout=300)
    print(run.stdout[-2000:], run.stderr[-2000:], flush=True)
    assert run.returncode == 0 and "RUN_DONE" in run.stdout
    out = _read_out(res, n)

    # uploads: one block and one grid for both MDPROP calls
    assert [int(x) for x in out["CACHE"][0]] == [1, 1, 1, 1]
    hm = lambda r: ob.hammat(r, ered)
    ymd, nmd = O.mdprop(y0.copy(), hm, rstart, nstep, dr)
    head, y = out["MDPROP"]
    y = y.reshape(n, n, order="F")
    assert int(head[0]) == int(head[1]) == nmd
    assert float(head[2]) == 0.0  # the second call (cached uploads) is bitwise the first
    err = np.abs(y - ymd).max() / np.abs(ymd).max()
    print(f"  MDPROP nodes={nmd} rel|dY|={err:.2e}", flush=True)
    assert err < 1e-10

    yld, nld = O.ldprop(y0.copy(), hm, rstart, nstep, dr)
    head, z = out["LDPROP"]
    err = np.abs(z.reshape(n, n, order="F") - yld).max() / np.abs(yld).max()
    print(f"  LDPROP nodes={nld} rel|dY|={err:.2e}", flush=True)
    assert int(head[0]) == nld and err < 1e-9

    yma, nma = O.maprop(y0.copy(), ob, ered, rstart, nstep, dr)
    head, z = out["MAPROP"]
    err = np.abs(z.reshape(n, n, order="F") - yma).max() / np.abs(yma).max()
    print(f"  MAPROP nodes={nma} rel|dY|={err:.2e}", flush=True)
    assert int(head[0]) == nma and err < 1e-9

    za, na = O.aiprop(y.copy(), ob, ered, rmid, rmax, nstepa, drnow, powr)
    head, z = out["AIPROP"]
    z = z.reshape(n, n, order="F")
    err = np.abs(z - za).max() / np.abs(za).max()
    print(f"  AIPROP nodes={na} steps={head[1]} rel|dZ|={err:.2e}", flush=True)
    assert int(head[0]) == na and int(head[1]) == nstepa and err < 1e-10

    kref = O.ytok(nb, wvec, n, nopen, z, rmax, blk.cent)
    head, k = out["YTOK"]
    assert int(head[0]) == nopen
    err = np.abs(k.reshape(nopen, nopen, order="F") - kref).max() / max(1.0, np.abs(kref).max())
    print(f"  YTOK NOPEN={nopen} rel|dK|={err:.2e}", flush=True)
    assert err < 1e-10

    head, ainv = out["SYMINV"]
    ainv = ainv.reshape(n, n, order="F")
    assert int(head[0]) == int((np.linalg.eigvalsh(y) < 0).sum())
    inv = np.linalg.inv(y)
    err = np.abs(ainv - inv).max() / np.abs(inv).max()
    print(f"  SYMINV IFAIL={int(head[0])} rel|dA|={err:.2e}", flush=True)
    assert err < 1e-11

    # VVPROP under IDIAG with DELRD steps (IALPHA = 0), RMID -> RMAX, from
    # the inverse just checked; its IREAD replay of the kept steps is bitwise
    # the IWRITE call
    opts = O.vivs_idiag_defaults(O.VivsOptions(idiag=True, ialpha=0))
    rv, steps = O.vvprop(ainv, ob, ered, rmid, rmax, 0.002, 0.064, 0.064, drmax=5.0, opts=opts)
    head, r = out["VVPROP"]
    r = r.reshape(n, n, order="F")
    err = np.abs(r - rv).max() / np.abs(rv).max()
    print(f"  VVPROP steps={int(head[0])} (oracle {len(steps)}) rel|dR|={err:.2e}", flush=True)
    assert int(head[0]) == int(head[1]) == len(steps)
    assert float(head[2]) == 0.0
    assert err < 1e-10

    # BDCTRL: MDPROP out to RMID with DRS = DR and in from RMAX with DRL = 2 DR
    (_, dro, no), (_, dri, ni) = S.bound_segments(rstart, rmid, rmax, dr, 2.0 * dr)
    og = O.BoundGrid(rstart, rmid, rmax, no, dro, ni, dri)
    m = O.bound_match(ob, ered, og, 6, 6)
    head, psi = out["BDCTRL"]
    print(f"  BDCTRL NODEC={int(head[0])} (oracle {m['NODEC']}) EIGMIN={float(head[2]):.6e} "
          f"(oracle {m['EIGMIN']:.6e})", flush=True)
    assert int(head[0]) == m["NODEC"] and int(head[1]) == m["IMIN"] + 1
    scale = max(1.0, float(np.abs(m["EVALS"]).max()))
    assert abs(float(head[2]) - m["EIGMIN"]) < 1e-9 * scale
    assert np.abs(psi - m["PSIMCH"]).max() < 1e-8
