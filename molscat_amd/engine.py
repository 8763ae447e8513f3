"""Host-side handle on the MI355X engine: one HIP context, device-resident
blocks (VL, EINT, CENT) and radial grids (R, P table), and the batched entry
points of the C ABI with numpy arrays.

Matrices cross the boundary column-major; numpy arrays here are
(batch, N, N) with element [m, i, j] = A_m(i, j), stored so that the
column-major view of each N x N slab is the matrix itself (the arrays are
transposed copies of C-order data where needed).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import lib as _lib
from .lib import McpError, dptr, iptr


def _colmajor(batch: np.ndarray) -> np.ndarray:
    """(m, N, N) logical matrices -> contiguous buffer whose m-th slab is the
    column-major storage of batch[m]."""
    return np.ascontiguousarray(np.transpose(batch, (0, 2, 1)), dtype=np.float64)


def _from_colmajor(buf: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.transpose(buf, (0, 2, 1)))


@dataclass
class BlockHandle:
    id: int
    n: int
    mxlam: int
    nconst: int
    nrsq: int
    cent: np.ndarray


@dataclass
class GridHandle:
    id: int
    r: np.ndarray
    mxlam: int


class Engine:
    """One MI355X device context (mcp_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self.lib.mcp_ctx_create(device, ctypes.byref(h))
        if rc != 0 or not h.value:
            raise McpError(f"mcp_ctx_create(device={device}) failed with code {rc} (is a gfx950 GPU visible?)")
        self.ctx = h

    def close(self):
        if getattr(self, "ctx", None) is not None and self.ctx.value:
            self.lib.mcp_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.mcp_last_error(self.ctx)
            raise McpError(f"{what}: code {rc}: {msg.decode() if msg else ''}")

    def info(self) -> dict:
        v = [ctypes.c_int32() for _ in range(5)]
        self._check(self.lib.mcp_ctx_info(self.ctx, *[ctypes.byref(x) for x in v]), "mcp_ctx_info")
        return dict(max_n_md=v[0].value, max_n_ld=v[1].value, max_n_eig=v[2].value, lds_bytes=v[3].value,
                    num_cu=v[4].value)

    # -- dense kernels ------------------------------------------------------
    def syminv(self, a: np.ndarray):
        """Batched SYMINV: a (m, N, N) symmetric (lower triangle read).
        Returns (inverse (m, N, N), kount (m,))."""
        a = np.asarray(a, dtype=np.float64)
        single = a.ndim == 2
        if single:
            a = a[None]
        m, n, _ = a.shape
        buf = _colmajor(a)
        out = np.empty_like(buf)
        k = np.empty(m, dtype=np.int32)
        self._check(self.lib.mcp_syminv_batched(self.ctx, n, m, dptr(buf), dptr(out), iptr(k)), "mcp_syminv_batched")
        inv = _from_colmajor(out)
        return (inv[0], int(k[0])) if single else (inv, k)

    def eigh(self, a: np.ndarray, vectors: bool = True):
        a = np.asarray(a, dtype=np.float64)
        single = a.ndim == 2
        if single:
            a = a[None]
        m, n, _ = a.shape
        buf = _colmajor(a)
        ev = np.empty((m, n))
        vec = np.empty_like(buf) if vectors else None
        self._check(self.lib.mcp_eigh_batched(self.ctx, n, m, dptr(buf), dptr(ev), dptr(vec)), "mcp_eigh_batched")
        v = _from_colmajor(vec) if vectors else None
        if single:
            return ev[0], (v[0] if vectors else None)
        return ev, v

    # -- device-resident data ----------------------------------------------
    def block(self, vl: np.ndarray, eint: np.ndarray, cent: np.ndarray, mxlam: int, nconst: int = 0,
              nrsq: int = 0, conlen: float = 0.0) -> BlockHandle:
        """Upload VL(NHAM, N(N+1)/2) (numpy shape (NHAM, ntri)), EINT, CENT."""
        vl = np.asarray(vl, dtype=np.float64)
        n = len(eint)
        nham = mxlam + nconst + nrsq
        if vl.shape != (nham, n * (n + 1) // 2):
            raise ValueError(f"VL shape {vl.shape}, expected {(nham, n * (n + 1) // 2)}")
        vlf = np.ascontiguousarray(vl.T)  # [ntri][nham]: lambda fastest
        bid = ctypes.c_int32()
        self._check(self.lib.mcp_block_create(self.ctx, n, mxlam, nconst, nrsq, dptr(vlf),
                                              dptr(np.ascontiguousarray(eint, dtype=np.float64)),
                                              dptr(np.ascontiguousarray(cent, dtype=np.float64)), float(conlen),
                                              ctypes.byref(bid)), "mcp_block_create")
        return BlockHandle(bid.value, n, mxlam, nconst, nrsq, np.array(cent, dtype=np.float64))

    def couple_rotor(self, jch, lch, jtot: int, lambdas) -> np.ndarray:
        """COUPLE for ITYPE = 1 on the device (mcp_couple_rotor): VL of the
        channels (j, L) = (jch[i], lch[i]) at total J, shape (NLAM, N(N+1)/2)
        as RotorBlock.vl."""
        jch = np.ascontiguousarray(jch, dtype=np.int32)
        lch = np.ascontiguousarray(lch, dtype=np.int32)
        lam = np.ascontiguousarray(lambdas, dtype=np.int32)
        n = len(jch)
        out = np.empty((n * (n + 1) // 2, len(lam)))
        self._check(self.lib.mcp_couple_rotor(self.ctx, n, iptr(jch), iptr(lch), int(jtot), len(lam), iptr(lam),
                                              dptr(out)), "mcp_couple_rotor")
        return np.ascontiguousarray(out.T)

    def block_rotor(self, jch, lch, jtot: int, lambdas, eint, cent, conlen: float = 0.0) -> BlockHandle:
        """mcp_block_create_rotor: an ITYPE = 1 block whose VL is made on the
        device from its channels (j, L) = (jch[i], lch[i])."""
        jch = np.ascontiguousarray(jch, dtype=np.int32)
        lch = np.ascontiguousarray(lch, dtype=np.int32)
        lam = np.ascontiguousarray(lambdas, dtype=np.int32)
        n = len(jch)
        bid = ctypes.c_int32()
        self._check(self.lib.mcp_block_create_rotor(self.ctx, n, iptr(jch), iptr(lch), int(jtot), len(lam),
                                                    iptr(lam), dptr(np.ascontiguousarray(eint, dtype=np.float64)),
                                                    dptr(np.ascontiguousarray(cent, dtype=np.float64)),
                                                    float(conlen), ctypes.byref(bid)), "mcp_block_create_rotor")
        return BlockHandle(bid.value, n, len(lam), 0, 0, np.array(cent, dtype=np.float64))

    def grid(self, r: np.ndarray, ptab: np.ndarray, dr: float = 0.0) -> GridHandle:
        """Upload radii r (npts,), the P(lambda, R) table ptab (npts, MXLAM) and,
        for an LD grid, its step DR."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        ptab = np.ascontiguousarray(ptab, dtype=np.float64)
        mxlam = ptab.shape[1] if ptab.ndim == 2 else 0
        gid = ctypes.c_int32()
        self._check(self.lib.mcp_grid_create(self.ctx, mxlam, len(r), dptr(r), dptr(ptab), float(dr),
                                             ctypes.byref(gid)),
                    "mcp_grid_create")
        return GridHandle(gid.value, r, mxlam)

    def airy_grid(self, r: np.ndarray, drnow: np.ndarray, ptab: np.ndarray) -> GridHandle:
        """Upload an AIRY grid: r = (RSTART, RA_0, RB_0, ...), step widths
        drnow, P table at r (see system.airy_grid)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        drnow = np.ascontiguousarray(drnow, dtype=np.float64)
        ptab = np.ascontiguousarray(ptab, dtype=np.float64)
        mxlam = ptab.shape[1] if ptab.ndim == 2 else 0
        gid = ctypes.c_int32()
        self._check(self.lib.mcp_grid_create_airy(self.ctx, mxlam, len(drnow), dptr(r), dptr(drnow), dptr(ptab),
                                                  ctypes.byref(gid)), "mcp_grid_create_airy")
        return GridHandle(gid.value, r, mxlam)

    def free_block(self, b: BlockHandle):
        self._check(self.lib.mcp_block_destroy(self.ctx, b.id), "mcp_block_destroy")

    def free_grid(self, g: GridHandle):
        self._check(self.lib.mcp_grid_destroy(self.ctx, g.id), "mcp_grid_destroy")

    # -- path entry points ----------------------------------------------------
    def hammat(self, b: BlockHandle, g: GridHandle, idx, ered: float, pconst=None) -> np.ndarray:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        w = np.empty((len(idx), b.n, b.n))
        pc = None if pconst is None else np.ascontiguousarray(pconst, dtype=np.float64)
        self._check(self.lib.mcp_hammat(self.ctx, b.id, g.id, len(idx), iptr(idx), dptr(pc), float(ered), dptr(w)),
                    "mcp_hammat")
        return _from_colmajor(w)

    def yinit(self, b: BlockHandle, g: GridHandle, ered, k0: int = 0, outward: bool = True, bcyo: float = 0.0,
              pconst=None) -> np.ndarray:
        ered = np.ascontiguousarray(np.atleast_1d(ered), dtype=np.float64)
        m = len(ered)
        y = np.empty((m, b.n, b.n))
        pc = None if pconst is None else np.ascontiguousarray(pconst, dtype=np.float64)
        self._check(self.lib.mcp_yinit(self.ctx, b.id, g.id, k0, 1 if outward else 0, float(bcyo), m, dptr(ered),
                                       dptr(pc), dptr(y)), "mcp_yinit")
        return _from_colmajor(y)

    def prop_ld(self, b: BlockHandle, g: GridHandle, iprop: int, ered, y0: np.ndarray, pconst=None):
        """Propagate ninst = len(ered) instances from y0 (m, N, N); returns
        (Y(RSTOP), nodes, status)."""
        ered = np.ascontiguousarray(np.atleast_1d(ered), dtype=np.float64)
        m = len(ered)
        y = _colmajor(np.asarray(y0, dtype=np.float64).reshape(m, b.n, b.n))
        nodes = np.empty(m, dtype=np.int32)
        st = np.empty(m, dtype=np.int32)
        pc = None if pconst is None else np.ascontiguousarray(pconst, dtype=np.float64)
        self._check(self.lib.mcp_prop_ld(self.ctx, b.id, g.id, iprop, m, dptr(ered), dptr(pc), dptr(y), iptr(nodes),
                                         iptr(st)), "mcp_prop_ld")
        return _from_colmajor(y), nodes, st

    def prop_ld_multi(self, blocks, g: GridHandle, ereds, y0s):
        """MDPROP for several blocks of uneven N in one persistent launch:
        blocks[b] with energies ereds[b] and starting matrices y0s[b]
        (m_b, N_b, N_b).  Returns [(Y(RSTOP), nodes, status) per block]."""
        nblk = len(blocks)
        ids = np.ascontiguousarray([b.id for b in blocks], dtype=np.int32)
        ninst = np.ascontiguousarray([len(np.atleast_1d(e)) for e in ereds], dtype=np.int32)
        ered = np.ascontiguousarray(np.concatenate([np.atleast_1d(e) for e in ereds]), dtype=np.float64)
        y = np.ascontiguousarray(np.concatenate([_colmajor(np.asarray(yy, dtype=np.float64).reshape(-1, b.n, b.n)).ravel()
                                                 for yy, b in zip(y0s, blocks)]))
        tot = int(ninst.sum())
        nodes = np.empty(tot, dtype=np.int32)
        st = np.empty(tot, dtype=np.int32)
        self._check(self.lib.mcp_prop_ld_multi(self.ctx, nblk, iptr(ids), g.id, iptr(ninst), dptr(ered), dptr(y),
                                               iptr(nodes), iptr(st)), "mcp_prop_ld_multi")
        out, yo, io = [], 0, 0
        for b, m in zip(blocks, ninst):
            m = int(m)
            cnt = m * b.n * b.n
            out.append((_from_colmajor(y[yo:yo + cnt].reshape(m, b.n, b.n)), nodes[io:io + m].copy(),
                        st[io:io + m].copy()))
            yo += cnt
            io += m
        return out

    def set_graph(self, on: bool):
        """mcp_prop_ld_multi through a captured HIP graph (replayed for the same key)."""
        self._check(self.lib.mcp_set_graph(self.ctx, 1 if on else 0), "mcp_set_graph")

    def prop_airy(self, b: BlockHandle, g: GridHandle, ered, y0: np.ndarray):
        """AIPROP over an AIRY grid for ninst = len(ered) instances; returns
        (Z at the grid's end, nodes, status)."""
        ered = np.ascontiguousarray(np.atleast_1d(ered), dtype=np.float64)
        m = len(ered)
        y = _colmajor(np.asarray(y0, dtype=np.float64).reshape(m, b.n, b.n))
        nodes = np.empty(m, dtype=np.int32)
        st = np.empty(m, dtype=np.int32)
        self._check(self.lib.mcp_prop_airy(self.ctx, b.id, g.id, m, dptr(ered), dptr(y), iptr(nodes), iptr(st)),
                    "mcp_prop_airy")
        return _from_colmajor(y), nodes, st

    def ytrans(self, b: BlockHandle, pconst, y0: Optional[np.ndarray] = None):
        """YTRANS for a block with NCONST > 0: the asymptotic thresholds
        (m, N) of each instance's constants pconst (m, nconst) and, with y0
        (m, N, N), Y in the asymptotic basis.  Returns (eint, y or None)."""
        pc = np.ascontiguousarray(np.atleast_2d(pconst), dtype=np.float64)
        m = pc.shape[0]
        eint = np.empty((m, b.n))
        buf = None if y0 is None else _colmajor(np.asarray(y0, dtype=np.float64).reshape(m, b.n, b.n))
        self._check(self.lib.mcp_ytrans(self.ctx, b.id, m, dptr(pc), dptr(buf), dptr(eint)), "mcp_ytrans")
        return eint, (None if buf is None else _from_colmajor(buf))

    def ytrans_cent(self, b: BlockHandle, pconst=None, y0: Optional[np.ndarray] = None, ninst: int = 1):
        """YTRANS for a block with NCONST > 0 and/or NRSQ > 0: (eint (m, N),
        cent (m, N), y or None) in the asymptotic basis; pconst (m, nconst)
        or None when NCONST = 0 (then m = ninst)."""
        pc = None if pconst is None else np.ascontiguousarray(np.atleast_2d(pconst), dtype=np.float64)
        m = pc.shape[0] if pc is not None else int(ninst)
        eint = np.empty((m, b.n))
        cent = np.empty((m, b.n))
        buf = None if y0 is None else _colmajor(np.asarray(y0, dtype=np.float64).reshape(m, b.n, b.n))
        self._check(self.lib.mcp_ytrans_cent(self.ctx, b.id, m, dptr(pc), dptr(buf), dptr(eint), dptr(cent)),
                    "mcp_ytrans_cent")
        return eint, cent, (None if buf is None else _from_colmajor(buf))

    def set_degtol(self, degtol: float):
        """YTRANS's DEGTOL (1e-10 by default)."""
        self._check(self.lib.mcp_set_degtol(self.ctx, float(degtol)), "mcp_set_degtol")

    def smatrix_segments(self, b: BlockHandle, grids, iprops, ered, wvec, nopen, nb, rval: float, pconst=None):
        """SCCTRL over several segments (grids with their propagator codes)."""
        ered = np.ascontiguousarray(np.atleast_1d(ered), dtype=np.float64)
        m = len(ered)
        n = b.n
        gids = np.ascontiguousarray([g.id for g in grids], dtype=np.int32)
        ips = np.ascontiguousarray(iprops, dtype=np.int32)
        nopen = np.ascontiguousarray(np.broadcast_to(nopen, (m,)), dtype=np.int32)
        nb = np.ascontiguousarray(np.broadcast_to(nb, (m, n)), dtype=np.int32)
        wv = np.ascontiguousarray(np.broadcast_to(wvec, (m, n)), dtype=np.float64)
        k = np.zeros((m, n * n))
        sr = np.zeros((m, n * n))
        si = np.zeros((m, n * n))
        nodes = np.empty(m, dtype=np.int32)
        st = np.empty(m, dtype=np.int32)
        pc = None if pconst is None else np.ascontiguousarray(pconst, dtype=np.float64)
        self._check(self.lib.mcp_smatrix_segments(self.ctx, b.id, len(gids), iptr(gids), iptr(ips), m, dptr(ered),
                                                  dptr(pc), dptr(wv), iptr(nopen), iptr(nb), float(rval), dptr(k),
                                                  dptr(sr), dptr(si), iptr(nodes), iptr(st)),
                    "mcp_smatrix_segments")
        return dict(K=_unpack_square(k, nopen), SR=_unpack_square(sr, nopen), SI=_unpack_square(si, nopen),
                    NODES=nodes, STATUS=st)

    def ytok_ktos(self, y: np.ndarray, wvec: np.ndarray, nopen, nb: np.ndarray, cent: np.ndarray, rval: float):
        y = np.asarray(y, dtype=np.float64)
        m, n, _ = y.shape
        nopen = np.ascontiguousarray(np.atleast_1d(nopen), dtype=np.int32)
        nb = np.ascontiguousarray(np.broadcast_to(nb, (m, n)), dtype=np.int32)
        wv = np.ascontiguousarray(np.broadcast_to(wvec, (m, n)), dtype=np.float64)
        k = np.zeros((m, n * n))
        sr = np.zeros((m, n * n))
        si = np.zeros((m, n * n))
        st = np.empty(m, dtype=np.int32)
        self._check(self.lib.mcp_ytok_ktos(self.ctx, n, m, dptr(_colmajor(y)), dptr(wv), iptr(nopen), iptr(nb),
                                           dptr(np.ascontiguousarray(cent, dtype=np.float64)), float(rval), dptr(k),
                                           dptr(sr), dptr(si), iptr(st)), "mcp_ytok_ktos")
        return _unpack_square(k, nopen), _unpack_square(sr, nopen), _unpack_square(si, nopen), st

    def smatrix_ld(self, b: BlockHandle, g: GridHandle, iprop: int, ered, wvec, nopen, nb, rval: float,
                   pconst=None):
        """YINIT + propagation + YTOK/KTOS on the device; rval = nominal RSTOP."""
        ered = np.ascontiguousarray(np.atleast_1d(ered), dtype=np.float64)
        m = len(ered)
        n = b.n
        nopen = np.ascontiguousarray(np.broadcast_to(nopen, (m,)), dtype=np.int32)
        nb = np.ascontiguousarray(np.broadcast_to(nb, (m, n)), dtype=np.int32)
        wv = np.ascontiguousarray(np.broadcast_to(wvec, (m, n)), dtype=np.float64)
        k = np.zeros((m, n * n))
        sr = np.zeros((m, n * n))
        si = np.zeros((m, n * n))
        nodes = np.empty(m, dtype=np.int32)
        st = np.empty(m, dtype=np.int32)
        pc = None if pconst is None else np.ascontiguousarray(pconst, dtype=np.float64)
        self._check(self.lib.mcp_smatrix_ld(self.ctx, b.id, g.id, iprop, m, dptr(ered), dptr(pc), dptr(wv),
                                            iptr(nopen), iptr(nb), float(rval), dptr(k), dptr(sr), dptr(si),
                                            iptr(nodes),
                                            iptr(st)), "mcp_smatrix_ld")
        return dict(K=_unpack_square(k, nopen), SR=_unpack_square(sr, nopen), SI=_unpack_square(si, nopen),
                    NODES=nodes, STATUS=st)

    def bound_match(self, b: BlockHandle, g_out: GridHandle, g_in: GridHandle, ered, iprop_out: int = 6,
                    iprop_in: int = 6, pconst=None):
        """BOUND matching for ninst = len(ered) energies: dict with NODEC,
        EIGMIN, IMIN (0-based), PSIMCH (m, N), EVALS (m, N), STATUS."""
        ered = np.ascontiguousarray(np.atleast_1d(ered), dtype=np.float64)
        m = len(ered)
        nodec = np.empty(m, dtype=np.int32)
        eigmin = np.empty(m)
        imin = np.empty(m, dtype=np.int32)
        psi = np.empty((m, b.n))
        ev = np.empty((m, b.n))
        st = np.empty(m, dtype=np.int32)
        pc = None if pconst is None else np.ascontiguousarray(pconst, dtype=np.float64)
        self._check(self.lib.mcp_bound_match(self.ctx, b.id, g_out.id, g_in.id, iprop_out, iprop_in, m, dptr(ered),
                                             dptr(pc), iptr(nodec), dptr(eigmin), iptr(imin), dptr(psi), dptr(ev),
                                             iptr(st)), "mcp_bound_match")
        return dict(NODEC=nodec, EIGMIN=eigmin, IMIN=imin, PSIMCH=psi, EVALS=ev, STATUS=st)

    def bound_wave(self, b: BlockHandle, g_out: GridHandle, g_in: GridHandle, ered: float, pconst=None):
        """BOUND's wavefunction at one energy (MDPROP both ways, WAVE records
        and WVBKSB on the device): dict with NODEC, EIGMIN, PSIMCH, PSI_OUT
        ((NSTEP_out + 1, N), row k k steps from RMATCH towards RMIN), PSI_IN
        (towards RMAX), STATUS; unnormalised (row 0 = PSIMCH)."""
        nout = (len(g_out.r) - 1) // 2
        nin = (len(g_in.r) - 1) // 2
        nodec = np.empty(1, dtype=np.int32)
        eigmin = np.empty(1)
        psi = np.empty(b.n)
        po = np.empty((nout + 1, b.n))
        pi = np.empty((nin + 1, b.n))
        st = np.empty(1, dtype=np.int32)
        pc = None if pconst is None else np.ascontiguousarray(pconst, dtype=np.float64)
        self._check(self.lib.mcp_bound_wave(self.ctx, b.id, g_out.id, g_in.id, float(ered), dptr(pc), iptr(nodec),
                                            dptr(eigmin), dptr(psi), dptr(po), dptr(pi), iptr(st)), "mcp_bound_wave")
        return dict(NODEC=int(nodec[0]), EIGMIN=float(eigmin[0]), PSIMCH=psi, PSI_OUT=po, PSI_IN=pi,
                    STATUS=int(st[0]))

    def set_team_size(self, g: int):
        """Workgroups per instance on the large-N path (0 = automatic)."""
        self._check(self.lib.mcp_set_team_size(self.ctx, int(g)), "mcp_set_team_size")

    def set_large_mode(self, mode: int):
        """Large-N MDPROP: 0 tree over the radial grid where memory allows
        (default), 1 sequential matrix-core kernel, 2 global-memory kernel;
        plus 4: DIAGVC and AIRY on the global-memory eigen-solver for every
        N."""
        self._check(self.lib.mcp_set_large_mode(self.ctx, int(mode)), "mcp_set_large_mode")

    def destroy_block(self, b: BlockHandle):
        self._check(self.lib.mcp_block_destroy(self.ctx, b.id), "mcp_block_destroy")

    def destroy_grid(self, g: GridHandle):
        self._check(self.lib.mcp_grid_destroy(self.ctx, g.id), "mcp_grid_destroy")

    def set_potl(self, b: BlockHandle, potl, rscale: float, ep2ru: float, scalam: float = 1.0):
        """POTENL's radial terms for the propagators that choose their radii
        as they go (VIVS, AIRY with TOLHI > 0): potl.terms[lam] = [(A, n) |
        (A, 0, E), ...] (A x^n or A exp(E x) at x = R * rscale), P = ep2ru *
        scalam * the sum clamped at 1e60.  SCAPOT's SCALAM (source_code/
        scapot-default.f:19) multiplies the clamped radial terms before EP2RU
        (source_code/hammat.f:55-73) and DERMAT's derivatives alike
        (source_code/dermat.f:86), so it travels folded into the library's
        EP2RU factor."""
        lam, npw, a, e = [], [], [], []
        for l, tl in enumerate(potl.terms):
            for t in tl:
                lam.append(l)
                npw.append(int(t[1]))
                a.append(float(t[0]))
                e.append(float(t[2]) if int(t[1]) == 0 else 0.0)
        lam = np.ascontiguousarray(lam, dtype=np.int32)
        npw = np.ascontiguousarray(npw, dtype=np.int32)
        a = np.ascontiguousarray(a, dtype=np.float64)
        e = np.ascontiguousarray(e, dtype=np.float64)
        self._check(self.lib.mcp_block_set_potl(self.ctx, b.id, len(lam), iptr(lam), iptr(npw), dptr(a), dptr(e),
                                                float(rscale), float(ep2ru) * float(scalam)), "mcp_block_set_potl")

    def prop_vivs(self, b: BlockHandle, ered, rmat0: np.ndarray, rstart: float, rstop: float, drnow: float,
                  tldiag: float, toff: float, drmax: float = 5.0, replay: int = 2, steps=None,
                  max_steps: int = 200000):
        """VVPROP for ninst = len(ered) instances: R(RSTART) -> R(RSTOP).
        replay 0: each instance its own steps; 1: every instance follows the
        record `steps` (max_steps x 4); 2: the first instance chooses the
        steps, the others follow them.  Returns (R, nstep, status, record)
        with the record of the steps chosen (replay 0 and 2) or given."""
        ered = np.ascontiguousarray(np.atleast_1d(ered), dtype=np.float64)
        m = len(ered)
        r = _colmajor(np.asarray(rmat0, dtype=np.float64).reshape(m, b.n, b.n))
        if steps is None:
            rec = np.zeros((max_steps, 4))
        else:
            rec = np.zeros((max_steps, 4))
            s = np.asarray(steps, dtype=np.float64).reshape(-1, 4)
            rec[: len(s)] = s
        nstep = np.empty(m, dtype=np.int32)
        st = np.empty(m, dtype=np.int32)
        self._check(self.lib.mcp_prop_vivs(self.ctx, b.id, m, dptr(ered), dptr(r), float(rstart), float(rstop),
                                           float(drnow), float(drmax), float(tldiag), float(toff), int(replay),
                                           int(max_steps), dptr(rec), iptr(nstep), iptr(st)), "mcp_prop_vivs")
        k = int(nstep.max()) if replay != 1 else len(rec)
        return _from_colmajor(r), nstep, st, rec[:k].copy()

    def prop_airy_tolhi(self, b: BlockHandle, ered, z0: np.ndarray, rstart: float, rstop: float, dr0: float,
                        powr: float, tolhi: float, replay: int = 2, steps=None, max_steps: int = 500000):
        """AIPROP with TOLHI > 0 for ninst = len(ered) instances: Z(RSTART)
        -> Z(RSTOP).  replay 0: each instance its own steps; 1: every
        instance follows `steps` ((k, 2) RNOW, DRNOW); 2: the first instance
        chooses the steps, the others follow them.  Returns (Z, nodes,
        nstep, status, the (k, 2) record of the steps chosen or given)."""
        ered = np.ascontiguousarray(np.atleast_1d(ered), dtype=np.float64)
        m = len(ered)
        z = _colmajor(np.asarray(z0, dtype=np.float64).reshape(m, b.n, b.n))
        rec = np.zeros((max_steps, 2))
        nrec = ctypes.c_int32(0)
        if steps is not None:
            s = np.asarray(steps, dtype=np.float64).reshape(-1, 2)
            rec[: len(s)] = s
            nrec.value = len(s)
        nodes = np.empty(m, dtype=np.int32)
        nstep = np.empty(m, dtype=np.int32)
        st = np.empty(m, dtype=np.int32)
        self._check(self.lib.mcp_prop_airy_tolhi(self.ctx, b.id, m, dptr(ered), dptr(z), float(rstart),
                                                 float(rstop), float(dr0), float(powr), float(tolhi), int(replay),
                                                 int(max_steps), dptr(rec), ctypes.byref(nrec), iptr(nodes),
                                                 iptr(nstep), iptr(st)), "mcp_prop_airy_tolhi")
        return _from_colmajor(z), nodes, nstep, st, rec[: nrec.value].copy()

    def mg_scheme(self, imgsel: int):
        """(A, B) of MGPROP's scheme IMGSEL (4: CS4, 5: MA5)."""
        m = ctypes.c_int32()
        a = np.empty(6)
        bb = np.empty(6)
        if self.lib.mcp_mg_scheme(int(imgsel), ctypes.byref(m), dptr(a), dptr(bb)) != 0:
            raise ValueError(f"IMGSEL {imgsel}: 4 (CS4) or 5 (MA5)")
        return a[: m.value].copy(), bb[: m.value].copy()

    def mg_radii(self, rstart: float, dr: float, nstep: int, imgsel: int) -> np.ndarray:
        """The radius of every W evaluation of MGPROP, in its order
        (source_code/mgprop.f:24-58)."""
        a, _ = self.mg_scheme(imgsel)
        m = len(a)
        out = []
        for istep in range(1, nstep + 1):
            r = rstart + (istep - 1) * dr
            for k in range(m):
                if k != 0 or istep == 1 or a[m - 1] != 0.0:
                    out.append(r)
                r = r + a[k] * dr
        return np.array(out)

    def prop_mg(self, b: BlockHandle, g: GridHandle, imgsel: int, nstep: int, dr: float, ered, y0: np.ndarray,
                pconst=None):
        """MGPROP over a grid of mg_radii for ninst = len(ered) instances;
        returns (Y(RSTOP), nodes, status)."""
        ered = np.ascontiguousarray(np.atleast_1d(ered), dtype=np.float64)
        m = len(ered)
        y = _colmajor(np.asarray(y0, dtype=np.float64).reshape(m, b.n, b.n))
        pc = None if pconst is None else np.ascontiguousarray(pconst, dtype=np.float64)
        nodes = np.empty(m, dtype=np.int32)
        st = np.empty(m, dtype=np.int32)
        self._check(self.lib.mcp_prop_mg(self.ctx, b.id, g.id, int(imgsel), int(nstep), float(dr), m, dptr(ered),
                                         dptr(pc), dptr(y), iptr(nodes), iptr(st)), "mcp_prop_mg")
        return _from_colmajor(y), nodes, st

    def set_md_kernel(self, mode: int):
        """MDPROP kernel for N <= 64: 0 automatic, 1 LDS-resident, 2 register-resident (one wave per instance)."""
        self._check(self.lib.mcp_set_md_kernel(self.ctx, int(mode)), "mcp_set_md_kernel")

    def set_lds_threads(self, nt: int):
        """Threads per instance of the LDS-resident MDPROP: 64, 256 or 0 (256)."""
        self._check(self.lib.mcp_set_lds_threads(self.ctx, int(nt)), "mcp_set_lds_threads")

    def set_inverse_kernel(self, mode: int):
        """Batched large-N inversions: 1 (default) 32-pivot passes for N <= 448, 0 16-pivot passes."""
        self._check(self.lib.mcp_set_inverse_kernel(self.ctx, int(mode)), "mcp_set_inverse_kernel")

    def set_vivs_options(self, iv: bool = True, ivp: bool = False, ivpp: bool = False, ishift: bool = False,
                         idiag: bool = False, ipert: bool = True, isym: bool = True, ialpha: int = 6,
                         ialfp: bool = False, alpha1: float = 1.0, alpha2: float = 1.5, xsqmax: float = 1.0e4):
        """The /LDVVCM/ controls of VVPROP (source_code/vvprop.f:57-82) for
        this context's later prop_vivs calls; the defaults are the
        reference's (source_code/mol.driver.f:277-349)."""
        self._check(self.lib.mcp_set_vivs_options(self.ctx, int(iv), int(ivp), int(ivpp), int(ishift), int(idiag),
                                                  int(ipert), int(isym), int(ialpha), int(ialfp), float(alpha1),
                                                  float(alpha2), float(xsqmax)), "mcp_set_vivs_options")

    def set_precision(self, newton_steps: int = 0, r_fp64: float = 0.0):
        """Mixed precision on the large-N tree's inversions: FP32 inverse +
        newton_steps FP64 Newton-Schulz steps, FP64 below r_fp64; 0 = FP64."""
        self._check(self.lib.mcp_set_precision(self.ctx, int(newton_steps), float(r_fp64)), "mcp_set_precision")

    def precision_stats(self) -> dict:
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.mcp_precision_stats(self.ctx, ctypes.byref(a), ctypes.byref(b)), "mcp_precision_stats")
        return {"mixed_matrices": a.value, "fp64_again": b.value}

    def set_team_launch(self, cooperative: bool):
        """Team kernels through the cooperative launch API (default) or as
        ordinary launches of the same grids (for tracing tools)."""
        self._check(self.lib.mcp_set_team_launch(self.ctx, 1 if cooperative else 0), "mcp_set_team_launch")

    def last_timing(self) -> dict:
        v = [ctypes.c_double() for _ in range(3)]
        self._check(self.lib.mcp_last_timing(self.ctx, *[ctypes.byref(x) for x in v]), "mcp_last_timing")
        return dict(ms_sector=v[0].value, ms_prop=v[1].value, ms_total=v[2].value)

    def last_segment_timing(self, k: int) -> dict:
        """Device times of segment k of the last smatrix_segments call."""
        v = [ctypes.c_double() for _ in range(2)]
        self._check(self.lib.mcp_last_segment_timing(self.ctx, int(k), *[ctypes.byref(x) for x in v]),
                    "mcp_last_segment_timing")
        return dict(ms_sector=v[0].value, ms_prop=v[1].value)


def _unpack_square(flat: np.ndarray, nopen: np.ndarray):
    out = []
    for m, nop in enumerate(nopen):
        a = flat[m, : nop * nop].reshape(nop, nop).T  # column-major nop x nop
        out.append(np.ascontiguousarray(a))
    return out
