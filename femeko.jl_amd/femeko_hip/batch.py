"""Batched solves over the C ABI (fk_batch_*): many meshes, several applied
fields each, solved together on one GPU."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import FemekoHIPError, dptr, iptr, uptr
from .device import Context, context


@dataclass
class BatchReport:
    iterations: np.ndarray   # (nmesh, nrhs)
    outer: np.ndarray
    relres: np.ndarray


class Batch:
    """nmesh meshes x nrhs Dirichlet right-hand sides (fk_batch)."""

    def __init__(self, meshes, nrhs: int = 3, ctx: Context | None = None, index_base: int = 1):
        self.ctx = ctx or context(0)
        self._L = self.ctx._L
        self.nmesh = len(meshes)
        self.nrhs = int(nrhs)
        self.nv = np.array([m.p.shape[1] for m in meshes], dtype=np.int64)
        self.nt = np.array([m.t.shape[1] for m in meshes], dtype=np.int64)
        self._p = [np.ascontiguousarray(np.asarray(m.p, dtype=np.float64).T) for m in meshes]
        self._t = [np.ascontiguousarray(np.asarray(m.t, dtype=np.int32).T) for m in meshes]
        P = (_lib._dp * self.nmesh)(*[dptr(a) for a in self._p])
        T = (_lib._ip * self.nmesh)(*[iptr(a) for a in self._t])
        h = ctypes.c_void_p()
        s = self._L.fk_batch_create(self.ctx.handle, self.nmesh, self.nv.ctypes.data_as(_lib._i64p),
                                    self.nt.ctypes.data_as(_lib._i64p), P, T, int(index_base), self.nrhs,
                                    ctypes.byref(h))
        self.ctx.check(s, "fk_batch_create")
        self.handle = h
        self._p = self._t = None
        nodes = ctypes.c_int64(0)
        nnz = ctypes.c_int64(0)
        nb = ctypes.c_int32(0)
        nc = ctypes.c_int32(0)
        self._L.fk_batch_info(h, ctypes.byref(nodes), ctypes.byref(nnz), ctypes.byref(nb), ctypes.byref(nc))
        self.total_nodes, self.total_nnz = int(nodes.value), int(nnz.value)
        self.nblocks, self.ncolours = int(nb.value), int(nc.value)

    def set_coefficient(self, mus) -> None:
        arrs = [np.ascontiguousarray(mu, dtype=np.float64) for mu in mus]
        for a, n in zip(arrs, self.nt):
            if a.size != n:
                raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, "coefficient size")
        F = (_lib._dp * self.nmesh)(*[dptr(a) for a in arrs])
        self.ctx.check(self._L.fk_batch_set_coefficient(self.handle, F), "fk_batch_set_coefficient")

    def assemble(self) -> None:
        self.ctx.check(self._L.fk_batch_assemble(self.handle), "fk_batch_assemble")

    def set_dirichlet(self, masks, gs) -> None:
        """masks[m]: (nv,), gs[m]: (nrhs, nv) Dirichlet values per right-hand side."""
        ms = [np.ascontiguousarray(m, dtype=np.uint8) for m in masks]
        gg = [np.ascontiguousarray(np.asarray(g, dtype=np.float64).reshape(self.nrhs, -1)) for g in gs]
        for a, g, n in zip(ms, gg, self.nv):
            if a.size != n or g.shape[1] != n:
                raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, "Dirichlet array size")
        M = (_lib._up * self.nmesh)(*[uptr(a) for a in ms])
        G = (_lib._dp * self.nmesh)(*[dptr(a) for a in gg])
        self.ctx.check(self._L.fk_batch_set_dirichlet(self.handle, M, G), "fk_batch_set_dirichlet")

    def set_precision(self, precision: str) -> None:
        """'f32' (FP32 PCG with FP64 reliable updates, default) or 'f32-restart'."""
        code = {"f32": _lib.FK_PREC_F32_REFINED, "f32-restart": _lib.FK_PREC_F32_RESTART}[precision]
        self.ctx.check(self._L.fk_batch_set_precision(self.handle, code), "fk_batch_set_precision")
        self.precision = precision

    def set_format(self, fmt: str) -> None:
        """SpMM copy: 'auto' (default: 'stencil' when every mesh's pattern is a
        stencil of at most 16 offsets, else 'csr'), 'stencil' or 'csr'. Set it
        before the first solve or autotune."""
        code = {"csr": 0, "stencil": 3, "auto": 4}[fmt]
        self.ctx.check(self._L.fk_batch_set_format(self.handle, code), "fk_batch_set_format")

    def solver_format(self) -> str | None:
        """'stencil' or 'csr' once a solve or autotune chose the copy, else None."""
        f = ctypes.c_int32(0)
        self.ctx.check(self._L.fk_batch_solver_format(self.handle, ctypes.byref(f)), "fk_batch_solver_format")
        return {-1: None, 0: "csr", 3: "stencil"}[int(f.value)]

    def solve(self, rtol: float = 1e-10, inner_rtol: float | None = None, max_outer: int = 50,
              max_iter: int = 1_000_000) -> BatchReport:
        """inner_rtol: replacement trigger (default 0.95) with 'f32', inner
        tolerance (default 1e-5) with 'f32-restart'."""
        if inner_rtol is None:
            inner_rtol = 1e-5 if getattr(self, "precision", "f32") == "f32-restart" else 0.95
        S = self.nmesh * self.nrhs
        it = np.zeros(S, np.int64)
        outer = np.zeros(S, np.int32)
        rel = np.zeros(S)
        self.ctx.check(self._L.fk_batch_solve(self.handle, float(rtol), float(inner_rtol), int(max_outer),
                                              int(max_iter), it.ctypes.data_as(_lib._i64p), iptr(outer), dptr(rel)),
                       "fk_batch_solve")
        sh = (self.nmesh, self.nrhs)
        return BatchReport(it.reshape(sh), outer.reshape(sh), rel.reshape(sh))

    def potential(self, m: int) -> np.ndarray:
        """(nrhs, nv[m]) potentials of mesh m."""
        out = np.empty(self.nrhs * int(self.nv[m]))
        self.ctx.check(self._L.fk_batch_potential(self.handle, int(m), dptr(out)), "fk_batch_potential")
        return out.reshape(self.nrhs, -1)

    def autotune(self):
        ln = ctypes.c_int32(0)
        bl = ctypes.c_int32(0)
        us = ctypes.c_double(0.0)
        self.ctx.check(self._L.fk_batch_autotune(self.handle, ctypes.byref(ln), ctypes.byref(bl), ctypes.byref(us)),
                       "fk_batch_autotune")
        return int(ln.value), int(bl.value), float(us.value)

    def set_launch(self, lanes: int, block: int) -> None:
        self.ctx.check(self._L.fk_batch_set_launch(self.handle, int(lanes), int(block)), "fk_batch_set_launch")

    def launch(self) -> dict:
        """The launch shape in use: CSR lanes per row, the update kernels'
        block, the stencil SpMM's block and rows per thread."""
        v = [ctypes.c_int32(0) for _ in range(4)]
        self.ctx.check(self._L.fk_batch_launch(self.handle, *[ctypes.byref(x) for x in v]), "fk_batch_launch")
        return {"lanes": v[0].value, "block": v[1].value, "spmm_block": v[2].value, "spmm_rows_per_thread": v[3].value}

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._L.fk_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
