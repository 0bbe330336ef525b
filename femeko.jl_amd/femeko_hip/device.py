"""Handles over the C ABI: a GPU context and a device-resident mesh."""
from __future__ import annotations

import atexit
import ctypes
import weakref
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import FemekoHIPError, dptr, fptr, iptr, uptr

PRECISIONS = {"f64": _lib.FK_PREC_F64, "f32": _lib.FK_PREC_F32_REFINED,
              "f32-restart": _lib.FK_PREC_F32_RESTART,
              "fp64": _lib.FK_PREC_F64, "f32-refined": _lib.FK_PREC_F32_REFINED}
METHODS = {"coloured": _lib.FK_ASM_COLOURED, "colored": _lib.FK_ASM_COLOURED, "atomic": _lib.FK_ASM_ATOMIC,
           "gather": _lib.FK_ASM_GATHER}


class Context:
    """A GPU and the stream every call on it runs on (fk_context)."""

    def __init__(self, device: int = 0):
        self._L = _lib.load()
        h = ctypes.c_void_p()
        s = self._L.fk_context_create(int(device), ctypes.byref(h))
        if s:
            raise FemekoHIPError(s, f"fk_context_create(device={device}): "
                                    f"{self._L.fk_status_string(s).decode()}")
        self.handle = h
        self.device = device

    def check(self, status: int, where: str) -> None:
        if status:
            msg = self._L.fk_context_last_error(self.handle).decode() or self._L.fk_status_string(status).decode()
            raise FemekoHIPError(status, f"{where}: {msg}")

    def synchronize(self) -> None:
        self.check(self._L.fk_context_synchronize(self.handle), "fk_context_synchronize")

    def set_timing(self, enable: bool) -> None:
        self.check(self._L.fk_set_timing(self.handle, int(bool(enable))), "fk_set_timing")

    def spmv_timing(self):
        n = ctypes.c_int64(0)
        ms = ctypes.c_double(0.0)
        self.check(self._L.fk_spmv_timing(self.handle, ctypes.byref(n), ctypes.byref(ms)), "fk_spmv_timing")
        return int(n.value), float(ms.value)

    def copy_probe(self, nbytes: int = 1 << 30, reps: int = 5) -> float:
        """Average milliseconds of one device-to-device copy of nbytes
        (read nbytes + write nbytes) with the library's streaming kernel."""
        ms = ctypes.c_double(0.0)
        self.check(self._L.fk_copy_probe(self.handle, int(nbytes), int(reps), ctypes.byref(ms)), "fk_copy_probe")
        return float(ms.value)

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._L.fk_context_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


_default = {}
_live_meshes = weakref.WeakSet()


@atexit.register
def _release_all():
    # free device memory while the HIP runtime is still up: meshes, then contexts
    for dm in list(_live_meshes):
        dm.close()
    for ctx in list(_default.values()):
        ctx.close()


def context(device: int = 0) -> Context:
    """Process-wide context of a device (created on first use)."""
    if device not in _default:
        _default[device] = Context(device)
    return _default[device]


@dataclass
class SolveReport:
    iterations: int
    outer: int
    relres: float


def _f64(a, n=None, name="array"):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, f"{name} has {a.size} entries, expected {n}")
    return a


def _spline_arrays(spl, name):
    """(knots, coefficients) of a cubic spline in FITPACK form: n knots and
    n - 4 coefficients (extra trailing coefficients, as scipy's splrep
    returns, are dropped)."""
    t = np.ascontiguousarray(getattr(spl, "t", spl[0] if isinstance(spl, tuple) else None), dtype=np.float64)
    c = np.ascontiguousarray(getattr(spl, "c", spl[1] if isinstance(spl, tuple) else None), dtype=np.float64)
    if t.ndim != 1 or c.ndim != 1 or t.size < 8 or c.size < t.size - 4:
        raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, f"{name}: need n >= 8 knots and n - 4 coefficients")
    return t, np.ascontiguousarray(c[:t.size - 4])


class DeviceMesh:
    """A MESH resident on the GPU: coordinates, connectivity in colour order,
    CSR pattern and element -> slot map, coefficients, values, Dirichlet data,
    potential (fk_mesh)."""

    def __init__(self, mesh, ctx: Context | None = None, check_volumes: bool = True, index_base: int = 1,
                 reorder: str = "auto"):
        """reorder: 'auto' (default: the library renumbers the nodes internally
        by reverse Cuthill-McKee unless the numbering is a stencil numbering or
        already compact), 'none' or 'rcm'. Results are always in the caller's
        numbering."""
        self.ctx = ctx or context(0)
        self._L = self.ctx._L
        self.nv = int(mesh.p.shape[1])
        self.nt = int(mesh.t.shape[1])
        p = np.ascontiguousarray(np.asarray(mesh.p, dtype=np.float64).T)   # (nv,3) == 3 x nv column-major
        t = np.ascontiguousarray(np.asarray(mesh.t, dtype=np.int32).T)     # (nt,4) == 4 x nt column-major
        VE = _f64(mesh.VE, self.nt, "VE") if (check_volumes and getattr(mesh, "VE", None) is not None) else None
        h = ctypes.c_void_p()
        code = {"auto": _lib.FK_REORDER_AUTO, "none": _lib.FK_REORDER_NONE, "rcm": _lib.FK_REORDER_RCM}[reorder]
        s = self._L.fk_mesh_create_ex(self.ctx.handle, dptr(p), self.nv, iptr(t), self.nt, int(index_base),
                                      dptr(VE), code, ctypes.byref(h))
        self.ctx.check(s, "fk_mesh_create")
        self.handle = h
        _live_meshes.add(self)
        nnz = ctypes.c_int64(0)
        nc = ctypes.c_int32(0)
        unused = ctypes.c_int64(0)
        self.ctx.check(self._L.fk_mesh_info(h, ctypes.byref(nnz), ctypes.byref(nc), ctypes.byref(unused)),
                       "fk_mesh_info")
        self.nnz = int(nnz.value)
        self.ncolours = int(nc.value)
        self.unused_nodes = int(unused.value)

    def numbering(self):
        """(renumbered, mean neighbour distance in the caller's numbering, in
        the internal one) of fk_mesh_numbering."""
        r = ctypes.c_int32(0)
        a, b = ctypes.c_double(0.0), ctypes.c_double(0.0)
        self.ctx.check(self._L.fk_mesh_numbering(self.handle, ctypes.byref(r), ctypes.byref(a), ctypes.byref(b)),
                       "fk_mesh_numbering")
        return bool(r.value), float(a.value), float(b.value)

    # -- pattern / assembly --------------------------------------------------
    def pattern(self, base: int = 1):
        rowptr = np.empty(self.nv + 1, np.int32)
        colidx = np.empty(self.nnz, np.int32)
        self.ctx.check(self._L.fk_pattern(self.handle, int(base), iptr(rowptr), iptr(colidx)), "fk_pattern")
        return rowptr, colidx

    def set_coefficient(self, f) -> None:
        f = _f64(f, self.nt, "f")
        self.ctx.check(self._L.fk_set_coefficient(self.handle, dptr(f)), "fk_set_coefficient")

    def assemble(self, method: str = "gather") -> None:
        self.ctx.check(self._L.fk_assemble(self.handle, METHODS[method]), "fk_assemble")

    def values_generation(self) -> int:
        """Counter of changes of the resident matrix values (fk_values_generation)."""
        g = ctypes.c_int64(0)
        self.ctx.check(self._L.fk_values_generation(self.handle, ctypes.byref(g)), "fk_values_generation")
        return int(g.value)

    def values(self) -> np.ndarray:
        v = np.empty(self.nnz)
        self.ctx.check(self._L.fk_values(self.handle, dptr(v)), "fk_values")
        return v

    def values_f32(self) -> np.ndarray:
        v = np.empty(self.nnz, np.float32)
        self.ctx.check(self._L.fk_values_f32(self.handle, fptr(v)), "fk_values_f32")
        return v

    def local_stiffness(self, f=None) -> np.ndarray:
        Ak = np.empty((self.nt, 16))
        fa = None if f is None else _f64(f, self.nt, "f")
        self.ctx.check(self._L.fk_local_stiffness(self.handle, dptr(fa), dptr(Ak)), "fk_local_stiffness")
        return Ak.T

    def tangential_values(self, H, dmu) -> np.ndarray:
        Hc = _f64(np.asarray(H, dtype=np.float64).T, 3 * self.nt, "H")
        d = _f64(dmu, self.nt, "dmu")
        v = np.empty(self.nnz)
        self.ctx.check(self._L.fk_tangential_stiffness(self.handle, dptr(Hc), dptr(d), dptr(v)),
                       "fk_tangential_stiffness")
        return v

    def lagrange(self) -> np.ndarray:
        C = np.empty(self.nv)
        self.ctx.check(self._L.fk_lagrange(self.handle, dptr(C)), "fk_lagrange")
        return C

    def volumes(self) -> np.ndarray:
        v = np.empty(self.nt)
        self.ctx.check(self._L.fk_mesh_volumes(self.handle, dptr(v)), "fk_mesh_volumes")
        return v

    # -- Dirichlet data and solve ----------------------------------------------
    def set_dirichlet(self, mask, g) -> None:
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        if m.size != self.nv:
            raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, "mask size")
        g = _f64(g, self.nv, "g")
        self.ctx.check(self._L.fk_set_dirichlet(self.handle, uptr(m), dptr(g)), "fk_set_dirichlet")

    def solve(self, precision: str = "f64", rtol: float = 1e-10, max_iter: int = 1_000_000,
              inner_rtol: float | None = None, max_outer: int = 50, warm_start: bool = False,
              raise_on_error: bool = True) -> SolveReport:
        """precision 'f64'; 'f32' (FP32 PCG with FP64 reliable updates, inner_rtol =
        residual drop that triggers an update, default 0.95); 'f32-restart'
        (restarted FP32 PCG to inner_rtol, default 1e-5, inside FP64 refinement)."""
        if inner_rtol is None:
            inner_rtol = 1e-5 if precision == "f32-restart" else 0.95
        it = ctypes.c_int64(0)
        outer = ctypes.c_int32(0)
        rel = ctypes.c_double(0.0)
        s = self._L.fk_solve(self.handle, PRECISIONS[precision], float(rtol), int(max_iter), float(inner_rtol),
                             int(max_outer), int(bool(warm_start)), ctypes.byref(it), ctypes.byref(outer),
                             ctypes.byref(rel))
        rep = SolveReport(int(it.value), int(outer.value), float(rel.value))
        if raise_on_error:
            self.ctx.check(s, "fk_solve")
        rep.status = s
        return rep

    def solve_rhs(self, b, rtol: float = 1e-10, max_iter: int = 1_000_000):
        b = _f64(b, self.nv, "b")
        x = np.empty(self.nv)
        it = ctypes.c_int64(0)
        rel = ctypes.c_double(0.0)
        self.ctx.check(self._L.fk_solve_rhs(self.handle, dptr(b), float(rtol), int(max_iter), dptr(x),
                                            ctypes.byref(it), ctypes.byref(rel)), "fk_solve_rhs")
        return x, int(it.value), float(rel.value)

    def potential(self) -> np.ndarray:
        u = np.empty(self.nv)
        self.ctx.check(self._L.fk_potential(self.handle, dptr(u)), "fk_potential")
        return u

    def set_potential(self, u) -> None:
        u = _f64(u, self.nv, "u")
        self.ctx.check(self._L.fk_set_potential(self.handle, dptr(u)), "fk_set_potential")

    def field(self, want_H: bool = True, want_Hn: bool = True, want_energy: bool = False):
        H = np.empty((self.nt, 3)) if want_H else None
        Hn = np.empty(self.nt) if want_Hn else None
        E = ctypes.c_double(0.0)
        self.ctx.check(self._L.fk_field(self.handle, dptr(H), dptr(Hn),
                                        ctypes.byref(E) if want_energy else None), "fk_field")
        return (H.T if H is not None else None), Hn, (float(E.value) if want_energy else None)

    def boundary_integral(self, surfaceT, normal, AE, F, shell_ids, index_base: int = 1) -> np.ndarray:
        sT = np.ascontiguousarray(np.asarray(surfaceT, dtype=np.int32).T)   # (ns,4) == 4 x ns column-major
        nr = np.ascontiguousarray(np.asarray(normal, dtype=np.float64).T)
        ae = _f64(AE, sT.shape[0], "AE")
        ids = np.ascontiguousarray(np.atleast_1d(shell_ids), dtype=np.int32)
        rhs = np.empty(self.nv)
        self.ctx.check(self._L.fk_boundary_integral(self.handle, sT.shape[0], iptr(sT), dptr(nr), dptr(ae),
                                                    dptr(_f64(F, 3, "F")), iptr(ids), ids.size, int(index_base),
                                                    dptr(rhs)), "fk_boundary_integral")
        return rhs

    def solve_neumann(self, b, rtol: float = 1e-10, max_iter: int = 1_000_000):
        """Saddle solve [A Lag; Lag' 0][u; m] = [b; 0]: returns (u, m, iterations, relres)."""
        b = _f64(b, self.nv, "b")
        u = np.empty(self.nv)
        mult = ctypes.c_double(0.0)
        it = ctypes.c_int64(0)
        rel = ctypes.c_double(0.0)
        self.ctx.check(self._L.fk_solve_neumann(self.handle, dptr(b), float(rtol), int(max_iter), dptr(u),
                                                ctypes.byref(mult), ctypes.byref(it), ctypes.byref(rel)),
                       "fk_solve_neumann")
        return u, float(mult.value), int(it.value), float(rel.value)

    def get_coefficient(self) -> np.ndarray:
        f = np.empty(self.nt)
        self.ctx.check(self._L.fk_get_coefficient(self.handle, dptr(f)), "fk_get_coefficient")
        return f

    def set_inner_precision(self, precision: str) -> None:
        """Precision of the solves inside the mu(|H|) loops, the Neumann
        solves (solve_neumann, demag_field, newton_saddle): 'f64' (default)
        or 'f32' (FP32-refined, fk_set_inner_precision)."""
        code = {"f64": _lib.FK_PREC_F64, "f32": _lib.FK_PREC_F32_REFINED}.get(precision)
        if code is None:
            raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, f"unknown inner precision {precision!r}")
        self.ctx.check(self._L.fk_set_inner_precision(self.handle, code), "fk_set_inner_precision")

    def nonlinear_froehlich(self, region, Ms: float, a: float, rtol: float = 1e-10, tol_mu: float = 1e-6,
                            max_outer: int = 20, switch_tol: float = 0.25, raise_on_error: bool = True):
        """Picard/Newton mu(|H|) loop on the GPU. Returns dict(outer, cg_iterations, change, status)."""
        reg = np.ascontiguousarray(region, dtype=np.uint8)
        if reg.size != self.nt:
            raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, "region size")
        outer = ctypes.c_int32(0)
        cg = ctypes.c_int64(0)
        ch = ctypes.c_double(0.0)
        s = self._L.fk_nonlinear_froehlich(self.handle, uptr(reg), float(Ms), float(a), float(rtol), float(tol_mu),
                                           int(max_outer), float(switch_tol), ctypes.byref(outer),
                                           ctypes.byref(cg), ctypes.byref(ch))
        if raise_on_error:
            self.ctx.check(s, "fk_nonlinear_froehlich")
        return {"outer": int(outer.value), "cg_iterations": int(cg.value), "change": float(ch.value), "status": s}

    def nonlinear_spline(self, region, spl_mu, spl_dmu, rtol: float = 1e-10, tol_H: float = 1e-4,
                         max_outer: int = 100, switch_tol: float = float("inf"), raise_on_error: bool = True):
        """The mu(|H|) loop with the tabulated law of NewtonRaphson.jl:204-207:
        spl_mu, spl_dmu are cubic splines in FITPACK form (objects with
        knots t and coefficients c, e.g. femeko_hip.material.Spline1D).
        Stops when mu0 max | |H| - |H|_prev | over the region < tol_H.
        Returns dict(outer, cg_iterations, change, status)."""
        reg = np.ascontiguousarray(region, dtype=np.uint8)
        if reg.size != self.nt:
            raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, "region size")
        tm, cm = _spline_arrays(spl_mu, "spl_mu")
        td, cd = _spline_arrays(spl_dmu, "spl_dmu")
        outer = ctypes.c_int32(0)
        cg = ctypes.c_int64(0)
        ch = ctypes.c_double(0.0)
        s = self._L.fk_nonlinear_spline(self.handle, uptr(reg), dptr(tm), tm.size, dptr(cm), dptr(td), td.size,
                                        dptr(cd), float(rtol), float(tol_H), int(max_outer), float(switch_tol),
                                        ctypes.byref(outer), ctypes.byref(cg), ctypes.byref(ch))
        if raise_on_error:
            self.ctx.check(s, "fk_nonlinear_spline")
        return {"outer": int(outer.value), "cg_iterations": int(cg.value), "change": float(ch.value), "status": s}

    def newton_saddle(self, region, rhs, law, rtol: float = 1e-10, tol_H: float = 1e-4, max_outer: int = 100,
                      min_step: float = 1e-4, raise_on_error: bool = True):
        """The reference's Neumann + Lagrange Newton loop with lineSearch
        (Magnetostatics/NewtonRaphson.jl:109-212) on the GPU. rhs: the
        driver's RHS = -BoundaryIntegral(mesh, Bext, shell_id); law:
        ("froehlich", Ms, a) or ("spline", spl_mu, spl_dmu) with FITPACK-form
        splines. Returns dict(outer, cg_iterations, change, residue, status);
        the potential and mu stay resident."""
        reg = np.ascontiguousarray(region, dtype=np.uint8)
        if reg.size != self.nt:
            raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, "region size")
        b = _f64(rhs, self.nv, "rhs")
        outer = ctypes.c_int32(0)
        cg = ctypes.c_int64(0)
        ch = ctypes.c_double(0.0)
        res = ctypes.c_double(0.0)
        if law[0] == "froehlich":
            Ms, a = float(law[1]), float(law[2])
            tm = cm = td = cd = None
            nm = nd = 0
        else:
            Ms, a = 0.0, 1.0
            tm, cm = _spline_arrays(law[1], "spl_mu")
            td, cd = _spline_arrays(law[2], "spl_dmu")
            nm, nd = tm.size, td.size
        s = self._L.fk_newton_saddle(self.handle, uptr(reg), dptr(b), Ms, a, dptr(tm), nm, dptr(cm), dptr(td), nd,
                                     dptr(cd), float(rtol), float(tol_H), int(max_outer), float(min_step),
                                     ctypes.byref(outer), ctypes.byref(cg), ctypes.byref(ch), ctypes.byref(res))
        if raise_on_error:
            self.ctx.check(s, "fk_newton_saddle")
        return {"outer": int(outer.value), "cg_iterations": int(cg.value), "change": float(ch.value),
                "residue": float(res.value), "status": s}

    def demag_field(self, inside, M, volumes, rtol: float = 1.4901161193847656e-08, max_iter: int = 100000,
                    warm: bool = False):
        """magnetostaticField (Micromagnetics/LL.jl:162-209) with the resident
        matrix A = stiffnessMatrix(mesh): M is 3 x nv, volumes nv; returns
        (Hd_nodes 3 x nv, PCG iterations, relative residual). rtol defaults to
        the reference cg's sqrt(eps); warm starts the PCG from the previous
        call's potential."""
        ins = np.ascontiguousarray(inside, dtype=np.uint8)
        if ins.size != self.nt:
            raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, "inside size")
        Mf = np.asfortranarray(np.asarray(M, dtype=np.float64))
        if Mf.shape != (3, self.nv):
            raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, "M must be 3 x nv")
        vol = _f64(volumes, self.nv, "volumes")
        Hd = np.empty((3, self.nv), order="F")
        it = ctypes.c_int64(0)
        rel = ctypes.c_double(0.0)
        self.ctx.check(self._L.fk_demag_field(self.handle, uptr(ins), dptr(Mf.reshape(-1, order="F")), dptr(vol),
                                              float(rtol), int(max_iter), 1 if warm else 0,
                                              dptr(Hd.reshape(-1, order="F")),
                                              ctypes.byref(it), ctypes.byref(rel)), "fk_demag_field")
        return Hd, int(it.value), float(rel.value)

    def adjoint_field_target(self, target, Hstar, rtol: float = 1e-10):
        """(J, dJ/dmu, adjoint PCG iterations) for J = sum_target VE |H - H*|^2 at the resident potential."""
        tg = np.ascontiguousarray(target, dtype=np.uint8)
        if tg.size != self.nt:
            raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, "target size")
        hs = _f64(Hstar, 3, "Hstar")
        J = ctypes.c_double(0.0)
        it = ctypes.c_int64(0)
        grad = np.empty(self.nt)
        self.ctx.check(self._L.fk_adjoint_field_target(self.handle, uptr(tg), dptr(hs), float(rtol), ctypes.byref(J),
                                                       dptr(grad), ctypes.byref(it)), "fk_adjoint_field_target")
        return float(J.value), grad, int(it.value)

    def line_search(self, u, du, rhs, min_step: float = 1e-4) -> float:
        """lineSearch(u, du, A, RHS, minStep): updates u in place, returns the residue."""
        if not (isinstance(u, np.ndarray) and u.dtype == np.float64 and u.flags.c_contiguous and u.size == self.nv):
            raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, "u must be a contiguous float64 array of nv entries")
        du = _f64(du, self.nv, "du")
        rhs = _f64(rhs, self.nv, "rhs")
        res = ctypes.c_double(0.0)
        self.ctx.check(self._L.fk_line_search(self.handle, dptr(u), dptr(du), dptr(rhs), float(min_step),
                                              ctypes.byref(res)), "fk_line_search")
        return float(res.value)

    def set_format(self, fmt: str) -> None:
        """Solver SpMV format: 'auto' (default: 'stencil' when the pattern is a
        stencil of at most 16 offsets, else 'sell64'), 'stencil' (symmetric
        diagonal storage), 'sell64' (shared slice offsets where the rows allow),
        'sell64-rowidx' (per-row indices only) or 'csr'."""
        code = {"csr": 0, "sell64": 1, "sell": 1, "sell64-rowidx": 2, "stencil": 3, "auto": 4}[fmt]
        self.ctx.check(self._L.fk_set_spmv_format(self.handle, code), "fk_set_spmv_format")

    def spmv_burst(self, precision: str = "f32", reps: int = 50) -> float:
        """Microseconds per launch of the solver's SpMV over reps back-to-back
        launches (overwrites the solver's direction vector)."""
        code = {"f64": _lib.FK_PREC_F64, "f32": _lib.FK_PREC_F32_REFINED}[precision]
        us = ctypes.c_double(0.0)
        self.ctx.check(self._L.fk_spmv_burst(self.handle, code, int(reps), ctypes.byref(us)), "fk_spmv_burst")
        return float(us.value)

    def solver_format(self):
        """(format name, stencil offsets or ()) the solver resolved for this mesh."""
        f = ctypes.c_int32(0)
        k = ctypes.c_int32(0)
        offs = (ctypes.c_int32 * 16)()
        self.ctx.check(self._L.fk_solver_format(self.handle, ctypes.byref(f), ctypes.byref(k), offs),
                       "fk_solver_format")
        name = {0: "csr", 1: "sell64", 3: "stencil"}[int(f.value)]
        return name, tuple(int(offs[i]) for i in range(int(k.value)))

    def sell_info(self):
        """(slices, shared-offset slices, stored entries) of the SELL copy."""
        a, b, c = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
        self.ctx.check(self._L.fk_sell_info(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                       "fk_sell_info")
        return int(a.value), int(b.value), int(c.value)

    def sell_layout(self):
        """(index bytes, batch groups, FP32 array bytes) of the SELL copy: 2
        (int16 offsets from the row) or 4 (int32 columns); 4 when the FP32
        product loads 4 groups of a row at once, 0 when it walks them one by
        one; the bytes one FP32 product streams besides x and y."""
        b, g, a = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int64(0)
        self.ctx.check(self._L.fk_sell_layout(self.handle, ctypes.byref(b), ctypes.byref(g), ctypes.byref(a)),
                       "fk_sell_layout")
        return int(b.value), int(g.value), int(a.value)

    def set_sell_shape(self, block: int, unroll: int = 1) -> None:
        """SELL SpMV launch shape: threads per block and slices in flight per wave."""
        self.ctx.check(self._L.fk_set_sell_shape(self.handle, int(block), int(unroll)), "fk_set_sell_shape")

    def set_lanes(self, lanes: int) -> None:
        self.ctx.check(self._L.fk_set_lanes_per_row(self.handle, int(lanes)), "fk_set_lanes_per_row")

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._L.fk_mesh_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
