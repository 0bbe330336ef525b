"""Python face of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

FP64 restatement of Femeko.jl's P1 magnetostatics path (liboracle.so, built
from femeko_oracle.c by oracle/Makefile) plus numpy pieces for the parts that
have no C twin (energy, Froehlich law, Newton loop, adjoint). It checks the HIP
library in tests/ and is timed as the CPU baseline in bench.py; the shipped
library never imports it.

Every function names the reference lines it restates. Pieces without a
reference counterpart (adjoint, Froehlich workload law) say so.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

MU0 = np.pi * 4e-7  # Magnetostatics/magneticField.jl:38

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_up = ctypes.POINTER(ctypes.c_uint8)
_i64 = ctypes.c_int64


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_abcd.argtypes = [_dp, _ip, ctypes.c_int32, _dp]
        L.or_element_volume.argtypes = [_dp, _ip]
        L.or_element_volume.restype = ctypes.c_double
        L.or_element_gradients.argtypes = [_dp, _ip, _i64, _dp, ctypes.c_int]
        L.or_local_stiffness.argtypes = [_dp, _ip, _i64, _dp, _dp, _dp, ctypes.c_int]
        L.or_tangent_local.argtypes = [_dp, _ip, _i64, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.or_sparse_csc.argtypes = [_i64, _i64, _ip, _dp, _dp, _ip, _ip, _dp]
        L.or_sparse_csc.restype = _i64
        L.or_spmv_csc.argtypes = [_i64, _ip, _ip, _dp, _dp, _dp]
        L.or_pcg_free.argtypes = [_i64, _ip, _ip, _dp, _up, _dp, ctypes.c_double, _i64, _dp,
                                  ctypes.POINTER(_i64), _dp, ctypes.c_int, ctypes.c_int]
        L.or_pcg_dirichlet.argtypes = L.or_pcg_free.argtypes
        L.or_pcg_iterations.argtypes = [_i64, _ip, _ip, _dp, _up, _dp, _i64, _dp, ctypes.c_int]
        L.or_field.argtypes = [_dp, _ip, _i64, _dp, _dp, _dp, ctypes.c_int]
        L.or_lagrange.argtypes = [_i64, _i64, _ip, _dp, _dp]
        L.or_boundary_integral.argtypes = [_i64, _i64, _ip, _dp, _dp, _dp, _ip, ctypes.c_int, _dp]
        L.or_num_threads_available.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _u(a):
    return a.ctypes.data_as(_up)


def _p(mesh):
    """p as a contiguous 3 x nv column-major buffer (x,y,z per node)."""
    return np.ascontiguousarray(mesh.p.T)


def _t(mesh):
    return np.ascontiguousarray(mesh.t.T, dtype=np.int32)


def threads() -> int:
    return int(lib().or_num_threads_available())


# --------------------------------------------------------------------------
def abcd(p: np.ndarray, nodes, nd: int):
    """src/FEM3D.jl:5-17 -> (a, b, c, d). p is (3, nv); nodes 1-based."""
    pp = np.ascontiguousarray(np.asarray(p, dtype=np.float64).T)
    nn = np.ascontiguousarray(np.asarray(nodes, dtype=np.int32))
    r = np.zeros(4)
    s = lib().or_abcd(_d(pp), _i(nn), int(nd), _d(r))
    if s:
        raise ValueError("singular element")
    return tuple(r)


def element_gradients(mesh, nthreads: int = 0) -> np.ndarray:
    """(nt, 4, 3) basis gradients of every element, via abcd."""
    G = np.empty((mesh.nt, 4, 3))
    s = lib().or_element_gradients(_d(_p(mesh)), _i(_t(mesh)), mesh.nt, _d(G), nthreads or threads())
    if s:
        raise ValueError("singular element")
    return G


def localStiffnessMatrix(mesh, f=None, nthreads: int = 0) -> np.ndarray:
    """src/FEM3D.jl:59-76 -> Ak (16, nt), column k = vec(VE f (bb'+cc'+dd'))."""
    Ak = np.empty((mesh.nt, 16))
    fp = None if f is None else _d(np.ascontiguousarray(f, dtype=np.float64))
    s = lib().or_local_stiffness(_d(_p(mesh)), _i(_t(mesh)), mesh.nt,
                                 _d(np.ascontiguousarray(mesh.VE)), fp, _d(Ak), nthreads or threads())
    if s:
        raise ValueError("singular element")
    return Ak.T


def tangential_local(mesh, H: np.ndarray, dmu: np.ndarray, nthreads: int = 0) -> np.ndarray:
    """Element matrices of tangentialStiffnessMatrix, src/FEM3D.jl:79-108. H is (3, nt)."""
    Ak = np.empty((mesh.nt, 16))
    Hc = np.ascontiguousarray(np.asarray(H, dtype=np.float64).T)
    s = lib().or_tangent_local(_d(_p(mesh)), _i(_t(mesh)), mesh.nt, _d(np.ascontiguousarray(mesh.VE)),
                               _d(Hc), _d(np.ascontiguousarray(dmu, dtype=np.float64)), _d(Ak),
                               nthreads or threads())
    if s:
        raise ValueError("singular element")
    return Ak.T


@dataclass
class CSC:
    """Julia SparseMatrixCSC arrays (1-based colptr / rowval)."""

    n: int
    colptr: np.ndarray
    rowval: np.ndarray
    nzval: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.nzval.size)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.nzval, self.rowval - 1, self.colptr - 1), shape=(self.n, self.n))


def sparse_from_local(mesh, Ak: np.ndarray, mu=None) -> CSC:
    """sparse(rowIDs, colIDs, Ak[n,k]*mu[k]) (Magnetostatics/NewtonRaphson.jl:115-127,154-161)."""
    nv, nt = mesh.nv, mesh.nt
    Akc = np.ascontiguousarray(np.asarray(Ak).T)  # (nt, 16)
    colptr = np.empty(nv + 1, np.int32)
    rowval = np.empty(16 * nt, np.int32)
    nzval = np.empty(16 * nt)
    mup = None if mu is None else _d(np.ascontiguousarray(mu, dtype=np.float64))
    nnz = lib().or_sparse_csc(nv, nt, _i(_t(mesh)), _d(Akc), mup, _i(colptr), _i(rowval), _d(nzval))
    if nnz < 0:
        raise MemoryError("oracle sparse build")
    return CSC(nv, colptr, rowval[:nnz].copy(), nzval[:nnz].copy())


def term_scale(mesh, f=None) -> np.ndarray:
    """Per CSC entry, sum over elements of |VE f| ||grad_i|| ||grad_j||: the
    magnitude of the operands of each element term VE f (grad_i . grad_j),
    which bounds its rounding error even where the term itself cancels to 0
    (e.g. face diagonals inside a flat outer face). The value-parity metric
    is |a - a_ref| <= 1e-12 * term_scale."""
    G = element_gradients(mesh)
    nrm = np.linalg.norm(G, axis=2)                     # (nt, 4)
    w = np.abs(mesh.VE * (1.0 if f is None else np.asarray(f, dtype=np.float64)))
    K = w[:, None, None] * nrm[:, :, None] * nrm[:, None, :]
    return sparse_from_local(mesh, K.reshape(mesh.nt, 16).T).nzval


def tangent_scale(mesh, H, dmu) -> np.ndarray:
    """term_scale of the tangent terms VE dmu/|H| (H.g_i)(H.g_j):
    sum over elements of VE |dmu| |H| ||grad_i|| ||grad_j||."""
    G = element_gradients(mesh)
    nrm = np.linalg.norm(G, axis=2)
    hn = np.linalg.norm(np.asarray(H, dtype=np.float64), axis=0)
    w = mesh.VE * np.abs(np.asarray(dmu, dtype=np.float64)) * hn
    K = w[:, None, None] * nrm[:, :, None] * nrm[:, None, :]
    return sparse_from_local(mesh, K.reshape(mesh.nt, 16).T).nzval


def stiffnessMatrix(mesh, f=None, nthreads: int = 0) -> CSC:
    """src/FEM3D.jl:40-56 with the zero-keeping single-call sparse convention."""
    return sparse_from_local(mesh, localStiffnessMatrix(mesh, f, nthreads))


def tangentialStiffnessMatrix(mesh, H, dmu, nthreads: int = 0) -> CSC:
    """src/FEM3D.jl:79-119 (zero-keeping convention, |H| = 0 elements contribute 0)."""
    return sparse_from_local(mesh, tangential_local(mesh, H, dmu, nthreads))


def spmv(A: CSC, x: np.ndarray) -> np.ndarray:
    y = np.empty(A.n)
    lib().or_spmv_csc(A.n, _i(A.colptr), _i(A.rowval), _d(A.nzval),
                      _d(np.ascontiguousarray(x, dtype=np.float64)), _d(y))
    return y


def pcg_free(A: CSC, mask: np.ndarray, b: np.ndarray, rtol: float, maxiter: int = 100000,
             x0=None, exact: bool = True, nthreads: int = 0):
    """Jacobi-PCG on A[free,free] x = b[free], x = 0 on fixed nodes."""
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64)
    it = _i64(0)
    rel = ctypes.c_double(0.0)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    s = lib().or_pcg_free(A.n, _i(A.colptr), _i(A.rowval), _d(A.nzval), _u(m),
                          _d(np.ascontiguousarray(b, dtype=np.float64)), rtol, maxiter, _d(x),
                          ctypes.byref(it), ctypes.byref(rel), nthreads or threads(), int(exact))
    return x, int(it.value), float(rel.value), int(s)


def pcg_dirichlet(A: CSC, mask: np.ndarray, g: np.ndarray, rtol: float, maxiter: int = 100000,
                  exact: bool = True, nthreads: int = 0):
    """Dirichlet elimination (Fluids/fluids3D.jl:142-174) + Jacobi-PCG. Returns phi, iters, relres, status."""
    phi = np.zeros(A.n)
    it = _i64(0)
    rel = ctypes.c_double(0.0)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    s = lib().or_pcg_dirichlet(A.n, _i(A.colptr), _i(A.rowval), _d(A.nzval), _u(m),
                               _d(np.ascontiguousarray(g, dtype=np.float64)), rtol, maxiter, _d(phi),
                               ctypes.byref(it), ctypes.byref(rel), nthreads or threads(), int(exact))
    return phi, int(it.value), float(rel.value), int(s)


def field(mesh, u: np.ndarray, nthreads: int = 0):
    """H recovery, magneticField.jl:111-130 (correct z): returns H (3, nt) and |H| (nt)."""
    H = np.empty((mesh.nt, 3))
    Hn = np.empty(mesh.nt)
    s = lib().or_field(_d(_p(mesh)), _i(_t(mesh)), mesh.nt, _d(np.ascontiguousarray(u, dtype=np.float64)),
                       _d(H), _d(Hn), nthreads or threads())
    if s:
        raise ValueError("singular element")
    return H.T, Hn


def energy(mesh, Hn: np.ndarray, mu: np.ndarray) -> float:
    """Linear magnetic energy 1/2 sum_e mu_e |H_e|^2 VE_e (the free-space density
    of src/magneticProperties.jl:268-273 with the element permeability)."""
    return float(np.sum(0.5 * mu * Hn * Hn * mesh.VE))


def lagrange(mesh) -> np.ndarray:
    """src/FEM3D.jl:274-281."""
    C = np.empty(mesh.nv)
    lib().or_lagrange(mesh.nv, mesh.nt, _i(_t(mesh)), _d(np.ascontiguousarray(mesh.VE)), _d(C))
    return C


def BoundaryIntegral(mesh, F, shell_id) -> np.ndarray:
    """src/FEM3D.jl:284-297."""
    RHS = np.empty(mesh.nv)
    sT = np.ascontiguousarray(mesh.surfaceT.T, dtype=np.int32)
    nrm = np.ascontiguousarray(mesh.normal.T, dtype=np.float64)
    ids = np.ascontiguousarray(np.atleast_1d(shell_id), dtype=np.int32)
    lib().or_boundary_integral(mesh.nv, mesh.ns, _i(sT), _d(nrm), _d(np.ascontiguousarray(mesh.AE)),
                               _d(np.ascontiguousarray(F, dtype=np.float64)), _i(ids), int(ids.size), _d(RHS))
    return RHS


def lineSearch(u: np.ndarray, du: np.ndarray, A: CSC, RHS: np.ndarray, minStep: float = 1e-4) -> float:
    """src/FEM3D.jl:123-152: backtracking alpha <- 0.9 alpha; mutates u."""
    alf = 1.0
    residual_old = np.linalg.norm(RHS - spmv(A, u))
    u_trial = u + du
    residue = np.linalg.norm(RHS - spmv(A, u_trial))
    while residue > residual_old and alf > minStep:
        alf *= 0.9
        u_trial = u + alf * du
        residue = np.linalg.norm(RHS - spmv(A, u_trial))
    if not alf < minStep:
        u[:] = u_trial
    return float(residue)


# --------------------------------------------------------------------------
# Froehlich permeability law of the nonlinear workload (no reference
# counterpart: Femeko evaluates tabulated splines, NewtonRaphson.jl:57-63).
def froehlich(Hn: np.ndarray, Ms: float, a: float):
    """mu = mu0 (1 + Ms/(a+|H|)), dmu/d|H| = -mu0 Ms/(a+|H|)^2."""
    d = a + Hn
    return MU0 * (1.0 + Ms / d), -MU0 * Ms / (d * d)


def _newton_loop(mesh, mask, g, magnet, law, deviation, tol, max_outer, newton, switch, rtol):
    """The loop of Magnetostatics/NewtonRaphson.jl:148-212 with Dirichlet data
    instead of the Neumann+Lagrange saddle system and the correct H_z.
    law(Hn) -> (mu, dmu) on the magnet; deviation(mu_new, mu_old, Hn_new,
    Hn_old) -> the scalar the loop stops on. Iteration 1 starts from mu = mu0
    with no tangent (the reference's att == 1 branch); later iterations add
    tangentialStiffnessMatrix (a Newton step) once the previous deviation is
    below `switch`, otherwise they are Picard steps."""
    Ak = localStiffnessMatrix(mesh)
    mu = np.full(mesh.nt, MU0)
    dmu = np.zeros(mesh.nt)
    u = np.where(mask.astype(bool), g, 0.0)
    H = np.zeros((3, mesh.nt))
    Hn = np.zeros(mesh.nt)
    hist = []
    steps = []
    cg_total = 0
    it = 0
    dev = np.inf
    for it in range(1, max_outer + 1):
        A = sparse_from_local(mesh, Ak, mu)
        use_newton = newton and dev < switch
        steps.append("N" if use_newton else "P")
        if use_newton:
            At = tangentialStiffnessMatrix(mesh, H, dmu)
            J = CSC(A.n, A.colptr, A.rowval, A.nzval + At.nzval)
        else:
            J = A
        res = -spmv(A, u)
        du, cgit, _rel, s = pcg_free(J, mask, res, rtol)
        if s:
            raise RuntimeError(f"oracle PCG status {s}")
        cg_total += cgit
        u = u + du
        Hn_old = Hn
        H, Hn = field(mesh, u)
        mu_new, dmu_new = law(Hn[magnet])
        mu_next = mu.copy()
        mu_next[magnet] = mu_new
        dmu = np.zeros(mesh.nt)
        dmu[magnet] = dmu_new
        dev = float(deviation(mu_next[magnet], mu[magnet], Hn[magnet], Hn_old[magnet])) if magnet.any() else 0.0
        hist.append(dev)
        mu = mu_next
        if not (np.all(np.isfinite(mu)) and np.all(np.isfinite(dmu))):
            raise FloatingPointError("Nans/Infs in mu")  # NewtonRaphson.jl:210-212
        if dev < tol:
            break
    H, Hn = field(mesh, u)
    return {"phi": u, "H": H, "Hn": Hn, "mu": mu, "dmu": dmu, "iters": it, "cg_iters": cg_total,
            "dmu_hist": hist, "steps": steps}


def newton_froehlich(mesh, mask, g, magnet: np.ndarray, Ms: float, a: float,
                     rtol: float = 1e-10, tol_mu: float = 1e-6, max_outer: int = 20,
                     newton: bool = True, switch: float = 0.25):
    """The loop with the Froehlich law (a workload law: no reference
    counterpart); far from the solution undamped Newton oscillates for it,
    which the reference counters with lineSearch, hence the Picard start.
    Stops when max |mu_new - mu| / mu_new over magnet elements < tol_mu.
    Returns dict(phi, H, Hn, mu, dmu, iters, cg_iters, dmu_hist, steps)."""
    return _newton_loop(mesh, mask, g, magnet, lambda h: froehlich(h, Ms, a),
                        lambda mn, mo, hn, ho: np.max(np.abs(mn - mo) / mn), tol_mu, max_outer, newton,
                        switch, rtol)


# --------------------------------------------------------------------------
# Tabulated permeability (NewtonRaphson.jl:40-63, 204-207)
def gradient1d(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """src/FEM.jl:66-90: one-sided differences at the two ends, central
    differences (y[i+1]-y[i-1])/(x[i+1]-x[i-1]) inside."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    if x.size != y.size or x.size < 2:
        raise ValueError("gradient needs two equal-length arrays of at least 2 points")
    z = np.empty_like(y)
    with np.errstate(divide="ignore", invalid="ignore"):
        z[0] = (y[1] - y[0]) / (x[1] - x[0])
        z[-1] = (y[-1] - y[-2]) / (x[-1] - x[-2])
        z[1:-1] = (y[2:] - y[:-2]) / (x[2:] - x[:-2])
    return z


def material_permeability(HofM: np.ndarray, B: np.ndarray):
    """src/magneticProperties.jl:163-188 (materialPermeability): mu = B/H with
    the non-finite entries set to the largest finite-or-zeroed value, dmu =
    gradient(H, mu) with the non-finite entries set to the smallest value.
    Returns (mu, dmu)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        mu = np.asarray(B, float) / np.asarray(HofM, float)
    bad = ~np.isfinite(mu)
    mu[bad] = 0.0
    mu[bad] = np.max(mu)
    dmu = gradient1d(HofM, mu)
    bad = ~np.isfinite(dmu)
    dmu[bad] = 0.0
    dmu[bad] = np.min(dmu)
    return mu, dmu


def bspline_eval(t: np.ndarray, c: np.ndarray, x: np.ndarray, k: int = 3) -> np.ndarray:
    """Value of the spline sum_j c[j] B_{j,k}(x) with knots t (length n,
    coefficients c[0:n-k-1]) -- what FITPACK's splev returns, and so what
    Dierckx.jl's Spline1D evaluates (NewtonRaphson.jl:57-63, 204-207). The
    basis comes from the de Boor-Cox recurrence B_{i,j} = w_{i,j} B_{i,j-1}
    + (1 - w_{i+1,j}) B_{i+1,j-1}, w_{i,j}(x) = (x - t_i)/(t_{i+j} - t_i),
    over the knot interval t[l] <= x < t[l+1] (l <= n-k-2, so the right end
    t[n-k-1] belongs to the last interval). x outside [t[k], t[n-k-1]] raises
    (Spline1D's bc = "error")."""
    t = np.asarray(t, float)
    c = np.asarray(c, float)
    x = np.atleast_1d(np.asarray(x, float))
    n = t.size
    lo, hi = t[k], t[n - k - 1]
    if np.any(~(x >= lo) | ~(x <= hi)):
        raise ValueError("spline argument outside the table range")
    out = np.empty_like(x)
    for q, xv in enumerate(x):
        l = int(np.searchsorted(t, xv, side="right")) - 1
        l = min(max(l, k), n - k - 2)
        # b[i] = B_{l-k+i, j} at xv; degree 0: only B_{l,0} = 1. Updating in
        # increasing i reads b[i+1] before it is overwritten.
        b = np.zeros(k + 2)
        b[k] = 1.0
        for j in range(1, k + 1):
            for i in range(k - j, k + 1):
                idx = l - k + i
                d0 = t[idx + j] - t[idx]
                d1 = t[idx + j + 1] - t[idx + 1]
                left = (xv - t[idx]) / d0 * b[i] if d0 > 0.0 else 0.0
                right = (t[idx + j + 1] - xv) / d1 * b[i + 1] if d1 > 0.0 else 0.0
                b[i] = left + right
        out[q] = float(np.dot(c[l - k:l + 1], b[:k + 1]))
    return out


def newton_spline(mesh, mask, g, magnet: np.ndarray, spl_mu, spl_dmu, rtol: float = 1e-10,
                  tol_H: float = 1e-4, max_outer: int = 100, newton: bool = True, switch: float = np.inf):
    """The loop with the tabulated law of NewtonRaphson.jl:204-207: mu and
    dmu from two splines (each (t, c), FITPACK form, cubic) of |H| on the
    magnet; stops when mu0 max |H_new - H_old| over magnet elements < tol_H
    (NewtonRaphson.jl:200, H_old = 0 before the first step). switch = inf
    takes Newton steps from iteration 2 on, as the reference (att > 1)."""
    law = lambda h: (bspline_eval(spl_mu[0], spl_mu[1], h), bspline_eval(spl_dmu[0], spl_dmu[1], h))
    return _newton_loop(mesh, mask, g, magnet, law, lambda mn, mo, hn, ho: MU0 * np.max(np.abs(hn - ho)),
                        tol_H, max_outer, newton, switch, rtol)


# --------------------------------------------------------------------------
# The reference's nonlinear driver: Neumann + Lagrange saddle Newton with
# lineSearch damping (Magnetostatics/NewtonRaphson.jl:109-212).
def newton_saddle(mesh, RHS: np.ndarray, magnet: np.ndarray, law, tol_H: float = 1e-4, max_att: int = 100,
                  min_step: float = 1e-4):
    """NewtonRaphson.jl:148-212 step by step, with the correct H_z of
    manyBodyIntersecting.jl:241-266 (the driver's Hfield[2,k] -= u[nd]*d is a
    typo). RHS = -BoundaryIntegral(mesh, Bext, shell_id) (:109); law(Hn) ->
    (mu, dmu) on the magnet elements (spl, spl_dmu of :57-63). Per attempt:
    A = A(mu) on the fixed pattern (:154-161), At = tangentialStiffnessMatrix
    of the previous field and dmu from the second attempt on (:164-168), du
    from the bordered system [A+At Lag; Lag' 0] \\ [RHS - A u; 0] (:171,
    direct solve), u updated by lineSearch(u, du, A, RHS) (:175), then H,
    div = mu0 max |H - Hold| over the magnet (:200), mu and dmu from the law
    (:204-207) and the non-finite check (:210-212); the loop runs while
    div > tol_H and att < max_att. Returns dict(phi, H, Hn, mu, dmu, iters,
    div, residues, divs)."""
    Ak = localStiffnessMatrix(mesh)
    Lag = lagrange(mesh)
    mu = np.full(mesh.nt, MU0)
    dmu = np.zeros(mesh.nt)
    u = np.zeros(mesh.nv)
    H = np.zeros((3, mesh.nt))
    Hn = np.zeros(mesh.nt)
    att = 0
    div = np.inf
    residues, divs = [], []
    while div > tol_H and att < max_att:
        att += 1
        Hold = Hn.copy()
        A = sparse_from_local(mesh, Ak, mu)
        if att > 1:
            At = tangentialStiffnessMatrix(mesh, H, dmu)
            J = CSC(A.n, A.colptr, A.rowval, A.nzval + At.nzval)
        else:
            J = A
        du, _lam = saddle_solve(J, Lag, RHS - spmv(A, u))
        residues.append(lineSearch(u, du, A, RHS, min_step))
        H, Hn = field(mesh, u)
        div = float(MU0 * np.max(np.abs(Hn[magnet] - Hold[magnet]))) if magnet.any() else 0.0
        divs.append(div)
        mu_m, dmu_m = law(Hn[magnet])
        mu[magnet] = mu_m
        dmu[magnet] = dmu_m
        if not (np.all(np.isfinite(mu)) and np.all(np.isfinite(dmu))):
            raise FloatingPointError("Nans/Infs in mu")
    return {"phi": u, "H": H, "Hn": Hn, "mu": mu, "dmu": dmu, "iters": att, "div": div, "residues": residues,
            "divs": divs}


# --------------------------------------------------------------------------
# Adjoint of the field-target objective (no reference counterpart).
def adjoint_gradient(mesh, mask, g, rho, target: np.ndarray, Hstar, rtol: float = 1e-13,
                     mur=lambda r: 1.0 + 999.0 * r ** 3, dmur=lambda r: 3.0 * 999.0 * r ** 2):
    """J = sum_{e in target} VE_e |H_e - H*|^2 with mu_e = mu0 mur(rho_e).

    State:   A(rho)[free,free] phi_f = -A[free,fixed] g.
    Adjoint: A[free,free] lam_f = -(dJ/dphi)_f, lam = 0 on fixed nodes.
    Grad:    dJ/drho_e = mu0 mur'(rho_e) VE_e (sum lam_i grad_i).(sum phi_i grad_i).
    Returns (J, grad, phi, lam)."""
    mu = MU0 * mur(rho)
    A = stiffnessMatrix(mesh, mu)
    phi, _it, _rel, s = pcg_dirichlet(A, mask, g, rtol)
    if s:
        raise RuntimeError(f"oracle PCG status {s}")
    G = element_gradients(mesh)  # (nt,4,3)
    t0 = mesh.t.T - 1  # (nt,4)
    H = -np.einsum("ki,kic->kc", phi[t0], G)
    d = H - np.asarray(Hstar)[None, :]
    w = np.where(target, mesh.VE, 0.0)
    J = float(np.sum(w * np.sum(d * d, axis=1)))
    # dJ/dphi_i = sum_e 2 VE (H-H*) . dH/dphi_i = sum_e -2 VE (H-H*).grad_i
    contrib = -2.0 * w[:, None] * np.einsum("kc,kic->ki", d, G)
    dJ = np.zeros(mesh.nv)
    np.add.at(dJ, t0.ravel(), contrib.ravel())
    lam, _it, _rel, s = pcg_free(A, mask, -dJ, rtol)
    if s:
        raise RuntimeError(f"oracle PCG status {s}")
    glam = np.einsum("ki,kic->kc", lam[t0], G)
    gphi = np.einsum("ki,kic->kc", phi[t0], G)
    grad = MU0 * dmur(rho) * mesh.VE * np.sum(glam * gphi, axis=1)
    return J, grad, phi, lam


def objective(mesh, mask, g, rho, target, Hstar, rtol: float = 1e-13,
              mur=lambda r: 1.0 + 999.0 * r ** 3) -> float:
    """J alone (for finite differences)."""
    mu = MU0 * mur(rho)
    A = stiffnessMatrix(mesh, mu)
    phi, _it, _rel, s = pcg_dirichlet(A, mask, g, rtol)
    if s:
        raise RuntimeError(f"oracle PCG status {s}")
    H, _Hn = field(mesh, phi)
    d = H.T - np.asarray(Hstar)[None, :]
    return float(np.sum(np.where(target, mesh.VE, 0.0) * np.sum(d * d, axis=1)))


def saddle_solve(A: CSC, Lag: np.ndarray, b: np.ndarray):
    """[A Lag; Lag' 0] \\ [b; 0] (Magnetostatics/magneticField.jl:107) by a
    direct sparse LU of the bordered matrix (scipy SuperLU in place of Julia's
    UMFPACK). Returns (u, multiplier)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    S = A.to_scipy()
    L = sp.csc_matrix(Lag.reshape(-1, 1))
    K = sp.bmat([[S, L], [L.T, None]], format="csc")
    sol = sla.spsolve(K, np.concatenate([b, [0.0]]))
    return sol[:-1], sol[-1]


# --------------------------------------------------------------------------
# Micromagnetics: demagnetizing field (Micromagnetics/LL.jl:162-209)
def node_volumes(mesh, inside: np.ndarray) -> np.ndarray:
    """Micromagnetics/permalloy.jl:92-96: sum of VE over the inside elements
    at each node."""
    vol = np.zeros(mesh.nv)
    t0 = np.asarray(mesh.t, dtype=np.int64) - 1
    for k in np.nonzero(inside)[0]:
        vol[t0[:, k]] += mesh.VE[k]
    return vol


def demag_load(mesh, inside: np.ndarray, M: np.ndarray) -> np.ndarray:
    """The load vector of magnetostaticField (LL.jl:170-183): over inside
    elements k, RHS[nds[i]] += VE_k (b_ik, c_ik, d_ik) . mean(M[:, nds])."""
    G = element_gradients(mesh)
    t0 = np.asarray(mesh.t, dtype=np.int64) - 1
    rhs = np.zeros(mesh.nv)
    for k in np.nonzero(inside)[0]:
        nds = t0[:, k]
        f = M[:, nds].sum(axis=1) / 4.0
        for i in range(4):
            rhs[nds[i]] += mesh.VE[k] * float(G[k, i] @ f)
    return rhs


def magnetostatic_field(mesh, inside: np.ndarray, M: np.ndarray, volumes: np.ndarray):
    """magnetostaticField (LL.jl:162-209) with A = stiffnessMatrix(mesh)
    (permalloy.jl:99). The reference's cg(A, RHS) converges on the consistent
    singular system to a solution defined up to a constant; here u comes from
    the direct bordered solve with Lag' u = 0, which has the same gradient.
    Returns (Hd_nodes (3, nv), u, RHS); nodes with volumes == 0 hold NaN
    (0/0, as the reference)."""
    inside = np.asarray(inside).astype(bool)
    A = stiffnessMatrix(mesh)
    rhs = demag_load(mesh, inside, M)
    u, _mult = saddle_solve(A, lagrange(mesh), rhs)
    H, _Hn = field(mesh, u)
    t0 = np.asarray(mesh.t, dtype=np.int64) - 1
    Hd = np.zeros((3, mesh.nv))
    for k in np.nonzero(inside)[0]:
        Hd[:, t0[:, k]] += mesh.VE[k] * H[:, k:k + 1]
    with np.errstate(divide="ignore", invalid="ignore"):
        Hd = Hd / volumes
    return Hd, u, rhs
