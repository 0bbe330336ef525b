"""Femeko-named drop-ins for the P1 path of src/FEM3D.jl, running on the GPU.

Same names, argument meaning and results as the reference functions; the
device copy of a MESH is made on first use and cached on the MESH object
(call ``invalidate(mesh)`` after changing mesh.p or mesh.t).

  stiffnessMatrix(mesh, f)                 src/FEM3D.jl:40-56
  localStiffnessMatrix(mesh, f)            src/FEM3D.jl:59-76
  tangentialStiffnessMatrix(mesh, H, dmu)  src/FEM3D.jl:79-119
  lagrange(mesh)                           src/FEM3D.jl:274-281
  magneticField(mesh, u)                   Magnetostatics/magneticField.jl:111-130
  solveDirichlet(mesh, mu, mask, g, ...)   the `\\` / cg solve of the drivers
                                           with Dirichlet elimination
                                           (Fluids/fluids3D.jl:142-174)
  scalarPotential(mesh, mu, Hext, shell)   the C++ bridge of
                                           Femeko_jl_C++/Magnetostatics/run.jl:66-79

Sparse results are scipy.sparse.csc_matrix with the arrays of Julia's
SparseMatrixCSC (0-based here, as scipy stores them).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import FK_ERR_STATE, FemekoHIPError, dptr, iptr
from .device import DeviceMesh, context


def device_mesh(mesh, device: int = 0) -> DeviceMesh:
    """The cached DeviceMesh of a MESH (uploaded on first use)."""
    dm = getattr(mesh, "_fk_device", None)
    if dm is None or dm.handle is None:
        dm = DeviceMesh(mesh, context(device))
        mesh._fk_device = dm
    return dm


def invalidate(mesh) -> None:
    dm = getattr(mesh, "_fk_device", None)
    if dm is not None:
        dm.close()
        mesh._fk_device = None


def _csc(dm: DeviceMesh, values: np.ndarray, resident: bool = False) -> sp.csc_matrix:
    rowptr, colidx = dm.pattern(base=0)
    # the pattern and the values are exactly symmetric: CSR arrays == CSC arrays
    A = sp.csc_matrix((values, colidx, rowptr), shape=(dm.nv, dm.nv))
    if resident:
        # the device holds these values: lineSearch runs there, as long as no
        # later call on the mesh has replaced them
        A.fk_mesh = dm
        A.fk_generation = dm.values_generation()
    return A


def lineSearch(u, du, A, RHS, minStep: float = 1e-4) -> float:
    """src/FEM3D.jl:123-152 on the GPU: alpha <- 0.9 alpha until
    ||RHS - A(u + alpha du)|| <= ||RHS - A u||; mutates u. A is a matrix
    returned by stiffnessMatrix (its values are resident on the GPU) or a
    MESH whose last assembled matrix is meant. A matrix whose resident values
    a later call on the same mesh has replaced (another stiffnessMatrix,
    solveSaddle, solveDirichlet, a nonlinear loop) raises FemekoHIPError
    rather than searching along another matrix."""
    dm = getattr(A, "fk_mesh", None)
    if dm is not None and dm.values_generation() != getattr(A, "fk_generation", None):
        raise FemekoHIPError(FK_ERR_STATE, "lineSearch: the matrix's values on the device were replaced by a "
                                           "later assembly on this mesh; call stiffnessMatrix again")
    if dm is None and hasattr(A, "p") and hasattr(A, "t"):
        dm = device_mesh(A)
    if dm is None:
        raise TypeError("lineSearch needs the matrix returned by stiffnessMatrix or a MESH")
    return dm.line_search(u, du, RHS, minStep)


def stiffnessMatrix(mesh, f=None) -> sp.csc_matrix:
    """Global sparse stiffness matrix sum_e VE_e f_e (b b' + c c' + d d')."""
    dm = device_mesh(mesh)
    dm.set_coefficient(np.ones(dm.nt) if f is None else f)
    dm.assemble()
    return _csc(dm, dm.values(), resident=True)


def localStiffnessMatrix(mesh, f=None) -> np.ndarray:
    """Ak (16, nt): column k = vec(VE_k f_k (b b' + c c' + d d'))."""
    return device_mesh(mesh).local_stiffness(f)


def tangentialStiffnessMatrix(mesh, H_vec, dmu) -> sp.csc_matrix:
    """Newton tangent: per element VE dmu/|H| (H . grad_i)(H . grad_j)."""
    dm = device_mesh(mesh)
    return _csc(dm, dm.tangential_values(H_vec, dmu))


def lagrange(mesh) -> np.ndarray:
    """C[nds] += VE/4 over all elements."""
    return device_mesh(mesh).lagrange()


def BoundaryIntegral(mesh, F, shell_id) -> np.ndarray:
    """RHS[nds] += dot(normal, F) * AE / 3 over the tagged surface triangles."""
    return device_mesh(mesh).boundary_integral(mesh.surfaceT, mesh.normal, mesh.AE, F, shell_id)


def solveSaddle(mesh, mu, b, rtol: float = 1e-10, precision: str = "f64"):
    """u of [A Lag; Lag' 0] \\ [b; 0] (Magnetostatics/magneticField.jl:99-108)
    with A = stiffnessMatrix(mesh, mu); precision 'f64' or 'f32' (FP32-refined
    PCG on the consistent singular system). Returns (u, multiplier)."""
    dm = device_mesh(mesh)
    dm.set_coefficient(mu)
    dm.assemble()
    dm.set_inner_precision(precision)
    u, mult, _it, _rel = dm.solve_neumann(b, rtol)
    return u, mult


def magneticField(mesh, u) -> np.ndarray:
    """H (3, nt) = -sum_nd u[nd] grad(lambda_nd) per element."""
    dm = device_mesh(mesh)
    dm.set_potential(u)
    H, _Hn, _E = dm.field(want_H=True, want_Hn=False)
    return H


def solveDirichlet(mesh, mu, mask, g, rtol: float = 1e-10, precision: str = "f64",
                   inner_rtol: float | None = None, max_iter: int = 1_000_000):
    """Assemble A(mu), fix phi = g on mask, Jacobi-PCG for the rest.
    Returns (u, SolveReport)."""
    dm = device_mesh(mesh)
    dm.set_coefficient(mu)
    dm.assemble()
    dm.set_dirichlet(mask, g)
    rep = dm.solve(precision=precision, rtol=rtol, inner_rtol=inner_rtol, max_iter=max_iter)
    return dm.potential(), rep


def scalarPotential(mesh, mu, Hext, shell_id) -> np.ndarray:
    """u of the reference's C++ bridge scalarPotential
    (Femeko_jl_C++/Magnetostatics/magnetostatics.cpp:49-93) through
    fk_scalar_potential, called as run.jl:47-79 does: 0-based copies of t and
    surfaceT (node numbers and tag) and shell_id - 1; triangle areas from p.
    Context-free: the mesh is uploaded, solved and freed in the call."""
    L = _lib.load()
    p = np.ascontiguousarray(np.asarray(mesh.p, dtype=np.float64).T)
    t = np.ascontiguousarray((np.asarray(mesh.t, dtype=np.int32) - 1).T)
    sT = np.ascontiguousarray((np.asarray(mesh.surfaceT, dtype=np.int32) - 1).T)
    nrm = np.ascontiguousarray(np.asarray(mesh.normal, dtype=np.float64).T)
    VE = np.ascontiguousarray(mesh.VE, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    H = np.ascontiguousarray(Hext, dtype=np.float64)
    nv, nt, ns = p.shape[0], t.shape[0], sT.shape[0]
    if mu.size != nt or H.size != 3 or VE.size != nt:
        raise FemekoHIPError(_lib.FK_ERR_INVALID_ARG, "scalarPotential: mu and VE need nt entries, Hext 3")
    u = np.zeros(nv)
    s = L.fk_scalar_potential(dptr(u), dptr(p), iptr(t), iptr(sT), dptr(nrm), nv, nt, ns, dptr(VE), dptr(mu),
                              dptr(H), int(shell_id) - 1)
    if s:
        raise FemekoHIPError(s, L.fk_scalar_potential_last_error().decode())
    return u
