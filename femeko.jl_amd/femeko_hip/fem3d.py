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

Sparse results are scipy.sparse.csc_matrix with the arrays of Julia's
SparseMatrixCSC (0-based here, as scipy stores them).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

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
        A.fk_mesh = dm  # the device holds these values: lineSearch runs there
    return A


def lineSearch(u, du, A, RHS, minStep: float = 1e-4) -> float:
    """src/FEM3D.jl:123-152 on the GPU: alpha <- 0.9 alpha until
    ||RHS - A(u + alpha du)|| <= ||RHS - A u||; mutates u. A is a matrix
    returned by stiffnessMatrix (its values are resident on the GPU) or a
    MESH whose last assembled matrix is meant."""
    dm = getattr(A, "fk_mesh", None)
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


def solveSaddle(mesh, mu, b, rtol: float = 1e-10):
    """u of [A Lag; Lag' 0] \\ [b; 0] (Magnetostatics/magneticField.jl:99-108)
    with A = stiffnessMatrix(mesh, mu). Returns (u, multiplier)."""
    dm = device_mesh(mesh)
    dm.set_coefficient(mu)
    dm.assemble()
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
