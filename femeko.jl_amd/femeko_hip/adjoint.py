"""Field-target objective and its reverse rule on the GPU.

fieldObjective(mesh, rho, target, Hstar, mask, g) = sum_{e in target} VE_e |H_e - H*|^2
with mu_e = mu0 * mur(rho_e) (default mur = 1 + 999 rho^3), H from the
Dirichlet problem phi = g on mask. ``rrule`` returns (J, pullback) in the
ChainRulesCore convention: pullback(dJ) -> (None, dJ * dJ/drho), the gradient
coming from the hand-written adjoint kernel (fk_adjoint_field_target) and the
chain factor mu0 * mur'(rho).
"""
from __future__ import annotations

import numpy as np

from .fem3d import device_mesh
from .mesh import MU0


def simp(r):
    return 1.0 + 999.0 * r ** 3


def dsimp(r):
    return 3.0 * 999.0 * r ** 2


def _state(mesh, rho, mask, g, mur, rtol, precision):
    """State solve and the precision of the adjoint solve: 'f64' (FP64 PCG
    for both) or 'f32' (FP32-refined, fk_solve and fk_set_inner_precision)."""
    dm = device_mesh(mesh)
    dm.set_coefficient(MU0 * mur(np.asarray(rho, dtype=np.float64)))
    dm.assemble()
    dm.set_dirichlet(mask, g)
    dm.solve(precision, rtol=rtol)
    dm.set_inner_precision(precision)
    return dm


def fieldObjective(mesh, rho, target, Hstar, mask, g, mur=simp, rtol: float = 1e-12,
                   precision: str = "f64") -> float:
    dm = _state(mesh, rho, mask, g, mur, rtol, precision)
    J, _grad, _it = dm.adjoint_field_target(target, Hstar, rtol)
    return J


def fieldObjectiveGradient(mesh, rho, target, Hstar, mask, g, mur=simp, dmur=dsimp, rtol: float = 1e-12,
                           precision: str = "f64"):
    """(J, dJ/drho, adjoint iterations)."""
    dm = _state(mesh, rho, mask, g, mur, rtol, precision)
    J, dJdmu, it = dm.adjoint_field_target(target, Hstar, rtol)
    return J, dJdmu * MU0 * dmur(np.asarray(rho, dtype=np.float64)), it


def rrule(f, mesh, rho, target, Hstar, mask, g, mur=simp, dmur=dsimp, rtol: float = 1e-12,
          precision: str = "f64"):
    """ChainRulesCore-style rrule of fieldObjective with respect to rho."""
    if f is not fieldObjective:
        raise TypeError("rrule is defined for fieldObjective")
    J, grad, _it = fieldObjectiveGradient(mesh, rho, target, Hstar, mask, g, mur, dmur, rtol, precision)

    def pullback(dJ):
        return None, dJ * grad

    return J, pullback
