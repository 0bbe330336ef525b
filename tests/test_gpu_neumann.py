"""GPU Neumann load and Lagrange saddle solve of Magnetostatics/magneticField.jl
against the oracle's direct solve of the bordered system.
Bars: RHS within 1e-13 relative; u within 1e-9 relative (PCG rtol 1e-11)."""
import numpy as np
import pytest

import femeko_hip as F
from femeko_hip.mesh import MU0, boundary_faces, centroids, elements_in_box, kuhn_box_mesh

pytestmark = pytest.mark.gpu


def test_saddle_solve_matches_direct(oracle):
    m = boundary_faces(kuhn_box_mesh(10, 0.2, seed=41))
    inside = elements_in_box(m, (-0.3,) * 3, (0.3,) * 3)
    mu = np.where(inside, MU0 * 3.0, MU0)          # mu0 (1 + 2) as magneticField.jl:91-92
    Hext = np.array([1.0, 0.0, 0.0]) / MU0
    RHS_ref = oracle.BoundaryIntegral(m, MU0 * Hext, m.shell_id)
    RHS = F.BoundaryIntegral(m, MU0 * Hext, m.shell_id)
    assert np.abs(RHS - RHS_ref).max() <= 1e-13 * np.abs(RHS_ref).max()
    A = oracle.stiffnessMatrix(m, mu)
    u_ref, mult_ref = oracle.saddle_solve(A, oracle.lagrange(m), -RHS_ref)
    u, mult = F.solveSaddle(m, mu, -RHS, rtol=1e-11)
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-9
    assert abs(mult - mult_ref) <= 1e-9 * max(1.0, abs(mult_ref))
    # Lag' u = 0
    assert abs(oracle.lagrange(m) @ u) <= 1e-12 * np.abs(u).max()
    F.invalidate(m)
