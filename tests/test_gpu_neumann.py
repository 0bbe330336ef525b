"""GPU Neumann load and Lagrange saddle solve of Magnetostatics/magneticField.jl
against the oracle's direct solve of the bordered system.
Bars: RHS within 1e-13 relative; u within 1e-9 relative (PCG rtol 1e-11, FP64
or FP32-refined)."""
import numpy as np
import pytest

import femeko_hip as F
from femeko_hip.mesh import MU0, boundary_faces, centroids, elements_in_box, kuhn_box_mesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_saddle_solve_matches_direct(oracle, precision):
    m = boundary_faces(kuhn_box_mesh(10, 0.2, seed=41))
    inside = elements_in_box(m, (-0.3,) * 3, (0.3,) * 3)
    mu = np.where(inside, MU0 * 3.0, MU0)          # mu0 (1 + 2) as magneticField.jl:91-92
    Hext = np.array([1.0, 0.0, 0.0]) / MU0
    RHS_ref = oracle.BoundaryIntegral(m, MU0 * Hext, m.shell_id)
    RHS = F.BoundaryIntegral(m, MU0 * Hext, m.shell_id)
    assert np.abs(RHS - RHS_ref).max() <= 1e-13 * np.abs(RHS_ref).max()
    A = oracle.stiffnessMatrix(m, mu)
    u_ref, mult_ref = oracle.saddle_solve(A, oracle.lagrange(m), -RHS_ref)
    u, mult = F.solveSaddle(m, mu, -RHS, rtol=1e-11, precision=precision)
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-9
    assert abs(mult - mult_ref) <= 1e-9 * max(1.0, abs(mult_ref))
    # Lag' u = 0
    assert abs(oracle.lagrange(m) @ u) <= 1e-12 * np.abs(u).max()
    F.invalidate(m)


def _bridge_reference(oracle, m, mu, Hext, shell):
    """scalarPotential's u from the oracle: BoundaryIntegral (areas from p),
    the bordered direct solve of A u = -RHS, then the constant that PCG from
    zero keeps, sum_i a_ii u_i = 0."""
    RHS = oracle.BoundaryIntegral(m, Hext, (shell,))
    A = oracle.stiffnessMatrix(m, mu)
    u_s, _mult = oracle.saddle_solve(A, oracle.lagrange(m), -RHS)
    d = A.to_scipy().diagonal()
    return u_s - (d @ u_s) / d.sum()


def test_scalar_potential_bridge_matches_oracle(oracle):
    """fk_scalar_potential (the drop-in for magnetostatics.cpp:49-93) called
    with run.jl's 0-based arrays, against the oracle's saddle solve of the
    same system; a sphere-free cube magnet of mu0 (1 + 2) in air (run.jl:53-55)."""
    m = boundary_faces(kuhn_box_mesh(10, 0.2, seed=42))
    inside = elements_in_box(m, (-0.3,) * 3, (0.3,) * 3)
    mu = np.where(inside, MU0 * 3.0, MU0)
    Hext = np.array([1.0, 0.0, 0.0])
    u = F.scalarPotential(m, mu, Hext, m.shell_id[0])
    u_ref = _bridge_reference(oracle, m, mu, Hext, m.shell_id[0])
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-8


def test_scalar_potential_bridge_patch_test():
    """Uniform mu: the weak form gives mu grad u = -Hext, so the P1 solution is
    exactly u = -Hext.x / mu + c on any mesh (jittered too), with c fixed by
    sum a_ii u_i = 0 (tag 1 shifted to 0 as run.jl does)."""
    m = boundary_faces(kuhn_box_mesh(8, 0.2, seed=43))
    mu = np.full(m.nt, MU0)
    Hext = np.array([0.3, -1.0, 0.5])
    u = F.scalarPotential(m, mu, Hext, 1)
    lin = -(Hext @ m.p) / MU0
    # the constant: compare differences from the mean
    du = (u - u.mean()) - (lin - lin.mean())
    assert np.abs(du).max() <= 1e-8 * np.abs(lin).max()

