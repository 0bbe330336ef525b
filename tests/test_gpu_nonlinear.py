"""GPU parity of the nonlinear mu(|H|) loop, lineSearch and the Femeko-named
drop-ins (femeko_hip.fem3d) against the oracle.

Bars: same Picard/Newton step sequence and outer iteration count as the
oracle; potential within 1e-8 relative, mu within 1e-6 relative (both loops
stop at a relative mu change below 1e-6, with each inner solve at rtol 1e-10).
"""
import numpy as np
import pytest

import femeko_hip as F
from femeko_hip import workloads as W
from femeko_hip.device import DeviceMesh, context

import linesearch_cases as LS

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("inner_precision", ["f64", "f32"])
@pytest.mark.parametrize("n,seed", [(8, 12), (12, 100)])
def test_froehlich_loop_matches_oracle(oracle, n, seed, inner_precision):
    # f32: the update solves run FP32-refined (FP64 true residual of each
    # step's own system) to the same rtol; same bars
    pr = W.froehlich(n=n, seed=seed)
    Ms, a = pr.meta["Ms"], pr.meta["a"]
    ref = oracle.newton_froehlich(pr.mesh, pr.mask, pr.g, pr.magnet, Ms, a)
    dm = DeviceMesh(pr.mesh, context(0))
    try:
        dm.set_dirichlet(pr.mask, pr.g)
        dm.set_inner_precision(inner_precision)
        out = dm.nonlinear_froehlich(pr.magnet.astype(np.uint8), Ms, a, rtol=1e-10, tol_mu=1e-6, max_outer=20)
        assert out["outer"] == ref["iters"]
        assert out["change"] < 1e-6
        phi = dm.potential()
        assert np.linalg.norm(phi - ref["phi"]) / np.linalg.norm(ref["phi"]) < 1e-8
        mu = dm.get_coefficient()
        np.testing.assert_allclose(mu, ref["mu"], rtol=1e-6)
        H, Hn, E = dm.field(want_energy=True)
        np.testing.assert_allclose(Hn, ref["Hn"], rtol=1e-6, atol=1e-9 * ref["Hn"].max())
        assert E == pytest.approx(oracle.energy(pr.mesh, ref["Hn"], ref["mu"]), rel=1e-6)
    finally:
        dm.close()


def test_line_search_matches_oracle(oracle):
    pr = W.linear(6)
    A = oracle.stiffnessMatrix(pr.mesh, pr.mu)
    S = A.to_scipy().tocsr()
    rng = np.random.default_rng(10)
    RHS = rng.normal(size=pr.mesh.nv) * 1e-8
    du = 50.0 * np.linalg.solve(S.toarray() + np.eye(pr.mesh.nv) * 1e-3 * S.diagonal().max(), RHS)
    u_ref = np.zeros(pr.mesh.nv)
    res_ref = oracle.lineSearch(u_ref, du, A, RHS)
    Ag = F.stiffnessMatrix(pr.mesh, pr.mu)
    u = np.zeros(pr.mesh.nv)
    res = F.lineSearch(u, du, Ag, RHS)
    assert res == pytest.approx(res_ref, rel=1e-10)
    np.testing.assert_allclose(u, u_ref, rtol=1e-12, atol=1e-12 * np.abs(u_ref).max())
    F.invalidate(pr.mesh)


@pytest.mark.parametrize("case", range(len(LS.CASES)))
def test_line_search_closed_form(oracle, case):
    """fk_line_search against the closed-form accepted step of
    tests/linesearch_cases.py (alpha* = 2 r0.w/||w||^2, src/FEM3D.jl:123-152)."""
    astar, min_step, k_expect = LS.CASES[case]
    pr = W.linear(6)
    A = oracle.stiffnessMatrix(pr.mesh, pr.mu)
    S = A.to_scipy().tocsr()
    rng = np.random.default_rng(100 + case)
    u, du, RHS = LS.make_case(S, rng, astar, pr.mesh.nv)
    k, _alpha, res_expect, u_expect = LS.expected(S, u, du, RHS, min_step)
    assert k == k_expect
    Ag = F.stiffnessMatrix(pr.mesh, pr.mu)
    u0 = u.copy()
    res = F.lineSearch(u, du, Ag, RHS, min_step)
    assert res == pytest.approx(res_expect, rel=1e-11)
    if k is None:
        assert np.array_equal(u, u0)
    else:
        np.testing.assert_allclose(u, u_expect, rtol=0, atol=1e-13 * np.abs(u_expect).max())
    F.invalidate(pr.mesh)


def test_femeko_named_dropins(oracle):
    pr = W.linear(10)
    m = pr.mesh
    A = F.stiffnessMatrix(m, pr.mu)
    Ar = oracle.stiffnessMatrix(m, pr.mu)
    assert np.array_equal(A.indptr + 1, Ar.colptr) and np.array_equal(A.indices + 1, Ar.rowval)
    scale = oracle.term_scale(m, pr.mu)
    assert np.all(np.abs(A.data - Ar.nzval) <= 1e-12 * scale)
    A1 = F.stiffnessMatrix(m)  # f defaults to ones
    assert np.all(np.abs(A1.data - oracle.stiffnessMatrix(m).nzval) <= 1e-12 * oracle.term_scale(m))
    Ak = F.localStiffnessMatrix(m, pr.mu)
    assert Ak.shape == (16, m.nt)
    np.testing.assert_allclose(F.lagrange(m), oracle.lagrange(m), rtol=1e-13)
    u, rep = F.solveDirichlet(m, pr.mu, pr.mask, pr.g, rtol=1e-12)
    ref, _it, _rel, _s = oracle.pcg_dirichlet(Ar, pr.mask, pr.g, 1e-13)
    assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-10
    H = F.magneticField(m, u)
    Href, _ = oracle.field(m, u)
    assert np.abs(H - Href).max() <= 1e-12 * np.abs(Href).max()
    rng = np.random.default_rng(4)
    Hv = rng.normal(size=(3, m.nt))
    dmu = -rng.random(m.nt) * 1e-9
    At = F.tangentialStiffnessMatrix(m, Hv, dmu)
    Atr = oracle.tangentialStiffnessMatrix(m, Hv, dmu)
    sc = oracle.tangent_scale(m, Hv, dmu)
    assert np.all(np.abs(At.data - Atr.nzval) <= 1e-12 * sc)
    F.invalidate(m)


def test_line_search_refuses_replaced_values(oracle):
    """A matrix object whose device values a later assembly replaced is not
    searched along: lineSearch raises instead of using the other matrix."""
    pr = W.linear(6)
    A1 = F.stiffnessMatrix(pr.mesh, pr.mu)
    A2 = F.stiffnessMatrix(pr.mesh, 2.0 * pr.mu)
    u = np.zeros(pr.mesh.nv)
    du = np.ones(pr.mesh.nv)
    RHS = np.zeros(pr.mesh.nv)
    with pytest.raises(F.FemekoHIPError):
        F.lineSearch(u, du, A1, RHS)
    F.lineSearch(u, du, A2, RHS)  # the current values
    F.invalidate(pr.mesh)


def _saddle_case(oracle, tabulated):
    import os
    from femeko_hip.material import Spline1D, material_permeability
    from femeko_hip.mesh import MU0, boundary_faces
    if tabulated:
        # the reference driver's own setting: Gd at 270 K, Bext = 0.5 T (NewtonRaphson.jl:30-36)
        pr = W.tabulated(n=8, seed=12)
        m = boundary_faces(pr.mesh)
        RHS = -oracle.BoundaryIntegral(m, np.array([0.5, 0.0, 0.0]), m.shell_id)
        tab = np.load(os.path.join(os.path.dirname(__file__), "golden", "gd270_bh.npz"))
        mu_t, dmu_t, _ = material_permeability(tab["H"], tab["B"])
        spl, spld = Spline1D(tab["H"], mu_t), Spline1D(tab["H"], dmu_t)
        law_ref = lambda h: (oracle.bspline_eval(spl.t, spl.c, h), oracle.bspline_eval(spld.t, spld.c, h))  # noqa
        return m, pr.magnet, RHS, law_ref, ("spline", spl, spld), 1e-6
    pr = W.froehlich(n=8, seed=12)
    m = boundary_faces(pr.mesh)
    RHS = -oracle.BoundaryIntegral(m, MU0 * np.array([1e4, 3e3, -2e3]), m.shell_id)
    return m, pr.magnet, RHS, (lambda h: oracle.froehlich(h, 1e5, 500.0)), ("froehlich", 1e5, 500.0), 1e-9


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("tabulated", [False, True])
def test_newton_saddle_matches_oracle(oracle, tabulated, precision):
    """fk_newton_saddle (NewtonRaphson.jl:109-212 on the GPU) against the
    oracle's restatement with direct bordered solves: the same number of
    attempts, potential within 1e-8, mu within 1e-6, the last lineSearch
    residue within 1e-6 relative."""
    m, magnet, RHS, law_ref, law, tol = _saddle_case(oracle, tabulated)
    ref = oracle.newton_saddle(m, RHS, magnet, law_ref, tol_H=tol)
    dm = DeviceMesh(m, context(0))
    try:
        dm.set_inner_precision(precision)
        out = dm.newton_saddle(magnet.astype(np.uint8), RHS, law, rtol=1e-12, tol_H=tol)
        assert out["outer"] == ref["iters"]
        assert out["change"] <= tol
        phi = dm.potential()
        assert np.linalg.norm(phi - ref["phi"]) / np.linalg.norm(ref["phi"]) < 1e-8
        np.testing.assert_allclose(dm.get_coefficient(), ref["mu"], rtol=1e-6)
        assert out["residue"] == pytest.approx(ref["residues"][-1], rel=1e-6, abs=1e-7 * ref["residues"][1])
    finally:
        dm.close()
