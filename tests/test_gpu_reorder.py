"""Meshes in a scattered node numbering, which the library renumbers inside
(reverse Cuthill-McKee): every entry point that takes or returns per-node
arrays or a matrix, through the C ABI, against the oracle in the caller's
numbering. Bars: pattern bit-exact, values to the 1e-12 term metric, solves
to their tolerance, loads and nodal fields to 1e-12, the nonlinear loop to
the bars of its structured-mesh test (its Newton assembly walking the
library's sorted copy of the elements); and the same results, within
rounding, as the same mesh kept in its own numbering."""
import numpy as np
import pytest

import femeko_hip as F
from femeko_hip import workloads as W
from femeko_hip.device import DeviceMesh, context
from femeko_hip.mesh import boundary_faces

pytestmark = pytest.mark.gpu


def _scattered(n=10, seed=7):
    pr = W.linear(n, seed=seed)
    perm = np.random.default_rng(seed).permutation(pr.mesh.nv)
    return W.renumbered(pr, perm, "random")


def test_scattered_mesh_is_renumbered_and_matches_oracle(oracle):
    pr = _scattered()
    m = pr.mesh
    dm = DeviceMesh(m, context(0))
    try:
        reordered, d_caller, d_int = dm.numbering()
        assert reordered and d_int < 0.2 * d_caller, (d_caller, d_int)
        A = oracle.stiffnessMatrix(m, pr.mu)
        rowptr, colidx = dm.pattern(base=1)
        assert np.array_equal(rowptr, A.colptr) and np.array_equal(colidx, A.rowval)
        dm.set_coefficient(pr.mu)
        dm.assemble()
        scale = oracle.term_scale(m, pr.mu)
        assert np.all(np.abs(dm.values() - A.nzval) <= 1e-12 * scale)
        v32 = dm.values_f32()
        assert np.all(np.abs(v32 - A.nzval) <= 1e-6 * np.abs(A.nzval) + 1e-12 * scale)
        np.testing.assert_allclose(dm.lagrange(), oracle.lagrange(m), rtol=1e-13)
        dm.set_dirichlet(pr.mask, pr.g)
        ref, _it, _rel, _s = oracle.pcg_dirichlet(A, pr.mask, pr.g, 1e-13)
        for prec in ("f64", "f32"):
            dm.solve(prec, rtol=1e-12)
            phi = dm.potential()
            assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) < 1e-10, prec
        # the potential round trip and the field from it
        rng = np.random.default_rng(3)
        u = rng.normal(size=m.nv)
        dm.set_potential(u)
        assert np.array_equal(dm.potential(), u)
        H, _Hn, _E = dm.field(want_H=True, want_Hn=False)
        Href, _ = oracle.field(m, u)
        assert np.abs(H - Href).max() <= 1e-12 * np.abs(Href).max()
        # a right-hand side solve on the free rows
        b = rng.normal(size=m.nv) * (pr.mask == 0)
        x, _its, _rel = dm.solve_rhs(b, rtol=1e-13)
        xr, _i2, _r2, _s2 = oracle.pcg_free(A, pr.mask, b, 1e-13)
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-10
        # tangent values in the caller's CSR order
        Hv = rng.normal(size=(3, m.nt))
        dmu = -rng.random(m.nt) * 1e-9
        At = dm.tangential_values(Hv, dmu)
        Atr = oracle.tangentialStiffnessMatrix(m, Hv, dmu)
        assert np.all(np.abs(At - Atr.nzval) <= 1e-12 * oracle.tangent_scale(m, Hv, dmu))
    finally:
        dm.close()


def test_scattered_mesh_line_search_and_neumann(oracle):
    pr = _scattered(8, seed=9)
    m = boundary_faces(pr.mesh)
    A = oracle.stiffnessMatrix(m, pr.mu)
    S = A.to_scipy().tocsr()
    rng = np.random.default_rng(9)
    RHS = rng.normal(size=m.nv)
    u0 = rng.normal(size=m.nv)
    du = -np.linalg.lstsq(S.toarray(), S @ u0 - RHS, rcond=None)[0] * 3.0
    Ag = F.stiffnessMatrix(m, pr.mu)
    assert F.device_mesh(m).numbering()[0]
    u_ref, u = u0.copy(), u0.copy()
    res_ref = oracle.lineSearch(u_ref, du, A, RHS)
    res = F.lineSearch(u, du, Ag, RHS)
    assert res == pytest.approx(res_ref, rel=1e-10)
    np.testing.assert_allclose(u, u_ref, rtol=0, atol=1e-12 * np.abs(u_ref).max())
    # Neumann load and saddle solve
    Fv = np.array([0.4, 1.0, -0.3])
    R = F.BoundaryIntegral(m, Fv, m.shell_id)
    Rr = oracle.BoundaryIntegral(m, Fv, m.shell_id)
    assert np.abs(R - Rr).max() <= 1e-13 * np.abs(Rr).max()
    us, mult = F.solveSaddle(m, pr.mu, -R, rtol=1e-12)
    ur, mr = oracle.saddle_solve(A, oracle.lagrange(m), -Rr)
    assert np.linalg.norm(us - ur) / np.linalg.norm(ur) < 1e-9
    F.invalidate(m)


def test_scattered_mesh_demag(oracle):
    pr = _scattered(8, seed=11)
    m = pr.mesh
    inside = pr.magnet
    vol = oracle.node_volumes(m, inside)
    dm = DeviceMesh(m, context(0))
    try:
        assert dm.numbering()[0]
        dm.set_coefficient(np.ones(m.nt))
        dm.assemble()
        M = np.random.default_rng(11).normal(size=(3, m.nv))
        ref, _u, _rhs = oracle.magnetostatic_field(m, inside, M, vol)
        Hd, _its, _rel = dm.demag_field(inside.astype(np.uint8), M, vol, rtol=1e-12)
        nan_ref = np.isnan(ref)
        assert np.array_equal(np.isnan(Hd), nan_ref)
        scale = np.abs(ref[~nan_ref]).max()
        assert np.max(np.abs(Hd[~nan_ref] - ref[~nan_ref])) < 1e-9 * scale
    finally:
        dm.close()


def test_renumbered_and_kept_numbering_agree():
    """The same scattered mesh with the renumbering on (auto) and off: the
    same potential to the solve tolerance, the same exported values."""
    pr = _scattered(9, seed=13)
    out = {}
    for mode in ("auto", "none"):
        dm = DeviceMesh(pr.mesh, context(0), reorder=mode)
        try:
            assert dm.numbering()[0] == (mode == "auto")
            dm.set_coefficient(pr.mu)
            dm.assemble()
            dm.set_dirichlet(pr.mask, pr.g)
            dm.solve("f64", rtol=1e-13)
            out[mode] = (dm.values(), dm.potential())
        finally:
            dm.close()
    va, pa = out["auto"]
    vn, pn = out["none"]
    assert np.abs(va - vn).max() <= 1e-13 * np.abs(vn).max()
    assert np.linalg.norm(pa - pn) / np.linalg.norm(pn) < 1e-11


def test_structured_numbering_is_kept():
    pr = W.linear(6)
    dm = DeviceMesh(pr.mesh, context(0))
    try:
        assert not dm.numbering()[0]
        assert dm.solver_format()[0] == "stencil"
    finally:
        dm.close()


@pytest.mark.parametrize("inner_precision", ["f64", "f32"])
def test_scattered_mesh_froehlich_loop(oracle, inner_precision):
    """The nonlinear loop on a scattered mesh: nodes renumbered inside and the
    row-gather Newton assembly walking its sorted copy of the elements (the
    caller's elements do not follow the internal numbering); the same bars
    as on the structured mesh."""
    pr = W.froehlich(n=8, seed=12)
    perm = np.random.default_rng(21).permutation(pr.mesh.nv)
    pr = W.renumbered(pr, perm, "random")
    Ms, a = pr.meta["Ms"], pr.meta["a"]
    ref = oracle.newton_froehlich(pr.mesh, pr.mask, pr.g, pr.magnet, Ms, a)
    dm = DeviceMesh(pr.mesh, context(0))
    try:
        assert dm.numbering()[0]
        dm.set_dirichlet(pr.mask, pr.g)
        dm.set_inner_precision(inner_precision)
        out = dm.nonlinear_froehlich(pr.magnet.astype(np.uint8), Ms, a, rtol=1e-10, tol_mu=1e-6, max_outer=20)
        assert out["outer"] == ref["iters"]
        phi = dm.potential()
        assert np.linalg.norm(phi - ref["phi"]) / np.linalg.norm(ref["phi"]) < 1e-8
        np.testing.assert_allclose(dm.get_coefficient(), ref["mu"], rtol=1e-6)
        # the plain assembly after the loop reads the loop's mu in the sorted
        # element order: the oracle's matrix of that mu
        dm.assemble()
        A = oracle.stiffnessMatrix(pr.mesh, dm.get_coefficient())
        assert np.all(np.abs(dm.values() - A.nzval) <= 1e-12 * oracle.term_scale(pr.mesh, dm.get_coefficient()))
    finally:
        dm.close()
