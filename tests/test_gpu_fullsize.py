"""The workload configurations at their full sizes, checked through
properties that hold at any size (the oracle is too slow at these sizes):

- n = 160 boxes (4.17M nodes, the headline): the pattern's closed-form size,
  zero row sums (rows of a P1 stiffness matrix sum to zero), symmetry
  x.Ay = y.Ax, the energy identity 1/2 phi.A phi = 1/2 sum mu |H|^2 VE, and
  both the FP32-refined (rtol 5e-11) and the FP64 (rtol 1e-10) potentials
  within 1e-9 of an FP64 solve at rtol 1e-13;
- n = 100 Froehlich magnet: the loop's fixed point mu = mu0 (1 + Ms/(a+|H|))
  at the returned potential;
- 64 meshes x 3 fields: every system below relres 1e-10, and sampled systems
  equal to single-mesh FP64 solves;
- n = 80 adjoint: the gradient against central differences on 32 seeded
  elements (the 16 largest gradients and 16 uniform draws).

test_boxes160_against_oracle checks the headline workload against the
oracle's FP64 Jacobi-PCG on the host (OpenMP, rtol 1e-13): phi, |H| and the
energy of the FP32-refined solve within 1e-9.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from femeko_hip import workloads as W
from femeko_hip.batch import Batch
from femeko_hip.device import DeviceMesh, context
from femeko_hip.mesh import MU0, kuhn_counts
from femeko_hip.workloads import dirichlet_values

pytestmark = pytest.mark.gpu


def test_boxes160_properties():
    pr = W.boxes(160)
    m = pr.mesh
    dm = DeviceMesh(m, context(0))
    try:
        assert dm.nnz == kuhn_counts(160)["nnz"]
        dm.set_coefficient(pr.mu)
        dm.assemble()
        rowptr, colidx = dm.pattern(base=0)
        val = dm.values()
        A = sp.csr_matrix((val, colidx, rowptr), shape=(m.nv, m.nv))
        rows = np.repeat(np.arange(m.nv), np.diff(rowptr))
        scale = np.bincount(rows, weights=np.abs(val), minlength=m.nv)
        assert np.max(np.abs(A @ np.ones(m.nv)) / scale) < 1e-12
        rng = np.random.default_rng(160)
        x, y = rng.normal(size=m.nv), rng.normal(size=m.nv)
        assert x @ (A @ y) == pytest.approx(y @ (A @ x), rel=1e-12)
        dm.set_dirichlet(pr.mask, pr.g)
        dm.solve("f64", rtol=1e-13)
        ref = dm.potential()
        _H, _Hn, E = dm.field(want_H=False, want_Hn=False, want_energy=True)
        assert 0.5 * ref @ (A @ ref) == pytest.approx(E, rel=1e-10)
        for prec, rtol in (("f32", 5e-11), ("f64", 1e-10)):
            rep = dm.solve(prec, rtol=rtol)
            assert rep.relres < rtol
            err = np.linalg.norm(dm.potential() - ref) / np.linalg.norm(ref)
            assert err < 1e-9, (prec, err)
    finally:
        dm.close()


def test_froehlich100_fixed_point():
    pr = W.froehlich()
    m = pr.mesh
    Ms, a = pr.meta["Ms"], pr.meta["a"]
    dm = DeviceMesh(m, context(0))
    try:
        dm.set_dirichlet(pr.mask, pr.g)
        out = dm.nonlinear_froehlich(pr.magnet.astype(np.uint8), Ms, a, rtol=1e-10, tol_mu=1e-6, max_outer=20)
        assert out["outer"] <= 20 and out["change"] < 1e-6
        mu = dm.get_coefficient()
        _H, Hn, _E = dm.field(want_H=False, want_Hn=True)
        law = MU0 * (1.0 + Ms / (a + Hn))
        # mu was updated from the field of the last step, which moved it by
        # less than the stop tolerance
        np.testing.assert_allclose(mu[pr.magnet], law[pr.magnet], rtol=2e-6)
        np.testing.assert_array_equal(mu[~pr.magnet], MU0)
    finally:
        dm.close()


def test_batch64_systems():
    probs = W.batch()
    gs = [np.stack([np.where(p.mask == 1, dirichlet_values(p.mesh, e), 0.0) for e in np.eye(3)]) for p in probs]
    b = Batch([p.mesh for p in probs], nrhs=3)
    try:
        b.set_coefficient([p.mu for p in probs])
        b.assemble()
        b.set_dirichlet([p.mask for p in probs], gs)
        rep = b.solve(rtol=1e-10)
        assert np.all(rep.relres < 1e-10)
        for i in (0, 17, 63):
            p = probs[i]
            phi = b.potential(i)
            dm = DeviceMesh(p.mesh, context(0))
            try:
                dm.set_coefficient(p.mu)
                dm.assemble()
                for j in range(3):
                    dm.set_dirichlet(p.mask, gs[i][j])
                    dm.solve("f64", rtol=1e-13)
                    ref = dm.potential()
                    assert np.linalg.norm(phi[j] - ref) / np.linalg.norm(ref) < 1e-9, (i, j)
            finally:
                dm.close()
    finally:
        b.close()


def test_adjoint80_central_differences():
    from femeko_hip import adjoint as ADJ
    from femeko_hip import fem3d
    pr = W.adjoint()
    m = pr.mesh
    rho = pr.meta["rho"]
    tg = pr.magnet.astype(np.uint8)
    Hs = pr.meta["Hstar"]
    try:
        _J, grad, _its = ADJ.fieldObjectiveGradient(m, rho, tg, Hs, pr.mask, pr.g)
        gmax = float(np.abs(grad).max())
        h = 1e-6
        rng = np.random.default_rng(80)
        idx = np.concatenate([np.argsort(-np.abs(grad))[:16], rng.choice(m.nt, 16, replace=False)])
        for e in idx:
            rp, rm = rho.copy(), rho.copy()
            rp[e] += h
            rm[e] -= h
            fd = (ADJ.fieldObjective(m, rp, tg, Hs, pr.mask, pr.g, rtol=1e-13)
                  - ADJ.fieldObjective(m, rm, tg, Hs, pr.mask, pr.g, rtol=1e-13)) / (2 * h)
            assert abs(fd - grad[e]) < 1e-4 * gmax, (e, fd, grad[e])
            if abs(grad[e]) >= 1e-2 * gmax:  # below, the differences' rounding noise dominates
                assert abs(fd - grad[e]) < 2e-4 * abs(grad[e]), (e, fd, grad[e])
    finally:
        fem3d.invalidate(m)


def test_boxes160_against_oracle(oracle):
    """The FP32-refined solve (rtol 5e-11) of the n = 160 workload against the
    oracle: its assembly (src/FEM3D.jl:40-76 restated, OpenMP) and its FP64
    Jacobi-PCG to rtol 1e-13 on the host cores; phi (relative L2), |H|
    (VE-weighted relative L2) and the energy within 1e-9."""
    pr = W.boxes(160)
    m = pr.mesh
    A = oracle.stiffnessMatrix(m, pr.mu, nthreads=oracle.threads())
    ref, it, rel, s = oracle.pcg_dirichlet(A, pr.mask, pr.g, 1e-13, exact=False)
    assert s == 0 and rel < 1e-13
    del A
    _H, Hn_ref = oracle.field(m, ref)
    E_ref = oracle.energy(m, Hn_ref, pr.mu)
    dm = DeviceMesh(m, context(0))
    try:
        dm.set_coefficient(pr.mu)
        dm.assemble()
        dm.set_dirichlet(pr.mask, pr.g)
        rep = dm.solve("f32", rtol=5e-11)
        assert rep.relres < 5e-11
        phi = dm.potential()
        _H, Hn, E = dm.field(want_H=False, want_Hn=True, want_energy=True)
    finally:
        dm.close()
    VE = m.VE
    err_phi = np.linalg.norm(phi - ref) / np.linalg.norm(ref)
    err_H = np.sqrt(np.sum(VE * (Hn - Hn_ref) ** 2) / np.sum(VE * Hn_ref ** 2))
    err_E = abs(E - E_ref) / abs(E_ref)
    print(f"n=160 vs oracle (PCG {it} its to {rel:.1e}): phi {err_phi:.2e} |H| {err_H:.2e} energy {err_E:.2e}")
    assert err_phi < 1e-9 and err_H < 1e-9 and err_E < 1e-9
