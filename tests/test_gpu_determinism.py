"""Bitwise reproducibility of the solver paths. Every reduction sums its
block partials in one fixed order (each block re-reduces the previous
kernel's partials, so all blocks hold the same alpha and beta) and no kernel
adds floating-point values atomically, so two runs on the same inputs must
give the same bits: the single-mesh FP32-refined and FP64 solves, the
nonlinear loop (row-gather Newton assembly, mu update) in both inner
precisions, Dirichlet and Neumann FP32 solves interleaved on one mesh, and
the batch."""
import numpy as np
import pytest

from femeko_hip import workloads as W
from femeko_hip.batch import Batch
from femeko_hip.device import DeviceMesh, context
from femeko_hip.workloads import dirichlet_values

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_single_mesh_solves_repeat_bit_for_bit():
    pr = W.boxes(40)
    dm = DeviceMesh(pr.mesh, context(0))
    try:
        dm.set_coefficient(pr.mu)
        dm.assemble()
        dm.set_dirichlet(pr.mask, pr.g)
        for prec, rtol in (("f32", 5e-11), ("f64", 1e-10)):
            r1 = dm.solve(prec, rtol=rtol)
            u1 = dm.potential()
            r2 = dm.solve(prec, rtol=rtol)
            u2 = dm.potential()
            assert (r1.iterations, r1.outer) == (r2.iterations, r2.outer), prec
            assert r1.relres == r2.relres, prec
            assert np.array_equal(_bits(u1), _bits(u2)), prec
    finally:
        dm.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_nonlinear_loop_repeats_bit_for_bit(precision):
    pr = W.froehlich(20)
    Ms, a = pr.meta["Ms"], pr.meta["a"]
    dm = DeviceMesh(pr.mesh, context(0))
    try:
        dm.set_dirichlet(pr.mask, pr.g)
        dm.set_inner_precision(precision)
        runs = []
        for _ in range(2):
            out = dm.nonlinear_froehlich(pr.magnet.astype(np.uint8), Ms, a, rtol=1e-10, tol_mu=1e-6, max_outer=20)
            runs.append((out, dm.get_coefficient(), dm.potential()))
        (o1, mu1, u1), (o2, mu2, u2) = runs
        assert (o1["outer"], o1["cg_iterations"]) == (o2["outer"], o2["cg_iterations"])
        assert o1["change"] == o2["change"]
        assert np.array_equal(_bits(mu1), _bits(mu2))
        assert np.array_equal(_bits(u1), _bits(u2))
    finally:
        dm.close()


def test_batch_solve_repeats_bit_for_bit():
    probs = W.batch(count=8, nmin=10, nmax=20)
    gs = [np.stack([np.where(p.mask == 1, dirichlet_values(p.mesh, e), 0.0) for e in np.eye(3)]) for p in probs]
    b = Batch([p.mesh for p in probs], nrhs=3)
    try:
        b.set_coefficient([p.mu for p in probs])
        b.assemble()
        b.set_dirichlet([p.mask for p in probs], gs)
        r1 = b.solve(rtol=1e-10)
        u1 = [b.potential(i).copy() for i in range(len(probs))]
        r2 = b.solve(rtol=1e-10)
        u2 = [b.potential(i) for i in range(len(probs))]
        assert np.array_equal(r1.iterations, r2.iterations)
        for i in range(len(probs)):
            assert np.array_equal(_bits(u1[i]), _bits(u2[i])), i
    finally:
        b.close()


def test_dirichlet_and_neumann_fp32_solves_interleave():
    # the FP32-refined Neumann solve runs the FP32 solver with every node
    # free (an all-zero mask swapped in for the Dirichlet one, D^-1 marked
    # stale, the captured replacement graph keyed on the mask): Dirichlet
    # and Neumann solves alternated on one mesh repeat bit for bit, and the
    # Neumann solution agrees with the FP64 path's
    pr = W.linear(12, mu_r=3.0)
    dm = DeviceMesh(pr.mesh, context(0))
    try:
        dm.set_coefficient(pr.mu)
        dm.assemble()
        dm.set_dirichlet(pr.mask, pr.g)
        dm.set_inner_precision("f32")
        b = np.random.default_rng(5).normal(size=pr.mesh.nv)
        runs = []
        for _ in range(2):
            r = dm.solve("f32", rtol=1e-10)
            u = dm.potential()
            un, mult, its, rel = dm.solve_neumann(b, 1e-10)
            runs.append((r, u, un, mult, its, rel))
        (r1, u1, n1, m1, i1, l1), (r2, u2, n2, m2, i2, l2) = runs
        assert (r1.iterations, r1.outer) == (r2.iterations, r2.outer)
        assert np.array_equal(_bits(u1), _bits(u2))
        assert (i1, l1, m1) == (i2, l2, m2) and l1 < 1e-10
        assert np.array_equal(_bits(n1), _bits(n2))
        dm.set_inner_precision("f64")
        n64, m64, _i, _l = dm.solve_neumann(b, 1e-12)
        assert np.linalg.norm(n1 - n64) / np.linalg.norm(n64) < 1e-8
        assert m1 == pytest.approx(m64, rel=1e-12)
        # and the Dirichlet FP32 solve after the FP64 Neumann one
        r3 = dm.solve("f32", rtol=1e-10)
        assert r3.iterations == r1.iterations
        assert np.array_equal(_bits(dm.potential()), _bits(u1))
    finally:
        dm.close()
