"""Bitwise reproducibility of the solver paths. Every reduction sums its
block partials in one fixed order (each block re-reduces the previous
kernel's partials, so all blocks hold the same alpha and beta) and no kernel
adds floating-point values atomically, so two runs on the same inputs must
give the same bits: the single-mesh FP32-refined and FP64 solves, the
nonlinear loop (row-gather Newton assembly, mu update) and the batch."""
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


def test_nonlinear_loop_repeats_bit_for_bit():
    pr = W.froehlich(20)
    Ms, a = pr.meta["Ms"], pr.meta["a"]
    dm = DeviceMesh(pr.mesh, context(0))
    try:
        dm.set_dirichlet(pr.mask, pr.g)
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
