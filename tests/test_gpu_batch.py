"""Batched solves (many meshes x 1 to 4 applied fields) against per-system oracle
solves. Bar: every system's potential within 1e-9 relative of the oracle's
direct-accuracy solution (GPU FP32-refined to rtol 1e-10), FP64 true residual
below 1e-10."""
import numpy as np
import pytest

from femeko_hip import workloads as W
from femeko_hip.batch import Batch
from femeko_hip.device import DeviceMesh, context
from femeko_hip.workloads import dirichlet_values

pytestmark = pytest.mark.gpu


def small_batch(count=5, seed=7):
    probs = W.batch(count=count, nmin=4, nmax=12, seed=seed)
    return probs


def _fields(probs):
    return [np.stack([np.where(p.mask == 1, dirichlet_values(p.mesh, e), 0.0) for e in np.eye(3)]) for p in probs]


def _check_oracle(oracle, b, probs, gs, mus=None):
    phis = []
    for i, p in enumerate(probs):
        A = oracle.stiffnessMatrix(p.mesh, p.mu if mus is None else mus[i])
        phi = b.potential(i)
        phis.append(phi)
        for j in range(3):
            ref, _it, _rel, s = oracle.pcg_dirichlet(A, p.mask, gs[i][j], 1e-13)
            assert s == 0
            err = np.linalg.norm(phi[j] - ref) / np.linalg.norm(ref)
            assert err < 1e-9, (i, j, err)
    return phis


@pytest.mark.parametrize("precision,inner,fmt", [("f32", None, "auto"), ("f32-restart", 1e-5, "auto"),
                                                 ("f32", None, "csr")])
def test_batch_matches_oracle(oracle, precision, inner, fmt):
    probs = small_batch()
    meshes = [p.mesh for p in probs]
    b = Batch(meshes, nrhs=3)
    try:
        b.set_format(fmt)
        assert b.solver_format() is None
        b.set_precision(precision)
        b.set_coefficient([p.mu for p in probs])
        b.assemble()
        gs = []
        for p in probs:
            gs.append(np.stack([np.where(p.mask == 1, dirichlet_values(p.mesh, e), 0.0) for e in np.eye(3)]))
        b.set_dirichlet([p.mask for p in probs], gs)
        lanes, block, us = b.autotune()
        assert lanes in (4, 8, 16) and block in (128, 256, 512) and us > 0
        launch = b.launch()
        if fmt == "auto":
            # the stencil SpMM's own block and rows per thread; the update
            # kernels keep 512 threads
            assert launch["spmm_block"] == block and launch["spmm_rows_per_thread"] in (1, 2)
            assert launch["block"] == 512
        else:
            assert launch["lanes"] == lanes and launch["block"] == block
        # Kuhn boxes: every mesh's pattern is a 7-offset stencil
        assert b.solver_format() == ("stencil" if fmt == "auto" else "csr")
        rep = b.solve(rtol=1e-10, inner_rtol=inner)
        assert np.all(rep.relres < 1e-10)
        assert np.all(rep.outer >= 1)
        phis = []
        for i, p in enumerate(probs):
            A = oracle.stiffnessMatrix(p.mesh, p.mu)
            phi = b.potential(i)
            phis.append(phi)
            for j in range(3):
                ref, _it, _rel, s = oracle.pcg_dirichlet(A, p.mask, gs[i][j], 1e-13)
                assert s == 0
                err = np.linalg.norm(phi[j] - ref) / np.linalg.norm(ref)
                assert err < 1e-9, (i, j, err)
        # a second solve with other launch parameters reaches the same
        # solution to the solve tolerance
        b.set_launch(4, 128)
        rep2 = b.solve(rtol=1e-10, inner_rtol=inner)
        assert np.all(rep2.relres < 1e-10)
        for i in range(len(probs)):
            d = np.linalg.norm(b.potential(i) - phis[i]) / np.linalg.norm(phis[i])
            assert d < 1e-9, (i, d)
    finally:
        b.close()


def test_batch_reliable_updates_fewer_iterations_than_restarts():
    # replacing the residual keeps the Krylov space; restarting loses it
    probs = small_batch(count=4, seed=11)
    b = Batch([p.mesh for p in probs], nrhs=3)
    try:
        b.set_coefficient([p.mu for p in probs])
        b.assemble()
        gs = [np.stack([np.where(p.mask == 1, dirichlet_values(p.mesh, e), 0.0) for e in np.eye(3)]) for p in probs]
        b.set_dirichlet([p.mask for p in probs], gs)
        ru = b.solve(rtol=1e-10)
        b.set_precision("f32-restart")
        rs = b.solve(rtol=1e-10)
        assert np.all(ru.relres < 1e-10) and np.all(rs.relres < 1e-10)
        assert ru.iterations.sum() < rs.iterations.sum()
    finally:
        b.close()


def test_batch_stencil_values_follow_reassembly(oracle):
    # a new coefficient and assembly after the first solve must reach the
    # stencil copy of the values
    probs = small_batch(count=3, seed=5)
    b = Batch([p.mesh for p in probs], nrhs=3)
    try:
        b.set_coefficient([p.mu for p in probs])
        b.assemble()
        gs = _fields(probs)
        b.set_dirichlet([p.mask for p in probs], gs)
        assert np.all(b.solve(rtol=1e-10).relres < 1e-10)
        assert b.solver_format() == "stencil"
        mus = [np.where(p.magnet, p.mu * 0.1, p.mu) for p in probs]
        b.set_coefficient(mus)
        b.assemble()
        assert np.all(b.solve(rtol=1e-10).relres < 1e-10)
        _check_oracle(oracle, b, probs, gs, mus)
    finally:
        b.close()


def test_batch_renumbered_mesh_falls_back_to_csr(oracle):
    # one randomly numbered mesh in the batch has no stencil pattern: 'auto'
    # resolves to CSR, 'stencil' is refused at the solve
    from femeko_hip.device import FemekoHIPError

    probs = small_batch(count=3, seed=9)
    rng = np.random.default_rng(3)
    probs[1] = W.renumbered(probs[1], rng.permutation(probs[1].mesh.nv), "random")
    gs = _fields(probs)
    for fmt in ("auto", "stencil"):
        b = Batch([p.mesh for p in probs], nrhs=3)
        try:
            b.set_format(fmt)
            b.set_coefficient([p.mu for p in probs])
            b.assemble()
            b.set_dirichlet([p.mask for p in probs], gs)
            if fmt == "stencil":
                with pytest.raises(FemekoHIPError):
                    b.solve(rtol=1e-10)
                continue
            assert np.all(b.solve(rtol=1e-10).relres < 1e-10)
            assert b.solver_format() == "csr"
            with pytest.raises(FemekoHIPError):
                b.set_format("stencil")
            _check_oracle(oracle, b, probs, gs)
        finally:
            b.close()


def test_batch_stencil_mixed_offset_counts(oracle):
    # two Kuhn boxes (7 offsets) and a two-element mesh (offsets 1, 2, 3):
    # the SpMM runs each mesh's own offset count
    from types import SimpleNamespace

    from femeko_hip.mesh import MESH

    probs = small_batch(count=2, seed=13)
    p = np.asfortranarray([[0, 1, 0, 0, 1], [0, 0, 1, 0, 1], [0, 0, 0, 1, 1]], dtype=float)
    t = np.asfortranarray([[1, 2], [2, 3], [3, 4], [4, 5]], dtype=np.int32)
    tiny = SimpleNamespace(mesh=MESH(p=p, t=t, VE=np.array([1 / 6, 1 / 3])),
                           mu=np.array([2.0e-6, 5.0e-6]), mask=np.array([1, 0, 0, 0, 1], np.uint8))
    probs.insert(1, tiny)
    gs = _fields([probs[0], probs[2]])
    gs.insert(1, np.stack([np.where(tiny.mask == 1, -p[j], 0.0) for j in range(3)]))
    b = Batch([q.mesh for q in probs], nrhs=3)
    try:
        b.set_format("stencil")
        b.set_coefficient([q.mu for q in probs])
        b.assemble()
        b.set_dirichlet([q.mask for q in probs], gs)
        assert np.all(b.solve(rtol=1e-10).relres < 1e-10)
        assert b.solver_format() == "stencil"
        _check_oracle(oracle, b, probs, gs)
    finally:
        b.close()


def test_batch_iterations_close_to_single_mesh_solves():
    # every row of a system must take each step: with several blocks per mesh
    # a block that skipped a step would leave x inconsistent and cost many
    # extra iterations after the replacements. The batch's count stays near
    # the single-mesh FP32 solver's on the same systems.
    probs = W.batch(count=3, nmin=18, nmax=22, seed=31)
    gs = _fields(probs)
    b = Batch([p.mesh for p in probs], nrhs=3)
    try:
        b.set_coefficient([p.mu for p in probs])
        b.assemble()
        b.set_dirichlet([p.mask for p in probs], gs)
        rep = b.solve(rtol=1e-10)
        assert np.all(rep.relres < 1e-10)
        single = []
        for i, p in enumerate(probs):
            dm = DeviceMesh(p.mesh, context(0))
            try:
                dm.set_coefficient(p.mu)
                dm.assemble()
                for j in range(3):
                    dm.set_dirichlet(p.mask, gs[i][j])
                    single.append(dm.solve("f32", rtol=1e-10, inner_rtol=0.9).iterations)
            finally:
                dm.close()
        single = np.array(single)
        its = np.asarray(rep.iterations).reshape(-1)  # system s = mesh * 3 + field
        assert np.all(its <= 1.5 * single + 64), (its, single)
    finally:
        b.close()


@pytest.mark.parametrize("nrhs", [1, 2, 4])
@pytest.mark.parametrize("fmt", ["auto", "csr"])
def test_batch_other_field_counts(oracle, nrhs, fmt):
    # the kernels are instantiated per number of right-hand sides (1..4);
    # every count against the oracle's per-system solves
    probs = small_batch(count=4, seed=11)
    dirs = [np.array(v, float) for v in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1))][:nrhs]
    gs = [np.stack([np.where(p.mask == 1, dirichlet_values(p.mesh, e), 0.0) for e in dirs]) for p in probs]
    b = Batch([p.mesh for p in probs], nrhs=nrhs)
    try:
        b.set_format(fmt)
        b.set_coefficient([p.mu for p in probs])
        b.assemble()
        b.set_dirichlet([p.mask for p in probs], gs)
        rep = b.solve(rtol=1e-10)
        assert rep.relres.shape == (len(probs), nrhs)
        assert np.all(rep.relres < 1e-10)
        assert b.solver_format() == ("stencil" if fmt == "auto" else "csr")
        for i, p in enumerate(probs):
            A = oracle.stiffnessMatrix(p.mesh, p.mu)
            phi = b.potential(i)
            assert phi.shape == (nrhs, p.mesh.nv)
            for j in range(nrhs):
                ref, _it, _rel, s = oracle.pcg_dirichlet(A, p.mask, gs[i][j], 1e-13)
                assert s == 0
                err = np.linalg.norm(phi[j] - ref) / np.linalg.norm(ref)
                assert err < 1e-9, (nrhs, fmt, i, j, err)
    finally:
        b.close()
