"""GPU parity of the demagnetizing field of Micromagnetics/LL.jl:162-209
(fk_demag_field) against the oracle's magnetostatic_field (direct bordered
solve), with FP64 and with FP32-refined Neumann solves. Bars: nodal field within 1e-9 of max|Hd| at rtol 1e-12, the same NaN
pattern (0/0 outside the magnet), the exact patch test H = -M for uniform M
everywhere, and repeated calls on the resident matrix (one per time step)."""
import numpy as np
import pytest

from femeko_hip import workloads as W
from femeko_hip.device import DeviceMesh, context

pytestmark = pytest.mark.gpu


def _device_A(mesh, precision="f64"):
    dm = DeviceMesh(mesh, context(0))
    dm.set_coefficient(np.ones(mesh.nt))
    dm.assemble()
    dm.set_inner_precision(precision)
    return dm


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_demag_matches_oracle_over_steps(oracle, precision):
    pr = W.linear(10)
    m = pr.mesh
    inside = pr.magnet
    vol = oracle.node_volumes(m, inside)
    dm = _device_A(m, precision)
    try:
        rng = np.random.default_rng(17)
        for step in range(3):
            M = rng.normal(size=(3, m.nv))
            ref, _u, _rhs = oracle.magnetostatic_field(m, inside, M, vol)
            Hd, its, rel = dm.demag_field(inside.astype(np.uint8), M, vol, rtol=1e-12)
            assert rel < 1e-12 and its > 0
            nan_ref = np.isnan(ref)
            assert np.array_equal(np.isnan(Hd), nan_ref), step
            scale = np.abs(ref[~nan_ref]).max()
            assert np.max(np.abs(Hd[~nan_ref] - ref[~nan_ref])) < 1e-9 * scale, step
        # the reference cg's default tolerance sqrt(eps): same field to ~1e-7
        Hd2, _its, rel2 = dm.demag_field(inside.astype(np.uint8), M, vol)
        assert rel2 < 1.5e-8
        ok = ~nan_ref
        assert np.max(np.abs(Hd2[ok] - ref[ok])) < 1e-6 * scale
    finally:
        dm.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_demag_uniform_magnetization_patch_test(oracle, precision):
    pr = W.linear(8)
    m = pr.mesh
    inside = np.ones(m.nt, np.uint8)
    Mvec = np.array([0.3, -1.2, 0.7])
    M = np.repeat(Mvec[:, None], m.nv, axis=1)
    vol = oracle.node_volumes(m, inside.astype(bool))
    dm = _device_A(m, precision)
    try:
        Hd, _its, _rel = dm.demag_field(inside, M, vol, rtol=1e-13)
        np.testing.assert_allclose(Hd, -M, rtol=0, atol=1e-9)
    finally:
        dm.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_demag_warm_start_same_field_fewer_iterations(oracle, precision):
    # slowly rotating magnetization, one call per step: the start from the
    # previous potential reaches the same field in fewer iterations
    pr = W.linear(10)
    m = pr.mesh
    inside = pr.magnet
    vol = oracle.node_volumes(m, inside)
    innodes = vol > 0
    dm = _device_A(m, precision)
    try:
        cold, warm = [], []
        for k in range(4):
            th = 0.05 * k
            M = np.zeros((3, m.nv))
            M[0, innodes], M[1, innodes] = np.cos(th), np.sin(th)
            ref, _u, _rhs = oracle.magnetostatic_field(m, inside, M, vol)
            ok = ~np.isnan(ref)
            scale = np.abs(ref[ok]).max()
            Hc, itc, _ = dm.demag_field(inside.astype(np.uint8), M, vol, rtol=1e-12)
            Hw, itw, relw = dm.demag_field(inside.astype(np.uint8), M, vol, rtol=1e-12, warm=True)
            assert relw < 1e-12
            for H in (Hc, Hw):
                assert np.array_equal(np.isnan(H), ~ok)
                assert np.max(np.abs(H[ok] - ref[ok])) < 1e-9 * scale
            cold.append(itc)
            warm.append(itw)
        # the warm call of each step starts from that step's own cold solution
        assert max(warm) < min(cold)
    finally:
        dm.close()


def test_demag_rejects_bad_input_and_keeps_state(oracle):
    # fk_demag_field checks its inputs (on the device, after the upload)
    # before anything it keeps changes: a non-finite M is FK_ERR_NONFINITE, a
    # negative or non-finite nodal volume FK_ERR_INVALID_ARG; the warm start
    # of the next good call then still starts from the last good potential
    from femeko_hip import _lib
    from femeko_hip._lib import FemekoHIPError
    pr = W.linear(8)
    m = pr.mesh
    inside = pr.magnet
    vol = oracle.node_volumes(m, inside)
    M = np.zeros((3, m.nv))
    M[2, vol > 0] = 1.0
    dm = _device_A(m)
    try:
        Hd0, _its, _rel = dm.demag_field(inside.astype(np.uint8), M, vol, rtol=1e-12)
        bad = M.copy()
        bad[1, 5] = np.nan
        with pytest.raises(FemekoHIPError) as e:
            dm.demag_field(inside.astype(np.uint8), bad, vol)
        assert e.value.status == _lib.FK_ERR_NONFINITE
        bad[1, 5] = np.inf
        with pytest.raises(FemekoHIPError) as e:
            dm.demag_field(inside.astype(np.uint8), bad, vol)
        assert e.value.status == _lib.FK_ERR_NONFINITE
        for v in (-1.0, np.nan):
            vb = vol.copy()
            vb[3] = v
            with pytest.raises(FemekoHIPError) as e:
                dm.demag_field(inside.astype(np.uint8), M, vb)
            assert e.value.status == _lib.FK_ERR_INVALID_ARG
        Hd1, its1, _rel = dm.demag_field(inside.astype(np.uint8), M, vol, rtol=1e-12, warm=True)
        assert its1 <= 2  # started from the potential of the same M
        ok = ~np.isnan(Hd0)
        np.testing.assert_allclose(Hd1[ok], Hd0[ok], rtol=0, atol=1e-9 * np.abs(Hd0[ok]).max())
    finally:
        dm.close()
