"""GPU parity of the mu(|H|) loop with the tabulated law of
Magnetostatics/NewtonRaphson.jl:57-63, 204-207 (cubic splines of the
permeability table, Newton from the second iteration) against the oracle.

Material: gadolinium at 270 K (tests/golden/gd270_bh.npz, made by
tests/golden/make_gd_table.py from the data files the reference ships),
applied field 0.5 T along x as NewtonRaphson.jl:33. Bars: the same number of
outer iterations as the oracle; potential within 1e-8 relative, mu within
1e-6 relative (both loops stop at mu0 max| |H| - |H|_prev | < 1e-8 T with each
inner solve at rtol 1e-10).
"""
import os

import numpy as np
import pytest

from femeko_hip import _lib
from femeko_hip import workloads as W
from femeko_hip._lib import FemekoHIPError
from femeko_hip.device import DeviceMesh, context
from femeko_hip.material import Spline1D, material_permeability

pytestmark = pytest.mark.gpu

GD270 = os.path.join(os.path.dirname(__file__), "golden", "gd270_bh.npz")


def gd_splines():
    tab = np.load(GD270)
    mu, dmu, _nu = material_permeability(tab["H"], tab["B"])
    return Spline1D(tab["H"], mu), Spline1D(tab["H"], dmu)


@pytest.mark.parametrize("inner_precision", ["f64", "f32"])
@pytest.mark.parametrize("n,seed", [(8, 21), (12, 22)])
def test_spline_loop_matches_oracle(oracle, n, seed, inner_precision):
    spl, spl_dmu = gd_splines()
    pr = W.tabulated(n=n, seed=seed)
    ref = oracle.newton_spline(pr.mesh, pr.mask, pr.g, pr.magnet, (spl.t, spl.c), (spl_dmu.t, spl_dmu.c),
                               rtol=1e-10, tol_H=1e-8, max_outer=30)
    dm = DeviceMesh(pr.mesh, context(0))
    try:
        dm.set_dirichlet(pr.mask, pr.g)
        dm.set_inner_precision(inner_precision)
        out = dm.nonlinear_spline(pr.magnet.astype(np.uint8), spl, spl_dmu, rtol=1e-10, tol_H=1e-8, max_outer=30)
        assert out["outer"] == ref["iters"]
        assert out["change"] < 1e-8
        phi = dm.potential()
        assert np.linalg.norm(phi - ref["phi"]) / np.linalg.norm(ref["phi"]) < 1e-8
        mu = dm.get_coefficient()
        np.testing.assert_allclose(mu, ref["mu"], rtol=1e-6)
        # the GPU's mu is the table's spline at the GPU's own |H|
        _H, Hn, _E = dm.field(want_energy=False)
        np.testing.assert_allclose(mu[pr.magnet], spl(Hn[pr.magnet]), rtol=1e-6)
    finally:
        dm.close()


def test_spline_loop_outside_table_is_a_domain_error():
    # 3 T applied: |H| in the magnet at the first step (mu = mu0) is beyond the
    # table's 1.59e6 A/m, which the reference's bc = "error" rejects
    spl, spl_dmu = gd_splines()
    pr = W.tabulated(n=6, seed=23, Bext=3.0)
    dm = DeviceMesh(pr.mesh, context(0))
    try:
        dm.set_dirichlet(pr.mask, pr.g)
        with pytest.raises(FemekoHIPError) as ei:
            dm.nonlinear_spline(pr.magnet.astype(np.uint8), spl, spl_dmu, max_outer=5)
        assert ei.value.status == _lib.FK_ERR_DOMAIN
        # the context still works afterwards
        pr2 = W.tabulated(n=6, seed=23)
        dm.set_dirichlet(pr2.mask, pr2.g)
        out = dm.nonlinear_spline(pr2.magnet.astype(np.uint8), spl, spl_dmu, tol_H=1e-6, max_outer=30)
        assert out["change"] < 1e-6
    finally:
        dm.close()


def test_spline_loop_rejects_bad_tables():
    spl, spl_dmu = gd_splines()
    pr = W.tabulated(n=4, seed=24)
    dm = DeviceMesh(pr.mesh, context(0))
    try:
        dm.set_dirichlet(pr.mask, pr.g)
        region = pr.magnet.astype(np.uint8)
        bad_t = spl.t.copy()
        bad_t[10], bad_t[11] = bad_t[11], bad_t[10]  # decreasing knots
        with pytest.raises(FemekoHIPError) as ei:
            dm.nonlinear_spline(region, (bad_t, spl.c), spl_dmu)
        assert ei.value.status == _lib.FK_ERR_INVALID_ARG
        bad_c = spl.c.copy()
        bad_c[3] = np.nan
        with pytest.raises(FemekoHIPError) as ei:
            dm.nonlinear_spline(region, spl, (spl_dmu.t, bad_c))
        assert ei.value.status == _lib.FK_ERR_NONFINITE
        with pytest.raises(FemekoHIPError):
            dm.nonlinear_spline(region, (spl.t[:6], spl.c[:2]), spl_dmu)
    finally:
        dm.close()
