"""GPU adjoint gradient against the oracle adjoint and central differences.

Bars: dJ/drho within 1e-7 relative of the oracle's FP64 adjoint (both solves
tight: GPU rtol 1e-12 in FP64 or FP32-refined, oracle 1e-13); J within 1e-10;
central differences
(step 1e-6) of the GPU objective within 2e-4 relative on seeded elements.
"""
import numpy as np
import pytest

from femeko_hip import adjoint as ADJ
from femeko_hip import fem3d
from femeko_hip import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("n,seed", [(8, 31), (20, 20)])
def test_adjoint_matches_oracle_and_fd(oracle, n, seed, precision):
    """n = 20, seed 20 is the config-0 mesh (Kuhn n=20, jitter 0.2h, seed 20)
    with the adjoint workload's design rho and target."""
    pr = W.adjoint(n=n, seed=seed)
    m = pr.mesh
    rho = pr.meta["rho"]
    Hs = pr.meta["Hstar"]
    target = pr.magnet.astype(np.uint8)
    J_ref, grad_ref, _phi, _lam = oracle.adjoint_gradient(m, pr.mask, pr.g, rho, pr.magnet, Hs)
    J, pullback = ADJ.rrule(ADJ.fieldObjective, m, rho, target, Hs, pr.mask, pr.g, precision=precision)
    _none, grad = pullback(1.0)
    assert J == pytest.approx(J_ref, rel=1e-10)
    scale = np.abs(grad_ref).max()
    assert np.abs(grad - grad_ref).max() <= 1e-7 * scale
    _n, g2 = pullback(2.5)
    np.testing.assert_allclose(g2, 2.5 * grad)
    rng = np.random.default_rng(seed)
    # largest-gradient elements plus random ones
    idx = np.concatenate([np.argsort(-np.abs(grad))[:4], rng.choice(m.nt, 4, replace=False)])
    h = 1e-6
    for e in idx:
        rp = rho.copy()
        rp[e] += h
        rm = rho.copy()
        rm[e] -= h
        fd = (ADJ.fieldObjective(m, rp, target, Hs, pr.mask, pr.g)
              - ADJ.fieldObjective(m, rm, target, Hs, pr.mask, pr.g)) / (2 * h)
        assert fd == pytest.approx(grad[e], rel=2e-4, abs=1e-6 * scale)
    fem3d.invalidate(m)
