"""Tabulated magnetic materials on the host: the permeability table of
src/magneticProperties.jl:163-188 and the cubic interpolating spline that
Magnetostatics/NewtonRaphson.jl:57-63 builds from it with Dierckx.jl.

Spline1D fits with scipy's FITPACK binding (splrep, the routine curfit that
Dierckx.jl wraps, k = 3, s = 0) and keeps the FITPACK form (knots t,
coefficients c) that the GPU loop evaluates (DeviceMesh.nonlinear_spline).
"""
from __future__ import annotations

import numpy as np

MU0 = np.pi * 4e-7


def gradient(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """src/FEM.jl:66-90: one-sided differences at the ends, central inside."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if x.size != y.size:
        raise ValueError("Input x and y must have the same length")
    if x.size < 2:
        raise ValueError("Input data must have more than 1 element")
    z = np.empty_like(y)
    with np.errstate(divide="ignore", invalid="ignore"):
        z[0] = (y[1] - y[0]) / (x[1] - x[0])
        z[-1] = (y[-1] - y[-2]) / (x[-1] - x[-2])
        z[1:-1] = (y[2:] - y[:-2]) / (x[2:] - x[:-2])
    return z


def material_permeability(HofM: np.ndarray, B: np.ndarray):
    """src/magneticProperties.jl:163-188 (materialPermeability): returns
    (mu, dmu, nu). mu = B/H with its non-finite entries replaced by the
    largest entry (after zeroing them), dmu = gradient(H, mu) with its
    non-finite entries replaced by the smallest, nu = mu0 H/B (relative
    reluctance) with non-finite entries 0."""
    H = np.asarray(HofM, dtype=np.float64)
    Bv = np.asarray(B, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        mu = Bv / H
        nu = MU0 * (H / Bv)
    idx = ~np.isfinite(mu)
    mu[idx] = 0.0
    mu[idx] = np.max(mu)
    dmu = gradient(H, mu)
    idx = ~np.isfinite(dmu)
    dmu[idx] = 0.0
    dmu[idx] = np.min(dmu)
    nu[~np.isfinite(nu)] = 0.0
    return mu, dmu, nu


class Spline1D:
    """Interpolating cubic spline of (x, y), as Dierckx.jl's
    Spline1D(x, y; k=3, s=0.0, bc="error"): knots t (n) and coefficients c
    (n - 4). Calling it evaluates on the host and raises outside [t[3],
    t[n-4]] (bc = "error"); the GPU loop evaluates the same (t, c)."""

    def __init__(self, x, y):
        from scipy.interpolate import splrep

        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        if np.any(np.diff(x) <= 0.0):
            raise ValueError("x must be strictly increasing")
        t, c, k = splrep(x, y, k=3, s=0)
        self.t = np.ascontiguousarray(t)
        self.c = np.ascontiguousarray(c[:t.size - 4])
        self.k = k

    def __call__(self, x):
        from scipy.interpolate import splev

        x = np.asarray(x, dtype=np.float64)
        if np.any(~(x >= self.t[3]) | ~(x <= self.t[-4])):
            raise ValueError("Spline1D: argument outside the table range")
        c = np.concatenate([self.c, np.zeros(4)])
        return splev(x, (self.t, c, 3))
