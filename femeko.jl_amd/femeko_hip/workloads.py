"""Seeded synthetic magnetostatics problems on Kuhn box meshes.

Each builder returns a ``Problem``: the mesh, per-element permeability mu
(= mu0 * mu_r, computed once on the host so every consumer sees identical
inputs), the Dirichlet node mask and values g = -H0 . x on the outer faces of
[-1, 1]^3, and workload metadata. Random draws use the counter-based
splitmix64 stream of ``mesh.uniform01`` with a per-problem seed and disjoint
counter ranges, so problems are reproducible bit for bit.

Problems (all with Dirichlet phi = -H0 . x on the box faces):
  linear(n=20)      mu_r = 1000 for centroids in [-0.25, 0.25]^3, else 1;
                    H0 = (1, 0, 0); jitter 0.2 h.
  froehlich(n=100)  same magnet box with the Froehlich law
                    mu_r(H) = 1 + Ms/(a + |H|), Ms = 1.2e6, a = 500;
                    |H0| = 1e4 along a seeded unit vector; jitter 0.2 h.
  boxes(n=160)      8 seeded axis-aligned boxes, log-uniform mu_r in [1, 1e3]
                    per element inside, 1 outside; H0 = (1, 0, 0); jitter 0.15 h.
  batch(64)         64 meshes with n ~ U{10..60}, one magnet box each with
                    mu_r ~ U[10, 1e3], H0 = x, y, z (three right-hand sides).
  adjoint(n=80)     rho_e ~ U[0, 1], mu_r = 1 + 999 rho^3; target box
                    [-0.1, 0.1]^3, H* = (2, 0, 0), H0 = (1, 0, 0); jitter 0.2 h.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .mesh import MESH, MU0, boundary_node_mask, centroids, elements_in_box, kuhn_box_mesh, set_inside, uniform01

# counter ranges of the random stream (node jitter uses 3*node+axis < 2^40)
_C_PROBLEM = 1 << 48


@dataclass
class Problem:
    name: str
    mesh: MESH
    mu: np.ndarray                 # (nt,) float64, mu0 * mu_r
    mask: np.ndarray               # (nv,) uint8, 1 = Dirichlet node
    g: np.ndarray                  # (nv,) float64, Dirichlet values (0 on free nodes)
    H0: np.ndarray                 # (3,) applied field
    magnet: np.ndarray             # (nt,) bool
    meta: dict = field(default_factory=dict)


def _draw(seed: int, k0: int, count: int) -> np.ndarray:
    return uniform01(seed, np.arange(_C_PROBLEM + k0, _C_PROBLEM + k0 + count, dtype=np.uint64))


def dirichlet_values(mesh: MESH, H0) -> np.ndarray:
    """g = -H0 . x (the scalar potential of a uniform applied field)."""
    H0 = np.asarray(H0, dtype=np.float64)
    return -(H0[0] * mesh.p[0] + H0[1] * mesh.p[1] + H0[2] * mesh.p[2])


def _finish(name, mesh, mu, H0, magnet, meta):
    mask = boundary_node_mask(mesh)
    g = np.where(mask.astype(bool), dirichlet_values(mesh, H0), 0.0)
    set_inside(mesh, magnet)
    return Problem(name, mesh, np.ascontiguousarray(mu), mask, g, np.asarray(H0, float), magnet, meta)


def renumbered(pr: Problem, perm: np.ndarray, tag: str) -> Problem:
    """The same problem with node k of the new numbering = node perm[k] of
    the old one (elements keep their order): a general, non-structured
    numbering of the same mesh."""
    perm = np.asarray(perm, dtype=np.int64)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    m = pr.mesh
    p = np.asfortranarray(m.p[:, perm])
    t = np.asfortranarray((inv[np.asarray(m.t, dtype=np.int64) - 1] + 1).astype(np.int32))
    mesh = MESH(p=p, t=t, VE=m.VE)
    set_inside(mesh, pr.magnet)
    meta = dict(pr.meta)
    meta["renumbered"] = tag
    return Problem(f"{pr.name}-{tag}", mesh, pr.mu, np.ascontiguousarray(pr.mask[perm]),
                   np.ascontiguousarray(pr.g[perm]), pr.H0, pr.magnet, meta)


def linear(n: int = 20, jitter: float = 0.2, seed: int = 20, mu_r: float = 1000.0,
           box=0.25, H0=(1.0, 0.0, 0.0)) -> Problem:
    mesh = kuhn_box_mesh(n, jitter, seed)
    magnet = elements_in_box(mesh, (-box,) * 3, (box,) * 3)
    mu = np.where(magnet, MU0 * mu_r, MU0)
    return _finish(f"linear{n}", mesh, mu, H0, magnet, {"n": n, "jitter": jitter, "seed": seed})


def random_unit(seed: int, k0: int = 0) -> np.ndarray:
    u = _draw(seed, k0, 2)
    z = 2.0 * u[0] - 1.0
    ph = 2.0 * np.pi * u[1]
    s = np.sqrt(max(0.0, 1.0 - z * z))
    return np.array([s * np.cos(ph), s * np.sin(ph), z])


def froehlich(n: int = 100, jitter: float = 0.2, seed: int = 100, Ms: float = 1.2e6,
              a: float = 500.0, H0mag: float = 1e4, box=0.25) -> Problem:
    mesh = kuhn_box_mesh(n, jitter, seed)
    magnet = elements_in_box(mesh, (-box,) * 3, (box,) * 3)
    H0 = H0mag * random_unit(seed)
    mu = np.full(mesh.nt, MU0)
    return _finish(f"froehlich{n}", mesh, mu, H0, magnet,
                   {"n": n, "jitter": jitter, "seed": seed, "Ms": Ms, "a": a})


def tabulated(n: int = 100, jitter: float = 0.2, seed: int = 101, Bext: float = 0.5, box=0.25) -> Problem:
    """Cube magnet [-box, box]^3 in the air box under the applied field
    Bext/mu0 along x (Magnetostatics/NewtonRaphson.jl:33-35: Bext = 0.5 T),
    for the loop with a tabulated mu(|H|) law."""
    mesh = kuhn_box_mesh(n, jitter, seed)
    magnet = elements_in_box(mesh, (-box,) * 3, (box,) * 3)
    mu = np.full(mesh.nt, MU0)
    return _finish(f"tabulated{n}", mesh, mu, (Bext / MU0, 0.0, 0.0), magnet,
                   {"n": n, "jitter": jitter, "seed": seed, "Bext": Bext})


def random_boxes(seed: int, count: int, k0: int = 16):
    """count boxes: centre U(-0.7, 0.7)^3, edge lengths U(0.1, 0.6)^3."""
    u = _draw(seed, k0, 6 * count).reshape(count, 6)
    c = -0.7 + 1.4 * u[:, :3]
    s = 0.1 + 0.5 * u[:, 3:]
    return c - s / 2, c + s / 2


def boxes(n: int = 160, jitter: float = 0.15, seed: int = 160, nboxes: int = 8,
          H0=(1.0, 0.0, 0.0)) -> Problem:
    mesh = kuhn_box_mesh(n, jitter, seed)
    cen = centroids(mesh)
    lo, hi = random_boxes(seed, nboxes)
    magnet = np.zeros(mesh.nt, dtype=bool)
    for b in range(nboxes):
        magnet |= elements_in_box(mesh, lo[b], hi[b], cen)
    del cen
    # log-uniform mu_r in [1, 1e3] per element inside the boxes
    u = uniform01(seed, np.arange(1 << 50, (1 << 50) + mesh.nt, dtype=np.uint64))
    mu = np.where(magnet, MU0 * 10.0 ** (3.0 * u), MU0)
    return _finish(f"boxes{n}", mesh, mu, H0, magnet,
                   {"n": n, "jitter": jitter, "seed": seed, "boxes": (lo.tolist(), hi.tolist())})


def batch(count: int = 64, nmin: int = 10, nmax: int = 60, jitter: float = 0.2, seed: int = 64):
    """List of Problems; each carries meta['H0s'] = identity rows (x, y, z)."""
    u = _draw(seed, 0, 4 * count).reshape(count, 4)
    sizes = nmin + np.minimum((u[:, 0] * (nmax - nmin + 1)).astype(int), nmax - nmin)
    out = []
    for i in range(count):
        n = int(sizes[i])
        s = seed * 1000 + i
        mesh = kuhn_box_mesh(n, jitter, s)
        lo, hi = random_boxes(s, 1)
        magnet = elements_in_box(mesh, lo[0], hi[0])
        mur = 10.0 + (1e3 - 10.0) * u[i, 1]
        mu = np.where(magnet, MU0 * mur, MU0)
        pr = _finish(f"batch{i}_n{n}", mesh, mu, (1.0, 0.0, 0.0), magnet,
                     {"n": n, "jitter": jitter, "seed": s, "mu_r": mur})
        pr.meta["H0s"] = np.eye(3)
        out.append(pr)
    return out


def adjoint(n: int = 80, jitter: float = 0.2, seed: int = 80, H0=(1.0, 0.0, 0.0),
            Hstar=(2.0, 0.0, 0.0), box=0.1) -> Problem:
    mesh = kuhn_box_mesh(n, jitter, seed)
    rho = uniform01(seed, np.arange(1 << 50, (1 << 50) + mesh.nt, dtype=np.uint64))
    target = elements_in_box(mesh, (-box,) * 3, (box,) * 3)
    mu = MU0 * (1.0 + 999.0 * rho ** 3)
    pr = _finish(f"adjoint{n}", mesh, mu, H0, target,
                 {"n": n, "jitter": jitter, "seed": seed, "Hstar": np.asarray(Hstar, float)})
    pr.meta["rho"] = rho
    return pr
