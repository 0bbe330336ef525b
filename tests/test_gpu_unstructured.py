"""Unstructured meshes: a 3D Delaunay tetrahedralization of random points
(irregular rows of up to ~80 entries, node valence up to 77, many colours),
uploaded as a Femeko MESH. Same bars as the structured parity tests:
pattern bit-exact, values within 1e-12 of sum|terms|, potential within 1e-8
of the oracle's tight solve (GPU rtol 1e-11)."""
import numpy as np
import pytest
from scipy.spatial import Delaunay

from femeko_hip.device import DeviceMesh, context
from femeko_hip.mesh import MESH, MU0, element_volumes, signed_dets

pytestmark = pytest.mark.gpu


def delaunay_mesh(npts=3000, seed=0):
    rng = np.random.default_rng(seed)
    corners = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)], float)
    faces = []
    for ax in range(3):
        for v in (0.0, 1.0):
            q = rng.random((120, 3))
            q[:, ax] = v
            faces.append(q)
    pts = np.vstack([rng.random((npts, 3)), corners] + faces)
    d = Delaunay(pts)
    p = np.asfortranarray(pts.T)
    t = np.asfortranarray((d.simplices.T + 1).astype(np.int32))
    m = MESH(p=p, t=t, VE=element_volumes(p, t))
    hull = np.zeros(m.nv, np.uint8)
    hull[np.unique(d.convex_hull)] = 1
    return m, hull


def test_delaunay_mesh_parity(oracle):
    m, hull = delaunay_mesh()
    assert np.any(signed_dets(m.p, m.t) < 0)  # mixed orientations, as the reference allows
    rng = np.random.default_rng(1)
    mu = MU0 * np.exp(rng.normal(size=m.nt))
    A = oracle.stiffnessMatrix(m, mu)
    dm = DeviceMesh(m, context(0))
    try:
        assert dm.ncolours > 40  # irregular valence: far more colours than the Kuhn grids' ~32
        rowptr, colidx = dm.pattern(base=1)
        assert np.array_equal(rowptr, A.colptr) and np.array_equal(colidx, A.rowval)
        dm.set_coefficient(mu)
        scale = oracle.term_scale(m, mu)
        for method in ("gather", "atomic", "coloured"):
            dm.assemble(method)
            v = dm.values()
            assert np.all(np.abs(v - A.nzval) <= 1e-12 * scale), method
        g = np.where(hull == 1, np.sin(3 * m.p[0]) + m.p[1] * m.p[2], 0.0)
        dm.set_dirichlet(hull, g)
        ref, _it, _rel, s = oracle.pcg_dirichlet(A, hull, g, 1e-13)
        assert s == 0
        # irregular rows: slices mostly fall back to per-row indices
        ns, nsh, _ent = dm.sell_info()
        assert nsh < ns // 2
        for fmt in ("sell64", "sell64-rowidx", "csr"):
            dm.set_format(fmt)
            for prec in ("f64", "f32"):
                rep = dm.solve(prec, rtol=1e-11)
                phi = dm.potential()
                assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) < 1e-8, (fmt, prec)
                assert rep.relres < 1e-11
    finally:
        dm.close()
