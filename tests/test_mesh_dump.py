"""Reading Femeko's text dump of a mesh (src/mesh.jl:232-242 through
save2file, src/geometry.jl:236-243: one matrix row per line, entries joined
by " , "; vectors one entry per line).

tests/golden/mesh_dump_2tet holds a dump of a two-element mesh written in
that layout by hand (Julia prints Float64 in shortest round-trip form, as
below): node 5 = (1, 1, 1) on the far side of the face (2, 3, 4) of the
reference tetrahedron, so the volumes are 1/6 and |det| / 6 = 2/6, the bottom
face (1, 3, 2) with tag 1 has area 1/2 and outward normal -z."""
import os

import numpy as np
import pytest

from femeko_hip import mesh as M

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mesh_dump_2tet")


def test_golden_dump_reads_as_written():
    m = M.load_mesh_dump(GOLDEN)
    np.testing.assert_array_equal(m.p, [[0, 1, 0, 0, 1], [0, 0, 1, 0, 1], [0, 0, 0, 1, 1]])
    assert m.p.dtype == np.float64 and m.p.flags.f_contiguous
    np.testing.assert_array_equal(m.t, [[1, 2], [2, 3], [3, 4], [4, 5]])
    assert m.t.dtype == np.int32 and (m.nv, m.nt) == (5, 2)
    # the dumped volumes are the closed-form ones and elementVolume's
    assert m.VE.tolist() == [1 / 6, 2 / 6]
    np.testing.assert_allclose(M.element_volumes(m.p, m.t), m.VE, rtol=1e-15)
    assert m.InsideElements.tolist() == [2] and m.InsideNodes.tolist() == [2, 3, 4, 5]
    np.testing.assert_array_equal(m.surfaceT, [[1], [3], [2], [1]])
    assert m.AE.tolist() == [0.5] and m.normal[:, 0].tolist() == [0.0, 0.0, -1.0]
    assert m.shell_id == (1,)


def test_dump_round_trip_is_bit_exact(tmp_path):
    m = M.boundary_faces(M.kuhn_box_mesh(4, jitter=0.2, seed=9))
    M.set_inside(m, np.arange(m.nt) % 3 == 0)
    M.save_mesh_dump(m, tmp_path)
    r = M.load_mesh_dump(tmp_path)
    for name in ("p", "VE", "AE", "normal"):
        a, b = np.asarray(getattr(m, name), np.float64), getattr(r, name)
        assert np.array_equal(a.view(np.uint64), np.ascontiguousarray(b).view(np.uint64).reshape(a.shape)), name
    for name in ("t", "surfaceT", "InsideElements", "InsideNodes"):
        np.testing.assert_array_equal(getattr(m, name), getattr(r, name))
    assert r.shell_id == (1,)


def test_dump_without_volumes_recomputes_them(tmp_path):
    m = M.kuhn_box_mesh(2, jitter=0.1, seed=3)
    M.save_mesh_dump(m, tmp_path)
    os.remove(tmp_path / "VE.txt")
    r = M.load_mesh_dump(tmp_path)
    np.testing.assert_allclose(r.VE, M.element_volumes(m.p, m.t), rtol=0, atol=0)


def test_julia_float_spellings(tmp_path):
    # Julia writes 1.0e-5 and -0.0 where Python writes 1e-05 and -0.0
    (tmp_path / "AE.txt").write_text("1.0e-5\n-0.0\n2.5\n")
    v = M._read_vector(str(tmp_path / "AE.txt"), np.float64)
    assert v.tolist() == [1e-5, -0.0, 2.5] and np.signbit(v[1])


@pytest.mark.parametrize("t_text,err", [
    ("1 , 2\n2 , 3\n3 , 4\n", "4 node rows"),               # a row missing
    ("1\n2\n3\n4\n5\n6\n7\n8\n9\n10\n", "second-order"),    # P2 (10 nodes)
    ("1 , 2\n2 , 3\n3 , 4\n4 , 6\n", "outside 1..nv"),       # node 6 of 5
    ("1 , 2\n2 , 3\n3\n4 , 5\n", "equal length"),             # ragged rows
])
def test_malformed_dumps_are_rejected(tmp_path, t_text, err):
    for name in ("p.txt", "VE.txt"):
        (tmp_path / name).write_text(open(os.path.join(GOLDEN, name)).read())
    (tmp_path / "t.txt").write_text(t_text)
    with pytest.raises(ValueError, match=err):
        M.load_mesh_dump(tmp_path)


@pytest.mark.gpu
def test_dumped_mesh_assembles_on_the_gpu_as_the_oracle(oracle, tmp_path):
    # a mesh taken through the text dump (save, then load as a Femeko user's
    # saved mesh would be) assembles on the GPU to the oracle's matrix
    from femeko_hip.device import DeviceMesh, context
    m0 = M.boundary_faces(M.kuhn_box_mesh(6, jitter=0.25, seed=21))
    M.save_mesh_dump(m0, tmp_path)
    m = M.load_mesh_dump(tmp_path)
    mu = 1.0 + np.arange(m.nt) % 5
    A = oracle.stiffnessMatrix(m, mu)
    dm = DeviceMesh(m, context(0))
    try:
        dm.set_coefficient(mu)
        dm.assemble()
        rowptr, colidx = dm.pattern(base=1)
        assert np.array_equal(rowptr, A.colptr) and np.array_equal(colidx, A.rowval)
        bound = oracle.term_scale(m, mu)
        assert np.all(np.abs(dm.values() - A.nzval) <= 1e-12 * bound)
    finally:
        dm.close()


def _with_interface(n=4):
    """A Kuhn box whose surfaceT also holds the interface between an inner
    body (elements with centroid in [-0.5,0.5]^3, tag 2 triangles) and the
    air around it, as Gmsh meshes every surface of a multi-body model."""
    m = M.boundary_faces(M.kuhn_box_mesh(n, jitter=0.0, seed=1))
    body = M.elements_in_box(m, (-0.5,) * 3, (0.5,) * 3)
    t0 = (m.t - 1).T.astype(np.int64)
    local = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
    faces = np.sort(t0[:, local].reshape(-1, 3), axis=1)
    owner = np.repeat(np.arange(m.nt), 4)
    _, inv, cnt = np.unique(faces, axis=0, return_inverse=True, return_counts=True)
    inv = inv.ravel()
    iface = []
    for key in np.nonzero(cnt == 2)[0]:
        e = owner[inv == key]
        if body[e[0]] != body[e[1]]:
            iface.append(faces[np.nonzero(inv == key)[0][0]])
    iface = np.array(iface, dtype=np.int64)
    P = m.p
    nrm = np.cross((P[:, iface[:, 1]] - P[:, iface[:, 0]]).T, (P[:, iface[:, 2]] - P[:, iface[:, 0]]).T).T
    area = np.linalg.norm(nrm, axis=0) / 2
    nrm /= 2 * area
    sT = np.vstack([(iface.T + 1), np.full((1, iface.shape[0]), 2)]).astype(np.int32)
    outer = (np.asarray(m.surfaceT).copy(), np.asarray(m.normal).copy(), np.asarray(m.AE).copy())
    m.surfaceT = np.asfortranarray(np.hstack([outer[0], sT]))
    m.normal = np.asfortranarray(np.hstack([outer[1], nrm]))
    m.AE = np.concatenate([outer[2], area])
    return m, outer, iface.shape[0]


def test_dump_shell_id_keeps_only_outer_surfaces(tmp_path, oracle):
    """surfaceT with an internal interface tag: the inferred shell_id is the
    outer boundary's tag only (src/geometry.jl:225-230), so BoundaryIntegral
    (src/FEM3D.jl:284-297) over it equals the load of the outer faces alone."""
    m, outer, n_iface = _with_interface()
    assert n_iface > 0
    M.save_mesh_dump(m, tmp_path)
    r = M.load_mesh_dump(tmp_path)
    assert r.shell_id == (1,)
    F = np.array([1.0, -2.0, 0.5])
    got = oracle.BoundaryIntegral(r, F, r.shell_id)
    ref_mesh = M.MESH(p=m.p, t=m.t, VE=m.VE)
    ref_mesh.surfaceT, ref_mesh.normal, ref_mesh.AE = (np.asfortranarray(outer[0]), np.asfortranarray(outer[1]),
                                                       outer[2])
    want = oracle.BoundaryIntegral(ref_mesh, F, (1,))
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)
    # all tags would add the interface's flux on the body's surface nodes
    both = oracle.BoundaryIntegral(r, F, (1, 2))
    assert np.abs(both - want).max() > 1e-3
