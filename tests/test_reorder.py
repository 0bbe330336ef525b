"""The library's node renumbering (reverse Cuthill-McKee, fk_rcm_order) on
the host: a permutation of the nodes, and as compact as scipy's RCM
(scipy.sparse.csgraph.reverse_cuthill_mckee, the published algorithm of
George & Liu) on a randomly numbered structured mesh and on a Delaunay
mesh, measured by the bandwidth max |i - j| and the mean |i - j| over
neighbouring nodes."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee

from femeko_hip import mesh as M


def _graph(mesh):
    t0 = np.asarray(mesh.t, dtype=np.int64) - 1
    pairs = [(a, b) for a in range(4) for b in range(4) if a != b]
    r = np.concatenate([t0[a] for a, _ in pairs])
    c = np.concatenate([t0[b] for _, b in pairs])
    G = sp.csr_matrix((np.ones(r.size, np.int8), (r, c)), shape=(mesh.nv, mesh.nv))
    G.sum_duplicates()
    return G


def _spread(G, perm):
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    coo = G.tocoo()
    d = np.abs(inv[coo.row] - inv[coo.col])
    return int(d.max()), float(d.mean())


def _delaunay(npts=3000, seed=3):
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    P = rng.random((npts, 3))
    tri = Delaunay(P)
    t = np.asfortranarray((tri.simplices.T + 1).astype(np.int32))
    p = np.asfortranarray(P.T)
    return M.MESH(p=p, t=t, VE=M.element_volumes(p, t))


@pytest.mark.parametrize("kind", ["kuhn_random", "delaunay"])
def test_rcm_order_is_a_compact_permutation(kind):
    if kind == "kuhn_random":
        m = M.kuhn_box_mesh(12, 0.2, seed=5)
        m = M.renumber_nodes(m, np.random.default_rng(5).permutation(m.nv))
    else:
        m = _delaunay()
    perm = M.rcm_order(m)
    assert np.array_equal(np.sort(perm), np.arange(m.nv))
    G = _graph(m)
    ours = _spread(G, perm)
    ref = _spread(G, np.asarray(reverse_cuthill_mckee(G, symmetric_mode=True), dtype=np.int64))
    caller = _spread(G, np.arange(m.nv))
    assert ours[0] <= 1.25 * ref[0] and ours[1] <= 1.25 * ref[1], (ours, ref)
    assert ours[1] < 0.5 * caller[1]


def test_rcm_order_keeps_isolated_nodes_last():
    # a node used by no element is numbered after all others
    m = M.kuhn_box_mesh(3, 0.0, seed=1)
    p = np.asfortranarray(np.hstack([m.p, [[5.0], [5.0], [5.0]]]))
    m2 = M.MESH(p=p, t=m.t, VE=m.VE)
    perm = M.rcm_order(m2)
    assert perm[-1] == m2.nv - 1 and np.array_equal(np.sort(perm), np.arange(m2.nv))
