"""Femeko ``MESH`` data model and the synthetic structured tetrahedral mesher.

``MESH`` mirrors the fields of Femeko's mesh struct that the P1
magnetostatics path reads (src/mesh.jl:9-59): ``p`` is 3 x nv column-major
Float64 (x, y, z of each node stored together), ``t`` is 4 x nt Int32 with
1-based node numbers, ``VE`` the element volumes filled by ``elementVolume``
(src/mesh.jl:216-219, 484-505), ``InsideElements`` / ``InsideNodes`` 1-based.

Femeko builds meshes with Gmsh (src/mesh.jl:62-245). The drop-in path does not
depend on where a mesh comes from; workloads here use ``kuhn_box_mesh``, a
structured n^3 grid of the cube [lo, hi]^3 with every cell split into six
tetrahedra around its v000 -> v111 diagonal (the Kuhn / Freudenthal split,
which is conforming across cells). Interior nodes are jittered per coordinate
by U(-j h, j h) drawn from splitmix64, so the same seed gives bit-identical
meshes on every machine.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

MU0 = np.pi * 4e-7  # vacuum permeability, Magnetostatics/magneticField.jl:38

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, counters: np.ndarray) -> np.ndarray:
    """k-th output of a splitmix64 generator started at ``seed``, for each k.

    Counter-based so any subset of the stream is computed independently:
    z = seed + (k + 1) * 0x9E3779B97F4A7C15, then the splitmix64 finaliser.
    """
    k = np.asarray(counters, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (k + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, counters: np.ndarray) -> np.ndarray:
    """Doubles in [0, 1): top 53 bits of splitmix64 times 2^-53."""
    return (splitmix64(seed, counters) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclass
class MESH:
    """Fields of Femeko's MESH (src/mesh.jl:9-59) used on the P1 path."""

    p: np.ndarray                     # (3, nv) float64, Fortran order (column-major)
    t: np.ndarray                     # (4, nt) int32, 1-based, Fortran order
    VE: np.ndarray                    # (nt,) float64
    InsideElements: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    InsideNodes: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    surfaceT: np.ndarray | None = None  # (4, ns) int32: 3 nodes + boundary tag, 1-based
    normal: np.ndarray | None = None    # (3, ns) float64 outward unit normals
    AE: np.ndarray | None = None        # (ns,) float64 triangle areas
    shell_id: tuple = ()
    order: int = 1
    n_cells: int = 0                    # cells per side for structured meshes (0 = unstructured)

    @property
    def nv(self) -> int:
        return int(self.p.shape[1])

    @property
    def nt(self) -> int:
        return int(self.t.shape[1])

    @property
    def ns(self) -> int:
        return 0 if self.surfaceT is None else int(self.surfaceT.shape[1])

    @property
    def nInside(self) -> int:
        return int(self.InsideElements.size)


# Axis permutations of the six Kuhn tetrahedra and their parity. Tet = v000,
# v000+e_a, v000+e_a+e_b, v111; odd permutations swap the last two nodes so every
# element is positively oriented.
_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
_ODD = (False, True, True, False, False, True)


def _kuhn_local_offsets() -> np.ndarray:
    """(6, 4, 3) integer corner offsets of the six tets of a unit cell."""
    out = np.zeros((6, 4, 3), dtype=np.int64)
    for m, (a, b, _c) in enumerate(_PERMS):
        v1 = np.zeros(3, np.int64)
        v1[a] = 1
        v2 = v1.copy()
        v2[b] = 1
        v3 = np.ones(3, np.int64)
        nodes = [np.zeros(3, np.int64), v1, v2, v3]
        if _ODD[m]:
            nodes[2], nodes[3] = nodes[3], nodes[2]
        out[m] = np.stack(nodes)
    return out


def element_volumes(p: np.ndarray, t: np.ndarray) -> np.ndarray:
    """elementVolume for every element, same operation order (src/mesh.jl:484-505)."""
    t0 = t - 1
    A = p[:, t0[0]]
    B = p[:, t0[1]]
    C = p[:, t0[2]]
    D = p[:, t0[3]]
    AB = B - A
    AC = C - A
    AD = D - A
    c0 = AC[1] * AD[2] - AC[2] * AD[1]
    c1 = AC[2] * AD[0] - AC[0] * AD[2]
    c2 = AC[0] * AD[1] - AC[1] * AD[0]
    triple = AB[0] * c0 + AB[1] * c1 + AB[2] * c2
    return np.abs(triple) / 6.0


def signed_dets(p: np.ndarray, t: np.ndarray) -> np.ndarray:
    t0 = t - 1
    A = p[:, t0[0]]
    AB = p[:, t0[1]] - A
    AC = p[:, t0[2]] - A
    AD = p[:, t0[3]] - A
    return (AB[0] * (AC[1] * AD[2] - AC[2] * AD[1])
            + AB[1] * (AC[2] * AD[0] - AC[0] * AD[2])
            + AB[2] * (AC[0] * AD[1] - AC[1] * AD[0]))


def kuhn_box_mesh(n: int, jitter: float = 0.0, seed: int = 0,
                  lo: float = -1.0, hi: float = 1.0) -> MESH:
    """Structured Kuhn-split tetrahedral mesh of [lo, hi]^3 with n cells per side.

    nt = 6 n^3 elements, nv = (n+1)^3 nodes. Node (i, j, k) has 1-based number
    1 + i + (n+1) (j + (n+1) k); cell (ci, cj, ck) owns elements
    6 (ci + n (cj + n ck)) + 1 .. + 6. Interior node coordinates are moved by
    (2u - 1) * jitter * h with u = uniform01(seed, 3 * node0 + axis).
    """
    if n < 1:
        raise ValueError("n must be >= 1")
    n1 = n + 1
    h = (hi - lo) / n
    idx = np.arange(n1, dtype=np.int64)
    # node coordinates, i fastest
    K, J, I = np.meshgrid(idx, idx, idx, indexing="ij")
    I = I.ravel()
    J = J.ravel()
    K = K.ravel()
    nv = n1 ** 3
    p = np.empty((3, nv), dtype=np.float64, order="F")
    p[0] = lo + I * h
    p[1] = lo + J * h
    p[2] = lo + K * h
    if jitter:
        interior = (I > 0) & (I < n) & (J > 0) & (J < n) & (K > 0) & (K < n)
        nodes = np.nonzero(interior)[0].astype(np.uint64)
        for axis in range(3):
            u = uniform01(seed, np.uint64(3) * nodes + np.uint64(axis))
            p[axis, nodes] += (2.0 * u - 1.0) * jitter * h
    # connectivity
    cidx = np.arange(n, dtype=np.int64)
    CK, CJ, CI = np.meshgrid(cidx, cidx, cidx, indexing="ij")
    CI = CI.ravel()
    CJ = CJ.ravel()
    CK = CK.ravel()
    off = _kuhn_local_offsets()  # (6,4,3)
    # node number of corner offset (di,dj,dk) for every cell: (ncell, 6, 4)
    ni = CI[:, None, None] + off[None, :, :, 0]
    nj = CJ[:, None, None] + off[None, :, :, 1]
    nk = CK[:, None, None] + off[None, :, :, 2]
    nodes0 = ni + n1 * (nj + n1 * nk)  # 0-based
    t = np.asfortranarray((nodes0.reshape(-1, 4).T + 1).astype(np.int32))
    VE = element_volumes(p, t)
    m = MESH(p=p, t=t, VE=VE, n_cells=n)
    return m


def boundary_faces(mesh: MESH, tag: int = 1) -> MESH:
    """Fill surfaceT / normal / AE / shell_id with the faces used by exactly one
    element (the outer shell), normals pointing out of the element, as Gmsh
    delivers them to Femeko (src/mesh.jl:220-229)."""
    t0 = (mesh.t - 1).T.astype(np.int64)  # (nt, 4)
    local = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
    faces = t0[:, local].reshape(-1, 3)        # (4 nt, 3), face f opposite local node f % 4
    opp = t0.reshape(-1)                       # opposite node of each face
    key = np.sort(faces, axis=1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    bnd = cnt[inv.ravel()] == 1
    F = faces[bnd]
    o = opp[bnd]
    P = mesh.p
    A = P[:, F[:, 0]]
    B = P[:, F[:, 1]]
    C = P[:, F[:, 2]]
    n = np.cross((B - A).T, (C - A).T).T
    area2 = np.linalg.norm(n, axis=0)
    n = n / area2
    # orient outward: away from the opposite node
    flip = np.einsum("ij,ij->j", n, P[:, o] - A) > 0
    n[:, flip] *= -1
    F[flip, 1], F[flip, 2] = F[flip, 2].copy(), F[flip, 1].copy()
    ns = F.shape[0]
    sT = np.empty((4, ns), dtype=np.int32, order="F")
    sT[:3] = (F.T + 1)
    sT[3] = tag
    mesh.surfaceT = sT
    mesh.normal = np.asfortranarray(n)
    mesh.AE = area2 / 2
    mesh.shell_id = (tag,)
    return mesh


def boundary_node_mask(mesh: MESH, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    """uint8 mask, 1 on nodes lying on the faces of the box [lo, hi]^3."""
    p = mesh.p
    on = np.zeros(mesh.nv, dtype=bool)
    for axis in range(3):
        on |= (p[axis] == lo) | (p[axis] == hi)
    return on.astype(np.uint8)


def centroids(mesh: MESH) -> np.ndarray:
    """(3, nt) element centroids, sum(p[c, nds]) / 4 (magneticField.jl:81-87)."""
    t0 = mesh.t - 1
    p = mesh.p
    c = np.empty((3, mesh.nt), dtype=np.float64)
    for axis in range(3):
        c[axis] = (((p[axis, t0[0]] + p[axis, t0[1]]) + p[axis, t0[2]]) + p[axis, t0[3]]) / 4
    return c


def elements_in_box(mesh: MESH, box_lo, box_hi, cen: np.ndarray | None = None) -> np.ndarray:
    """Boolean element mask: centroid inside the closed box."""
    c = centroids(mesh) if cen is None else cen
    inside = np.ones(mesh.nt, dtype=bool)
    for axis in range(3):
        inside &= (c[axis] >= box_lo[axis]) & (c[axis] <= box_hi[axis])
    return inside


def set_inside(mesh: MESH, inside: np.ndarray) -> None:
    """Fill InsideElements / InsideNodes from an element mask (src/mesh.jl:186-213)."""
    mesh.InsideElements = (np.nonzero(inside)[0] + 1).astype(np.int32)
    mesh.InsideNodes = np.unique(mesh.t[:, inside].ravel()).astype(np.int32)


def rcm_permutation(mesh: MESH) -> np.ndarray:
    """Reverse Cuthill-McKee order of the mesh's node graph (two nodes
    adjacent when they share an element): perm[k] = the old number of new
    node k (0-based). Renumbering an unstructured mesh this way before upload
    keeps a row's neighbours near it in memory, which the solver's SELL copy
    needs for cache reuse (the renumbering Gmsh offers at export)."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    t0 = np.asarray(mesh.t, dtype=np.int64) - 1
    pairs = [(a, b) for a in range(4) for b in range(4) if a != b]
    rows = np.concatenate([t0[a] for a, _ in pairs])
    cols = np.concatenate([t0[b] for _, b in pairs])
    G = sp.csr_matrix((np.ones(rows.size, np.int8), (rows, cols)), shape=(mesh.nv, mesh.nv))
    return np.asarray(reverse_cuthill_mckee(G, symmetric_mode=True), dtype=np.int64)


def renumber_nodes(mesh: MESH, perm: np.ndarray) -> MESH:
    """The same mesh with node k = old node perm[k] (0-based): p's columns
    permuted, t and surfaceT mapped, elements and their order unchanged,
    InsideNodes mapped. Node data of the caller maps the same way
    (new = old[perm])."""
    perm = np.asarray(perm, dtype=np.int64)
    if perm.shape != (mesh.nv,) or not np.array_equal(np.sort(perm), np.arange(mesh.nv)):
        raise ValueError("perm must be a permutation of the nodes")
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    t = np.asfortranarray((inv[np.asarray(mesh.t, dtype=np.int64) - 1] + 1).astype(np.int32))
    out = MESH(p=np.asfortranarray(mesh.p[:, perm]), t=t, VE=mesh.VE,
               InsideElements=mesh.InsideElements.copy(),
               InsideNodes=np.sort((inv[np.asarray(mesh.InsideNodes, dtype=np.int64) - 1] + 1).astype(np.int32)),
               normal=mesh.normal, AE=mesh.AE, shell_id=mesh.shell_id, order=mesh.order, n_cells=0)
    if mesh.surfaceT is not None:
        sT = np.array(mesh.surfaceT, dtype=np.int32, order="F", copy=True)
        sT[:3] = inv[sT[:3].astype(np.int64) - 1] + 1
        out.surfaceT = sT
    return out


def kuhn_counts(n: int) -> dict:
    """Closed-form sizes of the Kuhn grid: nt, nv, CSR nnz, free and boundary nodes."""
    n1 = n + 1
    edges = 3 * n * n1 * n1 + 3 * n * n * n1 + n ** 3
    return {
        "nt": 6 * n ** 3,
        "nv": n1 ** 3,
        "nnz": n1 ** 3 + 2 * edges,
        "free": (n - 1) ** 3,
        "dirichlet": n1 ** 3 - (n - 1) ** 3,
    }


# ---- text dumps of a mesh -------------------------------------------------
# Femeko's mesh builder can save a mesh as text files (src/mesh.jl:232-242):
# one file per field, written by save2file (src/geometry.jl:236-243), which
# prints each row of the matrix on one line with the entries joined by " , ".
# A 3 x nv or 4 x nt matrix therefore takes 3 or 4 lines; a vector (VE, AE,
# InsideElements, InsideNodes) takes one line per entry. These readers let a
# mesh built by Femeko (Gmsh) be loaded here without Julia.

_DUMP_MATRICES = {"p.txt": 3, "t.txt": 4, "surfaceT.txt": 4, "normal.txt": 3}
_DUMP_VECTORS = ("VE.txt", "AE.txt", "InsideElements.txt", "InsideNodes.txt")


def _read_rows(path, parse) -> list[list]:
    rows = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line:
                rows.append([parse(tok) for tok in line.split(",")])
    return rows


def _read_matrix(path, nrows: int, dtype) -> np.ndarray:
    parse = float if dtype == np.float64 else int
    rows = _read_rows(path, parse)
    if not rows:
        return np.zeros((nrows, 0), dtype=dtype, order="F")
    if len(rows) != nrows or len({len(r) for r in rows}) != 1:
        raise ValueError(f"{path}: expected {nrows} rows of equal length, got {[len(r) for r in rows]}")
    return np.asfortranarray(np.array(rows, dtype=dtype))


def _read_vector(path, dtype) -> np.ndarray:
    parse = float if dtype == np.float64 else int
    rows = _read_rows(path, parse)
    if any(len(r) != 1 for r in rows):
        raise ValueError(f"{path}: expected one value per line")
    return np.array([r[0] for r in rows], dtype=dtype)


def load_mesh_dump(folder) -> MESH:
    """A MESH from Femeko's text dump (t.txt and p.txt required; VE.txt,
    InsideElements.txt, InsideNodes.txt, surfaceT.txt, AE.txt, normal.txt
    read when present; VE recomputed by elementVolume when absent). Only P1
    meshes (4 nodes per element) are accepted; node numbers are 1-based as
    written. Raises ValueError on a malformed file or an out-of-range index."""
    import os
    path = lambda name: os.path.join(folder, name)  # noqa: E731
    rows_t = _dump_rows(path("t.txt"))
    if rows_t != 4:
        raise ValueError(f"t.txt: a P1 mesh has 4 node rows, found {rows_t} (second-order meshes are not supported)")
    t = _read_matrix(path("t.txt"), 4, np.int32)
    p = _read_matrix(path("p.txt"), 3, np.float64)
    nv, nt = p.shape[1], t.shape[1]
    if nt and (t.min() < 1 or t.max() > nv):
        raise ValueError("t.txt: node numbers outside 1..nv")
    if not np.all(np.isfinite(p)):
        raise ValueError("p.txt: non-finite coordinates")
    VE = _read_vector(path("VE.txt"), np.float64) if os.path.exists(path("VE.txt")) else element_volumes(p, t)
    if VE.size != nt:
        raise ValueError("VE.txt: one volume per element expected")
    mesh = MESH(p=p, t=t, VE=VE)
    if os.path.exists(path("InsideElements.txt")):
        ie = _read_vector(path("InsideElements.txt"), np.int32)
        if ie.size and (ie.min() < 1 or ie.max() > nt):
            raise ValueError("InsideElements.txt: element numbers outside 1..nt")
        mesh.InsideElements = ie
    if os.path.exists(path("InsideNodes.txt")):
        inn = _read_vector(path("InsideNodes.txt"), np.int32)
        if inn.size and (inn.min() < 1 or inn.max() > nv):
            raise ValueError("InsideNodes.txt: node numbers outside 1..nv")
        mesh.InsideNodes = inn
    if os.path.exists(path("surfaceT.txt")):
        sT = _read_matrix(path("surfaceT.txt"), 4, np.int32)
        if sT.shape[1] and (sT[:3].min() < 1 or sT[:3].max() > nv):
            raise ValueError("surfaceT.txt: node numbers outside 1..nv")
        mesh.surfaceT = sT
        ns = sT.shape[1]
        if os.path.exists(path("AE.txt")):
            mesh.AE = _read_vector(path("AE.txt"), np.float64)
            if mesh.AE.size != ns:
                raise ValueError("AE.txt: one area per surface triangle expected")
        if os.path.exists(path("normal.txt")):
            mesh.normal = _read_matrix(path("normal.txt"), 3, np.float64)
            if mesh.normal.shape[1] != ns:
                raise ValueError("normal.txt: one normal per surface triangle expected")
        mesh.shell_id = outer_shell_tags(t, sT)
    return mesh


def outer_shell_tags(t: np.ndarray, surfaceT: np.ndarray) -> tuple:
    """Surface tags (row 4 of surfaceT) whose triangles all lie on the outer
    boundary, i.e. are faces of exactly one element. Femeko's surfaceT holds
    every surface triangle Gmsh meshed, internal interfaces between bodies
    included (src/mesh.jl:147-166), while its shell_id keeps only the outer
    surfaces of the bounding shell (unifyModel's setdiff with the inner cells'
    boundary, src/geometry.jl:225-230): a dump does not store shell_id, so
    it is recovered from the connectivity."""
    sT = np.asarray(surfaceT)
    if sT.shape[1] == 0:
        return ()
    t0 = np.asarray(t, dtype=np.int64) - 1
    local = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
    faces = np.sort(t0.T[:, local].reshape(-1, 3), axis=1)
    keys, cnt = np.unique(faces, axis=0, return_counts=True)
    outer = keys[cnt == 1]
    nv = int(t0.max()) + 1 if t0.size else 0
    code = lambda f: (f[:, 0] * nv + f[:, 1]) * nv + f[:, 2]  # noqa: E731
    tri = np.sort(np.asarray(sT[:3], dtype=np.int64).T - 1, axis=1)
    on_outer = np.isin(code(tri), code(outer))
    tags = np.asarray(sT[3], dtype=np.int64)
    return tuple(int(g) for g in np.unique(tags) if on_outer[tags == g].all())


def _dump_rows(path) -> int:
    with open(path) as f:
        return sum(1 for line in f if line.strip())


def save_mesh_dump(mesh: MESH, folder) -> None:
    """Write a MESH in the text-dump layout that load_mesh_dump and Femeko's
    save2file use (rows on lines, entries joined by " , "; floats in their
    shortest round-trip form, so a dump read back gives the same bits)."""
    import os
    os.makedirs(folder, exist_ok=True)

    def put(name, rows):
        with open(os.path.join(folder, name), "w") as f:
            for row in rows:
                f.write(" , ".join(repr(float(v)) if isinstance(v, (float, np.floating)) else str(int(v))
                                   for v in row) + "\n")

    put("p.txt", np.asarray(mesh.p, dtype=np.float64))
    put("t.txt", np.asarray(mesh.t, dtype=np.int64))
    put("VE.txt", np.asarray(mesh.VE, dtype=np.float64).reshape(-1, 1))
    put("InsideElements.txt", np.asarray(mesh.InsideElements, dtype=np.int64).reshape(-1, 1))
    put("InsideNodes.txt", np.asarray(mesh.InsideNodes, dtype=np.int64).reshape(-1, 1))
    if mesh.surfaceT is not None:
        put("surfaceT.txt", np.asarray(mesh.surfaceT, dtype=np.int64))
        if mesh.AE is not None:
            put("AE.txt", np.asarray(mesh.AE, dtype=np.float64).reshape(-1, 1))
        if mesh.normal is not None:
            put("normal.txt", np.asarray(mesh.normal, dtype=np.float64))


def rcm_order(mesh: MESH) -> np.ndarray:
    """The library's reverse Cuthill-McKee order of the mesh's nodes (the
    internal numbering fk_mesh_create picks for a non-stencil mesh):
    perm[k] = the old number of new node k (0-based). Runs on the host."""
    from . import _lib
    L = _lib.load()
    t = np.ascontiguousarray(np.asarray(mesh.t, dtype=np.int32).T)
    perm = np.empty(mesh.nv, np.int32)
    s = L.fk_rcm_order(_lib.iptr(t), int(mesh.nv), int(t.shape[0]), 1, _lib.iptr(perm))
    if s:
        raise _lib.FemekoHIPError(s, "fk_rcm_order")
    return perm.astype(np.int64)
