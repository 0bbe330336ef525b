"""Pins of the CPU oracle (oracle/femeko_oracle.c, oracle/oracle.py).

The reference cannot run here (no Julia, empty Eigen submodule), so every
oracle function is checked against something it does not compute itself:
closed forms on the unjittered Kuhn tetrahedron, LAPACK solves of the abcd
system, brute-force pair sets, scipy's COO->CSC conversion and direct solver,
the P1 patch test, algebraic identities and finite differences.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from femeko_hip import workloads as W

import linesearch_cases as LS
from femeko_hip.mesh import MESH, MU0, boundary_faces, boundary_node_mask, element_volumes, kuhn_box_mesh, kuhn_counts


def kuhn_tet(h=0.25):
    p = np.asfortranarray(np.array([[0, 0, 0], [h, 0, 0], [h, h, 0], [h, h, h]], float).T)
    t = np.asfortranarray(np.array([[1, 2, 3, 4]], np.int32).T)
    return MESH(p=p, t=t, VE=element_volumes(p, t))


# ---------------------------------------------------------------------------
# A1/A2: abcd and element volumes
def test_abcd_closed_form_kuhn_tet(oracle):
    h = 0.25
    m = kuhn_tet(h)
    expect = np.array([[-1, 0, 0], [1, -1, 0], [0, 1, -1], [0, 0, 1]], float) / h
    for i in range(4):
        a, b, c, d = oracle.abcd(m.p, m.t[:, 0], i + 1)
        np.testing.assert_allclose([b, c, d], expect[i], rtol=0, atol=1e-14 / h)
        assert abs(a - (1.0 if i == 0 else 0.0)) < 1e-14  # lambda_i at the origin
    assert m.VE[0] == pytest.approx(h ** 3 / 6, rel=1e-15)


def test_abcd_matches_lapack_solve(oracle):
    rng = np.random.default_rng(1)
    for _ in range(50):
        P = rng.normal(size=(3, 4))
        p = np.asfortranarray(P)
        nodes = np.array([1, 2, 3, 4], np.int32)
        for nd in range(1, 5):
            order = [nd] + [k for k in nodes if k != nd]
            M = np.column_stack([np.ones(4), P[:, np.array(order) - 1].T])
            ref = np.linalg.solve(M, np.eye(4)[0])
            got = np.array(oracle.abcd(p, nodes, nd))
            np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())


def test_gradient_invariants(oracle):
    m = kuhn_box_mesh(4, 0.2, seed=3)
    G = oracle.element_gradients(m)           # (nt,4,3)
    t0 = m.t.T - 1
    scale = np.abs(G).max()
    np.testing.assert_allclose(G.sum(axis=1), 0.0, atol=1e-13 * scale)
    X = m.p.T[t0]                              # (nt,4,3)
    for j in range(4):
        lam = np.einsum("kic,kc->ki", G, X[:, j] - X[:, 0])  # grad_i . (x_j - x_0)
        expect = np.zeros(4)
        expect[j] += 1
        expect[0] -= 1
        np.testing.assert_allclose(lam, np.broadcast_to(expect, lam.shape), atol=1e-12)


def test_volumes_sum_to_box(oracle):
    m = kuhn_box_mesh(6, 0.2, seed=4)
    assert m.VE.sum() == pytest.approx(8.0, rel=1e-13)
    lib = oracle.lib()
    p = np.ascontiguousarray(m.p.T)
    for k in (0, 17, m.nt - 1):
        nds = np.ascontiguousarray(m.t[:, k], dtype=np.int32)
        assert lib.or_element_volume(oracle._d(p), oracle._i(nds)) == m.VE[k]


# ---------------------------------------------------------------------------
# A3: local stiffness
def test_local_stiffness_closed_form(oracle):
    h = 0.5
    m = kuhn_box_mesh(1, 0.0, lo=0.0, hi=h)
    Ak = oracle.localStiffnessMatrix(m)
    path = h / 6 * np.array([[1, -1, 0, 0], [-1, 2, -1, 0], [0, -1, 2, -1], [0, 0, -1, 1]], float)
    odd = (False, True, True, False, False, True)
    for k in range(6):
        K = Ak[:, k].reshape(4, 4, order="F")
        expect = path.copy()
        if odd[k]:  # last two nodes swapped to keep the orientation positive
            expect = expect[[0, 1, 3, 2]][:, [0, 1, 3, 2]]
        np.testing.assert_allclose(K, expect, atol=1e-15)


def test_local_stiffness_sym_rowsum_psd(oracle):
    m = kuhn_box_mesh(3, 0.2, seed=5)
    f = np.linspace(1, 3, m.nt)
    Ak = oracle.localStiffnessMatrix(m, f)
    for k in range(0, m.nt, 7):
        K = Ak[:, k].reshape(4, 4, order="F")
        assert np.array_equal(K, K.T)
        np.testing.assert_allclose(K.sum(axis=1), 0, atol=1e-14 * np.abs(K).max())
        ev = np.linalg.eigvalsh(K)
        assert ev.min() > -1e-13 * ev.max()
        assert np.sum(ev > 1e-10 * ev.max()) == 3   # null space = constants


# ---------------------------------------------------------------------------
# A4/A5: sparse pattern and values
@pytest.mark.parametrize("n,jit", [(1, 0.0), (2, 0.2), (3, 0.2), (5, 0.0)])
def test_sparse_pattern_bruteforce(oracle, n, jit):
    m = kuhn_box_mesh(n, jit, seed=n)
    A = oracle.stiffnessMatrix(m)
    pairs = set()
    for k in range(m.nt):
        for a in m.t[:, k]:
            for b in m.t[:, k]:
                pairs.add((int(b), int(a)))   # (col, row)
    pairs = sorted(pairs)
    cols = np.array([c for c, _ in pairs])
    rows = np.array([r for _, r in pairs], np.int32)
    colptr = np.concatenate([[1], 1 + np.cumsum(np.bincount(cols - 1, minlength=m.nv))]).astype(np.int32)
    assert np.array_equal(A.colptr, colptr)
    assert np.array_equal(A.rowval, rows)
    assert A.nnz == kuhn_counts(n)["nnz"]


def test_sparse_values_match_scipy_coo(oracle):
    pr = W.linear(8)
    m = pr.mesh
    Ak = oracle.localStiffnessMatrix(m)
    A = oracle.sparse_from_local(m, Ak, pr.mu)
    I = []
    J = []
    V = []
    for n in range(16):
        i, j = divmod(n, 4)
        I.append(m.t[i] - 1)
        J.append(m.t[j] - 1)
        V.append(Ak[n] * pr.mu)
    S = sp.coo_matrix((np.concatenate(V), (np.concatenate(I), np.concatenate(J))), shape=(m.nv, m.nv)).tocsc()
    S.sort_indices()
    assert np.array_equal(S.indptr + 1, A.colptr)
    assert np.array_equal(S.indices + 1, A.rowval)
    np.testing.assert_allclose(A.nzval, S.data, rtol=1e-14, atol=1e-14 * np.abs(S.data).max())


def test_unjittered_stencil_and_zero_rowsums(oracle):
    n = 4
    h = 2.0 / n
    m = kuhn_box_mesh(n, 0.0)
    S = oracle.stiffnessMatrix(m).to_scipy().tocsr()
    np.testing.assert_allclose(S @ np.ones(m.nv), 0.0, atol=1e-14)
    # interior node: 7-point Laplacian times h, face and body diagonals stored as exact zeros
    i = j = k = 2
    r = i + (n + 1) * (j + (n + 1) * k)
    row = S.getrow(r)
    vals = dict(zip(row.indices, row.data))
    assert len(vals) == 15
    assert vals[r] == pytest.approx(6 * h, rel=1e-14)
    for d in (1, n + 1, (n + 1) ** 2):
        assert vals[r + d] == pytest.approx(-h, rel=1e-14)
        assert vals[r - d] == pytest.approx(-h, rel=1e-14)
    others = [v for c, v in vals.items() if abs(c - r) not in (0, 1, n + 1, (n + 1) ** 2)]
    assert len(others) == 8 and max(abs(v) for v in others) < 1e-15


def test_config_mesh_exact_zeros_only_between_fixed_nodes(oracle):
    # Femeko's incremental A += sparse(...) (src/FEM3D.jl:47-53) drops entries
    # whose sum is exactly 0.0; the single-call convention kept here
    # (NewtonRaphson.jl:161) stores them. On the jittered workload mesh the
    # only exact zeros are face diagonals lying in the flat outer faces, i.e.
    # entries between two Dirichlet nodes that the solve never reads.
    pr = W.linear(20)
    A = oracle.stiffnessMatrix(pr.mesh, pr.mu)
    col = np.repeat(np.arange(A.n), np.diff(A.colptr))
    row = A.rowval - 1
    z = A.nzval == 0.0
    assert np.count_nonzero(z) == 468
    assert np.all(pr.mask[row[z]] == 1) and np.all(pr.mask[col[z]] == 1)


# ---------------------------------------------------------------------------
# A7/A8/A9: Dirichlet solve, fields, energy
def test_patch_test_linear_field(oracle):
    m = kuhn_box_mesh(6, 0.2, seed=6)
    H0 = np.array([0.3, -1.2, 0.7])
    mask = boundary_node_mask(m)
    g = -(H0 @ m.p)
    A = oracle.stiffnessMatrix(m, np.full(m.nt, 2.5))
    phi, it, rel, s = oracle.pcg_dirichlet(A, mask, np.where(mask == 1, g, 0.0), 1e-14)
    assert s == 0
    np.testing.assert_allclose(phi, g, atol=1e-12)
    H, Hn = oracle.field(m, phi)
    np.testing.assert_allclose(H, np.broadcast_to(H0[:, None], H.shape), atol=1e-10)


def test_layered_patch_test(oracle):
    # mu depends on x through slabs bounded by grid planes; x is not jittered,
    # so the piecewise-linear exact potential with mu_i phi'_i = const is in
    # the P1 space and the discrete solution reproduces it
    n = 8
    m = kuhn_box_mesh(n, 0.25, seed=7)
    h = 2.0 / n
    m.p[0] = -1.0 + h * np.round((m.p[0] + 1.0) / h)
    from femeko_hip.mesh import centroids
    cx = centroids(m)[0]
    edges = np.array([-1.0, -0.5, 0.25, 1.0])
    mus = np.array([1.0, 40.0, 3.0])
    slab = np.searchsorted(edges, cx) - 1
    mu = mus[slab]
    # exact potential: flux q = -mu phi' constant, phi(-1) = 0, phi(1) = -2
    L = np.diff(edges)
    q = 2.0 / np.sum(L / mus)
    breaks = np.concatenate([[0.0], np.cumsum(-q * L / mus)])

    def exact(x):
        s = np.clip(np.searchsorted(edges, x, side="right") - 1, 0, 2)
        return breaks[s] - q / mus[s] * (x - edges[s])

    m.VE = element_volumes(m.p, m.t)
    A = oracle.stiffnessMatrix(m, mu)
    mask = boundary_node_mask(m)
    ex = exact(m.p[0])
    phi, it, rel, s = oracle.pcg_dirichlet(A, mask, np.where(mask == 1, ex, 0.0), 1e-14)
    assert s == 0
    np.testing.assert_allclose(phi, ex, atol=1e-11)


def test_pcg_matches_direct_solve(oracle):
    pr = W.linear(20)
    A = oracle.stiffnessMatrix(pr.mesh, pr.mu)
    S = A.to_scipy().tocsr()
    free = pr.mask == 0
    b = -S[free][:, ~free] @ pr.g[~free]
    ref = pr.g.copy()
    ref[free] = sla.spsolve(S[free][:, free].tocsc(), b)
    phi, it, rel, s = oracle.pcg_dirichlet(A, pr.mask, pr.g, 1e-14)
    assert s == 0
    assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) < 1e-12
    phi10, it10, rel10, s10 = oracle.pcg_dirichlet(A, pr.mask, pr.g, 1e-10)
    assert s10 == 0 and rel10 < 1e-10
    assert np.linalg.norm(phi10 - ref) / np.linalg.norm(ref) < 1e-9


def test_pcg_iteration_count_matches_scipy_cg(oracle):
    pr = W.linear(12)
    A = oracle.stiffnessMatrix(pr.mesh, pr.mu)
    S = A.to_scipy().tocsr()
    free = pr.mask == 0
    Aff = S[free][:, free]
    b = -S[free][:, ~free] @ pr.g[~free]
    d = Aff.diagonal()
    M = sp.diags(1.0 / d)
    count = [0]

    def cb(_x):
        count[0] += 1

    _x, info = sla.cg(Aff, b, rtol=1e-10, atol=0.0, M=M, callback=cb, maxiter=10000)
    assert info == 0
    phi, it, rel, s = oracle.pcg_dirichlet(A, pr.mask, pr.g, 1e-10)
    assert abs(it - count[0]) <= 2


def test_energy_identity(oracle):
    # 1/2 phi' A phi == 1/2 sum mu |H|^2 VE for any nodal phi: assembly vs field
    pr = W.linear(8)
    A = oracle.stiffnessMatrix(pr.mesh, pr.mu).to_scipy()
    rng = np.random.default_rng(8)
    phi = rng.normal(size=pr.mesh.nv)
    H, Hn = oracle.field(pr.mesh, phi)
    e1 = 0.5 * phi @ (A @ phi)
    e2 = oracle.energy(pr.mesh, Hn, pr.mu)
    assert e1 == pytest.approx(e2, rel=1e-12)


def test_lagrange_and_boundary_integral(oracle):
    m = boundary_faces(kuhn_box_mesh(4, 0.2, seed=9))
    C = oracle.lagrange(m)
    assert C.sum() == pytest.approx(m.VE.sum(), rel=1e-14)
    # closed surface: sum_s n_s A_s = 0, so the flux load sums to zero
    F = np.array([1.0, 2.0, -0.5])
    R = oracle.BoundaryIntegral(m, F, m.shell_id)
    assert abs(R.sum()) < 1e-13
    # only the x = +1 face carries F = e_x flux: total = area 4
    Fx = np.array([1.0, 0, 0])
    Rx = oracle.BoundaryIntegral(m, Fx, m.shell_id)
    on_pos = m.p[0] == 1.0
    assert Rx[on_pos].sum() == pytest.approx(4.0, rel=1e-13)


def test_line_search(oracle):
    pr = W.linear(6)
    A = oracle.stiffnessMatrix(pr.mesh, pr.mu)
    S = A.to_scipy().tocsr()
    rng = np.random.default_rng(10)
    RHS = rng.normal(size=pr.mesh.nv) * 1e-8
    u = np.zeros(pr.mesh.nv)
    du = np.linalg.solve(S.toarray() + np.eye(pr.mesh.nv) * 1e-3 * S.diagonal().max(), RHS)
    r0 = np.linalg.norm(RHS)
    res = oracle.lineSearch(u, 50.0 * du, A, RHS)
    # a 50x overshoot is cut back until the residual drops below the old one
    assert res < r0
    assert np.linalg.norm(RHS - S @ u) == pytest.approx(res, rel=1e-12)



@pytest.mark.parametrize("case", range(len(LS.CASES)))
def test_line_search_closed_form(oracle, case):
    """The accepted step of lineSearch is fixed by alpha* = 2 r0.w/||w||^2
    (tests/linesearch_cases.py): a different step factor, a >= test or a
    flipped min_step branch changes k, u or the residue."""
    astar, min_step, k_expect = LS.CASES[case]
    pr = W.linear(6)
    A = oracle.stiffnessMatrix(pr.mesh, pr.mu)
    S = A.to_scipy().tocsr()
    rng = np.random.default_rng(100 + case)
    u, du, RHS = LS.make_case(S, rng, astar, pr.mesh.nv)
    k, alpha, res_expect, u_expect = LS.expected(S, u, du, RHS, min_step)
    assert k == k_expect
    u0 = u.copy()
    res = oracle.lineSearch(u, du, A, RHS, min_step)
    assert res == pytest.approx(res_expect, rel=1e-12)
    if k is None:
        assert np.array_equal(u, u0)
    else:
        np.testing.assert_allclose(u, u_expect, rtol=0, atol=1e-13 * np.abs(u_expect).max())
        assert not np.array_equal(u, u0) or k_expect is None


# ---------------------------------------------------------------------------
# the reference's saddle Newton loop (NewtonRaphson.jl:109-212)
def _neumann_case(oracle, n=8, seed=12, Hext=(1e4, 3e3, -2e3)):
    from femeko_hip.mesh import boundary_faces
    pr = W.froehlich(n=n, seed=seed)
    m = boundary_faces(pr.mesh)
    RHS = -oracle.BoundaryIntegral(m, MU0 * np.asarray(Hext, float), m.shell_id)
    return m, pr.magnet, RHS


def test_newton_saddle_linear_law_is_the_saddle_solve(oracle):
    """A law that does not depend on |H| (mu = 50 mu0, dmu = 0): attempt 1
    solves with mu0, attempt 2 with the law's mu exactly (the tangent is 0,
    lineSearch takes the full step), attempt 3 changes nothing, so the loop
    ends after 3 attempts at the bordered direct solve with that mu."""
    m, magnet, RHS = _neumann_case(oracle)
    law = lambda h: (np.full(h.size, 50.0 * MU0), np.zeros(h.size))  # noqa: E731
    r = oracle.newton_saddle(m, RHS, magnet, law, tol_H=1e-12)
    assert r["iters"] == 3 and r["div"] < 1e-15
    mu = np.where(magnet, 50.0 * MU0, MU0)
    u, _ = oracle.saddle_solve(oracle.stiffnessMatrix(m, mu), oracle.lagrange(m), RHS)
    assert np.linalg.norm(r["phi"] - u) / np.linalg.norm(u) < 1e-12


def test_newton_saddle_converges_to_the_nonlinear_fixed_point(oracle):
    """Froehlich law: at the end u solves A(mu(|grad u|)) u = RHS with mu from
    the law evaluated here at the final field (not the loop's arrays), keeps
    Lag'u = 0, and the last changes fall quadratically (Newton)."""
    m, magnet, RHS = _neumann_case(oracle)
    Ms, a = 1e5, 500.0
    r = oracle.newton_saddle(m, RHS, magnet, lambda h: oracle.froehlich(h, Ms, a), tol_H=1e-9)
    assert r["div"] <= 1e-9 and r["iters"] < 100
    _H, Hn = oracle.field(m, r["phi"])
    mu = np.where(magnet, MU0 * (1.0 + Ms / (a + Hn)), MU0)
    A = oracle.stiffnessMatrix(m, mu)
    assert np.linalg.norm(RHS - oracle.spmv(A, r["phi"])) <= 1e-9 * np.linalg.norm(RHS)
    assert abs(oracle.lagrange(m) @ r["phi"]) <= 1e-12 * np.abs(r["phi"]).max()
    d = r["divs"]
    assert d[-1] < 1e-3 * d[-2] and d[-2] < 1e-2 * d[-3]


# ---------------------------------------------------------------------------
# A6/A10: tangent and the mu update
def test_froehlich_derivative_fd(oracle):
    H = np.array([0.0, 10.0, 500.0, 3e4])
    Ms, a = 1.2e6, 500.0
    _mu, dmu = oracle.froehlich(H, Ms, a)
    eps = 1e-3
    fd = (oracle.froehlich(H + eps, Ms, a)[0] - oracle.froehlich(H - eps, Ms, a)[0]) / (2 * eps)
    np.testing.assert_allclose(dmu, fd, rtol=1e-6)


def test_tangent_directional_derivative(oracle):
    m = kuhn_box_mesh(4, 0.2, seed=11)
    Ms, a = 1.2e6, 500.0
    rng = np.random.default_rng(11)
    phi = rng.normal(size=m.nv) * 300.0
    delta = rng.normal(size=m.nv)
    Ak = oracle.localStiffnessMatrix(m)

    def R(u):
        _H, Hn = oracle.field(m, u)
        mu, _d = oracle.froehlich(Hn, Ms, a)
        return oracle.spmv(oracle.sparse_from_local(m, Ak, mu), u)

    H, Hn = oracle.field(m, phi)
    mu, dmu = oracle.froehlich(Hn, Ms, a)
    A = oracle.sparse_from_local(m, Ak, mu)
    At = oracle.tangentialStiffnessMatrix(m, H, dmu)
    J = oracle.CSC(A.n, A.colptr, A.rowval, A.nzval + At.nzval)
    eps = 1e-4
    fd = (R(phi + eps * delta) - R(phi - eps * delta)) / (2 * eps)
    Jd = oracle.spmv(J, delta)
    assert np.linalg.norm(fd - Jd) / np.linalg.norm(Jd) < 1e-6
    # without the tangent the derivative is wrong by far more
    Ad = oracle.spmv(A, delta)
    assert np.linalg.norm(fd - Ad) / np.linalg.norm(Jd) > 1e-3


def test_newton_froehlich_fixed_point(oracle):
    pr = W.froehlich(n=8, seed=12)
    out = oracle.newton_froehlich(pr.mesh, pr.mask, pr.g, pr.magnet, pr.meta["Ms"], pr.meta["a"])
    assert out["iters"] < 20
    assert out["dmu_hist"][-1] < 1e-6
    # at the fixed point mu == mu(|H|) and A(mu) phi == 0 on free rows
    mu, _d = oracle.froehlich(out["Hn"], pr.meta["Ms"], pr.meta["a"])
    mu = np.where(pr.magnet, mu, MU0)
    np.testing.assert_allclose(out["mu"], mu, rtol=1e-5)
    A = oracle.stiffnessMatrix(pr.mesh, out["mu"]).to_scipy()
    r = (A @ out["phi"])[pr.mask == 0]
    bref = (A @ pr.g)[pr.mask == 0]
    assert np.linalg.norm(r) / np.linalg.norm(bref) < 1e-6
    # Newton converges faster than the Picard iteration on the same problem
    pic = oracle.newton_froehlich(pr.mesh, pr.mask, pr.g, pr.magnet, pr.meta["Ms"], pr.meta["a"], newton=False,
                                  max_outer=40)
    assert pic["iters"] > out["iters"]


# ---------------------------------------------------------------------------
# A14: adjoint
def test_adjoint_gradient_central_differences(oracle):
    pr = W.adjoint(n=6, seed=13)
    m = pr.mesh
    rho = pr.meta["rho"]
    target = pr.magnet
    Hs = pr.meta["Hstar"]
    J, grad, _phi, _lam = oracle.adjoint_gradient(m, pr.mask, pr.g, rho, target, Hs)
    rng = np.random.default_rng(13)
    idx = rng.choice(m.nt, 8, replace=False)
    h = 1e-6
    for e in idx:
        rp = rho.copy()
        rp[e] += h
        rm = rho.copy()
        rm[e] -= h
        fd = (oracle.objective(m, pr.mask, pr.g, rp, target, Hs) - oracle.objective(m, pr.mask, pr.g, rm, target, Hs)) / (2 * h)
        assert fd == pytest.approx(grad[e], rel=2e-4, abs=1e-9 * np.abs(grad).max())


def test_adjoint_gradient_taylor_remainder(oracle):
    """Taylor test over all elements at once: along a random direction d,
    R(eps) = J(rho + eps d) - J(rho) - eps grad.d is O(eps^2) when grad is
    the gradient, so halving eps divides it by 4; with a wrong gradient (a
    dropped term, a flipped sign) R is O(eps) and halves instead."""
    pr = W.adjoint(n=5, seed=17)
    m = pr.mesh
    rho = pr.meta["rho"]
    target = pr.magnet
    Hs = pr.meta["Hstar"]
    J0, grad, _phi, _lam = oracle.adjoint_gradient(m, pr.mask, pr.g, rho, target, Hs, rtol=1e-14)
    d = np.random.default_rng(17).uniform(-1.0, 1.0, m.nt) * 0.1

    def remainder(eps, g):
        return oracle.objective(m, pr.mask, pr.g, rho + eps * d, target, Hs, rtol=1e-14) - J0 - eps * (g @ d)

    eps = [4e-3, 2e-3, 1e-3]
    r = [abs(remainder(e, grad)) for e in eps]
    for a, b in zip(r, r[1:]):
        assert 3.5 < a / b < 4.5, r
    # the same test rejects a gradient with a term dropped (half of it)
    rw = [abs(remainder(e, 0.5 * grad)) for e in eps]
    for a, b in zip(rw, rw[1:]):
        assert 1.8 < a / b < 2.2, rw


def test_kuhn_counts():
    for n in (1, 2, 7, 20):
        m = kuhn_box_mesh(n)
        c = kuhn_counts(n)
        assert m.nt == c["nt"] and m.nv == c["nv"]
        assert int(boundary_node_mask(m).sum()) == c["dirichlet"]


def test_renumbered_problem_is_the_same_problem(oracle):
    # a node renumbering permutes the solution and nothing else
    from femeko_hip import workloads as W
    pr = W.linear(6)
    perm = np.random.default_rng(3).permutation(pr.mesh.nv)
    pr2 = W.renumbered(pr, perm, "random")
    A = oracle.stiffnessMatrix(pr.mesh, pr.mu)
    A2 = oracle.stiffnessMatrix(pr2.mesh, pr2.mu)
    phi, _, _, s = oracle.pcg_dirichlet(A, pr.mask, pr.g, 1e-13)
    phi2, _, _, s2 = oracle.pcg_dirichlet(A2, pr2.mask, pr2.g, 1e-13)
    assert s == 0 and s2 == 0
    assert np.abs(phi2 - phi[perm]).max() < 1e-10 * np.abs(phi).max()
    assert np.array_equal(pr2.mesh.VE, pr.mesh.VE)


# ---------------------------------------------------------------------------
# Tabulated mu(|H|): NewtonRaphson.jl:57-63, 204-207; magneticProperties.jl:163-188
GD270 = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "gd270_bh.npz")


def _gd_splines(oracle):
    from scipy.interpolate import splrep
    tab = np.load(GD270)
    mu, dmu = oracle.material_permeability(tab["H"], tab["B"])
    t, c, _ = splrep(tab["H"], mu, k=3, s=0)
    td, cd, _ = splrep(tab["H"], dmu, k=3, s=0)
    return (t, c[:t.size - 4]), (td, cd[:td.size - 4]), tab


def test_gradient1d_known_answer(oracle):
    z = oracle.gradient1d(np.array([0.0, 1.0, 3.0, 6.0]), np.array([0.0, 1.0, 9.0, 36.0]))
    np.testing.assert_array_equal(z, [1.0, 3.0, 7.0, 9.0])


def test_material_permeability_known_answer(oracle):
    # mu = B/H: 0/... at H = 0 is non-finite -> zeroed, then set to the maximum
    mu, dmu = oracle.material_permeability(np.array([0.0, 1.0, 2.0, 4.0]), np.array([0.5, 1.0, 1.5, 2.0]))
    np.testing.assert_array_equal(mu, [1.0, 1.0, 0.75, 0.5])
    np.testing.assert_allclose(dmu, [0.0, -0.125, -1.0 / 6.0, -0.125], rtol=0, atol=1e-15)
    # H[0] == H[2] makes the central difference at 1 infinite -> zeroed, then
    # set to the minimum of the rest
    mu, dmu = oracle.material_permeability(np.array([1.0, 2.0, 1.0, 4.0]), np.array([1.0, 3.0, 2.0, 4.0]))
    np.testing.assert_array_equal(mu, [1.0, 1.5, 2.0, 1.0])
    np.testing.assert_allclose(dmu, [0.5, -1.0 / 3.0, -0.25, -1.0 / 3.0], rtol=1e-15)


def test_bspline_eval_matches_fitpack(oracle):
    from scipy.interpolate import splev
    (t, c), (td, cd), tab = _gd_splines(oracle)
    H = tab["H"]
    xs = np.concatenate([H, np.linspace(H[0], H[-1], 2001), t[3:-3]])
    for (tt, cc) in ((t, c), (td, cd)):
        ref = splev(xs, (tt, np.concatenate([cc, np.zeros(4)]), 3))
        got = oracle.bspline_eval(tt, cc, xs)
        assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))
    # the interpolant passes through the table
    mu, _ = oracle.material_permeability(tab["H"], tab["B"])
    np.testing.assert_allclose(oracle.bspline_eval(t, c, H), mu, rtol=1e-12)
    with pytest.raises(ValueError):
        oracle.bspline_eval(t, c, [H[-1] * 1.0001])
    with pytest.raises(ValueError):
        oracle.bspline_eval(t, c, [-1.0])


def test_product_material_matches_oracle(oracle):
    from femeko_hip import material
    tab = np.load(GD270)
    mu, dmu, nu = material.material_permeability(tab["H"], tab["B"])
    mu_o, dmu_o = oracle.material_permeability(tab["H"], tab["B"])
    np.testing.assert_array_equal(mu, mu_o)
    np.testing.assert_array_equal(dmu, dmu_o)
    np.testing.assert_allclose(nu[1:], MU0 * tab["H"][1:] / tab["B"][1:])
    spl = material.Spline1D(tab["H"], mu)
    (t, c), _, _ = _gd_splines(oracle)
    np.testing.assert_array_equal(spl.t, t)
    np.testing.assert_array_equal(spl.c, c)
    np.testing.assert_allclose(spl(tab["H"][5:9]), mu[5:9], rtol=1e-12)
    with pytest.raises(ValueError):
        spl(tab["H"][-1] * 2)


def test_newton_spline_gd_converges_quadratically(oracle):
    (sm, sd, _tab) = _gd_splines(oracle)
    pr = W.tabulated(n=6, seed=21)
    out = oracle.newton_spline(pr.mesh, pr.mask, pr.g, pr.magnet, sm, sd, tol_H=1e-8, max_outer=30)
    h = out["dmu_hist"]
    assert out["iters"] < 12 and h[-1] < 1e-8 and out["steps"][0] == "P" and set(out["steps"][1:]) == {"N"}
    # superlinear tail (Newton): the contraction ratio itself shrinks tenfold
    # per step over the last three steps
    r = [h[i + 1] / h[i] for i in range(len(h) - 4, len(h) - 1)]
    assert r[1] < r[0] / 10 and r[2] < r[1] / 10
    # Picard steps only: linear convergence, many more iterations
    pic = oracle.newton_spline(pr.mesh, pr.mask, pr.g, pr.magnet, sm, sd, tol_H=1e-8, max_outer=60, newton=False)
    assert pic["iters"] > 2 * out["iters"]
    # fixed point: mu = spl(|H|) on the magnet and A(mu) phi = 0 on free rows
    np.testing.assert_allclose(out["mu"][pr.magnet], oracle.bspline_eval(*sm, out["Hn"][pr.magnet]), rtol=1e-6)
    A = oracle.stiffnessMatrix(pr.mesh, out["mu"]).to_scipy()
    r = (A @ out["phi"])[pr.mask == 0]
    assert np.linalg.norm(r) / np.linalg.norm((A @ pr.g)[pr.mask == 0]) < 1e-8


def test_newton_spline_reproduces_froehlich_loop(oracle):
    # a densely tabulated Froehlich law through the spline loop lands on the
    # Froehlich loop's fixed point (to the interpolation error)
    from scipy.interpolate import splrep
    Ms, a = 1.2e6, 500.0
    Ht = np.concatenate([[0.0], np.geomspace(1.0, 1e6, 600)])
    mu_t, dmu_t = oracle.froehlich(Ht, Ms, a)
    t, c, _ = splrep(Ht, mu_t, k=3, s=0)
    td, cd, _ = splrep(Ht, dmu_t, k=3, s=0)
    pr = W.froehlich(n=6, seed=12)
    ref = oracle.newton_froehlich(pr.mesh, pr.mask, pr.g, pr.magnet, Ms, a, tol_mu=1e-10, max_outer=40)
    # Newton steps once mu0 |dH| < 1e-5 T (undamped Newton from the start
    # oscillates for this law)
    out = oracle.newton_spline(pr.mesh, pr.mask, pr.g, pr.magnet, (t, c[:t.size - 4]), (td, cd[:td.size - 4]),
                               tol_H=1e-12, max_outer=60, switch=1e-5)
    assert np.linalg.norm(out["phi"] - ref["phi"]) / np.linalg.norm(ref["phi"]) < 1e-7
    np.testing.assert_allclose(out["mu"], ref["mu"], rtol=1e-6)


# ---------------------------------------------------------------------------
# Demagnetizing field (Micromagnetics/LL.jl:162-209)
def test_demag_uniform_m_everywhere_is_exact(oracle):
    # uniform M over the whole mesh: u = M . x solves A u = RHS exactly (a P1
    # function), so H = -M on every element and every node
    pr = W.linear(5)
    m = pr.mesh
    inside = np.ones(m.nt, bool)
    Mvec = np.array([0.3, -1.2, 0.7])
    M = np.repeat(Mvec[:, None], m.nv, axis=1)
    vol = oracle.node_volumes(m, inside)
    Hd, u, rhs = oracle.magnetostatic_field(m, inside, M, vol)
    assert abs(rhs.sum()) < 1e-12 * np.abs(rhs).sum()
    np.testing.assert_allclose(Hd, -np.repeat(Mvec[:, None], m.nv, axis=1), rtol=0, atol=1e-10)
    lin = Mvec @ m.p
    np.testing.assert_allclose(u - u.mean(), lin - lin.mean(), rtol=0, atol=1e-10)


def test_demag_energy_identity_and_nan_outside(oracle):
    # partial magnet, random M: RHS . u = u' A u (the load and the operator
    # agree in sign and scale), and the reference's 0/0 outside the magnet
    pr = W.linear(6)
    m = pr.mesh
    inside = pr.magnet
    rng = np.random.default_rng(4)
    M = rng.normal(size=(3, m.nv))
    vol = oracle.node_volumes(m, inside)
    Hd, u, rhs = oracle.magnetostatic_field(m, inside, M, vol)
    A = oracle.stiffnessMatrix(m).to_scipy()
    assert rhs @ u == pytest.approx(u @ (A @ u), rel=1e-10)
    out = vol == 0
    assert out.any() and np.all(np.isnan(Hd[:, out])) and np.all(np.isfinite(Hd[:, ~out]))
    assert abs(rhs.sum()) < 1e-12 * np.abs(rhs).sum()


def test_rcm_renumbering_permutes_the_matrix(oracle):
    # a randomly numbered mesh renumbered by RCM: the bandwidth drops to the
    # structured numbering's order and the stiffness matrix is the same matrix
    # with rows and columns permuted (A' = P A P^T, bitwise)
    from femeko_hip.mesh import rcm_permutation, renumber_nodes
    pr = W.linear(6)
    rng = np.random.default_rng(5)
    rnd = renumber_nodes(pr.mesh, rng.permutation(pr.mesh.nv))
    perm = rcm_permutation(rnd)
    m2 = renumber_nodes(rnd, perm)

    def bandwidth(m):
        A = oracle.stiffnessMatrix(m, pr.mu).to_scipy().tocoo()
        return int(np.max(np.abs(A.row - A.col)))

    assert bandwidth(m2) < bandwidth(rnd) / 4
    A1 = oracle.stiffnessMatrix(rnd, pr.mu).to_scipy().toarray()
    A2 = oracle.stiffnessMatrix(m2, pr.mu).to_scipy().toarray()
    np.testing.assert_array_equal(A2, A1[np.ix_(perm, perm)])
    with pytest.raises(ValueError):
        renumber_nodes(rnd, np.zeros(rnd.nv, np.int64))
