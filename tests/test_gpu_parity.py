"""GPU parity of the HIP library against the CPU oracle, through the C ABI.

Bars (written next to each check):
  - CSR/CSC pattern: bit-exact against the oracle's Julia sparse() arrays.
  - FP64 values: |a_gpu - a_ref| <= 1e-12 * sum|element terms| per entry.
  - potential: the GPU solves to rtol 1e-10 and is compared with a direct
    solve; ||phi - phi_ref|| / ||phi_ref|| <= 1e-9.
  - H, |H|, energy from the same potential: 1e-12 relative.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from femeko_hip import _lib
from femeko_hip import workloads as W
from femeko_hip.device import DeviceMesh, context
from femeko_hip._lib import FemekoHIPError
from femeko_hip.mesh import MESH, boundary_node_mask, element_volumes, kuhn_box_mesh

pytestmark = pytest.mark.gpu

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "linear20.npz")


def direct_solution(A_csc_scipy, mask, g):
    S = A_csc_scipy.tocsr()
    free = mask == 0
    phi = g.copy()
    phi[free] = sla.spsolve(S[free][:, free].tocsc(), -S[free][:, ~free] @ g[~free])
    return phi


def abs_terms(oracle, mesh, f):
    """per CSC entry: sum over elements of |VE f| ||grad_i|| ||grad_j||, the
    operand magnitude of every element term (scale of the value metric)"""
    return oracle.term_scale(mesh, f)


@pytest.fixture(scope="module")
def lin20(oracle):
    pr = W.linear(20)
    A = oracle.stiffnessMatrix(pr.mesh, pr.mu)
    dm = DeviceMesh(pr.mesh, context(0))
    dm.set_coefficient(pr.mu)
    dm.assemble()
    dm.set_dirichlet(pr.mask, pr.g)
    yield pr, A, dm
    dm.close()


# ---------------------------------------------------------------------------
# pattern and values
def test_pattern_bit_exact_linear20(lin20):
    pr, A, dm = lin20
    rowptr, colidx = dm.pattern(base=1)
    assert np.array_equal(rowptr, A.colptr)
    assert np.array_equal(colidx, A.rowval)
    assert dm.nnz == 128581


@pytest.mark.parametrize("n,jit,shuffle", [(1, 0.0, False), (3, 0.0, True), (5, 0.2, True)])
def test_pattern_bit_exact_small_and_shuffled(oracle, n, jit, shuffle):
    m = kuhn_box_mesh(n, jit, seed=n)
    if shuffle:
        # renumber nodes and reorder elements: the library must not rely on
        # the structured numbering
        rng = np.random.default_rng(n)
        pn = rng.permutation(m.nv)
        inv = np.empty_like(pn)
        inv[pn] = np.arange(m.nv)
        p = np.asfortranarray(m.p[:, pn])
        t = np.asfortranarray((inv[m.t - 1] + 1).astype(np.int32)[:, rng.permutation(m.nt)])
        m = MESH(p=p, t=t, VE=element_volumes(p, t))
    A = oracle.stiffnessMatrix(m)
    dm = DeviceMesh(m, context(0))
    try:
        rowptr, colidx = dm.pattern(base=1)
        assert np.array_equal(rowptr, A.colptr) and np.array_equal(colidx, A.rowval)
        r0, c0 = dm.pattern(base=0)
        assert np.array_equal(r0 + 1, rowptr) and np.array_equal(c0 + 1, colidx)
        f = np.linspace(0.5, 2.0, m.nt)
        dm.set_coefficient(f)
        Af = oracle.stiffnessMatrix(m, f)
        bound = 1e-12 * abs_terms(oracle, m, f)
        for method in ("coloured", "gather", "atomic"):
            dm.assemble(method)
            v = dm.values()
            assert np.all(np.abs(v - Af.nzval) <= bound), method
    finally:
        dm.close()


def test_values_fp64_metric(oracle, lin20):
    pr, A, dm = lin20
    v = dm.values()
    bound = abs_terms(oracle, pr.mesh, pr.mu)
    err = np.abs(v - A.nzval)
    assert np.all(err <= 1e-12 * bound), float((err / bound).max())
    assert np.linalg.norm(v - A.nzval) / np.linalg.norm(A.nzval) < 1e-14
    golden = np.load(GOLDEN)
    assert np.all(np.abs(v - golden["nzval"]) <= 1e-12 * bound)


def test_values_symmetric_deterministic_and_atomic(oracle, lin20):
    pr, A, dm = lin20
    dm.assemble("coloured")
    v1 = dm.values()
    dm.assemble("coloured")
    v2 = dm.values()
    assert np.array_equal(v1.view(np.uint64), v2.view(np.uint64))  # bitwise reproducible
    rowptr, colidx = dm.pattern(base=0)
    S = sp.csr_matrix((v1, colidx, rowptr), shape=(dm.nv, dm.nv))
    assert (S != S.T).nnz == 0  # exactly symmetric
    dm.assemble("atomic")
    va = dm.values()
    bound = abs_terms(oracle, pr.mesh, pr.mu)
    assert np.all(np.abs(va - A.nzval) <= 1e-12 * bound)
    dm.assemble("coloured")
    f32 = dm.values_f32()
    assert np.array_equal(f32, v1.astype(np.float32))
    dm.assemble()


def test_values_row_gather_assembly(oracle, lin20):
    # the row-gather assembly: same metric against the oracle, exactly
    # symmetric, bitwise reproducible
    pr, A, dm = lin20
    try:
        dm.assemble("gather")
        vg = dm.values()
        bound = abs_terms(oracle, pr.mesh, pr.mu)
        assert np.all(np.abs(vg - A.nzval) <= 1e-12 * bound)
        rowptr, colidx = dm.pattern(base=0)
        S = sp.csr_matrix((vg, colidx, rowptr), shape=(dm.nv, dm.nv))
        assert (S != S.T).nnz == 0
        dm.assemble("gather")
        assert np.array_equal(dm.values().view(np.uint64), vg.view(np.uint64))
        golden = np.load(GOLDEN)
        rep = dm.solve("f64", rtol=1e-10)
        assert np.linalg.norm(dm.potential() - golden["phi"]) / np.linalg.norm(golden["phi"]) < 1e-9
        # against the coloured scatter: the same terms summed in another order
        dm.assemble("coloured")
        assert np.all(np.abs(dm.values() - vg) <= 1e-13 * bound)
    finally:
        dm.assemble()


def test_local_stiffness_volumes_lagrange(oracle, lin20):
    pr, A, dm = lin20
    m = pr.mesh
    Ak = dm.local_stiffness(pr.mu)
    Ak_ref = oracle.localStiffnessMatrix(m, pr.mu)
    scale = np.abs(Ak_ref).max(axis=0)
    assert np.all(np.abs(Ak - Ak_ref) <= 1e-12 * scale[None, :])
    Ak1 = dm.local_stiffness(None)
    assert np.all(np.abs(Ak1 - oracle.localStiffnessMatrix(m)) <= 1e-12 * np.abs(Ak1).max(axis=0)[None, :])
    VE = dm.volumes()
    np.testing.assert_allclose(VE, m.VE, rtol=1e-13)
    C = dm.lagrange()
    np.testing.assert_allclose(C, oracle.lagrange(m), rtol=1e-13)


def test_tangent_values(oracle, lin20):
    pr, A, dm = lin20
    m = pr.mesh
    rng = np.random.default_rng(3)
    H = rng.normal(size=(3, m.nt)) * 1e3
    H[:, ::50] = 0.0  # |H| = 0 elements contribute nothing
    dmu = -np.abs(rng.normal(size=m.nt)) * 1e-9
    v = dm.tangential_values(H, dmu)
    ref = oracle.tangentialStiffnessMatrix(m, H, dmu).nzval
    scale = oracle.tangent_scale(m, H, dmu)
    assert np.all(np.abs(v - ref) <= 1e-12 * scale)


# ---------------------------------------------------------------------------
# solve and field
def test_solve_f64_linear20(oracle, lin20):
    pr, A, dm = lin20
    golden = np.load(GOLDEN)
    rep = dm.solve("f64", rtol=1e-10)
    phi = dm.potential()
    ref = golden["phi"]
    assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) < 1e-9
    assert rep.relres < 1e-10
    # same algorithm as the oracle: iteration counts agree within the spread that
    # last-bit differences cause at this permeability contrast (about +-3)
    it_ref = int(golden["pcg_iters_1e10"])
    assert abs(rep.iterations - it_ref) <= max(3, 0.05 * it_ref)
    # fixed nodes carry g exactly
    assert np.array_equal(phi[pr.mask == 1], pr.g[pr.mask == 1])


@pytest.mark.parametrize("delta", [0.1, 1e-5])
def test_solve_f32_refined_linear20(lin20, delta):
    # delta = 1e-5 lets the FP32 recursion drift far below the true residual
    # between updates; the direction restart keeps it convergent
    pr, A, dm = lin20
    golden = np.load(GOLDEN)
    rep = dm.solve("f32", rtol=1e-10, inner_rtol=delta)
    phi = dm.potential()
    ref = golden["phi"]
    assert rep.relres < 1e-10 and rep.outer >= 2
    assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) < 1e-9


@pytest.mark.parametrize("fmt", ["csr", "sell64", "sell64-rowidx", "stencil", "auto"])
@pytest.mark.parametrize("prec", ["f64", "f32", "f32-restart"])
def test_spmv_formats_solve_agree(lin20, prec, fmt):
    # the solver's CSR kernels, its SELL-64 copies (shared slice offsets or
    # per-row indices) and its stencil copy give the same solution
    pr, A, dm = lin20
    golden = np.load(GOLDEN)
    try:
        dm.set_format(fmt)
        rep = dm.solve(prec, rtol=1e-10)
        phi = dm.potential()
        assert rep.relres < 1e-10
        assert np.linalg.norm(phi - golden["phi"]) / np.linalg.norm(golden["phi"]) < 1e-9
        it_ref = int(golden["pcg_iters_1e10"])
        if prec == "f64":
            assert abs(rep.iterations - it_ref) <= max(3, 0.05 * it_ref)
    finally:
        dm.set_format("auto")


@pytest.mark.parametrize("shape", [(256, 1), (256, 2), (512, 1), (512, 2), (1024, 1)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_sell_launch_shapes_agree(lin20, prec, shape):
    # every SELL SpMV launch shape (1 or 2 slices in flight per wave, 4 to 16
    # waves per block) solves to the golden potential; rows are computed
    # alike, only the grouping of the p.q partial sums differs
    pr, A, dm = lin20
    golden = np.load(GOLDEN)
    try:
        dm.set_sell_shape(*shape)
        rep = dm.solve(prec, rtol=1e-10)
        phi = dm.potential()
        assert rep.relres < 1e-10
        assert np.linalg.norm(phi - golden["phi"]) / np.linalg.norm(golden["phi"]) < 1e-9
        if prec == "f64":
            it_ref = int(golden["pcg_iters_1e10"])
            assert abs(rep.iterations - it_ref) <= max(3, 0.05 * it_ref)
        # fixed-order reductions: a repeat is bitwise equal
        dm.solve(prec, rtol=1e-10)
        assert np.array_equal(dm.potential(), phi)
    finally:
        dm.set_sell_shape(256, 2)  # the default


def test_sell_shape_rejects_unknown():
    pr = W.linear(4)
    dm = DeviceMesh(pr.mesh)
    for shape in [(128, 1), (256, 3), (256, 0), (1024, 2), (512, -1)]:
        with pytest.raises(FemekoHIPError):
            dm.set_sell_shape(*shape)
    dm.close()


def test_sell_layout_statistics(lin20):
    # the Kuhn mesh has translation-invariant connectivity: every slice of
    # 64 rows shares one offset set (boundary rows store zeros for the
    # neighbours they lack); per-row indices store exactly the padded
    # off-diagonal rows
    pr, A, dm = lin20
    try:
        dm.set_format("sell64")
        ns, nsh, ent = dm.sell_info()
        assert ns == (dm.nv + 63) // 64 and nsh == ns
        # streamed bytes: 4 B per stored value, the shared offsets (at least
        # one per stored column of a slice) and 20 B of slice tables per slice
        ib, _ng, ab = dm.sell_layout()
        assert ab >= 4 * ent + 20 * ns
        dm.set_format("sell64-rowidx")
        ns2, nsh2, ent2 = dm.sell_info()
        assert ns2 == ns and nsh2 == 0
        # per-row indices: one per stored value
        ib2, _ng2, ab2 = dm.sell_layout()
        assert ab2 == (4 + ib2) * ent2 + 20 * ns2 + 8
        counts = np.diff(A.colptr) - 1  # off-diagonal entries per row
        padded = sum(64 * counts[s * 64:(s + 1) * 64].max() for s in range(ns))
        assert ent2 == padded
        dm.set_format("csr")
        with pytest.raises(FemekoHIPError):
            dm.sell_info()
        # back to the first layout with no solve in between: the rebuilt
        # arrays live at new addresses and the captured iteration graphs
        # must follow them
        dm.set_format("sell64")
        for prec in ("f64", "f32"):
            rep = dm.solve(prec, rtol=1e-10)
            golden = np.load(GOLDEN)
            assert rep.relres < 1e-10
            assert np.linalg.norm(dm.potential() - golden["phi"]) / np.linalg.norm(golden["phi"]) < 1e-9
    finally:
        dm.set_format("auto")


def test_stencil_format_on_kuhn_meshes(lin20):
    # Kuhn meshes in their natural numbering: 7 positive offsets
    # 1, n, n+1, n^2, n^2+1, n^2+n, n^2+n+1 of (n+1)-point grid lines
    pr, A, dm = lin20
    name, offs = dm.solver_format()
    m = 21
    assert name == "stencil"
    assert set(offs) <= {1, m, m + 1, m * m, m * m + 1, m * m + m, m * m + m + 1}
    assert len(offs) == 7
    # every positive column offset of the CSR pattern is one of them
    rowptr, colidx = dm.pattern(base=0)
    rows = np.repeat(np.arange(dm.nv), np.diff(rowptr))
    d = colidx - rows
    assert set(np.unique(d[d > 0]).tolist()) == set(offs)


def test_stencil_rejected_on_renumbered_mesh(oracle):
    # a random node numbering is no stencil: 'stencil' fails, 'auto' uses SELL
    m = kuhn_box_mesh(4, 0.1, seed=9)
    rng = np.random.default_rng(9)
    pn = rng.permutation(m.nv)
    inv = np.empty_like(pn)
    inv[pn] = np.arange(m.nv)
    p = np.asfortranarray(m.p[:, pn])
    t = np.asfortranarray((inv[m.t - 1] + 1).astype(np.int32))
    mm = MESH(p=p, t=t, VE=element_volumes(p, t))
    dm = DeviceMesh(mm, context(0))
    try:
        assert dm.solver_format()[0] == "sell64"
        dm.set_format("stencil")
        with pytest.raises(FemekoHIPError):
            dm.solver_format()
    finally:
        dm.close()


def test_true_residual_of_gpu_solution(lin20):
    pr, A, dm = lin20
    dm.solve("f64", rtol=1e-10)
    phi = dm.potential()
    S = A.to_scipy().tocsr()
    free = pr.mask == 0
    r = (S @ phi)[free]
    b = (S @ pr.g)[free]
    assert np.linalg.norm(r) / np.linalg.norm(b) < 2e-10


def test_field_and_energy(oracle, lin20):
    pr, A, dm = lin20
    golden = np.load(GOLDEN)
    dm.set_potential(golden["phi"])
    H, Hn, E = dm.field(want_energy=True)
    Href, Hnref = oracle.field(pr.mesh, golden["phi"])
    scale = np.abs(Hnref).max()
    assert np.abs(H - Href).max() <= 1e-12 * scale
    assert np.abs(Hn - golden["Hn"]).max() <= 1e-12 * scale
    assert E == pytest.approx(float(golden["energy"]), rel=1e-12)
    # energy identity: 1/2 phi' A phi from the GPU matrix equals the field energy
    rowptr, colidx = dm.pattern(base=0)
    S = sp.csr_matrix((dm.values(), colidx, rowptr), shape=(dm.nv, dm.nv))
    assert 0.5 * golden["phi"] @ (S @ golden["phi"]) == pytest.approx(E, rel=1e-12)


def test_patch_test_on_gpu():
    m = kuhn_box_mesh(12, 0.2, seed=21)
    H0 = np.array([0.4, 1.0, -0.3])
    mask = boundary_node_mask(m)
    g = -(H0 @ m.p)
    dm = DeviceMesh(m, context(0))
    try:
        dm.set_coefficient(np.full(m.nt, 3.0))
        dm.assemble()
        dm.set_dirichlet(mask, np.where(mask == 1, g, 0.0))
        for prec in ("f64", "f32"):
            dm.solve(prec, rtol=1e-12)
            phi = dm.potential()
            assert np.abs(phi - g).max() < 1e-9
            H, Hn, _ = dm.field()
            assert np.abs(H - H0[:, None]).max() < 1e-7
    finally:
        dm.close()


def test_solve_rhs_matches_oracle(oracle, lin20):
    pr, A, dm = lin20
    rng = np.random.default_rng(5)
    b = rng.normal(size=dm.nv) * 1e-7
    x, it, rel = dm.solve_rhs(b, rtol=1e-12)
    xr, itr, relr, s = oracle.pcg_free(A, pr.mask, b, 1e-13)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-9
    assert np.all(x[pr.mask == 1] == 0.0)


# ---------------------------------------------------------------------------
# error paths (all forced from the host)
def test_error_statuses():
    m = kuhn_box_mesh(3, 0.1, seed=2)
    ctx = context(0)
    bad = MESH(p=m.p, t=np.asfortranarray(m.t.copy()), VE=m.VE)
    bad.t[2, 5] = m.nv + 7
    with pytest.raises(FemekoHIPError) as ei:
        DeviceMesh(bad, ctx)
    assert ei.value.status == _lib.FK_ERR_INDEX
    rep = MESH(p=m.p, t=np.asfortranarray(m.t.copy()), VE=m.VE)
    rep.t[1, 0] = rep.t[0, 0]
    with pytest.raises(FemekoHIPError) as ei:
        DeviceMesh(rep, ctx, check_volumes=False)
    assert ei.value.status == _lib.FK_ERR_DEGENERATE
    flat = MESH(p=np.asfortranarray(m.p.copy()), t=m.t, VE=m.VE)
    nds = m.t[:, 0] - 1
    flat.p[2, nds] = 0.0  # all four nodes of element 1 coplanar
    flat.p[2, nds] = flat.p[2, nds[0]]
    with pytest.raises(FemekoHIPError) as ei:
        DeviceMesh(flat, ctx, check_volumes=False)
    assert ei.value.status in (_lib.FK_ERR_DEGENERATE, _lib.FK_ERR_INDEX)
    wrongVE = MESH(p=m.p, t=m.t, VE=m.VE * 1.01)
    with pytest.raises(FemekoHIPError) as ei:
        DeviceMesh(wrongVE, ctx)
    assert ei.value.status == _lib.FK_ERR_INVALID_ARG
    dm = DeviceMesh(m, ctx)
    try:
        with pytest.raises(FemekoHIPError) as ei:
            dm.assemble()
        assert ei.value.status == _lib.FK_ERR_STATE
        mu = np.ones(m.nt)
        mu[3] = np.nan
        with pytest.raises(FemekoHIPError) as ei:
            dm.set_coefficient(mu)
        assert ei.value.status == _lib.FK_ERR_NONFINITE
        dm.set_coefficient(np.ones(m.nt))
        dm.assemble()
        with pytest.raises(FemekoHIPError) as ei:
            dm.solve("f64")
        assert ei.value.status == _lib.FK_ERR_STATE
        dm.set_dirichlet(np.zeros(m.nv, np.uint8), np.zeros(m.nv))
        with pytest.raises(FemekoHIPError) as ei:
            dm.solve("f64")
        assert ei.value.status == _lib.FK_ERR_SINGULAR
        mask = boundary_node_mask(m)
        dm.set_dirichlet(mask, np.where(mask == 1, m.p[0], 0.0))
        rep = dm.solve("f64", rtol=1e-14, max_iter=2, raise_on_error=False)
        assert rep.status == _lib.FK_ERR_NOT_CONVERGED and rep.iterations == 2
        rep = dm.solve("f64", rtol=1e-12)
        assert rep.relres < 1e-12
    finally:
        dm.close()


def test_copy_probe_bandwidth_is_physical():
    ctx = context(0)
    ms = ctx.copy_probe(1 << 28, 3)
    gbs = 2 * (1 << 28) / (ms * 1e-3) / 1e9
    assert 100.0 < gbs < 8000.0  # below the HBM peak, far above PCIe
    with pytest.raises(FemekoHIPError):
        ctx.copy_probe(8, 1)


@pytest.mark.parametrize("fmt", ["stencil", "sell64", "csr"])
def test_spmv_burst_leaves_solver_usable(lin20, fmt):
    # the burst timing swaps in a state of its own and zeroes the direction
    # vector: the next solve (graph replays included) must be unaffected
    pr, A, dm = lin20
    try:
        dm.set_format(fmt)
        golden = np.load(GOLDEN)
        for prec in ("f32", "f64"):
            dm.solve(prec, rtol=1e-10)
            us = dm.spmv_burst(prec, 5)
            assert 0.0 < us < 1e5
            rep = dm.solve(prec, rtol=1e-10)
            assert rep.relres < 1e-10
            assert np.linalg.norm(dm.potential() - golden["phi"]) / np.linalg.norm(golden["phi"]) < 1e-9
        with pytest.raises(FemekoHIPError):
            dm.spmv_burst("f32", 0)
    finally:
        dm.set_format("auto")
