"""Benchmark: assembly + Jacobi-PCG throughput on one MI355X.

Workload (the mixed-precision configuration, the largest that the headline
metric is quoted on): Kuhn box mesh n = 160 (24,576,000 tets, 4,173,281 nodes,
61,979,041 CSR entries), interior jitter 0.15 h, 8 seeded boxes with
log-uniform mu_r in [1, 1e3], Dirichlet phi = -x on the outer faces. A step is
value-only assembly of A(mu) on the GPU plus the FP32 Jacobi-PCG whose
recursive residual is replaced by the FP64 true residual b - A64 phi each time
it has fallen by --delta (reliable updates), to FP64 relres 5e-11 (--rtol);
mesh, pattern, mu and the Dirichlet data are resident in HBM before timing
starts. The plain FP64 path runs to the configured 1e-10 (--rtol64): at
those tolerances both potentials are within about 1.5e-10 of an FP64 solve
at 1e-13, while the FP32-refined solve at 1e-10 left up to 1.06e-9 (its
remaining error sits in smooth modes that weigh little in the residual).

Reported beside the headline: the plain FP64 GPU solve of the same system
(speed-up), the classic restarted refinement (FP32 PCG to 1e-5 inside an FP64
residual-correction loop), the SpMV kernel's achieved bandwidth (back-to-back
launches between two HIP events on the library stream), and, after the GPU
part, the oracle on the host cores: its full assembly and Jacobi-PCG solve of
the same system to the FP64 path's tolerance (the CPU baseline, plus a
1-thread sample), continued to rtol 1e-13 as the FP64 reference that the
errors of the GPU potentials, |H| and energy are reported against.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 160] [--no-cpu]
                [--shape 256,2] [--delta 0.95] [--rtol 5e-11] [--rtol64 1e-10]
                [--format auto|stencil|sell64|sell64-rowidx|csr]
                [--workload boxes|batch|froehlich|tabulated|demag|adjoint]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "femeko.jl_amd")]

import numpy as np  # noqa: E402

METRIC = "assembly+CG DOF/s at 4.1M nodes, FP32-refined vs FP64, ≤1e-9 rel err; SpMV GB/s"
HBM_PEAK_GBS = 8000.0


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def spmv_bytes(nnz, nv, s):
    """compulsory bytes of one CSR SpMV: values + int32 columns, int32 row
    pointers, x read once, y written once"""
    return nnz * (s + 4) + 4 * (nv + 1) + 2 * s * nv


def stencil_spmv_bytes(K, nv, s):
    """compulsory bytes of one difference-form SpMV on the symmetric stencil
    copy: K value columns of nv rows (each off-diagonal pair stored once, no
    index), x read once, y written once"""
    return K * s * nv + 2 * s * nv


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_legs(pr, nthreads, rtol64):
    """The oracle (the C restatement of the CPU path) on the host, untimed
    for the GPU line:
      1. the CPU baseline on `nthreads` cores: full assembly (OpenMP element
         matrices + the serial sparse() build, FEMc.cpp:102-133) and a full
         Jacobi-PCG solve to rtol64, the FP64 GPU path's tolerance -- both
         measured, nothing extrapolated;
      2. the same PCG continued to rtol 1e-13 from that potential: the FP64
         reference the GPU potentials, |H| and energy are checked against;
      3. a 1-thread sample: element matrices of 1/16 of the elements and 10
         PCG iterations on one thread, scaled to the full mesh and to the
         iteration count of 1."""
    from oracle import oracle as O
    L = O.lib()
    m = pr.mesh
    t0 = time.perf_counter()
    Ak = O.localStiffnessMatrix(m, pr.mu, nthreads=nthreads)
    t_local = time.perf_counter() - t0
    t0 = time.perf_counter()
    A = O.sparse_from_local(m, Ak)
    t_sparse = time.perf_counter() - t0
    del Ak
    mask = np.ascontiguousarray(pr.mask, dtype=np.uint8)
    g = np.ascontiguousarray(pr.g, dtype=np.float64)
    phi = np.zeros(m.nv)
    it = O._i64(0)
    rel = O.ctypes.c_double(0.0)
    t0 = time.perf_counter()
    s1 = L.or_pcg_dirichlet(A.n, O._i(A.colptr), O._i(A.rowval), O._d(A.nzval), O._u(mask), O._d(g), rtol64,
                            1000000, O._d(phi), O.ctypes.byref(it), O.ctypes.byref(rel), nthreads, 0)
    t_solve = time.perf_counter() - t0
    it1, rel1 = int(it.value), float(rel.value)
    log(f"oracle on {nthreads} threads: assembly {t_local + t_sparse:.2f} s, PCG to {rtol64:g} {t_solve:.2f} s "
        f"({it1} iterations, status {s1})")
    s2 = L.or_pcg_dirichlet(A.n, O._i(A.colptr), O._i(A.rowval), O._d(A.nzval), O._u(mask), O._d(g), 1e-13,
                            1000000, O._d(phi), O.ctypes.byref(it), O.ctypes.byref(rel), nthreads, 0)
    it2, rel2 = int(it.value), float(rel.value)
    _H, Hn_ref = O.field(m, phi, nthreads=nthreads)
    E_ref = O.energy(m, Hn_ref, pr.mu)
    log(f"oracle reference: PCG continued to 1e-13 ({it2} more iterations, relres {rel2:.1e}, status {s2})")
    # one thread: element matrices of 1/16 of the elements, 10 PCG iterations
    k = max(1, m.nt // 16)
    t_sub = np.ascontiguousarray(np.asarray(m.t, dtype=np.int32)[:, :k].T)
    Ak1 = np.empty((k, 16))
    t0 = time.perf_counter()
    L.or_local_stiffness(O._d(O._p(m)), O._i(t_sub), k, O._d(np.ascontiguousarray(m.VE[:k])),
                         O._d(np.ascontiguousarray(pr.mu[:k])), O._d(Ak1), 1)
    t_local1 = (time.perf_counter() - t0) * (m.nt / k)
    phi1 = np.zeros(m.nv)
    t0 = time.perf_counter()
    L.or_pcg_iterations(A.n, O._i(A.colptr), O._i(A.rowval), O._d(A.nzval), O._u(mask), O._d(g), 10, O._d(phi1), 1)
    t_it1 = (time.perf_counter() - t0) / 10
    t_1thread = t_local1 + t_sparse + t_it1 * it1
    cpu = {
        "value": m.nv / (t_local + t_sparse + t_solve),
        "unit": "DOF/s",
        "cores": nthreads,
        "kind": "port",
        "sample": (f"oracle C restatement (OpenMP, {nthreads} threads) on the same n={pr.meta['n']} mesh: full "
                   f"assembly {t_local + t_sparse:.2f} s (element matrices {t_local:.2f} s + serial sparse() build "
                   f"{t_sparse:.2f} s) + full Jacobi-PCG to rtol {rtol64:g} {t_solve:.2f} s ({it1} iterations), "
                   f"measured end to end"),
        "t_assembly_s": t_local + t_sparse,
        "t_solve_s": t_solve,
        "iterations": it1,
        "host": {"cpu_count": os.cpu_count(), "model": _cpu_model()},
        "single_thread": {"value": m.nv / t_1thread, "unit": "DOF/s", "cores": 1,
                          "sample": (f"element matrices of {k} elements ({t_local1 * k / m.nt:.2f} s, scaled "
                                     f"x{m.nt / k:.0f}) + the serial sparse() build + 10 PCG iterations "
                                     f"({t_it1 * 1e3:.0f} ms each) scaled to {it1}")},
    }
    ref = {"phi": phi, "Hn": Hn_ref, "energy": E_ref, "iterations": it1 + it2, "relres": rel2}
    return cpu, ref


def _pmc_traffic(args, solver_format):
    """HBM bytes per FP32 SpMV launch from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE + WRITE_SIZE, scaled by what the
    same counters read for the plain copy k_copy16 of known traffic), when
    one exists for this workload, format and launch shape."""
    tag = "" if args.renumber == "none" else "-" + args.renumber
    path = os.path.join(ROOT, "profiles", f"pmc_spmv_f32_boxes160{tag}.txt")
    if args.n != 160 or not os.path.exists(path):
        return None
    vals = {}
    with open(path) as fh:
        for line in fh:
            if ":" in line and line.split(":")[0].strip() in ("traffic_bytes_per_launch", "format", "shape"):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    if vals.get("format") != solver_format or "traffic_bytes_per_launch" not in vals:
        return None
    if solver_format == "sell64" and vals.get("shape") != args.shape:
        return None
    return float(vals["traffic_bytes_per_launch"])


def _common(args, metric, value, unit, ms, dtype, config, higher=True):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


def bench_batch(args):
    """64 Kuhn meshes (n ~ U{10..60}) x 3 applied fields, FP32 PCG with FP64
    reliable updates to rtol 1e-10, all 192 systems in one set of launches
    (the restarted refinement, FP32 PCG to 1e-5 inside an FP64 loop, is
    measured beside it)."""
    from femeko_hip import workloads as W
    from femeko_hip.batch import Batch
    from femeko_hip.workloads import dirichlet_values
    t0 = time.perf_counter()
    probs = W.batch()
    log(f"batch: {len(probs)} meshes, {sum(p.mesh.nv for p in probs)} nodes ({time.perf_counter()-t0:.1f} s)")
    b = Batch([p.mesh for p in probs], nrhs=3)
    b.set_format(args.format if args.format in ("auto", "stencil") else "csr")
    b.set_coefficient([p.mu for p in probs])
    b.set_dirichlet([p.mask for p in probs],
                    [np.stack([np.where(p.mask == 1, dirichlet_values(p.mesh, e), 0.0) for e in np.eye(3)])
                     for p in probs])
    b.assemble()
    lanes, block, us = b.autotune()
    if args.batch_update_block:
        b.set_launch(lanes, args.batch_update_block)
    fmt = b.solver_format()
    launch = b.launch()
    log(f"autotuner: SpMM {fmt} {launch} ({us:.1f} us per SpMM)")

    def step():
        b.assemble()
        return b.solve(rtol=1e-10, inner_rtol=args.delta)

    for _ in range(args.warmup):
        rep = step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rep = step()
    t = (time.perf_counter() - t0) / args.steps
    nodes = b.total_nodes
    dof_solved = 3 * nodes
    log(f"batch step {t*1e3:.1f} ms; iterations max {rep.iterations.max()} mean {rep.iterations.mean():.0f}")
    out = _common(args, "batched assembly+CG DOF/s (64 meshes x 3 fields, FP32-refined)", dof_solved / t,
                  "DOF/s", t * 1e3, "f32+f64",
                  {"workload": "batch64: 64 Kuhn meshes n~U{10..60}, one mu box each, H0 = x, y, z",
                   "meshes": len(probs), "systems": 3 * len(probs), "nodes": nodes, "nnz": b.total_nnz,
                   "solver_format": fmt, "autotuned_launch": launch})
    out["per_solve_ms"] = t * 1e3 / (3 * len(probs))
    out["iterations_max"] = int(rep.iterations.max())
    out["iterations_mean"] = float(rep.iterations.mean())
    out["relres_max"] = float(rep.relres.max())
    # the classic restarted refinement on the same batch, one step
    b.set_precision("f32-restart")
    b.assemble()
    t0 = time.perf_counter()
    rs = b.solve(rtol=1e-10, inner_rtol=1e-5)
    t_rs = time.perf_counter() - t0
    b.set_precision("f32")
    log(f"batch restarted refinement: {t_rs*1e3:.1f} ms; iterations max {rs.iterations.max()}")
    out["f32_restart_inner1e-5"] = {"ms_per_step": t_rs * 1e3, "iterations_max": int(rs.iterations.max()),
                                    "iterations_mean": float(rs.iterations.mean()),
                                    "relres_max": float(rs.relres.max())}
    b.close()
    return out


def bench_froehlich(args):
    """n=100 Froehlich magnet, Picard/Newton mu loop on the GPU, update solves
    to 1e-10 in FP64 PCG or (--inner-precision f32) FP32 PCG with FP64 residual
    replacement."""
    from femeko_hip import workloads as W
    from femeko_hip.device import DeviceMesh, context
    pr = W.froehlich()
    dm = DeviceMesh(pr.mesh, context(0))
    dm.set_dirichlet(pr.mask, pr.g)
    dm.set_inner_precision(args.inner_precision)
    reg = pr.magnet.astype(np.uint8)
    for _ in range(args.warmup):
        out = dm.nonlinear_froehlich(reg, pr.meta["Ms"], pr.meta["a"])
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = dm.nonlinear_froehlich(reg, pr.meta["Ms"], pr.meta["a"])
    t = (time.perf_counter() - t0) / args.steps
    log(f"froehlich: {t*1e3:.1f} ms, outer {out['outer']}, CG {out['cg_iterations']}, change {out['change']:.2e}")
    prec = "FP64" if args.inner_precision == "f64" else "FP32-refined update solves"
    res = _common(args, f"nonlinear magnetostatics DOF/s (Picard/Newton mu loop, {prec})", pr.mesh.nv / t, "DOF/s",
                  t * 1e3, "f64" if args.inner_precision == "f64" else "f32+f64",
                  {"workload": "froehlich100: Kuhn n=100, Froehlich mu_r = 1 + 1.2e6/(500+|H|) in [-0.25,0.25]^3, "
                               "|H0| = 1e4", "nodes": pr.mesh.nv, "tets": pr.mesh.nt})
    res.update({"outer_iterations": out["outer"], "cg_iterations": out["cg_iterations"],
                "final_mu_change": out["change"]})
    if args.inner_precision == "f32":
        _loop_vs_f64(dm, lambda: dm.nonlinear_froehlich(reg, pr.meta["Ms"], pr.meta["a"]), res)
    dm.close()
    return res


def _loop_vs_f64(dm, run, res):
    """After the timed region: the same loop run twice with FP64 update
    solves (the second timed), and the relative 2-norm difference of the
    potentials."""
    phi = dm.potential()
    dm.set_inner_precision("f64")
    run()
    t0 = time.perf_counter()
    run()
    t = time.perf_counter() - t0
    ref = dm.potential()
    d = float(np.linalg.norm(phi - ref) / np.linalg.norm(ref))
    log(f"FP64 update solves: {t*1e3:.1f} ms; potential difference {d:.2e}")
    res.update({"f64_loop_ms": t * 1e3, "speedup_vs_f64_loop": t * 1e3 / res["ms_per_step"],
                "phi_rel_diff_vs_f64_loop": d})


def bench_tabulated(args):
    """n=100 gadolinium cube (B(H) table at 270 K, the reference's material)
    under 0.5 T, mu(|H|) loop with the tabulated law, update solves to 1e-10
    (FP32-refined or FP64, --inner-precision), stop at mu0 max| |H| - |H|_prev | < 1e-4 T (NewtonRaphson.jl:36). Picard
    steps until that change is below 1e-2 T, Newton steps after: undamped
    Newton from the second step (the reference damps it with lineSearch)
    drives a few cube-corner elements, where the field is singular, past the
    table's end at this resolution."""
    from femeko_hip import workloads as W
    from femeko_hip.device import DeviceMesh, context
    from femeko_hip.material import Spline1D, material_permeability
    tab = np.load(os.path.join(ROOT, "tests", "golden", "gd270_bh.npz"))
    mu_t, dmu_t, _nu = material_permeability(tab["H"], tab["B"])
    spl, spl_dmu = Spline1D(tab["H"], mu_t), Spline1D(tab["H"], dmu_t)
    pr = W.tabulated()
    dm = DeviceMesh(pr.mesh, context(0))
    dm.set_dirichlet(pr.mask, pr.g)
    dm.set_inner_precision(args.inner_precision)
    reg = pr.magnet.astype(np.uint8)
    for _ in range(args.warmup):
        out = dm.nonlinear_spline(reg, spl, spl_dmu, rtol=1e-10, tol_H=1e-4, switch_tol=1e-2)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = dm.nonlinear_spline(reg, spl, spl_dmu, rtol=1e-10, tol_H=1e-4, switch_tol=1e-2)
    t = (time.perf_counter() - t0) / args.steps
    log(f"tabulated: {t*1e3:.1f} ms, outer {out['outer']}, CG {out['cg_iterations']}, change {out['change']:.2e}")
    prec = "FP64" if args.inner_precision == "f64" else "FP32-refined update solves"
    res = _common(args, f"nonlinear magnetostatics DOF/s (tabulated mu(|H|) Newton loop, {prec})", pr.mesh.nv / t,
                  "DOF/s", t * 1e3, "f64" if args.inner_precision == "f64" else "f32+f64",
                  {"workload": "tabulated100: Kuhn n=100, Gd B(H) at 270 K (62-point table, cubic splines of mu "
                               "and dmu) in [-0.25,0.25]^3, Bext = 0.5 T along x; Newton once the change "
                               "is below 1e-2 T", "nodes": pr.mesh.nv,
                   "tets": pr.mesh.nt})
    res.update({"outer_iterations": out["outer"], "cg_iterations": out["cg_iterations"],
                "final_H_change_T": out["change"]})
    if args.inner_precision == "f32":
        _loop_vs_f64(dm, lambda: dm.nonlinear_spline(reg, spl, spl_dmu, rtol=1e-10, tol_H=1e-4, switch_tol=1e-2),
                     res)
    dm.close()
    return res


def bench_demag(args):
    """Repeated-RHS Poisson solves of the micromagnetics driver
    (Micromagnetics/LL.jl:162-209): n=100 Kuhn box, magnet [-0.25,0.25]^3
    with a unit magnetization that rotates a little each step, A =
    stiffnessMatrix(mesh) resident, one fk_demag_field per step at the
    reference cg's tolerance sqrt(eps). A step = load vector + Neumann PCG
    (FP32-refined or FP64, --inner-precision) + nodal field, host copies of M
    and Hd included."""
    from femeko_hip import workloads as W
    from femeko_hip.device import DeviceMesh, context
    pr = W.linear(100, seed=102)
    m = pr.mesh
    inside = pr.magnet
    t0i = np.asarray(m.t, dtype=np.int64) - 1
    vol = np.zeros(m.nv)
    np.add.at(vol, t0i[:, inside].ravel(), np.repeat(m.VE[inside][None, :], 4, axis=0).ravel())
    dm = DeviceMesh(m, context(0))
    dm.set_coefficient(np.ones(m.nt))
    dm.assemble()
    dm.set_inner_precision(args.inner_precision)
    ins = inside.astype(np.uint8)
    innodes = vol > 0

    def M_at(k):
        # 3 x nv column-major (x, y, z of a node together), the layout the
        # library and Julia's Matrix take, so the wrapper passes it as is
        th = 0.05 * k
        M = np.zeros((3, m.nv), order="F")
        M[0, innodes] = np.cos(th)
        M[1, innodes] = np.sin(th)
        return M

    # the magnetizations of all steps made before the timed region: the step
    # is the library call (host copies of M and Hd included)
    Ms = [M_at(k) for k in range(args.warmup + args.steps + 1)]

    def run(warm):
        for k in range(args.warmup):
            dm.demag_field(ins, Ms[k], vol, warm=warm)
        its = []
        t0 = time.perf_counter()
        for k in range(args.steps):
            _Hd, it, _rel = dm.demag_field(ins, Ms[args.warmup + k], vol, warm=warm)
            its.append(it)
        return (time.perf_counter() - t0) / args.steps, its

    t_cold, it_cold = run(False)
    t, iters = run(True)
    log(f"demag: {t*1e3:.1f} ms per step starting from the previous potential (PCG iterations {iters}); "
        f"{t_cold*1e3:.1f} ms from zero as the reference's cg ({it_cold})")
    f32 = args.inner_precision == "f32"
    prec = "FP32-refined" if f32 else "FP64"
    res = _common(args, f"demagnetizing-field solves per second (fixed matrix, new right-hand side each step, {prec})",
                  1.0 / t, "solves/s", t * 1e3, "f32+f64" if f32 else "f64",
                  {"workload": "demag100: Kuhn n=100, magnet [-0.25,0.25]^3, unit M rotating 0.05 rad per step, "
                               "rtol sqrt(eps), PCG started from the previous step's potential", "nodes": m.nv,
                   "tets": m.nt})
    res.update({"pcg_iterations": iters, "from_zero": {"ms_per_step": t_cold * 1e3, "pcg_iterations": it_cold}})
    if f32:
        # after the timed region: the same steps from zero in both precisions
        # at rtol 1e-12, the fields compared, and the FP64 path timed
        k = args.warmup + args.steps
        H32, _i, _r = dm.demag_field(ins, Ms[k], vol, rtol=1e-12)
        dm.set_inner_precision("f64")
        t64, it64 = run(True)
        H64, _i, _r = dm.demag_field(ins, Ms[k], vol, rtol=1e-12)
        ok = ~np.isnan(H64)
        d = float(np.abs(H32[ok] - H64[ok]).max() / np.abs(H64[ok]).max())
        log(f"FP64 Neumann solves: {t64*1e3:.1f} ms per step ({it64}); field difference at rtol 1e-12 {d:.2e}")
        res.update({"f64_ms_per_step": t64 * 1e3, "speedup_vs_f64": t64 / t, "f64_pcg_iterations": it64,
                    "field_max_diff_vs_f64_at_1e-12": d})
    dm.close()
    return res


def bench_adjoint(args):
    """n=80 adjoint gradient of the field-target objective, state and adjoint
    solves FP32-refined or FP64 (--inner-precision); central differences
    (step 1e-6, FP64 solves to 1e-13) on 32 seeded elements."""
    from femeko_hip import adjoint as ADJ
    from femeko_hip import fem3d
    from femeko_hip import workloads as W
    pr = W.adjoint()
    m = pr.mesh
    rho = pr.meta["rho"]
    tg = pr.magnet.astype(np.uint8)
    Hs = pr.meta["Hstar"]
    prec = args.inner_precision

    def timed(precision):
        out = ADJ.fieldObjectiveGradient(m, rho, tg, Hs, pr.mask, pr.g, precision=precision)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = ADJ.fieldObjectiveGradient(m, rho, tg, Hs, pr.mask, pr.g, precision=precision)
        return (time.perf_counter() - t0) / args.steps, out

    t, (J, grad, its) = timed(prec)
    # 32 seeded elements: the 16 largest gradients and 16 uniform draws; the
    # differences use FP64 solves to rtol 1e-13 so their noise stays far
    # below the step's truncation error
    rng = np.random.default_rng(80)
    idx = np.concatenate([np.argsort(-np.abs(grad))[:16], rng.choice(m.nt, 16, replace=False)])
    h = 1e-6
    gmax = float(np.abs(grad).max())
    err_scaled, err_rel_large = [], []
    for e in idx:
        rp = rho.copy()
        rp[e] += h
        rm = rho.copy()
        rm[e] -= h
        fd = (ADJ.fieldObjective(m, rp, tg, Hs, pr.mask, pr.g, rtol=1e-13)
              - ADJ.fieldObjective(m, rm, tg, Hs, pr.mask, pr.g, rtol=1e-13)) / (2 * h)
        err_scaled.append(abs(fd - grad[e]) / gmax)
        if abs(grad[e]) >= 1e-2 * gmax:  # below, the difference's rounding noise dominates
            err_rel_large.append(abs(fd - grad[e]) / abs(grad[e]))
    log(f"adjoint: {t*1e3:.1f} ms per state+adjoint+gradient; FD max |fd-g|/max|g| {max(err_scaled):.2e}, "
        f"max rel err where |g| >= 1e-2 max|g| {max(err_rel_large):.2e} ({len(err_rel_large)} elements)")
    if prec == "f32":
        # after the timed region: the FP64 path timed, its gradient compared
        t64, (J64, grad64, its64) = timed("f64")
        dg = float(np.abs(grad - grad64).max() / gmax)
        log(f"FP64 state and adjoint solves: {t64*1e3:.1f} ms; gradient difference {dg:.2e} of max|g|")
    res = _common(args, "state+adjoint+gradient DOF/s (" + ("FP32-refined" if prec == "f32" else "FP64") + " solves)",
                  m.nv / t, "DOF/s", t * 1e3, "f32+f64" if prec == "f32" else "f64",
                  {"workload": "adjoint80: Kuhn n=80, mu_r = 1 + 999 rho^3, target [-0.1,0.1]^3, H* = (2,0,0)",
                   "nodes": m.nv, "tets": m.nt})
    if prec == "f32":
        res.update({"f64_ms": t64 * 1e3, "speedup_vs_f64": t64 / t, "grad_max_diff_vs_f64_over_max_grad": dg,
                    "J_rel_diff_vs_f64": abs(J - J64) / abs(J64), "f64_adjoint_iterations": its64})
    res.update({"J": J, "adjoint_iterations": its, "fd_elements": int(len(idx)), "fd_step": h,
                "fd_max_err_over_max_grad": float(max(err_scaled)),
                "fd_max_rel_err_large_grad": float(max(err_rel_large)),
                "fd_large_grad_elements": len(err_rel_large)})
    fem3d.invalidate(m)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=160)
    ap.add_argument("--lanes", type=int, default=16)
    ap.add_argument("--format", default="auto", choices=["auto", "stencil", "csr", "sell64", "sell64-rowidx"])
    ap.add_argument("--rtol", type=float, default=5e-11,
                    help="relative FP64 residual the FP32-refined solve stops at (5e-11: its phi error then "
                         "stays at or below the FP64 path's at 1e-10 on every seed measured; at 1e-10 it "
                         "ranged 1.4e-10 to 1.06e-9)")
    ap.add_argument("--rtol64", type=float, default=1e-10,
                    help="relative residual the plain FP64 solve stops at (the configured 1e-10)")
    ap.add_argument("--delta", type=float, default=0.95,
                    help="FP32 solve: residual drop that triggers an FP64 residual replacement")
    ap.add_argument("--shape", default="256,2",
                    help="FP32 SELL SpMV launch shape: threads per block, slices in flight per wave")
    ap.add_argument("--renumber", default="none", choices=["none", "rcm", "random"],
                    help="renumber the nodes (reverse Cuthill-McKee or random) to measure a general, "
                         "non-stencil numbering of the same mesh")
    ap.add_argument("--reorder", default="auto", choices=["auto", "none", "rcm"],
                    help="the library's internal node renumbering (fk_mesh_create_ex)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--inner-precision", default="f32", choices=["f64", "f32"],
                    help="froehlich, tabulated, demag and adjoint workloads: precision of their inner solves")
    ap.add_argument("--batch-update-block", type=int, default=0, choices=[0, 128, 256, 512],
                    help="batch workload: the update kernels' block size instead of the autotuner's")
    ap.add_argument("--workload", default="boxes",
                    choices=["boxes", "batch", "froehlich", "tabulated", "demag", "adjoint"],
                    help="boxes: the headline n=160 mixed-precision line (default); the others measure the "
                         "batched, nonlinear and adjoint configurations")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the multi-GPU path is a design "
              f"note only)", file=sys.stderr)
        sys.exit(2)
    if args.workload != "boxes":
        out = {"batch": bench_batch, "froehlich": bench_froehlich, "tabulated": bench_tabulated,
               "demag": bench_demag, "adjoint": bench_adjoint}[args.workload](args)
        print(json.dumps(out))
        return

    from femeko_hip import workloads as W
    from femeko_hip.device import DeviceMesh, context

    t0 = time.perf_counter()
    pr = W.boxes(n=args.n) if args.n != 160 else W.boxes()
    ctx = context(0)
    if args.renumber != "none":
        # a general numbering of the same mesh: the pattern is no stencil and
        # the solver uses its SELL copy
        if args.renumber == "random":
            perm = np.random.default_rng(args.n).permutation(pr.mesh.nv)
        else:
            from femeko_hip.mesh import rcm_permutation
            perm = rcm_permutation(pr.mesh)
        pr = W.renumbered(pr, perm, args.renumber)
        log(f"renumbered nodes ({args.renumber})")
    m = pr.mesh
    log(f"mesh n={args.n}: nt={m.nt} nv={m.nv} ({time.perf_counter()-t0:.1f} s)")
    t0 = time.perf_counter()
    dm = DeviceMesh(m, ctx, reorder=args.reorder)
    dm.set_lanes(args.lanes)
    dm.set_format(args.format)
    shape = tuple(int(v) for v in args.shape.split(","))
    dm.set_sell_shape(*shape)
    log(f"device mesh: nnz={dm.nnz} colours={dm.ncolours} ({time.perf_counter()-t0:.1f} s)")
    renumbered, dist_caller, dist_internal = dm.numbering()
    log(f"node numbering: {'renumbered internally (RCM)' if renumbered else 'caller order kept'}; "
        f"mean neighbour distance {dist_caller:.0f} (caller), {dist_internal:.0f} (internal)")
    fmt_name, stencil_offsets = dm.solver_format()
    sell = dm.sell_info() if fmt_name == "sell64" else None
    sell_index = sell_groups = sell_array_bytes = None
    if sell is not None:
        # int16 offsets from the row or int32 columns, and the batch width, as
        # the library chose them
        ib, sell_groups, sell_array_bytes = dm.sell_layout()
        sell_index = {2: "short", 4: "int"}[ib]
    log(f"solver format: {fmt_name} {stencil_offsets if stencil_offsets else ''}")
    if sell:
        log(f"SELL copy: {sell[0]} slices, {sell[1]} with shared offsets, {sell[2]} stored entries")
    dm.set_coefficient(pr.mu)
    dm.set_dirichlet(pr.mask, pr.g)
    free = int((pr.mask == 0).sum())

    def step(precision):
        dm.assemble()
        if precision == "f32":
            return dm.solve("f32", rtol=args.rtol, inner_rtol=args.delta)
        return dm.solve("f64", rtol=args.rtol64)

    # ---- headline: FP32-refined
    for _ in range(args.warmup):
        rep = step("f32")
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rep = step("f32")
    ctx.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    log(f"f32-refined step {t_step*1e3:.1f} ms, iterations {rep.iterations}, outer {rep.outer}, "
        f"relres {rep.relres:.2e}")
    phi32 = dm.potential()
    _H, Hn32, E32 = dm.field(want_H=False, want_Hn=True, want_energy=True)
    # assembly alone
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dm.assemble()
    ctx.synchronize()
    t_asm = (time.perf_counter() - t0) / args.steps

    # ---- plain FP64 GPU path of the same system
    rep64 = step("f64")
    ctx.synchronize()
    t0 = time.perf_counter()
    nrep64 = max(1, min(args.steps, 2))
    for _ in range(nrep64):
        rep64 = step("f64")
    ctx.synchronize()
    t_step64 = (time.perf_counter() - t0) / nrep64
    log(f"f64 step {t_step64*1e3:.1f} ms, iterations {rep64.iterations}, relres {rep64.relres:.2e}")
    phi64 = dm.potential()
    _H, Hn64, E64 = dm.field(want_H=False, want_Hn=True, want_energy=True)
    if args.no_cpu:
        # no oracle run: the error reference is the GPU's own FP64 solve at 1e-13
        dm.solve("f64", rtol=1e-13)
        _H, Hn_g, E_g = dm.field(want_H=False, want_Hn=True, want_energy=True)
        gpu_ref = {"phi": dm.potential(), "Hn": Hn_g, "energy": E_g}

    # ---- practical streaming ceiling: plain 1 GiB device copy
    copy_ms = ctx.copy_probe(1 << 30, 5)
    copy_gbs = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    log(f"copy probe: 1 GiB in {copy_ms:.3f} ms, {copy_gbs:.0f} GB/s read+write")

    # ---- SpMV kernel timing: HIP events around every SpMV launch of one instrumented solve
    ctx.set_timing(True)
    dm.assemble()
    t0 = time.perf_counter()
    rep_t = dm.solve("f32", rtol=args.rtol, inner_rtol=args.delta)
    ctx.synchronize()
    t_instr = time.perf_counter() - t0
    launches, total_ms = ctx.spmv_timing()
    ctx.set_timing(False)
    spmv_us_in_solve = total_ms / max(launches, 1) * 1e3
    # the same kernel back to back on the solver's data: its duration without
    # the dispatch gap that an event pair around every launch adds
    spmv_us = dm.spmv_burst("f32", 100)
    B_csr = spmv_bytes(dm.nnz, dm.nv, 4)
    # the compulsory bytes of the copy the solver streams
    # the format's own compulsory bytes: the stencil copy's columns, or the
    # SELL arrays (values with padding, indices, slice tables), plus x and y
    if fmt_name == "stencil":
        B32 = stencil_spmv_bytes(len(stencil_offsets), dm.nv, 4)
    elif fmt_name == "sell64":
        B32 = sell_array_bytes + 8 * dm.nv
    else:
        B32 = B_csr
    achieved = B32 / (spmv_us * 1e-6) / 1e9
    log(f"SpMV f32: {spmv_us:.1f} us per launch back to back, {achieved:.0f} GB/s; "
        f"{spmv_us_in_solve:.1f} us avg over {launches} launches in a solve with events around each")

    out = {
        "metric": METRIC,
        "value": m.nv / t_step,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": (f"boxes{args.n}{'' if args.renumber == 'none' else '-' + args.renumber}: "
                         f"Kuhn box mesh n={args.n}, jitter 0.15h, 8 log-uniform mu_r boxes in "
                         f"[1,1e3], Dirichlet phi=-x; step = value assembly + FP32 Jacobi-PCG whose residual is "
                         f"replaced by the FP64 true residual after every {args.delta:g}x drop, to FP64 "
                         f"relres {args.rtol:g} (FP64 comparison path to {args.rtol64:g})"),
            "nodes": m.nv, "tets": m.nt, "nnz": dm.nnz, "free_dofs": free, "colours": dm.ncolours,
            "lanes_per_row": args.lanes, "spmv_format": fmt_name, "stencil_offsets": list(stencil_offsets),
            "sell_shape": list(shape),
            "internal_renumbering": renumbered,
            "sell_slices": sell[0] if sell else None, "sell_shared_slices": sell[1] if sell else None,
            "sell_entries": sell[2] if sell else None,
            "update_trigger": args.delta,
        },
        "iterations_f32": rep.iterations,
        "outer_refinements": rep.outer,
        "relres": rep.relres,
        "ms_assembly": t_asm * 1e3,
        # value assembly against its compulsory FP64 bytes (t 16 nt, mu 8 nt,
        # coordinates 24 nv, values 8 nnz written)
        "assembly_roofline": {"bound": "hbm", "kernel": "k_assemble_gather",
                              "bytes": 16 * m.nt + 8 * m.nt + 24 * m.nv + 8 * dm.nnz,
                              "achieved": (16 * m.nt + 8 * m.nt + 24 * m.nv + 8 * dm.nnz) / t_asm / 1e9,
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": (16 * m.nt + 8 * m.nt + 24 * m.nv + 8 * dm.nnz) / t_asm / 1e9 / HBM_PEAK_GBS},
        "fp64": {"ms_per_step": t_step64 * 1e3, "iterations": rep64.iterations, "relres": rep64.relres,
                 "dof_per_s": m.nv / t_step64},
        "speedup_vs_fp64": t_step64 / t_step,
        "roofline": {
            "bound": "hbm",
            "kernel": {"csr": "k_cg_spmv<float,%d>" % args.lanes,
                       "sell64": "k_cg_spmv_sell<float,%s,%d,%d,%s>" % ((sell_index,) + shape + (sell_groups,)),
                       "stencil": "k_cg_spmv_stencil<float,256,%d>" % len(stencil_offsets)}[fmt_name],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_launch": B32,
            "bytes_basis": {"stencil": "stencil copy: K*4*nv values + 4*nv x + 4*nv y",
                            "sell64": "SELL copy: 4 B per stored value (padding included) + its indices, "
                                      "shared offsets and slice tables + 4*nv x + 4*nv y",
                            "csr": "CSR: nnz*(4+4) + 4*(nv+1) + 8*nv"}[fmt_name],
            "csr_formula_bytes": B_csr,
            "csr_formula_GBs": B_csr / (spmv_us * 1e-6) / 1e9,
            "copy_probe_GBs": copy_gbs,
            "avg_launch_us": spmv_us,
            "avg_launch_us_basis": "100 back-to-back launches on the solver's data between two HIP events",
            "avg_launch_us_in_solve": spmv_us_in_solve,
            "launches_in_solve": launches,
            "instrumented_solve_ms": t_instr * 1e3,
        },
    }
    # ---- classic restarted refinement (FP32 PCG to inner rtol 1e-5, FP64 residual correction)
    dm.assemble()
    ctx.synchronize()
    t0 = time.perf_counter()
    rep_rs = dm.solve("f32-restart", rtol=1e-10, inner_rtol=1e-5)
    ctx.synchronize()
    t_rs = time.perf_counter() - t0
    log(f"f32-restart step {t_rs*1e3:.1f} ms, iterations {rep_rs.iterations}, outer {rep_rs.outer}")
    out["f32_restart_inner1e-5"] = {"ms_per_step": t_rs * 1e3, "iterations": rep_rs.iterations,
                                    "outer": rep_rs.outer, "relres": rep_rs.relres}
    traffic = _pmc_traffic(args, fmt_name)
    if traffic is not None:
        # HBM bytes per launch from the committed counter profile, and the
        # bandwidth they imply at this run's launch time
        out["roofline"]["traffic"] = traffic
        out["roofline"]["traffic_GBs"] = traffic / (spmv_us * 1e-6) / 1e9
        out["roofline"]["traffic_frac"] = out["roofline"]["traffic_GBs"] / HBM_PEAK_GBS
    dm.close()
    if not args.no_cpu:
        nthreads = min(16, os.cpu_count() or 1)
        log(f"oracle legs with {nthreads} threads ...")
        out["cpu_baseline"], ref = oracle_legs(pr, nthreads, args.rtol64)
        basis = f"oracle FP64 Jacobi-PCG on the host to relres {ref['relres']:.1e} ({ref['iterations']} iterations)"
        log(f"CPU baseline {out['cpu_baseline']['value']:.3e} DOF/s")
    else:
        out["cpu_baseline"] = None
        ref = gpu_ref
        basis = "the GPU's own FP64 solve to rtol 1e-13 (no oracle run)"
    VE = m.VE

    def errors(phi, Hn, E):
        return {"phi_rel_l2": float(np.linalg.norm(phi - ref["phi"]) / np.linalg.norm(ref["phi"])),
                "Hn_rel_l2_VE": float(np.sqrt(np.sum(VE * (Hn - ref["Hn"]) ** 2) / np.sum(VE * ref["Hn"] ** 2))),
                "energy_rel": float(abs(E - ref["energy"]) / abs(ref["energy"]))}

    out["error_vs_oracle"] = {"basis": basis, "f32_refined": errors(phi32, Hn32, E32),
                              "fp64": errors(phi64, Hn64, E64)}
    e = out["error_vs_oracle"]["f32_refined"]
    log(f"errors of the FP32-refined solve vs {basis}: phi {e['phi_rel_l2']:.2e} |H| {e['Hn_rel_l2_VE']:.2e} "
        f"energy {e['energy_rel']:.2e}; FP64 path phi {out['error_vs_oracle']['fp64']['phi_rel_l2']:.2e}")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
