/*
 * femeko_hip.h -- C ABI of the MI355X (gfx950) P1 magnetostatics library.
 *
 * Drop-in native boundary for Femeko.jl's 3D P1 path: Julia (or any C FFI)
 * keeps its MESH and drivers and calls these functions where the drivers call
 * stiffnessMatrix / localStiffnessMatrix / tangentialStiffnessMatrix, the
 * sparse solve and the H-field loop. The model is the existing reference
 * bridge `scalarPotential` (Femeko_jl_C++/Magnetostatics/magnetostatics.cpp:
 * 49-60, bound by @ccall in Femeko_jl_C++/Magnetostatics/run.jl:66-79):
 * plain pointers and sizes, caller-owned host buffers in and out.
 *
 * Conventions
 *   - Every function returns an int status (FK_OK == 0). Nothing throws
 *     across this ABI (the reference throws C++ exceptions through extern "C",
 *     magnetostatics.cpp:29-30,36-37). fk_context_last_error() explains the
 *     last failure on a context.
 *   - Host arrays use Femeko's layouts: p is 3 x nv column-major Float64,
 *     t is 4 x nt column-major Int32 (src/mesh.jl:9-59). Node numbers may be
 *     1-based (Julia, index_base = 1) or 0-based (index_base = 0); the
 *     library never needs a shifted copy (the reference makes one,
 *     run.jl:50-52).
 *   - Per-element arrays (f, mu, H, |H|, gradients) are in the caller's
 *     element order. Per-node arrays are in the caller's node order.
 *   - Device memory is owned by the handles; the mesh, coefficients, matrix
 *     values, Dirichlet data and potential stay resident on the GPU between
 *     calls.
 *   - A context is not re-entrant: call it from one thread at a time.
 *     Different contexts may be used concurrently.
 */
#ifndef FEMEKO_HIP_H
#define FEMEKO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FK_OK 0
#define FK_ERR_INVALID_ARG 1   /* null pointer, bad size, bad enum value */
#define FK_ERR_HIP 2           /* a HIP runtime call failed */
#define FK_ERR_DEGENERATE 3    /* zero-volume / non-finite element, repeated node */
#define FK_ERR_INDEX 4         /* node number out of range */
#define FK_ERR_NONFINITE 5     /* NaN/Inf coefficient or field (NewtonRaphson.jl:210-212) */
#define FK_ERR_NOT_CONVERGED 6 /* iteration limit reached before the tolerance */
#define FK_ERR_SINGULAR 7      /* free node without stiffness, CG breakdown, no Dirichlet node */
#define FK_ERR_OOM 8           /* device or host allocation failed */
#define FK_ERR_STATE 9         /* call order: e.g. solve before assemble */
#define FK_ERR_UNSUPPORTED 10  /* mesh beyond the library's limits */
#define FK_ERR_DOMAIN 11       /* argument outside a table's range (Spline1D bc = "error") */

/* precision of the linear solve */
#define FK_PREC_F64 1          /* FP64 matrix, FP64 Jacobi-PCG */
#define FK_PREC_F32_REFINED 2  /* FP32 matrix and PCG, FP64 true residual replaced in every
                                  iteration chunk, FP64 solution (Krylov space kept) */
#define FK_PREC_F32_RESTART 3  /* FP32 PCG to inner_rtol restarted inside an FP64
                                  residual-correction loop (classic iterative refinement) */

/* method of the element assembly */
#define FK_ASM_COLOURED 0      /* mesh-coloured conflict-free adds (deterministic) */
#define FK_ASM_ATOMIC 1        /* hardware FP64 atomics (measured alternative) */
#define FK_ASM_GATHER 2        /* one thread per row over the node's elements, values summed in
                                  LDS and written once (deterministic, exactly symmetric; falls
                                  back to FK_ASM_COLOURED for rows beyond 256 entries) */

typedef struct fk_context fk_context;
typedef struct fk_mesh fk_mesh;

/* ---- context ------------------------------------------------------------ */
int fk_version(void);
const char *fk_status_string(int status);
/* Select GPU `device`, create the stream all work of this context runs on. */
int fk_context_create(int device, fk_context **out);
int fk_context_destroy(fk_context *ctx);
const char *fk_context_last_error(const fk_context *ctx);
int fk_context_synchronize(fk_context *ctx);

/* ---- mesh ----------------------------------------------------------------
 * Upload p (3 x nv) and t (4 x nt), colour the elements, build the CSR
 * pattern and the element -> CSR-slot map on the GPU. Replaces the mesh data
 * every FEM3D.jl P1 function reads (src/mesh.jl:9-59) and the fixed-sparsity
 * COO index build of Magnetostatics/NewtonRaphson.jl:115-127.
 * VE may be NULL; if given it is checked against the recomputed volumes
 * (elementVolume, src/mesh.jl:484-505) and a mismatch above 1e-10 relative
 * returns FK_ERR_INVALID_ARG. */
int fk_mesh_create(fk_context *ctx, const double *p, int64_t nv, const int32_t *t, int64_t nt,
                   int32_t index_base, const double *VE, fk_mesh **out);
/* Node numbering inside the library. The solver's matrix copies and the
 * gather kernels read node data at the columns of each row, so their speed
 * depends on how far apart neighbouring node numbers are; a mesh generator's
 * numbering need not keep them close. fk_mesh_create (FK_REORDER_AUTO)
 * renumbers the nodes internally by reverse Cuthill-McKee unless the
 * caller's numbering is a stencil numbering (a structured grid) or its mean
 * neighbour distance is within 2x of RCM's. Every per-node array of the
 * ABI stays in the caller's node order and every exported matrix in the
 * caller's numbering: the renumbering is not visible in any result.
 * FK_REORDER_NONE keeps the caller's numbering, FK_REORDER_RCM always
 * renumbers. No counterpart in the reference (Femeko solves in Gmsh's
 * numbering, src/mesh.jl:62-245). */
#define FK_REORDER_AUTO 0
#define FK_REORDER_NONE 1
#define FK_REORDER_RCM 2
int fk_mesh_create_ex(fk_context *ctx, const double *p, int64_t nv, const int32_t *t, int64_t nt,
                      int32_t index_base, const double *VE, int32_t reorder, fk_mesh **out);
/* Whether the mesh was renumbered, and the mean |i - j| over neighbouring
 * nodes (i, j) in the caller's numbering and in the internal one (both 0
 * when FK_REORDER_AUTO kept a stencil numbering without measuring). */
int fk_mesh_numbering(const fk_mesh *mesh, int32_t *reordered, double *mean_distance_caller,
                      double *mean_distance_internal);
/* The reverse Cuthill-McKee order fk_mesh_create uses, on the host (no
 * GPU): perm[k] = the 0-based node numbered k (nv entries). t as for
 * fk_mesh_create. */
int fk_rcm_order(const int32_t *t, int64_t nv, int64_t nt, int32_t index_base, int32_t *perm);
int fk_mesh_destroy(fk_mesh *mesh);
/* nnz of the pattern, number of element colours, nodes used by no element */
int fk_mesh_info(const fk_mesh *mesh, int64_t *nnz, int32_t *ncolours, int64_t *unused_nodes);
/* Element volumes |det|/6 (mesh.VE fill, src/mesh.jl:216-219). VE: nt. */
int fk_mesh_volumes(fk_mesh *mesh, double *VE);

/* ---- pattern and assembly ------------------------------------------------
 * CSR pattern of the stiffness matrix: the sorted unique node pairs of all
 * elements, explicit zeros kept -- exactly the structure of Julia's
 * sparse(rowIDs, colIDs, ...) in NewtonRaphson.jl:161. The pattern is
 * symmetric, so rowptr/colidx are also the CSC colptr/rowval of
 * SparseMatrixCSC. rowptr: nv+1, colidx: nnz. */
int fk_pattern(fk_mesh *mesh, int32_t index_base, int32_t *rowptr, int32_t *colidx);
/* Per-element coefficient f (mu), nt values, caller's element order.
 * Non-finite values return FK_ERR_NONFINITE. */
int fk_set_coefficient(fk_mesh *mesh, const double *f);
/* Assemble A(f) values on the GPU (FP64; FP32 copy made when a mixed solve
 * needs it). method: FK_ASM_GATHER (fastest: 1.0-1.15 ms vs 13.6 ms coloured at 24.6M
 * tets), FK_ASM_COLOURED or FK_ASM_ATOMIC; all three agree to rounding. The
 * first FK_ASM_GATHER call on a mesh builds its incidence lists on the host. */
int fk_assemble(fk_mesh *mesh, int32_t method);
/* Counter of changes of the resident matrix values: it grows with every
 * assembly and whenever a call leaves other values there (the nonlinear loops
 * leave A + At). A caller that keeps a matrix object for later calls on the
 * resident values (lineSearch, src/FEM3D.jl:123-152, takes A as an argument)
 * stores the counter with it and compares before use. No counterpart in the
 * reference, whose matrices are host objects. */
int fk_values_generation(const fk_mesh *mesh, int64_t *generation);
/* Copy the assembled FP64 values (nnz, CSR order == CSC order). */
int fk_values(fk_mesh *mesh, double *values);
int fk_values_f32(fk_mesh *mesh, float *values);
/* stiffnessMatrix(mesh, f) -> SparseMatrixCSC (src/FEM3D.jl:40-56,
 * src/FEMc.cpp:102-133): set f, assemble, copy pattern and values.
 * colptr: nv+1, rowval/nzval: nnz (fk_mesh_info). */
int fk_stiffness_matrix(fk_mesh *mesh, const double *f, int32_t index_base, int32_t *colptr,
                        int32_t *rowval, double *nzval);
/* localStiffnessMatrix(mesh, f) -> Ak 16 x nt column-major
 * (src/FEM3D.jl:59-76, src/FEMc.cpp:76-99). f may be NULL (ones). */
int fk_local_stiffness(fk_mesh *mesh, const double *f, double *Ak);
/* tangentialStiffnessMatrix(mesh, H_vec, dmu) (src/FEM3D.jl:79-119): values
 * on the same pattern. H is 3 x nt, dmu nt. Elements with |H| == 0
 * contribute 0 (the reference divides by |H| there, FEM3D.jl:98). */
int fk_tangential_stiffness(fk_mesh *mesh, const double *H, const double *dmu, double *nzval);
/* lagrange(mesh) (src/FEM3D.jl:274-281): C[nds] += VE/4. C: nv. */
int fk_lagrange(fk_mesh *mesh, double *C);

/* ---- Dirichlet data and solve --------------------------------------------
 * Fixed nodes (mask != 0) keep phi = g; free nodes solve
 * A[free,free] x = -A[free,fixed] g (Dirichlet elimination as in
 * Fluids/fluids3D.jl:142-174, done with a row/column mask, no submatrix). */
int fk_set_dirichlet(fk_mesh *mesh, const uint8_t *mask, const double *g);
/* Jacobi-preconditioned CG (the solver of magnetostatics.cpp:26-38 and
 * run.jl:141) on the assembled matrix, stopping at ||r||_2 < rtol ||b||_2.
 * precision FK_PREC_F64: one FP64 PCG to rtol.
 * precision FK_PREC_F32_REFINED: FP32 values and PCG recursion (difference-
 * form SpMV on the stencil and SELL copies); at the end of each chunk of iterations in
 * which the FP32 recursive residual has fallen by the factor inner_rtol
 * (0 < inner_rtol < 1, e.g. 0.95; other values mean 0.95) since the last update, the
 * FP32 increments are folded into the FP64 potential and the residual is
 * replaced by the FP64 true residual -mask(A64 phi), keeping the search
 * direction ("reliable updates"). The decisions are taken on the device;
 * the solve stops on the FP64 residual (max_outer unused). At the same
 * rtol its potential can be further from the exact solution than the FP64
 * solve's (measured at 4.2M nodes, rtol 1e-10: up to 10x, the remaining
 * error in smooth modes); half the rtol brought it level.
 * precision FK_PREC_F32_RESTART: FP32 PCG to inner_rtol restarted inside an
 * FP64 residual/update loop (at most max_outer restarts).
 * warm_start != 0 starts from the resident potential. The potential stays
 * on the GPU (fk_potential copies it out).
 * iterations: total PCG iterations; outer: residual replacements or restarts;
 * relres: final ||r||/||b|| (FP64 true residual for the mixed solves). */
int fk_solve(fk_mesh *mesh, int32_t precision, double rtol, int64_t max_iter, double inner_rtol,
             int32_t max_outer, int32_t warm_start, int64_t *iterations, int32_t *outer,
             double *relres);
/* Solve A[free,free] x = b[free] with x = 0 on fixed nodes (b: nv, x: nv,
 * FP64 PCG). The matrix is the last assembled one. Used for Newton updates
 * and adjoint solves. */
int fk_solve_rhs(fk_mesh *mesh, const double *b, double rtol, int64_t max_iter, double *x,
                 int64_t *iterations, double *relres);
int fk_potential(fk_mesh *mesh, double *phi);
int fk_set_potential(fk_mesh *mesh, const double *phi);

/* ---- field -----------------------------------------------------------------
 * H = -grad(phi) per element from the resident potential
 * (Magnetostatics/magneticField.jl:111-130, correct z component as in
 * manyBodyIntersecting.jl:241-266). H: 3 x nt (may be NULL), Hn: |H| (may be
 * NULL), energy: 1/2 sum mu |H|^2 VE with the current coefficient (may be
 * NULL; src/magneticProperties.jl:268-273 with the element permeability). */
int fk_field(fk_mesh *mesh, double *H, double *Hn, double *energy);

/* Allocation-free variant of fk_field, with no counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). All
 * pointers are DEVICE pointers: phi (nv, caller node order) in, H (3 x nt),
 * Hn (nt) and energy (one double) out, each output may be NULL. The same
 * kernel as fk_field writes H and |H|; the energy's block partials go to a
 * static per-GPU scratch of 4096 doubles and are summed on the device in the
 * order fk_field sums them, so every output equals fk_field's bit for bit.
 * Enqueued on the context's stream and not waited for (fk_context_synchronize
 * waits); no allocation, copy or synchronization. A mesh whose field kernel
 * needs more block partials than the scratch holds returns
 * FK_ERR_UNSUPPORTED. Calls that overlap on one GPU (two streams or
 * contexts) race on the shared scratch. On a mesh the library renumbered
 * (fk_mesh_numbering) phi is first gathered into the internal node order
 * in a buffer allocated with the mesh. */
int fk_field_noalloc(fk_mesh *mesh, const double *phi, double *H, double *Hn, double *energy);

/* ---- Neumann flux load and Lagrange saddle solve ----------------------------
 * BoundaryIntegral(mesh, F, shell_id) (src/FEM3D.jl:284-297): for surface
 * triangles whose tag (row 4 of surfaceT, 4 x ns) is one of shell_ids,
 * RHS[nodes] += dot(normal, F) * AE / 3. RHS: nv (overwritten). */
int fk_boundary_integral(fk_mesh *mesh, int64_t ns, const int32_t *surfaceT, const double *normal,
                         const double *AE, const double *F, const int32_t *shell_ids, int32_t nshell,
                         int32_t index_base, double *RHS);
/* The saddle solve [A Lag; Lag' 0][u; m] = [b; 0] of
 * Magnetostatics/magneticField.jl:99-108 with A the assembled matrix and Lag
 * = lagrange(mesh): m = sum(b)/sum(Lag), then Jacobi-PCG on the consistent
 * singular system A u = b - m Lag (no fixed node) to rtol, and the constant
 * projected out so that Lag' u = 0. u: nv. */
int fk_solve_neumann(fk_mesh *mesh, const double *b, double rtol, int64_t max_iter, double *u,
                     double *multiplier, int64_t *iterations, double *relres);

/* Drop-in for the reference's own native bridge scalarPotential
 * (Femeko_jl_C++/Magnetostatics/magnetostatics.cpp:49-93, called by @ccall in
 * Femeko_jl_C++/Magnetostatics/run.jl:66-79): the same argument list and
 * meaning, so run.jl switches by changing the library and symbol name only.
 * Context-free and synchronous like the reference: every call uploads the
 * mesh to GPU 0, solves and frees everything. Arrays are 0-based as run.jl
 * passes them (t, surfaceT and shell_id shifted by one, run.jl:50-52):
 * p 3 x nv, t 4 x nt, surfaceT 4 x ns (row 4: surface tag), normal 3 x ns,
 * VE and mu nt, Hext 3. RHS = BoundaryIntegral over the triangles tagged
 * shell_id with areas computed from p (src/FEMc.cpp:45-54, 170-200); u (nv,
 * caller-owned) solves stiffnessMatrix(p, t, VE, mu) u = -RHS by Jacobi-PCG
 * from zero on the singular system (its constant fixed by sum a_ii u_i = 0,
 * where PCG from zero stays) to relative residual 1e-10 (the reference's
 * Eigen default tolerance eps is below what FP64 CG reaches, and its failure
 * is thrown through extern "C", magnetostatics.cpp:29-37; here a status is
 * returned and fk_scalar_potential_last_error() explains it). */
int fk_scalar_potential(double *u, const double *p, const int32_t *t, const int32_t *surfaceT, const double *normal,
                        int32_t nv, int32_t nt, int32_t ns, const double *VE, const double *mu, const double *Hext,
                        int32_t shell_id);
const char *fk_scalar_potential_last_error(void);

/* Demagnetizing field of Micromagnetics/LL.jl:162-209 (magnetostaticField),
 * called once per Landau-Lifshitz step with the matrix fixed: with the
 * assembled A (stiffnessMatrix(mesh), permalloy.jl:99, no fixed node), the
 * load RHS_i = sum over elements k with inside[k] != 0 (caller order) of
 * VE_k grad(lambda_i) . mean(M over the 4 nodes of k), u from the projected
 * Neumann Jacobi-PCG from zero to rtol relative to ||RHS|| (the reference's
 * cg(A, RHS) defaults to sqrt(eps)), H_k = -grad(u), and Hd_nodes[:, i] =
 * sum over inside elements k at node i of VE_k H_k / volumes[i] (NaN where
 * volumes[i] == 0, as the reference's 0/0). M, Hd_nodes: 3 x nv column-major
 * (M[3 i + c]); volumes: nv (the driver's per-node sums of inside VE).
 * warm != 0: the PCG starts from the previous call's potential on this mesh
 * (kept on the device) instead of zero, same stopping test. A non-finite M
 * entry returns FK_ERR_NONFINITE, a negative or non-finite volume
 * FK_ERR_INVALID_ARG, both before the kept potential changes. */
int fk_demag_field(fk_mesh *mesh, const uint8_t *inside, const double *M, const double *volumes, double rtol,
                   int64_t max_iter, int32_t warm, double *Hd_nodes, int64_t *iterations, double *relres);

/* ---- nonlinear loop and line search ----------------------------------------
 * Picard/Newton mu(|H|) loop of Magnetostatics/NewtonRaphson.jl:148-212 with
 * the Froehlich law mu = mu0 (1 + Ms/(a+|H|)) on elements with region != 0
 * (caller order) and mu0 elsewhere, under the Dirichlet data already set.
 * Starts from mu = mu0 (no tangent, as att == 1 there); an iteration adds the
 * tangent of tangentialStiffnessMatrix (src/FEM3D.jl:79-119) once the last
 * relative mu change is below switch_tol. Each FP64 Jacobi-PCG update solve
 * runs to rtol relative to the current residual. Stops when
 * max |mu_new - mu| / mu_new over the region < tol_mu (FK_ERR_NOT_CONVERGED
 * after max_outer iterations). Non-finite mu returns FK_ERR_NONFINITE
 * (NewtonRaphson.jl:210-212). The potential and mu stay resident
 * (fk_potential, fk_get_coefficient, fk_field). */
int fk_nonlinear_froehlich(fk_mesh *mesh, const uint8_t *region, double Ms, double a, double rtol,
                           double tol_mu, int32_t max_outer, double switch_tol, int32_t *outer,
                           int64_t *cg_iterations, double *final_change);
/* Precision of the linear solves that the composite entry points run
 * inside: the update solves of fk_nonlinear_froehlich and
 * fk_nonlinear_spline, the adjoint solve of fk_adjoint_field_target, and the
 * singular Neumann solves of fk_solve_neumann, fk_demag_field and
 * fk_newton_saddle. FK_PREC_F64 (default, as the
 * reference: FP64 Jacobi-PCG) or FK_PREC_F32_REFINED (FP32 values and PCG
 * recursion, the FP64 residual b - A x of the solve's own system replaced as
 * in fk_solve, stopped on that FP64 residual at the same rtol; for the
 * Neumann solves every node is free and the constant is projected out as in
 * FP64). Iteration counts then count FP32 iterations. No counterpart in the
 * reference (its solves are FP64). */
int fk_set_inner_precision(fk_mesh *mesh, int32_t precision);
/* The same loop with the tabulated law of Magnetostatics/NewtonRaphson.jl:
 * 57-63, 204-207: mu = spl(|H|), dmu = spl_dmu(|H|) on the region, two cubic
 * splines in FITPACK form as Dierckx.jl's Spline1D(H, mu; k=3, s=0) holds
 * them (get_knots, get_coeffs): n knots t and n - 4 coefficients c each,
 * evaluated on the GPU. |H| outside [t[3], t[n-4]] returns FK_ERR_DOMAIN
 * (the reference's bc = "error"). Stops when mu0 max | |H| - |H|_prev | over
 * the region < tol_H (NewtonRaphson.jl:196-200, |H|_prev = 0 before the
 * first step); switch_tol = INFINITY takes the Newton tangent from the second
 * iteration on, as the reference (att > 1). final_change: that maximum. */
int fk_nonlinear_spline(fk_mesh *mesh, const uint8_t *region, const double *t_mu, int32_t n_mu,
                        const double *c_mu, const double *t_dmu, int32_t n_dmu, const double *c_dmu,
                        double rtol, double tol_H, int32_t max_outer, double switch_tol, int32_t *outer,
                        int64_t *cg_iterations, double *final_change);
/* The reference's own nonlinear driver, Magnetostatics/NewtonRaphson.jl:
 * 109-212, on the GPU: Neumann load and the Lagrange multiplier of the
 * constant, every Newton step damped by lineSearch. rhs (nv, caller node
 * order) is the driver's RHS = -BoundaryIntegral(mesh, Bext, shell_id)
 * (fk_boundary_integral, negated). From u = 0 and mu = mu0: per attempt,
 * A = A(mu) on the element region's law (mu0 elsewhere), on attempts after
 * the first plus the tangent of tangentialStiffnessMatrix at the current
 * field and dmu (src/FEM3D.jl:79-119); du from [A+At Lag; Lag' 0][du; l] =
 * [rhs - A u; 0] (projected Jacobi-PCG to rtol relative to the right-hand
 * side, Lag' du = 0); u updated by lineSearch(u, du, A, rhs, min_step)
 * (src/FEM3D.jl:123-152); then H, change = mu0 max | |H| - |H|_prev | over
 * the region, mu and dmu from the law. Loops while change > tol_H and
 * attempts < max_outer (FK_ERR_NOT_CONVERGED if the change is still above).
 * The law: with t_mu == NULL the Froehlich law mu0 (1 + Ms/(a+|H|)), else
 * the tabulated splines as for fk_nonlinear_spline (|H| outside the table:
 * FK_ERR_DOMAIN). outer: attempts; cg_iterations: PCG iterations in all;
 * final_change: the last change; residue: the last lineSearch residue. The
 * potential and mu stay resident (fk_potential, fk_get_coefficient). */
int fk_newton_saddle(fk_mesh *mesh, const uint8_t *region, const double *rhs, double Ms, double a, const double *t_mu,
                     int32_t n_mu, const double *c_mu, const double *t_dmu, int32_t n_dmu, const double *c_dmu,
                     double rtol, double tol_H, int32_t max_outer, double min_step, int32_t *outer,
                     int64_t *cg_iterations, double *final_change, double *residue);
/* lineSearch(u, du, A, RHS, minStep) (src/FEM3D.jl:123-152) with A the last
 * assembled matrix (all rows): alpha <- 0.9 alpha until
 * ||RHS - A(u + alpha du)|| <= ||RHS - A u||; u (nv) is updated in place
 * unless alpha fell below min_step. residue: the last trial's norm. */
int fk_line_search(fk_mesh *mesh, double *u, const double *du, const double *rhs, double min_step,
                   double *residue);
/* Current per-element coefficient (mu) in caller order. f: nt. */
int fk_get_coefficient(fk_mesh *mesh, double *f);

/* ---- adjoint ---------------------------------------------------------------
 * No counterpart in the reference: sensitivity of the field-target objective
 * J = sum_{e: target[e]} VE_e |H_e - Hstar|^2 (H from the resident potential
 * of the last fk_solve) with respect to the per-element coefficient mu:
 * adjoint solve A[free,free] lam = -(dJ/dphi)[free] (FP64 Jacobi-PCG to
 * rtol), then dJ/dmu_e = VE_e (sum lam_i grad_i).(sum phi_j grad_j).
 * target: nt (caller order), Hstar: 3, dJdmu: nt. The caller chains
 * dmu/drho (a ChainRulesCore rrule in the Julia wrapper). */
int fk_adjoint_field_target(fk_mesh *mesh, const uint8_t *target, const double *Hstar, double rtol,
                            double *J, double *dJdmu, int64_t *iterations);

/* ---- batched solves ----------------------------------------------------------
 * No counterpart in the reference (which solves one mesh and one applied
 * field per driver run, magneticField.jl:96-108): nmesh independent meshes,
 * nrhs (1..4) Dirichlet right-hand sides each, solved together. Mesh m has
 * p[m] (3 x nv[m]) and t[m] (4 x nt[m]) in Femeko layout with node numbers
 * local to the mesh. The batch is uploaded as one disconnected mesh (one
 * colouring, one pattern, one assembly launch per colour); the solver runs
 * nmesh*nrhs FP32 Jacobi-PCG systems with FP64 residual refinement in one set
 * of graph-captured launches, sharing each mesh's matrix across its
 * right-hand sides. */
typedef struct fk_batch fk_batch;
int fk_batch_create(fk_context *ctx, int32_t nmesh, const int64_t *nv, const int64_t *nt, const double *const *p,
                    const int32_t *const *t, int32_t index_base, int32_t nrhs, fk_batch **out);
int fk_batch_destroy(fk_batch *batch);
int fk_batch_info(const fk_batch *batch, int64_t *nodes, int64_t *nnz, int32_t *nblocks, int32_t *ncolours);
/* f[m]: nt[m] coefficients of mesh m */
int fk_batch_set_coefficient(fk_batch *batch, const double *const *f);
int fk_batch_assemble(fk_batch *batch);
/* mask[m]: nv[m]; g[m]: nrhs * nv[m], right-hand side major (g[m][j*nv[m] + i]) */
int fk_batch_set_dirichlet(fk_batch *batch, const uint8_t *const *mask, const double *const *g);
/* Solve every system to ||r|| < rtol ||b|| (FP64 true residual). With
 * FK_PREC_F32_REFINED (default): FP32 PCG per system with FP64 reliable
 * updates decided on the device, inner_rtol = the residual drop that triggers
 * a replacement (e.g. 0.95), max_outer unused. With FK_PREC_F32_RESTART: FP32
 * PCG to inner_rtol (e.g. 1e-5) restarted inside an FP64 residual-correction
 * loop of at most max_outer steps. Per system s = m*nrhs + j: iterations[s],
 * outer[s] (replacements or restarts), relres[s]. */
int fk_batch_solve(fk_batch *batch, double rtol, double inner_rtol, int32_t max_outer, int64_t max_iter,
                   int64_t *iterations, int32_t *outer, double *relres);
/* FK_PREC_F32_REFINED or FK_PREC_F32_RESTART for the following solves. */
int fk_batch_set_precision(fk_batch *batch, int32_t precision);
/* SpMM copy of the batched solves: FK_SPMV_CSR, FK_SPMV_STENCIL (every mesh's
 * pattern must be a stencil of at most 16 positive offsets with at most 25%
 * padding, else the next solve fails with FK_ERR_INVALID_ARG) or FK_SPMV_AUTO
 * (default: stencil when every mesh qualifies, else CSR). Must be set before
 * the first solve or autotune of the batch. No counterpart in the reference
 * (its solves run one mesh at a time). */
int fk_batch_set_format(fk_batch *batch, int32_t format);
/* The SpMM copy in use: FK_SPMV_CSR or FK_SPMV_STENCIL, or -1 before the
 * first solve or autotune. */
int fk_batch_solver_format(fk_batch *batch, int32_t *format);
/* phi: nrhs * nv[mesh], right-hand side major */
int fk_batch_potential(fk_batch *batch, int32_t mesh, double *phi);
/* Autotuner: times the batched SpMM on this batch (all systems active) and
 * keeps the fastest launch. With the CSR copy: lanes per row {4, 8, 16} x
 * block size {128, 256, 512} (the block size also of the update kernels).
 * With the stencil copy: the SpMM's block size {128, 256, 512} x rows per
 * thread and step {1, 2} (lanes returned unchanged; block = the SpMM's); the
 * update kernels keep their block (512 unless set). fk_batch_set_launch sets
 * lanes and the update kernels' block; fk_batch_launch reports the launch in
 * use. */
int fk_batch_autotune(fk_batch *batch, int32_t *lanes, int32_t *block, double *best_us);
int fk_batch_set_launch(fk_batch *batch, int32_t lanes, int32_t block);
int fk_batch_launch(fk_batch *batch, int32_t *lanes, int32_t *block, int32_t *spmm_block, int32_t *spmm_rows);

/* ---- instrumentation -------------------------------------------------------
 * Average device time of the sparse matrix-vector kernel over the last solve
 * (HIP events on the context stream around each launch), when enabled with
 * fk_set_timing(ctx, 1). Returns launches and total milliseconds. */
int fk_set_timing(fk_context *ctx, int32_t enable);
int fk_spmv_timing(fk_context *ctx, int64_t *launches, double *total_ms);
/* Streaming-bandwidth probe: copy `bytes` (a multiple of 16 is used) between
 * two device buffers `reps` times with a plain kernel, one 16-byte element
 * per thread (k_copy16), and return the average milliseconds per copy.
 * Reads and writes `bytes` each per copy: a reference streaming rate for
 * data larger than the 256 MB MALL, and the kernel of known traffic that
 * calibrates hardware byte counters. */
int fk_copy_probe(fk_context *ctx, int64_t bytes, int32_t reps, double *ms_per_copy);
/* Duration of the solver's SpMV kernel (FK_PREC_F64: the FP64 kernel,
 * FK_PREC_F32_REFINED: the FP32 one, in the format the solver uses for this
 * mesh) from reps back-to-back launches between two events on the library
 * stream, microseconds per launch. Overwrites the solver's direction vector;
 * call between solves. A measurement aid with no counterpart in the reference. */
int fk_spmv_burst(fk_mesh *mesh, int32_t precision, int32_t reps, double *us_per_launch);
/* SpMV tuning: lanes per row (4, 8, 16, 32, 64) of the CSR kernels. */
int fk_set_lanes_per_row(fk_mesh *mesh, int32_t lanes);
/* Matrix format of the solver's SpMV (copies are refreshed from the CSR
 * values after each assembly; the exported matrix is always CSR):
 *   FK_SPMV_CSR: the assembled CSR.
 *   FK_SPMV_SELL64: a sliced-ELL copy of the off-diagonal entries (64 rows
 *     per slice = one wavefront, diagonal kept apart) that stores the column
 *     offsets of a slice once when its rows share one offset set (structured
 *     mesh regions) and per row elsewhere; FK_SPMV_SELL64_ROWIDX always per
 *     row.
 *   FK_SPMV_STENCIL: for patterns whose column offsets j - i form one set of
 *     at most 16 +-offsets (node numberings of structured meshes): one value
 *     column per positive offset, a(i, i + o); a(i, i - o) = a(i - o, i) is
 *     read from the same column (the matrix is exactly symmetric), so each
 *     off-diagonal pair is stored once. FK_ERR_UNSUPPORTED on other patterns.
 *   FK_SPMV_AUTO (default): FK_SPMV_STENCIL when the pattern allows it (at
 *     most 25% padding), else FK_SPMV_SELL64.
 * The SELL and stencil kernels compute the PCG product in the difference
 * form sum_j a_ij (x_j - x_i) (rows of a P1 stiffness matrix sum to zero)
 * in both precisions, the FP64 true residual of the mixed solves in the
 * standard form with the assembled diagonal. Results agree to rounding. */
#define FK_SPMV_CSR 0
#define FK_SPMV_SELL64 1
#define FK_SPMV_SELL64_ROWIDX 2
#define FK_SPMV_STENCIL 3      /* symmetric diagonal storage, see below */
#define FK_SPMV_AUTO 4         /* default: FK_SPMV_STENCIL when the pattern allows, else FK_SPMV_SELL64 */
int fk_set_spmv_format(fk_mesh *mesh, int32_t format);
/* The format the solver uses for this mesh (FK_SPMV_CSR, FK_SPMV_SELL64 or
 * FK_SPMV_STENCIL, resolved on demand); for the stencil its number of
 * positive offsets and the offsets (offsets: room for 16; NULL allowed). */
int fk_solver_format(fk_mesh *mesh, int32_t *format, int32_t *stencil_offsets, int32_t *offsets);
/* SELL copy statistics (built on demand): slices, slices in the shared-offset
 * form, stored entries (values including padding). Any pointer may be NULL. */
int fk_sell_info(fk_mesh *mesh, int64_t *nslices, int64_t *shared_slices, int64_t *entries);
/* SELL copy layout (built on demand): bytes per stored per-row index, 2
 * (int16 offsets from the row, when every column lies within 32767 of its
 * row) or 4 (int32 columns); the groups of 4 entries the FP32 product loads
 * at once (4), or 0 when it walks the groups one by one (slices wider than
 * 19 entries); and the bytes of the arrays one FP32 product streams besides
 * x and y (values with their padding, indices, shared offsets, slice
 * tables), the format's own compulsory traffic with 8 nv for x and y. Any
 * pointer may be NULL. */
int fk_sell_layout(fk_mesh *mesh, int32_t *index_bytes, int32_t *batch_groups, int64_t *f32_array_bytes);
/* Launch shape of the FP32 SELL SpMV: threads per block (256, 512, 1024;
 * default 256, blocks swept in dispatch order per XCD) and 64-row slices in
 * flight per wavefront (1 or 2, default 2; 1024 threads only with 1); a
 * block covers block/64 x unroll slices and leaves one p.q partial. The
 * FP64 product runs on 1024-thread blocks of the solver grid. Rows are
 * computed alike in every shape; the p.q partial sums group differently, so
 * solves agree to rounding. */
int fk_set_sell_shape(fk_mesh *mesh, int32_t block, int32_t unroll);

#ifdef __cplusplus
}
#endif

#endif /* FEMEKO_HIP_H */
