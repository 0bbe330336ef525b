/*
 * femeko_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of Femeko.jl's P1 reduced-scalar-potential
 * magnetostatics path. It is the checker for the HIP library and the
 * CPU-baseline leg of the benchmark; nothing in the shipped library links,
 * loads or calls it. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may use it.
 *
 * Conventions follow the reference exactly:
 *   p  : 3 x nv column-major doubles (x, y, z of node k at p[3k..3k+2])
 *   t  : 4 x nt column-major int32 node numbers, 1-BASED (src/mesh.jl:17)
 *   sparse output: Julia SparseMatrixCSC arrays, 1-based colptr/rowval
 *
 * Each function cites the reference lines it restates. Julia and Eigen are
 * not available, so this file is pinned by closed forms, invariants and
 * scipy cross-checks in tests/ rather than by running the reference.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_OK 0
#define OR_ERR_ALLOC 8
#define OR_ERR_SINGULAR 7
#define OR_ERR_NOT_CONVERGED 6

/* ------------------------------------------------------------------------ */
/* abcd: coefficients of the P1 basis function of node nd on a tetrahedron.
 * src/FEM3D.jl:5-17: rows [1 x y z] of nd first, then the three other nodes
 * in element order; solve M r = e1 with Julia's dense `\` (LU with partial
 * pivoting). Returns a, b, c, d in r[0..3]; b, c, d = grad(lambda_nd). */
int or_abcd(const double *p, const int32_t *nodes, int32_t nd, double *r)
{
    int32_t order[4];
    int m = 1;
    order[0] = nd;
    for (int i = 0; i < 4; ++i)
        if (nodes[i] != nd && m < 4) order[m++] = nodes[i];
    if (m != 4) return OR_ERR_SINGULAR;
    double M[4][5];
    for (int i = 0; i < 4; ++i) {
        const double *q = p + 3 * (int64_t)(order[i] - 1);
        M[i][0] = 1.0;
        M[i][1] = q[0];
        M[i][2] = q[1];
        M[i][3] = q[2];
        M[i][4] = (i == 0) ? 1.0 : 0.0;
    }
    /* LU with partial pivoting (LAPACK getrf ordering), then solve */
    for (int k = 0; k < 4; ++k) {
        int piv = k;
        double best = fabs(M[k][k]);
        for (int i = k + 1; i < 4; ++i)
            if (fabs(M[i][k]) > best) { best = fabs(M[i][k]); piv = i; }
        if (best == 0.0) return OR_ERR_SINGULAR;
        if (piv != k)
            for (int j = 0; j < 5; ++j) { double tmp = M[k][j]; M[k][j] = M[piv][j]; M[piv][j] = tmp; }
        for (int i = k + 1; i < 4; ++i) {
            double l = M[i][k] / M[k][k];
            M[i][k] = l;
            for (int j = k + 1; j < 5; ++j) M[i][j] -= l * M[k][j];
        }
    }
    for (int i = 3; i >= 0; --i) {
        double s = M[i][4];
        for (int j = i + 1; j < 4; ++j) s -= M[i][j] * r[j];
        r[i] = s / M[i][i];
    }
    return OR_OK;
}

/* elementVolume, src/mesh.jl:484-505: |AB . (AC x AD)| / 6 */
double or_element_volume(const double *p, const int32_t *nds)
{
    const double *A = p + 3 * (int64_t)(nds[0] - 1);
    const double *B = p + 3 * (int64_t)(nds[1] - 1);
    const double *C = p + 3 * (int64_t)(nds[2] - 1);
    const double *D = p + 3 * (int64_t)(nds[3] - 1);
    double AB[3], AC[3], AD[3];
    for (int i = 0; i < 3; ++i) { AB[i] = B[i] - A[i]; AC[i] = C[i] - A[i]; AD[i] = D[i] - A[i]; }
    double c0 = AC[1] * AD[2] - AC[2] * AD[1];
    double c1 = AC[2] * AD[0] - AC[0] * AD[2];
    double c2 = AC[0] * AD[1] - AC[1] * AD[0];
    double tp = AB[0] * c0 + AB[1] * c1 + AB[2] * c2;
    return fabs(tp) / 6.0;
}

/* basis gradients of the 4 nodes of element k: g[i] = (b_i, c_i, d_i) */
static int element_grads(const double *p, const int32_t *nds, double g[4][3])
{
    double r[4];
    for (int i = 0; i < 4; ++i) {
        int s = or_abcd(p, nds, nds[i], r);
        if (s) return s;
        g[i][0] = r[1];
        g[i][1] = r[2];
        g[i][2] = r[3];
    }
    return OR_OK;
}

/* grad(lambda_i) of every element via abcd: G is 12 x nt, (b,c,d) of node i
 * at G[12k + 3i .. 3i+2]. */
int or_element_gradients(const double *p, const int32_t *t, int64_t nt, double *G, int nthreads)
{
    int status = OR_OK;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int64_t k = 0; k < nt; ++k) {
        double g[4][3];
        if (element_grads(p, t + 4 * k, g)) { status = OR_ERR_SINGULAR; continue; }
        for (int i = 0; i < 4; ++i)
            for (int c = 0; c < 3; ++c) G[12 * k + 3 * i + c] = g[i][c];
    }
    (void)nthreads;
    return status;
}

/* localStiffnessMatrix, src/FEM3D.jl:59-76 (C++ twin src/FEMc.cpp:76-99):
 * Ak[:,k] = vec(VE[k]*f[k]*(b b' + c c' + d d')), column-major vec, 16 x nt.
 * f may be NULL (ones). */
int or_local_stiffness(const double *p, const int32_t *t, int64_t nt, const double *VE,
                       const double *f, double *Ak, int nthreads)
{
    int status = OR_OK;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int64_t k = 0; k < nt; ++k) {
        double g[4][3];
        if (element_grads(p, t + 4 * k, g)) { status = OR_ERR_SINGULAR; continue; }
        double s = VE[k] * (f ? f[k] : 1.0);
        double *out = Ak + 16 * k;
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                double m = g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2];
                out[i + 4 * j] = s * m;
            }
    }
    (void)nthreads;
    return status;
}

/* Newton tangent, src/FEM3D.jl:79-119: per element
 * At_ij = VE*dmu/|H| * (H . grad_i)(H . grad_j); zero where |H| == 0 (the
 * reference would divide by zero there, FEM3D.jl:98). H is 3 x nt. */
int or_tangent_local(const double *p, const int32_t *t, int64_t nt, const double *VE,
                     const double *H, const double *dmu, double *Ak, int nthreads)
{
    int status = OR_OK;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int64_t k = 0; k < nt; ++k) {
        double g[4][3];
        double *out = Ak + 16 * k;
        if (element_grads(p, t + 4 * k, g)) { status = OR_ERR_SINGULAR; continue; }
        const double *h = H + 3 * k;
        double hn = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
        double hg[4];
        for (int i = 0; i < 4; ++i) hg[i] = h[0] * g[i][0] + h[1] * g[i][1] + h[2] * g[i][2];
        for (int i = 0; i < 4; ++i)
            for (int j = i; j < 4; ++j) {
                double v = (hn > 0.0) ? VE[k] * dmu[k] / hn * hg[i] * hg[j] : 0.0;
                out[i + 4 * j] = v;
                out[j + 4 * i] = v;
            }
    }
    (void)nthreads;
    return status;
}

/* ------------------------------------------------------------------------ */
/* Julia `sparse(rowIDs, colIDs, Acsc, nv, nv)` with the fixed-sparsity COO
 * order of Magnetostatics/NewtonRaphson.jl:115-127,154-161: triplet
 * (n-1)*nt + k has row t[i,k], col t[j,k], value Ak[n,k]*mu[k] (mu may be
 * NULL), n = 4(i-1)+j running i-major. Duplicates are summed in input order
 * and explicit zeros are kept (SparseArrays `sparse`; Eigen setFromTriplets,
 * src/FEMc.cpp:129, does the same). Output is CSC, 1-based. rowval / nzval
 * need room for 16*nt entries; returns nnz (< 0 on allocation failure). */
int64_t or_sparse_csc(int64_t nv, int64_t nt, const int32_t *t, const double *Ak,
                      const double *mu, int32_t *colptr, int32_t *rowval, double *nzval)
{
    const int64_t ntrip = 16 * nt;
    int64_t *cnt = (int64_t *)calloc((size_t)nv + 1, sizeof(int64_t));
    int32_t *brow = (int32_t *)malloc((size_t)ntrip * sizeof(int32_t));
    double *bval = (double *)malloc((size_t)ntrip * sizeof(double));
    int64_t *w = (int64_t *)malloc(((size_t)nv + 1) * sizeof(int64_t));
    int64_t *last = (int64_t *)malloc((size_t)nv * sizeof(int64_t));
    if (!cnt || !brow || !bval || !w || !last) {
        free(cnt); free(brow); free(bval); free(w); free(last);
        return -1;
    }
    /* pass 1: count triplets per column */
    for (int n = 0; n < 16; ++n) {
        int j = n % 4; /* n = 4i + j in 0-based: i-major like the Julia loops */
        for (int64_t k = 0; k < nt; ++k) cnt[t[4 * k + j]] += 1; /* t is 1-based */
    }
    for (int64_t c = 0; c < nv; ++c) cnt[c + 1] += cnt[c];
    memcpy(w, cnt, ((size_t)nv + 1) * sizeof(int64_t));
    /* pass 2: stable bucket by column in input order */
    for (int n = 0; n < 16; ++n) {
        int i = n / 4, j = n % 4;
        for (int64_t k = 0; k < nt; ++k) {
            int32_t col = t[4 * k + j] - 1;
            int64_t pos = w[col]++;
            brow[pos] = t[4 * k + i] - 1;
            /* Ak[n,k] with the i-major n of the Julia loop; Ak[:,k] is the
             * column-major vec of a symmetric 4x4, so this is entry (j,i) ==
             * entry (i,j) bit for bit */
            double v = Ak[16 * k + n];
            bval[pos] = mu ? v * mu[k] : v;
        }
    }
    /* pass 3: per column, combine duplicates (input order) and sort rows */
    for (int64_t r = 0; r < nv; ++r) last[r] = -1;
    int64_t nnz = 0;
    colptr[0] = 1;
    for (int64_t c = 0; c < nv; ++c) {
        int64_t start = nnz;
        for (int64_t q = cnt[c]; q < cnt[c + 1]; ++q) {
            int32_t r = brow[q];
            if (last[r] >= start) {
                nzval[last[r]] += bval[q];
            } else {
                last[r] = nnz;
                rowval[nnz] = r;
                nzval[nnz] = bval[q];
                ++nnz;
            }
        }
        /* insertion sort of this column by row (columns hold ~15 entries) */
        for (int64_t a = start + 1; a < nnz; ++a) {
            int32_t rr = rowval[a];
            double vv = nzval[a];
            int64_t b = a - 1;
            while (b >= start && rowval[b] > rr) {
                rowval[b + 1] = rowval[b];
                nzval[b + 1] = nzval[b];
                --b;
            }
            rowval[b + 1] = rr;
            nzval[b + 1] = vv;
        }
        for (int64_t a = start; a < nnz; ++a) rowval[a] += 1;
        colptr[c + 1] = (int32_t)(nnz + 1);
        (void)start;
    }
    free(cnt); free(brow); free(bval); free(w); free(last);
    return nnz;
}

/* y = A x for CSC arrays (1-based). Exact column scatter, serial. */
void or_spmv_csc(int64_t nv, const int32_t *colptr, const int32_t *rowval, const double *nzval,
                 const double *x, double *y)
{
    for (int64_t r = 0; r < nv; ++r) y[r] = 0.0;
    for (int64_t c = 0; c < nv; ++c) {
        double xc = x[c];
        for (int64_t q = colptr[c] - 1; q < colptr[c + 1] - 1; ++q) y[rowval[q] - 1] += nzval[q] * xc;
    }
}

/* Threaded row-gather product with the CSC arrays read as rows, i.e. A^T x.
 * The assembled A is symmetric up to summation-order rounding of mirrored
 * entries, so this is A x to ~1 ulp per entry; used where OpenMP speed
 * matters (CPU baseline). */
static void spmv_gather(int64_t nv, const int32_t *colptr, const int32_t *rowval,
                        const double *nzval, const double *x, double *y, int nthreads)
{
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int64_t c = 0; c < nv; ++c) {
        double s = 0.0;
        for (int64_t q = colptr[c] - 1; q < colptr[c + 1] - 1; ++q) s += nzval[q] * x[rowval[q] - 1];
        y[c] = s;
    }
    (void)nthreads;
}

static double dotp(int64_t n, const double *a, const double *b, int nthreads)
{
    double s = 0.0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) reduction(+ : s) schedule(static)
#endif
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    (void)nthreads;
    return s;
}

/* Jacobi-preconditioned CG on the free rows: solve A[free,free] x = b[free]
 * with x = 0 on fixed nodes (mask != 0). The loop is the textbook
 * preconditioned CG that Eigen's ConjugateGradient with its default diagonal
 * preconditioner runs (Femeko_jl_C++/Magnetostatics/magnetostatics.cpp:26-38):
 * stop when ||r||_2 < rtol ||b||_2. x holds the initial guess on entry.
 * exact != 0 uses the serial exact A x; otherwise the threaded gather. */
int or_pcg_free(int64_t nv, const int32_t *colptr, const int32_t *rowval, const double *nzval,
                const uint8_t *mask, const double *bin, double rtol, int64_t maxiter,
                double *x, int64_t *iters_out, double *relres_out, int nthreads, int exact)
{
    double *r = (double *)malloc((size_t)nv * sizeof(double));
    double *z = (double *)malloc((size_t)nv * sizeof(double));
    double *pv = (double *)calloc((size_t)nv, sizeof(double));
    double *q = (double *)malloc((size_t)nv * sizeof(double));
    double *dinv = (double *)malloc((size_t)nv * sizeof(double));
    double *b = (double *)malloc((size_t)nv * sizeof(double));
    if (!r || !z || !pv || !q || !dinv || !b) {
        free(r); free(z); free(pv); free(q); free(dinv); free(b);
        return OR_ERR_ALLOC;
    }
#define APPLY(xx, yy) do { if (exact) or_spmv_csc(nv, colptr, rowval, nzval, xx, yy); \
                           else spmv_gather(nv, colptr, rowval, nzval, xx, yy, nthreads); } while (0)
    for (int64_t c = 0; c < nv; ++c) {
        double d = 0.0;
        for (int64_t k = colptr[c] - 1; k < colptr[c + 1] - 1; ++k)
            if (rowval[k] - 1 == c) d = nzval[k];
        dinv[c] = (mask[c] || d == 0.0) ? 0.0 : 1.0 / d;
    }
    for (int64_t i = 0; i < nv; ++i) {
        b[i] = mask[i] ? 0.0 : bin[i];
        if (mask[i]) x[i] = 0.0;
    }
    APPLY(x, q);
    for (int64_t i = 0; i < nv; ++i) r[i] = mask[i] ? 0.0 : b[i] - q[i];
    double bnorm2 = dotp(nv, b, b, nthreads);
    double thresh = rtol * rtol * bnorm2;
    double rr = dotp(nv, r, r, nthreads);
    int64_t it = 0;
    int status = OR_OK;
    if (bnorm2 == 0.0) {
        for (int64_t i = 0; i < nv; ++i) x[i] = 0.0;
        rr = 0.0;
    } else if (rr >= thresh) {
        for (int64_t i = 0; i < nv; ++i) { z[i] = dinv[i] * r[i]; pv[i] = z[i]; }
        double rz = dotp(nv, r, z, nthreads);
        while (it < maxiter) {
            APPLY(pv, q);
            for (int64_t i = 0; i < nv; ++i) if (mask[i]) q[i] = 0.0;
            double pq = dotp(nv, pv, q, nthreads);
            if (!(pq > 0.0)) { status = OR_ERR_SINGULAR; break; }
            double alpha = rz / pq;
            for (int64_t i = 0; i < nv; ++i) {
                x[i] += alpha * pv[i];
                r[i] -= alpha * q[i];
            }
            ++it;
            rr = dotp(nv, r, r, nthreads);
            if (rr < thresh) break;
            for (int64_t i = 0; i < nv; ++i) z[i] = dinv[i] * r[i];
            double rz_new = dotp(nv, r, z, nthreads);
            double beta = rz_new / rz;
            rz = rz_new;
            for (int64_t i = 0; i < nv; ++i) pv[i] = z[i] + beta * pv[i];
        }
        if (status == OR_OK && rr >= thresh) status = OR_ERR_NOT_CONVERGED;
    }
#undef APPLY
    *iters_out = it;
    *relres_out = bnorm2 > 0.0 ? sqrt(rr / bnorm2) : 0.0;
    free(r); free(z); free(pv); free(q); free(dinv); free(b);
    return status;
}

/* Dirichlet elimination as in Fluids/fluids3D.jl:142-174: fixed nodes keep
 * phi = g, free nodes solve A[free,free] x = -A[free,fixed] g[fixed]. phi
 * (nv) is the output; its free entries are the initial guess. */
int or_pcg_dirichlet(int64_t nv, const int32_t *colptr, const int32_t *rowval, const double *nzval,
                     const uint8_t *mask, const double *g, double rtol, int64_t maxiter,
                     double *phi, int64_t *iters_out, double *relres_out, int nthreads, int exact)
{
    double *gd = (double *)malloc((size_t)nv * sizeof(double));
    double *b = (double *)malloc((size_t)nv * sizeof(double));
    if (!gd || !b) { free(gd); free(b); return OR_ERR_ALLOC; }
    for (int64_t i = 0; i < nv; ++i) gd[i] = mask[i] ? g[i] : 0.0;
    if (exact) or_spmv_csc(nv, colptr, rowval, nzval, gd, b);
    else spmv_gather(nv, colptr, rowval, nzval, gd, b, nthreads);
    for (int64_t i = 0; i < nv; ++i) b[i] = -b[i];
    int s = or_pcg_free(nv, colptr, rowval, nzval, mask, b, rtol, maxiter, phi, iters_out,
                        relres_out, nthreads, exact);
    for (int64_t i = 0; i < nv; ++i) if (mask[i]) phi[i] = g[i];
    free(gd); free(b);
    return s;
}

/* Fixed number of PCG iterations on the gather operator, for timing a
 * bounded sample of the CPU path. Returns seconds-free: caller times it. */
int or_pcg_iterations(int64_t nv, const int32_t *colptr, const int32_t *rowval, const double *nzval,
                      const uint8_t *mask, const double *g, int64_t niter, double *phi, int nthreads)
{
    int64_t it;
    double rel;
    int s = or_pcg_dirichlet(nv, colptr, rowval, nzval, mask, g, 0.0, niter, phi, &it, &rel, nthreads, 0);
    return (s == OR_ERR_NOT_CONVERGED) ? OR_OK : s;
}

/* Field recovery, Magnetostatics/magneticField.jl:111-130 with the correct
 * z component (manyBodyIntersecting.jl:241-266): H[:,k] = -sum_nd u[nd] grad
 * of the basis of nd, stored 3 x nt; Hn[k] = |H[:,k]|. */
int or_field(const double *p, const int32_t *t, int64_t nt, const double *u, double *H,
             double *Hn, int nthreads)
{
    int status = OR_OK;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int64_t k = 0; k < nt; ++k) {
        const int32_t *nds = t + 4 * k;
        double hx = 0.0, hy = 0.0, hz = 0.0;
        for (int i = 0; i < 4; ++i) {
            double r[4];
            if (or_abcd(p, nds, nds[i], r)) { status = OR_ERR_SINGULAR; }
            double un = u[nds[i] - 1];
            hx -= un * r[1];
            hy -= un * r[2];
            hz -= un * r[3];
        }
        H[3 * k] = hx;
        H[3 * k + 1] = hy;
        H[3 * k + 2] = hz;
        if (Hn) Hn[k] = sqrt(hx * hx + hy * hy + hz * hz);
    }
    (void)nthreads;
    return status;
}

/* lagrange, src/FEM3D.jl:274-281: C[nds] += VE/4 */
void or_lagrange(int64_t nv, int64_t nt, const int32_t *t, const double *VE, double *C)
{
    for (int64_t i = 0; i < nv; ++i) C[i] = 0.0;
    for (int64_t k = 0; k < nt; ++k)
        for (int i = 0; i < 4; ++i) C[t[4 * k + i] - 1] += VE[k] / 4;
}

/* BoundaryIntegral, src/FEM3D.jl:284-297: for surface triangles whose tag
 * (row 4 of surfaceT) is in shell_ids, RHS[nodes] += dot(normal, F)*AE/3 */
void or_boundary_integral(int64_t nv, int64_t ns, const int32_t *surfaceT, const double *normal,
                          const double *AE, const double *F, const int32_t *shell_ids, int nshell,
                          double *RHS)
{
    for (int64_t i = 0; i < nv; ++i) RHS[i] = 0.0;
    for (int64_t s = 0; s < ns; ++s) {
        int in = 0;
        for (int q = 0; q < nshell; ++q) in |= (surfaceT[4 * s + 3] == shell_ids[q]);
        if (!in) continue;
        const double *nrm = normal + 3 * s;
        double v = (nrm[0] * F[0] + nrm[1] * F[1] + nrm[2] * F[2]) * AE[s] / 3;
        for (int i = 0; i < 3; ++i) RHS[surfaceT[4 * s + i] - 1] += v;
    }
}

int or_num_threads_available(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
