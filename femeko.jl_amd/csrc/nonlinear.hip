// Nonlinear magnetostatics on the GPU: the Picard/Newton mu(|H|) loop of
// Magnetostatics/NewtonRaphson.jl:148-212 with either the Froehlich law
// mu = mu0 (1 + Ms/(a + |H|)) or the reference's tabulated law (two cubic
// splines of |H|, NewtonRaphson.jl:57-63, 204-207) in the magnet region, and
// lineSearch (src/FEM3D.jl:123-152).
//
// Per outer iteration, everything stays on the device:
//   k_assemble_newton_gather (row-gather form, assemble.hip; the coloured
//       k_assemble_newton below, one launch per colour, when the rows are too
//       long for its LDS segment): gradients, H = -grad(phi) of the current
//       potential, K_e(mu) plus (Newton step) the tangent
//       VE mu'/|H| (H.g_i)(H.g_j) of src/FEM3D.jl:79-119 into the CSR values,
//       and the element residual -K_e(mu) phi_e into r -- one pass, no
//       separate SpMV for the residual;
//   FP64 Jacobi-PCG for the update du (rtol relative to ||r||);
//   phi += du;
//   k_field_mu / k_field_spline: H of the new potential, mu and mu' from the
//       law (the "mu update fused into the element gradient kernel"), block
//       maxima of the law's change measure over the magnet (Froehlich:
//       |mu_new - mu| / mu_new; tabulated: mu0 | |H| - |H|_prev |, the
//       reference's) and a counter of non-finite or out-of-table results.
// The host reads one max per iteration to decide convergence and whether the
// next step is Newton (previous change below `switch_tol`) or Picard;
// iteration 1 starts from mu = mu0 with no tangent, as the reference's
// att == 1 branch.
#include <algorithm>
#include <cmath>

#include "device.h"
#include "internal.h"

namespace fk {

__global__ void k_fill_f64b(double *__restrict__ a, int64_t n, double v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_gather_u8(const uint8_t *__restrict__ src, const int32_t *__restrict__ perm, int64_t nt,
                            uint8_t *__restrict__ dst) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < nt) dst[e] = src[perm[e]] ? 1 : 0;
}

__global__ void k_scatter_coef(const double *__restrict__ coef, const int32_t *__restrict__ perm, int64_t nt,
                               double *__restrict__ out) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < nt) out[perm[e]] = coef[e];
}

__global__ __launch_bounds__(256) void k_assemble_newton(const int4 *__restrict__ tet, const double *__restrict__ X,
                                                         const double *__restrict__ Y, const double *__restrict__ Z,
                                                         const double *__restrict__ mu,
                                                         const double *__restrict__ dmu,
                                                         const double *__restrict__ phi,
                                                         const int32_t *__restrict__ slot, double *__restrict__ val,
                                                         double *__restrict__ r, int newton, int64_t e0, int64_t e1) {
    const int64_t e = e0 + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= e1) return;
    const int4 nd = tet[e];
    const Grad G = element_gradients(X, Y, Z, nd);
    const double vol = fabs(G.det) * (1.0 / 6.0);
    const double u[4] = {phi[nd.x], phi[nd.y], phi[nd.z], phi[nd.w]};
    const double w = vol * mu[e];
    double K[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j) {
            K[i][j] = w * dot3(G.g[i], G.g[j]);
            K[j][i] = K[i][j];
        }
    // element residual -K u (before adding the tangent)
    double ru[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) ru[i] = -(K[i][0] * u[0] + K[i][1] * u[1] + K[i][2] * u[2] + K[i][3] * u[3]);
    if (newton) {
        double h[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) h[c] = -(u[0] * G.g[0][c] + u[1] * G.g[1][c] + u[2] * G.g[2][c] + u[3] * G.g[3][c]);
        const double hn = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
        const double wt = hn > 0.0 ? vol * dmu[e] / hn : 0.0;
        double hg[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) hg[i] = dot3(h, G.g[i]);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) K[i][j] += wt * hg[i] * hg[j];
    }
    const int4 *s4 = reinterpret_cast<const int4 *>(slot + 16 * e);
    const int4 q0 = s4[0], q1 = s4[1], q2 = s4[2], q3 = s4[3];
    const int s[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                       q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) val[s[4 * i + j]] += K[i][j];
    r[nd.x] += ru[0];
    r[nd.y] += ru[1];
    r[nd.z] += ru[2];
    r[nd.w] += ru[3];
}

// H of phi, Froehlich mu and mu' in the region, block max of the relative
// change, count of non-finite results
__global__ __launch_bounds__(kBlock) void k_field_mu(const int4 *__restrict__ tet, const double *__restrict__ X,
                                                     const double *__restrict__ Y, const double *__restrict__ Z,
                                                     const double *__restrict__ phi,
                                                     const uint8_t *__restrict__ region, double mu0, double Ms,
                                                     double a, int64_t nt, double *__restrict__ mu,
                                                     double *__restrict__ dmu, double *__restrict__ hn_prev,
                                                     double *__restrict__ part, int32_t *__restrict__ bad) {
    __shared__ double sh[kBlock / 64 + 1];
    double dev = 0.0;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < nt;
         e += static_cast<int64_t>(gridDim.x) * kBlock) {
        if (!region[e]) continue;
        const int4 nd = tet[e];
        const Grad G = element_gradients(X, Y, Z, nd);
        const double u0 = phi[nd.x], u1 = phi[nd.y], u2 = phi[nd.z], u3 = phi[nd.w];
        double hh = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double h = -(u0 * G.g[0][c] + u1 * G.g[1][c] + u2 * G.g[2][c] + u3 * G.g[3][c]);
            hh += h * h;
        }
        const double hn = sqrt(hh);
        const double d = a + hn;
        const double mnew = mu0 * (1.0 + Ms / d);
        const double dnew = -mu0 * Ms / (d * d);
        if (!isfinite(mnew) || !isfinite(dnew)) atomicAdd(bad, 1);
        if (hn_prev) {
            // the reference's measure mu0 | |H| - |H|_prev | (NewtonRaphson.jl:200)
            dev = fmax(dev, mu0 * fabs(hn - hn_prev[e]));
            hn_prev[e] = hn;
        } else {
            dev = fmax(dev, fabs(mnew - mu[e]) / mnew);
        }
        mu[e] = mnew;
        dmu[e] = dnew;
    }
    const double m = block_max<kBlock>(dev, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// Cubic spline sum_j c[j] B_j(x) with n knots t (FITPACK form, c[0:n-4]), the
// value Dierckx.jl's Spline1D returns. The four B-splines that are nonzero on
// the knot interval t[l] <= x < t[l+1] (3 <= l <= n-5; the right end t[n-4]
// belongs to the last interval) come from the de Boor-Cox recurrence
// B_{i,j} = w_{i,j} B_{i,j-1} + (1 - w_{i+1,j}) B_{i+1,j-1} with
// w_{i,j}(x) = (x - t_i) / (t_{i+j} - t_i), starting from B_{l,0} = 1.
// Outside [t[3], t[n-4]] it sets *out and returns 0.
__device__ __forceinline__ double bspline3(const double *__restrict__ t, const double *__restrict__ c, int n,
                                           double x, int *out) {
    if (!(x >= t[3] && x <= t[n - 4])) {
        *out = 1;
        return 0.0;
    }
    int lo = 3, hi = n - 5;  // largest l in [3, n-5] with t[l] <= x
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t[mid] <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int l = lo;
    double b[5] = {0.0, 0.0, 0.0, 1.0, 0.0};  // b[i] = B_{l-3+i, j}
#pragma unroll
    for (int j = 1; j <= 3; ++j) {
#pragma unroll
        for (int i = 3 - j; i <= 3; ++i) {
            const int idx = l - 3 + i;
            const double d0 = t[idx + j] - t[idx];
            const double d1 = t[idx + j + 1] - t[idx + 1];
            const double left = d0 > 0.0 ? (x - t[idx]) / d0 * b[i] : 0.0;
            const double right = d1 > 0.0 ? (t[idx + j + 1] - x) / d1 * b[i + 1] : 0.0;
            b[i] = left + right;
        }
    }
    return c[l - 3] * b[0] + c[l - 2] * b[1] + c[l - 1] * b[2] + c[l] * b[3];
}

struct SplineLaw {
    const double *t_mu, *c_mu, *t_dmu, *c_dmu;
    int n_mu, n_dmu;
};

// H of phi, tabulated mu and mu' in the region, block max of
// mu0 | |H| - |H|_prev |, count of out-of-table or non-finite results
__global__ __launch_bounds__(kBlock) void k_field_spline(const int4 *__restrict__ tet,
                                                         const double *__restrict__ X,
                                                         const double *__restrict__ Y,
                                                         const double *__restrict__ Z,
                                                         const double *__restrict__ phi,
                                                         const uint8_t *__restrict__ region, double mu0,
                                                         const SplineLaw law, int64_t nt,
                                                         double *__restrict__ mu, double *__restrict__ dmu,
                                                         double *__restrict__ hn_prev, double *__restrict__ part,
                                                         int32_t *__restrict__ bad, int32_t *__restrict__ outside) {
    __shared__ double sh[kBlock / 64 + 1];
    double dev = 0.0;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < nt;
         e += static_cast<int64_t>(gridDim.x) * kBlock) {
        if (!region[e]) continue;
        const int4 nd = tet[e];
        const Grad G = element_gradients(X, Y, Z, nd);
        const double u0 = phi[nd.x], u1 = phi[nd.y], u2 = phi[nd.z], u3 = phi[nd.w];
        double hh = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double h = -(u0 * G.g[0][c] + u1 * G.g[1][c] + u2 * G.g[2][c] + u3 * G.g[3][c]);
            hh += h * h;
        }
        const double hn = sqrt(hh);
        int out = 0;
        const double mnew = bspline3(law.t_mu, law.c_mu, law.n_mu, hn, &out);
        const double dnew = bspline3(law.t_dmu, law.c_dmu, law.n_dmu, hn, &out);
        if (out) atomicAdd(outside, 1);
        if (!isfinite(mnew) || !isfinite(dnew)) atomicAdd(bad, 1);
        dev = fmax(dev, mu0 * fabs(hn - hn_prev[e]));
        hn_prev[e] = hn;
        mu[e] = mnew;
        dmu[e] = dnew;
    }
    const double m = block_max<kBlock>(dev, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// y += x
__global__ void k_add_into(double *__restrict__ y, const double *__restrict__ x, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

// y = u + alpha du (trial of the line search)
__global__ void k_axpy_to(const double *__restrict__ u, const double *__restrict__ du, double alpha,
                          double *__restrict__ y, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = u[i] + alpha * du[i];
}

// partials of ||rhs - A x||^2 over all rows (the unmasked residual of lineSearch)
__global__ __launch_bounds__(kBlock) void k_resid_norm(const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ col,
                                                       const double *__restrict__ val, const double *__restrict__ x,
                                                       const double *__restrict__ rhs, double *__restrict__ part,
                                                       int64_t n) {
    __shared__ double sh[kBlock / 64 + 1];
    constexpr int L = 16, RPB = kBlock / L;
    const int sub = threadIdx.x / L, lane = threadIdx.x % L;
    double acc = 0.0;
    for (int64_t rb = static_cast<int64_t>(blockIdx.x) * RPB; rb < n; rb += static_cast<int64_t>(gridDim.x) * RPB) {
        const int64_t row = rb + sub;
        double s = 0.0;
        if (row < n) {
            const int32_t a0 = rowptr[row], a1 = rowptr[row + 1];
            for (int32_t k = a0 + lane; k < a1; k += L) s += val[k] * x[col[k]];
        }
#pragma unroll
        for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (row < n && lane == 0) {
            const double d = rhs[row] - s;
            acc += d * d;
        }
    }
    const double tot = block_sum<kBlock>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// ---------------------------------------------------------------------------
static double host_max(const double *p, int n) {
    double m = 0.0;
    for (int i = 0; i < n; ++i) m = std::max(m, p[i]);
    return m;
}

static double host_sum(const double *p, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += p[i];
    return s;
}

// the law of one loop: Froehlich (spline == nullptr) or tabulated
struct LawArgs {
    double Ms = 0.0, a = 1.0;
    const SplineLaw *spline = nullptr;
};

static void nonlinear_loop_body(fk_mesh *m, const uint8_t *region_host, const LawArgs &law, double rtol,
                                double tol, int max_outer, double switch_tol, int32_t *outer_out, int64_t *cg_out,
                                double *dev_out);

// The loop leaves A + At of its last step in the values: they are marked
// stale on every exit, errors included, so a later solve reassembles A(mu).
static void nonlinear_loop(fk_mesh *m, const uint8_t *region_host, const LawArgs &law, double rtol, double tol,
                           int max_outer, double switch_tol, int32_t *outer_out, int64_t *cg_out,
                           double *dev_out) {
    try {
        nonlinear_loop_body(m, region_host, law, rtol, tol, max_outer, switch_tol, outer_out, cg_out, dev_out);
    } catch (...) {
        m->have_values = false;
        m->values_gen += 1;
        throw;
    }
    m->have_values = false;
    m->values_gen += 1;
}

// The law's region (caller element order) into m->region (colour order).
static void upload_region(fk_mesh *m, const uint8_t *region_host) {
    hipStream_t st = m->ctx->stream;
    DevBuf<uint8_t> tmp(m->nt);
    m->region.ensure(m->nt);
    FK_HIP(hipMemcpyAsync(tmp.get(), region_host, m->nt, hipMemcpyHostToDevice, st));
    k_gather_u8<<<grid_for(m->nt, 256), 256, 0, st>>>(tmp.get(), m->perm.get(), m->nt, m->region.get());
    FK_HIP(hipGetLastError());
    FK_HIP(hipStreamSynchronize(st));
}

// val64 = K(mu) (+ the tangent on a Newton step) at the resident potential and
// r = -K(mu) phi, all rows; the row-gather form, or one coloured launch per
// colour when the rows exceed its LDS segment.
static void newton_values(fk_mesh *m, int newton, double *r) {
    hipStream_t st = m->ctx->stream;
    if (!assemble_newton_gather(m, m->coef.get(), m->dmu.get(), m->phi.get(), newton, r)) {
        k_fill_f64b<<<grid_for(m->nnz, 256), 256, 0, st>>>(m->val64.get(), m->nnz, 0.0);
        k_fill_f64b<<<grid_for(m->nv, 256), 256, 0, st>>>(r, m->nv, 0.0);
        for (int c = 0; c < m->ncolours; ++c) {
            const int64_t e0 = m->colour_off[c], e1 = m->colour_off[c + 1];
            if (e1 <= e0) continue;
            k_assemble_newton<<<grid_for(e1 - e0, 256), 256, 0, st>>>(
                m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(), m->coef.get(), m->dmu.get(), m->phi.get(),
                m->slot.get(), m->val64.get(), r, newton, e0, e1);
        }
    }
    FK_HIP(hipGetLastError());
    values_changed(m);
}

static void nonlinear_loop_body(fk_mesh *m, const uint8_t *region_host, const LawArgs &law, double rtol,
                                double tol, int max_outer, double switch_tol, int32_t *outer_out, int64_t *cg_out,
                                double *dev_out) {
    if (!m->have_dirichlet) fail(FK_ERR_STATE, "no Dirichlet data set");
    if (m->nfixed == 0) fail(FK_ERR_SINGULAR, "no fixed node: the pure Neumann system is singular");
    hipStream_t st = m->ctx->stream;
    const int64_t nv = m->nv, nt = m->nt;
    const double mu0 = 3.141592653589793 * 4e-7;
    upload_region(m, region_host);
    m->coef.ensure(nt);
    m->dmu.ensure(nt);
    m->val64.ensure(m->nnz);
    m->phi.ensure(nv);
    m->r64.ensure(nv);
    ensure_work_f64(m);
    k_fill_f64b<<<grid_for(nt, 256), 256, 0, st>>>(m->coef.get(), nt, mu0);
    k_fill_f64b<<<grid_for(nt, 256), 256, 0, st>>>(m->dmu.get(), nt, 0.0);
    if (law.spline) {
        m->hn_prev.ensure(nt);
        k_fill_f64b<<<grid_for(nt, 256), 256, 0, st>>>(m->hn_prev.get(), nt, 0.0);
    }
    k_init_phi_dev(m, 0);  // phi = g on fixed nodes, 0 elsewhere
    m->have_coef = true;
    m->coef_nat_valid = false;  // the loop updates coef (colour order) on the device
    m->coef_g_valid = false;
    m->have_phi = true;
    const int nb = std::max(1, m->nblocks);
    m->part_misc.ensure(nb);
    DevBuf<int32_t> bad(2);  // non-finite, outside the table
    double dev = INFINITY;
    int32_t it = 0;
    int64_t cg = 0;
    PcgWork<double> &w = m->w64;
    for (it = 1; it <= max_outer; ++it) {
        const int newton = dev < switch_tol ? 1 : 0;
        newton_values(m, newton, w.r.get());
        if (m->inner_precision == FK_PREC_F32_REFINED) {
            // FP32 recursion on the step's system, FP64 true residual
            // w.r - (A + At) du replaced as the recursion falls
            m->inc64.ensure(nv);
            int64_t its = 0;
            const int s32 = solve_f32_increment(m, m->inc64.get(), w.r.get(), rtol, 1000000, &its);
            cg += its;
            if (s32 != FK_OK && s32 != FK_ERR_NOT_CONVERGED)
                fail(s32, std::string("inner solve: ") + fk_status_string(s32));
            phi_axpy_f64(m, m->inc64.get(), 1.0);
        } else {
            compute_dinv_f64(m);
            m->dinv64_valid = true;
            const CgState s = run_pcg_f64(m, rtol, 1000000, -1.0);
            cg += s.iters;
            if (s.status != FK_OK && s.status != FK_ERR_NOT_CONVERGED)
                fail(s.status, std::string("inner solve: ") + fk_status_string(s.status));
            phi_axpy_f64(m, w.x.get(), 1.0);
        }
        FK_HIP(hipMemsetAsync(bad.get(), 0, 2 * sizeof(int32_t), st));
        if (law.spline)
            k_field_spline<<<nb, kBlock, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(), m->phi.get(),
                                                  m->region.get(), mu0, *law.spline, nt, m->coef.get(),
                                                  m->dmu.get(), m->hn_prev.get(), m->part_misc.get(), bad.get(),
                                                  bad.get() + 1);
        else
            k_field_mu<<<nb, kBlock, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(), m->phi.get(),
                                              m->region.get(), mu0, law.Ms, law.a, nt, m->coef.get(),
                                              m->dmu.get(), nullptr, m->part_misc.get(), bad.get());
        FK_HIP(hipGetLastError());
        int32_t hbad[2] = {0, 0};
        FK_HIP(hipMemcpyAsync(m->hpart.ptr, m->part_misc.get(), nb * sizeof(double), hipMemcpyDeviceToHost, st));
        FK_HIP(hipMemcpyAsync(hbad, bad.get(), 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        FK_HIP(hipStreamSynchronize(st));
        if (hbad[1]) fail(FK_ERR_DOMAIN, std::to_string(hbad[1]) + " element(s) with |H| outside the table");
        if (hbad[0]) fail(FK_ERR_NONFINITE, "Nans/Infs in mu");  // NewtonRaphson.jl:210-212
        dev = host_max(m->hpart.ptr, nb);
        if (dev < tol) break;
    }
    if (outer_out) *outer_out = std::min(it, max_outer);
    if (cg_out) *cg_out = cg;
    if (dev_out) *dev_out = dev;
    if (!(dev < tol)) fail(FK_ERR_NOT_CONVERGED, "mu update did not reach its tolerance");
}

void nonlinear_froehlich(fk_mesh *m, const uint8_t *region_host, double Ms, double a, double rtol, double tol_mu,
                         int max_outer, double switch_tol, int32_t *outer_out, int64_t *cg_out, double *dev_out) {
    LawArgs law;
    law.Ms = Ms;
    law.a = a;
    nonlinear_loop(m, region_host, law, rtol, tol_mu, max_outer, switch_tol, outer_out, cg_out, dev_out);
}

// knots non-decreasing and finite, a non-empty domain, finite coefficients
static void check_spline(const double *t, int n, const double *c, const char *name) {
    if (n < 8) fail(FK_ERR_INVALID_ARG, std::string(name) + ": a cubic spline needs at least 8 knots");
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(t[i]) || (i > 0 && t[i] < t[i - 1]))
            fail(FK_ERR_INVALID_ARG, std::string(name) + ": knots must be finite and non-decreasing");
    if (!(t[3] < t[n - 4])) fail(FK_ERR_INVALID_ARG, std::string(name) + ": empty knot domain");
    for (int i = 0; i < n - 4; ++i)
        if (!std::isfinite(c[i])) fail(FK_ERR_NONFINITE, std::string(name) + ": non-finite coefficient");
}

void nonlinear_spline(fk_mesh *m, const uint8_t *region_host, const double *t_mu, int n_mu, const double *c_mu,
                      const double *t_dmu, int n_dmu, const double *c_dmu, double rtol, double tol_H, int max_outer,
                      double switch_tol, int32_t *outer_out, int64_t *cg_out, double *dev_out) {
    check_spline(t_mu, n_mu, c_mu, "mu spline");
    check_spline(t_dmu, n_dmu, c_dmu, "dmu spline");
    hipStream_t st = m->ctx->stream;
    DevBuf<double> tm(n_mu), cm(n_mu - 4), td(n_dmu), cd(n_dmu - 4);
    FK_HIP(hipMemcpyAsync(tm.get(), t_mu, n_mu * sizeof(double), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(cm.get(), c_mu, (n_mu - 4) * sizeof(double), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(td.get(), t_dmu, n_dmu * sizeof(double), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(cd.get(), c_dmu, (n_dmu - 4) * sizeof(double), hipMemcpyHostToDevice, st));
    const SplineLaw sl{tm.get(), cm.get(), td.get(), cd.get(), n_mu, n_dmu};
    LawArgs law;
    law.spline = &sl;
    nonlinear_loop(m, region_host, law, rtol, tol_H, max_outer, switch_tol, outer_out, cg_out, dev_out);
}

// lineSearch (src/FEM3D.jl:123-152) on device vectors with A the resident
// values (all rows): alpha <- 0.9 alpha while ||rhs - A(u + alpha du)|| >
// ||rhs - A u|| and alpha > min_step; u becomes u + alpha du unless alpha fell
// below min_step. y: nv scratch. Returns the last trial's residue; *alpha_out
// the last trial step and *moved whether u changed.
static double line_search_dev(fk_mesh *m, double *u, const double *du, const double *rhs, double min_step,
                              double *y, double *alpha_out, bool *moved) {
    hipStream_t st = m->ctx->stream;
    const int64_t nv = m->nv;
    const int nb = std::max(1, m->nblocks);
    m->part_misc.ensure(nb);
    auto norm_of = [&](const double *x) {
        k_resid_norm<<<nb, kBlock, 0, st>>>(m->rowptr.get(), m->colidx.get(), m->val64.get(), x, rhs,
                                            m->part_misc.get(), nv);
        FK_HIP(hipGetLastError());
        FK_HIP(hipMemcpyAsync(m->hpart.ptr, m->part_misc.get(), nb * sizeof(double), hipMemcpyDeviceToHost, st));
        FK_HIP(hipStreamSynchronize(st));
        return std::sqrt(host_sum(m->hpart.ptr, nb));
    };
    double alf = 1.0;
    const double residual_old = norm_of(u);
    k_axpy_to<<<grid_for(nv, 256), 256, 0, st>>>(u, du, 1.0, y, nv);
    double residue = norm_of(y);
    while (residue > residual_old && alf > min_step) {
        alf *= 0.9;
        k_axpy_to<<<grid_for(nv, 256), 256, 0, st>>>(u, du, alf, y, nv);
        residue = norm_of(y);
    }
    *moved = !(alf < min_step);
    if (*moved) FK_HIP(hipMemcpyAsync(u, y, nv * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (alpha_out) *alpha_out = alf;
    return residue;
}

double line_search(fk_mesh *m, double *u_host, const double *du_host, const double *rhs_host, double min_step) {
    if (!m->have_values) fail(FK_ERR_STATE, "matrix not assembled");
    hipStream_t st = m->ctx->stream;
    const int64_t nv = m->nv;
    DevBuf<double> u(nv), du(nv), rhs(nv), y(nv);
    {
        const NodesIn<double> ui(m, u_host), dui(m, du_host), ri(m, rhs_host);
        FK_HIP(hipMemcpyAsync(u.get(), ui.ptr, nv * sizeof(double), hipMemcpyHostToDevice, st));
        FK_HIP(hipMemcpyAsync(du.get(), dui.ptr, nv * sizeof(double), hipMemcpyHostToDevice, st));
        FK_HIP(hipMemcpyAsync(rhs.get(), ri.ptr, nv * sizeof(double), hipMemcpyHostToDevice, st));
        FK_HIP(hipStreamSynchronize(st));
    }
    bool moved = false;
    const double residue = line_search_dev(m, u.get(), du.get(), rhs.get(), min_step, y.get(), nullptr, &moved);
    if (moved) {
        NodesOut<double> uo(m, u_host);
        FK_HIP(hipMemcpyAsync(uo.ptr, u.get(), nv * sizeof(double), hipMemcpyDeviceToHost, st));
        FK_HIP(hipStreamSynchronize(st));
        uo.commit();
    }
    return residue;
}

// ---- the reference's nonlinear driver ---------------------------------------
// Magnetostatics/NewtonRaphson.jl:148-212: Neumann boundary load and the
// Lagrange multiplier of the constant, every step damped by lineSearch.
//   att = 1: A(mu0), no tangent; att > 1: A(mu) + At(H, dmu)
//   du from [A + At, Lag; Lag', 0] [du; l] = [RHS - A u; 0]   (:171)
//   u updated by lineSearch(u, du, A, RHS) with A = A(mu)      (:175)
//   H of u, div = mu0 max |H - Hold| over the region           (:200)
//   mu, dmu from the law at |H| on the region                  (:204-207)
// while div > tol_H and att < max_outer. The bordered solve is the projected
// Jacobi-PCG of solve_neumann_dev (A + At is singular with null space the
// constants, like A: the tangent's rows sum to zero too); everything stays on
// the device, the host reads the scalar decisions.
static void newton_saddle_body(fk_mesh *m, const uint8_t *region_host, const double *rhs_host, const LawArgs &law,
                               double rtol, double tol_H, int max_outer, double min_step, int32_t *outer_out,
                               int64_t *cg_out, double *div_out, double *residue_out) {
    hipStream_t st = m->ctx->stream;
    const int64_t nv = m->nv, nt = m->nt;
    const double mu0 = 3.141592653589793 * 4e-7;
    upload_region(m, region_host);
    m->coef.ensure(nt);
    m->dmu.ensure(nt);
    m->hn_prev.ensure(nt);
    m->val64.ensure(m->nnz);
    m->phi.ensure(nv);
    ensure_work_f64(m);
    k_fill_f64b<<<grid_for(nt, 256), 256, 0, st>>>(m->coef.get(), nt, mu0);
    k_fill_f64b<<<grid_for(nt, 256), 256, 0, st>>>(m->dmu.get(), nt, 0.0);
    k_fill_f64b<<<grid_for(nt, 256), 256, 0, st>>>(m->hn_prev.get(), nt, 0.0);
    k_fill_f64b<<<grid_for(nv, 256), 256, 0, st>>>(m->phi.get(), nv, 0.0);
    FK_HIP(hipGetLastError());
    m->have_coef = true;
    m->coef_nat_valid = false;
    m->coef_g_valid = false;
    m->have_phi = true;
    DevBuf<double> rhs(nv), y(nv), junk(nv);
    {
        const NodesIn<double> ri(m, rhs_host);
        FK_HIP(hipMemcpyAsync(rhs.get(), ri.ptr, nv * sizeof(double), hipMemcpyHostToDevice, st));
        FK_HIP(hipStreamSynchronize(st));
    }
    const int nb = std::max(1, m->nblocks);
    m->part_misc.ensure(nb);
    DevBuf<int32_t> bad(2);  // non-finite, outside the table
    PcgWork<double> &w = m->w64;
    double div = INFINITY, residue = 0.0;
    int32_t att = 0;
    int64_t cg = 0;
    while (div > tol_H && att < max_outer) {
        ++att;
        const int newton = att > 1 ? 1 : 0;
        newton_values(m, newton, w.r.get());  // A (+ At), r = -A u
        k_add_into<<<grid_for(nv, 256), 256, 0, st>>>(w.r.get(), rhs.get(), nv);
        FK_HIP(hipGetLastError());
        const NeumannResult nr = solve_neumann_dev(m, rtol, 1000000, nullptr);  // du in w.x
        cg += nr.iters;
        if (newton) newton_values(m, 0, junk.get());  // lineSearch takes A(mu) alone
        bool moved = false;
        residue = line_search_dev(m, m->phi.get(), w.x.get(), rhs.get(), min_step, y.get(), nullptr, &moved);
        FK_HIP(hipMemsetAsync(bad.get(), 0, 2 * sizeof(int32_t), st));
        if (law.spline)
            k_field_spline<<<nb, kBlock, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(), m->phi.get(),
                                                  m->region.get(), mu0, *law.spline, nt, m->coef.get(),
                                                  m->dmu.get(), m->hn_prev.get(), m->part_misc.get(), bad.get(),
                                                  bad.get() + 1);
        else
            k_field_mu<<<nb, kBlock, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(), m->phi.get(),
                                              m->region.get(), mu0, law.Ms, law.a, nt, m->coef.get(),
                                              m->dmu.get(), m->hn_prev.get(), m->part_misc.get(), bad.get());
        FK_HIP(hipGetLastError());
        int32_t hbad[2] = {0, 0};
        FK_HIP(hipMemcpyAsync(m->hpart.ptr, m->part_misc.get(), nb * sizeof(double), hipMemcpyDeviceToHost, st));
        FK_HIP(hipMemcpyAsync(hbad, bad.get(), 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        FK_HIP(hipStreamSynchronize(st));
        if (hbad[1]) fail(FK_ERR_DOMAIN, std::to_string(hbad[1]) + " element(s) with |H| outside the table");
        if (hbad[0]) fail(FK_ERR_NONFINITE, "Nans/Infs in mu");  // NewtonRaphson.jl:210-212
        div = host_max(m->hpart.ptr, nb);
    }
    if (outer_out) *outer_out = att;
    if (cg_out) *cg_out = cg;
    if (div_out) *div_out = div;
    if (residue_out) *residue_out = residue;
    if (div > tol_H) fail(FK_ERR_NOT_CONVERGED, "mu0 max |H - Hold| did not reach its tolerance");
}

void newton_saddle(fk_mesh *m, const uint8_t *region_host, const double *rhs_host, double Ms, double a,
                   const double *t_mu, int n_mu, const double *c_mu, const double *t_dmu, int n_dmu,
                   const double *c_dmu, double rtol, double tol_H, int max_outer, double min_step,
                   int32_t *outer_out, int64_t *cg_out, double *div_out, double *residue_out) {
    hipStream_t st = m->ctx->stream;
    LawArgs law;
    law.Ms = Ms;
    law.a = a;
    DevBuf<double> tm, cm, td, cd;
    SplineLaw sl{};
    if (t_mu) {
        check_spline(t_mu, n_mu, c_mu, "mu spline");
        check_spline(t_dmu, n_dmu, c_dmu, "dmu spline");
        tm.alloc(n_mu); cm.alloc(n_mu - 4); td.alloc(n_dmu); cd.alloc(n_dmu - 4);
        FK_HIP(hipMemcpyAsync(tm.get(), t_mu, n_mu * sizeof(double), hipMemcpyHostToDevice, st));
        FK_HIP(hipMemcpyAsync(cm.get(), c_mu, (n_mu - 4) * sizeof(double), hipMemcpyHostToDevice, st));
        FK_HIP(hipMemcpyAsync(td.get(), t_dmu, n_dmu * sizeof(double), hipMemcpyHostToDevice, st));
        FK_HIP(hipMemcpyAsync(cd.get(), c_dmu, (n_dmu - 4) * sizeof(double), hipMemcpyHostToDevice, st));
        sl = SplineLaw{tm.get(), cm.get(), td.get(), cd.get(), n_mu, n_dmu};
        law.spline = &sl;
    } else if (!(a > 0.0) || !std::isfinite(Ms)) {
        fail(FK_ERR_INVALID_ARG, "Froehlich law needs a > 0 and a finite Ms");
    }
    try {
        newton_saddle_body(m, region_host, rhs_host, law, rtol, tol_H, max_outer, min_step, outer_out, cg_out,
                           div_out, residue_out);
    } catch (...) {
        m->have_values = false;
        m->values_gen += 1;
        throw;
    }
    // the values hold A(mu) of the last step, not of the final mu
    m->have_values = false;
    m->values_gen += 1;
}

void get_coefficient(fk_mesh *m, double *f_host) {
    if (!m->have_coef) fail(FK_ERR_STATE, "no coefficient");
    hipStream_t st = m->ctx->stream;
    DevBuf<double> out(m->nt);
    k_scatter_coef<<<grid_for(m->nt, 256), 256, 0, st>>>(m->coef.get(), m->perm.get(), m->nt, out.get());
    FK_HIP(hipGetLastError());
    FK_HIP(hipMemcpyAsync(f_host, out.get(), m->nt * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
}

}  // namespace fk

using fk::guard;

extern "C" {

int fk_set_inner_precision(fk_mesh *m, int32_t precision) {
    if (!m || (precision != FK_PREC_F64 && precision != FK_PREC_F32_REFINED)) return FK_ERR_INVALID_ARG;
    m->inner_precision = precision;
    return FK_OK;
}

int fk_nonlinear_froehlich(fk_mesh *m, const uint8_t *region, double Ms, double a, double rtol, double tol_mu,
                           int32_t max_outer, double switch_tol, int32_t *outer, int64_t *cg_iterations,
                           double *final_change) {
    if (!m || !region || !(Ms >= 0.0) || !(a > 0.0) || !(rtol > 0.0) || !(tol_mu > 0.0) || max_outer < 1)
        return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_nonlinear_froehlich", [&] {
        fk::set_device(m->ctx);
        fk::nonlinear_froehlich(m, region, Ms, a, rtol, tol_mu, max_outer, switch_tol, outer, cg_iterations,
                                final_change);
    });
}

int fk_nonlinear_spline(fk_mesh *m, const uint8_t *region, const double *t_mu, int32_t n_mu, const double *c_mu,
                        const double *t_dmu, int32_t n_dmu, const double *c_dmu, double rtol, double tol_H,
                        int32_t max_outer, double switch_tol, int32_t *outer, int64_t *cg_iterations,
                        double *final_change) {
    if (!m || !region || !t_mu || !c_mu || !t_dmu || !c_dmu || n_mu < 8 || n_dmu < 8 || !(rtol > 0.0) ||
        !(tol_H > 0.0) || max_outer < 1)
        return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_nonlinear_spline", [&] {
        fk::set_device(m->ctx);
        fk::nonlinear_spline(m, region, t_mu, n_mu, c_mu, t_dmu, n_dmu, c_dmu, rtol, tol_H, max_outer, switch_tol,
                             outer, cg_iterations, final_change);
    });
}

int fk_line_search(fk_mesh *m, double *u, const double *du, const double *rhs, double min_step, double *residue) {
    if (!m || !u || !du || !rhs) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_line_search", [&] {
        fk::set_device(m->ctx);
        const double r = fk::line_search(m, u, du, rhs, min_step);
        if (residue) *residue = r;
    });
}

int fk_get_coefficient(fk_mesh *m, double *f) {
    if (!m || !f) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_get_coefficient", [&] {
        fk::set_device(m->ctx);
        fk::get_coefficient(m, f);
    });
}

int fk_newton_saddle(fk_mesh *m, const uint8_t *region, const double *rhs, double Ms, double a, const double *t_mu,
                     int32_t n_mu, const double *c_mu, const double *t_dmu, int32_t n_dmu, const double *c_dmu,
                     double rtol, double tol_H, int32_t max_outer, double min_step, int32_t *outer,
                     int64_t *cg_iterations, double *final_change, double *residue) {
    if (!m || !region || !rhs || !(rtol > 0.0) || !(tol_H >= 0.0) || max_outer < 1 || !(min_step >= 0.0) ||
        (t_mu && (!c_mu || !t_dmu || !c_dmu)))
        return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_newton_saddle", [&] {
        fk::set_device(m->ctx);
        fk::newton_saddle(m, region, rhs, Ms, a, t_mu, n_mu, c_mu, t_dmu, n_dmu, c_dmu, rtol, tol_H, max_outer,
                          min_step, outer, cg_iterations, final_change, residue);
    });
}

}  // extern "C"
