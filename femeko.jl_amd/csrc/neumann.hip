// Neumann flux load and the Lagrange-multiplier saddle solve of the
// reference's linear driver, Magnetostatics/magneticField.jl:96-108:
//   RHS = BoundaryIntegral(mesh, F, shell_id)      (src/FEM3D.jl:284-297)
//   Lag = lagrange(mesh)                           (src/FEM3D.jl:274-281)
//   [A Lag; Lag' 0] [u; m] = [b; 0]
// A is singular with null space span{1} (A 1 = 0) and Lag' 1 = total volume
// > 0, so the saddle system has the unique solution
//   m = (1'b) / (1'Lag),   A u = b - m Lag  (consistent),   Lag' u = 0.
// The GPU solves the consistent singular system with the same Jacobi-PCG
// kernels (every node free) and projects the constant out at the end:
//   u <- u - (Lag'u / Lag'1) 1.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "device.h"
#include "internal.h"

namespace fk {

// F . n * area / 3 onto the 3 nodes of every tagged surface triangle
// (FP64 atomics: a handful of triangles share a node)
__global__ void k_boundary_integral(int64_t ns, const int32_t *__restrict__ sT, const double *__restrict__ nrm,
                                    const double *__restrict__ AE, double fx, double fy, double fz,
                                    const int32_t *__restrict__ ids, int nids, int32_t base,
                                    double *__restrict__ rhs) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int32_t tag = sT[4 * s + 3];
    bool in = false;
    for (int k = 0; k < nids; ++k) in |= (tag == ids[k]);
    if (!in) return;
    const double v = (nrm[3 * s] * fx + nrm[3 * s + 1] * fy + nrm[3 * s + 2] * fz) * AE[s] / 3;
    for (int i = 0; i < 3; ++i) unsafeAtomicAdd(&rhs[sT[4 * s + i] - base], v);
}

__global__ void k_dinv_free(const double *__restrict__ val, const int32_t *__restrict__ diag, int64_t nv,
                            double *__restrict__ dinv, int32_t *__restrict__ bad) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const int32_t s = diag[i];
    const double d = s >= 0 ? val[s] : 0.0;
    if (!(d > 0.0) || !isfinite(d)) {
        dinv[i] = 0.0;
        atomicAdd(bad, 1);
        return;
    }
    dinv[i] = 1.0 / d;
}

// partials of (sum a, sum c, sum a.c)
__global__ __launch_bounds__(kBlock) void k_sums3(const double *__restrict__ a, const double *__restrict__ c,
                                                  int64_t n, double *__restrict__ part) {
    __shared__ double sh[kBlock / 64 + 1];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kBlock) {
        s0 += a[i];
        s1 += c[i];
        s2 += a[i] * c[i];
    }
    s0 = block_sum<kBlock>(s0, sh);
    s1 = block_sum<kBlock>(s1, sh);
    s2 = block_sum<kBlock>(s2, sh);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = s0;
        part[3 * blockIdx.x + 1] = s1;
        part[3 * blockIdx.x + 2] = s2;
    }
}

__global__ void k_axpy_inplace(double *__restrict__ y, const double *__restrict__ x, double a, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

// r -= A u over the FP64 CSR values (one row per thread group of 16 lanes)
__global__ __launch_bounds__(kBlock) void k_sub_spmv_csr(const int32_t *__restrict__ rowptr,
                                                         const int32_t *__restrict__ col,
                                                         const double *__restrict__ val,
                                                         const double *__restrict__ u, double *__restrict__ r,
                                                         int64_t n) {
    constexpr int L = 16, RPB = kBlock / L;
    const int sub = threadIdx.x / L, lane = threadIdx.x % L;
    for (int64_t rb = static_cast<int64_t>(blockIdx.x) * RPB; rb < n; rb += static_cast<int64_t>(gridDim.x) * RPB) {
        const int64_t row = rb + sub;
        double s = 0.0;
        if (row < n) {
            const int32_t a0 = rowptr[row], a1 = rowptr[row + 1];
            for (int32_t k = a0 + lane; k < a1; k += L) s += val[k] * u[col[k]];
        }
#pragma unroll
        for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (row < n && lane == 0) r[row] -= s;
    }
}

// partials of (sum d_i u_i, sum d_i) with d_i = a_ii
__global__ __launch_bounds__(kBlock) void k_diag_weighted(const double *__restrict__ val,
                                                          const int32_t *__restrict__ diag,
                                                          const double *__restrict__ u, int64_t n,
                                                          double *__restrict__ part) {
    __shared__ double sh[kBlock / 64 + 1];
    double s0 = 0.0, s1 = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kBlock) {
        const double d = diag[i] >= 0 ? val[diag[i]] : 0.0;
        s0 += d * u[i];
        s1 += d;
    }
    s0 = block_sum<kBlock>(s0, sh);
    s1 = block_sum<kBlock>(s1, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0;
        part[2 * blockIdx.x + 1] = s1;
    }
}

__global__ void k_shift(double *__restrict__ y, double c, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] -= c;
}

__global__ __launch_bounds__(256) void k_lagrange_n(const int4 *__restrict__ tet, const double *__restrict__ X,
                                                    const double *__restrict__ Y, const double *__restrict__ Z,
                                                    double *__restrict__ C, int64_t e0, int64_t e1) {
    const int64_t e = e0 + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= e1) return;
    const int4 nd = tet[e];
    const Grad G = element_gradients(X, Y, Z, nd);
    const double v = fabs(G.det) * (1.0 / 6.0) / 4;
    C[nd.x] += v;
    C[nd.y] += v;
    C[nd.z] += v;
    C[nd.w] += v;
}

void boundary_integral(fk_mesh *m, int64_t ns, const int32_t *sT_caller, const double *nrm, const double *AE,
                       const double *F, const int32_t *ids, int nids, int32_t base, double *rhs_host) {
    hipStream_t st = m->ctx->stream;
    for (int64_t s = 0; s < ns; ++s)
        for (int i = 0; i < 3; ++i) {
            const int64_t v = static_cast<int64_t>(sT_caller[4 * s + i]) - base;
            if (v < 0 || v >= m->nv) fail(FK_ERR_INDEX, "surface triangle node number out of range");
        }
    // triangle nodes in the internal numbering
    std::vector<int32_t> sT_int;
    const int32_t *sT = sT_caller;
    if (m->reordered && ns > 0) {
        sT_int.assign(sT_caller, sT_caller + 4 * ns);
        for (int64_t s = 0; s < ns; ++s)
            for (int i = 0; i < 3; ++i) sT_int[4 * s + i] = m->node_inv[sT_caller[4 * s + i] - base] + base;
        sT = sT_int.data();
    }
    NodesOut<double> rhs_out(m, rhs_host);
    DevBuf<int32_t> dT(4 * std::max<int64_t>(ns, 1)), dids(std::max(nids, 1));
    DevBuf<double> dn(3 * std::max<int64_t>(ns, 1)), dA(std::max<int64_t>(ns, 1)), rhs(m->nv);
    if (ns > 0) {
        FK_HIP(hipMemcpyAsync(dT.get(), sT, 4 * ns * sizeof(int32_t), hipMemcpyHostToDevice, st));
        FK_HIP(hipMemcpyAsync(dn.get(), nrm, 3 * ns * sizeof(double), hipMemcpyHostToDevice, st));
        FK_HIP(hipMemcpyAsync(dA.get(), AE, ns * sizeof(double), hipMemcpyHostToDevice, st));
    }
    if (nids > 0) FK_HIP(hipMemcpyAsync(dids.get(), ids, nids * sizeof(int32_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemsetAsync(rhs.get(), 0, m->nv * sizeof(double), st));
    if (ns > 0)
        k_boundary_integral<<<grid_for(ns, 256), 256, 0, st>>>(ns, dT.get(), dn.get(), dA.get(), F[0], F[1], F[2],
                                                               dids.get(), nids, base, rhs.get());
    FK_HIP(hipGetLastError());
    FK_HIP(hipMemcpyAsync(rhs_out.ptr, rhs.get(), m->nv * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    rhs_out.commit();
}

static void sums3(fk_mesh *m, const double *a, const double *c, double out[3]) {
    hipStream_t st = m->ctx->stream;
    const int nb = std::max(1, m->nblocks);
    DevBuf<double> part(3 * nb);
    k_sums3<<<nb, kBlock, 0, st>>>(a, c, m->nv, part.get());
    FK_HIP(hipGetLastError());
    std::vector<double> h(3 * nb);
    FK_HIP(hipMemcpyAsync(h.data(), part.get(), 3 * nb * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    out[0] = out[1] = out[2] = 0.0;
    for (int i = 0; i < nb; ++i) {
        out[0] += h[3 * i];
        out[1] += h[3 * i + 1];
        out[2] += h[3 * i + 2];
    }
}

// The saddle solve with b already in w64.r on the device; leaves u (Lag' u =
// 0) in w64.x. With u0 (device, nv) the PCG starts from u0: it solves for
// the correction of the residual b' - A u0 and still stops at rtol relative
// to ||b'||, the test of a start from zero.
NeumannResult solve_neumann_dev(fk_mesh *m, double rtol, int64_t max_iter, const double *u0) {
    if (!m->have_values) fail(FK_ERR_STATE, "matrix not assembled");
    hipStream_t st = m->ctx->stream;
    const int64_t nv = m->nv;
    PcgWork<double> &w = m->w64;
    // every node free
    {
        DevBuf<int32_t> bad(1);
        FK_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), st));
        k_dinv_free<<<grid_for(nv, 256), 256, 0, st>>>(m->val64.get(), m->diag.get(), nv, w.dinv.get(), bad.get());
        FK_HIP(hipGetLastError());
        int32_t h = 0;
        FK_HIP(hipMemcpyAsync(&h, bad.get(), sizeof(int32_t), hipMemcpyDeviceToHost, st));
        FK_HIP(hipStreamSynchronize(st));
        m->dinv64_valid = false;  // the Dirichlet solves recompute theirs
        if (h) fail(FK_ERR_SINGULAR, std::to_string(h) + " node(s) without a positive diagonal");
    }
    // Lag (geometry only: built once per mesh and kept) and b
    if (!m->have_lag) {
        m->lag.ensure(nv);
        FK_HIP(hipMemsetAsync(m->lag.get(), 0, nv * sizeof(double), st));
        for (int c = 0; c < m->ncolours; ++c) {
            const int64_t e0 = m->colour_off[c], e1 = m->colour_off[c + 1];
            if (e1 <= e0) continue;
            k_lagrange_n<<<grid_for(e1 - e0, 256), 256, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(),
                                                                 m->lag.get(), e0, e1);
        }
        FK_HIP(hipGetLastError());
        m->have_lag = true;
    }
    const DevBuf<double> &lag = m->lag;
    double s[3];
    sums3(m, w.r.get(), lag.get(), s);  // sum b, sum Lag
    const double mult = s[1] != 0.0 ? s[0] / s[1] : 0.0;
    // b' = b - m Lag, consistent with the null space of A
    k_axpy_inplace<<<grid_for(nv, 256), 256, 0, st>>>(w.r.get(), lag.get(), -mult, nv);
    FK_HIP(hipGetLastError());
    double bb_ref = -1.0;
    if (u0) {
        sums3(m, w.r.get(), w.r.get(), s);
        bb_ref = s[2];
        k_sub_spmv_csr<<<std::max(1, m->nblocks), kBlock, 0, st>>>(m->rowptr.get(), m->colidx.get(),
                                                                    m->val64.get(), u0, w.r.get(), nv);
        FK_HIP(hipGetLastError());
    }
    CgState cs{};
    if (m->inner_precision == FK_PREC_F32_REFINED) {
        // FP32 recursion with FP64 residual replacement on every row; the
        // solution (from zero, the correction when warm) into w.x
        double rr = 0.0, bb = 0.0;
        int64_t its = 0;
        cs.status = solve_f32_free(m, w.x.get(), w.r.get(), rtol, max_iter, bb_ref, &its, &rr, &bb);
        cs.iters = its;
        cs.rr = rr;
        cs.bb = bb;
    } else {
        cs = run_pcg_f64(m, rtol, max_iter, bb_ref);
    }
    if (cs.status != FK_OK) fail(cs.status, std::string("Neumann solve: ") + fk_status_string(cs.status));
    if (u0) {
        k_axpy_inplace<<<grid_for(nv, 256), 256, 0, st>>>(w.x.get(), u0, 1.0, nv);
        FK_HIP(hipGetLastError());
    }
    // project: Lag' u = 0
    sums3(m, w.x.get(), lag.get(), s);
    k_shift<<<grid_for(nv, 256), 256, 0, st>>>(w.x.get(), s[2] / s[1], nv);
    FK_HIP(hipGetLastError());
    return NeumannResult{mult, cs.iters, cs.bb > 0.0 ? std::sqrt(cs.rr / cs.bb) : 0.0};
}

void solve_neumann(fk_mesh *m, const double *b_host, double rtol, int64_t max_iter, double *u_host,
                   double *mult_out, int64_t *iters_out, double *relres_out) {
    if (!m->have_values) fail(FK_ERR_STATE, "matrix not assembled");
    hipStream_t st = m->ctx->stream;
    ensure_work_f64(m);
    const NodesIn<double> b_in(m, b_host);
    NodesOut<double> u_out(m, u_host);
    FK_HIP(hipMemcpyAsync(m->w64.r.get(), b_in.ptr, m->nv * sizeof(double), hipMemcpyHostToDevice, st));
    const NeumannResult res = solve_neumann_dev(m, rtol, max_iter, nullptr);
    FK_HIP(hipMemcpyAsync(u_out.ptr, m->w64.x.get(), m->nv * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    u_out.commit();
    if (mult_out) *mult_out = res.multiplier;
    if (iters_out) *iters_out = res.iters;
    if (relres_out) *relres_out = res.relres;
}

// scalarPotential of Femeko_jl_C++/Magnetostatics/magnetostatics.cpp:49-93
// (the C++ bridge run.jl:66-79 calls): RHS = BoundaryIntegral with the
// triangle areas from p (src/FEMc.cpp:45-54, 170-200), A = stiffnessMatrix(p,
// t, VE, mu) (FEMc.cpp:102-133), u solving A u = -RHS by Jacobi-PCG from zero
// (Eigen's ConjugateGradient, magnetostatics.cpp:26-38). A is singular (A 1 =
// 0) and -RHS consistent for a closed shell; PCG from zero with the diagonal
// preconditioner keeps sum_i a_ii u_i = 0, so the constant is fixed that way.
static void scalar_potential(double *u, const double *p, const int32_t *t, const int32_t *surfaceT,
                             const double *normal, int64_t nv, int64_t nt, int64_t ns, const double *VE,
                             const double *mu, const double *Hext, int32_t shell_id, double rtol) {
    struct Ctx {
        fk_context *c = nullptr;
        ~Ctx() { if (c) fk_context_destroy(c); }
    } ctx;
    struct Mesh {
        fk_mesh *m = nullptr;
        ~Mesh() { if (m) fk_mesh_destroy(m); }
    } mesh;
    int s = fk_context_create(0, &ctx.c);
    if (s != FK_OK) fail(s, "no GPU context on device 0");
    s = fk_mesh_create(ctx.c, p, nv, t, nt, 0, VE, &mesh.m);
    if (s != FK_OK) fail(s, std::string("mesh: ") + fk_context_last_error(ctx.c));
    fk_mesh *m = mesh.m;
    set_device(m->ctx);
    // triangle areas |AB x AC| / 2 from the node coordinates
    std::vector<double> AE(static_cast<size_t>(std::max<int64_t>(ns, 1)));
    for (int64_t k = 0; k < ns; ++k) {
        const int32_t *tri = surfaceT + 4 * k;
        for (int i = 0; i < 3; ++i)
            if (tri[i] < 0 || tri[i] >= nv) fail(FK_ERR_INDEX, "surface triangle node number out of range");
        const double *a = p + 3 * static_cast<int64_t>(tri[0]), *b = p + 3 * static_cast<int64_t>(tri[1]),
                     *c = p + 3 * static_cast<int64_t>(tri[2]);
        const double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double x = ab[1] * ac[2] - ab[2] * ac[1], y = ab[2] * ac[0] - ab[0] * ac[2],
                     z = ab[0] * ac[1] - ab[1] * ac[0];
        AE[k] = 0.5 * std::sqrt(x * x + y * y + z * z);
    }
    std::vector<double> rhs(static_cast<size_t>(nv));
    boundary_integral(m, ns, surfaceT, normal, AE.data(), Hext, &shell_id, 1, 0, rhs.data());
    for (auto &v : rhs) v = -v;
    const NodesIn<double> rhs_in(m, rhs.data());
    upload_coefficient(m, mu);
    assemble(m, FK_ASM_GATHER);
    ensure_work_f64(m);
    hipStream_t st = m->ctx->stream;
    FK_HIP(hipMemcpyAsync(m->w64.r.get(), rhs_in.ptr, nv * sizeof(double), hipMemcpyHostToDevice, st));
    solve_neumann_dev(m, rtol, std::max<int64_t>(1000, 20 * nv), nullptr);
    // fix the constant: sum_i a_ii u_i = 0
    const int nb = std::max(1, m->nblocks);
    DevBuf<double> part(2 * nb);
    k_diag_weighted<<<nb, kBlock, 0, st>>>(m->val64.get(), m->diag.get(), m->w64.x.get(), nv, part.get());
    FK_HIP(hipGetLastError());
    std::vector<double> h(2 * nb);
    FK_HIP(hipMemcpyAsync(h.data(), part.get(), 2 * nb * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    double du = 0.0, dd = 0.0;
    for (int i = 0; i < nb; ++i) {
        du += h[2 * i];
        dd += h[2 * i + 1];
    }
    k_shift<<<grid_for(nv, 256), 256, 0, st>>>(m->w64.x.get(), dd > 0.0 ? du / dd : 0.0, nv);
    FK_HIP(hipGetLastError());
    NodesOut<double> u_out(m, u);
    FK_HIP(hipMemcpyAsync(u_out.ptr, m->w64.x.get(), nv * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    u_out.commit();
}

thread_local std::string g_bridge_error;

}  // namespace fk

using fk::guard;

extern "C" {

int fk_boundary_integral(fk_mesh *m, int64_t ns, const int32_t *surfaceT, const double *normal, const double *AE,
                         const double *F, const int32_t *shell_ids, int32_t nshell, int32_t index_base,
                         double *RHS) {
    if (!m || ns < 0 || (ns > 0 && (!surfaceT || !normal || !AE)) || !F || (nshell > 0 && !shell_ids) ||
        nshell < 0 || !RHS || (index_base != 0 && index_base != 1))
        return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_boundary_integral", [&] {
        fk::set_device(m->ctx);
        fk::boundary_integral(m, ns, surfaceT, normal, AE, F, shell_ids, nshell, index_base, RHS);
    });
}

int fk_solve_neumann(fk_mesh *m, const double *b, double rtol, int64_t max_iter, double *u, double *multiplier,
                     int64_t *iterations, double *relres) {
    if (!m || !b || !u || !(rtol > 0.0) || max_iter < 1) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_solve_neumann", [&] {
        fk::set_device(m->ctx);
        fk::solve_neumann(m, b, rtol, max_iter, u, multiplier, iterations, relres);
    });
}

int fk_scalar_potential(double *u, const double *p, const int32_t *t, const int32_t *surfaceT, const double *normal,
                        int32_t nv, int32_t nt, int32_t ns, const double *VE, const double *mu, const double *Hext,
                        int32_t shell_id) {
    fk::g_bridge_error.clear();
    if (!u || !p || !t || nv < 1 || nt < 1 || ns < 0 || (ns > 0 && (!surfaceT || !normal)) || !mu || !Hext) {
        fk::g_bridge_error = "fk_scalar_potential: null pointer or bad size";
        return FK_ERR_INVALID_ARG;
    }
    try {
        fk::scalar_potential(u, p, t, surfaceT, normal, nv, nt, ns, VE, mu, Hext, shell_id, 1e-10);
        return FK_OK;
    } catch (const fk::Error &e) {
        fk::g_bridge_error = "fk_scalar_potential: " + e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        fk::g_bridge_error = "fk_scalar_potential: host allocation failed";
        return FK_ERR_OOM;
    } catch (...) {
        fk::g_bridge_error = "fk_scalar_potential: unexpected error";
        return FK_ERR_HIP;
    }
}

const char *fk_scalar_potential_last_error(void) { return fk::g_bridge_error.c_str(); }

}  // extern "C"
