// Adjoint sensitivity of a field-target objective (no counterpart in the
// reference; it differentiates the linear solve of
// Magnetostatics/magneticField.jl:96-130 with respect to the per-element
// permeability).
//
//   J(mu) = sum_{e in target} VE_e |H_e - H*|^2,  H_e = -sum_i phi_i grad_i
//   state:   A(mu)[free,free] phi_f = -A[free,fixed] g           (fk_solve)
//   adjoint: A[free,free] lam_f = -(dJ/dphi)_f, lam = 0 on fixed nodes
//   dJ/dmu_e = VE_e (sum_i lam_i grad_i) . (sum_j phi_j grad_j)
//
// dJ/dphi_i = sum_{e in target, i in e} -2 VE_e (H_e - H*) . grad_i is
// scattered colour by colour (conflict free, fixed order); the adjoint solve
// reuses the FP64 Jacobi-PCG; the contraction is one thread per element.
#include <algorithm>
#include <cmath>

#include "device.h"
#include "internal.h"

namespace fk {

__global__ void k_fill_zero(double *__restrict__ a, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

__global__ void k_gather_target(const uint8_t *__restrict__ src, const int32_t *__restrict__ perm, int64_t nt,
                                uint8_t *__restrict__ dst) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < nt) dst[e] = src[perm[e]] ? 1 : 0;
}

// b -= dJ/dphi over the target elements of one colour (b = -dJ/dphi)
__global__ __launch_bounds__(256) void k_adjoint_rhs(const int4 *__restrict__ tet, const double *__restrict__ X,
                                                     const double *__restrict__ Y, const double *__restrict__ Z,
                                                     const uint8_t *__restrict__ target,
                                                     const double *__restrict__ phi, double hx, double hy, double hz,
                                                     double *__restrict__ b, int64_t e0, int64_t e1) {
    const int64_t e = e0 + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= e1 || !target[e]) return;
    const int4 nd = tet[e];
    const Grad G = element_gradients(X, Y, Z, nd);
    const double vol = fabs(G.det) * (1.0 / 6.0);
    const double u[4] = {phi[nd.x], phi[nd.y], phi[nd.z], phi[nd.w]};
    double d[3];
    const double hs[3] = {hx, hy, hz};
#pragma unroll
    for (int c = 0; c < 3; ++c)
        d[c] = -(u[0] * G.g[0][c] + u[1] * G.g[1][c] + u[2] * G.g[2][c] + u[3] * G.g[3][c]) - hs[c];
    // -dJ/dphi_i = 2 VE d . grad_i
    b[nd.x] += 2.0 * vol * dot3(d, G.g[0]);
    b[nd.y] += 2.0 * vol * dot3(d, G.g[1]);
    b[nd.z] += 2.0 * vol * dot3(d, G.g[2]);
    b[nd.w] += 2.0 * vol * dot3(d, G.g[3]);
}

// block partials of J
__global__ __launch_bounds__(kBlock) void k_objective(const int4 *__restrict__ tet, const double *__restrict__ X,
                                                      const double *__restrict__ Y, const double *__restrict__ Z,
                                                      const uint8_t *__restrict__ target,
                                                      const double *__restrict__ phi, double hx, double hy,
                                                      double hz, int64_t nt, double *__restrict__ part) {
    __shared__ double sh[kBlock / 64 + 1];
    double acc = 0.0;
    const double hs[3] = {hx, hy, hz};
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < nt;
         e += static_cast<int64_t>(gridDim.x) * kBlock) {
        if (!target[e]) continue;
        const int4 nd = tet[e];
        const Grad G = element_gradients(X, Y, Z, nd);
        const double u0 = phi[nd.x], u1 = phi[nd.y], u2 = phi[nd.z], u3 = phi[nd.w];
        double dd = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double d = -(u0 * G.g[0][c] + u1 * G.g[1][c] + u2 * G.g[2][c] + u3 * G.g[3][c]) - hs[c];
            dd += d * d;
        }
        acc += fabs(G.det) * (1.0 / 6.0) * dd;
    }
    const double tot = block_sum<kBlock>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// dJ/dmu_e = VE (sum lam_i grad_i) . (sum phi_j grad_j), caller order
__global__ __launch_bounds__(256) void k_adjoint_grad(const int4 *__restrict__ tet, const double *__restrict__ X,
                                                      const double *__restrict__ Y, const double *__restrict__ Z,
                                                      const int32_t *__restrict__ perm,
                                                      const double *__restrict__ phi, const double *__restrict__ lam,
                                                      int64_t nt, double *__restrict__ grad) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nt) return;
    const int4 nd = tet[e];
    const Grad G = element_gradients(X, Y, Z, nd);
    const double u[4] = {phi[nd.x], phi[nd.y], phi[nd.z], phi[nd.w]};
    const double l[4] = {lam[nd.x], lam[nd.y], lam[nd.z], lam[nd.w]};
    double gu[3], gl[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        gu[c] = u[0] * G.g[0][c] + u[1] * G.g[1][c] + u[2] * G.g[2][c] + u[3] * G.g[3][c];
        gl[c] = l[0] * G.g[0][c] + l[1] * G.g[1][c] + l[2] * G.g[2][c] + l[3] * G.g[3][c];
    }
    grad[perm[e]] = fabs(G.det) * (1.0 / 6.0) * dot3(gu, gl);
}

void adjoint_field_target(fk_mesh *m, const uint8_t *target_host, const double *Hstar, double rtol, double *J_out,
                          double *grad_host, int64_t *iters_out) {
    if (!m->have_phi) fail(FK_ERR_STATE, "no state potential: call fk_solve first");
    if (!m->have_values) fail(FK_ERR_STATE, "matrix not assembled");
    if (!m->have_dirichlet) fail(FK_ERR_STATE, "no Dirichlet data set");
    hipStream_t st = m->ctx->stream;
    const int64_t nv = m->nv, nt = m->nt;
    {
        DevBuf<uint8_t> tmp(nt);
        m->target.ensure(nt);
        FK_HIP(hipMemcpyAsync(tmp.get(), target_host, nt, hipMemcpyHostToDevice, st));
        k_gather_target<<<grid_for(nt, 256), 256, 0, st>>>(tmp.get(), m->perm.get(), nt, m->target.get());
        FK_HIP(hipGetLastError());
        FK_HIP(hipStreamSynchronize(st));
    }
    ensure_work_f64(m);
    if (!m->dinv64_valid) {
        compute_dinv_f64(m);
        m->dinv64_valid = true;
    }
    const int nb = std::max(1, m->nblocks);
    m->part_misc.ensure(nb);
    // objective
    k_objective<<<nb, kBlock, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(), m->target.get(),
                                       m->phi.get(), Hstar[0], Hstar[1], Hstar[2], nt, m->part_misc.get());
    FK_HIP(hipGetLastError());
    FK_HIP(hipMemcpyAsync(m->hpart.ptr, m->part_misc.get(), nb * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    double J = 0.0;
    for (int i = 0; i < nb; ++i) J += m->hpart.ptr[i];
    // adjoint right-hand side into the PCG residual buffer
    PcgWork<double> &w = m->w64;
    k_fill_zero<<<grid_for(nv, 256), 256, 0, st>>>(w.r.get(), nv);
    for (int c = 0; c < m->ncolours; ++c) {
        const int64_t e0 = m->colour_off[c], e1 = m->colour_off[c + 1];
        if (e1 <= e0) continue;
        k_adjoint_rhs<<<grid_for(e1 - e0, 256), 256, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(),
                                                              m->target.get(), m->phi.get(), Hstar[0], Hstar[1],
                                                              Hstar[2], w.r.get(), e0, e1);
    }
    FK_HIP(hipGetLastError());
    // A lambda = rhs on the free rows, lambda = 0 on the fixed nodes: FP64
    // PCG into w.x, or the FP32-refined solve into inc64
    const double *lam = w.x.get();
    int64_t its = 0;
    if (m->inner_precision == FK_PREC_F32_REFINED) {
        m->inc64.ensure(nv);
        const int s32 = solve_f32_increment(m, m->inc64.get(), w.r.get(), rtol, 1000000, &its);
        if (s32 != FK_OK) fail(s32, std::string("adjoint solve: ") + fk_status_string(s32));
        lam = m->inc64.get();
    } else {
        const CgState s = run_pcg_f64(m, rtol, 1000000, -1.0);
        if (s.status != FK_OK) fail(s.status, std::string("adjoint solve: ") + fk_status_string(s.status));
        its = s.iters;
    }
    DevBuf<double> grad(nt);
    k_adjoint_grad<<<grid_for(nt, 256), 256, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(),
                                                      m->perm.get(), m->phi.get(), lam, nt, grad.get());
    FK_HIP(hipGetLastError());
    FK_HIP(hipMemcpyAsync(grad_host, grad.get(), nt * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    if (J_out) *J_out = J;
    if (iters_out) *iters_out = its;
}

}  // namespace fk

using fk::guard;

extern "C" {

int fk_adjoint_field_target(fk_mesh *m, const uint8_t *target, const double *Hstar, double rtol, double *J,
                            double *dJdmu, int64_t *iterations) {
    if (!m || !target || !Hstar || !dJdmu || !(rtol > 0.0)) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_adjoint_field_target", [&] {
        fk::set_device(m->ctx);
        fk::adjoint_field_target(m, target, Hstar, rtol, J, dJdmu, iterations);
    });
}

}  // extern "C"
