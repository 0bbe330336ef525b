// Demagnetizing field of the micromagnetics driver: magnetostaticField of
// Micromagnetics/LL.jl:162-209, one call per Landau-Lifshitz time step with
// the matrix fixed (A = stiffnessMatrix(mesh), Micromagnetics/permalloy.jl:99)
// and a new right-hand side each time:
//   k_demag_load (one launch per colour, no two elements of a colour share a
//       node): RHS_i += VE_k grad(lambda_i) . mean(M over the 4 nodes of k)
//       over the magnetic elements k;
//   the projected Neumann Jacobi-PCG (A has no fixed node; the load sums to
//       zero) from zero, as the reference's cg(A, RHS);
//   k_demag_nodal (per colour): H_k = -grad(u) and the volume-weighted sum
//       VE_k H_k into the nodes of the magnetic elements;
//   k_div_volumes: division by the caller's nodal volumes (0/0 = NaN on
//       nodes outside the magnet, as the reference).
// The matrix, its solver copy and the Jacobi diagonal stay resident across
// calls; each call moves M in and the nodal field out. With warm != 0 the
// PCG starts from the previous call's potential (kept on the device): the
// field of the next time step is close to the last one's, and the stopping
// test stays relative to ||RHS||, as for the start from zero.
#include <cmath>

#include "device.h"
#include "internal.h"

namespace fk {

__global__ void k_mask_to_colour(const uint8_t *__restrict__ src, const int32_t *__restrict__ perm, int64_t nt,
                                 uint8_t *__restrict__ dst) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < nt) dst[e] = src[perm[e]] ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_demag_load(const int4 *__restrict__ tet, const double *__restrict__ X,
                                                    const double *__restrict__ Y, const double *__restrict__ Z,
                                                    const uint8_t *__restrict__ inside,
                                                    const double *__restrict__ M, double *__restrict__ rhs,
                                                    int64_t e0, int64_t e1) {
    const int64_t e = e0 + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= e1 || !inside[e]) return;
    const int4 nd = tet[e];
    const Grad G = element_gradients(X, Y, Z, nd);
    const double vol = fabs(G.det) * (1.0 / 6.0);
    double f[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        f[c] = (M[3 * static_cast<int64_t>(nd.x) + c] + M[3 * static_cast<int64_t>(nd.y) + c] +
                M[3 * static_cast<int64_t>(nd.z) + c] + M[3 * static_cast<int64_t>(nd.w) + c]) /
               4.0;
    const int n[4] = {nd.x, nd.y, nd.z, nd.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) rhs[n[i]] += vol * (G.g[i][0] * f[0] + G.g[i][1] * f[1] + G.g[i][2] * f[2]);
}

__global__ __launch_bounds__(256) void k_demag_nodal(const int4 *__restrict__ tet, const double *__restrict__ X,
                                                     const double *__restrict__ Y, const double *__restrict__ Z,
                                                     const uint8_t *__restrict__ inside,
                                                     const double *__restrict__ u, double *__restrict__ Hd,
                                                     int64_t e0, int64_t e1) {
    const int64_t e = e0 + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= e1 || !inside[e]) return;
    const int4 nd = tet[e];
    const Grad G = element_gradients(X, Y, Z, nd);
    const double vol = fabs(G.det) * (1.0 / 6.0);
    const double u0 = u[nd.x], u1 = u[nd.y], u2 = u[nd.z], u3 = u[nd.w];
    double h[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) h[c] = -(u0 * G.g[0][c] + u1 * G.g[1][c] + u2 * G.g[2][c] + u3 * G.g[3][c]);
    const int n[4] = {nd.x, nd.y, nd.z, nd.w};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) Hd[3 * static_cast<int64_t>(n[i]) + c] += vol * h[c];
}

__global__ void k_div_volumes(double *__restrict__ Hd, const double *__restrict__ vol, int64_t nv) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < 3 * nv) Hd[t] = Hd[t] / vol[t / 3];
}

// input checks on the device: bad[0] counts non-finite magnetization
// entries, bad[1] non-finite or negative nodal volumes
__global__ void k_demag_check(const double *__restrict__ M, const double *__restrict__ vol, int64_t nv,
                              int32_t *__restrict__ bad) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < 3 * nv && !isfinite(M[t])) atomicAdd(bad, 1);
    if (t < nv && !(isfinite(vol[t]) && vol[t] >= 0.0)) atomicAdd(bad + 1, 1);
}

void demag_field(fk_mesh *m, const uint8_t *inside_host, const double *M_host, const double *vol_host, double rtol,
                 int64_t max_iter, int warm, double *Hd_host, int64_t *iters_out, double *relres_out) {
    if (!m->have_values) fail(FK_ERR_STATE, "matrix not assembled (A = stiffnessMatrix(mesh))");
    const int64_t nv = m->nv, nt = m->nt;
    hipStream_t st = m->ctx->stream;
    ensure_work_f64(m);
    m->demag_in_caller.ensure(nt);
    m->demag_inside.ensure(nt);
    m->demag_M.ensure(3 * nv);
    m->demag_vol.ensure(nv);
    m->demag_Hd.ensure(3 * nv);
    DevBuf<uint8_t> &in_caller = m->demag_in_caller, &inside = m->demag_inside;
    DevBuf<double> &M = m->demag_M, &vol = m->demag_vol, &Hd = m->demag_Hd;
    const NodesIn<double> M_in(m, M_host, 3), vol_in(m, vol_host);
    FK_HIP(hipMemcpyAsync(in_caller.get(), inside_host, nt, hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(M.get(), M_in.ptr, 3 * nv * sizeof(double), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(vol.get(), vol_in.ptr, nv * sizeof(double), hipMemcpyHostToDevice, st));
    {
        // checked before anything the call keeps (the previous potential) changes
        DevBuf<int32_t> bad(2);
        FK_HIP(hipMemsetAsync(bad.get(), 0, 2 * sizeof(int32_t), st));
        k_demag_check<<<grid_for(3 * nv, 256), 256, 0, st>>>(M.get(), vol.get(), nv, bad.get());
        FK_HIP(hipGetLastError());
        int32_t h[2] = {0, 0};
        FK_HIP(hipMemcpyAsync(h, bad.get(), 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        FK_HIP(hipStreamSynchronize(st));
        if (h[0]) fail(FK_ERR_NONFINITE, "non-finite magnetization");
        if (h[1]) fail(FK_ERR_INVALID_ARG, "nodal volumes");
    }
    k_mask_to_colour<<<grid_for(nt, 256), 256, 0, st>>>(in_caller.get(), m->perm.get(), nt, inside.get());
    double *rhs = m->w64.r.get();
    FK_HIP(hipMemsetAsync(rhs, 0, nv * sizeof(double), st));
    for (int c = 0; c < m->ncolours; ++c) {
        const int64_t e0 = m->colour_off[c], e1 = m->colour_off[c + 1];
        if (e1 <= e0) continue;
        k_demag_load<<<grid_for(e1 - e0, 256), 256, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(),
                                                             inside.get(), M.get(), rhs, e0, e1);
    }
    FK_HIP(hipGetLastError());
    const bool from_prev = warm && m->have_demag_u;
    const NeumannResult res = solve_neumann_dev(m, rtol, max_iter, from_prev ? m->demag_u.get() : nullptr);
    m->demag_u.ensure(nv);
    FK_HIP(hipMemcpyAsync(m->demag_u.get(), m->w64.x.get(), nv * sizeof(double), hipMemcpyDeviceToDevice, st));
    m->have_demag_u = true;
    FK_HIP(hipMemsetAsync(Hd.get(), 0, 3 * nv * sizeof(double), st));
    for (int c = 0; c < m->ncolours; ++c) {
        const int64_t e0 = m->colour_off[c], e1 = m->colour_off[c + 1];
        if (e1 <= e0) continue;
        k_demag_nodal<<<grid_for(e1 - e0, 256), 256, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(),
                                                              inside.get(), m->w64.x.get(), Hd.get(), e0, e1);
    }
    k_div_volumes<<<grid_for(3 * nv, 256), 256, 0, st>>>(Hd.get(), vol.get(), nv);
    FK_HIP(hipGetLastError());
    NodesOut<double> Hd_out(m, Hd_host, 3);
    FK_HIP(hipMemcpyAsync(Hd_out.ptr, Hd.get(), 3 * nv * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    Hd_out.commit();
    if (iters_out) *iters_out = res.iters;
    if (relres_out) *relres_out = res.relres;
}

}  // namespace fk

extern "C" {

int fk_demag_field(fk_mesh *m, const uint8_t *inside, const double *M, const double *volumes, double rtol,
                   int64_t max_iter, int32_t warm, double *Hd_nodes, int64_t *iterations, double *relres) {
    if (!m || !inside || !M || !volumes || !Hd_nodes || !(rtol > 0.0) || max_iter < 1) return FK_ERR_INVALID_ARG;
    return fk::guard(m->ctx, "fk_demag_field", [&] {
        fk::set_device(m->ctx);
        fk::demag_field(m, inside, M, volumes, rtol, max_iter, warm, Hd_nodes, iterations, relres);
    });
}

}  // extern "C"
