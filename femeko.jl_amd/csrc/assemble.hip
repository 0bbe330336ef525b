// Element kernels: value-only stiffness assembly, local stiffness export,
// Newton tangent, Lagrange vector.
//
// One thread per tetrahedron. Node coordinates are read from SoA arrays, the
// basis gradients and volume come from one 3x3 determinant (closed form of
// Femeko's abcd, src/FEM3D.jl:5-17), the 10 unique entries of
// K = VE f (b b' + c c' + d d') (src/FEM3D.jl:59-76) are formed in registers
// and added into the CSR values through the element's 16 precomputed slots.
// Elements are stored colour by colour; a launch per colour makes every add
// conflict free (no atomics, deterministic, and entry (i,j) and (j,i) receive
// identical sums, so the matrix is exactly symmetric). The atomic variant
// covers all elements in one launch with hardware FP64 atomics.
//
// The gather variant turns the scatter around: one thread per matrix row
// walks the elements incident to its node (in element order), recomputes each
// element's gradients and adds the element's row of K into the row's values
// held in LDS, then the block writes its rows' values once, coalesced. No
// colour passes (each colour pass of the scatter touches nearly every cache
// line of the value array), no atomics; entry (i,j) and (j,i) receive the
// same terms in the same (element) order, so the matrix stays exactly
// symmetric and runs are bitwise reproducible.
#include <algorithm>
#include <thread>

#include "device.h"
#include "internal.h"

namespace fk {

__global__ void k_fill_f64(double *__restrict__ a, int64_t n, double v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_f64_to_f32(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = static_cast<float>(a[i]);
}

// coef[e] = f[perm[e]]; counts non-finite entries
__global__ void k_gather_coef(const double *__restrict__ f, const int32_t *__restrict__ perm, int64_t nt,
                              double *__restrict__ coef, int32_t *__restrict__ bad) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nt) return;
    const double v = f[perm[e]];
    coef[e] = v;
    if (!isfinite(v)) atomicAdd(bad, 1);
}

// Local matrix of element e, 10 unique values in the order
// (0,0) (0,1) (0,2) (0,3) (1,1) (1,2) (1,3) (2,2) (2,3) (3,3).
__device__ __forceinline__ void local_values(const Grad &G, double w, double k[10]) {
    int q = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j) k[q++] = w * dot3(G.g[i], G.g[j]);
}

template <bool ATOMIC>
__device__ __forceinline__ void scatter16(double *__restrict__ val, const int32_t *__restrict__ slot, int64_t e,
                                          const double k[10]) {
    const int4 *s4 = reinterpret_cast<const int4 *>(slot + 16 * e);
    const int4 r0 = s4[0], r1 = s4[1], r2 = s4[2], r3 = s4[3];
    const int s[16] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w,
                       r2.x, r2.y, r2.z, r2.w, r3.x, r3.y, r3.z, r3.w};
    // upper-triangle index of (i,j)
    constexpr int U[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double v = k[U[i][j]];
            if (ATOMIC)
                unsafeAtomicAdd(&val[s[4 * i + j]], v);
            else
                val[s[4 * i + j]] += v;
        }
}

template <bool ATOMIC>
__global__ __launch_bounds__(256) void k_assemble(const int4 *__restrict__ tet, const double *__restrict__ X,
                                                  const double *__restrict__ Y, const double *__restrict__ Z,
                                                  const double *__restrict__ coef, const int32_t *__restrict__ slot,
                                                  double *__restrict__ val, int64_t e0, int64_t e1) {
    const int64_t e = e0 + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= e1) return;
    const Grad G = element_gradients(X, Y, Z, tet[e]);
    const double w = fabs(G.det) * (1.0 / 6.0) * coef[e];
    double k[10];
    local_values(G, w, k);
    scatter16<ATOMIC>(val, slot, e, k);
}

// Row-gather assembly: thread = row i (one wavefront = one slice of 64
// rows, kGatherRows rows per block). The incidences of slice s are stored
// transposed from sptr[s] (entry k of lane l at sptr[s] + 64 k + l, width[s]
// entries per lane, padding .x = kNoInc) so each load is coalesced; an
// incidence is one 8-byte word pair: .x = element << 2 | local index of the
// row's node, elements in caller order, ascending per node; .y packs the
// positions (0..255, relative to rowptr[i]) of the element's 4 nodes in row
// i. Elements are read in caller order, whose spatial locality (neighbouring
// rows share most elements) keeps them in L2. The block's value segment
// [rowptr[i0], rowptr[i0 + rows]) is accumulated in LDS (dynamic size) and
// written once.
//
// Node coordinates come from 32-byte records (x, y, z, w): two loads per
// node instead of three, and the record of the row's own node (local index
// la of every incidence) is loaded once per row, so an incidence loads three
// records. The cost of this kernel is the number of per-lane addresses its
// gathers issue more than their bytes: at n=160 (24.6 M elements) the
// separate X, Y, Z arrays took 1.21 ms, records 1.05 ms, records without the
// own node 1.01 ms, with bit-identical values (a numbering without locality
// pays for the records' larger footprint, see below). Staging the coordinates of the
// block's columns in LDS instead (no per-incidence coordinate gathers, 32
// bytes of LDS per value) took 1.27-1.88 ms: the LDS then holds about 5
// waves per CU.
//
// The incidences are taken kGatherUnroll at a time with every load
// unconditional (a padding entry or one past the width reads element 0 and
// adds nothing): the loads of the group (incidence words, then element nodes
// and coefficients, then coordinates) are in flight together, and no branch
// sits between a load and its use (a branch there makes the compiler wait
// for all outstanding loads). With records, the own node and the 8-byte
// incidence, 8 per group (210 VGPRs) measured 1.00 ms, 6 0.995 ms, 4 1.07
// ms; on the same mesh with its nodes numbered by reverse Cuthill-McKee
// (elements in cell order, so element and node order no longer follow each
// other) 1.88, 1.95 and 2.33 ms, where the separate X, Y, Z arrays took
// 1.76 ms. The sums run in the same (element) order either way.
constexpr uint32_t kNoInc = 0xffffffffu;
constexpr int kGatherUnroll = 8;
constexpr int kNewtonUnroll = 4;

// The 12 coordinates (and the 4 w fields when W) of an element's nodes in
// local order, node la being the row's own node whose record the caller
// holds: three record loads.
template <bool W>
__device__ __forceinline__ void element_records(const double4 *__restrict__ P, int4 nd, int la, const double4 &own,
                                                double c[12], double w[4]) {
    const double4 L0 = P[la == 0 ? nd.y : nd.x];
    const double4 L1 = P[la <= 1 ? nd.z : nd.y];
    const double4 L2 = P[la <= 2 ? nd.w : nd.z];
    // node b < la is other node b, node b > la other node b - 1
    c[0] = la == 0 ? own.x : L0.x;
    c[1] = la == 0 ? own.y : L0.y;
    c[2] = la == 0 ? own.z : L0.z;
    c[3] = la == 1 ? own.x : la == 0 ? L0.x : L1.x;
    c[4] = la == 1 ? own.y : la == 0 ? L0.y : L1.y;
    c[5] = la == 1 ? own.z : la == 0 ? L0.z : L1.z;
    c[6] = la == 2 ? own.x : la <= 1 ? L1.x : L2.x;
    c[7] = la == 2 ? own.y : la <= 1 ? L1.y : L2.y;
    c[8] = la == 2 ? own.z : la <= 1 ? L1.z : L2.z;
    c[9] = la == 3 ? own.x : L2.x;
    c[10] = la == 3 ? own.y : L2.y;
    c[11] = la == 3 ? own.z : L2.z;
    if constexpr (W) {
        w[0] = la == 0 ? own.w : L0.w;
        w[1] = la == 1 ? own.w : la == 0 ? L0.w : L1.w;
        w[2] = la == 2 ? own.w : la <= 1 ? L1.w : L2.w;
        w[3] = la == 3 ? own.w : L2.w;
    }
}

__global__ __launch_bounds__(kGatherRows) void k_assemble_gather(
    const int32_t *__restrict__ rowptr, const int64_t *__restrict__ sptr, const int32_t *__restrict__ width,
    const uint2 *__restrict__ inc, const int4 *__restrict__ tet, const double4 *__restrict__ P,
    const double *__restrict__ coef, int64_t nv, double *__restrict__ val) {
    extern __shared__ double acc[];
    constexpr int UG = kGatherUnroll;
    // XCD-aware: each XCD takes a contiguous eighth of the rows, so the
    // elements shared by rows one grid line or plane apart stay in its L2
    const int64_t i0 = static_cast<int64_t>(logical_block(static_cast<int>(gridDim.x))) * kGatherRows;
    if (i0 >= nv) return;  // padding of the grid to a multiple of 8 (whole block)
    const int64_t i = i0 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int64_t ns = (nv + 63) >> 6;
    const int64_t s = i >> 6;  // uniform over the wave
    const int64_t sc = min(s, ns - 1);
    const int32_t base = rowptr[i0];
    const int32_t seg = rowptr[min(i0 + kGatherRows, nv)] - base;
    for (int32_t k = threadIdx.x; k < seg; k += kGatherRows) acc[k] = 0.0;
    __syncthreads();
    double *a = acc + (rowptr[min(i, nv)] - base);
    const double4 own = P[min(i, nv - 1)];
    const int64_t sb = sptr[sc] + lane;
    const int W = s < ns ? width[sc] : 0;
    constexpr int U[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
    for (int k = 0; k < W; k += UG) {
        uint2 w[UG];
#pragma unroll
        for (int u = 0; u < UG; ++u) w[u] = inc[sb + 64 * static_cast<int64_t>(min(k + u, W - 1))];
        bool use[UG];
        int4 nd[UG];
        double cf[UG];
#pragma unroll
        for (int u = 0; u < UG; ++u) {
            use[u] = k + u < W && w[u].x != kNoInc;
            const int64_t e = use[u] ? static_cast<int64_t>(w[u].x >> 2) : 0;
            nd[u] = tet[e];
            cf[u] = coef[e];
        }
        double c[UG][12];
#pragma unroll
        for (int u = 0; u < UG; ++u) element_records<false>(P, nd[u], static_cast<int>(w[u].x & 3u), own, c[u], nullptr);
#pragma unroll
        for (int u = 0; u < UG; ++u) {
            const Grad G = gradients_of(c[u]);
            const double sc6 = fabs(G.det) * (1.0 / 6.0) * cf[u];
            // all 10 unique local values by one fixed instruction sequence;
            // row i and row j then add the very same value for (i,j), (j,i)
            double kv[10];
            local_values(G, sc6, kv);
            const int la = static_cast<int>(w[u].x & 3u);
            if (use[u]) {
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const double v = la == 0 ? kv[U[0][b]] : la == 1 ? kv[U[1][b]] : la == 2 ? kv[U[2][b]] : kv[U[3][b]];
                    a[(w[u].y >> (8 * b)) & 255u] += v;
                }
            }
        }
    }
    __syncthreads();
    for (int32_t k = threadIdx.x; k < seg; k += kGatherRows) val[base + k] = acc[k];
}

// Newton-step assembly of the nonlinear loop in row-gather form (the
// incidence lists and node records of k_assemble_gather, kNewtonUnroll
// incidences per step in the same way; the records' w holds phi): row i
// adds, over its elements e (local index la), K_e(mu)[la][b] =
// VE mu_e g_la.g_b and, on a Newton step, the tangent
// VE mu'_e/|H_e| (H_e.g_la)(H_e.g_b) of src/FEM3D.jl:79-119, with
// H_e = -grad(phi); and sums the element residual -sum_b K_e(mu)[la][b] phi_b
// into r[i]. Values and residual are written once (no zero fill); g_la.g_b
// and (H.g_la)(H.g_b) are formed in an order that does not depend on which of
// the two rows computes them, so the matrix is exactly symmetric. mu and dmu
// are in caller element order.
__global__ __launch_bounds__(kGatherRows) void k_assemble_newton_gather(
    const int32_t *__restrict__ rowptr, const int64_t *__restrict__ sptr, const int32_t *__restrict__ width,
    const uint2 *__restrict__ inc, const int4 *__restrict__ tet, const double4 *__restrict__ P,
    const double *__restrict__ mu, const double *__restrict__ dmu, int newton, int64_t nv,
    double *__restrict__ val, double *__restrict__ r) {
    extern __shared__ double acc[];
    constexpr int UN = kNewtonUnroll;
    const int64_t i0 = static_cast<int64_t>(logical_block(static_cast<int>(gridDim.x))) * kGatherRows;
    if (i0 >= nv) return;  // padding of the grid to a multiple of 8 (whole block)
    const int64_t i = i0 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int64_t ns = (nv + 63) >> 6;
    const int64_t s = i >> 6;  // uniform over the wave
    const int64_t sc = min(s, ns - 1);
    const int32_t base = rowptr[i0];
    const int32_t seg = rowptr[min(i0 + kGatherRows, nv)] - base;
    for (int32_t k = threadIdx.x; k < seg; k += kGatherRows) acc[k] = 0.0;
    __syncthreads();
    double res = 0.0;
    double *a = acc + (rowptr[min(i, nv)] - base);
    const double4 own = P[min(i, nv - 1)];
    const int64_t sb = sptr[sc] + lane;
    const int W = s < ns ? width[sc] : 0;
    for (int k = 0; k < W; k += UN) {
        uint2 w[UN];
#pragma unroll
        for (int u = 0; u < UN; ++u) w[u] = inc[sb + 64 * static_cast<int64_t>(min(k + u, W - 1))];
        bool use[UN];
        int4 nd[UN];
        double m1[UN], d1[UN];
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            use[u] = k + u < W && w[u].x != kNoInc;
            const int64_t e = use[u] ? static_cast<int64_t>(w[u].x >> 2) : 0;
            nd[u] = tet[e];
            m1[u] = mu[e];
            d1[u] = newton ? dmu[e] : 0.0;
        }
        double c[UN][12], ph[UN][4];
#pragma unroll
        for (int u = 0; u < UN; ++u) element_records<true>(P, nd[u], static_cast<int>(w[u].x & 3u), own, c[u], ph[u]);
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            const Grad G = gradients_of(c[u]);
            const double vol = fabs(G.det) * (1.0 / 6.0);
            const double wk = vol * m1[u];
            const int la = static_cast<int>(w[u].x & 3u);
            double gl[3];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
                gl[cc] = la == 0 ? G.g[0][cc] : la == 1 ? G.g[1][cc] : la == 2 ? G.g[2][cc] : G.g[3][cc];
            double kv[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) kv[b] = wk * dot3(gl, G.g[b]);
            const double ru = kv[0] * ph[u][0] + kv[1] * ph[u][1] + kv[2] * ph[u][2] + kv[3] * ph[u][3];
            if (newton) {
                double h[3];
#pragma unroll
                for (int cc = 0; cc < 3; ++cc)
                    h[cc] = -(ph[u][0] * G.g[0][cc] + ph[u][1] * G.g[1][cc] + ph[u][2] * G.g[2][cc] +
                              ph[u][3] * G.g[3][cc]);
                const double hn = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
                const double wt = hn > 0.0 ? vol * d1[u] / hn : 0.0;
                double hg[4];
#pragma unroll
                for (int b = 0; b < 4; ++b) hg[b] = dot3(h, G.g[b]);
                const double hl = la == 0 ? hg[0] : la == 1 ? hg[1] : la == 2 ? hg[2] : hg[3];
#pragma unroll
                for (int b = 0; b < 4; ++b) kv[b] += wt * (hl * hg[b]);
            }
            if (use[u]) {
                res -= ru;
#pragma unroll
                for (int b = 0; b < 4; ++b) a[(w[u].y >> (8 * b)) & 255u] += kv[b];
            }
        }
    }
    __syncthreads();
    for (int32_t k = threadIdx.x; k < seg; k += kGatherRows) val[base + k] = acc[k];
    if (i < nv) r[i] = res;
}

// Blocks of the row-gather kernels: one per kGatherRows rows, padded to a
// multiple of 8 so that logical_block gives each XCD a contiguous eighth of
// the rows (the padding blocks return at once).
static int64_t gather_blocks(int64_t nv) {
    const int64_t nb = (nv + kGatherRows - 1) / kGatherRows;
    return nb < 8 ? nb : (nb + 7) / 8 * 8;
}

// P[i] = (X[i], Y[i], Z[i], 0): the row-gather kernels' node records
__global__ void k_pack_xyzw(const double *__restrict__ X, const double *__restrict__ Y, const double *__restrict__ Z,
                            int64_t nv, double4 *__restrict__ P) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nv) P[i] = make_double4(X[i], Y[i], Z[i], 0.0);
}

// P[i].w = phi[i]
__global__ void k_records_w(const double *__restrict__ phi, int64_t nv, double4 *__restrict__ P) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nv) P[i].w = phi[i];
}

// coef_nat[perm[e]] = coef[e] (colour order -> caller order)
__global__ void k_coef_to_caller(const double *__restrict__ coef, const int32_t *__restrict__ perm, int64_t nt,
                                 double *__restrict__ coef_nat) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < nt) coef_nat[perm[e]] = coef[e];
}

// Newton tangent of src/FEM3D.jl:79-119:
// At_ij = VE dmu/|H| (H . grad_i)(H . grad_j), H and dmu in caller order.
__global__ __launch_bounds__(256) void k_tangent(const int4 *__restrict__ tet, const double *__restrict__ X,
                                                 const double *__restrict__ Y, const double *__restrict__ Z,
                                                 const int32_t *__restrict__ perm, const double *__restrict__ H,
                                                 const double *__restrict__ dmu, const int32_t *__restrict__ slot,
                                                 double *__restrict__ val, int64_t e0, int64_t e1) {
    const int64_t e = e0 + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= e1) return;
    const Grad G = element_gradients(X, Y, Z, tet[e]);
    const int64_t o = perm[e];
    const double h[3] = {H[3 * o], H[3 * o + 1], H[3 * o + 2]};
    const double hn = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    const double w = hn > 0.0 ? fabs(G.det) * (1.0 / 6.0) * dmu[o] / hn : 0.0;
    double hg[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) hg[i] = dot3(h, G.g[i]);
    double k[10];
    int q = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j) k[q++] = w * hg[i] * hg[j];
    scatter16<false>(val, slot, e, k);
}

// Ak[:,k] = vec(VE f (b b' + c c' + d d')), 16 x nt column-major, caller order
__global__ __launch_bounds__(256) void k_local_stiffness(const int4 *__restrict__ tet, const double *__restrict__ X,
                                                         const double *__restrict__ Y, const double *__restrict__ Z,
                                                         const int32_t *__restrict__ perm,
                                                         const double *__restrict__ coef, int64_t nt,
                                                         double *__restrict__ Ak) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nt) return;
    const Grad G = element_gradients(X, Y, Z, tet[e]);
    const double w = fabs(G.det) * (1.0 / 6.0) * (coef ? coef[e] : 1.0);
    double k[10];
    local_values(G, w, k);
    constexpr int U[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
    double *out = Ak + 16 * static_cast<int64_t>(perm[e]);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) out[i + 4 * j] = k[U[i][j]];
}

// C[nds] += VE/4 (src/FEM3D.jl:274-281), one colour per launch
__global__ __launch_bounds__(256) void k_lagrange(const int4 *__restrict__ tet, const double *__restrict__ X,
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

// ---------------------------------------------------------------------------
void upload_coefficient(fk_mesh *m, const double *f) {
    hipStream_t st = m->ctx->stream;
    m->coef_nat.ensure(m->nt);
    DevBuf<double> &df = m->coef_nat;  // caller order, read by the row-gather assembly
    DevBuf<int32_t> bad(1);
    FK_HIP(hipMemcpyAsync(df.get(), f, m->nt * sizeof(double), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), st));
    m->coef.ensure(m->nt);
    k_gather_coef<<<grid_for(m->nt, 256), 256, 0, st>>>(df.get(), m->perm.get(), m->nt, m->coef.get(), bad.get());
    FK_HIP(hipGetLastError());
    int32_t h = 0;
    FK_HIP(hipMemcpyAsync(&h, bad.get(), sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    if (h) {
        m->have_coef = false;
        fail(FK_ERR_NONFINITE, std::to_string(h) + " non-finite coefficient value(s)");
    }
    m->have_coef = true;
    m->coef_nat_valid = true;
    m->coef_g_valid = false;
}

// The row-gather kernels' elements: the caller's order, or the sorted copy
// when the caller's elements are not sorted by their lowest node.
static const int4 *gather_tet(const fk_mesh *m) { return m->gather_sorted ? m->tet_g.get() : m->tet_nat.get(); }

// Colour order -> the row-gather kernels' element order (the caller's, or
// the sorted one).
static void to_gather_order(fk_mesh *m, const double *src, double *dst) {
    k_coef_to_caller<<<grid_for(m->nt, 256), 256, 0, m->ctx->stream>>>(
        src, m->gather_sorted ? m->perm_g.get() : m->perm.get(), m->nt, dst);
    FK_HIP(hipGetLastError());
}

// Incidence lists of the row-gather assembly (host build, once per mesh).
// Unsupported (gather_ok = false) when a row has more than 256 entries or a
// block's value segment exceeds kGatherLdsMax doubles.
void build_gather(fk_mesh *m) {
    if (m->gather_built) return;
    m->gather_built = true;
    m->gather_ok = false;
    hipStream_t st = m->ctx->stream;
    const int64_t nv = m->nv, nt = m->nt, nnz = m->nnz;
    if (nt >= (int64_t(1) << 30)) return;
    std::vector<int4> tet(static_cast<size_t>(nt));
    std::vector<int32_t> rp(static_cast<size_t>(nv + 1)), ci(static_cast<size_t>(nnz));
    FK_HIP(hipMemcpyAsync(tet.data(), gather_tet(m), nt * sizeof(int4), hipMemcpyDeviceToHost, st));
    FK_HIP(hipMemcpyAsync(rp.data(), m->rowptr.get(), (nv + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipMemcpyAsync(ci.data(), m->colidx.get(), nnz * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < nv; ++i)
        if (rp[i + 1] - rp[i] > 256) return;
    int64_t seg_max = 0;
    for (int64_t i0 = 0; i0 < nv; i0 += kGatherRows)
        seg_max = std::max<int64_t>(seg_max, rp[std::min(i0 + kGatherRows, nv)] - rp[i0]);
    if (seg_max > kGatherLdsMax) return;
    // node -> incidences, ascending caller element order
    std::vector<int64_t> ptr(static_cast<size_t>(nv + 1), 0);
    for (int64_t e = 0; e < nt; ++e) {
        const int4 t = tet[e];
        ++ptr[t.x + 1];
        ++ptr[t.y + 1];
        ++ptr[t.z + 1];
        ++ptr[t.w + 1];
    }
    for (int64_t i = 0; i < nv; ++i) ptr[i + 1] += ptr[i];
    std::vector<uint32_t> incl(static_cast<size_t>(4 * nt));
    {
        std::vector<int64_t> cur(ptr.begin(), ptr.end() - 1);
        for (int64_t e = 0; e < nt; ++e) {
            const int4 t = tet[e];
            const int32_t nd[4] = {t.x, t.y, t.z, t.w};
            for (int a = 0; a < 4; ++a)
                incl[cur[nd[a]]++] = static_cast<uint32_t>(e << 2) | static_cast<uint32_t>(a);
        }
    }
    // slices of 64 rows, transposed
    const int64_t ns = (nv + 63) / 64;
    std::vector<int32_t> width(static_cast<size_t>(ns));
    std::vector<int64_t> sptr(static_cast<size_t>(ns + 1), 0);
    for (int64_t s = 0; s < ns; ++s) {
        int64_t w = 0;
        for (int64_t r = s * 64; r < std::min(nv, s * 64 + 64); ++r) w = std::max(w, ptr[r + 1] - ptr[r]);
        width[s] = static_cast<int32_t>(w);
        sptr[s + 1] = sptr[s] + 64 * w;
    }
    const int64_t len = sptr[ns];
    std::vector<uint2> inc(static_cast<size_t>(len), make_uint2(kNoInc, 0u));
    const int nth = static_cast<int>(std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency())));
    std::vector<std::thread> th;
    for (int q = 0; q < nth; ++q)
        th.emplace_back([&, q] {
            for (int64_t s = ns * q / nth; s < ns * (q + 1) / nth; ++s)
                for (int64_t r = s * 64; r < std::min(nv, s * 64 + 64); ++r) {
                    const int lane = static_cast<int>(r - s * 64);
                    const int32_t *b = ci.data() + rp[r], *en = ci.data() + rp[r + 1];
                    for (int64_t k = ptr[r]; k < ptr[r + 1]; ++k) {
                        const uint32_t w = incl[k];
                        const int4 t = tet[w >> 2];
                        const int32_t nd[4] = {t.x, t.y, t.z, t.w};
                        uint32_t o = 0;
                        for (int c = 0; c < 4; ++c) {
                            const int32_t *f = std::lower_bound(b, en, nd[c]);
                            o |= static_cast<uint32_t>(f - b) << (8 * c);
                        }
                        const int64_t pos = sptr[s] + 64 * (k - ptr[r]) + lane;
                        inc[pos] = make_uint2(w, o);
                    }
                }
        });
    for (auto &x : th) x.join();
    m->inc_sptr.alloc(ns + 1);
    m->inc_width.alloc(ns);
    m->inc.alloc(len);
    m->xyzw.alloc(nv);
    FK_HIP(hipMemcpyAsync(m->inc_sptr.get(), sptr.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(m->inc_width.get(), width.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(m->inc.get(), inc.data(), len * sizeof(uint2), hipMemcpyHostToDevice, st));
    k_pack_xyzw<<<grid_for(nv, 256), 256, 0, st>>>(m->X.get(), m->Y.get(), m->Z.get(), nv, m->xyzw.get());
    FK_HIP(hipGetLastError());
    FK_HIP(hipStreamSynchronize(st));
    m->gather_lds = static_cast<int32_t>(seg_max);
    m->gather_ok = true;
}

// Enqueue the nonlinear loop's Newton-step assembly in row-gather form into
// val64 and r (mu, dmu in colour order, converted here). Returns false, with
// nothing enqueued, when the mesh's rows are too long for the LDS segment;
// the caller then runs the coloured element kernel.
bool assemble_newton_gather(fk_mesh *m, const double *mu, const double *dmu, const double *phi, int newton,
                            double *r) {
    build_gather(m);
    if (!m->gather_ok) return false;
    hipStream_t st = m->ctx->stream;
    // mu and dmu in the kernels' element order (coef_nat or coef_g holds mu
    // until the loop rewrites coef right after this assembly)
    DevBuf<double> &mu_g = m->gather_sorted ? m->coef_g : m->coef_nat;
    mu_g.ensure(m->nt);
    m->dmu_nat.ensure(m->nt);
    to_gather_order(m, mu, mu_g.get());
    if (newton) to_gather_order(m, dmu, m->dmu_nat.get());
    m->coef_nat_valid = false;
    m->coef_g_valid = false;
    k_records_w<<<grid_for(m->nv, 256), 256, 0, st>>>(phi, m->nv, m->xyzw.get());
    const int64_t nb = gather_blocks(m->nv);
    k_assemble_newton_gather<<<static_cast<unsigned>(nb), kGatherRows,
                               static_cast<size_t>(std::max(m->gather_lds, 1)) * sizeof(double), st>>>(
        m->rowptr.get(), m->inc_sptr.get(), m->inc_width.get(), m->inc.get(), gather_tet(m), m->xyzw.get(),
        mu_g.get(), m->dmu_nat.get(), newton, m->nv, m->val64.get(), r);
    FK_HIP(hipGetLastError());
    return true;
}

// The coefficient in caller element order (coef_nat), converted from the
// colour-ordered copy when the device last wrote that one. With
// `may_allocate` false it returns nullptr instead of allocating the buffer.
const double *coef_caller_order(fk_mesh *m, bool may_allocate) {
    if (!m->coef_nat_valid) {
        if (m->coef_nat.n < static_cast<size_t>(m->nt)) {
            if (!may_allocate) return nullptr;
            m->coef_nat.ensure(m->nt);
        }
        k_coef_to_caller<<<grid_for(m->nt, 256), 256, 0, m->ctx->stream>>>(m->coef.get(), m->perm.get(), m->nt,
                                                                          m->coef_nat.get());
        FK_HIP(hipGetLastError());
        m->coef_nat_valid = true;
    }
    return m->coef_nat.get();
}

// The coefficient in the row-gather kernels' element order.
static const double *coef_gather_order(fk_mesh *m) {
    if (!m->gather_sorted) return coef_caller_order(m, true);
    if (!m->coef_g_valid) {
        m->coef_g.ensure(m->nt);
        to_gather_order(m, m->coef.get(), m->coef_g.get());
        m->coef_g_valid = true;
    }
    return m->coef_g.get();
}

// Enqueue the value assembly on the context stream (no host sync).
void assemble(fk_mesh *m, int method) {
    if (!m->have_coef) fail(FK_ERR_STATE, "no coefficient set");
    hipStream_t st = m->ctx->stream;
    m->val64.ensure(m->nnz);
    if (method == FK_ASM_GATHER) {
        build_gather(m);
        if (m->gather_ok) {
            const double *coef = coef_gather_order(m);
            const int64_t nb = gather_blocks(m->nv);
            k_assemble_gather<<<static_cast<unsigned>(nb), kGatherRows,
                                static_cast<size_t>(std::max(m->gather_lds, 1)) * sizeof(double), st>>>(
                m->rowptr.get(), m->inc_sptr.get(), m->inc_width.get(), m->inc.get(), gather_tet(m),
                m->xyzw.get(), coef, m->nv, m->val64.get());
            FK_HIP(hipGetLastError());
            values_changed(m);
            return;
        }
        method = FK_ASM_COLOURED;  // rows too long for the LDS segment
    }
    k_fill_f64<<<grid_for(m->nnz, 256), 256, 0, st>>>(m->val64.get(), m->nnz, 0.0);
    if (method == FK_ASM_ATOMIC) {
        k_assemble<true><<<grid_for(m->nt, 256), 256, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(),
                                                              m->coef.get(), m->slot.get(), m->val64.get(), 0,
                                                              m->nt);
    } else {
        for (int c = 0; c < m->ncolours; ++c) {
            const int64_t e0 = m->colour_off[c], e1 = m->colour_off[c + 1];
            if (e1 <= e0) continue;
            k_assemble<false><<<grid_for(e1 - e0, 256), 256, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(),
                                                                      m->Z.get(), m->coef.get(), m->slot.get(),
                                                                      m->val64.get(), e0, e1);
        }
    }
    FK_HIP(hipGetLastError());
    values_changed(m);
}

void make_val32(fk_mesh *m) {
    if (m->have_val32) return;
    hipStream_t st = m->ctx->stream;
    m->val32.ensure(m->nnz);
    k_f64_to_f32<<<grid_for(m->nnz, 256), 256, 0, st>>>(m->val64.get(), m->val32.get(), m->nnz);
    FK_HIP(hipGetLastError());
    m->have_val32 = true;
}

void local_stiffness(fk_mesh *m, const double *f, double *Ak_host) {
    hipStream_t st = m->ctx->stream;
    DevBuf<double> coef;
    if (f) {
        DevBuf<double> df(m->nt);
        DevBuf<int32_t> bad(1);
        coef.alloc(m->nt);
        FK_HIP(hipMemcpyAsync(df.get(), f, m->nt * sizeof(double), hipMemcpyHostToDevice, st));
        FK_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), st));
        k_gather_coef<<<grid_for(m->nt, 256), 256, 0, st>>>(df.get(), m->perm.get(), m->nt, coef.get(), bad.get());
        FK_HIP(hipGetLastError());
        FK_HIP(hipStreamSynchronize(st));
    }
    DevBuf<double> Ak(16 * m->nt);
    k_local_stiffness<<<grid_for(m->nt, 256), 256, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(),
                                                            m->perm.get(), f ? coef.get() : nullptr, m->nt,
                                                            Ak.get());
    FK_HIP(hipGetLastError());
    FK_HIP(hipMemcpyAsync(Ak_host, Ak.get(), 16 * m->nt * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
}

void tangential(fk_mesh *m, const double *H, const double *dmu, double *nzval_host) {
    hipStream_t st = m->ctx->stream;
    DevBuf<double> dH(3 * m->nt), ddmu(m->nt), val(m->nnz);
    FK_HIP(hipMemcpyAsync(dH.get(), H, 3 * m->nt * sizeof(double), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(ddmu.get(), dmu, m->nt * sizeof(double), hipMemcpyHostToDevice, st));
    k_fill_f64<<<grid_for(m->nnz, 256), 256, 0, st>>>(val.get(), m->nnz, 0.0);
    for (int c = 0; c < m->ncolours; ++c) {
        const int64_t e0 = m->colour_off[c], e1 = m->colour_off[c + 1];
        if (e1 <= e0) continue;
        k_tangent<<<grid_for(e1 - e0, 256), 256, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(),
                                                          m->perm.get(), dH.get(), ddmu.get(), m->slot.get(),
                                                          val.get(), e0, e1);
    }
    FK_HIP(hipGetLastError());
    FK_HIP(hipMemcpyAsync(nzval_host, val.get(), m->nnz * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    values_to_caller(m, nzval_host);
}

void lagrange(fk_mesh *m, double *C_host) {
    hipStream_t st = m->ctx->stream;
    DevBuf<double> C(m->nv);
    k_fill_f64<<<grid_for(m->nv, 256), 256, 0, st>>>(C.get(), m->nv, 0.0);
    for (int c = 0; c < m->ncolours; ++c) {
        const int64_t e0 = m->colour_off[c], e1 = m->colour_off[c + 1];
        if (e1 <= e0) continue;
        k_lagrange<<<grid_for(e1 - e0, 256), 256, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(),
                                                           C.get(), e0, e1);
    }
    FK_HIP(hipGetLastError());
    NodesOut<double> out(m, C_host);
    FK_HIP(hipMemcpyAsync(out.ptr, C.get(), m->nv * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    out.commit();
}

}  // namespace fk

// ---------------------------------------------------------------------------
// C ABI: coefficients and assembly
using fk::guard;

extern "C" {

int fk_set_coefficient(fk_mesh *m, const double *f) {
    if (!m || !f) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_set_coefficient", [&] {
        fk::set_device(m->ctx);
        fk::upload_coefficient(m, f);
    });
}

int fk_assemble(fk_mesh *m, int32_t method) {
    if (!m || (method != FK_ASM_COLOURED && method != FK_ASM_ATOMIC && method != FK_ASM_GATHER))
        return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_assemble", [&] {
        fk::set_device(m->ctx);
        fk::assemble(m, method);
    });
}

int fk_values(fk_mesh *m, double *values) {
    if (!m || !values) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_values", [&] {
        if (!m->have_values) fk::fail(FK_ERR_STATE, "matrix not assembled");
        fk::set_device(m->ctx);
        FK_HIP(hipMemcpyAsync(values, m->val64.get(), m->nnz * sizeof(double), hipMemcpyDeviceToHost,
                              m->ctx->stream));
        FK_HIP(hipStreamSynchronize(m->ctx->stream));
        fk::values_to_caller(m, values);
    });
}

int fk_values_f32(fk_mesh *m, float *values) {
    if (!m || !values) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_values_f32", [&] {
        if (!m->have_values) fk::fail(FK_ERR_STATE, "matrix not assembled");
        fk::set_device(m->ctx);
        fk::make_val32(m);
        FK_HIP(hipMemcpyAsync(values, m->val32.get(), m->nnz * sizeof(float), hipMemcpyDeviceToHost,
                              m->ctx->stream));
        FK_HIP(hipStreamSynchronize(m->ctx->stream));
        fk::values_to_caller(m, values);
    });
}

int fk_stiffness_matrix(fk_mesh *m, const double *f, int32_t base, int32_t *colptr, int32_t *rowval,
                        double *nzval) {
    if (!m || !f || !colptr || !rowval || !nzval || (base != 0 && base != 1)) return FK_ERR_INVALID_ARG;
    int s = fk_set_coefficient(m, f);
    if (s) return s;
    s = fk_assemble(m, FK_ASM_GATHER);
    if (s) return s;
    s = fk_pattern(m, base, colptr, rowval);
    if (s) return s;
    return fk_values(m, nzval);
}

int fk_local_stiffness(fk_mesh *m, const double *f, double *Ak) {
    if (!m || !Ak) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_local_stiffness", [&] {
        fk::set_device(m->ctx);
        fk::local_stiffness(m, f, Ak);
    });
}

int fk_tangential_stiffness(fk_mesh *m, const double *H, const double *dmu, double *nzval) {
    if (!m || !H || !dmu || !nzval) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_tangential_stiffness", [&] {
        fk::set_device(m->ctx);
        fk::tangential(m, H, dmu, nzval);
    });
}

int fk_lagrange(fk_mesh *m, double *C) {
    if (!m || !C) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_lagrange", [&] {
        fk::set_device(m->ctx);
        fk::lagrange(m, C);
    });
}

}  // extern "C"
