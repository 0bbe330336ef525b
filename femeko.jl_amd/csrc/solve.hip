// Jacobi-preconditioned CG on the Dirichlet-masked CSR matrix, FP64 and
// FP32-with-FP64-refinement, and the per-element field kernel.
//
// One PCG iteration is three kernels on a persistent grid (nblocks blocks of
// 256 threads, each owning a contiguous band of rows, bands laid out so every
// XCD streams one contiguous eighth of the vectors):
//   K1  q = A p (CSR, `lanes` lanes per row, sub-wave shuffle reduction)
//       fused with the block partials of p.q
//   K2  alpha = rho / p.q from the partials (every block reduces them in the
//       same order, so all blocks hold the same alpha bit for bit);
//       x += alpha p; r -= alpha q; z = D^-1 r (never stored);
//       partials of r.z and r.r
//   K3  beta = rho_new / rho; convergence test on r.r; p = D^-1 r + beta p
// Dot products accumulate in FP64 in a fixed order: runs are bitwise
// reproducible. Fixed (Dirichlet) rows carry D^-1 = 0, so r, z, p, x stay 0
// there and A p never picks up fixed columns: this is A[free,free] without a
// submatrix (Fluids/fluids3D.jl:142-174 eliminates by slicing).
// Chunks of iterations are captured once into a hipGraph; the host polls a
// device-side state flag between chunks, and kernels of a finished solve
// return at once.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "device.h"
#include "internal.h"

namespace fk {

// ---------------------------------------------------------------------------
// kernels

template <typename T>
__global__ void k_dinv(const T *__restrict__ val, const int32_t *__restrict__ diag, const uint8_t *__restrict__ mask,
                       int64_t nv, T *__restrict__ dinv, int32_t *__restrict__ bad) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    if (mask[i]) {
        dinv[i] = T(0);
        return;
    }
    const int32_t s = diag[i];
    const double d = s >= 0 ? static_cast<double>(val[s]) : 0.0;
    if (!(d > 0.0) || !isfinite(d)) {
        dinv[i] = T(0);
        atomicAdd(bad, 1);
        return;
    }
    dinv[i] = static_cast<T>(1.0 / d);
}

template <typename T, int LPR>
__device__ __forceinline__ T csr_row(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                     const T *__restrict__ val, const T *__restrict__ x, int64_t row, bool valid,
                                     int lane) {
    T s = T(0);
    if (valid) {
        const int32_t a = rowptr[row], b = rowptr[row + 1];
        for (int32_t k = a + lane; k < b; k += LPR) s += val[k] * x[col[k]];
    }
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s;
}

// Entry test of an iteration's first kernel: stop if the previous iteration
// stopped, and carry the flag to the other parity so the iteration after
// this one sees it too.
__device__ __forceinline__ bool stopped_entry(CgState *__restrict__ st, int parity) {
    if (!st->done_at[parity]) return false;
    if (blockIdx.x == 0 && threadIdx.x == 0) st->done_at[parity ^ 1] = 1;
    return true;
}

// K1: q = A p, partial p.q
template <typename T, int LPR>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ col, const T *__restrict__ val,
                                                    const T *__restrict__ p, T *__restrict__ q,
                                                    double *__restrict__ part, int64_t n, int64_t chunk,
                                                    int nblocks, CgState *__restrict__ st, int parity) {
    if (stopped_entry(st, parity)) return;
    __shared__ double sh[kBlock / 64 + 1];
    const int lb = logical_block(nblocks);
    const int64_t r0 = static_cast<int64_t>(lb) * chunk;
    const int64_t r1 = min(n, r0 + chunk);
    const int sub = threadIdx.x / LPR, lane = threadIdx.x % LPR;
    constexpr int RPB = kBlock / LPR;
    double acc = 0.0;
    for (int64_t rb = r0; rb < r1; rb += RPB) {
        const int64_t row = rb + sub;
        const bool valid = row < r1;
        const T s = csr_row<T, LPR>(rowptr, col, val, p, row, valid, lane);
        if (valid && lane == 0) {
            q[row] = s;
            acc += static_cast<double>(p[row]) * static_cast<double>(s);
        }
    }
    const double tot = block_sum<kBlock>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// K2: x += alpha p; r -= alpha q; partials of r.z and r.r
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update_xr(T *__restrict__ x, T *__restrict__ r,
                                                         const T *__restrict__ p, const T *__restrict__ q,
                                                         const T *__restrict__ dinv,
                                                         const double *__restrict__ part_pq,
                                                         const double *__restrict__ part_rz_cur,
                                                         double *__restrict__ part_rz_next,
                                                         double *__restrict__ part_rr, int64_t n, int64_t chunk,
                                                         int nblocks, CgState *__restrict__ st, int parity) {
    if (st->done_at[parity]) return;
    __shared__ double sh[kBlock / 64 + 1];
    const double rho = sum_partials<kBlock>(part_rz_cur, nblocks, sh);
    const double pq = sum_partials<kBlock>(part_pq, nblocks, sh);
    if (!(pq > 0.0) || !isfinite(pq)) {
        // breakdown: A not positive definite on the Krylov space (every
        // block sees the same pq and skips the update)
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->status = FK_ERR_SINGULAR;
            st->done_at[parity ^ 1] = 1;
        }
        return;
    }
    const T alpha = static_cast<T>(rho / pq);
    const int lb = logical_block(nblocks);
    const int64_t r0 = static_cast<int64_t>(lb) * chunk;  // chunk is a multiple of 64
    const int64_t r1 = min(n, r0 + chunk);
    double rz = 0.0, rr = 0.0;
    // 16-byte vector loads and stores over the band, scalar tail at n
    constexpr int VW = 16 / sizeof(T);
    using V = Vec16<T>;
    const int64_t v1 = r1 > r0 ? r1 / VW : r0 / VW;
    for (int64_t j = r0 / VW + threadIdx.x; j < v1; j += kBlock) {
        const V D = reinterpret_cast<const V *>(dinv)[j];
        const V P = reinterpret_cast<const V *>(p)[j];
        const V Q = reinterpret_cast<const V *>(q)[j];
        V X = reinterpret_cast<const V *>(x)[j];
        V R = reinterpret_cast<const V *>(r)[j];
#pragma unroll
        for (int e = 0; e < VW; ++e) {
            X.v[e] += alpha * P.v[e];
            const T ri = D.v[e] != T(0) ? R.v[e] - alpha * Q.v[e] : T(0);
            R.v[e] = ri;
            rz += static_cast<double>(ri) * static_cast<double>(D.v[e] * ri);
            rr += static_cast<double>(ri) * static_cast<double>(ri);
        }
        reinterpret_cast<V *>(x)[j] = X;
        reinterpret_cast<V *>(r)[j] = R;
    }
    for (int64_t i = max(r0, v1 * VW) + threadIdx.x; i < r1; i += kBlock) {
        const T d = dinv[i];
        x[i] += alpha * p[i];
        const T ri = d != T(0) ? r[i] - alpha * q[i] : T(0);
        r[i] = ri;
        const T z = d * ri;
        rz += static_cast<double>(ri) * static_cast<double>(z);
        rr += static_cast<double>(ri) * static_cast<double>(ri);
    }
    const double trz = block_sum<kBlock>(rz, sh);
    const double trr = block_sum<kBlock>(rr, sh);
    if (threadIdx.x == 0) {
        part_rz_next[blockIdx.x] = trz;
        part_rr[blockIdx.x] = trr;
    }
}

// K3: convergence test, p = D^-1 r + beta p
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update_p(T *__restrict__ p, const T *__restrict__ r,
                                                        const T *__restrict__ dinv,
                                                        const double *__restrict__ part_rz_next,
                                                        const double *__restrict__ part_rz_cur,
                                                        const double *__restrict__ part_rr, int64_t n,
                                                        int64_t chunk, int nblocks, CgState *__restrict__ st,
                                                        int parity) {
    if (st->done_at[parity]) return;
    __shared__ double sh[kBlock / 64 + 1];
    const double rho_new = sum_partials<kBlock>(part_rz_next, nblocks, sh);
    const double rho = sum_partials<kBlock>(part_rz_cur, nblocks, sh);
    if (blockIdx.x == 0) {
        const double rr = sum_partials<kBlock>(part_rr, nblocks, sh);
        if (threadIdx.x == 0) {
            const int64_t it = st->iters + 1;
            st->iters = it;
            st->rr = rr;
            int stop = 0;
            if (rr < st->thresh) {
                stop = 1;
            } else if (!isfinite(rr)) {
                st->status = FK_ERR_NONFINITE;
                stop = 1;
            } else if (it >= st->max_iter) {
                st->status = FK_ERR_NOT_CONVERGED;
                stop = 1;
            }
            // seen from the next iteration on; every block of this launch
            // still completes the p update below
            if (stop) st->done_at[parity ^ 1] = 1;
        }
    }
    const T beta = static_cast<T>(rho_new / rho);
    const int lb = logical_block(nblocks);
    const int64_t r0 = static_cast<int64_t>(lb) * chunk;  // chunk is a multiple of 64
    const int64_t r1 = min(n, r0 + chunk);
    constexpr int VW = 16 / sizeof(T);
    using V = Vec16<T>;
    const int64_t v1 = r1 > r0 ? r1 / VW : r0 / VW;
    for (int64_t j = r0 / VW + threadIdx.x; j < v1; j += kBlock) {
        const V D = reinterpret_cast<const V *>(dinv)[j];
        const V R = reinterpret_cast<const V *>(r)[j];
        V P = reinterpret_cast<const V *>(p)[j];
#pragma unroll
        for (int e = 0; e < VW; ++e) P.v[e] = D.v[e] * R.v[e] + beta * P.v[e];
        reinterpret_cast<V *>(p)[j] = P;
    }
    for (int64_t i = max(r0, v1 * VW) + threadIdx.x; i < r1; i += kBlock) p[i] = dinv[i] * r[i] + beta * p[i];
}

// ---- SELL-64x4: slices of 64 rows (one wavefront). Slice s with width W
// (its longest row) holds 64 W entries from sell_ptr[s]: the first
// 4 floor(W/4) entries of every row as groups of 4 consecutive entries per
// lane (group g of lane l at 4 (64 g + l)), so each lane reads 4 values and 4
// column indices with one 16-byte load, then the remaining W mod 4 entries
// one per lane (entry t of lane l at 256 floor(W/4) + 64 t + l). Padding
// entries carry value 0 and a column inside the row's own slice. Column
// indices are int32, or int16 offsets from the row when the matrix bandwidth
// allows (every column within 32767 of its row).
template <typename T>
__global__ void k_sell_gather(const T *__restrict__ src, const int32_t *__restrict__ map, int64_t n,
                              T *__restrict__ dst) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t k = map[i];
    dst[i] = k >= 0 ? src[k] : T(0);
}

template <typename T>
struct alignas(4 * sizeof(T)) Quad {
    T v[4];
};

__device__ __forceinline__ int64_t col_of(int32_t c, int64_t) { return c; }
__device__ __forceinline__ int64_t col_of(int16_t d, int64_t row) { return row + d; }

template <typename T, typename IDX>
__device__ __forceinline__ T sell_row(const int64_t *__restrict__ sptr, const int32_t *__restrict__ width,
                                      const IDX *__restrict__ col, const T *__restrict__ val,
                                      const T *__restrict__ x, int64_t s, int lane, int64_t row) {
    const int64_t base = sptr[s];
    const int W = width[s];
    const int ng = W >> 2;
    const Quad<T> *v4 = reinterpret_cast<const Quad<T> *>(val + base) + lane;
    const Quad<IDX> *c4 = reinterpret_cast<const Quad<IDX> *>(col + base) + lane;
    T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
    for (int g = 0; g < ng; ++g) {
        const Quad<T> V = v4[64 * g];
        const Quad<IDX> C = c4[64 * g];
        a0 += V.v[0] * x[col_of(C.v[0], row)];
        a1 += V.v[1] * x[col_of(C.v[1], row)];
        a2 += V.v[2] * x[col_of(C.v[2], row)];
        a3 += V.v[3] * x[col_of(C.v[3], row)];
    }
    const int64_t tb = base + 256 * static_cast<int64_t>(ng) + lane;
    const int tail = W - 4 * ng;
    if (tail > 0) a0 += val[tb] * x[col_of(col[tb], row)];
    if (tail > 1) a1 += val[tb + 64] * x[col_of(col[tb + 64], row)];
    if (tail > 2) a2 += val[tb + 128] * x[col_of(col[tb + 128], row)];
    return (a0 + a1) + (a2 + a3);
}

// K1 in SELL form: q = A p, partial p.q; each wave owns whole slices
template <typename T, typename IDX>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_sell(const int64_t *__restrict__ sptr,
                                                         const int32_t *__restrict__ width,
                                                         const IDX *__restrict__ col, const T *__restrict__ val,
                                                         const T *__restrict__ p, T *__restrict__ q,
                                                         double *__restrict__ part, int64_t n, int64_t nslices,
                                                         int64_t chunk_slices, int nblocks,
                                                         CgState *__restrict__ st, int parity) {
    if (stopped_entry(st, parity)) return;
    __shared__ double sh[kBlock / 64 + 1];
    const int lb = logical_block(nblocks);
    const int64_t s0 = static_cast<int64_t>(lb) * chunk_slices;
    const int64_t s1 = min(nslices, s0 + chunk_slices);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double acc = 0.0;
    for (int64_t s = s0 + wv; s < s1; s += kBlock / 64) {
        const int64_t row = s * 64 + lane;
        const T y = sell_row<T, IDX>(sptr, width, col, val, p, s, lane, row);
        if (row < n) {
            q[row] = y;
            acc += static_cast<double>(p[row]) * static_cast<double>(y);
        }
    }
    const double tot = block_sum<kBlock>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// FP64 r = -mask(A phi), SELL form
template <typename IDX>
__global__ __launch_bounds__(kBlock) void k_neg_residual_sell(const int64_t *__restrict__ sptr,
                                                              const int32_t *__restrict__ width,
                                                              const IDX *__restrict__ col,
                                                              const double *__restrict__ val,
                                                              const double *__restrict__ phi,
                                                              const uint8_t *__restrict__ mask,
                                                              double *__restrict__ r, double *__restrict__ part,
                                                              int64_t n, int64_t nslices, int64_t chunk_slices,
                                                              int nblocks) {
    __shared__ double sh[kBlock / 64 + 1];
    const int lb = logical_block(nblocks);
    const int64_t s0 = static_cast<int64_t>(lb) * chunk_slices;
    const int64_t s1 = min(nslices, s0 + chunk_slices);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double acc = 0.0;
    for (int64_t s = s0 + wv; s < s1; s += kBlock / 64) {
        const int64_t row = s * 64 + lane;
        const double y = sell_row<double, IDX>(sptr, width, col, val, phi, s, lane, row);
        if (row < n) {
            const double v = mask[row] ? 0.0 : -y;
            r[row] = v;
            acc += v * v;
        }
    }
    const double tot = block_sum<kBlock>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// x = 0, p = z = D^-1 r, partials of r.z and r.r (initial residual r given)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_init(T *__restrict__ x, T *__restrict__ r, T *__restrict__ p,
                                                    const T *__restrict__ dinv, double *__restrict__ part_rz,
                                                    double *__restrict__ part_rr, int64_t n, int64_t chunk,
                                                    int nblocks) {
    __shared__ double sh[kBlock / 64 + 1];
    const int lb = logical_block(nblocks);
    const int64_t r0 = static_cast<int64_t>(lb) * chunk;
    const int64_t r1 = min(n, r0 + chunk);
    double rz = 0.0, rr = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
        const T d = dinv[i];
        const T ri = d != T(0) ? r[i] : T(0);
        r[i] = ri;
        x[i] = T(0);
        const T z = d * ri;
        p[i] = z;
        rz += static_cast<double>(ri) * static_cast<double>(z);
        rr += static_cast<double>(ri) * static_cast<double>(ri);
    }
    const double trz = block_sum<kBlock>(rz, sh);
    const double trr = block_sum<kBlock>(rr, sh);
    if (threadIdx.x == 0) {
        part_rz[blockIdx.x] = trz;
        part_rr[blockIdx.x] = trr;
    }
}

// one block: threshold from the reference norm, stop at once if b == 0
__global__ __launch_bounds__(kBlock) void k_cg_start(const double *__restrict__ part_rr, int nblocks, double bb_ref,
                                                     double rtol, int64_t max_iter, CgState *__restrict__ st) {
    __shared__ double sh[kBlock / 64 + 1];
    const double rr = sum_partials<kBlock>(part_rr, nblocks, sh);
    if (threadIdx.x == 0) {
        const double bb = bb_ref >= 0.0 ? bb_ref : rr;
        st->bb = bb;
        st->thresh = rtol * rtol * bb;
        st->iters = 0;
        st->rr = rr;
        st->status = FK_OK;
        st->max_iter = max_iter;
        int stop = 0;
        if (!isfinite(rr)) {
            st->status = FK_ERR_NONFINITE;
            stop = 1;
        } else if (rr == 0.0 || rr < st->thresh) {
            stop = 1;
        } else if (max_iter <= 0) {
            st->status = FK_ERR_NOT_CONVERGED;
            stop = 1;
        }
        st->done_at[0] = stop;
        st->done_at[1] = stop;
    }
}

__host__ __device__ inline bool cg_stopped(const CgState &s) { return (s.done_at[0] | s.done_at[1]) != 0; }

// r = -(A phi) on free rows, 0 on fixed rows; partials of r.r. FP64.
template <int LPR>
__global__ __launch_bounds__(kBlock) void k_neg_residual(const int32_t *__restrict__ rowptr,
                                                         const int32_t *__restrict__ col,
                                                         const double *__restrict__ val,
                                                         const double *__restrict__ phi,
                                                         const uint8_t *__restrict__ mask, double *__restrict__ r,
                                                         double *__restrict__ part, int64_t n, int64_t chunk,
                                                         int nblocks) {
    __shared__ double sh[kBlock / 64 + 1];
    const int lb = logical_block(nblocks);
    const int64_t r0 = static_cast<int64_t>(lb) * chunk;
    const int64_t r1 = min(n, r0 + chunk);
    const int sub = threadIdx.x / LPR, lane = threadIdx.x % LPR;
    constexpr int RPB = kBlock / LPR;
    double acc = 0.0;
    for (int64_t rb = r0; rb < r1; rb += RPB) {
        const int64_t row = rb + sub;
        const bool valid = row < r1;
        const double s = csr_row<double, LPR>(rowptr, col, val, phi, row, valid, lane);
        if (valid && lane == 0) {
            const double v = mask[row] ? 0.0 : -s;
            r[row] = v;
            acc += v * v;
        }
    }
    const double tot = block_sum<kBlock>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void k_init_phi(double *__restrict__ phi, const double *__restrict__ g, const uint8_t *__restrict__ mask,
                           int64_t n, int keep_free) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mask[i])
        phi[i] = g[i];
    else if (!keep_free)
        phi[i] = 0.0;
}

__global__ void k_masked_copy(double *__restrict__ dst, const double *__restrict__ src,
                              const uint8_t *__restrict__ mask, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = mask[i] ? src[i] : 0.0;
}

__global__ void k_scale_to_f32(const double *__restrict__ a, float *__restrict__ b, double s, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = static_cast<float>(a[i] * s);
}

template <typename T>
__global__ void k_phi_axpy(double *__restrict__ phi, const T *__restrict__ x, double s, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) phi[i] += s * static_cast<double>(x[i]);
}

// H = -sum phi_i grad_i per element; |H|; energy partials 1/2 mu |H|^2 VE
__global__ __launch_bounds__(kBlock) void k_field(const int4 *__restrict__ tet, const double *__restrict__ X,
                                                  const double *__restrict__ Y, const double *__restrict__ Z,
                                                  const int32_t *__restrict__ perm, const double *__restrict__ phi,
                                                  const double *__restrict__ coef, int64_t nt,
                                                  double *__restrict__ H, double *__restrict__ Hn,
                                                  double *__restrict__ part) {
    __shared__ double sh[kBlock / 64 + 1];
    double en = 0.0;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < nt;
         e += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int4 nd = tet[e];
        const Grad G = element_gradients(X, Y, Z, nd);
        const double u0 = phi[nd.x], u1 = phi[nd.y], u2 = phi[nd.z], u3 = phi[nd.w];
        double h[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) h[c] = -(u0 * G.g[0][c] + u1 * G.g[1][c] + u2 * G.g[2][c] + u3 * G.g[3][c]);
        const double hh = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
        const int64_t o = perm[e];
        if (H) {
            H[3 * o] = h[0];
            H[3 * o + 1] = h[1];
            H[3 * o + 2] = h[2];
        }
        if (Hn) Hn[o] = sqrt(hh);
        if (coef) en += 0.5 * coef[e] * hh * fabs(G.det) * (1.0 / 6.0);
    }
    const double tot = block_sum<kBlock>(en, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// ---------------------------------------------------------------------------
// host side

template <typename T>
static PcgWork<T> &work(fk_mesh *m);
template <>
PcgWork<double> &work<double>(fk_mesh *m) { return m->w64; }
template <>
PcgWork<float> &work<float>(fk_mesh *m) { return m->w32; }

template <typename T>
static const T *values(fk_mesh *m);
template <>
const double *values<double>(fk_mesh *m) { return m->val64.get(); }
template <>
const float *values<float>(fk_mesh *m) { return m->val32.get(); }

template <typename T>
static void ensure_work(fk_mesh *m) {
    PcgWork<T> &w = work<T>(m);
    const size_t nv = static_cast<size_t>(m->nv), nb = static_cast<size_t>(m->nblocks);
    w.x.ensure(nv); w.r.ensure(nv); w.p.ensure(nv); w.q.ensure(nv); w.dinv.ensure(nv);
    w.part_pq.ensure(nb); w.part_rz.ensure(2 * nb); w.part_rr.ensure(nb); w.part_bb.ensure(nb);
    w.state.ensure(1);
}

template <typename T>
static void compute_dinv(fk_mesh *m) {
    hipStream_t st = m->ctx->stream;
    PcgWork<T> &w = work<T>(m);
    DevBuf<int32_t> bad(1);
    FK_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), st));
    k_dinv<T><<<grid_for(m->nv, 256), 256, 0, st>>>(values<T>(m), m->diag.get(), m->mask.get(), m->nv, w.dinv.get(),
                                                    bad.get());
    FK_HIP(hipGetLastError());
    int32_t h = 0;
    FK_HIP(hipMemcpyAsync(&h, bad.get(), sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    if (h) fail(FK_ERR_SINGULAR, std::to_string(h) + " free node(s) without a positive diagonal");
}

// fill the SELL-64x4 columns (int32 or int16 row offsets) and the SELL -> CSR
// position map
__global__ void k_sell_fill(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
                            const int64_t *__restrict__ sptr, const int32_t *__restrict__ width, int64_t nv,
                            int64_t nrows_padded, int32_t *__restrict__ scol, int16_t *__restrict__ sdcol,
                            int32_t *__restrict__ smap) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= nrows_padded) return;
    const int64_t s = r >> 6;
    const int lane = static_cast<int>(r & 63);
    const int W = width[s];
    const int ng = W >> 2;
    const int64_t base = sptr[s];
    const int32_t a = r < nv ? rowptr[r] : 0;
    const int32_t len = r < nv ? rowptr[r + 1] - a : 0;
    const int32_t self = static_cast<int32_t>(r < nv ? r : nv - 1);
    for (int k = 0; k < W; ++k) {
        const int64_t o = k < 4 * ng ? base + 4 * (64 * static_cast<int64_t>(k >> 2) + lane) + (k & 3)
                                     : base + 256 * static_cast<int64_t>(ng) + 64 * (k - 4 * ng) + lane;
        const int32_t c = k < len ? colidx[a + k] : self;
        if (scol) scol[o] = c;
        if (sdcol) sdcol[o] = static_cast<int16_t>(c - r);
        smap[o] = k < len ? a + k : -1;
    }
}

// largest |col - row| over the CSR pattern
__global__ void k_bandwidth(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colidx, int64_t nv,
                            int32_t *__restrict__ out) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= nv) return;
    int32_t b = 0;
    for (int32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) b = max(b, abs(colidx[k] - static_cast<int32_t>(r)));
    atomicMax(out, b);
}

static void build_sell(fk_mesh *m) {
    if (m->have_sell) return;
    hipStream_t st = m->ctx->stream;
    const int64_t nv = m->nv;
    std::vector<int32_t> rp(static_cast<size_t>(nv + 1));
    FK_HIP(hipMemcpyAsync(rp.data(), m->rowptr.get(), (nv + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    const int64_t ns = (nv + 63) / 64;
    std::vector<int32_t> width(static_cast<size_t>(ns));
    std::vector<int64_t> sptr(static_cast<size_t>(ns + 1), 0);
    for (int64_t s = 0; s < ns; ++s) {
        int32_t w = 0;
        for (int64_t r = s * 64; r < std::min(nv, s * 64 + 64); ++r) w = std::max(w, rp[r + 1] - rp[r]);
        width[s] = w;
        sptr[s + 1] = sptr[s] + 64 * static_cast<int64_t>(w);
    }
    m->nslices = ns;
    m->sell_len = sptr[ns];
    m->sell_chunk = (ns + m->nblocks - 1) / m->nblocks;
    // int16 row offsets when every column lies within 32767 of its row
    {
        DevBuf<int32_t> bw(1);
        FK_HIP(hipMemsetAsync(bw.get(), 0, sizeof(int32_t), st));
        k_bandwidth<<<grid_for(nv, 256), 256, 0, st>>>(m->rowptr.get(), m->colidx.get(), nv, bw.get());
        FK_HIP(hipGetLastError());
        int32_t h = 0;
        FK_HIP(hipMemcpyAsync(&h, bw.get(), sizeof(int32_t), hipMemcpyDeviceToHost, st));
        FK_HIP(hipStreamSynchronize(st));
        m->bandwidth = h;
        m->sell_short = m->allow_short && h + 64 <= 32767;
    }
    m->sell_ptr.alloc(ns + 1);
    m->sell_width.alloc(ns);
    m->sell_col.release();
    m->sell_dcol.release();
    if (m->sell_short)
        m->sell_dcol.alloc(m->sell_len);
    else
        m->sell_col.alloc(m->sell_len);
    m->sell_map.alloc(m->sell_len);
    FK_HIP(hipMemcpyAsync(m->sell_ptr.get(), sptr.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(m->sell_width.get(), width.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice, st));
    k_sell_fill<<<grid_for(ns * 64, 256), 256, 0, st>>>(m->rowptr.get(), m->colidx.get(), m->sell_ptr.get(),
                                                        m->sell_width.get(), nv, ns * 64, m->sell_col.get(),
                                                        m->sell_dcol.get(), m->sell_map.get());
    FK_HIP(hipGetLastError());
    FK_HIP(hipStreamSynchronize(st));
    m->have_sell = true;
}

// bring the solver's copy of the values up to date with val64 / val32
template <typename T>
static void sync_solver_values(fk_mesh *m);
template <>
void sync_solver_values<double>(fk_mesh *m) {
    if (m->format != 1 || m->sell64_valid) return;
    build_sell(m);
    m->sell_val64.ensure(m->sell_len);
    k_sell_gather<double><<<grid_for(m->sell_len, 256), 256, 0, m->ctx->stream>>>(
        m->val64.get(), m->sell_map.get(), m->sell_len, m->sell_val64.get());
    FK_HIP(hipGetLastError());
    m->sell64_valid = true;
}
template <>
void sync_solver_values<float>(fk_mesh *m) {
    if (m->format != 1 || m->sell32_valid) return;
    build_sell(m);
    make_val32(m);
    m->sell_val32.ensure(m->sell_len);
    k_sell_gather<float><<<grid_for(m->sell_len, 256), 256, 0, m->ctx->stream>>>(
        m->val32.get(), m->sell_map.get(), m->sell_len, m->sell_val32.get());
    FK_HIP(hipGetLastError());
    m->sell32_valid = true;
}

template <typename T>
static const T *sell_values(fk_mesh *m);
template <>
const double *sell_values<double>(fk_mesh *m) { return m->sell_val64.get(); }
template <>
const float *sell_values<float>(fk_mesh *m) { return m->sell_val32.get(); }

template <typename T, int LPR>
static void launch_spmv(fk_mesh *m, PcgWork<T> &w, int parity) {
    k_cg_spmv<T, LPR><<<m->nblocks, kBlock, 0, m->ctx->stream>>>(m->rowptr.get(), m->colidx.get(), values<T>(m),
                                                                 w.p.get(), w.q.get(), w.part_pq.get(), m->nv,
                                                                 m->chunk, m->nblocks, w.state.get(), parity);
}

template <typename T>
static void launch_spmv_lanes(fk_mesh *m, PcgWork<T> &w, int parity) {
    if (m->format == 1) {
        if (m->sell_short)
            k_cg_spmv_sell<T, int16_t><<<m->nblocks, kBlock, 0, m->ctx->stream>>>(
                m->sell_ptr.get(), m->sell_width.get(), m->sell_dcol.get(), sell_values<T>(m), w.p.get(), w.q.get(),
                w.part_pq.get(), m->nv, m->nslices, m->sell_chunk, m->nblocks, w.state.get(), parity);
        else
            k_cg_spmv_sell<T, int32_t><<<m->nblocks, kBlock, 0, m->ctx->stream>>>(
                m->sell_ptr.get(), m->sell_width.get(), m->sell_col.get(), sell_values<T>(m), w.p.get(), w.q.get(),
                w.part_pq.get(), m->nv, m->nslices, m->sell_chunk, m->nblocks, w.state.get(), parity);
        return;
    }
    switch (m->lanes) {
        case 4: launch_spmv<T, 4>(m, w, parity); break;
        case 8: launch_spmv<T, 8>(m, w, parity); break;
        case 32: launch_spmv<T, 32>(m, w, parity); break;
        case 64: launch_spmv<T, 64>(m, w, parity); break;
        default: launch_spmv<T, 16>(m, w, parity); break;
    }
}

template <typename T>
static void enqueue_iteration(fk_mesh *m, PcgWork<T> &w, int parity, bool timed) {
    hipStream_t st = m->ctx->stream;
    const int nb = m->nblocks;
    double *rz_cur = w.part_rz.get() + (parity ? nb : 0);
    double *rz_next = w.part_rz.get() + (parity ? 0 : nb);
    TimingEvents &te = m->ctx->timing;
    if (timed) {
        if (te.used + 2 > te.ev.size()) {
            for (int i = 0; i < 256; ++i) {
                hipEvent_t e;
                FK_HIP(hipEventCreate(&e));
                te.ev.push_back(e);
            }
        }
        FK_HIP(hipEventRecord(te.ev[te.used], st));
    }
    launch_spmv_lanes<T>(m, w, parity);
    if (timed) {
        FK_HIP(hipEventRecord(te.ev[te.used + 1], st));
        te.used += 2;
    }
    k_cg_update_xr<T><<<nb, kBlock, 0, st>>>(w.x.get(), w.r.get(), w.p.get(), w.q.get(), w.dinv.get(),
                                             w.part_pq.get(), rz_cur, rz_next, w.part_rr.get(), m->nv, m->chunk, nb,
                                             w.state.get(), parity);
    k_cg_update_p<T><<<nb, kBlock, 0, st>>>(w.p.get(), w.r.get(), w.dinv.get(), rz_next, rz_cur, w.part_rr.get(),
                                            m->nv, m->chunk, nb, w.state.get(), parity);
}

// collect the first `valid` event pairs recorded since the last harvest
// (launches after convergence return at once and are not counted)
static void harvest_timing(fk_context *ctx, int64_t valid) {
    TimingEvents &te = ctx->timing;
    for (size_t i = 0; i + 1 < te.used && static_cast<int64_t>(i / 2) < valid; i += 2) {
        float ms = 0.f;
        FK_HIP(hipEventSynchronize(te.ev[i + 1]));
        FK_HIP(hipEventElapsedTime(&ms, te.ev[i], te.ev[i + 1]));
        te.total_ms += ms;
        te.launches += 1;
    }
    te.used = 0;
}

// PCG on work<T>: r holds the initial residual on entry; x returns the
// correction. bb_ref < 0: tolerance relative to ||r0||, else to sqrt(bb_ref).
template <typename T>
static CgState run_pcg(fk_mesh *m, double rtol, int64_t max_iter, double bb_ref) {
    hipStream_t st = m->ctx->stream;
    PcgWork<T> &w = work<T>(m);
    const int nb = m->nblocks;
    sync_solver_values<T>(m);
    k_cg_init<T><<<nb, kBlock, 0, st>>>(w.x.get(), w.r.get(), w.p.get(), w.dinv.get(), w.part_rz.get(),
                                        w.part_bb.get(), m->nv, m->chunk, nb);
    k_cg_start<<<1, kBlock, 0, st>>>(w.part_bb.get(), nb, bb_ref, rtol, max_iter, w.state.get());
    FK_HIP(hipGetLastError());
    const bool timed = m->ctx->timing.enabled;
    CgState *hs = m->hstate.ptr;
    if (timed) {
        // direct launches with events around every SpMV; poll every 16 iterations
        int64_t it = 0, before = 0;
        while (true) {
            for (int k = 0; k < 16; ++k, ++it) enqueue_iteration<T>(m, w, static_cast<int>(it & 1), true);
            FK_HIP(hipMemcpyAsync(hs, w.state.get(), sizeof(CgState), hipMemcpyDeviceToHost, st));
            FK_HIP(hipStreamSynchronize(st));
            harvest_timing(m->ctx, hs->iters - before);
            before = hs->iters;
            if (cg_stopped(*hs)) break;
        }
        return *hs;
    }
    if (!m->ctx->graphs) {
        // direct launches, poll every 32 iterations
        int64_t it = 0;
        while (true) {
            for (int k = 0; k < 32; ++k, ++it) enqueue_iteration<T>(m, w, static_cast<int>(it & 1), false);
            FK_HIP(hipMemcpyAsync(hs, w.state.get(), sizeof(CgState), hipMemcpyDeviceToHost, st));
            FK_HIP(hipStreamSynchronize(st));
            if (cg_stopped(*hs)) break;
        }
        return *hs;
    }
    // graph of G iterations (G even keeps the parity sequence)
    const int G = m->nv > 200000 ? 32 : 16;
    const int cfg = m->lanes + 1000 * m->format;
    if (!w.graph || w.graph_iters != G || w.graph_lanes != cfg) {
        w.reset_graph();
        hipGraph_t graph;
        FK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < G; ++k) enqueue_iteration<T>(m, w, k & 1, false);
        FK_HIP(hipStreamEndCapture(st, &graph));
        hipError_t e = hipGraphInstantiate(&w.graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        FK_HIP(e);
        w.graph_iters = G;
        w.graph_lanes = cfg;
    }
    // launch chunk k+1 before waiting for the state copy after chunk k
    FK_HIP(hipGraphLaunch(w.graph, st));
    FK_HIP(hipMemcpyAsync(&hs[0], w.state.get(), sizeof(CgState), hipMemcpyDeviceToHost, st));
    hipEvent_t ev[2];
    FK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    FK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    FK_HIP(hipEventRecord(ev[0], st));
    int cur = 0;
    CgState out{};
    try {
        while (true) {
            FK_HIP(hipGraphLaunch(w.graph, st));
            FK_HIP(hipMemcpyAsync(&hs[1 - cur], w.state.get(), sizeof(CgState), hipMemcpyDeviceToHost, st));
            FK_HIP(hipEventRecord(ev[1 - cur], st));
            FK_HIP(hipEventSynchronize(ev[cur]));
            if (cg_stopped(hs[cur])) break;
            cur = 1 - cur;
        }
        FK_HIP(hipStreamSynchronize(st));
        out = hs[1 - cur];  // latest copy
    } catch (...) {
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        throw;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return out;
}

// Run exactly G iterations from parity 0 (kernels after a stop return at
// once), then return the state (host sync). G must be even.
template <typename T>
static CgState run_chunk(fk_mesh *m, PcgWork<T> &w, int G) {
    hipStream_t st = m->ctx->stream;
    CgState *hs = m->hstate.ptr;
    const bool timed = m->ctx->timing.enabled;
    if (timed || !m->ctx->graphs) {
        FK_HIP(hipMemcpyAsync(hs, w.state.get(), sizeof(CgState), hipMemcpyDeviceToHost, st));
        FK_HIP(hipStreamSynchronize(st));
        const int64_t before = hs->iters;
        for (int k = 0; k < G; ++k) enqueue_iteration<T>(m, w, k & 1, timed);
        FK_HIP(hipMemcpyAsync(hs, w.state.get(), sizeof(CgState), hipMemcpyDeviceToHost, st));
        FK_HIP(hipStreamSynchronize(st));
        if (timed) harvest_timing(m->ctx, hs->iters - before);
        return *hs;
    }
    const int cfg = m->lanes + 1000 * m->format;
    if (!w.graph || w.graph_iters != G || w.graph_lanes != cfg) {
        w.reset_graph();
        hipGraph_t graph;
        FK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < G; ++k) enqueue_iteration<T>(m, w, k & 1, false);
        FK_HIP(hipStreamEndCapture(st, &graph));
        hipError_t e = hipGraphInstantiate(&w.graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        FK_HIP(e);
        w.graph_iters = G;
        w.graph_lanes = cfg;
    }
    FK_HIP(hipGraphLaunch(w.graph, st));
    FK_HIP(hipMemcpyAsync(hs, w.state.get(), sizeof(CgState), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    return *hs;
}

static double neg_residual(fk_mesh *m, const double *phi, double *r);

// phi += x / s; x = 0
template <typename T>
__global__ void k_rr_fold(double *__restrict__ phi, T *__restrict__ x, double inv_s, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    phi[i] += inv_s * static_cast<double>(x[i]);
    x[i] = T(0);
}

// residual replacement: r = s r64 (free rows), partials of r.D^-1.r into the
// parity-0 slot the next iteration reads as rho
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rr_replace(T *__restrict__ r, const double *__restrict__ r64, double s,
                                                       const T *__restrict__ dinv, double *__restrict__ part_rz,
                                                       int64_t n, int64_t chunk, int nblocks) {
    __shared__ double sh[kBlock / 64 + 1];
    const int lb = logical_block(nblocks);
    const int64_t r0 = static_cast<int64_t>(lb) * chunk;
    const int64_t r1 = min(n, r0 + chunk);
    double rz = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
        const T d = dinv[i];
        const T ri = d != T(0) ? static_cast<T>(s * r64[i]) : T(0);
        r[i] = ri;
        rz += static_cast<double>(ri) * static_cast<double>(d * ri);
    }
    const double t = block_sum<kBlock>(rz, sh);
    if (threadIdx.x == 0) part_rz[blockIdx.x] = t;
}

__global__ void k_rr_resume(CgState *__restrict__ st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->done_at[0] = 0;
        st->done_at[1] = 0;
        st->status = FK_OK;
    }
}

// FP32 Jacobi-PCG whose residual is replaced by the FP64 true residual
// -mask(A64 phi), phi accumulating the FP32 increments in FP64 ("reliable
// updates"). A replacement happens at the end of an iteration chunk once the
// FP32 recursive residual has fallen by the factor `delta` since the last
// replacement, or once it claims convergence; the search direction is kept,
// so the Krylov space is not restarted. Convergence is decided on the FP64
// true residual: ||r|| < rtol ||b||. Returns (iterations, replacements,
// final ||r||^2).
static int solve_f32_replace(fk_mesh *m, double rtol, int64_t max_iter, double bb, double delta, int64_t *iters,
                             int32_t *nrep, double *rr_out) {
    hipStream_t st = m->ctx->stream;
    const int64_t nv = m->nv;
    PcgWork<float> &w = m->w32;
    const int nb = m->nblocks;
    const double thresh = rtol * rtol * bb;
    double rr = neg_residual(m, m->phi.get(), m->r64.get());
    *iters = 0;
    *nrep = 0;
    *rr_out = rr;
    if (rr < thresh) return FK_OK;
    const double s = 1.0 / std::sqrt(rr);
    sync_solver_values<float>(m);
    k_scale_to_f32<<<grid_for(nv, 256), 256, 0, st>>>(m->r64.get(), w.r.get(), s, nv);
    k_cg_init<float><<<nb, kBlock, 0, st>>>(w.x.get(), w.r.get(), w.p.get(), w.dinv.get(), w.part_rz.get(),
                                            w.part_bb.get(), nv, m->chunk, nb);
    // the FP32 recursion stops (mid-chunk) once it reaches the target; the
    // FP64 true residual then decides, and the solve resumes if it is short
    k_cg_start<<<1, kBlock, 0, st>>>(w.part_bb.get(), nb, -1.0, rtol * std::sqrt(bb) * s, max_iter,
                                     w.state.get());
    FK_HIP(hipGetLastError());
    const int G = nv > 200000 ? 32 : 16;
    const double target2 = rtol * rtol * bb * s * s;  // the FP64 target in scaled units
    double rr_last = 1.0;                              // scaled ||r||^2 at the last replacement
    int status = FK_OK;
    const char *tr = std::getenv("FEMEKO_HIP_TRACE");
    const bool trace = tr && tr[0] == '1';
    constexpr int kMaxRestarts = 32;
    int restarts = 0;
    // r32 = s r64, x32 = 0, p = D^-1 r32 (rho into the parity-0 slot)
    auto restart_direction = [&] {
        k_scale_to_f32<<<grid_for(nv, 256), 256, 0, st>>>(m->r64.get(), w.r.get(), s, nv);
        k_cg_init<float><<<nb, kBlock, 0, st>>>(w.x.get(), w.r.get(), w.p.get(), w.dinv.get(), w.part_rz.get(),
                                                w.part_bb.get(), nv, m->chunk, nb);
        FK_HIP(hipGetLastError());
    };
    while (true) {
        const CgState cs = run_chunk<float>(m, w, G);
        *iters = cs.iters;
        const bool need = cg_stopped(cs) || cs.rr < delta * delta * rr_last || cs.rr < target2 ||
                          cs.iters >= max_iter;
        if (trace)
            std::fprintf(stderr, "[fk trace] its %lld status %d stopped %d rr32 %.3e rr_last %.3e replace %d\n",
                         static_cast<long long>(cs.iters), cs.status, cg_stopped(cs) ? 1 : 0, cs.rr, rr_last,
                         need ? 1 : 0);
        if (!need) continue;
        const bool broken = cs.status == FK_ERR_SINGULAR || cs.status == FK_ERR_NONFINITE || !std::isfinite(cs.rr);
        if (broken) {
            // the FP32 recursion broke down since the last update: drop its
            // increment (phi and r64 still hold the last FP64 state) and
            // restart the direction from that residual
            if (++restarts > kMaxRestarts) {
                status = cs.status != FK_OK ? cs.status : FK_ERR_NONFINITE;
                break;
            }
            if (cs.iters >= max_iter) {
                status = FK_ERR_NOT_CONVERGED;
                break;
            }
            restart_direction();
            if (trace) std::fprintf(stderr, "[fk trace] FP32 breakdown: increment dropped, direction restarted\n");
            k_rr_resume<<<1, 64, 0, st>>>(w.state.get());
            FK_HIP(hipGetLastError());
            continue;
        }
        k_rr_fold<float><<<grid_for(nv, 256), 256, 0, st>>>(m->phi.get(), w.x.get(), 1.0 / s, nv);
        FK_HIP(hipGetLastError());
        rr = neg_residual(m, m->phi.get(), m->r64.get());
        *nrep += 1;
        *rr_out = rr;
        rr_last = rr * s * s;
        if (trace) std::fprintf(stderr, "[fk trace] replacement: true rr (scaled) %.3e\n", rr_last);
        if (rr < thresh) break;
        if (!std::isfinite(rr)) {
            status = FK_ERR_NONFINITE;
            break;
        }
        if (cs.iters >= max_iter) {
            status = FK_ERR_NOT_CONVERGED;
            break;
        }
        if (rr_last > 4.0 * cs.rr) {
            // the recursion drifted below the true residual: the search
            // direction no longer matches it, restart from p = D^-1 r
            if (++restarts > kMaxRestarts) {
                status = FK_ERR_NOT_CONVERGED;
                break;
            }
            restart_direction();
            if (trace) std::fprintf(stderr, "[fk trace] direction restarted\n");
        } else {
            k_rr_replace<float><<<nb, kBlock, 0, st>>>(w.r.get(), m->r64.get(), s, w.dinv.get(), w.part_rz.get(),
                                                       nv, m->chunk, nb);
        }
        k_rr_resume<<<1, 64, 0, st>>>(w.state.get());
        FK_HIP(hipGetLastError());
    }
    FK_HIP(hipStreamSynchronize(st));
    return status;
}

template <int LPR>
static void launch_neg_residual(fk_mesh *m, const double *phi, double *r, double *part) {
    k_neg_residual<LPR><<<m->nblocks, kBlock, 0, m->ctx->stream>>>(m->rowptr.get(), m->colidx.get(),
                                                                   m->val64.get(), phi, m->mask.get(), r, part,
                                                                   m->nv, m->chunk, m->nblocks);
}

// r = -mask(A64 phi); returns ||r||^2 (host sync)
static double neg_residual(fk_mesh *m, const double *phi, double *r) {
    hipStream_t st = m->ctx->stream;
    m->part_misc.ensure(m->nblocks);
    if (m->format == 1) {
        sync_solver_values<double>(m);
        if (m->sell_short)
            k_neg_residual_sell<int16_t><<<m->nblocks, kBlock, 0, st>>>(
                m->sell_ptr.get(), m->sell_width.get(), m->sell_dcol.get(), m->sell_val64.get(), phi, m->mask.get(),
                r, m->part_misc.get(), m->nv, m->nslices, m->sell_chunk, m->nblocks);
        else
            k_neg_residual_sell<int32_t><<<m->nblocks, kBlock, 0, st>>>(
                m->sell_ptr.get(), m->sell_width.get(), m->sell_col.get(), m->sell_val64.get(), phi, m->mask.get(),
                r, m->part_misc.get(), m->nv, m->nslices, m->sell_chunk, m->nblocks);
    } else switch (m->lanes) {
        case 4: launch_neg_residual<4>(m, phi, r, m->part_misc.get()); break;
        case 8: launch_neg_residual<8>(m, phi, r, m->part_misc.get()); break;
        case 32: launch_neg_residual<32>(m, phi, r, m->part_misc.get()); break;
        case 64: launch_neg_residual<64>(m, phi, r, m->part_misc.get()); break;
        default: launch_neg_residual<16>(m, phi, r, m->part_misc.get()); break;
    }
    FK_HIP(hipGetLastError());
    FK_HIP(hipMemcpyAsync(m->hpart.ptr, m->part_misc.get(), m->nblocks * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (int i = 0; i < m->nblocks; ++i) s += m->hpart.ptr[i];
    return s;
}

// ---- entry points for the other translation units (nonlinear loop, adjoint)
void ensure_work_f64(fk_mesh *m) { ensure_work<double>(m); }
void compute_dinv_f64(fk_mesh *m) { compute_dinv<double>(m); }
void compute_dinv_f32(fk_mesh *m) { compute_dinv<float>(m); }
CgState run_pcg_f64(fk_mesh *m, double rtol, int64_t max_iter, double bb_ref) {
    return run_pcg<double>(m, rtol, max_iter, bb_ref);
}
void phi_axpy_f64(fk_mesh *m, const double *x, double s) {
    k_phi_axpy<double><<<grid_for(m->nv, 256), 256, 0, m->ctx->stream>>>(m->phi.get(), x, s, m->nv);
    FK_HIP(hipGetLastError());
}
void k_init_phi_dev(fk_mesh *m, int keep_free) {
    k_init_phi<<<grid_for(m->nv, 256), 256, 0, m->ctx->stream>>>(m->phi.get(), m->g.get(), m->mask.get(), m->nv,
                                                                 keep_free);
    FK_HIP(hipGetLastError());
}

void set_dirichlet(fk_mesh *m, const uint8_t *mask, const double *g) {
    hipStream_t st = m->ctx->stream;
    int64_t nfixed = 0;
    for (int64_t i = 0; i < m->nv; ++i) {
        if (mask[i]) {
            ++nfixed;
            if (!std::isfinite(g[i])) fail(FK_ERR_NONFINITE, "non-finite Dirichlet value");
        }
    }
    m->mask.ensure(m->nv);
    m->g.ensure(m->nv);
    FK_HIP(hipMemcpyAsync(m->mask.get(), mask, m->nv, hipMemcpyHostToDevice, st));
    DevBuf<double> tmp(m->nv);
    FK_HIP(hipMemcpyAsync(tmp.get(), g, m->nv * sizeof(double), hipMemcpyHostToDevice, st));
    k_masked_copy<<<grid_for(m->nv, 256), 256, 0, st>>>(m->g.get(), tmp.get(), m->mask.get(), m->nv);
    FK_HIP(hipGetLastError());
    FK_HIP(hipStreamSynchronize(st));
    m->nfixed = nfixed;
    m->have_dirichlet = true;
    m->dinv64_valid = false;
    m->dinv32_valid = false;
}

static void check_ready(fk_mesh *m) {
    if (!m->have_values) fail(FK_ERR_STATE, "matrix not assembled");
    if (!m->have_dirichlet) fail(FK_ERR_STATE, "no Dirichlet data set");
    if (m->nfixed == 0) fail(FK_ERR_SINGULAR, "no fixed node: the pure Neumann system is singular");
}

void solve(fk_mesh *m, int precision, double rtol, int64_t max_iter, double inner_rtol, int max_outer, int warm,
           int64_t *iters, int32_t *outer, double *relres) {
    check_ready(m);
    hipStream_t st = m->ctx->stream;
    const int64_t nv = m->nv;
    m->phi.ensure(nv);
    m->r64.ensure(nv);
    ensure_work<double>(m);
    const bool mixed = precision == FK_PREC_F32_REFINED || precision == FK_PREC_F32_RESTART;
    if (mixed) {
        ensure_work<float>(m);
        make_val32(m);
    }
    if (!m->dinv64_valid) {
        compute_dinv<double>(m);
        m->dinv64_valid = true;
    }
    if (mixed && !m->dinv32_valid) {
        compute_dinv<float>(m);
        m->dinv32_valid = true;
    }
    k_init_phi<<<grid_for(nv, 256), 256, 0, st>>>(m->phi.get(), m->g.get(), m->mask.get(), nv,
                                                  (warm && m->have_phi) ? 1 : 0);
    FK_HIP(hipGetLastError());
    m->have_phi = true;
    // ||b||^2 with b = -A[free,fixed] g
    const double bb = neg_residual(m, m->g.get(), m->r64.get());
    int64_t total = 0;
    int32_t nout = 0;
    double rel = 0.0;
    int status = FK_OK;
    if (bb == 0.0) {
        k_init_phi<<<grid_for(nv, 256), 256, 0, st>>>(m->phi.get(), m->g.get(), m->mask.get(), nv, 0);
        FK_HIP(hipStreamSynchronize(st));
    } else if (precision == FK_PREC_F64) {
        PcgWork<double> &w = m->w64;
        (void)neg_residual(m, m->phi.get(), w.r.get());
        const CgState s = run_pcg<double>(m, rtol, max_iter, bb);
        k_phi_axpy<double><<<grid_for(nv, 256), 256, 0, st>>>(m->phi.get(), w.x.get(), 1.0, nv);
        FK_HIP(hipGetLastError());
        FK_HIP(hipStreamSynchronize(st));
        total = s.iters;
        nout = 1;
        rel = std::sqrt(s.rr / bb);
        status = s.status;
    } else if (precision == FK_PREC_F32_REFINED) {
        double rr = 0.0;
        // inner_rtol is the reliable-update trigger: replace once the FP32
        // residual fell by this factor (values outside (0,1) mean 0.5)
        const double delta = (inner_rtol > 0.0 && inner_rtol < 1.0) ? inner_rtol : 0.5;
        status = solve_f32_replace(m, rtol, max_iter, bb, delta, &total, &nout, &rr);
        rel = std::sqrt(rr / bb);
    } else {
        // restarted refinement: each FP32 PCG runs from scratch to inner_rtol
        PcgWork<float> &w = m->w32;
        double rr = neg_residual(m, m->phi.get(), m->r64.get());
        const double thresh = rtol * rtol * bb;
        while (!(rr < thresh) && nout < max_outer && total < max_iter) {
            const double s = std::sqrt(rr);
            k_scale_to_f32<<<grid_for(nv, 256), 256, 0, st>>>(m->r64.get(), w.r.get(), 1.0 / s, nv);
            FK_HIP(hipGetLastError());
            const CgState cs = run_pcg<float>(m, inner_rtol, max_iter - total, -1.0);
            total += cs.iters;
            ++nout;
            if (cs.status == FK_ERR_SINGULAR || cs.status == FK_ERR_NONFINITE) {
                status = cs.status;
                break;
            }
            k_phi_axpy<float><<<grid_for(nv, 256), 256, 0, st>>>(m->phi.get(), w.x.get(), s, nv);
            FK_HIP(hipGetLastError());
            rr = neg_residual(m, m->phi.get(), m->r64.get());
            if (!std::isfinite(rr)) {
                status = FK_ERR_NONFINITE;
                break;
            }
        }
        rel = std::sqrt(rr / bb);
        if (status == FK_OK && !(rr < thresh)) status = FK_ERR_NOT_CONVERGED;
    }
    if (iters) *iters = total;
    if (outer) *outer = nout;
    if (relres) *relres = rel;
    if (status != FK_OK) fail(status, std::string("solve stopped: ") + fk_status_string(status));
}

void solve_rhs(fk_mesh *m, const double *b, double rtol, int64_t max_iter, double *x, int64_t *iters,
               double *relres) {
    check_ready(m);
    hipStream_t st = m->ctx->stream;
    ensure_work<double>(m);
    if (!m->dinv64_valid) {
        compute_dinv<double>(m);
        m->dinv64_valid = true;
    }
    PcgWork<double> &w = m->w64;
    FK_HIP(hipMemcpyAsync(w.r.get(), b, m->nv * sizeof(double), hipMemcpyHostToDevice, st));
    const CgState s = run_pcg<double>(m, rtol, max_iter, -1.0);
    FK_HIP(hipMemcpyAsync(x, w.x.get(), m->nv * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    if (iters) *iters = s.iters;
    if (relres) *relres = s.bb > 0 ? std::sqrt(s.rr / s.bb) : 0.0;
    if (s.status != FK_OK) fail(s.status, std::string("solve stopped: ") + fk_status_string(s.status));
}

void field(fk_mesh *m, double *H, double *Hn, double *energy) {
    if (!m->have_phi) fail(FK_ERR_STATE, "no potential: solve or set one first");
    if (energy && !m->have_coef) fail(FK_ERR_STATE, "energy needs a coefficient");
    hipStream_t st = m->ctx->stream;
    DevBuf<double> dH, dHn;
    if (H) dH.alloc(3 * m->nt);
    if (Hn) dHn.alloc(m->nt);
    const int nb = std::max(1, m->nblocks);
    m->part_misc.ensure(nb);
    k_field<<<nb, kBlock, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(), m->perm.get(), m->phi.get(),
                                   energy ? m->coef.get() : nullptr, m->nt, H ? dH.get() : nullptr,
                                   Hn ? dHn.get() : nullptr, m->part_misc.get());
    FK_HIP(hipGetLastError());
    if (H) FK_HIP(hipMemcpyAsync(H, dH.get(), 3 * m->nt * sizeof(double), hipMemcpyDeviceToHost, st));
    if (Hn) FK_HIP(hipMemcpyAsync(Hn, dHn.get(), m->nt * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipMemcpyAsync(m->hpart.ptr, m->part_misc.get(), nb * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    if (energy) {
        double s = 0.0;
        for (int i = 0; i < nb; ++i) s += m->hpart.ptr[i];
        *energy = s;
    }
}

}  // namespace fk

// ---------------------------------------------------------------------------
// C ABI: Dirichlet data, solve, field, instrumentation
using fk::guard;

extern "C" {

int fk_set_dirichlet(fk_mesh *m, const uint8_t *mask, const double *g) {
    if (!m || !mask || !g) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_set_dirichlet", [&] {
        fk::set_device(m->ctx);
        fk::set_dirichlet(m, mask, g);
    });
}

int fk_solve(fk_mesh *m, int32_t precision, double rtol, int64_t max_iter, double inner_rtol, int32_t max_outer,
             int32_t warm_start, int64_t *iterations, int32_t *outer, double *relres) {
    if (!m ||
        (precision != FK_PREC_F64 && precision != FK_PREC_F32_REFINED && precision != FK_PREC_F32_RESTART) ||
        !(rtol >= 0.0) || max_iter < 0)
        return FK_ERR_INVALID_ARG;
    if (precision == FK_PREC_F32_RESTART && (!(inner_rtol > 0.0) || !(inner_rtol < 1.0) || max_outer < 1))
        return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_solve", [&] {
        fk::set_device(m->ctx);
        fk::solve(m, precision, rtol, max_iter, inner_rtol, max_outer, warm_start, iterations, outer, relres);
    });
}

int fk_solve_rhs(fk_mesh *m, const double *b, double rtol, int64_t max_iter, double *x, int64_t *iterations,
                 double *relres) {
    if (!m || !b || !x || !(rtol >= 0.0) || max_iter < 0) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_solve_rhs", [&] {
        fk::set_device(m->ctx);
        fk::solve_rhs(m, b, rtol, max_iter, x, iterations, relres);
    });
}

int fk_potential(fk_mesh *m, double *phi) {
    if (!m || !phi) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_potential", [&] {
        if (!m->have_phi) fk::fail(FK_ERR_STATE, "no potential");
        fk::set_device(m->ctx);
        FK_HIP(hipMemcpyAsync(phi, m->phi.get(), m->nv * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
        FK_HIP(hipStreamSynchronize(m->ctx->stream));
    });
}

int fk_set_potential(fk_mesh *m, const double *phi) {
    if (!m || !phi) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_set_potential", [&] {
        fk::set_device(m->ctx);
        m->phi.ensure(m->nv);
        FK_HIP(hipMemcpyAsync(m->phi.get(), phi, m->nv * sizeof(double), hipMemcpyHostToDevice, m->ctx->stream));
        FK_HIP(hipStreamSynchronize(m->ctx->stream));
        m->have_phi = true;
    });
}

int fk_field(fk_mesh *m, double *H, double *Hn, double *energy) {
    if (!m) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_field", [&] {
        fk::set_device(m->ctx);
        fk::field(m, H, Hn, energy);
    });
}

int fk_set_timing(fk_context *ctx, int32_t enable) {
    if (!ctx) return FK_ERR_INVALID_ARG;
    ctx->timing.enabled = enable != 0;
    ctx->timing.launches = 0;
    ctx->timing.total_ms = 0.0;
    ctx->timing.used = 0;
    return FK_OK;
}

int fk_spmv_timing(fk_context *ctx, int64_t *launches, double *total_ms) {
    if (!ctx) return FK_ERR_INVALID_ARG;
    if (launches) *launches = ctx->timing.launches;
    if (total_ms) *total_ms = ctx->timing.total_ms;
    return FK_OK;
}

int fk_set_spmv_format(fk_mesh *m, int32_t format) {
    if (!m || (format != FK_SPMV_CSR && format != FK_SPMV_SELL64)) return FK_ERR_INVALID_ARG;
    m->format = format;
    return FK_OK;
}

int fk_set_lanes_per_row(fk_mesh *m, int32_t lanes) {
    if (!m || (lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)) return FK_ERR_INVALID_ARG;
    m->lanes = lanes;
    return FK_OK;
}

}  // extern "C"
