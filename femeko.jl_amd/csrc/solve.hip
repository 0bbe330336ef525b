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
#include <thread>
#include <type_traits>

#include "device.h"
#include "internal.h"

namespace fk {

// ---------------------------------------------------------------------------
// kernels

// D^-1 of the FP64 diagonal, stored in T (0 on fixed rows)
template <typename T>
__global__ void k_dinv(const double *__restrict__ val, const int32_t *__restrict__ diag,
                       const uint8_t *__restrict__ mask, int64_t nv, T *__restrict__ dinv, int32_t *__restrict__ bad) {
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

// Vectors per thread that K2 and K3 load before they reduce the previous
// kernel's partials: the loads are in flight while the block waits on the
// reduction (a band of 4096 rows is 4 vectors per thread in FP32).
constexpr int kPrefetch = 4;
// Partials per thread that K2 and K3 load first (EarlyPartials): K2's SpMV
// partials as pairs, PQ per thread, instantiated for 2, 4, 8 and 16 and
// picked per launch from the SpMV's block count (16 cover the 8152 blocks of
// the FP32 stencil SpMV at n=160, 4 its 2016 at n=100, 2 the 1024 of the
// FP64 grid), and the vector kernels' own 1024 partials. Issued before the vector loads, the sums no longer wait for
// those: K2 alone on n=161^3 took 13.2 us against 15.4 us with the partials
// loaded after the vectors (12.1 us with no reduction at all).
constexpr int kEarlyBlk = 4;

// Clamped index of a prefetched vector: every prefetch load unconditional
// (a load under a branch makes the compiler wait for all outstanding loads
// where the paths join); the value is used only when j < v1. The index stays
// below nvec = n / VW, the whole vectors of the array, also for a block
// whose band is empty (r0 >= n, where v1 = r0 / VW lies past the array);
// n >= VW (a mesh has at least 4 nodes).
__device__ __forceinline__ int64_t clamp_vec(int64_t j, int64_t v1, int64_t nvec) {
    const int64_t lim = v1 < nvec ? v1 : nvec;
    const int64_t c = j < lim ? j : lim - 1;
    return c > 0 ? c : 0;
}

// K2: r -= alpha q; partials of r.z and r.r. alpha (every block computes the
// same value from the same partials) is left in st->alpha for K3, which
// applies x += alpha p while it reads p anyway: 12 bytes per row less than
// updating x here.
template <typename T, int PQ>
__global__ __launch_bounds__(kBlock) void k_cg_update_xr(T *__restrict__ r, const T *__restrict__ q,
                                                         const T *__restrict__ dinv,
                                                         const double *__restrict__ part_pq, int npq,
                                                         const double *__restrict__ part_rz_cur,
                                                         double *__restrict__ part_rz_next,
                                                         double *__restrict__ part_rr, int64_t n, int64_t chunk,
                                                         int nblocks, CgState *__restrict__ st, int parity) {
    if (st->done_at[parity]) return;
    __shared__ double sh[2 * (kBlock / 64 + 1)];
    EarlyPartials<kBlock, PQ, 2> epq;
    EarlyPartials<kBlock, kEarlyBlk, 1> erz;
    epq.load(part_pq, npq);
    erz.load(part_rz_cur, nblocks);
    __builtin_amdgcn_sched_barrier(0);
    const int lb = logical_block(nblocks);
    const int64_t r0 = static_cast<int64_t>(lb) * chunk;  // chunk is a multiple of 64
    const int64_t r1 = min(n, r0 + chunk);
    // 16-byte vector loads and stores over the band, scalar tail at n
    constexpr int VW = 16 / sizeof(T);
    using V = Vec16<T>;
    const int64_t v1 = r1 > r0 ? r1 / VW : r0 / VW;
    const int64_t j0 = r0 / VW + threadIdx.x;
    V D[kPrefetch], Q[kPrefetch], R[kPrefetch];
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
        const int64_t j = clamp_vec(j0 + u * kBlock, v1, n / VW);
        D[u] = reinterpret_cast<const V *>(dinv)[j];
        Q[u] = reinterpret_cast<const V *>(q)[j];
        R[u] = reinterpret_cast<const V *>(r)[j];
    }
    __builtin_amdgcn_sched_barrier(0);  // no use of a partial before the vector loads are issued
    double red[2] = {erz.sum(part_rz_cur, nblocks), epq.sum(part_pq, npq)};
    block_sum_n<kBlock, 2>(red, sh);
    const double rho = red[0], pq = red[1];
    if (!(pq > 0.0) || !isfinite(pq)) {
        // breakdown: A not positive definite on the Krylov space (every
        // block sees the same pq and skips the update; K3 then adds nothing
        // to x)
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->status = FK_ERR_SINGULAR;
            st->done_at[parity ^ 1] = 1;
            st->alpha = 0.0;
        }
        return;
    }
    const T alpha = static_cast<T>(rho / pq);
    if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha = static_cast<double>(alpha);
    double rz = 0.0, rr = 0.0;
    auto step = [&](const V &Dv, const V &Qv, V &Rv) {
#pragma unroll
        for (int e = 0; e < VW; ++e) {
            const T ri = Dv.v[e] != T(0) ? Rv.v[e] - alpha * Qv.v[e] : T(0);
            Rv.v[e] = ri;
            rz += static_cast<double>(ri) * static_cast<double>(Dv.v[e] * ri);
            rr += static_cast<double>(ri) * static_cast<double>(ri);
        }
    };
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
        const int64_t j = j0 + u * kBlock;
        if (j < v1) {
            step(D[u], Q[u], R[u]);
            reinterpret_cast<V *>(r)[j] = R[u];
        }
    }
    for (int64_t j = j0 + kPrefetch * kBlock; j < v1; j += kBlock) {
        const V Dv = reinterpret_cast<const V *>(dinv)[j];
        const V Qv = reinterpret_cast<const V *>(q)[j];
        V Rv = reinterpret_cast<const V *>(r)[j];
        step(Dv, Qv, Rv);
        reinterpret_cast<V *>(r)[j] = Rv;
    }
    for (int64_t i = max(r0, v1 * VW) + threadIdx.x; i < r1; i += kBlock) {
        const T d = dinv[i];
        const T ri = d != T(0) ? r[i] - alpha * q[i] : T(0);
        r[i] = ri;
        const T z = d * ri;
        rz += static_cast<double>(ri) * static_cast<double>(z);
        rr += static_cast<double>(ri) * static_cast<double>(ri);
    }
    double tot[2] = {rz, rr};
    block_sum_n<kBlock, 2>(tot, sh);
    if (threadIdx.x == 0) {
        part_rz_next[blockIdx.x] = tot[0];
        part_rr[blockIdx.x] = tot[1];
    }
}

// 16-byte accesses of x in K3. x is read and written once per iteration,
// by this kernel only, so in FP64 it carries the non-temporal hint: the
// FP64 iteration's working set (stencil copy 234 MB + five vectors 167 MB)
// exceeds the 256 MB MALL, and x then takes no room from the data the next
// kernels read again. Measured at n=160: FP64 step 269.9 against 276.6 ms
// (stencil copy), 375.9 against 384.1 ms (SELL copy, RCM numbering). The
// FP32 iteration's working set fits the MALL, and the hint made its step
// slower (123.3 against 120.6 ms; 229.4 against 225.7 ms), so FP32 keeps
// plain accesses.
template <typename T>
constexpr bool x_nt() {
    return sizeof(T) == 8;
}
template <typename T>
__device__ __forceinline__ Vec16<T> ld_x16(const Vec16<T> *p) {
    if constexpr (x_nt<T>()) {
        using E = T __attribute__((ext_vector_type(16 / sizeof(T))));
        const E v = __builtin_nontemporal_load(reinterpret_cast<const E *>(p));
        Vec16<T> o;
#pragma unroll
        for (int e = 0; e < 16 / static_cast<int>(sizeof(T)); ++e) o.v[e] = v[e];
        return o;
    }
    return *p;
}
template <typename T>
__device__ __forceinline__ void st_x16(Vec16<T> *p, const Vec16<T> &v) {
    if constexpr (x_nt<T>()) {
        using E = T __attribute__((ext_vector_type(16 / sizeof(T))));
        E e;
#pragma unroll
        for (int k = 0; k < 16 / static_cast<int>(sizeof(T)); ++k) e[k] = v.v[k];
        __builtin_nontemporal_store(e, reinterpret_cast<E *>(p));
        return;
    }
    *p = v;
}

// K3: convergence test, x += alpha p (alpha of K2; 0 after a breakdown),
// p = D^-1 r + beta p
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update_p(T *__restrict__ x, T *__restrict__ p,
                                                        const T *__restrict__ r, const T *__restrict__ dinv,
                                                        const double *__restrict__ part_rz_next,
                                                        const double *__restrict__ part_rz_cur,
                                                        const double *__restrict__ part_rr, int64_t n,
                                                        int64_t chunk, int nblocks, CgState *__restrict__ st,
                                                        int parity) {
    if (st->done_at[parity]) return;
    __shared__ double sh[3 * (kBlock / 64 + 1)];
    EarlyPartials<kBlock, kEarlyBlk, 1> ezn, ezc, err;
    ezn.load(part_rz_next, nblocks);
    ezc.load(part_rz_cur, nblocks);
    err.load(part_rr, nblocks);
    __builtin_amdgcn_sched_barrier(0);
    const int lb = logical_block(nblocks);
    const int64_t r0 = static_cast<int64_t>(lb) * chunk;  // chunk is a multiple of 64
    const int64_t r1 = min(n, r0 + chunk);
    constexpr int VW = 16 / sizeof(T);
    using V = Vec16<T>;
    const int64_t v1 = r1 > r0 ? r1 / VW : r0 / VW;
    const int64_t j0 = r0 / VW + threadIdx.x;
    V D[kPrefetch], R[kPrefetch], P[kPrefetch], X[kPrefetch];
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
        const int64_t j = clamp_vec(j0 + u * kBlock, v1, n / VW);
        D[u] = reinterpret_cast<const V *>(dinv)[j];
        R[u] = reinterpret_cast<const V *>(r)[j];
        P[u] = reinterpret_cast<const V *>(p)[j];
        X[u] = ld_x16(reinterpret_cast<const V *>(x) + j);
    }
    __builtin_amdgcn_sched_barrier(0);  // no use of a partial before the vector loads are issued
    const T alpha = static_cast<T>(st->alpha);
    // the three reductions in one pass (r.r used only in block 0, which
    // keeps the convergence state)
    double red[3] = {ezn.sum(part_rz_next, nblocks), ezc.sum(part_rz_cur, nblocks),
                     blockIdx.x == 0 ? err.sum(part_rr, nblocks) : 0.0};
    block_sum_n<kBlock, 3>(red, sh);
    const double rho_new = red[0], rho = red[1];
    if (blockIdx.x == 0) {
        const double rr = red[2];
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
    const bool addx = alpha != T(0);
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
        const int64_t j = j0 + u * kBlock;
        if (j < v1) {
            if (addx) {
#pragma unroll
                for (int e = 0; e < VW; ++e) X[u].v[e] += alpha * P[u].v[e];
                st_x16(reinterpret_cast<V *>(x) + j, X[u]);
            }
#pragma unroll
            for (int e = 0; e < VW; ++e) P[u].v[e] = D[u].v[e] * R[u].v[e] + beta * P[u].v[e];
            reinterpret_cast<V *>(p)[j] = P[u];
        }
    }
    for (int64_t j = j0 + kPrefetch * kBlock; j < v1; j += kBlock) {
        const V Dv = reinterpret_cast<const V *>(dinv)[j];
        const V Rv = reinterpret_cast<const V *>(r)[j];
        V Pv = reinterpret_cast<const V *>(p)[j];
        if (addx) {
            V Xv = ld_x16(reinterpret_cast<const V *>(x) + j);
#pragma unroll
            for (int e = 0; e < VW; ++e) Xv.v[e] += alpha * Pv.v[e];
            st_x16(reinterpret_cast<V *>(x) + j, Xv);
        }
#pragma unroll
        for (int e = 0; e < VW; ++e) Pv.v[e] = Dv.v[e] * Rv.v[e] + beta * Pv.v[e];
        reinterpret_cast<V *>(p)[j] = Pv;
    }
    for (int64_t i = max(r0, v1 * VW) + threadIdx.x; i < r1; i += kBlock) {
        const T pi = p[i];
        if (addx) x[i] += alpha * pi;
        p[i] = dinv[i] * r[i] + beta * pi;
    }
}

// ---- SELL-64x4 solver copy of the matrix: off-diagonal entries only (the
// diagonal is a separate array), slices of 64 rows (one wavefront). Slice s
// holds 64 W values from ptr[s] (W = |width[s]|): the first 4 floor(W/4)
// entries of every row as groups of 4 consecutive entries per lane (group g
// of lane l at 4 (64 g + l)), so each lane reads 4 values with one 16-byte
// load, then the remaining W mod 4 entries one per lane (entry t of lane l at
// 256 floor(W/4) + 64 t + l). Column indices come in one of two forms:
//   per row (width[s] = W >= 0): 64 W indices from cptr[s], laid out like
//     the values, as int32 columns or int16 offsets from the row (when every
//     column lies within 32767 of its row); padding entries point at the row
//     itself (or at n - 1 past the last row) and carry value 0;
//   shared (width[s] = -W < 0): the off-diagonal columns of every row of the
//     slice are among row + o_k for W offsets o_k stored once from
//     soff[cptr[s]]; a row without column row + o_k holds value 0 there and
//     its index is clamped into [0, n). Regions of a mesh with translation-
//     invariant connectivity (structured blocks) take this form: the
//     wavefront reads its offsets with scalar loads and streams values only.
template <typename T, typename IDX>
struct Sell {
    const int64_t *ptr;    // value offset per slice
    const int32_t *width;  // W (per-row indices) or -W (shared offsets)
    const int64_t *cptr;   // index offset per slice (col or soff)
    const IDX *col;        // per-row indices
    const int32_t *soff;   // shared offsets
    const T *val;          // off-diagonal values
    const T *diag;         // diagonal (FP64 standard form only)
    int64_t n;
};

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

// Loads of the SELL copy's values and indices in the FP32 product carry the
// non-temporal hint: read once per product, they then do not push the PCG
// vectors (5 x 16.7 MB at n=160) out of the 256 MB MALL, which the copy's
// 358 MB would otherwise sweep each iteration. Measured on the n=160 mesh in
// RCM numbering: SpMV 68.4 us against 78.1 us, FP32 step 225.8 against
// 250.7 ms. The same hint on the FP64 SELL copy made the FP64 step slower
// (406.8 against 383.8 ms), on the FP64 stencil copy too (299.5 against
// 276.0 ms), and on the batch's stencil SpMM (75.5 against 73.5 ms): those
// keep plain loads.
template <bool NT, typename T>
__device__ __forceinline__ T ld_s(const T *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ Quad<T> ld_s(const Quad<T> *p) {
    if constexpr (NT) {
        using V = T __attribute__((ext_vector_type(4)));
        const V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
        Quad<T> q;
        q.v[0] = v.x;
        q.v[1] = v.y;
        q.v[2] = v.z;
        q.v[3] = v.w;
        return q;
    }
    return *p;
}
template <typename T>
constexpr bool sell_nt() {
    return sizeof(T) == 4;
}

__device__ __forceinline__ int64_t col_of(int32_t c, int64_t) { return c; }
__device__ __forceinline__ int64_t col_of(int16_t d, int64_t row) { return row + d; }
__device__ __forceinline__ int64_t clamp_col(int64_t c, int64_t n) { return min(max(c, int64_t(0)), n - 1); }

// One gathered term: a x_j (standard form) or a (x_j - x_i) (difference form).
// P1 stiffness rows sum to zero (the basis gradients of an element sum to
// zero), so A x = sum_{j != i} a_ij (x_j - x_i) exactly; in FP32 the
// difference form keeps smooth x from cancelling (the terms are as small as
// the result) and makes the rounded operator an exact weighted graph
// Laplacian, spectrally within FP32 rounding of A.
template <bool DIFF, typename T>
__device__ __forceinline__ T term(T a, T xj, T xi) {
    return DIFF ? a * (xj - xi) : a * xj;
}

// off-diagonal part of row `row` (lane `lane` of slice s)
template <bool DIFF, typename T, typename IDX>
__device__ __forceinline__ T sell_row(const Sell<T, IDX> &A, const T *__restrict__ x, int64_t s, int lane,
                                      int64_t row, T xi) {
    const int64_t base = A.ptr[s];
    const int Wr = A.width[s];
    const int64_t cb = A.cptr[s];
    const int W = Wr < 0 ? -Wr : Wr;
    const int ng = W >> 2;
    const int tail = W - 4 * ng;
    const Quad<T> *v4 = reinterpret_cast<const Quad<T> *>(A.val + base) + lane;
    const int64_t tb = base + 256 * static_cast<int64_t>(ng) + lane;
    T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
    if (Wr < 0) {
        const int32_t *o = A.soff + cb;  // uniform over the wave
        for (int g = 0; g < ng; ++g) {
            const Quad<T> V = ld_s<sell_nt<T>()>(v4 + 64 * g);
            a0 += term<DIFF>(V.v[0], x[clamp_col(row + o[4 * g], A.n)], xi);
            a1 += term<DIFF>(V.v[1], x[clamp_col(row + o[4 * g + 1], A.n)], xi);
            a2 += term<DIFF>(V.v[2], x[clamp_col(row + o[4 * g + 2], A.n)], xi);
            a3 += term<DIFF>(V.v[3], x[clamp_col(row + o[4 * g + 3], A.n)], xi);
        }
        const int32_t *ot = o + 4 * ng;
        if (tail > 0) a0 += term<DIFF>(A.val[tb], x[clamp_col(row + ot[0], A.n)], xi);
        if (tail > 1) a1 += term<DIFF>(A.val[tb + 64], x[clamp_col(row + ot[1], A.n)], xi);
        if (tail > 2) a2 += term<DIFF>(A.val[tb + 128], x[clamp_col(row + ot[2], A.n)], xi);
    } else {
        const Quad<IDX> *c4 = reinterpret_cast<const Quad<IDX> *>(A.col + cb) + lane;
        for (int g = 0; g < ng; ++g) {
            const Quad<T> V = ld_s<sell_nt<T>()>(v4 + 64 * g);
            const Quad<IDX> C = ld_s<sell_nt<T>()>(c4 + 64 * g);
            a0 += term<DIFF>(V.v[0], x[col_of(C.v[0], row)], xi);
            a1 += term<DIFF>(V.v[1], x[col_of(C.v[1], row)], xi);
            a2 += term<DIFF>(V.v[2], x[col_of(C.v[2], row)], xi);
            a3 += term<DIFF>(V.v[3], x[col_of(C.v[3], row)], xi);
        }
        const int64_t ct = cb + 256 * static_cast<int64_t>(ng) + lane;
        if (tail > 0) a0 += term<DIFF>(A.val[tb], x[col_of(A.col[ct], row)], xi);
        if (tail > 1) a1 += term<DIFF>(A.val[tb + 64], x[col_of(A.col[ct + 64], row)], xi);
        if (tail > 2) a2 += term<DIFF>(A.val[tb + 128], x[col_of(A.col[ct + 128], row)], xi);
    }
    return (a0 + a1) + (a2 + a3);
}

// off-diagonal part of row `row` (lane `lane` of slice s), all loads of the
// row issued before any is used: the per-row indices of the NG groups and
// the tail, then the values and the gathers of x, then the sums. The loads
// are unconditional: a group or tail entry past the slice's width reads a
// valid stand-in (the first values of the array, column 0 of the row's
// range) and its term is dropped, so no branch splits the batch and a slice
// costs two dependent memory round trips (indices, then values and x)
// instead of two per group of 4 entries. The sums run in the order of the
// per-group loop (sell_row).
template <int NG, bool DIFF, typename T, typename IDX>
__device__ __forceinline__ T sell_row_batched(const Sell<T, IDX> &A, const T *__restrict__ x, int64_t s, int lane,
                                              int64_t row, T xi) {
    const int64_t base = A.ptr[s];
    const int Wr = A.width[s];
    const int64_t cb = A.cptr[s];
    const int W = Wr < 0 ? -Wr : Wr;
    const int ng = W >> 2;
    const int tail = W - 4 * ng;
    const int64_t n = A.n;
    const Quad<T> *v4 = reinterpret_cast<const Quad<T> *>(A.val + base) + lane;
    const Quad<T> *v4_any = reinterpret_cast<const Quad<T> *>(A.val);
    const int64_t tb = base + 256 * static_cast<int64_t>(ng) + lane;
    Quad<T> V[NG];
    T xv[NG][4];
    T vt[3], xt[3];
    if (Wr < 0) {
        const int32_t *o = A.soff + cb;  // uniform over the wave
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const bool ok = g < ng;
            V[g] = ld_s<sell_nt<T>()>(ok ? v4 + 64 * g : v4_any);
#pragma unroll
            for (int j = 0; j < 4; ++j) xv[g][j] = x[clamp_col(row + (ok ? o[4 * g + j] : 0), n)];
        }
        const int32_t *ot = o + 4 * ng;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const bool ok = t < tail;
            vt[t] = A.val[ok ? tb + 64 * t : 0];
            xt[t] = x[clamp_col(row + (ok ? ot[t] : 0), n)];
        }
    } else {
        const Quad<IDX> *c4 = reinterpret_cast<const Quad<IDX> *>(A.col + cb) + lane;
        const Quad<IDX> *c4_any = reinterpret_cast<const Quad<IDX> *>(A.col);
        const int64_t ct = cb + 256 * static_cast<int64_t>(ng) + lane;
        Quad<IDX> C[NG];
        IDX ci[3];
#pragma unroll
        for (int g = 0; g < NG; ++g) C[g] = ld_s<sell_nt<T>()>(g < ng ? c4 + 64 * g : c4_any);
#pragma unroll
        for (int t = 0; t < 3; ++t) ci[t] = A.col[t < tail ? ct + 64 * t : 0];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            V[g] = ld_s<sell_nt<T>()>(g < ng ? v4 + 64 * g : v4_any);
#pragma unroll
            for (int j = 0; j < 4; ++j) xv[g][j] = x[clamp_col(col_of(C[g].v[j], row), n)];
        }
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            vt[t] = A.val[t < tail ? tb + 64 * t : 0];
            xt[t] = x[clamp_col(col_of(ci[t], row), n)];
        }
    }
    T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
#pragma unroll
    for (int g = 0; g < NG; ++g)
        if (g < ng) {
            a0 += term<DIFF>(V[g].v[0], xv[g][0], xi);
            a1 += term<DIFF>(V[g].v[1], xv[g][1], xi);
            a2 += term<DIFF>(V[g].v[2], xv[g][2], xi);
            a3 += term<DIFF>(V[g].v[3], xv[g][3], xi);
        }
    if (tail > 0) a0 += term<DIFF>(vt[0], xt[0], xi);
    if (tail > 1) a1 += term<DIFF>(vt[1], xt[1], xi);
    if (tail > 2) a2 += term<DIFF>(vt[2], xt[2], xi);
    return (a0 + a1) + (a2 + a3);
}

// (A x)_row: the PCG products in the difference form (both precisions: no
// diagonal is read); the FP64 true residual (RESID) in the standard form
// with the assembled diagonal added last. NG > 0: rows of at most 4 NG + 3
// entries with all loads issued first (sell_row_batched); NG = 0: any width,
// group by group (sell_row).
template <int NG, bool RESID = false, typename T, typename IDX>
__device__ __forceinline__ T sell_apply(const Sell<T, IDX> &A, const T *__restrict__ x, int64_t s, int lane,
                                        int64_t row, T xi) {
    constexpr bool DIFF = !RESID;
    const T off = NG > 0 ? sell_row_batched<(NG > 0 ? NG : 1), DIFF>(A, x, s, lane, row, xi)
                         : sell_row<DIFF>(A, x, s, lane, row, xi);
    if (DIFF) return off;
    return row < A.n ? off + A.diag[row] * xi : off;
}

// K1 in SELL form: q = A p, partial p.q; each wave owns whole slices, U
// slices in flight per wave (independent loads interleaved)
template <typename T, typename IDX, int BLOCK, int U, int NG>
__global__ __launch_bounds__(BLOCK) void k_cg_spmv_sell(const Sell<T, IDX> A, const T *__restrict__ p,
                                                        T *__restrict__ q, double *__restrict__ part,
                                                        int64_t nslices, int64_t chunk_slices, int nblocks,
                                                        CgState *__restrict__ st, int parity) {
    if (stopped_entry(st, parity)) return;
    __shared__ double sh[BLOCK / 64 + 1];
    constexpr int NW = BLOCK / 64;
    const int lb = logical_block(nblocks);
    const int64_t s0 = static_cast<int64_t>(lb) * chunk_slices;
    const int64_t s1 = min(nslices, s0 + chunk_slices);
    // wave index made uniform for the compiler: slice tables and shared
    // offsets are then read with scalar loads
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t n = A.n;
    double acc = 0.0;
    int64_t s = s0 + wv;
    for (; s + (U - 1) * NW < s1; s += U * NW) {
        T y[U], xi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t row = (s + u * NW) * 64 + lane;
            xi[u] = row < n ? p[row] : T(0);
            y[u] = sell_apply<NG>(A, p, s + u * NW, lane, row, xi[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t row = (s + u * NW) * 64 + lane;
            if (row < n) {
                q[row] = y[u];
                acc += static_cast<double>(xi[u]) * static_cast<double>(y[u]);
            }
        }
    }
    for (; s < s1; s += NW) {
        const int64_t row = s * 64 + lane;
        const T xi = row < n ? p[row] : T(0);
        const T y = sell_apply<NG>(A, p, s, lane, row, xi);
        if (row < n) {
            q[row] = y;
            acc += static_cast<double>(xi) * static_cast<double>(y);
        }
    }
    const double tot = block_sum<BLOCK>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// FP64 r = -mask(A phi), SELL form
// (gate: when given, the kernel runs only in a replacement step, act == 1)
template <typename IDX, int NG>
__global__ __launch_bounds__(kBlock) void k_neg_residual_sell(const Sell<double, IDX> A,
                                                              const double *__restrict__ phi,
                                                              const uint8_t *__restrict__ mask,
                                                              const double *__restrict__ b,
                                                              double *__restrict__ r, double *__restrict__ part,
                                                              int64_t nslices, int64_t chunk_slices, int nblocks,
                                                              const RuState *__restrict__ gate) {
    if (gate && gate->act != 1) return;
    __shared__ double sh[kBlock / 64 + 1];
    const int lb = logical_block(nblocks);
    const int64_t s0 = static_cast<int64_t>(lb) * chunk_slices;
    const int64_t s1 = min(nslices, s0 + chunk_slices);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double acc = 0.0;
    for (int64_t s = s0 + wv; s < s1; s += kBlock / 64) {
        const int64_t row = s * 64 + lane;
        const double xi = row < A.n ? phi[row] : 0.0;
        const double y = sell_apply<NG, true>(A, phi, s, lane, row, xi);
        if (row < A.n) {
            const double v = mask[row] ? 0.0 : (b ? b[row] - y : -y);
            r[row] = v;
            acc += v * v;
        }
    }
    const double tot = block_sum<kBlock>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// ---- stencil copy: symmetric diagonal storage -------------------------------
// Column k (k < K, at most 16) of val holds a(i, i + off[k]) for the positive
// offsets off[k] (0 where row i has no such column, or past the last row);
// a(i, i - off[k]) = a(i - off[k], i) is the same column read at row
// i - off[k]. One thread per row; the shifted reads trail the streamed ones
// by off[k] rows of the same band and are served by L2.
template <typename T>
struct Stencil {
    const T *val;   // K x ld
    const T *diag;  // FP64 standard form only
    int64_t ld;
    int64_t n;
    int32_t K;
    int32_t off[16];
};

// K offsets known at compile time. FP32: all 4K loads are issued
// unconditionally (indices clamped into range; a clamped forward entry holds
// value 0, a clamped backward term is dropped), then the terms are summed.
// FP64: loads guarded by the row bounds over the runtime count, few
// registers, so 512-thread blocks run 8 waves per SIMD and the occupancy
// hides the latency. Measured at n=160 (PCG product, difference form, 72 B
// per row): 81 us, against 86 us with the FP32 form's unconditional loads
// (occupancy), 114 us with a compile-time offset loop, 82 us on the FP32
// kernel's short-lived 256-thread grid, and 88 us for the standard form,
// which also streams the diagonal.
template <bool DIFF, int K, typename T>
__device__ __forceinline__ T stencil_row(const Stencil<T> &A, const T *__restrict__ x, int64_t i, T xi) {
    if (sizeof(T) == 8) {
        T a0 = T(0), a1 = T(0);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (k >= A.K) break;
            const int64_t o = A.off[k];
            const T *__restrict__ v = A.val + k * A.ld;
            if (i + o < A.n) a0 += term<DIFF>(v[i], x[i + o], xi);
            if (i >= o) a1 += term<DIFF>(v[i - o], x[i - o], xi);
        }
        return a0 + a1;
    }
    T vf[K], vb[K], xf[K], xb[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int64_t o = A.off[k];
        const T *__restrict__ v = A.val + k * A.ld;
        const int64_t f = min(i + o, A.n - 1);
        const int64_t b = max(i - o, int64_t(0));
        vf[k] = v[i];
        vb[k] = v[b];
        xf[k] = x[f];
        xb[k] = x[b];
    }
    T a0 = T(0), a1 = T(0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        a0 += term<DIFF>(vf[k], xf[k], xi);
        if (i >= A.off[k]) a1 += term<DIFF>(vb[k], xb[k], xi);
    }
    return a0 + a1;
}

// the PCG product in the difference form (both precisions); the FP64 true
// residual (RESID) in the standard form with the assembled diagonal
template <int K, bool RESID = false, typename T>
__device__ __forceinline__ T stencil_apply(const Stencil<T> &A, const T *__restrict__ x, int64_t i, T xi) {
    constexpr bool DIFF = !RESID;
    const T off = stencil_row<DIFF, K>(A, x, i, xi);
    return DIFF ? off : off + A.diag[i] * xi;
}

// Rows of the stencil kernels: each XCD (blocks b = x mod 8) owns one
// contiguous eighth of the rows and its blocks sweep it together, block b
// taking every (nblocks/8)-th run of BLOCK rows. The XCD's active window is
// then nblocks/8 * BLOCK rows wide and moves through the band, so rows one
// grid plane back (the largest offsets) were read a window or two earlier
// and are still in its L2.
template <int BLOCK>
struct StencilRows {
    int64_t start, end, step;
    __device__ StencilRows(int64_t n, int nblocks) {
        if ((nblocks & 7) != 0) {  // too few blocks: one contiguous run each
            const int64_t c = (n + nblocks - 1) / nblocks;
            start = static_cast<int64_t>(blockIdx.x) * c;
            end = min(n, start + c);
            step = BLOCK;
            return;
        }
        const int x = blockIdx.x & 7, b = blockIdx.x >> 3, nbx = nblocks >> 3;
        const int64_t band = ((n + 7) / 8 + BLOCK - 1) / BLOCK * BLOCK;
        start = x * band + static_cast<int64_t>(b) * BLOCK;
        end = min(n, (x + 1) * band);
        step = static_cast<int64_t>(nbx) * BLOCK;
    }
};

// K1 in stencil form: q = A p, partial p.q
template <typename T, int BLOCK, int K>
__global__ __launch_bounds__(BLOCK) void k_cg_spmv_stencil(const Stencil<T> A, const T *__restrict__ p,
                                                           T *__restrict__ q, double *__restrict__ part,
                                                           int64_t chunk, int nblocks, CgState *__restrict__ st,
                                                           int parity) {
    if (stopped_entry(st, parity)) return;
    __shared__ double sh[BLOCK / 64 + 1];
    (void)chunk;
    const StencilRows<BLOCK> R(A.n, nblocks);
    double acc = 0.0;
    for (int64_t row0 = R.start; row0 < R.end; row0 += R.step) {
        const int64_t i = row0 + threadIdx.x;
        if (i >= R.end) break;
        const T xi = p[i];
        const T y = stencil_apply<K>(A, p, i, xi);
        q[i] = y;
        acc += static_cast<double>(xi) * static_cast<double>(y);
    }
    const double tot = block_sum<BLOCK>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// FP64 r = -mask(A phi), stencil form (gate as in k_neg_residual_sell)
template <int K, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_neg_residual_stencil(const Stencil<double> A,
                                                                const double *__restrict__ phi,
                                                                const uint8_t *__restrict__ mask,
                                                                const double *__restrict__ b,
                                                                double *__restrict__ r, double *__restrict__ part,
                                                                int64_t chunk, int nblocks,
                                                                const RuState *__restrict__ gate) {
    if (gate && gate->act != 1) return;
    __shared__ double sh[BLOCK / 64 + 1];
    (void)chunk;
    const StencilRows<BLOCK> R(A.n, nblocks);
    double acc = 0.0;
    for (int64_t row0 = R.start; row0 < R.end; row0 += R.step) {
        const int64_t i = row0 + threadIdx.x;
        if (i >= R.end) break;
        const double y = stencil_apply<K, true>(A, phi, i, phi[i]);
        const double v = mask[i] ? 0.0 : (b ? b[i] - y : -y);
        r[i] = v;
        acc += v * v;
    }
    const double tot = block_sum<BLOCK>(acc, sh);
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
        st->alpha = 0.0;
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
                                                         const uint8_t *__restrict__ mask,
                                                         const double *__restrict__ b, double *__restrict__ r,
                                                         double *__restrict__ part, int64_t n, int64_t chunk,
                                                         int nblocks, const RuState *__restrict__ gate) {
    if (gate && gate->act != 1) return;
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
            const double v = mask[row] ? 0.0 : (b ? b[row] - s : -s);
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

// plain 16-byte streaming copy (bandwidth probe)
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

template <typename T>
__global__ void k_phi_axpy(double *__restrict__ phi, const T *__restrict__ x, double s, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) phi[i] += s * static_cast<double>(x[i]);
}

// H = -sum phi_i grad_i per element; |H|; energy partials 1/2 mu |H|^2 VE.
// Elements in caller order (perm null: element e is written at e) or in
// colour order (written at perm[e]); coef in the same order as tet.
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
        const int64_t o = perm ? perm[e] : e;
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

// ---- allocation-free field recovery ------------------------------------------
// Block partials of the energy sum live in this file-scope array, one copy per
// GPU for the life of the process. Host code never passes `scratch` itself to
// a kernel (in host code the name is a host-side stand-in): it uses the
// address hipGetSymbolAddress returned for the device, kept per device in
// g_scratch. Every call on a GPU shares the array: the field kernel
// overwrites the partials of all its blocks and the summing kernel reads only
// those, so no clearing is needed; calls that overlap on one GPU (two
// streams, two contexts) race on it.
constexpr int kScratchPartials = 4096;
static __device__ __align__(256) double scratch[kScratchPartials];
constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices] = {};

// Called with `device` current (fk_context_create), once per device.
void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) fail(FK_ERR_UNSUPPORTED, "device index beyond the library's table");
    if (g_scratch[device]) return;
    void *ptr = nullptr;
    FK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[device] = static_cast<double *>(ptr);
}

// out[k] = in[perm[k]]: a per-node device array from the caller's node order
// to the internal one
__global__ void k_gather_nodes(const double *__restrict__ in, const int32_t *__restrict__ perm, int64_t n,
                               double *__restrict__ out) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) out[k] = in[perm[k]];
}

// energy = part[0] + part[1] + ... in index order (the order fk_field sums
// the same partials on the host), one thread
__global__ void k_sum_in_order(const double *__restrict__ part, int n, double *__restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    *out = s;
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

// key of the captured iteration graphs: every launch parameter they bake in
static int solver_config(const fk_mesh *m) {
    return m->lanes + 100 * m->format + 1000 * m->sell_unroll + 10000 * m->sell_block;
}

// Blocks of the FP32 stencil SpMV: two runs of 256 rows per block in each
// XCD's band (8152 blocks at n=160, 4.17M rows). Measured there: 26.2 us,
// against 27.8 us with runs of 512 rows and 35.9 us for the 1024-block grid
// of the other kernels, whose blocks sweep 8 runs each and drift apart, so
// the rows one grid plane back have left L2 more often.
constexpr int kStencilBlock = 256;
static int stencil_blocks(const fk_mesh *m) {
    const int64_t band = ((m->nv + 7) / 8 + kStencilBlock - 1) / kStencilBlock * kStencilBlock;
    const int64_t runs = band / kStencilBlock;
    return static_cast<int>(8 * std::max<int64_t>(1, (runs + 1) / 2));
}

// Blocks of the FP32 SELL SpMV: `unroll` 64-row slices per wave (in flight
// together), so a block is as short-lived as the stencil kernel's and the
// blocks of an XCD sweep its band in dispatch order; each block leaves one
// p.q partial, which the next kernel re-sums in every block.
static int sell_blocks(const fk_mesh *m) {
    const int64_t ns = (m->nv + 63) / 64;
    const int64_t per = std::max(1, m->sell_block / 64) * std::max(1, m->sell_unroll);
    return static_cast<int>((((ns + per - 1) / per) + 7) / 8 * 8);
}

// blocks of the SpMV kernel, which writes one p.q partial per block
template <typename T>
static int spmv_blocks(const fk_mesh *m) {
    if (sizeof(T) == 4 && m->kind == FK_SPMV_STENCIL) return stencil_blocks(m);
    if (sizeof(T) == 4 && m->kind == FK_SPMV_SELL64) return sell_blocks(m);
    return m->nblocks;
}

template <typename T>
static void ensure_work(fk_mesh *m) {
    PcgWork<T> &w = work<T>(m);
    const size_t nv = static_cast<size_t>(m->nv), nb = static_cast<size_t>(m->nblocks);
    w.x.ensure(nv); w.r.ensure(nv); w.p.ensure(nv); w.q.ensure(nv); w.dinv.ensure(nv);
    w.part_pq.ensure(std::max<size_t>({nb, static_cast<size_t>(stencil_blocks(m)), static_cast<size_t>(sell_blocks(m))})); w.part_rz.ensure(2 * nb); w.part_rr.ensure(nb); w.part_bb.ensure(nb);
    w.state.ensure(1);
}

template <typename T>
static void compute_dinv(fk_mesh *m) {
    hipStream_t st = m->ctx->stream;
    PcgWork<T> &w = work<T>(m);
    DevBuf<int32_t> bad(1);
    FK_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), st));
    k_dinv<T><<<grid_for(m->nv, 256), 256, 0, st>>>(m->val64.get(), m->diag.get(), m->mask.get(), m->nv,
                                                    w.dinv.get(), bad.get());
    FK_HIP(hipGetLastError());
    int32_t h = 0;
    FK_HIP(hipMemcpyAsync(&h, bad.get(), sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    if (h) fail(FK_ERR_SINGULAR, std::to_string(h) + " free node(s) without a positive diagonal");
}

// run f(i) for i in [0, n) on up to 16 host threads (contiguous ranges)
template <typename F>
static void parallel_for(int64_t n, F f) {
    int nt = static_cast<int>(std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency())));
    if (n < 4096) nt = 1;
    if (nt == 1) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&f, n, nt, t] {
            for (int64_t i = n * t / nt; i < n * (t + 1) / nt; ++i) f(i);
        });
    for (auto &x : th) x.join();
}

// Build the SELL-64x4 copy (layout above) on the host from the CSR pattern:
// per slice, the shared-offset form when the union of the rows' offsets is
// at most 1.25x the longest row (each extra offset costs one padded value
// per row, each per-row index it saves costs a quarter to half a value),
// else per-row indices. Uploads the slice tables, the indices and the
// SELL -> CSR position map (-1 on padding) that later value refreshes use.
static void build_sell(fk_mesh *m) {
    if (m->have_sell) return;
    // captured iteration graphs hold the addresses of the arrays rebuilt here
    m->w64.reset_graph();
    m->w32.reset_graph();
    hipStream_t st = m->ctx->stream;
    const int64_t nv = m->nv, nnz = m->nnz;
    std::vector<int32_t> rp(static_cast<size_t>(nv + 1)), ci(static_cast<size_t>(nnz));
    FK_HIP(hipMemcpyAsync(rp.data(), m->rowptr.get(), (nv + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipMemcpyAsync(ci.data(), m->colidx.get(), nnz * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    const int64_t ns = (nv + 63) / 64;
    const bool allow_shared = m->format != FK_SPMV_SELL64_ROWIDX;
    // pass 1: widths, shared offset sets, bandwidth
    std::vector<int32_t> width(static_cast<size_t>(ns));
    std::vector<std::vector<int32_t>> shared(static_cast<size_t>(ns));
    std::vector<int32_t> band(static_cast<size_t>(ns), 0);
    parallel_for(ns, [&](int64_t s) {
        const int64_t r0 = s * 64, r1 = std::min(nv, r0 + 64);
        int32_t w = 0, b = 0;
        std::vector<int32_t> off;
        for (int64_t r = r0; r < r1; ++r) {
            int32_t len = 0;
            for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
                const int32_t d = ci[k] - static_cast<int32_t>(r);
                if (d == 0) continue;
                ++len;
                b = std::max(b, d < 0 ? -d : d);
                if (allow_shared) off.push_back(d);
            }
            w = std::max(w, len);
        }
        band[s] = b;
        if (allow_shared && w > 0) {
            std::sort(off.begin(), off.end());
            off.erase(std::unique(off.begin(), off.end()), off.end());
            if (static_cast<int32_t>(off.size()) * 4 <= 5 * w) {
                shared[s] = std::move(off);
                width[s] = -static_cast<int32_t>(shared[s].size());
                return;
            }
        }
        width[s] = w;
    });
    int32_t bw = 0, maxg = 0;
    for (int64_t s = 0; s < ns; ++s) {
        bw = std::max(bw, band[s]);
        maxg = std::max(maxg, (width[s] < 0 ? -width[s] : width[s]) >> 2);
    }
    m->sell_max_groups = maxg;
    m->bandwidth = bw;
    m->sell_short = m->allow_short && bw + 64 <= 32767;
    // pass 2: offsets of values and indices
    std::vector<int64_t> vptr(static_cast<size_t>(ns + 1), 0), cptr(static_cast<size_t>(ns), 0);
    int64_t ncol = 0, nsoff = 0, nshared = 0;
    for (int64_t s = 0; s < ns; ++s) {
        const int32_t W = width[s] < 0 ? -width[s] : width[s];
        vptr[s + 1] = vptr[s] + 64 * static_cast<int64_t>(W);
        if (width[s] < 0) {
            cptr[s] = nsoff;
            nsoff += W;
            ++nshared;
        } else {
            cptr[s] = ncol;
            ncol += 64 * static_cast<int64_t>(W);
        }
    }
    const int64_t len = vptr[ns];
    // pass 3: map and indices
    std::vector<int32_t> map(static_cast<size_t>(len), -1), soff(static_cast<size_t>(nsoff));
    std::vector<int32_t> col32(m->sell_short ? 0 : static_cast<size_t>(ncol));
    std::vector<int16_t> col16(m->sell_short ? static_cast<size_t>(ncol) : 0);
    parallel_for(ns, [&](int64_t s) {
        const int32_t W = width[s] < 0 ? -width[s] : width[s];
        const int32_t ng = W >> 2;
        auto pos = [&](int lane, int32_t k) {
            return k < 4 * ng ? 4 * (64 * static_cast<int64_t>(k >> 2) + lane) + (k & 3)
                              : 256 * static_cast<int64_t>(ng) + 64 * static_cast<int64_t>(k - 4 * ng) + lane;
        };
        if (width[s] < 0) {
            const std::vector<int32_t> &off = shared[s];
            std::copy(off.begin(), off.end(), soff.begin() + cptr[s]);
            for (int lane = 0; lane < 64; ++lane) {
                const int64_t r = s * 64 + lane;
                if (r >= nv) break;
                const int32_t *b = ci.data() + rp[r], *e = ci.data() + rp[r + 1];
                for (int32_t k = 0; k < W; ++k) {
                    const int64_t c = r + off[k];
                    if (c < 0 || c >= nv) continue;
                    const int32_t *f = std::lower_bound(b, e, static_cast<int32_t>(c));
                    if (f != e && *f == c) map[vptr[s] + pos(lane, k)] = static_cast<int32_t>(f - ci.data());
                }
            }
            return;
        }
        for (int lane = 0; lane < 64; ++lane) {
            const int64_t r = s * 64 + lane;
            const int32_t self = static_cast<int32_t>(r < nv ? r : nv - 1);
            int32_t k = 0;
            if (r < nv)
                for (int32_t q = rp[r]; q < rp[r + 1]; ++q) {
                    if (ci[q] == r) continue;
                    const int64_t o = pos(lane, k++);
                    map[vptr[s] + o] = q;
                    if (m->sell_short)
                        col16[cptr[s] + o] = static_cast<int16_t>(ci[q] - r);
                    else
                        col32[cptr[s] + o] = ci[q];
                }
            for (; k < W; ++k) {
                const int64_t o = pos(lane, k);
                if (m->sell_short)
                    col16[cptr[s] + o] = static_cast<int16_t>(self - r);
                else
                    col32[cptr[s] + o] = self;
            }
        }
    });
    m->nslices = ns;
    m->nshared = nshared;
    m->sell_len = len;
    m->sell_chunk = (ns + m->nblocks - 1) / m->nblocks;
    m->sell_ptr.alloc(ns + 1);
    m->sell_width.alloc(ns);
    m->sell_cptr.alloc(ns);
    m->sell_col.alloc(col32.size());
    m->sell_dcol.alloc(col16.size());
    m->sell_soff.alloc(soff.size());
    m->sell_map.alloc(len);
    FK_HIP(hipMemcpyAsync(m->sell_ptr.get(), vptr.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(m->sell_width.get(), width.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(m->sell_cptr.get(), cptr.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice, st));
    if (!col32.empty())
        FK_HIP(hipMemcpyAsync(m->sell_col.get(), col32.data(), col32.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                              st));
    if (!col16.empty())
        FK_HIP(hipMemcpyAsync(m->sell_dcol.get(), col16.data(), col16.size() * sizeof(int16_t),
                              hipMemcpyHostToDevice, st));
    if (!soff.empty())
        FK_HIP(hipMemcpyAsync(m->sell_soff.get(), soff.data(), soff.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                              st));
    if (len > 0)
        FK_HIP(hipMemcpyAsync(m->sell_map.get(), map.data(), len * sizeof(int32_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipStreamSynchronize(st));  // host vectors go out of scope
    m->have_sell = true;
}

// Stencil values from the FP64 CSR values, one thread per row: the row's
// entries lie in one or two cache lines, read once for all K columns (and the
// diagonal for the FP64 standard form).
// Grid of the stencil fill: 512 blocks of 256 threads (two per CU). Measured
// at n=160: the step 0.6 ms shorter than with one thread per row over the
// whole grid, whose in-flight CSR lines overflow L2 and are fetched again for
// each column; 2048 blocks gained nothing.
constexpr int kFillBlocks = 512;
template <typename T>
__global__ void k_stencil_fill(const double *__restrict__ val64, const int32_t *__restrict__ map,
                               const int32_t *__restrict__ diagpos, int K, int64_t ld, int64_t n, T *__restrict__ out,
                               double *__restrict__ diag64) {
    // bounded grid (a multiple of 8): the blocks of XCD x sweep its eighth of
    // the rows together, so the CSR lines of the rows in flight fit its L2
    const int x = blockIdx.x & 7, bx = blockIdx.x >> 3, nbx = gridDim.x >> 3;
    const int64_t band = ((ld + 7) / 8 + 255) / 256 * 256;
    const int64_t end = min(ld, (x + 1) * band);
    for (int64_t i = x * band + static_cast<int64_t>(bx) * blockDim.x + threadIdx.x; i < end;
         i += static_cast<int64_t>(nbx) * blockDim.x) {
        for (int k = 0; k < K; ++k) {
            const int32_t pos = map[k * ld + i];
            out[k * ld + i] = pos >= 0 ? static_cast<T>(val64[pos]) : T(0);
        }
        if (diag64 && i < n) {
            const int32_t d = diagpos[i];
            diag64[i] = d >= 0 ? val64[d] : 0.0;
        }
    }
}

// Stencil copy (layout at k_cg_spmv_stencil), host build once per mesh:
// stencil_ok when the positive column offsets of all rows form one set of at
// most 16 and the columns hold at most 25% padding.
static void build_stencil(fk_mesh *m) {
    if (m->stencil_built) return;
    m->stencil_built = true;
    m->stencil_ok = false;
    m->w64.reset_graph();
    m->w32.reset_graph();
    hipStream_t st = m->ctx->stream;
    const int64_t nv = m->nv, nnz = m->nnz;
    std::vector<int32_t> rp(static_cast<size_t>(nv + 1)), ci(static_cast<size_t>(nnz));
    FK_HIP(hipMemcpyAsync(rp.data(), m->rowptr.get(), (nv + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipMemcpyAsync(ci.data(), m->colidx.get(), nnz * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    // union of the positive offsets, per thread then merged (gives up past 16)
    const int nth = static_cast<int>(std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency())));
    std::vector<std::vector<int32_t>> local(static_cast<size_t>(nth));
    {
        std::vector<std::thread> th;
        for (int q = 0; q < nth; ++q)
            th.emplace_back([&, q] {
                std::vector<int32_t> &P = local[q];
                for (int64_t i = nv * q / nth; i < nv * (q + 1) / nth; ++i)
                    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
                        const int32_t d = ci[k] - static_cast<int32_t>(i);
                        if (d <= 0) continue;
                        auto it = std::lower_bound(P.begin(), P.end(), d);
                        if (it == P.end() || *it != d) {
                            if (P.size() >= 16) {
                                P.assign(17, 0);  // marks "too many"
                                return;
                            }
                            P.insert(it, d);
                        }
                    }
            });
        for (auto &x : th) x.join();
    }
    std::vector<int32_t> P;
    for (auto &l : local) {
        if (l.size() > 16) return;
        for (int32_t d : l) {
            auto it = std::lower_bound(P.begin(), P.end(), d);
            if (it == P.end() || *it != d) P.insert(it, d);
        }
        if (P.size() > 16) return;
    }
    const int32_t K = static_cast<int32_t>(P.size());
    const int64_t ld = (nv + 63) / 64 * 64;
    const int64_t pairs = (nnz - nv) / 2;  // off-diagonal pairs
    if (K == 0 || 4 * K * nv > 5 * pairs + 256 * K) return;
    std::vector<int32_t> map(static_cast<size_t>(K * ld), -1);
    parallel_for(nv, [&](int64_t i) {
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const int32_t d = ci[k] - static_cast<int32_t>(i);
            if (d <= 0) continue;
            const int64_t c = std::lower_bound(P.begin(), P.end(), d) - P.begin();
            map[c * ld + i] = k;
        }
    });
    m->st_map.alloc(K * ld);
    FK_HIP(hipMemcpyAsync(m->st_map.get(), map.data(), K * ld * sizeof(int32_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipStreamSynchronize(st));
    m->st_K = K;
    for (int k = 0; k < 16; ++k) m->st_off[k] = k < K ? P[k] : 0;
    m->st_ld = ld;
    m->stencil_ok = true;
}

// the solver format actually used (FK_SPMV_CSR, FK_SPMV_SELL64, FK_SPMV_STENCIL)
static int solver_kind(fk_mesh *m) {
    if (m->kind >= 0) return m->kind;
    switch (m->format) {
        case FK_SPMV_CSR: m->kind = FK_SPMV_CSR; break;
        case FK_SPMV_STENCIL:
            build_stencil(m);
            if (!m->stencil_ok)
                fail(FK_ERR_UNSUPPORTED, "the matrix pattern is not a stencil of at most 16 offsets");
            m->kind = FK_SPMV_STENCIL;
            break;
        case FK_SPMV_AUTO:
            build_stencil(m);
            m->kind = m->stencil_ok ? FK_SPMV_STENCIL : FK_SPMV_SELL64;
            break;
        default: m->kind = FK_SPMV_SELL64; break;
    }
    return m->kind;
}

template <typename T>
static Stencil<T> stencil_view(fk_mesh *m, const T *val, const T *diag) {
    Stencil<T> A{val, diag, m->st_ld, m->nv, m->st_K, {}};
    for (int k = 0; k < 16; ++k) A.off[k] = m->st_off[k];
    return A;
}

// bring the solver's copy of the values up to date with val64 / val32
template <typename T>
static void sync_solver_values(fk_mesh *m);
template <>
void sync_solver_values<double>(fk_mesh *m) {
    if (m->format == FK_SPMV_CSR || m->sell64_valid) return;
    if (solver_kind(m) == FK_SPMV_STENCIL) {
        const int64_t len = m->st_K * m->st_ld;
        m->st_val64.ensure(len);
        m->sell_diag64.ensure(m->nv);
        k_stencil_fill<double><<<kFillBlocks, 256, 0, m->ctx->stream>>>(
            m->val64.get(), m->st_map.get(), m->diag.get(), m->st_K, m->st_ld, m->nv, m->st_val64.get(),
            m->sell_diag64.get());
        FK_HIP(hipGetLastError());
        m->sell64_valid = true;
        return;
    }
    build_sell(m);
    m->sell_val64.ensure(m->sell_len);
    m->sell_diag64.ensure(m->nv);
    if (m->sell_len > 0)
        k_sell_gather<double><<<grid_for(m->sell_len, 256), 256, 0, m->ctx->stream>>>(
            m->val64.get(), m->sell_map.get(), m->sell_len, m->sell_val64.get());
    k_sell_gather<double><<<grid_for(m->nv, 256), 256, 0, m->ctx->stream>>>(m->val64.get(), m->diag.get(), m->nv,
                                                                            m->sell_diag64.get());
    FK_HIP(hipGetLastError());
    m->sell64_valid = true;
}
template <>
void sync_solver_values<float>(fk_mesh *m) {
    if (m->format == FK_SPMV_CSR || m->sell32_valid) return;
    if (solver_kind(m) == FK_SPMV_STENCIL) {
        const int64_t len = m->st_K * m->st_ld;
        m->st_val32.ensure(len);
        k_stencil_fill<float><<<kFillBlocks, 256, 0, m->ctx->stream>>>(
            m->val64.get(), m->st_map.get(), m->diag.get(), m->st_K, m->st_ld, m->nv, m->st_val32.get(), nullptr);
        FK_HIP(hipGetLastError());
        m->sell32_valid = true;
        return;
    }
    make_val32(m);
    build_sell(m);
    m->sell_val32.ensure(m->sell_len);
    if (m->sell_len > 0)
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
static const T *sell_diag(fk_mesh *m);
template <>
const double *sell_diag<double>(fk_mesh *m) { return m->sell_diag64.get(); }
template <>
const float *sell_diag<float>(fk_mesh *) { return nullptr; }

template <typename T, typename IDX>
static Sell<T, IDX> sell_view(fk_mesh *m, const IDX *col) {
    return Sell<T, IDX>{m->sell_ptr.get(), m->sell_width.get(), m->sell_cptr.get(), col, m->sell_soff.get(),
                        sell_values<T>(m), sell_diag<T>(m), m->nv};
}

// Groups of 4 entries the batched SELL row loads at once (FP32 product):
// 4 when the widest slice has at most 4 groups (19 entries), else 0 (group
// by group). Measured at n=160 in RCM numbering (58.5M stored entries, at
// most 4 groups), FP32 / FP64 us per SpMV: 256 threads 80.1 / 145.7 batched,
// 84.2 / 148.1 group by group; 1024 threads 100.0 / 145.2 batched, 87.2 /
// 142.7 group by group; 8-group instances 103-126 / 206-598 (their padding
// loads). The FP32 product therefore runs batched on 256-thread blocks, the
// FP64 product and residual group by group on 1024 (the residual on the
// solver grid).
static int sell_batch_groups(const fk_mesh *m) { return m->sell_max_groups <= 4 ? 4 : 0; }

template <typename T, typename IDX, int BLOCK, int U>
static void launch_sell(fk_mesh *m, PcgWork<T> &w, const IDX *col, int parity) {
    const int nb = spmv_blocks<T>(m);
    const int64_t chunk = nb == m->nblocks ? m->sell_chunk : (m->nslices + nb - 1) / nb;
    const Sell<T, IDX> A = sell_view<T, IDX>(m, col);
    hipStream_t st = m->ctx->stream;
    if (sizeof(T) == 4 && sell_batch_groups(m) == 4)
        k_cg_spmv_sell<T, IDX, BLOCK, U, 4><<<nb, BLOCK, 0, st>>>(A, w.p.get(), w.q.get(), w.part_pq.get(),
                                                                 m->nslices, chunk, nb, w.state.get(), parity);
    else
        k_cg_spmv_sell<T, IDX, BLOCK, U, 0><<<nb, BLOCK, 0, st>>>(A, w.p.get(), w.q.get(), w.part_pq.get(),
                                                                 m->nslices, chunk, nb, w.state.get(), parity);
}

template <typename IDX>
static void launch_neg_residual_sell(fk_mesh *m, const IDX *col, const double *phi, const double *b, double *r,
                                     const RuState *gate) {
    const Sell<double, IDX> A = sell_view<double, IDX>(m, col);
    k_neg_residual_sell<IDX, 0><<<m->nblocks, kBlock, 0, m->ctx->stream>>>(
        A, phi, m->mask.get(), b, r, m->part_misc.get(), m->nslices, m->sell_chunk, m->nblocks, gate);
}

template <typename T, typename IDX>
static void launch_sell_shape(fk_mesh *m, PcgWork<T> &w, const IDX *col, int parity) {
    if (sizeof(T) == 8) {  // FP64: 1024-thread blocks on the solver grid
        launch_sell<T, IDX, 1024, 1>(m, w, col, parity);
        return;
    }
    const int shape = m->sell_block * 4 + m->sell_unroll;
    switch (shape) {
        case 256 * 4 + 2: launch_sell<T, IDX, 256, 2>(m, w, col, parity); break;
        case 512 * 4 + 1: launch_sell<T, IDX, 512, 1>(m, w, col, parity); break;
        case 512 * 4 + 2: launch_sell<T, IDX, 512, 2>(m, w, col, parity); break;
        case 1024 * 4 + 1: launch_sell<T, IDX, 1024, 1>(m, w, col, parity); break;
        default: launch_sell<T, IDX, 256, 1>(m, w, col, parity); break;
    }
}

template <typename T>
static const T *stencil_values(fk_mesh *m);
template <>
const double *stencil_values<double>(fk_mesh *m) { return m->st_val64.get(); }
template <>
const float *stencil_values<float>(fk_mesh *m) { return m->st_val32.get(); }

template <typename T, int BLOCK, int K>
static void launch_stencil_k(fk_mesh *m, PcgWork<T> &w, int parity) {
    const int nb = spmv_blocks<T>(m);
    k_cg_spmv_stencil<T, BLOCK, K><<<nb, BLOCK, 0, m->ctx->stream>>>(
        stencil_view<T>(m, stencil_values<T>(m), sell_diag<T>(m)), w.p.get(), w.q.get(), w.part_pq.get(), m->chunk,
        nb, w.state.get(), parity);
}

// dispatch on the number of positive offsets (1..16, a compile-time constant
// of the kernel)
template <int K = 1>
struct StencilK {
    template <typename F>
    static void run(int k, F &&f) {
        if (k == K)
            f(std::integral_constant<int, K>{});
        else
            StencilK<K + 1>::run(k, static_cast<F &&>(f));
    }
};
template <>
struct StencilK<17> {
    template <typename F>
    static void run(int, F &&) {
        fail(FK_ERR_UNSUPPORTED, "stencil with more than 16 offsets");
    }
};

template <typename T, int BLOCK>
static void launch_stencil(fk_mesh *m, PcgWork<T> &w, int parity) {
    StencilK<>::run(m->st_K, [&](auto kc) { launch_stencil_k<T, BLOCK, decltype(kc)::value>(m, w, parity); });
}

template <typename T>
static void launch_spmv_lanes(fk_mesh *m, PcgWork<T> &w, int parity) {
    if (m->kind == FK_SPMV_STENCIL) {
        // FP32: the short-lived blocks of stencil_blocks; FP64: 1024 blocks
        // of 512 threads (8 waves per SIMD for its guarded loads)
        if (sizeof(T) == 4)
            launch_stencil<T, kStencilBlock>(m, w, parity);
        else
            launch_stencil<T, 512>(m, w, parity);
        return;
    }
    if (m->format != FK_SPMV_CSR) {
        if (m->sell_short)
            launch_sell_shape<T, int16_t>(m, w, m->sell_dcol.get(), parity);
        else
            launch_sell_shape<T, int32_t>(m, w, m->sell_col.get(), parity);
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
    const int npq = spmv_blocks<T>(m);
    const int pairs = (npq + 1) / 2;
    auto xr = [&](auto kern) {
        kern<<<nb, kBlock, 0, st>>>(w.r.get(), w.q.get(), w.dinv.get(), w.part_pq.get(), npq, rz_cur, rz_next,
                                    w.part_rr.get(), m->nv, m->chunk, nb, w.state.get(), parity);
    };
    if (pairs <= 2 * kBlock)
        xr(k_cg_update_xr<T, 2>);
    else if (pairs <= 4 * kBlock)
        xr(k_cg_update_xr<T, 4>);
    else if (pairs <= 8 * kBlock)
        xr(k_cg_update_xr<T, 8>);
    else
        xr(k_cg_update_xr<T, 16>);
    k_cg_update_p<T><<<nb, kBlock, 0, st>>>(w.x.get(), w.p.get(), w.r.get(), w.dinv.get(), rz_next, rz_cur,
                                            w.part_rr.get(),
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
    const int cfg = solver_config(m);
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

static double neg_residual(fk_mesh *m, const double *phi, double *r, const double *b = nullptr);
static void enqueue_neg_residual(fk_mesh *m, const double *phi, double *r, const RuState *gate,
                                 const double *b = nullptr);

constexpr int kMaxRestarts = 32;  // direction restarts before the FP32 solve gives up

// ---- FP32 solve with FP64 reliable updates, decided on the device --------
// After every chunk of G iterations, five small kernels run (no-ops unless
// the check asks for a replacement):
//   check   the recursion's residual fell by delta since the last replacement,
//           or it stopped: fold (1) or, after an FP32 breakdown, drop (2) its
//           increment
//   fold    phi += x / s (act 1), x = 0
//   resid   r64 = -mask(A64 phi) and its partials (act 1 only)
//   decide  converged / failed / replace (keep p) / restart (p = D^-1 r), and
//           the next stop threshold max(target, delta^2 ||r||^2)
//   replace r = s r64, rho partials, and p on a restart; then resume
__global__ void k_ru_check(const CgState *__restrict__ cg, RuState *__restrict__ ru) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    ru->act = 0;
    if (ru->finished) return;
    const bool stopped = (cg->done_at[0] | cg->done_at[1]) != 0;
    if (!stopped) {
        // replace at the end of the chunk in which the residual fell by delta
        if (cg->rr < ru->delta2 * ru->rr_last) ru->act = 1;
        return;
    }
    const bool broken = cg->status == FK_ERR_SINGULAR || cg->status == FK_ERR_NONFINITE || !isfinite(cg->rr);
    if (!broken) {
        ru->act = 1;
        return;
    }
    if (++ru->restarts > kMaxRestarts || cg->iters >= ru->max_iter) {
        ru->finished = 1;
        ru->status = cg->iters >= ru->max_iter ? FK_ERR_NOT_CONVERGED
                                               : (cg->status != FK_OK ? cg->status : FK_ERR_NONFINITE);
        return;
    }
    ru->act = 2;
}

template <typename T>
__global__ void k_ru_fold(double *__restrict__ phi, T *__restrict__ x, const RuState *__restrict__ ru, int64_t n) {
    const int act = ru->act;
    if (act == 0) return;
    const double inv_s = 1.0 / ru->s;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (act == 1) phi[i] += inv_s * static_cast<double>(x[i]);
        x[i] = T(0);
    }
}

__global__ __launch_bounds__(kBlock) void k_ru_decide(const double *__restrict__ part, int nblocks,
                                                      CgState *__restrict__ cg, RuState *__restrict__ ru) {
    if (ru->act == 0) return;
    __shared__ double sh[kBlock / 64 + 1];
    const double rr = ru->act == 1 ? sum_partials<kBlock>(part, nblocks, sh) : 0.0;
    if (threadIdx.x != 0) return;
    if (ru->act == 2) {
        ru->mode = 2;  // r64 still holds the last true residual
        return;
    }
    ru->rr64 = rr;
    ru->nrep += 1;
    const double s = ru->s;
    if (rr < ru->thresh64) {
        ru->finished = 1;
        ru->status = FK_OK;
        return;
    }
    if (!isfinite(rr)) {
        ru->finished = 1;
        ru->status = FK_ERR_NONFINITE;
        return;
    }
    if (cg->iters >= ru->max_iter) {
        ru->finished = 1;
        ru->status = FK_ERR_NOT_CONVERGED;
        return;
    }
    const double scaled = rr * s * s;
    int mode = 1;
    if (scaled > 4.0 * cg->rr) {
        // the recursion drifted below the true residual: its direction no
        // longer matches it, restart from p = D^-1 r
        if (++ru->restarts > kMaxRestarts) {
            ru->finished = 1;
            ru->status = FK_ERR_NOT_CONVERGED;
            return;
        }
        mode = 2;
    }
    ru->mode = mode;
    ru->rr_last = scaled;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_ru_replace(T *__restrict__ r, T *__restrict__ p,
                                                       const double *__restrict__ r64, const T *__restrict__ dinv,
                                                       double *__restrict__ part_rz, const RuState *__restrict__ ru,
                                                       int64_t n, int64_t chunk, int nblocks) {
    if (ru->act == 0 || ru->finished) return;
    __shared__ double sh[kBlock / 64 + 1];
    const double s = ru->s;
    const bool restart = ru->mode == 2;
    const int lb = logical_block(nblocks);
    const int64_t r0 = static_cast<int64_t>(lb) * chunk;
    const int64_t r1 = min(n, r0 + chunk);
    double rz = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
        const T d = dinv[i];
        const T ri = d != T(0) ? static_cast<T>(s * r64[i]) : T(0);
        r[i] = ri;
        if (restart) p[i] = d * ri;
        rz += static_cast<double>(ri) * static_cast<double>(d * ri);
    }
    const double t = block_sum<kBlock>(rz, sh);
    if (threadIdx.x == 0) part_rz[blockIdx.x] = t;  // parity-0 slot: rho of the next iteration
}

__global__ void k_ru_resume(CgState *__restrict__ cg, RuState *__restrict__ ru) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (ru->act != 0 && !ru->finished) {
        cg->done_at[0] = 0;
        cg->done_at[1] = 0;
        cg->status = FK_OK;
    }
    ru->act = 0;
}

// G iterations (G even: the chunk ends on parity 0) and the replacement step
// of the system A u = b (b null: 0) on the free rows
static void enqueue_ru_chunk(fk_mesh *m, PcgWork<float> &w, int G, bool timed, double *u, const double *b) {
    hipStream_t st = m->ctx->stream;
    const int nb = m->nblocks;
    for (int k = 0; k < G; ++k) enqueue_iteration<float>(m, w, k & 1, timed);
    RuState *ru = m->ru.get();
    k_ru_check<<<1, 64, 0, st>>>(w.state.get(), ru);
    k_ru_fold<float><<<nb, kBlock, 0, st>>>(u, w.x.get(), ru, m->nv);
    enqueue_neg_residual(m, u, m->r64.get(), ru, b);
    k_ru_decide<<<1, kBlock, 0, st>>>(m->part_misc.get(), nb, w.state.get(), ru);
    k_ru_replace<float><<<nb, kBlock, 0, st>>>(w.r.get(), w.p.get(), m->r64.get(), w.dinv.get(), w.part_rz.get(), ru,
                                               m->nv, m->chunk, nb);
    k_ru_resume<<<1, 64, 0, st>>>(w.state.get(), ru);
    FK_HIP(hipGetLastError());
}

// FP32 Jacobi-PCG whose residual is replaced by the FP64 true residual
// mask(b - A64 u) (u = phi and b = 0 for the Dirichlet problem; b the
// right-hand side of an increment otherwise), u accumulating the FP32
// increments in FP64 ("reliable updates"). A replacement happens at the end
// of the iteration chunk in which the FP32 recursive residual has fallen by
// the factor `delta` since the last replacement, or once the recursion stops
// (target reached, breakdown); it keeps the search direction, so the Krylov
// space is not restarted. Convergence is decided on the FP64 true residual:
// ||r|| < rtol ||b||. Returns (iterations, replacements, final ||r||^2).
static int solve_f32_replace(fk_mesh *m, double *u, const double *b, double rtol, int64_t max_iter, double bb,
                             double delta, int64_t *iters, int32_t *nrep, double *rr_out) {
    hipStream_t st = m->ctx->stream;
    const int64_t nv = m->nv;
    PcgWork<float> &w = m->w32;
    const int nb = m->nblocks;
    const double thresh = rtol * rtol * bb;
    double rr = neg_residual(m, u, m->r64.get(), b);
    *iters = 0;
    *nrep = 0;
    *rr_out = rr;
    if (rr < thresh) return FK_OK;
    const double s = 1.0 / std::sqrt(rr);
    const double target2 = thresh * s * s;
    sync_solver_values<float>(m);
    k_scale_to_f32<<<grid_for(nv, 256), 256, 0, st>>>(m->r64.get(), w.r.get(), s, nv);
    k_cg_init<float><<<nb, kBlock, 0, st>>>(w.x.get(), w.r.get(), w.p.get(), w.dinv.get(), w.part_rz.get(),
                                            w.part_bb.get(), nv, m->chunk, nb);
    // the recursion itself stops only at the target (scaled units)
    k_cg_start<<<1, kBlock, 0, st>>>(w.part_bb.get(), nb, 1.0, std::sqrt(target2), max_iter, w.state.get());
    FK_HIP(hipGetLastError());
    if (!m->ru) {
        m->ru.alloc(1);
        m->hru.alloc(2);
    }
    RuState init{};
    init.max_iter = max_iter;
    init.s = s;
    init.rr64 = rr;
    init.rr_last = 1.0;
    init.thresh64 = thresh;
    init.target2 = target2;
    init.delta2 = delta * delta;
    FK_HIP(hipMemcpyAsync(m->ru.get(), &init, sizeof(RuState), hipMemcpyHostToDevice, st));
    m->part_misc.ensure(nb);
    const int G = nv > 200000 ? 32 : 16;
    const char *tr = std::getenv("FEMEKO_HIP_TRACE");
    const bool trace = tr && tr[0] == '1';
    const bool timed = m->ctx->timing.enabled;
    RuState *hr = m->hru.ptr;
    CgState *hs = m->hstate.ptr;
    auto report = [&](const RuState &x) {
        if (trace)
            std::fprintf(stderr, "[fk trace] replacements %d restarts %d rr64 %.3e rr_last(scaled) %.3e%s\n", x.nrep,
                         x.restarts, x.rr64, x.rr_last, x.finished ? " finished" : "");
    };
    if (timed) {
        // synchronous chunks, events around every SpMV
        int64_t before = 0;
        while (true) {
            enqueue_ru_chunk(m, w, G, true, u, b);
            FK_HIP(hipMemcpyAsync(hr, m->ru.get(), sizeof(RuState), hipMemcpyDeviceToHost, st));
            FK_HIP(hipMemcpyAsync(hs, w.state.get(), sizeof(CgState), hipMemcpyDeviceToHost, st));
            FK_HIP(hipStreamSynchronize(st));
            harvest_timing(m->ctx, hs->iters - before);
            before = hs->iters;
            report(hr[0]);
            if (hr[0].finished) break;
        }
    } else {
        hipGraphExec_t *gx = nullptr;
        if (m->ctx->graphs) {
            const int cfg = solver_config(m) * 64 + G;
            if (!w.ru_graph || w.ru_graph_cfg != cfg || w.ru_graph_u != u || w.ru_graph_b != b ||
                w.ru_graph_mask != m->mask.get()) {
                if (w.ru_graph) (void)hipGraphExecDestroy(w.ru_graph);
                w.ru_graph = nullptr;
                hipGraph_t graph;
                FK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                enqueue_ru_chunk(m, w, G, false, u, b);
                FK_HIP(hipStreamEndCapture(st, &graph));
                hipError_t e = hipGraphInstantiate(&w.ru_graph, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                FK_HIP(e);
                w.ru_graph_cfg = cfg;
                w.ru_graph_u = u;
                w.ru_graph_b = b;
                w.ru_graph_mask = m->mask.get();
            }
            gx = &w.ru_graph;
        }
        auto launch = [&] {
            if (gx)
                FK_HIP(hipGraphLaunch(*gx, st));
            else
                enqueue_ru_chunk(m, w, G, false, u, b);
        };
        // one chunk in flight ahead of the state copy being examined
        hipEvent_t ev[2];
        FK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        FK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
        try {
            launch();
            FK_HIP(hipMemcpyAsync(&hr[0], m->ru.get(), sizeof(RuState), hipMemcpyDeviceToHost, st));
            FK_HIP(hipEventRecord(ev[0], st));
            int cur = 0;
            int last_rep = -1;
            while (true) {
                launch();
                FK_HIP(hipMemcpyAsync(&hr[1 - cur], m->ru.get(), sizeof(RuState), hipMemcpyDeviceToHost, st));
                FK_HIP(hipEventRecord(ev[1 - cur], st));
                FK_HIP(hipEventSynchronize(ev[cur]));
                if (hr[cur].nrep != last_rep) {
                    report(hr[cur]);
                    last_rep = hr[cur].nrep;
                }
                if (hr[cur].finished) break;
                cur = 1 - cur;
            }
            FK_HIP(hipStreamSynchronize(st));
        } catch (...) {
            (void)hipEventDestroy(ev[0]);
            (void)hipEventDestroy(ev[1]);
            throw;
        }
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
    }
    FK_HIP(hipMemcpyAsync(&hr[0], m->ru.get(), sizeof(RuState), hipMemcpyDeviceToHost, st));
    FK_HIP(hipMemcpyAsync(hs, w.state.get(), sizeof(CgState), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    *iters = hs->iters;
    *nrep = hr[0].nrep;
    *rr_out = hr[0].rr64;
    return hr[0].status;
}

template <int LPR>
static void launch_neg_residual(fk_mesh *m, const double *phi, const double *b, double *r, double *part,
                                const RuState *gate) {
    k_neg_residual<LPR><<<m->nblocks, kBlock, 0, m->ctx->stream>>>(m->rowptr.get(), m->colidx.get(),
                                                                   m->val64.get(), phi, m->mask.get(), b, r, part,
                                                                   m->nv, m->chunk, m->nblocks, gate);
}

// enqueue r = mask(b - A64 phi) (b null: 0) and its block partials of r.r
// into part_misc (no host sync); gate as in the kernels
static void enqueue_neg_residual(fk_mesh *m, const double *phi, double *r, const RuState *gate, const double *b) {
    hipStream_t st = m->ctx->stream;
    if (m->kind == FK_SPMV_STENCIL) {
        StencilK<>::run(m->st_K, [&](auto kc) {
            // 512-thread blocks: 8 waves per SIMD, as the FP64 solver SpMV
            // (measured 97 us against 133 us with 256-thread blocks at n=160)
            k_neg_residual_stencil<decltype(kc)::value, 512><<<m->nblocks, 512, 0, st>>>(
                stencil_view<double>(m, m->st_val64.get(), m->sell_diag64.get()), phi, m->mask.get(), b, r,
                m->part_misc.get(), m->chunk, m->nblocks, gate);
        });
    } else if (m->format != FK_SPMV_CSR) {
        if (m->sell_short)
            launch_neg_residual_sell<int16_t>(m, m->sell_dcol.get(), phi, b, r, gate);
        else
            launch_neg_residual_sell<int32_t>(m, m->sell_col.get(), phi, b, r, gate);
    } else switch (m->lanes) {
        case 4: launch_neg_residual<4>(m, phi, b, r, m->part_misc.get(), gate); break;
        case 8: launch_neg_residual<8>(m, phi, b, r, m->part_misc.get(), gate); break;
        case 32: launch_neg_residual<32>(m, phi, b, r, m->part_misc.get(), gate); break;
        case 64: launch_neg_residual<64>(m, phi, b, r, m->part_misc.get(), gate); break;
        default: launch_neg_residual<16>(m, phi, b, r, m->part_misc.get(), gate); break;
    }
    FK_HIP(hipGetLastError());
}

// r = mask(b - A64 phi) (b null: 0); returns ||r||^2 (host sync)
static double neg_residual(fk_mesh *m, const double *phi, double *r, const double *b) {
    hipStream_t st = m->ctx->stream;
    m->part_misc.ensure(m->nblocks);
    if (m->format != FK_SPMV_CSR) sync_solver_values<double>(m);
    enqueue_neg_residual(m, phi, r, nullptr, b);
    FK_HIP(hipMemcpyAsync(m->hpart.ptr, m->part_misc.get(), m->nblocks * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (int i = 0; i < m->nblocks; ++i) s += m->hpart.ptr[i];
    return s;
}

// ---- entry points for the other translation units (nonlinear loop, adjoint)
// FP32-refined solve of A u = b on the free rows with u = 0 on the fixed
// nodes (u, b FP64 on the device, nv values; u is overwritten), to
// ||mask(b - A u)|| < rtol ||mask b||: the increment of a Picard or Newton
// step. Returns the status and the FP32 iterations.
static int solve_f32_from_zero(fk_mesh *m, double *u, const double *b, double rtol, int64_t max_iter,
                               double bb_ref, int64_t *iters, double *rr_out, double *bb_out) {
    hipStream_t st = m->ctx->stream;
    m->r64.ensure(m->nv);
    ensure_work<float>(m);
    if (m->format == FK_SPMV_CSR) make_val32(m);
    if (!m->dinv32_valid) {
        compute_dinv<float>(m);
        m->dinv32_valid = true;
    }
    FK_HIP(hipMemsetAsync(u, 0, m->nv * sizeof(double), st));
    const double bb0 = neg_residual(m, u, m->r64.get(), b);  // ||mask b||^2
    const double bb = bb_ref >= 0.0 ? bb_ref : bb0;
    *iters = 0;
    *rr_out = bb0;
    *bb_out = bb;
    if (!std::isfinite(bb0) || !std::isfinite(bb)) return FK_ERR_NONFINITE;
    if (!(bb0 > 0.0)) return FK_OK;
    int32_t nrep = 0;
    return solve_f32_replace(m, u, b, rtol, max_iter, bb, 0.95, iters, &nrep, rr_out);
}

int solve_f32_increment(fk_mesh *m, double *u, const double *b, double rtol, int64_t max_iter, int64_t *iters) {
    double rr = 0.0, bb = 0.0;
    return solve_f32_from_zero(m, u, b, rtol, max_iter, -1.0, iters, &rr, &bb);
}

// The same with every node free, for the consistent singular Neumann system
// (A 1 = 0, b summing to zero): the FP32 solver runs with an all-zero mask
// swapped in for the Dirichlet one (D^-1 and the FP64 residual on every row),
// which is swapped back, and D^-1 marked stale, on every way out. The
// difference-form product maps constants to zero exactly, as A does. Stops at
// rtol relative to sqrt(bb_ref) (bb_ref < 0: ||b||); *rr_out and *bb_out
// return the final FP64 ||b - A u||^2 and the reference.
int solve_f32_free(fk_mesh *m, double *u, const double *b, double rtol, int64_t max_iter, double bb_ref,
                   int64_t *iters, double *rr_out, double *bb_out) {
    if (m->mask_free.n < static_cast<size_t>(m->nv)) {
        m->mask_free.ensure(m->nv);
        FK_HIP(hipMemsetAsync(m->mask_free.get(), 0, m->nv, m->ctx->stream));
    }
    struct Swap {
        fk_mesh *m;
        explicit Swap(fk_mesh *mm) : m(mm) {
            std::swap(m->mask, m->mask_free);
            m->dinv32_valid = false;
        }
        ~Swap() {
            std::swap(m->mask, m->mask_free);
            m->dinv32_valid = false;
        }
    } swap(m);
    return solve_f32_from_zero(m, u, b, rtol, max_iter, bb_ref, iters, rr_out, bb_out);
}

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

void set_dirichlet(fk_mesh *m, const uint8_t *mask_caller, const double *g_caller) {
    hipStream_t st = m->ctx->stream;
    const NodesIn<uint8_t> mask_in(m, mask_caller);
    const NodesIn<double> g_in(m, g_caller);
    const uint8_t *mask = mask_in.ptr;
    const double *g = g_in.ptr;
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
        if (m->format == FK_SPMV_CSR) make_val32(m);  // the CSR kernels read the FP32 CSR copy
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
        // residual fell by this factor (values outside (0,1) mean 0.95)
        const double delta = (inner_rtol > 0.0 && inner_rtol < 1.0) ? inner_rtol : 0.95;
        status = solve_f32_replace(m, m->phi.get(), nullptr, rtol, max_iter, bb, delta, &total, &nout, &rr);
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
    const NodesIn<double> b_in(m, b);
    NodesOut<double> x_out(m, x);
    FK_HIP(hipMemcpyAsync(w.r.get(), b_in.ptr, m->nv * sizeof(double), hipMemcpyHostToDevice, st));
    const CgState s = run_pcg<double>(m, rtol, max_iter, -1.0);
    FK_HIP(hipMemcpyAsync(x_out.ptr, w.x.get(), m->nv * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    x_out.commit();
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
    // caller element order: coalesced H writes, node reads local to the mesh
    k_field<<<nb, kBlock, 0, st>>>(m->tet_nat.get(), m->X.get(), m->Y.get(), m->Z.get(), nullptr, m->phi.get(),
                                   energy ? coef_caller_order(m, true) : nullptr, m->nt, H ? dH.get() : nullptr,
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

// Back-to-back launches of the solver's SpMV (precision T, the format the
// solver resolved) on its own work vectors, timed by two events on the
// library stream: the kernel's duration without per-launch event overhead.
// p is zeroed first (the duration does not depend on the values); a state
// of its own keeps the kernels from returning at entry.
template <typename T>
static double spmv_burst(fk_mesh *m, int reps) {
    hipStream_t st = m->ctx->stream;
    ensure_work<T>(m);
    if (solver_kind(m) == FK_SPMV_CSR && sizeof(T) == 4) make_val32(m);
    sync_solver_values<T>(m);
    PcgWork<T> &w = work<T>(m);
    FK_HIP(hipMemsetAsync(w.p.get(), 0, m->nv * sizeof(T), st));
    DevBuf<CgState> fresh(1);
    FK_HIP(hipMemsetAsync(fresh.get(), 0, sizeof(CgState), st));
    std::swap(w.state, fresh);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto restore = [&] {
        std::swap(w.state, fresh);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    try {
        FK_HIP(hipEventCreate(&e0));
        FK_HIP(hipEventCreate(&e1));
        launch_spmv_lanes<T>(m, w, 0);  // warm-up
        FK_HIP(hipEventRecord(e0, st));
        for (int k = 0; k < reps; ++k) launch_spmv_lanes<T>(m, w, 0);
        FK_HIP(hipEventRecord(e1, st));
        FK_HIP(hipGetLastError());
        FK_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        FK_HIP(hipEventElapsedTime(&ms, e0, e1));
        restore();
        return static_cast<double>(ms) * 1e3 / reps;
    } catch (...) {
        restore();
        throw;
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
        fk::NodesOut<double> out(m, phi);
        FK_HIP(hipMemcpyAsync(out.ptr, m->phi.get(), m->nv * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
        FK_HIP(hipStreamSynchronize(m->ctx->stream));
        out.commit();
    });
}

int fk_set_potential(fk_mesh *m, const double *phi) {
    if (!m || !phi) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_set_potential", [&] {
        fk::set_device(m->ctx);
        m->phi.ensure(m->nv);
        const fk::NodesIn<double> in(m, phi);
        FK_HIP(hipMemcpyAsync(m->phi.get(), in.ptr, m->nv * sizeof(double), hipMemcpyHostToDevice, m->ctx->stream));
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

int fk_field_noalloc(fk_mesh *m, const double *phi, double *H, double *Hn, double *energy) {
    if (!m) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_field_noalloc", [&] {
        if (!phi) fk::fail(FK_ERR_INVALID_ARG, "phi (device, nv values) is required");
        if (energy && !m->have_coef) fk::fail(FK_ERR_STATE, "energy needs a coefficient");
        const int nb = std::max(1, m->nblocks);
        if (nb > fk::kScratchPartials)
            fk::fail(FK_ERR_UNSUPPORTED, "the field kernel's " + std::to_string(nb) + " block partials exceed the " +
                                             std::to_string(fk::kScratchPartials) + " the static scratch holds");
        const int dev = m->ctx->device;
        double *part = dev >= 0 && dev < fk::kMaxDevices ? fk::g_scratch[dev] : nullptr;
        if (!part) fk::fail(FK_ERR_STATE, "no scratch address for this device (context set-up failed)");
        fk::set_device(m->ctx);
        hipStream_t st = m->ctx->stream;
        if (m->reordered) {
            // the potential into the internal node order (buffers allocated with the mesh)
            fk::k_gather_nodes<<<fk::grid_for(m->nv, 256), 256, 0, st>>>(phi, m->d_node_perm.get(), m->nv,
                                                                          m->phi_internal.get());
            phi = m->phi_internal.get();
        }
        // caller element order as fk_field, unless the energy needs a
        // coefficient conversion buffer that was never allocated (this entry
        // allocates nothing): then the colour-ordered copies
        const double *coef_nat = energy ? fk::coef_caller_order(m, false) : nullptr;
        if (!energy || coef_nat)
            fk::k_field<<<nb, fk::kBlock, 0, st>>>(m->tet_nat.get(), m->X.get(), m->Y.get(), m->Z.get(), nullptr,
                                                   phi, coef_nat, m->nt, H, Hn, part);
        else
            fk::k_field<<<nb, fk::kBlock, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(),
                                                   m->perm.get(), phi, m->coef.get(), m->nt, H, Hn, part);
        if (energy) fk::k_sum_in_order<<<1, 64, 0, st>>>(part, nb, energy);
        FK_HIP(hipGetLastError());
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

int fk_copy_probe(fk_context *ctx, int64_t bytes, int32_t reps, double *ms_per_copy) {
    if (!ctx || bytes < 16 || reps < 1 || !ms_per_copy) return FK_ERR_INVALID_ARG;
    return fk::guard(ctx, "fk_copy_probe", [&] {
        hipStream_t st = ctx->stream;
        const int64_t n = bytes / 16;
        fk::DevBuf<uint4> src(static_cast<size_t>(n)), dst(static_cast<size_t>(n));
        FK_HIP(hipMemsetAsync(src.get(), 0, n * 16, st));
        // one 16-byte element per thread over the whole buffer: 6.1 TB/s for
        // 2 GiB measured, against 4.5-4.6 TB/s for a grid-stride loop over
        // 16-64 blocks per CU
        const int64_t grid = std::min<int64_t>((n + 255) / 256, 0x7fffffff);
        hipEvent_t e0, e1;
        FK_HIP(hipEventCreate(&e0));
        FK_HIP(hipEventCreate(&e1));
        fk::k_copy16<<<static_cast<unsigned>(grid), 256, 0, st>>>(src.get(), dst.get(), n);  // warm-up
        FK_HIP(hipEventRecord(e0, st));
        for (int i = 0; i < reps; ++i)
            fk::k_copy16<<<static_cast<unsigned>(grid), 256, 0, st>>>(src.get(), dst.get(), n);
        FK_HIP(hipEventRecord(e1, st));
        FK_HIP(hipGetLastError());
        FK_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        FK_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms_per_copy = static_cast<double>(ms) / reps;
    });
}

int fk_spmv_burst(fk_mesh *m, int32_t precision, int32_t reps, double *us_per_launch) {
    if (!m || !us_per_launch || reps < 1 || (precision != FK_PREC_F64 && precision != FK_PREC_F32_REFINED))
        return FK_ERR_INVALID_ARG;
    return fk::guard(m->ctx, "fk_spmv_burst", [&] {
        fk::set_device(m->ctx);
        if (!m->have_values) fk::fail(FK_ERR_STATE, "matrix not assembled");
        *us_per_launch = precision == FK_PREC_F64 ? fk::spmv_burst<double>(m, reps) : fk::spmv_burst<float>(m, reps);
    });
}

int fk_set_spmv_format(fk_mesh *m, int32_t format) {
    if (!m || format < FK_SPMV_CSR || format > FK_SPMV_AUTO) return FK_ERR_INVALID_ARG;
    if (format != m->format) {
        // the SELL forms differ in layout (rebuilt on next use); the
        // resolved format is decided again
        m->have_sell = false;
        m->sell64_valid = false;
        m->sell32_valid = false;
        m->kind = -1;
    }
    m->format = format;
    return FK_OK;
}

int fk_solver_format(fk_mesh *m, int32_t *format, int32_t *stencil_offsets, int32_t *offsets) {
    if (!m || !format) return FK_ERR_INVALID_ARG;
    return fk::guard(m->ctx, "fk_solver_format", [&] {
        fk::set_device(m->ctx);
        *format = fk::solver_kind(m);
        if (stencil_offsets) *stencil_offsets = *format == FK_SPMV_STENCIL ? m->st_K : 0;
        if (offsets && *format == FK_SPMV_STENCIL)
            for (int k = 0; k < m->st_K; ++k) offsets[k] = m->st_off[k];
    });
}

int fk_sell_info(fk_mesh *m, int64_t *nslices, int64_t *shared_slices, int64_t *entries) {
    if (!m) return FK_ERR_INVALID_ARG;
    return fk::guard(m->ctx, "fk_sell_info", [&] {
        if (fk::solver_kind(m) != FK_SPMV_SELL64) fk::fail(FK_ERR_STATE, "the solver format is not SELL-64");
        fk::build_sell(m);
        if (nslices) *nslices = m->nslices;
        if (shared_slices) *shared_slices = m->nshared;
        if (entries) *entries = m->sell_len;
    });
}

int fk_sell_layout(fk_mesh *m, int32_t *index_bytes, int32_t *batch_groups, int64_t *f32_array_bytes) {
    if (!m) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_sell_layout", [&] {
        if (fk::solver_kind(m) != FK_SPMV_SELL64) fk::fail(FK_ERR_STATE, "the solver format is not SELL-64");
        fk::build_sell(m);
        if (index_bytes) *index_bytes = m->sell_short ? 2 : 4;
        if (batch_groups) *batch_groups = fk::sell_batch_groups(m);
        if (f32_array_bytes) {
            // what one FP32 product streams besides x and y: values (padding
            // included), per-row indices, shared offsets, the slice tables
            const int64_t ns = m->nslices;
            *f32_array_bytes = m->sell_len * static_cast<int64_t>(sizeof(float)) +
                               static_cast<int64_t>(m->sell_col.n) * 4 + static_cast<int64_t>(m->sell_dcol.n) * 2 +
                               static_cast<int64_t>(m->sell_soff.n) * 4 + (ns + 1) * 8 + ns * 4 + ns * 8;
        }
    });
}

int fk_set_sell_shape(fk_mesh *m, int32_t block, int32_t unroll) {
    if (!m) return FK_ERR_INVALID_ARG;
    const bool ok = ((block == 256 || block == 512) && (unroll == 1 || unroll == 2)) || (block == 1024 && unroll == 1);
    if (!ok) return FK_ERR_INVALID_ARG;
    m->sell_block = block;
    m->sell_unroll = unroll;
    return FK_OK;
}

int fk_set_lanes_per_row(fk_mesh *m, int32_t lanes) {
    if (!m || (lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)) return FK_ERR_INVALID_ARG;
    m->lanes = lanes;
    return FK_OK;
}

}  // extern "C"
