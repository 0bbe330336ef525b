// Device helpers shared by the femeko_hip kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace fk {

// Basis gradients of a P1 tetrahedron in closed form. With edges
// e_k = P_k - P_0, det = e1 . (e2 x e3); grad(lambda_1) = (e2 x e3)/det,
// grad(lambda_2) = (e3 x e1)/det, grad(lambda_3) = (e1 x e2)/det and
// grad(lambda_0) = -(sum of the others). These are the b, c, d that Femeko's
// abcd obtains from a 4x4 solve per node (src/FEM3D.jl:5-17); |det|/6 is
// elementVolume (src/mesh.jl:484-505).
struct Grad {
    double g[4][3];
    double det;
};

// The 12 coordinates of an element's nodes (x, y, z per node), loaded apart
// from the arithmetic so that callers can issue the loads of several
// elements before using any.
__device__ __forceinline__ void element_coords(const double *__restrict__ X, const double *__restrict__ Y,
                                               const double *__restrict__ Z, int4 nd, double c[12]) {
    c[0] = X[nd.x]; c[1] = Y[nd.x]; c[2] = Z[nd.x];
    c[3] = X[nd.y]; c[4] = Y[nd.y]; c[5] = Z[nd.y];
    c[6] = X[nd.z]; c[7] = Y[nd.z]; c[8] = Z[nd.z];
    c[9] = X[nd.w]; c[10] = Y[nd.w]; c[11] = Z[nd.w];
}

__device__ __forceinline__ Grad gradients_of(const double c[12]) {
    const double x0 = c[0], y0 = c[1], z0 = c[2];
    const double ax = c[3] - x0, ay = c[4] - y0, az = c[5] - z0;
    const double bx = c[6] - x0, by = c[7] - y0, bz = c[8] - z0;
    const double cx = c[9] - x0, cy = c[10] - y0, cz = c[11] - z0;
    // cofactor vectors
    const double c1x = by * cz - bz * cy, c1y = bz * cx - bx * cz, c1z = bx * cy - by * cx;  // b x c
    const double c2x = cy * az - cz * ay, c2y = cz * ax - cx * az, c2z = cx * ay - cy * ax;  // c x a
    const double c3x = ay * bz - az * by, c3y = az * bx - ax * bz, c3z = ax * by - ay * bx;  // a x b
    Grad G;
    G.det = ax * c1x + ay * c1y + az * c1z;
    const double inv = 1.0 / G.det;
    G.g[1][0] = c1x * inv; G.g[1][1] = c1y * inv; G.g[1][2] = c1z * inv;
    G.g[2][0] = c2x * inv; G.g[2][1] = c2y * inv; G.g[2][2] = c2z * inv;
    G.g[3][0] = c3x * inv; G.g[3][1] = c3y * inv; G.g[3][2] = c3z * inv;
    G.g[0][0] = -(G.g[1][0] + G.g[2][0] + G.g[3][0]);
    G.g[0][1] = -(G.g[1][1] + G.g[2][1] + G.g[3][1]);
    G.g[0][2] = -(G.g[1][2] + G.g[2][2] + G.g[3][2]);
    return G;
}

__device__ __forceinline__ Grad element_gradients(const double *__restrict__ X, const double *__restrict__ Y,
                                                  const double *__restrict__ Z, int4 nd) {
    double c[12];
    element_coords(X, Y, Z, nd, c);
    return gradients_of(c);
}

// 16 bytes of T for one vector load / store
template <typename T>
struct alignas(16) Vec16 {
    T v[16 / sizeof(T)];
};

__device__ __forceinline__ double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// ---- deterministic reductions (fixed operation order) ---------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Sum over the block; every thread gets the result. sh: BLOCK/64 + 1 doubles.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double *sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < BLOCK / 64; ++i) s += sh[i];
        sh[BLOCK / 64] = s;
    }
    __syncthreads();
    const double s = sh[BLOCK / 64];
    __syncthreads();
    return s;
}

template <int BLOCK>
__device__ __forceinline__ double block_max(double v, double *sh) {
    v = wave_max(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = sh[0];
#pragma unroll
        for (int i = 1; i < BLOCK / 64; ++i) s = fmax(s, sh[i]);
        sh[BLOCK / 64] = s;
    }
    __syncthreads();
    const double s = sh[BLOCK / 64];
    __syncthreads();
    return s;
}

// Sums of n block partials, in the same order in every block. One thread's
// share: four accumulators over a stride of BLOCK, combined as
// (s0 + s1) + (s2 + s3); the four independent loads in flight measured 2 us
// less per launch than one accumulator for 4080 partials.
template <int BLOCK>
__device__ __forceinline__ double thread_partials(const double *__restrict__ part, int n) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = threadIdx.x;
    for (; i + 3 * BLOCK < n; i += 4 * BLOCK) {
        s0 += part[i];
        s1 += part[i + BLOCK];
        s2 += part[i + 2 * BLOCK];
        s3 += part[i + 3 * BLOCK];
    }
    for (; i < n; i += BLOCK) s0 += part[i];
    return (s0 + s1) + (s2 + s3);
}

// The same for a long partial array (16-byte aligned): pairs of partials per
// load, four pairs in flight per thread.
template <int BLOCK>
__device__ __forceinline__ double thread_partials_wide(const double *__restrict__ part, int n) {
    const double2 *__restrict__ p2 = reinterpret_cast<const double2 *>(part);
    const int n2 = n >> 1;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = threadIdx.x;
    for (; i + 3 * BLOCK < n2; i += 4 * BLOCK) {
        const double2 a = p2[i], b = p2[i + BLOCK], c = p2[i + 2 * BLOCK], d = p2[i + 3 * BLOCK];
        s0 += a.x + a.y;
        s1 += b.x + b.y;
        s2 += c.x + c.y;
        s3 += d.x + d.y;
    }
    for (; i < n2; i += BLOCK) {
        const double2 a = p2[i];
        s0 += a.x + a.y;
    }
    if ((n & 1) && threadIdx.x == 0) s1 += part[n - 1];
    return (s0 + s1) + (s2 + s3);
}

// Partials of the previous kernel loaded before anything else a block
// loads: PER values (P = 1) or 16-byte pairs (P = 2) per thread, every load
// unconditional (index clamped into the array, the value masked in sum()),
// so that the caller, after a scheduling barrier, issues its vector loads
// and the sums wait for these loads only (loads retire in issue order)
// while the vector loads are still in flight. Partials beyond the PER
// loads of every thread (grids larger than PER * BLOCK * P blocks) are
// added by a trailing loop. The order of the sums is fixed, the same in
// every block.
template <int BLOCK, int PER, int P>
struct EarlyPartials {
    using E = typename std::conditional<P == 2, double2, double>::type;
    E v[PER];
    __device__ __forceinline__ void load(const double *__restrict__ part, int n) {
        const E *__restrict__ pe = reinterpret_cast<const E *>(part);
        const int last = max(n / P - 1, 0);
#pragma unroll
        for (int u = 0; u < PER; ++u) v[u] = pe[min(static_cast<int>(threadIdx.x) + u * BLOCK, last)];
    }
    __device__ __forceinline__ double sum(const double *__restrict__ part, int n) const {
        const int ne = n / P;  // whole elements
        double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const bool in = static_cast<int>(threadIdx.x) + u * BLOCK < ne;
            if constexpr (P == 2)
                s[u & 3] += in ? v[u].x + v[u].y : 0.0;
            else
                s[u & 3] += in ? v[u] : 0.0;
        }
        for (int i = static_cast<int>(threadIdx.x) + PER * BLOCK; i < ne; i += BLOCK) {
            if constexpr (P == 2)
                s[0] += part[2 * i] + part[2 * i + 1];
            else
                s[0] += part[i];
        }
        if (P == 2 && (n & 1) && threadIdx.x == 0) s[1] += part[n - 1];
        return (s[0] + s[1]) + (s[2] + s[3]);
    }
};

template <int BLOCK>
__device__ __forceinline__ double sum_partials(const double *__restrict__ part, int n, double *sh) {
    return block_sum<BLOCK>(thread_partials<BLOCK>(part, n), sh);
}

// block_sum of N values in one LDS round: each value is reduced in the very
// order block_sum uses (wave sums, then thread k adds the BLOCK/64 wave sums
// of value k from 0.0), so the results are bit-identical to N block_sum
// calls. sh holds N * (BLOCK/64 + 1) doubles.
template <int BLOCK, int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N], double *sh) {
    constexpr int W = BLOCK / 64;
    static_assert(N <= 64, "one thread per value");
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double s = wave_sum(v[k]);
        if (l == 0) sh[k * (W + 1) + w] = s;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        double *row = sh + threadIdx.x * (W + 1);
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < W; ++i) s += row[i];
        row[W] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = sh[k * (W + 1) + W];
    __syncthreads();
}

// Blocks are handed to the 8 XCDs round-robin; give each XCD a contiguous
// 1/8 of the logical chunks so its L2 holds one band of rows (and the x
// entries those rows gather).
__device__ __forceinline__ int logical_block(int nblocks) {
    const int b = blockIdx.x;
    if ((nblocks & 7) != 0) return b;
    return (b & 7) * (nblocks >> 3) + (b >> 3);
}

}  // namespace fk
