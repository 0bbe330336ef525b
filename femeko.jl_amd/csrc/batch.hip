// Batched solves: many independent meshes, up to 4 right-hand sides each, in
// one set of launches.
//
// A batch of meshes is one disconnected mesh (node numbers offset per mesh),
// so upload, colouring, pattern, slot map and assembly are those of a single
// fk_mesh: its stiffness matrix is block diagonal and one launch per colour
// assembles every mesh of the batch. Only the solver differs: every
// (mesh, right-hand side) pair is its own PCG system with its own scalars.
//   - The persistent grid's logical blocks are cut at mesh boundaries
//     (table blk -> mesh, row range), so each block's partial sums belong to
//     one mesh and a mesh's blocks are contiguous; per-system alpha/beta are
//     reductions over that range, in a fixed order.
//   - The right-hand sides of a mesh share the matrix: vectors are stored node
//     major with the nrhs values of a node together, one gather of the column
//     serves all of them (SpMM), matrix traffic is paid once per nrhs solves.
//   - FP32 PCG per system with FP64 reliable updates decided on the device
//     after every chunk of iterations (per system: fold the FP32 increment,
//     replace the residual by the FP64 true residual, keep the direction),
//     or, as the alternative, FP32 PCG to an inner tolerance restarted inside
//     an FP64 residual-correction loop; systems that have finished are
//     skipped, blocks whose systems are all done return at once; chunks of
//     iterations run from a captured hipGraph.
//   - Lanes per row and block size come from an autotuner that times the
//     SpMM kernel on the batch itself.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "device.h"
#include "internal.h"

struct fk_batch {
    fk_context *ctx = nullptr;
    fk_mesh *mesh = nullptr;  // the concatenated mesh (owned)
    int32_t nmesh = 0, nrhs = 1, nsys = 0;
    std::vector<int64_t> node_off, elem_off;
    // logical block table
    int nblocks = 0;
    fk::DevBuf<int32_t> blk_mesh;
    fk::DevBuf<int64_t> blk_r0, blk_r1;
    fk::DevBuf<int32_t> mesh_b0, mesh_b1;
    std::vector<int32_t> h_mesh_b0, h_mesh_b1;
    // block table of the stencil SpMM (finer: 1024 rows per block)
    int sp_nblocks = 0;
    int sp_block = 256;         // threads per block of the stencil SpMM
    int sp_rpt = 1;             // rows per thread and step of the stencil SpMM (autotuned)
    fk::DevBuf<int32_t> sp_blk_mesh, sp_mesh_b0, sp_mesh_b1;
    fk::DevBuf<int64_t> sp_blk_r0, sp_blk_r1;
    // Dirichlet values and potentials, node major [node][rhs]
    fk::DevBuf<double> g, phi, r64;
    bool have_dirichlet = false, have_phi = false;
    // FP32 inner work, node major
    fk::DevBuf<float> x, r, p, q;
    // partials [logical block][rhs]
    fk::DevBuf<double> part_pq, part_rz, part_rr, part_bb;
    // per-system state
    fk::DevBuf<int32_t> done, status, active, ndone;
    fk::DevBuf<int64_t> iters;
    fk::DevBuf<double> thresh, scale;
    fk::HostBuf<double> hpart;
    fk::HostBuf<int32_t> hcount;
    hipGraphExec_t graph = nullptr;
    int64_t graph_cfg = -1;
    int lanes = 8, block = 512;  // block: the update kernels (and the CSR SpMM)
    // reliable updates, per system
    int precision = FK_PREC_F32_REFINED;
    fk::DevBuf<int32_t> ru_act, ru_mode, ru_fin, ru_nrep, ru_restarts, ru_status, nfin;
    fk::DevBuf<double> rr32, rr64, rr_last, thresh64, part_res;
    fk::DevBuf<float> alpha;    // per system: step length of the current iteration, for the p-update kernel
    hipGraphExec_t ru_graph = nullptr;
    int64_t ru_graph_cfg = -1;
    double ru_delta = -1.0;     // baked into the captured chunk
    int64_t ru_max_iter = -1;
    // per-mesh stencil copy (every mesh's pattern a stencil of <= 16 offsets)
    int stencil = -1;           // -1 not decided, 0 no (CSR SpMM), 1 yes
    int format = FK_SPMV_AUTO;  // FK_SPMV_CSR, FK_SPMV_STENCIL or FK_SPMV_AUTO
    fk::DevBuf<int64_t> node_off_dev;
    bool st_valid = false;      // FP32 stencil values match the assembly
    int64_t st_len = 0;
    int st_Kc = 0;              // common offset count of all meshes, 0 if they differ
    fk::DevBuf<int32_t> st_off, st_K, st_map;
    fk::DevBuf<int64_t> st_vbase;
    fk::DevBuf<float> st_val32;
    fk::DevBuf<double> st_val64, st_diag64;  // FP64 copy and diagonal for the true residuals
    ~fk_batch() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (ru_graph) (void)hipGraphExecDestroy(ru_graph);
        delete mesh;
    }
};

namespace fk {

struct BatchArgs {
    const int32_t *blk_mesh;
    const int64_t *blk_r0, *blk_r1;
    const int32_t *mesh_b0, *mesh_b1;
    int nblocks, nrhs;
};

__device__ __forceinline__ int batch_logical_block(int nblocks) { return logical_block(nblocks); }

__device__ __forceinline__ bool mesh_all_done(const int32_t *__restrict__ done, int mesh, int nrhs) {
    bool all = true;
    for (int k = 0; k < nrhs; ++k) all &= done[mesh * nrhs + k] != 0;
    return all;
}

// sum over the logical blocks [b0, b1) of part[b*nrhs + k], fixed order:
// one thread's share (two accumulators over a stride of BLOCK), then the
// block sum
template <int BLOCK>
__device__ __forceinline__ double seg_thread(const double *__restrict__ part, int b0, int b1, int k, int nrhs) {
    double s0 = 0.0, s1 = 0.0;
    int b = b0 + static_cast<int>(threadIdx.x);
    for (; b + BLOCK < b1; b += 2 * BLOCK) {
        s0 += part[b * nrhs + k];
        s1 += part[(b + BLOCK) * nrhs + k];
    }
    for (; b < b1; b += BLOCK) s0 += part[b * nrhs + k];
    return s0 + s1;
}
// seg_thread with the thread's first partial loaded ahead (seg_first) by
// the caller before its vector loads: the same sums in the same order, so
// the waits for the partials do not include the vector loads (loads retire
// in issue order). seg_first's index is clamped into the segment (b1 > b0).
template <int BLOCK>
__device__ __forceinline__ double seg_first(const double *__restrict__ part, int b0, int b1, int k, int nrhs) {
    const int b = b0 + static_cast<int>(threadIdx.x);
    return part[(b < b1 ? b : b1 - 1) * nrhs + k];
}
template <int BLOCK>
__device__ __forceinline__ double seg_thread_first(double first, const double *__restrict__ part, int b0, int b1,
                                                   int k, int nrhs) {
    double s0 = 0.0, s1 = 0.0;
    int b = b0 + static_cast<int>(threadIdx.x);
    if (b + BLOCK < b1) {
        // a segment longer than the block: seg_thread's loops, the head
        // element taken from `first`
        s0 += first;
        s1 += part[(b + BLOCK) * nrhs + k];
        for (b += 2 * BLOCK; b + BLOCK < b1; b += 2 * BLOCK) {
            s0 += part[b * nrhs + k];
            s1 += part[(b + BLOCK) * nrhs + k];
        }
        for (; b < b1; b += BLOCK) s0 += part[b * nrhs + k];
    } else if (b < b1) {
        s0 += first;
    }
    return s0 + s1;
}

// Clamped row of a prefetch load (the value is used only for i < r1): an
// unconditional load, no branch between the loads and the partial sums.
__device__ __forceinline__ int64_t clamp_row(int64_t i, int64_t r1) {
    const int64_t c = i < r1 ? i : r1 - 1;
    return c > 0 ? c : 0;
}

template <int BLOCK>
__device__ __forceinline__ double seg_sum(const double *__restrict__ part, int b0, int b1, int k, int nrhs,
                                          double *sh) {
    return block_sum<BLOCK>(seg_thread<BLOCK>(part, b0, b1, k, nrhs), sh);
}

// K1: q = A p for all right-hand sides of the block's mesh, partials of p.q.
// Difference form q_i = sum_j a_ij (p_j - p_i): the rows of every mesh's P1
// stiffness matrix sum to zero, and in FP32 this form does not cancel for
// smooth p (the diagonal term is zero and costs nothing but its load).
template <int BLOCK, int LPR>
__global__ __launch_bounds__(BLOCK) void kb_spmm(BatchArgs ba, const int32_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ col, const float *__restrict__ val,
                                                 const float *__restrict__ p, float *__restrict__ q,
                                                 double *__restrict__ part, const int32_t *__restrict__ done) {
    __shared__ double sh[BLOCK / 64 + 1];
    const int lb = batch_logical_block(ba.nblocks);
    const int mesh = ba.blk_mesh[lb];
    const int nrhs = ba.nrhs;
    if (mesh_all_done(done, mesh, nrhs)) return;
    const int64_t r0 = ba.blk_r0[lb], r1 = ba.blk_r1[lb];
    const int sub = threadIdx.x / LPR, lane = threadIdx.x % LPR;
    constexpr int RPB = BLOCK / LPR;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t rb = r0; rb < r1; rb += RPB) {
        const int64_t row = rb + sub;
        const bool valid = row < r1;
        float s[4] = {0.f, 0.f, 0.f, 0.f};
        if (valid) {
            float pi[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (j < nrhs) pi[j] = p[row * nrhs + j];
            const int32_t a = rowptr[row], b = rowptr[row + 1];
            for (int32_t k = a + lane; k < b; k += LPR) {
                const float v = val[k];
                const float *pc = p + static_cast<int64_t>(col[k]) * nrhs;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (j < nrhs) s[j] += v * (pc[j] - pi[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1) s[j] += __shfl_xor(s[j], o, 64);
        }
        if (valid && lane == 0) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (j < nrhs) {
                    q[row * nrhs + j] = s[j];
                    acc[j] += static_cast<double>(p[row * nrhs + j]) * static_cast<double>(s[j]);
                }
        }
    }
    for (int j = 0; j < nrhs; ++j) {
        const double t = block_sum<BLOCK>(acc[j], sh);
        if (threadIdx.x == 0) part[lb * nrhs + j] = t;
    }
}

// K1 from the per-mesh stencil copy: mesh m has K[m] positive offsets
// off[16 m + k] and value columns of its n_m rows from vbase[m]; column k
// holds a(i, i + o) (zero where i + o is past the mesh) and a(i, i - o) is
// read at row i - o (the matrix is exactly symmetric). Difference form; NR
// right-hand sides per row. KC > 0: every mesh has KC offsets (unrolled);
// KC = 0: the mesh's own count. Loads are unconditional: out-of-mesh
// neighbours are clamped to a row of the mesh and weighted zero. RPT rows per
// thread and step (rows BLOCK apart; those past the block's range are
// clamped to the first and not stored), so their loads are issued together:
// more loads in flight per wave against more registers (fewer waves).
template <int BLOCK, int NR, int KC, int RPT>
__global__ __launch_bounds__(BLOCK) void kb_spmm_stencil(BatchArgs ba, const int64_t *__restrict__ node_off,
                                                         const int32_t *__restrict__ Ks,
                                                         const int32_t *__restrict__ offs,
                                                         const int64_t *__restrict__ vbase,
                                                         const float *__restrict__ val, const float *__restrict__ p,
                                                         float *__restrict__ q, double *__restrict__ part,
                                                         const int32_t *__restrict__ done) {
    __shared__ double sh[NR * (BLOCK / 64 + 1)];
    const int lb = batch_logical_block(ba.nblocks);
    const int mesh = ba.blk_mesh[lb];
    if (mesh_all_done(done, mesh, NR)) return;
    const int64_t r0 = ba.blk_r0[lb], r1 = ba.blk_r1[lb];
    const int64_t a = node_off[mesh], e = node_off[mesh + 1], n = e - a;
    const int K = KC > 0 ? KC : Ks[mesh];
    const float *__restrict__ v0 = val + vbase[mesh];
    int32_t off[KC > 0 ? KC : 1];
#pragma unroll
    for (int k = 0; k < (KC > 0 ? KC : 1); ++k) off[k] = offs[16 * mesh + k];
    double acc[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) acc[j] = 0.0;
    // the mesh's rows through block-uniform base pointers and 32-bit row
    // indices (a mesh has fewer than 2^32 / NR nodes)
    const float *__restrict__ pm = p + a * NR;
    float *__restrict__ qm = q + a * NR;
    const uint32_t nn = static_cast<uint32_t>(n);
    const uint32_t l1 = static_cast<uint32_t>(r1 - a);
    for (uint32_t l0 = static_cast<uint32_t>(r0 - a) + threadIdx.x; l0 < l1; l0 += RPT * BLOCK) {
        bool own[RPT];
        uint32_t li[RPT];
        float xi[RPT][NR], s[RPT][NR];
#pragma unroll
        for (int h = 0; h < RPT; ++h) {
            const uint32_t lh = l0 + static_cast<uint32_t>(h) * BLOCK;
            own[h] = h == 0 || lh < l1;
            li[h] = own[h] ? lh : l0;
            const float *__restrict__ ph = pm + static_cast<size_t>(li[h]) * NR;
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                xi[h][j] = ph[j];
                s[h][j] = 0.f;
            }
        }
        auto term = [&](int k, uint32_t o) {
            const float *__restrict__ v = v0 + static_cast<uint64_t>(k) * nn;
#pragma unroll
            for (int h = 0; h < RPT; ++h) {
                const float wu = v[li[h]];
                const uint32_t lu = min(li[h] + o, nn - 1);
                const bool lo_ok = li[h] >= o;
                const uint32_t ll = lo_ok ? li[h] - o : 0u;
                const float wl = lo_ok ? v[ll] : 0.f;
                // a row's NR values at one 64-bit address: one vector load
                const float *__restrict__ pu = pm + static_cast<size_t>(lu) * NR;
                const float *__restrict__ pl = pm + static_cast<size_t>(ll) * NR;
#pragma unroll
                for (int j = 0; j < NR; ++j) s[h][j] += wu * (pu[j] - xi[h][j]) + wl * (pl[j] - xi[h][j]);
            }
        };
        if constexpr (KC > 0) {
#pragma unroll
            for (int k = 0; k < KC; ++k) term(k, static_cast<uint32_t>(off[k]));
        } else {
            for (int k = 0; k < K; ++k) term(k, static_cast<uint32_t>(offs[16 * mesh + k]));
        }
#pragma unroll
        for (int h = 0; h < RPT; ++h) {
            if (!own[h]) continue;
            float *__restrict__ qh = qm + static_cast<size_t>(li[h]) * NR;
            // q = A p is read once more, by kb_update_xr right after:
            // non-temporal on both sides keeps it from evicting the stencil
            // values and p from the MALL (64-mesh step 67.0-67.8 ms against
            // 71.4 ms with plain accesses)
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                __builtin_nontemporal_store(s[h][j], qh + j);
                acc[j] += static_cast<double>(xi[h][j]) * static_cast<double>(s[h][j]);
            }
        }
    }
    block_sum_n<BLOCK, NR>(acc, sh);  // the NR sums in one LDS round
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NR; ++j) part[lb * NR + j] = acc[j];
    }
}

__global__ void kb_stencil_fill(const double *__restrict__ val64, const int32_t *__restrict__ map, int64_t len,
                                float *__restrict__ out) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= len) return;
    const int32_t k = map[t];
    out[t] = k >= 0 ? static_cast<float>(val64[k]) : 0.f;
}

// K2: per system alpha, r -= alpha q, partials of r.z, r.r; the mesh's first
// block leaves alpha in alphas[] for K3, which applies x += alpha p while it
// reads p anyway. NR right-hand sides per node (contiguous). The block's
// first row is loaded before the per-mesh reductions so the loads overlap
// them.
template <int BLOCK, int NR>
__global__ __launch_bounds__(BLOCK) void kb_update_xr(BatchArgs ba, float *__restrict__ r,
                                                      const float *__restrict__ q, const float *__restrict__ dinv,
                                                      const double *__restrict__ part_pq,
                                                      const int32_t *__restrict__ pq_b0,
                                                      const int32_t *__restrict__ pq_b1,
                                                      const double *__restrict__ rz_cur, double *__restrict__ rz_next,
                                                      double *__restrict__ part_rr, int32_t *__restrict__ done,
                                                      int32_t *__restrict__ status, int32_t *__restrict__ ndone,
                                                      float *__restrict__ alphas) {
    __shared__ double sh[2 * NR * (BLOCK / 64 + 1)];
    const int lb = batch_logical_block(ba.nblocks);
    const int mesh = ba.blk_mesh[lb];
    if (mesh_all_done(done, mesh, NR)) {
        // K3 of this iteration then leaves the mesh alone
        if (lb == ba.mesh_b0[mesh] && threadIdx.x < NR) alphas[mesh * NR + threadIdx.x] = 0.f;
        return;
    }
    const int64_t r0 = ba.blk_r0[lb], r1 = ba.blk_r1[lb];
    const int64_t i0 = r0 + threadIdx.x;
    const int b0 = ba.mesh_b0[mesh], b1 = ba.mesh_b1[mesh];
    const int c0 = pq_b0[mesh], c1 = pq_b1[mesh];
    // the first partial of each sum, then the row's values (see seg_first)
    double f_rz[NR], f_pq[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        f_rz[j] = seg_first<BLOCK>(rz_cur, b0, b1, j, NR);
        f_pq[j] = seg_first<BLOCK>(part_pq, c0, c1, j, NR);
    }
    __builtin_amdgcn_sched_barrier(0);
    float d0, r0v[NR], q0[NR];
    {
        const int64_t ic = clamp_row(i0, r1);
        d0 = dinv[ic];
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            r0v[j] = r[ic * NR + j];
            q0[j] = __builtin_nontemporal_load(q + ic * NR + j);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    // the 2 NR per-mesh reductions in one pass (same order per value)
    double red[2 * NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        red[j] = seg_thread_first<BLOCK>(f_rz[j], rz_cur, b0, b1, j, NR);
        red[NR + j] = seg_thread_first<BLOCK>(f_pq[j], part_pq, c0, c1, j, NR);
    }
    block_sum_n<BLOCK, 2 * NR>(red, sh);
    float alpha[NR];
    bool act[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        const double rho = red[j];
        const double pq = red[NR + j];
        const bool live = !done[mesh * NR + j];
        act[j] = live && pq > 0.0 && isfinite(pq);
        alpha[j] = act[j] ? static_cast<float>(rho / pq) : 0.f;
        if (lb == b0 && threadIdx.x == 0) alphas[mesh * NR + j] = alpha[j];
        if (live && !act[j] && lb == b0 && threadIdx.x == 0) {
            // breakdown (every block of the mesh sees the same pq): stop the
            // system, its caller restarts or reports it
            status[mesh * NR + j] = FK_ERR_SINGULAR;
            done[mesh * NR + j] = 1;
            atomicAdd(ndone, 1);
        }
    }
    double rz[NR], rr[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) rz[j] = rr[j] = 0.0;
    // all NR values of a node are stored together (one vector store; a
    // system K2 does not step stores its value unchanged)
    auto row = [&](int64_t i, float d, float *rv, const float *qv) {
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            float ri = rv[j];
            if (act[j]) ri = d != 0.f ? ri - alpha[j] * qv[j] : 0.f;
            rv[j] = ri;
            const float z = d * ri;
            rz[j] += static_cast<double>(ri) * static_cast<double>(z);
            rr[j] += static_cast<double>(ri) * static_cast<double>(ri);
        }
#pragma unroll
        for (int j = 0; j < NR; ++j) r[i * NR + j] = rv[j];
    };
    if (i0 < r1) row(i0, d0, r0v, q0);
    for (int64_t i = i0 + BLOCK; i < r1; i += BLOCK) {
        float rv[NR], qv[NR];
        const float d = dinv[i];
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            rv[j] = r[i * NR + j];
            qv[j] = __builtin_nontemporal_load(q + i * NR + j);
        }
        row(i, d, rv, qv);
    }
    double tot[2 * NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        tot[j] = rz[j];
        tot[NR + j] = rr[j];
    }
    block_sum_n<BLOCK, 2 * NR>(tot, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            rz_next[lb * NR + j] = tot[j];
            part_rr[lb * NR + j] = tot[NR + j];
        }
    }
}

// x of the batch's p update is read and written once per iteration and by
// no other kernel of it: non-temporal, so it does not take MALL room from
// the stencil values, p, q and r that the next kernels read again (the
// iteration's 326 MB exceed the 256 MB MALL). 64-mesh step 71.4 ms against
// 73.4-73.5 ms with plain accesses.
__device__ __forceinline__ float ldx(const float *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void stx(float *p, float v) { __builtin_nontemporal_store(v, p); }

// K3: per system convergence test (first block of the mesh), x += alpha p,
// p = D^-1 r + beta p for the systems K2 stepped (alpha != 0). The mesh's
// first block marks converged systems done during this launch, so the other
// blocks decide from alpha, which no block of this launch writes, and all
// rows of a system finish the step together.
template <int BLOCK, int NR>
__global__ __launch_bounds__(BLOCK) void kb_update_p(BatchArgs ba, float *__restrict__ x, float *__restrict__ p,
                                                     const float *__restrict__ r,
                                                     const float *__restrict__ alphas,
                                                     const float *__restrict__ dinv,
                                                     const double *__restrict__ rz_next,
                                                     const double *__restrict__ rz_cur,
                                                     const double *__restrict__ part_rr, int32_t *__restrict__ done,
                                                     int32_t *__restrict__ status, int64_t *__restrict__ iters,
                                                     const double *__restrict__ thresh, int64_t max_iter,
                                                     int32_t *__restrict__ ndone, double *__restrict__ rr32) {
    __shared__ double sh[3 * NR * (BLOCK / 64 + 1)];
    const int lb = batch_logical_block(ba.nblocks);
    const int mesh = ba.blk_mesh[lb];
    float alpha[NR];
    bool upd[NR];
    bool any = false;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        alpha[j] = alphas[mesh * NR + j];
        upd[j] = alpha[j] != 0.f;
        any |= upd[j];
    }
    if (!any) return;
    const int64_t r0 = ba.blk_r0[lb], r1 = ba.blk_r1[lb];
    const int64_t i0 = r0 + threadIdx.x;
    const int b0 = ba.mesh_b0[mesh], b1 = ba.mesh_b1[mesh];
    // the first partial of each sum, then the row's values (see seg_first)
    double f_zn[NR], f_zc[NR], f_rr[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        f_zn[j] = seg_first<BLOCK>(rz_next, b0, b1, j, NR);
        f_zc[j] = seg_first<BLOCK>(rz_cur, b0, b1, j, NR);
        f_rr[j] = seg_first<BLOCK>(part_rr, b0, b1, j, NR);
    }
    __builtin_amdgcn_sched_barrier(0);
    float d0, rv0[NR], pv0[NR], xv0[NR];
    {
        const int64_t ic = clamp_row(i0, r1);
        d0 = dinv[ic];
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            rv0[j] = r[ic * NR + j];
            pv0[j] = p[ic * NR + j];
            xv0[j] = ldx(x + ic * NR + j);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    // the per-mesh reductions in one pass (r.r only in the mesh's first
    // block, which keeps the convergence state)
    double red[3 * NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        red[j] = seg_thread_first<BLOCK>(f_zn[j], rz_next, b0, b1, j, NR);
        red[NR + j] = seg_thread_first<BLOCK>(f_zc[j], rz_cur, b0, b1, j, NR);
        red[2 * NR + j] = lb == b0 ? seg_thread_first<BLOCK>(f_rr[j], part_rr, b0, b1, j, NR) : 0.0;
    }
    block_sum_n<BLOCK, 3 * NR>(red, sh);
    float beta[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        const int s = mesh * NR + j;
        const double rho_new = red[j];
        const double rho = red[NR + j];
        beta[j] = upd[j] ? static_cast<float>(rho_new / rho) : 0.f;
        if (lb == b0) {
            const double rr = red[2 * NR + j];
            if (threadIdx.x == 0 && upd[j]) {
                const int64_t it = iters[s] + 1;
                iters[s] = it;
                rr32[s] = rr;
                int fin = 0;
                if (rr < thresh[s]) {
                    fin = 1;
                } else if (!isfinite(rr)) {
                    status[s] = FK_ERR_NONFINITE;
                    fin = 1;
                } else if (it >= max_iter) {
                    status[s] = FK_ERR_NOT_CONVERGED;
                    fin = 1;
                }
                if (fin) {
                    done[s] = 1;
                    atomicAdd(ndone, 1);
                }
            }
        }
    }
    // all NR values of a node are loaded and stored together (vector
    // accesses; a system K2 did not step keeps its values)
    auto row = [&](int64_t i, float d, const float *rv, float *pv, float *xv) {
#pragma unroll
        for (int j = 0; j < NR; ++j)
            if (upd[j]) {
                xv[j] += alpha[j] * pv[j];
                pv[j] = d * rv[j] + beta[j] * pv[j];
            }
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            stx(x + i * NR + j, xv[j]);
            p[i * NR + j] = pv[j];
        }
    };
    if (i0 < r1) row(i0, d0, rv0, pv0, xv0);
    for (int64_t i = i0 + BLOCK; i < r1; i += BLOCK) {
        const float d = dinv[i];
        float rv[NR], pv[NR], xv[NR];
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            rv[j] = r[i * NR + j];
            pv[j] = p[i * NR + j];
            xv[j] = ldx(x + i * NR + j);
        }
        row(i, d, rv, pv, xv);
    }
}

// r64 = -mask(A64 phi) for every system, partials of |r|^2 per [block][rhs]
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void kb_neg_residual(BatchArgs ba, const int32_t *__restrict__ rowptr,
                                                         const int32_t *__restrict__ col,
                                                         const double *__restrict__ val,
                                                         const double *__restrict__ phi,
                                                         const uint8_t *__restrict__ mask, double *__restrict__ r,
                                                         double *__restrict__ part, const int32_t *__restrict__ gate) {
    __shared__ double sh[BLOCK / 64 + 1];
    constexpr int LPR = 16, RPB = BLOCK / LPR;
    const int lb = batch_logical_block(ba.nblocks);
    const int nrhs = ba.nrhs;
    if (gate) {
        // replacement step: only meshes with a system that replaces (act 1)
        const int mesh = ba.blk_mesh[lb];
        bool any = false;
        for (int j = 0; j < nrhs; ++j) any |= gate[mesh * nrhs + j] == 1;
        if (!any) return;
    }
    const int64_t r0 = ba.blk_r0[lb], r1 = ba.blk_r1[lb];
    const int sub = threadIdx.x / LPR, lane = threadIdx.x % LPR;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t rb = r0; rb < r1; rb += RPB) {
        const int64_t row = rb + sub;
        const bool valid = row < r1;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        if (valid) {
            const int32_t a = rowptr[row], b = rowptr[row + 1];
            for (int32_t k = a + lane; k < b; k += LPR) {
                const double v = val[k];
                const double *pc = phi + static_cast<int64_t>(col[k]) * nrhs;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (j < nrhs) s[j] += v * pc[j];
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1) s[j] += __shfl_xor(s[j], o, 64);
        }
        if (valid && lane == 0) {
            const bool fixed = mask[row] != 0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (j < nrhs) {
                    const double v = fixed ? 0.0 : -s[j];
                    r[row * nrhs + j] = v;
                    acc[j] += v * v;
                }
        }
    }
    for (int j = 0; j < nrhs; ++j) {
        const double t = block_sum<BLOCK>(acc[j], sh);
        if (threadIdx.x == 0) part[lb * nrhs + j] = t;
    }
}

// r64 = -mask(A64 phi) from the per-mesh FP64 stencil copy (standard form
// with the diagonal), partials of |r|^2 per [block][rhs] on the batch's main
// block table, gate as kb_neg_residual. Loads guarded by the mesh's rows.
template <int BLOCK, int NR, int KC>
__global__ __launch_bounds__(BLOCK) void kb_neg_residual_stencil(
    BatchArgs ba, const int64_t *__restrict__ node_off, const int32_t *__restrict__ Ks,
    const int32_t *__restrict__ offs, const int64_t *__restrict__ vbase, const double *__restrict__ val,
    const double *__restrict__ diag, const double *__restrict__ phi, const uint8_t *__restrict__ mask,
    double *__restrict__ r, double *__restrict__ part, const int32_t *__restrict__ gate) {
    __shared__ double sh[NR * (BLOCK / 64 + 1)];
    const int lb = batch_logical_block(ba.nblocks);
    const int mesh = ba.blk_mesh[lb];
    if (gate) {
        bool any = false;
#pragma unroll
        for (int j = 0; j < NR; ++j) any |= gate[mesh * NR + j] == 1;
        if (!any) return;
    }
    const int64_t r0 = ba.blk_r0[lb], r1 = ba.blk_r1[lb];
    const int64_t a = node_off[mesh], n = node_off[mesh + 1] - a;
    const int K = KC > 0 ? KC : Ks[mesh];
    const double *__restrict__ v0 = val + vbase[mesh];
    double acc[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) acc[j] = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += BLOCK) {
        const int64_t li = i - a;
        double y[NR];
        const double d = diag[i];
#pragma unroll
        for (int j = 0; j < NR; ++j) y[j] = d * phi[i * NR + j];
        if constexpr (KC > 0) {
            // unconditional loads (indices clamped into the mesh): column k
            // holds 0 where li + o is past the mesh and a backward term out
            // of range is weighted 0, so each y gains exactly the guarded
            // form's terms in the same order
#pragma unroll
            for (int k = 0; k < KC; ++k) {
                const int64_t o = offs[16 * mesh + k];
                const double *__restrict__ v = v0 + static_cast<int64_t>(k) * n;
                const int64_t lu = min(li + o, n - 1), ll = max(li - o, int64_t(0));
                const double wu = v[li];
                const double wl = li >= o ? v[ll] : 0.0;
#pragma unroll
                for (int j = 0; j < NR; ++j) y[j] += wu * phi[(a + lu) * NR + j];
#pragma unroll
                for (int j = 0; j < NR; ++j) y[j] += wl * phi[(a + ll) * NR + j];
            }
        } else {
            for (int k = 0; k < K; ++k) {
                const int64_t o = offs[16 * mesh + k];
                const double *__restrict__ v = v0 + static_cast<int64_t>(k) * n;
                if (li + o < n) {
                    const double w = v[li];
#pragma unroll
                    for (int j = 0; j < NR; ++j) y[j] += w * phi[(i + o) * NR + j];
                }
                if (li >= o) {
                    const double w = v[li - o];
#pragma unroll
                    for (int j = 0; j < NR; ++j) y[j] += w * phi[(i - o) * NR + j];
                }
            }
        }
        const bool fixed = mask[i] != 0;
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            const double rv = fixed ? 0.0 : -y[j];
            r[i * NR + j] = rv;
            acc[j] += rv * rv;
        }
    }
    block_sum_n<BLOCK, NR>(acc, sh);  // the NR sums in one LDS round
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NR; ++j) part[lb * NR + j] = acc[j];
    }
}

__global__ void kb_stencil_fill64(const double *__restrict__ val64, const int32_t *__restrict__ map,
                                  const int32_t *__restrict__ diag_pos, int64_t len, int64_t nv,
                                  double *__restrict__ out, double *__restrict__ diag) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < len) {
        const int32_t k = map[t];
        out[t] = k >= 0 ? val64[k] : 0.0;
    }
    if (t < nv) {
        const int32_t k = diag_pos[t];
        diag[t] = k >= 0 ? val64[k] : 0.0;
    }
}

// inner init: r32 = r64 / scale (active systems), x = 0, p = D^-1 r; partials
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void kb_init(BatchArgs ba, const double *__restrict__ r64,
                                                 const double *__restrict__ scale, const int32_t *__restrict__ active,
                                                 const float *__restrict__ dinv, float *__restrict__ x,
                                                 float *__restrict__ r, float *__restrict__ p,
                                                 double *__restrict__ part_rz, double *__restrict__ part_bb) {
    __shared__ double sh[BLOCK / 64 + 1];
    const int lb = batch_logical_block(ba.nblocks);
    const int mesh = ba.blk_mesh[lb];
    const int nrhs = ba.nrhs;
    const int64_t r0 = ba.blk_r0[lb], r1 = ba.blk_r1[lb];
    double rz[4] = {0.0, 0.0, 0.0, 0.0}, bb[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = r0 + threadIdx.x; i < r1; i += BLOCK) {
        const float d = dinv[i];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= nrhs) continue;
            const int s = mesh * nrhs + j;
            const int64_t o = i * nrhs + j;
            const float ri = (active[s] && d != 0.f) ? static_cast<float>(r64[o] / scale[s]) : 0.f;
            r[o] = ri;
            x[o] = 0.f;
            const float z = d * ri;
            p[o] = z;
            rz[j] += static_cast<double>(ri) * static_cast<double>(z);
            bb[j] += static_cast<double>(ri) * static_cast<double>(ri);
        }
    }
    for (int j = 0; j < nrhs; ++j) {
        const double a = block_sum<BLOCK>(rz[j], sh);
        const double b = block_sum<BLOCK>(bb[j], sh);
        if (threadIdx.x == 0) {
            part_rz[lb * nrhs + j] = a;
            part_bb[lb * nrhs + j] = b;
        }
    }
}

// one thread per system: thresholds and initial done flags
__global__ void kb_start(BatchArgs ba, const double *__restrict__ part_bb, const int32_t *__restrict__ active,
                         double rtol, int32_t *__restrict__ done, int32_t *__restrict__ status,
                         int64_t *__restrict__ iters, double *__restrict__ thresh, int32_t *__restrict__ ndone,
                         int nsys) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    const int mesh = s / ba.nrhs, j = s % ba.nrhs;
    double bb = 0.0;
    for (int b = ba.mesh_b0[mesh]; b < ba.mesh_b1[mesh]; ++b) bb += part_bb[b * ba.nrhs + j];
    thresh[s] = rtol * rtol * bb;
    iters[s] = 0;
    status[s] = FK_OK;
    const int d = (!active[s] || bb == 0.0) ? 1 : 0;
    done[s] = d;
    if (d) atomicAdd(ndone, 1);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void kb_phi_update(BatchArgs ba, double *__restrict__ phi,
                                                       const float *__restrict__ x, const double *__restrict__ scale,
                                                       const int32_t *__restrict__ active) {
    const int lb = batch_logical_block(ba.nblocks);
    const int mesh = ba.blk_mesh[lb];
    const int nrhs = ba.nrhs;
    const int64_t r0 = ba.blk_r0[lb], r1 = ba.blk_r1[lb];
    for (int64_t i = r0 + threadIdx.x; i < r1; i += BLOCK)
        for (int j = 0; j < nrhs; ++j) {
            const int s = mesh * nrhs + j;
            if (active[s]) phi[i * nrhs + j] += scale[s] * static_cast<double>(x[i * nrhs + j]);
        }
}

__global__ void kb_init_phi(double *__restrict__ phi, const double *__restrict__ g, const uint8_t *__restrict__ mask,
                            int64_t nv, int nrhs) {
    const int64_t o = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (o >= nv * nrhs) return;
    phi[o] = mask[o / nrhs] ? g[o] : 0.0;
}

// ---- reliable updates per system (after every chunk of iterations) ------
constexpr int kBatchMaxRestarts = 32;

struct BatchRu {
    int32_t *act, *mode, *fin, *nrep, *restarts, *fstatus, *nfin;
    double *rr32, *rr64, *rr_last, *thresh64;
    const double *scale;
    int32_t *done, *status;
    const int64_t *iters;
    double delta2;
    int64_t max_iter;
    int nsys;
};

// one thread per system: replace when the FP32 residual fell by delta since
// the last replacement, or the recursion stopped (target, limit, breakdown)
__global__ void kb_ru_check(BatchRu u) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= u.nsys) return;
    u.act[s] = 0;
    if (u.fin[s]) return;
    if (!u.done[s]) {
        if (u.rr32[s] < u.delta2 * u.rr_last[s] || u.iters[s] >= u.max_iter) u.act[s] = 1;
        return;
    }
    const bool broken = u.status[s] == FK_ERR_SINGULAR || u.status[s] == FK_ERR_NONFINITE || !isfinite(u.rr32[s]);
    if (!broken) {
        u.act[s] = 1;
        return;
    }
    if (++u.restarts[s] > kBatchMaxRestarts || u.iters[s] >= u.max_iter) {
        u.fin[s] = 1;
        u.fstatus[s] = u.iters[s] >= u.max_iter ? FK_ERR_NOT_CONVERGED : u.status[s];
        atomicAdd(u.nfin, 1);
        return;
    }
    u.act[s] = 2;  // drop the increment, restart from the last true residual
}

// phi += scale x (act 1), x = 0 (act 1 and 2); a node's NR values are
// loaded and stored together (a system that does not fold keeps its values)
template <int BLOCK, int NR>
__global__ __launch_bounds__(BLOCK) void kb_ru_fold(BatchArgs ba, BatchRu u, double *__restrict__ phi,
                                                    float *__restrict__ x) {
    const int lb = batch_logical_block(ba.nblocks);
    const int mesh = ba.blk_mesh[lb];
    int a[NR];
    double sc[NR];
    bool any = false;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        a[j] = u.act[mesh * NR + j];
        sc[j] = u.scale[mesh * NR + j];
        any |= a[j] != 0;
    }
    if (!any) return;
    const int64_t r0 = ba.blk_r0[lb], r1 = ba.blk_r1[lb];
    for (int64_t i = r0 + threadIdx.x; i < r1; i += BLOCK) {
        double ph[NR];
        float xv[NR];
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            ph[j] = phi[i * NR + j];
            xv[j] = x[i * NR + j];
        }
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            if (a[j] == 1) ph[j] += sc[j] * static_cast<double>(xv[j]);
            if (a[j]) xv[j] = 0.f;
        }
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            phi[i * NR + j] = ph[j];
            x[i * NR + j] = xv[j];
        }
    }
}

// one thread per system: the FP64 residual decides (block sums in order)
__global__ void kb_ru_decide(BatchArgs ba, BatchRu u, const double *__restrict__ part) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= u.nsys || u.act[s] == 0) return;
    if (u.act[s] == 2) {
        u.mode[s] = 2;
        return;
    }
    const int mesh = s / ba.nrhs, j = s % ba.nrhs;
    double rr = 0.0;
    for (int b = ba.mesh_b0[mesh]; b < ba.mesh_b1[mesh]; ++b) rr += part[b * ba.nrhs + j];
    u.rr64[s] = rr;
    u.nrep[s] += 1;
    int fin = 0, st = FK_OK;
    if (rr < u.thresh64[s]) {
        fin = 1;
    } else if (!isfinite(rr)) {
        fin = 1;
        st = FK_ERR_NONFINITE;
    } else if (u.iters[s] >= u.max_iter) {
        fin = 1;
        st = FK_ERR_NOT_CONVERGED;
    } else {
        const double sc = u.scale[s];
        const double scaled = rr / (sc * sc);
        int mode = 1;
        if (scaled > 4.0 * u.rr32[s]) {  // drifted: restart the direction
            if (++u.restarts[s] > kBatchMaxRestarts) {
                fin = 1;
                st = FK_ERR_NOT_CONVERGED;
            }
            mode = 2;
        }
        u.mode[s] = mode;
        u.rr_last[s] = scaled;
    }
    if (fin) {
        u.fin[s] = 1;
        u.fstatus[s] = st;
        u.done[s] = 1;  // the FP32 kernels skip it from now on
        atomicAdd(u.nfin, 1);
    }
}

// r32 = r64 / scale, p = D^-1 r on a restart, rho partials into the parity-0
// slot, for the replaced systems only. When every system of the block's mesh
// is replaced (and restarts), a node's values are stored together.
template <int BLOCK, int NR>
__global__ __launch_bounds__(BLOCK) void kb_ru_replace(BatchArgs ba, BatchRu u, const double *__restrict__ r64,
                                                       const float *__restrict__ dinv, float *__restrict__ r,
                                                       float *__restrict__ p, double *__restrict__ part_rz) {
    __shared__ double sh[NR * (BLOCK / 64 + 1)];
    const int lb = batch_logical_block(ba.nblocks);
    const int mesh = ba.blk_mesh[lb];
    bool rep[NR], restart[NR];
    bool any = false, all_rep = true, all_restart = true;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        const int s = mesh * NR + j;
        rep[j] = u.act[s] != 0 && !u.fin[s];
        restart[j] = rep[j] && u.mode[s] == 2;
        any |= rep[j];
        all_rep &= rep[j];
        all_restart &= restart[j];
    }
    if (!any) return;
    double dinvs[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) dinvs[j] = 1.0 / u.scale[mesh * NR + j];
    const int64_t r0 = ba.blk_r0[lb], r1 = ba.blk_r1[lb];
    double rz[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) rz[j] = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += BLOCK) {
        const float d = dinv[i];
        double rd[NR];
#pragma unroll
        for (int j = 0; j < NR; ++j) rd[j] = r64[i * NR + j];
        float rv[NR];
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            rv[j] = d != 0.f ? static_cast<float>(rd[j] * dinvs[j]) : 0.f;
            if (rep[j]) rz[j] += static_cast<double>(rv[j]) * static_cast<double>(d * rv[j]);
        }
        if (all_rep) {
#pragma unroll
            for (int j = 0; j < NR; ++j) r[i * NR + j] = rv[j];
            if (all_restart) {
#pragma unroll
                for (int j = 0; j < NR; ++j) p[i * NR + j] = d * rv[j];
            } else {
#pragma unroll
                for (int j = 0; j < NR; ++j)
                    if (restart[j]) p[i * NR + j] = d * rv[j];
            }
        } else {
#pragma unroll
            for (int j = 0; j < NR; ++j)
                if (rep[j]) {
                    r[i * NR + j] = rv[j];
                    if (restart[j]) p[i * NR + j] = d * rv[j];
                }
        }
    }
    block_sum_n<BLOCK, NR>(rz, sh);  // the NR sums in one LDS round
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NR; ++j)
            if (rep[j]) part_rz[lb * NR + j] = rz[j];
    }
}

__global__ void kb_ru_resume(BatchRu u) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= u.nsys) return;
    if (u.act[s] != 0 && !u.fin[s]) {
        u.done[s] = 0;
        u.status[s] = FK_OK;
    }
    u.act[s] = 0;
}

// initial state: FP32 stop threshold = the scaled FP64 target
__global__ void kb_ru_start(BatchRu u, const double *__restrict__ target2, double *__restrict__ thresh,
                            int64_t *__restrict__ iters) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= u.nsys) return;
    thresh[s] = target2[s];
    iters[s] = 0;
    u.status[s] = FK_OK;
    u.done[s] = u.fin[s];
    u.act[s] = 0;
    u.mode[s] = 1;
    u.nrep[s] = 0;
    u.restarts[s] = 0;
    u.rr_last[s] = 1.0;
    u.rr32[s] = 1.0;
}

__global__ void kb_zero_i32(int32_t *__restrict__ a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

// ---------------------------------------------------------------------------
static BatchArgs args_of(fk_batch *b) {
    return BatchArgs{b->blk_mesh.get(), b->blk_r0.get(), b->blk_r1.get(), b->mesh_b0.get(), b->mesh_b1.get(),
                     b->nblocks, b->nrhs};
}

// Logical blocks cut at mesh boundaries, at most `chunk` rows each; with
// pad8 the count is padded to a multiple of 8 with empty blocks of the last
// mesh, so that logical_block gives each XCD a contiguous eighth.
struct TableHost {
    std::vector<int32_t> bm, b0, b1;
    std::vector<int64_t> r0, r1;
};

static TableHost make_table(const fk_batch *b, int64_t chunk, bool pad8) {
    TableHost t;
    t.b0.assign(b->nmesh, 0);
    t.b1.assign(b->nmesh, 0);
    for (int m = 0; m < b->nmesh; ++m) {
        const int64_t a = b->node_off[m], e = b->node_off[m + 1];
        const int64_t nb = std::max<int64_t>(1, (e - a + chunk - 1) / chunk);
        const int64_t c = (e - a + nb - 1) / nb;
        t.b0[m] = static_cast<int32_t>(t.bm.size());
        for (int64_t k = 0; k < nb; ++k) {
            t.bm.push_back(m);
            t.r0.push_back(std::min(e, a + k * c));
            t.r1.push_back(std::min(e, a + (k + 1) * c));
        }
        t.b1[m] = static_cast<int32_t>(t.bm.size());
    }
    if (pad8) {
        const int last = b->nmesh - 1;
        const int64_t e = b->node_off[last + 1];
        while (t.bm.size() % 8 != 0) {
            t.bm.push_back(last);
            t.r0.push_back(e);
            t.r1.push_back(e);
        }
        t.b1[last] = static_cast<int32_t>(t.bm.size());
    }
    return t;
}

static void upload_table(const TableHost &t, int nmesh, fk::DevBuf<int32_t> &bm, fk::DevBuf<int64_t> &r0,
                         fk::DevBuf<int64_t> &r1, fk::DevBuf<int32_t> &b0, fk::DevBuf<int32_t> &b1) {
    bm.alloc(t.bm.size());
    r0.alloc(t.bm.size());
    r1.alloc(t.bm.size());
    b0.alloc(nmesh);
    b1.alloc(nmesh);
    FK_HIP(hipMemcpy(bm.get(), t.bm.data(), t.bm.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    FK_HIP(hipMemcpy(r0.get(), t.r0.data(), t.r0.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    FK_HIP(hipMemcpy(r1.get(), t.r1.data(), t.r1.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    FK_HIP(hipMemcpy(b0.get(), t.b0.data(), nmesh * sizeof(int32_t), hipMemcpyHostToDevice));
    FK_HIP(hipMemcpy(b1.get(), t.b1.data(), nmesh * sizeof(int32_t), hipMemcpyHostToDevice));
}

// the table of every kernel but the stencil SpMM: ~4 blocks per CU overall.
// Not padded to a multiple of 8, so logical_block keeps dispatch order: with
// XCD bands (padded) the 64-mesh step measured 76.8 ms against 72.3 ms, and
// with every kernel on the SpMM's banded table of 1024-row blocks 89.9 ms.
static void build_block_table(fk_batch *b) {
    const int64_t total = b->node_off.back();
    const int64_t target_blocks = std::max<int64_t>(b->nmesh, 4 * static_cast<int64_t>(b->ctx->num_cus));
    const int64_t chunk = std::max<int64_t>(256, (total + target_blocks - 1) / target_blocks);
    TableHost t = make_table(b, chunk, false);
    b->nblocks = static_cast<int>(t.bm.size());
    b->h_mesh_b0 = t.b0;
    b->h_mesh_b1 = t.b1;
    upload_table(t, b->nmesh, b->blk_mesh, b->blk_r0, b->blk_r1, b->mesh_b0, b->mesh_b1);
}

// Stencil SpMM table: 1024 rows per block. Measured on the single-mesh
// stencil kernel (4.17M rows): blocks of about two 512-row runs, swept in
// dispatch order inside each XCD's band, take 26-27 us against 36 us for
// blocks of 8 runs, whose rows one grid plane back have left L2 more often.
constexpr int64_t kSpmmRows = 1024;
static void build_spmm_table(fk_batch *b) {
    TableHost t = make_table(b, kSpmmRows, true);
    b->sp_nblocks = static_cast<int>(t.bm.size());
    upload_table(t, b->nmesh, b->sp_blk_mesh, b->sp_blk_r0, b->sp_blk_r1, b->sp_mesh_b0, b->sp_mesh_b1);
    const size_t need = static_cast<size_t>(std::max(b->sp_nblocks, b->nblocks)) * b->nrhs;
    b->part_pq.ensure(need);
}

static BatchArgs spmm_args_of(fk_batch *b) {
    return BatchArgs{b->sp_blk_mesh.get(), b->sp_blk_r0.get(), b->sp_blk_r1.get(), b->sp_mesh_b0.get(),
                     b->sp_mesh_b1.get(), b->sp_nblocks, b->nrhs};
}

template <int BLOCK, int LPR>
static void launch_spmm(fk_batch *b, int parity_unused) {
    (void)parity_unused;
    fk_mesh *m = b->mesh;
    kb_spmm<BLOCK, LPR><<<b->nblocks, BLOCK, 0, b->ctx->stream>>>(args_of(b), m->rowptr.get(), m->colidx.get(),
                                                                 m->val32.get(), b->p.get(), b->q.get(),
                                                                 b->part_pq.get(), b->done.get());
}

template <int BLOCK>
static void launch_spmm_lanes(fk_batch *b, int lanes) {
    switch (lanes) {
        case 4: launch_spmm<BLOCK, 4>(b, 0); break;
        case 16: launch_spmm<BLOCK, 16>(b, 0); break;
        default: launch_spmm<BLOCK, 8>(b, 0); break;
    }
}

template <int BLOCK, int NR, int KC>
static void launch_spmm_stencil_k(fk_batch *b) {
    if (b->sp_rpt == 1)
        kb_spmm_stencil<BLOCK, NR, KC, 1><<<b->sp_nblocks, BLOCK, 0, b->ctx->stream>>>(
            spmm_args_of(b), b->node_off_dev.get(), b->st_K.get(), b->st_off.get(), b->st_vbase.get(),
            b->st_val32.get(), b->p.get(), b->q.get(), b->part_pq.get(), b->done.get());
    else
        kb_spmm_stencil<BLOCK, NR, KC, 2><<<b->sp_nblocks, BLOCK, 0, b->ctx->stream>>>(
            spmm_args_of(b), b->node_off_dev.get(), b->st_K.get(), b->st_off.get(), b->st_vbase.get(),
            b->st_val32.get(), b->p.get(), b->q.get(), b->part_pq.get(), b->done.get());
}

// 7 offsets: the Kuhn pattern of structured boxes; other counts run the
// mesh's own count
template <int BLOCK, int NR>
static void launch_spmm_stencil_nr(fk_batch *b) {
    if (b->st_Kc == 7)
        launch_spmm_stencil_k<BLOCK, NR, 7>(b);
    else
        launch_spmm_stencil_k<BLOCK, NR, 0>(b);
}

template <int BLOCK>
static void launch_spmm_stencil(fk_batch *b) {
    switch (b->nrhs) {
        case 1: launch_spmm_stencil_nr<BLOCK, 1>(b); break;
        case 2: launch_spmm_stencil_nr<BLOCK, 2>(b); break;
        case 3: launch_spmm_stencil_nr<BLOCK, 3>(b); break;
        default: launch_spmm_stencil_nr<BLOCK, 4>(b); break;
    }
    FK_HIP(hipGetLastError());
}

// block: the CSR SpMM's threads per block, or with the stencil copy its own
static void spmm(fk_batch *b, int block, int lanes) {
    if (b->stencil == 1) {
        switch (block) {
            case 128: launch_spmm_stencil<128>(b); break;
            case 512: launch_spmm_stencil<512>(b); break;
            default: launch_spmm_stencil<256>(b); break;
        }
        return;
    }
    switch (block) {
        case 128: launch_spmm_lanes<128>(b, lanes); break;
        case 512: launch_spmm_lanes<512>(b, lanes); break;
        default: launch_spmm_lanes<256>(b, lanes); break;
    }
}

template <int BLOCK, int NR>
static void iteration_tail(fk_batch *b, int parity, int64_t max_iter) {
    hipStream_t st = b->ctx->stream;
    const int nb = b->nblocks, nr = b->nrhs;
    double *rz_cur = b->part_rz.get() + (parity ? nb * nr : 0);
    double *rz_next = b->part_rz.get() + (parity ? 0 : nb * nr);
    fk_mesh *m = b->mesh;
    const float *dinv = m->w32.dinv.get();  // set up by the solve
    if (!dinv) fail(FK_ERR_STATE, "batch iteration without a Jacobi diagonal");
    kb_update_xr<BLOCK, NR><<<nb, BLOCK, 0, st>>>(args_of(b), b->r.get(), b->q.get(),
                                              dinv, b->part_pq.get(),
                                              b->stencil == 1 ? b->sp_mesh_b0.get() : b->mesh_b0.get(),
                                              b->stencil == 1 ? b->sp_mesh_b1.get() : b->mesh_b1.get(), rz_cur, rz_next,
                                              b->part_rr.get(), b->done.get(), b->status.get(), b->ndone.get(),
                                              b->alpha.get());
    kb_update_p<BLOCK, NR><<<nb, BLOCK, 0, st>>>(args_of(b), b->x.get(), b->p.get(), b->r.get(), b->alpha.get(),
                                             dinv, rz_next, rz_cur,
                                             b->part_rr.get(), b->done.get(), b->status.get(), b->iters.get(),
                                             b->thresh.get(), max_iter, b->ndone.get(), b->rr32.get());
}

template <int BLOCK>
static void iteration_tail_nr(fk_batch *b, int parity, int64_t max_iter) {
    switch (b->nrhs) {
        case 1: iteration_tail<BLOCK, 1>(b, parity, max_iter); break;
        case 2: iteration_tail<BLOCK, 2>(b, parity, max_iter); break;
        case 3: iteration_tail<BLOCK, 3>(b, parity, max_iter); break;
        default: iteration_tail<BLOCK, 4>(b, parity, max_iter); break;
    }
}

// the launch shape a captured iteration chunk bakes in
static int64_t launch_config(const fk_batch *b) {
    return b->block + 1000 * static_cast<int64_t>(b->lanes) + 100000 * static_cast<int64_t>(b->sp_block) +
           100000000 * static_cast<int64_t>(b->sp_rpt);
}

static void enqueue_batch_iteration(fk_batch *b, int parity, int64_t max_iter) {
    spmm(b, b->stencil == 1 ? b->sp_block : b->block, b->lanes);
    switch (b->block) {
        case 128: iteration_tail_nr<128>(b, parity, max_iter); break;
        case 512: iteration_tail_nr<512>(b, parity, max_iter); break;
        default: iteration_tail_nr<256>(b, parity, max_iter); break;
    }
}

template <int NR, int KC>
static void launch_neg_residual_stencil(fk_batch *b, const double *phi, double *part, const int32_t *gate) {
    kb_neg_residual_stencil<256, NR, KC><<<b->nblocks, 256, 0, b->ctx->stream>>>(
        args_of(b), b->node_off_dev.get(), b->st_K.get(), b->st_off.get(), b->st_vbase.get(), b->st_val64.get(),
        b->st_diag64.get(), phi, b->mesh->mask.get(), b->r64.get(), part, gate);
}

template <int NR>
static void launch_neg_residual_stencil_nr(fk_batch *b, const double *phi, double *part, const int32_t *gate) {
    if (b->st_Kc == 7)
        launch_neg_residual_stencil<NR, 7>(b, phi, part, gate);
    else
        launch_neg_residual_stencil<NR, 0>(b, phi, part, gate);
}

// r64 = -mask(A64 phi): the FP64 stencil copy when the batch has one (filled
// with the FP32 copy after each assembly), else the CSR values
static void neg_residual_batch(fk_batch *b, const double *phi, double *part = nullptr, const int32_t *gate = nullptr) {
    fk_mesh *m = b->mesh;
    double *pt = part ? part : b->part_rr.get();
    if (b->stencil == 1 && b->st_valid) {
        switch (b->nrhs) {
            case 1: launch_neg_residual_stencil_nr<1>(b, phi, pt, gate); break;
            case 2: launch_neg_residual_stencil_nr<2>(b, phi, pt, gate); break;
            case 3: launch_neg_residual_stencil_nr<3>(b, phi, pt, gate); break;
            default: launch_neg_residual_stencil_nr<4>(b, phi, pt, gate); break;
        }
    } else {
        kb_neg_residual<256><<<b->nblocks, 256, 0, b->ctx->stream>>>(args_of(b), m->rowptr.get(), m->colidx.get(),
                                                                    m->val64.get(), phi, m->mask.get(),
                                                                    b->r64.get(), pt, gate);
    }
    FK_HIP(hipGetLastError());
}

// per-system sums of part_rr (host)
static void system_norms(fk_batch *b, std::vector<double> &out) {
    hipStream_t st = b->ctx->stream;
    const size_t n = static_cast<size_t>(b->nblocks) * b->nrhs;
    FK_HIP(hipMemcpyAsync(b->hpart.ptr, b->part_rr.get(), n * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    out.assign(b->nsys, 0.0);
    for (int m = 0; m < b->nmesh; ++m)
        for (int j = 0; j < b->nrhs; ++j) {
            double s = 0.0;
            for (int k = b->h_mesh_b0[m]; k < b->h_mesh_b1[m]; ++k) s += b->hpart.ptr[k * b->nrhs + j];
            out[m * b->nrhs + j] = s;
        }
}

static void inner_solve(fk_batch *b, double inner_rtol, int64_t max_iter) {
    hipStream_t st = b->ctx->stream;
    fk_mesh *m = b->mesh;
    const int nb = b->nblocks;
    kb_zero_i32<<<1, 64, 0, st>>>(b->ndone.get(), 1);
    kb_init<256><<<nb, 256, 0, st>>>(args_of(b), b->r64.get(), b->scale.get(), b->active.get(), m->w32.dinv.get(),
                                     b->x.get(), b->r.get(), b->p.get(), b->part_rz.get(), b->part_bb.get());
    kb_start<<<grid_for(b->nsys, 64), 64, 0, st>>>(args_of(b), b->part_bb.get(), b->active.get(), inner_rtol,
                                                   b->done.get(), b->status.get(), b->iters.get(), b->thresh.get(),
                                                   b->ndone.get(), b->nsys);
    FK_HIP(hipGetLastError());
    const int G = 16;
    const int64_t cfg = launch_config(b);
    if (!b->graph || b->graph_cfg != cfg) {
        if (b->graph) (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
        hipGraph_t graph;
        FK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        // max_iter lives in the kernel arguments: capture with a large cap and
        // enforce the caller's limit from the host between chunks
        for (int k = 0; k < G; ++k) enqueue_batch_iteration(b, k & 1, int64_t(1) << 40);
        FK_HIP(hipStreamEndCapture(st, &graph));
        hipError_t e = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        FK_HIP(e);
        b->graph_cfg = cfg;
    }
    int64_t launched = 0;
    while (true) {
        FK_HIP(hipGraphLaunch(b->graph, st));
        launched += G;
        FK_HIP(hipMemcpyAsync(b->hcount.ptr, b->ndone.get(), sizeof(int32_t), hipMemcpyDeviceToHost, st));
        FK_HIP(hipStreamSynchronize(st));
        if (b->hcount.ptr[0] >= b->nsys || launched >= max_iter) break;
    }
}

static BatchRu ru_of(fk_batch *b, double delta, int64_t max_iter) {
    return BatchRu{b->ru_act.get(),   b->ru_mode.get(),  b->ru_fin.get(),      b->ru_nrep.get(), b->ru_restarts.get(),
                   b->ru_status.get(), b->nfin.get(),     b->rr32.get(),        b->rr64.get(),    b->rr_last.get(),
                   b->thresh64.get(),  b->scale.get(),    b->done.get(),        b->status.get(),  b->iters.get(),
                   delta * delta,      max_iter,          b->nsys};
}

// G iterations and the per-system replacement step
static void enqueue_ru_chunk(fk_batch *b, const BatchRu &u, int G) {
    hipStream_t st = b->ctx->stream;
    fk_mesh *m = b->mesh;
    for (int k = 0; k < G; ++k) enqueue_batch_iteration(b, k & 1, int64_t(1) << 40);
    const unsigned gs = static_cast<unsigned>(grid_for(b->nsys, 64));
    kb_ru_check<<<gs, 64, 0, st>>>(u);
    switch (b->nrhs) {
        case 1: kb_ru_fold<256, 1><<<b->nblocks, 256, 0, st>>>(args_of(b), u, b->phi.get(), b->x.get()); break;
        case 2: kb_ru_fold<256, 2><<<b->nblocks, 256, 0, st>>>(args_of(b), u, b->phi.get(), b->x.get()); break;
        case 3: kb_ru_fold<256, 3><<<b->nblocks, 256, 0, st>>>(args_of(b), u, b->phi.get(), b->x.get()); break;
        default: kb_ru_fold<256, 4><<<b->nblocks, 256, 0, st>>>(args_of(b), u, b->phi.get(), b->x.get()); break;
    }
    neg_residual_batch(b, b->phi.get(), b->part_res.get(), b->ru_act.get());
    kb_ru_decide<<<gs, 64, 0, st>>>(args_of(b), u, b->part_res.get());
    switch (b->nrhs) {
#define FK_RU_REPLACE(NR)                                                                                          \
    kb_ru_replace<256, NR><<<b->nblocks, 256, 0, st>>>(args_of(b), u, b->r64.get(), m->w32.dinv.get(), b->r.get(), \
                                                       b->p.get(), b->part_rz.get())
        case 1: FK_RU_REPLACE(1); break;
        case 2: FK_RU_REPLACE(2); break;
        case 3: FK_RU_REPLACE(3); break;
        default: FK_RU_REPLACE(4); break;
#undef FK_RU_REPLACE
    }
    kb_ru_resume<<<gs, 64, 0, st>>>(u);
    FK_HIP(hipGetLastError());
}

// FP32 PCG per system with FP64 reliable updates decided on the device
static void batch_solve_ru(fk_batch *b, double rtol, double delta, int64_t max_iter, int64_t *iters_out,
                           int32_t *outer_out, double *relres_out) {
    fk_mesh *m = b->mesh;
    hipStream_t st = b->ctx->stream;
    const int64_t nv = m->nv;
    const int nr = b->nrhs, S = b->nsys;
    b->phi.ensure(nv * nr);
    b->r64.ensure(nv * nr);
    kb_init_phi<<<grid_for(nv * nr, 256), 256, 0, st>>>(b->phi.get(), b->g.get(), m->mask.get(), nv, nr);
    b->have_phi = true;
    std::vector<double> bb, rr;
    neg_residual_batch(b, b->g.get());  // b = -A[free,fixed] g per system
    system_norms(b, bb);
    neg_residual_batch(b, b->phi.get());  // r64 = residual of the start
    system_norms(b, rr);
    std::vector<int32_t> act(S), fin(S);
    std::vector<double> scale(S), th(S), t2(S);
    int nfin0 = 0;
    for (int s = 0; s < S; ++s) {
        th[s] = rtol * rtol * bb[s];
        act[s] = (bb[s] > 0.0 && !(rr[s] < th[s])) ? 1 : 0;
        fin[s] = !act[s];
        nfin0 += fin[s];
        scale[s] = act[s] ? std::sqrt(rr[s]) : 1.0;
        t2[s] = act[s] ? th[s] / rr[s] : 0.0;
    }
    DevBuf<double> target2(S);
    std::vector<int32_t> zeros(S, 0);
    FK_HIP(hipMemcpyAsync(b->active.get(), act.data(), S * sizeof(int32_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(b->ru_fin.get(), fin.data(), S * sizeof(int32_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(b->ru_status.get(), zeros.data(), S * sizeof(int32_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(b->nfin.get(), &nfin0, sizeof(int32_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(b->scale.get(), scale.data(), S * sizeof(double), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(b->thresh64.get(), th.data(), S * sizeof(double), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(b->rr64.get(), rr.data(), S * sizeof(double), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(target2.get(), t2.data(), S * sizeof(double), hipMemcpyHostToDevice, st));
    const BatchRu u = ru_of(b, delta, max_iter);
    kb_init<256><<<b->nblocks, 256, 0, st>>>(args_of(b), b->r64.get(), b->scale.get(), b->active.get(),
                                             m->w32.dinv.get(), b->x.get(), b->r.get(), b->p.get(), b->part_rz.get(),
                                             b->part_bb.get());
    kb_ru_start<<<grid_for(S, 64), 64, 0, st>>>(u, target2.get(), b->thresh.get(), b->iters.get());
    FK_HIP(hipGetLastError());
    const int G = 16;
    const int64_t cfg = launch_config(b);
    const double delta_key = delta;
    if (!b->ru_graph || b->ru_graph_cfg != cfg || b->ru_delta != delta_key || b->ru_max_iter != max_iter) {
        if (b->ru_graph) (void)hipGraphExecDestroy(b->ru_graph);
        b->ru_graph = nullptr;
        hipGraph_t graph;
        FK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        enqueue_ru_chunk(b, u, G);
        FK_HIP(hipStreamEndCapture(st, &graph));
        hipError_t e = hipGraphInstantiate(&b->ru_graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        FK_HIP(e);
        b->ru_graph_cfg = cfg;
        b->ru_delta = delta_key;
        b->ru_max_iter = max_iter;
    }
    // one chunk in flight ahead of the count being examined
    int32_t *hc = b->hcount.ptr;
    hipEvent_t ev[2];
    FK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    FK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    try {
        FK_HIP(hipGraphLaunch(b->ru_graph, st));
        FK_HIP(hipMemcpyAsync(&hc[0], b->nfin.get(), sizeof(int32_t), hipMemcpyDeviceToHost, st));
        FK_HIP(hipEventRecord(ev[0], st));
        int cur = 0;
        while (true) {
            FK_HIP(hipGraphLaunch(b->ru_graph, st));
            FK_HIP(hipMemcpyAsync(&hc[1 - cur], b->nfin.get(), sizeof(int32_t), hipMemcpyDeviceToHost, st));
            FK_HIP(hipEventRecord(ev[1 - cur], st));
            FK_HIP(hipEventSynchronize(ev[cur]));
            if (hc[cur] >= S) break;
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
    std::vector<int64_t> its(S);
    std::vector<int32_t> nrep(S), fst(S);
    std::vector<double> rr64(S);
    FK_HIP(hipMemcpyAsync(its.data(), b->iters.get(), S * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipMemcpyAsync(nrep.data(), b->ru_nrep.get(), S * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipMemcpyAsync(fst.data(), b->ru_status.get(), S * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipMemcpyAsync(rr64.data(), b->rr64.get(), S * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    int status = FK_OK;
    for (int s = 0; s < S; ++s) {
        if (iters_out) iters_out[s] = act[s] ? its[s] : 0;
        if (outer_out) outer_out[s] = act[s] ? nrep[s] : 0;
        if (relres_out) relres_out[s] = bb[s] > 0.0 ? std::sqrt(rr64[s] / bb[s]) : 0.0;
        if (fst[s] != FK_OK && status == FK_OK) status = fst[s];
    }
    if (status != FK_OK) fail(status, std::string("batch solve: ") + fk_status_string(status));
}

static void batch_solve_restart(fk_batch *b, double rtol, double inner_rtol, int max_outer, int64_t max_iter,
                                int64_t *iters_out, int32_t *outer_out, double *relres_out) {
    fk_mesh *m = b->mesh;
    hipStream_t st = b->ctx->stream;
    const int64_t nv = m->nv;
    const int nr = b->nrhs, S = b->nsys;
    b->phi.ensure(nv * nr);
    b->r64.ensure(nv * nr);
    kb_init_phi<<<grid_for(nv * nr, 256), 256, 0, st>>>(b->phi.get(), b->g.get(), m->mask.get(), nv, nr);
    b->have_phi = true;
    std::vector<double> bb, rr;
    neg_residual_batch(b, b->g.get());  // b = -A[free,fixed] g per system
    system_norms(b, bb);
    neg_residual_batch(b, b->phi.get());
    system_norms(b, rr);
    std::vector<int64_t> total(S, 0), its(S);
    std::vector<int32_t> outer(S, 0), act(S), stat(S, FK_OK);
    std::vector<double> scale(S, 1.0);
    for (int step = 0; step < max_outer; ++step) {
        int nact = 0;
        for (int s = 0; s < S; ++s) {
            act[s] = (bb[s] > 0.0 && !(rr[s] < rtol * rtol * bb[s]) && stat[s] == FK_OK) ? 1 : 0;
            scale[s] = act[s] ? std::sqrt(rr[s]) : 1.0;
            nact += act[s];
        }
        if (nact == 0) break;
        FK_HIP(hipMemcpyAsync(b->active.get(), act.data(), S * sizeof(int32_t), hipMemcpyHostToDevice, st));
        FK_HIP(hipMemcpyAsync(b->scale.get(), scale.data(), S * sizeof(double), hipMemcpyHostToDevice, st));
        inner_solve(b, inner_rtol, max_iter);
        kb_phi_update<256><<<b->nblocks, 256, 0, st>>>(args_of(b), b->phi.get(), b->x.get(), b->scale.get(),
                                                       b->active.get());
        FK_HIP(hipGetLastError());
        std::vector<int32_t> st_dev(S);
        FK_HIP(hipMemcpyAsync(its.data(), b->iters.get(), S * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        FK_HIP(hipMemcpyAsync(st_dev.data(), b->status.get(), S * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        neg_residual_batch(b, b->phi.get());
        system_norms(b, rr);
        for (int s = 0; s < S; ++s)
            if (act[s]) {
                total[s] += its[s];
                outer[s] += 1;
                if (st_dev[s] == FK_ERR_NONFINITE) stat[s] = FK_ERR_NONFINITE;
            }
    }
    int status = FK_OK;
    for (int s = 0; s < S; ++s) {
        const double rel = bb[s] > 0.0 ? std::sqrt(rr[s] / bb[s]) : 0.0;
        if (iters_out) iters_out[s] = total[s];
        if (outer_out) outer_out[s] = outer[s];
        if (relres_out) relres_out[s] = rel;
        if (stat[s] != FK_OK) status = stat[s];
        else if (bb[s] > 0.0 && !(rr[s] < rtol * rtol * bb[s]) && status == FK_OK) status = FK_ERR_NOT_CONVERGED;
    }
    if (status != FK_OK) fail(status, std::string("batch solve: ") + fk_status_string(status));
}

// Per-mesh stencil copy, decided once per batch: every mesh's positive column
// offsets form one set of at most 16 with at most 25% padding.
static void build_batch_stencil(fk_batch *b) {
    if (b->stencil >= 0) return;
    b->stencil = 0;
    if (b->format == FK_SPMV_CSR) return;
    const bool must = b->format == FK_SPMV_STENCIL;
    auto reject = [&](const char *why) {
        if (must) fail(FK_ERR_INVALID_ARG, std::string("batch stencil format: ") + why);
    };
    fk_mesh *m = b->mesh;
    hipStream_t st = b->ctx->stream;
    const int64_t nv = m->nv, nnz = m->nnz;
    std::vector<int32_t> rp(static_cast<size_t>(nv + 1)), ci(static_cast<size_t>(nnz));
    FK_HIP(hipMemcpyAsync(rp.data(), m->rowptr.get(), (nv + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipMemcpyAsync(ci.data(), m->colidx.get(), nnz * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    const int M = b->nmesh;
    std::vector<int32_t> K(M), offs(16 * static_cast<size_t>(M), 0);
    std::vector<int64_t> vbase(M + 1, 0);
    for (int q = 0; q < M; ++q) {
        const int64_t a = b->node_off[q], e = b->node_off[q + 1];
        std::vector<int32_t> P;
        for (int64_t i = a; i < e; ++i)
            for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
                const int32_t d = ci[k] - static_cast<int32_t>(i);
                if (d <= 0) continue;
                auto it = std::lower_bound(P.begin(), P.end(), d);
                if (it == P.end() || *it != d) {
                    if (P.size() >= 16) return reject("more than 16 column offsets");
                    P.insert(it, d);
                }
            }
        const int64_t n = e - a;
        const int64_t pairs = (rp[e] - rp[a] - n) / 2;
        if (P.empty() || 4 * static_cast<int64_t>(P.size()) * n > 5 * pairs + 256 * static_cast<int64_t>(P.size()))
            return reject("more than 25% padding");
        K[q] = static_cast<int32_t>(P.size());
        for (size_t k = 0; k < P.size(); ++k) offs[16 * q + k] = P[k];
        vbase[q + 1] = vbase[q] + K[q] * n;
    }
    const int64_t len = vbase[M];
    std::vector<int32_t> map(static_cast<size_t>(len), -1);
    for (int q = 0; q < M; ++q) {
        const int64_t a = b->node_off[q], e = b->node_off[q + 1], n = e - a;
        const int32_t *P = offs.data() + 16 * q;
        for (int64_t i = a; i < e; ++i)
            for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
                const int32_t d = ci[k] - static_cast<int32_t>(i);
                if (d <= 0) continue;
                const int64_t c = std::lower_bound(P, P + K[q], d) - P;
                map[vbase[q] + c * n + (i - a)] = k;
            }
    }
    b->st_K.alloc(M);
    b->st_off.alloc(16 * static_cast<size_t>(M));
    b->st_vbase.alloc(M);
    b->st_map.alloc(len);
    b->st_val32.alloc(len);
    b->st_val64.alloc(len);
    b->st_diag64.alloc(nv);
    b->node_off_dev.alloc(M + 1);
    FK_HIP(hipMemcpyAsync(b->st_K.get(), K.data(), M * sizeof(int32_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(b->st_off.get(), offs.data(), offs.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(b->st_vbase.get(), vbase.data(), M * sizeof(int64_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(b->st_map.get(), map.data(), len * sizeof(int32_t), hipMemcpyHostToDevice, st));
    FK_HIP(hipMemcpyAsync(b->node_off_dev.get(), b->node_off.data(), (M + 1) * sizeof(int64_t),
                          hipMemcpyHostToDevice, st));
    FK_HIP(hipStreamSynchronize(st));
    build_spmm_table(b);
    b->st_len = len;
    b->st_Kc = K[0];
    for (int q = 1; q < M; ++q)
        if (K[q] != K[0]) b->st_Kc = 0;
    b->st_valid = false;
    b->stencil = 1;
    // captured chunks launched the CSR SpMM
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_cfg = -1;
    if (b->ru_graph) (void)hipGraphExecDestroy(b->ru_graph);
    b->ru_graph = nullptr;
    b->ru_graph_cfg = -1;
}

// the SpMM copy for the current values
static void prepare_spmm(fk_batch *b) {
    fk_mesh *m = b->mesh;
    build_batch_stencil(b);
    if (b->stencil == 0) make_val32(m);
    if (b->stencil == 1 && !b->st_valid) {
        kb_stencil_fill<<<grid_for(b->st_len, 256), 256, 0, b->ctx->stream>>>(m->val64.get(), b->st_map.get(),
                                                                             b->st_len, b->st_val32.get());
        kb_stencil_fill64<<<grid_for(std::max(b->st_len, m->nv), 256), 256, 0, b->ctx->stream>>>(
            m->val64.get(), b->st_map.get(), m->diag.get(), b->st_len, m->nv, b->st_val64.get(),
            b->st_diag64.get());
        FK_HIP(hipGetLastError());
        b->st_valid = true;
    }
}

void batch_solve(fk_batch *b, double rtol, double inner_rtol, int max_outer, int64_t max_iter, int64_t *iters_out,
                 int32_t *outer_out, double *relres_out) {
    fk_mesh *m = b->mesh;
    if (!m->have_values) fail(FK_ERR_STATE, "batch not assembled");
    if (!b->have_dirichlet) fail(FK_ERR_STATE, "no Dirichlet data set");
    prepare_spmm(b);
    if (!m->dinv32_valid) {
        // dinv of the FP32 values, masked by the batch's fixed nodes
        m->w32.dinv.ensure(m->nv);
        compute_dinv_f32(m);
        m->dinv32_valid = true;
    }
    if (b->precision == FK_PREC_F32_REFINED)
        batch_solve_ru(b, rtol, inner_rtol, max_iter, iters_out, outer_out, relres_out);
    else
        batch_solve_restart(b, rtol, inner_rtol, max_outer, max_iter, iters_out, outer_out, relres_out);
}

// Autotuner, on the batch itself, with every system active: the SpMM's
// (block, lanes) candidates of the CSR kernel, or (block, rows per thread)
// of the stencil kernel. The update kernels keep 512 threads per block with
// the stencil SpMM: timed the same way (all systems active) 128 won, but in
// the solve, where the systems of small meshes finish first and the last
// iterations run a few large meshes' blocks, the 64-mesh step took 78.3 ms
// with 128 against 72.3 ms with 512 (72.9 with 256).
void batch_autotune(fk_batch *b, int32_t *lanes_out, int32_t *block_out, double *best_us_out) {
    fk_mesh *m = b->mesh;
    if (!m->have_values) fail(FK_ERR_STATE, "batch not assembled");
    prepare_spmm(b);
    hipStream_t st = b->ctx->stream;
    const int64_t nvr = m->nv * b->nrhs;
    // all systems active, p = 1 so the kernel does full work
    std::vector<int32_t> zero(b->nsys, 0);
    FK_HIP(hipMemcpyAsync(b->done.get(), zero.data(), b->nsys * sizeof(int32_t), hipMemcpyHostToDevice, st));
    std::vector<float> ones(static_cast<size_t>(nvr), 1.0f);
    FK_HIP(hipMemcpyAsync(b->p.get(), ones.data(), nvr * sizeof(float), hipMemcpyHostToDevice, st));
    FK_HIP(hipStreamSynchronize(st));
    hipEvent_t e0, e1;
    FK_HIP(hipEventCreate(&e0));
    FK_HIP(hipEventCreate(&e1));
    auto time_us = [&](auto &&launch) {
        launch();  // warm
        FK_HIP(hipEventRecord(e0, st));
        const int reps = 10;
        for (int k = 0; k < reps; ++k) launch();
        FK_HIP(hipEventRecord(e1, st));
        FK_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        FK_HIP(hipEventElapsedTime(&ms, e0, e1));
        return ms * 1e3 / reps;
    };
    const int blocks[3] = {128, 256, 512};
    double best = 1e30;
    if (b->stencil == 1) {
        int bb = b->sp_block, br = b->sp_rpt;
        for (int bi = 0; bi < 3; ++bi)
            for (int rpt = 1; rpt <= 2; ++rpt) {
                b->sp_rpt = rpt;
                const double us = time_us([&] { spmm(b, blocks[bi], b->lanes); });
                if (us < best) {
                    best = us;
                    bb = blocks[bi];
                    br = rpt;
                }
            }
        b->sp_block = bb;
        b->sp_rpt = br;
    } else {
        int bl = b->lanes, bb = b->block;
        const int lanes[3] = {4, 8, 16};
        for (int bi = 0; bi < 3; ++bi)
            for (int li = 0; li < 3; ++li) {
                const double us = time_us([&] { spmm(b, blocks[bi], lanes[li]); });
                if (us < best) {
                    best = us;
                    bl = lanes[li];
                    bb = blocks[bi];
                }
            }
        b->lanes = bl;
        b->block = bb;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    FK_HIP(hipStreamSynchronize(st));
    if (lanes_out) *lanes_out = b->lanes;
    if (block_out) *block_out = b->stencil == 1 ? b->sp_block : b->block;
    if (best_us_out) *best_us_out = best;
}

}  // namespace fk

// ---------------------------------------------------------------------------
using fk::guard;

extern "C" {

int fk_batch_create(fk_context *ctx, int32_t nmesh, const int64_t *nv, const int64_t *nt, const double *const *p,
                    const int32_t *const *t, int32_t index_base, int32_t nrhs, fk_batch **out) {
    if (!ctx || !out || nmesh < 1 || !nv || !nt || !p || !t || nrhs < 1 || nrhs > 4 ||
        (index_base != 0 && index_base != 1))
        return FK_ERR_INVALID_ARG;
    *out = nullptr;
    fk_batch *b = new (std::nothrow) fk_batch();
    if (!b) return FK_ERR_OOM;
    b->ctx = ctx;
    b->nmesh = nmesh;
    b->nrhs = nrhs;
    b->nsys = nmesh * nrhs;
    int s = guard(ctx, "fk_batch_create", [&] {
        fk::set_device(ctx);
        b->node_off.assign(nmesh + 1, 0);
        b->elem_off.assign(nmesh + 1, 0);
        for (int m = 0; m < nmesh; ++m) {
            if (nv[m] < 4 || nt[m] < 1 || !p[m] || !t[m]) fk::fail(FK_ERR_INVALID_ARG, "bad mesh sizes or pointers");
            b->node_off[m + 1] = b->node_off[m] + nv[m];
            b->elem_off[m + 1] = b->elem_off[m] + nt[m];
        }
        const int64_t NV = b->node_off.back(), NT = b->elem_off.back();
        if (NV >= (int64_t(1) << 31) || 16 * NT >= (int64_t(1) << 31))
            fk::fail(FK_ERR_UNSUPPORTED, "batch exceeds 2^31 nodes or 2^27 elements");
        // concatenate: node numbers of mesh m are offset by node_off[m]
        std::vector<double> P(static_cast<size_t>(3 * NV));
        std::vector<int32_t> T(static_cast<size_t>(4 * NT));
        for (int m = 0; m < nmesh; ++m) {
            std::copy(p[m], p[m] + 3 * nv[m], P.begin() + 3 * b->node_off[m]);
            for (int64_t k = 0; k < 4 * nt[m]; ++k) {
                const int32_t v = t[m][k];
                if (v < index_base || v - index_base >= nv[m])
                    fk::fail(FK_ERR_INDEX, "mesh " + std::to_string(m) + ": node number out of range");
                T[4 * b->elem_off[m] + k] = static_cast<int32_t>(v - index_base + b->node_off[m]);
            }
        }
        b->mesh = new fk_mesh();
        b->mesh->ctx = ctx;
        b->mesh->nv = NV;
        b->mesh->nt = NT;
        fk::build_mesh(b->mesh, P.data(), T.data(), 0, nullptr);
        fk::build_block_table(b);
        const size_t nvr = static_cast<size_t>(NV) * nrhs;
        b->x.alloc(nvr); b->r.alloc(nvr); b->p.alloc(nvr); b->q.alloc(nvr);
        const size_t np = static_cast<size_t>(b->nblocks) * nrhs;
        b->part_pq.alloc(np); b->part_rz.alloc(2 * np); b->part_rr.alloc(np); b->part_bb.alloc(np);
        b->done.alloc(b->nsys); b->status.alloc(b->nsys); b->active.alloc(b->nsys); b->ndone.alloc(1);
        b->iters.alloc(b->nsys); b->thresh.alloc(b->nsys); b->scale.alloc(b->nsys);
        // per-system state of the reliable updates; rr32 is written by the
        // iteration kernels of both solve modes
        const size_t S = static_cast<size_t>(b->nsys);
        b->ru_act.alloc(S); b->ru_mode.alloc(S); b->ru_fin.alloc(S); b->ru_nrep.alloc(S);
        b->ru_restarts.alloc(S); b->ru_status.alloc(S); b->nfin.alloc(1);
        b->rr32.alloc(S); b->rr64.alloc(S); b->rr_last.alloc(S); b->thresh64.alloc(S); b->part_res.alloc(np);
        b->alpha.alloc(S);
        b->hpart.alloc(np + 16);
        b->hcount.alloc(4);
    });
    if (s != FK_OK) {
        delete b;
        return s;
    }
    *out = b;
    return FK_OK;
}

int fk_batch_destroy(fk_batch *b) {
    if (!b) return FK_ERR_INVALID_ARG;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    return FK_OK;
}

int fk_batch_info(const fk_batch *b, int64_t *nodes, int64_t *nnz, int32_t *nblocks, int32_t *ncolours) {
    if (!b) return FK_ERR_INVALID_ARG;
    if (nodes) *nodes = b->mesh->nv;
    if (nnz) *nnz = b->mesh->nnz;
    if (nblocks) *nblocks = b->nblocks;
    if (ncolours) *ncolours = b->mesh->ncolours;
    return FK_OK;
}

int fk_batch_set_coefficient(fk_batch *b, const double *const *f) {
    if (!b || !f) return FK_ERR_INVALID_ARG;
    return guard(b->ctx, "fk_batch_set_coefficient", [&] {
        fk::set_device(b->ctx);
        std::vector<double> F(static_cast<size_t>(b->mesh->nt));
        for (int m = 0; m < b->nmesh; ++m) {
            if (!f[m]) fk::fail(FK_ERR_INVALID_ARG, "null coefficient array");
            std::copy(f[m], f[m] + (b->elem_off[m + 1] - b->elem_off[m]), F.begin() + b->elem_off[m]);
        }
        fk::upload_coefficient(b->mesh, F.data());
    });
}

int fk_batch_assemble(fk_batch *b) {
    if (!b) return FK_ERR_INVALID_ARG;
    return guard(b->ctx, "fk_batch_assemble", [&] {
        fk::set_device(b->ctx);
        fk::assemble(b->mesh, FK_ASM_GATHER);
        b->st_valid = false;
    });
}

int fk_batch_set_dirichlet(fk_batch *b, const uint8_t *const *mask, const double *const *g) {
    if (!b || !mask || !g) return FK_ERR_INVALID_ARG;
    return guard(b->ctx, "fk_batch_set_dirichlet", [&] {
        fk::set_device(b->ctx);
        const int64_t NV = b->mesh->nv;
        const int nr = b->nrhs;
        std::vector<uint8_t> M(static_cast<size_t>(NV));
        std::vector<double> G(static_cast<size_t>(NV) * nr, 0.0), G0(static_cast<size_t>(NV), 0.0);
        for (int m = 0; m < b->nmesh; ++m) {
            if (!mask[m] || !g[m]) fk::fail(FK_ERR_INVALID_ARG, "null Dirichlet array");
            const int64_t a = b->node_off[m], n = b->node_off[m + 1] - a;
            int64_t fixed = 0;
            for (int64_t i = 0; i < n; ++i) {
                M[a + i] = mask[m][i] ? 1 : 0;
                fixed += M[a + i];
                for (int j = 0; j < nr; ++j) {
                    const double v = g[m][j * n + i];  // right-hand side major per mesh
                    if (M[a + i] && !std::isfinite(v)) fk::fail(FK_ERR_NONFINITE, "non-finite Dirichlet value");
                    G[(a + i) * nr + j] = M[a + i] ? v : 0.0;
                }
            }
            if (fixed == 0) fk::fail(FK_ERR_SINGULAR, "mesh " + std::to_string(m) + " has no fixed node");
        }
        fk::set_dirichlet(b->mesh, M.data(), G0.data());
        b->g.ensure(static_cast<size_t>(NV) * nr);
        FK_HIP(hipMemcpy(b->g.get(), G.data(), G.size() * sizeof(double), hipMemcpyHostToDevice));
        b->have_dirichlet = true;
    });
}

int fk_batch_solve(fk_batch *b, double rtol, double inner_rtol, int32_t max_outer, int64_t max_iter,
                   int64_t *iterations, int32_t *outer, double *relres) {
    if (!b || !(rtol >= 0.0) || !(inner_rtol > 0.0) || !(inner_rtol < 1.0) || max_outer < 1 || max_iter < 1)
        return FK_ERR_INVALID_ARG;
    return guard(b->ctx, "fk_batch_solve", [&] {
        fk::set_device(b->ctx);
        fk::batch_solve(b, rtol, inner_rtol, max_outer, max_iter, iterations, outer, relres);
    });
}

int fk_batch_potential(fk_batch *b, int32_t mesh, double *phi) {
    if (!b || !phi || mesh < 0 || mesh >= b->nmesh) return FK_ERR_INVALID_ARG;
    return guard(b->ctx, "fk_batch_potential", [&] {
        if (!b->have_phi) fk::fail(FK_ERR_STATE, "no potential");
        fk::set_device(b->ctx);
        const int64_t a = b->node_off[mesh], n = b->node_off[mesh + 1] - a;
        const int nr = b->nrhs;
        std::vector<double> tmp(static_cast<size_t>(n) * nr);
        FK_HIP(hipMemcpyAsync(tmp.data(), b->phi.get() + a * nr, tmp.size() * sizeof(double), hipMemcpyDeviceToHost,
                              b->ctx->stream));
        FK_HIP(hipStreamSynchronize(b->ctx->stream));
        for (int64_t i = 0; i < n; ++i)
            for (int j = 0; j < nr; ++j) phi[j * n + i] = tmp[i * nr + j];
    });
}

int fk_batch_autotune(fk_batch *b, int32_t *lanes, int32_t *block, double *best_us) {
    if (!b) return FK_ERR_INVALID_ARG;
    return guard(b->ctx, "fk_batch_autotune", [&] {
        fk::set_device(b->ctx);
        fk::batch_autotune(b, lanes, block, best_us);
    });
}

int fk_batch_set_precision(fk_batch *b, int32_t precision) {
    if (!b || (precision != FK_PREC_F32_REFINED && precision != FK_PREC_F32_RESTART)) return FK_ERR_INVALID_ARG;
    b->precision = precision;
    return FK_OK;
}

int fk_batch_set_format(fk_batch *b, int32_t format) {
    if (!b || (format != FK_SPMV_CSR && format != FK_SPMV_STENCIL && format != FK_SPMV_AUTO))
        return FK_ERR_INVALID_ARG;
    if (b->stencil >= 0 && format != b->format) {
        b->ctx->err = "fk_batch_set_format: the SpMM copy is already chosen";
        return FK_ERR_STATE;
    }
    b->format = format;
    return FK_OK;
}

int fk_batch_solver_format(fk_batch *b, int32_t *format) {
    if (!b || !format) return FK_ERR_INVALID_ARG;
    *format = b->stencil < 0 ? -1 : (b->stencil == 1 ? FK_SPMV_STENCIL : FK_SPMV_CSR);
    return FK_OK;
}

int fk_batch_launch(fk_batch *b, int32_t *lanes, int32_t *block, int32_t *spmm_block, int32_t *spmm_rows) {
    if (!b) return FK_ERR_INVALID_ARG;
    if (lanes) *lanes = b->lanes;
    if (block) *block = b->block;
    if (spmm_block) *spmm_block = b->sp_block;
    if (spmm_rows) *spmm_rows = b->sp_rpt;
    return FK_OK;
}

int fk_batch_set_launch(fk_batch *b, int32_t lanes, int32_t block) {
    if (!b || (lanes != 4 && lanes != 8 && lanes != 16) || (block != 128 && block != 256 && block != 512))
        return FK_ERR_INVALID_ARG;
    b->lanes = lanes;
    b->block = block;
    return FK_OK;
}

}  // extern "C"
