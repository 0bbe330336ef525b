// Context, mesh upload, element colouring and the CSR pattern / slot map.
//
// The pattern is the set of sorted unique node pairs (t[i,k], t[j,k]) over
// all elements with explicit zeros kept: the structure Julia's
// sparse(rowIDs, colIDs, Acsc) produces from the fixed COO indices of
// Magnetostatics/NewtonRaphson.jl:115-127,161. It is built on the GPU by a
// radix sort of the 16*nt packed keys row*nv+col; the stable sort carries the
// triplet id 16e+n so the same pass yields the element -> CSR-slot map that
// value-only re-assembly scatters through.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "device.h"
#include "internal.h"

namespace fk {

void set_device(fk_context *ctx) { FK_HIP(hipSetDevice(ctx->device)); }

// ---------------------------------------------------------------------------
// kernels

__global__ void k_pair_keys(const int4 *__restrict__ tet, int64_t nt, uint64_t nv, uint64_t *__restrict__ keys,
                            uint32_t *__restrict__ vals) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 16 * nt) return;
    const int64_t e = i >> 4;
    const int n = static_cast<int>(i & 15);
    const int4 nd = tet[e];
    const int a = n >> 2, b = n & 3;  // n = 4a + b, a-major as the Julia COO loops
    const int na = a == 0 ? nd.x : a == 1 ? nd.y : a == 2 ? nd.z : nd.w;
    const int nb = b == 0 ? nd.x : b == 1 ? nd.y : b == 2 ? nd.z : nd.w;
    keys[i] = static_cast<uint64_t>(na) * nv + static_cast<uint64_t>(nb);
    vals[i] = static_cast<uint32_t>(i);
}

__global__ void k_head_flags(const uint64_t *__restrict__ keys, int64_t n, int32_t *__restrict__ flags) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ void k_scatter_pattern(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ vals,
                                  const int32_t *__restrict__ seg, int64_t n, uint64_t nv,
                                  int32_t *__restrict__ slot, int32_t *__restrict__ colidx,
                                  int32_t *__restrict__ rowcnt, int32_t *__restrict__ diag) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t s = seg[i] - 1;
    slot[vals[i]] = s;
    const uint64_t key = keys[i];
    if (i == 0 || key != keys[i - 1]) {
        const uint64_t row = key / nv, col = key - row * nv;
        colidx[s] = static_cast<int32_t>(col);
        atomicAdd(&rowcnt[row], 1);
        if (row == col) diag[row] = s;
    }
}

__global__ void k_fill_i32(int32_t *__restrict__ a, int64_t n, int32_t v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_count_unused(const int32_t *__restrict__ rowcnt, int64_t nv, int32_t *__restrict__ counter) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nv && rowcnt[i] == 0) atomicAdd(counter, 1);
}

// det / volume check of every element; counters[0] = degenerate elements,
// counters[1] = VE mismatches (VE given in caller order, may be null)
__global__ void k_check_elements(const int4 *__restrict__ tet, const double *__restrict__ X,
                                 const double *__restrict__ Y, const double *__restrict__ Z, int64_t nt,
                                 const int32_t *__restrict__ perm, const double *__restrict__ VE,
                                 int32_t *__restrict__ counters) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nt) return;
    const Grad G = element_gradients(X, Y, Z, tet[e]);
    if (!(G.det != 0.0) || !isfinite(G.det)) atomicAdd(&counters[0], 1);
    if (VE) {
        const double vol = fabs(G.det) / 6.0;
        const double ve = VE[perm[e]];
        if (!(fabs(ve - vol) <= 1e-10 * fabs(vol))) atomicAdd(&counters[1], 1);
    }
}

__global__ void k_volumes(const int4 *__restrict__ tet, const double *__restrict__ X, const double *__restrict__ Y,
                          const double *__restrict__ Z, int64_t nt, const int32_t *__restrict__ perm,
                          double *__restrict__ VE) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nt) return;
    const Grad G = element_gradients(X, Y, Z, tet[e]);
    VE[perm[e]] = fabs(G.det) / 6.0;
}

// ---------------------------------------------------------------------------
// host-side colouring: greedy, elements in caller order, node bit masks of
// used colours. Two elements of one colour share no node, so a colour's
// element scatter needs no atomics.
template <int W>
static bool greedy_colour(const int32_t *t, int64_t nt, int64_t nv, int32_t base, std::vector<uint8_t> &colour,
                          int32_t &ncolours) {
    std::vector<uint64_t> used(static_cast<size_t>(nv) * W, 0ull);
    int32_t maxc = -1;
    for (int64_t e = 0; e < nt; ++e) {
        const int32_t *nd = t + 4 * e;
        int c = -1;
        for (int w = 0; w < W && c < 0; ++w) {
            uint64_t m = 0;
            for (int i = 0; i < 4; ++i) m |= used[static_cast<size_t>(nd[i] - base) * W + w];
            if (~m) c = 64 * w + __builtin_ctzll(~m);
        }
        if (c < 0) return false;
        colour[e] = static_cast<uint8_t>(c);
        for (int i = 0; i < 4; ++i) used[static_cast<size_t>(nd[i] - base) * W + (c >> 6)] |= 1ull << (c & 63);
        maxc = std::max(maxc, c);
    }
    ncolours = maxc + 1;
    return true;
}

void setup_partition(fk_mesh *m) {
    // persistent grid of the solver kernels: ~4 blocks of 256 per CU
    const int64_t want = std::max<int64_t>(1, (m->nv + kBlock - 1) / kBlock);
    int64_t nb = std::min<int64_t>(want, 4 * static_cast<int64_t>(m->ctx->num_cus));
    if (nb >= kNumXcd) nb = (nb / kNumXcd) * kNumXcd;
    m->nblocks = static_cast<int>(nb);
    // bands of whole 64-row groups: 16-byte aligned vector access in every band
    m->chunk = ((m->nv + nb - 1) / nb + 63) / 64 * 64;
}

void validate_input(const double *p, const int32_t *t, int64_t nv, int64_t nt, int32_t base) {
    for (int64_t i = 0; i < 3 * nv; ++i)
        if (!std::isfinite(p[i])) fail(FK_ERR_NONFINITE, "non-finite node coordinate");
    for (int64_t e = 0; e < nt; ++e) {
        const int32_t *nd = t + 4 * e;
        for (int i = 0; i < 4; ++i) {
            if (nd[i] < base || static_cast<int64_t>(nd[i]) - base >= nv)
                fail(FK_ERR_INDEX, "element " + std::to_string(e + base) + " has node number out of range");
            for (int j = 0; j < i; ++j)
                if (nd[i] == nd[j]) fail(FK_ERR_DEGENERATE, "element " + std::to_string(e + base) + " repeats a node");
        }
    }
}

// p, t already validated (and in the internal numbering)
void build_mesh(fk_mesh *m, const double *p, const int32_t *t, int32_t base, const double *VE) {
    const int64_t nv = m->nv, nt = m->nt;
    hipStream_t st = m->ctx->stream;
    // ---- colour and order elements by colour
    std::vector<uint8_t> colour(static_cast<size_t>(nt));
    int32_t nc = 0;
    if (!greedy_colour<2>(t, nt, nv, base, colour, nc) && !greedy_colour<4>(t, nt, nv, base, colour, nc))
        fail(FK_ERR_UNSUPPORTED, "element colouring needs more than 256 colours");
    m->ncolours = nc;
    m->colour_off.assign(static_cast<size_t>(nc) + 1, 0);
    for (int64_t e = 0; e < nt; ++e) m->colour_off[colour[e] + 1]++;
    for (int c = 0; c < nc; ++c) m->colour_off[c + 1] += m->colour_off[c];
    std::vector<int64_t> cursor(m->colour_off.begin(), m->colour_off.end() - 1);
    std::vector<int32_t> perm(static_cast<size_t>(nt));
    std::vector<int4> tet(static_cast<size_t>(nt));
    for (int64_t e = 0; e < nt; ++e) {
        const int64_t pos = cursor[colour[e]]++;
        perm[pos] = static_cast<int32_t>(e);
        const int32_t *nd = t + 4 * e;
        tet[pos] = make_int4(nd[0] - base, nd[1] - base, nd[2] - base, nd[3] - base);
    }
    colour.clear();
    colour.shrink_to_fit();
    // ---- upload geometry
    {
        std::vector<double> xs(static_cast<size_t>(nv)), ys(static_cast<size_t>(nv)), zs(static_cast<size_t>(nv));
        for (int64_t i = 0; i < nv; ++i) { xs[i] = p[3 * i]; ys[i] = p[3 * i + 1]; zs[i] = p[3 * i + 2]; }
        m->X.alloc(nv); m->Y.alloc(nv); m->Z.alloc(nv);
        FK_HIP(hipMemcpy(m->X.get(), xs.data(), nv * sizeof(double), hipMemcpyHostToDevice));
        FK_HIP(hipMemcpy(m->Y.get(), ys.data(), nv * sizeof(double), hipMemcpyHostToDevice));
        FK_HIP(hipMemcpy(m->Z.get(), zs.data(), nv * sizeof(double), hipMemcpyHostToDevice));
    }
    m->tet.alloc(nt);
    m->perm.alloc(nt);
    m->tet_nat.alloc(nt);
    FK_HIP(hipMemcpy(m->tet.get(), tet.data(), nt * sizeof(int4), hipMemcpyHostToDevice));
    FK_HIP(hipMemcpy(m->perm.get(), perm.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice));
    // the same elements in caller order: field recovery and the row-gather
    // assembly walk this copy, whose node reads follow the mesh's own locality
    for (int64_t e = 0; e < nt; ++e) {
        const int32_t *nd = t + 4 * e;
        tet[e] = make_int4(nd[0] - base, nd[1] - base, nd[2] - base, nd[3] - base);
    }
    FK_HIP(hipMemcpy(m->tet_nat.get(), tet.data(), nt * sizeof(int4), hipMemcpyHostToDevice));
    // the row-gather kernels' element order: by lowest node (stable counting
    // sort), so that the elements of neighbouring rows lie near each other in
    // memory; kept only when the caller's order is not already sorted so
    {
        std::vector<int32_t> key(static_cast<size_t>(nt));
        bool sorted = true;
        for (int64_t e = 0; e < nt; ++e) {
            const int4 q = tet[e];
            key[e] = std::min(std::min(q.x, q.y), std::min(q.z, q.w));
            if (e > 0 && key[e] < key[e - 1]) sorted = false;
        }
        m->gather_sorted = !sorted;
        if (!sorted) {
            std::vector<int64_t> start(static_cast<size_t>(nv) + 1, 0);
            for (int64_t e = 0; e < nt; ++e) ++start[key[e] + 1];
            for (int64_t i = 0; i < nv; ++i) start[i + 1] += start[i];
            std::vector<int32_t> c2s(static_cast<size_t>(nt));  // caller -> sorted position
            std::vector<int4> tg(static_cast<size_t>(nt));
            for (int64_t e = 0; e < nt; ++e) {
                const int64_t s = start[key[e]]++;
                c2s[e] = static_cast<int32_t>(s);
                tg[s] = tet[e];
            }
            // colour position -> sorted position
            for (int64_t e = 0; e < nt; ++e) perm[e] = c2s[perm[e]];
            m->tet_g.alloc(nt);
            m->perm_g.alloc(nt);
            FK_HIP(hipMemcpy(m->tet_g.get(), tg.data(), nt * sizeof(int4), hipMemcpyHostToDevice));
            FK_HIP(hipMemcpy(m->perm_g.get(), perm.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice));
        }
    }
    tet.clear(); tet.shrink_to_fit();
    perm.clear(); perm.shrink_to_fit();
    // ---- element checks
    {
        DevBuf<int32_t> counters(2);
        DevBuf<double> dVE;
        FK_HIP(hipMemsetAsync(counters.get(), 0, 2 * sizeof(int32_t), st));
        if (VE) {
            dVE.alloc(nt);
            FK_HIP(hipMemcpyAsync(dVE.get(), VE, nt * sizeof(double), hipMemcpyHostToDevice, st));
        }
        k_check_elements<<<grid_for(nt, 256), 256, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(), nt,
                                                            m->perm.get(), VE ? dVE.get() : nullptr,
                                                            counters.get());
        FK_HIP(hipGetLastError());
        int32_t h[2];
        FK_HIP(hipMemcpyAsync(h, counters.get(), sizeof(h), hipMemcpyDeviceToHost, st));
        FK_HIP(hipStreamSynchronize(st));
        if (h[0]) fail(FK_ERR_DEGENERATE, std::to_string(h[0]) + " element(s) with zero or non-finite volume");
        if (h[1]) fail(FK_ERR_INVALID_ARG, std::to_string(h[1]) + " VE entries differ from the element volumes");
    }
    // ---- pattern and slot map
    const int64_t N = 16 * nt;
    {
        DevBuf<uint64_t> k0(N), k1(N);
        DevBuf<uint32_t> v0(N), v1(N);
        k_pair_keys<<<grid_for(N, 256), 256, 0, st>>>(m->tet.get(), nt, static_cast<uint64_t>(nv), k0.get(), v0.get());
        FK_HIP(hipGetLastError());
        const uint64_t maxkey = static_cast<uint64_t>(nv) * static_cast<uint64_t>(nv) - 1;
        const unsigned end_bit = maxkey == 0 ? 1u : 64u - static_cast<unsigned>(__builtin_clzll(maxkey));
        rocprim::double_buffer<uint64_t> kb(k0.get(), k1.get());
        rocprim::double_buffer<uint32_t> vb(v0.get(), v1.get());
        size_t tmp_bytes = 0;
        FK_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, kb, vb, static_cast<size_t>(N), 0u, end_bit, st));
        DevBuf<uint8_t> tmp(tmp_bytes + 16);
        FK_HIP(rocprim::radix_sort_pairs(tmp.get(), tmp_bytes, kb, vb, static_cast<size_t>(N), 0u, end_bit, st));
        const uint64_t *keys = kb.current();
        const uint32_t *vals = vb.current();
        // head flags and segment ids; reuse the alternate key buffer as int32 storage
        int32_t *flags = reinterpret_cast<int32_t *>(kb.alternate());
        int32_t *seg = flags + N;  // alternate key buffer holds 2N int32
        k_head_flags<<<grid_for(N, 256), 256, 0, st>>>(keys, N, flags);
        FK_HIP(hipGetLastError());
        size_t scan_bytes = 0;
        FK_HIP(rocprim::inclusive_scan(nullptr, scan_bytes, flags, seg, static_cast<size_t>(N),
                                       rocprim::plus<int32_t>(), st));
        DevBuf<uint8_t> stmp(scan_bytes + 16);
        FK_HIP(rocprim::inclusive_scan(stmp.get(), scan_bytes, flags, seg, static_cast<size_t>(N),
                                       rocprim::plus<int32_t>(), st));
        int32_t nnz = 0;
        FK_HIP(hipMemcpyAsync(&nnz, seg + N - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        FK_HIP(hipStreamSynchronize(st));
        m->nnz = nnz;
        m->slot.alloc(N);
        m->colidx.alloc(nnz);
        m->diag.alloc(nv);
        DevBuf<int32_t> rowcnt(nv + 1);
        k_fill_i32<<<grid_for(nv + 1, 256), 256, 0, st>>>(rowcnt.get(), nv + 1, 0);
        k_fill_i32<<<grid_for(nv, 256), 256, 0, st>>>(m->diag.get(), nv, -1);
        k_scatter_pattern<<<grid_for(N, 256), 256, 0, st>>>(keys, vals, seg, N, static_cast<uint64_t>(nv),
                                                            m->slot.get(), m->colidx.get(), rowcnt.get(),
                                                            m->diag.get());
        FK_HIP(hipGetLastError());
        DevBuf<int32_t> unused(1);
        k_fill_i32<<<1, 64, 0, st>>>(unused.get(), 1, 0);
        k_count_unused<<<grid_for(nv, 256), 256, 0, st>>>(rowcnt.get(), nv, unused.get());
        m->rowptr.alloc(nv + 1);
        size_t ex_bytes = 0;
        FK_HIP(rocprim::exclusive_scan(nullptr, ex_bytes, rowcnt.get(), m->rowptr.get(), int32_t(0),
                                       static_cast<size_t>(nv + 1), rocprim::plus<int32_t>(), st));
        DevBuf<uint8_t> etmp(ex_bytes + 16);
        FK_HIP(rocprim::exclusive_scan(etmp.get(), ex_bytes, rowcnt.get(), m->rowptr.get(), int32_t(0),
                                       static_cast<size_t>(nv + 1), rocprim::plus<int32_t>(), st));
        int32_t hu = 0;
        FK_HIP(hipMemcpyAsync(&hu, unused.get(), sizeof(int32_t), hipMemcpyDeviceToHost, st));
        FK_HIP(hipStreamSynchronize(st));
        m->unused_nodes = hu;
    }
    setup_partition(m);
    m->hstate.alloc(4);
    m->hpart.alloc(static_cast<size_t>(std::max(m->nblocks, 1)) * 4 + 64);
}

void element_volumes(fk_mesh *m, double *VE_host) {
    hipStream_t st = m->ctx->stream;
    DevBuf<double> d(m->nt);
    k_volumes<<<grid_for(m->nt, 256), 256, 0, st>>>(m->tet.get(), m->X.get(), m->Y.get(), m->Z.get(), m->nt,
                                                     m->perm.get(), d.get());
    FK_HIP(hipGetLastError());
    FK_HIP(hipMemcpyAsync(VE_host, d.get(), m->nt * sizeof(double), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
}

}  // namespace fk

// ---------------------------------------------------------------------------
// C ABI: context and mesh

using fk::guard;

extern "C" {

int fk_version(void) { return 100; }

const char *fk_status_string(int s) {
    switch (s) {
        case FK_OK: return "ok";
        case FK_ERR_INVALID_ARG: return "invalid argument";
        case FK_ERR_HIP: return "HIP runtime error";
        case FK_ERR_DEGENERATE: return "degenerate element";
        case FK_ERR_INDEX: return "node index out of range";
        case FK_ERR_NONFINITE: return "non-finite value";
        case FK_ERR_NOT_CONVERGED: return "not converged";
        case FK_ERR_SINGULAR: return "singular system";
        case FK_ERR_OOM: return "out of memory";
        case FK_ERR_STATE: return "wrong call order";
        case FK_ERR_UNSUPPORTED: return "unsupported mesh";
        case FK_ERR_DOMAIN: return "argument outside the table range";
        default: return "unknown status";
    }
}

int fk_context_create(int device, fk_context **out) {
    if (!out) return FK_ERR_INVALID_ARG;
    *out = nullptr;
    fk_context *ctx = nullptr;
    try {
        int count = 0;
        FK_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) return FK_ERR_INVALID_ARG;
        ctx = new fk_context();
        ctx->device = device;
        FK_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        FK_HIP(hipGetDeviceProperties(&prop, device));
        ctx->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        // Iteration chunks are captured in hipGraphs unless FEMEKO_HIP_GRAPHS=0,
        // or a rocprofiler-sdk tool (rocprofv3) is attached and
        // FEMEKO_HIP_GRAPHS=1 does not insist: under rocprofv3 --kernel-trace
        // a graph launch of the FP64 iteration chunk once ended the host
        // process with SIGSEGV, while the host code around the launch (the
        // pinned state copies, two of the four allocated, and the event
        // waits) reads only memory the context owns and the same path runs
        // clean without the tool; the kernels are the same either way.
        const char *g = std::getenv("FEMEKO_HIP_GRAPHS");
        const char *tool = std::getenv("ROCP_TOOL_LIBRARIES");
        const bool profiled = tool && tool[0] != '\0';
        ctx->graphs = g ? g[0] != '0' : !profiled;
        FK_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        fk::noalloc_setup(device);
    } catch (const fk::Error &e) {
        delete ctx;
        return e.code;
    } catch (...) {
        delete ctx;
        return FK_ERR_OOM;
    }
    *out = ctx;
    return FK_OK;
}

int fk_context_destroy(fk_context *ctx) {
    if (!ctx) return FK_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (hipEvent_t e : ctx->timing.ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FK_OK;
}

const char *fk_context_last_error(const fk_context *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int fk_context_synchronize(fk_context *ctx) {
    if (!ctx) return FK_ERR_INVALID_ARG;
    return guard(ctx, "fk_context_synchronize", [&] {
        fk::set_device(ctx);
        FK_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int fk_mesh_create(fk_context *ctx, const double *p, int64_t nv, const int32_t *t, int64_t nt, int32_t index_base,
                   const double *VE, fk_mesh **out) {
    return fk_mesh_create_ex(ctx, p, nv, t, nt, index_base, VE, FK_REORDER_AUTO, out);
}

int fk_mesh_create_ex(fk_context *ctx, const double *p, int64_t nv, const int32_t *t, int64_t nt,
                      int32_t index_base, const double *VE, int32_t reorder, fk_mesh **out) {
    if (!ctx || !out || reorder < FK_REORDER_AUTO || reorder > FK_REORDER_RCM) return FK_ERR_INVALID_ARG;
    *out = nullptr;
    if (!p || !t || nv < 4 || nt < 1 || (index_base != 0 && index_base != 1)) {
        ctx->err = "fk_mesh_create: bad arguments";
        return FK_ERR_INVALID_ARG;
    }
    if (nv >= (int64_t(1) << 31) || 16 * nt >= (int64_t(1) << 31)) {
        ctx->err = "fk_mesh_create: mesh exceeds 2^31 nodes or 2^27 elements";
        return FK_ERR_UNSUPPORTED;
    }
    fk_mesh *m = new (std::nothrow) fk_mesh();
    if (!m) return FK_ERR_OOM;
    m->ctx = ctx;
    m->nv = nv;
    m->nt = nt;
    int s = guard(ctx, "fk_mesh_create", [&] {
        fk::set_device(ctx);
        fk::validate_input(p, t, nv, nt, index_base);
        std::vector<int32_t> perm =
            fk::choose_numbering(t, nt, nv, index_base, reorder, &m->mean_dist_caller, &m->mean_dist_internal);
        if (perm.empty()) {
            fk::build_mesh(m, p, t, index_base, VE);
            return;
        }
        // the mesh in the internal numbering: p's columns permuted, t mapped
        std::vector<int32_t> inv(static_cast<size_t>(nv));
        for (int64_t k = 0; k < nv; ++k) inv[perm[k]] = static_cast<int32_t>(k);
        std::vector<double> pi(static_cast<size_t>(3 * nv));
        for (int64_t k = 0; k < nv; ++k)
            for (int c = 0; c < 3; ++c) pi[3 * k + c] = p[3 * static_cast<int64_t>(perm[k]) + c];
        std::vector<int32_t> ti(static_cast<size_t>(4 * nt));
        for (int64_t q = 0; q < 4 * nt; ++q) ti[q] = inv[t[q] - index_base];
        m->reordered = true;
        m->node_perm = std::move(perm);
        m->node_inv = std::move(inv);
        fk::build_mesh(m, pi.data(), ti.data(), 0, VE);
        // device copies for the entry points that take device arrays in the
        // caller's node order (allocated here: those calls allocate nothing)
        m->d_node_perm.alloc(nv);
        m->phi_internal.alloc(nv);
        FK_HIP(hipMemcpy(m->d_node_perm.get(), m->node_perm.data(), nv * sizeof(int32_t), hipMemcpyHostToDevice));
    });
    if (s != FK_OK) {
        delete m;
        return s;
    }
    *out = m;
    return FK_OK;
}

int fk_mesh_destroy(fk_mesh *m) {
    if (!m) return FK_ERR_INVALID_ARG;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    delete m;
    return FK_OK;
}

int fk_mesh_info(const fk_mesh *m, int64_t *nnz, int32_t *ncolours, int64_t *unused) {
    if (!m) return FK_ERR_INVALID_ARG;
    if (nnz) *nnz = m->nnz;
    if (ncolours) *ncolours = m->ncolours;
    if (unused) *unused = m->unused_nodes;
    return FK_OK;
}

int fk_values_generation(const fk_mesh *m, int64_t *generation) {
    if (!m || !generation) return FK_ERR_INVALID_ARG;
    *generation = m->values_gen;
    return FK_OK;
}

int fk_mesh_volumes(fk_mesh *m, double *VE) {
    if (!m || !VE) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_mesh_volumes", [&] {
        fk::set_device(m->ctx);
        fk::element_volumes(m, VE);
    });
}

int fk_pattern(fk_mesh *m, int32_t base, int32_t *rowptr, int32_t *colidx) {
    if (!m || !rowptr || !colidx || (base != 0 && base != 1)) return FK_ERR_INVALID_ARG;
    return guard(m->ctx, "fk_pattern", [&] {
        fk::set_device(m->ctx);
        hipStream_t st = m->ctx->stream;
        if (m->reordered) {
            fk::build_caller_csr(m);
            std::copy(m->caller_rowptr.begin(), m->caller_rowptr.end(), rowptr);
            std::copy(m->caller_colidx.begin(), m->caller_colidx.end(), colidx);
        } else {
            FK_HIP(hipMemcpyAsync(rowptr, m->rowptr.get(), (m->nv + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
            FK_HIP(hipMemcpyAsync(colidx, m->colidx.get(), m->nnz * sizeof(int32_t), hipMemcpyDeviceToHost, st));
            FK_HIP(hipStreamSynchronize(st));
        }
        if (base) {
            for (int64_t i = 0; i <= m->nv; ++i) rowptr[i] += 1;
            for (int64_t i = 0; i < m->nnz; ++i) colidx[i] += 1;
        }
    });
}

}  // extern "C"
