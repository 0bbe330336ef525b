// Internal declarations of the femeko_hip library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "femeko_hip.h"

namespace fk {

// ---------------------------------------------------------------------------
// errors: thrown inside the library, turned into status codes at the C ABI
struct Error {
    int code;
    std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error{code, msg}; }

#define FK_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            ::fk::fail(e_ == hipErrorOutOfMemory ? FK_ERR_OOM : FK_ERR_HIP,                   \
                       std::string(#call) + ": " + hipGetErrorString(e_));                    \
    } while (0)

// ---------------------------------------------------------------------------
// device buffer with ownership
template <typename T>
struct DevBuf {
    T *ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        FK_HIP(hipMalloc(reinterpret_cast<void **>(&ptr), count * sizeof(T)));
        n = count;
    }
    void ensure(size_t count) { if (n < count) alloc(count); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    T *get() const { return ptr; }
    explicit operator bool() const { return ptr != nullptr; }
};

// pinned host buffer
template <typename T>
struct HostBuf {
    T *ptr = nullptr;
    size_t n = 0;
    HostBuf() = default;
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    ~HostBuf() { if (ptr) (void)hipHostFree(ptr); }
    void alloc(size_t count) {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        FK_HIP(hipHostMalloc(reinterpret_cast<void **>(&ptr), count * sizeof(T), hipHostMallocDefault));
        n = count;
    }
};

// ---------------------------------------------------------------------------
// state of one PCG run, read by every kernel of the iteration
struct CgState {
    // stop flags by iteration parity: the kernels of an iteration with parity
    // c read done_at[c] on entry and only ever write done_at[c ^ 1], so no
    // kernel reads a flag that a block of the same launch may be writing
    int32_t done_at[2];
    int32_t status;     // FK_OK, FK_ERR_NOT_CONVERGED, FK_ERR_SINGULAR
    int64_t iters;      // completed iterations
    double thresh;      // rtol^2 * ||b||^2
    double bb;          // ||b||^2
    double rr;          // ||r||^2 at stop
    int64_t max_iter;   // iteration limit of this run
    double alpha;       // step length of the current iteration (0 after a breakdown)
    double pad;
};

// State of the FP32 solve's residual replacements. The decisions are taken
// on the device between iteration chunks, so the host only polls for the end.
struct RuState {
    int32_t act;        // this chunk: 0 none, 1 fold x and replace r, 2 drop x (FP32 breakdown)
    int32_t mode;       // decided: 1 keep the direction p, 2 restart p = D^-1 r
    int32_t finished;   // the solve is over (converged or failed)
    int32_t status;     // its status
    int32_t nrep;       // replacements done
    int32_t restarts;   // direction restarts done
    int64_t max_iter;   // FP32 iteration limit
    double s;           // residual scale 1 / ||r0||
    double rr64;        // last FP64 ||r||^2
    double rr_last;     // scaled ||r||^2 at the last replacement
    double thresh64;    // rtol^2 ||b||^2
    double target2;     // the same in scaled units
    double delta2;      // squared update trigger
    double pad[2];
};

constexpr int kBlock = 256;      // threads per block of the solver kernels
constexpr int kGatherRows = 128;      // rows per block of the row-gather assembly
constexpr int kGatherLdsMax = 8192;   // its largest per-block value segment (doubles, 64 KB)
constexpr int kNumXcd = 8;       // MI355X accelerator complex dies

// Solver workspace of one precision.
template <typename T>
struct PcgWork {
    DevBuf<T> x, r, p, q, dinv;
    DevBuf<double> part_pq, part_rz, part_rr, part_bb;  // part_rz holds 2 x nblocks
    DevBuf<CgState> state;
    // captured chunks of iterations (plain PCG, and PCG + residual
    // replacement); they bake in device addresses, so they are reset
    // whenever a buffer their kernels read is reallocated
    hipGraphExec_t graph = nullptr;
    int graph_iters = 0;
    int graph_lanes = 0;
    hipGraphExec_t ru_graph = nullptr;
    int ru_graph_cfg = -1;
    const double *ru_graph_u = nullptr, *ru_graph_b = nullptr;  // the unknown and RHS it baked in
    const uint8_t *ru_graph_mask = nullptr;                        // and the mask
    void reset_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_iters = 0;
        if (ru_graph) (void)hipGraphExecDestroy(ru_graph);
        ru_graph = nullptr;
        ru_graph_cfg = -1;
        ru_graph_u = ru_graph_b = nullptr;
        ru_graph_mask = nullptr;
    }
    ~PcgWork() { reset_graph(); }
};

struct TimingEvents {
    bool enabled = false;
    std::vector<hipEvent_t> ev;  // pairs
    size_t used = 0;
    int64_t launches = 0;
    double total_ms = 0.0;
};

}  // namespace fk

struct fk_context {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 256;
    bool graphs = true;  // capture PCG iteration chunks in hipGraphs (FEMEKO_HIP_GRAPHS=0 disables)
    std::string err;
    fk::TimingEvents timing;
};

struct fk_mesh {
    fk_context *ctx = nullptr;
    int64_t nv = 0, nt = 0, nnz = 0;
    int64_t unused_nodes = 0;
    int32_t ncolours = 0;
    std::vector<int64_t> colour_off;  // ncolours + 1, offsets in colour order

    // geometry and connectivity (elements in colour order)
    fk::DevBuf<double> X, Y, Z;       // node coordinates, SoA
    fk::DevBuf<int4> tet;             // 0-based nodes per element
    fk::DevBuf<int32_t> perm;         // colour position -> caller element index
    // CSR pattern
    fk::DevBuf<int32_t> rowptr, colidx, diag, slot;  // slot: 16 per element (colour order)
    // coefficients and values
    fk::DevBuf<double> coef;          // per element (colour order)
    bool have_coef = false;
    fk::DevBuf<double> dmu;           // d mu / d|H| per element (colour order), nonlinear loop
    fk::DevBuf<double> hn_prev;       // |H| of the previous step (colour order), tabulated-law loop
    int inner_precision = FK_PREC_F64; // solves inside the mu(|H|) loops, Neumann and demag
    fk::DevBuf<uint8_t> mask_free;    // all-zero mask: the FP32 solver with every node free
    fk::DevBuf<double> inc64;         // FP64 increment of an FP32-refined update solve
    fk::DevBuf<double> demag_u;       // potential of the last demagnetizing-field call
    bool have_demag_u = false;
    // buffers of the demagnetizing-field call, kept across calls
    fk::DevBuf<uint8_t> demag_in_caller, demag_inside;
    fk::DevBuf<double> demag_M, demag_vol, demag_Hd;
    fk::DevBuf<double> lag;           // Lagrange vector C_i = sum VE/4 (geometry only)
    bool have_lag = false;
    fk::DevBuf<uint8_t> region;       // nonlinear material region (colour order)
    fk::DevBuf<uint8_t> target;       // objective region of the adjoint (colour order)
    fk::DevBuf<double> val64;
    fk::DevBuf<float> val32;
    bool have_values = false, have_val32 = false;
    int64_t values_gen = 0;           // bumped whenever val64 changes or is dropped
    // Dirichlet data
    fk::DevBuf<uint8_t> mask;
    fk::DevBuf<double> g;
    bool have_dirichlet = false;
    int64_t nfixed = 0;
    bool dinv64_valid = false;        // dinv buffers match current values + mask
    bool dinv32_valid = false;
    // potential and residual
    fk::DevBuf<double> phi, r64;
    bool have_phi = false;
    fk::DevBuf<double> scratch_nv;    // nv doubles
    fk::DevBuf<double> scratch_nt;    // per element scratch
    fk::DevBuf<double> part_misc;     // block partials of field/energy/residual kernels
    fk::HostBuf<fk::CgState> hstate;  // pinned copies of the CG state
    fk::HostBuf<double> hpart;        // pinned copies of partials
    fk::DevBuf<fk::RuState> ru;       // FP32 solve replacement state
    fk::HostBuf<fk::RuState> hru;     // pinned copies (2, polled alternately)

    // SELL-64 copy of the matrix for the solver (CSR stays the storage format)
    int format = FK_SPMV_AUTO;        // requested solver format (FK_SPMV_*)
    bool have_sell = false, sell64_valid = false, sell32_valid = false;
    int64_t nslices = 0, sell_len = 0, sell_chunk = 0;
    int64_t nshared = 0;              // slices in the shared-offset form
    int32_t sell_max_groups = 0;      // widest slice, in groups of 4 entries
    fk::DevBuf<int64_t> sell_ptr, sell_cptr;
    fk::DevBuf<int32_t> sell_width, sell_col, sell_map, sell_soff;
    fk::DevBuf<int16_t> sell_dcol;    // int16 column offsets from the row (when sell_short)
    fk::DevBuf<double> sell_diag64;   // diagonal of the FP64 standard-form SpMV
    // stencil copy (symmetric diagonal storage) for patterns whose column
    // offsets are one set of at most 16 +-offsets: column k holds
    // a(i, i + off[k]) for off[k] > 0; a(i, i - off[k]) is read from row i - off[k]
    int kind = -1;                    // resolved solver format (FK_SPMV_*), -1 not resolved
    bool stencil_built = false, stencil_ok = false;
    int32_t st_K = 0;
    int32_t st_off[16] = {};
    int64_t st_ld = 0;
    fk::DevBuf<int32_t> st_map;       // K x ld: CSR position or -1
    fk::DevBuf<double> st_val64;
    fk::DevBuf<float> st_val32;
    bool sell_short = false, allow_short = true;
    int32_t bandwidth = 0;
    fk::DevBuf<double> sell_val64;
    fk::DevBuf<float> sell_val32;

    // row-gather assembly: per 64-row slice, transposed incidence lists
    // (element << 2 | local index, caller element order) and the element's 4
    // positions in the row; connectivity and coefficient in caller order
    bool gather_built = false, gather_ok = false;
    int32_t gather_lds = 0;           // largest block value segment (doubles)
    fk::DevBuf<double4> xyzw;         // row-gather node records (x, y, z, w = phi for the Newton step)
    fk::DevBuf<int64_t> inc_sptr;
    fk::DevBuf<int32_t> inc_width;
    fk::DevBuf<uint2> inc;            // row-gather incidences (element << 2 | local, 4 node positions)
    fk::DevBuf<int4> tet_nat;         // elements in caller order (uploaded with the mesh)
    fk::DevBuf<double> coef_nat;      // coefficient in caller order
    fk::DevBuf<double> dmu_nat;       // dmu in the row-gather kernels' element order
    bool coef_nat_valid = false;
    // The row-gather kernels' element order: the caller's when its elements
    // are already sorted by their lowest (internal) node, else that sort
    // (gather_sorted): tet_g, perm_g (colour position -> sorted position)
    // and coef_g (the coefficient in that order) exist only then.
    bool gather_sorted = false;
    fk::DevBuf<int4> tet_g;
    fk::DevBuf<int32_t> perm_g;
    fk::DevBuf<double> coef_g;
    bool coef_g_valid = false;

    // internal node numbering (reorder.hip): internal node k is the caller's
    // node node_perm[k], node_inv is the inverse; empty unless reordered
    bool reordered = false;
    std::vector<int32_t> node_perm, node_inv;
    fk::DevBuf<int32_t> d_node_perm;   // node_perm on the device
    fk::DevBuf<double> phi_internal;  // fk_field_noalloc's potential in the internal order
    double mean_dist_caller = 0.0, mean_dist_internal = 0.0;
    // the CSR pattern in the caller's numbering, built on the first export:
    // caller_map[q] = internal CSR position of caller position q
    bool caller_csr_built = false;
    std::vector<int32_t> caller_rowptr, caller_colidx, caller_map;

    int lanes = 16;                   // lanes per CSR row in SpMV
    int sell_block = 256;             // threads per block of the FP32 SELL SpMV
    int sell_unroll = 2;              // slices in flight per wave in the FP32 SELL SpMV
    int nblocks = 0;                  // persistent grid of the solver kernels
    int64_t chunk = 0;                // rows per logical block

    fk::PcgWork<double> w64;
    fk::PcgWork<float> w32;
};

namespace fk {

// ---- helpers implemented in the .hip files --------------------------------
void set_device(fk_context *ctx);
// per-device address of the allocation-free entry points' static scratch
// (call with the device current)
void noalloc_setup(int device);
void build_mesh(fk_mesh *m, const double *p, const int32_t *t, int32_t base, const double *VE);
void element_volumes(fk_mesh *m, double *VE_host);
void assemble(fk_mesh *m, int method);
bool assemble_newton_gather(fk_mesh *m, const double *mu, const double *dmu, const double *phi, int newton,
                            double *r);
void make_val32(fk_mesh *m);
void local_stiffness(fk_mesh *m, const double *f, double *Ak_host);
void tangential(fk_mesh *m, const double *H, const double *dmu, double *nzval_host);
void lagrange(fk_mesh *m, double *C_host);
void upload_coefficient(fk_mesh *m, const double *f);
const double *coef_caller_order(fk_mesh *m, bool may_allocate);
void set_dirichlet(fk_mesh *m, const uint8_t *mask, const double *g);
void solve(fk_mesh *m, int precision, double rtol, int64_t max_iter, double inner_rtol, int max_outer,
           int warm, int64_t *iters, int32_t *outer, double *relres);
void solve_rhs(fk_mesh *m, const double *b, double rtol, int64_t max_iter, double *x, int64_t *iters,
               double *relres);
void field(fk_mesh *m, double *H, double *Hn, double *energy);
void boundary_integral(fk_mesh *m, int64_t ns, const int32_t *sT, const double *nrm, const double *AE,
                       const double *F, const int32_t *ids, int nids, int32_t base, double *rhs_host);
void setup_partition(fk_mesh *m);
void ensure_work_f64(fk_mesh *m);
void compute_dinv_f64(fk_mesh *m);
void compute_dinv_f32(fk_mesh *m);
// call whenever val64 changed: drops every derived copy (FP32, SELL, D^-1)
inline void values_changed(fk_mesh *m) {
    m->have_values = true;
    m->values_gen += 1;
    m->have_val32 = false;
    m->dinv64_valid = false;
    m->dinv32_valid = false;
    m->sell64_valid = false;
    m->sell32_valid = false;
}
CgState run_pcg_f64(fk_mesh *m, double rtol, int64_t max_iter, double bb_ref);
int solve_f32_increment(fk_mesh *m, double *u, const double *b, double rtol, int64_t max_iter, int64_t *iters);
int solve_f32_free(fk_mesh *m, double *u, const double *b, double rtol, int64_t max_iter, double bb_ref,
                   int64_t *iters, double *rr_out, double *bb_out);
struct NeumannResult {
    double multiplier;
    int64_t iters;
    double relres;
};
// the projected Neumann solve of b in w64.r (device); u returned in w64.x;
// u0: start vector (device) or nullptr for a start from zero
NeumannResult solve_neumann_dev(fk_mesh *m, double rtol, int64_t max_iter, const double *u0);
void phi_axpy_f64(fk_mesh *m, const double *x, double s);
void k_init_phi_dev(fk_mesh *m, int keep_free);

// ---- node numbering (reorder.hip) ------------------------------------------
void validate_input(const double *p, const int32_t *t, int64_t nv, int64_t nt, int32_t base);
// empty: keep the caller's numbering; else perm[k] = caller node at internal k
std::vector<int32_t> choose_numbering(const int32_t *t, int64_t nt, int64_t nv, int32_t base, int mode,
                                      double *mean_caller, double *mean_internal);
void build_caller_csr(fk_mesh *m);

// Per-node host arrays (comps values per node, node-major) between the
// caller's node order and the internal one. NodesIn.ptr is the array in the
// internal order (the caller's own array when the mesh is not renumbered).
template <typename T>
struct NodesIn {
    std::vector<T> buf;
    const T *ptr;
    NodesIn(const fk_mesh *m, const T *caller, int comps = 1) : ptr(caller) {
        if (!m->reordered || !caller) return;
        buf.resize(static_cast<size_t>(m->nv) * comps);
        const int32_t *perm = m->node_perm.data();
        for (int64_t k = 0; k < m->nv; ++k)
            for (int c = 0; c < comps; ++c) buf[k * comps + c] = caller[static_cast<int64_t>(perm[k]) * comps + c];
        ptr = buf.data();
    }
};
// NodesOut.ptr receives the internal order; commit() writes the caller's
// array (call after the device copy into ptr has completed).
template <typename T>
struct NodesOut {
    const fk_mesh *m;
    T *caller;
    int comps;
    std::vector<T> buf;
    T *ptr;
    NodesOut(const fk_mesh *mesh, T *out, int c = 1) : m(mesh), caller(out), comps(c), ptr(out) {
        if (!m->reordered || !out) return;
        buf.resize(static_cast<size_t>(m->nv) * comps);
        ptr = buf.data();
    }
    void commit() {
        if (!m->reordered || !caller) return;
        const int32_t *perm = m->node_perm.data();
        for (int64_t k = 0; k < m->nv; ++k)
            for (int c = 0; c < comps; ++c) caller[static_cast<int64_t>(perm[k]) * comps + c] = buf[k * comps + c];
    }
};
// values of the internal CSR order -> the caller's CSR order (in place)
template <typename T>
void values_to_caller(fk_mesh *m, T *vals) {
    if (!m->reordered) return;
    build_caller_csr(m);
    std::vector<T> tmp(vals, vals + m->nnz);
    const int32_t *map = m->caller_map.data();
    for (int64_t q = 0; q < m->nnz; ++q) vals[q] = tmp[map[q]];
}

// Run body; map library errors to a status and remember the message.
template <typename F>
int guard(fk_context *ctx, const char *where, F &&body) {
    try {
        body();
        if (ctx) ctx->err.clear();
        return FK_OK;
    } catch (const Error &e) {
        if (ctx) ctx->err = std::string(where) + ": " + e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->err = std::string(where) + ": host allocation failed";
        return FK_ERR_OOM;
    } catch (...) {
        if (ctx) ctx->err = std::string(where) + ": unexpected error";
        return FK_ERR_HIP;
    }
}

// launch helpers
inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    return static_cast<unsigned>(g < 1 ? 1 : g);
}

}  // namespace fk
