// Internal node numbering of a mesh.
//
// The solver's matrix copies and the row-gather kernels read node data at
// the columns of each row, so their cache reuse depends on how far apart the
// numbers of neighbouring nodes are. Meshes from a mesh generator arrive in
// its numbering, which need not keep neighbours close. fk_mesh_create
// therefore renumbers the nodes internally (reverse Cuthill-McKee over the
// node graph: George & Liu, "Computer Solution of Large Sparse Positive
// Definite Systems", 1981, ch. 4, with a pseudo-peripheral start node) unless
// the caller's numbering is a stencil numbering (a structured mesh, whose
// pattern the solver streams as a stencil) or already as compact as RCM's.
// Every per-node array crossing the C ABI is mapped between the caller's
// order and the internal one, and exported matrices are returned in the
// caller's numbering, so the renumbering is invisible to the caller.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <thread>
#include <unordered_set>

#include "internal.h"

namespace fk {

namespace {

template <typename F>
void host_parallel(int64_t n, F f) {
    int nt = static_cast<int>(std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency())));
    if (n < 65536) nt = 1;
    if (nt == 1) {
        f(int64_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t) th.emplace_back([&f, n, nt, t] { f(n * t / nt, n * (t + 1) / nt); });
    for (auto &x : th) x.join();
}

}  // namespace

// Node adjacency (nodes sharing an element, no self loops), CSR with sorted
// neighbour lists.
void node_graph(const int32_t *t, int64_t nt, int64_t nv, int32_t base, std::vector<int64_t> &ptr,
                std::vector<int32_t> &adj) {
    // node -> incident elements
    std::vector<int64_t> iptr(static_cast<size_t>(nv + 1), 0);
    for (int64_t e = 0; e < nt; ++e)
        for (int i = 0; i < 4; ++i) iptr[t[4 * e + i] - base + 1]++;
    for (int64_t v = 0; v < nv; ++v) iptr[v + 1] += iptr[v];
    std::vector<int32_t> inc(static_cast<size_t>(4 * nt));
    {
        std::vector<int64_t> cur(iptr.begin(), iptr.end() - 1);
        for (int64_t e = 0; e < nt; ++e)
            for (int i = 0; i < 4; ++i) inc[cur[t[4 * e + i] - base]++] = static_cast<int32_t>(e);
    }
    auto neighbours = [&](int64_t v, std::vector<int32_t> &out) {
        out.clear();
        for (int64_t q = iptr[v]; q < iptr[v + 1]; ++q) {
            const int32_t *nd = t + 4 * static_cast<int64_t>(inc[q]);
            for (int i = 0; i < 4; ++i)
                if (nd[i] - base != v) out.push_back(nd[i] - base);
        }
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    };
    ptr.assign(static_cast<size_t>(nv + 1), 0);
    host_parallel(nv, [&](int64_t a, int64_t b) {
        std::vector<int32_t> tmp;
        for (int64_t v = a; v < b; ++v) {
            neighbours(v, tmp);
            ptr[v + 1] = static_cast<int64_t>(tmp.size());
        }
    });
    for (int64_t v = 0; v < nv; ++v) ptr[v + 1] += ptr[v];
    adj.resize(static_cast<size_t>(ptr[nv]));
    host_parallel(nv, [&](int64_t a, int64_t b) {
        std::vector<int32_t> tmp;
        for (int64_t v = a; v < b; ++v) {
            neighbours(v, tmp);
            std::copy(tmp.begin(), tmp.end(), adj.begin() + ptr[v]);
        }
    });
}

namespace {

// Breadth-first level structure from `root` over the unvisited part of its
// component: returns the nodes in BFS order, `depth` the number of levels.
int64_t bfs_levels(const std::vector<int64_t> &ptr, const std::vector<int32_t> &adj, int32_t root,
                   std::vector<int32_t> &level, std::vector<int32_t> &order, int32_t &last_level_start) {
    order.clear();
    order.push_back(root);
    level[root] = 0;
    size_t head = 0;
    int32_t depth = 0;
    last_level_start = 0;
    while (head < order.size()) {
        const int32_t u = order[head++];
        if (level[u] > depth) {
            depth = level[u];
            last_level_start = static_cast<int32_t>(head - 1);
        }
        for (int64_t q = ptr[u]; q < ptr[u + 1]; ++q) {
            const int32_t w = adj[q];
            if (level[w] < 0) {
                level[w] = level[u] + 1;
                order.push_back(w);
            }
        }
    }
    return depth;
}

}  // namespace

// perm[k] = the caller's node (0-based) at internal position k
std::vector<int32_t> rcm_permutation(const std::vector<int64_t> &ptr, const std::vector<int32_t> &adj, int64_t nv) {
    std::vector<int32_t> order;
    order.reserve(static_cast<size_t>(nv));
    std::vector<int32_t> level(static_cast<size_t>(nv), -1);  // BFS scratch, reset per search
    std::vector<uint8_t> done(static_cast<size_t>(nv), 0);
    std::vector<int32_t> isolated, bfs, bfs2, nb;
    auto deg = [&](int32_t v) { return ptr[v + 1] - ptr[v]; };
    for (int32_t seed = 0; seed < nv; ++seed) {
        if (done[seed]) continue;
        if (deg(seed) == 0) {
            done[seed] = 1;
            isolated.push_back(seed);
            continue;
        }
        // pseudo-peripheral node of seed's component: restart from a node of
        // the last BFS level with the smallest degree while the depth grows;
        // `bfs` always holds the search whose levels are set
        int32_t root = seed, l0 = 0;
        int64_t depth = bfs_levels(ptr, adj, root, level, bfs, l0);
        for (int it = 0; it < 4; ++it) {
            int32_t cand = bfs[l0];
            for (size_t k = l0; k < bfs.size(); ++k)
                if (deg(bfs[k]) < deg(cand)) cand = bfs[k];
            for (int32_t v : bfs) level[v] = -1;
            int32_t l2 = 0;
            const int64_t d2 = bfs_levels(ptr, adj, cand, level, bfs2, l2);
            bfs.swap(bfs2);
            if (d2 <= depth) break;
            root = cand;
            depth = d2;
            l0 = l2;
        }
        for (int32_t v : bfs) level[v] = -1;
        // Cuthill-McKee: breadth first from root, each node's unvisited
        // neighbours by increasing degree (ties by number)
        const size_t start = order.size();
        order.push_back(root);
        done[root] = 1;
        for (size_t head = start; head < order.size(); ++head) {
            const int32_t u = order[head];
            nb.clear();
            for (int64_t q = ptr[u]; q < ptr[u + 1]; ++q)
                if (!done[adj[q]]) {
                    done[adj[q]] = 1;
                    nb.push_back(adj[q]);
                }
            std::sort(nb.begin(), nb.end(), [&](int32_t a, int32_t b) {
                return deg(a) != deg(b) ? deg(a) < deg(b) : a < b;
            });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    order.insert(order.end(), isolated.begin(), isolated.end());
    return order;
}

// mean |i - j| over the graph's edges with node i numbered num[i] (identity
// when num is empty)
static double mean_distance(const std::vector<int64_t> &ptr, const std::vector<int32_t> &adj, int64_t nv,
                            const std::vector<int32_t> &num) {
    double s = 0.0;
    for (int64_t v = 0; v < nv; ++v) {
        const int64_t a = num.empty() ? v : num[v];
        for (int64_t q = ptr[v]; q < ptr[v + 1]; ++q) {
            const int64_t b = num.empty() ? adj[q] : num[adj[q]];
            s += static_cast<double>(a > b ? a - b : b - a);
        }
    }
    return ptr[nv] > 0 ? s / static_cast<double>(ptr[nv]) : 0.0;
}

// Stencil numbering: the positive node-number differences over all element
// node pairs form at most 16 values (a structured mesh numbered by its grid).
static bool stencil_numbering(const int32_t *t, int64_t nt) {
    std::unordered_set<int32_t> offs;
    for (int64_t e = 0; e < nt; ++e) {
        const int32_t *nd = t + 4 * e;
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) {
                const int32_t d = nd[a] > nd[b] ? nd[a] - nd[b] : nd[b] - nd[a];
                if (offs.insert(d).second && offs.size() > 16) return false;
            }
    }
    return true;
}

std::vector<int32_t> choose_numbering(const int32_t *t, int64_t nt, int64_t nv, int32_t base, int mode,
                                      double *mean_caller, double *mean_internal) {
    *mean_caller = *mean_internal = 0.0;
    if (mode == FK_REORDER_NONE) return {};
    if (mode == FK_REORDER_AUTO && stencil_numbering(t, nt)) return {};
    std::vector<int64_t> ptr;
    std::vector<int32_t> adj;
    node_graph(t, nt, nv, base, ptr, adj);
    std::vector<int32_t> perm = rcm_permutation(ptr, adj, nv);
    std::vector<int32_t> inv(static_cast<size_t>(nv));
    for (int64_t k = 0; k < nv; ++k) inv[perm[k]] = static_cast<int32_t>(k);
    *mean_caller = mean_distance(ptr, adj, nv, {});
    *mean_internal = mean_distance(ptr, adj, nv, inv);
    // AUTO keeps a numbering that is already about as compact
    if (mode == FK_REORDER_AUTO && !(*mean_internal < 0.5 * *mean_caller)) return {};
    return perm;
}

// ---- caller-order CSR of a renumbered mesh ----------------------------------
// Caller row i is internal row inv[i] with its columns mapped back and sorted;
// map[q] is the internal CSR position of caller position q.
void build_caller_csr(fk_mesh *m) {
    if (!m->reordered || m->caller_csr_built) return;
    const int64_t nv = m->nv, nnz = m->nnz;
    hipStream_t st = m->ctx->stream;
    std::vector<int32_t> rp(static_cast<size_t>(nv + 1)), ci(static_cast<size_t>(nnz));
    FK_HIP(hipMemcpyAsync(rp.data(), m->rowptr.get(), (nv + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipMemcpyAsync(ci.data(), m->colidx.get(), nnz * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FK_HIP(hipStreamSynchronize(st));
    const std::vector<int32_t> &perm = m->node_perm, &inv = m->node_inv;
    std::vector<int32_t> &crp = m->caller_rowptr, &cci = m->caller_colidx, &cmap = m->caller_map;
    crp.assign(static_cast<size_t>(nv + 1), 0);
    for (int64_t i = 0; i < nv; ++i) crp[i + 1] = crp[i] + (rp[inv[i] + 1] - rp[inv[i]]);
    cci.resize(static_cast<size_t>(nnz));
    cmap.resize(static_cast<size_t>(nnz));
    host_parallel(nv, [&](int64_t a, int64_t b) {
        std::vector<std::pair<int32_t, int32_t>> row;
        for (int64_t i = a; i < b; ++i) {
            const int32_t k = inv[i];
            row.clear();
            for (int32_t q = rp[k]; q < rp[k + 1]; ++q) row.emplace_back(perm[ci[q]], q);
            std::sort(row.begin(), row.end());
            int32_t o = crp[i];
            for (const auto &c : row) {
                cci[o] = c.first;
                cmap[o] = c.second;
                ++o;
            }
        }
    });
    m->caller_csr_built = true;
}

}  // namespace fk

extern "C" {

int fk_rcm_order(const int32_t *t, int64_t nv, int64_t nt, int32_t index_base, int32_t *perm) {
    if (!t || !perm || nv < 1 || nt < 0 || (index_base != 0 && index_base != 1) || nv >= (int64_t(1) << 31))
        return FK_ERR_INVALID_ARG;
    for (int64_t q = 0; q < 4 * nt; ++q)
        if (t[q] < index_base || static_cast<int64_t>(t[q]) - index_base >= nv) return FK_ERR_INDEX;
    try {
        std::vector<int64_t> ptr;
        std::vector<int32_t> adj;
        fk::node_graph(t, nt, nv, index_base, ptr, adj);
        const std::vector<int32_t> order = fk::rcm_permutation(ptr, adj, nv);
        std::copy(order.begin(), order.end(), perm);
        return FK_OK;
    } catch (const std::bad_alloc &) {
        return FK_ERR_OOM;
    }
}

int fk_mesh_numbering(const fk_mesh *m, int32_t *reordered, double *mean_distance_caller,
                      double *mean_distance_internal) {
    if (!m) return FK_ERR_INVALID_ARG;
    if (reordered) *reordered = m->reordered ? 1 : 0;
    if (mean_distance_caller) *mean_distance_caller = m->mean_dist_caller;
    if (mean_distance_internal) *mean_distance_internal = m->mean_dist_internal;
    return FK_OK;
}

}  // extern "C"
