// Host-side graph construction: synthetic generators, CSR build (symmetrise,
// self-loop removal, dedup), CSC transpose and per-edge weights.
//
// The reference builds its host CSR in utils::traversal::Context
// (include/utils/graphs/traversal.h:99-200) and its CSC with the serial
// counting transpose csr_tocsc (include/groute/graphs/csr_graph.h:330-397);
// dedup and self-loop removal live in the offline gtools
// (samples/tools/gtools.cpp:56,135). Here all of these are parallel over
// host threads and deterministic: every output depends only on the inputs and
// seeds, never on the thread count.
#include "../../include/sepgraph.h"
#include "sg_host.hpp"

#include <cstring>
#include <memory>
#include <new>

namespace sg {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
const char *get_error() { return g_err.c_str(); }

}  // namespace sg

using namespace sg;

#define SG_FAIL(code, msg)      \
    do {                        \
        sg::set_error(msg);     \
        return (code);          \
    } while (0)

extern "C" const char *sg_last_error(void) { return sg::get_error(); }
extern "C" int sg_version(void) { return 100; }

extern "C" int sg_gen_rmat_edges(int scale, int edge_factor, double a, double b, double c,
                                 uint64_t seed, uint32_t *src, uint32_t *dst)
{
    if (scale < 1 || scale > 31 || edge_factor < 1 || !src || !dst)
        SG_FAIL(SG_ERR_INVALID, "sg_gen_rmat_edges: scale must be in [1,31], edge_factor >= 1");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        SG_FAIL(SG_ERR_INVALID, "sg_gen_rmat_edges: probabilities must be >= 0 with a+b+c <= 1");
    uint64_t count = (uint64_t)edge_factor << scale;
    const double ab = a + b, abc = a + b + c;
    parallel_for(count, [&](uint64_t beg, uint64_t end, int) {
        for (uint64_t i = beg; i < end; i++) {
            uint32_t u = 0, v = 0;
            uint64_t k = i * (uint64_t)scale;
            for (int l = 0; l < scale; l++) {
                double r = unit(draw(seed, k + l));
                uint32_t bit = 1u << (scale - 1 - l);
                if (r < a) {
                } else if (r < ab) {
                    v |= bit;
                } else if (r < abc) {
                    u |= bit;
                } else {
                    u |= bit;
                    v |= bit;
                }
            }
            src[i] = u;
            dst[i] = v;
        }
    });
    return SG_OK;
}

extern "C" int sg_gen_grid_edges(uint32_t rows, uint32_t cols, double p_delete, uint64_t seed,
                                 uint32_t *src, uint32_t *dst, uint64_t *count)
{
    if (!rows || !cols || !src || !dst || !count)
        SG_FAIL(SG_ERR_INVALID, "sg_gen_grid_edges: empty grid or null output");
    if ((uint64_t)rows * cols >= (1ull << 31)) SG_FAIL(SG_ERR_RANGE, "sg_gen_grid_edges: grid too large");
    const uint64_t nv = (uint64_t)rows * cols;
    // Pass 1: per-row kept-edge counts; pass 2: write at row offsets, so the
    // output order is fixed (row-major, horizontal before vertical).
    std::vector<uint64_t> per_row(rows + 1, 0);
    auto kept = [&](uint64_t id) { return unit(draw(seed, id)) >= p_delete; };
    parallel_for(rows, [&](uint64_t rb, uint64_t re, int) {
        for (uint64_t r = rb; r < re; r++) {
            uint64_t k = 0;
            for (uint64_t cc = 0; cc < cols; cc++) {
                uint64_t id = 2 * (r * cols + cc);
                if (cc + 1 < cols && kept(id)) k++;
                if (r + 1 < rows && kept(id + 1)) k++;
            }
            per_row[r + 1] = k;
        }
    });
    for (uint32_t r = 0; r < rows; r++) per_row[r + 1] += per_row[r];
    parallel_for(rows, [&](uint64_t rb, uint64_t re, int) {
        for (uint64_t r = rb; r < re; r++) {
            uint64_t pos = per_row[r];
            for (uint64_t cc = 0; cc < cols; cc++) {
                uint64_t vtx = r * cols + cc, id = 2 * vtx;
                if (cc + 1 < cols && kept(id)) {
                    src[pos] = (uint32_t)vtx;
                    dst[pos] = (uint32_t)(vtx + 1);
                    pos++;
                }
                if (r + 1 < rows && kept(id + 1)) {
                    src[pos] = (uint32_t)vtx;
                    dst[pos] = (uint32_t)(vtx + cols);
                    pos++;
                }
            }
        }
    });
    (void)nv;
    *count = per_row[rows];
    return SG_OK;
}

extern "C" int sg_build_csr(uint32_t n, uint64_t count, const uint32_t *src, const uint32_t *dst,
                            int flags, uint32_t *row_start, uint32_t *edge_dst, uint64_t cap,
                            uint64_t *m_out)
{
    if (!row_start || !m_out || (count && (!src || !dst)))
        SG_FAIL(SG_ERR_INVALID, "sg_build_csr: null argument");
    const bool sym = flags & SG_CSR_SYMMETRIZE, noloop = flags & SG_CSR_DROP_SELF_LOOPS,
               dedup = flags & SG_CSR_DEDUP;
    // Degree count with relaxed atomics; validates ids on the way.
    std::unique_ptr<uint32_t[]> deg(new (std::nothrow) uint32_t[(size_t)n + 1]);
    if (!deg) SG_FAIL(SG_ERR_NOMEM, "sg_build_csr: host allocation failed");
    std::memset(deg.get(), 0, ((size_t)n + 1) * sizeof(uint32_t));
    std::atomic<int> bad{0};
    parallel_for(count, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t i = b; i < e; i++) {
            uint32_t u = src[i], v = dst[i];
            if (u >= n || v >= n) {
                bad.store(1, std::memory_order_relaxed);
                continue;
            }
            if (noloop && u == v) continue;
            __atomic_fetch_add(&deg[u], 1u, __ATOMIC_RELAXED);
            if (sym) __atomic_fetch_add(&deg[v], 1u, __ATOMIC_RELAXED);
        }
    });
    if (bad.load()) SG_FAIL(SG_ERR_RANGE, "sg_build_csr: edge endpoint >= n");
    // Exclusive scan into a 64-bit temporary offset array.
    std::unique_ptr<uint64_t[]> off(new (std::nothrow) uint64_t[(size_t)n + 1]);
    if (!off) SG_FAIL(SG_ERR_NOMEM, "sg_build_csr: host allocation failed");
    off[0] = 0;
    for (uint32_t v = 0; v < n; v++) off[v + 1] = off[v] + deg[v];
    const uint64_t total = off[n];
    std::unique_ptr<uint32_t[]> tmp(new (std::nothrow) uint32_t[total ? total : 1]);
    std::unique_ptr<uint64_t[]> cur(new (std::nothrow) uint64_t[(size_t)n + 1]);
    if (!tmp || !cur) SG_FAIL(SG_ERR_NOMEM, "sg_build_csr: host allocation failed");
    std::memcpy(cur.get(), off.get(), ((size_t)n + 1) * sizeof(uint64_t));
    parallel_for(count, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t i = b; i < e; i++) {
            uint32_t u = src[i], v = dst[i];
            if (noloop && u == v) continue;
            tmp[__atomic_fetch_add(&cur[u], 1ull, __ATOMIC_RELAXED)] = v;
            if (sym) tmp[__atomic_fetch_add(&cur[v], 1ull, __ATOMIC_RELAXED)] = u;
        }
    });
    // Sort each row (and dedup); rows are independent.
    parallel_for_dynamic(n, 4096, [&](uint64_t rb, uint64_t re) {
        for (uint64_t v = rb; v < re; v++) {
            uint32_t *b = tmp.get() + off[v], *e = tmp.get() + off[v + 1];
            std::sort(b, e);
            uint32_t k = (uint32_t)(e - b);
            if (dedup) k = (uint32_t)(std::unique(b, e) - b);
            deg[v] = k;
        }
    });
    uint64_t m = 0;
    for (uint32_t v = 0; v < n; v++) m += deg[v];
    if (m > 0xFFFFFFFFull) SG_FAIL(SG_ERR_RANGE, "sg_build_csr: more than 2^32-1 edges");
    if (m > cap) SG_FAIL(SG_ERR_CAPACITY, "sg_build_csr: edge_dst capacity too small");
    row_start[0] = 0;
    for (uint32_t v = 0; v < n; v++) row_start[v + 1] = row_start[v] + deg[v];
    parallel_for_dynamic(n, 4096, [&](uint64_t rb, uint64_t re) {
        for (uint64_t v = rb; v < re; v++)
            std::memcpy(edge_dst + row_start[v], tmp.get() + off[v], (size_t)deg[v] * sizeof(uint32_t));
    });
    *m_out = m;
    return SG_OK;
}

extern "C" int sg_gen_weights(uint32_t n, const uint32_t *row_start, const uint32_t *edge_dst,
                              uint32_t lo, uint32_t hi, uint64_t seed, int symmetric,
                              uint32_t *weights)
{
    if (!row_start || !weights || hi <= lo) SG_FAIL(SG_ERR_INVALID, "sg_gen_weights: need lo < hi");
    const uint32_t range = hi - lo;
    parallel_for_dynamic(n, 4096, [&](uint64_t rb, uint64_t re) {
        for (uint64_t u = rb; u < re; u++) {
            for (uint32_t e = row_start[u]; e < row_start[u + 1]; e++) {
                uint64_t v = edge_dst[e], a = u, b = v;
                if (symmetric && a > b) std::swap(a, b);
                weights[e] = lo + below(draw(seed, (a << 32) | b), range);
            }
        }
    });
    return SG_OK;
}

extern "C" int sg_csc_from_csr(uint32_t n, const uint32_t *row_start, const uint32_t *edge_dst,
                               const uint32_t *weights, uint32_t *col_start, uint32_t *edge_src,
                               uint32_t *csc_weights)
{
    if (!row_start || !col_start || !edge_src) SG_FAIL(SG_ERR_INVALID, "sg_csc_from_csr: null argument");
    const uint64_t m = row_start[n];
    std::unique_ptr<uint32_t[]> cnt(new (std::nothrow) uint32_t[(size_t)n + 1]);
    if (!cnt) SG_FAIL(SG_ERR_NOMEM, "sg_csc_from_csr: host allocation failed");
    std::memset(cnt.get(), 0, ((size_t)n + 1) * sizeof(uint32_t));
    std::atomic<int> bad{0};
    parallel_for(m, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t i = b; i < e; i++) {
            if (edge_dst[i] >= n) {
                bad.store(1);
                continue;
            }
            __atomic_fetch_add(&cnt[edge_dst[i]], 1u, __ATOMIC_RELAXED);
        }
    });
    if (bad.load()) SG_FAIL(SG_ERR_RANGE, "sg_csc_from_csr: edge endpoint >= n");
    col_start[0] = 0;
    for (uint32_t v = 0; v < n; v++) col_start[v + 1] = col_start[v] + cnt[v];
    // Scatter rows in parallel (atomic cursors), then sort each column by
    // source so the order matches the serial transpose exactly.
    std::unique_ptr<uint32_t[]> cur(new (std::nothrow) uint32_t[(size_t)n + 1]);
    std::unique_ptr<uint64_t[]> pairs;
    if (!cur) SG_FAIL(SG_ERR_NOMEM, "sg_csc_from_csr: host allocation failed");
    std::memcpy(cur.get(), col_start, ((size_t)n + 1) * sizeof(uint32_t));
    const bool wt = weights && csc_weights;
    if (wt) {
        pairs.reset(new (std::nothrow) uint64_t[m ? m : 1]);
        if (!pairs) SG_FAIL(SG_ERR_NOMEM, "sg_csc_from_csr: host allocation failed");
    }
    parallel_for_dynamic(n, 4096, [&](uint64_t rb, uint64_t re) {
        for (uint64_t u = rb; u < re; u++) {
            for (uint32_t e = row_start[u]; e < row_start[u + 1]; e++) {
                uint32_t pos = __atomic_fetch_add(&cur[edge_dst[e]], 1u, __ATOMIC_RELAXED);
                if (wt)
                    pairs[pos] = ((uint64_t)u << 32) | weights[e];
                else
                    edge_src[pos] = (uint32_t)u;
            }
        }
    });
    parallel_for_dynamic(n, 4096, [&](uint64_t vb, uint64_t ve) {
        for (uint64_t v = vb; v < ve; v++) {
            uint32_t b = col_start[v], e = col_start[v + 1];
            if (wt) {
                // stable by source: (source, weight) pairs sort by source first;
                // equal sources only occur for duplicate edges, kept in any order
                // with identical keys ordered by weight.
                std::sort(pairs.get() + b, pairs.get() + e);
                for (uint32_t i = b; i < e; i++) {
                    edge_src[i] = (uint32_t)(pairs[i] >> 32);
                    csc_weights[i] = (uint32_t)pairs[i];
                }
            } else {
                std::sort(edge_src + b, edge_src + e);
            }
        }
    });
    return SG_OK;
}

extern "C" int sg_pick_sources(uint32_t n, const uint32_t *row_start, int k, uint64_t seed,
                               uint32_t *out)
{
    if (!row_start || !out || k < 0 || n == 0) SG_FAIL(SG_ERR_INVALID, "sg_pick_sources: bad argument");
    uint64_t nonisolated = 0;
    for (uint32_t v = 0; v < n; v++) nonisolated += row_start[v + 1] > row_start[v];
    if (nonisolated < (uint64_t)k) SG_FAIL(SG_ERR_RANGE, "sg_pick_sources: not enough non-isolated vertices");
    std::vector<uint8_t> taken(n, 0);
    int got = 0;
    for (uint64_t i = 0; got < k; i++) {
        uint32_t v = (uint32_t)(draw(seed, i) % n);
        if (row_start[v + 1] == row_start[v] || taken[v]) continue;
        taken[v] = 1;
        out[got++] = v;
    }
    return SG_OK;
}
