// Device graph upload and engine life cycle (C ABI).
//
// Mirrors Engine::LoadGraph (include/framework/framework.cuh:99-133) and the
// single:: CSR/CSC allocators (include/groute/graphs/csr_graph.h:1596-1783):
// CSR and CSC are mirrored to HBM once and stay resident; an undirected graph
// aliases its CSC to the CSR. Unlike the reference, no call exits the process.
#include "sg_internal.hpp"

#include <cstring>
#include <mutex>
#include <vector>

using namespace sg;

namespace {

// bit v of `inert` is set when vertex v has in-degree 0.
__global__ void k_inert_bitmap(const uint32_t *cstart, uint32_t n, uint32_t words, uint64_t *inert)
{
    const uint32_t w = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (w >= words) return;
    const uint32_t v = w * 64 + lane();
    bool zero = v < n && cstart[v + 1] == cstart[v];
    uint64_t m = ballot(zero);
    if (lane() == 0) inert[w] = m;
}

// Row r of dst (caller's vertex order) from row r of src (device order):
// dst[v] = src[inv[v]]. Reads of inv and writes are sequential (on MI355X
// this gather beats the scatter dst[perm[i]] = src[i], whose partial-line
// random writes cost more than random reads).
__global__ void k_gather_rows(const uint32_t *src, uint32_t *dst, const uint32_t *inv, uint32_t n)
{
    const size_t off = (size_t)blockIdx.y * n;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        dst[off + v] = src[off + inv[v]];
}

// Exact reach statistics of a traversal result in the caller's vertex order:
// vertices whose 32-bit value is not the 0xFFFFFFFF sentinel (-1 level /
// UINT32_MAX distance) and the sum of their out-degrees, read from the device
// CSR through inv (null: the orders agree). out[0], out[1] must be zero on
// entry.
__global__ void k_reach_stats(const uint32_t *value, const uint32_t *row, const uint32_t *inv, uint32_t n,
                              unsigned long long *out)
{
    __shared__ unsigned long long s[2][16];
    unsigned long long cnt = 0, deg = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        if (value[v] != 0xFFFFFFFFu) {
            const uint32_t d = inv ? inv[v] : v;
            cnt++;
            deg += row[d + 1] - row[d];
        }
    }
    cnt = wave_sum_u64(cnt);
    deg = wave_sum_u64(deg);
    const int wid = threadIdx.x / kWave;
    if (lane() == 0) {
        s[0][wid] = cnt;
        s[1][wid] = deg;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x / kWave); i++) {
            cnt += s[0][i];
            deg += s[1][i];
        }
        atomicAdd(&out[0], cnt);
        atomicAdd(&out[1], deg);
    }
}

int dev_alloc(sg_graph *g, void **p, uint64_t bytes)
{
    if (g->nallocs >= 32) SG_FAIL(SG_ERR_STATE, "sg_graph_create: too many allocations");
    hipError_t e = hipMalloc(p, bytes ? bytes : 4);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        sg::set_error(std::string("sg_graph_create: hipMalloc failed: ") + hipGetErrorString(e));
        return SG_ERR_NOMEM;
    }
    g->allocs[g->nallocs++] = *p;
    g->device_bytes += bytes;
    return SG_OK;
}

}  // namespace

extern "C" int sg_device_count(void)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return c;
}

extern "C" int sg_graph_destroy(sg_graph *g)
{
    if (!g) return SG_OK;
    for (int i = 0; i < g->nallocs; i++) (void)hipFree(g->allocs[i]);
    delete g;
    return SG_OK;
}

namespace {

// Degree-ordered relabelling of a CSR: new vertex i is old vertex perm[i],
// ordered by total degree (out + in) descending, ties by id. Rows keep their
// neighbours sorted. The hot vertices of a skewed graph then share a few
// cache lines, so the random per-vertex accesses of every kernel (levels,
// distances, frontier bits, PageRank contributions) mostly hit L2.
void relabel_csr(uint32_t n, const uint32_t *row, const uint32_t *col, const uint32_t *w,
                 std::vector<uint32_t> &perm, std::vector<uint32_t> &inv, std::vector<uint32_t> &nrow,
                 std::vector<uint32_t> &ncol, std::vector<uint32_t> &nw)
{
    const uint64_t m = row[n];
    std::vector<uint32_t> deg(n);
    for (uint32_t v = 0; v < n; v++) deg[v] = row[v + 1] - row[v];
    for (uint64_t e = 0; e < m; e++) deg[col[e]]++;
    perm.resize(n);
    for (uint32_t v = 0; v < n; v++) perm[v] = v;
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return deg[a] > deg[b]; });
    inv.resize(n);
    for (uint32_t i = 0; i < n; i++) inv[perm[i]] = i;
    nrow.assign((size_t)n + 1, 0);
    for (uint32_t i = 0; i < n; i++) nrow[i + 1] = nrow[i] + (row[perm[i] + 1] - row[perm[i]]);
    ncol.resize(m ? m : 1);
    if (w) nw.resize(m ? m : 1);
    sg::parallel_for_dynamic(n, 4096, [&](uint64_t i0, uint64_t i1) {
        std::vector<std::pair<uint32_t, uint32_t>> t;
        for (uint64_t i = i0; i < i1; i++) {
            const uint32_t u = perm[i];
            const uint32_t b = row[u], d = row[u + 1] - b, o = nrow[i];
            if (!w) {
                for (uint32_t j = 0; j < d; j++) ncol[o + j] = inv[col[b + j]];
                std::sort(ncol.begin() + o, ncol.begin() + o + d);
            } else {
                t.resize(d);
                for (uint32_t j = 0; j < d; j++) t[j] = {inv[col[b + j]], w[b + j]};
                std::sort(t.begin(), t.end());
                for (uint32_t j = 0; j < d; j++) {
                    ncol[o + j] = t[j].first;
                    nw[o + j] = t[j].second;
                }
            }
        }
    });
}

// Relabel when the degree distribution is skewed (a hub far above the mean);
// uniform-degree graphs such as grids keep their id order and its locality.
bool want_relabel(uint32_t n, const uint32_t *row, uint32_t flags)
{
    if (flags & SG_GRAPH_NO_RELABEL) return false;
    if (flags & SG_GRAPH_RELABEL) return true;
    const uint64_t m = row[n];
    uint32_t dmax = 0;
    for (uint32_t v = 0; v < n; v++) dmax = std::max(dmax, row[v + 1] - row[v]);
    return n > 4096 && (double)dmax > 64.0 * ((double)m / n + 1.0);
}

}  // namespace

extern "C" int sg_graph_create(uint32_t n, uint64_t m, const uint32_t *row_start, const uint32_t *edge_dst,
                               const uint32_t *weights, const uint32_t *col_start, const uint32_t *edge_src,
                               const uint32_t *csc_weights, int undirected, sg_graph **out)
{
    return sg_graph_create_ex(n, m, row_start, edge_dst, weights, col_start, edge_src, csc_weights,
                              undirected ? SG_GRAPH_UNDIRECTED : 0u, out);
}

extern "C" int sg_graph_create_ex(uint32_t n, uint64_t m, const uint32_t *row_start, const uint32_t *edge_dst,
                                  const uint32_t *weights, const uint32_t *col_start, const uint32_t *edge_src,
                                  const uint32_t *csc_weights, uint32_t flags, sg_graph **out)
{
    const bool undirected = (flags & SG_GRAPH_UNDIRECTED) != 0;
    if (!out) SG_FAIL(SG_ERR_INVALID, "sg_graph_create: null output handle");
    *out = nullptr;
    if (n == 0 || n >= (1u << 31)) SG_FAIL(SG_ERR_RANGE, "sg_graph_create: need 0 < n < 2^31");
    if (m >= (1ull << 32)) SG_FAIL(SG_ERR_RANGE, "sg_graph_create: need m < 2^32");
    if (!row_start || (m && !edge_dst)) SG_FAIL(SG_ERR_INVALID, "sg_graph_create: null CSR array");
    if (row_start[0] != 0 || row_start[n] != m) SG_FAIL(SG_ERR_INVALID, "sg_graph_create: row_start[0] != 0 or row_start[n] != m");
    for (uint32_t v = 0; v < n; v++)
        if (row_start[v + 1] < row_start[v]) SG_FAIL(SG_ERR_INVALID, "sg_graph_create: row_start not monotone");
    for (uint64_t e = 0; e < m; e++)
        if (edge_dst[e] >= n) SG_FAIL(SG_ERR_RANGE, "sg_graph_create: edge_dst entry >= n");
    bool have_csc = col_start != nullptr;
    if (!undirected && have_csc && (!edge_src || (weights && !csc_weights)))
        SG_FAIL(SG_ERR_INVALID, "sg_graph_create: incomplete CSC");
    if (!undirected && have_csc) {
        // a caller's CSC is checked like the CSR before anything reads it,
        // and must give every vertex the in-degree the CSR gives it
        if (col_start[0] != 0 || col_start[n] != m)
            SG_FAIL(SG_ERR_INVALID, "sg_graph_create: col_start[0] != 0 or col_start[n] != m");
        for (uint32_t v = 0; v < n; v++)
            if (col_start[v + 1] < col_start[v]) SG_FAIL(SG_ERR_INVALID, "sg_graph_create: col_start not monotone");
        for (uint64_t e = 0; e < m; e++)
            if (edge_src[e] >= n) SG_FAIL(SG_ERR_RANGE, "sg_graph_create: edge_src entry >= n");
        std::vector<uint32_t> indeg(n, 0u);
        for (uint64_t e = 0; e < m; e++) indeg[edge_dst[e]]++;
        for (uint32_t v = 0; v < n; v++)
            if (indeg[v] != col_start[v + 1] - col_start[v])
                SG_FAIL(SG_ERR_INVALID, "sg_graph_create: CSC in-degrees differ from the CSR's");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        SG_FAIL(SG_ERR_HIP, "sg_graph_create: no HIP device visible");
    }
    SG_HIP(hipSetDevice(0));

    // device-side vertex order (the caller's ids are kept at the boundary)
    std::vector<uint32_t> perm, inv, r_row, r_col, r_w;
    const bool relabel = want_relabel(n, row_start, flags);
    if (relabel) {
        relabel_csr(n, row_start, edge_dst, weights, perm, inv, r_row, r_col, r_w);
        row_start = r_row.data();
        edge_dst = r_col.data();
        weights = weights ? r_w.data() : nullptr;
        have_csc = false;  // rebuilt from the relabelled CSR below
        col_start = edge_src = csc_weights = nullptr;
    }

    // Host CSC when the caller did not pass one (directed graphs only).
    std::vector<uint32_t> h_cs, h_src, h_cw;
    const uint32_t *cs = col_start, *src = edge_src, *cw = csc_weights;
    if (!undirected && !have_csc) {
        h_cs.resize((size_t)n + 1);
        h_src.resize(m ? m : 1);
        if (weights) h_cw.resize(m ? m : 1);
        int rc = sg_csc_from_csr(n, row_start, edge_dst, weights, h_cs.data(), h_src.data(),
                                 weights ? h_cw.data() : nullptr);
        if (rc) return rc;
        cs = h_cs.data();
        src = h_src.data();
        cw = weights ? h_cw.data() : nullptr;
    }

    sg_graph *g = new sg_graph();
    g->undirected = undirected != 0;
    g->weighted = weights != nullptr;
    g->relabeled = relabel;
    g->h_inv = std::move(inv);
    DevGraph &d = g->dev;
    d.n = n;
    d.m = m;
    d.words = (n + 63) / 64;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
        delete g;
        SG_FAIL(SG_ERR_HIP, "sg_graph_create: hipGetDeviceProperties failed");
    }
    g->num_cus = prop.multiProcessorCount;

    auto up = [&](const uint32_t *h, uint64_t count, const uint32_t **dptr) -> int {
        void *p = nullptr;
        int rc = dev_alloc(g, &p, count * sizeof(uint32_t));
        if (rc) return rc;
        if (count) SG_HIP(hipMemcpy(p, h, count * sizeof(uint32_t), hipMemcpyHostToDevice));
        *dptr = (const uint32_t *)p;
        return SG_OK;
    };
    int rc = up(row_start, (uint64_t)n + 1, &d.row);
    if (!rc) rc = up(edge_dst, m, &d.col);
    if (!rc && weights) rc = up(weights, m, &d.w);
    if (!rc) {
        if (undirected) {
            d.cstart = d.row;
            d.csrc = d.col;
            d.cw = d.w;
        } else {
            rc = up(cs, (uint64_t)n + 1, &d.cstart);
            if (!rc) rc = up(src, m, &d.csrc);
            if (!rc && cw) rc = up(cw, m, &d.cw);
        }
    }
    if (!rc) {
        const uint32_t *hcs = undirected ? row_start : cs;
        const uint32_t *hsrc = undirected ? edge_dst : src;
        std::vector<uint32_t> head(n);
        for (uint32_t v = 0; v < n; v++) head[v] = hcs[v + 1] > hcs[v] ? hsrc[hcs[v]] : 0xFFFFFFFFu;
        rc = up(head.data(), n, &d.chead);
        const uint32_t *hcw = undirected ? weights : cw;
        if (!rc && hcw) {
            // reuse `head` for the smallest in-edge weight of each vertex
            sg::parallel_for_dynamic(n, 65536, [&](uint64_t v0, uint64_t v1) {
                for (uint64_t v = v0; v < v1; v++) {
                    uint32_t mw = 0xFFFFFFFFu;
                    for (uint32_t e = hcs[v]; e < hcs[v + 1]; e++) mw = std::min(mw, hcw[e]);
                    head[v] = mw;
                }
            });
            rc = up(head.data(), n, &d.cminw);
        }
    }
    if (!rc && relabel) {
        const uint32_t *p = nullptr;
        rc = up(g->h_inv.data(), n, &p);
        g->d_inv = p;
        if (!rc) rc = up(perm.data(), n, &p);
        g->d_perm = p;
    }
    if (!rc) {
        void *p = nullptr;
        rc = dev_alloc(g, &p, (uint64_t)d.words * 8);
        if (!rc) {
            d.inert = (const uint64_t *)p;
            const int waves_per_block = 4;
            dim3 grid((d.words + waves_per_block - 1) / waves_per_block);
            hipLaunchKernelGGL(k_inert_bitmap, grid, dim3(waves_per_block * kWave), 0, 0, d.cstart, n, d.words,
                               (uint64_t *)p);
            hipError_t e = hipDeviceSynchronize();
            if (e != hipSuccess) {
                sg::set_error(std::string("sg_graph_create: inert bitmap kernel: ") + hipGetErrorString(e));
                rc = SG_ERR_HIP;
            }
        }
    }
    // static chunk lists of high-degree vertices (host side): out-edges from
    // the CSR, in-edges from the CSC (the same list when undirected)
    auto upload2 = [&](const std::vector<uint2> &h, const uint2 **dptr, uint32_t *count) -> int {
        void *p = nullptr;
        int r = dev_alloc(g, &p, (uint64_t)h.size() * sizeof(uint2));
        if (!r && !h.empty()) {
            hipError_t e = hipMemcpy(p, h.data(), h.size() * sizeof(uint2), hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                sg::set_error(std::string("sg_graph_create: chunk list upload: ") + hipGetErrorString(e));
                r = SG_ERR_HIP;
            }
        }
        *dptr = (const uint2 *)p;
        *count = (uint32_t)h.size();
        return r;
    };
    auto chunk_list = [&](const uint32_t *offs, uint32_t threshold, uint32_t chunk, std::vector<uint2> &huge,
                          std::vector<uint2> *hubs) {
        for (uint32_t u = 0; u < n; u++) {
            const uint32_t deg = offs[u + 1] - offs[u];
            if (deg <= threshold) continue;
            if (hubs) hubs->push_back(make_uint2((uint32_t)huge.size(), (deg + chunk - 1) / chunk));
            for (uint32_t c = 0; c < deg; c += chunk) huge.push_back(make_uint2(u, offs[u] + c));
        }
    };
    {
        std::vector<uint2> out_chunks, in_chunks, in_hubs;
        chunk_list(row_start, kHugeDegree, kHugeChunk, out_chunks, nullptr);
        chunk_list(undirected ? row_start : cs, kPullHubDegree, kPullChunk, in_chunks, &in_hubs);
        if (!rc) rc = upload2(out_chunks, &d.huge, &d.nhuge);
        if (!rc) {
            std::vector<uint32_t> ends(out_chunks.size() ? out_chunks.size() : 1);
            for (size_t c = 0; c < out_chunks.size(); c++)
                ends[c] = std::min(row_start[out_chunks[c].x + 1], out_chunks[c].y + kHugeChunk);
            rc = up(ends.data(), out_chunks.size(), &d.huge_end);
        }
        if (!rc) rc = upload2(in_chunks, &d.hugein, &d.nhugein);
        if (!rc) rc = upload2(in_hubs, &d.hubin, &d.nhubin);
        const uint32_t *hcw = undirected ? weights : cw;
        if (!rc && hcw && !in_chunks.empty() && (uint64_t)in_chunks.size() * kPullChunk < 0xFFFFFFFFull) {
            // weighted graphs: every in-hub's in-edges by increasing weight
            // (then source), chunk c at [c * kPullChunk, (c + 1) * kPullChunk),
            // the hub's last chunk padded with 0xFFFFFFFF sources; and each
            // chunk's smallest weight
            const uint32_t *hcs = undirected ? row_start : cs;
            const uint32_t *hsrc = undirected ? edge_dst : src;
            const uint64_t total = (uint64_t)in_chunks.size() * kPullChunk;
            std::vector<uint32_t> hs(total, 0xFFFFFFFFu), hw(total, 0xFFFFFFFFu), hmin(in_chunks.size());
            sg::parallel_for_dynamic(in_hubs.size(), 64, [&](uint64_t h0, uint64_t h1) {
                std::vector<uint64_t> e;
                for (uint64_t h = h0; h < h1; h++) {
                    const uint32_t c0 = in_hubs[h].x, nc = in_hubs[h].y;
                    const uint32_t v = in_chunks[c0].x, b = hcs[v], deg = hcs[v + 1] - b;
                    e.resize(deg);
                    for (uint32_t j = 0; j < deg; j++) e[j] = ((uint64_t)hcw[b + j] << 32) | hsrc[b + j];
                    std::sort(e.begin(), e.end());
                    const uint64_t at = (uint64_t)c0 * kPullChunk;
                    for (uint32_t j = 0; j < deg; j++) {
                        hs[at + j] = (uint32_t)e[j];
                        hw[at + j] = (uint32_t)(e[j] >> 32);
                    }
                    for (uint32_t k = 0; k < nc; k++) hmin[c0 + k] = hw[at + (uint64_t)k * kPullChunk];
                }
            });
            // the order an SSSP pull round visits them in: every hub's first
            // (lightest) chunk, then the others, each with its smallest weight
            std::vector<uint2> ord;
            std::vector<uint32_t> omin;
            ord.reserve(in_chunks.size());
            omin.reserve(in_chunks.size());
            for (const uint2 &h : in_hubs) {
                ord.push_back(make_uint2(in_chunks[h.x].x, h.x));
                omin.push_back(hmin[h.x]);
            }
            for (const uint2 &h : in_hubs)
                for (uint32_t k = 1; k < h.y; k++) {
                    ord.push_back(make_uint2(in_chunks[h.x + k].x, h.x + k));
                    omin.push_back(hmin[h.x + k]);
                }
            uint32_t nord = 0;
            rc = up(hs.data(), total, &d.hub_src);
            if (!rc) rc = up(hw.data(), total, &d.hub_w);
            if (!rc) rc = up(omin.data(), omin.size(), &d.hub_minw);
            if (!rc) rc = upload2(ord, &d.hub_ord, &nord);
        }
    }
    if (!rc) {
        // sliced in-edges (64 vertices per slice, in-hubs left to the chunks)
        const uint32_t *hcs = undirected ? row_start : cs;
        const uint32_t *hsrc = undirected ? edge_dst : src;
        const uint32_t *hcw = undirected ? weights : cw;
        const uint32_t W = d.words;
        std::vector<uint32_t> off((size_t)W + 1, 0u);
        sg::parallel_for_dynamic(W, 4096, [&](uint64_t w0, uint64_t w1) {
            for (uint64_t w = w0; w < w1; w++) {
                uint32_t len = 0;
                for (uint64_t v = w * 64; v < std::min<uint64_t>(n, w * 64 + 64); v++) {
                    const uint32_t deg = hcs[v + 1] - hcs[v];
                    if (deg <= kPullHubDegree) len = std::max(len, deg);
                }
                off[w + 1] = len * 64;
            }
        });
        uint64_t total = 0;
        for (uint32_t w = 0; w < W; w++) {
            total += off[w + 1];
            off[w + 1] = (uint32_t)std::min<uint64_t>(total, 0xFFFFFFFFull);
        }
        // offsets are 32-bit: a larger padded layout is not built (the pull
        // rounds then schedule in-edges per wave instead)
        if (total >= 0xFFFFFFFFull) total = 0;
        std::vector<uint32_t> ssrc(total ? total : 1, 0xFFFFFFFFu), sw;
        if (hcw) sw.assign(total ? total : 1, 0u);
        if (total) sg::parallel_for_dynamic(W, 4096, [&](uint64_t w0, uint64_t w1) {
            for (uint64_t w = w0; w < w1; w++)
                for (uint64_t v = w * 64; v < std::min<uint64_t>(n, w * 64 + 64); v++) {
                    const uint32_t deg = hcs[v + 1] - hcs[v];
                    if (deg > kPullHubDegree) continue;
                    // a weighted vertex's in-edges by increasing weight (then
                    // source): an SSSP pull round's test "this weight alone
                    // cannot beat the best so far" then holds for every later
                    // row once it holds for one
                    uint64_t e[kPullHubDegree];
                    for (uint32_t j = 0; j < deg; j++)
                        e[j] = ((uint64_t)(hcw ? hcw[hcs[v] + j] : 0u) << 32) | hsrc[hcs[v] + j];
                    if (hcw) std::sort(e, e + deg);
                    for (uint32_t j = 0; j < deg; j++) {
                        const uint64_t at = off[w] + (uint64_t)j * 64 + (v - w * 64);
                        ssrc[at] = (uint32_t)e[j];
                        if (hcw) sw[at] = (uint32_t)(e[j] >> 32);
                    }
                }
        });
        if (total) {
            rc = up(off.data(), (uint64_t)W + 1, &d.sell_off);
            if (!rc) rc = up(ssrc.data(), total, &d.sell_src);
            if (!rc && hcw) rc = up(sw.data(), total, &d.sell_w);
        }
    }
    if (rc) {
        sg_graph_destroy(g);
        return rc;
    }
    if (weights && m) {
        double s = 0;
        for (uint64_t e = 0; e < m; e++) s += weights[e];
        g->avg_weight = s / (double)m;
    }
    *out = g;
    return SG_OK;
}

extern "C" int sg_graph_device_view(const sg_graph *g, sg_graph_view *out)
{
    if (!g || !out) SG_FAIL(SG_ERR_INVALID, "sg_graph_device_view: null argument");
    const DevGraph &d = g->dev;
    out->n = d.n;
    out->relabeled = g->relabeled ? 1u : 0u;
    out->m = d.m;
    out->row_start = d.row;
    out->edge_dst = d.col;
    out->weights = d.w;
    out->col_start = d.cstart;
    out->edge_src = d.csrc;
    out->csc_weights = d.cw;
    return SG_OK;
}

extern "C" int sg_graph_info(const sg_graph *g, uint32_t *n, uint64_t *m, uint64_t *device_bytes)
{
    if (!g) SG_FAIL(SG_ERR_INVALID, "sg_graph_info: null graph");
    if (n) *n = g->dev.n;
    if (m) *m = g->dev.m;
    if (device_bytes) *device_bytes = g->device_bytes;
    return SG_OK;
}

extern "C" void sg_engine_options_default(sg_engine_options *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->force_variant = SG_VARIANT_AUTO;
    o->fused = 0;
    o->priority = SG_PRIO_NONE;
    o->priority_threshold = 0.f;
    o->cut_threshold = 0.f;
    o->lb_push = SG_LB_DEFAULT;
    o->lb_pull = SG_LB_DEFAULT;
    o->alpha = 15.0;
    o->beta = 18.0;
    o->edge_factor = 1.0;
    o->max_iteration = 0;
    o->block_size = 0;
    o->large_degree = 0;
}

extern "C" int sg_engine_create(sg_graph *g, sg_engine **out)
{
    if (!g || !out) SG_FAIL(SG_ERR_INVALID, "sg_engine_create: null argument");
    *out = nullptr;
    SG_HIP(hipSetDevice(0));
    sg_engine *e = new sg_engine();
    e->g = g;
    sg_engine_options_default(&e->opt);
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        SG_FAIL(SG_ERR_HIP, "sg_engine_create: hipStreamCreate failed");
    }
    e->own_stream = true;
    if (hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess) {
        sg_engine_destroy(e);
        SG_FAIL(SG_ERR_HIP, "sg_engine_create: hipEventCreate failed");
    }
    *out = e;
    return SG_OK;
}

extern "C" int sg_engine_destroy(sg_engine *e)
{
    if (!e) return SG_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    engine_forget_stream(e);
    void *ptrs[] = {e->value, e->bitmaps, e->queues, e->ctr,
                    e->abort_flag, e->stats, e->trace, e->aux, e->srcbuf, e->garena, e->work};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (int i = 0; i < e->ngrids; i++) (void)hipFree(e->grids[i].words);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->own_stream && e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return SG_OK;
}

extern "C" int sg_engine_set_stream(sg_engine *e, void *hip_stream)
{
    if (!e) SG_FAIL(SG_ERR_INVALID, "sg_engine_set_stream: null engine");
    SG_HIP(hipStreamSynchronize(e->stream));
    engine_forget_stream(e);
    if (e->own_stream) (void)hipStreamDestroy(e->stream);
    e->stream = (hipStream_t)hip_stream;
    e->own_stream = false;
    return SG_OK;
}

extern "C" int sg_engine_stream(sg_engine *e, void **hip_stream)
{
    if (!e || !hip_stream) SG_FAIL(SG_ERR_INVALID, "sg_engine_stream: null argument");
    *hip_stream = (void *)e->stream;
    return SG_OK;
}

extern "C" int sg_engine_order_launch(sg_engine *e)
{
    if (!e) SG_FAIL(SG_ERR_INVALID, "sg_engine_order_launch: null engine");
    // the caller's kernel follows on this stream: any persistent kernel the
    // library enqueues later waits for it like for one of its own
    e->launch_seq++;
    return engine_launch_persistent(e, []() { return SG_OK; });
}

extern "C" int sg_engine_set_options(sg_engine *e, const sg_engine_options *o)
{
    if (!e || !o) SG_FAIL(SG_ERR_INVALID, "sg_engine_set_options: null argument");
    if (o->force_variant < SG_VARIANT_AUTO || o->force_variant > SG_SYNC_PULL_DD)
        SG_FAIL(SG_ERR_INVALID, "sg_engine_set_options: unknown variant");
    if (o->alpha < 0 || o->beta <= 0) SG_FAIL(SG_ERR_INVALID, "sg_engine_set_options: need alpha >= 0, beta > 0");
    if (o->block_size != 0 && (o->block_size % 64 || o->block_size < 64 || o->block_size > 1024))
        SG_FAIL(SG_ERR_INVALID, "sg_engine_set_options: block_size must be a multiple of 64 in [64, 1024]");
    if (o->max_iteration < 0) SG_FAIL(SG_ERR_INVALID, "sg_engine_set_options: max_iteration < 0");
    if (o->grid_blocks < 0 || o->grid_blocks > (int)kMaxBlocks)
        SG_FAIL(SG_ERR_INVALID, "sg_engine_set_options: grid_blocks must be in [0, 8192]");
    if (o->lb_push < SG_LB_DEFAULT || o->lb_push > SG_LB_HYBRID || o->lb_pull < SG_LB_DEFAULT ||
        o->lb_pull > SG_LB_HYBRID)
        SG_FAIL(SG_ERR_INVALID, "sg_engine_set_options: unknown load-balancing mode");
    if (o->large_degree != 0 && o->large_degree != (int32_t)kHugeDegree)
        SG_FAIL(SG_ERR_UNSUPPORTED, "sg_engine_set_options: large_degree must be 0 or 256 (the split the graph "
                                    "was built with)");
    e->opt = *o;
    return SG_OK;
}

namespace sg {

static std::mutex g_pk_mu;
static hipEvent_t g_pk_last = nullptr;     // recorded on demand, see below
static hipStream_t g_pk_stream = nullptr;  // stream of the last persistent launch

int engine_launch_persistent(sg_engine *e, const std::function<int()> &launch)
{
    return engine_launch_persistent_on(e->stream, launch);
}

int engine_launch_persistent_on(hipStream_t s, const std::function<int()> &launch)
{
    std::lock_guard<std::mutex> lk(g_pk_mu);
    if (g_pk_stream && g_pk_stream != s) {
        // recorded now on the other stream, the event completes no earlier
        // than that stream's last persistent kernel; launches that stay on
        // one stream record nothing (a record costs ~5 us between
        // back-to-back kernels)
        if (!g_pk_last) SG_HIP(hipEventCreateWithFlags(&g_pk_last, hipEventDisableTiming));
        SG_HIP(hipEventRecord(g_pk_last, g_pk_stream));
        SG_HIP(hipStreamWaitEvent(s, g_pk_last, 0));
    }
    int rc = launch();
    if (rc) return rc;
    g_pk_stream = s;
    return SG_OK;
}

void engine_forget_stream(sg_engine *e)
{
    std::lock_guard<std::mutex> lk(g_pk_mu);
    if (g_pk_stream == e->stream) g_pk_stream = nullptr;
}

std::mutex &persistent_mutex() { return g_pk_mu; }

int persistent_wait_other(hipStream_t s)
{
    // the previous persistent kernel went to another stream: wait for that
    // stream on the host (no event is created, nothing is allocated); the
    // caller completes its own kernel before releasing the lock, so later
    // launches need not wait for it
    if (g_pk_stream && g_pk_stream != s) SG_HIP(hipStreamSynchronize(g_pk_stream));
    g_pk_stream = nullptr;
    return SG_OK;
}

// Probe one grid size: true when every block was resident at once. The
// barrier words are zeroed first, so the grid's launches start from clean
// counters (every launch takes its own per-XCD census).
static int probe_grid(sg_engine *e, sg_engine::PGrid &pg, bool *resident)
{
    SG_HIP(hipMemsetAsync(pg.words, 0, kBarWords * 4, e->stream));
    int rc = engine_launch_persistent(e, [&]() { return pg.probe(pg.blocks, pg.words); });
    if (rc) return rc;
    uint32_t counts[kMaxXcd + 2];
    SG_HIP(hipMemcpyAsync(counts, pg.words + bar_count(0), sizeof(counts), hipMemcpyDeviceToHost, e->stream));
    SG_HIP(hipStreamSynchronize(e->stream));
    uint32_t total = 0;
    for (int x = 0; x < kMaxXcd; x++) total += counts[x];
    *resident = counts[kMaxXcd + 1] == 0 && total == (uint32_t)pg.blocks;
    // the probe's words share none with the barrier's counters, which stay zero
    pg.launches = 0;
    return SG_OK;
}

static int size_grid(sg_engine *e, sg_engine::PGrid &pg, int per_cu)
{
    if (pg.cap > 0 && pg.cap < per_cu * e->g->num_cus) {
        // a capped grid: the cap itself when it is co-resident
        pg.blocks = pg.cap;
        bool resident = false;
        int rc = probe_grid(e, pg, &resident);
        if (rc) return rc;
        if (resident) return SG_OK;
        per_cu = std::max(1, pg.cap / e->g->num_cus);
    }
    for (int pc = per_cu; pc >= 1; pc--) {
        pg.blocks = pc * e->g->num_cus;
        bool resident = false;
        int rc = probe_grid(e, pg, &resident);
        if (rc) return rc;
        if (resident) return SG_OK;
    }
    SG_FAIL(SG_ERR_HIP, "persistent grid: one block per CU is not co-resident");
}

int engine_persistent_grid(sg_engine *e, const void *fn, int bs, const std::function<int(int, uint32_t *)> &probe,
                           sg_engine::PGrid **out)
{
    e->launch_seq++;  // any persistent kernel may overwrite the BFS state
    if (e->needs_bar_reset) {
        e->bfs_token = 0;
        // an aborted run leaves counters mid-way: zero and re-probe every
        // grid, and zero the concurrent-SSSP barrier areas
        for (int i = 0; i < e->ngrids; i++) {
            sg_engine::PGrid &pg = e->grids[i];
            int rc = size_grid(e, pg, std::max(1, pg.blocks / e->g->num_cus));
            if (rc) return rc;
        }
        if (e->garena && e->garena_bar_bytes) {
            SG_HIP(hipMemsetAsync((char *)e->garena + e->garena_bar_off, 0, (size_t)e->garena_bar_bytes, e->stream));
            e->group_launches = 0;
        }
        SG_HIP(hipMemsetAsync(e->abort_flag, 0, 4, e->stream));
        e->needs_bar_reset = false;
    }
    const int cap = std::max(0, e->opt.grid_blocks);
    for (int i = 0; i < e->ngrids; i++) {
        if (e->grids[i].fn == fn && e->grids[i].bs == bs && e->grids[i].cap == cap) {
            *out = &e->grids[i];
            return SG_OK;
        }
    }
    if (e->ngrids == sg_engine::kMaxGrids) SG_FAIL(SG_ERR_STATE, "persistent grid: too many kernel configurations");
    int per_cu = 0;
    SG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, bs, 0));
    if (per_cu < 1) SG_FAIL(SG_ERR_HIP, "persistent grid: kernel cannot be resident");
    sg_engine::PGrid pg;
    pg.fn = fn;
    pg.bs = bs;
    pg.cap = cap;
    pg.probe = probe;
    hipError_t err = hipMalloc((void **)&pg.words, kBarWords * 4);
    if (err != hipSuccess) {
        (void)hipGetLastError();
        SG_FAIL(SG_ERR_NOMEM, "persistent grid: hipMalloc failed");
    }
    int rc = size_grid(e, pg, per_cu);
    if (rc) {
        (void)hipFree(pg.words);
        return rc;
    }
    e->grids[e->ngrids] = pg;
    *out = &e->grids[e->ngrids++];
    return SG_OK;
}

int engine_alloc(void **p, uint64_t bytes)
{
    hipError_t err = hipMalloc(p, bytes ? bytes : 4);
    if (err != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        sg::set_error(std::string("engine state: hipMalloc failed: ") + hipGetErrorString(err));
        return SG_ERR_NOMEM;
    }
    return SG_OK;
}


uint32_t dev_vertex(const sg_graph *g, uint32_t v) { return g->relabeled ? g->h_inv[v] : v; }

int engine_result_rows(sg_engine *e, uint64_t rows, void **buf)
{
    const uint64_t bytes = rows * (uint64_t)e->g->dev.n * 4;
    int rc = engine_ensure_value(e, bytes);
    if (rc) return rc;
    if (!e->g->relabeled) {
        *buf = e->value;
        return SG_OK;
    }
    if ((rc = engine_ensure_work(e, bytes))) return rc;
    *buf = e->work;
    return SG_OK;
}

int engine_ensure_work(sg_engine *e, uint64_t bytes)
{
    int rc;
    if (!e->work || e->work_cap < bytes) {
        if (e->work) {
            SG_HIP(hipStreamSynchronize(e->stream));
            (void)hipFree(e->work);
            e->work = nullptr;
            e->work_cap = 0;
        }
        if ((rc = engine_alloc(&e->work, bytes))) return rc;
        e->work_cap = bytes;
    }
    return SG_OK;
}

int engine_gather_rows(sg_engine *e, const void *src, void *dst, uint64_t rows)
{
    if (!e->g->relabeled) return SG_OK;
    const uint32_t n = e->g->dev.n;
    for (uint64_t r0 = 0; r0 < rows; r0 += 65535) {
        const uint32_t nr = (uint32_t)std::min<uint64_t>(65535, rows - r0);
        dim3 grid(std::min<uint32_t>((n + 255) / 256, 4096u), nr);
        hipLaunchKernelGGL(k_gather_rows, grid, dim3(256), 0, e->stream, (const uint32_t *)src + r0 * n,
                           (uint32_t *)dst + r0 * n, e->g->d_inv, n);
        SG_HIP(hipGetLastError());
    }
    return SG_OK;
}

int engine_run_status(sg_engine *e)
{
    SG_HIP(hipStreamSynchronize(e->stream));
    int32_t st = 0;
    SG_HIP(hipMemcpy(&st, (const char *)e->stats + offsetof(DevStats, status), 4, hipMemcpyDeviceToHost));
    if (st != SG_OK) {
        e->needs_bar_reset = true;
        SG_FAIL(st, "engine run reported an error status");
    }
    return SG_OK;
}

int engine_finish(sg_engine *e, uint64_t rows)
{
    return engine_gather_rows(e, e->work, e->value, rows);
}

int engine_ensure_srcbuf(sg_engine *e, uint64_t entries)
{
    if (e->srcbuf && e->srcbuf_cap >= entries) return SG_OK;
    if (e->srcbuf) {
        SG_HIP(hipStreamSynchronize(e->stream));
        (void)hipFree(e->srcbuf);
        e->srcbuf = nullptr;
        e->srcbuf_cap = 0;
    }
    int rc = engine_alloc((void **)&e->srcbuf, entries * 4);
    if (rc) return rc;
    e->srcbuf_cap = entries;
    return SG_OK;
}

int engine_ensure_value(sg_engine *e, uint64_t bytes)
{
    if (!e->value || e->value_cap < bytes) {
        if (e->value) {
            SG_HIP(hipStreamSynchronize(e->stream));
            (void)hipFree(e->value);
            e->value = nullptr;
            e->value_cap = 0;
        }
        int rc = engine_alloc(&e->value, bytes);
        if (rc) return rc;
        e->value_cap = bytes;
    }
    e->value_bytes = bytes;
    return SG_OK;
}

int engine_ensure_aux(sg_engine *e, uint64_t bytes)
{
    if (e->aux && e->aux_bytes >= bytes) return SG_OK;
    if (e->aux) {
        SG_HIP(hipStreamSynchronize(e->stream));
        (void)hipFree(e->aux);
        e->aux = nullptr;
        e->aux_bytes = 0;
    }
    int rc = engine_alloc(&e->aux, bytes);
    if (rc) return rc;
    e->aux_bytes = bytes;
    return SG_OK;
}

int engine_ensure_bfs_state(sg_engine *e)
{
    const DevGraph &d = e->g->dev;
    if (e->bitmaps) return SG_OK;
    auto alloc = [&](void **p, uint64_t bytes) -> int {
        hipError_t err = hipMalloc(p, bytes ? bytes : 4);
        if (err != hipSuccess) {
            (void)hipGetLastError();
            sg::set_error(std::string("engine state: hipMalloc failed: ") + hipGetErrorString(err));
            return SG_ERR_NOMEM;
        }
        return SG_OK;
    };
    int rc;
    if ((rc = engine_ensure_value(e, (uint64_t)d.n * 4))) return rc;
    // BFS: 2 visited filters + 3 rotating frontier bitmaps (SSSP uses the
    // first 3 words-sized areas, PageRank the first 2)
    if ((rc = alloc((void **)&e->bitmaps, (uint64_t)d.words * 8 * 5))) return rc;
    // Frontier queues may hold a vertex more than once (non-atomic discovery
    // across XCDs); past qcap a round falls back to scanning the level array.
    e->qcap = (uint32_t)std::min<uint64_t>(2ull * d.n + 4096, 0xFFFFFFF0ull);
    if ((rc = alloc((void **)&e->queues, (uint64_t)e->qcap * 4 * 3))) return rc;
    if ((rc = alloc((void **)&e->ctr, kCtrBytes))) return rc;
    if ((rc = alloc((void **)&e->abort_flag, 4))) return rc;
    SG_HIP(hipMemsetAsync(e->abort_flag, 0, 4, e->stream));
    if ((rc = alloc((void **)&e->stats, sizeof(DevStats)))) return rc;
    if ((rc = alloc((void **)&e->trace, sizeof(DevTrace) * kTraceRows))) return rc;
    SG_HIP(hipMemsetAsync(e->stats, 0, sizeof(DevStats), e->stream));
    e->needs_bar_reset = false;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0) == hipSuccess && khz > 0)
        e->wall_khz = khz;
    return SG_OK;
}

}  // namespace sg

extern "C" int sg_engine_sync(sg_engine *e)
{
    if (!e) SG_FAIL(SG_ERR_INVALID, "sg_engine_sync: null engine");
    SG_HIP(hipStreamSynchronize(e->stream));
    if (e->last_algo == 10) return sssp_group_status(e);
    if (e->stats) {
        DevStats s;
        SG_HIP(hipMemcpy(&s, e->stats, sizeof(s), hipMemcpyDeviceToHost));
        if (s.status != SG_OK) {
            e->needs_bar_reset = true;
            SG_FAIL(s.status, s.status == SG_ERR_TIMEOUT ? "device: grid barrier wait expired"
                                                         : "device: run reported an error");
        }
    }
    return SG_OK;
}

extern "C" int sg_engine_value_ptr(sg_engine *e, void **dev_ptr, uint64_t *bytes)
{
    if (!e || !dev_ptr) SG_FAIL(SG_ERR_INVALID, "sg_engine_value_ptr: null argument");
    if (!e->value) SG_FAIL(SG_ERR_STATE, "sg_engine_value_ptr: nothing has run yet");
    *dev_ptr = e->value;
    if (bytes) *bytes = e->value_bytes;
    return SG_OK;
}

extern "C" int sg_engine_gather_value(sg_engine *e, void *host_out, uint64_t bytes)
{
    if (!e || !host_out) SG_FAIL(SG_ERR_INVALID, "sg_engine_gather_value: null argument");
    if (!e->value) SG_FAIL(SG_ERR_STATE, "sg_engine_gather_value: nothing has run yet");
    if (bytes < e->value_bytes) SG_FAIL(SG_ERR_CAPACITY, "sg_engine_gather_value: host buffer too small");
    int rc = sg_engine_sync(e);
    if (rc) return rc;
    SG_HIP(hipMemcpy(host_out, e->value, e->value_bytes, hipMemcpyDeviceToHost));
    return SG_OK;
}

extern "C" int sg_engine_get_stats(sg_engine *e, sg_stats *out)
{
    if (!e || !out) SG_FAIL(SG_ERR_INVALID, "sg_engine_get_stats: null argument");
    std::memset(out, 0, sizeof(*out));
    if (!e->stats) SG_FAIL(SG_ERR_STATE, "sg_engine_get_stats: nothing has run yet");
    SG_HIP(hipStreamSynchronize(e->stream));
    DevStats s;
    SG_HIP(hipMemcpy(&s, e->stats, sizeof(s), hipMemcpyDeviceToHost));
    out->rounds = s.rounds;
    out->push_rounds = s.push_rounds;
    out->pull_rounds = s.pull_rounds;
    out->reached = s.reached;
    out->reached_edges = s.reached_edges;
    out->policy_trace = s.trace;
    out->max_depth = s.max_depth;
    out->status = s.status;
    if (e->last_algo == 10 && e->group_status) {
        // concurrent SSSP: the worst status over the call's queries
        int32_t gs = 0;
        SG_HIP(hipMemcpy(&gs, e->group_status, 4, hipMemcpyDeviceToHost));
        out->status = gs;
    }
    out->iterations = s.iterations;
    out->residual = s.residual;
    if ((e->last_algo == 0 || e->last_algo == 1) && e->value) {
        // traversals: exact reach statistics from the result array, computed
        // here rather than inside the timed traversal
        unsigned long long *acc = (unsigned long long *)((char *)e->ctr + kCtrBytes - 64);
        SG_HIP(hipMemsetAsync(acc, 0, 16, e->stream));
        const DevGraph &d = e->g->dev;
        // few blocks: each adds into the same two words once
        const uint32_t blocks = std::min<uint32_t>((d.n + 255) / 256, (uint32_t)e->g->num_cus);
        // the result is in the caller's order; degrees come through inv
        hipLaunchKernelGGL(k_reach_stats, dim3(blocks), dim3(256), 0, e->stream, (const uint32_t *)e->value, d.row,
                           e->g->relabeled ? e->g->d_inv : nullptr, d.n, acc);
        SG_HIP(hipGetLastError());
        unsigned long long h[2];
        SG_HIP(hipMemcpyAsync(h, acc, 16, hipMemcpyDeviceToHost, e->stream));
        SG_HIP(hipStreamSynchronize(e->stream));
        out->reached = (uint32_t)h[0];
        out->reached_edges = h[1];
    }
    if (e->timed) {
        float ms = 0.f;
        SG_HIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        out->kernel_ms = ms;
    }
    return SG_OK;
}

extern "C" int sg_engine_get_trace(sg_engine *e, sg_round_trace *out, int max, int *n_out)
{
    if (!e || !out || !n_out || max < 0) SG_FAIL(SG_ERR_INVALID, "sg_engine_get_trace: bad argument");
    *n_out = 0;
    if (!e->stats || !e->trace) SG_FAIL(SG_ERR_STATE, "sg_engine_get_trace: nothing has run yet");
    SG_HIP(hipStreamSynchronize(e->stream));
    DevStats s;
    SG_HIP(hipMemcpy(&s, e->stats, sizeof(s), hipMemcpyDeviceToHost));
    uint32_t rows = std::min<uint32_t>(std::min<uint32_t>(s.rounds + 1, kTraceRows), (uint32_t)max);
    std::vector<DevTrace> t(rows ? rows : 1);
    if (rows) SG_HIP(hipMemcpy(t.data(), e->trace, rows * sizeof(DevTrace), hipMemcpyDeviceToHost));
    const double us_per_tick = 1000.0 / e->wall_khz;
    for (uint32_t i = 0; i < rows; i++) {
        out[i].round = t[i].round;
        out[i].variant = t[i].variant;
        out[i].in_nodes = t[i].in_nodes;
        out[i].pad = 0;
        out[i].in_edges = t[i].in_edges;
        out[i].time_us = (double)t[i].ticks * us_per_tick;
    }
    *n_out = (int)rows;
    return SG_OK;
}
