// Graph files: the reference's three input formats (src/utils/parser.cpp,
// dispatched by GetCachedGraph, src/utils/utils.cpp:39-58), the cleanups and
// text outputs of its gtools converter (samples/tools/gtools.cpp:56-346), and
// the --gen_weights generator (include/utils/graphs/traversal.h:174-186).
//
// A reader produces exactly the CSR the reference's host graph binds
// (traversal.h:147-158 binds xadj/adjncy/adjwgt as row_start/edge_dst/
// edge_weights): same vertex count, same row contents in the same order, the
// same weights. It also keeps the reference's recorded edge count
// (`nedges`), which differs from row_start[n] for MatrixMarket files with
// duplicate entries. Unlike the reference, malformed files are reported
// through status codes instead of exit(), and row ends and vertex ids are
// validated before anything reaches a device.
#include "../../include/sepgraph.h"
#include "sg_host.hpp"

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <random>
#include <string>
#include <vector>

struct sg_host_csr {
    uint32_t n = 0;
    uint64_t header_edges = 0;
    bool weighted = false;
    std::vector<uint32_t> row, col, w;  // w empty unless weighted
};

namespace {

using sg::set_error;

int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

struct File {
    FILE *f = nullptr;
    File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

// Whole file into memory, NUL-terminated, every '\n' replaced by '\0' so that
// strtol stops at the end of a line exactly where it stops at the '\n' of
// the reference's getline buffer.
int slurp_lines(const char *path, std::vector<char> &buf, std::vector<uint64_t> &starts)
{
    File fp(path, "rb");
    if (!fp.f) return fail(SG_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
    if (std::fseek(fp.f, 0, SEEK_END) != 0) return fail(SG_ERR_IO, std::string("cannot seek ") + path);
    long sz = std::ftell(fp.f);
    if (sz < 0) return fail(SG_ERR_IO, std::string("cannot size ") + path);
    std::rewind(fp.f);
    buf.resize((size_t)sz + 1);
    if (sz > 0 && std::fread(buf.data(), 1, (size_t)sz, fp.f) != (size_t)sz)
        return fail(SG_ERR_IO, std::string("short read of ") + path);
    buf[(size_t)sz] = '\0';
    starts.clear();
    uint64_t b = 0;
    for (uint64_t i = 0; i < (uint64_t)sz; i++) {
        if (buf[i] == '\n') {
            buf[i] = '\0';
            starts.push_back(b);
            b = i + 1;
        }
    }
    if (b < (uint64_t)sz) starts.push_back(b);  // last line without '\n'
    return SG_OK;
}

// Next line not starting with '%' (the readers' comment rule), or nullptr.
struct Lines {
    std::vector<char> buf;
    std::vector<uint64_t> starts;
    size_t next = 0;
    char *get()
    {
        while (next < starts.size()) {
            char *l = buf.data() + starts[next++];
            if (l[0] != '%') return l;
        }
        return nullptr;
    }
};

// ------------------------------------------------------------------ .gr v1
// ReadGraphGR (parser.cpp:409-505).
int read_gr(const char *path, sg_host_csr &g)
{
    File fp(path, "rb");
    if (!fp.f) return fail(SG_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
    uint64_t h[4];
    if (std::fread(h, sizeof(uint64_t), 4, fp.f) != 4) return fail(SG_ERR_INVALID, "gr: unable to read header");
    if (h[0] != 1) return fail(SG_ERR_INVALID, "gr: unknown file version " + std::to_string(h[0]));
    const uint64_t esize = h[1], n = h[2], m = h[3];
    // the reference keeps n and m in 32-bit idx_t
    if (n >= 0xFFFFFFFFull || m > 0xFFFFFFFFull)
        return fail(SG_ERR_RANGE, "gr: n or m does not fit the 32-bit vertex / edge ids");
    if (esize != 0 && esize != 4)
        return fail(SG_ERR_UNSUPPORTED, "gr: edge data of " + std::to_string(esize) + " bytes (only 4-byte weights)");
    g.n = (uint32_t)n;
    g.header_edges = m;
    g.row.assign(n + 1, 0);
    std::vector<uint64_t> ends(n);
    if (n && std::fread(ends.data(), sizeof(uint64_t), n, fp.f) != n)
        return fail(SG_ERR_INVALID, "gr: partial read of node data");
    uint64_t prev = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (ends[i] < prev || ends[i] > m) return fail(SG_ERR_INVALID, "gr: row ends not non-decreasing within [0, m]");
        prev = ends[i];
        g.row[i + 1] = (uint32_t)ends[i];
    }
    if (prev != m) return fail(SG_ERR_INVALID, "gr: last row end differs from m");
    g.col.resize(m);
    if (m && std::fread(g.col.data(), sizeof(uint32_t), m, fp.f) != m)
        return fail(SG_ERR_INVALID, "gr: partial read of edge destinations");
    for (uint64_t e = 0; e < m; e++)
        if (g.col[e] >= n) return fail(SG_ERR_RANGE, "gr: edge destination out of range");
    if (esize) {
        if ((m & 1) && std::fseek(fp.f, 4, SEEK_CUR) != 0) return fail(SG_ERR_INVALID, "gr: error when seeking");
        g.w.resize(m);
        if (m && std::fread(g.w.data(), sizeof(uint32_t), m, fp.f) != m)
            return fail(SG_ERR_INVALID, "gr: partial read of edge data");
        g.weighted = true;
    }
    return SG_OK;
}

// ------------------------------------------------------------------- METIS
// ReadGraph (parser.cpp:230-395). Numbers are converted as the reference
// converts them: strtol, then truncated to the 32-bit idx_t.
int read_metis(const char *path, sg_host_csr &g)
{
    Lines L;
    if (int rc = slurp_lines(path, L.buf, L.starts)) return rc;
    char *line = L.get();
    if (!line) return fail(SG_ERR_INVALID, "metis: premature end of input file");
    long long nv = 0, ne = 0, fmt = 0, ncon = 0;
    int nf = std::sscanf(line, "%lld %lld %lld %lld", &nv, &ne, &fmt, &ncon);
    if (nf < 2) return fail(SG_ERR_INVALID, "metis: the header does not give the number of vertices and edges");
    if (nv <= 0 || ne <= 0) return fail(SG_ERR_INVALID, "metis: nvtxs and nedges must be positive");
    if (nv >= 0xFFFFFFFFll || 2 * ne > 0xFFFFFFFFll) return fail(SG_ERR_RANGE, "metis: graph too large for 32-bit ids");
    if (fmt > 111) return fail(SG_ERR_INVALID, "metis: cannot read this type of file format [fmt=" + std::to_string(fmt) + "]");
    char fs[32];
    std::snprintf(fs, sizeof fs, "%03lld", fmt % 1000);
    const bool readvs = fs[0] == '1', readvw = fs[1] == '1', readew = fs[2] == '1';
    if (ncon > 0 && !readvw) return fail(SG_ERR_INVALID, "metis: ncon given but fmt has no vertex weights");
    if (ncon == 0) ncon = 1;
    const uint32_t n = (uint32_t)nv;
    const uint64_t nedges = 2 * (uint64_t)ne;
    g.n = n;
    g.header_edges = nedges;
    g.weighted = readew;
    g.row.assign((size_t)n + 1, 0);
    g.col.resize(nedges);
    if (readew) g.w.resize(nedges);
    uint64_t k = 0;
    for (uint32_t i = 0; i < n; i++) {
        char *cur = L.get();
        if (!cur) return fail(SG_ERR_INVALID, "metis: premature end of input file while reading vertex " + std::to_string(i + 1));
        char *nxt = nullptr;
        if (readvs) {
            std::strtol(cur, &nxt, 10);
            if (nxt == cur) return fail(SG_ERR_INVALID, "metis: vertex " + std::to_string(i + 1) + " has no size");
            cur = nxt;
        }
        if (readvw) {
            for (long long c = 0; c < ncon; c++) {
                std::strtol(cur, &nxt, 10);
                if (nxt == cur) return fail(SG_ERR_INVALID, "metis: vertex " + std::to_string(i + 1) + " lacks weights");
                cur = nxt;
            }
        }
        for (;;) {
            const uint32_t edge = (uint32_t)std::strtol(cur, &nxt, 10);
            if (nxt == cur) break;
            cur = nxt;
            if (edge < 1 || edge > n)
                return fail(SG_ERR_RANGE, "metis: edge " + std::to_string(edge) + " of vertex " + std::to_string(i + 1) + " out of bounds");
            uint32_t ew = 1;
            if (readew) {
                ew = (uint32_t)std::strtol(cur, &nxt, 10);
                if (nxt == cur) return fail(SG_ERR_INVALID, "metis: premature end of line for vertex " + std::to_string(i + 1));
                if (ew == 0) return fail(SG_ERR_INVALID, "metis: edge weights must be positive");
                cur = nxt;
            }
            if (k == nedges) return fail(SG_ERR_INVALID, "metis: more edges in the file than the header gives");
            g.col[k] = edge - 1;
            if (readew) g.w[k] = ew;
            k++;
        }
        g.row[i + 1] = (uint32_t)k;
    }
    if (k != nedges)
        return fail(SG_ERR_INVALID, "metis: the header gives " + std::to_string(nedges / 2) + " edges, the file holds " +
                                        std::to_string(k / 2));
    return SG_OK;
}

// ------------------------------------------------------------ MatrixMarket
// ReadGraphMarket (parser.cpp:508-665): a std::set per row gives sorted,
// duplicate-free rows; a sort of (row, head[, weight]) triples followed by a
// unique pass gives the same rows.
int read_market(const char *path, sg_host_csr &g)
{
    Lines L;
    if (int rc = slurp_lines(path, L.buf, L.starts)) return rc;
    char *line = L.get();
    if (!line) return fail(SG_ERR_INVALID, "market: premature end of input file");
    long long rows = 0, cols = 0, nnz = 0;
    if (std::sscanf(line, "%lld %lld %lld", &rows, &cols, &nnz) < 3)
        return fail(SG_ERR_INVALID, "market: the header does not give rows, columns and entries");
    if (cols < 0 || nnz < 0 || cols >= 0xFFFFFFFFll || nnz > 0xFFFFFFFFll)
        return fail(SG_ERR_RANGE, "market: sizes out of range");
    const uint32_t n = (uint32_t)cols;  // the reference reads both sizes into nvtxs
    struct Ent {
        uint32_t s, d, w;
    };
    std::vector<Ent> ents;
    ents.reserve((size_t)nnz);
    int weighted = -1;
    for (long long e = 0; e < nnz; e++) {
        char *l = L.get();
        if (!l) return fail(SG_ERR_INVALID, "market: premature end of input file while reading edge list");
        char *p = nullptr;
        long src = std::strtol(l, &p, 10) - 1;
        long dst = std::strtol(p, &p, 10) - 1;
        if (src < 0 || src >= (long)n || dst < 0 || dst >= (long)n)
            return fail(SG_ERR_RANGE, "market: entry '" + std::to_string(src) + " " + std::to_string(dst) + "' out of range");
        if (weighted < 0) {
            weighted = 0;
            if (*p == ' ' || *p == '\t') {
                weighted = 1;
                for (const char *q = p; *q; q++)
                    if (*q == '.') return fail(SG_ERR_UNSUPPORTED, "market: floating-point weights are not supported");
            }
        }
        uint32_t w = 0;
        if (weighted == 1) w = (uint32_t)std::strtol(p, &p, 10);
        ents.push_back({(uint32_t)src, (uint32_t)dst, w});
    }
    std::sort(ents.begin(), ents.end(), [](const Ent &a, const Ent &b) {
        if (a.s != b.s) return a.s < b.s;
        if (a.d != b.d) return a.d < b.d;
        return a.w < b.w;
    });
    const bool wt = weighted == 1;
    auto same = [wt](const Ent &a, const Ent &b) { return a.s == b.s && a.d == b.d && (!wt || a.w == b.w); };
    ents.erase(std::unique(ents.begin(), ents.end(), same), ents.end());
    g.n = n;
    g.header_edges = (uint64_t)nnz;
    g.weighted = wt;
    g.row.assign((size_t)n + 1, 0);
    g.col.resize(ents.size());
    if (wt) g.w.resize(ents.size());
    for (size_t i = 0; i < ents.size(); i++) {
        g.row[ents[i].s + 1]++;
        g.col[i] = ents[i].d;
        if (wt) g.w[i] = ents[i].w;
    }
    for (uint32_t v = 0; v < n; v++) g.row[v + 1] += g.row[v];
    return SG_OK;
}

// Rebuild every row through keep(u, b, e, out_col, out_w), rows split over
// host threads; the new CSR replaces g's.
template <typename F>
void rebuild_rows(sg_host_csr &g, F keep)
{
    const int nt = sg::host_threads();
    std::vector<std::vector<uint32_t>> tc(nt), tw(nt);
    std::vector<uint32_t> deg(g.n);
    std::vector<uint64_t> lo(nt + 1, g.n);
    sg::parallel_for(
        g.n,
        [&](uint64_t b, uint64_t e, int t) {
            lo[t] = b;
            for (uint64_t u = b; u < e; u++) {
                const size_t before = tc[t].size();
                keep((uint32_t)u, g.row[u], g.row[u + 1], tc[t], tw[t]);
                deg[u] = (uint32_t)(tc[t].size() - before);
            }
        },
        nt);
    std::vector<uint32_t> row((size_t)g.n + 1, 0);
    for (uint32_t u = 0; u < g.n; u++) row[u + 1] = row[u] + deg[u];
    std::vector<uint32_t> col(row[g.n]), w(g.weighted ? row[g.n] : 0);
    // thread t's rows are contiguous and start at lo[t]
    for (int t = 0; t < nt; t++) {
        if (tc[t].empty()) continue;
        const uint64_t at = row[lo[t]];
        std::copy(tc[t].begin(), tc[t].end(), col.begin() + at);
        if (g.weighted) std::copy(tw[t].begin(), tw[t].end(), w.begin() + at);
    }
    g.row.swap(row);
    g.col.swap(col);
    g.w.swap(w);
}

// gtools --deself_cycle (gtools.cpp:56-131).
void drop_self_loops(sg_host_csr &g)
{
    rebuild_rows(g, [&](uint32_t u, uint32_t b, uint32_t e, std::vector<uint32_t> &oc, std::vector<uint32_t> &ow) {
        for (uint32_t i = b; i < e; i++)
            if (g.col[i] != u) {
                oc.push_back(g.col[i]);
                if (g.weighted) ow.push_back(g.w[i]);
            }
    });
}

// gtools --deduplicate (gtools.cpp:133-212): first occurrence kept, row order
// otherwise unchanged.
void dedup_first(sg_host_csr &g)
{
    rebuild_rows(g, [&](uint32_t, uint32_t b, uint32_t e, std::vector<uint32_t> &oc, std::vector<uint32_t> &ow) {
        const uint32_t d = e - b;
        if (d <= 32) {
            for (uint32_t i = b; i < e; i++) {
                bool seen = false;
                for (uint32_t j = b; j < i && !seen; j++) seen = g.col[j] == g.col[i];
                if (!seen) {
                    oc.push_back(g.col[i]);
                    if (g.weighted) ow.push_back(g.w[i]);
                }
            }
            return;
        }
        std::vector<uint64_t> key(d);  // head << 32 | position: first occurrence sorts first
        for (uint32_t i = 0; i < d; i++) key[i] = (uint64_t)g.col[b + i] << 32 | i;
        std::sort(key.begin(), key.end());
        std::vector<uint32_t> firsts;
        for (uint32_t i = 0; i < d; i++)
            if (i == 0 || (key[i] >> 32) != (key[i - 1] >> 32)) firsts.push_back((uint32_t)key[i]);
        std::sort(firsts.begin(), firsts.end());
        for (uint32_t p : firsts) {
            oc.push_back(g.col[b + p]);
            if (g.weighted) ow.push_back(g.w[b + p]);
        }
    });
}

// gtools --d2ud (gtools.cpp:215-327): row x gets heads {y : x->y or y->x},
// sorted, each once. The weighted form inserts (y, w) into a std::map per row
// while walking the edges in row-major order, so the first edge between x
// and y in that order gives the pair's weight in both rows.
void to_undirected(sg_host_csr &g)
{
    const uint32_t n = g.n;
    const uint64_t m = g.row[n];
    // reverse adjacency: for each head, (tail, edge id) in edge-id order
    std::vector<uint32_t> rstart((size_t)n + 1, 0);
    for (uint64_t e = 0; e < m; e++) rstart[g.col[e] + 1]++;
    for (uint32_t v = 0; v < n; v++) rstart[v + 1] += rstart[v];
    std::vector<uint32_t> rsrc(m), reid(m);
    {
        std::vector<uint32_t> fill(rstart.begin(), rstart.end() - 1);
        for (uint32_t u = 0; u < n; u++)
            for (uint32_t e = g.row[u]; e < g.row[u + 1]; e++) {
                const uint32_t at = fill[g.col[e]]++;
                rsrc[at] = u;
                reid[at] = e;
            }
    }
    rebuild_rows(g, [&](uint32_t x, uint32_t b, uint32_t e, std::vector<uint32_t> &oc, std::vector<uint32_t> &ow) {
        std::vector<uint64_t> key;  // head << 32 | edge id
        key.reserve((size_t)(e - b) + (rstart[x + 1] - rstart[x]));
        for (uint32_t i = b; i < e; i++) key.push_back((uint64_t)g.col[i] << 32 | i);
        for (uint32_t i = rstart[x]; i < rstart[x + 1]; i++) key.push_back((uint64_t)rsrc[i] << 32 | reid[i]);
        std::sort(key.begin(), key.end());
        for (size_t i = 0; i < key.size(); i++) {
            if (i && (key[i] >> 32) == (key[i - 1] >> 32)) continue;
            oc.push_back((uint32_t)(key[i] >> 32));
            if (g.weighted) ow.push_back(g.w[(uint32_t)key[i]]);
        }
    });
}

int write_gr(const sg_host_csr &g, const char *path)
{
    File fp(path, "wb");
    if (!fp.f) return fail(SG_ERR_IO, std::string("cannot create ") + path + ": " + std::strerror(errno));
    const uint64_t m = g.row[g.n];
    const uint64_t h[4] = {1, g.weighted ? 4ull : 0ull, g.n, m};
    bool ok = std::fwrite(h, sizeof h, 1, fp.f) == 1;
    std::vector<uint64_t> ends(g.row.begin() + 1, g.row.end());
    ok = ok && (ends.empty() || std::fwrite(ends.data(), sizeof(uint64_t), ends.size(), fp.f) == ends.size());
    ok = ok && (m == 0 || std::fwrite(g.col.data(), sizeof(uint32_t), m, fp.f) == m);
    if (g.weighted) {
        const uint32_t pad = 0;
        if (m & 1) ok = ok && std::fwrite(&pad, 4, 1, fp.f) == 1;
        ok = ok && (m == 0 || std::fwrite(g.w.data(), sizeof(uint32_t), m, fp.f) == m);
    }
    return ok ? SG_OK : fail(SG_ERR_IO, std::string("short write to ") + path);
}

int write_metis(const sg_host_csr &g, const char *path)
{
    const uint64_t m = g.row[g.n];
    if (g.n == 0 || m == 0 || (m & 1))
        return fail(SG_ERR_INVALID, "metis: needs n > 0 and an even, positive number of edge entries");
    File fp(path, "w");
    if (!fp.f) return fail(SG_ERR_IO, std::string("cannot create ") + path + ": " + std::strerror(errno));
    bool ok = std::fprintf(fp.f, "%% METIS graph, %u vertices\n", g.n) > 0;
    ok = ok && (g.weighted ? std::fprintf(fp.f, "%u %llu 1\n", g.n, (unsigned long long)(m / 2))
                           : std::fprintf(fp.f, "%u %llu\n", g.n, (unsigned long long)(m / 2))) > 0;
    std::string ln;
    char num[32];
    for (uint32_t u = 0; u < g.n && ok; u++) {
        ln.clear();
        for (uint32_t e = g.row[u]; e < g.row[u + 1]; e++) {
            int k = g.weighted ? std::snprintf(num, sizeof num, "%s%u %u", ln.empty() ? "" : " ", g.col[e] + 1, g.w[e])
                               : std::snprintf(num, sizeof num, "%s%u", ln.empty() ? "" : " ", g.col[e] + 1);
            ln.append(num, (size_t)k);
        }
        ln.push_back('\n');
        ok = std::fwrite(ln.data(), 1, ln.size(), fp.f) == ln.size();
    }
    return ok ? SG_OK : fail(SG_ERR_IO, std::string("short write to ") + path);
}

int write_market(const sg_host_csr &g, const char *path)
{
    File fp(path, "w");
    if (!fp.f) return fail(SG_ERR_IO, std::string("cannot create ") + path + ": " + std::strerror(errno));
    const uint64_t m = g.row[g.n];
    bool ok = std::fprintf(fp.f, "%%%%MatrixMarket matrix coordinate %s general\n%u %u %llu\n",
                           g.weighted ? "integer" : "pattern", g.n, g.n, (unsigned long long)m) > 0;
    for (uint32_t u = 0; u < g.n && ok; u++)
        for (uint32_t e = g.row[u]; e < g.row[u + 1] && ok; e++)
            ok = (g.weighted ? std::fprintf(fp.f, "%u %u %u\n", u + 1, g.col[e] + 1, g.w[e])
                             : std::fprintf(fp.f, "%u %u\n", u + 1, g.col[e] + 1)) > 0;
    return ok ? SG_OK : fail(SG_ERR_IO, std::string("short write to ") + path);
}

}  // namespace

extern "C" int sg_gen_weights_std(uint64_t m, int32_t range, uint32_t *weights)
{
    if (range < 1 || (m && !weights)) return fail(SG_ERR_INVALID, "sg_gen_weights_std: range must be >= 1");
    std::default_random_engine gen;
    std::uniform_int_distribution<int> dist(1, range);
    for (uint64_t i = 0; i < m; i++) weights[i] = (uint32_t)dist(gen);
    return SG_OK;
}

extern "C" int sg_file_read(const char *path, int format, sg_host_csr **out)
{
    if (!path || !out) return fail(SG_ERR_INVALID, "sg_file_read: null argument");
    *out = nullptr;
    std::unique_ptr<sg_host_csr> g(new (std::nothrow) sg_host_csr);
    if (!g) return fail(SG_ERR_NOMEM, "sg_file_read: out of host memory");
    int rc;
    try {
        switch (format) {
        case SG_FORMAT_GR: rc = read_gr(path, *g); break;
        case SG_FORMAT_METIS: rc = read_metis(path, *g); break;
        case SG_FORMAT_MARKET: rc = read_market(path, *g); break;
        default: return fail(SG_ERR_INVALID, "sg_file_read: unsupported graph format");
        }
    } catch (const std::bad_alloc &) {
        return fail(SG_ERR_NOMEM, "sg_file_read: out of host memory");
    }
    if (rc) return rc;
    *out = g.release();
    return SG_OK;
}

extern "C" int sg_file_from_csr(uint32_t n, const uint32_t *row_start, const uint32_t *edge_dst,
                                const uint32_t *weights, sg_host_csr **out)
{
    if (!row_start || !out || n == 0xFFFFFFFFu) return fail(SG_ERR_INVALID, "sg_file_from_csr: bad argument");
    *out = nullptr;
    const uint64_t m = row_start[n];
    if (m && !edge_dst) return fail(SG_ERR_INVALID, "sg_file_from_csr: edge_dst is null");
    for (uint32_t v = 0; v < n; v++)
        if (row_start[v] > row_start[v + 1]) return fail(SG_ERR_INVALID, "sg_file_from_csr: row_start decreases");
    for (uint64_t e = 0; e < m; e++)
        if (edge_dst[e] >= n) return fail(SG_ERR_RANGE, "sg_file_from_csr: edge destination out of range");
    try {
        std::unique_ptr<sg_host_csr> g(new sg_host_csr);
        g->n = n;
        g->header_edges = m;
        g->row.assign(row_start, row_start + n + 1);
        g->col.assign(edge_dst, edge_dst + m);
        if (weights) {
            g->weighted = true;
            g->w.assign(weights, weights + m);
        }
        *out = g.release();
    } catch (const std::bad_alloc &) {
        return fail(SG_ERR_NOMEM, "sg_file_from_csr: out of host memory");
    }
    return SG_OK;
}

extern "C" int sg_file_free(sg_host_csr *g)
{
    delete g;
    return SG_OK;
}

extern "C" int sg_file_info(const sg_host_csr *g, uint32_t *n, uint64_t *m, int *weighted, uint64_t *header_edges)
{
    if (!g) return fail(SG_ERR_INVALID, "sg_file_info: null handle");
    if (n) *n = g->n;
    if (m) *m = g->row[g->n];
    if (weighted) *weighted = g->weighted ? 1 : 0;
    if (header_edges) *header_edges = g->header_edges;
    return SG_OK;
}

extern "C" int sg_file_copy(const sg_host_csr *g, uint32_t *row_start, uint32_t *edge_dst, uint32_t *weights)
{
    if (!g || !row_start || (g->row[g->n] && !edge_dst)) return fail(SG_ERR_INVALID, "sg_file_copy: null argument");
    std::copy(g->row.begin(), g->row.end(), row_start);
    std::copy(g->col.begin(), g->col.end(), edge_dst);
    if (weights && g->weighted) std::copy(g->w.begin(), g->w.end(), weights);
    return SG_OK;
}

extern "C" int sg_file_clean(sg_host_csr *g, int flags)
{
    if (!g) return fail(SG_ERR_INVALID, "sg_file_clean: null handle");
    if (flags & ~(SG_CLEAN_SELF_LOOPS | SG_CLEAN_DEDUP | SG_CLEAN_TO_UNDIRECTED))
        return fail(SG_ERR_INVALID, "sg_file_clean: unknown flag");
    try {
        if (flags & SG_CLEAN_SELF_LOOPS) drop_self_loops(*g);
        if (flags & SG_CLEAN_DEDUP) dedup_first(*g);
        if (flags & SG_CLEAN_TO_UNDIRECTED) {
            if ((uint64_t)g->row[g->n] * 2 > 0xFFFFFFFFull)
                return fail(SG_ERR_RANGE, "sg_file_clean: undirected graph would exceed 32-bit edge ids");
            to_undirected(*g);
        }
    } catch (const std::bad_alloc &) {
        return fail(SG_ERR_NOMEM, "sg_file_clean: out of host memory");
    }
    g->header_edges = g->row[g->n];
    return SG_OK;
}

extern "C" int sg_file_write(const sg_host_csr *g, const char *path, int format)
{
    if (!g || !path) return fail(SG_ERR_INVALID, "sg_file_write: null argument");
    switch (format) {
    case SG_FORMAT_GR: return write_gr(*g, path);
    case SG_FORMAT_METIS: return write_metis(*g, path);
    case SG_FORMAT_MARKET: return write_market(*g, path);
    default: return fail(SG_ERR_INVALID, "sg_file_write: unsupported graph format");
    }
}

extern "C" int sg_file_write_edgelist(const sg_host_csr *g, const char *path, const char *separator)
{
    if (!g || !path || !separator) return fail(SG_ERR_INVALID, "sg_file_write_edgelist: null argument");
    File fp(path, "w");
    if (!fp.f) return fail(SG_ERR_IO, std::string("cannot create ") + path + ": " + std::strerror(errno));
    bool ok = true;
    for (uint32_t u = 0; u < g->n && ok; u++)
        for (uint32_t e = g->row[u]; e < g->row[u + 1] && ok; e++)
            ok = (g->weighted ? std::fprintf(fp.f, "%u%s%u%s%u\n", u, separator, g->col[e], separator, g->w[e])
                              : std::fprintf(fp.f, "%u%s%u\n", u, separator, g->col[e])) > 0;
    return ok ? SG_OK : fail(SG_ERR_IO, std::string("short write to ") + path);
}

extern "C" int sg_file_write_out_degree(const sg_host_csr *g, const char *path, const char *separator)
{
    if (!g || !path || !separator) return fail(SG_ERR_INVALID, "sg_file_write_out_degree: null argument");
    File fp(path, "w");
    if (!fp.f) return fail(SG_ERR_IO, std::string("cannot create ") + path + ": " + std::strerror(errno));
    bool ok = true;
    for (uint32_t u = 0; u < g->n && ok; u++)
        ok = std::fprintf(fp.f, "%u%s%u\n", u, separator, g->row[u + 1] - g->row[u]) > 0;
    return ok ? SG_OK : fail(SG_ERR_IO, std::string("short write to ") + path);
}
