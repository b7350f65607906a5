/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle.h for the contract and
 * for how each routine is pinned). Plain C99, single-threaded.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_INF32 0xFFFFFFFFu

/* Transpose by counting sort over rows in ascending order, so each column
 * lists its sources ascending (csr_graph.h:330-397). */
int or_csc_from_csr(uint32_t n, const uint32_t *row, const uint32_t *col, const uint32_t *w,
                    uint32_t *col_start, uint32_t *src, uint32_t *w_out)
{
    uint32_t m = row[n];
    uint32_t *cursor = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    if (!cursor) return -1;
    for (uint32_t e = 0; e < m; e++) {
        if (col[e] >= n) { free(cursor); return -2; }
        cursor[col[e] + 1]++;
    }
    for (uint32_t v = 0; v < n; v++) cursor[v + 1] += cursor[v];
    memcpy(col_start, cursor, ((size_t)n + 1) * sizeof(uint32_t));
    for (uint32_t u = 0; u < n; u++) {
        for (uint32_t e = row[u]; e < row[u + 1]; e++) {
            uint32_t pos = cursor[col[e]]++;
            src[pos] = u;
            if (w && w_out) w_out[pos] = w[e];
        }
    }
    free(cursor);
    return 0;
}

/* BFSHost, hybrid_bfs_host.cpp:33-60: a FIFO queue; a vertex gets level+1
 * the first time it is seen. */
int or_bfs(uint32_t n, const uint32_t *row, const uint32_t *col, uint32_t source, int32_t *levels)
{
    if (source >= n) return -2;
    uint32_t *queue = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    if (!queue) return -1;
    for (uint32_t v = 0; v < n; v++) levels[v] = -1;
    size_t head = 0, tail = 0;
    levels[source] = 0;
    queue[tail++] = source;
    while (head < tail) {
        uint32_t u = queue[head++];
        int32_t next = levels[u] + 1;
        for (uint32_t e = row[u]; e < row[u + 1]; e++) {
            uint32_t v = col[e];
            if (levels[v] == -1) {
                levels[v] = next;
                queue[tail++] = v;
            }
        }
    }
    free(queue);
    return 0;
}

/* SSSPHostNaive, hybrid_sssp_host.cpp:34-60: FIFO label-correcting relaxation
 * in uint32 arithmetic; a vertex is re-queued every time its distance drops.
 * The queue is a growable ring buffer (std::queue in the reference). */
int or_sssp_fifo(uint32_t n, const uint32_t *row, const uint32_t *col, const uint32_t *w,
                 uint32_t source, uint32_t *dist)
{
    if (source >= n) return -2;
    size_t cap = (size_t)n + 16, head = 0, count = 0;
    uint32_t *ring = (uint32_t *)malloc(cap * sizeof(uint32_t));
    if (!ring) return -1;
    for (uint32_t v = 0; v < n; v++) dist[v] = OR_INF32;
    dist[source] = 0;
    ring[0] = source;
    count = 1;
    while (count) {
        uint32_t u = ring[head];
        head = (head + 1) % cap;
        count--;
        uint32_t du = dist[u];
        for (uint32_t e = row[u]; e < row[u + 1]; e++) {
            uint32_t v = col[e];
            uint32_t nd = du + w[e];
            if (dist[v] > nd) {
                dist[v] = nd;
                if (count == cap) { /* grow: unroll the ring into a bigger buffer */
                    size_t ncap = cap * 2;
                    uint32_t *nr = (uint32_t *)malloc(ncap * sizeof(uint32_t));
                    if (!nr) { free(ring); return -1; }
                    for (size_t i = 0; i < count; i++) nr[i] = ring[(head + i) % cap];
                    free(ring);
                    ring = nr;
                    cap = ncap;
                    head = 0;
                }
                ring[(head + count) % cap] = v;
                count++;
            }
        }
    }
    free(ring);
    return 0;
}

/* ---- Dijkstra (binary heap of (dist, vertex), lazy deletion) ---- */
typedef struct { uint64_t d; uint32_t v; } or_hitem;

static void heap_push(or_hitem *h, size_t *sz, or_hitem it)
{
    size_t i = (*sz)++;
    while (i > 0) {
        size_t p = (i - 1) / 2;
        if (h[p].d <= it.d) break;
        h[i] = h[p];
        i = p;
    }
    h[i] = it;
}

static or_hitem heap_pop(or_hitem *h, size_t *sz)
{
    or_hitem top = h[0];
    or_hitem last = h[--(*sz)];
    size_t i = 0, n = *sz;
    for (;;) {
        size_t l = 2 * i + 1, r = l + 1, s = i;
        uint64_t sd = last.d;
        if (l < n && h[l].d < sd) { s = l; sd = h[l].d; }
        if (r < n && h[r].d < sd) { s = r; }
        if (s == i) break;
        h[i] = h[s];
        i = s;
    }
    if (n) h[i] = last;
    return top;
}

int or_sssp_dijkstra(uint32_t n, const uint32_t *row, const uint32_t *col, const uint32_t *w,
                     uint32_t source, uint32_t *dist)
{
    if (source >= n) return -2;
    uint32_t m = row[n];
    or_hitem *h = (or_hitem *)malloc(((size_t)m + 2) * sizeof(or_hitem));
    uint64_t *d = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    if (!h || !d) { free(h); free(d); return -1; }
    for (uint32_t v = 0; v < n; v++) d[v] = UINT64_MAX;
    size_t sz = 0;
    d[source] = 0;
    or_hitem s0 = {0, source};
    heap_push(h, &sz, s0);
    while (sz) {
        or_hitem it = heap_pop(h, &sz);
        if (it.d != d[it.v]) continue;
        for (uint32_t e = row[it.v]; e < row[it.v + 1]; e++) {
            uint64_t nd = it.d + w[e];
            uint32_t v = col[e];
            if (nd < d[v]) {
                d[v] = nd;
                or_hitem ni = {nd, v};
                heap_push(h, &sz, ni);
            }
        }
    }
    for (uint32_t v = 0; v < n; v++) dist[v] = d[v] == UINT64_MAX ? OR_INF32 : (uint32_t)d[v];
    free(h);
    free(d);
    return 0;
}

/* FP64 power iteration with uniform teleport and uniform dangling redistribution. */
int or_pagerank(uint32_t n, const uint32_t *row, const uint32_t *col, double damping, double tol,
                int max_iter, double *x, int *iters, double *last_residual)
{
    if (n == 0) return -2;
    double *y = (double *)malloc((size_t)n * sizeof(double));
    if (!y) return -1;
    for (uint32_t v = 0; v < n; v++) x[v] = 1.0 / n;
    int it = 0;
    double res = INFINITY;
    while (it < max_iter) {
        double dangling = 0.0;
        for (uint32_t u = 0; u < n; u++)
            if (row[u + 1] == row[u]) dangling += x[u];
        double base = (1.0 - damping) / n + damping * dangling / n;
        for (uint32_t v = 0; v < n; v++) y[v] = base;
        for (uint32_t u = 0; u < n; u++) {
            uint32_t deg = row[u + 1] - row[u];
            if (!deg) continue;
            double c = damping * x[u] / deg;
            for (uint32_t e = row[u]; e < row[u + 1]; e++) y[col[e]] += c;
        }
        res = 0.0;
        for (uint32_t v = 0; v < n; v++) {
            res += fabs(y[v] - x[v]);
            x[v] = y[v];
        }
        it++;
        if (res < tol) break;
    }
    free(y);
    if (iters) *iters = it;
    if (last_residual) *last_residual = res;
    return 0;
}

/* PageRankHost, hybrid_pr_host.cpp:43-124, in FP64: ranks start at 1-a, the
 * residual is seeded by one push of (1-a)*a/outdeg along every edge, then a
 * FIFO worklist absorbs residual into the rank and pushes a*res/outdeg to
 * out-neighbours; a neighbour is enqueued when its residual crosses `error`. */
int or_pagerank_residual(uint32_t n, const uint32_t *row, const uint32_t *col, double alpha,
                         double error, double *rank)
{
    double *res = (double *)calloc((size_t)n, sizeof(double));
    uint32_t *qa = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
    uint32_t *qb = (uint32_t *)malloc(((size_t)row[n] + (size_t)n + 1) * sizeof(uint32_t));
    if (!res || !qa || !qb) { free(res); free(qa); free(qb); return -1; }
    /* the out-queue can receive each vertex at most once per crossing; bound by m + n */
    size_t cap_b = (size_t)row[n] + (size_t)n + 1;
    uint32_t *in = qa, *out = qb;
    size_t cap_in = (size_t)n + 1, cap_out = cap_b;
    for (uint32_t v = 0; v < n; v++) rank[v] = 1.0 - alpha;
    for (uint32_t u = 0; u < n; u++) {
        uint32_t deg = row[u + 1] - row[u];
        if (!deg) continue;
        double upd = ((1.0 - alpha) * alpha) / deg;
        for (uint32_t e = row[u]; e < row[u + 1]; e++) res[col[e]] += upd;
    }
    size_t nin = 0;
    for (uint32_t v = 0; v < n; v++) in[nin++] = v;
    while (nin) {
        size_t nout = 0;
        for (size_t i = 0; i < nin; i++) {
            uint32_t u = in[i];
            double r = res[u];
            rank[u] += r;
            res[u] = 0.0;
            uint32_t deg = row[u + 1] - row[u];
            if (!deg) continue;
            double upd = r * alpha / deg;
            for (uint32_t e = row[u]; e < row[u + 1]; e++) {
                uint32_t v = col[e];
                double prev = res[v];
                res[v] += upd;
                if (prev + upd > error && prev < error) {
                    if (nout == cap_out) {
                        size_t nc = cap_out * 2;
                        uint32_t *nb = (uint32_t *)realloc(out, nc * sizeof(uint32_t));
                        if (!nb) { free(res); free(in); free(out); return -1; }
                        out = nb;
                        cap_out = nc;
                    }
                    out[nout++] = v;
                }
            }
        }
        /* swap worklists */
        uint32_t *t = in; in = out; out = t;
        size_t tc = cap_in; cap_in = cap_out; cap_out = tc;
        nin = nout;
    }
    free(res);
    free(in);
    free(out);
    return 0;
}

/* BetweennessCentralityHost, hybrid_bc_host.cpp:16-92: level-synchronous BFS
 * accumulating sigma (path counts), then dependencies from the deepest level
 * upward, delta(v) += sigma(v)/sigma(w) * (1 + delta(w)) over successors w,
 * for depths max-1 .. 1 (the source keeps 0), finally halved. */
int or_bc_single(uint32_t n, const uint32_t *row, const uint32_t *col, uint32_t source,
                 double *half_delta, double *sigma_out)
{
    if (source >= n) return -2;
    int32_t *depth = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    double *sigma = (double *)calloc((size_t)n, sizeof(double));
    double *delta = (double *)calloc((size_t)n, sizeof(double));
    uint32_t *order = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    if (!depth || !sigma || !delta || !order) {
        free(depth); free(sigma); free(delta); free(order);
        return -1;
    }
    for (uint32_t v = 0; v < n; v++) depth[v] = -1;
    depth[source] = 0;
    sigma[source] = 1.0;
    size_t head = 0, tail = 0;
    order[tail++] = source;
    while (head < tail) {
        uint32_t u = order[head++];
        for (uint32_t e = row[u]; e < row[u + 1]; e++) {
            uint32_t v = col[e];
            if (depth[v] == -1) {
                depth[v] = depth[u] + 1;
                order[tail++] = v;
            }
            if (depth[v] == depth[u] + 1) sigma[v] += sigma[u];
        }
    }
    /* BFS order is non-decreasing in depth: walk it backwards. */
    for (size_t i = tail; i-- > 0;) {
        uint32_t u = order[i];
        if (depth[u] == 0) continue;
        double acc = 0.0;
        for (uint32_t e = row[u]; e < row[u + 1]; e++) {
            uint32_t v = col[e];
            if (depth[v] == depth[u] + 1) acc += sigma[u] / sigma[v] * (1.0 + delta[v]);
        }
        delta[u] = acc;
    }
    if (half_delta)
        for (uint32_t v = 0; v < n; v++) half_delta[v] = 0.5 * delta[v];
    if (sigma_out) memcpy(sigma_out, sigma, (size_t)n * sizeof(double));
    free(depth);
    free(sigma);
    free(delta);
    free(order);
    return 0;
}

int or_bc(uint32_t n, const uint32_t *row, const uint32_t *col, const uint32_t *sources, int k,
          double *bc)
{
    double *hd = (double *)malloc((size_t)n * sizeof(double));
    if (!hd) return -1;
    for (uint32_t v = 0; v < n; v++) bc[v] = 0.0;
    for (int i = 0; i < k; i++) {
        int rc = or_bc_single(n, row, col, sources[i], hd, NULL);
        if (rc) { free(hd); return rc; }
        for (uint32_t v = 0; v < n; v++) bc[v] += hd[v];
    }
    free(hd);
    return 0;
}
