// Betweenness centrality (Brandes) for up to 64 sources per launch, one
// persistent kernel per batch.
//
// Reference path: samples/hybrid_bc/hybrid_bc.cu. Stage 1 (the BC app,
// hybrid_bc.cu:23-131) runs the hybrid BFS engine accumulating path counts
// sigma and pushes each level's vertices onto a WorklistStack
// (include/framework/worklist_stack.cuh:16-111); stage 2 (StageTwoDDFused,
// hybrid_bc.cu:134-225) walks the stack back level by level adding
// sigma(v)/sigma(w) (1 + delta(w)) over successors w; the result is halved
// (hybrid_bc.cu:374-376). The reference handles one source per run, with
// sigma and delta in FP32.
//
// Here a batch of B <= 64 sources shares every traversal of the adjacency:
//   * each vertex carries 64-bit masks, bit s for source s: `seen` and one
//     mask per BFS level (a ring of three dense n-word arrays);
//   * per-source state is laid out source-minor, [vertex][64], so a wave
//     handles one vertex with lane s = source s and reads a neighbour's row
//     (512 B of sigma or psi) as one coalesced access;
//   * the forward pass is level-synchronous: discovery of the next level's
//     masks (push over the level's out-edges with atomicOr, or pull over
//     in-edges of vertices not yet seen by every source, switched by the
//     alpha rule on the level's out-degree), then sigma of every newly
//     reached (vertex, source) pair gathered from its predecessors on that
//     level (FP64, no atomics except for in-hubs split into chunks);
//   * every (vertex, level-mask) pair is pushed on a level stack, so the
//     backward pass visits each vertex once per distinct depth it has over
//     the batch;
//   * the backward pass stores psi(w) = (1 + delta(w)) / sigma(w) per
//     (vertex, source), so a predecessor v needs one load per successor:
//     delta(v) = sigma(v) * sum of psi(w) over successors w. delta is FP64 in
//     registers; bc(v) += delta(v) / 2 accumulates in FP64 over batches and
//     is written out as FP32.
#include "frontier.hpp"

using namespace sg;

namespace {

constexpr uint32_t kMaxLevels = 1u << 20;

// First element of source s's result row: rows are n elements, and `orow`
// (when given) maps a batch's sources onto rows of a larger result pool.
__device__ __forceinline__ size_t out_row(const uint32_t *orow, uint32_t s, uint32_t n)
{
    return (size_t)(orow ? orow[s] : s) * n;
}

struct BcArgs {
    DevGraph g;
    const uint32_t *src;  // sources of this batch
    uint32_t nsrc;
    uint64_t *seen;       // [n]
    uint64_t *F;          // 3 x n: level masks, level L in F + (L % 3) n
    double *sigma;        // [n][64]
    double *psi;          // [n][64]
    double *bc;           // [n], FP64 accumulator over batches
    float *out;           // [n], written by the last batch
    float *dep;           // optional: 0.5 delta_s(v) in row out_row(orow, s) of n floats
    const uint32_t *orow; // optional result row of each source (null: row s)
    uint32_t *stk_v;      // level stack: vertex
    uint64_t *stk_m;      // level stack: sources with that vertex at this level
    uint32_t cap;
    uint32_t *lvl;        // stack start of each level [kMaxLevels + 2]
    uint32_t *ctl;        // [0] stack length, [1] overflow
    FrontierCtr *ctr;     // 3 slots: per-level entry count / out-degree sum
    GridBarrier bar;
    uint32_t *abort;
    DevStats *stats;
    DevTrace *trace;
    double alpha, beta;
    double total_wl;
    int32_t force;
    uint32_t first, last;
    uint32_t probe;
};

__device__ __forceinline__ uint64_t wave_or_u64(uint64_t x)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) x |= __shfl_xor(x, d, kWave);
    return x;
}

__device__ __forceinline__ uint64_t seg_or(uint64_t s, int k)
{
    const int ln = lane();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint64_t y = __shfl_up(s, d, kWave);
        const int ko = __shfl_up(k, d, kWave);
        if (ln >= d && ko == k) s |= y;
    }
    return s;
}

// OR over in-edges [b, e) of each lane's vertex of Fc[csrc[edge]], the wave
// walking the concatenated lists 64 edges per step. A lane stops once its OR
// holds every bit of `need` (the reference's early break, generalised to a
// set of sources: with one source it stops at the first parent).
__device__ __forceinline__ uint64_t wave_gather_or(const DevGraph &g, const uint64_t *Fc, uint32_t b, uint32_t e,
                                                   uint64_t need)
{
    const int ln = lane();
    uint64_t acc = 0;
    for (;;) {
        const uint32_t r = e - b;
        const uint32_t incl = wave_incl_scan(r);
        const uint32_t total = shfl(incl, kWave - 1);
        if (total == 0) break;
        const uint32_t excl = incl - r;
        const uint32_t j = (uint32_t)ln;
        const int k = owner_lane(incl, j);
        const uint32_t bk = shfl(b, k), xk = shfl(excl, k);
        uint64_t cand = 0;
        if (j < total) cand = Fc[g.csrc[bk + (j - xk)]];
        const uint64_t s = seg_or(cand, k);
        const uint32_t lo = min(excl, 64u), hi = min(incl, 64u), cov = hi - lo;
        const uint64_t seg = __shfl(s, (int)(hi > 0 ? hi - 1 : 0), kWave);
        if (cov) {
            acc |= seg;
            b += cov;
            if ((acc & need) == need) b = e;
        }
    }
    return acc;
}

// Sum, in lane s, of row[u][s] over the edges [b, e) (edge ids into `adj`)
// whose neighbour u has bit s in (M[u] & mask). Whole wave, uniform b/e/mask.
__device__ __forceinline__ double masked_row_sum(const uint32_t *adj, const uint64_t *M, const double *row,
                                                 uint32_t b, uint32_t e, uint64_t mask)
{
    const int ln = lane();
    double acc = 0.0;
    for (uint32_t e0 = b; e0 < e; e0 += kWave) {
        const uint32_t ed = e0 + ln;
        uint32_t u = 0;
        uint64_t fm = 0;
        if (ed < e) {
            u = adj[ed];
            fm = M[u] & mask;
        }
        uint64_t mm = ballot(fm != 0);
        uint64_t srcs = wave_or_u64(fm);
        if (__popcll(srcs) < __popcll(mm)) {
            // fewer sources than matching neighbours (small batches): per
            // source, every lane loads its neighbour's value and the wave sums
            while (srcs) {
                const int s = first_lane(srcs);
                srcs &= srcs - 1;
                double x = ((fm >> s) & 1) ? row[(size_t)u * 64 + s] : 0.0;
                x = wave_sum_f64(x);
                if (ln == s) acc += x;
            }
            continue;
        }
        while (mm) {
            const int j = first_lane(mm);
            mm &= mm - 1;
            const uint64_t fj = __shfl(fm, j, kWave);
            const uint32_t uj = shfl(u, j);
            if ((fj >> ln) & 1) acc += row[(size_t)uj * 64 + ln];
        }
    }
    return acc;
}

// Discovery of the next level's masks: Fn[w] gets the bits of sources that
// reach w for the first time through a level-L vertex. Push expands the
// level's entries (seg_v/seg_m, len) over out-edges with atomicOr; pull has
// every vertex not yet seen by all sources OR its in-neighbours' level masks.
// Whole grid; `seen` holds every level up to L.
__device__ __forceinline__ void ms_discover(const DevGraph &g, const uint64_t *seen, const uint64_t *Fc, uint64_t *Fn,
                                            const uint32_t *seg_v, const uint64_t *seg_m, uint32_t len, uint64_t bm,
                                            bool pull, uint32_t gwave, uint32_t nwaves)
{
    const int ln = lane();
    const uint32_t n = g.n, W = g.words;
    if (!pull) {
        for (uint32_t c0 = 0; c0 < g.nhuge; c0 += nwaves * kWave) {
            const uint32_t c = c0 + (uint32_t)ln * nwaves + gwave;
            uint2 dsc = make_uint2(0u, 0u);
            uint64_t mk = 0;
            if (c < g.nhuge) {
                dsc = g.huge[c];
                mk = Fc[dsc.x];
            }
            uint64_t mm = ballot(mk != 0);
            while (mm) {
                const int t = first_lane(mm);
                mm &= mm - 1;
                const uint32_t u = shfl(dsc.x, t), beg = shfl(dsc.y, t);
                const uint64_t mu = __shfl(mk, t, kWave);
                const uint32_t end = min(g.row[u + 1], beg + kHugeChunk);
                for (uint32_t e0 = beg; e0 < end; e0 += kWave) {
                    const uint32_t e = e0 + ln;
                    if (e < end) {
                        const uint32_t w = g.col[e];
                        const uint64_t nm = mu & ~seen[w];
                        if (nm && (Fn[w] & nm) != nm) atomicOr((unsigned long long *)&Fn[w], (unsigned long long)nm);
                    }
                }
            }
        }
        const uint32_t per = (len + nwaves - 1) / nwaves;
        const uint32_t qb = min(len, gwave * per), qe = min(len, gwave * per + per);
        for (uint32_t i0 = qb; i0 < qe; i0 += kWave) {
            const uint32_t i = i0 + ln;
            uint32_t b = 0, d = 0;
            uint64_t mu = 0;
            if (i < qe) {
                const uint32_t u = seg_v[i];
                mu = seg_m[i];
                b = g.row[u];
                d = g.row[u + 1] - b;
                if (d > kHugeDegree) d = 0;
            }
            const uint32_t incl = wave_incl_scan(d);
            const uint32_t total = shfl(incl, kWave - 1);
            const uint32_t off = b - (incl - d);
            for (uint32_t jb = 0; jb < total; jb += kWave) {
                const uint32_t j = jb + ln;
                const int k = owner_lane(incl, j);
                const uint32_t e = shfl(off, k) + j;
                const uint64_t mk = __shfl(mu, k, kWave);
                if (j < total) {
                    const uint32_t w = g.col[e];
                    const uint64_t nm = mk & ~seen[w];
                    if (nm && (Fn[w] & nm) != nm) atomicOr((unsigned long long *)&Fn[w], (unsigned long long)nm);
                }
            }
        }
    } else {
        for (uint32_t c = gwave; c < g.nhugein; c += nwaves) {
            const uint2 dsc = g.hugein[c];
            const uint64_t unseen = bm & ~seen[dsc.x];
            if (!unseen) continue;
            const uint32_t end = min(g.cstart[dsc.x + 1], dsc.y + kPullChunk);
            uint64_t o = 0;
            for (uint32_t e0 = dsc.y; e0 < end; e0 += kWave) {
                const uint32_t e = e0 + ln;
                if (e < end) o |= Fc[g.csrc[e]];
            }
            o = wave_or_u64(o) & unseen;
            if (ln == 0 && o) atomicOr((unsigned long long *)&Fn[dsc.x], (unsigned long long)o);
        }
        for (uint32_t w = gwave; w < W; w += nwaves) {
            const uint32_t v = w * 64 + ln;
            uint32_t b = 0, e = 0;
            uint64_t unseen = 0;
            if (v < n) {
                unseen = bm & ~seen[v];
                b = g.cstart[v];
                e = g.cstart[v + 1];
                if (!unseen || e - b > kPullHubDegree) b = e;
            }
            const uint64_t o = wave_gather_or(g, Fc, b, e, unseen) & unseen;
            if (o) Fn[v] = o;
        }
    }
}

// Append (v, m) for lanes with `pred` to the pair list (vv, mm) whose length
// is *count (one atomic per wave); past `cap` entries *overflow is set.
__device__ __forceinline__ void append_pair(uint32_t *count, uint32_t *vv, uint64_t *mm, uint32_t cap,
                                            uint32_t *overflow, bool pred, uint32_t v, uint64_t m)
{
    const uint64_t bm = ballot(pred);
    if (!bm) return;
    const int leader = first_lane(bm);
    uint32_t base = 0;
    if (lane() == leader) base = atomicAdd(count, (uint32_t)__popcll(bm));
    base = shfl(base, leader);
    if (pred) {
        const uint32_t pos = base + rank_in(bm);
        if (pos < cap) {
            vv[pos] = v;
            mm[pos] = m;
        } else {
            *overflow = 1;
        }
    }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_bc(BcArgs a)
{
    if (a.probe) {
        residency_probe(a.bar.words, 1u << 20);
        return;
    }
    constexpr int WPB = BS / kWave;
    const DevGraph &g = a.g;
    const uint32_t n = g.n, W = g.words;
    const uint32_t gtid = blockIdx.x * BS + threadIdx.x;
    const uint32_t nthreads = gridDim.x * BS;
    const uint32_t gwave = blockIdx.x * WPB + threadIdx.x / kWave;
    const uint32_t nwaves = gridDim.x * WPB;
    const int ln = lane();
    const uint32_t xcc = xcc_id();
    BarState gen = bar_begin(a.bar, xcc);
    const long long t_start = wall_clock64();
    const uint64_t bm = a.nsrc >= 64 ? ~0ull : ((1ull << a.nsrc) - 1);

    __shared__ unsigned long long s_red[2][WPB];
    __shared__ unsigned long long s_lvl[5];
    __shared__ uint32_t s_len;

    // ---- prologue: clear masks; zero the rows hub chunks accumulate into
    for (uint32_t v = gtid; v < n; v += nthreads) {
        a.seen[v] = 0;
        a.F[v] = 0;
        a.F[(size_t)n + v] = 0;
        a.F[2 * (size_t)n + v] = 0;
        if (a.first) a.bc[v] = 0.0;
    }
    if (a.dep)
        for (uint32_t s = 0; s < a.nsrc; s++) {
            float *row = a.dep + out_row(a.orow, s, n);
            for (uint32_t v = gtid; v < n; v += nthreads) row[v] = 0.0f;
        }
    for (uint32_t h = gwave; h < g.nhubin; h += nwaves) a.sigma[(size_t)g.hugein[g.hubin[h].x].x * 64 + ln] = 0.0;
    for (uint32_t c = gwave; c < g.nhuge; c += nwaves) a.psi[(size_t)g.huge[c].x * 64 + ln] = 0.0;
    if (blockIdx.x == 0) {
        clear_ctr(&a.ctr[0]);
        clear_ctr(&a.ctr[1]);
        clear_ctr(&a.ctr[2]);
        if (threadIdx.x == 0) {
            a.ctl[0] = 0;
            a.ctl[1] = 0;
        }
    }
    bool ok = grid_sync(a.bar, xcc, gen, a.abort);
    if (ok && blockIdx.x == 0 && threadIdx.x < kWave) {
        // sources: lane s owns source s; a vertex listed twice gets both bits
        const bool has = (uint32_t)ln < a.nsrc;
        const uint32_t v = has ? a.src[ln] : 0u;
        uint64_t m = 0;
        int firstl = 64;
        for (int t = 0; t < 64; t++) {
            const uint32_t vt = shfl(v, t);
            if (has && (uint32_t)t < a.nsrc && vt == v) {
                m |= 1ull << t;
                firstl = min(firstl, t);
            }
        }
        const bool lead = has && firstl == ln;
        if (has) a.sigma[(size_t)v * 64 + ln] = 1.0;
        if (lead) {
            a.seen[v] = m;
            a.F[v] = m;
        }
        append_pair(&a.ctl[0], a.stk_v, a.stk_m, a.cap, &a.ctl[1], lead, v, m);
        const uint32_t dv = lead ? g.row[v + 1] - g.row[v] : 0u;
        const unsigned long long sc = wave_sum_u64(dv);
        const unsigned long long cn = wave_sum_u64(lead ? 1ull : 0ull);
        if (ln == 0) {
            a.ctr[0].stripe[0].scout = sc;
            a.ctr[0].stripe[0].count = cn;
            a.lvl[0] = 0;
        }
    }
    if (ok) ok = grid_sync(a.bar, xcc, gen, a.abort);

    // ---- forward: levels 0, 1, ...
    uint32_t L = 0, prev_beg = 0, seg_beg = 0, seg_end = 0;
    uint32_t pushes = 0, pulls = 0;
    unsigned long long ptrace = 0;
    DirectionPolicy pol{a.alpha, a.beta, (uint32_t)min(a.total_wl, 4294967295.0), n, a.force};
    while (ok) {
        read_counts(&a.ctr[L % 3], s_lvl);
        if (threadIdx.x == 0) s_len = __hip_atomic_load(&a.ctl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const RoundCounts rc = unpack_counts(s_lvl, 0xFFFFFFFFu);
        seg_end = min(s_len, a.cap);
        __syncthreads();
        if (rc.count == 0 || L + 1 >= kMaxLevels) break;
        if (gtid == 0) a.lvl[L + 1] = seg_end;
        const bool pull = pol.next(L, rc.count, rc.scout);
        if (pull) {
            pulls++;
            if (L < 64) ptrace |= 1ull << L;
        } else {
            pushes++;
        }
        if (gtid == 0) write_trace(a.trace, L, pull, rc.count, rc.scout, t_start);
        if (blockIdx.x == 0) clear_ctr(&a.ctr[(L + 2) % 3]);
        const uint64_t *Fc = a.F + (size_t)(L % 3) * n;
        uint64_t *Fn = a.F + (size_t)((L + 1) % 3) * n;
        uint64_t *Ff = a.F + (size_t)((L + 2) % 3) * n;
        // level L-1's masks are no longer read: clear them for level L+2
        for (uint32_t i = prev_beg + gtid; i < seg_beg; i += nthreads) Ff[a.stk_v[i]] = 0;

        // (1) discovery of level L+1
        ms_discover(g, a.seen, Fc, Fn, a.stk_v + seg_beg, a.stk_m + seg_beg, seg_end - seg_beg, bm, pull, gwave, nwaves);
        ok = grid_sync(a.bar, xcc, gen, a.abort);
        if (!ok) break;

        // (2) the new level: seen, stack, sigma from predecessors on level L
        unsigned long long scout = 0, cnt = 0;
        for (uint32_t w = gwave; w < W; w += nwaves) {
            const uint32_t v = w * 64 + ln;
            uint64_t nw = 0;
            uint32_t b = 0, e = 0;
            if (v < n) {
                nw = Fn[v];
                if (nw) {
                    a.seen[v] |= nw;
                    b = g.cstart[v];
                    e = g.cstart[v + 1];
                    scout += g.row[v + 1] - g.row[v];
                    cnt++;
                }
            }
            append_pair(&a.ctl[0], a.stk_v, a.stk_m, a.cap, &a.ctl[1], nw != 0, v, nw);
            uint64_t cand = ballot(nw != 0 && e - b <= kPullHubDegree);
            while (cand) {
                const int t = first_lane(cand);
                cand &= cand - 1;
                const uint32_t vt = w * 64 + t;
                const uint64_t mt = __shfl(nw, t, kWave);
                const double s = masked_row_sum(g.csrc, Fc, a.sigma, shfl(b, t), shfl(e, t), mt);
                if ((mt >> ln) & 1) a.sigma[(size_t)vt * 64 + ln] = s;
            }
        }
        for (uint32_t c = gwave; c < g.nhugein; c += nwaves) {
            const uint2 dsc = g.hugein[c];
            const uint64_t nw = Fn[dsc.x];
            if (!nw) continue;
            const uint32_t end = min(g.cstart[dsc.x + 1], dsc.y + kPullChunk);
            const double s = masked_row_sum(g.csrc, Fc, a.sigma, dsc.y, end, nw);
            if (s != 0.0) atomicAdd(&a.sigma[(size_t)dsc.x * 64 + ln], s);
        }
        publish_counts<WPB>(&a.ctr[(L + 1) % 3], scout, cnt, false, s_red);
        pol.pull_prev = pull;
        prev_beg = seg_beg;
        seg_beg = seg_end;
        L++;
        ok = grid_sync(a.bar, xcc, gen, a.abort);
    }
    // levels 0 .. D hold entries; lvl[D + 1] is the stack length
    const uint32_t D = L ? L - 1 : 0;
    const uint32_t levels = L;

    // ---- backward: levels D .. 1
    for (uint32_t l = D; ok && l >= 1; l--) {
        const uint32_t sb = a.lvl[l], se = a.lvl[l + 1];
        const uint64_t *An = a.F + (size_t)((l + 1) % 3) * n;  // level l+1
        const uint64_t *Ac = a.F + (size_t)(l % 3) * n;        // level l
        uint64_t *Af = a.F + (size_t)((l + 2) % 3) * n;        // level l+2, then l-1
        if (l + 2 <= D) {
            const uint32_t fb = a.lvl[l + 2], fe = a.lvl[l + 3];
            for (uint32_t i = fb + gtid; i < fe; i += nthreads) Af[a.stk_v[i]] = 0;
        }
        for (uint32_t i = sb + gwave; i < se; i += nwaves) {
            const uint32_t v = a.stk_v[i];
            const uint32_t rb = g.row[v], re = g.row[v + 1];
            if (re - rb > kHugeDegree) continue;
            const uint64_t mv = a.stk_m[i];
            const double acc = masked_row_sum(g.col, An, a.psi, rb, re, mv);
            double dl = 0.0;
            if ((mv >> ln) & 1) {
                const double sg = a.sigma[(size_t)v * 64 + ln];
                dl = sg * acc;
                a.psi[(size_t)v * 64 + ln] = (1.0 + dl) / sg;
                if (a.dep) a.dep[out_row(a.orow, ln, n) + v] = (float)(0.5 * dl);
            }
            dl = wave_sum_f64(dl);
            if (ln == 0) a.bc[v] += 0.5 * dl;
        }
        for (uint32_t c = gwave; c < g.nhuge; c += nwaves) {
            const uint2 dsc = g.huge[c];
            const uint64_t mv = Ac[dsc.x];
            if (!mv) continue;
            const uint32_t end = min(g.row[dsc.x + 1], dsc.y + kHugeChunk);
            const double acc = masked_row_sum(g.col, An, a.psi, dsc.y, end, mv);
            if (acc != 0.0) atomicAdd(&a.psi[(size_t)dsc.x * 64 + ln], acc);
        }
        ok = grid_sync(a.bar, xcc, gen, a.abort);
        if (!ok) break;
        for (uint32_t i = sb + gwave; i < se; i += nwaves) {
            const uint32_t v = a.stk_v[i];
            if (g.row[v + 1] - g.row[v] <= kHugeDegree) continue;
            const uint64_t mv = a.stk_m[i];
            double dl = 0.0;
            if ((mv >> ln) & 1) {
                const size_t k = (size_t)v * 64 + ln;
                const double sg = a.sigma[k];
                dl = sg * a.psi[k];
                a.psi[k] = (1.0 + dl) / sg;
                if (a.dep) a.dep[out_row(a.orow, ln, n) + v] = (float)(0.5 * dl);
            }
            dl = wave_sum_f64(dl);
            if (ln == 0) a.bc[v] += 0.5 * dl;
        }
        {
            const uint32_t pb = a.lvl[l - 1], pe = a.lvl[l];
            for (uint32_t i = pb + gtid; i < pe; i += nthreads) Af[a.stk_v[i]] = a.stk_m[i];
        }
        ok = grid_sync(a.bar, xcc, gen, a.abort);
    }
    if (ok && a.last) {
        for (uint32_t v = gtid; v < n; v += nthreads) a.out[v] = (float)a.bc[v];
    }
    if (gtid == 0) {
        DevStats s = a.first ? DevStats{} : *a.stats;
        s.rounds += levels;
        s.push_rounds += pushes;
        s.pull_rounds += pulls;
        s.trace = ptrace;
        s.max_depth = max(s.max_depth, D);
        s.iterations += 1;
        const bool overflow = a.ctl[1] != 0;
        if (s.status == SG_OK) s.status = !ok ? SG_ERR_TIMEOUT : (overflow || levels + 1 >= kMaxLevels) ? SG_ERR_CAPACITY : SG_OK;
        *a.stats = s;
    }
}

// ------------------------------------------------------------ multi-source BFS
// Levels of up to 64 independent BFS queries in one launch: the discovery of
// the BC forward pass (shared masks, one adjacency sweep per level for every
// query) with each newly reached (vertex, source) pair writing its level into
// that source's own int32 row, so each query's result is a contiguous n-vector
// equal to a single-source BFS from it. Level entries live in a two-half ring
// (each vertex appears at most once per level).
struct MsBfsArgs {
    DevGraph g;
    const uint32_t *src;
    uint32_t nsrc;
    uint64_t *seen;       // [n]
    uint64_t *F;          // 3 x n level masks
    int32_t *out;         // levels of source s in row out_row(orow, s), -1 when unreached
    const uint32_t *orow;
    uint32_t *ring_v;     // 2 x n: entries of level L in half L % 2
    uint64_t *ring_m;
    uint32_t *ctl;        // [1] overflow, [2 + h] entries in ring half h
    FrontierCtr *ctr;
    GridBarrier bar;
    uint32_t *abort;
    DevStats *stats;
    DevTrace *trace;
    double alpha, beta;
    double total_wl;
    int32_t force;
    uint32_t max_rounds;
    uint32_t probe;
};

template <int BS>
__global__ __launch_bounds__(BS) void k_msbfs(MsBfsArgs a)
{
    if (a.probe) {
        residency_probe(a.bar.words, 1u << 20);
        return;
    }
    constexpr int WPB = BS / kWave;
    const DevGraph &g = a.g;
    const uint32_t n = g.n, W = g.words;
    const uint32_t gtid = blockIdx.x * BS + threadIdx.x;
    const uint32_t nthreads = gridDim.x * BS;
    const uint32_t gwave = blockIdx.x * WPB + threadIdx.x / kWave;
    const uint32_t nwaves = gridDim.x * WPB;
    const int ln = lane();
    const uint32_t xcc = xcc_id();
    BarState gen = bar_begin(a.bar, xcc);
    const long long t_start = wall_clock64();
    const uint64_t bm = a.nsrc >= 64 ? ~0ull : ((1ull << a.nsrc) - 1);

    __shared__ unsigned long long s_red[2][WPB];
    __shared__ unsigned long long s_lvl[5];
    __shared__ uint32_t s_len;

    for (uint32_t v = gtid; v < n; v += nthreads) {
        a.seen[v] = 0;
        a.F[v] = 0;
        a.F[(size_t)n + v] = 0;
        a.F[2 * (size_t)n + v] = 0;
    }
    for (uint32_t s = 0; s < a.nsrc; s++) {
        int32_t *row = a.out + out_row(a.orow, s, n);
        for (uint32_t v = gtid; v < n; v += nthreads) row[v] = -1;
    }
    if (blockIdx.x == 0) {
        clear_ctr(&a.ctr[0]);
        clear_ctr(&a.ctr[1]);
        clear_ctr(&a.ctr[2]);
        if (threadIdx.x == 0) {
            a.ctl[1] = 0;
            a.ctl[2] = 0;
            a.ctl[3] = 0;
        }
    }
    bool ok = grid_sync(a.bar, xcc, gen, a.abort);
    if (ok && blockIdx.x == 0 && threadIdx.x < kWave) {
        const bool has = (uint32_t)ln < a.nsrc;
        const uint32_t v = has ? a.src[ln] : 0u;
        uint64_t m = 0;
        int firstl = 64;
        for (int t = 0; t < 64; t++) {
            const uint32_t vt = shfl(v, t);
            if (has && (uint32_t)t < a.nsrc && vt == v) {
                m |= 1ull << t;
                firstl = min(firstl, t);
            }
        }
        const bool lead = has && firstl == ln;
        if (has) a.out[out_row(a.orow, ln, n) + v] = 0;
        if (lead) {
            a.seen[v] = m;
            a.F[v] = m;
        }
        append_pair(&a.ctl[2], a.ring_v, a.ring_m, n, &a.ctl[1], lead, v, m);
        const uint32_t dv = lead ? g.row[v + 1] - g.row[v] : 0u;
        const unsigned long long sc = wave_sum_u64(dv);
        const unsigned long long cn = wave_sum_u64(lead ? 1ull : 0ull);
        if (ln == 0) {
            a.ctr[0].stripe[0].scout = sc;
            a.ctr[0].stripe[0].count = cn;
        }
    }
    if (ok) ok = grid_sync(a.bar, xcc, gen, a.abort);

    uint32_t L = 0, prev_len = 0;
    uint32_t pushes = 0, pulls = 0;
    unsigned long long ptrace = 0;
    DirectionPolicy pol{a.alpha, a.beta, (uint32_t)min(a.total_wl, 4294967295.0), n, a.force};
    while (ok) {
        read_counts(&a.ctr[L % 3], s_lvl);
        if (threadIdx.x == 0) s_len = __hip_atomic_load(&a.ctl[2 + (L & 1)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const RoundCounts rc = unpack_counts(s_lvl, 0xFFFFFFFFu);
        const uint32_t len = min(s_len, n);
        __syncthreads();
        if (rc.count == 0 || (a.max_rounds && L >= a.max_rounds)) break;
        // the other half held level L-1, whose count every block read a level ago
        if (gtid == 0) a.ctl[2 + ((L + 1) & 1)] = 0;
        const bool pull = pol.next(L, rc.count, rc.scout);
        if (pull) {
            pulls++;
            if (L < 64) ptrace |= 1ull << L;
        } else {
            pushes++;
        }
        if (gtid == 0) write_trace(a.trace, L, pull, rc.count, rc.scout, t_start);
        if (blockIdx.x == 0) clear_ctr(&a.ctr[(L + 2) % 3]);
        const uint64_t *Fc = a.F + (size_t)(L % 3) * n;
        uint64_t *Fn = a.F + (size_t)((L + 1) % 3) * n;
        uint64_t *Ff = a.F + (size_t)((L + 2) % 3) * n;
        const size_t half = (size_t)(L & 1) * n, other = (size_t)((L + 1) & 1) * n;
        for (uint32_t i = gtid; i < prev_len; i += nthreads) Ff[a.ring_v[other + i]] = 0;

        ms_discover(g, a.seen, Fc, Fn, a.ring_v + half, a.ring_m + half, len, bm, pull, gwave, nwaves);
        ok = grid_sync(a.bar, xcc, gen, a.abort);
        if (!ok) break;

        unsigned long long scout = 0, cnt = 0;
        for (uint32_t w = gwave; w < W; w += nwaves) {
            const uint32_t v = w * 64 + ln;
            uint64_t nw = 0;
            if (v < n) {
                nw = Fn[v];
                if (nw) {
                    a.seen[v] |= nw;
                    scout += g.row[v + 1] - g.row[v];
                    cnt++;
                }
            }
            append_pair(&a.ctl[2 + ((L + 1) & 1)], a.ring_v + other, a.ring_m + other, n, &a.ctl[1], nw != 0, v, nw);
            // one pass per source present in this word: lanes write consecutive vertices
            uint64_t any = wave_or_u64(nw);
            while (any) {
                const int s = __ffsll((unsigned long long)any) - 1;
                any &= any - 1;
                if ((nw >> s) & 1) a.out[out_row(a.orow, s, n) + v] = (int32_t)(L + 1);
            }
        }
        publish_counts<WPB>(&a.ctr[(L + 1) % 3], scout, cnt, false, s_red);
        pol.pull_prev = pull;
        prev_len = len;
        L++;
        ok = grid_sync(a.bar, xcc, gen, a.abort);
    }
    if (gtid == 0) {
        DevStats s = *a.stats;
        s.rounds += L;
        s.push_rounds += pushes;
        s.pull_rounds += pulls;
        s.trace = ptrace;
        s.max_depth = max(s.max_depth, L ? L - 1 : 0u);
        s.iterations += 1;
        if (s.status == SG_OK) s.status = ok ? (a.ctl[1] ? SG_ERR_CAPACITY : SG_OK) : SG_ERR_TIMEOUT;
        *a.stats = s;
    }
}

// ----------------------------------------------------------- multi-source SSSP
// Up to 64 independent SSSP queries in one launch. Distances live as rows
// dist[v][64] (lane q = query q), so relaxing an edge for every query whose
// frontier holds its tail is one coalesced row access and at most one
// atomicMin per improving lane. A round pushes from every vertex whose mask
// (queries that lowered its distance last round) is non-zero; a vertex enters
// the next round's entry list when the first bit of its next mask is set
// (the atomicOr returns 0). Results are written transposed, one contiguous
// n-vector per query, equal to sg_sssp from that source.
struct MsSsspArgs {
    DevGraph g;
    const uint32_t *src;
    uint32_t nsrc;
    uint32_t *dist;       // [n][64]
    uint64_t *F;          // 3 x n masks, round r in F + (r % 3) n
    uint32_t *out;        // distances of source q in row out_row(orow, q)
    const uint32_t *orow;
    uint32_t *ring;       // 3 x n: entries of round r in third r % 3
    uint32_t *ctl;        // [1] overflow, [2 + t] entries in third t
    FrontierCtr *ctr;     // 3 slots: count / out-degree of each round's entries
    GridBarrier bar;
    uint32_t *abort;
    DevStats *stats;
    DevTrace *trace;
    uint32_t max_rounds;
    uint32_t probe;
};

struct MsRelax {
    const MsSsspArgs *a;
    uint64_t *Fn;
    uint32_t *ring_next;
    uint32_t *cnt_next;
    uint32_t qbits;  // queries of the batch fit in 2^qbits lanes
    unsigned long long scout, count;
};

// Edge steps a group of lanes keeps in flight in the batched relaxation.
#ifndef SG_MS_RELAX_STEPS
#define SG_MS_RELAX_STEPS 16
#endif
constexpr int kMsRelaxSteps = SG_MS_RELAX_STEPS;

// Relax the out-edges [b, e) of u for the queries in `mask`. The wave is
// split into 64 / 2^qbits groups of 2^qbits lanes (lane q of a group = query
// q); each group takes kMsRelaxSteps edges per step, with the loads and
// atomics of all of them issued before any result is used, so a batch of B
// queries keeps about 64 / B x kMsRelaxSteps edges in flight per wave. A
// group's leader publishes the queries whose distance it lowered into the
// next round's mask of the edge's head; the wave queues the heads whose mask
// it was first to set, with one append per wave and step.
__device__ __forceinline__ void ms_relax_range(MsRelax &R, uint32_t u, uint32_t b, uint32_t e, uint64_t mask)
{
    constexpr int U = kMsRelaxSteps;
    const MsSsspArgs &a = *R.a;
    const DevGraph &g = a.g;
    const int ln = lane();
    const uint32_t qb = R.qbits;
    const uint32_t q = (uint32_t)ln & ((1u << qb) - 1), grp = (uint32_t)ln >> qb, ngrp = 64u >> qb;
    const uint64_t gmask = qb == 6 ? ~0ull : ((1ull << (1u << qb)) - 1);
    const bool mine = (mask >> q) & 1;
    const uint32_t du = mine ? a.dist[(size_t)u * 64 + q] : 0xFFFFFFFFu;
    for (uint32_t e0 = b; e0 < e; e0 += ngrp * U) {
        uint32_t v[U], nd[U], cur[U];
        bool act[U];
#pragma unroll
        for (int s = 0; s < U; s++) {
            const uint32_t ed = e0 + grp + (uint32_t)s * ngrp;
            act[s] = ed < e;
            v[s] = act[s] ? g.col[ed] : 0u;
            const uint32_t w = (act[s] && mine) ? (g.w ? g.w[ed] : 1u) : 0u;
            nd[s] = du > 0xFFFFFFFFu - w ? 0xFFFFFFFFu : du + w;
            act[s] = act[s] && mine;
        }
#pragma unroll
        for (int s = 0; s < U; s++) cur[s] = act[s] ? a.dist[(size_t)v[s] * 64 + q] : 0u;
        // every atomic of the step is issued before any result is read (a
        // result consumed inside its own branch makes each one wait in turn)
        uint32_t old[U];
#pragma unroll
        for (int s = 0; s < U; s++) {
            old[s] = 0xFFFFFFFFu;
            if (act[s] && nd[s] < cur[s]) old[s] = atomicMin(&a.dist[(size_t)v[s] * 64 + q], nd[s]);
        }
        unsigned long long fo[U];
#pragma unroll
        for (int s = 0; s < U; s++) {
            const bool imp = act[s] && nd[s] < cur[s] && nd[s] < old[s];
            const uint64_t im = ballot(imp);
            const uint64_t gm = (im >> (grp << qb)) & gmask;
            fo[s] = ~0ull;
            if (q == 0 && gm) fo[s] = atomicOr((unsigned long long *)&R.Fn[v[s]], (unsigned long long)gm);
        }
        // heads whose mask this wave set first: one append for the whole step
        uint64_t fm[U];
        uint32_t tot = 0;
#pragma unroll
        for (int s = 0; s < U; s++) {
            fm[s] = ballot(fo[s] == 0);
            tot += (uint32_t)__popcll(fm[s]);
        }
        if (!tot) continue;
        uint32_t base = 0;
        if (ln == 0) base = atomicAdd(R.cnt_next, tot);
        base = shfl(base, 0);
#pragma unroll
        for (int s = 0; s < U; s++) {
            if (fo[s] == 0) {
                const uint32_t pos = base + (uint32_t)rank_in(fm[s]);
                if (pos < g.n) R.ring_next[pos] = v[s];
                else a.ctl[1] = 1;
                R.scout += g.row[v[s] + 1] - g.row[v[s]];
                R.count++;
            }
            base += (uint32_t)__popcll(fm[s]);
        }
    }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_mssssp(MsSsspArgs a)
{
    if (a.probe) {
        residency_probe(a.bar.words, 1u << 20);
        return;
    }
    constexpr int WPB = BS / kWave;
    const DevGraph &g = a.g;
    const uint32_t n = g.n;
    const uint32_t gtid = blockIdx.x * BS + threadIdx.x;
    const uint32_t nthreads = gridDim.x * BS;
    const uint32_t gwave = blockIdx.x * WPB + threadIdx.x / kWave;
    const uint32_t nwaves = gridDim.x * WPB;
    const int ln = lane();
    const uint32_t xcc = xcc_id();
    BarState gen = bar_begin(a.bar, xcc);
    const long long t_start = wall_clock64();
    uint32_t qbits = 0;
    while ((1u << qbits) < a.nsrc) qbits++;

    __shared__ unsigned long long s_red[2][WPB];
    __shared__ unsigned long long s_lvl[5];
    __shared__ uint32_t s_len;

    for (size_t i = gtid; i < (size_t)n * 64; i += nthreads) a.dist[i] = 0xFFFFFFFFu;
    for (size_t i = gtid; i < 3 * (size_t)n; i += nthreads) a.F[i] = 0;
    if (blockIdx.x == 0) {
        clear_ctr(&a.ctr[0]);
        clear_ctr(&a.ctr[1]);
        clear_ctr(&a.ctr[2]);
        if (threadIdx.x == 0) {
            a.ctl[1] = 0;
            a.ctl[2] = 0;
            a.ctl[3] = 0;
            a.ctl[4] = 0;
        }
    }
    bool ok = grid_sync(a.bar, xcc, gen, a.abort);
    if (ok && blockIdx.x == 0 && threadIdx.x < kWave) {
        const bool has = (uint32_t)ln < a.nsrc;
        const uint32_t v = has ? a.src[ln] : 0u;
        uint64_t m = 0;
        int firstl = 64;
        for (int t = 0; t < 64; t++) {
            const uint32_t vt = shfl(v, t);
            if (has && (uint32_t)t < a.nsrc && vt == v) {
                m |= 1ull << t;
                firstl = min(firstl, t);
            }
        }
        const bool lead = has && firstl == ln;
        if (has) a.dist[(size_t)v * 64 + ln] = 0;
        if (lead) a.F[v] = m;
        const uint64_t lm = ballot(lead);
        if (lead) a.ring[rank_in(lm)] = v;
        const uint32_t dv = lead ? g.row[v + 1] - g.row[v] : 0u;
        const unsigned long long sc = wave_sum_u64(dv);
        if (ln == 0) {
            a.ctl[2] = (uint32_t)__popcll(lm);
            a.ctr[0].stripe[0].scout = sc;
            a.ctr[0].stripe[0].count = (unsigned long long)__popcll(lm);
        }
    }
    if (ok) ok = grid_sync(a.bar, xcc, gen, a.abort);

    uint32_t r = 0, prev_len = 0;
    while (ok) {
        read_counts(&a.ctr[r % 3], s_lvl);
        if (threadIdx.x == 0) s_len = __hip_atomic_load(&a.ctl[2 + r % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const RoundCounts rc = unpack_counts(s_lvl, 0xFFFFFFFFu);
        const uint32_t len = min(s_len, n);
        __syncthreads();
        if (rc.count == 0 || (a.max_rounds && r >= a.max_rounds)) break;
        if (gtid == 0) {
            a.ctl[2 + (r + 2) % 3] = 0;  // held round r-1, read by every block a round ago
            write_trace(a.trace, r, false, rc.count, rc.scout, t_start);
        }
        if (blockIdx.x == 0) clear_ctr(&a.ctr[(r + 2) % 3]);
        const uint64_t *Fc = a.F + (size_t)(r % 3) * n;
        uint64_t *Ff = a.F + (size_t)((r + 2) % 3) * n;
        const uint32_t *cur = a.ring + (size_t)(r % 3) * n;
        const uint32_t *old = a.ring + (size_t)((r + 2) % 3) * n;
        for (uint32_t i = gtid; i < prev_len; i += nthreads) Ff[old[i]] = 0;

        MsRelax R{&a, a.F + (size_t)((r + 1) % 3) * n, a.ring + (size_t)((r + 1) % 3) * n, &a.ctl[2 + (r + 1) % 3],
                  qbits, 0, 0};
        // hubs through the static chunk list
        for (uint32_t c = gwave; c < g.nhuge; c += nwaves) {
            const uint2 dsc = g.huge[c];
            const uint64_t mk = Fc[dsc.x];
            if (mk) ms_relax_range(R, dsc.x, dsc.y, min(g.row[dsc.x + 1], dsc.y + kHugeChunk), mk);
        }
        for (uint32_t i = gwave; i < len; i += nwaves) {
            const uint32_t u = cur[i];
            const uint32_t b = g.row[u], e = g.row[u + 1];
            if (e - b > kHugeDegree) continue;
            ms_relax_range(R, u, b, e, Fc[u]);
        }
        publish_counts<WPB>(&a.ctr[(r + 1) % 3], R.scout, R.count, false, s_red);
        prev_len = len;
        r++;
        ok = grid_sync(a.bar, xcc, gen, a.abort);
    }
    // results, one contiguous row per query
    if (ok) {
        for (uint32_t v = gtid; v < n; v += nthreads) {
            const uint32_t *row = a.dist + (size_t)v * 64;
            for (uint32_t q = 0; q < a.nsrc; q++) a.out[out_row(a.orow, q, n) + v] = row[q];
        }
    }
    if (gtid == 0) {
        DevStats s = *a.stats;
        s.rounds += r;
        s.push_rounds += r;
        s.max_depth = max(s.max_depth, r);
        s.iterations += 1;
        if (s.status == SG_OK) s.status = ok ? (a.ctl[1] ? SG_ERR_CAPACITY : SG_OK) : SG_ERR_TIMEOUT;
        *a.stats = s;
    }
}

}  // namespace


// ------------------------------------------------------------------ host side

namespace {

constexpr uint64_t al256(uint64_t x) { return (x + 255) / 256 * 256; }

// Offsets of each batched kind's working arrays inside the engine's arena.
struct MsLayout {
    uint64_t off[10];
    uint64_t total;
};

MsLayout ms_layout(int kind, uint64_t n)
{
    MsLayout L{};
    uint64_t o = 0;
    auto put = [&](int i, uint64_t bytes) {
        L.off[i] = o;
        o += al256(bytes);
    };
    if (kind == kMsBfs) {
        put(0, n * 8);      // seen
        put(1, 3 * n * 8);  // F
        put(2, 2 * n * 4);  // ring_v
        put(3, 2 * n * 8);  // ring_m
        put(4, 256);        // ctl
    } else if (kind == kMsSssp) {
        put(0, n * 64 * 4); // dist rows
        put(1, 3 * n * 8);  // F
        put(2, 3 * n * 4);  // ring
        put(4, 256);        // ctl
    } else {
        const uint64_t cap = std::min<uint64_t>(n * 64 + 4096, 0xFFFFFFF0ull);
        put(0, n * 8);                             // seen
        put(1, 3 * n * 8);                         // F
        put(2, n * 64 * 8);                        // sigma
        put(3, n * 64 * 8);                        // psi
        put(5, n * 8);                             // bc accumulator
        put(6, cap * 4);                           // stack vertices
        put(7, cap * 8);                           // stack masks
        put(8, (uint64_t)(kMaxLevels + 4) * 4);    // level starts
        put(4, 256);                               // ctl
        L.off[9] = cap;
    }
    L.total = o;
    return L;
}

template <typename Args>
void fill_common(sg_engine *e, Args &a)
{
    a.g = e->g->dev;
    a.ctr = (FrontierCtr *)e->ctr;
    a.bar.spin_limit = 1u << 22;
    a.abort = e->abort_flag;
    a.stats = e->stats;
    a.trace = e->trace;
}

template <typename Args>
int launch_grid(sg_engine *e, const void *fn, int bs, Args &a)
{
    sg_engine::PGrid *pg = nullptr;
    auto probe = [e, a, fn, bs](int blocks, uint32_t *words) -> int {
        Args p = a;
        p.probe = 1;
        p.bar.words = words;
        void *params[] = {&p};
        SG_HIP(hipLaunchKernel(fn, dim3(blocks), dim3(bs), params, 0, e->stream));
        return SG_OK;
    };
    int rc = engine_persistent_grid(e, fn, bs, probe, &pg);
    if (rc) return rc;
    bind_barrier(pg, a.bar);
    return engine_launch_persistent(e, [&]() -> int {
        void *params[] = {&a};
        SG_HIP(hipLaunchKernel(fn, dim3(pg->blocks), dim3(bs), params, 0, e->stream));
        pg->launches++;
        return SG_OK;
    });
}

}  // namespace

namespace sg {

int ms_prepare(sg_engine *e, int kind)
{
    int rc = engine_ensure_bfs_state(e);
    if (rc) return rc;
    return engine_ensure_aux(e, ms_layout(kind, e->g->dev.n).total);
}

int ms_launch(sg_engine *e, int kind, const uint32_t *dev_src, uint32_t nsrc, void *out, const uint32_t *dev_orow,
              bool first, bool last)
{
    if (nsrc < 1 || nsrc > 64) SG_FAIL(SG_ERR_RANGE, "batched launch: 1..64 sources");
    const uint64_t n = e->g->dev.n;
    const MsLayout L = ms_layout(kind, n);
    if (!e->aux || e->aux_bytes < L.total) SG_FAIL(SG_ERR_STATE, "batched launch: arena not prepared");
    char *base = (char *)e->aux;
    const sg_engine_options &o = e->opt;
    const int force = forced_direction(o);
    if (force < -1) SG_FAIL(SG_ERR_UNSUPPORTED, "batched launch: only data-driven push/pull variants apply");
    if (o.lb_push == SG_LB_NONE || o.lb_pull == SG_LB_NONE)
        SG_FAIL(SG_ERR_UNSUPPORTED, "batched launch: LoadBalancing::NONE applies to sg_bfs and sg_sssp only");
    const int bs = o.block_size ? o.block_size : 1024;
    if (bs != 256 && bs != 512 && bs != 1024) SG_FAIL(SG_ERR_UNSUPPORTED, "batched launch: block_size must be 256, 512 or 1024");
    const double alpha = o.alpha == 0.0 ? 1.0 : o.alpha;
    const double total_wl = (double)e->g->dev.m * o.edge_factor;
    if (kind == kMsBfs) {
        MsBfsArgs a{};
        fill_common(e, a);
        a.src = dev_src;
        a.nsrc = nsrc;
        a.seen = (uint64_t *)(base + L.off[0]);
        a.F = (uint64_t *)(base + L.off[1]);
        a.ring_v = (uint32_t *)(base + L.off[2]);
        a.ring_m = (uint64_t *)(base + L.off[3]);
        a.ctl = (uint32_t *)(base + L.off[4]);
        a.out = (int32_t *)out;
        a.orow = dev_orow;
        a.alpha = alpha;
        a.beta = o.beta;
        a.total_wl = total_wl;
        a.force = force;
        a.max_rounds = (uint32_t)o.max_iteration;
        const void *fn = bs == 256 ? (const void *)k_msbfs<256> : bs == 512 ? (const void *)k_msbfs<512> : (const void *)k_msbfs<1024>;
        return launch_grid(e, fn, bs, a);
    }
    if (kind == kMsSssp) {
        MsSsspArgs a{};
        fill_common(e, a);
        a.src = dev_src;
        a.nsrc = nsrc;
        a.dist = (uint32_t *)(base + L.off[0]);
        a.F = (uint64_t *)(base + L.off[1]);
        a.ring = (uint32_t *)(base + L.off[2]);
        a.ctl = (uint32_t *)(base + L.off[4]);
        a.out = (uint32_t *)out;
        a.orow = dev_orow;
        a.max_rounds = (uint32_t)o.max_iteration;
        const void *fn = bs == 256 ? (const void *)k_mssssp<256> : bs == 512 ? (const void *)k_mssssp<512> : (const void *)k_mssssp<1024>;
        return launch_grid(e, fn, bs, a);
    }
    BcArgs a{};
    fill_common(e, a);
    a.src = dev_src;
    a.nsrc = nsrc;
    a.seen = (uint64_t *)(base + L.off[0]);
    a.F = (uint64_t *)(base + L.off[1]);
    a.sigma = (double *)(base + L.off[2]);
    a.psi = (double *)(base + L.off[3]);
    a.bc = (double *)(base + L.off[5]);
    a.stk_v = (uint32_t *)(base + L.off[6]);
    a.stk_m = (uint64_t *)(base + L.off[7]);
    a.lvl = (uint32_t *)(base + L.off[8]);
    a.cap = (uint32_t)L.off[9];
    a.ctl = (uint32_t *)(base + L.off[4]);
    if (kind == kMsBcDep) {
        a.dep = (float *)out;
        a.orow = dev_orow;
        a.out = nullptr;
        a.last = 0;
    } else {
        a.out = (float *)out;
        a.last = last;
    }
    a.first = first;
    a.alpha = alpha;
    a.beta = o.beta;
    a.total_wl = total_wl;
    a.force = force;
    const void *fn = bs == 256 ? (const void *)k_bc<256> : bs == 512 ? (const void *)k_bc<512> : (const void *)k_bc<1024>;
    return launch_grid(e, fn, bs, a);
}

}  // namespace sg

namespace {

// C-ABI driver: k host sources in batches of `batch`, results in the value
// array (row i for sources[i], or the BC sum).
int ms_run(sg_engine *e, int kind, const uint32_t *sources, int32_t k, int32_t batch, const char *name)
{
    if (!e || !sources) SG_FAIL(SG_ERR_INVALID, std::string(name) + ": null argument");
    if (k < 1) SG_FAIL(SG_ERR_RANGE, std::string(name) + ": need at least one source");
    if (batch == 0) batch = 64;
    if (batch < 1 || batch > 64) SG_FAIL(SG_ERR_RANGE, std::string(name) + ": batch must be in [1, 64]");
    const DevGraph &d = e->g->dev;
    for (int32_t i = 0; i < k; i++)
        if (sources[i] >= d.n) SG_FAIL(SG_ERR_RANGE, std::string(name) + ": source >= n");
    if (forced_direction(e->opt) < -1)
        SG_FAIL(SG_ERR_UNSUPPORTED, std::string(name) + ": only data-driven push/pull variants apply");
    if (e->opt.lb_push == SG_LB_NONE || e->opt.lb_pull == SG_LB_NONE)
        SG_FAIL(SG_ERR_UNSUPPORTED, std::string(name) + ": LoadBalancing::NONE applies to sg_bfs and sg_sssp only");
    const bool single_bc = batch == 1 && (kind == kMsBcDep || kind == kMsBcSum);
    int rc = single_bc ? engine_ensure_bfs_state(e) : ms_prepare(e, kind);
    if (rc) return rc;
    const uint64_t n = d.n;
    const uint64_t rows = kind == kMsBcSum ? 1 : (uint64_t)k;
    void *res = nullptr;
    if ((rc = engine_result_rows(e, rows, &res))) return rc;
    if ((rc = engine_ensure_srcbuf(e, (uint64_t)k))) return rc;
    std::vector<uint32_t> dsrc(k);
    for (int32_t i = 0; i < k; i++) dsrc[i] = dev_vertex(e->g, sources[i]);
    // the source buffer may still be read by an earlier launch on the stream
    SG_HIP(hipStreamSynchronize(e->stream));
    SG_HIP(hipMemcpy(e->srcbuf, dsrc.data(), (size_t)k * 4, hipMemcpyHostToDevice));
    SG_HIP(hipMemsetAsync(e->stats, 0, sizeof(DevStats), e->stream));
    e->last_algo = 3 + kind;
    SG_HIP(hipEventRecord(e->ev0, e->stream));
    if (single_bc) {
        // one source per launch: the single-source kernel (no bit-packed rows)
        for (int32_t i = 0; i < k; i++) {
            float *dep = kind == kMsBcDep ? (float *)res + (size_t)i * n : nullptr;
            float *sum = kind == kMsBcSum ? (float *)res : nullptr;
            if ((rc = bc1_launch(e, dsrc[i], dep, nullptr, sum, i == 0, i == k - 1))) return rc;
        }
        SG_HIP(hipEventRecord(e->ev1, e->stream));
        if ((rc = engine_finish(e, rows))) return rc;
        e->timed = true;
        return SG_OK;
    }
    const int nb = (k + batch - 1) / batch;
    for (int i = 0; i < nb; i++) {
        const uint32_t ns = (uint32_t)std::min(batch, k - i * batch);
        void *out = kind == kMsBcSum ? res : (void *)((uint32_t *)res + (size_t)i * batch * n);
        rc = ms_launch(e, kind, e->srcbuf + (size_t)i * batch, ns, out, nullptr, i == 0, i == nb - 1);
        if (rc) return rc;
    }
    SG_HIP(hipEventRecord(e->ev1, e->stream));
    if ((rc = engine_finish(e, rows))) return rc;
    e->timed = true;
    return SG_OK;
}

}  // namespace

extern "C" int sg_bc(sg_engine *e, const uint32_t *sources, int32_t k, int32_t batch)
{
    return ms_run(e, kMsBcSum, sources, k, batch, "sg_bc");
}

extern "C" int sg_bc_dependency(sg_engine *e, const uint32_t *sources, int32_t k, int32_t batch)
{
    return ms_run(e, kMsBcDep, sources, k, batch, "sg_bc_dependency");
}

extern "C" int sg_bfs_multi(sg_engine *e, const uint32_t *sources, int32_t k, int32_t batch)
{
    return ms_run(e, kMsBfs, sources, k, batch, "sg_bfs_multi");
}

extern "C" int sg_sssp_multi(sg_engine *e, const uint32_t *sources, int32_t k, int32_t batch)
{
    return ms_run(e, kMsSssp, sources, k, batch, "sg_sssp_multi");
}
