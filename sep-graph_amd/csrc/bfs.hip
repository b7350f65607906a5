// Hybrid (direction-switching) BFS as one persistent cooperative kernel.
//
// Reference path (samples/hybrid_bfs/hybrid_bfs.cu:173-243 driving
// Engine::Start, include/framework/framework.cuh:315-375): each round the host
// converts the frontier between a queue and a bitmap, counts it, sums its
// out-degree, launches AsyncPushDD (queue, CSR out-edges,
// include/framework/variants/async_push_dd.cuh:79-147) or AsyncPullDD (bitmap,
// CSC in-edges with early exit, async_pull_dd.cuh:26-70), and asks
// PolicyDecisionMaker::GetNextPolicy (include/framework/hybrid_policy.h:237-291)
// for the next direction -- at least six blocking host round trips per round.
//
// Here the whole traversal is one launch. Every block evaluates the same
// alpha/beta rule on device from counters read after a grid barrier, so no
// round needs the host. Frontier state per round r:
//   fbm[r%3]   bitmap of the frontier (always maintained; pull reads it)
//   q[r%2]     queue of frontier vertices with out-degree <= large_degree
//              (only produced by push rounds)
//   lq[r%2]    list of frontier vertices with out-degree > large_degree, each
//              with a prefix of its chunk count, so a push round splits their
//              adjacency into fixed-size chunks spread over every wave
//   ctr[r%3]   counters (queue lengths, striped count/scout sums)
// Push round: large-list chunks, then the small queue (or, right after a pull
// round, the frontier bitmap) with wave-level prefix-sum edge scheduling;
// claims go through atomicOr on the visited bitmap, so every vertex enters a
// frontier exactly once (the reference instead tolerates duplicates from a
// non-atomic min, hybrid_bfs.cu:128-134). Pull round: one wave per 64-vertex
// bitmap word; unvisited lanes scan CSC in-edges serially for a few edges,
// then wave-cooperatively, stopping at the first parent in the frontier.
#include "sg_internal.hpp"

using namespace sg;

namespace {

constexpr int kStripes = 16;
constexpr uint32_t kPullSerial = 8;  // in-edges scanned per lane before going wave-cooperative

struct alignas(128) Stripe {
    unsigned long long scout;
    unsigned long long count;
    uint32_t pad[28];
};

struct alignas(128) BfsCtr {
    uint32_t nsmall;
    uint32_t pad0;
    unsigned long long large;  // (entries << 32) | chunks
    uint32_t pad1[28];
    Stripe stripe[kStripes];
};
static_assert(sizeof(BfsCtr) == 128 * (1 + kStripes), "BfsCtr layout");

struct BfsArgs {
    DevGraph g;
    int32_t *level;
    uint64_t *vis;
    uint64_t *fbm;  // 3 * words
    uint32_t *q0, *q1;
    uint2 *lq0, *lq1;
    BfsCtr *ctr;    // 3 slots
    GridBarrier bar;
    uint32_t *abort;
    DevStats *stats;
    uint32_t source;
    uint32_t large_deg;
    uint32_t chunk;
    int32_t force;  // -1 auto, 0 push, 1 pull
    double alpha, beta;
    uint32_t total_workload0;
    uint32_t max_rounds;
};

__device__ __forceinline__ unsigned long long ld_relaxed64(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Zero one counter slot (called by one whole block).
__device__ void clear_ctr(BfsCtr *c)
{
    uint32_t *w = (uint32_t *)c;
    for (uint32_t i = threadIdx.x; i < sizeof(BfsCtr) / 4; i += blockDim.x) w[i] = 0;
}

struct Emit {
    uint32_t *qout;
    uint2 *lout;
    BfsCtr *nxt;
    uint32_t large_deg, chunk;
};

// Route newly discovered vertices of this wave into the next frontier's small
// queue or large list (one atomic per wave per list). Wave-convergent.
__device__ __forceinline__ void emit_wave(const Emit &em, bool found, uint32_t v, uint32_t dv, bool want_queue)
{
    const bool big = found && dv > em.large_deg;
    if (want_queue) wave_append(found && !big, v, em.qout, &em.nxt->nsmall);
    uint64_t bm = ballot(big);
    if (bm) {
        uint32_t nch = big ? (dv + em.chunk - 1) / em.chunk : 0;
        uint32_t incl = wave_incl_scan(nch);
        uint32_t tot = shfl(incl, kWave - 1);
        int leader = first_lane(bm);
        unsigned long long old = 0;
        if (lane() == leader)
            old = atomicAdd(&em.nxt->large, ((unsigned long long)__popcll(bm) << 32) | tot);
        old = shfl(old, leader);
        if (big) em.lout[(uint32_t)(old >> 32) + rank_in(bm)] = make_uint2(v, (uint32_t)old + incl - nch);
    }
}

// Claim v (not yet visited) for the next level; returns true for the unique winner.
__device__ __forceinline__ bool claim(uint64_t *vis, uint32_t v)
{
    const uint64_t bit = 1ull << (v & 63);
    unsigned long long *wp = (unsigned long long *)&vis[v >> 6];
    if (*wp & bit) return false;
    unsigned long long old = atomicOr(wp, (unsigned long long)bit);
    return !(old & bit);
}

template <int BS>
__global__ __launch_bounds__(BS) void k_bfs(BfsArgs a)
{
    constexpr int WPB = BS / kWave;
    const DevGraph &g = a.g;
    const uint32_t n = g.n, W = g.words;
    const uint32_t gtid = blockIdx.x * BS + threadIdx.x;
    const uint32_t nthreads = gridDim.x * BS;
    const uint32_t wid = threadIdx.x / kWave;
    const uint32_t gwave = blockIdx.x * WPB + wid;
    const uint32_t nwaves = gridDim.x * WPB;
    const int ln = lane();
    uint64_t *fbm0 = a.fbm, *fbm1 = a.fbm + W, *fbm2 = a.fbm + 2 * (size_t)W;

    __shared__ unsigned long long s_red[2][WPB];
    __shared__ unsigned long long s_lvl[4];

    // ---- prologue: reset per-query state, seed the source (one barrier)
    const uint32_t src = a.source;
    for (uint32_t i = gtid; i < n; i += nthreads) a.level[i] = (i == src) ? 0 : -1;
    for (uint32_t w = gtid; w < W; w += nthreads) {
        const uint64_t sb = (w == (src >> 6)) ? (1ull << (src & 63)) : 0ull;
        a.vis[w] = (g.inert ? g.inert[w] : 0ull) | sb;
        fbm0[w] = sb;
        fbm1[w] = 0;
        fbm2[w] = 0;
    }
    if (blockIdx.x == 0) {
        clear_ctr(&a.ctr[0]);
        clear_ctr(&a.ctr[1]);
        clear_ctr(&a.ctr[2]);
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t deg = g.row[src + 1] - g.row[src];
            if (deg > a.large_deg) {
                a.lq0[0] = make_uint2(src, 0u);
                a.ctr[0].large = (1ull << 32) | ((deg + a.chunk - 1) / a.chunk);
            } else {
                a.q0[0] = src;
                a.ctr[0].nsmall = 1;
            }
            a.ctr[0].stripe[0].scout = deg;
            a.ctr[0].stripe[0].count = 1;
        }
    }
    bool ok = grid_sync(a.bar, a.abort);

    // ---- replicated policy state (identical in every block)
    bool pull_prev = false;
    uint32_t total_wl = a.total_workload0;
    unsigned long long trace = 0, reached_edges = 0;
    uint32_t pushes = 0, pulls = 0, reached = 0;
    uint32_t L = 0;
    while (ok) {
        BfsCtr *cur = &a.ctr[L % 3];
        if (threadIdx.x == 0) {
            unsigned long long cnt = 0, sc = 0;
            for (int s = 0; s < kStripes; s++) {
                cnt += ld_relaxed64(&cur->stripe[s].count);
                sc += ld_relaxed64(&cur->stripe[s].scout);
            }
            s_lvl[0] = cnt;
            s_lvl[1] = sc;
            s_lvl[2] = __hip_atomic_load(&cur->nsmall, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_lvl[3] = ld_relaxed64(&cur->large);
        }
        __syncthreads();
        const unsigned long long count = s_lvl[0], scout = s_lvl[1];
        const uint32_t nsmall = (uint32_t)s_lvl[2];
        const uint32_t ne = (uint32_t)(s_lvl[3] >> 32), nc = (uint32_t)s_lvl[3];
        __syncthreads();
        if (count == 0) break;
        reached += (uint32_t)count;
        reached_edges += scout;
        if (a.max_rounds && L >= a.max_rounds) break;

        // Direction for this round: GetInitPolicy / GetNextPolicy with the
        // reference's uint32 workload counters.
        bool pull;
        if (a.force >= 0) {
            pull = a.force == 1;
        } else if (L == 0) {
            pull = false;
        } else if (pull_prev && (double)count > (double)n / a.beta) {
            pull = true;
        } else {
            const uint32_t out_wl = pull_prev ? 1u : (uint32_t)scout;
            pull = (double)out_wl > (double)total_wl / a.alpha;
        }
        if (!pull) total_wl -= (uint32_t)scout;
        if (pull) {
            pulls++;
            if (L < 64) trace |= 1ull << L;
        } else {
            pushes++;
        }

        // Slot and bitmap for round L+2 are free now: clear them.
        uint64_t *fcur = (L % 3 == 0) ? fbm0 : (L % 3 == 1) ? fbm1 : fbm2;
        uint64_t *fnext = (L % 3 == 0) ? fbm1 : (L % 3 == 1) ? fbm2 : fbm0;
        uint64_t *fclr = (L % 3 == 0) ? fbm2 : (L % 3 == 1) ? fbm0 : fbm1;
        if (blockIdx.x == 0) clear_ctr(&a.ctr[(L + 2) % 3]);
        for (uint32_t w = gtid; w < W; w += nthreads) fclr[w] = 0;

        BfsCtr *nxt = &a.ctr[(L + 1) % 3];
        const uint32_t *qin = (L & 1) ? a.q1 : a.q0;
        const uint2 *lin = (L & 1) ? a.lq1 : a.lq0;
        Emit em{(L & 1) ? a.q0 : a.q1, (L & 1) ? a.lq0 : a.lq1, nxt, a.large_deg, a.chunk};
        const int32_t next_level = (int32_t)L + 1;
        unsigned long long my_scout = 0, my_count = 0;

        if (!pull) {
            // (1) large-degree vertices: fixed-size chunks spread over all waves
            for (uint32_t c = gwave; c < nc; c += nwaves) {
                // 64-ary search for the last entry whose chunk base is <= c
                // (entries are sorted by base; entry 0 has base 0).
                uint32_t lo = 0, hi = ne;
                while (hi - lo > 1) {
                    const uint32_t step = (hi - lo + kWave - 1) / kWave;
                    const uint32_t idx = lo + (uint32_t)ln * step;
                    const uint64_t le = ballot(idx < hi && lin[idx].y <= c);
                    const uint32_t k = (uint32_t)__popcll(le) - 1;
                    lo = lo + k * step;
                    hi = min(hi, lo + step);
                }
                const uint2 ent = lin[lo];
                const uint32_t u = ent.x;
                const uint32_t beg = g.row[u] + (c - ent.y) * a.chunk;
                const uint32_t end = min(g.row[u + 1], beg + a.chunk);
                for (uint32_t e0 = beg; e0 < end; e0 += kWave) {
                    const uint32_t e = e0 + ln;
                    const bool act = e < end;
                    const uint32_t v = act ? g.col[e] : 0u;
                    bool found = act && claim(a.vis, v);
                    uint32_t dv = 0;
                    if (found) {
                        a.level[v] = next_level;
                        atomicOr((unsigned long long *)&fnext[v >> 6], 1ull << (v & 63));
                        dv = g.row[v + 1] - g.row[v];
                        my_scout += dv;
                        my_count++;
                    }
                    emit_wave(em, found, v, dv, true);
                }
            }
            // (2) small-degree frontier: from the queue after a push round, from
            //     the bitmap after a pull round (large ones are in the list).
            const bool from_bitmap = pull_prev;
            const uint32_t nitems = from_bitmap ? W : (nsmall + kWave - 1) / kWave;
            for (uint32_t it = gwave; it < nitems; it += nwaves) {
                uint32_t u = 0;
                bool valid;
                if (from_bitmap) {
                    const uint64_t fw = fcur[it];
                    if (!fw) continue;  // wave-uniform
                    u = it * 64 + ln;
                    valid = (fw >> ln) & 1;
                } else {
                    const uint32_t i = it * kWave + ln;
                    valid = i < nsmall;
                    if (valid) u = qin[i];
                }
                uint32_t b = 0, d = 0;
                if (valid) {
                    b = g.row[u];
                    d = g.row[u + 1] - b;
                    if (from_bitmap && d > a.large_deg) d = 0;  // handled via the large list
                }
                const uint32_t incl = wave_incl_scan(d);
                const uint32_t total = shfl(incl, kWave - 1);
                const uint32_t off = b - (incl - d);
                for (uint32_t jb = 0; jb < total; jb += kWave) {
                    const uint32_t j = jb + ln;
                    const bool act = j < total;
                    int k = 0;
#pragma unroll
                    for (int step = 32; step > 0; step >>= 1) {
                        const uint32_t t = shfl(incl, k + step - 1);
                        if (t <= j) k += step;
                    }
                    const uint32_t e = shfl(off, k) + j;
                    const uint32_t v = act ? g.col[e] : 0u;
                    bool found = act && claim(a.vis, v);
                    uint32_t dv = 0;
                    if (found) {
                        a.level[v] = next_level;
                        atomicOr((unsigned long long *)&fnext[v >> 6], 1ull << (v & 63));
                        dv = g.row[v + 1] - g.row[v];
                        my_scout += dv;
                        my_count++;
                    }
                    emit_wave(em, found, v, dv, true);
                }
            }
        } else {
            // Pull: each wave owns one 64-vertex word of the visited bitmap.
            for (uint32_t w = gwave; w < W; w += nwaves) {
                const uint64_t vw = a.vis[w];
                if (vw == ~0ull) continue;  // wave-uniform
                const uint32_t v = w * 64 + ln;
                const bool want = v < n && !((vw >> ln) & 1);
                bool found = false;
                uint32_t b = 0, e = 0;
                if (want) {
                    b = g.cstart[v];
                    e = g.cstart[v + 1];
                    const uint32_t lim = min(e, b + kPullSerial);
                    for (; b < lim; b++) {
                        const uint32_t u = g.csrc[b];
                        if ((fcur[u >> 6] >> (u & 63)) & 1) {
                            found = true;
                            break;
                        }
                    }
                }
                uint64_t rest = ballot(want && !found && b < e);
                while (rest) {
                    const int t = first_lane(rest);
                    rest &= rest - 1;
                    const uint32_t tb = shfl(b, t), te = shfl(e, t);
                    bool hit = false;
                    for (uint32_t x = tb; x < te; x += kWave) {
                        const uint32_t y = x + ln;
                        bool h = false;
                        if (y < te) {
                            const uint32_t u = g.csrc[y];
                            h = (fcur[u >> 6] >> (u & 63)) & 1;
                        }
                        if (ballot(h)) {
                            hit = true;
                            break;
                        }
                    }
                    if (ln == t) found = hit;
                }
                const uint64_t fm = ballot(found);
                if (!fm) continue;
                if (ln == 0) {
                    a.vis[w] = vw | fm;
                    fnext[w] = fm;
                }
                uint32_t dv = 0;
                if (found) {
                    a.level[v] = next_level;
                    dv = g.row[v + 1] - g.row[v];
                    my_scout += dv;
                }
                if (ln == 0) my_count += __popcll(fm);
                emit_wave(em, found, v, dv, false);
            }
        }

        // Block-reduce count and scout, one atomic per block into a stripe.
        my_scout = wave_sum_u64(my_scout);
        my_count = wave_sum_u64(my_count);
        if (ln == 0) {
            s_red[0][wid] = my_scout;
            s_red[1][wid] = my_count;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long sc = 0, cn = 0;
            for (int i = 0; i < WPB; i++) {
                sc += s_red[0][i];
                cn += s_red[1][i];
            }
            Stripe &st = nxt->stripe[blockIdx.x % kStripes];
            if (sc) atomicAdd(&st.scout, sc);
            if (cn) atomicAdd(&st.count, cn);
        }
        pull_prev = pull;
        L++;
        ok = grid_sync(a.bar, a.abort);
    }

    if (gtid == 0) {
        DevStats s;
        s.rounds = L;
        s.push_rounds = pushes;
        s.pull_rounds = pulls;
        s.reached = reached;
        s.reached_edges = reached_edges;
        s.trace = trace;
        s.max_depth = L ? L - 1 : 0;
        s.status = ok ? SG_OK : SG_ERR_TIMEOUT;
        s.iterations = L;
        s.pad = 0;
        s.residual = 0.0;
        *a.stats = s;
    }
}

template <int BS>
int launch_bfs(sg_engine *e, BfsArgs &args)
{
    if (e->coop_blocks == 0) {
        int per_cu = 0;
        SG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bfs<BS>, BS, 0));
        if (per_cu < 1) SG_FAIL(SG_ERR_HIP, "sg_bfs: kernel cannot be resident");
        int blocks = per_cu * e->g->num_cus;
        const int max_blocks = (int)((e->bar_words_len / kBarLine - 3) * kBarGroup);
        e->coop_blocks = std::min(blocks, max_blocks);
    }
    args.bar.nblocks = (uint32_t)e->coop_blocks;
    void *params[] = {&args};
    SG_HIP(hipLaunchCooperativeKernel((const void *)k_bfs<BS>, dim3(e->coop_blocks), dim3(BS), params, 0,
                                      e->stream));
    return SG_OK;
}

}  // namespace

static_assert(sizeof(BfsCtr) * 3 <= kCtrBytes, "BFS counters exceed the engine's counter area");

extern "C" int sg_bfs(sg_engine *e, uint32_t source)
{
    if (!e) SG_FAIL(SG_ERR_INVALID, "sg_bfs: null engine");
    const DevGraph &d = e->g->dev;
    if (source >= d.n) SG_FAIL(SG_ERR_RANGE, "sg_bfs: source >= n");
    int rc = engine_ensure_bfs_state(e);
    if (rc) return rc;
    if (e->needs_bar_reset && (rc = engine_reset_barrier(e))) return rc;
    const sg_engine_options &o = e->opt;
    BfsArgs a{};
    a.g = d;
    a.level = (int32_t *)e->value;
    a.vis = e->bitmaps;
    a.fbm = e->bitmaps + d.words;
    a.q0 = e->queues;
    a.q1 = e->queues + d.n;
    a.lq0 = e->large;
    a.lq1 = e->large + e->large_cap;
    a.ctr = (BfsCtr *)e->ctr;
    a.bar.words = e->bar_words;
    a.bar.spin_limit = 1u << 24;
    a.abort = e->abort_flag;
    a.stats = e->stats;
    a.source = source;
    a.large_deg = o.large_degree > 0 ? (uint32_t)o.large_degree : 256u;
    a.chunk = 256;
    if (o.alpha == 0.0 || o.force_variant == SG_ASYNC_PUSH_DD || o.force_variant == SG_SYNC_PUSH_DD)
        a.force = 0;
    else if (o.force_variant == SG_ASYNC_PULL_DD || o.force_variant == SG_SYNC_PULL_DD)
        a.force = 1;
    else if (o.force_variant == SG_VARIANT_AUTO)
        a.force = -1;
    else
        SG_FAIL(SG_ERR_UNSUPPORTED, "sg_bfs: only data-driven push/pull variants apply to BFS");
    a.alpha = o.alpha == 0.0 ? 1.0 : o.alpha;
    a.beta = o.beta;
    a.total_workload0 = (uint32_t)((double)d.m * o.edge_factor);
    a.max_rounds = (uint32_t)o.max_iteration;
    e->last_algo = 0;
    SG_HIP(hipEventRecord(e->ev0, e->stream));
    const int bs = o.block_size ? o.block_size : 512;
    switch (bs) {
    case 256: rc = launch_bfs<256>(e, a); break;
    case 512: rc = launch_bfs<512>(e, a); break;
    case 1024: rc = launch_bfs<1024>(e, a); break;
    default: SG_FAIL(SG_ERR_UNSUPPORTED, "sg_bfs: block_size must be 256, 512 or 1024");
    }
    if (rc) return rc;
    SG_HIP(hipEventRecord(e->ev1, e->stream));
    e->timed = true;
    return SG_OK;
}
