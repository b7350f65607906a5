// Device side of the hybrid (direction-switching) BFS: one query runs as a
// persistent grid (a whole launch, or the blocks of one XCD).
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
// round needs the host. Per vertex, a push round writes the level and two
// bits:
//   * discovery: an edge whose head shows an unset bit in the visited filter
//     claims that bit (one returning atomicOr per run of a wave's candidates
//     in one word, none per edge); the claimer stores level L+1 and queues the
//     vertex. The reference's default non-atomic BFS min
//     (hybrid_bfs.cu:128-134, hybrid_bfs/main.cpp:37) lets racing writers
//     queue a vertex several times; claiming keeps the queue exact, which
//     matters on high-diameter graphs where duplicates compound round after
//     round. Candidates also set the next frontier's bitmap, so visited and
//     frontier tests read 512 KiB of L2-resident bits, never the level array;
//   * the next queue is staged in LDS per block and published with one
//     atomic per block per round (same-address atomics serialise on MI355X);
//   * vertices with out-degree > kHugeDegree are expanded through the graph's
//     static chunk list, filtered by membership, so hub adjacencies spread
//     over every wave without building any list at run time.
// A pull round sweeps 64-vertex words (dealt round-robin over the waves,
// kPullGroup words at a time): every unvisited lane tests its first
// in-neighbour from the graph's head array, then the group's words walk their
// remaining in-edge lists in lock step, 64 edges per word per step with all
// of the group's loads in flight together, re-planning after each step so a
// vertex stops at its first parent (the reference's early break,
// pull_functor.h:138-188). Pull rounds write whole words of the next frontier
// bitmap and of a visited filter, so they need no atomics.
#pragma once

#include "frontier.hpp"

using namespace sg;

// Words of 64 vertices a wave works on at once in a pull round.
#ifndef SG_PULL_GROUP
#define SG_PULL_GROUP 2
#endif
// Minimum waves per SIMD the compiler must allow for k_bfs (0: no bound);
// 8 caps registers at 64 so two 1024-thread blocks fit on a CU.
#ifndef SG_BFS_WAVES
#define SG_BFS_WAVES 0
#endif

namespace {
namespace bfsk {

constexpr int kPullGroup = SG_PULL_GROUP;

struct BfsArgs {
    DevGraph g;
    // levels in the caller's vertex order: device vertex v's level goes to
    // level[perm[v]] (perm null: the orders agree). Nothing reads it during
    // the traversal (visited tests use the filter), so a degree-ordered graph
    // needs no device-order copy and no gather at the end.
    int32_t *level;
    const uint32_t *perm;
    uint64_t *vis;    // W words: visited filter (set bit => visited or in-degree 0)
    uint64_t *vis_idle;  // the other filter, reset here for the next query
    uint64_t *fbit;   // 3 * W words: frontier L's bitmap in slot L % 3
    uint32_t *queue;  // 3 * qcap entries
    uint32_t qcap;
    FrontierCtr *ctr; // 3 slots
    GridBarrier bar;
    uint32_t *abort;
    DevStats *stats;
    DevTrace *trace;
    uint32_t source;     // device vertex id
    uint32_t source_id;  // the caller's id of the same vertex
    int32_t force;    // -1 auto, 0 push, 1 pull
    double alpha, beta;
    uint32_t total_workload0;
    uint32_t max_rounds;
    uint32_t probe;   // launch is a residency probe of the grid
    uint32_t full_init;  // bitmaps and counters not known to be clean
    // LoadBalancing::NONE (one lane per vertex, its edges in order, no hub
    // chunks) for push / pull rounds; otherwise the balanced schedulers
    uint32_t lb_push_none, lb_pull_none;
};

// Hub chunks whose heads a block stages in LDS at a time (with 64-edge
// chunks: 4096 heads, 16 KiB), and the mark of a staged slot past a chunk's end.
constexpr uint32_t kStageChunks = 64;
// Member chunks a block lists before staging them (16 KiB of LDS).
constexpr uint32_t kHubList = 2048;
constexpr uint32_t kNoHead = 0xFFFFFFFFu;
static_assert(kHugeChunk <= kWave, "a staged hub chunk is expanded by one wave in one step");

// Largest source out-degree whose expansion (round 0) block 0 does alone.
constexpr uint32_t kLocalRound0 = 4096;

__device__ __forceinline__ uint32_t caller_id(const BfsArgs &a, uint32_t v) { return a.perm ? a.perm[v] : v; }

struct Round {
    const BfsArgs *a;
    int32_t L;              // level being expanded
    const uint64_t *fcur;   // frontier L as a bitmap
    uint64_t *fnext;        // frontier L+1, built by this round
    Stager S;
    unsigned long long scout, count;  // this lane's share of the next frontier
    bool huge;                        // this lane discovered a vertex above kHugeDegree
    bool levels = true;               // write discovered levels now (else the next round does)
};

__device__ __forceinline__ bool in_frontier(const Round &R, uint32_t u) { return test_bit(R.fcur, u); }

// Push edge -> v for `act` lanes; wave-convergent.
__device__ __forceinline__ void push_visit(Round &R, bool act, uint32_t v)
{
    const BfsArgs &a = *R.a;
    // candidates: not visited as far as this lane can see (bits set by other
    // waves in this round may not be visible yet). Claiming the bit in the
    // visited filter picks one discoverer; every candidate was unvisited
    // before this round, so all of them belong in the next frontier's bitmap.
    // Visited tests here and membership tests in a pull round read 512 KiB
    // of L2-resident bits instead of the level array, with one atomic per
    // run of a wave's candidates in one word, none per edge.
    const bool cand = act && !test_bit(a.vis, v);
    const bool found = wave_claim_bits(a.vis, cand, v, R.fnext);
    if (found) {
        if (R.levels) a.level[caller_id(a, v)] = R.L + 1;
        const uint32_t dv = a.g.row[v + 1] - a.g.row[v];
        R.scout += dv;
        R.count++;
        R.huge |= dv > kHugeDegree;
    }
    stage(R.S, found, v);
}

// Expand the out-edges of up to 64 vertices (one per lane) with wave-level
// prefix-sum scheduling: 64 consecutive edges of the concatenated adjacency
// per step, so CSR reads are coalesced and lanes stay busy. Vertices above
// kHugeDegree are left to the chunk list.
__device__ __forceinline__ void push_expand(Round &R, bool valid, uint32_t u)
{
    const DevGraph &g = R.a->g;
    uint32_t b = 0, d = 0;
    if (valid) {
        b = g.row[u];
        d = g.row[u + 1] - b;
        if (d > kHugeDegree) d = 0;
    }
    const uint32_t incl = wave_incl_scan(d);
    const uint32_t total = shfl(incl, kWave - 1);
    const uint32_t off = b - (incl - d);
    const int ln = lane();
    for (uint32_t jb = 0; jb < total; jb += kWave) {
        const uint32_t j = jb + ln;
        const bool act = j < total;
        const int k = owner_lane(incl, j);
        const uint32_t e = shfl(off, k) + j;
        push_visit(R, act, act ? g.col[e] : 0u);
    }
}

// LoadBalancing::NONE push: each lane walks its own vertex's out-edges in
// CSR order, hubs included (the reference's thread-per-vertex relax,
// async_push_dd.cuh:24-78); the wave steps through the longest list.
__device__ __forceinline__ void push_expand_serial(Round &R, bool valid, uint32_t u)
{
    const DevGraph &g = R.a->g;
    uint32_t b = 0, d = 0;
    if (valid) {
        b = g.row[u];
        d = g.row[u + 1] - b;
    }
    uint32_t dmax = d;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) dmax = max(dmax, (uint32_t)__shfl_xor(dmax, o, kWave));
    for (uint32_t j = 0; j < dmax; j++) {
        const bool act = j < d;
        push_visit(R, act, act ? g.col[b + j] : 0u);
    }
}

// LoadBalancing::NONE pull for the word `w` (visited word vw): each
// unvisited lane walks its own in-edges in CSC order and stops at its first
// parent in the frontier (async_pull_dd.cuh:26-70).
__device__ __forceinline__ void pull_word_serial(Round &R, uint32_t w, uint64_t vw)
{
    const BfsArgs &a = *R.a;
    const DevGraph &g = a.g;
    const int ln = lane();
    const uint32_t v = w * 64 + ln;
    const uint64_t visited = ballot(v >= g.n) | vw;
    const bool want = !((visited >> ln) & 1);
    uint32_t b = want ? g.cstart[v] : 0u, e = want ? g.cstart[v + 1] : 0u;
    bool found = false;
    while (ballot(b < e && !found)) {
        if (b < e && !found) {
            found = in_frontier(R, g.csrc[b]);
            b++;
        }
    }
    const uint64_t fm = ballot(found);
    if (ln == 0) {
        a.vis[w] = visited | fm;
        R.fnext[w] = fm;
        R.count += __popcll(fm);
    }
    if (found) {
        a.level[caller_id(a, v)] = R.L + 1;
        const uint32_t dv = g.row[v + 1] - g.row[v];
        R.scout += dv;
        R.huge |= dv > kHugeDegree;
    }
}

// Bottom-up step for the 64 vertices of each of ng (<= G) words wg[i] (vwg[i]:
// visited filter word). Every load of a stage is issued for all G words
// before any is used, so the words' dependent chains overlap.
template <int G>
__device__ __forceinline__ void pull_group(Round &R, const uint32_t (&wg)[G], const uint64_t (&vwg)[G], int ng)
{
    const BfsArgs &a = *R.a;
    const DevGraph &g = a.g;
    const int ln = lane();
    uint32_t v[G], b[G], e[G], u[G];
    uint64_t visited[G];
    bool found[G];
#pragma unroll
    for (int i = 0; i < G; i++) v[i] = wg[i] * 64 + ln;
    bool want[G];
    uint32_t dv[G];
    // One step issues, for every unvisited lane, the first in-neighbour (from
    // the head array, coalesced across the word), the in-edge range and the
    // out-degree, so the parent test, the walk and the frontier statistics
    // wait on a single round of loads.
#pragma unroll
    for (int i = 0; i < G; i++) {
        // the filter holds every discovery (pull rounds write whole words,
        // push rounds set bits), the inert vertices and the source
        visited[i] = i < ng ? (ballot(v[i] >= g.n) | vwg[i]) : ~0ull;
        want[i] = !((visited[i] >> ln) & 1);
        u[i] = want[i] ? g.chead[v[i]] : 0xFFFFFFFFu;
        b[i] = want[i] ? g.cstart[v[i]] : 0u;
        e[i] = want[i] ? g.cstart[v[i] + 1] : 0u;
        dv[i] = want[i] ? g.row[v[i] + 1] - g.row[v[i]] : 0u;
    }
#pragma unroll
    for (int i = 0; i < G; i++) {
        found[i] = u[i] != 0xFFFFFFFFu && in_frontier(R, u[i]);
        // lanes still looking walk their in-edges past the head
        if (found[i] || u[i] == 0xFFFFFFFFu) b[i] = e[i];
        else b[i]++;
    }
    // Walk the remaining in-edge lists of every word of the group in lock
    // step: each iteration plans one 64-edge step per word (the concatenated
    // lists of its still-searching lanes), then issues all of the group's
    // loads before any test, so one memory latency covers G steps. A lane
    // stops at its first parent (re-planned every step).
    for (;;) {
        uint32_t idx[G], incl[G], excl[G];
        bool valid[G];
        bool any = false;
#pragma unroll
        for (int i = 0; i < G; i++) {
            const uint32_t r = found[i] ? 0u : e[i] - b[i];
            incl[i] = wave_incl_scan(r);
            excl[i] = incl[i] - r;
            const uint32_t total = shfl(incl[i], kWave - 1);
            any |= total != 0;
            valid[i] = (uint32_t)ln < total;
            const int k = owner_lane(incl[i], valid[i] ? (uint32_t)ln : 0u);
            idx[i] = shfl(b[i], k) + ((uint32_t)ln - shfl(excl[i], k));
        }
        if (!any) break;
        uint32_t src[G];
#pragma unroll
        for (int i = 0; i < G; i++) src[i] = valid[i] ? g.csrc[idx[i]] : 0u;
#pragma unroll
        for (int i = 0; i < G; i++) {
            const uint64_t H = ballot(valid[i] && in_frontier(R, src[i]));
            const uint32_t lo = min(excl[i], 64u), hi = min(incl[i], 64u);
            if (hi > lo) {
                const uint32_t c = hi - lo;
                const uint64_t msk = (c == 64 ? ~0ull : ((1ull << c) - 1)) << lo;
                if (H & msk) found[i] = true;
                b[i] += c;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < G; i++) {
        if (i >= ng) break;
        const uint64_t fm = ballot(found[i]);
        if (ln == 0) {
            a.vis[wg[i]] = visited[i] | fm;
            R.fnext[wg[i]] = fm;
            R.count += __popcll(fm);
        }
        if (found[i]) {
            a.level[caller_id(a, v[i])] = R.L + 1;
            R.scout += dv[i];
            R.huge |= dv[i] > kHugeDegree;
        }
    }
}

#if SG_BFS_WAVES
#define SG_BFS_BOUNDS(BS) __launch_bounds__(BS, SG_BFS_WAVES)
#else
#define SG_BFS_BOUNDS(BS) __launch_bounds__(BS)
#endif

// One BFS query over the `nblocks` blocks of a persistent grid (this block
// is number `brank` of them): the whole launch, or the blocks one XCD holds
// when a launch serves one query per XCD. `gen` is the block's barrier state
// for a.bar (from bar_begin or the caller's census), advanced by every
// barrier. Returns false when a bounded wait expired (the run then stopped).
// kStage: push rounds stage hub neighbour lists in LDS (the engine option
// stage_hubs; a separate instantiation, so the default path keeps its
// register budget).
template <int BS, bool kStage = false>
__device__ __forceinline__ bool bfs_run(const BfsArgs &a, uint32_t brank, uint32_t nblocks, uint32_t xcc,
                                        BarState &gen)
{
    constexpr int WPB = BS / kWave;
    constexpr uint32_t kLdsQ = 8u * BS;
    const DevGraph &g = a.g;
    const uint32_t n = g.n, W = g.words;
    const uint32_t gtid = brank * BS + threadIdx.x;
    const uint32_t nthreads = nblocks * BS;
    const uint32_t gwave = brank * WPB + threadIdx.x / kWave;
    const uint32_t nwaves = nblocks * WPB;
    const int ln = lane();
    const long long t_start = wall_clock64();

    __shared__ unsigned long long s_red[2][WPB];
    __shared__ unsigned long long s_lvl[5];
    __shared__ uint32_t s_q[kLdsQ];
    __shared__ uint32_t s_qn, s_flush[2];
    // hub round: the block's member chunks (first edge, end) and their staged heads
    __shared__ uint2 s_hc[kStage ? kHubList : 1];
    __shared__ uint32_t s_he[kStage ? kStageChunks * kHugeChunk : 1];
    __shared__ uint32_t s_hn;

    // ---- prologue: reset the levels (caller order) and seed the source.
    // The previous query left the frontier bitmaps and counters zeroed, and
    // visited filters alternate between two buffers (the idle one is reset
    // here for the next query), so with a narrow source block 0 expands it
    // (round 0) while the other blocks reset, and one barrier ends both.
    const uint32_t src = a.source;
    const uint32_t sbeg = g.row[src], sdeg = g.row[src + 1] - sbeg;
    const uint64_t sbit = 1ull << (src & 63);
    for (uint32_t i = gtid; i < n; i += nthreads) a.level[i] = (i == a.source_id) ? 0 : -1;
    for (uint32_t w = gtid; w < W; w += nthreads) a.vis_idle[w] = g.inert ? g.inert[w] : 0ull;
    DirectionPolicy pol{a.alpha, a.beta, a.total_workload0, n, a.force};  // replicated in every block
    unsigned long long ptrace = 0;
    uint32_t pushes = 0, pulls = 0;
    uint32_t L = 0;
    const bool local0 = !a.full_init && a.force != 1 && sdeg <= kLocalRound0;
    if (!local0) {
        // state of unknown content (first query, other work since), a wide
        // source or forced pull: reset everything, round 0 over the grid
        for (uint32_t w = gtid; w < W; w += nthreads) {
            const uint64_t sb = (w == (src >> 6)) ? sbit : 0ull;
            a.vis[w] = (g.inert ? g.inert[w] : 0ull) | sb;
            a.fbit[w] = sb;                    // frontier 0
            a.fbit[(size_t)W + w] = 0;         // frontier 1, built by round 0
            a.fbit[2 * (size_t)W + w] = 0;
        }
        if (brank == 0) {
            clear_ctr(&a.ctr[0]);
            clear_ctr(&a.ctr[1]);
            clear_ctr(&a.ctr[2]);
            block_sync_global();
            if (threadIdx.x == 0) {
                a.queue[0] = src;
                a.ctr[0].qlen = 1;
                a.ctr[0].has_huge = sdeg > kHugeDegree;
                a.ctr[0].stripe[0].scout = sdeg;
                a.ctr[0].stripe[0].count = 1;
            }
        }
    } else if (brank == 0) {
        // round 0 without its levels: other blocks are still resetting the
        // level array, so round 1 writes level 1 for the queue it reads
        if (threadIdx.x == 0) {
            atomicOr((unsigned long long *)&a.vis[src >> 6], (unsigned long long)sbit);
            s_qn = 0;
        }
        block_sync_global();
        Round R;
        R.a = &a;
        R.L = 0;
        R.fcur = nullptr;
        R.fnext = a.fbit + (size_t)W;
        R.S = Stager{a.queue + (size_t)a.qcap, a.qcap, &a.ctr[1], s_q, &s_qn, kLdsQ};
        R.scout = 0;
        R.count = 0;
        R.huge = false;
        R.levels = false;
        for (uint32_t e0 = (threadIdx.x / kWave) * kWave; e0 < sdeg; e0 += WPB * kWave) {
            const uint32_t e = e0 + (uint32_t)ln;
            const bool act = e < sdeg;
            push_visit(R, act, act ? g.col[sbeg + e] : 0u);
        }
        flush_stage(R.S, s_flush);
        publish_counts<WPB>(&a.ctr[1], R.scout, R.count, R.huge, s_red);
    }
    bool ok = grid_sync(a.bar, xcc, gen, a.abort);
    if (local0) {
        pol.next(0, 1, sdeg);  // push, as every round 0
        pushes = 1;
        if (gtid == 0) write_trace(a.trace, 0, false, 1, sdeg, t_start);
        L = 1;
    }

    while (ok) {
        read_counts(&a.ctr[L % 3], s_lvl);
        if (threadIdx.x == 0) s_qn = 0;
        __syncthreads();
        const RoundCounts rc = unpack_counts(s_lvl, a.qcap);
        __syncthreads();
        if (L == 1 && local0) {
            // the source's out-neighbours found by block 0 (at most
            // kLocalRound0, so the queue holds them all)
            const uint32_t *q1 = a.queue + (size_t)a.qcap;
            for (uint32_t i = gtid; i < rc.qlen; i += nthreads) a.level[caller_id(a, q1[i])] = 1;
        }
        if (rc.count == 0) break;
        if (a.max_rounds && L >= a.max_rounds) break;

        const bool pull = pol.next(L, rc.count, rc.scout);
        if (pull) {
            pulls++;
            if (L < 64) ptrace |= 1ull << L;
        } else {
            pushes++;
        }
        if (gtid == 0) write_trace(a.trace, L, pull, rc.count, rc.scout, t_start);
        if (brank == 0) clear_ctr(&a.ctr[(L + 2) % 3]);

        Round R;
        R.a = &a;
        R.L = (int32_t)L;
        R.fcur = a.fbit + (size_t)(L % 3) * W;
        R.fnext = a.fbit + (size_t)((L + 1) % 3) * W;
        {
            // frontier L-1 is no longer read: clear its bitmap for round L+1
            uint64_t *fclr = a.fbit + (size_t)((L + 2) % 3) * W;
            for (uint32_t w = gtid; w < W; w += nthreads) fclr[w] = 0;
        }
        R.S = Stager{a.queue + (size_t)((L + 1) % 3) * a.qcap, a.qcap, &a.ctr[(L + 1) % 3], s_q, &s_qn, kLdsQ};
        R.scout = 0;
        R.count = 0;
        R.huge = false;
        const uint32_t *qin = a.queue + (size_t)(L % 3) * a.qcap;

        if (!pull) {
            // (1) high-degree frontier vertices through the static chunk list.
            //     With kStage their neighbour lists are staged in LDS
            //     (the reference's CTA-level expansion of large vertices,
            //     cta_scheduler_hybrid.cuh:86-142): thread t of block b
            //     checks descriptors k * nblocks * BS + t * nblocks + b (the
            //     chunks of one hub land in different blocks) and lists the
            //     chunks whose vertex is in the frontier (first edge, end),
            //     up to kHubList at a time; the block then loads their heads
            //     into LDS, kStageChunks chunks per pass with every thread's
            //     loads in flight together, and its waves expand one staged
            //     chunk per step.
            const uint32_t nhuge = rc.has_huge && !a.lb_push_none ? g.nhuge : 0u;
            const uint32_t per = !kStage ? 0u : (nhuge + nblocks * BS - 1) / (nblocks * BS);  // descriptors per thread
            for (uint32_t k0 = 0; k0 < per;) {
                if (threadIdx.x == 0) s_hn = 0;
                __syncthreads();
                // each pass lists at most kHubList chunks: BS threads x kHubList / BS descriptors
                const uint32_t k1 = min(per, k0 + kHubList / BS);
                for (uint32_t k = k0; k < k1; k++) {
                    const uint32_t c = k * nblocks * BS + threadIdx.x * nblocks + brank;
                    if (c < nhuge) {
                        const uint2 dsc = g.huge[c];
                        const uint32_t end = g.huge_end[c];
                        if (in_frontier(R, dsc.x)) s_hc[atomicAdd(&s_hn, 1u)] = make_uint2(dsc.y, end);
                    }
                }
                k0 = k1;
                __syncthreads();
                const uint32_t nm = s_hn;
                for (uint32_t m0 = 0; m0 < nm; m0 += kStageChunks) {
                    const uint32_t mc = min(kStageChunks, nm - m0);
                    for (uint32_t i = threadIdx.x; i < mc * kHugeChunk; i += BS) {
                        const uint2 ch = s_hc[m0 + i / kHugeChunk];
                        const uint32_t e = ch.x + i % kHugeChunk;
                        s_he[i] = e < ch.y ? g.col[e] : kNoHead;
                    }
                    __syncthreads();
                    for (uint32_t q = threadIdx.x / kWave; q < mc; q += WPB) {
                        const uint32_t v = s_he[q * kHugeChunk + (uint32_t)ln];
                        push_visit(R, v != kNoHead, v != kNoHead ? v : 0u);
                    }
                    __syncthreads();
                }
                __syncthreads();  // every thread has read the count before it is reset
            }
            //     Otherwise (the default: on RMAT-22 staging measured 6 %
            //     slower for the whole query, its block barriers outweighing
            //     the overlapped loads) lane ln of wave gw takes chunk
            //     base + ln * nwaves + gw, so the chunks of one hub land on
            //     different waves, and each wave expands its member chunks
            //     one after another.
            for (uint32_t base = 0; !kStage && base < nhuge; base += nwaves * kWave) {
                const uint32_t c = base + (uint32_t)ln * nwaves + gwave;
                uint2 dsc = make_uint2(0u, 0u);
                uint32_t end = 0;
                bool mem = false;
                if (c < nhuge) {
                    dsc = g.huge[c];
                    end = g.huge_end[c];
                    mem = in_frontier(R, dsc.x);
                }
                uint64_t mm = ballot(mem);
                while (mm) {
                    const int t = first_lane(mm);
                    mm &= mm - 1;
                    const uint32_t beg = shfl(dsc.y, t), stop = shfl(end, t);
                    for (uint32_t e0 = beg; e0 < stop; e0 += kWave) {
                        const uint32_t e = e0 + ln;
                        const bool act = e < stop;
                        push_visit(R, act, act ? g.col[e] : 0u);
                    }
                }
            }
            // (2) the rest of the frontier: from the queue after a push round,
            //     from the bitmap after a pull round or if the queue overflowed.
            if (pol.pull_prev || rc.overflow) {
                // words dealt round-robin: in a degree-ordered graph adjacent
                // words hold vertices of similar degree
                for (uint32_t wb = 0; wb < W; wb += nwaves * kWave) {
                    const uint32_t w = wb + (uint32_t)ln * nwaves + gwave;
                    const uint64_t fl = w < W ? R.fcur[w] : 0ull;
                    wave_word_vertices(fl, w, [&](bool valid, uint32_t u) {
                        if (a.lb_push_none) push_expand_serial(R, valid, u);
                        else push_expand(R, valid, u);
                    });
                }
            } else {
                // entries dealt round-robin, lane ln of wave gw taking entry
                // base + ln * nwaves + gw: a short queue puts one vertex on each
                // of many waves, and no wave gets a run of entries staged
                // together (neighbours of one hub)
                for (uint32_t i0 = 0; i0 < rc.qlen; i0 += nwaves * kWave) {
                    const uint32_t i = i0 + (uint32_t)ln * nwaves + gwave;
                    const bool valid = i < rc.qlen;
                    if (!ballot(valid)) break;
                    if (a.lb_push_none) push_expand_serial(R, valid, valid ? qin[i] : 0u);
                    else push_expand(R, valid, valid ? qin[i] : 0u);
                }
            }
        } else {
            // Pull: sweep this wave's words, skipping ones the filter marks
            // full. Words are dealt round-robin (lane ln of wave gw takes word
            // base + ln * nwaves + gw): unreached high-in-degree vertices walk
            // long in-edge lists, and they cluster at low ids (RMAT,
            // degree-ordered graphs).
            for (uint32_t wb = 0; wb < W; wb += nwaves * kWave) {
                const uint32_t wl = wb + (uint32_t)ln * nwaves + gwave;
                const uint64_t vl = wl < W ? a.vis[wl] : ~0ull;
                if (wl < W && vl == ~0ull) R.fnext[wl] = 0;
                uint64_t todo = ballot(vl != ~0ull);
                if (a.lb_pull_none) {
                    while (todo) {
                        const int t = first_lane(todo);
                        todo &= todo - 1;
                        pull_word_serial(R, wb + (uint32_t)t * nwaves + gwave, shfl(vl, t));
                    }
                }
                while (todo) {
                    constexpr int G = kPullGroup;
                    uint32_t wg[G];
                    uint64_t vwg[G];
                    int ng = 0;
#pragma unroll
                    for (int i = 0; i < G; i++) {
                        wg[i] = wb;
                        vwg[i] = ~0ull;
                        if (todo) {
                            const int t = first_lane(todo);
                            todo &= todo - 1;
                            wg[i] = wb + (uint32_t)t * nwaves + gwave;
                            vwg[i] = shfl(vl, t);
                            ng++;
                        }
                    }
                    pull_group<G>(R, wg, vwg, ng);
                }
            }
        }
        // counters first: their atomics are then in flight while the queue
        // flush waits for its position
        publish_counts<WPB>(R.S.nxt, R.scout, R.count, R.huge, s_red);
        flush_stage(R.S, s_flush);
        pol.pull_prev = pull;
        L++;
        ok = grid_sync(a.bar, xcc, gen, a.abort);
    }

    // ---- epilogue: every block left the loop after the same barrier, which
    // made the last round's levels visible. Zero the frontier bitmaps and
    // counters for the next query.
    if (ok) {
        for (uint32_t w = gtid; w < 3 * W; w += nthreads) a.fbit[w] = 0;
        if (brank == 0) {
            clear_ctr(&a.ctr[0]);
            clear_ctr(&a.ctr[1]);
            clear_ctr(&a.ctr[2]);
        }
    }
    if (gtid == 0) write_final(a.trace, a.stats, L, pushes, pulls, ptrace, ok, t_start);
    return ok;
}

}  // namespace bfsk
}  // namespace
