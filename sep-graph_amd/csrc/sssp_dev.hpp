// Device side of the hybrid (direction-switching) single-source shortest
// paths: one query runs as a persistent grid (a whole launch, or the blocks
// of one XCD).
//
// Reference path: samples/hybrid_sssp/hybrid_sssp.cu:151-242 drives the same
// Engine round loop as BFS with the SSSP app: a push round relaxes the
// out-edges of queued vertices with atomicMin(buffer[dst], buffer[src] + w)
// and queues every dst it lowers (AccumulateBuffer, hybrid_sssp.cu:98-118);
// a pull round schedules every vertex and lowers it from in-neighbours that
// are in the frontier bitmap (CombineValueBuffer, hybrid_sssp.cu:72-95;
// pull_functor.h:138-188). Distances are uint32, UINT32_MAX when unreached.
//
// Here, as for BFS, the round loop, the alpha/beta direction rule and the
// frontier counters live on the device. Differences from the reference:
//   * a relaxation first compares against a plain load of dist[v] and only
//     issues atomicMin when it can lower it (atomics run at ~27 G/s on
//     MI355X, loads several times faster);
//   * round r's frontier is a bitmap; a vertex lowered in round r is queued
//     for round r+1 by the lane that claims its bit there (one returning
//     atomicOr per run of a wave's lanes in one word), so it is queued once
//     however many relaxations lowered it, and membership tests (hub chunks,
//     pull) read 512 KiB of L2-resident bits;
//   * a pull round owns each vertex with one lane, so the minimum over its
//     frontier in-neighbours is a segmented wave reduction, not atomics.
#pragma once

#include "frontier.hpp"

using namespace sg;

// Frontier bitmap words per thread that SSSP pull rounds keep in LDS.
#ifndef SG_FRONT_LDS
#define SG_FRONT_LDS 12
#endif

namespace {
namespace ssspk {

constexpr uint32_t kInf = 0xFFFFFFFFu;

// 64-edge batches a pull step keeps in flight per word, and hub in-edge
// chunks a wave pulls at once.
#ifndef SG_SSSP_PULL_STEPS
#define SG_SSSP_PULL_STEPS 4
#endif
#ifndef SG_SSSP_HUB_CHUNKS
#define SG_SSSP_HUB_CHUNKS 2
#endif
constexpr int kPullSteps = SG_SSSP_PULL_STEPS;
constexpr int kHubChunks = SG_SSSP_HUB_CHUNKS;

// Table of a launch with one query per XCD (passed by value). The XCDs that
// hold blocks in this launch take queries 0, 1, ... in XCD order.
struct SsspGroups {
    int32_t *status;           // worst status over a call's queries (atomicMin), SG_OK before it
    uint32_t source[kMaxXcd];
    uint32_t *out[kMaxXcd];    // distance row of each query
    uint32_t nq;
};

struct SsspArgs {
    DevGraph g;
    uint32_t *dist;
    uint64_t *fbit;   // 3 * W words: round r's frontier bitmap in slot r % 3
    uint32_t *queue;  // 3 * qcap entries
    uint32_t qcap;
    FrontierCtr *ctr;
    GridBarrier bar;
    uint32_t *abort;
    DevStats *stats;
    DevTrace *trace;
    uint32_t source;
    int32_t force;
    double alpha, beta;
    uint32_t total_workload0;
    uint32_t max_rounds;
    uint32_t probe;   // launch is a residency probe of the grid
    uint32_t delta;   // two-level priority bucket width (0: off)
    // degree-ordered graphs: dist is in device order, and the kernel ends by
    // writing out[v] = dist[inv[v]] (the caller's order); null otherwise
    uint32_t *out;
    const uint32_t *inv;
    uint32_t grouped;          // one query per XCD (gr), per-query state arrays
    SsspGroups gr;
    // LoadBalancing::NONE (one lane per vertex, its edges in order, no hub
    // chunks) for push / pull rounds; otherwise the balanced schedulers
    uint32_t lb_push_none, lb_pull_none;
};

struct Round {
    const SsspArgs *a;
    uint32_t L;
    const uint64_t *fcur;
    const uint64_t *sfb;  // LDS copy of fcur's first sfw words (pull rounds), else unused
    uint32_t sfw;
    uint64_t *fnext;
    Stager S;
    unsigned long long scout, count;
    bool huge;
    uint32_t thr;   // entries with dist >= thr are deferred (priority mode)
    uint32_t minv;  // smallest distance this lane queued for the next round
    uint32_t fmin;  // no distance in this round's frontier is below this
};

// The smallest distance queued for a round is kept complemented in the
// counter slot's first pad word, so that zeroed slots read as "none".
__device__ __forceinline__ uint32_t &slot_cmin(FrontierCtr *c) { return c->pad[0]; }

__device__ __forceinline__ uint32_t sat_add(uint32_t d, uint32_t w) { return d > kInf - w ? kInf : d + w; }

// Frontier membership; pull rounds read the low-id words (in a degree-ordered
// graph the sources of most in-edges) from the block's LDS copy instead of
// sending one L2 request per in-edge.
__device__ __forceinline__ bool in_frontier(const Round &R, uint32_t u)
{
    const uint32_t w = u >> 6;
    const uint64_t x = w < R.sfw ? R.sfb[w] : R.fcur[w];
    return (x >> (u & 63)) & 1;
}

__device__ __forceinline__ uint32_t edge_w(const uint32_t *w, uint32_t e) { return w ? w[e] : 1u; }

// True when no in-neighbour in this round's frontier can lower v: every one
// has a distance of at least R.fmin, and v's in-edges weigh at least its
// minimum in-weight, so a candidate is at least fmin + minw(v). A frontier
// vertex lowered within this same round below fmin is in the next frontier,
// from which v is reached then, so skipping v now only defers that.
__device__ __forceinline__ bool settled(const Round &R, uint32_t v, uint32_t dv)
{
    if (dv == kInf) return false;
    const DevGraph &g = R.a->g;
    const uint32_t mw = g.cminw ? g.cminw[v] : (g.cw ? 0u : 1u);
    return (uint64_t)dv <= (uint64_t)R.fmin + mw;
}

// Relax u->v with tentative distance nd for `act` lanes; wave-convergent.
__device__ __forceinline__ void relax(Round &R, bool act, uint32_t v, uint32_t nd)
{
    const SsspArgs &a = *R.a;
    bool lowered = false;
    if (act && nd < a.dist[v]) {
        lowered = nd < atomicMin(&a.dist[v], nd);
        if (lowered) R.minv = min(R.minv, nd);
    }
    // the lane claiming v's bit in the next frontier queues it
    const bool queued = wave_claim_bits(R.fnext, lowered, v);
    if (queued) {
        const uint32_t dv = a.g.row[v + 1] - a.g.row[v];
        R.scout += dv;
        R.count++;
        R.huge |= dv > kHugeDegree;
    }
    stage(R.S, queued, v);
}

// Priority mode: carry a frontier vertex whose distance is not below the
// threshold into the next round's frontier instead of expanding it.
__device__ __forceinline__ void defer(Round &R, bool d, uint32_t u, uint32_t du)
{
    const SsspArgs &a = *R.a;
    if (d) R.minv = min(R.minv, du);
    const bool queued = wave_claim_bits(R.fnext, d, u);
    if (queued) {
        const uint32_t dv = a.g.row[u + 1] - a.g.row[u];
        R.scout += dv;
        R.count++;
        R.huge |= dv > kHugeDegree;
    }
    stage(R.S, queued, u);
}

// Relax the out-edges of up to 64 queued vertices (one per lane) with
// wave-level prefix-sum edge scheduling; hubs are left to the chunk list.
__device__ __forceinline__ void push_expand(Round &R, bool valid, uint32_t u)
{
    const DevGraph &g = R.a->g;
    uint32_t b = 0, d = 0, du = kInf;
    bool later = false;
    if (valid) {
        b = g.row[u];
        d = g.row[u + 1] - b;
        if (d > kHugeDegree) d = 0;
        du = R.a->dist[u];
        if (du == kInf) d = 0;
        if (du != kInf && du >= R.thr) {
            later = true;
            d = 0;
        }
    }
    defer(R, later, u, du);
    const uint32_t incl = wave_incl_scan(d);
    const uint32_t total = shfl(incl, kWave - 1);
    const uint32_t off = b - (incl - d);
    const int ln = lane();
    for (uint32_t jb = 0; jb < total; jb += kWave) {
        const uint32_t j = jb + ln;
        const bool act = j < total;
        const int k = owner_lane(incl, j);
        const uint32_t e = shfl(off, k) + j;
        const uint32_t duk = shfl(du, k);
        uint32_t v = 0, nd = kInf;
        if (act) {
            v = g.col[e];
            nd = sat_add(duk, edge_w(g.w, e));
        }
        relax(R, act, v, nd);
    }
}

// LoadBalancing::NONE push: each lane relaxes its own vertex's out-edges in
// CSR order, hubs included (async_push_dd.cuh:24-78).
__device__ __forceinline__ void push_expand_serial(Round &R, bool valid, uint32_t u)
{
    const DevGraph &g = R.a->g;
    uint32_t b = 0, d = 0, du = kInf;
    bool later = false;
    if (valid) {
        b = g.row[u];
        d = g.row[u + 1] - b;
        du = R.a->dist[u];
        if (du == kInf) d = 0;
        if (du != kInf && du >= R.thr) {
            later = true;
            d = 0;
        }
    }
    defer(R, later, u, du);
    uint32_t dmax = d;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) dmax = max(dmax, (uint32_t)__shfl_xor(dmax, o, kWave));
    for (uint32_t j = 0; j < dmax; j++) {
        const bool act = j < d;
        uint32_t v = 0, nd = kInf;
        if (act) {
            v = g.col[b + j];
            nd = sat_add(du, edge_w(g.w, b + j));
        }
        relax(R, act, v, nd);
    }
}

// LoadBalancing::NONE pull for the word w: each lane's vertex takes the
// minimum of dist[u] + w over its in-neighbours u in the frontier, walking
// its whole in-edge list in CSC order (async_pull_dd.cuh:26-70).
__device__ __forceinline__ void pull_word_serial(Round &R, uint32_t w)
{
    const SsspArgs &a = *R.a;
    const DevGraph &g = a.g;
    const int ln = lane();
    const uint32_t v = w * 64 + ln;
    uint32_t b = 0, e = 0, best = kInf;
    if (v < g.n) {
        b = g.cstart[v];
        e = g.cstart[v + 1];
        best = a.dist[v];
    }
    const uint32_t dv0 = best;
    for (; b < e; b++) {
        const uint32_t u = g.csrc[b];
        if (in_frontier(R, u)) {
            const uint32_t du = a.dist[u];
            if (du != kInf) best = min(best, sat_add(du, edge_w(g.cw, b)));
        }
    }
    const bool improved = v < g.n && best < dv0;
    if (improved) {
        a.dist[v] = best;
        R.minv = min(R.minv, best);
        const uint32_t dv = g.row[v + 1] - g.row[v];
        R.scout += dv;
        R.huge |= dv > kHugeDegree;
    }
    const uint64_t fm = ballot(improved);
    if (ln == 0 && fm) {
        R.fnext[w] = fm;
        R.count += __popcll(fm);
    }
}

// Pull step for the 64 vertices of word w: each lane's vertex takes the
// minimum of dist[u] + w over in-neighbours u in the frontier. The word's
// concatenated in-edge lists are walked U x 64 edges per step, every load of
// a stage issued for all U batches before any is used, so one memory latency
// covers U batches.
template <int U>
__device__ __forceinline__ void pull_word(Round &R, uint32_t w)
{
    const SsspArgs &a = *R.a;
    const DevGraph &g = a.g;
    const int ln = lane();
    const uint32_t v = w * 64 + ln;
    uint32_t b = 0, e = 0, best = kInf;
    if (v < g.n) {
        b = g.cstart[v];
        e = g.cstart[v + 1];
        best = a.dist[v];
        if (e - b > kPullHubDegree) b = e;  // long in-edge lists go through the chunk list
        if (settled(R, v, best)) b = e;
    }
    const uint32_t dv0 = best;
    for (;;) {
        const uint32_t r = e - b;
        const uint32_t incl = wave_incl_scan(r);
        const uint32_t total = shfl(incl, kWave - 1);
        if (total == 0) break;
        const uint32_t excl = incl - r;
        int k[U];
        uint32_t ed[U], u[U];
#pragma unroll
        for (int s = 0; s < U; s++) {
            const uint32_t j = (uint32_t)(s * kWave + ln);
            k[s] = owner_lane(incl, j < total ? j : 0u);
            ed[s] = shfl(b, k[s]) + (j - shfl(excl, k[s]));
            u[s] = kInf;
            if (j < total) u[s] = g.csrc[ed[s]];
        }
        uint32_t du[U], wt[U];
#pragma unroll
        for (int s = 0; s < U; s++) {
            const bool f = u[s] != kInf && in_frontier(R, u[s]);
            du[s] = f ? a.dist[u[s]] : kInf;
            wt[s] = f ? edge_w(g.cw, ed[s]) : 0u;
        }
#pragma unroll
        for (int s = 0; s < U; s++) {
            // segmented minimum over lanes with the same owner (owners are
            // contiguous), then each vertex takes the last lane of its run
            uint32_t m = du[s] != kInf ? sat_add(du[s], wt[s]) : kInf;
#pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
                const uint32_t y = __shfl_up(m, d, kWave);
                const int ko = __shfl_up(k[s], d, kWave);
                if (ln >= d && ko == k[s]) m = min(m, y);
            }
            const uint32_t base = (uint32_t)(s * kWave);
            const uint32_t lo = min(max(excl, base), base + 64u) - base;
            const uint32_t hi = min(max(incl, base), base + 64u) - base;
            const uint32_t seg = shfl(m, (int)(hi > 0 ? hi - 1 : 0));
            if (hi > lo) best = min(best, seg);
        }
        b += min(incl, (uint32_t)(U * kWave)) - min(excl, (uint32_t)(U * kWave));
    }
    const bool improved = v < g.n && best < dv0;
    if (improved) {
        a.dist[v] = best;
        R.minv = min(R.minv, best);
        const uint32_t dv = g.row[v + 1] - g.row[v];
        R.scout += dv;
        R.huge |= dv > kHugeDegree;
    }
    const uint64_t fm = ballot(improved);
    if (ln == 0 && fm) {
        // hubs of this word may set bits concurrently from the chunk list
        atomicOr((unsigned long long *)&R.fnext[w], (unsigned long long)fm);
        R.count += __popcll(fm);
    }
}

// Pull step for the 64 vertices of word w through the sliced in-edges
// (DevGraph::sell_*): row j of the slice gives every lane the j-th in-edge of
// its own vertex in one coalesced load, so a lane keeps its vertex's running
// minimum in a register, with no edge scheduling or cross-lane reduction.
// kPullSteps rows are loaded before any is used. A slice whose vertices are
// all settled, hubs or without in-edges is skipped whole. A vertex's rows
// hold its in-edges by increasing weight, so once the last weight read
// cannot beat the lane's best (every frontier distance is at least fmin), no
// later row can: the lane stops, and the wave when all its lanes have.
__device__ __forceinline__ void pull_slice(Round &R, uint32_t w)
{
    const SsspArgs &a = *R.a;
    const DevGraph &g = a.g;
    const int ln = lane();
    const uint32_t v = w * 64 + ln;
    uint32_t best = kInf;
    bool want = false;
    uint32_t deg = 0;
    if (v < g.n) {
        deg = g.cstart[v + 1] - g.cstart[v];
        best = a.dist[v];
        want = deg != 0 && deg <= kPullHubDegree && !settled(R, v, best);
    }
    const uint32_t dv0 = best;
    if (ballot(want)) {
        const uint32_t beg = g.sell_off[w];
        const uint32_t rows = (g.sell_off[w + 1] - beg) / kWave;
        const uint32_t *S = g.sell_src + beg + ln;
        const uint32_t *Wt = g.sell_w ? g.sell_w + beg + ln : nullptr;
        constexpr int U = kPullSteps;
        for (uint32_t j0 = 0; j0 < rows; j0 += U) {
            uint32_t u[U], du[U], wt[U];
#pragma unroll
            for (int s = 0; s < U; s++) {
                const bool in = want && j0 + s < rows;
                u[s] = in ? S[(size_t)(j0 + s) * kWave] : kInf;
                wt[s] = in ? (Wt ? Wt[(size_t)(j0 + s) * kWave] : 1u) : 0u;
            }
            // an in-edge whose weight alone cannot beat the lane's best (every
            // frontier distance is at least fmin) skips its distance read
#pragma unroll
            for (int s = 0; s < U; s++) {
                const bool f = u[s] != kInf && sat_add(R.fmin, wt[s]) < best && in_frontier(R, u[s]);
                du[s] = f ? a.dist[u[s]] : kInf;
            }
#pragma unroll
            for (int s = 0; s < U; s++)
                if (du[s] != kInf) best = min(best, sat_add(du[s], wt[s]));
            want = want && j0 + U < deg && sat_add(R.fmin, wt[U - 1]) < best;
            if (!ballot(want)) break;
        }
    }
    const bool improved = v < g.n && best < dv0;
    if (improved) {
        a.dist[v] = best;
        R.minv = min(R.minv, best);
        const uint32_t dv = g.row[v + 1] - g.row[v];
        R.scout += dv;
        R.huge |= dv > kHugeDegree;
    }
    const uint64_t fm = ballot(improved);
    if (ln == 0 && fm) {
        // hubs of this word may set bits concurrently from the chunk list
        atomicOr((unsigned long long *)&R.fnext[w], (unsigned long long)fm);
        R.count += __popcll(fm);
    }
}

// The merges of a group of hub chunks issued by pull_hubs, finished: a
// chunk whose atomicMin found the hub above its candidate lowered it, and the
// chunk that sets the hub's next-frontier bit counts it (several may lower
// it). pv[i] is kInf where no atomic was issued; only lane 0 issued them.
template <int NC>
__device__ __forceinline__ void finish_merges(Round &R, const uint32_t (&pv)[NC], const uint32_t (&pm)[NC],
                                              const uint32_t (&pold)[NC])
{
    const DevGraph &g = R.a->g;
    if (lane() != 0) return;
#pragma unroll
    for (int i = 0; i < NC; i++) {
        const uint32_t v = pv[i];
        if (v == kInf || pm[i] >= pold[i]) continue;
        R.minv = min(R.minv, pm[i]);
        const uint64_t bit = 1ull << (v & 63);
        if (!(atomicOr((unsigned long long *)&R.fnext[v >> 6], (unsigned long long)bit) & bit)) {
            const uint32_t dv = g.row[v + 1] - g.row[v];
            R.scout += dv;
            R.count++;
            R.huge |= dv > kHugeDegree;
        }
    }
}

// Pull for the hub in-edge chunks dealt to this wave (first, first +
// stride, ...), kHubChunks at a time; each chunk's minimum candidate is merged
// into its hub with atomicMin, as several waves share a hub. The loads are
// staged so that each round trip serves as much as it can: the next group's
// descriptors are read while this group is worked; a chunk's hub distance,
// hub end, source ids and weights are read together (ids read past the hub's
// end, still inside the edge array, are dropped afterwards); and a chunk's
// atomicMin is not waited for: whether it lowered the hub is taken once the
// next group's loads are in flight.
__device__ __forceinline__ void pull_hubs(Round &R, uint32_t first, uint32_t stride, uint32_t nchunks)
{
    const SsspArgs &a = *R.a;
    const DevGraph &g = a.g;
    const int ln = lane();
    constexpr int U = kPullChunk / kWave;
    constexpr int NC = kHubChunks;
    uint2 nxt[NC];
#pragma unroll
    for (int i = 0; i < NC; i++) {
        const uint32_t c = first + (uint32_t)i * stride;
        nxt[i] = c < nchunks ? g.hugein[c] : make_uint2(kInf, 0u);
    }
    uint32_t pv[NC], pm[NC], pold[NC];
#pragma unroll
    for (int i = 0; i < NC; i++) {
        pv[i] = kInf;
        pm[i] = kInf;
        pold[i] = 0u;
    }
    for (uint32_t c0 = first; c0 < nchunks; c0 += NC * stride) {
        uint2 dsc[NC];
#pragma unroll
        for (int i = 0; i < NC; i++) {
            dsc[i] = nxt[i];
            const uint32_t c = c0 + (uint32_t)(NC + i) * stride;
            nxt[i] = c < nchunks ? g.hugein[c] : make_uint2(kInf, 0u);
        }
        uint32_t hend[NC], dv[NC], u[NC][U], wt[NC][U], du[NC][U];
#pragma unroll
        for (int i = 0; i < NC; i++) {
            const bool ok = dsc[i].x != kInf;
            hend[i] = ok ? g.cstart[dsc[i].x + 1] : 0u;
            dv[i] = ok ? a.dist[dsc[i].x] : 0u;
#pragma unroll
            for (int s = 0; s < U; s++) {
                const uint32_t ed = dsc[i].y + (uint32_t)(s * kWave + ln);
                const bool in = ok && ed < g.m;
                u[i][s] = in ? g.csrc[ed] : kInf;
                wt[i][s] = in ? edge_w(g.cw, ed) : 0u;
            }
        }
        finish_merges<NC>(R, pv, pm, pold);
        // an in-edge whose weight alone cannot beat the hub's distance (every
        // frontier distance is at least fmin) skips its distance read; so do
        // all in-edges of a hub no frontier vertex can lower
#pragma unroll
        for (int i = 0; i < NC; i++) {
            const uint32_t end = min(hend[i], dsc[i].y + kPullChunk);
#pragma unroll
            for (int s = 0; s < U; s++) {
                const uint32_t ed = dsc[i].y + (uint32_t)(s * kWave + ln);
                const bool f = ed < end && u[i][s] != kInf && sat_add(R.fmin, wt[i][s]) < dv[i] &&
                               in_frontier(R, u[i][s]);
                du[i][s] = f ? a.dist[u[i][s]] : kInf;
            }
        }
#pragma unroll
        for (int i = 0; i < NC; i++) {
            uint32_t m = kInf;
#pragma unroll
            for (int s = 0; s < U; s++)
                if (du[i][s] != kInf) m = min(m, sat_add(du[i][s], wt[i][s]));
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1) m = min(m, (uint32_t)__shfl_xor(m, d, kWave));
            // dv is this chunk's earlier read of the hub: a candidate not
            // below it cannot lower the hub
            const bool issue = dsc[i].x != kInf && m < dv[i];
            pv[i] = issue ? dsc[i].x : kInf;
            pm[i] = m;
            if (issue && ln == 0) pold[i] = atomicMin(&a.dist[dsc[i].x], m);
        }
    }
    finish_merges<NC>(R, pv, pm, pold);
}

// pull_hubs over the weight-ordered copy of the in-hub edges (weighted
// graphs, DevGraph::hub_src): every frontier distance is at least fmin, so a
// chunk whose lightest in-edge cannot lower its hub (fmin + minw >= the hub's
// distance) is skipped without reading it -- with the edges ordered, that is
// every chunk after the first few of a hub that already has a distance. The
// hub's distance is read first (one round trip more per group of chunks
// than pull_hubs); the chunk descriptors and smallest weights of the next
// group are read while this group is worked, and merges are finished after
// the next group's first loads are in flight, as in pull_hubs.
__device__ __forceinline__ void pull_hubs_sorted(Round &R, uint32_t first, uint32_t stride, uint32_t nchunks)
{
    // positions first, first + stride, ... < nchunks of DevGraph::hub_ord
    const SsspArgs &a = *R.a;
    const DevGraph &g = a.g;
    const int ln = lane();
    constexpr int U = kPullChunk / kWave;
    constexpr int NC = kHubChunks;
    uint2 nv[NC];
    uint32_t nmw[NC];
#pragma unroll
    for (int i = 0; i < NC; i++) {
        const uint32_t c = first + (uint32_t)i * stride;
        nv[i] = c < nchunks ? g.hub_ord[c] : make_uint2(kInf, 0u);
        nmw[i] = c < nchunks ? g.hub_minw[c] : kInf;
    }
    uint32_t pv[NC], pm[NC], pold[NC];
#pragma unroll
    for (int i = 0; i < NC; i++) {
        pv[i] = kInf;
        pm[i] = kInf;
        pold[i] = 0u;
    }
    for (uint32_t c0 = first; c0 < nchunks; c0 += NC * stride) {
        uint32_t hv[NC], hc[NC], mw[NC], dv[NC];
#pragma unroll
        for (int i = 0; i < NC; i++) {
            hv[i] = nv[i].x;
            hc[i] = nv[i].y;
            mw[i] = nmw[i];
            const uint32_t c = c0 + (uint32_t)(NC + i) * stride;
            nv[i] = c < nchunks ? g.hub_ord[c] : make_uint2(kInf, 0u);
            nmw[i] = c < nchunks ? g.hub_minw[c] : kInf;
        }
#pragma unroll
        for (int i = 0; i < NC; i++) dv[i] = hv[i] != kInf ? a.dist[hv[i]] : 0u;
        finish_merges<NC>(R, pv, pm, pold);
        uint32_t u[NC][U], wt[NC][U], du[NC][U];
#pragma unroll
        for (int i = 0; i < NC; i++) {
            const bool live = hv[i] != kInf && sat_add(R.fmin, mw[i]) < dv[i];
            const uint64_t at = (uint64_t)hc[i] * kPullChunk + (uint32_t)ln;
#pragma unroll
            for (int s = 0; s < U; s++) {
                u[i][s] = live ? g.hub_src[at + (uint64_t)s * kWave] : kInf;
                wt[i][s] = live ? g.hub_w[at + (uint64_t)s * kWave] : 0u;
            }
        }
        // an in-edge whose weight alone cannot beat the hub's distance skips
        // its distance read (past the hub's end the source is 0xFFFFFFFF)
#pragma unroll
        for (int i = 0; i < NC; i++)
#pragma unroll
            for (int s = 0; s < U; s++) {
                const bool f = u[i][s] != kInf && sat_add(R.fmin, wt[i][s]) < dv[i] && in_frontier(R, u[i][s]);
                du[i][s] = f ? a.dist[u[i][s]] : kInf;
            }
#pragma unroll
        for (int i = 0; i < NC; i++) {
            uint32_t m = kInf;
#pragma unroll
            for (int s = 0; s < U; s++)
                if (du[i][s] != kInf) m = min(m, sat_add(du[i][s], wt[i][s]));
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1) m = min(m, (uint32_t)__shfl_xor(m, d, kWave));
            const bool issue = hv[i] != kInf && m < dv[i];
            pv[i] = issue ? hv[i] : kInf;
            pm[i] = m;
            if (issue && ln == 0) pold[i] = atomicMin(&a.dist[hv[i]], m);
        }
    }
    finish_merges<NC>(R, pv, pm, pold);
}

// `gen`: the block's barrier state (from bar_begin, or the census of a
// grouped launch); `brank` of `nblocks` ranks the block in the query's grid.
// Returns false when a bounded wait expired (the run then stopped). `gen`
// is advanced by every barrier. kFrontWords: frontier bitmap words a pull
// round copies into LDS (the first kFrontWords * 64 vertex ids).
template <int BS, uint32_t kFrontWords = SG_FRONT_LDS * BS>
__device__ __forceinline__ bool sssp_run(const SsspArgs &a, uint32_t brank, uint32_t nblocks, uint32_t xcc,
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
    __shared__ uint32_t s_qn, s_flush[2], s_cmin, s_bmin;
    constexpr uint32_t kFrontLds = kFrontWords;  // frontier words cached in LDS by pull rounds
    __shared__ uint64_t s_fb[kFrontLds];

    const uint32_t src = a.source;
    for (uint32_t i = gtid; i < n; i += nthreads) a.dist[i] = (i == src) ? 0u : kInf;
    // frontier 0 = the source; frontier 1's bitmap starts clear (round L
    // clears L+2's)
    for (uint32_t w = gtid; w < 2 * W; w += nthreads) a.fbit[w] = w == (src >> 6) ? 1ull << (src & 63) : 0ull;
    if (brank == 0) {
        clear_ctr(&a.ctr[0]);
        clear_ctr(&a.ctr[1]);
        clear_ctr(&a.ctr[2]);
        block_sync_global();
        if (threadIdx.x == 0) {
            const uint32_t deg = g.row[src + 1] - g.row[src];
            a.queue[0] = src;
            a.ctr[0].qlen = 1;
            a.ctr[0].has_huge = deg > kHugeDegree;
            a.ctr[0].stripe[0].scout = deg;
            a.ctr[0].stripe[0].count = 1;
            slot_cmin(&a.ctr[0]) = ~0u;  // smallest queued distance: 0
        }
    }
    bool ok = grid_sync(a.bar, xcc, gen, a.abort);

    DirectionPolicy pol{a.alpha, a.beta, a.total_workload0, n, a.force};
    unsigned long long ptrace = 0;
    uint32_t pushes = 0, pulls = 0;
    uint32_t L = 0;
    while (ok) {
        read_counts(&a.ctr[L % 3], s_lvl);
        if (threadIdx.x == 0) {
            s_qn = 0;
            s_cmin = __hip_atomic_load(&slot_cmin(&a.ctr[L % 3]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const RoundCounts rc = unpack_counts(s_lvl, a.qcap);
        // priority mode: expand only the lowest delta-bucket present
        uint32_t thr = kInf;
        if (a.delta) {
            const uint64_t t = ((uint64_t)(~s_cmin) / a.delta + 1) * a.delta;
            thr = (uint32_t)min(t, (uint64_t)kInf);
        }
        __syncthreads();
        if (rc.count == 0) break;
        if (a.max_rounds && L >= a.max_rounds) break;

        const bool pull = pol.next(L, rc.count, rc.scout);
        if (pull) {
            pulls++;
            if (L < 64) ptrace |= 1ull << L;
        } else {
            pushes++;
        }
        if (gtid == 0 && a.trace) write_trace(a.trace, L, pull, rc.count, rc.scout, t_start);
        if (brank == 0) clear_ctr(&a.ctr[(L + 2) % 3]);

        Round R;
        R.a = &a;
        R.L = L;
        // every round leaves its next frontier in fnext (pull rounds by whole
        // words, push rounds one atomicOr per run of queued vertices)
        R.fcur = a.fbit + (size_t)(L % 3) * W;
        R.sfb = s_fb;
        R.sfw = 0;
        R.fnext = a.fbit + (size_t)((L + 1) % 3) * W;
        // the bitmap for round L+2 is free (read in round L-1): clear it
        uint64_t *fclr = a.fbit + (size_t)((L + 2) % 3) * W;
        for (uint32_t w = gtid; w < W; w += nthreads) fclr[w] = 0;
        R.S = Stager{a.queue + (size_t)((L + 1) % 3) * a.qcap, a.qcap, &a.ctr[(L + 1) % 3], s_q, &s_qn, kLdsQ};
        R.scout = 0;
        R.count = 0;
        R.huge = false;
        R.thr = thr;
        R.minv = kInf;
        // the smallest distance queued for this round (every path that adds
        // a vertex to a frontier records it), 0 when none was recorded
        R.fmin = s_cmin ? ~s_cmin : 0u;
        const uint32_t *qin = a.queue + (size_t)(L % 3) * a.qcap;

        if (!pull) {
            const uint32_t nhuge = rc.has_huge && !a.lb_push_none ? g.nhuge : 0u;
            for (uint32_t base = 0; base < nhuge; base += nwaves * kWave) {
                const uint32_t c = base + (uint32_t)ln * nwaves + gwave;
                uint2 dsc = make_uint2(0u, 0u);
                bool mem = false;
                if (c < nhuge) {
                    dsc = g.huge[c];
                    mem = in_frontier(R, dsc.x);
                }
                uint64_t mm = ballot(mem);
                while (mm) {
                    const int t = first_lane(mm);
                    mm &= mm - 1;
                    const uint32_t u = shfl(dsc.x, t), beg = shfl(dsc.y, t);
                    const uint32_t end = min(g.row[u + 1], beg + kHugeChunk);
                    const uint32_t du = a.dist[u];
                    // a deferred hub is carried by its queue entry
                    if (du >= thr) continue;
                    for (uint32_t e0 = beg; e0 < end; e0 += kWave) {
                        const uint32_t e = e0 + ln;
                        const bool act = e < end && du != kInf;
                        uint32_t v = 0, nd = kInf;
                        if (act) {
                            v = g.col[e];
                            nd = sat_add(du, edge_w(g.w, e));
                        }
                        relax(R, act, v, nd);
                    }
                }
            }
            if (pol.pull_prev) {
                // after a pull round, from its bitmap, words dealt round-robin
                // (in a degree-ordered graph adjacent words hold vertices of
                // similar degree)
                for (uint32_t wb = 0; wb < W; wb += nwaves * kWave) {
                    const uint32_t w = wb + (uint32_t)ln * nwaves + gwave;
                    const uint64_t fl = w < W ? R.fcur[w] : 0ull;
                    wave_word_vertices(fl, w, [&](bool valid, uint32_t u) {
                        if (a.lb_push_none) push_expand_serial(R, valid, u);
                        else push_expand(R, valid, u);
                    });
                }
            } else if (!rc.overflow) {
                // the queue, entries dealt round-robin: lane ln of wave gw
                // takes entry base + ln * nwaves + gw, so no wave gets a run of
                // entries staged together
                for (uint32_t i0 = 0; i0 < rc.qlen; i0 += nwaves * kWave) {
                    const uint32_t i = i0 + (uint32_t)ln * nwaves + gwave;
                    const bool valid = i < rc.qlen;
                    if (!ballot(valid)) break;
                    if (a.lb_push_none) push_expand_serial(R, valid, valid ? qin[i] : 0u);
                    else push_expand(R, valid, valid ? qin[i] : 0u);
                }
            } else {
                for (uint32_t w = gwave; w < W; w += nwaves) {
                    const uint32_t v = w * 64 + ln;
                    if (a.lb_push_none) push_expand_serial(R, v < n && test_bit(R.fcur, v), v);
                    else push_expand(R, v < n && test_bit(R.fcur, v), v);
                }
            }
        } else {
            R.sfw = min(W, kFrontLds);
            for (uint32_t i = threadIdx.x; i < R.sfw; i += BS) s_fb[i] = R.fcur[i];
            __syncthreads();
            // Pull: (1) hubs' in-edges through the static chunk list, spread
            // over all waves, kHubChunks chunks per wave at a time; (2) every
            // other vertex with in-edges, a word at a time.
            const uint32_t nhugein = a.lb_pull_none ? 0u : g.nhugein;
            if (nhugein && g.hub_src) {
                // every hub's first (lightest) chunk before the others: a hub
                // lowered by it lets its heavier chunks be skipped even in the
                // first pull round, when every hub starts unreached (any
                // distance read is an upper bound, so the order is only a
                // matter of how much is skipped)
                if (gwave < g.nhubin) pull_hubs_sorted(R, gwave, nwaves, g.nhubin);
                if (g.nhubin + gwave < nhugein) pull_hubs_sorted(R, g.nhubin + gwave, nwaves, nhugein);
            } else if (gwave < nhugein) {
                pull_hubs(R, gwave, nwaves, nhugein);
            }
            // words interleaved over waves: heavy vertices cluster at low ids
            // (RMAT, degree-ordered graphs), so contiguous ranges would leave a
            // few waves with most of the in-edges
            for (uint32_t w = gwave; w < W; w += nwaves) {
                if (g.inert && g.inert[w] == ~0ull) continue;
                if (a.lb_pull_none) pull_word_serial(R, w);
                else if (g.sell_src) pull_slice(R, w);
                else pull_word<kPullSteps>(R, w);
            }
        }
        // counters first: their atomics are then in flight while the queue
        // flush waits for its position
        publish_counts<WPB>(R.S.nxt, R.scout, R.count, R.huge, s_red);
        {
            // block minimum of the queued distances, one atomic per block
            // (the priority threshold and the settled test of pull rounds)
            uint32_t m = R.minv;
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1) m = min(m, (uint32_t)__shfl_xor(m, d, kWave));
            if (threadIdx.x == 0) s_bmin = kInf;
            __syncthreads();
            if (ln == 0 && m != kInf) atomicMin(&s_bmin, m);
            __syncthreads();
            if (threadIdx.x == 0 && s_bmin != kInf) atomicMax(&slot_cmin(R.S.nxt), ~s_bmin);
        }
        flush_stage(R.S, s_flush);
        pol.pull_prev = pull;
        L++;
        ok = grid_sync(a.bar, xcc, gen, a.abort);
    }

    // distances in the caller's vertex order: every block left the loop after
    // the same barrier, which made the last round's distances visible
    if (ok && a.out) gather_caller_order(a.out, a.dist, a.inv, n, gtid, nthreads);
    if (gtid == 0) write_final(a.trace, a.stats, L, pushes, pulls, ptrace, ok, t_start);
    return ok;
}

}  // namespace ssspk
}  // namespace
