// Frontier machinery shared by the persistent traversal kernels (BFS, SSSP,
// BC forward): per-round counter slots, LDS-staged queue appends and the
// wave-level owner search used by prefix-sum edge scheduling.
//
// These replace the reference's device worklist (groute::dev::Queue with one
// atomicAdd per appended item, include/groute/device/queue.cuh:84-88) and its
// per-round count/out-degree read-backs (framework.cuh:451-571, 700-796).
#pragma once

#include "sg_internal.hpp"

namespace sg {

constexpr int kStripes = 16;

struct alignas(128) Stripe {
    unsigned long long scout;  // sum of out-degree over appended vertices
    unsigned long long count;  // appended vertices
    uint32_t pad[28];
};

// Counters describing one round's frontier; three slots rotate.
struct alignas(128) FrontierCtr {
    uint32_t qlen;      // entries appended to this round's queue
    uint32_t overflow;  // set when appends ran past the queue capacity
    uint32_t has_huge;  // the frontier holds a vertex above kHugeDegree
    uint32_t pad[29];
    Stripe stripe[kStripes];
};
static_assert(sizeof(FrontierCtr) == 128 * (1 + kStripes), "FrontierCtr layout");
static_assert(sizeof(FrontierCtr) * 3 <= kCtrBytes, "frontier counters exceed the engine's counter area");

__device__ __forceinline__ unsigned long long ld_relaxed64(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Zero one counter slot (called by one whole block).
__device__ inline void clear_ctr(FrontierCtr *c)
{
    uint32_t *w = (uint32_t *)c;
    for (uint32_t i = threadIdx.x; i < sizeof(FrontierCtr) / 4; i += blockDim.x) w[i] = 0;
}

// Block barrier that also orders global memory between the block's waves:
// __syncthreads() alone waits only for LDS traffic, so a global store issued
// by one wave could otherwise land after another wave's later store to the
// same word.
__device__ __forceinline__ void block_sync_global()
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

__device__ __forceinline__ bool test_bit(const uint64_t *bm, uint32_t u) { return (bm[u >> 6] >> (u & 63)) & 1; }

// Index of the lane owning position j of a wave-wide concatenation whose
// inclusive prefix sums are `incl` (the number of lanes with incl <= j).
__device__ __forceinline__ int owner_lane(uint32_t incl, uint32_t j)
{
    int k = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        const uint32_t t = shfl(incl, k + step - 1);
        if (t <= j) k += step;
    }
    return k;
}

// Position of the r-th set bit of x (r < popcount(x)), by halving.
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t x, uint32_t r)
{
    uint32_t pos = 0;
    const uint32_t lo = (uint32_t)__popc((uint32_t)x);
    if (r >= lo) {
        r -= lo;
        x >>= 32;
        pos += 32;
    }
    uint32_t y = (uint32_t)x;
#pragma unroll
    for (int h = 16; h >= 2; h >>= 1) {
        const uint32_t c = (uint32_t)__popc(y & ((1u << h) - 1u));
        if (r >= c) {
            r -= c;
            y >>= h;
            pos += (uint32_t)h;
        }
    }
    if (r >= (y & 1u)) pos += 1;
    return pos;
}

// The vertices of a wave's bitmap words, up to 64 at a time: lane ln holds
// word fl of index wi (vertices wi * 64 + bit); f(valid, v) is called by the
// whole wave once per group of up to 64 of them. A frontier read from a
// bitmap is sparse after a pull round (mostly one or two vertices per word),
// and expanding word by word had left most lanes of every step idle.
template <class F>
__device__ __forceinline__ void wave_word_vertices(uint64_t fl, uint32_t wi, F &&f)
{
    const uint32_t c = (uint32_t)__popcll(fl);
    const uint32_t incl = wave_incl_scan(c);
    const uint32_t total = shfl(incl, kWave - 1);
    const int ln = lane();
    for (uint32_t j0 = 0; j0 < total; j0 += kWave) {
        const uint32_t j = j0 + (uint32_t)ln;
        const bool valid = j < total;
        const int k = owner_lane(incl, valid ? j : total - 1);
        const uint64_t fk = shfl(fl, k);
        const uint32_t first = shfl(incl, k) - shfl(c, k);  // this word's first position
        const uint32_t wk = shfl(wi, k);
        f(valid, valid ? wk * 64 + nth_set_bit(fk, j - first) : 0u);
    }
}

// Snapshot of a counter slot, read by wave 0 of each block after a barrier.
struct RoundCounts {
    unsigned long long count, scout;
    uint32_t qlen;
    bool overflow, has_huge;
};

// Wave 0 reads the slot into LDS; the caller must __syncthreads() before use.
__device__ __forceinline__ void read_counts(const FrontierCtr *c, unsigned long long *s_lvl)
{
    if (threadIdx.x < kWave) {
        // every load in flight at once (one memory round trip): stripes in
        // lanes 0..15, the slot's header words in lanes 16..18
        const int ln = lane();
        unsigned long long cnt = 0, sc = 0;
        uint32_t hdr = 0;
        if (ln < kStripes) {
            cnt = ld_relaxed64(&c->stripe[ln].count);
            sc = ld_relaxed64(&c->stripe[ln].scout);
        } else if (ln < kStripes + 3) {
            const uint32_t *h = ln == kStripes ? &c->qlen : ln == kStripes + 1 ? &c->overflow : &c->has_huge;
            hdr = __hip_atomic_load(h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        cnt = wave_sum_u64(cnt);
        sc = wave_sum_u64(sc);
        const uint32_t qlen = shfl(hdr, kStripes), ovf = shfl(hdr, kStripes + 1), hg = shfl(hdr, kStripes + 2);
        if (ln == 0) {
            s_lvl[0] = cnt;
            s_lvl[1] = sc;
            s_lvl[2] = qlen;
            s_lvl[3] = ovf;
            s_lvl[4] = hg;
        }
    }
}

__device__ __forceinline__ RoundCounts unpack_counts(const unsigned long long *s_lvl, uint32_t qcap)
{
    RoundCounts r;
    r.count = s_lvl[0];
    r.scout = s_lvl[1];
    r.qlen = (uint32_t)min((unsigned long long)qcap, s_lvl[2]);
    r.overflow = s_lvl[3] != 0;
    r.has_huge = s_lvl[4] != 0;
    return r;
}

// Set bit v of the bitmap `bits` for every lane with p, with one atomicOr per
// run of adjacent lanes whose v fall in the same 64-bit word: neighbour lists
// are sorted, so a wave's 64 edges often hit one word, and 64 same-address
// atomics from a wave (and from every wave expanding the same hubs)
// serialise. A segmented OR-scan gathers each run's bits into its last lane;
// lanes of one word that are not adjacent form runs of their own, so every
// bit is still set. bits2, when given, receives the same bits. Wave-convergent.
__device__ __forceinline__ void wave_set_bits(uint64_t *bits, bool p, uint32_t v, uint64_t *bits2 = nullptr)
{
    if (!ballot(p)) return;
    const int ln = lane();
    const uint32_t w = p ? (v >> 6) : 0xFFFFFFFFu;
    uint64_t m = p ? (1ull << (v & 63)) : 0ull;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint64_t y = __shfl_up(m, d, kWave);
        const uint32_t wy = __shfl_up(w, d, kWave);
        if (ln >= d && wy == w) m |= y;
    }
    const uint32_t wn = __shfl_down(w, 1, kWave);
    const bool last = ln == kWave - 1 || wn != w;
    if (p && last) {
        atomicOr((unsigned long long *)&bits[w], (unsigned long long)m);
        if (bits2) atomicOr((unsigned long long *)&bits2[w], (unsigned long long)m);
    }
}

// wave_set_bits on `bits`, returning for each lane with p whether this call
// set its bit (the bit was clear before): each run's atomicOr returns the old
// word, broadcast to the run's lanes; of the lanes of one run with the same v
// (adjacent or not: two frontier vertices' concatenated lists can give
// v = 5, 6, 5) only the first claims it, found from the run's exclusive
// prefix OR. Racing waves therefore claim a vertex once (an unclaimed copy
// would enter the next queue again, and duplicates compound over the rounds
// of a high-diameter traversal). bits2 gets every p lane's bit with a
// non-returning atomicOr. Wave-convergent.
__device__ __forceinline__ bool wave_claim_bits(uint64_t *bits, bool p, uint32_t v, uint64_t *bits2 = nullptr)
{
    if (!ballot(p)) return false;
    const int ln = lane();
    const uint32_t w = p ? (v >> 6) : 0xFFFFFFFFu;
    const uint64_t mine = p ? (1ull << (v & 63)) : 0ull;
    // runs are maximal groups of adjacent lanes with the same word; the scan
    // must not reach into an earlier run of the same word, whose lanes claim
    // through their own atomic (a bit ORed by both could be taken from them)
    const uint32_t wprev = __shfl_up(w, 1, kWave);
    const bool first = ln == 0 || wprev != w;
    const uint32_t run = (uint32_t)__popcll(ballot(first) & (~0ull >> (kWave - 1 - ln)));
    uint64_t m = mine;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint64_t y = __shfl_up(m, d, kWave);
        const uint32_t ry = __shfl_up(run, d, kWave);
        if (ln >= d && ry == run) m |= y;
    }
    const uint32_t wn = __shfl_down(w, 1, kWave);
    const bool last = ln == kWave - 1 || wn != w;
    uint64_t old = 0;
    if (p && last) {
        old = atomicOr((unsigned long long *)&bits[w], (unsigned long long)m);
        if (bits2) atomicOr((unsigned long long *)&bits2[w], (unsigned long long)m);
    }
    // a lane's run ends at the first lane at or above it flagged last
    const uint64_t lm = ballot(last) >> ln;
    old = __shfl(old, ln + __builtin_ctzll(lm), kWave);
    // bits of the earlier lanes of this lane's run
    const uint64_t mp = __shfl_up(m, 1, kWave);
    const uint32_t rp = __shfl_up(run, 1, kWave);
    const uint64_t before = (ln > 0 && rp == run) ? mp : 0ull;
    return p && !(old & mine) && !(before & mine);
}

// out[v] = src[inv[v]] for v = t, t + nt, ... < n: a kernel epilogue putting a
// result kept in device vertex order back into the caller's order. Four
// elements per thread and step, their index loads and then their gathers
// issued together, since each gather is a random read.
template <typename T>
__device__ __forceinline__ void gather_caller_order(T *out, const T *src, const uint32_t *inv, uint32_t n, uint32_t t,
                                                    uint32_t nt)
{
    constexpr int K = 4;
    for (uint64_t v0 = t; v0 < n; v0 += (uint64_t)K * nt) {
        uint32_t idx[K];
        T x[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint64_t v = v0 + (uint64_t)k * nt;
            idx[k] = v < n ? inv[v] : 0u;
        }
#pragma unroll
        for (int k = 0; k < K; k++) x[k] = v0 + (uint64_t)k * nt < n ? src[idx[k]] : T();
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint64_t v = v0 + (uint64_t)k * nt;
            if (v < n) out[v] = x[k];
        }
    }
}

// Per-block staging of the next queue in LDS.
struct Stager {
    uint32_t *qout;        // global queue of the next round
    uint32_t qcap;
    FrontierCtr *nxt;
    uint32_t *s_q;         // LDS buffer
    uint32_t *s_qn;        // LDS fill count
    uint32_t lds_cap;
};

// Stage this wave's `found` vertices: into LDS while the block buffer has
// room, the remainder straight to the global queue (spill), or, without
// spill, mark the next queue overflowed so the round after reads the frontier
// bitmap instead (pull rounds: a frontier too large for the LDS buffers is
// followed by another pull round, and same-address global appends from every
// wave would serialise). Wave-convergent.
__device__ __forceinline__ void stage(const Stager &S, bool found, uint32_t v, bool spill = true)
{
    const uint64_t m = ballot(found);
    if (!m) return;
    const int leader = first_lane(m);
    uint32_t base = 0;
    if (lane() == leader) base = atomicAdd(S.s_qn, (uint32_t)__popcll(m));
    base = shfl(base, leader);
    const uint32_t slot = base + rank_in(m);
    const bool to_lds = found && slot < S.lds_cap;
    if (to_lds) S.s_q[slot] = v;
    const bool to_glb = found && !to_lds;
    const uint64_t gm = ballot(to_glb);
    if (gm && !spill) {
        if (lane() == first_lane(gm)) S.nxt->overflow = 1;
        return;
    }
    if (gm) {
        const int l2 = first_lane(gm);
        uint32_t gb = 0;
        if (lane() == l2) gb = atomicAdd(&S.nxt->qlen, (uint32_t)__popcll(gm));
        gb = shfl(gb, l2);
        const uint32_t pos = gb + rank_in(gm);
        if (to_glb) {
            if (pos < S.qcap) S.qout[pos] = v;
            else S.nxt->overflow = 1;
        }
    }
}

// Publish the block's staged entries with one atomic (whole block).
__device__ __forceinline__ void flush_stage(const Stager &S, uint32_t *s_flush)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t k = min(*S.s_qn, S.lds_cap);
        s_flush[0] = k ? atomicAdd(&S.nxt->qlen, k) : 0u;
        s_flush[1] = k;
    }
    __syncthreads();
    const uint32_t fb = s_flush[0], fk = s_flush[1];
    for (uint32_t i = threadIdx.x; i < fk; i += blockDim.x) {
        const uint32_t pos = fb + i;
        if (pos < S.qcap) S.qout[pos] = S.s_q[i];
    }
    if (threadIdx.x == 0 && fb + fk > S.qcap) S.nxt->overflow = 1;
}

// Block-reduce this lane's (scout, count, huge) contribution and add it to a
// stripe of the next slot (one atomic per block per field). s_red: 2 x WPB.
template <int WPB>
__device__ __forceinline__ void publish_counts(FrontierCtr *nxt, unsigned long long scout, unsigned long long count,
                                               bool huge, unsigned long long (*s_red)[WPB])
{
    const unsigned long long bsc = wave_sum_u64(scout);
    const unsigned long long bcn = wave_sum_u64(count);
    const bool bhg = ballot(huge) != 0;
    const int wid = threadIdx.x / kWave;
    if (lane() == 0) {
        s_red[0][wid] = bsc;
        s_red[1][wid] = bcn | (bhg ? (1ull << 63) : 0ull);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long sc = 0, cn = 0;
        bool hg = false;
        for (int i = 0; i < WPB; i++) {
            sc += s_red[0][i];
            cn += s_red[1][i] & ~(1ull << 63);
            hg |= (s_red[1][i] >> 63) != 0;
        }
        Stripe &st = nxt->stripe[blockIdx.x % kStripes];
        if (sc) atomicAdd(&st.scout, sc);
        if (cn) atomicAdd(&st.count, cn);
        if (hg) nxt->has_huge = 1;  // same value from every writer
    }
}

// The reference's direction rule (PolicyDecisionMaker::GetInitPolicy /
// GetNextPolicy, include/framework/hybrid_policy.h:218-291) with its uint32
// workload counters. Returns true for pull.
struct DirectionPolicy {
    double alpha, beta;
    uint32_t total_wl;
    uint32_t n;
    int32_t force;  // -1 auto, 0 push, 1 pull
    bool pull_prev = false;

    __device__ __forceinline__ bool next(uint32_t round, unsigned long long count, unsigned long long scout)
    {
        bool pull;
        if (force >= 0) {
            pull = force == 1;
        } else if (round == 0) {
            pull = false;
        } else if (pull_prev && (double)count > (double)n / beta) {
            pull = true;
        } else {
            const uint32_t out_wl = pull_prev ? 1u : (uint32_t)scout;
            pull = (double)out_wl > (double)total_wl / alpha;
        }
        if (!pull) total_wl -= (uint32_t)scout;
        return pull;
    }
};

__device__ __forceinline__ void write_trace(DevTrace *trace, uint32_t L, bool pull, unsigned long long count,
                                            unsigned long long scout, long long t_start)
{
    if (trace && L < kTraceRows) {
        DevTrace t;
        t.round = L;
        t.variant = pull ? SG_ASYNC_PULL_DD : SG_ASYNC_PUSH_DD;
        t.in_nodes = (uint32_t)count;
        t.pad = 0;
        t.in_edges = scout;
        t.ticks = (unsigned long long)(wall_clock64() - t_start);
        trace[L] = t;
    }
}

__device__ __forceinline__ void write_final(DevTrace *trace, DevStats *stats, uint32_t L, uint32_t pushes,
                                            uint32_t pulls, unsigned long long ptrace, bool ok, long long t_start)
{
    if (trace && L < kTraceRows) {
        DevTrace t;
        t.round = L;
        t.variant = -1;
        t.in_nodes = 0;
        t.pad = 0;
        t.in_edges = 0;
        t.ticks = (unsigned long long)(wall_clock64() - t_start);
        trace[L] = t;
    }
    DevStats s;
    s.rounds = L;
    s.push_rounds = pushes;
    s.pull_rounds = pulls;
    s.reached = 0;  // exact reach statistics come from sg_engine_get_stats
    s.reached_edges = 0;
    s.trace = ptrace;
    s.max_depth = L ? L - 1 : 0;
    s.status = ok ? SG_OK : SG_ERR_TIMEOUT;
    s.iterations = L;
    s.pad = 0;
    s.residual = 0.0;
    *stats = s;
}

// Map the engine's variant option onto a forced direction (-1 = auto).
inline int forced_direction(const sg_engine_options &o)
{
    if (o.alpha == 0.0 || o.force_variant == SG_ASYNC_PUSH_DD || o.force_variant == SG_SYNC_PUSH_DD) return 0;
    if (o.force_variant == SG_ASYNC_PULL_DD || o.force_variant == SG_SYNC_PULL_DD) return 1;
    if (o.force_variant == SG_VARIANT_AUTO) return -1;
    return -2;  // not a data-driven traversal variant
}

}  // namespace sg
