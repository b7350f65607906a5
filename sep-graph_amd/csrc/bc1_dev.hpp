// Device side of single-source Brandes BC (one source per run), on the
// structure of the BFS kernel: one query runs as a persistent grid (a whole
// launch, or the blocks of one XCD).
//
// Reference path: samples/hybrid_bc/hybrid_bc.cu:303-411 runs one source per
// Engine::Start. Stage 1 (the BC app, hybrid_bc.cu:23-131) is a hybrid BFS
// that also sums path counts sigma and pushes every level onto a
// WorklistStack (include/framework/worklist_stack.cuh:16-111); stage 2
// (StageTwoDDFused, hybrid_bc.cu:134-225) pops the stack level by level and
// adds sigma(u)/sigma(w) (1 + delta(w)) over successors w with atomicAdd; the
// result is halved (hybrid_bc.cu:374-376). Sigma and delta are FP32 there.
//
// Here:
//   1. levels: bfs_run (push/pull by the alpha/beta rule, exact claims),
//      levels kept in device vertex order;
//   2. level lists: a counting sort of the reached vertices by level (block
//      histograms in LDS, one atomic per block and level, positions handed out
//      per wave and level), so each later phase visits a level's vertices
//      without scanning all n; the same pass sums the out- and in-degrees of
//      each of the first kBcVolLevels levels (edge volumes);
//   3. sigma, level by level: every vertex gathers sigma(u) over in-neighbours
//      one level up (pull over the CSC, wave prefix-sum scheduling), testing
//      "u is one level up" in a bitmap of that level (n/8 bytes, L2-resident)
//      rather than reading u's level from the n-word level array; three such
//      bitmaps rotate, each set while its level is processed and cleared
//      through its level list two levels later; vertices
//      with in-degree above kPullHubDegree are summed per 256-edge chunk of
//      the graph's static in-hub list and merged with an FP64 atomicAdd --
//      path counts are integers below 2^53, so the order of the additions
//      cannot change them; a level holding a sixteenth of the vertices or
//      more is pulled through the sliced in-edges (a lane per vertex), and a
//      level whose upper neighbour has at most 1/kBcPushRatio of its in-edges
//      as out-edges is pushed instead: sigma(u) added along u's out-edges into
//      the level with FP64 atomics (the BFS direction rule applied to the
//      sigma pass; the reference's stage 1 pushes or pulls with its BFS); a
//      forced push or pull variant fixes the direction of every level pass;
//   4. dependencies, level by level from the deepest (the same rotating
//      level bitmaps): acc(u) = sum over
//      successors w one level down of psi(w) = 1/sigma(w) + acc(w), which is
//      (1 + delta(w)) / sigma(w) with delta(w) = sigma(w) acc(w); sigma and
//      acc sit side by side (one 16-byte load per successor); vertices with
//      out-degree above kHugeDegree are summed per 64-edge chunk of the
//      static out-hub list with an FP64 atomicAdd; a level whose lower
//      neighbour has at most 1/kBcPushRatio of its out-edges as in-edges is pushed
//      instead, psi(w) added along w's in-edges into the level (the
//      reference's stage 2 always gathers over each vertex's out-edges,
//      hybrid_bc.cu:155-167);
//   5. 0.5 delta(v) = 0.5 sigma(v) acc(v) written out (0 for the source and
//      for unreached vertices, as BetweennessCentralityHost,
//      hybrid_bc_host.cpp:16-92).
// Sigma and acc are FP64; the output is FP32.
#pragma once

#include "bfs_dev.hpp"

namespace {
namespace bc1k {

using namespace bfsk;

// Deepest level a single-source BC run can hold, and the levels a block
// counts in LDS (deeper ones use global atomics).
constexpr uint32_t kBcMaxDepth = 1u << 16;
constexpr uint32_t kBcLdsLevels = 2048;
// Levels whose edge volumes (sum of out- and of in-degrees) are counted; the
// sigma and dependency passes of these levels pick pull or push by them,
// deeper levels always pull.
constexpr uint32_t kBcVolLevels = 64;
// A pass pushes when the other side's edges number at most 1/kBcPushRatio of
// this side's: a push adds one device-scope FP64 atomic per edge inside the
// DAG, which costs about as much as scanning a dozen edges in a pull.
constexpr uint32_t kBcPushRatio = 16;

struct Bc1Args {
    BfsArgs f;          // forward BFS: f.level (device order, f.perm null), f.stats
    double2 *sa;        // [n]: (sigma, acc) per device vertex
    uint32_t *order;    // [n]: reached vertices grouped by level
    uint32_t *lcnt;     // [kBcMaxDepth + 1]: vertices per level
    uint32_t *loff;     // [kBcMaxDepth + 2]: first position of each level in order
    uint32_t *lfill;    // [kBcMaxDepth + 1]: positions handed out per level
    uint32_t *ctl;      // [0]: deepest level reached, [1]: a level beyond kBcMaxDepth
    uint32_t *lvol;     // [4 * kBcVolLevels]: out-edge volume, in-edge volume, out-hubs, in-hubs per level
    uint64_t *lbm;      // 3 x words: bitmaps of three consecutive levels (level j in slot j % 3)
    // outputs (any may be null):
    float *dep;         // 0.5 delta(v) at dep[perm ? perm[v] : v]
    const uint32_t *perm;
    double *bcsum;      // FP64 sum over sources, device order: = (first) or += 0.5 delta(v)
    float *bcout;       // with `last`: FP32 copy of bcsum, device order
    uint32_t first, last;
};

// For each distinct level among the wave's `valid` lanes, one LDS atomicAdd
// of its lane count into hist[level] (levels below kBcLdsLevels; deeper
// levels go to the global counter glob[level]). With `take`, each lane gets
// the value hist held before its wave's add plus its rank among the wave's
// lanes of that level (a position). Wave-convergent.
__device__ __forceinline__ uint32_t wave_level_add(uint32_t *hist, uint32_t *glob, bool valid, uint32_t l, bool take)
{
    uint64_t todo = ballot(valid);
    uint32_t got = 0;
    while (todo) {
        // the lane indices are wave-uniform: read lanes, no cross-lane shuffle
        const uint32_t l0 = __builtin_amdgcn_readfirstlane(__builtin_amdgcn_readlane(l, first_lane(todo)));
        const uint64_t m = ballot(valid && l == l0);
        todo &= ~m;
        uint32_t base = 0;
        if (lane() == first_lane(m)) {
            const uint32_t c = (uint32_t)__popcll(m);
            base = l0 < kBcLdsLevels ? atomicAdd(&hist[l0], c) : atomicAdd(&glob[l0], c);
        }
        if (take) {
            base = __builtin_amdgcn_readlane(base, first_lane(m));
            if (valid && l == l0) got = base + (uint32_t)rank_in(m);
        }
    }
    return got;
}

// Sum over the edges [b, e) of each lane's vertex (edge ids into `adj`) of
// val(neighbour) for neighbours whose bit is set in `inlevel`, the wave walking the
// concatenated lists kLevelSteps x 64 edges per step with every load of a
// stage issued for all sub-steps before any is used (one memory latency
// covers them all); FP64.
constexpr int kLevelSteps = 4;

template <typename Val>
__device__ __forceinline__ double wave_level_sum(const uint32_t *adj, const uint64_t *inlevel, uint32_t b, uint32_t e,
                                                 Val val)
{
    constexpr int U = kLevelSteps;
    const int ln = lane();
    const uint32_t r = e - b;
    const uint32_t incl = wave_incl_scan(r);
    const uint32_t total = shfl(incl, kWave - 1);
    const uint32_t excl = incl - r;
    double acc = 0.0;
    for (uint32_t jb = 0; jb < total; jb += U * kWave) {
        int k[U];
        uint32_t u[U];
        bool act[U];
#pragma unroll
        for (int s = 0; s < U; s++) {
            const uint32_t j = jb + (uint32_t)(s * kWave + ln);
            act[s] = j < total;
            k[s] = owner_lane(incl, act[s] ? j : 0u);
            const uint32_t ed = shfl(b, k[s]) + (j - shfl(excl, k[s]));
            u[s] = act[s] ? adj[ed] : 0u;
        }
        bool on[U];
#pragma unroll
        for (int s = 0; s < U; s++) on[s] = act[s] && test_bit(inlevel, u[s]);
        double x[U];
#pragma unroll
        for (int s = 0; s < U; s++) x[s] = on[s] ? val(u[s]) : 0.0;
#pragma unroll
        for (int s = 0; s < U; s++) {
            // segmented sum over the lanes of one owner (owners are contiguous)
            double y = x[s];
#pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
                const double t = __shfl_up(y, d, kWave);
                const int ko = __shfl_up(k[s], d, kWave);
                if (ln >= d && ko == k[s]) y += t;
            }
            // this lane's vertex covers positions [excl, incl) of the concatenation
            const uint32_t base = jb + (uint32_t)(s * kWave);
            const uint32_t lo = min(max(excl, base), base + 64u) - base;
            const uint32_t hi = min(max(incl, base), base + 64u) - base;
            const double seg = __shfl(y, (int)(hi > 0 ? hi - 1 : 0), kWave);
            if (hi > lo) acc += seg;
        }
    }
    return acc;
}

// Sum over the edges [b, e) of each lane's hub chunk (at most kChunk edges,
// a multiple of 64) of val(neighbour) for neighbours whose bit is set in
// `inlevel`, returned to the chunk's lane. All 64 lanes walk one chunk's edges;
// Q chunks go together, every load of a step issued for all of them before
// any is used, and each chunk costs one wave reduction rather than the
// per-step owner search and segmented scan of wave_level_sum. FP64.
template <uint32_t kChunk, int Q, typename Val>
__device__ __forceinline__ double wave_chunk_sum(const uint32_t *adj, const uint64_t *inlevel, uint32_t b, uint32_t e,
                                                 Val val)
{
    static_assert(kChunk % kWave == 0, "chunks are whole wave steps");
    const int ln = lane();
    uint64_t todo = ballot(e > b);
    double mine = 0.0;
    while (todo) {
        int t[Q];
        uint32_t cb[Q], ce[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            t[q] = todo ? first_lane(todo) : -1;
            todo &= todo - 1;
            cb[q] = t[q] >= 0 ? shfl(b, t[q]) : 0u;
            ce[q] = t[q] >= 0 ? shfl(e, t[q]) : 0u;
        }
        double acc[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) acc[q] = 0.0;
#pragma unroll
        for (uint32_t o = 0; o < kChunk; o += kWave) {
            uint32_t u[Q];
            bool on[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const uint32_t pos = cb[q] + o + (uint32_t)ln;
                on[q] = pos < ce[q];
                u[q] = on[q] ? adj[pos] : 0u;
            }
#pragma unroll
            for (int q = 0; q < Q; q++) on[q] = on[q] && test_bit(inlevel, u[q]);
            double x[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) x[q] = on[q] ? val(u[q]) : 0.0;
#pragma unroll
            for (int q = 0; q < Q; q++) acc[q] += x[q];
        }
#pragma unroll
        for (int q = 0; q < Q; q++) {
            if (t[q] < 0) break;  // wave-uniform
            const double sum = wave_sum_f64(acc[q]);
            if (ln == t[q]) mine = sum;
        }
    }
    return mine;
}

// For every edge [b, e) of each lane's vertex (edge ids into `adj`), in the
// scheduling of wave_level_sum: when `to(neighbour)` holds, atomically adds the
// owning lane's x into dst[neighbour] (FP64). The push form of a level pass,
// for a level whose neighbouring level has far fewer edges.
template <typename To>
__device__ __forceinline__ void wave_level_push(const uint32_t *adj, uint32_t b, uint32_t e, double x, double *dst,
                                                size_t stride, To to)
{
    constexpr int U = kLevelSteps;
    const int ln = lane();
    const uint32_t r = e - b;
    const uint32_t incl = wave_incl_scan(r);
    const uint32_t total = shfl(incl, kWave - 1);
    const uint32_t excl = incl - r;
    for (uint32_t jb = 0; jb < total; jb += U * kWave) {
        uint32_t u[U];
        double xs[U];
        bool act[U];
#pragma unroll
        for (int s = 0; s < U; s++) {
            const uint32_t j = jb + (uint32_t)(s * kWave + ln);
            act[s] = j < total;
            const int k = owner_lane(incl, act[s] ? j : 0u);
            const uint32_t ed = shfl(b, k) + (j - shfl(excl, k));
            xs[s] = __shfl(x, k, kWave);
            u[s] = act[s] ? adj[ed] : 0u;
        }
        bool on[U];
#pragma unroll
        for (int s = 0; s < U; s++) on[s] = act[s] && to(u[s]);
#pragma unroll
        for (int s = 0; s < U; s++)
            if (on[s]) atomicAdd(dst + (size_t)u[s] * stride, xs[s]);
    }
}

// Out-edge and in-edge volume, out-hub and in-hub count of a lane's vertex
// at level l (below kBcVolLevels) added into vol[k * kBcVolLevels + l] for
// k = 0..3 (LDS). Per-lane LDS atomics: lanes of one level serialise on one
// bank, which costs less than wave reductions' chains of cross-lane shuffles.
__device__ __forceinline__ void lane_level_vol(uint32_t *vol, bool valid, uint32_t l, uint32_t od, uint32_t id)
{
    if (!valid) return;
    atomicAdd(&vol[l], od);
    atomicAdd(&vol[kBcVolLevels + l], id);
    if (od > kHugeDegree) atomicAdd(&vol[2 * kBcVolLevels + l], 1u);
    if (id > kPullHubDegree) atomicAdd(&vol[3 * kBcVolLevels + l], 1u);
}

template <int BS>
__device__ __forceinline__ bool bc1_run(const Bc1Args &a, uint32_t brank, uint32_t nblocks, uint32_t xcc,
                                        BarState &gen)
{
    constexpr int WPB = BS / kWave;
    const DevGraph &g = a.f.g;
    const uint32_t n = g.n;
    const uint32_t gtid = brank * BS + threadIdx.x;
    const uint32_t nthreads = nblocks * BS;
    const uint32_t gwave = brank * WPB + threadIdx.x / kWave;
    const uint32_t nwaves = nblocks * WPB;
    const int ln = lane();
    const int32_t *lvl = a.f.level;

    __shared__ uint32_t s_hist[kBcLdsLevels];
    __shared__ uint32_t s_max, s_scan[BS / kWave], s_D;
    __shared__ uint32_t s_vol[4 * kBcVolLevels];

    // level counters start at zero (the BFS's first barrier publishes this)
    for (uint32_t i = gtid; i <= kBcMaxDepth; i += nthreads) {
        a.lcnt[i] = 0;
        a.lfill[i] = 0;
    }
    for (uint32_t i = gtid; i < 4 * kBcVolLevels; i += nthreads) a.lvol[i] = 0;
    if (gtid == 0) {
        a.ctl[0] = 0;
        a.ctl[1] = 0;
    }
    bool ok = bfs_run<BS>(a.f, brank, nblocks, xcc, gen);
    int32_t status = ok ? SG_OK : SG_ERR_TIMEOUT;

    // ---- per-level counts: block histograms, one atomic per block and level
    if (ok) {
        for (uint32_t i = threadIdx.x; i < kBcLdsLevels; i += BS) s_hist[i] = 0;
        for (uint32_t i = threadIdx.x; i < 4 * kBcVolLevels; i += BS) s_vol[i] = 0;
        if (threadIdx.x == 0) s_max = 0;
        __syncthreads();
        uint32_t lmax = 0;
        for (uint32_t v0 = brank * BS; v0 < n; v0 += nthreads) {
            const uint32_t v = v0 + threadIdx.x;
            const int32_t l = v < n ? lvl[v] : -1;
            const bool valid = l >= 0 && (uint32_t)l <= kBcMaxDepth;
            if (valid) lmax = max(lmax, (uint32_t)l);
            if (l >= 0 && !valid) a.ctl[1] = 1u;
            wave_level_add(s_hist, a.lcnt, valid, valid ? (uint32_t)l : 0u, false);
            // edge volumes of the shallow levels (a level's edges number
            // below 2^32: edge offsets are 32-bit)
            const bool vv = valid && (uint32_t)l < kBcVolLevels;
            if (vv)
                lane_level_vol(s_vol, true, (uint32_t)l, g.row[v + 1] - g.row[v], g.cstart[v + 1] - g.cstart[v]);
        }
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) lmax = max(lmax, (uint32_t)__shfl_xor(lmax, d, kWave));
        if (ln == 0) atomicMax(&s_max, lmax);
        __syncthreads();
        const uint32_t top = min(s_max, kBcLdsLevels - 1);
        for (uint32_t l = threadIdx.x; l <= top; l += BS)
            if (s_hist[l]) atomicAdd(&a.lcnt[l], s_hist[l]);
        for (uint32_t i = threadIdx.x; i < 4 * kBcVolLevels; i += BS)
            if (s_vol[i]) atomicAdd(&a.lvol[i], s_vol[i]);
        if (threadIdx.x == 0) atomicMax(&a.ctl[0], s_max);
        ok = grid_sync(a.f.bar, xcc, gen, a.f.abort);
    }
    // ---- level offsets: one block scans the counts
    if (ok) {
        if (threadIdx.x == 0) {
            s_D = __hip_atomic_load(&a.ctl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_max = __hip_atomic_load(&a.ctl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        if (s_max) {
            // a level too deep for the level lists: every block sees the flag
            ok = false;
            status = SG_ERR_CAPACITY;
        }
    }
    if (ok) {
        const uint32_t D = s_D;
        if (brank == 0) {
            // thread t sums a contiguous run of levels, then a block scan
            const uint32_t per = (D + 1 + BS - 1) / BS;
            const uint32_t l0 = threadIdx.x * per, l1 = min(l0 + per, D + 1);
            uint32_t sum = 0;
            for (uint32_t l = l0; l < l1; l++) sum += a.lcnt[l];
            const uint32_t incl = wave_incl_scan(sum);
            if (ln == kWave - 1) s_scan[threadIdx.x / kWave] = incl;
            __syncthreads();
            uint32_t wbase = 0;
            for (uint32_t w = 0; w < threadIdx.x / kWave; w++) wbase += s_scan[w];
            uint32_t pos = wbase + incl - sum;
            for (uint32_t l = l0; l < l1; l++) {
                a.loff[l] = pos;
                pos += a.lcnt[l];
            }
            if (threadIdx.x == BS - 1) a.loff[D + 1] = pos;
        }
        ok = grid_sync(a.f.bar, xcc, gen, a.f.abort);
    }
    // ---- scatter the reached vertices into their levels' ranges; (sigma,
    // acc) start at (1, 0) for the source and (0, 0) elsewhere
    if (ok) {
        for (uint32_t i = threadIdx.x; i < kBcLdsLevels; i += BS) s_hist[i] = 0;
        __syncthreads();
        // the same vertices as the count pass, so the same block histogram
        for (uint32_t v0 = brank * BS; v0 < n; v0 += nthreads) {
            const uint32_t v = v0 + threadIdx.x;
            const int32_t l = v < n ? lvl[v] : -1;
            const bool valid = l >= 0 && (uint32_t)l < kBcLdsLevels;
            wave_level_add(s_hist, nullptr, valid, valid ? (uint32_t)l : 0u, false);
        }
        __syncthreads();
        const uint32_t top = min(s_D, kBcLdsLevels - 1);
        for (uint32_t l = threadIdx.x; l <= top; l += BS)
            if (s_hist[l]) s_hist[l] = a.loff[l] + atomicAdd(&a.lfill[l], s_hist[l]);
        __syncthreads();
        for (uint32_t v0 = brank * BS; v0 < n; v0 += nthreads) {
            const uint32_t v = v0 + threadIdx.x;
            const int32_t l = v < n ? lvl[v] : -1;
            const bool valid = l >= 0;
            // deep levels take positions from the global fill counter, offset below
            uint32_t pos = wave_level_add(s_hist, a.lfill, valid, valid ? (uint32_t)l : 0u, true);
            if (valid) {
                if ((uint32_t)l >= kBcLdsLevels) pos += a.loff[l];
                a.order[pos] = v;
                a.sa[v] = make_double2(l == 0 ? 1.0 : 0.0, 0.0);
            }
        }
        // level bitmaps: slot 0 holds level 0 (the source), slots 1 and 2 start clear
        const uint32_t W = g.words, src = a.f.source;
        for (uint32_t w = gtid; w < 3 * W; w += nthreads) a.lbm[w] = w == (src >> 6) ? 1ull << (src & 63) : 0ull;
        ok = grid_sync(a.f.bar, xcc, gen, a.f.abort);
    }
    const uint32_t D = ok ? s_D : 0u;
    // ---- sigma, level by level
    const uint32_t W = g.words;
    for (uint32_t L = 1; ok && L <= D; L++) {
        const uint32_t b0 = a.loff[L], e0 = a.loff[L + 1];
        const uint64_t *up = a.lbm + (size_t)((L - 1) % 3) * W;  // level L-1
        uint64_t *cur = a.lbm + (size_t)(L % 3) * W;             // level L, set now
        uint64_t *old = a.lbm + (size_t)((L + 1) % 3) * W;       // level L-2, cleared now
        if (L >= 2)
            for (uint32_t i = a.loff[L - 2] + gtid; i < a.loff[L - 1]; i += nthreads) old[a.order[i] >> 6] = 0ull;
        // push when level L-1's out-edges are far fewer than level L's in-edges
        // (or when a push variant is forced; a forced pull never pushes)
        const bool push = a.f.force == 0 || (a.f.force < 0 && L < kBcVolLevels &&
                                             (uint64_t)a.lvol[L - 1] * kBcPushRatio <= a.lvol[kBcVolLevels + L]);
        if (push) {
            auto at_level = [&](uint32_t w) { return lvl[w] == (int32_t)L; };
            for (uint32_t i0 = b0 + gwave * kWave; i0 < e0; i0 += nwaves * kWave) {
                const uint32_t i = i0 + (uint32_t)ln;
                wave_set_bits(cur, i < e0, i < e0 ? a.order[i] : 0u);
            }
            // sigma(u) of level L-1 along its out-edges into level L
            const uint32_t p0 = a.loff[L - 1];
            for (uint32_t i0 = p0; i0 < b0; i0 += nwaves * kWave) {
                const uint32_t i = i0 + (uint32_t)ln * nwaves + gwave;
                uint32_t u = 0, b = 0, e = 0;
                double x = 0.0;
                if (i < b0) {
                    u = a.order[i];
                    b = g.row[u];
                    e = g.row[u + 1];
                    if (e - b > kHugeDegree) e = b;  // out-hubs: the chunks below
                    x = a.sa[u].x;
                }
                wave_level_push(g.col, b, e, x, &a.sa[0].x, 2, at_level);
            }
            // (skipped for a level known to hold no out-hub)
            const uint32_t nch = L - 1 < kBcVolLevels && !a.lvol[2 * kBcVolLevels + L - 1] ? 0u : g.nhuge;
            for (uint32_t c0 = 0; c0 < nch; c0 += nwaves * kWave) {
                const uint32_t c = c0 + (uint32_t)ln * nwaves + gwave;
                uint32_t u = 0, b = 0, e = 0;
                if (c < g.nhuge) {
                    const uint2 dsc = g.huge[c];
                    if (lvl[dsc.x] == (int32_t)(L - 1)) {
                        u = dsc.x;
                        b = dsc.y;
                        e = min(g.row[u + 1], b + kHugeChunk);
                    }
                }
                if (!ballot(e > b)) continue;
                const double x = e > b ? a.sa[u].x : 0.0;
                wave_level_push(g.col, b, e, x, &a.sa[0].x, 2, at_level);
            }
        } else if (g.sell_src && (uint64_t)(e0 - b0) * 16 >= n) {
            // a level holding a sixteenth of the vertices or more: by the
            // 64-vertex slices of the sliced in-edges, one lane per vertex and
            // one coalesced row per in-edge index, instead of prefix-sum
            // scheduling over the level's list (path counts are integers below
            // 2^53, so the order of the additions cannot change them)
            for (uint32_t w = gwave; w < W; w += nwaves) {
                const uint32_t v = w * 64 + (uint32_t)ln;
                const bool at = v < n && lvl[v] == (int32_t)L;
                // one wave owns word w of this level's bitmap, clear until now
                const uint64_t m = ballot(at);
                if (!m) continue;
                if (ln == 0) cur[w] = m;
                const uint32_t deg = at ? g.cstart[v + 1] - g.cstart[v] : 0u;
                const bool light = at && deg <= kPullHubDegree;  // in-hubs: the chunk pass below
                uint32_t rows = light ? deg : 0u;
#pragma unroll
                for (int d = kWave / 2; d > 0; d >>= 1) rows = max(rows, (uint32_t)__shfl_xor(rows, d, kWave));
                if (!rows) continue;
                const uint32_t *S = g.sell_src + g.sell_off[w] + ln;
                constexpr int U = kLevelSteps;
                double acc = 0.0;
                for (uint32_t j0 = 0; j0 < rows; j0 += U) {
                    uint32_t u[U];
#pragma unroll
                    for (int s = 0; s < U; s++) u[s] = light && j0 + s < deg ? S[(size_t)(j0 + s) * kWave] : 0xFFFFFFFFu;
                    double x[U];
#pragma unroll
                    for (int s = 0; s < U; s++) x[s] = u[s] != 0xFFFFFFFFu && test_bit(up, u[s]) ? a.sa[u[s]].x : 0.0;
#pragma unroll
                    for (int s = 0; s < U; s++) acc += x[s];
                }
                if (light) a.sa[v].x = acc;
            }
        } else {
            for (uint32_t i0 = b0 + gwave * kWave; i0 < e0; i0 += nwaves * kWave) {
                const uint32_t i = i0 + (uint32_t)ln;
                wave_set_bits(cur, i < e0, i < e0 ? a.order[i] : 0u);
            }
            // a wave's 64 vertices strided over the list, so the waves' edge
            // totals even out (neighbouring positions hold vertices of close
            // ids, hence of close degrees)
            for (uint32_t i0 = b0; i0 < e0; i0 += nwaves * kWave) {
                const uint32_t i = i0 + (uint32_t)ln * nwaves + gwave;
                uint32_t v = 0, b = 0, e = 0;
                if (i < e0) {
                    v = a.order[i];
                    b = g.cstart[v];
                    e = g.cstart[v + 1];
                    if (e - b > kPullHubDegree) e = b;  // in-hubs: the chunk pass below
                }
                const double s = wave_level_sum(g.csrc, up, b, e, [&](uint32_t u) { return a.sa[u].x; });
                if (i < e0 && e > b) a.sa[v].x = s;
            }
        }
        // in-hub chunks of this level, kPullChunk in-edges each: a lane per
        // chunk, the wave's chunks walked together like a level list (every
        // load of a step in flight at once), one atomic per chunk
        // (skipped for a level known to hold no in-hub)
        const uint32_t nchin = push || (L < kBcVolLevels && !a.lvol[3 * kBcVolLevels + L]) ? 0u : g.nhugein;
        for (uint32_t c0 = 0; c0 < nchin; c0 += nwaves * kWave) {
            const uint32_t c = c0 + (uint32_t)ln * nwaves + gwave;
            uint32_t v = 0, b = 0, e = 0;
            if (c < g.nhugein) {
                const uint2 dsc = g.hugein[c];
                if (lvl[dsc.x] == (int32_t)L) {
                    v = dsc.x;
                    b = dsc.y;
                    e = min(g.cstart[v + 1], b + kPullChunk);
                }
            }
            if (!ballot(e > b)) continue;
            const double x = wave_chunk_sum<kPullChunk, 4>(g.csrc, up, b, e, [&](uint32_t u) { return a.sa[u].x; });
            if (e > b && x != 0.0) atomicAdd(&a.sa[v].x, x);
        }
        ok = grid_sync(a.f.bar, xcc, gen, a.f.abort);
    }
    // ---- dependencies, from the deepest level up (the deepest has no
    // successors: acc stays 0); the source's level needs none
    for (uint32_t L = D; ok && L-- > 1;) {
        const uint32_t b0 = a.loff[L], e0 = a.loff[L + 1];
        const uint64_t *down = a.lbm + (size_t)((L + 1) % 3) * W;  // level L+1
        uint64_t *cur = a.lbm + (size_t)(L % 3) * W;               // level L, set now
        uint64_t *old = a.lbm + (size_t)((L + 2) % 3) * W;         // level L+2, cleared now
        if (L + 2 <= D)
            for (uint32_t i = a.loff[L + 2] + gtid; i < a.loff[L + 3]; i += nthreads) old[a.order[i] >> 6] = 0ull;
        auto psi = [&](uint32_t w) {
            const double2 x = a.sa[w];
            return 1.0 / x.x + x.y;
        };
        // push when level L+1's in-edges are far fewer than level L's out-edges
        const bool push = a.f.force == 0 || (a.f.force < 0 && L + 1 < kBcVolLevels &&
                                             (uint64_t)a.lvol[kBcVolLevels + L + 1] * kBcPushRatio <= a.lvol[L]);
        if (push) {
            auto at_level = [&](uint32_t u) { return lvl[u] == (int32_t)L; };
            for (uint32_t i0 = b0 + gwave * kWave; i0 < e0; i0 += nwaves * kWave) {
                const uint32_t i = i0 + (uint32_t)ln;
                wave_set_bits(cur, i < e0, i < e0 ? a.order[i] : 0u);
            }
            // psi(w) of level L+1 along its in-edges into level L
            const uint32_t p1 = a.loff[L + 2];
            for (uint32_t i0 = e0; i0 < p1; i0 += nwaves * kWave) {
                const uint32_t i = i0 + (uint32_t)ln * nwaves + gwave;
                uint32_t w = 0, b = 0, e = 0;
                double x = 0.0;
                if (i < p1) {
                    w = a.order[i];
                    b = g.cstart[w];
                    e = g.cstart[w + 1];
                    if (e - b > kPullHubDegree) e = b;  // in-hubs: the chunks below
                    x = psi(w);
                }
                wave_level_push(g.csrc, b, e, x, &a.sa[0].y, 2, at_level);
            }
            const uint32_t nch = L + 1 < kBcVolLevels && !a.lvol[3 * kBcVolLevels + L + 1] ? 0u : g.nhugein;
            for (uint32_t c0 = 0; c0 < nch; c0 += nwaves * kWave) {
                const uint32_t c = c0 + (uint32_t)ln * nwaves + gwave;
                uint32_t w = 0, b = 0, e = 0;
                if (c < g.nhugein) {
                    const uint2 dsc = g.hugein[c];
                    if (lvl[dsc.x] == (int32_t)(L + 1)) {
                        w = dsc.x;
                        b = dsc.y;
                        e = min(g.cstart[w + 1], b + kPullChunk);
                    }
                }
                if (!ballot(e > b)) continue;
                const double x = e > b ? psi(w) : 0.0;
                wave_level_push(g.csrc, b, e, x, &a.sa[0].y, 2, at_level);
            }
            ok = grid_sync(a.f.bar, xcc, gen, a.f.abort);
            continue;
        }
        for (uint32_t i0 = b0 + gwave * kWave; i0 < e0; i0 += nwaves * kWave) {
            const uint32_t i = i0 + (uint32_t)ln;
            wave_set_bits(cur, i < e0, i < e0 ? a.order[i] : 0u);
        }
        // strided over the list like the sigma pass
        for (uint32_t i0 = b0; i0 < e0; i0 += nwaves * kWave) {
            const uint32_t i = i0 + (uint32_t)ln * nwaves + gwave;
            uint32_t u = 0, b = 0, e = 0;
            if (i < e0) {
                u = a.order[i];
                b = g.row[u];
                e = g.row[u + 1];
                if (e - b > kHugeDegree) e = b;  // out-hubs: the chunk pass below
            }
            const double s = wave_level_sum(g.col, down, b, e, psi);
            if (i < e0 && e > b) a.sa[u].y = s;
        }
        // out-hub chunks of this level, kHugeChunk out-edges each, walked
        // like the in-hub chunks of the sigma pass
        const uint32_t nch = L < kBcVolLevels && !a.lvol[2 * kBcVolLevels + L] ? 0u : g.nhuge;
        for (uint32_t c0 = 0; c0 < nch; c0 += nwaves * kWave) {
            const uint32_t c = c0 + (uint32_t)ln * nwaves + gwave;
            uint32_t u = 0, b = 0, e = 0;
            if (c < g.nhuge) {
                const uint2 dsc = g.huge[c];
                if (lvl[dsc.x] == (int32_t)L) {
                    u = dsc.x;
                    b = dsc.y;
                    e = min(g.row[u + 1], b + kHugeChunk);
                }
            }
            if (!ballot(e > b)) continue;
            const double x = wave_chunk_sum<kHugeChunk, 8>(g.col, down, b, e, psi);
            if (e > b && x != 0.0) atomicAdd(&a.sa[u].y, x);
        }
        ok = grid_sync(a.f.bar, xcc, gen, a.f.abort);
    }
    // ---- 0.5 delta(v) = 0.5 sigma(v) acc(v); 0 for the source and for
    // unreached vertices
    if (ok) {
        for (uint32_t v = gtid; v < n; v += nthreads) {
            const int32_t l = lvl[v];
            double h = 0.0;
            if (l > 0) {
                const double2 x = a.sa[v];
                h = 0.5 * x.x * x.y;
            }
            if (a.dep) a.dep[a.perm ? a.perm[v] : v] = (float)h;
            if (a.bcsum) {
                const double t = a.first ? h : a.bcsum[v] + h;
                a.bcsum[v] = t;
                if (a.last && a.bcout) a.bcout[v] = (float)t;
            }
        }
    }
    if (!ok && status == SG_OK) status = SG_ERR_TIMEOUT;
    if (gtid == 0 && status != SG_OK) a.f.stats->status = status;
    return ok;
}

inline Bc1Layout bc1_layout(uint64_t n)
{
    Bc1Layout L{};
    uint64_t o = 0;
    auto put = [&](uint64_t &off, uint64_t bytes) {
        off = o;
        o += (bytes + 255) / 256 * 256;
    };
    put(L.sa, n * 16);
    put(L.lvl, n * 4);
    put(L.order, n * 4);
    put(L.lcnt, (uint64_t)(kBcMaxDepth + 2) * 4);
    put(L.loff, (uint64_t)(kBcMaxDepth + 2) * 4);
    put(L.lfill, (uint64_t)(kBcMaxDepth + 2) * 4);
    put(L.ctl, 256);
    put(L.lvol, 4 * kBcVolLevels * 4);
    put(L.lbm, 3 * ((n + 63) / 64) * 8);
    put(L.fstats, sizeof(DevStats));
    put(L.bcsum, n * 8);
    L.total = o;
    return L;
}

inline int bc1_fill(const sg_engine *e, char *base, const Bc1Layout &L, uint32_t dev_source, Bc1Args &b)
{
    const DevGraph &d = e->g->dev;
    const sg_engine_options &o = e->opt;
    const int force = forced_direction(o);
    if (force < -1) SG_FAIL(SG_ERR_UNSUPPORTED, "BC: only data-driven push/pull variants apply");
    if (dev_source >= d.n) SG_FAIL(SG_ERR_RANGE, "BC: source >= n");
    b = Bc1Args{};
    BfsArgs &f = b.f;
    f.g = d;
    f.level = (int32_t *)(base + L.lvl);
    f.perm = nullptr;
    f.full_init = 1;
    f.qcap = e->qcap;
    f.bar.spin_limit = 1u << 22;
    f.abort = e->abort_flag;
    f.stats = (DevStats *)(base + L.fstats);
    f.trace = nullptr;
    f.source = dev_source;
    f.source_id = dev_source;  // levels stay in device order
    f.force = force;
    f.alpha = o.alpha == 0.0 ? 1.0 : o.alpha;
    f.beta = o.beta;
    f.total_workload0 = (uint32_t)((double)d.m * o.edge_factor);
    f.max_rounds = 0;  // dependencies need every level
    b.sa = (double2 *)(base + L.sa);
    b.order = (uint32_t *)(base + L.order);
    b.lcnt = (uint32_t *)(base + L.lcnt);
    b.loff = (uint32_t *)(base + L.loff);
    b.lfill = (uint32_t *)(base + L.lfill);
    b.ctl = (uint32_t *)(base + L.ctl);
    b.lvol = (uint32_t *)(base + L.lvol);
    b.lbm = (uint64_t *)(base + L.lbm);
    return SG_OK;
}

}  // namespace bc1k
}  // namespace
