// PageRank as persistent kernels: a topology-driven pull (power iteration)
// and a data-driven residual propagation that switches between push and pull
// per round.
//
// Reference path: samples/hybrid_pr/hybrid_pr.cu:26-125 (the PageRank app:
// residual push, AccumulateBuffer with float atomicAdd) driven by
// RunSyncPullTD (include/framework/algo_variants.cuh:753-823, topology-driven
// pull over the CSC) or RunAsyncPushDD (algo_variants.cuh:133-203, worklist of
// vertices whose residual crossed the threshold, hybrid_pr.cu:66-110).
//
// The ranks computed here are the normalised PageRank with uniform teleport
// and uniform redistribution of dangling mass,
//     x = (1-d)/n + d (P^T x + (sum of x over dangling vertices)/n),
// stored as FP32. Sums are compensated FP32 (hi + lo, Neumaier): in the pull
// mode the per-vertex in-edge sums and the dangling mass, in the residual
// mode the unnormalised ranks, the residuals (each atomic addition's
// rounding error, recomputed by the adding lane, goes to a second FP32
// array) and the normalising sum; only the L1 residual (pull) and the
// residual mass absorbed at the end (DD) are reduced in FP64. In the pull
// mode every reduction runs in a fixed order for a given grid, so repeated
// runs give identical bits.
//
// Pull TD: one lane per vertex sums x[u]/outdeg(u) (kept as an FP32 array
// `contrib`, ping-ponged between iterations) over its in-edges, reading them
// row by row from the graph's sliced in-edge layout (a lane's own entries,
// in order; without that layout, the wave walks the concatenated in-edge
// lists 64 edges per step with a segmented shuffle scan). Vertices with
// in-degree above kPullHubDegree are summed per 256-edge chunk into
// compensated slots and finished after a grid barrier, one wave per hub, in
// chunk order.
//
// Residual DD: solves the unnormalised system y = (1-d)/n + d P^T y by
// residual propagation (dangling mass leaks) and normalises y at the end; for
// uniform teleport this is the same ranking vector, x = y / |y|_1. A round
// activates every vertex whose residual reached theta = tol (1-d) / n (the
// rounds' worklist holds vertices whose residual crossed theta), moves the
// residual into y and spreads d r / outdeg along out-edges: with FP32 atomics
// when the active out-degree is small (push), or by a gather over in-edges
// from active sources when it exceeds m / alpha (pull).
#include "frontier.hpp"

using namespace sg;

namespace {

// ------------------------------------------------------------------ helpers

// ------------------------------------------- compensated FP32 accumulation
// A sum kept as hi + lo, where lo collects the rounding error of every
// addition into hi (Neumaier's variant of Kahan summation): FP32 storage and
// arithmetic with an error close to one rounding of the final sum.
struct CSum {
    float hi, lo;
};

__device__ __forceinline__ void csum_add(CSum &a, float x)
{
    const float s = a.hi + x;
    const float t = fabsf(a.hi) >= fabsf(x) ? (a.hi - s) + x : (x - s) + a.hi;
    a.hi = s;
    a.lo += t;
}

// a then b (the order fixes the rounding)
__device__ __forceinline__ CSum csum_merge(CSum a, const CSum &b)
{
    csum_add(a, b.hi);
    a.lo += b.lo;
    return a;
}

__device__ __forceinline__ float csum_value(const CSum &a) { return a.hi + a.lo; }

// Wave sum; partners merge lower lane first, so every lane ends with the
// same bits.
__device__ __forceinline__ CSum wave_csum(CSum a)
{
    const int ln = lane();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const CSum o{__shfl_xor(a.hi, d, kWave), __shfl_xor(a.lo, d, kWave)};
        a = (ln & d) ? csum_merge(o, a) : csum_merge(a, o);
    }
    return a;
}

// Segmented inclusive compensated sum (owners non-decreasing along the wave).
__device__ __forceinline__ CSum seg_csum(CSum s, int k)
{
    const int ln = lane();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const CSum y{__shfl_up(s.hi, d, kWave), __shfl_up(s.lo, d, kWave)};
        const int ko = __shfl_up(k, d, kWave);
        if (ln >= d && ko == k) s = csum_merge(y, s);
    }
    return s;
}

// Per-block partials of the pull iteration: the L1 residual in FP64 and the
// dangling mass as a compensated FP32 sum (16 bytes, the size of a double2).
struct PrPart {
    double res;
    float dhi, dlo;
};
static_assert(sizeof(PrPart) == sizeof(double2), "PrPart shares the partial-sum area with double2 partials");
static_assert(sizeof(CSum) == 8, "hub chunk sums are 8 bytes");

// Block sums of (residual, dangling) in a fixed order; valid in thread 0.
template <int WPB>
__device__ __forceinline__ PrPart block_part(double res, CSum dang, PrPart *s_red)
{
    res = wave_sum_f64(res);
    dang = wave_csum(dang);
    const int wid = threadIdx.x / kWave;
    if (lane() == 0) s_red[wid] = PrPart{res, dang.hi, dang.lo};
    __syncthreads();
    PrPart t{0.0, 0.f, 0.f};
    if (threadIdx.x == 0) {
        CSum c{0.f, 0.f};
        for (int i = 0; i < WPB; i++) {
            t.res += s_red[i].res;
            c = csum_merge(c, CSum{s_red[i].dhi, s_red[i].dlo});
        }
        t.dhi = c.hi;
        t.dlo = c.lo;
    }
    __syncthreads();
    return t;
}

// Sum of the per-block partials in a fixed order, returned to every thread.
template <int BS>
__device__ __forceinline__ PrPart grid_part(const PrPart *part, uint32_t nblocks, PrPart *s_red, PrPart *s_out)
{
    double res = 0.0;
    CSum dang{0.f, 0.f};
    for (uint32_t i = threadIdx.x; i < nblocks; i += BS) {
        const PrPart q = part[i];
        res += q.res;
        dang = csum_merge(dang, CSum{q.dhi, q.dlo});
    }
    const PrPart t = block_part<BS / kWave>(res, dang, s_red);
    if (threadIdx.x == 0) *s_out = t;
    __syncthreads();
    const PrPart r = *s_out;
    __syncthreads();
    return r;
}

// Sum over the in-edges [b, e) of each lane's vertex of val(edge) as a
// compensated FP32 sum, the wave walking the concatenated lists 4 x 64
// edges per iteration (all of an iteration's loads issued before its
// reductions); each 64-edge sub-step reduces its segments in a fixed order,
// so the result is deterministic.
template <typename F>
__device__ __forceinline__ CSum wave_gather_csum(uint32_t b, uint32_t e, F val)
{
    constexpr int Q = 4;
    const int ln = lane();
    CSum acc{0.f, 0.f};
    for (;;) {
        const uint32_t r = e - b;
        const uint32_t incl = wave_incl_scan(r);
        const uint32_t total = shfl(incl, kWave - 1);
        if (total == 0) break;
        const uint32_t excl = incl - r;
        int k[Q];
        float cand[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const uint32_t j = (uint32_t)(q * kWave + ln);
            k[q] = owner_lane(incl, j < total ? j : total - 1);
            const uint32_t bk = shfl(b, k[q]), xk = shfl(excl, k[q]);
            cand[q] = j < total ? val(bk + (j - xk)) : 0.f;
        }
        const uint32_t span = min(total, (uint32_t)(Q * kWave));
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const uint32_t s0 = (uint32_t)(q * kWave);
            if (s0 >= span) break;  // wave-uniform
            const CSum s = seg_csum(CSum{cand[q], 0.f}, k[q]);
            const uint32_t lo = min(max(excl, s0), s0 + kWave), hi = min(max(incl, s0), min(span, s0 + kWave));
            const int at = (int)(hi > s0 ? hi - 1 - s0 : 0);
            const CSum seg{__shfl(s.hi, at, kWave), __shfl(s.lo, at, kWave)};
            if (hi > lo) acc = csum_merge(acc, seg);
        }
        b += min(incl, span) - min(excl, span);
    }
    return acc;
}

// ------------------------------------------------------------ pull TD kernel

struct PrArgs {
    DevGraph g;
    float *x;          // ranks [n]
    float *contrib;    // 2 x n: x[u] / outdeg(u) of the current / next iteration
    CSum *hubpart;     // [nhugein] chunk sums
    PrPart *part;      // 2 x kMaxBlocks (residual, dangling) partials
    GridBarrier bar;
    uint32_t *abort;
    DevStats *stats;
    DevTrace *trace;
    double damping, tol;
    uint32_t max_iter;
    uint32_t probe;
};

template <int BS>
__global__ __launch_bounds__(BS) void k_pr_pull(PrArgs a)
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
    const float df = (float)a.damping;
    const float inv_nf = (float)(1.0 / (double)n);

    __shared__ PrPart s_red[WPB];
    __shared__ PrPart s_glob;
    // The lowest ids' contributions (in a degree-ordered graph the sources of
    // most in-edges) are copied to LDS each iteration: they are read-only
    // until the next barrier, and the gather reads them there instead of
    // sending one L2 request per in-edge.
#ifndef SG_PR_LDS
#define SG_PR_LDS 32
#endif
    constexpr uint32_t kCcLds = SG_PR_LDS * BS;
    __shared__ float s_cc[kCcLds];

    // x = 1/n; contributions of iteration 0; dangling mass
    CSum dang{0.f, 0.f};
    for (uint32_t v = gtid; v < n; v += nthreads) {
        a.x[v] = inv_nf;
        const uint32_t dout = g.row[v + 1] - g.row[v];
        a.contrib[v] = dout ? inv_nf / (float)dout : 0.0f;
        if (!dout) csum_add(dang, inv_nf);
    }
    {
        const PrPart t = block_part<WPB>(0.0, dang, s_red);
        if (threadIdx.x == 0) a.part[blockIdx.x] = t;
    }
    bool ok = grid_sync(a.bar, xcc, gen, a.abort);

    uint32_t it = 0;
    double res = INFINITY;
    while (ok) {
        const PrPart gs = grid_part<BS>(a.part + (size_t)(it & 1) * kMaxBlocks, gridDim.x, s_red, &s_glob);
        if (it > 0) res = gs.res;
        if ((it > 0 && res < a.tol) || it >= a.max_iter) break;
        const float base = (1.0f - df) * inv_nf + df * csum_value(CSum{gs.dhi, gs.dlo}) * inv_nf;
        const float *cc = a.contrib + (size_t)(it & 1) * n;
        float *cn = a.contrib + (size_t)((it + 1) & 1) * n;
        if (gtid == 0 && it < kTraceRows) {
            DevTrace t;
            t.round = it;
            t.variant = SG_SYNC_PULL_TD;
            t.in_nodes = n;
            t.pad = 0;
            t.in_edges = g.m;
            t.ticks = (unsigned long long)(wall_clock64() - t_start);
            a.trace[it] = t;
        }
        const uint32_t ncc = min(n, kCcLds);
        for (uint32_t i = threadIdx.x; i < ncc; i += BS) s_cc[i] = cc[i];
        __syncthreads();
        auto contrib = [&](uint32_t u) -> float { return u < ncc ? s_cc[u] : cc[u]; };
        double r_acc = 0.0;
        CSum d_acc{0.f, 0.f};
        // x(v) from its in-edge sum: FP32 ranks and contributions, the
        // residual |x' - x| summed in FP64
        auto finish = [&](uint32_t v, const CSum &s) {
            const float xn = base + df * csum_value(s);
            r_acc += fabs((double)xn - (double)a.x[v]);
            a.x[v] = xn;
            const uint32_t dout = g.row[v + 1] - g.row[v];
            if (dout) cn[v] = xn / (float)dout;
            else csum_add(d_acc, xn);
        };
        // (A1) chunk sums of the in-hubs, kQ chunks per wave together: every
        // in-edge load of the kQ chunks is issued before any contribution is
        // gathered, and every gather before any is added (each lane still
        // adds its edges of a chunk in edge order, so the sums are the same)
        constexpr int kQ = 4, kSteps = (int)(kPullChunk / kWave);
        for (uint32_t c0 = gwave * kQ; c0 < g.nhugein; c0 += nwaves * kQ) {
            uint32_t cb[kQ], ce[kQ];
#pragma unroll
            for (int q = 0; q < kQ; q++) {
                cb[q] = ce[q] = 0;
                if (c0 + q < g.nhugein) {
                    const uint2 dsc = g.hugein[c0 + q];
                    cb[q] = dsc.y;
                    ce[q] = min(g.cstart[dsc.x + 1], dsc.y + kPullChunk);
                }
            }
            uint32_t u[kQ][kSteps];
#pragma unroll
            for (int q = 0; q < kQ; q++)
#pragma unroll
                for (int t = 0; t < kSteps; t++) {
                    const uint32_t e = cb[q] + (uint32_t)(t * kWave + ln);
                    u[q][t] = e < ce[q] ? g.csrc[e] : 0xFFFFFFFFu;
                }
            float cv[kQ][kSteps];
#pragma unroll
            for (int q = 0; q < kQ; q++)
#pragma unroll
                for (int t = 0; t < kSteps; t++) cv[q][t] = u[q][t] != 0xFFFFFFFFu ? contrib(u[q][t]) : 0.f;
#pragma unroll
            for (int q = 0; q < kQ; q++) {
                if (c0 + q >= g.nhugein) break;  // wave-uniform
                CSum s{0.f, 0.f};
#pragma unroll
                for (int t = 0; t < kSteps; t++)
                    if (u[q][t] != 0xFFFFFFFFu) csum_add(s, cv[q][t]);
                s = wave_csum(s);
                if (ln == 0) a.hubpart[c0 + q] = s;
            }
        }
        // (A2) every other vertex, a 64-vertex word per wave at a time: from
        // the sliced in-edges (each lane sums its own vertex's row entries in
        // order) or, without them, by a scheduled walk of the word's lists
        for (uint32_t w = gwave; w < W; w += nwaves) {
            const uint32_t v = w * 64 + ln;
            uint32_t b = 0, e = 0;
            bool own = false;
            if (v < n) {
                b = g.cstart[v];
                e = g.cstart[v + 1];
                own = e - b <= kPullHubDegree;
                if (!own) b = e;
            }
            CSum s{0.f, 0.f};
            if (g.sell_src) {
                const uint32_t beg = g.sell_off[w];
                const uint32_t rows = (g.sell_off[w + 1] - beg) / kWave;
                const uint32_t *S = g.sell_src + beg + ln;
                constexpr int U = 4;
                for (uint32_t j0 = 0; j0 < rows; j0 += U) {
                    uint32_t u[U];
#pragma unroll
                    for (int q = 0; q < U; q++) u[q] = (own && j0 + q < rows) ? S[(size_t)(j0 + q) * kWave] : 0xFFFFFFFFu;
                    float c[U];
#pragma unroll
                    for (int q = 0; q < U; q++) c[q] = u[q] != 0xFFFFFFFFu ? contrib(u[q]) : 0.f;
#pragma unroll
                    for (int q = 0; q < U; q++)
                        if (u[q] != 0xFFFFFFFFu) csum_add(s, c[q]);
                }
            } else {
                s = wave_gather_csum(b, e, [&](uint32_t ed) { return contrib(g.csrc[ed]); });
            }
            if (own) finish(v, s);
        }
        ok = grid_sync(a.bar, xcc, gen, a.abort);
        if (!ok) break;
        // (B) in-hubs: chunk sums in chunk order, one wave per hub
        for (uint32_t h = gwave; h < g.nhubin; h += nwaves) {
            const uint2 hb = g.hubin[h];
            CSum s{0.f, 0.f};
            for (uint32_t i = ln; i < hb.y; i += kWave) s = csum_merge(s, a.hubpart[hb.x + i]);
            s = wave_csum(s);
            if (ln == 0) finish(g.hugein[hb.x].x, s);
        }
        {
            const PrPart t = block_part<WPB>(r_acc, d_acc, s_red);
            if (threadIdx.x == 0) a.part[(size_t)((it + 1) & 1) * kMaxBlocks + blockIdx.x] = t;
        }
        it++;
        ok = grid_sync(a.bar, xcc, gen, a.abort);
    }

    if (gtid == 0) {
        DevStats s{};
        s.rounds = it;
        s.pull_rounds = it;
        s.iterations = it;
        s.residual = res;
        s.status = ok ? SG_OK : SG_ERR_TIMEOUT;
        *a.stats = s;
        if (it < kTraceRows) {
            DevTrace t{};
            t.round = it;
            t.variant = -1;
            t.ticks = (unsigned long long)(wall_clock64() - t_start);
            a.trace[it] = t;
        }
    }
}

// ------------------------------------------------------- residual DD kernel

struct PrDdArgs {
    DevGraph g;
    float *x;          // normalised ranks [n] (output)
    CSum *y;           // unnormalised ranks [n], compensated FP32
    float *r;          // residuals [n]: r + rlo, the rounding error of every
    float *rlo;        // atomic addition into r collected in rlo
    float *p;          // d r / outdeg of this round's active vertices [n]
    uint64_t *abit;    // 2 x W: active vertices of rounds k (k % 2)
    uint32_t *queue;   // 2 x qcap: vertices whose residual crossed theta
    uint32_t qcap;
    FrontierCtr *ctr;  // [0,1]: activation totals; [2,3,4]: queue appends
    PrPart *part;      // kMaxBlocks partials for the final sums
    GridBarrier bar;
    uint32_t *abort;
    DevStats *stats;
    DevTrace *trace;
    double damping, theta, alpha;
    double total_wl;
    int32_t force;     // -1 auto, 0 push, 1 pull
    uint32_t max_rounds;
    uint32_t probe;
    uint32_t iterative;  // rounds chosen by the ITERATIVE scheme's predicted times
    int32_t it_force;    // ITERATIVE: -1, or the variant every round runs (kItPullTd...)
    float ticks_per_us;  // wall clock rate
    float cut;           // sample-based priority for ASYNC_PUSH_TD rounds (0: off)
};

// Sample-based priority (GetPriorityThreshold, include/framework/algo_variants.cuh:98-131;
// PRIORITY_SAMPLE_SIZE, graph_datum.cuh:18): a round's threshold is the
// residual at index size * (1 - cut) of size sampled residuals in ascending
// order, and only vertices at or above it push (IsHighPriority,
// samples/hybrid_pr/hybrid_pr.cu:117-121).
constexpr uint32_t kPrioSample = 1000;
constexpr uint32_t kPrioSort = 1024;  // sample padded to a power of two

__device__ __forceinline__ uint64_t sample_hash(uint64_t z)
{
    // SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014)
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Every block draws the same vertices (a hash of round and index) and sorts
// their residuals in LDS (bitonic); residuals other blocks have already moved
// out this round read as 0, which can only lower a block's threshold.
template <int BS>
__device__ double sampled_threshold(const PrDdArgs &a, uint32_t k, double *s_smp)
{
    const uint32_t n = a.g.n, size = min(kPrioSample, n);
    for (uint32_t i = threadIdx.x; i < kPrioSort; i += BS)
        s_smp[i] = i < size ? (double)a.r[(uint32_t)(sample_hash(((uint64_t)k << 32) | i) % n)] : __builtin_inf();
    __syncthreads();
    for (uint32_t len = 2; len <= kPrioSort; len <<= 1) {
        for (uint32_t j = len >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < kPrioSort; i += BS) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const double x = s_smp[i], y = s_smp[p];
                    if ((x > y) == ((i & len) == 0)) {
                        s_smp[i] = y;
                        s_smp[p] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    const uint32_t at = min((uint32_t)((float)size * (1.0f - a.cut)), size - 1);
    const double t = s_smp[at];
    __syncthreads();
    return t;
}

// The ITERATIVE scheme's variant set in its std::set order
// (include/framework/hybrid_policy.h:172-186; AlgoVariant::operator< compares
// model, message passing, scheduling with SYNC < ASYNC, PUSH < PULL, TD < DD).
enum { kItPullTd = 0, kItPushTd = 1, kItPushDd = 2 };

// Replicated in every block: each computes the same decisions from the same
// barrier stamps and counters (PolicyDecisionMaker::GetNextPolicy and
// PredicateRunningTime, hybrid_policy.h:130-169,246-264; times are float as
// in its RunningInfo).
struct IterativePolicy {
    float time[3] = {0.f, 0.f, 0.f};  // last measured round time of each variant (us)
    float time_per_work = 0.f;
    float last_in_wl = 0.f;  // input workload (out-degree sum) of the last round
    int last = -1;

    // Record the round just finished (variant `v`, `us` long, `in_wl` input edges).
    __device__ void record(int v, float us, float in_wl)
    {
        time[v] = us;
        last = v;
        last_in_wl = in_wl;
    }

    // Variant for round k, `out_wl` = out-degree sum of the current worklist.
    __device__ int next(uint32_t k, float out_wl)
    {
        if (k < 3) return (int)k;  // inspection
        if (last == kItPushDd) time_per_work = time[kItPushDd] / last_in_wl;
        const float pred[3] = {time[kItPullTd], time[kItPushTd], time_per_work * out_wl};
        int best = 0;
        for (int v = 1; v < 3; v++)
            if (pred[v] < pred[best]) best = v;  // std::min_element: first minimum
        return best;
    }
};

struct DdRound {
    const PrDdArgs *a;
    double theta;
    Stager S;
    unsigned long long count;
    unsigned long long qscout;  // out-degree sum of queued vertices (ITERATIVE)
    bool qdeg;
};

// Add `add` to v's residual (atomically when `atomic`), queueing v if this
// addition took its residual across theta. An atomic addition returns the
// old value, so the lane recomputes the rounded sum the atomic stored and
// adds its rounding error (TwoSum, Knuth) to rlo[v]. Wave-convergent.
__device__ __forceinline__ void dd_accumulate(DdRound &R, bool act, uint32_t v, CSum add, bool atomic)
{
    bool crossed = false;
    const float ad = csum_value(add);
    if (act && ad > 0.f) {
        float old;
        if (atomic) {
            old = atomicAdd(&R.a->r[v], add.hi);
            const float s = old + add.hi;
            const float bb = s - old;
            const float err = (old - (s - bb)) + (add.hi - bb) + add.lo;
            if (err != 0.f) atomicAdd(&R.a->rlo[v], err);
        } else {
            CSum c{R.a->r[v], R.a->rlo[v]};
            old = c.hi;
            c = csum_merge(c, add);
            R.a->r[v] = c.hi;
            R.a->rlo[v] = c.lo;
        }
        crossed = (double)old < R.theta && (double)old + (double)ad >= R.theta;
        R.count += crossed;
        if (crossed && R.qdeg) R.qscout += R.a->g.row[v + 1] - R.a->g.row[v];
    }
    stage(R.S, crossed, v);
}

// Push the active vertices of one word (one lane per vertex): wave-level
// prefix-sum scheduling of their out-edges; hubs go through the chunk list.
__device__ __forceinline__ void dd_push_word(DdRound &R, uint64_t word, uint32_t w)
{
    const PrDdArgs &a = *R.a;
    const DevGraph &g = a.g;
    const int ln = lane();
    const uint32_t u = w * 64 + ln;
    uint32_t b = 0, dg = 0;
    float pu = 0.f;
    if ((word >> ln) & 1) {
        b = g.row[u];
        dg = g.row[u + 1] - b;
        if (dg > kHugeDegree) dg = 0;
        if (dg) pu = a.p[u];
    }
    const uint32_t incl = wave_incl_scan(dg);
    const uint32_t total = shfl(incl, kWave - 1);
    const uint32_t off = b - (incl - dg);
    for (uint32_t jb = 0; jb < total; jb += kWave) {
        const uint32_t j = jb + ln;
        const bool act = j < total;
        const int k = owner_lane(incl, j);
        const uint32_t e = shfl(off, k) + j;
        const float pk = __shfl(pu, k, kWave);
        const uint32_t v = act ? g.col[e] : 0u;
        dd_accumulate(R, act, v, CSum{pk, 0.f}, true);
    }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_pr_dd(PrDdArgs a)
{
    if (a.probe) {
        residency_probe(a.bar.words, 1u << 20);
        return;
    }
    constexpr int WPB = BS / kWave;
    constexpr uint32_t kLdsQ = 4u * BS;
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
    const double d = a.damping, theta = a.theta;
    FrontierCtr *act_ctr = a.ctr;       // 2 slots
    FrontierCtr *q_ctr = a.ctr + 2;     // 3 slots

    __shared__ double s_smp[kPrioSort];
    // pull rounds test the active bits of the lowest ids (in a degree-ordered
    // graph the sources of most in-edges) in an LDS copy of the bitmap
#ifndef SG_PRDD_LDS
#define SG_PRDD_LDS 8
#endif
    constexpr uint32_t kAbLds = SG_PRDD_LDS * BS;
    __shared__ uint64_t s_ab[kAbLds];
    __shared__ PrPart s_red2[WPB];
    __shared__ PrPart s_glob;
    __shared__ unsigned long long s_red[2][WPB];
    __shared__ unsigned long long s_lvl[5];
    __shared__ uint32_t s_q[kLdsQ];
    __shared__ uint32_t s_qn, s_flush[2];

    const float r0 = (float)((1.0 - d) / (double)n);
    const float df = (float)d;
    for (uint32_t v = gtid; v < n; v += nthreads) {
        a.y[v] = CSum{0.f, 0.f};
        a.r[v] = r0;
        a.rlo[v] = 0.f;
    }
    for (uint32_t w = gtid; w < 2 * W; w += nthreads) a.abit[w] = 0;
    if (blockIdx.x == 0) {
        for (int s = 0; s < 5; s++) clear_ctr(&a.ctr[s]);
    }
    // stamped barriers: the block completing each one records the wall clock,
    // so every block derives the same round times for the ITERATIVE scheme
    bool ok = grid_sync<true>(a.bar, xcc, gen, a.abort);
    __shared__ unsigned long long s_stamp;
    IterativePolicy ipol;
    unsigned long long stamp0 = 0, stamp_prev = 0;
    float in_wl_prev = 0.f;
    int v_prev = -1;
    bool held = false;  // the last activation held residuals >= theta back (priority)

    uint32_t k = 0, pushes = 0, pulls = 0;
    unsigned long long ptrace = 0;
    while (ok) {
        if (a.max_rounds && k >= a.max_rounds) break;
        // ---- activation: move residual >= theta into y, set p and the bit
        if (threadIdx.x == 0) s_stamp = barrier_stamp(a.bar);
        read_counts(&q_ctr[k % 3], s_lvl);
        __syncthreads();
        const RoundCounts qc = unpack_counts(s_lvl, a.qcap);
        const unsigned long long stamp = s_stamp;
        __syncthreads();
        // ITERATIVE: time the round that just ended, then pick this round's
        // variant (topology-driven ones rebuild the active set by a scan)
        int var = -1;
        if (a.iterative) {
            if (k == 0) {
                stamp0 = stamp;
            } else {
                const float us = (float)((double)(stamp - stamp_prev) / (double)a.ticks_per_us);
                ipol.record(v_prev, us, in_wl_prev);
                if (gtid == 0 && k - 1 < kTraceRows) a.trace[k - 1].ticks = stamp - stamp0;
            }
            stamp_prev = stamp;
            var = a.it_force >= 0 ? a.it_force : ipol.next(k, (float)qc.scout);
        }
        uint64_t *ab = a.abit + (size_t)(k & 1) * W;
        unsigned long long scout = 0, cnt = 0;
        auto activate = [&](uint32_t u, float ru) {
            CSum yu = a.y[u];
            csum_add(yu, ru);
            a.y[u] = yu;
            const uint32_t dout = g.row[u + 1] - g.row[u];
            if (dout) a.p[u] = df * ru / (float)dout;
            scout += dout;
            cnt++;
        };
        bool waiting = false;  // this lane saw a residual >= theta left below the priority threshold
        // a scan also after a round that held residuals back: they did not
        // cross theta again, so no queue holds them
        if (k == 0 || qc.overflow || held || var == kItPullTd || var == kItPushTd) {
            double thr = theta;
            if (var == kItPushTd && a.cut > 0.f) thr = fmax(theta, sampled_threshold<BS>(a, k, s_smp));
            for (uint32_t w = gwave; w < W; w += nwaves) {
                const uint32_t v = w * 64 + ln;
                bool act = false;
                if (v < n) {
                    const float rv = a.r[v] + a.rlo[v];
                    act = (double)rv >= thr;
                    waiting |= !act && (double)rv >= theta;
                    if (act) {
                        a.r[v] = 0.f;
                        a.rlo[v] = 0.f;
                        activate(v, rv);
                    }
                }
                const uint64_t m = ballot(act);
                if (ln == 0) ab[w] = m;
            }
        } else {
            const uint32_t *qin = a.queue + (size_t)(k & 1) * a.qcap;
            for (uint32_t i = gtid; i < qc.qlen; i += nthreads) {
                const uint32_t u = qin[i];
                // claim the residual: a vertex queued twice is activated once
                // (no propagation runs in this phase; a claim split between
                // two copies returns the part it took)
                const float hi = atomicExch(&a.r[u], 0.f);
                const float lo = atomicExch(&a.rlo[u], 0.f);
                const float ru = hi + lo;
                if ((double)ru >= theta) {
                    atomicOr((unsigned long long *)&ab[u >> 6], 1ull << (u & 63));
                    activate(u, ru);
                } else if (hi != 0.f || lo != 0.f) {
                    atomicAdd(&a.r[u], hi);
                    atomicAdd(&a.rlo[u], lo);
                }
            }
        }
        // the slot's has_huge flag carries "residuals >= theta were held back"
        publish_counts<WPB>(&act_ctr[k & 1], scout, cnt, waiting, s_red);
        ok = grid_sync(a.bar, xcc, gen, a.abort);
        if (!ok) break;

        // ---- propagation
        read_counts(&act_ctr[k & 1], s_lvl);
        if (threadIdx.x == 0) s_qn = 0;
        __syncthreads();
        const RoundCounts ac = unpack_counts(s_lvl, a.qcap);
        __syncthreads();
        if (ac.count == 0 && !ac.has_huge) break;  // every residual is below theta
        held = ac.has_huge;
        bool pull;
        if (var >= 0) pull = var == kItPullTd;
        else if (a.force >= 0) pull = a.force == 1;
        else pull = (double)ac.scout > a.total_wl / a.alpha;
        if (pull) {
            pulls++;
            if (k < 64) ptrace |= 1ull << k;
        } else {
            pushes++;
        }
        if (var >= 0) {
            // the reference's input workload: edges for a topology-driven
            // round, the worklist's out-degree sum for a data-driven one
            in_wl_prev = var == kItPushDd ? (float)ac.scout : (float)g.m;
            v_prev = var;
            if (gtid == 0 && k < kTraceRows) {
                DevTrace t{};
                t.round = k;
                t.variant = var == kItPullTd ? SG_SYNC_PULL_TD : var == kItPushTd ? SG_ASYNC_PUSH_TD : SG_ASYNC_PUSH_DD;
                t.in_nodes = (uint32_t)ac.count;
                t.in_edges = ac.scout;
                a.trace[k] = t;  // ticks filled in when the round's end is known
            }
        } else if (gtid == 0) {
            write_trace(a.trace, k, pull, ac.count, ac.scout, t_start);
        }
        if (blockIdx.x == 0) {
            clear_ctr(&act_ctr[(k + 1) & 1]);
            clear_ctr(&q_ctr[(k + 2) % 3]);
        }
        uint64_t *ab_next = a.abit + (size_t)((k + 1) & 1) * W;
        for (uint32_t w = gtid; w < W; w += nthreads) ab_next[w] = 0;

        DdRound R;
        R.a = &a;
        R.theta = theta;
        R.S = Stager{a.queue + (size_t)((k + 1) & 1) * a.qcap, a.qcap, &q_ctr[(k + 1) % 3], s_q, &s_qn, kLdsQ};
        R.count = 0;
        R.qscout = 0;
        R.qdeg = a.iterative != 0;
        if (!pull) {
            for (uint32_t c0 = 0; c0 < g.nhuge; c0 += nwaves * kWave) {
                const uint32_t c = c0 + (uint32_t)ln * nwaves + gwave;
                uint2 dsc = make_uint2(0u, 0u);
                bool mem = false;
                if (c < g.nhuge) {
                    dsc = g.huge[c];
                    mem = test_bit(ab, dsc.x);
                }
                uint64_t mm = ballot(mem);
                while (mm) {
                    const int t = first_lane(mm);
                    mm &= mm - 1;
                    const uint32_t u = shfl(dsc.x, t), beg = shfl(dsc.y, t);
                    const uint32_t end = min(g.row[u + 1], beg + kHugeChunk);
                    const float pu = a.p[u];
                    for (uint32_t e0 = beg; e0 < end; e0 += kWave) {
                        const uint32_t e = e0 + ln;
                        const bool act = e < end;
                        dd_accumulate(R, act, act ? g.col[e] : 0u, CSum{pu, 0.f}, true);
                    }
                }
            }
            for (uint32_t w = gwave; w < W; w += nwaves) {
                const uint64_t word = ab[w];
                if (word) dd_push_word(R, word, w);
            }
        } else {
            const uint32_t nab = min(W, kAbLds);
            for (uint32_t i = threadIdx.x; i < nab; i += BS) s_ab[i] = ab[i];
            __syncthreads();
            auto active = [&](uint32_t u) -> bool {
                const uint32_t w = u >> 6;
                return ((w < nab ? s_ab[w] : ab[w]) >> (u & 63)) & 1;
            };
            for (uint32_t c = gwave; c < g.nhugein; c += nwaves) {
                const uint2 dsc = g.hugein[c];
                const uint32_t end = min(g.cstart[dsc.x + 1], dsc.y + kPullChunk);
                CSum s{0.f, 0.f};
                for (uint32_t e0 = dsc.y; e0 < end; e0 += kWave) {
                    const uint32_t e = e0 + ln;
                    if (e < end) {
                        const uint32_t u = g.csrc[e];
                        if (active(u)) csum_add(s, a.p[u]);
                    }
                }
                s = wave_csum(s);
                dd_accumulate(R, ln == 0, dsc.x, s, true);
            }
            for (uint32_t w = gwave; w < W; w += nwaves) {
                const uint32_t v = w * 64 + ln;
                uint32_t b = 0, e = 0;
                bool own = false;
                if (v < n) {
                    b = g.cstart[v];
                    e = g.cstart[v + 1];
                    own = e - b <= kPullHubDegree;
                    if (!own) b = e;
                }
                const CSum s = wave_gather_csum(b, e, [&](uint32_t ed) {
                    const uint32_t u = g.csrc[ed];
                    return active(u) ? a.p[u] : 0.f;
                });
                dd_accumulate(R, own, v, s, false);
            }
        }
        publish_counts<WPB>(&q_ctr[(k + 1) % 3], R.qscout, R.count, false, s_red);
        flush_stage(R.S, s_flush);
        k++;
        ok = grid_sync<true>(a.bar, xcc, gen, a.abort);
    }

    // ---- absorb what residual is left, then normalise
    // the sum of y (compensated FP32) normalises; the residual mass absorbed
    // here is reduced in FP64
    CSum ysum{0.f, 0.f};
    double rsum = 0.0;
    if (ok) {
        for (uint32_t v = gtid; v < n; v += nthreads) {
            CSum yv = a.y[v];
            const float rv = a.r[v], rl = a.rlo[v];
            csum_add(yv, rv);
            csum_add(yv, rl);
            a.y[v] = yv;
            csum_add(ysum, csum_value(yv));
            rsum += (double)rv + (double)rl;
        }
        const PrPart t = block_part<WPB>(rsum, ysum, s_red2);
        if (threadIdx.x == 0) a.part[blockIdx.x] = t;
        ok = grid_sync(a.bar, xcc, gen, a.abort);
    }
    PrPart tot{0.0, 1.f, 0.f};
    if (ok) {
        tot = grid_part<BS>(a.part, gridDim.x, s_red2, &s_glob);
        const float ys = csum_value(CSum{tot.dhi, tot.dlo});
        const float inv = ys > 0.f ? 1.0f / ys : 0.f;
        for (uint32_t v = gtid; v < n; v += nthreads) a.x[v] = csum_value(a.y[v]) * inv;
    }
    if (gtid == 0) {
        DevStats s{};
        s.rounds = k;
        s.push_rounds = pushes;
        s.pull_rounds = pulls;
        s.trace = ptrace;
        s.iterations = k;
        s.residual = tot.res;  // residual mass absorbed at the end (unnormalised)
        s.status = ok ? SG_OK : SG_ERR_TIMEOUT;
        *a.stats = s;
        if (k < kTraceRows) {
            DevTrace t{};
            t.round = k;
            t.variant = -1;
            t.ticks = (unsigned long long)(wall_clock64() - t_start);
            a.trace[k] = t;
        }
    }
}

template <typename Args, void (*K)(Args)>
int launch_persistent(sg_engine *e, Args &args, int bs)
{
    sg_engine::PGrid *pg = nullptr;
    auto probe = [e, args, bs](int blocks, uint32_t *words) -> int {
        Args p = args;
        p.probe = 1;
        p.bar.words = words;
        void *params[] = {&p};
        SG_HIP(hipLaunchKernel((const void *)K, dim3(blocks), dim3(bs), params, 0, e->stream));
        return SG_OK;
    };
    int rc = engine_persistent_grid(e, (const void *)K, bs, probe, &pg);
    if (rc) return rc;
    if (pg->blocks > (int)kMaxBlocks) SG_FAIL(SG_ERR_STATE, "sg_pagerank: grid exceeds the partial-sum area");
    bind_barrier(pg, args.bar);
    return engine_launch_persistent(e, [&]() -> int {
        SG_HIP(hipEventRecord(e->ev0, e->stream));
        void *params[] = {&args};
        SG_HIP(hipLaunchKernel((const void *)K, dim3(pg->blocks), dim3(bs), params, 0, e->stream));
        pg->launches++;
        SG_HIP(hipEventRecord(e->ev1, e->stream));
        return SG_OK;
    });
}

}  // namespace

extern "C" int sg_pagerank(sg_engine *e, double damping, double tol, int32_t max_iter, int32_t mode)
{
    if (!e) SG_FAIL(SG_ERR_INVALID, "sg_pagerank: null engine");
    if (!(damping >= 0.0 && damping < 1.0)) SG_FAIL(SG_ERR_RANGE, "sg_pagerank: damping must be in [0, 1)");
    if (!(tol > 0.0)) SG_FAIL(SG_ERR_RANGE, "sg_pagerank: tol must be > 0");
    if (max_iter < 0) SG_FAIL(SG_ERR_RANGE, "sg_pagerank: max_iter < 0");
    const sg_engine_options &o = e->opt;
    if (o.lb_push == SG_LB_NONE || o.lb_pull == SG_LB_NONE)
        SG_FAIL(SG_ERR_UNSUPPORTED, "sg_pagerank: LoadBalancing::NONE applies to sg_bfs and sg_sssp only");
    if (mode == SG_PR_AUTO) {
        const int fv = o.force_variant;
        mode = (fv == SG_ASYNC_PUSH_DD || fv == SG_ASYNC_PULL_DD || fv == SG_SYNC_PUSH_DD || fv == SG_SYNC_PULL_DD)
                   ? SG_PR_PUSH_DD
                   : SG_PR_PULL_TD;
    }
    if (mode != SG_PR_PULL_TD && mode != SG_PR_PUSH_DD && mode != SG_PR_ITERATIVE)
        SG_FAIL(SG_ERR_INVALID, "sg_pagerank: unknown mode");
    int32_t it_force = -1;
    if (mode == SG_PR_ITERATIVE) {
        switch (o.force_variant) {
        case SG_VARIANT_AUTO: break;
        case SG_SYNC_PULL_TD: it_force = kItPullTd; break;
        case SG_ASYNC_PUSH_TD: it_force = kItPushTd; break;
        case SG_ASYNC_PUSH_DD: it_force = kItPushDd; break;
        default:
            SG_FAIL(SG_ERR_UNSUPPORTED,
                    "sg_pagerank: the ITERATIVE scheme runs SYNC_PULL_TD, ASYNC_PUSH_TD or ASYNC_PUSH_DD only");
        }
    }
    const DevGraph &d = e->g->dev;
    int rc = engine_ensure_bfs_state(e);
    if (rc) return rc;
    void *res = nullptr;
    if ((rc = engine_result_rows(e, 1, &res))) return rc;
    const uint64_t n = d.n;
    const int bs = o.block_size ? o.block_size : 1024;
    if (bs != 256 && bs != 512 && bs != 1024) SG_FAIL(SG_ERR_UNSUPPORTED, "sg_pagerank: block_size must be 256, 512 or 1024");
    e->last_algo = 2;
    const uint64_t part_bytes = 2ull * kMaxBlocks * sizeof(double2);
    if (mode == SG_PR_PULL_TD) {
        // contrib (2n floats) | hubpart | partials
        const uint64_t off_hub = (2 * n * 4 + 15) / 16 * 16;
        const uint64_t off_part = off_hub + ((uint64_t)d.nhugein * 8 + 15) / 16 * 16;
        if ((rc = engine_ensure_aux(e, off_part + part_bytes))) return rc;
        char *base = (char *)e->aux;
        PrArgs a{};
        a.g = d;
        a.x = (float *)res;
        a.contrib = (float *)base;
        a.hubpart = (CSum *)(base + off_hub);
        a.part = (PrPart *)(base + off_part);
        a.bar.spin_limit = 1u << 22;
        a.abort = e->abort_flag;
        a.stats = e->stats;
        a.trace = e->trace;
        a.damping = damping;
        a.tol = tol;
        a.max_iter = max_iter ? (uint32_t)max_iter : 1000u;
        switch (bs) {
        case 256: rc = launch_persistent<PrArgs, k_pr_pull<256>>(e, a, 256); break;
        case 512: rc = launch_persistent<PrArgs, k_pr_pull<512>>(e, a, 512); break;
        default: rc = launch_persistent<PrArgs, k_pr_pull<1024>>(e, a, 1024); break;
        }
    } else {
        const int force = mode == SG_PR_ITERATIVE ? -1 : forced_direction(o);
        if (force < -1) SG_FAIL(SG_ERR_UNSUPPORTED, "sg_pagerank: data-driven mode takes push/pull DD variants only");
        // y (n compensated floats) | r | rlo | p (n floats each) | partials
        const uint64_t off_part = (n * 8 + 3 * n * 4 + 15) / 16 * 16;
        if ((rc = engine_ensure_aux(e, off_part + part_bytes))) return rc;
        char *base = (char *)e->aux;
        PrDdArgs a{};
        a.g = d;
        a.x = (float *)res;
        a.y = (CSum *)base;
        a.r = (float *)(base + n * 8);
        a.rlo = (float *)(base + n * 12);
        a.p = (float *)(base + n * 16);
        a.part = (PrPart *)(base + off_part);
        a.abit = e->bitmaps;
        a.queue = e->queues;
        a.qcap = e->qcap;
        a.ctr = (FrontierCtr *)e->ctr;
        a.bar.spin_limit = 1u << 22;
        a.abort = e->abort_flag;
        a.stats = e->stats;
        a.trace = e->trace;
        a.damping = damping;
        a.theta = tol * (1.0 - damping) / (double)n;
        a.alpha = o.alpha == 0.0 ? 1.0 : o.alpha;
        a.total_wl = (double)d.m * o.edge_factor;
        a.force = force;
        a.max_rounds = max_iter ? (uint32_t)max_iter : 100000u;
        a.iterative = mode == SG_PR_ITERATIVE;
        a.it_force = it_force;
        a.ticks_per_us = (float)(e->wall_khz / 1000.0);
        a.cut = 0.f;
        if (mode == SG_PR_ITERATIVE && o.priority == SG_PRIO_SAMPLING) {
            if (!(o.cut_threshold > 0.f && o.cut_threshold < 1.f))
                SG_FAIL(SG_ERR_RANGE, "sg_pagerank: sample-based priority needs 0 < cut_threshold < 1");
            a.cut = o.cut_threshold;
        }
        switch (bs) {
        case 256: rc = launch_persistent<PrDdArgs, k_pr_dd<256>>(e, a, 256); break;
        case 512: rc = launch_persistent<PrDdArgs, k_pr_dd<512>>(e, a, 512); break;
        default: rc = launch_persistent<PrDdArgs, k_pr_dd<1024>>(e, a, 1024); break;
        }
    }
    if (rc) return rc;
    if ((rc = engine_finish(e, 1))) return rc;
    e->timed = true;
    return SG_OK;
}
