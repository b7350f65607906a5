// Device primitives for gfx950 (CDNA4, wave64): lane/ballot helpers,
// wave-aggregated queue appends and a bounded, two-level grid barrier for
// cooperatively launched persistent kernels.
//
// These replace the reference's cub warp primitives (ShuffleIndex<32>,
// LaneId, LaneMaskLt, WarpReduce) and cub::GridBarrier
// (include/groute/device/cta_scheduler_hybrid.cuh:150-253,
// include/groute/device/queue.cuh:60-114,
// include/framework/algo_variants.cuh:353-355). Ballots here are 64-bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sg {

constexpr int kWave = 64;

__device__ __forceinline__ int lane() { return (int)__lane_id(); }

// Number of lanes below this one whose bit is set in `mask`.
__device__ __forceinline__ int rank_in(uint64_t mask)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int first_lane(uint64_t mask) { return __ffsll((unsigned long long)mask) - 1; }

template <typename T>
__device__ __forceinline__ T shfl(T v, int src) { return __shfl(v, src, kWave); }

// Inclusive prefix sum across the wave (Hillis-Steele over __shfl_up).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
    const int l = lane();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t y = __shfl_up(x, d, kWave);
        if (l >= d) x += y;
    }
    return x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

__device__ __forceinline__ double wave_sum_f64(double x)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

// Append `v` to q (counter *cnt) for every lane with `pred`, one atomic per wave.
// Must be reached by the whole wave (convergent), with per-lane predicates.
__device__ __forceinline__ void wave_append(bool pred, uint32_t v, uint32_t *q, uint32_t *cnt)
{
    uint64_t m = ballot(pred);
    if (!m) return;
    int leader = first_lane(m);
    uint32_t base = 0;
    if (lane() == leader) base = atomicAdd(cnt, (uint32_t)__popcll(m));
    base = shfl(base, leader);
    if (pred) q[base + rank_in(m)] = v;
}

// ---------------------------------------------------------------- grid barrier
// A grid barrier for persistent kernels whose blocks are all resident.
//
// Blocks are grouped by the XCD they physically run on (hardware register
// XCC_ID; each XCD has its own L2). Data produced before the barrier reaches
// the writer's L2 when the block's waves pass __syncthreads (vector L1 is
// write-through); the last block to arrive on an XCD writes that L2 back once
// (agent-scope release) and signals the top counter, so the barrier costs one
// L2 write-back per XCD instead of one per block. The last XCD flips the
// generation word, which every block polls (system scope, so no cache can
// answer) before its agent-scope acquire. Relaying the flip through a per-XCD
// flag set by the XCD's leader measured slower on MI355X (one more store/poll
// hop on the critical path).
//
// Counters come in two sets used by alternate barriers (generation parity).
// A set's counters are zeroed by the blocks that finished with them -- the
// XCD leader after its top arrival, the last XCD after the flip -- and those
// stores drain before the same blocks arrive at the next barrier, so they
// land before the set is used again two barriers on. No store has to
// complete between the last arrival and the flip.
//
// How many blocks each XCD holds is counted on every launch: each block
// registers on its XCD when it starts (bar_begin), and before its first
// barrier a block waits until the whole grid has registered and reads the
// per-XCD counts (the census). Registration counters come in two sets chosen
// by the launch's parity (GridBarrier::rset); block 0 zeroes the other set,
// which only the previous launch on these words used, when it starts. So no
// launch depends on how blocks were dispatched in an earlier one.
constexpr int kBarLine = 32;      // uint32 per 128-byte line
constexpr int kMaxXcd = 16;       // XCC_ID is 4 bits wide
constexpr int kBarWords = kBarLine * (6 + 4 * kMaxXcd);

// Word offsets inside the barrier area.
constexpr int kBarGen = 0;
constexpr int kBarStamp = 2;  // 64-bit: wall clock when a stamped barrier completed
__host__ __device__ constexpr int bar_top(int set) { return kBarLine * (1 + set); }
__host__ __device__ constexpr int bar_arrive(int set, int x) { return kBarLine * (3 + set * kMaxXcd + x); }
__host__ __device__ constexpr int bar_count(int x) { return kBarLine * (3 + 2 * kMaxXcd) + x; }
// residency probe: blocks registered / blocks whose wait for the others expired
constexpr int kBarProbeTotal = kBarLine * (3 + 2 * kMaxXcd) + kMaxXcd;
constexpr int kBarProbeFail = kBarProbeTotal + 1;
// per-launch registration: blocks on XCD x, and blocks in all, per parity set
__host__ __device__ constexpr int bar_reg(int set, int x) { return kBarLine * (4 + 2 * kMaxXcd + set * kMaxXcd + x); }
__host__ __device__ constexpr int bar_regtotal(int set) { return kBarLine * (4 + 4 * kMaxXcd + set); }
static_assert(bar_regtotal(1) + kBarLine <= kBarWords, "barrier layout");

struct GridBarrier {
    uint32_t *words;          // kBarWords, zero-initialised counters
    uint32_t spin_limit;      // max polls before giving up (sets *abort)
    uint32_t rset;            // registration set of this launch (launch parity)
};

// Barrier state kept by thread 0 of each block.
struct BarState {
    uint32_t gen;   // generation the next barrier completes from
    uint32_t cnt;   // blocks of the grid on this block's XCD (0 before the census)
    uint32_t nxcd;  // XCDs holding at least one block of the grid
    uint32_t rank;  // this block's arrival order among its XCD's blocks
};

__device__ __forceinline__ uint32_t xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 15; }

__device__ __forceinline__ uint32_t atomic_load_relaxed(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void atomic_store_relaxed(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t atomic_add_relaxed(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Loads and stores of flags that another block changes while this one waits.
// An agent-scope atomic load (sc1 only) can keep hitting a cached copy across
// polls on gfx950 -- observed as barrier waits that never saw a release -- so
// polling uses system scope (sc0 sc1), which no cache level may satisfy.
__device__ __forceinline__ uint32_t poll_load(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ void flag_store(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Bounded wait until *p != v (want_ne) or *p == v (!want_ne).
__device__ __forceinline__ bool spin_wait(const uint32_t *p, uint32_t v, bool want_ne, uint32_t limit,
                                          uint32_t *abort)
{
    uint32_t polls = 0;
    for (;;) {
        const uint32_t x = poll_load(p);
        if (want_ne ? x != v : x == v) return true;
        if ((++polls & 63) == 0 && poll_load(abort)) return false;
        if (polls > limit) {
            flag_store(abort, 1u);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Start of a persistent kernel (thread 0's copy is the one used): read the
// generation, register this block on its XCD, and let block 0 zero the
// registration set of the other parity for the next launch.
__device__ __forceinline__ BarState bar_begin(const GridBarrier &gb, uint32_t xcc)
{
    BarState s{0u, 0u, 0u, 0u};
    if (threadIdx.x == 0) {
        uint32_t *w = gb.words;
        const int set = (int)(gb.rset & 1u);
        s.gen = poll_load(w + kBarGen);
        s.rank = atomic_add_relaxed(w + bar_reg(set, (int)xcc), 1u);
        atomic_add_relaxed(w + bar_regtotal(set), 1u);
        if (blockIdx.x == 0) {
            for (int x = 0; x < kMaxXcd; x++) flag_store(w + bar_reg(set ^ 1, x), 0u);
            flag_store(w + bar_regtotal(set ^ 1), 0u);
        }
    }
    return s;
}

// The census (thread 0): bounded wait until every block of the grid has
// registered, then the blocks per XCD (all[x], when given) and this block's
// share. The total and the per-XCD counts are read together and accepted only
// when they agree, so a registration whose two atomics land in either order
// is never half seen.
__device__ inline bool bar_census(const GridBarrier &gb, uint32_t xcc, BarState &s, uint32_t *abort,
                                  uint32_t *all = nullptr)
{
    const uint32_t *w = gb.words;
    const int set = (int)(gb.rset & 1u);
    uint32_t polls = 0;
    for (;;) {
        uint32_t c[kMaxXcd];
        const uint32_t tot = poll_load(w + bar_regtotal(set));
#pragma unroll
        for (int x = 0; x < kMaxXcd; x++) c[x] = poll_load(w + bar_reg(set, x));
        uint32_t sum = 0, nx = 0;
#pragma unroll
        for (int x = 0; x < kMaxXcd; x++) {
            sum += c[x];
            nx += c[x] != 0u;
        }
        if (tot == gridDim.x && sum == gridDim.x) {
            s.cnt = c[xcc];
            s.nxcd = nx;
            if (all)
                for (int x = 0; x < kMaxXcd; x++) all[x] = c[x];
            return true;
        }
        if ((++polls & 63) == 0 && poll_load(abort)) return false;
        if (polls > gb.spin_limit) {
            flag_store(abort, 1u);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Residency probe, run by a persistent kernel launched with its real grid and
// resources before first use: every block counts itself on its XCD and then
// waits (bounded) until all blocks of the grid have done so. A block whose
// wait expires records a failure: the grid is then larger than what can be
// co-resident and the host shrinks it. Every block exits either way. All of
// a block's waves stay alive until thread 0 is done (the closing
// __syncthreads), so a probe block occupies exactly the wave slots, registers
// and LDS of a working block.
__device__ inline void residency_probe(uint32_t *words, uint32_t limit)
{
    if (threadIdx.x == 0) {
        atomicAdd(&words[bar_count(xcc_id())], 1u);
        atomicAdd(&words[kBarProbeTotal], 1u);
        uint32_t polls = 0;
        while (poll_load(&words[kBarProbeTotal]) < gridDim.x) {
            if (++polls > limit) {
                atomicAdd(&words[kBarProbeFail], 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
}

// Returns false if the wait timed out or another block aborted; the caller
// must then leave its loop. `st` is this block's state from bar_begin (only
// thread 0's copy is used): the first barrier of a launch takes the census,
// and each barrier advances the generation.
// kStamp: the block that completes the barrier also records the wall clock
// (barrier_stamp), so every block can time the grid's rounds identically.
template <bool kStamp = false>
__device__ __forceinline__ bool grid_sync(const GridBarrier &gb, uint32_t xcc, BarState &st, uint32_t *abort)
{
    __shared__ uint32_t s_ok;
    // __syncthreads() alone only waits for LDS traffic: every wave must first
    // drain its outstanding global stores/atomics (vmcnt) so they are in the
    // XCD's L2 before the leader writes that L2 back.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t *w = gb.words;
        uint32_t ok = st.cnt != 0u || bar_census(gb, xcc, st, abort);
        if (ok) {
            const uint32_t gen0 = st.gen;
            st.gen = gen0 + 1;
            const int set = (int)(gen0 & 1u);
            uint32_t *arr = w + bar_arrive(set, (int)xcc);
            uint32_t *top = w + bar_top(set);
            const uint32_t a = atomic_add_relaxed(arr, 1u);
            if (a == st.cnt - 1) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                const uint32_t t = atomic_add_relaxed(top, 1u);
                flag_store(arr, 0u);  // this set is used again two barriers on
                if (t == st.nxcd - 1) {
                    if (kStamp) {
                        // the stamp must be visible before the flip that publishes it
                        __hip_atomic_store((uint64_t *)(w + kBarStamp), (uint64_t)wall_clock64(), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_SYSTEM);
                        __builtin_amdgcn_s_waitcnt(0);
                    }
                    flag_store(w + kBarGen, gen0 + 1);
                    flag_store(top, 0u);
                }
            }
            // every block polls the generation word itself (a relay through a
            // per-XCD flag costs one more store/poll hop on the critical path)
            ok = spin_wait(w + kBarGen, gen0, true, gb.spin_limit, abort);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __builtin_amdgcn_s_waitcnt(0);
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// Wall clock recorded by the last stamped barrier (read after it).
__device__ __forceinline__ uint64_t barrier_stamp(const GridBarrier &gb)
{
    return __hip_atomic_load((const uint64_t *)(gb.words + kBarStamp), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace sg
