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
// The number of blocks each XCD receives for a given grid is measured once per
// engine by a census kernel (dispatch is deterministic); a launch whose blocks
// land differently cannot complete a group, runs into the bounded spin and
// reports a timeout instead of hanging.
constexpr int kBarLine = 32;      // uint32 per 128-byte line
constexpr int kMaxXcd = 16;       // XCC_ID is 4 bits wide
constexpr int kBarWords = kBarLine * (4 + 2 * kMaxXcd);

// Word offsets inside the barrier area.
constexpr int kBarGen = 0;
constexpr int kBarStamp = 2;  // 64-bit: wall clock when a stamped barrier completed
__host__ __device__ constexpr int bar_top(int set) { return kBarLine * (1 + set); }
__host__ __device__ constexpr int bar_arrive(int set, int x) { return kBarLine * (3 + set * kMaxXcd + x); }
__host__ __device__ constexpr int bar_count(int x) { return kBarLine * (3 + 2 * kMaxXcd) + x; }
// residency probe: blocks registered / blocks whose wait for the others expired
constexpr int kBarProbeTotal = kBarLine * (3 + 2 * kMaxXcd) + kMaxXcd;
constexpr int kBarProbeFail = kBarProbeTotal + 1;

struct GridBarrier {
    uint32_t *words;          // kBarWords, zero-initialised counters
    uint32_t nxcd;            // XCDs that hold at least one block of the grid
    uint32_t spin_limit;      // max polls before giving up (sets *abort)
    uint32_t cnt[kMaxXcd];    // blocks of the grid on each XCD (passed by value)
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

// Generation the barrier words are at when a kernel starts (thread 0 keeps it
// in a register and advances it per barrier, so no barrier re-reads it).
__device__ __forceinline__ uint32_t grid_gen(const GridBarrier &gb) { return poll_load(gb.words + kBarGen); }

// Returns false if the wait timed out or another block aborted; the caller
// must then leave its loop. `gen` is this block's generation (grid_gen at
// kernel start), advanced on success; only thread 0's copy is used.
// kStamp: the block that completes the barrier also records the wall clock
// (barrier_stamp), so every block can time the grid's rounds identically.
template <bool kStamp = false>
__device__ __forceinline__ bool grid_sync(const GridBarrier &gb, uint32_t xcc, uint32_t &gen, uint32_t *abort)
{
    __shared__ uint32_t s_ok;
    // __syncthreads() alone only waits for LDS traffic: every wave must first
    // drain its outstanding global stores/atomics (vmcnt) so they are in the
    // XCD's L2 before the leader writes that L2 back.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t *w = gb.words;
        const uint32_t gen0 = gen;
        gen = gen0 + 1;
        const uint32_t cnt = gb.cnt[xcc];
        const int set = (int)(gen0 & 1u);
        uint32_t ok = 1;
        uint32_t *arr = w + bar_arrive(set, (int)xcc);
        uint32_t *top = w + bar_top(set);
        const uint32_t a = atomic_add_relaxed(arr, 1u);
        if (a == cnt - 1) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            const uint32_t t = atomic_add_relaxed(top, 1u);
            flag_store(arr, 0u);  // this set is used again two barriers on
            if (t == gb.nxcd - 1) {
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
