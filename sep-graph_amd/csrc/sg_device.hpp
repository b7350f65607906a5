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
// Blocks are grouped by kBarGroup; the last arrival of a group signals the top
// counter, the last group flips the generation word. Counters sit on separate
// 128-byte lines. All memory effects before the barrier are released at agent
// scope and acquired after it, which on CDNA writes back / invalidates the
// per-XCD L2 so data produced on one XCD is visible on every other.
constexpr int kBarGroup = 16;
constexpr int kBarLine = 32;  // uint32 per 128-byte line

struct GridBarrier {
    uint32_t *words;  // [0]: generation, [kBarLine]: top counter, [kBarLine*(2+g)]: group g
    uint32_t nblocks;
    uint32_t spin_limit;  // max polls before giving up (sets *abort)
};

__device__ __forceinline__ uint32_t atomic_load_relaxed(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void atomic_store_relaxed(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Returns false if the wait timed out (the caller must then stop its loop);
// *abort is set so every block observes the failure.
__device__ __forceinline__ bool grid_sync(const GridBarrier &gb, uint32_t *abort)
{
    __shared__ uint32_t s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t *gen = gb.words;
        uint32_t *top = gb.words + kBarLine;
        const uint32_t g = blockIdx.x / kBarGroup;
        const uint32_t ngroups = (gb.nblocks + kBarGroup - 1) / kBarGroup;
        const uint32_t gsize = min((uint32_t)kBarGroup, gb.nblocks - g * kBarGroup);
        uint32_t *grp = gb.words + kBarLine * (2 + g);
        const uint32_t gen0 = atomic_load_relaxed(gen);
        uint32_t ok = 1;
        bool released = false;
        uint32_t a = __hip_atomic_fetch_add(grp, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (a == gsize - 1) {
            atomic_store_relaxed(grp, 0u);
            uint32_t t = __hip_atomic_fetch_add(top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (t == ngroups - 1) {
                atomic_store_relaxed(top, 0u);
                __hip_atomic_store(gen, gen0 + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                released = true;
            }
        }
        if (!released) {
            uint32_t polls = 0;
            while (atomic_load_relaxed(gen) == gen0) {
                if (++polls > gb.spin_limit || atomic_load_relaxed(abort)) {
                    atomic_store_relaxed(abort, 1u);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

}  // namespace sg
