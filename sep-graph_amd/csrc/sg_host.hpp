// Host-side helpers shared by the library's C++ sources: a thread pool-free
// parallel_for, the counter-based SplitMix64 generator and the thread-local
// error message behind sg_last_error().
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

namespace sg {

// Number of host worker threads: SG_THREADS if set, else hardware concurrency
// capped at 16 (the host share a GPU process gets on an MI355X node).
inline int host_threads()
{
    if (const char *s = std::getenv("SG_THREADS")) {
        int t = std::atoi(s);
        if (t > 0) return t;
    }
    unsigned hc = std::thread::hardware_concurrency();
    if (hc == 0) hc = 1;
    return (int)std::min(hc, 16u);
}

// Calls fn(begin, end, thread_index) on [0, n) split into contiguous ranges.
template <typename F>
void parallel_for(uint64_t n, F fn, int nthreads = 0)
{
    if (nthreads <= 0) nthreads = host_threads();
    if (n < 65536 || nthreads == 1) {
        fn((uint64_t)0, n, 0);
        return;
    }
    std::vector<std::thread> pool;
    uint64_t step = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; t++) {
        uint64_t b = std::min<uint64_t>(n, step * t), e = std::min<uint64_t>(n, b + step);
        if (b >= e) break;
        pool.emplace_back([=, &fn]() { fn(b, e, t); });
    }
    for (auto &th : pool) th.join();
}

// Dynamic schedule: workers grab `grain`-sized blocks of [0, n).
template <typename F>
void parallel_for_dynamic(uint64_t n, uint64_t grain, F fn, int nthreads = 0)
{
    if (nthreads <= 0) nthreads = host_threads();
    if (n <= grain || nthreads == 1) {
        fn((uint64_t)0, n);
        return;
    }
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++) {
        pool.emplace_back([&]() {
            for (;;) {
                uint64_t b = next.fetch_add(grain);
                if (b >= n) break;
                fn(b, std::min(n, b + grain));
            }
        });
    }
    for (auto &th : pool) th.join();
}

// SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014), used counter-based: draw k of the stream seeded
// with s is mix(s + (k + 1) * gamma). Any element of the stream can be
// computed independently, so parallel generation is bit-reproducible.
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t draw(uint64_t seed, uint64_t k) { return mix64(seed + (k + 1) * kGamma); }

// Uniform double in [0, 1) from the top 53 bits.
inline double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

// Uniform integer in [0, range) by multiply-shift of the top 32 bits.
inline uint32_t below(uint64_t h, uint32_t range) { return (uint32_t)(((h >> 32) * (uint64_t)range) >> 32); }

void set_error(const std::string &msg);
const char *get_error();

}  // namespace sg
