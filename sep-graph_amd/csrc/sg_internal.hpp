// Internal host/device structures shared by the library's HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/sepgraph.h"
#include "sg_device.hpp"
#include "sg_host.hpp"

#define SG_FAIL(code, msg)      \
    do {                        \
        sg::set_error(msg);     \
        return (code);          \
    } while (0)

#define SG_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess) {                                                            \
            sg::set_error(std::string(#call) + ": " + hipGetErrorString(_e));              \
            return SG_ERR_HIP;                                                             \
        }                                                                                  \
    } while (0)

namespace sg {

// Read-only view of a device-resident graph, passed by value to kernels
// (the reference's dev::CSRGraph / dev::CSCGraph, csr_graph.h:745-871).
struct DevGraph {
    uint32_t n;
    uint32_t words;           // ceil(n / 64): length of per-vertex bitmaps
    uint64_t m;
    const uint32_t *row;      // CSR row_start [n+1]
    const uint32_t *col;      // CSR edge_dst [m]
    const uint32_t *w;        // CSR weights [m] or null
    const uint32_t *cstart;   // CSC col_start [n+1] (== row when undirected)
    const uint32_t *csrc;     // CSC edge_src [m]
    const uint32_t *cw;       // CSC weights [m] or null
    const uint64_t *inert;    // bitmap: in-degree 0 (never reached unless source)
};

}  // namespace sg

struct sg_graph {
    sg::DevGraph dev{};
    bool undirected = false;
    bool weighted = false;
    uint64_t device_bytes = 0;
    void *allocs[8] = {};
    int nallocs = 0;
    int device = 0;
    int num_cus = 0;
    double avg_weight = 0.0;  // for the SSSP delta heuristic
};

namespace sg {

// Bytes reserved per engine for per-round counters of the persistent kernels.
constexpr uint64_t kCtrBytes = 65536;

// Device-side run summary written by the persistent kernels.
struct DevStats {
    uint32_t rounds, push_rounds, pull_rounds, reached;
    unsigned long long reached_edges;
    unsigned long long trace;
    uint32_t max_depth;
    int32_t status;
    uint32_t iterations;
    uint32_t pad;
    double residual;
};

}  // namespace sg

struct sg_engine {
    sg_graph *g = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    sg_engine_options opt{};
    int last_algo = -1;

    // device allocations
    void *value = nullptr;        // per-vertex result (levels / distances / ranks)
    uint64_t value_bytes = 0;
    uint64_t *bitmaps = nullptr;  // visited + 3 frontier bitmaps (4 * words)
    uint32_t *queues = nullptr;   // 2 small queues [n] each
    uint2 *large = nullptr;       // 2 large lists
    uint32_t large_cap = 0;
    void *ctr = nullptr;          // kCtrBytes of per-round counters
    uint32_t *bar_words = nullptr;
    uint32_t *abort_flag = nullptr;
    sg::DevStats *stats = nullptr;
    uint32_t bar_words_len = 0;

    int coop_blocks = 0;          // persistent grid size for the BFS kernel
    int block_size = 256;
    bool needs_bar_reset = true;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
};

namespace sg {
int engine_ensure_bfs_state(sg_engine *e);
int engine_reset_barrier(sg_engine *e);
}  // namespace sg
