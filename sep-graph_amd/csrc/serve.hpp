// Host interface of the serving kernel (serve_kernel.hip): one persistent
// launch in which each XCD is a query slot fed from a ring in pinned host
// memory.
#pragma once

#include "sg_internal.hpp"

namespace sg {

// Ring entries in flight at most (published and not yet completed).
constexpr uint32_t kServeRing = 1024;
// Graphs one serving kernel answers queries on.
constexpr int kServeGraphs = 4;
// A launch ends after this long without a query to claim; the host launches
// again when more arrive.
constexpr uint32_t kServeIdleMs = 20;

// One published query, written by the host before it advances the tail.
struct ServeEntry {
    uint32_t graph;      // index into the server's graphs
    uint32_t kind;       // SG_QUERY_*
    uint32_t source;     // device vertex id
    uint32_t source_id;  // the caller's id of the same vertex
    uint32_t row;        // result row in the pool
};

// Byte offsets of one slot's working arena.
struct ServeLayout {
    uint64_t bitmaps, queue, ctr;      // BFS (and BC forward): 2 filters + 3 frontier bitmaps, queues, counters
    uint64_t sbitmaps, squeue, sctr;   // SSSP: 3 frontier bitmaps, queues, counters
    uint64_t stats;
    uint64_t dist;                     // SSSP distances in device order (relabelled graphs)
    uint64_t bc;                       // the BC arena, laid out by bc1
    Bc1Layout bc1;
    uint64_t total;
};

ServeLayout serve_layout(uint32_t n, uint32_t qcap);

struct ServeDev;

// Device state, the pinned ring and the co-resident grid of the serving
// kernel for up to kServeGraphs graphs on one device, writing results into
// `pool` (rows `row_bytes` apart, at least 4 n of the largest graph; the
// caller's vertex order). Launches go to `stream`. SSSP on graph i uses the
// two-level priority schedule when opt asks for it or bit i of prio_graphs
// is set.
int serve_create(sg_graph *const *graphs, int ngraphs, const sg_engine_options &opt, uint32_t prio_graphs,
                 hipStream_t stream, void *pool, uint64_t row_bytes, ServeDev **out);
// Asks a running launch to finish once the ring is drained, waits, frees.
void serve_destroy(ServeDev *d);
// Publish a query (at most kServeRing may be unfinished); returns its
// sequence number.
uint32_t serve_publish(ServeDev *d, int graph, int kind, uint32_t dev_source, uint32_t source_id, uint32_t row);
// Whether query `seq` has finished, and its device status.
bool serve_poll(ServeDev *d, uint32_t seq, int32_t *status);
// Keep the kernel running while published queries are unclaimed: notes a
// launch that ended and launches again when needed. *running tells whether a
// launch is in flight afterwards. SG_ERR_TIMEOUT when the launch that ended
// had a bounded wait expire (its state is then reset; queries it had claimed
// and not finished are lost, see serve_claimed).
int serve_pump(ServeDev *d, bool *running);
// Entries claimed by the kernel so far (only meaningful when no launch runs).
uint32_t serve_claimed(ServeDev *d);
uint64_t serve_device_bytes(const ServeDev *d);

}  // namespace sg
