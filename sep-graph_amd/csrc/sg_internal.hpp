// Internal host/device structures shared by the library's HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <mutex>
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
    // First in-neighbour of every vertex (csrc[cstart[v]], 0xFFFFFFFF when it
    // has none): a pull round's first parent test streams this n-word array
    // instead of touching one csrc cache line per vertex.
    const uint32_t *chead;
    // Smallest in-edge weight of every vertex (UINT32_MAX when it has no
    // in-edges; null for an unweighted graph, whose weights are all 1): a
    // vertex whose distance is at most the frontier's smallest distance plus
    // this cannot be lowered by an SSSP pull round.
    const uint32_t *cminw;
    // In-edges in 64-vertex slices (sliced ELLPACK with the wave width as
    // slice height): slice w holds the vertices of bitmap word w, its j-th
    // row the j-th in-edge of each of them, 64 consecutive entries
    // (0xFFFFFFFF past a vertex's in-degree), so a wave reads one in-edge per
    // lane per coalesced load, with no scheduling. Vertices with in-degree
    // above kPullHubDegree are left out (the in-hub chunk list holds them).
    // sell_off[w] is slice w's first entry (words + 1 entries); sell_w holds
    // the weights (null for an unweighted graph).
    const uint32_t *sell_off;
    const uint32_t *sell_src;
    const uint32_t *sell_w;
    // Static split of high-degree vertices: every vertex with out-degree >
    // kHugeDegree contributes ceil(deg / kHugeChunk) chunk descriptors
    // (vertex, first edge), so a push round can spread its edges over the grid.
    const uint2 *huge;
    uint32_t nhuge;           // number of chunk descriptors
    const uint32_t *huge_end; // end (one past the last edge) of each chunk
    // Split of the in-edges of every vertex with in-degree > kPullHubDegree
    // over the CSC (vertex, first in-edge), for pull rounds that visit all
    // in-edges (SSSP), so no wave walks a long in-edge list alone.
    const uint2 *hugein;
    uint32_t nhugein;
    // One entry per such in-hub: (index of its first descriptor in hugein,
    // number of descriptors); its chunks are contiguous there.
    const uint2 *hubin;
    uint32_t nhubin;
    // Weighted graphs: the same in-hub in-edges by increasing weight (then
    // source), chunk c of hugein at [c * kPullChunk, (c + 1) * kPullChunk)
    // (0xFFFFFFFF sources past the hub's end); the chunks as (hub, chunk) in
    // the order an SSSP pull round visits them -- every hub's first chunk
    // (nhubin of them), then the others -- with each one's smallest weight
    // beside it, so a round skips every chunk whose lightest in-edge cannot
    // lower its hub. Null otherwise.
    const uint32_t *hub_src;
    const uint32_t *hub_w;
    const uint2 *hub_ord;
    const uint32_t *hub_minw;
};

// A vertex whose out-degree exceeds this is expanded through the static chunk
// list instead of by a single wave: a wave expands the out-edges of up to 64
// vertices together, one 64-edge step (a few dependent memory round trips)
// at a time, so the bound keeps any wave's share of a push round short.
#ifndef SG_HUGE_DEGREE
#define SG_HUGE_DEGREE 256
#endif
#ifndef SG_HUGE_CHUNK
#define SG_HUGE_CHUNK 64
#endif
constexpr uint32_t kHugeDegree = SG_HUGE_DEGREE;
constexpr uint32_t kHugeChunk = SG_HUGE_CHUNK;
// In-degree above which a pull round splits a vertex's in-edges into chunks
// of kPullChunk.
constexpr uint32_t kPullHubDegree = 64;
constexpr uint32_t kPullChunk = 256;

}  // namespace sg

struct sg_graph {
    sg::DevGraph dev{};
    bool undirected = false;
    bool weighted = false;
    uint64_t device_bytes = 0;
    void *allocs[32] = {};
    int nallocs = 0;
    int device = 0;
    int num_cus = 0;
    double avg_weight = 0.0;  // for the SSSP delta heuristic
    // Degree-ordered device ids: the caller's vertex v is device vertex
    // h_inv[v]; results are gathered back into the caller's order.
    bool relabeled = false;
    std::vector<uint32_t> h_inv;
    const uint32_t *d_inv = nullptr;
    // device vertex i is the caller's vertex d_perm[i] (for kernels that
    // write results in the caller's order as they go)
    const uint32_t *d_perm = nullptr;
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

// Rows of the per-round device trace kept per engine.
constexpr uint32_t kTraceRows = 4096;
// Upper bound on the blocks of a persistent grid (sg_engine_options.grid_blocks).
constexpr uint32_t kMaxBlocks = 8192;

struct DevTrace {
    uint32_t round;
    int32_t variant;
    uint32_t in_nodes;
    uint32_t pad;
    unsigned long long in_edges;
    unsigned long long ticks;  // wall_clock64() - kernel start
};

}  // namespace sg

struct sg_engine {
    sg_graph *g = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    sg_engine_options opt{};
    int last_algo = -1;

    // device allocations
    void *value = nullptr;        // result of the last run (levels / distances / ranks / BC)
    uint64_t value_bytes = 0;     // bytes of that result
    uint64_t value_cap = 0;       // allocated bytes
    uint64_t *bitmaps = nullptr;  // 5 * words: 2 visited filters + 3 rotating frontier bitmaps
    // BFS state left clean by the last BFS (frontier bitmaps and counters
    // zeroed, the idle visited filter reset): valid while no other kernel ran
    uint64_t launch_seq = 0;      // persistent-kernel launches on this engine
    uint64_t bfs_token = 0;       // launch_seq right after the last BFS launch
    uint32_t bfs_parity = 0;      // visited filter the next BFS uses
    uint32_t *queues = nullptr;   // 3 frontier queues of qcap entries
    uint32_t qcap = 0;
    void *ctr = nullptr;          // kCtrBytes of per-round counters
    uint32_t *abort_flag = nullptr;
    sg::DevStats *stats = nullptr;
    sg::DevTrace *trace = nullptr;    // kTraceRows rows
    void *aux = nullptr;              // per-algorithm working arrays (PageRank, BC), grown on demand
    uint64_t aux_bytes = 0;
    uint32_t *srcbuf = nullptr;       // device copy of a batched call's sources
    uint64_t srcbuf_cap = 0;          // entries
    void *work = nullptr;             // results in device vertex order (relabelled graphs)
    uint64_t work_cap = 0;
    void *garena = nullptr;           // per-query state of concurrent (one per XCD) SSSP
    uint64_t garena_bytes = 0;
    uint64_t garena_bar_off = 0;      // its barrier areas (re-zeroed after an aborted run)
    uint64_t garena_bar_bytes = 0;
    uint64_t group_launches = 0;      // launches on those barrier areas (registration parity)
    int32_t *group_status = nullptr;  // worst status over the last concurrent-SSSP call
    uint32_t trace_rows = 0;          // rows written by the last run (host copy)
    double wall_khz = 100000.0;       // device wall clock rate
    // Persistent-grid configurations (one per kernel and block size): grid
    // size from occupancy checked by a residency probe, barrier words.
    struct PGrid {
        const void *fn = nullptr;
        int bs = 0;
        int cap = 0;       // sg_engine_options.grid_blocks it was sized for
        int blocks = 0;
        uint64_t launches = 0;  // kernels run on `words` (their parity picks the registration set)
        uint32_t *words = nullptr;
        // launches the kernel in residency-probe mode with a given grid
        std::function<int(int blocks, uint32_t *words)> probe;
    };
    static constexpr int kMaxGrids = 32;
    PGrid grids[kMaxGrids];
    int ngrids = 0;
    bool needs_bar_reset = false;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
};

namespace sg {
int engine_ensure_bfs_state(sg_engine *e);
// Make the value array hold `bytes` (grown when needed, synchronising the
// stream first) and record that the next result has that size.
int engine_ensure_value(sg_engine *e, uint64_t bytes);
// Grow the engine's auxiliary arena to at least `bytes` (synchronises the
// stream before replacing a smaller one; contents are not preserved).
int engine_ensure_aux(sg_engine *e, uint64_t bytes);
int engine_ensure_srcbuf(sg_engine *e, uint64_t entries);

// Relabelled graphs keep vertices in degree order on the device.
// The device id of the caller's vertex v.
uint32_t dev_vertex(const sg_graph *g, uint32_t v);
// Where a run writes `rows` rows of n 32-bit results: the value array, or,
// for a relabelled graph, the engine's work array (device order), which
// engine_finish then gathers into the value array in the caller's order.
int engine_result_rows(sg_engine *e, uint64_t rows, void **buf);
// The work array alone, at least `bytes` (grown when needed).
int engine_ensure_work(sg_engine *e, uint64_t bytes);
int engine_finish(sg_engine *e, uint64_t rows);
// dst rows (caller's order) from src rows (device order); no-op for graphs
// kept in the caller's order. Asynchronous on the engine's stream.
int engine_gather_rows(sg_engine *e, const void *src, void *dst, uint64_t rows);

// Batched (bit-packed, up to 64 sources per launch) query kinds.
enum MsKind { kMsBfs = 0, kMsSssp = 1, kMsBcDep = 2, kMsBcSum = 3 };
// Size the engine's arena for a kind (may synchronise and reallocate).
int ms_prepare(sg_engine *e, int kind);
// One batch of nsrc <= 64 device-resident sources, asynchronously on the
// engine's stream. Results: row dev_orow[i] (or i when null) of `out`, n
// elements per row (int32 levels, uint32 distances, FP32 dependencies); for
// kMsBcSum, `out` is the n-float BC sum written by the batch with `last`.
int ms_launch(sg_engine *e, int kind, const uint32_t *dev_src, uint32_t nsrc, void *out, const uint32_t *dev_orow,
              bool first, bool last);

// Working arena of one single-source BC run (byte offsets).
struct Bc1Layout {
    uint64_t sa, lvl, order, lcnt, loff, lfill, ctl, lvol, lbm, fstats, bcsum, total;
};

// Single-source BC from device vertex `dev_source` over the whole grid
// (asynchronous): `dep` (optional) receives 0.5 delta, in device order, or at
// dep[perm[v]] when perm is given; with `bcout`, an FP64 sum kept in the
// engine's arena is started (first) or extended and the last source writes it
// to bcout as FP32 in device order. The engine's stats sum rounds and count
// sources over the launches from `first` on.
int bc1_launch(sg_engine *e, uint32_t dev_source, float *dep, const uint32_t *perm, float *bcout, bool first,
               bool last);

// Up to one SSSP query per XCD in one launch (the XCD's blocks form the
// query's grid); query i writes its distances to the device row out[i].
int sssp_group_launch(sg_engine *e, const uint32_t *sources, int nq, uint32_t *const *out);
// Worst status over the queries of the last concurrent-SSSP call
// (synchronises the stream).
int sssp_group_status(sg_engine *e);
// One SSSP query over the whole grid (device source id) computing in the
// device row `dist` (device order); a relabelled graph's result is gathered
// into the row `out` (caller's order) by the kernel. Asynchronous;
// engine_run_status synchronises and reports it.
int sssp_single_launch(sg_engine *e, uint32_t source, uint32_t *dist, uint32_t *out);
// One BFS query from the caller's vertex `source` writing int32 levels in the
// caller's vertex order to the device row `out`; asynchronous. `traced` also
// fills the engine's round trace and kernel timing (sg_bfs).
int bfs_launch(sg_engine *e, uint32_t source, int32_t *out, bool traced);
// Synchronise the engine's stream and return the status the last run wrote
// to its device stats (marking the barriers for a re-probe after a failure).
int engine_run_status(sg_engine *e);

// Device allocation owned by the engine (freed with it); reports SG_ERR_NOMEM.
int engine_alloc(void **p, uint64_t bytes);
// Grid for a persistent kernel: the largest blocks-per-CU x CUs (starting
// from the occupancy API's figure) whose blocks all turn out to be resident
// when the kernel is launched in probe mode, with the barrier's per-XCD block
// counts from that probe. Barriers are re-zeroed after an aborted run.
int engine_persistent_grid(sg_engine *e, const void *fn, int bs,
                           const std::function<int(int, uint32_t *)> &probe, sg_engine::PGrid **out);
// Point a launch's barrier at the grid's words, with the registration set of
// the grid's next launch; the caller counts the launch (pg->launches++) once
// it has been enqueued.
inline void bind_barrier(const sg_engine::PGrid *pg, GridBarrier &bar)
{
    bar.words = pg->words;
    bar.rset = (uint32_t)(pg->launches & 1u);
}
// Concurrent SSSP: size its arena and reset the call's status word
// (asynchronous), before the call's first sssp_group_launch.
int sssp_group_begin(sg_engine *e);
// Enqueue one persistent kernel (or residency probe) through `launch`, after
// the device's previous persistent kernel when that was enqueued on another
// stream. Persistent grids fill the GPU: two that overlap can each hold CUs
// the other's blocks need and wait at a grid barrier until their bounded
// waits expire, so persistent kernels of all engines run one at a time.
int engine_launch_persistent(sg_engine *e, const std::function<int()> &launch);
// Drop the engine's stream from that ordering (the stream was synchronised:
// its persistent kernels are done), before the stream is destroyed or replaced.
void engine_forget_stream(sg_engine *e);
// The lock behind that ordering, for a caller that enqueues a persistent
// kernel on a stream of its own and waits for it to finish before releasing
// the lock (the allocation-free BFS).
std::mutex &persistent_mutex();
// With persistent_mutex() held: wait on the host for the device's previous
// persistent kernel when it went to a stream other than `s`, and forget it.
int persistent_wait_other(hipStream_t s);
// engine_launch_persistent for a stream that belongs to no engine.
int engine_launch_persistent_on(hipStream_t s, const std::function<int()> &launch);
}  // namespace sg
