/*
 * sepgraph.h — C ABI of the MI355X-native hybrid graph engine (libsepgraph.so).
 *
 * The reference exposes its hot path as C++ templates: an app derives from
 * sepgraph::api::AppBase (include/framework/variants/api.cuh:22-120) and is
 * driven by sepgraph::engine::Engine (include/framework/framework.cuh:56-797)
 * configured through EngineOptions (include/framework/common.h:203-321).
 * Templates with device hooks cannot cross a C ABI, so this header exports the
 * same life cycle over opaque handles:
 *
 *   Engine ctor + LoadGraph          -> sg_graph_create + sg_engine_create
 *   EngineOptions / SetOptions        -> sg_engine_options + sg_engine_set_options
 *   InitGraph(source) + Start()       -> sg_bfs / sg_sssp / sg_pagerank / sg_bc
 *   GatherValue() / GatherBuffer()    -> sg_engine_gather_value / sg_engine_value_ptr
 *   PrintInfo() / JSON running data   -> sg_engine_get_stats
 *
 * Conventions (all entry points):
 *   - every call returns an int status (SG_OK == 0, negative on error) and
 *     never exits the process; sg_last_error() gives the thread's last message.
 *     (The reference prints and calls exit(1) on any CUDA error,
 *     include/groute/internal/cuda_utils.h:41-65.)
 *   - vertex ids and CSR/CSC offsets are uint32 (the reference's index_t);
 *     graphs are limited to n < 2^31 vertices and m < 2^32 edges.
 *   - BFS levels are int32 with -1 for unreached (the int32 view of the
 *     reference's UINT32_MAX sentinel, samples/hybrid_bfs/hybrid_bfs_common.h:38);
 *     SSSP distances are uint32 with UINT32_MAX for unreached.
 *   - runs are asynchronous on the engine's HIP stream (sg_engine_set_stream
 *     takes a hipStream_t as void*); gather/stats calls synchronise it.
 *   - a handle is not safe for concurrent use from several threads; distinct
 *     handles are independent.
 *   - the traversal kernels are persistent (every block resident, grid
 *     barriers between rounds) and each fills the GPU, so two of them must
 *     not overlap: the library makes each engine's stream wait for the
 *     device's previous persistent kernel when that went to another stream.
 *     Runs of different engines therefore execute one after another even
 *     when their streams are otherwise independent.
 */
#ifndef SEPGRAPH_H
#define SEPGRAPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
    SG_OK = 0,
    SG_ERR_INVALID = -1,     /* bad argument or malformed graph */
    SG_ERR_RANGE = -2,       /* vertex id / size out of range */
    SG_ERR_NOMEM = -3,       /* host or device allocation failed */
    SG_ERR_HIP = -4,         /* HIP runtime error */
    SG_ERR_CAPACITY = -5,    /* a fixed-capacity buffer would overflow */
    SG_ERR_TIMEOUT = -6,     /* a device-side bounded wait expired */
    SG_ERR_STATE = -7,       /* call not valid in the handle's current state */
    SG_ERR_UNSUPPORTED = -8, /* option combination not implemented */
    SG_ERR_IO = -9           /* a file could not be opened, read or written */
};

/* Thread-local message describing the last failing call on this thread. */
const char *sg_last_error(void);
/* Library version, e.g. 100 for 0.1.0. */
int sg_version(void);
/* Number of visible HIP devices (0 on a host without GPU; never an error). */
int sg_device_count(void);

/* ------------------------------------------------- host graph construction
 * Synthetic inputs and CSR/CSC building. The reference loads .gr/.mtx/METIS
 * files or generates uniform/chain/clique graphs (include/utils/graphs/traversal.h:99-200);
 * it has no RMAT or grid generator. All generators are counter-based
 * (SplitMix64 draw k of the seed's stream), so output is independent of the
 * number of host threads. */

/* R-MAT quadrant descent (Chakrabarti, Zhan & Faloutsos 2004) without noise
 * or relabelling: edge i, level l (l = 0 is the top bit) picks quadrant
 * a/b/c/d from draw(seed, i*scale + l). Writes edge_factor << scale tuples. */
int sg_gen_rmat_edges(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                      uint32_t *src, uint32_t *dst);

/* rows x cols 4-neighbour grid, vertex id r*cols + c. Each undirected edge is
 * kept unless draw(seed, edge_id) falls below p_delete (horizontal edge of
 * (r,c) has id 2*(r*cols+c), vertical edge id 2*(r*cols+c)+1). Writes each
 * kept edge once as (u, v) with u < v; *count gets the number written.
 * Capacity needed: 2*rows*cols. */
int sg_gen_grid_edges(uint32_t rows, uint32_t cols, double p_delete, uint64_t seed, uint32_t *src,
                      uint32_t *dst, uint64_t *count);

/* Build CSR (row_start[n+1], edge_dst[*m_out]) from an edge list. flags:
 * SG_CSR_SYMMETRIZE adds the reverse of every edge; SG_CSR_DROP_SELF_LOOPS;
 * SG_CSR_DEDUP removes duplicate (u,v). Neighbours end up sorted ascending.
 * edge_dst must hold `cap` entries (2*count is always enough with
 * SYMMETRIZE, count without). */
enum { SG_CSR_SYMMETRIZE = 1, SG_CSR_DROP_SELF_LOOPS = 2, SG_CSR_DEDUP = 4 };
int sg_build_csr(uint32_t n, uint64_t count, const uint32_t *src, const uint32_t *dst, int flags,
                 uint32_t *row_start, uint32_t *edge_dst, uint64_t cap, uint64_t *m_out);

/* Per-edge uint32 weights uniform in [lo, hi), weight of (u,v) from
 * draw(seed, u<<32|v), or from the unordered pair when `symmetric` (so both
 * directions of an undirected edge agree). */
int sg_gen_weights(uint32_t n, const uint32_t *row_start, const uint32_t *edge_dst, uint32_t lo,
                   uint32_t hi, uint64_t seed, int symmetric, uint32_t *weights);

/* CSC (col_start[n+1], edge_src[m], optional weights) with sources ascending in
 * each column, as csr_tocsc (include/groute/graphs/csr_graph.h:330-397). */
int sg_csc_from_csr(uint32_t n, const uint32_t *row_start, const uint32_t *edge_dst,
                    const uint32_t *weights, uint32_t *col_start, uint32_t *edge_src,
                    uint32_t *csc_weights);

/* k distinct vertices with out-degree >= 1, drawn as draw(seed, i) mod n with
 * rejection. Fails with SG_ERR_RANGE if fewer than k such vertices exist. */
int sg_pick_sources(uint32_t n, const uint32_t *row_start, int k, uint64_t seed, uint32_t *out);

/* Missing edge weights generated as the reference's drivers do with
 * --gen_weights (include/utils/graphs/traversal.h:174-186): weight i is the
 * i-th draw of std::uniform_int_distribution<int>(1, range) over a
 * default-constructed std::default_random_engine, in edge order. Compiled
 * against the same C++ standard library, the sequence is identical. */
int sg_gen_weights_std(uint64_t m, int32_t range, uint32_t *weights);

/* ------------------------------------------------------------- graph files
 * The reference's three input formats, chosen by --format
 * (src/utils/utils.cpp:39-58), read into a host CSR, plus the cleanups and
 * outputs of its gtools converter (samples/tools/gtools.cpp:56-346). The
 * result goes to sg_graph_create like any other host CSR. Readers validate
 * what the reference trusts (row ends, vertex ids), so a malformed file is an
 * error rather than an out-of-range access later. */
typedef struct sg_host_csr sg_host_csr;
enum { SG_FORMAT_GR = 0, SG_FORMAT_METIS = 1, SG_FORMAT_MARKET = 2 };

/* Galois .gr version 1 (ReadGraphGR, src/utils/parser.cpp:409-505): u64 header
 * {1, edge data size, n, m}, u64 row ends, u32 heads, 4 pad bytes when m is
 * odd and edge data follows, u32 edge data. Edge data of a size other than 0
 * or 4 bytes is SG_ERR_UNSUPPORTED (the reference reads 4-byte words
 * regardless).
 * METIS text (ReadGraph, parser.cpp:230-395): '%' comment lines, header
 * "n m [fmt [ncon]]" with m undirected edges (every edge listed from both
 * ends, so 2m entries), fmt digits vertex-size / vertex-weight / edge-weight,
 * 1-based neighbour ids; vertex sizes and weights are parsed and dropped.
 * MatrixMarket coordinate (ReadGraphMarket, parser.cpp:508-665): '%' lines,
 * header "rows cols nnz" (n = cols), nnz 1-based "row col [weight]" lines;
 * integer weights when the first entry has a third field; each row's entries
 * deduplicated and sorted as a std::set (of heads, or of (head, weight)
 * pairs when weighted); symmetric files are not expanded. */
int sg_file_read(const char *path, int format, sg_host_csr **out);
/* Wrap a caller CSR (copied) so it can be cleaned or written. */
int sg_file_from_csr(uint32_t n, const uint32_t *row_start, const uint32_t *edge_dst,
                     const uint32_t *weights, sg_host_csr **out);
int sg_file_free(sg_host_csr *g);
/* n; m = row_start[n]; whether edge weights are present; and the edge count
 * the reference's reader records (header_edges): 2 x the METIS header's m,
 * the MatrixMarket header's nnz (not lowered by the deduplication,
 * parser.cpp:541-665), the .gr header's m. */
int sg_file_info(const sg_host_csr *g, uint32_t *n, uint64_t *m, int *weighted, uint64_t *header_edges);
/* Copy out row_start[n+1], edge_dst[m] and, when present and weights != NULL,
 * weights[m]. */
int sg_file_copy(const sg_host_csr *g, uint32_t *row_start, uint32_t *edge_dst, uint32_t *weights);

/* gtools cleanups, applied in its order: self-loops dropped (--deself_cycle),
 * duplicates dropped keeping each row's first occurrence and its weight, row
 * order otherwise unchanged (--deduplicate), then directed to undirected
 * (--d2ud): both directions of every edge, each row sorted by head; a pair's
 * weight, both ways, is that of the first edge between the two vertices in
 * row-major order. */
enum { SG_CLEAN_SELF_LOOPS = 1, SG_CLEAN_DEDUP = 2, SG_CLEAN_TO_UNDIRECTED = 4 };
int sg_file_clean(sg_host_csr *g, int flags);

/* Write in one of the three formats, so that sg_file_read gives back the
 * same CSR (MatrixMarket: when rows are sorted and duplicate-free; METIS: m
 * must be even, as every edge is listed from both ends). */
int sg_file_write(const sg_host_csr *g, const char *path, int format);
/* gtools --output: one "src<sep>dst[<sep>weight]\n" line per edge, 0-based.
 * gtools --out_degree: one "v<sep>out-degree\n" line per vertex. */
int sg_file_write_edgelist(const sg_host_csr *g, const char *path, const char *separator);
int sg_file_write_out_degree(const sg_host_csr *g, const char *path, const char *separator);

/* ------------------------------------------------------------ device graph */
typedef struct sg_graph sg_graph;

/* Upload a graph to device 0 (the reference's LoadGraph: CSR and CSC mirrors,
 * include/framework/framework.cuh:99-133). weights may be NULL. When
 * `undirected` is set the CSC aliases the CSR (FLAGS_undirected,
 * framework.cuh:111-118); otherwise col_start/edge_src/csc_weights give the
 * CSC, or are all NULL to have the library build it. */
int sg_graph_create(uint32_t n, uint64_t m, const uint32_t *row_start, const uint32_t *edge_dst,
                    const uint32_t *weights, const uint32_t *col_start, const uint32_t *edge_src,
                    const uint32_t *csc_weights, int undirected, sg_graph **out);

/* sg_graph_create with flags. By default the device copy orders vertices by
 * total degree (hot vertices first, so per-vertex state shares cache lines)
 * when the degree distribution is skewed; every entry point still takes and
 * returns the caller's vertex ids, and results are identical.
 * SG_GRAPH_RELABEL forces that order, SG_GRAPH_NO_RELABEL keeps the caller's.
 * With a relabelled graph a caller-given CSC is rebuilt from the CSR. */
enum { SG_GRAPH_UNDIRECTED = 1, SG_GRAPH_RELABEL = 2, SG_GRAPH_NO_RELABEL = 4 };
int sg_graph_create_ex(uint32_t n, uint64_t m, const uint32_t *row_start, const uint32_t *edge_dst,
                       const uint32_t *weights, const uint32_t *col_start, const uint32_t *edge_src,
                       const uint32_t *csc_weights, uint32_t flags, sg_graph **out);
int sg_graph_destroy(sg_graph *g);

/* Device arrays of a graph, for kernels compiled outside the library (the
 * generic engine of sepgraph_app.hpp; the reference passes its
 * groute::graphs::dev::CSRGraph / CSCGraph to app kernels by value,
 * csr_graph.h:745-871). Pointers stay valid until sg_graph_destroy; the
 * arrays must not be written. Vertex ids are device ids, which equal the
 * caller's ids unless `relabeled` is set (create with SG_GRAPH_NO_RELABEL to
 * keep them). CSC arrays alias the CSR ones for an undirected graph; weight
 * arrays are NULL for an unweighted graph. */
typedef struct sg_graph_view {
    uint32_t n;
    uint32_t relabeled;
    uint64_t m;
    const uint32_t *row_start;     /* [n + 1] */
    const uint32_t *edge_dst;      /* [m] */
    const uint32_t *weights;       /* [m] or NULL */
    const uint32_t *col_start;     /* [n + 1] */
    const uint32_t *edge_src;      /* [m] */
    const uint32_t *csc_weights;   /* [m] or NULL */
} sg_graph_view;
int sg_graph_device_view(const sg_graph *g, sg_graph_view *out);
/* n, m, and bytes of device memory held by the graph. */
int sg_graph_info(const sg_graph *g, uint32_t *n, uint64_t *m, uint64_t *device_bytes);

/* ------------------------------------------------------------------ engine */
typedef struct sg_engine sg_engine;

/* AlgoVariant (include/framework/common.h:43-177): model x passing x scheduling. */
enum {
    SG_VARIANT_AUTO = -1,
    SG_ASYNC_PUSH_TD = 0,
    SG_ASYNC_PUSH_DD = 1,
    SG_ASYNC_PULL_TD = 2,
    SG_ASYNC_PULL_DD = 3,
    SG_SYNC_PUSH_TD = 4,
    SG_SYNC_PUSH_DD = 5,
    SG_SYNC_PULL_TD = 6,
    SG_SYNC_PULL_DD = 7
};
/* LoadBalancing (common.h): NONE / COARSE_GRAINED / FINE_GRAINED / HYBRID. */
enum { SG_LB_DEFAULT = -1, SG_LB_NONE = 0, SG_LB_COARSE = 1, SG_LB_FINE = 2, SG_LB_HYBRID = 3 };
/* Priority (common.h): NONE / LOW_HIGH (two-level) / SAMPLING. */
enum { SG_PRIO_NONE = 0, SG_PRIO_LOW_HIGH = 1, SG_PRIO_SAMPLING = 2 };

typedef struct sg_engine_options {
    int32_t force_variant;      /* EngineOptions::ForceVariant, or SG_VARIANT_AUTO */
    int32_t fused;              /* EngineOptions::SetFused. Every run here is one persistent
                                   multi-round kernel (the reference's fused form), so the
                                   field only records the request */
    int32_t priority;           /* SG_PRIO_* */
    float priority_threshold;   /* SetTwoLevelBasedPriority delta (SSSP); 0 = heuristic */
    float cut_threshold;        /* SetSampleBasedPriority */
    int32_t lb_push, lb_pull;   /* SetLoadBalancing, SG_LB_*, for push / pull rounds.
                                   SG_LB_NONE: one lane per vertex walks its edges in
                                   order, no hub splitting (sg_bfs, sg_sssp only; other
                                   entry points return SG_ERR_UNSUPPORTED). DEFAULT,
                                   COARSE, FINE and HYBRID select the balanced schedulers:
                                   hub edge lists split into chunks over the grid, the
                                   rest expanded by waves with prefix-sum scheduling */
    double alpha;               /* FLAGS_alpha, default 15 (include/utils/app_skeleton.h:50) */
    double beta;                /* FLAGS_beta, default 18 (app_skeleton.h:51) */
    double edge_factor;         /* FLAGS_edge_factor, default 1 (app_skeleton.h:49) */
    int32_t max_iteration;      /* FLAGS_max_iteration; 0 = no cap (the reference's 200
                                   cuts high-diameter traversals short, framework.cuh:359-364) */
    int32_t block_size;         /* threads per workgroup for traversal kernels; 0 = tuned */
    int32_t large_degree;       /* out-degree above which a push vertex is split across
                                   the grid: 0 or 256 (the split built with the graph);
                                   other values are SG_ERR_UNSUPPORTED */
    int32_t grid_blocks;        /* cap on the persistent grid's workgroups (0: fill the
                                   GPU); small grids make rounds cheaper on high-diameter
                                   graphs whose frontiers stay small */
    int32_t stage_hubs;         /* sg_bfs push rounds: 1 stages the neighbour lists of
                                   high-degree frontier vertices in LDS per workgroup (the
                                   reference's CTA-level expansion of large vertices,
                                   cta_scheduler_hybrid.cuh:86-142); 0 (default) expands
                                   their 64-edge chunks wave by wave. Results are the same */
    int32_t reserved[6];
} sg_engine_options;

void sg_engine_options_default(sg_engine_options *opt);

typedef struct sg_stats {
    uint32_t rounds;            /* traversal rounds (levels) executed */
    uint32_t push_rounds, pull_rounds;
    uint32_t reached;           /* n_r: vertices with a finite result */
    uint64_t reached_edges;     /* m_r: sum of out-degree over reached vertices */
    uint64_t policy_trace;      /* bit r = 1 if round r ran pull (first 64 rounds) */
    double kernel_ms;           /* device time of the last run (HIP events) */
    double residual;            /* PageRank: final L1 change (FP64 on device) */
    uint32_t iterations;        /* PageRank iterations */
    uint32_t max_depth;         /* deepest level reached */
    int32_t status;             /* device-side status of the last run (SG_OK or error) */
    uint32_t reserved[7];
} sg_stats;

/* One row of the per-round trace (the reference's --trace line:
 * round, policy, in-nodes, in-edges, time; framework.cuh:334-348). */
typedef struct sg_round_trace {
    uint32_t round;
    int32_t variant;      /* SG_ASYNC_PUSH_DD / SG_ASYNC_PULL_DD ...; -1 on the final row */
    uint32_t in_nodes;    /* frontier size at the start of the round */
    uint32_t pad;
    uint64_t in_edges;    /* sum of out-degree over that frontier */
    double time_us;       /* device wall clock from kernel start to the start of the round */
} sg_round_trace;

int sg_engine_create(sg_graph *g, sg_engine **out);
int sg_engine_destroy(sg_engine *e);
int sg_engine_set_stream(sg_engine *e, void *hip_stream);
int sg_engine_set_options(sg_engine *e, const sg_engine_options *opt);
/* The HIP stream the engine's runs are enqueued on. */
int sg_engine_stream(sg_engine *e, void **hip_stream);
/* Order a kernel that the caller is about to enqueue on the engine's stream
 * and that needs the whole GPU (a persistent or cooperative grid) after every
 * persistent kernel the library enqueued before, on any engine, and before
 * the ones it enqueues later. No counterpart in the reference (which has one
 * stream per engine and no persistent kernels outside its own). */
int sg_engine_order_launch(sg_engine *e);

/* BFS from `source` (hybrid_bfs, samples/hybrid_bfs/hybrid_bfs.cu:173-243):
 * int32 levels, direction chosen per round by the reference's alpha/beta rule
 * (include/framework/hybrid_policy.h:237-291). Asynchronous. */
int sg_bfs(sg_engine *e, uint32_t source);

/* sg_bfs for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. It takes no engine and allocates nothing: the
 * caller passes the graph, the source and `levels`, a device buffer of n
 * int32 that receives the levels in the caller's vertex order (-1 when
 * unreached), and the HIP stream (as void*; NULL for the default stream) the
 * kernel runs on. opt may be NULL (the defaults: alpha 15, beta 18, auto
 * direction); block_size must be 0 or 1024 and grid_blocks 0. The kernel's
 * working state lives in a fixed device array of about 17 MiB that the
 * library holds on each GPU it runs on until the process ends, so graphs are
 * limited to n <= 2^23 vertices (SG_ERR_CAPACITY beyond). The first call on a
 * device sizes the persistent grid (a residency probe). Synchronous: returns
 * once the levels are written, with the device-side status. Calls from one
 * process are serialised per device, since they share that array; each call
 * resets the state it reads. Results equal sg_bfs's. */
int sg_bfs_noalloc(const sg_graph *g, uint32_t source, int32_t *levels, void *hip_stream,
                   const sg_engine_options *opt);

/* Single-source shortest paths from `source` (hybrid_sssp,
 * samples/hybrid_sssp/hybrid_sssp.cu:151-242): uint32 distances, UINT32_MAX
 * for unreached, relaxation d(u)+w(u,v) with atomicMin as the reference's
 * AccumulateBuffer (hybrid_sssp.cu:98-118). A graph without weights uses
 * weight 1 on every edge (the reference's all-ones fallback,
 * include/groute/graphs/csr_graph.h:1840-1846). Direction per round follows
 * the same alpha/beta rule as BFS. With options.priority == SG_PRIO_LOW_HIGH
 * the two-level (near/far) schedule of SetTwoLevelBasedPriority is used
 * (push rounds only): each round expands the frontier vertices whose distance
 * lies in the lowest delta-wide bucket present and carries the others over,
 * with delta = options.priority_threshold, or the reference's heuristic
 * c * mean weight / mean degree (hybrid_sssp.cu:183-199, c = 64 lanes here
 * for its 32-lane warp) when that is below 1. Results are identical either way.
 * Asynchronous. */
int sg_sssp(sg_engine *e, uint32_t source);

/* PageRank schedules. SG_PR_PULL_TD is the reference's SYNC_PULL_TD
 * (topology-driven pull, RunSyncPullTD, include/framework/algo_variants.cuh:753-823);
 * SG_PR_PUSH_DD is its ASYNC_PUSH_DD residual push (hybrid_pr.cu:66-110),
 * each round pushing or pulling by the alpha rule on the active out-degree.
 * SG_PR_ITERATIVE is the reference's ITERATIVE_SCHEME on the same residual
 * system (PolicyDecisionMaker, include/framework/hybrid_policy.h:130-264):
 * rounds 0, 1, 2 run SYNC_PULL_TD, ASYNC_PUSH_TD, ASYNC_PUSH_DD (the order of
 * its variant set); every later round runs the variant with the least
 * predicted time -- a topology-driven variant's last measured round time, or
 * for ASYNC_PUSH_DD the time per input edge of its last round times the
 * out-degree sum of the current worklist (ties to the earlier variant).
 * Round times are taken on the device between grid barriers. A
 * topology-driven round finds its active vertices by scanning every residual
 * (RebuildArrayWorklist), a data-driven one from the queue the previous round
 * filled. options.force_variant SYNC_PULL_TD / ASYNC_PUSH_TD / ASYNC_PUSH_DD
 * runs that variant every round (ForceVariant); other forced variants are
 * SG_ERR_UNSUPPORTED. With options.priority SG_PRIO_SAMPLING
 * (SetSampleBasedPriority, 0 < cut_threshold < 1) an ASYNC_PUSH_TD round
 * pushes only residuals at or above the value at index 1000 * (1 - cut) of
 * 1000 sampled residuals in ascending order (GetPriorityThreshold,
 * include/framework/algo_variants.cuh:98-131). The trace rows give each round's variant, active
 * vertices and out-degree, and the barrier clock at the round's end. */
enum { SG_PR_AUTO = 0, SG_PR_PULL_TD = 1, SG_PR_PUSH_DD = 2, SG_PR_ITERATIVE = 3 };

/* PageRank with damping `damping`, uniform teleport and uniform redistribution
 * of dangling mass, ranks summing to 1 (FP32 per vertex, the engine's value
 * array). SG_PR_PULL_TD iterates x' = (1-d)/n + d (P^T x + dangling(x)/n)
 * from x = 1/n and stops after the first iteration whose L1 change (reduced in
 * FP64 on the device) is below tol, or after max_iter iterations.
 * SG_PR_PUSH_DD and SG_PR_ITERATIVE push residuals until every residual is
 * below tol * (1 - d) / n and normalise. Per-vertex sums are accumulated in FP64.
 * sg_stats.iterations / .residual report the run. Asynchronous.
 * The reference's PR app solves the unnormalised system without dangling
 * redistribution (samples/hybrid_pr/hybrid_pr.cu:26-125); for uniform
 * teleport its solution normalised to sum 1 is this ranking vector. */
int sg_pagerank(sg_engine *e, double damping, double tol, int32_t max_iter, int32_t mode);

/* Betweenness centrality from k sources (host array), `batch` sources per
 * launch (1..64; 0 means 64): value[v] = 0.5 * sum over sources s of the
 * Brandes dependency delta_s(v), FP32 (accumulated in FP64; path counts
 * sigma in FP64). With k = 1 this is the reference's single-source BC
 * (samples/hybrid_bc/hybrid_bc.cu:303-376, BetweennessCentralityHost,
 * hybrid_bc_host.cpp:16-92). Graphs may be directed (forward over
 * out-edges, dependencies over successors). With batch 1 each level's
 * sigma and dependency pass pulls or pushes by the level's edge volumes; a
 * forced push or pull variant (options.force_variant) fixes that direction
 * as it fixes the forward traversal's. Asynchronous; sg_stats.rounds
 * counts forward levels over all batches, .max_depth the deepest level,
 * .iterations the batches. */
int sg_bc(sg_engine *e, const uint32_t *sources, int32_t k, int32_t batch);

/* Single-source BC dependencies for k sources, `batch` per launch: the value
 * array becomes k rows of n FP32, row i = 0.5 * delta_{sources[i]}(v), i.e.
 * the reference's single-source BC result for that source
 * (samples/hybrid_bc/hybrid_bc.cu:303-376). Serves BC-dependency queries. */
int sg_bc_dependency(sg_engine *e, const uint32_t *sources, int32_t k, int32_t batch);

/* BFS from k sources, `batch` (1..64, 0 = 64) per launch sharing one sweep
 * of the adjacency per level (bit-packed frontiers, bit s = source s). The
 * value array becomes k rows of n int32 levels, row i equal to sg_bfs from
 * sources[i] (-1 when unreached). No reference counterpart: the reference
 * runs one source per Engine::Start (framework.cuh:315-375); this serves
 * batched BFS queries. */
int sg_bfs_multi(sg_engine *e, const uint32_t *sources, int32_t k, int32_t batch);

/* SSSP from k sources, `batch` (1..64, 0 = 64) per launch: distance rows
 * [vertex][64] relaxed for every query of the batch at once (data-driven
 * push rounds). The value array becomes k rows of n uint32 distances, row i
 * equal to sg_sssp from sources[i] (UINT32_MAX when unreached). No reference
 * counterpart (one source per run there); serves batched SSSP queries. */
int sg_sssp_multi(sg_engine *e, const uint32_t *sources, int32_t k, int32_t batch);

/* SSSP from k sources, up to one concurrent query per XCD of the GPU (8 on
 * MI355X): each query runs the single-source kernel on the blocks one XCD
 * holds, with a barrier local to that XCD, so its rounds stay within one L2.
 * The value array becomes k rows of n uint32 distances, row i equal to
 * sg_sssp from sources[i]. The engine's options apply (with
 * SG_PRIO_LOW_HIGH, every query uses the two-level schedule). No reference
 * counterpart; serves batched SSSP queries with per-query latency close to
 * a lone query's. */
int sg_sssp_concurrent(sg_engine *e, const uint32_t *sources, int32_t k);

/* ---------------------------------------------------------------- serving
 * No reference counterpart: the reference answers one query per process
 * (samples/hybrid_bfs/main.cpp and its siblings). A server keeps a graph
 * resident and answers BFS / SSSP / BC-dependency queries with continuous
 * batching, in one of two forms:
 *   SG_SERVE_SLOTS (default): one persistent serving kernel in which every
 *   XCD of the GPU is a query slot. Queries are published, oldest first, into
 *   a ring in pinned host memory as soon as a result row is free; a slot that
 *   finishes its query claims the next entry, so arrivals join while other
 *   queries run and up to one query per XCD runs at a time, each on the
 *   single-source kernels (BFS, SSSP, BC dependency) over its XCD's blocks.
 *   The kernel ends after a stretch without work and is launched again when
 *   queries arrive.
 *   SG_SERVE_BATCHES: a worker forms each batch (up to 64 sources of one
 *   kind, the kind whose oldest query has waited longest) when the GPU frees
 *   up, from whatever is queued. BFS and BC-dependency batches run as one
 *   bit-packed multi-source launch, as do SSSP batches of 8 or more queries;
 *   a smaller SSSP batch runs its queries back to back, each over the whole
 *   GPU.
 * Admission is by device memory: each query holds one result row of n
 * 32-bit values (int32 levels, uint32 distances, FP32 0.5*delta) from the
 * start of its batch until released; rows = memory_budget / (4 n). Results
 * equal those of sg_bfs / sg_sssp / sg_bc_dependency for the same source.
 * Thread-safe: any thread may submit, wait, read and release. */
enum { SG_QUERY_BFS = 0, SG_QUERY_SSSP = 1, SG_QUERY_BC_DEP = 2, SG_QUERY_KINDS = 3 };
enum { SG_SERVE_SLOTS = 0, SG_SERVE_BATCHES = 1 };

typedef struct sg_server sg_server;

typedef struct sg_server_config {
    uint64_t memory_budget;  /* device bytes for result rows */
    int32_t max_rows;        /* cap on result rows (0: the budget alone) */
    int32_t max_batch[3];    /* per kind: queries per batch, 1..64 (0: 64) */
    /* smallest BFS / SSSP batch run as one bit-packed launch; smaller ones run
     * whole-GPU single queries back to back (0: 16 for BFS, 8 for SSSP);
     * SG_SERVE_BATCHES only */
    int32_t bfs_packed_min, sssp_packed_min;
    int32_t mode;            /* SG_SERVE_SLOTS (default) or SG_SERVE_BATCHES */
    /* bit i: SSSP queries on graph i use the two-level priority schedule
     * (bucket width options.priority_threshold, or the heuristic when 0), as
     * options.priority = SG_PRIO_LOW_HIGH does for every graph; the
     * reference's hybrid_sssp sets it per dataset with --sparse
     * (samples/hybrid_sssp/hybrid_sssp.cu:201-206). SG_SERVE_SLOTS only. */
    uint32_t sssp_prio_graphs;
    int32_t reserved[4];
} sg_server_config;

typedef struct sg_server_stats {
    uint64_t submitted, completed, failed, batches, queued;
    uint64_t device_bytes;   /* graph + working arenas + result rows */
    double mean_batch;       /* queries per launch */
    double busy_ms;          /* wall time with a batch on the GPU */
    uint32_t rows, free_rows;
    uint32_t reserved[8];
} sg_server_stats;

void sg_server_config_default(sg_server_config *c);
/* opt may be NULL (defaults); the server owns an engine on its own stream. */
int sg_server_create(sg_graph *g, const sg_server_config *cfg, const sg_engine_options *opt, sg_server **out);
/* A server over 1..4 resident graphs on one device (SG_SERVE_SLOTS only):
 * queries name their graph by its index in `graphs`; result rows are sized
 * for the largest graph (memory_budget counts rows of 4 n_max bytes). The
 * graphs must outlive the server. */
int sg_server_create_multi(sg_graph *const *graphs, int32_t ngraphs, const sg_server_config *cfg,
                           const sg_engine_options *opt, sg_server **out);
int sg_server_destroy(sg_server *s);
int sg_server_submit(sg_server *s, int32_t kind, uint32_t source, uint64_t *ticket);
/* sg_server_submit on graph `graph` of a multi-graph server (0 is the first). */
int sg_server_submit_to(sg_server *s, int32_t graph, int32_t kind, uint32_t source, uint64_t *ticket);
/* timeout_ms < 0 waits without limit; SG_ERR_TIMEOUT if not finished in time */
int sg_server_wait(sg_server *s, uint64_t ticket, int32_t timeout_ms);
/* device pointer to the finished query's row (valid until released), its
 * size (4 n of the query's graph), time queued before its batch (or slot)
 * started and arrival-to-done time */
int sg_server_result(sg_server *s, uint64_t ticket, void **dev_ptr, uint64_t *bytes, double *queue_ms,
                     double *total_ms);
int sg_server_copy_result(sg_server *s, uint64_t ticket, void *host_out, uint64_t bytes);
int sg_server_release(sg_server *s, uint64_t ticket);
int sg_server_get_stats(sg_server *s, sg_server_stats *out);
/* Open-loop load: query i (kind, source) is submitted arrival_s[i] seconds
 * after the call starts (non-decreasing), results are released on
 * completion, and latency_ms[i] = completion - scheduled arrival. Returns
 * when all have finished. */
int sg_server_replay(sg_server *s, const int32_t *kinds, const uint32_t *sources, const double *arrival_s,
                     int64_t count, double *latency_ms);
/* sg_server_replay with query i on graph graphs[i] (NULL: all on graph 0). */
int sg_server_replay_multi(sg_server *s, const int32_t *graphs, const int32_t *kinds, const uint32_t *sources,
                           const double *arrival_s, int64_t count, double *latency_ms);

/* ---------------------------------------------------------------- autotuner
 * No reference counterpart (the reference fixes alpha = 15, beta = 18 and
 * --block_size for every graph, app_skeleton.h:50-51). Times the real kernels
 * on k sample sources of the engine's graph and picks: alpha/beta and block
 * size for sg_bfs, alpha and block size for sg_sssp, and per query kind the
 * sources per launch with the highest queries/s whose launch time stays
 * within latency_budget_ms (for SSSP, between whole-GPU queries run back to
 * back, 1, 2 or 4 per batch, and bit-packed launches of 8 to 64); feed
 * out->max_batch to sg_server_config. Leaves
 * the engine on the tuned BFS options. Results never depend on these knobs. */
typedef struct sg_tune_config {
    double latency_budget_ms;  /* per batched launch (default 25) */
    int32_t tune_bfs, tune_sssp, tune_batches;
    int32_t reserved[8];
} sg_tune_config;

typedef struct sg_tune_result {
    double bfs_alpha, bfs_beta;
    int32_t bfs_block_size;
    int32_t sssp_block_size;
    double sssp_alpha;
    int32_t bfs_grid_blocks, sssp_grid_blocks;  /* 0: the full GPU */
    int32_t sssp_priority;                      /* SG_PRIO_NONE or SG_PRIO_LOW_HIGH */
    int32_t pad0;
    double bfs_ms, bfs_default_ms;    /* mean per query: tuned / alpha 15, beta 18, 1024 threads */
    double sssp_ms, sssp_default_ms;
    int32_t max_batch[3];             /* per SG_QUERY_* kind */
    int32_t pad;
    double batch_ms[3], batch_qps[3];
    int32_t reserved[8];
} sg_tune_result;

void sg_tune_config_default(sg_tune_config *c);
int sg_autotune(sg_engine *e, const uint32_t *sources, int32_t k, const sg_tune_config *cfg, sg_tune_result *out);

/* Synchronise the engine's stream and report the device-side status. */
int sg_engine_sync(sg_engine *e);
/* Device pointer and size in bytes of the last run's per-vertex result. */
int sg_engine_value_ptr(sg_engine *e, void **dev_ptr, uint64_t *bytes);
/* Copy the last run's per-vertex result to host memory (synchronises). */
int sg_engine_gather_value(sg_engine *e, void *host_out, uint64_t bytes);
/* Statistics of the last run (synchronises). */
int sg_engine_get_stats(sg_engine *e, sg_stats *out);
/* Per-round trace of the last traversal (synchronises): up to `max` rows,
 * the last one (variant -1) marking the end; *n_out gets the rows written. */
int sg_engine_get_trace(sg_engine *e, sg_round_trace *out, int max, int *n_out);

#ifdef __cplusplus
}
#endif
#endif /* SEPGRAPH_H */
