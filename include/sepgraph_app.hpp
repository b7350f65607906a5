// The reference's app plugin surface and a generic engine that runs any app
// written against it (compile with hipcc; needs libsepgraph.so).
//
// An app derives from sepgraph::api::AppBase<TValue, TBuffer, TWeight>
// (include/framework/variants/api.cuh:22-120) and supplies the device hooks
//   GetInitValue, GetInitBuffer, GetIdentityElement, CombineValueBuffer,
//   AccumulateBuffer (unweighted and/or weighted), IsActiveNode,
//   PostComputation, IsHighPriority
// with the reference's signatures, status codes (ACCUMULATE_SUCCESS_BREAK ...)
// and members (m_csr_graph, m_csc_graph, m_nnodes, m_nedges, m_model,
// m_msg_passing, m_scheduling, m_p_current_round). A driver then runs it with
//   sepgraph::engine::Engine<TValue, TBuffer, TWeight, MyApp, CtorArgs...>
// exactly like the built-in apps (include/sepgraph.hpp).
//
// The generic engine is one cooperative persistent kernel per run that keeps
// the reference's TRAVERSAL-scheme semantics (framework.cuh:315-796):
//   * InitGraph: value/buffer from GetInitValue/GetInitBuffer, the first
//     worklist from IsActiveNode (driver.cuh:75-91, RebuildWorklist);
//   * a push round (ASYNC_PUSH_DD, async_push_dd.cuh:24-78): for each queued
//     vertex, CombineValueBuffer, then AccumulateBuffer along its out-edges
//     until a *_BREAK status; a SUCCESS status queues the head;
//   * a pull round (ASYNC_PULL_DD, algo_variants.cuh:825-903): every vertex's
//     CombineValueBuffer writes its buffer and nominates it, then each
//     nominee accumulates the buffers of its in-neighbours that are in the
//     frontier bitmap, in CSC order, until a *_BREAK status;
//   * the direction of each round from the reference's alpha/beta rule
//     (hybrid_policy.h:237-291), or the forced variant; PostComputation by one
//     thread after each round, then *m_p_current_round += 1
//     (driver.cuh:587-591).
// Frontiers are exact: a vertex enters the next worklist once per round (the
// lane whose atomicOr sets its bit in the next frontier bitmap queues it),
// where the reference's queue may hold it several times; apps whose results
// do not depend on duplicate pushes (all of the reference's) are unaffected.
// Edges of one vertex are visited by one lane in CSR/CSC order (the
// reference's LoadBalancing::NONE), so *_BREAK is exact. Topology-driven
// variants and the ITERATIVE scheme are the built-in PageRank's; asked of a
// generic app they fail with SG_ERR_UNSUPPORTED.
#pragma once

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <vector>

#include "sepgraph.hpp"

// groute/graphs/common.h:36
typedef uint32_t index_t;

namespace utils {
// include/utils/cuda_utils.h: the (buffer, schedule) pair CombineValueBuffer returns
template <typename T1, typename T2>
struct alignas(8) pair {
    T1 first;
    T2 second;
    __host__ __device__ pair() {}
    __host__ __device__ pair(const T1 a, const T2 b) : first(a), second(b) {}
};
}  // namespace utils

namespace groute {
namespace graphs {
namespace dev {
// Device views of the resident graph with the reference's accessors
// (csr_graph.h:745-871).
struct CSRGraph {
    uint32_t nnodes = 0, nedges = 0;
    const uint32_t *row_start = nullptr, *edge_dst = nullptr;
    __device__ __forceinline__ index_t begin_edge(index_t node) const { return row_start[node]; }
    __device__ __forceinline__ index_t end_edge(index_t node) const { return row_start[node + 1]; }
    __device__ __forceinline__ index_t edge_dest(index_t edge) const { return edge_dst[edge]; }
};
struct CSCGraph {
    uint32_t nnodes = 0, nedges = 0;
    const uint32_t *col_start = nullptr, *edge_source = nullptr, *row_start = nullptr;
    __device__ __forceinline__ index_t begin_edge(index_t node) const { return col_start[node]; }
    __device__ __forceinline__ index_t end_edge(index_t node) const { return col_start[node + 1]; }
    __device__ __forceinline__ index_t edge_src(index_t edge) const { return edge_source[edge]; }
    __device__ __forceinline__ index_t out_degree(index_t node) const { return row_start[node + 1] - row_start[node]; }
};
}  // namespace dev
}  // namespace graphs
}  // namespace groute

namespace sepgraph {
namespace api {

template <typename TValue, typename TBuffer, typename TWeight>
struct AppBase {
    // [ACCUMULATE] [CONTINUE/BREAK]
    const int ACCUMULATE_SUCCESS_BREAK = 0;
    const int ACCUMULATE_SUCCESS_CONTINUE = 1;
    const int ACCUMULATE_FAILURE_BREAK = 2;
    const int ACCUMULATE_FAILURE_CONTINUE = 3;

    groute::graphs::dev::CSRGraph m_csr_graph;
    groute::graphs::dev::CSCGraph m_csc_graph;
    uint32_t m_nnodes = 0;
    uint32_t m_nedges = 0;
    common::Model m_model = common::Model::ASYNC;
    common::MsgPassing m_msg_passing = common::MsgPassing::PUSH;
    common::Scheduling m_scheduling = common::Scheduling::DATA_DRIVEN;
    uint32_t *m_p_current_round = nullptr;

    __device__ virtual TValue GetInitValue(index_t node) const = 0;
    __device__ virtual TBuffer GetInitBuffer(index_t node) const = 0;
    __host__ __device__ virtual TBuffer GetIdentityElement() const = 0;
    __device__ virtual utils::pair<TBuffer, bool> CombineValueBuffer(index_t node, TValue *p_value,
                                                                     TBuffer *p_buffer) = 0;
    // An app overrides the overload its graph uses (weighted when TWeight is
    // not graphs::NoWeight and the graph has weights); the other reports a
    // failure that stops the edge loop.
    __device__ virtual int AccumulateBuffer(index_t src, index_t dst, TBuffer *p_buffer, TBuffer buffer)
    {
        return ACCUMULATE_FAILURE_BREAK;
    }
    __device__ virtual int AccumulateBuffer(index_t src, index_t dst, TWeight weight, TBuffer *p_buffer,
                                            TBuffer buffer)
    {
        return ACCUMULATE_FAILURE_BREAK;
    }
    // Topology-driven variants only (RebuildWorklist); also seeds the first
    // worklist.
    __device__ virtual bool IsActiveNode(index_t node, TBuffer buffer) const { return false; }
    __device__ virtual void PostComputation() {}
    __device__ virtual bool IsHighPriority(TBuffer current_priority, TBuffer buffer) const { return true; }

    __host__ __device__ void SetVariant(common::Model model, common::MsgPassing msg_passing,
                                        common::Scheduling scheduling)
    {
        m_model = model;
        m_msg_passing = msg_passing;
        m_scheduling = scheduling;
    }
    __host__ __device__ void SetVariant(common::AlgoVariant v)
    {
        m_model = v.m_model;
        m_msg_passing = v.m_msg_passing;
        m_scheduling = v.m_scheduling;
    }
};

}  // namespace api

namespace detail {

// Counters of one round's frontier (three slots rotate) and of a pull
// round's nominees.
struct AppRound {
    unsigned long long count;  // vertices queued
    unsigned long long scout;  // their out-degree sum
    uint32_t qlen;             // queue entries (== count)
    uint32_t nominees;         // pull round: vertices CombineValueBuffer scheduled
    uint32_t pad[26];
};

struct AppRunStats {
    uint32_t rounds, push_rounds, pull_rounds, pad;
    unsigned long long policy_trace;  // bit r: round r pulled
};

template <typename TApp, typename TValue, typename TBuffer, typename TWeight>
struct AppArgs {
    TApp app;
    sg_graph_view g;
    TValue *value;
    TBuffer *buffer;
    uint32_t *queue;    // 3 x n
    uint64_t *bits;     // 3 x ceil(n / 64)
    uint32_t *nominee;  // n
    AppRound *ctr;      // 3
    uint32_t *round;    // the app's m_p_current_round
    AppRunStats *stats;
    double alpha, beta;
    uint32_t total_workload0;
    int32_t force;  // -1 auto, 0 push, 1 pull
    uint32_t max_rounds;
    uint32_t weighted;
};

template <typename TApp, typename TValue, typename TBuffer, typename TWeight>
__device__ __forceinline__ int app_accumulate(TApp &app, const AppArgs<TApp, TValue, TBuffer, TWeight> &a,
                                              const uint32_t *w, index_t src, index_t dst, uint32_t e, TBuffer *p,
                                              TBuffer b)
{
    if constexpr (std::is_same<TWeight, graphs::NoWeight>::value) {
        return app.AccumulateBuffer(src, dst, p, b);
    } else {
        if (a.weighted) return app.AccumulateBuffer(src, dst, (TWeight)w[e], p, b);
        return app.AccumulateBuffer(src, dst, p, b);
    }
}

// Queue v for the next round once: the lane whose atomicOr sets its bit.
template <typename TApp, typename TValue, typename TBuffer, typename TWeight>
__device__ __forceinline__ void app_emit(const AppArgs<TApp, TValue, TBuffer, TWeight> &a, uint64_t *bnext,
                                         uint32_t *qnext, AppRound *cnext, uint32_t v)
{
    const unsigned long long bit = 1ull << (v & 63);
    const unsigned long long old = atomicOr((unsigned long long *)&bnext[v >> 6], bit);
    if (old & bit) return;
    const uint32_t pos = atomicAdd(&cnext->qlen, 1u);
    qnext[pos] = v;
    atomicAdd(&cnext->count, 1ull);
    atomicAdd(&cnext->scout, (unsigned long long)(a.g.row_start[v + 1] - a.g.row_start[v]));
}

template <typename TApp, typename TValue, typename TBuffer, typename TWeight>
__global__ void k_app_run(AppArgs<TApp, TValue, TBuffer, TWeight> a)
{
    namespace cg = cooperative_groups;
    cg::grid_group grid = cg::this_grid();
    TApp app = a.app;
    const uint32_t n = a.g.n, W = (n + 63) / 64;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nt = gridDim.x * blockDim.x;
    const int ok_continue[2] = {app.ACCUMULATE_SUCCESS_CONTINUE, app.ACCUMULATE_FAILURE_CONTINUE};
    const int ok_success[2] = {app.ACCUMULATE_SUCCESS_BREAK, app.ACCUMULATE_SUCCESS_CONTINUE};

    // InitGraph
    for (uint32_t v = tid; v < n; v += nt) {
        a.value[v] = app.GetInitValue(v);
        a.buffer[v] = app.GetInitBuffer(v);
    }
    for (uint32_t w = tid; w < 3 * W; w += nt) a.bits[w] = 0;
    for (uint32_t i = tid; i < 3 * (uint32_t)(sizeof(AppRound) / 4); i += nt) ((uint32_t *)a.ctr)[i] = 0;
    if (tid == 0) *a.round = 0;
    grid.sync();
    // the first worklist: RebuildWorklist over every vertex
    for (uint32_t v = tid; v < n; v += nt)
        if (app.IsActiveNode(v, a.buffer[v])) app_emit(a, a.bits, a.queue, &a.ctr[0], v);
    grid.sync();

    uint32_t total_wl = a.total_workload0;
    bool pull_prev = false;
    uint32_t L = 0, pushes = 0, pulls = 0;
    unsigned long long ptrace = 0;
    for (;;) {
        const uint32_t cur = L % 3, nxt = (L + 1) % 3, idle = (L + 2) % 3;
        const unsigned long long count = a.ctr[cur].count, scout = a.ctr[cur].scout;
        const uint32_t qlen = a.ctr[cur].qlen;
        if (count == 0) break;
        if (a.max_rounds && L >= a.max_rounds) break;
        // GetNextPolicy (hybrid_policy.h:237-291), evaluated identically by
        // every thread
        bool pull;
        if (a.force >= 0) {
            pull = a.force == 1;
        } else if (L == 0) {
            pull = false;
        } else if (pull_prev && (double)count > (double)n / a.beta) {
            pull = true;
        } else {
            const uint32_t out_wl = pull_prev ? 1u : (uint32_t)scout;
            pull = (double)out_wl > (double)total_wl / a.alpha;
        }
        if (!pull) total_wl -= (uint32_t)scout;
        app.SetVariant(common::Model::ASYNC, pull ? common::MsgPassing::PULL : common::MsgPassing::PUSH,
                       common::Scheduling::DATA_DRIVEN);
        const uint64_t *bcur = a.bits + (size_t)cur * W;
        uint64_t *bnext = a.bits + (size_t)nxt * W;
        uint32_t *qnext = a.queue + (size_t)nxt * n;
        AppRound *cnext = &a.ctr[nxt];
        // slot L + 2 (round L - 1's input) is free: clear it for round L + 1
        for (uint32_t w = tid; w < W; w += nt) a.bits[(size_t)idle * W + w] = 0;
        if (tid < sizeof(AppRound) / 4) ((uint32_t *)&a.ctr[idle])[tid] = 0;
        if (!pull) {
            const uint32_t *qcur = a.queue + (size_t)cur * n;
            for (uint32_t i = tid; i < qlen; i += nt) {
                const index_t u = qcur[i];
                const utils::pair<TBuffer, bool> p = app.CombineValueBuffer(u, &a.value[u], &a.buffer[u]);
                if (!p.second) continue;
                for (uint32_t e = a.g.row_start[u], end = a.g.row_start[u + 1]; e < end; e++) {
                    const index_t v = a.g.edge_dst[e];
                    const int st = app_accumulate(app, a, a.g.weights, u, v, e, &a.buffer[v], p.first);
                    if (st == ok_success[0] || st == ok_success[1]) app_emit(a, bnext, qnext, cnext, v);
                    if (!(st == ok_continue[0] || st == ok_continue[1])) break;
                }
            }
        } else {
            for (uint32_t v = tid; v < n; v += nt) {
                const utils::pair<TBuffer, bool> p = app.CombineValueBuffer(v, &a.value[v], &a.buffer[v]);
                a.buffer[v] = p.first;
                if (p.second) a.nominee[atomicAdd(&a.ctr[cur].nominees, 1u)] = v;
            }
            grid.sync();
            const uint32_t nn = a.ctr[cur].nominees;
            for (uint32_t i = tid; i < nn; i += nt) {
                const index_t v = a.nominee[i];
                for (uint32_t e = a.g.col_start[v], end = a.g.col_start[v + 1]; e < end; e++) {
                    const index_t u = a.g.edge_src[e];
                    if (!((bcur[u >> 6] >> (u & 63)) & 1)) continue;
                    const int st = app_accumulate(app, a, a.g.csc_weights, u, v, e, &a.buffer[v], a.buffer[u]);
                    if (st == ok_success[0] || st == ok_success[1]) app_emit(a, bnext, qnext, cnext, v);
                    if (!(st == ok_continue[0] || st == ok_continue[1])) break;
                }
            }
        }
        grid.sync();
        // PostComputation, then the round counter (driver.cuh:587-591)
        if (tid == 0) {
            app.PostComputation();
            *a.round += 1;
        }
        if (pull) {
            pulls++;
            if (L < 64) ptrace |= 1ull << L;
        } else {
            pushes++;
        }
        pull_prev = pull;
        L++;
        grid.sync();
    }
    if (tid == 0) {
        AppRunStats s;
        s.rounds = L;
        s.push_rounds = pushes;
        s.pull_rounds = pulls;
        s.pad = 0;
        s.policy_trace = ptrace;
        *a.stats = s;
    }
}

inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw Error(SG_ERR_HIP, std::string("sepgraph: ") + what + ": " + hipGetErrorString(e));
}

template <typename TApp, typename TValue, typename TBuffer, typename TWeight>
struct AppRunner {
    // Runs `app` on the engine's graph and stream; fills value and buffer
    // (n entries each) and returns the run's statistics.
    static sg_stats Start(sg_engine *e, sg_graph *g, TApp app, bool iterative, const sg_engine_options &o,
                          std::vector<TValue> &value, std::vector<TBuffer> &buffer)
    {
        if (iterative) throw Error(SG_ERR_UNSUPPORTED, "sepgraph: the ITERATIVE scheme is not available to AppBase apps");
        int force = -1;
        if (o.alpha == 0.0 || o.force_variant == SG_ASYNC_PUSH_DD || o.force_variant == SG_SYNC_PUSH_DD) force = 0;
        else if (o.force_variant == SG_ASYNC_PULL_DD || o.force_variant == SG_SYNC_PULL_DD) force = 1;
        else if (o.force_variant != SG_VARIANT_AUTO)
            throw Error(SG_ERR_UNSUPPORTED, "sepgraph: AppBase apps run data-driven variants only");
        sg_graph_view gv;
        check(sg_graph_device_view(g, &gv));
        if (gv.relabeled) throw Error(SG_ERR_STATE, "sepgraph: AppBase apps need a graph in the caller's order");
        const uint32_t n = gv.n, W = (n + 63) / 64;

        app.m_csr_graph.nnodes = n;
        app.m_csr_graph.nedges = (uint32_t)gv.m;
        app.m_csr_graph.row_start = gv.row_start;
        app.m_csr_graph.edge_dst = gv.edge_dst;
        app.m_csc_graph.nnodes = n;
        app.m_csc_graph.nedges = (uint32_t)gv.m;
        app.m_csc_graph.col_start = gv.col_start;
        app.m_csc_graph.edge_source = gv.edge_src;
        app.m_csc_graph.row_start = gv.row_start;
        app.m_nnodes = n;
        app.m_nedges = (uint32_t)gv.m;

        AppArgs<TApp, TValue, TBuffer, TWeight> a{app, gv};
        // one allocation for the run's working state
        auto al = [](size_t x) { return (x + 255) / 256 * 256; };
        const size_t o_value = 0, o_buffer = o_value + al(sizeof(TValue) * n);
        const size_t o_queue = o_buffer + al(sizeof(TBuffer) * n);
        const size_t o_bits = o_queue + al(sizeof(uint32_t) * 3 * (size_t)n);
        const size_t o_nom = o_bits + al(sizeof(uint64_t) * 3 * (size_t)W);
        const size_t o_ctr = o_nom + al(sizeof(uint32_t) * n);
        const size_t o_round = o_ctr + al(sizeof(AppRound) * 3);
        const size_t o_stats = o_round + 256;
        const size_t total = o_stats + al(sizeof(AppRunStats));
        char *mem = nullptr;
        hip_check(hipMalloc((void **)&mem, total), "hipMalloc");
        struct Free {
            char *p;
            ~Free() { (void)hipFree(p); }
        } guard{mem};
        a.value = (TValue *)(mem + o_value);
        a.buffer = (TBuffer *)(mem + o_buffer);
        a.queue = (uint32_t *)(mem + o_queue);
        a.bits = (uint64_t *)(mem + o_bits);
        a.nominee = (uint32_t *)(mem + o_nom);
        a.ctr = (AppRound *)(mem + o_ctr);
        a.round = (uint32_t *)(mem + o_round);
        a.stats = (AppRunStats *)(mem + o_stats);
        a.app.m_p_current_round = a.round;
        a.alpha = o.alpha == 0.0 ? 1.0 : o.alpha;
        a.beta = o.beta;
        a.total_workload0 = (uint32_t)((double)gv.m * o.edge_factor);
        a.force = force;
        a.max_rounds = (uint32_t)std::max(0, o.max_iteration);
        a.weighted = gv.weights != nullptr;

        // a cooperative grid: every block resident, so grid.sync() is valid
        const int bs = 256;
        auto *fn = k_app_run<TApp, TValue, TBuffer, TWeight>;
        int per_cu = 0, dev = 0;
        hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, bs, 0), "occupancy");
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        const uint32_t need = (std::max<uint32_t>(n, 1) + bs - 1) / bs;
        const int blocks = (int)std::max<uint32_t>(1, std::min<uint32_t>(need, (uint32_t)(per_cu * prop.multiProcessorCount)));
        void *stream = nullptr;
        check(sg_engine_stream(e, &stream));
        check(sg_engine_order_launch(e));
        hipEvent_t ev0, ev1;
        hip_check(hipEventCreate(&ev0), "hipEventCreate");
        hip_check(hipEventCreate(&ev1), "hipEventCreate");
        void *params[] = {&a};
        hip_check(hipEventRecord(ev0, (hipStream_t)stream), "hipEventRecord");
        hip_check(hipLaunchCooperativeKernel((const void *)fn, dim3(blocks), dim3(bs), params, 0, (hipStream_t)stream),
                  "hipLaunchCooperativeKernel");
        hip_check(hipEventRecord(ev1, (hipStream_t)stream), "hipEventRecord");
        hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, ev0, ev1), "hipEventElapsedTime");
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
        value.resize(n);
        buffer.resize(n);
        hip_check(hipMemcpy(value.data(), a.value, sizeof(TValue) * n, hipMemcpyDeviceToHost), "hipMemcpy");
        hip_check(hipMemcpy(buffer.data(), a.buffer, sizeof(TBuffer) * n, hipMemcpyDeviceToHost), "hipMemcpy");
        AppRunStats rs;
        hip_check(hipMemcpy(&rs, a.stats, sizeof(rs), hipMemcpyDeviceToHost), "hipMemcpy");
        sg_stats st{};
        st.rounds = rs.rounds;
        st.push_rounds = rs.push_rounds;
        st.pull_rounds = rs.pull_rounds;
        st.policy_trace = rs.policy_trace;
        st.max_depth = rs.rounds ? rs.rounds - 1 : 0;
        st.kernel_ms = ms;
        st.status = SG_OK;
        return st;
    }
};

}  // namespace detail
}  // namespace sepgraph
