// Header-only C++ facade with the reference's driver-side names over the C ABI
// of libsepgraph.so, so a hybrid_* sample driver keeps its structure:
//
//   typedef sepgraph::engine::Engine<level_t, level_t, sepgraph::graphs::NoWeight,
//                                    sepgraph::apps::BFS, index_t, bool> HybridEngine;
//   HybridEngine engine(sepgraph::policy::AlgoType::TRAVERSAL_SCHEME);
//   engine.LoadGraph(csr);                   // reference: LoadGraph() from --graphfile
//   sepgraph::common::EngineOptions opt;
//   opt.ForceVariant(sepgraph::common::AlgoVariant::ASYNC_PUSH_DD);
//   engine.SetOptions(opt);
//   engine.InitGraph(source, non_atomic);
//   engine.Start();
//   const std::vector<level_t> &levels = engine.GatherValue();
//
// Reference interfaces mirrored (names and argument meaning):
//   sepgraph::common enums and AlgoVariant      include/framework/common.h:13-177
//   sepgraph::common::EngineOptions             include/framework/common.h:203-321
//   sepgraph::policy::AlgoType                  include/framework/hybrid_policy.h
//   sepgraph::engine::Engine<TValue, TBuffer, TWeight, TApp, InitArgs...>
//       SetOptions / LoadGraph / InitGraph / Start / PrintInfo / GatherValue /
//       CSRGraph                                include/framework/framework.cuh:56-397
// The built-in apps (sepgraph::apps) select the library's kernels, whose
// per-edge hooks are those of samples/hybrid_{bfs,sssp,pr,bc}/hybrid_*.cu.
// An app written against sepgraph::api::AppBase (the reference's hooks,
// api.cuh:22-120) runs on a generic engine: include sepgraph_app.hpp and
// compile the driver with hipcc. Errors throw sepgraph::Error carrying the
// SG_ERR_* code instead of exiting the process.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "sepgraph.h"

namespace sepgraph {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc)
{
    if (rc != SG_OK) throw Error(rc, std::string("sepgraph: ") + sg_last_error());
}

namespace common {

enum class LoadBalancing { NONE, COARSE_GRAINED, FINE_GRAINED, HYBRID };
enum class Priority { NONE, LOW_HIGH, SAMPLING };
enum class Model { SYNC, ASYNC };
enum class MsgPassing { PUSH, PULL };
enum class Scheduling { TOPOLOGY_DRIVEN, DATA_DRIVEN };

class AlgoVariant {
public:
    Model m_model = Model::ASYNC;
    MsgPassing m_msg_passing = MsgPassing::PUSH;
    Scheduling m_scheduling = Scheduling::DATA_DRIVEN;

    static const AlgoVariant ASYNC_PUSH_TD, ASYNC_PUSH_DD, ASYNC_PULL_TD, ASYNC_PULL_DD;
    static const AlgoVariant SYNC_PUSH_TD, SYNC_PUSH_DD, SYNC_PULL_TD, SYNC_PULL_DD;

    AlgoVariant() = default;
    AlgoVariant(Model model, MsgPassing msg_passing, Scheduling scheduling)
        : m_model(model), m_msg_passing(msg_passing), m_scheduling(scheduling)
    {
    }
    bool operator==(const AlgoVariant &o) const
    {
        return m_model == o.m_model && m_msg_passing == o.m_msg_passing && m_scheduling == o.m_scheduling;
    }
    MsgPassing GetMsgPassing() const { return m_msg_passing; }
    Scheduling GetScheduling() const { return m_scheduling; }
    // SG_ASYNC_PUSH_TD .. SG_SYNC_PULL_DD
    int Code() const
    {
        return (m_model == Model::SYNC ? 4 : 0) + (m_msg_passing == MsgPassing::PULL ? 2 : 0) +
               (m_scheduling == Scheduling::DATA_DRIVEN ? 1 : 0);
    }
    std::string ToString() const
    {
        std::string s = m_model == Model::SYNC ? "SYNC_" : "ASYNC_";
        s += m_msg_passing == MsgPassing::PUSH ? "PUSH_" : "PULL_";
        s += m_scheduling == Scheduling::TOPOLOGY_DRIVEN ? "TD" : "DD";
        return s;
    }
};

inline const AlgoVariant AlgoVariant::ASYNC_PUSH_TD{Model::ASYNC, MsgPassing::PUSH, Scheduling::TOPOLOGY_DRIVEN};
inline const AlgoVariant AlgoVariant::ASYNC_PUSH_DD{Model::ASYNC, MsgPassing::PUSH, Scheduling::DATA_DRIVEN};
inline const AlgoVariant AlgoVariant::ASYNC_PULL_TD{Model::ASYNC, MsgPassing::PULL, Scheduling::TOPOLOGY_DRIVEN};
inline const AlgoVariant AlgoVariant::ASYNC_PULL_DD{Model::ASYNC, MsgPassing::PULL, Scheduling::DATA_DRIVEN};
inline const AlgoVariant AlgoVariant::SYNC_PUSH_TD{Model::SYNC, MsgPassing::PUSH, Scheduling::TOPOLOGY_DRIVEN};
inline const AlgoVariant AlgoVariant::SYNC_PUSH_DD{Model::SYNC, MsgPassing::PUSH, Scheduling::DATA_DRIVEN};
inline const AlgoVariant AlgoVariant::SYNC_PULL_TD{Model::SYNC, MsgPassing::PULL, Scheduling::TOPOLOGY_DRIVEN};
inline const AlgoVariant AlgoVariant::SYNC_PULL_DD{Model::SYNC, MsgPassing::PULL, Scheduling::DATA_DRIVEN};

class EngineOptions {
public:
    EngineOptions() { sg_engine_options_default(&o_); }

    void SetSampleBasedPriority(float cut_threshold = 0.2f)
    {
        o_.priority = SG_PRIO_SAMPLING;
        o_.cut_threshold = cut_threshold;
    }
    void SetTwoLevelBasedPriority(float priority_threshold)
    {
        o_.priority = SG_PRIO_LOW_HIGH;
        o_.priority_threshold = priority_threshold;
    }
    void ForceVariant(AlgoVariant v) { o_.force_variant = v.Code(); }
    bool IsForceVariant() const { return o_.force_variant != SG_VARIANT_AUTO; }
    void SetFused() { o_.fused = 1; }
    bool IsFused() const { return o_.fused != 0; }
    void SetLoadBalancing(MsgPassing msg_passing, LoadBalancing lb)
    {
        (msg_passing == MsgPassing::PUSH ? o_.lb_push : o_.lb_pull) = (int32_t)lb;
    }
    // direction-switch thresholds (the reference's --alpha / --beta flags)
    void SetAlphaBeta(double alpha, double beta)
    {
        o_.alpha = alpha;
        o_.beta = beta;
    }
    void SetBlockSize(int block_size) { o_.block_size = block_size; }
    void SetMaxIteration(int max_iteration) { o_.max_iteration = max_iteration; }
    // FLAGS_edge_factor: the traversal's unvisited-edge budget is edge_factor * nedges
    void SetEdgeFactor(double edge_factor) { o_.edge_factor = edge_factor; }
    const sg_engine_options &c_options() const { return o_; }

private:
    sg_engine_options o_{};
};

}  // namespace common

namespace policy {
enum class AlgoType { TRAVERSAL_SCHEME, ITERATIVE_SCHEME };
}

namespace graphs {
struct NoWeight {
};
// A host CSR (the reference's host::CSRGraph, csr_graph.h:104-476): row_start
// has n + 1 entries, edge_dst m; edge_weights empty or m entries.
struct CSRGraph {
    uint32_t nnodes = 0;
    uint64_t nedges = 0;
    std::vector<uint32_t> row_start, edge_dst, edge_weights;
    bool undirected = false;
};

// --graphfile / --format loading (GetCachedGraph, src/utils/utils.cpp:39-58,
// bound as in include/utils/graphs/traversal.h:147-158): format "gr",
// "metis" or "market"; `clean` takes SG_CLEAN_* flags (the gtools cleanups).
// nedges is row_start[nnodes], the count the engine traverses.
inline CSRGraph ReadGraph(const std::string &path, const std::string &format, int clean = 0,
                          bool undirected = false)
{
    int fmt = format == "gr" ? SG_FORMAT_GR : format == "metis" ? SG_FORMAT_METIS : format == "market" ? SG_FORMAT_MARKET : -1;
    if (fmt < 0) throw Error(SG_ERR_INVALID, "sepgraph: unsupported graph format " + format);
    sg_host_csr *h = nullptr;
    check(sg_file_read(path.c_str(), fmt, &h));
    struct Free {
        sg_host_csr *h;
        ~Free() { sg_file_free(h); }
    } guard{h};
    if (clean) check(sg_file_clean(h, clean));
    CSRGraph g;
    int weighted = 0;
    check(sg_file_info(h, &g.nnodes, &g.nedges, &weighted, nullptr));
    g.row_start.resize((size_t)g.nnodes + 1);
    g.edge_dst.resize(g.nedges);
    if (weighted) g.edge_weights.resize(g.nedges);
    check(sg_file_copy(h, g.row_start.data(), g.edge_dst.data(), weighted ? g.edge_weights.data() : nullptr));
    g.undirected = undirected;
    return g;
}

// --gen_weights: fill missing weights with the reference drivers' sequence
// (include/utils/graphs/traversal.h:174-186); a graph with weights is kept.
inline void GenWeights(CSRGraph &g, int32_t weight_range = 100)
{
    if (!g.edge_weights.empty()) return;
    g.edge_weights.resize(g.nedges);
    check(sg_gen_weights_std(g.nedges, weight_range, g.edge_weights.data()));
}
}  // namespace graphs

namespace apps {
// The library's built-in apps, selected rather than re-implemented: the
// per-edge hooks of samples/hybrid_{bfs,sssp,pr,bc}/hybrid_*.cu are compiled
// into its kernels. Templates with the reference's app parameters, so that
// Engine<level_t, level_t, NoWeight, apps::BFS, index_t> reads as the
// reference's Engine<..., hybrid_bfs::BFS, index_t, bool>.
template <typename TValue, typename TBuffer, typename TWeight, typename... InitArgs>
struct BFS {};       // levels from a source: InitGraph(source[, non_atomic])
template <typename TValue, typename TBuffer, typename TWeight, typename... InitArgs>
struct SSSP {};      // distances from a source: InitGraph(source[, ...])
template <typename TValue, typename TBuffer, typename TWeight, typename... InitArgs>
struct PageRank {};  // ranks: InitGraph(tolerance) (damping 0.85)
template <typename TValue, typename TBuffer, typename TWeight, typename... InitArgs>
struct BC {};        // single-source BC dependency: InitGraph(source)
}  // namespace apps

namespace api {
// The reference's app base (include/framework/variants/api.cuh:22-120), with
// device hooks: defined in sepgraph_app.hpp, which needs hipcc.
template <typename TValue, typename TBuffer, typename TWeight>
struct AppBase;
}  // namespace api

namespace detail {
template <typename T>
struct always_false : std::false_type {
};

template <template <typename, typename, typename, typename...> class A, template <typename, typename, typename, typename...> class B>
struct same_app : std::false_type {
};
template <template <typename, typename, typename, typename...> class A>
struct same_app<A, A> : std::true_type {
};

// Runs an app derived from api::AppBase on the generic persistent engine;
// defined in sepgraph_app.hpp (compile the driver with hipcc and include it).
template <typename TApp, typename TValue, typename TBuffer, typename TWeight>
struct AppRunner;
}  // namespace detail

namespace engine {

// The reference's Engine (include/framework/framework.cuh:56-397):
// TAppImpl<TValue, TBuffer, TWeight, UnusedData...> is the app; UnusedData are
// its constructor arguments, given to InitGraph. The built-in apps::* run the
// library's kernels through the C ABI; an app derived from api::AppBase runs
// its own hooks on the generic engine of sepgraph_app.hpp.
template <typename TValue, typename TBuffer, typename TWeight,
          template <typename, typename, typename, typename...> class TAppImpl, typename... UnusedData>
class Engine {
    typedef TAppImpl<TValue, TBuffer, TWeight, UnusedData...> AppImplDeviceObject;
    static constexpr bool kBfs = detail::same_app<TAppImpl, apps::BFS>::value;
    static constexpr bool kSssp = detail::same_app<TAppImpl, apps::SSSP>::value;
    static constexpr bool kPr = detail::same_app<TAppImpl, apps::PageRank>::value;
    static constexpr bool kBc = detail::same_app<TAppImpl, apps::BC>::value;
    static constexpr bool kBuiltin = kBfs || kSssp || kPr || kBc;

public:
    // Graph sizes (the reference's GraphDatum nnodes / nedges).
    struct GraphDatum {
        uint32_t nnodes = 0;
        uint64_t nedges = 0;
    };

    explicit Engine(policy::AlgoType algo_type) : algo_type_(algo_type) {}
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;
    ~Engine()
    {
        if (e_) sg_engine_destroy(e_);
        if (g_) sg_graph_destroy(g_);
    }

    void SetOptions(common::EngineOptions &options)
    {
        options_ = options;
        if (e_) check(sg_engine_set_options(e_, &options_.c_options()));
    }

    void LoadGraph(const graphs::CSRGraph &csr)
    {
        // a second LoadGraph replaces the first graph and its engine
        if (e_) {
            sg_engine_destroy(e_);
            e_ = nullptr;
        }
        if (g_) {
            sg_graph_destroy(g_);
            g_ = nullptr;
        }
        csr_ = csr;
        datum_.nnodes = csr_.nnodes;
        datum_.nedges = csr_.nedges;
        const uint32_t *w = (!std::is_same<TWeight, graphs::NoWeight>::value && !csr_.edge_weights.empty())
                                ? csr_.edge_weights.data()
                                : nullptr;
        // an app's own hooks see vertex ids, so its graph keeps the caller's order
        const uint32_t flags = (csr_.undirected ? (uint32_t)SG_GRAPH_UNDIRECTED : 0u) |
                               (kBuiltin ? 0u : (uint32_t)SG_GRAPH_NO_RELABEL);
        check(sg_graph_create_ex(csr_.nnodes, csr_.nedges, csr_.row_start.data(), csr_.edge_dst.data(), w, nullptr,
                                 nullptr, nullptr, flags, &g_));
        check(sg_engine_create(g_, &e_));
        check(sg_engine_set_options(e_, &options_.c_options()));
    }

    // The app's initial state: the constructor arguments of the app (the
    // source vertex for BFS, SSSP and BC; the tolerance for PageRank).
    void InitGraph(UnusedData... args) { init_ = std::make_tuple(args...); }

    void Start()
    {
        if (!e_) throw Error(SG_ERR_STATE, "sepgraph: Start before LoadGraph");
        if constexpr (kBfs) {
            check(sg_bfs(e_, source()));
        } else if constexpr (kSssp) {
            check(sg_sssp(e_, source()));
        } else if constexpr (kPr) {
            check(sg_pagerank(e_, 0.85, (double)std::get<0>(init_), 0,
                              algo_type_ == policy::AlgoType::ITERATIVE_SCHEME ? SG_PR_ITERATIVE : SG_PR_PULL_TD));
        } else if constexpr (kBc) {
            const uint32_t s = source();
            check(sg_bc_dependency(e_, &s, 1, 1));
        } else if constexpr (std::is_base_of<api::AppBase<TValue, TBuffer, TWeight>, AppImplDeviceObject>::value) {
            const AppImplDeviceObject app = std::make_from_tuple<AppImplDeviceObject>(init_);
            stats_ = detail::AppRunner<AppImplDeviceObject, TValue, TBuffer, TWeight>::Start(
                e_, g_, app, algo_type_ == policy::AlgoType::ITERATIVE_SCHEME, options_.c_options(), value_,
                buffer_);
            gathered_ = true;
            return;
        } else {
            static_assert(detail::always_false<AppImplDeviceObject>::value,
                          "sepgraph::engine::Engine: the app must be one of apps::BFS, apps::SSSP, apps::PageRank, "
                          "apps::BC, or derive from sepgraph::api::AppBase (include sepgraph_app.hpp)");
        }
        check(sg_engine_sync(e_));
        check(sg_engine_get_stats(e_, &stats_));
    }

    void PrintInfo() const
    {
        std::printf("rounds %u (push %u, pull %u), time %.3f ms\n", stats_.rounds, stats_.push_rounds,
                    stats_.pull_rounds, stats_.kernel_ms);
    }

    const std::vector<TValue> &GatherValue()
    {
        if constexpr (kBuiltin) {
            static_assert(sizeof(TValue) == 4, "values are 32-bit");
            value_.resize(csr_.nnodes);
            check(sg_engine_gather_value(e_, value_.data(), (uint64_t)value_.size() * sizeof(TValue)));
        } else {
            if (!gathered_) throw Error(SG_ERR_STATE, "sepgraph: GatherValue before Start");
        }
        return value_;
    }

    // The app's buffers after the run (apps derived from api::AppBase).
    const std::vector<TBuffer> &GatherBuffer()
    {
        if (kBuiltin || !gathered_) throw Error(SG_ERR_STATE, "sepgraph: GatherBuffer needs an AppBase app that ran");
        return buffer_;
    }

    const graphs::CSRGraph &CSRGraph() const { return csr_; }
    const GraphDatum &GetGraphDatum() const { return datum_; }
    const sg_stats &Stats() const { return stats_; }
    sg_engine *Handle() const { return e_; }

private:
    // The first InitGraph argument as a vertex id, clamped to [0, n - 1] as
    // the reference's drivers clamp --source_node (hybrid_bfs.cu:180,
    // hybrid_sssp.cu:158).
    uint32_t source() const
    {
        const long long s = (long long)std::get<0>(init_);
        const long long hi = csr_.nnodes ? (long long)csr_.nnodes - 1 : 0;
        return (uint32_t)(s < 0 ? 0 : s > hi ? hi : s);
    }

    policy::AlgoType algo_type_;
    common::EngineOptions options_;
    graphs::CSRGraph csr_;
    GraphDatum datum_;
    sg_graph *g_ = nullptr;
    sg_engine *e_ = nullptr;
    std::tuple<UnusedData...> init_{};
    std::vector<TValue> value_;
    std::vector<TBuffer> buffer_;
    bool gathered_ = false;
    sg_stats stats_{};
};

}  // namespace engine
}  // namespace sepgraph
