// Command-line surface of the hybrid_* drivers: the reference's flags with
// their names and defaults (include/utils/app_skeleton.h:42-87 and the
// per-sample flags of samples/hybrid_*), its run-info JSON
// (Engine::SaveToJson, include/framework/framework.cuh:205-257, written by
// utils::JsonWriter, src/utils/to_json.cpp:13-28: one object, keys sorted,
// four-space indent), its output files (samples/hybrid_*/hybrid_*_host.cpp)
// and its --check host verification, so existing run scripts and result
// parsers work unchanged.
//
// Flags are parsed the way gflags parses them: --name=value, --name value,
// -name; a boolean flag alone means true and --noname means false; an
// unknown flag is an error.
#pragma once

#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <queue>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "../../include/sepgraph.hpp"

namespace cli {

// ------------------------------------------------------------------ flags
class Flags {
public:
    void Bool(const char *name, bool *p, bool def, const char *help) { add(name, 'b', p, help), *p = def; }
    void Int(const char *name, int *p, int def, const char *help) { add(name, 'i', p, help), *p = def; }
    void Double(const char *name, double *p, double def, const char *help) { add(name, 'd', p, help), *p = def; }
    void String(const char *name, std::string *p, const char *def, const char *help)
    {
        add(name, 's', p, help);
        *p = def;
    }
    bool Given(const std::string &name) const
    {
        auto it = index_.find(name);
        return it != index_.end() && flags_[it->second].given;
    }

    // Returns false (after printing why) on a malformed command line; exits
    // 0 after --help.
    bool Parse(int argc, char **argv)
    {
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i];
            if (a.size() < 2 || a[0] != '-') continue;  // positional arguments are left alone
            a = a.substr(a[1] == '-' ? 2 : 1);
            if (a == "help") {
                Help(argv[0]);
                std::exit(0);
            }
            std::string name = a, value;
            bool has_value = false;
            const size_t eq = a.find('=');
            if (eq != std::string::npos) {
                name = a.substr(0, eq);
                value = a.substr(eq + 1);
                has_value = true;
            }
            auto it = index_.find(name);
            if (it == index_.end() && !has_value && name.compare(0, 2, "no") == 0) {
                auto jt = index_.find(name.substr(2));
                if (jt != index_.end() && flags_[jt->second].kind == 'b') {
                    *(bool *)flags_[jt->second].p = false;
                    flags_[jt->second].given = true;
                    continue;
                }
            }
            if (it == index_.end()) {
                std::fprintf(stderr, "ERROR: unknown command line flag '%s'\n", name.c_str());
                return false;
            }
            Flag &f = flags_[it->second];
            if (!has_value && f.kind != 'b') {
                if (i + 1 >= argc) {
                    std::fprintf(stderr, "ERROR: flag '%s' is missing its argument\n", name.c_str());
                    return false;
                }
                value = argv[++i];
                has_value = true;
            }
            if (!set(f, has_value, value)) {
                std::fprintf(stderr, "ERROR: illegal value '%s' specified for flag '%s'\n", value.c_str(),
                             name.c_str());
                return false;
            }
            f.given = true;
        }
        return true;
    }

    void Help(const char *prog) const
    {
        std::printf("%s: flags\n", prog);
        for (const Flag &f : flags_) std::printf("  -%s (%s)\n", f.name.c_str(), f.help.c_str());
    }

private:
    struct Flag {
        std::string name;
        char kind;
        void *p;
        std::string help;
        bool given = false;
    };
    void add(const char *name, char kind, void *p, const char *help)
    {
        index_[name] = flags_.size();
        flags_.push_back(Flag{name, kind, p, help});
    }
    static bool set(Flag &f, bool has_value, const std::string &v)
    {
        char *end = nullptr;
        switch (f.kind) {
        case 'b':
            if (!has_value || v == "true" || v == "1" || v == "yes" || v == "t" || v == "y") *(bool *)f.p = true;
            else if (v == "false" || v == "0" || v == "no" || v == "f" || v == "n") *(bool *)f.p = false;
            else return false;
            return true;
        case 'i': {
            const long x = std::strtol(v.c_str(), &end, 10);
            if (v.empty() || *end) return false;
            *(int *)f.p = (int)x;
            return true;
        }
        case 'd': {
            const double x = std::strtod(v.c_str(), &end);
            if (v.empty() || *end) return false;
            *(double *)f.p = x;
            return true;
        }
        default:
            *(std::string *)f.p = v;
            return true;
        }
    }
    std::vector<Flag> flags_;
    std::map<std::string, size_t> index_;
};

// The flags every driver takes (include/utils/app_skeleton.h:42-87).
struct Common {
    bool interactive, check, verbose, trace, undirected, stats, gen_graph, gen_weights, priority, wl_sort, wl_unique;
    std::string cmdfile, output, lb_push, lb_pull, graphfile, format, json;
    int max_iteration, prio_delta, gen_nnodes, gen_factor, gen_method, gen_weight_range, block_size, grid_size;
    double edge_factor, alpha, beta, wl_alloc_factor;

    void Register(Flags &f)
    {
        f.Bool("interactive", &interactive, false, "Run an interactive session");
        f.String("cmdfile", &cmdfile, "", "A file with commands to execute");
        f.Int("max_iteration", &max_iteration, 200, "Maximum iterated rounds");
        f.Bool("priority", &priority, false, "Enable priority scheduling (may not available for all algorithms)");
        f.Int("prio_delta", &prio_delta, 0, "The soft priority delta");
        f.Double("edge_factor", &edge_factor, 1.0, "Unvisited edges = edge_factor * nedges");
        f.Double("alpha", &alpha, 15, "DO BFS/SSSP/BC heuristic parameter");
        f.Double("beta", &beta, 18, "DO BFS/SSSP/BC heuristic parameter");
        f.String("output", &output, "", "File to store output to");
        f.Bool("check", &check, false, "Check results");
        f.Bool("verbose", &verbose, true, "Verbose prints");
        f.Bool("trace", &trace, false, "Trace prints (effects performance)");
        f.String("lb_push", &lb_push, "", "load-balancing policy for push (none/coarse/fine/hybrid)");
        f.String("lb_pull", &lb_pull, "", "load-balancing policy for pull (none/coarse/fine/hybrid)");
        f.String("graphfile", &graphfile, "", "A file with a graph in Dimacs 10 format");
        f.String("format", &format, "gr", "graph format (gr/market/metis formats are supported)");
        f.Bool("undirected", &undirected, false, "treat input graph as undirected graph");
        f.String("json", &json, "", "output running data as json format");
        f.Bool("stats", &stats, false, "Print graph statistics");
        f.Bool("gen_graph", &gen_graph, false, "Generate a random graph");
        f.Int("gen_nnodes", &gen_nnodes, 100000, "Number of nodes for random graph generation");
        f.Int("gen_factor", &gen_factor, 10, "A factor number for graph generation");
        f.Int("gen_method", &gen_method, 0,
              "graph generation method: 0 random, 1 two-way chain without segment intersection, 2 two-way chain, "
              "3 full cliques per device");
        f.Bool("gen_weights", &gen_weights, false, "Generate edge weights if missing in graph input");
        f.Int("gen_weight_range", &gen_weight_range, 100, "The range to generate edge weights from");
        f.Int("block_size", &block_size, 256, "Block size for traversal kernels");
        f.Int("grid_size", &grid_size, 20, "Grid size for traversal kernels");
        f.Double("wl_alloc_factor", &wl_alloc_factor, 0.2, "Worklist allocation factor");
        f.Bool("wl_sort", &wl_sort, false, "Sort worklist by node id");
        f.Bool("wl_unique", &wl_unique, false, "Unique duplicated nodes");
    }
};

// ------------------------------------------------------------------ JSON
// One flat object, keys sorted, four-space indent (what utils::JsonWriter
// prints through nlohmann::json with std::setw(4)).
class Json {
public:
    void Write(const std::string &k, double v) { nums_[k] = fmt_num(v); }
    void Write(const std::string &k, int v) { nums_[k] = std::to_string(v); }
    void Write(const std::string &k, const std::string &v) { nums_[k] = quote(v); }
    void Write(const std::string &k, const char *v) { nums_[k] = quote(v); }
    void Write(const std::string &k, const std::vector<std::string> &v)
    {
        std::string s = "[";
        for (size_t i = 0; i < v.size(); i++) s += (i ? ",\n        " : "\n        ") + quote(v[i]);
        s += v.empty() ? "]" : "\n    ]";
        nums_[k] = s;
    }
    bool Save(const std::string &path) const
    {
        if (path.empty()) return true;
        FILE *f = std::fopen(path.c_str(), "w");
        if (!f) return false;
        std::fprintf(f, "{");
        size_t i = 0;
        for (const auto &kv : nums_)
            std::fprintf(f, "%s\n    %s: %s", i++ ? "," : "", quote(kv.first).c_str(), kv.second.c_str());
        std::fprintf(f, "\n}\n");
        return std::fclose(f) == 0;
    }

private:
    static std::string fmt_num(double v)
    {
        if (!std::isfinite(v)) return "null";
        char b[64];
        std::snprintf(b, sizeof(b), "%.9g", v);
        std::string s = b;
        if (s.find_first_of(".eE") == std::string::npos) s += ".0";
        return s;
    }
    static std::string quote(const std::string &v)
    {
        std::string s = "\"";
        for (char c : v) {
            if (c == '"' || c == '\\') s += '\\', s += c;
            else if (c == '\n') s += "\\n";
            else if ((unsigned char)c < 0x20) {
                char b[8];
                std::snprintf(b, sizeof(b), "\\u%04x", c);
                s += b;
            } else s += c;
        }
        return s + "\"";
    }
    std::map<std::string, std::string> nums_;
};

// ------------------------------------------------------------------ graph
// --graphfile/--format, or --gen_graph with --gen_method/--gen_nnodes/
// --gen_factor (the generators of include/groute/graphs/csr_graph.h:483-673
// on one device), then --gen_weights (include/utils/graphs/traversal.h:173-187).
// Returns 0, or the exit status after printing why not.
inline int LoadGraph(const Common &c, sepgraph::graphs::CSRGraph &g)
{
    using sepgraph::graphs::CSRGraph;
    if (c.interactive || !c.cmdfile.empty()) {
        std::fprintf(stderr, "interactive and command-file sessions are not supported; run one command per process\n");
        return 1;
    }
    if (c.gen_graph) {
        std::printf("\nGenerating graph, nnodes: %d, gen_factor: %d", c.gen_nnodes, c.gen_factor);
        const uint32_t n = (uint32_t)std::max(c.gen_nnodes, 0);
        if (n < 2 || c.gen_factor < 1) {
            std::fprintf(stderr, "\ngen_graph needs gen_nnodes >= 2 and gen_factor >= 1\n");
            return 1;
        }
        g = CSRGraph();
        g.nnodes = n;
        g.row_start.assign((size_t)n + 1, 0);
        if (c.gen_method == 1 || c.gen_method == 2) {
            // a two-way chain (one segment on one device)
            for (uint32_t v = 0; v < n; v++) {
                g.row_start[v] = (uint32_t)g.edge_dst.size();
                if (v >= 1) g.edge_dst.push_back(v - 1);
                if (v + 1 < n) g.edge_dst.push_back(v + 1);
            }
        } else if (c.gen_method == 3) {
            // one full clique
            if ((uint64_t)n * (n - 1) >= (1ull << 32)) {
                std::fprintf(stderr, "\nclique too large\n");
                return 1;
            }
            for (uint32_t v = 0; v < n; v++) {
                g.row_start[v] = (uint32_t)g.edge_dst.size();
                for (uint32_t u = 0; u < n; u++)
                    if (u != v) g.edge_dst.push_back(u);
            }
        } else {
            // each vertex draws 1..gen_factor distinct neighbours other than itself
            std::default_random_engine gen;
            std::uniform_int_distribution<int> nnb(1, c.gen_factor);
            std::uniform_int_distribution<uint32_t> node(0, n - 1);
            for (uint32_t v = 0; v < n; v++) {
                g.row_start[v] = (uint32_t)g.edge_dst.size();
                const int k = nnb(gen);
                std::set<uint32_t> seen;
                for (int i = 0; i < k; i++) {
                    uint32_t u;
                    do u = node(gen);
                    while (u == v || seen.count(u));
                    seen.insert(u);
                    g.edge_dst.push_back(u);
                }
            }
        }
        g.row_start[n] = (uint32_t)g.edge_dst.size();
        g.nedges = g.edge_dst.size();
        g.undirected = c.undirected;
    } else {
        if (c.graphfile.empty()) {
            std::printf("A Graph File must be provided\n");
            return 1;
        }
        std::printf("\nLoading graph %s (%s)\n", c.graphfile.c_str(), c.format.c_str());
        try {
            g = sepgraph::graphs::ReadGraph(c.graphfile, c.format, 0, c.undirected);
        } catch (const sepgraph::Error &e) {
            std::fprintf(stderr, "%s\n", e.what());
            return 2;
        }
        if (g.nnodes == 0) {
            std::printf("Empty graph!\n");
            return 1;
        }
        if (c.stats) {
            uint32_t dmax = 0;
            for (uint32_t v = 0; v < g.nnodes; v++) dmax = std::max(dmax, g.row_start[v + 1] - g.row_start[v]);
            std::printf("The graph has %u vertices, and %" PRIu64 " edges (avg. degree: %f, max. degree: %u)\n",
                        g.nnodes, g.nedges, (double)g.nedges / g.nnodes, dmax);
            std::exit(0);
        }
    }
    if (g.edge_weights.empty() && c.gen_weights) {
        if (c.verbose)
            std::printf("\nNo edge data in the input graph, generating edge weights from the range [%d, %d]\n", 1,
                        c.gen_weight_range);
        sepgraph::graphs::GenWeights(g, c.gen_weight_range);
    }
    return 0;
}

inline bool ParseLb(const std::string &s, sepgraph::common::LoadBalancing *lb)
{
    using sepgraph::common::LoadBalancing;
    if (s == "none") *lb = LoadBalancing::NONE;
    else if (s == "coarse") *lb = LoadBalancing::COARSE_GRAINED;
    else if (s == "fine") *lb = LoadBalancing::FINE_GRAINED;
    else if (s == "hybrid") *lb = LoadBalancing::HYBRID;
    else return false;
    return true;
}

inline const char *LbName(int lb)
{
    switch (lb) {
    case SG_LB_NONE: return "NONE";
    case SG_LB_COARSE: return "COARSE_GRAINED";
    case SG_LB_FINE: return "FINE_GRAINED";
    case SG_LB_HYBRID: return "HYBRID";
    default: return "DEFAULT";
    }
}

// EngineOptions from the common flags and what a sample already set
// (Engine::LoadOptions, framework.cuh:400-449): --lb_push / --lb_pull when
// the sample forced none, FINE_GRAINED for graphs of average degree >= 5,
// alpha = 0 forcing ASYNC_PUSH_DD; --block_size when given (256, 512 or 1024
// here; otherwise the library's tuned size), --max_iteration, --alpha,
// --beta, --edge_factor.
inline bool ApplyCommon(const Common &c, const Flags &f, const sepgraph::graphs::CSRGraph &g,
                        sepgraph::common::EngineOptions &opt, bool lb_push_forced, bool lb_pull_forced)
{
    using namespace sepgraph::common;
    const double avg_degree = g.nnodes ? (double)g.nedges / g.nnodes : 0.0;
    LoadBalancing lb;
    if (!lb_push_forced) {
        if (!c.lb_push.empty()) {
            if (!ParseLb(c.lb_push, &lb)) {
                std::fprintf(stderr, "unknown push load-balancing policy");
                return false;
            }
            opt.SetLoadBalancing(MsgPassing::PUSH, lb);
        } else if (avg_degree >= 5) {
            opt.SetLoadBalancing(MsgPassing::PUSH, LoadBalancing::FINE_GRAINED);
        }
    }
    if (!lb_pull_forced) {
        if (!c.lb_pull.empty()) {
            if (!ParseLb(c.lb_pull, &lb)) {
                std::fprintf(stderr, "unknown pull load-balancing policy");
                return false;
            }
            opt.SetLoadBalancing(MsgPassing::PULL, lb);
        } else if (avg_degree >= 5) {
            opt.SetLoadBalancing(MsgPassing::PULL, LoadBalancing::FINE_GRAINED);
        }
    }
    if (c.alpha == 0) {
        std::fprintf(stderr, "Warning: alpha = 0, A general method AsyncPushDD is used\n");
        opt.ForceVariant(AlgoVariant::ASYNC_PUSH_DD);
    }
    opt.SetAlphaBeta(c.alpha, c.beta);
    opt.SetEdgeFactor(c.edge_factor);
    opt.SetMaxIteration(c.max_iteration);
    if (f.Given("block_size")) opt.SetBlockSize(c.block_size);
    return true;
}

inline double NowMs()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// The engine-level keys of Engine::SaveToJson (framework.cuh:205-257). The
// per-round overheads the reference measures on the host are those of its
// host round loop; here the loop runs on the device, so they are 0.
inline void EngineJson(Json &j, const Common &c, const sepgraph::common::EngineOptions &opt, const sg_stats &st,
                       double total_ms, uint32_t nnodes, uint64_t nedges, bool iterative)
{
    const sg_engine_options &o = opt.c_options();
    for (const char *k : {"time_input_active_node", "time_output_active_node", "time_input_workload",
                          "time_output_workload", "time_queue2bitmap", "time_bitmap2queue", "time_rebuild_worklist",
                          "time_priority_sample", "time_sort_worklist", "time_unique_worklist"})
        j.Write(k, 0.0);
    const int rounds = (int)st.rounds;
    j.Write("time_kernel", st.kernel_ms);
    j.Write("time_total", total_ms);
    j.Write("time_per_round", rounds ? total_ms / rounds : 0.0);
    j.Write("num_iteration", rounds);
    if (opt.IsForceVariant()) {
        static const char *names[] = {"ASYNC_PUSH_TD", "ASYNC_PUSH_DD", "ASYNC_PULL_TD", "ASYNC_PULL_DD",
                                      "SYNC_PUSH_TD", "SYNC_PUSH_DD", "SYNC_PULL_TD", "SYNC_PULL_DD"};
        j.Write("force_variant", names[o.force_variant & 7]);
    }
    if (o.lb_push != SG_LB_DEFAULT) j.Write("force_push_load_balancing", LbName(o.lb_push));
    if (o.lb_pull != SG_LB_DEFAULT) j.Write("force_pull_load_balancing", LbName(o.lb_pull));
    if (o.priority == SG_PRIO_LOW_HIGH) {
        j.Write("priority_type", "low_high");
        j.Write("priority_delta", (double)o.priority_threshold);
    } else if (o.priority == SG_PRIO_SAMPLING) {
        j.Write("priority_type", "sampling");
        j.Write("cut_threshold", (double)o.cut_threshold);
    } else {
        j.Write("priority_type", "none");
    }
    // every run is one persistent multi-round kernel
    j.Write("fused_kernel", "YES");
    j.Write("max_iteration_reached", rounds == c.max_iteration ? "YES" : "NO");
    j.Write("device", "AMD Instinct MI355X");
    j.Write("dataset", c.graphfile);
    j.Write("nnodes", (int)nnodes);
    j.Write("nedges", (int)nedges);
    j.Write("algo_type", iterative ? "ITERATIVE_SCHEME" : "TRAVERSAL_SCHEME");
}

// --trace rows (the reference's per-round line, framework.cuh:337-345),
// rebuilt from the device trace; the total workload is replayed with the
// reference's uint32 decrement on push rounds (framework.cuh:562).
inline std::vector<std::string> TraceRows(sg_engine *e, uint64_t nedges, double edge_factor)
{
    std::vector<sg_round_trace> t(4096);
    int nt = 0;
    std::vector<std::string> out;
    if (sg_engine_get_trace(e, t.data(), (int)t.size(), &nt) != SG_OK) return out;
    uint32_t total = (uint32_t)((double)nedges * edge_factor);
    static const char *names[] = {"ASYNC_PUSH_TD", "ASYNC_PUSH_DD", "ASYNC_PULL_TD", "ASYNC_PULL_DD",
                                  "SYNC_PUSH_TD", "SYNC_PUSH_DD", "SYNC_PULL_TD", "SYNC_PULL_DD"};
    for (int i = 0; i + 1 < nt; i++) {
        const sg_round_trace &r = t[i], &nx = t[i + 1];
        const bool push = r.variant == SG_ASYNC_PUSH_DD || r.variant == SG_SYNC_PUSH_DD;
        if (push) total -= (uint32_t)r.in_edges;
        char b[512];
        std::snprintf(b, sizeof(b),
                      "Round: %d Policy to execute: %s Time: %f In-nodes: %u Out-nodes: %u Input-edges: %u "
                      "Output-edges: %u Total-workload: %u",
                      (int)r.round, r.variant >= 0 ? names[r.variant & 7] : "NONE", (nx.time_us - r.time_us) / 1000.0,
                      r.in_nodes, nx.variant >= 0 ? nx.in_nodes : 0u, (uint32_t)r.in_edges,
                      nx.variant >= 0 ? (uint32_t)nx.in_edges : 0u, total);
        out.push_back(b);
    }
    return out;
}

// ------------------------------------------------------------------ host checks
// What --check compares against: the reference's host verification
// (samples/hybrid_*/hybrid_*_host.cpp), restated.
inline std::vector<uint32_t> HostBfs(const sepgraph::graphs::CSRGraph &g, uint32_t s)
{
    std::vector<uint32_t> lv(g.nnodes, UINT32_MAX);
    std::deque<uint32_t> q;
    lv[s] = 0;
    q.push_back(s);
    while (!q.empty()) {
        const uint32_t u = q.front();
        q.pop_front();
        for (uint32_t e = g.row_start[u]; e < g.row_start[u + 1]; e++) {
            const uint32_t v = g.edge_dst[e];
            if (lv[v] == UINT32_MAX) {
                lv[v] = lv[u] + 1;
                q.push_back(v);
            }
        }
    }
    return lv;
}

inline std::vector<uint32_t> HostSssp(const sepgraph::graphs::CSRGraph &g, uint32_t s)
{
    std::vector<uint32_t> d(g.nnodes, UINT32_MAX);
    std::deque<uint32_t> q;
    std::vector<char> inq(g.nnodes, 0);
    d[s] = 0;
    q.push_back(s);
    inq[s] = 1;
    while (!q.empty()) {
        const uint32_t u = q.front();
        q.pop_front();
        inq[u] = 0;
        for (uint32_t e = g.row_start[u]; e < g.row_start[u + 1]; e++) {
            const uint32_t v = g.edge_dst[e];
            const uint32_t w = g.edge_weights.empty() ? 1u : g.edge_weights[e];
            if (d[u] + w < d[v]) {
                d[v] = d[u] + w;
                if (!inq[v]) {
                    inq[v] = 1;
                    q.push_back(v);
                }
            }
        }
    }
    return d;
}

// Mismatch count as BFSCheckErrors / SSSPCheckErrors (over and miss errors).
inline int CountErrors(const std::vector<uint32_t> &got, const std::vector<uint32_t> &want)
{
    if (got.size() != want.size()) return (int)std::abs((long long)got.size() - (long long)want.size());
    int over = 0, miss = 0;
    for (size_t i = 0; i < want.size(); i++) {
        if (got[i] > want[i]) over++;
        else if (got[i] < want[i]) miss++;
    }
    if (miss) std::printf("Miss errors: %d\n\n", miss);
    if (over) std::printf("Over errors: %d\n\n", over);
    return over + miss;
}

// FP64 power iteration: damping d, uniform teleport and dangling mass, until
// the L1 change is below tol.
inline std::vector<double> HostPageRank(const sepgraph::graphs::CSRGraph &g, double d, double tol)
{
    const uint32_t n = g.nnodes;
    std::vector<double> x(n, 1.0 / n), y(n);
    for (int it = 0; it < 10000; it++) {
        double dangling = 0.0;
        std::fill(y.begin(), y.end(), 0.0);
        for (uint32_t u = 0; u < n; u++) {
            const uint32_t deg = g.row_start[u + 1] - g.row_start[u];
            if (!deg) {
                dangling += x[u];
                continue;
            }
            const double c = x[u] / deg;
            for (uint32_t e = g.row_start[u]; e < g.row_start[u + 1]; e++) y[g.edge_dst[e]] += c;
        }
        double diff = 0.0;
        for (uint32_t v = 0; v < n; v++) {
            const double nv = (1.0 - d) / n + d * (y[v] + dangling / n);
            diff += std::fabs(nv - x[v]);
            x[v] = nv;
        }
        if (diff < tol) break;
    }
    return x;
}

// Single-source Brandes dependencies, halved as the reference's BC driver
// reports them (hybrid_bc_host.cpp:16-92, hybrid_bc.cu:374-376).
inline std::vector<double> HostBc(const sepgraph::graphs::CSRGraph &g, uint32_t s)
{
    const uint32_t n = g.nnodes;
    std::vector<double> sigma(n, 0.0), delta(n, 0.0);
    std::vector<int64_t> dist(n, -1);
    std::vector<uint32_t> order;
    std::deque<uint32_t> q;
    sigma[s] = 1.0;
    dist[s] = 0;
    q.push_back(s);
    while (!q.empty()) {
        const uint32_t u = q.front();
        q.pop_front();
        order.push_back(u);
        for (uint32_t e = g.row_start[u]; e < g.row_start[u + 1]; e++) {
            const uint32_t v = g.edge_dst[e];
            if (dist[v] < 0) {
                dist[v] = dist[u] + 1;
                q.push_back(v);
            }
            if (dist[v] == dist[u] + 1) sigma[v] += sigma[u];
        }
    }
    // the source's own dependency is not part of its BC (the reference's
    // backward loop stops above depth 0, hybrid_bc_host.cpp:64-77)
    for (size_t i = order.size(); i-- > 1;) {
        const uint32_t u = order[i];
        for (uint32_t e = g.row_start[u]; e < g.row_start[u + 1]; e++) {
            const uint32_t v = g.edge_dst[e];
            if (dist[v] == dist[u] + 1) delta[u] += sigma[u] / sigma[v] * (1.0 + delta[v]);
        }
    }
    for (double &x : delta) x *= 0.5;
    return delta;
}

// Relative-error count with the reference's ERROR_THRESHOLD (0.05) and its
// near-zero rule (hybrid_bc_host.cpp:94-126, hybrid_pr_host.cpp:125-160).
inline int CountRelErrors(const std::vector<float> &got, const std::vector<double> &want)
{
    if (got.size() != want.size()) return (int)std::abs((long long)got.size() - (long long)want.size());
    int errors = 0;
    for (size_t i = 0; i < want.size(); i++) {
        const double g = got[i], w = want[i];
        const bool bad = std::fabs(g) < 0.01 ? std::fabs(g - w) > 0.05 : std::fabs((g - w) / w) > 0.05;
        errors += bad ? 1 : 0;
    }
    return errors;
}

// --output of BFS levels / SSSP distances: "node value" or "node INF"
// (hybrid_bfs_host.cpp:103-125, hybrid_sssp_host.cpp:97-115).
inline bool OutputLevels(const std::string &path, const std::vector<uint32_t> &v)
{
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "Could not open '%s' for writing\n", path.c_str());
        return false;
    }
    for (size_t i = 0; i < v.size(); i++) {
        if (v[i] == UINT32_MAX) std::fprintf(f, "%u INF\n", (unsigned)i);
        else std::fprintf(f, "%u %u\n", (unsigned)i, v[i]);
    }
    return std::fclose(f) == 0;
}

}  // namespace cli
