// hybrid_bfs: the reference's BFS driver (samples/hybrid_bfs/main.cpp and
// hybrid_bfs.cu:173-243) over libsepgraph.so, with its flags, JSON keys,
// --output format and --check.
#include "cli.hpp"

using namespace sepgraph;

int main(int argc, char **argv)
{
    cli::Flags flags;
    cli::Common c;
    c.Register(flags);
    int source_node, top_ranks;
    bool sparse, non_atomic, print_ranks;
    std::string variant;
    flags.Int("source_node", &source_node, 0, "The source node for the BFS traversal (clamped to [0, nnodes-1])");
    flags.Bool("sparse", &sparse, false, "use async/push/dd + fusion for high-diameter");
    flags.String("variant", &variant, "", "force variant as async/sync push/pull dd/td");
    flags.Bool("non_atomic", &non_atomic, true, "using non-atomic operation for accumulate");
    flags.Int("top_ranks", &top_ranks, 10, "The number of top ranks to compare for PR regression");
    flags.Bool("print_ranks", &print_ranks, false, "Write out ranks to output");
    if (!flags.Parse(argc, argv)) return 1;

    std::printf("\nTesting single GPU BFS\n--------------------\n\n");
    graphs::CSRGraph g;
    if (int rc = cli::LoadGraph(c, g)) return rc;
    const uint32_t source = (uint32_t)std::min<long long>(std::max<long long>(0, source_node), g.nnodes - 1);

    common::EngineOptions opt;
    if (!variant.empty()) {
        if (variant == "async_push_dd") opt.ForceVariant(common::AlgoVariant::ASYNC_PUSH_DD);
        else if (variant == "async_pull_dd") opt.ForceVariant(common::AlgoVariant::ASYNC_PULL_DD);
        else {
            std::fprintf(stderr, "unsupported variant\n");
            return 1;
        }
    }
    if (sparse) {
        opt.SetFused();
        opt.ForceVariant(common::AlgoVariant::ASYNC_PUSH_DD);
        opt.SetLoadBalancing(common::MsgPassing::PUSH, common::LoadBalancing::NONE);
    } else {
        opt.SetLoadBalancing(common::MsgPassing::PULL, common::LoadBalancing::NONE);
    }
    if (!cli::ApplyCommon(c, flags, g, opt, sparse, !sparse)) return 1;

    cli::Json json;
    bool success = true;
    try {
        typedef engine::Engine<uint32_t, uint32_t, graphs::NoWeight, apps::BFS, uint32_t, bool> HybridEngine;
        HybridEngine engine(policy::AlgoType::TRAVERSAL_SCHEME);
        engine.LoadGraph(g);
        engine.SetOptions(opt);
        // levels are exact whatever non_atomic says: a vertex is claimed once
        engine.InitGraph(source, non_atomic);
        const double t0 = cli::NowMs();
        engine.Start();
        const double total = cli::NowMs() - t0;
        engine.PrintInfo();
        cli::EngineJson(json, c, opt, engine.Stats(), total, g.nnodes, g.nedges, false);
        json.Write("non_atomic", non_atomic ? "YES" : "NO");
        if (c.trace) json.Write("round", cli::TraceRows(engine.Handle(), g.nedges, c.edge_factor));
        const std::vector<uint32_t> &levels = engine.GatherValue();
        if (c.check) {
            const int errors = cli::CountErrors(levels, cli::HostBfs(g, source));
            success = errors == 0;
            std::printf("total errors: %d\n", errors);
        } else {
            std::printf("Warning: Result not checked\n");
        }
        if (!c.output.empty() && !cli::OutputLevels(c.output, levels)) success = false;
    } catch (const Error &e) {
        std::fprintf(stderr, "%s (code %d)\n", e.what(), e.code());
        return 1;
    }
    if (!json.Save(c.json)) std::fprintf(stderr, "could not write %s\n", c.json.c_str());
    std::printf("Overall: Test %s\n", success ? "passed" : "FAILED");
    return 0;
}
