// hybrid_sssp: the reference's SSSP driver (samples/hybrid_sssp/main.cpp and
// hybrid_sssp.cu:151-242) over libsepgraph.so, with its flags (including the
// two-level priority delta heuristic of --sparse), JSON keys, --output format
// and --check.
#include "cli.hpp"

using namespace sepgraph;

int main(int argc, char **argv)
{
    cli::Flags flags;
    cli::Common c;
    c.Register(flags);
    int source_node;
    bool sparse;
    std::string variant;
    flags.Int("source_node", &source_node, 0, "The source node for the SSSP traversal (clamped to [0, nnodes-1])");
    flags.Bool("sparse", &sparse, false, "use async/push/dd + fusion for high-diameter");
    flags.String("variant", &variant, "", "force variant as async/sync push/pull dd/td");
    if (!flags.Parse(argc, argv)) return 1;

    std::printf("\nTesting single GPU SSSP\n--------------------\n\n");
    graphs::CSRGraph g;
    if (int rc = cli::LoadGraph(c, g)) return rc;
    const uint32_t source = (uint32_t)std::min<long long>(std::max<long long>(0, source_node), g.nnodes - 1);

    common::EngineOptions opt;
    if (!variant.empty()) {
        if (variant == "async_push_dd") opt.ForceVariant(common::AlgoVariant::ASYNC_PUSH_DD);
        else if (variant == "async_pull_dd") opt.ForceVariant(common::AlgoVariant::ASYNC_PULL_DD);
        else if (variant == "async_push_td") opt.ForceVariant(common::AlgoVariant::ASYNC_PUSH_TD);
        else {
            std::fprintf(stderr, "unsupported variant\n");
            return 1;
        }
    }
    // delta = c * mean weight / mean degree with c = 32 (hybrid_sssp.cu:183-199)
    double weight_sum = 0;
    for (uint64_t e = 0; e < g.nedges; e++) weight_sum += g.edge_weights.empty() ? 1.0 : g.edge_weights[e];
    const int init_prio = g.nedges ? (int)(32 * (weight_sum / g.nedges) / (1.0 * g.nedges / g.nnodes)) : 0;
    std::printf("Priority delta: %u\n", init_prio);
    if (sparse) {
        opt.SetFused();
        opt.SetTwoLevelBasedPriority((float)init_prio);
        opt.ForceVariant(common::AlgoVariant::ASYNC_PUSH_DD);
        opt.SetLoadBalancing(common::MsgPassing::PUSH, common::LoadBalancing::NONE);
    }
    if (c.prio_delta > 0) {
        std::printf("Enable priority for scale-free dataset\n");
        opt.SetTwoLevelBasedPriority((float)c.prio_delta);
    }
    if (!cli::ApplyCommon(c, flags, g, opt, sparse, false)) return 1;

    cli::Json json;
    bool success = true;
    try {
        typedef engine::Engine<uint32_t, uint32_t, uint32_t, apps::SSSP, uint32_t> HybridEngine;
        HybridEngine engine(policy::AlgoType::TRAVERSAL_SCHEME);
        engine.LoadGraph(g);
        engine.SetOptions(opt);
        engine.InitGraph(source);
        const double t0 = cli::NowMs();
        engine.Start();
        const double total = cli::NowMs() - t0;
        engine.PrintInfo();
        cli::EngineJson(json, c, opt, engine.Stats(), total, g.nnodes, g.nedges, false);
        if (c.trace) json.Write("round", cli::TraceRows(engine.Handle(), g.nedges, c.edge_factor));
        const std::vector<uint32_t> &dist = engine.GatherValue();
        if (c.check) {
            const int errors = cli::CountErrors(dist, cli::HostSssp(g, source));
            success = errors == 0;
            std::printf("total errors: %d\n", errors);
        } else {
            std::printf("Warning: Result not checked\n");
        }
        if (!c.output.empty() && !cli::OutputLevels(c.output, dist)) success = false;
    } catch (const Error &e) {
        std::fprintf(stderr, "%s (code %d)\n", e.what(), e.code());
        return 1;
    }
    if (!json.Save(c.json)) std::fprintf(stderr, "could not write %s\n", c.json.c_str());
    std::printf("Overall: Test %s\n", success ? "passed" : "FAILED");
    return 0;
}
