// hybrid_pr: the reference's PageRank driver (samples/hybrid_pr/main.cpp and
// hybrid_pr.cu:128-210) over libsepgraph.so, with its flags, JSON keys,
// --output format and --check. Ranks are the normalised PageRank (damping
// 0.85, uniform teleport and dangling mass), a scalar multiple of the
// reference's unnormalised residual-push ranks, so their order is the same;
// --error is the run's L1 tolerance.
#include <cfloat>

#include "cli.hpp"

using namespace sepgraph;

int main(int argc, char **argv)
{
    cli::Flags flags;
    cli::Common c;
    c.Register(flags);
    double error, cut_threshold;
    bool sparse, print_ranks, norm;
    int top_ranks;
    std::string variant;
    flags.Double("error", &error, 0.01, "PR error tolerance");
    flags.Double("cut_threshold", &cut_threshold, 0, "Cut threshold for index calculation");
    flags.Bool("sparse", &sparse, false, "disable load-balancing for sparse graph");
    flags.String("variant", &variant, "", "force variant as async/sync push/pull dd/td");
    flags.Int("top_ranks", &top_ranks, 10, "The number of top ranks to compare for PR regression");
    flags.Bool("print_ranks", &print_ranks, false, "Write out ranks to output");
    flags.Bool("norm", &norm, false, "Normalize PR output ranks (L1)");
    if (!flags.Parse(argc, argv)) return 1;

    std::printf("\nTesting single GPU PR\n--------------------\n\n");
    graphs::CSRGraph g;
    if (int rc = cli::LoadGraph(c, g)) return rc;

    common::EngineOptions opt;
    if (cut_threshold > 0) opt.SetSampleBasedPriority((float)cut_threshold);
    if (sparse) {
        opt.SetLoadBalancing(common::MsgPassing::PUSH, common::LoadBalancing::NONE);
        opt.SetLoadBalancing(common::MsgPassing::PULL, common::LoadBalancing::NONE);
    }
    if (!variant.empty()) {
        if (variant == "async_push_dd") opt.ForceVariant(common::AlgoVariant::ASYNC_PUSH_DD);
        else if (variant == "async_push_td") opt.ForceVariant(common::AlgoVariant::ASYNC_PUSH_TD);
        else if (variant == "sync_pull_td") opt.ForceVariant(common::AlgoVariant::SYNC_PULL_TD);
        else {
            std::fprintf(stderr, "unsupported variant\n");
            return 1;
        }
    }
    if (!cli::ApplyCommon(c, flags, g, opt, sparse, sparse)) return 1;

    cli::Json json;
    bool success = true;
    try {
        typedef engine::Engine<float, float, graphs::NoWeight, apps::PageRank, double> HybridEngine;
        HybridEngine engine(policy::AlgoType::ITERATIVE_SCHEME);
        engine.SetOptions(opt);
        engine.LoadGraph(g);
        engine.InitGraph(error);
        const double t0 = cli::NowMs();
        engine.Start();
        const double total = cli::NowMs() - t0;
        engine.PrintInfo();
        cli::EngineJson(json, c, opt, engine.Stats(), total, g.nnodes, g.nedges, true);
        json.Write("error_tolerance", (double)(float)error);
        const std::vector<float> &ranks = engine.GatherValue();
        double sum = 0;
        for (float r : ranks) sum += r;
        std::printf("Total rank : %f\n", sum);
        std::printf("max residual: %f\n", engine.Stats().residual);
        if (c.check) {
            const int errors = cli::CountRelErrors(ranks, cli::HostPageRank(g, 0.85, 1e-10));
            success = errors == 0;
            std::printf("total errors: %d\n", errors);
        } else {
            std::printf("Warning: Result not checked\n");
        }
        if (!c.output.empty()) {
            // PageRankOutput (hybrid_pr_host.cpp:163-225): stable sort by rank,
            // listed from the highest
            FILE *f = std::fopen(c.output.c_str(), "w");
            if (!f) {
                std::fprintf(stderr, "Could not open '%s' for writing\n", c.output.c_str());
                success = false;
            } else {
                std::vector<uint32_t> order(ranks.size());
                for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
                std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return ranks[a] < ranks[b]; });
                const int n = (int)ranks.size();
                std::fprintf(f, "ALPHA %*e EPSILON %*e\n", FLT_DIG, 0.85, FLT_DIG, error);
                std::fprintf(f, "RANKS 1--%d of %d\n", n, n);
                for (int i = 1; i <= n; i++) {
                    const uint32_t node = order[n - i];
                    if (!print_ranks) std::fprintf(f, "%d %u\n", i, node);
                    else std::fprintf(f, "%u %*e\n", node, FLT_DIG, norm ? ranks[node] / sum : (double)ranks[node]);
                }
                std::fclose(f);
            }
        }
    } catch (const Error &e) {
        std::fprintf(stderr, "%s (code %d)\n", e.what(), e.code());
        return 1;
    }
    if (!json.Save(c.json)) std::fprintf(stderr, "could not write %s\n", c.json.c_str());
    std::printf("Overall: Test %s\n", success ? "passed" : "FAILED");
    return 0;
}
