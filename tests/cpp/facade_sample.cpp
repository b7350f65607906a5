// A hybrid_* style driver written against the C++ facade (include/sepgraph.hpp):
// loads a graph file the way --graphfile/--format do, runs BFS, SSSP, PageRank
// and single-source BC the way samples/hybrid_{bfs,sssp,pr,bc}/*.cu drive the
// reference Engine, and writes each result vector to <out>.<app>.bin.
//
// usage: facade_sample graphfile format out_prefix source undirected(0|1) [gen_weight_range]
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "sepgraph.hpp"

using namespace sepgraph;

template <typename T>
static void dump(const std::string &path, const std::vector<T> &v)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) std::exit(3);
    std::fclose(f);
}

int main(int argc, char **argv)
{
    if (argc != 6 && argc != 7) {
        std::fprintf(stderr, "usage: %s graphfile format out_prefix source undirected [gen_weight_range]\n", argv[0]);
        return 2;
    }
    graphs::CSRGraph g;
    try {
        g = graphs::ReadGraph(argv[1], argv[2], 0, std::atoi(argv[5]) != 0);
        if (argc == 7) graphs::GenWeights(g, std::atoi(argv[6]));
    } catch (const Error &e) {
        std::fprintf(stderr, "%s (code %d)\n", e.what(), e.code());
        return 2;
    }
    const uint32_t source = (uint32_t)std::strtoul(argv[4], nullptr, 10);
    const std::string out = argv[3];
    try {
        {
            typedef engine::Engine<int32_t, int32_t, graphs::NoWeight, apps::BFS, uint32_t> HybridEngine;
            HybridEngine engine(policy::AlgoType::TRAVERSAL_SCHEME);
            common::EngineOptions opt;
            opt.SetAlphaBeta(15, 18);
            engine.SetOptions(opt);
            engine.LoadGraph(g);
            engine.InitGraph(source);
            engine.Start();
            engine.PrintInfo();
            dump(out + ".bfs.bin", engine.GatherValue());
        }
        {
            typedef engine::Engine<uint32_t, uint32_t, uint32_t, apps::SSSP, uint32_t> HybridEngine;
            HybridEngine engine(policy::AlgoType::TRAVERSAL_SCHEME);
            common::EngineOptions opt;
            opt.ForceVariant(common::AlgoVariant::ASYNC_PUSH_DD);
            engine.LoadGraph(g);
            engine.SetOptions(opt);
            engine.InitGraph(source);
            engine.Start();
            dump(out + ".sssp.bin", engine.GatherValue());
        }
        {
            typedef engine::Engine<float, float, graphs::NoWeight, apps::PageRank, double> HybridEngine;
            HybridEngine engine(policy::AlgoType::TRAVERSAL_SCHEME);
            engine.LoadGraph(g);
            engine.InitGraph(1e-6);
            engine.Start();
            dump(out + ".pr.bin", engine.GatherValue());
        }
        {
            typedef engine::Engine<float, float, graphs::NoWeight, apps::BC, uint32_t> HybridEngine;
            HybridEngine engine(policy::AlgoType::TRAVERSAL_SCHEME);
            engine.LoadGraph(g);
            engine.InitGraph(source);
            engine.Start();
            dump(out + ".bc.bin", engine.GatherValue());
        }
    } catch (const Error &e) {
        std::fprintf(stderr, "%s (code %d)\n", e.what(), e.code());
        return 1;
    }
    return 0;
}
