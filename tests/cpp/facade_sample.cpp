// A hybrid_* style driver written against the C++ facade (include/sepgraph.hpp):
// loads a binary CSR, runs BFS, SSSP, PageRank and single-source BC the way
// samples/hybrid_{bfs,sssp,pr,bc}/*.cu drive the reference Engine, and writes
// each result vector to <out>.<app>.bin.
//
// graph file: u32 n, u64 m, u32 undirected, u32 weighted, u32 row_start[n+1],
//             u32 edge_dst[m], u32 weights[m] (when weighted)
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
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s graph.bin out_prefix source\n", argv[0]);
        return 2;
    }
    graphs::CSRGraph g;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t und = 0, weighted = 0;
    bool ok = std::fread(&g.nnodes, 4, 1, f) == 1 && std::fread(&g.nedges, 8, 1, f) == 1 &&
              std::fread(&und, 4, 1, f) == 1 && std::fread(&weighted, 4, 1, f) == 1;
    g.undirected = und != 0;
    g.row_start.resize(g.nnodes + 1);
    g.edge_dst.resize(g.nedges);
    ok = ok && std::fread(g.row_start.data(), 4, g.nnodes + 1, f) == g.nnodes + 1 &&
         std::fread(g.edge_dst.data(), 4, g.nedges, f) == g.nedges;
    if (weighted) {
        g.edge_weights.resize(g.nedges);
        ok = ok && std::fread(g.edge_weights.data(), 4, g.nedges, f) == g.nedges;
    }
    std::fclose(f);
    if (!ok) return 2;
    const uint32_t source = (uint32_t)std::strtoul(argv[3], nullptr, 10);
    const std::string out = argv[2];
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
