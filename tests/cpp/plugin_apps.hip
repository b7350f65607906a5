// Apps written against the reference's plugin API (sepgraph::api::AppBase,
// include/framework/variants/api.cuh:22-120), run through the facade's Engine
// on the generic engine of include/sepgraph_app.hpp:
//   bfs   -- the hooks of samples/hybrid_bfs/hybrid_bfs.cu:29-170 (levels;
//            push schedules improved vertices, pull stops at the first parent
//            with ACCUMULATE_SUCCESS_BREAK; atomic or non-atomic min)
//   sssp  -- the hooks of samples/hybrid_sssp/hybrid_sssp.cu:30-140 (weighted
//            AccumulateBuffer)
//   cc    -- an app the library has no kernel for: connected components by
//            minimum-label propagation on an undirected graph
//
// usage: plugin_apps app graph.gr out.bin source undirected(0|1) variant(auto|push|pull)
// Writes the app's value array (uint32 per vertex) to out.bin and prints the
// round counts.
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sepgraph_app.hpp"

#define IDENTITY_ELEMENT UINT_MAX

namespace hybrid_bfs {
template <typename TValue, typename TBuffer, typename TWeight, typename... UnusedData>
struct BFS : sepgraph::api::AppBase<TValue, TBuffer, TWeight> {
    using sepgraph::api::AppBase<TValue, TBuffer, TWeight>::AccumulateBuffer;
    index_t m_source_node;
    bool m_non_atomic;

    BFS(index_t source, bool non_atomic) : m_source_node(source), m_non_atomic(non_atomic) {}

    __forceinline__ __device__ TValue GetInitValue(index_t node) const override { return (TValue)IDENTITY_ELEMENT; }

    __forceinline__ __device__ TBuffer GetInitBuffer(index_t node) const override
    {
        return node == m_source_node ? (TBuffer)0 : (TBuffer)UINT32_MAX;
    }

    __forceinline__ __host__ __device__ TBuffer GetIdentityElement() const override { return IDENTITY_ELEMENT; }

    __forceinline__ __device__ utils::pair<TBuffer, bool> CombineValueBuffer(index_t node, TValue *p_value,
                                                                             TBuffer *p_buffer) override
    {
        TBuffer b = *p_buffer;
        const bool improved = *p_value > b;
        if (improved) {
            *p_value = b;
            b += 1;
        }
        if (this->m_msg_passing == sepgraph::common::MsgPassing::PUSH) return utils::pair<TBuffer, bool>(b, improved);
        // pull: only vertices not reached yet look for a parent
        return utils::pair<TBuffer, bool>(b, !improved && *p_buffer == IDENTITY_ELEMENT);
    }

    __forceinline__ __device__ int AccumulateBuffer(index_t src, index_t dst, TBuffer *p_buffer,
                                                    TBuffer buffer) override
    {
        TBuffer prev;
        if (m_non_atomic) {
            prev = *p_buffer;
            if (prev > buffer) *p_buffer = buffer;
        } else {
            prev = atomicMin(p_buffer, buffer);
        }
        const bool lowered = buffer < prev;
        if (this->m_msg_passing == sepgraph::common::MsgPassing::PUSH)
            return lowered ? this->ACCUMULATE_SUCCESS_CONTINUE : this->ACCUMULATE_FAILURE_CONTINUE;
        return lowered ? this->ACCUMULATE_SUCCESS_BREAK : this->ACCUMULATE_FAILURE_CONTINUE;
    }

    __forceinline__ __device__ bool IsActiveNode(index_t node, TBuffer buffer) const override
    {
        return buffer != IDENTITY_ELEMENT;
    }
};
}  // namespace hybrid_bfs

namespace hybrid_sssp {
template <typename TValue, typename TBuffer, typename TWeight, typename... UnusedData>
struct SSSP : sepgraph::api::AppBase<TValue, TBuffer, TWeight> {
    using sepgraph::api::AppBase<TValue, TBuffer, TWeight>::AccumulateBuffer;
    index_t m_source_node;

    explicit SSSP(index_t source_node) : m_source_node(source_node) {}

    __forceinline__ __device__ TValue GetInitValue(index_t node) const override { return (TValue)IDENTITY_ELEMENT; }
    __forceinline__ __device__ TBuffer GetInitBuffer(index_t node) const override
    {
        return node == m_source_node ? (TBuffer)0 : (TBuffer)IDENTITY_ELEMENT;
    }
    __forceinline__ __host__ __device__ TBuffer GetIdentityElement() const override { return IDENTITY_ELEMENT; }

    __forceinline__ __device__ utils::pair<TBuffer, bool> CombineValueBuffer(index_t node, TValue *p_value,
                                                                             TBuffer *p_buffer) override
    {
        const TBuffer b = *p_buffer;
        const bool improved = *p_value > b;
        if (improved) *p_value = b;
        // push: expand improved vertices; pull: every vertex looks again
        const bool pull = this->m_msg_passing == sepgraph::common::MsgPassing::PULL;
        return utils::pair<TBuffer, bool>(b, pull || improved);
    }

    __forceinline__ __device__ int AccumulateBuffer(index_t src, index_t dst, TWeight weight, TBuffer *p_buffer,
                                                    TBuffer buffer) override
    {
        if (buffer == IDENTITY_ELEMENT) return this->ACCUMULATE_FAILURE_CONTINUE;
        const TBuffer nd = buffer + weight;
        return nd < atomicMin(p_buffer, nd) ? this->ACCUMULATE_SUCCESS_CONTINUE : this->ACCUMULATE_FAILURE_CONTINUE;
    }

    __forceinline__ __device__ bool IsActiveNode(index_t node, TBuffer buffer) const override
    {
        return buffer != IDENTITY_ELEMENT;
    }

    __forceinline__ __device__ bool IsHighPriority(TBuffer current_priority, TBuffer buffer) const override
    {
        return current_priority > buffer;
    }
};
}  // namespace hybrid_sssp

namespace cc {
// Each vertex ends with the smallest id of its connected component.
template <typename TValue, typename TBuffer, typename TWeight, typename... UnusedData>
struct MinLabel : sepgraph::api::AppBase<TValue, TBuffer, TWeight> {
    using sepgraph::api::AppBase<TValue, TBuffer, TWeight>::AccumulateBuffer;
    MinLabel() {}

    __device__ TValue GetInitValue(index_t node) const override { return (TValue)IDENTITY_ELEMENT; }
    __device__ TBuffer GetInitBuffer(index_t node) const override { return (TBuffer)node; }
    __host__ __device__ TBuffer GetIdentityElement() const override { return IDENTITY_ELEMENT; }

    __device__ utils::pair<TBuffer, bool> CombineValueBuffer(index_t node, TValue *p_value, TBuffer *p_buffer) override
    {
        const TBuffer b = *p_buffer;
        const bool improved = *p_value > b;
        if (improved) *p_value = b;
        return utils::pair<TBuffer, bool>(b, improved || this->m_msg_passing == sepgraph::common::MsgPassing::PULL);
    }

    __device__ int AccumulateBuffer(index_t src, index_t dst, TBuffer *p_buffer, TBuffer buffer) override
    {
        return buffer < atomicMin(p_buffer, buffer) ? this->ACCUMULATE_SUCCESS_CONTINUE
                                                    : this->ACCUMULATE_FAILURE_CONTINUE;
    }

    __device__ bool IsActiveNode(index_t node, TBuffer buffer) const override { return true; }
};
}  // namespace cc

using namespace sepgraph;

template <typename T>
static int dump(const char *path, const std::vector<T> &v)
{
    FILE *f = std::fopen(path, "wb");
    if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) return 3;
    std::fclose(f);
    return 0;
}

static common::EngineOptions options(const std::string &variant)
{
    common::EngineOptions opt;
    if (variant == "push") opt.ForceVariant(common::AlgoVariant::ASYNC_PUSH_DD);
    else if (variant == "pull") opt.ForceVariant(common::AlgoVariant::ASYNC_PULL_DD);
    return opt;
}

template <typename E>
static int report(E &engine, const char *out)
{
    const sg_stats &s = engine.Stats();
    std::printf("rounds %u push %u pull %u\n", s.rounds, s.push_rounds, s.pull_rounds);
    return dump(out, engine.GatherValue());
}

int main(int argc, char **argv)
{
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s bfs|sssp|cc graph.gr out.bin source undirected(0|1) auto|push|pull\n", argv[0]);
        return 2;
    }
    const std::string app = argv[1], variant = argv[6];
    const index_t source = (index_t)std::strtoul(argv[4], nullptr, 10);
    try {
        graphs::CSRGraph g = graphs::ReadGraph(argv[2], "gr", 0, std::atoi(argv[5]) != 0);
        common::EngineOptions opt = options(variant);
        if (app == "bfs") {
            engine::Engine<uint32_t, uint32_t, graphs::NoWeight, hybrid_bfs::BFS, index_t, bool> e(
                policy::AlgoType::TRAVERSAL_SCHEME);
            e.LoadGraph(g);
            e.SetOptions(opt);
            e.InitGraph(source, false);
            e.Start();
            return report(e, argv[3]);
        }
        if (app == "sssp") {
            engine::Engine<uint32_t, uint32_t, uint32_t, hybrid_sssp::SSSP, index_t> e(
                policy::AlgoType::TRAVERSAL_SCHEME);
            e.LoadGraph(g);
            e.SetOptions(opt);
            e.InitGraph(source);
            e.Start();
            return report(e, argv[3]);
        }
        if (app == "cc") {
            engine::Engine<uint32_t, uint32_t, graphs::NoWeight, cc::MinLabel> e(policy::AlgoType::TRAVERSAL_SCHEME);
            e.LoadGraph(g);
            e.SetOptions(opt);
            e.InitGraph();
            e.Start();
            return report(e, argv[3]);
        }
    } catch (const Error &err) {
        std::fprintf(stderr, "%s (code %d)\n", err.what(), err.code());
        return 1;
    }
    std::fprintf(stderr, "unknown app %s\n", app.c_str());
    return 2;
}
