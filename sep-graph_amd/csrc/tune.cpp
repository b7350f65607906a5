// Autotuner (C ABI sg_autotune): picks, for one resident graph, the
// direction-switch thresholds and block size of the single-query traversals
// and the sources-per-launch of each batched query kind, by timing the real
// kernels on sample sources.
//
// The reference fixes alpha = 15, beta = 18 (app_skeleton.h:50-51) and the
// block size (--block_size 256, traversal.h:73-82) for every graph; its
// ITERATIVE scheme only predicts per-round times from sampled rounds
// (hybrid_policy.h:130-169). Here the choice is made per graph by
// measurement, with results bit-identical for every candidate (the direction
// rule and block shape change the schedule, never the levels or distances).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "../../include/sepgraph.h"

namespace {

// Mean kernel time (ms) of `run` over the sample sources; < 0 on error.
template <typename F>
double mean_ms(sg_engine *e, const uint32_t *src, int k, F run)
{
    double tot = 0.0;
    for (int i = 0; i < k; i++) {
        if (run(src[i]) != SG_OK) return -1.0;
        sg_stats st;
        if (sg_engine_get_stats(e, &st) != SG_OK || st.status != SG_OK) return -1.0;
        tot += st.kernel_ms;
    }
    return tot / k;
}

}  // namespace

extern "C" void sg_tune_config_default(sg_tune_config *c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->latency_budget_ms = 25.0;
    c->tune_bfs = 1;
    c->tune_sssp = 1;
    c->tune_batches = 1;
}

extern "C" int sg_autotune(sg_engine *e, const uint32_t *sources, int32_t k, const sg_tune_config *cfg,
                           sg_tune_result *out)
{
    if (!e || !sources || !out || k < 1) return SG_ERR_INVALID;
    sg_tune_config c;
    if (cfg) c = *cfg;
    else sg_tune_config_default(&c);
    std::memset(out, 0, sizeof(*out));
    sg_engine_options base;
    sg_engine_options_default(&base);
    const double alphas[] = {4, 8, 15, 30, 60};
    const double betas[] = {9, 18, 36};
    const int blocks[] = {256, 512, 1024};
    int rc = SG_OK;

    if (c.tune_bfs) {
        double best = -1.0;
        for (int bs : blocks)
            for (double a : alphas)
                for (double b : betas) {
                    sg_engine_options o = base;
                    o.alpha = a;
                    o.beta = b;
                    o.block_size = bs;
                    if ((rc = sg_engine_set_options(e, &o))) return rc;
                    const double t = mean_ms(e, sources, k, [&](uint32_t s) { return sg_bfs(e, s); });
                    if (t < 0) return SG_ERR_HIP;
                    if (best < 0 || t < best) {
                        best = t;
                        out->bfs_alpha = a;
                        out->bfs_beta = b;
                        out->bfs_block_size = bs;
                    }
                    if (a == 15 && b == 18 && bs == 1024) out->bfs_default_ms = t;
                }
        out->bfs_ms = best;
    }
    if (c.tune_bfs) {
        // then the grid size for the best BFS configuration
        for (int gb : {128, 32}) {
            sg_engine_options o = base;
            o.alpha = out->bfs_alpha;
            o.beta = out->bfs_beta;
            o.block_size = out->bfs_block_size;
            o.grid_blocks = gb;
            if ((rc = sg_engine_set_options(e, &o))) return rc;
            const double t = mean_ms(e, sources, k, [&](uint32_t s) { return sg_bfs(e, s); });
            if (t < 0) return SG_ERR_HIP;
            if (t < out->bfs_ms) {
                out->bfs_ms = t;
                out->bfs_grid_blocks = gb;
            }
        }
    }
    if (c.tune_sssp) {
        // SSSP queries cost ~20x a BFS: a staged search on up to 4 sources
        const int ks = std::min(k, 4);
        auto time_sssp = [&](const sg_engine_options &o) -> double {
            if (sg_engine_set_options(e, &o) != SG_OK) return -1.0;
            return mean_ms(e, sources, ks, [&](uint32_t s) { return sg_sssp(e, s); });
        };
        double best = -1.0;
        sg_engine_options bo = base;
        for (int prio : {SG_PRIO_NONE, SG_PRIO_LOW_HIGH})
            for (int bs : {512, 1024}) {
                sg_engine_options o = base;
                o.priority = prio;
                o.block_size = bs;
                const double t = time_sssp(o);
                if (t < 0) return SG_ERR_HIP;
                if (prio == SG_PRIO_NONE && bs == 1024) out->sssp_default_ms = t;
                if (best < 0 || t < best) {
                    best = t;
                    bo = o;
                }
            }
        if (bo.priority == SG_PRIO_NONE)
            for (double a : {4.0, 30.0}) {
                sg_engine_options o = bo;
                o.alpha = a;
                const double t = time_sssp(o);
                if (t < 0) return SG_ERR_HIP;
                if (t < best) {
                    best = t;
                    bo = o;
                }
            }
        for (int gb : {128, 32}) {
            sg_engine_options o = bo;
            o.grid_blocks = gb;
            const double t = time_sssp(o);
            if (t < 0) return SG_ERR_HIP;
            if (t < best) {
                best = t;
                bo = o;
            }
        }
        out->sssp_alpha = bo.alpha;
        out->sssp_block_size = bo.block_size;
        out->sssp_priority = bo.priority;
        out->sssp_grid_blocks = bo.grid_blocks;
        out->sssp_ms = best;
    }
    if (c.tune_batches) {
        // per kind: the batch size with the highest queries/s whose launch
        // stays within the latency budget (the smallest batch if none does)
        if ((rc = sg_engine_set_options(e, &base))) return rc;
        const int sizes[] = {4, 8, 16, 32, 64};
        std::vector<uint32_t> src(64);
        for (int i = 0; i < 64; i++) src[i] = sources[i % k];
        for (int kind = 0; kind < SG_QUERY_KINDS; kind++) {
            double best_qps = -1.0;
            if (kind == SG_QUERY_SSSP) {
                // the server runs an SSSP batch below 8 queries as whole-grid
                // queries one after another (throughput: one query's; the
                // limit bounds how long the batch holds the GPU), and from 8
                // queries as one bit-packed launch; the best queries/s within
                // the budget wins
                const double t = mean_ms(e, sources, std::min(k, 4), [&](uint32_t s) { return sg_sssp(e, s); });
                if (t < 0) return SG_ERR_HIP;
                int b = 1;
                while (b < 4 && 2 * b * t <= c.latency_budget_ms) b *= 2;
                out->max_batch[kind] = b;
                out->batch_ms[kind] = b * t;
                out->batch_qps[kind] = t > 0 ? 1e3 / t : 0.0;
                for (int bb : {8, 16, 32, 64}) {
                    if ((rc = sg_sssp_multi(e, src.data(), bb, bb))) return rc;  // warm
                    if ((rc = sg_sssp_multi(e, src.data(), bb, bb))) return rc;
                    sg_stats st;
                    if ((rc = sg_engine_get_stats(e, &st))) return rc;
                    if (st.kernel_ms > c.latency_budget_ms) break;
                    const double qps = bb / (st.kernel_ms / 1e3);
                    if (qps > out->batch_qps[kind]) {
                        out->max_batch[kind] = bb;
                        out->batch_ms[kind] = st.kernel_ms;
                        out->batch_qps[kind] = qps;
                    }
                }
                continue;
            }
            for (int b : sizes) {
                auto run = [&]() {
                    return kind == SG_QUERY_BFS ? sg_bfs_multi(e, src.data(), b, b)
                                                : sg_bc_dependency(e, src.data(), b, b);
                };
                if ((rc = run())) return rc;  // warm (grid probe, arena)
                if ((rc = run())) return rc;
                sg_stats st;
                if ((rc = sg_engine_get_stats(e, &st))) return rc;
                const double ms = st.kernel_ms;
                const double qps = b / (ms / 1e3);
                const bool fits = ms <= c.latency_budget_ms;
                if ((fits && qps > best_qps) || best_qps < 0) {
                    best_qps = qps;
                    out->max_batch[kind] = b;
                    out->batch_ms[kind] = ms;
                    out->batch_qps[kind] = qps;
                }
                if (!fits) break;
            }
        }
    }
    // leave the engine on the tuned BFS configuration
    sg_engine_options o = base;
    if (c.tune_bfs) {
        o.alpha = out->bfs_alpha;
        o.beta = out->bfs_beta;
        o.block_size = out->bfs_block_size;
        o.grid_blocks = out->bfs_grid_blocks;
    }
    return sg_engine_set_options(e, &o);
}
