// Hybrid (direction-switching) BFS as one persistent kernel per query: the
// kernel, its launch and the C ABI entry point (sg_bfs).
// The traversal itself is bfs_run in bfs_dev.hpp.
#include "bfs_dev.hpp"

using namespace bfsk;

namespace {

template <int BS>
__global__ SG_BFS_BOUNDS(BS) void k_bfs(BfsArgs a)
{
    if (a.probe) {
        residency_probe(a.bar.words, 1u << 20);
        return;
    }
    const uint32_t xcc = xcc_id();
    BarState gen = bar_begin(a.bar, xcc);
    bfs_run<BS>(a, blockIdx.x, gridDim.x, xcc, gen);
}


template <int BS>
int launch_bfs(sg_engine *e, BfsArgs &args, bool timed)
{
    sg_engine::PGrid *pg = nullptr;
    auto probe = [e, args](int blocks, uint32_t *words) -> int {
        BfsArgs p = args;
        p.probe = 1;
        p.bar.words = words;
        hipLaunchKernelGGL(k_bfs<BS>, dim3(blocks), dim3(BS), 0, e->stream, p);
        SG_HIP(hipGetLastError());
        return SG_OK;
    };
    int rc = engine_persistent_grid(e, (const void *)k_bfs<BS>, BS, probe, &pg);
    if (rc) return rc;
    bind_barrier(pg, args.bar);
    return engine_launch_persistent(e, [&]() -> int {
        if (timed) SG_HIP(hipEventRecord(e->ev0, e->stream));
        hipLaunchKernelGGL(k_bfs<BS>, dim3(pg->blocks), dim3(BS), 0, e->stream, args);
        SG_HIP(hipGetLastError());
        pg->launches++;
        if (timed) SG_HIP(hipEventRecord(e->ev1, e->stream));
        return SG_OK;
    });
}

}  // namespace

namespace sg {

// One BFS query from the caller's vertex `source` writing int32 levels, in the
// caller's vertex order, to the device row `out` (n entries). Asynchronous;
// with `traced`, the run also fills the engine's trace and kernel timing.
int bfs_launch(sg_engine *e, uint32_t source, int32_t *out, bool traced)
{
    const DevGraph &d = e->g->dev;
    if (source >= d.n || !out) SG_FAIL(SG_ERR_RANGE, "sg_bfs: source >= n");
    const sg_engine_options &o = e->opt;
    const int force = forced_direction(o);
    if (force < -1) SG_FAIL(SG_ERR_UNSUPPORTED, "sg_bfs: only data-driven push/pull variants apply to BFS");
    int rc = engine_ensure_bfs_state(e);
    if (rc) return rc;
    BfsArgs a{};
    a.g = d;
    a.level = out;
    a.perm = e->g->relabeled ? e->g->d_perm : nullptr;
    // state left clean by the previous BFS unless another kernel ran since
    // (each persistent launch advances launch_seq) or an abort reset it
    const bool clean = e->bfs_token != 0 && e->bfs_token == e->launch_seq && !e->needs_bar_reset;
    a.full_init = clean ? 0u : 1u;
    a.vis = e->bitmaps + (size_t)(e->bfs_parity & 1) * d.words;
    a.vis_idle = e->bitmaps + (size_t)((e->bfs_parity & 1) ^ 1) * d.words;
    a.fbit = e->bitmaps + 2 * (size_t)d.words;
    a.queue = e->queues;
    a.qcap = e->qcap;
    a.ctr = (FrontierCtr *)e->ctr;
    a.bar.spin_limit = 1u << 22;
    a.abort = e->abort_flag;
    a.stats = e->stats;
    a.trace = traced ? e->trace : nullptr;
    a.source = dev_vertex(e->g, source);
    a.source_id = source;
    a.force = force;
    a.alpha = o.alpha == 0.0 ? 1.0 : o.alpha;
    a.beta = o.beta;
    a.total_workload0 = (uint32_t)((double)d.m * o.edge_factor);
    a.max_rounds = (uint32_t)o.max_iteration;
    a.lb_push_none = o.lb_push == SG_LB_NONE;
    a.lb_pull_none = o.lb_pull == SG_LB_NONE;
    e->last_algo = 0;
    const int bs = o.block_size ? o.block_size : 1024;
    switch (bs) {
    case 256: rc = launch_bfs<256>(e, a, traced); break;
    case 512: rc = launch_bfs<512>(e, a, traced); break;
    case 1024: rc = launch_bfs<1024>(e, a, traced); break;
    default: SG_FAIL(SG_ERR_UNSUPPORTED, "sg_bfs: block_size must be 256, 512 or 1024");
    }
    if (rc) return rc;
    e->bfs_token = e->launch_seq;
    e->bfs_parity ^= 1;
    if (traced) e->timed = true;
    return SG_OK;
}

}  // namespace sg

extern "C" int sg_bfs(sg_engine *e, uint32_t source)
{
    if (!e) SG_FAIL(SG_ERR_INVALID, "sg_bfs: null engine");
    const DevGraph &d = e->g->dev;
    if (source >= d.n) SG_FAIL(SG_ERR_RANGE, "sg_bfs: source >= n");
    int rc = engine_ensure_value(e, (uint64_t)d.n * 4);
    if (rc) return rc;
    return bfs_launch(e, source, (int32_t *)e->value, true);
}

