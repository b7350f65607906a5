// Single-source Brandes BC as one persistent kernel per source: the kernel,
// its working arena and its launch. The traversal is bc1_run in bc1_dev.hpp.
// sg_bc and sg_bc_dependency run it when called with one source per launch
// (the reference's form, samples/hybrid_bc/hybrid_bc.cu:303-411).
#include "bc1_dev.hpp"

using namespace bc1k;

namespace {

// Per-launch run summary kept beside the BFS's own, folded into the engine's
// stats the way the batched BC kernel reports (sums over sources).
struct Bc1KArgs {
    Bc1Args b;
    DevStats *acc;   // the engine's stats: rounds / depth / iterations over the call
    uint32_t probe;
};

template <int BS>
__global__ __launch_bounds__(BS) void k_bc1(Bc1KArgs a)
{
    if (a.probe) {
        residency_probe(a.b.f.bar.words, 1u << 20);
        return;
    }
    const uint32_t xcc = xcc_id();
    BarState gen = bar_begin(a.b.f.bar, xcc);
    bc1_run<BS>(a.b, blockIdx.x, gridDim.x, xcc, gen);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const DevStats f = *a.b.f.stats;  // written by this thread
        DevStats s = a.b.first ? DevStats{} : *a.acc;
        s.rounds += f.rounds;
        s.push_rounds += f.push_rounds;
        s.pull_rounds += f.pull_rounds;
        s.trace = f.trace;
        s.max_depth = max(s.max_depth, f.max_depth);
        s.iterations += 1;
        if (s.status == SG_OK) s.status = f.status;
        *a.acc = s;
    }
}

template <int BS>
int launch_bc1(sg_engine *e, Bc1KArgs &a)
{
    sg_engine::PGrid *pg = nullptr;
    auto probe = [e, a](int blocks, uint32_t *words) -> int {
        Bc1KArgs p = a;
        p.probe = 1;
        p.b.f.bar.words = words;
        hipLaunchKernelGGL(k_bc1<BS>, dim3(blocks), dim3(BS), 0, e->stream, p);
        SG_HIP(hipGetLastError());
        return SG_OK;
    };
    int rc = engine_persistent_grid(e, (const void *)k_bc1<BS>, BS, probe, &pg);
    if (rc) return rc;
    bind_barrier(pg, a.b.f.bar);
    return engine_launch_persistent(e, [&]() -> int {
        hipLaunchKernelGGL(k_bc1<BS>, dim3(pg->blocks), dim3(BS), 0, e->stream, a);
        SG_HIP(hipGetLastError());
        pg->launches++;
        return SG_OK;
    });
}

}  // namespace

namespace sg {

// One source (device id) over the whole grid. `dep` (optional) receives
// 0.5 delta in device order (or at dep[perm[v]] with perm); with `bcout`,
// the FP64 sum in the arena is started (first) or extended, and the last
// source writes it to bcout as FP32 in device order. Asynchronous.
int bc1_launch(sg_engine *e, uint32_t dev_source, float *dep, const uint32_t *perm, float *bcout, bool first,
               bool last)
{
    int rc = engine_ensure_bfs_state(e);
    if (rc) return rc;
    const Bc1Layout L = bc1_layout(e->g->dev.n);
    if ((rc = engine_ensure_aux(e, L.total))) return rc;
    char *base = (char *)e->aux;
    Bc1KArgs a{};
    if ((rc = bc1_fill(e, base, L, dev_source, a.b))) return rc;
    a.b.f.vis = e->bitmaps;
    a.b.f.vis_idle = e->bitmaps + e->g->dev.words;
    a.b.f.fbit = e->bitmaps + 2 * (size_t)e->g->dev.words;
    a.b.f.queue = e->queues;
    a.b.f.ctr = (FrontierCtr *)e->ctr;
    a.b.dep = dep;
    a.b.perm = perm;
    if (bcout) {
        a.b.bcsum = (double *)(base + L.bcsum);
        a.b.bcout = bcout;
    }
    a.b.first = first;
    a.b.last = last;
    a.acc = e->stats;
    const int bs = e->opt.block_size ? e->opt.block_size : 1024;
    switch (bs) {
    case 256: return launch_bc1<256>(e, a);
    case 512: return launch_bc1<512>(e, a);
    case 1024: return launch_bc1<1024>(e, a);
    default: SG_FAIL(SG_ERR_UNSUPPORTED, "BC: block_size must be 256, 512 or 1024");
    }
}

}  // namespace sg
