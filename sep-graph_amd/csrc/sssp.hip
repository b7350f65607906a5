// Hybrid (direction-switching) single-source shortest paths as one
// persistent kernel per query (or one query per XCD): the kernel, its
// launches and the C ABI entry points. The traversal itself is sssp_run in
// sssp_dev.hpp.
#include "sssp_dev.hpp"

using namespace ssspk;

namespace {

// One query over the whole grid, or (a.grouped) one query per XCD: the
// blocks an XCD holds in this launch form that query's grid, ranked by their
// registration order, with a barrier of their own. Concurrent single-source
// queries each keep their working set in one XCD's L2 and synchronise only
// within it, which suits SSSP, whose rounds are latency-bound at far below
// the bandwidth of the chip.
template <int BS>
__global__ __launch_bounds__(BS) void k_sssp(SsspArgs a)
{
    if (a.probe) {
        residency_probe(a.bar.words, 1u << 20);
        return;
    }
    const uint32_t xcc = xcc_id();
    BarState st = bar_begin(a.bar, xcc);
    if (!a.grouped) {
        sssp_run<BS>(a, blockIdx.x, gridDim.x, xcc, st);
        return;
    }
    // Grouped: the census of this launch decides which XCD runs which query.
    __shared__ uint32_t s_q, s_rank, s_cnt, s_ok;
    if (threadIdx.x == 0) {
        uint32_t all[kMaxXcd];
        const bool ok = bar_census(a.bar, xcc, st, a.abort, all);
        uint32_t q = 0, held = 0;
        for (uint32_t x = 0; x < (uint32_t)kMaxXcd; x++) {
            q += (x < xcc && all[x] != 0u) ? 1u : 0u;
            held += all[x] != 0u ? 1u : 0u;
        }
        // queries no XCD holds fail the call; the first block of the
        // first XCD holding blocks reports them
        if (ok && held < a.gr.nq && q == 0 && st.rank == 0) atomicMin(a.gr.status, (int32_t)SG_ERR_CAPACITY);
        if (!ok) atomicMin(a.gr.status, (int32_t)SG_ERR_TIMEOUT);
        s_q = q;
        s_rank = st.rank;
        s_cnt = st.cnt;
        s_ok = ok;
    }
    __syncthreads();
    const uint32_t q = s_q;
    if (!s_ok || q >= a.gr.nq) return;  // no query for this XCD in this launch
    const size_t W = a.g.words;
    SsspArgs b = a;
    b.dist = a.gr.out[q];
    b.fbit = a.fbit + (size_t)q * 3 * W;
    b.queue = a.queue + (size_t)q * 3 * a.qcap;
    b.ctr = a.ctr + (size_t)q * 3;
    b.stats = a.stats + q;
    b.trace = nullptr;
    b.source = a.gr.source[q];
    // the XCD's own barrier area (after the whole-grid one), whose census is
    // the one just taken: all s_cnt blocks of the XCD, on one XCD
    b.bar.words = a.bar.words + (size_t)(xcc + 1) * kBarWords;
    BarState xs{0u, s_cnt, 1u, s_rank};
    if (threadIdx.x == 0) xs.gen = poll_load(b.bar.words + kBarGen);
    sssp_run<BS>(b, s_rank, s_cnt, xcc, xs);
    if (s_rank == 0 && threadIdx.x == 0) atomicMin(a.gr.status, b.stats->status);
}

// Persistent grid for k_sssp<BS>; grouped launches take the barrier area
// `group_words` ((1 + kMaxXcd) areas) and give the XCDs holding blocks one
// query each, in XCD order.
template <int BS>
int launch_sssp(sg_engine *e, SsspArgs &args, uint32_t *group_words)
{
    if (!args.abort || !args.stats || (!args.dist && !args.grouped) || !args.fbit || !args.queue ||
        !args.ctr || (args.grouped && (!group_words || !args.gr.status)))
        SG_FAIL(SG_ERR_STATE, "sg_sssp: launch with unset device state");
    sg_engine::PGrid *pg = nullptr;
    auto probe = [e, args](int blocks, uint32_t *words) -> int {
        SsspArgs p = args;
        p.probe = 1;
        p.grouped = 0;
        p.bar.words = words;
        hipLaunchKernelGGL(k_sssp<BS>, dim3(blocks), dim3(BS), 0, e->stream, p);
        SG_HIP(hipGetLastError());
        return SG_OK;
    };
    int rc = engine_persistent_grid(e, (const void *)k_sssp<BS>, BS, probe, &pg);
    if (rc) return rc;
    if (args.grouped) {
        args.bar.words = group_words;
        args.bar.rset = (uint32_t)(e->group_launches & 1u);
    } else {
        bind_barrier(pg, args.bar);
    }
    return engine_launch_persistent(e, [&]() -> int {
        hipLaunchKernelGGL(k_sssp<BS>, dim3(pg->blocks), dim3(BS), 0, e->stream, args);
        SG_HIP(hipGetLastError());
        if (args.grouped) e->group_launches++;
        else pg->launches++;
        return SG_OK;
    });
}

int launch_sssp_bs(sg_engine *e, SsspArgs &a, uint32_t *group_words)
{
    const int bs = e->opt.block_size ? e->opt.block_size : 1024;
    switch (bs) {
    case 256: return launch_sssp<256>(e, a, group_words);
    case 512: return launch_sssp<512>(e, a, group_words);
    case 1024: return launch_sssp<1024>(e, a, group_words);
    default: SG_FAIL(SG_ERR_UNSUPPORTED, "sg_sssp: block_size must be 256, 512 or 1024");
    }
}

// Arguments shared by every SSSP launch of an engine (graph, options).
int sssp_args(sg_engine *e, SsspArgs &a)
{
    const DevGraph &d = e->g->dev;
    const sg_engine_options &o = e->opt;
    const int force = forced_direction(o);
    if (force < -1) SG_FAIL(SG_ERR_UNSUPPORTED, "sg_sssp: only data-driven push/pull variants apply to SSSP");
    a = SsspArgs{};
    a.g = d;
    a.bar.spin_limit = 1u << 22;
    a.abort = e->abort_flag;
    a.force = force;
    a.alpha = o.alpha == 0.0 ? 1.0 : o.alpha;
    a.beta = o.beta;
    a.total_workload0 = (uint32_t)((double)d.m * o.edge_factor);
    a.max_rounds = (uint32_t)o.max_iteration;
    a.lb_push_none = o.lb_push == SG_LB_NONE;
    a.lb_pull_none = o.lb_pull == SG_LB_NONE;
    if (o.priority == SG_PRIO_LOW_HIGH) {
        // two-level priority (the reference's SetTwoLevelBasedPriority,
        // algo_variants.cuh:605-685) runs push rounds only; the bucket width
        // is the option's threshold, or c * mean weight / mean degree with c
        // the wave width (hybrid_sssp.cu:183-199 uses the 32-lane warp)
        double dl = o.priority_threshold;
        if (!(dl >= 1.0)) {
            const double avg_deg = d.n ? (double)d.m / d.n : 1.0;
            const double avg_w = e->g->weighted ? e->g->avg_weight : 1.0;
            dl = avg_deg > 0 ? kWave * avg_w / avg_deg : 1.0;
        }
        a.delta = (uint32_t)std::max(1.0, std::min(dl, 2147483647.0));
        a.force = 0;
    }
    return SG_OK;
}

// Working state of concurrent SSSP (one query per XCD): per-query bitmaps,
// queues, counters and stats, the barrier areas (the whole grid's, then one
// per XCD) and the call's status word. Offsets into the engine's garena.
struct GroupLayout {
    uint64_t fbit, queue, ctr, stats, bar, status, total;
};

GroupLayout group_layout(const sg_engine *e)
{
    const uint64_t W = e->g->dev.words, qc = e->qcap;
    auto al = [](uint64_t x) { return (x + 255) / 256 * 256; };
    GroupLayout L;
    L.fbit = 0;
    L.queue = L.fbit + al(kMaxXcd * 3 * W * 8);
    L.ctr = L.queue + al(kMaxXcd * 3 * qc * 4);
    L.stats = L.ctr + al(kMaxXcd * 3 * sizeof(FrontierCtr));
    L.bar = L.stats + al(kMaxXcd * sizeof(DevStats));
    L.status = L.bar + al((uint64_t)(1 + kMaxXcd) * kBarWords * 4);
    L.total = L.status + 256;
    return L;
}

__global__ void k_set_status(int32_t *status, int32_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *status = v;
}

}  // namespace

namespace sg {

// Size the concurrent-SSSP arena and set the call's status word to SG_OK
// (by a kernel, in order on the stream). Every launch of the call folds the
// status of each of its queries into that word, so a failure in any launch
// of a multi-launch call stays visible until the next call.
int sssp_group_begin(sg_engine *e)
{
    int rc = engine_ensure_bfs_state(e);
    if (rc) return rc;
    const GroupLayout L = group_layout(e);
    if (!e->garena || e->garena_bytes < L.total) {
        if (e->garena) {
            SG_HIP(hipStreamSynchronize(e->stream));
            (void)hipFree(e->garena);
            e->garena = nullptr;
            e->garena_bytes = 0;
        }
        if ((rc = engine_alloc(&e->garena, L.total))) return rc;
        e->garena_bytes = L.total;
        // barrier counters start at zero; the engine re-zeroes them after an
        // aborted run (engine_persistent_grid)
        SG_HIP(hipMemsetAsync((char *)e->garena + L.bar, 0, (size_t)(L.status - L.bar), e->stream));
        e->garena_bar_off = L.bar;
        e->garena_bar_bytes = L.status - L.bar;
        e->group_launches = 0;
    }
    e->group_status = (int32_t *)((char *)e->garena + L.status);
    hipLaunchKernelGGL(k_set_status, dim3(1), dim3(64), 0, e->stream, e->group_status, (int32_t)SG_OK);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

// Up to one SSSP query per XCD in one launch; query i writes distances to
// out[i] (device rows of n). The XCDs holding blocks in the launch take the
// queries in XCD order; queries beyond their number fail the call with
// SG_ERR_CAPACITY. Asynchronous; sssp_group_begin must have run first.
int sssp_group_launch(sg_engine *e, const uint32_t *sources, int nq, uint32_t *const *out)
{
    if (nq < 1 || nq > kMaxXcd) SG_FAIL(SG_ERR_RANGE, "concurrent SSSP: 1..16 queries per launch");
    if (!e->garena || !e->group_status) SG_FAIL(SG_ERR_STATE, "concurrent SSSP: arena not prepared");
    const DevGraph &d = e->g->dev;
    for (int i = 0; i < nq; i++)
        if (sources[i] >= d.n) SG_FAIL(SG_ERR_RANGE, "concurrent SSSP: source >= n");
    int rc;
    const GroupLayout L = group_layout(e);
    char *base = (char *)e->garena;
    SsspArgs a;
    if ((rc = sssp_args(e, a))) return rc;
    a.fbit = (uint64_t *)(base + L.fbit);
    a.queue = (uint32_t *)(base + L.queue);
    a.qcap = e->qcap;
    a.ctr = (FrontierCtr *)(base + L.ctr);
    a.stats = (DevStats *)(base + L.stats);
    a.grouped = 1;
    a.gr.status = e->group_status;
    a.gr.nq = (uint32_t)nq;
    for (int i = 0; i < nq; i++) {
        a.gr.source[i] = sources[i];
        a.gr.out[i] = out[i];
    }
    e->last_algo = 10;  // concurrent SSSP (status in group_status)
    return launch_sssp_bs(e, a, (uint32_t *)(base + L.bar));
}

// One SSSP query over the whole grid from device vertex `source`, with its
// distances in the device row `dist` (device vertex order) and, for a
// relabelled graph, gathered into the row `out` in the caller's order.
// Asynchronous; its status lands in the engine's stats, overwritten by the
// next run (see engine_run_status).
int sssp_single_launch(sg_engine *e, uint32_t source, uint32_t *dist, uint32_t *out)
{
    if (source >= e->g->dev.n || !dist || (e->g->relabeled && !out))
        SG_FAIL(SG_ERR_RANGE, "sssp_single_launch: bad source or row");
    int rc = engine_ensure_bfs_state(e);
    if (rc) return rc;
    SsspArgs a;
    if ((rc = sssp_args(e, a))) return rc;
    a.dist = dist;
    if (e->g->relabeled) {
        a.out = out;
        a.inv = e->g->d_inv;
    }
    a.fbit = e->bitmaps;
    a.queue = e->queues;
    a.qcap = e->qcap;
    a.ctr = (FrontierCtr *)e->ctr;
    a.stats = e->stats;
    a.trace = nullptr;
    a.source = source;
    e->last_algo = 1;
    return launch_sssp_bs(e, a, nullptr);
}

// Worst status over the queries of the last concurrent-SSSP call
// (synchronises the stream).
int sssp_group_status(sg_engine *e)
{
    if (!e->group_status) return SG_OK;
    SG_HIP(hipStreamSynchronize(e->stream));
    int32_t st = 0;
    SG_HIP(hipMemcpy(&st, e->group_status, 4, hipMemcpyDeviceToHost));
    if (st != SG_OK) {
        e->needs_bar_reset = true;
        SG_FAIL(st, st == SG_ERR_CAPACITY ? "concurrent SSSP: more queries than XCDs holding blocks"
                                          : "concurrent SSSP: a query reported an error");
    }
    return SG_OK;
}

}  // namespace sg

extern "C" int sg_sssp(sg_engine *e, uint32_t source)
{
    if (!e) SG_FAIL(SG_ERR_INVALID, "sg_sssp: null engine");
    const DevGraph &d = e->g->dev;
    if (source >= d.n) SG_FAIL(SG_ERR_RANGE, "sg_sssp: source >= n");
    // engine state first: sssp_args copies its device pointers
    int rc = engine_ensure_bfs_state(e);
    if (rc) return rc;
    SsspArgs a;
    if ((rc = sssp_args(e, a))) return rc;
    void *res = nullptr;
    if ((rc = engine_result_rows(e, 1, &res))) return rc;
    a.dist = (uint32_t *)res;
    a.fbit = e->bitmaps;  // three rotating frontier bitmaps
    a.queue = e->queues;
    a.qcap = e->qcap;
    a.ctr = (FrontierCtr *)e->ctr;
    a.stats = e->stats;
    a.trace = e->trace;
    a.source = dev_vertex(e->g, source);
    if (e->g->relabeled) {
        // the kernel gathers the result into the value array itself
        a.out = (uint32_t *)e->value;
        a.inv = e->g->d_inv;
    }
    e->last_algo = 1;
    SG_HIP(hipEventRecord(e->ev0, e->stream));
    if ((rc = launch_sssp_bs(e, a, nullptr))) return rc;
    SG_HIP(hipEventRecord(e->ev1, e->stream));
    e->timed = true;
    return SG_OK;
}

extern "C" int sg_sssp_concurrent(sg_engine *e, const uint32_t *sources, int32_t k)
{
    if (!e || !sources) SG_FAIL(SG_ERR_INVALID, "sg_sssp_concurrent: null argument");
    if (k < 1) SG_FAIL(SG_ERR_RANGE, "sg_sssp_concurrent: need at least one source");
    const DevGraph &d = e->g->dev;
    for (int32_t i = 0; i < k; i++)
        if (sources[i] >= d.n) SG_FAIL(SG_ERR_RANGE, "sg_sssp_concurrent: source >= n");
    int rc = engine_ensure_bfs_state(e);
    if (rc) return rc;
    void *res = nullptr;
    if ((rc = engine_result_rows(e, (uint64_t)k, &res))) return rc;
    std::vector<uint32_t> dsrc(k);
    for (int32_t i = 0; i < k; i++) dsrc[i] = dev_vertex(e->g, sources[i]);
    // one query per XCD: MI355X has 8 (a launch on fewer XCDs reports
    // SG_ERR_CAPACITY)
    const int per = 8;
    if ((rc = sssp_group_begin(e))) return rc;
    SG_HIP(hipEventRecord(e->ev0, e->stream));
    for (int32_t i = 0; i < k; i += per) {
        const int nq = std::min<int32_t>(per, k - i);
        uint32_t *out[kMaxXcd];
        for (int j = 0; j < nq; j++) out[j] = (uint32_t *)res + (size_t)(i + j) * d.n;
        if ((rc = sssp_group_launch(e, dsrc.data() + i, nq, out))) return rc;
    }
    SG_HIP(hipEventRecord(e->ev1, e->stream));
    if ((rc = engine_finish(e, (uint64_t)k))) return rc;
    e->timed = true;
    return SG_OK;
}
