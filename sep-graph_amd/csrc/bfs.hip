// Hybrid (direction-switching) BFS as one persistent kernel per query: the
// kernel, its launch and the C ABI entry points (sg_bfs, sg_bfs_noalloc).
// The traversal itself is bfs_run in bfs_dev.hpp.
#include "bfs_dev.hpp"

using namespace bfsk;

namespace {

template <int BS, bool kStage>
__global__ SG_BFS_BOUNDS(BS) void k_bfs(BfsArgs a)
{
    if (a.probe) {
        residency_probe(a.bar.words, 1u << 20);
        return;
    }
    const uint32_t xcc = xcc_id();
    BarState gen = bar_begin(a.bar, xcc);
    bfs_run<BS, kStage>(a, blockIdx.x, gridDim.x, xcc, gen);
}


template <int BS, bool kStage>
int launch_bfs(sg_engine *e, BfsArgs &args, bool timed)
{
    sg_engine::PGrid *pg = nullptr;
    auto probe = [e, args](int blocks, uint32_t *words) -> int {
        BfsArgs p = args;
        p.probe = 1;
        p.bar.words = words;
        hipLaunchKernelGGL((k_bfs<BS, kStage>), dim3(blocks), dim3(BS), 0, e->stream, p);
        SG_HIP(hipGetLastError());
        return SG_OK;
    };
    int rc = engine_persistent_grid(e, (const void *)k_bfs<BS, kStage>, BS, probe, &pg);
    if (rc) return rc;
    bind_barrier(pg, args.bar);
    return engine_launch_persistent(e, [&]() -> int {
        if (timed) SG_HIP(hipEventRecord(e->ev0, e->stream));
        hipLaunchKernelGGL((k_bfs<BS, kStage>), dim3(pg->blocks), dim3(BS), 0, e->stream, args);
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
    case 256: rc = o.stage_hubs ? launch_bfs<256, true>(e, a, traced) : launch_bfs<256, false>(e, a, traced); break;
    case 512: rc = o.stage_hubs ? launch_bfs<512, true>(e, a, traced) : launch_bfs<512, false>(e, a, traced); break;
    case 1024: rc = o.stage_hubs ? launch_bfs<1024, true>(e, a, traced) : launch_bfs<1024, false>(e, a, traced); break;
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

// ------------------------------------------------ allocation-free BFS
// sg_bfs_noalloc: the same kernel as sg_bfs, for callers that cannot allocate
// device memory. Its working state -- visited filters, frontier bitmaps and
// queues, round counters, grid barrier words, abort flag and run status --
// lives in `scratch`, a file-scope device array of fixed size that each GPU
// holds while the library is loaded. Host code never passes the array's name
// to a kernel: the array's device address comes from hipGetSymbolAddress,
// taken once per device with that device current, and only that address is
// used. Calls from one process are serialised per device (the call holds a
// lock until its levels are written); the array keeps the last call's
// contents, so every call resets what it reads.
namespace {

// Largest graph the array is sized for, and the queue length per round
// (a longer frontier marks the queue overflowed and the next round reads the
// frontier bitmap instead).
constexpr uint32_t kNaMaxVertices = 1u << 23;
constexpr uint32_t kNaWords = kNaMaxVertices / 64;
constexpr uint32_t kNaQcap = 1u << 20;
constexpr size_t na_align(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNaOffBitmaps = 0;  // 5 x kNaWords: 2 visited filters, 3 frontier bitmaps
constexpr size_t kNaOffQueues = kNaOffBitmaps + na_align(5ull * kNaWords * 8);
constexpr size_t kNaOffCtr = kNaOffQueues + na_align(3ull * kNaQcap * 4);
constexpr size_t kNaOffProbe = kNaOffCtr + na_align(kCtrBytes);   // residency probe's barrier words
constexpr size_t kNaOffBar = kNaOffProbe + na_align(kBarWords * 4);  // the run's barrier words
constexpr size_t kNaOffAbort = kNaOffBar + na_align(kBarWords * 4);
constexpr size_t kNaOffStats = kNaOffAbort + 256;
constexpr size_t kNaBytes = kNaOffStats + na_align(sizeof(DevStats));
static_assert(kNaBytes < (18ull << 20), "allocation-free BFS scratch stays under 18 MiB");

}  // namespace

static __device__ __align__(256) uint64_t scratch[kNaBytes / 8];

namespace {

constexpr int kNaBlock = 1024;
constexpr int kNaMaxDevices = 64;

struct NaDevice {
    char *ptr = nullptr;   // device address of `scratch` on this device
    int blocks = 0;        // co-resident grid of k_bfs<kNaBlock, false>
    uint64_t launches = 0; // runs on the barrier words (registration parity)
};
std::mutex g_na_mu[kNaMaxDevices];
NaDevice g_na[kNaMaxDevices];

__global__ void k_na_zero(uint32_t *p, uint32_t words)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x) p[i] = 0u;
}

int na_zero(hipStream_t s, char *p, size_t bytes)
{
    hipLaunchKernelGGL(k_na_zero, dim3(64), dim3(256), 0, s, (uint32_t *)p, (uint32_t)(bytes / 4));
    SG_HIP(hipGetLastError());
    return SG_OK;
}

// Once per device: the array's address, cleared counters, and the largest
// blocks-per-CU grid of k_bfs<kNaBlock, false> whose blocks are all resident.
int na_setup(NaDevice &d, const sg_graph *g, hipStream_t s, const BfsArgs &proto)
{
    void *p = nullptr;
    SG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) SG_FAIL(SG_ERR_HIP, "sg_bfs_noalloc: no device address for the scratch array");
    d.ptr = (char *)p;
    int rc = na_zero(s, d.ptr + kNaOffCtr, kNaBytes - kNaOffCtr);
    if (rc) return rc;
    int per_cu = 0;
    SG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_bfs<kNaBlock, false>, kNaBlock, 0));
    if (per_cu < 1) SG_FAIL(SG_ERR_HIP, "sg_bfs_noalloc: kernel cannot be resident");
    uint32_t *words = (uint32_t *)(d.ptr + kNaOffProbe);
    for (int pc = per_cu; pc >= 1; pc--) {
        const int blocks = pc * g->num_cus;
        if ((rc = na_zero(s, (char *)words, kBarWords * 4))) return rc;
        BfsArgs pa = proto;
        pa.probe = 1;
        pa.bar.words = words;
        hipLaunchKernelGGL((k_bfs<kNaBlock, false>), dim3(blocks), dim3(kNaBlock), 0, s, pa);
        SG_HIP(hipGetLastError());
        uint32_t counts[kMaxXcd + 2];
        SG_HIP(hipMemcpyAsync(counts, words + bar_count(0), sizeof(counts), hipMemcpyDeviceToHost, s));
        SG_HIP(hipStreamSynchronize(s));
        uint32_t total = 0;
        for (int x = 0; x < kMaxXcd; x++) total += counts[x];
        if (counts[kMaxXcd + 1] == 0 && total == (uint32_t)blocks) {
            d.blocks = blocks;
            return SG_OK;
        }
    }
    SG_FAIL(SG_ERR_HIP, "sg_bfs_noalloc: one block per CU is not co-resident");
}

}  // namespace

extern "C" int sg_bfs_noalloc(const sg_graph *g, uint32_t source, int32_t *levels, void *hip_stream,
                              const sg_engine_options *opt)
{
    if (!g || !levels) SG_FAIL(SG_ERR_INVALID, "sg_bfs_noalloc: null argument");
    const DevGraph &dg = g->dev;
    if (source >= dg.n) SG_FAIL(SG_ERR_RANGE, "sg_bfs_noalloc: source >= n");
    if (dg.n > kNaMaxVertices)
        SG_FAIL(SG_ERR_CAPACITY, "sg_bfs_noalloc: graph has more than 2^23 vertices (the scratch array's limit)");
    if (g->device < 0 || g->device >= kNaMaxDevices) SG_FAIL(SG_ERR_RANGE, "sg_bfs_noalloc: device id out of range");
    sg_engine_options o;
    if (opt) o = *opt;
    else sg_engine_options_default(&o);
    const int force = forced_direction(o);
    if (force < -1) SG_FAIL(SG_ERR_UNSUPPORTED, "sg_bfs_noalloc: only data-driven push/pull variants apply to BFS");
    if (o.block_size != 0 && o.block_size != kNaBlock)
        SG_FAIL(SG_ERR_UNSUPPORTED, "sg_bfs_noalloc: block_size must be 0 or 1024");
    if (o.grid_blocks != 0) SG_FAIL(SG_ERR_UNSUPPORTED, "sg_bfs_noalloc: grid_blocks must be 0");
    if (o.alpha < 0 || o.beta <= 0) SG_FAIL(SG_ERR_INVALID, "sg_bfs_noalloc: need alpha >= 0, beta > 0");
    const hipStream_t s = (hipStream_t)hip_stream;

    std::lock_guard<std::mutex> dlk(g_na_mu[g->device]);
    SG_HIP(hipSetDevice(g->device));
    NaDevice &d = g_na[g->device];
    std::lock_guard<std::mutex> plk(persistent_mutex());
    int rc = persistent_wait_other(s);
    if (rc) return rc;

    BfsArgs a{};
    a.g = dg;
    a.level = levels;
    a.perm = g->relabeled ? g->d_perm : nullptr;
    a.full_init = 1;  // the array's contents are not trusted from call to call
    a.qcap = kNaQcap;
    a.bar.spin_limit = 1u << 22;
    a.trace = nullptr;
    a.source = dev_vertex(g, source);
    a.source_id = source;
    a.force = force;
    a.alpha = o.alpha == 0.0 ? 1.0 : o.alpha;
    a.beta = o.beta;
    a.total_workload0 = (uint32_t)((double)dg.m * o.edge_factor);
    a.max_rounds = (uint32_t)o.max_iteration;
    a.lb_push_none = o.lb_push == SG_LB_NONE;
    a.lb_pull_none = o.lb_pull == SG_LB_NONE;
    if (!d.ptr && (rc = na_setup(d, g, s, a))) return rc;

    uint64_t *bm = (uint64_t *)(d.ptr + kNaOffBitmaps);
    a.vis = bm;
    a.vis_idle = bm + dg.words;
    a.fbit = bm + 2 * (size_t)dg.words;
    a.queue = (uint32_t *)(d.ptr + kNaOffQueues);
    a.ctr = (FrontierCtr *)(d.ptr + kNaOffCtr);
    a.bar.words = (uint32_t *)(d.ptr + kNaOffBar);
    a.bar.rset = (uint32_t)(d.launches & 1u);
    a.abort = (uint32_t *)(d.ptr + kNaOffAbort);
    a.stats = (DevStats *)(d.ptr + kNaOffStats);
    hipLaunchKernelGGL((k_bfs<kNaBlock, false>), dim3(d.blocks), dim3(kNaBlock), 0, s, a);
    SG_HIP(hipGetLastError());
    d.launches++;
    int32_t status = 0;
    SG_HIP(hipMemcpyAsync(&status, (const char *)a.stats + offsetof(DevStats, status), 4, hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    if (status != SG_OK) {
        // an expired wait leaves barrier counters mid-way and the abort flag
        // set: clear both (and the status) before the next call
        if ((rc = na_zero(s, d.ptr + kNaOffBar, kNaBytes - kNaOffBar))) return rc;
        SG_HIP(hipStreamSynchronize(s));
        d.launches = 0;
        SG_FAIL(status, "sg_bfs_noalloc: device: grid barrier wait expired");
    }
    return SG_OK;
}
