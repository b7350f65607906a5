// The serving kernel: one persistent launch in which every XCD of the GPU is
// a query slot. No reference counterpart (the reference answers one query
// per process, samples/hybrid_*/main.cpp).
//
// The host publishes queries into a ring in pinned host memory (kind,
// source, result row) and advances a tail counter. In the kernel, the blocks
// an XCD holds form one slot: the slot's leader claims the next ring entry
// (compare-and-swap on a head counter in device memory), the slot's blocks
// run that query -- BFS (bfs_run), SSSP (sssp_run) or a single-source BC
// dependency (bc1_run) -- over their own state arrays with a barrier local
// to the XCD, the leader publishes completion (a 64-bit word per ring entry
// in host memory, after a system-scope release), and the slot takes the next
// entry. So queries that arrive while others run are admitted as soon as a
// slot frees up, and up to one query per XCD runs at a time, each keeping
// its working set in its XCD's L2. Slots leave when the host asks them to
// and the ring is drained, when every slot of the launch has found no work
// for a stretch (the host relaunches the kernel when more arrives), or when
// any bounded wait expires. A slot never leaves alone while others work: a
// launch whose idle slots had left would serve later arrivals with fewer
// slots until its last query ended.
#include "bc1_dev.hpp"
#include "serve.hpp"
#include "sssp_dev.hpp"

#include <atomic>
#include <cstring>
#include <memory>

namespace {

constexpr uint32_t kExit = 0xFFFFFFFFu;
// Frontier bitmap words an SSSP pull round keeps in LDS in the serving
// kernel (whose LDS also holds the BFS and BC kernels' buffers).
constexpr int kServeBlock = 1024;
constexpr uint32_t kServeFront = 2u * kServeBlock;

// Per-slot mailbox in device memory: the leader fills it before the slot's
// barrier, every block of the slot reads it after.
struct alignas(64) Mail {
    uint32_t seq;       // claimed ring entry, kExit to leave
    uint32_t graph, kind, source, source_id, row;
    uint32_t full_init, parity;
};

// Per-slot BFS bookkeeping kept across queries and launches: the slot's BFS
// state is clean for graph `graph` (frontier bitmaps and counters zeroed,
// idle filter reset to that graph's inert bits) after a BFS (or a BC forward
// BFS) on it ended normally.
struct alignas(64) SlotCtl {
    uint32_t clean, parity, graph;
};

struct ServeArgs {
    DevGraph g[sg::kServeGraphs];
    const uint32_t *perm[sg::kServeGraphs];  // device vertex -> caller's id (relabelled graphs), else null
    const uint32_t *inv[sg::kServeGraphs];   // caller's id -> device vertex, else null
    uint32_t total_wl[sg::kServeGraphs];     // the direction rule's workload (edge count x edge factor)
    uint32_t delta[sg::kServeGraphs];        // SSSP two-level bucket width (0: off)
    char *pool;             // result rows in the caller's vertex order, row_bytes apart
    uint64_t row_bytes;
    const sg::ServeEntry *ring;
    const uint32_t *tail;   // host memory: entries published
    const uint32_t *stop;   // host memory: leave once the ring is drained
    unsigned long long *done;  // host memory: per entry, (seq + 1) | status << 32
    uint32_t *head;         // entries claimed
    uint32_t *idle;         // [0]: slots voting to end this launch, [1]: the launch is closing (zeroed per launch)
    Mail *mail;             // kMaxXcd
    SlotCtl *ctl;           // kMaxXcd
    char *slots;            // one arena of sg::ServeLayout per slot
    int32_t slot_of[kMaxXcd];  // slot of each XCD (-1: the XCD's blocks leave)
    sg::ServeLayout L;
    uint32_t qcap;
    GridBarrier bar;        // the launch's census area, then one area per XCD
    uint32_t claim_limit;   // polls a slot's blocks wait for its leader's claim
    uint32_t *abort;
    // options
    double alpha, beta;
    int32_t force;
    unsigned long long idle_ticks;
    uint32_t probe;
};

__device__ __forceinline__ uint32_t sys_load(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The slot leader waits for a ring entry and claims it; kExit when the
// host asked to stop and nothing is left, after an abort, or when the launch
// is closing. A slot that has found nothing for the idle limit votes to
// close; the vote that makes every one of the launch's `nlive` slots idle
// closes the launch. A slot withdraws its vote before claiming, so a claim
// can only race with the closing vote, and a query claimed then still runs.
__device__ Mail claim_next(const ServeArgs &a, uint32_t xcc, uint32_t nlive)
{
    Mail m{};
    m.seq = kExit;
    const unsigned long long t0 = wall_clock64();
    bool voted = false;
    for (;;) {
        if (poll_load(a.abort) || poll_load(a.idle + 1)) return m;
        const uint32_t t = sys_load(a.tail);
        uint32_t h = poll_load(a.head);
        if ((int32_t)(t - h) > 0) {
            if (voted) {
                atomicSub(a.idle, 1u);
                voted = false;
            }
            if (atomicCAS(a.head, h, h + 1u) != h) continue;
            const sg::ServeEntry *en = a.ring + (h % sg::kServeRing);
            m.seq = h;
            m.graph = sys_load(&en->graph);
            m.kind = sys_load(&en->kind);
            m.source = sys_load(&en->source);
            m.source_id = sys_load(&en->source_id);
            m.row = sys_load(&en->row);
            const SlotCtl c = a.ctl[xcc];
            m.full_init = c.clean && c.graph == m.graph ? 0u : 1u;
            m.parity = c.parity & 1u;
            return m;
        }
        if (sys_load(a.stop)) return m;
        if (!voted && wall_clock64() - t0 > a.idle_ticks) {
            voted = true;
            if (atomicAdd(a.idle, 1u) + 1u >= nlive) {
                flag_store(a.idle + 1, 1u);
                return m;
            }
        }
        if (voted && poll_load(a.idle) >= nlive) {
            flag_store(a.idle + 1, 1u);
            return m;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// The BFS arguments of a slot (levels written by the caller).
__device__ __forceinline__ bfsk::BfsArgs slot_bfs(const ServeArgs &a, char *base, const Mail &m, uint32_t xcc)
{
    const uint32_t W = a.g[m.graph].words;
    uint64_t *bm = (uint64_t *)(base + a.L.bitmaps);
    bfsk::BfsArgs b{};
    b.g = a.g[m.graph];
    b.vis = bm + (size_t)m.parity * W;
    b.vis_idle = bm + (size_t)(m.parity ^ 1u) * W;
    b.fbit = bm + 2 * (size_t)W;
    b.queue = (uint32_t *)(base + a.L.queue);
    b.qcap = a.qcap;
    b.ctr = (FrontierCtr *)(base + a.L.ctr);
    b.bar = a.bar;
    b.bar.words = a.bar.words + (size_t)(xcc + 1) * kBarWords;
    b.abort = a.abort;
    b.stats = (DevStats *)(base + a.L.stats);
    b.trace = nullptr;
    b.source = m.source;
    b.source_id = m.source_id;
    b.force = a.force;
    b.alpha = a.alpha;
    b.beta = a.beta;
    b.total_workload0 = a.total_wl[m.graph];
    b.max_rounds = 0;
    b.full_init = m.full_init;
    return b;
}

// One query of each kind on a slot (the slot's blocks, its barrier area and
// arena). Not inlined: each keeps the register budget it has alone.
template <int BS>
__device__ __noinline__ bool serve_bfs(const ServeArgs &a, char *base, const Mail m, uint32_t brank, uint32_t nb,
                                       uint32_t xcc, BarState &gen)
{
    bfsk::BfsArgs b = slot_bfs(a, base, m, xcc);
    b.level = (int32_t *)(a.pool + (size_t)m.row * a.row_bytes);
    b.perm = a.perm[m.graph];
    return bfsk::bfs_run<BS>(b, brank, nb, xcc, gen);
}

template <int BS>
__device__ __noinline__ bool serve_sssp(const ServeArgs &a, char *base, const Mail m, uint32_t brank, uint32_t nb,
                                        uint32_t xcc, BarState &gen)
{
    ssspk::SsspArgs s{};
    s.g = a.g[m.graph];
    uint32_t *row = (uint32_t *)(a.pool + (size_t)m.row * a.row_bytes);
    if (a.perm[m.graph]) {
        // device-order distances, gathered into the row at the end
        s.dist = (uint32_t *)(base + a.L.dist);
        s.out = row;
        s.inv = a.inv[m.graph];
    } else {
        s.dist = row;
    }
    s.fbit = (uint64_t *)(base + a.L.sbitmaps);
    s.queue = (uint32_t *)(base + a.L.squeue);
    s.qcap = a.qcap;
    s.ctr = (FrontierCtr *)(base + a.L.sctr);
    s.bar = a.bar;
    s.bar.words = a.bar.words + (size_t)(xcc + 1) * kBarWords;
    s.abort = a.abort;
    s.stats = (DevStats *)(base + a.L.stats);
    s.source = m.source;
    s.force = a.delta[m.graph] ? 0 : a.force;
    s.alpha = a.alpha;
    s.beta = a.beta;
    s.total_workload0 = a.total_wl[m.graph];
    s.delta = a.delta[m.graph];
    return ssspk::sssp_run<BS, kServeFront>(s, brank, nb, xcc, gen);
}

template <int BS>
__device__ __noinline__ bool serve_bc(const ServeArgs &a, char *base, const Mail m, uint32_t brank, uint32_t nb,
                                      uint32_t xcc, BarState &gen)
{
    bc1k::Bc1Args c{};
    c.f = slot_bfs(a, base, m, xcc);
    char *bb = base + a.L.bc;
    c.f.level = (int32_t *)(bb + a.L.bc1.lvl);
    c.f.perm = nullptr;
    c.f.source_id = m.source;  // levels stay in device order
    c.sa = (double2 *)(bb + a.L.bc1.sa);
    c.order = (uint32_t *)(bb + a.L.bc1.order);
    c.lcnt = (uint32_t *)(bb + a.L.bc1.lcnt);
    c.loff = (uint32_t *)(bb + a.L.bc1.loff);
    c.lfill = (uint32_t *)(bb + a.L.bc1.lfill);
    c.ctl = (uint32_t *)(bb + a.L.bc1.ctl);
    c.lvol = (uint32_t *)(bb + a.L.bc1.lvol);
    c.lbm = (uint64_t *)(bb + a.L.bc1.lbm);
    c.dep = (float *)(a.pool + (size_t)m.row * a.row_bytes);
    c.perm = a.perm[m.graph];
    return bc1k::bc1_run<BS>(c, brank, nb, xcc, gen);
}

template <int BS>
__global__ __launch_bounds__(BS) void k_serve(ServeArgs a)
{
    if (a.probe) {
        residency_probe(a.bar.words, 1u << 20);
        return;
    }
    const uint32_t xcc = xcc_id();
    BarState st = bar_begin(a.bar, xcc);
    __shared__ uint32_t s_rank, s_cnt, s_ok, s_live;
    __shared__ Mail s_mail;
    if (threadIdx.x == 0) {
        uint32_t all[kMaxXcd];
        const bool ok = bar_census(a.bar, xcc, st, a.abort, all);
        // slots of this launch: XCDs holding blocks that have an arena
        uint32_t live = 0;
        for (int x = 0; x < kMaxXcd; x++) live += (ok && all[x] != 0u && a.slot_of[x] >= 0) ? 1u : 0u;
        s_rank = st.rank;
        s_cnt = st.cnt;
        s_ok = ok;
        s_live = live;
    }
    __syncthreads();
    if (!s_ok) return;
    // the blocks of this XCD form its slot, ranked by registration order,
    // with the XCD's own barrier area
    // values read from LDS are uniform: keep them in scalar registers
    const uint32_t brank = __builtin_amdgcn_readfirstlane(s_rank), nb = __builtin_amdgcn_readfirstlane(s_cnt);
    GridBarrier xb = a.bar;
    xb.words = a.bar.words + (size_t)(xcc + 1) * kBarWords;
    BarState gen{0u, nb, 1u, brank};
    if (threadIdx.x == 0) gen.gen = poll_load(xb.words + kBarGen);
    if (a.slot_of[xcc] < 0) return;  // an XCD the residency probe found empty holds no slot
    char *base = a.slots + (size_t)a.slot_of[xcc] * a.L.total;
    DevStats *stats = (DevStats *)(base + a.L.stats);
    const bool leader = brank == 0 && threadIdx.x == 0;
    // The slot's other blocks wait for the leader's claim as long as no slot
    // has work for it, i.e. as long as the longest query running in another
    // slot: that wait gets a bound far above any query (the leader itself
    // always returns: on a claim, a stop, an abort or the idle vote).
    GridBarrier xclaim = xb;
    xclaim.spin_limit = a.claim_limit;
    for (;;) {
        if (leader) a.mail[xcc] = claim_next(a, xcc, s_live);
        if (!grid_sync(xclaim, xcc, gen, a.abort)) break;
        if (threadIdx.x == 0) s_mail = a.mail[xcc];
        __syncthreads();
        Mail m;
        m.seq = __builtin_amdgcn_readfirstlane(s_mail.seq);
        m.graph = __builtin_amdgcn_readfirstlane(s_mail.graph);
        m.kind = __builtin_amdgcn_readfirstlane(s_mail.kind);
        m.source = __builtin_amdgcn_readfirstlane(s_mail.source);
        m.source_id = __builtin_amdgcn_readfirstlane(s_mail.source_id);
        m.row = __builtin_amdgcn_readfirstlane(s_mail.row);
        m.full_init = __builtin_amdgcn_readfirstlane(s_mail.full_init);
        m.parity = __builtin_amdgcn_readfirstlane(s_mail.parity);
        __syncthreads();
        if (m.seq == kExit) break;
        // each kind in a function of its own, so each gets the registers
        // it needs alone (inlined together, the three spill)
        const bool ok = m.kind == SG_QUERY_BFS    ? serve_bfs<BS>(a, base, m, brank, nb, xcc, gen)
                        : m.kind == SG_QUERY_SSSP ? serve_sssp<BS>(a, base, m, brank, nb, xcc, gen)
                                                  : serve_bc<BS>(a, base, m, brank, nb, xcc, gen);
        if (!ok) break;
        // every block's writes of the result are done before it is published
        if (!grid_sync(xb, xcc, gen, a.abort)) break;
        if (leader) {
            const int32_t status = stats->status;  // written by this thread
            if (m.kind != SG_QUERY_SSSP && status == SG_OK) {
                // the BFS left its state clean with the filters swapped
                a.ctl[xcc].clean = 1u;
                a.ctl[xcc].parity = m.parity ^ 1u;
                a.ctl[xcc].graph = m.graph;
            } else if (m.kind != SG_QUERY_SSSP) {
                a.ctl[xcc].clean = 0u;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");  // system scope: the row reaches memory first
            const unsigned long long word =
                (unsigned long long)(m.seq + 1u) | ((unsigned long long)(uint32_t)status << 32);
            __hip_atomic_store(&a.done[m.seq % sg::kServeRing], word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

__global__ void k_serve_zero(uint32_t *p, uint64_t words)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

}  // namespace

namespace sg {

ServeLayout serve_layout(uint32_t n, uint32_t qcap)
{
    ServeLayout L{};
    const uint64_t W = (n + 63) / 64;
    uint64_t o = 0;
    auto put = [&](uint64_t &off, uint64_t bytes) {
        off = o;
        o += (bytes + 255) / 256 * 256;
    };
    put(L.bitmaps, 5 * W * 8);
    put(L.queue, 3ull * qcap * 4);
    put(L.ctr, 3 * sizeof(FrontierCtr));
    put(L.sbitmaps, 3 * W * 8);
    put(L.squeue, 3ull * qcap * 4);
    put(L.sctr, 3 * sizeof(FrontierCtr));
    put(L.stats, sizeof(DevStats));
    put(L.dist, (uint64_t)n * 4);
    L.bc1 = bc1k::bc1_layout(n);
    put(L.bc, L.bc1.total);
    L.total = o;
    return L;
}

struct ServeDev {
    sg_graph *g = nullptr;
    hipStream_t stream = nullptr;
    ServeArgs a{};
    // host-visible control
    ServeEntry *ring = nullptr;
    uint32_t *tail = nullptr, *stop = nullptr;
    unsigned long long *done = nullptr;
    // device control and state
    char *dev = nullptr;
    uint64_t dev_bytes = 0;
    uint64_t off_bar = 0, bar_bytes = 0;
    int blocks = 0;
    uint64_t launches = 0;
    bool running = false;
    uint32_t published = 0;
    int nslots = 0;
    void *arenas = nullptr;
};

namespace {

int serve_zero(ServeDev *d, void *p, uint64_t bytes)
{
    hipLaunchKernelGGL(k_serve_zero, dim3(256), dim3(256), 0, d->stream, (uint32_t *)p, bytes / 4);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int size_serve_grid(ServeDev *d)
{
    int per_cu = 0;
    SG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_serve<kServeBlock>, kServeBlock, 0));
    if (per_cu < 1) SG_FAIL(SG_ERR_HIP, "serving kernel cannot be resident");
    uint32_t *words = (uint32_t *)(d->dev + d->off_bar);
    for (int pc = per_cu; pc >= 1; pc--) {
        const int blocks = pc * d->g->num_cus;
        int rc = serve_zero(d, words, kBarWords * 4);
        if (rc) return rc;
        ServeArgs p = d->a;
        p.probe = 1;
        p.bar.words = words;
        rc = engine_launch_persistent_on(d->stream, [&]() -> int {
            hipLaunchKernelGGL(k_serve<kServeBlock>, dim3(blocks), dim3(kServeBlock), 0, d->stream, p);
            SG_HIP(hipGetLastError());
            return SG_OK;
        });
        if (rc) return rc;
        uint32_t counts[kMaxXcd + 2];
        SG_HIP(hipMemcpyAsync(counts, words + bar_count(0), sizeof(counts), hipMemcpyDeviceToHost, d->stream));
        SG_HIP(hipStreamSynchronize(d->stream));
        uint32_t total = 0;
        for (int x = 0; x < kMaxXcd; x++) total += counts[x];
        if (counts[kMaxXcd + 1] == 0 && total == (uint32_t)blocks) {
            d->blocks = blocks;
            // one slot per XCD that held blocks
            d->nslots = 0;
            for (int x = 0; x < kMaxXcd; x++) d->a.slot_of[x] = counts[x] ? d->nslots++ : -1;
            // the probe's counters are not the barrier's: clear the whole area
            return serve_zero(d, d->dev + d->off_bar, d->bar_bytes);
        }
    }
    SG_FAIL(SG_ERR_HIP, "serving kernel: one block per CU is not co-resident");
}

}  // namespace

int serve_create(sg_graph *const *graphs, int ngraphs, const sg_engine_options &o, uint32_t prio_graphs,
                 hipStream_t stream, void *pool, uint64_t row_bytes, ServeDev **out)
{
    *out = nullptr;
    if (ngraphs < 1 || ngraphs > kServeGraphs) SG_FAIL(SG_ERR_RANGE, "server: 1..4 graphs");
    std::unique_ptr<ServeDev> d(new ServeDev());
    sg_graph *g = graphs[0];
    d->g = g;
    d->stream = stream;
    const int force = forced_direction(o);
    if (force < -1) SG_FAIL(SG_ERR_UNSUPPORTED, "server: only data-driven push/pull variants apply");
    if (o.lb_push == SG_LB_NONE || o.lb_pull == SG_LB_NONE)
        SG_FAIL(SG_ERR_UNSUPPORTED, "server: LoadBalancing::NONE applies to sg_bfs and sg_sssp only");
    // slot arenas are sized for the largest graph
    uint32_t nmax = 0;
    for (int i = 0; i < ngraphs; i++) {
        if (graphs[i]->device != g->device) SG_FAIL(SG_ERR_INVALID, "server: graphs on different devices");
        nmax = std::max(nmax, graphs[i]->dev.n);
    }
    if (row_bytes < (uint64_t)nmax * 4) SG_FAIL(SG_ERR_INVALID, "server: result rows shorter than a graph");
    // a frontier queue of n/2 entries per round: a longer one marks the queue
    // overflowed and the next round reads the frontier bitmap instead
    const uint32_t qcap = (uint32_t)std::max<uint64_t>(4096, (uint64_t)nmax / 2);
    const ServeLayout L = serve_layout(nmax, qcap);
    const uint64_t o_head = 0, o_mail = 256, o_ctl = o_mail + kMaxXcd * sizeof(Mail), o_abort = o_ctl + kMaxXcd * sizeof(SlotCtl);
    const uint64_t o_bar = (o_abort + 256 + 255) / 256 * 256;
    const uint64_t bar_bytes = (uint64_t)(1 + kMaxXcd) * kBarWords * 4;
    d->dev_bytes = o_bar + bar_bytes;
    int rc = engine_alloc((void **)&d->dev, d->dev_bytes);
    if (rc) return rc;
    d->off_bar = o_bar;
    d->bar_bytes = bar_bytes;
    if (hipHostMalloc((void **)&d->ring, kServeRing * sizeof(ServeEntry), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&d->done, kServeRing * 8, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&d->tail, 256, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        serve_destroy(d.release());
        SG_FAIL(SG_ERR_NOMEM, "server: pinned ring allocation failed");
    }
    std::memset(d->ring, 0, kServeRing * sizeof(ServeEntry));
    std::memset((void *)d->done, 0, kServeRing * 8);
    std::memset(d->tail, 0, 256);
    d->stop = d->tail + 16;  // its own cache line
    // the control words and barrier areas start at zero
    if ((rc = serve_zero(d.get(), d->dev, d->dev_bytes))) {
        serve_destroy(d.release());
        return rc;
    }
    ServeArgs &a = d->a;
    for (int i = 0; i < ngraphs; i++) {
        const sg_graph *gi = graphs[i];
        const DevGraph &dg = gi->dev;
        a.g[i] = dg;
        a.perm[i] = gi->relabeled ? gi->d_perm : nullptr;
        a.inv[i] = gi->relabeled ? gi->d_inv : nullptr;
        a.total_wl[i] = (uint32_t)((double)dg.m * o.edge_factor);
        a.delta[i] = 0;
        if (o.priority == SG_PRIO_LOW_HIGH || ((prio_graphs >> i) & 1u)) {
            // the SSSP two-level bucket width, as for sg_sssp
            double dl = o.priority_threshold;
            if (!(dl >= 1.0)) {
                const double avg_deg = dg.n ? (double)dg.m / dg.n : 1.0;
                const double avg_w = gi->weighted ? gi->avg_weight : 1.0;
                dl = avg_deg > 0 ? kWave * avg_w / avg_deg : 1.0;
            }
            a.delta[i] = (uint32_t)std::max(1.0, std::min(dl, 2147483647.0));
        }
    }
    a.pool = (char *)pool;
    a.row_bytes = row_bytes;
    a.ring = d->ring;
    a.tail = d->tail;
    a.stop = d->stop;
    a.done = d->done;
    a.head = (uint32_t *)(d->dev + o_head);
    a.idle = (uint32_t *)(d->dev + o_head + 128);  // its own cache line, before the mailboxes
    a.mail = (Mail *)(d->dev + o_mail);
    a.ctl = (SlotCtl *)(d->dev + o_ctl);
    a.abort = (uint32_t *)(d->dev + o_abort);
    a.L = L;
    a.qcap = qcap;
    a.bar.words = (uint32_t *)(d->dev + o_bar);
    a.bar.spin_limit = 1u << 22;
    a.claim_limit = 1u << 31;  // about a minute of polling
    a.alpha = o.alpha == 0.0 ? 1.0 : o.alpha;
    a.beta = o.beta;
    a.force = force;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g->device) != hipSuccess || khz <= 0) khz = 100000;
    a.idle_ticks = (unsigned long long)khz * kServeIdleMs;
    for (int x = 0; x < kMaxXcd; x++) a.slot_of[x] = -1;
    if ((rc = size_serve_grid(d.get()))) {
        serve_destroy(d.release());
        return rc;
    }
    // slot arenas need no initial contents: a slot's first BFS resets its
    // state (slot control starts "not clean"), SSSP and BC reset theirs
    if ((rc = engine_alloc(&d->arenas, (uint64_t)d->nslots * L.total))) {
        serve_destroy(d.release());
        return rc;
    }
    a.slots = (char *)d->arenas;
    d->dev_bytes += (uint64_t)d->nslots * L.total;
    if (hipStreamSynchronize(stream) != hipSuccess) {
        (void)hipGetLastError();
        serve_destroy(d.release());
        SG_FAIL(SG_ERR_HIP, "server: serving kernel set-up failed");
    }
    *out = d.release();
    return SG_OK;
}

void serve_destroy(ServeDev *d)
{
    if (!d) return;
    if (d->stream) {
        if (d->tail) {
            std::atomic_thread_fence(std::memory_order_seq_cst);
            __atomic_store_n(d->stop, 1u, __ATOMIC_SEQ_CST);
        }
        (void)hipStreamSynchronize(d->stream);
    }
    if (d->dev) (void)hipFree(d->dev);
    if (d->arenas) (void)hipFree(d->arenas);
    if (d->ring) (void)hipHostFree(d->ring);
    if (d->done) (void)hipHostFree((void *)d->done);
    if (d->tail) (void)hipHostFree(d->tail);
    delete d;
}

uint32_t serve_publish(ServeDev *d, int graph, int kind, uint32_t dev_source, uint32_t source_id, uint32_t row)
{
    const uint32_t seq = d->published;
    ServeEntry &en = d->ring[seq % kServeRing];
    en.graph = (uint32_t)graph;
    en.kind = (uint32_t)kind;
    en.source = dev_source;
    en.source_id = source_id;
    en.row = row;
    __atomic_store_n(&d->done[seq % kServeRing], 0ull, __ATOMIC_RELAXED);
    d->published = seq + 1;
    // the entry's stores are ordered before the tail's (x86 TSO and a
    // release store), and the device reads both at system scope
    __atomic_store_n(d->tail, d->published, __ATOMIC_RELEASE);
    return seq;
}

bool serve_poll(ServeDev *d, uint32_t seq, int32_t *status)
{
    const unsigned long long w = __atomic_load_n(&d->done[seq % kServeRing], __ATOMIC_ACQUIRE);
    if ((uint32_t)w != seq + 1u) return false;
    *status = (int32_t)(uint32_t)(w >> 32);
    return true;
}

int serve_pump(ServeDev *d, bool *running)
{
    if (d->running) {
        const hipError_t q = hipStreamQuery(d->stream);
        if (q == hipErrorNotReady) {
            *running = true;
            return SG_OK;
        }
        d->running = false;
        if (q != hipSuccess) {
            (void)hipGetLastError();
            SG_FAIL(SG_ERR_HIP, "serving kernel failed");
        }
        uint32_t ab = 0;
        SG_HIP(hipMemcpy(&ab, d->a.abort, 4, hipMemcpyDeviceToHost));
        if (ab) {
            // an expired wait: barrier counters are mid-way and slot state is
            // unknown; clear them (and the flag) before the next launch
            int rc = serve_zero(d, d->dev + d->off_bar, d->bar_bytes);
            if (!rc) rc = serve_zero(d, d->a.abort, 4);
            if (!rc) rc = serve_zero(d, d->a.ctl, kMaxXcd * sizeof(SlotCtl));
            if (rc) return rc;
            SG_HIP(hipStreamSynchronize(d->stream));
            d->launches = 0;
            *running = false;
            SG_FAIL(SG_ERR_TIMEOUT, "serving kernel: a grid barrier wait expired");
        }
    }
    *running = false;
    uint32_t head = 0;
    SG_HIP(hipMemcpy(&head, d->a.head, 4, hipMemcpyDeviceToHost));
    if ((int32_t)(d->published - head) <= 0) return SG_OK;  // nothing unclaimed
    ServeArgs a = d->a;
    a.bar.rset = (uint32_t)(d->launches & 1u);
    a.probe = 0;
    // the launch starts with no idle votes and not closing (in stream order)
    int zr = serve_zero(d, d->a.idle, 8);
    if (zr) return zr;
    int rc = engine_launch_persistent_on(d->stream, [&]() -> int {
        hipLaunchKernelGGL(k_serve<kServeBlock>, dim3(d->blocks), dim3(kServeBlock), 0, d->stream, a);
        SG_HIP(hipGetLastError());
        return SG_OK;
    });
    if (rc) return rc;
    d->launches++;
    d->running = true;
    *running = true;
    return SG_OK;
}

uint32_t serve_claimed(ServeDev *d)
{
    uint32_t head = 0;
    if (hipMemcpy(&head, d->a.head, 4, hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipGetLastError();
        return d->published;
    }
    return head;
}

uint64_t serve_device_bytes(const ServeDev *d) { return d->dev_bytes; }

}  // namespace sg
