// serve.hip -- continuous-batching query server over device-resident tables.
//
// The reference executes one query at a time through its pull-based engine
// (src/execution/execution_engine.cpp:67-142) and re-transfers tables per
// query (include/data/value_vector.hpp:248-272).  Here tables stay in HBM
// for the server's lifetime and a run replays a list of queries with
// arrival times through n_slots worker slots, each a library context with
// its own HIP stream:
//   * admission: a free slot takes the earliest query (in arrival order) that
//     has arrived, when its padded HBM footprint fits in the budget left by
//     the queries in flight (FIFO, no reordering);
//   * the admitted query's operators run on that slot's stream, so queries in
//     flight overlap on the device;
//   * completion frees the slot and the footprint; latency = end - arrival.
// n_slots = 1 is the plain one-query-at-a-time loop.  Everything the
// schedule looks at (arrival times, table sizes) is public.
#include "resident_tables.hpp"
#include "serve_kernels.hpp"

#include <map>
#include <new>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <thread>

struct vo_server;

namespace vo {
namespace {

using namespace serve_k;

const std::pair<const void *, BodyInfo> kBodies[] = {VO_SERVE_BODIES(VO_BODY_ENTRY)};
const bool kRegistered = (register_bodies(kBodies, sizeof kBodies / sizeof kBodies[0]), true);

// Every table also keeps a pinned host copy: the transfer-mode run ships it
// to the device per query, as the reference's engine does.
struct FactTable {
    uint64_t n, n_pad;
    int32_t *day;
    int64_t *qty, *price, *disc;
    int32_t *h_day;
    int64_t *h_qty, *h_price, *h_disc;
};

struct JoinTable {
    uint64_t n_dim, n_fact, n_pad;
    int32_t *dk, *dg, *fk;
    int64_t *fv;
    int32_t *h_dk, *h_dg, *h_fk;
    int64_t *h_fv;
};

struct SortTable {
    uint64_t n, n_pad;
    uint64_t *keys;
    uint32_t *pay;
    uint64_t *h_keys;
    uint32_t *h_pay;
};

uint64_t align256(uint64_t b) { return (b + 255) & ~255ull; }

// Where a query leaves its whole result on the device: the filter's
// compacted (price, disc) columns, the sort's sorted (keys, payload).
struct OutLoc {
    const void *c0 = nullptr, *c1 = nullptr;
    uint64_t b0 = 0, b1 = 0; // bytes of each
};

uint64_t pow2_at_least(uint64_t n, uint64_t floor_)
{
    uint64_t p = floor_;
    while (p < n) p <<= 1;
    return p;
}

__global__ void k_gen_join_cols(uint64_t seed, const uint32_t *__restrict__ perm, uint64_t n_dim, uint64_t n_fact,
                                int32_t *__restrict__ dk, int32_t *__restrict__ dg, int32_t *__restrict__ fk,
                                int64_t *__restrict__ fv)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_dim; i += stride) {
        dk[i] = (int32_t)perm[i];
        dg[i] = (int32_t)dev::rand_range(seed, 6, i, 0, 25);
    }
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_fact; i += stride) {
        // foreign keys uniform over the dimension keys (generation reads perm at a random row)
        fk[i] = (int32_t)perm[dev::rand_range(seed, 7, i, 0, n_dim)];
        fv[i] = dev::rand_range(seed, 8, i, 100, 10000000 - 100 + 1);
    }
}

__global__ void k_iota1(uint32_t *__restrict__ p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)(i + 1);
}


constexpr uint64_t kMaxGroups = 64;
constexpr uint64_t kJoinGroups = 32; // public bound of the generated group attribute (0..24)
// pinned result words of one query: filter 32 B, join groups, sort partial sums
constexpr size_t kResBytes = 2 * kChecksumBlocks * 8 > kMaxGroups * 12 + 8 ? 2 * kChecksumBlocks * 8 : kMaxGroups * 12 + 8;

// Query intermediates that live outside the slot context's scratch arena.
// In transfer mode the buffer also holds the query's input columns.
uint64_t query_buffer_bytes(uint32_t kind, uint64_t n_pad, bool transfer)
{
    switch (kind) {
    case VO_Q_FILTER_SUM: return (transfer ? 28 : 16) * n_pad + 64;
    case VO_Q_SORT: return 12 * n_pad + 2 * kChecksumBlocks * 8;
    default: return kMaxGroups * 12 + 64 + (transfer ? 4 * 256 + 20 * n_pad : 0);
    }
}

// Scratch-arena bytes the query's operators take (upper bounds of compact.hip / join.hip).
uint64_t query_scratch_bytes(uint32_t kind, uint64_t n_pad)
{
    switch (kind) {
    case VO_Q_FILTER_SUM: return 2 * n_pad + (1u << 20);
    case VO_Q_JOIN_GROUPSUM: return 34 * n_pad + (4u << 20);
    default: return 1u << 20;
    }
}

} // namespace
} // namespace vo

using namespace vo;

struct vo_server {
    int device = 0;
    uint32_t n_slots = 1;
    uint32_t queue_lanes = 0; // VO_RUN_QUEUE lanes (0: min(slots, 8))
    uint64_t budget = 0;
    std::string last_error;
    std::vector<vo_ctx *> slots;
    std::vector<void *> slot_buf;
    std::vector<uint64_t> slot_buf_bytes;
    std::vector<FactTable> facts;
    std::vector<JoinTable> joins;
    std::vector<SortTable> sorts;
    std::vector<void *> allocations;
    std::vector<void *> host_allocations;
    std::vector<void *> slot_hbuf; // pinned host staging per slot (transfer mode)
    std::vector<uint64_t> slot_hbuf_bytes;
    std::vector<void *> slot_res; // pinned result words per slot (the last node of a query)
    // per slot: captured query graphs by (kind, table), valid while the slot's
    // scratch arena and query buffer are the allocations they were captured on
    struct Graph {
        hipGraphExec_t exec = nullptr;
        void *scr = nullptr, *buf = nullptr;
        bool failed = false;
        vo::OutLoc loc; // where a replay leaves the query's whole result
    };
    std::vector<std::map<uint64_t, Graph>> graphs;
    // queue mode: one persistent context whose resident kernel runs a lane
    // per slot (slot i on lane i mod lanes)
    vo_ctx *qctx = nullptr;
    uint32_t q_lanes = 0;
};

namespace {

template <typename F>
int sguard(vo_server *s, F &&f)
{
    if (!s) return VO_ERR_INVALID;
    try {
        VO_HIP(hipSetDevice(s->device));
        f();
        return VO_OK;
    } catch (const Error &e) {
        s->last_error = e.msg;
        return e.status;
    } catch (const std::exception &e) {
        s->last_error = e.what();
        return VO_ERR_INVALID;
    }
}

void *dalloc(vo_server *s, size_t bytes)
{
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error{VO_ERR_NOMEM, "server table allocation failed"};
    }
    s->allocations.push_back(p);
    return p;
}

void *halloc(vo_server *s, size_t bytes)
{
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, std::max<size_t>(bytes, 256), hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error{VO_ERR_NOMEM, "server host table allocation failed"};
    }
    s->host_allocations.push_back(p);
    return p;
}

vo_ctx *slot_ctx(vo_server *s, uint32_t i)
{
    while (s->slots.size() <= i) {
        vo_config c{};
        c.device = s->device;
        vo_ctx *ctx = nullptr;
        int st = vo_ctx_create(&c, &ctx);
        if (st != VO_OK) throw Error{st, "server slot context creation failed"};
        s->slots.push_back(ctx);
        s->slot_buf.push_back(nullptr);
        s->slot_buf_bytes.push_back(0);
        s->slot_hbuf.push_back(nullptr);
        s->slot_hbuf_bytes.push_back(0);
        void *res = nullptr;
        VO_HIP(hipHostMalloc(&res, kResBytes, hipHostMallocDefault));
        s->slot_res.push_back(res);
        s->graphs.emplace_back();
    }
    return s->slots[i];
}

uint64_t table_n_pad(vo_server *s, const vo_query &q)
{
    switch (q.kind) {
    case VO_Q_FILTER_SUM:
        VO_CHECK(q.table < s->facts.size(), VO_ERR_INVALID, "unknown fact table");
        return s->facts[q.table].n_pad;
    case VO_Q_JOIN_GROUPSUM:
        VO_CHECK(q.table < s->joins.size(), VO_ERR_INVALID, "unknown join table");
        return s->joins[q.table].n_pad;
    case VO_Q_SORT:
        VO_CHECK(q.table < s->sorts.size(), VO_ERR_INVALID, "unknown sort table");
        return s->sorts[q.table].n_pad;
    default: throw Error{VO_ERR_INVALID, "unknown query kind"};
    }
}

void h2d(vo_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    if (bytes) VO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
}

// Enqueues one query on a slot's stream, ending with the copy of its result
// words to the slot's pinned result buffer `res`; nothing here waits.  With
// `transfer`, the input columns are first copied from the table's pinned host
// copy into the slot buffer and the query runs on those copies; a sort also
// copies its sorted columns back to host memory (hbuf).  *loc (may be NULL)
// receives where the whole result stays until the slot's next query.
void enqueue_query(vo_server *s, vo_ctx *ctx, void *buf, void *hbuf, const vo_query &q, bool transfer, void *res,
                   OutLoc *loc = nullptr)
{
    char *b = (char *)buf;
    OutLoc l;
    switch (q.kind) {
    case VO_Q_FILTER_SUM: {
        const FactTable &t = s->facts[q.table];
        vo_filter_args a{};
        if (transfer) {
            // private copies: the network compacts price/disc in place
            a.day = (int32_t *)b;
            a.qty = (int64_t *)(b + 4 * t.n_pad);
            a.price = (int64_t *)(b + 12 * t.n_pad);
            a.disc = (int64_t *)(b + 20 * t.n_pad);
            h2d(ctx, (void *)a.day, t.h_day, 4 * t.n);
            h2d(ctx, (void *)a.qty, t.h_qty, 8 * t.n);
            h2d(ctx, (void *)a.price, t.h_price, 8 * t.n);
            h2d(ctx, (void *)a.disc, t.h_disc, 8 * t.n);
            a.out = (uint64_t *)(b + 28 * t.n_pad);
        } else {
            a.day = t.day;
            a.qty = t.qty;
            a.price = t.price;
            a.disc = t.disc;
            a.out_price = (int64_t *)b;
            a.out_disc = (int64_t *)(b + 8 * t.n_pad);
            a.out = (uint64_t *)(b + 16 * t.n_pad);
        }
        a.n = t.n;
        a.n_pad = t.n_pad;
        a.day_lo = 8766;
        a.day_hi = 9131;
        a.disc_lo = 5;
        a.disc_hi = 7;
        a.qty_lt = 24;
        op_filter_sum(ctx, a);
        to_host_async(ctx, res, a.out, 32);
        // without output columns the network compacts the private copies in place
        l.c0 = transfer ? (const void *)a.price : (const void *)a.out_price;
        l.c1 = transfer ? (const void *)a.disc : (const void *)a.out_disc;
        l.b0 = l.b1 = 8 * t.n;
        break;
    }
    case VO_Q_JOIN_GROUPSUM: {
        const JoinTable &t = s->joins[q.table];
        vo_join_args a{};
        char *o = b;
        if (transfer) {
            char *in = b + align256(kMaxGroups * 12 + 64);
            a.dim_key = (const int32_t *)in;
            in += align256(4 * t.n_dim);
            a.dim_group = (const int32_t *)in;
            in += align256(4 * t.n_dim);
            a.fact_fk = (const int32_t *)in;
            in += align256(4 * t.n_fact);
            a.fact_val = (const int64_t *)in;
            h2d(ctx, (void *)a.dim_key, t.h_dk, 4 * t.n_dim);
            h2d(ctx, (void *)a.dim_group, t.h_dg, 4 * t.n_dim);
            h2d(ctx, (void *)a.fact_fk, t.h_fk, 4 * t.n_fact);
            h2d(ctx, (void *)a.fact_val, t.h_fv, 8 * t.n_fact);
        } else {
            a.dim_key = t.dk;
            a.dim_group = t.dg;
            a.fact_fk = t.fk;
            a.fact_val = t.fv;
        }
        a.n_dim = t.n_dim;
        a.n_fact = t.n_fact;
        a.n_pad = t.n_pad;
        a.out_group = (int32_t *)o;
        a.out_sum = (int64_t *)(o + kMaxGroups * 4);
        a.out_n_groups = (uint64_t *)(o + kMaxGroups * 12);
        // the tables' group attribute lies in [0, 25): dense GROUP BY over 32 ids
        a.max_groups = kJoinGroups;
        a.dense_groups = 1;
        op_join_groupsum(ctx, a);
        to_host_async(ctx, res, o, kMaxGroups * 12 + 8);
        break;
    }
    case VO_Q_SORT: {
        const SortTable &t = s->sorts[q.table];
        uint64_t *k = (uint64_t *)b;
        uint32_t *p = (uint32_t *)(b + 8 * t.n_pad);
        uint64_t *part = (uint64_t *)(b + 12 * t.n_pad);
        if (transfer) {
            h2d(ctx, k, t.h_keys, 8 * t.n);
            h2d(ctx, p, t.h_pay, 4 * t.n);
        } else {
            // the resident table stays unsorted: the query sorts its own copy
            copy_async(ctx, k, t.keys, 8 * t.n_pad);
            copy_async(ctx, p, t.pay, 4 * t.n_pad);
        }
        op_sort_kp(ctx, k, p, t.n, t.n_pad);
        launch(ctx, LaunchInfo{"sort_checksum", dim3(kChecksumBlocks), dim3(256), 0, t.n_pad}, k_sort_checksum,
               (const uint64_t *)k, (const uint32_t *)p, t.n_pad, part);
        if (transfer) { // the sorted relation goes back to the host
            char *hb = (char *)hbuf;
            VO_HIP(hipMemcpyAsync(hb, k, 8 * t.n, hipMemcpyDeviceToHost, ctx->stream));
            VO_HIP(hipMemcpyAsync(hb + 8 * t.n, p, 4 * t.n, hipMemcpyDeviceToHost, ctx->stream));
        }
        to_host_async(ctx, res, part, 2 * kChecksumBlocks * 8);
        l.c0 = k;
        l.c1 = p;
        l.b0 = 8 * t.n;
        l.b1 = 4 * t.n;
        break;
    }
    default: throw Error{VO_ERR_INVALID, "unknown query kind"};
    }
    if (loc) *loc = l;
}

// Bytes of a query's whole result (vo_query.out).
uint64_t out_bytes(vo_server *s, const vo_query &q)
{
    switch (q.kind) {
    case VO_Q_FILTER_SUM: return q.table < s->facts.size() ? 32 + 16 * s->facts[q.table].n : 0;
    case VO_Q_JOIN_GROUPSUM: return q.table < s->joins.size() ? kMaxGroups * 12 + 8 : 0;
    case VO_Q_SORT: return q.table < s->sorts.size() ? 12 * s->sorts[q.table].n : 0;
    default: return 0;
    }
}

// Copies a finished query's whole result into q.out (host memory): the
// result words from the pinned buffer, the columns from the slot's device
// buffer, which holds them until the slot's next query.  The copies run on
// the slot's stream: a legacy-stream copy would have to wait for every
// blocking stream, and another slot may be capturing its stream as a graph.
void copy_out(vo_server *s, vo_ctx *ctx, const vo_query &q, const void *res, const OutLoc &l)
{
    if (!q.out) return;
    char *o = (char *)q.out;
    switch (q.kind) {
    case VO_Q_FILTER_SUM:
        std::memcpy(o, res, 32);
        o += 32;
        break;
    case VO_Q_JOIN_GROUPSUM: std::memcpy(o, res, kMaxGroups * 12 + 8); return;
    default: break;
    }
    if (l.b0) VO_HIP(hipMemcpyAsync(o, l.c0, l.b0, hipMemcpyDeviceToHost, ctx->stream));
    if (l.b1) VO_HIP(hipMemcpyAsync(o + l.b0, l.c1, l.b1, hipMemcpyDeviceToHost, ctx->stream));
    if (l.b0 || l.b1) VO_HIP(hipStreamSynchronize(ctx->stream));
    (void)s;
}

// The result words of a finished query from its result buffer.
void finish_query(vo_server *s, const vo_query &q, const void *res, uint64_t r[4])
{
    switch (q.kind) {
    case VO_Q_FILTER_SUM: {
        const uint64_t *h = (const uint64_t *)res;
        r[0] = h[0];
        r[1] = h[2];
        r[2] = h[3];
        r[3] = 0;
        break;
    }
    case VO_Q_JOIN_GROUPSUM: {
        const char *h = (const char *)res;
        const int32_t *g = (const int32_t *)h;
        const int64_t *sm = (const int64_t *)(h + kMaxGroups * 4);
        const uint64_t G = *(const uint64_t *)(h + kMaxGroups * 12);
        uint64_t w = 0, tot = 0;
        for (uint64_t i = 0; i < std::min<uint64_t>(G, kMaxGroups); ++i) {
            w += (uint64_t)(g[i] + 1) * (uint64_t)sm[i];
            tot += (uint64_t)sm[i];
        }
        r[0] = G;
        r[1] = w;
        r[2] = tot;
        r[3] = 0;
        break;
    }
    default: { // VO_Q_SORT
        const uint64_t *h = (const uint64_t *)res;
        uint64_t x = 0, y = 0;
        for (uint32_t i = 0; i < kChecksumBlocks; ++i) {
            x += h[2 * i];
            y += h[2 * i + 1];
        }
        r[0] = s->sorts[q.table].n;
        r[1] = x;
        r[2] = y;
        r[3] = 0;
        break;
    }
    }
}

// Records the query's launches and copies on the slot stream as a graph
// (nothing runs); false if the stream cannot be captured for it.
bool capture_query(vo_server *s, vo_ctx *ctx, void *buf, void *res, const vo_query &q, hipGraphExec_t *exec,
                   OutLoc *loc)
{
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    bool ok = true;
    try {
        enqueue_query(s, ctx, buf, nullptr, q, false, res, loc);
    } catch (const Error &) {
        ok = false;
    }
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    ok = ok && e == hipSuccess && graph != nullptr;
    if (ok) ok = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (!ok) {
        (void)hipGetLastError();
        *exec = nullptr;
    }
    return ok;
}

// Queue mode: the query's operators are recorded as one program and run by
// the resident kernel on the slot's lane, beside the other lanes' queries
// (blocking until this query's own completion flag is set).
void run_query_queue(vo_server *s, uint32_t slot, uint32_t lane, const vo_query &q, uint64_t r[4])
{
    vo_ctx *ctx = s->qctx;
    void *buf = s->slot_buf[slot], *res = s->slot_res[slot];
    OutLoc loc;
    uint64_t seq;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        Recorder rec;
        rec.lane = lane;
        ctx->rec = &rec;
        try {
            enqueue_query(s, ctx, buf, nullptr, q, false, res, &loc);
        } catch (...) {
            ctx->rec = nullptr;
            throw;
        }
        ctx->rec = nullptr;
        seq = queue_publish_locked(ctx, lane, rec);
    }
    queue_wait_throw(ctx, seq, 0); // several slot threads wait at once: no shared last_error
    finish_query(s, q, res, r);
    copy_out(s, ctx, q, res, loc);
}

// The queue-mode context with `lanes` lanes, its resident kernel started.
void queue_ctx(vo_server *s, uint32_t lanes)
{
    if (s->qctx && s->q_lanes == lanes) return;
    if (s->qctx) vo_ctx_destroy(s->qctx);
    s->qctx = nullptr;
    s->q_lanes = 0;
    vo_config c{};
    c.device = s->device;
    c.flags = VO_CFG_PERSISTENT;
    c.queue_capacity = 1024;
    c.queue_lanes = lanes;
    vo_ctx *ctx = nullptr;
    int st = vo_ctx_create(&c, &ctx);
    if (st != VO_OK) throw Error{st, "server queue context creation failed"};
    st = vo_queue_start(ctx);
    if (st != VO_OK) {
        const std::string m = vo_last_error(ctx);
        vo_ctx_destroy(ctx);
        throw Error{st, "server queue: " + m};
    }
    s->qctx = ctx;
    s->q_lanes = lanes;
}

// Runs one query on a slot and fills its result words (blocking).  Over the
// resident tables, a (kind, table) pair seen before on this slot replays the
// graph captured after its first run: one launch instead of one per kernel.
bool run_query(vo_server *s, uint32_t slot, const vo_query &q, bool transfer, uint64_t r[4])
{
    vo_ctx *ctx = s->slots[slot];
    void *buf = s->slot_buf[slot], *res = s->slot_res[slot];
    OutLoc loc;
    if (transfer) {
        enqueue_query(s, ctx, buf, s->slot_hbuf[slot], q, true, res, &loc);
        VO_HIP(hipStreamSynchronize(ctx->stream));
        finish_query(s, q, res, r);
        copy_out(s, ctx, q, res, loc);
        return false;
    }
    auto &cache = s->graphs[slot];
    const uint64_t key = ((uint64_t)q.kind << 32) | q.table;
    auto it = cache.find(key);
    const bool fresh = it != cache.end() && it->second.scr == ctx->scratch && it->second.buf == buf;
    if (fresh && it->second.exec) {
        VO_HIP(hipGraphLaunch(it->second.exec, ctx->stream));
        VO_HIP(hipStreamSynchronize(ctx->stream));
        finish_query(s, q, res, r);
        copy_out(s, ctx, q, res, it->second.loc);
        return true;
    }
    enqueue_query(s, ctx, buf, nullptr, q, false, res, &loc);
    VO_HIP(hipStreamSynchronize(ctx->stream));
    finish_query(s, q, res, r);
    copy_out(s, ctx, q, res, loc);
    if (fresh) return false; // a capture of this pair failed before: it stays eager
    vo_server::Graph g;
    g.scr = ctx->scratch;
    g.buf = buf;
    if (it != cache.end() && it->second.exec) (void)hipGraphExecDestroy(it->second.exec);
    g.failed = !capture_query(s, ctx, buf, res, q, &g.exec, &g.loc);
    cache[key] = g;
    return false;
}

} // namespace

extern "C" {

int vo_server_create(const vo_server_config *cfg, vo_server **out)
{
    if (!cfg || !out || cfg->queue_lanes > 8) return VO_ERR_INVALID;
    *out = nullptr;
    vo_server *s = new vo_server();
    s->device = (int)cfg->device;
    s->n_slots = std::max<uint32_t>(1, cfg->n_slots);
    s->queue_lanes = cfg->queue_lanes;
    s->budget = cfg->hbm_budget_bytes ? cfg->hbm_budget_bytes : (200ull << 30);
    int st = sguard(s, [&] {
        for (uint32_t i = 0; i < s->n_slots; ++i) slot_ctx(s, i);
    });
    if (st != VO_OK) {
        vo_server_destroy(s);
        return st;
    }
    *out = s;
    return VO_OK;
}

int vo_server_destroy(vo_server *s)
{
    if (!s) return VO_ERR_INVALID;
    (void)hipSetDevice(s->device);
    if (s->qctx) vo_ctx_destroy(s->qctx);
    for (auto &cache : s->graphs)
        for (auto &kv : cache)
            if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    for (vo_ctx *c : s->slots) vo_ctx_destroy(c);
    for (void *p : s->slot_res)
        if (p) (void)hipHostFree(p);
    for (void *p : s->slot_buf)
        if (p) (void)hipFree(p);
    for (void *p : s->allocations) (void)hipFree(p);
    for (void *p : s->slot_hbuf)
        if (p) (void)hipHostFree(p);
    for (void *p : s->host_allocations) (void)hipHostFree(p);
    delete s;
    return VO_OK;
}

const char *vo_server_last_error(const vo_server *s) { return s ? s->last_error.c_str() : "null server"; }

int vo_server_add_fact_table(vo_server *s, uint64_t seed, uint64_t n_rows, uint32_t *id)
{
    return sguard(s, [&] {
        VO_CHECK(n_rows > 0 && n_rows <= (1ull << 29), VO_ERR_INVALID, "fact rows out of range");
        FactTable t;
        t.n = n_rows;
        t.n_pad = pow2_at_least(n_rows, 4096);
        t.day = (int32_t *)dalloc(s, 4 * t.n_pad);
        t.qty = (int64_t *)dalloc(s, 8 * t.n_pad);
        t.price = (int64_t *)dalloc(s, 8 * t.n_pad);
        t.disc = (int64_t *)dalloc(s, 8 * t.n_pad);
        vo_ctx *ctx = slot_ctx(s, 0);
        op_gen_fact(ctx, seed, t.day, t.qty, t.price, t.disc, t.n);
        t.h_day = (int32_t *)halloc(s, 4 * t.n);
        t.h_qty = (int64_t *)halloc(s, 8 * t.n);
        t.h_price = (int64_t *)halloc(s, 8 * t.n);
        t.h_disc = (int64_t *)halloc(s, 8 * t.n);
        VO_HIP(hipMemcpyAsync(t.h_day, t.day, 4 * t.n, hipMemcpyDeviceToHost, ctx->stream));
        VO_HIP(hipMemcpyAsync(t.h_qty, t.qty, 8 * t.n, hipMemcpyDeviceToHost, ctx->stream));
        VO_HIP(hipMemcpyAsync(t.h_price, t.price, 8 * t.n, hipMemcpyDeviceToHost, ctx->stream));
        VO_HIP(hipMemcpyAsync(t.h_disc, t.disc, 8 * t.n, hipMemcpyDeviceToHost, ctx->stream));
        VO_HIP(hipStreamSynchronize(ctx->stream));
        s->facts.push_back(t);
        if (id) *id = (uint32_t)(s->facts.size() - 1);
    });
}

int vo_server_add_join_tables(vo_server *s, uint64_t seed, uint64_t n_dim, uint64_t n_fact, uint32_t *id)
{
    return sguard(s, [&] {
        VO_CHECK(n_dim > 0 && n_fact > 0 && n_dim + n_fact <= (1ull << 29), VO_ERR_INVALID, "join rows out of range");
        JoinTable t;
        t.n_dim = n_dim;
        t.n_fact = n_fact;
        t.n_pad = pow2_at_least(n_dim + n_fact, 4096);
        const uint64_t dpad = pow2_at_least(n_dim, 1);
        t.dk = (int32_t *)dalloc(s, 4 * n_dim);
        t.dg = (int32_t *)dalloc(s, 4 * n_dim);
        t.fk = (int32_t *)dalloc(s, 4 * n_fact);
        t.fv = (int64_t *)dalloc(s, 8 * n_fact);
        uint32_t *perm = (uint32_t *)dalloc(s, 4 * dpad);
        vo_ctx *ctx = slot_ctx(s, 0);
        // unique dimension keys: a Philox-tag shuffle of 1..dpad, first n_dim of them
        const dim3 g((uint32_t)std::min<uint64_t>(65536, (dpad + 255) / 256));
        launch(ctx, LaunchInfo{"server_iota", g, dim3(256), 0, dpad}, k_iota1, perm, dpad);
        vo_shuffle_args sh{perm, nullptr, dpad, seed};
        op_shuffle(ctx, sh);
        const dim3 g2((uint32_t)std::min<uint64_t>(65536, (std::max(n_dim, n_fact) + 255) / 256));
        launch(ctx, LaunchInfo{"server_gen_join", g2, dim3(256), 0, n_dim}, k_gen_join_cols, seed,
               (const uint32_t *)perm, n_dim, n_fact, t.dk, t.dg, t.fk, t.fv);
        t.h_dk = (int32_t *)halloc(s, 4 * n_dim);
        t.h_dg = (int32_t *)halloc(s, 4 * n_dim);
        t.h_fk = (int32_t *)halloc(s, 4 * n_fact);
        t.h_fv = (int64_t *)halloc(s, 8 * n_fact);
        VO_HIP(hipMemcpyAsync(t.h_dk, t.dk, 4 * n_dim, hipMemcpyDeviceToHost, ctx->stream));
        VO_HIP(hipMemcpyAsync(t.h_dg, t.dg, 4 * n_dim, hipMemcpyDeviceToHost, ctx->stream));
        VO_HIP(hipMemcpyAsync(t.h_fk, t.fk, 4 * n_fact, hipMemcpyDeviceToHost, ctx->stream));
        VO_HIP(hipMemcpyAsync(t.h_fv, t.fv, 8 * n_fact, hipMemcpyDeviceToHost, ctx->stream));
        VO_HIP(hipStreamSynchronize(ctx->stream));
        s->joins.push_back(t);
        if (id) *id = (uint32_t)(s->joins.size() - 1);
    });
}

int vo_server_add_sort_table(vo_server *s, uint64_t seed, uint64_t n_rows, uint32_t *id)
{
    return sguard(s, [&] {
        VO_CHECK(n_rows > 0 && n_rows <= (1ull << 29), VO_ERR_INVALID, "sort rows out of range");
        SortTable t;
        t.n = n_rows;
        t.n_pad = pow2_at_least(n_rows, 1);
        t.keys = (uint64_t *)dalloc(s, 8 * t.n_pad);
        t.pay = (uint32_t *)dalloc(s, 4 * t.n_pad);
        vo_ctx *ctx = slot_ctx(s, 0);
        op_gen_sort_input(ctx, seed, t.keys, t.pay, t.n);
        t.h_keys = (uint64_t *)halloc(s, 8 * t.n);
        t.h_pay = (uint32_t *)halloc(s, 4 * t.n);
        VO_HIP(hipMemcpyAsync(t.h_keys, t.keys, 8 * t.n, hipMemcpyDeviceToHost, ctx->stream));
        VO_HIP(hipMemcpyAsync(t.h_pay, t.pay, 4 * t.n, hipMemcpyDeviceToHost, ctx->stream));
        VO_HIP(hipStreamSynchronize(ctx->stream));
        s->sorts.push_back(t);
        if (id) *id = (uint32_t)(s->sorts.size() - 1);
    });
}

uint64_t vo_server_query_out_bytes(vo_server *s, const vo_query *q)
{
    return s && q ? out_bytes(s, *q) : 0;
}

int vo_server_table_column(vo_server *s, uint32_t kind, uint32_t table, uint32_t column, const void **host,
                           uint64_t *rows)
{
    return sguard(s, [&] {
        VO_CHECK(host && rows, VO_ERR_INVALID, "null output");
        switch (kind) {
        case VO_Q_FILTER_SUM: {
            VO_CHECK(table < s->facts.size() && column < 4, VO_ERR_INVALID, "unknown fact table column");
            const FactTable &t = s->facts[table];
            const void *c[4] = {t.h_day, t.h_qty, t.h_price, t.h_disc};
            *host = c[column];
            *rows = t.n;
            break;
        }
        case VO_Q_JOIN_GROUPSUM: {
            VO_CHECK(table < s->joins.size() && column < 4, VO_ERR_INVALID, "unknown join table column");
            const JoinTable &t = s->joins[table];
            const void *c[4] = {t.h_dk, t.h_dg, t.h_fk, t.h_fv};
            *host = c[column];
            *rows = column < 2 ? t.n_dim : t.n_fact;
            break;
        }
        case VO_Q_SORT: {
            VO_CHECK(table < s->sorts.size() && column < 2, VO_ERR_INVALID, "unknown sort table column");
            const SortTable &t = s->sorts[table];
            *host = column == 0 ? (const void *)t.h_keys : (const void *)t.h_pay;
            *rows = t.n;
            break;
        }
        default: throw Error{VO_ERR_INVALID, "unknown query kind"};
        }
    });
}

int vo_server_run(vo_server *s, const vo_query *q, uint64_t n, uint32_t n_slots_override, vo_query_result *out,
                  vo_run_stats *stats)
{
    return vo_server_run_mode(s, q, n, n_slots_override, VO_RUN_RESIDENT, out, stats);
}

int vo_server_run_mode(vo_server *s, const vo_query *q, uint64_t n, uint32_t n_slots_override, uint32_t mode,
                       vo_query_result *out, vo_run_stats *stats)
{
    return sguard(s, [&] {
        VO_CHECK(q && out, VO_ERR_INVALID, "null query list");
        VO_CHECK(mode == VO_RUN_RESIDENT || mode == VO_RUN_TRANSFER || mode == VO_RUN_QUEUE, VO_ERR_INVALID,
                 "unknown run mode");
        const bool transfer = mode == VO_RUN_TRANSFER;
        const bool queued = mode == VO_RUN_QUEUE;
        const uint32_t slots = n_slots_override ? n_slots_override : s->n_slots;
        const uint32_t lanes = std::min<uint32_t>(slots, s->queue_lanes ? s->queue_lanes : 8);
        // size every slot for the largest query it may take (no allocation while timing)
        std::vector<uint64_t> fp(n);
        uint64_t max_buf = 64, max_scr = 1 << 20, max_hbuf = 0;
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t np = table_n_pad(s, q[i]);
            const uint64_t b = vo::query_buffer_bytes(q[i].kind, np, transfer);
            const uint64_t c = vo::query_scratch_bytes(q[i].kind, np);
            fp[i] = b + c;
            max_buf = std::max(max_buf, b);
            max_scr = std::max(max_scr, c);
            if (transfer && q[i].kind == VO_Q_SORT) max_hbuf = std::max(max_hbuf, 12 * s->sorts[q[i].table].n);
        }
        for (uint32_t i = 0; i < slots; ++i) {
            vo_ctx *ctx = slot_ctx(s, i);
            if (s->slot_buf_bytes[i] < max_buf) {
                if (s->slot_buf[i]) VO_HIP(hipFree(s->slot_buf[i]));
                s->slot_buf[i] = nullptr;
                VO_HIP(hipMalloc(&s->slot_buf[i], max_buf));
                s->slot_buf_bytes[i] = max_buf;
            }
            if (!queued) scratch(ctx, max_scr);
            if (s->slot_hbuf_bytes[i] < max_hbuf) {
                if (s->slot_hbuf[i]) VO_HIP(hipHostFree(s->slot_hbuf[i]));
                s->slot_hbuf[i] = nullptr;
                VO_HIP(hipHostMalloc(&s->slot_hbuf[i], max_hbuf, hipHostMallocDefault));
                s->slot_hbuf_bytes[i] = max_hbuf;
            }
        }
        if (queued) {
            queue_ctx(s, lanes);
            std::lock_guard<std::mutex> lk(s->qctx->mu);
            for (uint32_t l = 0; l < lanes; ++l) reserve_lane_scratch(s->qctx, l, max_scr);
        }
        VO_HIP(hipDeviceSynchronize());

        std::mutex mu;
        std::condition_variable cv;
        uint64_t next = 0, in_flight = 0, peak = 0;
        uint32_t running = 0, max_running = 0;
        std::atomic<uint32_t> replays{0};
        int first_error = VO_OK;
        std::string err_msg;
        using clk = std::chrono::steady_clock;
        const auto t0 = clk::now();
        auto now_s = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };

        auto worker = [&](uint32_t slot) {
            (void)hipSetDevice(s->device);
            for (;;) {
                uint64_t i;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    for (;;) {
                        if (next >= n || first_error != VO_OK) return;
                        const vo_query &qq = q[next];
                        const double wait = qq.arrival_s - now_s();
                        if (wait > 0) {
                            cv.wait_for(lk, std::chrono::duration<double>(wait));
                            continue;
                        }
                        if (fp[next] > s->budget) {
                            // can never fit: reject instead of blocking the queue
                            out[next] = vo_query_result{};
                            out[next].status = VO_ERR_ADMISSION;
                            out[next].footprint_bytes = fp[next];
                            ++next;
                            continue;
                        }
                        if (in_flight + fp[next] > s->budget) {
                            cv.wait(lk);
                            continue;
                        }
                        break;
                    }
                    i = next++;
                    in_flight += fp[i];
                    peak = std::max(peak, in_flight);
                    max_running = std::max(max_running, ++running);
                }
                vo_query_result r{};
                r.admit_s = now_s();
                r.footprint_bytes = fp[i];
                r.slot = slot;
                try {
                    if (queued)
                        run_query_queue(s, slot, slot % lanes, q[i], r.r);
                    else if (run_query(s, slot, q[i], transfer, r.r))
                        replays.fetch_add(1);
                    r.status = VO_OK;
                } catch (const Error &e) {
                    r.status = e.status;
                    std::lock_guard<std::mutex> lk(mu);
                    if (first_error == VO_OK) {
                        first_error = e.status;
                        err_msg = e.msg;
                    }
                } catch (const std::exception &e) {
                    // nothing may leave a worker thread (std::terminate); the
                    // status follows sguard's
                    const int st = dynamic_cast<const std::bad_alloc *>(&e) ? VO_ERR_NOMEM : VO_ERR_INVALID;
                    r.status = st;
                    std::lock_guard<std::mutex> lk(mu);
                    if (first_error == VO_OK) {
                        first_error = st;
                        err_msg = std::string("serving worker: ") + e.what();
                    }
                }
                r.end_s = now_s();
                r.latency_s = r.end_s - q[i].arrival_s;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    out[i] = r;
                    in_flight -= fp[i];
                    --running;
                }
                cv.notify_all();
            }
        };
        std::vector<std::thread> th;
        for (uint32_t i = 0; i < slots; ++i) th.emplace_back(worker, i);
        for (auto &t : th) t.join();
        if (first_error != VO_OK) throw Error{first_error, err_msg};
        if (stats) {
            std::vector<double> lat;
            double first_arrival = 1e300, last_end = 0;
            for (uint64_t i = 0; i < n; ++i) {
                if (out[i].status != VO_OK) continue;
                lat.push_back(out[i].latency_s);
                first_arrival = std::min(first_arrival, q[i].arrival_s);
                last_end = std::max(last_end, out[i].end_s);
            }
            std::sort(lat.begin(), lat.end());
            vo_run_stats st{};
            if (!lat.empty()) {
                auto pct = [&](double p) { return lat[std::min(lat.size() - 1, (size_t)(p * (lat.size() - 1) + 0.5))]; };
                double sum = 0;
                for (double x : lat) sum += x;
                st.makespan_s = last_end - first_arrival;
                st.qps = st.makespan_s > 0 ? lat.size() / st.makespan_s : 0;
                st.p50_s = pct(0.50);
                st.p99_s = pct(0.99);
                st.mean_s = sum / lat.size();
            }
            st.peak_bytes_in_flight = peak;
            st.max_concurrent = max_running;
            st.graph_replays = replays.load();
            *stats = st;
        }
    });
}

} // extern "C"
