// api.hip -- context lifecycle, command submission/wait, scratch arena,
// launch tracing and kernel timing behind the C ABI of include/vo_api.h.
//
// Launch mode (default): vo_submit enqueues the command's kernels on the
// context stream and records a completion event; vo_wait waits on it with a
// timeout.  This replaces CommandQueueController::push/wait
// (src/cuda/commands.cpp:47-91) for callers that do not run the resident
// queue kernel (queue.hip).
#include "vo_internal.hpp"

#include <chrono>
#include <cstdio>
#include <thread>

using namespace vo;

namespace vo {

// The arena of the current command: the context's own in launch mode and
// on queue lane 0, a lane's own on the other lanes (commands of different
// lanes run at the same time).
void *scratch(vo_ctx *ctx, size_t bytes)
{
    const uint32_t lane = ctx->rec ? ctx->rec->lane : 0;
    void *&arena = lane ? ctx->lane_scratch[lane] : ctx->scratch;
    size_t &have = lane ? ctx->lane_scratch_bytes[lane] : ctx->scratch_bytes;
    if (bytes <= have) return arena;
    // commands already submitted may still use the old arena
    if (ctx->rec)
        queue_drain_locked(ctx);
    else
        VO_HIP(hipStreamSynchronize(ctx->stream));
    if (arena) VO_HIP(hipFree(arena));
    arena = nullptr;
    have = 0;
    size_t want = bytes + (bytes >> 3);
    want = (want + 0xFFFFF) & ~(size_t)0xFFFFF;
    hipError_t e = hipMalloc(&arena, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        arena = nullptr;
        throw Error{VO_ERR_NOMEM, "scratch arena: hipMalloc of " + std::to_string(want) + " bytes failed"};
    }
    have = want;
    return arena;
}

void reserve_lane_scratch(vo_ctx *ctx, uint32_t lane, size_t bytes)
{
    Recorder r;
    r.lane = lane;
    Recorder *saved = ctx->rec;
    ctx->rec = &r;
    try {
        (void)scratch(ctx, bytes);
    } catch (...) {
        ctx->rec = saved;
        throw;
    }
    ctx->rec = saved;
}

// ------------------------------------------------------ resident programs

namespace {
std::mutex g_body_mu;
std::vector<std::pair<const void *, BodyInfo>> &body_table()
{
    static std::vector<std::pair<const void *, BodyInfo>> t;
    return t;
}
} // namespace

void register_bodies(const std::pair<const void *, BodyInfo> *entries, size_t n)
{
    std::lock_guard<std::mutex> lk(g_body_mu);
    auto &t = body_table();
    t.insert(t.end(), entries, entries + n);
}

const BodyInfo *find_body(const void *kernel)
{
    std::lock_guard<std::mutex> lk(g_body_mu);
    for (const auto &e : body_table())
        if (e.first == kernel) return &e.second;
    return nullptr;
}

namespace {
// Blocks a recorded fill or copy uses: one per 64 KiB, at most 4096 (the lane
// strides over the bytes with at most all of its blocks).
uint32_t byte_step_blocks(uint64_t bytes)
{
    return (uint32_t)std::min<uint64_t>(4096, std::max<uint64_t>(1, (bytes + 65535) >> 16));
}
} // namespace

void fill_async(vo_ctx *ctx, void *dst, int value, uint64_t bytes)
{
    if (bytes == 0) return;
    if (!ctx->rec) {
        VO_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
        return;
    }
    Step s{};
    s.id = kStepFill;
    s.grid = byte_step_blocks(bytes);
    s.slots = 1;
    s.aux = (uint16_t)(value & 0xFF);
    const FillArgs a{dst, bytes};
    std::memcpy(s.blob, &a, sizeof a);
    ctx->rec->steps.push_back(s);
}

void copy_async(vo_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    if (bytes == 0) return;
    if (!ctx->rec) {
        VO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        return;
    }
    Step s{};
    s.id = kStepCopy;
    s.grid = byte_step_blocks(bytes);
    s.slots = 1;
    const CopyArgs a{dst, src, bytes};
    std::memcpy(s.blob, &a, sizeof a);
    ctx->rec->steps.push_back(s);
}

void to_host_async(vo_ctx *ctx, void *pinned_dst, const void *src, uint64_t bytes)
{
    if (bytes == 0) return;
    if (!ctx->rec) {
        VO_HIP(hipMemcpyAsync(pinned_dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        return;
    }
    copy_async(ctx, pinned_dst, src, bytes);
}

void upload_async(vo_ctx *ctx, void *dst, const void *host_src, uint64_t bytes)
{
    if (bytes == 0) return;
    if (!ctx->rec) {
        VO_HIP(hipMemcpyAsync(dst, host_src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return;
    }
    // the bytes travel with the program (host-mapped); a copy step moves them
    Recorder &r = *ctx->rec;
    const size_t off = (r.data.size() + 15) & ~(size_t)15;
    r.data.resize(off + bytes);
    std::memcpy(r.data.data() + off, host_src, bytes);
    Step s{};
    s.id = kStepCopy;
    s.grid = byte_step_blocks(bytes);
    s.slots = 1;
    const CopyArgs a{dst, nullptr, bytes};
    std::memcpy(s.blob, &a, sizeof a);
    r.steps.push_back(s);
    r.data_refs.emplace_back(r.steps.size() - 1, off);
}

} // namespace vo

namespace {
void dispatch(vo_ctx *ctx, const vo_command &cmd);
} // namespace

namespace vo {

// Launch mode: piece i runs on the context stream (i mod 4 == 0) or on
// helper stream i mod 4 - 1, forked from the context stream before the first
// piece and joined back into it after the last, so that one piece's small or
// latency-bound launches overlap another's.  The pieces share no memory they
// write, and none of them allocates (the caller's scratch is in place).
static void run_tracks_launch(vo_ctx *ctx, size_t n, const std::function<void(size_t)> &f)
{
    constexpr size_t kHelpers = 3;
    const size_t used = std::min(kHelpers, n - 1);
    for (size_t h = 0; h < used; ++h)
        if (!ctx->track_streams[h]) VO_HIP(hipStreamCreateWithFlags(&ctx->track_streams[h], hipStreamNonBlocking));
    for (size_t e = 0; e <= used; ++e)
        if (!ctx->track_events[e]) VO_HIP(hipEventCreateWithFlags(&ctx->track_events[e], hipEventDisableTiming));
    hipStream_t const main = ctx->stream;
    VO_HIP(hipEventRecord(ctx->track_events[0], main));
    for (size_t h = 0; h < used; ++h) VO_HIP(hipStreamWaitEvent(ctx->track_streams[h], ctx->track_events[0], 0));
    auto join = [&] {
        ctx->stream = main;
        for (size_t h = 0; h < used; ++h) {
            (void)hipEventRecord(ctx->track_events[h + 1], ctx->track_streams[h]);
            (void)hipStreamWaitEvent(main, ctx->track_events[h + 1], 0);
        }
    };
    try {
        for (size_t i = 0; i < n; ++i) {
            ctx->stream = i % (kHelpers + 1) == 0 ? main : ctx->track_streams[i % (kHelpers + 1) - 1];
            f(i);
        }
    } catch (...) {
        join();
        throw;
    }
    join();
}

void run_tracks(vo_ctx *ctx, size_t n, const std::function<void(size_t)> &f)
{
    if (!ctx->rec) {
        if (n < 2) {
            for (size_t i = 0; i < n; ++i) f(i);
            return;
        }
        run_tracks_launch(ctx, n, f);
        return;
    }
    Recorder *const parent = ctx->rec;
    std::vector<Recorder> tr(n);
    try {
        for (size_t i = 0; i < n; ++i) {
            tr[i].lane = parent->lane;
            ctx->rec = &tr[i];
            f(i);
        }
    } catch (...) {
        ctx->rec = parent;
        throw;
    }
    ctx->rec = parent;
    // each track's steps (a step and the slots its arguments run on over)
    std::vector<std::vector<size_t>> starts(n);
    size_t phases = 0;
    for (size_t i = 0; i < n; ++i) {
        for (size_t s = 0; s < tr[i].steps.size(); s += std::max<uint16_t>(1, tr[i].steps[s].slots))
            starts[i].push_back(s);
        phases = std::max(phases, starts[i].size());
    }
    // data: each track's bytes appended, its references moved with them
    std::vector<size_t> data_at(n);
    for (size_t i = 0; i < n; ++i) {
        const size_t off = (parent->data.size() + 15) & ~(size_t)15;
        parent->data.resize(off + tr[i].data.size());
        if (!tr[i].data.empty()) std::memcpy(parent->data.data() + off, tr[i].data.data(), tr[i].data.size());
        data_at[i] = off;
    }
    std::vector<std::vector<size_t>> new_index(n);
    for (size_t i = 0; i < n; ++i) new_index[i].assign(tr[i].steps.size(), 0);
    for (size_t k = 0; k < phases; ++k) {
        size_t last = SIZE_MAX;
        for (size_t i = 0; i < n; ++i) {
            if (k >= starts[i].size()) continue;
            const size_t s0 = starts[i][k];
            const size_t s1 = k + 1 < starts[i].size() ? starts[i][k + 1] : tr[i].steps.size();
            last = parent->steps.size();
            for (size_t s = s0; s < s1; ++s) {
                new_index[i][s] = parent->steps.size();
                parent->steps.push_back(tr[i].steps[s]);
            }
            parent->steps[last].aux |= kStepNoBarrier;
        }
        if (last != SIZE_MAX) parent->steps[last].aux &= (uint16_t)~kStepNoBarrier; // barrier after the phase
    }
    for (size_t i = 0; i < n; ++i)
        for (const auto &ref : tr[i].data_refs) parent->data_refs.emplace_back(new_index[i][ref.first], data_at[i] + ref.second);
}

void record_command(vo_ctx *ctx, const vo_command &cmd, Recorder &rec)
{
    VO_CHECK(!ctx->tracing, VO_ERR_UNSUPPORTED, "tracing runs in launch mode only");
    ctx->rec = &rec;
    try {
        dispatch(ctx, cmd);
    } catch (...) {
        ctx->rec = nullptr;
        throw;
    }
    ctx->rec = nullptr;
}

TraceSlot *trace_slot(vo_ctx *ctx)
{
    if (!ctx->tracing) return nullptr;
    size_t i = ctx->trace_launches.size(); // slot of the launch being recorded next
    if (i >= ctx->trace_capacity) {
        size_t cap = std::max<size_t>(1024, ctx->trace_capacity * 2);
        TraceSlot *ns = nullptr;
        VO_HIP(hipMalloc(&ns, cap * sizeof(TraceSlot)));
        VO_HIP(hipMemsetAsync(ns, 0, cap * sizeof(TraceSlot), ctx->stream));
        if (ctx->trace_slots) {
            VO_HIP(hipMemcpyAsync(ns, ctx->trace_slots, ctx->trace_capacity * sizeof(TraceSlot),
                                  hipMemcpyDeviceToDevice, ctx->stream));
            VO_HIP(hipStreamSynchronize(ctx->stream));
            VO_HIP(hipFree(ctx->trace_slots));
        }
        ctx->trace_slots = ns;
        ctx->trace_capacity = cap;
    }
    return ctx->trace_slots + i;
}

void begin_launch(vo_ctx *ctx, const LaunchInfo &li, hipEvent_t *ev_stop)
{
    if (ctx->tracing) {
        LaunchRecord r;
        r.name = li.family;
        r.grid[0] = li.grid.x;
        r.grid[1] = li.grid.y;
        r.grid[2] = li.grid.z;
        r.block[0] = li.block.x;
        r.block[1] = li.block.y;
        r.block[2] = li.block.z;
        r.lds = li.lds;
        r.scalar_hash = li.scalar_hash;
        r.addr_digest = 0;
        r.n_addr = 0;
        ctx->trace_launches.push_back(r);
    }
    if (ctx->flags & VO_CFG_KERNEL_TIMING) {
        TimingRecord t;
        t.family = li.family;
        VO_HIP(hipEventCreate(&t.start));
        VO_HIP(hipEventCreate(&t.stop));
        VO_HIP(hipEventRecord(t.start, ctx->stream));
        *ev_stop = t.stop;
        ctx->timings.push_back(t);
    }
}

void end_launch(vo_ctx *ctx, hipEvent_t ev_stop)
{
    if (ev_stop) VO_HIP(hipEventRecord(ev_stop, ctx->stream));
}

} // namespace vo

// ---------------------------------------------------------------- helpers

namespace {

int fail(vo_ctx *ctx, const Error &e)
{
    if (ctx) {
        ctx->last_error = e.msg;
        ctx->last_status = e.status;
    }
    return e.status;
}

template <typename F>
int guarded(vo_ctx *ctx, F &&f)
{
    if (!ctx) return VO_ERR_INVALID;
    try {
        f();
        return VO_OK;
    } catch (const Error &e) {
        return fail(ctx, e);
    } catch (const std::exception &e) {
        return fail(ctx, Error{VO_ERR_INVALID, e.what()});
    }
}

void dispatch(vo_ctx *ctx, const vo_command &cmd)
{
    const vo_command_args &a = cmd.args;
    switch (cmd.opcode) {
    case VO_OP_NOP: break;
    case VO_OP_SORT: op_sort_multi(ctx, a.sort); break;
    case VO_OP_SORT_KP: op_sort_kp(ctx, a.sort_kp.keys, a.sort_kp.payload, a.sort_kp.n, a.sort_kp.n_pad); break;
    case VO_OP_PERMUTE: op_permute(ctx, a.permute); break;
    case VO_OP_SCATTER: op_scatter(ctx, a.scatter); break;
    case VO_OP_GATHER: op_gather(ctx, a.gather); break;
    case VO_OP_COMPARE_EQ: op_compare_eq(ctx, a.compare_eq); break;
    case VO_OP_COMPOSE_POSITION_MAP: op_compose(ctx, a.compose_position_map); break;
    case VO_OP_ENRICH_POSITIONS: op_enrich(ctx, a.enrich_positions); break;
    case VO_OP_COMPACT: op_compact(ctx, a.compact); break;
    case VO_OP_SHUFFLE: op_shuffle(ctx, a.shuffle); break;
    case VO_OP_FILTER_SUM: op_filter_sum(ctx, a.filter); break;
    case VO_OP_PKFK_JOIN_GROUPSUM: op_join_groupsum(ctx, a.join); break;
    case VO_OP_SORT_BATCH: op_sort_batch(ctx, a.sort_batch); break;
    case VO_OP_COMPACT_BATCH: op_compact_batch(ctx, a.compact_batch); break;
    case VO_OP_SORT_MERGE_JOIN_COUNT: (void)op_join_pairs_count(ctx, a.join_pairs); break;
    case VO_OP_SORT_MERGE_JOIN_WRITE: op_join_pairs_write(ctx, a.join_pairs); break;
    case VO_OP_GATHER_OBLIVIOUS: op_gather_oblivious(ctx, a.gather_oblivious); break;
    case VO_OP_PREDICATE_MASK: op_predicate_mask(ctx, a.predicate); break;
    case VO_OP_TERMINATE: break;
    default:
        // the reference ignores unknown opcodes (persistent_kernel.cu:69-72); report them instead
        throw Error{VO_ERR_UNSUPPORTED, "opcode " + std::to_string(cmd.opcode) + " is not supported"};
    }
}

hipEvent_t get_event(vo_ctx *ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    VO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
}

// Retires completed commands (in order).  Returns a HIP error if the stream
// reported one.
void reap(vo_ctx *ctx)
{
    size_t k = 0;
    for (; k < ctx->pending.size(); ++k) {
        hipError_t e = hipEventQuery(ctx->pending[k].second);
        if (e == hipErrorNotReady) break;
        if (e != hipSuccess) throw Error{VO_ERR_HIP, std::string("device error: ") + hipGetErrorString(e)};
        ctx->done_seq = ctx->pending[k].first;
        ctx->event_pool.push_back(ctx->pending[k].second);
    }
    ctx->pending.erase(ctx->pending.begin(), ctx->pending.begin() + k);
}

int wait_seq(vo_ctx *ctx, uint64_t seq, uint32_t timeout_ms)
{
    if (timeout_ms == 0) timeout_ms = ctx->wait_timeout_ms;
    auto t0 = std::chrono::steady_clock::now();
    int spins = 0;
    for (;;) {
        {
            std::lock_guard<std::mutex> lk(ctx->mu);
            reap(ctx);
            if (ctx->done_seq >= seq || seq >= ctx->next_seq) {
                if (seq >= ctx->next_seq) throw Error{VO_ERR_INVALID, "wait on a sequence id never submitted"};
                return VO_OK;
            }
        }
        auto el = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0);
        if ((uint64_t)el.count() >= timeout_ms)
            throw Error{VO_ERR_TIMEOUT, "vo_wait: command " + std::to_string(seq) + " not finished after " +
                                            std::to_string(timeout_ms) + " ms"};
        if (++spins < 2000)
            std::this_thread::yield();
        else
            std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

} // namespace

// -------------------------------------------------------------- extern C

extern "C" {

int vo_version(void) { return VO_API_VERSION; }

const char *vo_status_string(int s)
{
    switch (s) {
    case VO_OK: return "ok";
    case VO_ERR_INVALID: return "invalid argument";
    case VO_ERR_HIP: return "HIP error";
    case VO_ERR_TIMEOUT: return "timeout";
    case VO_ERR_NOMEM: return "out of memory";
    case VO_ERR_UNSUPPORTED: return "unsupported";
    case VO_ERR_QUEUE: return "queue error";
    case VO_ERR_ADMISSION: return "admission rejected";
    default: return "unknown status";
    }
}

int vo_ctx_create(const vo_config *cfg, vo_ctx **out)
{
    if (!out) return VO_ERR_INVALID;
    *out = nullptr;
    vo_ctx *ctx = new vo_ctx();
    try {
        vo_config c{};
        if (cfg) c = *cfg;
        int dev = c.device;
        if (dev < 0) VO_HIP(hipGetDevice(&dev));
        VO_HIP(hipSetDevice(dev));
        ctx->device = dev;
        ctx->flags = c.flags;
        ctx->queue_capacity = c.queue_capacity ? c.queue_capacity : 1024;
        VO_CHECK(is_pow2(ctx->queue_capacity), VO_ERR_INVALID, "queue_capacity must be a power of two");
        ctx->wait_timeout_ms = c.wait_timeout_ms ? c.wait_timeout_ms : 60000;
        ctx->queue_idle_us = c.queue_idle_us ? c.queue_idle_us : 20000;
        ctx->queue_lanes = c.queue_lanes ? c.queue_lanes : 1;
        if (c.stream) {
            ctx->stream = (hipStream_t)c.stream;
        } else {
            // a blocking stream: work the caller queues on the legacy null
            // stream (a framework's default stream) stays ordered with it
            VO_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamDefault));
            ctx->own_stream = true;
        }
        hipDeviceProp_t prop;
        VO_HIP(hipGetDeviceProperties(&prop, dev));
        ctx->num_cus = prop.multiProcessorCount;
        ctx->tuning = apply_tuning_word(ctx->tuning, c.tuning);
        noalloc_scratch_init(dev);
    } catch (const Error &e) {
        delete ctx;
        return e.status;
    }
    *out = ctx;
    return VO_OK;
}

int vo_ctx_destroy(vo_ctx *ctx)
{
    if (!ctx) return VO_ERR_INVALID;
    if (ctx->queue) vo_queue_stop(ctx, 5000);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : ctx->pending) (void)hipEventDestroy(p.second);
    ctx->pending.clear();
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto &t : ctx->timings) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    for (void *p : ctx->lane_scratch)
        if (p) (void)hipFree(p);
    join_pairs_state_free(ctx);
    staging_free(ctx);
    if (ctx->trace_slots) (void)hipFree(ctx->trace_slots);
    for (auto &s : ctx->track_streams)
        if (s) (void)hipStreamDestroy(s);
    for (auto &e : ctx->track_events)
        if (e) (void)hipEventDestroy(e);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return VO_OK;
}

const char *vo_last_error(const vo_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

static int run_sync(vo_ctx *ctx, vo_command &cmd);

// Persistent mode, a command whose program cannot be recorded (a kernel
// without a resident body): the submitted commands finish, the resident
// kernel leaves, the command runs on the context stream to completion, and a
// NOP through the ring gives it the next sequence id, so waits and ordering
// stay those of the queue.
static int run_parked(vo_ctx *ctx, const vo_command &cmd)
{
    const int st = queue_park(ctx);
    if (st != VO_OK) return st;
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        dispatch(ctx, cmd);
        VO_HIP(hipStreamSynchronize(ctx->stream));
    });
}

uint64_t vo_submit(vo_ctx *ctx, vo_command *cmd)
{
    if (!ctx || !cmd) return 0;
    if (ctx->queue) {
        const uint64_t seq = queue_submit(ctx, cmd);
        if (seq != 0 || ctx->last_status != VO_ERR_UNSUPPORTED) return seq;
        if (run_parked(ctx, *cmd) != VO_OK) return 0;
        vo_command nop{};
        nop.opcode = VO_OP_NOP;
        const uint64_t s2 = queue_submit(ctx, &nop);
        cmd->sequence_id = s2;
        return s2;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    try {
        dispatch(ctx, *cmd);
        hipEvent_t ev = get_event(ctx);
        VO_HIP(hipEventRecord(ev, ctx->stream));
        uint64_t seq = ctx->next_seq++;
        ctx->pending.emplace_back(seq, ev);
        cmd->sequence_id = seq;
        return seq;
    } catch (const Error &e) {
        fail(ctx, e);
        return 0;
    }
}

int vo_wait(vo_ctx *ctx, uint64_t seq, uint32_t timeout_ms)
{
    if (ctx && ctx->queue) return queue_wait(ctx, seq, timeout_ms);
    int rc = VO_OK;
    int g = guarded(ctx, [&] { rc = wait_seq(ctx, seq, timeout_ms); });
    return g != VO_OK ? g : rc;
}

int vo_poll(vo_ctx *ctx, uint64_t seq)
{
    if (!ctx) return -VO_ERR_INVALID;
    if (ctx->queue) return queue_poll(ctx, seq);
    try {
        std::lock_guard<std::mutex> lk(ctx->mu);
        reap(ctx);
        return ctx->done_seq >= seq ? 1 : 0;
    } catch (const Error &e) {
        return -fail(ctx, e);
    }
}

int vo_sync(vo_ctx *ctx)
{
    return guarded(ctx, [&] {
        if (ctx->queue) {
            int st = queue_sync(ctx);
            if (st != VO_OK) throw Error{st, ctx->last_error};
            // and the asynchronous launch-mode calls (they park the queue)
            VO_HIP(hipStreamSynchronize(ctx->stream));
            return;
        }
        VO_HIP(hipStreamSynchronize(ctx->stream));
        std::lock_guard<std::mutex> lk(ctx->mu);
        reap(ctx);
    });
}

static int run_sync(vo_ctx *ctx, vo_command &cmd)
{
    if (!ctx) return VO_ERR_INVALID;
    uint64_t s = vo_submit(ctx, &cmd);
    if (s == 0) return ctx->last_status ? ctx->last_status : VO_ERR_INVALID;
    return vo_wait(ctx, s, 0);
}

int vo_sort_kp(vo_ctx *ctx, uint64_t *keys, uint32_t *payload, uint64_t n, uint64_t n_pad)
{
    vo_command c{};
    c.opcode = VO_OP_SORT_KP;
    c.args.sort_kp = {keys, payload, n, n_pad};
    return run_sync(ctx, c);
}

int vo_compact_noalloc(vo_ctx *ctx, const vo_compact_args *args)
{
    if (!args) return VO_ERR_INVALID;
    if (ctx && ctx->queue) {
        const int st = queue_park(ctx);
        if (st != VO_OK) return st;
    }
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        op_compact_noalloc(ctx, *args);
    });
}

int vo_sort_kp_async(vo_ctx *ctx, uint64_t *keys, uint32_t *payload, uint64_t n, uint64_t n_pad)
{
    if (ctx && ctx->queue) {
        const int st = queue_park(ctx);
        if (st != VO_OK) return st;
    }
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        op_sort_kp(ctx, keys, payload, n, n_pad);
    });
}

#define VO_SYNC_OP(fn, OPC, field, T)                                                                          \
    int fn(vo_ctx *ctx, const T *args)                                                                         \
    {                                                                                                          \
        if (!args) return VO_ERR_INVALID;                                                                      \
        vo_command c{};                                                                                        \
        c.opcode = OPC;                                                                                        \
        c.args.field = *args;                                                                                  \
        return run_sync(ctx, c);                                                                               \
    }

VO_SYNC_OP(vo_sort, VO_OP_SORT, sort, vo_sort_args)
VO_SYNC_OP(vo_permute, VO_OP_PERMUTE, permute, vo_permute_args)
VO_SYNC_OP(vo_compact, VO_OP_COMPACT, compact, vo_compact_args)
VO_SYNC_OP(vo_shuffle, VO_OP_SHUFFLE, shuffle, vo_shuffle_args)
VO_SYNC_OP(vo_filter_sum, VO_OP_FILTER_SUM, filter, vo_filter_args)
VO_SYNC_OP(vo_pkfk_join_groupsum, VO_OP_PKFK_JOIN_GROUPSUM, join, vo_join_args)
VO_SYNC_OP(vo_sort_batch, VO_OP_SORT_BATCH, sort_batch, vo_sort_batch_args)
VO_SYNC_OP(vo_compact_batch, VO_OP_COMPACT_BATCH, compact_batch, vo_compact_batch_args)
VO_SYNC_OP(vo_join_pairs_write, VO_OP_SORT_MERGE_JOIN_WRITE, join_pairs, vo_join_pairs_args)
VO_SYNC_OP(vo_gather_oblivious, VO_OP_GATHER_OBLIVIOUS, gather_oblivious, vo_gather_oblivious_args)
VO_SYNC_OP(vo_predicate_mask, VO_OP_PREDICATE_MASK, predicate, vo_predicate_args)

int vo_transfer_to_device(vo_ctx *ctx, void *dst, const void *src, uint64_t bytes, uint64_t padded_bytes,
                          uint64_t *link_bytes)
{
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        op_transfer_to_device(ctx, dst, src, bytes, padded_bytes);
        if (link_bytes) *link_bytes = last_transfer_bytes(ctx);
    });
}

int vo_transfer_to_host(vo_ctx *ctx, void *dst, const void *src, uint64_t bytes, uint64_t padded_bytes,
                        uint64_t *link_bytes)
{
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        op_transfer_to_host(ctx, dst, src, bytes, padded_bytes);
        if (link_bytes) *link_bytes = last_transfer_bytes(ctx);
    });
}

int vo_join_pairs_count(vo_ctx *ctx, const vo_join_pairs_args *args, uint64_t *n_out)
{
    if (!args) return VO_ERR_INVALID;
    vo_command c{};
    c.opcode = VO_OP_SORT_MERGE_JOIN_COUNT;
    c.args.join_pairs = *args;
    const int st = run_sync(ctx, c);
    if (st != VO_OK || !n_out) return st;
    if (!args->state) {
        *n_out = join_pairs_last_count(ctx);
        return VO_OK;
    }
    // the count sits in the caller's state buffer, after the three copies
    return guarded(ctx, [&] {
        const uint64_t off = join_pairs_count_offset(args->n_left, args->n_right);
        VO_HIP(hipMemcpy(n_out, (const char *)args->state + off, 8, hipMemcpyDeviceToHost));
    });
}

uint64_t vo_join_pairs_state_bytes(uint64_t n_left, uint64_t n_right)
{
    if (n_left >= (1ull << 31) || n_right >= (1ull << 31)) return 0;
    return join_pairs_state_bytes(n_left, n_right);
}

// ------------------------------------------------------------ diagnostics

int vo_trace_begin(vo_ctx *ctx)
{
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        VO_CHECK(!ctx->queue, VO_ERR_UNSUPPORTED, "tracing runs in launch mode only");
        ctx->trace_launches.clear();
        if (ctx->trace_slots) VO_HIP(hipMemsetAsync(ctx->trace_slots, 0, ctx->trace_capacity * sizeof(TraceSlot), ctx->stream));
        ctx->tracing = true;
    });
}

int vo_trace_end(vo_ctx *ctx, vo_trace *out)
{
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        VO_CHECK(ctx->tracing, VO_ERR_INVALID, "no trace open");
        VO_HIP(hipStreamSynchronize(ctx->stream));
        const size_t n = ctx->trace_launches.size();
        std::vector<TraceSlot> slots(n);
        if (n) VO_HIP(hipMemcpy(slots.data(), ctx->trace_slots, n * sizeof(TraceSlot), hipMemcpyDeviceToHost));
        uint64_t ld = 0x243F6A8885A308D3ull, ad = 0x13198A2E03707344ull, na = 0;
        for (size_t i = 0; i < n; ++i) {
            LaunchRecord &r = ctx->trace_launches[i];
            r.addr_digest = slots[i].digest;
            r.n_addr = slots[i].count;
            for (char ch : r.name) ld = hash_mix(ld, (uint64_t)(unsigned char)ch);
            for (int d = 0; d < 3; ++d) ld = hash_mix(ld, hash_mix(r.grid[d], r.block[d]));
            ld = hash_mix(ld, hash_mix(r.lds, r.scalar_hash));
            ad = hash_mix(ad, hash_mix(r.addr_digest, r.n_addr));
            na += r.n_addr;
        }
        ctx->tracing = false;
        ctx->last_trace = ctx->trace_launches;
        ctx->trace_launches.clear();
        if (out) {
            out->n_launches = n;
            out->launch_digest = ld;
            out->address_digest = ad;
            out->n_addresses = na;
        }
    });
}

int vo_trace_launch(vo_ctx *ctx, uint64_t i, char *name, uint32_t name_cap, uint32_t grid[3], uint32_t block[3],
                    uint64_t *addr_digest)
{
    return guarded(ctx, [&] {
        VO_CHECK(i < ctx->last_trace.size(), VO_ERR_INVALID, "trace launch index out of range");
        const LaunchRecord &r = ctx->last_trace[i];
        if (name && name_cap) {
            std::snprintf(name, name_cap, "%s", r.name.c_str());
        }
        for (int d = 0; d < 3; ++d) {
            if (grid) grid[d] = r.grid[d];
            if (block) block[d] = r.block[d];
        }
        if (addr_digest) *addr_digest = r.addr_digest;
    });
}

int vo_kernel_stats(vo_ctx *ctx, const char *family, uint64_t *launches, double *total_ms)
{
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        VO_HIP(hipStreamSynchronize(ctx->stream));
        uint64_t cnt = 0;
        double ms = 0;
        for (auto &t : ctx->timings) {
            if (family) {
                // "name*" matches every family that starts with "name"
                const size_t fl = std::strlen(family);
                const bool prefix = fl > 0 && family[fl - 1] == '*';
                if (prefix ? t.family.compare(0, fl - 1, family, fl - 1) != 0 : t.family != family) continue;
            }
            float f = 0;
            VO_HIP(hipEventElapsedTime(&f, t.start, t.stop));
            ms += f;
            ++cnt;
        }
        if (launches) *launches = cnt;
        if (total_ms) *total_ms = ms;
    });
}

int vo_kernel_stats_reset(vo_ctx *ctx)
{
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        VO_HIP(hipStreamSynchronize(ctx->stream));
        for (auto &t : ctx->timings) {
            (void)hipEventDestroy(t.start);
            (void)hipEventDestroy(t.stop);
        }
        ctx->timings.clear();
    });
}

int vo_sort_plan(vo_ctx *ctx, uint64_t n_pad, uint32_t *n_passes)
{
    return guarded(ctx, [&] {
        VO_CHECK(is_pow2(n_pad), VO_ERR_INVALID, "n_pad must be a power of two");
        if (n_passes) *n_passes = (uint32_t)sort_plan_passes(ctx, n_pad);
    });
}

int vo_autotune_sort(vo_ctx *ctx, uint32_t log_min, uint32_t log_max, uint64_t rows_per_class, uint32_t *chosen)
{
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        VO_CHECK(!ctx->queue && !ctx->tracing, VO_ERR_INVALID, "autotune runs in launch mode without a trace");
        autotune_sort(ctx, (int)log_min, (int)log_max, rows_per_class, chosen);
    });
}

int vo_set_class_tuning(vo_ctx *ctx, uint32_t log_n, uint32_t tuning_word)
{
    return guarded(ctx, [&] {
        VO_CHECK(log_n <= 32, VO_ERR_INVALID, "log_n above 32");
        ctx->class_tuning[log_n] = tuning_word;
    });
}

int vo_get_class_tuning(vo_ctx *ctx, uint32_t log_n, uint32_t *tuning_word)
{
    return guarded(ctx, [&] {
        VO_CHECK(log_n <= 32 && tuning_word, VO_ERR_INVALID, "bad arguments");
        *tuning_word = ctx->class_tuning[log_n];
    });
}

int vo_hbm_copy(vo_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        op_hbm_copy(ctx, dst, src, bytes);
        VO_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int vo_gen_sort_input(vo_ctx *ctx, uint64_t seed, uint64_t *keys, uint32_t *payload, uint64_t n)
{
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        op_gen_sort_input(ctx, seed, keys, payload, n);
        VO_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int vo_gen_fact(vo_ctx *ctx, uint64_t seed, int32_t *day, int64_t *qty, int64_t *price, int64_t *disc, uint64_t n)
{
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        op_gen_fact(ctx, seed, day, qty, price, disc, n);
        VO_HIP(hipStreamSynchronize(ctx->stream));
    });
}

} // extern "C"
