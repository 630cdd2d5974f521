// vo_internal.hpp -- context, launch bookkeeping and device helpers shared by
// every translation unit of libvelodb_amd.so.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vo_api.h"

namespace vo {

// ------------------------------------------------------------------ errors

struct Error {
    int status;
    std::string msg;
};

// A failed call's error stays the thread's last HIP error until read: it is
// read here, so the next launch's hipGetLastError check does not report it
// again as its own.
#define VO_HIP(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            (void)hipGetLastError();                                                              \
            throw ::vo::Error{VO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)};     \
        }                                                                                         \
    } while (0)

#define VO_CHECK(cond, status, msg)                                                               \
    do {                                                                                          \
        if (!(cond)) throw ::vo::Error{(status), (msg)};                                          \
    } while (0)

// ------------------------------------------------------------------ tracing

// Per-launch record kept while a trace is open.
struct LaunchRecord {
    std::string name;
    uint32_t grid[3];
    uint32_t block[3];
    uint32_t lds;
    uint64_t scalar_hash; // hash of the launch's non-pointer arguments
    uint64_t addr_digest; // filled at trace end from the device slot
    uint64_t n_addr;
};

// Per-launch device slot written by instrumented kernels: {sum of mixed
// per-thread address hashes, number of addresses}.
struct TraceSlot {
    unsigned long long digest;
    unsigned long long count;
};

struct TimingRecord {
    std::string family;
    hipEvent_t start, stop;
};

// ------------------------------------------------------------------ tuning

// Public-size-only schedule parameters (chosen by the autotuner or defaults).
struct SortTuning {
    int max_fused_stages = 5;  // stages per register-only global pass (1..5)
    int max_hybrid_stages = 8; // stages per LDS-staged global pass (0 = off, 4..8; used above max_fused_stages)
    int pipe_merge = 0;        // resident merge blocks with next-tile prefetch
    int wide_block = 0;        // 12-byte sort: 1024-thread blocks, 8 records per thread
    int flip_hybrid_block = 0; // 12-byte sort's hybrid passes: the other block shape (twice / half the threads)
    int flip_streaming = 0;    // sort and compaction passes: the other cache policy than the size rule picks
};

struct Recorder;

} // namespace vo

// ------------------------------------------------------------------ context

struct vo_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    uint32_t flags = 0;
    uint32_t queue_capacity = 1024;
    uint32_t wait_timeout_ms = 60000;
    uint32_t queue_idle_us = 20000;
    uint32_t queue_lanes = 1;
    int num_cus = 0;
    std::string last_error;
    int last_status = 0;
    std::mutex mu; // guards submission and bookkeeping

    // completion tracking (launch mode)
    uint64_t next_seq = 1;
    uint64_t done_seq = 0;
    std::vector<std::pair<uint64_t, hipEvent_t>> pending;
    std::vector<hipEvent_t> event_pool;

    // scratch arena (launch mode and queue lane 0), and the other lanes' own
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    void *lane_scratch[8] = {};
    size_t lane_scratch_bytes[8] = {};

    // tracing
    bool tracing = false;
    std::vector<vo::LaunchRecord> trace_launches;
    vo::TraceSlot *trace_slots = nullptr; // device
    size_t trace_capacity = 0;
    std::vector<vo::LaunchRecord> last_trace;

    // timing
    std::vector<vo::TimingRecord> timings;

    vo::SortTuning tuning;
    uint32_t class_tuning[33] = {}; // per log2(n_pad) tuning word from the autotuner (0 = context default)

    // persistent queue (queue.hip)
    void *queue = nullptr;
    vo::Recorder *rec = nullptr; // set while a command's program is recorded (resident.hpp)

    // equi-join routing state kept from COUNT to WRITE (join_pairs.hip)
    void *join_pairs_state = nullptr;

    // pinned staging ring and copy threads of the host transfers (transfer.hip)
    void *staging = nullptr;

    // helper streams (default priority) and their fork/join events for the
    // independent pieces of one command in launch mode (run_tracks)
    hipStream_t track_streams[3] = {};
    hipEvent_t track_events[4] = {};
};

namespace vo {

// Scratch arena: returns at least `bytes` of device memory, valid until the
// next call on this context.  Synchronises the stream before growing.
void *scratch(vo_ctx *ctx, size_t bytes);

// Trace slot for the next launch (nullptr when not tracing).
TraceSlot *trace_slot(vo_ctx *ctx);

// Record a launch (trace + timing bookkeeping).  `pre` is called before the
// kernel launch (after the timing start event); returns nothing.
struct LaunchInfo {
    const char *family;
    dim3 grid, block;
    uint32_t lds;
    uint64_t scalar_hash;
};

void begin_launch(vo_ctx *ctx, const LaunchInfo &li, hipEvent_t *ev_stop);
void end_launch(vo_ctx *ctx, hipEvent_t ev_stop);


inline uint64_t hash_mix(uint64_t h, uint64_t v)
{
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ull;
    return h ^ (h >> 32);
}

inline int ilog2(uint64_t v)
{
    int r = -1;
    while (v) {
        v >>= 1;
        ++r;
    }
    return r;
}

inline bool is_pow2(uint64_t v) { return v && !(v & (v - 1)); }

// Tuning word (vo_config.tuning): bits 0-3 register-pass stages, bits 4-7
// hybrid-pass stages + 1; bits 8-11 toggle the prefetching merge, the
// 1024-thread 12-byte sort, its hybrid block shape and the sort passes'
// cache policy.  0 keeps the defaults.
inline SortTuning apply_tuning_word(SortTuning t, uint32_t w)
{
    if (w & 0xF) t.max_fused_stages = (int)(w & 0xF);
    if (w & 0xF0) t.max_hybrid_stages = (int)((w >> 4) & 0xF) - 1;
    if (w & 0x100) t.pipe_merge = !t.pipe_merge;
    if (w & 0x200) t.wide_block = !t.wide_block;
    if (w & 0x400) t.flip_hybrid_block = !t.flip_hybrid_block;
    if (w & 0x800) t.flip_streaming = !t.flip_streaming;
    return t;
}

// ---- op entry points implemented in the .hip files (enqueue only)
void op_sort_kp(vo_ctx *ctx, uint64_t *keys, uint32_t *payload, uint64_t n, uint64_t n_pad);
void op_sort_multi(vo_ctx *ctx, const vo_sort_args &a);
void op_sort_batch(vo_ctx *ctx, const vo_sort_batch_args &a);
void op_compact_batch(vo_ctx *ctx, const vo_compact_batch_args &a);
void op_permute(vo_ctx *ctx, const vo_permute_args &a);
void op_scatter(vo_ctx *ctx, const vo_scatter_args &a);
void op_gather(vo_ctx *ctx, const vo_gather_args &a);
void op_compare_eq(vo_ctx *ctx, const vo_compare_eq_args &a);
void op_compose(vo_ctx *ctx, const vo_compose_args &a);
void op_enrich(vo_ctx *ctx, const vo_enrich_args &a);
void op_compact(vo_ctx *ctx, const vo_compact_args &a);
// op_compact over the fixed device scratch of compact.hip (allocates nothing)
void op_compact_noalloc(vo_ctx *ctx, const vo_compact_args &a);
// looks up `device`'s copy of that scratch (device current); once per device
void noalloc_scratch_init(int device);
void op_shuffle(vo_ctx *ctx, const vo_shuffle_args &a);
void op_filter_sum(vo_ctx *ctx, const vo_filter_args &a);
void op_join_groupsum(vo_ctx *ctx, const vo_join_args &a);
uint64_t op_join_pairs_count(vo_ctx *ctx, const vo_join_pairs_args &a); // returns n_out (synchronises)
void op_join_pairs_write(vo_ctx *ctx, const vo_join_pairs_args &a);
void join_pairs_state_free(vo_ctx *ctx);
uint64_t join_pairs_last_count(vo_ctx *ctx); // n_out of the last COUNT (context state)
uint64_t join_pairs_state_bytes(uint64_t n_left, uint64_t n_right);
uint64_t join_pairs_count_offset(uint64_t n_left, uint64_t n_right); // of n_out inside the state
void op_gather_oblivious(vo_ctx *ctx, const vo_gather_oblivious_args &a);
void op_predicate_mask(vo_ctx *ctx, const vo_predicate_args &a);
void op_transfer_to_device(vo_ctx *ctx, void *dst, const void *src, uint64_t bytes, uint64_t padded_bytes);
void op_transfer_to_host(vo_ctx *ctx, void *dst, const void *src, uint64_t bytes, uint64_t padded_bytes);
uint64_t last_transfer_bytes(vo_ctx *ctx);
void staging_free(vo_ctx *ctx);
void op_gen_sort_input(vo_ctx *ctx, uint64_t seed, uint64_t *keys, uint32_t *payload, uint64_t n);
void op_gen_fact(vo_ctx *ctx, uint64_t seed, int32_t *day, int64_t *qty, int64_t *price, int64_t *disc,
                 uint64_t n);
int sort_plan_passes(vo_ctx *ctx, uint64_t n_pad);
void op_hbm_copy(vo_ctx *ctx, void *dst, const void *src, uint64_t bytes);
void autotune_sort(vo_ctx *ctx, int log_min, int log_max, uint64_t rows_per_class, uint32_t *chosen);

// ---- resident command queue (queue.hip); used when ctx->queue is set
uint64_t queue_submit(vo_ctx *ctx, vo_command *cmd);
int queue_wait(vo_ctx *ctx, uint64_t seq, uint32_t timeout_ms);
// The same wait, throwing instead of setting the context's last error (for
// callers that wait from several threads on one context).
void queue_wait_throw(vo_ctx *ctx, uint64_t seq, uint32_t timeout_ms);
int queue_poll(vo_ctx *ctx, uint64_t seq);
int queue_sync(vo_ctx *ctx);
int queue_park(vo_ctx *ctx);          // finish submitted commands, let the resident kernel leave
uint64_t queue_submit_lane(vo_ctx *ctx, vo_command *cmd, uint32_t lane);
void queue_drain_locked(vo_ctx *ctx); // every submitted command complete (caller holds ctx->mu)
uint32_t queue_lanes(vo_ctx *ctx);
uint64_t queue_next_seq(vo_ctx *ctx);                  // id of the command being recorded
// A recorded program of several operators as one ring entry (internal opcode
// kOpProgram); caller holds ctx->mu.
constexpr uint32_t kOpProgram = 0x1000;
uint64_t queue_publish_locked(vo_ctx *ctx, uint32_t lane, Recorder &rec);
// Sizes a queue lane's scratch arena ahead of its first command.
void reserve_lane_scratch(vo_ctx *ctx, uint32_t lane, size_t bytes);
void queue_wait_seq_locked(vo_ctx *ctx, uint64_t seq); // caller holds ctx->mu

// ---- shared device helpers (sort.hip) used by other ops
// Sort (u64 key, u64 value) records by (key, value) ascending, n_pad = 2^k.
void sort_kv(vo_ctx *ctx, uint64_t *keys, uint64_t *vals, uint64_t n_pad);
// Sort (u64 key, u32 payload) without padding (all n_pad rows are live).
void sort_kp_full(vo_ctx *ctx, uint64_t *keys, uint32_t *payload, uint64_t n_pad);

// ---- compaction helpers (compact.hip) used by the join / group-by
// ORCompact of two 8-byte columns by a device mask over n_pad live rows;
// the selected count lands in *count_dev, followed by
// compact_rows2_scratch(n_pad) - 8 bytes of tile arrays the call owns.
void compact_rows2(vo_ctx *ctx, uint64_t *c0, uint64_t *c1, const uint8_t *mask, uint64_t n_pad, uint64_t *count_dev);
void compact_rows1(vo_ctx *ctx, uint64_t *c0, const uint8_t *mask, uint64_t n_pad, uint64_t *count_dev);
size_t compact_rows2_scratch(uint64_t n_pad);

} // namespace vo

// ------------------------------------------------------------ device side

namespace vo {
namespace dev {

// Address-trace folding used by the instrumented kernel builds.
__device__ __forceinline__ uint64_t tfold(uint64_t h, uint64_t tag, uint64_t idx)
{
    uint64_t v = (idx << 4) ^ tag;
    h ^= v * 0x9E3779B97F4A7C15ull;
    h = (h << 27) | (h >> 37);
    return h * 0xBF58476D1CE4E5B9ull + 0x94D049BB133111EBull;
}

// Fold a thread's final hash into the launch slot (order independent sum).
__device__ __forceinline__ void tcommit(TraceSlot *slot, uint64_t gtid, uint64_t h, uint64_t n)
{
    uint64_t z = h ^ (gtid * 0xD1B54A32D192ED03ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    atomicAdd(&slot->digest, (unsigned long long)z);
    atomicAdd(&slot->count, (unsigned long long)n);
}

// One thread's address trace, gated at run time (slot == nullptr when the
// context is not tracing): at() folds the index of each global access where
// it happens, done() commits the thread's hash to the launch slot.
struct Atr {
    TraceSlot *slot;
    uint64_t h = 0, n = 0;
    __device__ explicit Atr(TraceSlot *s) : slot(s) {}
    __device__ __forceinline__ void at(uint64_t tag, uint64_t idx)
    {
        if (slot) {
            h = tfold(h, tag, idx);
            ++n;
        }
    }
    __device__ __forceinline__ void done(uint64_t gtid)
    {
        if (slot) tcommit(slot, gtid, h, n);
    }
};

// The same with the gate fixed at compile time, for kernels whose register
// budget is tight (ON = false compiles to nothing).
template <bool ON>
struct AtrT {
    TraceSlot *slot;
    uint64_t h = 0, n = 0;
    __device__ explicit AtrT(TraceSlot *s) : slot(s) {}
    __device__ __forceinline__ void at(uint64_t tag, uint64_t idx)
    {
        if constexpr (ON) {
            h = tfold(h, tag, idx);
            ++n;
        }
    }
    __device__ __forceinline__ void done(uint64_t gtid)
    {
        if constexpr (ON) tcommit(slot, gtid, h, n);
    }
};

__device__ __forceinline__ uint64_t gtid() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t rand64(uint64_t seed, uint32_t stream, uint64_t i)
{
    const uint64_t g = 0x9E3779B97F4A7C15ull;
    uint64_t s0 = mix64(seed ^ (g * ((uint64_t)stream + 1)));
    return mix64(s0 + (i + 1) * g);
}

__device__ __forceinline__ int64_t rand_range(uint64_t seed, uint32_t stream, uint64_t i, int64_t lo,
                                              uint64_t range)
{
    uint64_t r32 = rand64(seed, stream, i) >> 32;
    return lo + (int64_t)((r32 * range) >> 32);
}

} // namespace dev
} // namespace vo

#include "resident.hpp"
