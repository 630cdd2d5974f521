// sort_kp.hip -- the (u64 key, u32 payload) oblivious sort (the headline
// path) and the (u64 key, u64 value) sort used by the join.
#include "resident_tables.hpp"
#include "sort_kernels.hpp"

namespace vo {

using namespace sortimpl;

namespace {

const std::pair<const void *, BodyInfo> kBodies[] = {VO_SORT_KP_BODIES(VO_BODY_ENTRY) VO_SORT_KV_BODIES(VO_BODY_ENTRY)};
const bool kRegistered = (register_bodies(kBodies, sizeof kBodies / sizeof kBodies[0]), true);

// Register-only global passes take up to 5 stages below 2^27 records and up
// to 3 from 2^27 on, where groups of 4 and 5 stages run faster as LDS-staged
// passes (2^28: 5455 vs 5162 and 5165 vs 4955 GB/s); below 2^27 the register
// passes win (2^25: 5.06 vs 5.13 ms per sort, 2^23: 1.107 vs 1.128 ms).
// Hybrid passes of 4-6 stages with 1024-thread blocks: at 2^28, 5494 -> 5611,
// 5147 -> 5216 and 5319 -> 5483 GB/s (7 and 8 stages are faster at 512).
using SortKP = SortImpl<RecKP, 13, 4, 512, 5, 8, 3, 27, 6>;
// the 16-byte sort likewise: at 2^28 its 4-5 stage register passes ran at
// 4.80-4.99 TB/s against 5.11-5.36 for its LDS-staged passes
// (512-thread hybrid passes measured no faster for it: 78.04-78.05 vs 78.06-78.14 ms per
// 2^28-row shuffle with permutation output)
using SortKV = SortImpl<RecKV, 12, 4, 256, 5, 7, 3, 27>;
// The resident queue kernel runs 256-thread blocks: the 12-byte sort runs
// there on 2^12-record tiles (the same network, so the same output).
using SortKP_R = SortImpl<RecKP, 12, 4, 256, 5, 7, 3, 27>;
// The same tiles with 1024-thread blocks of 8 records per thread (tuning
// bit 0x200): the same passes at up to twice the waves per SIMD.
using SortKP_W = SortImpl<RecKP, 13, 3, 1024, 5, 8, 3, 27>;

void run_kp(vo_ctx *ctx, Bufs<RecKP> b, int n_log, uint64_t n_valid, int fill)
{
    if (ctx->rec)
        SortKP_R::run(ctx, b, n_log, n_valid, fill);
    else if (ctx->tuning.wide_block)
        SortKP_W::run(ctx, b, n_log, n_valid, fill);
    else
        SortKP::run(ctx, b, n_log, n_valid, fill);
}
} // namespace

void op_sort_kp(vo_ctx *ctx, uint64_t *keys, uint32_t *payload, uint64_t n, uint64_t n_pad)
{
    VO_CHECK(is_pow2(n_pad) && n <= n_pad, VO_ERR_INVALID, "n_pad must be a power of two >= n");
    VO_CHECK(n_pad <= kMaxSortRows, VO_ERR_INVALID, "n_pad above 2^29 (32-bit buffer offsets of 8-byte columns)");
    VO_CHECK(keys && payload, VO_ERR_INVALID, "null buffer");
    Bufs<RecKP> b;
    b.k[0] = keys;
    b.p = payload;
    run_kp(ctx, b, ilog2(n_pad), n, 1);
}

void sort_kp_full(vo_ctx *ctx, uint64_t *keys, uint32_t *payload, uint64_t n_pad)
{
    Bufs<RecKP> b;
    b.k[0] = keys;
    b.p = payload;
    run_kp(ctx, b, ilog2(n_pad), n_pad, 0);
}

void sort_kv(vo_ctx *ctx, uint64_t *keys, uint64_t *vals, uint64_t n_pad)
{
    Bufs<RecKV> b;
    b.k[0] = keys;
    b.p = vals;
    SortKV::run(ctx, b, ilog2(n_pad), n_pad, 0);
}

int sort_plan_passes(vo_ctx *ctx, uint64_t n_pad) { return SortKP::passes(ctx, ilog2(n_pad)); }

// ------------------------------------------------------------ batched sorts

namespace {

// Runs one size class with the class's tuning word (autotuned or default).
void run_class(vo_ctx *ctx, Bufs<RecKP> b, int c, uint64_t count, const uint64_t *offsets, const uint64_t *ns)
{
    const SortTuning saved = ctx->tuning;
    ctx->tuning = apply_tuning_word(saved, ctx->class_tuning[c]);
    try {
        if (ctx->rec)
            SortKP_R::run_mapped(ctx, b, c, count, offsets, ns);
        else
            SortKP::run_mapped(ctx, b, c, count, offsets, ns);
    } catch (...) {
        ctx->tuning = saved;
        throw;
    }
    ctx->tuning = saved;
}

} // namespace

// Independent sorts of uneven public sizes, bucketed by power-of-two size
// class: every class is one schedule (one launch per pass for all of its
// instances), instead of one schedule per instance.
void op_sort_batch(vo_ctx *ctx, const vo_sort_batch_args &a)
{
    VO_CHECK(a.keys && a.payload && (a.items || a.n_items == 0), VO_ERR_INVALID, "null buffer");
    std::vector<std::vector<uint64_t>> off(33), ns(33);
    for (uint64_t i = 0; i < a.n_items; ++i) {
        const vo_batch_item &it = a.items[i];
        VO_CHECK(is_pow2(it.n_pad) && it.n <= it.n_pad && it.offset % it.n_pad == 0, VO_ERR_INVALID,
                 "batch item: n_pad must be a power of two >= n dividing offset");
        VO_CHECK(it.n_pad <= kMaxSortRows, VO_ERR_INVALID, "batch item above 2^29 rows");
        const int c = ilog2(it.n_pad);
        off[c].push_back(it.offset);
        ns[c].push_back(it.n);
    }
    // tables for every class in one upload (offsets then live counts)
    std::vector<uint64_t> host;
    std::vector<size_t> at(33, 0);
    for (int c = 0; c <= 32; ++c) {
        at[c] = host.size();
        host.insert(host.end(), off[c].begin(), off[c].end());
        host.insert(host.end(), ns[c].begin(), ns[c].end());
    }
    if (host.empty()) return;
    uint64_t *dtab = (uint64_t *)scratch(ctx, host.size() * 8);
    upload_async(ctx, dtab, host.data(), host.size() * 8);
    Bufs<RecKP> b;
    b.k[0] = a.keys;
    b.p = a.payload;
    // size classes are independent pieces (in the resident queue they share
    // the lane between barriers)
    std::vector<int> classes;
    for (int c = 32; c >= 0; --c)
        if (!off[c].empty()) classes.push_back(c);
    run_tracks(ctx, classes.size(), [&](size_t t) {
        const int c = classes[t];
        const uint64_t cnt = off[c].size();
        if (c < 4) {
            // instances of at most 8 rows: one small launch each
            for (uint64_t i = 0; i < cnt; ++i) op_sort_kp(ctx, a.keys + off[c][i], a.payload + off[c][i], ns[c][i], 1ull << c);
            return;
        }
        run_class(ctx, b, c, cnt, dtab + at[c], dtab + at[c] + cnt);
    });
    // the host table must outlive its upload (recorded: it travels with the program)
    if (!ctx->rec) VO_HIP(hipStreamSynchronize(ctx->stream));
}

// Autotuner: for each size class in [log_min, log_max], times every
// schedule variant on a synthetic batch of that class (public sizes only)
// and keeps the fastest tuning word in ctx->class_tuning.
void autotune_sort(vo_ctx *ctx, int log_min, int log_max, uint64_t rows_per_class, uint32_t *chosen)
{
    VO_CHECK(log_min >= 4 && log_max <= 29 && log_min <= log_max, VO_ERR_INVALID, "autotune class range 4..29");
    static const uint32_t candidates[] = {
        0x5 | (9u << 4),         0x5 | (9u << 4) | 0x100, 0x5 | (8u << 4), 0x5 | (7u << 4),
        0x5 | (1u << 4),         0x4 | (9u << 4),         0x4 | (1u << 4), 0x3 | (9u << 4),
    };
    for (int c = log_min; c <= log_max; ++c) {
        const uint64_t inst = 1ull << c;
        const uint64_t count = std::max<uint64_t>(1, rows_per_class / inst);
        const uint64_t total = count * inst;
        uint64_t *keys = nullptr;
        uint32_t *pay = nullptr;
        VO_HIP(hipMalloc(&keys, total * 8));
        hipError_t e = hipMalloc(&pay, total * 4);
        if (e != hipSuccess) {
            (void)hipFree(keys);
            VO_HIP(e);
        }
        std::vector<uint64_t> host(2 * count);
        for (uint64_t i = 0; i < count; ++i) {
            host[i] = i * inst;
            host[count + i] = inst;
        }
        uint64_t *dtab = nullptr;
        VO_HIP(hipMalloc(&dtab, host.size() * 8));
        VO_HIP(hipMemcpy(dtab, host.data(), host.size() * 8, hipMemcpyHostToDevice));
        Bufs<RecKP> b;
        b.k[0] = keys;
        b.p = pay;
        hipEvent_t e0, e1;
        VO_HIP(hipEventCreate(&e0));
        VO_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        uint32_t best_w = 0;
        const uint32_t saved_word = ctx->class_tuning[c];
        for (uint32_t w : candidates) {
            ctx->class_tuning[c] = w;
            op_gen_sort_input(ctx, 7, keys, pay, total);
            run_class(ctx, b, c, count, dtab, dtab + count); // warm-up
            float ms_total = 0;
            for (int rep = 0; rep < 3; ++rep) {
                op_gen_sort_input(ctx, 8 + rep, keys, pay, total);
                VO_HIP(hipEventRecord(e0, ctx->stream));
                run_class(ctx, b, c, count, dtab, dtab + count);
                VO_HIP(hipEventRecord(e1, ctx->stream));
                VO_HIP(hipEventSynchronize(e1));
                float ms = 0;
                VO_HIP(hipEventElapsedTime(&ms, e0, e1));
                ms_total += ms;
            }
            if (ms_total < best) {
                best = ms_total;
                best_w = w;
            }
        }
        (void)saved_word;
        ctx->class_tuning[c] = best_w;
        if (chosen) chosen[c - log_min] = best_w;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(dtab);
        (void)hipFree(pay);
        (void)hipFree(keys);
    }
}

} // namespace vo
