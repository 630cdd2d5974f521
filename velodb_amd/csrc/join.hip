// join.hip -- oblivious primary-key / foreign-key join + GROUP BY SUM.
//
// Reference method: the union -> sort -> expand -> compact join of
// obliviator/join/enclave/scalable_oblivious_join.c:159-317, specialised to
// a PK/FK join (every fact row matches at most one dimension row, so the
// m0 x m1 expansion of :211-224 / oblivious_distribute.c is a forward fill),
// followed by the sort + segmented-scan GROUP BY of
// obliviator/tpch_q3/our_3/enclave/scalable_oblivious_join.c:257-321
// (one row per group is kept by the compaction network instead of a second
// sort).  Sums are exact int64.
//
//   1. union of n_dim + n_fact rows padded to n_pad:
//        key = (key32 << 32) | (is_fact << 31) | row,  val = group | value
//   2. oblivious sort of (key, val)
//   3. forward fill of the last dimension row's (key, group); a fact row is
//      matched when its key equals it
//   4. ORCompact of the matched rows' (group, value)
//   5. sort by (group << 32 | position), rows past the match count get
//      group 0xFFFFFFFF
//   6. segmented inclusive sum by group; the last row of each group selected
//   7. ORCompact -> the first G rows are (group, sum), ascending by group.
// With dense_groups (group ids known to lie in [0, max_groups), max_groups
// <= 64), steps 4-7 become one oblivious dense aggregation fused with the
// fill of step 3: every union row updates all G per-thread sums by selects,
// so the matched rows need no compaction.
// Every pass reads and writes every row; only the group count G and the
// match count leave the device.
#include "join_kernels.hpp"
#include "resident_tables.hpp"

#include <algorithm>

namespace vo {
namespace {

using namespace join_k;

const std::pair<const void *, BodyInfo> kBodies[] = {VO_JOIN_BODIES(VO_BODY_ENTRY)};
const bool kRegistered = (register_bodies(kBodies, sizeof kBodies / sizeof kBodies[0]), true);

dim3 grid_for(uint64_t n) { return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 65536))); }

} // namespace

void op_join_groupsum(vo_ctx *ctx, const vo_join_args &a)
{
    const uint64_t n_pad = a.n_pad;
    VO_CHECK(is_pow2(n_pad) && a.n_dim + a.n_fact <= n_pad, VO_ERR_INVALID, "n_pad must be a power of two >= n_dim+n_fact");
    VO_CHECK(n_pad >= (uint64_t)JT, VO_ERR_INVALID, "join needs n_pad >= 4096");
    VO_CHECK(a.n_dim < (1ull << 31) && a.n_fact < (1ull << 31), VO_ERR_INVALID, "row counts must be below 2^31");
    VO_CHECK(a.dim_key && a.dim_group && a.fact_fk && a.fact_val && a.out_group && a.out_sum, VO_ERR_INVALID,
             "null buffer");
    VO_CHECK(!a.dense_groups || (a.max_groups >= 1 && a.max_groups <= 64), VO_ERR_INVALID,
             "dense GROUP BY takes 1 <= max_groups <= 64");
    const uint32_t nt = (uint32_t)(n_pad / JT);
    // arena: keys, vals, col_a, col_b (8 B each), mask (1 B), tile arrays, compaction scratch
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_k = take(n_pad * 8), o_v = take(n_pad * 8), o_a = take(n_pad * 8), o_b = take(n_pad * 8),
                 o_m = take(n_pad), o_fl = take(nt * sizeof(Fill)), o_fc = take(nt * sizeof(Fill)),
                 o_sg = take(nt * sizeof(Seg)), o_sc = take(nt * 8), o_ss = take(nt * sizeof(SegSum)),
                 o_sp = take(nt * sizeof(SegSum)), o_c1 = take(compact_rows2_scratch(n_pad)),
                 o_c2 = take(compact_rows2_scratch(n_pad)), o_part = take((size_t)nt * 64 * 16),
                 o_bad = take((size_t)65536 * 4);
    char *base = (char *)scratch(ctx, off);
    uint64_t *keys = (uint64_t *)(base + o_k), *vals = (uint64_t *)(base + o_v);
    uint64_t *col_a = (uint64_t *)(base + o_a), *col_b = (uint64_t *)(base + o_b);
    uint8_t *mask = (uint8_t *)(base + o_m);
    Fill *tile_last = (Fill *)(base + o_fl), *fcarry = (Fill *)(base + o_fc);
    Seg *seg = (Seg *)(base + o_sg);
    int64_t *scarry = (int64_t *)(base + o_sc);
    SegSum *ssum = (SegSum *)(base + o_ss), *spre = (SegSum *)(base + o_sp);
    uint64_t *cnt1 = (uint64_t *)(base + o_c1), *cnt2 = (uint64_t *)(base + o_c2);

    // 1-2. union and sort: 12-byte records for the dense GROUP BY when the
    // caller wants no sorted union and max_groups <= 63 (public arguments)
    const bool kp = a.dense_groups && a.max_groups <= 63 && !a.sorted_keys && !a.sorted_vals;
    uint32_t *bad = (uint32_t *)(base + o_bad);
    const dim3 bgrid = grid_for(n_pad);
    if (kp) {
        launch(ctx, LaunchInfo{"join_build", bgrid, dim3(256), 0, hash_mix(n_pad, 1)},
               ctx->tracing ? k_join_build_kp<true> : k_join_build_kp<false>, a.dim_key, a.dim_group, a.n_dim,
               a.fact_fk, a.fact_val, a.n_fact, n_pad, keys, (uint32_t *)vals, bad, trace_slot(ctx));
        sort_kp_full(ctx, keys, (uint32_t *)vals, n_pad);
    } else {
        launch(ctx, LaunchInfo{"join_build", bgrid, dim3(256), 0, n_pad},
               ctx->tracing ? k_join_build<true> : k_join_build<false>,
               a.dim_key, a.dim_group,
               a.n_dim, a.fact_fk, a.fact_val, a.n_fact, n_pad, keys, vals, trace_slot(ctx));
        sort_kv(ctx, keys, vals, n_pad);
    }
    if (a.sorted_keys) copy_async(ctx, a.sorted_keys, keys, n_pad * 8);
    if (a.sorted_vals) copy_async(ctx, a.sorted_vals, vals, n_pad * 8);
    // 3. forward fill
    launch(ctx, LaunchInfo{"join_fill_tiles", dim3(nt), dim3(JB), 0, n_pad},
           kp ? (ctx->tracing ? k_fill_tiles<true, true> : k_fill_tiles<false, true>)
              : (ctx->tracing ? k_fill_tiles<true, false> : k_fill_tiles<false, false>),
           (const uint64_t *)keys, (const void *)vals, tile_last, trace_slot(ctx));
    launch(ctx, LaunchInfo{"join_fill_carry", dim3(1), dim3(kScanThreads), 0, nt},
           ctx->tracing ? k_fill_carry<true> : k_fill_carry<false>,
           (const Fill *)tile_last, nt, fcarry, trace_slot(ctx));
    if (a.dense_groups) {
        // 4'-5'. dense GROUP BY fused with the fill: one partial row per tile
        int64_t *part = (int64_t *)(base + o_part);
        const uint32_t G = a.max_groups <= 32 ? 32 : 64;
        const uint32_t n_bad = kp ? bgrid.x : 0u;
        if (G == 32) {
            auto agg = kp ? (ctx->tracing ? k_fill_agg<32, true, true> : k_fill_agg<32, false, true>)
                          : (ctx->tracing ? k_fill_agg<32, true, false> : k_fill_agg<32, false, false>);
            launch(ctx, LaunchInfo{"join_fill_agg", dim3(nt), dim3(JB), 0, hash_mix(n_pad, 32)}, agg,
                   (const uint64_t *)keys, (const void *)vals, (const Fill *)fcarry, part,
                   part + (uint64_t)nt * 32, trace_slot(ctx));
            launch(ctx, LaunchInfo{"group_dense_out", dim3(1), dim3(kOutThreads), 0, a.max_groups}, k_dense_out<32>,
                   (const int64_t *)part, (const int64_t *)(part + (uint64_t)nt * 32), nt, a.max_groups, a.out_group,
                   a.out_sum, a.out_n_groups, (const uint32_t *)bad, n_bad, trace_slot(ctx));
        } else {
            auto agg = kp ? (ctx->tracing ? k_fill_agg<64, true, true> : k_fill_agg<64, false, true>)
                          : (ctx->tracing ? k_fill_agg<64, true, false> : k_fill_agg<64, false, false>);
            launch(ctx, LaunchInfo{"join_fill_agg", dim3(nt), dim3(JB), 0, hash_mix(n_pad, 64)}, agg,
                   (const uint64_t *)keys, (const void *)vals, (const Fill *)fcarry, part,
                   part + (uint64_t)nt * 64, trace_slot(ctx));
            launch(ctx, LaunchInfo{"group_dense_out", dim3(1), dim3(kOutThreads), 0, a.max_groups}, k_dense_out<64>,
                   (const int64_t *)part, (const int64_t *)(part + (uint64_t)nt * 64), nt, a.max_groups, a.out_group,
                   a.out_sum, a.out_n_groups, (const uint32_t *)bad, n_bad, trace_slot(ctx));
        }
        return;
    }
    launch(ctx, LaunchInfo{"join_fill_rows", dim3(nt), dim3(JB), 0, n_pad},
           ctx->tracing ? k_fill_rows<true> : k_fill_rows<false>,
           (const uint64_t *)keys,
           (const uint64_t *)vals, (const Fill *)fcarry, col_a, col_b, mask, trace_slot(ctx));
    // 4. compaction of matched rows
    compact_rows2(ctx, col_a, col_b, mask, n_pad, cnt1);
    // 5. group-by sort
    launch(ctx, LaunchInfo{"group_keys", grid_for(n_pad), dim3(256), 0, n_pad},
           ctx->tracing ? k_group_keys<true> : k_group_keys<false>,
           (const uint64_t *)col_a,
           (const uint64_t *)col_b, (const uint64_t *)cnt1, n_pad, keys, vals, trace_slot(ctx));
    sort_kv(ctx, keys, vals, n_pad);
    // 6. segmented sum
    launch(ctx, LaunchInfo{"group_seg_tiles", dim3(nt), dim3(JB), 0, n_pad},
           ctx->tracing ? k_seg_tiles<true> : k_seg_tiles<false>,
           (const uint64_t *)keys,
           (const uint64_t *)vals, seg, trace_slot(ctx));
    launch(ctx, LaunchInfo{"group_seg_summ", grid_for(nt), dim3(256), 0, nt},
           ctx->tracing ? k_seg_summ<true> : k_seg_summ<false>,
           (const Seg *)seg,
           (const uint64_t *)keys, nt, ssum, trace_slot(ctx));
    launch(ctx, LaunchInfo{"group_seg_carry", dim3(1), dim3(kScanThreads), 0, nt},
           ctx->tracing ? k_seg_carry<true> : k_seg_carry<false>,
           (const SegSum *)ssum,
           nt, spre, scarry, trace_slot(ctx));
    launch(ctx, LaunchInfo{"group_seg_rows", dim3(nt), dim3(JB), 0, n_pad},
           ctx->tracing ? k_seg_rows<true> : k_seg_rows<false>,
           (const uint64_t *)keys,
           (const uint64_t *)vals, (const int64_t *)scarry, n_pad, col_a, col_b, mask, trace_slot(ctx));
    // 7. one row per group
    compact_rows2(ctx, col_a, col_b, mask, n_pad, cnt2);
    launch(ctx, LaunchInfo{"group_out", grid_for(std::max<uint64_t>(a.max_groups, 1)), dim3(256), 0, a.max_groups},

           ctx->tracing ? k_group_out<true> : k_group_out<false>,
           (const uint64_t *)col_a, (const uint64_t *)col_b, (const uint64_t *)cnt2, n_pad, a.max_groups,
           a.out_group, a.out_sum, a.out_n_groups, trace_slot(ctx));
}

} // namespace vo
