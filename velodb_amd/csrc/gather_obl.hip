// gather_obl.hip -- oblivious gather of table values by record id, the
// device form of the host oblivious materialization of
// src/operator/materialization_operator.cpp:362-517 (obliviousRecoverSingle:
// bitonic sort of (record id, position, pairing index) tuples, reads in
// sorted order with a dummy read per duplicate, restore by the pairing
// index).
//
// Here no address depends on a record id:
//   1. union of the table rows (key = row << 33) and the requests
//      (key = id << 33 | 1 << 32 | request index), padded to N = 2^k;
//   2. sort by key: each request follows the table row it asks for (values
//      of at most 4 bytes ride as the u32 payload of 12-byte records, wider
//      ones as a u64 column);
//   3. forward fill of the last table row (row, value) to the requests;
//   4. ORCompact of the request rows: the first n rows hold
//      (request index << 1 | found, value) in record-id order;
//   5. sort of those n rows (padded to 2^k) by request index, as
//      (u64 key, u32 payload) records: the value's low word rides below the
//      index in the key and its high word is the payload.
// Every pass reads and writes every row of its padded size; the record ids
// and their repetitions never reach an address.
#include "gather_obl_kernels.hpp"
#include "resident_tables.hpp"

#include <algorithm>

namespace vo {
namespace {

using namespace og_k;

const std::pair<const void *, BodyInfo> kBodies[] = {VO_OG_BODIES(VO_BODY_ENTRY)};
const bool kRegistered = (register_bodies(kBodies, sizeof kBodies / sizeof kBodies[0]), true);

dim3 grid_for(uint64_t n) { return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 65536))); }

} // namespace

void op_gather_oblivious(vo_ctx *ctx, const vo_gather_oblivious_args &a)
{
    const uint32_t eb = a.elem_bytes;
    VO_CHECK(eb == 1 || eb == 2 || eb == 4 || eb == 8, VO_ERR_INVALID, "elem_bytes must be 1, 2, 4 or 8");
    VO_CHECK(a.n_table < (1ull << 31) && a.n < (1ull << 32), VO_ERR_INVALID, "table below 2^31 rows, requests below 2^32");
    VO_CHECK((a.table || a.n_table == 0) && (a.record_ids || a.n == 0) && (a.out || a.n == 0), VO_ERR_INVALID,
             "null buffer");
    if (a.n == 0) return;
    uint64_t n_pad = GT;
    while (n_pad < a.n_table + a.n) n_pad <<= 1;
    uint64_t m_pad = 1;
    while (m_pad < a.n) m_pad <<= 1;
    VO_CHECK(n_pad <= (1ull << 29), VO_ERR_INVALID, "table + requests above 2^29 rows");
    const uint32_t nt = (uint32_t)(n_pad / GT);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_k = take(n_pad * 8), o_v = take(n_pad * 8), o_c0 = take(n_pad * 8), o_c1 = take(n_pad * 8),
                 o_m = take(n_pad), o_t = take(nt * sizeof(Last)), o_c = take(nt * sizeof(Last)),
                 o_cnt = take(compact_rows2_scratch(n_pad));
    char *base = (char *)scratch(ctx, off);
    uint64_t *keys = (uint64_t *)(base + o_k), *vals = (uint64_t *)(base + o_v);
    uint64_t *c0 = (uint64_t *)(base + o_c0), *c1 = (uint64_t *)(base + o_c1);
    uint8_t *mask = (uint8_t *)(base + o_m);
    Last *tiles = (Last *)(base + o_t), *carry = (Last *)(base + o_c);
    uint64_t *cnt = (uint64_t *)(base + o_cnt);
    const bool tr = ctx->tracing;
    // values of at most 4 bytes (a public width) ride as the u32 payload of
    // 12-byte union records
    const bool kp = eb <= 4;
    auto pick = [&](auto f0, auto f1, auto k0, auto k1) { return kp ? (tr ? k1 : k0) : (tr ? f1 : f0); };
    launch(ctx, LaunchInfo{"gather_obl_build", grid_for(n_pad), dim3(256), 0, hash_mix(hash_mix(eb, a.n_table), a.n)},
           pick(k_og_build<false, false>, k_og_build<true, false>, k_og_build<false, true>, k_og_build<true, true>),
           a.table, a.n_table, eb, a.record_ids, a.n, n_pad, keys, (void *)vals, trace_slot(ctx));
    if (kp) sort_kp_full(ctx, keys, reinterpret_cast<uint32_t *>(vals), n_pad);
    else sort_kv(ctx, keys, vals, n_pad);
    launch(ctx, LaunchInfo{"gather_obl_tiles", dim3(nt), dim3(GB), 0, n_pad},
           pick(k_og_tiles<false, false>, k_og_tiles<true, false>, k_og_tiles<false, true>, k_og_tiles<true, true>),
           (const uint64_t *)keys, (const void *)vals, tiles, trace_slot(ctx));
    launch(ctx, LaunchInfo{"gather_obl_carry", dim3(1), dim3(kScanThreads), 0, nt},
           tr ? k_og_carry<true> : k_og_carry<false>, (const Last *)tiles, nt, carry, trace_slot(ctx));
    launch(ctx, LaunchInfo{"gather_obl_rows", dim3(nt), dim3(GB), 0, n_pad},
           pick(k_og_rows<false, false>, k_og_rows<true, false>, k_og_rows<false, true>, k_og_rows<true, true>),
           (const uint64_t *)keys, (const void *)vals, (const Last *)carry, c0, c1, mask, trace_slot(ctx));
    compact_rows2(ctx, c0, c1, mask, n_pad, cnt);
    // the first n rows are the requests; reuse keys (and vals as the u32
    // payload column) for the final sort
    uint32_t *hi = reinterpret_cast<uint32_t *>(vals);
    launch(ctx, LaunchInfo{"gather_obl_order", grid_for(m_pad), dim3(256), 0, hash_mix(a.n, m_pad)},
           tr ? k_og_order<true> : k_og_order<false>, (const uint64_t *)c0, (const uint64_t *)c1, a.n, m_pad, keys,
           hi, trace_slot(ctx));
    sort_kp_full(ctx, keys, hi, m_pad);
    launch(ctx, LaunchInfo{"gather_obl_out", grid_for(a.n), dim3(256), 0, hash_mix(a.n, eb)},
           tr ? k_og_out<true> : k_og_out<false>, (const uint64_t *)keys, (const uint32_t *)hi, a.n, eb, a.out,
           a.out_found, trace_slot(ctx));
}

} // namespace vo
