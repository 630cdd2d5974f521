// shuffle.hip -- oblivious shuffle: sort on Philox4x32-10 tags.
//
// Replaces the reference's position-map shuffle, a CPU Fisher-Yates over
// std::mt19937_64 (src/data/oblivious_table.cpp:240-255) composed on the
// device (include/cuda/oblivious_shuffle.cuh:10-21).  Row i gets the tag
// Philox4x32-10(counter = (i, 0), key = seed) (Salmon et al., SC'11); the rows
// are sorted by (tag, i) with the oblivious bitonic network, and the payload
// travels with its row, so the permutation is never materialised as
// addresses.  perm_out[j] receives the source row of output row j (without
// perm_out the payload is shuffled alone, as 12-byte records).  With
// n < n_pad live rows, rows >= n get the largest tag and stay in place, so the
// first n rows receive a uniform permutation of the live rows (a position map
// of n records padded to a power of two).
#include "shuffle_kernels.hpp"
#include "resident_tables.hpp"

#include <algorithm>

namespace vo {
namespace {

using namespace shuffle_k;

const std::pair<const void *, BodyInfo> kBodies[] = {VO_SHUFFLE_BODIES(VO_BODY_ENTRY)};
const bool kRegistered = (register_bodies(kBodies, sizeof kBodies / sizeof kBodies[0]), true);

dim3 grid_for(uint64_t n) { return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 65536))); }

} // namespace

void op_shuffle(vo_ctx *ctx, const vo_shuffle_args &a)
{
    VO_CHECK(is_pow2(a.n_pad) && a.n_pad <= (1ull << 32), VO_ERR_INVALID, "n_pad must be a power of two <= 2^32");
    VO_CHECK(a.payload != nullptr, VO_ERR_INVALID, "null payload");
    VO_CHECK(a.n <= a.n_pad, VO_ERR_INVALID, "n above n_pad");
    const uint64_t n = a.n ? a.n : a.n_pad;
    if (a.perm_out == nullptr) {
        uint64_t *keys = (uint64_t *)scratch(ctx, 8 * a.n_pad + 256);
        launch(ctx, LaunchInfo{"shuffle_tags", grid_for(a.n_pad), dim3(256), 0, hash_mix(a.n_pad, 1)}, k_shuffle_keys,
               a.seed, n, a.n_pad, keys, trace_slot(ctx));
        sort_kp_full(ctx, keys, (uint32_t *)a.payload, a.n_pad);
        return;
    }
    uint64_t *keys = (uint64_t *)scratch(ctx, 16 * a.n_pad + 256);
    uint64_t *vals = keys + a.n_pad;
    launch(ctx, LaunchInfo{"shuffle_tags", grid_for(a.n_pad), dim3(256), 0, a.n_pad}, k_shuffle_tags, a.seed,
           (const uint32_t *)a.payload, n, a.n_pad, keys, vals, trace_slot(ctx));
    sort_kv(ctx, keys, vals, a.n_pad);
    launch(ctx, LaunchInfo{"shuffle_unpack", grid_for(a.n_pad), dim3(256), 0, a.n_pad}, k_shuffle_unpack,
           (const uint64_t *)vals, a.n_pad, a.payload, a.perm_out, trace_slot(ctx));
}

} // namespace vo
