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
#include "vo_internal.hpp"

#include <algorithm>

namespace vo {
namespace {

__device__ __forceinline__ uint64_t philox_tag(uint64_t seed, uint64_t i)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = (uint32_t)i, c1 = (uint32_t)(i >> 32), c2 = 0, c3 = 0;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
        const uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    return ((uint64_t)c1 << 32) | c0;
}

__global__ void k_shuffle_tags(uint64_t seed, const uint32_t *__restrict__ payload, uint64_t n, uint64_t n_pad,
                               uint64_t *__restrict__ keys, uint64_t *__restrict__ vals, TraceSlot *slot)
{
    dev::Atr tr(slot);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pad; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t t = philox_tag(seed, i);
        keys[i] = i < n ? t : ~0ull; // a tie with a live tag of ~0 breaks by row: live rows first
        vals[i] = (i << 32) | payload[i];
        tr.at(1, i);
    }
    tr.done(dev::gtid());
}

__global__ void k_shuffle_unpack(const uint64_t *__restrict__ vals, uint64_t n_pad, uint32_t *__restrict__ payload,
                                 uint32_t *__restrict__ perm_out, TraceSlot *slot)
{
    dev::Atr tr(slot);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pad; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = vals[i];
        payload[i] = (uint32_t)v;
        if (perm_out) perm_out[i] = (uint32_t)(v >> 32);
        tr.at(1, i);
    }
    tr.done(dev::gtid());
}

// Payload-only form (no perm_out): the payload column itself rides as the
// u32 payload of (u64 key, u32 payload) records, 12 bytes per row instead of
// 16, sorted in place.  Live rows get key = tag >> 1, pad rows 2^63 | row, so
// pads stay in place after every live row.  The order is the (tag, row) order
// of the 16-byte form unless two live tags share their top 63 bits (a tie
// then falls to the payload values).
__global__ void k_shuffle_keys(uint64_t seed, uint64_t n, uint64_t n_pad, uint64_t *__restrict__ keys, TraceSlot *slot)
{
    dev::Atr tr(slot);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pad; i += (uint64_t)gridDim.x * blockDim.x) {
        keys[i] = i < n ? (philox_tag(seed, i) >> 1) : ((1ull << 63) | i);
        tr.at(1, i);
    }
    tr.done(dev::gtid());
}

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
