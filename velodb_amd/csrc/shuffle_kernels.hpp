// shuffle_kernels.hpp -- Philox4x32-10 tag generation and unpacking of the
// oblivious shuffle (see shuffle.hip) as bodies plus wrappers, for launch
// mode and the resident queue kernel.
#pragma once

#include "vo_internal.hpp"

namespace vo {
namespace shuffle_k {
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

__device__ __forceinline__ void b_shuffle_tags(uint64_t seed, const uint32_t *__restrict__ payload, uint64_t n, uint64_t n_pad,
                               uint64_t *__restrict__ keys, uint64_t *__restrict__ vals, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::Atr tr(slot);
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n_pad; i += (uint64_t)nvb * 256) {
        const uint64_t t = philox_tag(seed, i);
        keys[i] = i < n ? t : ~0ull; // a tie with a live tag of ~0 breaks by row: live rows first
        vals[i] = (i << 32) | payload[i];
        tr.at(1, i);
    }
    tr.done(dev::gtid());
}

__global__ void k_shuffle_tags(uint64_t seed, const uint32_t *__restrict__ payload, uint64_t n, uint64_t n_pad,
                               uint64_t *__restrict__ keys, uint64_t *__restrict__ vals, TraceSlot *slot)
{
    b_shuffle_tags(seed, payload, n, n_pad, keys, vals, slot, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void b_shuffle_unpack(const uint64_t *__restrict__ vals, uint64_t n_pad, uint32_t *__restrict__ payload,
                                 uint32_t *__restrict__ perm_out, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::Atr tr(slot);
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n_pad; i += (uint64_t)nvb * 256) {
        const uint64_t v = vals[i];
        payload[i] = (uint32_t)v;
        if (perm_out) perm_out[i] = (uint32_t)(v >> 32);
        tr.at(1, i);
    }
    tr.done(dev::gtid());
}

__global__ void k_shuffle_unpack(const uint64_t *__restrict__ vals, uint64_t n_pad, uint32_t *__restrict__ payload,
                                 uint32_t *__restrict__ perm_out, TraceSlot *slot)
{
    b_shuffle_unpack(vals, n_pad, payload, perm_out, slot, blockIdx.x, gridDim.x);
}

// Payload-only form (no perm_out): the payload column itself rides as the
// u32 payload of (u64 key, u32 payload) records, 12 bytes per row instead of
// 16, sorted in place.  Live rows get key = tag >> 1, pad rows 2^63 | row, so
// pads stay in place after every live row.  The order is the (tag, row) order
// of the 16-byte form unless two live tags share their top 63 bits (a tie
// then falls to the payload values).
__device__ __forceinline__ void b_shuffle_keys(uint64_t seed, uint64_t n, uint64_t n_pad, uint64_t *__restrict__ keys, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::Atr tr(slot);
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n_pad; i += (uint64_t)nvb * 256) {
        keys[i] = i < n ? (philox_tag(seed, i) >> 1) : ((1ull << 63) | i);
        tr.at(1, i);
    }
    tr.done(dev::gtid());
}

__global__ void k_shuffle_keys(uint64_t seed, uint64_t n, uint64_t n_pad, uint64_t *__restrict__ keys, TraceSlot *slot)
{
    b_shuffle_keys(seed, n, n_pad, keys, slot, blockIdx.x, gridDim.x);
}

} // namespace
} // namespace shuffle_k
} // namespace vo
