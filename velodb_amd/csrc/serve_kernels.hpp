// serve_kernels.hpp -- the serving loop's result checksum of a sorted table
// as a device body plus wrapper, for launch mode (serve.hip) and the
// resident queue kernel.
#pragma once

#include "vo_internal.hpp"

namespace vo {
namespace serve_k {
namespace {

constexpr uint32_t kChecksumBlocks = 256;

// Order-sensitive checksums of a sorted table: sum key[i]*(2i+1), sum
// pay[i]*(i+1) mod 2^64, one partial pair per block (ab: 512 words of LDS).
__device__ __forceinline__ void b_sort_checksum(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ pay,
                                                uint64_t n, uint64_t *__restrict__ partial, uint64_t *ab, uint32_t vb,
                                                uint32_t nvb)
{
    uint64_t *a = ab, *b = ab + 256;
    uint64_t x = 0, y = 0;
    for (uint64_t i = (uint64_t)vb * 256 + threadIdx.x; i < n; i += (uint64_t)nvb * 256) {
        x += keys[i] * (2 * i + 1);
        y += (uint64_t)pay[i] * (i + 1);
    }
    a[threadIdx.x] = x;
    b[threadIdx.x] = y;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (uint32_t)off) {
            a[threadIdx.x] += a[threadIdx.x + off];
            b[threadIdx.x] += b[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * vb] = a[0];
        partial[2 * vb + 1] = b[0];
    }
}

__global__ __launch_bounds__(256) void k_sort_checksum(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ pay,
                                                       uint64_t n, uint64_t *__restrict__ partial)
{
    __shared__ uint64_t ab[512];
    b_sort_checksum(keys, pay, n, partial, ab, blockIdx.x, gridDim.x);
}

} // namespace
} // namespace serve_k
} // namespace vo
