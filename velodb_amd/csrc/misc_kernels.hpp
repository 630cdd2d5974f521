// misc_kernels.hpp -- the reference's element-wise device ops as device
// bodies plus __global__ wrappers (used by misc.hip in launch mode and by
// the resident queue kernel).
//   OP_PERMUTE             include/cuda/permute.cuh:10-42
//   OP_GATHER              include/cuda/gather.cuh:11-45
//   OP_COMPARE_EQ          include/cuda/compare.cuh:25-91
//   OP_COMPOSE_POSITION_MAP / OP_ENRICH_POSITIONS
//                          include/cuda/oblivious_shuffle.cuh:10-34
#pragma once

#include "vo_internal.hpp"

namespace vo {
namespace misc_k {
namespace {

constexpr uint32_t EB = 256;

// Element-wise bodies take their virtual block and grid explicitly (the
// resident queue kernel runs them over its own blocks); every launch uses
// 256-thread blocks.
template <typename T>
__device__ __forceinline__ void b_permute(T *__restrict__ out, const T *__restrict__ in,
                                          const int64_t *__restrict__ idx, uint64_t n, uint32_t vb, uint32_t nvb)
{
    for (uint64_t i = (uint64_t)vb * EB + threadIdx.x; i < n; i += (uint64_t)nvb * EB) out[i] = in[idx[i]];
}

template <typename T>
__global__ __launch_bounds__(EB) void k_permute(T *__restrict__ out, const T *__restrict__ in,
                                                const int64_t *__restrict__ idx, uint64_t n)
{
    b_permute<T>(out, in, idx, n, blockIdx.x, gridDim.x);
}

template <typename T>
__device__ __forceinline__ void b_gather(T *__restrict__ out, const T *__restrict__ in, const int32_t *__restrict__ pos,
                                         const uint8_t *__restrict__ mask, uint64_t n, uint32_t vb, uint32_t nvb)
{
    for (uint64_t i = (uint64_t)vb * EB + threadIdx.x; i < n; i += (uint64_t)nvb * EB) {
        const bool keep = (mask == nullptr) || (mask[i] != 0);
        if (keep) out[pos[i]] = in[i];
    }
}

template <typename T>
__global__ __launch_bounds__(EB) void k_gather(T *__restrict__ out, const T *__restrict__ in,
                                               const int32_t *__restrict__ pos, const uint8_t *__restrict__ mask,
                                               uint64_t n)
{
    b_gather<T>(out, in, pos, mask, n, blockIdx.x, gridDim.x);
}

template <typename T>
__device__ __forceinline__ void b_compare_eq(const T *__restrict__ a, const T *__restrict__ b, uint8_t *__restrict__ out,
                                             uint64_t n, uint32_t vb, uint32_t nvb)
{
    for (uint64_t i = (uint64_t)vb * EB + threadIdx.x; i < n; i += (uint64_t)nvb * EB) out[i] = a[i] == b[i] ? 1 : 0;
}

template <typename T>
__global__ __launch_bounds__(EB) void k_compare_eq(const T *__restrict__ a, const T *__restrict__ b,
                                                   uint8_t *__restrict__ out, uint64_t n)
{
    b_compare_eq<T>(a, b, out, n, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void b_compose(const uint32_t *__restrict__ old_map, const uint32_t *__restrict__ sigma,
                                          uint32_t *__restrict__ new_map, uint64_t n, uint32_t vb, uint32_t nvb)
{
    for (uint64_t i = (uint64_t)vb * EB + threadIdx.x; i < n; i += (uint64_t)nvb * EB) new_map[i] = old_map[sigma[i]];
}

__global__ __launch_bounds__(EB) void k_compose(const uint32_t *__restrict__ old_map, const uint32_t *__restrict__ sigma,
                                                uint32_t *__restrict__ new_map, uint64_t n)
{
    b_compose(old_map, sigma, new_map, n, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void b_enrich(const uint32_t *__restrict__ rid, const uint32_t *__restrict__ map,
                                         uint32_t *__restrict__ pos, uint64_t n, uint32_t vb, uint32_t nvb)
{
    for (uint64_t i = (uint64_t)vb * EB + threadIdx.x; i < n; i += (uint64_t)nvb * EB) pos[i] = map[rid[i]];
}

__global__ __launch_bounds__(EB) void k_enrich(const uint32_t *__restrict__ rid, const uint32_t *__restrict__ map,
                                               uint32_t *__restrict__ pos, uint64_t n)
{
    b_enrich(rid, map, pos, n, blockIdx.x, gridDim.x);
}

} // namespace
} // namespace misc_k
} // namespace vo
