// predicate_kernels.hpp -- the device pass of OP_PREDICATE_MASK (column-vs-constant
// terms joined by AND / OR, NULL rows unselected) as a body plus wrapper,
// for predicate.hip in launch mode and the resident queue kernel.
#pragma once

#include "vo_internal.hpp"

namespace vo {
namespace pred_k {
namespace {

struct PredTerms {
    const void *col[VO_MAX_PREDICATE_TERMS];
    const uint8_t *valid[VO_MAX_PREDICATE_TERMS];
    uint64_t k[VO_MAX_PREDICATE_TERMS];
    uint32_t type[VO_MAX_PREDICATE_TERMS];
    uint32_t cmp[VO_MAX_PREDICATE_TERMS];
};

template <typename T>
__device__ __forceinline__ bool compare(T v, T k, uint32_t cmp)
{
    switch (cmp) {
    case VO_CMP_EQ: return v == k;
    case VO_CMP_NE: return v != k;
    case VO_CMP_LT: return v < k;
    case VO_CMP_LE: return v <= k;
    case VO_CMP_GT: return v > k;
    default: return v >= k;
    }
}

template <typename T>
__device__ __forceinline__ T from_bits(uint64_t b)
{
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, b);
    else if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, (uint32_t)b);
    else if constexpr (sizeof(T) == 2) return __builtin_bit_cast(T, (uint16_t)b);
    else return __builtin_bit_cast(T, (uint8_t)b);
}

// Comparison results of rows i0 .. i0+R-1 (rows >= n are not read).
template <typename T, int R>
__device__ __forceinline__ void term_rows(const void *col, uint64_t kb, uint32_t cmp, uint64_t i0, uint64_t n,
                                          bool (&out)[R])
{
    const T *c = static_cast<const T *>(col);
    const T k = from_bits<T>(kb);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint64_t i = i0 + r;
        out[r] = i < n ? compare<T>(c[i], k, cmp) : false;
    }
}

// R consecutive rows per thread; the type switch is uniform over the grid.
template <int R>
__device__ __forceinline__ void b_predicate_mask(PredTerms t, uint32_t n_terms, uint32_t conj, uint64_t n,
                                                        uint64_t n_pad, uint8_t *__restrict__ out, TraceSlot *slot, uint32_t vb, uint32_t nvb)
{
    dev::Atr tr(slot);
    const uint64_t stride = (uint64_t)nvb * 256 * R;
    for (uint64_t i0 = ((uint64_t)vb * 256 + threadIdx.x) * R; i0 < n_pad; i0 += stride) {
        bool acc[R], any_null[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            acc[r] = conj != 0; // AND starts true, OR starts false
            any_null[r] = false;
        }
        for (uint32_t j = 0; j < n_terms; ++j) {
            bool v[R];
            switch (t.type[j]) {
            case VO_T_BOOLEAN: term_rows<uint8_t, R>(t.col[j], t.k[j], t.cmp[j], i0, n, v); break;
            case VO_T_TINYINT: term_rows<int8_t, R>(t.col[j], t.k[j], t.cmp[j], i0, n, v); break;
            case VO_T_SMALLINT: term_rows<int16_t, R>(t.col[j], t.k[j], t.cmp[j], i0, n, v); break;
            case VO_T_INTEGER: term_rows<int32_t, R>(t.col[j], t.k[j], t.cmp[j], i0, n, v); break;
            case VO_T_DATE: term_rows<uint32_t, R>(t.col[j], t.k[j], t.cmp[j], i0, n, v); break;
            case VO_T_BIGINT: term_rows<int64_t, R>(t.col[j], t.k[j], t.cmp[j], i0, n, v); break;
            case VO_T_FLOAT: term_rows<float, R>(t.col[j], t.k[j], t.cmp[j], i0, n, v); break;
            case VO_T_DOUBLE: term_rows<double, R>(t.col[j], t.k[j], t.cmp[j], i0, n, v); break;
            default: term_rows<uint64_t, R>(t.col[j], t.k[j], t.cmp[j], i0, n, v); break; // VARCHAR, UBIGINT
            }
            const uint8_t *vb = t.valid[j];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint64_t i = i0 + r;
                if (vb) any_null[r] = any_null[r] | (i < n && vb[i] == 0);
                acc[r] = conj ? (acc[r] & v[r]) : (acc[r] | v[r]);
            }
        }
        uint32_t word = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool sel = (i0 + r < n) & acc[r] & !any_null[r];
            word |= (uint32_t)sel << (8 * r);
        }
        if constexpr (R == 4) {
            *reinterpret_cast<uint32_t *>(out + i0) = word; // n_pad is a multiple of 4 on this path
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (i0 + r < n_pad) out[i0 + r] = (uint8_t)(word >> (8 * r));
        }
        tr.at(1, i0);
    }
    tr.done(dev::gtid());
}

template <int R>
__global__ __launch_bounds__(256) void k_predicate_mask(PredTerms t, uint32_t n_terms, uint32_t conj, uint64_t n,
                                                        uint64_t n_pad, uint8_t *__restrict__ out, TraceSlot *slot)
{
    b_predicate_mask<R>(t, n_terms, conj, n, n_pad, out, slot, blockIdx.x, gridDim.x);
}

} // namespace
} // namespace pred_k
} // namespace vo
