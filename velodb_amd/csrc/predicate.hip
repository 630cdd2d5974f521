// predicate.hip -- WHERE predicates evaluated on the device into a byte mask.
//
// The reference evaluates predicates on the host, row batch by row batch:
// ComparisonExpression::evaluateBatch compares a column with a constant
// (src/expression/comparison_expression.cpp:13-92) and
// BinaryLogicalExpression::evaluateBatch joins the comparisons with AND / OR
// (src/expression/logical_expression.cpp:61-150); the resulting BOOLEAN column
// is the mask FilterCompactionOperator compacts by.  Here the same mask is
// computed in one device pass over the columns, so the predicate runs where
// the columns live and only the mask (device memory) feeds the compaction.
//
// Semantics kept from the reference:
//   - comparisons EQ, NE, LT, LE, GT, GE in the column's own type (signed,
//     unsigned or IEEE floating point as the type is), against a constant of
//     that type;
//   - a term whose row is NULL gives NULL, and AND / OR with a NULL operand
//     give NULL (:61-150 set the result to NULL with data 0 whenever an
//     operand is NULL); a NULL row is not selected, so its mask byte is 0.
// All terms of one call are joined by one connective (a flat AND chain, as
// the reference's fused path evaluates, or a flat OR chain).
//
// Every row of every term column is read and every mask byte written: the
// address stream depends on n and the terms' types only.
#include "vo_internal.hpp"

#include <algorithm>

namespace vo {
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
__global__ __launch_bounds__(256) void k_predicate_mask(PredTerms t, uint32_t n_terms, uint32_t conj, uint64_t n,
                                                        uint64_t n_pad, uint8_t *__restrict__ out, TraceSlot *slot)
{
    dev::Atr tr(slot);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * R;
    for (uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * R; i0 < n_pad; i0 += stride) {
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

} // namespace

void op_predicate_mask(vo_ctx *ctx, const vo_predicate_args &a)
{
    VO_CHECK(a.n_terms >= 1 && a.n_terms <= VO_MAX_PREDICATE_TERMS, VO_ERR_INVALID, "1..8 predicate terms");
    VO_CHECK(a.terms != nullptr, VO_ERR_INVALID, "null terms");
    VO_CHECK(a.connective == VO_CONN_AND || a.connective == VO_CONN_OR, VO_ERR_INVALID, "unknown connective");
    VO_CHECK(a.n <= a.n_pad, VO_ERR_INVALID, "n above n_pad");
    if (a.n_pad == 0) return;
    VO_CHECK(a.out_mask != nullptr, VO_ERR_INVALID, "null mask");
    PredTerms t{};
    for (uint32_t j = 0; j < a.n_terms; ++j) {
        const vo_predicate_term &s = a.terms[j];
        switch (s.type_id) {
        case VO_T_BOOLEAN: case VO_T_TINYINT: case VO_T_SMALLINT: case VO_T_INTEGER: case VO_T_DATE:
        case VO_T_BIGINT: case VO_T_FLOAT: case VO_T_DOUBLE: case VO_T_VARCHAR: case VO_T_UBIGINT: break;
        default: throw Error{VO_ERR_UNSUPPORTED, "predicate term " + std::to_string(j) + ": unsupported column type"};
        }
        VO_CHECK(s.cmp <= VO_CMP_GE, VO_ERR_INVALID, "predicate term " + std::to_string(j) + ": unknown comparison");
        VO_CHECK(s.col != nullptr || a.n == 0, VO_ERR_INVALID, "predicate term " + std::to_string(j) + ": null column");
        t.col[j] = s.col;
        t.valid[j] = s.valid;
        t.k[j] = s.constant;
        t.type[j] = s.type_id;
        t.cmp[j] = s.cmp;
    }
    uint64_t h = hash_mix(a.n_pad, a.n_terms);
    for (uint32_t j = 0; j < a.n_terms; ++j) h = hash_mix(h, ((uint64_t)t.type[j] << 1) | (t.valid[j] != nullptr));
    constexpr uint32_t kBlock = 256;
    // 4 rows and one 32-bit mask store per thread when n_pad is a multiple of
    // 4 and the mask starts on a 4-byte boundary; else 1 row and byte stores.
    const bool words = a.n_pad % 4 == 0 && reinterpret_cast<uintptr_t>(a.out_mask) % 4 == 0;
    const uint32_t rows = words ? 4 : 1;
    const uint64_t threads = a.n_pad / rows;
    const dim3 grid((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((threads + kBlock - 1) / kBlock, 1u << 16)));
    if (rows == 4)
        launch(ctx, LaunchInfo{"predicate_mask", grid, dim3(kBlock), 0, h}, k_predicate_mask<4>, t, a.n_terms,
               (uint32_t)(a.connective == VO_CONN_AND), a.n, a.n_pad, a.out_mask, trace_slot(ctx));
    else
        launch(ctx, LaunchInfo{"predicate_mask", grid, dim3(kBlock), 0, h}, k_predicate_mask<1>, t, a.n_terms,
               (uint32_t)(a.connective == VO_CONN_AND), a.n, a.n_pad, a.out_mask, trace_slot(ctx));
}

} // namespace vo
