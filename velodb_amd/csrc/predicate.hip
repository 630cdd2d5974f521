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
#include "predicate_kernels.hpp"
#include "resident_tables.hpp"

#include <algorithm>

namespace vo {
namespace {

using namespace pred_k;

const std::pair<const void *, BodyInfo> kBodies[] = {VO_PRED_BODIES(VO_BODY_ENTRY)};
const bool kRegistered = (register_bodies(kBodies, sizeof kBodies / sizeof kBodies[0]), true);

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
