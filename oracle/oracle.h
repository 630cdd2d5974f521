/*
 * oracle.h -- CPU restatement of the reference's oblivious-operator algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (velodb_amd/libvelodb_amd.so) never links,
 * loads or calls anything here, and has no CPU fallback.
 *
 * Every function names the reference file:line it restates.  Pinning of each
 * function (what independent check fixes its output) is listed next to it.
 */
#ifndef VELODB_ORACLE_H
#define VELODB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Seeded generators (workload definition, shared with the device generator) */
/* ------------------------------------------------------------------------ */

/* SplitMix64 output mix (Steele, Lea, Flood, "Fast splittable pseudorandom
 * number generators", OOPSLA 2014).  Pinned: known-answer vector of the
 * published sequence from state 0 (tests/test_oracle.py). */
uint64_t orc_mix64(uint64_t z);

/* i-th output of the SplitMix64 stream whose start state is
 * mix64(seed ^ golden*(stream+1)).  Counter-based, so any i can be produced
 * independently on the device. */
uint64_t orc_rand64(uint64_t seed, uint32_t stream, uint64_t i);

/* Uniform integer in [lo, lo+range) from the high 32 bits of orc_rand64
 * (multiply-shift).  range must be < 2^32. */
int64_t orc_rand_range(uint64_t seed, uint32_t stream, uint64_t i, int64_t lo, uint64_t range);

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11).  Pinned: the Random123 known-answer vectors
 * (tests/test_oracle.py). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Shuffle tag of row i: 64 bits of Philox4x32-10 on counter (i_lo, i_hi, 0, 0)
 * and key (seed_lo, seed_hi). */
uint64_t orc_philox_tag(uint64_t seed, uint64_t i);

/* Sort input (u64 keys, u32 row-index payloads): key[i] = orc_rand64(seed, 0, i), payload[i] = i. */
void orc_gen_sort_input(uint64_t seed, size_t n, uint64_t *keys, uint32_t *payload);

/* ------------------------------------------------------------------------ */
/* Bitonic sort                                                              */
/* ------------------------------------------------------------------------ */

/* Address-stream trace modelled on TraceMem (reference
 * obliviator/join_kks_INT_INT/IncludeL3/trace_mem.h:23-116): every element
 * read and write folds its index into an FNV-1a-64 hash; reads fold i, writes
 * fold -1-i (trace_mem.h:60-85).  NULL disables recording. */
typedef struct orc_trace {
    uint64_t hash;
    uint64_t n_ops;
    uint64_t n_stages;
} orc_trace;

/* Bitonic network with the reference's (k, j, i) schedule
 * (include/cuda/sort.cuh:104-117, src/operator/materialization_operator.cpp:156-169)
 * over (u64 key, u32 payload) records, ascending by key with ties by payload.
 * Rows [n, npad) are padded first with key = UINT64_MAX, payload = row index
 * (sort.cuh:79-100 pads with numeric_limits::max; the payload tie-break is
 * this build's fix of the pad/real tie).  Pinned: equals the unique sorted
 * order of the (key, payload) pairs (numpy lexsort) and the KKS known-answer
 * arrays (obliviator/join_kks_INT_INT/test_sort/test_sorts.cpp:12-44). */
void orc_bitonic_sort_kp(uint64_t *keys, uint32_t *payload, size_t n, size_t npad, orc_trace *trace);

/* Same network over (u64 key, u64 value) records, compared by (key, value). */
void orc_bitonic_sort_kv(uint64_t *keys, uint64_t *vals, size_t npad);

/* Reference data-type ids (include/data/data_type.hpp:12-27). */
enum {
    ORC_T_BOOLEAN = 1, ORC_T_TINYINT = 2, ORC_T_SMALLINT = 3, ORC_T_INTEGER = 4,
    ORC_T_BIGINT = 5, ORC_T_FLOAT = 6, ORC_T_DOUBLE = 7, ORC_T_VARCHAR = 8,
    ORC_T_DATE = 10,
    ORC_T_UBIGINT = 64 /* new: u64 keys (the reference has no unsigned 64-bit type) */
};

/* Multi-column OP_SORT with the reference's exact semantics
 * (include/cuda/sort.cuh:11-118): pads each column of rows [n, npad) with
 * max (ASC) / lowest (DESC) (:79-100), sets indices[i] = i, then runs the
 * (k, j, i) network comparing column by column (needSwap :31-40,
 * compareSingle :11-29) and swapping every column plus the index (swapRows
 * :50-68).  Ties over all columns are not swapped.  Pinned: brute-force
 * equality with a stable sort on unique keys and the ORDER BY known answers
 * of tests/query/test_order_by.cpp:196-272. Returns 0 or -1 on bad args. */
int orc_bitonic_sort_multi(void *const *cols, const uint32_t *types, const uint8_t *asc,
                           size_t ncols, int64_t *indices, size_t n, size_t npad);

/* OP_PERMUTE (include/cuda/permute.cuh:10-22): out[i] = in[idx[i]], i < n. */
void orc_permute(void *out, const void *in, const int64_t *idx, size_t n, size_t elem_size);

/* Multi-threaded bitonic sort of (key, payload) records, restating the
 * recursive parallel bitonic of obliviator/join/enclave/bitonic.c:72-242
 * (bitonic_sort_new: halves sorted in opposite directions on half the
 * threads each, then bitonic_merge: the compare pass at distance
 * prev_pow_two(len) split over the threads, then both halves merged in
 * parallel).  Rows [n, npad) become pads (key max, payload = row) first.
 * Pinned: equal to orc_bitonic_sort_kp (the same unique sorted order) for
 * every thread count.  Used as the multi-core CPU baseline. */
void orc_bitonic_sort_kp_mt(uint64_t *keys, uint32_t *payload, size_t n, size_t npad, int nthreads);

/* ------------------------------------------------------------------------ */
/* Oblivious compaction (ORCompact)                                          */
/* ------------------------------------------------------------------------ */

/* Order-preserving oblivious compaction restating
 * oblivious_compact_elem / oblivious_compact_inner_elem /
 * oblivious_compact_2power_inner_elem
 * (obliviator/join/enclave/oblivious_compact.c:59-153, 327-338):
 * prefix counts, split at pow2_lt(N) (common/util.c:58-65), offset routing
 * network of conditional swaps.  Elements are elem_size opaque bytes.
 * Pinned: the selected prefix equals a stable partition (brute force), the
 * rest is a permutation of the unselected elements, and the whole buffer
 * equals the independent level-by-level closed form orc_compact_levels(). */
void orc_compact(void *buf, size_t elem_size, const uint8_t *selected, size_t n);

/* Closed form of the same network for power-of-two n: for block size
 * s = 2, 4, ..., n and every aligned block [b, b+s), offset o = C[b] mod s
 * (C = exclusive prefix count of selected), and the 2-power merge step of
 * oblivious_compact.c:88-105.  Used to cross-check orc_compact. */
void orc_compact_levels(void *buf, size_t elem_size, const uint8_t *selected, size_t n);

/* ------------------------------------------------------------------------ */
/* Filter + aggregate (Q6-style predicate, exact int128 SUM)                 */
/* ------------------------------------------------------------------------ */

typedef struct orc_filter_spec {
    int32_t day_lo, day_hi;   /* day_lo <= day < day_hi */
    int64_t disc_lo, disc_hi; /* disc_lo <= disc <= disc_hi */
    int64_t qty_lt;           /* qty < qty_lt */
} orc_filter_spec;

/* Fact-table generator: day in [8036,10561], qty in [1,50], price cents in
 * [90000,10495000], disc % in [0,10], streams 1..4 of orc_rand64. */
void orc_gen_fact(uint64_t seed, size_t n, int32_t *day, int64_t *qty, int64_t *price, int64_t *disc);

/* Branch-free mask = AND of the range comparisons, the host predicate of
 * src/operator/seq_scan_operator.cpp:38-45 with the batch AND of
 * src/expression/logical_expression.cpp:61-104.  Rows >= n are 0. */
void orc_filter_mask(const int32_t *day, const int64_t *qty, const int64_t *disc,
                     size_t n, size_t npad, const orc_filter_spec *spec, uint8_t *mask);

/* Full pipeline over npad rows: mask, ORCompact of (price, disc) pairs,
 * exact int128 SUM(price*disc) over the first m compacted rows.
 * Writes the compacted pairs (npad each) when out_price/out_disc != NULL.
 * Returns m; the sum is returned as two 64-bit halves. */
size_t orc_filter_sum(const int32_t *day, const int64_t *qty, const int64_t *price, const int64_t *disc,
                      size_t n, size_t npad, const orc_filter_spec *spec,
                      int64_t *out_price, int64_t *out_disc, uint64_t *sum_lo, int64_t *sum_hi);

/* ------------------------------------------------------------------------ */
/* PK/FK join + GROUP BY                                                     */
/* ------------------------------------------------------------------------ */

/* Dimension: keys are a permutation of 1..n_dim (sort of Philox-free
 * SplitMix tags, ties by index), group uniform in [0, n_groups).
 * Fact: fk uniform in [1, n_dim], value uniform in [vlo, vhi]. */
void orc_gen_join(uint64_t seed, size_t n_dim, size_t n_fact, uint32_t n_groups,
                  int64_t vlo, int64_t vhi,
                  int32_t *dim_key, int32_t *dim_group, int32_t *fact_fk, int64_t *fact_val);

/* Union record sort key: (key << 32) | (is_fact << 31) | row, so that a
 * dimension row precedes the fact rows of its key and every key is unique. */
uint64_t orc_union_key(int32_t key, int is_fact, uint32_t row);

/* The join + group-by pipeline, restating
 * obliviator/join/enclave/scalable_oblivious_join.c:159-317 specialised to a
 * primary-key / foreign-key join (each fact row matches at most one
 * dimension row, so the m0 x m1 expansion degenerates to a forward fill),
 * then the GROUP BY of obliviator/tpch_q3/our_3/enclave/scalable_oblivious_join.c:257-321
 * (sort by group, forward segmented sum, keep one row per group) with exact
 * int64 sums:
 *   1. union of dim and fact records padded to npad = pow2 (pad key = max),
 *   2. bitonic sort by union key,
 *   3. forward fill of the last dimension (key, group) to each fact row,
 *      matched = fact && key equal,
 *   4. ORCompact of matched fact rows -> join rows (group, value),
 *   5. bitonic sort of join rows by (group << 32 | position),
 *   6. segmented inclusive sum by group, select the last row of each group,
 *   7. ORCompact -> first G rows are (group, sum) ascending by group.
 * Intermediates are written when the pointers are non-NULL (each npad long).
 * Returns G (groups present) and fills out_group/out_sum[0..G).
 * Pinned: per-group sums equal a brute-force lookup, and every intermediate
 * equals the numpy closed form (sorted union, fill, stable partition). */
size_t orc_join_groupsum(const int32_t *dim_key, const int32_t *dim_group, size_t n_dim,
                         const int32_t *fact_fk, const int64_t *fact_val, size_t n_fact,
                         size_t npad,
                         uint64_t *sorted_keys, uint64_t *sorted_vals,
                         int32_t *out_group, int64_t *out_sum, size_t max_groups);


/* ------------------------------------------------------------------------ */
/* General equi-join into row-index pairs                                    */
/* ------------------------------------------------------------------------ */

/* Oblivious distribution, restating obliv_distribute
 * (obliviator/join_kks_INT_INT/Include/sort.h:87-126): sort the n elements
 * by target (dummies, target -1, last), extend to m slots of dummies, then
 * for j = prev_pow_two(m) (the largest power of two below m, sort.h:18-22)
 * down to 1 and i = m-j-1 down to 0, move the element at i to i+j when its
 * target >= i+j; the slot it moves to must hold a dummy (the network's
 * invariant, asserted in the reference).  Every valid element ends at its
 * target.  Returns 0, or -1 when the invariant fails (targets not distinct
 * or >= m).  Pinned against the reference's own known-answer program
 * (join_kks_INT_INT/test_sort/test_obliv_distribute.cpp) built in _ref/. */
int orc_distribute(const int64_t *target, const uint64_t *payload, size_t n, size_t m,
                   int64_t *out_target, uint64_t *out_payload);

/* Equi-join of two key columns into row-index pairs, restating
 * scalable_oblivious_join (obliviator/join/enclave/scalable_oblivious_join.c:159-317):
 *   1. union of the left (table_0) and right rows, padded to a power of two,
 *      bitonic sort (the reference sorts by key only, :205; here ties are
 *      broken by (table, row) so that the output order is canonical),
 *   2. forward / backward scans giving m0 / m1 = left / right rows of the
 *      row's key (:211-224),
 *   3. control bits table_0 && m1 > 0 and !table_0 && m0 > 0, two copies
 *      (soj_scan_1 :104-122), ORCompact of each (:251-252),
 *   4. index targets = exclusive prefix sums of m1 (left) and m0 (right)
 *      over the compacted rows (aggregation_tree_i, aggregation_tree.c:339-386),
 *      n_out = their total,
 *   5. oblivious distribution of each copy to n_out slots and forward fill
 *      of the empty slots (:305-314),
 *   6. j_order = (j mod m0) * m1 + j / m0 for the right copy, j its index in
 *      the key's block (aggregation_tree.c:200-221), sort by (key, j_order).
 * Keys are order-preserving 64-bit codes.  Output row i pairs out_left[i]
 * with out_right[i]; pairs come ordered by (key, left row, right row).
 * Returns n_out; the outputs are written only when n_out <= cap.  *status is
 * 0, or -1 when a distribution invariant failed.
 * Pinned: the reference's golden join (obliviator/join/test.txt ->
 * test_output.txt) and brute-force nested loops over random inputs. */
size_t orc_join_pairs(const uint64_t *lkey, size_t n_left, const uint64_t *rkey, size_t n_right,
                      int64_t *out_left, int64_t *out_right, size_t cap, int *status);


/* Oblivious materialization of one column, restating obliviousRecoverSingle
 * (src/operator/materialization_operator.cpp:362-455) with
 * obliviousGatherTyped (:461-517): tuples (record id, position, pairing
 * index), bitonic sort by record id padded with INT64_MAX
 * (bitonicSortTuples :137-173), reads in sorted order with a one-entry cache
 * for repeated ids, and the restore permutation by pairing index.  Positions
 * are the record ids (no position map).  An id outside [0, n_table) reads
 * nothing and gives value 0, found 0 (the reference wraps it modulo the
 * table size, :395).  Values are 64-bit.
 * Pinned: numpy take on random ids with repetitions. */
void orc_gather_recover(const uint64_t *table, size_t n_table, const int64_t *ids, size_t n,
                        uint64_t *out, uint8_t *found);

/* Predicate mask over column-vs-constant terms, restating the host
 * ComparisonExpression::evaluateBatch with a scalar
 * (src/expression/comparison_expression.cpp:13-92: the comparison in the
 * column's type; a NULL row gives NULL with data 0) and the AND / OR of
 * BinaryLogicalExpression::evaluateBatch (src/expression/logical_expression.cpp:61-150:
 * a NULL operand makes the result NULL with data 0), evaluated as the
 * left-deep chain (((t0 op t1) op t2) ...) a parser builds.  mask[i] = data
 * byte of row i (0 for NULL), rows >= n are 0.  types: vo_type ids;
 * cmps: 0..5 in ComparisonType order; consts: bit patterns in the column
 * type; valid[j] may be NULL (no NULLs), else 0 marks a NULL row; conn 0 =
 * AND, 1 = OR.  Returns 0, or -1 for an unsupported type or comparison.
 * Pinned: the reference's test_expression_batch.cpp vectors and the Q6 mask
 * of orc_filter_mask. */
int orc_predicate_mask(const void *const *cols, const uint8_t *const *valid, const uint32_t *types,
                       const uint32_t *cmps, const uint64_t *consts, uint32_t n_terms, uint32_t conn,
                       size_t n, size_t npad, uint8_t *mask);

#ifdef __cplusplus
}
#endif
#endif
