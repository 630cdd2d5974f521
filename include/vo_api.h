/*
 * vo_api.h -- C ABI of the MI355X (gfx950) oblivious-operator library.
 *
 * This is the drop-in boundary for VelODB's device runtime: the
 * TaskManager / Command interface that the reference's operators call
 * (include/cuda/task_manager.hpp:12-30, include/cuda/commands.hpp:16-201).
 * Everything crossing it is plain C: pointers, sizes, POD structs.
 *
 * Ownership (as in the reference, src/operator/sort_operator.cpp:54-56):
 * callers own every device buffer they pass.  Buffers named "n_pad" long
 * must hold at least n_pad elements; the library writes the pad region
 * [n, n_pad) exactly as include/cuda/sort.cuh:79-100 does.  Temporary device
 * memory (scan partials, union arrays, trace slots) comes from a scratch
 * arena owned by the context and grown on demand.
 *
 * Errors: every call returns a vo_status (0 = VO_OK).  vo_last_error()
 * gives the message of the context's last failure.  Device faults surface
 * through the stream status, as src/cuda/commands.cpp:78-82 does, and waits
 * carry a timeout (the reference waits forever).
 *
 * Threading: vo_submit is safe to call from several host threads on one
 * context (the ring is guarded; the reference ring is single-producer,
 * src/cuda/commands.cpp:47-67).  Commands complete in submission order on one
 * context; separate contexts run concurrently on separate HIP streams.
 */
#ifndef VO_API_H
#define VO_API_H

#include <stddef.h>
#include <stdint.h>

#if defined(VO_BUILD)
#define VO_EXPORT __attribute__((visibility("default")))
#else
#define VO_EXPORT
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define VO_API_VERSION 1
#define VO_MAX_SORT_COLUMNS 8 /* include/common/constants.hpp.in:17 */

typedef enum vo_status {
    VO_OK = 0,
    VO_ERR_INVALID = 1,     /* bad argument (size, alignment, type id) */
    VO_ERR_HIP = 2,         /* HIP runtime error or device fault */
    VO_ERR_TIMEOUT = 3,     /* vo_wait / vo_queue_stop timed out */
    VO_ERR_NOMEM = 4,       /* device or pinned allocation failed */
    VO_ERR_UNSUPPORTED = 5, /* opcode or type the library does not implement */
    VO_ERR_QUEUE = 6,       /* command queue not running / stopped / full */
    VO_ERR_ADMISSION = 7    /* serving loop: request exceeds the HBM budget */
} vo_status;

/* Column type ids: the reference's DataTypeId values
 * (include/data/data_type.hpp:12-27, device types include/data/type_traits.hpp:10-19),
 * plus VO_T_UBIGINT for the unsigned 64-bit keys the reference lacks. */
typedef enum vo_type {
    VO_T_BOOLEAN = 1,  /* uint8_t  */
    VO_T_TINYINT = 2,  /* int8_t   */
    VO_T_SMALLINT = 3, /* int16_t  */
    VO_T_INTEGER = 4,  /* int32_t  */
    VO_T_BIGINT = 5,   /* int64_t  */
    VO_T_FLOAT = 6,    /* float    */
    VO_T_DOUBLE = 7,   /* double   */
    VO_T_VARCHAR = 8,  /* size_t ordinal */
    VO_T_DATE = 10,    /* uint32_t */
    VO_T_UBIGINT = 64  /* uint64_t (new) */
} vo_type;

/* Opcodes.  Values 0..12 and 0xFFFFFFFF are the reference's OpCode
 * (include/cuda/commands.hpp:16-31); 0x100+ are the oblivious operators. */
typedef enum vo_opcode {
    VO_OP_NOP = 0,
    VO_OP_SCATTER = 1,                  /* compaction positions (stable prefix count) */
    VO_OP_GATHER = 2,                   /* out[pos[i]] = in[i] if mask (data-dependent writes) */
    VO_OP_SORT = 3,                     /* multi-column bitonic sort, reference semantics */
    VO_OP_PERMUTE = 4,                  /* out[i] = in[idx[i]] (data-dependent reads) */
    VO_OP_SORT_MERGE_JOIN_COUNT = 5,    /* oblivious equi-join: sort, count the pairs (vo_join_pairs_args) */
    VO_OP_SORT_MERGE_JOIN_WRITE = 6,    /* oblivious equi-join: write the row-index pairs */
    VO_OP_HASH_JOIN_BUILD = 7,          /* unsupported: non-oblivious on the device */
    VO_OP_HASH_JOIN_COUNT = 8,          /* unsupported */
    VO_OP_HASH_JOIN_WRITE = 9,          /* unsupported */
    VO_OP_COMPARE_EQ = 10,              /* out_mask[i] = a[i] == b[i] */
    VO_OP_COMPOSE_POSITION_MAP = 11,    /* new_map[i] = old_map[sigma[i]] */
    VO_OP_ENRICH_POSITIONS = 12,        /* positions[i] = map[record_ids[i]] */
    VO_OP_SORT_KP = 0x100,              /* (u64 key, u32 payload) oblivious sort */
    VO_OP_COMPACT = 0x101,              /* ORCompact routing-network compaction */
    VO_OP_SHUFFLE = 0x102,              /* Philox-tag oblivious shuffle */
    VO_OP_FILTER_SUM = 0x103,           /* predicate + compaction + int128 SUM */
    VO_OP_PKFK_JOIN_GROUPSUM = 0x104,   /* PK/FK join + GROUP BY SUM */
    VO_OP_SORT_BATCH = 0x105,           /* many independent (u64,u32) sorts */
    VO_OP_GATHER_OBLIVIOUS = 0x106,     /* table values by record id, no data-dependent address */
    VO_OP_PREDICATE_MASK = 0x107,       /* WHERE terms (column vs constant, AND / OR) -> u8 mask */
    VO_OP_COMPACT_BATCH = 0x108,        /* many independent ORCompact compactions */
    VO_OP_TERMINATE = 0xFFFFFFFFu
} vo_opcode;

/* ------------------------------------------------------------------ args */

/* OP_SORT: mirror of CommandArgs::SortArgs (include/cuda/commands.hpp:64-72).
 * Sorts the columns in place over n_padded_rows rows and writes
 * indices[i] = original row, with the reference's semantics
 * (include/cuda/sort.cuh:11-118), bit-exact:
 *   - pad rows [n_rows, n_padded_rows) are filled with numeric max (ASC
 *     column) or lowest (DESC column) first (sort.cuh:79-100);
 *   - rows equal in every key column are not exchanged: ties, including a
 *     real row whose keys equal the pad values, are decided by the bitonic
 *     network, so a pad row can precede such a real row and the first n_rows
 *     indices need not all be < n_rows (as in the reference);
 *   - FLOAT/DOUBLE: -0.0 and +0.0 compare equal and every value keeps its own
 *     bits (the sign of a zero travels with its row).  With more than
 *     32 - log2(n_padded_rows) FLOAT/DOUBLE key columns the call returns
 *     VO_ERR_UNSUPPORTED;
 *   - NaN (deviation): the reference's compareSingle has no order for NaN
 *     (it compares neither equal nor less); here a NaN takes its IEEE
 *     total-order place, above +inf (below -inf when its sign bit is set). */
typedef struct vo_sort_args {
    void *sort_cols[VO_MAX_SORT_COLUMNS];
    int64_t *indices;
    uint64_t n_sort_columns;
    uint64_t n_rows;
    uint64_t n_padded_rows;
    uint32_t col_types[VO_MAX_SORT_COLUMNS];
    uint8_t ascending_flags[VO_MAX_SORT_COLUMNS];
} vo_sort_args;

/* OP_SORT_KP: records (keys[i], payload[i]); ascending by key, ties by
 * payload.  Rows [n, n_pad) are set to (UINT64_MAX, i) first. */
typedef struct vo_sort_kp_args {
    uint64_t *keys;
    uint32_t *payload;
    uint64_t n;
    uint64_t n_pad;
} vo_sort_kp_args;

/* OP_PERMUTE: mirror of CommandArgs::PermuteArgs (commands.hpp:74-80). */
typedef struct vo_permute_args {
    void *out_data;
    const void *in_data;
    const int64_t *in_indices;
    uint64_t n;
    uint32_t type_id;
} vo_permute_args;

/* OP_SCATTER: mirror of CommandArgs::ScatterArgs (commands.hpp:48-53).
 * out_indices[i] = number of selected rows before i (for selected i);
 * *out_count = total.  Stable, unlike the reference's per-tile atomicAdd. */
typedef struct vo_scatter_args {
    int32_t *out_indices;
    uint64_t *out_count;
    const uint8_t *in_mask;
    uint64_t n;
} vo_scatter_args;

/* OP_GATHER: mirror of CommandArgs::GatherArgs (commands.hpp:55-62). */
typedef struct vo_gather_args {
    void *out_data;
    const void *in_data;
    const int32_t *in_indices;
    const uint8_t *in_mask;
    uint64_t n;
    uint32_t type_id;
} vo_gather_args;

/* OP_COMPARE_EQ: mirror of CommandArgs::CompareEqArgs (commands.hpp:129-135). */
typedef struct vo_compare_eq_args {
    const void *left_data;
    const void *right_data;
    uint8_t *out_mask;
    uint64_t n;
    uint32_t type_id;
} vo_compare_eq_args;

/* OP_COMPOSE_POSITION_MAP / OP_ENRICH_POSITIONS (commands.hpp:137-149). */
typedef struct vo_compose_args {
    const uint32_t *old_map;
    const uint32_t *sigma;
    uint32_t *new_map;
    uint64_t n;
} vo_compose_args;

typedef struct vo_enrich_args {
    const uint32_t *record_ids;
    const uint32_t *position_map;
    uint32_t *positions;
    uint64_t n;
} vo_enrich_args;

/* OP_COMPACT: order-preserving oblivious compaction of up to 8 columns of
 * n_pad rows in place (rows with mask[i] != 0 move to the front in order;
 * rows >= n count as unselected).  Every row is read and written at every
 * level of the routing network regardless of the mask.  *out_count (device
 * pointer, may be NULL) receives the number of selected rows. */
typedef struct vo_compact_args {
    const uint8_t *mask;
    void *cols[VO_MAX_SORT_COLUMNS];
    uint32_t col_bytes[VO_MAX_SORT_COLUMNS]; /* 1, 2, 4 or 8 */
    uint32_t n_cols;
    uint64_t n;
    uint64_t n_pad;
    uint64_t *out_count;
} vo_compact_args;

/* OP_SHUFFLE: oblivious shuffle of n_pad rows: tag_i = Philox4x32-10(i, seed),
 * sort by (tag, i) and apply the permutation to the payload column.
 * perm_out (optional) receives the source row of each output row.  Without
 * perm_out the payload is sorted alone as (tag >> 1, payload) records, 12
 * bytes per row instead of 16: the same permutation unless two live tags
 * share their top 63 bits.  With 0 < n < n_pad, rows >= n keep their place
 * and the first n rows receive a permutation of themselves. */
typedef struct vo_shuffle_args {
    uint32_t *payload; /* n_pad rows, shuffled in place */
    uint32_t *perm_out;
    uint64_t n_pad;
    uint64_t seed;
    uint64_t n;        /* live rows; 0 -> n_pad */
} vo_shuffle_args;

/* OP_FILTER_SUM: Q6-style predicate
 *   day_lo <= day < day_hi && disc_lo <= disc <= disc_hi && qty < qty_lt
 * over n rows padded to n_pad, oblivious compaction of (price, disc) into
 * out_price / out_disc (n_pad each), and exact int128 SUM(price * disc) over
 * the selected rows.  With out_price and out_disc both NULL the compaction
 * runs in place: price and disc (n_pad each) are overwritten with the
 * compacted rows.
 * out (device pointer, 32 bytes): { u64 count, u64 pad, u64 sum_lo, i64 sum_hi }. */
typedef struct vo_filter_args {
    const int32_t *day;
    const int64_t *qty;
    const int64_t *price;
    const int64_t *disc;
    int64_t *out_price;
    int64_t *out_disc;
    uint64_t n;
    uint64_t n_pad;
    int32_t day_lo, day_hi;
    int64_t disc_lo, disc_hi;
    int64_t qty_lt;
    uint64_t *out;
} vo_filter_args;

/* OP_PKFK_JOIN_GROUPSUM: dimension (dim_key unique, dim_group in
 * [0, n_groups)) joined with fact (fact_fk, fact_val) on key, then
 * SUM(fact_val) GROUP BY dim_group.  Union of n_dim + n_fact rows padded to
 * n_pad.  out_group / out_sum (device, max_groups each) receive the groups
 * ascending; *out_n_groups (device) the group count; sorted_keys /
 * sorted_vals (device, n_pad each, optional) the sorted union. */
typedef struct vo_join_args {
    const int32_t *dim_key;
    const int32_t *dim_group;
    uint64_t n_dim;
    const int32_t *fact_fk;
    const int64_t *fact_val;
    uint64_t n_fact;
    uint64_t n_pad;
    int32_t *out_group;
    int64_t *out_sum;
    uint64_t *out_n_groups;
    uint64_t max_groups;
    uint64_t *sorted_keys;
    uint64_t *sorted_vals;
    uint32_t dense_groups; /* 1: group ids lie in [0, max_groups) (a public bound, max_groups <= 64);
                              the GROUP BY is an oblivious dense aggregation fused with the join's
                              fill instead of sort + segmented scan + compaction.  Rows whose group
                              is outside the bound are not counted.  With max_groups <= 63 and no
                              sorted_keys / sorted_vals, the union is sorted as 12-byte records;
                              there fact values must be <= INT64_MAX - 65, and a larger one makes
                              *out_n_groups = UINT64_MAX (no result). */
} vo_join_args;

/* OP_SORT_MERGE_JOIN_COUNT / OP_SORT_MERGE_JOIN_WRITE: oblivious equi-join
 * of two key columns into row-index pairs; replaces the non-oblivious
 * SortMergeJoinCountArgs / SortMergeJoinWriteArgs pair
 * (include/cuda/commands.hpp:82-98, include/cuda/sort_merge_join.cuh:104-172).
 * The output is every (l, r) with left_keys[l] == right_keys[r], as row
 * indices (the rowids the reference's operator fills with 0..n-1,
 * src/operator/sort_merge_join_operator.cpp:75-92), ordered by (key, l, r).
 * Inputs need not be sorted.  Key types: BOOLEAN, TINYINT, SMALLINT,
 * INTEGER, BIGINT, DATE, VARCHAR (ordinal) and UBIGINT.
 *   COUNT: sorts the union of the n_left + n_right rows (padded to a power
 *     of two >= 4096, at most 2^29), derives every row's match counts and
 *     output positions, and compacts the matched rows; *out_count (device)
 *     receives the number of pairs n_out.  The routing state goes into
 *     `state`, a device buffer of vo_join_pairs_state_bytes(n_left, n_right)
 *     bytes the caller owns (as the reference passes MergeJoinBlock buffers
 *     from COUNT to WRITE, commands.hpp:82-98), or, with state NULL, stays in
 *     the context until the next COUNT.
 *   WRITE: expands a COUNT's state into out_left / out_right (device,
 *     out_cap rows each, out_cap >= n_out, n_out <= 2^29); rows
 *     [n_out, out_cap) are -1.  With `state`: the buffer the COUNT of the same
 *     key columns and sizes filled, and n_out = the pair count it produced
 *     (the caller reads *out_count, as the reference's operator reads its
 *     counts, sort_merge_join_operator.cpp:113-133), so joins with their own
 *     state buffers may interleave their COUNTs and WRITEs.  With state NULL:
 *     the context's last COUNT with the same key pointers and sizes (type_id
 *     0: the COUNT's).  out_cap 0 means the power of two >= n_out, the size
 *     the reference's operator allocates (sort_merge_join_operator.cpp:126-133).
 * Every pass reads and writes every row of its padded size; n_out is the
 * only data-dependent value that leaves the device (the reference returns
 * it too).  Uses about 46 bytes of device memory per padded union row and
 * 44 per padded output row (at least 8192 output rows). */
typedef struct vo_join_pairs_args {
    const void *left_keys;
    uint64_t n_left;
    const void *right_keys;
    uint64_t n_right;
    uint32_t type_id;    /* vo_type of both key columns */
    uint32_t reserved;   /* 0 */
    uint64_t *out_count; /* COUNT: device, receives n_out (may be NULL) */
    int64_t *out_left;   /* WRITE: device, out_cap rows */
    int64_t *out_right;  /* WRITE: device, out_cap rows */
    uint64_t out_cap;
    void *state;         /* device, vo_join_pairs_state_bytes() bytes, caller-owned; NULL: context state */
    uint64_t n_out;      /* WRITE with state: the COUNT's pair count */
} vo_join_pairs_args;

/* OP_PREDICATE_MASK: the device form of the host predicate evaluation
 * (ComparisonExpression::evaluateBatch, src/expression/comparison_expression.cpp:13-92,
 * and BinaryLogicalExpression::evaluateBatch, src/expression/logical_expression.cpp:61-150).
 * out_mask[i] = 1 when row i < n satisfies every term (VO_CONN_AND) or any
 * term (VO_CONN_OR) and no term's row is NULL, else 0; rows [n, n_pad) get 0.
 * A term compares its column with a constant in the column's own type:
 * `cmp` takes the reference's ComparisonType order (EQUAL, NOT_EQUAL,
 * LESS_THAN, LESS_THAN_OR_EQUAL, GREATER_THAN, GREATER_THAN_OR_EQUAL);
 * `constant` holds the constant's bit pattern in its low bytes (an integer
 * of the column's width, a FLOAT's 32 bits, a DOUBLE's 64 bits); `valid` is
 * an optional byte per row, 0 = NULL.  Every row of every term column is
 * read.  `terms` is host memory, read before vo_submit / vo_predicate_mask
 * returns. */
#define VO_MAX_PREDICATE_TERMS 8
typedef enum vo_cmp {
    VO_CMP_EQ = 0,
    VO_CMP_NE = 1,
    VO_CMP_LT = 2,
    VO_CMP_LE = 3,
    VO_CMP_GT = 4,
    VO_CMP_GE = 5
} vo_cmp;
typedef enum vo_connective { VO_CONN_AND = 0, VO_CONN_OR = 1 } vo_connective;
typedef struct vo_predicate_term {
    const void *col;      /* n rows of the column type */
    const uint8_t *valid; /* optional, n bytes, 0 = NULL */
    uint64_t constant;    /* bit pattern of the constant */
    uint32_t type_id;     /* vo_type */
    uint32_t cmp;         /* vo_cmp */
} vo_predicate_term;
typedef struct vo_predicate_args {
    const vo_predicate_term *terms; /* host array of n_terms (1..8) */
    uint32_t n_terms;
    uint32_t connective;            /* vo_connective */
    uint8_t *out_mask;              /* device, n_pad bytes */
    uint64_t n;
    uint64_t n_pad;
} vo_predicate_args;

/* OP_GATHER_OBLIVIOUS: out[i] = table[record_ids[i]] for i < n, the device
 * form of the host oblivious materialization
 * (src/operator/materialization_operator.cpp:362-517).  The table rows and
 * the requests are sorted together by record id, each value is filled forward
 * to its requests, the requests are compacted and sorted back to request
 * order: no address depends on a record id or on how often it repeats.  An
 * id outside [0, n_table) gives 0 and out_found[i] = 0 (out_found optional).
 * n_table + n padded to a power of two must be at most 2^29. */
typedef struct vo_gather_oblivious_args {
    const void *table; /* n_table rows of elem_bytes */
    uint64_t n_table;
    const int64_t *record_ids; /* n requests */
    uint64_t n;
    void *out;                 /* n rows of elem_bytes, request order */
    uint8_t *out_found;        /* optional, n bytes */
    uint32_t elem_bytes;       /* 1, 2, 4 or 8 */
    uint32_t reserved;         /* 0 */
} vo_gather_oblivious_args;

/* OP_SORT_BATCH: independent sorts in one schedule.  Item i sorts
 * (keys + offset, payload + offset) over n rows padded to n_pad (a power of
 * two; offset must be a multiple of n_pad).  Items are bucketed by n_pad. */
typedef struct vo_batch_item {
    uint64_t offset;
    uint64_t n;
    uint64_t n_pad;
} vo_batch_item;

typedef struct vo_sort_batch_args {
    uint64_t *keys;
    uint32_t *payload;
    const vo_batch_item *items; /* host array */
    uint64_t n_items;
} vo_sort_batch_args;

/* OP_COMPACT_BATCH: independent compactions (OP_COMPACT's network) in one
 * schedule.  Item i compacts rows [offset, offset + n_pad) of every column
 * by the same rows of `mask`: the rows r < n with mask[offset + r] != 0 move,
 * in order, to the front of the item's rows, as vo_compact on that item
 * alone would leave them, and out_counts[i] receives their count.  n_pad is
 * a power of two in [4096, 2^32] and offset a multiple of n_pad.  Items are
 * bucketed by n_pad: one launch per pass serves every item of a class.  The
 * virtual rows of one class (items x n_pad) are at most 2^32. */
typedef struct vo_compact_batch_args {
    const uint8_t *mask;                      /* device */
    void *cols[VO_MAX_SORT_COLUMNS];          /* device */
    uint32_t col_bytes[VO_MAX_SORT_COLUMNS];  /* 1, 2, 4 or 8 */
    uint32_t n_cols;
    const vo_batch_item *items;               /* host array */
    uint64_t n_items;
    uint64_t *out_counts;                     /* device, n_items (may be NULL) */
} vo_compact_batch_args;

typedef union vo_command_args {
    vo_sort_args sort;
    vo_sort_kp_args sort_kp;
    vo_permute_args permute;
    vo_scatter_args scatter;
    vo_gather_args gather;
    vo_compare_eq_args compare_eq;
    vo_compose_args compose_position_map;
    vo_enrich_args enrich_positions;
    vo_compact_args compact;
    vo_shuffle_args shuffle;
    vo_filter_args filter;
    vo_join_args join;
    vo_sort_batch_args sort_batch;
    vo_compact_batch_args compact_batch;
    vo_join_pairs_args join_pairs; /* OP_SORT_MERGE_JOIN_COUNT / _WRITE */
    vo_gather_oblivious_args gather_oblivious;
    vo_predicate_args predicate;   /* OP_PREDICATE_MASK */
} vo_command_args;

/* Mirror of Command (commands.hpp:152-156). */
typedef struct vo_command {
    uint64_t sequence_id; /* set by vo_submit */
    uint32_t opcode;      /* vo_opcode */
    uint32_t flags;       /* reserved, 0 */
    vo_command_args args;
} vo_command;

/* ------------------------------------------------------------- context */

enum {
    VO_CFG_PERSISTENT = 1u << 0,    /* run commands in the resident queue kernel */
    VO_CFG_KERNEL_TIMING = 1u << 1, /* HIP events around every launch (vo_kernel_stats) */
};

typedef struct vo_config {
    uint32_t queue_capacity; /* ring entries, power of two; 0 -> 1024 (task_manager.hpp:14) */
    uint32_t flags;          /* VO_CFG_* */
    void *stream;            /* hipStream_t to run on; NULL -> the context creates a blocking stream,
                                ordered with work on the null (default) stream */
    int32_t device;          /* HIP device ordinal; -1 -> current */
    uint32_t wait_timeout_ms;/* default timeout of vo_wait when 0 is passed; 0 -> 60000 */
    uint32_t tuning;         /* 0 -> autotuned table / built-in defaults.  Sort schedule word:
                                bits 0-3 stages per register pass (1..5), bits 4-7 stages per
                                LDS-staged pass + 1, 0x100 merge blocks that prefetch their
                                next tile, 0x200 the (u64, u32) sort with 1024-thread blocks,
                                0x400 its LDS-staged passes with the other block shape, 0x800
                                the other cache policy of the sort and compaction passes
                                (streaming loads and stores are chosen above 256 MB of rows).
                                Every word applies the same network: the output does not
                                change */
    uint32_t queue_idle_us;  /* the resident queue kernel leaves after this long without a
                                command (0 -> 20000); the next vo_submit starts it again */
    uint32_t queue_lanes;    /* lanes of the resident queue kernel, 1..8 (0 -> 1): the resident
                                blocks are split into lanes that run their own commands side by
                                side (vo_submit_lane) */
} vo_config;

typedef struct vo_ctx vo_ctx;

VO_EXPORT int vo_ctx_create(const vo_config *cfg, vo_ctx **out);
VO_EXPORT int vo_ctx_destroy(vo_ctx *ctx);
VO_EXPORT const char *vo_last_error(const vo_ctx *ctx);
VO_EXPORT const char *vo_status_string(int status);
VO_EXPORT int vo_version(void);

/* TaskManager::start / stop (src/cuda/task_manager.cu:26-124).  start()
 * launches the resident queue kernel when VO_CFG_PERSISTENT is set; without
 * it commands are launched directly on the context stream.
 * While the resident kernel runs it holds registers and LDS on every CU, so
 * kernels of other streams and libraries may wait until it leaves: it
 * leaves after queue_idle_us without commands, on vo_queue_stop, and when a
 * synchronous call of this context needs an operator the resident kernel
 * does not run (the kernel first finishes every submitted command). */
VO_EXPORT int vo_queue_start(vo_ctx *ctx);
VO_EXPORT int vo_queue_stop(vo_ctx *ctx, uint32_t timeout_ms);
VO_EXPORT int vo_queue_running(const vo_ctx *ctx);

/* TaskManager::submitCommand / waitCommand.  vo_submit returns the sequence
 * id (> 0) and stores it in cmd->sequence_id, or 0 on error (see
 * vo_last_error).  vo_wait returns VO_ERR_TIMEOUT after timeout_ms
 * (0 -> the context default).  In persistent mode an opcode the resident
 * kernel does not run is executed when vo_submit is called: the commands
 * submitted before it finish, it runs on the context stream, and its
 * sequence id is that of a NOP queued after it, so the order of commands
 * and the meaning of vo_wait are unchanged. */
VO_EXPORT uint64_t vo_submit(vo_ctx *ctx, vo_command *cmd);
/* vo_submit on one lane of the resident queue (0 <= lane < vo_queue_lanes).
 * Commands of one lane run in submission order; commands of different lanes
 * run side by side and may finish in any order (vo_wait / vo_poll take each
 * command's own sequence id).  Lanes of one context share its buffers, so
 * commands that touch the same memory go on the same lane.  No counterpart
 * in the reference, whose queue runs one command at a time
 * (persistent_kernel.cu:18-124); used by the serving loop. */
VO_EXPORT uint64_t vo_submit_lane(vo_ctx *ctx, vo_command *cmd, uint32_t lane);
VO_EXPORT int vo_queue_lanes(const vo_ctx *ctx); /* 0 when the resident queue is not running */
VO_EXPORT int vo_wait(vo_ctx *ctx, uint64_t seq, uint32_t timeout_ms);
VO_EXPORT int vo_poll(vo_ctx *ctx, uint64_t seq); /* 1 done, 0 pending, <0 -error */
VO_EXPORT int vo_sync(vo_ctx *ctx);               /* wait for everything submitted */

/* Synchronous conveniences: submit + wait. */
VO_EXPORT int vo_sort_kp(vo_ctx *ctx, uint64_t *keys, uint32_t *payload, uint64_t n, uint64_t n_pad);
VO_EXPORT int vo_sort(vo_ctx *ctx, const vo_sort_args *args);
VO_EXPORT int vo_permute(vo_ctx *ctx, const vo_permute_args *args);
VO_EXPORT int vo_compact(vo_ctx *ctx, const vo_compact_args *args);
VO_EXPORT int vo_shuffle(vo_ctx *ctx, const vo_shuffle_args *args);
VO_EXPORT int vo_filter_sum(vo_ctx *ctx, const vo_filter_args *args);
VO_EXPORT int vo_pkfk_join_groupsum(vo_ctx *ctx, const vo_join_args *args);
VO_EXPORT int vo_sort_batch(vo_ctx *ctx, const vo_sort_batch_args *args);
/* OP_COMPACT_BATCH, synchronous.  No reference counterpart: the batched
 * form of ORCompact (obliviator/join/enclave/oblivious_compact.c:59-338)
 * for thousands of compactions of uneven public sizes. */
VO_EXPORT int vo_compact_batch(vo_ctx *ctx, const vo_compact_batch_args *args);
/* OP_SORT_MERGE_JOIN_COUNT then OP_SORT_MERGE_JOIN_WRITE; *n_out (host, may
 * be NULL) receives the pair count after COUNT. */
VO_EXPORT int vo_join_pairs_count(vo_ctx *ctx, const vo_join_pairs_args *args, uint64_t *n_out);
VO_EXPORT int vo_join_pairs_write(vo_ctx *ctx, const vo_join_pairs_args *args);
/* Bytes of the caller-owned COUNT -> WRITE state for n_left + n_right rows. */
VO_EXPORT uint64_t vo_join_pairs_state_bytes(uint64_t n_left, uint64_t n_right);
VO_EXPORT int vo_gather_oblivious(vo_ctx *ctx, const vo_gather_oblivious_args *args);
/* OP_PREDICATE_MASK (replaces the host ComparisonExpression /
 * BinaryLogicalExpression::evaluateBatch, comparison_expression.cpp:13-92,
 * logical_expression.cpp:61-150, for column-vs-constant terms). */
VO_EXPORT int vo_predicate_mask(vo_ctx *ctx, const vo_predicate_args *args);

/* Host <-> device wire format (StagedTransfer, include/cuda/staged_transfer.hpp:101-225,
 * and the padded moves of ValueVector::to, include/data/value_vector.hpp:140-294).
 * Copies `bytes` between pageable host memory and a device buffer of at
 * least padded_bytes, and always moves padded_bytes over the link: to the
 * device the pad region receives zeros, from the device the pad rows are read
 * and discarded, so the link shows the padded size only (callers pad to the
 * power-of-two capacity of the column).  Staged through a ring of pinned
 * chunks owned by the context, with the host copies on worker threads
 * overlapping the DMA.  Ordered after the work enqueued on the context
 * stream and complete on return.  *link_bytes (may be NULL) receives the
 * bytes moved over the link. */
VO_EXPORT int vo_transfer_to_device(vo_ctx *ctx, void *dst_dev, const void *src_host, uint64_t bytes,
                                    uint64_t padded_bytes, uint64_t *link_bytes);
VO_EXPORT int vo_transfer_to_host(vo_ctx *ctx, void *dst_host, const void *src_dev, uint64_t bytes,
                                  uint64_t padded_bytes, uint64_t *link_bytes);

/* Asynchronous forms: enqueue on the context stream and return. */
VO_EXPORT int vo_sort_kp_async(vo_ctx *ctx, uint64_t *keys, uint32_t *payload, uint64_t n, uint64_t n_pad);

/* Allocation-free compaction.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as vo_compact, but asynchronous (enqueued on the context stream;
 * synchronise the stream or call vo_sync before reading the results), and
 * its tile arrays live in a fixed 2 MiB device array of the library instead
 * of the context's scratch arena, so the call allocates nothing.
 * Limits: n_pad <= 2^29 (VO_ERR_INVALID above); tracing and kernel timing
 * must be off (they allocate).  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other contexts,
 * other threads) race and must be serialised by the caller. */
VO_EXPORT int vo_compact_noalloc(vo_ctx *ctx, const vo_compact_args *args);


/* --------------------------------------------------------- diagnostics */

/* Obliviousness trace (modelled on TraceMem,
 * obliviator/join_kks_INT_INT/IncludeL3/trace_mem.h:23-116).  Between begin
 * and end every launch is recorded (kernel id, grid, block, LDS bytes and its
 * non-pointer arguments) and runs an instrumented build of the kernel that
 * folds every global and LDS address it touches into a per-thread hash. */
typedef struct vo_trace {
    uint64_t n_launches;
    uint64_t launch_digest;  /* hash of the launch sequence and shapes */
    uint64_t address_digest; /* hash of all per-launch address digests */
    uint64_t n_addresses;    /* number of addresses folded */
} vo_trace;

VO_EXPORT int vo_trace_begin(vo_ctx *ctx);
VO_EXPORT int vo_trace_end(vo_ctx *ctx, vo_trace *out);
/* Per-launch record i of the last trace: kernel id string, grid/block dims,
 * that launch's address digest. */
VO_EXPORT int vo_trace_launch(vo_ctx *ctx, uint64_t i, char *name, uint32_t name_cap,
                    uint32_t grid[3], uint32_t block[3], uint64_t *addr_digest);

/* Kernel timing (VO_CFG_KERNEL_TIMING): launches and summed milliseconds
 * per kernel family since the last reset. */
VO_EXPORT int vo_kernel_stats(vo_ctx *ctx, const char *family, uint64_t *launches, double *total_ms);
VO_EXPORT int vo_kernel_stats_reset(vo_ctx *ctx);

/* Sort schedule for n_pad rows: number of HBM passes (launches) and the
 * per-pass record of (kernel family, level, stages), for planning and docs. */
VO_EXPORT int vo_sort_plan(vo_ctx *ctx, uint64_t n_pad, uint32_t *n_passes);

/* Autotuner (no counterpart in the reference, which hard-codes its launch
 * shape, src/cuda/task_manager.cu:32-55).  For each size class
 * 2^log_min .. 2^log_max it times the schedule variants (stages per
 * register pass, stages per LDS-staged pass, resident merge blocks) on a
 * synthetic batch of about rows_per_class rows and keeps the fastest for
 * that class in the context; vo_sort_batch then uses it.  The choice
 * depends only on public sizes.  chosen (optional, log_max-log_min+1
 * entries) receives the tuning words (vo_config.tuning format). */
VO_EXPORT int vo_autotune_sort(vo_ctx *ctx, uint32_t log_min, uint32_t log_max, uint64_t rows_per_class,
                               uint32_t *chosen);
VO_EXPORT int vo_set_class_tuning(vo_ctx *ctx, uint32_t log_n, uint32_t tuning_word);
VO_EXPORT int vo_get_class_tuning(vo_ctx *ctx, uint32_t log_n, uint32_t *tuning_word);

/* Plain device-to-device copy of `bytes` (a multiple of 16) with a
 * streaming kernel: the reference point for HBM traffic counters and the
 * practical copy bandwidth.  No counterpart in the reference. */
VO_EXPORT int vo_hbm_copy(vo_ctx *ctx, void *dst, const void *src, uint64_t bytes);

/* ----------------------------------------------------------- serving loop */

/* A continuous-batching query server (no counterpart in the reference,
 * whose ExecutionEngine runs one query at a time,
 * src/execution/execution_engine.cpp:67-142).  Tables live on the device
 * for the server's lifetime.  A run replays a list of queries with arrival
 * times: a query is admitted when it has arrived, a slot is free and its
 * padded HBM footprint fits in the remaining budget; admitted queries run
 * concurrently, one HIP stream per slot.  The schedule depends only on
 * public sizes and arrival times. */
enum { VO_Q_FILTER_SUM = 1, VO_Q_JOIN_GROUPSUM = 2, VO_Q_SORT = 3 };

typedef struct vo_server_config {
    uint32_t n_slots;          /* concurrent queries (1 = the plain loop) */
    uint32_t device;
    uint64_t hbm_budget_bytes; /* admission budget for query intermediates */
    uint32_t queue_lanes;      /* VO_RUN_QUEUE: lanes of the resident queue, 1..8 (0: min(slots, 8));
                                  slot i runs on lane i mod lanes */
    uint32_t reserved;         /* 0 */
} vo_server_config;

typedef struct vo_query {
    uint32_t kind;  /* VO_Q_* */
    uint32_t table; /* table id of that kind */
    double arrival_s;
    void *out;      /* optional host memory of vo_server_query_out_bytes() bytes: the query's whole
                       result, copied out of the device after the query ends (outside its latency):
                         filter: { u64 count, u64 0, u64 sum_lo, i64 sum_hi } then price[n], disc[n]
                                 (the compacted columns: the selected rows first, in row order);
                         join:   { i32 group[64], i64 sum[64], u64 n_groups } (groups ascending);
                         sort:   keys[n] (u64) then payload[n] (u32), sorted */
} vo_query;

typedef struct vo_query_result {
    double admit_s, end_s, latency_s; /* relative to the run start; latency = end - arrival */
    uint64_t footprint_bytes;
    uint64_t r[4]; /* filter: count, sum_lo, sum_hi; join: groups, sum-weighted checksum, total;
                      sort: rows, key checksum, payload checksum */
    int32_t status;
    uint32_t slot;
} vo_query_result;

typedef struct vo_run_stats {
    double makespan_s, qps, p50_s, p99_s, mean_s;
    uint64_t peak_bytes_in_flight;
    uint32_t max_concurrent;
    uint32_t graph_replays; /* resident queries run as one replay of a captured HIP graph */
} vo_run_stats;

typedef struct vo_server vo_server;

VO_EXPORT int vo_server_create(const vo_server_config *cfg, vo_server **out);
VO_EXPORT int vo_server_destroy(vo_server *s);
VO_EXPORT const char *vo_server_last_error(const vo_server *s);
/* Device-resident tables; each returns its id through *id. */
VO_EXPORT int vo_server_add_fact_table(vo_server *s, uint64_t seed, uint64_t n_rows, uint32_t *id);
VO_EXPORT int vo_server_add_join_tables(vo_server *s, uint64_t seed, uint64_t n_dim, uint64_t n_fact, uint32_t *id);
VO_EXPORT int vo_server_add_sort_table(vo_server *s, uint64_t seed, uint64_t n_rows, uint32_t *id);
/* Replays n queries; n_slots_override (0 = the server's) picks the
 * concurrency, 1 gives the one-query-at-a-time loop.  Tables are read in
 * place in HBM (VO_RUN_RESIDENT). */
VO_EXPORT int vo_server_run(vo_server *s, const vo_query *q, uint64_t n, uint32_t n_slots_override,
                            vo_query_result *out, vo_run_stats *stats);

/* How a run reaches the tables. */
enum {
    VO_RUN_RESIDENT = 0, /* queries read the device-resident tables in place */
    VO_RUN_TRANSFER = 1, /* the reference engine's per-query path: each query copies its input
                            columns from pinned host memory into device buffers
                            (ValueVector::to(CUDA), include/data/value_vector.hpp:248-272), and a
                            sort copies its sorted columns back to host memory */
    VO_RUN_QUEUE = 2     /* over the resident tables, every admitted query's operators go into the
                            resident command queue as one program, on its slot's lane of a
                            persistent context (min(n_slots, 8) lanes): the resident kernel runs
                            the lanes' queries side by side */
};
VO_EXPORT int vo_server_run_mode(vo_server *s, const vo_query *q, uint64_t n, uint32_t n_slots_override,
                                 uint32_t mode, vo_query_result *out, vo_run_stats *stats);
/* Bytes of a query's whole result (vo_query.out), 0 for an unknown table. */
VO_EXPORT uint64_t vo_server_query_out_bytes(vo_server *s, const vo_query *q);
/* The host copy of a table column (the data the device copy was generated
 * with), for checking results: fact tables day (i32), qty, price, disc (i64);
 * join tables dim_key, dim_group, fact_fk (i32), fact_val (i64); sort tables
 * keys (u64), payload (u32).  *rows receives the column's row count. */
VO_EXPORT int vo_server_table_column(vo_server *s, uint32_t kind, uint32_t table, uint32_t column,
                                     const void **host, uint64_t *rows);

/* ------------------------------------------------------ data generators */

/* Device generators of the synthetic workloads (same streams as the
 * host generator; key[i] = SplitMix64 stream 0 of seed, payload[i] = i). */
VO_EXPORT int vo_gen_sort_input(vo_ctx *ctx, uint64_t seed, uint64_t *keys, uint32_t *payload, uint64_t n);
VO_EXPORT int vo_gen_fact(vo_ctx *ctx, uint64_t seed, int32_t *day, int64_t *qty, int64_t *price, int64_t *disc, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
