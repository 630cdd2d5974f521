"""ctypes mirror of include/vo_api.h (the library's C ABI).

Every struct here matches the C layout field for field; tests and the bench
call the library only through these prototypes, the way a reference-side
binding would.
"""
import ctypes as C

VO_OK = 0
VO_ERR_INVALID = 1
VO_ERR_HIP = 2
VO_ERR_TIMEOUT = 3
VO_ERR_NOMEM = 4
VO_ERR_UNSUPPORTED = 5
VO_ERR_QUEUE = 6
VO_ERR_ADMISSION = 7

# DataTypeId values (reference include/data/data_type.hpp:12-27) + UBIGINT
T_BOOLEAN, T_TINYINT, T_SMALLINT, T_INTEGER, T_BIGINT = 1, 2, 3, 4, 5
T_FLOAT, T_DOUBLE, T_VARCHAR, T_DATE, T_UBIGINT = 6, 7, 8, 10, 64

OP_NOP = 0
OP_SCATTER = 1
OP_GATHER = 2
OP_SORT = 3
OP_PERMUTE = 4
OP_SORT_MERGE_JOIN_COUNT = 5
OP_SORT_MERGE_JOIN_WRITE = 6
OP_HASH_JOIN_BUILD = 7
OP_HASH_JOIN_COUNT = 8
OP_HASH_JOIN_WRITE = 9
OP_COMPARE_EQ = 10
OP_COMPOSE_POSITION_MAP = 11
OP_ENRICH_POSITIONS = 12
OP_SORT_KP = 0x100
OP_COMPACT = 0x101
OP_SHUFFLE = 0x102
OP_FILTER_SUM = 0x103
OP_PKFK_JOIN_GROUPSUM = 0x104
OP_SORT_BATCH = 0x105
OP_GATHER_OBLIVIOUS = 0x106
OP_PREDICATE_MASK = 0x107
OP_COMPACT_BATCH = 0x108
OP_TERMINATE = 0xFFFFFFFF

CFG_PERSISTENT = 1 << 0
CFG_KERNEL_TIMING = 1 << 1

MAX_SORT_COLUMNS = 8

P = C.c_void_p
u8, u32, u64, i32, i64 = C.c_uint8, C.c_uint32, C.c_uint64, C.c_int32, C.c_int64


class SortArgs(C.Structure):
    _fields_ = [("sort_cols", P * 8), ("indices", P), ("n_sort_columns", u64), ("n_rows", u64),
                ("n_padded_rows", u64), ("col_types", u32 * 8), ("ascending_flags", u8 * 8)]


class SortKpArgs(C.Structure):
    _fields_ = [("keys", P), ("payload", P), ("n", u64), ("n_pad", u64)]


class PermuteArgs(C.Structure):
    _fields_ = [("out_data", P), ("in_data", P), ("in_indices", P), ("n", u64), ("type_id", u32)]


class ScatterArgs(C.Structure):
    _fields_ = [("out_indices", P), ("out_count", P), ("in_mask", P), ("n", u64)]


class GatherArgs(C.Structure):
    _fields_ = [("out_data", P), ("in_data", P), ("in_indices", P), ("in_mask", P), ("n", u64), ("type_id", u32)]


class CompareEqArgs(C.Structure):
    _fields_ = [("left_data", P), ("right_data", P), ("out_mask", P), ("n", u64), ("type_id", u32)]


class ComposeArgs(C.Structure):
    _fields_ = [("old_map", P), ("sigma", P), ("new_map", P), ("n", u64)]


class EnrichArgs(C.Structure):
    _fields_ = [("record_ids", P), ("position_map", P), ("positions", P), ("n", u64)]


class CompactArgs(C.Structure):
    _fields_ = [("mask", P), ("cols", P * 8), ("col_bytes", u32 * 8), ("n_cols", u32), ("n", u64),
                ("n_pad", u64), ("out_count", P)]


class ShuffleArgs(C.Structure):
    _fields_ = [("payload", P), ("perm_out", P), ("n_pad", u64), ("seed", u64), ("n", u64)]


class FilterArgs(C.Structure):
    _fields_ = [("day", P), ("qty", P), ("price", P), ("disc", P), ("out_price", P), ("out_disc", P),
                ("n", u64), ("n_pad", u64), ("day_lo", i32), ("day_hi", i32), ("disc_lo", i64),
                ("disc_hi", i64), ("qty_lt", i64), ("out", P)]


class JoinArgs(C.Structure):
    _fields_ = [("dim_key", P), ("dim_group", P), ("n_dim", u64), ("fact_fk", P), ("fact_val", P),
                ("n_fact", u64), ("n_pad", u64), ("out_group", P), ("out_sum", P), ("out_n_groups", P),
                ("max_groups", u64), ("sorted_keys", P), ("sorted_vals", P), ("dense_groups", u32)]


class BatchItem(C.Structure):
    _fields_ = [("offset", u64), ("n", u64), ("n_pad", u64)]


class SortBatchArgs(C.Structure):
    _fields_ = [("keys", P), ("payload", P), ("items", P), ("n_items", u64)]


class CompactBatchArgs(C.Structure):
    _fields_ = [("mask", P), ("cols", P * 8), ("col_bytes", u32 * 8), ("n_cols", u32), ("items", P),
                ("n_items", u64), ("out_counts", P)]


class JoinPairsArgs(C.Structure):
    _fields_ = [("left_keys", P), ("n_left", u64), ("right_keys", P), ("n_right", u64), ("type_id", u32),
                ("reserved", u32), ("out_count", P), ("out_left", P), ("out_right", P), ("out_cap", u64),
                ("state", P), ("n_out", u64)]


class GatherObliviousArgs(C.Structure):
    _fields_ = [("table", P), ("n_table", u64), ("record_ids", P), ("n", u64), ("out", P), ("out_found", P),
                ("elem_bytes", u32), ("reserved", u32)]


# ComparisonType order (reference include/expression/expression.hpp:34-41)
CMP_EQ, CMP_NE, CMP_LT, CMP_LE, CMP_GT, CMP_GE = 0, 1, 2, 3, 4, 5
CONN_AND, CONN_OR = 0, 1
MAX_PREDICATE_TERMS = 8


class PredicateTerm(C.Structure):
    _fields_ = [("col", P), ("valid", P), ("constant", u64), ("type_id", u32), ("cmp", u32)]


class PredicateArgs(C.Structure):
    _fields_ = [("terms", C.POINTER(PredicateTerm)), ("n_terms", u32), ("connective", u32), ("out_mask", P),
                ("n", u64), ("n_pad", u64)]


def constant_bits(type_id, value):
    """Bit pattern of a constant in a column type (the `constant` field)."""
    import struct
    if type_id == T_FLOAT:
        return struct.unpack("<I", struct.pack("<f", value))[0]
    if type_id == T_DOUBLE:
        return struct.unpack("<Q", struct.pack("<d", value))[0]
    width = {T_BOOLEAN: 1, T_TINYINT: 1, T_SMALLINT: 2, T_INTEGER: 4, T_DATE: 4}.get(type_id, 8)
    return int(value) & ((1 << (8 * width)) - 1)


class CommandArgs(C.Union):
    _fields_ = [("sort", SortArgs), ("sort_kp", SortKpArgs), ("permute", PermuteArgs), ("scatter", ScatterArgs),
                ("gather", GatherArgs), ("compare_eq", CompareEqArgs), ("compose_position_map", ComposeArgs),
                ("enrich_positions", EnrichArgs), ("compact", CompactArgs), ("shuffle", ShuffleArgs),
                ("filter", FilterArgs), ("join", JoinArgs), ("sort_batch", SortBatchArgs),
                ("join_pairs", JoinPairsArgs), ("gather_oblivious", GatherObliviousArgs),
                ("predicate", PredicateArgs), ("compact_batch", CompactBatchArgs)]


class Command(C.Structure):
    _fields_ = [("sequence_id", u64), ("opcode", u32), ("flags", u32), ("args", CommandArgs)]


class Config(C.Structure):
    _fields_ = [("queue_capacity", u32), ("flags", u32), ("stream", P), ("device", i32),
                ("wait_timeout_ms", u32), ("tuning", u32), ("queue_idle_us", u32),
                ("queue_lanes", u32)]


Q_FILTER_SUM, Q_JOIN_GROUPSUM, Q_SORT = 1, 2, 3
RUN_RESIDENT, RUN_TRANSFER, RUN_QUEUE = 0, 1, 2


class ServerConfig(C.Structure):
    _fields_ = [("n_slots", u32), ("device", u32), ("hbm_budget_bytes", u64), ("queue_lanes", u32),
                ("reserved", u32)]


class Query(C.Structure):
    _fields_ = [("kind", u32), ("table", u32), ("arrival_s", C.c_double), ("out", P)]


class QueryResult(C.Structure):
    _fields_ = [("admit_s", C.c_double), ("end_s", C.c_double), ("latency_s", C.c_double),
                ("footprint_bytes", u64), ("r", u64 * 4), ("status", i32), ("slot", u32)]


class RunStats(C.Structure):
    _fields_ = [("makespan_s", C.c_double), ("qps", C.c_double), ("p50_s", C.c_double), ("p99_s", C.c_double),
                ("mean_s", C.c_double), ("peak_bytes_in_flight", u64), ("max_concurrent", u32),
                ("graph_replays", u32)]


class Trace(C.Structure):
    _fields_ = [("n_launches", u64), ("launch_digest", u64), ("address_digest", u64), ("n_addresses", u64)]


# (name, restype, argtypes) for every function declared in include/vo_api.h
PROTOTYPES = [
    ("vo_ctx_create", C.c_int, [C.POINTER(Config), C.POINTER(P)]),
    ("vo_ctx_destroy", C.c_int, [P]),
    ("vo_last_error", C.c_char_p, [P]),
    ("vo_status_string", C.c_char_p, [C.c_int]),
    ("vo_version", C.c_int, []),
    ("vo_queue_start", C.c_int, [P]),
    ("vo_queue_stop", C.c_int, [P, u32]),
    ("vo_queue_running", C.c_int, [P]),
    ("vo_submit", u64, [P, C.POINTER(Command)]),
    ("vo_submit_lane", u64, [P, C.POINTER(Command), u32]),
    ("vo_queue_lanes", C.c_int, [P]),
    ("vo_wait", C.c_int, [P, u64, u32]),
    ("vo_poll", C.c_int, [P, u64]),
    ("vo_sync", C.c_int, [P]),
    ("vo_sort_kp", C.c_int, [P, P, P, u64, u64]),
    ("vo_sort", C.c_int, [P, C.POINTER(SortArgs)]),
    ("vo_permute", C.c_int, [P, C.POINTER(PermuteArgs)]),
    ("vo_compact", C.c_int, [P, C.POINTER(CompactArgs)]),
    ("vo_compact_noalloc", C.c_int, [P, C.POINTER(CompactArgs)]),
    ("vo_shuffle", C.c_int, [P, C.POINTER(ShuffleArgs)]),
    ("vo_filter_sum", C.c_int, [P, C.POINTER(FilterArgs)]),
    ("vo_pkfk_join_groupsum", C.c_int, [P, C.POINTER(JoinArgs)]),
    ("vo_sort_batch", C.c_int, [P, C.POINTER(SortBatchArgs)]),
    ("vo_compact_batch", C.c_int, [P, C.POINTER(CompactBatchArgs)]),
    ("vo_join_pairs_count", C.c_int, [P, C.POINTER(JoinPairsArgs), C.POINTER(u64)]),
    ("vo_transfer_to_device", C.c_int, [P, P, P, u64, u64, C.POINTER(u64)]),
    ("vo_gather_oblivious", C.c_int, [P, C.POINTER(GatherObliviousArgs)]),
    ("vo_predicate_mask", C.c_int, [P, C.POINTER(PredicateArgs)]),
    ("vo_transfer_to_host", C.c_int, [P, P, P, u64, u64, C.POINTER(u64)]),
    ("vo_join_pairs_write", C.c_int, [P, C.POINTER(JoinPairsArgs)]),
    ("vo_join_pairs_state_bytes", u64, [u64, u64]),
    ("vo_sort_kp_async", C.c_int, [P, P, P, u64, u64]),
    ("vo_trace_begin", C.c_int, [P]),
    ("vo_trace_end", C.c_int, [P, C.POINTER(Trace)]),
    ("vo_trace_launch", C.c_int, [P, u64, C.c_char_p, u32, C.POINTER(u32), C.POINTER(u32), C.POINTER(u64)]),
    ("vo_kernel_stats", C.c_int, [P, C.c_char_p, C.POINTER(u64), C.POINTER(C.c_double)]),
    ("vo_kernel_stats_reset", C.c_int, [P]),
    ("vo_sort_plan", C.c_int, [P, u64, C.POINTER(u32)]),
    ("vo_autotune_sort", C.c_int, [P, u32, u32, u64, C.POINTER(u32)]),
    ("vo_set_class_tuning", C.c_int, [P, u32, u32]),
    ("vo_get_class_tuning", C.c_int, [P, u32, C.POINTER(u32)]),
    ("vo_hbm_copy", C.c_int, [P, P, P, u64]),
    ("vo_server_create", C.c_int, [C.POINTER(ServerConfig), C.POINTER(P)]),
    ("vo_server_destroy", C.c_int, [P]),
    ("vo_server_last_error", C.c_char_p, [P]),
    ("vo_server_add_fact_table", C.c_int, [P, u64, u64, C.POINTER(u32)]),
    ("vo_server_add_join_tables", C.c_int, [P, u64, u64, u64, C.POINTER(u32)]),
    ("vo_server_add_sort_table", C.c_int, [P, u64, u64, C.POINTER(u32)]),
    ("vo_server_run", C.c_int, [P, C.POINTER(Query), u64, u32, C.POINTER(QueryResult), C.POINTER(RunStats)]),
    ("vo_server_run_mode", C.c_int, [P, C.POINTER(Query), u64, u32, u32, C.POINTER(QueryResult),
                                     C.POINTER(RunStats)]),
    ("vo_server_query_out_bytes", u64, [P, C.POINTER(Query)]),
    ("vo_server_table_column", C.c_int, [P, u32, u32, u32, C.POINTER(P), C.POINTER(u64)]),
    ("vo_gen_sort_input", C.c_int, [P, u64, P, P, u64]),
    ("vo_gen_fact", C.c_int, [P, u64, P, P, P, P, u64]),
]


def bind(lib):
    for name, res, args in PROTOTYPES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
