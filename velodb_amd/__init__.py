"""velodb_amd -- MI355X (gfx950) oblivious-operator library, Python host side.

The compute path is libvelodb_amd.so (hand-written HIP kernels behind the C
ABI of include/vo_api.h).  This module only loads it through ctypes and wraps
the calls; there is no CPU fallback: if the library is missing, loading fails
with an error.
"""
import ctypes as C
import os

import numpy as np

from . import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvelodb_amd.so")

_lib = None


class VoError(RuntimeError):
    """A non-zero vo_status from the library, with vo_last_error's text."""

    def __init__(self, status, msg):
        super().__init__(f"velodb_amd status {status}: {msg}")
        self.status = status


def load_library(path=None):
    """Load and bind libvelodb_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"{p} is missing: build it with __graft_entry__.build() "
                          f"(make -C velodb_amd/csrc)")
    # One HIP runtime per process: when torch (which ships its own HIP and HSA
    # runtimes) is installed, load it first so that the library binds to the
    # runtime torch uses.  Loaded the other way round, the process would hold
    # two runtimes and torch could no longer open the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = _abi.bind(C.CDLL(p))
    if path is None:
        _lib = lib
    return lib


def _ptr(x):
    """Device pointer of a torch tensor (or an int address / None)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


class Server:
    """Continuous-batching query server over device-resident tables
    (vo_server_* in include/vo_api.h)."""

    FILTER_SUM, JOIN_GROUPSUM, SORT = _abi.Q_FILTER_SUM, _abi.Q_JOIN_GROUPSUM, _abi.Q_SORT

    def __init__(self, n_slots=16, hbm_budget_bytes=200 << 30, device=0, queue_lanes=0):
        self.lib = load_library()
        cfg = _abi.ServerConfig(n_slots, device, hbm_budget_bytes, queue_lanes, 0)
        h = C.c_void_p()
        st = self.lib.vo_server_create(C.byref(cfg), C.byref(h))
        if st != 0:
            raise VoError(st, "vo_server_create failed")
        self.h = h

    def _check(self, st):
        if st != 0:
            raise VoError(st, self.lib.vo_server_last_error(self.h).decode())

    def add_fact_table(self, seed, n_rows):
        i = C.c_uint32()
        self._check(self.lib.vo_server_add_fact_table(self.h, seed, n_rows, C.byref(i)))
        return i.value

    def add_join_tables(self, seed, n_dim, n_fact):
        i = C.c_uint32()
        self._check(self.lib.vo_server_add_join_tables(self.h, seed, n_dim, n_fact, C.byref(i)))
        return i.value

    def add_sort_table(self, seed, n_rows):
        i = C.c_uint32()
        self._check(self.lib.vo_server_add_sort_table(self.h, seed, n_rows, C.byref(i)))
        return i.value

    def run(self, queries, n_slots=0, transfer=False, outputs=None, queue=False):
        """queries: list of (kind, table, arrival_s).  Returns (results, stats).
        transfer=True ships each query's inputs from pinned host memory (the
        reference engine's per-query path) instead of reading the resident
        tables.  queue=True runs every query as one program in the resident
        command queue, a lane per slot.  outputs: optional list (one entry per
        query) of numpy uint8 arrays of query_out_bytes() bytes, or None, that
        receive each query's whole result (vo_query.out)."""
        qs = []
        for i, (k, t, a) in enumerate(queries):
            o = outputs[i] if outputs is not None else None
            qs.append(_abi.Query(k, t, a, o.ctypes.data if o is not None else None))
        arr = (_abi.Query * len(queries))(*qs)
        res = (_abi.QueryResult * len(queries))()
        st = _abi.RunStats()
        mode = _abi.RUN_QUEUE if queue else _abi.RUN_TRANSFER if transfer else _abi.RUN_RESIDENT
        self._check(self.lib.vo_server_run_mode(self.h, arr, len(queries), n_slots, mode, res, C.byref(st)))
        return list(res), st

    def query_out_bytes(self, kind, table):
        q = _abi.Query(kind, table, 0.0, None)
        return int(self.lib.vo_server_query_out_bytes(self.h, C.byref(q)))

    def table_column(self, kind, table, column):
        """A copy of a table column's host data (numpy): fact day/qty/price/disc,
        join dim_key/dim_group/fact_fk/fact_val, sort keys/payload."""
        import numpy as np
        dts = {_abi.Q_FILTER_SUM: (np.int32, np.int64, np.int64, np.int64),
               _abi.Q_JOIN_GROUPSUM: (np.int32, np.int32, np.int32, np.int64),
               _abi.Q_SORT: (np.uint64, np.uint32)}
        ptr, rows = C.c_void_p(), C.c_uint64()
        self._check(self.lib.vo_server_table_column(self.h, kind, table, column, C.byref(ptr), C.byref(rows)))
        dt = np.dtype(dts[kind][column])
        buf = (C.c_char * (rows.value * dt.itemsize)).from_address(ptr.value)
        return np.frombuffer(buf, dtype=dt).copy()

    def close(self):
        if self.h:
            self.lib.vo_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Context:
    """One vo_ctx: a HIP stream plus submission bookkeeping.

    stream: a hipStream_t address (e.g. torch.cuda.current_stream().cuda_stream)
    so that library work is ordered with torch's copies; None lets the library
    create its own stream.
    """

    def __init__(self, stream=None, flags=0, device=-1, queue_capacity=0, wait_timeout_ms=0, tuning=0,
                 queue_idle_us=0, queue_lanes=0):
        self.lib = load_library()
        cfg = _abi.Config(queue_capacity, flags, stream, device, wait_timeout_ms, tuning, queue_idle_us,
                          queue_lanes)
        h = C.c_void_p()
        st = self.lib.vo_ctx_create(C.byref(cfg), C.byref(h))
        if st != 0:
            raise VoError(st, self.lib.vo_status_string(st).decode())
        self.h = h

    def close(self):
        if self.h:
            self.lib.vo_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st):
        if st != 0:
            raise VoError(st, self.lib.vo_last_error(self.h).decode())
        return st

    # ---- ops -------------------------------------------------------------
    def sort_kp(self, keys, payload, n, n_pad):
        return self.check(self.lib.vo_sort_kp(self.h, _ptr(keys), _ptr(payload), n, n_pad))

    def sort_kp_async(self, keys, payload, n, n_pad):
        return self.check(self.lib.vo_sort_kp_async(self.h, _ptr(keys), _ptr(payload), n, n_pad))

    def sort(self, cols, types, ascending, indices, n_rows, n_padded_rows):
        a = _abi.SortArgs()
        for i, c in enumerate(cols):
            a.sort_cols[i] = _ptr(c)
            a.col_types[i] = types[i]
            a.ascending_flags[i] = 1 if ascending[i] else 0
        a.indices = _ptr(indices)
        a.n_sort_columns = len(cols)
        a.n_rows = n_rows
        a.n_padded_rows = n_padded_rows
        return self.check(self.lib.vo_sort(self.h, C.byref(a)))

    def permute(self, out, inp, idx, n, type_id):
        a = _abi.PermuteArgs(_ptr(out), _ptr(inp), _ptr(idx), n, type_id)
        return self.check(self.lib.vo_permute(self.h, C.byref(a)))

    def compact(self, mask, cols, col_bytes, n, n_pad, out_count=None, noalloc=False):
        """vo_compact; with noalloc, vo_compact_noalloc (asynchronous: the
        results are ready after sync())."""
        a = _abi.CompactArgs()
        a.mask = _ptr(mask)
        for i, c in enumerate(cols):
            a.cols[i] = _ptr(c)
            a.col_bytes[i] = col_bytes[i]
        a.n_cols = len(cols)
        a.n, a.n_pad = n, n_pad
        a.out_count = _ptr(out_count)
        fn = self.lib.vo_compact_noalloc if noalloc else self.lib.vo_compact
        return self.check(fn(self.h, C.byref(a)))

    def shuffle(self, payload, n_pad, seed, perm_out=None, n=0):
        a = _abi.ShuffleArgs(_ptr(payload), _ptr(perm_out), n_pad, seed, n)
        return self.check(self.lib.vo_shuffle(self.h, C.byref(a)))

    def filter_sum(self, day, qty, price, disc, n, n_pad, out, out_price=None, out_disc=None,
                   day_lo=8766, day_hi=9131, disc_lo=5, disc_hi=7, qty_lt=24):
        a = _abi.FilterArgs(_ptr(day), _ptr(qty), _ptr(price), _ptr(disc), _ptr(out_price), _ptr(out_disc),
                            n, n_pad, day_lo, day_hi, disc_lo, disc_hi, qty_lt, _ptr(out))
        return self.check(self.lib.vo_filter_sum(self.h, C.byref(a)))

    def join_groupsum(self, dim_key, dim_group, n_dim, fact_fk, fact_val, n_fact, n_pad,
                      out_group, out_sum, out_n_groups, max_groups, sorted_keys=None, sorted_vals=None,
                      dense_groups=False):
        a = _abi.JoinArgs(_ptr(dim_key), _ptr(dim_group), n_dim, _ptr(fact_fk), _ptr(fact_val), n_fact, n_pad,
                          _ptr(out_group), _ptr(out_sum), _ptr(out_n_groups), max_groups,
                          _ptr(sorted_keys), _ptr(sorted_vals), 1 if dense_groups else 0)
        return self.check(self.lib.vo_pkfk_join_groupsum(self.h, C.byref(a)))

    def join_pairs_count(self, left_keys, n_left, right_keys, n_right, type_id, out_count=None, state=None):
        """OP_SORT_MERGE_JOIN_COUNT: returns the number of (left, right) row
        pairs with equal keys; the routing state goes into `state` (a device
        buffer of join_pairs_state_bytes() bytes) or stays in the context."""
        a = _abi.JoinPairsArgs(_ptr(left_keys), n_left, _ptr(right_keys), n_right, type_id, 0, _ptr(out_count),
                               None, None, 0, _ptr(state), 0)
        n = C.c_uint64(0)
        self.check(self.lib.vo_join_pairs_count(self.h, C.byref(a), C.byref(n)))
        return n.value

    def join_pairs_write(self, left_keys, n_left, right_keys, n_right, type_id, out_left, out_right, out_cap,
                         state=None, n_out=0):
        """OP_SORT_MERGE_JOIN_WRITE after the COUNT of the same columns:
        out_left/out_right (int64, out_cap rows) get the pairs ordered by
        (key, left row, right row), then -1.  With `state`, n_out is the
        count that COUNT returned."""
        a = _abi.JoinPairsArgs(_ptr(left_keys), n_left, _ptr(right_keys), n_right, type_id, 0, None,
                               _ptr(out_left), _ptr(out_right), out_cap, _ptr(state), n_out)
        return self.check(self.lib.vo_join_pairs_write(self.h, C.byref(a)))

    def join_pairs_state_bytes(self, n_left, n_right):
        return int(self.lib.vo_join_pairs_state_bytes(n_left, n_right))

    def join_pairs(self, left_keys, right_keys, type_id):
        """COUNT, allocate next_pow2(n_out) rows as the reference's operator
        does (sort_merge_join_operator.cpp:126-133), WRITE.  Returns
        (n_out, out_left, out_right) as int64 device tensors."""
        import torch
        nl, nr = int(left_keys.numel()), int(right_keys.numel())
        n = self.join_pairs_count(left_keys, nl, right_keys, nr, type_id)
        cap = 1
        while cap < n:
            cap <<= 1
        ol = torch.empty(cap, dtype=torch.int64, device=left_keys.device)
        orr = torch.empty(cap, dtype=torch.int64, device=left_keys.device)
        self.join_pairs_write(left_keys, nl, right_keys, nr, type_id, ol, orr, cap)
        return n, ol, orr

    def gather_oblivious(self, table, n_table, record_ids, n, out, elem_bytes, out_found=None):
        """OP_GATHER_OBLIVIOUS: out[i] = table[record_ids[i]] (0 and
        found = 0 for ids outside the table)."""
        a = _abi.GatherObliviousArgs(_ptr(table), n_table, _ptr(record_ids), n, _ptr(out), _ptr(out_found),
                                     elem_bytes, 0)
        return self.check(self.lib.vo_gather_oblivious(self.h, C.byref(a)))

    def predicate_mask(self, terms, out_mask, n, n_pad, connective=_abi.CONN_AND):
        """OP_PREDICATE_MASK: out_mask[i] = 1 for rows i < n meeting all
        (CONN_AND) or any (CONN_OR) of the terms with no NULL term, else 0.
        terms: (column tensor, type id, cmp, constant value[, valid bytes])."""
        arr = (_abi.PredicateTerm * max(1, len(terms)))()
        for j, t in enumerate(terms):
            col, type_id, cmp, value = t[:4]
            valid = t[4] if len(t) > 4 else None
            arr[j] = _abi.PredicateTerm(_ptr(col), _ptr(valid), _abi.constant_bits(type_id, value), type_id, cmp)
        a = _abi.PredicateArgs(arr, len(terms), connective, _ptr(out_mask), n, n_pad)
        return self.check(self.lib.vo_predicate_mask(self.h, C.byref(a)))

    def transfer_to_device(self, dst, src_host, padded_bytes=None):
        """Pageable numpy array -> device tensor, padded_bytes over the link
        (default: the tensor's size); returns the bytes moved."""
        src_host = np.ascontiguousarray(src_host)
        nb = src_host.nbytes
        pb = dst.numel() * dst.element_size() if padded_bytes is None else padded_bytes
        moved = C.c_uint64(0)
        self.check(self.lib.vo_transfer_to_device(self.h, _ptr(dst), src_host.ctypes.data_as(C.c_void_p), nb, pb,
                                                  C.byref(moved)))
        return moved.value

    def transfer_to_host(self, dst_host, src, padded_bytes=None):
        """Device tensor -> pageable numpy array (dst_host.nbytes bytes),
        padded_bytes over the link; returns the bytes moved."""
        nb = dst_host.nbytes
        pb = src.numel() * src.element_size() if padded_bytes is None else padded_bytes
        moved = C.c_uint64(0)
        self.check(self.lib.vo_transfer_to_host(self.h, dst_host.ctypes.data_as(C.c_void_p), _ptr(src), nb, pb,
                                                C.byref(moved)))
        return moved.value

    def sort_batch(self, keys, payload, items):
        arr = (_abi.BatchItem * len(items))(*[_abi.BatchItem(o, n, p) for (o, n, p) in items])
        a = _abi.SortBatchArgs(_ptr(keys), _ptr(payload), C.cast(arr, C.c_void_p), len(items))
        return self.check(self.lib.vo_sort_batch(self.h, C.byref(a)))

    @staticmethod
    def compact_batch_args(mask, cols, col_bytes, items, out_counts=None):
        """vo_compact_batch_args for items [(offset, n, n_pad), ...]; the
        returned tuple keeps the host item array alive with the struct."""
        arr = (_abi.BatchItem * len(items))(*[_abi.BatchItem(o, n, p) for (o, n, p) in items])
        a = _abi.CompactBatchArgs()
        a.mask = _ptr(mask)
        for i, c in enumerate(cols):
            a.cols[i] = _ptr(c)
            a.col_bytes[i] = col_bytes[i]
        a.n_cols = len(cols)
        a.items = C.cast(arr, C.c_void_p)
        a.n_items = len(items)
        a.out_counts = _ptr(out_counts)
        return a, arr

    def compact_batch(self, mask, cols, col_bytes, items, out_counts=None):
        a, _arr = self.compact_batch_args(mask, cols, col_bytes, items, out_counts)
        return self.check(self.lib.vo_compact_batch(self.h, C.byref(a)))

    def queue_start(self):
        return self.check(self.lib.vo_queue_start(self.h))

    def queue_stop(self, timeout_ms=5000):
        return self.check(self.lib.vo_queue_stop(self.h, timeout_ms))

    def queue_running(self):
        return bool(self.lib.vo_queue_running(self.h))

    def poll(self, seq):
        r = self.lib.vo_poll(self.h, seq)
        if r < 0:
            raise VoError(-r, self.lib.vo_last_error(self.h).decode())
        return bool(r)

    def submit(self, cmd, lane=None):
        """vo_submit, or vo_submit_lane on one lane of the resident queue."""
        if lane is None:
            s = self.lib.vo_submit(self.h, C.byref(cmd))
        else:
            s = self.lib.vo_submit_lane(self.h, C.byref(cmd), lane)
        if s == 0:
            raise VoError(-1, self.lib.vo_last_error(self.h).decode())
        return s

    def queue_lanes(self):
        return self.lib.vo_queue_lanes(self.h)

    def wait(self, seq, timeout_ms=0):
        return self.check(self.lib.vo_wait(self.h, seq, timeout_ms))

    def sync(self):
        return self.check(self.lib.vo_sync(self.h))

    def autotune_sort(self, log_min, log_max, rows_per_class=1 << 24):
        out = (C.c_uint32 * (log_max - log_min + 1))()
        self.check(self.lib.vo_autotune_sort(self.h, log_min, log_max, rows_per_class, out))
        return {log_min + i: int(w) for i, w in enumerate(out)}

    def set_class_tuning(self, log_n, word):
        return self.check(self.lib.vo_set_class_tuning(self.h, log_n, word))

    def get_class_tuning(self, log_n):
        w = C.c_uint32()
        self.check(self.lib.vo_get_class_tuning(self.h, log_n, C.byref(w)))
        return w.value

    def hbm_copy(self, dst, src, nbytes):
        return self.check(self.lib.vo_hbm_copy(self.h, _ptr(dst), _ptr(src), nbytes))

    def gen_sort_input(self, seed, keys, payload, n):
        return self.check(self.lib.vo_gen_sort_input(self.h, seed, _ptr(keys), _ptr(payload), n))

    def gen_fact(self, seed, day, qty, price, disc, n):
        return self.check(self.lib.vo_gen_fact(self.h, seed, _ptr(day), _ptr(qty), _ptr(price), _ptr(disc), n))

    # ---- diagnostics -----------------------------------------------------
    def trace_begin(self):
        return self.check(self.lib.vo_trace_begin(self.h))

    def trace_end(self):
        t = _abi.Trace()
        self.check(self.lib.vo_trace_end(self.h, C.byref(t)))
        return t

    def trace_launches(self, n):
        out = []
        name = C.create_string_buffer(64)
        g = (C.c_uint32 * 3)()
        b = (C.c_uint32 * 3)()
        d = C.c_uint64()
        for i in range(n):
            self.check(self.lib.vo_trace_launch(self.h, i, name, 64, g, b, C.byref(d)))
            out.append((name.value.decode(), tuple(g), tuple(b), d.value))
        return out

    def kernel_stats(self, family=None):
        n = C.c_uint64()
        ms = C.c_double()
        fam = family.encode() if family else None
        self.check(self.lib.vo_kernel_stats(self.h, fam, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def kernel_stats_reset(self):
        return self.check(self.lib.vo_kernel_stats_reset(self.h))

    def sort_plan(self, n_pad):
        n = C.c_uint32()
        self.check(self.lib.vo_sort_plan(self.h, n_pad, C.byref(n)))
        return n.value
