"""Python access to the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / CPU baseline.  The product
library never imports or calls this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
REF_KAT = os.path.join(_HERE, "_ref", "kks_test_sorts")
REF_DIST = os.path.join(_HERE, "_ref", "kks_test_obliv_distribute")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        u64, u32, i64, i32, sz = C.c_uint64, C.c_uint32, C.c_int64, C.c_int32, C.c_size_t
        L.orc_mix64.restype = u64
        L.orc_mix64.argtypes = [u64]
        L.orc_rand64.restype = u64
        L.orc_rand64.argtypes = [u64, u32, u64]
        L.orc_rand_range.restype = i64
        L.orc_rand_range.argtypes = [u64, u32, u64, i64, u64]
        L.orc_philox4x32_10.restype = None
        L.orc_philox4x32_10.argtypes = [P, P, P]
        L.orc_philox_tag.restype = u64
        L.orc_philox_tag.argtypes = [u64, u64]
        L.orc_gen_sort_input.restype = None
        L.orc_gen_sort_input.argtypes = [u64, sz, P, P]
        L.orc_bitonic_sort_kp.restype = None
        L.orc_bitonic_sort_kp.argtypes = [P, P, sz, sz, P]
        L.orc_bitonic_sort_kp_mt.restype = None
        L.orc_bitonic_sort_kp_mt.argtypes = [P, P, sz, sz, C.c_int]
        L.orc_bitonic_sort_kv.restype = None
        L.orc_bitonic_sort_kv.argtypes = [P, P, sz]
        L.orc_bitonic_sort_multi.restype = C.c_int
        L.orc_bitonic_sort_multi.argtypes = [P, P, P, sz, P, sz, sz]
        L.orc_permute.restype = None
        L.orc_permute.argtypes = [P, P, P, sz, sz]
        L.orc_compact.restype = None
        L.orc_compact.argtypes = [P, sz, P, sz]
        L.orc_compact_levels.restype = None
        L.orc_compact_levels.argtypes = [P, sz, P, sz]
        L.orc_gen_fact.restype = None
        L.orc_gen_fact.argtypes = [u64, sz, P, P, P, P]
        L.orc_filter_mask.restype = None
        L.orc_filter_mask.argtypes = [P, P, P, sz, sz, P, P]
        L.orc_filter_sum.restype = sz
        L.orc_filter_sum.argtypes = [P, P, P, P, sz, sz, P, P, P, P, P]
        L.orc_gen_join.restype = None
        L.orc_gen_join.argtypes = [u64, sz, sz, u32, i64, i64, P, P, P, P]
        L.orc_union_key.restype = u64
        L.orc_union_key.argtypes = [i32, C.c_int, u32]
        L.orc_join_groupsum.restype = sz
        L.orc_join_groupsum.argtypes = [P, P, sz, P, P, sz, sz, P, P, P, P, sz]
        L.orc_distribute.restype = C.c_int
        L.orc_distribute.argtypes = [P, P, sz, sz, P, P]
        L.orc_gather_recover.restype = None
        L.orc_gather_recover.argtypes = [P, sz, P, sz, P, P]
        L.orc_join_pairs.restype = sz
        L.orc_join_pairs.argtypes = [P, sz, P, sz, P, P, sz, P]
        L.orc_predicate_mask.restype = C.c_int
        L.orc_predicate_mask.argtypes = [P, P, P, P, P, u32, u32, sz, sz, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Trace(C.Structure):
    _fields_ = [("hash", C.c_uint64), ("n_ops", C.c_uint64), ("n_stages", C.c_uint64)]


class FilterSpec(C.Structure):
    _fields_ = [("day_lo", C.c_int32), ("day_hi", C.c_int32), ("disc_lo", C.c_int64), ("disc_hi", C.c_int64),
                ("qty_lt", C.c_int64)]


Q6 = dict(day_lo=8766, day_hi=9131, disc_lo=5, disc_hi=7, qty_lt=24)


def gen_sort_input(seed, n):
    k = np.empty(n, np.uint64)
    p = np.empty(n, np.uint32)
    lib().orc_gen_sort_input(seed, n, _p(k), _p(p))
    return k, p


def bitonic_sort_kp(keys, payload, n, npad, trace=False):
    """Returns (keys, payload[, Trace]) sorted over npad rows (copies)."""
    k = np.empty(npad, np.uint64)
    p = np.empty(npad, np.uint32)
    k[:n] = keys[:n]
    p[:n] = payload[:n]
    t = Trace() if trace else None
    lib().orc_bitonic_sort_kp(_p(k), _p(p), n, npad, C.byref(t) if trace else None)
    return (k, p, t) if trace else (k, p)


def bitonic_sort_kp_mt(keys, payload, n, npad, threads):
    """Multi-threaded recursive bitonic (the CPU baseline); same output as
    bitonic_sort_kp."""
    k = np.empty(npad, np.uint64)
    p = np.empty(npad, np.uint32)
    k[:n] = keys[:n]
    p[:n] = payload[:n]
    lib().orc_bitonic_sort_kp_mt(_p(k), _p(p), n, npad, threads)
    return k, p


def bitonic_sort_kp_mt_inplace(k, p, n, threads):
    """bitonic_sort_kp_mt on caller arrays (uint64 keys, uint32 payload, both
    of the padded length), sorted in place with no copy."""
    assert k.dtype == np.uint64 and p.dtype == np.uint32 and len(k) == len(p)
    assert k.flags.c_contiguous and p.flags.c_contiguous and n <= len(k)
    lib().orc_bitonic_sort_kp_mt(_p(k), _p(p), n, len(k), threads)


def bitonic_sort_kv(keys, vals):
    k = np.ascontiguousarray(keys, np.uint64).copy()
    v = np.ascontiguousarray(vals, np.uint64).copy()
    lib().orc_bitonic_sort_kv(_p(k), _p(v), len(k))
    return k, v


def bitonic_sort_multi(cols, types, asc, n, npad):
    """cols: list of numpy arrays of length >= npad (copied).  Returns (cols, idx)."""
    cs = [np.ascontiguousarray(c).copy() for c in cols]
    ptrs = (C.c_void_p * 8)(*[c.ctypes.data for c in cs])
    ty = (C.c_uint32 * 8)(*types)
    fl = (C.c_uint8 * 8)(*[1 if a else 0 for a in asc])
    idx = np.empty(npad, np.int64)
    rc = lib().orc_bitonic_sort_multi(ptrs, ty, fl, len(cs), _p(idx), n, npad)
    if rc != 0:
        raise ValueError("orc_bitonic_sort_multi rejected its arguments")
    return cs, idx


def compact(buf, selected):
    """buf: 2-D array (rows x bytes-per-row view) compacted in place copy."""
    b = np.ascontiguousarray(buf).copy()
    s = np.ascontiguousarray(selected, np.uint8)
    lib().orc_compact(_p(b), b.itemsize * (b.shape[1] if b.ndim == 2 else 1), _p(s), len(s))
    return b


def compact_levels(buf, selected):
    b = np.ascontiguousarray(buf).copy()
    s = np.ascontiguousarray(selected, np.uint8)
    lib().orc_compact_levels(_p(b), b.itemsize * (b.shape[1] if b.ndim == 2 else 1), _p(s), len(s))
    return b


def gen_fact(seed, n):
    day = np.empty(n, np.int32)
    qty = np.empty(n, np.int64)
    price = np.empty(n, np.int64)
    disc = np.empty(n, np.int64)
    lib().orc_gen_fact(seed, n, _p(day), _p(qty), _p(price), _p(disc))
    return day, qty, price, disc


def filter_sum(day, qty, price, disc, n, npad, spec=Q6, want_rows=False):
    """Returns (count, int128 sum as Python int[, out_price, out_disc])."""
    def padded(a, dt):
        out = np.zeros(npad, dt)
        out[:n] = a[:n]
        return out
    d, q, pr, di = padded(day, np.int32), padded(qty, np.int64), padded(price, np.int64), padded(disc, np.int64)
    op = np.empty(npad, np.int64) if want_rows else None
    od = np.empty(npad, np.int64) if want_rows else None
    lo, hi = C.c_uint64(), C.c_int64()
    fs = FilterSpec(**spec)
    m = lib().orc_filter_sum(_p(d), _p(q), _p(pr), _p(di), n, npad, C.byref(fs), _p(op), _p(od),
                             C.byref(lo), C.byref(hi))
    s = (hi.value << 64) | lo.value
    return (m, s, op, od) if want_rows else (m, s)


def filter_mask(day, qty, disc, n, npad, spec=Q6):
    m = np.empty(npad, np.uint8)
    fs = FilterSpec(**spec)

    def padded(a, dt):
        out = np.zeros(npad, dt)
        out[:n] = a[:n]
        return out
    lib().orc_filter_mask(_p(padded(day, np.int32)), _p(padded(qty, np.int64)), _p(padded(disc, np.int64)),
                          n, npad, C.byref(fs), _p(m))
    return m


def gen_join(seed, n_dim, n_fact, n_groups=25, vlo=100, vhi=10_000_000):
    dk = np.empty(n_dim, np.int32)
    dg = np.empty(n_dim, np.int32)
    fk = np.empty(n_fact, np.int32)
    fv = np.empty(n_fact, np.int64)
    lib().orc_gen_join(seed, n_dim, n_fact, n_groups, vlo, vhi, _p(dk), _p(dg), _p(fk), _p(fv))
    return dk, dg, fk, fv


def join_groupsum(dk, dg, fk, fv, npad, max_groups=64, want_sorted=False):
    sk = np.empty(npad, np.uint64) if want_sorted else None
    sv = np.empty(npad, np.uint64) if want_sorted else None
    og = np.empty(max_groups, np.int32)
    os_ = np.empty(max_groups, np.int64)
    G = lib().orc_join_groupsum(_p(dk), _p(dg), len(dk), _p(fk), _p(fv), len(fk), npad, _p(sk), _p(sv),
                                _p(og), _p(os_), max_groups)
    out = (og[:G].copy(), os_[:G].copy())
    return out + (sk, sv) if want_sorted else out


# Order-preserving 64-bit codes of the join key types (vo_type ids): signed
# types flip the sign bit, unsigned ones widen.
_SIGNED = {2: np.int8, 3: np.int16, 4: np.int32, 5: np.int64}
_UNSIGNED = {1: np.uint8, 10: np.uint32, 8: np.uint64, 64: np.uint64}


def key_codes(keys, type_id):
    if type_id in _SIGNED:
        v = np.asarray(keys, dtype=_SIGNED[type_id]).astype(np.int64).view(np.uint64)
        return v ^ np.uint64(1 << 63)
    if type_id in _UNSIGNED:
        return np.asarray(keys, dtype=_UNSIGNED[type_id]).astype(np.uint64)
    raise ValueError("join keys are integer types")


def distribute(targets, payload, m):
    """KKS oblivious distribution; returns (targets, payload) of the m
    slots (-1 / 0 for dummies), or raises if the invariant fails."""
    t = np.ascontiguousarray(targets, dtype=np.int64)
    p = np.ascontiguousarray(payload, dtype=np.uint64)
    ot = np.empty(m, np.int64)
    op = np.empty(m, np.uint64)
    if lib().orc_distribute(_p(t), _p(p), len(t), m, _p(ot), _p(op)) != 0:
        raise ValueError("distribution invariant failed")
    return ot, op


def join_pairs(lcodes, rcodes):
    """Equi-join of two key-code columns: (left rows, right rows), int64,
    ordered by (key, left row, right row)."""
    lk = np.ascontiguousarray(lcodes, dtype=np.uint64)
    rk = np.ascontiguousarray(rcodes, dtype=np.uint64)
    st = C.c_int(0)
    n_out = lib().orc_join_pairs(_p(lk), len(lk), _p(rk), len(rk), None, None, 0, C.byref(st))
    ol = np.empty(max(n_out, 1), np.int64)
    or_ = np.empty(max(n_out, 1), np.int64)
    n2 = lib().orc_join_pairs(_p(lk), len(lk), _p(rk), len(rk), _p(ol), _p(or_), n_out, C.byref(st))
    if st.value != 0 or n2 != n_out:
        raise ValueError("join oracle invariant failed")
    return ol[:n_out], or_[:n_out]


def gather_recover(table, ids):
    """Oblivious materialization of one column: (values as uint64, found)."""
    t = np.ascontiguousarray(table, dtype=np.uint64)
    r = np.ascontiguousarray(ids, dtype=np.int64)
    out = np.empty(max(len(r), 1), np.uint64)
    found = np.empty(max(len(r), 1), np.uint8)
    lib().orc_gather_recover(_p(t), len(t), _p(r), len(r), _p(out), _p(found))
    return out[:len(r)], found[:len(r)]


def rand64(seed, stream, i):
    return lib().orc_rand64(seed, stream, i)


def philox4x32_10(ctr, key):
    c = np.array(ctr, np.uint32)
    k = np.array(key, np.uint32)
    o = np.empty(4, np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(o))
    return [int(x) for x in o]


def philox_tag(seed, i):
    return lib().orc_philox_tag(seed, i)


def const_bits(type_id, value):
    """Bit pattern of a constant in a column type (vo_type ids)."""
    import struct
    if type_id == 6:
        return struct.unpack("<I", struct.pack("<f", value))[0]
    if type_id == 7:
        return struct.unpack("<Q", struct.pack("<d", value))[0]
    width = {1: 1, 2: 1, 3: 2, 4: 4, 10: 4}.get(type_id, 8)
    return int(value) & ((1 << (8 * width)) - 1)


def predicate_mask(terms, n, npad, conn=0):
    """terms: (numpy column, type id, cmp 0..5, constant value[, valid bytes or
    None]); conn 0 = AND, 1 = OR.  Returns the uint8 mask of npad rows."""
    k = len(terms)
    cols = [np.ascontiguousarray(t[0]) for t in terms]
    valids = [None if len(t) < 5 or t[4] is None else np.ascontiguousarray(t[4], np.uint8) for t in terms]
    cp = (C.c_void_p * max(1, k))(*[c.ctypes.data for c in cols])
    vp = (C.c_void_p * max(1, k))(*[None if v is None else v.ctypes.data for v in valids])
    ty = (C.c_uint32 * max(1, k))(*[t[1] for t in terms])
    cm = (C.c_uint32 * max(1, k))(*[t[2] for t in terms])
    kb = (C.c_uint64 * max(1, k))(*[const_bits(t[1], t[3]) for t in terms])
    m = np.empty(npad, np.uint8)
    rc = lib().orc_predicate_mask(cp, vp, ty, cm, kb, k, conn, n, npad, _p(m))
    if rc != 0:
        raise ValueError("unsupported predicate term")
    return m


# ------------------------------------------------------------ closed forms
# The results the oracle's restatements are pinned against (tests/test_oracle.py
# checks each one equal to its restatement on random inputs).  They run in
# O(n log n) numpy, so tests can check results at the configs' full sizes.

def filter_closed_form(day, qty, price, disc, spec=Q6):
    """Filter + ORCompact + SUM as a stable partition: (count, exact sum,
    selected price, selected disc), the selected rows in row order
    (oblivious_compact.c:59-153 keeps their order; the sum is exact in Python
    integers)."""
    sel = ((day >= spec["day_lo"]) & (day < spec["day_hi"]) & (disc >= spec["disc_lo"]) &
           (disc <= spec["disc_hi"]) & (qty < spec["qty_lt"]))
    p, d = price[sel], disc[sel]
    hi = np.abs(p).max(initial=0) * np.abs(d).max(initial=0)
    if hi < (1 << 62) // max(1, len(p)):  # the int64 dot product cannot overflow
        s = int(np.dot(p, d))
    else:
        s = sum(int(a) * int(b) for a, b in zip(p, d))
    return int(sel.sum()), s, p, d


def join_groupsum_closed_form(dk, dg, fk, fv, max_groups=64):
    """PK/FK join + SUM GROUP BY as an array lookup: every fact row joins the
    one dimension row with its key (keys are unique); (groups ascending,
    int64 sums) over the joined rows whose group lies in [0, max_groups)."""
    dk = np.asarray(dk, np.int64)
    order = np.argsort(dk, kind="stable")
    pos = np.searchsorted(dk[order], fk)
    pos = np.minimum(pos, len(dk) - 1)
    hit = dk[order][pos] == fk
    g = np.asarray(dg, np.int64)[order][pos][hit]
    v = np.asarray(fv, np.int64)[hit]
    keep = (g >= 0) & (g < max_groups)
    g, v = g[keep], v[keep]
    groups = np.unique(g)
    sums = np.array([int(v[g == x].sum()) for x in groups], np.int64)
    return groups.astype(np.int32), sums


def sort_closed_form(keys, payload):
    """Ascending by key, ties by payload: the unique order of distinct
    (key, payload) pairs, which the bitonic network produces."""
    order = np.lexsort((payload, keys))
    return keys[order], payload[order]
