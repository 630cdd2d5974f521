"""GPU parity of the oblivious bitonic sort (through the C ABI) against the
CPU oracle: bit-exact outputs, value-independent launch/address traces."""
import numpy as np
import pytest

import oracle
from tests.gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu

M64 = np.uint64((1 << 64) - 1)


def _run_sort(vctx, k, p, n, npad):
    dk = to_dev(k, npad)
    dp = to_dev(p, npad)
    vctx.sort_kp(dk, dp, n, npad)
    return from_dev(dk, np.uint64), from_dev(dp, np.uint32)


@pytest.mark.parametrize("n,npad", [(0, 1), (1, 1), (2, 2), (3, 4), (7, 8), (16, 16), (17, 32), (1000, 1024),
                                    (4095, 4096), (8192, 8192), (8193, 16384), (50000, 65536),
                                    (1 << 17, 1 << 17), (300000, 1 << 19)])
def test_sort_kp_bit_exact(vctx, n, npad):
    print(f"sort_kp n={n} npad={npad}", flush=True)
    k, p = oracle.gen_sort_input(1000 + n, n)
    gk, gp = _run_sort(vctx, k, p, n, npad)
    ek, ep = oracle.bitonic_sort_kp(k, p, n, npad)
    assert np.array_equal(gk, ek)
    assert np.array_equal(gp, ep)


def test_sort_2p20_two_seeds(vctx):
    """2^20 records, u64 uniform keys, payload = row index, seeds 1 and 2."""
    n = 1 << 20
    for seed in (1, 2):
        print(f"sort 2^20 seed={seed}", flush=True)
        k, p = oracle.gen_sort_input(seed, n)
        gk, gp = _run_sort(vctx, k, p, n, n)
        ek, ep = oracle.bitonic_sort_kp(k, p, n, n)
        assert np.array_equal(gk, ek) and np.array_equal(gp, ep)


def test_sort_duplicate_keys_ties_by_payload(vctx):
    rng = np.random.default_rng(2)
    n = 1 << 15
    k = rng.integers(0, 5, n).astype(np.uint64)
    p = rng.permutation(n).astype(np.uint32)
    gk, gp = _run_sort(vctx, k, p, n, n)
    order = np.lexsort((p, k))
    assert np.array_equal(gk, k[order]) and np.array_equal(gp, p[order])


def test_sort_extreme_keys(vctx):
    # keys at 0 and 2^64-1 next to pad rows; payload = row index keeps them ahead of pads
    n, npad = 1000, 1024
    k = np.where(np.arange(n) % 3 == 0, M64, np.uint64(0)).astype(np.uint64)
    p = np.arange(n, dtype=np.uint32)
    gk, gp = _run_sort(vctx, k, p, n, npad)
    ek, ep = oracle.bitonic_sort_kp(k, p, n, npad)
    assert np.array_equal(gk, ek) and np.array_equal(gp, ep)
    assert gp[:n].max() < n


def test_device_generator_matches_oracle(vctx):
    n = 100003
    dk = empty_dev(n, np.uint64)
    dp = empty_dev(n, np.uint32)
    vctx.gen_sort_input(7, dk, dp, n)
    k, p = oracle.gen_sort_input(7, n)
    assert np.array_equal(from_dev(dk, np.uint64), k) and np.array_equal(from_dev(dp, np.uint32), p)


def _trace(vctx, seed, n, npad):
    k, p = oracle.gen_sort_input(seed, n)
    dk = to_dev(k, npad)
    dp = to_dev(p, npad)
    vctx.trace_begin()
    vctx.sort_kp(dk, dp, n, npad)
    t = vctx.trace_end()
    return t, vctx.trace_launches(t.n_launches)


@pytest.mark.parametrize("n,npad", [(1 << 20, 1 << 20), (5000, 8192), (100, 128)])
def test_sort_trace_value_independent(vctx, n, npad):
    """Launch sequence, grid shapes and per-launch address digests are equal
    for two inputs of equal padded size and different values."""
    print(f"trace n={n}", flush=True)
    t1, l1 = _trace(vctx, 1, n, npad)
    t2, l2 = _trace(vctx, 2, n, npad)
    assert t1.n_launches == t2.n_launches > 0
    assert l1 == l2
    assert t1.launch_digest == t2.launch_digest
    assert t1.address_digest == t2.address_digest
    # every launch folded addresses: each record is read and written once per pass
    assert t1.n_addresses >= 2 * npad * t1.n_launches


def test_sort_trace_sees_a_different_shape(vctx):
    # sanity: the digest is sensitive (a different padded size changes it)
    t1, _ = _trace(vctx, 1, 4000, 4096)
    t2, _ = _trace(vctx, 1, 4000, 8192)
    assert t1.launch_digest != t2.launch_digest


def test_sort_large_properties(vctx):
    """2^24 records: size-independent properties (sorted, payload a permutation)."""
    import torch
    n = 1 << 24
    dk = empty_dev(n, np.uint64)
    dp = empty_dev(n, np.uint32)
    vctx.gen_sort_input(3, dk, dp, n)
    vctx.sort_kp(dk, dp, n, n)
    k = from_dev(dk, np.uint64)
    p = from_dev(dp, np.uint32)
    assert np.all(k[1:] >= k[:-1])
    assert np.array_equal(np.sort(p), np.arange(n, dtype=np.uint32))
    # keys travelled with their payloads
    kk, _ = oracle.gen_sort_input(3, n)
    assert np.array_equal(kk[p], k)
    del torch


# ----------------------------------------------------- multi-column OP_SORT

TYPES = {1: np.uint8, 2: np.int8, 3: np.int16, 4: np.int32, 5: np.int64, 6: np.float32, 7: np.float64,
         8: np.uint64, 10: np.uint32, 64: np.uint64}


def _rand_col(rng, t, n, distinct):
    dt = TYPES[t]
    if t in (6, 7):
        return (rng.integers(-distinct, distinct, n) * 0.5).astype(dt)
    info = np.iinfo(dt)
    lo, hi = max(int(info.min), -distinct), min(int(info.max), distinct)
    return rng.integers(lo, hi, n, endpoint=True).astype(dt)


@pytest.mark.parametrize("types,asc,n,npad,distinct", [
    ([4], [True], 1000, 1024, 50),
    ([5, 4], [True, False], 3000, 4096, 10),
    ([7, 8], [False, True], 5000, 8192, 20),
    ([1, 3, 2], [True, True, False], 20000, 32768, 3),
    ([6, 10, 5, 4], [False, True, True, False], 9000, 16384, 4),
    ([4, 4, 4, 4, 4, 4, 4, 4], [True, False] * 4, 2000, 2048, 2),
    ([64], [True], 70000, 1 << 17, 1 << 30),
    # levels of 4-5 stages above the tile run as hybrid passes: 2 and 4 key words
    ([5, 4], [True, False], 100000, 1 << 17, 10),
    ([6, 10, 5, 4], [False, True, True, False], 70000, 1 << 17, 4),
])
def test_sort_multi_bit_exact_with_ties(vctx, types, asc, n, npad, distinct):
    """Reference OP_SORT semantics incl. ties (decided by the network) and the
    max/lowest pad rows; indices and every column bit-exact."""
    print(f"sort_multi types={types} n={n}", flush=True)
    rng = np.random.default_rng(len(types) * 1000 + n)
    cols = [_rand_col(rng, t, npad, distinct) for t in types]
    ecols, eidx = oracle.bitonic_sort_multi(cols, types, asc, n, npad)
    dcols = [to_dev(c) for c in cols]
    didx = empty_dev(npad, np.int64)
    vctx.sort(dcols, types, asc, didx, n, npad)
    assert np.array_equal(from_dev(didx, np.int64), eidx)
    for dc, ec, t in zip(dcols, ecols, types):
        assert np.array_equal(from_dev(dc, TYPES[t]).view(np.uint8), ec.view(np.uint8))


@pytest.mark.parametrize("types,asc,n,npad,distinct", [
    ([5, 4], [True, False], 100000, 1 << 17, 10),
    ([6, 10, 5, 4], [False, True, True, False], 70000, 1 << 17, 4),
    ([4, 4, 4, 4, 4, 4, 4, 4], [True, False] * 4, 20000, 1 << 15, 2),
])
def test_sort_multi_streaming_bit_exact(types, asc, n, npad, distinct):
    """OP_SORT with the sort passes' streaming loads and stores (tuning 0x800
    picks them below the size that would): indices and every column equal
    the oracle's, ties included."""
    import torch
    import velodb_amd
    print(f"sort_multi streaming types={types} n={n}", flush=True)
    rng = np.random.default_rng(len(types) * 1000 + n + 7)
    cols = [_rand_col(rng, t, npad, distinct) for t in types]
    ecols, eidx = oracle.bitonic_sort_multi(cols, types, asc, n, npad)
    dcols = [to_dev(c) for c in cols]
    didx = empty_dev(npad, np.int64)
    with velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream, tuning=0x800) as ctx:
        ctx.sort(dcols, types, asc, didx, n, npad)
    assert np.array_equal(from_dev(didx, np.int64), eidx)
    for dc, ec, t in zip(dcols, ecols, types):
        assert np.array_equal(from_dev(dc, TYPES[t]).view(np.uint8), ec.view(np.uint8))


@pytest.mark.parametrize("tuning", [0x100, 0x5 | (1 << 4), 0x3 | (7 << 4), 0x1 | (1 << 4), 0x4 | (8 << 4),
                                    0x200, 0x200 | 0x3 | (7 << 4), 0x200 | 0x100, 0x400, 0x800, 0x800 | 0x200])
def test_sort_schedule_variants_bit_exact(tuning):
    """Every schedule variant (fused-stage widths, hybrid on/off, resident
    merge blocks, 1024-thread blocks, streaming loads and stores below the
    size that picks them) applies the same network: bit-exact at 2^20,
    2^17+3, one full tile and a partial tile."""
    import torch
    import velodb_amd
    print(f"schedule tuning={tuning:#x}", flush=True)
    with velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream, tuning=tuning) as ctx:
        for n, npad in [(1 << 20, 1 << 20), ((1 << 17) + 3, 1 << 18), (1 << 13, 1 << 13), (5000, 8192)]:
            k, p = oracle.gen_sort_input(77 + n, n)
            gk, gp = _run_sort(ctx, k, p, n, npad)
            ek, ep = oracle.bitonic_sort_kp(k, p, n, npad)
            assert np.array_equal(gk, ek) and np.array_equal(gp, ep)


# ------------------------------------- OP_SORT edge cases vs the oracle

def _extremes(t):
    """Values at and next to the pad values of a type: lowest, lowest+1, 0,
    max-1, max (and +-0.0, +-inf for floating point)."""
    dt = TYPES[t]
    if t in (6, 7):
        fi = np.finfo(dt)
        return np.array([fi.min, np.nextafter(fi.min, dt(0)), -0.0, 0.0, np.nextafter(fi.max, dt(0)), fi.max,
                         -np.inf, np.inf], dtype=dt)
    ii = np.iinfo(dt)
    return np.array([ii.min, ii.min + 1, 0, ii.max - 1, ii.max], dtype=dt)


@pytest.mark.parametrize("t", [1, 2, 3, 4, 5, 6, 7, 8, 10, 64])
@pytest.mark.parametrize("asc", [True, False])
def test_sort_multi_rows_at_pad_values(vctx, t, asc):
    """Real rows at numeric max / lowest tie with the pad rows (sort.cuh:79-100):
    the network decides where they land, so the indices and the columns must
    match the reference's rule bit for bit, for either direction."""
    print(f"sort_multi extremes type={t} asc={asc}", flush=True)
    rng = np.random.default_rng(t * 2 + asc)
    n, npad = 3000, 4096
    vals = _extremes(t)
    c0 = vals[rng.integers(0, len(vals), npad)]
    c1 = rng.integers(0, 3, npad).astype(np.int32)  # second key: partial tie-breaker
    for types, cols, ascs in (([t], [c0], [asc]), ([t, 4], [c0, c1], [asc, not asc])):
        ecols, eidx = oracle.bitonic_sort_multi(cols, types, ascs, n, npad)
        dcols = [to_dev(c) for c in cols]
        didx = empty_dev(npad, np.int64)
        vctx.sort(dcols, types, ascs, didx, n, npad)
        assert np.array_equal(from_dev(didx, np.int64), eidx)
        for dc, ec, tt in zip(dcols, ecols, types):
            assert np.array_equal(from_dev(dc, TYPES[tt]).view(np.uint8), ec.view(np.uint8))


@pytest.mark.parametrize("types,asc", [([6], [True]), ([7], [False]), ([6, 7], [True, False]),
                                       ([4, 7, 6], [False, True, True])])
def test_sort_multi_negative_zero_keeps_its_bits(vctx, types, asc):
    """-0.0 and +0.0 compare equal (compareSingle, sort.cuh:11-29) and every
    row keeps its own bits: the output equals the oracle's byte for byte."""
    print(f"sort_multi -0.0 types={types}", flush=True)
    rng = np.random.default_rng(len(types))
    n, npad = 7000, 8192
    cols = []
    for t in types:
        if t in (6, 7):
            cols.append(np.array([-0.0, 0.0, 1.5, -1.5], dtype=TYPES[t])[rng.integers(0, 4, npad)])
        else:
            cols.append(rng.integers(0, 3, npad).astype(TYPES[t]))
    ecols, eidx = oracle.bitonic_sort_multi(cols, types, asc, n, npad)
    dcols = [to_dev(c) for c in cols]
    didx = empty_dev(npad, np.int64)
    vctx.sort(dcols, types, asc, didx, n, npad)
    gidx = from_dev(didx, np.int64)
    assert np.array_equal(gidx, eidx)
    for dc, ec, c, t in zip(dcols, ecols, cols, types):
        g = from_dev(dc, TYPES[t])
        assert np.array_equal(g.view(np.uint8), ec.view(np.uint8))
        if t in (6, 7):  # every real row carries its own sign of zero
            live = gidx < n
            assert np.array_equal(np.signbit(g[live]), np.signbit(c[gidx[live]]))


def test_sort_multi_nan_takes_its_total_order_place(vctx):
    """The documented deviation: NaN (unordered in the reference) sorts by its
    IEEE total-order image, above +inf and, with the sign bit, below -inf."""
    neg_nan = np.array([0xFFC00000], np.uint32).view(np.float32)[0]
    vals = np.array([np.nan, np.inf, -np.inf, 1.0, -2.0, neg_nan, 0.5, 3.0], np.float32)
    n = npad = len(vals)  # no pad rows: every value is distinct, so the order is unique
    for asc in (True, False):
        dcol = to_dev(vals, npad)
        didx = empty_dev(npad, np.int64)
        vctx.sort([dcol], [6], [asc], didx, n, npad)
        gidx = from_dev(didx, np.int64)
        order = [5, 2, 4, 6, 3, 7, 1, 0]  # -NaN, -inf, -2, 0.5, 1, 3, inf, NaN
        expect = order if asc else order[::-1]
        assert list(gidx) == expect
        got = from_dev(dcol, np.float32)
        assert np.array_equal(got.view(np.uint32), vals[expect].view(np.uint32))


def test_sort_multi_rejects_more_fp_columns_than_sign_bits(vctx):
    """At 2^29 rows the payload has 3 spare bits: a fourth FLOAT/DOUBLE key is
    refused on the host before any launch."""
    import velodb_amd
    dummy = [empty_dev(16, np.float32) for _ in range(4)]
    didx = empty_dev(16, np.int64)
    with pytest.raises(velodb_amd.VoError) as e:
        vctx.sort(dummy, [6, 6, 7, 6], [True] * 4, didx, 0, 1 << 29)
    assert e.value.status == 5  # VO_ERR_UNSUPPORTED
