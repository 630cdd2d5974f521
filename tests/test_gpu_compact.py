"""GPU parity of the oblivious compaction (ORCompact network), the Q6-style
filter + exact int128 SUM and OP_SCATTER, against the CPU oracle."""
import numpy as np
import pytest

import oracle
from tests.gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu

WIDTH_DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def _rows(cols):
    """Pack columns into one byte-row array for the oracle (element = all columns)."""
    return np.concatenate([c.view(np.uint8).reshape(len(c), -1) for c in cols], axis=1)


@pytest.mark.parametrize("widths,n,n_pad,p", [
    ([8], 4096, 4096, 0.3),
    ([8, 8], 4000, 4096, 0.02),
    ([4], 10000, 16384, 0.5),
    ([8, 4], 1 << 17, 1 << 17, 0.9),
    ([1, 2, 8], 150000, 1 << 18, 0.1),
    ([8, 8], (1 << 20) - 5, 1 << 20, 0.0181),
    ([8, 8, 4, 2, 1], 1 << 20, 1 << 21, 0.6),
])
def test_compact_bit_exact(vctx, widths, n, n_pad, p):
    print(f"compact widths={widths} n={n} n_pad={n_pad}", flush=True)
    rng = np.random.default_rng(n + len(widths))
    cols = [rng.integers(0, np.iinfo(WIDTH_DT[w]).max, n_pad, dtype=WIDTH_DT[w], endpoint=True) for w in widths]
    mask = (rng.random(n_pad) < p).astype(np.uint8)
    sel = mask.copy()
    sel[n:] = 0
    expect = oracle.compact(_rows(cols), sel)
    dcols = [to_dev(c) for c in cols]
    dmask = to_dev(mask)
    dcount = empty_dev(1, np.uint64)
    vctx.compact(dmask, dcols, widths, n, n_pad, out_count=dcount)
    got = _rows([from_dev(d, WIDTH_DT[w]) for d, w in zip(dcols, widths)])
    assert int(from_dev(dcount, np.uint64)[0]) == int(sel.sum())
    assert np.array_equal(got, expect)


@pytest.mark.parametrize("widths,n,n_pad,p", [
    ([8, 8], (1 << 20) - 5, 1 << 20, 0.0181),
    ([8, 8, 4, 2, 1], 1 << 20, 1 << 21, 0.6),
    ([8], 70000, 1 << 17, 0.5),
])
def test_compact_streaming_bit_exact(widths, n, n_pad, p):
    """The network passes with streaming (non-temporal) row accesses, which
    tuning 0x800 picks below the size that would: same rows and count as
    the oracle (every pass width: local, LDS-staged and register passes)."""
    import torch
    import velodb_amd
    print(f"compact streaming widths={widths} n={n} n_pad={n_pad}", flush=True)
    rng = np.random.default_rng(n + len(widths) + 11)
    cols = [rng.integers(0, np.iinfo(WIDTH_DT[w]).max, n_pad, dtype=WIDTH_DT[w], endpoint=True) for w in widths]
    mask = (rng.random(n_pad) < p).astype(np.uint8)
    sel = mask.copy()
    sel[n:] = 0
    expect = oracle.compact(_rows(cols), sel)
    dcols = [to_dev(c) for c in cols]
    dmask = to_dev(mask)
    dcount = empty_dev(1, np.uint64)
    with velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream, tuning=0x800) as ctx:
        ctx.compact(dmask, dcols, widths, n, n_pad, out_count=dcount)
    got = _rows([from_dev(d, WIDTH_DT[w]) for d, w in zip(dcols, widths)])
    assert int(from_dev(dcount, np.uint64)[0]) == int(sel.sum())
    assert np.array_equal(got, expect)


@pytest.mark.parametrize("n_pad,p", [(1 << 26, 0.3), (1 << 27, 0.02)])
def test_compact_many_tiles_scan(vctx, n_pad, p):
    """Above 8192 tiles (n_pad > 2^25) the tile counts are scanned in three
    passes (segment sums, their scan, segment prefixes) instead of one
    block: the count and the compacted column equal the mask-ordered
    selection (a stable compaction's closed form), with a live-row bound
    that cuts a segment."""
    print(f"compact many tiles n_pad={n_pad}", flush=True)
    rng = np.random.default_rng(n_pad + 3)
    n = n_pad - 12345
    vals = rng.integers(0, 2 ** 63, n_pad, dtype=np.uint64)
    mask = (rng.random(n_pad) < p).astype(np.uint8)
    sel = mask.astype(bool)
    sel[n:] = False
    dv, dm = to_dev(vals), to_dev(mask)
    dcount = empty_dev(1, np.uint64)
    vctx.compact(dm, [dv], [8], n, n_pad, out_count=dcount)
    m = int(sel.sum())
    assert int(from_dev(dcount, np.uint64)[0]) == m
    assert np.array_equal(from_dev(dv, np.uint64)[:m], vals[sel])


@pytest.mark.parametrize("mask_kind", ["all", "none", "alternate", "last"])
def test_compact_edge_masks(vctx, mask_kind):
    n_pad = 1 << 16
    vals = np.arange(n_pad, dtype=np.uint64) * 7 + 3
    mask = {"all": np.ones(n_pad, np.uint8), "none": np.zeros(n_pad, np.uint8),
            "alternate": (np.arange(n_pad) % 2).astype(np.uint8),
            "last": (np.arange(n_pad) == n_pad - 1).astype(np.uint8)}[mask_kind]
    expect = oracle.compact(vals.reshape(-1, 1), mask).reshape(-1)
    dv = to_dev(vals)
    vctx.compact(to_dev(mask), [dv], [8], n_pad, n_pad)
    got = from_dev(dv, np.uint64)
    assert np.array_equal(got, expect)
    m = int(mask.sum())
    assert np.array_equal(got[:m], vals[mask != 0])


def _filter_case(vctx, n, n_pad, seed, want_rows=True):
    day, qty, price, disc = oracle.gen_fact(seed, n)
    d = {k: to_dev(v, n_pad, fill_byte=0) for k, v in dict(day=day, qty=qty, price=price, disc=disc).items()}
    out = empty_dev(4, np.uint64)
    op = empty_dev(n_pad, np.int64) if want_rows else None
    od = empty_dev(n_pad, np.int64) if want_rows else None
    vctx.filter_sum(d["day"], d["qty"], d["price"], d["disc"], n, n_pad, out, out_price=op, out_disc=od)
    o = from_dev(out, np.uint64)
    s = int(o[2]) | (int(o[3]) << 64)
    if s >= 1 << 127:
        s -= 1 << 128
    return (day, qty, price, disc), int(o[0]), s, (from_dev(op, np.int64) if want_rows else None), \
        (from_dev(od, np.int64) if want_rows else None)


@pytest.mark.parametrize("n,n_pad", [(4096, 4096), (100000, 131072), (1 << 20, 1 << 20)])
def test_filter_sum_bit_exact(vctx, n, n_pad):
    print(f"filter n={n}", flush=True)
    (day, qty, price, disc), m, s, op, od = _filter_case(vctx, n, n_pad, 5)
    em, es, eop, eod = oracle.filter_sum(day, qty, price, disc, n, n_pad, want_rows=True)
    assert (m, s) == (em, es)
    assert np.array_equal(op, eop) and np.array_equal(od, eod)


@pytest.mark.parametrize("n,n_pad", [(100000, 131072), (1 << 20, 1 << 20)])
def test_filter_sum_streaming_bit_exact(n, n_pad):
    """Filter + compaction + int128 SUM with streaming loads and stores in
    every pass (tuning 0x800 picks them below the size that would): count,
    sum and rows equal the oracle's."""
    import torch
    import velodb_amd
    print(f"filter streaming n={n}", flush=True)
    with velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream, tuning=0x800) as ctx:
        (day, qty, price, disc), m, s, op, od = _filter_case(ctx, n, n_pad, 9)
    em, es, eop, eod = oracle.filter_sum(day, qty, price, disc, n, n_pad, want_rows=True)
    assert (m, s) == (em, es)
    assert np.array_equal(op, eop) and np.array_equal(od, eod)


def test_filter_sum_6M_rows_2p23(vctx):
    """6,000,000 fact rows padded to 2^23, Q6-style predicate,
    stable oblivious compaction, exact int128 SUM(price*disc)."""
    n, n_pad = 6_000_000, 1 << 23
    (day, qty, price, disc), m, s, op, od = _filter_case(vctx, n, n_pad, 1)
    em, es, eop, eod = oracle.filter_sum(day, qty, price, disc, n, n_pad, want_rows=True)
    assert (m, s) == (em, es)
    assert np.array_equal(op, eop) and np.array_equal(od, eod)
    sel = (day >= 8766) & (day < 9131) & (disc >= 5) & (disc <= 7) & (qty < 24)
    assert m == int(sel.sum()) and abs(m / n - 0.0181) < 0.001


def test_filter_sum_int128_range(vctx):
    """Products near 2^125 and a total beyond 2^64: the int128 path is exact."""
    n = 4096
    day = np.full(n, 8800, np.int32)
    qty = np.ones(n, np.int64)
    price = np.full(n, (1 << 62) + 12345, np.int64)
    disc = np.full(n, (1 << 62) - 1, np.int64)
    disc[::2] = -((1 << 62) - 7)
    dd = [to_dev(a) for a in (day, qty, price, disc)]
    out = empty_dev(4, np.uint64)
    # widen the disc range so every row passes the predicate
    vctx.filter_sum(*dd, n, n, out, disc_lo=-(1 << 63), disc_hi=(1 << 63) - 1)
    o = from_dev(out, np.uint64)
    s = int(o[2]) | (int(o[3]) << 64)
    if s >= 1 << 127:
        s -= 1 << 128
    assert int(o[0]) == n
    assert s == sum(int(a) * int(b) for a, b in zip(price, disc))


def test_scatter_positions_stable(vctx):
    import ctypes
    from velodb_amd import _abi
    rng = np.random.default_rng(3)
    n = 100_003
    mask = (rng.random(n) < 0.3).astype(np.uint8)
    dmask = to_dev(mask)
    dout = to_dev(np.full(n, -1, np.int32))
    dcnt = empty_dev(1, np.uint64)
    cmd = _abi.Command()
    cmd.opcode = _abi.OP_SCATTER
    cmd.args.scatter = _abi.ScatterArgs(dout.data_ptr(), dcnt.data_ptr(), dmask.data_ptr(), n)
    vctx.wait(vctx.submit(cmd))
    got = from_dev(dout, np.int32)
    pos = np.cumsum(mask) - mask
    assert np.array_equal(got[mask != 0], pos[mask != 0].astype(np.int32))
    assert np.all(got[mask == 0] == -1)
    assert int(from_dev(dcnt, np.uint64)[0]) == int(mask.sum())
    del ctypes


def test_filter_sum_3e8_rows_2p29(vctx):
    """3.0e8 fact rows padded to 2^29 (the roofline size), generated
    on the device by the generator the oracle pins (orc_gen_fact).  The
    selected count, the exact int128 SUM(price*disc) and the compacted
    (price, disc) prefix equal their closed forms computed from the same
    columns: the count and sum of a mask, and the masked rows in row order
    (a stable partition, the ORCompact invariant).  Every product is below
    1.05e9 and the total below 2^63, so int64 torch arithmetic is exact."""
    import torch
    n, n_pad = 300_000_000, 1 << 29
    print(f"filter_sum n={n} n_pad={n_pad}", flush=True)
    day = empty_dev(n_pad, np.int32)
    qty, price, disc = (empty_dev(n_pad, np.int64) for _ in range(3))
    vctx.gen_fact(11, day, qty, price, disc, n)
    out = empty_dev(4, np.uint64)
    op, od = empty_dev(n_pad, np.int64), empty_dev(n_pad, np.int64)
    vctx.filter_sum(day, qty, price, disc, n, n_pad, out, op, od)
    o = from_dev(out, np.uint64)
    d, q, p, c = day[:n], qty[:n], price[:n], disc[:n]
    sel = (d >= 8766) & (d < 9131) & (c >= 5) & (c <= 7) & (q < 24)
    m = int(sel.sum().item())
    s = int((p * c)[sel].sum().item())
    assert int(o[0]) == m and abs(m / n - 0.0181) < 0.0005
    assert int(o[2]) == s and int(o[3]) == 0
    assert torch.equal(op[:m], p[sel]) and torch.equal(od[:m], c[sel])
    del day, qty, price, disc, op, od, d, q, p, c, sel
    torch.cuda.empty_cache()
