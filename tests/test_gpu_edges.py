"""Edge cases of every operator: empty and ragged inputs, the largest
supported sort, sizes that must be rejected (a status code, no launch),
inputs whose every row is a pad or unmatched, and the pairs join, oblivious
gather and wire-format transfers at their smallest and degenerate sizes."""
import numpy as np
import pytest

import oracle
from tests.gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu


def _status(fn):
    from velodb_amd import VoError
    try:
        fn()
    except VoError as e:
        return e.status
    return 0


def test_sort_rejects_bad_sizes(vctx):
    from velodb_amd import _abi
    dk, dp = empty_dev(16, np.uint64), empty_dev(16, np.uint32)
    # n_pad not a power of two, n > n_pad, n_pad above 2^29: rejected before any launch
    assert _status(lambda: vctx.sort_kp(dk, dp, 10, 12)) != 0
    assert _status(lambda: vctx.sort_kp(dk, dp, 17, 16)) != 0
    assert _status(lambda: vctx.sort_kp(dk, dp, 5, 1 << 30)) != 0
    # the context stays usable
    k, p = oracle.gen_sort_input(1, 10)
    dk, dp = to_dev(k, 16), to_dev(p, 16)
    vctx.sort_kp(dk, dp, 10, 16)
    ek, ep = oracle.bitonic_sort_kp(k, p, 10, 16)
    assert np.array_equal(from_dev(dk, np.uint64), ek) and np.array_equal(from_dev(dp, np.uint32), ep)
    assert _abi.CFG_PERSISTENT  # module constants present


def test_sort_all_pad_rows(vctx):
    """n = 0: every row is a pad (key max, payload = row index)."""
    dk, dp = to_dev(np.arange(8192, dtype=np.uint64)), to_dev(np.zeros(8192, np.uint32))
    vctx.sort_kp(dk, dp, 0, 8192)
    assert np.all(from_dev(dk, np.uint64) == np.uint64(2 ** 64 - 1))
    assert np.array_equal(from_dev(dp, np.uint32), np.arange(8192, dtype=np.uint32))


def test_sort_largest_supported_size(vctx):
    """2^29 records (the largest n_pad: 8-byte columns at 32-bit buffer
    offsets), ragged n: sorted, payload a permutation of the live rows, pads
    last."""
    import torch
    n_pad = 1 << 29
    n = n_pad - 12345
    dk = empty_dev(n_pad, np.uint64)
    dp = empty_dev(n_pad, np.uint32)
    vctx.gen_sort_input(11, dk, dp, n)
    vctx.sort_kp(dk, dp, n, n_pad)
    torch.cuda.synchronize()
    sign = -(2 ** 63)
    k = dk ^ sign
    assert bool(torch.all(k[1:] >= k[:-1]))
    assert bool(torch.all(dk[n:] == -1))
    live = dp[:n].to(torch.int64) & 0xFFFFFFFF
    assert int(live.max()) == n - 1 and int(live.min()) == 0
    assert int(torch.unique(live).numel()) == n
    del dk, dp, k, live
    torch.cuda.empty_cache()


def test_compact_no_live_rows(vctx):
    n_pad = 4096
    mask = to_dev(np.ones(n_pad, np.uint8))
    col = to_dev(np.arange(n_pad, dtype=np.uint64))
    cnt = empty_dev(1, np.uint64)
    vctx.compact(mask, [col], [8], 0, n_pad, cnt)  # n = 0: rows >= n count as unselected
    assert int(from_dev(cnt, np.uint64)[0]) == 0
    assert sorted(from_dev(col, np.uint64).tolist()) == list(range(n_pad))


def test_compact_rejects_small_and_odd_sizes(vctx):
    mask = to_dev(np.ones(4096, np.uint8))
    col = to_dev(np.arange(4096, dtype=np.uint64))
    assert _status(lambda: vctx.compact(mask, [col], [8], 100, 2048)) != 0   # below one 4096-row tile
    assert _status(lambda: vctx.compact(mask, [col], [8], 100, 6000)) != 0   # not a power of two
    assert _status(lambda: vctx.compact(mask, [col], [3], 100, 4096)) != 0   # 3-byte column


def test_filter_sum_empty_and_nothing_selected(vctx):
    n_pad = 4096
    day, qty, price, disc = (to_dev(x, n_pad) for x in oracle.gen_fact(3, 100))
    out = empty_dev(4, np.uint64)
    vctx.filter_sum(day, qty, price, disc, 0, n_pad, out)  # no rows
    assert from_dev(out, np.uint64).tolist() == [0, 0, 0, 0]
    # a predicate no row meets
    day, qty, price, disc = (to_dev(x, n_pad) for x in oracle.gen_fact(3, 3000))
    vctx.filter_sum(day, qty, price, disc, 3000, n_pad, out, day_lo=0, day_hi=0)
    assert from_dev(out, np.uint64).tolist() == [0, 0, 0, 0]


@pytest.mark.parametrize("dense", [False, True])
def test_join_without_matches(vctx, dense):
    """Every fact key misses the dimension table: no groups, zero sums."""
    dk = np.arange(1, 101, dtype=np.int32)
    dg = np.zeros(100, np.int32)
    fk = np.arange(1000, 1500, dtype=np.int32)
    fv = np.ones(500, np.int64)
    d = [to_dev(x) for x in (dk, dg, fk, fv)]
    og, os_, on = empty_dev(8, np.int32), empty_dev(8, np.int64), empty_dev(1, np.uint64)
    vctx.join_groupsum(d[0], d[1], 100, d[2], d[3], 500, 4096, og, os_, on, 8, dense_groups=dense)
    assert int(from_dev(on, np.uint64)[0]) == 0
    assert np.all(from_dev(og, np.int32) == -1) and np.all(from_dev(os_, np.int64) == 0)


def test_join_rejects_bad_sizes(vctx):
    d = [to_dev(np.ones(10, t)) for t in (np.int32, np.int32, np.int32, np.int64)]
    og, os_, on = empty_dev(8, np.int32), empty_dev(8, np.int64), empty_dev(1, np.uint64)
    assert _status(lambda: vctx.join_groupsum(d[0], d[1], 10, d[2], d[3], 10, 2048, og, os_, on, 8)) != 0
    assert _status(lambda: vctx.join_groupsum(d[0], d[1], 4000, d[2], d[3], 4000, 4096, og, os_, on, 8)) != 0


@pytest.mark.parametrize("n_pad", [1, 2, 4])
def test_shuffle_tiny(vctx, n_pad):
    pay = to_dev(np.arange(n_pad, dtype=np.uint32) + 7)
    perm = empty_dev(n_pad, np.uint32)
    vctx.shuffle(pay, n_pad, 3, perm)
    got = from_dev(pay, np.uint32)
    pm = from_dev(perm, np.uint32)
    assert sorted(got.tolist()) == list(range(7, 7 + n_pad))
    assert np.array_equal(got, pm + 7)


def test_sort_batch_empty_and_degenerate(vctx):
    dk, dp = empty_dev(64, np.uint64), empty_dev(64, np.uint32)
    vctx.sort_batch(dk, dp, [])  # nothing to do
    k, p = oracle.gen_sort_input(4, 3)
    dk, dp = to_dev(k, 4), to_dev(p, 4)
    vctx.sort_batch(dk, dp, [(0, 3, 4)])
    ek, ep = oracle.bitonic_sort_kp(k, p, 3, 4)
    assert np.array_equal(from_dev(dk, np.uint64), ek) and np.array_equal(from_dev(dp, np.uint32), ep)
    # an item whose offset is not a multiple of its padded size is rejected
    from velodb_amd import VoError
    with pytest.raises(VoError):
        vctx.sort_batch(dk, dp, [(2, 3, 4)])


def test_join_pairs_edges(vctx):
    """Pairs join: WRITE without a COUNT, the reference's write arguments
    (type 0, capacity 0), one row per side, a single key on both sides."""
    import velodb_amd
    from velodb_amd import VoError, _abi
    ctx2 = velodb_amd.Context()
    try:
        d = to_dev(np.array([1], np.int32))
        o = empty_dev(4, np.int64)
        with pytest.raises(VoError):  # no COUNT on this context yet
            ctx2.join_pairs_write(d, 1, d, 1, _abi.T_INTEGER, o, o, 4)
    finally:
        ctx2.close()
    one = to_dev(np.array([5], np.int64))
    assert vctx.join_pairs_count(one, 1, one, 1, _abi.T_BIGINT) == 1
    ol, orr = empty_dev(1, np.int64), empty_dev(1, np.int64)
    vctx.join_pairs_write(one, 1, one, 1, 0, ol, orr, 0)  # type and capacity of the COUNT
    assert from_dev(ol, np.int64).tolist() == [0] and from_dev(orr, np.int64).tolist() == [0]
    # 3 x 4 rows of one key: 12 pairs in (left, right) order
    lk, rk = to_dev(np.full(3, -9, np.int16)), to_dev(np.full(4, -9, np.int16))
    assert vctx.join_pairs_count(lk, 3, rk, 4, _abi.T_SMALLINT) == 12
    ol, orr = empty_dev(16, np.int64), empty_dev(16, np.int64)
    vctx.join_pairs_write(lk, 3, rk, 4, _abi.T_SMALLINT, ol, orr, 16)
    assert from_dev(ol, np.int64)[:12].tolist() == [a for a in range(3) for _ in range(4)]
    assert from_dev(orr, np.int64)[:12].tolist() == [b for _ in range(3) for b in range(4)]


def test_gather_oblivious_edges(vctx):
    """Oblivious gather: no requests, an empty table, a one-row table, and a
    bad element width."""
    from velodb_amd import VoError
    t = to_dev(np.array([77], np.uint64))
    ids = to_dev(np.array([0, 0, 1, -1], np.int64))
    out = empty_dev(4, np.uint64)
    found = empty_dev(4, np.uint8)
    vctx.gather_oblivious(t, 1, ids, 0, out, 8)  # nothing to do
    vctx.gather_oblivious(t, 1, ids, 4, out, 8, out_found=found)
    assert from_dev(out, np.uint64).tolist() == [77, 77, 0, 0]
    assert from_dev(found, np.uint8).tolist() == [1, 1, 0, 0]
    vctx.gather_oblivious(t, 0, ids, 4, out, 8, out_found=found)  # empty table: nothing found
    assert from_dev(out, np.uint64).tolist() == [0, 0, 0, 0]
    assert from_dev(found, np.uint8).tolist() == [0, 0, 0, 0]
    with pytest.raises(VoError):
        vctx.gather_oblivious(t, 1, ids, 4, out, 3)


def test_transfer_edges(vctx):
    """Wire format: zero bytes, a padded size that is not a multiple of the
    staging chunk, an odd byte count."""
    import torch
    dev = torch.full((1000,), 0x55, dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()
    assert vctx.transfer_to_device(dev, np.zeros(0, np.uint8), padded_bytes=0) == 0
    src = (np.arange(777) % 251).astype(np.uint8)
    assert vctx.transfer_to_device(dev, src, padded_bytes=1000) == 1000
    got = dev.cpu().numpy()
    assert np.array_equal(got[:777], src) and np.all(got[777:] == 0)
    back = np.empty(333, np.uint8)
    assert vctx.transfer_to_host(back, dev, padded_bytes=1000) == 1000
    assert np.array_equal(back, src[:333])
