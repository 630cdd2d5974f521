"""Oblivious equi-join into row-index pairs (OP_SORT_MERGE_JOIN_COUNT /
_WRITE, the SortMergeJoinOperator path): bit-exact against the oracle and
the reference's golden join, at full size through properties, edge cases,
and trace identity."""
import os

import numpy as np
import pytest

import oracle
from tests.gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
T_INTEGER, T_BIGINT = 4, 5
_DT = {1: np.uint8, 2: np.int8, 3: np.int16, 4: np.int32, 5: np.int64, 10: np.uint32, 64: np.uint64}


def _read_join_file(path):
    lines = open(path).read().split("\n")
    n1, n2 = map(int, lines[0].split())
    rows = [ln.split() for ln in lines[1:] if ln.strip()]
    return rows[:n1], rows[n1:n1 + n2]


def _run(vctx, lk, rk, t, cap_extra=0):
    dl = to_dev(np.asarray(lk, _DT[t])) if len(lk) else empty_dev(1, np.int64)
    dr = to_dev(np.asarray(rk, _DT[t])) if len(rk) else empty_dev(1, np.int64)
    n = vctx.join_pairs_count(dl, len(lk), dr, len(rk), t)
    cap = 1
    while cap < n:
        cap <<= 1
    cap += cap_extra
    ol, orr = empty_dev(cap, np.int64), empty_dev(cap, np.int64)
    vctx.join_pairs_write(dl, len(lk), dr, len(rk), t, ol, orr, cap)
    a, b = from_dev(ol, np.int64), from_dev(orr, np.int64)
    assert np.all(a[n:cap] == -1) and np.all(b[n:cap] == -1)
    return n, a[:n], b[:n]


def test_reference_golden_join(vctx):
    """obliviator/join/test.txt -> test_output.txt."""
    print("join_pairs golden", flush=True)
    left, right = _read_join_file(os.path.join(GOLDEN, "obliviator_join_input.txt"))
    expected = sorted(tuple(ln.split()) for ln in open(os.path.join(GOLDEN, "obliviator_join_expected.txt"))
                      if ln.strip())
    lk = [int(r[0]) for r in left]
    rk = [int(r[0]) for r in right]
    n, a, b = _run(vctx, lk, rk, T_INTEGER)
    got = sorted((left[i][0], left[i][1], right[j][0], right[j][1]) for i, j in zip(a, b))
    assert n == 512 and got == expected
    wl, wr = oracle.join_pairs(oracle.key_codes(lk, T_INTEGER), oracle.key_codes(rk, T_INTEGER))
    assert np.array_equal(a, wl) and np.array_equal(b, wr)


def test_fk_join_fixture(vctx):
    left, right = _read_join_file(os.path.join(GOLDEN, "obliviator_fk_join_input.txt"))
    lk = [int(r[0]) for r in left]
    rk = [int(r[0]) for r in right]
    n, a, b = _run(vctx, lk, rk, T_INTEGER)
    wl, wr = oracle.join_pairs(oracle.key_codes(lk, T_INTEGER), oracle.key_codes(rk, T_INTEGER))
    assert n == 1024 and np.array_equal(a, wl) and np.array_equal(b, wr)


@pytest.mark.parametrize("t", [1, 2, 3, 4, 5, 10, 64])
@pytest.mark.parametrize("shape", ["dense_dups", "pk_fk", "sparse", "one_key"])
def test_parity_with_oracle(vctx, t, shape):
    print(f"join_pairs parity type={t} shape={shape}", flush=True)
    rng = np.random.default_rng(t * 31 + len(shape))
    info = np.iinfo(_DT[t])
    if shape == "dense_dups":     # m:n blocks, many keys on both sides
        span = min(60, int(info.max) - int(info.min))
        lk = rng.integers(0, span, 3000) + info.min
        rk = rng.integers(0, span, 2000) + info.min
    elif shape == "pk_fk":        # unique left keys, right keys drawn from them (and some misses)
        hi = min(int(info.max), 20000)
        lo = max(int(info.min), hi - 3000)
        lk = rng.permutation(np.arange(lo, hi))[:2500]
        rk = rng.integers(max(int(info.min), lo - 500), hi, 7000)
    elif shape == "sparse":       # full-range values, few matches
        lk = rng.integers(info.min, info.max, 4000, dtype=_DT[t], endpoint=True)
        rk = np.concatenate([rng.integers(info.min, info.max, 3000, dtype=_DT[t], endpoint=True), lk[:50]])
        rk[:2] = [info.max, info.min]
        lk[:2] = [info.max, info.min]
    else:                         # one key on both sides: n_left * n_right pairs
        lk = np.full(70, info.max)
        rk = np.full(130, info.max)
    lk, rk = np.asarray(lk, _DT[t]), np.asarray(rk, _DT[t])
    n, a, b = _run(vctx, lk, rk, t)
    wl, wr = oracle.join_pairs(oracle.key_codes(lk, t), oracle.key_codes(rk, t))
    assert n == len(wl)
    assert np.array_equal(a, wl) and np.array_equal(b, wr)


def _check_pairs_torch(lk, rk, n, ol, orr):
    """Size-independent checks on device: every pair joins equal keys, the
    pairs are strictly increasing in (key, left, right) (so distinct), and n
    equals the number of equal-key pairs from the key histograms: together
    the output is exactly the join."""
    import torch
    l, r = ol[:n], orr[:n]
    assert int(l.min()) >= 0 and int(r.min()) >= 0
    assert int(l.max()) < lk.numel() and int(r.max()) < rk.numel()
    kl, kr = lk[l], rk[r]
    assert bool(torch.all(kl == kr))
    same = kl[1:] == kl[:-1]
    inc = (kl[1:] > kl[:-1]) | (same & (l[1:] > l[:-1])) | (same & (l[1:] == l[:-1]) & (r[1:] > r[:-1]))
    assert bool(torch.all(inc))
    ul, cl = torch.unique(lk, return_counts=True)
    ur, cr = torch.unique(rk, return_counts=True)
    common, il, ir = np.intersect1d(ul.cpu().numpy(), ur.cpu().numpy(), return_indices=True)
    want = int((cl.cpu().numpy()[il].astype(np.int64) * cr.cpu().numpy()[ir].astype(np.int64)).sum())
    assert n == want
    assert bool(torch.all(ol[n:] == -1)) and bool(torch.all(orr[n:] == -1))


def test_full_size_pk_fk(vctx):
    """1.5M unique dimension keys x 6M foreign keys (the sizes of the join
    workload): 6M pairs."""
    import torch
    print("join_pairs full-size pk/fk", flush=True)
    g = torch.Generator(device="cuda:0").manual_seed(7)
    lk = (torch.randperm(1_500_000, device="cuda:0", generator=g) + 1).to(torch.int32)
    rk = torch.randint(1, 1_500_001, (6_000_000,), device="cuda:0", generator=g, dtype=torch.int32)
    n, ol, orr = vctx.join_pairs(lk, rk, T_INTEGER)
    assert n == 6_000_000
    _check_pairs_torch(lk.to(torch.int64), rk.to(torch.int64), n, ol, orr)
    # every fact row appears once
    assert int(torch.unique(orr[:n]).numel()) == n


def test_full_size_many_to_many(vctx):
    """2M x 40K rows over 20000 keys (about 4M pairs) with one heavy key of
    500 right rows."""
    import torch
    print("join_pairs full-size m:n", flush=True)
    g = torch.Generator(device="cuda:0").manual_seed(11)
    lk = torch.randint(0, 20000, (2_000_000,), device="cuda:0", generator=g, dtype=torch.int64)
    rk = torch.randint(0, 20000, (40_000,), device="cuda:0", generator=g, dtype=torch.int64)
    rk[:500] = 123  # a heavy key
    n, ol, orr = vctx.join_pairs(lk, rk, T_BIGINT)
    assert n > 3_000_000
    _check_pairs_torch(lk, rk, n, ol, orr)


def test_edges(vctx):
    from velodb_amd import VoError
    print("join_pairs edges", flush=True)
    # empty sides and no matches: n = 0, the output is all -1
    for lk, rk in (([], [1, 2, 3]), ([1, 2], []), ([1, 2, 3], [4, 5]), ([], [])):
        n, a, b = _run(vctx, lk, rk, T_INTEGER, cap_extra=5)
        assert n == 0
    # out_cap beyond next_pow2(n): the tail is -1 (checked in _run)
    n, a, b = _run(vctx, [5, 5, 6], [5, 6, 6, 7], T_INTEGER, cap_extra=3000)
    assert n == 4 and a.tolist() == [0, 1, 2, 2] and b.tolist() == [0, 0, 1, 2]
    dl, dr = to_dev(np.array([1, 2], np.int32)), to_dev(np.array([2, 2], np.int32))
    ol, orr = empty_dev(4, np.int64), empty_dev(4, np.int64)
    # WRITE for other columns than the last COUNT, or too small an output
    assert vctx.join_pairs_count(dl, 2, dr, 2, T_INTEGER) == 2
    with pytest.raises(VoError):
        vctx.join_pairs_write(dr, 2, dl, 2, T_INTEGER, ol, orr, 4)
    with pytest.raises(VoError):
        vctx.join_pairs_write(dl, 2, dr, 2, T_INTEGER, ol, orr, 1)
    vctx.join_pairs_write(dl, 2, dr, 2, T_INTEGER, ol, orr, 4)
    assert from_dev(ol, np.int64).tolist() == [1, 1, -1, -1] and from_dev(orr, np.int64).tolist() == [0, 1, -1, -1]
    # floating-point keys are not supported (equality of -0.0 / NaN)
    with pytest.raises(VoError) as e:
        vctx.join_pairs_count(dl, 2, dr, 2, 7)
    assert e.value.status == 5


def test_through_command_queue(vctx):
    """OP_SORT_MERGE_JOIN_COUNT / _WRITE submitted as commands (the
    TaskManager path of sort_merge_join_operator.cpp:113-149)."""
    from velodb_amd import _abi
    lk = np.array([3, 1, 2, 3], np.int32)
    rk = np.array([3, 3, 2, 9], np.int32)
    dl, dr = to_dev(lk), to_dev(rk)
    cnt = empty_dev(1, np.uint64)
    cmd = _abi.Command()
    cmd.opcode = _abi.OP_SORT_MERGE_JOIN_COUNT
    cmd.args.join_pairs = _abi.JoinPairsArgs(dl.data_ptr(), 4, dr.data_ptr(), 4, T_INTEGER, 0, cnt.data_ptr(),
                                             None, None, 0)
    vctx.wait(vctx.submit(cmd))
    n = int(from_dev(cnt, np.uint64)[0])
    assert n == 5
    ol, orr = empty_dev(8, np.int64), empty_dev(8, np.int64)
    cmd = _abi.Command()
    cmd.opcode = _abi.OP_SORT_MERGE_JOIN_WRITE
    # as translated from the reference's SortMergeJoinWriteArgs: no type, no
    # capacity (0: the power of two >= n_out the reference's operator allocates)
    cmd.args.join_pairs = _abi.JoinPairsArgs(dl.data_ptr(), 4, dr.data_ptr(), 4, 0, 0, None,
                                             ol.data_ptr(), orr.data_ptr(), 0)
    vctx.wait(vctx.submit(cmd))
    assert from_dev(ol, np.int64).tolist() == [2, 0, 0, 3, 3, -1, -1, -1]
    assert from_dev(orr, np.int64).tolist() == [2, 0, 1, 0, 1, -1, -1, -1]


@pytest.mark.parametrize("ty", [T_BIGINT, T_INTEGER])  # 16-byte and 12-byte union records
def test_trace_value_independent(vctx, ty):
    """Two inputs of equal sizes and equal pair count (the one public output)
    but different key structure: identical launches and address digests."""
    dt = np.int64 if ty == T_BIGINT else np.int32

    def run(lk, rk):
        dl, dr = to_dev(np.asarray(lk, dt)), to_dev(np.asarray(rk, dt))
        vctx.trace_begin()
        n = vctx.join_pairs_count(dl, len(lk), dr, len(rk), ty)
        ol, orr = empty_dev(16384, np.int64), empty_dev(16384, np.int64)
        vctx.join_pairs_write(dl, len(lk), dr, len(rk), ty, ol, orr, 16384)
        t = vctx.trace_end()
        return n, t, vctx.trace_launches(t.n_launches)

    rng = np.random.default_rng(2)
    # 9000 pairs both ways: one 90 x 100 block, or 9000 one-to-one keys
    n1, t1, l1 = run(np.full(3000, 7), np.concatenate([np.full(3, 7), rng.integers(100, 200, 9997)]))
    lk2 = rng.permutation(np.arange(3000))
    rk2 = np.concatenate([np.repeat(lk2[:3000], 3), rng.integers(5000, 6000, 1000)])
    n2, t2, l2 = run(lk2, rk2)
    assert n1 == n2 == 9000
    assert t1.n_launches == t2.n_launches > 0 and l1 == l2
    assert t1.launch_digest == t2.launch_digest and t1.address_digest == t2.address_digest
