"""Oblivious shuffle (Philox4x32-10 tags + bitonic sort): exact permutation
against the oracle's tags, permutation validity, movement, a uniformity
chi-square test and downstream-result equality."""
import numpy as np
import pytest

import oracle
from tests.gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu


def _shuffle(vctx, payload, seed):
    n_pad = len(payload)
    dp = to_dev(payload)
    dperm = empty_dev(n_pad, np.uint32)
    vctx.shuffle(dp, n_pad, seed, perm_out=dperm)
    return from_dev(dp, np.uint32), from_dev(dperm, np.uint32)


@pytest.mark.parametrize("n_pad,seed", [(8, 1), (1024, 2), (1 << 16, 3), (1 << 20, 12345678901)])
def test_shuffle_permutation_matches_oracle_tags(vctx, n_pad, seed):
    payload = (np.arange(n_pad, dtype=np.uint64) * 2654435761 % (1 << 32)).astype(np.uint32)
    out, perm = _shuffle(vctx, payload, seed)
    tags = np.array([oracle.philox_tag(seed, i) for i in range(n_pad)], np.uint64) if n_pad <= 1 << 16 else None
    if tags is None:
        # large case: spot-check tags through the order they induce
        sample = np.arange(0, n_pad, 997)
        tags_s = np.array([oracle.philox_tag(seed, int(i)) for i in perm[sample]], np.uint64)
        assert np.all(np.diff(tags_s.astype(object)) >= 0)
    else:
        expect = np.lexsort((np.arange(n_pad), tags))
        assert np.array_equal(perm, expect.astype(np.uint32))
    assert np.array_equal(np.sort(perm), np.arange(n_pad, dtype=np.uint32))
    assert np.array_equal(out, payload[perm])


def test_shuffle_moves_most_rows(vctx):
    # reference tests/data/test_oblivious_table.cpp:61-95: > 90% of positions change
    n = 1024
    _, perm = _shuffle(vctx, np.arange(n, dtype=np.uint32), 99)
    assert (perm != np.arange(n)).mean() > 0.9


def test_shuffle_uniformity_chi_square(vctx):
    from scipy import stats
    n, trials = 8, 4000
    table = np.zeros((n, n), np.int64)
    for s in range(1, trials + 1):
        out, _ = _shuffle(vctx, np.arange(n, dtype=np.uint32), s)
        table[out, np.arange(n)] += 1  # item out[j] landed at position j
    chi2 = ((table - trials / n) ** 2 / (trials / n)).sum()
    p = stats.chi2.sf(chi2, (n - 1) * (n - 1))
    assert p > 1e-4, (chi2, p)


def test_shuffle_downstream_results_unchanged(vctx):
    """Sorting shuffled rows gives the same result as sorting the originals."""
    n = 1 << 15
    k, _ = oracle.gen_sort_input(5, n)
    payload = np.arange(n, dtype=np.uint32)
    shuffled, perm = _shuffle(vctx, payload, 7)
    dk = to_dev(k[shuffled])
    dp = to_dev(shuffled)
    vctx.sort_kp(dk, dp, n, n)
    ek, ep = oracle.bitonic_sort_kp(k, payload, n, n)
    assert np.array_equal(from_dev(dk, np.uint64), ek) and np.array_equal(from_dev(dp, np.uint32), ep)


def test_shuffle_is_seed_deterministic_and_seed_sensitive(vctx):
    p = np.arange(4096, dtype=np.uint32)
    a, _ = _shuffle(vctx, p, 11)
    b, _ = _shuffle(vctx, p, 11)
    c, _ = _shuffle(vctx, p, 12)
    assert np.array_equal(a, b) and not np.array_equal(a, c)


@pytest.mark.parametrize("n,n_pad", [(1000, 1024), (3, 4), (5000, 8192), (1, 1)])
def test_shuffle_live_rows_only(vctx, n, n_pad):
    """With n < n_pad live rows (a position map of n records), rows >= n stay
    in place and the first n rows are the live rows in the order of their
    sorted Philox tags."""
    pay = np.arange(n_pad, dtype=np.uint32) * 3 + 1
    d = to_dev(pay)
    perm = empty_dev(n_pad, np.uint32)
    vctx.shuffle(d, n_pad, 21, perm, n=n)
    got = from_dev(d, np.uint32)
    pm = from_dev(perm, np.uint32)
    tags = [oracle.philox_tag(21, i) for i in range(n)]
    want = sorted(range(n), key=lambda i: (tags[i], i))
    assert pm[:n].tolist() == want
    assert pm[n:].tolist() == list(range(n, n_pad))
    assert np.array_equal(got, pay[pm])


@pytest.mark.parametrize("n,n_pad,seed", [(8, 8, 1), (1024, 1024, 2), (1000, 1024, 21), (5000, 8192, 3),
                                          (1 << 16, 1 << 16, 4), (3, 4, 5), (1, 1, 6)])
def test_payload_only_shuffle_is_the_same_permutation(vctx, n, n_pad, seed):
    """Without perm_out the payload is shuffled as 12-byte (tag >> 1,
    payload) records: the same rows land in the same places as with the
    16-byte (tag, row) form whenever no two live tags share their top 63
    bits (checked here), and rows >= n stay in place."""
    tags = [oracle.philox_tag(seed, i) for i in range(n)]
    assert len({t >> 1 for t in tags}) == n  # no 63-bit tie in this case
    pay = (np.arange(n_pad, dtype=np.uint64) * 2654435761 % (1 << 32)).astype(np.uint32)
    a = to_dev(pay)
    vctx.shuffle(a, n_pad, seed, n=n)
    b = to_dev(pay)
    perm = empty_dev(n_pad, np.uint32)
    vctx.shuffle(b, n_pad, seed, perm, n=n)
    got, want = from_dev(a, np.uint32), from_dev(b, np.uint32)
    assert np.array_equal(got, want)
    order = sorted(range(n), key=lambda i: (tags[i], i))
    assert np.array_equal(got[:n], pay[order]) and np.array_equal(got[n:], pay[n:])


def test_payload_only_shuffle_trace_value_independent(vctx):
    n = 1 << 14

    def run(vals):
        d = to_dev(vals)
        vctx.trace_begin()
        vctx.shuffle(d, n, 99)
        t = vctx.trace_end()
        return t, vctx.trace_launches(t.n_launches)

    (t1, l1), (t2, l2) = run(np.zeros(n, np.uint32)), run(np.random.default_rng(1).integers(0, 2**32, n, np.uint32))
    assert t1.n_launches == t2.n_launches > 0 and l1 == l2
    assert t1.launch_digest == t2.launch_digest and t1.address_digest == t2.address_digest


@pytest.mark.parametrize("tuning", [0x3 | (8 << 4), 0x5 | (8 << 4), 0x2 | (6 << 4), 0x800])
def test_kv_sort_schedules_give_the_same_shuffle(vctx, tuning):
    """The 16-byte sort behind perm_out with register passes capped at 3
    (groups of 4-7 stages as LDS-staged passes), 5, or 2 (hybrid up to 5),
    or with streaming loads and stores (0x800): every schedule applies the
    same network, so payload and permutation equal the default context's at
    2^20 and 2^17."""
    import torch
    import velodb_amd
    print(f"kv schedule tuning={tuning:#x}", flush=True)
    with velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream, tuning=tuning) as ctx:
        for n_pad, seed in [(1 << 20, 99), (1 << 17, 5)]:
            payload = np.random.default_rng(seed).integers(0, 2 ** 32, n_pad, dtype=np.uint64).astype(np.uint32)
            want_p, want_perm = _shuffle(vctx, payload, seed)
            got_p, got_perm = _shuffle(ctx, payload, seed)
            assert np.array_equal(got_p, want_p) and np.array_equal(got_perm, want_perm)
