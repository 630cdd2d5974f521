"""GPU parity of the oblivious PK/FK join + GROUP BY SUM against the oracle."""
import numpy as np
import pytest

import oracle
from tests.gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu


def _run(vctx, dk, dg, fk, fv, n_pad, max_groups=64, dense=False):
    d = [to_dev(x) for x in (dk, dg, fk, fv)]
    og = empty_dev(max_groups, np.int32)
    os_ = empty_dev(max_groups, np.int64)
    on = empty_dev(1, np.uint64)
    sk = empty_dev(n_pad, np.uint64)
    sv = empty_dev(n_pad, np.uint64)
    vctx.join_groupsum(d[0], d[1], len(dk), d[2], d[3], len(fk), n_pad, og, os_, on, max_groups, sk, sv,
                       dense_groups=dense)
    G = int(from_dev(on, np.uint64)[0])
    g = from_dev(og, np.int32)
    s = from_dev(os_, np.int64)
    assert np.all(g[G:] == -1) and np.all(s[G:] == 0)
    return g[:G], s[:G], from_dev(sk, np.uint64), from_dev(sv, np.uint64)


@pytest.mark.parametrize("dense", [False, True])
@pytest.mark.parametrize("n_dim,n_fact,n_pad", [(100, 400, 4096), (1500, 6000, 8192), (20000, 50000, 1 << 17),
                                                (150000, 600000, 1 << 20)])
def test_join_groupsum_bit_exact(vctx, n_dim, n_fact, n_pad, dense):
    print(f"join n_dim={n_dim} n_fact={n_fact} dense={dense}", flush=True)
    dk, dg, fk, fv = oracle.gen_join(4, n_dim, n_fact)
    g, s, sk, sv = _run(vctx, dk, dg, fk, fv, n_pad, dense=dense)
    eg, es, esk, esv = oracle.join_groupsum(dk, dg, fk, fv, n_pad, want_sorted=True)
    assert np.array_equal(sk, esk) and np.array_equal(sv, esv)
    assert np.array_equal(g, eg) and np.array_equal(s, es)


@pytest.mark.parametrize("dense", [False, True])
def test_join_1p5M_dim_6M_fact(vctx, dense):
    """1.5M dimension rows (unique keys 1..1.5M, 25 groups), 6M fact
    rows, union padded to 2^23; 25 exact int64 sums."""
    n_dim, n_fact, n_pad = 1_500_000, 6_000_000, 1 << 23
    dk, dg, fk, fv = oracle.gen_join(1, n_dim, n_fact)
    g, s, _, _ = _run(vctx, dk, dg, fk, fv, n_pad, max_groups=25 if dense else 64, dense=dense)
    # closed form: every fact row matches exactly one dimension row
    g_of_key = np.zeros(n_dim + 1, np.int64)
    g_of_key[dk] = dg
    want = np.bincount(g_of_key[fk], weights=None, minlength=25)
    sums = np.zeros(25, np.int64)
    np.add.at(sums, g_of_key[fk], fv)
    assert g.tolist() == list(range(25))
    assert s.tolist() == sums.tolist()
    assert want.sum() == n_fact


@pytest.mark.parametrize("dense", [False, True])
def test_join_unmatched_and_empty_groups(vctx, dense):
    dk = np.array([1, 2, 3, 10], np.int32)
    dg = np.array([0, 1, 0, 7], np.int32)
    fk = np.array([1, 4, 3, 3, 9, 2, 11, 11], np.int32)
    fv = np.array([10, 1000, 5, 7, 1000, 3, 99, 99], np.int64)
    g, s, _, _ = _run(vctx, dk, dg, fk, fv, 4096, max_groups=8, dense=dense)
    assert g.tolist() == [0, 1] and s.tolist() == [22, 3]


def test_join_dense_groups_all_64_present(vctx):
    """Every group id 0..63 present: the dense path fills all 64 output slots
    in ascending order, equal to the sort-based path."""
    rng = np.random.default_rng(9)
    n_dim, n_fact = 6400, 40000
    dk = (rng.permutation(n_dim) + 1).astype(np.int32)
    dg = (np.arange(n_dim) % 64).astype(np.int32)
    fk = dk[rng.integers(0, n_dim, n_fact)]
    fv = rng.integers(100, 10_000_000, n_fact).astype(np.int64)
    gd, sd, _, _ = _run(vctx, dk, dg, fk, fv, 1 << 16, max_groups=64, dense=True)
    gs, ss, _, _ = _run(vctx, dk, dg, fk, fv, 1 << 16, max_groups=64, dense=False)
    eg, es = oracle.join_groupsum(dk, dg, fk, fv, 1 << 16)
    assert gd.tolist() == list(range(64)) and np.array_equal(gd, eg) and np.array_equal(sd, es)
    assert np.array_equal(gs, eg) and np.array_equal(ss, es)


def test_join_single_group_spans_tiles(vctx):
    # one group: the segmented sum must carry across every tile
    n_dim, n_fact = 1000, 70000
    dk = np.arange(1, n_dim + 1, dtype=np.int32)
    dg = np.zeros(n_dim, np.int32)
    rng = np.random.default_rng(1)
    fk = rng.integers(1, n_dim + 1, n_fact).astype(np.int32)
    fv = rng.integers(100, 10_000_000, n_fact).astype(np.int64)
    g, s, _, _ = _run(vctx, dk, dg, fk, fv, 1 << 17)
    assert g.tolist() == [0] and s.tolist() == [int(fv.sum())]


# ------------------------------------------------ 12-byte union records
# dense_groups with max_groups <= 63 and no sorted-union outputs sorts the
# union as (u64 key, u32 payload) records; fact values up to INT64_MAX - 65.

_I64_MAX = np.iinfo(np.int64).max
_I64_MIN = np.iinfo(np.int64).min


def _run_compact(vctx, dk, dg, fk, fv, n_pad, max_groups):
    """Returns (groups, sums), or None when the library reports UINT64_MAX
    groups (a fact value the 12-byte union cannot hold)."""
    d = [to_dev(x) for x in (dk, dg, fk, fv)]
    og = empty_dev(max_groups, np.int32)
    os_ = empty_dev(max_groups, np.int64)
    on = empty_dev(1, np.uint64)
    vctx.join_groupsum(d[0], d[1], len(dk), d[2], d[3], len(fk), n_pad, og, os_, on, max_groups, dense_groups=True)
    G = int(from_dev(on, np.uint64)[0])
    if G == (1 << 64) - 1:
        return None
    g = from_dev(og, np.int32)
    s = from_dev(os_, np.int64)
    assert G <= max_groups and np.all(g[G:] == -1) and np.all(s[G:] == 0)
    return g[:G], s[:G]


@pytest.mark.parametrize("n_dim,n_fact,n_pad", [(100, 400, 4096), (1500, 6000, 8192), (20000, 50000, 1 << 17),
                                                (150000, 600000, 1 << 20)])
def test_join_compact_union_bit_exact(vctx, n_dim, n_fact, n_pad):
    print(f"join compact union n_dim={n_dim} n_fact={n_fact}", flush=True)
    dk, dg, fk, fv = oracle.gen_join(4, n_dim, n_fact)
    g, s = _run_compact(vctx, dk, dg, fk, fv, n_pad, 25)
    eg, es = oracle.join_groupsum(dk, dg, fk, fv, n_pad, max_groups=25)
    assert np.array_equal(g, eg) and np.array_equal(s, es)


def test_join_compact_union_1p5M_dim_6M_fact(vctx):
    """1.5M dimension rows, 6M fact rows, union padded to 2^23, 25 groups:
    the closed-form sums."""
    n_dim, n_fact, n_pad = 1_500_000, 6_000_000, 1 << 23
    dk, dg, fk, fv = oracle.gen_join(1, n_dim, n_fact)
    g, s = _run_compact(vctx, dk, dg, fk, fv, n_pad, 25)
    g_of_key = np.zeros(n_dim + 1, np.int64)
    g_of_key[dk] = dg
    sums = np.zeros(25, np.int64)
    np.add.at(sums, g_of_key[fk], fv)
    assert g.tolist() == list(range(25)) and s.tolist() == sums.tolist()


def test_join_compact_union_value_extremes_and_bounds(vctx):
    """INT64_MIN and INT64_MAX - 65 are exact; groups outside [0, max_groups)
    (negative, 63 and above) are not counted, as on the 16-byte path; a fact
    value of INT64_MAX - 64, matched or not, gives UINT64_MAX groups."""
    dk = np.array([1, 2, 3, 4, 5, 6, 7], np.int32)
    dg = np.array([0, 1, 2, 3, 70, -5, 63], np.int32)
    fk = np.array([1, 2, 3, 4, 1, 5, 6, 7, 9], np.int32)
    fv = np.array([_I64_MIN, _I64_MAX - 65, -1, 0, 5, 11, 12, 13, 77], np.int64)
    g, s = _run_compact(vctx, dk, dg, fk, fv, 4096, 25)
    assert g.tolist() == [0, 1, 2, 3]
    assert s.tolist() == [_I64_MIN + 5, _I64_MAX - 65, -1, 0]
    # the same rows through the 16-byte path (sorted outputs requested)
    gk, sk, _, _ = _run(vctx, dk, dg, fk, fv, 4096, max_groups=25, dense=True)
    assert np.array_equal(g, gk) and np.array_equal(s, sk)
    for bad_key in (1, 9):  # matched, unmatched
        fk2 = np.append(fk, np.int32(bad_key))
        fv2 = np.append(fv, np.int64(_I64_MAX - 64))
        assert _run_compact(vctx, dk, dg, fk2, fv2, 4096, 25) is None
    # the context is still usable and the next result is exact
    g, s = _run_compact(vctx, dk, dg, fk, fv, 4096, 25)
    assert s.tolist() == [_I64_MIN + 5, _I64_MAX - 65, -1, 0]


def test_join_compact_union_trace_value_independent(vctx):
    """Two inputs of equal sizes, different keys, groups and values: the
    same launches and address digests on the 12-byte path."""
    def run(seed):
        dk, dg, fk, fv = oracle.gen_join(seed, 3000, 12000)
        d = [to_dev(x) for x in (dk, dg, fk, fv)]
        og, os_, on = empty_dev(25, np.int32), empty_dev(25, np.int64), empty_dev(1, np.uint64)
        vctx.trace_begin()
        vctx.join_groupsum(d[0], d[1], 3000, d[2], d[3], 12000, 1 << 14, og, os_, on, 25, dense_groups=True)
        t = vctx.trace_end()
        return t, vctx.trace_launches(t.n_launches)

    (t1, l1), (t2, l2) = run(5), run(6)
    assert t1.n_launches == t2.n_launches > 0 and l1 == l2
    assert t1.launch_digest == t2.launch_digest and t1.address_digest == t2.address_digest
