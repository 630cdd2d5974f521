"""Obliviousness of the compaction, filter + SUM, join + GROUP BY, multi-column
sort and shuffle paths: the launch sequence, grid shapes and per-launch address digests are
identical for two inputs of equal padded size and different values (the
reference's check, obliviator test_scripts/test.py:20-51 over trace_mem.h).
A positive control shows the digests see data-dependent addressing: the
reference-compatible OP_SCATTER writes only selected rows, and its trace
changes with the mask."""
import numpy as np
import pytest

import oracle
from tests.gpu_util import empty_dev, to_dev

pytestmark = pytest.mark.gpu


def _traced(vctx, fn):
    vctx.trace_begin()
    fn()
    t = vctx.trace_end()
    return t, vctx.trace_launches(t.n_launches)


def _same(a, b):
    (t1, l1), (t2, l2) = a, b
    assert t1.n_launches == t2.n_launches > 0
    assert l1 == l2
    assert t1.launch_digest == t2.launch_digest
    assert t1.address_digest == t2.address_digest
    # every launch folded its addresses
    assert all(d != 0 for (_, _, _, d) in l1), [x for x in l1 if x[3] == 0]


@pytest.mark.parametrize("n_pad,widths", [(1 << 16, (8, 8)), (1 << 17, (8, 8)), (1 << 20, (8, 8)), (1 << 18, (4, 1))])
def test_compaction_trace_value_independent(vctx, n_pad, widths):
    print(f"compaction trace n_pad={n_pad} widths={widths}", flush=True)
    rng = np.random.default_rng(n_pad)
    n = n_pad - 1000
    dt = {8: np.uint64, 4: np.uint32, 1: np.uint8}

    def run(p):
        mask = to_dev((rng.random(n) < p).astype(np.uint8), n_pad)
        cols = [to_dev(rng.integers(0, 200, n_pad).astype(dt[w])) for w in widths]
        cnt = empty_dev(1, np.uint64)
        return _traced(vctx, lambda: vctx.compact(mask, cols, list(widths), n, n_pad, cnt))

    _same(run(0.02), run(0.9))


def test_filter_sum_trace_value_independent(vctx):
    n, n_pad = 100000, 1 << 17

    def run(seed):
        day, qty, price, disc = (to_dev(x, n_pad) for x in oracle.gen_fact(seed, n))
        out = empty_dev(4, np.uint64)
        op, od = empty_dev(n_pad, np.int64), empty_dev(n_pad, np.int64)
        return _traced(vctx, lambda: vctx.filter_sum(day, qty, price, disc, n, n_pad, out, out_price=op,
                                                     out_disc=od))

    _same(run(1), run(2))


@pytest.mark.parametrize("dense", [False, True])
def test_join_trace_value_independent(vctx, dense):
    n_dim, n_fact, n_pad = 3000, 12000, 1 << 14

    def run(seed):
        dk, dg, fk, fv = oracle.gen_join(seed, n_dim, n_fact)
        if seed == 2:  # a very different shape of data: half the fact rows unmatched, one group
            fk = np.where(np.arange(n_fact) % 2 == 0, fk, n_dim + 100 + np.arange(n_fact)).astype(np.int32)
            dg = np.zeros_like(dg)
        d = [to_dev(x) for x in (dk, dg, fk, fv)]
        og, os_, on = empty_dev(32, np.int32), empty_dev(32, np.int64), empty_dev(1, np.uint64)
        return _traced(vctx, lambda: vctx.join_groupsum(d[0], d[1], n_dim, d[2], d[3], n_fact, n_pad, og, os_, on,
                                                        32, dense_groups=dense))

    _same(run(1), run(2))


def test_multi_column_sort_trace_value_independent(vctx):
    """OP_SORT (SortOperator path): INTEGER ASC + DOUBLE DESC keys."""
    from velodb_amd import _abi
    n, n_pad = 5000, 8192

    def run(seed):
        rng = np.random.default_rng(seed)
        a = to_dev(rng.integers(-5, 5, n).astype(np.int32), n_pad)
        b = to_dev(rng.normal(size=n), n_pad)
        idx = empty_dev(n_pad, np.int64)
        return _traced(vctx, lambda: vctx.sort([a, b], [_abi.T_INTEGER, _abi.T_DOUBLE], [True, False], idx, n, n_pad))

    _same(run(1), run(2))


def test_shuffle_trace_value_independent(vctx):
    n_pad = 1 << 16

    def run(seed):
        pay = to_dev(np.random.default_rng(seed).integers(0, 1 << 32, n_pad, dtype=np.uint64).astype(np.uint32))
        return _traced(vctx, lambda: vctx.shuffle(pay, n_pad, 77))

    _same(run(1), run(2))


def test_trace_detects_data_dependent_writes(vctx):
    """Positive control: OP_SCATTER (reference semantics, only selected rows
    written) gives different address digests for different masks."""
    from velodb_amd import _abi
    n = 1 << 14

    def run(p):
        mask = to_dev((np.random.default_rng(5).random(n) < p).astype(np.uint8))
        out = empty_dev(n, np.int32)
        cnt = empty_dev(1, np.uint64)
        cmd = _abi.Command()
        cmd.opcode = _abi.OP_SCATTER
        cmd.args.scatter = _abi.ScatterArgs(out.data_ptr(), cnt.data_ptr(), mask.data_ptr(), n)
        return _traced(vctx, lambda: vctx.wait(vctx.submit(cmd)))

    (t1, l1), (t2, l2) = run(0.1), run(0.6)
    assert [x[:3] for x in l1] == [x[:3] for x in l2]  # same launches and shapes
    assert t1.address_digest != t2.address_digest     # different addresses
