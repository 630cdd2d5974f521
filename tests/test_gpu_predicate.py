"""WHERE predicates on the device (OP_PREDICATE_MASK, the device form of
ComparisonExpression / BinaryLogicalExpression::evaluateBatch,
src/expression/comparison_expression.cpp:13-92, logical_expression.cpp:61-150):
bit-exact against the oracle's restatement for every column type and
comparison, with NULLs, AND and OR chains and pad rows; the Q6 predicate
against the fused filter's count at 2^23; trace identity across values."""
import numpy as np
import pytest

import oracle
from tests.gpu_util import empty_dev, from_dev, to_dev
from velodb_amd import _abi

pytestmark = pytest.mark.gpu

_TYPES = {  # type id -> numpy dtype of the column
    _abi.T_BOOLEAN: np.uint8, _abi.T_TINYINT: np.int8, _abi.T_SMALLINT: np.int16, _abi.T_INTEGER: np.int32,
    _abi.T_DATE: np.uint32, _abi.T_BIGINT: np.int64, _abi.T_FLOAT: np.float32, _abi.T_DOUBLE: np.float64,
    _abi.T_VARCHAR: np.uint64, _abi.T_UBIGINT: np.uint64}


def _column(rng, type_id, n):
    dt = np.dtype(_TYPES[type_id])
    if dt.kind == "f":
        v = rng.integers(-20, 20, n).astype(dt) / dt.type(4)  # repeats, so EQ / LE / GE see ties
        v[::97] = -0.0
        v[5::211] = np.nan  # IEEE: only NE holds for NaN
        return v
    if type_id == _abi.T_BOOLEAN:
        return rng.integers(0, 2, n).astype(dt)
    info = np.iinfo(dt)
    v = rng.integers(-30, 30, n) if info.min < 0 else rng.integers(0, 60, n)
    v = v.astype(dt)
    v[::101] = info.max
    v[1::103] = info.min
    return v


def _constant(type_id):
    return {_abi.T_BOOLEAN: 1, _abi.T_FLOAT: 0.25, _abi.T_DOUBLE: -1.5, _abi.T_DATE: 17,
            _abi.T_VARCHAR: 9, _abi.T_UBIGINT: 40}.get(type_id, -3)


def _run(vctx, terms, n, n_pad, conn):
    """terms: (numpy column, type, cmp, value, valid or None)."""
    dev = [(to_dev(t[0], max(n, 1)), t[1], t[2], t[3], None if t[4] is None else to_dev(t[4], max(n, 1)))
           for t in terms]
    out = to_dev(np.full(max(n_pad, 1), 0xEE, np.uint8))
    vctx.predicate_mask(dev, out, n, n_pad, connective=conn)
    return from_dev(out, np.uint8)[:n_pad]


@pytest.mark.parametrize("type_id", sorted(_TYPES))
@pytest.mark.parametrize("cmp", range(6))
def test_single_term_every_type_and_comparison(vctx, type_id, cmp):
    n, n_pad = 5000, 8192
    rng = np.random.default_rng(type_id * 10 + cmp)
    col = _column(rng, type_id, n)
    k = _constant(type_id)
    got = _run(vctx, [(col, type_id, cmp, k, None)], n, n_pad, _abi.CONN_AND)
    want = oracle.predicate_mask([(col, type_id, cmp, k)], n, n_pad)
    assert np.array_equal(got, want)
    assert not got[n:].any()


@pytest.mark.parametrize("conn", [_abi.CONN_AND, _abi.CONN_OR])
@pytest.mark.parametrize("n,n_pad", [(0, 4), (1, 1), (2, 2), (3, 4), (1000, 1001), (70000, 1 << 17),
                                     (1 << 16, 1 << 16), (300001, 300001)])
def test_chains_with_nulls(vctx, conn, n, n_pad):
    """Up to eight terms of mixed types, some with NULL rows, against the
    left-deep chain of the reference's binary AND / OR."""
    print(f"predicate chain conn={conn} n={n} n_pad={n_pad}", flush=True)
    rng = np.random.default_rng(n + 7 * conn)
    types = [_abi.T_INTEGER, _abi.T_BIGINT, _abi.T_DOUBLE, _abi.T_BOOLEAN, _abi.T_SMALLINT, _abi.T_DATE,
             _abi.T_FLOAT, _abi.T_TINYINT]
    for n_terms in (1, 2, 3, 8):
        terms = []
        for j in range(n_terms):
            t = types[j]
            col = _column(rng, t, max(n, 1))
            valid = (rng.random(max(n, 1)) > 0.1).astype(np.uint8) if j % 3 == 1 else None
            terms.append((col, t, int(rng.integers(0, 6)), _constant(t), valid))
        got = _run(vctx, terms, n, n_pad, conn)
        want = oracle.predicate_mask(terms, n, n_pad, conn)
        assert np.array_equal(got, want), n_terms


def test_unaligned_mask_pointer(vctx):
    """A mask that starts one byte into its buffer (n_pad a multiple of 4)
    takes the byte-store path and leaves the byte before it alone."""
    import torch
    n = 4096
    col = np.random.default_rng(3).integers(-50, 50, n).astype(np.int64)
    buf = torch.full((n + 8,), 0xEE, dtype=torch.uint8, device="cuda:0")
    vctx.predicate_mask([(to_dev(col), _abi.T_BIGINT, _abi.CMP_GE, 0)], buf[1:1 + n], n, n)
    got = buf.cpu().numpy()
    assert got[0] == 0xEE and got[1 + n] == 0xEE
    assert np.array_equal(got[1:1 + n], (col >= 0).astype(np.uint8))


def test_reference_expression_vectors(vctx):
    """tests/expression/test_expression_batch.cpp:46-48, 97-184 of the
    reference: val > 20 and (val > 10) AND active."""
    val = np.array([10, 20, 5, 30, 25], np.int32)
    act = np.array([1, 0, 1, 1, 0], np.uint8)
    got = _run(vctx, [(val, _abi.T_INTEGER, _abi.CMP_GT, 20, None)], 5, 8, _abi.CONN_AND)
    assert got.tolist() == [0, 0, 0, 1, 1, 0, 0, 0]
    got = _run(vctx, [(val, _abi.T_INTEGER, _abi.CMP_GT, 10, None), (act, _abi.T_BOOLEAN, _abi.CMP_NE, 0, None)],
               5, 8, _abi.CONN_AND)
    assert got.tolist() == [0, 0, 0, 1, 0, 0, 0, 0]


def test_null_operand_makes_row_unselected(vctx):
    """A NULL operand gives NULL under AND and OR alike
    (logical_expression.cpp:138-143), so TRUE OR NULL is not selected."""
    a = np.array([1, 1, 0, 0], np.int32)
    b = np.array([1, 0, 1, 0], np.int32)
    bv = np.array([1, 0, 0, 1], np.uint8)
    t = [(a, _abi.T_INTEGER, _abi.CMP_EQ, 1, None), (b, _abi.T_INTEGER, _abi.CMP_EQ, 1, bv)]
    assert _run(vctx, t, 4, 4, _abi.CONN_OR).tolist() == [1, 0, 0, 0]
    assert _run(vctx, t, 4, 4, _abi.CONN_AND).tolist() == [1, 0, 0, 0]


def test_q6_predicate_matches_fused_filter_at_2p23(vctx):
    """The Q6 WHERE clause as five generic terms selects exactly the rows of
    the fused filter: the same count (6e6 rows padded to 2^23) and the
    same mask as the oracle's Q6 mask on a 2^20 prefix."""
    import torch
    n, n_pad = 6_000_000, 1 << 23
    day = torch.empty(n_pad, dtype=torch.int32, device="cuda:0")
    qty, price, disc = (torch.empty(n_pad, dtype=torch.int64, device="cuda:0") for _ in range(3))
    vctx.gen_fact(1, day, qty, price, disc, n)
    out = torch.empty(4, dtype=torch.int64, device="cuda:0")
    op, od = torch.empty_like(price), torch.empty_like(disc)  # the inputs stay as generated
    vctx.filter_sum(day, qty, price, disc, n, n_pad, out, out_price=op, out_disc=od)
    mask = torch.empty(n_pad, dtype=torch.uint8, device="cuda:0")
    terms = [(day, _abi.T_INTEGER, _abi.CMP_GE, 8766), (day, _abi.T_INTEGER, _abi.CMP_LT, 9131),
             (disc, _abi.T_BIGINT, _abi.CMP_GE, 5), (disc, _abi.T_BIGINT, _abi.CMP_LE, 7),
             (qty, _abi.T_BIGINT, _abi.CMP_LT, 24)]
    vctx.predicate_mask(terms, mask, n, n_pad)
    torch.cuda.synchronize()
    assert int(mask.sum()) == int(out[0].item()) > 0
    m = 1 << 20
    d, q, _, di = oracle.gen_fact(1, m)
    want = oracle.filter_mask(d, q, di, m, m)
    assert np.array_equal(mask[:m].cpu().numpy(), want)


def test_trace_value_independent(vctx):
    """Two columns of different values, same shape: identical launches and
    address digests."""
    n = 50000

    def run(seed):
        col = np.random.default_rng(seed).integers(-100, 100, n).astype(np.int64)
        d = to_dev(col)
        out = empty_dev(n, np.uint8)
        vctx.trace_begin()
        vctx.predicate_mask([(d, _abi.T_BIGINT, _abi.CMP_LT, 0)], out, n, n)
        t = vctx.trace_end()
        return t, vctx.trace_launches(t.n_launches), from_dev(out, np.uint8)[:n], col

    (t1, l1, o1, c1), (t2, l2, o2, c2) = run(1), run(2)
    assert np.array_equal(o1, (c1 < 0).astype(np.uint8)) and np.array_equal(o2, (c2 < 0).astype(np.uint8))
    assert t1.n_launches == t2.n_launches > 0 and l1 == l2
    assert t1.launch_digest == t2.launch_digest and t1.address_digest == t2.address_digest


def test_rejected_arguments(vctx):
    from velodb_amd import VoError
    col = to_dev(np.arange(16, dtype=np.int32))
    out = empty_dev(16, np.uint8)
    with pytest.raises(VoError):
        vctx.predicate_mask([], out, 16, 16)
    with pytest.raises(VoError):
        vctx.predicate_mask([(col, _abi.T_INTEGER, _abi.CMP_EQ, 1)] * 9, out, 16, 16)
    with pytest.raises(VoError):
        vctx.predicate_mask([(col, _abi.T_INTEGER, 6, 1)], out, 16, 16)
    with pytest.raises(VoError):
        vctx.predicate_mask([(col, 99, _abi.CMP_EQ, 1)], out, 16, 16)
    with pytest.raises(VoError):
        vctx.predicate_mask([(col, _abi.T_INTEGER, _abi.CMP_EQ, 1)], out, 17, 16)
    with pytest.raises(VoError):
        vctx.predicate_mask([(col, _abi.T_INTEGER, _abi.CMP_EQ, 1)], out, 16, 16, connective=2)
    # still usable afterwards
    vctx.predicate_mask([(col, _abi.T_INTEGER, _abi.CMP_GE, 8)], out, 16, 16)
    assert from_dev(out, np.uint8)[:16].tolist() == [0] * 8 + [1] * 8
