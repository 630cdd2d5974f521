"""Oblivious gather (the device form of the host oblivious materialization,
src/operator/materialization_operator.cpp:362-517): bit-exact against the
oracle's restatement, at full size against an indexing closed form, and
trace identity between repeated and distinct record ids."""
import numpy as np
import pytest

import oracle
from tests.gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu
_DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


@pytest.mark.parametrize("eb", [1, 2, 4, 8])
@pytest.mark.parametrize("n_table,n", [(1, 1), (1000, 5000), (5000, 300), (65536, 65536), (0, 10), (3, 4096)])
def test_parity_with_oracle(vctx, eb, n_table, n):
    print(f"gather_oblivious eb={eb} n_table={n_table} n={n}", flush=True)
    rng = np.random.default_rng(eb * 1000 + n_table + n)
    info = np.iinfo(_DT[eb])
    table = rng.integers(0, info.max, n_table, dtype=_DT[eb], endpoint=True)
    ids = rng.integers(-2, max(n_table, 1) + 2, n).astype(np.int64)
    if n > 20:
        ids[5:20] = ids[4]
    dt = to_dev(table) if n_table else empty_dev(1, np.uint64)
    di = to_dev(ids)
    out = empty_dev(n, _DT[eb] if eb != 2 else np.int16)
    found = empty_dev(n, np.uint8)
    vctx.gather_oblivious(dt, n_table, di, n, out, eb, out_found=found)
    got = from_dev(out, _DT[eb])[:n]
    gf = from_dev(found, np.uint8)[:n]
    want, wf = oracle.gather_recover(table.astype(np.uint64), ids)
    assert np.array_equal(gf, wf)
    assert np.array_equal(got.astype(np.uint64), want)


def test_full_size_against_indexing(vctx):
    """2^22 table rows, 2^23 requests with heavy repetition: equals
    table[ids] (torch indexing as the closed form)."""
    import torch
    print("gather_oblivious full size", flush=True)
    g = torch.Generator(device="cuda:0").manual_seed(3)
    nt, n = 1 << 22, 1 << 23
    table = torch.randint(-(1 << 62), 1 << 62, (nt,), device="cuda:0", generator=g, dtype=torch.int64)
    ids = torch.randint(0, nt, (n,), device="cuda:0", generator=g, dtype=torch.int64)
    ids[: n // 4] = 12345
    out = torch.empty(n, dtype=torch.int64, device="cuda:0")
    found = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    vctx.gather_oblivious(table, nt, ids, n, out, 8, out_found=found)
    torch.cuda.synchronize()
    assert bool(torch.all(out == table[ids]))
    assert bool(torch.all(found == 1))


def test_trace_value_independent(vctx):
    """All requests for one record vs all distinct records: identical
    launches and address digests."""
    n_table, n = 5000, 8000

    def run(ids):
        dt = to_dev(np.arange(n_table, dtype=np.uint64) * 3)
        di = to_dev(ids.astype(np.int64))
        out = empty_dev(n, np.uint64)
        vctx.trace_begin()
        vctx.gather_oblivious(dt, n_table, di, n, out, 8)
        t = vctx.trace_end()
        return t, vctx.trace_launches(t.n_launches), from_dev(out, np.uint64)[:n]

    a = np.full(n, 17)
    b = np.random.default_rng(1).permutation(np.arange(n) % n_table)
    (t1, l1, o1), (t2, l2, o2) = run(a), run(b)
    assert np.all(o1 == 51) and np.array_equal(o2, b.astype(np.uint64) * 3)
    assert t1.n_launches == t2.n_launches > 0 and l1 == l2
    assert t1.launch_digest == t2.launch_digest and t1.address_digest == t2.address_digest
