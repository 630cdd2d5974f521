"""Every opcode in the resident queue kernel (TaskManager / persistentKernel,
reference src/cuda/persistent_kernel.cu:18-124, which dispatches every
command inside one resident kernel): a command submitted to the resident
queue gives the same bytes as the same command in launch mode, and lanes of
the resident grid run commands side by side, each checked against the
oracle.

The resident kernel keeps every CU while it runs, so these tests only ever
synchronize torch's stream, never the device."""
import numpy as np
import pytest

import oracle
from tests.gpu_util import to_dev

pytestmark = pytest.mark.gpu


def _sync():
    import torch
    torch.cuda.current_stream().synchronize()


def _host(t, dtype):
    return t.cpu().numpy().view(dtype).copy()


@pytest.fixture
def rctx(torch_dev):
    import velodb_amd
    from velodb_amd import _abi
    ctx = velodb_amd.Context(flags=_abi.CFG_PERSISTENT, queue_capacity=64)
    ctx.queue_start()
    yield ctx
    ctx.queue_stop()
    ctx.close()


def _run(ctx, cmd, timeout_ms=60000):
    ctx.wait(ctx.submit(cmd), timeout_ms)


# ------------------------------------------------------------------ OP_SORT

_NP = {1: np.uint8, 2: np.int8, 3: np.int16, 4: np.int32, 5: np.int64, 6: np.float32, 7: np.float64, 10: np.uint32,
       64: np.uint64}


def _col(rng, t, n):
    dt = np.dtype(_NP[t])
    if dt.kind == "f":
        return rng.integers(-50, 50, n).astype(dt) / 4
    if t == 1:
        return rng.integers(0, 2, n).astype(dt)
    info = np.iinfo(dt)
    return rng.integers(max(info.min, -30), min(info.max, 30), n, endpoint=True).astype(dt)


@pytest.mark.parametrize("types,asc,n,n_pad", [
    ([5], [True], 100000, 1 << 17),            # one key word
    ([4, 7], [True, False], 5000, 8192),       # two columns in one word pair
    ([5, 5], [False, True], 70000, 1 << 17),   # two words
    ([5, 5, 5], [True, True, False], 3000, 4096),  # four words
    ([5] * 5, [True] * 5, 2000, 2048),         # eight words
    ([1, 3, 10], [False, True, True], 777, 1024),  # small: local pass below the tile
])
def test_resident_sort_multi_equals_launch_mode(rctx, vctx, types, asc, n, n_pad):
    import torch
    from velodb_amd import _abi
    print(f"resident OP_SORT types={types} n={n} n_pad={n_pad}", flush=True)
    rng = np.random.default_rng(n + len(types))
    cols = [_col(rng, t, n_pad) for t in types]
    la = [to_dev(c) for c in cols]
    ra = [to_dev(c) for c in cols]
    li = torch.empty(n_pad, dtype=torch.int64, device="cuda:0")
    ri = torch.empty(n_pad, dtype=torch.int64, device="cuda:0")
    vctx.sort(la, types, asc, li, n, n_pad)
    _sync()
    cmd = _abi.Command()
    cmd.opcode = _abi.OP_SORT
    a = cmd.args.sort
    for i, c in enumerate(ra):
        a.sort_cols[i] = c.data_ptr()
        a.col_types[i] = types[i]
        a.ascending_flags[i] = 1 if asc[i] else 0
    a.indices = ri.data_ptr()
    a.n_sort_columns = len(types)
    a.n_rows, a.n_padded_rows = n, n_pad
    _run(rctx, cmd)
    _sync()
    assert torch.equal(li, ri)
    for x, y in zip(la, ra):
        assert torch.equal(x, y)
    # and both are the oracle's
    ecols, eidx = oracle.bitonic_sort_multi(cols, types, asc, n, n_pad)
    assert np.array_equal(_host(ri, np.int64), eidx)
    for c, e in zip(ra, ecols):
        assert np.array_equal(_host(c, np.uint8), e.view(np.uint8))


# ------------------------------------------------------------ OP_SORT_BATCH

def test_resident_sort_batch_equals_launch_mode(rctx, vctx):
    import ctypes as C

    import torch
    from velodb_amd import _abi
    rng = np.random.default_rng(12)
    sizes = [3, 8, 9, 100, 1000, 4096, 5000, 30000, 70000, 1 << 17] + list(rng.integers(10, 20000, 40))
    items, off = [], 0
    for n in sizes:
        p = 1
        while p < n:
            p <<= 1
        off = (off + p - 1) // p * p
        items.append((off, int(n), p))
        off += p
    total = off
    k = rng.integers(0, 2**63, total, dtype=np.uint64)
    k[::7] = k[3]  # ties on the key
    pay = rng.integers(0, 2**32, total, dtype=np.uint64).astype(np.uint32)
    lk, lp = to_dev(k), to_dev(pay)
    rk, rp = to_dev(k), to_dev(pay)
    vctx.sort_batch(lk, lp, items)
    _sync()
    arr = (_abi.BatchItem * len(items))(*[_abi.BatchItem(o, n, p) for (o, n, p) in items])
    cmd = _abi.Command()
    cmd.opcode = _abi.OP_SORT_BATCH
    cmd.args.sort_batch = _abi.SortBatchArgs(rk.data_ptr(), rp.data_ptr(), C.cast(arr, C.c_void_p), len(items))
    _run(rctx, cmd)
    _sync()
    assert torch.equal(lk, rk) and torch.equal(lp, rp)


# -------------------------------------------------------------------- lanes

def test_lanes_run_commands_side_by_side(torch_dev):
    """Four lanes: a 2^22-record sort on lane 0 and small sorts on the other
    lanes, submitted together; every result matches the oracle, and each
    command's own completion flag reports it."""
    import torch

    import velodb_amd
    from velodb_amd import _abi
    ctx = velodb_amd.Context(flags=_abi.CFG_PERSISTENT, queue_capacity=64, queue_lanes=4)
    try:
        ctx.queue_start()
        assert ctx.queue_lanes() == 4
        big_n = 1 << 22
        bk = torch.empty(big_n, dtype=torch.int64, device="cuda:0")
        bp = torch.empty(big_n, dtype=torch.int32, device="cuda:0")
        with velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream) as g:
            g.gen_sort_input(3, bk, bp, big_n)
        small = []
        for i in range(12):
            n = 3000 + 83 * i
            k, p = oracle.gen_sort_input(900 + i, n)
            small.append((k, p, n, to_dev(k, 4096), to_dev(p, 4096)))
        _sync()

        def cmd(dk, dp, n, npad):
            c = _abi.Command()
            c.opcode = _abi.OP_SORT_KP
            c.args.sort_kp = _abi.SortKpArgs(dk.data_ptr(), dp.data_ptr(), n, npad)
            return c

        seqs = [ctx.submit(cmd(bk, bp, big_n, big_n), lane=0)]
        for i, (k, p, n, dk, dp) in enumerate(small):
            seqs.append(ctx.submit(cmd(dk, dp, n, 4096), lane=1 + i % 3))
        early = sum(ctx.poll(s) for s in seqs[1:])
        print(f"lanes: {early} of {len(seqs) - 1} small sorts done when polled right after submission", flush=True)
        for s in seqs:
            ctx.wait(s, 60000)
        ks = _host(bk, np.uint64)
        assert np.all(ks[1:] >= ks[:-1])
        assert np.array_equal(np.sort(_host(bp, np.uint32)), np.arange(big_n, dtype=np.uint32))
        for (k, p, n, dk, dp) in small:
            ek, ep = oracle.bitonic_sort_kp(k, p, n, 4096)
            assert np.array_equal(_host(dk, np.uint64), ek) and np.array_equal(_host(dp, np.uint32), ep)
        ctx.queue_stop()
    finally:
        ctx.close()


def test_lane_out_of_range_is_rejected(rctx):
    from velodb_amd import VoError, _abi
    c = _abi.Command()
    c.opcode = _abi.OP_NOP
    with pytest.raises(VoError):
        rctx.submit(c, lane=3)
    _run(rctx, c)
