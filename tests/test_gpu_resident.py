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


def test_resident_compact_batch_equals_launch_mode(rctx, vctx):
    """OP_COMPACT_BATCH as one resident program (its size classes share the
    lane between barriers) gives the launch-mode bytes and counts."""
    from velodb_amd import Context, _abi
    rng = np.random.default_rng(13)
    sizes = [0, 4096, 5000, 30000, 70000, 1 << 17] + list(rng.integers(10, 40000, 40))
    items, off = [], 0
    for n in sizes:
        p = 4096
        while p < n:
            p <<= 1
        off = (off + p - 1) // p * p
        items.append((off, int(n), p))
        off += p
    c0 = rng.integers(0, 2**63, off, dtype=np.uint64)
    c1 = rng.integers(0, 2**16, off, dtype=np.uint64).astype(np.uint16)
    mask = (rng.random(off) < 0.3).astype(np.uint8)

    def make():
        return {"c0": to_dev(c0), "c1": to_dev(c1), "m": to_dev(mask),
                "cnt": to_dev(np.zeros(len(items), np.uint64))}

    def run(ctx, b):
        if ctx is vctx:
            ctx.compact_batch(b["m"], [b["c0"], b["c1"]], [8, 2], items, b["cnt"])
            return
        a, _arr = Context.compact_batch_args(b["m"], [b["c0"], b["c1"]], [8, 2], items, b["cnt"])
        cmd = _abi.Command()
        cmd.opcode = _abi.OP_COMPACT_BATCH
        cmd.args.compact_batch = a
        _run(ctx, cmd)

    d = _both(vctx, rctx, make, run)
    cnt = _host(d["cnt"], np.uint64)
    for k, (o, n, p) in enumerate(items):
        assert int(cnt[k]) == int(mask[o:o + min(n, p)].sum()), k


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


# ------------------------------------------- every other opcode, both modes

def _both(vctx, rctx, make, run):
    """make() -> fresh device buffers (deterministic); run(ctx, bufs) runs the
    operator.  The launch-mode and resident results must be the same bytes."""
    import torch
    dl = make()
    run(vctx, dl)
    _sync()
    dr = make()
    _sync()
    run(rctx, dr)
    _sync()
    for k in dl:
        assert torch.equal(dl[k], dr[k]), k
    return dr


def test_resident_compact(rctx, vctx):
    n, n_pad = 150000, 1 << 18
    rng = np.random.default_rng(21)
    c0 = rng.integers(0, 2**63, n_pad, dtype=np.uint64)
    c1 = rng.integers(0, 2**32, n_pad, dtype=np.uint64).astype(np.uint32)
    mask = (rng.random(n_pad) < 0.1).astype(np.uint8)

    def make():
        return {"c0": to_dev(c0), "c1": to_dev(c1), "m": to_dev(mask), "cnt": to_dev(np.zeros(1, np.uint64))}

    d = _both(vctx, rctx, make, lambda ctx, b: ctx.compact(b["m"], [b["c0"], b["c1"]], [8, 4], n, n_pad,
                                                           out_count=b["cnt"]))
    sel = mask.copy()
    sel[n:] = 0
    assert int(_host(d["cnt"], np.uint64)[0]) == int(sel.sum())


@pytest.mark.parametrize("in_place", [False, True])
def test_resident_filter_sum(rctx, vctx, in_place):
    n, n_pad = 300001, 1 << 19
    day, qty, price, disc = oracle.gen_fact(31, n)

    def make():
        b = {k: to_dev(v, n_pad) for k, v in (("day", day), ("qty", qty), ("price", price), ("disc", disc))}
        b["out"] = to_dev(np.zeros(4, np.uint64))
        if not in_place:
            b["op"], b["od"] = to_dev(np.zeros(n_pad, np.int64)), to_dev(np.zeros(n_pad, np.int64))
        return b

    def run(ctx, b):
        ctx.filter_sum(b["day"], b["qty"], b["price"], b["disc"], n, n_pad, b["out"],
                       out_price=b.get("op"), out_disc=b.get("od"))

    d = _both(vctx, rctx, make, run)
    m, s, _, _ = oracle.filter_closed_form(day, qty, price, disc)
    head = _host(d["out"], np.uint64)
    assert int(head[0]) == m and (int(head[2]) | (int(head[3]) << 64)) == s


@pytest.mark.parametrize("dense,sorted_out", [(False, True), (True, True), (True, False)])
def test_resident_join_groupsum(rctx, vctx, dense, sorted_out):
    n_dim, n_fact, n_pad = 20000, 50000, 1 << 17
    dk, dg, fk, fv = oracle.gen_join(41, n_dim, n_fact)

    def make():
        b = {k: to_dev(v) for k, v in (("dk", dk), ("dg", dg), ("fk", fk), ("fv", fv))}
        b["og"] = to_dev(np.zeros(64, np.int32))
        b["os"] = to_dev(np.zeros(64, np.int64))
        b["on"] = to_dev(np.zeros(1, np.uint64))
        if sorted_out:
            b["sk"], b["sv"] = to_dev(np.zeros(n_pad, np.uint64)), to_dev(np.zeros(n_pad, np.uint64))
        return b

    def run(ctx, b):
        ctx.join_groupsum(b["dk"], b["dg"], n_dim, b["fk"], b["fv"], n_fact, n_pad, b["og"], b["os"], b["on"], 32,
                          b.get("sk"), b.get("sv"), dense_groups=dense)

    d = _both(vctx, rctx, make, run)
    g, s = oracle.join_groupsum_closed_form(dk, dg, fk, fv, max_groups=32)
    G = int(_host(d["on"], np.uint64)[0])
    assert G == len(g)
    assert np.array_equal(_host(d["og"], np.int32)[:G], g) and np.array_equal(_host(d["os"], np.int64)[:G], s)


@pytest.mark.parametrize("with_perm", [False, True])
def test_resident_shuffle(rctx, vctx, with_perm):
    n, n_pad = 70000, 1 << 17
    pay = np.arange(n_pad, dtype=np.uint32) * 7

    def make():
        b = {"p": to_dev(pay)}
        if with_perm:
            b["perm"] = to_dev(np.zeros(n_pad, np.uint32))
        return b

    d = _both(vctx, rctx, make, lambda ctx, b: ctx.shuffle(b["p"], n_pad, 1234, perm_out=b.get("perm"), n=n))
    got = _host(d["p"], np.uint32)
    assert sorted(got[:n].tolist()) == sorted(pay[:n].tolist()) and np.array_equal(got[n:], pay[n:])


def test_resident_predicate_mask(rctx, vctx):
    from velodb_amd import _abi
    n, n_pad = 99999, 1 << 17
    rng = np.random.default_rng(51)
    a = rng.integers(-50, 50, n).astype(np.int32)
    b = rng.random(n) * 10
    c = rng.integers(0, 1000, n).astype(np.int64)
    valid = (rng.random(n) < 0.9).astype(np.uint8)

    def make():
        return {"a": to_dev(a), "b": to_dev(b), "c": to_dev(c), "v": to_dev(valid),
                "out": to_dev(np.full(n_pad, 0xEE, np.uint8))}

    def run(ctx, d):
        terms = [(d["a"], _abi.T_INTEGER, _abi.CMP_GE, -10), (d["b"], _abi.T_DOUBLE, _abi.CMP_LT, 7.5, d["v"]),
                 (d["c"], _abi.T_BIGINT, _abi.CMP_NE, 500)]
        ctx.predicate_mask(terms, d["out"], n, n_pad)

    d = _both(vctx, rctx, make, run)
    want = ((a >= -10) & (b < 7.5) & (valid != 0) & (c != 500)).astype(np.uint8)
    got = _host(d["out"], np.uint8)
    assert np.array_equal(got[:n], want) and not got[n:].any()


@pytest.mark.parametrize("eb", [4, 8])
def test_resident_gather_oblivious(rctx, vctx, eb):
    dt = {4: np.uint32, 8: np.uint64}[eb]
    rng = np.random.default_rng(61 + eb)
    table = rng.integers(0, np.iinfo(dt).max, 30000, dtype=dt, endpoint=True)
    ids = rng.integers(-2, 30002, 50000).astype(np.int64)

    def make():
        return {"t": to_dev(table), "i": to_dev(ids), "o": to_dev(np.zeros(50000, dt)),
                "f": to_dev(np.zeros(50000, np.uint8))}

    d = _both(vctx, rctx, make, lambda ctx, b: ctx.gather_oblivious(b["t"], len(table), b["i"], len(ids), b["o"], eb,
                                                                    out_found=b["f"]))
    want, wf = oracle.gather_recover(table.astype(np.uint64), ids)
    assert np.array_equal(_host(d["o"], dt).astype(np.uint64), want) and np.array_equal(_host(d["f"], np.uint8), wf)


def test_resident_scatter(rctx, vctx):
    from velodb_amd import _abi
    n = 100003
    mask = (np.random.default_rng(71).random(n) < 0.3).astype(np.uint8)

    def make():
        return {"m": to_dev(mask), "o": to_dev(np.full(n, -1, np.int32)), "c": to_dev(np.zeros(1, np.uint64))}

    def run(ctx, b):
        cmd = _abi.Command()
        cmd.opcode = _abi.OP_SCATTER
        cmd.args.scatter = _abi.ScatterArgs(b["o"].data_ptr(), b["c"].data_ptr(), b["m"].data_ptr(), n)
        ctx.wait(ctx.submit(cmd), 60000)

    d = _both(vctx, rctx, make, run)
    pos = np.cumsum(mask) - mask
    got = _host(d["o"], np.int32)
    assert np.array_equal(got[mask != 0], pos[mask != 0].astype(np.int32)) and np.all(got[mask == 0] == -1)


@pytest.mark.parametrize("type_id", [4, 5])
def test_resident_join_pairs_context_state(rctx, vctx, type_id):
    rng = np.random.default_rng(81 + type_id)
    dt = {4: np.int32, 5: np.int64}[type_id]
    lk = rng.integers(0, 3000, 20000).astype(dt)
    rk = rng.integers(0, 3000, 9000).astype(dt)
    out = {}
    for name, ctx in (("launch", vctx), ("resident", rctx)):
        dl, dr = to_dev(lk), to_dev(rk)
        n = ctx.join_pairs_count(dl, len(lk), dr, len(rk), type_id)
        cap = 1
        while cap < n:
            cap <<= 1
        ol, orr = to_dev(np.zeros(cap, np.int64)), to_dev(np.zeros(cap, np.int64))
        ctx.join_pairs_write(dl, len(lk), dr, len(rk), type_id, ol, orr, cap)
        _sync()
        out[name] = (n, _host(ol, np.int64), _host(orr, np.int64))
    assert out["launch"][0] == out["resident"][0]
    assert np.array_equal(out["launch"][1], out["resident"][1]) and np.array_equal(out["launch"][2], out["resident"][2])
    n = out["launch"][0]
    wl, wr = oracle.join_pairs(oracle.key_codes(lk, type_id), oracle.key_codes(rk, type_id))
    assert np.array_equal(out["resident"][1][:n], wl) and np.array_equal(out["resident"][2][:n], wr)


def test_resident_interleaved_joins_with_caller_state(rctx):
    """Two joins on one resident queue, COUNT A, COUNT B, WRITE A, WRITE B:
    each keeps its routing state in its own buffer (as the reference's
    MergeJoinBlock buffers), so the interleaving is correct."""
    import torch
    rng = np.random.default_rng(91)
    jobs = []
    for nl, nr, kmax in ((15000, 6000, 2000), (4000, 30000, 500)):
        lk = rng.integers(0, kmax, nl).astype(np.int32)
        rk = rng.integers(0, kmax, nr).astype(np.int32)
        st = torch.empty(rctx.join_pairs_state_bytes(nl, nr), dtype=torch.uint8, device="cuda:0")
        cnt = to_dev(np.zeros(1, np.uint64))
        jobs.append({"lk": lk, "rk": rk, "dl": to_dev(lk), "dr": to_dev(rk), "st": st, "cnt": cnt})
    _sync()
    for j in jobs:
        j["n"] = rctx.join_pairs_count(j["dl"], len(j["lk"]), j["dr"], len(j["rk"]), 4, out_count=j["cnt"],
                                       state=j["st"])
    for j in jobs:
        n = j["n"]
        assert int(_host(j["cnt"], np.uint64)[0]) == n
        cap = 1
        while cap < n:
            cap <<= 1
        j["ol"], j["or"] = to_dev(np.zeros(cap, np.int64)), to_dev(np.zeros(cap, np.int64))
        _sync()
        rctx.join_pairs_write(j["dl"], len(j["lk"]), j["dr"], len(j["rk"]), 4, j["ol"], j["or"], cap, state=j["st"],
                              n_out=n)
    for j in jobs:
        n = j["n"]
        wl, wr = oracle.join_pairs(oracle.key_codes(j["lk"], 4), oracle.key_codes(j["rk"], 4))
        assert n == len(wl)
        assert np.array_equal(_host(j["ol"], np.int64)[:n], wl) and np.array_equal(_host(j["or"], np.int64)[:n], wr)
