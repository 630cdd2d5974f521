"""Resident command queue (TaskManager / persistentKernel, reference
src/cuda/task_manager.cu:26-124, src/cuda/persistent_kernel.cu:18-124):
commands run inside one cooperative kernel and give the same bytes as launch
mode and the oracle; ring back-pressure, restart after an idle exit,
commands run in order, rejection of invalid commands, an entry point that
parks the queue, and stop.

The queue kernel stays resident between commands, so these tests never call
a device-wide synchronize (it would wait for the resident kernel to leave):
they synchronize torch's stream only."""
import ctypes as C
import time

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
def qctx(torch_dev):
    import velodb_amd
    from velodb_amd import _abi
    ctx = velodb_amd.Context(flags=_abi.CFG_PERSISTENT, queue_capacity=16)
    ctx.queue_start()
    assert ctx.queue_running()
    yield ctx
    ctx.queue_stop()
    assert not ctx.queue_running()
    ctx.close()


def _sort_cmd(dk, dp, n, npad):
    from velodb_amd import _abi
    cmd = _abi.Command()
    cmd.opcode = _abi.OP_SORT_KP
    cmd.args.sort_kp = _abi.SortKpArgs(dk.data_ptr(), dp.data_ptr(), n, npad)
    return cmd


@pytest.mark.parametrize("n,npad", [(1, 1), (3, 4), (7, 8), (17, 32), (1000, 1024), (4095, 4096), (8192, 8192),
                                    (8193, 16384), (50000, 65536), (300000, 1 << 19), (1 << 21, 1 << 21)])
def test_queue_sort_kp_matches_oracle(qctx, n, npad):
    print(f"queue sort_kp n={n} npad={npad}", flush=True)
    k, p = oracle.gen_sort_input(500 + n, n)
    dk, dp = to_dev(k, npad), to_dev(p, npad)
    _sync()
    qctx.wait(qctx.submit(_sort_cmd(dk, dp, n, npad)), 20000)
    ek, ep = oracle.bitonic_sort_kp(k, p, n, npad)
    assert np.array_equal(_host(dk, np.uint64), ek)
    assert np.array_equal(_host(dp, np.uint32), ep)


def test_queue_equals_launch_mode_2p24(qctx, vctx):
    """Same input sorted by the resident kernel and by launch mode."""
    import torch
    n = 1 << 24
    dk = torch.empty(n, dtype=torch.int64, device="cuda:0")
    dp = torch.empty(n, dtype=torch.int32, device="cuda:0")
    vctx.gen_sort_input(7, dk, dp, n)
    dk2, dp2 = dk.clone(), dp.clone()
    _sync()
    qctx.wait(qctx.submit(_sort_cmd(dk, dp, n, n)), 30000)
    vctx.sort_kp(dk2, dp2, n, n)
    _sync()
    assert torch.equal(dk, dk2) and torch.equal(dp, dp2)
    ks = _host(dk, np.uint64)
    assert np.all(ks[1:] >= ks[:-1])


def test_queue_many_commands_back_pressure(qctx):
    """40 commands through a 16-entry ring: submit blocks until slots free."""
    bufs, seqs = [], []
    for i in range(40):
        n = 1000 + 70 * i
        k, p = oracle.gen_sort_input(i, n)
        dk, dp = to_dev(k, 4096), to_dev(p, 4096)
        bufs.append((k, p, n, dk, dp))
    _sync()
    for (k, p, n, dk, dp) in bufs:
        seqs.append(qctx.submit(_sort_cmd(dk, dp, n, 4096)))
    assert seqs == list(range(seqs[0], seqs[0] + 40))
    qctx.sync()
    for (k, p, n, dk, dp) in bufs:
        ek, ep = oracle.bitonic_sort_kp(k, p, n, 4096)
        assert np.array_equal(_host(dk, np.uint64), ek) and np.array_equal(_host(dp, np.uint32), ep)


def test_queue_elementwise_ops(qctx):
    import torch
    from velodb_amd import _abi
    rng = np.random.default_rng(3)
    n = 100003
    src = rng.integers(-2**62, 2**62, n, dtype=np.int64)
    idx = rng.permutation(n).astype(np.int64)
    pos = rng.permutation(n).astype(np.int32)
    mask = (rng.random(n) < 0.5).astype(np.uint8)
    a32 = rng.integers(0, 4, n).astype(np.int32)
    b32 = rng.integers(0, 4, n).astype(np.int32)
    old = rng.permutation(n).astype(np.uint32)
    sig = rng.permutation(n).astype(np.uint32)
    d_src, d_idx, d_pos = to_dev(src), to_dev(idx), to_dev(pos)
    d_mask, d_a, d_b = to_dev(mask), to_dev(a32), to_dev(b32)
    d_old, d_sig = to_dev(old), to_dev(sig)
    d_perm = torch.zeros(n, dtype=torch.int64, device="cuda:0")
    d_gath = torch.zeros(n, dtype=torch.int64, device="cuda:0")
    d_eq = torch.zeros(n, dtype=torch.uint8, device="cuda:0")
    d_new = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    d_enr = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    _sync()
    cmds = []
    c = _abi.Command(); c.opcode = _abi.OP_PERMUTE
    c.args.permute = _abi.PermuteArgs(d_perm.data_ptr(), d_src.data_ptr(), d_idx.data_ptr(), n, _abi.T_BIGINT)
    cmds.append(c)
    c = _abi.Command(); c.opcode = _abi.OP_GATHER
    c.args.gather = _abi.GatherArgs(d_gath.data_ptr(), d_src.data_ptr(), d_pos.data_ptr(), d_mask.data_ptr(), n,
                                    _abi.T_BIGINT)
    cmds.append(c)
    c = _abi.Command(); c.opcode = _abi.OP_COMPARE_EQ
    c.args.compare_eq = _abi.CompareEqArgs(d_a.data_ptr(), d_b.data_ptr(), d_eq.data_ptr(), n, _abi.T_INTEGER)
    cmds.append(c)
    c = _abi.Command(); c.opcode = _abi.OP_COMPOSE_POSITION_MAP
    c.args.compose_position_map = _abi.ComposeArgs(d_old.data_ptr(), d_sig.data_ptr(), d_new.data_ptr(), n)
    cmds.append(c)
    c = _abi.Command(); c.opcode = _abi.OP_ENRICH_POSITIONS
    c.args.enrich_positions = _abi.EnrichArgs(d_sig.data_ptr(), d_old.data_ptr(), d_enr.data_ptr(), n)
    cmds.append(c)
    last = 0
    for c in cmds:
        last = qctx.submit(c)
    qctx.wait(last, 20000)
    assert np.array_equal(_host(d_perm, np.int64), src[idx])
    g = np.zeros(n, np.int64)
    g[pos[mask != 0]] = src[mask != 0]
    assert np.array_equal(_host(d_gath, np.int64), g)
    assert np.array_equal(_host(d_eq, np.uint8), (a32 == b32).astype(np.uint8))
    assert np.array_equal(_host(d_new, np.uint32), old[sig])
    assert np.array_equal(_host(d_enr, np.uint32), old[sig])


def test_queue_runs_commands_in_order(qctx):
    """An oblivious operator (OP_SHUFFLE) and element-wise commands run in
    submission order in the resident kernel; invalid commands are rejected
    at submission and the queue keeps running."""
    import torch

    from velodb_amd import VoError, _abi
    n = 4096
    rng = np.random.default_rng(8)
    src = rng.integers(-1000, 1000, n).astype(np.int64)
    pay = torch.arange(n, dtype=torch.int32, device="cuda:0")
    d_src = torch.from_numpy(src).to("cuda:0")
    idx = torch.empty(n, dtype=torch.int64, device="cuda:0")
    out = torch.empty(n, dtype=torch.int64, device="cuda:0")
    torch.cuda.current_stream().synchronize()
    sh = _abi.Command()
    sh.opcode = _abi.OP_SHUFFLE
    sh.args.shuffle = _abi.ShuffleArgs(pay.data_ptr(), 0, n, 99, 0)
    # the shuffled rows as a permutation: COMPOSE with the identity map
    ident = torch.arange(n, dtype=torch.int32, device="cuda:0")
    perm32 = torch.empty(n, dtype=torch.int32, device="cuda:0")
    torch.cuda.current_stream().synchronize()
    s1 = qctx.submit(sh)
    comp = _abi.Command()
    comp.opcode = _abi.OP_COMPOSE_POSITION_MAP
    comp.args.compose_position_map = _abi.ComposeArgs(ident.data_ptr(), pay.data_ptr(), perm32.data_ptr(), n)
    s2 = qctx.submit(comp)                    # after the shuffle, on the same lane
    qctx.wait(s2, 20000)
    assert s2 > s1
    p = perm32.cpu().numpy()
    assert sorted(p.tolist()) == list(range(n)) and np.count_nonzero(p != np.arange(n)) > 0.9 * n
    idx.copy_(perm32.to(torch.int64))
    torch.cuda.current_stream().synchronize()
    pm = _abi.Command()
    pm.opcode = _abi.OP_PERMUTE
    pm.args.permute = _abi.PermuteArgs(out.data_ptr(), d_src.data_ptr(), idx.data_ptr(), n, _abi.T_BIGINT)
    qctx.wait(qctx.submit(pm), 20000)
    assert np.array_equal(out.cpu().numpy(), src[p])
    c = _abi.Command()
    c.opcode = _abi.OP_COMPACT                 # null buffers
    with pytest.raises(VoError):
        qctx.submit(c)
    c = _abi.Command()
    c.opcode = _abi.OP_PERMUTE
    c.args.permute = _abi.PermuteArgs(0, 0, 0, 10, 999)  # unknown type id
    with pytest.raises(VoError):
        qctx.submit(c)
    c = _abi.Command()
    c.opcode = _abi.OP_NOP
    qctx.wait(qctx.submit(c), 10000)  # the queue still runs


def test_queue_restarts_after_idle_exit(torch_dev):
    import velodb_amd
    from velodb_amd import _abi
    ctx = velodb_amd.Context(flags=_abi.CFG_PERSISTENT, queue_idle_us=2000)
    try:
        ctx.queue_start()
        for rnd in range(3):
            n = 5000 + rnd
            k, p = oracle.gen_sort_input(rnd, n)
            dk, dp = to_dev(k, 8192), to_dev(p, 8192)
            _sync()
            ctx.wait(ctx.submit(_sort_cmd(dk, dp, n, 8192)), 20000)
            ek, ep = oracle.bitonic_sort_kp(k, p, n, 8192)
            assert np.array_equal(_host(dk, np.uint64), ek) and np.array_equal(_host(dp, np.uint32), ep)
            t0 = time.time()
            while ctx.queue_running() and time.time() - t0 < 5:
                time.sleep(0.01)
            assert not ctx.queue_running()  # left after 2 ms without commands
        ctx.queue_stop()
    finally:
        ctx.close()


def test_queue_poll_alone_restarts_an_idle_kernel(torch_dev):
    """A caller that only polls (never waits) still sees its command finish
    after the resident kernel left on idle: vo_poll restarts it."""
    import velodb_amd
    from velodb_amd import _abi
    ctx = velodb_amd.Context(flags=_abi.CFG_PERSISTENT, queue_idle_us=2000)
    try:
        ctx.queue_start()
        for rnd in range(2):
            t0 = time.time()
            while ctx.queue_running() and time.time() - t0 < 5:
                time.sleep(0.01)
            assert not ctx.queue_running()
            n = 3000 + rnd
            k, p = oracle.gen_sort_input(40 + rnd, n)
            dk, dp = to_dev(k, 4096), to_dev(p, 4096)
            _sync()
            seq = ctx.submit(_sort_cmd(dk, dp, n, 4096))
            t0 = time.time()
            while not ctx.poll(seq):
                assert time.time() - t0 < 20, "command never finished under polling alone"
                time.sleep(0.001)
            ek, ep = oracle.bitonic_sort_kp(k, p, n, 4096)
            assert np.array_equal(_host(dk, np.uint64), ek) and np.array_equal(_host(dp, np.uint32), ep)
        ctx.queue_stop()
    finally:
        ctx.close()


def test_queue_parks_for_launch_only_call(qctx):
    """An entry point that runs in launch mode only (here the allocation-free
    compaction) first finishes the submitted commands and lets the resident
    kernel leave; vo_sync also waits for its work; the next submit starts the
    resident kernel again."""
    n = 6000
    k, p = oracle.gen_sort_input(77, n)
    dk, dp = to_dev(k, 8192), to_dev(p, 8192)
    vals = np.arange(1 << 12, dtype=np.uint64) * 3
    mask = (np.arange(1 << 12) % 3 == 0).astype(np.uint8)
    dv, dm = to_dev(vals), to_dev(mask)
    _sync()
    s1 = qctx.submit(_sort_cmd(dk, dp, n, 8192))
    qctx.compact(dm, [dv], [8], 1 << 12, 1 << 12, noalloc=True)  # launch mode: parks the queue first
    assert qctx.poll(s1)
    assert not qctx.queue_running()
    qctx.sync()
    assert np.array_equal(_host(dv, np.uint64)[:int(mask.sum())], vals[mask != 0])
    ek, ep = oracle.bitonic_sort_kp(k, p, n, 8192)
    assert np.array_equal(_host(dk, np.uint64), ek) and np.array_equal(_host(dp, np.uint32), ep)
    k2, p2 = oracle.gen_sort_input(78, n)
    dk2, dp2 = to_dev(k2, 8192), to_dev(p2, 8192)
    _sync()
    qctx.wait(qctx.submit(_sort_cmd(dk2, dp2, n, 8192)), 20000)
    ek2, ep2 = oracle.bitonic_sort_kp(k2, p2, n, 8192)
    assert np.array_equal(_host(dk2, np.uint64), ek2) and np.array_equal(_host(dp2, np.uint32), ep2)


def test_queue_start_without_persistent_flag_is_launch_mode(vctx):
    vctx.queue_start()  # no-op: commands keep launching on the context stream
    assert not vctx.queue_running()


_SECOND_PROCESS = r'''
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import torch
import oracle
import velodb_amd
from velodb_amd import _abi
from tests.gpu_util import to_dev

ctx = velodb_amd.Context(flags=_abi.CFG_PERSISTENT)
try:
    ctx.queue_start()
    print("started", flush=True)
    ctx.queue_stop()
except velodb_amd.VoError as e:
    print("refused:", e, flush=True)
n, npad = 3000, 4096
k, p = oracle.gen_sort_input(9, n)
dk, dp = to_dev(k, npad), to_dev(p, npad)
torch.cuda.current_stream().synchronize()
ctx.sort_kp(dk, dp, n, npad)  # launch mode when the queue did not start
ek, ep = oracle.bitonic_sort_kp(k, p, n, npad)
ok = np.array_equal(dk.cpu().numpy().view(np.uint64), ek) and np.array_equal(dp.cpu().numpy().view(np.uint32), ep)
print("sort ok" if ok else "sort differs", flush=True)
ctx.close()
'''


def test_queue_start_in_a_second_process(torch_dev):
    """After this process has made a cooperative launch, a child process's
    vo_queue_start either starts its own queue or (on this runtime, where
    cooperative launches are one process per GPU) is refused with an error
    that names the likely cause.  Either way the child's context then runs
    launch-mode work with the oracle's bytes: the refused launch leaves no
    stale HIP error behind for the next launch to report."""
    import os
    import subprocess
    import sys

    import velodb_amd
    from velodb_amd import _abi
    ctx = velodb_amd.Context(flags=_abi.CFG_PERSISTENT)
    ctx.queue_start()
    cmd = _abi.Command()
    cmd.opcode = _abi.OP_NOP
    ctx.wait(ctx.submit(cmd), 20000)
    ctx.queue_stop()
    ctx.close()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("second process: queue_start, then a launch-mode sort", flush=True)
    r = subprocess.run([sys.executable, "-c", _SECOND_PROCESS, root], capture_output=True, text=True, timeout=150)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "sort ok" in r.stdout
    if "refused" in r.stdout:
        assert "another process" in r.stdout
