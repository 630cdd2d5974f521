"""Batched oblivious sorts: many independent (u64, u32) sorts of uneven
public sizes in one schedule per size class, bit-exact per instance against
the oracle and against one vo_sort_kp call per instance."""
import numpy as np
import pytest

import oracle
from tests.gpu_util import from_dev, to_dev

pytestmark = pytest.mark.gpu


def _layout(sizes, shuffle_seed=None):
    """Offsets aligned to each instance's padded size; optional shuffled order."""
    items = []
    order = list(range(len(sizes)))
    if shuffle_seed is not None:
        np.random.default_rng(shuffle_seed).shuffle(order)
    off = 0
    place = {}
    for i in order:
        n = sizes[i]
        n_pad = 1 << max(0, int(np.ceil(np.log2(max(n, 1)))))
        off = (off + n_pad - 1) // n_pad * n_pad
        place[i] = (off, n, n_pad)
        off += n_pad
    items = [place[i] for i in range(len(sizes))]
    return items, off


def _fill(items, total, seed):
    keys = np.zeros(total, np.uint64)
    pay = np.zeros(total, np.uint32)
    for k, (off, n, n_pad) in enumerate(items):
        kk, pp = oracle.gen_sort_input(seed * 100000 + k, n)
        keys[off:off + n] = kk
        pay[off:off + n] = pp
    return keys, pay


@pytest.mark.parametrize("shuffle_seed", [None, 3])
def test_batch_bit_exact_per_instance(vctx, shuffle_seed):
    rng = np.random.default_rng(11)
    sizes = list(np.exp(rng.uniform(np.log(2 ** 4), np.log(2 ** 17), 300)).astype(int)) + [0, 1, 2, 3, 5, 8, 9, 16]
    items, total = _layout(sizes, shuffle_seed)
    keys, pay = _fill(items, total, 1)
    dk, dp = to_dev(keys), to_dev(pay)
    vctx.sort_batch(dk, dp, items)
    gk, gp = from_dev(dk, np.uint64), from_dev(dp, np.uint32)
    for (off, n, n_pad) in items:
        ek, ep = oracle.bitonic_sort_kp(keys[off:off + n], pay[off:off + n], n, n_pad)
        assert np.array_equal(gk[off:off + n_pad], ek) and np.array_equal(gp[off:off + n_pad], ep), (off, n, n_pad)


def test_batch_equals_one_sort_per_instance(vctx):
    import torch
    rng = np.random.default_rng(0)
    sizes = list(np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 20), 64)).astype(int))
    items, total = _layout(sizes)
    keys, pay = _fill(items, total, 2)
    dk1, dp1 = to_dev(keys), to_dev(pay)
    dk2, dp2 = dk1.clone(), dp1.clone()
    vctx.sort_batch(dk1, dp1, items)
    for (off, n, n_pad) in items:
        vctx.sort_kp_async(dk2[off:], dp2[off:], n, n_pad)
    vctx.sync()
    torch.cuda.synchronize()
    assert torch.equal(dk1, dk2) and torch.equal(dp1, dp2)


def test_autotuner_choices_are_valid_and_used(vctx):
    import velodb_amd
    import torch
    with velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream) as ctx:
        chosen = ctx.autotune_sort(13, 15, rows_per_class=1 << 18)
        assert set(chosen) == {13, 14, 15}
        for c, w in chosen.items():
            assert w != 0 and ctx.get_class_tuning(c) == w
        sizes = [1 << 13, (1 << 14) - 7, 1 << 15, 20000]
        items, total = _layout(sizes)
        keys, pay = _fill(items, total, 3)
        dk, dp = to_dev(keys), to_dev(pay)
        ctx.sort_batch(dk, dp, items)
        gk, gp = from_dev(dk, np.uint64), from_dev(dp, np.uint32)
        for (off, n, n_pad) in items:
            ek, ep = oracle.bitonic_sort_kp(keys[off:off + n], pay[off:off + n], n, n_pad)
            assert np.array_equal(gk[off:off + n_pad], ek) and np.array_equal(gp[off:off + n_pad], ep)


def test_batch_4096_sorts_2p10_to_2p20_equals_per_instance(vctx):
    """At full size: 4096 sorts, sizes log-uniform in [2^10, 2^20]
    (numpy default_rng(0)), 1.27e9 buffer rows: the batched schedule gives
    the same bytes as one sort per instance (both against device-generated
    inputs), and every instance comes out sorted."""
    import torch
    rng = np.random.default_rng(0)
    sizes = list(np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 20), 4096)).astype(int))
    items, total = _layout(sizes)
    keys = torch.empty(total, dtype=torch.int64, device="cuda:0")
    pay = torch.empty(total, dtype=torch.int32, device="cuda:0")
    vctx.gen_sort_input(7, keys, pay, total)
    vctx.sort_batch(keys, pay, items)
    torch.cuda.synchronize()
    kb, pb = keys.clone(), pay.clone()
    torch.cuda.synchronize()
    vctx.gen_sort_input(7, keys, pay, total)
    for (off, n, n_pad) in items:
        vctx.sort_kp_async(keys[off:], pay[off:], n, n_pad)
    vctx.sync()
    torch.cuda.synchronize()
    assert torch.equal(kb, keys) and torch.equal(pb, pay)
    # sortedness of the live rows of a spread of instances (unsigned order)
    sign = -(2 ** 63)
    for (off, n, n_pad) in items[::97]:
        k = kb[off:off + n] ^ sign
        assert bool(torch.all(k[1:] >= k[:-1]))
    del kb, pb, keys, pay
    torch.cuda.empty_cache()
