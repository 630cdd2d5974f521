"""Batched oblivious compactions (OP_COMPACT_BATCH): many independent
ORCompact compactions of uneven public sizes, one schedule per power-of-two
size class.  Each item must end exactly as the oracle's compaction of that
item alone (obliviator/join/enclave/oblivious_compact.c:59-338) and as one
vo_compact call on it, with its selected count in out_counts."""
import numpy as np
import pytest

import oracle
from tests.gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu

WIDTH_DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def _rows(cols):
    return np.concatenate([c.view(np.uint8).reshape(len(c), -1) for c in cols], axis=1)


def _layout(sizes, rng=None):
    """(offset, n, n_pad) per item: n_pad = max(4096, next power of two),
    offsets aligned to n_pad; with rng, items are placed in shuffled order."""
    order = list(range(len(sizes)))
    if rng is not None:
        rng.shuffle(order)
    place, off = {}, 0
    for i in order:
        n = int(sizes[i])
        n_pad = 4096
        while n_pad < n:
            n_pad <<= 1
        off = (off + n_pad - 1) // n_pad * n_pad
        place[i] = (off, n, n_pad)
        off += n_pad
    return [place[i] for i in range(len(sizes))], off


def _data(items, total, widths, seed):
    rng = np.random.default_rng(seed)
    cols = [rng.integers(0, np.iinfo(WIDTH_DT[w]).max, total, dtype=WIDTH_DT[w], endpoint=True) for w in widths]
    mask = np.zeros(total, np.uint8)
    for k, (off, n, n_pad) in enumerate(items):
        p = [0.0, 1.0, 0.02, 0.5, 0.9][k % 5]
        mask[off:off + n_pad] = rng.random(n_pad) < p  # rows >= n may be set: they must not count
    return cols, mask


@pytest.mark.parametrize("widths,shuffled", [([8, 4], False), ([8, 8, 2], True), ([1], True)])
def test_compact_batch_bit_exact_vs_oracle(vctx, widths, shuffled):
    rng = np.random.default_rng(len(widths))
    sizes = list(np.exp(rng.uniform(np.log(100), np.log(1 << 17), 90)).astype(int)) + [0, 1, 4095, 4096, 4097,
                                                                                          1 << 17]
    items, total = _layout(sizes, np.random.default_rng(5) if shuffled else None)
    print(f"compact batch widths={widths} items={len(items)} rows={total}", flush=True)
    cols, mask = _data(items, total, widths, 7)
    dcols = [to_dev(c) for c in cols]
    dcounts = empty_dev(len(items), np.uint64)
    vctx.compact_batch(to_dev(mask), dcols, widths, items, dcounts)
    got = _rows([from_dev(d, WIDTH_DT[w]) for d, w in zip(dcols, widths)])
    counts = from_dev(dcounts, np.uint64)
    rows = _rows(cols)
    for k, (off, n, n_pad) in enumerate(items):
        sel = mask[off:off + n_pad].copy()
        sel[n:] = 0
        assert int(counts[k]) == int(sel.sum()), (k, off, n, n_pad)
        expect = oracle.compact(rows[off:off + n_pad], sel)
        assert np.array_equal(got[off:off + n_pad], expect), (k, off, n, n_pad)


def test_compact_batch_leaves_gaps_untouched(vctx):
    """Rows between items (alignment gaps) are never read or written."""
    items = [(0, 3000, 4096), (16384, 9000, 16384), (8192, 4096, 4096)]
    total = 1 << 15
    rng = np.random.default_rng(2)
    c = rng.integers(0, 2**63, total, dtype=np.uint64)
    mask = (rng.random(total) < 0.4).astype(np.uint8)
    d = to_dev(c)
    vctx.compact_batch(to_dev(mask), [d], [8], items)
    got = from_dev(d, np.uint64)
    assert np.array_equal(got[4096:8192], c[4096:8192]) and np.array_equal(got[12288:16384], c[12288:16384])


def test_compact_batch_equals_one_compaction_per_item(vctx):
    """1024 items log-uniform in [2^12, 2^20] (the batched-compaction
    workload): the batch leaves every byte as one vo_compact per item."""
    import torch
    rng = np.random.default_rng(3)
    sizes = np.exp(rng.uniform(np.log(1 << 12), np.log(1 << 20), 1024)).astype(int)
    items, total = _layout(sizes)
    print(f"compact batch 1024 items rows={total}", flush=True)
    g = torch.Generator(device="cuda:0").manual_seed(4)
    c0 = torch.randint(-2**62, 2**62, (total,), device="cuda:0", generator=g)
    c1 = torch.randint(-2**31, 2**31 - 1, (total,), device="cuda:0", dtype=torch.int32, generator=g)
    mask = (torch.rand(total, device="cuda:0", generator=g) < 0.3).to(torch.uint8)
    b0, b1 = c0.clone(), c1.clone()
    counts = torch.zeros(len(items), dtype=torch.int64, device="cuda:0")
    one = torch.zeros(len(items), dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    vctx.compact_batch(mask, [b0, b1], [8, 4], items, counts)
    for k, (off, n, n_pad) in enumerate(items):
        vctx.compact(mask[off:], [c0[off:], c1[off:]], [8, 4], n, n_pad, out_count=one[k:])
    torch.cuda.synchronize()
    assert torch.equal(b0, c0) and torch.equal(b1, c1)
    assert torch.equal(counts, one)


@pytest.mark.parametrize("bad", ["small", "misaligned", "n_above"])
def test_compact_batch_rejects_bad_items(vctx, bad):
    import velodb_amd
    item = {"small": (0, 100, 2048), "misaligned": (4096, 10, 8192), "n_above": (0, 5000, 4096)}[bad]
    d = to_dev(np.zeros(1 << 14, np.uint64))
    with pytest.raises(velodb_amd.VoError):
        vctx.compact_batch(to_dev(np.zeros(1 << 14, np.uint8)), [d], [8], [item])
