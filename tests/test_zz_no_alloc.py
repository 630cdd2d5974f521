"""vo_compact_noalloc, the allocation-free compaction: its output must equal
vo_compact's (and the oracle's) bit for bit, on consecutive calls that share
the fixed device scratch, the second smaller than the first, and at the
largest size it accepts; sizes above that are rejected before any launch."""
import numpy as np
import pytest

import oracle
from tests.gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu

WIDTH_DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def _rows(cols):
    return np.concatenate([c.view(np.uint8).reshape(len(c), -1) for c in cols], axis=1)


def _both(vctx, widths, n, n_pad, p, seed):
    rng = np.random.default_rng(seed)
    cols = [rng.integers(0, np.iinfo(WIDTH_DT[w]).max, n_pad, dtype=WIDTH_DT[w], endpoint=True) for w in widths]
    mask = (rng.random(n_pad) < p).astype(np.uint8)
    dmask = to_dev(mask)
    ref_cols = [to_dev(c) for c in cols]
    na_cols = [to_dev(c) for c in cols]
    ref_cnt, na_cnt = empty_dev(1, np.uint64), empty_dev(1, np.uint64)
    vctx.compact(dmask, ref_cols, widths, n, n_pad, out_count=ref_cnt)
    vctx.compact(dmask, na_cols, widths, n, n_pad, out_count=na_cnt, noalloc=True)
    vctx.sync()
    got = _rows([from_dev(d, WIDTH_DT[w]) for d, w in zip(na_cols, widths)])
    ref = _rows([from_dev(d, WIDTH_DT[w]) for d, w in zip(ref_cols, widths)])
    sel = mask.copy()
    sel[n:] = 0
    assert int(from_dev(na_cnt, np.uint64)[0]) == int(from_dev(ref_cnt, np.uint64)[0]) == int(sel.sum())
    assert np.array_equal(got, ref)
    assert np.array_equal(got, oracle.compact(_rows(cols), sel))


def test_noalloc_equals_vo_compact_on_two_calls_second_smaller(vctx):
    print("noalloc compact n_pad=2^20 then 2^14", flush=True)
    _both(vctx, [8, 4], (1 << 20) - 7, 1 << 20, 0.3, 1)
    _both(vctx, [8, 8, 2], 10000, 1 << 14, 0.6, 2)


def test_noalloc_at_its_largest_size(vctx):
    import torch
    n_pad = 1 << 29
    print("noalloc compact n_pad=2^29 one 1-byte column", flush=True)
    g = torch.Generator(device="cuda:0").manual_seed(3)
    mask = (torch.rand(n_pad, device="cuda:0", generator=g) < 0.02).to(torch.uint8)
    col = torch.arange(n_pad, device="cuda:0", dtype=torch.int64).to(torch.uint8)
    a, b = col.clone(), col.clone()
    ca, cb = empty_dev(1, np.uint64), empty_dev(1, np.uint64)
    vctx.compact(mask, [a], [1], n_pad - 3, n_pad, out_count=ca)
    vctx.compact(mask, [b], [1], n_pad - 3, n_pad, out_count=cb, noalloc=True)
    vctx.sync()
    m = int(mask[:n_pad - 3].sum().item())
    assert int(cb.item()) == int(ca.item()) == m
    assert torch.equal(a, b)
    assert torch.equal(b[:m], col[:n_pad - 3][mask[:n_pad - 3] != 0])
    del a, b, col, mask
    torch.cuda.empty_cache()


def test_noalloc_rejects_sizes_above_its_scratch(vctx):
    import velodb_amd
    dmask, dcol = empty_dev(4096, np.uint8), empty_dev(4096, np.uint64)
    with pytest.raises(velodb_amd.VoError, match="above 2\\^29"):
        vctx.compact(dmask, [dcol], [8], 4096, 1 << 30, noalloc=True)
