"""Helpers for GPU tests: move numpy arrays to device buffers (torch is only
the allocator/copy engine here) and back, bit for bit."""
import numpy as np

_SIGNED = {np.dtype(np.uint64): np.int64, np.dtype(np.uint32): np.int32, np.dtype(np.uint16): np.int16,
           np.dtype(np.uint8): np.uint8}


def to_dev(a, length=None, fill_byte=0xA5):
    """Device copy of `a`; with `length`, a buffer of that many elements whose
    tail beyond len(a) holds fill_byte garbage."""
    import torch
    a = np.ascontiguousarray(a)
    n = len(a) if length is None else length
    host = np.empty(n, a.dtype)
    host.view(np.uint8)[:] = fill_byte
    host[:len(a)] = a
    v = host.view(_SIGNED.get(host.dtype, host.dtype))
    return torch.from_numpy(v.copy()).to("cuda:0")


def from_dev(t, dtype):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().view(dtype).copy()


def empty_dev(n, dtype):
    import torch
    tmap = {np.dtype(np.uint64): torch.int64, np.dtype(np.int64): torch.int64, np.dtype(np.uint32): torch.int32,
            np.dtype(np.int32): torch.int32, np.dtype(np.uint8): torch.uint8, np.dtype(np.int16): torch.int16,
            np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32}
    return torch.empty(max(n, 1), dtype=tmap[np.dtype(dtype)], device="cuda:0")
