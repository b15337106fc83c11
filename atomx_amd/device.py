"""Device memory and stream plumbing (PyTorch-ROCm tensors as HBM buffers)."""

from __future__ import annotations

import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def require_gpu():
    """Raises unless a HIP device is visible; the engine has no CPU path."""
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("atomx_amd requires a HIP GPU (torch.cuda.is_available() is False)")


def device():
    require_gpu()
    return torch().device("cuda", torch().cuda.current_device())


def to_dev(a, dtype=None):
    """Host array (or tensor) -> contiguous device tensor.

    Host data is staged through pinned memory and copied asynchronously on
    the current stream, so an upload does not wait for work already queued
    there (torch keeps the staging block until the copy has run)."""
    t = torch()
    if isinstance(a, t.Tensor):
        out = a.to(device=device(), dtype=dtype or a.dtype)
        return out.contiguous()
    arr = np.ascontiguousarray(a, dtype=dtype if dtype is not None else None)
    return t.from_numpy(arr).pin_memory().to(device(), non_blocking=True)


def f64(a):
    return to_dev(np.ascontiguousarray(a, dtype=np.float64)) if not _is_tensor(a) \
        else a.to(device=device(), dtype=torch().float64).contiguous()


def i32(a):
    return to_dev(np.ascontiguousarray(a, dtype=np.int32)) if not _is_tensor(a) \
        else a.to(device=device(), dtype=torch().int32).contiguous()


def empty(shape, dtype=None):
    t = torch()
    return t.empty(shape, dtype=dtype or t.float64, device=device())


def zeros(shape, dtype=None):
    t = torch()
    return t.zeros(shape, dtype=dtype or t.float64, device=device())


def ld(x):
    """Leading dimension (row pitch in elements) of a contiguous 2-D tensor.

    Taken from the shape, not from stride(0): a size-1 leading axis may carry
    any stride (numpy gives 0 for an inserted axis, and torch keeps it)."""
    if not x.is_contiguous():
        raise ValueError("device operand must be contiguous")
    return int(x.shape[-1])


def ptr(x):
    return None if x is None else x.data_ptr()


def stream_handle(stream=None):
    """hipStream_t of the given (or current) torch stream, as an integer."""
    t = torch()
    s = stream if stream is not None else t.cuda.current_stream()
    return s.cuda_stream


def _is_tensor(a):
    return isinstance(a, torch().Tensor)


def host(x):
    return x.detach().cpu().numpy()
