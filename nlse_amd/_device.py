"""Device-memory plumbing on top of PyTorch-ROCm (allocation, streams,
host<->device copies).  No arithmetic of the solver happens here."""

from __future__ import annotations

import numpy as np
import torch

from . import _lib

_NP2TORCH = {
    np.dtype(np.complex64): torch.complex64,
    np.dtype(np.complex128): torch.complex128,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.float64): torch.float64,
}


def require_gpu() -> None:
    """Fail loudly when the engine cannot run: library missing or no GPU."""
    _lib.load()
    if not torch.cuda.is_available():
        raise ImportError("no ROCm GPU visible to PyTorch; the HIP engine needs one")


def device() -> torch.device:
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def is_device_tensor(x) -> bool:
    return isinstance(x, torch.Tensor) and x.is_cuda


def to_device(x, dtype=None) -> torch.Tensor:
    """numpy array / tensor -> contiguous device tensor (copy if needed)."""
    if isinstance(x, torch.Tensor):
        t = x
    else:
        a = np.ascontiguousarray(np.asarray(x))
        t = torch.from_numpy(a)
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device()).contiguous()


def to_host(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return x


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "expected a contiguous device tensor"
    return t.data_ptr()


def bparam(values, items: int):
    """Per-item parameter -> (device float64 tensor, BParam).  Scalars
    broadcast with stride 0; arrays must hold one value per item."""
    arr = np.asarray(values, dtype=np.float64).reshape(-1)
    if arr.size == 1:
        t = torch.tensor(arr, dtype=torch.float64, device=device())
        return t, _lib.BParam(t.data_ptr(), 0)
    if arr.size != items:
        raise ValueError(f"parameter with {arr.size} values for {items} items")
    t = torch.from_numpy(np.ascontiguousarray(arr)).to(device())
    return t, _lib.BParam(t.data_ptr(), 1)
