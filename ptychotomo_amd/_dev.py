"""Device-buffer plumbing on top of PyTorch-ROCm: raw pointers, the current
HIP stream, and allocation helpers.  PyTorch is used only for memory and
streams; every computation goes through the HIP library."""
import ctypes

import numpy as np
import torch

c64 = torch.complex64
f32 = torch.float32
f64 = torch.float64


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("ptychotomo_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def ptr(t):
    """Raw device address of a contiguous tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def to_dev(x, dtype):
    """numpy array / tensor -> contiguous device tensor of `dtype`."""
    if isinstance(x, torch.Tensor):
        return x.to(device=device(), dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x)).to(device=device(), dtype=dtype).contiguous()


def empty(shape, dtype=c64):
    return torch.empty(shape, dtype=dtype, device=device())


def zeros(shape, dtype=c64):
    return torch.zeros(shape, dtype=dtype, device=device())
