"""Device selection and host/device transfer helpers.

The product path needs the HIP device: ``device()`` raises when it is absent,
so nothing silently runs on the CPU.
"""
import numpy as np


def device():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('snoopy_amd needs a ROCm GPU (torch.cuda.is_available() is False)')
    return torch.device('cuda', torch.cuda.current_device())


def as_dev(a, dtype):
    import torch
    if isinstance(a, torch.Tensor):
        return a.to(device=device(), dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=device(), dtype=dtype)


def as_dev_f64(a):
    import torch
    return as_dev(a, torch.float64)


def as_dev_i32(a):
    import torch
    return as_dev(a, torch.int32)
