"""Host/device array plumbing for the drop-in layer.

Field data is either a NumPy array (as the caller gave it) or a torch
tensor on the HIP device (complex64, produced by the kernels).  Device
memory and streams come from torch; all arithmetic on field data goes
through libmftx.
"""
import numpy as np
import torch

from . import _lib


def is_dev(x):
    """A torch tensor in the library's device memory.  Anything else (NumPy
    arrays, CPU tensors, tensors on another device) is host-side data that is
    copied to the device where a kernel reads it, so no kernel is ever handed
    a pointer it cannot dereference."""
    return isinstance(x, torch.Tensor) and x.is_cuda and x.device == _lib.device()


def shape_of(x):
    return tuple(x.shape)


def size_of(x):
    return int(x.numel()) if is_dev(x) else int(np.size(x))


def to_dev(x, real=False):
    """Device tensor of x: complex128/float64 stay exact; the kernels read them directly."""
    dev = _lib.device()
    if isinstance(x, torch.Tensor):
        t = x.to(dev)
    else:
        a = np.asarray(x)
        if np.iscomplexobj(a):
            a = a.astype(np.complex128, copy=False)
        else:
            a = a.astype(np.float64, copy=False)
        a = np.ascontiguousarray(a)
        if not a.flags.writeable:
            a = a.copy()
        t = torch.from_numpy(a).to(dev)
    if real and t.is_complex():
        raise TypeError('expected a real array')
    if t.dim() == 2 and t.stride(1) != 1:
        t = t.contiguous()
    return t


def write_back(dst, src):
    """Copy a device tensor into a host-side destination (a NumPy array or a
    view of one, or a CPU or other-device torch tensor), in place."""
    if isinstance(dst, torch.Tensor):
        dst.copy_(src)
    else:
        dst[...] = src.cpu().numpy()


def to_host(x, dtype=complex):
    """NumPy copy (complex128 by default) of a host or device array."""
    if is_dev(x):
        x = x.detach().cpu().numpy()
    return np.asarray(x, dtype=dtype)


def row_stride(t):
    if t.dim() != 2 or t.stride(1) != 1:
        raise ValueError('expected a 2-D array with unit column stride')
    return t.stride(0)
