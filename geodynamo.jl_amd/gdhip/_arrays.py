"""Array plumbing: numpy host arrays (copied through GD_HOST) or torch device
tensors (passed as device pointers, ordered on torch's current stream)."""
import ctypes

import numpy as np

from ._lib import GD_HOST, check, lib

_ctx = {}


def context(device=0):
    """Library context for a HIP device (created once per process)."""
    if device not in _ctx:
        h = ctypes.c_void_p()
        check(lib().gd_ctx_create(device, ctypes.byref(h)))
        _ctx[device] = h
    return _ctx[device]


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def prepare(*arrays, device=0, shapes=None):
    """Return (ctx, flags, [pointers]) for arrays that are all host numpy
    float64 C-contiguous, or all torch float64 contiguous device tensors.
    shapes (optional): the exact shape each array must have (None: any), checked
    before any pointer reaches the library, which reads and writes exactly
    that many values."""
    kinds = {_is_torch(a) for a in arrays if a is not None}
    if len(kinds) > 1:
        raise TypeError("mix of host and device arrays in one call")
    on_device = kinds == {True}
    ptrs = []
    for k, a in enumerate(arrays):
        if a is None:
            ptrs.append(None)
            continue
        if on_device:
            import torch

            if not a.is_cuda or not a.is_contiguous() or a.dtype != torch.float64:
                raise TypeError("device arrays must be contiguous float64 torch tensors on the GPU")
            ptrs.append(ctypes.c_void_p(a.data_ptr()))
        else:
            if not isinstance(a, np.ndarray) or not a.flags.c_contiguous or a.dtype != np.float64:
                raise TypeError("host arrays must be C-contiguous float64 numpy arrays")
            ptrs.append(a.ctypes.data_as(ctypes.c_void_p))
        if shapes is not None and shapes[k] is not None and tuple(a.shape) != tuple(shapes[k]):
            raise ValueError("array %d has shape %s, the call needs %s" % (k, tuple(a.shape), tuple(shapes[k])))
    if on_device:
        import torch

        dev = next(a for a in arrays if a is not None).device.index or 0
        ctx = context(dev)
        check(lib().gd_ctx_set_stream(ctx, ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)))
        return ctx, 0, ptrs
    ctx = context(device)
    check(lib().gd_ctx_set_stream(ctx, None))
    return ctx, GD_HOST, ptrs


def f64(shape):
    return np.zeros(shape, dtype=np.float64)
