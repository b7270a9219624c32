"""Per-part loads and imbalance of a partition on the GPU.

Mirrors coupe::imbalance (coupe/src/imbalance.rs): ``compute_parts_load``
(:14-39), ``imbalance`` (:41-77), ``imbalance_target`` (:79-92),
``max_imbalance`` (:94-105), with the same argument order.  ``partition`` and ``weights`` are torch tensors on the GPU;
the computation runs in libcoupe_hip.so (coupe_hip_parts_load).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from . import CoupeError, DeviceContext, _context_for, _weight_type

_NP = {_native.COUPE_INT: np.int32, _native.COUPE_INT64: np.int64, _native.COUPE_DOUBLE: np.float64}


def _call(partition, weights, context, stream):
    import torch

    if not (getattr(partition, "is_cuda", False) and getattr(weights, "is_cuda", False)):
        raise TypeError("partition and weights must be GPU tensors")
    if partition.dtype not in (torch.int64, torch.uint64) or not partition.is_contiguous():
        raise TypeError("partition must be a contiguous 64-bit integer tensor")
    if not weights.is_contiguous():
        raise TypeError("weights must be contiguous")
    if partition.shape[0] != weights.shape[0]:
        raise ValueError("partition and weights differ in length")
    ctx = context or _context_for(partition.device.index or 0)
    st = stream if stream is not None else torch.cuda.current_stream(partition.device).cuda_stream
    return ctx, st, _weight_type(weights.dtype)


def compute_parts_load(partition, num_parts: int, weights, *, context: DeviceContext = None, stream=None):
    """Per-part sums of the weights (numpy array of the weight type)."""
    ctx, st, wt = _call(partition, weights, context, stream)
    out = np.zeros(max(int(num_parts), 1), dtype=_NP[wt])[: int(num_parts)]
    err = _native.lib().coupe_hip_parts_load(ctx.handle, C.c_void_p(st), partition.data_ptr(), partition.shape[0],
                                             wt, weights.data_ptr(), int(num_parts), out.ctypes.data)
    if err != _native.OK:
        raise CoupeError(err)
    return out


def imbalance(num_parts: int, partition, weights, *, context: DeviceContext = None, stream=None) -> float:
    """max over parts of (load - ideal) / ideal (0 for no part or zero load)."""
    ctx, st, wt = _call(partition, weights, context, stream)
    res = C.c_double()
    err = _native.lib().coupe_hip_imbalance(ctx.handle, C.c_void_p(st), partition.data_ptr(), partition.shape[0],
                                            wt, weights.data_ptr(), int(num_parts), C.byref(res))
    if err != _native.OK:
        raise CoupeError(err)
    return res.value


def imbalance_target(targets, partition, weights, *, context: DeviceContext = None, stream=None):
    """Largest load - target over the parts, in the weight type (0 without parts)."""
    t = np.asarray(targets)
    loads = compute_parts_load(partition, t.size, weights, context=context, stream=stream)
    if t.size == 0:
        return loads.dtype.type(0)
    return (loads - t.astype(loads.dtype)).max()


def max_imbalance(num_parts: int, partition, weights, *, context: DeviceContext = None, stream=None):
    """Largest minus smallest part load, in the weight type."""
    loads = compute_parts_load(partition, num_parts, weights, context=context, stream=stream)
    if loads.size == 0:
        return loads.dtype.type(0)
    return loads.max() - loads.min()
