"""coupe_amd -- coupe's Recursive Coordinate Bisection on AMD Instinct MI355X.

Host-side mirror of coupe's RCB interface (coupe/src/lib.rs:76-91,
coupe/src/algorithms/recursive_bisection.rs:779-812):

    >>> import numpy as np, coupe_amd
    >>> points = np.array([[1., 1.], [-1., 1.], [1., -1.], [-1., -1.]])
    >>> weights = np.ones(4, dtype=np.int32)
    >>> part_ids = np.zeros(4, dtype=np.uint64)
    >>> coupe_amd.Rcb(iter_count=2).partition(part_ids, (points, weights))

``partition`` writes the part ids into ``part_ids`` like coupe's
``&mut [usize]`` and raises :class:`InputLenMismatch` like coupe's
``Error::InputLenMismatch``.  Host (numpy) inputs go through the C entry point
``coupe_rcb``; torch tensors already on the GPU go through
``coupe_hip_rcb_device`` with no PCIe copy.  The computation always runs in
the HIP library; nothing here computes ids.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Any, Tuple

import numpy as np

from . import _native
from ._native import strerror

__all__ = ["Rcb", "Rib", "Grid", "Partition", "Error", "InputLenMismatch", "CoupeError", "DeviceContext", "strerror",
           "partition_batch", "partition_batch_mixed", "autotune", "Server", "Mesh"]


class Error(Exception):
    """coupe::Error (coupe/src/algorithms.rs:45-73)."""


class InputLenMismatch(Error):
    """Input sets don't have matching lengths (coupe::Error::InputLenMismatch)."""

    def __init__(self, expected: int, actual: int):
        super().__init__(f"input sets don't have the same length (expected {expected} items, got {actual})")
        self.expected = expected
        self.actual = actual


class CoupeError(Error):
    """A status code of the C ABI other than the ones coupe::Error models."""

    def __init__(self, code: int):
        super().__init__(f"{strerror(code)} (coupe_err {code})")
        self.code = code


class Partition:
    """coupe::Partition: ``partition(part_ids, data) -> metadata``."""

    def partition(self, part_ids, data):  # pragma: no cover - interface
        raise NotImplementedError


_WTYPES = {np.dtype(np.int32): _native.COUPE_INT, np.dtype(np.int64): _native.COUPE_INT64,
           np.dtype(np.float64): _native.COUPE_DOUBLE}


def _is_device_tensor(x) -> bool:
    return hasattr(x, "data_ptr") and getattr(x, "is_cuda", False)


def _common_device(**tensors) -> int:
    """The device index shared by every tensor; TypeError when one is not a
    GPU tensor or they sit on different devices (a host pointer handed to a
    kernel would fault the GPU)."""
    devs = set()
    for name, t in tensors.items():
        if not _is_device_tensor(t):
            raise TypeError(f"{name} must be a GPU tensor when any input is (got {type(t).__name__}"
                            f"{' on ' + str(t.device) if hasattr(t, 'device') else ''})")
        devs.add(t.device.index or 0)
    if len(devs) != 1:
        raise TypeError(f"inputs on several devices: {sorted(devs)}")
    return devs.pop()


def _check_out_ids(part_ids, device=None) -> bool:
    """Checks a buffer the ids are written into as 64-bit integers: a
    contiguous int64/uint64 GPU tensor (on ``device`` when given) or numpy
    array.  True for a GPU tensor; TypeError otherwise (a narrower or strided
    buffer would be written past its end)."""
    import torch

    if _is_device_tensor(part_ids):
        if part_ids.dtype not in (torch.int64, torch.uint64) or not part_ids.is_contiguous():
            raise TypeError("part_ids must be a contiguous 64-bit integer tensor")
        if device is not None and (part_ids.device.index or 0) != device:
            raise TypeError(f"part_ids is on device {part_ids.device.index or 0}, the work runs on device {device}")
        return True
    if (not isinstance(part_ids, np.ndarray) or part_ids.dtype not in (np.int64, np.uint64)
            or not part_ids.flags.c_contiguous):
        raise TypeError("part_ids must be a contiguous 64-bit integer numpy array or GPU tensor")
    return False


def _check_device_points(points, device):
    """Device coordinates are read as contiguous float64 on ``device``."""
    import torch

    if points.dtype != torch.float64 or not points.is_contiguous():
        raise TypeError("points must be a contiguous float64 tensor")
    if (points.device.index or 0) != device:
        raise TypeError(f"points are on device {points.device.index or 0}, the work runs on device {device}")


def _weight_type(dtype) -> int:
    key = np.dtype(str(dtype).replace("torch.", ""))
    if key not in _WTYPES:
        raise TypeError(f"weights must be int32, int64 or float64, got {dtype}")
    return _WTYPES[key]


class DeviceContext:
    """Owns a coupe_hip_ctx (device work buffers reused across calls)."""

    def __init__(self, device: int = 0):
        self._ctx = _native.lib().coupe_hip_ctx_create(int(device))
        if not self._ctx:
            raise RuntimeError(f"no usable HIP device {device}")
        self.device = device

    @property
    def handle(self):
        return self._ctx

    def stats(self) -> dict:
        st = _native.Stats()
        _native.lib().coupe_hip_last_stats(self._ctx, C.byref(st))
        return st.as_dict()

    def last_rib(self) -> dict:
        """Rotation of this context's last Rib call (coupe_hip_last_rib)."""
        info = _native.RibInfo()
        err = _native.lib().coupe_hip_last_rib(self.handle, C.byref(info))
        if err != _native.OK:
            raise CoupeError(err)
        return info.as_dict()

    def set_profiling(self, enable: bool):
        _native.lib().coupe_hip_set_profiling(self._ctx, 1 if enable else 0)

    def kernel_times(self) -> dict:
        """Per-kernel-class launches, total event-timed ms and algorithmic
        bytes of the last call (profiling must be enabled)."""
        n = len(_native.KERNELS)
        arr = (_native.KernelTime * n)()
        _native.lib().coupe_hip_kernel_times(self._ctx, arr, n)
        return {k: {"launches": arr[i].launches, "total_ms": arr[i].total_ms, "algo_bytes": arr[i].algo_bytes}
                for i, k in enumerate(_native.KERNELS)}

    def close(self):
        if self._ctx:
            _native.lib().coupe_hip_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = {}


def _context_for(device: int) -> DeviceContext:
    if device not in _default_ctx:
        _default_ctx[device] = DeviceContext(device)
    return _default_ctx[device]


@dataclass
class Rcb(Partition):
    """coupe::Rcb { iter_count, tolerance } (recursive_bisection.rs:779-792).

    ``iter_count``: number of bisection levels (at most 2^iter_count parts).
    ``tolerance``: per-split tolerance on the normalised imbalance.
    """

    iter_count: int = 0
    tolerance: float = 0.0

    _host_fn = "coupe_rcb"
    _device_fn = "coupe_hip_rcb_device"

    def partition(self, part_ids, data: Tuple[Any, Any], *, stream=None, context: DeviceContext = None):
        points, weights = data
        if _is_device_tensor(points) or _is_device_tensor(part_ids):
            return self._partition_device(part_ids, points, weights, stream, context)
        return self._partition_host(part_ids, points, weights)

    # -- host memory: coupe_rcb (include/coupe.h) --------------------------------
    def _partition_host(self, part_ids, points, weights):
        pts = np.ascontiguousarray(points, dtype=np.float64)
        w = np.ascontiguousarray(weights)
        if w.dtype == np.bool_ or w.dtype.kind == "u":
            w = w.astype(np.int64)
        n_part = len(part_ids)
        # rcb(): weights first, then points (recursive_bisection.rs:661-672)
        if w.shape[0] != n_part:
            raise InputLenMismatch(n_part, w.shape[0])
        if pts.shape[0] != n_part:
            raise InputLenMismatch(n_part, pts.shape[0])
        if not isinstance(part_ids, np.ndarray) or part_ids.dtype != np.uint64 or not part_ids.flags.c_contiguous:
            raise TypeError("part_ids must be a contiguous numpy uint64 array (coupe's &mut [usize])")
        if n_part == 0:
            return None
        dim = pts.shape[1] if pts.ndim == 2 else 0
        wtype = _weight_type(w.dtype)
        L = _native.lib()
        dp = L.coupe_data_array(n_part, _native.COUPE_DOUBLE, pts.ctypes.data)
        dw = L.coupe_data_array(n_part, wtype, w.ctypes.data)
        try:
            err = getattr(L, self._host_fn)(part_ids.ctypes.data, dim, dp, dw, int(self.iter_count),
                                            float(self.tolerance))
        finally:
            L.coupe_data_free(dp)
            L.coupe_data_free(dw)
        if err != _native.OK:
            raise CoupeError(err)
        return None

    # -- device memory: coupe_hip_rcb_device (include/coupe_hip.h) --------------
    def _partition_device(self, part_ids, points, weights, stream, context):
        import torch

        n_part = part_ids.shape[0]
        if weights.shape[0] != n_part:
            raise InputLenMismatch(n_part, weights.shape[0])
        if points.shape[0] != n_part:
            raise InputLenMismatch(n_part, points.shape[0])
        if part_ids.dtype not in (torch.int64, torch.uint64) or not part_ids.is_contiguous():
            raise TypeError("part_ids must be a contiguous 64-bit integer tensor")
        if points.dtype != torch.float64 or not points.is_contiguous():
            raise TypeError("points must be a contiguous float64 tensor of shape (N, D)")
        if not weights.is_contiguous():
            raise TypeError("weights must be contiguous")
        dev = _common_device(part_ids=part_ids, points=points, weights=weights)
        ctx = context or _context_for(dev)
        if ctx.device != dev:
            raise TypeError(f"the context is on device {ctx.device}, the tensors on device {dev}")
        st = stream if stream is not None else torch.cuda.current_stream(points.device).cuda_stream
        dim = points.shape[1] if points.dim() == 2 else 0
        wtype = _weight_type(weights.dtype)
        err = getattr(_native.lib(), self._device_fn)(ctx.handle, C.c_void_p(st), part_ids.data_ptr(), n_part, dim,
                                                      points.data_ptr(), wtype, weights.data_ptr(),
                                                      int(self.iter_count), float(self.tolerance))
        if err != _native.OK:
            raise CoupeError(err)
        return None


@dataclass
class Rib(Rcb):
    """coupe::Rib { iter_count, tolerance } (recursive_bisection.rs:901-930):
    RCB after a change of basis whose first axis is the points' inertia axis.

    As coupe's ``rib`` (:829-851), an empty point set returns before any
    length check; otherwise the checks and errors are those of Rcb.  The
    rotation of the last device call is ``DeviceContext.last_rib()``.
    """

    _host_fn = "coupe_rib"
    _device_fn = "coupe_hip_rib_device"

    def partition(self, part_ids, data: Tuple[Any, Any], *, stream=None, context: DeviceContext = None):
        points, _ = data
        if len(points) == 0:   # OrientedBoundingBox::from_points is None
            return None
        return super().partition(part_ids, data, stream=stream, context=context)


def partition_batch(part_ids, points, weights, instances, *, stream=None, context: DeviceContext = None):
    """Batched RCB on device tensors (coupe_hip_rcb_batch, include/coupe_hip.h).

    ``instances`` is a sequence of ``(n, iter_count, tolerance)`` (or of
    ``(n, Rcb)``).  Instance q owns the next n rows of ``points`` (N, D)
    float64, of ``weights`` and of ``part_ids`` (int64/uint64), in order;
    its ids are those ``Rcb(iter_count, tolerance)`` gives it alone.
    """
    import torch

    specs = []
    for it in instances:
        if len(it) == 2 and isinstance(it[1], Rcb):
            specs.append((int(it[0]), int(it[1].iter_count), float(it[1].tolerance)))
        else:
            n, iters, tol = it
            specs.append((int(n), int(iters), float(tol)))
    total = sum(n for n, _, _ in specs)
    dev = _common_device(part_ids=part_ids, points=points, weights=weights)
    _check_out_ids(part_ids, dev)
    for name, t in (("part_ids", part_ids), ("points", points), ("weights", weights)):
        if t.shape[0] != total:
            raise InputLenMismatch(total, t.shape[0])
        if not t.is_contiguous():
            raise TypeError(f"{name} must be contiguous")
    if points.dtype != torch.float64:
        raise TypeError("points must be float64")
    arr = (_native.Instance * max(len(specs), 1))()
    for q, (n, iters, tol) in enumerate(specs):
        arr[q].n, arr[q].iter_count, arr[q].tolerance = n, iters, tol
    ctx = context or _context_for(dev)
    if ctx.device != dev:
        raise TypeError(f"the context is on device {ctx.device}, the tensors on device {dev}")
    st = stream if stream is not None else torch.cuda.current_stream(points.device).cuda_stream
    dim = points.shape[1] if points.dim() == 2 else 0
    err = _native.lib().coupe_hip_rcb_batch(ctx.handle, C.c_void_p(st), part_ids.data_ptr(), dim,
                                            points.data_ptr(), _weight_type(weights.dtype), weights.data_ptr(),
                                            arr, len(specs))
    if err != _native.OK:
        raise CoupeError(err)
    return None


def _mixed_instances(part_ids, points, weights, instances):
    """The instance table of a batch of both dimensions and the tensors'
    device, after the checks every such call needs (lengths, dtypes,
    layout, one device)."""
    import torch

    specs = [(int(n), int(d), int(it), float(tol)) for n, d, it, tol in instances]
    total = sum(n for n, _, _, _ in specs)
    coords = sum(n * d for n, d, _, _ in specs)
    dev = _common_device(part_ids=part_ids, points=points, weights=weights)
    _check_out_ids(part_ids, dev)
    for name, t, want in (("part_ids", part_ids, total), ("weights", weights, total)):
        if t.shape[0] != want:
            raise InputLenMismatch(want, t.shape[0])
        if not t.is_contiguous():
            raise TypeError(f"{name} must be contiguous")
    if points.dtype != torch.float64 or not points.is_contiguous() or points.numel() != coords:
        raise TypeError(f"points must be a contiguous float64 tensor of {coords} coordinates")
    arr = (_native.InstanceND * max(len(specs), 1))()
    for q, (n, d, it, tol) in enumerate(specs):
        arr[q].n, arr[q].dimension, arr[q].iter_count, arr[q].tolerance = n, d, it, tol
    return arr, len(specs), dev


def partition_batch_mixed(part_ids, points, weights, instances, *, stream=None, context: DeviceContext = None):
    """Batched RCB of instances of dimension 2 and 3 in one call
    (coupe_hip_rcb_batch_mixed, include/coupe_hip.h).

    ``instances`` is a sequence of ``(n, dimension, iter_count, tolerance)``.
    ``points`` is a flat float64 GPU tensor: instance q's n * dimension
    coordinates follow instance q-1's; ``weights`` and ``part_ids`` hold one
    entry per point in the same order.  Instance q's ids are those
    ``Rcb(iter_count, tolerance)`` gives it alone.
    """
    import torch

    arr, count, dev = _mixed_instances(part_ids, points, weights, instances)
    ctx = context or _context_for(dev)
    if ctx.device != dev:
        raise TypeError(f"the context is on device {ctx.device}, the tensors on device {dev}")
    st = stream if stream is not None else torch.cuda.current_stream(points.device).cuda_stream
    err = _native.lib().coupe_hip_rcb_batch_mixed(ctx.handle, C.c_void_p(st), part_ids.data_ptr(),
                                                  points.data_ptr(), _weight_type(weights.dtype),
                                                  weights.data_ptr(), arr, count)
    if err != _native.OK:
        raise CoupeError(err)
    return None


def autotune(part_ids, points, weights, instances, *, trials: int = 3, stream=None,
             context: DeviceContext = None) -> dict:
    """Chooses the engine's switches for a batch's class (coupe_hip_autotune,
    include/coupe_hip.h): same arguments as :func:`partition_batch_mixed`;
    later calls of that class on the context run with the switches kept.
    Returns them with the fastest mean call time (``best_ms``)."""
    import torch

    arr, count, dev = _mixed_instances(part_ids, points, weights, instances)
    ctx = context or _context_for(dev)
    if ctx.device != dev:
        raise TypeError(f"the context is on device {ctx.device}, the tensors on device {dev}")
    st = stream if stream is not None else torch.cuda.current_stream(points.device).cuda_stream
    chosen, ms = _native.Tuning(), C.c_double()
    err = _native.lib().coupe_hip_autotune(ctx.handle, C.c_void_p(st), part_ids.data_ptr(), points.data_ptr(),
                                           _weight_type(weights.dtype), weights.data_ptr(), arr, count,
                                           int(trials), C.byref(chosen), C.byref(ms))
    if err != _native.OK:
        raise CoupeError(err)
    out = chosen.as_dict()
    out["best_ms"] = ms.value
    return out


class Server:
    """Request server with continuous batching and resident meshes
    (coupe_hip_server_*, coupe_hip_mesh_*, include/coupe_hip.h).

    ``submit(part_ids, points, weights, iter_count, tolerance)`` takes numpy
    arrays (host memory) or torch tensors on the GPU and returns a ticket;
    ``wait(ticket)`` blocks until the ids are written and returns the
    request's latency figures; ``poll`` and ``cancel`` look at or withdraw a
    queued request.  ``create_mesh(points, weights)`` keeps a mesh resident in
    HBM; ``submit_mesh(mesh, part_ids, iter_count, tolerance)`` partitions it
    in place.  The arrays are kept alive until the request is waited for.
    Device tensors are read after the work queued on the current torch
    stream at submit time.
    """

    def __init__(self, device: int = 0, slots: int = 4, memory_budget: int = 64 << 30,
                 max_batch_points: int = 1 << 26, max_batch_requests: int = 256):
        cfg = _native.ServerConfig(device, slots, memory_budget, max_batch_points, max_batch_requests)
        self._s = _native.lib().coupe_hip_server_create(C.byref(cfg))
        if not self._s:
            raise RuntimeError(f"cannot start a server on device {device}")
        self.device = device
        self._keep = {}

    @staticmethod
    def _stream_of(t):
        import torch
        return torch.cuda.current_stream(t.device).cuda_stream

    def submit(self, part_ids, points, weights, iter_count: int, tolerance: float) -> int:
        dev_in = _is_device_tensor(points)
        dev_out = _is_device_tensor(part_ids)
        if dev_in != _is_device_tensor(weights):
            raise TypeError("points and weights must both be host arrays or both device tensors")
        if dev_in:
            _common_device(points=points, weights=weights)
            _check_device_points(points, self.device)
            if not weights.is_contiguous():
                raise TypeError("weights must be contiguous")
        else:
            points = np.ascontiguousarray(points, dtype=np.float64)
            weights = np.ascontiguousarray(weights)
        _check_out_ids(part_ids, self.device)
        n = points.shape[0]
        if weights.shape[0] != n:
            raise InputLenMismatch(n, weights.shape[0])
        if part_ids.shape[0] != n:
            raise InputLenMismatch(part_ids.shape[0], n)
        ptr = (lambda a: a.data_ptr()) if dev_in else (lambda a: a.ctypes.data)
        optr = part_ids.data_ptr() if dev_out else part_ids.ctypes.data
        stream = self._stream_of(points) if dev_in else None
        req = _native.Request(ptr(points), ptr(weights), _weight_type(weights.dtype), n,
                              points.shape[1] if points.ndim == 2 else 0, int(iter_count), float(tolerance),
                              optr, 1 if dev_in else 0, 1 if dev_out else 0, stream, 0)
        return self._submit(req, (part_ids, points, weights))

    def _submit(self, req, keep) -> int:
        t = C.c_uint64()
        err = _native.lib().coupe_hip_server_submit(self._s, C.byref(req), C.byref(t))
        if err != _native.OK:
            raise CoupeError(err)
        self._keep[t.value] = keep
        return t.value

    def create_mesh(self, points, weights) -> "Mesh":
        """A mesh resident in HBM (its working copy: f32 coordinates and the
        weights), from numpy arrays or GPU tensors."""
        return Mesh(self, points, weights)

    def submit_mesh(self, mesh: "Mesh", part_ids, iter_count: int, tolerance: float) -> int:
        """Partitions a resident mesh; ``part_ids`` (n ids, numpy or GPU
        tensor, or None to keep them on the mesh only) receives its ids."""
        if part_ids is not None and part_ids.shape[0] != mesh.n:
            raise InputLenMismatch(mesh.n, part_ids.shape[0])
        dev_out = part_ids is not None and _check_out_ids(part_ids, self.device)
        optr = None if part_ids is None else (part_ids.data_ptr() if dev_out else part_ids.ctypes.data)
        stream = self._stream_of(part_ids) if dev_out else None
        req = _native.Request(None, None, 0, 0, 0, int(iter_count), float(tolerance), optr, 0,
                              1 if dev_out else 0, stream, mesh.handle)
        return self._submit(req, (part_ids,))

    def wait(self, ticket: int) -> dict:
        st = _native.RequestStats()
        err = _native.lib().coupe_hip_server_wait(self._s, int(ticket), C.byref(st))
        self._keep.pop(ticket, None)
        if err != _native.OK:
            raise CoupeError(err)
        return {k: getattr(st, k) for k, _ in st._fields_}

    def poll(self, ticket: int) -> bool:
        done = C.c_int()
        err = _native.lib().coupe_hip_server_poll(self._s, int(ticket), C.byref(done))
        if err != _native.OK:
            raise CoupeError(err)
        return bool(done.value)

    def cancel(self, ticket: int) -> bool:
        """True when the request was still queued and is withdrawn (its
        ticket is then unknown); False when it started (wait for it)."""
        c = C.c_int()
        err = _native.lib().coupe_hip_server_cancel(self._s, int(ticket), C.byref(c))
        if err != _native.OK:
            raise CoupeError(err)
        if c.value:
            self._keep.pop(ticket, None)
        return bool(c.value)

    def stats(self) -> dict:
        st = _native.ServerStats()
        err = _native.lib().coupe_hip_server_stats(self._s, C.byref(st))
        if err != _native.OK:
            raise CoupeError(err)
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self._s:
            _native.lib().coupe_hip_server_destroy(self._s)
            self._s = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mesh:
    """A mesh resident in HBM on a Server (coupe_hip_mesh_*)."""

    def __init__(self, server: Server, points, weights):
        dev_in = _is_device_tensor(points)
        if dev_in != _is_device_tensor(weights):
            raise TypeError("points and weights must both be host arrays or both device tensors")
        if dev_in:
            _common_device(points=points, weights=weights)
            points, weights = points.contiguous(), weights.contiguous()
            _check_device_points(points, server.device)
            ptr = lambda a: a.data_ptr()
            stream = Server._stream_of(points)
        else:
            points = np.ascontiguousarray(points, dtype=np.float64)
            weights = np.ascontiguousarray(weights)
            ptr = lambda a: a.ctypes.data
            stream = None
        n = points.shape[0]
        if weights.shape[0] != n:
            raise InputLenMismatch(n, weights.shape[0])
        self.server, self.n = server, n
        self.dimension = points.shape[1] if points.ndim == 2 else 0
        self.weight_type = _weight_type(weights.dtype)
        h = C.c_uint64()
        err = _native.lib().coupe_hip_mesh_create(server._s, ptr(points), ptr(weights), self.weight_type, n,
                                                  self.dimension, 1 if dev_in else 0, stream, C.byref(h))
        if err != _native.OK:
            raise CoupeError(err)
        self.handle = h.value

    def update(self, points=None, weights=None, first: int = 0):
        """Replaces points [first, first + len) and/or their weights (numpy or
        GPU tensors, after the mesh's queued requests)."""
        ref = points if points is not None else weights
        if ref is None:
            return
        dev_in = _is_device_tensor(ref)
        count = ref.shape[0]
        for name, a in (("points", points), ("weights", weights)):
            if a is not None and a.shape[0] != count:
                raise InputLenMismatch(count, a.shape[0])
            if a is not None and _is_device_tensor(a) != dev_in:
                raise TypeError("points and weights must both be host arrays or both device tensors")
        if dev_in:
            conv = lambda a: None if a is None else a.contiguous()
            ptr = lambda a: None if a is None else a.data_ptr()
        else:
            conv = lambda a: None if a is None else np.ascontiguousarray(a)
            ptr = lambda a: None if a is None else a.ctypes.data
        points, weights = conv(points), conv(weights)
        if dev_in:
            for a in (points, weights):
                if a is not None and (a.device.index or 0) != self.server.device:
                    raise TypeError(f"the mesh lives on device {self.server.device}, an update on "
                                    f"device {a.device.index or 0}")
        if points is not None and (str(points.dtype).replace("torch.", "") != "float64"
                                   or points.ndim != 2 or points.shape[1] != self.dimension):
            raise TypeError(f"points must be float64 of shape (k, {self.dimension})")
        if weights is not None and _weight_type(weights.dtype) != self.weight_type:
            raise TypeError("weights must keep the mesh's weight type")
        stream = Server._stream_of(ref) if dev_in else None
        err = _native.lib().coupe_hip_mesh_update(self.server._s, self.handle, int(first), count, ptr(points),
                                                  ptr(weights), 1 if dev_in else 0, stream)
        if err != _native.OK:
            raise CoupeError(err)

    def parts(self):
        """(perm, offsets): the point indices of its last partition grouped by
        part (numpy uint32 / uint64)."""
        k = C.c_uint64()
        L = _native.lib()
        err = L.coupe_hip_mesh_parts(self.server._s, self.handle, None, None, 0, C.byref(k), 0)
        if err != _native.OK:
            raise CoupeError(err)
        perm = np.empty(self.n, dtype=np.uint32)
        off = np.empty(k.value + 1, dtype=np.uint64)
        err = L.coupe_hip_mesh_parts(self.server._s, self.handle, perm.ctypes.data, off.ctypes.data, off.shape[0],
                                     C.byref(k), 0)
        if err != _native.OK:
            raise CoupeError(err)
        return perm, off

    def destroy(self):
        if self.handle:
            err = _native.lib().coupe_hip_mesh_destroy(self.server._s, self.handle)
            self.handle = 0
            if err != _native.OK:
                raise CoupeError(err)


class Grid:
    """coupe::Grid (coupe/src/cartesian/mod.rs:43-180): a 2D or 3D cartesian
    grid partitioned by RCB on the GPU (coupe_hip_grid_rcb).  Weights and ids
    are in memory order, x fastest."""

    def __init__(self, size):
        size = tuple(int(s) for s in size)
        if len(size) not in (2, 3) or min(size) < 1:
            raise ValueError("a grid has 2 or 3 axes of at least one cell")
        self.size = size

    @classmethod
    def new_2d(cls, width: int, height: int) -> "Grid":
        return cls((width, height))

    @classmethod
    def new_3d(cls, width: int, height: int, depth: int) -> "Grid":
        return cls((width, height, depth))

    def __len__(self):
        n = 1
        for s in self.size:
            n *= s
        return n

    def rcb(self, partition, weights, iter_count: int, *, threads: int = 0, stream=None,
            context: DeviceContext = None):
        """Grid::rcb(&mut partition, &weights, iter_count).  ``partition`` and
        ``weights`` are numpy arrays (host) or GPU tensors.  ``threads``: the
        rayon pool size coupe's median chunking follows (0: RAYON_NUM_THREADS,
        else the logical CPU count)."""
        import torch

        n = len(self)
        if len(weights) != n or len(partition) != n:
            raise ValueError(f"a {self.size} grid has {n} cells")
        device_io = _is_device_tensor(weights)
        device_out = _is_device_tensor(partition)
        # the ids are written as 64-bit integers straight into a device
        # partition, so it must be one, on the device the weights are on
        if device_out and (partition.dtype not in (torch.int64, torch.uint64) or not partition.is_contiguous()):
            raise TypeError("partition must be a contiguous 64-bit integer tensor")
        if device_io and device_out:
            index = _common_device(partition=partition, weights=weights)
        elif device_io or device_out:
            index = (weights if device_io else partition).device.index or 0
        else:
            index = context.device if context is not None else 0
        if context is not None and context.device != index:
            raise TypeError(f"the context is on device {context.device}, the grid's tensors on device {index}")
        dev = torch.device("cuda", index)
        if device_io:
            w = weights.contiguous()
            wtype = _weight_type(w.dtype)
        else:
            wn = np.ascontiguousarray(weights)
            wtype = _weight_type(wn.dtype)
            w = torch.from_numpy(wn).to(dev)
        out = partition if device_out else torch.empty(n, dtype=torch.int64, device=dev)
        ctx = context or _context_for(index)
        st = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
        sz = (C.c_size_t * 3)(*(list(self.size) + [1] * (3 - len(self.size))))
        err = _native.lib().coupe_hip_grid_rcb(ctx.handle, C.c_void_p(st), len(self.size), sz, wtype, w.data_ptr(),
                                                int(iter_count), int(threads), out.data_ptr())
        if err != _native.OK:
            raise CoupeError(err)
        if not device_out:
            torch.cuda.synchronize(dev)
            partition[:] = out.cpu().numpy().astype(partition.dtype, copy=False)
        return None
