"""ctypes bindings of libcoupe_hip.so (include/coupe.h, include/coupe_hip.h).

The library is loaded from this package directory only; a missing library is
an error (there is no fallback implementation).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# COUPE_HIP_LIBRARY selects another build of the library (A/B measurements).
LIB_PATH = os.environ.get("COUPE_HIP_LIBRARY") or os.path.join(HERE, "libcoupe_hip.so")
WORKLOADS_PATH = os.path.join(HERE, "libcoupe_workloads.so")

# enum coupe_err (include/coupe.h)
OK, ALLOC, CRASH, BAD_DIMENSION, BAD_TYPE, BIPART_ONLY, LEN_MISMATCH, NOT_FOUND, NEG_VALUES = range(9)
# enum coupe_type
COUPE_INT, COUPE_INT64, COUPE_DOUBLE = 0, 1, 2
# limits of coupe_hip_rcb_device_noalloc (include/coupe_hip.h)
NOALLOC_MAX_POINTS, NOALLOC_MAX_ITER_COUNT = 1 << 19, 12

# Every symbol the headers declare, in declaration order.
EXPORTS = [
    "coupe_strerror", "coupe_data_free", "coupe_data_array", "coupe_data_constant", "coupe_data_fn",
    "coupe_rcb", "coupe_rib",
    "coupe_hip_ctx_create", "coupe_hip_ctx_destroy", "coupe_hip_rcb_device", "coupe_hip_rcb_device_noalloc",
    "coupe_hip_rcb_batch", "coupe_hip_rcb_batch_mixed", "coupe_hip_autotune",
    "coupe_hip_rib_device", "coupe_hip_last_rib", "coupe_hip_grid_rcb",
    "coupe_hip_barycentres", "coupe_hip_parts_load", "coupe_hip_imbalance",
    "coupe_hip_server_create", "coupe_hip_server_destroy", "coupe_hip_server_submit", "coupe_hip_server_wait",
    "coupe_hip_server_poll", "coupe_hip_server_cancel", "coupe_hip_server_stats", "coupe_hip_mesh_create",
    "coupe_hip_mesh_update", "coupe_hip_mesh_parts", "coupe_hip_mesh_destroy",
    "coupe_hip_last_stats", "coupe_hip_set_profiling", "coupe_hip_kernel_times",
]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("slots", C.c_uint32), ("memory_budget", C.c_uint64),
                ("max_batch_points", C.c_uint64), ("max_batch_requests", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("points", C.c_void_p), ("weights", C.c_void_p), ("weight_type", C.c_int),
                ("n", C.c_size_t), ("dimension", C.c_size_t), ("iter_count", C.c_size_t),
                ("tolerance", C.c_double), ("partition", C.c_void_p), ("device_input", C.c_int),
                ("device_output", C.c_int), ("stream", C.c_void_p), ("mesh", C.c_uint64)]


class RequestStats(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("run_ms", C.c_double), ("total_ms", C.c_double),
                ("batch_requests", C.c_uint32), ("batch_points", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("memory_budget", C.c_uint64), ("device_bytes", C.c_uint64), ("mesh_bytes", C.c_uint64),
                ("peak_device_bytes", C.c_uint64), ("queued", C.c_uint32), ("running", C.c_uint32),
                ("meshes", C.c_uint32), ("completed", C.c_uint64), ("cancelled", C.c_uint64)]


class RibInfo(C.Structure):
    _fields_ = [("dimension", C.c_size_t), ("centroid", C.c_double * 3), ("inertia", C.c_double * 9),
                ("axis", C.c_double * 3), ("aabb_to_obb", C.c_double * 9), ("obb_to_aabb", C.c_double * 9)]

    def as_dict(self):
        d = self.dimension
        mat = lambda a: [[a[i * d + j] for j in range(d)] for i in range(d)]
        return {"dimension": d, "centroid": list(self.centroid[:d]), "inertia": mat(self.inertia),
                "axis": list(self.axis[:d]), "aabb_to_obb": mat(self.aabb_to_obb),
                "obb_to_aabb": mat(self.obb_to_aabb)}


class Instance(C.Structure):
    _fields_ = [("n", C.c_size_t), ("iter_count", C.c_size_t), ("tolerance", C.c_double)]


class InstanceND(C.Structure):
    _fields_ = [("n", C.c_size_t), ("dimension", C.c_size_t), ("iter_count", C.c_size_t), ("tolerance", C.c_double)]

KERNELS = ["prologue", "hist", "decide", "partition", "ids", "other", "moments"]


class KernelTime(C.Structure):
    _fields_ = [("launches", C.c_uint32), ("total_ms", C.c_double), ("algo_bytes", C.c_double)]

ITH_FN = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_size_t)


class Tuning(C.Structure):
    _fields_ = [("bins", C.c_uint32), ("workgroups_per_cu", C.c_uint32), ("search_passes", C.c_uint32),
                ("full_first", C.c_uint32), ("graphs", C.c_uint32), ("search_chunk", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [
        ("points", C.c_uint64),
        ("levels", C.c_uint32),
        ("passes", C.c_uint32),
        ("max_passes", C.c_uint32),
        ("weight_mode", C.c_uint32),
        ("weight_bytes", C.c_uint32),
        ("bins", C.c_uint32),
        ("reruns", C.c_uint32),
        ("unit_points", C.c_uint64),
        ("device_bytes", C.c_uint64),
        ("first_read", C.c_uint64),
        ("forwarded", C.c_uint32),
        ("full_first", C.c_uint32),
        ("graph", C.c_uint32),
        ("tuned", C.c_uint32),
        ("tuning", Tuning),
    ]

    def as_dict(self):
        return {k: (getattr(self, k).as_dict() if k == "tuning" else getattr(self, k)) for k, _ in self._fields_}


_lib = None


def lib():
    """The loaded product library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m coupe_amd.build`")
        # torch ships its own libamdhip64.so.7.  Loading it first makes this
        # library bind to the same HIP runtime instead of a second copy from
        # /opt/rocm (two runtimes in one process hide the GPU from torch).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        L.coupe_strerror.argtypes = [C.c_int]
        L.coupe_strerror.restype = C.c_char_p
        L.coupe_data_free.argtypes = [C.c_void_p]
        L.coupe_data_free.restype = None
        L.coupe_data_array.argtypes = [C.c_size_t, C.c_int, C.c_void_p]
        L.coupe_data_array.restype = C.c_void_p
        L.coupe_data_constant.argtypes = [C.c_size_t, C.c_int, C.c_void_p]
        L.coupe_data_constant.restype = C.c_void_p
        L.coupe_data_fn.argtypes = [C.c_void_p, C.c_size_t, C.c_int, ITH_FN]
        L.coupe_data_fn.restype = C.c_void_p
        L.coupe_rcb.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_size_t, C.c_double]
        L.coupe_rcb.restype = C.c_int
        L.coupe_rib.argtypes = L.coupe_rcb.argtypes
        L.coupe_rib.restype = C.c_int
        L.coupe_hip_ctx_create.argtypes = [C.c_int]
        L.coupe_hip_ctx_create.restype = C.c_void_p
        L.coupe_hip_ctx_destroy.argtypes = [C.c_void_p]
        L.coupe_hip_ctx_destroy.restype = None
        L.coupe_hip_rcb_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t,
                                           C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.c_double]
        L.coupe_hip_rcb_device.restype = C.c_int
        L.coupe_hip_rcb_device_noalloc.argtypes = L.coupe_hip_rcb_device.argtypes[1:]
        L.coupe_hip_rcb_device_noalloc.restype = C.c_int
        L.coupe_hip_rib_device.argtypes = L.coupe_hip_rcb_device.argtypes
        L.coupe_hip_rib_device.restype = C.c_int
        L.coupe_hip_last_rib.argtypes = [C.c_void_p, C.POINTER(RibInfo)]
        L.coupe_hip_last_rib.restype = C.c_int
        L.coupe_hip_rcb_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_int,
                                          C.c_void_p, C.POINTER(Instance), C.c_size_t]
        L.coupe_hip_rcb_batch.restype = C.c_int
        L.coupe_hip_rcb_batch_mixed.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                                C.POINTER(InstanceND), C.c_size_t]
        L.coupe_hip_rcb_batch_mixed.restype = C.c_int
        L.coupe_hip_autotune.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                         C.POINTER(InstanceND), C.c_size_t, C.c_uint32, C.POINTER(Tuning),
                                         C.POINTER(C.c_double)]
        L.coupe_hip_autotune.restype = C.c_int
        L.coupe_hip_grid_rcb.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t), C.c_int,
                                         C.c_void_p, C.c_size_t, C.c_uint32, C.c_void_p]
        L.coupe_hip_grid_rcb.restype = C.c_int
        L.coupe_hip_barycentres.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t,
                                            C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p]
        L.coupe_hip_barycentres.restype = C.c_int
        L.coupe_hip_parts_load.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p,
                                           C.c_size_t, C.c_void_p]
        L.coupe_hip_parts_load.restype = C.c_int
        L.coupe_hip_imbalance.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p,
                                          C.c_size_t, C.POINTER(C.c_double)]
        L.coupe_hip_imbalance.restype = C.c_int
        L.coupe_hip_server_create.argtypes = [C.POINTER(ServerConfig)]
        L.coupe_hip_server_create.restype = C.c_void_p
        L.coupe_hip_server_destroy.argtypes = [C.c_void_p]
        L.coupe_hip_server_destroy.restype = None
        L.coupe_hip_server_submit.argtypes = [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]
        L.coupe_hip_server_submit.restype = C.c_int
        L.coupe_hip_server_wait.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(RequestStats)]
        L.coupe_hip_server_wait.restype = C.c_int
        L.coupe_hip_server_poll.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_int)]
        L.coupe_hip_server_poll.restype = C.c_int
        L.coupe_hip_server_cancel.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_int)]
        L.coupe_hip_server_cancel.restype = C.c_int
        L.coupe_hip_server_stats.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
        L.coupe_hip_server_stats.restype = C.c_int
        L.coupe_hip_mesh_create.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_size_t, C.c_size_t,
                                            C.c_int, C.c_void_p, C.POINTER(C.c_uint64)]
        L.coupe_hip_mesh_create.restype = C.c_int
        L.coupe_hip_mesh_update.argtypes = [C.c_void_p, C.c_uint64, C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p,
                                            C.c_int, C.c_void_p]
        L.coupe_hip_mesh_update.restype = C.c_int
        L.coupe_hip_mesh_parts.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_size_t,
                                           C.POINTER(C.c_uint64), C.c_int]
        L.coupe_hip_mesh_parts.restype = C.c_int
        L.coupe_hip_mesh_destroy.argtypes = [C.c_void_p, C.c_uint64]
        L.coupe_hip_mesh_destroy.restype = C.c_int
        L.coupe_hip_last_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
        L.coupe_hip_last_stats.restype = C.c_int
        L.coupe_hip_set_profiling.argtypes = [C.c_void_p, C.c_int]
        L.coupe_hip_set_profiling.restype = C.c_int
        L.coupe_hip_kernel_times.argtypes = [C.c_void_p, C.POINTER(KernelTime), C.c_int]
        L.coupe_hip_kernel_times.restype = C.c_int
        _lib = L
    return _lib


def strerror(code: int) -> str:
    return lib().coupe_strerror(int(code)).decode()


_wl = None


def workloads_lib():
    global _wl
    if _wl is None:
        if not os.path.exists(WORKLOADS_PATH):
            raise RuntimeError(f"{WORKLOADS_PATH} is missing: build it with `python -m coupe_amd.build`")
        L = C.CDLL(WORKLOADS_PATH)
        P = C.c_void_p
        L.wl_uniform.argtypes = [P, P, C.c_uint64, C.c_int, C.c_uint64, C.c_int]
        L.wl_gauss_mix.argtypes = [P, P, C.c_uint64, C.c_uint64, C.c_double]
        L.wl_jgrid.argtypes = [P, P, C.c_uint64, C.c_uint64, C.c_int]
        L.wl_box.argtypes = [P, P, C.c_uint64, C.c_int, P, P, C.c_uint64, C.c_int]
        L.wl_batch_header.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                      P, P, P, P, P]
        L.wl_instance_seed.argtypes = [C.c_uint64, C.c_uint64]
        L.wl_instance_seed.restype = C.c_uint64
        L.wl_splitmix64.argtypes = [C.c_uint64, C.c_uint64]
        L.wl_splitmix64.restype = C.c_uint64
        L.wl_uniform01.argtypes = [C.c_uint64, C.c_uint64]
        L.wl_uniform01.restype = C.c_double
        _wl = L
    return _wl
