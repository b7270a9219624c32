"""Synthetic RCB workloads (thin wrappers over libcoupe_workloads.so).

See coupe_amd/workloads.c for the exact generator definitions.  Sizes and
parameters of the standard workloads:

* ``uniform2d``    2^20 points uniform in [0,1)^2, unit weights, 8 parts
                   (iter_count 3), tolerance 0.05, seed 0
* ``gauss3d``      2^24 points from 8 Gaussians (sigma 0.05), weights in
                   [1, 100], 1024 parts (iter_count 10), tolerance 0.01, seed 1
* ``jgrid512``     512^3 = 2^27 jittered hex-cell centroids, lognormal integer
                   weights, 4096 parts (iter_count 12), tolerance 0.01, seed 2
* ``batch4096``    4096 instances, sizes log-uniform in [10^3, 10^6], D in
                   {2, 3}, integer weights in [1, 16], iter_count in 4..8,
                   tolerance 0.02, seed 3
"""
from __future__ import annotations

import numpy as np

from ._native import workloads_lib


def uniform(n: int, dim: int, seed: int, wmax: int = 0):
    pts = np.empty((n, dim), dtype=np.float64)
    w = np.empty(n, dtype=np.float64)
    workloads_lib().wl_uniform(pts.ctypes.data, w.ctypes.data, n, dim, seed, wmax)
    return pts, w


def gauss_mix(n: int, seed: int = 1, sigma: float = 0.05):
    pts = np.empty((n, 3), dtype=np.float64)
    w = np.empty(n, dtype=np.float64)
    workloads_lib().wl_gauss_mix(pts.ctypes.data, w.ctypes.data, n, seed, sigma)
    return pts, w


def jgrid(m: int, seed: int = 2, jitter: bool = True):
    n = m * m * m
    pts = np.empty((n, 3), dtype=np.float64)
    w = np.empty(n, dtype=np.float64)
    workloads_lib().wl_jgrid(pts.ctypes.data, w.ctypes.data, m, seed, 1 if jitter else 0)
    return pts, w


def box(n: int, dim: int, lo, ext, seed: int, wmax: int):
    pts = np.empty((n, dim), dtype=np.float64)
    w = np.empty(n, dtype=np.float64)
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    ext = np.ascontiguousarray(ext, dtype=np.float64)
    workloads_lib().wl_box(pts.ctypes.data, w.ctypes.data, n, dim, lo.ctypes.data, ext.ctypes.data, seed, wmax)
    return pts, w


def batch_header(count: int, seed: int = 3, lg_lo: float = 3.0, lg_hi: float = 6.0, l_lo: int = 4, l_hi: int = 8,
                 force_dim: int = 0):
    n = np.empty(count, dtype=np.uint64)
    dim = np.empty(count, dtype=np.int32)
    lo = np.empty((count, 3), dtype=np.float64)
    ext = np.empty((count, 3), dtype=np.float64)
    iters = np.empty(count, dtype=np.int32)
    workloads_lib().wl_batch_header(count, seed, lg_lo, lg_hi, l_lo, l_hi, force_dim, n.ctypes.data,
                                    dim.ctypes.data, lo.ctypes.data, ext.ctypes.data, iters.ctypes.data)
    return n, dim, lo, ext, iters


def instance_seed(seed: int, q: int) -> int:
    return workloads_lib().wl_instance_seed(seed, q)


# Standard workloads: (name, generator, iter_count, tolerance)
def standard(name: str, scale: int = 0):
    """Returns ``(points, weights, iter_count, tolerance)``.  ``scale`` > 0
    shrinks the point count by 2^scale (grid side by 2^(scale/3) for jgrid)."""
    if name == "uniform2d":
        pts, w = uniform((1 << 20) >> scale, 2, 0, 0)
        return pts, w, 3, 0.05
    if name == "gauss3d":
        pts, w = gauss_mix((1 << 24) >> scale, 1, 0.05)
        return pts, w, 10, 0.01
    if name == "jgrid512":
        m = 512 >> (scale // 3)
        pts, w = jgrid(m, 2, True)
        return pts, w, 12, 0.01
    if name == "grid512_unit":
        m = 512 >> (scale // 3)
        pts, w = jgrid(m, 2, False)
        return pts, w, 12, 0.01
    raise KeyError(name)
