"""ctypes wrapper of the CPU oracle (oracle/rcb_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by ``bench.py``'s cpu_baseline leg.  The product
(``coupe_amd``) never imports this module.

The oracle restates coupe's ``Rcb`` (coupe/src/algorithms/recursive_bisection.rs
:456-705, coupe/src/geometry.rs:33-78) in C.  How each function is pinned:

* ``rcb`` / ``rcb_par`` -- the golden vectors in tests/golden/rcb_vectors.json
  (the reference's ``test_rcb_basic`` :1077-1115, the ``Rcb`` doctest
  :739-768, ``coupe-ffi/examples/rcb.c``, and hand-traced cases that hit
  every exit of the split loop); the closed form of dyadic grids (interleaved
  cell-index bits); an independent pure-Python transcription on small seeded
  inputs (tests/pyref.py); invariants of ``test_par_rcb_split``.
* ``bbox`` -- the reference's ``test_aabb_2d``/``test_aabb_3d`` values
  (coupe/src/geometry.rs:341-355, :489-503) and numpy min/max.
* ``split_1d`` -- the invariants of ``test_par_rcb_split`` (:1041-1075).
* ``reorder_1d`` -- the invariants of ``test_reorder_split_scalar``
  (:952-983).
* ``parts_load`` / ``imbalance`` / ``imbalance_target`` / ``max_imbalance``
  (numpy restatement of coupe/src/imbalance.rs:14-105) -- closed forms on
  hand-made partitions (equal loads give 0, one part holding everything gives
  num_parts - 1), i32 wrapping sums, the zero-part and zero-weight guards,
  and ``np.bincount`` for the per-part sums.
* ``rib_frame`` (inertia matrix, axis, Householder reflection, inverse;
  oracle/rib_oracle.c) -- the reference's geometry tests
  (geometry.rs:363-393 inertia matrix [[18,-18],[-18,18]] and axis
  parallel to (1,-1); :505-518 axis parallel to (1,-1,-1); :395-408 the box
  centre (3,3); :460-487 reflection columns orthonormal, first column
  parallel to the axis), numpy.linalg.eigh (LAPACK) on random symmetric
  matrices, and H H^-1 = I.  ``rib`` -- the ``Rib`` doctest
  (recursive_bisection.rs:866-894) and, given a rotation, ``rcb`` on the
  mapped points.  nalgebra's iterative eigen-solver is not restated, so the
  last bits of the rotation, and ids that depend on them, are "parity
  unpinned" against the reference; the GPU tests compare ids given the
  rotation the GPU used.
* ``barycentres`` (numpy restatement of tools/lib/lib.rs:511-539) -- the
  reference's MEDIT test mesh (mesh-io/src/medit/parser.rs:578-595) with its
  barycentres computed by hand, centres of unit squares / cubes, edges and
  lower-dimension elements left out, out-of-range nodes rejected.
* ``grid_rcb`` / ``weighted_median`` (oracle/grid_oracle.c, coupe's
  ``Grid::rcb``) -- the ``Grid`` doctest (mod.rs:21-41: a 2x2 grid, 2
  iterations, ids {0,1,2,3}), ``test_3d`` (rcb.rs:316-361: a 4^3 grid, 3
  iterations, 2x2x2 blocks), ``test_weighted_median`` (rcb.rs:300-314:
  left weight = sum of the weights before the position) on seeded inputs,
  and the closed form of power-of-two grids with unit weights (bit-halving
  splits, first axis 1, ids = interleaved bits).  The result depends on
  the rayon pool size (the median's chunk size); ``threads`` sets it.
* Non-integer f64 weights: the reference's summation order is rayon's and
  unspecified, so ids for real weights are "parity unpinned" beyond the
  1e-12 relative per-part weight tolerance used by the GPU tests.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "librcb_oracle.so")

W_INT, W_INT64, W_DOUBLE = 0, 1, 2


class OracleNode(C.Structure):
    _fields_ = [
        ("heap_id", C.c_uint64),
        ("count", C.c_uint64),
        ("count_left", C.c_uint64),
        ("weight_left", C.c_double),
        ("split_pos", C.c_float),
        ("steps", C.c_uint32),
        ("level", C.c_uint32),
        ("all_left", C.c_uint32),
    ]


class OracleStats(C.Structure):
    _fields_ = [
        ("ambiguous_pivots", C.c_uint64),
        ("candidates", C.c_uint64),
        ("nodes", C.c_uint64),
        ("trace", C.POINTER(OracleNode)),
        ("trace_cap", C.c_uint64),
        ("trace_len", C.c_uint64),
    ]


def build() -> None:
    """Compile the oracle library in place (gcc, see oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        sig = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p,
               C.c_size_t, C.c_int, C.c_size_t, C.c_double, C.POINTER(OracleStats)]
        L.oracle_rcb.argtypes = sig
        L.oracle_rcb.restype = C.c_int
        L.oracle_rcb_par.argtypes = sig + [C.c_int]
        L.oracle_rcb_par.restype = C.c_int
        L.oracle_rib.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p,
                                 C.c_size_t, C.c_int, C.c_size_t, C.c_double, C.c_void_p, C.c_void_p]
        L.oracle_rib.restype = C.c_int
        L.oracle_rib_frame.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t] + [C.c_void_p] * 5
        L.oracle_rib_frame.restype = C.c_int
        L.oracle_inertia_vector.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p]
        L.oracle_inertia_vector.restype = C.c_double
        L.oracle_householder.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p]
        L.oracle_householder.restype = None
        L.oracle_map_points.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p]
        L.oracle_map_points.restype = None
        L.oracle_grid_rcb.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_int, C.c_void_p, C.c_size_t,
                                      C.c_int, C.c_void_p]
        L.oracle_grid_rcb.restype = C.c_int
        L.oracle_weighted_median.argtypes = [C.c_int, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_void_p]
        L.oracle_weighted_median.restype = None
        L.oracle_bbox.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p]
        L.oracle_bbox.restype = C.c_int
        L.oracle_split_1d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_double, C.c_float,
                                      C.c_float, C.POINTER(C.c_uint64), C.POINTER(C.c_int64),
                                      C.POINTER(C.c_float)]
        L.oracle_split_1d.restype = C.c_int
        L.oracle_reorder_1d.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t]
        L.oracle_reorder_1d.restype = C.c_size_t
        L.oracle_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _kind_of(weights) -> int:
    dt = np.asarray(weights).dtype
    if dt == np.int32:
        return W_INT
    if dt == np.int64:
        return W_INT64
    if dt == np.float64:
        return W_DOUBLE
    raise TypeError(f"weights must be int32, int64 or float64, got {dt}")


def rcb(points, weights, iter_count, tolerance, *, n_part=None, constant=False,
        trace=False, parallel=False, nthreads=0):
    """Run the oracle.  Returns ``(err, ids, stats)``.

    ``points`` is (N, D) float64 (AoS), ``weights`` an int32/int64/float64
    array of N values (or of one value with ``constant=True``, like
    ``coupe_data_constant``).  ``ids`` is a uint64 array (coupe's usize ids).
    """
    pts = np.ascontiguousarray(points, dtype=np.float64)
    n = pts.shape[0]
    dim = pts.shape[1] if pts.ndim == 2 else 0
    w = np.ascontiguousarray(weights)
    kind = _kind_of(w)
    n_w = n if constant else w.shape[0]
    if n_part is None:
        n_part = n
    ids = np.zeros(max(n_part, 1), dtype=np.uint64)[:n_part]
    st = OracleStats()
    nodes = None
    if trace:
        cap = 1 << 16
        nodes = (OracleNode * cap)()
        st.trace = C.cast(nodes, C.POINTER(OracleNode))
        st.trace_cap = cap
    args = [ids.ctypes.data, n_part, dim, pts.ctypes.data, n, kind, w.ctypes.data, n_w,
            1 if constant else 0, iter_count, float(tolerance), C.byref(st)]
    if parallel:
        err = lib().oracle_rcb_par(*args, int(nthreads))
    else:
        err = lib().oracle_rcb(*args)
    out = {"ambiguous_pivots": st.ambiguous_pivots, "candidates": st.candidates, "nodes": st.nodes}
    if trace:
        out["trace"] = [
            dict(heap_id=x.heap_id, count=x.count, count_left=x.count_left,
                 weight_left=x.weight_left, split_pos=x.split_pos, steps=x.steps,
                 level=x.level, all_left=x.all_left)
            for x in nodes[: st.trace_len]
        ]
    return err, ids, out


def bbox(points):
    pts = np.ascontiguousarray(points, dtype=np.float64)
    dim = pts.shape[1]
    lo = np.zeros(dim)
    hi = np.zeros(dim)
    ok = lib().oracle_bbox(pts.ctypes.data, pts.shape[0], dim, lo.ctypes.data, hi.ctypes.data)
    return (lo, hi) if ok else None


def split_1d(coords, weights, tolerance, vmin, vmax):
    """One split of one node (``par_rcb_split``).  Returns the permuted
    coordinates and weights plus ``(count_left, weight_left, split_pos)``."""
    c = np.ascontiguousarray(coords, dtype=np.float32).copy()
    w = np.ascontiguousarray(weights, dtype=np.int64).copy()
    cl = C.c_uint64()
    wl = C.c_int64()
    sp = C.c_float()
    lib().oracle_split_1d(c.ctypes.data, w.ctypes.data, c.shape[0], float(tolerance),
                          float(vmin), float(vmax), C.byref(cl), C.byref(wl), C.byref(sp))
    return c, w, cl.value, wl.value, sp.value


def reorder_1d(coords, pivot):
    c = np.ascontiguousarray(coords, dtype=np.float32).copy()
    left = lib().oracle_reorder_1d(c.ctypes.data, c.shape[0], int(pivot))
    return c, left


def num_threads() -> int:
    return lib().oracle_num_threads()


def parts_load(partition, num_parts, weights):
    """coupe::imbalance::compute_parts_load (coupe/src/imbalance.rs:14-39):
    per-part sums of the weights in the weight type (i32 sums wrap mod 2^32
    as Rust's release arithmetic does).  Raises IndexError for an id >=
    num_parts (coupe's ``acc[part]`` panics)."""
    ids = np.asarray(partition, dtype=np.uint64)
    w = np.asarray(weights)
    if ids.size and int(ids.max()) >= num_parts:
        raise IndexError("part id out of range")
    if w.dtype == np.float64:
        out = np.zeros(num_parts, dtype=np.float64)
        for p, x in zip(ids.tolist(), w.tolist()):
            out[p] += x
        return out
    acc = [0] * num_parts
    for p, x in zip(ids.tolist(), w.tolist()):
        acc[p] += int(x)
    bits = 32 if w.dtype == np.int32 else 64
    wrap = [((a + (1 << (bits - 1))) % (1 << bits)) - (1 << (bits - 1)) for a in acc]
    return np.array(wrap, dtype=np.int32 if bits == 32 else np.int64)


def _wsum(vals, dtype):
    if dtype == np.float64:
        s = 0.0
        for x in vals.tolist():
            s += x
        return s
    bits = 32 if dtype == np.int32 else 64
    s = sum(int(x) for x in vals.tolist())
    return ((s + (1 << (bits - 1))) % (1 << bits)) - (1 << (bits - 1))


def imbalance(num_parts, partition, weights):
    """coupe::imbalance::imbalance (coupe/src/imbalance.rs:41-77)."""
    if num_parts == 0:
        return 0.0
    loads = parts_load(partition, num_parts, weights)
    total = _wsum(loads, loads.dtype)
    ideal = float(total) / float(num_parts)
    if ideal == 0.0:
        return 0.0
    return max((float(x) - ideal) / ideal for x in loads.tolist())


def imbalance_target(targets, partition, weights):
    """coupe::imbalance::imbalance_target (coupe/src/imbalance.rs:79-92):
    the largest load - target, in the weight type (0 without parts)."""
    t = np.asarray(targets)
    loads = parts_load(partition, t.size, weights)
    if t.size == 0:
        return loads.dtype.type(0)
    return max(loads[i] - t[i] for i in range(t.size))


def max_imbalance(num_parts, partition, weights):
    """coupe::imbalance::max_imbalance (coupe/src/imbalance.rs:94-105):
    largest minus smallest part load, in the weight type."""
    loads = parts_load(partition, num_parts, weights)
    if loads.size == 0:
        return loads.dtype.type(0)
    return loads.max() - loads.min()


def rib_frame(points):
    """Rotation of ``Rib`` for these points: dict with centroid, inertia,
    axis, aabb_to_obb (Householder reflection) and obb_to_aabb (its inverse),
    or None for no points."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    n, dim = pts.shape
    c, m, a, h, inv = (np.zeros(dim), np.zeros(dim * dim), np.zeros(dim), np.zeros(dim * dim), np.zeros(dim * dim))
    rc = lib().oracle_rib_frame(pts.ctypes.data, n, dim, c.ctypes.data, m.ctypes.data, a.ctypes.data,
                                h.ctypes.data, inv.ctypes.data)
    if rc == 0:
        return None
    return {"centroid": c, "inertia": m.reshape(dim, dim), "axis": a, "aabb_to_obb": h.reshape(dim, dim),
            "obb_to_aabb": inv.reshape(dim, dim), "singular": rc < 0}


def inertia_vector(matrix):
    m = np.ascontiguousarray(matrix, dtype=np.float64)
    dim = m.shape[0]
    v = np.zeros(dim)
    lam = lib().oracle_inertia_vector(m.ctypes.data, dim, v.ctypes.data)
    return v, lam


def householder(axis):
    e = np.ascontiguousarray(axis, dtype=np.float64)
    h = np.zeros(e.size * e.size)
    lib().oracle_householder(e.ctypes.data, e.size, h.ctypes.data)
    return h.reshape(e.size, e.size)


def map_points(points, obb_to_aabb):
    pts = np.ascontiguousarray(points, dtype=np.float64)
    m = np.ascontiguousarray(obb_to_aabb, dtype=np.float64)
    out = np.empty_like(pts)
    lib().oracle_map_points(pts.ctypes.data, pts.shape[0], pts.shape[1], m.ctypes.data, out.ctypes.data)
    return out


def rib(points, weights, iter_count, tolerance, *, obb_to_aabb=None, n_part=None):
    """``Rib`` on the CPU: (err, ids).  With ``obb_to_aabb`` the given
    rotation replaces the oracle's own."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    w = np.ascontiguousarray(weights)
    kind = _kind_of(w)
    n = pts.shape[0]
    n_part = n if n_part is None else n_part
    ids = np.zeros(max(n_part, 1), dtype=np.uint64)
    dim = pts.shape[1] if pts.ndim == 2 else 0
    m = None if obb_to_aabb is None else np.ascontiguousarray(obb_to_aabb, dtype=np.float64)
    err = lib().oracle_rib(ids.ctypes.data, n_part, dim, pts.ctypes.data, n, kind, w.ctypes.data, w.shape[0], 0,
                           int(iter_count), float(tolerance), None, None if m is None else m.ctypes.data)
    return err, ids[:n_part]


# element name -> (dimension, nodes per element), mesh-io/src/lib.rs:24-56
_ELEMENTS = {"edges": (1, 2), "triangles": (2, 3), "quadrilaterals": (2, 4), "quadrangles": (2, 4),
             "tetrahedra": (3, 4), "hexahedra": (3, 8)}


def barycentres(dimension, coordinates, topology):
    """barycentres() of tools/lib/lib.rs:511-539: for the elements of the
    highest element dimension except edges, in topology order, the sum of
    their nodes' coordinates taken in node order from 0.0, divided by the
    node count.  ``topology`` is a list of (type name, 0-based node index
    array (n x k)).  Raises IndexError for a node index out of range."""
    coords = np.asarray(coordinates, dtype=np.float64).reshape(-1, dimension)
    if not topology:
        return np.zeros((0, dimension))
    top = max(_ELEMENTS[name][0] for name, *_ in topology)
    out = []
    for name, nodes, *_ in topology:
        if _ELEMENTS[name][0] != top or name == "edges":
            continue
        nodes = np.asarray(nodes, dtype=np.int64)
        if nodes.size and (nodes.min() < 0 or nodes.max() >= coords.shape[0]):
            raise IndexError("node index out of range")
        bc = np.zeros((nodes.shape[0], dimension))
        for j in range(nodes.shape[1]):
            bc = bc + coords[nodes[:, j]]
        out.append(bc / float(nodes.shape[1]))
    return np.concatenate(out) if out else np.zeros((0, dimension))


def grid_rcb(size, weights, iter_count, threads=16, splits=False):
    """``Grid::rcb`` on the CPU: ids in memory order (x fastest).  With
    ``splits`` also the heap of split positions (index 1 = root)."""
    size = [int(x) for x in size]
    dim = len(size)
    w = np.ascontiguousarray(weights)
    kind = _kind_of(w)
    n = int(np.prod(size))
    part = np.zeros(max(n, 1), dtype=np.uint64)
    sz = (C.c_size_t * 3)(*(size + [1] * (3 - dim)))
    sp = np.zeros(1 << iter_count, dtype=np.int64) if splits and iter_count < 30 else None
    err = lib().oracle_grid_rcb(part.ctypes.data, dim, sz, kind, w.ctypes.data, int(iter_count), int(threads),
                                None if sp is None else sp.ctypes.data)
    if err:
        raise RuntimeError(f"oracle_grid_rcb: error {err}")
    return (part[:n], sp) if splits else part[:n]


def weighted_median(weights, threads=16):
    w = np.ascontiguousarray(weights)
    kind = _kind_of(w)
    pos = C.c_size_t()
    left = np.zeros(1, dtype=np.float64 if kind == W_DOUBLE else np.int64)
    lib().oracle_weighted_median(kind, w.ctypes.data, w.shape[0], int(threads), C.byref(pos), left.ctypes.data)
    return pos.value, left[0]
