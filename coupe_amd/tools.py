"""mesh-part style runs of the GPU partitioners on real meshes.

Mirrors the part of the coupe tools that feeds RCB (tools/lib/lib.rs):

* ``barycentres(mesh)`` -- the points the tools partition: one per element of
  the mesh's highest element dimension, edges excluded (:511-539), computed
  on the GPU (``coupe_hip_barycentres``).
* ``used_element_count(mesh)`` -- the length of the partition the tools
  allocate (:493-509).
* ``parse_algorithm(spec)`` -- "rcb,ITER[,TOL]" with TOL defaulting to 0.05
  (:418-421); "rib,ITER[,TOL]" runs ``coupe_amd.Rib`` the same way.
* ``mesh_part(mesh, weights, specs)`` -- ``ToRunner for Rcb`` (:213-231): the
  first weight criterion, integers as int64 and floats as float64, then the
  algorithms in order on one partition array (tools/bins/mesh-part.rs:21-78).
* ``part_info(...)`` -- per-criterion totals, imbalance, heaviest and lightest
  parts (tools/bins/part-info.rs:64-153), from GPU per-part loads.

Command line (the reference's mesh-part options that apply here):

    python -m coupe_amd.tools -m mesh.mesh -w weights.mewe -a rcb,12 [out.part]
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys
from typing import List, Sequence

import numpy as np

from . import CoupeError, DeviceContext, Rcb, Rib, _context_for, _native
from . import imbalance as _imb
from .meshio import ELEMENT_TYPES, Mesh, read_mesh, read_partition, read_weights, write_partition  # noqa: F401


def _element_dim(mesh: Mesh):
    if not mesh.topology:
        return None
    return max(ELEMENT_TYPES[name][0] for name, _, _ in mesh.topology)


def used_element_count(mesh: Mesh) -> int:
    """Elements of the highest dimension, edges included (lib.rs:493-509)."""
    d = _element_dim(mesh)
    if d is None:
        return 0
    return sum(int(nodes.shape[0]) for name, nodes, _ in mesh.topology if ELEMENT_TYPES[name][0] == d)


def barycentres(mesh: Mesh, *, device: int = 0, context: DeviceContext = None, stream=None, as_tensor=False):
    """Barycentres of the elements of the highest dimension except edges, in
    topology order (lib.rs:511-539), computed on the GPU.  Returns an
    (n, dimension) float64 array (a device tensor with ``as_tensor``)."""
    import torch

    dim = mesh.dimension
    d = _element_dim(mesh)
    blocks = [] if d is None else [(name, nodes) for name, nodes, _ in mesh.topology
                                   if ELEMENT_TYPES[name][0] == d and name != "edges"]
    total = sum(int(nodes.shape[0]) for _, nodes in blocks)
    dev = torch.device("cuda", device)
    out = torch.empty((total, dim), dtype=torch.float64, device=dev)
    if total:
        coords = torch.from_numpy(np.ascontiguousarray(mesh.coordinates, dtype=np.float64)).to(dev)
        ctx = context or _context_for(device)
        st = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
        off = 0
        for name, nodes in blocks:
            k = ELEMENT_TYPES[name][1]
            d_nodes = torch.from_numpy(np.ascontiguousarray(nodes, dtype=np.int64)).to(dev)
            n = int(nodes.shape[0])
            err = _native.lib().coupe_hip_barycentres(ctx.handle, C.c_void_p(st), coords.data_ptr(),
                                                       mesh.node_count, dim, d_nodes.data_ptr(), n, k,
                                                       out[off:off + n].data_ptr())
            if err != _native.OK:
                raise CoupeError(err)
            off += n
    if as_tensor:
        return out
    return out.cpu().numpy()


def parse_algorithm(spec: str):
    """"rcb,ITER[,TOL]" (tools/lib/lib.rs:418-421) or "rib,ITER[,TOL]"."""
    name, *args = spec.split(",")
    if name not in ("rcb", "rib"):
        raise ValueError(f"unsupported algorithm {name!r} (this build runs rcb and rib)")
    if not args:
        raise ValueError(f"{name}: missing required argument ITER")
    if len(args) > 2:
        raise ValueError(f"{name}: too many arguments")
    iters = int(args[0])
    tol = float(args[1]) if len(args) > 1 else 0.05
    return (Rcb if name == "rcb" else Rib)(iter_count=iters, tolerance=tol)


def _first_criterion(weights: np.ndarray):
    w = np.asarray(weights)
    if w.ndim == 1:
        w = w[:, None]
    if w.shape[1] == 0:
        return np.zeros(w.shape[0], dtype=np.int64)
    col = np.ascontiguousarray(w[:, 0])
    return col.astype(np.int64) if np.issubdtype(col.dtype, np.integer) else col.astype(np.float64)


def mesh_part(mesh: Mesh, weights, specs: Sequence[str], *, device: int = 0) -> np.ndarray:
    """Partition ``mesh`` as mesh-part does: points = barycentres, weights =
    the first criterion, algorithms applied in order to one partition
    array of ``used_element_count`` ids."""
    import torch

    if mesh.dimension not in (2, 3):
        raise ValueError(f"expected 2D or 3D mesh, got a {mesh.dimension}D mesh")
    algos = [parse_algorithm(s) for s in specs]
    n_part = used_element_count(mesh)
    dev = torch.device("cuda", device)
    pts = barycentres(mesh, device=device, as_tensor=True)
    w = torch.from_numpy(_first_criterion(weights)).to(dev)
    part = torch.zeros(n_part, dtype=torch.int64, device=dev)
    for a in algos:
        a.partition(part, (pts, w))
    torch.cuda.synchronize(dev)
    return part.cpu().numpy().astype(np.uint64)


def part_info(part_ids, weights, *, device: int = 0) -> dict:
    """part-info's statistics (part-info.rs:64-153): part count (max id + 1),
    empty parts, and per criterion the total, the imbalance
    max(0, (max load - ideal) / ideal), and the heaviest / lightest loads
    with how many parts share them.  Loads are summed on the GPU."""
    import torch

    ids = np.asarray(part_ids, dtype=np.uint64)
    w = np.asarray(weights)
    if w.ndim == 1:
        w = w[:, None]
    n_parts = int(ids.max()) + 1 if ids.size else 0
    res = {"part_count": n_parts,
           "empty_parts": n_parts - int(np.unique(ids).size) if ids.size else 0, "criteria": []}
    if n_parts == 0 or w.shape[1] == 0:
        return res
    dev = torch.device("cuda", device)
    d_ids = torch.from_numpy(ids.view(np.int64)).to(dev)
    for c in range(w.shape[1]):
        col = _first_criterion(w[:, c:c + 1])
        loads = _imb.compute_parts_load(d_ids, n_parts, torch.from_numpy(col).to(dev))
        total = loads.sum()
        ideal = float(total) / n_parts
        mx, mn = loads.max(), loads.min()
        imb = max(0.0, (float(mx) - ideal) / ideal) if ideal != 0.0 else 0.0
        res["criteria"].append({"total": total.item(), "imbalance": imb, "max_part": mx.item(),
                                "max_part_count": int((loads == mx).sum()), "min_part": mn.item(),
                                "min_part_count": int((loads == mn).sum())})
    return res


def main(argv: List[str] = None) -> int:
    ap = argparse.ArgumentParser(prog="python -m coupe_amd.tools",
                                 description="mesh-part: partition a mesh with the GPU RCB / RIB")
    ap.add_argument("-m", "--mesh", required=True)
    ap.add_argument("-w", "--weights", required=True)
    ap.add_argument("-a", "--algorithm", action="append", default=[])
    ap.add_argument("output", nargs="?", default="-")
    args = ap.parse_args(argv)
    mesh = read_mesh(args.mesh)
    weights = read_weights(args.weights)
    part = mesh_part(mesh, weights, args.algorithm)
    if args.output == "-":
        write_partition(sys.stdout.buffer, part)
    else:
        write_partition(args.output, part)
    return 0


if __name__ == "__main__":
    sys.exit(main())
