"""Seeded RCB inputs shared by the oracle tests and the GPU parity tests."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rcb_vectors.json")
_DT = {"int": np.int32, "int64": np.int64, "double": np.float64}


def golden():
    with open(GOLDEN) as f:
        data = json.load(f)
    out = []
    for v in data["rcb"]:
        pts = np.array(v["points"], dtype=np.float64)
        w = np.array(v["weights"], dtype=_DT[v["weight_type"]])
        out.append((v, pts, w))
    return out, data["aabb"]


def random_case(seed: int, n_max: int = 300):
    """A small input exercising ties, duplicates, clusters, negative values,
    every weight kind and a spread of iter_count / tolerance."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, n_max + 1))
    dim = int(rng.choice([2, 3]))
    shape = rng.integers(0, 6)
    if shape == 0:
        pts = rng.random((n, dim))
    elif shape == 1:   # few distinct values: many ties and duplicates
        pts = rng.integers(0, 5, (n, dim)).astype(np.float64)
    elif shape == 2:   # clusters, negative coordinates
        ctr = rng.normal(0, 10, (3, dim))
        pts = ctr[rng.integers(0, 3, n)] + rng.normal(0, 0.01, (n, dim))
    elif shape == 3:   # wide dynamic range
        pts = rng.normal(0, 1, (n, dim)) * 10.0 ** rng.integers(-3, 4, (n, dim))
    elif shape == 4:   # all points equal
        pts = np.tile(rng.random(dim), (n, 1))
    else:              # dyadic grid with ties on the candidates
        pts = rng.integers(0, 16, (n, dim)).astype(np.float64) / 8.0
    kind = ["int", "int64", "double", "double_real"][int(rng.integers(0, 4))]
    if kind == "int":
        w = rng.integers(-5, 100, n).astype(np.int32)
    elif kind == "int64":
        w = rng.integers(1, 1000, n).astype(np.int64)
    elif kind == "double":
        w = rng.integers(1, 50, n).astype(np.float64)
    else:
        w = rng.random(n) * 10.0
    iters = int(rng.integers(0, 9))
    tol = float(rng.choice([0.0, 0.001, 0.01, 0.05, 0.2, -1.0]))
    return pts, w, iters, tol


def kind_name(w) -> str:
    return {np.dtype(np.int32): "int", np.dtype(np.int64): "int64", np.dtype(np.float64): "double"}[w.dtype]


def integral(w) -> bool:
    return w.dtype != np.float64 or bool(np.all(w == np.trunc(w)))


def grid_closed_form(m_bits: int, dim: int, iter_count: int):
    """Ids of the unit-weight dyadic grid: cell centres (i + 1/2) / 2^m of an
    (2^m)^dim grid, index i0 + n (i1 + n i2), iter_count <= dim * m.  Level l
    splits axis l mod dim at the midpoint of its box, between the two central
    cells, so the id is the interleave of the cell-index bits, MSB first."""
    n = 1 << m_bits
    idx = np.arange(n ** dim, dtype=np.int64)
    coords = [(idx // n ** a) % n for a in range(dim)]
    part = np.zeros(n ** dim, dtype=np.uint64)
    for level in range(iter_count):
        a = level % dim
        bit = (coords[a] >> (m_bits - 1 - level // dim)) & 1
        part |= bit.astype(np.uint64) << np.uint64(iter_count - 1 - level)
    pts = np.stack([(c + 0.5) / n for c in coords], axis=1)
    return pts, part
