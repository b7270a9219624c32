"""Test meshes: the reference's MEDIT test inputs and structured meshes."""
import numpy as np

# mesh-io/src/medit/parser.rs:580-593 (test_parse)
REF_PARSE = """MeshVersionFormatted 1
        Dimension 3
        Vertices
        4
        2.3 0.0 1 0
        1231 2.00 3.14 0
        -21.2 21 0.0001 0
        -0.2 -0.2 -0.2 0
        Triangles
        2
        1 2 3 0
        2 3 4 1
        End
        """

# mesh-io/src/medit/serializer.rs:144-159 (test_serialize)
REF_SERIALIZE = ("MeshVersionFormatted 2\nDimension 3\n\nVertices\n\t4\n 2.3 0 1 0\n 1231 2 3.14 0\n"
                 " -21.2 21 0.0001 0\n -0.2 -0.2 -0.2 0\n\nTriangles\n\t2\n 1 2 3 0\n 2 3 4 0\n\nEnd")


def hex_grid(m, jitter=0.0, seed=0):
    """m^3 hexahedra on an (m+1)^3 node lattice; nodes jittered by up to
    `jitter` cell widths."""
    rng = np.random.default_rng(seed)
    g = np.arange(m + 1, dtype=np.float64)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    coords = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1) / m
    coords += rng.uniform(-jitter, jitter, coords.shape) / m
    nid = lambda i, j, k: (i * (m + 1) + j) * (m + 1) + k   # noqa: E731
    i, j, k = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    hexes = np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                      nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)], axis=1)
    return coords, hexes


def tri_grid(m, seed=0):
    """2 m^2 triangles on an (m+1)^2 lattice plus its boundary edges."""
    rng = np.random.default_rng(seed)
    g = np.arange(m + 1, dtype=np.float64)
    x, y = np.meshgrid(g, g, indexing="ij")
    coords = np.stack([x.ravel(), y.ravel()], axis=1) * 0.1 + rng.normal(0, 1e-3, ((m + 1) ** 2, 2))
    nid = lambda i, j: i * (m + 1) + j   # noqa: E731
    i, j = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    i, j = i.ravel(), j.ravel()
    t1 = np.stack([nid(i, j), nid(i + 1, j), nid(i + 1, j + 1)], axis=1)
    t2 = np.stack([nid(i, j), nid(i + 1, j + 1), nid(i, j + 1)], axis=1)
    edges = np.stack([np.arange(m), np.arange(1, m + 1)], axis=1)
    return coords, np.concatenate([t1, t2]), edges


def to_medit(dim, coords, blocks):
    """ASCII MEDIT text: blocks = [(keyword, 0-based nodes)]."""
    out = [f"MeshVersionFormatted 2\nDimension {dim}\nVertices\n{coords.shape[0]}"]
    out += [" ".join(repr(float(c)) for c in row) + " 0" for row in coords]
    for name, nodes in blocks:
        out.append(f"{name}\n{nodes.shape[0]}")
        out += [" ".join(str(int(v) + 1) for v in row) + " 1" for row in nodes]
    out.append("End\n")
    return "\n".join(out)
