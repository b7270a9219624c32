"""MEDIT / MeWe / MePe formats and the tools' element selection, on the CPU."""
import struct

import numpy as np
import pytest

from coupe_amd import meshio, tools
from tests import meshes


def test_reference_parse_case():
    m = meshio.read_mesh(meshes.REF_PARSE.encode())
    assert m.dimension == 3
    assert m.coordinates.tolist() == [[2.3, 0.0, 1.0], [1231.0, 2.0, 3.14], [-21.2, 21.0, 0.0001], [-0.2, -0.2, -0.2]]
    assert m.node_refs.tolist() == [0, 0, 0, 0]
    (name, nodes, refs), = m.topology
    assert name == "triangles" and nodes.tolist() == [[0, 1, 2], [1, 2, 3]] and refs.tolist() == [0, 1]


def test_reference_serialize_case_parses():
    m = meshio.parse_ascii(meshes.REF_SERIALIZE)
    assert m.node_count == 4 and m.element_count == 2
    assert m.topology[0][1].tolist() == [[0, 1, 2], [1, 2, 3]]


def test_ascii_details():
    text = ("  meshversionformatted 1\nDIMENSION 2\nVertices 3\n0 0 7\n1 0 7\n0 1\n"
            "Corners\n1\n1\nTriangles junk 1\n1 2 3 5 extra\nEdges\n1\n1 2 0\nEnd\n")
    m = meshio.parse_ascii(text)
    assert m.node_refs.tolist() == [7, 7]
    assert [t[0] for t in m.topology] == ["triangles", "edges"]
    assert m.topology[0][2].tolist() == [5]
    with pytest.raises(meshio.FormatError):
        meshio.parse_ascii("MeshVersionFormatted 1\nDimension 2\nVertices\n1\n0 0 0 9\nEnd\n")
    with pytest.raises(meshio.FormatError):
        meshio.parse_ascii("MeshVersionFormatted 1\nDimension 2\nPolygons\n1\nEnd\n")
    with pytest.raises(meshio.FormatError):
        meshio.read_mesh(b"Dimension 2\n")


@pytest.mark.parametrize("version,end", [(1, "<"), (2, "<"), (3, ">"), (4, "<")])
def test_binary(version, end):
    fi, ff, fp = {1: ("i", "f", "i"), 2: ("i", "d", "i"), 3: ("i", "d", "q"), 4: ("q", "d", "q")}[version]
    p = lambda f, *v: struct.pack(end + f, *v)   # noqa: E731
    b = p("i", 1) + p("i", version) + p("i", 3) + p(fp, 0) + p("i", 3)
    coords = [(0.5, 1.0, 2.0), (1.5, 1.0, 2.0), (0.5, 3.0, 2.0), (0.5, 1.0, 6.0)]
    b += p("i", 4) + p(fp, 0) + p(fi, 4) + b"".join(p(ff * 3 + fi, *c, 3) for c in coords)
    b += p("i", 8) + p(fp, 0) + p(fi, 1) + p(fi * 5, 1, 2, 3, 4, 9)
    b += p("i", 54)
    m = meshio.read_mesh(b)
    assert m.dimension == 3 and m.coordinates.tolist() == [list(c) for c in coords]
    assert m.node_refs.tolist() == [3] * 4
    assert m.topology[0][0] == "tetrahedra" and m.topology[0][1].tolist() == [[0, 1, 2, 3]]
    assert m.topology[0][2].tolist() == [9]


def test_weights_format():
    w = np.array([[3], [-1], [7]], dtype=np.int64)
    raw = meshio.write_weights(None, w)
    assert raw == b"MeWe\x01\x01" + struct.pack("<HQ", 1, 3) + struct.pack("<3q", 3, -1, 7)
    assert meshio.read_weights(raw).tolist() == w.tolist()
    f = np.array([[0.5, 2.0], [1.25, -3.0]])
    got = meshio.read_weights(meshio.write_weights(None, f))
    assert got.dtype == np.float64 and got.tolist() == f.tolist()
    assert meshio.write_weights(None, np.zeros((0, 1), dtype=np.int64)) == b"MeWe\x01\x01" + bytes(10)
    assert meshio.read_weights(b"MeWe\x01\x00\x00\x00").shape == (0, 0)
    with pytest.raises(meshio.FormatError):
        meshio.read_weights(b"MeWx\x01\x00\x00\x00")
    with pytest.raises(meshio.FormatError):
        meshio.read_weights(b"MeWe\x02\x00\x01\x00" + bytes(8))


def test_partition_format():
    raw = meshio.write_partition(None, [3, 0, 1])
    assert raw == b"MePe" + struct.pack("<Q", 3) + struct.pack("<3Q", 3, 0, 1)
    assert meshio.read_partition(raw).tolist() == [3, 0, 1]
    with pytest.raises(meshio.FormatError):
        meshio.read_partition(b"MeWe" + bytes(8))


def test_algorithm_specs():
    a = tools.parse_algorithm("rcb,12")
    assert (type(a).__name__, a.iter_count, a.tolerance) == ("Rcb", 12, 0.05)
    b = tools.parse_algorithm("rib,3,0.2")
    assert (type(b).__name__, b.iter_count, b.tolerance) == ("Rib", 3, 0.2)
    for bad in ("rcb", "kmeans,3", "rcb,1,2,3"):
        with pytest.raises(ValueError):
            tools.parse_algorithm(bad)


def test_used_element_count():
    coords, tris, edges = meshes.tri_grid(4)
    m = meshio.Mesh(2, coords, np.zeros(len(coords), dtype=np.int64),
                    [("triangles", tris, np.zeros(len(tris))), ("edges", edges, np.zeros(len(edges)))])
    assert tools.used_element_count(m) == 32
    m1 = meshio.Mesh(2, coords, np.zeros(len(coords)), [("edges", edges, np.zeros(len(edges)))])
    assert tools.used_element_count(m1) == 4   # edges count, but have no barycentre


def test_oracle_barycentres(oracle_mod):
    m = meshio.read_mesh(meshes.REF_PARSE.encode())
    bc = oracle_mod.barycentres(3, m.coordinates, [(n, v) for n, v, _ in m.topology])
    exp = [[((0.0 + 2.3) + 1231.0 + -21.2) / 3.0, ((0.0 + 0.0) + 2.0 + 21.0) / 3.0, ((0.0 + 1.0) + 3.14 + 0.0001) / 3.0],
           [((0.0 + 1231.0) + -21.2 + -0.2) / 3.0, ((0.0 + 2.0) + 21.0 + -0.2) / 3.0, ((0.0 + 3.14) + 0.0001 + -0.2) / 3.0]]
    assert bc.tolist() == exp
    coords, hexes = meshes.hex_grid(4)
    bc = oracle_mod.barycentres(3, coords, [("hexahedra", hexes), ("triangles", hexes[:, :3])])
    assert bc.shape == (64, 3)
    np.testing.assert_allclose(bc[0], [0.125, 0.125, 0.125], rtol=0, atol=1e-15)
    assert bc.min() > 0 and bc.max() < 1
    coords2, tris, edges = meshes.tri_grid(3)
    assert oracle_mod.barycentres(2, coords2, [("edges", edges)]).shape == (0, 2)
    with pytest.raises(IndexError):
        oracle_mod.barycentres(2, coords2, [("triangles", tris + 100)])
