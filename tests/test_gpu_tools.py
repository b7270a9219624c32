"""mesh-part style runs on the GPU: barycentres (coupe_hip_barycentres)
bit for bit against the oracle, and the whole harness (MEDIT + MeWe in,
MePe out) against the oracle's RCB of the oracle's barycentres."""
import os
import sys

import numpy as np
import pytest

import coupe_amd
from coupe_amd import meshio, tools
from tests import meshes

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def test_barycentres_reference_mesh(oracle_mod):
    say("barycentres of the reference's MEDIT test mesh")
    m = meshio.read_mesh(meshes.REF_PARSE.encode())
    got = tools.barycentres(m)
    exp = oracle_mod.barycentres(3, m.coordinates, [(n, v) for n, v, _ in m.topology])
    assert got.tolist() == exp.tolist()


@pytest.mark.parametrize("m,jitter", [(8, 0.0), (40, 0.3)])
def test_barycentres_hex(oracle_mod, m, jitter):
    say(f"barycentres of a {m}^3 hex mesh")
    coords, hexes = meshes.hex_grid(m, jitter, seed=m)
    mesh = meshio.Mesh(3, coords, np.zeros(len(coords), dtype=np.int64),
                       [("hexahedra", hexes, np.zeros(len(hexes), dtype=np.int64)),
                        ("triangles", hexes[:10, :3], np.zeros(10, dtype=np.int64))])
    got = tools.barycentres(mesh)
    exp = oracle_mod.barycentres(3, coords, [("hexahedra", hexes), ("triangles", hexes[:10, :3])])
    assert np.array_equal(got, exp)


def test_barycentres_2d_blocks_and_errors(oracle_mod):
    say("barycentres of mixed 2D blocks")
    coords, tris, edges = meshes.tri_grid(30, seed=3)
    quads = np.stack([tris[:50, 0], tris[:50, 1], tris[:50, 2], tris[:50, 0]], axis=1)
    topo = [("edges", edges), ("triangles", tris), ("quadrilaterals", quads)]
    mesh = meshio.Mesh(2, coords, np.zeros(len(coords)), [(n, v, np.zeros(len(v))) for n, v in topo])
    assert np.array_equal(tools.barycentres(mesh), oracle_mod.barycentres(2, coords, topo))
    bad = meshio.Mesh(2, coords, np.zeros(len(coords)), [("triangles", tris + len(coords), np.zeros(len(tris)))])
    with pytest.raises(coupe_amd.CoupeError):
        tools.barycentres(bad)


@pytest.mark.parametrize("integer", [True, False])
def test_mesh_part_end_to_end(oracle_mod, tmp_path, integer):
    say(f"mesh-part on a jittered 48^3 hex mesh, integer weights {integer}")
    coords, hexes = meshes.hex_grid(48, 0.25, seed=5)
    text = meshes.to_medit(3, coords, [("Hexahedra", hexes)])
    rng = np.random.default_rng(9)
    if integer:
        w = rng.integers(1, 50, (len(hexes), 2)).astype(np.int64)
    else:
        w = np.round(rng.random((len(hexes), 1)) * 40) + 1.0
    (tmp_path / "m.mesh").write_text(text)
    meshio.write_weights(str(tmp_path / "w.mewe"), w)
    out = tmp_path / "out.part"
    assert tools.main(["-m", str(tmp_path / "m.mesh"), "-w", str(tmp_path / "w.mewe"), "-a", "rcb,9,0.02",
                       str(out)]) == 0
    got = meshio.read_partition(str(out))
    mesh = meshio.read_mesh(str(tmp_path / "m.mesh"))
    pts = oracle_mod.barycentres(3, mesh.coordinates, [(n, v) for n, v, _ in mesh.topology])
    err, exp, _ = oracle_mod.rcb(pts, w[:, 0], 9, 0.02)
    assert err == 0 and np.array_equal(got, exp)
    info = tools.part_info(got, w)
    assert info["part_count"] == 512 and info["empty_parts"] == 0
    loads = np.bincount(got.astype(np.int64), weights=w[:, 0].astype(np.float64), minlength=512)
    assert info["criteria"][0]["max_part"] == loads.max() and info["criteria"][0]["min_part"] == loads.min()
    ideal = loads.sum() / 512
    assert info["criteria"][0]["imbalance"] == pytest.approx(max(0.0, (loads.max() - ideal) / ideal), rel=1e-12)


def test_mesh_part_rib_chain(oracle_mod, tmp_path):
    say("mesh-part with rcb then rib on one partition")
    coords, tris, edges = meshes.tri_grid(60, seed=2)
    mesh = meshio.Mesh(2, coords, np.zeros(len(coords)), [("triangles", tris, np.zeros(len(tris))),
                                                          ("edges", edges, np.zeros(len(edges)))])
    w = np.ones((len(tris), 1), dtype=np.int64)
    got = tools.mesh_part(mesh, w, ["rcb,3", "rib,5,0.01"])
    assert got.shape == (len(tris),) and int(got.max()) < 32
