"""The CPU restatement of coupe's Rib (oracle/rib_oracle.c) pinned by the
reference's geometry tests, the Rib doctest, LAPACK and invariants."""
import numpy as np
import pytest



def _parallel(u, v, tol):
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    return abs(abs(u @ v) - np.linalg.norm(u) * np.linalg.norm(v)) <= tol


def test_inertia_matrix_2d(oracle_mod):
    # coupe/src/geometry.rs:363-375
    f = oracle_mod.rib_frame(np.array([[3., 0.], [0., 3.], [6., -3.]]))
    assert f["inertia"].tolist() == [[18., -18.], [-18., 18.]]
    assert f["centroid"].tolist() == [3., 0.]


def test_inertia_vector_2d(oracle_mod):
    # geometry.rs:377-393: parallel to (1, -1)
    f = oracle_mod.rib_frame(np.array([[3., 0.], [0., 3.], [6., -3.]]))
    assert _parallel(f["axis"], [1., -1.], 1e-15)
    assert np.linalg.norm(f["axis"]) == pytest.approx(1.0, abs=1e-15)


def test_inertia_vector_3d(oracle_mod):
    # geometry.rs:505-518: cross((1,-1,-1), v) == 0 to 1e-15
    f = oracle_mod.rib_frame(np.array([[3., 0., 0.], [0., 3., 3.], [6., -3., -3.]]))
    v = f["axis"]
    assert np.linalg.norm(np.cross([1., -1., -1.], v)) <= 1e-15 * np.sqrt(3.0)


def test_obb_center(oracle_mod):
    # geometry.rs:395-408: the centre of the oriented box is (3, 3)
    pts = np.array([[0., 1.], [1., 0.], [5., 6.], [6., 5.]])
    f = oracle_mod.rib_frame(pts)
    mapped = oracle_mod.map_points(pts, f["obb_to_aabb"])
    centre = 0.5 * (mapped.min(axis=0) + mapped.max(axis=0))
    back = f["aabb_to_obb"] @ centre
    np.testing.assert_allclose(back, [3., 3.], rtol=0, atol=4 * np.finfo(float).eps * 6)


@pytest.mark.parametrize("seed", range(8))
def test_householder_reflection(oracle_mod, seed):
    # geometry.rs:460-487 (there in 6-D; 2-D and 3-D here): orthonormal
    # columns, first column parallel to the vector.
    rng = np.random.default_rng(seed)
    for dim in (2, 3):
        e = rng.standard_normal(dim)
        h = oracle_mod.householder(e)
        for c in range(dim):
            assert np.linalg.norm(h[:, c]) == pytest.approx(1.0, abs=4e-16 * dim)
            for c2 in range(dim):
                if c2 != c:
                    assert abs(h[:, c] @ h[:, c2]) <= 1e-14
        u = e / np.linalg.norm(e)
        np.testing.assert_allclose(u * (u @ h[:, 0]), h[:, 0], atol=1e-14)


def test_householder_identity_cases(oracle_mod):
    assert oracle_mod.householder(np.array([2.0, 0.0, 0.0])).tolist() == np.eye(3).tolist()
    assert oracle_mod.householder(np.array([1.0, 1e-17])).tolist() == np.eye(2).tolist()


@pytest.mark.parametrize("seed", range(20))
def test_axis_matches_lapack(oracle_mod, seed):
    rng = np.random.default_rng(100 + seed)
    dim = 2 + seed % 2
    a = rng.standard_normal((dim, dim)) * 10 ** rng.uniform(-3, 3)
    m = a @ a.T
    v, lam = oracle_mod.inertia_vector(m)
    w, vec = np.linalg.eigh(m)
    gap = (w[-1] - w[-2]) / max(abs(w[-1]), 1e-300)
    if gap < 1e-6:
        pytest.skip("nearly repeated eigenvalue")
    assert lam == pytest.approx(w[-1], rel=1e-12)
    assert _parallel(v, vec[:, -1], 1e-10 / gap)


def test_diagonal_matrix_takes_first_largest_axis(oracle_mod):
    v, _ = oracle_mod.inertia_vector(np.diag([2.0, 5.0, 5.0]))
    assert v.tolist() == [0.0, 1.0, 0.0]
    v, _ = oracle_mod.inertia_vector(np.diag([4.0, 4.0]))
    assert v.tolist() == [1.0, 0.0]


@pytest.mark.parametrize("seed", range(6))
def test_inverse_and_map(oracle_mod, seed):
    rng = np.random.default_rng(seed)
    pts = rng.standard_normal((500, 2 + seed % 2)) * [5.0, 1.0, 0.3][: 2 + seed % 2]
    f = oracle_mod.rib_frame(pts)
    d = pts.shape[1]
    np.testing.assert_allclose(f["aabb_to_obb"] @ f["obb_to_aabb"], np.eye(d), atol=1e-15 * 8)
    np.testing.assert_allclose(f["obb_to_aabb"], f["aabb_to_obb"], atol=1e-15 * 8)   # a reflection
    np.testing.assert_allclose(oracle_mod.map_points(pts, f["obb_to_aabb"]), pts @ f["obb_to_aabb"].T,
                               rtol=0, atol=1e-13)
    # the mapped cloud's first axis carries the largest spread
    mapped = oracle_mod.map_points(pts, f["obb_to_aabb"])
    var = mapped.var(axis=0)
    assert var[0] == var.max()


def test_rib_doctest(oracle_mod):
    # recursive_bisection.rs:866-894
    pts = np.array([[1., 10.], [-1., 10.], [1., -10.], [-1., -10.]])
    err, ids = oracle_mod.rib(pts, np.ones(4, dtype=np.int32), 1, 0.0)
    assert err == 0
    assert ids[0] == ids[1] and ids[2] == ids[3] and ids[1] != ids[2]


def test_rib_is_rcb_of_mapped_points(oracle_mod):
    rng = np.random.default_rng(7)
    pts = rng.standard_normal((3000, 3)) @ np.array([[3., 1., 0.], [0., 1., 0.5], [0., 0., 0.2]])
    w = rng.integers(1, 5, 3000).astype(np.int64)
    err, ids = oracle_mod.rib(pts, w, 5, 0.05)
    assert err == 0
    f = oracle_mod.rib_frame(pts)
    err2, ids2, _ = oracle_mod.rcb(oracle_mod.map_points(pts, f["obb_to_aabb"]), w, 5, 0.05)
    assert err2 == 0 and np.array_equal(ids, ids2)
    err3, ids3 = oracle_mod.rib(pts, w, 5, 0.05, obb_to_aabb=f["obb_to_aabb"])
    assert err3 == 0 and np.array_equal(ids, ids3)


def test_first_split_is_across_the_long_axis(oracle_mod):
    # A cloud stretched along u: Rib's first split separates the points by
    # their projection on u (up to f32 rounding of the mapped coordinate).
    rng = np.random.default_rng(11)
    u = np.array([0.6, 0.8])
    t = rng.uniform(-50, 50, 4001)
    pts = np.outer(t, u) + rng.standard_normal((4001, 2)) * 0.5 @ np.array([[0., 0.], [-0.8, 0.6]])
    err, ids = oracle_mod.rib(pts, np.ones(4001, dtype=np.int64), 1, 0.01)
    assert err == 0
    proj = pts @ u
    a, b = proj[ids == 0], proj[ids == 1]
    assert a.size and b.size
    lo, hi = (a, b) if a.mean() < b.mean() else (b, a)
    assert lo.max() <= hi.min() + 1e-4


def test_rib_empty_and_errors(oracle_mod):
    err, ids = oracle_mod.rib(np.zeros((0, 2)), np.zeros(0, dtype=np.int64), 3, 0.1)
    assert err == 0
    err, _ = oracle_mod.rib(np.zeros((3, 4)), np.ones(3, dtype=np.int64), 3, 0.1)
    assert err == 3
