"""The CPU restatement of coupe's Grid::rcb (oracle/grid_oracle.c) pinned by
the reference's doctest and tests and by the closed form of power-of-two
grids."""
import numpy as np
import pytest


def closed_form(bits, iters):
    """Unit weights on a grid of 2^bits[k] cells per axis: every split halves
    the axis at a power of two, first axis 1, then cycling."""
    dim = len(bits)
    size = [1 << b for b in bits]
    idx = np.arange(int(np.prod(size)), dtype=np.int64)
    pos = [idx % size[0], (idx // size[0]) % size[1]] + ([idx // size[0] // size[1]] if dim == 3 else [])
    ids = np.zeros(idx.size, dtype=np.int64)
    used = [0] * dim
    for level in range(iters):
        a = (1 + level) % dim
        if used[a] >= bits[a]:
            # one cell wide: the median of one value is position 0, the
            # left side is empty and every cell goes right
            bit = np.ones_like(ids)
        else:
            bit = (pos[a] >> (bits[a] - 1 - used[a])) & 1
        used[a] += 1
        ids = 2 * ids + bit
    return size, ids.astype(np.uint64)


def test_grid_doctest(oracle_mod):
    # coupe/src/cartesian/mod.rs:21-41
    p = oracle_mod.grid_rcb([2, 2], np.ones(4), 2)
    assert sorted(p.tolist()) == [0, 1, 2, 3]


def test_reference_test_3d(oracle_mod):
    # coupe/src/cartesian/rcb.rs:316-361
    n = 4
    p = oracle_mod.grid_rcb([n, n, n], np.ones(64), 3)
    idx = lambda x, y, z: x + n * (y + n * z)   # noqa: E731
    for off in [(0, 0, 0), (2, 0, 0), (0, 2, 0), (2, 2, 0), (0, 0, 2), (2, 0, 2), (0, 2, 2), (2, 2, 2)]:
        x, y, z = off
        block = {p[idx(x + a, y + b, z + c)] for a in (0, 1) for b in (0, 1) for c in (0, 1)}
        assert len(block) == 1
    assert len(set(p.tolist())) == 8


@pytest.mark.parametrize("seed", range(30))
@pytest.mark.parametrize("threads", [2, 3, 16, 40])
def test_weighted_median_property(oracle_mod, seed, threads):
    # rcb.rs:300-314: the left weight is the sum of the weights before the position
    rng = np.random.default_rng(seed)
    w = rng.integers(0, 1_000_000, int(rng.integers(2, 200))).astype(np.int64)
    pos, left = oracle_mod.weighted_median(w, threads)
    assert left == w[:pos].sum()
    assert 0 <= pos < w.size


@pytest.mark.parametrize("bits,iters", [((3, 3), 6), ((4, 2), 5), ((3, 3, 3), 9), ((2, 3, 4), 7), ((5, 5), 3)])
def test_power_of_two_closed_form(oracle_mod, bits, iters):
    size, exp = closed_form(bits, iters)
    for w in (np.ones(int(np.prod(size))), np.full(int(np.prod(size)), 7, dtype=np.int32)):
        assert np.array_equal(oracle_mod.grid_rcb(size, w, iters), exp)


def test_threads_change_the_result_only_through_chunks(oracle_mod):
    rng = np.random.default_rng(4)
    w = rng.integers(1, 100, 60 * 50).astype(np.int64)
    a = oracle_mod.grid_rcb([60, 50], w, 6, threads=2)
    b = oracle_mod.grid_rcb([60, 50], w, 6, threads=2)
    assert np.array_equal(a, b)
    ids = {t: oracle_mod.grid_rcb([60, 50], w, 6, threads=t) for t in (2, 5, 16)}
    for t, p in ids.items():
        assert p.max() < 64
        counts = np.bincount(p.astype(np.int64), weights=w.astype(np.float64), minlength=64)
        assert counts.min() > 0


def test_i32_sums_wrap(oracle_mod):
    w = np.full(16, 2**30, dtype=np.int32)   # total wraps to 0 in i32
    p = oracle_mod.grid_rcb([4, 4], w, 2)
    assert p.shape == (16,)
