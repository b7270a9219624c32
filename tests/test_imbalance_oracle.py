"""The numpy restatement of coupe's imbalance module pinned by closed forms."""
import numpy as np
import pytest


def test_equal_loads_give_zero(oracle_mod):
    ids = np.repeat(np.arange(8, dtype=np.uint64), 5)
    w = np.full(40, 3, dtype=np.int64)
    assert oracle_mod.parts_load(ids, 8, w).tolist() == [15] * 8
    assert oracle_mod.imbalance(8, ids, w) == 0.0
    assert oracle_mod.max_imbalance(8, ids, w) == 0


def test_everything_in_one_part(oracle_mod):
    # ideal = W/k, one part holds W: (W - W/k)/(W/k) = k - 1.
    for k in (1, 2, 7, 64):
        ids = np.zeros(10, dtype=np.uint64)
        w = np.arange(1, 11, dtype=np.float64)
        assert oracle_mod.imbalance(k, ids, w) == pytest.approx(k - 1, rel=1e-15)
        assert oracle_mod.max_imbalance(k, ids, w) == (55.0 if k > 1 else 0.0)


def test_hand_computed(oracle_mod):
    ids = np.array([0, 1, 1, 2, 2, 2], dtype=np.uint64)
    w = np.array([1, 2, 3, 4, 5, 6], dtype=np.int32)
    assert oracle_mod.parts_load(ids, 4, w).tolist() == [1, 5, 15, 0]
    # ideal = 21/4 = 5.25; max is part 2: (15 - 5.25) / 5.25
    assert oracle_mod.imbalance(4, ids, w) == (15 - 5.25) / 5.25
    assert oracle_mod.max_imbalance(4, ids, w) == 15
    assert oracle_mod.imbalance_target([2, 2, 10, 0], ids, w) == 5


def test_bincount_agrees(oracle_mod):
    rng = np.random.default_rng(3)
    ids = rng.integers(0, 37, 5000).astype(np.uint64)
    w = rng.integers(-1000, 1000, 5000).astype(np.int64)
    exp = np.bincount(ids.astype(np.int64), weights=w.astype(np.float64), minlength=37)
    assert oracle_mod.parts_load(ids, 37, w).tolist() == exp.astype(np.int64).tolist()


def test_i32_wraps(oracle_mod):
    ids = np.zeros(3, dtype=np.uint64)
    w = np.array([2**31 - 1, 2**31 - 1, 5], dtype=np.int32)
    assert oracle_mod.parts_load(ids, 1, w).tolist() == [3]   # 2^32 + 3 mod 2^32


def test_guards(oracle_mod):
    assert oracle_mod.imbalance(0, np.zeros(0, dtype=np.uint64), np.zeros(0)) == 0.0
    ids = np.array([0, 1], dtype=np.uint64)
    assert oracle_mod.imbalance(2, ids, np.array([5, -5], dtype=np.int64)) == 0.0   # ideal 0
    with pytest.raises(IndexError):
        oracle_mod.parts_load(np.array([3], dtype=np.uint64), 3, np.ones(1))
