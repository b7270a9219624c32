"""Per-part loads and imbalance on the GPU (coupe_hip_parts_load,
coupe_hip_imbalance) against the oracle's restatement of coupe's imbalance
module.  Integer weights: exact.  f64 weights: the reference adds in rayon's
unspecified order, so loads agree to 1e-12 relative."""
import sys

import numpy as np
import pytest

import coupe_amd
from coupe_amd import imbalance as gim

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("num_parts", [1, 7, 4096, 4097, 70000])
@pytest.mark.parametrize("wdt", [np.int32, np.int64, np.float64])
def test_loads_match_oracle(oracle_mod, num_parts, wdt):
    say(f"parts_load k={num_parts} w={np.dtype(wdt).name}")
    rng = np.random.default_rng(num_parts)
    n = 200_000
    ids = rng.integers(0, num_parts, n).astype(np.uint64)
    if wdt == np.float64:
        w = rng.random(n) * 10
    else:
        w = rng.integers(-(1 << 20), 1 << 20, n).astype(wdt)
    got = gim.compute_parts_load(_dev(ids), num_parts, _dev(w))
    exp = oracle_mod.parts_load(ids, num_parts, w)
    assert got.dtype == exp.dtype
    if wdt == np.float64:
        np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-9)
        assert gim.imbalance(num_parts, _dev(ids), _dev(w)) == pytest.approx(
            oracle_mod.imbalance(num_parts, ids, w), rel=1e-9, abs=1e-12)
    else:
        assert np.array_equal(got, exp)
        assert gim.imbalance(num_parts, _dev(ids), _dev(w)) == oracle_mod.imbalance(num_parts, ids, w)
        assert gim.max_imbalance(num_parts, _dev(ids), _dev(w)) == oracle_mod.max_imbalance(num_parts, ids, w)


def test_i32_sums_wrap(oracle_mod):
    say("parts_load i32 wrap")
    ids = np.zeros(1000, dtype=np.uint64)
    w = np.full(1000, 2**30, dtype=np.int32)
    got = gim.compute_parts_load(_dev(ids), 1, _dev(w))
    assert got.tolist() == oracle_mod.parts_load(ids, 1, w).tolist() == [0]   # 1000 * 2^30 = 250 * 2^32


def test_guards_and_errors():
    say("imbalance guards")
    ids = np.array([0, 1, 1], dtype=np.uint64)
    assert gim.imbalance(0, _dev(ids), _dev(np.ones(3))) == 0.0
    assert gim.imbalance(2, _dev(ids), _dev(np.zeros(3))) == 0.0
    empty = np.zeros(0, dtype=np.uint64)
    assert gim.compute_parts_load(_dev(empty), 0, _dev(np.zeros(0, dtype=np.int64))).size == 0
    assert gim.compute_parts_load(_dev(empty), 3, _dev(np.zeros(0))).tolist() == [0.0, 0.0, 0.0]
    with pytest.raises(coupe_amd.CoupeError):   # id 1 >= num_parts: coupe's index panics
        gim.compute_parts_load(_dev(ids), 1, _dev(np.ones(3)))


def test_rcb_partition_imbalance(oracle_mod):
    say("imbalance of an RCB partition, 2^20 points, 64 parts")
    import torch
    from coupe_amd import workloads
    pts, w, _, _ = workloads.standard("gauss3d", scale=4)
    d_pts, d_w = _dev(pts), _dev(w)
    d_ids = torch.zeros(pts.shape[0], dtype=torch.int64, device="cuda")
    coupe_amd.Rcb(iter_count=6, tolerance=0.05).partition(d_ids, (d_pts, d_w))
    ids = d_ids.cpu().numpy().view(np.uint64)
    got = gim.imbalance(64, d_ids, d_w)
    assert got == pytest.approx(oracle_mod.imbalance(64, ids, w), rel=1e-12, abs=1e-15)
    assert 0.0 <= got < 1.0
