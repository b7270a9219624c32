"""Grid::rcb on the GPU (coupe_hip_grid_rcb) against the oracle's restatement,
id for id, for integer weights; at the reference's tests and the closed form
of power-of-two grids."""
import sys

import numpy as np
import pytest

import coupe_amd
from tests.test_grid_oracle import closed_form

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def gpu_grid(size, w, iters, threads):
    part = np.zeros(int(np.prod(size)), dtype=np.uint64)
    coupe_amd.Grid(size).rcb(part, w, iters, threads=threads)
    return part


def test_grid_doctest():
    part = gpu_grid([2, 2], np.ones(4), 2, 16)
    assert sorted(part.tolist()) == [0, 1, 2, 3]


def test_reference_test_3d(oracle_mod):
    p = gpu_grid([4, 4, 4], np.ones(64), 3, 16)
    assert np.array_equal(p, oracle_mod.grid_rcb([4, 4, 4], np.ones(64), 3))
    assert len(set(p.tolist())) == 8


@pytest.mark.parametrize("bits,iters", [((3, 3), 6), ((4, 2), 5), ((3, 3, 3), 9), ((2, 3, 4), 7), ((9, 9), 12),
                                        ((6, 6, 6), 12)])
def test_power_of_two_closed_form(bits, iters):
    size, exp = closed_form(bits, iters)
    say(f"closed form {size} L={iters}")
    for w in (np.ones(int(np.prod(size))), np.full(int(np.prod(size)), 3, dtype=np.int64)):
        assert np.array_equal(gpu_grid(size, w, iters, 16), exp)


CASES = [([37, 53], "i32", 6, 2), ([37, 53], "i64", 7, 3), ([100, 7, 13], "f64int", 8, 16),
         ([256, 256], "i64", 10, 40), ([64, 64, 64], "i32", 12, 16), ([5000, 3], "f64int", 9, 7),
         ([1, 300], "i64", 5, 16), ([3, 1, 700], "i32", 6, 4), ([1000, 1000], "f64int", 12, 16),
         ([129, 65, 33], "i64", 11, 24)]


@pytest.mark.parametrize("size,kind,iters,threads", CASES)
def test_grid_matches_oracle(oracle_mod, size, kind, iters, threads):
    say(f"grid {size} {kind} L={iters} threads={threads}")
    rng = np.random.default_rng(sum(size) + iters)
    n = int(np.prod(size))
    if kind == "i32":
        w = rng.integers(-5, 1000, n).astype(np.int32)
    elif kind == "i64":
        w = rng.integers(0, 1 << 40, n).astype(np.int64)
    else:
        w = rng.integers(0, 100, n).astype(np.float64)
    got = gpu_grid(size, w, iters, threads)
    exp = oracle_mod.grid_rcb(size, w, iters, threads=threads)
    assert np.array_equal(got, exp), f"{(got != exp).sum()} ids differ"


WIDE = [([4500, 40], "i32", 10, 16), ([10000, 300], "f64int", 12, 16), ([2100, 5, 20], "i64", 9, 8),
        ([4097, 33], "f64int", 11, 5), ([9, 2049, 3], "i32", 8, 16), ([100, 9000], "i64", 10, 16),
        ([3, 4, 9000], "i32", 9, 16), ([1100000, 3], "f64int", 8, 16)]


@pytest.mark.parametrize("size,kind,iters,threads", WIDE)
def test_grid_wide_rows_and_z_bands(oracle_mod, size, kind, iters, threads):
    """Rows wider than one tile of the summed-area table (2048 columns),
    more than one band of 16 planes along z, odd and even widths, and tall
    narrow shapes whose scans over bands, segments or planes run in chunks."""
    test_grid_matches_oracle(oracle_mod, size, kind, iters, threads)


@pytest.mark.parametrize("sat", ["1", "0"])
def test_both_paths_agree(oracle_mod, monkeypatch, sat):
    """The summed-area path and the per-level path (COUPE_HIP_GRID_SAT=0)
    give the oracle's ids."""
    monkeypatch.setenv("COUPE_HIP_GRID_SAT", sat)
    for size, kind, iters, threads in (([300, 200], "i64", 9, 16), ([40, 30, 20], "i32", 8, 3),
                                       ([2500, 20], "f64int", 7, 16)):
        test_grid_matches_oracle(oracle_mod, size, kind, iters, threads)


def test_non_integral_weights_take_the_level_path(oracle_mod):
    """Quarter weights are exactly summable in any order, but not integers:
    the summed-area path hands the call to the per-level path."""
    rng = np.random.default_rng(11)
    w = rng.integers(0, 400, 300 * 170).astype(np.float64) / 4.0
    got = gpu_grid([300, 170], w, 10, 16)
    assert np.array_equal(got, oracle_mod.grid_rcb([300, 170], w, 10, threads=16))


def test_deeper_than_the_dense_heap(oracle_mod):
    """iter_count above 20 takes the per-level path (most leaves empty)."""
    rng = np.random.default_rng(12)
    w = rng.integers(1, 9, 48 * 40).astype(np.int64)
    got = gpu_grid([48, 40], w, 22, 16)
    assert np.array_equal(got, oracle_mod.grid_rcb([48, 40], w, 22, threads=16))


def test_no_levels(oracle_mod):
    got = gpu_grid([30, 20], np.ones(600), 0, 16)
    assert not got.any()


def test_i32_sums_wrap(oracle_mod):
    w = np.full(64 * 64, 2**24, dtype=np.int32)   # the total (2^36) wraps in i32
    assert np.array_equal(gpu_grid([64, 64], w, 6, 16), oracle_mod.grid_rcb([64, 64], w, 6, threads=16))


def test_device_tensors_and_reuse(oracle_mod):
    import torch
    rng = np.random.default_rng(3)
    ctx = coupe_amd.DeviceContext(0)
    try:
        for size, iters in (([300, 200], 9), ([50, 40, 30], 7), ([300, 200], 4)):
            n = int(np.prod(size))
            w = rng.integers(1, 50, n).astype(np.int64)
            out = torch.zeros(n, dtype=torch.int64, device="cuda")
            coupe_amd.Grid(size).rcb(out, torch.from_numpy(w).cuda(), iters, threads=16, context=ctx)
            torch.cuda.synchronize()
            exp = oracle_mod.grid_rcb(size, w, iters, threads=16)
            assert np.array_equal(out.cpu().numpy().astype(np.uint64), exp)
    finally:
        ctx.close()


def test_reference_bench_shape_scaled(oracle_mod):
    """The shape of coupe's Grid benchmark (weights i as f64, iter_count 12),
    at 2000 x 2000."""
    say("grid 2000x2000, weights i, L=12")
    n = 2000 * 2000
    w = np.arange(n, dtype=np.float64)
    got = gpu_grid([2000, 2000], w, 12, 16)
    assert np.array_equal(got, oracle_mod.grid_rcb([2000, 2000], w, 12, threads=16))


def test_errors():
    with pytest.raises(ValueError):
        coupe_amd.Grid([0, 4])
    with pytest.raises(ValueError):
        coupe_amd.Grid([4, 4]).rcb(np.zeros(15, dtype=np.uint64), np.ones(15), 2)


def test_sizes_rejected_before_device_work():
    """An axis of 0 or past u32, a cell count past 64 bits and iter_count
    above 62 are refused before any device work (null weights and ids are
    never touched)."""
    import ctypes as C

    from coupe_amd import _native

    ctx = coupe_amd._context_for(0)
    lib = _native.lib()
    for dim, size, iters in [(3, (2**32 - 1, 2**32 - 1, 2**32 - 1), 4), (2, (0, 4, 1), 2),
                             (2, (2**32, 4, 1), 2), (2, (4, 4, 1), 63)]:
        say(f"sizes {size[:dim]} iter_count {iters}")
        sz = (C.c_size_t * 3)(*size)
        err = lib.coupe_hip_grid_rcb(ctx.handle, None, dim, sz, _native.COUPE_DOUBLE, None, iters, 0, None)
        assert err == _native.CRASH
