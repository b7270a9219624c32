"""Parity of the HIP path with the oracle, through the C ABI.

Integer and integer-valued weights: ids bit for bit.  Non-integer f64
weights: ids equal unless a split decision lies within rounding of its
threshold (then per-part weights must agree to 1e-12 relative).
"""
import ctypes as C
import sys

import numpy as np
import pytest

import coupe_amd
from coupe_amd import _native, workloads
from tests import cases

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def gpu_rcb(pts, w, iters, tol):
    part = np.zeros(pts.shape[0], dtype=np.uint64)
    coupe_amd.Rcb(iter_count=iters, tolerance=tol).partition(part, (pts, w))
    return part


def oracle_rcb(oracle, pts, w, iters, tol, parallel=False):
    err, ids, st = oracle.rcb(pts, w, iters, tol, parallel=parallel, nthreads=16 if parallel else 0)
    assert err == 0
    return ids, st


def part_weights(ids, w):
    k = int(ids.max()) + 1 if ids.size else 0
    return np.bincount(ids.astype(np.int64), weights=w.astype(np.float64), minlength=k)


def assert_parity(oracle, pts, w, iters, tol, label, parallel_oracle=False):
    say(f"{label}: n={pts.shape[0]} d={pts.shape[1]} L={iters} tol={tol} w={w.dtype}")
    got = gpu_rcb(pts, w, iters, tol)
    exp, st = oracle_rcb(oracle, pts, w, iters, tol, parallel=parallel_oracle)
    if cases.integral(w):
        if not np.array_equal(got, exp):
            bad = np.nonzero(got != exp)[0]
            raise AssertionError(f"{label}: {bad.size} ids differ, first at {bad[:5]}: "
                                 f"gpu {got[bad[:5]]} oracle {exp[bad[:5]]}")
    else:
        if not np.array_equal(got, exp):
            a, b = part_weights(got, w), part_weights(exp, w)
            assert a.shape == b.shape, label
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=0, err_msg=label)
    return got


def test_golden_vectors(oracle_mod):
    vecs, _ = cases.golden()
    L = _native.lib()
    for v, pts, w in vecs:
        say(v["id"])
        n = pts.shape[0]
        part = np.zeros(n, dtype=np.uint64)
        dp = L.coupe_data_array(n, _native.COUPE_DOUBLE, pts.ctypes.data)
        wt = {np.dtype(np.int32): 0, np.dtype(np.int64): 1, np.dtype(np.float64): 2}[w.dtype]
        dw = (L.coupe_data_constant(n, wt, w.ctypes.data) if v["constant"]
              else L.coupe_data_array(n, wt, w.ctypes.data))
        err = L.coupe_rcb(part.ctypes.data, 2, dp, dw, v["iter_count"], v["tolerance"])
        L.coupe_data_free(dp)
        L.coupe_data_free(dw)
        assert err == 0
        assert part.tolist() == v["expected"], v["id"]


def test_callback_data_sets():
    # coupe_data_fn for points and weights (coupe-ffi/src/data.rs:97-144).
    L = _native.lib()
    pts = np.random.default_rng(3).random((500, 3))
    w = np.arange(1, 501, dtype=np.int64)
    keep = []
    p_cb = _native.ITH_FN(lambda ctx, i: pts.ctypes.data + 24 * i)
    w_cb = _native.ITH_FN(lambda ctx, i: w.ctypes.data + 8 * i)
    keep += [p_cb, w_cb]
    dp = L.coupe_data_fn(None, 500, _native.COUPE_DOUBLE, p_cb)
    dw = L.coupe_data_fn(None, 500, _native.COUPE_INT64, w_cb)
    part = np.zeros(500, dtype=np.uint64)
    assert L.coupe_rcb(part.ctypes.data, 3, dp, dw, 5, 0.02) == 0
    L.coupe_data_free(dp)
    L.coupe_data_free(dw)
    assert np.array_equal(part, gpu_rcb(pts, w, 5, 0.02))


def test_host_entry_staged_in_chunks(oracle_mod):
    """coupe_rcb on inputs of several pinned chunks: arrays, a constant data
    set and a callback data set (called from several host threads), ids equal
    to the oracle's."""
    L = _native.lib()
    rng = np.random.default_rng(33)
    n = 3_000_000   # 72 MB of points: three chunks and a half
    pts = rng.random((n, 3))
    w = rng.integers(1, 1000, n).astype(np.int64)
    part = np.zeros(n, dtype=np.uint64)
    say("host entry, arrays in chunks")
    dp = L.coupe_data_array(n, _native.COUPE_DOUBLE, pts.ctypes.data)
    dw = L.coupe_data_array(n, _native.COUPE_INT64, w.ctypes.data)
    assert L.coupe_rcb(part.ctypes.data, 3, dp, dw, 11, 0.01) == 0
    L.coupe_data_free(dp)
    L.coupe_data_free(dw)
    exp, _ = oracle_rcb(oracle_mod, pts, w, 11, 0.01, parallel=True)
    assert np.array_equal(part, exp)
    # a callback for the points, a constant for the weights
    m = 400_000
    sub = np.ascontiguousarray(pts[:m])
    p_cb = _native.ITH_FN(lambda ctx, i: sub.ctypes.data + 24 * i)
    one = np.array([3], dtype=np.int32)
    say("host entry, callback points and constant weights")
    dp = L.coupe_data_fn(None, m, _native.COUPE_DOUBLE, p_cb)
    dw = L.coupe_data_constant(m, _native.COUPE_INT, one.ctypes.data)
    part = np.zeros(m, dtype=np.uint64)
    assert L.coupe_rcb(part.ctypes.data, 3, dp, dw, 9, 0.02) == 0
    L.coupe_data_free(dp)
    L.coupe_data_free(dw)
    exp, _ = oracle_rcb(oracle_mod, sub, np.full(m, 3, dtype=np.int32), 9, 0.02)
    assert np.array_equal(part, exp)


@pytest.mark.parametrize("seed", range(160))
def test_random_small_cases(oracle_mod, seed):
    pts, w, iters, tol = cases.random_case(seed, n_max=300)
    assert_parity(oracle_mod, pts, w, iters, tol, f"random seed {seed}")


@pytest.mark.parametrize("seed", range(12))
def test_random_medium_cases(oracle_mod, seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(10_000, 300_000))
    dim = 2 + seed % 2
    if seed % 3 == 0:
        pts = rng.random((n, dim))
    elif seed % 3 == 1:
        pts = rng.normal(0, 1, (n, dim)) ** 3
    else:
        pts = np.round(rng.random((n, dim)) * 50) / 50
    w = rng.integers(1, 1000, n).astype([np.int32, np.int64, np.float64][seed % 3])
    assert_parity(oracle_mod, pts, w, int(rng.integers(3, 14)), float(rng.choice([0.0, 0.01, 0.05])),
                  f"medium seed {seed}")


def test_consecutive_calls_on_one_context(oracle_mod):
    """A call must not see per-tile state left by the previous, deeper one."""
    torch = pytest.importorskip("torch")
    ctx = coupe_amd.DeviceContext(0)
    try:
        rng = np.random.default_rng(77)
        runs = [(rng.random((400_000, 3)), 12, 0.01), (rng.random((221_239, 3)), 9, 0.05),
                (rng.random((150_000, 2)), 11, 0.0), (rng.random((90_000, 3)), 5, 0.05)]
        for q, (pts, iters, tol) in enumerate(runs):
            w = rng.integers(1, 1000, pts.shape[0]).astype(np.int32)
            say(f"consecutive call {q}: n={pts.shape[0]} L={iters}")
            out = torch.zeros(pts.shape[0], dtype=torch.int64, device="cuda")
            coupe_amd.Rcb(iter_count=iters, tolerance=tol).partition(
                out, (torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda()), context=ctx)
            torch.cuda.synchronize()
            exp, _ = oracle_rcb(oracle_mod, pts, w, iters, tol)
            assert np.array_equal(out.cpu().numpy().astype(np.uint64), exp), f"call {q}"
    finally:
        ctx.close()


def test_weight_modes_change_between_calls(oracle_mod):
    """A context takes its last call's weight mode for the next call and reads
    the prologue's flags at the end; a call whose weights do not fit that
    mode (16-bit after wide integers, integers after fractional doubles, ...)
    runs again in its own mode.  Ids equal the oracle's on every call."""
    torch = pytest.importorskip("torch")
    ctx = coupe_amd.DeviceContext(0)
    try:
        rng = np.random.default_rng(1700)
        kinds = [("u16", np.int64), ("wide", np.int64), ("u16", np.float64), ("frac", np.float64),
                 ("u16", np.int32), ("neg", np.int32), ("wide", np.int64), ("u16", np.int64)]
        for q, (kind, dt) in enumerate(kinds):
            n = int(rng.integers(20_000, 120_000))
            pts = rng.random((n, 3))
            if kind == "u16":
                w = rng.integers(0, 65536, n).astype(dt)
            elif kind == "wide":
                w = rng.integers(1, 1 << 33, n).astype(dt)
            elif kind == "neg":
                w = rng.integers(-50, 50, n).astype(dt)
            else:
                w = rng.random(n) * 10.0 + 0.5
            iters, tol = 9, 0.02
            own = {"u16": 2, "wide": 8, "neg": 4, "frac": 8}[kind]
            exp, _ = oracle_rcb(oracle_mod, pts, w, iters, tol)
            # twice: the first call may keep a wider integer mode from the
            # previous data (ids are the same); the second uses the data's own
            for rep in range(2):
                say(f"weight mode change, call {q}.{rep}: {kind} {np.dtype(dt).name} n={n}")
                out = torch.zeros(n, dtype=torch.int64, device="cuda")
                coupe_amd.Rcb(iter_count=iters, tolerance=tol).partition(
                    out, (torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda()), context=ctx)
                torch.cuda.synchronize()
                got = out.cpu().numpy().astype(np.uint64)
                if kind == "frac":
                    a, b = part_weights(got, w), part_weights(exp, w)
                    np.testing.assert_allclose(a, b, rtol=1e-12, atol=0, err_msg=f"call {q}.{rep}")
                else:
                    assert np.array_equal(got, exp), f"call {q}.{rep} ({kind})"
                st = ctx.stats()
                if rep == 0:
                    assert st["weight_bytes"] >= own, (q, kind, st)
                else:
                    assert st["weight_bytes"] == own, (q, kind, st)
    finally:
        ctx.close()


@pytest.mark.parametrize("queued", [0, 1, 3])
def test_split_search_queue_depths(oracle_mod, monkeypatch, queued):
    """Split searches queued without host checks (COUPE_HIP_SEARCH_PASSES):
    a search that needs more passes than were queued makes the call run
    again with a host check after every pass; ids never change."""
    torch = pytest.importorskip("torch")
    monkeypatch.setenv("COUPE_HIP_SEARCH_PASSES", str(queued))
    ctx = coupe_amd.DeviceContext(0)
    try:
        rng = np.random.default_rng(500 + queued)
        reruns = 0
        for q in range(6):
            n = int(rng.integers(20_000, 200_000))
            dim = 2 + q % 2
            pts = rng.normal(0, 1, (n, dim)) ** 3 if q % 2 else rng.random((n, dim))
            w = rng.integers(1, 1000, n).astype(np.int64)
            iters, tol = int(rng.integers(4, 12)), [0.0, 0.001][q % 2]
            say(f"queued {queued}, case {q}: n={n} d={dim} L={iters} tol={tol}")
            out = torch.zeros(n, dtype=torch.int64, device="cuda")
            coupe_amd.Rcb(iter_count=iters, tolerance=tol).partition(
                out, (torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda()), context=ctx)
            torch.cuda.synchronize()
            exp, _ = oracle_rcb(oracle_mod, pts, w, iters, tol)
            assert np.array_equal(out.cpu().numpy().astype(np.uint64), exp), f"queued {queued} case {q}"
            st = ctx.stats()
            reruns += st["reruns"]
            if queued == 0:
                assert st["reruns"] == 0
            elif st["reruns"] == 0:
                assert st["max_passes"] <= queued + 1
        if queued == 1:
            assert reruns > 0, "tol 0 searches should outrun one queued pass"
    finally:
        ctx.close()


@pytest.mark.parametrize("pairs", ["0", "1"])
def test_move_every_level_or_every_other(oracle_mod, monkeypatch, pairs):
    """COUPE_HIP_PAIRS=0 moves the points at every level, 1 (the default)
    once per two levels with the children searched in place; same ids."""
    torch = pytest.importorskip("torch")
    monkeypatch.setenv("COUPE_HIP_PAIRS", pairs)
    ctx = coupe_amd.DeviceContext(0)
    try:
        rng = np.random.default_rng(900)
        for q in range(8):
            n = int(rng.integers(1_000, 150_000))
            dim = 2 + q % 2
            pts = [rng.random((n, dim)), rng.normal(0, 1, (n, dim)) ** 3,
                   np.round(rng.random((n, dim)) * 20) / 20][q % 3]
            w = rng.integers(1, 1000, n).astype([np.int32, np.int64, np.float64][q % 3])
            iters, tol = int(rng.integers(1, 14)), float(rng.choice([0.0, 0.01, 0.05]))
            say(f"pairs {pairs}, case {q}: n={n} d={dim} L={iters} tol={tol}")
            out = torch.zeros(n, dtype=torch.int64, device="cuda")
            coupe_amd.Rcb(iter_count=iters, tolerance=tol).partition(
                out, (torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda()), context=ctx)
            torch.cuda.synchronize()
            exp, _ = oracle_rcb(oracle_mod, pts, w, iters, tol)
            assert np.array_equal(out.cpu().numpy().astype(np.uint64), exp), f"pairs {pairs} case {q}"
    finally:
        ctx.close()


@pytest.mark.parametrize("forward", ["0", "1"])
def test_move_forwards_the_next_first_pass(oracle_mod, monkeypatch, forward):
    """The move to level l + 2 forwards the first pass of that level's
    children (COUPE_HIP_FORWARD, default on): parts whose split ends at their
    first candidate skip the children's read pass, the others read it.  Cases
    mix both (grids and uniform clouds accept the midpoint, skewed clouds do
    not), with parts of many tiles whose runs cross destination tiles."""
    torch = pytest.importorskip("torch")
    monkeypatch.setenv("COUPE_HIP_FORWARD", forward)
    ctx = coupe_amd.DeviceContext(0)
    try:
        rng = np.random.default_rng(1300)
        seen_forwarded = seen_read = 0
        for q in range(8):
            n = int(rng.integers(20_000, 700_000))
            dim = 2 + q % 2
            pts = [rng.random((n, dim)), rng.normal(0, 1, (n, dim)) ** 3,
                   (np.round(rng.random((n, dim)) * 64) + 0.5) / 64,
                   np.concatenate([rng.random((n // 2, dim)), rng.normal(0.3, 0.01, (n - n // 2, dim))])][q % 4]
            w = rng.integers(1, 1000, n).astype([np.int32, np.int64, np.float64][q % 3])
            iters, tol = int(rng.integers(4, 14)), float(rng.choice([0.0, 0.01, 0.05, 0.2]))
            say(f"forward {forward}, case {q}: n={n} d={dim} L={iters} tol={tol}")
            out = torch.zeros(n, dtype=torch.int64, device="cuda")
            coupe_amd.Rcb(iter_count=iters, tolerance=tol).partition(
                out, (torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda()), context=ctx)
            torch.cuda.synchronize()
            st = ctx.stats()
            seen_forwarded += st["forwarded"]
            seen_read += st["first_read"]
            exp, _ = oracle_rcb(oracle_mod, pts, w, iters, tol)
            assert np.array_equal(out.cpu().numpy().astype(np.uint64), exp), f"forward {forward} case {q}"
        assert seen_read > 0
        if forward == "1":
            assert seen_forwarded > 0
        else:
            assert seen_forwarded == 0
    finally:
        ctx.close()


@pytest.mark.parametrize("full", ["0", "1"])
def test_last_pair_full_first_pass(oracle_mod, monkeypatch, full):
    """A second call of the same shape takes the last pair's children's first
    pass with every candidate (COUPE_HIP_FULL_FIRST, default on) when, in the
    first call, most parts two levels up were still searching after their
    first decision: uniform clouds at tolerance 0, skewed clouds at tight
    tolerances.  Both calls equal the
    oracle, and with the switch on the second call takes that pass (the
    `full_first` statistic)."""
    torch = pytest.importorskip("torch")
    monkeypatch.setenv("COUPE_HIP_FULL_FIRST", full)
    ctx = coupe_amd.DeviceContext(0)
    try:
        rng = np.random.default_rng(1500)
        for q in range(6):
            n = int(rng.integers(30_000, 400_000))
            dim = 2 + q % 2
            # uniform clouds at tolerance 0: no split ends at its first candidate
            pts = rng.normal(0, 1, (n, dim)) ** 3 if q % 3 else rng.random((n, dim))
            w = rng.integers(1, 1000, n).astype([np.int32, np.int64, np.float64][q % 3])
            iters = [4, 6, 8][q % 3]
            tol = float([0.0, 0.001, 0.01][q % 3])
            say(f"full_first {full}, case {q}: n={n} d={dim} L={iters} tol={tol}")
            exp, _ = oracle_rcb(oracle_mod, pts, w, iters, tol)
            used = []
            for call in range(2):
                out = torch.zeros(n, dtype=torch.int64, device="cuda")
                coupe_amd.Rcb(iter_count=iters, tolerance=tol).partition(
                    out, (torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda()), context=ctx)
                torch.cuda.synchronize()
                used.append(ctx.stats()["full_first"])
                assert np.array_equal(out.cpu().numpy().astype(np.uint64), exp), f"full {full} case {q} call {call}"
            assert used[0] == 0, f"case {q}: no hint on the first call"
            if full == "0":
                assert used[1] == 0
            elif tol == 0.0:   # no part splits at its first candidate
                assert used[1] == 1, f"case {q}: the second call's last pair took the 2-bin first pass"
    finally:
        ctx.close()


@pytest.mark.parametrize("w16", ["0", "1"])
def test_stored_weight_widths(oracle_mod, monkeypatch, w16):
    """Integer weights in [0, 65535] are stored in 16 bits (COUPE_HIP_WEIGHTS16,
    default on), other integers in 32 or 64, non-integers as f64; ids equal
    either way.  Cases at and past the 16-bit edges: 65535, 65536, 0, -1."""
    torch = pytest.importorskip("torch")
    monkeypatch.setenv("COUPE_HIP_WEIGHTS16", w16)
    ctx = coupe_amd.DeviceContext(0)
    try:
        rng = np.random.default_rng(1500)
        widths = []
        for q in range(8):
            n = int(rng.integers(5_000, 200_000))
            dim = 2 + q % 2
            pts = rng.random((n, dim))
            hi = [65535, 65536, 1000, 40000, 1 << 40, 3][q % 6]
            lo = [0, 0, 1, -1, 1, 0][q % 6]
            dt = [np.int32, np.int64, np.float64][q % 3]
            if dt == np.int32:
                hi = min(hi, (1 << 31) - 1)
            w = rng.integers(lo, hi, n, endpoint=True).astype(dt)
            if q % 6 == 0:
                w[rng.integers(0, n, 10)] = 65535   # the top of the 16-bit range
            iters, tol = int(rng.integers(3, 12)), float(rng.choice([0.0, 0.01, 0.05]))
            say(f"weights16 {w16}, case {q}: n={n} d={dim} w in [{w.min()}, {w.max()}] {w.dtype} L={iters}")
            out = torch.zeros(n, dtype=torch.int64, device="cuda")
            coupe_amd.Rcb(iter_count=iters, tolerance=tol).partition(
                out, (torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda()), context=ctx)
            torch.cuda.synchronize()
            widths.append(ctx.stats()["weight_bytes"])
            exp, _ = oracle_rcb(oracle_mod, pts, w, iters, tol)
            assert np.array_equal(out.cpu().numpy().astype(np.uint64), exp), f"weights16 {w16} case {q}"
        assert (2 in widths) == (w16 == "1")
        assert any(x > 2 for x in widths)
    finally:
        ctx.close()


def test_edge_cases(oracle_mod):
    rng = np.random.default_rng(7)
    one = rng.random((1, 3))
    assert_parity(oracle_mod, one, np.ones(1), 6, 0.05, "single point")
    assert_parity(oracle_mod, rng.random((2, 2)), np.ones(2, np.int32), 3, 0.0, "two points")
    same = np.tile([[0.25, -3.0, 7.0]], (1000, 1))
    assert_parity(oracle_mod, same, np.ones(1000), 5, 0.05, "all points equal")
    assert_parity(oracle_mod, rng.random((5000, 3)), np.ones(5000), 0, 0.05, "iter_count 0")
    assert_parity(oracle_mod, rng.random((300, 2)), np.ones(300), 20, 0.05, "deep: 2^20 leaves, 300 points")
    assert_parity(oracle_mod, rng.random((3000, 3)), np.ones(3000), 9, -1.0, "negative tolerance")
    big = rng.integers(2 ** 28, 2 ** 31 - 1, 4000).astype(np.int32)
    assert_parity(oracle_mod, rng.random((4000, 2)), big, 6, 0.01, "i32 sum wraps")
    huge = rng.integers(2 ** 40, 2 ** 41, 4000).astype(np.int64)
    assert_parity(oracle_mod, rng.random((4000, 3)), huge, 6, 0.01, "int64 beyond int32")
    wide = rng.normal(0, 1, (4000, 3)) * 10.0 ** rng.integers(-30, 30, (4000, 3))
    assert_parity(oracle_mod, wide, np.ones(4000), 8, 0.01, "wide dynamic range")
    tiny = rng.random((4000, 2)) * 1e-40   # f32 denormals after the cast
    assert_parity(oracle_mod, tiny, np.ones(4000), 8, 0.01, "denormal coordinates")
    zw = np.zeros(2000)
    assert_parity(oracle_mod, rng.random((2000, 2)), zw, 4, 0.05, "zero weights (NaN imbalance)")
    negw = rng.integers(-50, 50, 2000).astype(np.float64)
    assert_parity(oracle_mod, rng.random((2000, 3)), negw, 5, 0.05, "negative weights")


def test_real_weights(oracle_mod):
    rng = np.random.default_rng(9)
    for k in range(4):
        n = 50_000
        pts = rng.random((n, 2 + k % 2))
        w = rng.random(n) * 3.0 + 0.1
        assert_parity(oracle_mod, pts, w, 8, 0.01, f"real weights {k}")


def test_uniform2d_full(oracle_mod):
    pts, w, it, tol = workloads.standard("uniform2d")
    assert_parity(oracle_mod, pts, w, it, tol, "uniform2d 2^20")


def test_gauss3d_scaled(oracle_mod):
    pts, w, it, tol = workloads.standard("gauss3d", scale=4)
    assert_parity(oracle_mod, pts, w, it, tol, "gauss3d 2^20")


def test_jgrid_scaled(oracle_mod):
    pts, w, it, tol = workloads.standard("jgrid512", scale=6)   # 128^3
    assert_parity(oracle_mod, pts, w, it, tol, "jgrid 128^3")


@pytest.mark.slow
def test_gauss3d_full(oracle_mod):
    pts, w, it, tol = workloads.standard("gauss3d")
    assert_parity(oracle_mod, pts, w, it, tol, "gauss3d 2^24", parallel_oracle=True)


@pytest.mark.slow
def test_grid512_closed_form():
    # 512^3 cell centres, unit weights, 4096 parts: ids are the interleaved
    # cell-index bits (axis 0, 1, 2, 0, ...), MSB first.
    say("grid512 closed form: n=2^27 L=12")
    pts, w = workloads.jgrid(512, jitter=False)
    got = gpu_rcb(pts, w, 12, 0.01)
    del pts
    idx = np.arange(512 ** 3, dtype=np.int64)
    exp = np.zeros(512 ** 3, dtype=np.uint64)
    for level in range(12):
        a = level % 3
        coord = (idx // 512 ** a) % 512
        bit = ((coord >> (8 - level // 3)) & 1).astype(np.uint64)
        exp |= bit << np.uint64(11 - level)
    assert np.array_equal(got, exp)


def test_device_entry_matches_host_entry():
    torch = pytest.importorskip("torch")
    pts, w, it, tol = workloads.standard("gauss3d", scale=6)
    host = gpu_rcb(pts, w, it, tol)
    dp = torch.from_numpy(pts).cuda()
    dw = torch.from_numpy(w).cuda()
    for _ in range(2):   # the context's buffers are reused across calls
        out = torch.zeros(pts.shape[0], dtype=torch.int64, device="cuda")
        coupe_amd.Rcb(iter_count=it, tolerance=tol).partition(out, (dp, dw))
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy().astype(np.uint64), host)


# ---------------------------------------------------------------------------
# Batched RCB (coupe_hip_rcb_batch): every instance equals its lone run.
# ---------------------------------------------------------------------------
def _batch_case(seed, dim, count):
    rng = np.random.default_rng(seed)
    specs, pts, ws = [], [], []
    for q in range(count):
        n = int(rng.choice([0, 1, 2, 17, 300, 2048, 2049, 5000, 40000]))
        kind = int(rng.integers(0, 3))
        if kind == 0:
            p = rng.random((n, dim))
        elif kind == 1:
            p = rng.integers(0, 7, (n, dim)).astype(np.float64)
        else:
            p = rng.normal(0, 1, (n, dim)) * 10.0 ** rng.integers(-2, 3)
        pts.append(p)
        ws.append(rng.integers(1, 50, n).astype(np.float64))
        specs.append((n, int(rng.integers(0, 11)), float(rng.choice([0.0, 0.02, 0.05]))))
    return specs, np.concatenate(pts).reshape(-1, dim), np.concatenate(ws)


@pytest.mark.parametrize("seed,dim", [(0, 2), (1, 3), (2, 2), (3, 3)])
def test_batch_matches_each_instance(oracle_mod, seed, dim):
    torch = pytest.importorskip("torch")
    specs, pts, w = _batch_case(seed, dim, 40)
    say(f"batch seed {seed} dim {dim}: {len(specs)} instances, {pts.shape[0]} points")
    out = torch.zeros(pts.shape[0], dtype=torch.int64, device="cuda")
    coupe_amd.partition_batch(out, torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda(), specs)
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.uint64)
    off = 0
    for q, (n, iters, tol) in enumerate(specs):
        exp, _ = oracle_rcb(oracle_mod, pts[off:off + n], w[off:off + n], iters, tol)
        assert np.array_equal(got[off:off + n], exp), f"instance {q}: n={n} L={iters} tol={tol}"
        off += n


def test_batch_of_one_equals_plain_call():
    torch = pytest.importorskip("torch")
    pts, w, it, tol = workloads.standard("gauss3d", scale=6)
    plain = gpu_rcb(pts, w, it, tol)
    out = torch.zeros(pts.shape[0], dtype=torch.int64, device="cuda")
    coupe_amd.partition_batch(out, torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda(),
                              [(pts.shape[0], it, tol)])
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().astype(np.uint64), plain)


@pytest.mark.parametrize("seed", [0, 1])
def test_batch_of_both_dimensions(oracle_mod, seed):
    """coupe_hip_rcb_batch_mixed: 2D and 3D instances alternating in one call."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(60 + seed)
    specs, flat, ws = [], [], []
    for q in range(50):
        d = int(rng.integers(2, 4))
        n = int(rng.choice([0, 1, 2, 17, 300, 2048, 2049, 5000, 40000]))
        p = rng.random((n, d)) if q % 3 else rng.normal(0, 1, (n, d)) ** 3
        flat.append(p.ravel())
        ws.append(rng.integers(1, 50, n).astype([np.int64, np.float64, np.int32][seed]))
        specs.append((n, d, int(rng.integers(0, 11)), float(rng.choice([0.0, 0.02, 0.05]))))
    flat, w = np.concatenate(flat), np.concatenate(ws)
    say(f"mixed batch seed {seed}: {len(specs)} instances, {w.shape[0]} points")
    out = torch.full((w.shape[0],), -1, dtype=torch.int64, device="cuda")
    coupe_amd.partition_batch_mixed(out, torch.from_numpy(flat).cuda(), torch.from_numpy(w).cuda(), specs)
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.uint64)
    off = poff = 0
    for q, (n, d, iters, tol) in enumerate(specs):
        exp, _ = oracle_rcb(oracle_mod, flat[poff:poff + n * d].reshape(n, d), w[off:off + n], iters, tol)
        assert np.array_equal(got[off:off + n], exp), f"instance {q}: n={n} D={d} L={iters} tol={tol}"
        off += n
        poff += n * d


def test_null_arrays_refused():
    """A null array with points to read or ids to write is refused before
    any device work, by every entry that takes device arrays."""
    import torch

    lib = _native.lib()
    ctx = coupe_amd._context_for(0)
    n = 1000
    pts = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    w = torch.ones(n, dtype=torch.float64, device="cuda")
    out = torch.zeros(n, dtype=torch.int64, device="cuda")
    P, W, O = pts.data_ptr(), w.data_ptr(), out.data_ptr()
    for p, ww, o in [(None, W, O), (P, None, O), (P, W, None)]:
        print(f"[gpu] rcb_device with a null array: {p is None, ww is None, o is None}", flush=True)
        assert lib.coupe_hip_rcb_device(ctx.handle, None, o, n, 3, p, _native.COUPE_DOUBLE, ww, 4,
                                        0.05) == _native.CRASH
        inst = (_native.Instance * 1)()
        inst[0].n, inst[0].iter_count, inst[0].tolerance = n, 4, 0.05
        assert lib.coupe_hip_rcb_batch(ctx.handle, None, o, 3, p, _native.COUPE_DOUBLE, ww, inst,
                                       1) == _native.CRASH
    loads = (C.c_double * 4)()
    assert lib.coupe_hip_parts_load(ctx.handle, None, None, n, _native.COUPE_DOUBLE, W, 4, loads) == _native.CRASH
    imb = C.c_double()
    assert lib.coupe_hip_imbalance(ctx.handle, None, O, n, _native.COUPE_DOUBLE, None, 4,
                                   C.byref(imb)) == _native.CRASH
    # no points: nothing to read or write, so null arrays are fine
    assert lib.coupe_hip_rcb_device(ctx.handle, None, None, 0, 3, None, _native.COUPE_DOUBLE, None, 4,
                                    0.05) == _native.OK
