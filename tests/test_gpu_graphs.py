"""Calls replayed as captured graphs: with graphs on (COUPE_HIP_GRAPHS=1, or
chosen by the autotuner), a context captures a call of a shape it has run
before (hints known) once its graph key repeats, and replays the graph for
later calls of that shape.  Every replay must give the eager path's ids -- the data is read at
replay time, a weight mode the graph cannot hold or a search longer than its
queued passes falls back to eager runs -- with eager work and
synchronisations of the program between the replays."""
import sys

import numpy as np
import pytest

import coupe_amd

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _ids(oracle, pts, w, iters, tol):
    err, ids, _ = oracle.rcb(pts, w, iters, tol)
    assert err == 0
    return ids


def test_replays_read_the_current_data(oracle_mod, monkeypatch):
    torch = pytest.importorskip("torch")
    monkeypatch.setenv("COUPE_HIP_GRAPHS", "1")
    rng = np.random.default_rng(70)
    ctx = coupe_amd.DeviceContext(0)
    try:
        n, iters, tol = 300_000, 10, 0.01
        d_pts = torch.empty((n, 3), dtype=torch.float64, device="cuda")
        d_w = torch.empty(n, dtype=torch.float64, device="cuda")
        out = torch.empty(n, dtype=torch.int64, device="cuda")
        seen = []
        # data kinds: uniform, clustered (longer searches), weights past 16
        # bits (another weight mode), then uniform again
        kinds = ["uniform", "uniform", "uniform", "cluster", "uniform", "wide", "wide", "uniform", "cluster"]
        for q, kind in enumerate(kinds):
            pts = rng.random((n, 3)) if kind != "cluster" else rng.normal(0, 1, (n, 3)) ** 5
            w = (rng.integers(1, 100, n) if kind != "wide" else rng.integers(1, 1 << 30, n)).astype(np.float64)
            d_pts.copy_(torch.from_numpy(pts))
            d_w.copy_(torch.from_numpy(w))
            torch.cuda.synchronize()
            out.fill_(-1)   # the program's own eager work between replays
            coupe_amd.Rcb(iter_count=iters, tolerance=tol).partition(out, (d_pts, d_w), context=ctx)
            torch.cuda.synchronize()
            st = ctx.stats()
            seen.append(st["graph"])
            say(f"graph replay call {q} ({kind}): graph {st['graph']} reruns {st['reruns']}")
            got = out.cpu().numpy().astype(np.uint64)
            assert np.array_equal(got, _ids(oracle_mod, pts, w, iters, tol)), f"call {q} ({kind})"
        assert seen[0] == 0, seen
        # the same data again and again: captured once its key repeats, then replayed
        exp = _ids(oracle_mod, pts, w, iters, tol)
        again = []
        for rep in range(5):
            out.fill_(-1)
            coupe_amd.Rcb(iter_count=iters, tolerance=tol).partition(out, (d_pts, d_w), context=ctx)
            torch.cuda.synchronize()
            again.append(ctx.stats()["graph"])
            assert np.array_equal(out.cpu().numpy().astype(np.uint64), exp), f"repeat {rep}"
        say(f"graph use over repeated calls: {again}")
        assert 2 in again and again[-1] == 1 and again[-2] == 1, again
    finally:
        ctx.close()


def test_shapes_alternating_on_one_context(oracle_mod, monkeypatch):
    torch = pytest.importorskip("torch")
    monkeypatch.setenv("COUPE_HIP_GRAPHS", "1")
    rng = np.random.default_rng(71)
    ctx = coupe_amd.DeviceContext(0)
    try:
        shapes = [(100_000, 2, 7, 0.05), (250_000, 3, 12, 0.01), (4_096, 3, 5, 0.0)]
        data = []
        for n, d, it, tol in shapes:
            pts = rng.random((n, d))
            w = rng.integers(1, 1000, n).astype(np.int64)
            data.append((torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda(),
                         torch.empty(n, dtype=torch.int64, device="cuda"), _ids(oracle_mod, pts, w, it, tol)))
        graphs = []
        for rep in range(5):
            for k, (n, d, it, tol) in enumerate(shapes):
                dp, dw, out, exp = data[k]
                out.fill_(-1)
                coupe_amd.Rcb(iter_count=it, tolerance=tol).partition(out, (dp, dw), context=ctx)
                torch.cuda.synchronize()
                graphs.append(ctx.stats()["graph"])
                assert np.array_equal(out.cpu().numpy().astype(np.uint64), exp), f"rep {rep} shape {k}"
        say(f"graph use, shapes alternating: {graphs}")
        assert graphs[-3:].count(1) >= 2, graphs
    finally:
        ctx.close()


def test_batches_replayed(oracle_mod, monkeypatch):
    """Batches of both dimensions, each call replayed; COUPE_HIP_GRAPHS=0
    gives the same ids kernel by kernel."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(72)
    specs, flat, ws = [], [], []
    for q in range(80):
        d = 2 + q % 2
        n = int(10 ** rng.uniform(2, 4.5))
        flat.append(rng.random((n, d)).ravel())
        ws.append(rng.integers(1, 16, n).astype(np.float64))
        specs.append((n, d, int(rng.integers(4, 9)), 0.02))
    flat, w = np.concatenate(flat), np.concatenate(ws)
    d_flat, d_w = torch.from_numpy(flat).cuda(), torch.from_numpy(w).cuda()
    expect = []
    off = poff = 0
    for n, d, it, tol in specs:
        expect.append(_ids(oracle_mod, flat[poff:poff + n * d].reshape(n, d), w[off:off + n], it, tol))
        off += n
        poff += n * d
    expect = np.concatenate(expect)
    for graphs in ("1", "0"):
        monkeypatch.setenv("COUPE_HIP_GRAPHS", graphs)
        ctx = coupe_amd.DeviceContext(0)
        try:
            out = torch.empty(w.shape[0], dtype=torch.int64, device="cuda")
            for rep in range(4):
                out.fill_(-1)
                coupe_amd.partition_batch_mixed(out, d_flat, d_w, specs, context=ctx)
                torch.cuda.synchronize()
                say(f"mixed batch, graphs {graphs}, call {rep}: graph {ctx.stats()['graph']}")
                assert np.array_equal(out.cpu().numpy().astype(np.uint64), expect), f"graphs {graphs} call {rep}"
                if graphs == "0":
                    assert ctx.stats()["graph"] == 0
        finally:
            ctx.close()


def test_autotuned_switches_keep_the_ids(oracle_mod):
    """The autotuner's candidates (bins, workgroups per CU, full first pass)
    change the work, never the ids; the switches it keeps apply to later
    calls of the batch's class."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(73)
    specs, flat, ws = [], [], []
    for q in range(40):
        d = 3 if q % 3 else 2
        n = int(10 ** rng.uniform(3, 5))
        flat.append((rng.normal(0, 1, (n, d)) ** 3).ravel())
        ws.append(rng.integers(1, 100, n).astype(np.int64))
        specs.append((n, d, int(rng.integers(4, 11)), 0.01))
    flat, w = np.concatenate(flat), np.concatenate(ws)
    expect = []
    off = poff = 0
    for n, d, it, tol in specs:
        expect.append(_ids(oracle_mod, flat[poff:poff + n * d].reshape(n, d), w[off:off + n], it, tol))
        off += n
        poff += n * d
    expect = np.concatenate(expect)
    ctx = coupe_amd.DeviceContext(0)
    try:
        d_flat, d_w = torch.from_numpy(flat).cuda(), torch.from_numpy(w).cuda()
        out = torch.full((w.shape[0],), -1, dtype=torch.int64, device="cuda")
        chosen = coupe_amd.autotune(out, d_flat, d_w, specs, trials=2, context=ctx)
        torch.cuda.synchronize()
        say(f"autotune: {chosen}")
        assert np.array_equal(out.cpu().numpy().astype(np.uint64), expect)
        out.fill_(-1)
        coupe_amd.partition_batch_mixed(out, d_flat, d_w, specs, context=ctx)
        torch.cuda.synchronize()
        st = ctx.stats()
        assert st["tuned"] == 1 and st["tuning"]["bins"] == chosen["bins"], st
        assert st["tuning"]["search_chunk"] == chosen["search_chunk"], st
        assert np.array_equal(out.cpu().numpy().astype(np.uint64), expect)
    finally:
        ctx.close()


@pytest.mark.parametrize("bins,wg,chunk", [(16, 1, 1), (32, 2, 4), (64, 0, 3), (128, 4, 2), (256, 0, 1)])
def test_switches_from_the_environment(oracle_mod, monkeypatch, bins, wg, chunk):
    """COUPE_HIP_BINS / COUPE_HIP_WG_PER_CU / COUPE_HIP_SEARCH_CHUNK: the
    values the tuner tries, on clustered clouds whose searches take several
    passes."""
    torch = pytest.importorskip("torch")
    monkeypatch.setenv("COUPE_HIP_BINS", str(bins))
    monkeypatch.setenv("COUPE_HIP_WG_PER_CU", str(wg))
    monkeypatch.setenv("COUPE_HIP_SEARCH_CHUNK", str(chunk))
    rng = np.random.default_rng(74 + bins)
    ctx = coupe_amd.DeviceContext(0)
    try:
        for q in range(3):
            n = int(rng.integers(50_000, 400_000))
            pts = rng.normal(0, 1, (n, 3)) ** 5
            w = rng.integers(1, 300, n).astype(np.int64)
            it, tol = 10, [0.0, 0.001, 0.01][q]
            say(f"bins {bins} wg {wg} case {q}: n={n}")
            out = torch.full((n,), -1, dtype=torch.int64, device="cuda")
            coupe_amd.Rcb(iter_count=it, tolerance=tol).partition(
                out, (torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda()), context=ctx)
            torch.cuda.synchronize()
            assert ctx.stats()["tuning"]["bins"] == bins and ctx.stats()["tuning"]["search_chunk"] == chunk
            assert np.array_equal(out.cpu().numpy().astype(np.uint64), _ids(oracle_mod, pts, w, it, tol))
    finally:
        ctx.close()


def test_level_with_no_queued_pass_reruns_when_new_data_needs_one(oracle_mod):
    """A level whose search ended at its first decision in the context's last
    call of the same shape queues no search pass: data of that shape which
    needs passes there is closed by that first decision, and the call runs
    again with host checks.  Grid cell centres (every split at its box
    midpoint) and clustered points alternate on one context; every call has
    the oracle's ids."""
    torch = pytest.importorskip("torch")
    m = 32
    g = (np.arange(m) + 0.5) / m
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    grid = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    n = grid.shape[0]
    rng = np.random.default_rng(91)
    clustered = rng.normal(0, 1, (n, 3)) ** 5
    ctx = coupe_amd.DeviceContext(0)
    try:
        for q, pts in enumerate([grid, clustered, grid, clustered, grid]):
            w = np.ones(n, dtype=np.int64) if q % 2 == 0 else rng.integers(1, 50, n).astype(np.int64)
            out = torch.full((n,), -1, dtype=torch.int64, device="cuda")
            coupe_amd.Rcb(iter_count=9, tolerance=0.01).partition(
                out, (torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda()), context=ctx)
            torch.cuda.synchronize()
            st = ctx.stats()
            say(f"alternating data, call {q}: reruns {st['reruns']}, passes {st['passes']}")
            assert np.array_equal(out.cpu().numpy().astype(np.uint64), _ids(oracle_mod, pts, w, 9, 0.01)), f"call {q}"
            if q == 1:
                assert st["reruns"] == 1, st   # no pass was queued where the grid needed none
    finally:
        ctx.close()


@pytest.mark.parametrize("nt_bytes", ["0", "1e15"])
def test_move_store_modes(oracle_mod, monkeypatch, nt_bytes):
    """The move stores non-temporally when its size passes the switch and the
    shape's last call moved tiles to few destinations (grid order), else
    through the caches, and the id descent reads and writes non-temporally
    when its size passes the switch: both forms give the oracle's ids, on
    data in grid order and in random order, on the first and the later calls
    of a shape."""
    torch = pytest.importorskip("torch")
    monkeypatch.setenv("COUPE_HIP_MOVE_NT_BYTES", nt_bytes)
    from coupe_amd import workloads
    ctx = coupe_amd.DeviceContext(0)
    try:
        pts, w = workloads.jgrid(48, 2, True)
        perm = np.random.default_rng(9).permutation(len(w))
        for name, p, ww in (("grid order", pts, w), ("random order", pts[perm], w[perm])):
            exp = _ids(oracle_mod, p, ww, 10, 0.01)
            d_p, d_w = torch.from_numpy(np.ascontiguousarray(p)).cuda(), torch.from_numpy(ww).cuda()
            out = torch.empty(len(ww), dtype=torch.int64, device="cuda")
            for call in range(3):
                out.fill_(-1)
                coupe_amd.Rcb(iter_count=10, tolerance=0.01).partition(out, (d_p, d_w), context=ctx)
                torch.cuda.synchronize()
                say(f"store mode switch {nt_bytes}, {name}, call {call}")
                assert np.array_equal(out.cpu().numpy().astype(np.uint64), exp), f"{name} call {call}"
    finally:
        ctx.close()
